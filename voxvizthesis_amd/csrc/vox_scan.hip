// vox_scan.hip -- C ABI of the device-wide exclusive prefix sum
// (include/voxmi.h: vox_scan_u32).
//
// The voxelizer sizes its (group, tile) bins, its per-level tile lists and
// its brick addresses with one single-pass decoupled look-back scan
// (k_scan in vox_kernels.hip).  This entry point runs the same kernel on
// caller-owned device arrays: the look-back scratch (one 64-bit status per
// 8192-value tile, zeroed before first use, left reusable by every scan) and
// the time-out word are allocated and freed by the call.
//
// vox_scan_u32_noalloc runs the same kernel with its scratch in a
// file-scope device array (`scratch` below) instead: no device allocation,
// for callers that cannot allocate.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/voxmi.h"
#include "vox_internal.h"

using namespace voxmi;

namespace {

thread_local std::string t_scanErr;

bool scan_fail(const std::string& m) {
    t_scanErr = m;
    return false;
}

// Arguments the scan kernel assumes: out holds n + 1 values, both arrays are
// 16-byte aligned (whole tiles load and store 16 bytes per lane).
bool scan_args_ok(const uint32_t* in, const uint32_t* out, uint64_t n) {
    if (!out) return scan_fail("vox_scan_u32: out is null");
    if (n && !in) return scan_fail("vox_scan_u32: in is null");
    if ((reinterpret_cast<uintptr_t>(in) & 15u) || (reinterpret_cast<uintptr_t>(out) & 15u))
        return scan_fail("vox_scan_u32: in and out must be 16-byte aligned");
    if (n >= (1ull << 40)) return scan_fail("vox_scan_u32: n above 2^40 values");
    return true;
}

// Scratch of vox_scan_u32_noalloc: the look-back area (scan_tmp_words(n)
// words for n <= kNoAllocMaxN) followed, at word kNoAllocErrWord, by the
// time-out word.  One copy per GPU, held until the process ends; every call
// on a GPU shares it, so calls that overlap in time on one GPU race.
constexpr uint64_t kNoAllocMaxN = 1ull << 32;
constexpr uint64_t kNoAllocErrWord = ((2 * ((kNoAllocMaxN + 8191) / 8192 + 1)) + 63) & ~63ull;
constexpr uint64_t kNoAllocWords = kNoAllocErrWord + 64;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (the host-side name is not a device address).
constexpr int kMaxDevices = 64;
std::mutex g_scratchMu;
uint32_t* g_scratchPtr[kMaxDevices] = {};

uint32_t* scratch_ptr(int dev) {
    std::lock_guard<std::mutex> lk(g_scratchMu);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            scan_fail(std::string("vox_scan_u32_noalloc: hipGetSymbolAddress: ") +
                      hipGetErrorString(e));
            return nullptr;
        }
        g_scratchPtr[dev] = static_cast<uint32_t*>(p);
    }
    return g_scratchPtr[dev];
}

}  // namespace

extern "C" {

int vox_scan_u32(uint32_t* in, uint32_t* out, uint64_t n, int32_t clear_input) {
    t_scanErr.clear();
    if (!scan_args_ok(in, out, n)) return 0;
    std::stringstream err;
    DevBuf tmp, word;
    // the scratch is zeroed by ensure_scan_tmp when allocated; the time-out
    // word by a fill kernel (no memset: the same kernels run under capture)
    if (!ensure_scan_tmp(tmp, n, nullptr, err, "scan scratch") ||
        !word.ensure(4, err, "scan time-out word")) {
        tmp.release();
        return scan_fail(err.str()), 0;
    }
    launch_fill_u32(nullptr, word.as<uint32_t>(), 1, 0u);
    if (clear_input)
        launch_scan_clear(nullptr, in, out, n, tmp.as<uint32_t>(), word.as<uint32_t>());
    else
        launch_scan(nullptr, in, out, n, tmp.as<uint32_t>(), word.as<uint32_t>());
    uint32_t timedOut = 0;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(&timedOut, word.p, 4, hipMemcpyDeviceToHost);
    tmp.release();
    word.release();
    if (e != hipSuccess) return scan_fail(std::string("vox_scan_u32: ") + hipGetErrorString(e)), 0;
    if (timedOut) return scan_fail("vox_scan_u32: a workgroup gave up waiting on a predecessor"), 0;
    return 1;
}

int vox_scan_u32_noalloc(uint32_t* in, uint32_t* out, uint64_t n, int32_t clear_input) {
    t_scanErr.clear();
    if (!scan_args_ok(in, out, n)) return 0;
    if (n > kNoAllocMaxN)
        return scan_fail("vox_scan_u32_noalloc: n above 2^32 values (the size of its fixed "
                         "scratch)"), 0;
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return scan_fail(std::string("vox_scan_u32_noalloc: ") + hipGetErrorString(e)), 0;
    if (dev < 0 || dev >= kMaxDevices) return scan_fail("vox_scan_u32_noalloc: device index above 63"), 0;
    uint32_t* base = scratch_ptr(dev);
    if (!base) return 0;
    uint32_t* word = base + kNoAllocErrWord;
    // the array keeps the last call's contents: clear the look-back words
    // this call uses and the time-out word (fill kernels, no memset)
    launch_fill_u32(nullptr, base, scan_tmp_words(n), 0u);
    launch_fill_u32(nullptr, word, 1, 0u);
    if (clear_input)
        launch_scan_clear(nullptr, in, out, n, base, word);
    else
        launch_scan(nullptr, in, out, n, base, word);
    uint32_t timedOut = 0;
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(&timedOut, word, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return scan_fail(std::string("vox_scan_u32_noalloc: ") + hipGetErrorString(e)), 0;
    if (timedOut) return scan_fail("vox_scan_u32_noalloc: a workgroup gave up waiting on a predecessor"), 0;
    return 1;
}

const char* vox_scan_get_error(void) { return t_scanErr.c_str(); }

}  // extern "C"

// vox_scan.hip -- C ABI of the device-wide exclusive prefix sum
// (include/voxmi.h: vox_scan_u32).
//
// The voxelizer sizes its (group, tile) bins, its per-level tile lists and
// its brick addresses with one single-pass decoupled look-back scan
// (k_scan in vox_kernels.hip).  This entry point runs the same kernel on
// caller-owned device arrays: the look-back scratch (one 64-bit status per
// 8192-value tile, zeroed before first use, left reusable by every scan) and
// the time-out word are allocated and freed by the call.
#include <hip/hip_runtime.h>

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

const char* vox_scan_get_error(void) { return t_scanErr.c_str(); }

}  // extern "C"

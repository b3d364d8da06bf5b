// vox_sort.hip -- key sort of the bricks written with several chunks (the
// chunk that completes a brick, then inside-one-chunk before across-chunks,
// then node order): a device radix sort of (64-bit key, node) pairs.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "vox_internal.h"

namespace voxmi {

hipError_t chunk_sort_pairs(hipStream_t st, void* tmp, size_t* tmpBytes,
                            const unsigned long long* keysIn, unsigned long long* keysOut,
                            const uint32_t* nodesIn, uint32_t* nodesOut, uint32_t n, int keyBits) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, *tmpBytes, keysIn, keysOut, nodesIn, nodesOut,
                                              n, 0, keyBits, st);
}

}  // namespace voxmi

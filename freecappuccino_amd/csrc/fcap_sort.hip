// Device radix sort of (64-bit key, int value) pairs and exclusive scan of ints
// (hipCUB over rocPRIM), used by the set-up of the network's cluster layout and
// of the coarse space.  Kept in its own translation
// unit: the library templates are heavy to compile.
#include <hipcub/hipcub.hpp>

#include "fcap_internal.h"
#include "fcap_kernels.h"

namespace fcap {

void sort_pairs_u64(hipStream_t s, const unsigned long long *key_in, unsigned long long *key_out, const int *val_in,
                    int *val_out, int n, DevBuf &tmp) {
    size_t bytes = 0;
    FC_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, key_in, key_out, val_in, val_out, n, 0, 64, s));
    tmp.ensure(bytes);
    FC_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp.get(), bytes, key_in, key_out, val_in, val_out, n, 0, 64, s));
}

void exclusive_scan_i32(hipStream_t s, const int *in, int *out, int n, DevBuf &tmp) {
    size_t bytes = 0;
    FC_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s));
    tmp.ensure(bytes);
    FC_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp.get(), bytes, in, out, n, s));
}

}  // namespace fcap

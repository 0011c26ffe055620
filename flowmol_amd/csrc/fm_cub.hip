// fm_cub.hip -- device-wide radix sort (hipCUB) used by sort_mols.
#include "fm_internal.h"

#include <hipcub/hipcub.hpp>

// stable sort of (key, value) pairs by key (radix sort over bits [0, end_bit))
size_t cub_sort_pairs_bytes(int n, int end_bit) {
    size_t b = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const int*)nullptr, (int*)nullptr, (const int*)nullptr,
                                       (int*)nullptr, n, 0, end_bit);
    return b + 256;
}

int cub_sort_pairs(fm_ctx* c, const int* kin, int* kout, const int* vin, int* vout, int n, int end_bit) {
    size_t bytes = c->sort_tmp_bytes;
    FM_HIP(hipcub::DeviceRadixSort::SortPairs(c->sort_tmp, bytes, kin, kout, vin, vout, n, 0, end_bit, c->stream));
    return 0;
}

// fm_cub.hip -- device-wide scan and ordered compaction (hipCUB) used at list rebuilds.
#include "fm_internal.h"

#include <hipcub/hipcub.hpp>

size_t cub_temp_bytes(int n) {
    size_t a = 0, b = 0;
    hipcub::CountingInputIterator<int> it(0);
    hipcub::DeviceScan::ExclusiveSum(nullptr, a, (const int*)nullptr, (int*)nullptr, n);
    hipcub::DeviceSelect::Flagged(nullptr, b, it, (const int*)nullptr, (int*)nullptr, (int*)nullptr, n);
    return (a > b ? a : b) + 256;
}

int cub_exclusive_scan(fm_ctx* c, const int* in, int* out, int n) {
    size_t bytes = c->cub_bytes;
    FM_HIP(hipcub::DeviceScan::ExclusiveSum(c->cub_tmp, bytes, in, out, n, c->stream));
    return 0;
}

// out[0..count) = indices e (ascending) with flags[e] != 0
int cub_select_flagged_index(fm_ctx* c, const int* flags, int* out, int* count, int n) {
    size_t bytes = c->cub_bytes;
    hipcub::CountingInputIterator<int> it(0);
    FM_HIP(hipcub::DeviceSelect::Flagged(c->cub_tmp, bytes, it, flags, out, count, n, c->stream));
    return 0;
}

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

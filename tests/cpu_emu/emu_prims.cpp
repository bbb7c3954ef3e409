// CPU stand-ins for the rocPRIM calls of ikd_prims.hip (stable radix-sort and
// exclusive-scan semantics) in the emulated debugging build.
#include <algorithm>
#include <numeric>
#include <vector>

#include "hip_emu.h"

namespace ikd {

static unsigned long long bits(unsigned long long k, int b0, int b1) {
    unsigned long long v = k >> b0;
    int w = b1 - b0;
    return w >= 64 ? v : (v & ((1ull << w) - 1));
}

hipError_t prim_sort_pairs_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* kin, unsigned long long* kout,
                               const uint32_t* vin, uint32_t* vout, int n, int begin_bit, int end_bit, hipStream_t) {
    if (!tmp) {
        tmp_bytes = 16;
        return hipSuccess;
    }
    emu_tally("prim:sort_pairs_u64");
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return bits(kin[a], begin_bit, end_bit) < bits(kin[b], begin_bit, end_bit); });
    std::vector<unsigned long long> k2(n);
    std::vector<uint32_t> v2(n);
    for (int i = 0; i < n; ++i) {
        k2[i] = kin[idx[i]];
        v2[i] = vin[idx[i]];
    }
    std::copy(k2.begin(), k2.end(), kout);
    std::copy(v2.begin(), v2.end(), vout);
    return hipSuccess;
}

hipError_t prim_sort_keys_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* kin, unsigned long long* kout,
                              int n, int begin_bit, int end_bit, hipStream_t) {
    if (!tmp) {
        tmp_bytes = 16;
        return hipSuccess;
    }
    emu_tally("prim:sort_keys_u64");
    std::vector<unsigned long long> k2(kin, kin + n);
    std::stable_sort(k2.begin(), k2.end(), [&](unsigned long long a, unsigned long long b) {
        return bits(a, begin_bit, end_bit) < bits(b, begin_bit, end_bit);
    });
    std::copy(k2.begin(), k2.end(), kout);
    return hipSuccess;
}

template <class T>
static hipError_t exscan(void* tmp, size_t& tmp_bytes, const T* in, T* out, int n) {
    if (!tmp) {
        tmp_bytes = 16;
        return hipSuccess;
    }
    emu_tally("prim:exscan");
    std::vector<T> v(in, in + n);
    T acc = 0;
    for (int i = 0; i < n; ++i) {
        out[i] = acc;
        acc += v[i];
    }
    return hipSuccess;
}
hipError_t prim_exscan_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* in, unsigned long long* out, int n,
                           hipStream_t) {
    return exscan(tmp, tmp_bytes, in, out, n);
}
hipError_t prim_exscan_i32(void* tmp, size_t& tmp_bytes, const int* in, int* out, int n, hipStream_t) {
    return exscan(tmp, tmp_bytes, in, out, n);
}
hipError_t prim_exscan_i64(void* tmp, size_t& tmp_bytes, const long long* in, long long* out, int n, hipStream_t) {
    return exscan(tmp, tmp_bytes, in, out, n);
}

}  // namespace ikd

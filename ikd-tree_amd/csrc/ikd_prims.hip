// ikd_prims.hip — device-wide sort and scan primitives (rocPRIM), kept in one
// translation unit so the heavy templates compile once.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "ikd_internal.h"

namespace ikd {

hipError_t prim_sort_pairs_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* kin,
                               unsigned long long* kout, const uint32_t* vin, uint32_t* vout, int n, int begin_bit,
                               int end_bit, hipStream_t s) {
    return rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, vin, vout, (unsigned int)n, (unsigned int)begin_bit,
                                     (unsigned int)end_bit, s);
}

hipError_t prim_sort_keys_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* kin, unsigned long long* kout,
                              int n, int begin_bit, int end_bit, hipStream_t s) {
    return rocprim::radix_sort_keys(tmp, tmp_bytes, kin, kout, (unsigned int)n, (unsigned int)begin_bit,
                                    (unsigned int)end_bit, s);
}

hipError_t prim_exscan_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* in, unsigned long long* out, int n,
                           hipStream_t s) {
    return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, 0ull, (size_t)n, rocprim::plus<unsigned long long>(), s);
}

hipError_t prim_exscan_i32(void* tmp, size_t& tmp_bytes, const int* in, int* out, int n, hipStream_t s) {
    return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), s);
}

hipError_t prim_exscan_i64(void* tmp, size_t& tmp_bytes, const long long* in, long long* out, int n, hipStream_t s) {
    return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, 0ll, (size_t)n, rocprim::plus<long long>(), s);
}

}  // namespace ikd

// ikd_gload.h — 16-byte loads that the compiler issues as global (not flat)
// memory instructions.  Pointers the kernels read through arrive inside
// by-value argument structs, where the compiler no longer knows they point
// to global memory and falls back to flat loads (which also wait on the LDS
// counter); casting to address space 1 restores global_load_dwordx4.
#pragma once

namespace ikd {

typedef float gv4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ gv4 gload4(const void* p) {
    return *(const __attribute__((address_space(1))) gv4*)p;
}

}  // namespace ikd

namespace ikd {

// the 64-byte hot record of a node, as four global 16-byte loads
__device__ __forceinline__ void load_record(const void* rec, float f[16]) {
    const gv4* r = reinterpret_cast<const gv4*>(rec);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const gv4 v = gload4(r + i);
        f[4 * i] = v.x;
        f[4 * i + 1] = v.y;
        f[4 * i + 2] = v.z;
        f[4 * i + 3] = v.w;
    }
}

}  // namespace ikd

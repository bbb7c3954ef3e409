// Emulation stand-in for ikd_gload.h: a plain 16-byte load.
#pragma once
namespace ikd {
typedef float4 gv4;
inline gv4 gload4(const void* p) {
    gv4 v;
    __builtin_memcpy(&v, p, sizeof v);
    return v;
}
}  // namespace ikd
namespace ikd {
inline void load_record(const void* rec, float f[16]) { __builtin_memcpy(f, rec, 64); }
}  // namespace ikd

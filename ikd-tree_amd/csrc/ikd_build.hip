// ikd_build.hip — level-synchronous GPU build of one or more subtrees.
//
// Restates BuildTree (ikd_Tree.cpp:573-622) for the device.  Every segment
// [l, r] of a level is split exactly as the reference splits it:
//   axis = longest extent of the segment (strict '>', ties go to the lower
//          axis, cpp:594-595),
//   node = element of rank mid-l on that axis with mid = (l+r)>>1 (cpp:581,
//          599-614),
//   left = the mid-l smaller elements, right = the rest (cpp:616-617).
// The three coordinate orders are sorted once (stable, so equal keys keep
// their input order); each level then stably partitions all three lists
// around the pivot, so the segment's bbox is always its first and last
// element per list and the rank-(mid-l) element is read directly.  For keys
// without ties on the split axis the resulting tree is the reference's tree;
// with ties the placement of equal keys is deterministic (input order) where
// the reference's depends on libstdc++'s nth_element.
//
// Nodes are written top-down: a segment's node gets its point, axis, size and
// depth, and its subtree AABB (the segment bbox) goes into the parent's child
// box slot.  Children are allocated as sibling pairs in segment order, so
// the layout is deterministic.
#include "ikd_internal.h"

namespace ikd {

__global__ void k_build_keys(const float4* __restrict__ pts, const int* __restrict__ seg_of, int m, int axis,
                             unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    float4 p = pts[i];
    float c = axis == 0 ? p.x : (axis == 1 ? p.y : p.z);
    keys[i] = ((unsigned long long)(uint32_t)seg_of[i] << 32) | float_key(c);
    vals[i] = (uint32_t)i;
}

__global__ void k_build_init_sid(const unsigned long long* __restrict__ k0, int m, int* __restrict__ sid) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    sid[i] = (int)(k0[i] >> 32);
}

__device__ __forceinline__ float coord_of(float4 p, int a) { return a == 0 ? p.x : (a == 1 ? p.y : p.z); }

// Per segment: longest axis, mid, and what it allocates (one pair if it has a
// child; one next-level segment per non-empty side).
__global__ void k_build_seg_a(BuildLevel L, int bound, DevHdr* hdr) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) hdr->seg_count[L.cur ^ 1] = 0;
    if (s >= bound) return;
    int count = hdr->seg_count[L.cur];
    if (s >= count) {
        L.seg_flags[s] = 0ull;
        return;
    }
    int4 sg = L.seg[s];
    int l = sg.x, r = sg.y;
    float ext[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float lo = coord_of(L.pts[L.perm[a][l]], a);
        float hi = coord_of(L.pts[L.perm[a][r]], a);
        ext[a] = hi - lo;
    }
    int axis = 0;
    if (ext[1] > ext[axis]) axis = 1;
    if (ext[2] > ext[axis]) axis = 2;
    int mid = (l + r) >> 1;
    int has_l = mid > l, has_r = mid < r;
    L.seg_axis[s] = axis;
    L.seg_mid[s] = mid;
    unsigned long long need_pair = (has_l || has_r) ? 1ull : 0ull;
    L.seg_flags[s] = need_pair | ((unsigned long long)(has_l + has_r) << 32);
}

__device__ __forceinline__ void write_child_box(NodeHot* hot, int parent, int side, const float* b) {
    float* f = hot[parent].f;
    if (side == 0) {
        *reinterpret_cast<float4*>(f + 4) = make_float4(b[0], b[1], b[2], b[3]);
        *reinterpret_cast<float2*>(f + 8) = make_float2(b[4], b[5]);
    } else {
        *reinterpret_cast<float2*>(f + 10) = make_float2(b[0], b[1]);
        *reinterpret_cast<float4*>(f + 12) = make_float4(b[2], b[3], b[4], b[5]);
    }
}

// Per segment: write the node and emit the child segments.
__global__ void k_build_seg_b(BuildLevel L, int bound, Pool pool, DevHdr* hdr) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= bound) return;
    int count = hdr->seg_count[L.cur];
    if (s >= count) return;
    unsigned long long sc = L.seg_scan[s];
    unsigned long long fl = L.seg_flags[s];
    int pairs_before = (int)(uint32_t)sc;
    int child_before = (int)(sc >> 32);
    int need_pair = (int)(fl & 1ull);
    int nch = (int)(fl >> 32);
    int4 sg = L.seg[s];
    int l = sg.x, r = sg.y, node = sg.z, parent = sg.w;
    int depth = L.seg_depth[s];
    int axis = L.seg_axis[s];
    int mid = L.seg_mid[s];
    int has_l = mid > l, has_r = mid < r;
    int pair = need_pair ? hdr->pair_used + pairs_before : 0;

    float4 P = L.pts[L.perm[axis][mid]];
    uint32_t link = ((uint32_t)pair << LK_PAIR_SHIFT) | (has_l ? LK_HASL : 0u) | (has_r ? LK_HASR : 0u);
    *reinterpret_cast<float4*>(pool.hot[node].f) = make_float4(P.x, P.y, P.z, bitsf(link));
    pool.ids[node] = fbits(P.w);
    pool.par[node] = parent;
    pool.cnt[node] = make_int4(r - l + 1, 0, 0, depth);
    pool.meta[node] = (uint32_t)axis;
    atomicMax(&hdr->max_depth, depth);

    float box[6];
    box[0] = L.pts[L.perm[0][l]].x;
    box[1] = L.pts[L.perm[1][l]].y;
    box[2] = L.pts[L.perm[2][l]].z;
    box[3] = L.pts[L.perm[0][r]].x;
    box[4] = L.pts[L.perm[1][r]].y;
    box[5] = L.pts[L.perm[2][r]].z;
    write_child_box(pool.hot, parent, node & 1, box);

    int c = child_before;
    int2 ch = make_int2(-1, -1);
    if (has_l) {
        L.seg_next[c] = make_int4(l, mid - 1, 2 * pair, node);
        L.seg_depth_next[c] = depth + 1;
        ch.x = c++;
    }
    if (has_r) {
        L.seg_next[c] = make_int4(mid + 1, r, 2 * pair + 1, node);
        L.seg_depth_next[c] = depth + 1;
        ch.y = c++;
    }
    L.seg_child[s] = ch;
    if (s == count - 1) {
        hdr->seg_count[L.cur ^ 1] = child_before + nch;
        hdr->pair_pending = pairs_before + need_pair;
    }
}

__global__ void k_build_side(BuildLevel L, DevHdr* hdr) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j == 0) {
        hdr->pair_used += hdr->pair_pending;
        hdr->pair_pending = 0;
    }
    if (j >= L.m) return;
    int s = L.sid[j];
    if (s < 0) return;
    int a = L.seg_axis[s];
    int mid = L.seg_mid[s];
    uint32_t p = L.perm[a][j];
    L.side[p] = (uint8_t)(j < mid ? 0 : (j == mid ? 2 : 1));
}

__global__ void k_build_flags(BuildLevel L) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int b = blockIdx.y;
    if (j >= L.m) return;
    int s = L.sid[j];
    unsigned long long v = 0ull;
    if (s >= 0) {
        uint8_t sd = L.side[L.perm[b][j]];
        v = sd == 0 ? 1ull : (sd == 1 ? (1ull << 32) : 0ull);
    }
    L.pos_flags[(size_t)b * L.m + j] = v;
}

__global__ void k_build_scatter(BuildLevel L) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int b = blockIdx.y;
    if (j >= L.m) return;
    int s = L.sid[j];
    if (s < 0) {
        if (b == 0) L.sid_next[j] = -1;
        return;
    }
    int l = L.seg[s].x;
    int mid = L.seg_mid[s];
    uint32_t p = L.perm[b][j];
    uint8_t sd = L.side[p];
    if (sd == 2) {
        if (b == 0) L.sid_next[mid] = -1;
        return;
    }
    size_t base = (size_t)b * L.m;
    unsigned long long ex = L.pos_scan[base + j] - L.pos_scan[base + l];
    int np = sd == 0 ? l + (int)(uint32_t)ex : mid + 1 + (int)(ex >> 32);
    L.perm_next[b][np] = p;
    if (b == 0) {
        int2 ch = L.seg_child[s];
        L.sid_next[np] = sd == 0 ? ch.x : ch.y;
    }
}

static inline int blocks_for(long long n, int bs) { return (int)((n + bs - 1) / bs); }

// out[j] = pts[perm[j]]: a level's points in list order, so every segment's
// points are contiguous for the one-workgroup build of the levels below.
// Positions whose point is already a node of an upper level (sid < 0) are
// never written by the partitions and hold no valid index: skip them.
// out[j] = the point at list position j; inv[point] = j (positions already
// placed as nodes are skipped: no kernel wrote their list entries).
__global__ void k_gather_list(const float4* __restrict__ pts, const uint32_t* __restrict__ perm,
                              const int* __restrict__ sid, int m, float4* __restrict__ out, int* __restrict__ inv) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    if (sid[j] < 0) {
        out[j] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const uint32_t p = perm[j];
    out[j] = pts[p];
    if (inv) inv[p] = j;
}
void launch_gather_list(const float4* pts, const uint32_t* perm, const int* sid, int m, float4* out, int* inv,
                        hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_gather_list, dim3(blocks_for(m, 256)), dim3(256), 0, s, pts, perm, sid, m, out, inv);
}

void launch_build_keys(const float4* pts, const int* seg_of, int m, int axis, unsigned long long* keys,
                       uint32_t* vals, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_build_keys, dim3(blocks_for(m, 256)), dim3(256), 0, s, pts, seg_of, m, axis, keys, vals);
}
void launch_build_init_sid(const unsigned long long* k0, int m, int* sid, hipStream_t s) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_build_init_sid, dim3(blocks_for(m, 256)), dim3(256), 0, s, k0, m, sid);
}
void launch_build_seg_a(const BuildLevel& L, int bound, DevHdr* hdr, hipStream_t s) {
    hipLaunchKernelGGL(k_build_seg_a, dim3(blocks_for(bound > 0 ? bound : 1, 256)), dim3(256), 0, s, L, bound, hdr);
}
void launch_build_seg_b(const BuildLevel& L, int bound, Pool pool, DevHdr* hdr, hipStream_t s) {
    if (bound <= 0) return;
    hipLaunchKernelGGL(k_build_seg_b, dim3(blocks_for(bound, 256)), dim3(256), 0, s, L, bound, pool, hdr);
}
void launch_build_side(const BuildLevel& L, DevHdr* hdr, hipStream_t s) {
    hipLaunchKernelGGL(k_build_side, dim3(blocks_for(L.m > 0 ? L.m : 1, 256)), dim3(256), 0, s, L, hdr);
}
void launch_build_flags(const BuildLevel& L, hipStream_t s) {
    if (L.m <= 0) return;
    hipLaunchKernelGGL(k_build_flags, dim3(blocks_for(L.m, 256), 3), dim3(256), 0, s, L);
}
void launch_build_scatter(const BuildLevel& L, hipStream_t s) {
    if (L.m <= 0) return;
    hipLaunchKernelGGL(k_build_scatter, dim3(blocks_for(L.m, 256), 3), dim3(256), 0, s, L);
}

}  // namespace ikd

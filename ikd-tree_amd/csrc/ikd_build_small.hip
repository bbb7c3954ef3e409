// ikd_build_small.hip — one-launch build of many small subtrees.
//
// The same construction as ikd_build.hip (BuildTree, ikd_Tree.cpp:573-622:
// longest axis with ties to the lower axis, the element of rank mid-l with
// mid = (l+r)>>1 as the node, stable presorted partitions so that equal keys
// keep their input order), but one 256-thread workgroup builds one whole
// segment of at most SMALL_MAX points in LDS: a bitonic sort per axis, then
// one level per block-synchronous step.  Segments, pairs and children are
// numbered in the same order as the global build, so a subtree built here is
// the one the level-synchronous global build makes for the same input.
//
// Node allocation is deterministic: segment g takes the pair range that
// starts at pair_used + exclusive_scan(pairs)[g], where pairs(m) is the number
// of nodes with children in the median tree of m points.
#include <algorithm>

#include "ikd_internal.h"

namespace ikd {

// Three shapes: one lane per segment of at most MICRO_MAX points (most insert
// groups hold one to a few points), a 64-thread workgroup for segments of at
// most TINY_MAX points, a 256-thread one for the rest up to SMALL_MAX.  The
// lane kernel also lists the larger segments for the two workgroup kernels,
// which loop over their list, so no workgroup is launched only to find its
// segment belongs to the other shape.
constexpr int MICRO_MAX = 8;
constexpr int TINY_MAX = 128;
constexpr int SMALL_SB = 256;  // threads of the workgroup build of segments above TINY_MAX points

// Nodes with at least one child in the median-split tree of m points.  A node
// of size x >= 2 leaves children of sizes (x-1)>>1 and x-1-((x-1)>>1), so a
// level holds at most two distinct sizes.
__device__ int median_tree_pairs(int m) {
    int sz[2] = {m, -1};
    long long cnt[2] = {1, 0};
    long long internal = 0;
    while (cnt[0] > 0 || cnt[1] > 0) {
        int nsz[2] = {-1, -1};
        long long ncnt[2] = {0, 0};
        for (int i = 0; i < 2; ++i) {
            if (cnt[i] == 0 || sz[i] < 2) continue;
            internal += cnt[i];
            int parts[2] = {(sz[i] - 1) >> 1, sz[i] - 1 - ((sz[i] - 1) >> 1)};
            for (int p = 0; p < 2; ++p) {
                int x = parts[p];
                if (x < 2) continue;
                int slot = (nsz[0] == x || nsz[0] < 0) ? 0 : 1;
                nsz[slot] = x;
                ncnt[slot] += cnt[i];
            }
        }
        sz[0] = nsz[0];
        sz[1] = nsz[1];
        cnt[0] = ncnt[0];
        cnt[1] = ncnt[1];
    }
    return (int)internal;
}

// count_dev (optional) holds the number of live segments among the S slots;
// slots past it take no pairs.
__global__ void k_small_pairs(const int4* __restrict__ seg, int S, const int* __restrict__ count_dev,
                              int* __restrict__ pairs, DevHdr* hdr) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g == 0) {  // the size-class lists of this build start empty
        hdr->small_lists[0] = 0;
        hdr->small_lists[1] = 0;
    }
    if (g >= S) return;
    const int count = count_dev ? *count_dev : S;
    pairs[g] = g < count ? median_tree_pairs(seg[g].y - seg[g].x + 1) : 0;
}

__global__ void k_small_commit(const int* __restrict__ pairs, const int* __restrict__ scan, int S, DevHdr* hdr) {
    hdr->pair_used += scan[S - 1] + pairs[S - 1];
}

__device__ __forceinline__ void write_child_box_s(NodeHot* hot, int parent, int side, const float* b) {
    float* f = hot[parent].f;
    if (side == 0) {
        *reinterpret_cast<float4*>(f + 4) = make_float4(b[0], b[1], b[2], b[3]);
        *reinterpret_cast<float2*>(f + 8) = make_float2(b[4], b[5]);
    } else {
        *reinterpret_cast<float2*>(f + 10) = make_float2(b[0], b[1]);
        *reinterpret_cast<float4*>(f + 12) = make_float4(b[2], b[3], b[4], b[5]);
    }
}

// Exclusive scan of one 64-bit value per thread across the workgroup
// (*total gets the sum): an inclusive scan inside each wavefront by
// shuffles, then the wavefront totals through LDS (two barriers).  Several
// small counters can ride in one value as bit fields that never overflow.
template <int SB>
__device__ unsigned long long block_exscan_u64(unsigned long long v, unsigned long long* wtot,
                                               unsigned long long* total) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    unsigned long long x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int src = (lane - off) & 63;  // lanes below off ignore what they read
        const unsigned lo = (unsigned)__shfl((int)(unsigned)x, src);
        const unsigned hi = (unsigned)__shfl((int)(unsigned)(x >> 32), src);
        const unsigned long long y = ((unsigned long long)hi << 32) | lo;
        if (lane >= off) x += y;
    }
    if (lane == 63) wtot[w] = x;
    __syncthreads();
    unsigned long long base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < SB / 64; ++i) {
        if (i < w) base += wtot[i];
        tot += wtot[i];
    }
    *total = tot;
    __syncthreads();
    return base + x - v;
}

// Ascending bitonic sort of P keys (P a power of two, at most E * SB) held in
// registers: key i lives in x[i / SB] of thread i % SB.  Stages whose partner
// is in the same wavefront exchange by shuffles, stages across the thread's
// own registers by swaps, and only the stages across wavefronts go through
// LDS (xch, SB * E entries).  Keys must be distinct.
template <int SB, int E>
__device__ void reg_bitonic(unsigned long long (&x)[E], int P, unsigned long long* xch) {
    const int t = threadIdx.x;
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= SB) {
                const int ej = j / SB;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int e2 = e ^ ej;
                    if (e2 <= e || e2 >= E) continue;
                    const bool up = ((e * SB + t) & k) == 0;
                    const unsigned long long a = x[e], b = x[e2];
                    if ((a > b) == up) {
                        x[e] = b;
                        x[e2] = a;
                    }
                }
            } else if (j >= 64) {
                __syncthreads();
#pragma unroll
                for (int e = 0; e < E; ++e) xch[e * SB + t] = x[e];
                __syncthreads();
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int i = e * SB + t;
                    const unsigned long long y = xch[i ^ j];
                    const bool keep_min = ((i & j) == 0) == ((i & k) == 0);
                    x[e] = keep_min ? (x[e] < y ? x[e] : y) : (x[e] > y ? x[e] : y);
                }
            } else {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const int i = e * SB + t;
                    const unsigned lo = (unsigned)__shfl_xor((int)(unsigned)x[e], j);
                    const unsigned hi = (unsigned)__shfl_xor((int)(unsigned)(x[e] >> 32), j);
                    const unsigned long long y = ((unsigned long long)hi << 32) | lo;
                    const bool keep_min = ((i & j) == 0) == ((i & k) == 0);
                    x[e] = keep_min ? (x[e] < y ? x[e] : y) : (x[e] > y ? x[e] : y);
                }
            }
        }
    }
    __syncthreads();  // xch may be reused at once
}

// orig (optional) gives each point's index in the caller's original input:
// equal coordinates are then ordered by it, as the global build's stable
// presorted lists order them; without it, by position in the segment.
template <int SB, int SMAX>
__global__ __launch_bounds__(SB) void k_build_small(const float4* __restrict__ pts, const int4* __restrict__ seg,
                                                     const int* __restrict__ seg_depth,
                                                     const int* __restrict__ pair_scan,
                                                     const int* __restrict__ list,
                                                     const int* __restrict__ list_count,
                                                     const uint32_t* __restrict__ orig, SmallLists pre, Pool pool,
                                                     DevHdr* hdr) {
    constexpr int SMAX_SEG = SMAX / 2 + 1;
    __shared__ float px[SMAX], py[SMAX], pz[SMAX];
    __shared__ unsigned long long skey[SMAX];  // sort keys, then running counts
    __shared__ unsigned short lst[2][3][SMAX];
    __shared__ short sid[2][SMAX];
    __shared__ unsigned char side[SMAX];
    __shared__ short seg_l[2][SMAX_SEG], seg_r[2][SMAX_SEG];
    __shared__ int seg_node[2][SMAX_SEG], seg_par[2][SMAX_SEG];
    __shared__ unsigned char seg_axis[SMAX_SEG];
    __shared__ short seg_mid[SMAX_SEG], seg_cl[SMAX_SEG], seg_cr[SMAX_SEG];
    __shared__ unsigned long long wtot[SB / 64];
    __shared__ unsigned short rnk[SMAX];

    const int t = threadIdx.x;
    const int nlist = *list_count;
    for (int li = blockIdx.x; li < nlist; li += gridDim.x) {
    const int g = list[li];
    const int4 sg = seg[g];
    const int l0 = sg.x;
    const int m = sg.y - sg.x + 1;
    // the list holds this shape's size class; the host guarantees m <= SMALL_MAX
    if (m <= 0 || m > SMAX) continue;
    const int depth0 = seg_depth[g];
    int next_pair = hdr->pair_used + pair_scan[g];

    for (int i = t; i < m; i += SB) {
        float4 p = pts[l0 + i];
        px[i] = p.x;
        py[i] = p.y;
        pz[i] = p.z;
    }
    int P = 1;
    while (P < m) P <<= 1;
    if (pre.inv0) {
        // the caller's lists are already sorted (coordinate, original index)
        // per segment, and the points were gathered in x order
        for (int i = t; i < m; i += SB) {
            lst[0][0][i] = (unsigned short)i;
            lst[0][1][i] = (unsigned short)(pre.inv0[pre.list1[l0 + i]] - l0);
            lst[0][2][i] = (unsigned short)(pre.inv0[pre.list2[l0 + i]] - l0);
        }
    } else {
    // tie-break rank of every point: its order by original index, or its position
    constexpr int E = SMAX / SB;
    unsigned long long x[E];
    if (orig) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = e * SB + t;
            x[e] = i < m ? (((unsigned long long)orig[l0 + i] << 16) | (unsigned)i) : ~0ull;
        }
        reg_bitonic<SB, E>(x, P, skey);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = e * SB + t;
            if (i < m) rnk[x[e] & 0xffffu] = (unsigned short)i;
        }
    } else {
        for (int i = t; i < m; i += SB) rnk[i] = (unsigned short)i;
    }
    __syncthreads();
    // three stable sorted index lists: key (coordinate, rank, position)
    for (int a = 0; a < 3; ++a) {
        const float* c = a == 0 ? px : (a == 1 ? py : pz);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = e * SB + t;
            x[e] = i < m ? (((unsigned long long)float_key(c[i]) << 32) | ((unsigned)rnk[i] << 16) | (unsigned)i)
                         : ~0ull;
        }
        reg_bitonic<SB, E>(x, P, skey);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = e * SB + t;
            if (i < m) lst[0][a][i] = (unsigned short)(x[e] & 0xffffu);
        }
    }
    __syncthreads();
    }
    for (int i = t; i < m; i += SB) sid[0][i] = 0;
    if (t == 0) {
        seg_l[0][0] = 0;
        seg_r[0][0] = (short)(m - 1);
        seg_node[0][0] = sg.z;
        seg_par[0][0] = sg.w;
    }
    __syncthreads();
    unsigned long long* run = skey;  // the sort keys' space, reused for running counts
    const int chunk = (m + SB - 1) / SB;
    int ns = 1;
    int cur = 0;
    for (int level = 0; ns > 0; ++level) {
        const int depth = depth0 + level;
        const int per = (ns + SB - 1) / SB;  // contiguous segments per thread keep the order
        int my_pairs = 0, my_children = 0;
        for (int q = 0; q < per; ++q) {
            const int s = t * per + q;
            if (s >= ns) break;
            const int l = seg_l[cur][s], r = seg_r[cur][s];
            float ext[3];
            ext[0] = px[lst[cur][0][r]] - px[lst[cur][0][l]];
            ext[1] = py[lst[cur][1][r]] - py[lst[cur][1][l]];
            ext[2] = pz[lst[cur][2][r]] - pz[lst[cur][2][l]];
            int axis = 0;
            if (ext[1] > ext[axis]) axis = 1;
            if (ext[2] > ext[axis]) axis = 2;
            const int mid = (l + r) >> 1;
            seg_axis[s] = (unsigned char)axis;
            seg_mid[s] = (short)mid;
            const int hl = mid > l, hr = mid < r;
            my_pairs += (hl || hr) ? 1 : 0;
            my_children += hl + hr;
        }
        unsigned long long tot2;
        const unsigned long long off2 =
            block_exscan_u64<SB>((unsigned long long)my_pairs | ((unsigned long long)my_children << 32), wtot, &tot2);
        int pair_off = (int)(unsigned)off2, child_off = (int)(off2 >> 32);
        const int tot_pairs = (int)(unsigned)tot2, tot_children = (int)(tot2 >> 32);
        for (int q = 0; q < per; ++q) {
            const int s = t * per + q;
            if (s >= ns) break;
            const int l = seg_l[cur][s], r = seg_r[cur][s], mid = seg_mid[s], axis = seg_axis[s];
            const int node = seg_node[cur][s], parent = seg_par[cur][s];
            const int hl = mid > l, hr = mid < r;
            const int pair = (hl || hr) ? next_pair + pair_off++ : 0;
            const int piv = lst[cur][axis][mid];
            const uint32_t link = ((uint32_t)pair << LK_PAIR_SHIFT) | (hl ? LK_HASL : 0u) | (hr ? LK_HASR : 0u);
            const float4 P4 = pts[l0 + piv];
            *reinterpret_cast<float4*>(pool.hot[node].f) = make_float4(P4.x, P4.y, P4.z, bitsf(link));
            pool.ids[node] = fbits(P4.w);
            pool.par[node] = parent;
            pool.cnt[node] = make_int4(r - l + 1, 0, 0, depth);
            pool.meta[node] = (uint32_t)axis;
            float box[6] = {px[lst[cur][0][l]], py[lst[cur][1][l]], pz[lst[cur][2][l]],
                            px[lst[cur][0][r]], py[lst[cur][1][r]], pz[lst[cur][2][r]]};
            write_child_box_s(pool.hot, parent, node & 1, box);
            int cl = -1, cr = -1;
            if (hl) {
                cl = child_off++;
                seg_l[cur ^ 1][cl] = (short)l;
                seg_r[cur ^ 1][cl] = (short)(mid - 1);
                seg_node[cur ^ 1][cl] = 2 * pair;
                seg_par[cur ^ 1][cl] = node;
            }
            if (hr) {
                cr = child_off++;
                seg_l[cur ^ 1][cr] = (short)(mid + 1);
                seg_r[cur ^ 1][cr] = (short)r;
                seg_node[cur ^ 1][cr] = 2 * pair + 1;
                seg_par[cur ^ 1][cr] = node;
            }
            seg_cl[s] = (short)cl;
            seg_cr[s] = (short)cr;
        }
        if (t == 0) atomicMax(&hdr->max_depth, depth);
        next_pair += tot_pairs;
        __syncthreads();
        if (tot_children == 0) break;
        // side of every live position, read off its segment's axis list
        for (int j = t; j < m; j += SB) {
            const int s = sid[cur][j];
            if (s < 0) continue;
            const int mid = seg_mid[s];
            side[lst[cur][seg_axis[s]][j]] = (unsigned char)(j < mid ? 0 : (j == mid ? 2 : 1));
        }
        __syncthreads();
        // stable partition of the three lists at once: for every position, the
        // numbers of left and right entries of each list before it, as six
        // 10-bit fields of one value (no field reaches 1024: at most m - 1
        // entries move), so one scan serves all three lists
        unsigned long long cnt6 = 0;
        for (int q = 0; q < chunk; ++q) {
            const int j = t * chunk + q;
            if (j >= m) break;
            if (sid[cur][j] < 0) continue;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const unsigned char sd = side[lst[cur][b][j]];
                if (sd < 2) cnt6 += 1ull << (20 * b + 10 * sd);
            }
        }
        unsigned long long tot6;
        unsigned long long acc = block_exscan_u64<SB>(cnt6, wtot, &tot6);
        for (int q = 0; q < chunk; ++q) {
            const int j = t * chunk + q;
            if (j >= m) break;
            run[j] = acc;
            if (sid[cur][j] < 0) continue;
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const unsigned char sd = side[lst[cur][b][j]];
                if (sd < 2) acc += 1ull << (20 * b + 10 * sd);
            }
        }
        __syncthreads();
        for (int b = 0; b < 3; ++b) {
            for (int q = 0; q < chunk; ++q) {
                const int j = t * chunk + q;
                if (j >= m) break;
                const int s = sid[cur][j];
                if (s < 0) {
                    if (b == 0) sid[cur ^ 1][j] = -1;
                    continue;
                }
                const int p = lst[cur][b][j];
                const unsigned char sd = side[p];
                if (sd == 2) {
                    if (b == 0) sid[cur ^ 1][seg_mid[s]] = -1;
                    continue;
                }
                const int l = seg_l[cur][s], mid = seg_mid[s];
                // fields only grow along the list, so the packed difference
                // subtracts field by field without borrows
                const int before = (int)(((run[j] - run[l]) >> (20 * b + 10 * sd)) & 1023ull);
                const int np = sd == 0 ? l + before : mid + 1 + before;
                lst[cur ^ 1][b][np] = (unsigned short)p;
                if (b == 0) sid[cur ^ 1][np] = sd == 0 ? seg_cl[s] : seg_cr[s];
            }
        }
        __syncthreads();
        ns = tot_children;
        cur ^= 1;
    }
    __syncthreads();  // the next segment of the list reuses the LDS arrays
    }
}

// One lane builds one segment of at most MICRO_MAX points: the median tree of
// k_build_small (axis of the longest extent, ties to the lower axis; node =
// the element of rank (n-1)>>1 in (coordinate, position) order; both sides
// keep the segment's order).  The lane keeps the work order as 4-bit
// positions packed in one register and its pending ranges in LDS (nothing
// indexed at run time lives in private memory); pairs are taken from the
// segment's range one internal node at a time.  Larger segments (or every
// segment when micro is off) go to the size-class lists.
constexpr int MICRO_BS = 256;
constexpr int MICRO_STACK = 6;  // pending ranges: at most one per level below the current one
__device__ __forceinline__ float micro_coord(const float4& q, int axis) {
    return axis == 0 ? q.x : (axis == 1 ? q.y : q.z);
}
__global__ __launch_bounds__(MICRO_BS) void k_build_micro(const float4* __restrict__ pts, const int4* __restrict__ seg,
                                                          const int* __restrict__ seg_depth,
                                                          const int* __restrict__ pair_scan, int S,
                                                          const int* __restrict__ count_dev, int micro, Pool pool,
                                                          DevHdr* hdr, int* __restrict__ tiny_list,
                                                          int* __restrict__ small_list) {
    // per lane: (lo | hi << 4 | level << 8), slot node, slot parent
    __shared__ int stk_r[MICRO_STACK][MICRO_BS], stk_node[MICRO_STACK][MICRO_BS], stk_par[MICRO_STACK][MICRO_BS];
    const int tid = threadIdx.x;
    const int g = blockIdx.x * blockDim.x + tid;
    if (g >= S || (count_dev && g >= *count_dev)) return;
    const int4 sg = seg[g];
    const int m = sg.y - sg.x + 1;
    if (m <= 0) return;
    if (!micro || m > MICRO_MAX) {
        if (m <= TINY_MAX) tiny_list[atomicAdd(&hdr->small_lists[0], 1)] = g;
        else small_list[atomicAdd(&hdr->small_lists[1], 1)] = g;
        return;
    }
    const float4* P = pts + sg.x;
    uint64_t perm = 0x76543210ull;  // nibble i: the position at work-order slot i
    int next_pair = hdr->pair_used + pair_scan[g];
    const int depth0 = seg_depth[g];
    int nw = 1;
    stk_r[0][tid] = m << 4;
    stk_node[0][tid] = sg.z;
    stk_par[0][tid] = sg.w;
    int deepest = depth0;
    while (nw > 0) {
        --nw;
        const int rr = stk_r[nw][tid];
        const int lo = rr & 15, hi = (rr >> 4) & 15, depth = depth0 + (rr >> 8);
        const int node = stk_node[nw][tid], parent = stk_par[nw][tid];
        const int cnt = hi - lo;
        // bounds in key order, as the sorted lists give them (-0 below +0)
        const float4 q0 = P[(perm >> (4 * lo)) & 15];
        float mn[3] = {q0.x, q0.y, q0.z}, mx[3] = {q0.x, q0.y, q0.z};
        for (int i = lo + 1; i < hi; ++i) {
            const float4 q = P[(perm >> (4 * i)) & 15];
            const float c[3] = {q.x, q.y, q.z};
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (float_key(c[a]) < float_key(mn[a])) mn[a] = c[a];
                if (float_key(c[a]) > float_key(mx[a])) mx[a] = c[a];
            }
        }
        const float ext[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
        int axis = 0;
        if (ext[1] > ext[axis]) axis = 1;
        if (ext[2] > ext[axis]) axis = 2;
        // the element of rank (cnt-1)>>1 in (coordinate, position) order
        const int want = (cnt - 1) >> 1;
        int pos = -1;
        uint32_t kp = 0;
        for (int i = lo; i < hi; ++i) {
            const int pi = (int)((perm >> (4 * i)) & 15);
            const uint32_t ka = float_key(micro_coord(P[pi], axis));
            int rank = 0;
            for (int j = lo; j < hi; ++j) {
                const int pj = (int)((perm >> (4 * j)) & 15);
                const uint32_t kb = float_key(micro_coord(P[pj], axis));
                rank += (kb < ka || (kb == ka && pj < pi)) ? 1 : 0;
            }
            if (rank == want) {
                pos = pi;
                kp = ka;
            }
        }
        // stable split: ranks below the pivot's first, then the pivot, then the ones above
        uint64_t left = 0, right = 0;
        int nl = 0, nr = 0;
        for (int i = lo; i < hi; ++i) {
            const int pi = (int)((perm >> (4 * i)) & 15);
            if (pi == pos) continue;
            const uint32_t kb = float_key(micro_coord(P[pi], axis));
            if (kb < kp || (kb == kp && pi < pos)) left |= (uint64_t)pi << (4 * nl++);
            else right |= (uint64_t)pi << (4 * nr++);
        }
        const uint64_t run = left | ((uint64_t)pos << (4 * nl)) | (right << (4 * (nl + 1)));
        const uint64_t mask = ((1ull << (4 * cnt)) - 1) << (4 * lo);
        perm = (perm & ~mask) | (run << (4 * lo));
        const int hl = nl > 0, hr = nr > 0;
        const int pair = (hl || hr) ? next_pair++ : 0;
        const uint32_t link = ((uint32_t)pair << LK_PAIR_SHIFT) | (hl ? LK_HASL : 0u) | (hr ? LK_HASR : 0u);
        const float4 P4 = P[pos];
        *reinterpret_cast<float4*>(pool.hot[node].f) = make_float4(P4.x, P4.y, P4.z, bitsf(link));
        pool.ids[node] = fbits(P4.w);
        pool.par[node] = parent;
        pool.cnt[node] = make_int4(cnt, 0, 0, depth);
        pool.meta[node] = (uint32_t)axis;
        const float box[6] = {mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]};
        write_child_box_s(pool.hot, parent, node & 1, box);
        deepest = depth > deepest ? depth : deepest;
        const int lv = (rr >> 8) + 1;
        if (hr) {
            stk_r[nw][tid] = (lo + nl + 1) | (hi << 4) | (lv << 8);
            stk_node[nw][tid] = 2 * pair + 1;
            stk_par[nw][tid] = node;
            ++nw;
        }
        if (hl) {
            stk_r[nw][tid] = lo | ((lo + nl) << 4) | (lv << 8);
            stk_node[nw][tid] = 2 * pair;
            stk_par[nw][tid] = node;
            ++nw;
        }
    }
    atomicMax(&hdr->max_depth, deepest);
}

static inline int blocks_for(long long n, int bs) { return (int)((n + bs - 1) / bs); }

void launch_small_pairs(const int4* seg, int S, const int* count_dev, int* pairs, DevHdr* hdr, hipStream_t s) {
    if (S <= 0) return;
    hipLaunchKernelGGL(k_small_pairs, dim3(blocks_for(S, 256)), dim3(256), 0, s, seg, S, count_dev, pairs, hdr);
}
void launch_build_small(const float4* pts, const int4* seg, const int* seg_depth, const int* pair_scan, int S,
                        const int* count_dev, const uint32_t* orig, SmallLists pre, int* lists, Pool pool,
                        DevHdr* hdr, hipStream_t s) {
    if (S <= 0) return;
    // the lane build follows the position tie rule only: presorted or
    // original-order inputs go to the workgroup builds
    const int micro = (!orig && !pre.inv0) ? 1 : 0;
    hipLaunchKernelGGL(k_build_micro, dim3(blocks_for(S, MICRO_BS)), dim3(MICRO_BS), 0, s, pts, seg, seg_depth, pair_scan, S,
                       count_dev, micro, pool, hdr, lists, lists + S);
    // each workgroup loops over its shape's list; the grids only need to be
    // wide enough to fill the device when a list is long
    hipLaunchKernelGGL((k_build_small<64, TINY_MAX>), dim3(std::min(S, 2048)), dim3(64), 0, s, pts, seg, seg_depth,
                       pair_scan, lists, &hdr->small_lists[0], orig, pre, pool, hdr);
    hipLaunchKernelGGL((k_build_small<SMALL_SB, SMALL_MAX>), dim3(std::min(S, 512)), dim3(SMALL_SB), 0, s, pts, seg,
                       seg_depth, pair_scan, lists + S, &hdr->small_lists[1], orig, pre, pool, hdr);
}
void launch_small_commit(const int* pairs, const int* pair_scan, int S, DevHdr* hdr, hipStream_t s) {
    if (S <= 0) return;
    hipLaunchKernelGGL(k_small_commit, dim3(1), dim3(1), 0, s, pairs, pair_scan, S, hdr);
}

}  // namespace ikd

// ikd_query.hip — read-only traversals: batched kNN (Nearest_Search), box
// search (Box_Search) and radius search (Radius_Search).
//
// kNN (ikd_Tree.cpp:366-397, 868-1013).  One lane per query, 64 queries per
// wavefront, a per-lane explicit stack in LDS, and the best-k list in
// registers.  The traversal visits a node only if it is not tree_deleted and
// its box distance is within max_dist (cpp:870-873), offers the node's own
// point when it is live (cpp:885-892), and descends into the nearer child
// first (left on ties, cpp:896) while the heap is not full or the child's box
// distance can still improve the k-th candidate.
//
// Selection order.  Candidates are ranked by the total order
// (d2, x, y, z, id), and a child is pruned only when its box distance is
// strictly greater than the k-th d2.  The result is therefore the k smallest
// live points under that order, whatever the tree's shape: identical for any
// layout, batch composition or launch configuration.  Against the reference
// the squared distances agree bit for bit; only which of several points tied
// at the k-th distance is returned can differ (its choice follows its
// traversal order).  Output is ascending in the same order, which on exact
// distance ties is the reference comparator's x order (ikd_Tree.h:95-106).
//
// max_dist: the reference compares in double (cpp:872, 887).  The host
// passes thr = the largest float <= max_dist*max_dist, for which
// (double)d <= max_dist*max_dist  <=>  d <= thr  for every float d.
#include "ikd_internal.h"

namespace ikd {

static inline int blocks_for(long long n, int bs) { return (int)((n + bs - 1) / bs); }

// a strictly precedes b in (d2, x, y, z, id) order, given equal d2
__device__ __noinline__ bool tie_less(const NodeHot* __restrict__ hot, const uint32_t* __restrict__ ids, int a, int b) {
    const float* fa = hot[a].f;
    const float* fb = hot[b].f;
    if (fa[0] != fb[0]) return fa[0] < fb[0];
    if (fa[1] != fb[1]) return fa[1] < fb[1];
    if (fa[2] != fb[2]) return fa[2] < fb[2];
    return ids[a] < ids[b];
}

// entry (ed, en) ranks after candidate (d, n)?  Empty slots (+inf, -1) rank
// after everything, padding slots (-inf, -2) before everything.
__device__ __forceinline__ bool after(const NodeHot* hot, const uint32_t* ids, float ed, int en, float d, int n) {
    if (ed != d) return ed > d;
    if (en < 0) return en == -1;
    return tie_less(hot, ids, n, en);
}

template <int KM, int STACK>
__global__ __launch_bounds__(64) void k_knn(KnnArgs a) {
    __shared__ int st_n[STACK][64];
    __shared__ float st_d[STACK][64];
    const int lane = threadIdx.x;
    const int gq = blockIdx.x * 64 + lane;
    if (gq >= a.nq) return;
    const int qi = a.order ? (int)a.order[gq] : gq;
    const float qx = a.q[3 * qi], qy = a.q[3 * qi + 1], qz = a.q[3 * qi + 2];
    const NodeHot* __restrict__ hot = a.hot;
    const float thr = a.thr;
    const int k = a.k;

    // descending best list: slot 0 is the current k-th candidate
    float bd_[KM];
    int bn_[KM];
#pragma unroll
    for (int i = 0; i < KM; ++i) {
        bd_[i] = i < k ? INFINITY : -INFINITY;
        bn_[i] = i < k ? -1 : -2;
    }
    unsigned int visits = 0, boxes = 0;

    const uint32_t lk0 = fbits(hot[SENTINEL].f[3]);
    int sp = 0;
    int n = -1;
    float bd = 0.0f;
    if ((lk0 & LK_HASL) && !(lk0 & LK_TDL)) {
        n = ROOT;
        bd = calc_box_dist(hot[SENTINEL].f + 4, qx, qy, qz);
    }
    while (true) {
        if ((int)visits > a.max_steps) {
            atomicOr(a.overflow, 2);
            break;
        }
        if (n >= 0 && bd <= thr) {
            const float4* rec = reinterpret_cast<const float4*>(hot[n].f);
            const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3];
            const uint32_t lk = fbits(r0.w);
            ++visits;
            if (!(lk & LK_PDEL)) {
                float d = calc_dist(qx, qy, qz, r0.x, r0.y, r0.z);
                if (d <= thr && after(hot, a.ids, bd_[0], bn_[0], d, n)) {
                    // drop slot 0, insert (d, n) keeping descending order
                    bool gnext[KM];
#pragma unroll
                    for (int i = 0; i < KM; ++i) gnext[i] = (i + 1 < KM) ? after(hot, a.ids, bd_[i + 1], bn_[i + 1], d, n) : false;
#pragma unroll
                    for (int i = 0; i < KM; ++i) {
                        bool gcur = (i == 0) ? true : gnext[i - 1];
                        if (gnext[i]) {
                            bd_[i] = bd_[i + 1 < KM ? i + 1 : i];
                            bn_[i] = bn_[i + 1 < KM ? i + 1 : i];
                        } else if (gcur) {
                            bd_[i] = d;
                            bn_[i] = n;
                        }
                    }
                }
            }
            const int pair = (int)lk_pair(lk);
            const bool hl = (lk & LK_HASL) != 0, hr = (lk & LK_HASR) != 0;
            boxes += (hl ? 1u : 0u) + (hr ? 1u : 0u);
            const bool el = hl && !(lk & LK_TDL), er = hr && !(lk & LK_TDR);
            float bl[6] = {r1.x, r1.y, r1.z, r1.w, r2.x, r2.y};
            float br[6] = {r2.z, r2.w, r3.x, r3.y, r3.z, r3.w};
            const float dl = el ? calc_box_dist(bl, qx, qy, qz) : INFINITY;
            const float dr = er ? calc_box_dist(br, qx, qy, qz) : INFINITY;
            if (el || er) {
                const bool near_left = el && (!er || dl <= dr);
                const int nn = near_left ? 2 * pair : 2 * pair + 1;
                const float dn = near_left ? dl : dr;
                const bool has_far = near_left ? er : el;
                const int fn = near_left ? 2 * pair + 1 : 2 * pair;
                const float df = near_left ? dr : dl;
                const float top = bd_[0];
                if (has_far && df <= top && df <= thr) {
                    if (sp < STACK) {
                        st_n[sp][lane] = fn;
                        st_d[sp][lane] = df;
                        ++sp;
                    } else {
                        atomicOr(a.overflow, 1);
                    }
                }
                if (dn <= top && dn <= thr) {
                    n = nn;
                    bd = dn;
                    continue;
                }
            }
        }
        // backtrack to the deepest deferred child that can still improve
        n = -1;
        while (sp > 0) {
            --sp;
            float d = st_d[sp][lane];
            if (d <= bd_[0]) {
                n = st_n[sp][lane];
                bd = d;
                break;
            }
        }
        if (n < 0) break;
    }

    // ascending output: real entries sit in slots [k-found, k-1]
    int found = 0;
#pragma unroll
    for (int i = 0; i < KM; ++i) found += bn_[i] >= 0 ? 1 : 0;
    a.out_counts[qi] = found;
#pragma unroll
    for (int i = 0; i < KM; ++i) {
        if (i >= k) break;
        const int j = k - 1 - i;
        const long long o = (long long)qi * k + j;
        const int node = bn_[i];
        if (node >= 0) {
            const float* f = hot[node].f;
            a.out_xyz[3 * o] = f[0];
            a.out_xyz[3 * o + 1] = f[1];
            a.out_xyz[3 * o + 2] = f[2];
            a.out_d2[o] = bd_[i];
            a.out_ids[o] = a.ids[node];
        } else {
            a.out_xyz[3 * o] = 0.0f;
            a.out_xyz[3 * o + 1] = 0.0f;
            a.out_xyz[3 * o + 2] = 0.0f;
            a.out_d2[o] = INFINITY;
            a.out_ids[o] = 0xffffffffu;
        }
    }
    if (a.visits) a.visits[qi] = visits;
    if (a.boxes) a.boxes[qi] = boxes;
}

template <int KM>
static void launch_knn_km(const KnnArgs& a, int stack, hipStream_t s) {
    dim3 grid(blocks_for(a.nq, 64)), block(64);
    if (stack <= 32) hipLaunchKernelGGL((k_knn<KM, 32>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((k_knn<KM, 64>), grid, block, 0, s, a);
}

// kmax_bucket unused: the bucket is chosen from k; stack is the LDS stack depth.
int launch_knn(const KnnArgs& a, int stack, int block, hipStream_t s) {
    (void)block;
    if (a.nq <= 0) return 0;
    int k = a.k;
    if (k <= 1) launch_knn_km<1>(a, stack, s);
    else if (k <= 2) launch_knn_km<2>(a, stack, s);
    else if (k <= 4) launch_knn_km<4>(a, stack, s);
    else if (k <= 5) launch_knn_km<5>(a, stack, s);
    else if (k <= 8) launch_knn_km<8>(a, stack, s);
    else if (k <= 10) launch_knn_km<10>(a, stack, s);
    else if (k <= 16) launch_knn_km<16>(a, stack, s);
    else if (k <= 20) launch_knn_km<20>(a, stack, s);
    else if (k <= 32) launch_knn_km<32>(a, stack, s);
    else if (k <= 64) launch_knn_km<64>(a, stack, s);
    else return -1;
    return 0;
}

// ------------------------------------------------------------- Morton order
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_morton(const float* __restrict__ q, int nq, const NodeHot* __restrict__ hot,
                         unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const float* b = hot[SENTINEL].f + 4;  // root AABB
    uint32_t code = 0;
    float c[3] = {q[3 * i], q[3 * i + 1], q[3 * i + 2]};
    for (int a = 0; a < 3; ++a) {
        float ext = b[3 + a] - b[a];
        float t = ext > 0.0f ? (c[a] - b[a]) / ext : 0.0f;
        t = fminf(fmaxf(t, 0.0f), 1.0f);
        uint32_t u = (uint32_t)(t * 1023.0f);
        code |= spread10(u) << (2 - a);
    }
    keys[i] = code;
    vals[i] = (uint32_t)i;
}

void launch_morton(const float* q, int nq, const NodeHot* hot, unsigned long long* keys, uint32_t* vals,
                   hipStream_t s) {
    if (nq <= 0) return;
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(nq, 256)), dim3(256), 0, s, q, nq, hot, keys, vals);
}

// ------------------------------------------------------- box / radius search
// Search_by_range (cpp:1015-1044) and Search_by_radius (cpp:1046-1087).  One
// lane per query; a child subtree is classified from its AABB as pruned,
// wholly inside (flattened without further tests, as the reference calls
// flatten) or partial.  Pass 0 counts (a whole subtree contributes
// TreeSize - invalid_point_num without being walked), pass 1 writes.
//
// Radius classification is the reference's float arithmetic: centre
// (lo+hi)*0.5, dist = sqrtf(calc_dist(centre, q)), rs = sqrtf(radius_sq);
// pruned if dist > r + rs, whole if dist <= r - rs; a node's own point is
// taken when calc_dist(p, q) <= r*r.
__device__ __forceinline__ int classify(const RangeArgs& a, int qi, const float* b, float qx, float qy, float qz,
                                        float r) {
    if (a.mode == 0) {
        const float* box = a.boxes + 6 * qi;
        if (box_disjoint(box, b)) return 0;
        return box_covers(box, b) ? 1 : 2;
    }
    float cx = (float)((double)(b[0] + b[3]) * 0.5);
    float cy = (float)((double)(b[1] + b[4]) * 0.5);
    float cz = (float)((double)(b[2] + b[5]) * 0.5);
    float dist = sqrtf(calc_dist(cx, cy, cz, qx, qy, qz));
    float rs = sqrtf(box_radius_sq(b));
    if (dist > r + rs) return 0;
    if (dist <= r - rs) return 1;
    return 2;
}

__global__ void k_range(RangeArgs a, int fill) {
    int qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= a.nq) return;
    const NodeHot* __restrict__ hot = a.pool.hot;
    float qx = 0, qy = 0, qz = 0, r = 0;
    if (a.mode == 1) {
        qx = a.centers[3 * qi];
        qy = a.centers[3 * qi + 1];
        qz = a.centers[3 * qi + 2];
        r = a.radii[qi];
    }
    long long total = 0;
    long long pos = fill ? a.offsets[qi] : 0;
    long long lim_end = fill ? a.offsets[qi + 1] : 0;
    auto emit = [&](int u, const float* f) {
        if (fill) {
            if (pos >= lim_end) {
                atomicOr(a.overflow, 8);
                return;
            }
            a.out_xyz[3 * pos] = f[0];
            a.out_xyz[3 * pos + 1] = f[1];
            a.out_xyz[3 * pos + 2] = f[2];
            a.out_ids[pos] = a.pool.ids[u];
            ++pos;
        } else {
            ++total;
        }
    };
    uint32_t lk0 = fbits(hot[SENTINEL].f[3]);
    if ((lk0 & LK_HASL) && !(lk0 & LK_TDL)) {
        int stack[2 * MAX_DEPTH + 4];
        int sp = 0;
        int c0 = classify(a, qi, hot[SENTINEL].f + 4, qx, qy, qz, r);
        if (c0 == 1 && !fill) {
            int4 cc = a.pool.cnt[ROOT];
            total += cc.x - cc.y;
        } else if (c0 != 0) {
            stack[sp++] = ROOT | (c0 == 1 ? (int)0x80000000 : 0);
        }
        float r2 = r * r;
        long long steps = 0;
        while (sp > 0) {
            if (++steps > a.max_steps) {
                atomicOr(a.overflow, 8);
                break;
            }
            int e = stack[--sp];
            bool whole = e < 0;
            int u = e & 0x7fffffff;
            const float* f = hot[u].f;
            uint32_t lk = fbits(f[3]);
            if (!(lk & LK_PDEL)) {
                bool take;
                if (whole) take = true;
                else if (a.mode == 0) take = box_contains(a.boxes + 6 * qi, f[0], f[1], f[2]);
                else take = calc_dist(f[0], f[1], f[2], qx, qy, qz) <= r2;
                if (take) emit(u, f);
            }
            int pair = (int)lk_pair(lk);
#pragma unroll
            for (int s = 1; s >= 0; --s) {
                if (!(lk & lk_has(s)) || (lk & lk_td(s))) continue;
                int c = 2 * pair + s;
                int cls = whole ? 1 : classify(a, qi, f + 4 + 6 * s, qx, qy, qz, r);
                if (cls == 0) continue;
                if (cls == 1 && !fill) {
                    int4 cc = a.pool.cnt[c];
                    total += cc.x - cc.y;
                    continue;
                }
                if (sp < 2 * MAX_DEPTH + 4) stack[sp++] = c | (cls == 1 ? (int)0x80000000 : 0);
                else atomicOr(a.overflow, 8);
            }
        }
    }
    if (!fill) a.counts[qi] = total;
}

void launch_range(const RangeArgs& a, int fill, hipStream_t s) {
    if (a.nq <= 0) return;
    hipLaunchKernelGGL(k_range, dim3(blocks_for(a.nq, 64)), dim3(64), 0, s, a, fill);
}

}  // namespace ikd

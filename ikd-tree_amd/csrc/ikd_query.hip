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
#include <algorithm>
#include <atomic>

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

// Sorted best-k list of one lane: bd_[0] is the current k-th candidate,
// slots past k hold (-inf, -2) padding.
template <int KM>
struct BestList {
    float d[KM];
    int n[KM];
    __device__ __forceinline__ void init(int k) {
#pragma unroll
        for (int i = 0; i < KM; ++i) {
            d[i] = i < k ? INFINITY : -INFINITY;
            n[i] = i < k ? -1 : -2;
        }
    }
    // offer (dist, node); keeps the k smallest in (d2, x, y, z, id) order.
    // Without an exact distance tie in the list the ranks are plain float
    // comparisons (branch-free selects); a tie takes the full comparator.
    __device__ __forceinline__ void offer(const NodeHot* hot, const uint32_t* ids, float dist, int node) {
        if (dist > d[0]) return;  // strictly worse than the k-th: the common case
        bool tie = false;
#pragma unroll
        for (int i = 0; i < KM; ++i) tie = tie || d[i] == dist;
        bool gnext[KM];
        if (!tie) {
#pragma unroll
            for (int i = 0; i < KM; ++i) gnext[i] = (i + 1 < KM) ? d[i + 1] > dist : false;
        } else {
            if (!after(hot, ids, d[0], n[0], dist, node)) return;
#pragma unroll
            for (int i = 0; i < KM; ++i)
                gnext[i] = (i + 1 < KM) ? after(hot, ids, d[i + 1], n[i + 1], dist, node) : false;
        }
#pragma unroll
        for (int i = 0; i < KM; ++i) {
            bool gcur = (i == 0) ? true : gnext[i - 1];
            if (gnext[i]) {
                d[i] = d[i + 1 < KM ? i + 1 : i];
                n[i] = n[i + 1 < KM ? i + 1 : i];
            } else if (gcur) {
                d[i] = dist;
                n[i] = node;
            }
        }
    }
};

struct QueryCtx {
    const NodeHot* hot;
    const uint32_t* ids;
    int qo;      // query index in the caller's arrays
    int k;
    float thr;
    float qx, qy, qz;
};

__device__ __forceinline__ QueryCtx load_query(const KnnArgs& a, int qi) {
    QueryCtx c;
    c.qo = a.order ? (int)a.order[qi] : qi;
    c.k = a.qk ? min(max(a.qk[c.qo], 0), a.k) : a.k;
    if (a.trees) {
        const int ti = a.qtree[c.qo];
        const bool ok = ti >= 0 && ti < a.ntrees;
        const TreeRef t = a.trees[ok ? ti : 0];
        c.hot = t.hot;
        c.ids = t.ids;
        if (!ok) c.k = 0;
    } else {
        c.hot = a.hot;
        c.ids = a.ids;
    }
    c.thr = a.qthr ? a.qthr[c.qo] : a.thr;
    c.qx = a.q[3 * c.qo];
    c.qy = a.q[3 * c.qo + 1];
    c.qz = a.q[3 * c.qo + 2];
    return c;
}

__device__ __forceinline__ void write_slot(const KnnArgs& a, const QueryCtx& c, int j, float d, int node) {
    const long long o = (long long)c.qo * a.out_stride + j;
    if (node >= 0) {
        const float* f = c.hot[node].f;
        a.out_xyz[3 * o] = f[0];
        a.out_xyz[3 * o + 1] = f[1];
        a.out_xyz[3 * o + 2] = f[2];
        a.out_d2[o] = d;
        a.out_ids[o] = c.ids[node];
    } else {
        a.out_xyz[3 * o] = 0.0f;
        a.out_xyz[3 * o + 1] = 0.0f;
        a.out_xyz[3 * o + 2] = 0.0f;
        a.out_d2[o] = INFINITY;
        a.out_ids[o] = 0xffffffffu;
    }
}

// Ascending output: the list holds the k best in descending order with the
// real entries in its last `found` slots.  Slots [found, out_stride) of the
// query's row are written empty.
template <int KM>
__device__ __forceinline__ void finish_query(const KnnArgs& a, const QueryCtx& c, const BestList<KM>& best,
                                             unsigned int visits, unsigned int boxes) {
    int found = 0;
#pragma unroll
    for (int i = 0; i < KM; ++i) found += best.n[i] >= 0 ? 1 : 0;
    a.out_counts[c.qo] = found;
#pragma unroll
    for (int i = 0; i < KM; ++i) {
        if (i < c.k) write_slot(a, c, c.k - 1 - i, best.d[i], best.n[i]);
    }
    for (int j = c.k; j < a.out_stride; ++j) write_slot(a, c, j, INFINITY, -1);
    if (a.visits) a.visits[c.qo] = visits;
    if (a.boxes) a.boxes[c.qo] = boxes;
}

// One visit of node n: offer its point, then compute both children's box
// distances.  Returns the near child (or -1) and, through far/df, the far one.
template <int KM>
__device__ __forceinline__ void visit_node(const QueryCtx& c, BestList<KM>& best, int n, int& near_n, float& dn,
                                           int& far_n, float& df, unsigned int& boxes) {
    const gv4* rec = reinterpret_cast<const gv4*>(c.hot[n].f);
    const gv4 r0 = gload4(rec), r1 = gload4(rec + 1), r2 = gload4(rec + 2), r3 = gload4(rec + 3);
    const uint32_t lk = fbits(r0.w);
    if (!(lk & LK_PDEL)) {
        float d = calc_dist(c.qx, c.qy, c.qz, r0.x, r0.y, r0.z);
        if (d <= c.thr) best.offer(c.hot, c.ids, d, n);
    }
    const int pair = (int)lk_pair(lk);
    const bool hl = (lk & LK_HASL) != 0, hr = (lk & LK_HASR) != 0;
    boxes += (hl ? 1u : 0u) + (hr ? 1u : 0u);
    const bool el = hl && !(lk & LK_TDL), er = hr && !(lk & LK_TDR);
    float bl[6] = {r1.x, r1.y, r1.z, r1.w, r2.x, r2.y};
    float br[6] = {r2.z, r2.w, r3.x, r3.y, r3.z, r3.w};
    const float dl = el ? calc_box_dist(bl, c.qx, c.qy, c.qz) : INFINITY;
    const float dr = er ? calc_box_dist(br, c.qx, c.qy, c.qz) : INFINITY;
    near_n = -1;
    far_n = -1;
    if (el || er) {
        const bool near_left = el && (!er || dl <= dr);
        near_n = near_left ? 2 * pair : 2 * pair + 1;
        dn = near_left ? dl : dr;
        if (near_left ? er : el) {
            far_n = near_left ? 2 * pair + 1 : 2 * pair;
            df = near_left ? dr : dl;
        }
    }
}

// ---- LDS traversal stacks --------------------------------------------------
// Each lane's deferred far children live in dynamic LDS as [stack][64] node
// indices followed by [stack][64] 16-bit box distances.  A stored distance is
// the float's upper half (sign, exponent, 7 mantissa bits): for the
// non-negative box distances that is a round-down, so a popped entry is
// admitted whenever the exact distance would be (the visit then offers
// nothing the exact test would have rejected, it only costs the visit).  Six
// bytes per entry instead of eight, and the stack depth follows the tree depth
// at run time, keep more wavefronts resident per CU.
struct LdsStack {
    int* n;
    unsigned short* d;
};
__device__ __forceinline__ LdsStack lds_stack(int stack) {
    extern __shared__ int lds_dyn[];
    return LdsStack{lds_dyn, reinterpret_cast<unsigned short*>(lds_dyn + stack * 64)};
}
__device__ __forceinline__ unsigned short dist_key(float d) { return (unsigned short)(fbits(d) >> 16); }
__device__ __forceinline__ float dist_val(unsigned short k) { return bitsf((uint32_t)k << 16); }
static inline size_t lds_stack_bytes(int stack) { return (size_t)stack * 64 * (sizeof(int) + sizeof(unsigned short)); }

// ---- one lane per query -------------------------------------------------
// Workgroups are dispatched to the 8 XCDs round-robin; map block b to query
// chunk (b % 8) * per + b / 8 so each XCD (and its L2) works on one
// contiguous eighth of the batch, i.e. on one spatial region of a scan.
constexpr int NUM_XCD = 8;
__device__ __forceinline__ int xcd_chunk(int b, int nchunks) {
    const int per = (nchunks + NUM_XCD - 1) / NUM_XCD;
    return (b % NUM_XCD) * per + b / NUM_XCD;
}
static inline int xcd_grid(int nchunks) {
    const int per = (nchunks + 7) / 8;
    return per * 8;
}

// Compact stacks (trees of at most 2^24 nodes): one 4-byte entry per
// deferred child, the node index in the upper 24 bits and the box distance's
// exponent bits below, i.e. the largest power of two not above it (0 for 0
// and subnormals).  That lower bound admits a popped entry whenever the exact
// distance would, and at most one visit more than the 16-bit key (the visit
// evaluates the children's boxes exactly, and a child's box lies inside the
// node's).  Two thirds of the LDS per lane keep more wavefronts resident.
constexpr int COMPACT_NODE_BITS = 24;
__device__ __forceinline__ uint32_t compact_entry(int n, float d) {
    return ((uint32_t)n << 8) | (fbits(d) >> 23);
}
__device__ __forceinline__ float compact_dist(uint32_t e) { return bitsf((e & 0xffu) << 23); }
static inline size_t lds_compact_bytes(int stack) { return (size_t)stack * 64 * sizeof(uint32_t); }

template <int KM, bool COMPACT>
__global__ __launch_bounds__(64) void k_knn(KnnArgs a) {
    const LdsStack st = lds_stack(a.stack);
    extern __shared__ int lds_dyn[];
    uint32_t* const cst = reinterpret_cast<uint32_t*>(lds_dyn);
    const int lane = threadIdx.x;
    const int chunk = xcd_chunk(blockIdx.x, (a.nq + 63) / 64);
    const int gq = chunk * 64 + lane;
    if (gq >= a.nq) return;
    const QueryCtx c = load_query(a, gq);
    BestList<KM> best;
    best.init(c.k);
    unsigned int visits = 0, boxes = 0;

    const uint32_t lk0 = fbits(c.hot[SENTINEL].f[3]);
    int sp = 0;
    int n = -1;
    float bd = 0.0f;
    if ((lk0 & LK_HASL) && !(lk0 & LK_TDL)) {
        n = ROOT;
        bd = calc_box_dist(c.hot[SENTINEL].f + 4, c.qx, c.qy, c.qz);
    }
    while (true) {
        if ((int)visits > a.max_steps) {
            atomicOr(a.overflow, 2);
            break;
        }
        if (n >= 0 && bd <= c.thr) {
            ++visits;
            int nn, fn;
            float dn = 0.0f, df = 0.0f;
            visit_node<KM>(c, best, n, nn, dn, fn, df, boxes);
            const float top = best.d[0];
            if (fn >= 0 && df <= top && df <= c.thr) {
                if (sp < a.stack) {
                    if (COMPACT) {
                        cst[sp * 64 + lane] = compact_entry(fn, df);
                    } else {
                        st.n[sp * 64 + lane] = fn;
                        st.d[sp * 64 + lane] = dist_key(df);
                    }
                    ++sp;
                } else {
                    atomicOr(a.overflow, 1);
                }
            }
            if (nn >= 0 && dn <= top && dn <= c.thr) {
                n = nn;
                bd = dn;
                continue;
            }
        }
        // backtrack to the deepest deferred child that can still improve
        n = -1;
        while (sp > 0) {
            --sp;
            if (COMPACT) {
                const uint32_t e = cst[sp * 64 + lane];
                const float d = compact_dist(e);
                if (d <= best.d[0]) {
                    n = (int)(e >> 8);
                    bd = d;
                    break;
                }
            } else {
                const float d = dist_val(st.d[sp * 64 + lane]);
                if (d <= best.d[0]) {
                    n = st.n[sp * 64 + lane];
                    bd = d;
                    break;
                }
            }
        }
        if (n < 0) break;
    }

    finish_query<KM>(a, c, best, visits, boxes);
}

// ---- one lane per query, top levels staged in LDS -------------------------
// The same walk as k_knn; the first levels of the tree are read from a copy in
// the workgroup's LDS instead of global memory.  A node handle is its index
// (>= 0), -1 for none, or -2 - s for slot s of the staged table, whose
// children are slots 2s+1 and 2s+2 down to the table's last level and node
// indices below it.  The best list keeps node indices (top_idx), so results
// and output are those of k_knn.
__global__ void k_top_table(const NodeHot* __restrict__ hot, int slots, NodeHot* __restrict__ top,
                            int* __restrict__ top_idx) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= slots) return;
    // the path from the root: the bits of s + 1 below its leading one, top down
    const uint32_t p = (uint32_t)s + 1u;
    const int depth = 31 - __builtin_clz(p);
    const uint32_t lk0 = fbits(hot[SENTINEL].f[3]);
    int n = (lk0 & LK_HASL) ? ROOT : -1;
    for (int d = depth - 1; d >= 0 && n >= 0; --d) {
        const int side = (int)((p >> d) & 1u);
        const uint32_t lk = fbits(hot[n].f[3]);
        n = (lk & lk_has(side)) ? (int)(2 * lk_pair(lk)) + side : -1;
    }
    top_idx[s] = n;
    float4* o = reinterpret_cast<float4*>(top[s].f);
    if (n >= 0) {
        const float4* r = reinterpret_cast<const float4*>(hot[n].f);
        o[0] = r[0];
        o[1] = r[1];
        o[2] = r[2];
        o[3] = r[3];
    } else {
        o[0] = o[1] = o[2] = o[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

void launch_top_table(const NodeHot* hot, int levels, NodeHot* top, int* top_idx, hipStream_t s) {
    const int slots = (1 << levels) - 1;
    hipLaunchKernelGGL(k_top_table, dim3(blocks_for(slots, 256)), dim3(256), 0, s, hot, slots, top, top_idx);
}

// bytes of the LDS stacks rounded up to whole 16-byte rows (the table follows)
IKD_HD size_t lds_stack_bytes16(int stack) { return ((size_t)stack * 64 * (sizeof(int) + sizeof(unsigned short)) + 15) & ~(size_t)15; }
static inline size_t lds_top_bytes(int slots) { return (size_t)slots * (sizeof(NodeHot) + sizeof(int)); }

template <int KM>
__global__ __launch_bounds__(64) void k_knn_staged(KnnArgs a) {
    extern __shared__ int lds_dyn[];
    const LdsStack st = lds_stack(a.stack);
    const int slots = a.top_slots;
    float4* t4 = reinterpret_cast<float4*>(reinterpret_cast<char*>(lds_dyn) + lds_stack_bytes16(a.stack));
    int* tix = reinterpret_cast<int*>(t4 + 4 * slots);
    const int lane = threadIdx.x;
    {
        const gv4* src = reinterpret_cast<const gv4*>(a.top);
        for (int i = lane; i < 4 * slots; i += 64) {
            const gv4 v = gload4(src + i);
            t4[i] = make_float4(v.x, v.y, v.z, v.w);
        }
        for (int i = lane; i < slots; i += 64) tix[i] = a.top_idx[i];
    }
    __syncthreads();
    const int chunk = xcd_chunk(blockIdx.x, (a.nq + 63) / 64);
    const int gq = chunk * 64 + lane;
    if (gq >= a.nq) return;
    const QueryCtx c = load_query(a, gq);
    BestList<KM> best;
    best.init(c.k);
    unsigned int visits = 0, boxes = 0;

    const uint32_t lk0 = fbits(c.hot[SENTINEL].f[3]);
    int sp = 0;
    int n = -1;
    float bd = 0.0f;
    if ((lk0 & LK_HASL) && !(lk0 & LK_TDL)) {
        n = -2;  // the root: slot 0
        bd = calc_box_dist(c.hot[SENTINEL].f + 4, c.qx, c.qy, c.qz);
    }
    while (true) {
        if ((int)visits > a.max_steps) {
            atomicOr(a.overflow, 2);
            break;
        }
        if (n != -1 && bd <= c.thr) {
            ++visits;
            float4 r0, r1, r2, r3;
            int node, cl, cr;
            if (n <= -2) {
                const int s = -2 - n;
                r0 = t4[4 * s];
                r1 = t4[4 * s + 1];
                r2 = t4[4 * s + 2];
                r3 = t4[4 * s + 3];
                node = tix[s];
                const uint32_t lk = fbits(r0.w);
                const bool inner = 2 * s + 2 < slots;
                cl = inner ? -2 - (2 * s + 1) : (int)(2 * lk_pair(lk));
                cr = inner ? -2 - (2 * s + 2) : (int)(2 * lk_pair(lk)) + 1;
            } else {
                const gv4* rec = reinterpret_cast<const gv4*>(c.hot[n].f);
                const gv4 g0 = gload4(rec), g1 = gload4(rec + 1), g2 = gload4(rec + 2), g3 = gload4(rec + 3);
                r0 = make_float4(g0.x, g0.y, g0.z, g0.w);
                r1 = make_float4(g1.x, g1.y, g1.z, g1.w);
                r2 = make_float4(g2.x, g2.y, g2.z, g2.w);
                r3 = make_float4(g3.x, g3.y, g3.z, g3.w);
                node = n;
                const uint32_t lk = fbits(r0.w);
                cl = (int)(2 * lk_pair(lk));
                cr = cl + 1;
            }
            const uint32_t lk = fbits(r0.w);
            if (!(lk & LK_PDEL)) {
                const float d = calc_dist(c.qx, c.qy, c.qz, r0.x, r0.y, r0.z);
                if (d <= c.thr) best.offer(c.hot, c.ids, d, node);
            }
            const bool hl = (lk & LK_HASL) != 0, hr = (lk & LK_HASR) != 0;
            boxes += (hl ? 1u : 0u) + (hr ? 1u : 0u);
            const bool el = hl && !(lk & LK_TDL), er = hr && !(lk & LK_TDR);
            const float bl[6] = {r1.x, r1.y, r1.z, r1.w, r2.x, r2.y};
            const float br[6] = {r2.z, r2.w, r3.x, r3.y, r3.z, r3.w};
            const float dl = el ? calc_box_dist(bl, c.qx, c.qy, c.qz) : INFINITY;
            const float dr = er ? calc_box_dist(br, c.qx, c.qy, c.qz) : INFINITY;
            int nn = -1, fn = -1;
            float dn = 0.0f, df = 0.0f;
            if (el || er) {
                const bool near_left = el && (!er || dl <= dr);
                nn = near_left ? cl : cr;
                dn = near_left ? dl : dr;
                if (near_left ? er : el) {
                    fn = near_left ? cr : cl;
                    df = near_left ? dr : dl;
                }
            }
            const float top = best.d[0];
            if (fn != -1 && df <= top && df <= c.thr) {
                if (sp < a.stack) {
                    st.n[sp * 64 + lane] = fn;
                    st.d[sp * 64 + lane] = dist_key(df);
                    ++sp;
                } else {
                    atomicOr(a.overflow, 1);
                }
            }
            if (nn != -1 && dn <= top && dn <= c.thr) {
                n = nn;
                bd = dn;
                continue;
            }
        }
        n = -1;
        while (sp > 0) {
            --sp;
            const float d = dist_val(st.d[sp * 64 + lane]);
            if (d <= best.d[0]) {
                n = st.n[sp * 64 + lane];
                bd = d;
                break;
            }
        }
        if (n == -1) break;
    }

    finish_query<KM>(a, c, best, visits, boxes);
}

// ---- a team of T lanes per query ------------------------------------------
// Each lane of a team runs its own depth-first walk with its own deque of
// deferred far children (a ring buffer in LDS holding at most one entry per
// level of its current walk, so never more than the tree depth).  Whenever a
// lane is idle, a busy teammate donates the bottom (shallowest) entry of its
// deque.  Each lane keeps its own best-k list; the pruning bound is the
// smallest k-th distance of the team, which is never below the true k-th
// distance, so nothing of the answer is pruned.  At the end the T lists are
// merged by rank in the total order.  Every node is reached through exactly
// one parent, so no point is offered twice and the merge sees distinct
// candidates.
__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// A deque slot of either layout: 6-byte (index, 16-bit distance) or compact.
template <bool COMPACT>
__device__ __forceinline__ void slot_put(const LdsStack& st, uint32_t* cst, int i, int node, float d) {
    if (COMPACT) {
        cst[i] = compact_entry(node, d);
    } else {
        st.n[i] = node;
        st.d[i] = dist_key(d);
    }
}
template <bool COMPACT>
__device__ __forceinline__ int slot_get(const LdsStack& st, const uint32_t* cst, int i, float* d) {
    if (COMPACT) {
        const uint32_t e = cst[i];
        *d = compact_dist(e);
        return (int)(e >> 8);
    }
    *d = dist_val(st.d[i]);
    return st.n[i];
}

template <int KM, int T, bool COMPACT>
__global__ __launch_bounds__(64) void k_knn_team(KnnArgs a) {
    constexpr int TEAMS = 64 / T;
    const LdsStack st = lds_stack(a.stack);
    extern __shared__ int lds_dyn[];
    uint32_t* const cst = reinterpret_cast<uint32_t*>(lds_dyn);
    const int smask = a.stack - 1;  // team launches use a power-of-two stack
    __shared__ int dn_n[TEAMS][T];
    __shared__ float dn_d[TEAMS][T];
    __shared__ float mg_d[TEAMS][T * KM];
    __shared__ int mg_n[TEAMS][T * KM];
    const int lane = threadIdx.x;
    const int team = lane / T, tl = lane % T;
    const unsigned long long team_mask = (T == 64) ? ~0ull : (((1ull << T) - 1ull) << (team * T));
    const unsigned long long below = (1ull << lane) - 1ull;
    const int qi = xcd_chunk(blockIdx.x, (a.nq + TEAMS - 1) / TEAMS) * TEAMS + team;
    const bool valid = qi < a.nq;
    QueryCtx c = load_query(a, valid ? qi : 0);
    if (!valid) c.k = 1;
    BestList<KM> best;
    best.init(c.k);
    unsigned int visits = 0, boxes = 0;
    int head = 0, tail = 0;  // deque [head, tail) modulo a.stack

    int n = -1;
    float bd = 0.0f;
    if (valid && tl == 0) {
        const uint32_t lk0 = fbits(c.hot[SENTINEL].f[3]);
        if ((lk0 & LK_HASL) && !(lk0 & LK_TDL)) {
            n = ROOT;
            bd = calc_box_dist(c.hot[SENTINEL].f + 4, c.qx, c.qy, c.qz);
        }
    }
    float B = INFINITY;  // team bound: min over the team's k-th distances
    int iter = 0;
    while (true) {
        // an idle lane resumes the deepest entry of its own deque
        while (n < 0 && tail != head) {
            --tail;
            const int s = tail & smask;
            float d;
            const int e = slot_get<COMPACT>(st, cst, s * 64 + lane, &d);
            if (d <= fminf(B, best.d[0])) {
                n = e;
                bd = d;
            }
        }
        // lanes still idle receive the bottom entry of a busy teammate
        const unsigned long long imask = __ballot(valid && n < 0) & team_mask;
        const unsigned long long dmask = __ballot(tail != head) & team_mask;
        // teams with idle and busy lanes hand work over; the fences run on
        // every lane of the wavefront whenever any team does, so the wave
        // barriers are never inside team-divergent control flow
        const bool donate = imask != 0 && dmask != 0;
        if (__ballot(donate) != 0) {
            const int ni = __popcll(imask), nd = __popcll(dmask);
            if (donate && tail != head) {
                const int r = __popcll(dmask & below);
                if (r < ni) {
                    const int s = head & smask;
                    float d;
                    dn_n[team][r] = slot_get<COMPACT>(st, cst, s * 64 + lane, &d);
                    dn_d[team][r] = d;
                    ++head;
                }
            }
            wave_lds_fence();
            if (donate && valid && n < 0) {
                const int r = __popcll(imask & below);
                if (r < nd) {
                    n = dn_n[team][r];
                    bd = dn_d[team][r];
                }
            }
            wave_lds_fence();
        }
        if (__ballot(n >= 0 || tail != head) == 0) break;
        if (++iter > a.max_steps) {
            if (lane == 0) atomicOr(a.overflow, 2);
            break;
        }

        if (n >= 0) {
            float bound = fminf(B, best.d[0]);
            if (bd > c.thr || bd > bound) {
                n = -1;
            } else {
                ++visits;
                int nn, fn;
                float dn = 0.0f, df = 0.0f;
                visit_node<KM>(c, best, n, nn, dn, fn, df, boxes);
                bound = fminf(B, best.d[0]);
                if (fn >= 0 && df <= bound && df <= c.thr) {
                    if (tail - head < a.stack) {
                        const int s = tail & smask;
                        slot_put<COMPACT>(st, cst, s * 64 + lane, fn, df);
                        ++tail;
                    } else {
                        atomicOr(a.overflow, 1);
                    }
                }
                if (nn >= 0 && dn <= bound && dn <= c.thr) {
                    n = nn;
                    bd = dn;
                } else {
                    n = -1;
                }
            }
        }
        float v = best.d[0];
#pragma unroll
        for (int off = T / 2; off >= 1; off >>= 1) v = fminf(v, __shfl_xor(v, off, T));
        B = v;
    }

    // ---- merge the team's lists by rank in (d2, x, y, z, id) order
#pragma unroll
    for (int i = 0; i < KM; ++i) {
        mg_d[team][tl * KM + i] = best.d[i];
        mg_n[team][tl * KM + i] = best.n[i];
    }
    wave_lds_fence();
    int real = 0;
    for (int cidx = tl; cidx < T * KM; cidx += T) real += mg_n[team][cidx] >= 0 ? 1 : 0;
#pragma unroll
    for (int off = T / 2; off >= 1; off >>= 1) real += __shfl_xor(real, off, T);
    unsigned int vsum = visits, bsum = boxes;
#pragma unroll
    for (int off = T / 2; off >= 1; off >>= 1) {
        vsum += __shfl_xor(vsum, off, T);
        bsum += __shfl_xor(bsum, off, T);
    }
    if (!valid) return;
    const int found = min(real, c.k);
    for (int cidx = tl; cidx < T * KM; cidx += T) {
        const int node = mg_n[team][cidx];
        if (node < 0) continue;
        const float d = mg_d[team][cidx];
        int rank = 0;
        for (int j = 0; j < T * KM; ++j) {
            const int nj = mg_n[team][j];
            if (nj < 0 || j == cidx) continue;
            const float dj = mg_d[team][j];
            if (dj < d || (dj == d && tie_less(c.hot, c.ids, nj, node))) ++rank;
        }
        if (rank < c.k) write_slot(a, c, rank, d, node);
    }
    for (int j = found + tl; j < a.out_stride; j += T) write_slot(a, c, j, INFINITY, -1);
    if (tl == 0) {
        a.out_counts[c.qo] = found;
        if (a.visits) a.visits[c.qo] = vsum;
        if (a.boxes) a.boxes[c.qo] = bsum;
    }
}

// ---- persistent lanes -----------------------------------------------------
// A fixed grid (a multiple of 8 workgroups, so workgroup b stays on XCD
// b % 8) walks the batch; a lane that finishes its query takes the next one
// of its XCD's contiguous eighth of the batch.  Idle lanes of a wavefront
// claim their queries with one atomic per wavefront.
// idle lanes a persistent wavefront collects before it takes new queries
constexpr int PERSIST_REFILL = 16;

template <int KM, bool COMPACT>
__global__ __launch_bounds__(64) void k_knn_persist(KnnArgs a, int* __restrict__ next) {
    const LdsStack st = lds_stack(a.stack);
    extern __shared__ int lds_dyn[];
    uint32_t* const cst = reinterpret_cast<uint32_t*>(lds_dyn);
    const int lane = threadIdx.x;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int xcd = blockIdx.x % NUM_XCD;
    const int per = (a.nq + NUM_XCD - 1) / NUM_XCD;
    const int lo = min(xcd * per, a.nq), hi = min(lo + per, a.nq);
    QueryCtx c;
    BestList<KM> best;
    bool has = false;
    int n = -1, sp = 0;
    float bd = 0.0f;
    unsigned int visits = 0, boxes = 0;
    bool drained = false;
    while (true) {
        // refill idle lanes in bunches (one counter update per bunch), and
        // when every lane is idle
        const unsigned long long imask = __ballot(!has);
        const int idle = __popcll(imask);
        if (idle == 64 && drained) break;
        if (!drained && (idle >= PERSIST_REFILL || idle == 64)) {
            const int leader = __ffsll((long long)imask) - 1;
            int base = 0;
            if (lane == leader) base = atomicAdd(&next[xcd], __popcll(imask));
            base = __shfl(base, leader);
            if (base + lo >= hi) drained = true;
            if (!has) {
                const int qi = lo + base + __popcll(imask & below);
                if (qi < hi) {
                    c = load_query(a, qi);
                    best.init(c.k);
                    visits = boxes = 0;
                    sp = 0;
                    n = -1;
                    const uint32_t lk0 = fbits(c.hot[SENTINEL].f[3]);
                    if ((lk0 & LK_HASL) && !(lk0 & LK_TDL)) {
                        n = ROOT;
                        bd = calc_box_dist(c.hot[SENTINEL].f + 4, c.qx, c.qy, c.qz);
                    }
                    has = true;
                }
            }
        }
        if (!has) continue;
        if ((int)visits > a.max_steps) {
            atomicOr(a.overflow, 2);
            n = -1;
            sp = 0;
        }
        if (n >= 0 && bd <= c.thr) {
            ++visits;
            int nn, fn;
            float dn = 0.0f, df = 0.0f;
            visit_node<KM>(c, best, n, nn, dn, fn, df, boxes);
            const float top = best.d[0];
            if (fn >= 0 && df <= top && df <= c.thr) {
                if (sp < a.stack) {
                    if (COMPACT) {
                        cst[sp * 64 + lane] = compact_entry(fn, df);
                    } else {
                        st.n[sp * 64 + lane] = fn;
                        st.d[sp * 64 + lane] = dist_key(df);
                    }
                    ++sp;
                } else {
                    atomicOr(a.overflow, 1);
                }
            }
            if (nn >= 0 && dn <= top && dn <= c.thr) {
                n = nn;
                bd = dn;
                continue;
            }
        }
        n = -1;
        while (sp > 0) {
            --sp;
            if (COMPACT) {
                const uint32_t e = cst[sp * 64 + lane];
                const float d = compact_dist(e);
                if (d <= best.d[0]) {
                    n = (int)(e >> 8);
                    bd = d;
                    break;
                }
            } else {
                const float d = dist_val(st.d[sp * 64 + lane]);
                if (d <= best.d[0]) {
                    n = st.n[sp * 64 + lane];
                    bd = d;
                    break;
                }
            }
        }
        if (n >= 0) continue;
        finish_query<KM>(a, c, best, visits, boxes);
        has = false;
    }
}

// compute units of the current device (cached per device)
static int device_cus() {
    static std::atomic<int> cus[16];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return 256;
    int v = cus[dev].load(std::memory_order_relaxed);
    if (v <= 0) {
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0) v = 256;
        cus[dev].store(v, std::memory_order_relaxed);
    }
    return v;
}

template <int KM>
static int launch_knn_km(const KnnArgs& a0, int lanes, int stack, hipStream_t s, int* persist_next,
                         int persist_waves) {
    KnnArgs a = a0;
    dim3 block(64);
    if (lanes >= 2) {  // team deques are ring buffers: round the stack up to a power of two
        int p = 1;
        while (p < stack) p <<= 1;
        stack = p;
    }
    a.stack = stack;
    const size_t lds = lds_stack_bytes(stack);
    switch (lanes) {
        case 1:
            if (a.top_slots > 0 && a.top && !a.trees) {
                const size_t lds_t = lds_stack_bytes16(stack) + lds_top_bytes(a.top_slots);
                hipLaunchKernelGGL((k_knn_staged<KM>), dim3(xcd_grid(blocks_for(a.nq, 64))), block, lds_t, s, a);
                return 0;
            }
            if (a.compact_stack) {
                hipLaunchKernelGGL((k_knn<KM, true>), dim3(xcd_grid(blocks_for(a.nq, 64))), block,
                                   lds_compact_bytes(stack), s, a);
                return 0;
            }
            hipLaunchKernelGGL((k_knn<KM, false>), dim3(xcd_grid(blocks_for(a.nq, 64))), block, lds, s, a);
            return 0;
        case 0: {  // persistent lanes; the 8 per-XCD query counters were zeroed by the caller
            // as many single-wavefront workgroups as the LDS stacks let each CU hold
            const size_t lb = a.compact_stack ? lds_compact_bytes(stack) : lds;
            const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(32, (160u << 10) / std::max<size_t>(lb, 1)));
            int waves = std::min(blocks_for(a.nq, 64), std::max(persist_waves, device_cus() * per_cu));
            if (a.compact_stack)
                hipLaunchKernelGGL((k_knn_persist<KM, true>), dim3(xcd_grid(waves)), block, lb, s, a, persist_next);
            else
                hipLaunchKernelGGL((k_knn_persist<KM, false>), dim3(xcd_grid(waves)), block, lb, s, a, persist_next);
            return 0;
        }
        case 2:
            if (a.compact_stack)
                hipLaunchKernelGGL((k_knn_team<KM, 2, true>), dim3(xcd_grid(blocks_for(a.nq, 32))), block,
                                   lds_compact_bytes(stack), s, a);
            else
                hipLaunchKernelGGL((k_knn_team<KM, 2, false>), dim3(xcd_grid(blocks_for(a.nq, 32))), block, lds,
                                   s, a);
            return 0;
        case 4:
            if (a.compact_stack)
                hipLaunchKernelGGL((k_knn_team<KM, 4, true>), dim3(xcd_grid(blocks_for(a.nq, 16))), block,
                                   lds_compact_bytes(stack), s, a);
            else
                hipLaunchKernelGGL((k_knn_team<KM, 4, false>), dim3(xcd_grid(blocks_for(a.nq, 16))), block, lds,
                                   s, a);
            return 0;
        case 8:
            if (a.compact_stack)
                hipLaunchKernelGGL((k_knn_team<KM, 8, true>), dim3(xcd_grid(blocks_for(a.nq, 8))), block,
                                   lds_compact_bytes(stack), s, a);
            else
                hipLaunchKernelGGL((k_knn_team<KM, 8, false>), dim3(xcd_grid(blocks_for(a.nq, 8))), block, lds,
                                   s, a);
            return 0;
        case 16:
            if (a.compact_stack)
                hipLaunchKernelGGL((k_knn_team<KM, 16, true>), dim3(xcd_grid(blocks_for(a.nq, 4))), block,
                                   lds_compact_bytes(stack), s, a);
            else
                hipLaunchKernelGGL((k_knn_team<KM, 16, false>), dim3(xcd_grid(blocks_for(a.nq, 4))), block, lds,
                                   s, a);
            return 0;
        default:
            return -1;
    }
}

int launch_knn(const KnnArgs& a, int kmax, int lanes, int stack, hipStream_t s, int* persist_next,
               int persist_waves) {
    if (a.nq <= 0) return 0;
    if (lanes == 0) launch_fill_i32(persist_next, 0, NUM_XCD, s);
#define IKD_KNN(KMV) return launch_knn_km<KMV>(a, lanes, stack, s, persist_next, persist_waves)
    if (kmax <= 1) IKD_KNN(1);
    if (kmax <= 2) IKD_KNN(2);
    if (kmax <= 4) IKD_KNN(4);
    if (kmax <= 5) IKD_KNN(5);
    if (kmax <= 8) IKD_KNN(8);
    if (kmax <= 10) IKD_KNN(10);
    if (kmax <= 16) IKD_KNN(16);
    if (kmax <= 20) IKD_KNN(20);
    if (kmax <= 32) IKD_KNN(32);
#undef IKD_KNN
    if (kmax <= 64) return launch_knn_km<64>(a, lanes <= 1 ? lanes : (lanes > 4 ? 4 : lanes), stack, s, persist_next,
                                             persist_waves);
    return -1;
}

// ------------------------------------------------ per-query max_dist thresholds
// The device form of the host conversion: m = max_dist^2 in double, the float
// nearest to m, stepped down one ulp when it lies above m.
__global__ void k_thresholds(const double* __restrict__ md, int n, float* __restrict__ thr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = md[i] * md[i];
    float f;
    if (m != m) {
        f = -INFINITY;
    } else {
        f = __double2float_rn(m);
        if ((double)f > m) f = bitsf(fbits(f) - 1u);  // f > m >= 0: one ulp toward zero
    }
    thr[i] = f;
}

void launch_thresholds(const double* max_dist, int n, float* thr, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_thresholds, dim3(blocks_for(n, 256)), dim3(256), 0, s, max_dist, n, thr);
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

// Range search in items.  A walk per query (one lane each, depth-first
// stack in LDS as [stack][64] entries) goes down only through subtrees of
// more than task_nodes nodes; everything at or below that size, and the own
// point of every node it walks, becomes an item of the query's list, in
// pre-order:
//   RANGE_OWN   the node's own point (the walk checked the predicate),
//   RANGE_COVER a subtree the query covers whole,
//   RANGE_PART  a subtree the query cuts.
// The count walk (k_range_top, list 0) counts each query's items, the list
// walk (list 1) writes them; one lane per item then counts its points
// (k_range_item_count: 1, the subtree's live count, or a walk of the cut
// subtree with covered parts counted from their stored counts), a scan over
// the items gives every item its output offset and every query its first,
// and one lane per item writes its points (k_range_item_fill) in pre-order.
// The output is the pre-order sequence of one walk per query.
//
// Stack entries: node index (< 2^28) with bit 31 = covered whole, bit 30 =
// an item left for its pop (so items keep pre-order), bit 29 = the item is
// a cut subtree.
constexpr int RANGE_WHOLE = (int)0x80000000;
constexpr int RANGE_TASK = 0x40000000;
constexpr int RANGE_CUT = 0x20000000;
constexpr int RANGE_NODE = 0x1fffffff;
constexpr int RANGE_OWN = 0, RANGE_COVER = 1, RANGE_PART = 2;

__device__ __forceinline__ Pool range_pool(const RangeArgs& a, int qi, bool* ok) {
    *ok = true;
    if (!a.pools) return a.pool;
    const int ti = a.qtree[qi];
    if (ti < 0 || ti >= a.ntrees) {
        *ok = false;
        return a.pool;
    }
    return a.pools[ti];
}

struct RangeQuery {
    float qx, qy, qz, r, r2;
};
__device__ __forceinline__ RangeQuery range_query(const RangeArgs& a, int qi) {
    RangeQuery q{0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (a.mode == 1) {
        q.qx = a.centers[3 * qi];
        q.qy = a.centers[3 * qi + 1];
        q.qz = a.centers[3 * qi + 2];
        q.r = a.radii[qi];
        q.r2 = q.r * q.r;
    }
    return q;
}
__device__ __forceinline__ bool range_takes(const RangeArgs& a, int qi, const RangeQuery& q, const float* f) {
    if (a.mode == 0) return box_contains(a.boxes + 6 * qi, f[0], f[1], f[2]);
    return calc_dist(f[0], f[1], f[2], q.qx, q.qy, q.qz) <= q.r2;
}

__global__ __launch_bounds__(64) void k_range_top(RangeArgs a, int list) {
    extern __shared__ int lds_dyn[];
    int* stk = lds_dyn;
    const int lane = threadIdx.x;
    const int qi = blockIdx.x * 64 + lane;
    if (qi >= a.nq) return;
    bool ok;
    const Pool P = range_pool(a, qi, &ok);
    if (!ok) {
        if (!list) a.tcounts[qi] = 0;
        return;
    }
    const NodeHot* __restrict__ hot = P.hot;
    const int G = a.task_nodes;
    const RangeQuery q = range_query(a, qi);
    long long items = 0;
    long long tpos = list ? a.toffsets[qi] : 0;
    const long long tlim = list ? a.toffsets[qi + 1] : 0;
    auto item = [&](int u, int kind) {
        if (!list) {
            ++items;
        } else if (tpos < tlim) {
            a.tasks[tpos++] = make_int4(u, qi, kind, 0);
        } else {
            atomicOr(a.overflow, 8);
        }
    };
    // an entry for child c of class cls (1 covered, 2 cut)
    auto entry = [&](int c, int cls) -> int {
        if (P.cnt[c].x <= G) return c | RANGE_TASK | (cls == 2 ? RANGE_CUT : 0);
        return c | (cls == 1 ? RANGE_WHOLE : 0);
    };
    const uint32_t lk0 = fbits(hot[SENTINEL].f[3]);
    if ((lk0 & LK_HASL) && !(lk0 & LK_TDL)) {
        int sp = 0;
        const int c0 = classify(a, qi, hot[SENTINEL].f + 4, q.qx, q.qy, q.qz, q.r);
        if (c0 != 0) {
            stk[lane] = entry(ROOT, c0);
            sp = 1;
        }
        long long steps = 0;
        while (sp > 0) {
            if (++steps > a.max_steps) {
                atomicOr(a.overflow, 8);
                break;
            }
            const int e = stk[--sp * 64 + lane];
            const int u = e & RANGE_NODE;
            if (e & RANGE_TASK) {
                item(u, (e & RANGE_CUT) ? RANGE_PART : RANGE_COVER);
                continue;
            }
            const bool whole = (e & RANGE_WHOLE) != 0;
            float f[16];
            load_record(hot[u].f, f);
            const uint32_t lk = fbits(f[3]);
            if (!(lk & LK_PDEL) && (whole || range_takes(a, qi, q, f))) item(u, RANGE_OWN);
            const int pair = (int)lk_pair(lk);
#pragma unroll
            for (int s = 1; s >= 0; --s) {
                if (!(lk & lk_has(s)) || (lk & lk_td(s))) continue;
                const int c = 2 * pair + s;
                const int cls = whole ? 1 : classify(a, qi, f + 4 + 6 * s, q.qx, q.qy, q.qz, q.r);
                if (cls == 0) continue;
                if (sp < a.stack) stk[sp++ * 64 + lane] = entry(c, cls);
                else atomicOr(a.overflow, 8);
            }
        }
    }
    if (!list) a.tcounts[qi] = items;
}

// The points one item holds: counted (fill false) or written from pos on in
// pre-order (fill true, at most lim).  Returns the count / the end position.
__device__ long long range_item(const RangeArgs& a, const Pool& P, int qi, int u, int kind, int* stk, int lane,
                                long long pos, long long lim, bool fill) {
    const RangeQuery q = range_query(a, qi);
    const NodeHot* __restrict__ hot = P.hot;
    auto emit = [&](int n, const float* f) {
        if (!fill) {
            ++pos;
        } else if (pos < lim) {
            a.out_xyz[3 * pos] = f[0];
            a.out_xyz[3 * pos + 1] = f[1];
            a.out_xyz[3 * pos + 2] = f[2];
            a.out_ids[pos] = P.ids[n];
            ++pos;
        } else {
            atomicOr(a.overflow, 8);
        }
    };
    if (kind == RANGE_OWN) {
        float f[16];
        load_record(hot[u].f, f);
        emit(u, f);
        return pos;
    }
    if (kind == RANGE_COVER && !fill) {
        const int4 cc = P.cnt[u];
        return pos + (cc.x - cc.y);
    }
    int sp = 1;
    stk[lane] = u | (kind == RANGE_COVER ? RANGE_WHOLE : 0);
    long long steps = 0;
    while (sp > 0) {
        if (++steps > a.max_steps) {
            atomicOr(a.overflow, 8);
            break;
        }
        const int e = stk[--sp * 64 + lane];
        const int n = e & RANGE_NODE;
        const bool whole = (e & RANGE_WHOLE) != 0;
        float f[16];
        load_record(hot[n].f, f);
        const uint32_t lk = fbits(f[3]);
        if (!(lk & LK_PDEL) && (whole || range_takes(a, qi, q, f))) emit(n, f);
        const int pair = (int)lk_pair(lk);
#pragma unroll
        for (int s = 1; s >= 0; --s) {
            if (!(lk & lk_has(s)) || (lk & lk_td(s))) continue;
            const int c = 2 * pair + s;
            const int cls = whole ? 1 : classify(a, qi, f + 4 + 6 * s, q.qx, q.qy, q.qz, q.r);
            if (cls == 0) continue;
            if (cls == 1 && !fill) {  // counted from the stored counts
                const int4 cc = P.cnt[c];
                pos += cc.x - cc.y;
                continue;
            }
            if (sp < a.stack) stk[sp++ * 64 + lane] = c | (cls == 1 ? RANGE_WHOLE : 0);
            else atomicOr(a.overflow, 8);
        }
    }
    return pos;
}

// counts[i] = points of item i (counts[ntask], the scan's trailing zero, set by the caller)
__global__ __launch_bounds__(64) void k_range_item_count(RangeArgs a, long long ntask, long long* __restrict__ counts) {
    extern __shared__ int lds_dyn[];
    const int lane = threadIdx.x;
    const long long ti = (long long)blockIdx.x * 64 + lane;
    if (ti >= ntask) return;
    const int4 t = a.tasks[ti];
    bool ok;
    const Pool P = range_pool(a, t.y, &ok);
    counts[ti] = ok ? range_item(a, P, t.y, t.x, t.z, lds_dyn, lane, 0, 0, false) : 0;
}

// offsets[q] = output offset of query q's first item (items of a query are
// consecutive; toffsets[nq] = ntask gives offsets[nq] = the total)
__global__ void k_range_query_offsets(const long long* __restrict__ toffsets, const long long* __restrict__ item_off,
                                      int nq, long long* __restrict__ offsets) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q <= nq) offsets[q] = item_off[toffsets[q]];
}

__global__ __launch_bounds__(64) void k_range_item_fill(RangeArgs a, long long ntask,
                                                        const long long* __restrict__ item_off) {
    extern __shared__ int lds_dyn[];
    const int lane = threadIdx.x;
    const long long ti = (long long)blockIdx.x * 64 + lane;
    if (ti >= ntask) return;
    const int4 t = a.tasks[ti];
    bool ok;
    const Pool P = range_pool(a, t.y, &ok);
    if (!ok) return;
    const long long end = range_item(a, P, t.y, t.x, t.z, lds_dyn, lane, item_off[ti], item_off[ti + 1], true);
    if (end != item_off[ti + 1]) atomicOr(a.overflow, 8);  // counts and walk disagree
}

void launch_range_top(const RangeArgs& a, int list, hipStream_t s) {
    if (a.nq <= 0) return;
    hipLaunchKernelGGL(k_range_top, dim3(blocks_for(a.nq, 64)), dim3(64), (size_t)a.stack * 64 * sizeof(int), s, a,
                       list);
}
void launch_range_item_count(const RangeArgs& a, long long ntask, long long* counts, hipStream_t s) {
    if (ntask <= 0) return;
    hipLaunchKernelGGL(k_range_item_count, dim3((unsigned)blocks_for(ntask, 64)), dim3(64),
                       (size_t)a.stack * 64 * sizeof(int), s, a, ntask, counts);
}
void launch_range_query_offsets(const long long* toffsets, const long long* item_off, int nq, long long* offsets,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_range_query_offsets, dim3(blocks_for(nq + 1, 256)), dim3(256), 0, s, toffsets, item_off, nq,
                       offsets);
}
void launch_range_item_fill(const RangeArgs& a, long long ntask, const long long* item_off, hipStream_t s) {
    if (ntask <= 0) return;
    hipLaunchKernelGGL(k_range_item_fill, dim3((unsigned)blocks_for(ntask, 64)), dim3(64),
                       (size_t)a.stack * 64 * sizeof(int), s, a, ntask, item_off);
}

}  // namespace ikd

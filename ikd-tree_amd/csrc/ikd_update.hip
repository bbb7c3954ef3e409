// ikd_update.hip — device kernels for the update path: lazy delete labels,
// bottom-up refit (Update), Criterion_Check, subtree enumeration (flatten),
// batched insertion (Add_by_point), box and point deletion, and the voxel
// downsample fold of Add_Points.
//
// Lazy labels.  As in the reference (Delete_by_range, cpp:656-669), a box
// that covers a whole subtree only marks the subtree root tree_deleted; the
// nodes below keep stale labels and are never read by a traversal, because
// every traversal stops at a tree_deleted node.  The only walk that enters a
// deleted subtree is insertion (Add_by_point descends through it, cpp:830-
// 858); it pushes the labels down along its path exactly as Push_Down does
// (cpp:1109-1181), so the refit above the new leaf sees correct children.
#include "ikd_internal.h"

namespace ikd {

static inline int blocks_for(long long n, int bs) { return (int)((n + bs - 1) / bs); }

__device__ __forceinline__ uint32_t* link_ptr(NodeHot* hot, int n) {
    return reinterpret_cast<uint32_t*>(&hot[n].f[3]);
}
__device__ __forceinline__ uint32_t load_link(const NodeHot* hot, int n) { return fbits(hot[n].f[3]); }

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

// std::min / std::max as the reference's Update uses them (cpp:1204-1297):
// min(a,b) = b<a ? b : a, max(a,b) = a<b ? b : a.
__device__ __forceinline__ float smin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return a < b ? b : a; }

// ============================================================ refit / Update
__global__ void k_mark_dirty(Pool pool, const int* __restrict__ seeds, const int* nseeds_dev, int bound,
                             int stop_at_tdel, int* __restrict__ dirty, int* __restrict__ pend, DevHdr* hdr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bound) return;
    if (nseeds_dev && i >= *nseeds_dev) return;
    int v = seeds[i];
    if (v < ROOT) return;
    if (stop_at_tdel) {
        // a seed under a subtree that this batch deleted whole is hidden:
        // refitting it (or its covered ancestor) would read stale labels.
        int u = v;
        for (int step = 0;; ++step) {
            if (pool.meta[u] & M_TDEL) return;
            if (u == ROOT) break;
            u = pool.par[u];
            if (step > WALK_LIMIT || u < ROOT) {
                atomicOr(&hdr->overflow, 16);
                return;
            }
        }
    }
    int u = v;
    for (int step = 0;; ++step) {
        uint32_t old = atomicOr(&pool.meta[u], M_DIRTY);
        if (old & M_DIRTY) break;
        int pos = atomicAdd(&hdr->dirty_count, 1);
        dirty[pos] = u;
        pend[u] = 0;
        if (u == ROOT) break;
        u = pool.par[u];
        if (step > WALK_LIMIT || u < ROOT) {
            atomicOr(&hdr->overflow, 16);
            return;
        }
    }
}

// Update (cpp:1183-1323) of one node from its children, then
// Criterion_Check (cpp:1089-1107).
__device__ void refit_node(Pool pool, Params prm, int u, int* viol, int* defer, DevHdr* hdr) {
    const NodeHot& h = pool.hot[u];
    uint32_t lk = fbits(h.f[3]);
    uint32_t m = pool.meta[u];
    bool pdel = (m & M_PDEL) != 0, pds = (m & M_PDS) != 0;
    int size = 1, inv = pdel ? 1 : 0, dd = pds ? 1 : 0;
    bool all_td = true, all_tds = true, any_td = false;
    int pair = (int)lk_pair(lk);
    bool has[2] = {(lk & LK_HASL) != 0, (lk & LK_HASR) != 0};
    bool ctd[2] = {false, false};
    int csize[2] = {0, 0};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (!has[s]) continue;
        int c = 2 * pair + s;
        int4 cc = pool.cnt[c];
        uint32_t mc = pool.meta[c];
        size += cc.x;
        inv += cc.y;
        dd += cc.z;
        csize[s] = cc.x;
        ctd[s] = (mc & M_TDEL) != 0;
        all_td = all_td && ctd[s];
        all_tds = all_tds && ((mc & M_TDS) != 0);
        any_td = any_td || ctd[s];
    }
    bool tdel = all_td && pdel;
    bool tds = all_tds && pds;
    // a node Delete_Points deleted and no later point of the call passed
    // through: its range is the one of its last Update, own point live
    const bool frozen = (m & M_FROZEN) != 0;
    const bool rdel = frozen ? false : pdel;
    bool everything = (frozen ? false : tdel) || (!any_td && !rdel);
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (has[s] && (everything || !ctd[s])) {
            const float* b = h.f + 4 + 6 * s;
            lo[0] = smin(lo[0], b[0]); hi[0] = smax(hi[0], b[3]);
            lo[1] = smin(lo[1], b[1]); hi[1] = smax(hi[1], b[4]);
            lo[2] = smin(lo[2], b[2]); hi[2] = smax(hi[2], b[5]);
        }
    }
    if (everything || !rdel) {
        lo[0] = smin(lo[0], h.f[0]); hi[0] = smax(hi[0], h.f[0]);
        lo[1] = smin(lo[1], h.f[1]); hi[1] = smax(hi[1], h.f[1]);
        lo[2] = smin(lo[2], h.f[2]); hi[2] = smax(hi[2], h.f[2]);
    }
    int depth = pool.cnt[u].w;
    pool.cnt[u] = make_int4(size, inv, dd, depth);
    if (u == ROOT && size > 3 && !frozen) {
        // Update's root statistics, "for paper data" (cpp:1315-1321)
        const float bal = (float)(has[0] ? csize[0] : csize[1]) / (float)(size - 1);
        hdr->alpha_del = (float)inv / (float)size;
        hdr->alpha_bal = ((double)bal >= 0.5 - EPSS) ? bal : 1.0f - bal;
    }
    uint32_t nm = (m & ~(M_TDEL | M_TDS | M_DIRTY | M_VIOL | M_TOP | M_FROZEN | M_KEEPBOX)) | (tdel ? M_TDEL : 0u) |
                  (tds ? M_TDS : 0u);
    if (size > MIN_UNBALANCED_SIZE && !frozen) {
        int son = has[0] ? csize[0] : csize[1];
        float del_eval = (float)inv / (float)size;
        float bal_eval = (float)son / (float)(size - 1);
        bool violates = del_eval > prm.delete_param || bal_eval > prm.balance_param ||
                        bal_eval < 1.0f - prm.balance_param;
        if (violates && prm.defer_points > 0 && size - inv > prm.defer_points && defer) {
            // Large subtree: keep its true counts and list it (once) for a
            // later rebuild, as the reference hands subtrees of >= 1500 nodes
            // to its background rebuild thread (cpp:627-633).
            violates = false;
            if (!(m & M_DEFER)) {
                int pos = atomicAdd(&hdr->defer_count, 1);
                if (pos < DEFER_CAP) {
                    defer[pos] = u;
                    nm |= M_DEFER;
                }
            }
        }
        if (violates) {
            nm |= M_VIOL;
            int pos = atomicAdd(&hdr->viol_count, 1);
            viol[2 * pos] = u;
            viol[2 * pos + 1] = size;
            // The subtree will be rebuilt from its live points: hand the parent
            // the post-rebuild counts now, so one bottom-up pass already sees
            // the cascade the rebuild causes above (the reference re-checks the
            // parent after each rebuild on its way up, cpp:860-863).
            pool.cnt[u] = make_int4(size - inv, 0, 0, depth);
            nm &= ~M_TDS;
        }
    }
    pool.meta[u] = nm;
    int p = pool.par[u];
    float box[6] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
    if (!(m & M_KEEPBOX)) write_child_box(pool.hot, p, u & 1, box);
    uint32_t tdbit = lk_td(u & 1);
    if (tdel) atomicOr(link_ptr(pool.hot, p), tdbit);
    else atomicAnd(link_ptr(pool.hot, p), ~tdbit);
}

// Single-pass bottom-up refit.  The dirty set is closed under ancestors, so
// every dirty node but the root has a dirty parent: count each node's dirty
// children, start one lane at every dirty node without any, and let the lane
// that finishes a node's last dirty child go on to refit the node itself
// (the fences order each child's writes before its parent reads them).  The
// same order as one launch per depth level, in three launches.
__global__ void k_refit_count(Pool pool, const int* __restrict__ dirty, DevHdr* hdr, int* __restrict__ pend) {
    const int n = hdr->dirty_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int u = dirty[i];
        if (u != ROOT) atomicAdd(&pend[pool.par[u]], 1);
    }
}

__global__ void k_refit_starts(const int* __restrict__ dirty, DevHdr* hdr, const int* __restrict__ pend,
                               int* __restrict__ start) {
    const int n = hdr->dirty_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        start[i] = pend[dirty[i]] == 0 ? 1 : 0;
}

__global__ void k_refit_up(Pool pool, Params prm, const int* __restrict__ dirty, const int* __restrict__ start,
                           DevHdr* hdr, int* __restrict__ pend, int* __restrict__ viol, int* __restrict__ defer) {
    const int n = hdr->dirty_count;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!start[i]) continue;
        int u = dirty[i];
        for (int step = 0; step <= WALK_LIMIT; ++step) {
            refit_node(pool, prm, u, viol, defer, hdr);
            if (u == ROOT) break;
            const int p = pool.par[u];
            // release this node's results to the lane that finishes the
            // parent; that lane acquires them after the last decrement
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            if (atomicSub(&pend[p], 1) != 1) break;  // a sibling is still on its way
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            u = p;
        }
    }
}

void launch_refit_up(Pool pool, Params prm, const int* dirty, int* pend, int* start, DevHdr* hdr, int* viol,
                     int* defer, hipStream_t s) {
    hipLaunchKernelGGL(k_refit_count, dim3(512), dim3(256), 0, s, pool, dirty, hdr, pend);
    hipLaunchKernelGGL(k_refit_starts, dim3(512), dim3(256), 0, s, dirty, hdr, pend, start);
    hipLaunchKernelGGL(k_refit_up, dim3(512), dim3(256), 0, s, pool, prm, dirty, start, hdr, pend, viol, defer);
}

// Deferred rebuilds: an entry is still due if its node is linked into the
// tree (every parent link on the way to the root points back at it), no
// subtree on the way is deleted, no ancestor is itself listed (the ancestor's
// rebuild covers it) and it still fails Criterion_Check on its current counts.
// keep[i] = 1 for due entries, with info (node, live, depth, parent) and the
// node count before the rebuild.
__global__ void k_defer_check(Pool pool, Params prm, const int* __restrict__ defer, int n, int4* __restrict__ info,
                              int* __restrict__ before, int* __restrict__ keep) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keep[i] = 0;
    int u = defer[i];
    uint32_t mu = pool.meta[u];
    if (!(mu & M_DEFER) || (mu & M_TDEL)) return;
    int v = u;
    for (int step = 0; step <= WALK_LIMIT; ++step) {
        int p = pool.par[v];
        if (p < 0) return;
        uint32_t lk = fbits(pool.hot[p].f[3]);
        int side = v & 1;
        if (!(lk & lk_has(side)) || (lk & lk_td(side)) || (int)(2 * lk_pair(lk)) + side != v) return;
        if (p == SENTINEL) break;
        uint32_t mp = pool.meta[p];
        if (mp & (M_TDEL | M_DEFER)) return;
        v = p;
    }
    if (v != ROOT) return;
    int4 c = pool.cnt[u];
    int size = c.x, inv = c.y;
    if (size <= MIN_UNBALANCED_SIZE) return;
    uint32_t lk = fbits(pool.hot[u].f[3]);
    int pair = (int)lk_pair(lk);
    int son = (lk & LK_HASL) ? pool.cnt[2 * pair].x : ((lk & LK_HASR) ? pool.cnt[2 * pair + 1].x : 0);
    float del_eval = (float)inv / (float)size;
    float bal_eval = (float)son / (float)(size - 1);
    if (!(del_eval > prm.delete_param || bal_eval > prm.balance_param || bal_eval < 1.0f - prm.balance_param)) return;
    keep[i] = 1;
    info[i] = make_int4(u, size - inv, c.w, pool.par[u]);
    before[i] = size;
}

__global__ void k_defer_clear(Pool pool, const int* __restrict__ defer, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAnd(&pool.meta[defer[i]], ~M_DEFER);
}

void launch_defer_check(Pool pool, Params prm, const int* defer, int n, int4* info, int* before, int* keep,
                        hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_defer_check, dim3(blocks_for(n, 256)), dim3(256), 0, s, pool, prm, defer, n, info, before,
                       keep);
}
void launch_defer_clear(Pool pool, const int* defer, int n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_defer_clear, dim3(blocks_for(n, 256)), dim3(256), 0, s, pool, defer, n);
}

// A violator is rebuilt only if no ancestor violates too: the reference
// checks bottom-up on the way out of the recursion and ends up rebuilding
// the highest violating node of the path (cpp:862-863, 706-707).
__global__ void k_find_top(Pool pool, const int* __restrict__ viol, DevHdr* hdr, int* __restrict__ top) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hdr->viol_count) return;
    int v = viol[2 * i];
    int u = v;
    for (int step = 0; u != ROOT; ++step) {
        u = pool.par[u];
        if (step > WALK_LIMIT || u < ROOT) {
            atomicOr(&hdr->overflow, 16);
            return;
        }
        if (pool.meta[u] & M_VIOL) return;
    }
    atomicOr(&pool.meta[v], M_TOP);
    int pos = atomicAdd(&hdr->top_count, 1);
    top[2 * pos] = v;
    top[2 * pos + 1] = viol[2 * i + 1];  // nodes in the subtree before the rebuild
}

__global__ void k_clear_viol(Pool pool, const int* __restrict__ viol, DevHdr* hdr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hdr->viol_count) return;
    atomicAnd(&pool.meta[viol[2 * i]], ~M_VIOL);
}

// A root the build just made is a new node: BuildTree's Update of it sets the
// statistics when it has more than 3 nodes (cpp:1315); a smaller new root
// has none (0 here, the reference leaves them unset).
__global__ void k_root_alpha_fresh(Pool pool, DevHdr* hdr) {
    float ab = 0.0f, ad = 0.0f;
    const uint32_t lk0 = fbits(pool.hot[SENTINEL].f[3]);
    if (lk0 & LK_HASL) {
        const int4 c = pool.cnt[ROOT];
        if (c.x > 3) {
            const uint32_t lk = fbits(pool.hot[ROOT].f[3]);
            const int pair = (int)lk_pair(lk);
            const int son = (lk & LK_HASL) ? pool.cnt[2 * pair].x : pool.cnt[2 * pair + 1].x;
            const float bal = (float)son / (float)(c.x - 1);
            ad = (float)c.y / (float)c.x;
            ab = ((double)bal >= 0.5 - EPSS) ? bal : 1.0f - bal;
        }
    }
    hdr->alpha_bal = ab;
    hdr->alpha_del = ad;
}

void launch_root_alpha_fresh(Pool pool, DevHdr* hdr, hipStream_t s) {
    hipLaunchKernelGGL(k_root_alpha_fresh, dim3(1), dim3(1), 0, s, pool, hdr);
}

void launch_mark_dirty(Pool pool, const int* seeds, const int* nseeds_dev, int bound, int stop_at_tdel,
                       int* dirty, int* pend, DevHdr* hdr, hipStream_t s) {
    if (bound <= 0) return;
    hipLaunchKernelGGL(k_mark_dirty, dim3(blocks_for(bound, 256)), dim3(256), 0, s, pool, seeds, nseeds_dev, bound,
                       stop_at_tdel, dirty, pend, hdr);
}
void launch_find_top(Pool pool, const int* viol, int bound, DevHdr* hdr, int* top, hipStream_t s) {
    if (bound <= 0) return;
    hipLaunchKernelGGL(k_find_top, dim3(blocks_for(bound, 256)), dim3(256), 0, s, pool, viol, hdr, top);
}
void launch_clear_viol(Pool pool, const int* viol, int bound, DevHdr* hdr, hipStream_t s) {
    if (bound <= 0) return;
    hipLaunchKernelGGL(k_clear_viol, dim3(blocks_for(bound, 256)), dim3(256), 0, s, pool, viol, hdr);
}

// ======================================================= enumerate (flatten)
// Live points of several subtrees in the reference's flatten order (own
// point, left subtree, right subtree; ikd_Tree.cpp:1326-1333), one depth
// level per launch.  A frontier entry (node, base, segment, flags) knows where
// the first live point of its subtree goes: its own point takes base when
// live, the left subtree starts after it and the right subtree after the
// left's live count (TreeSize - invalid_point_num of the left child).
// Tree-deleted subtrees hold no live points: without a record they are not
// entered; with one (rec) they are walked too, flags bit 0 marking a deleted
// subtree and bit 1 a downsample-deleted one, and every deleted point the
// downsample did not delete is recorded, as flatten's DELETE_POINTS_REC mode
// does during Rebuild (cpp:1338-1341, 638).  Three rotating frontier counters
// let a level reset the counter of the level after next.
struct EnumStep {
    int pushes;       // children pushed
    int4 child[2];
};

__device__ __forceinline__ void enum_record(const Pool& pool, int u, DevHdr* hdr, float4* rec, int rec_cap) {
    const int pos = atomicAdd(&hdr->removed_count, 1);
    if (pos < rec_cap) {
        const float4 p = *reinterpret_cast<const float4*>(pool.hot[u].f);
        rec[pos] = make_float4(p.x, p.y, p.z, bitsf(pool.ids[u]));
    } else {
        atomicOr(&hdr->overflow, 32);
    }
}

// One frontier entry: writes its own live point (returns 1) or records it,
// and lists the children to push.
__device__ __forceinline__ int enum_visit(const Pool& pool, int4 e, float4* __restrict__ out_pts,
                                          int* __restrict__ out_seg, DevHdr* hdr, int out_cap, float4* rec,
                                          int rec_cap, EnumStep& st) {
    st.pushes = 0;
    const int u = e.x;
    const uint32_t m = pool.meta[u];
    const bool dead = (e.w & 1) || (m & M_TDEL);
    const bool dsub = (e.w & 2) || (m & M_TDS);
    if (dead && !rec) return 0;
    const uint32_t lk = load_link(pool.hot, u);
    const int pair = (int)lk_pair(lk);
    if (dead) {
        // a deleted subtree: every node is deleted; walk it for the record
        if (!(m & M_PDS) && !(e.w & 2)) enum_record(pool, u, hdr, rec, rec_cap);
#pragma unroll
        for (int s = 0; s < 2; ++s)
            if (lk & lk_has(s)) st.child[st.pushes++] = make_int4(2 * pair + s, 0, e.z, 1 | (dsub ? 2 : 0));
        return 0;
    }
    int next = e.y, wrote = 0;
    if (!(m & M_PDEL)) {
        if (next < out_cap) {
            const float4 p = *reinterpret_cast<const float4*>(pool.hot[u].f);
            out_pts[next] = make_float4(p.x, p.y, p.z, bitsf(pool.ids[u]));
            out_seg[next] = e.z;
            wrote = 1;
        } else {
            atomicOr(&hdr->overflow, 32);
        }
        ++next;
    } else if (rec && !(m & M_PDS)) {
        enum_record(pool, u, hdr, rec, rec_cap);
    }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (!(lk & lk_has(s))) continue;
        const int c = 2 * pair + s;
        if (lk & lk_td(s)) {
            if (rec) st.child[st.pushes++] = make_int4(c, 0, e.z, 1);
            continue;
        }
        const int4 cc = pool.cnt[c];
        st.child[st.pushes++] = make_int4(c, next, e.z, 0);
        next += cc.x - cc.y;
    }
    return wrote;
}

__global__ void k_enum_preorder(Pool pool, const int4* __restrict__ front, int lvl, int4* __restrict__ front_next,
                                float4* __restrict__ out_pts, int* __restrict__ out_seg, DevHdr* hdr, int bound,
                                int out_cap, float4* __restrict__ rec, int rec_cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) hdr->front3[(lvl + 2) % 3] = 0;
    const int nf = min(hdr->front3[lvl % 3], bound);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nf; i += gridDim.x * blockDim.x) {
        EnumStep st;
        if (enum_visit(pool, front[i], out_pts, out_seg, hdr, out_cap, rec, rec_cap, st))
            atomicAdd(&hdr->rec_count, 1);
        for (int c = 0; c < st.pushes; ++c) {
            const int pos = atomicAdd(&hdr->front3[(lvl + 1) % 3], 1);
            if (pos < bound) front_next[pos] = st.child[c];
            else atomicOr(&hdr->overflow, 32);
        }
    }
}

void launch_enum_preorder(Pool pool, const int4* front, int lvl, int4* front_next, float4* out_pts, int* out_seg,
                          DevHdr* hdr, int bound, int out_cap, float4* rec, int rec_cap, hipStream_t s) {
    // a grid-stride loop: the frontier size is known only on the device
    const int blocks = std::min(blocks_for(bound > 0 ? bound : 1, 256), 2048);
    hipLaunchKernelGGL(k_enum_preorder, dim3(blocks), dim3(256), 0, s, pool, front, lvl, front_next, out_pts, out_seg,
                       hdr, bound, out_cap, rec, rec_cap);
}

// The same walk by one workgroup over every level, for subtrees of a few
// hundred nodes: the levels are separated by workgroup barriers instead of
// launches, and the walk ends as soon as a level is empty.  f0 holds the
// n_roots root entries; f0 and f1 alternate as frontiers of `bound` entries.
__global__ __launch_bounds__(1024) void k_enum_preorder_block(Pool pool, int4* __restrict__ f0, int4* __restrict__ f1,
                                                              int n_roots, float4* __restrict__ out_pts,
                                                              int* __restrict__ out_seg, DevHdr* hdr, int bound,
                                                              int out_cap, int max_levels, float4* __restrict__ rec,
                                                              int rec_cap) {
    __shared__ int next_count, written;
    if (threadIdx.x == 0) written = 0;
    int nf = min(n_roots, bound);
    int4* fin = f0;
    int4* fout = f1;
    int lvl = 0;
    for (; lvl < max_levels && nf > 0; ++lvl) {
        if (threadIdx.x == 0) next_count = 0;
        __syncthreads();
        int mine = 0;
        for (int i = threadIdx.x; i < nf; i += blockDim.x) {
            EnumStep st;
            mine += enum_visit(pool, fin[i], out_pts, out_seg, hdr, out_cap, rec, rec_cap, st);
            for (int c = 0; c < st.pushes; ++c) {
                const int pos = atomicAdd(&next_count, 1);
                if (pos < bound) fout[pos] = st.child[c];
                else atomicOr(&hdr->overflow, 32);
            }
        }
        if (mine) atomicAdd(&written, mine);
        __syncthreads();
        nf = min(next_count, bound);
        int4* tmp = fin;
        fin = fout;
        fout = tmp;
        __syncthreads();  // every thread has read next_count before it is reset
    }
    if (threadIdx.x == 0) {
        atomicAdd(&hdr->rec_count, written);
        if (nf > 0) atomicOr(&hdr->overflow, 32);  // deeper than the levels allowed
    }
}

void launch_enum_preorder_block(Pool pool, int4* f0, int4* f1, int n_roots, float4* out_pts, int* out_seg,
                                DevHdr* hdr, int bound, int out_cap, int max_levels, float4* rec, int rec_cap,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_enum_preorder_block, dim3(1), dim3(1024), 0, s, pool, f0, f1, n_roots, out_pts, out_seg, hdr,
                       bound, out_cap, max_levels, rec, rec_cap);
}

// The same walk for many small subtrees in one launch: workgroup g walks the
// roots groups[g].x .. groups[g].x + groups[g].y - 1 (entries of `roots`)
// level by level in its own frontier area, entries [groups[g].z,
// groups[g].z + groups[g].w) of f0 and f1 (the group's node count, which
// bounds every level).  Output positions come with the root entries, so the
// groups never meet.
__global__ __launch_bounds__(256) void k_enum_groups(Pool pool, const int4* __restrict__ roots,
                                                     const int4* __restrict__ groups, int4* __restrict__ f0,
                                                     int4* __restrict__ f1, float4* __restrict__ out_pts,
                                                     int* __restrict__ out_seg, DevHdr* hdr, int out_cap,
                                                     int max_levels, float4* __restrict__ rec, int rec_cap) {
    __shared__ int next_count, written;
    const int4 g = groups[blockIdx.x];
    const int bound = g.w;
    int4* fin = f0 + g.z;
    int4* fout = f1 + g.z;
    if (threadIdx.x == 0) written = 0;
    int nf = min(g.y, bound);
    for (int i = threadIdx.x; i < nf; i += blockDim.x) fin[i] = roots[g.x + i];
    if (g.y > bound && threadIdx.x == 0) atomicOr(&hdr->overflow, 32);
    __syncthreads();
    for (int lvl = 0; lvl < max_levels && nf > 0; ++lvl) {
        if (threadIdx.x == 0) next_count = 0;
        __syncthreads();
        int mine = 0;
        for (int i = threadIdx.x; i < nf; i += blockDim.x) {
            EnumStep st;
            mine += enum_visit(pool, fin[i], out_pts, out_seg, hdr, out_cap, rec, rec_cap, st);
            for (int c = 0; c < st.pushes; ++c) {
                const int pos = atomicAdd(&next_count, 1);
                if (pos < bound) fout[pos] = st.child[c];
                else atomicOr(&hdr->overflow, 32);
            }
        }
        if (mine) atomicAdd(&written, mine);
        __syncthreads();
        nf = min(next_count, bound);
        int4* tmp = fin;
        fin = fout;
        fout = tmp;
        __syncthreads();  // every thread has read next_count before it is reset
    }
    if (threadIdx.x == 0) {
        atomicAdd(&hdr->rec_count, written);
        if (nf > 0) atomicOr(&hdr->overflow, 32);  // deeper than the levels allowed
    }
}

void launch_enum_groups(Pool pool, const int4* roots, const int4* groups, int n_groups, int4* f0, int4* f1,
                        float4* out_pts, int* out_seg, DevHdr* hdr, int out_cap, int max_levels, float4* rec,
                        int rec_cap, hipStream_t s) {
    if (n_groups <= 0) return;
    hipLaunchKernelGGL(k_enum_groups, dim3(n_groups), dim3(256), 0, s, pool, roots, groups, f0, f1, out_pts, out_seg,
                       hdr, out_cap, max_levels, rec, rec_cap);
}

// ================================================================== insert
// Marks node u as wholly deleted, as Push_Down writes a child of a
// tree_deleted parent (cpp:1125-1131).
__device__ void push_deleted(Pool pool, int u, bool ds) {
    uint32_t bits = M_PDEL | M_TDEL | (ds ? (M_PDS | M_TDS) : 0u);
    atomicOr(&pool.meta[u], bits);
    atomicOr(link_ptr(pool.hot, u), LK_PDEL);
    int sz = pool.cnt[u].x;
    pool.cnt[u].y = sz;
    if (ds) pool.cnt[u].z = sz;
}

// Add_by_point's descent (cpp:830-858): p[axis] < node[axis] goes left,
// everything else right, through deleted nodes too, down to an empty child
// slot.  Emits (node << 1 | side) of that slot.
__global__ void k_insert_descend(Pool pool, const float4* __restrict__ pts, int n,
                                 unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals, DevHdr* hdr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 P = pts[i];
    vals[i] = (uint32_t)i;
    uint32_t lk0 = load_link(pool.hot, SENTINEL);
    if (!(lk0 & LK_HASL)) {
        keys[i] = 0ull;  // empty tree: the slot is the sentinel's left child (the root)
        return;
    }
    int u = ROOT;
    bool del = false, ds = false;
    int side = 0;
    for (int step = 0;; ++step) {
        if (step > WALK_LIMIT) {
            atomicOr(&hdr->overflow, 64);
            break;
        }
        uint32_t m = pool.meta[u];
        uint32_t lk = load_link(pool.hot, u);
        if (del || (m & M_TDEL)) {
            del = true;
            ds = ds || (m & M_TDS);
            push_deleted(pool, u, ds);
            int pair = (int)lk_pair(lk);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (lk & lk_has(s)) {
                    push_deleted(pool, 2 * pair + s, ds);
                    atomicOr(link_ptr(pool.hot, u), lk_td(s));
                }
            }
        }
        int axis = (int)(m & M_AXIS);
        const float* f = pool.hot[u].f;
        float pc = axis == 0 ? P.x : (axis == 1 ? P.y : P.z);
        side = pc < f[axis] ? 0 : 1;
        if (!(lk & lk_has(side))) break;
        u = 2 * (int)lk_pair(lk) + side;
    }
    keys[i] = ((unsigned long long)(uint32_t)u << 1) | (unsigned long long)side;
}

__global__ void k_insert_groups(const unsigned long long* __restrict__ keys, int n, int* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// One segment per attach slot; a slot under a childless node needs a new
// sibling pair (one per node even when both sides receive points).
__global__ void k_insert_segments(Pool pool, const unsigned long long* __restrict__ keys,
                                  const int* __restrict__ gscan, const int* __restrict__ gflag, int n,
                                  int4* __restrict__ seg, int* __restrict__ seg_depth, int* __restrict__ seg_of,
                                  int* __restrict__ attach, int* __restrict__ pair_flag, DevHdr* hdr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int g = gscan[i] + gflag[i] - 1;
    seg_of[i] = g;
    unsigned long long key = keys[i];
    int X = (int)(key >> 1);
    int side = (int)(key & 1ull);
    if (gflag[i]) {
        seg[g].x = i;
        seg[g].w = X;
        attach[g] = (int)key;
        seg_depth[g] = X == SENTINEL ? 0 : pool.cnt[X].w + 1;
        uint32_t lk = load_link(pool.hot, X);
        bool first_of_node = (i == 0) || ((int)(keys[i - 1] >> 1) != X);
        pair_flag[g] = (lk_pair(lk) == 0 && first_of_node) ? 1 : 0;
        (void)side;
    }
    if (i == n - 1 || keys[i + 1] != key) seg[g].y = i;
    if (i == n - 1) hdr->group_count = g + 1;
}

__global__ void k_insert_pairs(Pool pool, const int* __restrict__ attach, const int* __restrict__ pair_flag,
                               const int* __restrict__ pair_scan, DevHdr* hdr, int bound, int4* __restrict__ seg) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    int ng = hdr->group_count;
    if (g == 0 && ng > 0) hdr->new_pairs = pair_scan[ng - 1] + pair_flag[ng - 1];
    if (g >= bound || g >= ng) return;
    int key = attach[g];
    int X = key >> 1, side = key & 1;
    uint32_t lk = load_link(pool.hot, X);
    int pair = (int)lk_pair(lk);
    if (pair == 0) pair = hdr->pair_used + (pair_flag[g] ? pair_scan[g] : pair_scan[g] - 1);
    seg[g].z = 2 * pair + side;
}

__global__ void k_commit_pairs(DevHdr* hdr) {
    hdr->pair_used += hdr->new_pairs;
    hdr->new_pairs = 0;
}

// After the build: hook each new subtree into its slot.
__global__ void k_insert_link(Pool pool, const int* __restrict__ attach, const int4* __restrict__ seg, DevHdr* hdr,
                              int bound) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= bound || g >= hdr->group_count) return;
    int key = attach[g];
    int X = key >> 1, side = key & 1;
    int slot = seg[g].z;
    if (seg[g].y == seg[g].x) {
        // a single point becomes a leaf split on its parent's axis + 1
        // (Add_by_point, cpp:823; the sentinel's axis is 0)
        const uint32_t axis = ((pool.meta[X] & M_AXIS) + 1) % 3;
        pool.meta[slot] = (pool.meta[slot] & ~M_AXIS) | axis;
    }
    uint32_t bits = lk_has(side) | ((uint32_t)(slot >> 1) << LK_PAIR_SHIFT);
    atomicOr(link_ptr(pool.hot, X), bits);
    atomicAnd(link_ptr(pool.hot, X), ~lk_td(side));
}

void launch_insert_descend(Pool pool, const float4* pts, int n, unsigned long long* keys, uint32_t* vals,
                           DevHdr* hdr, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_insert_descend, dim3(blocks_for(n, 256)), dim3(256), 0, s, pool, pts, n, keys, vals, hdr);
}
void launch_insert_groups(const unsigned long long* keys, int n, int* flag, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_insert_groups, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, flag);
}
void launch_insert_segments2(Pool pool, const unsigned long long* keys, const int* gscan, const int* gflag, int n,
                             int4* seg, int* seg_depth, int* seg_of, int* attach, int* pair_flag, DevHdr* hdr,
                             hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_insert_segments, dim3(blocks_for(n, 256)), dim3(256), 0, s, pool, keys, gscan, gflag, n, seg,
                       seg_depth, seg_of, attach, pair_flag, hdr);
}
void launch_insert_pairs(Pool pool, const int* attach, const int* pair_flag, const int* pair_scan, DevHdr* hdr,
                         int bound, int4* seg, hipStream_t s) {
    if (bound <= 0) return;
    hipLaunchKernelGGL(k_insert_pairs, dim3(blocks_for(bound, 256)), dim3(256), 0, s, pool, attach, pair_flag,
                       pair_scan, hdr, bound, seg);
    hipLaunchKernelGGL(k_commit_pairs, dim3(1), dim3(1), 0, s, hdr);
}
void launch_insert_link(Pool pool, const int* attach, const int4* seg, DevHdr* hdr, int bound, hipStream_t s) {
    if (bound <= 0) return;
    hipLaunchKernelGGL(k_insert_link, dim3(blocks_for(bound, 256)), dim3(256), 0, s, pool, attach, seg, hdr, bound);
}

// ============================================================ box deletion
// Delete_by_range (cpp:647-710) for many boxes at once.
// Pass 1 records what each box deletes: COVER for a subtree whose AABB lies
// inside the half-open box (the subtree root takes the lazy label), POINT
// for a single live node point inside it.
constexpr uint32_t REC_COVER = 0x80000000u;
constexpr uint32_t REC_PARTIAL = 0x40000000u;  // a node the box cuts: Updated, labels unchanged

// Pass 1 runs breadth-first over all boxes at once: a frontier of (node, box)
// pairs, one lane per pair, so a box that cuts through much of the tree (a
// sliding-window trim) is walked by many lanes instead of one.  Each launch
// (a step) takes every pair BOXDEL_LEVELS tree levels down, depth first with
// a small stack per lane, and hands the pairs below on as the next frontier;
// front3[step % 3] holds the step's frontier size.  Every (node, box) pair the
// box cuts is visited once, whatever the step depth.
__global__ void k_box_delete_init(Pool pool, const float* __restrict__ boxes, int nb, int2* __restrict__ front,
                                  unsigned int* __restrict__ recs, DevHdr* hdr, int cap) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    uint32_t lk0 = load_link(pool.hot, SENTINEL);
    if (!(lk0 & LK_HASL) || (lk0 & LK_TDL)) return;
    const float* box = boxes + 6 * b;
    const float* rb = pool.hot[SENTINEL].f + 4;
    if (box_disjoint(box, rb)) return;
    if (box_covers(box, rb)) {
        int pos = atomicAdd(&hdr->rec_count, 1);
        if (pos < cap) recs[pos] = REC_COVER | (uint32_t)ROOT;
        return;
    }
    int pos = atomicAdd(&hdr->front3[0], 1);
    front[pos] = make_int2(ROOT, b);  // nb <= the frontier capacity
}

__global__ void k_box_delete_level(Pool pool, const float* __restrict__ boxes, const int2* __restrict__ front,
                                   int step, int2* __restrict__ front_next, unsigned int* __restrict__ recs,
                                   DevHdr* hdr, int fcap, int cap) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) hdr->front3[(step + 2) % 3] = 0;
    const int count = min(hdr->front3[step % 3], fcap);
    for (int i = tid; i < count; i += gridDim.x * blockDim.x) {
        const int2 e = front[i];
        const float* box = boxes + 6 * e.y;
        // depth-first below e.x: a node at local depth d pushes its cut
        // children at d + 1 < BOXDEL_LEVELS (at most one pending sibling per
        // level, so BOXDEL_LEVELS entries hold the stack)
        int stk[BOXDEL_LEVELS], sdep[BOXDEL_LEVELS];
        int top = 0;
        int u = e.x, d = 0;
        while (true) {
            float f[16];
            load_record(pool.hot[u].f, f);
            const uint32_t lk = fbits(f[3]);
            // a node the box cuts is Updated on the way out (cpp:703-704)
            int pos = atomicAdd(&hdr->rec_count, 1);
            if (pos < cap) recs[pos] = REC_PARTIAL | (uint32_t)u;
            if (!(lk & LK_PDEL) && box_contains(box, f[0], f[1], f[2])) {
                pos = atomicAdd(&hdr->rec_count, 1);
                if (pos < cap) recs[pos] = (uint32_t)u;
            }
            const int pair = (int)lk_pair(lk);
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (!(lk & lk_has(s)) || (lk & lk_td(s))) continue;
                const float* cb = f + 4 + 6 * s;
                if (box_disjoint(box, cb)) continue;
                const int c = 2 * pair + s;
                if (box_covers(box, cb)) {
                    pos = atomicAdd(&hdr->rec_count, 1);
                    if (pos < cap) recs[pos] = REC_COVER | (uint32_t)c;
                } else if (d + 1 < BOXDEL_LEVELS && top < BOXDEL_LEVELS) {
                    stk[top] = c;
                    sdep[top++] = d + 1;
                } else {
                    pos = atomicAdd(&hdr->front3[(step + 1) % 3], 1);
                    if (pos < fcap) front_next[pos] = make_int2(c, e.y);
                    else hdr->front_over = 1;
                }
            }
            if (top == 0) break;
            --top;
            u = stk[top];
            d = sdep[top];
        }
    }
}

void launch_box_delete_bfs(Pool pool, const float* boxes, int nb, int2* f0, int2* f1, int fcap, int S0, int S1,
                           unsigned int* recs, DevHdr* hdr, int cap, hipStream_t s) {
    if (nb <= 0) return;
    if (S0 == 0)
        hipLaunchKernelGGL(k_box_delete_init, dim3(blocks_for(nb, 64)), dim3(64), 0, s, pool, boxes, nb, f0, recs,
                           hdr, cap);
    for (int S = S0; S < S1; ++S) {
        const int2* fin = (S & 1) ? f1 : f0;
        int2* fout = (S & 1) ? f0 : f1;
        hipLaunchKernelGGL(k_box_delete_level, dim3(512), dim3(256), 0, s, pool, boxes, fin, S, fout, recs, hdr, fcap,
                           cap);
    }
}

// Pass 2 applies the records (idempotent, so overlapping boxes are fine) and
// names the refit seeds: the node itself for POINT, the parent for COVER.
__global__ void k_box_delete_apply(Pool pool, const unsigned int* __restrict__ recs, DevHdr* hdr, int bound,
                                   int ds, int* __restrict__ seeds) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bound) return;
    if (i >= hdr->rec_count) return;
    uint32_t r = recs[i];
    int u = (int)(r & ~(REC_COVER | REC_PARTIAL));
    if (r & REC_PARTIAL) {
        seeds[i] = u;
    } else if (r & REC_COVER) {
        uint32_t bits = M_PDEL | M_TDEL | (ds ? (M_PDS | M_TDS) : 0u);
        atomicOr(&pool.meta[u], bits);
        atomicOr(link_ptr(pool.hot, u), LK_PDEL);
        int sz = pool.cnt[u].x;
        pool.cnt[u].y = sz;
        if (ds) pool.cnt[u].z = sz;
        int p = pool.par[u];
        atomicOr(link_ptr(pool.hot, p), lk_td(u & 1));
        seeds[i] = p;
    } else {
        atomicOr(&pool.meta[u], M_PDEL | (ds ? M_PDS : 0u));
        atomicOr(link_ptr(pool.hot, u), LK_PDEL);
        seeds[i] = u;
    }
}

void launch_box_delete_apply(Pool pool, const unsigned int* recs, DevHdr* hdr, int bound, int ds, int* seeds,
                             hipStream_t s) {
    if (bound <= 0) return;
    hipLaunchKernelGGL(k_box_delete_apply, dim3(blocks_for(bound, 256)), dim3(256), 0, s, pool, recs, hdr, bound, ds,
                       seeds);
}

// ====================================================== box re-insertion
// Add_by_range (cpp:762-815) for one box.  Breadth-first from the root over
// the nodes whose stored range the box cuts, entering deleted subtrees too
// (the reference does not stop at tree_deleted here).  A cut node first
// hands its deletion label to its children (Push_Down, cpp:1109-1181, which
// the device applies only along the paths it walks), gets its own point back
// if the box holds it and the downsample did not delete it (cpp:784-786),
// and becomes a refit seed (it is Updated on the way out, cpp:809-812).  A
// child whose range the box covers is recorded for the reset below; a child
// the box cuts goes to the next level.  front3[lvl % 3] holds the level's
// frontier size, rec_count the records.
__global__ void k_addbox_init(Pool pool, const float* __restrict__ box, int2* __restrict__ front,
                              unsigned int* __restrict__ recs, DevHdr* hdr) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    hdr->front3[0] = 0;
    hdr->front3[1] = 0;
    hdr->front3[2] = 0;
    hdr->rec_count = 0;
    hdr->front_over = 0;
    const uint32_t lk0 = load_link(pool.hot, SENTINEL);
    if (!(lk0 & LK_HASL)) return;
    const float* rb = pool.hot[SENTINEL].f + 4;
    if (box_disjoint(box, rb)) return;
    if (box_covers(box, rb)) {
        recs[0] = REC_COVER | (uint32_t)ROOT;
        hdr->rec_count = 1;
        return;
    }
    front[0] = make_int2(ROOT, 0);
    hdr->front3[0] = 1;
}

__global__ void k_addbox_level(Pool pool, const float* __restrict__ box, const int2* __restrict__ front, int lvl,
                               int2* __restrict__ front_next, unsigned int* __restrict__ recs, DevHdr* hdr, int fcap,
                               int cap) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) hdr->front3[(lvl + 2) % 3] = 0;
    const int count = min(hdr->front3[lvl % 3], fcap);
    for (int i = tid; i < count; i += gridDim.x * blockDim.x) {
        const int u = front[i].x;
        const uint32_t m = pool.meta[u];
        const uint32_t lk = load_link(pool.hot, u);
        const int pair = (int)lk_pair(lk);
        if (m & M_TDEL) {
            // Push_Down of a deleted node: the children take its labels
            const bool ds = (m & M_TDS) != 0;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (!(lk & lk_has(s))) continue;
                push_deleted(pool, 2 * pair + s, ds);
                atomicOr(link_ptr(pool.hot, u), lk_td(s));
            }
        }
        const float* f = pool.hot[u].f;
        if (box_contains(box, f[0], f[1], f[2])) {
            // point_deleted := point_downsample_deleted
            if (m & M_PDS) {
                atomicOr(&pool.meta[u], M_PDEL);
                atomicOr(link_ptr(pool.hot, u), LK_PDEL);
            } else {
                atomicAnd(&pool.meta[u], ~M_PDEL);
                atomicAnd(link_ptr(pool.hot, u), ~LK_PDEL);
            }
        }
        int pos = atomicAdd(&hdr->rec_count, 1);
        if (pos < cap) recs[pos] = REC_PARTIAL | (uint32_t)u;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            if (!(lk & lk_has(s))) continue;
            const float* cb = f + 4 + 6 * s;
            if (box_disjoint(box, cb)) continue;
            const int c = 2 * pair + s;
            if (box_covers(box, cb)) {
                pos = atomicAdd(&hdr->rec_count, 1);
                if (pos < cap) recs[pos] = REC_COVER | (uint32_t)c;
            } else {
                pos = atomicAdd(&hdr->front3[(lvl + 1) % 3], 1);
                if (pos < fcap) front_next[pos] = make_int2(c, 0);
                else hdr->front_over = 1;
            }
        }
    }
}

// Records -> refit seeds (cut nodes) and the first level of the reset
// (covered subtree roots, bit 0 set).  front3 must be zero.
__global__ void k_addbox_split(const unsigned int* __restrict__ recs, DevHdr* hdr, int bound, int* __restrict__ seeds,
                               int2* __restrict__ rfront) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= bound || i >= hdr->rec_count) return;
    const uint32_t r = recs[i];
    const int u = (int)(r & ~(REC_COVER | REC_PARTIAL));
    if (r & REC_COVER) {
        seeds[i] = -1;
        const int pos = atomicAdd(&hdr->front3[0], 1);
        rfront[pos] = make_int2(u, 1);  // at most one record per node: pos < bound
    } else {
        seeds[i] = u;
    }
}

// Add_by_range's covered case (cpp:777-783) and what Push_Down then hands
// down the subtree (cpp:1115-1131): every node goes back to its downsample
// state, deleted iff the downsample deleted it or a subtree it lies in, with
// invalid_point_num := down_del_num.  The reference hands the labels down
// lazily as later calls walk the subtree; the result each walk sees is the
// same as rewriting the whole subtree now, level by level.  Ranges stay as
// they are (nothing below the covered node is Updated).  Entry (node, bits):
// bit 0 the covered root itself, bit 1 the parent's new
// tree_downsample_deleted, bit 2 the parent's new tree_deleted.
__global__ void k_addbox_reset(Pool pool, const int2* __restrict__ front, int lvl, int2* __restrict__ front_next,
                               DevHdr* hdr, int fcap) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0) hdr->front3[(lvl + 2) % 3] = 0;
    const int count = min(hdr->front3[lvl % 3], fcap);
    for (int i = tid; i < count; i += gridDim.x * blockDim.x) {
        const int2 e = front[i];
        const int v = e.x;
        const uint32_t m = pool.meta[v];
        const int4 c = pool.cnt[v];
        bool tds, pds, td, pd;
        int dd = c.z, inv;
        if (e.y & 1) {
            // tree_deleted := tree_downsample_deleted, point_deleted :=
            // point_downsample_deleted, invalid := down_del (cpp:778-782)
            tds = (m & M_TDS) != 0;
            pds = (m & M_PDS) != 0;
            td = tds;
            pd = pds;
            inv = dd;
            const int p = pool.par[v];
            if (td) atomicOr(link_ptr(pool.hot, p), lk_td(v & 1));
            else atomicAnd(link_ptr(pool.hot, p), ~lk_td(v & 1));
        } else {
            // Push_Down from the reset parent (cpp:1115-1123)
            const bool ptds = (e.y & 2) != 0, ptd = (e.y & 4) != 0;
            tds = (m & M_TDS) || ptds;
            pds = (m & M_PDS) || ptds;
            td = ptd || tds;
            pd = td || pds;
            if (ptds) dd = c.x;
            inv = ptd ? c.x : dd;
        }
        pool.meta[v] = (m & ~(M_PDEL | M_TDEL | M_PDS | M_TDS)) | (pd ? M_PDEL : 0u) | (td ? M_TDEL : 0u) |
                       (pds ? M_PDS : 0u) | (tds ? M_TDS : 0u);
        pool.cnt[v] = make_int4(c.x, inv, dd, c.w);
        const uint32_t lk = load_link(pool.hot, v);
        const int pair = (int)lk_pair(lk);
        uint32_t nl = lk & ~(LK_PDEL | LK_TDL | LK_TDR);
        if (pd) nl |= LK_PDEL;
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            if (!(lk & lk_has(s))) continue;
            const int ch = 2 * pair + s;
            const bool ch_td = td || (pool.meta[ch] & M_TDS) || tds;
            if (ch_td) nl |= lk_td(s);
            const int pos = atomicAdd(&hdr->front3[(lvl + 1) % 3], 1);
            if (pos < fcap) front_next[pos] = make_int2(ch, (tds ? 2 : 0) | (td ? 4 : 0));
            else hdr->front_over = 1;
        }
        // only this lane writes v's own link word in this launch
        *link_ptr(pool.hot, v) = nl;
    }
}

void launch_addbox_walk(Pool pool, const float* box, int2* f0, int2* f1, int fcap, int L0, int L1,
                        unsigned int* recs, DevHdr* hdr, int cap, hipStream_t s) {
    if (L0 == 0) hipLaunchKernelGGL(k_addbox_init, dim3(1), dim3(64), 0, s, pool, box, f0, recs, hdr);
    for (int L = L0; L < L1; ++L) {
        const int2* fin = (L & 1) ? f1 : f0;
        int2* fout = (L & 1) ? f0 : f1;
        hipLaunchKernelGGL(k_addbox_level, dim3(512), dim3(256), 0, s, pool, box, fin, L, fout, recs, hdr, fcap, cap);
    }
}

void launch_addbox_reset(Pool pool, const unsigned int* recs, int nrec, int* seeds, int2* f0, int2* f1, int fcap,
                         int levels, DevHdr* hdr, hipStream_t s) {
    if (nrec <= 0) return;
    launch_fill_i32(hdr->front3, 0, 3, s);
    hipLaunchKernelGGL(k_addbox_split, dim3(blocks_for(nrec, 256)), dim3(256), 0, s, recs, hdr, nrec, seeds, f0);
    for (int L = 0; L < levels; ++L) {
        const int2* fin = (L & 1) ? f1 : f0;
        int2* fout = (L & 1) ? f0 : f1;
        hipLaunchKernelGGL(k_addbox_reset, dim3(512), dim3(256), 0, s, pool, fin, L, fout, hdr, fcap);
    }
}

// ========================================================== point deletion
// Delete_by_point (cpp:712-760): follow the insertion rule from the root,
// stop at a missing or tree_deleted node, take the first live node that is
// same_point() with the query.
__device__ int find_point(Pool pool, float x, float y, float z) {
    uint32_t lk0 = load_link(pool.hot, SENTINEL);
    if (!(lk0 & LK_HASL)) return -1;
    int u = ROOT;
    for (int step = 0; step <= WALK_LIMIT; ++step) {
        uint32_t m = pool.meta[u];
        if (m & M_TDEL) return -1;
        const float* f = pool.hot[u].f;
        if (!(m & M_PDEL) && same_point(f[0], f[1], f[2], x, y, z)) return u;
        uint32_t lk = fbits(f[3]);
        int axis = (int)(m & M_AXIS);
        float pc = axis == 0 ? x : (axis == 1 ? y : z);
        int side = pc < f[axis] ? 0 : 1;
        if (!(lk & lk_has(side))) return -1;
        u = 2 * (int)lk_pair(lk) + side;
    }
    return -1;
}

__global__ void k_delete_points_find(Pool pool, const float* __restrict__ xyz, int n, int* __restrict__ cand,
                                     DevHdr* hdr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int u = find_point(pool, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    cand[i] = u;
    if (u >= 0) {
        uint32_t old = atomicOr(&pool.meta[u], M_CLAIM);
        if (old & M_CLAIM) atomicAdd(&hdr->conflicts, 1);
    }
}

// Which nodes the reference's walks Update.  Point i's walk (cpp:712-760)
// Updates every node it leaves for a child slot; it does not Update the node
// it deletes (cpp:717-721) and stops at a node that is tree_deleted by then,
// which includes nodes the call's earlier points emptied.  With the
// conflict-free candidates known, node w is emptied by this call at
// max_del[w] (the last batch index deleting inside its subtree) iff its
// pre-call invalid count plus the call's deletions below it reaches TreeSize.
__global__ void k_delete_points_times(Pool pool, const int* __restrict__ cand, int n, int* __restrict__ del_below,
                                      int* __restrict__ max_del, DevHdr* hdr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int u = cand[i];
    for (int step = 0; u >= ROOT; ++step) {
        if (step > WALK_LIMIT) {
            atomicOr(&hdr->overflow, 16);
            return;
        }
        atomicAdd(&del_below[u], 1);
        atomicMax(&max_del[u], i);
        u = pool.par[u];
    }
}

// last_pass[w]: the highest batch index whose walk Updates w (-1: none);
// path_seed[i]: the deepest node point i's walk Updates (-1: none).
__device__ void walk_updates(Pool pool, float x, float y, float z, int i, int found, const int* del_below,
                             const int* max_del, int* last_pass, int* path_seed) {
    int seed = -1;
    uint32_t lk0 = load_link(pool.hot, SENTINEL);
    int u = (lk0 & LK_HASL) ? ROOT : -1;
    for (int step = 0; u >= 0 && step <= WALK_LIMIT; ++step) {
        if (u == found) break;
        const uint32_t m = pool.meta[u];
        if (m & M_TDEL) break;
        if (del_below) {
            const int4 c = pool.cnt[u];
            if (c.y + del_below[u] == c.x && max_del[u] < i) break;  // emptied earlier in the call
        }
        atomicMax(&last_pass[u], i);
        seed = u;
        const float* f = pool.hot[u].f;
        const uint32_t lk = fbits(f[3]);
        const int axis = (int)(m & M_AXIS);
        const float pc = axis == 0 ? x : (axis == 1 ? y : z);
        const int side = pc < f[axis] ? 0 : 1;
        u = (lk & lk_has(side)) ? 2 * (int)lk_pair(lk) + side : -1;
    }
    path_seed[i] = seed;
}

__global__ void k_delete_points_walk(Pool pool, const float* __restrict__ xyz, const int* __restrict__ cand, int n,
                                     const int* __restrict__ del_below, const int* __restrict__ max_del,
                                     int* __restrict__ last_pass, int* __restrict__ path_seed) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    walk_updates(pool, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], i, cand[i], del_below, max_del, last_pass,
                 path_seed);
}

// The labels a deleted node u (batch index i) takes for the refit: frozen
// range unless a later point's walk Updates it; the stored range kept as is
// if no walk of the call Updates it at all.
__device__ __forceinline__ uint32_t freeze_bits(const int* last_pass, int u, int i) {
    const int lp = last_pass[u];
    if (lp > i) return 0u;
    return M_FROZEN | (lp < 0 ? M_KEEPBOX : 0u);
}

// mode 0: apply the (conflict-free) candidates; mode 1: only drop the claims.
// seeds[2i] = deleted node, seeds[2i+1] = deepest Updated node of the walk.
__global__ void k_delete_points_apply(Pool pool, const int* __restrict__ cand, const int* __restrict__ last_pass,
                                      const int* __restrict__ path_seed, int n, int* __restrict__ seeds, int mode) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int u = cand[i];
    seeds[2 * i] = -1;
    seeds[2 * i + 1] = mode == 0 ? path_seed[i] : -1;
    if (u < 0) return;
    atomicAnd(&pool.meta[u], ~M_CLAIM);
    if (mode == 0) {
        atomicOr(&pool.meta[u], M_PDEL | freeze_bits(last_pass, u, i));
        atomicOr(link_ptr(pool.hot, u), LK_PDEL);
        seeds[2 * i] = u;
    }
}

// Ordered fallback when two points of one batch reach the same node: one
// lane walks the batch in order, as the reference's loop does (cpp:514-531).
// A deletion raises invalid_point_num up the path and labels every node it
// empties tree_deleted, so later walks stop there; the refit recomputes
// both from the children afterwards.
__global__ void k_delete_points_serial(Pool pool, const float* __restrict__ xyz, int n, int* __restrict__ last_pass,
                                       int* __restrict__ path_seed, int* __restrict__ seeds, DevHdr* hdr) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int i = 0; i < n; ++i) {
        const float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
        const int u = find_point(pool, x, y, z);
        walk_updates(pool, x, y, z, i, u, nullptr, nullptr, last_pass, path_seed);
        seeds[2 * i] = u;
        seeds[2 * i + 1] = path_seed[i];
        if (u < 0) continue;
        pool.meta[u] |= M_PDEL;
        *link_ptr(pool.hot, u) |= LK_PDEL;
        int a = u;
        for (int step = 0; a >= ROOT; ++step) {
            if (step > WALK_LIMIT) {
                hdr->overflow |= 16;
                return;
            }
            int4 c = pool.cnt[a];
            c.y += 1;
            pool.cnt[a] = c;
            if (c.y == c.x) pool.meta[a] |= M_TDEL;
            a = pool.par[a];
        }
        __threadfence();
    }
    for (int i = 0; i < n; ++i) {
        const int u = seeds[2 * i];
        if (u >= 0) pool.meta[u] |= freeze_bits(last_pass, u, i);
    }
}

void launch_delete_points_find(Pool pool, const float* xyz, int n, int* cand, DevHdr* hdr, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_delete_points_find, dim3(blocks_for(n, 256)), dim3(256), 0, s, pool, xyz, n, cand, hdr);
}
void launch_delete_points_apply(Pool pool, const float* xyz, const int* cand, int* del_below, int* max_del,
                                int* last_pass, int* path_seed, int n, int* seeds, DevHdr* hdr, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_delete_points_times, dim3(blocks_for(n, 256)), dim3(256), 0, s, pool, cand, n, del_below,
                       max_del, hdr);
    hipLaunchKernelGGL(k_delete_points_walk, dim3(blocks_for(n, 256)), dim3(256), 0, s, pool, xyz, cand, n, del_below,
                       max_del, last_pass, path_seed);
    hipLaunchKernelGGL(k_delete_points_apply, dim3(blocks_for(n, 256)), dim3(256), 0, s, pool, cand, last_pass,
                       path_seed, n, seeds, 0);
}
void launch_delete_points_unclaim(Pool pool, const int* cand, int n, int* seeds, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_delete_points_apply, dim3(blocks_for(n, 256)), dim3(256), 0, s, pool, cand, nullptr, nullptr,
                       n, seeds, 1);
}
void launch_delete_points_serial(Pool pool, const float* xyz, int n, int* last_pass, int* path_seed, int* seeds,
                                 DevHdr* hdr, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_delete_points_serial, dim3(1), dim3(64), 0, s, pool, xyz, n, last_pass, path_seed, seeds,
                       hdr);
}

// ================================================================ misc
__global__ void k_fill_i32(int* p, int v, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_i32(int* p, int v, long long n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_i32, dim3(blocks_for(n, 256)), dim3(256), 0, s, p, v, n);
}

__global__ void k_pack_points(const float* __restrict__ xyz, const uint32_t* __restrict__ ids, uint32_t base, int n,
                              float4* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t id = ids ? ids[i] : base + (uint32_t)i;
    out[i] = make_float4(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], bitsf(id));
}
void launch_pack_points(const float* xyz, const uint32_t* ids, uint32_t base, int n, float4* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pack_points, dim3(blocks_for(n, 256)), dim3(256), 0, s, xyz, ids, base, n, out);
}

}  // namespace ikd

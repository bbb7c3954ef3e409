// ikd_voxel.hip — the voxel downsample of Add_Points (ikd_Tree.cpp:422-448)
// as one ordered fold per voxel.
//
// For every input point the reference takes the voxel box
// lo = floor(p/ds)*ds, hi = lo+ds (float, cpp:424-429) and its centre
// m = lo + (hi-lo)/2.0 (double, narrowed to float, cpp:430-432), collects the
// live points S of the box (Search_by_range, pre-order), keeps the point
// nearest to m (the new point unless some s is strictly nearer, the first
// such s on ties), and when |S| > 1 or that point is same_point() with the new
// point, deletes S (downsample labels) and inserts the kept point, counting one.
//
// When no point of the batch or of the map lies in the boxes of two different
// voxels of the batch, voxels are independent and each one is a sequential
// fold over its own points (in input order, kept by a stable sort): after its
// first insertion a voxel holds exactly one live point c, and a later point p
// replaces c iff dist(p,m) <= dist(c,m), or re-inserts c iff it is
// same_point() with it.  Only the final state is materialised (one box delete
// of the original S, one insert of c); the intermediate nodes the reference
// creates and deletes again only change size().
//
// For power-of-two ds every box is exact and the boxes tile space, so the
// independence holds by construction.  For other ds the boxes can overlap or
// miss their own points by an ulp; the fold then reports the first input index
// at which independence breaks and the host runs the batch in ordered chunks.
#include "ikd_internal.h"

namespace ikd {

static inline int blocks_for(long long n, int bs) { return (int)((n + bs - 1) / bs); }

constexpr int VOX_BIAS = 1 << 20;

__device__ __forceinline__ unsigned long long vox_key(int ix, int iy, int iz) {
    return ((unsigned long long)(uint32_t)(ix + VOX_BIAS) << 42) | ((unsigned long long)(uint32_t)(iy + VOX_BIAS) << 21) |
           (unsigned long long)(uint32_t)(iz + VOX_BIAS);
}

__global__ void k_voxel_keys(const float* __restrict__ xyz, const uint32_t* __restrict__ ids, uint32_t id_base, int n,
                             float ds, unsigned long long* __restrict__ keys, uint32_t* __restrict__ vals,
                             float4* __restrict__ pts, int* __restrict__ bad) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = xyz[3 * i], y = xyz[3 * i + 1], z = xyz[3 * i + 2];
    float fx = floorf(x / ds), fy = floorf(y / ds), fz = floorf(z / ds);
    const float lim = (float)VOX_BIAS;
    if (!(fabsf(fx) < lim && fabsf(fy) < lim && fabsf(fz) < lim)) {
        atomicOr(bad, 1);
        fx = fy = fz = 0.0f;
    }
    keys[i] = vox_key((int)fx, (int)fy, (int)fz);
    vals[i] = (uint32_t)i;
    uint32_t id = ids ? ids[i] : id_base + (uint32_t)i;
    pts[i] = make_float4(x, y, z, bitsf(id));
}

__global__ void k_voxel_flags(const unsigned long long* __restrict__ keys, int n, int* __restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void k_voxel_starts(const int* __restrict__ flag, const int* __restrict__ scan, int n,
                               int* __restrict__ vstart, DevHdr* hdr) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) vstart[scan[i]] = i;
    if (i == n - 1) {
        int nv = scan[i] + flag[i];
        vstart[nv] = n;
        hdr->group_count = nv;
    }
}

// first sorted position of a key, or -1
__device__ int find_key(const unsigned long long* keys, int n, unsigned long long k) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (keys[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && keys[lo] == k) ? lo : -1;
}

struct VoxGeo {
    float fidx[3];  // floor(p/ds) of the voxel
    float ds;
};

// Offsets o in {-1,0,1}^3 \ {0} whose voxel box also holds point (x,y,z);
// returns the smallest first-input-index among such voxels present in the
// batch (or INT_MAX).
__device__ int foreign_voxel_first_index(const VoxGeo& g, float x, float y, float z,
                                         const unsigned long long* keys, const uint32_t* order, int n) {
    float p[3] = {x, y, z};
    bool in[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            float lo = (g.fidx[a] + (float)(o - 1)) * g.ds;
            float hi = lo + g.ds;
            in[a][o] = lo <= p[a] && p[a] < hi;
        }
    }
    int best = 0x7fffffff;
    for (int ox = 0; ox < 3; ++ox) {
        if (!in[0][ox]) continue;
        for (int oy = 0; oy < 3; ++oy) {
            if (!in[1][oy]) continue;
            for (int oz = 0; oz < 3; ++oz) {
                if (!in[2][oz] || (ox == 1 && oy == 1 && oz == 1)) continue;
                unsigned long long k = vox_key((int)g.fidx[0] + ox - 1, (int)g.fidx[1] + oy - 1, (int)g.fidx[2] + oz - 1);
                int pos = find_key(keys, n, k);
                if (pos >= 0) best = min(best, (int)order[pos]);
            }
        }
    }
    return best;
}

__global__ void k_voxel_fold(Pool pool, const float4* __restrict__ pts, const uint32_t* __restrict__ order,
                             const unsigned long long* __restrict__ keys, const int* __restrict__ vstart, int n,
                             int bound, float ds, int check, float* __restrict__ del_boxes, int* __restrict__ del_flag,
                             float4* __restrict__ ins_pts, int* __restrict__ ins_flag, DevHdr* hdr,
                             long long max_steps, int* __restrict__ del_out, int del_cap, int stack_cap,
                             const int* __restrict__ bad) {
    // per-lane depth-first stacks in LDS, [stack_cap][64]
    extern __shared__ int lds_dyn[];
    const int lane = threadIdx.x;
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (*bad) return;  // a voxel index out of range: the call fails and must change nothing
    if (v >= bound || v >= hdr->group_count) return;
    int j0 = vstart[v], j1 = vstart[v + 1];
    float4 P0 = pts[order[j0]];
    VoxGeo g;
    g.ds = ds;
    g.fidx[0] = floorf(P0.x / ds);
    g.fidx[1] = floorf(P0.y / ds);
    g.fidx[2] = floorf(P0.z / ds);
    float box[6];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        box[a] = g.fidx[a] * ds;
        box[3 + a] = box[a] + ds;
    }
    float mx = (float)((double)box[0] + (double)(box[3] - box[0]) / 2.0);
    float my = (float)((double)box[1] + (double)(box[4] - box[1]) / 2.0);
    float mz = (float)((double)box[2] + (double)(box[5] - box[2]) / 2.0);
    int first_idx = (int)order[j0];
    int split = 0x7fffffff;  // 2*index (+1 when the split is not a singleton)

    // ---- S0: live map points in the box, pre-order (own, left, right) ----
    int n0 = 0, best_s = -1;
    float best_sd = INFINITY;
    uint32_t lk0 = fbits(pool.hot[SENTINEL].f[3]);
    if ((lk0 & LK_HASL) && !(lk0 & LK_TDL) && !box_disjoint(box, pool.hot[SENTINEL].f + 4)) {
        int sp = 0;
        lds_dyn[lane] = ROOT;
        sp = 1;
        long long steps = 0;
        while (sp > 0) {
            if (++steps > max_steps) {
                atomicOr(&hdr->overflow, 4);
                break;
            }
            int u = lds_dyn[--sp * 64 + lane];
            float f[16];
            load_record(pool.hot[u].f, f);
            uint32_t lk = fbits(f[3]);
            if (!(lk & LK_PDEL) && box_contains(box, f[0], f[1], f[2])) {
                ++n0;
                float d = calc_dist(f[0], f[1], f[2], mx, my, mz);
                if (d < best_sd) {
                    best_sd = d;
                    best_s = u;
                }
                if (check) {
                    int other = foreign_voxel_first_index(g, f[0], f[1], f[2], keys, order, n);
                    if (other != 0x7fffffff) split = min(split, 2 * max(other, first_idx) + 1);
                }
            }
            int pair = (int)lk_pair(lk);
            // push right first so the left subtree is visited first
#pragma unroll
            for (int s = 1; s >= 0; --s) {
                if (!(lk & lk_has(s)) || (lk & lk_td(s))) continue;
                if (box_disjoint(box, f + 4 + 6 * s)) continue;
                if (sp < stack_cap) lds_dyn[sp++ * 64 + lane] = 2 * pair + s;
                else atomicOr(&hdr->overflow, 4);
            }
        }
    }

    // ---- the ordered fold over this voxel's batch points ----
    bool have_new = false;
    float cx = 0, cy = 0, cz = 0;
    uint32_t cid = 0;
    float ex = 0, ey = 0, ez = 0;  // the single original point when |S0| == 1
    if (best_s >= 0) {
        const float* f = pool.hot[best_s].f;
        ex = f[0]; ey = f[1]; ez = f[2];
    }
    uint32_t eid = best_s >= 0 ? pool.ids[best_s] : 0u;
    int added = 0;
    for (int j = j0; j < j1; ++j) {
        int idx = (int)order[j];
        float4 P = pts[idx];
        if (check) {
            if (!box_contains(box, P.x, P.y, P.z)) split = min(split, 2 * idx);
            int other = foreign_voxel_first_index(g, P.x, P.y, P.z, keys, order, n);
            if (other != 0x7fffffff) split = min(split, 2 * idx);
        }
        float dp = calc_dist(P.x, P.y, P.z, mx, my, mz);
        float bx = P.x, by = P.y, bz = P.z;
        uint32_t bid = fbits(P.w);
        int s_count;
        if (j == j0) {
            s_count = n0;
            if (n0 > 0 && best_sd < dp) {
                bx = ex; by = ey; bz = ez; bid = eid;
            }
        } else {
            s_count = 1;
            float qx = have_new ? cx : ex, qy = have_new ? cy : ey, qz = have_new ? cz : ez;
            uint32_t qid = have_new ? cid : eid;
            float de = calc_dist(qx, qy, qz, mx, my, mz);
            if (de < dp) {
                bx = qx; by = qy; bz = qz; bid = qid;
            }
        }
        if (s_count > 1 || same_point(P.x, P.y, P.z, bx, by, bz)) {
            have_new = true;
            cx = bx; cy = by; cz = bz; cid = bid;
            ++added;
        }
    }
    if (split != 0x7fffffff) atomicMin(&hdr->split_index, split);
    atomicAdd(&hdr->added, added);
    if (have_new) atomicAdd(&hdr->ins_count, 1);
    del_flag[v] = (have_new && n0 > 0) ? 1 : 0;
    ins_flag[v] = have_new ? 1 : 0;
#pragma unroll
    for (int a = 0; a < 6; ++a) del_boxes[6 * v + a] = box[a];
    ins_pts[v] = make_float4(cx, cy, cz, bitsf(cid));

    // The voxel's original points S0 are deleted (downsample labels) when a
    // new point replaces them: Delete_by_range(voxel, downsample) of cpp:446.
    // Voxel boxes are disjoint here, so each S0 node is written by one lane.
    if (del_out && have_new && n0 > 0 && split == 0x7fffffff) {
        int sp = 0;
        lds_dyn[lane] = ROOT;
        sp = 1;
        long long steps = 0;
        while (sp > 0) {
            if (++steps > max_steps) {
                atomicOr(&hdr->overflow, 4);
                break;
            }
            int u = lds_dyn[--sp * 64 + lane];
            float f[16];
            load_record(pool.hot[u].f, f);
            uint32_t lk = fbits(f[3]);
            if (!(lk & LK_PDEL) && box_contains(box, f[0], f[1], f[2])) {
                atomicOr(&pool.meta[u], M_PDEL | M_PDS);
                atomicOr(reinterpret_cast<uint32_t*>(&pool.hot[u].f[3]), LK_PDEL);
                int pos = atomicAdd(&hdr->rec_count, 1);
                if (pos < del_cap) del_out[pos] = u;
                else atomicOr(&hdr->overflow, 4);
            }
            int pair = (int)lk_pair(lk);
#pragma unroll
            for (int s2 = 1; s2 >= 0; --s2) {
                if (!(lk & lk_has(s2)) || (lk & lk_td(s2))) continue;
                if (box_disjoint(box, f + 4 + 6 * s2)) continue;
                if (sp < stack_cap) lds_dyn[sp++ * 64 + lane] = 2 * pair + s2;
                else atomicOr(&hdr->overflow, 4);
            }
        }
    }
}

// ---------------------------------------------------------------- team fold
// The same fold with T lanes per voxel.  Maps built from raw scans hold
// hundreds to thousands of points in one voxel near the trajectory, and one
// lane walking all of them sets the launch time; a team splits the walk.
// Every lane runs its own depth-first walk over a deque in LDS; a lane that
// runs dry takes the shallowest entry of a busy teammate.  The walk order is
// no longer pre-order, so the tie on distance to the centre (the reference
// keeps the first point of Search_by_range's pre-order, own-left-right) is
// broken by each node's pre-order rank, read off its path from the root.

// Pre-order position of node u: the sides (left 0, right 1) of its path from
// the root as bits from the top, and its depth.  Node a precedes node b in
// pre-order iff (path_a, depth_a) < (path_b, depth_b) lexicographically: an
// ancestor's path is a prefix (aligned, no larger; ties go to the shallower),
// and below the first differing side the left subtree comes first.
__device__ void preorder_pos(const Pool& pool, int u, unsigned long long& path, int& depth) {
    unsigned long long bits = 0;
    int d = 0;
    for (int v = u; v != ROOT && v > SENTINEL + 1 && d < 64; v = pool.par[v]) {
        bits = (bits >> 1) | ((unsigned long long)(v & 1) << 63);
        ++d;
    }
    path = bits;
    depth = d;
}
__device__ bool preorder_less(const Pool& pool, int a, int b) {
    unsigned long long pa, pb;
    int da, db;
    preorder_pos(pool, a, pa, da);
    preorder_pos(pool, b, pb, db);
    return pa < pb || (pa == pb && da < db);
}

__device__ __forceinline__ void team_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Visits every node of the tree whose point is live and inside box, once per
// team, calling visit(node, record) on the lane that reaches it.  Every lane
// of the wavefront calls this (active marks the teams with work); stk is the
// lane's column of a [cap][64] LDS array, cap a power of two.
template <int T, class Visit>
__device__ void team_box_walk(const Pool& pool, const float* box, bool active, int* lds, int cap, int (*dn)[T],
                              long long max_steps, DevHdr* hdr, Visit visit) {
    const int lane = threadIdx.x;
    const int team = lane / T;
    const unsigned long long team_mask = (T == 64) ? ~0ull : (((1ull << T) - 1ull) << (team * T));
    const unsigned long long below = (1ull << lane) - 1ull;
    const int mask = cap - 1;
    int head = 0, tail = 0;  // the lane's deque [head, tail)
    int n = -1;
    if (active && lane % T == 0) {
        const uint32_t lk0 = fbits(pool.hot[SENTINEL].f[3]);
        if ((lk0 & LK_HASL) && !(lk0 & LK_TDL) && !box_disjoint(box, pool.hot[SENTINEL].f + 4)) n = ROOT;
    }
    long long iter = 0;
    while (true) {
        if (n < 0 && tail != head) {
            --tail;
            n = lds[(tail & mask) * 64 + lane];
        }
        const unsigned long long imask = __ballot(active && n < 0) & team_mask;
        const unsigned long long dmask = __ballot(tail != head) & team_mask;
        // the fences run on every lane whenever any team hands work over, so
        // the wave barriers are never inside team-divergent control flow
        const bool donate = imask != 0 && dmask != 0;
        if (__ballot(donate) != 0) {
            const int ni = __popcll(imask), nd = __popcll(dmask);
            if (donate && tail != head) {
                const int r = __popcll(dmask & below);
                if (r < ni) {
                    dn[team][r] = lds[(head & mask) * 64 + lane];
                    ++head;
                }
            }
            team_lds_fence();
            if (donate && active && n < 0) {
                const int r = __popcll(imask & below);
                if (r < nd) n = dn[team][r];
            }
            team_lds_fence();
        }
        if (__ballot(n >= 0 || tail != head) == 0) break;
        if (++iter > max_steps) {
            if (lane == 0) atomicOr(&hdr->overflow, 4);
            break;
        }
        if (n >= 0) {
            float f[16];
            load_record(pool.hot[n].f, f);
            const uint32_t lk = fbits(f[3]);
            if (!(lk & LK_PDEL) && box_contains(box, f[0], f[1], f[2])) visit(n, f);
            const int pair = (int)lk_pair(lk);
            int next = -1;
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (!(lk & lk_has(s)) || (lk & lk_td(s))) continue;
                if (box_disjoint(box, f + 4 + 6 * s)) continue;
                if (next < 0) {
                    next = 2 * pair + s;
                } else if (tail - head < cap) {
                    lds[(tail & mask) * 64 + lane] = 2 * pair + s;
                    ++tail;
                } else {
                    atomicOr(&hdr->overflow, 4);
                }
            }
            n = next;
        }
    }
}

template <int T>
__global__ __launch_bounds__(64) void k_voxel_fold_team(Pool pool, const float4* __restrict__ pts,
                                                        const uint32_t* __restrict__ order,
                                                        const unsigned long long* __restrict__ keys,
                                                        const int* __restrict__ vstart, int n, int bound, float ds,
                                                        int check, float* __restrict__ del_boxes,
                                                        int* __restrict__ del_flag, float4* __restrict__ ins_pts,
                                                        int* __restrict__ ins_flag, DevHdr* hdr, long long max_steps,
                                                        int* __restrict__ del_out, int del_cap, int stack_cap,
                                                        const int* __restrict__ bad) {
    constexpr int TEAMS = 64 / T;
    if (*bad) return;  // a voxel index out of range: the call fails and must change nothing (uniform exit)
    extern __shared__ int lds_dyn[];  // per-lane deques, [stack_cap][64]
    __shared__ int dn[TEAMS][T];
    const int lane = threadIdx.x;
    const int team = lane / T, tl = lane % T;
    const int v = blockIdx.x * TEAMS + team;
    const bool valid = v < bound && v < hdr->group_count;
    const int vv = valid ? v : 0;
    const int j0 = valid ? vstart[vv] : 0, j1 = valid ? vstart[vv + 1] : 0;
    const float4 P0 = pts[order[valid ? j0 : 0]];
    VoxGeo g;
    g.ds = ds;
    g.fidx[0] = floorf(P0.x / ds);
    g.fidx[1] = floorf(P0.y / ds);
    g.fidx[2] = floorf(P0.z / ds);
    float box[6];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        box[a] = g.fidx[a] * ds;
        box[3 + a] = box[a] + ds;
    }
    const float mx = (float)((double)box[0] + (double)(box[3] - box[0]) / 2.0);
    const float my = (float)((double)box[1] + (double)(box[4] - box[1]) / 2.0);
    const float mz = (float)((double)box[2] + (double)(box[5] - box[2]) / 2.0);
    const int first_idx = (int)order[valid ? j0 : 0];

    // ---- S0: live map points in the box; the nearest to the centre, first in pre-order on ties
    int n0 = 0, best_s = -1, split = 0x7fffffff;
    float best_sd = INFINITY;
    team_box_walk<T>(pool, box, valid, lds_dyn, stack_cap, dn, max_steps, hdr, [&](int u, const float* f) {
        ++n0;
        const float d = calc_dist(f[0], f[1], f[2], mx, my, mz);
        if (best_s < 0 || d < best_sd || (d == best_sd && preorder_less(pool, u, best_s))) {
            best_sd = d;
            best_s = u;
        }
        if (check) {
            const int other = foreign_voxel_first_index(g, f[0], f[1], f[2], keys, order, n);
            if (other != 0x7fffffff) split = min(split, 2 * max(other, first_idx) + 1);
        }
    });
#pragma unroll
    for (int off = T / 2; off >= 1; off >>= 1) {
        n0 += __shfl_xor(n0, off, T);
        split = min(split, __shfl_xor(split, off, T));
        const float od = __shfl_xor(best_sd, off, T);
        const int os = __shfl_xor(best_s, off, T);
        if (os >= 0 && (best_s < 0 || od < best_sd || (od == best_sd && preorder_less(pool, os, best_s)))) {
            best_sd = od;
            best_s = os;
        }
    }

    // ---- the ordered fold over this voxel's batch points (every lane of the team alike)
    bool have_new = false;
    float cx = 0, cy = 0, cz = 0;
    uint32_t cid = 0;
    float ex = 0, ey = 0, ez = 0;  // the single original point when |S0| == 1
    if (best_s >= 0) {
        const float* f = pool.hot[best_s].f;
        ex = f[0];
        ey = f[1];
        ez = f[2];
    }
    const uint32_t eid = best_s >= 0 ? pool.ids[best_s] : 0u;
    int added = 0;
    for (int j = j0; j < j1; ++j) {
        const int idx = (int)order[j];
        const float4 P = pts[idx];
        if (check) {
            if (!box_contains(box, P.x, P.y, P.z)) split = min(split, 2 * idx);
            const int other = foreign_voxel_first_index(g, P.x, P.y, P.z, keys, order, n);
            if (other != 0x7fffffff) split = min(split, 2 * idx);
        }
        const float dp = calc_dist(P.x, P.y, P.z, mx, my, mz);
        float bx = P.x, by = P.y, bz = P.z;
        uint32_t bid = fbits(P.w);
        int s_count;
        if (j == j0) {
            s_count = n0;
            if (n0 > 0 && best_sd < dp) {
                bx = ex;
                by = ey;
                bz = ez;
                bid = eid;
            }
        } else {
            s_count = 1;
            const float qx = have_new ? cx : ex, qy = have_new ? cy : ey, qz = have_new ? cz : ez;
            const uint32_t qid = have_new ? cid : eid;
            const float de = calc_dist(qx, qy, qz, mx, my, mz);
            if (de < dp) {
                bx = qx;
                by = qy;
                bz = qz;
                bid = qid;
            }
        }
        if (s_count > 1 || same_point(P.x, P.y, P.z, bx, by, bz)) {
            have_new = true;
            cx = bx;
            cy = by;
            cz = bz;
            cid = bid;
            ++added;
        }
    }
    if (valid && tl == 0) {
        if (split != 0x7fffffff) atomicMin(&hdr->split_index, split);
        atomicAdd(&hdr->added, added);
        if (have_new) atomicAdd(&hdr->ins_count, 1);
        del_flag[v] = (have_new && n0 > 0) ? 1 : 0;
        ins_flag[v] = have_new ? 1 : 0;
#pragma unroll
        for (int a = 0; a < 6; ++a) del_boxes[6 * v + a] = box[a];
        ins_pts[v] = make_float4(cx, cy, cz, bitsf(cid));
    }

    // The voxel's original points S0 are deleted (downsample labels) when a
    // new point replaces them: Delete_by_range(voxel, downsample) of cpp:446.
    // Voxel boxes are disjoint here, so each S0 node is written by one lane.
    const bool del = valid && del_out && have_new && n0 > 0 && split == 0x7fffffff;
    team_box_walk<T>(pool, box, del, lds_dyn, stack_cap, dn, max_steps, hdr, [&](int u, const float*) {
        atomicOr(&pool.meta[u], M_PDEL | M_PDS);
        atomicOr(reinterpret_cast<uint32_t*>(&pool.hot[u].f[3]), LK_PDEL);
        const int pos = atomicAdd(&hdr->rec_count, 1);
        if (pos < del_cap) del_out[pos] = u;
        else atomicOr(&hdr->overflow, 4);
    });
}

// stream compaction of the flagged voxel boxes / points
__global__ void k_compact_boxes(const int* __restrict__ flag, const int* __restrict__ scan, const float* __restrict__ in,
                                int bound, DevHdr* hdr, float* __restrict__ out, int* __restrict__ out_count) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    int nv = hdr->group_count;
    if (v >= bound || v >= nv) return;
    if (v == nv - 1) *out_count = scan[v] + flag[v];
    if (!flag[v]) return;
    int o = scan[v];
#pragma unroll
    for (int a = 0; a < 6; ++a) out[6 * o + a] = in[6 * v + a];
}

__global__ void k_compact_pts(const int* __restrict__ flag, const int* __restrict__ scan, const float4* __restrict__ in,
                              int bound, DevHdr* hdr, float4* __restrict__ out, int* __restrict__ out_count) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    int nv = hdr->group_count;
    if (v >= bound || v >= nv) return;
    if (v == nv - 1) *out_count = scan[v] + flag[v];
    if (!flag[v]) return;
    out[scan[v]] = in[v];
}

void launch_voxel_keys2(const float* xyz, const uint32_t* ids, uint32_t id_base, int n, float ds,
                        unsigned long long* keys, uint32_t* vals, float4* pts, int* bad, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_voxel_keys, dim3(blocks_for(n, 256)), dim3(256), 0, s, xyz, ids, id_base, n, ds, keys, vals,
                       pts, bad);
}
void launch_voxel_flags(const unsigned long long* keys, int n, int* flag, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_voxel_flags, dim3(blocks_for(n, 256)), dim3(256), 0, s, keys, n, flag);
}
void launch_voxel_starts(const int* flag, const int* scan, int n, int* vstart, DevHdr* hdr, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_voxel_starts, dim3(blocks_for(n, 256)), dim3(256), 0, s, flag, scan, n, vstart, hdr);
}
void launch_voxel_fold2(Pool pool, const float4* pts, const uint32_t* order, const unsigned long long* keys,
                        const int* vstart, int n, int bound, float ds, int check, float* del_boxes, int* del_flag,
                        float4* ins_pts, int* ins_flag, DevHdr* hdr, long long max_steps, int* del_out, int del_cap,
                        int stack_cap, int lanes, const int* bad, hipStream_t s) {
    if (bound <= 0) return;
    const size_t lds = (size_t)stack_cap * 64 * sizeof(int);
#define IKD_FOLD_TEAM(T)                                                                                            \
    hipLaunchKernelGGL(k_voxel_fold_team<T>, dim3(blocks_for(bound, 64 / (T))), dim3(64), lds, s, pool, pts, order, \
                       keys, vstart, n, bound, ds, check, del_boxes, del_flag, ins_pts, ins_flag, hdr, max_steps,     \
                       del_out, del_cap, stack_cap, bad)
    switch (lanes) {
        case 4: IKD_FOLD_TEAM(4); break;
        case 8: IKD_FOLD_TEAM(8); break;
        case 16: IKD_FOLD_TEAM(16); break;
        default:
            hipLaunchKernelGGL(k_voxel_fold, dim3(blocks_for(bound, 64)), dim3(64), lds, s, pool, pts, order, keys,
                               vstart, n, bound, ds, check, del_boxes, del_flag, ins_pts, ins_flag, hdr, max_steps,
                               del_out, del_cap, stack_cap, bad);
    }
#undef IKD_FOLD_TEAM
}
void launch_compact_boxes(const int* flag, const int* scan, const float* in, int bound, DevHdr* hdr, float* out,
                          int* out_count, hipStream_t s) {
    if (bound <= 0) return;
    hipLaunchKernelGGL(k_compact_boxes, dim3(blocks_for(bound, 256)), dim3(256), 0, s, flag, scan, in, bound, hdr, out,
                       out_count);
}
void launch_compact_pts(const int* flag, const int* scan, const float4* in, int bound, DevHdr* hdr, float4* out,
                        int* out_count, hipStream_t s) {
    if (bound <= 0) return;
    hipLaunchKernelGGL(k_compact_pts, dim3(blocks_for(bound, 256)), dim3(256), 0, s, flag, scan, in, bound, hdr, out,
                       out_count);
}

}  // namespace ikd

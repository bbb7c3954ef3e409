// ikd_internal.h — declarations shared between the kernel translation units
// and the host orchestration (not part of the public C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ikd_common.h"
#include "ikd_gload.h"

namespace ikd {

// Small device-resident header: counters the kernels advance and the host
// reads back when it needs a size.
struct DevHdr {
    int pair_used;      // next free sibling pair
    int max_depth;      // deepest node written so far
    int seg_count[2];   // build: active segments at the current / next level
    int dirty_count;    // refit: nodes on modified paths
    int viol_count;     // Criterion_Check violators found by the refit
    int top_count;      // topmost violators (subtrees to rebuild)
    int rec_count;      // delete records (box traversal) / enumerate output
    int front_count[2]; // enumerate: frontier sizes
    int conflicts;      // Delete_Points lanes that hit the same node
    int group_count;    // insert: distinct attach slots
    int new_pairs;      // insert: pairs allocated for leaf attach points
    int split_index;    // downsample: first batch index that needs ordered handling
    int overflow;       // a traversal stack overflowed (must stay 0)
    int pair_pending;   // build: pairs taken by the level just written
    int front3[3];      // enumerate: rotating frontier counters
    int added;          // downsample: Add_Points return value of the batch
    int defer_count;    // deferred rebuilds listed (kept across calls)
    int front_over;     // a box-deletion frontier outgrew its buffer
    int ins_count;      // downsample: voxels whose fold keeps a point to insert
    int removed_count;  // points rebuilds removed since the last acquire (kept across calls)
    float alpha_bal;    // the root's alpha_bal / alpha_del as of its last Update (cpp:1315-1321)
    float alpha_del;
    int small_lists[2]; // small build: segments listed for the 64- and 256-thread workgroup builds
};
static_assert(sizeof(DevHdr) % sizeof(int) == 0, "the header is cleared as ints");
constexpr int DEFER_CAP = 4096;  // deferred-rebuild list entries
constexpr int BOXDEL_LEVELS = 4;  // tree levels one box-delete launch takes each frontier pair down

struct Pool {
    NodeHot* hot;
    uint32_t* ids;
    int32_t* par;
    int4* cnt;  // TreeSize, invalid_point_num, down_del_num, depth
    uint32_t* meta;
};

struct Params {
    float delete_param;
    float balance_param;
    float box_length;
    int defer_points;  // violators with more live points are listed for a later rebuild (0: never)
};

// ---------------------------------------------------------------- build
struct BuildLevel {
    const float4* pts;     // x, y, z, id bits
    const uint32_t* perm[3];
    uint32_t* perm_next[3];
    const int* sid;        // active segment per position, -1 when placed
    int* sid_next;
    uint8_t* side;         // per point: 0 left, 1 right, 2 pivot
    const int4* seg;       // l, r, node, parent
    const int* seg_depth;
    int4* seg_next;
    int* seg_depth_next;
    int* seg_axis;
    int* seg_mid;
    int2* seg_child;       // next-level indices of the two children, -1 if none
    unsigned long long* seg_flags;  // scan input: pairs | children<<32
    unsigned long long* seg_scan;
    unsigned long long* pos_flags;  // 3m: isL | isR<<32
    unsigned long long* pos_scan;
    int m;
    int cur;               // 0/1: which DevHdr::seg_count holds this level's count
};

void launch_build_keys(const float4* pts, const int* seg_of, int m, int axis, unsigned long long* keys,
                       uint32_t* vals, hipStream_t s);
void launch_build_init_sid(const unsigned long long* sorted_keys0, int m, int* sid, hipStream_t s);
void launch_build_seg_a(const BuildLevel& L, int bound, DevHdr* hdr, hipStream_t s);
void launch_build_seg_b(const BuildLevel& L, int bound, Pool pool, DevHdr* hdr, hipStream_t s);
void launch_build_side(const BuildLevel& L, DevHdr* hdr, hipStream_t s);
void launch_build_flags(const BuildLevel& L, hipStream_t s);
void launch_build_scatter(const BuildLevel& L, hipStream_t s);

void launch_small_pairs(const int4* seg, int S, const int* count_dev, int* pairs, DevHdr* hdr, hipStream_t s);
// Presorted lists for the small build (all null: the kernel sorts): the
// segments' y- and z-ordered lists of point indices (list1, list2) and, for
// every point, its position in the x-ordered list the points were gathered in.
struct SmallLists {
    const uint32_t* list1 = nullptr;
    const uint32_t* list2 = nullptr;
    const int* inv0 = nullptr;
};
// lists: 2 S ints of scratch (the size-class lists)
void launch_build_small(const float4* pts, const int4* seg, const int* seg_depth, const int* pair_scan, int S,
                        const int* count_dev, const uint32_t* orig, SmallLists pre, int* lists, Pool pool,
                        DevHdr* hdr, hipStream_t s);
void launch_gather_list(const float4* pts, const uint32_t* perm, const int* sid, int m, float4* out, int* inv,
                        hipStream_t s);
void launch_small_commit(const int* pairs, const int* pair_scan, int S, DevHdr* hdr, hipStream_t s);

// ---------------------------------------------------------------- refit
void launch_refit_up(Pool pool, Params prm, const int* dirty, int* pend, int* start, DevHdr* hdr, int* viol,
                     int* defer, hipStream_t s);
void launch_mark_dirty(Pool pool, const int* seeds, const int* nseeds_dev, int bound, int stop_at_tdel,
                       int* dirty, int* pend, DevHdr* hdr, hipStream_t s);
void launch_defer_check(Pool pool, Params prm, const int* defer, int n, int4* info, int* before, int* keep,
                        hipStream_t s);
void launch_defer_clear(Pool pool, const int* defer, int n, hipStream_t s);
void launch_find_top(Pool pool, const int* viol, int bound, DevHdr* hdr, int* top, hipStream_t s);
void launch_clear_viol(Pool pool, const int* viol, int bound, DevHdr* hdr, hipStream_t s);
// root statistics of a root the build just created (0 when it has <= 3 nodes)
void launch_root_alpha_fresh(Pool pool, DevHdr* hdr, hipStream_t s);

// ---------------------------------------------------------------- enumerate
// rec (nullable): points of deleted nodes that the downsample did not delete
// are appended at hdr->removed_count (rec_cap entries).
void launch_enum_preorder(Pool pool, const int4* front, int lvl, int4* front_next, float4* out_pts, int* out_seg,
                          DevHdr* hdr, int bound, int out_cap, float4* rec, int rec_cap, hipStream_t s);
void launch_enum_preorder_block(Pool pool, int4* f0, int4* f1, int n_roots, float4* out_pts, int* out_seg,
                                DevHdr* hdr, int bound, int out_cap, int max_levels, float4* rec, int rec_cap,
                                hipStream_t s);
// groups: (first root, roots, frontier offset, frontier entries) per workgroup
void launch_enum_groups(Pool pool, const int4* roots, const int4* groups, int n_groups, int4* f0, int4* f1,
                        float4* out_pts, int* out_seg, DevHdr* hdr, int out_cap, int max_levels, float4* rec,
                        int rec_cap, hipStream_t s);

// ---------------------------------------------------------------- insert
void launch_insert_descend(Pool pool, const float4* pts, int n, unsigned long long* keys, uint32_t* vals,
                           DevHdr* hdr, hipStream_t s);
void launch_insert_groups(const unsigned long long* sorted_keys, int n, int* grp_flag, hipStream_t s);
void launch_insert_segments2(Pool pool, const unsigned long long* keys, const int* gscan, const int* gflag, int n,
                             int4* seg, int* seg_depth, int* seg_of, int* attach, int* pair_flag, DevHdr* hdr,
                             hipStream_t s);
void launch_insert_pairs(Pool pool, const int* attach, const int* pair_flag, const int* pair_scan, DevHdr* hdr,
                         int bound, int4* seg, hipStream_t s);
void launch_insert_link(Pool pool, const int* attach, const int4* seg, DevHdr* hdr, int bound, hipStream_t s);

// ---------------------------------------------------------------- deletes
// levels [L0, L1) of the breadth-first box deletion (L0 == 0 seeds the frontier)
void launch_box_delete_bfs(Pool pool, const float* boxes, int nb, int2* f0, int2* f1, int fcap, int L0, int L1,
                           unsigned int* recs, DevHdr* hdr, int cap, hipStream_t s);
// Add_by_range for one box: levels [L0, L1) of the walk (L0 == 0 starts it);
// then the reset of the covered subtrees (`levels` levels) and the refit
// seeds (-1 or a cut node) in seeds[0, nrec).
void launch_addbox_walk(Pool pool, const float* box, int2* f0, int2* f1, int fcap, int L0, int L1,
                        unsigned int* recs, DevHdr* hdr, int cap, hipStream_t s);
void launch_addbox_reset(Pool pool, const unsigned int* recs, int nrec, int* seeds, int2* f0, int2* f1, int fcap,
                         int levels, DevHdr* hdr, hipStream_t s);
void launch_box_delete_apply(Pool pool, const unsigned int* recs, DevHdr* hdr, int bound, int ds, int* seeds,
                             hipStream_t s);
// Delete_Points: find claims each point's node (conflicts counted); apply
// (conflict-free batches) works out which nodes the reference's walks Update
// (per-node del_below = 0, max_del = -1, last_pass = -1 before the call;
// per-point path_seed) and labels the deleted nodes; seeds get 2n entries
// (deleted node, deepest Updated node) per point.  serial: the ordered
// fallback, same outputs.
void launch_delete_points_find(Pool pool, const float* xyz, int n, int* cand, DevHdr* hdr, hipStream_t s);
void launch_delete_points_apply(Pool pool, const float* xyz, const int* cand, int* del_below, int* max_del,
                                int* last_pass, int* path_seed, int n, int* seeds, DevHdr* hdr, hipStream_t s);
void launch_delete_points_unclaim(Pool pool, const int* cand, int n, int* seeds, hipStream_t s);
void launch_delete_points_serial(Pool pool, const float* xyz, int n, int* last_pass, int* path_seed, int* seeds,
                                 DevHdr* hdr, hipStream_t s);

// ---------------------------------------------------------------- downsample
void launch_voxel_keys2(const float* xyz, const uint32_t* ids, uint32_t id_base, int n, float ds,
                        unsigned long long* keys, uint32_t* vals, float4* pts, int* bad, hipStream_t s);
void launch_voxel_flags(const unsigned long long* keys, int n, int* flag, hipStream_t s);
void launch_voxel_starts(const int* flag, const int* scan, int n, int* vstart, DevHdr* hdr, hipStream_t s);
void launch_voxel_fold2(Pool pool, const float4* pts, const uint32_t* order, const unsigned long long* keys,
                        const int* vstart, int n, int bound, float ds, int check, float* del_boxes, int* del_flag,
                        float4* ins_pts, int* ins_flag, DevHdr* hdr, long long max_steps, int* del_out, int del_cap,
                        int stack_cap, int lanes, const int* bad, hipStream_t s);
void launch_compact_boxes(const int* flag, const int* scan, const float* in, int bound, DevHdr* hdr, float* out,
                          int* out_count, hipStream_t s);
void launch_compact_pts(const int* flag, const int* scan, const float4* in, int bound, DevHdr* hdr, float4* out,
                        int* out_count, hipStream_t s);

// ---------------------------------------------------------------- queries
struct TreeRef {
    const NodeHot* hot;
    const uint32_t* ids;
};

struct KnnArgs {
    const NodeHot* hot;     // single-tree launch
    const uint32_t* ids;
    const TreeRef* trees;   // multi-tree launch: tree table, with qtree[q] per query (else null)
    const int* qtree;
    int ntrees;             // entries of trees[]; a query naming no valid tree finds nothing
    const float* q;         // packed xyz
    const uint32_t* order;  // optional query permutation (Morton order), may be null
    int nq;
    int k;                  // uniform k, or the bound on per-query qk[q] (clamped to [0, k])
    float thr;              // largest float <= max_dist^2 (double); or per-query qthr[q]
    const int* qk;
    const float* qthr;
    int out_stride;         // result slots per query
    float* out_xyz;
    float* out_d2;
    uint32_t* out_ids;
    int* out_counts;
    unsigned int* visits;   // optional instrumentation: visited nodes per query
    unsigned int* boxes;    // optional: child boxes evaluated per query
    int* overflow;
    int max_steps;          // visit bound per query (the pool size): a corrupt tree cannot hang a lane
    int stack;              // LDS stack entries per lane (set by launch_knn)
    // optional, one-lane single-tree launches: the tree's top levels staged in
    // each workgroup's LDS (top_slots = 2^levels - 1 records, launch_top_table)
    const NodeHot* top;
    const int* top_idx;
    int top_slots;
    // one-lane launches: 4-byte LDS stack entries (every tree of the launch
    // has at most 2^24 node slots)
    int compact_stack;
};
// Copies the tree's first `levels` levels into a complete binary layout (slot 0
// the root, slot s's children at 2s+1 and 2s+2): top[s] the node's record,
// top_idx[s] its index or -1 where the tree has no node.
void launch_top_table(const NodeHot* hot, int levels, NodeHot* top, int* top_idx, hipStream_t s);
// lanes_per_query 0 (persistent lanes), 1 (one lane per query) or 2, 4, 8, 16 (a team of lanes
// shares one query's traversal); stack = LDS entries per lane (1-lane) or
// per-lane share of the team stack.
int launch_knn(const KnnArgs& a, int kmax, int lanes_per_query, int stack, hipStream_t s, int* persist_next,
               int persist_waves);

// per-query thresholds: thr[i] = largest float <= max_dist[i]^2 (double); NaN gives -inf
void launch_thresholds(const double* max_dist, int n, float* thr, hipStream_t s);

void launch_morton(const float* q, int nq, const NodeHot* hot, unsigned long long* keys, uint32_t* vals,
                   hipStream_t s);

struct RangeArgs {
    Pool pool;
    const float* boxes;   // box search: 6 floats per query
    const float* centers; // radius search: xyz per query
    const float* radii;
    int nq;
    int mode;             // 0 box, 1 radius
    float* out_xyz;
    uint32_t* out_ids;
    int* overflow;
    long long max_steps;
    int stack;            // LDS stack entries per lane (>= tree depth + 2)
    const Pool* pools;    // multi-tree launch: pool table with qtree[q] per query (else null)
    const int* qtree;
    int ntrees;           // a query naming no valid tree finds nothing
    // items: the walks stop at subtrees of at most task_nodes nodes; a
    // query's items (own points of walked nodes, covered and cut subtrees)
    // are counted per query (tcounts), listed from toffsets[q] (tasks:
    // node, query, kind) and written from item_offsets[i]
    int task_nodes;
    long long* tcounts;
    const long long* toffsets;
    int4* tasks;
    const long long* item_offsets;
};
// list 0: tcounts[q] = items of query q; list 1: the items at toffsets[q]
void launch_range_top(const RangeArgs& a, int list, hipStream_t s);
// counts[i] = points of item i
void launch_range_item_count(const RangeArgs& a, long long ntask, long long* counts, hipStream_t s);
// offsets[q] = item_off[toffsets[q]] for q in [0, nq]
void launch_range_query_offsets(const long long* toffsets, const long long* item_off, int nq, long long* offsets,
                                hipStream_t s);
void launch_range_item_fill(const RangeArgs& a, long long ntask, const long long* item_off, hipStream_t s);

// ---------------------------------------------------------------- misc
void launch_fill_i32(int* p, int v, long long n, hipStream_t s);
void launch_pack_points(const float* xyz, const uint32_t* ids, uint32_t id_base, int n, float4* out,
                        hipStream_t s);

// ---------------------------------------------------------------- primitives (rocPRIM)
hipError_t prim_sort_pairs_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* kin,
                               unsigned long long* kout, const uint32_t* vin, uint32_t* vout, int n,
                               int begin_bit, int end_bit, hipStream_t s);
hipError_t prim_sort_keys_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* kin,
                              unsigned long long* kout, int n, int begin_bit, int end_bit, hipStream_t s);
hipError_t prim_exscan_u64(void* tmp, size_t& tmp_bytes, const unsigned long long* in,
                           unsigned long long* out, int n, hipStream_t s);
hipError_t prim_exscan_i32(void* tmp, size_t& tmp_bytes, const int* in, int* out, int n, hipStream_t s);
hipError_t prim_exscan_i64(void* tmp, size_t& tmp_bytes, const long long* in, long long* out, int n,
                           hipStream_t s);

}  // namespace ikd

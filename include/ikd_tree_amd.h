/*
 * ikd_tree_amd.h — C ABI of the MI355X ikd-Tree library (libikd_hip.so).
 *
 * Each entry point replaces one method of hku-mars/ikd-Tree's
 * KD_TREE<PointType> (ikd-Tree/ikd_Tree.h:225-248); the citation above each
 * declaration names the reference interface it stands in for.  The tree lives
 * in device memory (HBM) and all work runs as HIP kernels on gfx950, on one
 * stream per tree.
 *
 * Conventions
 *  - Points are packed float triples (x, y, z), 12 bytes each.  Every point
 *    carries a uint32 payload id; queries return ids so callers can map
 *    results back to full point records (the reference returns PointType
 *    copies, ikd_Tree.cpp:392).  When an ids pointer is NULL, Build numbers
 *    points 0..n-1 and Add_Points continues from the last id handed out.
 *  - Boxes are 6 floats (min x, y, z, max x, y, z) and are half-open
 *    [min, max) on every axis (BoxPointType, ikd_Tree.h:32-35).
 *  - Entry points without the _device suffix take host pointers and return
 *    when the work is finished.  _device entry points take device pointers,
 *    are ordered on the tree's stream and return without waiting.
 *  - Errors are negative return codes; the reference has no error channel
 *    and is undefined on the same inputs (k <= 0, Add_Points on an empty tree).
 *  - One writer per tree; calls on one tree are serialized by the caller.
 */
#ifndef IKD_TREE_AMD_H
#define IKD_TREE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IKD_OK 0
#define IKD_ERR_ARG -1         /* invalid argument (null handle, k <= 0, negative size) */
#define IKD_ERR_HIP -2         /* a HIP runtime call failed */
#define IKD_ERR_NOMEM -3       /* device allocation failed */
#define IKD_ERR_CAPACITY -4    /* caller's output buffer too small; counts are valid */
#define IKD_ERR_UNSUPPORTED -5 /* k above 64, or voxel indices beyond +-2^20 */
#define IKD_ERR_INTERNAL -6    /* a device traversal stack overflowed */

typedef struct ikd_tree_s* ikd_tree_t;

/* Event counts of one tree since ikd_create, kept on the host (no
 * counterpart in the reference; reading them waits for nothing). */
typedef struct ikd_counters {
    int64_t rebuilds;       /* subtrees rebuilt (Rebuild, ikd_Tree.cpp:500-533), deferred ones included */
    int64_t rebuilt_points; /* live points those rebuilds re-sorted */
    int64_t compactions;    /* whole-tree rebuilds: pool mostly garbage, or the tree deeper than 62 levels */
    int64_t pool_grows;     /* node-pool reallocations */
    int64_t scratch_grows;  /* scratch-buffer allocations and reallocations */
    int64_t host_waits;     /* host waits on the tree's stream inside the library's calls */
} ikd_counters_t;

/* KD_TREE(float delete_param=0.5, float balance_param=0.6, float box_length=0.2)
 * ikd_Tree.h:226, ikd_Tree.cpp:9-17.  Uses the current HIP device. */
int ikd_create(float delete_param, float balance_param, float box_length, ikd_tree_t* out);

/* ~KD_TREE, ikd_Tree.cpp:19-27 */
int ikd_destroy(ikd_tree_t t);

/* Set_delete_criterion_param / Set_balance_criterion_param /
 * set_downsample_param, ikd_Tree.cpp:29-42 */
int ikd_set_delete_criterion_param(ikd_tree_t t, float delete_param);
int ikd_set_balance_criterion_param(ikd_tree_t t, float balance_param);
int ikd_set_downsample_param(ikd_tree_t t, float box_length);

/* InitializeKDTree(delete_param, balance_param, box_length), ikd_Tree.cpp:44-49 */
int ikd_initialize(ikd_tree_t t, float delete_param, float balance_param, float box_length);

/* Build(PointVector point_cloud), ikd_Tree.cpp:352-364 */
int ikd_build(ikd_tree_t t, const float* xyz, const uint32_t* ids, int64_t n);

/* Nearest_Search(point, k, Nearest_Points, Point_Distance, max_dist),
 * ikd_Tree.cpp:366-397, batched over nq queries.  Query i's results are at
 * [i*k, i*k + out_counts[i]) in ascending squared distance; unused slots get
 * d2 = +inf and id = 0xffffffff.  max_dist may be INFINITY.  k <= 64. */
int ikd_nearest_search(ikd_tree_t t, const float* queries, int64_t nq, int k, double max_dist,
                       float* out_xyz, float* out_d2, uint32_t* out_ids, int32_t* out_counts);

/* Same, with every pointer in device memory; ordered on the tree's stream. */
int ikd_nearest_search_device(ikd_tree_t t, const float* d_queries, int64_t nq, int k, double max_dist,
                              float* d_out_xyz, float* d_out_d2, uint32_t* d_out_ids, int32_t* d_out_counts);

/* Instrumented Nearest_Search: returns, per query, the number of nodes visited
 * and of child AABBs evaluated by the one-lane depth-first walk (nearer child
 * first, the reference's traversal, ikd_Tree.cpp:868-1013), whatever launch
 * shape the tree is tuned to (no counterpart in the reference; used to compute
 * algorithmic bytes for the roofline).  Host pointers. */
int ikd_nearest_search_counters(ikd_tree_t t, const float* queries, int64_t nq, int k, double max_dist,
                                uint32_t* out_visits, uint32_t* out_boxes);

/* Box_Search(Box_of_Point, Storage), ikd_Tree.cpp:399-404, batched over nb
 * boxes.  out_offsets (nb+1 entries) is always filled; box i's points are at
 * [out_offsets[i], out_offsets[i+1]).  If out_xyz is NULL or capacity <
 * out_offsets[nb], only the offsets are produced (IKD_ERR_CAPACITY when
 * out_xyz was given).  Result order within a box is unspecified. */
int ikd_box_search(ikd_tree_t t, const float* boxes, int64_t nb, int64_t* out_offsets, float* out_xyz,
                   uint32_t* out_ids, int64_t capacity);

/* Radius_Search(point, radius, Storage), ikd_Tree.cpp:406-411, batched;
 * same output convention as ikd_box_search. */
int ikd_radius_search(ikd_tree_t t, const float* centers, const float* radii, int64_t nq, int64_t* out_offsets,
                      float* out_xyz, uint32_t* out_ids, int64_t capacity);

/* Add_Points(PointToAdd, downsample_on), ikd_Tree.cpp:413-489.  *out_added
 * receives the reference's return value (0 when downsample_on is 0). */
int ikd_add_points(ikd_tree_t t, const float* xyz, const uint32_t* ids, int64_t n, int downsample_on,
                   int32_t* out_added);

/* Same with xyz/ids in device memory (ids may be NULL). */
int ikd_add_points_device(ikd_tree_t t, const float* d_xyz, const uint32_t* d_ids, int64_t n, int downsample_on,
                          int32_t* out_added);

/* Delete_Points(PointToDel), ikd_Tree.cpp:513-533 */
int ikd_delete_points(ikd_tree_t t, const float* xyz, int64_t n);

/* Delete_Point_Boxes(BoxPoints), ikd_Tree.cpp:535-556; *out_deleted receives
 * the number of live points removed. */
int ikd_delete_point_boxes(ikd_tree_t t, const float* boxes, int64_t nb, int32_t* out_deleted);

/* Add_Point_Boxes(BoxPoints), ikd_Tree.cpp:491-511 (Add_by_range, cpp:762-815):
 * gives back the deleted points the boxes reach that the downsample did not
 * delete, box by box.  As in the reference, the walk decides by the nodes'
 * stored ranges, which leave deleted subtrees out: a box that covers a node's
 * range restores that whole subtree (points outside the box included), and
 * deleted points inside a box stay deleted where an ancestor's range misses
 * them. */
int ikd_add_point_boxes(ikd_tree_t t, const float* boxes, int64_t nb);

/* size(), ikd_Tree.cpp:78-96: nodes in the tree including lazily deleted ones */
int ikd_size(ikd_tree_t t, int64_t* out);

/* validnum(), ikd_Tree.cpp:128-145: live points */
int ikd_validnum(ikd_tree_t t, int64_t* out);

/* tree_range(), ikd_Tree.cpp:98-126: root AABB (min xyz, max xyz); zeros when empty */
int ikd_tree_range(ikd_tree_t t, float* out6);

/* root_alpha(alpha_bal, alpha_del), ikd_Tree.cpp:147-165: the values the root's
 * last Update stored (cpp:1315-1321); 0 after Build, whose BuildTree runs
 * before the root is the tree's root (the reference leaves them unset). */
int ikd_root_alpha(ikd_tree_t t, float* alpha_bal, float* alpha_del);

/* flatten(Root_Node, Storage, NOT_RECORD), ikd_Tree.cpp:1325-1352: all live
 * points.  *out_n receives the count; at most capacity points are written
 * (IKD_ERR_CAPACITY if more exist). */
int ikd_flatten(ikd_tree_t t, float* out_xyz, uint32_t* out_ids, int64_t capacity, int64_t* out_n);

/* Wait for the tree's stream; return its hipStream_t (as void*). */
int ikd_synchronize(ikd_tree_t t);
int ikd_stream(ikd_tree_t t, void** out_stream);

/* Diagnostics (no counterpart in the reference): pool nodes in use, node
 * capacity, deepest node, rebuilds run so far. */
int ikd_pool_stats(ikd_tree_t t, int64_t* used_nodes, int64_t* capacity_nodes, int32_t* max_depth,
                   int64_t* rebuilds);
/* acquire_removed_points, ikd_Tree.cpp:558-571: the points rebuilds removed
 * since the last call (recorded while IKD_TUNE_RECORD_REMOVED is 1), then the
 * record is cleared.  *out_n gets the count; when it exceeds capacity the call
 * returns IKD_ERR_CAPACITY and keeps the record.  Which deleted points a
 * rebuild has dropped so far depends on when rebuilds ran, which differs from
 * the reference (rebuilds run after each call here, inside the per-point
 * loops there). */
int ikd_acquire_removed_points(ikd_tree_t t, float* out_xyz, uint32_t* out_ids, int64_t capacity, int64_t* out_n);
/* Diagnostics (no counterpart in the reference): the tree's event counts. */
int ikd_get_counters(ikd_tree_t t, ikd_counters_t* out);
/* Device memory the tree holds: node pool, scratch, removed-point record and
 * status block, in bytes (a host-side figure; no device call).  No
 * counterpart in the reference, whose nodes live in host memory; used by the
 * server's HBM admission. */
int ikd_device_bytes(ikd_tree_t t, int64_t* out_bytes);

/* Launch tuning (no counterpart in the reference; results never depend on it):
 *  IKD_TUNE_QUERY_ORDER 1 sorts batches of >= 4096 kNN queries into
 *                       Morton order before the launch, 0 (default) keeps input order;
 *  IKD_TUNE_KNN_STACK   minimum LDS stack entries per lane, 0..64 (default 0:
 *                       the tree depth + 2, enough for any query);
 *  IKD_TUNE_LANES_PER_QUERY -1 (default: from the batch size); 0 persistent
 *                       lanes (one query per lane at a time, idle lanes fetch
 *                       the next query); or 1, 2, 4, 8, 16 lanes sharing one
 *                       query's traversal (64 / lanes queries per wavefront). */
#define IKD_TUNE_QUERY_ORDER 1
#define IKD_TUNE_KNN_STACK 2
#define IKD_TUNE_LANES_PER_QUERY 3
/*  IKD_TUNE_SMALL_BUILD 1 (default) builds every subtree of <= 1024 points
 *                       with one workgroup in one launch; 0 always uses the
 *                       level-synchronous build (identical trees). */
#define IKD_TUNE_SMALL_BUILD 4
/*  IKD_TUNE_DEFER_REBUILD N > 0: a subtree with more than N live points that
 *                       fails Criterion_Check is not rebuilt inside the update
 *                       call but listed for ikd_rebuild_deferred (the reference
 *                       hands subtrees of >= 1500 nodes to a background thread,
 *                       ikd_Tree.cpp:627-633); 0 (default) rebuilds at once.
 *                       Query results and live sets never depend on it. */
#define IKD_TUNE_DEFER_REBUILD 5
/*  IKD_TUNE_FOLD_LANES  lanes that walk one voxel's box in Add_Points'
 *                       downsample: 1, 4, 8 or 16; 0 (default) takes 8 for
 *                       maps of >= 4096 nodes and 1 below (identical results). */
#define IKD_TUNE_FOLD_LANES 6
/*  IKD_TUNE_RECORD_REMOVED 1: rebuilds record every deleted point they drop
 *                       that the downsample did not delete, for
 *                       ikd_acquire_removed_points (flatten's DELETE_POINTS_REC
 *                       mode, ikd_Tree.cpp:638,1338-1341); 0 (default): no
 *                       record, and rebuilds skip deleted subtrees. */
#define IKD_TUNE_RECORD_REMOVED 7
/*  IKD_TUNE_KNN_STAGE   levels of the tree (0..12, default 0) that one-lane
 *                       kNN launches copy into each workgroup's LDS and read
 *                       from there (capped so that a workgroup's LDS stays
 *                       within 64 KiB); the results are the same. */
#define IKD_TUNE_KNN_STAGE 8
/*  IKD_TUNE_KNN_COMPACT 1 (default): one-lane kNN launches over trees of at
 *                       most 2^24 node slots keep 4-byte traversal stack
 *                       entries in LDS (more wavefronts resident); 0: the
 *                       6-byte entries.  Results are the same.  Also a
 *                       batch key. */
#define IKD_TUNE_KNN_COMPACT 9
int ikd_set_tuning(ikd_tree_t t, int key, int value);

/* Run the deferred rebuilds that are still due, in the order they were
 * listed, until their live points exceed max_points (at least one; <= 0: no
 * limit); *out_rebuilt gets the number run.  Rebuild_Thread's work
 * (ikd_Tree.cpp:200-350), run when the caller has time, e.g. between scans.
 * With IKD_TUNE_DEFER_REBUILD on, an update that leaves the node pool more
 * than 4x live points + 4M nodes leaves its reclamation (a whole-tree
 * rebuild) to this call, which then does only that, whatever max_points
 * says; past 8x live + 8M nodes, or when the tree is too deep for the
 * traversal stacks, the update reclaims at once. */
int ikd_rebuild_deferred(ikd_tree_t t, int64_t max_points, int32_t* out_rebuilt);

/* Payload-id housekeeping (no counterpart in the reference; for callers that
 * keep point records by payload id, as the C++ facade does).
 * ikd_list_payload_ids: the ids the tree can still return, ascending and
 * unique: every node, live or deleted, and every point in the removed-point
 * record; *out_n gets the count, out_ids may be NULL to query it.
 * ikd_relabel_payload_ids: gives the id old_ids[i] the new id i (old_ids
 * strictly ascending, typically the list above) and continues automatic ids
 * from n.  Together they let a caller drop the records of points the tree
 * no longer holds and keep ids below 2^32 however many points pass through. */
int ikd_list_payload_ids(ikd_tree_t t, uint32_t* out_ids, int64_t capacity, int64_t* out_n);
int ikd_relabel_payload_ids(ikd_tree_t t, const uint32_t* old_ids, int64_t n);

/* Pre-order dump of the tree (diagnostic, no counterpart in the reference):
 * per node x, y, z, id (4 x 4 bytes), axis, TreeSize, invalid_point_num, flags
 * (bit 0 point_deleted, 1 tree_deleted, 2/3 downsample labels), range[6]
 * (min xyz, max xyz), left and right row index or -1: 64 bytes per row.
 * *out_n gets the row count; out_rows may be NULL to query it. */
int ikd_dump_nodes(ikd_tree_t t, void* out_rows, int64_t capacity, int64_t* out_n);

/* Text of the last error on this tree (never NULL). */
const char* ikd_last_error(ikd_tree_t t);

/* ---- Multi-tree batches (no counterpart in the reference: the batched form
 * of Nearest_Search, ikd_Tree.cpp:366-397, over many trees in one launch). ---
 * A batch owns a stream.  A launch waits for the pending work of every tree
 * it reads, and every such tree's later work waits for the launch. */
typedef struct ikd_batch_s* ikd_batch_t;
int ikd_batch_create(ikd_batch_t* out);
int ikd_batch_destroy(ikd_batch_t b);

/* Nearest_Search for nq queries spread over trees[0..ntrees): query i searches
 * trees[d_query_tree[i]] (an index outside the list finds nothing) with
 * k = d_k[i] clamped to [0, k_max] (d_k NULL: k_max for all) and max_dist =
 * d_max_dist[i] (d_max_dist NULL: max_dist for all).  Row i of the outputs
 * has k_max slots: results ascending in [0, count), then empty slots (d2 =
 * +inf, id = 0xffffffff).  All pointers are device memory; ordered on the
 * batch stream; results are identical to one ikd_nearest_search_device per
 * tree.  k_max <= 64. */
int ikd_batch_nearest_search_device(ikd_batch_t b, const ikd_tree_t* trees, int ntrees, const int32_t* d_query_tree,
                                    const float* d_queries, int64_t nq, const int32_t* d_k, int k_max,
                                    const double* d_max_dist, double max_dist, float* d_out_xyz, float* d_out_d2,
                                    uint32_t* d_out_ids, int32_t* d_out_counts);

/* Allocation-free form of ikd_batch_nearest_search_device, with the same
 * arguments and results.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It calls no allocating
 * routine: the tree table, the per-query thresholds and the launch counters
 * live in a static device array of each GPU (about 4 MiB, held until the
 * process ends) that every call on that GPU shares, so calls that overlap on
 * one GPU race; ikd_batch_synchronize reports a traversal overflow of the
 * batch's last such call.  Limits: ntrees <= IKD_NOALLOC_MAX_TREES and, with
 * d_max_dist, nq <= IKD_NOALLOC_MAX_QUERIES (IKD_ERR_CAPACITY beyond them);
 * a tree deeper than the traversal stacks gives IKD_ERR_UNSUPPORTED (any
 * update call compacts it). */
#define IKD_NOALLOC_MAX_TREES 128
#define IKD_NOALLOC_MAX_QUERIES (1 << 20)
int ikd_batch_nearest_search_noalloc(ikd_batch_t b, const ikd_tree_t* trees, int ntrees,
                                     const int32_t* d_query_tree, const float* d_queries, int64_t nq,
                                     const int32_t* d_k, int k_max, const double* d_max_dist, double max_dist,
                                     float* d_out_xyz, float* d_out_d2, uint32_t* d_out_ids,
                                     int32_t* d_out_counts);

/* The same launch recorded once into a HIP graph and replayed (no
 * counterpart in the reference; for fixed serving loops that re-run one query
 * layout: the replay costs one graph launch instead of the entry point's host
 * work).  Create records the kNN of ikd_batch_nearest_search_device with
 * these exact arguments (the device buffers are read and written at every
 * replay, so the caller refills d_queries etc. between replays).  Launch
 * replays it on the batch stream, ordered after the trees' pending work and
 * before their later work, with the tree table refreshed from the trees'
 * current node pools; IKD_ERR_UNSUPPORTED when a tree has grown more than 8
 * levels deeper than at the recording (record again then).  Overflow is
 * reported by ikd_batch_synchronize.  The recording holds only kernels. */
typedef struct ikd_knn_graph_s* ikd_knn_graph_t;
int ikd_batch_knn_graph_create(ikd_batch_t b, const ikd_tree_t* trees, int ntrees, const int32_t* d_query_tree,
                               const float* d_queries, int64_t nq, const int32_t* d_k, int k_max,
                               const double* d_max_dist, double max_dist, float* d_out_xyz, float* d_out_d2,
                               uint32_t* d_out_ids, int32_t* d_out_counts, ikd_knn_graph_t* out);
int ikd_batch_knn_graph_launch(ikd_knn_graph_t g);
int ikd_batch_knn_graph_destroy(ikd_knn_graph_t g);

/* Radius_Search (ikd_Tree.cpp:406-411) for nq queries spread over trees
 * (count walks, then one lane per piece of each query's result): query i searches
 * trees[d_query_tree[i]] around d_centers[3i..] with radius d_radii[i] (an
 * index outside the list finds nothing).  d_out_offsets (device, nq+1) gets
 * the exclusive prefix of the counts and *out_total (host) the total; the
 * points and ids are written to d_out_xyz / d_out_ids when both are given and
 * capacity >= total (else IKD_ERR_CAPACITY if they were given).  Result order
 * within a query is unspecified.  Waits twice on the host (the piece count,
 * then the total). */
int ikd_batch_radius_search_device(ikd_batch_t b, const ikd_tree_t* trees, int ntrees, const int32_t* d_query_tree,
                                   const float* d_centers, const float* d_radii, int64_t nq, int64_t* d_out_offsets,
                                   float* d_out_xyz, uint32_t* d_out_ids, int64_t capacity, int64_t* out_total);

/* Wait for the batch stream; IKD_ERR_INTERNAL if a traversal overflowed. */
int ikd_batch_synchronize(ikd_batch_t b);
int ikd_batch_stream(ikd_batch_t b, void** out_stream);
/* IKD_TUNE_LANES_PER_QUERY and IKD_TUNE_KNN_STACK, as for a tree. */
int ikd_batch_set_tuning(ikd_batch_t b, int key, int value);
const char* ikd_batch_last_error(ikd_batch_t b);

/* ---- Serving (no counterpart in the reference): a request loop over many
 * device-resident maps.  A request is one scan for one map: Nearest_Search of
 * every scan point (ikd_Tree.cpp:366-397), then Add_Points of the scan
 * (cpp:413-489), then an optional Delete_Point_Boxes trim (cpp:535-556).
 * `slots` worker threads each take the oldest pending request whose map is
 * idle and whose device memory fits the HBM budget; requests for one map run
 * in submission order, one at a time, so each result is identical to running
 * the same requests alone in that order; requests for different maps overlap
 * on the device (one stream per map). ---- */
typedef struct ikd_server_s* ikd_server_t;

typedef struct ikd_request {
    /* set by the caller; the request and every buffer it names stay owned by
     * the caller and untouched until ikd_server_wait hands the request back */
    int32_t map;              /* index returned by ikd_server_add_map */
    int32_t n;                /* scan points */
    const float* xyz;         /* host, 3n floats */
    int32_t k;                /* kNN per scan point before the insert; 0 skips it (k <= 64) */
    double max_dist;
    int32_t downsample;       /* Add_Points(scan, downsample); -1 skips the insert */
    int32_t n_trim;           /* boxes deleted after the insert */
    const float* trim_boxes;  /* host, 6 n_trim floats */
    float* out_xyz;           /* host, n*k*3 floats, or NULL */
    float* out_d2;            /* host, n*k floats, or NULL */
    uint32_t* out_ids;        /* host, n*k, or NULL */
    int32_t* out_counts;      /* host, n, or NULL */
    int64_t t_arrival_ns;     /* caller's ikd_now_ns() at arrival */
    int64_t tag;              /* caller's */
    /* set by the server */
    int64_t t_start_ns, t_done_ns;
    int32_t added, trimmed, status;
    int64_t t_insert_done_ns; /* when Add_Points (and the kNN before it) had finished */
    ikd_counters_t delta;     /* the map's counters over this request */
    /* set by the caller */
    int32_t flags;            /* IKD_REQ_* bits */
} ikd_request_t;

/* out_xyz, out_d2, out_ids and out_counts are page-locked memory from
 * ikd_host_alloc: the results are copied into them by DMA, with no host copy
 * through the server's own buffers. */
#define IKD_REQ_PINNED_RESULTS 1

/* Page-locked host memory (hipHostMalloc) for request buffers; free it with
 * ikd_host_free once no request that names it is in the server.  No
 * counterpart in the reference, whose searches write host vectors. */
int ikd_host_alloc(int64_t bytes, void** out);
int ikd_host_free(void* p);

/* Monotonic clock of the server (ns). */
int64_t ikd_now_ns(void);
int ikd_server_create(int slots, int64_t hbm_budget_bytes, ikd_server_t* out);
/* Register a map; the tree must outlive the server and get no other calls
 * while the server runs. */
int ikd_server_add_map(ikd_server_t s, ikd_tree_t tree, int32_t* out_map);
int ikd_server_submit(ikd_server_t s, ikd_request_t* req);
/* Hand back up to max_out finished requests (timeout_ms < 0: wait for one, or
 * until nothing is pending or running; 0: do not wait). */
int ikd_server_wait(ikd_server_t s, ikd_request_t** out, int32_t max_out, int32_t timeout_ms, int32_t* out_n);
/* Requests waiting; requests running plus maintenance steps running or
 * queued (both 0: the server is idle). */
int ikd_server_pending(ikd_server_t s, int64_t* out_pending, int64_t* out_in_flight);
/* A map that has served a request is queued for maintenance: while no
 * request is waiting, one worker at a time runs ikd_rebuild_deferred(map,
 * max_points) on the oldest queued map that is not busy, and queues the map
 * again while such steps still find work.  A request for a map in
 * maintenance waits for at most that one step (-1, the default: one call
 * with no limit; 0: never).  Deferral itself is per tree
 * (IKD_TUNE_DEFER_REBUILD). */
int ikd_server_set_maintenance(ikd_server_t s, int64_t max_points);
/* Deferred rebuilds the workers have run so far. */
int ikd_server_stats(ikd_server_t s, int64_t* out_rebuilds);
/* Device bytes the registered maps hold (as of their last request or
 * maintenance step) and the estimate of the requests now running.  A request
 * is admitted while resident + running + its own estimate fits the budget;
 * with nothing running the oldest runnable request is always admitted. */
int ikd_server_memory(ikd_server_t s, int64_t* out_resident, int64_t* out_in_flight_bytes);
/* Finish every submitted request, then stop the workers. */
int ikd_server_destroy(ikd_server_t s);

#ifdef __cplusplus
}
#endif

#endif /* IKD_TREE_AMD_H */

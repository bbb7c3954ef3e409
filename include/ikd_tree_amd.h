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

/* Instrumented Nearest_Search: also returns, per query, the number of nodes
 * visited and of child AABBs evaluated (no counterpart in the reference; used
 * to compute algorithmic bytes for the roofline).  Host pointers. */
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

/* size(), ikd_Tree.cpp:78-96: nodes in the tree including lazily deleted ones */
int ikd_size(ikd_tree_t t, int64_t* out);

/* validnum(), ikd_Tree.cpp:128-145: live points */
int ikd_validnum(ikd_tree_t t, int64_t* out);

/* tree_range(), ikd_Tree.cpp:98-126: root AABB (min xyz, max xyz); zeros when empty */
int ikd_tree_range(ikd_tree_t t, float* out6);

/* root_alpha(alpha_bal, alpha_del), ikd_Tree.cpp:147-165 */
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

/* Launch tuning (no counterpart in the reference; results never depend on it):
 *  IKD_TUNE_QUERY_ORDER 1 sorts batches of >= 4096 kNN queries into
 *                       Morton order before the launch, 0 (default) keeps input order;
 *  IKD_TUNE_KNN_STACK   0 (default: from the tree depth), 32 or 64 LDS stack
 *                       entries per lane. */
#define IKD_TUNE_QUERY_ORDER 1
#define IKD_TUNE_KNN_STACK 2
int ikd_set_tuning(ikd_tree_t t, int key, int value);

/* Text of the last error on this tree (never NULL). */
const char* ikd_last_error(ikd_tree_t t);

#ifdef __cplusplus
}
#endif

#endif /* IKD_TREE_AMD_H */

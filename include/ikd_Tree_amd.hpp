// ikd_Tree_amd.hpp — header-only drop-in for hku-mars/ikd-Tree's
// KD_TREE<PointType> (ikd-Tree/ikd_Tree.h:59-249), backed by the MI355X
// library through its C ABI (ikd_tree_amd.h, libikd_hip.so).
//
// The public methods keep the reference's names, parameter lists, defaults and
// output conventions (ikd_Tree.h:226-248): outputs go into caller-owned
// vectors that are cleared first, Nearest_Search returns points in ascending
// squared distance, Box/Radius_Search return the live points found.  The
// public members Root_Node, PCL_Storage and max_queue_size, the Ptr alias and
// the delete_point_storage_set modes are there too (ikd_Tree.h:39, 62-63,
// 245-248).  The differences a caller can observe:
//   - errors throw std::runtime_error where the reference is undefined
//     (k <= 0, k > 64);
//   - Add_Points works on an empty tree (the reference dereferences a null
//     root there, ikd_Tree.cpp:447);
//   - Root_Node is an opaque handle: null while the tree has no nodes (as the
//     reference's is before Build), else the one node flatten() accepts,
//     standing for the whole tree; flatten() supports the NOT_RECORD mode;
//   - the library's rebuilds do not use PCL_Storage as scratch (the
//     reference's Rebuild leaves the last rebuilt subtree's points in it);
//   - PointVector uses std::allocator unless Eigen is included first (then
//     Eigen::aligned_allocator, as the reference, so PCL point vectors and
//     PointVector assign to each other).
// Payload types (e.g. pcl::PointXYZI) round-trip unchanged: the facade keeps
// every point the tree may still return host-side under its payload id and
// looks results up by id, so fields beyond x, y, z come back as they went in.
// Records of points the tree no longer holds (dropped by the downsample,
// removed by rebuilds and handed out) are reclaimed once the store has grown
// to twice what the tree can still return, and ids are renumbered from 0 then,
// so a long mapping run neither grows the store nor wraps the 32-bit ids.
//
// Threading: one writer per tree, as in the reference; reads (Nearest_Search,
// Box_Search, Radius_Search) may come from many threads at once, e.g. an
// OpenMP loop of Nearest_Search calls: the library serializes calls on one
// tree onto its stream.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ikd_tree_amd.h"

#if defined(EIGEN_WORLD_VERSION) && !defined(IKD_AMD_STD_ALLOCATOR)
template <class T>
using ikd_point_allocator = Eigen::aligned_allocator<T>;
#else
template <class T>
using ikd_point_allocator = std::allocator<T>;
#endif

#ifndef IKD_AMD_NO_STORAGE_ENUM
// flatten's record modes, ikd_Tree.h:39
enum delete_point_storage_set { NOT_RECORD, DELETE_POINTS_REC, MULTI_THREAD_REC };
#endif

#ifndef IKD_AMD_NO_BASIC_TYPES
// ikdTree_PointType and BoxPointType, ikd_Tree.h:22-35
struct ikdTree_PointType {
    float x, y, z;
    ikdTree_PointType(float px = 0.0f, float py = 0.0f, float pz = 0.0f) : x(px), y(py), z(pz) {}
};

struct BoxPointType {
    float vertex_min[3];
    float vertex_max[3];
};
#endif

template <typename PointType>
class KD_TREE {
public:
    using PointVector = std::vector<PointType, ikd_point_allocator<PointType>>;
    using Ptr = std::shared_ptr<KD_TREE<PointType>>;  // ikd_Tree.h:63

    // Opaque stand-in for the reference's node type (ikd_Tree.h:64-86): the
    // tree's nodes live in device memory, so the only node a caller can hold
    // is Root_Node, which stands for the whole tree.
    struct KD_TREE_NODE {
        const KD_TREE* tree;
    };

    KD_TREE(float delete_param = 0.5, float balance_param = 0.6, float box_length = 0.2) : root_handle_{this} {
        check(ikd_create(delete_param, balance_param, box_length, &h_), "KD_TREE");
        check(ikd_set_tuning(h_, IKD_TUNE_RECORD_REMOVED, 1), "KD_TREE");
    }
    ~KD_TREE() {
        if (h_) ikd_destroy(h_);
    }
    KD_TREE(const KD_TREE&) = delete;
    KD_TREE& operator=(const KD_TREE&) = delete;

    void Set_delete_criterion_param(float delete_param) { check(ikd_set_delete_criterion_param(h_, delete_param)); }
    void Set_balance_criterion_param(float balance_param) { check(ikd_set_balance_criterion_param(h_, balance_param)); }
    void set_downsample_param(float box_length) { check(ikd_set_downsample_param(h_, box_length)); }
    void InitializeKDTree(float delete_param = 0.5, float balance_param = 0.7, float box_length = 0.2) {
        check(ikd_initialize(h_, delete_param, balance_param, box_length));
    }

    int size() {
        int64_t v = 0;
        check(ikd_size(h_, &v));
        return (int)v;
    }
    int validnum() {
        int64_t v = 0;
        check(ikd_validnum(h_, &v));
        return (int)v;
    }
    void root_alpha(float& alpha_bal, float& alpha_del) { check(ikd_root_alpha(h_, &alpha_bal, &alpha_del)); }
    BoxPointType tree_range() {
        float r[6];
        check(ikd_tree_range(h_, r));
        BoxPointType b;
        for (int a = 0; a < 3; ++a) {
            b.vertex_min[a] = r[a];
            b.vertex_max[a] = r[3 + a];
        }
        return b;
    }

    // Build(PointVector), ikd_Tree.cpp:352-364
    void Build(PointVector point_cloud) {
        // The removed-point record outlives Build (the reference keeps
        // Points_deleted): turn its ids into records before the id table goes.
        take_removed(pending_removed_);
        payload_ = std::move(point_cloud);
        std::vector<float> xyz = pack(payload_);
        std::vector<uint32_t> ids(payload_.size());
        for (size_t i = 0; i < ids.size(); ++i) ids[i] = (uint32_t)i;
        check(ikd_build(h_, xyz.data(), ids.data(), (int64_t)payload_.size()), "Build");
        reclaim_at_ = std::max(2 * payload_.size(), min_reclaim_) + min_reclaim_;
        sync_root();
    }

    // Nearest_Search, ikd_Tree.cpp:366-397 (one query; see Nearest_Search_Batch)
    void Nearest_Search(PointType point, int k_nearest, PointVector& Nearest_Points, std::vector<float>& Point_Distance,
                        double max_dist = INFINITY) {
        if (k_nearest <= 0) throw std::runtime_error("Nearest_Search: k_nearest must be positive");
        float q[3] = {point.x, point.y, point.z};
        std::vector<float> xyz(3 * (size_t)k_nearest), d2((size_t)k_nearest);
        std::vector<uint32_t> ids((size_t)k_nearest);
        int32_t cnt = 0;
        check(ikd_nearest_search(h_, q, 1, k_nearest, max_dist, xyz.data(), d2.data(), ids.data(), &cnt),
              "Nearest_Search");
        PointVector().swap(Nearest_Points);
        std::vector<float>().swap(Point_Distance);
        for (int i = 0; i < cnt; ++i) {
            Nearest_Points.push_back(payload_[ids[i]]);
            Point_Distance.push_back(d2[i]);
        }
    }

    // Many queries in one launch (the batched form of Nearest_Search).
    void Nearest_Search_Batch(const PointVector& queries, int k_nearest, std::vector<PointVector>& Nearest_Points,
                              std::vector<std::vector<float>>& Point_Distance, double max_dist = INFINITY) {
        if (k_nearest <= 0) throw std::runtime_error("Nearest_Search_Batch: k_nearest must be positive");
        const size_t nq = queries.size();
        std::vector<float> q = pack(queries);
        std::vector<float> xyz(3 * nq * k_nearest), d2(nq * k_nearest);
        std::vector<uint32_t> ids(nq * k_nearest);
        std::vector<int32_t> cnt(nq);
        if (nq)
            check(ikd_nearest_search(h_, q.data(), (int64_t)nq, k_nearest, max_dist, xyz.data(), d2.data(), ids.data(),
                                     cnt.data()),
                  "Nearest_Search_Batch");
        Nearest_Points.assign(nq, PointVector());
        Point_Distance.assign(nq, std::vector<float>());
        for (size_t i = 0; i < nq; ++i)
            for (int j = 0; j < cnt[i]; ++j) {
                Nearest_Points[i].push_back(payload_[ids[i * k_nearest + j]]);
                Point_Distance[i].push_back(d2[i * k_nearest + j]);
            }
    }

    // Box_Search, ikd_Tree.cpp:399-404
    void Box_Search(const BoxPointType& Box_of_Point, PointVector& Storage) {
        Storage.clear();
        float b[6];
        for (int a = 0; a < 3; ++a) {
            b[a] = Box_of_Point.vertex_min[a];
            b[3 + a] = Box_of_Point.vertex_max[a];
        }
        int64_t off[2];
        check(ikd_box_search(h_, b, 1, off, nullptr, nullptr, 0), "Box_Search");
        fetch_range(off[1], [&](float* xyz, uint32_t* ids, int64_t cap) {
            return ikd_box_search(h_, b, 1, off, xyz, ids, cap);
        }, Storage);
    }

    // Radius_Search, ikd_Tree.cpp:406-411
    void Radius_Search(PointType point, const float radius, PointVector& Storage) {
        Storage.clear();
        float c[3] = {point.x, point.y, point.z};
        float r = radius;
        int64_t off[2];
        check(ikd_radius_search(h_, c, &r, 1, off, nullptr, nullptr, 0), "Radius_Search");
        fetch_range(off[1], [&](float* xyz, uint32_t* ids, int64_t cap) {
            return ikd_radius_search(h_, c, &r, 1, off, xyz, ids, cap);
        }, Storage);
    }

    // Add_Points, ikd_Tree.cpp:413-489
    int Add_Points(PointVector& PointToAdd, bool downsample_on) {
        if (PointToAdd.empty()) return 0;
        std::vector<float> xyz = pack(PointToAdd);
        std::vector<uint32_t> ids(PointToAdd.size());
        for (size_t i = 0; i < ids.size(); ++i) ids[i] = (uint32_t)(payload_.size() + i);
        payload_.insert(payload_.end(), PointToAdd.begin(), PointToAdd.end());
        int32_t added = 0;
        check(ikd_add_points(h_, xyz.data(), ids.data(), (int64_t)PointToAdd.size(), downsample_on ? 1 : 0, &added),
              "Add_Points");
        sync_root();
        reclaim_payload();
        return added;
    }

    // Delete_Points, ikd_Tree.cpp:513-533
    void Delete_Points(PointVector& PointToDel) {
        if (PointToDel.empty()) return;
        std::vector<float> xyz = pack(PointToDel);
        check(ikd_delete_points(h_, xyz.data(), (int64_t)PointToDel.size()), "Delete_Points");
        sync_root();
    }

    // Delete_Point_Boxes, ikd_Tree.cpp:535-556
    int Delete_Point_Boxes(std::vector<BoxPointType>& BoxPoints) {
        if (BoxPoints.empty()) return 0;
        std::vector<float> b = pack_boxes(BoxPoints);
        int32_t n = 0;
        check(ikd_delete_point_boxes(h_, b.data(), (int64_t)BoxPoints.size(), &n), "Delete_Point_Boxes");
        sync_root();
        return n;
    }

    // Add_Point_Boxes, ikd_Tree.cpp:491-511
    void Add_Point_Boxes(std::vector<BoxPointType>& BoxPoints) {
        if (BoxPoints.empty()) return;
        std::vector<float> b = pack_boxes(BoxPoints);
        check(ikd_add_point_boxes(h_, b.data(), (int64_t)BoxPoints.size()), "Add_Point_Boxes");
        sync_root();
    }

    // flatten(root, Storage, storage_type), ikd_Tree.cpp:1325-1352: appends the
    // live points under root (Root_Node: the whole tree) in pre-order.
    void flatten(KD_TREE_NODE* root, PointVector& Storage, delete_point_storage_set storage_type) {
        if (root == nullptr) return;
        if (root != &root_handle_) throw std::runtime_error("flatten: only this tree's Root_Node can be flattened");
        if (storage_type != NOT_RECORD)
            throw std::runtime_error("flatten: only the NOT_RECORD mode is supported");
        append_live(Storage);
    }

    // flatten(Root_Node, Storage, NOT_RECORD) into a cleared vector
    void flatten(PointVector& Storage) {
        Storage.clear();
        append_live(Storage);
    }

    // acquire_removed_points, ikd_Tree.cpp:558-571 (the facade records from
    // construction on, as the reference does)
    void acquire_removed_points(PointVector& removed_points) {
        for (const PointType& p : pending_removed_) removed_points.push_back(p);
        PointVector().swap(pending_removed_);
        take_removed(removed_points);
    }

    ikd_tree_t handle() const { return h_; }

    // Point records the facade holds (for callers watching memory), and the
    // store size below which reclaiming never runs (default 1M records).
    size_t payload_records() const { return payload_.size(); }
    void set_payload_reclaim_floor(size_t records) {
        min_reclaim_ = records;
        reclaim_at_ = std::max(2 * payload_.size(), min_reclaim_) + min_reclaim_;
    }

    PointVector PCL_Storage;          // ikd_Tree.h:246 (public, unused by the library)
    KD_TREE_NODE* Root_Node = nullptr;  // ikd_Tree.h:247 (null while the tree has no nodes)
    int max_queue_size = 0;           // ikd_Tree.h:248

private:
    ikd_tree_t h_ = nullptr;
    KD_TREE_NODE root_handle_;
    PointVector payload_;          // payload id -> point record
    PointVector pending_removed_;  // removed points recorded before the last Build
    size_t min_reclaim_ = size_t(1) << 20;
    size_t reclaim_at_ = size_t(2) << 20;  // payload_ size that triggers reclaim_payload()

    // Keeps only the records of ids the tree can still return, renumbered
    // 0..n-1 on both sides.
    void reclaim_payload() {
        if (payload_.size() < reclaim_at_) return;
        int64_t n = 0;
        check(ikd_list_payload_ids(h_, nullptr, 0, &n), "list_payload_ids");
        std::vector<uint32_t> ids((size_t)n);
        if (n) check(ikd_list_payload_ids(h_, ids.data(), n, &n), "list_payload_ids");
        PointVector kept;
        kept.reserve((size_t)n);
        for (uint32_t id : ids) kept.push_back(payload_[id]);
        check(ikd_relabel_payload_ids(h_, ids.data(), n), "relabel_payload_ids");
        payload_.swap(kept);
        reclaim_at_ = std::max(2 * payload_.size(), min_reclaim_) + min_reclaim_;
    }

    void sync_root() {
        int64_t n = 0;
        check(ikd_size(h_, &n), "size");
        Root_Node = n > 0 ? &root_handle_ : nullptr;
    }

    void append_live(PointVector& Storage) {
        int64_t n = 0;
        check(ikd_flatten(h_, nullptr, nullptr, 0, &n), "flatten");
        if (n == 0) return;
        std::vector<float> xyz(3 * n);
        std::vector<uint32_t> ids(n);
        check(ikd_flatten(h_, xyz.data(), ids.data(), n, &n), "flatten");
        for (int64_t i = 0; i < n; ++i) Storage.push_back(payload_[ids[i]]);
    }

    // moves the library's removed-point record (ids) into out as records
    template <class Vec>
    void take_removed(Vec& out) {
        int64_t n = 0;
        int rc = ikd_acquire_removed_points(h_, nullptr, nullptr, 0, &n);
        if (rc != IKD_OK && rc != IKD_ERR_CAPACITY) check(rc, "acquire_removed_points");
        if (n == 0) return;
        std::vector<float> xyz(3 * n);
        std::vector<uint32_t> ids(n);
        check(ikd_acquire_removed_points(h_, xyz.data(), ids.data(), n, &n), "acquire_removed_points");
        for (int64_t i = 0; i < n; ++i) out.push_back(payload_[ids[i]]);
    }

    static std::vector<float> pack_boxes(const std::vector<BoxPointType>& boxes) {
        std::vector<float> b(6 * boxes.size());
        for (size_t i = 0; i < boxes.size(); ++i)
            for (int a = 0; a < 3; ++a) {
                b[6 * i + a] = boxes[i].vertex_min[a];
                b[6 * i + 3 + a] = boxes[i].vertex_max[a];
            }
        return b;
    }

    void check(int rc, const char* what = "ikd") {
        if (rc != IKD_OK) throw std::runtime_error(std::string(what) + ": " + ikd_last_error(h_));
    }

    static std::vector<float> pack(const PointVector& v) {
        std::vector<float> xyz(3 * v.size());
        for (size_t i = 0; i < v.size(); ++i) {
            xyz[3 * i] = v[i].x;
            xyz[3 * i + 1] = v[i].y;
            xyz[3 * i + 2] = v[i].z;
        }
        return xyz;
    }

    template <class Fill>
    void fetch_range(int64_t total, Fill fill, PointVector& Storage) {
        if (total <= 0) return;
        std::vector<float> xyz(3 * total);
        std::vector<uint32_t> ids(total);
        check(fill(xyz.data(), ids.data(), total), "range search");
        for (int64_t i = 0; i < total; ++i) Storage.push_back(payload_[ids[i]]);
    }
};

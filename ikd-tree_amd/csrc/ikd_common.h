// ikd_common.h — node-pool layout, flag bits and the bit-exact geometry
// helpers shared by the HIP kernels and the host orchestration.
//
// Node pool (device-resident, one per tree)
// -----------------------------------------
// Node 0 is a sentinel (the reference's STATIC_ROOT_NODE, ikd_Tree.cpp:358-363):
// its left child slot holds the root, so the root's AABB lives in the
// sentinel's child-box slot like every other node's AABB.  Node 1 is unused.
// The root, when present, is node 2.  Children are always allocated as a
// sibling pair (2p, 2p+1): left child at the even index, right at the odd one,
// and a node's side is its index parity.
//
// Hot record (64 B, what traversals read):
//   f[0..2]  own point x, y, z
//   f[3]     link word: bits 0..4 flags (below), bits 5..31 child pair index p
//   f[4..9]  AABB of the left child  (lo x,y,z, hi x,y,z)
//   f[10..15] AABB of the right child
// A visit therefore needs exactly one 64-byte record: the node's own point and
// both children's boxes (calc_box_dist of the children, cpp:894-895).
//
// Cold arrays: id[n] (payload id), par[n] (parent), cnt[n] = {TreeSize,
// invalid_point_num, down_del_num, depth}, meta[n] (axis + lazy labels).
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define IKD_HD __host__ __device__ __forceinline__
#else
#define IKD_HD inline
#endif

namespace ikd {

// ---- link word (hot) ----------------------------------------------------
constexpr uint32_t LK_PDEL = 1u << 0;  // own point deleted
constexpr uint32_t LK_HASL = 1u << 1;  // left child present
constexpr uint32_t LK_HASR = 1u << 2;  // right child present
constexpr uint32_t LK_TDL = 1u << 3;   // left child's subtree fully deleted
constexpr uint32_t LK_TDR = 1u << 4;   // right child's subtree fully deleted
constexpr uint32_t LK_FLAGS = 0x1fu;
constexpr int LK_PAIR_SHIFT = 5;

IKD_HD uint32_t lk_pair(uint32_t lk) { return lk >> LK_PAIR_SHIFT; }
IKD_HD uint32_t lk_has(int side) { return side ? LK_HASR : LK_HASL; }
IKD_HD uint32_t lk_td(int side) { return side ? LK_TDR : LK_TDL; }

// ---- meta word (cold) ---------------------------------------------------
constexpr uint32_t M_AXIS = 3u;        // division_axis
constexpr uint32_t M_PDEL = 1u << 2;   // point_deleted
constexpr uint32_t M_TDEL = 1u << 3;   // tree_deleted
constexpr uint32_t M_PDS = 1u << 4;    // point_downsample_deleted
constexpr uint32_t M_TDS = 1u << 5;    // tree_downsample_deleted
constexpr uint32_t M_DIRTY = 1u << 8;  // on a modified path (refit pending)
constexpr uint32_t M_VIOL = 1u << 9;   // violates Criterion_Check after refit
constexpr uint32_t M_TOP = 1u << 10;   // topmost violator: subtree is rebuilt
constexpr uint32_t M_CLAIM = 1u << 11; // claimed by a Delete_Points lane
constexpr uint32_t M_DEFER = 1u << 12; // violator whose rebuild was deferred (listed once)
// Delete_Points leaves the node it deletes without Update (cpp:717-721); the
// refit then keeps its range as of its last Update, when its point was live:
constexpr uint32_t M_FROZEN = 1u << 13;  // range from the children and the own point as if live, no Criterion_Check
constexpr uint32_t M_KEEPBOX = 1u << 14; // not Updated at all in this call: the stored range stays

constexpr int SENTINEL = 0;
constexpr int ROOT = 2;
constexpr int MAX_DEPTH = 64;  // deepest tree the traversal stacks accept
constexpr int WALK_LIMIT = 4096;  // bound on any root-to-leaf or leaf-to-root walk
constexpr int SMALL_MAX = 1024;   // largest subtree built by one workgroup in LDS

// ---- tunables from the reference (ikd_Tree.h:13-18) ----------------------
constexpr double EPSS = 1e-6;
constexpr int MIN_UNBALANCED_SIZE = 10;

struct alignas(16) NodeHot {
    float f[16];
};

// order-preserving map float -> uint32 (ascending floats give ascending keys)
IKD_HD uint32_t float_key(float v) {
    uint32_t u;
    __builtin_memcpy(&u, &v, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

IKD_HD uint32_t fbits(float v) {
    uint32_t u;
    __builtin_memcpy(&u, &v, 4);
    return u;
}
IKD_HD float bitsf(uint32_t u) {
    float v;
    __builtin_memcpy(&v, &u, 4);
    return v;
}

// calc_dist, ikd_Tree.cpp:1373-1378: ((dx*dx + dy*dy) + dz*dz), every
// operation rounded to float (the library is built with -ffp-contract=off).
IKD_HD float calc_dist(float ax, float ay, float az, float bx, float by, float bz) {
    float dx = ax - bx, dy = ay - by, dz = az - bz;
    float d = dx * dx;
    d = d + dy * dy;
    d = d + dz * dz;
    return d;
}

// calc_box_dist, ikd_Tree.cpp:1380-1391: b = {lo x,y,z, hi x,y,z}; the six
// terms are added in x-lo, x-hi, y-lo, y-hi, z-lo, z-hi order.
IKD_HD float calc_box_dist(const float* b, float px, float py, float pz) {
    float s = 0.0f;
    if (px < b[0]) { float t = px - b[0]; s = s + t * t; }
    if (px > b[3]) { float t = px - b[3]; s = s + t * t; }
    if (py < b[1]) { float t = py - b[1]; s = s + t * t; }
    if (py > b[4]) { float t = py - b[4]; s = s + t * t; }
    if (pz < b[2]) { float t = pz - b[2]; s = s + t * t; }
    if (pz > b[5]) { float t = pz - b[5]; s = s + t * t; }
    return s;
}

// same_point, ikd_Tree.cpp:1368-1371: |a-b| < 1e-6 per axis, the float
// difference compared with the double constant.
IKD_HD bool same_point(float ax, float ay, float az, float bx, float by, float bz) {
    float dx = ax - bx, dy = ay - by, dz = az - bz;
    dx = dx < 0 ? -dx : dx;
    dy = dy < 0 ? -dy : dy;
    dz = dz < 0 ? -dz : dz;
    return (double)dx < EPSS && (double)dy < EPSS && (double)dz < EPSS;
}

// Half-open box rules of Delete_by_range / Search_by_range (cpp:653-656,
// 670, 1019-1026).  box = {min x,y,z, max x,y,z}; r = node AABB.
IKD_HD bool box_disjoint(const float* box, const float* r) {
    return box[3] <= r[0] || box[0] > r[3] || box[4] <= r[1] || box[1] > r[4] || box[5] <= r[2] ||
           box[2] > r[5];
}
IKD_HD bool box_covers(const float* box, const float* r) {
    return box[0] <= r[0] && box[3] > r[3] && box[1] <= r[1] && box[4] > r[4] && box[2] <= r[2] &&
           box[5] > r[5];
}
IKD_HD bool box_contains(const float* box, float x, float y, float z) {
    return box[0] <= x && box[3] > x && box[1] <= y && box[4] > y && box[2] <= z && box[5] > z;
}

// Update's radius_sq, cpp:1309-1312: the halvings are double products
// narrowed to float; the sum is float.
IKD_HD float box_radius_sq(const float* r) {
    float xl = (float)((double)(r[3] - r[0]) * 0.5);
    float yl = (float)((double)(r[4] - r[1]) * 0.5);
    float zl = (float)((double)(r[5] - r[2]) * 0.5);
    float s = xl * xl;
    s = s + yl * yl;
    s = s + zl * zl;
    return s;
}

}  // namespace ikd

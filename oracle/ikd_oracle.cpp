// ikd_oracle.cpp — CPU restatement of hku-mars/ikd-Tree's KD_TREE<PointType>
// (ikd-Tree/ikd_Tree.h, ikd-Tree/ikd_Tree.cpp), written from the reference's
// documented behaviour for use as a CHECKER.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product path links or calls this
// file: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load the shared library built from it.
//
// What it restates (one function per reference routine, each citing the
// reference lines it follows):
//   - the pointer tree with lazy delete labels (h:64-86), Push_Down / Update,
//   - BuildTree's longest-axis median split via std::nth_element,
//   - the bounded heap and the nearest-first recursive kNN Search,
//   - Search_by_range / Search_by_radius / flatten,
//   - Add_Points with voxel downsample, Add_by_point, Delete_by_point,
//     Delete_by_range, Add_by_range (Add_Point_Boxes), Criterion_Check and
//     Rebuild.
// Deliberate differences (none of them changes a live set or a query result):
//   - every Rebuild runs synchronously, also for subtrees of >= 1500 nodes
//     that the reference hands to its background thread (cpp:627-633); the
//     thread, its operation log and its locks are not restated;
//   - points carry a uint32 payload id so tests can compare identities; the
//     id never takes part in a decision.
//
// Floating point follows the reference's C++ typing exactly: float where the
// reference computes in float, double where it promotes (max_dist, the
// voxel centre, the *0.5 halvings).  Build with -ffp-contract=off.
//
// Parity pinning (see tests/test_oracle_*.py): kNN against numpy brute force
// (bit-exact squared distances), box search and box delete against the
// half-open box filter, radius search against brute force outside a 4-ulp band,
// Add_Points against an independent flat-list model of the per-point
// downsample rule, Build structure against a recursive median definition on
// tie-free inputs, and tree invariants (counts, tight live AABBs, criterion).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr double kEpss = 1e-6;                 // EPSS, h:13
constexpr int kMinimalUnbalancedTreeSize = 10; // h:14

struct Pt {
    float x, y, z;
    uint32_t id;
};

struct Box {
    float vmin[3];
    float vmax[3];
};

// KD_TREE_NODE, h:64-86 (mutex and working flag dropped: single-threaded).
struct Node {
    Pt point{};
    uint8_t axis = 0;
    int tree_size = 1;
    int invalid = 0;
    int down_del = 0;
    bool point_deleted = false;
    bool tree_deleted = false;
    bool point_ds_deleted = false;
    bool tree_ds_deleted = false;
    bool push_left = false;
    bool push_right = false;
    float radius_sq = 0.0f;
    float rx[2] = {0, 0}, ry[2] = {0, 0}, rz[2] = {0, 0};
    Node* left = nullptr;
    Node* right = nullptr;
    Node* father = nullptr;
    float alpha_del = 0.0f;
    float alpha_bal = 0.0f;
};

// same_point, cpp:1368-1371: float differences, fabs on float, compared with
// the double constant EPSS.
inline bool same_point(const Pt& a, const Pt& b) {
    return std::fabs(a.x - b.x) < kEpss && std::fabs(a.y - b.y) < kEpss &&
           std::fabs(a.z - b.z) < kEpss;
}

// calc_dist, cpp:1373-1378: ((dx*dx + dy*dy) + dz*dz) in float.
inline float calc_dist(const Pt& a, const Pt& b) {
    float dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
    float d = dx * dx;
    d = d + dy * dy;
    d = d + dz * dz;
    return d;
}

// calc_box_dist, cpp:1380-1391: terms added in x-lo, x-hi, y-lo, y-hi, z-lo,
// z-hi order; INFINITY for a null node.
inline float calc_box_dist(const Node* n, const Pt& p) {
    if (n == nullptr) return INFINITY;
    float s = 0.0f;
    if (p.x < n->rx[0]) { float t = p.x - n->rx[0]; s += t * t; }
    if (p.x > n->rx[1]) { float t = p.x - n->rx[1]; s += t * t; }
    if (p.y < n->ry[0]) { float t = p.y - n->ry[0]; s += t * t; }
    if (p.y > n->ry[1]) { float t = p.y - n->ry[1]; s += t * t; }
    if (p.z < n->rz[0]) { float t = p.z - n->rz[0]; s += t * t; }
    if (p.z > n->rz[1]) { float t = p.z - n->rz[1]; s += t * t; }
    return s;
}

inline float coord(const Pt& p, int axis) { return axis == 0 ? p.x : (axis == 1 ? p.y : p.z); }

// PointType_CMP, h:95-106: distances closer than 1e-10 compare by x.
struct HeapEntry {
    Pt point;
    float dist = INFINITY;
};
inline bool heap_less(const HeapEntry& a, const HeapEntry& b) {
    float diff = std::fabs(a.dist - b.dist);
    if ((double)diff < 1e-10) return a.point.x < b.point.x;
    return a.dist < b.dist;
}

// MANUAL_HEAP, h:108-172: array max-heap of fixed capacity; push is ignored
// when full; pop moves the last entry to the root and sifts it down.
class BoundedHeap {
public:
    explicit BoundedHeap(int cap) : cap_(cap), h_(cap > 0 ? cap : 1) {}
    int size() const { return n_; }
    const HeapEntry& top() const { return h_[0]; }
    void push(const HeapEntry& e) {
        if (n_ >= cap_) return;
        int i = n_++;
        HeapEntry t = e;
        while (i > 0) {
            int up = (i - 1) / 2;
            if (heap_less(h_[up], t)) { h_[i] = h_[up]; i = up; } else break;
        }
        h_[i] = t;
    }
    void pop() {
        if (n_ == 0) return;
        h_[0] = h_[n_ - 1];
        --n_;
        int i = 0;
        HeapEntry t = h_[0];
        int c = 1;
        while (c < n_) {
            if (c + 1 < n_ && heap_less(h_[c], h_[c + 1])) ++c;
            if (heap_less(t, h_[c])) { h_[i] = h_[c]; i = c; c = 2 * i + 1; } else break;
        }
        h_[i] = t;
    }

private:
    int cap_;
    int n_ = 0;
    std::vector<HeapEntry> h_;
};

struct Counters {
    long long visits = 0;     // nodes that passed Search's entry checks
    long long child_boxes = 0; // calc_box_dist calls on non-null children
};

class Tree {
public:
    float delete_param = 0.5f;
    float balance_param = 0.6f;
    float downsample_size = 0.2f;
    Node* root = nullptr;
    Node sentinel;  // STATIC_ROOT_NODE, cpp:358-363
    // every pending label has been pushed down since the last change (a
    // multi-threaded kNN may then search without writing the tree)
    bool labels_pushed = false;

    ~Tree() { destroy(root); }

    // ----- lazy labels --------------------------------------------------
    // Push_Down, cpp:1109-1181.
    static void push_one(Node* parent, Node* c) {
        c->tree_ds_deleted |= parent->tree_ds_deleted;
        c->point_ds_deleted |= parent->tree_ds_deleted;
        c->tree_deleted = parent->tree_deleted || c->tree_ds_deleted;
        c->point_deleted = c->tree_deleted || c->point_ds_deleted;
        if (parent->tree_ds_deleted) c->down_del = c->tree_size;
        if (parent->tree_deleted) c->invalid = c->tree_size;
        else c->invalid = c->down_del;
        c->push_left = true;
        c->push_right = true;
    }
    static void push_down(Node* n) {
        if (n == nullptr) return;
        if (n->push_left && n->left != nullptr) { push_one(n, n->left); n->push_left = false; }
        if (n->push_right && n->right != nullptr) { push_one(n, n->right); n->push_right = false; }
    }

    // Update, cpp:1183-1323 (the three child cases share one formula).
    void update(Node* n) {
        Node* kids[2] = {n->left, n->right};
        int size = 1, inv = n->point_deleted ? 1 : 0, dd = n->point_ds_deleted ? 1 : 0;
        bool all_td = true, all_tds = true, any_child_td = false;
        for (Node* c : kids) {
            if (!c) continue;
            size += c->tree_size;
            inv += c->invalid;
            dd += c->down_del;
            all_td = all_td && c->tree_deleted;
            all_tds = all_tds && c->tree_ds_deleted;
            any_child_td = any_child_td || c->tree_deleted;
        }
        n->tree_size = size;
        n->invalid = inv;
        n->down_del = dd;
        n->tree_ds_deleted = all_tds && n->point_ds_deleted;
        n->tree_deleted = all_td && n->point_deleted;
        float lo[3] = {INFINITY, INFINITY, INFINITY};
        float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        auto take_box = [&](const Node* c) {
            lo[0] = std::min(lo[0], c->rx[0]); hi[0] = std::max(hi[0], c->rx[1]);
            lo[1] = std::min(lo[1], c->ry[0]); hi[1] = std::max(hi[1], c->ry[1]);
            lo[2] = std::min(lo[2], c->rz[0]); hi[2] = std::max(hi[2], c->rz[1]);
        };
        auto take_point = [&](const Pt& p) {
            lo[0] = std::min(lo[0], p.x); hi[0] = std::max(hi[0], p.x);
            lo[1] = std::min(lo[1], p.y); hi[1] = std::max(hi[1], p.y);
            lo[2] = std::min(lo[2], p.z); hi[2] = std::max(hi[2], p.z);
        };
        bool everything = n->tree_deleted || (!any_child_td && !n->point_deleted);
        for (Node* c : kids)
            if (c && (everything || !c->tree_deleted)) take_box(c);
        if (everything || !n->point_deleted) take_point(n->point);
        n->rx[0] = lo[0]; n->rx[1] = hi[0];
        n->ry[0] = lo[1]; n->ry[1] = hi[1];
        n->rz[0] = lo[2]; n->rz[1] = hi[2];
        // cpp:1309-1312: the halvings are double products narrowed to float.
        float xl = (n->rx[1] - n->rx[0]) * 0.5;
        float yl = (n->ry[1] - n->ry[0]) * 0.5;
        float zl = (n->rz[1] - n->rz[0]) * 0.5;
        float r = xl * xl;
        r = r + yl * yl;
        r = r + zl * zl;
        n->radius_sq = r;
        if (n->left) n->left->father = n;
        if (n->right) n->right->father = n;
        // cpp:1315-1321: root statistics recorded "for paper data".
        if (n == root && n->tree_size > 3) {
            Node* son = n->left ? n->left : n->right;
            float bal = float(son->tree_size) / (n->tree_size - 1);
            n->alpha_del = float(n->invalid) / n->tree_size;
            n->alpha_bal = (bal >= 0.5 - kEpss) ? bal : 1 - bal;
        }
    }

    // ----- build --------------------------------------------------------
    // BuildTree, cpp:573-622: longest extent (strict >, ties to the lower
    // axis), nth_element at mid=(l+r)>>1, node = Storage[mid].
    void build_tree(Node** slot, int l, int r, std::vector<Pt>& st) {
        if (l > r) return;
        Node* n = new Node();
        n->tree_size = 0;
        *slot = n;
        int mid = (l + r) >> 1;
        float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int i = l; i <= r; ++i) {
            mn[0] = std::min(mn[0], st[i].x); mx[0] = std::max(mx[0], st[i].x);
            mn[1] = std::min(mn[1], st[i].y); mx[1] = std::max(mx[1], st[i].y);
            mn[2] = std::min(mn[2], st[i].z); mx[2] = std::max(mx[2], st[i].z);
        }
        float ext[3] = {mx[0] - mn[0], mx[1] - mn[1], mx[2] - mn[2]};
        int axis = 0;
        for (int a = 1; a < 3; ++a)
            if (ext[a] > ext[axis]) axis = a;
        n->axis = (uint8_t)axis;
        auto first = st.begin() + l, nth = st.begin() + mid, last = st.begin() + r + 1;
        if (axis == 0) std::nth_element(first, nth, last, [](const Pt& a, const Pt& b) { return a.x < b.x; });
        else if (axis == 1) std::nth_element(first, nth, last, [](const Pt& a, const Pt& b) { return a.y < b.y; });
        else std::nth_element(first, nth, last, [](const Pt& a, const Pt& b) { return a.z < b.z; });
        n->point = st[mid];
        Node *ls = nullptr, *rs = nullptr;
        build_tree(&ls, l, mid - 1, st);
        build_tree(&rs, mid + 1, r, st);
        n->left = ls;
        n->right = rs;
        update(n);
    }

    // Build, cpp:352-364.
    void build(std::vector<Pt> cloud) {
        if (root != nullptr) destroy(root);
        root = nullptr;
        if (cloud.empty()) return;
        sentinel = Node();
        sentinel.tree_size = 0;
        build_tree(&sentinel.left, 0, (int)cloud.size() - 1, cloud);
        update(&sentinel);
        sentinel.tree_size = 0;
        root = sentinel.left;
    }

    static void destroy(Node* n) {
        if (!n) return;
        destroy(n->left);
        destroy(n->right);
        delete n;
    }

    // ----- queries ------------------------------------------------------
    // Search, cpp:868-1013.  The double comparisons with max_dist*max_dist
    // follow cpp:871-873 and cpp:887.
    void search(Node* n, int k, const Pt& q, BoundedHeap& h, double max_dist, Counters* ctr) {
        if (n == nullptr || n->tree_deleted) return;
        double cur = calc_box_dist(n, q);
        double mds = max_dist * max_dist;
        if (cur > mds) return;
        if (n->push_left || n->push_right) push_down(n);
        if (ctr) ctr->visits++;
        if (!n->point_deleted) {
            float d = calc_dist(q, n->point);
            if (d <= mds && (h.size() < k || d < h.top().dist)) {
                if (h.size() >= k) h.pop();
                h.push(HeapEntry{n->point, d});
            }
        }
        if (ctr) ctr->child_boxes += (n->left != nullptr) + (n->right != nullptr);
        float dl = calc_box_dist(n->left, q);
        float dr = calc_box_dist(n->right, q);
        if (h.size() < k || (dl < h.top().dist && dr < h.top().dist)) {
            if (dl <= dr) {
                search(n->left, k, q, h, max_dist, ctr);
                if (h.size() < k || dr < h.top().dist) search(n->right, k, q, h, max_dist, ctr);
            } else {
                search(n->right, k, q, h, max_dist, ctr);
                if (h.size() < k || dl < h.top().dist) search(n->left, k, q, h, max_dist, ctr);
            }
        } else {
            if (dl < h.top().dist) search(n->left, k, q, h, max_dist, ctr);
            if (dr < h.top().dist) search(n->right, k, q, h, max_dist, ctr);
        }
    }

    // Nearest_Search, cpp:366-397: heap of capacity 2k; results popped from
    // the top and inserted at the front, i.e. ascending heap order.
    int nearest(const Pt& q, int k, double max_dist, Pt* out, float* out_d, Counters* ctr) {
        BoundedHeap h(2 * k);
        search(root, k, q, h, max_dist, ctr);
        int found = std::min(k, h.size());
        for (int i = found - 1; i >= 0; --i) {
            out[i] = h.top().point;
            out_d[i] = h.top().dist;
            h.pop();
        }
        return found;
    }

    static bool disjoint(const Box& b, const Node* n) {
        return b.vmax[0] <= n->rx[0] || b.vmin[0] > n->rx[1] || b.vmax[1] <= n->ry[0] ||
               b.vmin[1] > n->ry[1] || b.vmax[2] <= n->rz[0] || b.vmin[2] > n->rz[1];
    }
    static bool covers(const Box& b, const Node* n) {
        return b.vmin[0] <= n->rx[0] && b.vmax[0] > n->rx[1] && b.vmin[1] <= n->ry[0] &&
               b.vmax[1] > n->ry[1] && b.vmin[2] <= n->rz[0] && b.vmax[2] > n->rz[1];
    }
    static bool inside(const Box& b, const Pt& p) {
        return b.vmin[0] <= p.x && b.vmax[0] > p.x && b.vmin[1] <= p.y && b.vmax[1] > p.y &&
               b.vmin[2] <= p.z && b.vmax[2] > p.z;
    }

    // Points_deleted (cpp:558-571): points removed from the tree by rebuilds.
    std::vector<Pt> removed;

    // flatten, cpp:1325-1352: live points in pre-order; with rec (the
    // DELETE_POINTS_REC mode, cpp:1338-1341) every deleted point that was not
    // removed by the downsample is recorded after the node's children.
    void flatten(Node* n, std::vector<Pt>& out, std::vector<Pt>* rec = nullptr) {
        if (!n) return;
        push_down(n);
        if (!n->point_deleted) out.push_back(n->point);
        flatten(n->left, out, rec);
        flatten(n->right, out, rec);
        if (rec && n->point_deleted && !n->point_ds_deleted) rec->push_back(n->point);
    }

    // Search_by_range, cpp:1015-1044.
    void search_by_range(Node* n, const Box& b, std::vector<Pt>& out) {
        if (!n) return;
        push_down(n);
        if (disjoint(b, n)) return;
        if (covers(b, n)) { flatten(n, out); return; }
        if (inside(b, n->point) && !n->point_deleted) out.push_back(n->point);
        search_by_range(n->left, b, out);
        search_by_range(n->right, b, out);
    }

    // Search_by_radius, cpp:1046-1087: float arithmetic with float sqrt.
    void search_by_radius(Node* n, const Pt& q, float radius, std::vector<Pt>& out) {
        if (!n) return;
        push_down(n);
        Pt c;
        c.x = (n->rx[0] + n->rx[1]) * 0.5;
        c.y = (n->ry[0] + n->ry[1]) * 0.5;
        c.z = (n->rz[0] + n->rz[1]) * 0.5;
        c.id = 0;
        float dist = std::sqrt(calc_dist(c, q));
        float rs = std::sqrt(n->radius_sq);
        if (dist > radius + rs) return;
        if (dist <= radius - rs) { flatten(n, out); return; }
        if (!n->point_deleted && calc_dist(n->point, q) <= radius * radius) out.push_back(n->point);
        search_by_radius(n->left, q, radius, out);
        search_by_radius(n->right, q, radius, out);
    }

    // ----- updates ------------------------------------------------------
    // Criterion_Check, cpp:1089-1107.
    bool criterion(const Node* n) const {
        if (n->tree_size <= kMinimalUnbalancedTreeSize) return false;
        const Node* son = n->left ? n->left : n->right;
        float del_eval = float(n->invalid) / n->tree_size;
        float bal_eval = float(son->tree_size) / (n->tree_size - 1);
        if (del_eval > delete_param) return true;
        if (bal_eval > balance_param || bal_eval < 1 - balance_param) return true;
        return false;
    }

    // Rebuild, cpp:624-645, always on the caller's thread.
    void rebuild(Node** slot) {
        Node* father = (*slot)->father;
        std::vector<Pt> st;
        flatten(*slot, st, &removed);
        destroy(*slot);
        *slot = nullptr;
        if (!st.empty()) build_tree(slot, 0, (int)st.size() - 1, st);
        if (*slot) (*slot)->father = father;
        if (slot == &root) sentinel.left = root;
    }

    void finish(Node** slot, bool allow_rebuild) {
        update(*slot);
        if (allow_rebuild && criterion(*slot)) rebuild(slot);
    }

    // Add_by_point, cpp:817-866.
    void add_by_point(Node** slot, const Pt& p, bool allow_rebuild, int father_axis) {
        if (*slot == nullptr) {
            Node* n = new Node();
            n->point = p;
            n->axis = (uint8_t)((father_axis + 1) % 3);
            *slot = n;
            update(n);
            return;
        }
        Node* n = *slot;
        push_down(n);
        if (coord(p, n->axis) < coord(n->point, n->axis))
            add_by_point(&n->left, p, allow_rebuild, n->axis);
        else
            add_by_point(&n->right, p, allow_rebuild, n->axis);
        finish(slot, allow_rebuild);
    }

    // Delete_by_point, cpp:712-760.
    void delete_by_point(Node** slot, const Pt& p, bool allow_rebuild) {
        Node* n = *slot;
        if (n == nullptr || n->tree_deleted) return;
        push_down(n);
        if (same_point(n->point, p) && !n->point_deleted) {
            n->point_deleted = true;
            n->invalid += 1;
            if (n->invalid == n->tree_size) n->tree_deleted = true;
            return;
        }
        if (coord(p, n->axis) < coord(n->point, n->axis))
            delete_by_point(&n->left, p, allow_rebuild);
        else
            delete_by_point(&n->right, p, allow_rebuild);
        finish(slot, allow_rebuild);
    }

    // Add_by_range, cpp:762-815: enters deleted subtrees too (no tree_deleted
    // check).  A subtree whose range the box covers gets back every point the
    // downsample did not delete, through the labels Push_Down hands on
    // (invalid := down_del, deletion := downsample deletion); a partly
    // covered node gets back its own point if it lies in the box.  The
    // decisions use the nodes' stored ranges, which leave out deleted
    // subtrees, so a covered node may bring back points outside the box and
    // points inside it stay deleted where an ancestor's range misses them.
    void add_by_range(Node** slot, const Box& b, bool allow_rebuild) {
        Node* n = *slot;
        if (n == nullptr) return;
        push_down(n);
        if (disjoint(b, n)) return;
        if (covers(b, n)) {
            n->tree_deleted = n->tree_ds_deleted;
            n->point_deleted = n->point_ds_deleted;
            n->push_left = true;
            n->push_right = true;
            n->invalid = n->down_del;
            return;
        }
        if (inside(b, n->point)) n->point_deleted = n->point_ds_deleted;
        add_by_range(&n->left, b, allow_rebuild);
        add_by_range(&n->right, b, allow_rebuild);
        finish(slot, allow_rebuild);
    }

    // Delete_by_range, cpp:647-710.
    int delete_by_range(Node** slot, const Box& b, bool allow_rebuild, bool is_ds) {
        Node* n = *slot;
        if (n == nullptr || n->tree_deleted) return 0;
        push_down(n);
        if (disjoint(b, n)) return 0;
        if (covers(b, n)) {
            n->tree_deleted = true;
            n->point_deleted = true;
            n->push_left = true;
            n->push_right = true;
            int cnt = n->tree_size - n->invalid;
            n->invalid = n->tree_size;
            if (is_ds) {
                n->tree_ds_deleted = true;
                n->point_ds_deleted = true;
                n->down_del = n->tree_size;
            }
            return cnt;
        }
        int cnt = 0;
        if (!n->point_deleted && inside(b, n->point)) {
            n->point_deleted = true;
            cnt += 1;
            if (is_ds) n->point_ds_deleted = true;
        }
        cnt += delete_by_range(&n->left, b, allow_rebuild, is_ds);
        cnt += delete_by_range(&n->right, b, allow_rebuild, is_ds);
        finish(slot, allow_rebuild);
        return cnt;
    }

    // Add_Points, cpp:413-489.
    int add_points(const std::vector<Pt>& pts, bool downsample_on) {
        int added = 0;
        for (const Pt& p : pts) {
            if (downsample_on) {
                Box b;
                b.vmin[0] = std::floor(p.x / downsample_size) * downsample_size;
                b.vmax[0] = b.vmin[0] + downsample_size;
                b.vmin[1] = std::floor(p.y / downsample_size) * downsample_size;
                b.vmax[1] = b.vmin[1] + downsample_size;
                b.vmin[2] = std::floor(p.z / downsample_size) * downsample_size;
                b.vmax[2] = b.vmin[2] + downsample_size;
                Pt mid;
                mid.x = b.vmin[0] + (b.vmax[0] - b.vmin[0]) / 2.0;
                mid.y = b.vmin[1] + (b.vmax[1] - b.vmin[1]) / 2.0;
                mid.z = b.vmin[2] + (b.vmax[2] - b.vmin[2]) / 2.0;
                mid.id = 0;
                std::vector<Pt> found;
                search_by_range(root, b, found);
                float best_d = calc_dist(p, mid);
                Pt best = p;
                for (const Pt& s : found) {
                    float d = calc_dist(s, mid);
                    if (d < best_d) { best_d = d; best = s; }
                }
                if (found.size() > 1 || same_point(p, best)) {
                    if (!found.empty()) delete_by_range(&root, b, true, true);
                    add_by_point(&root, best, true, root ? root->axis : 0);
                    ++added;
                }
            } else {
                add_by_point(&root, p, true, root ? root->axis : 0);
            }
            if (root) { root->father = &sentinel; sentinel.left = root; }
        }
        return added;
    }

    void delete_points(const std::vector<Pt>& pts) {
        for (const Pt& p : pts) delete_by_point(&root, p, true);
        if (root) { root->father = &sentinel; sentinel.left = root; }
    }

    int delete_boxes(const std::vector<Box>& boxes) {
        int cnt = 0;
        for (const Box& b : boxes) cnt += delete_by_range(&root, b, true, false);
        if (root) { root->father = &sentinel; sentinel.left = root; }
        return cnt;
    }

    // Add_Point_Boxes, cpp:491-511.
    void add_boxes(const std::vector<Box>& boxes) {
        for (const Box& b : boxes) add_by_range(&root, b, true);
        if (root) { root->father = &sentinel; sentinel.left = root; }
    }

    // size / validnum / tree_range, cpp:78-145 (no rebuild in flight here).
    int size() const { return root ? root->tree_size : 0; }
    int validnum() const { return root ? root->tree_size - root->invalid : 0; }
};

struct DumpRow {
    float x, y, z;
    uint32_t id;
    int32_t axis, tree_size, invalid, flags;  // flags: pd | td<<1 | pds<<2 | tds<<3
    float range[6];
    int32_t left, right;  // pre-order row index, -1 for null
};

int dump_rec(Node* n, std::vector<DumpRow>& rows) {
    if (!n) return -1;
    Tree::push_down(n);
    int me = (int)rows.size();
    rows.push_back(DumpRow{});
    DumpRow r{};
    r.x = n->point.x; r.y = n->point.y; r.z = n->point.z; r.id = n->point.id;
    r.axis = n->axis; r.tree_size = n->tree_size; r.invalid = n->invalid;
    r.flags = (n->point_deleted ? 1 : 0) | (n->tree_deleted ? 2 : 0) | (n->point_ds_deleted ? 4 : 0) |
              (n->tree_ds_deleted ? 8 : 0);
    r.range[0] = n->rx[0]; r.range[1] = n->ry[0]; r.range[2] = n->rz[0];
    r.range[3] = n->rx[1]; r.range[4] = n->ry[1]; r.range[5] = n->rz[1];
    r.left = dump_rec(n->left, rows);
    r.right = dump_rec(n->right, rows);
    rows[me] = r;
    return me;
}

std::vector<Pt> make_points(const float* xyz, const uint32_t* ids, int64_t n) {
    std::vector<Pt> v((size_t)n);
    for (int64_t i = 0; i < n; ++i)
        v[i] = Pt{xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], ids ? ids[i] : (uint32_t)i};
    return v;
}

}  // namespace

// ---------------------------------------------------------------------------
// C API for the test harness (ctypes).  Points are packed float xyz triples.
// ---------------------------------------------------------------------------
extern "C" {

void* ora_create(float delete_param, float balance_param, float box_length) {
    Tree* t = new Tree();
    t->delete_param = delete_param;
    t->balance_param = balance_param;
    t->downsample_size = box_length;
    return t;
}

void ora_destroy(void* h) { delete (Tree*)h; }

void ora_set_params(void* h, float delete_param, float balance_param, float box_length) {
    Tree* t = (Tree*)h;
    t->delete_param = delete_param;
    t->balance_param = balance_param;
    t->downsample_size = box_length;
}

void ora_build(void* h, const float* xyz, const uint32_t* ids, int64_t n) {
    ((Tree*)h)->labels_pushed = false;
    ((Tree*)h)->build(make_points(xyz, ids, n));
}

// Batched Nearest_Search.  out_* hold k slots per query; out_counts the
// number found.  visits/child_boxes (nullable) receive per-query counters.
// threads > 1 runs queries in parallel: Search is read-only once every lazy
// label has been pushed down, which ora_push_all() does first.
int ora_knn(void* h, const float* q, int64_t nq, int k, double max_dist, float* out_xyz, float* out_d2,
            uint32_t* out_ids, int32_t* out_counts, int64_t* visits, int64_t* child_boxes, int threads) {
    Tree* t = (Tree*)h;
    if (k <= 0) return -1;
    if (threads > 1 && !t->labels_pushed) {
        std::vector<Pt> tmp;
        t->flatten(t->root, tmp);  // pushes every pending label down once per change
        t->labels_pushed = true;
    }
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads > 1 ? threads : 1)
    for (int64_t i = 0; i < nq; ++i) {
        Pt qp{q[3 * i], q[3 * i + 1], q[3 * i + 2], 0};
        std::vector<Pt> pts((size_t)k);
        std::vector<float> ds((size_t)k);
        Counters c;
        int f = t->nearest(qp, k, max_dist, pts.data(), ds.data(), (visits || child_boxes) ? &c : nullptr);
        out_counts[i] = f;
        for (int j = 0; j < k; ++j) {
            int64_t o = i * k + j;
            if (j < f) {
                out_xyz[3 * o] = pts[j].x; out_xyz[3 * o + 1] = pts[j].y; out_xyz[3 * o + 2] = pts[j].z;
                out_d2[o] = ds[j];
                out_ids[o] = pts[j].id;
            } else {
                out_xyz[3 * o] = out_xyz[3 * o + 1] = out_xyz[3 * o + 2] = 0.0f;
                out_d2[o] = INFINITY;
                out_ids[o] = 0xffffffffu;
            }
        }
        if (visits) visits[i] = c.visits;
        if (child_boxes) child_boxes[i] = c.child_boxes;
    }
    return 0;
}

// Box_Search, cpp:399-404.  Returns the number of points; writes at most cap.
int64_t ora_box_search(void* h, const float* box6, float* out_xyz, uint32_t* out_ids, int64_t cap) {
    Tree* t = (Tree*)h;
    Box b;
    for (int a = 0; a < 3; ++a) { b.vmin[a] = box6[a]; b.vmax[a] = box6[3 + a]; }
    std::vector<Pt> out;
    t->search_by_range(t->root, b, out);
    for (int64_t i = 0; i < (int64_t)out.size() && i < cap; ++i) {
        out_xyz[3 * i] = out[i].x; out_xyz[3 * i + 1] = out[i].y; out_xyz[3 * i + 2] = out[i].z;
        if (out_ids) out_ids[i] = out[i].id;
    }
    return (int64_t)out.size();
}

// Radius_Search, cpp:406-411.
int64_t ora_radius_search(void* h, const float* q3, float radius, float* out_xyz, uint32_t* out_ids,
                          int64_t cap) {
    Tree* t = (Tree*)h;
    Pt q{q3[0], q3[1], q3[2], 0};
    std::vector<Pt> out;
    t->search_by_radius(t->root, q, radius, out);
    for (int64_t i = 0; i < (int64_t)out.size() && i < cap; ++i) {
        out_xyz[3 * i] = out[i].x; out_xyz[3 * i + 1] = out[i].y; out_xyz[3 * i + 2] = out[i].z;
        if (out_ids) out_ids[i] = out[i].id;
    }
    return (int64_t)out.size();
}

int ora_add_points(void* h, const float* xyz, const uint32_t* ids, int64_t n, int downsample_on) {
    ((Tree*)h)->labels_pushed = false;
    return ((Tree*)h)->add_points(make_points(xyz, ids, n), downsample_on != 0);
}

void ora_delete_points(void* h, const float* xyz, int64_t n) {
    ((Tree*)h)->labels_pushed = false;
    ((Tree*)h)->delete_points(make_points(xyz, nullptr, n));
}

int ora_delete_point_boxes(void* h, const float* boxes6, int64_t nb) {
    ((Tree*)h)->labels_pushed = false;
    std::vector<Box> v((size_t)nb);
    for (int64_t i = 0; i < nb; ++i)
        for (int a = 0; a < 3; ++a) { v[i].vmin[a] = boxes6[6 * i + a]; v[i].vmax[a] = boxes6[6 * i + 3 + a]; }
    return ((Tree*)h)->delete_boxes(v);
}

// Add_Point_Boxes, cpp:491-511.
void ora_add_point_boxes(void* h, const float* boxes6, int64_t nb) {
    ((Tree*)h)->labels_pushed = false;
    std::vector<Box> v((size_t)nb);
    for (int64_t i = 0; i < nb; ++i)
        for (int a = 0; a < 3; ++a) { v[i].vmin[a] = boxes6[6 * i + a]; v[i].vmax[a] = boxes6[6 * i + 3 + a]; }
    ((Tree*)h)->add_boxes(v);
}

int64_t ora_size(void* h) { return ((Tree*)h)->size(); }
int64_t ora_validnum(void* h) { return ((Tree*)h)->validnum(); }

// tree_range, cpp:98-126: root AABB as (min xyz, max xyz); zeros when empty.
void ora_tree_range(void* h, float* out6) {
    Tree* t = (Tree*)h;
    if (!t->root) { std::memset(out6, 0, 6 * sizeof(float)); return; }
    out6[0] = t->root->rx[0]; out6[1] = t->root->ry[0]; out6[2] = t->root->rz[0];
    out6[3] = t->root->rx[1]; out6[4] = t->root->ry[1]; out6[5] = t->root->rz[1];
}

void ora_root_alpha(void* h, float* alpha_bal, float* alpha_del) {
    Tree* t = (Tree*)h;
    *alpha_bal = t->root ? t->root->alpha_bal : 0.0f;
    *alpha_del = t->root ? t->root->alpha_del : 0.0f;
}

// flatten(Root_Node, ..., NOT_RECORD): live points in pre-order.
int64_t ora_flatten(void* h, float* out_xyz, uint32_t* out_ids, int64_t cap) {
    Tree* t = (Tree*)h;
    std::vector<Pt> out;
    t->flatten(t->root, out);
    for (int64_t i = 0; i < (int64_t)out.size() && i < cap; ++i) {
        out_xyz[3 * i] = out[i].x; out_xyz[3 * i + 1] = out[i].y; out_xyz[3 * i + 2] = out[i].z;
        if (out_ids) out_ids[i] = out[i].id;
    }
    return (int64_t)out.size();
}

// Pre-order dump of every node (after pushing all lazy labels down) for the
// structure and invariant tests.  Returns the node count; writes at most cap.
int64_t ora_dump(void* h, DumpRow* rows, int64_t cap) {
    Tree* t = (Tree*)h;
    std::vector<DumpRow> v;
    dump_rec(t->root, v);
    for (int64_t i = 0; i < (int64_t)v.size() && i < cap; ++i) rows[i] = v[i];
    return (int64_t)v.size();
}

int ora_dump_row_bytes() { return (int)sizeof(DumpRow); }

// acquire_removed_points, cpp:558-571: the points rebuilds removed since the
// last call (deleted, not by the downsample), then the record is cleared.
// Returns the count; copies only when cap holds them all (and only then clears).
int64_t ora_acquire_removed(void* h, float* out_xyz, uint32_t* out_ids, int64_t cap) {
    Tree* t = (Tree*)h;
    int64_t n = (int64_t)t->removed.size();
    if (n > cap) return n;
    for (int64_t i = 0; i < n; ++i) {
        out_xyz[3 * i] = t->removed[i].x;
        out_xyz[3 * i + 1] = t->removed[i].y;
        out_xyz[3 * i + 2] = t->removed[i].z;
        if (out_ids) out_ids[i] = t->removed[i].id;
    }
    t->removed.clear();
    return n;
}

}  // extern "C"

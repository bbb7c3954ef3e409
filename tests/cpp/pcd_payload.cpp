// The Search demo's flow without PCL (examples/ikd_Tree_Search_demo.cpp:64-76):
// load a .pcd map into a PCL-layout point type, Build, then Nearest_Search and
// Box_Search.  Every result must carry the full record it was loaded with
// (intensity, normals, curvature), the kNN distances must equal brute force,
// and the box must return exactly the points inside it.  argv[1]: the PCD file.
// Prints "ok".
#include <algorithm>
#include <cstdio>
#include <map>
#include <tuple>

#include "ikd_Tree_amd.hpp"
#include "ikd_pcd.hpp"
#include "ikd_pcl_types.hpp"

using P = ikd_pcl::PointXYZINormal;

static bool same(const P& a, const P& b) {
    return a.x == b.x && a.y == b.y && a.z == b.z && a.intensity == b.intensity && a.normal_x == b.normal_x &&
           a.normal_y == b.normal_y && a.normal_z == b.normal_z && a.curvature == b.curvature;
}

int main(int argc, char** argv) {
    KD_TREE<P>::PointVector cloud;
    if (argc < 2 || ikd_pcd::loadPCDFile(argv[1], cloud) == -1) return 1;
    std::map<std::tuple<float, float, float>, P> by_xyz;
    for (const P& p : cloud) by_xyz[{p.x, p.y, p.z}] = p;
    KD_TREE<P>::Ptr kdtree_ptr(new KD_TREE<P>(0.3, 0.6, 0.2));
    KD_TREE<P>& tree = *kdtree_ptr;
    tree.Build(cloud);
    if (tree.validnum() != (int)cloud.size()) {
        std::printf("validnum %d of %zu\n", tree.validnum(), cloud.size());
        return 2;
    }
    KD_TREE<P>::PointVector nearest;
    std::vector<float> dist;
    for (int q = 0; q < 50; ++q) {
        const P& c = cloud[(q * 97) % cloud.size()];
        P query;
        query.x = c.x + 0.05f * (q % 5);
        query.y = c.y - 0.03f * (q % 3);
        query.z = c.z;
        tree.Nearest_Search(query, 5, nearest, dist);
        std::vector<float> bf;
        for (const P& p : cloud) {
            const float dx = p.x - query.x, dy = p.y - query.y, dz = p.z - query.z;
            bf.push_back(dx * dx + dy * dy + dz * dz);
        }
        std::partial_sort(bf.begin(), bf.begin() + 5, bf.end());
        if (nearest.size() != 5) return 3;
        for (int j = 0; j < 5; ++j) {
            auto it = by_xyz.find({nearest[j].x, nearest[j].y, nearest[j].z});
            if (it == by_xyz.end() || !same(it->second, nearest[j])) {
                std::printf("query %d: result %d lost its record\n", q, j);
                return 4;
            }
            if (dist[j] != bf[j]) {
                std::printf("query %d: distance %d %g vs %g\n", q, j, dist[j], bf[j]);
                return 5;
            }
        }
    }
    BoxPointType box;
    for (int a = 0; a < 3; ++a) {
        box.vertex_min[a] = -3.0f;
        box.vertex_max[a] = 4.0f;
    }
    KD_TREE<P>::PointVector found;
    tree.Box_Search(box, found);
    size_t want = 0;
    for (const P& p : cloud) {
        const float c[3] = {p.x, p.y, p.z};
        bool in = true;
        for (int a = 0; a < 3; ++a) in = in && c[a] >= box.vertex_min[a] && c[a] < box.vertex_max[a];
        want += in ? 1 : 0;
    }
    if (found.size() != want) return 6;
    for (const P& p : found) {
        auto it = by_xyz.find({p.x, p.y, p.z});
        if (it == by_xyz.end() || !same(it->second, p)) return 7;
    }
    std::printf("ok %zu points, %zu in the box\n", cloud.size(), found.size());
    return 0;
}

// A long downsampled mapping run through the facade keeps a bounded record
// store: records of points the downsample drops or rebuilds remove are
// reclaimed and ids renumbered, while every result still carries the record
// (intensity payload) its point came with.  Prints "ok".
#include <cstdio>
#include <random>

#include "ikd_Tree_amd.hpp"

struct P {
    float x, y, z, intensity;
};

static float tag(float x, float y, float z) { return x * 7.0f + y * 13.0f + z * 17.0f; }

int main() {
    std::mt19937 gen(5);
    std::uniform_real_distribution<float> u(-6.0f, 6.0f);
    KD_TREE<P> tree(0.5, 0.6, 0.5);
    tree.set_payload_reclaim_floor(4000);
    auto pt = [&]() {
        P p{u(gen), u(gen), u(gen), 0.0f};
        p.intensity = tag(p.x, p.y, p.z);
        return p;
    };
    KD_TREE<P>::PointVector cloud(3000);
    for (auto& p : cloud) p = pt();
    tree.Build(cloud);
    size_t most = 0;
    KD_TREE<P>::PointVector res, removed;
    std::vector<float> dist;
    for (int round = 0; round < 40; ++round) {
        KD_TREE<P>::PointVector scan(2000);
        for (auto& p : scan) p = pt();
        tree.Add_Points(scan, true);
        most = std::max(most, tree.payload_records());
        if (round % 10 == 9) {
            std::vector<BoxPointType> box(1);
            for (int a = 0; a < 3; ++a) {
                box[0].vertex_min[a] = -2.0f;
                box[0].vertex_max[a] = 1.0f;
            }
            tree.Delete_Point_Boxes(box);
        }
        tree.acquire_removed_points(removed);
        for (int q = 0; q < 50; ++q) {
            P p = pt();
            tree.Nearest_Search(p, 5, res, dist);
            for (const P& r : res)
                if (r.intensity != tag(r.x, r.y, r.z)) {
                    std::printf("round %d: a result carries another point's record\n", round);
                    return 1;
                }
        }
    }
    for (const P& r : removed)
        if (r.intensity != tag(r.x, r.y, r.z)) {
            std::printf("a removed point carries another point's record\n");
            return 1;
        }
    KD_TREE<P>::PointVector live;
    tree.flatten(live);
    for (const P& r : live)
        if (r.intensity != tag(r.x, r.y, r.z)) {
            std::printf("a live point carries another point's record\n");
            return 1;
        }
    // 83k points went in; the store stays within a few times the tree
    if (most > 4 * (size_t)tree.size() + 12000) {
        std::printf("record store grew to %zu for a tree of %d nodes\n", most, tree.size());
        return 1;
    }
    std::printf("ok store peak %zu, tree %d nodes, %d live\n", most, tree.size(), tree.validnum());
    return 0;
}

// Build 4000 points, delete 2500 of them (rebuilds drop them and record the
// ones not deleted by the downsample), Build a second, small cloud, then
// acquire_removed_points: the records from before the second Build must come
// back unchanged (payload field included), and nothing may read past the
// facade's id table.  Prints "ok".
#include <cstdio>
#include <random>
#include <set>
#include <tuple>

#include "ikd_Tree_amd.hpp"

struct P {
    float x, y, z;
    int tag;
};

int main() {
    std::mt19937 gen(3);
    std::uniform_real_distribution<float> u(-30.0f, 30.0f);
    KD_TREE<P>::PointVector cloud(4000);
    for (int i = 0; i < 4000; ++i) cloud[i] = {u(gen), u(gen), u(gen), i};
    KD_TREE<P> tree(0.5, 0.6, 0.2);
    tree.Build(cloud);
    KD_TREE<P>::PointVector del(cloud.begin(), cloud.begin() + 2500);
    tree.Delete_Points(del);
    KD_TREE<P>::PointVector small(10);
    for (int i = 0; i < 10; ++i) small[i] = {u(gen), u(gen), u(gen), 100000 + i};
    tree.Build(small);
    KD_TREE<P>::PointVector removed;
    tree.acquire_removed_points(removed);
    if (removed.empty()) {
        std::printf("no removed points recorded\n");
        return 1;
    }
    for (const P& p : removed) {
        if (p.tag < 0 || p.tag >= 2500) {
            std::printf("removed point with tag %d was never deleted\n", p.tag);
            return 1;
        }
        const P& o = cloud[p.tag];
        if (o.x != p.x || o.y != p.y || o.z != p.z) {
            std::printf("removed point %d does not match its record\n", p.tag);
            return 1;
        }
    }
    KD_TREE<P>::PointVector again;
    tree.acquire_removed_points(again);
    if (!again.empty()) {
        std::printf("the record was not cleared\n");
        return 1;
    }
    std::printf("ok %zu removed points\n", removed.size());
    return 0;
}

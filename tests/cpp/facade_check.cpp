// Drives the KD_TREE<PointType> drop-in facade (include/ikd_Tree_amd.hpp) the
// way reference callers do (examples/ikd_Tree_demo.cpp:204-256): Build, one
// Nearest_Search per point, Box_Search, Radius_Search, Delete_Point_Boxes,
// Add_Points with downsample, flatten.  Inputs and outputs are raw float32
// files so the Python test can compare every result with the oracle.
// The point type carries an intensity payload (its input index) to show that
// payload fields round-trip through the facade.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ikd_Tree_amd.hpp"

struct PointXYZI {
    float x, y, z, intensity;
};

static std::vector<float> read_f32(const char* path) {
    std::vector<float> v;
    FILE* f = std::fopen(path, "rb");
    if (!f) { std::perror(path); std::exit(2); }
    float buf[4096];
    size_t n;
    while ((n = std::fread(buf, sizeof(float), 4096, f)) > 0) v.insert(v.end(), buf, buf + n);
    std::fclose(f);
    return v;
}

static void write_f32(const std::string& path, const std::vector<float>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    std::fwrite(v.data(), sizeof(float), v.size(), f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s points.f32 queries.f32 boxes.f32 out_prefix\n", argv[0]);
        return 2;
    }
    std::vector<float> p = read_f32(argv[1]), q = read_f32(argv[2]), b = read_f32(argv[3]);
    std::string out = argv[4];
    using Tree = KD_TREE<PointXYZI>;
    Tree::PointVector cloud, queries;
    for (size_t i = 0; i < p.size() / 3; ++i) cloud.push_back({p[3 * i], p[3 * i + 1], p[3 * i + 2], (float)i});
    for (size_t i = 0; i < q.size() / 3; ++i) queries.push_back({q[3 * i], q[3 * i + 1], q[3 * i + 2], -1.0f - i});
    std::vector<BoxPointType> boxes(b.size() / 6);
    for (size_t i = 0; i < boxes.size(); ++i)
        for (int a = 0; a < 3; ++a) {
            boxes[i].vertex_min[a] = b[6 * i + a];
            boxes[i].vertex_max[a] = b[6 * i + 3 + a];
        }

    Tree tree(0.5, 0.6, 0.5);
    tree.Build(cloud);
    // plain per-point Nearest_Search, k = 5, no max_dist
    std::vector<float> knn;  // per query: 5 x (x, y, z, intensity, d2), padded with -1
    Tree::PointVector res;
    std::vector<float> dist;
    for (const auto& pt : queries) {
        tree.Nearest_Search(pt, 5, res, dist);
        for (int j = 0; j < 5; ++j) {
            if (j < (int)res.size()) {
                knn.insert(knn.end(), {res[j].x, res[j].y, res[j].z, res[j].intensity, dist[j]});
            } else {
                knn.insert(knn.end(), {-1, -1, -1, -1, -1});
            }
        }
    }
    write_f32(out + "_knn.f32", knn);
    auto dump = [&](const Tree::PointVector& v, const std::string& name) {
        std::vector<float> f;
        for (const auto& x : v) f.insert(f.end(), {x.x, x.y, x.z, x.intensity});
        write_f32(out + name, f);
    };
    tree.Box_Search(boxes[0], res);
    dump(res, "_box.f32");
    tree.Radius_Search(queries[0], 7.5f, res);
    dump(res, "_radius.f32");
    int deleted = tree.Delete_Point_Boxes(boxes);
    int added = tree.Add_Points(queries, true);
    tree.flatten(res);
    dump(res, "_flat.f32");
    float alpha_bal = 0, alpha_del = 0;
    tree.root_alpha(alpha_bal, alpha_del);
    BoxPointType range = tree.tree_range();
    std::printf("deleted %d added %d validnum %d size %d range %g %g\n", deleted, added, tree.validnum(), tree.size(),
                range.vertex_min[0], range.vertex_max[0]);
    return 0;
}

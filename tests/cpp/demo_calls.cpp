// The call sequences of the reference's three demos, typed against the
// facade (include/ikd_Tree_amd.hpp) with only the include changed:
//   - examples/ikd_Tree_demo.cpp:43, 204-266: a KD_TREE<ikdTree_PointType>
//     object, Build, then rounds of Add_Points (no downsample), Delete_Points,
//     Delete_Point_Boxes, Add_Point_Boxes, Nearest_Search per target, size(),
//     acquire_removed_points;
//   - examples/ikd_Tree_Search_demo.cpp:61-105: KD_TREE<PointType>::Ptr,
//     Build, validnum, Box_Search, Radius_Search;
//   - examples/ikd_Tree_Async_demo.cpp:54-106: Ptr, Build, Delete_Point_Boxes,
//     acquire_removed_points, flatten(Root_Node, PCL_Storage, NOT_RECORD).
// Inputs come from raw float32 files written by the test (no PCD data here);
// every result the demos compute is written back for comparison with the
// oracle.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ikd_Tree_amd.hpp"

// a pcl::PointXYZ-like record: 16 bytes, x y z first
struct PointXYZ {
    float x, y, z, pad;
};

static std::vector<float> read_f32(const std::string& path) {
    std::vector<float> v;
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        std::perror(path.c_str());
        std::exit(2);
    }
    float buf[4096];
    size_t n;
    while ((n = std::fread(buf, sizeof(float), 4096, f)) > 0) v.insert(v.end(), buf, buf + n);
    std::fclose(f);
    return v;
}

static FILE* out_file(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::perror(path.c_str());
        std::exit(2);
    }
    return f;
}

template <class V>
static V points(const std::vector<float>& f) {
    V v;
    for (size_t i = 0; i + 2 < f.size(); i += 3) {
        typename V::value_type p;
        p.x = f[i];
        p.y = f[i + 1];
        p.z = f[i + 2];
        v.push_back(p);
    }
    return v;
}

static std::vector<BoxPointType> boxes(const std::vector<float>& f) {
    std::vector<BoxPointType> b(f.size() / 6);
    for (size_t i = 0; i < b.size(); ++i)
        for (int a = 0; a < 3; ++a) {
            b[i].vertex_min[a] = f[6 * i + a];
            b[i].vertex_max[a] = f[6 * i + 3 + a];
        }
    return b;
}

template <class V>
static void write_points(FILE* f, const V& v) {
    for (const auto& p : v) {
        float r[3] = {p.x, p.y, p.z};
        std::fwrite(r, sizeof(float), 3, f);
    }
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s in_dir out_dir\n", argv[0]);
        return 2;
    }
    const std::string in = std::string(argv[1]) + "/", out = std::string(argv[2]) + "/";
    const int rounds = (int)read_f32(in + "rounds.f32")[0];

    // ---- ikd_Tree_demo.cpp
    {
        using PointVector = KD_TREE<ikdTree_PointType>::PointVector;
        KD_TREE<ikdTree_PointType> ikd_Tree(0.3, 0.6, 0.2);
        PointVector point_cloud = points<PointVector>(read_f32(in + "cloud.f32"));
        ikd_Tree.Build(point_cloud);
        FILE* log = out_file(out + "demo_rounds.f32");
        FILE* knn = out_file(out + "demo_knn.f32");
        FILE* removed = out_file(out + "demo_removed.f32");
        for (int counter = 0; counter < rounds; ++counter) {
            const std::string r = std::to_string(counter);
            PointVector cloud_increment = points<PointVector>(read_f32(in + "inc" + r + ".f32"));
            ikd_Tree.Add_Points(cloud_increment, false);
            PointVector cloud_decrement = points<PointVector>(read_f32(in + "dec" + r + ".f32"));
            ikd_Tree.Delete_Points(cloud_decrement);
            float n_box_deleted = -1;
            if (counter % 2 == 1) {
                std::vector<BoxPointType> Delete_Boxes = boxes(read_f32(in + "delbox" + r + ".f32"));
                n_box_deleted = (float)ikd_Tree.Delete_Point_Boxes(Delete_Boxes);
            }
            if (counter % 2 == 0) {
                std::vector<BoxPointType> Add_Boxes = boxes(read_f32(in + "addbox" + r + ".f32"));
                ikd_Tree.Add_Point_Boxes(Add_Boxes);
            }
            PointVector targets = points<PointVector>(read_f32(in + "target" + r + ".f32"));
            PointVector search_result;
            std::vector<float> PointDist;
            for (const ikdTree_PointType& target : targets) {
                PointVector().swap(search_result);
                ikd_Tree.Nearest_Search(target, 5, search_result, PointDist);
                for (int j = 0; j < 5; ++j) {
                    float rec[4] = {-1, -1, -1, -1};
                    if (j < (int)search_result.size()) {
                        rec[0] = search_result[j].x;
                        rec[1] = search_result[j].y;
                        rec[2] = search_result[j].z;
                        rec[3] = PointDist[j];
                    }
                    std::fwrite(rec, sizeof(float), 4, knn);
                }
            }
            PointVector removed_points;
            ikd_Tree.acquire_removed_points(removed_points);
            write_points(removed, removed_points);
            float line[4] = {(float)ikd_Tree.validnum(), (float)ikd_Tree.size(), n_box_deleted,
                             (float)removed_points.size()};
            std::fwrite(line, sizeof(float), 4, log);
        }
        std::fclose(log);
        std::fclose(knn);
        std::fclose(removed);
        PointVector rest;
        ikd_Tree.flatten(rest);
        FILE* f = out_file(out + "demo_live.f32");
        write_points(f, rest);
        std::fclose(f);
    }

    // ---- ikd_Tree_Search_demo.cpp
    using PointType = PointXYZ;
    using PointVector = KD_TREE<PointType>::PointVector;
    {
        KD_TREE<PointType>::Ptr kdtree_ptr(new KD_TREE<PointType>(0.3, 0.6, 0.2));
        KD_TREE<PointType>& ikd_Tree = *kdtree_ptr;
        PointVector src = points<PointVector>(read_f32(in + "cloud.f32"));
        ikd_Tree.Build(src);
        std::vector<BoxPointType> b = boxes(read_f32(in + "searchbox.f32"));
        PointVector Searched_Points;
        ikd_Tree.Box_Search(b[0], Searched_Points);
        FILE* f = out_file(out + "search_box.f32");
        write_points(f, Searched_Points);
        std::fclose(f);
        std::vector<float> c = read_f32(in + "ball.f32");
        PointType ball_center_pt;
        ball_center_pt.x = c[0];
        ball_center_pt.y = c[1];
        ball_center_pt.z = c[2];
        float radius = c[3];
        PointVector Searched_Points_radius;
        ikd_Tree.Radius_Search(ball_center_pt, radius, Searched_Points_radius);
        f = out_file(out + "search_radius.f32");
        write_points(f, Searched_Points_radius);
        std::fclose(f);
        std::printf("search: valid %d box %d radius %d\n", ikd_Tree.validnum(), (int)Searched_Points.size(),
                    (int)Searched_Points_radius.size());
    }

    // ---- ikd_Tree_Async_demo.cpp
    {
        KD_TREE<PointType>::Ptr kdtree_ptr(new KD_TREE<PointType>(0.3, 0.6, 0.2));
        KD_TREE<PointType>& ikd_Tree = *kdtree_ptr;
        if (ikd_Tree.Root_Node != nullptr) return 3;  // nothing built yet
        PointVector src = points<PointVector>(read_f32(in + "cloud.f32"));
        ikd_Tree.Build(src);
        if (ikd_Tree.Root_Node == nullptr) return 4;
        std::vector<BoxPointType> bx = boxes(read_f32(in + "asyncbox.f32"));
        int num_deleted = ikd_Tree.Delete_Point_Boxes(bx);
        PointVector Removed_Points;
        ikd_Tree.acquire_removed_points(Removed_Points);
        ikd_Tree.flatten(ikd_Tree.Root_Node, ikd_Tree.PCL_Storage, NOT_RECORD);
        PointVector Remaining_Points = ikd_Tree.PCL_Storage;
        FILE* f = out_file(out + "async_live.f32");
        write_points(f, Remaining_Points);
        std::fclose(f);
        std::printf("async: deleted %d removed %d remain %d\n", num_deleted, (int)Removed_Points.size(),
                    (int)Remaining_Points.size());
    }
    return 0;
}

// ikd_pcl_types.hpp — point types with the memory layout and field names of
// PCL's pcl::PointXYZ, pcl::PointXYZI and pcl::PointXYZINormal, the payload
// types the reference instantiates KD_TREE with (ikd_Tree.cpp:1446-1449), for
// builds without PCL.  With PCL present, KD_TREE<pcl::PointXYZI> etc. work
// directly (the facade reads only x, y, z and keeps every other field as the
// point's payload); these types give code without PCL the same records, and
// ikd_pcd.hpp reads and writes their fields by PCD field name.
//
// Layout (PCL's PCL_ADD_POINT4D / PCL_ADD_NORMAL4D unions, 16-byte aligned):
//   PointXYZ        16 bytes: x y z (pad)
//   PointXYZI       32 bytes: x y z (pad) | intensity (pad x3)
//   PointXYZINormal 48 bytes: x y z (pad) | normal_x normal_y normal_z (pad) |
//                             intensity curvature (pad x2)
#pragma once

namespace ikd_pcl {

struct alignas(16) PointXYZ {
    union {
        float data[4];
        struct {
            float x, y, z;
        };
    };
    PointXYZ() : data{0.0f, 0.0f, 0.0f, 1.0f} {}
    PointXYZ(float px, float py, float pz) : data{px, py, pz, 1.0f} {}
};

struct alignas(16) PointXYZI {
    union {
        float data[4];
        struct {
            float x, y, z;
        };
    };
    union {
        float data_c[4];
        struct {
            float intensity;
        };
    };
    PointXYZI() : data{0.0f, 0.0f, 0.0f, 1.0f}, data_c{0.0f, 0.0f, 0.0f, 0.0f} {}
    PointXYZI(float px, float py, float pz, float i) : data{px, py, pz, 1.0f}, data_c{i, 0.0f, 0.0f, 0.0f} {}
};

struct alignas(16) PointXYZINormal {
    union {
        float data[4];
        struct {
            float x, y, z;
        };
    };
    union {
        float data_n[4];
        float normal[3];
        struct {
            float normal_x, normal_y, normal_z;
        };
    };
    union {
        float data_c[4];
        struct {
            float intensity, curvature;
        };
    };
    PointXYZINormal() : data{0.0f, 0.0f, 0.0f, 1.0f}, data_n{0.0f, 0.0f, 0.0f, 0.0f}, data_c{0.0f, 0.0f, 0.0f, 0.0f} {}
};

static_assert(sizeof(PointXYZ) == 16, "PCL layout");
static_assert(sizeof(PointXYZI) == 32, "PCL layout");
static_assert(sizeof(PointXYZINormal) == 48, "PCL layout");

}  // namespace ikd_pcl

// Device operations (kernel launches, copies, fills, primitive calls) that one
// call of the update and search paths issues, counted by the CPU emulation
// build's tally.  These calls are latency-bound chains of dependent launches,
// so the count is their cost; the Python test holds it to a budget.
#include <cmath>
#include <cstdio>
#include <vector>

#include "ikd_tree_amd.h"

extern "C" void emu_dispatch_reset();
extern "C" long long emu_dispatch_tally(char* buf, int len);

static unsigned s = 12345;
static float rnd(float lo, float hi) {
    s = s * 1664525u + 1013904223u;
    return lo + (hi - lo) * (float)((s >> 8) & 0xffffff) / 16777216.0f;
}

// a scan-like batch: a ground plane and a few walls around (cx, cy)
static std::vector<float> scan(float cx, float cy, int n) {
    std::vector<float> p(3 * n);
    for (int i = 0; i < n; ++i) {
        const float a = rnd(0.0f, 6.2831853f), r = rnd(2.0f, 40.0f);
        float x = cx + r * std::cos(a), y = cy + r * std::sin(a), z = 0.0f;
        if (i % 4 == 0) {  // walls
            x = cx + 20.0f * ((i / 4) % 2 ? 1.0f : -1.0f) + rnd(-0.05f, 0.05f);
            z = rnd(0.0f, 5.0f);
        }
        p[3 * i] = x;
        p[3 * i + 1] = y;
        p[3 * i + 2] = z + rnd(-0.02f, 0.02f);
    }
    return p;
}

static long long take(const char* what) {
    static char buf[1 << 16];
    long long n = emu_dispatch_tally(buf, sizeof(buf));
    std::printf("%s %lld\n", what, n);
    emu_dispatch_reset();
    return n;
}

int main() {
    ikd_tree_t t;
    if (ikd_create(0.5f, 0.6f, 0.5f, &t) != IKD_OK) return 2;
    std::vector<float> map;
    for (int f = 0; f < 6; ++f) {
        std::vector<float> p = scan(2.0f * f, 0.0f, 10000);
        map.insert(map.end(), p.begin(), p.end());
    }
    if (ikd_build(t, map.data(), nullptr, (int64_t)(map.size() / 3)) != IKD_OK) return 3;
    const int nq = 8000;
    std::vector<float> ox(nq * 5 * 3), od(nq * 5);
    std::vector<uint32_t> oi(nq * 5);
    std::vector<int32_t> oc(nq);
    for (int f = 6; f < 10; ++f) {
        std::vector<float> p = scan(2.0f * f, 0.0f, nq);
        emu_dispatch_reset();
        if (ikd_nearest_search(t, p.data(), nq, 5, 5.0, ox.data(), od.data(), oi.data(), oc.data()) != IKD_OK)
            return 4;
        take("knn");
        int32_t added = 0;
        if (ikd_add_points(t, p.data(), nullptr, nq, 1, &added) != IKD_OK) return 5;
        take("add_points");
    }
    // a sliding-window trim: everything outside a 30 m cube around the last pose
    const float big = 1e7f, cx = 18.0f, h = 15.0f;
    const float boxes[6][6] = {{-big, -big, -big, cx - h, big, big}, {cx + h, -big, -big, big, big, big},
                               {-big, -big, -big, big, -h, big},     {-big, h, -big, big, big, big},
                               {-big, -big, -big, big, big, -h},     {-big, -big, h, big, big, big}};
    emu_dispatch_reset();
    int32_t del = 0;
    if (ikd_delete_point_boxes(t, &boxes[0][0], 6, &del) != IKD_OK) return 6;
    take("trim");
    ikd_destroy(t);
    std::printf("ok\n");
    return 0;
}

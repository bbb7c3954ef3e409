// Regression check (CPU, emulation build with AddressSanitizer): a request
// may be released by its caller as soon as ikd_server_wait hands it back,
// while the worker goes on to maintain the request's map.  The program frees
// every request right after getting it back; a worker that still read a
// request after handing it back is reported by AddressSanitizer.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ikd_tree_amd.h"

int main() {
    const int n = 4000;
    std::vector<float> pts(3 * n);
    unsigned s = 12345;
    for (auto& v : pts) {
        s = s * 1664525u + 1013904223u;
        v = (float)((s >> 8) & 0xffff) / 65536.0f * 40.0f - 20.0f;
    }
    ikd_tree_t t;
    if (ikd_create(0.5f, 0.6f, 0.5f, &t) != IKD_OK) return 2;
    if (ikd_build(t, pts.data(), nullptr, n) != IKD_OK) return 3;
    ikd_set_tuning(t, IKD_TUNE_DEFER_REBUILD, 50);  // give the idle-map maintenance work to do
    ikd_server_t srv;
    if (ikd_server_create(2, 1ll << 30, &srv) != IKD_OK) return 4;
    int32_t map = -1;
    ikd_server_add_map(srv, t, &map);
    for (int round = 0; round < 12; ++round) {
        // a clustered scan: unbalances one side of the tree
        const int m = 1500;
        auto* scan = static_cast<float*>(std::malloc(sizeof(float) * 3 * m));
        for (int i = 0; i < 3 * m; ++i) {
            s = s * 1664525u + 1013904223u;
            scan[i] = (float)((s >> 8) & 0xffff) / 65536.0f * 4.0f + 5.0f + (float)round;
        }
        auto* r = static_cast<ikd_request_t*>(std::calloc(1, sizeof(ikd_request_t)));
        r->map = map;
        r->n = m;
        r->xyz = scan;
        r->k = 5;
        r->max_dist = 3.0;
        r->downsample = 1;
        if (ikd_server_submit(srv, r) != IKD_OK) return 5;
        ikd_request_t* back[4];
        int32_t got = 0;
        while (got == 0) ikd_server_wait(srv, back, 4, -1, &got);
        if (got != 1 || back[0] != r || r->status != IKD_OK) return 6;
        std::free(r);  // released right away, as a caller may do
        std::free(scan);
    }
    ikd_server_destroy(srv);
    int64_t live = 0;
    ikd_validnum(t, &live);
    ikd_destroy(t);
    std::printf("ok %lld\n", (long long)live);
    return 0;
}

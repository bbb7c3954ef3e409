// Server stress on the CPU emulation build with AddressSanitizer: 8 workers,
// 8 maps with deferred large-subtree rebuilds and chunked maintenance while
// no request waits, many requests in flight at once (kNN + Add_Points + a trim on
// every third request; every other request's results in page-locked
// buffers), maps added while requests run.  Every request must
// succeed and nothing may touch freed or out-of-bounds memory.  This is the
// path (worker -> ikd_rebuild_deferred -> rebuild -> build -> scan) of a host
// segmentation fault once seen under a profiler on the GPU.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "ikd_tree_amd.h"

static unsigned s = 777;
static float rnd(float lo, float hi) {
    s = s * 1664525u + 1013904223u;
    return lo + (hi - lo) * (float)((s >> 8) & 0xffffff) / 16777216.0f;
}

int main() {
    const int n_maps = 8, per_map = 6, n0 = 3000;
    std::vector<ikd_tree_t> trees(n_maps);
    ikd_server_t srv;
    if (ikd_server_create(8, 1ll << 30, &srv) != IKD_OK) return 2;
    ikd_server_set_maintenance(srv, 300);
    std::vector<int32_t> map_id(n_maps, -1);
    std::vector<ikd_request_t*> reqs;
    std::vector<float*> bufs;
    for (int m = 0; m < n_maps; ++m) {
        std::vector<float> pts(3 * n0);
        for (auto& v : pts) v = rnd(-20.0f, 20.0f);
        if (ikd_create(0.5f, 0.6f, 0.5f, &trees[m]) != IKD_OK) return 3;
        if (ikd_build(trees[m], pts.data(), nullptr, n0) != IKD_OK) return 4;
        ikd_set_tuning(trees[m], IKD_TUNE_DEFER_REBUILD, 40);
        ikd_server_add_map(srv, trees[m], &map_id[m]);
        // the first maps get requests before the later ones are added
        for (int j = 0; j < per_map; ++j) {
            const int n = 500 + 400 * ((m + j) % 4);
            float* scan = (float*)std::malloc(sizeof(float) * 3 * n);
            const float cx = rnd(-15.0f, 15.0f), cy = rnd(-15.0f, 15.0f);
            for (int i = 0; i < n; ++i) {
                scan[3 * i] = cx + rnd(-3.0f, 3.0f);
                scan[3 * i + 1] = cy + rnd(-3.0f, 3.0f);
                scan[3 * i + 2] = rnd(-1.0f, 1.0f);
            }
            float* trim = (float*)std::malloc(sizeof(float) * 6);
            trim[0] = cx - 1; trim[1] = cy - 1; trim[2] = -5; trim[3] = cx + 1; trim[4] = cy + 1; trim[5] = 5;
            auto* r = (ikd_request_t*)std::calloc(1, sizeof(ikd_request_t));
            r->map = map_id[m];
            r->n = n;
            r->xyz = scan;
            r->k = 5;
            r->max_dist = 2.0;
            r->downsample = 1;
            r->n_trim = (j % 3 == 2) ? 1 : 0;
            r->trim_boxes = trim;
            if (j % 2) {
                // page-locked result buffers, filled by DMA
                void *d2 = nullptr, *ids = nullptr, *cnt = nullptr;
                if (ikd_host_alloc(sizeof(float) * n * 5, &d2) != IKD_OK ||
                    ikd_host_alloc(sizeof(uint32_t) * n * 5, &ids) != IKD_OK ||
                    ikd_host_alloc(sizeof(int32_t) * n, &cnt) != IKD_OK)
                    return 9;
                r->out_d2 = (float*)d2;
                r->out_ids = (uint32_t*)ids;
                r->out_counts = (int32_t*)cnt;
                r->flags = IKD_REQ_PINNED_RESULTS;
            } else {
                r->out_d2 = (float*)std::malloc(sizeof(float) * n * 5);
                r->out_ids = (uint32_t*)std::malloc(sizeof(uint32_t) * n * 5);
                r->out_counts = (int32_t*)std::malloc(sizeof(int32_t) * n);
            }
            if (ikd_server_submit(srv, r) != IKD_OK) return 5;
            reqs.push_back(r);
            bufs.push_back(scan);
            bufs.push_back(trim);
        }
    }
    int done = 0;
    while (done < (int)reqs.size()) {
        ikd_request_t* back[16];
        int32_t got = 0;
        ikd_server_wait(srv, back, 16, -1, &got);
        for (int i = 0; i < got; ++i) {
            if (back[i]->status != IKD_OK) {
                std::printf("request failed: %d\n", back[i]->status);
                return 6;
            }
        }
        done += got;
    }
    // let the workers finish the maps' maintenance, then require that some ran
    for (int spin = 0;; ++spin) {
        int64_t pend = 0, running = 0;
        ikd_server_pending(srv, &pend, &running);
        if (pend == 0 && running == 0) break;
        if (spin > 600000) return 7;  // 10 minutes
        std::this_thread::sleep_for(std::chrono::milliseconds(1));
    }
    int64_t rebuilt = 0;
    ikd_server_stats(srv, &rebuilt);
    if (rebuilt == 0) {
        std::printf("no deferred rebuild ran\n");
        return 8;
    }
    ikd_server_destroy(srv);
    for (auto* r : reqs) {
        if (r->flags & IKD_REQ_PINNED_RESULTS) {
            ikd_host_free(r->out_d2);
            ikd_host_free(r->out_ids);
            ikd_host_free(r->out_counts);
        } else {
            std::free(r->out_d2);
            std::free(r->out_ids);
            std::free(r->out_counts);
        }
        std::free(r);
    }
    for (float* b : bufs) std::free(b);
    for (auto t : trees) ikd_destroy(t);
    std::printf("ok %d requests, %lld deferred rebuilds\n", done, (long long)rebuilt);
    return 0;
}

// ikd_serve.hip — request loop serving many device-resident maps at once.
//
// A request is one sensor scan for one map: Nearest_Search of every scan point
// against the map (ikd_Tree.cpp:366-397), then Add_Points of the scan
// (cpp:413-489), then optionally Delete_Point_Boxes (cpp:535-556) for a
// sliding-window trim.  The server runs `slots` worker threads; each takes
// the oldest pending request whose map is idle and whose device memory fits
// the HBM budget, so
//   - requests for one map run one at a time in submission order (one writer
//     per tree), which makes every result identical to running the same
//     requests alone, one after another;
//   - requests for different maps run concurrently, each on its map's stream,
//     so their kernels overlap on the device;
//   - a request is admitted as soon as a slot and memory free up.
// The server only uses the public C ABI of the tree library and the HIP
// runtime; it owns per-slot device buffers for scans and kNN results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ikd_tree_amd.h"

namespace {

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct SlotBufs {
    float* xyz = nullptr;  // scan on the device
    float* out_xyz = nullptr;
    float* out_d2 = nullptr;
    uint32_t* out_ids = nullptr;
    int32_t* out_cnt = nullptr;
    long long cap_pts = 0, cap_res = 0;
    // pinned host mirrors: the scan goes up and the results come down by
    // DMA without stopping the host (pageable copies wait for the stream)
    float* h_xyz_in = nullptr;
    float* h_xyz = nullptr;
    float* h_d2 = nullptr;
    uint32_t* h_ids = nullptr;
    int32_t* h_cnt = nullptr;
    long long hcap_pts = 0, hcap_res = 0;
    hipEvent_t ev = nullptr;  // behind the result copies
};

// device bytes a request needs beyond its map: the scan, the kNN results and
// the tree's insert scratch (about 160 bytes per inserted point)
int64_t request_bytes(const ikd_request_t* r) {
    const int64_t n = r->n, k = std::max(r->k, 0);
    return n * 12 + n * k * 20 + n * 4 + (r->downsample >= 0 ? n * 160 : 0) + (1 << 20);
}

}  // namespace

struct ikd_server_s {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<ikd_tree_t> maps;
    std::vector<char> busy;
    std::vector<int64_t> map_bytes;  // each map's device bytes, as of its last request or maintenance
    int64_t resident = 0;            // their sum
    std::deque<ikd_request_t*> pending;
    std::deque<ikd_request_t*> done;
    std::vector<std::thread> workers;
    std::vector<SlotBufs> bufs;
    int64_t budget = 0;
    int64_t in_flight_bytes = 0;
    int64_t in_flight = 0;
    int64_t maintain_points = -1;  // deferred-rebuild budget per maintenance step (-1: no limit, 0: never)
    int64_t maintenance_runs = 0;
    std::deque<int> maint_due;     // maps that may have deferred rebuilds left, oldest first
    std::vector<char> maint_queued;  // map is in maint_due
    int64_t maint_running = 0;     // maintenance steps on the device now (at most one)
    bool stop = false;
    std::string err;

    // Slot buffers grow with stream-ordered allocations on the request's
    // stream (hipFree would wait for every stream of the device).  The slot's
    // previous request ended with a host wait on its own stream, so nothing
    // still reads a buffer freed here.
    template <class T>
    static bool regrow(T** p, long long count, size_t elem, hipStream_t st) {
        if (*p) hipFreeAsync(*p, st);
        *p = nullptr;
        return hipMallocAsync((void**)p, (size_t)count * elem, st) == hipSuccess;
    }
    template <class T>
    static bool regrow_host(T** p, long long count, size_t elem) {
        if (*p) hipHostFree(*p);
        *p = nullptr;
        return hipHostMalloc((void**)p, (size_t)count * elem, hipHostMallocDefault) == hipSuccess;
    }
    // The slot's previous request ended with a host wait, so no copy still
    // uses a pinned buffer freed here.
    int grow_host(SlotBufs& b, long long n, int k, ikd_request_t* r) {
        if (!b.ev && hipEventCreateWithFlags(&b.ev, hipEventDisableTiming) != hipSuccess) {
            r->status = IKD_ERR_HIP;
            return IKD_ERR_HIP;
        }
        if (n > b.hcap_pts) {
            long long c = std::max(n, b.hcap_pts + b.hcap_pts / 2);
            if (!regrow_host(&b.h_xyz_in, c, 12) || !regrow_host(&b.h_cnt, c, 4)) {
                b.hcap_pts = 0;
                r->status = IKD_ERR_NOMEM;
                return IKD_ERR_NOMEM;
            }
            b.hcap_pts = c;
        }
        long long res = n * std::max(k, 1);
        if (res > b.hcap_res) {
            long long c = std::max(res, b.hcap_res + b.hcap_res / 2);
            if (!regrow_host(&b.h_xyz, c, 12) || !regrow_host(&b.h_d2, c, 4) || !regrow_host(&b.h_ids, c, 4)) {
                b.hcap_res = 0;
                r->status = IKD_ERR_NOMEM;
                return IKD_ERR_NOMEM;
            }
            b.hcap_res = c;
        }
        return IKD_OK;
    }
    static void release(SlotBufs& b) {
        for (void* p : {(void*)b.xyz, (void*)b.out_xyz, (void*)b.out_d2, (void*)b.out_ids, (void*)b.out_cnt})
            if (p) hipFree(p);
        for (void* p : {(void*)b.h_xyz_in, (void*)b.h_xyz, (void*)b.h_d2, (void*)b.h_ids, (void*)b.h_cnt})
            if (p) hipHostFree(p);
        if (b.ev) hipEventDestroy(b.ev);
        b = SlotBufs{};
    }
    int grow(SlotBufs& b, long long n, int k, ikd_request_t* r, hipStream_t st) {
        if (grow_host(b, n, k, r) != IKD_OK) return r->status;
        if (n > b.cap_pts) {
            long long c = std::max(n, b.cap_pts + b.cap_pts / 2);
            if (!regrow(&b.xyz, c, 12, st) || !regrow(&b.out_cnt, c, 4, st)) {
                b.cap_pts = 0;
                r->status = IKD_ERR_NOMEM;
                return IKD_ERR_NOMEM;
            }
            b.cap_pts = c;
        }
        long long res = n * std::max(k, 1);
        if (res > b.cap_res) {
            long long c = std::max(res, b.cap_res + b.cap_res / 2);
            if (!regrow(&b.out_xyz, c, 12, st) || !regrow(&b.out_d2, c, 4, st) || !regrow(&b.out_ids, c, 4, st)) {
                b.cap_res = 0;
                r->status = IKD_ERR_NOMEM;
                return IKD_ERR_NOMEM;
            }
            b.cap_res = c;
        }
        return IKD_OK;
    }

    // the request's work, on its map's stream (t: the request's map, read
    // under the server lock at admission)
    int run(SlotBufs& b, ikd_request_t* r, ikd_tree_t t) {
        void* sv = nullptr;
        int rc = ikd_stream(t, &sv);
        if (rc != IKD_OK) return rc;
        hipStream_t st = (hipStream_t)sv;
        const long long n = r->n;
        if (n > 0) {
            rc = grow(b, n, r->k, r, st);
            if (rc != IKD_OK) return rc;
            std::memcpy(b.h_xyz_in, r->xyz, n * 12);
            if (hipMemcpyAsync(b.xyz, b.h_xyz_in, n * 12, hipMemcpyHostToDevice, st) != hipSuccess) return IKD_ERR_HIP;
        }
        const bool knn = r->k > 0 && n > 0;
        const long long res = knn ? n * r->k : 0;
        if (knn) {
            rc = ikd_nearest_search_device(t, b.xyz, n, r->k, r->max_dist, b.out_xyz, b.out_d2, b.out_ids,
                                           b.out_cnt);
            if (rc != IKD_OK) return rc;
            // page-locked caller buffers take the results straight from the
            // device; others through the slot's page-locked mirrors
            const bool direct = (r->flags & IKD_REQ_PINNED_RESULTS) != 0;
            float* hd2 = direct ? r->out_d2 : b.h_d2;
            uint32_t* hids = direct ? r->out_ids : b.h_ids;
            float* hxyz = direct ? r->out_xyz : b.h_xyz;
            int32_t* hcnt = direct ? r->out_counts : b.h_cnt;
            if (r->out_d2 && hipMemcpyAsync(hd2, b.out_d2, res * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
                return IKD_ERR_HIP;
            if (r->out_ids && hipMemcpyAsync(hids, b.out_ids, res * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
                return IKD_ERR_HIP;
            if (r->out_xyz && hipMemcpyAsync(hxyz, b.out_xyz, res * 12, hipMemcpyDeviceToHost, st) != hipSuccess)
                return IKD_ERR_HIP;
            if (r->out_counts && hipMemcpyAsync(hcnt, b.out_cnt, n * 4, hipMemcpyDeviceToHost, st) != hipSuccess)
                return IKD_ERR_HIP;
            if (hipEventRecord(b.ev, st) != hipSuccess) return IKD_ERR_HIP;
            // The results are handed over before the insert is queued.  (Queuing
            // the insert behind the search without this wait lets every map's
            // work pile up in the shared hardware queues, which measured
            // worse for the slowest requests.)
            if (hipEventSynchronize(b.ev) != hipSuccess) return IKD_ERR_HIP;
            if (!direct) {
                if (r->out_d2) std::memcpy(r->out_d2, b.h_d2, res * 4);
                if (r->out_ids) std::memcpy(r->out_ids, b.h_ids, res * 4);
                if (r->out_xyz) std::memcpy(r->out_xyz, b.h_xyz, res * 12);
                if (r->out_counts) std::memcpy(r->out_counts, b.h_cnt, n * 4);
            }
        }
        if (r->downsample >= 0 && n > 0) {
            int32_t added = 0;
            rc = ikd_add_points_device(t, b.xyz, nullptr, n, r->downsample, &added);
            if (rc != IKD_OK) return rc;
            r->added = added;
        }
        r->t_insert_done_ns = now_ns();
        if (r->n_trim > 0) {
            int32_t del = 0;
            rc = ikd_delete_point_boxes(t, r->trim_boxes, r->n_trim, &del);
            if (rc != IKD_OK) return rc;
            r->trimmed = del;
        }
        return ikd_synchronize(t);
    }

    // The map's device bytes after a request or a maintenance step (its pool
    // and scratch may have grown or been compacted); called with lk held.
    void account(std::unique_lock<std::mutex>& lk, int map, ikd_tree_t tree) {
        int64_t mb = 0;
        lk.unlock();
        ikd_device_bytes(tree, &mb);
        lk.lock();
        resident += mb - map_bytes[map];
        map_bytes[map] = mb;
    }

    void queue_maintenance(int map) {
        if (maintain_points != 0 && !maint_queued[map]) {
            maint_queued[map] = 1;
            maint_due.push_back(map);
        }
    }

    // One step of a map's deferred rebuilds (ikd_rebuild_deferred with the
    // maintenance budget), run only while no request is waiting and one step
    // at a time, so that the rebuilds take the device's idle time rather than
    // the requests'.  A request for the map that arrives meanwhile waits for
    // this one step.  Called with lk held and busy[map] set.
    void maintain(std::unique_lock<std::mutex>& lk, int map) {
        ikd_tree_t tree = maps[map];
        const int64_t budget_pts = maintain_points;
        lk.unlock();
        int32_t nreb = 0;
        const int mrc = ikd_rebuild_deferred(tree, budget_pts, &nreb);
        lk.lock();
        maintenance_runs += nreb;
        account(lk, map, tree);
        // a bounded step that ran something may have left more
        if (mrc == IKD_OK && nreb > 0 && budget_pts > 0) queue_maintenance(map);
        busy[map] = 0;
        --maint_running;
        cv_done.notify_all();
        cv_work.notify_all();
    }

    void worker(int slot) {
        SlotBufs& b = bufs[slot];
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            ikd_request_t* r = nullptr;
            int64_t need = 0;
            int mmap = -1;
            while (true) {
                if (stop && pending.empty()) return;
                for (auto it = pending.begin(); it != pending.end(); ++it) {
                    ikd_request_t* c = *it;
                    if (busy[c->map]) continue;
                    const int64_t nb = request_bytes(c);
                    // admit when the resident maps, the running requests and
                    // this one fit; an idle server always admits its oldest
                    // runnable request
                    if (resident + in_flight_bytes + nb > budget && in_flight > 0) continue;
                    r = c;
                    need = nb;
                    pending.erase(it);
                    break;
                }
                if (r) break;
                // no request waits: the oldest idle map with deferred rebuilds
                if (pending.empty() && maint_running == 0 && !stop) {
                    for (auto it = maint_due.begin(); it != maint_due.end(); ++it) {
                        if (busy[*it]) continue;
                        mmap = *it;
                        maint_due.erase(it);
                        maint_queued[mmap] = 0;
                        break;
                    }
                    if (mmap >= 0) break;
                }
                cv_work.wait(lk);
            }
            if (mmap >= 0) {
                busy[mmap] = 1;
                ++maint_running;
                maintain(lk, mmap);
                continue;
            }
            busy[r->map] = 1;
            in_flight_bytes += need;
            ++in_flight;
            // the map list may grow (and move) while the lock is released
            const int map = r->map;
            ikd_tree_t tree = maps[map];
            lk.unlock();
            // the caller may release the request as soon as it is handed back:
            // nothing below reads it after that
            r->t_start_ns = now_ns();
            ikd_counters_t c0{}, c1{};
            ikd_get_counters(tree, &c0);
            int rc = run(b, r, tree);
            ikd_get_counters(tree, &c1);
            r->t_done_ns = now_ns();
            r->delta.rebuilds = c1.rebuilds - c0.rebuilds;
            r->delta.rebuilt_points = c1.rebuilt_points - c0.rebuilt_points;
            r->delta.compactions = c1.compactions - c0.compactions;
            r->delta.pool_grows = c1.pool_grows - c0.pool_grows;
            r->delta.scratch_grows = c1.scratch_grows - c0.scratch_grows;
            r->delta.host_waits = c1.host_waits - c0.host_waits;
            if (rc != IKD_OK && r->status == IKD_OK) r->status = rc;
            lk.lock();
            done.push_back(r);
            r = nullptr;
            cv_done.notify_all();
            queue_maintenance(map);
            account(lk, map, tree);
            busy[map] = 0;
            in_flight_bytes -= need;
            --in_flight;
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

int64_t ikd_now_ns(void) { return now_ns(); }

int ikd_host_alloc(int64_t bytes, void** out) {
    if (!out || bytes <= 0) return IKD_ERR_ARG;
    *out = nullptr;
    return hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault) == hipSuccess ? IKD_OK : IKD_ERR_NOMEM;
}

int ikd_host_free(void* p) {
    if (!p) return IKD_OK;
    return hipHostFree(p) == hipSuccess ? IKD_OK : IKD_ERR_HIP;
}

int ikd_server_create(int slots, int64_t hbm_budget_bytes, ikd_server_t* out) {
    if (!out || slots < 1 || slots > 64 || hbm_budget_bytes <= 0) return IKD_ERR_ARG;
    ikd_server_t s = new ikd_server_s();
    s->budget = hbm_budget_bytes;
    s->bufs.resize(slots);
    for (int i = 0; i < slots; ++i) s->workers.emplace_back([s, i] { s->worker(i); });
    *out = s;
    return IKD_OK;
}

int ikd_server_add_map(ikd_server_t s, ikd_tree_t tree, int32_t* out_map) {
    if (!s || !tree || !out_map) return IKD_ERR_ARG;
    int64_t mb = 0;
    const int rc = ikd_device_bytes(tree, &mb);
    if (rc != IKD_OK) return rc;
    std::lock_guard<std::mutex> g(s->mu);
    s->maps.push_back(tree);
    s->busy.push_back(0);
    s->maint_queued.push_back(0);
    s->map_bytes.push_back(mb);
    s->resident += mb;
    *out_map = (int32_t)(s->maps.size() - 1);
    return IKD_OK;
}

int ikd_server_submit(ikd_server_t s, ikd_request_t* r) {
    if (!s || !r) return IKD_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    if (r->map < 0 || r->map >= (int32_t)s->maps.size() || r->n < 0 || r->k > 64 || (r->n > 0 && !r->xyz) ||
        (r->n_trim > 0 && !r->trim_boxes) || s->stop) {
        s->err = "server: bad request";
        return IKD_ERR_ARG;
    }
    r->status = IKD_OK;
    r->added = r->trimmed = 0;
    r->t_start_ns = r->t_done_ns = 0;
    s->pending.push_back(r);
    s->cv_work.notify_one();
    return IKD_OK;
}

int ikd_server_wait(ikd_server_t s, ikd_request_t** out, int32_t max_out, int32_t timeout_ms, int32_t* out_n) {
    if (!s || !out || !out_n || max_out < 1) return IKD_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && timeout_ms != 0) {
        auto pred = [s] { return !s->done.empty() || (s->pending.empty() && s->in_flight == 0); };
        if (timeout_ms < 0) s->cv_done.wait(lk, pred);
        else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
    }
    int n = 0;
    while (n < max_out && !s->done.empty()) {
        out[n++] = s->done.front();
        s->done.pop_front();
    }
    *out_n = n;
    return IKD_OK;
}

int ikd_server_set_maintenance(ikd_server_t s, int64_t max_points) {
    if (!s) return IKD_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    s->maintain_points = max_points;
    return IKD_OK;
}

int ikd_server_stats(ikd_server_t s, int64_t* out_rebuilds) {
    if (!s) return IKD_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    if (out_rebuilds) *out_rebuilds = s->maintenance_runs;
    return IKD_OK;
}

int ikd_server_memory(ikd_server_t s, int64_t* out_resident, int64_t* out_in_flight_bytes) {
    if (!s) return IKD_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    if (out_resident) *out_resident = s->resident;
    if (out_in_flight_bytes) *out_in_flight_bytes = s->in_flight_bytes;
    return IKD_OK;
}

int ikd_server_pending(ikd_server_t s, int64_t* out_pending, int64_t* out_in_flight) {
    if (!s) return IKD_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    if (out_pending) *out_pending = (int64_t)s->pending.size();
    if (out_in_flight) *out_in_flight = s->in_flight + s->maint_running + (int64_t)s->maint_due.size();
    return IKD_OK;
}

int ikd_server_destroy(ikd_server_t s) {
    if (!s) return IKD_ERR_ARG;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    for (auto& w : s->workers) w.join();
    // the workers' last requests ended with a wait on their streams
    hipDeviceSynchronize();
    for (auto& b : s->bufs) ikd_server_s::release(b);
    delete s;
    return IKD_OK;
}

}  // extern "C"

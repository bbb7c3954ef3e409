// ikd_api.hip — the C ABI (include/ikd_tree_amd.h) and the host-side
// orchestration of the device kernels: node-pool management, the build
// driver, refit + Criterion_Check + Rebuild, and one function per public
// KD_TREE method.
//
// Every public call leaves the tree consistent: lazy labels only below
// tree_deleted nodes, counts and AABBs refit up every modified path, and no
// node on a modified path violating Criterion_Check (the topmost violator of
// each path is rebuilt, then its ancestors are re-checked, as the reference's
// unwinding does, ikd_Tree.cpp:704-707, 860-863).
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ikd_internal.h"
#include "ikd_tree_amd.h"

using namespace ikd;

namespace {

constexpr int BS = 256;
inline int nblk(long long n, int bs = BS) { return (int)((n + bs - 1) / bs); }
inline int bits_for(long long v) {
    int b = 0;
    while ((1ll << b) <= v) ++b;
    return b < 1 ? 1 : b;
}

// ----------------------------------------------------------- small kernels
__global__ void k_reset_counters(DevHdr* h, int which) {
    if (which & 1) { h->dirty_count = 0; h->viol_count = 0; h->top_count = 0; }
    if (which & 2) h->rec_count = 0;
    if (which & 4) h->conflicts = 0;
    if (which & 8) { h->split_index = INT_MAX; h->added = 0; h->ins_count = 0; }
    if (which & 16) h->group_count = 0;
}

__global__ void k_link_defer_flags(Pool pool, const int* __restrict__ defer, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicOr(&pool.meta[defer[i]], M_DEFER);
}

__global__ void k_link_bits(NodeHot* hot, int node, uint32_t set, uint32_t clear) {
    uint32_t* p = reinterpret_cast<uint32_t*>(&hot[node].f[3]);
    uint32_t v = *p;
    *p = (v | set) & ~clear;
}

// items[i] = (parent, link bits to clear): empty child slots, several of
// which may share a parent's link word
__global__ void k_clear_links(NodeHot* hot, const int2* __restrict__ items, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAnd(reinterpret_cast<uint32_t*>(&hot[items[i].x].f[3]), ~(uint32_t)items[i].y);
}

__global__ void k_init_pool(Pool pool, DevHdr* hdr) {
    // sentinel (node 0): its left child slot is the root (node 2, pair 1)
    float* f = pool.hot[SENTINEL].f;
    f[0] = f[1] = f[2] = 0.0f;
    f[3] = bitsf(1u << LK_PAIR_SHIFT);
    for (int i = 4; i < 16; ++i) f[i] = 0.0f;
    pool.meta[SENTINEL] = M_PDEL;
    pool.cnt[SENTINEL] = make_int4(0, 0, 0, -1);
    pool.par[SENTINEL] = -1;
    pool.ids[SENTINEL] = 0xffffffffu;
    pool.meta[1] = M_PDEL;
    pool.cnt[1] = make_int4(0, 0, 0, -1);
    pool.par[1] = -1;
    DevHdr z = {};
    z.pair_used = 2;
    z.split_index = INT_MAX;
    z.removed_count = hdr->removed_count;  // the removed-point record outlives the nodes
    *hdr = z;
}

__global__ void k_gather_pts(const float4* __restrict__ in, const uint32_t* __restrict__ idx, int n,
                             float4* __restrict__ out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[idx[i]];
}

__global__ void k_seeds_from_attach(const int* __restrict__ attach, DevHdr* hdr, int bound, int* __restrict__ seeds) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= bound) return;
    seeds[g] = g < hdr->group_count ? (attach[g] >> 1) : -1;
}

__global__ void k_group_maxlen(const int4* __restrict__ seg, DevHdr* hdr, int bound, int* __restrict__ out) {
    int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= bound || g >= hdr->group_count) return;
    atomicMax(out, seg[g].y - seg[g].x + 1);
}

// (node, live points, depth, parent) of the topmost violators, plus the node
// count of the subtree before its rebuild
__global__ void k_node_info(Pool pool, const int* __restrict__ top, int n, int4* __restrict__ out,
                            int* __restrict__ nodes_before) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int u = top[2 * i];
    int4 c = pool.cnt[u];
    out[i] = make_int4(u, c.x - c.y, c.w, pool.par[u]);
    nodes_before[i] = top[2 * i + 1];
}


// id := its index in the ascending list old (ids not in it are left alone:
// nodes a rebuild left behind in the pool)
__device__ __forceinline__ uint32_t relabel(uint32_t id, const uint32_t* old, int n) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (old[mid] < id) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && old[lo] == id) ? (uint32_t)lo : id;
}
__global__ void k_relabel_ids(uint32_t* __restrict__ ids, int used, const uint32_t* __restrict__ old, int n) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= ROOT && u < used) ids[u] = relabel(ids[u], old, n);
}
__global__ void k_relabel_records(float4* __restrict__ rec, int m, const uint32_t* __restrict__ old, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) rec[i].w = bitsf(relabel(fbits(rec[i].w), old, n));
}

__global__ void k_unpack_points(const float4* __restrict__ in, int n, float* __restrict__ xyz,
                                uint32_t* __restrict__ ids) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = in[i];
    xyz[3 * i] = p.x;
    xyz[3 * i + 1] = p.y;
    xyz[3 * i + 2] = p.z;
    if (ids) ids[i] = fbits(p.w);
}

// root summary for the host: cnt[ROOT], sentinel link, root AABB
struct RootInfo {
    int4 cnt;
    int4 child;  // TreeSize of left, right child (-1 if absent), link word, pad
    float box[8];
};
// Subtrees to rebuild whose summaries come back with the header in one copy
// (more than this: a second read).
constexpr int TOPINFO_CAP = 1024;
// What one host wait brings back: the header, the root summary and, after a
// refit, (node, live points, depth, parent) and the node count of the first
// TOPINFO_CAP topmost violators.  One device block and its pinned mirror.
struct DevStatus {
    DevHdr hdr;
    RootInfo root;
    int4 top_info[TOPINFO_CAP];
    int top_before[TOPINFO_CAP];
};
constexpr size_t STATUS_SHORT = offsetof(DevStatus, top_info);  // header + root summary
__global__ void k_top_info(Pool pool, const int* __restrict__ top, DevStatus* s) {
    const int n = min(s->hdr.top_count, TOPINFO_CAP);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int u = top[2 * i];
        const int4 c = pool.cnt[u];
        s->top_info[i] = make_int4(u, c.x - c.y, c.w, pool.par[u]);
        s->top_before[i] = top[2 * i + 1];
    }
}
// Enumerate: the walk must have written exactly the live points the
// subtrees' counts promise (overflow bit 32 otherwise, reported by the next
// header read).
__global__ void k_enum_check(DevHdr* hdr, int expect) {
    if (hdr->rec_count != expect) atomicOr(&hdr->overflow, 32);
}
__global__ void k_root_info(Pool pool, RootInfo* out) {
    RootInfo r = {};
    uint32_t lk0 = fbits(pool.hot[SENTINEL].f[3]);
    r.child.z = (int)lk0;
    if (lk0 & LK_HASL) {
        r.cnt = pool.cnt[ROOT];
        uint32_t lk = fbits(pool.hot[ROOT].f[3]);
        int pair = (int)lk_pair(lk);
        r.child.x = (lk & LK_HASL) ? pool.cnt[2 * pair].x : -1;
        r.child.y = (lk & LK_HASR) ? pool.cnt[2 * pair + 1].x : -1;
        for (int i = 0; i < 6; ++i) r.box[i] = pool.hot[SENTINEL].f[4 + i];
    }
    *out = r;
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

// Pinned staging for small host-to-device copies of host values (segment
// lists, counters, tree tables): hipMemcpyAsync from pageable memory may read
// the source after the call returns, so values that live on the host stack or
// in a local vector are first copied here.  A bump allocator; a region is
// reused only after the event recorded behind the last staged copy completed.
struct Stager {
    char* buf = nullptr;
    size_t cap = 0, used = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};

}  // namespace

struct ikd_tree_s {
    Params prm{0.5f, 0.6f, 0.2f};
    hipStream_t st = nullptr;
    std::string err;
    // node pool
    long long cap = 0;
    Pool pool{};
    DevStatus* d_status = nullptr;  // header, root summary, rebuild list
    DevStatus* h_status = nullptr;  // its pinned mirror
    DevHdr* d_hdr = nullptr;        // &d_status->hdr
    DevHdr* h_hdr = nullptr;
    RootInfo* d_root = nullptr;  // &d_status->root
    RootInfo* h_root = nullptr;
    uint32_t next_id = 0;
    long long rebuilds = 0;
    bool compact_due = false;  // a sparse pool left to ikd_rebuild_deferred
    int max_depth = 0;
    long long pool_used = 4;
    int query_order = 0;  // 1: Morton-sort large query batches before the kNN launch
    int knn_stack = 0;    // minimum LDS stack entries per lane (0: the tree depth + 2)
    int knn_lanes = -1;   // lanes per query: -1 picks from the batch size; 0 persistent; else 1, 2, 4, 8, 16
    int small_build = 1;  // build subtrees of <= SMALL_MAX points with one workgroup each
    int fold_lanes = 0;   // lanes per voxel in the downsample fold: 0 picks by map size; else 1, 4, 8, 16
    int knn_stage = 0;    // one-lane kNN: tree levels staged in each workgroup's LDS (0: none)
    int knn_compact = 1;  // one-lane kNN: 4-byte LDS stack entries when the pool has <= 2^24 slots
    int record_removed = 0;     // 1: rebuilds record the deleted points they drop (acquire_removed_points)
    float4* removed = nullptr;  // that record (x, y, z, id bits), hdr->removed_count entries
    long long removed_cap = 0;
    hipEvent_t ev = nullptr;  // orders batch launches against this tree's stream
    ikd_counters_t ctr{};     // event counts since creation (ikd_get_counters)
    hipMemPool_t mp = nullptr;  // stream-ordered allocations (null: hipMalloc / hipFree)
    Stager stg;                 // pinned staging of small host-to-device copies
    // The host copies of the header and root summary match the device: the
    // last call that changed the tree ended with a refresh and succeeded.
    // Add_Points, the update path of every scan, skips its opening wait then.
    bool hdr_valid = false;
    // One call at a time per tree: the C ABI locks this for the whole call, so
    // reads from many host threads (FAST-LIO's OpenMP loop of Nearest_Search)
    // serialize onto the tree's stream instead of sharing its scratch slots.
    // Recursive: public entry points call one another.
    std::recursive_mutex mu;
    // scratch
    DevBuf b[96];
};

// A multi-tree launch context: its own stream, scratch and error slot.  A
// launch waits for every tree it reads (their streams' pending work) and each
// of those trees' later work waits for the launch.
struct ikd_batch_s {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    std::string err;
    int knn_lanes = -1;
    int knn_stack = 0;
    int knn_compact = 1;
    int* h_flag = nullptr;  // pinned mirror of the overflow flags (the batch's, the scratch one)
    char* na_scratch = nullptr;  // this device's allocation-free scratch (ikd_batch_nearest_search_noalloc)
    int* na_flag = nullptr;      // its overflow flag, once the entry point ran on this batch
    ikd_counters_t ctr{};   // only scratch_grows and host_waits move
    hipMemPool_t mp = nullptr;
    Stager stg;
    std::mutex mu;
    DevBuf b[16];  // BatchBuf slots
};

namespace {

enum Buf {
    B_TMP = 0, B_PTS, B_PTS2, B_SEGOF, B_KEYA, B_KEYB, B_VALA, B_PERM0, B_PERM1 = B_PERM0 + 3, B_SID = B_PERM1 + 3,
    B_SID2, B_SIDE, B_SEG0, B_SEG1, B_SEGD0, B_SEGD1, B_SEGAX, B_SEGMID, B_SEGCH, B_SEGFL, B_SEGSC, B_POSFL, B_POSSC,
    B_SEEDS, B_DIRTY, B_VIOL, B_TOP, B_INFO, B_FRONT0, B_FRONT1, B_EKEY0, B_EKEY1, B_RECS, B_GFLAG, B_GSCAN, B_ATTACH,
    B_PFLAG, B_PSCAN, B_ISEG, B_ISEGD, B_IO0, B_IO1, B_IO2, B_IO3, B_IO4, B_VSTART, B_DELBOX, B_DELFL, B_INSPT,
    B_INSFL, B_CBOX, B_CPTS, B_MISC, B_MORTK, B_MORTV, B_MORTK2, B_MORTV2, B_CNT64, B_OFF64, B_ADDX, B_ADDI, B_PNEXT, B_SPAIRS, B_SPSCAN,
    B_DEFER, B_DINFO, B_DBEFORE, B_DKEEP, B_PTS3, B_PEND, B_START, B_INV0, B_LASTP, B_PSEED, B_DBELOW, B_MAXDEL,
    B_SLISTS, B_ENUMROOT, B_ENUMGRP, B_TOPT, B_TOPI, B_TCNT64, B_TOFF64, B_RTASKS, B_ICNT64, B_EMPTYL, B_COUNT
};
static_assert(B_COUNT <= 96, "scratch table");

#define HIPCHK(x)                                                   \
    do {                                                            \
        hipError_t e_ = (x);                                        \
        if (e_ != hipSuccess) {                                     \
            t->err = std::string(#x) + ": " + hipGetErrorString(e_); \
            return IKD_ERR_HIP;                                     \
        }                                                           \
    } while (0)
#define CHK(x)                  \
    do {                        \
        int r_ = (x);           \
        if (r_ != IKD_OK) return r_; \
    } while (0)

// The library's device memory pool on device `dev`: stream-ordered
// allocation, so growing a buffer never waits for the device (hipFree waits
// for every stream of the device, which stalls every other tree's work
// queued behind it).  Freed memory stays in the pool for reuse.  Null when
// the runtime offers no memory pools.
hipMemPool_t library_pool(int dev) {
    static std::mutex mu;
    static hipMemPool_t pools[64];
    static bool tried[64];
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> g(mu);
    if (!tried[dev]) {
        tried[dev] = true;
        hipMemPoolProps props{};
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t mp = nullptr;
        if (hipMemPoolCreate(&mp, &props) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            if (hipMemPoolSetAttribute(mp, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess) {
                pools[dev] = mp;
            } else {
                hipMemPoolDestroy(mp);
            }
        }
        (void)hipGetLastError();
    }
    return pools[dev];
}

// Device memory ordered on the owner's stream: usable by work queued after
// the call on that stream (other streams order themselves with events).
template <class Owner>
hipError_t dev_alloc(Owner* t, void** p, size_t bytes) {
    if (t->mp) return hipMallocFromPoolAsync(p, bytes, t->mp, t->st);
    return hipMalloc(p, bytes);
}
// Frees after the work already queued on the owner's stream.
template <class Owner>
void dev_free(Owner* t, void* p) {
    if (!p) return;
    if (t->mp) {
        hipFreeAsync(p, t->st);
        return;
    }
    hipStreamSynchronize(t->st);  // queued kernels may still read the buffer
    hipFree(p);
}

template <class T, class Owner>
int ensure(Owner* t, int slot, long long count, T** out) {
    size_t need = (size_t)std::max<long long>(count, 1) * sizeof(T);
    DevBuf& d = t->b[slot];
    if (d.bytes < need) {
        dev_free(t, d.p);
        d.p = nullptr;
        size_t nb = std::max(need, 2 * d.bytes);
        t->ctr.scratch_grows += 1;
        if (dev_alloc(t, &d.p, nb) != hipSuccess) {
            d.p = nullptr;
            d.bytes = 0;
            t->err = "device allocation of scratch failed";
            return IKD_ERR_NOMEM;
        }
        d.bytes = nb;
    }
    *out = reinterpret_cast<T*>(d.p);
    return IKD_OK;
}

// Host waits for the tree's stream; counted because each one is a round trip
// on the latency path of an update.
hipError_t wait_stream(ikd_tree_t t) {
    t->ctr.host_waits += 1;
    const hipError_t e = hipStreamSynchronize(t->st);
    if (e == hipSuccess) {
        // every staged copy has run: the staging buffer is free again
        t->stg.used = 0;
        t->stg.pending = false;
    }
    return e;
}

// Copies `bytes` of host memory to device `dst` on the owner's stream through
// the owner's pinned staging buffer, so `src` may go away as soon as this
// returns.
template <class Owner>
int stage_copy(Owner* t, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return IKD_OK;
    Stager& g = t->stg;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (g.used + need > g.cap) {
        // every staged region is about to be reused: wait for the last copy
        if (g.pending) HIPCHK(hipEventSynchronize(g.ev));
        g.pending = false;
        g.used = 0;
        if (need > g.cap) {
            if (g.buf) HIPCHK(hipHostFree(g.buf));
            g.buf = nullptr;
            g.cap = 0;
            const size_t nc = std::max<size_t>(need, 64 << 10);
            HIPCHK(hipHostMalloc((void**)&g.buf, nc, hipHostMallocDefault));
            g.cap = nc;
        }
    }
    if (!g.ev) HIPCHK(hipEventCreateWithFlags(&g.ev, hipEventDisableTiming));
    char* p = g.buf + g.used;
    std::memcpy(p, src, bytes);
    g.used += need;
    HIPCHK(hipMemcpyAsync(dst, p, bytes, hipMemcpyHostToDevice, t->st));
    HIPCHK(hipEventRecord(g.ev, t->st));
    g.pending = true;
    return IKD_OK;
}

template <class Owner>
void stage_release(Owner* t) {
    Stager& g = t->stg;
    if (g.pending) hipEventSynchronize(g.ev);
    if (g.ev) hipEventDestroy(g.ev);
    if (g.buf) hipHostFree(g.buf);
    g = Stager{};
}

// A count the host read back from the device (and sizes launches or host
// loops by) must lie in [lo, hi]; anything else means a kernel misbehaved,
// and the call stops with IKD_ERR_INTERNAL instead of launching with it.
template <class Owner>
int check_count(Owner* t, const char* what, long long v, long long lo, long long hi) {
    if (v >= lo && v <= hi) return IKD_OK;
    char msg[200];
    snprintf(msg, sizeof msg, "device count out of range: %s = %lld, expected [%lld, %lld]", what, v, lo, hi);
    t->err = msg;
    return IKD_ERR_INTERNAL;
}

// Copies the first `bytes` of the status block back and waits.
int read_status(ikd_tree_t t, size_t bytes) {
    HIPCHK(hipMemcpyAsync(t->h_status, t->d_status, bytes, hipMemcpyDeviceToHost, t->st));
    HIPCHK(wait_stream(t));
    t->pool_used = 2ll * t->h_hdr->pair_used;
    t->max_depth = t->h_hdr->max_depth;
    if (const int o = t->h_hdr->overflow) {
        // each bit names the kernel family whose bounded buffer ran out
        t->err = (o & 32)   ? "enumerate: the walk disagrees with the subtree counts or outgrew its frontier"
                 : (o & 4)  ? "Add_Points: a downsample walk outgrew its stack"
                 : (o & 16) ? "update: a walk outgrew its stack or buffer"
                            : "device traversal stack overflow";
        return IKD_ERR_INTERNAL;
    }
    return IKD_OK;
}
int sync_hdr(ikd_tree_t t) { return read_status(t, sizeof(DevHdr)); }

int read_root(ikd_tree_t t) {
    hipLaunchKernelGGL(k_root_info, dim3(1), dim3(1), 0, t->st, t->pool, t->d_root);
    HIPCHK(hipMemcpyAsync(t->h_root, t->d_root, sizeof(RootInfo), hipMemcpyDeviceToHost, t->st));
    HIPCHK(wait_stream(t));
    return IKD_OK;
}
// Header and root summary in one copy and one host wait.
int refresh(ikd_tree_t t) {
    hipLaunchKernelGGL(k_root_info, dim3(1), dim3(1), 0, t->st, t->pool, t->d_root);
    return read_status(t, STATUS_SHORT);
}
bool has_root(ikd_tree_t t) { return (t->h_root->child.z & (int)LK_HASL) != 0; }
long long live_count(ikd_tree_t t) {
    if (!has_root(t)) return 0;
    return (long long)t->h_root->cnt.x - t->h_root->cnt.y;
}

// rocPRIM temporary storage lives in scratch slot 0 of its owner (tree or batch)
template <class T, class Owner>
int prim_tmp(Owner* t, size_t bytes, void** out) {
    (void)sizeof(T);
    char* p;
    CHK(ensure(t, 0, (long long)bytes, &p));
    *out = p;
    return IKD_OK;
}

int sort_pairs(ikd_tree_t t, const unsigned long long* kin, unsigned long long* kout, const uint32_t* vin,
               uint32_t* vout, int n, int end_bit) {
    if (n <= 0) return IKD_OK;
    size_t bytes = 0;
    HIPCHK(prim_sort_pairs_u64(nullptr, bytes, kin, kout, vin, vout, n, 0, end_bit, t->st));
    void* tmp;
    CHK(prim_tmp<char>(t, bytes, &tmp));
    HIPCHK(prim_sort_pairs_u64(tmp, bytes, kin, kout, vin, vout, n, 0, end_bit, t->st));
    return IKD_OK;
}
int sort_keys(ikd_tree_t t, const unsigned long long* kin, unsigned long long* kout, int n, int end_bit) {
    if (n <= 0) return IKD_OK;
    size_t bytes = 0;
    HIPCHK(prim_sort_keys_u64(nullptr, bytes, kin, kout, n, 0, end_bit, t->st));
    void* tmp;
    CHK(prim_tmp<char>(t, bytes, &tmp));
    HIPCHK(prim_sort_keys_u64(tmp, bytes, kin, kout, n, 0, end_bit, t->st));
    return IKD_OK;
}
// Exclusive scans of up to SCAN_ONE_MAX values run as one workgroup in one
// launch: each of the 1024 threads sums a run of consecutive values, the run
// totals are scanned across the wavefronts, then each thread rewrites its run.
// (The device-wide scan costs two launches and a temporary-size query, which
// dominates at these sizes; from a few tiles of values on, its many
// workgroups finish first, measured on the serving workload.)  in == out is
// allowed.
constexpr int SCAN_ONE_MAX = 8192;
__device__ __forceinline__ int scan_shfl(int x, int src) { return __shfl(x, src); }
__device__ __forceinline__ long long scan_shfl(long long x, int src) {
    const unsigned lo = (unsigned)__shfl((int)(unsigned)x, src);
    const unsigned hi = (unsigned)__shfl((int)(unsigned)((unsigned long long)x >> 32), src);
    return (long long)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ unsigned long long scan_shfl(unsigned long long x, int src) {
    return (unsigned long long)scan_shfl((long long)x, src);
}
template <class T>
__global__ __launch_bounds__(1024) void k_scan_one(const T* in, T* out, int n) {
    __shared__ T wave_tot[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int chunk = (n + 1023) >> 10;
    const int b = min(t * chunk, n), e = min(b + chunk, n);
    T sum = 0;
#pragma unroll 8
    for (int i = b; i < e; ++i) sum += in[i];
    T inc = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T o = scan_shfl(inc, (lane - d) & 63);  // lanes below d ignore what they read
        if (lane >= d) inc += o;
    }
    if (lane == 63) wave_tot[w] = inc;
    __syncthreads();
    T run = inc - sum;
    for (int i = 0; i < w; ++i) run += wave_tot[i];
#pragma unroll 8
    for (int i = b; i < e; ++i) {
        const T v = in[i];
        out[i] = run;
        run += v;
    }
}

int exscan_u64(ikd_tree_t t, const unsigned long long* in, unsigned long long* out, int n) {
    if (n <= 0) return IKD_OK;
    if (n <= SCAN_ONE_MAX) {
        hipLaunchKernelGGL(k_scan_one<unsigned long long>, dim3(1), dim3(1024), 0, t->st, in, out, n);
        return IKD_OK;
    }
    size_t bytes = 0;
    HIPCHK(prim_exscan_u64(nullptr, bytes, in, out, n, t->st));
    void* tmp;
    CHK(prim_tmp<char>(t, bytes, &tmp));
    HIPCHK(prim_exscan_u64(tmp, bytes, in, out, n, t->st));
    return IKD_OK;
}
int exscan_i32(ikd_tree_t t, const int* in, int* out, int n) {
    if (n <= 0) return IKD_OK;
    if (n <= SCAN_ONE_MAX) {
        hipLaunchKernelGGL(k_scan_one<int>, dim3(1), dim3(1024), 0, t->st, in, out, n);
        return IKD_OK;
    }
    size_t bytes = 0;
    HIPCHK(prim_exscan_i32(nullptr, bytes, in, out, n, t->st));
    void* tmp;
    CHK(prim_tmp<char>(t, bytes, &tmp));
    HIPCHK(prim_exscan_i32(tmp, bytes, in, out, n, t->st));
    return IKD_OK;
}
template <class Owner>
int exscan_i64(Owner* t, const long long* in, long long* out, int n) {
    if (n <= 0) return IKD_OK;
    if (n <= SCAN_ONE_MAX) {
        hipLaunchKernelGGL(k_scan_one<long long>, dim3(1), dim3(1024), 0, t->st, in, out, n);
        return IKD_OK;
    }
    size_t bytes = 0;
    HIPCHK(prim_exscan_i64(nullptr, bytes, in, out, n, t->st));
    void* tmp;
    CHK(prim_tmp<char>(t, bytes, &tmp));
    HIPCHK(prim_exscan_i64(tmp, bytes, in, out, n, t->st));
    return IKD_OK;
}

// ------------------------------------------------------------- node pool
int alloc_pool(ikd_tree_t t, long long cap, Pool* p) {
    if (dev_alloc(t, (void**)&p->hot, cap * sizeof(NodeHot)) != hipSuccess) return IKD_ERR_NOMEM;
    if (dev_alloc(t, (void**)&p->ids, cap * sizeof(uint32_t)) != hipSuccess) return IKD_ERR_NOMEM;
    if (dev_alloc(t, (void**)&p->par, cap * sizeof(int32_t)) != hipSuccess) return IKD_ERR_NOMEM;
    if (dev_alloc(t, (void**)&p->cnt, cap * sizeof(int4)) != hipSuccess) return IKD_ERR_NOMEM;
    if (dev_alloc(t, (void**)&p->meta, cap * sizeof(uint32_t)) != hipSuccess) return IKD_ERR_NOMEM;
    return IKD_OK;
}
void free_pool(ikd_tree_t t, Pool* p) {
    dev_free(t, p->hot);
    dev_free(t, p->ids);
    dev_free(t, p->par);
    dev_free(t, p->cnt);
    dev_free(t, p->meta);
    *p = Pool{};
}

// Make room for `extra` more nodes (pool_used is exact after the last sync).
int reserve_nodes(ikd_tree_t t, long long extra) {
    long long need = t->pool_used + extra + 8;
    if (need <= t->cap) return IKD_OK;
    if (need >= (1ll << 28)) {
        t->err = "node pool would exceed 2^28 nodes";
        return IKD_ERR_NOMEM;
    }
    long long ncap = std::max(need + need / 2, 2 * t->cap);
    ncap = std::min(ncap, 1ll << 28);
    if (t->cap > 0) t->ctr.pool_grows += 1;
    Pool np{};
    if (alloc_pool(t, ncap, &np) != IKD_OK) {
        (void)hipGetLastError();
        free_pool(t, &np);
        t->err = "node pool allocation failed";
        return IKD_ERR_NOMEM;
    }
    long long used = t->pool_used;
    if (t->cap > 0) {
        HIPCHK(hipMemcpyAsync(np.hot, t->pool.hot, used * sizeof(NodeHot), hipMemcpyDeviceToDevice, t->st));
        HIPCHK(hipMemcpyAsync(np.ids, t->pool.ids, used * sizeof(uint32_t), hipMemcpyDeviceToDevice, t->st));
        HIPCHK(hipMemcpyAsync(np.par, t->pool.par, used * sizeof(int32_t), hipMemcpyDeviceToDevice, t->st));
        HIPCHK(hipMemcpyAsync(np.cnt, t->pool.cnt, used * sizeof(int4), hipMemcpyDeviceToDevice, t->st));
        HIPCHK(hipMemcpyAsync(np.meta, t->pool.meta, used * sizeof(uint32_t), hipMemcpyDeviceToDevice, t->st));
        free_pool(t, &t->pool);  // after the copies
    }
    t->pool = np;
    t->cap = ncap;
    return IKD_OK;
}

int reset_pool(ikd_tree_t t) {
    t->pool_used = 4;
    t->max_depth = 0;
    CHK(reserve_nodes(t, 4));
    hipLaunchKernelGGL(k_init_pool, dim3(1), dim3(1), 0, t->st, t->pool, t->d_hdr);
    HIPCHK(hipGetLastError());
    return IKD_OK;
}

// -------------------------------------------------------------- the build
// Builds S subtrees at once.  pts: m points grouped by segment (segment g at
// positions [seg[g].x, seg[g].y]); seg_of: segment of each point; seg /
// seg_depth: device arrays (l, r, slot node, slot parent) + depth, S entries.
int run_build(ikd_tree_t t, const float4* pts, const int* seg_of, int m, const int4* d_seg, const int* d_seg_depth,
              int S, int maxlen) {
    if (m <= 0 || S <= 0) return IKD_OK;
    CHK(reserve_nodes(t, 2ll * m + 4));
    hipStream_t st = t->st;
    if (maxlen <= SMALL_MAX && t->small_build) {
        // every subtree fits one workgroup: one launch builds them all
        int *pairs, *pscan, *lists;
        CHK(ensure(t, B_SPAIRS, S, &pairs));
        CHK(ensure(t, B_SPSCAN, S, &pscan));
        CHK(ensure(t, B_SLISTS, 2ll * S, &lists));
        launch_small_pairs(d_seg, S, nullptr, pairs, t->d_hdr, st);
        CHK(exscan_i32(t, pairs, pscan, S));
        launch_build_small(pts, d_seg, d_seg_depth, pscan, S, nullptr, nullptr, SmallLists{}, lists, t->pool,
                           t->d_hdr, st);
        launch_small_commit(pairs, pscan, S, t->d_hdr, st);
        HIPCHK(hipGetLastError());
        return IKD_OK;
    }
    unsigned long long *keya, *keyb;
    uint32_t* vala;
    CHK(ensure(t, B_KEYA, m, &keya));
    CHK(ensure(t, B_KEYB, m, &keyb));
    CHK(ensure(t, B_VALA, m, &vala));
    uint32_t* perm[2][3];
    for (int a = 0; a < 3; ++a) {
        CHK(ensure(t, B_PERM0 + a, m, &perm[0][a]));
        CHK(ensure(t, B_PERM1 + a, m, &perm[1][a]));
    }
    int *sid[2];
    CHK(ensure(t, B_SID, m, &sid[0]));
    CHK(ensure(t, B_SID2, m, &sid[1]));
    uint8_t* side;
    CHK(ensure(t, B_SIDE, m, &side));
    long long segcap = (long long)m + S + 1;
    int4* seg[2];
    int* segd[2];
    CHK(ensure(t, B_SEG0, segcap, &seg[0]));
    CHK(ensure(t, B_SEG1, segcap, &seg[1]));
    CHK(ensure(t, B_SEGD0, segcap, &segd[0]));
    CHK(ensure(t, B_SEGD1, segcap, &segd[1]));
    int *segax, *segmid;
    int2* segch;
    unsigned long long *segfl, *segsc, *posfl, *possc;
    CHK(ensure(t, B_SEGAX, segcap, &segax));
    CHK(ensure(t, B_SEGMID, segcap, &segmid));
    CHK(ensure(t, B_SEGCH, segcap, &segch));
    CHK(ensure(t, B_SEGFL, segcap, &segfl));
    CHK(ensure(t, B_SEGSC, segcap, &segsc));
    CHK(ensure(t, B_POSFL, 3ll * m, &posfl));
    CHK(ensure(t, B_POSSC, 3ll * m, &possc));

    int end_bit = 32 + bits_for(S);
    for (int a = 0; a < 3; ++a) {
        launch_build_keys(pts, seg_of, m, a, keya, vala, st);
        CHK(sort_pairs(t, keya, keyb, vala, perm[0][a], m, end_bit));
        if (a == 0) launch_build_init_sid(keyb, m, sid[0], st);
    }
    HIPCHK(hipMemcpyAsync(seg[0], d_seg, S * sizeof(int4), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(segd[0], d_seg_depth, S * sizeof(int), hipMemcpyDeviceToDevice, st));
    CHK(stage_copy(t, &t->d_hdr->seg_count[0], &S, sizeof(int)));

    int levels = bits_for(maxlen);  // floor(log2(maxlen)) + 1
    // A segment of x points leaves children of at most x/2: from level Ls on
    // every segment fits one workgroup, and one launch builds all of them.
    int Ls = levels;
    if (t->small_build) {
        Ls = 0;
        while ((maxlen >> Ls) > SMALL_MAX) ++Ls;
    }
    for (int L = 0; L < std::min(levels, Ls); ++L) {
        int cur = L & 1;
        long long bound = std::min<long long>((long long)S << std::min(L, 30), m);
        BuildLevel BL;
        BL.pts = pts;
        for (int a = 0; a < 3; ++a) {
            BL.perm[a] = perm[cur][a];
            BL.perm_next[a] = perm[cur ^ 1][a];
        }
        BL.sid = sid[cur];
        BL.sid_next = sid[cur ^ 1];
        BL.side = side;
        BL.seg = seg[cur];
        BL.seg_depth = segd[cur];
        BL.seg_next = seg[cur ^ 1];
        BL.seg_depth_next = segd[cur ^ 1];
        BL.seg_axis = segax;
        BL.seg_mid = segmid;
        BL.seg_child = segch;
        BL.seg_flags = segfl;
        BL.seg_scan = segsc;
        BL.pos_flags = posfl;
        BL.pos_scan = possc;
        BL.m = m;
        BL.cur = cur;
        launch_build_seg_a(BL, (int)bound, t->d_hdr, st);
        CHK(exscan_u64(t, segfl, segsc, (int)bound));
        launch_build_seg_b(BL, (int)bound, t->pool, t->d_hdr, st);
        launch_build_side(BL, t->d_hdr, st);
        if (L + 1 < levels) {
            launch_build_flags(BL, st);
            CHK(exscan_u64(t, posfl, possc, 3 * m));
            launch_build_scatter(BL, st);
        }
    }
    if (Ls < levels) {
        // level Ls: its segments (count on the device) and its x-ordered list
        const int cur = Ls & 1;
        const int bound = (int)std::min<long long>((long long)S << std::min(Ls, 30), m);
        float4* gpts;
        int *pairs, *pscan, *lists;
        CHK(ensure(t, B_PTS3, m, &gpts));
        CHK(ensure(t, B_SPAIRS, bound, &pairs));
        CHK(ensure(t, B_SPSCAN, bound, &pscan));
        CHK(ensure(t, B_SLISTS, 2ll * bound, &lists));
        int* inv0;
        CHK(ensure(t, B_INV0, m, &inv0));
        launch_gather_list(pts, perm[cur][0], sid[cur], m, gpts, inv0, st);
        const int* count_dev = &t->d_hdr->seg_count[cur];
        launch_small_pairs(seg[cur], bound, count_dev, pairs, t->d_hdr, st);
        CHK(exscan_i32(t, pairs, pscan, bound));
        // the level-synchronous lists are already sorted per segment
        SmallLists pre;
        pre.list1 = perm[cur][1];
        pre.list2 = perm[cur][2];
        pre.inv0 = inv0;
        launch_build_small(gpts, seg[cur], segd[cur], pscan, bound, count_dev, perm[cur][0], pre, lists, t->pool,
                           t->d_hdr, st);
        launch_small_commit(pairs, pscan, bound, t->d_hdr, st);
    }
    HIPCHK(hipGetLastError());
    return IKD_OK;
}

// ---------------------------------------------------------- enumerate live
// Live points of the subtrees rooted at roots[] (host list), grouped by root
// in root order, each group in the reference's flatten order (own, left,
// right).  live[i] is root i's live point count (its group's size).  Returns
// the points in B_PTS2 (float4) and their root index in B_SEGOF.
// Subtree nodes one workgroup enumerates: a level costs it about what a
// launch costs, so it only pays while every level fits one pass of 1024 threads.
constexpr long long ENUM_BLOCK_MAX = 1024;
// Grouped walk: nodes per workgroup, and the largest subtree it takes (a
// larger one walks one level per launch over the whole grid instead).
constexpr long long ENUM_GROUP_NODES = 2048;
constexpr long long ENUM_GROUP_ROOT_MAX = 16384;
// Range search: subtrees a query covers whole with at most this many nodes are
// written by one lane each after the walks (the walks stop there).
constexpr int RANGE_TASK_NODES = 256;

// Room for `extra` more records beyond the current count (fresh in h_hdr):
// the record survives the growth, unlike scratch.
int reserve_removed(ikd_tree_t t, long long extra) {
    const long long have = t->h_hdr->removed_count;
    const long long need = have + extra;
    if (need <= t->removed_cap) return IKD_OK;
    if (need >= (1ll << 31)) {
        t->err = "removed-point record would exceed 2^31 entries";
        return IKD_ERR_NOMEM;
    }
    const long long ncap = std::max(need + need / 2, 2 * t->removed_cap);
    float4* np = nullptr;
    if (dev_alloc(t, (void**)&np, (size_t)ncap * sizeof(float4)) != hipSuccess) {
        (void)hipGetLastError();
        t->err = "device allocation of the removed-point record failed";
        return IKD_ERR_NOMEM;
    }
    if (have > 0) HIPCHK(hipMemcpyAsync(np, t->removed, have * sizeof(float4), hipMemcpyDeviceToDevice, t->st));
    dev_free(t, t->removed);
    t->removed = np;
    t->removed_cap = ncap;
    return IKD_OK;
}

// min_depth: the smallest depth among the roots (their subtrees end by the
// tree's deepest level, so max_depth - min_depth + 1 levels cover them).
// record: a rebuild's flatten (DELETE_POINTS_REC, cpp:638); plain flattens
// (NOT_RECORD) pass false.
// nodes (nullable): each root's subtree node count; when every subtree is
// small, one launch walks them all (a workgroup per group of subtrees).
// Queues the walk without waiting: a count that disagrees with the walk is
// reported by the next header read.
int enumerate_live(ikd_tree_t t, const std::vector<int>& roots, const std::vector<long long>& live,
                   long long total_nodes, long long total_live, int min_depth, bool record, float4** out_pts,
                   int** out_seg, const std::vector<int>* nodes = nullptr) {
    hipStream_t st = t->st;
    int R = (int)roots.size();
    long long cap = std::max<long long>(total_nodes, 1);
    // grouped walk: consecutive roots share a workgroup up to ENUM_GROUP_NODES nodes
    bool grouped = nodes != nullptr && R > 0;
    std::vector<int4> groups;
    if (grouped) {
        long long off = 0, gn = 0;
        int first = 0;
        for (int i = 0; i < R && grouped; ++i) {
            const long long ni = std::max<long long>((*nodes)[i], 1);
            if (ni > ENUM_GROUP_ROOT_MAX) grouped = false;
            if (i > first && gn + ni > ENUM_GROUP_NODES) {
                groups.push_back(make_int4(first, i - first, (int)off, (int)gn));
                off += gn;
                first = i;
                gn = 0;
            }
            gn += ni;
        }
        if (grouped) {
            groups.push_back(make_int4(first, R - first, (int)off, (int)gn));
            cap = std::max(cap, off + gn);
        }
    }
    int4 *f0, *f1;
    CHK(ensure(t, B_FRONT0, cap, &f0));
    CHK(ensure(t, B_FRONT1, cap, &f1));
    const int n = (int)total_live;
    float4* pts;
    int* segof;
    CHK(ensure(t, B_PTS2, std::max(n, 1), &pts));
    CHK(ensure(t, B_SEGOF, std::max(n, 1), &segof));
    std::vector<int4> h(R);
    long long base = 0;
    for (int i = 0; i < R; ++i) {
        h[i] = make_int4(roots[i], (int)base, i, 0);
        base += live[i];
    }
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, st, t->d_hdr, 2);
    const int levels = std::max(t->max_depth - std::max(min_depth, 0), 0) + 2;
    float4* rec = nullptr;
    int rec_cap = 0;
    if (record && t->record_removed) {
        // every node of the subtrees may be a deleted one
        CHK(reserve_removed(t, cap));
        rec = t->removed;
        rec_cap = (int)t->removed_cap;
    }
    if (grouped) {
        int4 *droots, *dgroups;
        CHK(ensure(t, B_ENUMROOT, R, &droots));
        CHK(ensure(t, B_ENUMGRP, (long long)groups.size(), &dgroups));
        CHK(stage_copy(t, droots, h.data(), R * sizeof(int4)));
        CHK(stage_copy(t, dgroups, groups.data(), groups.size() * sizeof(int4)));
        launch_enum_groups(t->pool, droots, dgroups, (int)groups.size(), f0, f1, pts, segof, t->d_hdr, std::max(n, 1),
                           levels, rec, rec_cap, st);
    } else {
        CHK(stage_copy(t, f0, h.data(), R * sizeof(int4)));
        int init[3] = {R, 0, 0};
        CHK(stage_copy(t, t->d_hdr->front3, init, sizeof(init)));
        if (cap <= ENUM_BLOCK_MAX) {
            // small subtrees: one workgroup walks all levels
            launch_enum_preorder_block(t->pool, f0, f1, R, pts, segof, t->d_hdr, (int)cap, std::max(n, 1), levels,
                                       rec, rec_cap, st);
        } else {
            for (int L = 0; L < levels; ++L) {
                int4* fin = (L & 1) ? f1 : f0;
                int4* fout = (L & 1) ? f0 : f1;
                launch_enum_preorder(t->pool, fin, L, fout, pts, segof, t->d_hdr, (int)cap, std::max(n, 1), rec,
                                     rec_cap, st);
            }
        }
    }
    hipLaunchKernelGGL(k_enum_check, dim3(1), dim3(1), 0, st, t->d_hdr, n);
    HIPCHK(hipGetLastError());
    *out_pts = pts;
    *out_seg = segof;
    return IKD_OK;
}

// ------------------------------------------------- refit, criterion, rebuild
// Refit = mark the ancestors of every seed (one or more batches of seeds),
// then Update them bottom-up one depth level per launch, then list the
// topmost Criterion_Check violators.
int refit_begin(ikd_tree_t t) {
    long long cap = std::max<long long>(t->pool_used, 16);
    int *dirty, *viol, *top;
    CHK(ensure(t, B_DIRTY, cap, &dirty));
    CHK(ensure(t, B_VIOL, 2 * cap, &viol));
    CHK(ensure(t, B_TOP, 2 * cap, &top));
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, t->st, t->d_hdr, 1);
    return IKD_OK;
}
int refit_mark(ikd_tree_t t, const int* seeds, const int* nseeds_dev, int bound, int stop_at_tdel) {
    int *dirty, *pend;
    CHK(ensure(t, B_DIRTY, std::max<long long>(t->pool_used, 16), &dirty));
    CHK(ensure(t, B_PEND, std::max<long long>(t->cap, 16), &pend));  // indexed by node
    launch_mark_dirty(t->pool, seeds, nseeds_dev, bound, stop_at_tdel, dirty, pend, t->d_hdr, t->st);
    return IKD_OK;
}
int refit_finish(ikd_tree_t t) {
    long long cap = std::max<long long>(t->pool_used, 16);
    int *dirty, *viol, *top;
    CHK(ensure(t, B_DIRTY, cap, &dirty));
    CHK(ensure(t, B_VIOL, 2 * cap, &viol));
    CHK(ensure(t, B_TOP, 2 * cap, &top));
    int *defer, *pend, *start;
    CHK(ensure(t, B_DEFER, DEFER_CAP, &defer));
    CHK(ensure(t, B_PEND, std::max<long long>(t->cap, 16), &pend));
    CHK(ensure(t, B_START, cap, &start));
    launch_refit_up(t->pool, t->prm, dirty, pend, start, t->d_hdr, viol, defer, t->st);
    launch_find_top(t->pool, viol, (int)cap, t->d_hdr, top, t->st);
    launch_clear_viol(t->pool, viol, (int)cap, t->d_hdr, t->st);
    HIPCHK(hipGetLastError());
    return IKD_OK;
}
int refit(ikd_tree_t t, const int* seeds, int nseeds, int stop_at_tdel) {
    CHK(refit_begin(t));
    CHK(refit_mark(t, seeds, nullptr, nseeds, stop_at_tdel));
    return refit_finish(t);
}

// Rebuild the listed subtrees (info = node, live points, depth, parent;
// before = nodes in the subtree) from their live points.
int rebuild_listed(ikd_tree_t t, const std::vector<int4>& hinfo, const std::vector<int>& hbefore) {
    hipStream_t st = t->st;
    const int ntop = (int)hinfo.size();
    std::vector<int> roots;
    std::vector<long long> lives;
    int min_depth = INT_MAX;
    std::vector<int4> segs;
    std::vector<int> segd;
    std::vector<int> nodes;
    long long total_nodes = 0, total_live = 0;
    int maxlen = 0;
    bool root_rebuilt = false;
    std::vector<int2> empties;  // (parent, child-slot bits) of all-deleted subtrees
    for (int i = 0; i < ntop; ++i) {
        int u = hinfo[i].x, live = hinfo[i].y, depth = hinfo[i].z, parent = hinfo[i].w;
        CHK(check_count(t, "rebuild: subtree root", u, ROOT, t->cap - 1));
        CHK(check_count(t, "rebuild: parent", parent, SENTINEL, t->cap - 1));
        CHK(check_count(t, "rebuild: subtree nodes", hbefore[i], 1, t->pool_used));
        CHK(check_count(t, "rebuild: live points", live, 0, hbefore[i]));
        CHK(check_count(t, "rebuild: depth", depth, 0, 4 * MAX_DEPTH));
        if (u == ROOT) root_rebuilt = true;
        if (live <= 0) {
            // Rebuild of an all-deleted subtree leaves an empty slot (BuildTree of
            // an empty range, cpp:574).
            empties.push_back(make_int2(parent, (int)(lk_has(u & 1) | lk_td(u & 1))));
            continue;
        }
        segs.push_back(make_int4((int)total_live, (int)(total_live + live - 1), u, parent));
        segd.push_back(depth);
        roots.push_back(u);
        lives.push_back(live);
        nodes.push_back(hbefore[i]);
        min_depth = std::min(min_depth, depth);
        total_nodes += hbefore[i];
        total_live += live;
        t->ctr.rebuilt_points += live;
        maxlen = std::max(maxlen, live);
    }
    if (!empties.empty()) {
        const int ne = (int)empties.size();
        int2* de;
        CHK(ensure(t, B_EMPTYL, ne, &de));
        CHK(stage_copy(t, de, empties.data(), ne * sizeof(int2)));
        hipLaunchKernelGGL(k_clear_links, dim3(nblk(ne)), dim3(BS), 0, st, t->pool.hot, de, ne);
    }
    if (!roots.empty()) {
        float4* pts;
        int* segof;
        CHK(enumerate_live(t, roots, lives, total_nodes, total_live, min_depth, true, &pts, &segof, &nodes));
        int4* dseg;
        int* dsegd;
        CHK(ensure(t, B_ISEG, (long long)segs.size(), &dseg));
        CHK(ensure(t, B_ISEGD, (long long)segs.size(), &dsegd));
        CHK(stage_copy(t, dseg, segs.data(), segs.size() * sizeof(int4)));
        CHK(stage_copy(t, dsegd, segd.data(), segd.size() * sizeof(int)));
        if (root_rebuilt) {
            int zero = 0;
            CHK(stage_copy(t, &t->d_hdr->max_depth, &zero, sizeof(int)));
        }
        CHK(run_build(t, pts, segof, (int)total_live, dseg, dsegd, (int)segs.size(), maxlen));
        if (root_rebuilt) launch_root_alpha_fresh(t->pool, t->d_hdr, st);
    }
    return IKD_OK;
}

// Returns with the header and root summary fresh (refresh()).
int rebuild_tops(ikd_tree_t t) {
    hipStream_t st = t->st;
    for (int iter = 0; iter < 8; ++iter) {
        // the header, the root summary and the first subtrees to rebuild in one wait
        // (the refit's list; after a pass of rebuilds the list is empty)
        const int* top = iter == 0 ? reinterpret_cast<const int*>(t->b[B_TOP].p) : nullptr;
        hipLaunchKernelGGL(k_root_info, dim3(1), dim3(1), 0, st, t->pool, t->d_root);
        if (top) hipLaunchKernelGGL(k_top_info, dim3(1), dim3(256), 0, st, t->pool, top, t->d_status);
        CHK(read_status(t, top ? sizeof(DevStatus) : STATUS_SHORT));
        int ntop = t->h_hdr->top_count;
        CHK(check_count(t, "rebuild: subtrees", ntop, 0, t->pool_used));
        if (ntop == 0) return IKD_OK;
        std::vector<int4> hinfo(ntop);
        std::vector<int> hbefore(ntop);
        if (top && ntop <= TOPINFO_CAP) {
            std::memcpy(hinfo.data(), t->h_status->top_info, ntop * sizeof(int4));
            std::memcpy(hbefore.data(), t->h_status->top_before, ntop * sizeof(int));
        } else {
            int* topb;
            int4* info;
            int* before;
            CHK(ensure(t, B_TOP, 2 * ntop, &topb));
            CHK(ensure(t, B_INFO, ntop, &info));
            CHK(ensure(t, B_SEEDS, ntop, &before));
            hipLaunchKernelGGL(k_node_info, dim3(nblk(ntop)), dim3(BS), 0, st, t->pool, topb, ntop, info, before);
            HIPCHK(hipMemcpyAsync(hinfo.data(), info, ntop * sizeof(int4), hipMemcpyDeviceToHost, st));
            HIPCHK(hipMemcpyAsync(hbefore.data(), before, ntop * sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHK(wait_stream(t));
        }
        CHK(rebuild_listed(t, hinfo, hbefore));
        t->rebuilds += ntop;
        // The refit already handed every ancestor the post-rebuild counts, so
        // the rebuilt subtrees leave nothing to re-check.
        hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, st, t->d_hdr, 1);
    }
    t->err = "rebuild did not converge";
    return IKD_ERR_INTERNAL;
}

// Rebuild the whole tree from its live points when the pool holds mostly
// garbage or the tree got deeper than the traversal stacks allow.
// fresh: the header and root summary were refreshed after the last change.
// Returns with both fresh.
bool pool_sparse(ikd_tree_t t, long long live) { return t->pool_used > 4 * live + (1ll << 22); }

int compact_if_needed(ikd_tree_t t, bool force, bool fresh = false) {
    if (!fresh) CHK(refresh(t));
    long long live = live_count(t);
    bool deep = t->max_depth + 2 > MAX_DEPTH;
    bool sparse = pool_sparse(t, live);
    if (!force && !deep && !sparse) return IKD_OK;
    // With deferred rebuilds on, a sparse pool is reclaimed by
    // ikd_rebuild_deferred (off the update's latency) until it holds twice
    // the garbage; a tree too deep for the traversal stacks is rebuilt now.
    if (!force && !deep && t->prm.defer_points > 0 && t->pool_used <= 8 * live + (1ll << 23)) {
        t->compact_due = true;
        return IKD_OK;
    }
    t->compact_due = false;
    t->ctr.compactions += 1;
    if (!has_root(t)) {
        CHK(reset_pool(t));
        return refresh(t);
    }
    float4* pts;
    int* segof;
    std::vector<int> roots = {ROOT};
    std::vector<long long> lives = {live};
    CHK(enumerate_live(t, roots, lives, t->h_root->cnt.x, live, 0, true, &pts, &segof));
    // move the points out of the scratch slot the build reuses
    float4* keep;
    CHK(ensure(t, B_PTS, live, &keep));
    HIPCHK(hipMemcpyAsync(keep, pts, live * sizeof(float4), hipMemcpyDeviceToDevice, t->st));
    CHK(reset_pool(t));
    if (live == 0) return refresh(t);
    int* segof0;
    CHK(ensure(t, B_SEGOF, live, &segof0));
    launch_fill_i32(segof0, 0, live, t->st);
    int4 s0 = make_int4(0, (int)live - 1, ROOT, SENTINEL);
    int d0 = 0;
    int4* dseg;
    int* dsegd;
    CHK(ensure(t, B_ISEG, 1, &dseg));
    CHK(ensure(t, B_ISEGD, 1, &dsegd));
    CHK(stage_copy(t, dseg, &s0, sizeof(s0)));
    CHK(stage_copy(t, dsegd, &d0, sizeof(d0)));
    CHK(run_build(t, keep, segof0, (int)live, dseg, dsegd, 1, (int)live));
    hipLaunchKernelGGL(k_link_bits, dim3(1), dim3(1), 0, t->st, t->pool.hot, SENTINEL, LK_HASL, LK_TDL);
    launch_root_alpha_fresh(t->pool, t->d_hdr, t->st);
    t->rebuilds += 1;
    return refresh(t);
}

// ---------------------------------------------------------------- insert
// Inserts m packed points (device, float4 x,y,z,id), then refits the paths
// to the attach points together with extra_seeds (nodes whose own labels the
// caller changed; their count is read from *n_extra_dev on the device).
int insert_points(ikd_tree_t t, const float4* pts, int m, const int* extra_seeds = nullptr,
                  const int* n_extra_dev = nullptr, int extra_bound = 0) {
    if (m <= 0) {
        if (extra_bound <= 0) return refresh(t);  // callers rely on a fresh header and root summary
        CHK(refit_begin(t));
        CHK(refit_mark(t, extra_seeds, n_extra_dev, extra_bound, 0));
        CHK(refit_finish(t));
        return rebuild_tops(t);
    }
    hipStream_t st = t->st;
    // a root made by this insert is a new node (its statistics are set below)
    const bool new_root = !has_root(t);
    CHK(reserve_nodes(t, 4ll * m + 16));
    unsigned long long *ka, *kb;
    uint32_t *va, *vb;
    CHK(ensure(t, B_IO0, m, &ka));
    CHK(ensure(t, B_IO1, m, &kb));
    CHK(ensure(t, B_IO2, m, &va));
    CHK(ensure(t, B_IO3, m, &vb));
    launch_insert_descend(t->pool, pts, m, ka, va, t->d_hdr, st);
    CHK(sort_pairs(t, ka, kb, va, vb, m, 1 + bits_for(t->pool_used)));
    int *gflag, *gscan, *attach, *pflag, *pscan, *segof, *segd;
    int4* seg;
    CHK(ensure(t, B_GFLAG, m, &gflag));
    CHK(ensure(t, B_GSCAN, m, &gscan));
    CHK(ensure(t, B_ATTACH, m, &attach));
    CHK(ensure(t, B_PFLAG, m, &pflag));
    CHK(ensure(t, B_PSCAN, m, &pscan));
    CHK(ensure(t, B_SEGOF, m, &segof));
    CHK(ensure(t, B_ISEG, m, &seg));
    CHK(ensure(t, B_ISEGD, m, &segd));
    launch_insert_groups(kb, m, gflag, st);
    CHK(exscan_i32(t, gflag, gscan, m));
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, st, t->d_hdr, 16);
    launch_insert_segments2(t->pool, kb, gscan, gflag, m, seg, segd, segof, attach, pflag, t->d_hdr, st);
    CHK(exscan_i32(t, pflag, pscan, m));
    launch_insert_pairs(t->pool, attach, pflag, pscan, t->d_hdr, m, seg, st);
    int* misc;
    CHK(ensure(t, B_MISC, 4, &misc));
    launch_fill_i32(misc, 0, 1, st);
    hipLaunchKernelGGL(k_group_maxlen, dim3(nblk(m)), dim3(BS), 0, st, seg, t->d_hdr, m, misc);
    float4* spts;
    CHK(ensure(t, B_PTS2, m, &spts));
    hipLaunchKernelGGL(k_gather_pts, dim3(nblk(m)), dim3(BS), 0, st, pts, vb, m, spts);
    int maxlen = 0;
    HIPCHK(hipMemcpyAsync(&maxlen, misc, sizeof(int), hipMemcpyDeviceToHost, st));
    CHK(sync_hdr(t));
    int S = t->h_hdr->group_count;
    CHK(check_count(t, "insert: attach groups", S, 1, m));
    CHK(check_count(t, "insert: largest group", maxlen, 1, m));
    CHK(run_build(t, spts, segof, m, seg, segd, S, maxlen));
    launch_insert_link(t->pool, attach, seg, t->d_hdr, S, st);
    if (new_root) launch_root_alpha_fresh(t->pool, t->d_hdr, st);
    int* seeds;
    CHK(ensure(t, B_SEEDS, S, &seeds));
    hipLaunchKernelGGL(k_seeds_from_attach, dim3(nblk(S)), dim3(BS), 0, st, attach, t->d_hdr, S, seeds);
    // no wait for the exact count: the insert took at most the 4m + 16 nodes
    // reserved above, and the refit buffers only need an upper bound
    t->pool_used = std::min(t->cap, t->pool_used + 4ll * m + 16);
    CHK(refit_begin(t));
    CHK(refit_mark(t, seeds, nullptr, S, 0));
    if (extra_bound > 0) CHK(refit_mark(t, extra_seeds, n_extra_dev, extra_bound, 0));
    CHK(refit_finish(t));
    return rebuild_tops(t);
}

// ---------------------------------------------------------- box deletion
int delete_boxes_dev(ikd_tree_t t, const float* d_boxes, int nb, int ds) {
    if (nb <= 0) return IKD_OK;
    hipStream_t st = t->st;
    unsigned int* recs;
    // per box: the nodes it cuts, the points it deletes and the subtrees it covers
    long long cap = 2 * std::max<long long>(t->pool_used, 16) + 4ll * nb;
    long long fcap = std::max<long long>(std::max<long long>(t->pool_used, 1 << 16), nb);
    int2 *f0, *f1;
    CHK(ensure(t, B_RECS, cap, &recs));
    CHK(ensure(t, B_FRONT0, fcap, &f0));
    CHK(ensure(t, B_FRONT1, fcap, &f1));
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, st, t->d_hdr, 2);
    launch_fill_i32(t->d_hdr->front3, 0, 3, st);
    launch_fill_i32(&t->d_hdr->front_over, 0, 1, st);
    // steps of BOXDEL_LEVELS levels until the frontier is empty (the host's
    // depth may lag by a call)
    int S0 = 0, S1 = (t->max_depth + 2 + BOXDEL_LEVELS - 1) / BOXDEL_LEVELS;
    while (true) {
        launch_box_delete_bfs(t->pool, d_boxes, nb, f0, f1, (int)fcap, S0, S1, recs, t->d_hdr, (int)cap, st);
        CHK(sync_hdr(t));
        if (t->h_hdr->front_over || t->h_hdr->front3[S1 % 3] == 0) break;
        if (S1 * BOXDEL_LEVELS > 4 * MAX_DEPTH) {
            t->err = "box delete: frontier did not drain";
            return IKD_ERR_INTERNAL;
        }
        S0 = S1;
        S1 += 2;
    }
    int nrec = t->h_hdr->rec_count;
    CHK(check_count(t, "box delete: records", nrec, 0, INT_MAX));
    if (nrec > cap || t->h_hdr->front_over) {
        // overlapping boxes recorded more than one entry per node, or their
        // frontiers together outgrew the buffer: the union is order-free, so
        // apply the boxes in halves
        if (nb == 1) {
            t->err = "box delete: record buffer overflow";
            return IKD_ERR_INTERNAL;
        }
        int h = nb / 2;
        CHK(delete_boxes_dev(t, d_boxes, h, ds));
        return delete_boxes_dev(t, d_boxes + 6ll * h, nb - h, ds);
    }
    if (nrec == 0) return IKD_OK;
    int* seeds;
    CHK(ensure(t, B_SEEDS, nrec, &seeds));
    launch_box_delete_apply(t->pool, recs, t->d_hdr, nrec, ds, seeds, st);
    CHK(refit(t, seeds, nrec, 1));
    return rebuild_tops(t);
}

// ------------------------------------------------------- box re-insertion
// Add_Point_Boxes (cpp:491-511): the boxes one after another, as the
// reference's loop runs Add_by_range once per box (each box's walk reads the
// ranges the previous box's Updates and rebuilds left).  Per box: the walk
// over the cut nodes, the reset of the covered subtrees, then Update and
// Criterion_Check of the cut nodes bottom-up and the rebuilds.
int add_boxes_dev(ikd_tree_t t, const float* d_boxes, int nb) {
    hipStream_t st = t->st;
    for (int b = 0; b < nb; ++b) {
        const float* box = d_boxes + 6ll * b;
        // the walk records each node at most once, as cut or as covered
        const long long cap = std::max<long long>(t->pool_used + 16, 1 << 16);
        unsigned int* recs;
        int2 *f0, *f1;
        CHK(ensure(t, B_RECS, cap, &recs));
        CHK(ensure(t, B_FRONT0, cap, &f0));
        CHK(ensure(t, B_FRONT1, cap, &f1));
        int L0 = 0, L1 = t->max_depth + 2;
        while (true) {
            launch_addbox_walk(t->pool, box, f0, f1, (int)cap, L0, L1, recs, t->d_hdr, (int)cap, st);
            CHK(sync_hdr(t));
            if (t->h_hdr->front_over || t->h_hdr->front3[L1 % 3] == 0) break;
            if (L1 > 4 * MAX_DEPTH) {
                t->err = "Add_Point_Boxes: the walk did not drain";
                return IKD_ERR_INTERNAL;
            }
            L0 = L1;
            L1 += 8;
        }
        const int nrec = t->h_hdr->rec_count;
        if (t->h_hdr->front_over || nrec > cap) {
            t->err = "Add_Point_Boxes: the walk outgrew its buffers";
            return IKD_ERR_INTERNAL;
        }
        CHK(check_count(t, "box re-insertion: records", nrec, 0, cap));
        if (nrec == 0) continue;
        int* seeds;
        CHK(ensure(t, B_SEEDS, nrec, &seeds));
        const int levels = t->max_depth + 2;
        launch_addbox_reset(t->pool, recs, nrec, seeds, f0, f1, (int)cap, levels, t->d_hdr, st);
        HIPCHK(hipGetLastError());
        CHK(sync_hdr(t));
        if (t->h_hdr->front_over || t->h_hdr->front3[levels % 3] != 0) {
            t->err = "Add_Point_Boxes: the reset of covered subtrees did not drain";
            return IKD_ERR_INTERNAL;
        }
        CHK(refit(t, seeds, nrec, 0));
        CHK(rebuild_tops(t));
    }
    return IKD_OK;
}

// ---------------------------------------------------- Add_Points pipeline
int add_points_ds_chunk(ikd_tree_t t, const float* d_xyz, const uint32_t* d_ids, uint32_t id_base, int n, int check,
                        bool apply, int* split_out, int* added_out) {
    hipStream_t st = t->st;
    unsigned long long *ka, *kb;
    uint32_t *va, *vb;
    float4* pts;
    int *flag, *scan, *vstart, *bad;
    CHK(ensure(t, B_MORTK, n, &ka));
    CHK(ensure(t, B_MORTK2, n, &kb));
    CHK(ensure(t, B_MORTV, n, &va));
    CHK(ensure(t, B_MORTV2, n, &vb));
    CHK(ensure(t, B_PTS, n, &pts));
    CHK(ensure(t, B_GFLAG, n, &flag));
    CHK(ensure(t, B_GSCAN, n, &scan));
    CHK(ensure(t, B_VSTART, n + 1, &vstart));
    CHK(ensure(t, B_MISC, 4, &bad));
    launch_fill_i32(bad, 0, 4, st);
    launch_voxel_keys2(d_xyz, d_ids, id_base, n, t->prm.box_length, ka, va, pts, bad, st);
    CHK(sort_pairs(t, ka, kb, va, vb, n, 63));
    launch_voxel_flags(kb, n, flag, st);
    CHK(exscan_i32(t, flag, scan, n));
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, st, t->d_hdr, 16);
    launch_voxel_starts(flag, scan, n, vstart, t->d_hdr, st);
    float* delbox;
    float4* inspt;
    int *delfl, *insfl, *dseeds;
    long long dcap = std::max<long long>(t->pool_used, 16);
    CHK(ensure(t, B_DELBOX, 6ll * n, &delbox));
    CHK(ensure(t, B_DELFL, n, &delfl));
    CHK(ensure(t, B_INSPT, n, &inspt));
    CHK(ensure(t, B_INSFL, n, &insfl));
    CHK(ensure(t, B_RECS, dcap, &dseeds));
    // With conflict checks on, a first pass only decides whether the batch is
    // independent; the deleting pass runs only when it is.
    const int passes = (check && apply) ? 2 : 1;
    // lanes per voxel: teams split the walks of dense voxels (maps of raw
    // scans hold up to thousands of points per voxel); the team deques index
    // their LDS columns with a power-of-two mask
    const int fold_lanes = t->fold_lanes > 0 ? t->fold_lanes : (t->pool_used >= 4096 ? 8 : 1);
    int fold_stack = std::min(t->max_depth + 2, MAX_DEPTH);
    if (fold_lanes > 1) {
        int p2 = 1;
        while (p2 < fold_stack) p2 <<= 1;
        fold_stack = p2;
    }
    for (int pass = 0; pass < passes; ++pass) {
        const bool deleting = apply && pass == passes - 1;
        hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, st, t->d_hdr, 2 | 8);
        launch_voxel_fold2(t->pool, pts, vb, kb, vstart, n, n, t->prm.box_length, check, delbox, delfl, inspt, insfl,
                           t->d_hdr, t->pool_used + 16, deleting ? dseeds : nullptr, (int)dcap, fold_stack, fold_lanes,
                           bad, st);
        int hbad = 0;
        HIPCHK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st));
        CHK(sync_hdr(t));
        if (hbad) {
            t->err = "Add_Points: a voxel index floor(p/box_length) is beyond +-2^20";
            return IKD_ERR_UNSUPPORTED;
        }
        *split_out = t->h_hdr->split_index;
        *added_out = t->h_hdr->added;
        if (!apply || t->h_hdr->split_index != INT_MAX) return IKD_OK;
    }
    int nv = t->h_hdr->group_count;
    int ndel = t->h_hdr->rec_count;
    CHK(check_count(t, "downsample: voxels", nv, 0, n));
    CHK(check_count(t, "downsample: deleted points", ndel, 0, dcap));
    CHK(check_count(t, "downsample: kept points", t->h_hdr->ins_count, 0, nv));
    // compact the points to insert (voxel order), then insert them and refit
    // the insert paths and the deleted points' paths in one pass
    float4* cpts;
    int* counts;
    CHK(ensure(t, B_CPTS, nv, &cpts));
    CHK(ensure(t, B_CNT64, 2, &counts));
    launch_fill_i32(counts, 0, 2, st);
    CHK(exscan_i32(t, insfl, scan, nv));
    launch_compact_pts(insfl, scan, inspt, nv, t->d_hdr, cpts, counts + 1, st);
    // the fold counted the voxels that keep a point (the compacted count)
    const int hc = t->h_hdr->ins_count;
    return insert_points(t, cpts, hc, dseeds, nullptr, ndel);
}

bool exact_voxels(float ds) {
    int e;
    float m = std::frexp(ds, &e);
    return ds > 0.0f && m == 0.5f;
}

// Ends with the header and root summary fresh.  hdr_valid: they already are
// at the start.
int add_points_dev(ikd_tree_t t, const float* d_xyz, const uint32_t* d_ids, int n, int downsample, int* added,
                   bool hdr_valid) {
    *added = 0;
    if (n <= 0) return hdr_valid ? IKD_OK : refresh(t);
    uint32_t base = t->next_id;
    if (!d_ids) t->next_id += (uint32_t)n;
    if (!hdr_valid) CHK(sync_hdr(t));
    if (!downsample) {
        float4* pts;
        CHK(ensure(t, B_PTS, n, &pts));
        launch_pack_points(d_xyz, d_ids, base, n, pts, t->st);
        CHK(insert_points(t, pts, n));
        return compact_if_needed(t, false, true);  // the insert's rebuild pass refreshed both
    }
    if (!(t->prm.box_length > 0.0f)) {
        t->err = "Add_Points: downsample needs box_length > 0";
        return IKD_ERR_ARG;
    }
    int check = exact_voxels(t->prm.box_length) ? 0 : 1;
    int start = 0;
    while (start < n) {
        int len = n - start;
        const float* x = d_xyz + 3ll * start;
        const uint32_t* ids = d_ids ? d_ids + start : nullptr;
        uint32_t b0 = base + (uint32_t)start;
        int split = INT_MAX, a = 0;
        CHK(add_points_ds_chunk(t, x, ids, b0, len, check, true, &split, &a));
        if (split == INT_MAX) {
            *added += a;
            break;
        }
        // ordered handling: the conflict-free prefix, then (for a point that
        // sits in two voxel boxes or outside its own) that point alone
        int idx = split >> 1;
        bool single = (split & 1) == 0;
        if (idx > 0) {
            CHK(add_points_ds_chunk(t, x, ids, b0, idx, check, true, &split, &a));
            if (split != INT_MAX) {
                t->err = "Add_Points: ordered prefix still conflicts";
                return IKD_ERR_INTERNAL;
            }
            *added += a;
        }
        if (single) {
            CHK(add_points_ds_chunk(t, x + 3ll * idx, ids ? ids + idx : nullptr, b0 + (uint32_t)idx, 1, 0, true,
                                    &split, &a));
            *added += a;
            start += idx + 1;
        } else {
            start += idx;
        }
    }
    // the last chunk ended with its insert's rebuild pass, which refreshed both
    return compact_if_needed(t, false, true);
}

// Lanes per query so that a launch has about 2^17 lanes in flight: a small
// batch (one scan) shares each query among a team of lanes, a large batch
// keeps one lane per query (the most work-efficient form).
int auto_lanes(int nq) {
    int lanes = 1;
    while (lanes < 16 && (long long)nq * lanes < (1ll << 17)) lanes *= 2;
    return lanes;
}

float max_dist_threshold(double max_dist) {
    double m = max_dist * max_dist;
    if (std::isnan(m)) return -INFINITY;
    float f = (float)m;
    if ((double)f > m) f = std::nextafter(f, -INFINITY);
    return f;
}

int knn_dev(ikd_tree_t t, const float* d_q, int nq, int k, double max_dist, float* d_xyz, float* d_d2, uint32_t* d_ids,
            int* d_cnt, unsigned int* d_visits, unsigned int* d_boxes) {
    if (k < 1) {
        t->err = "Nearest_Search: k must be >= 1";
        return IKD_ERR_ARG;
    }
    if (k > 64) {
        t->err = "Nearest_Search: k above 64 is not supported";
        return IKD_ERR_UNSUPPORTED;
    }
    if (nq <= 0) return IKD_OK;
    if (t->max_depth + 2 > MAX_DEPTH) {
        t->hdr_valid = false;
        CHK(compact_if_needed(t, true));
    }
    KnnArgs a{};
    a.hot = t->pool.hot;
    a.ids = t->pool.ids;
    a.q = d_q;
    a.order = nullptr;
    a.nq = nq;
    a.k = k;
    a.out_stride = k;
    a.thr = max_dist_threshold(max_dist);
    a.out_xyz = d_xyz;
    a.out_d2 = d_d2;
    a.out_ids = d_ids;
    a.out_counts = d_cnt;
    a.visits = d_visits;
    a.boxes = d_boxes;
    a.overflow = &t->d_hdr->overflow;
    a.max_steps = (int)std::min<long long>(t->pool_used + 16, INT_MAX);
    a.compact_stack = (t->knn_compact && t->cap <= (1ll << 24)) ? 1 : 0;
    if (t->query_order && nq >= 4096) {
        unsigned long long *mk, *mk2;
        uint32_t *mv, *mv2;
        CHK(ensure(t, B_MORTK, nq, &mk));
        CHK(ensure(t, B_MORTK2, nq, &mk2));
        CHK(ensure(t, B_MORTV, nq, &mv));
        CHK(ensure(t, B_MORTV2, nq, &mv2));
        launch_morton(d_q, nq, t->pool.hot, mk, mv, t->st);
        CHK(sort_pairs(t, mk, mk2, mv, mv2, nq, 30));
        a.order = mv2;
    }
    // a lane defers at most one far child per level above its node
    int stack = t->max_depth + 2;
    if (t->knn_stack > stack) stack = t->knn_stack;
    if (stack > MAX_DEPTH) stack = MAX_DEPTH;
    int lanes = t->knn_lanes >= 0 ? t->knn_lanes : auto_lanes(nq);
    if (t->knn_stage > 0 && lanes == 1) {
        // the staged levels, capped so that one workgroup's LDS (stacks and
        // table) stays within 64 KiB
        int levels = t->knn_stage;
        const long long stack_b = ((long long)stack * 64 * 6 + 15) / 16 * 16;
        while (levels > 0 && stack_b + ((1ll << levels) - 1) * (long long)(sizeof(NodeHot) + sizeof(int)) > 65536)
            --levels;
        if (levels > 0) {
            const int slots = (1 << levels) - 1;
            NodeHot* top;
            int* top_idx;
            CHK(ensure(t, B_TOPT, slots, &top));
            CHK(ensure(t, B_TOPI, slots, &top_idx));
            launch_top_table(t->pool.hot, levels, top, top_idx, t->st);
            a.top = top;
            a.top_idx = top_idx;
            a.top_slots = slots;
        }
    }
    int* next;
    CHK(ensure(t, B_PNEXT, 8, &next));
    if (launch_knn(a, k, lanes, stack, t->st, next, 256 * 10) != 0) return IKD_ERR_UNSUPPORTED;
    HIPCHK(hipGetLastError());
    return IKD_OK;
}

// Range search plan on the owner's stream (count walk, item lists, item
// counts, offsets): d_offsets (device, nq + 1) gets every query's first
// output position; *total and *ntask come back to the host (two waits: the
// item count sizes the lists, the total sizes the output).  slots: scratch
// slots for the per-query item counts and offsets, the items, the item
// counts and the item offsets.
hipError_t owner_wait(ikd_tree_t t) { return wait_stream(t); }
template <class Owner>
hipError_t owner_wait(Owner* b) { return hipStreamSynchronize(b->st); }

template <class Owner>
int range_plan(Owner* t, RangeArgs& a, int nq, long long* d_offsets, const int (&slots)[5], long long pool_bound,
               long long* total, long long* ntask_out) {
    hipStream_t st = t->st;
    long long *tcnt, *toff, *icnt, *ioff;
    int4* items;
    CHK(ensure(t, slots[0], nq + 1, &tcnt));
    CHK(ensure(t, slots[1], nq + 1, &toff));
    launch_fill_i32(reinterpret_cast<int*>(tcnt + nq), 0, 2, st);  // the scan's trailing zero
    a.task_nodes = RANGE_TASK_NODES;
    a.tcounts = tcnt;
    launch_range_top(a, 0, st);
    CHK(exscan_i64(t, tcnt, toff, nq + 1));
    long long ntask = 0;
    HIPCHK(hipMemcpyAsync(&ntask, toff + nq, sizeof(long long), hipMemcpyDeviceToHost, st));
    HIPCHK(owner_wait(t));
    CHK(check_count(t, "range search: items", ntask, 0, std::min<long long>((long long)nq * pool_bound, INT_MAX - 1)));
    CHK(ensure(t, slots[2], std::max<long long>(ntask, 1), &items));
    CHK(ensure(t, slots[3], ntask + 1, &icnt));
    CHK(ensure(t, slots[4], ntask + 1, &ioff));
    a.toffsets = toff;
    a.tasks = items;
    launch_range_top(a, 1, st);
    launch_fill_i32(reinterpret_cast<int*>(icnt + ntask), 0, 2, st);
    launch_range_item_count(a, ntask, icnt, st);
    CHK(exscan_i64(t, icnt, ioff, (int)ntask + 1));
    launch_range_query_offsets(toff, ioff, nq, d_offsets, st);
    a.item_offsets = ioff;
    HIPCHK(hipMemcpyAsync(total, d_offsets + nq, sizeof(long long), hipMemcpyDeviceToHost, st));
    HIPCHK(owner_wait(t));
    *ntask_out = ntask;
    return IKD_OK;
}

int range_dev(ikd_tree_t t, int mode, const float* d_boxes, const float* d_centers, const float* d_radii, int nq,
              int64_t* h_offsets, float* h_xyz, uint32_t* h_ids, int64_t capacity) {
    if (t->max_depth + 2 > MAX_DEPTH) {
        t->hdr_valid = false;
        CHK(compact_if_needed(t, true));
    }
    hipStream_t st = t->st;
    long long* off;
    CHK(ensure(t, B_OFF64, nq + 1, &off));
    RangeArgs a{};
    a.max_steps = t->pool_used + 16;
    a.pool = t->pool;
    a.boxes = d_boxes;
    a.centers = d_centers;
    a.radii = d_radii;
    a.nq = nq;
    a.mode = mode;
    a.overflow = &t->d_hdr->overflow;
    a.stack = std::min(t->max_depth + 2, MAX_DEPTH);
    long long total = 0, ntask = 0;
    static const int slots[5] = {B_TCNT64, B_TOFF64, B_RTASKS, B_CNT64, B_ICNT64};
    CHK(range_plan(t, a, nq, off, slots, std::max<long long>(t->pool_used, 1), &total, &ntask));
    HIPCHK(hipMemcpyAsync(h_offsets, off, (nq + 1) * sizeof(long long), hipMemcpyDeviceToHost, st));
    CHK(sync_hdr(t));
    CHK(check_count(t, "range search: results", total, 0, (long long)nq * std::max<long long>(t->pool_used, 1)));
    if (!h_xyz) return IKD_OK;
    if (capacity < total) {
        t->err = "output capacity below the result count";
        return IKD_ERR_CAPACITY;
    }
    if (total == 0) return IKD_OK;
    float* dx;
    uint32_t* di;
    CHK(ensure(t, B_IO0, 3 * total, &dx));
    CHK(ensure(t, B_IO1, total, &di));
    a.out_xyz = dx;
    a.out_ids = di;
    launch_range_item_fill(a, ntask, a.item_offsets, st);
    HIPCHK(hipMemcpyAsync(h_xyz, dx, 3 * total * sizeof(float), hipMemcpyDeviceToHost, st));
    if (h_ids) HIPCHK(hipMemcpyAsync(h_ids, di, total * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    return sync_hdr(t);
}

template <class T>
int upload(ikd_tree_t t, int slot, const T* h, long long n, T** d) {
    CHK(ensure(t, slot, std::max<long long>(n, 1), d));
    if (n > 0) HIPCHK(hipMemcpyAsync(*d, h, n * sizeof(T), hipMemcpyHostToDevice, t->st));
    return IKD_OK;
}

}  // namespace

// =================================================================== C ABI
// Every entry point on a tree holds the tree's lock for the whole call.
#define CHECK_T()                    \
    if (!t) return IKD_ERR_ARG;      \
    std::lock_guard<std::recursive_mutex> call_lock_(t->mu)

extern "C" {

int ikd_create(float delete_param, float balance_param, float box_length, ikd_tree_t* out) {
    if (!out) return IKD_ERR_ARG;
    *out = nullptr;
    ikd_tree_t t = new ikd_tree_s();
    t->prm = Params{delete_param, balance_param, box_length};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        delete t;
        return IKD_ERR_HIP;
    }
    t->mp = library_pool(dev);
    if (hipStreamCreateWithFlags(&t->st, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&t->d_status, sizeof(DevStatus)) != hipSuccess ||
        hipHostMalloc(&t->h_status, sizeof(DevStatus), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&t->ev, hipEventDisableTiming) != hipSuccess) {
        ikd_destroy(t);
        return IKD_ERR_HIP;
    }
    t->d_hdr = &t->d_status->hdr;
    t->h_hdr = &t->h_status->hdr;
    t->d_root = &t->d_status->root;
    t->h_root = &t->h_status->root;
    // a zero header first: the pool's initialisation keeps only the record count
    launch_fill_i32(reinterpret_cast<int*>(t->d_hdr), 0, sizeof(DevHdr) / sizeof(int), t->st);
    int rc = reset_pool(t);
    if (rc == IKD_OK) rc = sync_hdr(t);
    if (rc == IKD_OK) rc = read_root(t);
    if (rc != IKD_OK) {
        ikd_destroy(t);
        return rc;
    }
    *out = t;
    return IKD_OK;
}

int ikd_destroy(ikd_tree_t t) {
    if (!t) return IKD_ERR_ARG;
    {
        // wait for a call still running on another thread
        std::lock_guard<std::recursive_mutex> g(t->mu);
    }
    if (t->st) {
        if (t->cap > 0) free_pool(t, &t->pool);
        for (auto& b : t->b) dev_free(t, b.p);
        dev_free(t, t->removed);
        hipStreamSynchronize(t->st);
    }
    stage_release(t);
    if (t->d_status) hipFree(t->d_status);
    if (t->h_status) hipHostFree(t->h_status);
    if (t->ev) hipEventDestroy(t->ev);
    if (t->st) hipStreamDestroy(t->st);
    delete t;
    return IKD_OK;
}

int ikd_set_delete_criterion_param(ikd_tree_t t, float v) {
    CHECK_T();
    t->prm.delete_param = v;
    return IKD_OK;
}
int ikd_set_balance_criterion_param(ikd_tree_t t, float v) {
    CHECK_T();
    t->prm.balance_param = v;
    return IKD_OK;
}
int ikd_set_downsample_param(ikd_tree_t t, float v) {
    CHECK_T();
    t->prm.box_length = v;
    return IKD_OK;
}
int ikd_initialize(ikd_tree_t t, float d, float b, float l) {
    CHECK_T();
    t->prm = Params{d, b, l};
    return IKD_OK;
}

int ikd_build(ikd_tree_t t, const float* xyz, const uint32_t* ids, int64_t n) {
    CHECK_T();
    t->hdr_valid = false;
    if (n < 0 || (n > 0 && !xyz) || n >= (1ll << 27)) return IKD_ERR_ARG;
    CHK(reset_pool(t));
    t->next_id = ids ? t->next_id : (uint32_t)n;
    if (n > 0) {
        float* dx;
        uint32_t* di = nullptr;
        CHK(upload(t, B_IO4, xyz, 3 * n, &dx));
        if (ids) CHK(upload(t, B_IO2, ids, n, &di));
        float4* pts;
        int* segof;
        CHK(ensure(t, B_PTS, n, &pts));
        CHK(ensure(t, B_SEGOF, n, &segof));
        launch_pack_points(dx, di, 0u, (int)n, pts, t->st);
        launch_fill_i32(segof, 0, n, t->st);
        int4 s0 = make_int4(0, (int)n - 1, ROOT, SENTINEL);
        int d0 = 0;
        int4* dseg;
        int* dsegd;
        CHK(ensure(t, B_ISEG, 1, &dseg));
        CHK(ensure(t, B_ISEGD, 1, &dsegd));
        CHK(stage_copy(t, dseg, &s0, sizeof(s0)));
        CHK(stage_copy(t, dsegd, &d0, sizeof(d0)));
        CHK(run_build(t, pts, segof, (int)n, dseg, dsegd, 1, (int)n));
        hipLaunchKernelGGL(k_link_bits, dim3(1), dim3(1), 0, t->st, t->pool.hot, SENTINEL, LK_HASL, LK_TDL);
    }
    CHK(sync_hdr(t));
    return read_root(t);
}

int ikd_nearest_search_device(ikd_tree_t t, const float* d_q, int64_t nq, int k, double max_dist, float* d_xyz,
                              float* d_d2, uint32_t* d_ids, int32_t* d_cnt) {
    CHECK_T();
    if (nq < 0 || nq >= (1ll << 31)) return IKD_ERR_ARG;
    return knn_dev(t, d_q, (int)nq, k, max_dist, d_xyz, d_d2, d_ids, d_cnt, nullptr, nullptr);
}

int ikd_nearest_search(ikd_tree_t t, const float* q, int64_t nq, int k, double max_dist, float* out_xyz,
                       float* out_d2, uint32_t* out_ids, int32_t* out_counts) {
    CHECK_T();
    if (nq < 0 || nq >= (1ll << 31) || (nq > 0 && (!q || !out_xyz || !out_d2 || !out_ids || !out_counts)))
        return IKD_ERR_ARG;
    if (k < 1) {
        t->err = "Nearest_Search: k must be >= 1";
        return IKD_ERR_ARG;
    }
    if (nq == 0) return IKD_OK;
    float *dq, *dx, *dd;
    uint32_t* di;
    int* dc;
    CHK(upload(t, B_IO0, q, 3 * nq, &dq));
    CHK(ensure(t, B_IO1, 3 * nq * k, &dx));
    CHK(ensure(t, B_IO2, nq * k, &dd));
    CHK(ensure(t, B_IO3, nq * k, &di));
    CHK(ensure(t, B_IO4, nq, &dc));
    CHK(knn_dev(t, dq, (int)nq, k, max_dist, dx, dd, di, dc, nullptr, nullptr));
    HIPCHK(hipMemcpyAsync(out_xyz, dx, 3 * nq * k * sizeof(float), hipMemcpyDeviceToHost, t->st));
    HIPCHK(hipMemcpyAsync(out_d2, dd, nq * k * sizeof(float), hipMemcpyDeviceToHost, t->st));
    HIPCHK(hipMemcpyAsync(out_ids, di, nq * k * sizeof(uint32_t), hipMemcpyDeviceToHost, t->st));
    HIPCHK(hipMemcpyAsync(out_counts, dc, nq * sizeof(int32_t), hipMemcpyDeviceToHost, t->st));
    return sync_hdr(t);
}

int ikd_nearest_search_counters(ikd_tree_t t, const float* q, int64_t nq, int k, double max_dist,
                                uint32_t* out_visits, uint32_t* out_boxes) {
    CHECK_T();
    if (nq < 0 || nq >= (1ll << 31) || (nq > 0 && (!q || !out_visits || !out_boxes))) return IKD_ERR_ARG;
    if (k < 1) return IKD_ERR_ARG;
    if (nq == 0) return IKD_OK;
    float *dq, *dx, *dd;
    uint32_t *di, *dv, *db;
    int* dc;
    CHK(upload(t, B_IO0, q, 3 * nq, &dq));
    CHK(ensure(t, B_IO1, 3 * nq * k, &dx));
    CHK(ensure(t, B_IO2, nq * k, &dd));
    CHK(ensure(t, B_IO3, nq * k, &di));
    CHK(ensure(t, B_IO4, nq, &dc));
    CHK(ensure(t, B_DELBOX, nq, &dv));
    CHK(ensure(t, B_DELFL, nq, &db));
    // the counts are those of the one-lane depth-first walk (the reference's
    // traversal), whatever launch shape the tree is tuned to
    const int lanes = t->knn_lanes, order = t->query_order;
    t->knn_lanes = 1;
    t->query_order = 0;
    int rc = knn_dev(t, dq, (int)nq, k, max_dist, dx, dd, di, dc, dv, db);
    t->knn_lanes = lanes;
    t->query_order = order;
    CHK(rc);
    HIPCHK(hipMemcpyAsync(out_visits, dv, nq * sizeof(uint32_t), hipMemcpyDeviceToHost, t->st));
    HIPCHK(hipMemcpyAsync(out_boxes, db, nq * sizeof(uint32_t), hipMemcpyDeviceToHost, t->st));
    return sync_hdr(t);
}

int ikd_box_search(ikd_tree_t t, const float* boxes, int64_t nb, int64_t* out_offsets, float* out_xyz,
                   uint32_t* out_ids, int64_t capacity) {
    CHECK_T();
    if (nb < 0 || nb >= (1ll << 31) || !out_offsets || (nb > 0 && !boxes)) return IKD_ERR_ARG;
    if (nb == 0) {
        out_offsets[0] = 0;
        return IKD_OK;
    }
    float* db;
    CHK(upload(t, B_IO4, boxes, 6 * nb, &db));
    return range_dev(t, 0, db, nullptr, nullptr, (int)nb, out_offsets, out_xyz, out_ids, capacity);
}

int ikd_radius_search(ikd_tree_t t, const float* centers, const float* radii, int64_t nq, int64_t* out_offsets,
                      float* out_xyz, uint32_t* out_ids, int64_t capacity) {
    CHECK_T();
    if (nq < 0 || nq >= (1ll << 31) || !out_offsets || (nq > 0 && (!centers || !radii))) return IKD_ERR_ARG;
    if (nq == 0) {
        out_offsets[0] = 0;
        return IKD_OK;
    }
    float *dc, *dr;
    CHK(upload(t, B_IO4, centers, 3 * nq, &dc));
    CHK(upload(t, B_IO3, radii, nq, &dr));
    return range_dev(t, 1, nullptr, dc, dr, (int)nq, out_offsets, out_xyz, out_ids, capacity);
}

int ikd_add_points_device(ikd_tree_t t, const float* d_xyz, const uint32_t* d_ids, int64_t n, int downsample_on,
                          int32_t* out_added) {
    CHECK_T();
    if (n < 0 || n >= (1ll << 27) || (n > 0 && !d_xyz)) return IKD_ERR_ARG;
    int added = 0;
    const bool valid = t->hdr_valid;
    t->hdr_valid = false;
    int rc = add_points_dev(t, d_xyz, d_ids, (int)n, downsample_on, &added, valid);
    if (out_added) *out_added = added;
    if (rc == IKD_OK) t->hdr_valid = true;  // add_points_dev ends with both fresh
    return rc;
}

int ikd_add_points(ikd_tree_t t, const float* xyz, const uint32_t* ids, int64_t n, int downsample_on,
                   int32_t* out_added) {
    CHECK_T();
    if (n < 0 || n >= (1ll << 27) || (n > 0 && !xyz)) return IKD_ERR_ARG;
    if (out_added) *out_added = 0;
    if (n == 0) return IKD_OK;
    float* dx;
    uint32_t* di = nullptr;
    // the caller's buffers outlive the call, which ends with a wait
    CHK(upload(t, B_ADDX, xyz, 3 * n, &dx));
    if (ids) CHK(upload(t, B_ADDI, ids, n, &di));
    return ikd_add_points_device(t, dx, di, n, downsample_on, out_added);
}

int ikd_delete_points(ikd_tree_t t, const float* xyz, int64_t n) {
    CHECK_T();
    t->hdr_valid = false;
    if (n < 0 || n >= (1ll << 31) || (n > 0 && !xyz)) return IKD_ERR_ARG;
    if (n == 0) return IKD_OK;
    float* dx;
    int *cand, *seeds, *last_pass, *path_seed, *del_below, *max_del;
    CHK(sync_hdr(t));
    const long long nodes = std::max<long long>(t->pool_used, 16);
    CHK(upload(t, B_IO4, xyz, 3 * n, &dx));
    CHK(ensure(t, B_IO0, n, &cand));
    CHK(ensure(t, B_SEEDS, 2 * n, &seeds));
    CHK(ensure(t, B_PSEED, n, &path_seed));
    CHK(ensure(t, B_LASTP, nodes, &last_pass));
    CHK(ensure(t, B_DBELOW, nodes, &del_below));
    CHK(ensure(t, B_MAXDEL, nodes, &max_del));
    launch_fill_i32(last_pass, -1, nodes, t->st);
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(1), 0, t->st, t->d_hdr, 4);
    launch_delete_points_find(t->pool, dx, (int)n, cand, t->d_hdr, t->st);
    CHK(sync_hdr(t));
    if (t->h_hdr->conflicts == 0) {
        launch_fill_i32(del_below, 0, nodes, t->st);
        launch_fill_i32(max_del, -1, nodes, t->st);
        launch_delete_points_apply(t->pool, dx, cand, del_below, max_del, last_pass, path_seed, (int)n, seeds,
                                   t->d_hdr, t->st);
    } else {
        launch_delete_points_unclaim(t->pool, cand, (int)n, seeds, t->st);
        launch_delete_points_serial(t->pool, dx, (int)n, last_pass, path_seed, seeds, t->d_hdr, t->st);
    }
    // the deleted nodes and every node a walk Updates (cpp:756-759)
    CHK(refit(t, seeds, 2 * (int)n, 0));
    CHK(rebuild_tops(t));
    return compact_if_needed(t, false, true);
}

int ikd_delete_point_boxes(ikd_tree_t t, const float* boxes, int64_t nb, int32_t* out_deleted) {
    CHECK_T();
    const bool valid = t->hdr_valid;
    t->hdr_valid = false;
    if (nb < 0 || nb >= (1ll << 31) || (nb > 0 && !boxes)) return IKD_ERR_ARG;
    if (out_deleted) *out_deleted = 0;
    if (nb == 0) {
        t->hdr_valid = valid;
        return IKD_OK;
    }
    if (!valid) CHK(read_root(t));
    long long before = live_count(t);
    float* db;
    CHK(upload(t, B_IO4, boxes, 6 * nb, &db));
    CHK(delete_boxes_dev(t, db, (int)nb, 0));
    CHK(compact_if_needed(t, false));  // ends with the header and root summary fresh
    if (out_deleted) *out_deleted = (int32_t)(before - live_count(t));
    t->hdr_valid = true;
    return IKD_OK;
}

int ikd_add_point_boxes(ikd_tree_t t, const float* boxes, int64_t nb) {
    CHECK_T();
    t->hdr_valid = false;
    if (nb < 0 || nb >= (1ll << 31) || (nb > 0 && !boxes)) return IKD_ERR_ARG;
    if (nb == 0) return IKD_OK;
    CHK(sync_hdr(t));
    float* db;
    CHK(upload(t, B_IO4, boxes, 6 * nb, &db));
    CHK(add_boxes_dev(t, db, (int)nb));
    return compact_if_needed(t, false);
}

int ikd_size(ikd_tree_t t, int64_t* out) {
    CHECK_T();
    if (!out) return IKD_ERR_ARG;
    CHK(read_root(t));
    *out = has_root(t) ? t->h_root->cnt.x : 0;
    return IKD_OK;
}

int ikd_validnum(ikd_tree_t t, int64_t* out) {
    CHECK_T();
    if (!out) return IKD_ERR_ARG;
    CHK(read_root(t));
    *out = live_count(t);
    return IKD_OK;
}

int ikd_tree_range(ikd_tree_t t, float* out6) {
    CHECK_T();
    if (!out6) return IKD_ERR_ARG;
    CHK(read_root(t));
    for (int i = 0; i < 6; ++i) out6[i] = has_root(t) ? t->h_root->box[i] : 0.0f;
    return IKD_OK;
}

// root_alpha, cpp:147-165: the root's statistics as of its last Update
// (cpp:1315-1321), kept in the device header by the refit and the builds.
int ikd_root_alpha(ikd_tree_t t, float* alpha_bal, float* alpha_del) {
    CHECK_T();
    if (!alpha_bal || !alpha_del) return IKD_ERR_ARG;
    CHK(sync_hdr(t));
    *alpha_bal = t->h_hdr->alpha_bal;
    *alpha_del = t->h_hdr->alpha_del;
    return IKD_OK;
}

int ikd_flatten(ikd_tree_t t, float* out_xyz, uint32_t* out_ids, int64_t capacity, int64_t* out_n) {
    CHECK_T();
    if (!out_n) return IKD_ERR_ARG;
    CHK(read_root(t));
    CHK(sync_hdr(t));
    long long live = live_count(t);
    *out_n = live;
    if (live == 0) return IKD_OK;
    if (!out_xyz) return IKD_OK;
    if (capacity < live) {
        t->err = "flatten: capacity below validnum()";
        return IKD_ERR_CAPACITY;
    }
    float4* pts;
    int* segof;
    std::vector<int> roots = {ROOT};
    std::vector<long long> lives = {live};
    CHK(enumerate_live(t, roots, lives, t->h_root->cnt.x, live, 0, false, &pts, &segof));
    float* dx;
    uint32_t* di;
    CHK(ensure(t, B_IO0, 3 * live, &dx));
    CHK(ensure(t, B_IO1, live, &di));
    hipLaunchKernelGGL(k_unpack_points, dim3(nblk(live)), dim3(BS), 0, t->st, pts, (int)live, dx, di);
    HIPCHK(hipMemcpyAsync(out_xyz, dx, 3 * live * sizeof(float), hipMemcpyDeviceToHost, t->st));
    if (out_ids) HIPCHK(hipMemcpyAsync(out_ids, di, live * sizeof(uint32_t), hipMemcpyDeviceToHost, t->st));
    return sync_hdr(t);
}

// Payload ids still reachable: every node of the tree (live or deleted: a
// deleted point can come back through Add_Point_Boxes or leave through a
// rebuild's removed-point record) and every point in that record.
int ikd_list_payload_ids(ikd_tree_t t, uint32_t* out_ids, int64_t capacity, int64_t* out_n) {
    CHECK_T();
    if (!out_n || capacity < 0 || (capacity > 0 && !out_ids)) return IKD_ERR_ARG;
    CHK(sync_hdr(t));
    const long long used = t->pool_used;
    std::vector<NodeHot> hot(used);
    std::vector<uint32_t> ids(used);
    HIPCHK(hipMemcpy(hot.data(), t->pool.hot, used * sizeof(NodeHot), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ids.data(), t->pool.ids, used * sizeof(uint32_t), hipMemcpyDeviceToHost));
    std::vector<uint32_t> out;
    if (fbits(hot[SENTINEL].f[3]) & LK_HASL) {
        std::vector<int> stack = {ROOT};
        while (!stack.empty()) {
            const int u = stack.back();
            stack.pop_back();
            if (u < 0 || u >= used || (long long)out.size() > used) {
                t->err = "list_payload_ids: node links out of range";
                return IKD_ERR_INTERNAL;
            }
            out.push_back(ids[u]);
            const uint32_t lk = fbits(hot[u].f[3]);
            const int pair = (int)lk_pair(lk);
            for (int s = 0; s < 2; ++s)
                if (lk & lk_has(s)) stack.push_back(2 * pair + s);
        }
    }
    const long long nrem = t->h_hdr->removed_count;
    if (nrem > 0) {
        std::vector<float4> rec(nrem);
        HIPCHK(hipMemcpy(rec.data(), t->removed, nrem * sizeof(float4), hipMemcpyDeviceToHost));
        for (const float4& r : rec) out.push_back(fbits(r.w));
    }
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
    *out_n = (int64_t)out.size();
    if (!out_ids) return IKD_OK;
    if (capacity < (int64_t)out.size()) {
        t->err = "list_payload_ids: capacity below the id count";
        return IKD_ERR_CAPACITY;
    }
    std::memcpy(out_ids, out.data(), out.size() * sizeof(uint32_t));
    return IKD_OK;
}

int ikd_relabel_payload_ids(ikd_tree_t t, const uint32_t* old_ids, int64_t n) {
    CHECK_T();
    t->hdr_valid = false;
    if (n < 0 || n >= (1ll << 31) || (n > 0 && !old_ids)) return IKD_ERR_ARG;
    for (int64_t i = 1; i < n; ++i)
        if (old_ids[i] <= old_ids[i - 1]) {
            t->err = "relabel_payload_ids: old ids must be strictly ascending";
            return IKD_ERR_ARG;
        }
    CHK(sync_hdr(t));
    if (n > 0) {
        uint32_t* d_old;
        CHK(upload(t, B_IO4, old_ids, n, &d_old));
        hipLaunchKernelGGL(k_relabel_ids, dim3(nblk(t->pool_used)), dim3(BS), 0, t->st, t->pool.ids,
                           (int)t->pool_used, d_old, (int)n);
        if (t->h_hdr->removed_count > 0)
            hipLaunchKernelGGL(k_relabel_records, dim3(nblk(t->h_hdr->removed_count)), dim3(BS), 0, t->st, t->removed,
                               t->h_hdr->removed_count, d_old, (int)n);
        HIPCHK(hipGetLastError());
    }
    t->next_id = (uint32_t)n;
    return sync_hdr(t);
}

// Pre-order dump of the live structure (diagnostic; the row layout is the
// oracle's: x, y, z, id, axis, TreeSize, invalid, flags, range[6], left, right).
// Lazy labels are resolved on the way down, as flatten() does.
int ikd_dump_nodes(ikd_tree_t t, void* out_rows, int64_t capacity, int64_t* out_n) {
    CHECK_T();
    if (!out_n) return IKD_ERR_ARG;
    CHK(sync_hdr(t));
    long long used = t->pool_used;
    std::vector<NodeHot> hot(used);
    std::vector<uint32_t> ids(used), meta(used);
    std::vector<int4> cnt(used);
    HIPCHK(hipMemcpy(hot.data(), t->pool.hot, used * sizeof(NodeHot), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ids.data(), t->pool.ids, used * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(meta.data(), t->pool.meta, used * sizeof(uint32_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cnt.data(), t->pool.cnt, used * sizeof(int4), hipMemcpyDeviceToHost));
    struct Row {
        float x, y, z;
        uint32_t id;
        int32_t axis, tree_size, invalid, flags;
        float range[6];
        int32_t left, right;
    };
    std::vector<Row> rows;
    uint32_t lk0 = fbits(hot[SENTINEL].f[3]);
    if (lk0 & LK_HASL) {
        struct Item {
            int node, parent_row, side;
            bool dead;
            const float* box;
        };
        std::vector<Item> stack;
        stack.push_back({ROOT, -1, 0, (lk0 & LK_TDL) != 0, hot[SENTINEL].f + 4});
        while (!stack.empty()) {
            Item it = stack.back();
            stack.pop_back();
            int u = it.node;
            const float* f = hot[u].f;
            uint32_t lk = fbits(f[3]);
            uint32_t m = meta[u];
            bool dead = it.dead || (m & M_TDEL);
            Row r{};
            r.x = f[0];
            r.y = f[1];
            r.z = f[2];
            r.id = ids[u];
            r.axis = (int)(m & M_AXIS);
            r.tree_size = cnt[u].x;
            r.invalid = dead ? cnt[u].x : cnt[u].y;
            bool pd = dead || (m & M_PDEL);
            r.flags = (pd ? 1 : 0) | (dead ? 2 : 0) | ((m & M_PDS) ? 4 : 0) | ((m & M_TDS) ? 8 : 0);
            for (int i = 0; i < 6; ++i) r.range[i] = it.box[i];
            r.left = r.right = -1;
            int me = (int)rows.size();
            rows.push_back(r);
            if (it.parent_row >= 0) {
                if (it.side == 0) rows[it.parent_row].left = me;
                else rows[it.parent_row].right = me;
            }
            int pair = (int)lk_pair(lk);
            for (int sd = 1; sd >= 0; --sd)
                if (lk & lk_has(sd))
                    stack.push_back({2 * pair + sd, me, sd, dead || (lk & lk_td(sd)), f + 4 + 6 * sd});
        }
    }
    *out_n = (int64_t)rows.size();
    if (out_rows) {
        if (capacity < (int64_t)rows.size()) return IKD_ERR_CAPACITY;
        std::memcpy(out_rows, rows.data(), rows.size() * sizeof(Row));
    }
    return IKD_OK;
}

int ikd_synchronize(ikd_tree_t t) {
    CHECK_T();
    return sync_hdr(t);
}

int ikd_stream(ikd_tree_t t, void** out) {
    CHECK_T();
    if (!out) return IKD_ERR_ARG;
    *out = (void*)t->st;
    return IKD_OK;
}

int ikd_rebuild_deferred(ikd_tree_t t, int64_t max_points, int32_t* out_rebuilt) {
    CHECK_T();
    t->hdr_valid = false;
    if (out_rebuilt) *out_rebuilt = 0;
    if (t->compact_due) {
        // the whole tree, whatever max_points says; it empties the list too
        t->compact_due = false;
        CHK(refresh(t));
        if (pool_sparse(t, live_count(t))) {
            CHK(compact_if_needed(t, true, true));
            if (out_rebuilt) *out_rebuilt = 1;
            return IKD_OK;
        }
    }
    CHK(sync_hdr(t));
    CHK(check_count(t, "deferred rebuilds", t->h_hdr->defer_count, 0, INT_MAX));
    int n = std::min(t->h_hdr->defer_count, DEFER_CAP);
    if (n <= 0) return IKD_OK;
    hipStream_t st = t->st;
    int *defer, *before, *keep;
    int4* info;
    CHK(ensure(t, B_DEFER, DEFER_CAP, &defer));
    CHK(ensure(t, B_DINFO, n, &info));
    CHK(ensure(t, B_DBEFORE, n, &before));
    CHK(ensure(t, B_DKEEP, n, &keep));
    launch_defer_check(t->pool, t->prm, defer, n, info, before, keep, st);
    std::vector<int> hdef(n), hkeep(n), hbef(n);
    std::vector<int4> hinf(n);
    HIPCHK(hipMemcpyAsync(hdef.data(), defer, n * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hkeep.data(), keep, n * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hinf.data(), info, n * sizeof(int4), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hbef.data(), before, n * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCHK(wait_stream(t));
    // take due entries in list order while the point budget lasts (at least one)
    std::vector<int4> take;
    std::vector<int> take_before, rest;
    long long budget = max_points > 0 ? max_points : LLONG_MAX, used = 0;
    for (int i = 0; i < n; ++i) {
        if (!hkeep[i]) continue;
        if (!take.empty() && used + hinf[i].y > budget) {
            rest.push_back(hdef[i]);
            continue;
        }
        take.push_back(hinf[i]);
        take_before.push_back(hbef[i]);
        used += hinf[i].y;
    }
    // every processed entry leaves the list; entries over the budget stay listed
    launch_defer_clear(t->pool, defer, n, st);
    int cnt = (int)rest.size();
    if (cnt > 0) {
        CHK(stage_copy(t, defer, rest.data(), cnt * sizeof(int)));
        hipLaunchKernelGGL(k_link_defer_flags, dim3(nblk(cnt)), dim3(BS), 0, st, t->pool, defer, cnt);
    }
    CHK(stage_copy(t, &t->d_hdr->defer_count, &cnt, sizeof(int)));
    HIPCHK(wait_stream(t));
    if (take.empty()) return IKD_OK;
    CHK(rebuild_listed(t, take, take_before));
    t->rebuilds += (long long)take.size();
    // the ancestors still hold the pre-rebuild counts: refit the paths above
    std::vector<int> seeds;
    for (const int4& e : take) seeds.push_back(e.y > 0 ? e.x : e.w);
    int* dseeds;
    CHK(ensure(t, B_DKEEP, (long long)seeds.size(), &dseeds));
    CHK(stage_copy(t, dseeds, seeds.data(), seeds.size() * sizeof(int)));
    CHK(sync_hdr(t));
    CHK(refit_begin(t));
    CHK(refit_mark(t, dseeds, nullptr, (int)seeds.size(), 0));
    CHK(refit_finish(t));
    CHK(rebuild_tops(t));
    CHK(read_root(t));
    if (out_rebuilt) *out_rebuilt = (int32_t)take.size();
    return IKD_OK;
}

int ikd_set_tuning(ikd_tree_t t, int key, int value) {
    CHECK_T();
    switch (key) {
        case IKD_TUNE_DEFER_REBUILD:
            if (value < 0) return IKD_ERR_ARG;
            t->prm.defer_points = value;
            return IKD_OK;
        case IKD_TUNE_QUERY_ORDER:
            t->query_order = value ? 1 : 0;
            return IKD_OK;
        case IKD_TUNE_KNN_STACK:
            if (value < 0 || value > MAX_DEPTH) return IKD_ERR_ARG;
            t->knn_stack = value;
            return IKD_OK;
        case IKD_TUNE_SMALL_BUILD:
            t->small_build = value ? 1 : 0;
            return IKD_OK;
        case IKD_TUNE_LANES_PER_QUERY:
            if (value < -1 || value > 16 || (value > 0 && (value & (value - 1)))) return IKD_ERR_ARG;
            t->knn_lanes = value;
            return IKD_OK;
        case IKD_TUNE_RECORD_REMOVED:
            t->record_removed = value ? 1 : 0;
            return IKD_OK;
        case IKD_TUNE_FOLD_LANES:
            if (!(value == 0 || value == 1 || value == 4 || value == 8 || value == 16)) return IKD_ERR_ARG;
            t->fold_lanes = value;
            return IKD_OK;
        case IKD_TUNE_KNN_STAGE:
            if (value < 0 || value > 12) return IKD_ERR_ARG;
            t->knn_stage = value;
            return IKD_OK;
        case IKD_TUNE_KNN_COMPACT:
            t->knn_compact = value ? 1 : 0;
            return IKD_OK;
        default:
            return IKD_ERR_ARG;
    }
}

int ikd_pool_stats(ikd_tree_t t, int64_t* used, int64_t* capacity, int32_t* max_depth, int64_t* rebuilds) {
    CHECK_T();
    CHK(sync_hdr(t));
    if (used) *used = t->pool_used;
    if (capacity) *capacity = t->cap;
    if (max_depth) *max_depth = t->max_depth;
    if (rebuilds) *rebuilds = t->rebuilds;
    return IKD_OK;
}

int ikd_acquire_removed_points(ikd_tree_t t, float* out_xyz, uint32_t* out_ids, int64_t capacity, int64_t* out_n) {
    CHECK_T();
    t->hdr_valid = false;
    if (!out_n || capacity < 0 || (capacity > 0 && !out_xyz)) return IKD_ERR_ARG;
    CHK(sync_hdr(t));
    const long long n = t->h_hdr->removed_count;
    *out_n = n;
    if (n > capacity) return IKD_ERR_CAPACITY;
    if (n > 0) {
        float* dx;
        uint32_t* di;
        CHK(ensure(t, B_IO0, 3 * n, &dx));
        CHK(ensure(t, B_IO1, n, &di));
        hipLaunchKernelGGL(k_unpack_points, dim3(nblk(n)), dim3(BS), 0, t->st, t->removed, (int)n, dx, di);
        HIPCHK(hipMemcpyAsync(out_xyz, dx, 3 * n * sizeof(float), hipMemcpyDeviceToHost, t->st));
        if (out_ids) HIPCHK(hipMemcpyAsync(out_ids, di, n * sizeof(uint32_t), hipMemcpyDeviceToHost, t->st));
    }
    launch_fill_i32(&t->d_hdr->removed_count, 0, 1, t->st);
    return sync_hdr(t);
}

int ikd_get_counters(ikd_tree_t t, ikd_counters_t* out) {
    CHECK_T();
    if (!out) return IKD_ERR_ARG;
    *out = t->ctr;
    out->rebuilds = t->rebuilds;
    return IKD_OK;
}

int ikd_device_bytes(ikd_tree_t t, int64_t* out_bytes) {
    CHECK_T();
    if (!out_bytes) return IKD_ERR_ARG;
    int64_t b = (int64_t)sizeof(DevStatus);
    b += t->cap * (int64_t)(sizeof(NodeHot) + sizeof(uint32_t) + sizeof(int32_t) + sizeof(int4) + sizeof(uint32_t));
    for (const DevBuf& d : t->b) b += (int64_t)d.bytes;
    b += t->removed_cap * (int64_t)sizeof(float4);
    *out_bytes = b;
    return IKD_OK;
}

const char* ikd_last_error(ikd_tree_t t) {
    if (!t) return "null tree handle";
    return t->err.c_str();
}

// ------------------------------------------------------------ multi-tree batches
enum BatchBuf { BB_TMP = 0, BB_TREES, BB_THR, BB_NEXT, BB_FLAG, BB_CNT, BB_POOLS, BB_TCNT, BB_TOFF, BB_TASKS,
                BB_ICNT, BB_COUNT };
static_assert(BB_COUNT <= 16, "batch scratch table");

// ------------------------------------- allocation-free multi-tree kNN scratch
// ikd_batch_nearest_search_noalloc keeps everything beyond its inputs and
// outputs in this array: 8 persistent-lane query counters and the overflow
// flag, the tree table, and the per-query max_dist thresholds.  Each GPU holds
// its own copy for the life of the process; every call on a GPU shares it.
constexpr size_t NA_CTRL_BYTES = 256;
constexpr size_t NA_TREES_OFF = NA_CTRL_BYTES;
constexpr size_t NA_THR_OFF = NA_TREES_OFF + ((IKD_NOALLOC_MAX_TREES * sizeof(TreeRef) + 255) & ~(size_t)255);
constexpr size_t NA_BYTES = NA_THR_OFF + (size_t)IKD_NOALLOC_MAX_QUERIES * sizeof(float);
static __device__ __align__(256) char scratch[NA_BYTES];

// The device address of `scratch` on the current device, resolved once per
// device (host code never uses the array's name as an address).
static char* noalloc_scratch(int dev) {
    static std::mutex mu;
    static char* ptr[64];
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> g(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        ptr[dev] = static_cast<char*>(p);
    }
    return ptr[dev];
}

// The tree table travels as a kernel argument and is stored into the scratch.
struct TreeTable {
    TreeRef t[IKD_NOALLOC_MAX_TREES];
};
__global__ void k_store_trees(TreeTable tab, int n, TreeRef* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = tab.t[i];
}

// shared prologue of batch launches: compact over-deep trees, make the batch
// stream wait for the trees' pending work, report the deepest tree and the
// largest pool
static int batch_begin(ikd_batch_t b, const ikd_tree_t* trees, int ntrees, int* depth, long long* steps) {
    auto* t = b;
    *depth = 0;
    *steps = 16;
    for (int i = 0; i < ntrees; ++i) {
        ikd_tree_t tr = trees[i];
        if (!tr) {
            b->err = "batch: null tree handle";
            return IKD_ERR_ARG;
        }
        std::lock_guard<std::recursive_mutex> g(tr->mu);
        if (tr->max_depth + 2 > MAX_DEPTH) {
            tr->hdr_valid = false;
            int rc = compact_if_needed(tr, true);
            if (rc != IKD_OK) {
                b->err = "batch: " + tr->err;
                return rc;
            }
        }
        *depth = std::max(*depth, tr->max_depth);
        *steps = std::max(*steps, tr->pool_used + 16);
        HIPCHK(hipEventRecord(tr->ev, tr->st));
        HIPCHK(hipStreamWaitEvent(b->st, tr->ev, 0));
    }
    return IKD_OK;
}

static int batch_end(ikd_batch_t b, const ikd_tree_t* trees, int ntrees) {
    auto* t = b;
    HIPCHK(hipEventRecord(b->ev, b->st));
    for (int i = 0; i < ntrees; ++i) {
        std::lock_guard<std::recursive_mutex> g(trees[i]->mu);
        HIPCHK(hipStreamWaitEvent(trees[i]->st, b->ev, 0));
    }
    return IKD_OK;
}


int ikd_batch_create(ikd_batch_t* out) {
    if (!out) return IKD_ERR_ARG;
    *out = nullptr;
    ikd_batch_t b = new ikd_batch_s();
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        delete b;
        return IKD_ERR_HIP;
    }
    b->mp = library_pool(dev);
    if (hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(&b->h_flag, 2 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
        ikd_batch_destroy(b);
        return IKD_ERR_HIP;
    }
    // set-up of the allocation-free entry point on this device (its absence
    // only disables that entry point)
    b->na_scratch = noalloc_scratch(dev);
    int* flag;
    int rc = ensure(b, BB_FLAG, 1, &flag);
    if (rc == IKD_OK) {
        launch_fill_i32(flag, 0, 1, b->st);
        if (hipStreamSynchronize(b->st) != hipSuccess) rc = IKD_ERR_HIP;
    }
    if (rc != IKD_OK) {
        ikd_batch_destroy(b);
        return rc;
    }
    *out = b;
    return IKD_OK;
}

int ikd_batch_destroy(ikd_batch_t b) {
    if (!b) return IKD_ERR_ARG;
    if (b->st) {
        for (auto& d : b->b) dev_free(b, d.p);
        hipStreamSynchronize(b->st);
    }
    stage_release(b);
    if (b->h_flag) hipHostFree(b->h_flag);
    if (b->ev) hipEventDestroy(b->ev);
    if (b->st) hipStreamDestroy(b->st);
    delete b;
    return IKD_OK;
}

int ikd_batch_set_tuning(ikd_batch_t b, int key, int value) {
    if (!b) return IKD_ERR_ARG;
    std::lock_guard<std::mutex> batch_lock_(b->mu);
    switch (key) {
        case IKD_TUNE_KNN_STACK:
            if (value < 0 || value > MAX_DEPTH) return IKD_ERR_ARG;
            b->knn_stack = value;
            return IKD_OK;
        case IKD_TUNE_LANES_PER_QUERY:
            if (value < -1 || value > 16 || (value > 0 && (value & (value - 1)))) return IKD_ERR_ARG;
            b->knn_lanes = value;
            return IKD_OK;
        case IKD_TUNE_KNN_COMPACT:
            b->knn_compact = value ? 1 : 0;
            return IKD_OK;
        default:
            return IKD_ERR_ARG;
    }
}

int ikd_batch_nearest_search_device(ikd_batch_t b, const ikd_tree_t* trees, int ntrees, const int32_t* d_query_tree,
                                    const float* d_queries, int64_t nq, const int32_t* d_k, int k_max,
                                    const double* d_max_dist, double max_dist, float* d_out_xyz, float* d_out_d2,
                                    uint32_t* d_out_ids, int32_t* d_out_counts) {
    if (!b) return IKD_ERR_ARG;
    std::lock_guard<std::mutex> batch_lock_(b->mu);
    if (!trees || ntrees <= 0 || nq < 0 || nq >= (1ll << 31)) {
        b->err = "batch: bad tree list or query count";
        return IKD_ERR_ARG;
    }
    if (k_max < 1) {
        b->err = "batch: k_max must be >= 1";
        return IKD_ERR_ARG;
    }
    if (k_max > 64) {
        b->err = "batch: k above 64 is not supported";
        return IKD_ERR_UNSUPPORTED;
    }
    if (nq == 0) return IKD_OK;
    if (!d_query_tree || !d_queries || !d_out_xyz || !d_out_d2 || !d_out_ids || !d_out_counts) {
        b->err = "batch: null device pointer";
        return IKD_ERR_ARG;
    }
    auto* t = b;  // HIPCHK reports into b->err
    int depth;
    long long steps;
    CHK(batch_begin(b, trees, ntrees, &depth, &steps));
    std::vector<TreeRef> refs((size_t)ntrees);
    bool small_pools = true;
    for (int i = 0; i < ntrees; ++i) {
        refs[i] = TreeRef{trees[i]->pool.hot, trees[i]->pool.ids};
        small_pools = small_pools && trees[i]->cap <= (1ll << 24);
    }
    TreeRef* d_refs;
    int *next, *flag;
    CHK(ensure(b, BB_TREES, ntrees, &d_refs));
    CHK(ensure(b, BB_NEXT, 8, &next));
    CHK(ensure(b, BB_FLAG, 1, &flag));
    CHK(stage_copy(b, d_refs, refs.data(), ntrees * sizeof(TreeRef)));
    KnnArgs a{};
    a.trees = d_refs;
    a.compact_stack = (small_pools && b->knn_compact) ? 1 : 0;
    a.qtree = d_query_tree;
    a.ntrees = ntrees;
    a.q = d_queries;
    a.nq = (int)nq;
    a.k = k_max;
    a.qk = d_k;
    a.thr = max_dist_threshold(max_dist);
    if (d_max_dist) {
        float* thr;
        CHK(ensure(b, BB_THR, nq, &thr));
        launch_thresholds(d_max_dist, (int)nq, thr, b->st);
        a.qthr = thr;
    }
    a.out_stride = k_max;
    a.out_xyz = d_out_xyz;
    a.out_d2 = d_out_d2;
    a.out_ids = d_out_ids;
    a.out_counts = d_out_counts;
    a.overflow = flag;
    a.max_steps = (int)std::min<long long>(steps, INT_MAX);
    int stack = std::min(std::max(depth + 2, b->knn_stack), MAX_DEPTH);
    int lanes = b->knn_lanes >= 0 ? b->knn_lanes : auto_lanes((int)nq);
    if (launch_knn(a, k_max, lanes, stack, b->st, next, 256 * 10) != 0) {
        b->err = "batch: unsupported launch shape";
        return IKD_ERR_UNSUPPORTED;
    }
    HIPCHK(hipGetLastError());
    return batch_end(b, trees, ntrees);
}

int ikd_batch_nearest_search_noalloc(ikd_batch_t b, const ikd_tree_t* trees, int ntrees,
                                     const int32_t* d_query_tree, const float* d_queries, int64_t nq,
                                     const int32_t* d_k, int k_max, const double* d_max_dist, double max_dist,
                                     float* d_out_xyz, float* d_out_d2, uint32_t* d_out_ids,
                                     int32_t* d_out_counts) {
    if (!b) return IKD_ERR_ARG;
    std::lock_guard<std::mutex> batch_lock_(b->mu);
    auto* t = b;  // HIPCHK reports into b->err
    if (!trees || ntrees <= 0 || nq < 0 || nq >= (1ll << 31)) {
        b->err = "batch (no-alloc): bad tree list or query count";
        return IKD_ERR_ARG;
    }
    if (ntrees > IKD_NOALLOC_MAX_TREES) {
        b->err = "batch (no-alloc): more than IKD_NOALLOC_MAX_TREES trees";
        return IKD_ERR_CAPACITY;
    }
    if (d_max_dist && nq > IKD_NOALLOC_MAX_QUERIES) {
        b->err = "batch (no-alloc): per-query max_dist for more than IKD_NOALLOC_MAX_QUERIES queries";
        return IKD_ERR_CAPACITY;
    }
    if (k_max < 1 || k_max > 64) {
        b->err = "batch (no-alloc): k_max must be in 1..64";
        return k_max < 1 ? IKD_ERR_ARG : IKD_ERR_UNSUPPORTED;
    }
    if (nq == 0) return IKD_OK;
    if (!d_query_tree || !d_queries || !d_out_xyz || !d_out_d2 || !d_out_ids || !d_out_counts) {
        b->err = "batch (no-alloc): null device pointer";
        return IKD_ERR_ARG;
    }
    if (!b->na_scratch) {
        b->err = "batch (no-alloc): the scratch array's device address is unavailable";
        return IKD_ERR_HIP;
    }
    int* ctrl = reinterpret_cast<int*>(b->na_scratch);  // [0, 8) query counters, [8] overflow flag
    TreeRef* table = reinterpret_cast<TreeRef*>(b->na_scratch + NA_TREES_OFF);
    float* thr = reinterpret_cast<float*>(b->na_scratch + NA_THR_OFF);
    TreeTable tab;
    int depth = 0;
    long long steps = 16;
    bool small_pools = true;
    for (int i = 0; i < ntrees; ++i) {
        ikd_tree_t tr = trees[i];
        if (!tr) {
            b->err = "batch (no-alloc): null tree handle";
            return IKD_ERR_ARG;
        }
        std::lock_guard<std::recursive_mutex> g(tr->mu);
        if (tr->max_depth + 2 > MAX_DEPTH) {
            // the compaction that fixes this allocates: leave it to any update call
            b->err = "batch (no-alloc): a tree is deeper than the traversal stacks; update it once first";
            return IKD_ERR_UNSUPPORTED;
        }
        depth = std::max(depth, tr->max_depth);
        steps = std::max(steps, tr->pool_used + 16);
        small_pools = small_pools && tr->cap <= (1ll << 24);
        tab.t[i] = TreeRef{tr->pool.hot, tr->pool.ids};
        HIPCHK(hipEventRecord(tr->ev, tr->st));
        HIPCHK(hipStreamWaitEvent(b->st, tr->ev, 0));
    }
    hipLaunchKernelGGL(k_store_trees, dim3(1), dim3(IKD_NOALLOC_MAX_TREES), 0, b->st, tab, ntrees, table);
    launch_fill_i32(ctrl + 8, 0, 1, b->st);
    KnnArgs a{};
    a.trees = table;
    a.compact_stack = (small_pools && b->knn_compact) ? 1 : 0;
    a.qtree = d_query_tree;
    a.ntrees = ntrees;
    a.q = d_queries;
    a.nq = (int)nq;
    a.k = k_max;
    a.qk = d_k;
    a.thr = max_dist_threshold(max_dist);
    if (d_max_dist) {
        launch_thresholds(d_max_dist, (int)nq, thr, b->st);
        a.qthr = thr;
    }
    a.out_stride = k_max;
    a.out_xyz = d_out_xyz;
    a.out_d2 = d_out_d2;
    a.out_ids = d_out_ids;
    a.out_counts = d_out_counts;
    a.overflow = ctrl + 8;
    a.max_steps = (int)std::min<long long>(steps, INT_MAX);
    const int stack = std::min(std::max(depth + 2, b->knn_stack), MAX_DEPTH);
    const int lanes = b->knn_lanes >= 0 ? b->knn_lanes : auto_lanes((int)nq);
    if (launch_knn(a, k_max, lanes, stack, b->st, ctrl, 256 * 10) != 0) {
        b->err = "batch (no-alloc): unsupported launch shape";
        return IKD_ERR_UNSUPPORTED;
    }
    HIPCHK(hipGetLastError());
    b->na_flag = ctrl + 8;
    return batch_end(b, trees, ntrees);
}

// ------------------------------------------------------- recorded kNN launch
// One multi-tree kNN launch recorded into a HIP graph: the graph holds only
// kernels (the overflow flag's clear, per-query thresholds, the kNN), reads
// the tree table from a device buffer of its own, and has its LDS stack fixed
// at recording.  Each replay refreshes the table from the trees' current node
// pools (they move when a pool grows), orders itself after the trees' pending
// work and orders the trees' later work after itself, as the entry points do.
struct ikd_knn_graph_s {
    ikd_batch_t b = nullptr;
    std::vector<ikd_tree_t> trees;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    TreeRef* d_table = nullptr;
    TreeRef* h_table = nullptr;     // pinned staging of the table refresh
    hipEvent_t ev_table = nullptr;  // behind the last refresh's copy
    bool table_pending = false;
    float* d_thr = nullptr;         // per-query thresholds (per-query max_dist)
    int* d_ctrl = nullptr;          // persistent-lane query counters
    int stack = 0;                  // LDS stack entries per lane of the recording
};

static void knn_graph_free(ikd_knn_graph_t g) {
    if (!g) return;
    if (g->table_pending) hipEventSynchronize(g->ev_table);
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    if (g->b && g->b->st) hipStreamSynchronize(g->b->st);  // no replay still reads the buffers
    if (g->d_table) hipFree(g->d_table);
    if (g->h_table) hipHostFree(g->h_table);
    if (g->d_thr) hipFree(g->d_thr);
    if (g->d_ctrl) hipFree(g->d_ctrl);
    if (g->ev_table) hipEventDestroy(g->ev_table);
    delete g;
}

// Levels a tree may grow by before a replay asks for a new recording.
constexpr int GRAPH_DEPTH_HEADROOM = 8;

int ikd_batch_knn_graph_create(ikd_batch_t b, const ikd_tree_t* trees, int ntrees, const int32_t* d_query_tree,
                               const float* d_queries, int64_t nq, const int32_t* d_k, int k_max,
                               const double* d_max_dist, double max_dist, float* d_out_xyz, float* d_out_d2,
                               uint32_t* d_out_ids, int32_t* d_out_counts, ikd_knn_graph_t* out) {
    if (!b || !out) return IKD_ERR_ARG;
    *out = nullptr;
    std::lock_guard<std::mutex> batch_lock_(b->mu);
    auto* t = b;  // HIPCHK reports into b->err
    if (!trees || ntrees <= 0 || nq <= 0 || nq >= (1ll << 31)) {
        b->err = "recorded kNN: bad tree list or query count";
        return IKD_ERR_ARG;
    }
    if (k_max < 1 || k_max > 64) {
        b->err = "recorded kNN: k_max must be in 1..64";
        return k_max < 1 ? IKD_ERR_ARG : IKD_ERR_UNSUPPORTED;
    }
    if (!d_query_tree || !d_queries || !d_out_xyz || !d_out_d2 || !d_out_ids || !d_out_counts) {
        b->err = "recorded kNN: null device pointer";
        return IKD_ERR_ARG;
    }
    int depth = 0;
    for (int i = 0; i < ntrees; ++i) {
        if (!trees[i]) {
            b->err = "recorded kNN: null tree handle";
            return IKD_ERR_ARG;
        }
        std::lock_guard<std::recursive_mutex> g(trees[i]->mu);
        if (trees[i]->max_depth + 2 > MAX_DEPTH) {
            b->err = "recorded kNN: a tree is deeper than the traversal stacks; update it once first";
            return IKD_ERR_UNSUPPORTED;
        }
        depth = std::max(depth, trees[i]->max_depth);
    }
    ikd_knn_graph_t g = new ikd_knn_graph_s();
    g->b = b;
    g->trees.assign(trees, trees + ntrees);
    g->stack = std::min(std::max(depth + 2 + GRAPH_DEPTH_HEADROOM, b->knn_stack), MAX_DEPTH);
    int* flag;
    if (ensure(b, BB_FLAG, 1, &flag) != IKD_OK || hipMalloc(&g->d_table, ntrees * sizeof(TreeRef)) != hipSuccess ||
        hipHostMalloc(&g->h_table, ntrees * sizeof(TreeRef), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&g->d_ctrl, 8 * sizeof(int)) != hipSuccess ||
        (d_max_dist && hipMalloc(&g->d_thr, nq * sizeof(float)) != hipSuccess) ||
        hipEventCreateWithFlags(&g->ev_table, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        knn_graph_free(g);
        b->err = "recorded kNN: allocation failed";
        return IKD_ERR_NOMEM;
    }
    KnnArgs a{};
    a.trees = g->d_table;
    a.qtree = d_query_tree;
    a.ntrees = ntrees;
    a.q = d_queries;
    a.nq = (int)nq;
    a.k = k_max;
    a.qk = d_k;
    a.thr = max_dist_threshold(max_dist);
    a.qthr = d_max_dist ? g->d_thr : nullptr;
    a.out_stride = k_max;
    a.out_xyz = d_out_xyz;
    a.out_d2 = d_out_d2;
    a.out_ids = d_out_ids;
    a.out_counts = d_out_counts;
    a.overflow = flag;
    // the visit bound only stops a walk on a corrupt tree; the recording must
    // outlive the trees' growth
    a.max_steps = 1 << 30;
    const int lanes = b->knn_lanes >= 0 ? b->knn_lanes : auto_lanes((int)nq);
    if (hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        knn_graph_free(g);
        b->err = "recorded kNN: stream capture could not begin";
        return IKD_ERR_HIP;
    }
    launch_fill_i32(flag, 0, 1, b->st);
    if (d_max_dist) launch_thresholds(d_max_dist, (int)nq, g->d_thr, b->st);
    const int lrc = launch_knn(a, k_max, lanes, g->stack, b->st, g->d_ctrl, 256 * 10);
    const hipError_t le = hipGetLastError();
    const hipError_t ce = hipStreamEndCapture(b->st, &g->graph);
    if (lrc != 0 || le != hipSuccess || ce != hipSuccess || !g->graph ||
        hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGetLastError();
        knn_graph_free(g);
        b->err = lrc != 0 ? "recorded kNN: unsupported launch shape" : "recorded kNN: recording the launch failed";
        return lrc != 0 ? IKD_ERR_UNSUPPORTED : IKD_ERR_HIP;
    }
    *out = g;
    return IKD_OK;
}

int ikd_batch_knn_graph_launch(ikd_knn_graph_t g) {
    if (!g) return IKD_ERR_ARG;
    ikd_batch_t b = g->b;
    std::lock_guard<std::mutex> batch_lock_(b->mu);
    auto* t = b;
    // the previous refresh's copy has read its staging before it is rewritten
    if (g->table_pending) HIPCHK(hipEventSynchronize(g->ev_table));
    g->table_pending = false;
    const int n = (int)g->trees.size();
    for (int i = 0; i < n; ++i) {
        ikd_tree_t tr = g->trees[i];
        std::lock_guard<std::recursive_mutex> lk(tr->mu);
        if (tr->max_depth + 2 > g->stack) {
            b->err = "recorded kNN: a tree grew deeper than the recorded launch allows; record it again";
            return IKD_ERR_UNSUPPORTED;
        }
        g->h_table[i] = TreeRef{tr->pool.hot, tr->pool.ids};
        HIPCHK(hipEventRecord(tr->ev, tr->st));
        HIPCHK(hipStreamWaitEvent(b->st, tr->ev, 0));
    }
    HIPCHK(hipMemcpyAsync(g->d_table, g->h_table, n * sizeof(TreeRef), hipMemcpyHostToDevice, b->st));
    HIPCHK(hipEventRecord(g->ev_table, b->st));
    g->table_pending = true;
    HIPCHK(hipGraphLaunch(g->exec, b->st));
    return batch_end(b, g->trees.data(), n);
}

int ikd_batch_knn_graph_destroy(ikd_knn_graph_t g) {
    if (!g) return IKD_ERR_ARG;
    std::lock_guard<std::mutex> batch_lock_(g->b->mu);
    knn_graph_free(g);
    return IKD_OK;
}

int ikd_batch_radius_search_device(ikd_batch_t b, const ikd_tree_t* trees, int ntrees, const int32_t* d_query_tree,
                                   const float* d_centers, const float* d_radii, int64_t nq, int64_t* d_out_offsets,
                                   float* d_out_xyz, uint32_t* d_out_ids, int64_t capacity, int64_t* out_total) {
    if (!b) return IKD_ERR_ARG;
    std::lock_guard<std::mutex> batch_lock_(b->mu);
    if (!trees || ntrees <= 0 || nq < 0 || nq >= (1ll << 31) || !out_total) {
        b->err = "batch: bad tree list, query count or total pointer";
        return IKD_ERR_ARG;
    }
    *out_total = 0;
    if (nq == 0) return IKD_OK;
    if (!d_query_tree || !d_centers || !d_radii || !d_out_offsets) {
        b->err = "batch: null device pointer";
        return IKD_ERR_ARG;
    }
    auto* t = b;
    int depth;
    long long steps;
    CHK(batch_begin(b, trees, ntrees, &depth, &steps));
    std::vector<Pool> pools((size_t)ntrees);
    for (int i = 0; i < ntrees; ++i) pools[i] = trees[i]->pool;
    Pool* d_pools;
    int* flag;
    CHK(ensure(b, BB_POOLS, ntrees, &d_pools));
    CHK(ensure(b, BB_FLAG, 1, &flag));
    CHK(stage_copy(b, d_pools, pools.data(), ntrees * sizeof(Pool)));
    RangeArgs a{};
    a.pools = d_pools;
    a.qtree = d_query_tree;
    a.ntrees = ntrees;
    a.centers = d_centers;
    a.radii = d_radii;
    a.nq = (int)nq;
    a.mode = 1;
    a.overflow = flag;
    a.max_steps = steps;
    a.stack = std::min(depth + 2, MAX_DEPTH);
    long long total = 0, ntask = 0;
    static const int slots[5] = {BB_TCNT, BB_TOFF, BB_TASKS, BB_CNT, BB_ICNT};
    CHK(range_plan(b, a, (int)nq, reinterpret_cast<long long*>(d_out_offsets), slots, std::max<long long>(steps, 1),
                   &total, &ntask));
    *out_total = total;
    if (total < 0) {
        b->err = "batch: device result count out of range";
        return IKD_ERR_INTERNAL;
    }
    if (d_out_xyz && d_out_ids && capacity >= total && total > 0) {
        a.out_xyz = d_out_xyz;
        a.out_ids = d_out_ids;
        launch_range_item_fill(a, ntask, a.item_offsets, b->st);
        HIPCHK(hipGetLastError());
    }
    CHK(batch_end(b, trees, ntrees));
    if (d_out_xyz && capacity < total) {
        b->err = "batch: output capacity below the result count";
        return IKD_ERR_CAPACITY;
    }
    return IKD_OK;
}

int ikd_batch_synchronize(ikd_batch_t b) {
    if (!b) return IKD_ERR_ARG;
    std::lock_guard<std::mutex> batch_lock_(b->mu);
    auto* t = b;
    int* flag;
    CHK(ensure(b, BB_FLAG, 1, &flag));
    b->h_flag[1] = 0;
    HIPCHK(hipMemcpyAsync(b->h_flag, flag, sizeof(int), hipMemcpyDeviceToHost, b->st));
    if (b->na_flag) HIPCHK(hipMemcpyAsync(b->h_flag + 1, b->na_flag, sizeof(int), hipMemcpyDeviceToHost, b->st));
    HIPCHK(hipStreamSynchronize(b->st));
    if (b->h_flag[0] || b->h_flag[1]) {
        launch_fill_i32(flag, 0, 1, b->st);
        HIPCHK(hipStreamSynchronize(b->st));
        b->err = "batch: device traversal stack overflow";
        return IKD_ERR_INTERNAL;
    }
    return IKD_OK;
}

int ikd_batch_stream(ikd_batch_t b, void** out_stream) {
    if (!b || !out_stream) return IKD_ERR_ARG;
    *out_stream = (void*)b->st;
    return IKD_OK;
}

const char* ikd_batch_last_error(ikd_batch_t b) {
    if (!b) return "null batch handle";
    return b->err.c_str();
}

}  // extern "C"

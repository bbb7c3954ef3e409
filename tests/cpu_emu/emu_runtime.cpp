// CPU emulation runtime for the tree library's debugging build (see hip_emu.h).
//
// A workgroup runs on the launching host thread as one fiber (ucontext) per
// GPU thread: a fiber runs until it reaches __syncthreads or returns, and the
// barrier opens once every fiber still running has reached it.  Workgroups of
// a launch run one after another; launches from several host threads are
// serialized by a lock.
#include "hip_emu.h"

#include <setjmp.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <string>
#include <set>
#include <typeindex>
#include <vector>

thread_local dim3 threadIdx;
thread_local dim3 blockIdx;
dim3 blockDim;
dim3 gridDim;

// Minimal x86-64 context switch: save the callee-saved registers and the stack
// pointer of the running context into *save_sp, resume the one at load_sp.
extern "C" void emu_switch(void** save_sp, void* load_sp);
asm(R"(
.text
.globl emu_switch
.type emu_switch,@function
emu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
.size emu_switch, .-emu_switch
)");

namespace {

constexpr int MAX_THREADS = 1024;
constexpr size_t FIBER_STACK = 256 * 1024;

enum { RUNNABLE = 0, AT_BARRIER = 1, DONE = 2 };

struct Fiber {
    void* sp = nullptr;
    char* stack = nullptr;
    int state = DONE;
};

struct BlockState {
    const std::function<void()>* body = nullptr;
    int nthreads = 0;
    int cur = -1;
    void* sched_sp = nullptr;
    Fiber fib[MAX_THREADS];
    int vote[MAX_THREADS];
    int ival[MAX_THREADS];
    float fval[MAX_THREADS];
    void* dyn = nullptr;
};
BlockState g_blk;
std::mutex g_launch_mu;
// Launch sites whose kernels use workgroup barriers run as fibers; the others
// run each GPU thread as a plain call.  A site is moved to fibers the first
// time a directly-run thread reaches a barrier (before any other thread of
// its workgroup ran).
std::set<std::type_index> g_fiber_sites;
bool g_direct = false;
int g_direct_tid = -1;
jmp_buf g_direct_jb;

void set_tid(int i) {
    threadIdx = dim3(i % blockDim.x, (i / blockDim.x) % blockDim.y, i / (blockDim.x * blockDim.y));
}

void fiber_entry() {
    const int i = g_blk.cur;
    (*g_blk.body)();
    g_blk.fib[i].state = DONE;
    emu_switch(&g_blk.fib[i].sp, g_blk.sched_sp);  // never resumed
    std::abort();
}

void run_block() {
    const int n = g_blk.nthreads;
    for (int i = 0; i < n; ++i) {
        Fiber& f = g_blk.fib[i];
        if (!f.stack) f.stack = static_cast<char*>(std::malloc(FIBER_STACK));
        // initial frame: six zeroed registers, then fiber_entry as the return
        // address, with the stack 16-byte aligned at fiber_entry's entry
        uintptr_t top = (reinterpret_cast<uintptr_t>(f.stack) + FIBER_STACK) & ~uintptr_t(15);
        void** fr = reinterpret_cast<void**>(top - 64);
        for (int r = 0; r < 6; ++r) fr[r] = nullptr;
        fr[6] = reinterpret_cast<void*>(&fiber_entry);
        fr[7] = nullptr;
        f.sp = fr;
        f.state = RUNNABLE;
    }
    while (true) {
        bool any = false;
        for (int i = 0; i < n; ++i) {
            Fiber& f = g_blk.fib[i];
            if (f.state != RUNNABLE) continue;
            g_blk.cur = i;
            set_tid(i);
            emu_switch(&g_blk.sched_sp, f.sp);
        }
        for (int i = 0; i < n; ++i) {
            if (g_blk.fib[i].state == AT_BARRIER) {
                g_blk.fib[i].state = RUNNABLE;
                any = true;
            }
        }
        if (!any) break;
    }
    g_blk.cur = -1;
}

}  // namespace

namespace ikd {
// dynamic LDS of the current workgroup, sized exactly as the launch asked
void* emu_dyn_lds() { return g_blk.dyn; }
}  // namespace ikd

void __syncthreads() {
    if (g_direct) longjmp(g_direct_jb, 1);
    const int i = g_blk.cur;
    g_blk.fib[i].state = AT_BARRIER;
    emu_switch(&g_blk.fib[i].sp, g_blk.sched_sp);
    set_tid(i);
}

static bool lane_active(int idx) { return idx < g_blk.nthreads && g_blk.fib[idx].state != DONE; }

unsigned long long __ballot(int pred) {
    const int t = g_blk.cur;
    g_blk.vote[t] = pred ? 1 : 0;
    __syncthreads();
    const int w0 = (t / 64) * 64;
    unsigned long long m = 0;
    for (int l = 0; l < 64; ++l)
        if (lane_active(w0 + l) && g_blk.vote[w0 + l]) m |= 1ull << l;
    __syncthreads();
    return m;
}

int emu_shfl_i(int v, int src_lane) {
    const int t = g_blk.cur;
    g_blk.ival[t] = v;
    __syncthreads();
    const int src = (t / 64) * 64 + src_lane;
    int r = src < g_blk.nthreads ? g_blk.ival[src] : v;
    __syncthreads();
    return r;
}

float emu_shfl_f(float v, int src_lane) {
    const int t = g_blk.cur;
    g_blk.fval[t] = v;
    __syncthreads();
    const int src = (t / 64) * 64 + src_lane;
    float r = src < g_blk.nthreads ? g_blk.fval[src] : v;
    __syncthreads();
    return r;
}

struct emu_node {
    dim3 grid, block;
    size_t shmem;
    std::function<void()> body;
};
struct emu_graph {
    std::vector<emu_node> nodes;
};
struct emu_graph_exec {
    std::vector<emu_node> nodes;
};
static thread_local emu_graph* t_capture = nullptr;

hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    if (t_capture) return hipErrorInvalidValue;
    t_capture = new emu_graph();
    return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) {
    if (!t_capture) return hipErrorInvalidValue;
    *g = t_capture;
    t_capture = nullptr;
    return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, void*, char*, size_t) {
    *e = new emu_graph_exec{g->nodes};
    return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (const emu_node& n : e->nodes) emu_launch(n.grid, n.block, n.shmem, n.body);
    return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}

static std::mutex g_tally_mu;
static std::map<std::string, long long> g_tally;
void emu_tally(const char* name) {
    std::lock_guard<std::mutex> g(g_tally_mu);
    ++g_tally[name];
}
extern "C" void emu_dispatch_reset() {
    std::lock_guard<std::mutex> g(g_tally_mu);
    g_tally.clear();
}
// "name count" lines into buf (truncated to len); returns the total count
extern "C" long long emu_dispatch_tally(char* buf, int len) {
    std::lock_guard<std::mutex> g(g_tally_mu);
    std::string out;
    long long total = 0;
    for (auto& kv : g_tally) {
        out += kv.first + " " + std::to_string(kv.second) + "\n";
        total += kv.second;
    }
    if (buf && len > 0) {
        std::snprintf(buf, (size_t)len, "%s", out.c_str());
    }
    return total;
}

void emu_launch(dim3 grid, dim3 block, size_t shmem, const std::function<void()>& body) {
    if (t_capture) {
        t_capture->nodes.push_back(emu_node{grid, block, shmem, body});
        return;
    }
    std::lock_guard<std::mutex> g(g_launch_mu);
    const int n = (int)(block.x * block.y * block.z);
    if (n <= 0 || n > MAX_THREADS) {
        std::fprintf(stderr, "emu: bad block size %d\n", n);
        std::abort();
    }
    gridDim = grid;
    blockDim = block;
    g_blk.body = &body;
    g_blk.nthreads = n;
    const std::type_index site(body.target_type());
    for (unsigned bz = 0; bz < grid.z; ++bz)
        for (unsigned by = 0; by < grid.y; ++by)
            for (unsigned bx = 0; bx < grid.x; ++bx) {
                blockIdx = dim3(bx, by, bz);
                g_blk.dyn = shmem ? std::malloc(shmem) : nullptr;
                if (shmem) std::memset(g_blk.dyn, 0xcd, shmem);
                bool fiber = g_fiber_sites.count(site) != 0;
                if (!fiber) {
                    g_direct = true;
                    for (int i = 0; i < n; ++i) {
                        g_direct_tid = i;
                        g_blk.cur = i;
                        set_tid(i);
                        if (setjmp(g_direct_jb) != 0) {
                            if (i != 0) {
                                std::fprintf(stderr, "emu: thread %d reached a barrier after threads of its "
                                                     "workgroup finished without one\n", i);
                                std::abort();
                            }
                            fiber = true;
                            g_fiber_sites.insert(site);
                            if (shmem) std::memset(g_blk.dyn, 0xcd, shmem);
                            break;
                        }
                        body();
                    }
                    g_direct = false;
                }
                if (fiber) run_block();
                std::free(g_blk.dyn);
                g_blk.dyn = nullptr;
            }
}

// ---- memory, streams, events ---------------------------------------------------
hipError_t hipMalloc(void** p, size_t n) {
    *p = std::malloc(n ? n : 1);
    if (!*p) return hipErrorOutOfMemory;
    std::memset(*p, 0xab, n);  // device memory starts as garbage
    return hipSuccess;
}
hipError_t hipFree(void* p) {
    std::free(p);
    return hipSuccess;
}
struct emu_pool { int dummy; };
hipError_t hipMemPoolCreate(hipMemPool_t* mp, const hipMemPoolProps*) {
    *mp = new emu_pool{0};
    return hipSuccess;
}
hipError_t hipMemPoolDestroy(hipMemPool_t mp) {
    delete mp;
    return hipSuccess;
}
hipError_t hipMemPoolSetAttribute(hipMemPool_t, hipMemPoolAttr, void*) { return hipSuccess; }
hipError_t hipMallocFromPoolAsync(void** p, size_t n, hipMemPool_t, hipStream_t) { return hipMalloc(p, n); }
hipError_t hipFreeAsync(void* p, hipStream_t) { return hipFree(p); }
hipError_t hipMallocAsync(void** p, size_t n, hipStream_t) { return hipMalloc(p, n); }
hipError_t hipGetDevice(int* dev) {
    *dev = 0;
    return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
    *v = 4;  // a few compute units: persistent launches stay small here
    return hipSuccess;
}
hipError_t hipHostMalloc(void** p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void* p) { return hipFree(p); }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
    if (n) std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
    emu_tally(k == hipMemcpyHostToDevice ? "copy:h2d" : k == hipMemcpyDeviceToHost ? "copy:d2h" : "copy:d2d");
    return hipMemcpy(d, s, n, k);
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
    emu_tally("memset");
    std::memset(d, v, n);
    return hipSuccess;
}
struct emu_stream { int dummy; };
struct emu_event { int dummy; };
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    *s = new emu_stream{0};
    return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
    delete s;
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) {
    emu_tally("wait:stream");
    return hipSuccess;
}
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) {
    *e = new emu_event{0};
    return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) {
    emu_tally("wait:event");
    return hipSuccess;
}
hipError_t hipGetSymbolAddress(void** p, const void* symbol) {
    *p = const_cast<void*>(symbol);
    return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "emulated HIP error"; }

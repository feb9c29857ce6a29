// CPU emulation runtime (see emu_hip.h) -- test infrastructure only.
#include "emu_hip.h"

#include <map>
#include <memory>

thread_local emu_dim3 threadIdx, blockIdx, blockDim, gridDim;

namespace emu {

uint64_t g_slots[1024];
static Barrier g_block;
static Barrier g_groups[11][1024];
static std::once_flag g_once;

static void init_barriers() {
    for (int l = 0; l <= 10; ++l)
        for (int g = 0; g < 1024; ++g) g_groups[l][g].count = 1 << l;
}

Barrier& group_barrier(int width, int group) { return g_groups[__builtin_ctz((unsigned)width)][group]; }
void syncthreads() { g_block.wait(); }

struct Pool {
    std::vector<std::thread> th;
    std::mutex m;
    std::condition_variable cv, cv_done;
    uint64_t gen = 0;
    int remaining = 0;
    const std::function<void()>* fn = nullptr;
    uint32_t bid = 0, gdim = 0;
    const uint32_t bdim;
    explicit Pool(uint32_t n) : bdim(n) {
        for (uint32_t i = 0; i < n; ++i) th.emplace_back([this, i] { worker(i); });
    }
    void worker(uint32_t i) {
        uint64_t seen = 0;
        for (;;) {
            std::unique_lock<std::mutex> l(m);
            cv.wait(l, [&] { return gen != seen; });
            seen = gen;
            const std::function<void()>* f = fn;
            const uint32_t b = bid, gd = gdim;
            l.unlock();
            threadIdx.x = i; blockIdx.x = b; blockDim.x = bdim; gridDim.x = gd;
            (*f)();
            std::lock_guard<std::mutex> g(m);
            if (--remaining == 0) cv_done.notify_one();
        }
    }
    void run(uint32_t b, uint32_t gd, const std::function<void()>& f) {
        std::unique_lock<std::mutex> l(m);
        fn = &f; bid = b; gdim = gd; remaining = (int)bdim; ++gen;
        cv.notify_all();
        cv_done.wait(l, [&] { return remaining == 0; });
    }
};

static std::map<uint32_t, Pool*> g_pools;
static std::mutex g_launch;

void launch_impl(uint32_t grid, uint32_t block, const std::function<void()>& fn, const char* name) {
    static const bool trace = getenv("EMU_TRACE") != nullptr;
    if (trace) { fprintf(stderr, "[emu] %s grid=%u block=%u\n", name, grid, block); fflush(stderr); }
    std::call_once(g_once, init_barriers);
    std::lock_guard<std::mutex> lk(g_launch);
    if (block == 0 || block > 1024 || (block % 64) != 0) { fprintf(stderr, "emu: bad block size %u\n", block); abort(); }
    Pool*& p = g_pools[block];
    if (!p) p = new Pool(block);     // lives until process exit
    g_block.count = (int)block;
    for (uint32_t b = 0; b < grid; ++b) p->run(b, grid, fn);
}

}  // namespace emu

hipError_t hipMalloc(void** p, size_t bytes) {
    void* q = nullptr;
    if (posix_memalign(&q, 256, bytes ? bytes : 1)) return 2;
    std::memset(q, 0xA5, bytes);     // device memory starts as garbage
    *p = q;
    return 0;
}
hipError_t hipFree(void* p) { free(p); return 0; }

struct emu_event { std::chrono::steady_clock::time_point t; };
hipError_t hipEventCreate(hipEvent_t* e) { *e = new emu_event(); return 0; }
hipError_t hipEventDestroy(hipEvent_t e) { delete e; return 0; }
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) { e->t = std::chrono::steady_clock::now(); return 0; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
    return 0;
}

// inverse complementary error function for the emulator (Newton on erfc)
float erfcinvf(float yf) {
    double y = yf;
    if (y <= 0.0) return INFINITY;
    if (y >= 2.0) return -INFINITY;
    double x = 0.0;
    for (int it = 0; it < 60; ++it) {
        double f = std::erfc(x) - y, fp = -2.0 / std::sqrt(M_PI) * std::exp(-x * x);
        double dx = f / fp;
        x -= dx;
        if (std::fabs(dx) < 1e-15 * (1.0 + std::fabs(x))) break;
    }
    return (float)x;
}


// Host side of the MI355X expander-decomposition engine: device state, the
// per-iteration launch sequence and the C ABI of include/parex_amd/pxd.h.
//
// One engine holds one graph or a block-diagonal batch of graphs.  Clusters
// are contiguous ranges of the active-position list; all label arithmetic of
// the reference (dense relabelling, deactivation labels, smallest label) is
// carried per graph, so a graph's result does not depend on its batch-mates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "parex_amd/pxd.h"
#include "pxd_kernels.h"
#include "pxd_ingest.h"
#include "pxd_small.h"

using namespace pxd;

static thread_local std::string g_last_error;

static int set_err(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    try {
        g_last_error = buf;
    } catch (...) {                   // no room for the message: keep the code
    }
    return code;
}

// Every C entry point runs its body through this: host allocations (vectors
// sized by the graph, the serving worker thread) may throw, and no exception
// may cross the C ABI.  Allocation failures become PXD_ERR_OOM.
template <class F>
static int guarded(F&& f) noexcept {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return set_err(PXD_ERR_OOM, "host allocation failed");
    } catch (const std::exception& ex) {
        return set_err(PXD_ERR_STATE, "%s", ex.what());
    } catch (...) {
        return set_err(PXD_ERR_STATE, "unexpected exception");
    }
}

#define HIPC(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess)                                                                     \
            return set_err(PXD_ERR_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_),       \
                           __FILE__, __LINE__);                                                   \
    } while (0)

#define CHK(x)                        \
    do {                              \
        int r_ = (x);                 \
        if (r_ != 0) return r_;       \
    } while (0)

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }
// Grid for grid-stride kernels: enough workgroups to fill 256 CUs many times
// over, never more (each extra workgroup costs dispatch time and, for kernels
// that end with a flag update, one more atomic on a shared word).
constexpr uint64_t kMaxGrid = 8192;
static inline uint32_t grid_for(uint64_t n, uint32_t per_block) {
    uint64_t g = (n + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > kMaxGrid) g = kMaxGrid;
    return (uint32_t)g;
}

// grid for kernels that split their items per XCD (xcd_span): a multiple of 8
static inline uint32_t grid_xcd(uint64_t n, uint32_t per_block) {
    const uint32_t g = grid_for(n, per_block);
    return (g + NXCD - 1) / NXCD * NXCD;
}

struct ClusterTable {
    uint32_t *cs = nullptr, *cgraph = nullptr, *cmin = nullptr, *cflag = nullptr, *cvol = nullptr;
    int32_t* clabel = nullptr;
    ClusterCols cols() const { return ClusterCols{cs, cgraph, cmin, cflag, cvol, clabel}; }
};

// cuda/devPartition.h:798-823 for one graph's labels: id = rank of the label
// among the distinct labels (sort, unique, rank); counting over the label
// range when it is small, a sort otherwise.  Returns the number of clusters.
static int32_t dense_rank(const int32_t* lab, uint32_t n, int32_t* ids) {
    if (n == 0) return 0;
    int32_t lo = lab[0], hi = lab[0];
    for (uint32_t i = 1; i < n; ++i) { lo = std::min(lo, lab[i]); hi = std::max(hi, lab[i]); }
    const uint64_t range = (uint64_t)((int64_t)hi - (int64_t)lo) + 1u;
    if (range <= 4ull * n + 64u) {
        std::vector<int32_t> rank(range, -1);
        for (uint32_t i = 0; i < n; ++i) rank[lab[i] - lo] = 0;
        int32_t k = 0;
        for (uint64_t r = 0; r < range; ++r)
            if (rank[r] == 0) rank[r] = ++k;
        for (uint32_t i = 0; i < n; ++i) ids[i] = rank[lab[i] - lo] - 1;
        return k;
    }
    std::vector<int32_t> u(lab, lab + n);
    std::sort(u.begin(), u.end());
    u.erase(std::unique(u.begin(), u.end()), u.end());
    for (uint32_t i = 0; i < n; ++i) ids[i] = (int32_t)(std::lower_bound(u.begin(), u.end(), lab[i]) - u.begin());
    return (int32_t)u.size();
}

struct pxd_engine {
    pxd_params P{};
    int dev = 0;
    hipStream_t st = nullptr;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;                 // size of each entry of allocs
    // buffers of the previous input, kept for the next setup on this engine:
    // (bytes, pointer), taken back best-fit; the rest is freed once setup ends
    std::vector<std::pair<size_t, void*>> pool;

    // sizes
    uint32_t G = 0, N = 0, A = 0;
    uint64_t M_total = 0;
    std::vector<uint32_t> h_gnode;    // G+1 node offsets
    std::vector<uint32_t> h_garc;     // G+1 arc offsets
    std::vector<uint32_t> h_gm;       // undirected edges per graph
    bool inited = false;

    // graph
    uint32_t *d_row = nullptr, *d_deg = nullptr, *d_nbr = nullptr, *d_nbr0 = nullptr, *d_gnode = nullptr,
             *d_garc = nullptr;
    // per node
    uint32_t* d_dint = nullptr;
    int32_t* d_label = nullptr;
    double* d_x[2] = {nullptr, nullptr};
    uint64_t* d_y[2] = {nullptr, nullptr};   // packed walk terms of x/dInt (fix_pack)
    uint32_t* d_xkey = nullptr;
    uint2* d_ediff = nullptr;   // per node: cut contribution and dInt of the current sweep
    uint32_t *d_ped = nullptr, *d_pvol = nullptr;   // cut contribution and dInt in position order (sweep)
    uint32_t *d_root = nullptr, *d_part = nullptr, *d_stamp = nullptr;
    uint8_t* d_alive = nullptr;
    // positions
    uint32_t *d_pnix[2] = {nullptr, nullptr}, *d_pkey[2] = {nullptr, nullptr};
    uint32_t *d_pclu[2] = {nullptr, nullptr};
    uint32_t* d_alist[2] = {nullptr, nullptr};
    pxd_node_data* d_nd = nullptr;
    // clusters
    ClusterTable ct[2];
    int ctc = 0;
    uint32_t *d_coff = nullptr, *d_ckeep = nullptr, *d_kscan = nullptr, *d_gfirst = nullptr;
    uint32_t *d_cmaxo = nullptr, *d_cmino = nullptr;
    int32_t* d_cK = nullptr;
    float* d_csp = nullptr;
    double* d_cavg = nullptr;
    unsigned long long *d_cbest = nullptr, *d_csum = nullptr;
    // graphs
    int32_t *d_gsmall = nullptr, *d_gsmall_next = nullptr, *d_giter = nullptr;
    uint32_t *d_gsurv = nullptr, *d_gseen = nullptr;
    // scratch
    uint32_t *d_tA = nullptr, *d_tB = nullptr, *d_tC = nullptr;
    uint32_t* d_cclist = nullptr;
    uint32_t *d_lwave = nullptr, *d_lblock = nullptr, *d_llarge = nullptr;
    uint32_t *d_ltiles = nullptr, *d_ltoff = nullptr;   // tiles per large segment, their scan
    LargeSeg* d_ls = nullptr;
    uint32_t *d_tseg = nullptr, *d_tidx = nullptr, *d_hist = nullptr, *d_hscan = nullptr;
    uint32_t tiles_cap = 0, ls_cap = 0;
    // single-pass scans: status words, ticket counter, the epoch of the last launch
    unsigned long long* d_lb_scan = nullptr;
    uint32_t* d_lb_ticket = nullptr;
    uint32_t* d_sout = nullptr;        // final counts of the one-workgroup run
    uint32_t lb_scan_cap = 0, lb_epoch = 0;
    uint32_t *d_tail_ed = nullptr, *d_tail_vol = nullptr, *d_hasst = nullptr, *d_carry_ed = nullptr,
             *d_carry_vol = nullptr;
    uint32_t* d_oneseg = nullptr;
    uint32_t* d_gclu = nullptr;
    uint32_t* d_giant = nullptr;
    uint32_t* d_nclu = nullptr;       // cluster of every active node (apply pass)
    uint32_t* d_ngraph = nullptr;     // graph (slot) of every node
    uint32_t* d_gstart = nullptr;     // iteration a graph joined at (0 for a batch)
    // heavy rows: chunk counts, their scan, chunk -> row, chunk partial sums
    uint32_t *d_hcnt = nullptr, *d_hoff = nullptr, *d_hrow = nullptr;
    Fix128* d_hpart = nullptr;
    uint32_t maxdeg = 0, heavy_thr = 0xFFFFFFFFu, heavy_cap = 0;
    bool heavy_valid = false;
    uint32_t* d_cclist2 = nullptr;
    uint32_t* d_ccsamp = nullptr;                    // CC sampling: arcs 1.. of every member
    unsigned long long* d_u64 = nullptr;
    Counters* d_cnt = nullptr;
    Counters hc{};

    // host mirrors
    int cur = 0, acur = 0, pcur = 0;
    uint32_t na = 0, nc = 0, ncut = 0, ncc = 0, nalist = 0;
    uint32_t ncc_members = 0;         // positions of the clusters the sweep flagged for CC
    bool giter_read = false;          // the cut's end-of-iteration read brought h_giter (serving)
    uint32_t iter1 = 0;
    bool sweep_done = false;
    int gs = 16;

    // timing
    double last_run_ms = 0.0;
    bool walk_timing = false;
    uint64_t walk_launches = 0;
    double walk_ms = 0.0, walk_bytes = 0.0;
    std::vector<double> walk_log_ms, walk_log_bytes;   // per timed launch
    hipEvent_t ev0 = nullptr, ev1 = nullptr;

    // kernel choice for the two arc passes of an iteration (0 = walk, 1 = cut
    // contribution): variants 0..5 = one row per 2, 4, ..., 64 lanes
    // (k_walk / k_contrib), 6..8 = the LDS-window kernels with 1, 2, 4 lanes
    // per row (4096-entry chunks), 9 = 2 lanes per row, 1024-entry chunks and
    // a quarter of the window, 10 = variant 7 with clamped slots (walk only;
    // the cut contribution runs variant 7 for it), 11 = 2 lanes per row over
    // node-range chunks of RANGE_R indices, 12 = the same over chunks of
    // RANGE_R / 2 indices with half the window (two workgroups a CU, so one
    // stages its window while the other computes), 13 = variant 7 with half
    // the window over 2048-entry chunks (two workgroups a CU), 14 / 15 = variants
    // 11 / 12 with one lane per row (short rows: one round of 64 rows a wave).
    // walk_force >= 0 pins one
    // variant for both passes; -1 = online
    // tuning: the first time the node list has a length in [2^b, 2^(b+1)),
    // every candidate is launched three times on the current state (both passes are
    // pure functions of it, all variants give the same bits) and timed, and
    // the fastest is kept for that pass and length class; -2 = the degree rule
    // without tuning.  The tables survive pxd_reset, so repeated
    // decompositions of the same input reuse them.
    static constexpr int WV_LANES2 = 0, WV_WIN1 = 6, WV_WIN2 = 7, WV_WIN4 = 8, WV_WIN2S = 9, WV_WIN2C = 10,
                         WV_WIN2R = 11, WV_WIN2RH = 12, WV_WIN2H = 13, WV_WIN1R = 14, WV_WIN1RH = 15,
                         WV_COUNT = 16;
    static constexpr uint32_t RANGE_R = 8192;        // node-range chunks of variant 11 (12: half)
    uint32_t* d_rmap = nullptr;                      // k_range_offsets of the current list, RANGE_R
    uint32_t* d_rmap_h = nullptr;                    // the same for RANGE_R / 2
    bool rmap_valid = false;                         // both maps belong to the current list
    bool rmap_h_built = false;
    static constexpr uint32_t TUNE_MIN = 1u << 16;   // shorter lists: launch cost dominates, no tuning
    int walk_force = -1;
    int small_kernel = 1;              // one-workgroup runs: 0 never, 1 up to SMALL_AUTO_ARCS, 2 always
    // tuned[pass][b][q]: the list lengths in [2^b, 2^(b+1)) in four quarters by
    // the two bits after the leading one (a long class such as 2^24 spans a
    // dozen iterations whose best kernel changes)
    static constexpr int TUNE_Q = 4;
    int tuned[2][33][TUNE_Q];
    int& tuned_slot(int pass, uint32_t n) {
        const int b = 31 - __builtin_clz(n);
        const int q = b >= 2 ? (int)((n >> (b - 2)) & 3u) : 0;
        return tuned[pass][b][q];
    }
    uint64_t tuned_for[3] = {~0ull, ~0ull, ~0ull};   // (graphs, nodes, arcs) the tables were tuned on
    uint64_t walk_tunings = 0;
    double avg_deg = 0.0;
    hipEvent_t tev0 = nullptr, tev1 = nullptr;

    pxd_engine() { std::fill(&tuned[0][0][0], &tuned[0][0][0] + 2 * 33 * TUNE_Q, -1); }
    ~pxd_engine() {
        release();
        trim_pool();
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (tev0) (void)hipEventDestroy(tev0);
        if (tev1) (void)hipEventDestroy(tev1);
        if (st) (void)hipStreamDestroy(st);
        if (h_pin) (void)hipHostFree(h_pin);
    }

    // The input's buffers go to the pool.  Nothing queued on the engine's
    // stream may still use one when it is handed out again, so the stream is
    // drained first (hipFree used to do that).  Every buffer is written by
    // setup and init_state before it is read, as on a fresh allocation.
    void release() {
        stop_ranker();
        if (!allocs.empty() && st) (void)hipStreamSynchronize(st);
        for (size_t i = 0; i < allocs.size(); ++i) pool.emplace_back(alloc_bytes[i], allocs[i]);
        allocs.clear();
        alloc_bytes.clear();
        d_hcnt = d_hoff = d_hrow = nullptr;   // heavy-row buffers are allocated lazily
        d_hpart = nullptr;
        heavy_cap = 0;
        ing_n = 0;
        ing_m = 0;
        d_ifx = d_itx = nullptr;
        ing = IngestWork{};
        inited = false;
    }

    void trim_pool() {
        for (auto& b : pool) (void)hipFree(b.second);
        pool.clear();
    }

    template <typename T>
    int alloc(T** p, size_t count) {
        void* q = nullptr;
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        size_t have = bytes, best = pool.size();
        for (size_t i = 0; i < pool.size(); ++i)
            if (pool[i].first >= bytes && (best == pool.size() || pool[i].first < pool[best].first)) best = i;
        if (best < pool.size()) {
            have = pool[best].first;
            q = pool[best].second;
            pool.erase(pool.begin() + best);
        } else {
            hipError_t e = hipMalloc(&q, bytes);
            if (e != hipSuccess && !pool.empty()) {   // give the kept buffers back and retry
                (void)hipGetLastError();
                trim_pool();
                e = hipMalloc(&q, bytes);
            }
            if (e != hipSuccess)
                return set_err(PXD_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
        }
        allocs.push_back(q);
        alloc_bytes.push_back(have);
        *p = (T*)q;
        return 0;
    }

    int sync() {
        HIPC(hipStreamSynchronize(st));
        return 0;
    }

    // pinned host words for the reads of device counts (a direct copy, no staging)
    uint32_t* h_pin = nullptr;
    static constexpr size_t PIN_BYTES = 256;
    int pinned() {
        if (!h_pin) HIPC(hipHostMalloc((void**)&h_pin, PIN_BYTES, hipHostMallocDefault));
        return 0;
    }
    int read_counters() {
        static_assert(sizeof(Counters) + 16 <= PIN_BYTES, "pinned read buffer");
        CHK(pinned());
        HIPC(hipMemcpyAsync(h_pin, d_cnt, sizeof(Counters), hipMemcpyDeviceToHost, st));
        CHK(sync());
        std::memcpy(&hc, h_pin, sizeof(Counters));
        if (hc.err & ERR_FIX_RANGE)
            return set_err(PXD_ERR_NUMERIC, "walk value outside the exact fixed-point range (|x/dInt| >= 2^31)");
        return 0;
    }

    int read_u32(const uint32_t* d, uint32_t* h) {
        CHK(pinned());
        uint32_t* slot = h_pin + (PIN_BYTES - 16) / 4;
        HIPC(hipMemcpyAsync(slot, d, 4, hipMemcpyDeviceToHost, st));
        CHK(sync());
        *h = *slot;
        return 0;
    }
    // device arrays for N nodes, A arcs and G graphs (capacities when serving)
    int allocate() {
        const size_t n1 = (size_t)N + 1;
        CHK(alloc(&d_row, n1));
        CHK(alloc(&d_deg, N));
        CHK(alloc(&d_nbr, A));
        if (P.keep_original) CHK(alloc(&d_nbr0, A));
        CHK(alloc(&d_gnode, G + 1));
        CHK(alloc(&d_garc, G + 1));
        CHK(alloc(&d_dint, N));
        CHK(alloc(&d_label, N));
        for (int b = 0; b < 2; ++b) {
            CHK(alloc(&d_x[b], N));
            CHK(alloc(&d_y[b], N));
            CHK(alloc(&d_pnix[b], N));
            CHK(alloc(&d_pkey[b], N));
            CHK(alloc(&d_pclu[b], N));
            CHK(alloc(&d_alist[b], N));
            CHK(alloc(&ct[b].cs, n1 + 1));
            CHK(alloc(&ct[b].cgraph, n1));
            CHK(alloc(&ct[b].cmin, n1));
            CHK(alloc(&ct[b].cflag, n1));
            CHK(alloc(&ct[b].cvol, n1));
            CHK(alloc(&ct[b].clabel, n1));
        }
        CHK(alloc(&d_xkey, N));
        CHK(alloc(&d_ediff, N));
        CHK(alloc(&d_ped, N));
        // chunk offsets (nch + 1) and the XCD split (NXCD + 1)
        CHK(alloc(&d_rmap, (size_t)N / RANGE_R + 3 + NXCD));
        CHK(alloc(&d_rmap_h, (size_t)N / (RANGE_R / 2) + 3 + NXCD));
        CHK(alloc(&d_pvol, N));
        CHK(alloc(&d_root, N));
        CHK(alloc(&d_part, N));
        CHK(alloc(&d_nclu, N));
        CHK(alloc(&d_ngraph, N));
        CHK(alloc(&d_stamp, N));
        CHK(alloc(&d_alive, N));
        if (P.record_node_data) CHK(alloc(&d_nd, N));
        CHK(alloc(&d_coff, n1));
        CHK(alloc(&d_ckeep, n1));
        CHK(alloc(&d_kscan, n1 + 1));
        CHK(alloc(&d_cmaxo, n1));
        CHK(alloc(&d_cmino, n1));
        CHK(alloc(&d_cK, n1));
        CHK(alloc(&d_csp, n1));
        CHK(alloc(&d_cavg, n1));
        CHK(alloc(&d_cbest, n1));
        CHK(alloc(&d_csum, n1));
        CHK(alloc(&d_gfirst, G + 1));
        CHK(alloc(&d_gsmall, G));
        CHK(alloc(&d_gsmall_next, G));
        CHK(alloc(&d_giter, G));
        CHK(alloc(&d_gsurv, G));
        CHK(alloc(&d_gseen, G));
        CHK(alloc(&d_gstart, G));
        CHK(alloc(&d_tA, n1 + 1));
        CHK(alloc(&d_tB, n1 + 1));
        CHK(alloc(&d_tC, n1 + 1));
        CHK(alloc(&d_cclist, N));
        CHK(alloc(&d_lwave, n1));
        CHK(alloc(&d_lblock, n1));
        CHK(alloc(&d_llarge, n1));
        ls_cap = N / TILE + 2;
        tiles_cap = 2 * (N / TILE) + 4;
        CHK(alloc(&d_ltiles, ls_cap));
        CHK(alloc(&d_ltoff, (size_t)ls_cap + 1));
        CHK(alloc(&d_ls, ls_cap));
        CHK(alloc(&d_tseg, tiles_cap));
        CHK(alloc(&d_tidx, tiles_cap));
        CHK(alloc(&d_hist, (size_t)tiles_cap * 256 + 1));
        CHK(alloc(&d_hscan, (size_t)tiles_cap * 256 + 2));
        // the largest scans: N + 1 elements, or the tile histograms
        lb_scan_cap = cdiv(std::max<uint64_t>((uint64_t)n1 + 1, (uint64_t)tiles_cap * 256 + 1), LB_SCAN_TILE) + 1;
        // the selections' short tiles (lists shorter than SE_LONG)
        lb_scan_cap = std::max<uint32_t>(lb_scan_cap, cdiv(n1, BS * SE_ROWS_SHORT) + 1);
        CHK(alloc(&d_lb_scan, lb_scan_cap));
        CHK(alloc(&d_lb_ticket, 1));
        const uint32_t nbk = cdiv(n1, BS * SCAN_ITEMS) + 1;
        CHK(alloc(&d_tail_ed, nbk));
        CHK(alloc(&d_tail_vol, nbk));
        CHK(alloc(&d_hasst, nbk));
        CHK(alloc(&d_carry_ed, nbk));
        CHK(alloc(&d_carry_vol, nbk));
        CHK(lb_clear());
        CHK(alloc(&d_oneseg, 2));
        CHK(alloc(&d_gclu, G));
        CHK(alloc(&d_giant, 2 * n1 + 2));
        CHK(alloc(&d_cclist2, N));
        CHK(alloc(&d_ccsamp, (size_t)(CC_SAMPLE_ROUNDS - 1) * N));
        CHK(alloc(&d_u64, 1));
        CHK(alloc(&d_sout, 8));
        CHK(alloc(&d_cnt, 1));
        return 0;
    }

    // ---------------------------------------------------------------- setup
    int setup(uint32_t g, const uint32_t* ncnt, const uint32_t* ecnt, const uint32_t* const* rptr,
              const uint32_t* const* eptr) {
        release();
        serving = false;
        HIPC(hipSetDevice(dev));
        G = g;
        h_gnode.assign(G + 1, 0);
        h_garc.assign(G + 1, 0);
        h_gm.assign(G, 0);
        uint64_t nn = 0, aa = 0;
        for (uint32_t i = 0; i < G; ++i) {
            h_gnode[i] = (uint32_t)nn;
            h_garc[i] = (uint32_t)aa;
            nn += ncnt[i];
            aa += 2ull * ecnt[i];
            h_gm[i] = ecnt[i];
            if (rptr[i][ncnt[i]] != 2ull * ecnt[i] || rptr[i][0] != 0)
                return set_err(PXD_ERR_ARG, "graph %u: ranges[n] must equal 2m and ranges[0] == 0", i);
        }
        if (nn >= 0x7FFFFFFFull || aa >= 0xFFFFFFFFull)
            return set_err(PXD_ERR_ARG, "batch too large for 32-bit indices (%llu nodes, %llu arcs)",
                           (unsigned long long)nn, (unsigned long long)aa);
        h_gnode[G] = (uint32_t)nn;
        h_garc[G] = (uint32_t)aa;
        // the online tuner's tables belong to the input's sizes: a new input of
        // the same sizes (the same workload uploaded again) keeps them
        if (tuned_for[0] != G || tuned_for[1] != nn || tuned_for[2] != aa) {
            std::fill(&tuned[0][0][0], &tuned[0][0][0] + 2 * 33 * TUNE_Q, -1);
            tuned_for[0] = G;
            tuned_for[1] = nn;
            tuned_for[2] = aa;
        }
        N = (uint32_t)nn;
        A = (uint32_t)aa;
        M_total = aa / 2;

        CHK(allocate());
        // Every graph's ranges (first n entries) and arcs go up as they are:
        // host work is O(graphs).  One pass over the rows rebases them to the
        // batch's arc offsets, checks that they do not decrease and finds the
        // largest degree; one pass over the arcs checks each below its graph's
        // n and rebases it to global node ids.  One sync reads both results.
        HIPC(hipMemsetAsync(d_cnt, 0, sizeof(Counters), st));
        for (uint32_t i = 0; i < G; ++i) {
            if (ncnt[i]) HIPC(hipMemcpyAsync(d_row + h_gnode[i], rptr[i], (size_t)ncnt[i] * 4, hipMemcpyHostToDevice, st));
            const uint64_t arcs = 2ull * ecnt[i];
            if (arcs) HIPC(hipMemcpyAsync(d_nbr + h_garc[i], eptr[i], arcs * 4, hipMemcpyHostToDevice, st));
        }
        HIPC(hipMemcpyAsync(d_gnode, h_gnode.data(), (G + 1) * 4, hipMemcpyHostToDevice, st));
        HIPC(hipMemcpyAsync(d_garc, h_garc.data(), (G + 1) * 4, hipMemcpyHostToDevice, st));
        k_fill_u32<<<1, 64, 0, st>>>(1, d_row + N, A);
        if (N) {
            k_rebase_rows<<<grid_for(N, BS), BS, 0, st>>>(N, G, d_gnode, d_garc, d_row);
            k_check_rows<<<grid_for(N, BS), BS, 0, st>>>(N, d_row, d_cnt);
        }
        if (A) k_rebase_arcs<<<grid_for(A, REBASE_CHUNK), BS, 0, st>>>(A, G, d_gnode, d_garc, d_nbr, &d_cnt->admit_bad);
        HIPC(hipGetLastError());
        Counters hc0{};
        HIPC(hipMemcpyAsync(&hc0, d_cnt, sizeof(Counters), hipMemcpyDeviceToHost, st));
        CHK(sync());                  // also: the host vectors go out of scope after the sync
        maxdeg = hc0.maxdeg;
        if (hc0.admit_bad) {          // name the first bad range or arc
            for (uint32_t i = 0; i < G; ++i)
                for (uint32_t u = 0; u < ncnt[i]; ++u)
                    if (rptr[i][u + 1] < rptr[i][u]) return set_err(PXD_ERR_ARG, "graph %u: ranges decrease at %u", i, u);
            for (uint32_t i = 0; i < G; ++i)
                for (uint64_t j = 0; j < 2ull * ecnt[i]; ++j)
                    if (eptr[i][j] >= ncnt[i])
                        return set_err(PXD_ERR_ARG, "graph %u: arc %llu points to %u >= n", i,
                                       (unsigned long long)j, eptr[i][j]);
            return set_err(PXD_ERR_STATE, "input check failed");
        }
        HIPC(hipMemsetAsync(d_cnt, 0, sizeof(Counters), st));
        if (P.keep_original && A)
            HIPC(hipMemcpyAsync(d_nbr0, d_nbr, (size_t)A * 4, hipMemcpyDeviceToDevice, st));
        CHK(sync());
        return init_state();
    }

    // one graph from an undirected edge list, CSR built on the device
    int setup_edges(uint32_t n, uint64_t m, const uint32_t* fx, const uint32_t* tx) {
        release();
        serving = false;
        HIPC(hipSetDevice(dev));
        std::fill(&tuned[0][0][0], &tuned[0][0][0] + 2 * 33 * TUNE_Q, -1);
        tuned_for[0] = tuned_for[1] = tuned_for[2] = ~0ull;
        if (n >= 0x7FFFFFFFu || 2 * m >= 0x7FFFFFFFull)
            return set_err(PXD_ERR_ARG, "device CSR build takes below 2^31 nodes and 2^31 arcs (%u nodes, %llu edges)",
                           n, (unsigned long long)m);
        if (m && (!fx || !tx)) return set_err(PXD_ERR_ARG, "null edge arrays");
        G = 1;
        h_gnode.assign({0u, n});
        h_garc.assign({0u, (uint32_t)(2 * m)});
        h_gm.assign(1, (uint32_t)m);
        N = n;
        A = (uint32_t)(2 * m);
        M_total = m;
        CHK(allocate());
        // the edge list and the build's work area: engine buffers, back to the
        // pool with the rest of this input
        uint32_t *fxd = nullptr, *txd = nullptr, *work = nullptr;
        CHK(alloc(&fxd, m));
        CHK(alloc(&txd, m));
        CHK(alloc(&work, ingest_work_words(n, 2 * m)));
        if (m) {
            HIPC(hipMemcpyAsync(fxd, fx, m * 4, hipMemcpyHostToDevice, st));
            HIPC(hipMemcpyAsync(txd, tx, m * 4, hipMemcpyHostToDevice, st));
        }
        HIPC(hipMemsetAsync(d_cnt, 0, sizeof(Counters), st));
        HIPC(ingest_csr_async(n, m, fxd, txd, d_row, d_nbr, ingest_carve(work, n, 2 * m), &d_cnt->admit_bad, st));
        uint32_t bad = 0;
        CHK(read_u32(&d_cnt->admit_bad, &bad));
        if (bad) {
            for (uint64_t i = 0; i < m; ++i)
                if (fx[i] >= n || tx[i] >= n)
                    return set_err(PXD_ERR_ARG, "edge %llu (%u, %u) has an end >= n = %u", (unsigned long long)i,
                                   fx[i], tx[i], n);
            return set_err(PXD_ERR_STATE, "edge check failed");
        }
        std::vector<uint32_t> row((size_t)N + 1);
        HIPC(hipMemcpyAsync(row.data(), d_row, ((size_t)N + 1) * 4, hipMemcpyDeviceToHost, st));
        HIPC(hipMemsetAsync(d_cnt, 0, sizeof(Counters), st));
        HIPC(hipMemcpyAsync(d_gnode, h_gnode.data(), (G + 1) * 4, hipMemcpyHostToDevice, st));
        if (P.keep_original && A)
            HIPC(hipMemcpyAsync(d_nbr0, d_nbr, (size_t)A * 4, hipMemcpyDeviceToDevice, st));
        CHK(sync());
        maxdeg = 0;
        for (uint32_t u = 0; u < N; ++u) maxdeg = std::max(maxdeg, row[u + 1] - row[u]);
        return init_state();
    }

    // back to the state right after initialize, from the device-resident input CSR
    int reset() {
        // only a completed initialize leaves a valid device copy of the input:
        // a failed one may have freed or never written it
        if (!inited) return set_err(PXD_ERR_ARG, "engine not initialized");
        if (serving) return set_err(PXD_ERR_ARG, "a serving engine has no reset");
        if (!P.keep_original) return set_err(PXD_ERR_ARG, "reset needs keep_original = 1");
        HIPC(hipSetDevice(dev));
        if (A) HIPC(hipMemcpyAsync(d_nbr, d_nbr0, (size_t)A * 4, hipMemcpyDeviceToDevice, st));
        HIPC(hipMemsetAsync(d_cnt, 0, sizeof(Counters), st));
        return init_state();
    }

    // initial clusters: one per non-empty graph (label 0, connectivity unknown)
    int init_state() {
        std::vector<uint32_t> cs, cg, cmin, cvol, gclu(G, 0);
        for (uint32_t g = 0; g < G; ++g) {
            gclu[g] = (uint32_t)cs.size();
            if (h_gnode[g + 1] == h_gnode[g]) continue;
            cs.push_back(h_gnode[g]);
            cg.push_back(g);
            cmin.push_back(h_gnode[g]);
            cvol.push_back(h_garc[g + 1] - h_garc[g]);
        }
        nc = (uint32_t)cs.size();
        cs.push_back(N);
        std::vector<uint32_t> zeros(nc + 1, 0);
        ctc = 0;
        ClusterTable& T = ct[0];
        HIPC(hipMemcpyAsync(T.cs, cs.data(), cs.size() * 4, hipMemcpyHostToDevice, st));
        HIPC(hipMemcpyAsync(T.cgraph, cg.data(), nc * 4, hipMemcpyHostToDevice, st));
        HIPC(hipMemcpyAsync(T.cmin, cmin.data(), nc * 4, hipMemcpyHostToDevice, st));
        HIPC(hipMemcpyAsync(T.cvol, cvol.data(), nc * 4, hipMemcpyHostToDevice, st));
        HIPC(hipMemcpyAsync(T.cflag, zeros.data(), nc * 4, hipMemcpyHostToDevice, st));
        HIPC(hipMemcpyAsync(T.clabel, zeros.data(), nc * 4, hipMemcpyHostToDevice, st));
        HIPC(hipMemcpyAsync(d_gclu, gclu.data(), G * 4, hipMemcpyHostToDevice, st));
        HIPC(hipMemsetAsync(d_gsmall, 0, G * 4, st));
        HIPC(hipMemsetAsync(d_gsmall_next, 0, G * 4, st));
        HIPC(hipMemsetAsync(d_giter, 0, G * 4, st));
        HIPC(hipMemsetAsync(d_gsurv, 0, G * 4, st));
        HIPC(hipMemsetAsync(d_gseen, 0, G * 4, st));
        HIPC(hipMemsetAsync(d_gstart, 0, G * 4, st));
        cur = acur = pcur = 0;
        if (N) {
            k_init_nodes<<<grid_for(N, BS), BS, 0, st>>>(N, G, d_gnode, d_row, P.seed, d_deg, d_dint, d_x[0], d_y[0],
                                                          d_xkey, d_label, d_alive, d_stamp, d_pnix[0], d_alist[0],
                                                          d_pclu[0], d_gclu, d_ngraph, &d_cnt->err);
            HIPC(hipGetLastError());
        }
        CHK(sync());
        na = N;
        nalist = N;
        ncut = 0;
        ncc = 0;
        iter1 = 0;
        sweep_done = false;
        const double avg = N ? (double)A / (double)N : 0.0;
        avg_deg = avg;
        // lanes per node: about four arcs per lane (the arc loops keep four in flight)
        gs = avg <= 24.0 ? 4 : avg <= 48.0 ? 8 : avg <= 96.0 ? 16 : avg <= 192.0 ? 32 : 64;
        // rows longer than 32 arcs per lane leave the lane-group kernels
        heavy_thr = maxdeg > 32u * gs ? 32u * gs : 0xFFFFFFFFu;
        heavy_valid = false;
        rmap_valid = false;
        if (heavy_thr != 0xFFFFFFFFu && !d_hcnt) {
            heavy_cap = A / HEAVY_CHUNK + A / heavy_thr + 1u;
            CHK(alloc(&d_hcnt, N));
            CHK(alloc(&d_hoff, (size_t)N + 1));
            CHK(alloc(&d_hrow, heavy_cap));
            CHK(alloc(&d_hpart, heavy_cap));
        }
        inited = true;
        return 0;
    }

    // ---------------------------------------------------------------- scan
    // Status words and tickets start cleared (buffers may come back from the
    // pool holding anything); epochs then keep launches apart.
    int lb_clear() {
        HIPC(hipMemsetAsync(d_lb_scan, 0, (size_t)lb_scan_cap * 8, st));
        HIPC(hipMemsetAsync(d_lb_ticket, 0, sizeof(uint32_t), st));
        lb_epoch = 0;
        return 0;
    }
    int next_epoch(uint32_t* e) {
        if (lb_epoch == LB_EPOCH_MAX) CHK(lb_clear());
        *e = ++lb_epoch;
        return 0;
    }
    // out[0..n) exclusive prefix sums, out[n] = total.  With pn the length is
    // *pn, read on the device, and n is a bound on it that sizes the grid;
    // total (when given) receives the sum too, at an address the host knows.
    int scan(const uint32_t* in, uint32_t* out, uint32_t n, const uint32_t* pn = nullptr, uint32_t* total = nullptr) {
        if (n == 0) {
            k_fill_u32<<<1, 64, 0, st>>>(1, out, 0u);
            if (total) k_fill_u32<<<1, 64, 0, st>>>(1, total, 0u);
        } else if (n <= (uint32_t)(BS * SCAN_ITEMS)) {
            k_scan_down<<<1, BS, 0, st>>>(n, in, out, nullptr, pn, total);
        } else if (n <= SMALL_SCAN_N) {
            k_scan_small<<<1, 1024, 0, st>>>(n, in, out, pn, total);
        } else {
            const uint32_t nb = cdiv(n, LB_SCAN_TILE);
            if (nb > lb_scan_cap) return set_err(PXD_ERR_STATE, "scan of %u elements exceeds the status capacity", n);
            uint32_t ep;
            CHK(next_epoch(&ep));
            k_scan_lb<<<nb, BS, 0, st>>>(n, in, out, d_lb_scan, d_lb_ticket, ep, pn, total);
        }
        HIPC(hipGetLastError());
        return 0;
    }

    // one launch of k_scan_emit over n elements (n = *pn on the device when pn
    // is given; n bounds it and sizes the grid); the flagged count to *total
    template <class F>
    int select(uint32_t n, const F& f, const uint32_t* pn = nullptr, uint32_t* total = nullptr) {
        const bool lng = n >= SE_LONG;
        const uint32_t nb = std::max(cdiv(n, BS * (lng ? SE_ROWS : SE_ROWS_SHORT)), 1u);
        if (nb > lb_scan_cap) return set_err(PXD_ERR_STATE, "selection over %u elements exceeds the status capacity", n);
        uint32_t ep;
        CHK(next_epoch(&ep));
        if (lng) k_scan_emit<SE_ROWS, F><<<nb, BS, 0, st>>>(n, f, d_lb_scan, d_lb_ticket, ep, pn, total);
        else k_scan_emit<SE_ROWS_SHORT, F><<<nb, BS, 0, st>>>(n, f, d_lb_scan, d_lb_ticket, ep, pn, total);
        HIPC(hipGetLastError());
        return 0;
    }

    // ---------------------------------------------------------------- segmented sort
    // Stable sort of (pkey[0], pnix[0]) inside every segment [segs[c], segs[c+1])
    // of the n items.  The size classes and the large segments' tile plan stay
    // on the device (no host read): grids are sized from what n and nseg bound,
    // and every kernel reads its list length from the counters.
    int seg_sort(const uint32_t* segs, uint32_t nseg, uint32_t n) {
        if (nseg == 0 || n < 2) return 0;
        if (n <= SORT_SMALL_USE) {    // one workgroup sorts every segment
            k_sort_small<<<1, SS_NT, 0, st>>>(nseg, segs, d_pkey[0], d_pnix[0]);
            HIPC(hipGetLastError());
            return 0;
        }
        HIPC(hipMemsetAsync(&d_cnt->nwave, 0, 3 * sizeof(uint32_t), st));
        k_classify<<<grid_for(nseg, BS), BS, 0, st>>>(nseg, segs, d_lwave, d_lblock, d_llarge, d_cnt);
        HIPC(hipGetLastError());
        // class sizes: 2..64, 65..TILE and more than TILE items per segment
        const uint32_t nwave = std::min(nseg, n / 2u), nblock = std::min(nseg, n / (uint32_t)(WAVE + 1)),
                       nlarge = std::min(nseg, n / (uint32_t)(TILE + 1));
        if (nwave) {
            k_sort_wave<<<std::min(cdiv(nwave, BS / WAVE), 8192u), BS, 0, st>>>(&d_cnt->nwave, d_lwave, segs,
                                                                                d_pkey[0], d_pnix[0]);
            HIPC(hipGetLastError());
        }
        if (nblock) {
            k_sort_block<<<std::min(nblock, 2048u), BS, 0, st>>>(&d_cnt->nblock, d_lblock, segs, d_pkey[0], d_pnix[0]);
            HIPC(hipGetLastError());
        }
        if (nlarge) {
            // a large segment of len items has ceil(len / TILE) <= len / TILE + 1 tiles
            const uint32_t tiles = n / (uint32_t)TILE + nlarge;
            if (nlarge > ls_cap || tiles > tiles_cap) return set_err(PXD_ERR_STATE, "tile capacity exceeded");
            k_large_plan<<<grid_for(nlarge, BS), BS, 0, st>>>(&d_cnt->nlarge, d_llarge, segs, d_ls, d_ltiles);
            HIPC(hipGetLastError());
            CHK(scan(d_ltiles, d_ltoff, nlarge, &d_cnt->nlarge, &d_cnt->tiles));
            k_large_expand<<<grid_for(tiles, BS), BS, 0, st>>>(&d_cnt->nlarge, d_ltoff, d_ls, d_tseg, d_tidx, d_cnt);
            HIPC(hipGetLastError());
            int src = 0;
            for (int pass = 0; pass < 4; ++pass) {   // four passes: the result is back in buffer 0
                const int sh = 8 * pass;
                k_tile_hist<<<(tiles + NXCD - 1) / NXCD * NXCD, BS, 0, st>>>(&d_cnt->tiles, d_tseg, d_tidx, d_ls, d_pkey[src], sh, d_hist);
                HIPC(hipGetLastError());
                CHK(scan(d_hist, d_hscan, tiles * 256u, &d_cnt->tiles256));
                k_tile_scatter<<<(tiles + NXCD - 1) / NXCD * NXCD, BS, 0, st>>>(&d_cnt->tiles, d_tseg, d_tidx, d_ls, d_hscan, d_pkey[src],
                                                     d_pnix[src], sh, d_pkey[src ^ 1], d_pnix[src ^ 1]);
                HIPC(hipGetLastError());
                src ^= 1;
            }
        }
        return 0;
    }

    // ---------------------------------------------------------------- steps
    template <int GSZ>
    void launch_walk() {
        const uint32_t groups_per_block = BS / GSZ;
        k_walk<GSZ><<<grid_xcd(nalist, groups_per_block), BS, 0, st>>>(nalist, d_alist[acur], d_row, d_dint, d_nbr,
                                                                         d_y[cur], d_x[cur], d_x[cur ^ 1],
                                                                         d_y[cur ^ 1], d_xkey, heavy_thr, d_cnt);
    }
    template <int GSZ, int U, int WIN, int CHUNK, bool CLAMP = false, bool RANGE = false>
    void launch_walk_win() {
        const uint32_t nch = RANGE ? cdiv(N, CHUNK) : cdiv(nalist, CHUNK);
        const uint32_t grid = (std::min<uint32_t>(std::max<uint32_t>(nch, 1u), 4096u) + NXCD - 1) / NXCD * NXCD;
        const uint32_t* rm = RANGE ? range_map(CHUNK) : nullptr;
        k_walk_win<GSZ, U, WIN, CHUNK, CLAMP, RANGE><<<grid, WBS, 0, st>>>(
            nalist, d_alist[acur], d_row, d_dint, d_nbr, d_y[cur], d_x[cur], d_x[cur ^ 1], d_y[cur ^ 1], d_xkey,
            heavy_thr, N, d_cnt, rm);
    }
    // list entries of the node-range chunks of R indices, R = RANGE_R or
    // RANGE_R / 2 (valid until the list changes)
    const uint32_t* range_map(uint32_t R) {
        if (!rmap_valid) {
            rmap_valid = true;
            rmap_h_built = false;
            const uint32_t nch = cdiv(N, RANGE_R);
            k_range_offsets<<<grid_for((uint64_t)nch + 1, BS), BS, 0, st>>>(nch, RANGE_R, d_alist[acur], nalist, d_rmap);
        }
        if (R == RANGE_R) return d_rmap;
        if (!rmap_h_built) {
            rmap_h_built = true;
            const uint32_t nch = cdiv(N, RANGE_R / 2);
            k_range_offsets<<<grid_for((uint64_t)nch + 1, BS), BS, 0, st>>>(nch, RANGE_R / 2, d_alist[acur], nalist,
                                                                             d_rmap_h);
        }
        return d_rmap_h;
    }
    void launch_walk_variant(int v) {
        switch (v) {
            case 0: launch_walk<2>(); break;
            case 1: launch_walk<4>(); break;
            case 2: launch_walk<8>(); break;
            case 3: launch_walk<16>(); break;
            case 4: launch_walk<32>(); break;
            case 5: launch_walk<64>(); break;
            case WV_WIN1: launch_walk_win<1, 8, 20480, 4096>(); break;
            case WV_WIN2: launch_walk_win<2, 8, 20480, 4096>(); break;
            case WV_WIN2S: launch_walk_win<2, 6, 8192, 1024>(); break;
            case WV_WIN2C: launch_walk_win<2, 6, 20480, 4096, true>(); break;
            case WV_WIN2R:
                if (serving) launch_walk_win<2, 8, 20480, 4096>();   // list not in index order
                else launch_walk_win<2, 8, 20480, RANGE_R, false, true>();
                break;
            case WV_WIN2RH:
                if (serving) launch_walk_win<2, 8, 20480, 4096>();
                else launch_walk_win<2, 6, 10240, RANGE_R / 2, false, true>();
                break;
            case WV_WIN2H: launch_walk_win<2, 6, 10240, 2048>(); break;
            case WV_WIN1R:
                if (serving) launch_walk_win<1, 8, 20480, 4096>();   // list not in index order
                else launch_walk_win<1, 8, 20480, RANGE_R, false, true>();
                break;
            case WV_WIN1RH:
                if (serving) launch_walk_win<1, 8, 20480, 4096>();
                else launch_walk_win<1, 8, 10240, RANGE_R / 2, false, true>();
                break;
            default: launch_walk_win<4, 8, 16384, 4096>(); break;
        }
    }
    static int lanes_variant(int lanes) {
        int v = 0;
        while ((2 << v) < lanes && v < 5) ++v;
        return v;
    }
    // degree rule (measured on one MI355X): very sparse rows take 2 lanes
    // each; up to 24 arcs per row the LDS window with 2 lanes once the list
    // fills the chip with 4096-entry chunks (2^20 entries: 256 chunks, one
    // workgroup per CU), else 4 lanes; longer rows about four arcs per lane
    int default_variant() const {
        if (avg_deg <= 6.0) return WV_LANES2;
        if (avg_deg <= 24.0) return nalist >= (1u << 20) ? WV_WIN2 : lanes_variant(4);
        return lanes_variant(gs);
    }
    // variant for `pass` at the current list length; launch(v) runs candidate v
    template <class F>
    int pick_variant(int pass, F&& launch) {
        if (walk_force >= 0) return walk_force;
        const int def = default_variant();
        if (walk_force == -2 || nalist < TUNE_MIN) return def;
        int& slot = tuned_slot(pass, nalist);
        if (slot >= 0) return slot;
        std::vector<int> cand;
        const int lv = lanes_variant(gs);
        for (int v = std::max(0, lv - 1); v <= std::min(5, lv + 1); ++v) cand.push_back(v);
        if (gs <= 8) cand.push_back(WV_WIN2);
        if (gs <= 4) cand.push_back(WV_WIN1);
        if (gs >= 4 && gs <= 8) cand.push_back(WV_WIN4);
        // short chunks, small window: mid-size lists that 4096-entry chunks leave CUs idle on
        if (gs <= 8 && nalist < (1u << 22)) cand.push_back(WV_WIN2S);
        // the walk's clamped-slot form, for long lists of rows that fill their slots
        if (pass == 0 && gs <= 8 && nalist >= (1u << 22)) cand.push_back(WV_WIN2C);
        // node-range chunks: windows that cover sparse lists too (they need the
        // list in index order, which a serving engine's appended graphs break)
        if (gs <= 8 && !serving) {
            cand.push_back(WV_WIN2R);
            cand.push_back(WV_WIN2RH);
        }
        // half windows (two workgroups a CU) over 2048-entry chunks
        if (gs <= 8) cand.push_back(WV_WIN2H);
        // (14 and 15, one lane per row over node ranges, are pinnable only:
        // pinned on the 512-SBM batch they were the slowest window kernels at
        // every list length)
        int best = def;
        float best_ms = 0.f;
        for (int v : cand) {
            for (int rep = 0; rep < 3; ++rep) {          // the fastest of three launches
                if (hipEventRecord(tev0, st) != hipSuccess) return def;
                launch(v);
                if (hipGetLastError() != hipSuccess) break;   // a candidate that cannot launch drops out
                if (hipEventRecord(tev1, st) != hipSuccess || hipEventSynchronize(tev1) != hipSuccess)
                    return def;                               // a failed stream shows at the caller's launch
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, tev0, tev1) != hipSuccess) return def;
                if (best_ms == 0.f || ms < best_ms) {
                    best = v;
                    best_ms = ms;
                }
            }
        }
        slot = best;
        ++walk_tunings;
        return best;
    }
    template <int GSZ>
    void launch_contrib() {
        k_contrib<GSZ><<<grid_xcd(nalist, BS / GSZ), BS, 0, st>>>(nalist, d_alist[acur], d_row, d_dint, d_nbr,
                                                                     d_xkey, d_ediff, heavy_thr);
    }
    template <int GSZ, int U, int WIN, int CHUNK, bool RANGE = false>
    void launch_contrib_win() {
        const uint32_t nch = RANGE ? cdiv(N, CHUNK) : cdiv(nalist, CHUNK);
        const uint32_t grid = (std::min<uint32_t>(std::max<uint32_t>(nch, 1u), 4096u) + NXCD - 1) / NXCD * NXCD;
        const uint32_t* rm = RANGE ? range_map(CHUNK) : nullptr;
        k_contrib_win<GSZ, U, WIN, CHUNK, RANGE><<<grid, WBS, 0, st>>>(nalist, d_alist[acur], d_row, d_dint, d_nbr,
                                                                      d_xkey, d_ediff, heavy_thr, N, rm);
    }
    void launch_contrib_variant(int v) {
        switch (v) {
            case 0: launch_contrib<2>(); break;
            case 1: launch_contrib<4>(); break;
            case 2: launch_contrib<8>(); break;
            case 3: launch_contrib<16>(); break;
            case 4: launch_contrib<32>(); break;
            case 5: launch_contrib<64>(); break;
            case WV_WIN1: launch_contrib_win<1, 8, 40960, 4096>(); break;
            case WV_WIN2: launch_contrib_win<2, 8, 40960, 4096>(); break;
            case WV_WIN2S: launch_contrib_win<2, 8, 16384, 1024>(); break;
            case WV_WIN2C: launch_contrib_win<2, 8, 40960, 4096>(); break;   // no clamped form
            case WV_WIN2R:
                if (serving) launch_contrib_win<2, 8, 40960, 4096>();   // list not in index order
                else launch_contrib_win<2, 8, 40960, RANGE_R, true>();
                break;
            case WV_WIN2RH:
                if (serving) launch_contrib_win<2, 8, 40960, 4096>();
                else launch_contrib_win<2, 8, 20480, RANGE_R / 2, true>();
                break;
            case WV_WIN2H: launch_contrib_win<2, 8, 20480, 2048>(); break;
            case WV_WIN1R:
                if (serving) launch_contrib_win<1, 8, 40960, 4096>();
                else launch_contrib_win<1, 8, 40960, RANGE_R, true>();
                break;
            case WV_WIN1RH:
                if (serving) launch_contrib_win<1, 8, 40960, 4096>();
                else launch_contrib_win<1, 8, 20480, RANGE_R / 2, true>();
                break;
            default: launch_contrib_win<4, 8, 32768, 4096>(); break;
        }
    }
    template <int GSZ>
    void launch_link(const uint32_t* list, uint32_t nl, const uint32_t* pn) {
        k_cc_link<GSZ><<<grid_xcd(nl, BS / GSZ), BS, 0, st>>>(nl, list, d_row, d_dint, d_nbr, d_part, d_root,
                                                                heavy_thr, d_cnt, pn, &d_cnt->cc_done);
    }
    void launch_disable_win(int lanes, uint32_t nl) {
        constexpr uint32_t CHUNK = 4096;
        const uint32_t grid = (std::min<uint32_t>(std::max<uint32_t>(cdiv(nl, CHUNK), 1u), 4096u) + NXCD - 1) / NXCD * NXCD;
        if (lanes == 1)
            k_disable_win<1, 4, 40960, CHUNK><<<grid, WBS, 0, st>>>(nl, d_cclist, d_alive, d_row, d_dint, d_nbr,
                                                                    d_root, d_x[cur], d_y[cur], heavy_thr, N,
                                                                    &d_cnt->err);
        else if (lanes == 2)
            k_disable_win<2, 8, 40960, CHUNK><<<grid, WBS, 0, st>>>(nl, d_cclist, d_alive, d_row, d_dint, d_nbr,
                                                                    d_root, d_x[cur], d_y[cur], heavy_thr, N,
                                                                    &d_cnt->err);
        else
            k_disable_win<4, 4, 32768, CHUNK><<<grid, WBS, 0, st>>>(nl, d_cclist, d_alive, d_row, d_dint, d_nbr,
                                                                    d_root, d_x[cur], d_y[cur], heavy_thr, N,
                                                                    &d_cnt->err);
    }
    template <int GSZ>
    void launch_disable(uint32_t nl) {
        k_disable<GSZ><<<grid_xcd(nl, BS / GSZ), BS, 0, st>>>(nl, d_cclist, d_alive, d_row, d_dint, d_nbr, d_root,
                                                                d_x[cur], d_y[cur], heavy_thr, &d_cnt->err);
    }

    // chunk map of the active rows above heavy_thr (valid until the next cut)
    int heavy_map() {
        if (heavy_thr == 0xFFFFFFFFu || heavy_valid || nalist == 0) return 0;
        k_heavy_count<<<grid_for(nalist, BS), BS, 0, st>>>(nalist, d_alist[acur], d_dint, heavy_thr, d_hcnt);
        HIPC(hipGetLastError());
        CHK(scan(d_hcnt, d_hoff, nalist));
        k_heavy_expand<<<grid_for(nalist, BS), BS, 0, st>>>(nalist, d_hcnt, d_hoff, d_hrow);
        HIPC(hipGetLastError());
        heavy_valid = true;
        return 0;
    }
    uint32_t heavy_grid() const { return std::min(heavy_cap, 4096u); }

#define PXD_DISPATCH_GS(fn, ...)                          \
    switch (gs) {                                         \
        case 4: fn<4>(__VA_ARGS__); break;                \
        case 8: fn<8>(__VA_ARGS__); break;                \
        case 16: fn<16>(__VA_ARGS__); break;              \
        case 32: fn<32>(__VA_ARGS__); break;              \
        default: fn<64>(__VA_ARGS__); break;              \
    }

    int walk() {
        if (!inited) return set_err(PXD_ERR_ARG, "engine not initialized");
        if (nalist) {
            const int wv = pick_variant(0, [this](int v) { launch_walk_variant(v); });
            if (walk_timing) HIPC(hipEventRecord(ev0, st));
            CHK(heavy_map());
            launch_walk_variant(wv);
            HIPC(hipGetLastError());
            if (heavy_thr != 0xFFFFFFFFu) {
                const uint32_t* H = d_hoff + nalist;   // total chunk count, on the device
                k_walk_heavy<<<heavy_grid(), BS, 0, st>>>(H, d_hrow, d_hoff, d_alist[acur], d_row, d_dint, d_nbr,
                                                           d_y[cur], d_hpart, d_cnt);
                k_walk_heavy_finish<<<grid_for(heavy_cap, BS), BS, 0, st>>>(H, d_hrow, d_hoff, d_alist[acur], d_dint,
                                                                             d_hpart, d_x[cur], d_x[cur ^ 1],
                                                                             d_y[cur ^ 1], d_xkey, d_cnt);
                HIPC(hipGetLastError());
            }
            if (walk_timing) {
                HIPC(hipEventRecord(ev1, st));
                HIPC(hipEventSynchronize(ev1));
                float ms = 0.f;
                HIPC(hipEventElapsedTime(&ms, ev0, ev1));
                // valid arcs of active nodes: sum of dInt over the node list
                HIPC(hipMemsetAsync(d_u64, 0, 8, st));
                // (dInt of inactive nodes is stale; sum only over the active list)
                k_gather<<<grid_xcd(nalist, BS * GATHER_ITEMS), BS, 0, st>>>(nalist, d_alist[acur], d_dint, d_tC);
                k_sum_u32<<<grid_for(nalist, BS), BS, 0, st>>>(nalist, d_tC, d_u64);
                unsigned long long arcs = 0;
                HIPC(hipMemcpyAsync(&arcs, d_u64, 8, hipMemcpyDeviceToHost, st));
                CHK(sync());
                walk_launches++;
                walk_ms += ms;
                walk_bytes += 12.0 * (double)arcs + 32.0 * (double)nalist;
                walk_log_ms.push_back(ms);
                walk_log_bytes.push_back(12.0 * (double)arcs + 32.0 * (double)nalist);
            }
            cur ^= 1;                 // the walk wrote x' and y' = x'/dInt into the other buffers
        } else {
            cur ^= 1;
        }
        sweep_done = false;
        return 0;
    }

    int sweep() {
        if (!inited) return set_err(PXD_ERR_ARG, "engine not initialized");
        ClusterTable& T = ct[ctc];
        ncut = 0;
        ncc = 0;
        ncc_members = 0;
        if (na == 0 || nc == 0) { sweep_done = true; return 0; }
        // keys in position order, then stable sort inside every cluster
        k_gather<<<grid_xcd(na, BS * GATHER_ITEMS), BS, 0, st>>>(na, d_pnix[0], d_xkey, d_pkey[0]);
        HIPC(hipGetLastError());
        CHK(seg_sort(T.cs, nc, na));
        CHK(heavy_map());
        launch_contrib_variant(pick_variant(1, [this](int v) { launch_contrib_variant(v); }));
        HIPC(hipGetLastError());
        if (heavy_thr != 0xFFFFFFFFu) {
            k_contrib_heavy<<<heavy_grid(), BS, 0, st>>>(d_hoff + nalist, d_hrow, d_hoff, d_alist[acur], d_row, d_dint,
                                                          d_nbr, d_xkey, d_ediff);
            HIPC(hipGetLastError());
        }
        k_sweep_reset<<<grid_for(nc, BS), BS, 0, st>>>(nc, d_cbest, d_cnt);
        const uint32_t nb = cdiv(na, BS * SCAN_ITEMS);
        k_sweep_a<<<nb, BS, 0, st>>>(na, d_pnix[0], d_pclu[0], T.cs, d_ediff, d_ped, d_pvol, d_tail_ed,
                                     d_tail_vol, d_hasst);
        HIPC(hipGetLastError());
        k_sweep_b<<<1, 1024, 0, st>>>(nb, d_tail_ed, d_tail_vol, d_hasst, d_carry_ed, d_carry_vol);
        HIPC(hipGetLastError());
        k_sweep_c<<<nb, BS, 0, st>>>(na, d_pnix[0], d_pclu[0], T.cs, T.cvol, d_ped, d_pvol, d_carry_ed, d_carry_vol,
                                     d_cbest, P.record_node_data ? d_nd : nullptr);
        HIPC(hipGetLastError());
        k_decide<<<grid_for(nc, BS), BS, 0, st>>>(nc, d_cbest, (double)P.sc_threshold, T.cflag, d_coff, d_csp, T.cs,
                                                   d_cnt);
        HIPC(hipGetLastError());
        CHK(read_counters());
        ncut = hc.ncut;
        ncc = hc.ncc;
        ncc_members = hc.ccn;
        sweep_done = true;
        return 0;
    }

    // One host read per cut: after the compaction of positions and clusters
    // (kept positions and clusters; in serving also every graph's iteration
    // count).  Until then the cluster count after a CC regroup is known on the
    // device only: the kernels read it from the counters and their grids are
    // sized from the bound nc + (CC members).
    int cut() {
        if (!inited) return set_err(PXD_ERR_ARG, "engine not initialized");
        if (!sweep_done) return set_err(PXD_ERR_ARG, "cutClusters needs a preceding computeSweepCuts");
        sweep_done = false;
        heavy_valid = false;          // the cut changes the active list and internal degrees
        rmap_valid = false;
        giter_read = false;
        ++iter1;
        if (na == 0 || nc == 0) return 0;
        ClusterTable* T = &ct[ctc];
        uint32_t nl = 0;
        uint32_t ncb = nc;                    // cluster count, or a bound on it
        const uint32_t* pnc = nullptr;        // the device count when ncb is a bound
        if (ncc > 0) {
            nl = ncc_members;
            k_mark<<<grid_xcd(na, BS), BS, 0, st>>>(na, d_pnix[0], d_pclu[0], T->cs, T->cflag, d_coff, iter1, d_part,
                                                    d_root, d_stamp);
            // the CC list in node order (the active list holds the positions' nodes);
            // its length is the member count the sweep returned
            HIPC(hipGetLastError());
            CHK(select(nalist, SelMembers{d_alist[acur], d_stamp, iter1, d_cclist, d_root}));
            CHK(connected_components(nl, *T));
            // regroup: stable sort of all positions by component / cluster minimum
            k_cc_flatten<<<grid_for(nl, BS), BS, 0, st>>>(nl, d_cclist, d_root);
            k_rid<<<grid_xcd(na, BS), BS, 0, st>>>(na, d_pnix[0], d_pclu[0], T->cflag, T->cmin, d_root, d_pkey[0]);
            k_one_segment<<<1, 64, 0, st>>>(d_oneseg, na, nullptr);
            HIPC(hipGetLastError());
            CHK(seg_sort(d_oneseg, 1, na));
            CHK(select(na, SelRuns{na, d_pkey[0], d_ngraph, T->cs, T->cmin, T->cgraph, T->cflag, d_pclu[0]}, nullptr,
                       &d_cnt->ncr));
            // components of the CC clusters are at most their members
            ncb = (uint32_t)std::min<uint64_t>(na, (uint64_t)nc + nl);
            pnc = &d_cnt->ncr;
        }
        // statistics of the (possibly new) clusters
        const uint32_t gc = grid_for(ncb, BS);
        // positions of every cluster are in ascending walk-key order (sweep sort,
        // then the stable regroup): min and max are the first and last keys
        k_stats_prep<<<gc, BS, 0, st>>>(ncb, T->cs, T->cgraph, d_pnix[0], d_xkey, d_gfirst, d_cmaxo, d_cmino, d_cK,
                                        d_csum, pnc);
        k_stats_sum<<<grid_xcd(na, BS * SUM_ITEMS), BS, 0, st>>>(na, d_pnix[0], d_pclu[0], d_x[cur], d_cK, d_csum,
                                                                 d_nclu);
        k_stats_final<<<gc, BS, 0, st>>>(ncb, T->cs, d_cmaxo, d_cmino, d_cK, d_csum, P.rw_threshold, d_ckeep, d_cavg,
                                         pnc);
        HIPC(hipGetLastError());
        CHK(scan(d_ckeep, d_kscan, ncb, pnc, &d_cnt->nc_new));
        k_labels<<<gc, BS, 0, st>>>(ncb, T->cgraph, d_gfirst, d_ckeep, d_kscan, d_gsmall, d_gsmall_next, d_gsurv,
                                    T->clabel, d_gseen, pnc);
        k_graph_update<<<grid_for(G, BS), BS, 0, st>>>(G, d_gseen, d_gsmall, d_gsmall_next, d_gsurv, d_giter,
                                                       d_gstart, iter1);
        // the active node list holds the same nodes as the positions: apply in
        // index order (k_stats_sum recorded every node's cluster)
        k_apply_nodes<<<grid_for(nalist, BS), BS, 0, st>>>(nalist, d_alist[acur], d_nclu, d_ckeep, T->clabel, d_cavg,
                                                           d_dint, d_label, d_x[cur], d_y[cur], d_alive, &d_cnt->err);
        HIPC(hipGetLastError());
        // drop deactivated clusters from the position list, the cluster table and
        // the node list (d_tB: every kept cluster's new first position)
        CHK(select(na, SelPositions{d_pnix[0], d_pclu[0], d_ckeep, d_kscan, T->cs, d_pnix[1], d_pclu[1], d_tB},
                   nullptr, &d_cnt->na_new));
        ClusterTable& T2 = ct[ctc ^ 1];
        k_clu_compact<<<gc, BS, 0, st>>>(ncb, 0u, d_ckeep, d_kscan, d_tB, T->cols(), T2.cols(), pnc, &d_cnt->na_new);
        HIPC(hipGetLastError());
        if (serving && G) {
            HIPC(hipMemcpyAsync(h_giter.data(), d_giter, G * 4, hipMemcpyDeviceToHost, st));
            giter_read = true;
        }
        CHK(read_counters());
        const uint32_t na_new = hc.na_new, nc_new = hc.nc_new;
        std::swap(d_pnix[0], d_pnix[1]);
        std::swap(d_pclu[0], d_pclu[1]);
        ctc ^= 1;
        if (na_new != na) {
            CHK(select(nalist, SelAlive{d_alist[acur], d_alive, d_alist[acur ^ 1]}));
            acur ^= 1;
            nalist = na_new;
        }
        na = na_new;
        nc = nc_new;
        if (nl) {
            // the LDS-window kernel when pinned, else for long lists of short
            // rows (the pass mutates the rows, so it cannot be timed by rerunning)
            const int wg = walk_force >= WV_WIN1
                               ? (walk_force == WV_WIN1 ? 1 : walk_force == WV_WIN4 ? 4 : 2)
                           : (walk_force < 0 && avg_deg <= 24.0 && avg_deg > 6.0 && nl >= (1u << 20)) ? 2 : 0;
            if (wg) launch_disable_win(wg, nl);
            else PXD_DISPATCH_GS(launch_disable, nl);
            if (heavy_thr != 0xFFFFFFFFu) {
                heavy_valid = false;
                rmap_valid = false;
                k_heavy_flags<<<grid_for(nl, BS), BS, 0, st>>>(nl, d_cclist, d_alive, d_dint, heavy_thr, d_hcnt);
                CHK(scan(d_hcnt, d_hoff, nl));
                k_compact_u32<<<grid_for(nl, BS), BS, 0, st>>>(nl, d_hcnt, d_hoff, d_cclist, d_hrow);
                k_disable_heavy<<<heavy_grid(), BS, 0, st>>>(d_hoff + nl, d_hrow, d_row, d_dint, d_nbr, d_root,
                                                              d_x[cur], d_y[cur], &d_cnt->err);
            }
            HIPC(hipGetLastError());
        }
        if (nc) {                     // k_clu_compact cleared the volumes
            ClusterTable& T3 = ct[ctc];
            k_volume<<<grid_xcd(na, BS * SUM_ITEMS), BS, 0, st>>>(na, d_pnix[0], d_pclu[0], d_dint, T3.cvol);
            HIPC(hipGetLastError());
        }
        return 0;
    }


    // Components of every CC part (clusters flagged CF_CC in table T):
    // sampling rounds over the first arcs of each node, each followed by one
    // pointer-jumping launch (trees need not end flat: roots are found by
    // following pointers), then exact linking rounds over the members outside
    // their part's giant tree.  The regroup pass (k_rid) leaves every tree flat.
    // The off-giant count stays on the device; the linking rounds run in
    // batches, each round ended by k_cc_round, which stops the later rounds
    // of a batch once one changed nothing: one host read per batch.
    int connected_components(uint32_t nl, const ClusterTable& T) {
        if (nl == 0) return 0;
        // the LDS window when the walk's choice for this list length is a window
        // kernel (neighbours near in index order) and the list fills the chip;
        // at any length when a window variant is pinned
        const int wv = walk_force >= 0 ? walk_force
                                       : (nalist >= TUNE_MIN ? tuned_slot(0, nalist) : -1);
        const bool win = wv >= WV_WIN1 && (walk_force >= 0 || nl >= (1u << 20));
        for (uint32_t r = 0; r < CC_SAMPLE_ROUNDS; ++r) {
            if (win) {
                // 8192-entry chunks: the staged window (160 KB) is shared by twice
                // the entries of a 4096 chunk (C4: 382 us instead of 419 us a launch)
                const uint32_t grid = (std::min<uint32_t>(cdiv(nl, 8192), 4096u) + NXCD - 1) / NXCD * NXCD;
                k_cc_sample_win<20480, 8192><<<grid, WBS, 0, st>>>(nl, d_cclist, d_row, d_dint, d_nbr, d_part,
                                                                   d_root, r, N, d_ccsamp);
            } else {
                k_cc_sample<<<grid_xcd(nl, BS), BS, 0, st>>>(nl, d_cclist, d_row, d_dint, d_nbr, d_part, d_root, r,
                                                             d_ccsamp);
            }
            // no jump after the last round: the giant roots and the off-giant
            // selection follow pointers themselves
            if (r + 1 < CC_SAMPLE_ROUNDS) k_cc_jump<<<grid_for(nl, BS), BS, 0, st>>>(nl, d_cclist, d_root);
        }
        k_cc_giant<<<grid_for(nc, BS), BS, 0, st>>>(nc, T.cs, T.cflag, d_coff, d_pnix[0], d_root, d_giant, d_cnt);
        HIPC(hipGetLastError());
        // the members outside the giant trees, in list order; their count stays on the device
        const uint32_t* pnu = &d_cnt->nu;
        CHK(select(nl, SelOffGiant{d_cclist, d_part, d_giant, d_root, d_cclist2}, nullptr, &d_cnt->nu));
        const bool heavy = heavy_thr != 0xFFFFFFFFu;
        if (heavy) {                  // chunk map of the list's long rows (the walk's map is not needed now)
            heavy_valid = false;
            rmap_valid = false;
            k_heavy_count<<<grid_for(nl, BS), BS, 0, st>>>(nl, d_cclist2, d_dint, heavy_thr, d_hcnt, pnu);
            CHK(scan(d_hcnt, d_hoff, nl, pnu, &d_cnt->hchunks));
            k_heavy_expand<<<grid_for(nl, BS), BS, 0, st>>>(nl, d_hcnt, d_hoff, d_hrow, pnu);
            HIPC(hipGetLastError());
        }
        uint32_t r = 0;
        for (uint32_t batch = 3;; batch = std::min(2 * batch, 24u)) {
            for (uint32_t b = 0; b < batch; ++b) {
                PXD_DISPATCH_GS(launch_link, d_cclist2, nl, pnu);
                if (heavy)
                    k_cc_link_heavy<<<heavy_grid(), BS, 0, st>>>(&d_cnt->hchunks, d_hrow, d_hoff, d_cclist2, d_row,
                                                                  d_dint, d_nbr, d_part, d_root, d_cnt,
                                                                  &d_cnt->cc_done);
                // shorten the trees the links deepened (finds stay exact without it)
                k_cc_jump<<<grid_for(nl, BS), BS, 0, st>>>(nl, d_cclist2, d_root, pnu, &d_cnt->cc_done);
                k_cc_round<<<1, 64, 0, st>>>(d_cnt, ++r);
                HIPC(hipGetLastError());
            }
            uint32_t done = 0;
            CHK(read_u32(&d_cnt->cc_done, &done));
            if (done) break;
            // every changed round lowers a pointer, so this bound is never reached
            if (r >= 100000) return set_err(PXD_ERR_STATE, "connected components did not converge (%u members)", nl);
        }
        return 0;                     // k_cc_flatten then sets every member's pointer to its root
    }

    // ---------------------------------------------------------------- serving
    // Graphs join a running engine between iterations and leave when they
    // finish.  Node ranges, arc ranges and graph slots are taken first fit
    // from fixed capacities (admission by device footprint) and reused.
    bool serving = false;
    std::map<uint64_t, uint64_t> free_nodes, free_arcs;   // start -> length
    struct Slot {
        uint64_t tag = 0;
        uint32_t base = 0, n = 0, abase = 0, arcs = 0;
        bool used = false;
    };
    std::vector<Slot> slots;
    std::vector<uint32_t> free_slots;
    uint32_t live_graphs = 0;
    uint64_t live_nodes = 0, live_arcs = 0;
    static constexpr int32_t RANK_FAILED = -2;   // no host memory to rank the labels
    struct Done {
        uint64_t tag;
        int32_t k, it;                // k = -1: the graph failed the device check; RANK_FAILED
        std::vector<int32_t> lab;     // raw labels, then cluster ids once ranked
        bool ranked;
        uint32_t n;                   // node count (lab is moved out while it is ranked)
    };
    // Finished results in completion order.  A worker thread ranks the labels
    // (cuda/devPartition.h:798-823) off the stepping thread; next/take wait for
    // the oldest one only.
    std::deque<Done> done;
    std::mutex done_mu;
    std::condition_variable done_cv;
    std::thread ranker;
    bool ranker_stop = false;

    void push_done(Done&& D) {
        {
            std::lock_guard<std::mutex> lk(done_mu);
            done.push_back(std::move(D));
        }
        done_cv.notify_all();
    }
    void rank_loop() {
        std::unique_lock<std::mutex> lk(done_mu);
        for (;;) {
            Done* D = nullptr;
            for (Done& d : done)
                if (!d.ranked) { D = &d; break; }
            if (!D) {
                if (ranker_stop) return;
                done_cv.wait(lk);
                continue;
            }
            std::vector<int32_t> lab = std::move(D->lab);   // deque elements keep their address
            lk.unlock();
            std::vector<int32_t> ids;
            int32_t k = 0;
            bool ok = true;
            try {                     // an exception here would end the process
                ids.resize(lab.size());
                k = dense_rank(lab.data(), (uint32_t)lab.size(), ids.data());
            } catch (...) {
                ok = false;
            }
            lk.lock();
            if (ok) {
                D->lab = std::move(ids);
                if (D->k >= 0) D->k = k;
            } else {                  // no room to rank: an error result without ids
                D->lab.clear();
                D->n = 0;
                D->k = RANK_FAILED;
            }
            D->ranked = true;
            done_cv.notify_all();
        }
    }
    void stop_ranker() {
        if (!ranker.joinable()) return;
        {
            std::lock_guard<std::mutex> lk(done_mu);
            ranker_stop = true;
        }
        done_cv.notify_all();
        ranker.join();
        ranker_stop = false;
    }
    // admissions wait as `pending` until the next step checks and appends them
    struct Pending {
        uint64_t tag;
        uint32_t base, n, abase, arcs, slot, stage;
    };
    std::vector<Pending> pending;
    // edge-list admissions: the CSR is built on the device in a workspace for
    // graphs of up to ing_n nodes and ing_m edges (pxd_serve_enable_edges)
    uint32_t ing_n = 0;
    uint64_t ing_m = 0;
    uint32_t *d_ifx = nullptr, *d_itx = nullptr;
    IngestWork ing{};
    uint32_t stage_used = 0;          // bump pointer into d_stage (N + G + 1 entries)
    uint32_t* d_stage = nullptr;      // raw ranges of pending graphs
    uint32_t* d_pbad = nullptr;       // device check result per pending graph
    std::vector<uint32_t> h_pbad;
    std::vector<int32_t> h_giter;

    static bool take_range(std::map<uint64_t, uint64_t>& fl, uint64_t len, uint64_t* start) {
        if (len == 0) { *start = 0; return true; }
        for (auto it = fl.begin(); it != fl.end(); ++it) {
            if (it->second < len) continue;
            *start = it->first;
            const uint64_t rest = it->second - len, ns = it->first + len;
            fl.erase(it);
            if (rest) fl[ns] = rest;
            return true;
        }
        return false;
    }
    static void give_range(std::map<uint64_t, uint64_t>& fl, uint64_t start, uint64_t len) {
        if (len == 0) return;
        auto it = fl.emplace(start, len).first;
        auto nx = std::next(it);
        if (nx != fl.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            fl.erase(nx);
        }
        if (it != fl.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                fl.erase(it);
            }
        }
    }

    int serve_create(uint32_t node_cap, uint64_t arc_cap, uint32_t graph_cap) {
        release();
        HIPC(hipSetDevice(dev));
        if (node_cap == 0 || node_cap >= 0x7FFFFFFFu || arc_cap >= 0xFFFFFFFFull || graph_cap == 0)
            return set_err(PXD_ERR_ARG, "serving capacities out of range (nodes < 2^31, arcs < 2^32, slots > 0)");
        N = node_cap;
        A = (uint32_t)arc_cap;
        G = graph_cap;
        M_total = 0;
        h_gnode.assign(G + 1, 0);
        h_garc.assign(G + 1, 0);
        h_gm.assign(G, 0);
        P.keep_original = 0;          // no reset: the input CSR is not kept
        CHK(allocate());
        CHK(alloc(&d_stage, (size_t)N + G + 1));
        CHK(alloc(&d_pbad, G));
        h_pbad.assign(G, 0);
        pending.clear();
        stage_used = 0;
        // heavy-row buffers sized for the smallest threshold (4 lanes per node)
        heavy_cap = A / HEAVY_CHUNK + A / 128u + 1u;
        CHK(alloc(&d_hcnt, N));
        CHK(alloc(&d_hoff, (size_t)N + 1));
        CHK(alloc(&d_hrow, heavy_cap));
        CHK(alloc(&d_hpart, heavy_cap));
        HIPC(hipMemsetAsync(d_cnt, 0, sizeof(Counters), st));
        HIPC(hipMemsetAsync(d_gsmall, 0, G * 4, st));
        HIPC(hipMemsetAsync(d_gsmall_next, 0, G * 4, st));
        HIPC(hipMemsetAsync(d_giter, 0, G * 4, st));
        HIPC(hipMemsetAsync(d_gsurv, 0, G * 4, st));
        HIPC(hipMemsetAsync(d_gseen, 0, G * 4, st));
        HIPC(hipMemsetAsync(d_gstart, 0, G * 4, st));
        free_nodes.clear();
        free_nodes[0] = N;
        free_arcs.clear();
        if (A) free_arcs[0] = A;
        slots.assign(G, Slot{});
        free_slots.clear();
        for (uint32_t g = G; g-- > 0;) free_slots.push_back(g);
        live_graphs = 0;
        live_nodes = live_arcs = 0;
        stop_ranker();
        done.clear();
        ranker = std::thread([this] { rank_loop(); });
        h_giter.assign(G, 0);
        cur = acur = pcur = 0;
        ctc = 0;
        na = nalist = nc = 0;
        ncut = ncc = 0;
        iter1 = 0;
        sweep_done = false;
        maxdeg = 0;
        gs = 4;
        heavy_thr = 128u;
        heavy_valid = false;
        rmap_valid = false;
        std::fill(&tuned[0][0][0], &tuned[0][0][0] + 2 * 33 * TUNE_Q, -1);
        tuned_for[0] = tuned_for[1] = tuned_for[2] = ~0ull;
        serving = true;
        inited = true;
        return sync();
    }

    // PXD_FULL when the graph does not fit the free capacity right now.  The
    // uploads are queued; the graph is checked on the device and joins at the
    // next step (the caller's buffers must stay valid until that step returns).
    int serve_admit(uint32_t n, uint32_t m, const uint32_t* ranges, const uint32_t* edges, uint64_t tag,
                    int32_t* slot_out) {
        if (!serving) return set_err(PXD_ERR_ARG, "not a serving engine");
        if (!slot_out || (!ranges) || (m && !edges)) return set_err(PXD_ERR_ARG, "null argument");
        const uint64_t arcs = 2ull * m;
        *slot_out = -1;
        if (ranges[0] != 0 || ranges[n] != arcs)
            return set_err(PXD_ERR_ARG, "ranges[n] must equal 2m and ranges[0] == 0");
        if (n == 0) {                 // nothing to decompose: an empty partition right away
            push_done(Done{tag, 0, 1, {}, true, 0u});
            return PXD_OK;
        }
        if (n > N || arcs > A) return set_err(PXD_ERR_ARG, "graph larger than the serving capacity");
        if (free_slots.empty()) return PXD_FULL;
        uint64_t nb = 0, ab = 0;
        if (!take_range(free_nodes, n, &nb)) return PXD_FULL;
        if (!take_range(free_arcs, arcs, &ab)) {
            give_range(free_nodes, nb, n);
            return PXD_FULL;
        }
        const uint32_t g = free_slots.back();
        free_slots.pop_back();
        slots[g] = Slot{tag, (uint32_t)nb, n, (uint32_t)ab, (uint32_t)arcs, true};
        // pending node counts stay within the free capacity, so the arena (N + G + 1) has room
        const uint32_t k = (uint32_t)pending.size();
        if (k == 0) HIPC(hipMemsetAsync(d_pbad, 0, G * 4, st));
        HIPC(hipMemcpyAsync(d_stage + stage_used, ranges, ((size_t)n + 1) * 4, hipMemcpyHostToDevice, st));
        if (arcs) HIPC(hipMemcpyAsync(d_nbr + ab, edges, arcs * 4, hipMemcpyHostToDevice, st));
        k_admit_check<<<grid_for(std::max<uint64_t>(n, arcs), BS), BS, 0, st>>>(n, d_stage + stage_used, arcs,
                                                                              d_nbr + ab, d_pbad + k);
        HIPC(hipGetLastError());
        pending.push_back(Pending{tag, (uint32_t)nb, n, (uint32_t)ab, (uint32_t)arcs, g, stage_used});
        stage_used += n + 1;
        live_graphs++;
        live_nodes += n;
        live_arcs += arcs;
        *slot_out = (int32_t)g;
        return PXD_OK;
    }

    int serve_enable_edges(uint32_t max_n, uint64_t max_m) {
        if (!serving) return set_err(PXD_ERR_ARG, "not a serving engine");
        if (ing_n) return set_err(PXD_ERR_ARG, "edge admissions are already enabled");
        if (max_n == 0 || max_n > N || 2 * max_m > A || 2 * max_m > 0x7FFFFFFFull)
            return set_err(PXD_ERR_ARG, "edge-admission limits beyond the serving capacity or 2^31 arcs");
        CHK(alloc(&d_ifx, max_m));
        CHK(alloc(&d_itx, max_m));
        uint32_t* work = nullptr;
        CHK(alloc(&work, ingest_work_words(max_n, 2 * max_m)));
        ing = ingest_carve(work, max_n, 2 * max_m);
        ing_n = max_n;
        ing_m = max_m;
        return 0;
    }

    // admission of an undirected edge list: room as for serve_admit, then the
    // CSR is built on the device straight into the staged ranges and the arc
    // range; an end >= n fails the graph's device check like a bad CSR
    int serve_admit_edges(uint32_t n, uint64_t m, const uint32_t* fx, const uint32_t* tx, uint64_t tag,
                          int32_t* slot_out) {
        if (!serving) return set_err(PXD_ERR_ARG, "not a serving engine");
        if (!ing_n) return set_err(PXD_ERR_ARG, "edge admissions need pxd_serve_enable_edges first");
        if (!slot_out || (m && (!fx || !tx))) return set_err(PXD_ERR_ARG, "null argument");
        *slot_out = -1;
        if (n == 0) {
            if (m) return set_err(PXD_ERR_ARG, "edges in a graph without nodes");
            push_done(Done{tag, 0, 1, {}, true, 0u});
            return PXD_OK;
        }
        if (n > ing_n || m > ing_m) return set_err(PXD_ERR_ARG, "graph larger than the edge-admission limits");
        const uint64_t arcs = 2 * m;
        if (free_slots.empty()) return PXD_FULL;
        uint64_t nb = 0, ab = 0;
        if (!take_range(free_nodes, n, &nb)) return PXD_FULL;
        if (!take_range(free_arcs, arcs, &ab)) {
            give_range(free_nodes, nb, n);
            return PXD_FULL;
        }
        const uint32_t g = free_slots.back();
        free_slots.pop_back();
        slots[g] = Slot{tag, (uint32_t)nb, n, (uint32_t)ab, (uint32_t)arcs, true};
        const uint32_t k = (uint32_t)pending.size();
        if (k == 0) HIPC(hipMemsetAsync(d_pbad, 0, G * 4, st));
        if (m) {
            HIPC(hipMemcpyAsync(d_ifx, fx, m * 4, hipMemcpyHostToDevice, st));
            HIPC(hipMemcpyAsync(d_itx, tx, m * 4, hipMemcpyHostToDevice, st));
        }
        HIPC(ingest_csr_async(n, m, d_ifx, d_itx, d_stage + stage_used, d_nbr + ab, ing, d_pbad + k, st));
        pending.push_back(Pending{tag, (uint32_t)nb, n, (uint32_t)ab, (uint32_t)arcs, g, stage_used});
        stage_used += n + 1;
        live_graphs++;
        live_nodes += n;
        live_arcs += arcs;
        *slot_out = (int32_t)g;
        return PXD_OK;
    }

    // checked pending graphs join (positions, active list, one cluster each);
    // failed ones return their room and leave an error result
    int join_pending() {
        if (pending.empty()) return 0;
        HIPC(hipMemcpyAsync(h_pbad.data(), d_pbad, pending.size() * 4, hipMemcpyDeviceToHost, st));
        CHK(sync());                  // also: every queued upload has completed
        for (size_t k = 0; k < pending.size(); ++k) {
            const Pending& Q = pending[k];
            if (h_pbad[k]) {
                give_range(free_nodes, Q.base, Q.n);
                give_range(free_arcs, Q.abase, Q.arcs);
                slots[Q.slot].used = false;
                free_slots.push_back(Q.slot);
                live_graphs--;
                live_nodes -= Q.n;
                live_arcs -= Q.arcs;
                push_done(Done{Q.tag, -1, 0, {}, true, 0u});
                continue;
            }
            k_admit_nodes<<<grid_for(Q.n, BS), BS, 0, st>>>(Q.base, Q.n, Q.abase, Q.slot, nc, d_stage + Q.stage, P.seed,
                                                            d_row, d_deg, d_dint, d_x[cur], d_y[cur], d_xkey, d_label,
                                                            d_alive, d_stamp, d_ngraph, d_pnix[0] + na,
                                                            d_pclu[0] + na, d_alist[acur] + nalist, &d_cnt->err);
            if (Q.arcs) k_rebase<<<grid_for(Q.arcs, BS), BS, 0, st>>>(Q.arcs, d_nbr + Q.abase, Q.base);
            k_admit_cluster<<<1, 64, 0, st>>>(nc, na, Q.n, Q.slot, Q.base, Q.arcs, iter1, ct[ctc].cols(), d_gsmall,
                                              d_gsmall_next, d_gsurv, d_giter, d_gseen, d_gstart);
            HIPC(hipGetLastError());
            na += Q.n;
            nalist += Q.n;
            nc += 1;
        }
        pending.clear();
        stage_used = 0;
        heavy_valid = false;
        rmap_valid = false;
        sweep_done = false;
        return 0;
    }

    // one iteration over every live graph; graphs that finished move to `done`
    int serve_step(uint32_t* finished) {
        if (!serving) return set_err(PXD_ERR_ARG, "not a serving engine");
        uint32_t nf = 0;
        CHK(join_pending());
        if (na > 0) {
            const double avg = live_nodes ? (double)live_arcs / (double)live_nodes : 0.0;
            avg_deg = avg;
            gs = avg <= 24.0 ? 4 : avg <= 48.0 ? 8 : avg <= 96.0 ? 16 : avg <= 192.0 ? 32 : 64;
            if (heavy_thr != 32u * gs) {
                heavy_thr = 32u * gs;
                heavy_valid = false;
                rmap_valid = false;
            }
            CHK(walk());
            CHK(sweep());
            CHK(cut());
            if (!giter_read) {
                HIPC(hipMemcpyAsync(h_giter.data(), d_giter, G * 4, hipMemcpyDeviceToHost, st));
                CHK(sync());
            }
            for (uint32_t g = 0; g < G; ++g) {
                Slot& S = slots[g];
                if (!S.used || h_giter[g] == 0) continue;
                Done D{S.tag, 0, h_giter[g], std::vector<int32_t>(S.n), false, S.n};
                HIPC(hipMemcpy(D.lab.data(), d_label + S.base, (size_t)S.n * 4, hipMemcpyDeviceToHost));
                push_done(std::move(D));
                give_range(free_nodes, S.base, S.n);
                give_range(free_arcs, S.abase, S.arcs);
                S.used = false;
                free_slots.push_back(g);
                live_graphs--;
                live_nodes -= S.n;
                live_arcs -= S.arcs;
                ++nf;
            }
        }
        if (finished) *finished = nf;
        return PXD_OK;
    }

    // One graph of at most SMALL_MAX_N nodes, straight after initialize or
    // reset, with the per-arc kernels left to the engine's choice and no
    // per-launch walk timing: the whole run is one workgroup's launch.
    bool small_run_applies() const {
        return small_kernel != 0 && !serving && G == 1 && N > 0 && N <= SMALL_MAX_N && A <= SMALL_MAX_ARCS &&
               (small_kernel == 2 || A <= SMALL_AUTO_ARCS) &&
               walk_force < 0 && !walk_timing && iter1 == 0 && cur == 0 && acur == 0 && ctc == 0 &&
               na == N && nalist == N && nc == 1 && !sweep_done;
    }
    int run_small(uint32_t* loops) {
        SmallArgs sa{};
        sa.n = N;
        sa.maxit = P.max_iter;
        sa.na = na;
        sa.nc = nc;
        sa.nalist = nalist;
        sa.sc_thr = (double)P.sc_threshold;
        sa.rw_thr = P.rw_threshold;
        sa.row = d_row;
        sa.dint = d_dint;
        sa.nbr = d_nbr;
        sa.xkey = d_xkey;
        sa.stamp = d_stamp;
        sa.root = d_root;
        sa.x[0] = d_x[0];
        sa.x[1] = d_x[1];
        sa.y[0] = d_y[0];
        sa.y[1] = d_y[1];
        sa.label = d_label;
        sa.alive = d_alive;
        for (int b = 0; b < 2; ++b) {
            sa.pnix[b] = d_pnix[b];
            sa.pclu[b] = d_pclu[b];
            sa.alist[b] = d_alist[b];
            sa.ct[b] = ct[b].cols();
        }
        sa.pkey = d_pkey[0];
        sa.cclist = d_cclist;
        sa.nclu = d_nclu;
        sa.flags = d_tB;
        sa.edv = d_ediff;
        sa.coff = d_coff;
        sa.ckeep = d_ckeep;
        sa.kscan = d_kscan;
        sa.cmaxo = d_cmaxo;
        sa.cmino = d_cmino;
        sa.csp = d_csp;
        sa.cavg = d_cavg;
        sa.cK = d_cK;
        sa.gsmall = d_gsmall;
        sa.giter = d_giter;
        sa.nd = P.record_node_data ? d_nd : nullptr;
        sa.cnt = d_cnt;
        sa.out = d_sout;
        k_decompose_small<<<1, SM_NT, 0, st>>>(sa);
        HIPC(hipGetLastError());
        CHK(pinned());
        static_assert(sizeof(Counters) + 64 <= PIN_BYTES, "pinned words of the run's counts");
        uint32_t* o = h_pin + (PIN_BYTES - 64) / 4;
        HIPC(hipMemcpyAsync(o, d_sout, 8 * 4, hipMemcpyDeviceToHost, st));
        CHK(read_counters());         // syncs; reports a walk value out of range
        na = o[0];
        nc = o[1];
        nalist = o[2];
        cur = (int)o[3];
        iter1 = o[4];
        *loops = o[5];
        if (o[6]) {                   // the current positions in buffer 0, as after a cut
            std::swap(d_pnix[0], d_pnix[1]);
            std::swap(d_pclu[0], d_pclu[1]);
        }
        ctc = (int)(o[7] >> 1);
        acur = (int)(o[7] & 1u);
        heavy_valid = false;
        rmap_valid = false;
        sweep_done = false;
        return 0;
    }

    int run(int32_t* iterations) {
        if (!inited) return set_err(PXD_ERR_ARG, "engine not initialized");
        if (serving) return set_err(PXD_ERR_ARG, "a serving engine advances with pxd_serve_step");
        HIPC(hipEventRecord(ev0, st));
        const uint32_t maxit = P.max_iter;
        uint32_t i = 0;
        if (small_run_applies()) {
            CHK(run_small(&i));
        } else {
            while (i++ < maxit && nc > 0 && na > 0) {
                CHK(walk());
                CHK(sweep());
                if (na == 0) break;
                CHK(cut());
            }
        }
        HIPC(hipEventRecord(ev1, st));
        HIPC(hipEventSynchronize(ev1));
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, ev0, ev1));
        last_run_ms = ms;
        // graphs still unfinished (max_iter) or never started report the final counter
        std::vector<int32_t> gi(G);
        if (G) HIPC(hipMemcpy(gi.data(), d_giter, G * 4, hipMemcpyDeviceToHost));
        int32_t mx = 0;
        for (uint32_t g = 0; g < G; ++g) {
            if (gi[g] == 0) gi[g] = (int32_t)i;
            mx = std::max(mx, gi[g]);
        }
        if (G) HIPC(hipMemcpy(d_giter, gi.data(), G * 4, hipMemcpyHostToDevice));
        if (iterations) *iterations = G ? mx : (int32_t)i;
        return 0;
    }
};

__global__ void k_set_all(uint32_t n, const double* __restrict__ v, const uint32_t* __restrict__ dint,
                          double* __restrict__ x, uint64_t* __restrict__ y, uint32_t* __restrict__ xkey,
                          uint32_t* err) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        const double a = v[u];
        x[u] = a;
        put_y(y, u, a, dint[u], err);
        xkey[u] = ord_f32((float)a);
    }
}


// =================================================================== allocation-free CSR build

// The work area of pxd_csr_from_edges_noalloc: a static device array, so the
// call allocates nothing.  Every GPU the library runs on holds its own copy
// (about 25 MB) until the process ends.  Host code never passes `scratch`
// itself (on the host it only names the symbol): it uses the device address
// that hipGetSymbolAddress gives for the current device.
constexpr uint32_t NOALLOC_MAX_NODES = PXD_NOALLOC_MAX_NODES;
constexpr uint64_t NOALLOC_MAX_EDGES = PXD_NOALLOC_MAX_EDGES;
constexpr uint64_t NOALLOC_WORDS = ingest_work_words(NOALLOC_MAX_NODES, 2 * NOALLOC_MAX_EDGES);
static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

static std::mutex g_scratch_mu;
static uint32_t* g_scratch_dev[64];   // device address of `scratch`, per device ordinal

// the current device's copy of `scratch`, looked up once per device
static int scratch_on_current_device(uint32_t** out) {
    int dev = -1;
    HIPC(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return set_err(PXD_ERR_ARG, "device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_err(PXD_ERR_HIP, "hipGetSymbolAddress returned no address");
        g_scratch_dev[dev] = (uint32_t*)p;
    }
    *out = g_scratch_dev[dev];
    return 0;
}

// =================================================================== C ABI

extern "C" {

void pxd_default_params(pxd_params* p) {
    if (!p) return;
    p->seed = 6582u;
    p->sc_threshold = 0.3f;
    p->rw_threshold = 1e-4f;
    p->max_iter = 200000u;
    p->keep_original = 1u;
    p->record_node_data = 1u;
    p->device = 0;
}

const char* pxd_last_error(void) { return g_last_error.c_str(); }

int pxd_create(const pxd_params* params, pxd_engine** out) {
    return guarded([&]() -> int {
        if (!out) return set_err(PXD_ERR_ARG, "out is NULL");
        *out = nullptr;
        pxd_engine* e = new (std::nothrow) pxd_engine();
        if (!e) return set_err(PXD_ERR_OOM, "host allocation failed");
        if (params) e->P = *params; else pxd_default_params(&e->P);
        e->dev = e->P.device;
        int ndev = 0;
        hipError_t err = hipGetDeviceCount(&ndev);
        if (err != hipSuccess || ndev <= e->dev) {
            delete e;
            return set_err(PXD_ERR_HIP, "no HIP device %d available (%s)", params ? params->device : 0,
                           hipGetErrorString(err));
        }
        if (hipSetDevice(e->dev) != hipSuccess || hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess ||
            hipEventCreate(&e->tev0) != hipSuccess || hipEventCreate(&e->tev1) != hipSuccess) {
            delete e;
            return set_err(PXD_ERR_HIP, "stream/event creation failed");
        }
        *out = e;
        return 0;
    });
}

int pxd_destroy(pxd_engine* e) {
    return guarded([&]() -> int {
        if (e) {
            (void)hipSetDevice(e->dev);
            delete e;
        }
        return 0;
    });
}

int pxd_initialize(pxd_engine* e, uint32_t n, uint32_t m, const uint32_t* ranges, const uint32_t* edges) {
    return guarded([&]() -> int {
        if (!e || !ranges || (m && !edges)) return set_err(PXD_ERR_ARG, "null argument");
        const uint32_t* r[1] = {ranges};
        const uint32_t* ed[1] = {edges};
        const int rc = e->setup(1, &n, &m, r, ed);
        e->trim_pool();   // buffers of the previous input that this one did not reuse
        return rc;
    });
}

int pxd_initialize_edges(pxd_engine* e, uint32_t n, uint64_t m, const uint32_t* fx, const uint32_t* tx) {
    return guarded([&]() -> int {
        if (!e) return set_err(PXD_ERR_ARG, "null engine");
        const int rc = e->setup_edges(n, m, fx, tx);
        e->trim_pool();
        return rc;
    });
}

int pxd_initialize_batch(pxd_engine* e, uint32_t num_graphs, const uint32_t* node_counts, const uint32_t* edge_counts,
                         const uint32_t* const* ranges_ptrs, const uint32_t* const* edges_ptrs) {
    return guarded([&]() -> int {
        if (!e || !node_counts || !edge_counts || !ranges_ptrs || !edges_ptrs || num_graphs == 0)
            return set_err(PXD_ERR_ARG, "null argument or empty batch");
        const int rc = e->setup(num_graphs, node_counts, edge_counts, ranges_ptrs, edges_ptrs);
        e->trim_pool();
        return rc;
    });
}

int pxd_set_thresholds(pxd_engine* e, float sc_threshold, float rw_threshold) {
    return guarded([&]() -> int {
        if (!e) return set_err(PXD_ERR_ARG, "null engine");
        e->P.sc_threshold = sc_threshold;
        e->P.rw_threshold = rw_threshold;
        return 0;
    });
}

int pxd_reset(pxd_engine* e) {
    return guarded([&]() -> int { return e ? e->reset() : set_err(PXD_ERR_ARG, "null engine"); });
}

int pxd_iterate_random_walk(pxd_engine* e) {
    return guarded([&]() -> int { return e ? e->walk() : set_err(PXD_ERR_ARG, "null engine"); });
}
int pxd_compute_sweep_cuts(pxd_engine* e) {
    return guarded([&]() -> int { return e ? e->sweep() : set_err(PXD_ERR_ARG, "null engine"); });
}
int pxd_cut_clusters(pxd_engine* e) {
    return guarded([&]() -> int { return e ? e->cut() : set_err(PXD_ERR_ARG, "null engine"); });
}

int pxd_expander_decomposition(pxd_engine* e, int32_t* iterations) {
    return guarded([&]() -> int {
        return e ? e->run(iterations) : set_err(PXD_ERR_ARG, "null engine");
    });
}

int pxd_graph_iterations(pxd_engine* e, int32_t* out, uint32_t num_graphs) {
    return guarded([&]() -> int {
        if (!e || !e->inited || !out || num_graphs < e->G) return set_err(PXD_ERR_ARG, "bad arguments");
        HIPC(hipSetDevice(e->dev));
        HIPC(hipMemcpy(out, e->d_giter, e->G * 4, hipMemcpyDeviceToHost));
        return 0;
    });
}

static int require(pxd_engine* e) {
    if (!e || !e->inited) return set_err(PXD_ERR_ARG, "engine not initialized");
    HIPC(hipSetDevice(e->dev));
    HIPC(hipStreamSynchronize(e->st));
    return 0;
}

int pxd_get_final_partition(pxd_engine* e, int32_t* ids_out, int32_t* num_clusters) {
    return guarded([&]() -> int {
        CHK(require(e));
        if ((e->N && !ids_out) || !num_clusters) return set_err(PXD_ERR_ARG, "null output");
        if (e->serving) return set_err(PXD_ERR_ARG, "a serving engine returns results with pxd_serve_take");
        std::vector<int32_t> lab(e->N);
        if (e->N) HIPC(hipMemcpy(lab.data(), e->d_label, e->N * 4, hipMemcpyDeviceToHost));
        // cuda/devPartition.h:798-823 per graph: sort, unique, rank
        for (uint32_t g = 0; g < e->G; ++g) {
            const uint32_t b = e->h_gnode[g], en = e->h_gnode[g + 1];
            num_clusters[g] = dense_rank(lab.data() + b, en - b, ids_out + b);
        }
        return 0;
    });
}

int pxd_read_random_walk_values(pxd_engine* e, float* out) {
    return guarded([&]() -> int {
        CHK(require(e));
        std::vector<double> v(e->N);
        if (e->N) HIPC(hipMemcpy(v.data(), e->d_x[e->cur], e->N * 8, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < e->N; ++i) out[i] = (float)v[i];
        return 0;
    });
}

int pxd_read_random_walk_values_f64(pxd_engine* e, double* out) {
    return guarded([&]() -> int {
        CHK(require(e));
        if (e->N) HIPC(hipMemcpy(out, e->d_x[e->cur], e->N * 8, hipMemcpyDeviceToHost));
        return 0;
    });
}

int pxd_set_random_walk_values_f64(pxd_engine* e, const double* values) {
    return guarded([&]() -> int {
        CHK(require(e));
        if (!e->N) return 0;
        double* tmp = nullptr;
        HIPC(hipMalloc(&tmp, (size_t)e->N * 8));
        HIPC(hipMemcpy(tmp, values, (size_t)e->N * 8, hipMemcpyHostToDevice));
        k_set_all<<<grid_for(e->N, BS), BS, 0, e->st>>>(e->N, tmp, e->d_dint, e->d_x[e->cur], e->d_y[e->cur], e->d_xkey,
                                                         &e->d_cnt->err);
        HIPC(hipGetLastError());
        HIPC(hipStreamSynchronize(e->st));
        HIPC(hipFree(tmp));
        e->sweep_done = false;            // the sweep order belongs to the old values
        return 0;
    });
}

int pxd_download_degrees(pxd_engine* e, uint32_t* out) {
    return guarded([&]() -> int {
        CHK(require(e));
        if (e->N) HIPC(hipMemcpy(out, e->d_dint, e->N * 4, hipMemcpyDeviceToHost));
        return 0;
    });
}

int pxd_download_labels(pxd_engine* e, int32_t* out) {
    return guarded([&]() -> int {
        CHK(require(e));
        if (e->N) HIPC(hipMemcpy(out, e->d_label, e->N * 4, hipMemcpyDeviceToHost));
        return 0;
    });
}

// rebuild the reference's in-place INVALID_EDGE view: the valid prefix of a
// row is an order-preserving subsequence of the original row
static int edge_view(pxd_engine* e, std::vector<uint32_t>& row, std::vector<uint32_t>& orig,
                     std::vector<uint8_t>& valid) {
    if (!e->P.keep_original) return set_err(PXD_ERR_ARG, "engine created with keep_original = 0");
    row.resize(e->N + 1);
    orig.resize(e->A);
    std::vector<uint32_t> cur(e->A), dint(e->N);
    HIPC(hipMemcpy(row.data(), e->d_row, (e->N + 1) * 4, hipMemcpyDeviceToHost));
    if (e->A) {
        HIPC(hipMemcpy(orig.data(), e->d_nbr0, e->A * 4, hipMemcpyDeviceToHost));
        HIPC(hipMemcpy(cur.data(), e->d_nbr, e->A * 4, hipMemcpyDeviceToHost));
    }
    if (e->N) HIPC(hipMemcpy(dint.data(), e->d_dint, e->N * 4, hipMemcpyDeviceToHost));
    valid.assign(e->A, 0);
    for (uint32_t u = 0; u < e->N; ++u) {
        uint32_t p = row[u];
        const uint32_t pe = row[u] + dint[u];
        for (uint32_t j = row[u]; j < row[u + 1] && p < pe; ++j)
            if (orig[j] == cur[p]) { valid[j] = 1; ++p; }
        if (p != pe) return set_err(PXD_ERR_STATE, "valid prefix of node %u is not a subsequence of its row", u);
    }
    return 0;
}

int pxd_download_graph(pxd_engine* e, uint32_t* ranges_out, uint32_t* edges_out) {
    return guarded([&]() -> int {
        CHK(require(e));
        std::vector<uint32_t> row, orig;
        std::vector<uint8_t> valid;
        CHK(edge_view(e, row, orig, valid));
        for (uint32_t g = 0; g < e->G; ++g) {
            const uint32_t nb = e->h_gnode[g], ab = e->h_garc[g];
            for (uint32_t u = nb; u <= e->h_gnode[g + 1]; ++u) ranges_out[u] = row[u] - ab;
            for (uint32_t j = ab; j < e->h_garc[g + 1]; ++j) edges_out[j] = valid[j] ? orig[j] - nb : PXD_INVALID_EDGE;
        }
        return 0;
    });
}

int pxd_download_active_edge_map(pxd_engine* e, int32_t* out) {
    return guarded([&]() -> int {
        CHK(require(e));
        std::vector<uint32_t> row, orig;
        std::vector<uint8_t> valid;
        CHK(edge_view(e, row, orig, valid));
        for (uint32_t j = 0; j < e->A; ++j) out[j] = valid[j];
        return 0;
    });
}

int pxd_download_partition(pxd_engine* e, pxd_node_data* out, uint32_t capacity, uint32_t* count) {
    return guarded([&]() -> int {
        CHK(require(e));
        if (!e->P.record_node_data) return set_err(PXD_ERR_ARG, "engine created with record_node_data = 0");
        const uint32_t c = std::min(capacity, e->sweep_done ? e->na : 0u);
        if (c) HIPC(hipMemcpy(out, e->d_nd, c * sizeof(pxd_node_data), hipMemcpyDeviceToHost));
        if (count) *count = c;
        return 0;
    });
}

int pxd_read_sweep_cuts(pxd_engine* e, int32_t* cluster_ids, pxd_sweep_cut* cuts, uint32_t capacity, uint32_t* count) {
    return guarded([&]() -> int {
        CHK(require(e));
        const uint32_t c = std::min(capacity, e->nc);
        ClusterTable& T = e->ct[e->ctc];
        std::vector<int32_t> lab(c);
        std::vector<float> sp(c);
        std::vector<uint32_t> off(c);
        if (c) {
            HIPC(hipMemcpy(lab.data(), T.clabel, c * 4, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(sp.data(), e->d_csp, c * 4, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(off.data(), e->d_coff, c * 4, hipMemcpyDeviceToHost));
        }
        for (uint32_t i = 0; i < c; ++i) {
            if (cluster_ids) cluster_ids[i] = lab[i];
            if (cuts) { cuts[i].clusterId = lab[i]; cuts[i].sparsity = sp[i]; cuts[i].offset = off[i]; }
        }
        if (count) *count = c;
        return 0;
    });
}

int pxd_get_num_cut_edges(pxd_engine* e, uint32_t* out) {
    return guarded([&]() -> int {
        CHK(require(e));
        HIPC(hipMemsetAsync(e->d_u64, 0, 8, e->st));
        if (e->N) k_sum_u32<<<grid_for(e->N, BS), BS, 0, e->st>>>(e->N, e->d_dint, e->d_u64);
        HIPC(hipGetLastError());
        unsigned long long s = 0;
        HIPC(hipMemcpyAsync(&s, e->d_u64, 8, hipMemcpyDeviceToHost, e->st));
        HIPC(hipStreamSynchronize(e->st));
        *out = (uint32_t)(e->M_total - s / 2);
        return 0;
    });
}

int pxd_download_active_nodes(pxd_engine* e, pxd_labeled_node* out, uint32_t capacity, uint32_t* count) {
    return guarded([&]() -> int {
        CHK(require(e));
        const uint32_t c = std::min(capacity, e->na);
        std::vector<uint32_t> nix(c), clu(c);
        std::vector<int32_t> lab(e->nc);
        if (c) {
            HIPC(hipMemcpy(nix.data(), e->d_pnix[0], c * 4, hipMemcpyDeviceToHost));
            HIPC(hipMemcpy(clu.data(), e->d_pclu[0], c * 4, hipMemcpyDeviceToHost));
        }
        if (e->nc) HIPC(hipMemcpy(lab.data(), e->ct[e->ctc].clabel, e->nc * 4, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < c; ++i) {
            out[i].nix = nix[i] - e->h_gnode[0];
            out[i].clusterId = clu[i] < e->nc ? lab[clu[i]] : -1;
        }
        if (count) *count = c;
        return 0;
    });
}

int pxd_get_counts(pxd_engine* e, uint32_t* num_active_nodes, uint32_t* num_active_clusters,
                   uint32_t* num_clusters_with_cut) {
    return guarded([&]() -> int {
        if (!e) return set_err(PXD_ERR_ARG, "null engine");
        if (num_active_nodes) *num_active_nodes = e->na;
        if (num_active_clusters) *num_active_clusters = e->nc;
        if (num_clusters_with_cut) *num_clusters_with_cut = e->ncut;
        return 0;
    });
}

int pxd_last_run_ms(pxd_engine* e, double* ms) {
    return guarded([&]() -> int {
        if (!e || !ms) return set_err(PXD_ERR_ARG, "null argument");
        *ms = e->last_run_ms;
        return 0;
    });
}

int pxd_enable_walk_timing(pxd_engine* e, int enable) {
    return guarded([&]() -> int {
        if (!e) return set_err(PXD_ERR_ARG, "null engine");
        e->walk_timing = enable != 0;
        e->walk_launches = 0;
        e->walk_ms = 0.0;
        e->walk_bytes = 0.0;
        e->walk_log_ms.clear();
        e->walk_log_bytes.clear();
        return 0;
    });
}

int pxd_walk_launch_log(pxd_engine* e, double* ms, double* alg_bytes, uint32_t capacity, uint32_t* count) {
    return guarded([&]() -> int {
        if (!e) return set_err(PXD_ERR_ARG, "null engine");
        const uint32_t n = (uint32_t)e->walk_log_ms.size();
        if (count) *count = n;
        if (capacity < n && (ms || alg_bytes)) return set_err(PXD_ERR_ARG, "capacity %u < %u launches", capacity, n);
        for (uint32_t i = 0; i < n; ++i) {
            if (ms) ms[i] = e->walk_log_ms[i];
            if (alg_bytes) alg_bytes[i] = e->walk_log_bytes[i];
        }
        return 0;
    });
}

int pxd_walk_stats(pxd_engine* e, uint64_t* launches, double* total_ms, double* alg_bytes) {
    return guarded([&]() -> int {
        if (!e) return set_err(PXD_ERR_ARG, "null engine");
        if (launches) *launches = e->walk_launches;
        if (total_ms) *total_ms = e->walk_ms;
        if (alg_bytes) *alg_bytes = e->walk_bytes;
        return 0;
    });
}

int pxd_set_small_graph_kernel(pxd_engine* e, int enable) {
    return guarded([&]() -> int {
        if (!e) return set_err(PXD_ERR_ARG, "null engine");
        if (enable < 0 || enable > 2) return set_err(PXD_ERR_ARG, "small-graph mode %d is not 0, 1 or 2", enable);
        e->small_kernel = enable;
        return 0;
    });
}

int pxd_set_walk_variant(pxd_engine* e, int variant) {
    return guarded([&]() -> int {
        if (!e) return set_err(PXD_ERR_ARG, "null engine");
        if (variant < -2 || variant >= pxd_engine::WV_COUNT) return set_err(PXD_ERR_ARG, "walk variant %d out of range", variant);
        e->walk_force = variant;
        return 0;
    });
}

int pxd_walk_tuning(pxd_engine* e, int32_t* table, uint32_t capacity, uint64_t* tunings) {
    return guarded([&]() -> int {
        if (!e) return set_err(PXD_ERR_ARG, "null engine");
        // per class, the choice for its longest lists tuned so far (the
        // highest quarter that has one)
        for (uint32_t b = 0; b < capacity && b < 66; ++b) {
            int v = -1;
            for (int q = pxd_engine::TUNE_Q - 1; q >= 0 && v < 0; --q) v = e->tuned[b / 33][b % 33][q];
            table[b] = v;
        }
        if (tunings) *tunings = e->walk_tunings;
        return 0;
    });
}

int pxd_serve_create(const pxd_params* params, uint32_t node_capacity, uint64_t arc_capacity, uint32_t graph_slots,
                     pxd_engine** out) {
    return guarded([&]() -> int {
        CHK(pxd_create(params, out));
        const int rc = (*out)->serve_create(node_capacity, arc_capacity, graph_slots);
        if (rc != 0) {
            delete *out;
            *out = nullptr;
        }
        return rc;
    });
}

int pxd_serve_admit(pxd_engine* e, uint32_t n, uint32_t m, const uint32_t* ranges, const uint32_t* edges,
                    uint64_t tag, int32_t* slot) {
    return guarded([&]() -> int {
        return e ? e->serve_admit(n, m, ranges, edges, tag, slot) : set_err(PXD_ERR_ARG, "null engine");
    });
}

int pxd_serve_enable_edges(pxd_engine* e, uint32_t max_nodes, uint64_t max_edges) {
    return guarded([&]() -> int {
        return e ? e->serve_enable_edges(max_nodes, max_edges) : set_err(PXD_ERR_ARG, "null engine");
    });
}

int pxd_serve_admit_edges(pxd_engine* e, uint32_t n, uint64_t m, const uint32_t* fx, const uint32_t* tx, uint64_t tag,
                          int32_t* slot) {
    return guarded([&]() -> int {
        return e ? e->serve_admit_edges(n, m, fx, tx, tag, slot) : set_err(PXD_ERR_ARG, "null engine");
    });
}

int pxd_serve_step(pxd_engine* e, uint32_t* finished) {
    return guarded([&]() -> int {
        return e ? e->serve_step(finished) : set_err(PXD_ERR_ARG, "null engine");
    });
}

int pxd_serve_next(pxd_engine* e, uint64_t* tag, uint32_t* n) {
    return guarded([&]() -> int {
        if (!e || !tag || !n) return set_err(PXD_ERR_ARG, "null argument");
        std::lock_guard<std::mutex> lk(e->done_mu);
        if (e->done.empty() || !e->done.front().ranked) return PXD_EMPTY;
        *tag = e->done.front().tag;
        *n = e->done.front().n;
        return PXD_OK;
    });
}

int pxd_serve_wait(pxd_engine* e) {
    return guarded([&]() -> int {
        if (!e) return set_err(PXD_ERR_ARG, "null engine");
        std::unique_lock<std::mutex> lk(e->done_mu);
        e->done_cv.wait(lk, [e] {
            for (const pxd_engine::Done& d : e->done)
                if (!d.ranked) return false;
            return true;
        });
        return PXD_OK;
    });
}

int pxd_serve_take(pxd_engine* e, int32_t* ids_out, int32_t* num_clusters, int32_t* iterations) {
    return guarded([&]() -> int {
        if (!e || !num_clusters || !iterations) return set_err(PXD_ERR_ARG, "null argument");
        std::unique_lock<std::mutex> lk(e->done_mu);
        if (e->done.empty()) return PXD_EMPTY;
        e->done_cv.wait(lk, [e] { return e->done.front().ranked; });
        pxd_engine::Done& D = e->done.front();
        if (D.n) {
            if (!ids_out) return set_err(PXD_ERR_ARG, "null output");
            std::memcpy(ids_out, D.lab.data(), (size_t)D.n * 4);
        }
        *num_clusters = D.k;
        *iterations = D.it;
        e->done.pop_front();
        return PXD_OK;
    });
}

int pxd_serve_stats(pxd_engine* e, uint32_t* live_graphs, uint64_t* live_nodes, uint64_t* live_arcs,
                    uint32_t* active_nodes, uint32_t* waiting_results) {
    return guarded([&]() -> int {
        if (!e) return set_err(PXD_ERR_ARG, "null engine");
        if (live_graphs) *live_graphs = e->live_graphs;
        if (live_nodes) *live_nodes = e->live_nodes;
        if (live_arcs) *live_arcs = e->live_arcs;
        if (active_nodes) *active_nodes = e->na;
        if (waiting_results) {
            std::lock_guard<std::mutex> lk(e->done_mu);
            *waiting_results = (uint32_t)e->done.size();
        }
        return PXD_OK;
    });
}

int pxd_csr_from_edges_noalloc(uint32_t n, uint64_t m, const uint32_t* d_fx, const uint32_t* d_tx,
                               uint32_t* d_ranges, uint32_t* d_edges, uint32_t* d_bad, void* stream) {
    return guarded([&]() -> int {
        if (n > NOALLOC_MAX_NODES || m > NOALLOC_MAX_EDGES)
            return set_err(PXD_ERR_ARG,
                           "pxd_csr_from_edges_noalloc takes at most %u nodes and %llu edges (%u nodes, %llu edges "
                           "given)",
                           NOALLOC_MAX_NODES, (unsigned long long)NOALLOC_MAX_EDGES, n, (unsigned long long)m);
        if (!d_ranges || !d_bad || (m && (!d_fx || !d_tx || !d_edges))) return set_err(PXD_ERR_ARG, "null argument");
        uint32_t* base = nullptr;
        CHK(scratch_on_current_device(&base));
        const hipStream_t st = (hipStream_t)stream;
        // every word the build reads is written first (counters and degrees by
        // fill kernels, not memset): the last call's contents do not matter
        k_fill_u32<<<1, 64, 0, st>>>(1, d_bad, 0u);
        HIPC(hipGetLastError());
        HIPC(ingest_csr_async(n, m, d_fx, d_tx, d_ranges, d_edges, ingest_carve(base, n, 2 * m), d_bad, st));
        return 0;
    });
}

}  // extern "C"

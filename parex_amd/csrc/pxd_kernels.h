// HIP kernels of the expander-decomposition engine (gfx950, wave64).
//
// State layout in HBM (one engine may hold a block-diagonal batch of graphs):
//   per node u      row[u] (u32 arc offset), deg[u], dint[u] (valid-prefix length),
//                   nbr[row[u] .. row[u]+dint[u]) = the still-valid arcs of u, kept
//                   as an order-preserving prefix of the original row;
//                   x[2][u] (FP64 walk double buffer), y[u] = the walk term of
//                   x[u]/dint[u] packed for exact summation (fix_pack),
//                   xkey[u] (ordered float of the last walk value), label[u]
//   per position k  pnix[k] (node), pkey[k] (sort key), pclu[k] (cluster id);
//                   clusters are contiguous position ranges [cs[c], cs[c+1])
//   per cluster c   cs, cgraph, cvol, cmin (min node index), clabel, cflag, ...
//   node list       alist = active nodes in ascending index order (walk and
//                   cut-contribution passes stream rows in memory order)
#pragma once

#include "pxd_device.h"
#include "parex_amd/pxd.h"

namespace pxd {

constexpr int BS = 256;                 // default block size (4 waves)
constexpr int WBS = 1024;               // LDS-window kernels: one 16-wave workgroup per CU
constexpr int TILE = 2048;              // LDS sort tile: 256 threads x 8 items
constexpr int TILE_ITEMS = 8;
constexpr int SCAN_ITEMS = 4;           // scans: 1024 elements per block

constexpr uint32_t CF_CONN = 1u;        // cluster known to be connected
constexpr uint32_t CF_CUT = 2u;         // sweep cut applies this iteration
constexpr uint32_t CF_CC = 4u;          // needs connected components this iteration

struct Counters {
    uint32_t err;
    uint32_t nwave, nblock, nlarge;
    uint32_t ncut, ncc;
    uint32_t cclist_n;
    uint32_t link_changed;
    uint32_t admit_bad;
    uint32_t maxdeg;          // largest degree of an uploaded batch
    // counts that stay on the device within an iteration (their kernels read
    // them; the host sizes grids from the iteration's upper bounds)
    uint32_t nl, nu;          // CC members, members outside the giant trees
    uint32_t hchunks;         // chunks of the current long-row map
    uint32_t tiles, tiles256; // tiles of the large-segment sort plan (and x256)
    uint32_t cc_done;         // the exact CC rounds have converged
    uint32_t ccn;             // positions of the clusters that need CC (= CC members)
    uint32_t ncr;             // clusters after the CC regroup
    uint32_t na_new, nc_new;  // positions and clusters kept by the cut, read once per iteration
};

// n from the device when a count pointer is given, else the value passed
__device__ __forceinline__ uint32_t dev_count(uint32_t n, const uint32_t* p) { return p ? *p : n; }

// y[u] = the packed walk term of x / d (0 for d = 0); a value outside the
// exact fixed-point range sets the error bit of the engine's status word
__device__ __forceinline__ void put_y(uint64_t* __restrict__ y, uint32_t u, double x, uint32_t d, uint32_t* err) {
    uint32_t ovf = 0;
    y[u] = fix_pack(d ? x / (double)d : 0.0, ovf);
    if (ovf) atomicOr(err, ERR_FIX_RANGE);
}

// ------------------------------------------------------------------ init

__device__ float init_value(uint32_t idx, uint32_t deg, uint32_t seed) {
    // cuda/devRandomWalk.h:22-47: hash -> minstd_rand -> normal_distribution<float>
    uint32_t x = seed ^ idx;
    x = ((x >> 16) ^ x) * 0x45d9f3bu;
    x = ((x >> 16) ^ x) * 0x45d9f3bu;
    x = (x >> 16) ^ x;
    const uint64_t M = 2147483647u;
    uint64_t s = x % M;
    if (s == 0) s = 1;
    s = (s * 48271u) % M;
    const uint32_t range = 2147483645u;
    const float S1 = (float)(1.0 / (double)range);
    const float S2 = S1 / 2;
    float S3 = (float)(-1.4142135623730950488016887242097);
    uint32_t u = (uint32_t)s - 1u;
    if (u > range / 2) { u = range - u; S3 = -S3; }
    float p = fmaf((float)u, S1, S2);
    float e = erfcinvf(2.0f * p);
    float sigma = (float)sqrt((double)deg);
    return (sigma * S3) * e;
}

// columns of a cluster table (device pointers)
struct ClusterCols {
    uint32_t *cs, *cgraph, *cmin, *cflag, *cvol;
    int32_t* clabel;
};

__global__ void k_init_nodes(uint32_t N, uint32_t G, const uint32_t* __restrict__ gnode,
                             const uint32_t* __restrict__ row, uint32_t seed,
                             uint32_t* __restrict__ deg, uint32_t* __restrict__ dint,
                             double* __restrict__ x, uint64_t* __restrict__ y,
                             uint32_t* __restrict__ xkey, int32_t* __restrict__ label,
                             uint8_t* __restrict__ alive, uint32_t* __restrict__ stamp,
                             uint32_t* __restrict__ pnix, uint32_t* __restrict__ alist,
                             uint32_t* __restrict__ pclu, const uint32_t* __restrict__ gclu,
                             uint32_t* __restrict__ ngraph, uint32_t* err) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < N; u += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = G;                  // graph of u: last g with gnode[g] <= u
        while (hi - lo > 1) {
            uint32_t mid = (lo + hi) >> 1;
            if (gnode[mid] <= u) lo = mid; else hi = mid;
        }
        uint32_t d = row[u + 1] - row[u];
        deg[u] = d;
        dint[u] = d;
        double v = (double)init_value(u - gnode[lo], d, seed);
        x[u] = v;
        put_y(y, u, v, d, err);
        xkey[u] = ord_f32((float)v);
        label[u] = 0;
        alive[u] = 1;
        stamp[u] = 0;
        pnix[u] = u;
        alist[u] = u;
        pclu[u] = gclu[lo];
        ngraph[u] = lo;
    }
}

// One admitted graph (serving): nodes [base, base + n) with raw ranges rg
// (rg[0] = 0) and arcs at abase.  Same initial state as k_init_nodes gives a
// graph of a batch; its nodes are appended to the positions (cluster c, in
// node order) and to the active list.
__global__ void k_admit_nodes(uint32_t base, uint32_t n, uint32_t abase, uint32_t slot, uint32_t c,
                              const uint32_t* __restrict__ rg, uint32_t seed, uint32_t* __restrict__ row,
                              uint32_t* __restrict__ deg, uint32_t* __restrict__ dint, double* __restrict__ x,
                              uint64_t* __restrict__ y, uint32_t* __restrict__ xkey, int32_t* __restrict__ label,
                              uint8_t* __restrict__ alive, uint32_t* __restrict__ stamp,
                              uint32_t* __restrict__ ngraph, uint32_t* __restrict__ pnix,
                              uint32_t* __restrict__ pclu, uint32_t* __restrict__ alist, uint32_t* err) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t u = base + i;
        const uint32_t d = rg[i + 1] - rg[i];
        row[u] = abase + rg[i];
        deg[u] = d;
        dint[u] = d;
        const double v = (double)init_value(i, d, seed);
        x[u] = v;
        put_y(y, u, v, d, err);
        xkey[u] = ord_f32((float)v);
        label[u] = 0;
        alive[u] = 1;
        stamp[u] = 0;
        ngraph[u] = slot;
        pnix[i] = u;
        pclu[i] = c;
        alist[i] = u;
    }
}

// checks of an uploaded graph before it joins: ranges non-decreasing (rg[0] =
// 0 and rg[n] = 2m are checked on the host) and every arc below n
__global__ void k_admit_check(uint32_t n, const uint32_t* __restrict__ rg, uint64_t arcs,
                              const uint32_t* __restrict__ nbr, uint32_t* __restrict__ bad) {
    uint32_t b = 0u;
    const uint64_t t0 = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x, step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = t0; i < n; i += step) b |= rg[i + 1] < rg[i] ? 1u : 0u;
    for (uint64_t j = t0; j < arcs; j += step) b |= nbr[j] >= n ? 2u : 0u;
    b = group_or<WAVE>(b);
    if (b && lane_id() == 0) atomicOr(bad, b);
}

// arcs of an admitted graph: local -> global node ids
__global__ void k_rebase(uint64_t n, uint32_t* __restrict__ nbr, uint32_t base) {
    for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
        nbr[j] += base;
}

// graph of position i in a batch: the last g with off[g] <= i (off ascending,
// G + 1 entries)
__device__ __forceinline__ uint32_t batch_graph(const uint32_t* __restrict__ off, uint32_t G, uint32_t i) {
    uint32_t lo = 0, hi = G;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// Row offsets of a batch uploaded graph by graph as local ranges: u of graph
// g gets the graph's arc offset added
__global__ void k_rebase_rows(uint32_t N, uint32_t G, const uint32_t* __restrict__ gnode,
                              const uint32_t* __restrict__ garc, uint32_t* __restrict__ row) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < N; u += gridDim.x * blockDim.x)
        row[u] += garc[batch_graph(gnode, G, u)];
}

// Rebased offsets must not decrease (each graph's ranges run from 0 to its
// 2m, checked on the host, so this is every graph's check); the largest row
// length goes to cnt->maxdeg
__global__ void k_check_rows(uint32_t N, const uint32_t* __restrict__ row, Counters* cnt) {
    uint32_t b = 0u, mx = 0u;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < N; u += gridDim.x * blockDim.x) {
        const uint32_t a = row[u], e = row[u + 1];
        if (e < a) b = 1u;
        else mx = max(mx, e - a);
    }
    b = group_or<WAVE>(b);
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, off));
    if (lane_id() == 0) {
        if (b) atomicOr(&cnt->admit_bad, 1u);
        if (mx) atomicMax(&cnt->maxdeg, mx);
    }
}

// Arcs of a batch uploaded as they are: every arc checked below its graph's
// n, then local -> global node ids.  A workgroup takes REBASE_CHUNK
// consecutive arcs (coalesced, each thread every BS-th arc); the graph of its
// first arc is searched once, and each thread moves forward through the
// graphs as its arcs do.
constexpr uint32_t REBASE_CHUNK = BS * 16;
__global__ void __launch_bounds__(BS) k_rebase_arcs(uint32_t A, uint32_t G, const uint32_t* __restrict__ gnode,
                                                    const uint32_t* __restrict__ garc, uint32_t* __restrict__ nbr,
                                                    uint32_t* __restrict__ bad) {
    uint32_t b = 0u;
    for (uint32_t c0 = blockIdx.x * REBASE_CHUNK; c0 < A; c0 += gridDim.x * REBASE_CHUNK) {
        uint32_t g = batch_graph(garc, G, c0);
        const uint32_t c1 = min(A, c0 + REBASE_CHUNK);
        for (uint32_t j = c0 + threadIdx.x; j < c1; j += BS) {
            while (j >= garc[g + 1]) ++g;
            const uint32_t v = nbr[j];
            const uint32_t n = gnode[g + 1] - gnode[g];
            b |= v >= n ? 1u : 0u;
            nbr[j] = v + gnode[g];
        }
    }
    b = group_or<WAVE>(b);
    if (b && lane_id() == 0) atomicOr(bad, 1u);
}

// the admitted graph's single cluster, appended as cluster c at positions
// [na, na + n), and its per-graph state
__global__ void k_admit_cluster(uint32_t c, uint32_t na, uint32_t n, uint32_t slot, uint32_t base, uint32_t vol,
                                uint32_t iter, ClusterCols T, int32_t* gsmall, int32_t* gsmall_next,
                                uint32_t* gsurv, int32_t* giter, uint32_t* gseen, uint32_t* gstart) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        T.cs[c] = na;
        T.cs[c + 1] = na + n;
        T.cgraph[c] = slot;
        T.cmin[c] = base;
        T.cflag[c] = 0u;
        T.clabel[c] = 0;
        T.cvol[c] = vol;
        gsmall[slot] = 0;
        gsmall_next[slot] = 0;
        gsurv[slot] = 0u;
        giter[slot] = 0;
        gseen[slot] = 0u;
        gstart[slot] = iter;
    }
}

// ------------------------------------------------------------------ walk (K1)

// Lazy walk step for every active node, GS lanes per node, nodes visited in
// index order: x'[u] = 0.9 * sum_{v in valid(u)} x[v]/dInt[v] + 0.1 * x[u]
// (cuda/devRandomWalk.h:117-203).  The neighbour sum is exact (fixed point),
// so its bits do not depend on GS or on the lane that adds a term.
// x' = s * (1 - 0.1) + x * 0.1 with s the exact neighbour sum, y' = x'/dInt,
// walk key (cuda/devRandomWalk.h:186-187)
__device__ __forceinline__ void walk_finish(uint32_t u, uint32_t d, Fix128 acc, const double* __restrict__ xc,
                                            double* __restrict__ xn, uint64_t* __restrict__ yn,
                                            uint32_t* __restrict__ xkey, uint32_t& ovf) {
    if (d > (1u << 24)) ovf = 1;
    const double s = fix_to_f64(acc);
    const double v = s * (1.0 - RW_STAY) + xc[u] * RW_STAY;
    xn[u] = v;
    yn[u] = fix_pack(d ? v / (double)d : 0.0, ovf);
    xkey[u] = ord_f32((float)v);
}

template <int GS>
__global__ void __launch_bounds__(BS) k_walk(uint32_t na, const uint32_t* __restrict__ alist,
                                             const uint32_t* __restrict__ row,
                                             const uint32_t* __restrict__ dint,
                                             const uint32_t* __restrict__ nbr,
                                             const uint64_t* __restrict__ y,
                                             const double* __restrict__ xc, double* __restrict__ xn,
                                             uint64_t* __restrict__ yn, uint32_t* __restrict__ xkey,
                                             uint32_t heavy, Counters* cnt) {
    const uint32_t lane = threadIdx.x & (GS - 1);
    const XcdSpan sp = xcd_span(na, BS / GS, threadIdx.x / GS);
    uint32_t ovf = 0;
    for (uint32_t i = sp.first; i < sp.end; i += sp.stride) {
        const uint32_t u = alist[i];
        const uint32_t start = row[u], d = dint[u];
        if (d > heavy) continue;      // rows above the threshold go to k_walk_heavy
        FixAcc fa = {0ull, 0};
        // four independent index loads, then four term gathers in flight per
        // lane; slots past the row end re-read the row's last arc (no branch,
        // same line) and add a zero term
        const uint32_t* p = nbr + start;
        for (uint32_t j = lane; j < d; j += 4 * GS) {
            const bool b1 = j + GS < d, b2 = j + 2 * GS < d, b3 = j + 3 * GS < d;
            const uint32_t v0 = p[j];
            const uint32_t v1 = p[min(j + GS, d - 1)];
            const uint32_t v2 = p[min(j + 2 * GS, d - 1)];
            const uint32_t v3 = p[min(j + 3 * GS, d - 1)];
            const uint64_t y0 = y[v0], y1 = y[v1], y2 = y[v2], y3 = y[v3];
            fixacc_add(fa, y0);
            fixacc_add(fa, b1 ? y1 : 0ull);
            fixacc_add(fa, b2 ? y2 : 0ull);
            fixacc_add(fa, b3 ? y3 : 0ull);
            fixacc_norm(fa);
        }
        const Fix128 acc = fixacc_fix(group_sum_fixacc<GS>(fa));
        if (lane == 0) walk_finish(u, d, acc, xc, xn, yn, xkey, ovf);
    }
    if (ovf) atomicOr(&cnt->err, ERR_FIX_RANGE);
}

// ------------------------------------------------------------------ cut contribution (K4)

// edv[u].x = sum over valid arcs of (neighbour sorts before u ? -1 : +1), where
// "before" compares walk keys and breaks ties by node index
// (cuda/devSweepCut.h:20-112).
template <int GS>
__global__ void __launch_bounds__(BS) k_contrib(uint32_t na, const uint32_t* __restrict__ alist,
                                                const uint32_t* __restrict__ row,
                                                const uint32_t* __restrict__ dint,
                                                const uint32_t* __restrict__ nbr,
                                                const uint32_t* __restrict__ xkey,
                                                uint2* __restrict__ edv, uint32_t heavy) {
    const uint32_t lane = threadIdx.x & (GS - 1);
    const XcdSpan sp = xcd_span(na, BS / GS, threadIdx.x / GS);
    for (uint32_t i = sp.first; i < sp.end; i += sp.stride) {
        const uint32_t u = alist[i];
        const uint32_t start = row[u], d = dint[u];
        if (d > heavy) {              // k_contrib_heavy adds the chunks' counts
            if (lane == 0) edv[u] = make_uint2(0u, d);
            continue;
        }
        const uint32_t mk = xkey[u];
        int32_t c = 0;
        const uint32_t* p = nbr + start;
        for (uint32_t j = lane; j < d; j += 4 * GS) {
            uint32_t v[4], k[4];
            bool b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {           // clamped: no branch around the loads
                b[q] = j + q * GS < d;
                v[q] = p[min(j + q * GS, d - 1)];
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) k[q] = xkey[v[q]];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const bool before = (k[q] != mk) ? (k[q] < mk) : (v[q] < u);
                if (b[q]) c += before ? -1 : 1;
            }
        }
        c = group_sum<GS>(c);
        if (lane == 0) edv[u] = make_uint2((uint32_t)c, d);
    }
}

// Cut contribution with an LDS window of xkey: the layout of k_walk_win
// (chunks of CH list entries per workgroup, WIN keys around a chunk staged
// with coalesced 16-byte loads when its nodes span less than WIN/2, 64 rows
// per wave in GS rounds, each row's count returned to its owner lane).
template <int GS, int U, int WIN, int CH, bool RANGE = false>
__global__ void __launch_bounds__(WBS) k_contrib_win(uint32_t na, const uint32_t* __restrict__ alist,
                                                     const uint32_t* __restrict__ row,
                                                     const uint32_t* __restrict__ dint,
                                                     const uint32_t* __restrict__ nbr,
                                                     const uint32_t* __restrict__ xkey,
                                                     uint2* __restrict__ edv, uint32_t heavy, uint32_t N,
                                                     const uint32_t* __restrict__ coff = nullptr) {
    static_assert(WIN % 4 == 0 && 64 % GS == 0, "window of key quads, lane groups within a wave");
    __shared__ uint32_t sk[WIN];
    constexpr uint32_t NPR = 64 / GS;
    constexpr uint32_t NW = WBS / 64;
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    const uint32_t g = lane / GS, gl = lane % GS;
    const uint32_t nch = RANGE ? (N + CH - 1) / CH : (na + CH - 1) / CH;
    const uint32_t x = blockIdx.x % NXCD, lb = blockIdx.x / NXCD, nb = gridDim.x / NXCD;
    const uint32_t cpx = nch / NXCD + (nch % NXCD != 0);
    // range chunks: the XCD split by active entries (k_range_offsets)
    const uint32_t cb = RANGE ? coff[nch + 1 + x] : min(nch, x * cpx);
    const uint32_t ce = RANGE ? coff[nch + 2 + x] : min(nch, cb + cpx);
    for (uint32_t c = cb + lb; c < ce; c += nb) {
        uint32_t c0, c1, mid;
        bool use;
        if constexpr (RANGE) {
            c0 = coff[c];
            c1 = coff[c + 1];
            if (c0 == c1) continue;
            use = true;
            mid = c * CH + CH / 2;
        } else {
            c0 = c * CH;
            c1 = min(c0 + CH, na);
            const uint32_t uf = alist[c0], ul = alist[c1 - 1];
            use = ul - uf < (uint32_t)WIN / 2;
            mid = uf / 2 + ul / 2;
        }
        uint32_t wb = 0;
        if (use) {
            wb = mid > (uint32_t)WIN / 2 ? mid - (uint32_t)WIN / 2 : 0u;
            if (N <= (uint32_t)WIN) wb = 0;
            else if (wb > N - (uint32_t)WIN) wb = N - (uint32_t)WIN;
            wb &= ~3u;
            for (uint32_t k = 4 * threadIdx.x; k < (uint32_t)WIN; k += 4 * WBS) {
                if (wb + k + 3 < N) {
                    const uint4 t = *(const uint4*)(xkey + wb + k);
                    sk[k] = t.x;
                    sk[k + 1] = t.y;
                    sk[k + 2] = t.z;
                    sk[k + 3] = t.w;
                } else {
                    for (uint32_t q = 0; q < 4; ++q)
                        if (wb + k + q < N) sk[k + q] = xkey[wb + k + q];
                }
            }
        }
        __syncthreads();
        for (uint32_t ib = c0 + w * 64; ib < c1; ib += NW * 64) {
            const uint32_t i = ib + lane;
            const bool ok = i < c1;
            const uint32_t u = ok ? alist[i] : alist[c0];
            const uint32_t st = row[u];
            const uint32_t dd = dint[u];
            const bool light = dd <= heavy;
            const uint32_t d = ok && light ? dd : 0u;
            const uint32_t mk = xkey[u];
            int32_t res = 0;
#pragma unroll
            for (int r = 0; r < GS; ++r) {
                const uint32_t s = r * NPR + g;
                const uint32_t rs = __shfl(st, s), rd = __shfl(d, s), ru = __shfl(u, s), rk = __shfl(mk, s);
                int32_t cnt = 0;
                for (uint32_t j = gl; j < rd; j += U * GS) {
                    uint32_t v[U], k[U];
#pragma unroll
                    for (int q = 0; q < U; ++q) {
                        const uint32_t jj = j + q * GS;
                        v[q] = jj < rd ? nbr[rs + jj] : 0u;
                    }
                    bool valid[U];
#pragma unroll
                    for (int q = 0; q < U; ++q) valid[q] = j + q * GS < rd;
                    win_gather<U, false>(k, v, valid, sk, xkey, wb, (uint32_t)WIN, use);
#pragma unroll
                    for (int q = 0; q < U; ++q) {
                        const bool before = (k[q] != rk) ? (k[q] < rk) : (v[q] < ru);
                        if (valid[q]) cnt += before ? -1 : 1;
                    }
                }
                cnt = group_sum<GS>(cnt);
                const int32_t got = __shfl(cnt, (lane % NPR) * GS);
                if (lane / NPR == (uint32_t)r) res = got;
            }
            if (ok) edv[u] = make_uint2(light ? (uint32_t)res : 0u, dd);   // k_contrib_heavy adds long rows' chunks
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ heavy rows

// Rows above the light kernels' threshold are split into chunks of
// HEAVY_CHUNK arcs, one workgroup per chunk.  hcnt[i] = chunk count of the
// i-th active node (0 for light rows); its exclusive scan hoff gives each
// heavy row its chunk range, hrow[chunk] = the row's index in the list.
constexpr uint32_t HEAVY_CHUNK = BS * 4;

__global__ void k_heavy_count(uint32_t n, const uint32_t* __restrict__ alist, const uint32_t* __restrict__ dint,
                              uint32_t heavy, uint32_t* __restrict__ hcnt, const uint32_t* pn = nullptr) {
    n = dev_count(n, pn);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t d = dint[alist[i]];
        hcnt[i] = d > heavy ? (d + HEAVY_CHUNK - 1u) / HEAVY_CHUNK : 0u;
    }
}

__global__ void k_heavy_expand(uint32_t n, const uint32_t* __restrict__ hcnt, const uint32_t* __restrict__ hoff,
                               uint32_t* __restrict__ hrow, const uint32_t* pn = nullptr) {
    n = dev_count(n, pn);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t c = hcnt[i], o = hoff[i];
        for (uint32_t k = 0; k < c; ++k) hrow[o + k] = i;
    }
}

// coff[c] = first entry of the ascending node list at or after node c*R, for
// c = 0..nch (coff[nch] = n): the list entries of the node range chunks.
// Then the XCD split by entries: coff[nch + 1 + x] (x = 0..NXCD) = the first
// chunk whose entries start at or after x/NXCD of the list (x = NXCD: nch),
// so that every XCD gets about the same number of active nodes however they
// cluster in index order.
__device__ __forceinline__ uint32_t range_lower(uint32_t c, uint32_t nch, uint32_t R,
                                                const uint32_t* __restrict__ alist, uint32_t n) {
    if (c >= nch) return n;
    const uint64_t key = (uint64_t)c * R;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t m = lo + (hi - lo) / 2;
        if ((uint64_t)alist[m] < key) lo = m + 1; else hi = m;
    }
    return lo;
}
__global__ void k_range_offsets(uint32_t nch, uint32_t R, const uint32_t* __restrict__ alist, uint32_t n,
                                uint32_t* __restrict__ coff) {
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c <= nch; c += gridDim.x * blockDim.x) {
        const uint32_t lo = range_lower(c, nch, R, alist, n);
        const uint32_t prev = c ? range_lower(c - 1, nch, R, alist, n) : 0u;
        coff[c] = lo;
        // chunk c is XCD x's first when target_x lies in (prev, lo] (c = 0: target_x = 0)
        for (uint32_t x = 0; x <= NXCD; ++x) {
            const uint32_t t = (uint32_t)((uint64_t)x * n / NXCD);
            const bool first = x == NXCD ? c == nch : (c == 0 ? t == 0u : (prev < t && t <= lo));
            if (first) coff[nch + 1 + x] = c;
        }
    }
}

// Walk with an LDS window of y.  A workgroup of WBS threads takes chunks of CH
// consecutive node-list entries.  When a chunk's nodes span less than WIN/2
// indices, the WIN values of y around them are staged into LDS with coalesced
// 16-byte loads, and every neighbour inside the window is read from LDS;
// the rest (and every neighbour of a wider chunk) from global memory.  Graphs
// whose neighbours sit near in index order (planted blocks, grids) then turn
// most per-arc 128-byte line fetches into LDS reads.  Each wave takes 64
// entries at a time: lane l loads the metadata of entry l (coalesced), the 64
// rows are summed in GS rounds of 64/GS rows with GS lanes per row and U
// gathers in flight per lane, each exact row sum goes back to its owner lane,
// and all 64 lanes round and store together.  Same arithmetic as k_walk, so
// the same bits.  Rows above `heavy` are left to k_walk_heavy.
// CLAMP: slots past a row's end load the row's last arc again (no exec-mask
// region per slot), the LDS value is read unconditionally and only window
// misses branch to a global load; the slot's term is then dropped.  Fewer
// branch regions when rows fill their slots, more wasted loads when they do not.
//
// RANGE: chunks are node-index ranges [c*CH, (c+1)*CH) instead of CH list
// entries; coff[c] (k_range_offsets) is the first list entry of chunk c.  The
// window always covers its chunk, however sparse the list (a serving engine's
// node space, late iterations); chunks without active nodes are skipped.
//
// A window of at most 80 KB asks for 8 waves per SIMD (at most 64 VGPRs), so
// that two workgroups share a CU: one stages its window while the other sums.
template <int GS, int U, int WIN, int CH, bool CLAMP = false, bool RANGE = false>
__global__ void __launch_bounds__(WBS, WIN * 8 <= 81920 ? 8 : 4) k_walk_win(uint32_t na, const uint32_t* __restrict__ alist,
                                                  const uint32_t* __restrict__ row,
                                                  const uint32_t* __restrict__ dint,
                                                  const uint32_t* __restrict__ nbr,
                                                  const uint64_t* __restrict__ y,
                                                  const double* __restrict__ xc, double* __restrict__ xn,
                                                  uint64_t* __restrict__ yn, uint32_t* __restrict__ xkey,
                                                  uint32_t heavy, uint32_t N, Counters* cnt,
                                                  const uint32_t* __restrict__ coff = nullptr) {
    static_assert(WIN % 2 == 0 && 64 % GS == 0, "window of term pairs, lane groups within a wave");
    __shared__ uint64_t sy[WIN];
    constexpr uint32_t NPR = 64 / GS;           // rows per round
    constexpr uint32_t NW = WBS / 64;
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    const uint32_t g = lane / GS, gl = lane % GS;
    // XCD-aware chunk split (see xcd_span): XCD x takes the x-th eighth of the
    // chunks; range chunks are split by active entries (k_range_offsets)
    const uint32_t nch = RANGE ? (N + CH - 1) / CH : (na + CH - 1) / CH;
    const uint32_t x = blockIdx.x % NXCD, lb = blockIdx.x / NXCD, nb = gridDim.x / NXCD;
    const uint32_t cpx = nch / NXCD + (nch % NXCD != 0);
    const uint32_t cb = RANGE ? coff[nch + 1 + x] : min(nch, x * cpx);
    const uint32_t ce = RANGE ? coff[nch + 2 + x] : min(nch, cb + cpx);
    uint32_t ovf = 0;
    for (uint32_t c = cb + lb; c < ce; c += nb) {
        uint32_t c0, c1, mid;
        bool use;
        if constexpr (RANGE) {
            c0 = coff[c];
            c1 = coff[c + 1];
            if (c0 == c1) continue;   // no active node in the range (uniform over the workgroup)
            use = true;
            mid = c * CH + CH / 2;
        } else {
            c0 = c * CH;
            c1 = min(c0 + CH, na);
            const uint32_t uf = alist[c0], ul = alist[c1 - 1];   // the list is ascending
            use = ul - uf < (uint32_t)WIN / 2;
            mid = uf / 2 + ul / 2;
        }
        uint32_t wb = 0;
        if (use) {
            wb = mid > (uint32_t)WIN / 2 ? mid - (uint32_t)WIN / 2 : 0u;
            if (N <= (uint32_t)WIN) wb = 0;
            else if (wb > N - (uint32_t)WIN) wb = N - (uint32_t)WIN;
            wb &= ~1u;
            for (uint32_t k = 2 * threadIdx.x; k < (uint32_t)WIN; k += 2 * WBS) {
                if (wb + k + 1 < N) {
                    const ulonglong2 t = *(const ulonglong2*)(y + wb + k);
                    sy[k] = t.x;
                    sy[k + 1] = t.y;
                } else if (wb + k < N) {
                    sy[k] = y[wb + k];
                }
            }
        }
        __syncthreads();
        for (uint32_t ib = c0 + w * 64; ib < c1; ib += NW * 64) {
            const uint32_t i = ib + lane;
            const bool ok = i < c1;
            const uint32_t u = ok ? alist[i] : alist[c0];
            const uint32_t st = row[u];
            const uint32_t dd = dint[u];
            const bool mine = ok && dd <= heavy;
            const uint32_t d = mine ? dd : 0u;
            const double xo = xc[u];
            FixAcc res = {0ull, 0};
#pragma unroll
            for (int r = 0; r < GS; ++r) {
                const uint32_t s = r * NPR + g;
                const uint32_t rs = __shfl(st, s), rd = __shfl(d, s);
                FixAcc acc = {0ull, 0};
                for (uint32_t j = gl; j < rd; j += U * GS) {
                    uint32_t q[U];
                    uint64_t yv[U];
                    if constexpr (CLAMP) {
#pragma unroll
                        for (int k = 0; k < U; ++k) q[k] = nbr[rs + min(j + k * GS, rd - 1)];
                    } else {
#pragma unroll
                        for (int k = 0; k < U; ++k) {
                            const uint32_t jj = j + k * GS;
                            q[k] = jj < rd ? nbr[rs + jj] : 0u;
                        }
                    }
                    // window words from LDS, misses from global memory, slots
                    // past the row's end 0
                    bool valid[U];
#pragma unroll
                    for (int k = 0; k < U; ++k) valid[k] = j + k * GS < rd;
                    win_gather<U, true>(yv, q, valid, sy, y, wb, (uint32_t)WIN, use);
#pragma unroll
                    for (int k = 0; k < U; ++k) fixacc_add(acc, yv[k]);
                    fixacc_norm(acc);
                }
                acc = group_sum_fixacc<GS>(acc);
                const uint32_t src = (lane % NPR) * GS;
                const uint64_t lo = __shfl(acc.lo, src);
                const int64_t hi = __shfl(acc.hi, src);
                if (lane / NPR == (uint32_t)r) res = FixAcc{lo, hi};
            }
            if (mine) {
                if (d > (1u << 24)) ovf = 1;
                const double sum = fix_to_f64(fixacc_fix(res));
                const double v = sum * (1.0 - RW_STAY) + xo * RW_STAY;
                xn[u] = v;
                yn[u] = fix_pack(d ? v / (double)d : 0.0, ovf);
                xkey[u] = ord_f32((float)v);
            }
        }
        __syncthreads();              // the window is restaged for the next chunk
    }
    if (ovf) atomicOr(&cnt->err, ERR_FIX_RANGE);
}

// sum of a Fix128 over the workgroup, valid in thread 0
__device__ __forceinline__ Fix128 block_sum_fix(Fix128 a, Fix128* lds) {
    a = group_sum_fix<WAVE>(a);
    if (lane_id() == 0) lds[threadIdx.x / WAVE] = a;
    __syncthreads();
    if (threadIdx.x == 0)
        for (uint32_t w = 1; w < BS / WAVE; ++w) fix_add128(a, lds[w].lo, lds[w].hi);
    return a;
}

// partial exact sum of one chunk of a heavy row; chunk count read on the device
__global__ void __launch_bounds__(BS) k_walk_heavy(const uint32_t* __restrict__ nchunks,
                                                   const uint32_t* __restrict__ hrow,
                                                   const uint32_t* __restrict__ hoff,
                                                   const uint32_t* __restrict__ alist,
                                                   const uint32_t* __restrict__ row,
                                                   const uint32_t* __restrict__ dint,
                                                   const uint32_t* __restrict__ nbr,
                                                   const uint64_t* __restrict__ y, Fix128* __restrict__ part,
                                                   Counters* cnt) {
    __shared__ Fix128 lds[BS / WAVE];
    const uint32_t H = *nchunks;
    uint32_t ovf = 0;
    for (uint32_t b = blockIdx.x; b < H; b += gridDim.x) {
        const uint32_t i = hrow[b];
        const uint32_t u = alist[i];
        const uint32_t d = dint[u];
        const uint32_t j0 = (b - hoff[i]) * HEAVY_CHUNK;
        const uint32_t* p = nbr + row[u];
        uint32_t v[4];
        bool ok[4];
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t j = j0 + q * BS + threadIdx.x;
            ok[q] = j < d;
            v[q] = p[min(j, d - 1u)];
        }
        uint64_t yv[4];
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) yv[q] = y[v[q]];
        FixAcc fa = {0ull, 0};
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) fixacc_add(fa, ok[q] ? yv[q] : 0ull);
        Fix128 acc = block_sum_fix(fixacc_fix(fa), lds);
        if (threadIdx.x == 0) part[b] = acc;
        __syncthreads();              // lds reused by the next chunk
    }
    if (ovf) atomicOr(&cnt->err, ERR_FIX_RANGE);
}

// per heavy row (its first chunk): exact sum of the chunk partials, then the
// same rounding and stores as the light kernel
__global__ void k_walk_heavy_finish(const uint32_t* __restrict__ nchunks, const uint32_t* __restrict__ hrow,
                                    const uint32_t* __restrict__ hoff, const uint32_t* __restrict__ alist,
                                    const uint32_t* __restrict__ dint, const Fix128* __restrict__ part,
                                    const double* __restrict__ xc, double* __restrict__ xn,
                                    uint64_t* __restrict__ yn, uint32_t* __restrict__ xkey, Counters* cnt) {
    const uint32_t H = *nchunks;
    uint32_t ovf = 0;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < H; b += gridDim.x * blockDim.x) {
        const uint32_t i = hrow[b];
        if (hoff[i] != b) continue;
        const uint32_t u = alist[i];
        const uint32_t d = dint[u];
        const uint32_t nch = (d + HEAVY_CHUNK - 1u) / HEAVY_CHUNK;
        Fix128 acc = {0ull, 0ull};
        for (uint32_t k = 0; k < nch; ++k) fix_add128(acc, part[b + k].lo, part[b + k].hi);
        walk_finish(u, d, acc, xc, xn, yn, xkey, ovf);
    }
    if (ovf) atomicOr(&cnt->err, ERR_FIX_RANGE);
}

// cut contribution of one chunk of a heavy row, added to edv[u].x (exact integers)
__global__ void __launch_bounds__(BS) k_contrib_heavy(const uint32_t* __restrict__ nchunks,
                                                      const uint32_t* __restrict__ hrow,
                                                      const uint32_t* __restrict__ hoff,
                                                      const uint32_t* __restrict__ alist,
                                                      const uint32_t* __restrict__ row,
                                                      const uint32_t* __restrict__ dint,
                                                      const uint32_t* __restrict__ nbr,
                                                      const uint32_t* __restrict__ xkey,
                                                      uint2* __restrict__ edv) {
    __shared__ int32_t lds[BS / WAVE];
    const uint32_t H = *nchunks;
    for (uint32_t b = blockIdx.x; b < H; b += gridDim.x) {
        const uint32_t i = hrow[b];
        const uint32_t u = alist[i];
        const uint32_t d = dint[u];
        const uint32_t mk = xkey[u];
        const uint32_t j0 = (b - hoff[i]) * HEAVY_CHUNK;
        const uint32_t* p = nbr + row[u];
        uint32_t v[4], k[4];
        bool ok[4];
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t j = j0 + q * BS + threadIdx.x;
            ok[q] = j < d;
            v[q] = p[min(j, d - 1u)];
        }
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) k[q] = xkey[v[q]];
        int32_t c = 0;
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
            const bool before = (k[q] != mk) ? (k[q] < mk) : (v[q] < u);
            if (ok[q]) c += before ? -1 : 1;
        }
        c = group_sum<WAVE>(c);
        if (lane_id() == 0) lds[threadIdx.x / WAVE] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            int32_t t = 0;
            for (uint32_t w = 0; w < BS / WAVE; ++w) t += lds[w];
            atomicAdd(&edv[u].x, (uint32_t)t);   // two's complement: the signed sum
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ generic scan

// out[i] = sum_{j<i} in[j] for i < n, out[n] = total (3-level reduce-then-scan).
__global__ void __launch_bounds__(BS) k_scan_up(uint32_t n, const uint32_t* __restrict__ in,
                                                uint32_t* __restrict__ bsum, const uint32_t* pn = nullptr) {
    __shared__ uint32_t lds[BS / WAVE];
    n = dev_count(n, pn);             // blocks past n add nothing (their sums are 0)
    const uint32_t base = blockIdx.x * (BS * SCAN_ITEMS) + threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) s += (base + i < n) ? in[base + i] : 0u;
    uint32_t tot;
    block_excl_scan<BS>(s, lds, tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(BS) k_scan_down(uint32_t n, const uint32_t* __restrict__ in,
                                                  uint32_t* __restrict__ out,
                                                  const uint32_t* __restrict__ boff, const uint32_t* pn = nullptr,
                                                  uint32_t* total = nullptr) {
    __shared__ uint32_t lds[BS / WAVE];
    n = dev_count(n, pn);
    const uint32_t base = blockIdx.x * (BS * SCAN_ITEMS) + threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) { v[i] = (base + i < n) ? in[base + i] : 0u; s += v[i]; }
    uint32_t tot;
    uint32_t run = block_excl_scan<BS>(s, lds, tot) + (boff ? boff[blockIdx.x] : 0u);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    // the block holding the last element (block 0 when n == 0) writes the total
    const uint32_t last = n ? (n - 1) / (BS * SCAN_ITEMS) : 0u;
    if (blockIdx.x == last && threadIdx.x == BS - 1) {
        out[n] = run;
        if (total) *total = run;
    }
}

// Exclusive scan of up to SMALL_SCAN_N elements in one workgroup (one launch
// instead of the three of the two-level scan); out[n] = total.
constexpr uint32_t SMALL_SCAN_ITEMS = 16;
constexpr uint32_t SMALL_SCAN_N = 1024 * SMALL_SCAN_ITEMS;
__global__ void __launch_bounds__(1024) k_scan_small(uint32_t n, const uint32_t* __restrict__ in,
                                                     uint32_t* __restrict__ out, const uint32_t* pn = nullptr,
                                                     uint32_t* total = nullptr) {
    __shared__ uint32_t lds[1024 / WAVE];
    n = dev_count(n, pn);
    const uint32_t base = threadIdx.x * SMALL_SCAN_ITEMS;
    uint32_t v[SMALL_SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t i = 0; i < SMALL_SCAN_ITEMS; ++i) {
        v[i] = base + i < n ? in[base + i] : 0u;
        s += v[i];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan<1024>(s, lds, tot);
#pragma unroll
    for (uint32_t i = 0; i < SMALL_SCAN_ITEMS; ++i) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == 0) {
        out[n] = tot;
        if (total) *total = tot;
    }
}

// ------------------------------------------------------------------ decoupled look-back

// Status words of the single-pass kernels below (Merrill & Garland, "Single-
// pass Parallel Prefix Scan with Decoupled Look-back", 2016): bits 0-31 a
// count, bits 32-33 the state (LB_AGG = the tile's own total, LB_INC = the
// inclusive prefix up to and including the tile), bits 34-63 the launch's
// epoch.  A word of another epoch reads as not written yet, so the status
// arrays are never cleared between launches (only once, when allocated).
// Tiles are handed out in ticket order (an atomic counter), so a tile only
// waits for tiles whose workgroups have started, and those publish without
// waiting for later ones: every wait ends.  The workgroup that draws the
// grid's last ticket sets the counter back to 0 for the next launch.
constexpr uint32_t LB_AGG = 1u, LB_INC = 2u;
constexpr uint32_t LB_EPOCH_MAX = (1u << 30) - 1u;

__device__ __forceinline__ unsigned long long lb_word(uint32_t epoch, uint32_t state, uint32_t v) {
    return ((unsigned long long)epoch << 34) | ((unsigned long long)state << 32) | v;
}
__device__ __forceinline__ void lb_publish(unsigned long long* p, unsigned long long w) {
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long lb_read(unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// state of a word in this epoch (0 = not written yet)
__device__ __forceinline__ uint32_t lb_state(unsigned long long w, uint32_t epoch) {
    return (uint32_t)(w >> 34) == epoch ? (uint32_t)(w >> 32) & 3u : 0u;
}
// called by one thread of the workgroup
__device__ __forceinline__ uint32_t lb_ticket(uint32_t* ticket) {
    const uint32_t t = atomicAdd(ticket, 1u);
    if (t == gridDim.x - 1u) atomicExch(ticket, 0u);   // every workgroup has drawn its ticket
    return t;
}

// Single-pass exclusive scan: out[i] = sum_{j<i} in[j], out[n] = total (also
// to *total when given); n = *pn when pn is given, the grid then covers the
// bound n.  One tile of LB_SCAN_TILE elements per workgroup.  A status hop
// costs a device-scope load (about a microsecond across the XCDs), and the
// inclusive prefixes advance by one look-back window per hop, so tiles are
// large (8192 elements) and the whole workgroup reads a window of BS earlier
// tiles at a time.
constexpr uint32_t LB_SCAN_ITEMS = 32;
constexpr uint32_t LB_SCAN_TILE = BS * LB_SCAN_ITEMS;
static_assert(LB_SCAN_ITEMS % 32u == 0u || 32u % LB_SCAN_ITEMS == 0u, "padded LDS index of a thread's run");
__global__ void __launch_bounds__(BS) k_scan_lb(uint32_t n, const uint32_t* __restrict__ in,
                                                uint32_t* __restrict__ out, unsigned long long* status,
                                                uint32_t* ticket, uint32_t epoch, const uint32_t* pn = nullptr,
                                                uint32_t* total = nullptr) {
    constexpr uint32_t NW = BS / WAVE;
    // the tile in LDS, element k at k + k / 32: coalesced loads and stores
    // in element order, each thread's LB_SCAN_ITEMS consecutive elements
    // read without bank conflicts
    __shared__ uint32_t sm[LB_SCAN_TILE + LB_SCAN_TILE / 32];
    __shared__ uint32_t lds[NW];
    __shared__ uint32_t s_tile, s_first;
    if (threadIdx.x == 0) s_tile = lb_ticket(ticket);
    __syncthreads();
    const uint32_t tile = s_tile;
    n = dev_count(n, pn);
    const uint32_t ntiles = n ? (n - 1u) / LB_SCAN_TILE + 1u : 1u;   // tile 0 writes out[0] when n == 0
    if (tile >= ntiles) return;
    const uint32_t base = tile * LB_SCAN_TILE;
    {
        // every element into registers first (elements past the end read the
        // last one and count 0), so the loads are in flight together
        uint32_t t[LB_SCAN_ITEMS];
#pragma unroll
        for (uint32_t i = 0; i < LB_SCAN_ITEMS; ++i) t[i] = n ? in[min(base + i * BS + threadIdx.x, n - 1u)] : 0u;
#pragma unroll
        for (uint32_t i = 0; i < LB_SCAN_ITEMS; ++i) {
            const uint32_t k = i * BS + threadIdx.x;
            sm[k + (k >> 5)] = base + k < n ? t[i] : 0u;
        }
    }
    __syncthreads();
    const uint32_t k0 = threadIdx.x * LB_SCAN_ITEMS + threadIdx.x * LB_SCAN_ITEMS / 32u;
    uint32_t v[LB_SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (uint32_t i = 0; i < LB_SCAN_ITEMS; ++i) {
        v[i] = sm[k0 + i + i / 32u];
        s += v[i];
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan<BS>(s, lds, tot);
    if (threadIdx.x == 0) lb_publish(&status[tile], lb_word(epoch, tile == 0 ? LB_INC : LB_AGG, tot));
    uint32_t prefix = 0;
    if (tile != 0) {
        // windows of BS earlier tiles; the nearest inclusive one ends the walk
        for (int64_t top = (int64_t)tile - 1;; top -= BS) {
            const int64_t j = top - (int64_t)threadIdx.x;
            unsigned long long w = 0ull;
            uint32_t stt = LB_INC;    // tiles before 0 read as an inclusive prefix of 0
            if (j >= 0) {
                while ((stt = lb_state(w = lb_read(&status[j]), epoch)) == 0u) __builtin_amdgcn_s_sleep(1);
            }
            if (threadIdx.x == 0) s_first = BS;
            __syncthreads();
            if (stt == LB_INC) atomicMin(&s_first, threadIdx.x);
            __syncthreads();
            const uint32_t first = s_first;
            uint32_t part;
            (void)block_excl_scan<BS>(threadIdx.x <= first && j >= 0 ? (uint32_t)w : 0u, lds, part);
            prefix += part;
            if (first < BS) break;
        }
        if (threadIdx.x == 0) lb_publish(&status[tile], lb_word(epoch, LB_INC, prefix + tot));
    }
    uint32_t run = prefix + ex;
#pragma unroll
    for (uint32_t i = 0; i < LB_SCAN_ITEMS; ++i) {
        sm[k0 + i + i / 32u] = run;
        run += v[i];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t i = 0; i < LB_SCAN_ITEMS; ++i) {
        const uint32_t k = i * BS + threadIdx.x;
        if (base + k < n) out[base + k] = sm[k + (k >> 5)];
    }
    if (tile == ntiles - 1u && threadIdx.x == 0) {
        out[n] = prefix + tot;
        if (total) *total = prefix + tot;
    }
}

// Single-pass ranked selection, the status protocol of k_scan_lb: element i
// of [0, n) has the values d = f.load(i) and the flag f.flag(i, d), and every
// element gets f.emit(i, r, flag, d) with r = the number of flagged elements
// before it; the flagged count goes to *total (when given).  load() does all
// of an element's reads and emit() only writes: the kernel loads every row of
// its tile before the first flag and emits from registers, so the rows' loads
// are in flight together (rows past the end load the last element and are
// dropped).  One launch replaces a flag kernel, a scan and a
// scatter kernel (the flags and prefixes never go to memory).  A tile is
// ROWS rows of BS consecutive elements, row q read by all threads (element
// base + q * BS + t), so the predicate's loads coalesce; an element's rank
// inside the tile is the flagged count of the earlier (row, wave) pairs plus
// the wave ballot's count of flagged lanes below it.  Long lists take tiles
// of SE_ROWS rows (few tiles for the look-back to cross), shorter ones
// SE_ROWS_SHORT rows (enough workgroups to occupy the chip).
constexpr uint32_t SE_ROWS = 16, SE_ROWS_SHORT = 8;
constexpr uint32_t SE_LONG = 1u << 23;             // lists of at least this many elements: SE_ROWS
template <uint32_t SE_ROWS, class F>
__global__ void __launch_bounds__(BS) k_scan_emit(uint32_t n, F f, unsigned long long* status, uint32_t* ticket,
                                                  uint32_t epoch, const uint32_t* pn, uint32_t* total) {
    constexpr uint32_t NW = BS / WAVE;
    constexpr uint32_t SE_TILE = BS * SE_ROWS;
    static_assert(SE_ROWS * NW <= BS, "one (row, wave) count a thread");
    __shared__ uint32_t s_cnt[SE_ROWS * NW];
    __shared__ uint32_t lds[NW];
    __shared__ uint32_t s_tile, s_first;
    if (threadIdx.x == 0) s_tile = lb_ticket(ticket);
    __syncthreads();
    const uint32_t tile = s_tile;
    n = dev_count(n, pn);
    const uint32_t ntiles = n ? (n - 1u) / SE_TILE + 1u : 1u;   // tile 0 writes the total when n == 0
    if (tile >= ntiles) return;
    const uint32_t base = tile * SE_TILE;
    const uint32_t w = threadIdx.x / WAVE, l = lane_id();
    const uint64_t lt = (1ull << l) - 1ull;
    // all rows' values and flags first (no branch per row: rows past the end
    // load the last element), then the ballots
    typename F::D d[SE_ROWS];
    uint32_t fl = 0u;
    if (n != 0u) {
#pragma unroll
        for (uint32_t q = 0; q < SE_ROWS; ++q) d[q] = f.load(min(base + q * BS + threadIdx.x, n - 1u));
#pragma unroll
        for (uint32_t q = 0; q < SE_ROWS; ++q) {
            const uint32_t i = base + q * BS + threadIdx.x;
            const bool fv = f.flag(min(i, n - 1u), d[q]);
            if (i < n && fv) fl |= 1u << q;
        }
    }
    uint32_t rk[SE_ROWS];
#pragma unroll
    for (uint32_t q = 0; q < SE_ROWS; ++q) {
        const uint64_t bal = __ballot((fl >> q) & 1u);
        rk[q] = (uint32_t)__popcll(bal & lt);
        if (l == 0) s_cnt[q * NW + w] = (uint32_t)__popcll(bal);
    }
    __syncthreads();
    uint32_t tot;
    {
        const uint32_t v = threadIdx.x < SE_ROWS * NW ? s_cnt[threadIdx.x] : 0u;
        const uint32_t ex = block_excl_scan<BS>(v, lds, tot);
        if (threadIdx.x < SE_ROWS * NW) s_cnt[threadIdx.x] = ex;
    }
    if (threadIdx.x == 0) lb_publish(&status[tile], lb_word(epoch, tile == 0 ? LB_INC : LB_AGG, tot));
    uint32_t prefix = 0;
    if (tile != 0) {
        // windows of BS earlier tiles; the nearest inclusive one ends the walk
        for (int64_t top = (int64_t)tile - 1;; top -= BS) {
            const int64_t j = top - (int64_t)threadIdx.x;
            unsigned long long wd = 0ull;
            uint32_t stt = LB_INC;    // tiles before 0 read as an inclusive prefix of 0
            if (j >= 0) {
                while ((stt = lb_state(wd = lb_read(&status[j]), epoch)) == 0u) __builtin_amdgcn_s_sleep(1);
            }
            if (threadIdx.x == 0) s_first = BS;
            __syncthreads();
            if (stt == LB_INC) atomicMin(&s_first, threadIdx.x);
            __syncthreads();
            const uint32_t first = s_first;
            uint32_t part;
            (void)block_excl_scan<BS>(threadIdx.x <= first && j >= 0 ? (uint32_t)wd : 0u, lds, part);
            prefix += part;
            if (first < BS) break;
        }
        if (threadIdx.x == 0) lb_publish(&status[tile], lb_word(epoch, LB_INC, prefix + tot));
    }
    __syncthreads();                  // s_cnt holds the (row, wave) prefixes
#pragma unroll
    for (uint32_t q = 0; q < SE_ROWS; ++q) {
        const uint32_t i = base + q * BS + threadIdx.x;
        if (i < n) f.emit(i, prefix + s_cnt[q * NW + w] + rk[q], ((fl >> q) & 1u) != 0u, d[q]);
    }
    if (tile == ntiles - 1u && threadIdx.x == 0 && total) *total = prefix + tot;
}

// k_scan_emit selections of the iteration:
// the CC members in node order: active nodes stamped by this cut
// (and every member's pointer set to itself, in node order)
struct SelMembers {
    const uint32_t* alist;
    const uint32_t* stamp;
    uint32_t iter;
    uint32_t* out;
    uint32_t* root;
    struct D {
        uint32_t u, s;
    };
    __device__ D load(uint32_t i) const {
        const uint32_t u = alist[i];
        return D{u, stamp[u]};
    }
    __device__ bool flag(uint32_t, const D& d) const { return d.s == iter; }
    __device__ void emit(uint32_t, uint32_t r, bool f, const D& d) const {
        if (f) {
            out[r] = d.u;
            root[d.u] = d.u;
        }
    }
};
// the members outside their part's giant tree
struct SelOffGiant {
    const uint32_t* list;
    const uint32_t* part;
    const uint32_t* giant;
    const uint32_t* root;
    uint32_t* out;
    struct D {
        uint32_t u, g, p;             // member, its part's giant root, its pointer
    };
    __device__ D load(uint32_t i) const {
        const uint32_t u = list[i];
        return D{u, giant[part[u]], root[u]};
    }
    __device__ bool flag(uint32_t i, const D& d) const;
    __device__ void emit(uint32_t, uint32_t r, bool f, const D& d) const {
        if (f) out[r] = d.u;
    }
};
// the positions of the kept clusters, their new cluster ids,
// and every kept cluster's new first position (cstart, by old cluster id)
struct SelPositions {
    const uint32_t* pnix;
    const uint32_t* pclu;
    const uint32_t* ckeep;
    const uint32_t* kscan;
    const uint32_t* cs;
    uint32_t *pnix_o, *pclu_o, *cstart;
    struct D {
        uint32_t c, u, keep, nc, first;   // cluster, node, kept, new cluster id, cluster start
    };
    __device__ D load(uint32_t k) const {
        const uint32_t c = pclu[k];
        return D{c, pnix[k], ckeep[c], kscan[c], cs[c]};
    }
    __device__ bool flag(uint32_t, const D& d) const { return d.keep != 0u; }
    __device__ void emit(uint32_t k, uint32_t r, bool f, const D& d) const {
        if (!f) return;
        pnix_o[r] = d.u;
        pclu_o[r] = d.nc;
        if (d.first == k) cstart[d.c] = r;
    }
};
// the active nodes still alive
struct SelAlive {
    const uint32_t* alist;
    const uint8_t* alive;
    uint32_t* out;
    struct D {
        uint32_t u, a;
    };
    __device__ D load(uint32_t i) const {
        const uint32_t u = alist[i];
        return D{u, alive[u]};
    }
    __device__ bool flag(uint32_t, const D& d) const { return d.a != 0u; }
    __device__ void emit(uint32_t, uint32_t r, bool f, const D& d) const {
        if (f) out[r] = d.u;
    }
};
// the cluster table from runs of equal regroup keys (sorted by minimum node):
// a run starts a cluster, every position gets its run's id
struct SelRuns {
    uint32_t na;
    const uint32_t* key;
    const uint32_t* ngraph;
    uint32_t *cs, *cmin, *cgraph, *cflag, *pclu;
    struct D {
        uint32_t key, prev;
    };
    __device__ D load(uint32_t k) const { return D{key[k], key[k ? k - 1u : 0u]}; }
    __device__ bool flag(uint32_t k, const D& d) const { return (k == 0u) | (d.key != d.prev); }
    __device__ void emit(uint32_t k, uint32_t r, bool f, const D& d) const {
        const uint32_t c = r + (f ? 1u : 0u) - 1u;
        pclu[k] = c;
        if (f) {
            cs[c] = k;
            cmin[c] = d.key;
            cgraph[c] = ngraph[d.key];
            cflag[c] = CF_CONN;
        }
        if (k == na - 1u) cs[c + 1u] = na;
    }
};

// ------------------------------------------------------------------ segmented sort

// Sort classes: size <= 1 nothing, <= 64 one wave, <= TILE one block in LDS,
// larger: LSD radix over 8-bit digits with per-tile histograms.
__global__ void __launch_bounds__(BS) k_classify(uint32_t nseg, const uint32_t* __restrict__ segs,
                                                 uint32_t* __restrict__ lwave, uint32_t* __restrict__ lblock,
                                                 uint32_t* __restrict__ llarge, Counters* cnt) {
    for (uint32_t c0 = blockIdx.x * blockDim.x; c0 < nseg; c0 += gridDim.x * blockDim.x) {
        const uint32_t c = c0 + threadIdx.x;
        const uint32_t s = c < nseg ? segs[c + 1] - segs[c] : 0u;
        const bool w = s > 1u && s <= (uint32_t)WAVE;
        const bool b = s > (uint32_t)WAVE && s <= (uint32_t)TILE;
        const bool l = s > (uint32_t)TILE;
        const uint32_t sw = wave_append(&cnt->nwave, w);
        const uint32_t sb = wave_append(&cnt->nblock, b);
        const uint32_t sl = wave_append(&cnt->nlarge, l);
        if (w) lwave[sw] = c;
        if (b) lblock[sb] = c;
        if (l) llarge[sl] = c;
    }
}

// one wave per segment (<= 64 items): rank by counting, stable on input order;
// waves stride over the list (its length read on the device)
__global__ void __launch_bounds__(BS) k_sort_wave(const uint32_t* __restrict__ pnlist, const uint32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ segs,
                                                  uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
  const uint32_t nlist = *pnlist;
  for (uint32_t w = blockIdx.x * (BS / WAVE) + threadIdx.x / WAVE; w < nlist; w += gridDim.x * (BS / WAVE)) {
    const uint32_t lane = lane_id();
    const uint32_t c = list[w];
    const uint32_t s0 = segs[c], s = segs[c + 1] - s0;
    const bool valid = lane < s;
    const uint32_t k = valid ? key[s0 + lane] : 0xFFFFFFFFu;
    const uint32_t v = valid ? val[s0 + lane] : 0u;
    uint32_t rank = 0;
    for (uint32_t j = 0; j < s; ++j) {
        const uint32_t kj = __shfl(k, (int)j);
        rank += (kj < k || (kj == k && j < lane)) ? 1u : 0u;
    }
    if (valid) {
        key[s0 + rank] = k;
        val[s0 + rank] = v;
    }
  }
}

// Stable rank of a tile's items by the 8-bit digit at `shift`.  Item i of
// lane l of wave w is tile position w*WAVE*TILE_ITEMS + i*WAVE + l (valid
// below cnt), so a wave's rounds visit the tile in order: an item's rank is
// the number of same-digit lanes below it in its round (eight ballots give the
// peer mask) plus the wave's running count of that digit, which the lowest peer
// reads and advances in LDS (the others get it by shuffle).  Per-digit prefixes
// over the waves and the digit starts hs[] give pos[i], the item's position in
// the digit-sorted tile.  Block-wide (every thread calls it); ends with a
// barrier; hs stays valid until the next call.
static_assert(BS == 256, "one thread per 8-bit digit");
__device__ __forceinline__ void rank8(const uint64_t (&e)[TILE_ITEMS], uint32_t cnt, int shift,
                                      uint32_t (*wcnt)[256], uint32_t* hs, uint32_t* scratch,
                                      uint32_t (&pos)[TILE_ITEMS]) {
    constexpr uint32_t NW = BS / WAVE;
    const uint32_t w = threadIdx.x / WAVE, l = lane_id();
    for (uint32_t q = threadIdx.x; q < NW * 256u; q += BS) (&wcnt[0][0])[q] = 0u;
    __syncthreads();
    const uint64_t lt = (1ull << l) - 1ull;
    uint32_t dg[TILE_ITEMS], rk[TILE_ITEMS];
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; ++i) {
        const uint32_t p = w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l;
        const bool ok = p < cnt;
        const uint32_t d = (uint32_t)(e[i] >> (32 + shift)) & 255u;
        uint64_t peer = __ballot(ok);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool one = (d >> bit) & 1u;
            const uint64_t bb = __ballot(one);
            peer &= one ? bb : ~bb;
        }
        const bool lead = ok && (peer & lt) == 0ull;
        const uint32_t leader = ok ? (uint32_t)__builtin_ctzll(peer) : l;
        const uint32_t c0 = lead ? wcnt[w][d] : 0u;
        const uint32_t c = __shfl(c0, (int)leader);
        if (lead) wcnt[w][d] = c0 + (uint32_t)__popcll(peer);
        dg[i] = d;
        rk[i] = c + (uint32_t)__popcll(peer & lt);
    }
    __syncthreads();
    {
        const uint32_t d = threadIdx.x;   // wcnt[w][d] becomes the count of digit d in waves < w
        uint32_t run = 0;
#pragma unroll
        for (uint32_t ww = 0; ww < NW; ++ww) {
            const uint32_t c = wcnt[ww][d];
            wcnt[ww][d] = run;
            run += c;
        }
        uint32_t tot;
        hs[d] = block_excl_scan<BS>(run, scratch, tot);
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; ++i) pos[i] = hs[dg[i]] + wcnt[w][dg[i]] + rk[i];
    __syncthreads();
}

// one block per segment of 65..TILE items, sorted in place; blocks stride
// over the list (its length read on the device)
__global__ void __launch_bounds__(BS) k_sort_block(const uint32_t* __restrict__ pnlist, const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ segs,
                                                   uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
    constexpr uint32_t NW = BS / WAVE;
    __shared__ uint64_t buf[TILE];
    __shared__ uint32_t wcnt[NW][256];
    __shared__ uint32_t hs[256];
    __shared__ uint32_t sred[NW];
    __shared__ uint32_t s_or, s_and;
    const uint32_t nlist = *pnlist;
  for (uint32_t blk = blockIdx.x; blk < nlist; blk += gridDim.x) {
    const uint32_t c = list[blk];
    const uint32_t s0 = segs[c], cnt = segs[c + 1] - s0;
    const uint32_t w = threadIdx.x / WAVE, l = lane_id();
    if (threadIdx.x == 0) { s_or = 0u; s_and = 0xFFFFFFFFu; }
    __syncthreads();
    uint64_t e[TILE_ITEMS];
    uint32_t lor = 0u, land = 0xFFFFFFFFu;
    {
        // every item loaded first (past the end: the last item, dropped), so
        // the loads are in flight together
        uint32_t kv[TILE_ITEMS], vv[TILE_ITEMS];
#pragma unroll
        for (int i = 0; i < TILE_ITEMS; ++i) {
            const uint32_t p = min(w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l, cnt - 1u);
            kv[i] = key[s0 + p];
            vv[i] = val[s0 + p];
        }
#pragma unroll
        for (int i = 0; i < TILE_ITEMS; ++i) {
            const bool ok = w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l < cnt;
            const uint32_t k = ok ? kv[i] : 0u;
            e[i] = ((uint64_t)k << 32) | (ok ? vv[i] : 0u);
            if (ok) { lor |= k; land &= k; }
        }
    }
    atomicOr(&s_or, lor);
    atomicAnd(&s_and, land);
    __syncthreads();
    const uint32_t bits = s_or ^ s_and;   // key bits that vary inside the segment
    uint32_t pos[TILE_ITEMS];
    for (int pass = 0; pass < 4; ++pass) {
        if (((bits >> (8 * pass)) & 255u) == 0u) continue;     // a constant digit keeps the order
        rank8(e, cnt, 8 * pass, wcnt, hs, sred, pos);
#pragma unroll
        for (int i = 0; i < TILE_ITEMS; ++i)
            if (w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l < cnt) buf[pos[i]] = e[i];
        __syncthreads();
#pragma unroll
        for (int i = 0; i < TILE_ITEMS; ++i) {
            const uint32_t p = w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l;
            if (p < cnt) e[i] = buf[p];
        }
    }
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; ++i) {
        const uint32_t p = w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l;
        if (p < cnt) {
            key[s0 + p] = (uint32_t)(e[i] >> 32);
            val[s0 + p] = (uint32_t)e[i];
        }
    }
    __syncthreads();                  // s_or / s_and and buf are reused by the next segment
  }
}

// Stable rank of NT * ITEMS items by an 8-bit digit per item: the item
// layout and the ballot ranking of rank8 (item i of lane l of wave w is
// position w*WAVE*ITEMS + i*WAVE + l, valid below cnt), with the digits given
// by the caller.  Block-wide; ends with a barrier.
template <int NT, int ITEMS>
__device__ __forceinline__ void rank_digits(const uint32_t (&dg)[ITEMS], uint32_t cnt, uint32_t (*wcnt)[256],
                                            uint32_t* hs, uint32_t* scratch, uint32_t (&pos)[ITEMS]) {
    constexpr uint32_t NW = NT / WAVE;
    static_assert(NT >= 256, "one thread per digit");
    const uint32_t w = threadIdx.x / WAVE, l = lane_id();
    for (uint32_t q = threadIdx.x; q < NW * 256u; q += NT) (&wcnt[0][0])[q] = 0u;
    __syncthreads();
    const uint64_t lt = (1ull << l) - 1ull;
    uint32_t rk[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t p = w * (WAVE * ITEMS) + (uint32_t)i * WAVE + l;
        const bool ok = p < cnt;
        const uint32_t d = dg[i];
        uint64_t peer = __ballot(ok);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool one = (d >> bit) & 1u;
            const uint64_t bb = __ballot(one);
            peer &= one ? bb : ~bb;
        }
        const bool lead = ok && (peer & lt) == 0ull;
        const uint32_t leader = ok ? (uint32_t)__builtin_ctzll(peer) : l;
        const uint32_t c0 = lead ? wcnt[w][d] : 0u;
        const uint32_t c = __shfl(c0, (int)leader);
        if (lead) wcnt[w][d] = c0 + (uint32_t)__popcll(peer);
        rk[i] = c + (uint32_t)__popcll(peer & lt);
    }
    __syncthreads();
    uint32_t run = 0;
    if (threadIdx.x < 256u) {
        const uint32_t d = threadIdx.x;   // wcnt[w][d] becomes the count of digit d in waves < w
#pragma unroll
        for (uint32_t ww = 0; ww < NW; ++ww) {
            const uint32_t c = wcnt[ww][d];
            wcnt[ww][d] = run;
            run += c;
        }
    }
    uint32_t tot;
    const uint32_t ex = block_excl_scan<NT>(run, scratch, tot);
    if (threadIdx.x < 256u) hs[threadIdx.x] = ex;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) pos[i] = hs[dg[i] & 255u] + wcnt[w][dg[i] & 255u] + rk[i];
    __syncthreads();
}

// Segmented stable sort of up to NT * ITEMS items [segs[0], segs[nseg]) by one
// workgroup of NT threads: every item gets a label that grows with its segment
// (the number of segment starts at or before it), then LSD passes over the key
// digits that vary and over the label digits sort by (label, key); stability
// keeps key ties in input order.  Workspace in LDS: buf and lab of NT * ITEMS
// entries, wcnt[NT / 64][256], hs[256], sred[NT / 64], sbits[2].
template <int NT, int ITEMS>
__device__ __forceinline__ void wg_seg_sort(uint32_t nseg, const uint32_t* __restrict__ segs, uint32_t* __restrict__ key,
                            uint32_t* __restrict__ val, uint64_t* buf, uint32_t* lab, uint32_t (*wcnt)[256],
                            uint32_t* hs, uint32_t* sred, uint32_t* sbits) {
    constexpr uint32_t CAP = (uint32_t)NT * ITEMS;
    const uint32_t base = segs[0];
    const uint32_t cnt = min(segs[nseg] - base, CAP);   // the caller keeps it within the bound
    const uint32_t w = threadIdx.x / WAVE, l = lane_id();
    for (uint32_t q = threadIdx.x; q < CAP; q += NT) lab[q] = 0u;
    if (threadIdx.x == 0) { sbits[0] = 0u; sbits[1] = 0xFFFFFFFFu; }
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < nseg; c += NT) {
        const uint32_t st = segs[c] - base;
        if (st < cnt) atomicAdd(&lab[st], 1u);
    }
    __syncthreads();
    {   // inclusive scan of the start counts, ITEMS consecutive entries a thread
        const uint32_t q0 = threadIdx.x * ITEMS;
        uint32_t v[ITEMS], sum = 0;
#pragma unroll
        for (uint32_t i = 0; i < (uint32_t)ITEMS; ++i) { v[i] = lab[q0 + i]; sum += v[i]; }
        uint32_t tot;
        uint32_t run = block_excl_scan<NT>(sum, sred, tot);
#pragma unroll
        for (uint32_t i = 0; i < (uint32_t)ITEMS; ++i) { run += v[i]; lab[q0 + i] = run; }
    }
    __syncthreads();
    uint64_t e[ITEMS];
    uint32_t lb[ITEMS];
    uint32_t lor = 0u, land = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t p = w * (WAVE * ITEMS) + (uint32_t)i * WAVE + l;
        const bool ok = p < cnt;
        const uint32_t k = ok ? key[base + p] : 0u;
        e[i] = ((uint64_t)k << 32) | (ok ? val[base + p] : 0u);
        lb[i] = ok ? lab[p] : 0u;
        if (ok) { lor |= k; land &= k; }
    }
    atomicOr(&sbits[0], lor);
    atomicAnd(&sbits[1], land);
    __syncthreads();
    const uint32_t kbits = sbits[0] ^ sbits[1];               // key bits that vary
    const uint32_t lbits = nseg > 1u ? 32u - (uint32_t)__clz((int)nseg) : 0u;   // labels are <= nseg
    for (int pass = 0; pass < 8; ++pass) {
        const bool on_label = pass >= 4;
        const int sh = 8 * (pass & 3);
        if (!on_label && ((kbits >> sh) & 255u) == 0u) continue;   // a constant digit keeps the order
        if (on_label && (uint32_t)sh >= lbits) break;
        uint32_t dg[ITEMS], pos[ITEMS];
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
            dg[i] = on_label ? (lb[i] >> sh) & 255u : (uint32_t)(e[i] >> (32 + sh)) & 255u;
        rank_digits<NT, ITEMS>(dg, cnt, wcnt, hs, sred, pos);
#pragma unroll
        for (int i = 0; i < ITEMS; ++i)
            if (w * (WAVE * ITEMS) + (uint32_t)i * WAVE + l < cnt) {
                buf[pos[i]] = e[i];
                lab[pos[i]] = lb[i];
            }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < ITEMS; ++i) {
            const uint32_t p = w * (WAVE * ITEMS) + (uint32_t)i * WAVE + l;
            if (p < cnt) { e[i] = buf[p]; lb[i] = lab[p]; }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
        const uint32_t p = w * (WAVE * ITEMS) + (uint32_t)i * WAVE + l;
        if (p < cnt) {
            key[base + p] = (uint32_t)(e[i] >> 32);
            val[base + p] = (uint32_t)e[i];
        }
    }
    __syncthreads();
}

// The segmented sort of up to SORT_SMALL_N items in one launch (instead of
// the size classes and the tile passes).
constexpr uint32_t SS_NT = 512, SS_ITEMS = 8;
constexpr uint32_t SORT_SMALL_N = SS_NT * SS_ITEMS;
constexpr uint32_t SORT_SMALL_USE = SORT_SMALL_N;   // sorts of up to this many items take k_sort_small
__global__ void __launch_bounds__(SS_NT) k_sort_small(uint32_t nseg, const uint32_t* __restrict__ segs,
                                                      uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
    constexpr uint32_t NW = SS_NT / WAVE;
    __shared__ uint64_t buf[SORT_SMALL_N];
    __shared__ uint32_t lab[SORT_SMALL_N];
    __shared__ uint32_t wcnt[NW][256];
    __shared__ uint32_t hs[256];
    __shared__ uint32_t sred[NW];
    __shared__ uint32_t sbits[2];
    wg_seg_sort<SS_NT, SS_ITEMS>(nseg, segs, key, val, buf, lab, wcnt, hs, sred, sbits);
}

// Large segments (more than TILE items): LSD radix sort over the four 8-bit
// digits, per digit a histogram of every tile, an exclusive scan of the
// histograms (segment by segment, digit-major) and the scatter.  The tile
// plan is built on the device.
struct LargeSeg {
    uint32_t start, len, ntiles, hbase;
};

// plan, first pass: start, length and tile count of every large segment (the
// list's length read on the device)
__global__ void k_large_plan(const uint32_t* __restrict__ pnlarge, const uint32_t* __restrict__ llarge,
                             const uint32_t* __restrict__ segs, LargeSeg* __restrict__ ls,
                             uint32_t* __restrict__ tcount) {
    const uint32_t nl = *pnlarge;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nl; i += gridDim.x * blockDim.x) {
        const uint32_t c = llarge[i];
        const uint32_t st = segs[c], len = segs[c + 1] - st;
        const uint32_t nt = (len + TILE - 1) / TILE;
        ls[i] = LargeSeg{st, len, nt, 0u};
        tcount[i] = nt;
    }
}

// second pass, after the tile counts are scanned (toff, toff[nl] = tiles):
// every segment's histogram base, and the segment and index of every tile
// (the segment by binary search over toff)
__global__ void k_large_expand(const uint32_t* __restrict__ pnlarge, const uint32_t* __restrict__ toff,
                               LargeSeg* __restrict__ ls, uint32_t* __restrict__ tseg, uint32_t* __restrict__ tidx,
                               Counters* cnt) {
    const uint32_t nl = *pnlarge;
    const uint32_t T = toff[nl];
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt->tiles256 = T * 256u;   // length of the histogram scans
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < T; j += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nl;     // last i with toff[i] <= j
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (toff[mid] <= j) lo = mid;
            else hi = mid;
        }
        tseg[j] = lo;
        tidx[j] = j - toff[lo];
        if (j == toff[lo]) ls[lo].hbase = j * 256u;
    }
}

__global__ void __launch_bounds__(BS) k_tile_hist(const uint32_t* __restrict__ pntiles, const uint32_t* __restrict__ tseg,
                                                  const uint32_t* __restrict__ tidx,
                                                  const LargeSeg* __restrict__ ls,
                                                  const uint32_t* __restrict__ key, int shift,
                                                  uint32_t* __restrict__ hist) {
    __shared__ uint32_t h[256];
    // XCD x takes the x-th eighth of the tiles (a segment's tiles are
    // consecutive, so its lines stay in one L2); the grid is a multiple of 8
    const XcdSpan sp = xcd_span(*pntiles, 1, 0);
  for (uint32_t ti = sp.first; ti < sp.end; ti += sp.stride) {
    h[threadIdx.x] = 0u;
    __syncthreads();
    const LargeSeg L = ls[tseg[ti]];
    const uint32_t t = tidx[ti];
    const uint32_t b = L.start + t * TILE;
    const uint32_t e = min(L.start + L.len, b + (uint32_t)TILE);
    // a wave whose lanes share one digit adds its count with one atomic: the
    // regroup's keys come in runs (a cluster's positions share its minimum),
    // and 64 lanes on one LDS counter would serialise
    uint32_t kv[TILE / BS];           // the tile's keys first, their loads in flight together
#pragma unroll
    for (uint32_t q = 0; q < (uint32_t)(TILE / BS); ++q) kv[q] = key[min(b + q * BS + threadIdx.x, e - 1u)];
#pragma unroll
    for (uint32_t q = 0; q < (uint32_t)(TILE / BS); ++q) {
        const uint32_t p = b + q * BS + threadIdx.x;
        const bool ok = p < e;
        const uint32_t d = ok ? (kv[q] >> shift) & 255u : 256u;
        const uint32_t d0 = __shfl(d, 0);
        const uint64_t same = __ballot(ok && d == d0);
        if (same == __ballot(ok)) {
            if (lane_id() == 0 && d0 < 256u) atomicAdd(&h[d0], (uint32_t)__popcll(same));
        } else if (ok) {
            atomicAdd(&h[d], 1u);
        }
    }
    __syncthreads();
    hist[L.hbase + threadIdx.x * L.ntiles + t] = h[threadIdx.x];
  }
}

// One LSD pass over one tile (rank8), then the scatter to the digit's global
// offset: items are staged in LDS at their rank and written out in tile order.
__global__ void __launch_bounds__(BS) k_tile_scatter(const uint32_t* __restrict__ pntiles, const uint32_t* __restrict__ tseg,
                                                     const uint32_t* __restrict__ tidx,
                                                     const LargeSeg* __restrict__ ls,
                                                     const uint32_t* __restrict__ hscan,
                                                     const uint32_t* __restrict__ ksrc,
                                                     const uint32_t* __restrict__ vsrc, int shift,
                                                     uint32_t* __restrict__ kdst, uint32_t* __restrict__ vdst) {
    constexpr uint32_t NW = BS / WAVE;
    __shared__ uint64_t buf[TILE];
    __shared__ uint32_t wcnt[NW][256];
    __shared__ uint32_t hs[256], goff[256];
    __shared__ uint32_t sred[NW];
    // tiles split per XCD as in k_tile_hist; the grid is a multiple of 8
    const XcdSpan sp = xcd_span(*pntiles, 1, 0);
  for (uint32_t ti = sp.first; ti < sp.end; ti += sp.stride) {
    const LargeSeg L = ls[tseg[ti]];
    const uint32_t t = tidx[ti];
    const uint32_t b = L.start + t * TILE;
    const uint32_t cnt = min(L.start + L.len, b + (uint32_t)TILE) - b;
    const uint32_t w = threadIdx.x / WAVE, l = lane_id();
    goff[threadIdx.x] = hscan[L.hbase + threadIdx.x * L.ntiles + t] - hscan[L.hbase] + L.start;
    uint64_t e[TILE_ITEMS];
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; ++i) {
        const uint32_t p = w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l;
        const bool ok = p < cnt;
        e[i] = ok ? (((uint64_t)ksrc[b + p] << 32) | vsrc[b + p]) : 0ull;
    }
    uint32_t pos[TILE_ITEMS];
    rank8(e, cnt, shift, wcnt, hs, sred, pos);
#pragma unroll
    for (int i = 0; i < TILE_ITEMS; ++i)
        if (w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l < cnt) buf[pos[i]] = e[i];
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < cnt; p += BS) {
        const uint64_t x = buf[p];
        const uint32_t d = (uint32_t)(x >> (32 + shift)) & 255u;
        const uint32_t dst = goff[d] + (p - hs[d]);
        kdst[dst] = (uint32_t)(x >> 32);
        vdst[dst] = (uint32_t)x;
    }
    __syncthreads();                  // buf, goff and hs are rewritten for the next tile
  }
}

// ------------------------------------------------------------------ sweep scan

// Phase A: per block of 1024 positions, the (edgeDiff, volume) total of the
// block's last segment and whether the block contains a segment start.
__global__ void __launch_bounds__(BS) k_sweep_a(uint32_t na, const uint32_t* __restrict__ pnix,
                                                const uint32_t* __restrict__ pclu,
                                                const uint32_t* __restrict__ cs,
                                                const uint2* __restrict__ edv,
                                                uint32_t* __restrict__ ped, uint32_t* __restrict__ pvol,
                                                uint32_t* __restrict__ tail_ed, uint32_t* __restrict__ tail_vol,
                                                uint32_t* __restrict__ has_start) {
    __shared__ uint32_t s_ed, s_vol, s_start;
    const uint32_t tb = xcd_block();          // this block's tile
    const uint32_t b0 = tb * (BS * SCAN_ITEMS);
    const uint32_t b1 = min(na, b0 + BS * SCAN_ITEMS);
    if (threadIdx.x == 0) { s_ed = 0u; s_vol = 0u; s_start = 0u; }
    __syncthreads();
    const uint32_t lastc = pclu[b1 - 1];
    uint32_t ed = 0u, vol = 0u, st = 0u;
    // every item's cluster and node, then their cluster starts and node values
    // (loads in flight together; items past the end read the last position)
    uint32_t c[SCAN_ITEMS], u[SCAN_ITEMS], cst[SCAN_ITEMS];
    uint2 ev[SCAN_ITEMS];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const uint32_t k = min(b0 + threadIdx.x * SCAN_ITEMS + i, b1 - 1u);
        c[i] = pclu[k];
        u[i] = pnix[k];
    }
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        cst[i] = cs[c[i]];
        ev[i] = edv[u[i]];            // the cut contribution and dInt, one gather
    }
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const uint32_t k = b0 + threadIdx.x * SCAN_ITEMS + i;
        if (k >= b1) break;
        if (cst[i] == k) st = 1u;
        // the node's values gathered once, kept in position order for k_sweep_c
        const uint32_t e = ev[i].x, d = ev[i].y;
        ped[k] = e;
        pvol[k] = d;
        if (c[i] == lastc) {
            ed += e;
            vol += d;
        }
    }
    ed = group_sum<WAVE>(ed);
    vol = group_sum<WAVE>(vol);
    st = group_or<WAVE>(st);
    if (lane_id() == 0) {
        atomicAdd(&s_ed, ed);
        atomicAdd(&s_vol, vol);
        atomicOr(&s_start, st);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        tail_ed[tb] = s_ed;
        tail_vol[tb] = s_vol;
        has_start[tb] = s_start;
    }
}

// Phase B: carry into every block = inclusive segmented sum of the previous
// blocks' tails (restart after a block with a start).  One block of 1024.
__global__ void __launch_bounds__(1024) k_sweep_b(uint32_t nb, const uint32_t* __restrict__ tail_ed,
                                                  const uint32_t* __restrict__ tail_vol,
                                                  const uint32_t* __restrict__ has_start,
                                                  uint32_t* __restrict__ carry_ed,
                                                  uint32_t* __restrict__ carry_vol) {
    __shared__ uint32_t f[1024], e[1024], v[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t lo = t * per, hi = min(nb, lo + per);
    uint32_t lf = 0u, le = 0u, lv = 0u;
    for (uint32_t i = lo; i < hi; ++i) {
        if (has_start[i]) { lf = 1u; le = tail_ed[i]; lv = tail_vol[i]; }
        else { le += tail_ed[i]; lv += tail_vol[i]; }
    }
    f[t] = lf; e[t] = le; v[t] = lv;
    __syncthreads();
    // inclusive segmented Hillis-Steele scan over the 1024 chunk aggregates
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        uint32_t pf = 0u, pe = 0u, pv = 0u;
        const bool take = t >= off;
        if (take) { pf = f[t - off]; pe = e[t - off]; pv = v[t - off]; }
        __syncthreads();
        if (take && !f[t]) { e[t] += pe; v[t] += pv; f[t] = pf; }
        __syncthreads();
    }
    uint32_t ce = t ? e[t - 1] : 0u, cv = t ? v[t - 1] : 0u;
    for (uint32_t i = lo; i < hi; ++i) {
        carry_ed[i] = ce;
        carry_vol[i] = cv;
        if (has_start[i]) { ce = tail_ed[i]; cv = tail_vol[i]; }
        else { ce += tail_ed[i]; cv += tail_vol[i]; }
    }
}

// Phase C: inclusive segmented prefix of (edgeDiff, volume, 1) per position,
// conductance (cuda/devSweepCut.h:403-432) and per-cluster rightmost argmin
// packed as (ordered sparsity << 32) | ~offset for atomicMin.
__global__ void __launch_bounds__(BS) k_sweep_c(uint32_t na, const uint32_t* __restrict__ pnix,
                                                const uint32_t* __restrict__ pclu,
                                                const uint32_t* __restrict__ cs,
                                                const uint32_t* __restrict__ cvol,
                                                const uint32_t* __restrict__ ped,
                                                const uint32_t* __restrict__ pvol,
                                                const uint32_t* __restrict__ carry_ed,
                                                const uint32_t* __restrict__ carry_vol,
                                                unsigned long long* __restrict__ cbest,
                                                pxd_node_data* __restrict__ nd) {
    __shared__ uint32_t f[BS], e[BS], v[BS];
    const uint32_t tb = xcd_block();          // this block's tile
    const uint32_t b0 = tb * (BS * SCAN_ITEMS);
    const uint32_t t = threadIdx.x;
    uint32_t ed[SCAN_ITEMS], vo[SCAN_ITEMS], cc[SCAN_ITEMS], uu[SCAN_ITEMS];
    bool st[SCAN_ITEMS], ok[SCAN_ITEMS];
    uint32_t lf = 0u, le = 0u, lv = 0u;
    // every item's values, then their clusters' starts, before any use (loads
    // in flight together; items past the end read the last position, na >= 1)
    uint32_t cst[SCAN_ITEMS];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const uint32_t k = min(b0 + t * SCAN_ITEMS + i, na - 1u);
        cc[i] = pclu[k];
        uu[i] = nd ? pnix[k] : 0u;
        ed[i] = ped[k];
        vo[i] = pvol[k];
    }
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) cst[i] = cs[cc[i]];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        const uint32_t k = b0 + t * SCAN_ITEMS + i;
        ok[i] = k < na;
        if (ok[i]) {
            st[i] = cst[i] == k;
        } else {
            cc[i] = 0u; uu[i] = 0u; st[i] = true; ed[i] = 0u; vo[i] = 0u;
        }
        if (st[i]) { lf = 1u; le = ed[i]; lv = vo[i]; }
        else { le += ed[i]; lv += vo[i]; }
    }
    f[t] = lf; e[t] = le; v[t] = lv;
    __syncthreads();
    for (uint32_t off = 1; off < BS; off <<= 1) {
        uint32_t pf = 0u, pe = 0u, pv = 0u;
        const bool take = t >= off;
        if (take) { pf = f[t - off]; pe = e[t - off]; pv = v[t - off]; }
        __syncthreads();
        if (take && !f[t]) { e[t] += pe; v[t] += pv; f[t] = pf; }
        __syncthreads();
    }
    // exclusive prefix for this thread's first item, carrying across blocks
    uint32_t re, rv;
    bool open;                         // prefix still continues a segment from before the block
    if (t == 0) { re = 0u; rv = 0u; open = true; }
    else { re = e[t - 1]; rv = v[t - 1]; open = !f[t - 1]; }
    if (open) { re += carry_ed[tb]; rv += carry_vol[tb]; }
    uint32_t runc = 0xFFFFFFFFu;
    unsigned long long runmin = ~0ull;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        if (st[i]) { re = ed[i]; rv = vo[i]; }
        else { re += ed[i]; rv += vo[i]; }
        if (!ok[i]) continue;
        const uint32_t k = b0 + t * SCAN_ITEMS + i;
        const uint32_t c = cc[i];
        const uint32_t off = k - cs[c] + 1u;
        const uint32_t tot = cvol[c];
        const uint32_t denom = (rv >= tot) ? 0u : ((rv < tot - rv) ? rv : (tot - rv));
        const float sp = denom > 0u ? ((float)re / (float)denom) : 2.0f;
        const unsigned long long pk = ((unsigned long long)ord_f32(sp) << 32) | (0xFFFFFFFFu - off);
        if (c != runc) {
            if (runc != 0xFFFFFFFFu) atomicMin(&cbest[runc], runmin);
            runc = c;
            runmin = pk;
        } else if (pk < runmin) {
            runmin = pk;
        }
        if (nd) {
            pxd_node_data o;
            o.nix = uu[i]; o.prefixVolume = rv; o.prefixEdgeDiff = (int32_t)re; o.offsetInCluster = off;
            nd[k] = o;
        }
    }
    // the last runs: one atomic per workgroup when all of them are one cluster
    // (a giant cluster would otherwise take one atomic per wave on one
    // address), else one per wave when the wave's are, else one per thread
    __shared__ unsigned long long s_min[BS / WAVE];
    __shared__ uint32_t s_c[BS / WAVE];
    const uint32_t c0 = __shfl(runc, 0);
    const bool wu = __all(runc == c0) && c0 != 0xFFFFFFFFu;
    if (wu) {
        for (int offs = WAVE / 2; offs > 0; offs >>= 1) {
            const unsigned long long o = __shfl_xor(runmin, offs);
            runmin = o < runmin ? o : runmin;
        }
    }
    if (lane_id() == 0) {
        s_c[t / WAVE] = wu ? c0 : 0xFFFFFFFFu;
        s_min[t / WAVE] = runmin;
    }
    __syncthreads();
    bool bu = s_c[0] != 0xFFFFFFFFu;
    for (uint32_t w = 1; w < BS / WAVE; ++w) bu = bu && s_c[w] == s_c[0];
    if (bu) {
        if (t == 0) {
            unsigned long long m = s_min[0];
            for (uint32_t w = 1; w < BS / WAVE; ++w) m = s_min[w] < m ? s_min[w] : m;
            atomicMin(&cbest[s_c[0]], m);
        }
    } else if (wu) {
        if (lane_id() == 0) atomicMin(&cbest[c0], runmin);
    } else if (runc != 0xFFFFFFFFu) {
        atomicMin(&cbest[runc], runmin);
    }
}


// ------------------------------------------------------------------ cut decision

// cuda/devSweepCut.h:701-709 (threshold filter) for every cluster; flags the
// clusters that need connected components this iteration.
__global__ void k_decide(uint32_t nc, const unsigned long long* __restrict__ cbest, double thr,
                         uint32_t* __restrict__ cflag, uint32_t* __restrict__ coff,
                         float* __restrict__ csp, const uint32_t* __restrict__ cs, Counters* cnt) {
    for (uint32_t c0 = blockIdx.x * blockDim.x; c0 < nc; c0 += gridDim.x * blockDim.x) {
        const uint32_t c = c0 + threadIdx.x;
        uint32_t f = 0u;
        if (c < nc) {
            const unsigned long long b = cbest[c];
            const float sp = unord_f32((uint32_t)(b >> 32));
            const uint32_t off = 0xFFFFFFFFu - (uint32_t)b;
            csp[c] = sp;
            coff[c] = off;
            f = cflag[c] & CF_CONN;
            if ((double)sp < thr) f |= CF_CUT;
            if ((f & CF_CUT) || !(f & CF_CONN)) f |= CF_CC;
            cflag[c] = f;
        }
        // counts: one atomic per wave; the CC clusters' sizes add up to the
        // number of CC members of the cut
        const uint32_t ncut = (uint32_t)__popcll(__ballot(c < nc && (f & CF_CUT)));
        const uint32_t ncc = (uint32_t)__popcll(__ballot(c < nc && (f & CF_CC)));
        const uint32_t ccn = group_sum<WAVE>((c < nc && (f & CF_CC)) ? cs[c + 1] - cs[c] : 0u);
        if (lane_id() == 0) {
            if (ncut) atomicAdd(&cnt->ncut, ncut);
            if (ncc) atomicAdd(&cnt->ncc, ncc);
            if (ccn) atomicAdd(&cnt->ccn, ccn);
        }
    }
}

// cuda/devPartition.h:545-622: positions beyond the cut offset form the second
// part.  Members of CC clusters get a part id and this iteration's stamp (the
// CC list is built from the stamps in node order, and SelMembers then gives
// every member a fresh root, a write in node order instead of a scattered one).
__global__ void __launch_bounds__(BS) k_mark(uint32_t na, const uint32_t* __restrict__ pnix,
                                             const uint32_t* __restrict__ pclu,
                                             const uint32_t* __restrict__ cs, const uint32_t* __restrict__ cflag,
                                             const uint32_t* __restrict__ coff, uint32_t iter,
                                             uint32_t* __restrict__ part, uint32_t* __restrict__ root,
                                             uint32_t* __restrict__ stamp) {
    const XcdSpan sp = xcd_span(na, BS, threadIdx.x);
    for (uint32_t k = sp.first; k < sp.end; k += sp.stride) {
        const uint32_t c = pclu[k];
        const uint32_t f = cflag[c];
        if (f & CF_CC) {
            const uint32_t u = pnix[k];
            const uint32_t side = ((f & CF_CUT) && (k - cs[c] + 1u > coff[c])) ? 1u : 0u;
            part[u] = (c << 1) | side;
            stamp[u] = iter;          // SelMembers sets root[u] = u, in node order
        }
    }
}

// Connected components of every CC part (parent pointers `root`, members in
// `list`).  Invariant of every kernel below: a pointer only ever moves to a
// smaller node of the same component, so trees stay acyclic and the root of a
// tree is its minimum node.  Every write to `root` is atomic (or, in
// k_cc_jump, to the thread's own slot): threads on different XCDs write the
// same slots within one launch.
//
// Sampling, one round per arc index r (Afforest): for the r-th arc of every
// member, if both ends are in the same part, the slot of the larger endpoint
// pointer is lowered to the smaller (atomicMin).  A pointer-jumping launch
// between rounds makes most endpoints of the next round share a pointer, so
// later rounds issue few atomics.  Plain reads may be stale; a stale pointer
// is still a smaller node of the same component, so it only weakens a link.
// Round 0 reads the first CC_SAMPLE_ROUNDS arcs of every member (one line of
// its row) and keeps arcs 1.. in samp[(r - 1) * nl + i] (INVALID past the
// row's end), so that later rounds read them coalesced instead of a line of
// every row again.
constexpr uint32_t CC_SAMPLE_ROUNDS = 4;

// the sampled neighbour of list entry i in round r (INVALID: none)
__device__ __forceinline__ uint32_t cc_sample_arc(uint32_t r, uint32_t i, uint32_t u, uint32_t nl,
                                                  const uint32_t* __restrict__ row,
                                                  const uint32_t* __restrict__ dint,
                                                  const uint32_t* __restrict__ nbr, uint32_t* __restrict__ samp) {
    if (r != 0u) return samp[(r - 1u) * nl + i];
    const uint32_t d = dint[u], st = row[u];
    uint32_t v0 = INVALID;
    if (d > 0u) v0 = nbr[st];
#pragma unroll
    for (uint32_t k = 1; k < CC_SAMPLE_ROUNDS; ++k) samp[(k - 1u) * nl + i] = d > k ? nbr[st + k] : INVALID;
    return v0;
}

__global__ void __launch_bounds__(BS) k_cc_sample(uint32_t nl, const uint32_t* __restrict__ list,
                                                  const uint32_t* __restrict__ row,
                                                  const uint32_t* __restrict__ dint,
                                                  const uint32_t* __restrict__ nbr,
                                                  const uint32_t* __restrict__ part,
                                                  uint32_t* __restrict__ root, uint32_t r,
                                                  uint32_t* __restrict__ samp) {
    const XcdSpan sp = xcd_span(nl, BS, threadIdx.x);
    for (uint32_t i = sp.first; i < sp.end; i += sp.stride) {
        const uint32_t u = list[i];
        const uint32_t v = cc_sample_arc(r, i, u, nl, row, dint, nbr, samp);
        if (v == INVALID) continue;
        if (part[v] != part[u]) continue;
        const uint32_t a = root[u], b = root[v];
        if (a == b) continue;
        const uint32_t hi = max(a, b), lo = min(a, b);
        if (root[hi] > lo) atomicMin(&root[hi], lo);
    }
}

// k_cc_sample with an LDS window of the parts and pointers (the layout of
// k_walk_win: chunks of CH entries of the member list, WIN values around a
// chunk staged when its nodes span less than WIN/2).  The staged pointers are
// those after the preceding jump launch; a stale pointer is still a smaller
// node of the same component, so links stay valid.  The slot lowered is
// checked against its current value, as in k_cc_sample.
template <int WIN, int CH>
__global__ void __launch_bounds__(WBS) k_cc_sample_win(uint32_t nl, const uint32_t* __restrict__ list,
                                                       const uint32_t* __restrict__ row,
                                                       const uint32_t* __restrict__ dint,
                                                       const uint32_t* __restrict__ nbr,
                                                       const uint32_t* __restrict__ part,
                                                       uint32_t* __restrict__ root, uint32_t r, uint32_t N,
                                                       uint32_t* __restrict__ samp) {
    __shared__ uint32_t sp_[WIN], sr_[WIN];
    const uint32_t nch = (nl + CH - 1) / CH;
    const uint32_t x = blockIdx.x % NXCD, lb = blockIdx.x / NXCD, nb = gridDim.x / NXCD;
    const uint32_t cpx = nch / NXCD + (nch % NXCD != 0);
    const uint32_t cb = min(nch, x * cpx), ce = min(nch, cb + cpx);
    for (uint32_t c = cb + lb; c < ce; c += nb) {
        const uint32_t c0 = c * CH, c1 = min(c0 + CH, nl);
        const uint32_t uf = list[c0], ul = list[c1 - 1];
        const bool use = ul >= uf && ul - uf < (uint32_t)WIN / 2;
        uint32_t wb = 0;
        if (use) {
            const uint32_t mid = uf / 2 + ul / 2;
            wb = mid > (uint32_t)WIN / 2 ? mid - (uint32_t)WIN / 2 : 0u;
            if (N <= (uint32_t)WIN) wb = 0;
            else if (wb > N - (uint32_t)WIN) wb = N - (uint32_t)WIN;
            for (uint32_t k = threadIdx.x; k < (uint32_t)WIN; k += WBS) {
                if (wb + k < N) {
                    sp_[k] = part[wb + k];
                    sr_[k] = ld_dev(&root[wb + k]);
                }
            }
        }
        __syncthreads();
        // SI members per thread at a time, each step's loads in flight together
        constexpr int SI = 4;
        for (uint32_t i0 = c0 + threadIdx.x; i0 < c1; i0 += SI * WBS) {
            uint32_t u[SI], v[SI], pu[SI], pv[SI], a[SI], b[SI];
            bool ok[SI];
#pragma unroll
            for (int k = 0; k < SI; ++k) {
                ok[k] = i0 + k * WBS < c1;
                u[k] = list[ok[k] ? i0 + k * WBS : c0];
            }
            if (r == 0u) {
                // every member's row bounds, then all their sampled arcs, then the
                // kept ones out (loads in flight together, as cc_sample_arc does
                // for one member)
                uint32_t dd[SI], rw[SI], a[SI][CC_SAMPLE_ROUNDS];
#pragma unroll
                for (int k = 0; k < SI; ++k) {
                    dd[k] = ok[k] ? dint[u[k]] : 0u;
                    rw[k] = row[u[k]];
                }
#pragma unroll
                for (int k = 0; k < SI; ++k)
#pragma unroll
                    for (uint32_t j = 0; j < CC_SAMPLE_ROUNDS; ++j) a[k][j] = dd[k] > j ? nbr[rw[k] + j] : INVALID;
#pragma unroll
                for (int k = 0; k < SI; ++k) {
                    if (ok[k]) {
#pragma unroll
                        for (uint32_t j = 1; j < CC_SAMPLE_ROUNDS; ++j) samp[(j - 1u) * nl + i0 + k * WBS] = a[k][j];
                    }
                    v[k] = ok[k] ? a[k][0] : INVALID;
                }
            } else {
#pragma unroll
                for (int k = 0; k < SI; ++k) v[k] = ok[k] ? samp[(r - 1u) * nl + i0 + k * WBS] : INVALID;
            }
#pragma unroll
            for (int k = 0; k < SI; ++k) {
                ok[k] = v[k] != INVALID;
                if (!ok[k]) v[k] = u[k];
            }
            win_gather<SI, false>(pu, u, ok, sp_, part, wb, (uint32_t)WIN, use);
            win_gather<SI, false>(pv, v, ok, sp_, part, wb, (uint32_t)WIN, use);
#pragma unroll
            for (int k = 0; k < SI; ++k) ok[k] = ok[k] && pu[k] == pv[k];
            win_gather<SI, false>(a, u, ok, sr_, (const uint32_t*)root, wb, (uint32_t)WIN, use);
            win_gather<SI, false>(b, v, ok, sr_, (const uint32_t*)root, wb, (uint32_t)WIN, use);
            uint32_t cur[SI], hi[SI], lo[SI];
#pragma unroll
            for (int k = 0; k < SI; ++k) {
                ok[k] = ok[k] && a[k] != b[k];
                hi[k] = max(a[k], b[k]);
                lo[k] = min(a[k], b[k]);
                cur[k] = ok[k] ? root[hi[k]] : 0u;
            }
#pragma unroll
            for (int k = 0; k < SI; ++k)
                if (ok[k] && cur[k] > lo[k]) atomicMin(&root[hi[k]], lo[k]);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint32_t cc_find(uint32_t* root, uint32_t r) {
    for (;;) {                        // strictly decreasing: ends at a root
        const uint32_t p = ld_dev(&root[r]);
        if (p == r) return r;
        r = p;
    }
}

// find with plain (cached) loads, for launches in which no root moves: the
// only pointer writes are a thread's own slot set to that slot's root, so
// an old value read from a cache is still on the same path to the same root
__device__ __forceinline__ uint32_t cc_find_stable(const uint32_t* __restrict__ root, uint32_t r) {
    for (;;) {
        const uint32_t p = root[r];
        if (p == r) return r;
        r = p;
    }
}

// Giant root of every part after sampling: the root of the part's middle
// member (by position).  Members already in that tree skip the exact phase.
__global__ void k_cc_giant(uint32_t nc, const uint32_t* __restrict__ cs, const uint32_t* __restrict__ cflag,
                           const uint32_t* __restrict__ coff, const uint32_t* __restrict__ pnix,
                           uint32_t* __restrict__ root, uint32_t* __restrict__ giant, Counters* cnt) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {   // the exact rounds that follow start unconverged
        cnt->cc_done = 0u;
        cnt->link_changed = 0u;
    }
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        const uint32_t f = cflag[c];
        if (!(f & CF_CC)) continue;
        const uint32_t b = cs[c], len = cs[c + 1] - b;
        const uint32_t o = (f & CF_CUT) ? min(coff[c], len) : len;
        giant[c << 1] = o ? cc_find_stable(root, pnix[b + o / 2]) : INVALID;
        giant[(c << 1) | 1] = (len > o) ? cc_find_stable(root, pnix[b + o + (len - o) / 2]) : INVALID;
    }
}

// a member outside its part's giant tree
__device__ __forceinline__ bool SelOffGiant::flag(uint32_t, const D& d) const {
    return (d.p == d.u ? d.u : cc_find_stable(root, d.p)) != d.g;
}

// Exact rounds over the members outside the giant trees, GS lanes per node:
// for every same-part arc (u, v) with find(u) != find(v) the larger root's
// slot is lowered to the smaller root.  A round counts as changed only when an
// atomic lowered a slot, so the sum of pointers falls in every changed round
// and the loop ends.  In a round without change every pointer is constant and
// every processed arc joins two nodes with the same root (else the larger root
// would have been lowered); arcs between giant members lie inside the giant's
// tree, and an arc with one giant end is processed from its other end.  Trees
// need not be flat: every root is found by following pointers.
template <int GS>
__global__ void __launch_bounds__(BS) k_cc_link(uint32_t nl, const uint32_t* __restrict__ list,
                                                const uint32_t* __restrict__ row,
                                                const uint32_t* __restrict__ dint,
                                                const uint32_t* __restrict__ nbr,
                                                const uint32_t* __restrict__ part,
                                                uint32_t* __restrict__ root, uint32_t heavy, Counters* cnt,
                                                const uint32_t* pn = nullptr, const uint32_t* stop = nullptr) {
    if (stop && *stop) return;        // the rounds have converged (uniform over the grid)
    nl = dev_count(nl, pn);
    const uint32_t lane = threadIdx.x & (GS - 1);
    const XcdSpan sp = xcd_span(nl, BS / GS, threadIdx.x / GS);
    uint32_t changed = 0;
    for (uint32_t i = sp.first; i < sp.end; i += sp.stride) {
        const uint32_t u = list[i];
        const uint32_t start = row[u], d = dint[u];
        if (d > heavy) continue;      // k_cc_link_heavy takes rows above the threshold
        const uint32_t pp = part[u];
        uint32_t ru = cc_find(root, u);
        for (uint32_t j = lane; j < d; j += GS) {
            const uint32_t v = nbr[start + j];
            if (part[v] != pp) continue;
            const uint32_t rv = cc_find(root, v);
            if (rv == ru) continue;
            const uint32_t hi = max(ru, rv), lo = min(ru, rv);
            if (ld_dev(&root[hi]) > lo && atomicMin(&root[hi], lo) > lo) {
                changed = 1;
                if (hi == ru) ru = lo;    // u's tree now hangs below lo
            }
        }
    }
    if (__any(changed) && lane_id() == 0 && __atomic_load_n(&cnt->link_changed, __ATOMIC_RELAXED) == 0u)
        atomicOr(&cnt->link_changed, 1u);
}

// the link rounds for rows above the threshold: one workgroup per chunk of
// HEAVY_CHUNK arcs (chunk map over the link list)
__global__ void __launch_bounds__(BS) k_cc_link_heavy(const uint32_t* __restrict__ nchunks,
                                                      const uint32_t* __restrict__ hrow,
                                                      const uint32_t* __restrict__ hoff,
                                                      const uint32_t* __restrict__ list,
                                                      const uint32_t* __restrict__ row,
                                                      const uint32_t* __restrict__ dint,
                                                      const uint32_t* __restrict__ nbr,
                                                      const uint32_t* __restrict__ part,
                                                      uint32_t* __restrict__ root, Counters* cnt,
                                                      const uint32_t* stop = nullptr) {
    if (stop && *stop) return;
    const uint32_t H = *nchunks;
    uint32_t changed = 0;
    for (uint32_t b = blockIdx.x; b < H; b += gridDim.x) {
        const uint32_t i = hrow[b];
        const uint32_t u = list[i];
        const uint32_t d = dint[u], pp = part[u];
        const uint32_t* p = nbr + row[u];
        const uint32_t j0 = (b - hoff[i]) * HEAVY_CHUNK;
        uint32_t ru = cc_find(root, u);
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t j = j0 + q * BS + threadIdx.x;
            if (j >= d) break;
            const uint32_t v = p[j];
            if (part[v] != pp) continue;
            const uint32_t rv = cc_find(root, v);
            if (rv == ru) continue;
            const uint32_t hi = max(ru, rv), lo = min(ru, rv);
            if (ld_dev(&root[hi]) > lo && atomicMin(&root[hi], lo) > lo) {
                changed = 1;
                if (hi == ru) ru = lo;
            }
        }
    }
    if (__any(changed) && lane_id() == 0 && __atomic_load_n(&cnt->link_changed, __ATOMIC_RELAXED) == 0u)
        atomicOr(&cnt->link_changed, 1u);
}

// pointer jumping, up to 8 steps per node and launch (own slot only)
__global__ void __launch_bounds__(BS) k_cc_jump(uint32_t nl, const uint32_t* __restrict__ list,
                                                uint32_t* __restrict__ root, const uint32_t* pn = nullptr,
                                                const uint32_t* stop = nullptr) {
    if (stop && *stop) return;
    nl = dev_count(nl, pn);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nl; i += gridDim.x * blockDim.x) {
        const uint32_t u = list[i];
        const uint32_t r0 = ld_dev(&root[u]);
        uint32_t r = r0;
        for (int s = 0; s < 8; ++s) {
            const uint32_t rr = ld_dev(&root[r]);
            if (rr == r) break;
            r = rr;
        }
        if (r != r0) st_dev(&root[u], r);
    }
}

// After the components have converged: every member's pointer set to its
// root, over the member list in node order (the first read of every chase and
// the write stream; the next hops mostly hit one cached root).  No root moves
// any more, so cached loads see pointers on the same path to the same root; a
// write sets the thread's own slot (agent scope), and a reader sees the old or
// the new pointer.  The trees end flat, for k_rid and arc disabling.
__global__ void k_cc_flatten(uint32_t nl, const uint32_t* __restrict__ list, uint32_t* __restrict__ root) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nl; i += gridDim.x * blockDim.x) {
        const uint32_t u = list[i];
        const uint32_t p = root[u];
        const uint32_t r = p == u ? u : cc_find_stable(root, p);
        if (r != p) st_dev(&root[u], r);
    }
}

// regroup key: component root for CC members (their pointers, flat after
// k_cc_flatten), the cluster's minimum node otherwise
__global__ void k_rid(uint32_t na, const uint32_t* __restrict__ pnix, const uint32_t* __restrict__ pclu,
                      const uint32_t* __restrict__ cflag, const uint32_t* __restrict__ cmin,
                      const uint32_t* __restrict__ root, uint32_t* __restrict__ key) {
    const XcdSpan sp = xcd_span(na, BS, threadIdx.x);
    for (uint32_t k = sp.first; k < sp.end; k += sp.stride) {
        const uint32_t c = pclu[k];
        key[k] = (cflag[c] & CF_CC) ? root[pnix[k]] : cmin[c];
    }
}

// ------------------------------------------------------------------ cluster statistics

// before the sweep: best cut of every cluster unset, and the counters the
// sweep and the cut fill cleared (all but the sticky error word)
__global__ void k_sweep_reset(uint32_t nc, unsigned long long* __restrict__ cbest, Counters* cnt) {
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) cbest[c] = ~0ull;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        uint32_t* w = reinterpret_cast<uint32_t*>(cnt);
        for (uint32_t i = 1; i < sizeof(Counters) / 4; ++i) w[i] = 0u;
    }
}

__device__ __forceinline__ int ceil_log2_u32(uint32_t k) {
    return k <= 1u ? 0 : 32 - __clz((int)(k - 1u));
}

// Per cluster before its sums: first cluster of every graph (gfirst); max /
// min of (float)x as ordered bits (-0 < +0) -- the positions of a cluster are
// sorted by that key, so they are its last and first keys; the fixed-point
// scale K (amax < 2^E, K = 62 - E - ceil(log2(count))); the sum cleared.
__global__ void k_stats_prep(uint32_t nc, const uint32_t* __restrict__ cs, const uint32_t* __restrict__ cgraph,
                             const uint32_t* __restrict__ pnix, const uint32_t* __restrict__ xkey,
                             uint32_t* __restrict__ gfirst, uint32_t* __restrict__ cmaxo,
                             uint32_t* __restrict__ cmino, int32_t* __restrict__ cK,
                             unsigned long long* __restrict__ csum, const uint32_t* pn = nullptr) {
    nc = dev_count(nc, pn);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        const uint32_t g = cgraph[c];
        if (c == 0 || g != cgraph[c - 1]) gfirst[g] = c;
        const uint32_t b = cs[c], e = cs[c + 1];
        const uint32_t lo = xkey[pnix[b]], hi = xkey[pnix[e - 1u]];
        cmino[c] = lo;
        cmaxo[c] = hi;
        const float amax = fmaxf(fabsf(unord_f32(hi)), fabsf(unord_f32(lo)));
        int E = 0;
        if (amax != 0.0f) frexpf(amax, &E);
        cK[c] = 62 - E - ceil_log2_u32(e - b);
        csum[c] = 0ull;
    }
}

// Per-cluster sums over the positions.  Clusters are contiguous position
// ranges, so a workgroup takes SUM_ITEMS * BS consecutive positions at a time:
// when the first and last are in one cluster, the whole range is, and it is
// summed over the workgroup and added with one atomic (a cluster of millions
// of positions would otherwise take one atomic per wave on one address);
// otherwise each wave adds per cluster run as before.  Integer sums: the
// order of the additions does not matter.  Four positions a thread: twice
// the workgroups of eight and half the atomics of two (C4: k_stats_sum 91 us
// and k_volume 73 us a launch, against 99 and 81 with eight).
// k_stats_sum also records every position's cluster by node (nclu, for the
// node-order apply pass), while it has the position's node and cluster.
constexpr uint32_t SUM_ITEMS = 4;
__global__ void __launch_bounds__(BS) k_stats_sum(uint32_t na, const uint32_t* __restrict__ pnix,
                                                  const uint32_t* __restrict__ pclu,
                                                  const double* __restrict__ x, const int32_t* __restrict__ cK,
                                                  unsigned long long* __restrict__ csum,
                                                  uint32_t* __restrict__ nclu) {
    __shared__ int64_t lds[BS / WAVE];
    constexpr uint32_t TR = BS * SUM_ITEMS;
    const XcdSpan sp = xcd_tiles(na, TR);
    for (uint32_t t = sp.first; t < sp.end; t += sp.stride) {
        const uint32_t r0 = t * TR;
        const uint32_t r1 = min(r0 + TR, na);
        const uint32_t cf = pclu[r0];
        // every item's node and cluster, then every item's value (loads in
        // flight together; items past the end read the last position)
        uint32_t u[SUM_ITEMS], ci[SUM_ITEMS];
        double xv[SUM_ITEMS];
#pragma unroll
        for (uint32_t i = 0; i < SUM_ITEMS; ++i) {
            const uint32_t k = min(r0 + i * BS + threadIdx.x, r1 - 1u);
            u[i] = pnix[k];
            ci[i] = pclu[k];
        }
#pragma unroll
        for (uint32_t i = 0; i < SUM_ITEMS; ++i) xv[i] = x[u[i]];
        if (cf == pclu[r1 - 1]) {
            const int32_t K = cK[cf];
            int64_t acc = 0;
#pragma unroll
            for (uint32_t i = 0; i < SUM_ITEMS; ++i) {
                const uint32_t k = r0 + i * BS + threadIdx.x;
                if (k < r1) {
                    acc += (int64_t)ldexp((double)(float)xv[i], K);
                    nclu[u[i]] = cf;
                }
            }
            acc = group_sum<WAVE>(acc);
            if (lane_id() == 0) lds[threadIdx.x / WAVE] = acc;
            __syncthreads();
            if (threadIdx.x == 0) {
                int64_t t = 0;
                for (uint32_t w = 0; w < BS / WAVE; ++w) t += lds[w];
                atomicAdd(&csum[cf], (unsigned long long)t);
            }
            __syncthreads();
            continue;
        }
        for (uint32_t i = 0; i < SUM_ITEMS; ++i) {
            const uint32_t k = r0 + i * BS + threadIdx.x;
            const bool ok = k < r1;
            const uint32_t c = ok ? ci[i] : 0xFFFFFFFFu;
            int64_t term = 0;
            if (ok) {
                term = (int64_t)ldexp((double)(float)xv[i], cK[c]);
                nclu[u[i]] = c;
            }
            const uint32_t c0 = __shfl(c, 0);
            if (c0 == 0xFFFFFFFFu) {
                continue;                      // whole wave past the end: nothing to add
            } else if (__all(c == c0 || c == 0xFFFFFFFFu)) {
                term = group_sum<WAVE>(term);
                if (lane_id() == 0) atomicAdd(&csum[c0], (unsigned long long)term);
            } else if (ok) {
                atomicAdd(&csum[c], (unsigned long long)term);
            }
        }
    }
}

// cuda/devPartition.h:173-184, 624-650: keep flag, average, FP32 sum
__global__ void k_stats_final(uint32_t nc, const uint32_t* __restrict__ cs, const uint32_t* __restrict__ cmaxo,
                              const uint32_t* __restrict__ cmino, const int32_t* __restrict__ cK,
                              const unsigned long long* __restrict__ csum, float rw_thr,
                              uint32_t* __restrict__ ckeep, double* __restrict__ cavg, const uint32_t* pn = nullptr) {
    nc = dev_count(nc, pn);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        const uint32_t count = cs[c + 1] - cs[c];
        const float rw_sum = (float)ldexp(i64_to_f64((int64_t)csum[c]), -cK[c]);
        const float pot = unord_f32(cmaxo[c]) - unord_f32(cmino[c]);
        ckeep[c] = (pot >= rw_thr && count >= 2u) ? 1u : 0u;
        cavg[c] = (double)rw_sum / (double)count;
    }
}

// labels: survivors dense per graph in cluster order, deactivated
// smallest - index - 1 (cuda/devPartition.h:686-698)
__global__ void k_labels(uint32_t nc, const uint32_t* __restrict__ cgraph, const uint32_t* __restrict__ gfirst,
                         const uint32_t* __restrict__ ckeep, const uint32_t* __restrict__ kscan,
                         const int32_t* __restrict__ gsmall, int32_t* __restrict__ gsmall_next,
                         uint32_t* __restrict__ gsurv, int32_t* __restrict__ clabel, uint32_t* __restrict__ gseen,
                         const uint32_t* pn = nullptr) {
    nc = dev_count(nc, pn);
    // a graph's clusters are contiguous: the per-graph survivor count and
    // smallest new label are reduced over the wave first (one atomic per graph
    // run in the wave instead of one per cluster on a single address)
    for (uint32_t c0 = blockIdx.x * blockDim.x; c0 < nc; c0 += gridDim.x * blockDim.x) {
        const uint32_t c = c0 + threadIdx.x;
        const bool ok = c < nc;
        const uint32_t g = ok ? cgraph[c] : 0xFFFFFFFFu;
        uint32_t surv = 0u;
        int32_t lab = 0x7FFFFFFF;
        if (ok) {
            const uint32_t f = gfirst[g];
            if (c == f) gseen[g] = 1u;    // the graph has clusters this iteration
            if (ckeep[c]) {
                clabel[c] = (int32_t)(kscan[c] - kscan[f]);
                surv = 1u;
            } else {
                lab = gsmall[g] - (int32_t)(c - f) - 1;
                clabel[c] = lab;
            }
        }
        const uint32_t g0 = __shfl(g, 0);
        if (__all(g == g0 || g == 0xFFFFFFFFu)) {
            if (g0 == 0xFFFFFFFFu) continue;
            surv = group_sum<WAVE>(surv);
            for (int off = WAVE / 2; off > 0; off >>= 1) lab = min(lab, __shfl_xor(lab, off));
            if (lane_id() == 0) {
                if (surv) atomicAdd(&gsurv[g0], surv);
                if (lab != 0x7FFFFFFF) atomicMin(&gsmall_next[g0], lab);
            }
        } else if (ok) {
            if (surv) atomicAdd(&gsurv[g], 1u);
            if (lab != 0x7FFFFFFF) atomicMin(&gsmall_next[g], lab);
        }
    }
}

// iteration counts relative to the iteration a graph joined at (gstart)
__global__ void k_graph_update(uint32_t G, const uint32_t* __restrict__ gactive, int32_t* __restrict__ gsmall,
                               int32_t* __restrict__ gsmall_next, uint32_t* __restrict__ gsurv,
                               int32_t* __restrict__ giter, const uint32_t* __restrict__ gstart, uint32_t iter1) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += gridDim.x * blockDim.x) {
        if (gactive[g]) {
            gsmall[g] = gsmall_next[g];
            if (gsurv[g] == 0u && giter[g] == 0) giter[g] = (int32_t)(iter1 - gstart[g]) + 1;
        }
        gsmall_next[g] = 0;
        gsurv[g] = 0u;
    }
}

// per position: the cluster of its node (nclu[pnix[k]] = pclu[k])
// per active node, in index order: write labels, recentre survivors
// (x -= avg), refresh y; deactivated nodes lose `alive`
__global__ void k_apply_nodes(uint32_t n, const uint32_t* __restrict__ alist, const uint32_t* __restrict__ nclu,
                              const uint32_t* __restrict__ ckeep, const int32_t* __restrict__ clabel,
                              const double* __restrict__ cavg, const uint32_t* __restrict__ dint,
                              int32_t* __restrict__ label, double* __restrict__ x, uint64_t* __restrict__ y,
                              uint8_t* __restrict__ alive, uint32_t* err) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t u = alist[i];
        const uint32_t c = nclu[u];
        label[u] = clabel[c];
        if (ckeep[c]) {
            const double v = x[u] - cavg[c];
            x[u] = v;
            put_y(y, u, v, dint[u], err);
        } else {
            alive[u] = 0;
        }
    }
}

// ------------------------------------------------------------------ compaction

// cstart[c]: the new first position of kept cluster c (SelPositions)
__global__ void k_clu_compact(uint32_t nc, uint32_t na_new, const uint32_t* __restrict__ ckeep,
                              const uint32_t* __restrict__ cscan, const uint32_t* __restrict__ cstart,
                              ClusterCols in, ClusterCols out, const uint32_t* pnc = nullptr,
                              const uint32_t* pna_new = nullptr) {
    nc = dev_count(nc, pnc);
    na_new = dev_count(na_new, pna_new);
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        if (!ckeep[c]) continue;
        const uint32_t d = cscan[c];
        out.cs[d] = cstart[c];
        out.cgraph[d] = in.cgraph[c];
        out.cmin[d] = in.cmin[c];
        out.cflag[d] = in.cflag[c] & CF_CONN;
        out.clabel[d] = in.clabel[c];
        out.cvol[d] = 0u;             // k_volume adds the new volumes
        if (d + 1 == cscan[nc]) out.cs[d + 1] = na_new;
    }
}

__global__ void k_compact_u32(uint32_t n, const uint32_t* __restrict__ f, const uint32_t* __restrict__ fs,
                              const uint32_t* __restrict__ in, uint32_t* __restrict__ out, const uint32_t* pn = nullptr) {
    n = dev_count(n, pn);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (f[i]) out[fs[i]] = in[i];
}

// ------------------------------------------------------------------ disable arcs (K23)

// For surviving CC members: keep arcs whose endpoint has the same component
// root, compacting them order-preservingly to the row front
// (cuda/devPartition.h:80-146); refresh dInt and y when the row changed.
template <int GS>
__global__ void __launch_bounds__(BS) k_disable(uint32_t nl, const uint32_t* __restrict__ list,
                                                const uint8_t* __restrict__ alive,
                                                const uint32_t* __restrict__ row,
                                                uint32_t* __restrict__ dint, uint32_t* __restrict__ nbr,
                                                const uint32_t* __restrict__ root,
                                                const double* __restrict__ x, uint64_t* __restrict__ y,
                                                uint32_t heavy, uint32_t* err, const uint32_t* pn = nullptr) {
    nl = dev_count(nl, pn);
    const uint32_t lane = threadIdx.x & (GS - 1);
    const XcdSpan sp = xcd_span(nl, BS / GS, threadIdx.x / GS);
    for (uint32_t i = sp.first; i < sp.end; i += sp.stride) {
        const uint32_t u = list[i];
        if (!alive[u]) continue;                      // uniform within the group
        const uint32_t start = row[u], d = dint[u];
        if (d > heavy) continue;                      // k_disable_heavy
        const uint32_t ru = root[u];
        uint32_t kept = 0;
        // four chunks of GS arcs loaded (indices, then roots) before the
        // in-order compaction; a write goes to an index at most that of the
        // arc written, all of which are already loaded
        for (uint32_t j0 = 0; j0 < d; j0 += 4 * GS) {
            uint32_t v[4];
            bool keep[4];
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) v[q] = nbr[start + min(j0 + q * GS + lane, d - 1)];
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) keep[q] = (j0 + q * GS + lane < d) && root[v[q]] == ru;
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) {
                const uint32_t j = j0 + q * GS + lane;
                const uint64_t bal = group_ballot<GS>(keep[q]);
                const uint32_t before = (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                if (keep[q] && kept + before != j) nbr[start + kept + before] = v[q];
                kept += (uint32_t)__popcll(bal);
            }
        }
        if (lane == 0 && kept != d) {
            dint[u] = kept;
            put_y(y, u, x[u], kept, err);
        }
    }
}

// k_disable with an LDS window of the component roots (the layout of
// k_walk_win: chunks of CH entries of the ascending member list, WIN roots
// around a chunk staged when its nodes span less than WIN/2, 64 rows per wave
// in GS rounds).  The per-row compaction is k_disable's: GS lanes load U
// chunks of GS arcs, then write the kept arcs in order, each to an index at
// most its own (all already loaded).  The kept count goes back to the owner
// lane, which updates dInt and y.
template <int GS, int U, int WIN, int CH>
__global__ void __launch_bounds__(WBS) k_disable_win(uint32_t nl, const uint32_t* __restrict__ list,
                                                     const uint8_t* __restrict__ alive,
                                                     const uint32_t* __restrict__ row,
                                                     uint32_t* __restrict__ dint, uint32_t* __restrict__ nbr,
                                                     const uint32_t* __restrict__ root,
                                                     const double* __restrict__ x, uint64_t* __restrict__ y,
                                                     uint32_t heavy, uint32_t N, uint32_t* err,
                                                     const uint32_t* pn = nullptr) {
    static_assert(WIN % 4 == 0 && 64 % GS == 0, "window of root quads, lane groups within a wave");
    __shared__ uint32_t sr[WIN];
    nl = dev_count(nl, pn);
    constexpr uint32_t NPR = 64 / GS;
    constexpr uint32_t NW = WBS / 64;
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    const uint32_t g = lane / GS, gl = lane % GS;
    const uint32_t nch = (nl + CH - 1) / CH;
    const uint32_t xx = blockIdx.x % NXCD, lb = blockIdx.x / NXCD, nb = gridDim.x / NXCD;
    const uint32_t cpx = nch / NXCD + (nch % NXCD != 0);
    const uint32_t cb = min(nch, xx * cpx), ce = min(nch, cb + cpx);
    for (uint32_t c = cb + lb; c < ce; c += nb) {
        const uint32_t c0 = c * CH, c1 = min(c0 + CH, nl);
        const uint32_t uf = list[c0], ul = list[c1 - 1];
        const bool use = ul - uf < (uint32_t)WIN / 2;
        uint32_t wb = 0;
        if (use) {
            const uint32_t mid = uf / 2 + ul / 2;
            wb = mid > (uint32_t)WIN / 2 ? mid - (uint32_t)WIN / 2 : 0u;
            if (N <= (uint32_t)WIN) wb = 0;
            else if (wb > N - (uint32_t)WIN) wb = N - (uint32_t)WIN;
            wb &= ~3u;
            for (uint32_t k = 4 * threadIdx.x; k < (uint32_t)WIN; k += 4 * WBS) {
                if (wb + k + 3 < N) {
                    const uint4 t = *(const uint4*)(root + wb + k);
                    sr[k] = t.x;
                    sr[k + 1] = t.y;
                    sr[k + 2] = t.z;
                    sr[k + 3] = t.w;
                } else {
                    for (uint32_t q = 0; q < 4; ++q)
                        if (wb + k + q < N) sr[k + q] = root[wb + k + q];
                }
            }
        }
        __syncthreads();
        for (uint32_t ib = c0 + w * 64; ib < c1; ib += NW * 64) {
            const uint32_t i = ib + lane;
            const bool ok = i < c1;
            const uint32_t u = ok ? list[i] : list[c0];
            const uint32_t st = row[u];
            const uint32_t dd = dint[u];
            const bool mine = ok && alive[u] && dd <= heavy;   // long rows: k_disable_heavy
            const uint32_t d = mine ? dd : 0u;
            const uint32_t ru = root[u];
            uint32_t res = 0;
#pragma unroll
            for (int r = 0; r < GS; ++r) {
                const uint32_t s = r * NPR + g;
                const uint32_t rs = __shfl(st, s), rd = __shfl(d, s), rr = __shfl(ru, s);
                uint32_t kept = 0;
                for (uint32_t j0 = 0; j0 < rd; j0 += U * GS) {
                    uint32_t v[U];
                    bool keep[U];
#pragma unroll
                    for (int q = 0; q < U; ++q) {
                        const uint32_t j = j0 + q * GS + gl;
                        v[q] = j < rd ? nbr[rs + j] : 0u;
                    }
                    bool valid[U];
                    uint32_t rv[U];
#pragma unroll
                    for (int q = 0; q < U; ++q) valid[q] = j0 + q * GS + gl < rd;
                    win_gather<U, false>(rv, v, valid, sr, root, wb, (uint32_t)WIN, use);
#pragma unroll
                    for (int q = 0; q < U; ++q) keep[q] = valid[q] && rv[q] == rr;
#pragma unroll
                    for (int q = 0; q < U; ++q) {
                        const uint32_t j = j0 + q * GS + gl;
                        const uint64_t bal = group_ballot<GS>(keep[q]);
                        const uint32_t before = (uint32_t)__popcll(bal & ((1ull << gl) - 1ull));
                        if (keep[q] && kept + before != j) nbr[rs + kept + before] = v[q];
                        kept += (uint32_t)__popcll(bal);
                    }
                }
                const uint32_t got = __shfl(kept, (lane % NPR) * GS);
                if (lane / NPR == (uint32_t)r) res = got;
            }
            if (mine && res != d) {
                dint[u] = res;
                put_y(y, u, x[u], res, err);
            }
        }
        __syncthreads();
    }
}

// 1 for the alive rows of the list above the threshold
__global__ void k_heavy_flags(uint32_t n, const uint32_t* __restrict__ list, const uint8_t* __restrict__ alive,
                              const uint32_t* __restrict__ dint, uint32_t heavy, uint32_t* __restrict__ f,
                              const uint32_t* pn = nullptr) {
    n = dev_count(n, pn);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t u = list[i];
        f[i] = (alive[u] && dint[u] > heavy) ? 1u : 0u;
    }
}

// arc disabling for rows above the threshold: one workgroup per row, chunks of
// HEAVY_CHUNK arcs in order; each thread owns four consecutive arcs, a block
// scan ranks the kept ones, and every write (after the barrier that ends the
// chunk's loads) goes to an index no larger than the arc's own
__global__ void __launch_bounds__(BS) k_disable_heavy(const uint32_t* __restrict__ nrows,
                                                      const uint32_t* __restrict__ hnode,
                                                      const uint32_t* __restrict__ row,
                                                      uint32_t* __restrict__ dint, uint32_t* __restrict__ nbr,
                                                      const uint32_t* __restrict__ root,
                                                      const double* __restrict__ x, uint64_t* __restrict__ y,
                                                      uint32_t* err) {
    __shared__ uint32_t lds[BS / WAVE];
    const uint32_t R = *nrows;
    for (uint32_t r = blockIdx.x; r < R; r += gridDim.x) {
        const uint32_t u = hnode[r];
        const uint32_t start = row[u], d = dint[u], ru = root[u];
        uint32_t kept = 0;
        for (uint32_t j0 = 0; j0 < d; j0 += HEAVY_CHUNK) {
            uint32_t v[4], nk = 0;
            bool k[4];
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q) {
                const uint32_t j = j0 + 4u * threadIdx.x + q;
                v[q] = nbr[start + min(j, d - 1u)];
                k[q] = j < d && root[v[q]] == ru;
                nk += k[q] ? 1u : 0u;
            }
            uint32_t tot;
            uint32_t w = kept + block_excl_scan<BS>(nk, lds, tot);
#pragma unroll
            for (uint32_t q = 0; q < 4; ++q)
                if (k[q]) nbr[start + w++] = v[q];
            kept += tot;
            __syncthreads();
        }
        if (threadIdx.x == 0 && kept != d) {
            dint[u] = kept;
            put_y(y, u, x[u], kept, err);
        }
        __syncthreads();
    }
}

// cvol[c] = sum of dInt over the cluster's positions (the k_stats_sum layout)
__global__ void __launch_bounds__(BS) k_volume(uint32_t na, const uint32_t* __restrict__ pnix,
                                               const uint32_t* __restrict__ pclu,
                                               const uint32_t* __restrict__ dint, uint32_t* __restrict__ cvol,
                                               const uint32_t* pn = nullptr) {
    __shared__ uint32_t lds[BS / WAVE];
    na = dev_count(na, pn);
    constexpr uint32_t TR = BS * SUM_ITEMS;
    const XcdSpan sp = xcd_tiles(na, TR);
    for (uint32_t t = sp.first; t < sp.end; t += sp.stride) {
        const uint32_t r0 = t * TR;
        const uint32_t r1 = min(r0 + TR, na);
        const uint32_t cf = pclu[r0];
        // every item's node, then every item's degree (loads in flight
        // together; items past the end read the last position and add 0)
        uint32_t u[SUM_ITEMS], dv[SUM_ITEMS];
#pragma unroll
        for (uint32_t i = 0; i < SUM_ITEMS; ++i) u[i] = pnix[min(r0 + i * BS + threadIdx.x, r1 - 1u)];
#pragma unroll
        for (uint32_t i = 0; i < SUM_ITEMS; ++i) dv[i] = r0 + i * BS + threadIdx.x < r1 ? dint[u[i]] : 0u;
        if (cf == pclu[r1 - 1]) {
            uint32_t acc = 0;
#pragma unroll
            for (uint32_t i = 0; i < SUM_ITEMS; ++i) acc += dv[i];
            acc = group_sum<WAVE>(acc);
            if (lane_id() == 0) lds[threadIdx.x / WAVE] = acc;
            __syncthreads();
            if (threadIdx.x == 0) {
                uint32_t t = 0;
                for (uint32_t w = 0; w < BS / WAVE; ++w) t += lds[w];
                atomicAdd(&cvol[cf], t);
            }
            __syncthreads();
            continue;
        }
        for (uint32_t i = 0; i < SUM_ITEMS; ++i) {
            const uint32_t k = r0 + i * BS + threadIdx.x;
            const bool ok = k < r1;
            const uint32_t c = ok ? pclu[k] : 0xFFFFFFFFu;
            uint32_t d = dv[i];
            const uint32_t c0 = __shfl(c, 0);
            if (c0 == 0xFFFFFFFFu) {
                continue;                      // whole wave past the end: nothing to add
            } else if (__all(c == c0 || c == 0xFFFFFFFFu)) {
                d = group_sum<WAVE>(d);
                if (lane_id() == 0) atomicAdd(&cvol[c0], d);
            } else if (ok) {
                atomicAdd(&cvol[c], d);
            }
        }
    }
}

// dst[i] = src[idx[i]]: GATHER_ITEMS elements a thread, all index loads, then
// all gathers in flight together (launched with BS threads a workgroup)
constexpr uint32_t GATHER_ITEMS = 4;
__global__ void __launch_bounds__(BS) k_gather(uint32_t n, const uint32_t* __restrict__ idx,
                                               const uint32_t* __restrict__ src, uint32_t* __restrict__ dst) {
    constexpr uint32_t TR = BS * GATHER_ITEMS;
    const XcdSpan sp = xcd_tiles(n, TR);
    for (uint32_t t = sp.first; t < sp.end; t += sp.stride) {
        const uint32_t i0 = t * TR + threadIdx.x;
        uint32_t j[GATHER_ITEMS], v[GATHER_ITEMS];
#pragma unroll
        for (uint32_t q = 0; q < GATHER_ITEMS; ++q) j[q] = idx[min(i0 + q * BS, n - 1u)];
#pragma unroll
        for (uint32_t q = 0; q < GATHER_ITEMS; ++q) v[q] = src[j[q]];
#pragma unroll
        for (uint32_t q = 0; q < GATHER_ITEMS; ++q)
            if (i0 + q * BS < n) dst[i0 + q * BS] = v[q];
    }
}

__global__ void k_fill_u32(uint32_t n, uint32_t* __restrict__ p, uint32_t v, const uint32_t* pn = nullptr) {
    n = dev_count(n, pn);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// segment bounds {0, n} of a one-segment sort (n read on the device when given)
__global__ void k_one_segment(uint32_t* __restrict__ seg, uint32_t n, const uint32_t* pn) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        seg[0] = 0u;
        seg[1] = dev_count(n, pn);
    }
}

// start of a batch of exact CC rounds: round r > 0 without a change ends them
// (later launches of the batch see cc_done and return at once)
__global__ void k_cc_round(Counters* cnt, uint32_t r) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && !cnt->cc_done) {
        if (r > 0 && cnt->link_changed == 0u) cnt->cc_done = 1u;
        else cnt->link_changed = 0u;
    }
}

// count of still-valid arcs over all nodes (getNumCutEdges)
__global__ void k_sum_u32(uint32_t n, const uint32_t* __restrict__ a, unsigned long long* __restrict__ out) {
    unsigned long long s = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) s += a[i];
    s = group_sum<WAVE>(s);
    if (lane_id() == 0) atomicAdd(out, s);
}

}  // namespace pxd

// Device-side CSR build from an undirected edge list (graph ingest), used by
// pxd_initialize_edges, the serving engine's edge-list admissions and
// pxd_csr_from_edges_noalloc.  Included by pxd_engine.hip after pxd_kernels.h
// (it reuses the engine's scans and the LDS digit ranking of its sorts).
//
// Same result as the reference's host build: every edge (u, v) gives the arcs
// u -> v and v -> u, rows are ordered by target, ranges[u] is the first arc of
// u (lib/utils/graph_io.h:49-84 reads the edges, lib/core/graph.h:147-184
// sorts them and fills the ranges).  Loops and duplicate edges are kept.
//
//   1. degrees by atomics (an end >= n sets the caller's flag and the edge is
//      left out), 2. exclusive scan -> ranges, 3. both arcs of every edge
//      written into their rows at atomic cursors (any order within a row),
//   4. every row sorted ascending by target: rows of up to 64 arcs by one
//      wave (rank by counting), up to TILE by one workgroup in LDS (ballot-
//      ranked 8-bit digit passes), longer ones by one workgroup each (LSD over
//      8-bit digits, tiles ranked in LDS, through a work copy of the row).
// Only target ids are sorted, so equal targets are interchangeable and the
// result does not depend on the scatter order.  Grid sizes come from the
// host-known sizes and every count is read on the device: a build issues no
// host synchronization.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace pxd {

constexpr uint32_t INGEST_WAVE_ROW = WAVE;   // rows up to this many arcs: one wave
constexpr uint32_t INGEST_BLOCK_ROW = TILE;  // up to this many: one workgroup in LDS

// Work buffers of one build, carved from one array of 32-bit words.
struct IngestWork {
    uint32_t* deg = nullptr;    // n + 1: degree counts, then the scatter cursors
    uint32_t* list = nullptr;   // n: rows of 2..64 arcs from the front, 65..TILE from the back
    uint32_t* llong = nullptr;  // rows above TILE arcs
    uint32_t* tmp = nullptr;    // 2m: copies of long rows between digit passes
    uint32_t* ctr = nullptr;    // 4 counters: short, block and long rows
    uint32_t* lvl = nullptr;    // scan partial sums
    uint32_t lvl_words = 0;
    uint32_t nmax = 0;
    uint64_t amax = 0;
};

// partial-sum words of the three-level scan of up to n + 1 elements
constexpr uint32_t ingest_scan_words(uint32_t n) {
    uint64_t len = (uint64_t)n + 2, words = 0;
    const uint64_t per = (uint64_t)BS * SCAN_ITEMS;
    for (;;) {
        const uint64_t nb = (len + per - 1) / per;
        words += 2 * nb + 3;
        if (nb <= 1) break;
        len = nb;
    }
    return (uint32_t)words;
}

constexpr uint64_t ingest_long_rows(uint64_t arcs) { return arcs / (INGEST_BLOCK_ROW + 1) + 1; }

// 32-bit words of the work area for graphs of up to n nodes and `arcs` arcs
constexpr uint64_t ingest_work_words(uint32_t n, uint64_t arcs) {
    return ((uint64_t)n + 1) + n + ingest_long_rows(arcs) + arcs + 4 + ingest_scan_words(n) + 64;
}

inline IngestWork ingest_carve(uint32_t* base, uint32_t n, uint64_t arcs) {
    IngestWork w;
    uint32_t* p = base;
    w.deg = p;
    p += (size_t)n + 1;
    w.list = p;
    p += n;
    w.llong = p;
    p += ingest_long_rows(arcs);
    w.tmp = p;
    p += arcs;
    w.ctr = p;
    p += 4;
    w.lvl = p;
    w.lvl_words = ingest_scan_words(n);
    w.nmax = n;
    w.amax = arcs;
    return w;
}

__global__ void k_ingest_degrees(uint64_t m, const uint32_t* __restrict__ fx, const uint32_t* __restrict__ tx,
                                 uint32_t n, uint32_t* __restrict__ deg, uint32_t* __restrict__ bad) {
    uint32_t b = 0u;
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = fx[e], v = tx[e];
        if (u >= n || v >= n) {       // reported; the edge is left out
            b = 1u;
            continue;
        }
        atomicAdd(&deg[u], 1u);
        atomicAdd(&deg[v], 1u);
    }
    b = group_or<WAVE>(b);
    if (b && lane_id() == 0) atomicOr(bad, 1u);
}

__global__ void k_ingest_cursors(uint32_t n, const uint32_t* __restrict__ row, uint32_t* __restrict__ cur) {
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) cur[u] = row[u];
}

__global__ void k_ingest_scatter(uint64_t m, const uint32_t* __restrict__ fx, const uint32_t* __restrict__ tx,
                                 uint32_t n, uint32_t* __restrict__ cur, uint32_t* __restrict__ nbr) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < m; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = fx[e], v = tx[e];
        if (u >= n || v >= n) continue;
        nbr[atomicAdd(&cur[u], 1u)] = v;
        nbr[atomicAdd(&cur[v], 1u)] = u;
    }
}

// rows by length: 2..64 -> list from the front (ctr[0]), 65..TILE -> list
// from the back (ctr[1]), longer -> llong (ctr[2])
__global__ void __launch_bounds__(BS) k_ingest_classify(uint32_t n, const uint32_t* __restrict__ row,
                                                        uint32_t* __restrict__ list, uint32_t* __restrict__ llong,
                                                        uint32_t* __restrict__ ctr) {
    for (uint32_t u0 = blockIdx.x * blockDim.x; u0 < n; u0 += gridDim.x * blockDim.x) {
        const uint32_t u = u0 + threadIdx.x;
        const uint32_t len = u < n ? row[u + 1] - row[u] : 0u;
        const bool w = len > 1u && len <= INGEST_WAVE_ROW;
        const bool b = len > INGEST_WAVE_ROW && len <= INGEST_BLOCK_ROW;
        const bool l = len > INGEST_BLOCK_ROW;
        const uint32_t sw = wave_append(&ctr[0], w);
        const uint32_t sb = wave_append(&ctr[1], b);
        const uint32_t sl = wave_append(&ctr[2], l);
        if (w) list[sw] = u;
        if (b) list[n - 1u - sb] = u;
        if (l) llong[sl] = u;
    }
}

// one wave per row of 2..64 arcs: rank by counting (ties by lane)
__global__ void __launch_bounds__(BS) k_ingest_sort_wave(const uint32_t* __restrict__ list,
                                                         const uint32_t* __restrict__ ctr,
                                                         const uint32_t* __restrict__ row, uint32_t* __restrict__ nbr) {
    const uint32_t cnt = ctr[0];
    const uint32_t lane = lane_id();
    for (uint32_t w = blockIdx.x * (BS / WAVE) + threadIdx.x / WAVE; w < cnt; w += gridDim.x * (BS / WAVE)) {
        const uint32_t u = list[w];
        const uint32_t s0 = row[u], s = row[u + 1] - s0;
        const bool valid = lane < s;
        const uint32_t k = valid ? nbr[s0 + lane] : 0xFFFFFFFFu;
        uint32_t rank = 0;
        for (uint32_t j = 0; j < s; ++j) {
            const uint32_t kj = __shfl(k, (int)j);
            rank += (kj < k || (kj == k && j < lane)) ? 1u : 0u;
        }
        if (valid) nbr[s0 + rank] = k;
    }
}

// one workgroup per row of 65..TILE arcs, sorted in LDS (digits that are
// constant over the row are skipped)
__global__ void __launch_bounds__(BS) k_ingest_sort_block(uint32_t n, const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ ctr,
                                                          const uint32_t* __restrict__ row,
                                                          uint32_t* __restrict__ nbr) {
    constexpr uint32_t NW = BS / WAVE;
    __shared__ uint64_t buf[TILE];
    __shared__ uint32_t wcnt[NW][256];
    __shared__ uint32_t hs[256];
    __shared__ uint32_t sred[NW];
    __shared__ uint32_t s_or, s_and;
    const uint32_t cnt_rows = ctr[1];
    const uint32_t w = threadIdx.x / WAVE, l = lane_id();
    for (uint32_t r = blockIdx.x; r < cnt_rows; r += gridDim.x) {
        const uint32_t u = list[n - 1u - r];
        const uint32_t s0 = row[u], cnt = row[u + 1] - s0;
        if (threadIdx.x == 0) {
            s_or = 0u;
            s_and = 0xFFFFFFFFu;
        }
        __syncthreads();
        uint64_t e[TILE_ITEMS];
        uint32_t lor = 0u, land = 0xFFFFFFFFu;
#pragma unroll
        for (int i = 0; i < TILE_ITEMS; ++i) {
            const uint32_t p = w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l;
            const bool ok = p < cnt;
            const uint32_t k = ok ? nbr[s0 + p] : 0u;
            e[i] = (uint64_t)k << 32;
            if (ok) {
                lor |= k;
                land &= k;
            }
        }
        atomicOr(&s_or, lor);
        atomicAnd(&s_and, land);
        __syncthreads();
        const uint32_t bits = s_or ^ s_and;
        uint32_t pos[TILE_ITEMS];
        for (int pass = 0; pass < 4; ++pass) {
            if (((bits >> (8 * pass)) & 255u) == 0u) continue;   // uniform over the workgroup
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
            __syncthreads();          // buf is written again by the next pass
        }
#pragma unroll
        for (int i = 0; i < TILE_ITEMS; ++i) {
            const uint32_t p = w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l;
            if (p < cnt) nbr[s0 + p] = (uint32_t)(e[i] >> 32);
        }
        __syncthreads();              // s_or / s_and are reset for the next row
    }
}

// one workgroup per row above TILE arcs: LSD over 8-bit digits, each pass a
// row histogram, then the row's tiles in order, each ranked stably in LDS and
// written to its digits' running offsets in the other buffer (the row or its
// copy in tmp); a final copy back when the passes end in tmp
__global__ void __launch_bounds__(BS) k_ingest_sort_long(const uint32_t* __restrict__ llong,
                                                         const uint32_t* __restrict__ ctr,
                                                         const uint32_t* __restrict__ row, uint32_t* __restrict__ nbr,
                                                         uint32_t* __restrict__ tmp) {
    constexpr uint32_t NW = BS / WAVE;
    __shared__ uint32_t wcnt[NW][256];
    __shared__ uint32_t hs[256], hist[256], base[256];
    __shared__ uint32_t sred[NW];
    __shared__ uint32_t s_or, s_and;
    const uint32_t cnt_rows = ctr[2];
    const uint32_t w = threadIdx.x / WAVE, l = lane_id();
    for (uint32_t r = blockIdx.x; r < cnt_rows; r += gridDim.x) {
        const uint32_t u = llong[r];
        const uint32_t s0 = row[u], len = row[u + 1] - s0;
        if (threadIdx.x == 0) {
            s_or = 0u;
            s_and = 0xFFFFFFFFu;
        }
        __syncthreads();
        uint32_t lor = 0u, land = 0xFFFFFFFFu;
        for (uint32_t p = threadIdx.x; p < len; p += BS) {
            const uint32_t k = nbr[s0 + p];
            lor |= k;
            land &= k;
        }
        atomicOr(&s_or, lor);
        atomicAnd(&s_and, land);
        __syncthreads();
        const uint32_t bits = s_or ^ s_and;
        uint32_t* src = nbr + s0;
        uint32_t* dst = tmp + s0;
        for (int pass = 0; pass < 4; ++pass) {
            const int sh = 8 * pass;
            if (((bits >> sh) & 255u) == 0u) continue;          // uniform over the workgroup
            hist[threadIdx.x] = 0u;
            __syncthreads();
            for (uint32_t p = threadIdx.x; p < len; p += BS) atomicAdd(&hist[(src[p] >> sh) & 255u], 1u);
            __syncthreads();
            {
                uint32_t tot;
                base[threadIdx.x] = block_excl_scan<BS>(hist[threadIdx.x], sred, tot);
            }
            __syncthreads();
            for (uint32_t t0 = 0; t0 < len; t0 += TILE) {
                const uint32_t cnt = min((uint32_t)TILE, len - t0);
                uint64_t e[TILE_ITEMS];
#pragma unroll
                for (int i = 0; i < TILE_ITEMS; ++i) {
                    const uint32_t p = w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l;
                    e[i] = p < cnt ? ((uint64_t)src[t0 + p] << 32) : 0ull;
                }
                uint32_t pos[TILE_ITEMS];
                rank8(e, cnt, sh, wcnt, hs, sred, pos);
#pragma unroll
                for (int i = 0; i < TILE_ITEMS; ++i) {
                    const uint32_t p = w * (WAVE * TILE_ITEMS) + (uint32_t)i * WAVE + l;
                    if (p < cnt) {
                        const uint32_t k = (uint32_t)(e[i] >> 32);
                        const uint32_t d = (k >> sh) & 255u;
                        dst[base[d] + (pos[i] - hs[d])] = k;
                    }
                }
                __syncthreads();          // every write of the tile used base[] before it advances
                {
                    const uint32_t d = threadIdx.x;
                    const uint32_t next = d < 255u ? hs[d + 1] : cnt;
                    base[d] += next - hs[d];
                }
                __syncthreads();
            }
            uint32_t* t = src;
            src = dst;
            dst = t;
        }
        if (src != nbr + s0)
            for (uint32_t p = threadIdx.x; p < len; p += BS) nbr[s0 + p] = src[p];
        __syncthreads();
    }
}

// exclusive scan out[0..n) of in[0..n), out[n] = total, partial sums in lvl
inline hipError_t ingest_scan(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* lvl, uint32_t lvl_words,
                              hipStream_t st) {
    if (n == 0) {
        k_fill_u32<<<1, 64, 0, st>>>(1, out, 0u);
        return hipGetLastError();
    }
    const uint32_t nb = (n + BS * SCAN_ITEMS - 1) / (BS * SCAN_ITEMS);
    if (nb <= 1) {
        k_scan_down<<<1, BS, 0, st>>>(n, in, out, nullptr);
        return hipGetLastError();
    }
    if (2 * nb + 3 > lvl_words) return hipErrorInvalidValue;
    uint32_t* sums = lvl;
    uint32_t* offs = lvl + nb + 1;
    k_scan_up<<<nb, BS, 0, st>>>(n, in, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = ingest_scan(sums, offs, nb, lvl + 2 * nb + 3, lvl_words - (2 * nb + 3), st);
    if (e != hipSuccess) return e;
    k_scan_down<<<nb, BS, 0, st>>>(n, in, out, offs);
    return hipGetLastError();
}

inline uint32_t ingest_grid(uint64_t items, uint32_t per, uint32_t cap) {
    uint64_t g = (items + per - 1) / per;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

// The build, queued on st: ranges d_row[0..n], arcs d_nbr[0..2m) from the
// device edge arrays (d_fx[i], d_tx[i]), i < m.  The work area w must hold n
// nodes and 2m arcs.  *d_bad (zeroed by the caller) becomes nonzero when an
// end is >= n; the CSR is then meaningless.
inline hipError_t ingest_csr_async(uint32_t n, uint64_t m, const uint32_t* d_fx, const uint32_t* d_tx,
                                   uint32_t* d_row, uint32_t* d_nbr, const IngestWork& w, uint32_t* d_bad,
                                   hipStream_t st) {
    const uint64_t arcs = 2 * m;
    if (n > w.nmax || arcs > w.amax) return hipErrorInvalidValue;
    k_fill_u32<<<ingest_grid((uint64_t)n + 1, BS, 8192), BS, 0, st>>>(n + 1, w.deg, 0u);
    k_fill_u32<<<1, 64, 0, st>>>(4, w.ctr, 0u);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (m) {
        k_ingest_degrees<<<ingest_grid(m, BS, 8192), BS, 0, st>>>(m, d_fx, d_tx, n, w.deg, d_bad);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if ((e = ingest_scan(w.deg, d_row, n, w.lvl, w.lvl_words, st)) != hipSuccess) return e;
    if (m && n) {
        k_ingest_cursors<<<ingest_grid(n, BS, 8192), BS, 0, st>>>(n, d_row, w.deg);
        k_ingest_scatter<<<ingest_grid(m, BS, 8192), BS, 0, st>>>(m, d_fx, d_tx, n, w.deg, d_nbr);
        k_ingest_classify<<<ingest_grid(n, BS, 4096), BS, 0, st>>>(n, d_row, w.list, w.llong, w.ctr);
        k_ingest_sort_wave<<<ingest_grid(n, BS / WAVE, 8192), BS, 0, st>>>(w.list, w.ctr, d_row, d_nbr);
        k_ingest_sort_block<<<ingest_grid(arcs / (INGEST_WAVE_ROW + 1) + 1, 1, 4096), BS, 0, st>>>(n, w.list, w.ctr,
                                                                                                   d_row, d_nbr);
        k_ingest_sort_long<<<ingest_grid(ingest_long_rows(arcs), 1, 1024), BS, 0, st>>>(w.llong, w.ctr, d_row, d_nbr,
                                                                                        w.tmp);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace pxd

// Whole decomposition of one small graph in one workgroup (gfx950, 16 waves):
// the iteration loop of the engine (walk, sweep, cut) runs on the device, with
// a workgroup barrier between phases instead of a kernel boundary and without
// host reads of counts.  For graphs of up to SMALL_MAX_N nodes every launch of
// the multi-kernel path is a few microseconds of latency, so one launch for
// the whole run replaces some 40 launches and three host reads per iteration.
//
// The phases restate the engine's kernels for one graph (cluster ids and
// positions are the graph's own, starting at 0) with the same arithmetic, so
// the final state -- labels, walk values, internal degrees, arcs, iteration
// count -- has the same bits as the multi-kernel path, and the engine's host
// side continues from it (further step calls, downloads).  Connected
// components use a union-find in LDS (hooks by compare-and-swap of a root
// under a smaller root), which gives every component its minimum node as root,
// the label the engine's linking rounds compute.  Per-cluster reductions (best
// cut, exact sums, volumes) are LDS atomics: integer and min reductions do not
// depend on their order.
#pragma once

#include "pxd_kernels.h"

namespace pxd {

// 16 waves: the phases are latency-bound, and four items a thread keep the
// sort within the 128 registers a thread of a 1024-thread workgroup has
constexpr int SM_NT = 1024;                          // threads of the workgroup
constexpr int SM_ITEMS = 4;                          // items per thread in tile passes
constexpr uint32_t SMALL_MAX_N = (uint32_t)SM_NT * SM_ITEMS;   // 4096 nodes
constexpr uint32_t SMALL_MAX_ARCS = 1u << 19;        // capacity of the one-workgroup run
// The automatic choice takes the one-workgroup run up to this many arcs: it
// costs about 20 us + 1.8 ns an arc per iteration on one CU (4096-node SBM,
// 52,918 arcs: 116 us; a 1,714-node level of the expander hierarchy, 291,712
// arcs: 550 us), the per-phase kernels about 170-270 us an iteration at
// these sizes.
constexpr uint32_t SMALL_AUTO_ARCS = 1u << 16;
constexpr uint32_t SM_LONG = 32;                     // longer rows: 16 lanes each, else one thread
constexpr uint32_t SM_GL = 16;                       // lanes per long row: four long rows a wave
constexpr uint32_t SM_NW = SM_NT / WAVE;

struct SmallArgs {
    uint32_t n, maxit;
    uint32_t na, nc, nalist;             // state at the call (after initialize: n, 1, n)
    double sc_thr;
    float rw_thr;
    const uint32_t* row;
    uint32_t *dint, *nbr, *xkey, *stamp, *root;
    double* x[2];
    uint64_t* y[2];
    int32_t* label;
    uint8_t* alive;
    uint32_t *pnix[2], *pclu[2], *alist[2];   // positions and active list: current and next
    uint32_t *pkey, *cclist, *nclu, *flags;
    uint2* edv;
    ClusterCols ct[2];                   // cluster table: current and next
    uint32_t *coff, *ckeep, *kscan, *cmaxo, *cmino;
    float* csp;
    double* cavg;
    int32_t* cK;
    int32_t *gsmall, *giter;             // graph 0
    pxd_node_data* nd;                   // NodeData per position, or null
    Counters* cnt;
    uint32_t* out;                       // na, nc, nalist, cur, iter1, loop counter, buffers
};

// exclusive scan of one value per thread over the workgroup
__device__ __forceinline__ uint32_t sm_scan(uint32_t v, uint32_t* sred, uint32_t& total) {
    return block_excl_scan<SM_NT>(v, sred, total);
}

// Keeps the entries of a[0..n) with keep(i) (in order, in place) and returns
// the count; n <= SMALL_MAX_N.  Each thread holds its ITEMS consecutive
// entries in registers across the barrier, so the writes cannot overwrite an
// entry before it is read.
template <class K>
__device__ uint32_t sm_compact(uint32_t* a, uint32_t n, K&& keep, uint32_t* sred) {
    const uint32_t i0 = threadIdx.x * SM_ITEMS;
    uint32_t v[SM_ITEMS], f = 0u, c = 0u;
#pragma unroll
    for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
        const uint32_t i = i0 + q;
        v[q] = i < n ? a[i] : 0u;
        if (i < n && keep(i, v[q])) { f |= 1u << q; ++c; }
    }
    uint32_t tot;
    uint32_t o = sm_scan(c, sred, tot);
    __syncthreads();
#pragma unroll
    for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q)
        if (f & (1u << q)) a[o++] = v[q];
    __syncthreads();
    return tot;
}

// The long rows among a wave's 64 entries (set lanes of the wave-uniform
// mask m) go four at a time to the wave's four 16-lane groups: group g takes
// the g-th lowest set lane.  Returns that lane, or 64 when the group has no
// row this round, and clears the taken lanes from m.
__device__ __forceinline__ uint32_t sm_take4(uint64_t& m) {
    const uint32_t g = lane_id() / SM_GL;
    uint32_t src = WAVE;
#pragma unroll
    for (uint32_t q = 0; q < WAVE / SM_GL; ++q) {
        const uint32_t r = m ? (uint32_t)__builtin_ctzll(m) : (uint32_t)WAVE;
        if (q == g) src = r;
        if (m) m &= m - 1ull;
    }
    return src;
}

// union-find in LDS: parents only ever point to smaller nodes of the same
// component; path halving lowers a pointer to an ancestor (atomicMin keeps it
// monotone under concurrent updates)
__device__ __forceinline__ uint32_t sm_find(uint32_t* p, uint32_t x) {
    volatile uint32_t* vp = p;
    uint32_t px = vp[x];
    while (px != x) {
        const uint32_t gp = vp[px];
        if (gp != px) atomicMin(&p[x], gp);
        x = px;
        px = vp[x];
    }
    return x;
}
__device__ __forceinline__ void sm_union(uint32_t* p, uint32_t a, uint32_t b) {
    for (;;) {
        a = sm_find(p, a);
        b = sm_find(p, b);
        if (a == b) return;
        if (a > b) { const uint32_t t = a; a = b; b = t; }
        if (atomicCAS(&p[b], b, a) == b) return;   // b was still a root: hooked under a
    }
}

__global__ void __launch_bounds__(SM_NT) k_decompose_small(SmallArgs args) {
    // The arguments are read from an LDS copy: a use reloads its pointer
    // there instead of the kernel holding some fifty pointers in registers
    // for its whole run (which 1024 threads' register budget cannot hold).
    __shared__ SmallArgs a;
    if (threadIdx.x == 0) a = args;
    // LDS: the sort's buffers; between sorts the same bytes hold the
    // per-cluster reductions and the union-find
    __shared__ uint64_t s_buf[SMALL_MAX_N];
    __shared__ uint32_t s_lab[SMALL_MAX_N];
    __shared__ uint32_t s_wcnt[SM_NW][256];
    __shared__ uint32_t s_hs[256];
    __shared__ uint32_t s_red[SM_NW];
    __shared__ uint32_t s_bits[2];
    __shared__ uint32_t s_cnt[4];
    __shared__ int32_t s_lab32;
    __shared__ uint32_t s_segs[2];
    uint32_t* const s_u32 = reinterpret_cast<uint32_t*>(s_buf);   // 2 x SMALL_MAX_N words

    __syncthreads();
    const uint32_t t = threadIdx.x, l = lane_id(), w = t / WAVE;
    uint32_t na = a.na, nc = a.nc, nalist = a.nalist;
    uint32_t cur = 0u, iter1 = 0u, ovf = 0u;
    uint32_t pb = 0u, tb = 0u, ab = 0u;          // current position, table and list buffers
    int32_t gsmall = *a.gsmall, giter = *a.giter;
    const uint32_t maxit = a.maxit;
    uint32_t i = 0;
    while (i++ < maxit && nc > 0 && na > 0) {
        uint32_t* const pnix = a.pnix[pb];
        uint32_t* const pclu = a.pclu[pb];
        uint32_t* const alist = a.alist[ab];
        const ClusterCols T = a.ct[tb];
        // ------------------------------------------------------------ walk (k_walk)
        {
            const double* xc = a.x[cur];
            double* xn = a.x[cur ^ 1u];
            const uint64_t* yc = a.y[cur];
            uint64_t* yn = a.y[cur ^ 1u];
            // rows of at most SM_LONG arcs: one thread each, four arcs in flight
            for (uint32_t e = t; e < nalist; e += SM_NT) {
                const uint32_t u = alist[e];
                const uint32_t st = a.row[u], d = a.dint[u];
                if (d > SM_LONG) continue;
                FixAcc fa = {0ull, 0};
                for (uint32_t j = 0; j < d; j += 4) {
                    uint32_t v[4];
                    uint64_t q[4];
#pragma unroll
                    for (uint32_t z = 0; z < 4; ++z) v[z] = a.nbr[st + min(j + z, d - 1u)];
#pragma unroll
                    for (uint32_t z = 0; z < 4; ++z) q[z] = yc[v[z]];
#pragma unroll
                    for (uint32_t z = 0; z < 4; ++z) fixacc_add(fa, j + z < d ? q[z] : 0ull);
                }
                walk_finish(u, d, fixacc_fix(fa), xc, xn, yn, a.xkey, ovf);
            }
            // longer rows: 16 lanes each, four rows of a wave at a time
            for (uint32_t b = w * WAVE; b < nalist; b += SM_NT) {
                const uint32_t e = b + l;
                const uint32_t u = e < nalist ? alist[e] : 0u;
                const uint32_t d = e < nalist ? a.dint[u] : 0u;
                for (uint64_t m = __ballot(d > SM_LONG); m;) {
                    const uint32_t src = sm_take4(m);
                    const uint32_t ur = __shfl(u, (int)(src & 63u)), dr = __shfl(d, (int)(src & 63u));
                    if (src < (uint32_t)WAVE) {            // uniform within the 16-lane group
                        const uint32_t st = a.row[ur], gl = l % SM_GL;
                        FixAcc fa = {0ull, 0};
                        for (uint32_t j = gl; j < dr; j += SM_GL) {
                            fixacc_add(fa, yc[a.nbr[st + j]]);
                            fixacc_norm(fa);
                        }
                        const Fix128 acc = fixacc_fix(group_sum_fixacc<SM_GL>(fa));
                        if (gl == 0) walk_finish(ur, dr, acc, xc, xn, yn, a.xkey, ovf);
                    }
                }
            }
            cur ^= 1u;
        }
        __syncthreads();
        // ------------------------------------------------------------ sweep
        // keys in position order, stable sort inside every cluster
        for (uint32_t k = t; k < na; k += SM_NT) a.pkey[k] = a.xkey[pnix[k]];
        __syncthreads();
        wg_seg_sort<SM_NT, SM_ITEMS>(nc, T.cs, a.pkey, pnix, s_buf, s_lab, s_wcnt, s_hs, s_red, s_bits);
        // cut contribution (k_contrib): neighbours before u in key order count -1
        for (uint32_t e = t; e < nalist; e += SM_NT) {
            const uint32_t u = alist[e];
            const uint32_t st = a.row[u], d = a.dint[u], mk = a.xkey[u];
            if (d > SM_LONG) continue;
            int32_t c = 0;
            for (uint32_t j = 0; j < d; j += 4) {
                uint32_t v[4], kv[4];
#pragma unroll
                for (uint32_t z = 0; z < 4; ++z) v[z] = a.nbr[st + min(j + z, d - 1u)];
#pragma unroll
                for (uint32_t z = 0; z < 4; ++z) kv[z] = a.xkey[v[z]];
#pragma unroll
                for (uint32_t z = 0; z < 4; ++z) {
                    const bool before = (kv[z] != mk) ? (kv[z] < mk) : (v[z] < u);
                    if (j + z < d) c += before ? -1 : 1;
                }
            }
            a.edv[u] = make_uint2((uint32_t)c, d);
        }
        for (uint32_t b = w * WAVE; b < nalist; b += SM_NT) {
            const uint32_t e = b + l;
            const uint32_t u = e < nalist ? alist[e] : 0u;
            const uint32_t d = e < nalist ? a.dint[u] : 0u;
            for (uint64_t m = __ballot(d > SM_LONG); m;) {
                const uint32_t src = sm_take4(m);
                const uint32_t ur = __shfl(u, (int)(src & 63u)), dr = __shfl(d, (int)(src & 63u));
                if (src < (uint32_t)WAVE) {                // uniform within the 16-lane group
                    const uint32_t st = a.row[ur], mk = a.xkey[ur], gl = l % SM_GL;
                    int32_t c = 0;
                    for (uint32_t j = gl; j < dr; j += SM_GL) {
                        const uint32_t v = a.nbr[st + j], kv = a.xkey[v];
                        c += ((kv != mk) ? (kv < mk) : (v < ur)) ? -1 : 1;
                    }
                    c = group_sum<SM_GL>(c);
                    if (gl == 0) a.edv[ur] = make_uint2((uint32_t)c, dr);
                }
            }
        }
        unsigned long long* s_best = reinterpret_cast<unsigned long long*>(s_buf);   // per cluster: best cut
        for (uint32_t c = t; c < nc; c += SM_NT) s_best[c] = ~0ull;
        if (t < 4) s_cnt[t] = 0u;
        __syncthreads();
        {   // segmented prefix of (edgeDiff, volume) per cluster (k_sweep_a/b/c in one tile)
            uint32_t* f = s_lab;
            uint32_t* ev = s_lab + SM_NT;
            uint32_t* vv = s_lab + 2 * SM_NT;
            const uint32_t k0 = t * SM_ITEMS;
            uint32_t lf = 0u, le = 0u, lv = 0u;
            for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
                const uint32_t k = k0 + q;
                if (k >= na) break;
                const uint2 w = a.edv[pnix[k]];
                if (T.cs[pclu[k]] == k) { lf = 1u; le = w.x; lv = w.y; }
                else { le += w.x; lv += w.y; }
            }
            f[t] = lf; ev[t] = le; vv[t] = lv;
            __syncthreads();
            for (uint32_t off = 1; off < (uint32_t)SM_NT; off <<= 1) {
                uint32_t pf = 0u, pe = 0u, pv = 0u;
                const bool take = t >= off;
                if (take) { pf = f[t - off]; pe = ev[t - off]; pv = vv[t - off]; }
                __syncthreads();
                if (take && !f[t]) { ev[t] += pe; vv[t] += pv; f[t] = pf; }
                __syncthreads();
            }
            uint32_t re = t ? ev[t - 1] : 0u, rv = t ? vv[t - 1] : 0u;
            uint32_t runc = 0xFFFFFFFFu;
            unsigned long long runmin = ~0ull;
            for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
                const uint32_t k = k0 + q;
                if (k >= na) break;
                const uint32_t u = pnix[k], c = pclu[k];
                const uint2 w = a.edv[u];
                const uint32_t cb = T.cs[c];
                if (cb == k) { re = w.x; rv = w.y; }
                else { re += w.x; rv += w.y; }
                const uint32_t off = k - cb + 1u;
                const uint32_t tot = T.cvol[c];
                const uint32_t denom = (rv >= tot) ? 0u : ((rv < tot - rv) ? rv : (tot - rv));
                const float sp = denom > 0u ? ((float)re / (float)denom) : 2.0f;
                const unsigned long long pk = ((unsigned long long)ord_f32(sp) << 32) | (0xFFFFFFFFu - off);
                if (c != runc) {
                    if (runc != 0xFFFFFFFFu) atomicMin(&s_best[runc], runmin);
                    runc = c;
                    runmin = pk;
                } else if (pk < runmin) {
                    runmin = pk;
                }
                if (a.nd) {
                    pxd_node_data o;
                    o.nix = u; o.prefixVolume = rv; o.prefixEdgeDiff = (int32_t)re; o.offsetInCluster = off;
                    a.nd[k] = o;
                }
            }
            if (runc != 0xFFFFFFFFu) atomicMin(&s_best[runc], runmin);
        }
        __syncthreads();
        // threshold filter and CC flags (k_decide)
        for (uint32_t c = t; c < nc; c += SM_NT) {
            const unsigned long long b = s_best[c];
            const float sp = unord_f32((uint32_t)(b >> 32));
            const uint32_t off = 0xFFFFFFFFu - (uint32_t)b;
            a.csp[c] = sp;
            a.coff[c] = off;
            uint32_t fl = T.cflag[c] & CF_CONN;
            if ((double)sp < a.sc_thr) fl |= CF_CUT;
            if ((fl & CF_CUT) || !(fl & CF_CONN)) fl |= CF_CC;
            T.cflag[c] = fl;
            if (fl & CF_CC) atomicAdd(&s_cnt[0], 1u);
        }
        __syncthreads();
        const uint32_t ncc = s_cnt[0];
        __syncthreads();
        // ------------------------------------------------------------ cut
        ++iter1;
        uint32_t nl = 0;
        if (ncc > 0) {
            uint32_t* s_part = s_u32;
            uint32_t* s_root = s_u32 + SMALL_MAX_N;
            for (uint32_t u = t; u < a.n; u += SM_NT) s_part[u] = INVALID;
            __syncthreads();
            // parts of the CC clusters (k_mark)
            for (uint32_t k = t; k < na; k += SM_NT) {
                const uint32_t c = pclu[k];
                const uint32_t fl = T.cflag[c];
                if (fl & CF_CC) {
                    const uint32_t u = pnix[k];
                    const uint32_t side = ((fl & CF_CUT) && (k - T.cs[c] + 1u > a.coff[c])) ? 1u : 0u;
                    s_part[u] = (c << 1) | side;
                    s_root[u] = u;
                    a.stamp[u] = iter1;
                }
            }
            __syncthreads();
            // members in node order
            for (uint32_t e = t; e < nalist; e += SM_NT) a.cclist[e] = alist[e];
            __syncthreads();
            nl = sm_compact(a.cclist, nalist, [&](uint32_t, uint32_t u) { return s_part[u] != INVALID; }, s_red);
            // components: every same-part arc joins its ends' trees
            for (uint32_t e = t; e < nl; e += SM_NT) {
                const uint32_t u = a.cclist[e];
                const uint32_t st = a.row[u], d = a.dint[u], pu = s_part[u];
                if (d > SM_LONG) continue;
                for (uint32_t j = 0; j < d; j += 4) {
                    uint32_t v[4];
#pragma unroll
                    for (uint32_t z = 0; z < 4; ++z) v[z] = a.nbr[st + min(j + z, d - 1u)];
#pragma unroll
                    for (uint32_t z = 0; z < 4; ++z)
                        if (j + z < d && s_part[v[z]] == pu) sm_union(s_root, u, v[z]);
                }
            }
            for (uint32_t b = w * WAVE; b < nl; b += SM_NT) {
                const uint32_t e = b + l;
                const uint32_t u = e < nl ? a.cclist[e] : 0u;
                const uint32_t d = e < nl ? a.dint[u] : 0u;
                for (uint64_t m = __ballot(d > SM_LONG); m;) {
                    const uint32_t src = sm_take4(m);
                    const uint32_t ur = __shfl(u, (int)(src & 63u)), dr = __shfl(d, (int)(src & 63u));
                    if (src < (uint32_t)WAVE) {
                        const uint32_t st = a.row[ur], pu = s_part[ur];
                        for (uint32_t j = l % SM_GL; j < dr; j += SM_GL) {
                            const uint32_t v = a.nbr[st + j];
                            if (s_part[v] == pu) sm_union(s_root, ur, v);
                        }
                    }
                }
            }
            __syncthreads();
            for (uint32_t e = t; e < nl; e += SM_NT) {
                const uint32_t u = a.cclist[e];
                a.root[u] = sm_find(s_root, u);
            }
            __syncthreads();
            // regroup (k_rid): stable sort of all positions by component /
            // cluster minimum, clusters from the runs of equal keys
            for (uint32_t k = t; k < na; k += SM_NT) {
                const uint32_t c = pclu[k];
                a.pkey[k] = (T.cflag[c] & CF_CC) ? a.root[pnix[k]] : T.cmin[c];
            }
            if (t == 0) { s_segs[0] = 0u; s_segs[1] = na; }
            __syncthreads();
            wg_seg_sort<SM_NT, SM_ITEMS>(1u, s_segs, a.pkey, pnix, s_buf, s_lab, s_wcnt, s_hs, s_red, s_bits);
            {
                const uint32_t k0 = t * SM_ITEMS;
                uint32_t c = 0u;
                for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
                    const uint32_t k = k0 + q;
                    if (k < na && (k == 0 || a.pkey[k] != a.pkey[k - 1])) ++c;
                }
                uint32_t tot;
                uint32_t id = sm_scan(c, s_red, tot);
                for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
                    const uint32_t k = k0 + q;
                    if (k >= na) break;
                    const uint32_t r = a.pkey[k];
                    if (k == 0 || r != a.pkey[k - 1]) {
                        T.cs[id] = k;
                        T.cmin[id] = r;
                        T.cgraph[id] = 0u;
                        T.cflag[id] = CF_CONN;
                        ++id;
                    }
                    pclu[k] = id - 1u;
                }
                if (t == 0) T.cs[tot] = na;
                nc = tot;
            }
            __syncthreads();
        }
        // ------------------------------------------------------------ cluster statistics
        unsigned long long* s_sum = reinterpret_cast<unsigned long long*>(s_buf);
        for (uint32_t c = t; c < nc; c += SM_NT) {
            const uint32_t b = T.cs[c], e = T.cs[c + 1];
            const uint32_t lo = a.xkey[pnix[b]], hi = a.xkey[pnix[e - 1u]];
            a.cmino[c] = lo;
            a.cmaxo[c] = hi;
            const float amax = fmaxf(fabsf(unord_f32(hi)), fabsf(unord_f32(lo)));
            int E = 0;
            if (amax != 0.0f) frexpf(amax, &E);
            a.cK[c] = 62 - E - ceil_log2_u32(e - b);
            s_sum[c] = 0ull;
        }
        __syncthreads();
        {
            const double* xc = a.x[cur];
            const uint32_t k0 = t * SM_ITEMS;
            uint32_t runc = 0xFFFFFFFFu;
            int64_t run = 0;
            for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
                const uint32_t k = k0 + q;
                if (k >= na) break;
                const uint32_t c = pclu[k];
                const int64_t term = (int64_t)ldexp((double)(float)xc[pnix[k]], a.cK[c]);
                if (c != runc) {
                    if (runc != 0xFFFFFFFFu) atomicAdd(&s_sum[runc], (unsigned long long)run);
                    runc = c;
                    run = term;
                } else {
                    run += term;
                }
            }
            if (runc != 0xFFFFFFFFu) atomicAdd(&s_sum[runc], (unsigned long long)run);
        }
        __syncthreads();
        uint32_t nsurv = 0u;
        int32_t labmin = 0x7FFFFFFF;
        {   // keep flags, averages (k_stats_final), survivors' ranks, labels (k_labels)
            const uint32_t c0 = t * SM_ITEMS;
            uint32_t kp = 0u, cntk = 0u;
            for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
                const uint32_t c = c0 + q;
                if (c >= nc) break;
                const uint32_t count = T.cs[c + 1] - T.cs[c];
                const float rw_sum = (float)ldexp(i64_to_f64((int64_t)s_sum[c]), -a.cK[c]);
                const float pot = unord_f32(a.cmaxo[c]) - unord_f32(a.cmino[c]);
                const uint32_t keep = (pot >= a.rw_thr && count >= 2u) ? 1u : 0u;
                a.ckeep[c] = keep;
                a.cavg[c] = (double)rw_sum / (double)count;
                if (keep) { kp |= 1u << q; ++cntk; }
            }
            uint32_t tot;
            uint32_t r = sm_scan(cntk, s_red, tot);
            for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
                const uint32_t c = c0 + q;
                if (c >= nc) break;
                a.kscan[c] = r;
                if (kp & (1u << q)) {
                    T.clabel[c] = (int32_t)r;
                    ++r;
                } else {
                    const int32_t lab = gsmall - (int32_t)c - 1;
                    T.clabel[c] = lab;
                    labmin = min(labmin, lab);
                }
            }
            if (t == 0) { a.kscan[nc] = tot; s_lab32 = 0; }
            nsurv = tot;
        }
        __syncthreads();
        if (labmin != 0x7FFFFFFF) atomicMin(&s_lab32, labmin);
        __syncthreads();
        // graph counters (k_graph_update; the graph had clusters this
        // iteration, it joined at iteration 0): the smallest new label, and
        // the iteration count once no cluster survives
        gsmall = s_lab32;
        if (nsurv == 0u && giter == 0) giter = (int32_t)iter1 + 1;
        const uint32_t nc_new = nsurv;
        // cluster of every node, labels, recentring, deactivation (the node
        // clusters k_stats_sum records, k_apply_nodes)
        for (uint32_t k = t; k < na; k += SM_NT) a.nclu[pnix[k]] = pclu[k];
        __syncthreads();
        for (uint32_t e = t; e < nalist; e += SM_NT) {
            const uint32_t u = alist[e];
            const uint32_t c = a.nclu[u];
            a.label[u] = T.clabel[c];
            if (a.ckeep[c]) {
                const double v = a.x[cur][u] - a.cavg[c];
                a.x[cur][u] = v;
                const uint32_t d = a.dint[u];
                a.y[cur][u] = fix_pack(d ? v / (double)d : 0.0, ovf);
            } else {
                a.alive[u] = 0;
            }
        }
        __syncthreads();
        // compaction into the other buffers (SelPositions, k_clu_compact):
        // positions and clusters of the kept clusters, the alive nodes
        uint32_t na_new;
        {
            const uint32_t k0 = t * SM_ITEMS;
            uint32_t ck = 0u;
            for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
                const uint32_t k = k0 + q;
                if (k >= na) break;
                ck += a.ckeep[pclu[k]];
            }
            uint32_t tot;
            uint32_t o = sm_scan(ck, s_red, tot);
            uint32_t* const pnix2 = a.pnix[pb ^ 1u];
            uint32_t* const pclu2 = a.pclu[pb ^ 1u];
            for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
                const uint32_t k = k0 + q;
                if (k >= na) break;
                const uint32_t c = pclu[k];
                a.flags[k] = o;                // new index of a kept position
                if (a.ckeep[c]) {
                    pnix2[o] = pnix[k];
                    pclu2[o] = a.kscan[c];
                    ++o;
                }
            }
            na_new = tot;
            __syncthreads();
            const ClusterCols T2 = a.ct[tb ^ 1u];
            for (uint32_t c = t; c < nc; c += SM_NT) {
                if (!a.ckeep[c]) continue;
                const uint32_t d = a.kscan[c];
                T2.cs[d] = a.flags[T.cs[c]];
                T2.cgraph[d] = T.cgraph[c];
                T2.cmin[d] = T.cmin[c];
                T2.cflag[d] = T.cflag[c] & CF_CONN;
                T2.clabel[d] = T.clabel[c];
                T2.cvol[d] = 0u;
            }
            if (t == 0) T2.cs[nc_new] = na_new;
            pb ^= 1u;
            tb ^= 1u;
        }
        if (na_new != na) {
            const uint32_t k0 = t * SM_ITEMS;
            uint32_t ck = 0u;
            for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
                const uint32_t e = k0 + q;
                if (e >= nalist) break;
                ck += a.alive[alist[e]];
            }
            uint32_t tot;
            uint32_t o = sm_scan(ck, s_red, tot);
            uint32_t* const alist2 = a.alist[ab ^ 1u];
            for (uint32_t q = 0; q < (uint32_t)SM_ITEMS; ++q) {
                const uint32_t e = k0 + q;
                if (e >= nalist) break;
                const uint32_t u = alist[e];
                if (a.alive[u]) alist2[o++] = u;
            }
            nalist = tot;
            ab ^= 1u;
        }
        na = na_new;
        nc = nc_new;
        __syncthreads();
        // arc disabling for the surviving members (k_disable): keep arcs whose
        // ends have one root, compacted in order to the row front
        if (nl) {
            // short rows: one thread walks the row in order; a kept arc moves
            // to an index at most its own, all of which are already read
            for (uint32_t e = t; e < nl; e += SM_NT) {
                const uint32_t u = a.cclist[e];
                const uint32_t d = a.dint[u];
                if (!a.alive[u] || d > SM_LONG) continue;
                const uint32_t st = a.row[u], ru = a.root[u];
                uint32_t kept = 0;
                for (uint32_t j = 0; j < d; j += 4) {
                    uint32_t v[4], rv[4];
#pragma unroll
                    for (uint32_t z = 0; z < 4; ++z) v[z] = a.nbr[st + min(j + z, d - 1u)];
#pragma unroll
                    for (uint32_t z = 0; z < 4; ++z) rv[z] = a.root[v[z]];
#pragma unroll
                    for (uint32_t z = 0; z < 4; ++z) {
                        if (j + z < d && rv[z] == ru) {
                            if (kept != j + z) a.nbr[st + kept] = v[z];
                            ++kept;
                        }
                    }
                }
                if (kept != d) {
                    a.dint[u] = kept;
                    a.y[cur][u] = fix_pack(kept ? a.x[cur][u] / (double)kept : 0.0, ovf);
                }
            }
            // long rows: 16 lanes each, 16 arcs at a time ranked by a group ballot
            for (uint32_t b = w * WAVE; b < nl; b += SM_NT) {
                const uint32_t e = b + l;
                const uint32_t u = e < nl ? a.cclist[e] : 0u;
                const uint32_t d = e < nl ? a.dint[u] : 0u;
                const bool mine = e < nl && a.alive[u] && d > SM_LONG;
                for (uint64_t m = __ballot(mine); m;) {
                    const uint32_t src = sm_take4(m);
                    const uint32_t ur = __shfl(u, (int)(src & 63u)), dr = __shfl(d, (int)(src & 63u));
                    if (src < (uint32_t)WAVE) {            // uniform within the 16-lane group
                        const uint32_t st = a.row[ur], ru = a.root[ur], gl = l % SM_GL;
                        uint32_t kept = 0;
                        for (uint32_t j0 = 0; j0 < dr; j0 += SM_GL) {
                            const uint32_t j = j0 + gl;
                            const uint32_t v = a.nbr[st + min(j, dr - 1u)];
                            const bool keep = j < dr && a.root[v] == ru;
                            const uint64_t bal = group_ballot<SM_GL>(keep);
                            const uint32_t before = (uint32_t)__popcll(bal & ((1ull << gl) - 1ull));
                            if (keep && kept + before != j) a.nbr[st + kept + before] = v;
                            kept += (uint32_t)__popcll(bal);
                        }
                        if (gl == 0 && kept != dr) {
                            a.dint[ur] = kept;
                            a.y[cur][ur] = fix_pack(kept ? a.x[cur][ur] / (double)kept : 0.0, ovf);
                        }
                    }
                }
            }
            __syncthreads();
        }
        // cluster volumes (k_volume)
        {
            uint32_t* s_vol = s_u32;
            const uint32_t* pn = a.pnix[pb];
            const uint32_t* pc = a.pclu[pb];
            const ClusterCols Tn = a.ct[tb];
            for (uint32_t c = t; c < nc; c += SM_NT) s_vol[c] = 0u;
            __syncthreads();
            for (uint32_t k = t; k < na; k += SM_NT) atomicAdd(&s_vol[pc[k]], a.dint[pn[k]]);
            __syncthreads();
            for (uint32_t c = t; c < nc; c += SM_NT) Tn.cvol[c] = s_vol[c];
            __syncthreads();
        }
    }
    if (ovf) atomicOr(&a.cnt->err, ERR_FIX_RANGE);
    if (t == 0) {
        *a.gsmall = gsmall;
        *a.giter = giter;
        a.out[0] = na;
        a.out[1] = nc;
        a.out[2] = nalist;
        a.out[3] = cur;
        a.out[4] = iter1;
        a.out[5] = i;
        a.out[6] = pb;
        a.out[7] = (tb << 1) | ab;
    }
}

}  // namespace pxd

/*
 * parex expander-decomposition ORACLE -- test infrastructure only (see
 * oracle.h for scope, the canonical-arithmetic contract and what pins it).
 * Every function cites the reference lines it restates.  Build: oracle/Makefile.
 */
#include "oracle.h"

#include <omp.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef unsigned __int128 u128;

#define RW_STAY 0.1                      /* lib/core/definitions.h:48 */
#define FIX_SHIFT 72                     /* W1 scale 2^72 */

struct ox_state {
    uint32_t n, m;
    uint32_t* ranges;     /* u32[n+1]                         cuda/devGraph.h:89-91 */
    uint32_t* nbr;        /* u32[2m], INVALID_EDGE marks      cuda/devGraph.h:88 */
    uint32_t* deg;        /* original degree                  cuda/devGraph.h:135 */
    uint32_t* dint;       /* allInternalDegrees               cuda/devPartition.h:235 */
    int32_t* labels;      /* allLabels                        cuda/devPartition.h:234 */
    double* dist[2];      /* walk double buffer               cuda/devRandomWalk.h:341-343 */
    int cur;
    ox_lnode* act;        /* activeNodes                      cuda/devPartition.h:237 */
    uint32_t na;          /* numActiveNodes */
    uint32_t nac;         /* numActiveClusters */
    uint64_t* keys;       /* packedKeys (per active position) */
    ox_ndata* nd_pre;     /* scNodeData before sort (stale entries kept like the reference) */
    ox_ndata* nd;         /* scanned NodeData aligned with act */
    int32_t* clabels;     /* clusterLabels */
    uint32_t* cvol;       /* clusterVolumes */
    ox_cdata* cdata;      /* clusterData */
    int32_t* lookup;      /* labelLookup */
    ox_sweepcut* scbuf;   /* sweepCutsBuffer */
    ox_sweepcut* sccut;   /* sweepCuts (filtered) */
    uint32_t ncut;        /* numClustersWithCut */
    int32_t smallest;     /* *d_smallestLabel */
    float sc_threshold, rw_threshold;
    int err;
    /* scratch */
    uint32_t* idx;
    uint32_t* tmp;
    uint32_t* uf;
    double* y;            /* x[v] / dInt[v] of the current walk step */
    uint64_t* nkey;       /* packed key of every node for the cut contribution */
};

/* OpenMP threads of the calling thread's later oracle calls (each per-node
 * loop writes only its own node's outputs: any count gives the same result) */
void ox_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

/* ------------------------------------------------------------------ init RNG */

/* cuda/devRandomWalk.h:29-34 */
uint32_t ox_hash(uint32_t x) {
    x = ((x >> 16) ^ x) * 0x45d9f3bu;
    x = ((x >> 16) ^ x) * 0x45d9f3bu;
    x = (x >> 16) ^ x;
    return x;
}

/* thrust::minstd_rand (default_random_engine) seeded with s, first draw:
 * linear_congruential_engine<u32, 48271, 0, 2^31-1>; seed(s) = s mod m, 0 -> 1 */
uint32_t ox_minstd_first(uint32_t seed) {
    const uint64_t M = 2147483647u;
    uint64_t x = seed % M;
    if (x == 0) x = 1;
    x = (x * 48271u) % M;
    return (uint32_t)x;
}

/* erfcinv in FP64: Abramowitz-Stegun 26.2.23 start + Halley steps on erfc.
 * Own implementation from the mathematics (not CUDA's erfcinvf). */
double ox_erfcinv(double y) {
    if (y <= 0.0) return INFINITY;
    if (y >= 2.0) return -INFINITY;
    int flip = 0;
    if (y > 1.0) { y = 2.0 - y; flip = 1; }
    double p = 0.5 * y;                       /* upper tail of the standard normal */
    double t = sqrt(-2.0 * log(p));
    double z = t - (2.515517 + 0.802853 * t + 0.010328 * t * t) /
                   (1.0 + 1.432788 * t + 0.189269 * t * t + 0.001308 * t * t * t);
    double x = z / sqrt(2.0);
    for (int it = 0; it < 8; ++it) {
        double f = erfc(x) - y;
        double fp = -2.0 / sqrt(M_PI) * exp(-x * x);
        double dx = f / fp;
        /* Halley correction: f'' = -2x f' */
        double denom = 1.0 + dx * x;
        x = x - dx / denom;
    }
    return flip ? -x : x;
}

/* cuda/devRandomWalk.h:36-47 with thrust normal_distribution_nvcc<float>::sample
 * (rocThrust random/detail/normal_distribution_base.h): mean 0,
 * stddev = (float)sqrt((double)deg). */
float ox_init_value(uint32_t idx, uint32_t deg, uint32_t seed) {
    const uint32_t range = 2147483645u;                 /* max - min of minstd */
    const float S1 = (float)(1.0 / (double)range);
    const float S2 = S1 / 2;
    float S3 = (float)(-1.4142135623730950488016887242097);
    uint32_t s = ox_hash(seed ^ idx);
    uint32_t u = ox_minstd_first(s) - 1u;
    if (u > range / 2) { u = range - u; S3 = -S3; }
    float p = fmaf((float)u, S1, S2);                   /* contracted as nvcc does */
    float e = (float)ox_erfcinv((double)(2.0f * p));
    float sigma = (float)sqrt((double)deg);
    return (sigma * S3) * e;
}

/* cuda/devRandomWalk.h:509-519: dist.Current() <- functor(i), stored as double */
void ox_init_vector(uint32_t n, const uint32_t* degrees, uint32_t seed, double* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = (double)ox_init_value(i, degrees[i], seed);
}

/* cuda/devRandomWalk.h:50-61 */
uint64_t ox_pack_key(int32_t label, float v) {
    uint32_t i;
    memcpy(&i, &v, 4);
    uint32_t mask = (i & 0x80000000u) ? 0xffffffffu : 0x80000000u;
    uint32_t ordered = i ^ mask;
    uint32_t olab = (uint32_t)label ^ 0x80000000u;
    return ((uint64_t)olab << 32) | ordered;
}

static int32_t key_label(uint64_t k) {           /* cuda/devSweepCut.h:308-312 */
    return (int32_t)((uint32_t)(k >> 32) ^ 0x80000000u);
}

/* ------------------------------------------------------------- W1 fixed point */

static void fix_add(i128* acc, double y, int* ovf) {
    uint64_t b;
    memcpy(&b, &y, 8);
    uint32_t E = (uint32_t)((b >> 52) & 0x7FFu);
    if (E == 0) return;                          /* zero / subnormal -> truncates to 0 */
    if (E == 0x7FF) { *ovf = 1; return; }
    uint64_t M = (b & ((1ull << 52) - 1)) | (1ull << 52);
    int sh = (int)E - (1075 - FIX_SHIFT);        /* y*2^72 = M*2^(E-1075+72) */
    u128 q;
    if (sh >= 0) {
        if (sh > 50) { *ovf = 1; return; }       /* |y| >= 2^31 */
        q = (u128)M << sh;
    } else {
        int r = -sh;
        q = (r >= 53) ? 0 : (u128)(M >> r);
    }
    if (b >> 63) *acc -= (i128)q; else *acc += (i128)q;
}

static double fix_to_double(i128 acc) {
    int neg = acc < 0;
    u128 a = neg ? (u128)0 - (u128)acc : (u128)acc;
    double d;
    if ((a >> 63) == 0) {
        d = (double)(int64_t)(uint64_t)a;        /* correctly rounded */
    } else {
        int nb = 0;
        u128 t = a;
        while (t) { ++nb; t >>= 1; }
        int k = nb - 63;
        uint64_t top = (uint64_t)(a >> k);
        if (a & (((u128)1 << k) - 1)) top |= 1;  /* sticky bit */
        d = ldexp((double)(int64_t)top, k);
    }
    if (neg) d = -d;
    return ldexp(d, -FIX_SHIFT);
}

double ox_fixsum(const double* terms, uint32_t k, int* overflow) {
    i128 acc = 0;
    int ovf = 0;
    for (uint32_t i = 0; i < k; ++i) fix_add(&acc, terms[i], &ovf);
    if (overflow) *overflow = ovf;
    return fix_to_double(acc);
}

/* ------------------------------------------------------------- C1 cluster sum */

static int ceil_log2(uint32_t k) {
    int L = 0;
    while ((1ull << L) < (uint64_t)k) ++L;
    return L;
}

float ox_cluster_sum(const float* vals, uint32_t k) {
    float amax = 0.0f;
    for (uint32_t i = 0; i < k; ++i) {
        float a = fabsf(vals[i]);
        if (a > amax) amax = a;
    }
    if (k == 0 || amax == 0.0f) return 0.0f;
    int E;
    frexpf(amax, &E);                            /* amax < 2^E */
    int K = 62 - E - ceil_log2(k);
    int64_t S = 0;
    for (uint32_t i = 0; i < k; ++i) S += (int64_t)ldexp((double)vals[i], K);
    return (float)ldexp((double)S, -K);
}

/* ------------------------------------------------------------- stable sorts */

/* Stable sort of idx[0..n) by key[idx[i]] (LSD radix over 16-bit digits,
 * digits constant over the input skipped).  Any stable sort gives the same
 * order, which is what the reference's CUB radix sort produces. */
static void rsort_u64(uint32_t* idx, uint32_t* tmp, uint32_t n, const uint64_t* key) {
    if (n < 2) return;
    uint64_t orv = 0, andv = ~0ull;
    for (uint32_t i = 0; i < n; ++i) { orv |= key[idx[i]]; andv &= key[idx[i]]; }
    const uint64_t vary = orv ^ andv;
    uint32_t* cnt = (uint32_t*)malloc(65536u * 4u);
    uint32_t *src = idx, *dst = tmp;
    for (int sh = 0; sh < 64; sh += 16) {
        if (((vary >> sh) & 0xFFFFu) == 0) continue;
        memset(cnt, 0, 65536u * 4u);
        for (uint32_t i = 0; i < n; ++i) ++cnt[(key[src[i]] >> sh) & 0xFFFFu];
        uint32_t run = 0;
        for (uint32_t d = 0; d < 65536u; ++d) { uint32_t c = cnt[d]; cnt[d] = run; run += c; }
        for (uint32_t i = 0; i < n; ++i) dst[cnt[(key[src[i]] >> sh) & 0xFFFFu]++] = src[i];
        uint32_t* t = src; src = dst; dst = t;
    }
    if (src != idx) memcpy(idx, src, (size_t)n * sizeof(uint32_t));
    free(cnt);
}

/* stable sort of idx by a[idx[i]].clusterId (signed) */
static void rsort_lab(uint32_t* idx, uint32_t* tmp, uint32_t n, const ox_lnode* a) {
    uint64_t* k = (uint64_t*)malloc((size_t)(n ? n : 1) * 8);
    for (uint32_t i = 0; i < n; ++i) k[i] = (uint64_t)((uint32_t)a[i].clusterId ^ 0x80000000u);
    rsort_u64(idx, tmp, n, k);
    free(k);
}

/* ------------------------------------------------------------- construction */

ox_state* ox_create(uint32_t n, uint32_t m, const uint32_t* ranges, const uint32_t* edges,
                    uint32_t seed, float sc_threshold, float rw_threshold,
                    const double* init_values) {
    ox_state* s = (ox_state*)calloc(1, sizeof(ox_state));
    size_t nn = n ? n : 1, aa = (m != 0) ? 2ull * m : 1;
    s->n = n; s->m = m;
    s->ranges = (uint32_t*)malloc((nn + 1) * 4);
    s->nbr = (uint32_t*)malloc(aa * 4);
    s->deg = (uint32_t*)malloc(nn * 4);
    s->dint = (uint32_t*)malloc(nn * 4);
    s->labels = (int32_t*)calloc(nn, 4);
    s->dist[0] = (double*)calloc(nn, 8);
    s->dist[1] = (double*)calloc(nn, 8);
    s->act = (ox_lnode*)malloc(nn * sizeof(ox_lnode));
    s->keys = (uint64_t*)calloc(nn, 8);
    s->nd_pre = (ox_ndata*)calloc(nn, sizeof(ox_ndata));
    s->nd = (ox_ndata*)calloc(nn, sizeof(ox_ndata));
    s->clabels = (int32_t*)calloc(nn, 4);
    s->cvol = (uint32_t*)malloc(nn * 4);
    s->cdata = (ox_cdata*)calloc(nn, sizeof(ox_cdata));
    s->lookup = (int32_t*)malloc(nn * 4);
    s->scbuf = (ox_sweepcut*)calloc(nn, sizeof(ox_sweepcut));
    s->sccut = (ox_sweepcut*)calloc(nn, sizeof(ox_sweepcut));
    s->idx = (uint32_t*)malloc(nn * 4);
    s->tmp = (uint32_t*)malloc(nn * 4);
    s->uf = (uint32_t*)malloc(nn * 4);
    s->y = (double*)malloc(nn * 8);
    s->nkey = (uint64_t*)malloc(nn * 8);
    memcpy(s->ranges, ranges, (size_t)(n + 1) * 4);
    if (m) memcpy(s->nbr, edges, 2ull * m * 4);
    /* GraphManager ctor, cuda/devGraph.h:125-139 */
    for (uint32_t i = 0; i < n; ++i) s->deg[i] = ranges[i + 1] - ranges[i];
    /* PartitionManager ctor, cuda/devPartition.h:260-329 */
    for (uint32_t i = 0; i < n; ++i) {
        s->act[i].nix = i;
        s->act[i].clusterId = 0;
        s->dint[i] = s->deg[i];
        s->cvol[i] = 2 * m;
        s->lookup[i] = -1;
    }
    s->na = n;
    s->nac = 1;
    /* SweepCutManager ctor, cuda/devSweepCut.h:232-249 */
    s->smallest = 0x7fffffff;
    s->sc_threshold = sc_threshold;
    s->rw_threshold = rw_threshold;
    /* RandomWalkManager ctor, cuda/devRandomWalk.h:351-359 */
    s->cur = 0;
    if (init_values) memcpy(s->dist[0], init_values, (size_t)n * 8);
    else ox_init_vector(n, s->deg, seed, s->dist[0]);
    return s;
}

void ox_destroy(ox_state* s) {
    if (!s) return;
    free(s->ranges); free(s->nbr); free(s->deg); free(s->dint); free(s->labels);
    free(s->dist[0]); free(s->dist[1]); free(s->act); free(s->keys); free(s->nd_pre);
    free(s->nd); free(s->clabels); free(s->cvol); free(s->cdata); free(s->lookup);
    free(s->scbuf); free(s->sccut); free(s->idx); free(s->tmp); free(s->uf);
    free(s->y); free(s->nkey);
    free(s);
}

/* ------------------------------------------------------------- walk (K1) */

/* cuda/devRandomWalk.h:117-203 (fusedRandomWalk_WarpParallel), launch 369-392 */
void ox_walk(ox_state* s) {
    const double* x = s->dist[s->cur];
    double* xn = s->dist[s->cur ^ 1];
    /* the term x[nb] / (double)dInt[nb] of every arc into nb, divided once per
     * node (the same IEEE quotient; terms with dInt = 0 are skipped below) */
    const int64_t n = s->n, na = s->na;
    int err = 0;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < n; ++v) s->y[v] = s->dint[v] ? x[v] / (double)s->dint[v] : 0.0;
    /* every position writes only its own node's value and its own key, and
     * the neighbour sum is exact: any thread split gives the same bits */
#pragma omp parallel for schedule(dynamic, 2048) reduction(| : err)
    for (int64_t i = 0; i < na; ++i) {
        uint32_t nix = s->act[i].nix;
        int32_t lab = s->act[i].clusterId;
        if (lab < 0) { s->keys[i] = ox_pack_key(lab, 0.0f); continue; }
        uint32_t start = s->ranges[nix], end = start + s->deg[nix];
        if (s->deg[nix] > (1u << 24)) err |= 2;
        i128 acc = 0;
        int ovf = 0;
        for (uint32_t j = start; j < end; ++j) {
            uint32_t nb = s->nbr[j];
            if (nb == OX_INVALID_EDGE) continue;
            if (s->dint[nb] == 0) { err |= 4; continue; }   /* reference prints ERROR and skips */
            fix_add(&acc, s->y[nb], &ovf);
        }
        if (ovf) err |= 1;
        double sum = fix_to_double(acc);
        double v = sum * (1.0 - RW_STAY) + x[nix] * RW_STAY;
        xn[nix] = v;
        s->keys[i] = ox_pack_key(lab, (float)v);
    }
    s->err |= err;
    s->cur ^= 1;
}

/* ------------------------------------------------------------- sweep */

/* SweepCutManager::prepare_nodes, cuda/devSweepCut.h:469-584 and
 * SweepCutManager::compute_sweep_cuts, cuda/devSweepCut.h:587-730 */
void ox_sweep(ox_state* s) {
    const double* x = s->dist[s->cur];
    uint32_t na = s->na;
    /* every neighbour's key (cuda/devSweepCut.h:78-81), packed once per node */
    const int64_t nn = s->n, nna = na;
    int err = 0;
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < nn; ++v) s->nkey[v] = ox_pack_key(s->labels[v], (float)x[v]);
    /* K4 nodeDiffKernel_Sparse_WarpParallel, cuda/devSweepCut.h:20-112 (each
     * position writes only its own NodeData) */
#pragma omp parallel for schedule(dynamic, 2048) reduction(| : err)
    for (int64_t i = 0; i < nna; ++i) {
        uint32_t nix = s->act[i].nix;
        int32_t lab = s->act[i].clusterId;
        if (lab < 0) continue;                            /* stale NodeData stays */
        uint64_t myKey = s->keys[i];
        uint32_t start = s->ranges[nix], end = start + s->deg[nix];
        int32_t c = 0;
        for (uint32_t j = start; j < end; ++j) {
            uint32_t nb = s->nbr[j];
            if (nb == OX_INVALID_EDGE) continue;
            uint64_t nk = s->nkey[nb];
            if (s->labels[nb] != lab) err |= 8;           /* reference prints ERROR */
            int before = (nk != myKey) ? (nk < myKey) : (nb < nix);
            c += before ? -1 : 1;
        }
        s->nd_pre[i].nix = nix;
        s->nd_pre[i].prefixEdgeDiff = c;
        s->nd_pre[i].prefixVolume = s->dint[nix];
        s->nd_pre[i].offsetInCluster = 1;
    }
    s->err |= err;
    /* K5: stable LSD radix sort of (key, NodeData) pairs over all 64 bits */
    for (uint32_t i = 0; i < na; ++i) s->idx[i] = i;
    rsort_u64(s->idx, s->tmp, na, s->keys);
    /* apply permutation to keys and NodeData */
    uint64_t* sk = (uint64_t*)malloc((size_t)(na ? na : 1) * 8);
    ox_ndata* snd = (ox_ndata*)malloc((size_t)(na ? na : 1) * sizeof(ox_ndata));
    for (uint32_t i = 0; i < na; ++i) { sk[i] = s->keys[s->idx[i]]; snd[i] = s->nd_pre[s->idx[i]]; }
    memcpy(s->keys, sk, (size_t)na * 8);
    memcpy(s->nd_pre, snd, (size_t)na * sizeof(ox_ndata));
    free(sk); free(snd);
    /* K6: smallest label, lower_bound(0), compaction (530-575) */
    if (na > 0) s->smallest = key_label(s->keys[0]);
    uint32_t inactive = 0;
    while (inactive < na && key_label(s->keys[inactive]) < 0) ++inactive;
    na -= inactive;
    s->na = na;
    for (uint32_t k = 0; k < na; ++k) {
        s->act[k].nix = s->nd_pre[k + inactive].nix;
        s->act[k].clusterId = key_label(s->keys[k + inactive]);
    }
    /* K7 inclusive_scan_by_key with NodeDataScanOp (598-610) */
    for (uint32_t k = 0; k < na; ++k) {
        ox_ndata v = s->nd_pre[k + inactive];
        if (k > 0 && s->act[k].clusterId == s->act[k - 1].clusterId) {
            v.prefixEdgeDiff += s->nd[k - 1].prefixEdgeDiff;
            v.prefixVolume += s->nd[k - 1].prefixVolume;
            v.offsetInCluster += s->nd[k - 1].offsetInCluster;
        }
        s->nd[k] = v;
    }
    /* K8 reduce_by_key of internal degrees -> clusterLabels, clusterVolumes (627-641) */
    uint32_t C = 0;
    for (uint32_t k = 0; k < na; ++k) {
        int32_t lab = s->act[k].clusterId;
        uint32_t d = s->dint[s->act[k].nix];
        if (k == 0 || lab != s->act[k - 1].clusterId) { s->clabels[C] = lab; s->cvol[C] = d; ++C; }
        else s->cvol[C - 1] += d;
    }
    s->nac = C;
    /* K9 updateLabelLookup (cuda/devPartition.h:714-732) */
    for (uint32_t c = 0; c < C; ++c) {
        int32_t lab = s->clabels[c];
        if (lab >= 0 && (uint32_t)lab < s->n) s->lookup[lab] = (int32_t)c;
        else s->err |= 16;
    }
    /* K10 SweepCutTransform + ArgMinOp reduce_by_key (374-433, 356-361, 665-688):
     * (a.s < b.s) ? a : b folds to the rightmost minimum */
    C = 0;
    for (uint32_t k = 0; k < na; ++k) {
        int32_t lab = s->act[k].clusterId;
        ox_ndata v = s->nd[k];
        ox_sweepcut sc;
        int32_t seg = -1;
        if (lab >= 0 && (uint32_t)lab < s->n) seg = s->lookup[lab];
        if (seg < 0 || (uint32_t)seg >= s->nac || s->clabels[seg] != lab) {
            s->err |= 32;
            sc.clusterId = lab; sc.sparsity = 2.0f; sc.offset = v.offsetInCluster;
        } else {
            uint32_t tot = s->cvol[seg];
            uint32_t ed = (uint32_t)v.prefixEdgeDiff;
            uint32_t pv = v.prefixVolume;
            uint32_t denom = (pv >= tot) ? 0 : ((pv < tot - pv) ? pv : (tot - pv));
            sc.clusterId = lab;
            sc.sparsity = (denom > 0) ? ((float)ed / (float)denom) : 2.0f;
            sc.offset = v.offsetInCluster;
        }
        if (k == 0 || lab != s->act[k - 1].clusterId) { s->scbuf[C++] = sc; }
        else {
            ox_sweepcut a = s->scbuf[C - 1];
            s->scbuf[C - 1] = (a.sparsity < sc.sparsity) ? a : sc;
        }
    }
    /* K11 copy_if sparsity < (double)sc_threshold (701-709) */
    double thr = (double)s->sc_threshold;
    s->ncut = 0;
    for (uint32_t c = 0; c < s->nac; ++c)
        if ((double)s->scbuf[c].sparsity < thr) s->sccut[s->ncut++] = s->scbuf[c];
}

/* ------------------------------------------------------------- cut + CC + stats */

static uint32_t uf_find(uint32_t* p, uint32_t a) {
    uint32_t r = a;
    while (p[r] != r) r = p[r];
    while (p[a] != r) { uint32_t nx = p[a]; p[a] = r; a = nx; }
    return r;
}

/* CudaDeviceManager::Impl::cutClusters, cuda/device.h:101-168 */
void ox_cut(ox_state* s) {
    uint32_t na = s->na;
    /* K12 PartitionManager::cutClusters, cuda/devPartition.h:545-622 */
    int err = 0;
    if (s->ncut > 0) {
        int32_t maxLabel = s->clabels[s->nac - 1];
        /* each position updates only itself and its own node */
#pragma omp parallel for schedule(static) reduction(| : err)
        for (int64_t k = 0; k < (int64_t)na; ++k) {
            int32_t lab = s->act[k].clusterId;
            if (lab < 0) { err |= 64; continue; }
            /* lower_bound over sweepCuts by clusterId */
            uint32_t lo = 0, hi = s->ncut;
            while (lo < hi) {
                uint32_t mid = (lo + hi) / 2;
                if (s->sccut[mid].clusterId < lab) lo = mid + 1; else hi = mid;
            }
            if (lo == s->ncut || s->sccut[lo].clusterId != lab) continue;
            if (s->nd[k].offsetInCluster > s->sccut[lo].offset) {
                int64_t upd = (int64_t)lo + (int64_t)maxLabel + 1;
                if (upd > 2147483647) err |= 128;
                s->act[k].clusterId = (int32_t)upd;
                s->labels[s->act[k].nix] = (int32_t)upd;
            }
        }
    }
    s->err |= err;
    /* K13-K17 ConsolidationManager::consolidate, cuda/devConsolidate.h:168-204:
     * components of same-label valid arcs of active nodes; label = min index.
     * Union-find that always hangs the larger root below the smaller: every
     * tree's root is its minimum node whatever the order the arcs are taken
     * in, so the active nodes are visited in index order (memory order). */
    uint8_t* mark = (uint8_t*)calloc(s->n ? s->n : 1, 1);
    for (uint32_t k = 0; k < na; ++k) {
        if (s->act[k].clusterId < 0) continue;
        mark[s->act[k].nix] = 1;
        s->uf[s->act[k].nix] = s->act[k].nix;
    }
    for (uint32_t u = 0; u < s->n; ++u) {
        if (!mark[u]) continue;
        int32_t lab = s->labels[u];
        uint32_t start = s->ranges[u], end = start + s->deg[u];
        for (uint32_t j = start; j < end; ++j) {
            uint32_t v = s->nbr[j];
            if (v == OX_INVALID_EDGE || s->labels[v] != lab) continue;
            uint32_t ru = uf_find(s->uf, u), rv = uf_find(s->uf, v);
            if (ru < rv) s->uf[rv] = ru; else if (rv < ru) s->uf[ru] = rv;
        }
    }
    for (uint32_t u = 0; u < s->n; ++u)
        if (mark[u]) s->labels[u] = (int32_t)uf_find(s->uf, u);
    free(mark);
    for (uint32_t k = 0; k < na; ++k)
        if (s->act[k].clusterId >= 0) s->act[k].clusterId = s->labels[s->act[k].nix];
    /* K19-K20 computeClusterData, cuda/devPartition.h:380-411 */
    for (uint32_t k = 0; k < na; ++k) s->idx[k] = k;
    rsort_lab(s->idx, s->tmp, na, s->act);
    ox_lnode* sorted = (ox_lnode*)malloc((size_t)(na ? na : 1) * sizeof(ox_lnode));
    for (uint32_t k = 0; k < na; ++k) sorted[k] = s->act[s->idx[k]];
    memcpy(s->act, sorted, (size_t)na * sizeof(ox_lnode));
    free(sorted);
    const double* x = s->dist[s->cur];
    float* vals = (float*)malloc((size_t)(na ? na : 1) * sizeof(float));
    uint32_t C = 0, k0 = 0;
    for (uint32_t k = 0; k <= na; ++k) {
        if (k == na || (k > 0 && s->act[k].clusterId != s->act[k - 1].clusterId)) {
            if (k > k0) {
                ox_cdata cd;
                float mx = vals[k0], mn = vals[k0];
                for (uint32_t q = k0; q < k; ++q) {
                    /* ordered-float max/min: -0 < +0 */
                    uint64_t a = ox_pack_key(0, vals[q]) & 0xffffffffu;
                    if (a > (ox_pack_key(0, mx) & 0xffffffffu)) mx = vals[q];
                    if (a < (ox_pack_key(0, mn) & 0xffffffffu)) mn = vals[q];
                }
                cd.rwSum = ox_cluster_sum(vals + k0, k - k0);
                cd.maxPotential = mx;
                cd.minPotential = mn;
                cd.totalElements = k - k0;
                s->clabels[C] = s->act[k0].clusterId;
                s->cdata[C] = cd;
                ++C;
            }
            k0 = k;
        }
        if (k < na) vals[k] = (float)x[s->act[k].nix];
    }
    free(vals);
    s->nac = C;
    for (uint32_t c = 0; c < C; ++c) {
        int32_t lab = s->clabels[c];
        if (lab >= 0 && (uint32_t)lab < s->n) s->lookup[lab] = (int32_t)c; else s->err |= 16;
    }
    /* K21-K22 recenterAndDeactivateClusters, cuda/devPartition.h:624-711 */
    double* xm = s->dist[s->cur];
    int32_t* newlab = (int32_t*)malloc((size_t)(C ? C : 1) * 4);
    int32_t surv = 0;
    for (uint32_t c = 0; c < C; ++c) {
        float pot = s->cdata[c].maxPotential - s->cdata[c].minPotential;
        int keep = (pot >= s->rw_threshold && s->cdata[c].totalElements >= 2);
        newlab[c] = surv;
        surv += keep;
    }
    err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
    for (int64_t k = 0; k < (int64_t)na; ++k) {
        int32_t lab = s->act[k].clusterId;
        if (lab < 0) { err |= 64; continue; }
        int32_t c = s->lookup[lab];
        const ox_cdata* cd = &s->cdata[c];
        float pot = cd->maxPotential - cd->minPotential;
        if (pot < s->rw_threshold || cd->totalElements < 2) {
            int32_t upd = s->smallest - c - 1;
            s->act[k].clusterId = upd;
            s->labels[s->act[k].nix] = upd;
        } else {
            double avg = (double)cd->rwSum / (double)cd->totalElements;
            s->act[k].clusterId = newlab[c];
            s->labels[s->act[k].nix] = newlab[c];
            xm[s->act[k].nix] -= avg;
        }
    }
    s->err |= err;
    free(newlab);
    s->nac = (uint32_t)surv;
    /* K23 disableEdgesKernel_node, cuda/devPartition.h:80-146, 503-522: each
     * position rewrites only its own node's row and internal degree */
#pragma omp parallel for schedule(dynamic, 2048)
    for (int64_t k = 0; k < (int64_t)na; ++k) {
        uint32_t u = s->act[k].nix;
        int32_t lab = s->act[k].clusterId;
        if (lab < 0) continue;
        uint32_t start = s->ranges[u], end = start + s->deg[u];
        uint32_t d = 0;
        for (uint32_t j = start; j < end; ++j) {
            uint32_t v = s->nbr[j];
            if (v == OX_INVALID_EDGE) continue;
            if (s->labels[v] != lab) s->nbr[j] = OX_INVALID_EDGE; else ++d;
        }
        s->dint[u] = d;
    }
}

/* CudaDeviceManager::Impl::expanderDecomposition, cuda/device.h:301-350 */
int ox_run(ox_state* s, int max_iter) {
    int i = 0;
    while (i++ < max_iter && s->nac > 0 && s->na > 0) {
        ox_walk(s);
        ox_sweep(s);
        if (s->na == 0) break;
        ox_cut(s);
    }
    return i;
}

/* cuda/devPartition.h:798-823 */
int ox_final_partition(ox_state* s, int32_t* ids_out) {
    uint32_t n = s->n;
    int32_t* u = (int32_t*)malloc((size_t)(n ? n : 1) * 4);
    memcpy(u, s->labels, (size_t)n * 4);
    /* sort ascending (insertion-free: use merge sort through idx on u32 keys) */
    uint64_t* k = (uint64_t*)calloc((size_t)(n ? n : 1), 8);
    for (uint32_t i = 0; i < n; ++i) { k[i] = (uint64_t)((uint32_t)u[i] ^ 0x80000000u); s->idx[i] = i; }
    rsort_u64(s->idx, s->tmp, n, k);
    int32_t* uniq = (int32_t*)malloc((size_t)(n ? n : 1) * 4);
    int nu = 0;
    for (uint32_t i = 0; i < n; ++i) {
        int32_t v = u[s->idx[i]];
        if (nu == 0 || uniq[nu - 1] != v) uniq[nu++] = v;
    }
    for (uint32_t i = 0; i < n; ++i) {
        int lo = 0, hi = nu;
        while (lo < hi) { int mid = (lo + hi) / 2; if (uniq[mid] < s->labels[i]) lo = mid + 1; else hi = mid; }
        ids_out[i] = lo;
    }
    free(u); free(k); free(uniq);
    return nu;
}

/* cuda/devPartition.h:741-759: m - validArcs/2 */
uint32_t ox_num_cut_edges(ox_state* s) {
    uint64_t valid = 0;
    for (uint64_t j = 0; j < 2ull * s->m; ++j) valid += (s->nbr[j] != OX_INVALID_EDGE);
    return s->m - (uint32_t)(valid / 2);
}

uint32_t ox_num_nodes(ox_state* s) { return s->n; }
uint32_t ox_num_active_nodes(ox_state* s) { return s->na; }
uint32_t ox_num_active_clusters(ox_state* s) { return s->nac; }
uint32_t ox_num_clusters_with_cut(ox_state* s) { return s->ncut; }
int32_t ox_smallest_label(ox_state* s) { return s->smallest; }
int ox_error_flags(ox_state* s) { return s->err; }
void ox_get_values(ox_state* s, double* out) { memcpy(out, s->dist[s->cur], (size_t)s->n * 8); }
void ox_get_labels(ox_state* s, int32_t* out) { memcpy(out, s->labels, (size_t)s->n * 4); }
void ox_get_internal_degrees(ox_state* s, uint32_t* out) { memcpy(out, s->dint, (size_t)s->n * 4); }
void ox_get_neighbors(ox_state* s, uint32_t* out) { memcpy(out, s->nbr, 2ull * s->m * 4); }
void ox_get_active(ox_state* s, ox_lnode* out) { memcpy(out, s->act, (size_t)s->na * sizeof(ox_lnode)); }
void ox_get_ndata(ox_state* s, ox_ndata* out) { memcpy(out, s->nd, (size_t)s->na * sizeof(ox_ndata)); }
void ox_get_sweepcuts(ox_state* s, ox_sweepcut* out) { memcpy(out, s->scbuf, (size_t)s->nac * sizeof(ox_sweepcut)); }
void ox_get_cluster_labels(ox_state* s, int32_t* out) { memcpy(out, s->clabels, (size_t)s->nac * 4); }
void ox_get_cluster_data(ox_state* s, ox_cdata* out) { memcpy(out, s->cdata, (size_t)s->nac * sizeof(ox_cdata)); }

/* ------------------------------------------------------------- graph helpers */

typedef struct { uint32_t fx, tx; } ox_edge;

static int edge_cmp(const void* a, const void* b) {
    const ox_edge* x = (const ox_edge*)a;
    const ox_edge* y = (const ox_edge*)b;
    if (x->fx != y->fx) return x->fx < y->fx ? -1 : 1;
    if (x->tx != y->tx) return x->tx < y->tx ? -1 : 1;
    return 0;
}

/* lib/core/graph.h:147-184.  Ranges of trailing isolated nodes are written as
 * 2m (the reference leaves them 0, which corrupts degrees; documented). */
uint32_t ox_finalize(uint32_t n, uint32_t m, const uint32_t* fx, const uint32_t* tx,
                     uint32_t* ranges_out, uint32_t* edges_out) {
    if (m == 0) return 0;
    ox_edge* full = (ox_edge*)malloc(2ull * m * sizeof(ox_edge));
    for (uint32_t e = 0; e < m; ++e) {
        uint32_t a = fx[e] < tx[e] ? fx[e] : tx[e], b = fx[e] < tx[e] ? tx[e] : fx[e];
        full[2 * e].fx = a; full[2 * e].tx = b;
        full[2 * e + 1].fx = b; full[2 * e + 1].tx = a;
    }
    qsort(full, 2ull * m, sizeof(ox_edge), edge_cmp);
    uint32_t eix = 0;
    for (uint32_t nix = 0; nix <= n; ++nix) {
        ranges_out[nix] = eix;
        while (nix < n && eix < 2 * m && full[eix].fx == nix) ++eix;
    }
    for (uint32_t e = 0; e < 2 * m; ++e) edges_out[e] = full[e].tx;
    free(full);
    return 2 * m;
}

/* lib/core/partition.h:233-295 with requiresSort=false, consolidate=false:
 * arcs whose endpoints carry different ids, halved */
uint32_t ox_verify_cut_edges(uint32_t n, const uint32_t* ranges, const uint32_t* edges,
                             const int32_t* ids) {
    uint64_t cross = 0;
    for (uint32_t u = 0; u < n; ++u)
        for (uint32_t j = ranges[u]; j < ranges[u + 1]; ++j)
            cross += (ids[u] != ids[edges[j]]);
    return (uint32_t)(cross / 2);
}

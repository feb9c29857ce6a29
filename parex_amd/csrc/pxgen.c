/*
 * Host-side graph ingest: deterministic generators, Matrix-Market reader,
 * CSR finalize, .pxg I/O and the partition verifier (include/parex_amd/pxgen.h).
 */
#include "parex_amd/pxgen.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_err[256];

static int fail(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

const char* pxg_last_error(void) { return g_err; }

void pxg_free(pxg_graph* g) {
    if (!g) return;
    free(g->ranges);
    free(g->edges);
    g->ranges = g->edges = NULL;
    g->n = g->m = 0;
}

/* ------------------------------------------------------------------ RNG */

static inline uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* uniform in (0, 1] */
static inline double urand(uint64_t* s) {
    return ((double)(splitmix64(s) >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

static uint64_t mix_seed(uint64_t a, uint64_t b) {
    uint64_t s = a * 0x9E3779B97F4A7C15ull + b;
    splitmix64(&s);
    return splitmix64(&s);
}

/* ------------------------------------------------------------------ buffers */

typedef struct {
    uint32_t* u;
    uint32_t* v;
    uint64_t len, cap;
} edge_buf;

static int eb_push(edge_buf* b, uint32_t u, uint32_t v) {
    if (b->len == b->cap) {
        uint64_t nc = b->cap ? b->cap * 2 : 1024;
        uint32_t* nu = (uint32_t*)realloc(b->u, nc * 4);
        uint32_t* nv = (uint32_t*)realloc(b->v, nc * 4);
        if (!nu || !nv) { free(nu ? nu : b->u); free(nv ? nv : b->v); b->u = b->v = NULL; return -1; }
        b->u = nu; b->v = nv; b->cap = nc;
    }
    b->u[b->len] = u;
    b->v[b->len] = v;
    ++b->len;
    return 0;
}

static void eb_free(edge_buf* b) { free(b->u); free(b->v); memset(b, 0, sizeof *b); }

/* CSR from an edge list sorted by (u, v) with u < v: rows come out sorted. */
static int csr_from_sorted(uint32_t n, const edge_buf* b, pxg_graph* out) {
    if (2 * b->len > 0xFFFFFFFFull) return fail("too many arcs for u32 CSR (%llu edges)", (unsigned long long)b->len);
    uint32_t* ranges = (uint32_t*)calloc((size_t)n + 1, 4);
    uint32_t* edges = (uint32_t*)malloc((size_t)(b->len ? 2 * b->len : 1) * 4);
    uint32_t* pos = (uint32_t*)malloc(((size_t)n + 1) * 4);
    if (!ranges || !edges || !pos) { free(ranges); free(edges); free(pos); return fail("out of memory"); }
    for (uint64_t e = 0; e < b->len; ++e) { ranges[b->u[e] + 1]++; ranges[b->v[e] + 1]++; }
    for (uint32_t i = 0; i < n; ++i) ranges[i + 1] += ranges[i];
    memcpy(pos, ranges, ((size_t)n + 1) * 4);
    for (uint64_t e = 0; e < b->len; ++e) {
        edges[pos[b->u[e]]++] = b->v[e];
        edges[pos[b->v[e]]++] = b->u[e];
    }
    free(pos);
    out->n = n;
    out->m = (uint32_t)b->len;
    out->ranges = ranges;
    out->edges = edges;
    return 0;
}

/* LSD radix sort of u64 keys, 16-bit digits, skipping constant digits */
static int radix_u64(uint64_t* a, uint64_t len) {
    uint64_t* tmp = (uint64_t*)malloc((size_t)(len ? len : 1) * 8);
    uint64_t* cnt = (uint64_t*)malloc(65536 * 8);
    if (!tmp || !cnt) { free(tmp); free(cnt); return -1; }
    uint64_t *src = a, *dst = tmp;
    for (int pass = 0; pass < 4; ++pass) {
        int sh = 16 * pass;
        memset(cnt, 0, 65536 * 8);
        for (uint64_t i = 0; i < len; ++i) cnt[(src[i] >> sh) & 0xFFFF]++;
        int constant = 0;
        for (int d = 0; d < 65536; ++d) if (cnt[d] == len) { constant = 1; break; }
        if (constant) continue;
        uint64_t sum = 0;
        for (int d = 0; d < 65536; ++d) { uint64_t c = cnt[d]; cnt[d] = sum; sum += c; }
        for (uint64_t i = 0; i < len; ++i) dst[cnt[(src[i] >> sh) & 0xFFFF]++] = src[i];
        uint64_t* t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, (size_t)len * 8);
    free(tmp);
    free(cnt);
    return 0;
}

/* keys (u<<32|v), sorted + deduplicated -> CSR */
static int csr_from_keys(uint32_t n, uint64_t* keys, uint64_t len, int dedup, pxg_graph* out) {
    if (radix_u64(keys, len)) return fail("out of memory in sort");
    edge_buf b = {0};
    for (uint64_t i = 0; i < len; ++i) {
        if (dedup && i > 0 && keys[i] == keys[i - 1]) continue;
        if (eb_push(&b, (uint32_t)(keys[i] >> 32), (uint32_t)keys[i])) { eb_free(&b); return fail("out of memory"); }
    }
    int rc = csr_from_sorted(n, &b, out);
    eb_free(&b);
    return rc;
}

/* ------------------------------------------------------------------ SBM */

/* geometric run over candidate targets [lo, hi) with probability p */
static int run_pairs(edge_buf* b, uint32_t i, uint64_t lo, uint64_t hi, double p, uint64_t* rng) {
    if (p <= 0.0 || lo >= hi) return 0;
    if (p >= 1.0) {
        for (uint64_t j = lo; j < hi; ++j) if (eb_push(b, i, (uint32_t)j)) return -1;
        return 0;
    }
    const double lq = log1p(-p);
    uint64_t j = lo;
    for (;;) {
        double g = floor(log(urand(rng)) / lq);
        if (g >= (double)(hi - j)) break;
        j += (uint64_t)g;
        if (eb_push(b, i, (uint32_t)j)) return -1;
        ++j;
        if (j >= hi) break;
    }
    return 0;
}

int pxg_sbm(uint32_t n, uint32_t k, double p_in, double p_out, uint64_t seed, pxg_graph* out) {
    memset(out, 0, sizeof *out);
    if (k == 0 || k > n) return fail("sbm: need 1 <= k <= n (k=%u, n=%u)", k, n);
    uint64_t rng = mix_seed(seed, 0x5B11ull);
    edge_buf b = {0};
    uint32_t blk = 0;
    uint64_t bend = (uint64_t)n / k;                   /* end of block 0 */
    for (uint32_t i = 0; i < n; ++i) {
        while ((uint64_t)i >= bend) { ++blk; bend = ((uint64_t)(blk + 1) * n) / k; }
        if (run_pairs(&b, i, (uint64_t)i + 1, bend, p_in, &rng) ||
            run_pairs(&b, i, bend, n, p_out, &rng)) { eb_free(&b); return fail("out of memory"); }
    }
    int rc = csr_from_sorted(n, &b, out);
    eb_free(&b);
    return rc;
}

int pxg_sbm_specs(uint32_t count, uint32_t nmin, uint32_t nmax, uint32_t kmin, uint32_t kmax,
                  double din0, double din1, double dout0, double dout1, uint64_t seed,
                  pxg_sbm_spec* out) {
    if (nmin < 2 || nmax < nmin || kmin < 1 || kmax < kmin) return fail("sbm_specs: bad ranges");
    uint64_t rng = mix_seed(seed, 0xBA7Cull);
    double l0 = log((double)nmin), l1 = log((double)nmax);
    for (uint32_t g = 0; g < count; ++g) {
        double un = urand(&rng), uk = urand(&rng), ui = urand(&rng), uo = urand(&rng);
        uint32_t n = (uint32_t)llround(exp(l0 + (l1 - l0) * (1.0 - un)));
        if (n < nmin) n = nmin;
        if (n > nmax) n = nmax;
        uint32_t k = kmin + (uint32_t)floor((1.0 - uk) * (double)(kmax - kmin + 1));
        if (k > kmax) k = kmax;
        if (k > n) k = n;
        double din = din0 + (din1 - din0) * (1.0 - ui);
        double dout = dout0 + (dout1 - dout0) * (1.0 - uo);
        double bs = (double)n / (double)k;
        out[g].n = n;
        out[g].k = k;
        out[g].p_in = bs > 1.0 ? fmin(1.0, din / (bs - 1.0)) : 0.0;
        out[g].p_out = (n - bs) > 0.0 ? fmin(1.0, dout / ((double)n - bs)) : 0.0;
        out[g].seed = mix_seed(seed, 1000003ull * (g + 1));
    }
    return 0;
}

int pxg_sbm_batch(uint32_t count, const pxg_sbm_spec* specs, pxg_graph* out) {
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : bad)
    for (int64_t g = 0; g < (int64_t)count; ++g) {
        if (pxg_sbm(specs[g].n, specs[g].k, specs[g].p_in, specs[g].p_out, specs[g].seed, &out[g])) bad |= 1;
    }
    return bad ? fail("sbm_batch: generation failed") : 0;
}

/* ------------------------------------------------------------------ torus */

static uint64_t hkey(uint32_t u, uint32_t v) {
    return u < v ? ((uint64_t)u << 32) | v : ((uint64_t)v << 32) | u;
}

typedef struct { uint64_t* slot; uint64_t mask; } hset;

static int hs_init(hset* h, uint64_t expect) {
    uint64_t cap = 1;
    while (cap < expect * 2 + 16) cap <<= 1;
    h->slot = (uint64_t*)malloc(cap * 8);
    if (!h->slot) return -1;
    memset(h->slot, 0xFF, cap * 8);
    h->mask = cap - 1;
    return 0;
}

/* returns 1 if inserted, 0 if present */
static int hs_insert(hset* h, uint64_t k) {
    uint64_t s = k;
    uint64_t i = splitmix64(&s) & h->mask;
    for (;;) {
        if (h->slot[i] == ~0ull) { h->slot[i] = k; return 1; }
        if (h->slot[i] == k) return 0;
        i = (i + 1) & h->mask;
    }
}

int pxg_torus(uint32_t L, uint64_t extra, uint64_t seed, pxg_graph* out) {
    memset(out, 0, sizeof *out);
    if (L < 3) return fail("torus: need L >= 3");
    uint64_t n = (uint64_t)L * L;
    if (n > 0x7FFFFFFFull) return fail("torus too large");
    uint64_t base = 2 * n;
    uint64_t total = base + extra;
    if (extra > n * (n - 1) / 2 - base) return fail("torus: too many extra edges");
    uint64_t* keys = (uint64_t*)malloc(total * 8);
    hset h;
    if (!keys || hs_init(&h, total)) { free(keys); return fail("out of memory"); }
    uint64_t t = 0;
    for (uint32_t r = 0; r < L; ++r)
        for (uint32_t c = 0; c < L; ++c) {
            uint32_t id = r * L + c;
            uint32_t right = r * L + (c + 1) % L, down = ((r + 1) % L) * L + c;
            keys[t++] = hkey(id, right);
            keys[t++] = hkey(id, down);
            hs_insert(&h, hkey(id, right));
            hs_insert(&h, hkey(id, down));
        }
    uint64_t rng = mix_seed(seed, 0x70C5ull);
    while (t < total) {
        uint32_t u = (uint32_t)(splitmix64(&rng) % n), v = (uint32_t)(splitmix64(&rng) % n);
        if (u == v) continue;
        uint64_t k = hkey(u, v);
        if (!hs_insert(&h, k)) continue;
        keys[t++] = k;
    }
    free(h.slot);
    int rc = csr_from_keys((uint32_t)n, keys, total, 0, out);
    free(keys);
    return rc;
}

/* ------------------------------------------------------------------ RMAT */

int pxg_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, pxg_graph* out) {
    memset(out, 0, sizeof *out);
    if (scale == 0 || scale > 30) return fail("rmat: scale out of range");
    uint64_t n = 1ull << scale;
    uint64_t samples = n * ef;
    uint64_t* keys = (uint64_t*)malloc((size_t)(samples ? samples : 1) * 8);
    if (!keys) return fail("out of memory");
    const uint64_t CH = 1u << 16;                       /* fixed chunking -> deterministic */
    uint64_t nch = (samples + CH - 1) / CH;
    uint8_t* keep = (uint8_t*)malloc((size_t)(samples ? samples : 1));
    if (!keep) { free(keys); return fail("out of memory"); }
#pragma omp parallel for schedule(static)
    for (int64_t ch = 0; ch < (int64_t)nch; ++ch) {
        uint64_t rng = mix_seed(seed, 0xA11ull + (uint64_t)ch);
        uint64_t lo = (uint64_t)ch * CH, hi = lo + CH < samples ? lo + CH : samples;
        for (uint64_t i = lo; i < hi; ++i) {
            uint32_t u = 0, v = 0;
            for (uint32_t lv = 0; lv < scale; ++lv) {
                double r = urand(&rng);
                uint32_t bu = 0, bv = 0;
                if (r <= a) { }
                else if (r <= a + b) { bv = 1; }
                else if (r <= a + b + c) { bu = 1; }
                else { bu = 1; bv = 1; }
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            keep[i] = (u != v);
            keys[i] = hkey(u, v);
        }
    }
    uint64_t w = 0;
    for (uint64_t i = 0; i < samples; ++i) if (keep[i]) keys[w++] = keys[i];
    free(keep);
    int rc = csr_from_keys((uint32_t)n, keys, w, 1, out);
    free(keys);
    return rc;
}

/* ------------------------------------------------------------------ edge list */

int pxg_from_edges(uint32_t n, uint64_t m, const uint32_t* fx, const uint32_t* tx, pxg_graph* out) {
    memset(out, 0, sizeof *out);
    if (2 * m > 0xFFFFFFFFull) return fail("too many edges");
    uint64_t* keys = (uint64_t*)malloc((size_t)(m ? 2 * m : 1) * 8);
    if (!keys) return fail("out of memory");
    for (uint64_t e = 0; e < m; ++e) {
        if (fx[e] >= n || tx[e] >= n) { free(keys); return fail("edge %llu out of range", (unsigned long long)e); }
        keys[2 * e] = ((uint64_t)fx[e] << 32) | tx[e];
        keys[2 * e + 1] = ((uint64_t)tx[e] << 32) | fx[e];
    }
    if (radix_u64(keys, 2 * m)) { free(keys); return fail("out of memory"); }
    uint32_t* ranges = (uint32_t*)calloc((size_t)n + 1, 4);
    uint32_t* edges = (uint32_t*)malloc((size_t)(m ? 2 * m : 1) * 4);
    if (!ranges || !edges) { free(keys); free(ranges); free(edges); return fail("out of memory"); }
    uint64_t eix = 0;
    for (uint32_t nix = 0; nix <= n; ++nix) {          /* lib/core/graph.h:173-182 */
        ranges[nix] = (uint32_t)eix;
        while (nix < n && eix < 2 * m && (uint32_t)(keys[eix] >> 32) == nix) ++eix;
    }
    for (uint64_t e = 0; e < 2 * m; ++e) edges[e] = (uint32_t)keys[e];
    free(keys);
    out->n = n;
    out->m = (uint32_t)m;
    out->ranges = ranges;
    out->edges = edges;
    return 0;
}

/* ------------------------------------------------------------------ MTX */

int pxg_read_mtx(const char* path, pxg_graph* out) {
    memset(out, 0, sizeof *out);
    FILE* f = fopen(path, "r");
    if (!f) return fail("Unable to open file.");
    char line[4096];
    do {
        if (!fgets(line, sizeof line, f)) { fclose(f); return 0; }    /* empty graph */
    } while (line[0] == '%');
    unsigned long long rows = 0, n = 0, m = 0;
    if (sscanf(line, "%llu %llu %llu", &rows, &n, &m) != 3) { fclose(f); return fail("bad header"); }
    uint32_t* fx = (uint32_t*)malloc((size_t)(m ? m : 1) * 4);
    uint32_t* tx = (uint32_t*)malloc((size_t)(m ? m : 1) * 4);
    if (!fx || !tx) { fclose(f); free(fx); free(tx); return fail("out of memory"); }
    for (unsigned long long e = 0; e < m; ++e) {
        unsigned long long a, b;
        if (fscanf(f, "%llu %llu", &a, &b) != 2 || a == 0 || b == 0 || a > n || b > n) {
            fclose(f); free(fx); free(tx);
            return fail("bad edge line %llu", e + 1);
        }
        fx[e] = (uint32_t)(a - 1);
        tx[e] = (uint32_t)(b - 1);
    }
    fclose(f);
    int rc = 0;
    if (m == 0) {                                        /* finalize() of no edges -> {} */
        out->n = 0; out->m = 0;
        out->ranges = (uint32_t*)calloc(1, 4);
        out->edges = (uint32_t*)malloc(4);
    } else {
        rc = pxg_from_edges((uint32_t)n, m, fx, tx, out);
    }
    free(fx);
    free(tx);
    return rc;
}

/* ------------------------------------------------------------------ .pxg */

int pxg_write(const char* path, const pxg_graph* g) {
    FILE* f = fopen(path, "wb");
    if (!f) return fail("cannot open %s", path);
    uint32_t hdr[2] = {g->n, g->m};
    int ok = fwrite("PXG1", 1, 4, f) == 4 && fwrite(hdr, 4, 2, f) == 2 &&
             fwrite(g->ranges, 4, (size_t)g->n + 1, f) == (size_t)g->n + 1 &&
             fwrite(g->edges, 4, 2ull * g->m, f) == 2ull * g->m;
    fclose(f);
    return ok ? 0 : fail("short write to %s", path);
}

int pxg_read(const char* path, pxg_graph* out) {
    memset(out, 0, sizeof *out);
    FILE* f = fopen(path, "rb");
    if (!f) return fail("cannot open %s", path);
    char magic[4];
    uint32_t hdr[2];
    if (fread(magic, 1, 4, f) != 4 || memcmp(magic, "PXG1", 4) || fread(hdr, 4, 2, f) != 2) {
        fclose(f);
        return fail("bad .pxg header in %s", path);
    }
    out->n = hdr[0];
    out->m = hdr[1];
    out->ranges = (uint32_t*)malloc(((size_t)hdr[0] + 1) * 4);
    const uint64_t narcs = 2ull * hdr[1];
    out->edges = (uint32_t*)malloc((narcs ? narcs : 1) * 4);
    int ok = out->ranges && out->edges &&
             fread(out->ranges, 4, (size_t)hdr[0] + 1, f) == (size_t)hdr[0] + 1 &&
             fread(out->edges, 4, 2ull * hdr[1], f) == 2ull * hdr[1];
    fclose(f);
    if (!ok) { pxg_free(out); return fail("short read from %s", path); }
    return 0;
}

/* ------------------------------------------------------------------ verify */

int pxg_verify_partition(const pxg_graph* g, const int32_t* ids, int32_t num_clusters,
                         uint64_t* cut_edges, uint64_t* vol, uint64_t* internal_vol,
                         double* conductance) {
    if (num_clusters < 0) return fail("negative cluster count");
    uint64_t* lv = (uint64_t*)calloc((size_t)num_clusters + 1, 8);
    uint64_t* li = (uint64_t*)calloc((size_t)num_clusters + 1, 8);
    if (!lv || !li) { free(lv); free(li); return fail("out of memory"); }
    uint64_t cross = 0, total = 0;
    int rc = 0;
    for (uint32_t u = 0; u < g->n && !rc; ++u) {
        int32_t cu = ids[u];
        if (cu < 0 || cu >= num_clusters) { rc = fail("node %u has id %d outside [0,%d)", u, cu, num_clusters); break; }
        for (uint32_t j = g->ranges[u]; j < g->ranges[u + 1]; ++j) {
            uint32_t v = g->edges[j];
            lv[cu]++;
            ++total;
            if (ids[v] == cu) li[cu]++; else ++cross;
        }
    }
    if (!rc) {
        if (cut_edges) *cut_edges = cross / 2;
        for (int32_t c = 0; c < num_clusters; ++c) {
            if (vol) vol[c] = lv[c];
            if (internal_vol) internal_vol[c] = li[c];
            if (conductance) {
                uint64_t cut = lv[c] - li[c];
                uint64_t den = lv[c] < total - lv[c] ? lv[c] : total - lv[c];
                conductance[c] = den ? (double)cut / (double)den : 0.0;
            }
        }
    }
    free(lv);
    free(li);
    return rc;
}

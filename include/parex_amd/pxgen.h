/*
 * Host-side graph ingest for the expander-decomposition engine: deterministic
 * synthetic generators, the Matrix-Market reader, CSR finalize, the .pxg binary
 * CSR format and the host partition verifier.  Plain C ABI; no GPU calls.
 *
 * CSR convention (lib/core/graph.h:145-158): ranges u32[n+1], edges u32[2m];
 * every undirected edge appears in both rows, rows are sorted ascending.
 * Buffers returned through pxg_graph are owned by the library; free them with
 * pxg_free.
 */
#ifndef PAREX_AMD_PXGEN_H
#define PAREX_AMD_PXGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t n;          /* nodes */
    uint32_t m;          /* undirected edges */
    uint32_t* ranges;    /* n+1 */
    uint32_t* edges;     /* 2m */
} pxg_graph;

void pxg_free(pxg_graph* g);

/* Planted-partition SBM: n nodes in k contiguous blocks (block b holds
 * [b*n/k, (b+1)*n/k)); each intra-block pair independently with p_in, each
 * inter-block pair with p_out.  Simple, undirected. */
int pxg_sbm(uint32_t n, uint32_t k, double p_in, double p_out, uint64_t seed, pxg_graph* out);

/* L x L torus (id = r*L + c, 4-neighbour wrap-around) plus `extra` uniformly
 * random non-loop, non-duplicate long-range edges drawn from `seed`. */
int pxg_torus(uint32_t L, uint64_t extra, uint64_t seed, pxg_graph* out);

/* RMAT: 2^scale nodes, ef * 2^scale sampled edges with quadrant probabilities
 * (a, b, c, 1-a-b-c); symmetrised, self-loops and duplicates removed,
 * isolated nodes kept. */
int pxg_rmat(uint32_t scale, uint32_t ef, double a, double b, double c, uint64_t seed, pxg_graph* out);

/* Parameters of the batch of independent SBMs (one row per graph). */
typedef struct {
    uint32_t n, k;
    double p_in, p_out;
    uint64_t seed;
} pxg_sbm_spec;

/* Draw `count` SBM specs: n = round(exp(U(ln nmin, ln nmax))), k ~ U{kmin..kmax},
 * intra degree ~ U(din0, din1) -> p_in = d/(n/k - 1), inter degree ~
 * U(dout0, dout1) -> p_out = d/(n - n/k); per-graph seeds derived from `seed`. */
int pxg_sbm_specs(uint32_t count, uint32_t nmin, uint32_t nmax, uint32_t kmin, uint32_t kmax,
                  double din0, double din1, double dout0, double dout1, uint64_t seed,
                  pxg_sbm_spec* out);

/* Generate all graphs of a spec list (OpenMP over graphs when available). */
int pxg_sbm_batch(uint32_t count, const pxg_sbm_spec* specs, pxg_graph* out);

/* Matrix Market reader with the reference's parse rules
 * (lib/utils/graph_io.h:49-84): skip '%' lines, header "rows n m", then m
 * 1-based pairs; then DynamicGraph::finalize (lib/core/graph.h:147-184). */
int pxg_read_mtx(const char* path, pxg_graph* out);

/* .pxg binary CSR: "PXG1", u32 n, u32 m, u32 ranges[n+1], u32 edges[2m] */
int pxg_write(const char* path, const pxg_graph* g);
int pxg_read(const char* path, pxg_graph* out);

/* CSR from an undirected edge list (any order; loops/duplicates kept as the
 * reference's finalize keeps them). */
int pxg_from_edges(uint32_t n, uint64_t m, const uint32_t* fx, const uint32_t* tx, pxg_graph* out);

/* Host re-verification of a final partition (main.cpp:53-60 ->
 * lib/core/partition.h:233-295): ids must lie in [0, num_clusters).
 * Returns 0 if valid; writes the cut-edge count and, when the arrays are
 * non-NULL, per-cluster volume, internal volume (arcs) and conductance
 * cut/min(vol, total-vol) (0 when the denominator is 0). */
int pxg_verify_partition(const pxg_graph* g, const int32_t* ids, int32_t num_clusters,
                         uint64_t* cut_edges, uint64_t* vol, uint64_t* internal_vol,
                         double* conductance);

const char* pxg_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

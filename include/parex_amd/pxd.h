/*
 * pxd -- C ABI of the MI355X expander-decomposition engine (HIP, gfx950).
 *
 * Drop-in boundary for the reference's CUDA expander decomposition: every
 * entry point below replaces one member of `class CudaDeviceManager`
 * (cuda/interface.h:16-57, forwarders cuda/device.h:394-427 of the reference).
 * The reference reads its parameters from mutable globals (randSeed,
 * sc_threshold, rw_threshold; lib/core/definitions.h:46-50); here they are
 * passed explicitly in pxd_params.
 *
 * Conventions: plain pointers and sizes only; every call returns 0 on success
 * or a negative pxd_status; pxd_last_error() gives the message of the last
 * failure on the calling thread.  A handle is not thread-safe; use one per
 * host thread.  All device work of a handle runs on the handle's own stream.
 * Output buffers are caller-owned host memory.
 */
#ifndef PAREX_AMD_PXD_H
#define PAREX_AMD_PXD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PXD_INVALID_EDGE 0xFFFFFFFFu   /* lib/core/definitions.h:52 */

typedef enum {
    PXD_OK = 0,
    PXD_ERR_ARG = -1,        /* bad argument / call order */
    PXD_ERR_HIP = -2,        /* HIP runtime error */
    PXD_ERR_OOM = -3,        /* device allocation failed */
    PXD_ERR_NUMERIC = -4,    /* walk value outside the exact fixed-point range */
    PXD_ERR_STATE = -5,      /* internal invariant violated */
    PXD_FULL = 1,            /* serving: the graph does not fit the free capacity now */
    PXD_EMPTY = 1            /* serving: no finished result is waiting */
} pxd_status;

/* lib/core/definitions.h:57-84 -- identical byte layouts */
typedef struct { int32_t clusterId; float sparsity; uint32_t offset; } pxd_sweep_cut;       /* SweepCutData */
typedef struct { uint32_t nix; int32_t clusterId; } pxd_labeled_node;                       /* LabeledNode */
typedef struct { uint32_t nix; uint32_t prefixVolume; int32_t prefixEdgeDiff; uint32_t offsetInCluster; } pxd_node_data; /* NodeData */

typedef struct {
    uint32_t seed;           /* randSeed, default 6582 */
    float sc_threshold;      /* sweep-cut (conductance) threshold, default 0.3 */
    float rw_threshold;      /* random-walk potential threshold, default 1e-4 */
    uint32_t max_iter;       /* MAX_NUM_ITER, default 200000 (cuda/device.h:309) */
    uint32_t keep_original;  /* keep a copy of the input CSR for pxd_download_graph /
                                pxd_download_active_edge_map (default 1) */
    uint32_t record_node_data; /* keep per-position NodeData for pxd_download_partition (default 1) */
    int32_t device;          /* HIP device ordinal, default 0 */
} pxd_params;

typedef struct pxd_engine pxd_engine;

void pxd_default_params(pxd_params* p);
const char* pxd_last_error(void);

/* CudaDeviceManager() / ~CudaDeviceManager()   (cuda/interface.h:18-19) */
int pxd_create(const pxd_params* params, pxd_engine** out);
int pxd_destroy(pxd_engine* e);

/* initialize(const Graph&)  (cuda/interface.h:24, cuda/device.h:24-38):
 * copies the CSR (ranges u32[n+1], edges u32[2m], rows sorted, both
 * directions present) to the device and builds all state, including the
 * initial walk vector. */
int pxd_initialize(pxd_engine* e, uint32_t n, uint32_t m, const uint32_t* ranges, const uint32_t* edges);

/* Batched variant (no reference counterpart: the reference handles one graph
 * per CudaDeviceManager).  num_graphs independent graphs, each given in its
 * own local numbering; graph g has node_counts[g] nodes, edge_counts[g]
 * undirected edges, ranges_ptrs[g] (n_g+1 entries) and edges_ptrs[g] (2 m_g).
 * Every graph's result equals what pxd_initialize + run gives for it alone. */
int pxd_initialize_batch(pxd_engine* e, uint32_t num_graphs, const uint32_t* node_counts,
                         const uint32_t* edge_counts, const uint32_t* const* ranges_ptrs,
                         const uint32_t* const* edges_ptrs);

/* The reference's input path, readGraph + Graph::finalize + initialize
 * (lib/utils/graph_io.h:49-84, lib/core/graph.h:147-184, cuda/interface.h:24),
 * with the CSR built on the device: m undirected edges (fx[i], tx[i]), both
 * arcs of each, rows ordered by target; loops and duplicates kept as given.
 * Arrays are host memory; an end >= n is refused (PXD_ERR_ARG).  Below 2^31
 * arcs. */
int pxd_initialize_edges(pxd_engine* e, uint32_t n, uint64_t m, const uint32_t* fx, const uint32_t* tx);

/* The reference reads sc_threshold / rw_threshold from globals on every step
 * call (cuda/devSweepCut.h:701, cuda/devPartition.h:555, 632); callers that
 * change them between steps update the engine with this call. */
int pxd_set_thresholds(pxd_engine* e, float sc_threshold, float rw_threshold);

/* Return to the state pxd_initialize / pxd_initialize_batch left, from the
 * device-resident copy of the input (needs keep_original = 1).  Equivalent to
 * calling initialize again with the same graphs, without the host transfer. */
int pxd_reset(pxd_engine* e);

/* iterateRandomWalk()   (cuda/interface.h:29, cuda/device.h:40-48) */
int pxd_iterate_random_walk(pxd_engine* e);
/* computeSweepCuts()    (cuda/interface.h:32, cuda/device.h:77-81) */
int pxd_compute_sweep_cuts(pxd_engine* e);
/* cutClusters()         (cuda/interface.h:36, cuda/device.h:101-168) */
int pxd_cut_clusters(pxd_engine* e);
/* expanderDecomposition() (cuda/interface.h:40, cuda/device.h:301-350):
 * runs the loop to completion; *iterations receives the value the reference
 * prints ("iterations: %d"); with a batch, the largest per-graph value. */
int pxd_expander_decomposition(pxd_engine* e, int32_t* iterations);
/* per-graph "iterations" values after pxd_expander_decomposition */
int pxd_graph_iterations(pxd_engine* e, int32_t* out, uint32_t num_graphs);

/* getFinalPartition()   (cuda/interface.h:42, cuda/devPartition.h:798-823):
 * ids_out gets n dense cluster ids; *num_clusters the number of ids.
 * For a batch, graph g's slice holds ids local to graph g and
 * num_clusters[g] is per graph (pass num_graphs entries). */
int pxd_get_final_partition(pxd_engine* e, int32_t* ids_out, int32_t* num_clusters);

/* readRandomWalkValues() (cuda/interface.h:30): n floats (FP64 state cast) */
int pxd_read_random_walk_values(pxd_engine* e, float* out);
/* the FP64 state itself (no reference counterpart; used for parity tests) */
int pxd_read_random_walk_values_f64(pxd_engine* e, double* out);
/* downloadDegrees() (cuda/interface.h:26): internal degrees, n entries */
int pxd_download_degrees(pxd_engine* e, uint32_t* out);
/* downloadLabels() (cuda/interface.h:45): allLabels, n entries */
int pxd_download_labels(pxd_engine* e, int32_t* out);
/* downloadGraph() (cuda/interface.h:25): ranges n+1, edges 2m with disabled
 * arcs written as PXD_INVALID_EDGE in their original positions */
int pxd_download_graph(pxd_engine* e, uint32_t* ranges_out, uint32_t* edges_out);
/* downloadActiveEdgeMap() (cuda/interface.h:44): 2m ints, 1 = arc still valid */
int pxd_download_active_edge_map(pxd_engine* e, int32_t* out);
/* downloadPartition() (cuda/interface.h:27): NodeData of the last sweep for the
 * first *count active positions (capacity entries available in out) */
int pxd_download_partition(pxd_engine* e, pxd_node_data* out, uint32_t capacity, uint32_t* count);
/* readSweepCuts() (cuda/interface.h:34, cuda/device.h:170-179): labels and
 * best cuts of the active clusters of the last sweep */
int pxd_read_sweep_cuts(pxd_engine* e, int32_t* cluster_ids, pxd_sweep_cut* cuts,
                        uint32_t capacity, uint32_t* count);
/* getNumCutEdges() (cuda/interface.h:38, cuda/devPartition.h:741-759): m - validArcs/2 */
int pxd_get_num_cut_edges(pxd_engine* e, uint32_t* out);
/* active list (activeNodes) in device order; capacity entries */
int pxd_download_active_nodes(pxd_engine* e, pxd_labeled_node* out, uint32_t capacity, uint32_t* count);
/* counters: numActiveNodes, numActiveClusters, numClustersWithCut */
int pxd_get_counts(pxd_engine* e, uint32_t* num_active_nodes, uint32_t* num_active_clusters,
                   uint32_t* num_clusters_with_cut);
/* overwrite the FP64 walk state of all n nodes (the reference's tests seed the
 * CPU walk from the device; this is the converse, for parity runs) */
int pxd_set_random_walk_values_f64(pxd_engine* e, const double* values);

/* device time spent inside the last pxd_expander_decomposition (ms, HIP events
 * on the handle's stream) and per-kernel-class accumulated times */
int pxd_last_run_ms(pxd_engine* e, double* ms);

/* Per-kernel timing of the walk kernel: enable HIP-event timing around every
 * walk launch; pxd_walk_stats returns launches, total ms and the algorithmic
 * bytes (4+8 per valid arc + 32 per active node) summed over those launches. */
int pxd_enable_walk_timing(pxd_engine* e, int enable);
int pxd_walk_stats(pxd_engine* e, uint64_t* launches, double* total_ms, double* alg_bytes);
/* The same per timed launch, in launch order: ms[i] and alg_bytes[i] for
 * i < *count (count alone when both arrays are null). */
int pxd_walk_launch_log(pxd_engine* e, double* ms, double* alg_bytes, uint32_t capacity, uint32_t* count);

/* One-workgroup runs (no counterpart in the reference, whose decomposition
 * loop is driven from the host, cuda/device.h:301-350): pxd_expander_decomposition
 * of a single graph of at most 4096 nodes and 2^19 arcs, started right after
 * pxd_initialize or pxd_reset, runs every iteration inside one workgroup in
 * one launch, with the same results as the per-phase kernels.  enable = 1
 * (default) takes it for graphs of at most 2^16 arcs (denser graphs run
 * faster on the per-phase kernels), 2 for every graph within its capacity,
 * 0 never. */
int pxd_set_small_graph_kernel(pxd_engine* e, int enable);

/* Kernel choice for the walk and the cut contribution, the two per-arc
 * passes of an iteration (no counterpart in the reference, whose kernels have
 * one launch shape each, cuda/devRandomWalk.h:369-392, cuda/devSweepCut.h:
 * 486-499).  variant 0..5: one row per 2, 4, 8, 16, 32, 64 lanes; 6, 7, 8:
 * the LDS-window kernels with 1, 2, 4 lanes per row; 9: the window kernel
 * with 2 lanes, 1024-entry chunks and a quarter window; 10: variant 7 with
 * clamped slots in the walk (variant 7 for the cut contribution); 11: the
 * window kernel with 2 lanes over node-index-range chunks (windows that
 * cover sparse lists; it needs the node list in index order, so a serving
 * engine, whose admitted graphs are appended, runs variant 7 for it); 12:
 * variant 11 over half-size ranges with half the window (two workgroups a
 * CU; a serving engine runs variant 7 for it); 13: the window kernel with 2
 * lanes, 2048-entry chunks and half the window (two workgroups a CU); 14,
 * 15: variants 11, 12 with one lane per row (a serving engine runs variant 6
 * for them); -1 (default): online
 * tuning -- the first time the active list has a length in a quarter of
 * [2^b, 2^(b+1)) (b >= 16; quarters by the two bits after the leading one)
 * every candidate is launched three times on the current state and timed,
 * the fastest is kept for that pass and length (the tables survive
 * pxd_reset); -2: the degree rule without tuning.  Every variant computes the
 * same bits.  pxd_walk_tuning copies the tables (entries 0..32: walk per
 * class b, 33..65: cut contribution, each the choice for the class's
 * highest tuned quarter; -1 = not tuned) and the number of tunings run. */
int pxd_set_walk_variant(pxd_engine* e, int variant);
int pxd_walk_tuning(pxd_engine* e, int32_t* table, uint32_t capacity, uint64_t* tunings);

/* ------------------------------------------------------------------ serving
 * No counterpart in the reference (cuda/interface.h runs one graph per
 * manager, main.cpp:44-50 one file per process).  Graphs join a running engine
 * between iterations and leave when they finish: continuous batching.  Each
 * graph's result equals its solo pxd_expander_decomposition result.
 * Capacities fix the device footprint: node ranges, arc ranges and graph
 * slots are taken first fit and reused. */

/* Engine with room for node_capacity nodes (< 2^31), arc_capacity arcs
 * (2 per undirected edge, < 2^32) and graph_slots graphs in flight. */
int pxd_serve_create(const pxd_params* params, uint32_t node_capacity, uint64_t arc_capacity, uint32_t graph_slots,
                     pxd_engine** out);
/* Admit one graph (CSR as for pxd_initialize) under the caller's tag.
 * PXD_FULL: it does not fit now (retry after steps free room).  The uploads
 * are queued (pinned host buffers make them plain DMA): the buffers must stay
 * valid until the next pxd_serve_step returns.  The graph is checked on the
 * device (non-decreasing ranges, arcs below n) and joins at that step; one
 * that fails comes back as a result with num_clusters = -1. */
int pxd_serve_admit(pxd_engine* e, uint32_t n, uint32_t m, const uint32_t* ranges, const uint32_t* edges,
                    uint64_t tag, int32_t* slot);
/* Edge-list admissions (the CSR built on the device, as pxd_initialize_edges
 * does): enable once with the largest graph to expect (a device workspace of
 * about 40 B per edge and 4 B per node), then admit undirected edge lists
 * (fx[i], tx[i]) like CSRs -- same PXD_FULL rule, buffers valid until the next
 * pxd_serve_step; an end >= n comes back as a result with num_clusters = -1. */
int pxd_serve_enable_edges(pxd_engine* e, uint32_t max_nodes, uint64_t max_edges);
int pxd_serve_admit_edges(pxd_engine* e, uint32_t n, uint64_t m, const uint32_t* fx, const uint32_t* tx,
                          uint64_t tag, int32_t* slot);
/* One iteration (walk, sweep, cut) of every graph in flight; *finished =
 * graphs that completed in it (their results wait for pxd_serve_take). */
int pxd_serve_step(pxd_engine* e, uint32_t* finished);
/* Tag and node count of the oldest result; PXD_EMPTY if none is ready (the
 * dense ranking runs on a worker thread of the engine). */
int pxd_serve_next(pxd_engine* e, uint64_t* tag, uint32_t* n);
/* Block until every finished result is ready. */
int pxd_serve_wait(pxd_engine* e);
/* Pop that result: cluster id per node (ids_out[n]), cluster count and the
 * iteration count its solo run reports (cuda/device.h:343-350), the ids
 * densely ranked as cuda/devPartition.h:798-823 does.  A failed result has a
 * negative cluster count and no ids (n = 0 in pxd_serve_next): -1 when the
 * graph failed the device check, -2 when the host had no memory to rank it. */
int pxd_serve_take(pxd_engine* e, int32_t* ids_out, int32_t* num_clusters, int32_t* iterations);
int pxd_serve_stats(pxd_engine* e, uint32_t* live_graphs, uint64_t* live_nodes, uint64_t* live_arcs,
                    uint32_t* active_nodes, uint32_t* waiting_results);

/* ------------------------------------------------------------------ allocation-free CSR build
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: the device CSR build of pxd_initialize_edges (the
 * reference's readGraph + Graph::finalize, lib/utils/graph_io.h:49-84 and
 * lib/core/graph.h:147-184) from a device-resident undirected edge list
 * (d_fx[i], d_tx[i]), i < m, into caller-owned device buffers d_ranges[n+1]
 * and d_edges[2m], with the same kernels.  The work area is a static device
 * array of the library (about 25 MB per GPU, held until the process ends), so
 * the call allocates nothing: n <= PXD_NOALLOC_MAX_NODES and m <=
 * PXD_NOALLOC_MAX_EDGES, larger graphs are refused with PXD_ERR_ARG.
 * Asynchronous: queued on `stream` (a hipStream_t; NULL = the null stream),
 * nothing is waited for.  *d_bad (device memory) becomes nonzero when an
 * end is >= n (the CSR is then meaningless).  Every call on one GPU shares
 * that work area: calls that overlap on one GPU (other streams or threads)
 * race and must be serialized by the caller. */
#define PXD_NOALLOC_MAX_NODES (1u << 20)
#define PXD_NOALLOC_MAX_EDGES (1ull << 21)
int pxd_csr_from_edges_noalloc(uint32_t n, uint64_t m, const uint32_t* d_fx, const uint32_t* d_tx,
                               uint32_t* d_ranges, uint32_t* d_edges, uint32_t* d_bad, void* stream);

#ifdef __cplusplus
}
#endif
#endif

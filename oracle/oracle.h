/*
 * parex expander-decomposition ORACLE -- test infrastructure only.
 *
 * A single-threaded CPU restatement, in plain C, of the reference's CUDA
 * expander decomposition (cuda/device.h, cuda/devRandomWalk.h,
 * cuda/devSweepCut.h, cuda/devPartition.h, cuda/devConsolidate.h).  It keeps
 * the reference's own data flow (one LabeledNode list, 64-bit sort keys,
 * stable sorts, reduce-by-key over labels, in-place INVALID_EDGE marking) so
 * that it is independent of the segment-based HIP engine it checks.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker.  The product never links it.
 *
 * Canonical arithmetic (the parity contract shared with the HIP engine; the
 * reference's own CUDA arithmetic is order-dependent and not reproducible
 * off-NVIDIA, so these choices pin one exact answer):
 *   W1  walk neighbour sum: each term y = x[nb] / (double)dInt[nb] (IEEE
 *       division) is truncated to a signed int128 fixed-point value with
 *       scale 2^72, the terms are summed exactly, and the sum is converted
 *       back with one correct rounding.  Any summation order gives the same
 *       bits.  |y| must stay below 2^31 (checked; error flag otherwise).
 *   W2  lazy step x' = s*(1-rw_stay) + x*rw_stay, no fused multiply-add.
 *   C1  cluster rwSum: values f = (float)x, e = exponent of max|f|,
 *       K = 62-(e+1)-ceil(log2(count)); terms trunc(f*2^K) summed in int64;
 *       rwSum = (float)ldexp((double)S, -K).  max/min over ordered floats
 *       (-0 < +0).
 *   S1  sparsity = (float)edgeDiff / (float)denom with IEEE division.
 *   CC  connected-component label = minimum node index of the component
 *       (cuda/devConsolidate.h:168-204 intends this; its racy hooking only
 *       reaches it when parent trees are flat at convergence).
 *
 * Parity pinning: see tests/test_oracle.py (closed forms, brute force,
 * golden vectors of the reference CPU code).  The init vector's erfcinv is
 * an FP64 Newton iteration on erfc, not CUDA's erfcinvf; it is pinned only
 * within a few float ulp, so end-to-end parity runs seed the oracle with the
 * device's initial vector (the reference's own test pattern,
 * tests/cudaTest.cpp:50-56).
 */
#ifndef PAREX_ORACLE_H
#define PAREX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OX_INVALID_EDGE 0xFFFFFFFFu

/* lib/core/definitions.h:57-84 -- byte-identical layouts */
typedef struct { int32_t clusterId; float sparsity; uint32_t offset; } ox_sweepcut;
typedef struct { uint32_t nix; int32_t clusterId; } ox_lnode;
typedef struct { uint32_t nix; uint32_t prefixVolume; int32_t prefixEdgeDiff; uint32_t offsetInCluster; } ox_ndata;
typedef struct { float rwSum; float maxPotential; float minPotential; uint32_t totalElements; } ox_cdata;

typedef struct ox_state ox_state;

/* Build state: CSR upload + PartitionManager/RandomWalk/SweepCut init
 * (cuda/device.h:24-38).  init_values may be NULL (own init vector). */
ox_state* ox_create(uint32_t n, uint32_t m, const uint32_t* ranges, const uint32_t* edges,
                    uint32_t seed, float sc_threshold, float rw_threshold,
                    const double* init_values);
void ox_destroy(ox_state* s);

/* one step each, cuda/device.h:40-48, 77-81, 101-168 */
void ox_walk(ox_state* s);
void ox_sweep(ox_state* s);
void ox_cut(ox_state* s);
/* full loop, cuda/device.h:301-350; returns printed "iterations" */
int  ox_run(ox_state* s, int max_iter);

/* finalize (cuda/devPartition.h:798-823) */
int  ox_final_partition(ox_state* s, int32_t* ids_out);
uint32_t ox_num_cut_edges(ox_state* s);       /* cuda/devPartition.h:741-759 */

/* state readers */
uint32_t ox_num_nodes(ox_state* s);
uint32_t ox_num_active_nodes(ox_state* s);
uint32_t ox_num_active_clusters(ox_state* s);
uint32_t ox_num_clusters_with_cut(ox_state* s);
int32_t  ox_smallest_label(ox_state* s);
int      ox_error_flags(ox_state* s);
void ox_get_values(ox_state* s, double* out);           /* dist.Current(), n */
void ox_get_labels(ox_state* s, int32_t* out);          /* allLabels, n */
void ox_get_internal_degrees(ox_state* s, uint32_t* out);
void ox_get_neighbors(ox_state* s, uint32_t* out);      /* mutable CSR, 2m */
void ox_get_active(ox_state* s, ox_lnode* out);         /* activeNodes[0..na) */
void ox_get_ndata(ox_state* s, ox_ndata* out);          /* scanned NodeData[0..na) */
void ox_get_sweepcuts(ox_state* s, ox_sweepcut* out);   /* sweepCutsBuffer[0..C) */
void ox_get_cluster_labels(ox_state* s, int32_t* out);  /* clusterLabels[0..C) */
void ox_get_cluster_data(ox_state* s, ox_cdata* out);   /* clusterData[0..C) */

/* OpenMP threads used by the calling thread's later calls (results do not
 * depend on it) */
void ox_set_threads(int n);

/* building blocks, exported for the oracle's own pinning tests */
void   ox_init_vector(uint32_t n, const uint32_t* degrees, uint32_t seed, double* out);
float  ox_init_value(uint32_t idx, uint32_t deg, uint32_t seed);
double ox_erfcinv(double y);
uint32_t ox_hash(uint32_t x);
uint32_t ox_minstd_first(uint32_t seed);
uint64_t ox_pack_key(int32_t label, float v);
double ox_fixsum(const double* terms, uint32_t k, int* overflow);
float  ox_cluster_sum(const float* vals, uint32_t k);

/* lib/core/graph.h:147-184: DynamicGraph::finalize on an edge list; returns
 * number of arcs written (2*m) or 0 for an empty list (then ranges untouched) */
uint32_t ox_finalize(uint32_t n, uint32_t m, const uint32_t* fx, const uint32_t* tx,
                     uint32_t* ranges_out, uint32_t* edges_out);

/* main.cpp:53-60 -> lib/core/partition.h:233-295 host verifier */
uint32_t ox_verify_cut_edges(uint32_t n, const uint32_t* ranges, const uint32_t* edges,
                             const int32_t* ids);

#ifdef __cplusplus
}
#endif
#endif

/* Expander hierarchy and normalized cut over the engine's decompositions.
 *
 * C ABI of libpxh.so (host C++, links libpxd.so).  It replaces the reference's
 * CPU-only consumers of a decomposition:
 *   ExpanderHierarchy   lib/algorithms/expander_hierarchy.h:13-160
 *   NormalizedCut       lib/algorithms/normalized_cut.h:12-145
 * and fills the empty useExternal hook of ExpanderHierarchy::build
 * (expander_hierarchy.h:91-101) with the HIP engine (pxd.h).
 *
 * A level l holds the partition of the graph below it (the input graph for
 * l = 0, the graph of level l-1 otherwise) and the multigraph it contracts to:
 * one node per cluster, one arc per cut arc, to the cluster of its far end.
 *   volumes[c]   level 0: sum of the input degrees of cluster c; above: sum of
 *                the volumes of its members (expander_hierarchy.h:49-58)
 *   cut_sizes[c] arcs leaving cluster c in the graph below (= its degree in
 *                the contracted graph, expander_hierarchy.h:64-76)
 *   parents[v]   cluster of node v of the graph below (clusterLookup)
 * Every function returns 0 on success, nonzero with pxh_last_error() set. */
#ifndef PAREX_AMD_PXH_H
#define PAREX_AMD_PXH_H

#include <stdint.h>

#include "pxd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pxh_hierarchy pxh_hierarchy;

const char* pxh_last_error(void);

/* ExpanderHierarchy(Graph*) (expander_hierarchy.h:105): keeps a copy of the
 * input CSR (n nodes, m undirected edges, 2m arcs). */
int pxh_create(uint32_t n, uint32_t m, const uint32_t* ranges, const uint32_t* edges, pxh_hierarchy** out);
/* Frees the hierarchy and the engine its last pxh_build used. */
int pxh_destroy(pxh_hierarchy* h);

/* ExpanderHierarchy::build<true> (expander_hierarchy.h:107-140) with the engine
 * as the decomposition: decompose the top graph on the GPU with
 * (target_sparsity, rw_threshold) as (sc_threshold, rw_threshold); when more
 * than 95 % of its edges are cut, multiply target_sparsity by 0.8 and retry;
 * otherwise contract and push a level.  Stops when the top graph has at most
 * one node, or no edges (the reference loops forever there), and fails after
 * max_retries sparsity cuts in a row.  params may be NULL (engine defaults).
 * One engine serves every level (each pxd_initialize reuses the device
 * buffers of the level before) and is kept until pxh_destroy or the next
 * pxh_build.  Levels pushed before a failure are kept. */
int pxh_build(pxh_hierarchy* h, const pxd_params* params, float target_sparsity, float rw_threshold,
              uint32_t max_retries);

/* ExpanderHierarchy::contract (expander_hierarchy.h:29-88) for a given
 * partition of the top graph: cluster c's members are
 * members[member_off[c] .. member_off[c+1]), in the order their cut arcs are
 * listed; every node must appear exactly once. */
int pxh_add_level(pxh_hierarchy* h, uint32_t num_clusters, const uint32_t* member_off, const uint32_t* members);

/* ExpanderHierarchy::height (expander_hierarchy.h:146-148). */
int pxh_height(pxh_hierarchy* h, uint32_t* height);

/* Sizes of level l: nodes (clusters), arcs of its graph (2 x cut edges) and
 * nodes of the graph below. */
int pxh_level_size(pxh_hierarchy* h, uint32_t level, uint32_t* nodes, uint64_t* arcs, uint32_t* below_nodes);

/* Copies level l out (any pointer may be NULL): ranges[nodes+1], edges[arcs],
 * volumes[nodes], cut_sizes[nodes], parents[below_nodes],
 * member_off[nodes+1], members[below_nodes]. */
int pxh_level(pxh_hierarchy* h, uint32_t level, uint32_t* ranges, uint32_t* edges, uint32_t* volumes,
              uint32_t* cut_sizes, uint32_t* parents, uint32_t* member_off, uint32_t* members);

/* ExpanderHierarchy::subtractVolume / addCut (expander_hierarchy.h:154-159):
 * u32 arithmetic on the volume or cut size of node nix of level l. */
int pxh_subtract_volume(pxh_hierarchy* h, uint32_t level, uint32_t nix, uint32_t volume);
int pxh_add_cut(pxh_hierarchy* h, uint32_t level, uint32_t nix, uint32_t cut);

/* NormalizedCut::compute(k) (normalized_cut.h:108-144): k greedy splits of the
 * hierarchy tree from its single top node, in the reference's float
 * arithmetic and tie order.  labels[n] receives each input node's part in
 * [0, k).  value (may be NULL) receives compute_normalized_cut of that
 * partition (lib/core/partition.h:418-427): the sum over parts in label order
 * of cut / volume, parts of volume 0 skipped.  The reference's compute
 * subtracts the split volumes from the hierarchy itself; this one works on a
 * copy, so repeated calls give the same labels.  Needs a top level with one
 * node. */
int pxh_normalized_cut(pxh_hierarchy* h, uint32_t k, int32_t* labels, float* value);

#ifdef __cplusplus
}
#endif

#endif

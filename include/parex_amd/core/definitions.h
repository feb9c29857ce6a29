// Core types of the expander-decomposition API, layout-compatible with the
// reference's lib/core/definitions.h:23-107 so that callers written against
// it compile and link against this engine unchanged.
//
// The include guard is the reference header's own (lib/core/definitions.h:5):
// a translation unit that reaches both this header and the reference's (its
// lib/core/partition.h includes "graph.h" by a relative path) gets whichever
// comes first and skips the other.  Both declare the same names, types and
// values, so either order compiles to the same program.
#ifndef RCUT_DEFINITIONS_H
#define RCUT_DEFINITIONS_H

#include <cassert>
#include <cstdint>
#include <iostream>
#include <vector>

// Logging macros of the reference (lib/core/definitions.h:9-19); its
// algorithm headers use them.
#ifdef DEBUG
#define INFO(msg) (std::cout << "[INFO] " << msg << std::endl)
#define LOG(msg) (std::cout << "[LOG] " << msg << std::endl)
#else
#define INFO(msg) ((void)0)
#define LOG(msg) ((void)0)
#endif
#define WARN(msg) (std::cout << "[WARNING] " << msg << "!" << std::endl)

using NodeIx = unsigned int;
using EdgeIx = unsigned int;
using EdgeWeight = unsigned int;
using frac_t = float;

// Launch geometry of the reference's CUDA kernels (lib/core/definitions.h:
// 26-41).  Only its cuda/ implementation, which this engine replaces, reads
// them; they stay for source compatibility.  The HIP engine launches 64-lane
// wavefronts and sizes its own grids.
constexpr int threads = 256;
constexpr int WARP = 32;
constexpr int warpsPerBlock = threads / WARP;
constexpr int WARPMASK = WARP - 1;
constexpr unsigned int BASE_SUBWARP_MASK = 0xFFFFFFFFu;
inline unsigned int getGridSize(unsigned int n) { return (n + threads - 1) / threads; }
inline unsigned int getGridSizeWarpParallel(unsigned int n) { return (n + warpsPerBlock - 1) / warpsPerBlock; }

// Process-wide parameters (lib/core/definitions.h:46-50): the CLI writes them,
// CudaDeviceManager reads randSeed at initialize() and both thresholds on
// every step call.
inline unsigned int randSeed{6582};
constexpr double rw_stay = 0.1;
inline frac_t rw_threshold = 1e-4f;
inline float sc_threshold = 0.3f;

constexpr NodeIx INVALID_EDGE = static_cast<NodeIx>(-1);

using label_t = int32_t;

struct SweepCutData {      // lib/core/definitions.h:57-61
    label_t clusterId;
    float sparsity;
    NodeIx offset;
};

struct LabeledNode {       // lib/core/definitions.h:63-66
    NodeIx nix;
    label_t clusterId;
};

struct NodeData {          // lib/core/definitions.h:68-77
    NodeIx nix;
    EdgeIx prefixVolume;
    int32_t prefixEdgeDiff;
    NodeIx offsetInCluster;
};

struct ClusterData {       // lib/core/definitions.h:79-84
    float rwSum;
    float maxPotential;
    float minPotential;
    NodeIx totalElements;
};

struct AllSweepCuts {      // lib/core/definitions.h:98-102
    std::vector<label_t> clusterIds;
    std::vector<SweepCutData> cuts;
};

struct FinalPartition {    // lib/core/definitions.h:104-107
    std::vector<label_t> clusterIds;
    int numClusters;
};

#endif

// Core types of the expander-decomposition API, layout-compatible with the
// reference's lib/core/definitions.h:23-107 so that callers written against
// it compile and link against this engine unchanged.
#ifndef PAREX_AMD_CORE_DEFINITIONS_H
#define PAREX_AMD_CORE_DEFINITIONS_H

#include <cstdint>
#include <vector>

using NodeIx = unsigned int;
using EdgeIx = unsigned int;
using EdgeWeight = unsigned int;
using frac_t = float;
using label_t = int32_t;

// Process-wide parameters (lib/core/definitions.h:46-50): the CLI writes them,
// CudaDeviceManager reads randSeed at initialize() and both thresholds on
// every step call.
inline unsigned int randSeed{6582};
constexpr double rw_stay = 0.1;
inline frac_t rw_threshold = 1e-4f;
inline float sc_threshold = 0.3f;

constexpr NodeIx INVALID_EDGE = static_cast<NodeIx>(-1);

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

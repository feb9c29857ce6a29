// Drop-in C++ host API of the expander decomposition: the same class, members
// and types as the reference's cuda/interface.h:11-60, implemented over the
// MI355X engine's C ABI (include/parex_amd/pxd.h) by parex_amd/csrc/facade.cpp.
// The include guard is the reference header's own (cuda/interface.h:5), so a
// caller that reaches both declares the class once.
#ifndef PAREX_INTERFACE_H
#define PAREX_INTERFACE_H

#include <memory>
#include <vector>

#include "../core/graph.h"

struct NodeUpdate {
    NodeIx nix;
    EdgeIx diff;
};

class CudaDeviceManager {
public:
    CudaDeviceManager();
    ~CudaDeviceManager();

    CudaDeviceManager(const CudaDeviceManager&) = delete;
    CudaDeviceManager& operator=(const CudaDeviceManager&) = delete;

    void initialize(const Graph& graph);
    Graph downloadGraph();
    std::vector<EdgeIx> downloadDegrees();
    std::vector<NodeData> downloadPartition();

    void iterateRandomWalk();
    std::vector<frac_t> readRandomWalkValues();

    void computeSweepCuts();
    AllSweepCuts readSweepCuts();

    void cutClusters();

    EdgeIx getNumCutEdges();

    void expanderDecomposition();

    FinalPartition getFinalPartition();

    std::vector<int> downloadActiveEdgeMap();
    std::vector<int> downloadLabels();
    int getNumClusters();

    void updateLabels(std::vector<NodeIx>& nodeLabels, NodeIx activeClusters);

private:
    struct Impl;
    std::unique_ptr<Impl> impl;
};

// declared by the reference (cuda/interface.h:60) and never defined there
int testCuda();

#endif

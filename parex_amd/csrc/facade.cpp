// CudaDeviceManager (include/parex_amd/cuda/interface.h) over the engine's C
// ABI: the member-by-member drop-in for the reference's cuda/device.h:394-427.
// Parameters come from the process-wide globals exactly when the reference
// reads them: randSeed at initialize(), the two thresholds on every step.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <stdexcept>
#include <string>

#include "parex_amd/cuda/interface.h"
#include "parex_amd/pxd.h"

namespace {
void check(int rc) {
    if (rc != 0) throw std::runtime_error(std::string("pxd: ") + pxd_last_error());
}
constexpr uint32_t kMaxIter = 200000;   // MAX_NUM_ITER, cuda/device.h:309

// The reference's cap is a constant; PAREX_MAX_ITER (a positive integer)
// lowers or raises it for this process, e.g. to exercise the capped exit.
uint32_t max_iterations() {
    const char* s = std::getenv("PAREX_MAX_ITER");
    if (!s || !*s) return kMaxIter;
    char* end = nullptr;
    const unsigned long v = std::strtoul(s, &end, 10);
    if (*end != '\0' || v == 0 || v > 0x7FFFFFFFul)
        throw std::runtime_error("PAREX_MAX_ITER must be a positive integer below 2^31");
    return (uint32_t)v;
}
}  // namespace

struct CudaDeviceManager::Impl {
    pxd_engine* e = nullptr;
    NodeIx n = 0;
    EdgeIx m = 0;
    uint32_t max_iter = kMaxIter;

    ~Impl() { pxd_destroy(e); }

    void create() {
        pxd_destroy(e);
        e = nullptr;
        pxd_params p;
        pxd_default_params(&p);
        p.seed = randSeed;
        p.sc_threshold = sc_threshold;
        p.rw_threshold = rw_threshold;
        max_iter = max_iterations();
        p.max_iter = max_iter;
        check(pxd_create(&p, &e));
    }
    pxd_engine* eng() {
        if (!e) throw std::runtime_error("CudaDeviceManager: initialize() was not called");
        check(pxd_set_thresholds(e, sc_threshold, rw_threshold));
        return e;
    }
};

CudaDeviceManager::CudaDeviceManager() : impl(std::make_unique<Impl>()) {}
CudaDeviceManager::~CudaDeviceManager() = default;

void CudaDeviceManager::initialize(const Graph& graph) {
    impl->create();
    impl->n = graph.numNodes;
    impl->m = graph.numEdges;
    static const uint32_t zero = 0;
    check(pxd_initialize(impl->e, graph.numNodes, graph.numEdges,
                         graph.ranges.empty() ? &zero : graph.ranges.data(),
                         graph.edges.empty() ? &zero : graph.edges.data()));
}

Graph CudaDeviceManager::downloadGraph() {
    Graph g(impl->n, impl->m);
    check(pxd_download_graph(impl->eng(), g.ranges.data(), g.edges.data()));
    return g;
}

std::vector<EdgeIx> CudaDeviceManager::downloadDegrees() {
    std::vector<EdgeIx> d(impl->n);
    check(pxd_download_degrees(impl->eng(), d.data()));
    return d;
}

std::vector<NodeData> CudaDeviceManager::downloadPartition() {
    std::vector<NodeData> nd(impl->n, NodeData{0, 0, 0, 0});
    uint32_t cnt = 0;
    static_assert(sizeof(NodeData) == sizeof(pxd_node_data), "NodeData layout");
    check(pxd_download_partition(impl->eng(), reinterpret_cast<pxd_node_data*>(nd.data()), impl->n, &cnt));
    return nd;
}

void CudaDeviceManager::iterateRandomWalk() { check(pxd_iterate_random_walk(impl->eng())); }

std::vector<frac_t> CudaDeviceManager::readRandomWalkValues() {
    std::vector<frac_t> v(impl->n);
    check(pxd_read_random_walk_values(impl->eng(), v.data()));
    return v;
}

void CudaDeviceManager::computeSweepCuts() { check(pxd_compute_sweep_cuts(impl->eng())); }

AllSweepCuts CudaDeviceManager::readSweepCuts() {
    uint32_t na = 0, nc = 0, ncut = 0;
    check(pxd_get_counts(impl->eng(), &na, &nc, &ncut));
    AllSweepCuts r;
    r.clusterIds.resize(nc);
    r.cuts.resize(nc);
    uint32_t cnt = 0;
    static_assert(sizeof(SweepCutData) == sizeof(pxd_sweep_cut), "SweepCutData layout");
    check(pxd_read_sweep_cuts(impl->e, r.clusterIds.data(), reinterpret_cast<pxd_sweep_cut*>(r.cuts.data()), nc,
                              &cnt));
    r.clusterIds.resize(cnt);
    r.cuts.resize(cnt);
    return r;
}

void CudaDeviceManager::cutClusters() { check(pxd_cut_clusters(impl->eng())); }

EdgeIx CudaDeviceManager::getNumCutEdges() {
    uint32_t c = 0;
    check(pxd_get_num_cut_edges(impl->eng(), &c));
    return c;
}

// cuda/device.h:301-350, including its console output
void CudaDeviceManager::expanderDecomposition() {
    std::cout << "Begin Expander Decomposition" << std::endl;
    int32_t it = 0;
    check(pxd_expander_decomposition(impl->eng(), &it));
    if ((uint32_t)it >= impl->max_iter) printf("Not terminating. ABORT.\n");
    printf("==================================================================================== \n"
           "iterations: %d\n", it);
}

FinalPartition CudaDeviceManager::getFinalPartition() {
    FinalPartition fp;
    fp.clusterIds.resize(impl->n);
    int32_t k = 0;
    check(pxd_get_final_partition(impl->eng(), fp.clusterIds.data(), &k));
    fp.numClusters = k;
    return fp;
}

std::vector<int> CudaDeviceManager::downloadActiveEdgeMap() {
    std::vector<int> a(2ull * impl->m);
    check(pxd_download_active_edge_map(impl->eng(), a.data()));
    return a;
}

std::vector<int> CudaDeviceManager::downloadLabels() {
    std::vector<int> l(impl->n);
    check(pxd_download_labels(impl->eng(), l.data()));
    return l;
}

// the reference returns 0 here (cuda/device.h:58-61)
int CudaDeviceManager::getNumClusters() { return 0; }

// the reference's updateLabels does nothing (cuda/devGraph.h:191-206)
void CudaDeviceManager::updateLabels(std::vector<NodeIx>&, NodeIx) {}

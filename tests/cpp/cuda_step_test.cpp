// Step-by-step test of the CudaDeviceManager facade against the oracle, in
// the pattern of the reference's tests/cudaTest.cpp:33-232 (UploadGraph,
// Degrees, RandomWalkTest, SweepCutTest, SweepCut, CutTest), driven through
// the C++ class only.  The reference checks its device path against its CPU
// classes within tolerances; here the engine is checked against the oracle
// (a CPU restatement of the device algorithm), which makes every comparison
// exact.  The oracle is seeded with the engine's initial walk vector: the
// initial values are floats stored as doubles, so readRandomWalkValues()
// returns them exactly (the reference's own seeding, tests/cudaTest.cpp:50-56).
//
//   cuda_step_test <graph.mtx> [seed] [sc_threshold] [rw_threshold]
//
// Prints gtest-like lines and exits with the number of failed tests.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include "parex_amd/cuda/interface.h"
#include "oracle.h"

namespace {

Graph graph;
int failures_in_test = 0;

#define EXPECT_TRUE(cond, ...)                                                   \
    do {                                                                         \
        if (!(cond)) {                                                           \
            ++failures_in_test;                                                  \
            std::printf("  %s:%d: expected %s: ", __FILE__, __LINE__, #cond);    \
            std::printf(__VA_ARGS__);                                            \
            std::printf("\n");                                                   \
        }                                                                        \
    } while (0)
#define ASSERT_TRUE(cond, ...)                 \
    do {                                       \
        const int before_ = failures_in_test;  \
        EXPECT_TRUE(cond, __VA_ARGS__);        \
        if (failures_in_test != before_) return; \
    } while (0)

// lib/utils/graph_io.h:49-84: '%' lines, "rows n m", m one-based pairs, then
// DynamicGraph::finalize
Graph read_mtx(const std::string& path) {
    std::ifstream in(path);
    if (!in) {
        std::fprintf(stderr, "cannot open %s\n", path.c_str());
        std::exit(100);
    }
    std::string line;
    while (std::getline(in, line) && line.rfind('%', 0) == 0) {
    }
    std::stringstream ss(line);
    NodeIx rows = 0, n = 0;
    EdgeIx m = 0;
    ss >> rows >> n >> m;
    DynamicGraph dg;
    dg.numNodes = n;
    for (EdgeIx i = 0; i < m; ++i) {
        NodeIx a, b;
        in >> a >> b;
        dg.addEdge(a - 1, b - 1);
    }
    return dg.finalize();
}

bool same_f32(float a, float b) {
    uint32_t x, y;
    std::memcpy(&x, &a, 4);
    std::memcpy(&y, &b, 4);
    return x == y;
}

struct OracleRun {
    ox_state* s = nullptr;
    explicit OracleRun(const std::vector<frac_t>& x0) {
        std::vector<double> init(x0.begin(), x0.end());
        s = ox_create(graph.numNodes, graph.numEdges, graph.ranges.data(),
                      graph.edges.empty() ? nullptr : graph.edges.data(), randSeed, sc_threshold, rw_threshold,
                      init.empty() ? nullptr : init.data());
    }
    ~OracleRun() { ox_destroy(s); }
    std::vector<double> values() const {
        std::vector<double> v(graph.numNodes);
        ox_get_values(s, v.data());
        return v;
    }
};

// the walk state after each step: equal as floats (equal FP64 bits round alike)
void expect_values(CudaDeviceManager& cuda, const OracleRun& ox, const char* what) {
    const std::vector<frac_t> y = cuda.readRandomWalkValues();
    const std::vector<double> z = ox.values();
    size_t bad = 0, first = 0;
    for (NodeIx u = 0; u < graph.numNodes; ++u)
        if (!same_f32(y[u], (float)z[u]) && bad++ == 0) first = u;
    EXPECT_TRUE(bad == 0, "%s: %zu walk values differ, first at node %zu (%g vs %g)", what, bad, first,
                (double)y[first], z[first]);
}

void expect_sweep_cuts(CudaDeviceManager& cuda, const OracleRun& ox, const char* what) {
    const AllSweepCuts r = cuda.readSweepCuts();
    const uint32_t nc = ox_num_active_clusters(ox.s);
    std::vector<ox_sweepcut> want(nc ? nc : 1);
    ox_get_sweepcuts(ox.s, want.data());
    EXPECT_TRUE(r.clusterIds.size() == nc && r.cuts.size() == nc, "%s: %zu sweep cuts, oracle %u", what,
                r.cuts.size(), nc);
    if (r.cuts.size() != nc) return;
    for (uint32_t c = 0; c < nc; ++c) {
        EXPECT_TRUE(r.clusterIds[c] == want[c].clusterId && r.cuts[c].clusterId == want[c].clusterId,
                    "%s: cluster %u label %d vs %d", what, c, r.clusterIds[c], want[c].clusterId);
        EXPECT_TRUE(same_f32(r.cuts[c].sparsity, want[c].sparsity), "%s: cluster %u sparsity %.9g vs %.9g", what, c,
                    (double)r.cuts[c].sparsity, (double)want[c].sparsity);
        EXPECT_TRUE(r.cuts[c].offset == want[c].offset, "%s: cluster %u offset %u vs %u", what, c, r.cuts[c].offset,
                    want[c].offset);
    }
}

void expect_node_data(CudaDeviceManager& cuda, const OracleRun& ox, const char* what) {
    const std::vector<NodeData> nd = cuda.downloadPartition();
    const uint32_t na = ox_num_active_nodes(ox.s);
    std::vector<ox_ndata> want(na ? na : 1);
    ox_get_ndata(ox.s, want.data());
    size_t bad = 0;
    for (uint32_t k = 0; k < na && k < nd.size(); ++k)
        bad += nd[k].nix != want[k].nix || nd[k].prefixVolume != want[k].prefixVolume ||
               nd[k].prefixEdgeDiff != want[k].prefixEdgeDiff || nd[k].offsetInCluster != want[k].offsetInCluster;
    EXPECT_TRUE(na <= nd.size() && bad == 0, "%s: %zu NodeData entries differ", what, bad);
}

void expect_degrees(CudaDeviceManager& cuda, const OracleRun& ox, const char* what) {
    std::vector<uint32_t> want(graph.numNodes);
    ox_get_internal_degrees(ox.s, want.data());
    const std::vector<EdgeIx> deg = cuda.downloadDegrees();
    EXPECT_TRUE(deg == want, "%s: internal degrees differ", what);
}

// -------------------------------------------------------------- tests

void UploadGraph(CudaDeviceManager& cuda) {
    const Graph g2 = cuda.downloadGraph();
    EXPECT_TRUE(g2 == graph, "downloaded graph differs from the uploaded one");
}

void Degrees(CudaDeviceManager& cuda) {
    std::vector<EdgeIx> expected(graph.numNodes);
    for (NodeIx u = 0; u < graph.numNodes; ++u) expected[u] = graph.degree(u);
    EXPECT_TRUE(cuda.downloadDegrees() == expected, "degrees differ");
}

void RandomWalkTest(CudaDeviceManager& cuda, int steps) {
    cuda.initialize(graph);
    OracleRun ox(cuda.readRandomWalkValues());
    for (int i = 0; i < steps; ++i) {
        cuda.iterateRandomWalk();
        ox_walk(ox.s);
    }
    expect_values(cuda, ox, "walk");
}

void SweepCutTest(CudaDeviceManager& cuda) {
    cuda.initialize(graph);
    OracleRun ox(cuda.readRandomWalkValues());
    for (int i = 0; i < 20; ++i) {
        cuda.iterateRandomWalk();
        ox_walk(ox.s);
    }
    expect_degrees(cuda, ox, "20 steps");
    expect_values(cuda, ox, "20 steps");
    cuda.computeSweepCuts();
    ox_sweep(ox.s);
    const AllSweepCuts r = cuda.readSweepCuts();
    EXPECT_TRUE(r.clusterIds.size() == 1 && r.cuts.size() == 1 && r.clusterIds[0] == 0,
                "one cluster with label 0 before any cut (%zu)", r.cuts.size());
    expect_sweep_cuts(cuda, ox, "sweep after 20 steps");
    expect_node_data(cuda, ox, "sweep after 20 steps");
}

void SweepCut(CudaDeviceManager& cuda, int steps) {
    cuda.initialize(graph);
    OracleRun ox(cuda.readRandomWalkValues());
    for (int i = 0; i < steps; ++i) {
        cuda.iterateRandomWalk();
        ox_walk(ox.s);
        cuda.computeSweepCuts();
        ox_sweep(ox.s);
        char what[64];
        std::snprintf(what, sizeof what, "step %d", i + 1);
        expect_sweep_cuts(cuda, ox, what);
    }
}

void CutTest(CudaDeviceManager& cuda) {
    cuda.initialize(graph);
    OracleRun ox(cuda.readRandomWalkValues());
    for (int i = 0; i < 10; ++i) {
        cuda.iterateRandomWalk();
        ox_walk(ox.s);
    }
    expect_values(cuda, ox, "10 steps");
    cuda.computeSweepCuts();
    ox_sweep(ox.s);
    expect_sweep_cuts(cuda, ox, "sweep");
    cuda.cutClusters();
    ox_cut(ox.s);
    std::vector<int32_t> lab(graph.numNodes);
    ox_get_labels(ox.s, lab.data());
    EXPECT_TRUE(cuda.downloadLabels() == lab, "labels after the cut differ");
    expect_degrees(cuda, ox, "after the cut");
    std::vector<int32_t> ids(graph.numNodes ? graph.numNodes : 1);
    const int k = ox_final_partition(ox.s, ids.data());
    ids.resize(graph.numNodes);
    const FinalPartition fp = cuda.getFinalPartition();
    EXPECT_TRUE(fp.numClusters == k && fp.clusterIds == ids, "final partition after one cut: %d clusters vs %d",
                fp.numClusters, k);
    EXPECT_TRUE(cuda.getNumCutEdges() == ox_num_cut_edges(ox.s), "getNumCutEdges %u vs %u", cuda.getNumCutEdges(),
                ox_num_cut_edges(ox.s));
    std::vector<uint32_t> nbr(2ull * graph.numEdges);
    ox_get_neighbors(ox.s, nbr.data());
    const Graph g2 = cuda.downloadGraph();
    EXPECT_TRUE(g2.edges == nbr && g2.ranges == graph.ranges, "INVALID-marked graph differs");
    const std::vector<int> map = cuda.downloadActiveEdgeMap();
    size_t bad = 0;
    for (size_t j = 0; j < nbr.size(); ++j) bad += map[j] != (nbr[j] != OX_INVALID_EDGE ? 1 : 0);
    EXPECT_TRUE(map.size() == nbr.size() && bad == 0, "active edge map: %zu entries differ", bad);
}

void ExpanderDecomposition(CudaDeviceManager& cuda) {
    cuda.initialize(graph);
    OracleRun ox(cuda.readRandomWalkValues());
    cuda.expanderDecomposition();
    const int it = ox_run(ox.s, 200000);
    std::vector<int32_t> ids(graph.numNodes ? graph.numNodes : 1);
    const int k = ox_final_partition(ox.s, ids.data());
    ids.resize(graph.numNodes);
    const FinalPartition fp = cuda.getFinalPartition();
    EXPECT_TRUE(fp.numClusters == k && fp.clusterIds == ids, "final partition: %d clusters vs %d (oracle %d iterations)",
                fp.numClusters, k, it);
    EXPECT_TRUE(cuda.getNumClusters() == 0, "getNumClusters() returns 0 as the reference does");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <graph.mtx> [seed] [sc_threshold] [rw_threshold]\n", argv[0]);
        return 100;
    }
    if (argc > 2) randSeed = (unsigned)std::strtoul(argv[2], nullptr, 10);
    if (argc > 3) sc_threshold = std::strtof(argv[3], nullptr);
    if (argc > 4) rw_threshold = std::strtof(argv[4], nullptr);
    graph = read_mtx(argv[1]);
    std::printf("graph %s: %u nodes, %u edges\n", argv[1], graph.numNodes, graph.numEdges);

    // one instance for every test, as in tests/cudaTest.cpp:18-31 (a local
    // here, so that it is destroyed before the HIP runtime shuts down)
    CudaDeviceManager cuda;
    cuda.initialize(graph);
    struct Case {
        std::string name;
        std::function<void()> run;
    };
    std::vector<Case> cases = {
        {"CudaTest.UploadGraph", [&] { UploadGraph(cuda); }},
        {"CudaTest.Degrees", [&] { Degrees(cuda); }},
        {"CudaTest.SweepCutTest", [&] { SweepCutTest(cuda); }},
        {"CudaTest.CutTest", [&] { CutTest(cuda); }},
        {"CudaTest.ExpanderDecomposition", [&] { ExpanderDecomposition(cuda); }},
    };
    for (int s : {0, 1, 2, 4, 8, 16, 64, 128})
        cases.push_back({"Iterations/CudaTest.RandomWalkTest/" + std::to_string(s), [&cuda, s] { RandomWalkTest(cuda, s); }});
    for (int s : {1, 4, 16})
        cases.push_back({"Iterations/CudaTest.SweepCut/" + std::to_string(s), [&cuda, s] { SweepCut(cuda, s); }});

    int failed = 0;
    for (const Case& c : cases) {
        std::printf("[ RUN      ] %s\n", c.name.c_str());
        std::fflush(stdout);
        failures_in_test = 0;
        try {
            c.run();
        } catch (const std::exception& ex) {
            ++failures_in_test;
            std::printf("  exception: %s\n", ex.what());
        }
        std::printf(failures_in_test ? "[  FAILED  ] %s\n" : "[       OK ] %s\n", c.name.c_str());
        std::fflush(stdout);
        failed += failures_in_test != 0;
    }
    std::printf("[==========] %zu tests, %d failed\n", cases.size(), failed);
    return failed;
}

// Parex CLI (main.cpp:15-64 of the reference): positional
//   parex <graph.mtx|graph.pxg> [seed] [sweep_cut_threshold] [random_walk_threshold]
// with the reference's console output, byte for byte.
#include <chrono>
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "parex_amd/cuda/interface.h"
#include "parex_amd/pxgen.h"

static bool ends_with(const std::string& s, const std::string& t) {
    return s.size() >= t.size() && s.compare(s.size() - t.size(), t.size(), t) == 0;
}

int main(int argc, char* argv[]) {
    if (argc < 2) {
        std::cout << "missing arguments. expected <filename> <?seed> <sc_threshold> <rw_threshold>. Exiting."
                  << std::endl;
        return 1;   // the reference goes on to read argv[1]
    }
    std::string filename = argv[1];
    if (argc > 2) randSeed = (unsigned int)std::stoull(argv[2]);
    if (argc > 3) sc_threshold = std::stof(argv[3]);
    if (argc > 4) rw_threshold = std::stof(argv[4]);

    printf("sparsity: %f\nthreshold: %f\n", sc_threshold, rw_threshold);

    pxg_graph pg{};
    int rc = ends_with(filename, ".pxg") ? pxg_read(filename.c_str(), &pg) : pxg_read_mtx(filename.c_str(), &pg);
    if (rc != 0) {
        std::cerr << pxg_last_error() << std::endl;
        return 1;
    }
    std::vector<NodeIx> edges(pg.edges, pg.edges + 2ull * pg.m);
    std::vector<EdgeIx> ranges(pg.ranges, pg.ranges + pg.n + 1ull);
    Graph G(std::move(edges), std::move(ranges), pg.n, pg.m);

    auto t0 = std::chrono::steady_clock::now();
    CudaDeviceManager cuda;
    cuda.initialize(G);
    cuda.expanderDecomposition();
    long long timeMillis =
        std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0).count();
    FinalPartition fpt = cuda.getFinalPartition();

    uint64_t cut = 0;
    if (pxg_verify_partition(&pg, fpt.clusterIds.data(), fpt.numClusters, &cut, nullptr, nullptr, nullptr) != 0) {
        std::cerr << "partition check failed: " << pxg_last_error() << std::endl;
        pxg_free(&pg);
        return 2;
    }
    printf("time: %lld\nnumClusters: %d\ncutEdges: %d\nnodes: %d\nedges: %d\n", timeMillis, fpt.numClusters, (int)cut,
           G.numNodes, G.numEdges);
    pxg_free(&pg);
    return 0;
}

// Driver for the reference's own CPU code (test infrastructure / CPU baseline).
// Everything algorithmic comes from the reference headers included below; this
// file only reads graphs in the .pxg binary CSR format and prints results.
//
//   parex_cpu_ref decompose <budget_s> <seed> <sc> <rw> <g1.pxg> [g2.pxg ...]
//       lib/algorithms/expander_decomposition.h:70-168 with Config{sc, rw, 1, 1}
//       and random_source seeded as main.cpp:22-23 does; stops starting new
//       graphs once budget_s seconds are spent.  Prints one JSON line.
//   parex_cpu_ref walk <g.pxg> <in.f32> <steps> <out.f32>
//       lib/algorithms/random_walk.h:51-70 (RandomWalk::iterate over cluster 0)
//   parex_cpu_ref sweep <g.pxg> <in.f32>
//       lib/core/partition.h:309-362 (Partition::sweepCut on cluster 0)
//   parex_cpu_ref verify <g.pxg> <ids.i32> <numClusters>
//       lib/core/partition.h:233-295 (splitByIndices<false,false>, main.cpp:54-60)
//   parex_cpu_ref hierarchy <g.pxg> <seed> <k>
//       lib/algorithms/expander_hierarchy.h:107-140 (build<false>, CPU
//       decompositions) then lib/algorithms/normalized_cut.h:108-144 (compute(k))
//       and lib/core/partition.h:418-427; prints every level, the size, volume
//       and internal volume of each of the k parts and the objective as one
//       JSON line.  The graph must be connected.
#include "algorithms/expander_decomposition.h"
#include "algorithms/normalized_cut.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

static Graph load_pxg(const char* path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { fprintf(stderr, "cannot open %s\n", path); exit(2); }
    char magic[4];
    uint32_t n = 0, m = 0;
    f.read(magic, 4);
    f.read(reinterpret_cast<char*>(&n), 4);
    f.read(reinterpret_cast<char*>(&m), 4);
    if (std::memcmp(magic, "PXG1", 4) != 0) { fprintf(stderr, "bad magic in %s\n", path); exit(2); }
    std::vector<NodeIx> edges(2ull * m);
    std::vector<EdgeIx> ranges(n + 1ull);
    f.read(reinterpret_cast<char*>(ranges.data()), (n + 1ull) * 4);
    f.read(reinterpret_cast<char*>(edges.data()), 2ull * m * 4);
    return Graph(std::move(edges), std::move(ranges), n, m);
}

template <typename T>
static std::vector<T> load_vec(const char* path, size_t count) {
    std::vector<T> v(count);
    std::ifstream f(path, std::ios::binary);
    f.read(reinterpret_cast<char*>(v.data()), count * sizeof(T));
    return v;
}

int main(int argc, char** argv) {
    if (argc < 2) { fprintf(stderr, "usage: see source\n"); return 2; }
    std::string mode = argv[1];
    if (mode == "decompose" && argc >= 7) {
        double budget = std::atof(argv[2]);
        unsigned seed = (unsigned)std::strtoul(argv[3], nullptr, 10);
        float sc = std::strtof(argv[4], nullptr), rw = std::strtof(argv[5], nullptr);
        double spent = 0.0;
        unsigned long long edges = 0;
        int graphs = 0;
        std::string per;
        for (int a = 6; a < argc && spent < budget; ++a) {
            Graph G = load_pxg(argv[a]);
            random_source.seed(seed);
            auto t0 = std::chrono::steady_clock::now();
            Partition p = expanderDecomposition(G, Config{sc, rw, 1, 1});
            auto t1 = std::chrono::steady_clock::now();
            double dt = std::chrono::duration<double>(t1 - t0).count();
            spent += dt;
            edges += G.numEdges;
            ++graphs;
            char buf[160];
            snprintf(buf, sizeof buf, "%s{\"m\":%u,\"s\":%.6f,\"clusters\":%u,\"cut\":%u}",
                     per.empty() ? "" : ",", G.numEdges, dt, p.numClusters(), p.getNumCutEdges());
            per += buf;
        }
        printf("{\"graphs\":%d,\"edges\":%llu,\"seconds\":%.6f,\"per_graph\":[%s]}\n",
               graphs, edges, spent, per.c_str());
        return 0;
    }
    if (mode == "walk" && argc >= 6) {
        Graph G = load_pxg(argv[2]);
        std::vector<frac_t> x = load_vec<frac_t>(argv[3], G.numNodes);
        int steps = std::atoi(argv[4]);
        RandomWalk rw(G.numNodes);
        rw.setData(x);
        Partition part(&G);
        for (int i = 0; i < steps; ++i) rw.iterate(part, {0});
        std::ofstream o(argv[5], std::ios::binary);
        for (NodeIx i = 0; i < G.numNodes; ++i) { float v = rw.values()[i]; o.write(reinterpret_cast<char*>(&v), 4); }
        return 0;
    }
    if (mode == "sweep" && argc >= 4) {
        Graph G = load_pxg(argv[2]);
        std::vector<frac_t> x = load_vec<frac_t>(argv[3], G.numNodes);
        Partition part(&G);
        SweepCut sc = part.sweepCut(0, x);
        printf("{\"sparsity\":%.9g,\"offset\":%u}\n", sc.sparsity, sc.offset);
        return 0;
    }
    if (mode == "verify" && argc >= 5) {
        Graph G = load_pxg(argv[2]);
        std::vector<int32_t> ids = load_vec<int32_t>(argv[3], G.numNodes);
        NodeIx k = (NodeIx)std::strtoul(argv[4], nullptr, 10);
        Partition part(&G);
        std::vector<NodeIx> mod;
        part.splitByIndices<false, false>(0, ids, k, mod);
        printf("{\"numClusters\":%u,\"cutEdges\":%u}\n", part.numClusters(), part.getNumCutEdges());
        return 0;
    }
    if (mode == "hierarchy" && argc >= 5) {
        Graph G = load_pxg(argv[2]);
        random_source.seed((unsigned)std::strtoul(argv[3], nullptr, 10));
        NodeIx k = (NodeIx)std::strtoul(argv[4], nullptr, 10);
        ExpanderHierarchy eh{&G};
        auto t0 = std::chrono::steady_clock::now();
        eh.build();
        auto t1 = std::chrono::steady_clock::now();
        auto arr = [](const char* name, const auto& v, bool comma) {
            printf("\"%s\":[", name);
            for (size_t i = 0; i < v.size(); ++i) printf(i ? ",%u" : "%u", (unsigned)v[i]);
            printf(comma ? "]," : "]");
        };
        printf("{\"levels\":[");
        for (NodeIx l = 0; l < eh.height(); ++l) {
            const auto& L = eh.lvl(l);
            std::vector<NodeIx> off{0}, members;
            for (const Cluster& c : L.partition) {
                for (const ClusterVertex& cv : c) members.push_back(cv.nix);
                off.push_back((NodeIx)members.size());
            }
            printf(l ? ",{" : "{");
            arr("member_off", off, true);
            arr("members", members, true);
            arr("ranges", L.graph.ranges, true);
            arr("edges", L.graph.edges, true);
            arr("volumes", L.volumes, true);
            arr("cut_sizes", L.cutSizes, true);
            arr("parents", L.parents, false);
            printf("}");
        }
        NormalizedCut nc{&eh};
        auto t2 = std::chrono::steady_clock::now();
        Partition kp = nc.compute(k);
        auto t3 = std::chrono::steady_clock::now();
        // splitByIndices<false,false> keeps the node order, so a part's member
        // range is positional; its size, volume and internal volume are those
        // of label i (partition.h:239-270), cluster i holding label i
        std::vector<NodeIx> size, vol, inner;
        for (const Cluster& c : kp) {
            size.push_back(c.size());
            vol.push_back(c.volume);
            inner.push_back(c.internalVolume);
        }
        printf("],");
        arr("nc_size", size, true);
        arr("nc_volume", vol, true);
        arr("nc_internal", inner, true);
        printf("\"build_s\":%.6f,\"nc_s\":%.6f,", std::chrono::duration<double>(t1 - t0).count(),
               std::chrono::duration<double>(t3 - t2).count());
        printf("\"nc_value\":%.9g}\n", compute_normalized_cut(kp));
        return 0;
    }
    fprintf(stderr, "bad arguments\n");
    return 2;
}

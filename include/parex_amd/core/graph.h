// CSR graph containers with the reference's layout and semantics
// (lib/core/graph.h:22-184): Graph holds both directions of every edge with
// rows sorted ascending; DynamicGraph::finalize builds one from an edge list.
#ifndef PAREX_AMD_CORE_GRAPH_H
#define PAREX_AMD_CORE_GRAPH_H

#include <algorithm>
#include <utility>
#include <vector>

#include "definitions.h"

struct Edge {
    NodeIx fx{0}, tx{0};
    Edge() = default;
    Edge(NodeIx f, NodeIx t) : fx(f), tx(t) {}
    bool operator==(const Edge& o) const { return fx == o.fx && tx == o.tx; }
};

class Graph {
public:
    std::vector<NodeIx> edges;        // 2 * numEdges
    std::vector<EdgeIx> ranges;       // numNodes + 1
    std::vector<EdgeWeight> weights;  // all 1
    NodeIx numNodes;
    EdgeIx numEdges;

    Graph() : numNodes{0}, numEdges{0} {}
    Graph(NodeIx n, EdgeIx m) : edges(2ull * m), ranges(n + 1ull), weights(2ull * m, 1), numNodes{n}, numEdges{m} {}
    Graph(std::vector<NodeIx>&& ed, std::vector<EdgeIx>&& rg, NodeIx n, EdgeIx m)
        : edges(std::move(ed)), ranges(std::move(rg)), weights(2ull * m, 1), numNodes{n}, numEdges{m} {}

    EdgeIx degree(NodeIx u) const { return ranges[u + 1] - ranges[u]; }
    bool operator==(const Graph& o) const {
        return edges == o.edges && ranges == o.ranges && weights == o.weights && numNodes == o.numNodes &&
               numEdges == o.numEdges;
    }
};

class DynamicGraph {
public:
    std::vector<Edge> edges{};
    NodeIx numNodes{0};
    EdgeIx numEdges{0};

    DynamicGraph() = default;
    DynamicGraph(std::vector<Edge>&& e, NodeIx n, EdgeIx m) : edges(std::move(e)), numNodes(n), numEdges(m) {}

    void addEdge(NodeIx f, NodeIx t) {
        edges.emplace_back(std::min(f, t), std::max(f, t));
        ++numEdges;
    }

    // both directions, sorted by (from, to); trailing isolated nodes get
    // ranges equal to 2m (lib/core/graph.h:163-172 leaves them at 0)
    Graph finalize() const {
        if (edges.empty()) return {};
        std::vector<Edge> full;
        full.reserve(2 * edges.size());
        for (const Edge& e : edges) {
            full.emplace_back(e.fx, e.tx);
            full.emplace_back(e.tx, e.fx);
        }
        std::sort(full.begin(), full.end(),
                  [](const Edge& a, const Edge& b) { return a.fx == b.fx ? a.tx < b.tx : a.fx < b.fx; });
        std::vector<EdgeIx> rg(numNodes + 1ull, 0);
        std::vector<NodeIx> eg(full.size());
        size_t eix = 0;
        for (NodeIx nix = 0; nix <= numNodes; ++nix) {
            rg[nix] = (EdgeIx)eix;
            while (nix < numNodes && eix < full.size() && full[eix].fx == nix) ++eix;
        }
        for (size_t i = 0; i < full.size(); ++i) eg[i] = full[i].tx;
        return Graph(std::move(eg), std::move(rg), numNodes, (EdgeIx)edges.size());
    }
};

#endif

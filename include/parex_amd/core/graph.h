// CSR graph containers with the reference's layout and semantics
// (lib/core/graph.h:22-184): Graph holds both directions of every edge with
// rows sorted ascending; DynamicGraph::finalize builds one from an edge list.
// Every member, constructor and macro of the reference's header is here, so
// its callers (lib/core/partition.h, lib/algorithms/*.h, main.cpp) compile
// against this header as they do against the reference's.
//
// The include guard is the reference header's own (lib/core/graph.h:5): a
// translation unit that reaches both headers uses whichever comes first.
#ifndef RCUT_GRAPH_H
#define RCUT_GRAPH_H

#include <algorithm>
#include <cstdint>
#include <iterator>
#include <limits>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include "definitions.h"

// neighbour and arc-index loops of the reference (lib/core/graph.h:16-19)
#define NEIGHBORS(__g, __i) auto __end_it = __g.nend(__i); for(auto _it = __g.nbegin(__i); _it != __end_it; ++_it)
#define NEIGHBORSPTR(__g, __i) auto __end_it = __g->nend(__i); for(auto _it = __g->nbegin(__i); _it != __end_it; ++_it)
#define ITER_EIX_P(__g, __i, __eix) for(EdgeIx __eix = __g->ranges.at(__i); __eix < __g->ranges.at(__i+1); __eix++)
#define ITER_EIX(__g, __i, __eix) for(EdgeIx __eix = __g.ranges.at(__i); __eix < __g.ranges.at(__i+1); __eix++)

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

    // row of `node` as iterators into edges (bounds-checked like the
    // reference's ranges.at)
    [[nodiscard]] std::vector<NodeIx>::iterator nbegin(NodeIx node) { return edges.begin() + ranges.at(node); }
    [[nodiscard]] std::vector<NodeIx>::iterator nend(NodeIx node) { return edges.begin() + ranges.at(node + 1); }
    [[nodiscard]] std::vector<NodeIx>::const_iterator nbegin(NodeIx node) const {
        return edges.begin() + ranges.at(node);
    }
    [[nodiscard]] std::vector<NodeIx>::const_iterator nend(NodeIx node) const {
        return edges.begin() + ranges.at(node + 1);
    }

    // rows are sorted, so membership is a binary search
    [[nodiscard]] bool hasNeighbor(NodeIx node, NodeIx nb) const {
        return std::binary_search(nbegin(node), nend(node), nb);
    }
    [[nodiscard]] EdgeIx startIndex(NodeIx node) const { return ranges.at(node); }
    [[nodiscard]] EdgeIx degree(NodeIx node) const {
        return static_cast<EdgeIx>(std::distance(nbegin(node), nend(node)));
    }

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
    // the edge is stored with its smaller end first; the last edge takes its place
    void removeEdge(NodeIx f, NodeIx t) {
        if (f > t) std::swap(f, t);
        auto it = std::find(edges.begin(), edges.end(), Edge{f, t});
        std::swap(*it, edges.back());
        popEdge();
    }
    void popEdge() {
        edges.pop_back();
        --numEdges;
    }
    void removeEdgeByIx(EdgeIx eix) {
        std::swap(edges.at(eix), edges.back());
        popEdge();
    }
    NodeIx addNode() { return numNodes++; }

    // both directions, sorted by (from, to); trailing isolated nodes get
    // ranges equal to 2m (lib/core/graph.h:163-172 leaves them at 0)
    [[nodiscard]] Graph finalize() const {
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

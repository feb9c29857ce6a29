// Expander hierarchy and normalized cut (host C++, C ABI in include/parex_amd/pxh.h).
//
// The hierarchy re-decomposes its own top graph with the HIP engine (pxd.h)
// until one node is left; contraction and the normalized cut are O(arcs) and
// O(k * tree size) host passes over the level arrays.
//
//   ExpanderHierarchy::contract   lib/algorithms/expander_hierarchy.h:29-88
//   ExpanderHierarchy::build      lib/algorithms/expander_hierarchy.h:107-140
//   NormalizedCut                 lib/algorithms/normalized_cut.h:12-145
//   compute_normalized_cut        lib/core/partition.h:418-427
#include "parex_amd/pxh.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
    g_err = msg;
    return PXD_ERR_ARG;
}

// No C++ exception may cross the C ABI: host allocation failures become
// PXD_ERR_OOM.
template <class F>
int guarded(const char* what, F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        g_err = std::string(what) + ": out of host memory";
        return PXD_ERR_OOM;
    } catch (const std::exception& ex) {
        g_err = std::string(what) + ": " + ex.what();
        return PXD_ERR_STATE;
    }
}

struct Level {
    uint32_t nodes = 0;                 // clusters of the partition = nodes of this level's graph
    std::vector<uint32_t> ranges;       // contracted multigraph, rows ascending
    std::vector<uint32_t> edges;
    std::vector<uint32_t> volumes;      // EdgeIx arithmetic (u32), as the reference keeps it
    std::vector<uint32_t> cut_sizes;
    std::vector<uint32_t> parents;      // cluster of each node of the graph below
    std::vector<uint32_t> member_off;   // partition of the graph below, members in listed order
    std::vector<uint32_t> members;
};

}  // namespace

struct pxh_hierarchy {
    uint32_t n = 0;
    std::vector<uint32_t> ranges, edges;   // the input graph
    std::vector<Level> levels;
    // the engine of the last pxh_build: every level re-initializes it, so the
    // device buffers of the first (largest) level serve all later ones; freed
    // by pxh_destroy
    pxd_engine* engine = nullptr;
    pxd_params engine_params{};         // the parameters `engine` was created with

    ~pxh_hierarchy() {
        if (engine) pxd_destroy(engine);
    }

    uint32_t top_nodes() const { return levels.empty() ? n : levels.back().nodes; }
    const std::vector<uint32_t>& top_ranges() const { return levels.empty() ? ranges : levels.back().ranges; }
    const std::vector<uint32_t>& top_edges() const { return levels.empty() ? edges : levels.back().edges; }
};

namespace {

// expander_hierarchy.h:29-88.  Level 0 volumes are input degrees (the cluster
// volume of a partition of the input graph); above, the members' volumes add
// up.  Each cut arc of a member becomes one arc to the cluster of its far end.
int contract(pxh_hierarchy* h, uint32_t k, const uint32_t* member_off, const uint32_t* members) {
    const uint32_t nb = h->top_nodes();
    const std::vector<uint32_t>& row = h->top_ranges();
    const std::vector<uint32_t>& nbr = h->top_edges();
    if (member_off[0] != 0 || member_off[k] != nb) return fail("pxh_add_level: member_off must run from 0 to the node count");
    // every offset checked before any member is read: ascending from 0 to nb
    // keeps every read of members below nb
    for (uint32_t c = 0; c < k; ++c)
        if (member_off[c] > member_off[c + 1]) return fail("pxh_add_level: member_off is not ascending");
    Level L;
    L.nodes = k;
    L.parents.assign(nb, UINT32_MAX);
    for (uint32_t c = 0; c < k; ++c) {
        for (uint32_t i = member_off[c]; i < member_off[c + 1]; ++i) {
            uint32_t v = members[i];
            if (v >= nb || L.parents[v] != UINT32_MAX) return fail("pxh_add_level: members is not a partition of the nodes");
            L.parents[v] = c;
        }
    }
    L.member_off.assign(member_off, member_off + k + 1);
    L.members.assign(members, members + nb);
    L.volumes.assign(k, 0);
    L.cut_sizes.assign(k, 0);
    L.ranges.assign(k + 1, 0);
    const bool first = h->levels.empty();
    const uint32_t* lower_vol = first ? nullptr : h->levels.back().volumes.data();
    // clusters are independent: each thread owns whole clusters
    #pragma omp parallel for schedule(dynamic, 64)
    for (uint32_t c = 0; c < k; ++c) {
        uint32_t vol = 0, cut = 0;
        for (uint32_t i = member_off[c]; i < member_off[c + 1]; ++i) {
            uint32_t v = members[i];
            vol += lower_vol ? lower_vol[v] : row[v + 1] - row[v];
            for (uint32_t a = row[v]; a < row[v + 1]; ++a)
                if (L.parents[nbr[a]] != c) cut++;
        }
        L.volumes[c] = vol;
        L.cut_sizes[c] = cut;
    }
    for (uint32_t c = 0; c < k; ++c) L.ranges[c + 1] = L.ranges[c] + L.cut_sizes[c];
    L.edges.resize(L.ranges[k]);
    #pragma omp parallel for schedule(dynamic, 64)
    for (uint32_t c = 0; c < k; ++c) {
        uint32_t o = L.ranges[c];
        for (uint32_t i = member_off[c]; i < member_off[c + 1]; ++i) {
            uint32_t v = members[i];
            for (uint32_t a = row[v]; a < row[v + 1]; ++a) {
                uint32_t pc = L.parents[nbr[a]];
                if (pc != c) L.edges[o++] = pc;
            }
        }
        // rows ascending, as the engine's CSR input convention has them; the
        // reference lists the same arcs in member order
        std::sort(L.edges.begin() + L.ranges[c], L.edges.begin() + L.ranges[c + 1]);
    }
    h->levels.push_back(std::move(L));
    return 0;
}

// One GPU decomposition of the top graph on the hierarchy's engine: dense
// cluster ids.
int decompose(const pxh_hierarchy* h, float sc, float rw, std::vector<int32_t>& ids, int32_t& k) {
    const uint32_t n = h->top_nodes();
    const std::vector<uint32_t>& row = h->top_ranges();
    pxd_engine* e = h->engine;
    int rc = 0;
    int32_t it = 0;
    ids.assign(n, 0);
    if ((rc = pxd_set_thresholds(e, sc, rw)) ||
        (rc = pxd_initialize(e, n, row[n] / 2, row.data(), h->top_edges().data())) ||
        (rc = pxd_expander_decomposition(e, &it)) || (rc = pxd_get_final_partition(e, ids.data(), &k))) {
        g_err = std::string("engine: ") + pxd_last_error();
        return rc;
    }
    return 0;
}

float as_frac(uint32_t v) { return static_cast<float>(v); }

// normalized_cut.h:12-145 on a copy of the hierarchy's volumes and cut sizes.
class NormalizedCutRun {
    struct TreeNode {
        uint32_t level, nix;
        bool operator<(const TreeNode& o) const { return level != o.level ? level < o.level : nix < o.nix; }
    };
    struct Candidate {
        TreeNode tn;
        float penalty;
    };

    const pxh_hierarchy* h;
    std::vector<std::vector<uint32_t>> vol, cut;
    std::vector<int32_t> labels;
    std::set<TreeNode> roots;                   // iterated in (level, nix) order
    std::vector<std::vector<uint8_t>> is_root;  // the same set, for O(1) membership
    bool write_labels = false;

    void add_root(const TreeNode& tn) {
        roots.insert(tn);
        is_root[tn.level][tn.nix] = 1;
    }

    // normalized_cut.h:43-51
    float edge_penalty(const TreeNode& tn, uint32_t cluster_cut, uint32_t cluster_vol) const {
        const uint32_t cut_a = cut[tn.level][tn.nix];
        const uint32_t vol_a = vol[tn.level][tn.nix];
        const uint32_t cut_b = cluster_cut + cut_a;
        const uint32_t vol_b = cluster_vol - vol_a;
        return as_frac(cut_a) / as_frac(vol_a) + as_frac(cut_b) / as_frac(vol_b);
    }

    // normalized_cut.h:56-82: children in member order; ties keep the first.
    // A level-0 cluster has no tree-node children: its call only labels the
    // input nodes.  Every round labels every input node again (each lies under
    // exactly one nearest root), so only the last round's labels are kept and
    // earlier rounds skip those calls.
    Candidate candidate(const TreeNode& tn, int32_t label, uint32_t cluster_cut, uint32_t cluster_vol) {
        Candidate best{TreeNode{0, 0}, 100000.0f};
        const Level& L = h->levels[tn.level];
        if (tn.level == 0) {
            if (write_labels)
                for (uint32_t i = L.member_off[tn.nix]; i < L.member_off[tn.nix + 1]; ++i) labels[L.members[i]] = label;
            return best;
        }
        for (uint32_t i = L.member_off[tn.nix]; i < L.member_off[tn.nix + 1]; ++i) {
            const uint32_t v = L.members[i];
            const TreeNode child{tn.level - 1, v};
            if (is_root[child.level][child.nix]) continue;
            if (vol[child.level][child.nix] < cluster_vol) {
                float pen = edge_penalty(child, cluster_cut, cluster_vol);
                if (pen < best.penalty) best = {child, pen};
            }
            Candidate cd = candidate(child, label, cluster_cut, cluster_vol);
            if (cd.penalty < best.penalty) best = cd;
        }
        return best;
    }

    // candidate() of a root with its children's subtrees evaluated in
    // parallel.  The serial loop keeps the first candidate of least penalty in
    // child order (strict <), a fold that the in-order reduction of the
    // per-child results reproduces; the final round's label writes are
    // disjoint between subtrees.
    Candidate root_candidate(const TreeNode& tn, int32_t label, uint32_t cluster_cut, uint32_t cluster_vol) {
        if (tn.level == 0) return candidate(tn, label, cluster_cut, cluster_vol);
        const Level& L = h->levels[tn.level];
        const uint32_t b = L.member_off[tn.nix], e = L.member_off[tn.nix + 1];
        std::vector<Candidate> part(e - b);
        #pragma omp parallel for schedule(dynamic, 16)
        for (uint32_t i = b; i < e; ++i) {
            const TreeNode child{tn.level - 1, L.members[i]};
            Candidate best{TreeNode{0, 0}, 100000.0f};
            if (!is_root[child.level][child.nix]) {
                if (vol[child.level][child.nix] < cluster_vol) {
                    float pen = edge_penalty(child, cluster_cut, cluster_vol);
                    if (pen < best.penalty) best = {child, pen};
                }
                Candidate cd = candidate(child, label, cluster_cut, cluster_vol);
                if (cd.penalty < best.penalty) best = cd;
            }
            part[i - b] = best;
        }
        Candidate best{TreeNode{0, 0}, 100000.0f};
        for (const Candidate& c : part)
            if (c.penalty < best.penalty) best = c;
        return best;
    }

    // normalized_cut.h:84-99: the split-off volume leaves every ancestor, its
    // cut joins every ancestor's
    void split(const TreeNode& tn) {
        const uint32_t vol_a = vol[tn.level][tn.nix];
        const uint32_t cut_a = cut[tn.level][tn.nix];
        uint32_t child = tn.nix;
        for (uint32_t l = tn.level + 1; l < h->levels.size(); ++l) {
            uint32_t parent = h->levels[l].parents[child];
            vol[l][parent] -= vol_a;
            cut[l][parent] += cut_a;
            child = parent;
        }
    }

public:
    explicit NormalizedCutRun(const pxh_hierarchy* hh) : h(hh), labels(hh->n, 0) {
        for (const Level& L : h->levels) {
            vol.push_back(L.volumes);
            cut.push_back(L.cut_sizes);
            is_root.emplace_back(L.nodes, 0);
        }
    }

    // normalized_cut.h:108-137
    const std::vector<int32_t>& compute(uint32_t k) {
        add_root(TreeNode{static_cast<uint32_t>(h->levels.size() - 1), 0});
        for (uint32_t i = 1; i <= k; ++i) {
            write_labels = i == k;
            Candidate best{TreeNode{0, 0}, 1000.0f};
            int32_t root_ix = 0;
            float best_delta = 100000.0f;
            for (const TreeNode& root : roots) {
                const uint32_t root_vol = vol[root.level][root.nix];
                const uint32_t root_cut = cut[root.level][root.nix];
                Candidate cd = root_candidate(root, root_ix, root_cut, root_vol);
                const float root_bias = as_frac(root_cut) / as_frac(root_vol);
                const float delta = cd.penalty - root_bias;
                if (delta < best_delta) {
                    best_delta = delta;
                    best = cd;
                }
                root_ix++;
            }
            split(best.tn);
            add_root(best.tn);
        }
        return labels;
    }
};

}  // namespace

extern "C" {

const char* pxh_last_error(void) { return g_err.c_str(); }

int pxh_create(uint32_t n, uint32_t m, const uint32_t* ranges, const uint32_t* edges, pxh_hierarchy** out) {
    if (!out || !ranges || (m && !edges)) return fail("pxh_create: null argument");
    if (ranges[0] != 0 || ranges[n] != 2ull * m) return fail("pxh_create: ranges must run from 0 to 2m");
    for (uint32_t v = 0; v < n; ++v)
        if (ranges[v] > ranges[v + 1]) return fail("pxh_create: ranges is not ascending");
    for (uint64_t a = 0; a < 2ull * m; ++a)
        if (edges[a] >= n) return fail("pxh_create: an arc ends past n");
    return guarded("pxh_create", [&] {
        auto h = std::make_unique<pxh_hierarchy>();
        h->n = n;
        h->ranges.assign(ranges, ranges + n + 1);
        h->edges.assign(edges, edges + 2ull * m);
        *out = h.release();
        return 0;
    });
}

int pxh_destroy(pxh_hierarchy* h) {
    delete h;
    return 0;
}

int pxh_add_level(pxh_hierarchy* h, uint32_t num_clusters, const uint32_t* member_off, const uint32_t* members) {
    if (!h || !member_off || (h->top_nodes() && !members)) return fail("pxh_add_level: null argument");
    return guarded("pxh_add_level", [&] { return contract(h, num_clusters, member_off, members); });
}

static int build_levels(pxh_hierarchy* h, const pxd_params* params, float target_sparsity, float rw_threshold,
                        uint32_t max_retries);

int pxh_build(pxh_hierarchy* h, const pxd_params* params, float target_sparsity, float rw_threshold,
              uint32_t max_retries) {
    if (!h) return fail("pxh_build: null hierarchy");
    return guarded("pxh_build", [&] { return build_levels(h, params, target_sparsity, rw_threshold, max_retries); });
}

static int build_levels(pxh_hierarchy* h, const pxd_params* params, float target_sparsity, float rw_threshold,
                        uint32_t max_retries) {
    pxd_params p;
    if (params) p = *params;
    else pxd_default_params(&p);
    // a later build with the same parameters keeps the engine (its device
    // buffers and kernel tuning): every level re-initializes it anyway
    if (h->engine && std::memcmp(&h->engine_params, &p, sizeof p) != 0) {
        pxd_destroy(h->engine);
        h->engine = nullptr;
    }
    if (!h->engine) {
        if (int rc = pxd_create(&p, &h->engine)) {
            h->engine = nullptr;
            g_err = std::string("pxd_create: ") + pxd_last_error();
            return rc;
        }
        h->engine_params = p;
    }
    float sparsity = target_sparsity;
    uint32_t retries = 0;
    std::vector<int32_t> ids;
    // expander_hierarchy.h:113-132 (do ... while top graph has more than one node)
    while (h->top_nodes() > 1 && h->top_ranges().back() > 0) {
        int32_t k = 0;
        if (int rc = decompose(h, sparsity, rw_threshold, ids, k)) return rc;
        const uint32_t n = h->top_nodes();
        const std::vector<uint32_t>& row = h->top_ranges();
        const std::vector<uint32_t>& nbr = h->top_edges();
        uint64_t cut_arcs = 0;
        #pragma omp parallel for schedule(static, 4096) reduction(+ : cut_arcs)
        for (uint32_t v = 0; v < n; ++v)
            for (uint32_t a = row[v]; a < row[v + 1]; ++a) cut_arcs += ids[v] != ids[nbr[a]];
        const uint32_t cut_edges = static_cast<uint32_t>(cut_arcs / 2), m = row[n] / 2;
        if (cut_edges > 0.95 * m) {
            if (++retries > max_retries) {
                char buf[160];
                snprintf(buf, sizeof buf, "pxh_build: %u sparsity reductions in a row still cut more than 95%% of %u edges",
                         max_retries, m);
                return fail(buf);
            }
            sparsity *= 0.8f;
            continue;
        }
        retries = 0;
        // clusters numbered by dense id, members ascending within each
        std::vector<uint32_t> off(static_cast<size_t>(k) + 1, 0), members(n);
        for (uint32_t v = 0; v < n; ++v) off[ids[v] + 1]++;
        for (int32_t c = 0; c < k; ++c) off[c + 1] += off[c];
        std::vector<uint32_t> fill(off.begin(), off.end() - 1);
        for (uint32_t v = 0; v < n; ++v) members[fill[ids[v]]++] = v;
        if (int rc = contract(h, static_cast<uint32_t>(k), off.data(), members.data())) return rc;
    }
    return 0;
}

int pxh_height(pxh_hierarchy* h, uint32_t* height) {
    if (!h || !height) return fail("pxh_height: null argument");
    *height = static_cast<uint32_t>(h->levels.size());
    return 0;
}

int pxh_level_size(pxh_hierarchy* h, uint32_t level, uint32_t* nodes, uint64_t* arcs, uint32_t* below_nodes) {
    if (!h || level >= h->levels.size()) return fail("pxh_level_size: no such level");
    const Level& L = h->levels[level];
    if (nodes) *nodes = L.nodes;
    if (arcs) *arcs = L.edges.size();
    if (below_nodes) *below_nodes = static_cast<uint32_t>(L.parents.size());
    return 0;
}

int pxh_level(pxh_hierarchy* h, uint32_t level, uint32_t* ranges, uint32_t* edges, uint32_t* volumes,
              uint32_t* cut_sizes, uint32_t* parents, uint32_t* member_off, uint32_t* members) {
    if (!h || level >= h->levels.size()) return fail("pxh_level: no such level");
    const Level& L = h->levels[level];
    auto put = [](uint32_t* dst, const std::vector<uint32_t>& src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(uint32_t));
    };
    put(ranges, L.ranges);
    put(edges, L.edges);
    put(volumes, L.volumes);
    put(cut_sizes, L.cut_sizes);
    put(parents, L.parents);
    put(member_off, L.member_off);
    put(members, L.members);
    return 0;
}

int pxh_subtract_volume(pxh_hierarchy* h, uint32_t level, uint32_t nix, uint32_t volume) {
    if (!h || level >= h->levels.size() || nix >= h->levels[level].nodes) return fail("pxh_subtract_volume: no such node");
    h->levels[level].volumes[nix] -= volume;
    return 0;
}

int pxh_add_cut(pxh_hierarchy* h, uint32_t level, uint32_t nix, uint32_t cut) {
    if (!h || level >= h->levels.size() || nix >= h->levels[level].nodes) return fail("pxh_add_cut: no such node");
    h->levels[level].cut_sizes[nix] += cut;
    return 0;
}

static int normalized_cut(pxh_hierarchy* h, uint32_t k, int32_t* labels, float* value);

int pxh_normalized_cut(pxh_hierarchy* h, uint32_t k, int32_t* labels, float* value) {
    if (!h || !labels) return fail("pxh_normalized_cut: null argument");
    if (h->levels.empty() || h->levels.back().nodes != 1)
        return fail("pxh_normalized_cut: the hierarchy needs a top level with exactly one node");
    if (k == 0) return fail("pxh_normalized_cut: k must be at least 1");
    return guarded("pxh_normalized_cut", [&] { return normalized_cut(h, k, labels, value); });
}

static int normalized_cut(pxh_hierarchy* h, uint32_t k, int32_t* labels, float* value) {
    NormalizedCutRun run(h);
    const std::vector<int32_t>& lab = run.compute(k);
    std::memcpy(labels, lab.data(), lab.size() * sizeof(int32_t));
    if (value) {
        // partition.h:418-427 over the parts in label order
        std::vector<uint32_t> pvol(k, 0), pint(k, 0);
        uint32_t* pv = pvol.data();
        uint32_t* pi = pint.data();
        // u32 sums wrap like the reference's and are order-free
        #pragma omp parallel for schedule(static, 4096) reduction(+ : pv[:k], pi[:k])
        for (uint32_t v = 0; v < h->n; ++v) {
            const uint32_t p = static_cast<uint32_t>(lab[v]);
            pv[p] += h->ranges[v + 1] - h->ranges[v];
            for (uint32_t a = h->ranges[v]; a < h->ranges[v + 1]; ++a) pi[p] += lab[h->edges[a]] == lab[v];
        }
        float nc = 0.0f;
        for (uint32_t p = 0; p < k; ++p)
            if (pvol[p]) nc += as_frac(pvol[p] - pint[p]) / as_frac(pvol[p]);
        *value = nc;
    }
    return 0;
}

}  // extern "C"

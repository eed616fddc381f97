// CPU oracle for graph search (LySearch) and the DPG layer builder
// (LyNNDescent::dpg / DelaunayGraph).
//
// TEST INFRASTRUCTURE ONLY, like oracle.cpp: restatements the tests check the
// HIP product against.  Paths are relative to the reference repository root.
//
// Deliberate, documented deviations:
//   * Ordering: the reference's priority queues compare distances only
//     (LySearch::CompareByFirst, src/lysearch.h:37-42), so equal distances
//     are ordered arbitrarily.  Here every comparison is on (dist, id), so
//     results are reproducible; the two agree whenever distances are
//     distinct.
//   * Entry points and seeds are arguments: the reference draws the entry
//     point with rand() (src/lysearch.cpp:378) and flat-search seeds from a
//     time(NULL)-seeded mt19937 (src/lysearch.cpp:455-458).
//   * Layers are CSR (offs[rows + 1], adj[]) instead of the reference's
//     "augmented" arrays (deg, ids...) + knnLocs (src/lysearch.cpp:112-186);
//     the content is the same.
//
// Pins (tests/test_oracle_search.py): the reference's own DPG code cannot be
// built here (it needs Boost), so orc_dpg_ref is pinned by closed forms:
// collinear rows give the path graph, a centre with axis rows gives a star
// whose centre keeps its maxM nearest, and a four-row case with ties,
// derived by hand, shows the in-place reverse pass dropping the one edge
// the order-free pass (orc_dpg_par, what the GPU builds) adds.  On sampled
// data the two passes share at least 99% of edges (measured 100%).

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

extern "C" float orc_l2_seq(const float* a, const float* b, uint32_t d);

namespace {

struct Cand {
    float d;
    uint32_t id;
};
inline bool less_c(const Cand& a, const Cand& b) { return a.d < b.d || (a.d == b.d && a.id < b.id); }

// nndHNSWSrch, src/lysearch.cpp:236-322 (and the seeded variant nndFlatSrch,
// src/lysearch.cpp:439-544): best-first search on one layer.  `top` is the
// result list of at most `beam` entries (the reference's topLst), `cand` the
// frontier (candis).  A node's distance is computed once (visited flags).
// start_d: the start's distance when the caller already has it (the layered
// search passes cur_dst, src/lysearch.cpp:415), else null (computed here).
void best_first(const float* X, uint32_t d, const uint64_t* offs, const uint32_t* adj, const float* q,
                const uint32_t* starts, const float* start_d, uint32_t nstart, uint32_t beam,
                std::vector<uint8_t>& flags, std::vector<uint32_t>& touched, std::vector<Cand>& out,
                uint64_t& ncmps) {
    std::vector<Cand> top, cand;  // top: max-heap on (d, id); cand: min-heap
    auto top_cmp = [](const Cand& a, const Cand& b) { return less_c(a, b); };
    auto cand_cmp = [](const Cand& a, const Cand& b) { return less_c(b, a); };
    for (uint32_t s = 0; s < nstart; ++s) {
        const uint32_t id = starts[s];
        if (flags[id]) continue;
        flags[id] = 1;
        touched.push_back(id);
        Cand c{start_d ? start_d[s] : orc_l2_seq(q, X + (uint64_t)id * d, d), id};
        if (!start_d) ++ncmps;
        cand.push_back(c);
        std::push_heap(cand.begin(), cand.end(), cand_cmp);
        top.push_back(c);
        std::push_heap(top.begin(), top.end(), top_cmp);
        if (top.size() > beam) {
            std::pop_heap(top.begin(), top.end(), top_cmp);
            top.pop_back();
        }
    }
    while (!cand.empty()) {
        const Cand c = cand.front();
        if (top.size() >= beam && less_c(top.front(), c)) break;  // src/lysearch.cpp:264-267
        std::pop_heap(cand.begin(), cand.end(), cand_cmp);
        cand.pop_back();
        for (uint64_t e = offs[c.id]; e < offs[c.id + 1]; ++e) {
            const uint32_t nb = adj[e];
            if (flags[nb]) continue;
            flags[nb] = 1;
            touched.push_back(nb);
            const Cand x{orc_l2_seq(q, X + (uint64_t)nb * d, d), nb};
            ++ncmps;
            if (top.size() < beam || less_c(x, top.front())) {  // src/lysearch.cpp:287-298
                cand.push_back(x);
                std::push_heap(cand.begin(), cand.end(), cand_cmp);
                top.push_back(x);
                std::push_heap(top.begin(), top.end(), top_cmp);
                if (top.size() > beam) {
                    std::pop_heap(top.begin(), top.end(), top_cmp);
                    top.pop_back();
                }
            }
        }
    }
    out = top;
    std::sort(out.begin(), out.end(), less_c);
}

}  // namespace

extern "C" {

// LySearch::knnSearch, src/lysearch.cpp:363-437, for each query: from
// `entry`, greedy descent through layers 0 .. nlayers-2 (move to the closest
// neighbour while it improves, src/lysearch.cpp:392-412), then best-first on
// the last layer with a result list of `beam` entries.  Writes the first k
// results (sorted by (dist, id)) and the number of distances computed.
void orc_search_layered(const float* X, uint32_t n, uint32_t d, uint32_t nlayers, const uint64_t* const* offs,
                        const uint32_t* const* adj, const float* Q, uint32_t nq, uint32_t entry, uint32_t beam,
                        uint32_t k, uint32_t* ids, float* dist, uint64_t* ncmps_out) {
    std::vector<uint8_t> flags(n, 0);
    std::vector<uint32_t> touched;
    std::vector<Cand> res;
    uint64_t ncmps = 0;
    for (uint32_t qi = 0; qi < nq; ++qi) {
        const float* q = Q + (uint64_t)qi * d;
        uint32_t cur = entry;
        float cd = orc_l2_seq(q, X + (uint64_t)cur * d, d);
        ++ncmps;
        for (uint32_t g = 0; g + 1 < nlayers; ++g) {
            bool moved = true;
            while (moved) {
                moved = false;
                const uint64_t b = offs[g][cur], e = offs[g][cur + 1];
                for (uint64_t j = b; j < e; ++j) {
                    const uint32_t nb = adj[g][j];
                    const float dd = orc_l2_seq(q, X + (uint64_t)nb * d, d);
                    ++ncmps;
                    if (dd < cd || (dd == cd && nb < cur)) {
                        cd = dd;
                        cur = nb;
                        moved = true;
                    }
                }
            }
        }
        touched.clear();
        best_first(X, d, offs[nlayers - 1], adj[nlayers - 1], q, &cur, &cd, 1, beam, flags, touched, res, ncmps);
        for (uint32_t t : touched) flags[t] = 0;
        for (uint32_t j = 0; j < k; ++j) {
            ids[(uint64_t)qi * k + j] = j < res.size() ? res[j].id : 0xFFFFFFFFu;
            dist[(uint64_t)qi * k + j] = j < res.size() ? res[j].d : 3.402823466e38f;
        }
    }
    if (ncmps_out) *ncmps_out = ncmps;
}

// LySearch::nndFlatSrch, src/lysearch.cpp:439-544: the result list starts
// from `beam` given seeds per query (seeds: nq x beam), then best-first.
void orc_search_flat(const float* X, uint32_t n, uint32_t d, const uint64_t* offs, const uint32_t* adj,
                     const float* Q, uint32_t nq, const uint32_t* seeds, uint32_t beam, uint32_t k, uint32_t* ids,
                     float* dist, uint64_t* ncmps_out) {
    std::vector<uint8_t> flags(n, 0);
    std::vector<uint32_t> touched;
    std::vector<Cand> res;
    uint64_t ncmps = 0;
    for (uint32_t qi = 0; qi < nq; ++qi) {
        const float* q = Q + (uint64_t)qi * d;
        touched.clear();
        best_first(X, d, offs, adj, q, seeds + (uint64_t)qi * beam, nullptr, beam, beam, flags, touched, res,
                   ncmps);
        for (uint32_t t : touched) flags[t] = 0;
        for (uint32_t j = 0; j < k; ++j) {
            ids[(uint64_t)qi * k + j] = j < res.size() ? res[j].id : 0xFFFFFFFFu;
            dist[(uint64_t)qi * k + j] = j < res.size() ? res[j].d : 3.402823466e38f;
        }
    }
    if (ncmps_out) *ncmps_out = ncmps;
}

}  // extern "C"

// ------------------------------------------------------------------ DPG

namespace {

struct Nb {
    uint32_t id;
    float d;
};

// LyNNDescent::insert2topk, src/lynndescent.cpp:548-623: bounded sorted
// insert; an entry equal in distance and id to one already present is
// rejected; with a full list, an entry equal in distance to the last one
// replaces it (the reference's shift loop stops at maxSize - 1).
bool insert2topk(std::vector<Nb>& L, uint32_t nid, float dst, uint32_t maxSize) {
    const size_t topk = L.size();
    if (topk == 0) {
        L.push_back({nid, dst});
        return true;
    }
    if (L[topk - 1].d < dst && topk == maxSize) return false;
    if (topk < maxSize && L[topk - 1].d < dst) {
        L.push_back({nid, dst});
        return true;
    }
    size_t i = topk;
    while (i > 0 && L[i - 1].d > dst) --i;
    for (size_t l = i; l > 0; --l) {
        if (L[l - 1].d < dst) break;
        if (L[l - 1].id == nid) return false;
    }
    size_t j;
    if (topk < maxSize) {
        L.resize(topk + 1);
        j = topk;
    } else {
        j = maxSize - 1;
    }
    while (j > i) {
        L[j] = L[j - 1];
        --j;
    }
    L[j] = {nid, dst};
    return true;
}

// Phase 1 of DelaunayGraph (src/lynndescent.cpp:648-676) for one row: walk
// the k-NN list in order and keep a neighbour unless an entry already kept
// is closer to it than the row is.
void occlusion_row(const float* X, uint32_t d, uint32_t i, const uint32_t* knn, uint32_t K, uint32_t M_,
                   std::vector<Nb>& L) {
    for (uint32_t j = 0; j < K; ++j) {
        if (L.size() > M_) break;
        const uint32_t nb = knn[(uint64_t)i * K + j];
        const float dst = orc_l2_seq(X + (uint64_t)i * d, X + (uint64_t)nb * d, d);
        bool good = true;
        for (const Nb& e : L)
            if (orc_l2_seq(X + (uint64_t)e.id * d, X + (uint64_t)nb * d, d) < dst) good = false;
        if (good) insert2topk(L, nb, dst, M_);
    }
}

// The reverse-edge rule of src/lynndescent.cpp:685-705: add source s (at
// distance dst) to target list T if T has room, else only if no entry of T
// is closer to s than dst.
void reverse_insert(const float* X, uint32_t d, std::vector<Nb>& T, uint32_t s, float dst, uint32_t maxM) {
    if (T.size() < maxM) {
        insert2topk(T, s, dst, maxM);
        return;
    }
    for (const Nb& e : T)
        if (orc_l2_seq(X + (uint64_t)e.id * d, X + (uint64_t)s * d, d) < dst) return;
    insert2topk(T, s, dst, maxM);
}

void to_csr(const std::vector<std::vector<Nb>>& G, uint64_t* offs, uint32_t* adj, float* adjd) {
    offs[0] = 0;
    for (size_t i = 0; i < G.size(); ++i) {
        offs[i + 1] = offs[i] + G[i].size();
        for (size_t j = 0; j < G[i].size(); ++j) {
            adj[offs[i] + j] = G[i][j].id;
            if (adjd) adjd[offs[i] + j] = G[i][j].d;
        }
    }
}

void from_csr(std::vector<std::vector<Nb>>& G, uint32_t rows, const uint64_t* offs, const uint32_t* adj,
              const float* adjd) {
    for (uint32_t i = 0; i < rows; ++i)
        for (uint64_t e = offs[i]; e < offs[i + 1]; ++e) G[i].push_back({adj[e], adjd[e]});
}

}  // namespace

extern "C" {

// LyNNDescent::DelaunayGraph, src/lynndescent.cpp:640-710, exactly as the
// reference runs it: phase 1 per row in order, then the reverse pass in row
// order, where each row's current list (grown by earlier reverse inserts)
// is walked.  prev_* (may be null): lists the rows start from, the
// reference's LyNNDescent::graph member carried over from the previous layer
// (it is never cleared between dpg calls).  Output CSR: offs[n + 1], adj and
// adjd with room for n * max(M_, maxM) entries.
void orc_dpg_ref(const float* X, uint32_t n, uint32_t d, const uint32_t* knn, uint32_t K, uint32_t M_,
                 uint32_t maxM, uint32_t prev_rows, const uint64_t* prev_offs, const uint32_t* prev_adj,
                 const float* prev_adjd, uint64_t* offs, uint32_t* adj, float* adjd) {
    std::vector<std::vector<Nb>> G(n);
    if (prev_offs) from_csr(G, prev_rows, prev_offs, prev_adj, prev_adjd);
    for (uint32_t i = 0; i < n; ++i) occlusion_row(X, d, i, knn, K, M_, G[i]);
    for (uint32_t i = 0; i < n; ++i) {
        const size_t num = G[i].size();
        for (size_t j = 0; j < num; ++j) {
            const Nb nn = G[i][j];
            reverse_insert(X, d, G[nn.id], i, nn.d, maxM);
        }
    }
    to_csr(G, offs, adj, adjd);
}

// The order-free restatement the GPU builds: phase 1 as above; the reverse
// pass reads the phase-1 lists only, and each target takes its reverse
// candidates in increasing source order.  It differs from orc_dpg_ref only
// where a row's phase-1 list gained reverse entries before its own turn in
// the reference's pass.
void orc_dpg_par(const float* X, uint32_t n, uint32_t d, const uint32_t* knn, uint32_t K, uint32_t M_,
                 uint32_t maxM, uint32_t prev_rows, const uint64_t* prev_offs, const uint32_t* prev_adj,
                 const float* prev_adjd, uint64_t* offs, uint32_t* adj, float* adjd) {
    std::vector<std::vector<Nb>> G(n);
    if (prev_offs) from_csr(G, prev_rows, prev_offs, prev_adj, prev_adjd);
    for (uint32_t i = 0; i < n; ++i) occlusion_row(X, d, i, knn, K, M_, G[i]);
    const std::vector<std::vector<Nb>> P1 = G;
    std::vector<std::vector<std::pair<uint32_t, float>>> rev(n);
    for (uint32_t i = 0; i < n; ++i)
        for (const Nb& e : P1[i]) rev[e.id].push_back({i, e.d});
    for (uint32_t u = 0; u < n; ++u)
        for (const auto& sv : rev[u]) reverse_insert(X, d, G[u], sv.first, sv.second, maxM);
    to_csr(G, offs, adj, adjd);
}

}  // extern "C"

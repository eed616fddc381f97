// Host classes over the C ABI; see nnmerge_host.hpp for the reference map.
#include "nnmerge_host.hpp"

#include <cmath>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <map>

namespace nnmerge {

float l2norm(const float* v1, const unsigned long s1, const float* v2, const unsigned long s2, const unsigned int d0) {
    // VMath::l2norm, src/vmath.cpp:56-88: one accumulator, elements in order
    const float* a = v1 + s1 * d0;
    const float* b = v2 + s2 * d0;
    float res = 0.f;
    for (unsigned i = 0; i < d0; ++i) {
        float t = a[i] - b[i];
        res += t * t;
    }
    return res;
}

float l1norm(const float* v1, const unsigned long s1, const float* v2, const unsigned long s2, const unsigned int d0) {
    // DistOpt::l1norm, src/distopt.cpp:82-110
    const float* a = v1 + s1 * d0;
    const float* b = v2 + s2 * d0;
    float res = 0.f;
    for (unsigned i = 0; i < d0; ++i) res += std::fabs(a[i] - b[i]);
    return res;
}

float ksquare(const float* v1, const unsigned long s1, const float* v2, const unsigned long s2, const unsigned int d0) {
    // DistOpt::KSquare, src/distopt.cpp:163-184
    const float* a = v1 + s1 * d0;
    const float* b = v2 + s2 * d0;
    float w2 = 0.f;
    for (unsigned i = 0; i < d0; ++i) {
        const float w1 = a[i] - b[i];
        if (a[i] != 0.f || b[i] != 0.f) w2 += w1 * w1 / (a[i] + b[i]);
    }
    return w2;
}

float cosine(const float* v1, const unsigned long s1, const float* v2, const unsigned long s2, const unsigned int d0) {
    // DistOpt::cosine, src/distopt.cpp:113-160
    const float* a = v1 + s1 * d0;
    const float* b = v2 + s2 * d0;
    float res = 0.f, l1 = 0.f, l2 = 0.f;
    for (unsigned i = 0; i < d0; ++i) {
        l1 += a[i] * a[i];
        l2 += b[i] * b[i];
        res += a[i] * b[i];
    }
    res = (l1 != 0.f && l2 != 0.f) ? res / (std::sqrt(l1) * std::sqrt(l2)) : 0.f;
    return 2.f - 2.f * res;
}

AbstractGraph::AbstractGraph(SpaceInterface<float>* space, int device) {
    nnm_default_params(&params);
    int rc = nnm_create(&ctx_, device);
    if (rc != NNM_OK) throw Error(rc, "nnm_create failed");
    if (space) {
        rc = nnm_set_metric(ctx_, space->metric());
        if (rc != NNM_OK) {
            const std::string msg = nnm_last_error(ctx_);
            nnm_destroy(ctx_);
            ctx_ = nullptr;
            throw Error(rc, msg);
        }
    }
}

AbstractGraph::~AbstractGraph() {
    if (ctx_) nnm_destroy(ctx_);
}

void AbstractGraph::check(int rc) const {
    if (rc != NNM_OK) throw Error(rc, nnm_last_error(ctx_));
}

void AbstractGraph::fetch() {
    ids_.resize((size_t)row * K0);
    dist_.resize((size_t)row * K0);
    flag_.resize((size_t)row * K0);
    check(nnm_get_graph(ctx_, ids_.data(), dist_.data(), flag_.data()));
}

unsigned* AbstractGraph::exportIds() const {
    // caller-owned copy, as KGraph::buildKGraph returns (src/kgraph.cpp:228-238)
    unsigned* out = new unsigned[(size_t)row * K0];
    std::copy(ids_.begin(), ids_.end(), out);
    return out;
}

std::vector<Neighbor> AbstractGraph::getNeighbors(unsigned i) const {
    std::vector<Neighbor> v;
    if (i >= row) return v;
    v.reserve(K0);
    for (unsigned j = 0; j < K0; ++j) {
        const size_t k = (size_t)i * K0 + j;
        v.push_back(Neighbor{ids_[k], dist_[k], 0u, flag_[k] != 0});
    }
    return v;
}

void AbstractGraph::saveKGraph(const char* dst) {
    // text format of AbstractMerge::saveKGraph, src/abstractmerge.cpp:278-291
    std::ofstream out(dst);
    if (!out) throw Error(NNM_ERR_INVALID, std::string("cannot open ") + dst);
    out << row << " " << K0 << "\n";
    for (unsigned i = 0; i < row; ++i) {
        out << i << " " << K0 << " ";
        for (unsigned j = 0; j < K0; ++j) out << ids_[(size_t)i * K0 + j] << " ";
        out << "\n";
    }
}

unsigned* KGraph::buildKGraph(float* mat, const unsigned K, const unsigned N, const unsigned D) {
    K0 = K;
    row = N;
    col = D;
    params.K = K;
    params.rule = NNM_RULE_ALL;
    check(nnm_set_data(ctx_, mat, N, D));
    nnm_stats s0{}, s1{};
    check(nnm_graph_random_init(ctx_, K, params.seed, &s0));
    check(nnm_nndescent(ctx_, &params, &s1));
    lastStats = s1;
    nCmps = s0.init_dists + s1.dists;
    buildSecs = s0.secs + s1.secs;
    mergeSecs = 0.0;
    fetch();
    return exportIds();
}

unsigned AbstractMerge::nnDescent() {
    nnm_params p = params;
    p.rule = NNM_RULE_MERGE;
    nnm_stats st{};
    check(nnm_nndescent(ctx_, &p, &st));
    lastStats = st;
    nCmps += st.dists;
    mergeSecs += st.secs;
    return 0;
}

unsigned AbstractMerge::localJoin() {
    nnm_params p = params;
    p.rule = NNM_RULE_MERGE;
    nnm_stats st{};
    check(nnm_local_join(ctx_, &p, iter_++, &st));
    lastStats = st;
    nCmps += st.dists;
    mergeSecs += st.secs;
    return 0;
}

unsigned AbstractMerge::merge2one() {
    check(nnm_merge_finish(ctx_, 1));
    return 1;
}

void AbstractMerge::mergePartitions(float* mat, unsigned K, unsigned N, unsigned D,
                                    const std::vector<uint32_t>& begin, const std::vector<uint8_t>& built) {
    K0 = K;
    row = N;
    col = D;
    params.K = K;
    const uint32_t P = (uint32_t)built.size();
    check(nnm_set_data(ctx_, mat, N, D));
    // 1. NN-Descent over every built partition at once: each partition its
    //    own group (KGraph::buildKGraph per subgraph, src/smerge.cpp:59-66)
    std::vector<uint32_t> grp(P);
    for (uint32_t s = 0; s < P; ++s) grp[s] = s;
    std::vector<uint8_t> all(P, 1);
    check(nnm_set_partition(ctx_, begin.data(), all.data(), grp.data(), P));
    nnm_stats s0{}, s1{};
    check(nnm_graph_random_init(ctx_, K, params.seed ^ 0xB01DULL, &s0));
    nnm_params bp = params;
    bp.rule = NNM_RULE_ALL;
    bp.max_iters = 0;  // subgraphs are built to convergence, as KGraph does
    bp.rho = 1.0f;
    check(nnm_nndescent(ctx_, &bp, &s1));
    buildSecs = s0.secs + s1.secs;
    // 2. the timed merge (src/smerge.cpp:75-81, src/jmerge.cpp:63-70)
    std::fill(grp.begin(), grp.end(), 0u);
    check(nnm_set_partition(ctx_, begin.data(), built.data(), grp.data(), P));
    nnm_stats r{}, i{};
    check(nnm_graph_save(ctx_));
    check(nnm_graph_restore(ctx_, 1, &r));  // InitSGraph: recompute distances
    nnm_params mp = params;
    mp.rule = NNM_RULE_MERGE;
    check(nnm_merge_init(ctx_, &mp, &i));
    nCmps = i.init_dists;
    mergeSecs = r.secs + i.secs;
    nnDescent();
    merge2one();
    fetch();
}

unsigned* Smerge::buildKGraph(float* mat, const unsigned K, const unsigned N, const unsigned D) {
    const uint32_t pre = N * 5 / 10;  // preSize, src/smerge.cpp:53
    mergePartitions(mat, K, N, D, {0u, pre, N}, {1, 1});
    return exportIds();
}

unsigned* Jmerge::buildKGraph(float* mat, const unsigned K, const unsigned N, const unsigned D) {
    const uint32_t pre = N * 5 / 10;  // preSize, src/jmerge.cpp:47
    mergePartitions(mat, K, N, D, {0u, pre, N}, {1, 0});
    return exportIds();
}

unsigned* PMerge::buildKGraph(float* mat, const unsigned K, const unsigned N, const unsigned D) {
    if (parts_ < 2 || parts_ > N / (K + 1)) throw Error(NNM_ERR_INVALID, "PMerge: bad partition count");
    std::vector<uint32_t> b(parts_ + 1);
    for (unsigned s = 0; s <= parts_; ++s) b[s] = (uint32_t)((uint64_t)N * s / parts_);
    mergePartitions(mat, K, N, D, b, std::vector<uint8_t>(parts_, 1));
    return exportIds();
}

void LyNNDescent::runLevel(const std::vector<uint32_t>& begin, const std::vector<uint8_t>& built,
                           const std::vector<uint32_t>& group, unsigned cut) {
    check(nnm_set_partition(ctx_, begin.data(), built.data(), group.data(), (uint32_t)built.size()));
    nnm_params mp = params;
    mp.rule = NNM_RULE_MERGE;
    mp.cut = cut;
    nnm_stats i{}, st{};
    check(nnm_merge_init(ctx_, &mp, &i));
    check(nnm_nndescent(ctx_, &mp, &st));
    check(nnm_merge_finish(ctx_, 1));
    LevelStat ls{begin.back(), group.empty() ? 0u : group.back() + 1, st.iters, i.init_dists + st.dists,
                 i.secs + st.secs};
    levels.push_back(ls);
    nCmps += ls.dists;
    mergeSecs += ls.secs;
    lastStats = st;
}

// LyNNDescent::dpg (src/lynndescent.cpp:502-540) on the current lists:
// occlusion bound K0/2, reverse bound maxM, lists carried from the previous
// layer; the layer is kept and, with a file, appended as saveDPGraph writes it.
void LyNNDescent::emitLayer(unsigned maxM, bool carry, const char* layerFile) {
    nnm_stats st{};
    check(nnm_dpg_build(ctx_, K0 / 2, maxM, carry ? 1 : 0, &st));
    dpgSecs += st.secs;
    DPGLayer L;
    L.rows = levels.empty() ? 0u : levels.back().rows;
    L.offs.resize((size_t)L.rows + 1);
    check(nnm_dpg_get(ctx_, L.offs.data(), nullptr, nullptr));
    L.adj.resize(L.offs.back());
    check(nnm_dpg_get(ctx_, L.offs.data(), L.adj.data(), nullptr));
    if (layerFile)
        check(nnm_io_write_layer(layerFile, layers.empty() ? 0 : 1, L.offs.data(), L.adj.data(), L.rows));
    layers.push_back(std::move(L));
}

unsigned LyNNDescent::buildhgraph(float* mat, unsigned N, unsigned D, unsigned topk, unsigned M,
                                  const char* dstFn) {
    K0 = topk;
    row = N;
    col = D;
    params.K = topk;
    levels.clear();
    layers.clear();
    nCmps = 0;
    mergeSecs = 0.0;
    dpgSecs = 0.0;
    if (N < N0 + 1) throw Error(NNM_ERR_INVALID, "buildhgraph needs more than N0 rows");
    const std::string layerPath = dstFn ? std::string(dstFn) + ".txt" : std::string();
    const char* layerFile = dstFn ? layerPath.c_str() : nullptr;
    check(nnm_set_data(ctx_, mat, N, D));
    // seed graph: plain NN-Descent over the first N0 rows (init + nnDescent(verse = 1))
    {
        uint32_t b[2] = {0u, N0};
        uint8_t bu[1] = {1};
        uint32_t g[1] = {0};
        check(nnm_set_partition(ctx_, b, bu, g, 1));
        nnm_stats s0{}, s1{};
        check(nnm_graph_random_init(ctx_, topk, params.seed, &s0));
        nnm_params bp = params;
        bp.rule = NNM_RULE_ALL;
        check(nnm_nndescent(ctx_, &bp, &s1));
        levels.push_back(LevelStat{N0, 1, s1.iters, s0.init_dists + s1.dists, s0.secs + s1.secs});
        nCmps += s0.init_dists + s1.dists;
        mergeSecs += s0.secs + s1.secs;
    }
    // the layer base M: _M on the base layer (all N rows), _M / 2 above
    auto layerM = [&](uint32_t rows) { return rows == N ? M : M / 2; };
    if (M) emitLayer(layerM(N0), false, layerFile);
    // appdNwSmpl(K/2, smplNum) + nnDescent(0) + merge2one, smplNum doubling;
    // setflag() relabels the whole graph as one built subgraph (the next
    // level's partition {built [0, cur), raw [cur, cur + b)})
    static const unsigned char level[10] = {1, 0, 0, 1, 0, 0, 1, 0, 0, 1};
    unsigned count_level = 1;
    uint32_t cur = N0, smpl = N0;
    while (cur < N) {
        const bool last = N - cur <= smpl;
        const uint32_t b = last ? (N - cur) : smpl;
        runLevel({0u, cur, cur + b}, {1, 0}, {0, 0}, topk - topk / 2);
        cur += b;
        smpl = cur;
        // the reference tests level[count_level] before count_level < 10
        // (reading past the table from level 10 on); the bound comes first here
        if (M && (last || (count_level < 10 && level[count_level]))) emitLayer(layerM(cur), true, layerFile);
        ++count_level;
    }
    fetch();
    if (dstFn) saveKGraph((std::string(dstFn) + "_baselayer.txt").c_str());
    return 0;
}

unsigned LyNNDescent::buildLeafHierarchy(float* mat, unsigned N, unsigned D, unsigned K,
                                         const std::vector<uint32_t>& leaf_begin) {
    K0 = K;
    row = N;
    col = D;
    params.K = K;
    levels.clear();
    nCmps = 0;
    mergeSecs = 0.0;
    if (leaf_begin.size() < 2 || leaf_begin.front() != 0 || leaf_begin.back() != N)
        throw Error(NNM_ERR_INVALID, "leaf_begin must run from 0 to N");
    check(nnm_set_data(ctx_, mat, N, D));
    std::vector<uint32_t> begin = leaf_begin;
    const uint32_t L = (uint32_t)begin.size() - 1;
    // every leaf built by NN-Descent in the same launches, each its own group
    {
        std::vector<uint8_t> bu(L, 1);
        std::vector<uint32_t> g(L);
        for (uint32_t s = 0; s < L; ++s) g[s] = s;
        check(nnm_set_partition(ctx_, begin.data(), bu.data(), g.data(), L));
        nnm_stats s0{}, s1{};
        check(nnm_graph_random_init(ctx_, K, params.seed ^ 0x1EAFull, &s0));
        nnm_params bp = params;
        bp.rule = NNM_RULE_ALL;
        check(nnm_nndescent(ctx_, &bp, &s1));
        buildSecs = s0.secs + s1.secs;
        levels.push_back(LevelStat{N, L, s1.iters, s0.init_dists + s1.dists, s0.secs + s1.secs});
    }
    // levels: adjacent pairs merged together (an odd last subgraph waits)
    while (begin.size() > 2) {
        const uint32_t S = (uint32_t)begin.size() - 1;
        std::vector<uint8_t> bu(S, 1);
        std::vector<uint32_t> g(S);
        for (uint32_t s = 0; s < S; ++s) g[s] = s / 2;
        runLevel(begin, bu, g, K / 2);
        std::vector<uint32_t> nb;
        for (uint32_t s = 0; s < S; s += 2) nb.push_back(begin[s]);
        nb.push_back(N);
        begin.swap(nb);
    }
    fetch();
    return 0;
}

unsigned* LyNNDescent::buildKGraph(float* mat, const unsigned K, const unsigned N, const unsigned D) {
    buildhgraph(mat, N, D, K);
    return exportIds();
}

const std::vector<unsigned> LySearch::Ks1 = {60, 50, 40, 36, 32, 28, 24, 20, 18, 15, 14, 12, 10, 8, 6, 4, 2};

void LySearch::check(int rc) const {
    if (rc != NNM_OK) throw Error(rc, ctx_ ? nnm_last_error(ctx_) : "nnmerge error");
}

void LySearch::init(SpaceInterface<float>* space, int device) {
    int rc = nnm_create(&ctx_, device);
    if (rc != NNM_OK) throw Error(rc, "nnm_create failed");
    if (space) check(nnm_set_metric(ctx_, space->metric()));
}

void LySearch::setLayers(const std::vector<DPGLayer>& layers) {
    if (layers.empty()) throw Error(NNM_ERR_INVALID, "LySearch: no layers");
    std::vector<uint32_t> rows;
    std::vector<const uint32_t*> offs, adj;
    static const uint32_t none = 0;
    for (const DPGLayer& L : layers) {
        if (L.offs.size() != (size_t)L.rows + 1) throw Error(NNM_ERR_INVALID, "LySearch: layer offsets");
        rows.push_back(L.rows);
        offs.push_back(L.offs.data());
        adj.push_back(L.adj.empty() ? &none : L.adj.data());
    }
    check(nnm_search_set_layers(ctx_, (uint32_t)layers.size(), rows.data(), offs.data(), adj.data()));
}

LySearch::LySearch(const char* Rsrc, const char* Gsrc, SpaceInterface<float>* space, int device) {
    init(space, device);
    try {
        uint64_t n = 0;
        uint32_t d = 0;
        check(nnm_io_matrix_shape(Rsrc, &n, &d));
        if (n == 0 || n > 0xFFFFFFF0ull) throw Error(NNM_ERR_INVALID, "LySearch: reference set size");
        std::vector<float> X(n * d);
        check(nnm_io_read_matrix(Rsrc, X.data(), n, d));
        check(nnm_set_data(ctx_, X.data(), n, d));
        row = (unsigned)n;
        col = d;
        uint32_t nl = 0, lr[8];
        uint64_t le[8];
        check(nnm_io_layers_shape(Gsrc, 8, &nl, lr, le));
        if (nl == 0) throw Error(NNM_ERR_INVALID, "LySearch: no layer in the graph file");
        std::vector<DPGLayer> layers(nl);
        std::vector<uint32_t*> po(nl), pa(nl);
        for (uint32_t g = 0; g < nl; ++g) {
            layers[g].rows = lr[g];
            layers[g].offs.resize((size_t)lr[g] + 1);
            layers[g].adj.resize(le[g]);
            po[g] = layers[g].offs.data();
            pa[g] = layers[g].adj.data();
        }
        check(nnm_io_read_layers(Gsrc, nl, po.data(), pa.data()));
        setLayers(layers);
    } catch (...) {
        nnm_destroy(ctx_);
        ctx_ = nullptr;
        throw;
    }
}

LySearch::LySearch(const float* mat, unsigned N, unsigned D, const std::vector<DPGLayer>& layers,
                   SpaceInterface<float>* space, int device) {
    init(space, device);
    try {
        check(nnm_set_data(ctx_, mat, N, D));
        row = N;
        col = D;
        setLayers(layers);
    } catch (...) {
        nnm_destroy(ctx_);
        ctx_ = nullptr;
        throw;
    }
}

LySearch::~LySearch() {
    if (ctx_) nnm_destroy(ctx_);
}

std::vector<SearchRecord> LySearch::knnSearch(const float* queries, unsigned nQry, const std::vector<unsigned>& beams,
                                              unsigned k, std::vector<uint32_t>* ids, std::vector<float>* dists) {
    if (beams.empty()) throw Error(NNM_ERR_INVALID, "LySearch::knnSearch: no beam");
    records.clear();
    stats.clear();
    std::vector<SearchRecord> out;
    // the launch takes at most 32 beams: longer lists run in chunks of 32
    std::vector<uint32_t> I;
    std::vector<float> Dd;
    if (ids) ids->assign((size_t)beams.size() * nQry * k, 0xFFFFFFFFu);
    if (dists) dists->assign((size_t)beams.size() * nQry * k, 0.f);
    nnm_search_params p{};
    p.mode = NNM_SEARCH_LAYERED;
    p.k = k;
    p.width = width;
    p.prec = prec;
    p.entry = entry;
    for (size_t b0 = 0; b0 < beams.size(); b0 += 32) {
        const size_t nb = std::min<size_t>(32, beams.size() - b0);
        std::vector<uint32_t> bl(beams.begin() + b0, beams.begin() + b0 + nb);
        I.resize(nb * nQry * k);
        Dd.resize(nb * nQry * k);
        std::vector<nnm_search_stats> st(nb);
        check(nnm_search_beams(ctx_, queries, nQry, &p, bl.data(), (uint32_t)nb, I.data(), Dd.data(), st.data()));
        if (ids) std::copy(I.begin(), I.end(), ids->begin() + b0 * nQry * k);
        if (dists) std::copy(Dd.begin(), Dd.end(), dists->begin() + b0 * nQry * k);
        for (size_t b = 0; b < nb; ++b) {
            stats.push_back(st[b]);
            out.push_back(SearchRecord{bl[b], 0.0, st[b].secs, nQry ? (unsigned long)(st[b].dists / nQry) : 0ul});
        }
    }
    records = out;
    return out;
}

int LySearch::knnSearch(const char* qryFn, const char* dstFn, const char* grdFn, const char* record) {
    uint64_t nq = 0;
    uint32_t qd = 0;
    check(nnm_io_matrix_shape(qryFn, &nq, &qd));
    if (qd != col) throw Error(NNM_ERR_INVALID, "LySearch::knnSearch: query dimension differs from the reference set");
    if (nq == 0 || nq > 0xFFFFFFF0ull) throw Error(NNM_ERR_INVALID, "LySearch::knnSearch: query count");
    std::vector<float> Q(nq * qd);
    check(nnm_io_read_matrix(qryFn, Q.data(), nq, qd));
    // ground truth: "query nn" pairs; the first pair of a query counts, and
    // every query named counts in the total (a std::map insert, as
    // Evaluator::load_NNmap does)
    std::map<uint64_t, uint64_t> nnmap;
    {
        std::ifstream g(grdFn);
        if (!g) throw Error(NNM_ERR_INVALID, "LySearch::knnSearch: cannot read the ground truth");
        uint64_t key = 0, val = 0;
        while (g >> key >> val) nnmap.insert(std::make_pair(key, val));
    }
    std::vector<uint32_t> ids;
    std::vector<float> dd;
    std::vector<SearchRecord> rec = knnSearch(Q.data(), (unsigned)nq, Ks1, 1, &ids, &dd);
    const std::vector<nnm_search_stats> all = stats;
    std::ofstream out(record);
    if (!out) throw Error(NNM_ERR_INVALID, "LySearch::knnSearch: cannot write the record");
    size_t last = 0;
    for (size_t b = 0; b < rec.size(); ++b) {
        uint64_t hits = 0;
        const uint32_t* r = ids.data() + b * nq;
        for (uint64_t q = 0; q < nq; ++q) {
            auto it = nnmap.find(q);
            if (it != nnmap.end() && it->second == r[q]) ++hits;
        }
        rec[b].recall = nnmap.empty() ? 0.0 : (double)hits / (double)nnmap.size();
        out << rec[b].topk << "\t" << rec[b].recall << "\t" << rec[b].secs << "\t" << rec[b].cmpsPerQuery << "\n";
        last = b;
        if (rec[b].recall == 1.0) break;
    }
    records.assign(rec.begin(), rec.begin() + last + 1);
    stats.assign(all.begin(), all.begin() + last + 1);
    // the result file after the sweep holds the last evaluated beam's top-1
    std::ofstream dst(dstFn);
    if (!dst) throw Error(NNM_ERR_INVALID, "LySearch::knnSearch: cannot write the results");
    for (uint64_t q = 0; q < nq; ++q) dst << q << "\t" << ids[last * nq + q] << "\t" << dd[last * nq + q] << "\n";
    return 0;
}

}  // namespace nnmerge

extern "C" int nnm_host_hmerge(const float* X, uint64_t n, uint32_t d, const nnm_params* p, uint32_t nleaves,
                               const uint32_t* leaf_begin, uint32_t* ids, float* dist, nnm_stats* st,
                               uint32_t* level_rows, uint32_t* level_groups, double* level_secs,
                               uint64_t* level_dists, uint32_t max_levels, uint32_t* nlevels) {
    if (!X || !p || !ids || n == 0 || n > 0xFFFFFFF0ull) return NNM_ERR_INVALID;
    try {
        nnmerge::LyNNDescent g;
        g.params = *p;
        if (nleaves && leaf_begin)
            g.buildLeafHierarchy(const_cast<float*>(X), (unsigned)n, d, p->K,
                                 std::vector<uint32_t>(leaf_begin, leaf_begin + nleaves + 1));
        else
            g.buildhgraph(const_cast<float*>(X), (unsigned)n, d, p->K);
        std::memcpy(ids, g.ids().data(), n * p->K * sizeof(uint32_t));
        if (dist) std::memcpy(dist, g.dists().data(), n * p->K * sizeof(float));
        if (st) {
            *st = g.lastStats;
            st->dists = g.nCmps;
            st->secs = g.mergeSecs;
        }
        const uint32_t L = (uint32_t)g.levels.size();
        if (nlevels) *nlevels = L;
        for (uint32_t i = 0; i < L && i < max_levels; ++i) {
            if (level_rows) level_rows[i] = g.levels[i].rows;
            if (level_groups) level_groups[i] = g.levels[i].groups;
            if (level_secs) level_secs[i] = g.levels[i].secs;
            if (level_dists) level_dists[i] = g.levels[i].dists;
        }
        return NNM_OK;
    } catch (const nnmerge::Error& e) {
        return e.code;
    } catch (...) {
        return NNM_ERR_HIP;
    }
}

extern "C" int nnm_host_hgraph(const float* X, uint64_t n, uint32_t d, const nnm_params* p, uint32_t M,
                               const char* dst_prefix, uint32_t* ids, float* dist, nnm_stats* st,
                               uint32_t* layer_rows, uint32_t max_layers, uint32_t* nlayers, double* dpg_secs) {
    if (!X || !p || !ids || n == 0 || n > 0xFFFFFFF0ull || M == 0) return NNM_ERR_INVALID;
    try {
        nnmerge::LyNNDescent g;
        g.params = *p;
        g.buildhgraph(const_cast<float*>(X), (unsigned)n, d, p->K, M, dst_prefix);
        std::memcpy(ids, g.ids().data(), n * p->K * sizeof(uint32_t));
        if (dist) std::memcpy(dist, g.dists().data(), n * p->K * sizeof(float));
        if (st) {
            *st = g.lastStats;
            st->dists = g.nCmps;
            st->secs = g.mergeSecs;
        }
        const uint32_t L = (uint32_t)g.layers.size();
        if (nlayers) *nlayers = L;
        for (uint32_t i = 0; i < L && i < max_layers; ++i)
            if (layer_rows) layer_rows[i] = g.layers[i].rows;
        if (dpg_secs) *dpg_secs = g.dpgSecs;
        return NNM_OK;
    } catch (const nnmerge::Error& e) {
        return e.code;
    } catch (...) {
        return NNM_ERR_HIP;
    }
}

extern "C" int nnm_host_lysearch(const char* refer, const char* graph, const char* query, const char* dst,
                                 const char* grd, const char* record, uint32_t entry, uint32_t prec,
                                 uint32_t max_records, uint32_t* nrecords, uint32_t* topk, double* recall, double* secs,
                                 uint64_t* cmps) {
    if (!refer || !graph || !query || !dst || !grd || !record) return NNM_ERR_INVALID;
    try {
        nnmerge::LySearch s(refer, graph);
        s.entry = entry;
        s.prec = prec;
        s.knnSearch(query, dst, grd, record);
        const uint32_t R = (uint32_t)s.records.size();
        if (nrecords) *nrecords = R;
        for (uint32_t i = 0; i < R && i < max_records; ++i) {
            if (topk) topk[i] = s.records[i].topk;
            if (recall) recall[i] = s.records[i].recall;
            if (secs) secs[i] = s.records[i].secs;
            if (cmps) cmps[i] = s.records[i].cmpsPerQuery;
        }
        return NNM_OK;
    } catch (const nnmerge::Error& e) {
        return e.code;
    } catch (...) {
        return NNM_ERR_HIP;
    }
}

extern "C" int nnm_host_build(int kind, const float* X, uint64_t n, uint32_t d, const nnm_params* p, uint32_t parts,
                              uint32_t* ids, float* dist, nnm_stats* st, double* build_secs) {
    if (!X || !p || !ids || n == 0 || n > 0xFFFFFFF0ull) return NNM_ERR_INVALID;
    try {
        nnmerge::AbstractGraph* g = nullptr;
        switch (kind) {
            case 0: g = new nnmerge::KGraph(); break;
            case 1: g = new nnmerge::Smerge(); break;
            case 2: g = new nnmerge::Jmerge(); break;
            case 3: g = new nnmerge::PMerge(parts); break;
            default: return NNM_ERR_INVALID;
        }
        g->params = *p;
        unsigned* out = g->buildKGraph(const_cast<float*>(X), p->K, (unsigned)n, d);
        std::memcpy(ids, out, n * p->K * sizeof(uint32_t));
        delete[] out;
        if (dist) std::memcpy(dist, g->dists().data(), n * p->K * sizeof(float));
        if (st) {
            *st = g->lastStats;
            st->dists = g->nCmps;
            st->secs = g->mergeSecs;
        }
        if (build_secs) *build_secs = g->buildSecs;
        delete g;
        return NNM_OK;
    } catch (const nnmerge::Error& e) {
        return e.code;
    } catch (...) {
        return NNM_ERR_HIP;
    }
}

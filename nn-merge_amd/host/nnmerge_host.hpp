// Host classes with the reference's entry points, running the local-join
// path on the GPU through the C ABI (include/nnmerge.h).
//
//   KGraph::buildKGraph(mat, K, N, D)      src/kgraph.cpp:217-245
//   Smerge::buildKGraph(mat, K, N, D)      src/smerge.cpp:46-91
//   Jmerge::buildKGraph(mat, K, N, D)      src/jmerge.cpp:43-80
//   AbstractMerge::nnDescent / localJoin   src/abstractmerge.cpp:78-192,
//                                          src/abstractmerge.h:29-30
//   AbstractMerge::merge2one               src/abstractmerge.cpp:232-277
//   saveKGraph (text "N K" / "i K ids")    src/abstractmerge.cpp:278-291
//   getNeighbors -> the Neighbor list view src/pvetex.h:18-30
//   LySearch(Rsrc, Gsrc, space)            src/lysearch.cpp:32-39
//   LySearch::knnSearch(qry, dst, grd, rec) src/lysearch.cpp:363-437
//
// PMerge generalises the two-way merges to P contiguous partitions merged
// jointly (pairs inside one built partition skipped), all partition builds
// in one batched launch.
//
// Ownership differs on purpose: the reference's graph objects delete the
// caller's matrix (src/abstractgraph.cpp:43-47, src/kgraph.cpp:28-32); these
// classes copy it to the device and never free caller memory.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nnmerge.h"

namespace nnmerge {

// DISTFUNC and SpaceInterface of src/abstractgraph.h:11-19, kept for CPU
// callers that pass a distance to the host classes.  A space also names its
// device metric (NNM_METRIC_*), which the host classes hand to the library;
// the reference's interface has only the function pointer.
template <typename MTYPE>
using DISTFUNC = MTYPE (*)(const float*, const unsigned long, const float*, const unsigned long, const unsigned int);

template <typename MTYPE>
class SpaceInterface {
public:
    virtual ~SpaceInterface() {}
    virtual DISTFUNC<MTYPE> get_dist_func() = 0;
    virtual unsigned metric() const { return NNM_METRIC_L2; }
};

// CPU distances of the spaces, in the reference's operation order.
float l2norm(const float* v1, const unsigned long s1, const float* v2, const unsigned long s2, const unsigned int d0);
float l1norm(const float* v1, const unsigned long s1, const float* v2, const unsigned long s2, const unsigned int d0);
float ksquare(const float* v1, const unsigned long s1, const float* v2, const unsigned long s2, const unsigned int d0);
float cosine(const float* v1, const unsigned long s1, const float* v2, const unsigned long s2, const unsigned int d0);

// L2space, src/space_lp.h:9-23.
class L2space : public SpaceInterface<float> {
public:
    DISTFUNC<float> get_dist_func() override { return l2norm; }
};

// L1space, src/space_lp.h:25-38.
class L1space : public SpaceInterface<float> {
public:
    DISTFUNC<float> get_dist_func() override { return l1norm; }
    unsigned metric() const override { return NNM_METRIC_L1; }
};

// Ksquare (chi-square), src/space_lp.h:40-51.
class Ksquare : public SpaceInterface<float> {
public:
    DISTFUNC<float> get_dist_func() override { return ksquare; }
    unsigned metric() const override { return NNM_METRIC_CHI2; }
};

// Cosine distance 2 - 2 cos (DistOpt::cosine, src/distopt.cpp:113-160; the
// reference defines the routine but no space class for it).
class CosineSpace : public SpaceInterface<float> {
public:
    DISTFUNC<float> get_dist_func() override { return cosine; }
    unsigned metric() const override { return NNM_METRIC_COSINE; }
};

// Host view of one list entry (mergeDataType::Neighbor, src/pvetex.h:18-30).
struct Neighbor {
    unsigned id;
    float dist;
    unsigned gd;
    bool flag;
};

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& m) : std::runtime_error(m), code(code) {}
    int code;
};

class AbstractGraph {
public:
    explicit AbstractGraph(SpaceInterface<float>* space = nullptr, int device = 0);
    virtual ~AbstractGraph();
    AbstractGraph(const AbstractGraph&) = delete;
    AbstractGraph& operator=(const AbstractGraph&) = delete;

    virtual unsigned* buildKGraph(float* mat, const unsigned K, const unsigned N, const unsigned D) = 0;
    virtual void saveKGraph(const char* dst);
    std::vector<Neighbor> getNeighbors(unsigned i) const;
    const std::vector<uint32_t>& ids() const { return ids_; }
    const std::vector<float>& dists() const { return dist_; }

    nnm_params params;          // K, rho, rev_cap, T, delta, prec, seed ...
    unsigned long nCmps = 0;    // distances of the timed region (reference nCmps)
    double mergeSecs = 0.0;     // device time of the timed region
    double buildSecs = 0.0;     // device time of the subgraph builds (not timed by the reference)
    nnm_stats lastStats{};

protected:
    void check(int rc) const;
    void fetch();
    unsigned* exportIds() const;

    nnm_ctx* ctx_ = nullptr;
    unsigned K0 = 0, row = 0, col = 0;
    std::vector<uint32_t> ids_;
    std::vector<float> dist_;
    std::vector<uint8_t> flag_;
};

class KGraph : public AbstractGraph {
public:
    using AbstractGraph::AbstractGraph;
    unsigned* buildKGraph(float* mat, const unsigned K, const unsigned N, const unsigned D) override;
};

class AbstractMerge : public AbstractGraph {
public:
    using AbstractGraph::AbstractGraph;
    // One NN-Descent loop over the current merge state (list build, local
    // join, delta test), as AbstractMerge::nnDescent.
    virtual unsigned nnDescent();
    // One iteration of the loop: the per-iteration local join.
    virtual unsigned localJoin();
    unsigned merge2one();

protected:
    // Builds every partition with NN-Descent in one launch, then runs the
    // timed merge: InitSGraph + InitialGraph + nnDescent + merge2one.
    void mergePartitions(float* mat, unsigned K, unsigned N, unsigned D, const std::vector<uint32_t>& begin,
                         const std::vector<uint8_t>& built);
    unsigned iter_ = 0;
};

class Smerge : public AbstractMerge {
public:
    using AbstractMerge::AbstractMerge;
    unsigned* buildKGraph(float* mat, const unsigned K, const unsigned N, const unsigned D) override;
};

class Jmerge : public AbstractMerge {
public:
    using AbstractMerge::AbstractMerge;
    unsigned* buildKGraph(float* mat, const unsigned K, const unsigned N, const unsigned D) override;
};

class PMerge : public AbstractMerge {
public:
    PMerge(unsigned parts, SpaceInterface<float>* space = nullptr, int device = 0)
        : AbstractMerge(space, device), parts_(parts) {}
    unsigned* buildKGraph(float* mat, const unsigned K, const unsigned N, const unsigned D) override;

private:
    unsigned parts_;
};

// LyNNDescent (src/lynndescent.h:19-86): hierarchical construction.
//   buildhgraph: the reference's doubling schedule (src/lynndescent.cpp:735-831):
//     NN-Descent over the first N0 = 64 rows, then repeatedly append a block
//     as large as the current graph (the rest at the end), J-Merge it in
//     (appdNwSmpl + nnDescent(verse = 0) + merge2one) until every row is in.
//   buildLeafHierarchy: uneven leaves built together in one batched launch,
//     then levels that merge adjacent pairs of subgraphs, every pair of a
//     level in the same launches (S-Merge rule: both sides built).
struct LevelStat {
    uint32_t rows;       // rows covered after the level
    uint32_t groups;     // merges run together
    uint32_t iters;
    uint64_t dists;
    double secs;         // device time of the level's merge
};

// One DPG layer as CSR over rows [0, rows).
struct DPGLayer {
    uint32_t rows;
    std::vector<uint32_t> offs;  // rows + 1
    std::vector<uint32_t> adj;
};

class LyNNDescent : public AbstractGraph {
public:
    static const unsigned N0 = 64;  // src/lynndescent.cpp:23
    using AbstractGraph::AbstractGraph;
    unsigned* buildKGraph(float* mat, const unsigned K, const unsigned N, const unsigned D) override;
    // buildhgraph, src/lynndescent.cpp:735-831.  M > 0 also emits the DPG
    // layers (LyNNDescent::dpg after the seed graph, after the levels the
    // level[] table marks and after the last level) into `layers`, and with
    // dstFn writes them to dstFn + ".txt" and the final k-NN graph to
    // dstFn + "_baselayer.txt".
    unsigned buildhgraph(float* mat, unsigned N, unsigned D, unsigned topk, unsigned M = 0,
                         const char* dstFn = nullptr);
    unsigned buildLeafHierarchy(float* mat, unsigned N, unsigned D, unsigned K, const std::vector<uint32_t>& leaf_begin);
    std::vector<LevelStat> levels;
    std::vector<DPGLayer> layers;  // emitted DPG layers, upper layer first
    double dpgSecs = 0.0;          // device time of the layer builds (not in mergeSecs)

private:
    void emitLayer(unsigned maxM, bool carry, const char* layerFile);
    void runLevel(const std::vector<uint32_t>& begin, const std::vector<uint8_t>& built,
                  const std::vector<uint32_t>& group, unsigned cut);
};

// One beam's line of the search record (src/lysearch.cpp:427-429): beam,
// recall@1 against the ground-truth file, seconds, distances per query.
struct SearchRecord {
    unsigned topk;
    double recall;
    double secs;
    unsigned long cmpsPerQuery;  // nCmps / nQry, integer division as in the reference
};

// LySearch (src/lysearch.h:23-93): layered graph search over a reference
// set.  The reference walks its beam list one beam after another, each a
// loop over the queries; here the whole list is one batched launch
// (nnm_search_beams: one wavefront per (beam, query)), and the files and
// records are produced afterwards in the reference's order.
class LySearch {
public:
    // the reference's beam list Ks1 (src/lysearch.cpp:380)
    static const std::vector<unsigned> Ks1;

    // Reference set (.txt / .fvecs / .bvecs) and layered graph file (upper
    // layer first, LySearch::loadGraphs src/lysearch.cpp:112-186).
    LySearch(const char* Rsrc, const char* Gsrc, SpaceInterface<float>* space = nullptr, int device = 0);
    // In memory: row-major N x D rows and CSR layers (upper layer first).
    LySearch(const float* mat, unsigned N, unsigned D, const std::vector<DPGLayer>& layers,
             SpaceInterface<float>* space = nullptr, int device = 0);
    ~LySearch();
    LySearch(const LySearch&) = delete;
    LySearch& operator=(const LySearch&) = delete;

    // src/lysearch.cpp:363-437: queries from qryFn, the Ks1 sweep; dstFn
    // ends with the top-1 result per query of the last beam evaluated
    // ("i\tid\tdist" lines, printTops :684-691), record gets one line per
    // beam ("topk\trecall\tsecs\tnCmps/nQry"), recall@1 against grdFn
    // ("query nn" pairs, Evaluator::load_NNmap src/evaluator.cpp:488-507;
    // Evaluator::NNRecall :199-256); the sweep stops after the first beam
    // whose recall is 1.
    int knnSearch(const char* qryFn, const char* dstFn, const char* grdFn, const char* record);
    // Batched search of nQry queries at every beam of `beams` in one launch;
    // k results per query (k <= the smallest beam), ids / dists as
    // beams x nQry x k.  Records carry recall 0 (no ground truth here).
    std::vector<SearchRecord> knnSearch(const float* queries, unsigned nQry, const std::vector<unsigned>& beams,
                                        unsigned k, std::vector<uint32_t>* ids = nullptr,
                                        std::vector<float>* dists = nullptr);

    unsigned entry = 0;             // entry row of layer 0 (the reference draws rand() % rows, :378)
    unsigned prec = NNM_PREC_FP32;  // NNM_PREC_BF16: bf16 rows with FP32 re-rank of the results
    unsigned width = 1;             // list entries expanded per step (1 = the reference's order)
    std::vector<SearchRecord> records;   // of the last knnSearch
    std::vector<nnm_search_stats> stats;  // of the last knnSearch, per beam
    unsigned rows() const { return row; }
    unsigned dim() const { return col; }

private:
    void check(int rc) const;
    void init(SpaceInterface<float>* space, int device);
    void setLayers(const std::vector<DPGLayer>& layers);

    nnm_ctx* ctx_ = nullptr;
    unsigned row = 0, col = 0;
};

}  // namespace nnmerge

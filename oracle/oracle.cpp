// CPU oracle for the nn-merge local-join path.
//
// TEST INFRASTRUCTURE ONLY.  This file restates the reference algorithms on
// the CPU so that tests, the smoke check and the CPU baseline leg of the
// benchmark can check the HIP product against them.  Nothing in the product
// library links, loads or calls it.
//
// Every function cites the reference file:line it follows (paths relative to
// the reference repository root).  Deliberate, documented deviations:
//   * RNG: the reference seeds std::mt19937 with time(NULL) and shuffles
//     reverse lists with std::random_shuffle (driven by rand()).  Here every
//     generator is std::mt19937 seeded from the caller's seed, and the shuffle
//     is std::shuffle with that generator, so runs are reproducible.
//   * Forward sampling rho and an iteration cap T are parameters; rho = 1 and
//     T = 0 (no cap) reproduce the reference loop exactly.
//   * Partitions generalise the two-subgraph merges to P contiguous
//     subgraphs, each either built (has a k-NN list) or raw.  With two
//     subgraphs this is Smerge (both built) or Jmerge (built + raw).
//   * When fewer than `cut` fresh cross samples are found, the reference
//     reads past tp_seeds (undefined behaviour); here the original entry is
//     kept for the missing slots.
//   * merge2one reads cpnGraph[i][n] past its end in the reference; here a
//     missing compensation entry compares as +inf.  `dedup` additionally
//     drops repeated ids (the product's default), `dedup = 0` keeps the
//     reference's duplicates.
//
// Pins (tests/test_oracle.py).  The delta test (the mean fraction of new
// list entries against 0.002, src/abstractmerge.cpp:171-180) uses a Boost
// accumulator in the reference, which is not vendored; it is restated as the
// plain arithmetic mean and pinned by closed forms: on a complete graph
// with forward sampling rho, delta after one iteration is (K - S) / K,
// equals the flagged entries over N K, and the loop ends when it falls
// strictly below the threshold.  The join's pair rules (KGraph, S-Merge,
// J-Merge) are pinned by closed-form one-iteration pair counts, the reverse
// lists by a hand-derived four-point case, merge2one by brute force over
// pool and parked entries (both dedup settings), getRndSds against numpy's
// MT19937 stream over many (seed, k, N), distances bit for bit against the
// reference's own routines built from its sources.

// Build: see oracle/Makefile (g++ -O2 -std=c++11 -ffp-contract=off, like the
// reference's src/Makefile:33 flags, contraction off so every float op rounds
// as the reference's non-FMA x86-64 code does).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>
#include <chrono>
#include <mutex>
#include <omp.h>

extern "C" {

// ---------------------------------------------------------------- distances

// VMath::l2norm, src/vmath.cpp:56-88.  One accumulator, elements in order
// 0..d-1 (the 4-way unroll does not change the order of additions).
float orc_l2_seq(const float* a, const float* b, uint32_t d) {
    float res = 0.0f;
    for (uint32_t i = 0; i < d; ++i) {
        float diff = a[i] - b[i];
        res += diff * diff;
    }
    return res;
}

// DistOpt::L2SqrSIMDExt, src/distopt.cpp:11-80.  Four lane partials over the
// first 4*floor(d/4) elements (lane l takes elements congruent to l mod 4),
// lanes summed 0+1+2+3 left to right, then the scalar tail added in order.
float orc_l2_simd4(const float* a, const float* b, uint32_t d) {
    float lane[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t q4 = (d >> 2) << 2;
    for (uint32_t i = 0; i < q4; i += 4) {
        for (int l = 0; l < 4; ++l) {
            float diff = a[i + l] - b[i + l];
            float sq = diff * diff;
            lane[l] = lane[l] + sq;
        }
    }
    float res = lane[0] + lane[1];
    res = res + lane[2];
    res = res + lane[3];
    for (uint32_t i = q4; i < d; ++i) {
        float delta = a[i] - b[i];
        res += delta * delta;
    }
    return res;
}

// FP64 squared L2 (VMath::l2 double variant, src/vmath.cpp:34-54).
double orc_l2_f64(const float* a, const float* b, uint32_t d) {
    double w = 0.0;
    for (uint32_t i = 0; i < d; ++i) {
        double t = (double)a[i] - (double)b[i];
        w += t * t;
    }
    return w;
}

// DistOpt::l1norm, src/distopt.cpp:82-110: sum of |a-b| in element order.
float orc_l1(const float* a, const float* b, uint32_t d) {
    float res = 0.0f;
    for (uint32_t i = 0; i < d; ++i) {
        float diff = a[i] - b[i];
        res += std::fabs(diff);
    }
    return res;
}

// DistOpt::KSquare, src/distopt.cpp:163-184: (a-b)^2 / (a+b) summed over the
// elements where a or b is nonzero.
float orc_chi2(const float* a, const float* b, uint32_t d) {
    float w2 = 0.0f;
    for (uint32_t i = 0; i < d; ++i) {
        float w1 = a[i] - b[i];
        if (a[i] != 0 || b[i] != 0) w2 += w1 * w1 / (a[i] + b[i]);
    }
    return w2;
}

// DistOpt::cosine, src/distopt.cpp:113-160: three accumulators in element
// order, res = res / (sqrt(l1) sqrt(l2)) (float square roots), 0 when a norm
// is zero, then 2 - 2 res.
float orc_cosine(const float* a, const float* b, uint32_t d) {
    float res = 0.0f, l1 = 0.0f, l2 = 0.0f;
    for (uint32_t i = 0; i < d; ++i) {
        l1 += a[i] * a[i];
        l2 += b[i] * b[i];
        res += a[i] * b[i];
    }
    if (l1 != 0 && l2 != 0)
        res = res / (std::sqrt(l1) * std::sqrt(l2));
    else
        res = 0;
    return 2 - 2 * res;
}

// FP64 forms of the same metrics, for exact ground truth.
double orc_l1_f64(const float* a, const float* b, uint32_t d) {
    double w = 0.0;
    for (uint32_t i = 0; i < d; ++i) w += std::fabs((double)a[i] - (double)b[i]);
    return w;
}
double orc_chi2_f64(const float* a, const float* b, uint32_t d) {
    double w = 0.0;
    for (uint32_t i = 0; i < d; ++i) {
        const double x = a[i], y = b[i], t = x - y;
        if (x != 0 || y != 0) w += t * t / (x + y);
    }
    return w;
}
double orc_cosine_f64(const float* a, const float* b, uint32_t d) {
    double r = 0.0, l1 = 0.0, l2 = 0.0;
    for (uint32_t i = 0; i < d; ++i) {
        l1 += (double)a[i] * a[i];
        l2 += (double)b[i] * b[i];
        r += (double)a[i] * b[i];
    }
    r = (l1 != 0 && l2 != 0) ? r / (std::sqrt(l1) * std::sqrt(l2)) : 0.0;
    return 2.0 - 2.0 * r;
}

// Metric of the oracle's graph routines and brute force (0 L2, 1 L1,
// 2 cosine, 3 chi-square): the SpaceInterface a graph class is built with
// (src/abstractgraph.h:14-19, src/space_lp.h).
static int g_metric = 0;
void orc_set_metric(int m) { g_metric = m; }


static float dist32(const float* a, const float* b, uint32_t d) {
    switch (g_metric) {
        case 1: return orc_l1(a, b, d);
        case 2: return orc_cosine(a, b, d);
        case 3: return orc_chi2(a, b, d);
        default: return orc_l2_seq(a, b, d);
    }
}
static double dist64(const float* a, const float* b, uint32_t d) {
    switch (g_metric) {
        case 1: return orc_l1_f64(a, b, d);
        case 2: return orc_cosine_f64(a, b, d);
        case 3: return orc_chi2_f64(a, b, d);
        default: return orc_l2_f64(a, b, d);
    }
}

// Brute-force k-NN in FP64 with ties broken by id (ascending).  For each of
// the nq query rows (given as row indices into X, or as separate vectors Q
// when Q != NULL) the k nearest rows of X are returned.  exclude_self drops
// the row whose index equals the query's row index.  This is the exact
// ground truth the reference reads from its (missing) 30-NN file
// (src/evaluator.cpp:343-378).
void orc_bruteforce_f64(const float* X, uint64_t n, uint32_t d,
                        const float* Q, const uint32_t* qrows, uint64_t nq,
                        uint32_t k, int exclude_self,
                        uint32_t* out_ids, double* out_dist) {
    std::vector<std::pair<double, uint32_t>> best;
    for (uint64_t q = 0; q < nq; ++q) {
        const float* qv = Q ? Q + q * d : X + (uint64_t)qrows[q] * d;
        best.clear();
        for (uint64_t r = 0; r < n; ++r) {
            if (exclude_self && !Q && r == qrows[q]) continue;
            double dd = dist64(qv, X + r * d, d);
            std::pair<double, uint32_t> c(dd, (uint32_t)r);
            if (best.size() < k) {
                best.push_back(c);
                std::push_heap(best.begin(), best.end());
            } else if (c < best.front()) {
                std::pop_heap(best.begin(), best.end());
                best.back() = c;
                std::push_heap(best.begin(), best.end());
            }
        }
        std::sort(best.begin(), best.end());
        for (uint32_t j = 0; j < k; ++j) {
            out_ids[q * k + j] = j < best.size() ? best[j].second : 0xFFFFFFFFu;
            out_dist[q * k + j] = j < best.size() ? best[j].first : INFINITY;
        }
    }
}

// ---------------------------------------------------------------- list model

struct ONb {          // mergeDataType::Neighbor, src/pvetex.h:18-30
    uint32_t id;
    float dist;
    uint32_t gd;
    uint8_t flag;
};

// UpdateKnnList, src/pvetex.h:40-85 (same rule as src/kgraph.h:42-85).
// Returns 1 when the entry was inserted.
static uint32_t update_knn_list(ONb* addr, uint32_t K, ONb nn) {
    if (addr[K - 1].dist <= nn.dist) return 0;              // :47
    uint32_t i = K;
    while (i > 0) {                                          // :51-56
        uint32_t j = i - 1;
        if (addr[j].dist <= nn.dist) break;
        i = j;
    }
    uint32_t l = i;
    while (l > 0) {                                          // :59-65
        uint32_t j = l - 1;
        if (addr[j].dist < nn.dist) break;
        if (addr[j].id == nn.id) return 0;
        l = j;
    }
    uint32_t j = K - 1;                                      // :67-81
    while (j > i) { addr[j] = addr[j - 1]; --j; }
    addr[j] = nn;
    return 1;
}

// Array form of UpdateKnnList for unit tests: lists are SoA (ids, dists,
// flags); the new entry is inserted with flag = 1.
uint32_t orc_update_knn_list(uint32_t* ids, float* dists, uint8_t* flags,
                             uint32_t K, uint32_t id, float dist) {
    std::vector<ONb> l(K);
    for (uint32_t i = 0; i < K; ++i) l[i] = ONb{ids[i], dists[i], 0, flags[i]};
    uint32_t r = update_knn_list(l.data(), K, ONb{id, dist, 0, 1});
    for (uint32_t i = 0; i < K; ++i) { ids[i] = l[i].id; dists[i] = l[i].dist; flags[i] = l[i].flag; }
    return r;
}

// AbstractGraph::getRndSds, src/abstractgraph.cpp:7-35: k distinct values in
// [0, N) (sorted draw in [0, N-k), bumped apart, rotated by a random offset).
static void get_rnd_sds(std::mt19937& rng, uint32_t k, uint32_t* addr, uint32_t N) {
    if (N == k) { for (uint32_t i = 0; i < k; ++i) addr[i] = i; return; }
    for (uint32_t i = 0; i < k; ++i) addr[i] = rng() % (N - k);
    std::sort(addr, addr + k);
    for (uint32_t i = 1; i < k; ++i)
        if (addr[i] <= addr[i - 1]) addr[i] = addr[i - 1] + 1;
    uint32_t off = rng() % N;
    for (uint32_t i = 0; i < k; ++i) addr[i] = (addr[i] + off) % N;
}

void orc_get_rnd_sds(uint32_t seed, uint32_t k, uint32_t N, uint32_t* out) {
    std::mt19937 rng(seed);
    get_rnd_sds(rng, k, out, N);
}

// Stats shared by the builders (AbstractGraph::nCmps, src/abstractgraph.h:33).
struct orc_stats {
    uint32_t iters;
    uint32_t converged;
    uint64_t dists;       // join distances (nCmps accumulated by localJoin)
    uint64_t init_dists;  // distances spent by graph/merge set-up
    double secs;          // wall time of the timed region
    double last_delta;
};

struct OParams {
    uint32_t K;
    float rho;
    uint32_t rev_cap;     // smplNum = 100, src/abstractmerge.cpp:81
    uint32_t max_iters;   // 0 = until delta (reference)
    float delta;          // default_delta = 0.002, src/abstractgraph.cpp:4
    uint32_t seed;
};

struct OInd { uint32_t id; uint32_t gid; };   // Ind, src/pvetex.h:35-38

struct OList {                                // NNList, src/pvetex.h:88-126
    std::vector<ONb> pool;
    std::vector<OInd> newl, oldl, rnew, rold;
};

static const uint32_t RAW_GID = 0xFFFFFFFFu;

// The generic NN-Descent loop shared by KGraph::nndescent
// (src/kgraph.cpp:98-189), AbstractMerge::nnDescent
// (src/abstractmerge.cpp:78-192) and LyNNDescent::nnDescent
// (src/lynndescent.cpp:201-310).  gid_of(v) gives the subgraph label the
// reverse entries carry (src/abstractmerge.cpp:107-111); rule_all = 1 keeps
// every pair with distinct ids (KGraph::join, src/kgraph.cpp:191-210, and
// verse = 1 in src/lynndescent.cpp:330), rule_all = 0 skips pairs whose two
// ends lie in the same built subgraph (Smerge::localJoin
// src/smerge.cpp:212-235, Jmerge::localJoin src/jmerge.cpp:178-203).
static void nn_descent(const float* X, uint32_t d, std::vector<OList>& g,
                       const std::vector<uint32_t>& gid, int rule_all,
                       const OParams& p, orc_stats* st, std::mt19937& rng) {
    const uint32_t N = (uint32_t)g.size();
    const uint32_t K = p.K;
    uint32_t S = (uint32_t)std::lround((double)p.rho * K);
    if (S < 1) S = 1;
    if (S > K) S = K;
    uint32_t it = 0;
    st->converged = 0;
    for (;;) {
        // list build, src/abstractmerge.cpp:86-137
        for (uint32_t v = 0; v < N; ++v) {
            OList& L = g[v];
            uint32_t taken = 0;
            for (uint32_t e = 0; e < K; ++e) {
                ONb& nb = L.pool[e];
                OList& o = g[nb.id];
                float maxDst = o.pool[K - 1].dist;
                if (nb.flag) {
                    if (taken >= S) continue;            // forward sampling (rho)
                    ++taken;
                    L.newl.push_back(OInd{nb.id, nb.gd});
                    if (maxDst < nb.dist) o.rnew.push_back(OInd{v, gid[v]});
                    nb.flag = 0;
                } else {
                    L.oldl.push_back(OInd{nb.id, nb.gd});
                    if (maxDst < nb.dist) o.rold.push_back(OInd{v, gid[v]});
                }
            }
        }
        // reverse sampling, src/abstractmerge.cpp:139-166
        for (uint32_t v = 0; v < N; ++v) {
            OList& L = g[v];
            std::shuffle(L.rnew.begin(), L.rnew.end(), rng);
            if (L.rnew.size() > p.rev_cap) L.rnew.resize(p.rev_cap);
            std::shuffle(L.rold.begin(), L.rold.end(), rng);
            if (L.rold.size() > p.rev_cap) L.rold.resize(p.rev_cap);
            L.newl.insert(L.newl.end(), L.rnew.begin(), L.rnew.end());
            L.oldl.insert(L.oldl.end(), L.rold.begin(), L.rold.end());
        }
        // local join, NNList::join src/pvetex.h:111-125 + the localJoin lambdas
        uint64_t cc = 0;
        for (uint32_t v = 0; v < N; ++v) {
            const OList& L = g[v];
            const size_t ns = L.newl.size(), os = L.oldl.size();
            for (size_t i = 0; i < ns; ++i) {
                const OInd si = L.newl[i];
                for (size_t j = i + 1; j < ns + os; ++j) {
                    const OInd sj = j < ns ? L.newl[j] : L.oldl[j - ns];
                    if (si.id == sj.id) continue;
                    if (!rule_all && si.gid == sj.gid && si.gid != RAW_GID) continue;
                    float dst = dist32(X + (uint64_t)si.id * d, X + (uint64_t)sj.id * d, d);
                    ++cc;
                    update_knn_list(g[si.id].pool.data(), K, ONb{sj.id, dst, sj.gid, 1});
                    update_knn_list(g[sj.id].pool.data(), K, ONb{si.id, dst, si.gid, 1});
                }
            }
        }
        st->dists += cc;
        ++it;
        // EvaluateDelta + mean, src/abstractmerge.cpp:171-180, :194-202
        double acc = 0.0;
        for (uint32_t v = 0; v < N; ++v) {
            uint32_t c = 0;
            for (uint32_t e = 0; e < K; ++e) c += g[v].pool[e].flag ? 1 : 0;
            acc += (double)((float)c / (float)K);
        }
        double mean = N ? acc / N : 0.0;
        st->last_delta = mean;
        for (uint32_t v = 0; v < N; ++v) {
            g[v].newl.clear(); g[v].oldl.clear(); g[v].rnew.clear(); g[v].rold.clear();
        }
        if (mean < p.delta) { st->converged = 1; break; }
        if (p.max_iters && it >= p.max_iters) break;
    }
    st->iters = it;
}

// nn_descent on every core: the reference's parallel form, which its
// sources carry as commented OpenMP pragmas (schedule(dynamic, 100) over
// the rows in Jmerge::localJoin, src/jmerge.cpp:185, and
// Smerge::localJoin, src/smerge.cpp:220) with one lock per list
// (NNList::parallel_try_insert, src/pvetex.h:99-109, and the LockGuard
// around the reverse pushes, src/abstractmerge.cpp:103-104).  The same
// algorithm as nn_descent; the reverse lists are shuffled with a generator
// per row and iteration (seed, iteration, row), and the insertion order
// of concurrent offers is the scheduler's, so results differ from the
// sequential loop's in their random choices only.
static void nn_descent_mt(const float* X, uint32_t d, std::vector<OList>& g,
                          const std::vector<uint32_t>& gid, int rule_all,
                          const OParams& p, orc_stats* st) {
    const uint32_t N = (uint32_t)g.size();
    const uint32_t K = p.K;
    uint32_t S = (uint32_t)std::lround((double)p.rho * K);
    if (S < 1) S = 1;
    if (S > K) S = K;
    std::vector<std::mutex> lock(N);
    uint32_t it = 0;
    st->converged = 0;
    for (;;) {
#pragma omp parallel for schedule(dynamic, 256)
        for (uint32_t v = 0; v < N; ++v) {
            OList& L = g[v];
            uint32_t taken = 0;
            for (uint32_t e = 0; e < K; ++e) {
                ONb& nb = L.pool[e];
                OList& o = g[nb.id];
                // pools change only in the join: no lock for reading them here
                const float maxDst = o.pool[K - 1].dist;
                if (nb.flag) {
                    if (taken >= S) continue;
                    ++taken;
                    L.newl.push_back(OInd{nb.id, nb.gd});
                    if (maxDst < nb.dist) {
                        std::lock_guard<std::mutex> gd(lock[nb.id]);
                        o.rnew.push_back(OInd{v, gid[v]});
                    }
                    nb.flag = 0;
                } else {
                    L.oldl.push_back(OInd{nb.id, nb.gd});
                    if (maxDst < nb.dist) {
                        std::lock_guard<std::mutex> gd(lock[nb.id]);
                        o.rold.push_back(OInd{v, gid[v]});
                    }
                }
            }
        }
#pragma omp parallel for schedule(dynamic, 256)
        for (uint32_t v = 0; v < N; ++v) {
            OList& L = g[v];
            std::mt19937 rng(p.seed * 0x9E3779B9u ^ (it * 0x85EBCA6Bu) ^ (v * 0xC2B2AE35u));
            std::shuffle(L.rnew.begin(), L.rnew.end(), rng);
            if (L.rnew.size() > p.rev_cap) L.rnew.resize(p.rev_cap);
            std::shuffle(L.rold.begin(), L.rold.end(), rng);
            if (L.rold.size() > p.rev_cap) L.rold.resize(p.rev_cap);
            L.newl.insert(L.newl.end(), L.rnew.begin(), L.rnew.end());
            L.oldl.insert(L.oldl.end(), L.rold.begin(), L.rold.end());
        }
        uint64_t cc = 0;
#pragma omp parallel for schedule(dynamic, 100) reduction(+ : cc)
        for (uint32_t v = 0; v < N; ++v) {
            const OList& L = g[v];
            const size_t ns = L.newl.size(), os = L.oldl.size();
            for (size_t i = 0; i < ns; ++i) {
                const OInd si = L.newl[i];
                for (size_t j = i + 1; j < ns + os; ++j) {
                    const OInd sj = j < ns ? L.newl[j] : L.oldl[j - ns];
                    if (si.id == sj.id) continue;
                    if (!rule_all && si.gid == sj.gid && si.gid != RAW_GID) continue;
                    float dst = dist32(X + (uint64_t)si.id * d, X + (uint64_t)sj.id * d, d);
                    ++cc;
                    {
                        std::lock_guard<std::mutex> gd(lock[si.id]);
                        update_knn_list(g[si.id].pool.data(), K, ONb{sj.id, dst, sj.gid, 1});
                    }
                    {
                        std::lock_guard<std::mutex> gd(lock[sj.id]);
                        update_knn_list(g[sj.id].pool.data(), K, ONb{si.id, dst, si.gid, 1});
                    }
                }
            }
        }
        st->dists += cc;
        ++it;
        double acc = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
        for (uint32_t v = 0; v < N; ++v) {
            uint32_t c = 0;
            for (uint32_t e = 0; e < K; ++e) c += g[v].pool[e].flag ? 1 : 0;
            acc += (double)((float)c / (float)K);
            g[v].newl.clear(); g[v].oldl.clear(); g[v].rnew.clear(); g[v].rold.clear();
        }
        double mean = N ? acc / N : 0.0;
        st->last_delta = mean;
        if (mean < p.delta) { st->converged = 1; break; }
        if (p.max_iters && it >= p.max_iters) break;
    }
    st->iters = it;
}

static bool nb_less(const ONb& a, const ONb& b) { return a.dist < b.dist; }   // Neighbor::ComNeighbor

// KGraph::InitialGraph, src/kgraph.cpp:56-96: K random distinct neighbours
// (self excluded) for every vertex of [b, e), ids local to the block.
static void random_init(const float* X, uint32_t d, std::vector<OList>& g,
                        uint32_t b, uint32_t e, uint32_t K, const std::vector<uint32_t>& gid,
                        std::mt19937& rng, uint64_t* nd) {
    std::vector<uint32_t> seeds(K + 1);
    const uint32_t n = e - b;
    for (uint32_t i = b; i < e; ++i) {
        g[i].pool.assign(K, ONb{0, INFINITY, 0, 1});
        get_rnd_sds(rng, K + 1, seeds.data(), n);
        uint32_t L = 0;
        for (uint32_t j = 0; j < K + 1 && L < K; ++j) {
            uint32_t id = b + seeds[j];
            if (id == i) continue;
            ONb& nb = g[i].pool[L++];
            nb.id = id;
            nb.dist = dist32(X + (uint64_t)i * d, X + (uint64_t)id * d, d);
            nb.flag = 1;
            nb.gd = gid[id];
            ++*nd;
        }
        std::sort(g[i].pool.begin(), g[i].pool.end(), nb_less);
    }
}

static double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// KGraph::buildKGraph, src/kgraph.cpp:217-245 (InitialGraph + nndescent).
// Lists come back sorted by distance, ids local to X.
void orc_kgraph(const float* X, uint32_t n, uint32_t d, uint32_t K, float rho,
                uint32_t rev_cap, uint32_t max_iters, float delta, uint32_t seed,
                uint32_t* out_ids, float* out_dist, orc_stats* st) {
    std::memset(st, 0, sizeof(*st));
    OParams p{K, rho, rev_cap, max_iters, delta, seed};
    std::mt19937 rng(seed);
    std::vector<OList> g(n);
    std::vector<uint32_t> gid(n, 0);
    double t0 = now_s();
    random_init(X, d, g, 0, n, K, gid, rng, &st->init_dists);
    nn_descent(X, d, g, gid, 1, p, st, rng);
    st->secs = now_s() - t0;
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < K; ++j) {
            out_ids[(uint64_t)i * K + j] = g[i].pool[j].id;
            out_dist[(uint64_t)i * K + j] = g[i].pool[j].dist;
        }
}

// merge2one, src/abstractmerge.cpp:232-277: the first K entries of the
// two-way merge of the pool (K entries, sorted) and the parked entries c
// (sorted); the pool wins distance ties.  dedup drops an id already taken.
static void merge2one(std::vector<ONb>& pool, const std::vector<ONb>& c, uint32_t K, int dedup) {
    std::vector<ONb> tmp(K);
    uint32_t a = 0, b = 0, k = 0;
    while (k < K && (a < K || b < c.size())) {
        const float da = a < K ? pool[a].dist : INFINITY;
        const float db = b < c.size() ? c[b].dist : INFINITY;
        ONb pick;
        if (a < K && da <= db) pick = pool[a++];     // pool wins ties
        else pick = c[b++];
        if (dedup) {
            bool dup = false;
            for (uint32_t t = 0; t < k; ++t)
                if (tmp[t].id == pick.id) { dup = true; break; }
            if (dup) continue;
        }
        tmp[k++] = pick;
    }
    for (uint32_t j = 0; j < k; ++j) pool[j] = tmp[j];
}

// merge2one over arrays: pool n x K and parked entries n x cut (ids, dists,
// each row sorted by distance; a parked id of 0xFFFFFFFF ends its row).
void orc_merge2one(uint32_t n, uint32_t K, uint32_t* pool_ids, float* pool_dist, uint32_t cut,
                   const uint32_t* cpn_ids, const float* cpn_dist, int dedup) {
    std::vector<ONb> pool(K), c;
    for (uint32_t v = 0; v < n; ++v) {
        for (uint32_t j = 0; j < K; ++j)
            pool[j] = ONb{pool_ids[(uint64_t)v * K + j], pool_dist[(uint64_t)v * K + j], 0, 0};
        c.clear();
        for (uint32_t j = 0; j < cut && cpn_ids[(uint64_t)v * cut + j] != 0xFFFFFFFFu; ++j)
            c.push_back(ONb{cpn_ids[(uint64_t)v * cut + j], cpn_dist[(uint64_t)v * cut + j], 0, 0});
        merge2one(pool, c, K, dedup);
        for (uint32_t j = 0; j < K; ++j) {
            pool_ids[(uint64_t)v * K + j] = pool[j].id;
            pool_dist[(uint64_t)v * K + j] = pool[j].dist;
        }
    }
}

// The NN-Descent loop (nn_descent) from lists given by the caller: ids n x K
// (each row's entries distinct, self excluded), flags n x K, and the
// subgraph label of every row (gid, 0xFFFFFFFF for raw rows); distances are
// computed here and each row is sorted by distance first.  rule_all = 1 is
// KGraph::join's rule, 0 the merge rule.  The lists after the loop come
// back in ids / dist / flags.
void orc_nndescent_lists(const float* X, uint32_t n, uint32_t d, uint32_t K, uint32_t* ids, float* dist,
                         uint8_t* flags, const uint32_t* gid_of, int rule_all, float rho, uint32_t rev_cap,
                         uint32_t max_iters, float delta, uint32_t seed, orc_stats* st) {
    std::memset(st, 0, sizeof(*st));
    OParams p{K, rho, rev_cap, max_iters, delta, seed};
    std::mt19937 rng(seed);
    std::vector<OList> g(n);
    std::vector<uint32_t> gid(gid_of, gid_of + n);
    for (uint32_t v = 0; v < n; ++v) {
        g[v].pool.resize(K);
        for (uint32_t j = 0; j < K; ++j) {
            ONb& nb = g[v].pool[j];
            nb.id = ids[(uint64_t)v * K + j];
            nb.dist = dist32(X + (uint64_t)v * d, X + (uint64_t)nb.id * d, d);
            nb.flag = flags[(uint64_t)v * K + j];
            nb.gd = gid[nb.id];
        }
        std::stable_sort(g[v].pool.begin(), g[v].pool.end(), nb_less);
    }
    double t0 = now_s();
    nn_descent(X, d, g, gid, rule_all, p, st, rng);
    st->secs = now_s() - t0;
    for (uint32_t v = 0; v < n; ++v)
        for (uint32_t j = 0; j < K; ++j) {
            ids[(uint64_t)v * K + j] = g[v].pool[j].id;
            dist[(uint64_t)v * K + j] = g[v].pool[j].dist;
            flags[(uint64_t)v * K + j] = g[v].pool[j].flag;
        }
}

// Partitioned merge: AbstractMerge::InitSGraph (src/abstractmerge.cpp:55-76),
// Smerge::InitialGraph (src/smerge.cpp:94-209) / Jmerge::InitialGraph
// (src/jmerge.cpp:83-176), AbstractMerge::nnDescent with the merge pair rule,
// then merge2one (src/abstractmerge.cpp:232-277).  The timed region matches
// src/smerge.cpp:75-81 / src/jmerge.cpp:63-70.
//
// Subgraph s covers rows [sub_begin[s], sub_begin[s+1]); sub_built[s] says
// whether it arrives with a k-NN graph.  init_ids holds, for every vertex of a
// built subgraph, its K neighbour ids (global, sorted by distance); rows of
// raw subgraphs are ignored.  cut is the far part replaced by cross samples
// (cutSize = K/2 in src/smerge.cpp:99 and K*3/6 in src/jmerge.cpp:87).
static void merge_impl(const float* X, uint32_t n, uint32_t d, uint32_t K,
                       uint32_t nsub, const uint32_t* sub_begin, const uint8_t* sub_built,
                       const uint32_t* init_ids, uint32_t cut, float rho, uint32_t rev_cap,
                       uint32_t max_iters, float delta, uint32_t seed, int dedup,
                       uint32_t* out_ids, float* out_dist, orc_stats* st, bool mt) {
    std::memset(st, 0, sizeof(*st));
    OParams p{K, rho, rev_cap, max_iters, delta, seed};
    std::mt19937 rng(seed);
    std::vector<OList> g(n);
    std::vector<std::vector<ONb>> cpn(n);
    std::vector<uint32_t> gid(n), sub(n);
    for (uint32_t s = 0; s < nsub; ++s)
        for (uint32_t v = sub_begin[s]; v < sub_begin[s + 1]; ++v) {
            sub[v] = s;
            gid[v] = sub_built[s] ? s : RAW_GID;
        }
    const uint32_t remain = K - cut;
    double t0 = now_s();
    // InitSGraph: recompute distances of the given lists, all flagged new.
    uint64_t nd = 0;
#pragma omp parallel for schedule(static) reduction(+ : nd) if (mt)
    for (uint32_t v = 0; v < n; ++v) {
        if (!sub_built[sub[v]]) continue;
        g[v].pool.resize(K);
        for (uint32_t j = 0; j < K; ++j) {
            ONb& nb = g[v].pool[j];
            nb.id = init_ids[(uint64_t)v * K + j];
            nb.dist = dist32(X + (uint64_t)v * d, X + (uint64_t)nb.id * d, d);
            nb.flag = 1;
            nb.gd = gid[nb.id];
            ++nd;
        }
    }
    auto setup_row = [&](uint32_t v, std::mt19937& rng, std::vector<uint32_t>& seeds,
                         std::vector<uint32_t>& fresh, uint64_t& nd) {
        const uint32_t s = sub[v];
        if (sub_built[s]) {
            // Cross samples from the other subgraphs (src/smerge.cpp:96-142).
            const uint32_t own_b = sub_begin[s], own_n = sub_begin[s + 1] - sub_begin[s];
            const uint32_t others = n - own_n;
            cpn[v].assign(g[v].pool.begin() + remain, g[v].pool.end());
            if (others == 0) return;
            uint32_t draw = 2 * cut < others ? 2 * cut : others;
            get_rnd_sds(rng, draw, seeds.data(), others);
            fresh.clear();
            for (uint32_t m = 0; m < draw && fresh.size() < cut; ++m) {
                uint32_t r = seeds[m];
                uint32_t id = r < own_b ? r : r + own_n;
                bool dup = false;
                for (uint32_t l = 0; l < K; ++l)
                    if (g[v].pool[l].id == id) { dup = true; break; }
                if (!dup) fresh.push_back(id);
            }
            for (uint32_t j = 0; j < cut && j < fresh.size(); ++j) {
                ONb& nb = g[v].pool[remain + j];
                nb.id = fresh[j];
                nb.gd = gid[nb.id];
                nb.flag = 1;
                nb.dist = dist32(X + (uint64_t)nb.id * d, X + (uint64_t)v * d, d);
                ++nd;
            }
            std::sort(g[v].pool.begin(), g[v].pool.end(), nb_less);
        } else {
            // Raw rows: K random neighbours over all rows (src/jmerge.cpp:147-171).
            g[v].pool.assign(K, ONb{0, INFINITY, 0, 1});
            get_rnd_sds(rng, K + 1, seeds.data(), n);
            uint32_t L = 0;
            for (uint32_t j = 0; j < K + 1 && L < K; ++j) {
                if (seeds[j] == v) continue;
                ONb& nb = g[v].pool[L++];
                nb.id = seeds[j];
                nb.gd = gid[nb.id];
                nb.flag = 1;
                nb.dist = dist32(X + (uint64_t)nb.id * d, X + (uint64_t)v * d, d);
                ++nd;
            }
            std::sort(g[v].pool.begin(), g[v].pool.end(), nb_less);
        }
    };
    if (!mt) {
        std::vector<uint32_t> seeds(2 * K + 2), fresh;
        for (uint32_t v = 0; v < n; ++v) setup_row(v, rng, seeds, fresh, nd);
    } else {
        // one generator per row: the draws no longer depend on row order
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : nd)
        for (uint32_t v = 0; v < n; ++v) {
            std::mt19937 r(seed * 0x9E3779B9u ^ (v * 0xC2B2AE35u) ^ 0x5EEDu);
            std::vector<uint32_t> seeds(2 * K + 2), fresh;
            setup_row(v, r, seeds, fresh, nd);
        }
    }
    st->init_dists = nd;
    if (mt) nn_descent_mt(X, d, g, gid, 0, p, st);
    else nn_descent(X, d, g, gid, 0, p, st, rng);
#pragma omp parallel for schedule(static) if (mt)
    for (uint32_t v = 0; v < n; ++v)
        if (!cpn[v].empty()) merge2one(g[v].pool, cpn[v], K, dedup);
    st->secs = now_s() - t0;
    for (uint32_t v = 0; v < n; ++v)
        for (uint32_t j = 0; j < K; ++j) {
            out_ids[(uint64_t)v * K + j] = g[v].pool[j].id;
            out_dist[(uint64_t)v * K + j] = g[v].pool[j].dist;
        }
}

void orc_merge(const float* X, uint32_t n, uint32_t d, uint32_t K,
               uint32_t nsub, const uint32_t* sub_begin, const uint8_t* sub_built,
               const uint32_t* init_ids, uint32_t cut, float rho, uint32_t rev_cap,
               uint32_t max_iters, float delta, uint32_t seed, int dedup,
               uint32_t* out_ids, float* out_dist, orc_stats* st) {
    merge_impl(X, n, d, K, nsub, sub_begin, sub_built, init_ids, cut, rho, rev_cap, max_iters, delta, seed, dedup,
               out_ids, out_dist, st, false);
}

// orc_merge on `threads` host threads (0 = OpenMP's default); returns the
// number of threads used.
int orc_merge_mt(const float* X, uint32_t n, uint32_t d, uint32_t K,
                 uint32_t nsub, const uint32_t* sub_begin, const uint8_t* sub_built,
                 const uint32_t* init_ids, uint32_t cut, float rho, uint32_t rev_cap,
                 uint32_t max_iters, float delta, uint32_t seed, int dedup,
                 uint32_t* out_ids, float* out_dist, orc_stats* st, int threads) {
    const int nt = threads > 0 ? threads : omp_get_max_threads();
    omp_set_num_threads(nt);
    merge_impl(X, n, d, K, nsub, sub_begin, sub_built, init_ids, cut, rho, rev_cap, max_iters, delta, seed, dedup,
               out_ids, out_dist, st, true);
    return nt;
}

// Evaluator::getKGraphRecall, src/evaluator.cpp:258-295 with intersect
// (src/evaluator.cpp:568-579): hits of the first k ids of each list in the
// ground-truth top-k set, over nrows*k.  unique = 1 counts a repeated id once
// (the reference counts it each time).
double orc_recall(const uint32_t* ids, uint32_t K, const uint32_t* rows, uint64_t nrows,
                  const uint32_t* gt, uint32_t gtK, uint32_t k, int unique) {
    uint64_t hits = 0;
    for (uint64_t r = 0; r < nrows; ++r) {
        const uint32_t* l = ids + (uint64_t)rows[r] * K;
        const uint32_t* t = gt + r * gtK;
        for (uint32_t i = 0; i < k; ++i) {
            if (unique) {
                bool seen = false;
                for (uint32_t j = 0; j < i; ++j) if (l[j] == l[i]) { seen = true; break; }
                if (seen) continue;
            }
            for (uint32_t j = 0; j < k; ++j)
                if (t[j] == l[i]) { ++hits; break; }
        }
    }
    return nrows ? (double)hits / ((double)nrows * k) : 0.0;
}

}  // extern "C"

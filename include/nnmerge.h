/*
 * nnmerge.h -- C ABI of the MI355X (gfx950) NN-Descent local-join library.
 *
 * The reference (wlzhao/nn-merge) is a single C++11 program without an FFI:
 * its merge classes call the local join through a virtual method
 * (AbstractMerge::localJoin, src/abstractmerge.h:30) once per iteration of
 * AbstractMerge::nnDescent (src/abstractmerge.cpp:78-192), and every pair goes
 * through NNList::join (src/pvetex.h:111-125), the distance pointer DISTFUNC
 * (src/abstractgraph.h:11-12) and NNList::parallel_try_insert
 * (src/pvetex.h:99-109).  This header is the boundary the host classes
 * (Smerge, Jmerge, KGraph, LyNNDescent, LySearch in nn-merge_amd/host) call in
 * place of those in-process calls.  Each entry point names the reference
 * routine it replaces.
 *
 * Conventions: plain pointers and sizes, no exceptions across the ABI, every
 * call returns an int status (NNM_OK = 0, negative = error) and
 * nnm_last_error() returns the text of the last failure on that context.
 * Host arrays are copied into device memory owned by the context; the library
 * never takes ownership of caller memory.  Calls are synchronous.
 * A context is used by one host thread at a time.
 */
#ifndef NNMERGE_H
#define NNMERGE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NNM_OK 0
#define NNM_ERR_INVALID -1     /* bad argument or shape */
#define NNM_ERR_HIP -2         /* HIP runtime failure */
#define NNM_ERR_NOMEM -3       /* device allocation failed or over the cap */
#define NNM_ERR_STATE -4       /* call out of order (no data, no graph, ...) */
#define NNM_ERR_UNSUPPORTED -5 /* valid request this build does not handle */

/* Subgraph label of rows that arrive without a k-NN graph (Gid == 1 rows of
 * Jmerge, src/jmerge.cpp:160-163; the new block of LyNNDescent::appdNwSmpl,
 * src/lynndescent.cpp:418-445). */
#define NNM_RAW 0xFFFFFFFFu

/* Distance precision of the local join. */
#define NNM_PREC_FP32 0    /* FP32 difference form on the vector ALUs */
#define NNM_PREC_BF16 1    /* BF16 MFMA norm form, FP32 re-rank */
#define NNM_PREC_LEARNED 2 /* learned projection prefilter, FP32 re-rank;
                              without a trained model every iteration uses
                              the plain BF16 join */

/* Pair rule of the local join. */
#define NNM_RULE_ALL 0   /* every pair with distinct ids: KGraph::join
                            (src/kgraph.cpp:191-210), verse = 1 in
                            src/lynndescent.cpp:330 */
#define NNM_RULE_MERGE 1 /* skip pairs inside one built subgraph:
                            Smerge::localJoin (src/smerge.cpp:224),
                            Jmerge::localJoin (src/jmerge.cpp:189-191),
                            LyNNDescent::localJoin (src/lynndescent.cpp:330) */

/* nnm_params.flags */
#define NNM_FLAG_GRAPH 1u    /* replay each iteration from a captured HIP graph */
#define NNM_FLAG_AUTOTUNE 2u /* time the join launch shapes on the first
                                iteration and keep the fastest (results do not
                                depend on the shape) */
#define NNM_FLAG_LEARNED_RENEW 4u /* NNM_PREC_LEARNED: when the lists leave the
                                     learned space, every entry is marked new so
                                     the plain join re-judges all of them */
#define NNM_FLAG_LEARNED_FP8 8u /* NNM_PREC_LEARNED: the learned phase works on
                                   FP8 (OCP e4m3) rows of the projection,
                                   scaled by 448 / max |p|, on the 16x16x32
                                   fp8 MFMA (d_out = 32); estimates and the
                                   plain tail are unchanged otherwise */
#define NNM_FLAG_LEARNED_MARK_FOUND 16u /* NNM_PREC_LEARNED: when the lists
                                   leave the learned space, the entries the
                                   learned phase found (ids not in the row when
                                   it started) are marked new, so the plain
                                   join re-judges them; the others keep their
                                   flags (RENEW marks every entry) */

/* Re-rank policy for approximate precisions. */
#define NNM_RERANK_AUTO 0   /* skip when the approximate form is exact */
#define NNM_RERANK_ALWAYS 1
#define NNM_RERANK_NEVER 2

/* Distance metrics (SpaceInterface, src/abstractgraph.h:14-19; the spaces
 * of src/space_lp.h:9-53 and the routines of src/distopt.cpp).  Each is
 * evaluated in FP32 with the reference's operation order: one accumulator,
 * elements 0..d-1, every multiply and add rounded separately (except the
 * library's L2, which uses fused multiply-add, exact on integer rows). */
#define NNM_METRIC_L2 0     /* squared L2: VMath::l2norm, src/vmath.cpp:56-88 */
#define NNM_METRIC_L1 1     /* sum |a-b|: DistOpt::l1norm, src/distopt.cpp:82-110 */
#define NNM_METRIC_COSINE 2 /* 2 - 2 a.b / (|a| |b|), 2 when a norm is 0:
                               DistOpt::cosine, src/distopt.cpp:113-160 */
#define NNM_METRIC_CHI2 3   /* sum (a-b)^2 / (a+b) over elements with a or b
                               nonzero: DistOpt::KSquare, src/distopt.cpp:163-184 */

typedef struct nnm_ctx nnm_ctx;

/* NN-Descent parameters.  Defaults (nnm_default_params) equal the reference
 * constants: K = 20, rho = 1 (every new entry joins), reverse cap 100
 * (smplNum, src/abstractmerge.cpp:81), no iteration cap, delta = 0.002
 * (default_delta, src/abstractgraph.cpp:4). */
typedef struct nnm_params {
    uint32_t K;         /* list length */
    float rho;          /* forward sampling: at most round(rho*K) new entries */
    uint32_t rev_cap;   /* reverse list cap (new and old each) */
    uint32_t max_iters; /* iteration cap, 0 = until delta */
    float delta;        /* stop when mean fraction of new entries < delta */
    uint32_t rule;      /* NNM_RULE_* */
    uint32_t prec;      /* NNM_PREC_* */
    uint32_t rerank;    /* NNM_RERANK_* */
    uint32_t offer_cap; /* offer-log capacity: records per row the join may
                           write per delivery round; 0 = 256 per row plus
                           one 4096-record chunk per join wave.  Every
                           offer reaches its target's list: when the log
                           fills, the vertices it could not take are joined
                           again in another round (slower, same result) */
    uint32_t cut;       /* merge set-up: far entries replaced by cross
                           samples; 0 = K/2 (cutSize, src/smerge.cpp:99) */
    uint32_t plain_tail;/* NNM_PREC_LEARNED with max_iters: the last
                           plain_tail iterations use the plain BF16 join */
    uint32_t flags;     /* NNM_FLAG_* */
    float learn_tau;    /* NNM_PREC_LEARNED: the learned-phase join offers a
                           pair to a target when its estimate is below
                           learn_tau x the target's estimated k-th distance
                           (0 = 0.6); the plain tail uses the exact k-th */
    float learn_rho;    /* NNM_PREC_LEARNED: forward sampling rate of the
                           learned phase (at most round(learn_rho*K) new
                           entries; 0 = rho); the plain tail samples at rho */
    uint64_t seed;      /* every random choice derives from this seed */
} nnm_params;

/* Counters of one call.  dists follows the reference's nCmps: one per pair
 * that passes the pair rule (src/smerge.cpp:226-227). */
typedef struct nnm_stats {
    uint32_t iters;
    uint32_t converged;    /* 1 when the delta test ended the loop */
    uint64_t dists;        /* join distances evaluated */
    uint64_t init_dists;   /* distances spent by set-up calls */
    uint64_t updates;      /* list entries replaced */
    uint64_t reranked;     /* candidates recomputed in exact FP32 */
    uint64_t list_entries; /* candidate-list entries joined */
    uint64_t offers;       /* offers that passed the k-th prefilter (each is
                              written to the offer log and delivered) */
    double last_delta;     /* mean fraction of new entries after the loop */
    double secs;           /* device time of the call */
    double join_secs;      /* device time spent in the join kernels */
    uint32_t learned_iters;/* iterations joined through the learned prefilter */
    uint32_t plain_iters;  /* iterations joined with the plain join */
    uint32_t tile_js;      /* join window rows used (autotuned or default) */
    uint32_t tile_chunk;   /* rows per work-queue grab used */
    uint32_t graph_iters;  /* iterations replayed from a HIP graph */
    uint32_t pad_;
    uint64_t long_dists;   /* part of dists / list_entries / offers that the */
    uint64_t long_entries; /* long-list join kernels account for; the rest  */
    uint64_t long_offers;  /* is the short-list (packed) join kernel's      */
    double join_main_secs; /* short-list join kernel time (eager iterations) */
    uint64_t join_main_bytes; /* its algorithmic bytes: per list entry the row,
                                 id, label, k-th distance (+ norm in the norm
                                 form); 8 per offer; 8 per vertex */
    uint64_t join_main_flops; /* its distance flops: 3d per FP32 distance
                                 (difference form), 2d per MFMA distance */
    uint64_t near_ties;    /* k-th entry decisions re-made on FP64 distances
                              (FP32 distances within rounding of each other) */
    uint64_t tie_swaps;    /* ... where FP64 kept the other entry */
    uint64_t redo_rows;    /* vertices joined again because the offer log
                              was full (0 when offer_cap suffices) */
    uint64_t learned_dists;/* part of dists joined on learned estimates
                              (projected distances, not exact ones) */
    uint32_t redo_rounds;  /* extra join rounds those vertices took */
    uint32_t pad2_;
    double deliver_secs;   /* device time of offer delivery (partition passes
                              and list owners; eager iterations) */
} nnm_stats;

/* FP64 error of the stored FP32 list distances (nnm_graph_dist_error). */
typedef struct nnm_dist_error {
    double max_abs;       /* max |d32 - d64| */
    double max_rel;       /* max |d32 - d64| / d64 */
    double mean_rel;      /* mean relative error over all entries */
    uint64_t entries;     /* list entries checked (active rows x K) */
    uint64_t order_flips; /* adjacent entries whose FP64 order disagrees */
} nnm_dist_error;

void nnm_default_params(nnm_params* p);

/* --- context ---------------------------------------------------------- */
int nnm_create(nnm_ctx** out, int device);
int nnm_destroy(nnm_ctx* ctx);
const char* nnm_last_error(const nnm_ctx* ctx);
/* Device-memory cap in bytes for this context (0 = no cap).  Allocations
 * past the cap fail with NNM_ERR_NOMEM. */
int nnm_set_mem_cap(nnm_ctx* ctx, uint64_t bytes);
uint64_t nnm_mem_used(const nnm_ctx* ctx);

/* --- data ------------------------------------------------------------- */
/* Row-major n x d float32 matrix (AbstractGraph::mat, src/abstractgraph.h:28,
 * loaded by IODelegator::load_refSet, src/iodelegator.cpp:159-182).
 * Replaces the matrix every distance call indexes through DISTFUNC
 * (src/abstractgraph.h:11-12).  Resets graph and partition state. */
int nnm_set_data(nnm_ctx* ctx, const float* X, uint64_t n, uint32_t d);

/* Metric of every later distance (NNM_METRIC_*; the default is L2).  The
 * choice of the space object in the reference (L2space / L1space / Ksquare,
 * src/space_lp.h:9-53, passed as SpaceInterface to the graph classes).
 * Metrics other than L2 run the FP32 join (NNM_PREC_FP32), random and merge
 * set-up, list recompute, brute force and pair distances; the BF16 and
 * learned joins, search, DPG and the FP64 error report return
 * NNM_ERR_UNSUPPORTED for them. */
int nnm_set_metric(nnm_ctx* ctx, uint32_t metric);

/* Partition of the rows into nsub contiguous subgraphs: subgraph s covers
 * rows [sub_begin[s], sub_begin[s+1]) (sub_begin has nsub+1 entries, the
 * last equal to n).  sub_built[s] = 1 for a subgraph that has a k-NN graph
 * (its rows carry label s), 0 for raw rows (label NNM_RAW).  sub_group[s]
 * puts subgraphs into merge groups: random samples are drawn inside a group
 * and groups never exchange neighbours, so several merges (or several
 * independent builds) run in one launch.  Replaces the Gid bookkeeping of
 * src/abstractmerge.cpp:107-111 and LyNNDescent::getGraphId
 * (src/lynndescent.cpp:189-199).  Groups must be contiguous runs of
 * subgraphs. */
int nnm_set_partition(nnm_ctx* ctx, const uint32_t* sub_begin, const uint8_t* sub_built,
                      const uint32_t* sub_group, uint32_t nsub);

/* --- graph ------------------------------------------------------------ */
/* Random initial lists: K distinct random neighbours per row inside its
 * group, self excluded, sorted by distance, all flagged new.  Replaces
 * KGraph::InitialGraph (src/kgraph.cpp:56-96). */
int nnm_graph_random_init(nnm_ctx* ctx, uint32_t K, uint64_t seed, nnm_stats* st);

/* Upload lists (n x K ids, global row numbers).  dist == NULL recomputes the
 * distances (AbstractMerge::InitSGraph, src/abstractmerge.cpp:55-76); flags
 * == NULL marks every entry new.  Rows are sorted by (dist, id) on upload. */
int nnm_set_graph(nnm_ctx* ctx, const uint32_t* ids, const float* dist, const uint8_t* flags,
                  uint32_t K, nnm_stats* st);

/* Download lists: ids and dist are n x K (either may be NULL), flags n x K
 * (may be NULL).  Replaces the pool read-out of AbstractMerge::saveKGraph
 * (src/abstractmerge.cpp:278-291) and KGraph::buildKGraph
 * (src/kgraph.cpp:228-238). */
int nnm_get_graph(nnm_ctx* ctx, uint32_t* ids, float* dist, uint8_t* flags);

/* Device-side checkpoint of the lists: nnm_graph_save copies ids, dists and
 * flags into a second device buffer, nnm_graph_restore copies them back
 * (recompute = 1 also recomputes the distances of every row not in a raw
 * subgraph from the FP32 rows, as InitSGraph does, src/abstractmerge.cpp:55-76;
 * recompute = 2 the same from the bf16 rows where those are exact, which
 * gives the same bits).  The reference's only checkpoint is the text
 * dump of AbstractMerge::saveKGraph (src/abstractmerge.cpp:278-291); this
 * keeps the copy in HBM so repeated merges start without a host transfer. */
int nnm_graph_save(nnm_ctx* ctx);
int nnm_graph_restore(nnm_ctx* ctx, int recompute, nnm_stats* st);

/* --- NN-Descent ------------------------------------------------------- */
/* The NN-Descent loop: list build + reverse sampling + local join + delta
 * test, until delta or max_iters.  Replaces AbstractMerge::nnDescent
 * (src/abstractmerge.cpp:78-192), KGraph::nndescent (src/kgraph.cpp:98-189)
 * and LyNNDescent::nnDescent (src/lynndescent.cpp:201-310), with the pair
 * rule of the matching localJoin. */
int nnm_nndescent(nnm_ctx* ctx, const nnm_params* p, nnm_stats* st);

/* One iteration of the loop above (list build, localJoin, delta).  The
 * per-iteration counterpart of AbstractMerge::localJoin
 * (src/abstractmerge.h:30). */
int nnm_local_join(nnm_ctx* ctx, const nnm_params* p, uint32_t iter, nnm_stats* st);

/* --- merges ----------------------------------------------------------- */
/* Merge set-up over the current partition.  Rows of built subgraphs park
 * their far `cut` entries and take `cut` fresh random samples from the other
 * subgraphs of their group (Smerge::InitialGraph src/smerge.cpp:94-209,
 * Jmerge::InitialGraph src/jmerge.cpp:95-140, LyNNDescent::appdNwSmpl
 * src/lynndescent.cpp:371-415); raw rows get K random neighbours over their
 * group (src/jmerge.cpp:147-171, src/lynndescent.cpp:421-445).  Requires a
 * graph for the built rows (nnm_set_graph or an earlier build). */
int nnm_merge_init(nnm_ctx* ctx, const nnm_params* p, nnm_stats* st);

/* Merge the parked entries back: AbstractMerge::merge2one
 * (src/abstractmerge.cpp:232-277), LyNNDescent::merge2one
 * (src/lynndescent.cpp:455-500).  dedup = 1 drops repeated ids (the
 * reference keeps them). */
int nnm_merge_finish(nnm_ctx* ctx, int dedup);

/* --- brute force ------------------------------------------------------ */
/* Exact k-NN (FP32 difference form, ties broken by id) of nq query rows
 * against all rows.  Q == NULL: the queries are rows `rows[0..nq)` of the
 * data and exclude_self drops each query's own row.  Q != NULL: nq x d query
 * vectors.  Output nq x k.  The reference reads this from a precomputed
 * ground-truth file (Evaluator::loadGrdKNNGraph, src/evaluator.cpp:343-378). */
int nnm_bruteforce_knn(nnm_ctx* ctx, const float* Q, const uint32_t* rows, uint64_t nq,
                       uint32_t k, int exclude_self, uint32_t* ids, float* dist);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid allocation): nnm_bruteforce_knn with every buffer on the device and
 * no allocation of any kind.  The nq query rows are row numbers of the
 * context's data (rows_dev, device memory; a row past the data gets ids
 * 0xFFFFFFFF and infinite distances), excluding each query's own row when
 * exclude_self is set; the nq x k results go to ids_dev / dist_dev (device
 * memory).  The per-range partial lists live in a fixed 16 MiB device array
 * per GPU, so nq is at most 65536 (larger requests return NNM_ERR_INVALID).
 * Calls that overlap on one GPU race on that array.  Returns after the
 * results are written. */
int nnm_bruteforce_knn_noalloc(nnm_ctx* ctx, const uint32_t* rows_dev, uint64_t nq, uint32_t k, int exclude_self,
                               uint32_t* ids_dev, float* dist_dev);

/* Squared L2 between row pairs (a[i], b[i]) in the library's FP32 order:
 * the DISTFUNC call of src/abstractgraph.h:11-12 batched. */
int nnm_pair_dists(nnm_ctx* ctx, const uint32_t* a, const uint32_t* b, uint64_t m, float* out);

/* --- learned candidate pruning ---------------------------------------- */
/* The reference has no learned component; these entry points serve the
 * NNM_PREC_LEARNED join.  A model f(x) = W2 relu(W1 a x + b1) + b2
 * (d -> hidden -> d_out, FP32 master weights) is trained on the GPU so that
 * |f(x) - f(y)|^2 ~ a^2 |x - y|^2 on pairs taken from the current lists (list
 * neighbours and two-hop candidates, the pairs the local join evaluates).
 * nnm_mlp_init: He-initialised model, input scale a from the data.
 * nnm_mlp_train: `steps` Adam steps on `batch` fresh pairs each; afterwards
 *   tau (read back by nnm_mlp_get_weights) is the `keep` quantile of p / kth
 *   over true improvements of a fresh sample, a diagnostic of how well the
 *   model ranks near the k-th distance.  The learned join's offer threshold
 *   is nnm_params.learn_tau, not tau.
 * nnm_mlp_loss_grad: loss and gradients of one batch given by the caller
 *   (rows already scaled; bf16 = 0 uses FP32 GEMMs), for checking.
 * nnm_mlp_project: projection of every row (used by the learned join);
 *   out (n x d_out floats) may be NULL.
 * nnm_mlp_project_fp8: the FP8 rows of NNM_FLAG_LEARNED_FP8 (made from the
 *   current projection when missing): out (n x d_out bytes, OCP e4m3 of
 *   clamp(s p, +-448), round to nearest even) may be NULL; *scale = s. */
int nnm_mlp_init(nnm_ctx* ctx, uint32_t hidden, uint32_t d_out, uint64_t seed);
int nnm_mlp_set_weights(nnm_ctx* ctx, const float* W1, const float* b1, const float* W2, const float* b2,
                        float alpha);
int nnm_mlp_get_weights(nnm_ctx* ctx, float* W1, float* b1, float* W2, float* b2, float* alpha, float* tau);
int nnm_mlp_loss_grad(nnm_ctx* ctx, const float* xa, const float* xb, const float* t, const float* s, uint32_t B,
                      int bf16, float* loss, float* gW1, float* gb1, float* gW2, float* gb2);
int nnm_mlp_train(nnm_ctx* ctx, uint32_t steps, uint32_t batch, float lr, float keep, uint64_t seed,
                  float* loss_first, float* loss_last, double* secs);
int nnm_mlp_project(nnm_ctx* ctx, float* out, double* secs);
int nnm_mlp_project_fp8(nnm_ctx* ctx, uint8_t* out, float* scale);

/* --- host-class drivers ----------------------------------------------- */
/* Runs one host class end to end (nn-merge_amd/host): kind 0 = KGraph
 * (KGraph::buildKGraph, src/kgraph.cpp:217-245), 1 = Smerge
 * (Smerge::buildKGraph, src/smerge.cpp:46-91), 2 = Jmerge
 * (Jmerge::buildKGraph, src/jmerge.cpp:43-80), 3 = P-way merge of `parts`
 * equal partitions.  The merges build their partitions first (all in one
 * launch, not timed, as the reference does not time them) and then run the
 * timed merge.  Outputs n x K ids (and dists when dist != NULL); st gets the
 * merge's counters (dists = the reference's nCmps, secs = device time of
 * the timed region); build_secs the device time of the partition builds.
 * Replaces the test() drivers' buildKGraph calls (src/smerge.cpp:238-261,
 * src/jmerge.cpp:205-224, src/kgraph.cpp:263-285). */
int nnm_host_build(int kind, const float* X, uint64_t n, uint32_t d, const nnm_params* p, uint32_t parts,
                   uint32_t* ids, float* dist, nnm_stats* st, double* build_secs);

/* Hierarchical construction through the LyNNDescent host class.
 * leaf_begin == NULL: the reference's doubling schedule,
 * LyNNDescent::buildhgraph (src/lynndescent.cpp:735-831): NN-Descent over the
 * first N0 = 64 rows, then blocks as large as the current graph J-Merged in
 * (appdNwSmpl + nnDescent(verse = 0) + merge2one).  Otherwise nleaves
 * contiguous leaves [leaf_begin[i], leaf_begin[i+1]) are built together and
 * merged pairwise level by level, every pair of a level in one launch per
 * iteration.  Per-level rows covered, merges run together, device seconds
 * and distances go to the level_* arrays (up to max_levels entries). */
int nnm_host_hmerge(const float* X, uint64_t n, uint32_t d, const nnm_params* p, uint32_t nleaves,
                    const uint32_t* leaf_begin, uint32_t* ids, float* dist, nnm_stats* st, uint32_t* level_rows,
                    uint32_t* level_groups, double* level_secs, uint64_t* level_dists, uint32_t max_levels,
                    uint32_t* nlevels);

/* LyNNDescent::buildhgraph with its search layers (src/lynndescent.cpp:735-831):
 * the doubling schedule of nnm_host_hmerge, plus a DPG layer
 * (LyNNDescent::dpg, src/lynndescent.cpp:502-540: occlusion bound K/2,
 * reverse bound M on the base layer and M/2 above, lists carried from layer
 * to layer) after the seed graph, after the levels that the reference's
 * level[] table marks (3, 6 and 9) and after the last level.  dst_prefix !=
 * NULL writes the layers to dst_prefix + ".txt" (upper layer first, each
 * "N" then "i deg ids...") and the final k-NN graph to dst_prefix +
 * "_baselayer.txt", as the reference does.  layer_rows gets each layer's row
 * count (up to max_layers); st->secs excludes the layer builds, as the
 * reference's construction time does, and dpg_secs reports them. */
int nnm_host_hgraph(const float* X, uint64_t n, uint32_t d, const nnm_params* p, uint32_t M, const char* dst_prefix,
                    uint32_t* ids, float* dist, nnm_stats* st, uint32_t* layer_rows, uint32_t max_layers,
                    uint32_t* nlayers, double* dpg_secs);

/* LySearch through its host class: LySearch(refer, graph) loads the
 * reference set and the layered graph file (src/lysearch.cpp:32-39,
 * :112-186), then knnSearch(query, dst, grd, record) (src/lysearch.cpp:363-437)
 * runs the reference's beam list {60, 50, ..., 2} from entry row `entry` of
 * layer 0 as one batched launch, writes one record line per beam
 * ("topk\trecall@1\tsecs\tnCmps/nQry", stopping after the first beam with
 * recall 1) and leaves the last beam's top-1 results in dst.  prec:
 * NNM_PREC_FP32 or NNM_PREC_BF16.  The records also come back in topk /
 * recall / secs / cmps (up to max_records; nrecords = lines written). */
int nnm_host_lysearch(const char* refer, const char* graph, const char* query, const char* dst, const char* grd,
                      const char* record, uint32_t entry, uint32_t prec, uint32_t max_records, uint32_t* nrecords,
                      uint32_t* topk, double* recall, double* secs, uint64_t* cmps);

/* Per-result FP64 error of the graph's FP32 distances: every list entry of
 * the active rows is recomputed in FP64 on the GPU.  No reference
 * counterpart (the reference stores FP32 distances only,
 * src/pvetex.h:18-30); the chosen precision reports it. */
int nnm_graph_dist_error(nnm_ctx* ctx, nnm_dist_error* out);

/* ------------------------------------------------------------ DPG layers */

/* Build a DPG layer (LyNNDescent::dpg + DelaunayGraph,
 * src/lynndescent.cpp:502-540, 640-710) from the context's k-NN lists over
 * the active rows: M is the occlusion list bound (the reference's M_ =
 * K0/2), maxM the reverse-edge bound (_M on the base layer, _M/2 above).
 * carry = 1 starts rows present in the previous nnm_dpg_build from that
 * layer's lists (the reference keeps LyNNDescent::graph across layers).
 * The reverse pass takes every offer from the occlusion-pruned lists in
 * increasing source order per target (order-free restatement of the
 * reference's in-place pass).  st->dists counts the distances computed. */
int nnm_dpg_build(nnm_ctx* ctx, uint32_t M, uint32_t maxM, int carry, nnm_stats* st);

/* The last DPG layer as CSR: offs[rows + 1] (rows = active rows at build
 * time), adj[offs[rows]] and dist[offs[rows]] (each list sorted by
 * distance; adj/dist may be null).  The reference writes the same lists as
 * text, "N" then "i deg ids..." (LyNNDescent::saveDPGraph,
 * src/lynndescent.cpp:625-638). */
int nnm_dpg_get(nnm_ctx* ctx, uint32_t* offs, uint32_t* adj, float* dist);

/* ------------------------------------------------------------ search */

#define NNM_SEARCH_LAYERED 0 /* LySearch::knnSearch: greedy through upper layers, best-first on the last */
#define NNM_SEARCH_FLAT 1    /* LySearch::nndFlatSrch: best-first on the last layer from seeds */
#define NNM_SEARCH_SRC_KNN 0
#define NNM_SEARCH_SRC_DPG 1

typedef struct nnm_search_params {
    uint32_t mode;   /* NNM_SEARCH_LAYERED or NNM_SEARCH_FLAT */
    uint32_t beam;   /* result-list size, the reference's topk (<= 256) */
    uint32_t k;      /* results per query (<= min(beam, 64)) */
    uint32_t width;  /* list entries expanded per step (1..4; 1 = the reference's order) */
    uint32_t prec;   /* NNM_PREC_FP32, or NNM_PREC_BF16: bf16 rows during the search,
                        the first `rerank` list entries recomputed in FP32 at the end */
    uint32_t rerank; /* 0 = min(beam, 64) */
    uint32_t entry;  /* layered mode: entry row of layer 0 (the reference draws
                        rand() % rows, src/lysearch.cpp:378) */
    uint32_t pad_;
    uint64_t seed;   /* flat mode without explicit seeds: seeds drawn from it */
} nnm_search_params;

typedef struct nnm_search_stats {
    uint64_t dists;       /* distances computed (the reference's nCmps, src/lysearch.cpp:281) */
    uint64_t reranked;    /* FP32 re-rank distances (bf16 mode) */
    uint64_t hash_resets; /* visited-hash refills (revisited rows are recomputed) */
    uint64_t capped;      /* queries stopped by the step cap (an error) */
    uint64_t steps;       /* list entries expanded, all queries */
    double secs;          /* search kernel time (a batched launch: this beam's share, by wave time) */
    double wave_secs;     /* summed per-query wave time on the device wall clock */
    double launch_secs;   /* time of the whole launch (all beams of a batched launch) */
} nnm_search_stats;

/* Search layers as CSR (the reference's augKnnGraph + knnLocs,
 * src/lysearch.cpp:112-186): layer g covers rows [0, rows[g]), offs[g] has
 * rows[g] + 1 entries, adj[g] the neighbour ids.  Layers grow (rows[g] <=
 * rows[g + 1]); the last one is the base layer. */
int nnm_search_set_layers(nnm_ctx* ctx, uint32_t nlayers, const uint32_t* rows, const uint32_t* const* offs,
                          const uint32_t* const* adj);

/* The context's own graph as the single search layer: NNM_SEARCH_SRC_DPG
 * (the last nnm_dpg_build) or NNM_SEARCH_SRC_KNN (the k-NN lists). */
int nnm_search_use_graph(nnm_ctx* ctx, int source);

/* Batched search of nq queries (nq x d float32, host): LySearch::knnSearch
 * (src/lysearch.cpp:363-437) + nndHNSWSrch (src/lysearch.cpp:236-322), or
 * nndFlatSrch (src/lysearch.cpp:439-544) in flat mode with seeds (nq x
 * beam, may be null).  Writes the first k results per query sorted by
 * (dist, id) into ids / dist (nq x k); missing results are 0xFFFFFFFF. */
int nnm_search(nnm_ctx* ctx, const float* Q, uint32_t nq, const nnm_search_params* p, const uint32_t* seeds,
               uint32_t* ids, float* dist, nnm_search_stats* st);

/* Every beam of a beam list in one launch: the reference's beam sweep in
 * LySearch::knnSearch (for topk in Ks1, src/lysearch.cpp:380-384), one
 * search per (beam, query) from the same entry row.  nbeams in [1, 32];
 * p->beam is ignored and p->k must not exceed the smallest beam; no explicit
 * seeds (flat mode draws them from p->seed).  ids / dist: nbeams x nq x k in
 * the order of `beams`; st: nbeams records, the same as nnm_search gives for
 * each beam alone, with secs the beam's share of the launch. */
int nnm_search_beams(nnm_ctx* ctx, const float* Q, uint32_t nq, const nnm_search_params* p, const uint32_t* beams,
                     uint32_t nbeams, uint32_t* ids, float* dist, nnm_search_stats* st);

/* ------------------------------------------------------------ file formats */
/* Host-side readers and writers of the reference's files (no device work).
 * Readers are two-phase (shape, then read into caller memory) and return
 * NNM_ERR_INVALID for unreadable or malformed files. */

/* Data matrix by extension: .txt "row col" then values
 * (IODelegator::loadMatrix, src/iodelegator.cpp:186-220); .fvecs records
 * [int32 d][d float32] (load_fvecs, :222-265); .bvecs records [int32 d][d
 * uint8] as floats, at most 10M rows, empty records skipped (loadBvec10m,
 * :319-385). */
int nnm_io_matrix_shape(const char* path, uint64_t* rows, uint32_t* dim);
int nnm_io_read_matrix(const char* path, float* out, uint64_t rows, uint32_t dim);
int nnm_io_write_matrix(const char* path, const float* X, uint64_t rows, uint32_t dim);

/* k-NN graph text "N K" then "i k id..." per row (IODelegator::loadKNNGraph,
 * src/iodelegator.cpp:65-102; AbstractMerge::saveKGraph,
 * src/abstractmerge.cpp:278-291).  nnm_io_read_knn keeps the first topk ids
 * of each row. */
int nnm_io_knn_shape(const char* path, uint64_t* rows, uint32_t* k);
int nnm_io_read_knn(const char* path, uint32_t* ids, uint64_t rows, uint32_t topk);
int nnm_io_write_knn(const char* path, const uint32_t* ids, uint64_t rows, uint32_t k);

/* DPG / layered search graph text: blocks of "N" then "i deg id..." per row
 * (LyNNDescent::saveDPGraph, src/lynndescent.cpp:625-638;
 * IODelegator::loadDPGraph, src/iodelegator.cpp:1173-1217), upper layer
 * first in a layered file (LySearch::loadGraphs, src/lysearch.cpp:112-186).
 * rows / edges: per block; offs[g] holds rows[g] + 1 entries. */
int nnm_io_layers_shape(const char* path, uint32_t max_layers, uint32_t* nlayers, uint32_t* rows, uint64_t* edges);
int nnm_io_read_layers(const char* path, uint32_t nlayers, uint32_t* const* offs, uint32_t* const* adj);
int nnm_io_write_layer(const char* path, int append, const uint32_t* offs, const uint32_t* adj, uint32_t rows);

#ifdef __cplusplus
}
#endif

#endif /* NNMERGE_H */

// Internal state and device helpers of the gfx950 NN-Descent library.
//
// Device layout (structure of arrays, row-major per vertex):
//   X      n x ld  float32, rows zero-padded to ld (multiple of 64) so every
//                  row starts on a 256-byte boundary and padding adds 0 to L2
//   g_ids  n x K   uint32  neighbour ids, each row sorted by (dist, id)
//   g_dist n x K   float32 squared L2
//   g_flag n x K   uint8   1 = entry is new (not yet joined)
//   label  n       uint32  subgraph label: s for rows of built subgraph s,
//                          NNM_RAW for raw rows
//   offer log      records {key, target} the join writes in order,
//                  grouped by target by partition passes before the owners
//                  merge them into their lists (nnm_offers.hip)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nnmerge.h"

#define NNM_WAVE 64
#define NNM_MAX_K 64
// Learned-phase offer threshold (fraction of the target's estimated k-th distance).
#define NNM_LEARN_TAU_DEFAULT 0.6f

// Candidate keys: squared distances are >= 0, so their IEEE bits order like
// the values; the low word breaks ties by id.  All ones marks no candidate.
#define NNM_EMPTY_KEY 0xFFFFFFFFFFFFFFFFull

// Offer log: records {key = dist bits << 32 | id, target} in two arrays,
// written by each join wave into chunks of NNM_OCH records it reserves from
// a shared counter.
#define NNM_OCH 8192u
// Owner bins: the partition passes group offers by target >> NNM_OWN_BITS.
#define NNM_OWN_BITS 5u
// Default offer-log capacity per active row (nnm_params.offer_cap = 0).
#define NNM_OFFER_CAP_DEFAULT 256u

// Reverse-list bucket capacity per target (entries beyond it go to the
// overflow CSR).
#define NNM_RB 64u

enum CounterIdx {
    CNT_DISTS = 0,
    CNT_UPDATES = 1,
    CNT_FLAGS = 2,
    CNT_RERANKED = 3,
    CNT_LIST_ENTRIES = 4,
    CNT_WORK = 5,      // work-queue cursor of the persistent join
    CNT_TOTAL_REV = 6,
    CNT_OFFERS = 7,    // offers the join wrote to the offer log
    CNT_WORK2 = 8,     // work-queue cursor of the long-list join
    CNT_WORK3 = 9,     // work-queue cursor of the window join
    CNT_ITER = 10,     // iteration number (salts derive from it on the device)
    CNT_LONG_DISTS = 11,    // share of DISTS / LIST_ENTRIES / OFFERS that the
    CNT_LONG_ENTRIES = 12,  // long-list join kernels account for (the rest is
    CNT_LONG_OFFERS = 13,   // the short-list kernel's)
    CNT_NEAR_TIES = 14,     // k-th entry decisions re-made in FP64
    CNT_TIE_SWAPS = 15,     // ... that changed which entry stays
    CNT_OCHUNKS = 16,       // offer-log chunks reserved in this round
    CNT_REDO = 17,          // vertices whose offers did not fit the log (joined again)
    CNT_ROUNDS = 18,        // extra join rounds run for such vertices
    CNT_N = 24
};

struct nnm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evj0 = nullptr, evj1 = nullptr, evjm = nullptr, evd1 = nullptr;
    std::string err;
    uint64_t mem_cap = 0;
    uint64_t mem_used = 0;
    int num_cus = 256;

    // data
    uint64_t n = 0;       // rows held
    uint64_t n_act = 0;   // rows covered by the partition (a prefix of the rows)
    uint32_t d = 0, ld = 0;
    float* X = nullptr;
    uint16_t* Xb = nullptr;   // bf16 copy (lazy)
    float* norms = nullptr;   // squared norms of the bf16 rows (lazy)
    int data_exact_bf16 = 0;  // every value an integer in [-256, 256]
    int data_i8 = 0;          // exact, rows of 128, values spanning at most 255: int8 rows exact
    float i8_shift = 0.f;     // the int8 rows hold value - i8_shift (in [-128, 127])
    int8_t* Xi8 = nullptr;    // int8 copy (lazy; padding 0)
    float* norms_i8 = nullptr;  // squared norms of the int8 rows (exact integers)
    float data_min = 0.f, data_max = 0.f;  // value range of the rows (with the exactness scan)
    uint32_t metric = NNM_METRIC_L2;
    float* sumsq = nullptr;   // cosine: squared row norms (with the data)

    // partition
    uint32_t nsub = 0, ngrp = 0;
    std::vector<uint32_t> h_sub_begin, h_sub_group, h_grp_first_sub;
    std::vector<uint8_t> h_sub_built;
    uint32_t* label = nullptr;      // n
    uint32_t* vsub = nullptr;       // n: subgraph of each row
    uint32_t* sub_begin = nullptr;  // nsub + 1
    uint32_t* sub_grp = nullptr;    // nsub
    uint8_t* sub_built = nullptr;   // nsub
    uint32_t* grp_begin = nullptr;  // ngrp
    uint32_t* grp_end = nullptr;    // ngrp

    // graph
    uint32_t K = 0;
    uint32_t* g_ids = nullptr;
    float* g_dist = nullptr;
    uint8_t* g_flag = nullptr;
    uint32_t* s_ids = nullptr;      // checkpoint copy
    float* s_dist = nullptr;
    uint8_t* s_flag = nullptr;
    uint32_t* p_ids = nullptr;      // lists at the learned phase's start (NNM_FLAG_LEARNED_MARK_FOUND)
    float* p_dist = nullptr;        // ... and their exact distances

    // parked far entries of a merge (cpnGraph)
    uint32_t cut = 0;
    uint32_t* c_ids = nullptr;
    float* c_dist = nullptr;
    uint8_t* c_flag = nullptr;
    uint8_t* has_cpn = nullptr;     // n
    int cpn_live = 0;

    // iteration scratch
    uint32_t cap_new = 0, cap_old = 0, rev_cap = 0;
    uint32_t* l_new = nullptr;      // n x cap_new
    uint32_t* l_old = nullptr;      // n x cap_old
    uint32_t* n_new = nullptr;      // n
    uint32_t* n_old = nullptr;      // n
    uint64_t* newmask = nullptr;    // n: list entries taken as new samples
    uint32_t* rcnt = nullptr;       // 2n reverse counts (new, old)
    uint32_t* rbuf = nullptr;       // 2n x NNM_RB reverse buckets
    uint32_t* ocnt = nullptr;       // 2n overflow sizes (count past NNM_RB)
    uint32_t* roff = nullptr;       // 2n + 1 overflow offsets
    uint32_t* rcur = nullptr;       // 2n overflow cursors
    uint32_t* obuf = nullptr;       // n*K overflow entries, grouped by bucket
    uint64_t* spill = nullptr;      // n: per row, the list entries past a full bucket (bit j: entry j)
    uint32_t* scan_tmp = nullptr;   // block sums
    // offer log and its partition (nnm_offers.hip)
    uint64_t* okey[2] = {nullptr, nullptr};  // 2 x (ochunks x NNM_OCH) record keys (dist bits << 32 | id)
    uint32_t* otgt[2] = {nullptr, nullptr};  // ... and their targets
    uint32_t* ochunk_used = nullptr;  // records written per log chunk
    uint32_t ochunks = 0;             // log capacity in chunks
    uint32_t ocap_row = 0;            // capacity per row it was sized for
    uint32_t* opart_cnt = nullptr;    // per pass: bin counts (largest pass)
    uint32_t* opart_cur[3] = {nullptr, nullptr, nullptr};  // per pass: bin cursors (bins + 1)
    uint32_t* work[2] = {nullptr, nullptr};  // join work lists for redo rounds (n each)
    uint32_t* opart_blk = nullptr;    // first pass: per-block bin counts (bins x grid)
    uint32_t* opart_boff = nullptr;   // ... their exclusive scan
    uint32_t* opart_tmp = nullptr;    // scan scratch of the passes
    uint32_t* opart_p1[2] = {nullptr, nullptr};  // second pass: counts per (owner bin, segment), their scan
    uint64_t* counters = nullptr;   // CNT_N
    float* kth = nullptr;           // n: k-th distance of every list at iteration start
    uint4* vmeta = nullptr;         // n: join record {label, kth x offer window, norm, 0}
    uint4* lrec = nullptr;          // n x 128 bytes: learned row (bytes 0..63) + join record (64..79)
    uint64_t lrec_n = 0;            // rows lrec holds
    uint64_t lrec_gen = 0;          // projection generation its rows came from (0: none)
    int lrec_live = 0;              // this iteration's learned join reads lrec
    uint32_t* grp_flags = nullptr;  // ngrp: per-group new-flag counts
    uint8_t* grp_active = nullptr;  // ngrp

    void* mlp = nullptr;            // learned projection state (nnm_mlp.hip)
    void* search = nullptr;         // search layers (nnm_search.hip)
    void* dpg = nullptr;            // DPG layer lists (nnm_dpg.hip)

    // join launch configuration (autotuned): window rows of the short-list
    // kernels and rows taken from the work queue at a time
    // per precision (NNM_PREC_*); tune_ld = 0 means not tuned yet
    uint32_t tune_js[3] = {0, 0, 0}, tune_chunk[3] = {0, 0, 0}, tune_ld[3] = {0, 0, 0};

    std::vector<std::pair<void*, size_t>> allocs;

    // pinned staging buffers for host <-> device transfers (see nnm_h2d)
    void* stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

// ------------------------------------------------------------ host helpers

int nnm_fail(nnm_ctx* c, int code, const std::string& msg);
int nnm_hip_fail(nnm_ctx* c, hipError_t e, const char* what);
int nnm_alloc(nnm_ctx* c, void** p, size_t bytes);
void nnm_free(nnm_ctx* c, void* p);
// Host <-> device copies on the context stream through the context's pinned
// staging buffers.  nnm_h2d returns once the source has been read (the copy
// may still be in flight); nnm_d2h returns with the data in dst.
// nnm_h2d_rows uploads n rows of d floats into rows of pitch ld (zero padded).
int nnm_h2d(nnm_ctx* c, void* dst, const void* src, size_t bytes);
int nnm_d2h(nnm_ctx* c, void* dst, const void* src, size_t bytes);
int nnm_h2d_rows(nnm_ctx* c, float* dst, uint32_t ld, const float* src, uint32_t d, uint64_t n);
int nnm_ensure_bf16(nnm_ctx* c);
int nnm_ensure_i8(nnm_ctx* c);
// The plain BF16 joins run on the int8 rows (exact integer Gram on MFMA,
// the same keys bit for bit: every distance is an exact integer below 2^24)
// once the rows exist (made before a descent's iterations).
inline bool nnm_use_i8(const nnm_ctx* c, const nnm_params* p) {
    return p->prec == NNM_PREC_BF16 && c->data_i8 && c->metric == NNM_METRIC_L2 && c->Xi8;
}
// Offer window above the k-th distance: 1 + 2 * (d + 2) * 2^-24 (twice the
// relative error bound of an FP32 squared distance of d terms), so that
// candidates within rounding of the k-th entry reach the FP64 tie test; 1
// where distances are exact (integer rows) and for learned estimates.
float nnm_near_tie_window(nnm_ctx* c, const nnm_params* p);
int nnm_require_l2(nnm_ctx* c, const char* what);
int nnm_ensure_iter_buffers(nnm_ctx* c, const nnm_params* p);

#define NNM_HIP(c, expr)                                  \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return nnm_hip_fail((c), _e, #expr); \
    } while (0)

#define NNM_LAUNCH_CHECK(c)                                  \
    do {                                                     \
        hipError_t _e = hipGetLastError();                   \
        if (_e != hipSuccess) return nnm_hip_fail((c), _e, "kernel launch"); \
    } while (0)

#define NNM_TRY(expr)                \
    do {                             \
        int _r = (expr);             \
        if (_r != NNM_OK) return _r; \
    } while (0)
#define NNM_TRY_MLP NNM_TRY

// Kernel launchers shared between translation units.
// round 0 joins every active row; round r > 0 the nwork vertices of the
// redo list round r - 1 left (work[(r - 1) & 1]).
int nnm_launch_join(nnm_ctx* c, const nnm_params* p, uint32_t js, uint32_t chunk, hipEvent_t mid = nullptr,
                    uint32_t round = 0, uint64_t nwork = 0);
// Longest candidate list (new + old) the learned join handles.
uint32_t nnm_join_learned_max_list();
void nnm_join_tune_space(nnm_ctx* c, uint32_t prec, std::vector<std::pair<uint32_t, uint32_t>>* out);
uint32_t nnm_join_default_js(nnm_ctx* c, uint32_t prec);
uint32_t nnm_join_default_chunk();

// Learned projection (nnm_mlp.hip)
void nnm_mlp_free(nnm_ctx* c);
void nnm_search_free(nnm_ctx* c);
int nnm_search_adopt_layer(nnm_ctx* c, uint32_t rows, uint32_t* offs, uint32_t* adj);
void nnm_dpg_free(nnm_ctx* c);
int nnm_fill_u64(nnm_ctx* c, uint64_t* p, uint64_t n, uint64_t v);
int nnm_fill_u32(nnm_ctx* c, uint32_t* p, uint64_t n, uint32_t v);
// exclusive scan of m u32 counts into out[0..m] (out[m] = total); tmp holds
// (m + 2047) / 2048 + 1 words
int nnm_scan_u32(nnm_ctx* c, const uint32_t* in, uint64_t m, uint32_t* out, uint32_t* tmp);
int nnm_mlp_ensure_projection(nnm_ctx* c, double* secs);
void nnm_mlp_invalidate_projection(nnm_ctx* c);
const uint16_t* nnm_mlp_rows(nnm_ctx* c);
uint64_t nnm_mlp_proj_gen(nnm_ctx* c);
const float* nnm_mlp_norms(nnm_ctx* c);
uint32_t nnm_mlp_dout(nnm_ctx* c);
float nnm_mlp_scale(nnm_ctx* c);
int nnm_mlp_trained(nnm_ctx* c);
// FP8 (OCP e4m3) rows of the projection for the learned join's FP8 prefilter
int nnm_mlp_ensure_fp8(nnm_ctx* c);
const uint8_t* nnm_mlp_rows8(nnm_ctx* c);
const float* nnm_mlp_norms8(nnm_ctx* c);
float nnm_mlp_scale8(nnm_ctx* c);

// ---------------------------------------------------------- device helpers

__host__ __device__ inline uint64_t nnm_mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Counter-based random stream: the value for (seed, a, b) is fixed, so the
// same draws come out whatever the launch shape or thread order.
__host__ __device__ inline uint64_t nnm_rand(uint64_t seed, uint64_t a, uint64_t b) {
    return nnm_mix64(seed ^ nnm_mix64(a * 0xD1B54A32D192ED03ull + b));
}

__host__ __device__ inline uint32_t nnm_hash32(uint32_t x, uint32_t salt) {
    x ^= salt;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

__device__ inline uint64_t nnm_key(float dist, uint32_t id) {
    return ((uint64_t)__float_as_uint(dist) << 32) | id;
}
__device__ inline float nnm_key_dist(uint64_t k) { return __uint_as_float((uint32_t)(k >> 32)); }
__device__ inline uint32_t nnm_key_id(uint64_t k) { return (uint32_t)k; }

__device__ inline uint32_t nnm_lane() { return __lane_id(); }

// Value of lane `src` (wave-uniform) through a scalar register: reads the lane
// whatever the exec mask, unlike a shuffle (whose inactive source lanes read 0).
__device__ inline uint32_t wave_read32(uint32_t x, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, (int)src);
}
__device__ inline uint64_t wave_read64(uint64_t x, uint32_t src) {
    return ((uint64_t)wave_read32((uint32_t)(x >> 32), src) << 32) | wave_read32((uint32_t)x, src);
}

// 64-bit value of lane src (per lane; the wave must be converged).
__device__ inline uint64_t shfl64(uint64_t x, uint32_t src) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)(uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Position of the j-th set bit (from 0) of m; some position in [0, 63] when
// m has j bits or fewer.
__device__ inline uint32_t nth_set_bit(uint64_t m, uint32_t j) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t w = 32; w > 0; w >>= 1) {
        const uint64_t low = m & ((1ull << w) - 1ull);
        const uint32_t c = (uint32_t)__popcll(low);
        if (j >= c) {
            j -= c;
            m >>= w;
            pos += w;
        } else {
            m = low;
        }
    }
    return pos;
}

// The library's FP32 squared L2: elements in order 0..d-1, one accumulator,
// each step acc = fma(a-b, a-b, acc).  Every kernel that produces a stored
// distance uses this order, so one pair always gives the same bits.
// B 16-byte loads of each row are issued before their (in-order) FMAs; any B
// gives the same bits, a larger B hides more latency for more registers.
template <int B = 8>
__device__ inline float nnm_l2_rows(const float* __restrict__ a, const float* __restrict__ b,
                                    uint32_t d) {
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    float acc = 0.f;
    const uint32_t n4 = d / 4;
    uint32_t i = 0;
    for (; i + B <= n4; i += B) {
        float4 x[B], y[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            x[u] = a4[i + u];
            y[u] = b4[i + u];
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            float t;
            t = x[u].x - y[u].x; acc = __builtin_fmaf(t, t, acc);
            t = x[u].y - y[u].y; acc = __builtin_fmaf(t, t, acc);
            t = x[u].z - y[u].z; acc = __builtin_fmaf(t, t, acc);
            t = x[u].w - y[u].w; acc = __builtin_fmaf(t, t, acc);
        }
    }
    for (; i < n4; ++i) {
        float4 x = a4[i], y = b4[i];
        float t;
        t = x.x - y.x; acc = __builtin_fmaf(t, t, acc);
        t = x.y - y.y; acc = __builtin_fmaf(t, t, acc);
        t = x.z - y.z; acc = __builtin_fmaf(t, t, acc);
        t = x.w - y.w; acc = __builtin_fmaf(t, t, acc);
    }
    for (uint32_t j = n4 * 4; j < d; ++j) {
        float t = a[j] - b[j];
        acc = __builtin_fmaf(t, t, acc);
    }
    return acc;
}

// Exact FP32 distance of rows a, b in the canonical order from bf16 rows:
// used where every value is exact in bf16 (integer rows), where it equals
// nnm_l2_rows on the FP32 rows bit for bit while reading half the bytes.
template <int B = 8>
__device__ inline float nnm_l2_rows_bf16(const uint16_t* __restrict__ a, const uint16_t* __restrict__ b, uint32_t ld) {
    const uint4* a4 = reinterpret_cast<const uint4*>(a);
    const uint4* b4 = reinterpret_cast<const uint4*>(b);
    float acc = 0.f;
    const uint32_t n8 = ld / 8;
    auto step = [&](const uint4 x, const uint4 y) {
        const uint32_t xs[4] = {x.x, x.y, x.z, x.w}, ys[4] = {y.x, y.y, y.z, y.w};
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            float t = __uint_as_float(xs[h] << 16) - __uint_as_float(ys[h] << 16);
            acc = __builtin_fmaf(t, t, acc);
            t = __uint_as_float(xs[h] & 0xFFFF0000u) - __uint_as_float(ys[h] & 0xFFFF0000u);
            acc = __builtin_fmaf(t, t, acc);
        }
    };
    uint32_t i = 0;
    for (; i + B <= n8; i += B) {
        uint4 x[B], y[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            x[u] = a4[i + u];
            y[u] = b4[i + u];
        }
#pragma unroll
        for (int u = 0; u < B; ++u) step(x[u], y[u]);
    }
    for (; i < n8; ++i) step(a4[i], b4[i]);
    return acc;
}

// Exact squared L2 distance of int8 rows (value - shift, padding 0):
// sum a^2 + sum b^2 - 2 sum ab in integers by four-way int8 dot products.
// For integer rows of 128 values spanning at most 255 every such distance
// is an integer below 2^24, so the float is exact: the bits of nnm_l2_rows.
__device__ inline float nnm_l2_rows_i8(const int8_t* __restrict__ a, const int8_t* __restrict__ b, uint32_t ld) {
    const uint4* a4 = reinterpret_cast<const uint4*>(a);
    const uint4* b4 = reinterpret_cast<const uint4*>(b);
    int aa = 0, bb = 0, ab = 0;
    const uint32_t n16 = ld / 16;
    for (uint32_t i0 = 0; i0 < n16; i0 += 4) {
        uint4 x[4], y[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t i = min(i0 + u, n16 - 1);  // past the row: its last chunk again, not summed
            x[u] = a4[i];
            y[u] = b4[i];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (i0 + u >= n16) break;
            const int xs[4] = {(int)x[u].x, (int)x[u].y, (int)x[u].z, (int)x[u].w};
            const int ys[4] = {(int)y[u].x, (int)y[u].y, (int)y[u].z, (int)y[u].w};
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                aa = __builtin_amdgcn_sdot4(xs[h], xs[h], aa, false);
                bb = __builtin_amdgcn_sdot4(ys[h], ys[h], bb, false);
                ab = __builtin_amdgcn_sdot4(xs[h], ys[h], ab, false);
            }
        }
    }
    return (float)(aa + bb - 2 * ab);
}

// Row distance from the bf16 rows when given (values exact in bf16), else
// from the FP32 rows: the same bits either way.
__device__ inline float nnm_row_dist(const float* __restrict__ X, const uint16_t* __restrict__ Xb, uint32_t ld,
                                     uint64_t a, uint64_t b) {
    return Xb ? nnm_l2_rows_bf16(Xb + a * ld, Xb + b * ld, ld) : nnm_l2_rows(X + a * ld, X + b * ld, ld);
}

// ---- the other metrics: DistOpt's operation order (src/distopt.cpp), one
// accumulator over elements 0..d-1, every multiply, divide and add rounded on
// its own (no contraction).  Rows are zero padded to d = ld; a padded element
// adds an exact zero (L1, cosine) or is skipped (chi-square), so the padded
// loops give the unpadded values bit for bit.

// DistOpt::l1norm, src/distopt.cpp:82-110.
__device__ inline float nnm_l1_rows(const float* __restrict__ a, const float* __restrict__ b, uint32_t d) {
#pragma clang fp contract(off)
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    float res = 0.f;
    for (uint32_t i = 0; i < d / 4; ++i) {
        const float4 x = a4[i], y = b4[i];
        res = res + fabsf(x.x - y.x);
        res = res + fabsf(x.y - y.y);
        res = res + fabsf(x.z - y.z);
        res = res + fabsf(x.w - y.w);
    }
    return res;
}

// DistOpt::KSquare, src/distopt.cpp:163-184: (a-b)^2 / (a+b) where a or b is
// nonzero.
__device__ inline float nnm_chi2_step(float w2, float x, float y) {
#pragma clang fp contract(off)
    const float w1 = x - y;
    return (x != 0.f || y != 0.f) ? w2 + (w1 * w1) / (x + y) : w2;
}
__device__ inline float nnm_chi2_rows(const float* __restrict__ a, const float* __restrict__ b, uint32_t d) {
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    float w2 = 0.f;
    for (uint32_t i = 0; i < d / 4; ++i) {
        const float4 x = a4[i], y = b4[i];
        w2 = nnm_chi2_step(w2, x.x, y.x);
        w2 = nnm_chi2_step(w2, x.y, y.y);
        w2 = nnm_chi2_step(w2, x.z, y.z);
        w2 = nnm_chi2_step(w2, x.w, y.w);
    }
    return w2;
}

// DistOpt::cosine, src/distopt.cpp:113-160, split in two: the dot product
// res and the squared norms l1, l2 are three separate accumulators there, so
// the norms are computed once per row (nnm_sumsq_rows) and combined with the
// dot product by nnm_cos_finish.
__device__ inline float nnm_dot_rows(const float* __restrict__ a, const float* __restrict__ b, uint32_t d) {
#pragma clang fp contract(off)
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* b4 = reinterpret_cast<const float4*>(b);
    float res = 0.f;
    for (uint32_t i = 0; i < d / 4; ++i) {
        const float4 x = a4[i], y = b4[i];
        res = res + x.x * y.x;
        res = res + x.y * y.y;
        res = res + x.z * y.z;
        res = res + x.w * y.w;
    }
    return res;
}
__device__ inline float nnm_sumsq_rows(const float* __restrict__ a, uint32_t d) { return nnm_dot_rows(a, a, d); }
__device__ inline float nnm_cos_finish(float dot, float l1, float l2) {
#pragma clang fp contract(off)
    const float r = (l1 != 0.f && l2 != 0.f) ? dot / (sqrtf(l1) * sqrtf(l2)) : 0.f;
    return 2.f - 2.f * r;
}

// Metric of a distance call: which function, and for cosine the squared row
// norms (sumsq, one per row).
struct nnm_metric_args {
    uint32_t metric;
    const float* sumsq;
};
inline nnm_metric_args metric_args(const nnm_ctx* c) { return nnm_metric_args{c->metric, c->sumsq}; }

__device__ inline float nnm_metric_rows(uint32_t metric, const float* __restrict__ a, const float* __restrict__ b,
                                        uint32_t d, float sa, float sb) {
    switch (metric) {
        case NNM_METRIC_L1: return nnm_l1_rows(a, b, d);
        case NNM_METRIC_COSINE: return nnm_cos_finish(nnm_dot_rows(a, b, d), sa, sb);
        case NNM_METRIC_CHI2: return nnm_chi2_rows(a, b, d);
        default: return nnm_l2_rows(a, b, d);
    }
}

// One element of a distance accumulation under metric M (the library's L2
// with fused multiply-add; the others with every operation rounded on its
// own, in the reference's order).
__device__ inline float step_l1(float acc, float x, float y) {
#pragma clang fp contract(off)
    return acc + fabsf(x - y);
}
__device__ inline float step_dot(float acc, float x, float y) {
#pragma clang fp contract(off)
    return acc + x * y;
}
template <uint32_t M>
__device__ inline float metric_step(float acc, float x, float y) {
    if constexpr (M == NNM_METRIC_L1) {
        return step_l1(acc, x, y);
    } else if constexpr (M == NNM_METRIC_COSINE) {
        return step_dot(acc, x, y);
    } else if constexpr (M == NNM_METRIC_CHI2) {
        return nnm_chi2_step(acc, x, y);
    } else {
        const float t = x - y;
        return __builtin_fmaf(t, t, acc);
    }
}

// Distance of rows a, b of X under the call's metric (bf16 rows only for L2).
__device__ inline float nnm_row_dist(const float* __restrict__ X, const uint16_t* __restrict__ Xb, uint32_t ld,
                                     uint64_t a, uint64_t b, const nnm_metric_args& m) {
    if (m.metric == NNM_METRIC_L2) return nnm_row_dist(X, Xb, ld, a, b);
    return nnm_metric_rows(m.metric, X + a * ld, X + b * ld, ld, m.sumsq ? m.sumsq[a] : 0.f,
                           m.sumsq ? m.sumsq[b] : 0.f);
}

// Value of lane (lane ^ J) for J in {1, 2, 4, 8, 16, 32}, without the LDS
// crossbar: quad permutes for 1 and 2, row shifts for 4 and 8 (a lane takes
// lane + J or lane - J inside its 16-lane row), lane swaps across rows for
// 16 and halves for 32 (v_permlane16_swap / v_permlane32_swap).
template <uint32_t J>
__device__ inline uint32_t wave_xor32(uint32_t x) {
    const uint32_t lane = nnm_lane();
    if constexpr (J == 1) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    } else if constexpr (J == 2) {
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    } else if constexpr (J == 4) {
        const uint32_t up = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x104, 0xF, 0xF, false);  // row_shl:4
        const uint32_t dn = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, false);  // row_shr:4
        return (lane & 4u) ? dn : up;
    } else if constexpr (J == 8) {
        const uint32_t up = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x108, 0xF, 0xF, false);  // row_shl:8
        const uint32_t dn = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, false);  // row_shr:8
        return (lane & 8u) ? dn : up;
    } else if constexpr (J == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
        return (lane & 16u) ? (uint32_t)r[0] : (uint32_t)r[1];
    } else {
        static_assert(J == 32, "lane xor distance");
        const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        return (lane & 32u) ? (uint32_t)r[0] : (uint32_t)r[1];
    }
}
template <uint32_t J>
__device__ inline uint64_t wave_xor64(uint64_t x) {
    return ((uint64_t)wave_xor32<J>((uint32_t)(x >> 32)) << 32) | wave_xor32<J>((uint32_t)x);
}

// One compare-exchange step of the bitonic network (block size KB, distance J).
template <uint32_t KB, uint32_t J, typename T>
__device__ inline T bitonic_step(T key, T other) {
    const uint32_t lane = nnm_lane();
    const bool up = ((lane & KB) == 0);
    const bool lower = ((lane & J) == 0);
    const T mn = key < other ? key : other;
    const T mx = key < other ? other : key;
    return (lower == up) ? mn : mx;
}

template <uint32_t KB, uint32_t J>
__device__ inline uint64_t bitonic_pass64(uint64_t key) {
    key = bitonic_step<KB, J>(key, wave_xor64<J>(key));
    if constexpr (J > 1) return bitonic_pass64<KB, J / 2>(key);
    else return key;
}
template <uint32_t KB, uint32_t J>
__device__ inline uint32_t bitonic_pass32(uint32_t key) {
    key = bitonic_step<KB, J>(key, wave_xor32<J>(key));
    if constexpr (J > 1) return bitonic_pass32<KB, J / 2>(key);
    else return key;
}

// Wave-wide bitonic sort of one 64-bit key per lane, ascending by lane.
__device__ inline uint64_t nnm_wave_sort64(uint64_t key) {
    key = bitonic_pass64<2, 1>(key);
    key = bitonic_pass64<4, 2>(key);
    key = bitonic_pass64<8, 4>(key);
    key = bitonic_pass64<16, 8>(key);
    key = bitonic_pass64<32, 16>(key);
    return bitonic_pass64<64, 32>(key);
}

__device__ inline uint32_t nnm_wave_sort32(uint32_t key) {
    key = bitonic_pass32<2, 1>(key);
    key = bitonic_pass32<4, 2>(key);
    key = bitonic_pass32<8, 4>(key);
    key = bitonic_pass32<16, 8>(key);
    key = bitonic_pass32<32, 16>(key);
    return bitonic_pass32<64, 32>(key);
}

__device__ inline uint32_t nnm_wave_sum32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, NNM_WAVE);
    return v;
}

__device__ inline uint64_t nnm_ballot(bool p) { return __ballot(p); }

__device__ inline uint32_t nnm_mask_rank(uint64_t mask) {
    // number of set bits of mask below this lane
    const uint32_t lane = nnm_lane();
    uint64_t below = lane ? (mask & ((~0ull) >> (64 - lane))) : 0ull;
    return (uint32_t)__popcll(below);
}

// Distinct random values in [0, M) for lanes [0, k), k <= 64, k <= M: the
// draw of AbstractGraph::getRndSds (src/abstractgraph.cpp:7-35) -- k values
// in [0, M-k), sorted, bumped apart, rotated by a random offset -- computed
// across the wave (sort, prefix max, rotate).  Lanes >= k return M.
__device__ inline uint32_t nnm_wave_rnd_sds(uint64_t seed, uint64_t stream, uint32_t k, uint32_t M) {
    const uint32_t lane = nnm_lane();
    if (M == k) return lane < k ? lane : M;
    uint32_t r = 0xFFFFFFFFu;
    if (lane < k) r = (uint32_t)(nnm_rand(seed, stream, lane) % (uint64_t)(M - k));
    r = nnm_wave_sort32(r);
    // bump: a'[i] = i + max_{j<=i}(a[j] - j)   (a[i] sorted, i < k)
    int32_t t = lane < k ? (int32_t)r - (int32_t)lane : -0x7FFFFFFF;
#pragma unroll
    for (int o = 1; o < NNM_WAVE; o <<= 1) {
        int32_t u = __shfl_up(t, (unsigned)o, NNM_WAVE);
        if ((int)lane >= o) t = t > u ? t : u;
    }
    uint32_t off = (uint32_t)(nnm_rand(seed, stream, 0xFFFFu) % M);
    if (lane >= k) return M;
    uint64_t v = (uint64_t)(t + (int32_t)lane) + off;
    return (uint32_t)(v % M);
}

// ---------------------------------------------------------- offer log
//
// The local join hands every offer (target, id, dist) to the target's owner
// through the offer log: each wave writes its offers in order into a log
// chunk of NNM_OCH records that it reserved with one atomic on a shared
// counter, so a wave store covers consecutive records and no offer ever
// overwrites another.  Partition passes then group the records by
// target and the owners merge them (nnm_offers.hip).  When the log has no
// chunk left the wave records the vertices it could not finish in the redo
// list; they are joined again in a later round (their earlier offers then
// arrive twice, which the owners' de-dup by id absorbs).
struct OfferSink {
    uint64_t* key;          // ochunks x NNM_OCH record keys
    uint32_t* tgt;          // ... and targets
    uint32_t* chunk_used;   // records written per chunk
    uint64_t* counters;     // CNT_OCHUNKS, CNT_REDO
    uint32_t* redo;         // vertices to join again
    uint32_t nchunks;
};

// Per-wave writer state (wave-uniform).  A writer without a chunk counts its
// chunk as full, so its first offers reserve one.
struct OfferWriter {
    uint32_t chunk = 0xFFFFFFFFu;
    uint32_t used = NNM_OCH;
    bool full = false;      // the log ran out of chunks for this wave
};

// Append the offers of the lanes with p set: record {key(dist, id), target}.
__device__ inline void offer_emit(const OfferSink& s, OfferWriter& w, bool p, uint32_t target, uint32_t id,
                                  float dist) {
    const uint64_t m = __ballot(p);
    if (!m || w.full) return;
    const uint32_t k = (uint32_t)__popcll(m);
    if (w.used + k > NNM_OCH) {
        const uint32_t lane = __lane_id();
        if (w.chunk != 0xFFFFFFFFu && lane == 0) s.chunk_used[w.chunk] = w.used;
        uint32_t ch = 0;
        if (lane == 0) ch = (uint32_t)atomicAdd((unsigned long long*)&s.counters[CNT_OCHUNKS], 1ull);
        ch = (uint32_t)__builtin_amdgcn_readfirstlane((int)ch);
        if (ch >= s.nchunks) {
            w.full = true;
            w.chunk = 0xFFFFFFFFu;
            w.used = NNM_OCH;
            return;
        }
        w.chunk = ch;
        w.used = 0;
    }
    if (p) {
        const uint64_t below = __lane_id() ? (m & (~0ull >> (64 - __lane_id()))) : 0ull;
        const uint64_t pos = (uint64_t)w.chunk * NNM_OCH + w.used + (uint32_t)__popcll(below);
        s.key[pos] = ((uint64_t)__float_as_uint(dist) << 32) | id;
        s.tgt[pos] = target;
    }
    w.used += k;
}

// Number of set bits of m below this lane (v_mbcnt).
__device__ inline uint32_t nnm_mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Reserve cnt (wave-uniform, at most NNM_OCH) consecutive records of the
// wave's log chunk for a batch of offers; *first is the batch's first record.
// A chunk that cannot hold the whole batch is closed at its current count
// (the passes read chunk_used records of every chunk, so its tail is never
// read) and a new chunk is taken.  Returns false, with w.full set, when the
// log is out of chunks.
__device__ inline bool offer_reserve(const OfferSink& s, OfferWriter& w, uint32_t cnt, uint64_t* first) {
    if (w.full) return false;
    // the writer is wave-uniform: keep it in scalar registers
    w.used = (uint32_t)__builtin_amdgcn_readfirstlane((int)w.used);
    w.chunk = (uint32_t)__builtin_amdgcn_readfirstlane((int)w.chunk);
    if (w.used + cnt > NNM_OCH) {
        const uint32_t lane = __lane_id();
        if (w.chunk != 0xFFFFFFFFu && lane == 0) s.chunk_used[w.chunk] = w.used;
        uint32_t ch = 0;
        if (lane == 0) ch = (uint32_t)atomicAdd((unsigned long long*)&s.counters[CNT_OCHUNKS], 1ull);
        ch = (uint32_t)__builtin_amdgcn_readfirstlane((int)ch);
        if (ch >= s.nchunks) {
            w.full = true;
            w.chunk = 0xFFFFFFFFu;
            w.used = NNM_OCH;
            return false;
        }
        w.chunk = ch;
        w.used = 0;
    }
    *first = (uint64_t)w.chunk * NNM_OCH + w.used;
    w.used += cnt;
    return true;
}

// Close the wave's chunk (its record count becomes visible to the passes).
__device__ inline void offer_close(const OfferSink& s, OfferWriter& w) {
    if (w.chunk != 0xFFFFFFFFu && __lane_id() == 0) s.chunk_used[w.chunk] = w.used;
    w.chunk = 0xFFFFFFFFu;
    w.used = NNM_OCH;
}

// Record vertices whose offers did not all reach the log (lanes with p set).
__device__ inline void offer_redo(const OfferSink& s, bool p, uint32_t v) {
    const uint64_t m = __ballot(p);
    if (!m) return;
    uint32_t base = 0;
    if (__lane_id() == 0) base = (uint32_t)atomicAdd((unsigned long long*)&s.counters[CNT_REDO], (unsigned long long)__popcll(m));
    base = (uint32_t)__builtin_amdgcn_readfirstlane((int)base);
    if (p) {
        const uint64_t below = __lane_id() ? (m & (~0ull >> (64 - __lane_id()))) : 0ull;
        s.redo[base + (uint32_t)__popcll(below)] = v;
    }
}

// Device scratch of nnm_bruteforce_knn_noalloc (bytes, one array per GPU)
// and the set-up of its per-device address (nnm_bruteforce.hip).
#define NNM_NOALLOC_SCRATCH_BYTES (16u << 20)
int nnm_bf_scratch_init(nnm_ctx* c);

// Offer log and delivery (nnm_offers.hip).
int nnm_offer_buffers(nnm_ctx* c, const nnm_params* p);
void nnm_offer_free(nnm_ctx* c);
OfferSink nnm_offer_sink(nnm_ctx* c, uint32_t round);
// Before a join round: no chunk reserved, empty redo list.
int nnm_offer_reset(nnm_ctx* c);
// Group the round's offers by target and merge them into the lists.
int nnm_offer_deliver(nnm_ctx* c, const nnm_params* p);

// Exact k-NN by brute force (FP32 difference form, ties broken by id).
//
// The reference has no such routine: it reads exact k-NN lists from a
// precomputed file (Evaluator::loadGrdKNNGraph, src/evaluator.cpp:343-378)
// and compares merged graphs against it (Evaluator::getKGraphRecall,
// src/evaluator.cpp:258-295).  This kernel produces those lists on the GPU.
//
// Layout: each lane owns one query row (held in registers) and a sorted
// top-KMAX list in registers; database rows stream through LDS in tiles of
// BT rows, read by every lane of the block at the same address.  The
// database is split into nsplit ranges over blockIdx.y; a second kernel
// merges the per-range lists.  Distances use the library order (0..d-1; L2
// with fused multiply-add, the other metrics with every operation rounded on
// its own), so they match the list kernels bit for bit.
//
// MFMA form (L2 on integer rows): when every row value is an integer whose
// squared norms and squared differences stay exact in FP32 (d max|v|^2 <=
// 2^23, d (max - min)^2 < 2^24; SIFT-like rows in [0, 255] with d <= 128),
// distances are n_q + n_x - 2 q.x with the dot products on
// v_mfma_f32_16x16x32_bf16: every product and partial sum is an integer below
// 2^24, so the value equals the FP32 difference form bit for bit.  A block
// holds 256 queries (64 per wave, their bf16 rows as MFMA operands in
// registers) and streams 64-row database tiles through LDS; each query keeps
// a sorted top-KM list in LDS and a threshold (its KM-th key) in registers.
// A 16x16 result block gives every query 16 candidates; those below the
// threshold are sorted within their 16 lanes and merged into the list by one
// bitonic merge across the wave.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "nnm_internal.hpp"

#define BF_BT 32
#define BF_KMAX 32

__device__ inline bool bf_lt(float d1, uint32_t i1, float d2, uint32_t i2) {
    return d1 < d2 || (d1 == d2 && i1 < i2);
}

template <int KM>
__device__ inline void bf_insert(float (&td)[KM], uint32_t (&ti)[KM], float d, uint32_t id) {
    if (!bf_lt(d, id, td[KM - 1], ti[KM - 1])) return;
    td[KM - 1] = d;
    ti[KM - 1] = id;
#pragma unroll
    for (int p = KM - 1; p > 0; --p) {
        const bool sw = bf_lt(td[p], ti[p], td[p - 1], ti[p - 1]);
        const float a = td[p], b = td[p - 1];
        const uint32_t ia = ti[p], ib = ti[p - 1];
        td[p] = sw ? b : a;
        td[p - 1] = sw ? a : b;
        ti[p] = sw ? ib : ia;
        ti[p - 1] = sw ? ia : ib;
    }
}

template <int LDMAX, uint32_t M>
__global__ __launch_bounds__(256) void k_bruteforce(const float* __restrict__ X, uint32_t ld, uint64_t n,
                                                    const float* __restrict__ Q, const uint32_t* __restrict__ rows,
                                                    uint64_t nq, int exclude_self, uint32_t nsplit,
                                                    float* __restrict__ part_d, uint32_t* __restrict__ part_i,
                                                    const float* __restrict__ sumsq) {
    __shared__ __align__(16) float tile[BF_BT * LDMAX];
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t split = blockIdx.y;
    const uint64_t per = (n + nsplit - 1) / nsplit;
    const uint64_t lo = split * per;
    const uint64_t hi = std::min<uint64_t>(n, lo + per);
    const bool active = q < nq;

    float qv[LDMAX];
    uint32_t self = 0xFFFFFFFFu;
    bool rowok = active;  // a query row past the data gets an empty list
    {
        const float* src = nullptr;
        if (active) {
            if (Q) {
                src = Q + q * ld;
            } else {
                const uint32_t r = rows[q];
                rowok = r < n;
                src = X + (uint64_t)(rowok ? r : 0u) * ld;
                if (exclude_self) self = r;
            }
        }
#pragma unroll
        for (int c = 0; c < LDMAX; c += 4) {
            float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
            if (rowok && (uint32_t)c < ld) t = *reinterpret_cast<const float4*>(src + c);
            qv[c] = t.x; qv[c + 1] = t.y; qv[c + 2] = t.z; qv[c + 3] = t.w;
        }
    }
    float qs = 0.f;  // cosine: the query's squared norm, in the metric's order
    if constexpr (M == NNM_METRIC_COSINE) {
#pragma unroll
        for (int c = 0; c < LDMAX; ++c) qs = metric_step<M>(qs, qv[c], qv[c]);
    }
    float td[BF_KMAX];
    uint32_t ti[BF_KMAX];
#pragma unroll
    for (int i = 0; i < BF_KMAX; ++i) { td[i] = INFINITY; ti[i] = 0xFFFFFFFFu; }

    const uint32_t row_f4 = ld / 4;
    for (uint64_t r0 = lo; r0 < hi; r0 += BF_BT) {
        const uint32_t nt = (uint32_t)std::min<uint64_t>(BF_BT, hi - r0);
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < nt * row_f4; e += blockDim.x) {
            const uint32_t rr = e / row_f4, cc = e % row_f4;
            *reinterpret_cast<float4*>(&tile[rr * LDMAX + cc * 4]) =
                *reinterpret_cast<const float4*>(X + (r0 + rr) * ld + cc * 4);
        }
        __syncthreads();
        if (!rowok) continue;
        for (uint32_t rr = 0; rr < nt; ++rr) {
            const float* tr = &tile[rr * LDMAX];
            float acc = 0.f;
#pragma unroll
            for (int c = 0; c < LDMAX; c += 4) {
                if ((uint32_t)c < ld) {
                    const float4 y = *reinterpret_cast<const float4*>(tr + c);
                    acc = metric_step<M>(acc, qv[c], y.x);
                    acc = metric_step<M>(acc, qv[c + 1], y.y);
                    acc = metric_step<M>(acc, qv[c + 2], y.z);
                    acc = metric_step<M>(acc, qv[c + 3], y.w);
                }
            }
            const uint32_t id = (uint32_t)(r0 + rr);
            if (id == self) continue;
            if constexpr (M == NNM_METRIC_COSINE) acc = nnm_cos_finish(acc, qs, sumsq[id]);
            bf_insert(td, ti, acc, id);
        }
    }
    if (!active) return;
    float* od = part_d + (q * nsplit + split) * BF_KMAX;
    uint32_t* oi = part_i + (q * nsplit + split) * BF_KMAX;
#pragma unroll
    for (int i = 0; i < BF_KMAX; ++i) { od[i] = td[i]; oi[i] = ti[i]; }
}

__global__ void k_bf_merge(uint64_t nq, uint32_t nsplit, uint32_t k, const float* __restrict__ part_d,
                           const uint32_t* __restrict__ part_i, float* __restrict__ out_d,
                           uint32_t* __restrict__ out_i) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    float td[BF_KMAX];
    uint32_t ti[BF_KMAX];
#pragma unroll
    for (int i = 0; i < BF_KMAX; ++i) { td[i] = INFINITY; ti[i] = 0xFFFFFFFFu; }
    for (uint32_t s = 0; s < nsplit; ++s) {
        const float* pd = part_d + (q * nsplit + s) * BF_KMAX;
        const uint32_t* pi = part_i + (q * nsplit + s) * BF_KMAX;
        for (uint32_t i = 0; i < BF_KMAX; ++i) {
            if (pi[i] == 0xFFFFFFFFu) break;
            bf_insert(td, ti, pd[i], pi[i]);
        }
    }
#pragma unroll
    for (int i = 0; i < BF_KMAX; ++i)
        if ((uint32_t)i < k) { out_d[q * k + i] = td[i]; out_i[q * k + i] = ti[i]; }
}

// ------------------------------------------------------------ MFMA form

#define BFM_Q 256   // queries per block: 4 waves x 64
#define BFM_TR 64   // database rows per LDS tile

typedef __bf16 bfm_bf16x8 __attribute__((ext_vector_type(8)));
typedef float bfm_f32x4 __attribute__((ext_vector_type(4)));

template <int KM>
struct BfmLds {
    uint4 tile[BFM_TR * 17];       // 64 rows of up to 16 uint4 (128 bf16) + 1 uint4 pad
    float xn[BFM_TR];
    uint64_t list[BFM_Q][KM];      // per query: sorted keys dist_bits << 32 | id
};

// KS: 32-element steps per row (ld = 32 KS); KM: list length (16 or 32).
// Q == nullptr: queries are rows[] of the database.
template <int KS, int KM>
__global__ __launch_bounds__(256) void k_bf_mfma(const uint16_t* __restrict__ Xb, const float* __restrict__ xnorm,
                                                 uint64_t n, const uint16_t* __restrict__ Qb,
                                                 const float* __restrict__ qnorm, const uint32_t* __restrict__ rows,
                                                 uint64_t nq, int exclude_self, uint32_t nsplit,
                                                 float* __restrict__ part_d, uint32_t* __restrict__ part_i) {
    constexpr uint32_t U4 = 4 * KS;   // uint4 per row
    constexpr uint32_t P = U4 + 1;    // LDS row pitch in uint4 (the pad spreads 16 rows over the banks)
    constexpr uint32_t LD = 32 * KS;  // bf16 per row
    __shared__ __align__(16) BfmLds<KM> L;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
    const uint32_t g = lane >> 4, c = lane & 15u;
    const uint64_t qbase = (uint64_t)blockIdx.x * BFM_Q + wv * 64u;
    const uint32_t split = blockIdx.y;
    const uint64_t per = (n + nsplit - 1) / nsplit;
    const uint64_t lo = std::min<uint64_t>(n, (uint64_t)split * per);
    const uint64_t hi = std::min<uint64_t>(n, lo + per);

    // A operands: lane (g, c) holds elements [32 ks + 8 g, +8) of query qb*16 + c
    uint4 A[4][KS];
#pragma unroll
    for (int qb = 0; qb < 4; ++qb) {
        const uint64_t q = qbase + qb * 16 + c;
        const uint16_t* src = nullptr;
        if (q < nq) {
            if (Qb) src = Qb + q * LD;
            else if (rows[q] < n) src = Xb + (uint64_t)rows[q] * LD;
        }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks)
            A[qb][ks] = src ? *reinterpret_cast<const uint4*>(src + ks * 32 + 8 * g) : make_uint4(0, 0, 0, 0);
    }
    // per lane: queries qb*16 + 4 g + i (the rows of its result registers)
    float qn[4][4];
    uint32_t self[4][4];
    uint64_t thr[4][4];
#pragma unroll
    for (int qb = 0; qb < 4; ++qb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint64_t q = qbase + qb * 16 + 4 * g + i;
            bool ok = q < nq;
            uint32_t r = 0xFFFFFFFFu;
            if (ok && !Qb) {
                r = rows[q];
                ok = r < n;
            }
            qn[qb][i] = ok ? (Qb ? qnorm[q] : xnorm[r]) : 0.f;
            self[qb][i] = (ok && !Qb && exclude_self) ? r : 0xFFFFFFFFu;
            thr[qb][i] = ok ? NNM_EMPTY_KEY : 0ull;  // 0: nothing passes
        }
    for (uint32_t e = tid; e < BFM_Q * KM; e += 256) L.list[e / KM][e % KM] = NNM_EMPTY_KEY;
    const uint4* X4 = reinterpret_cast<const uint4*>(Xb);

    for (uint64_t r0 = lo; r0 < hi; r0 += BFM_TR) {
        const uint32_t nt = (uint32_t)std::min<uint64_t>(BFM_TR, hi - r0);
        __syncthreads();  // the previous tile is consumed
        for (uint32_t e = tid; e < BFM_TR * U4; e += 256) {
            const uint32_t rr = e / U4, cc = e % U4;
            L.tile[rr * P + cc] = rr < nt ? X4[(r0 + rr) * U4 + cc] : make_uint4(0, 0, 0, 0);
        }
        if (tid < BFM_TR) L.xn[tid] = tid < nt ? xnorm[r0 + tid] : 0.f;
        __syncthreads();
        for (uint32_t sb = 0; sb * 16 < nt; ++sb) {
            uint4 B[KS];
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) B[ks] = L.tile[(sb * 16 + c) * P + ks * 4 + g];
            bfm_f32x4 acc[4];
#pragma unroll
            for (int qb = 0; qb < 4; ++qb) {
                acc[qb] = bfm_f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int ks = 0; ks < KS; ++ks)
                    acc[qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bfm_bf16x8, A[qb][ks]),
                                                                      __builtin_bit_cast(bfm_bf16x8, B[ks]),
                                                                      acc[qb], 0, 0, 0);
            }
            const uint32_t col = sb * 16 + c;
            const bool valid = col < nt;
            const uint32_t id = (uint32_t)(r0 + col);
            const float xn = L.xn[col];
#pragma unroll
            for (int qb = 0; qb < 4; ++qb)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const float dd = fmaxf((qn[qb][i] + xn) - 2.f * acc[qb][i], 0.f);
                    const uint64_t key = nnm_key(dd, id);
                    const bool pass = valid && id != self[qb][i] && key < thr[qb][i];
                    const uint64_t m = nnm_ballot(pass);
                    if (!m) continue;
                    // sort each 16-lane group (groups alternate direction:
                    // even ascending, odd descending), then merge group by group
                    uint64_t kk = pass ? key : NNM_EMPTY_KEY;
                    kk = bitonic_pass64<2, 1>(kk);
                    kk = bitonic_pass64<4, 2>(kk);
                    kk = bitonic_pass64<8, 4>(kk);
                    kk = bitonic_pass64<16, 8>(kk);
#pragma unroll
                    for (uint32_t gg = 0; gg < 4; ++gg) {
                        if (((m >> (16 * gg)) & 0xFFFFull) == 0) continue;
                        const uint32_t ql = wv * 64 + qb * 16 + 4 * gg + i;
                        // lanes [0, KM): the list ascending; [KM, 48): empty;
                        // [48, 64): the group's candidates descending
                        const uint32_t t = 63u - lane;  // ascending rank for lanes 48..63
                        const uint32_t src = lane >= 48 ? gg * 16 + ((gg & 1u) ? 15u - t : t) : lane;
                        const uint64_t cv = shfl64(kk, src);
                        uint64_t v = lane < (uint32_t)KM ? L.list[ql][lane] : NNM_EMPTY_KEY;
                        if (lane >= 48) v = cv;
                        v = bitonic_pass64<64, 32>(v);
                        if (lane < (uint32_t)KM) L.list[ql][lane] = v;
                        const uint64_t nthr = wave_read64(v, KM - 1);
                        if (g == gg) thr[qb][i] = nthr;
                    }
                }
        }
    }
    __syncthreads();
    for (uint32_t e = tid; e < BFM_Q * BF_KMAX; e += 256) {
        const uint32_t ql = e / BF_KMAX, j = e % BF_KMAX;
        const uint64_t q = (uint64_t)blockIdx.x * BFM_Q + ql;
        if (q >= nq) continue;
        const uint64_t key = j < (uint32_t)KM ? L.list[ql][j] : NNM_EMPTY_KEY;
        const bool empty = key == NNM_EMPTY_KEY;
        part_d[(q * nsplit + split) * BF_KMAX + j] = empty ? INFINITY : nnm_key_dist(key);
        part_i[(q * nsplit + split) * BF_KMAX + j] = empty ? 0xFFFFFFFFu : nnm_key_id(key);
    }
}

// Split count of a brute-force launch: enough (query block, range) blocks to
// fill the CUs a few times, at most 512 ranges, at most one per tile.
static uint32_t bf_splits(nnm_ctx* c, uint64_t nq) {
    const uint64_t qblocks = (nq + 255) / 256;
    const uint64_t want = (uint64_t)c->num_cus * 4;
    uint32_t nsplit = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want / qblocks, 512));
    return (uint32_t)std::min<uint64_t>(nsplit, (c->n + BF_BT - 1) / BF_BT);
}

// The MFMA form applies: L2, integer rows whose squared norms (d max|v|^2 <=
// 2^23) and squared distances (d (max - min)^2 < 2^24) are exact in FP32,
// d <= 128; with query vectors [qmin, qmax] widens the range.
static bool bf_mfma_ok(const nnm_ctx* c, bool with_q = false, float qmin = 0.f, float qmax = 0.f) {
    if (c->metric != NNM_METRIC_L2 || !c->data_exact_bf16 || c->ld > 128) return false;
    const double lo = with_q ? std::min(c->data_min, qmin) : c->data_min;
    const double hi = with_q ? std::max(c->data_max, qmax) : c->data_max;
    const double a = std::max(std::fabs(lo), std::fabs(hi)), r = hi - lo;
    return (double)c->d * a * a <= 8388608.0 && (double)c->d * r * r < 16777216.0;
}

static uint32_t bfm_splits(nnm_ctx* c, uint64_t nq) {
    const uint64_t qblocks = (nq + BFM_Q - 1) / BFM_Q;
    const uint64_t want = (uint64_t)c->num_cus * 2;
    uint32_t nsplit = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want / qblocks, 512));
    return (uint32_t)std::min<uint64_t>(nsplit, (c->n + BFM_TR - 1) / BFM_TR);
}

// MFMA range kernel + merge: queries are rows dR (dQb == nullptr) or bf16
// vectors dQb with squared norms dqn.
static int bfm_launch(nnm_ctx* c, const uint16_t* dQb, const float* dqn, const uint32_t* dR, uint64_t nq, uint32_t k,
                      int exclude_self, uint32_t nsplit, float* pd, uint32_t* pi, float* od, uint32_t* oi) {
    dim3 grid((uint32_t)((nq + BFM_Q - 1) / BFM_Q), nsplit);
    const uint32_t ks = c->ld / 32;
#define NNM_BFM(KS, KM)                                                                                        \
    hipLaunchKernelGGL((k_bf_mfma<KS, KM>), grid, dim3(256), 0, c->stream, c->Xb, c->norms, c->n, dQb, dqn, dR, \
                       nq, exclude_self, nsplit, pd, pi)
#define NNM_BFM_KS(KM)                \
    switch (ks) {                     \
        case 1: NNM_BFM(1, KM); break; \
        case 2: NNM_BFM(2, KM); break; \
        case 3: NNM_BFM(3, KM); break; \
        default: NNM_BFM(4, KM); break; \
    }
    if (k <= 16) { NNM_BFM_KS(16) } else { NNM_BFM_KS(32) }
#undef NNM_BFM_KS
#undef NNM_BFM
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return nnm_hip_fail(c, e, "k_bf_mfma launch");
    hipLaunchKernelGGL(k_bf_merge, dim3((uint32_t)((nq + 255) / 256)), dim3(256), 0, c->stream, nq, nsplit, k, pd,
                       pi, od, oi);
    e = hipGetLastError();
    if (e != hipSuccess) return nnm_hip_fail(c, e, "k_bf_merge launch");
    return NNM_OK;
}

// Launch of the range kernel and the merge over caller-provided buffers.
static int bf_launch(nnm_ctx* c, const float* dQ, const uint32_t* dR, uint64_t nq, uint32_t k, int exclude_self,
                     uint32_t nsplit, float* pd, uint32_t* pi, float* od, uint32_t* oi) {
    const uint64_t qblocks = (nq + 255) / 256;
    dim3 grid((uint32_t)qblocks, nsplit);
    if (c->metric == NNM_METRIC_COSINE && !c->sumsq) return nnm_fail(c, NNM_ERR_STATE, "no row norms");
#define NNM_BF(LDM, M)                                                                                         \
    hipLaunchKernelGGL((k_bruteforce<LDM, M>), grid, dim3(256), 0, c->stream, c->X, c->ld, c->n, dQ, dR, nq,   \
                       exclude_self, nsplit, pd, pi, c->sumsq)
#define NNM_BF_LD(M)                       \
    if (c->ld <= 64) NNM_BF(64, M);        \
    else if (c->ld <= 128) NNM_BF(128, M); \
    else NNM_BF(256, M);
    switch (c->metric) {
        case NNM_METRIC_L1: NNM_BF_LD(NNM_METRIC_L1) break;
        case NNM_METRIC_COSINE: NNM_BF_LD(NNM_METRIC_COSINE) break;
        case NNM_METRIC_CHI2: NNM_BF_LD(NNM_METRIC_CHI2) break;
        default: NNM_BF_LD(NNM_METRIC_L2) break;
    }
#undef NNM_BF_LD
#undef NNM_BF
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return nnm_hip_fail(c, e, "k_bruteforce launch");
    hipLaunchKernelGGL(k_bf_merge, dim3((uint32_t)((nq + 255) / 256)), dim3(256), 0, c->stream, nq, nsplit, k, pd,
                       pi, od, oi);
    e = hipGetLastError();
    if (e != hipSuccess) return nnm_hip_fail(c, e, "k_bf_merge launch");
    return NNM_OK;
}

extern "C" int nnm_bruteforce_knn(nnm_ctx* c, const float* Q, const uint32_t* rows, uint64_t nq, uint32_t k,
                                  int exclude_self, uint32_t* ids, float* dist) {
    if (!c || !ids || !dist || nq == 0) return nnm_fail(c, NNM_ERR_INVALID, "nnm_bruteforce_knn: bad arguments");
    if (!c->X) return nnm_fail(c, NNM_ERR_STATE, "no data");
    if (k == 0 || k > BF_KMAX) return nnm_fail(c, NNM_ERR_INVALID, "k must be in [1, 32]");
    if (!Q && !rows) return nnm_fail(c, NNM_ERR_INVALID, "need query vectors or query rows");
    if (c->ld > 256) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "brute force supports d <= 256");
    if (!Q)
        for (uint64_t i = 0; i < nq; ++i)
            if (rows[i] >= c->n) return nnm_fail(c, NNM_ERR_INVALID, "query row out of range");
    NNM_HIP(c, hipSetDevice(c->device));
    // query vectors: integers in a range that keeps the MFMA form exact?
    bool mfma = false;
    if (Q) {
        bool qint = true;
        float qmin = Q[0], qmax = Q[0];
        for (uint64_t i = 0; i < nq * c->d; ++i) {
            const float v = Q[i];
            qmin = std::min(qmin, v);
            qmax = std::max(qmax, v);
            if (qint && (!(v >= -256.f && v <= 256.f) || v != std::nearbyint(v))) qint = false;
        }
        mfma = qint && bf_mfma_ok(c, true, qmin, qmax);
    } else {
        mfma = bf_mfma_ok(c);
    }
    if (mfma) NNM_TRY(nnm_ensure_bf16(c));
    const uint32_t nsplit = mfma ? bfm_splits(c, nq) : bf_splits(c, nq);
    float* dQ = nullptr;
    uint16_t* dQb = nullptr;
    float* dqn = nullptr;
    uint32_t* dR = nullptr;
    float *pd = nullptr, *od = nullptr;
    uint32_t *pi = nullptr, *oi = nullptr;
    int rc = NNM_OK;
    auto cleanup = [&]() {
        nnm_free(c, dQ); nnm_free(c, dQb); nnm_free(c, dqn); nnm_free(c, dR); nnm_free(c, pd); nnm_free(c, pi);
        nnm_free(c, od); nnm_free(c, oi);
    };
    if (Q && mfma) {
        // bf16 rows of the integer queries (exact: the top 16 bits of each
        // float) and their squared norms (exact integers)
        std::vector<uint16_t> hb(nq * c->ld, 0);
        std::vector<float> hn(nq);
        for (uint64_t q = 0; q < nq; ++q) {
            double s2 = 0.0;
            for (uint32_t j = 0; j < c->d; ++j) {
                const float v = Q[q * c->d + j];
                uint32_t u;
                std::memcpy(&u, &v, 4);
                hb[q * c->ld + j] = (uint16_t)(u >> 16);
                s2 += (double)v * v;
            }
            hn[q] = (float)s2;
        }
        if ((rc = nnm_alloc(c, (void**)&dQb, nq * c->ld * 2)) || (rc = nnm_alloc(c, (void**)&dqn, nq * 4)) ||
            (rc = nnm_h2d(c, dQb, hb.data(), nq * c->ld * 2)) || (rc = nnm_h2d(c, dqn, hn.data(), nq * 4))) {
            cleanup();
            return rc;
        }
        NNM_HIP(c, hipStreamSynchronize(c->stream));  // the host staging buffers go out of scope
    } else if (Q) {
        rc = nnm_alloc(c, (void**)&dQ, nq * c->ld * sizeof(float));
        if (rc) { cleanup(); return rc; }
        rc = nnm_h2d_rows(c, dQ, c->ld, Q, c->d, nq);
        if (rc) { cleanup(); return rc; }
    } else {
        rc = nnm_alloc(c, (void**)&dR, nq * sizeof(uint32_t));
        if (rc) { cleanup(); return rc; }
        rc = nnm_h2d(c, dR, rows, nq * 4);
        if (rc) { cleanup(); return rc; }
    }
    if ((rc = nnm_alloc(c, (void**)&pd, nq * nsplit * BF_KMAX * 4)) ||
        (rc = nnm_alloc(c, (void**)&pi, nq * nsplit * BF_KMAX * 4)) ||
        (rc = nnm_alloc(c, (void**)&od, nq * k * 4)) || (rc = nnm_alloc(c, (void**)&oi, nq * k * 4))) {
        cleanup();
        return rc;
    }
    rc = mfma ? bfm_launch(c, dQb, dqn, dR, nq, k, exclude_self, nsplit, pd, pi, od, oi)
              : bf_launch(c, dQ, dR, nq, k, exclude_self, nsplit, pd, pi, od, oi);
    if (rc != NNM_OK) { cleanup(); return rc; }
    rc = nnm_d2h(c, ids, oi, nq * k * 4);
    if (rc == NNM_OK) rc = nnm_d2h(c, dist, od, nq * k * 4);
    cleanup();
    if (rc != NNM_OK) return rc;
    return NNM_OK;
}

// ------------------------------------------------------------ no allocation
//
// nnm_bruteforce_knn_noalloc keeps the per-range partial lists in this
// file-scope device array instead of allocations.  Host code uses the array
// only through its device address, taken once per device with that device
// current (nnm_bf_scratch_init, called by nnm_create); the name itself is a
// host stand-in, not a device address.  Every range kernel writes all its
// partial entries before the merge kernel reads them, so nothing in the
// array has to start at zero.
static __device__ __align__(256) uint8_t scratch[NNM_NOALLOC_SCRATCH_BYTES];

#define NNM_MAX_DEVICES 64
static void* g_scratch_ptr[NNM_MAX_DEVICES];

int nnm_bf_scratch_init(nnm_ctx* c) {
    if (c->device < 0 || c->device >= NNM_MAX_DEVICES) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "device index");
    if (g_scratch_ptr[c->device]) return NNM_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return nnm_hip_fail(c, e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[c->device] = p;
    return NNM_OK;
}

extern "C" int nnm_bruteforce_knn_noalloc(nnm_ctx* c, const uint32_t* rows_dev, uint64_t nq, uint32_t k,
                                          int exclude_self, uint32_t* ids_dev, float* dist_dev) {
    if (!c || !rows_dev || !ids_dev || !dist_dev || nq == 0)
        return nnm_fail(c, NNM_ERR_INVALID, "nnm_bruteforce_knn_noalloc: bad arguments");
    if (!c->X) return nnm_fail(c, NNM_ERR_STATE, "no data");
    if (k == 0 || k > BF_KMAX) return nnm_fail(c, NNM_ERR_INVALID, "k must be in [1, 32]");
    if (c->ld > 256) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "brute force supports d <= 256");
    void* base = c->device >= 0 && c->device < NNM_MAX_DEVICES ? g_scratch_ptr[c->device] : nullptr;
    if (!base) return nnm_fail(c, NNM_ERR_STATE, "no scratch address for this device (context not created)");
    const uint64_t per_split = nq * BF_KMAX * (sizeof(float) + sizeof(uint32_t));
    if (per_split > NNM_NOALLOC_SCRATCH_BYTES)
        return nnm_fail(c, NNM_ERR_INVALID, "nnm_bruteforce_knn_noalloc: nq above the scratch limit (65536 queries)");
    uint32_t nsplit = bf_splits(c, nq);
    nsplit = (uint32_t)std::min<uint64_t>(nsplit, NNM_NOALLOC_SCRATCH_BYTES / per_split);
    NNM_HIP(c, hipSetDevice(c->device));
    float* pd = reinterpret_cast<float*>(base);
    uint32_t* pi = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(base) + nq * nsplit * BF_KMAX * sizeof(float));
    NNM_TRY(bf_launch(c, nullptr, rows_dev, nq, k, exclude_self, nsplit, pd, pi, dist_dev, ids_dev));
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    return NNM_OK;
}

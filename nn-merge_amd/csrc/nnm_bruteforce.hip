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
#include <hip/hip_runtime.h>

#include <algorithm>

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

// Split count of a brute-force launch: enough (query block, range) blocks to
// fill the CUs a few times, at most 512 ranges, at most one per tile.
static uint32_t bf_splits(nnm_ctx* c, uint64_t nq) {
    const uint64_t qblocks = (nq + 255) / 256;
    const uint64_t want = (uint64_t)c->num_cus * 4;
    uint32_t nsplit = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want / qblocks, 512));
    return (uint32_t)std::min<uint64_t>(nsplit, (c->n + BF_BT - 1) / BF_BT);
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
    const uint32_t nsplit = bf_splits(c, nq);
    float* dQ = nullptr;
    uint32_t* dR = nullptr;
    float *pd = nullptr, *od = nullptr;
    uint32_t *pi = nullptr, *oi = nullptr;
    int rc = NNM_OK;
    auto cleanup = [&]() {
        nnm_free(c, dQ); nnm_free(c, dR); nnm_free(c, pd); nnm_free(c, pi); nnm_free(c, od); nnm_free(c, oi);
    };
    if (Q) {
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
    rc = bf_launch(c, dQ, dR, nq, k, exclude_self, nsplit, pd, pi, od, oi);
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

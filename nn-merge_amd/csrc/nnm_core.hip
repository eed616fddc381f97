// Context, data, partition, graph set-up and the NN-Descent driver loop.
//
// Kernel map (reference routine each one replaces):
//   k_random_init     KGraph::InitialGraph        src/kgraph.cpp:56-96
//   k_set_graph       AbstractMerge::InitSGraph   src/abstractmerge.cpp:55-76
//   k_merge_init      Smerge::InitialGraph        src/smerge.cpp:94-209
//                     Jmerge::InitialGraph        src/jmerge.cpp:83-176
//                     LyNNDescent::appdNwSmpl     src/lynndescent.cpp:354-452
//   k_merge_finish    AbstractMerge::merge2one    src/abstractmerge.cpp:232-277
//   k_lb_build        list build of nnDescent     src/abstractmerge.cpp:86-137
//                     (+ the new flags it leaves, EvaluateDelta src/abstractmerge.cpp:194-202)
//   k_lb_select       reverse sampling            src/abstractmerge.cpp:139-166
//   (nnm_offers.hip)  UpdateKnnList per pair      src/pvetex.h:40-85, 99-109
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "nnm_internal.hpp"

// ================================================================ host utils

int nnm_fail(nnm_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

int nnm_hip_fail(nnm_ctx* c, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) return nnm_fail(c, NNM_ERR_NOMEM, m);
    return nnm_fail(c, NNM_ERR_HIP, m);
}

int nnm_alloc(nnm_ctx* c, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    bytes = (bytes + 255) & ~size_t(255);
    if (c->mem_cap && c->mem_used + bytes > c->mem_cap)
        return nnm_fail(c, NNM_ERR_NOMEM, "device memory cap exceeded (" +
                                              std::to_string(c->mem_used + bytes) + " > " +
                                              std::to_string(c->mem_cap) + " bytes)");
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) { *p = nullptr; return nnm_hip_fail(c, e, "hipMalloc"); }
    c->mem_used += bytes;
    c->allocs.emplace_back(*p, bytes);
    return NNM_OK;
}

void nnm_free(nnm_ctx* c, void* p) {
    if (!p) return;
    for (size_t i = 0; i < c->allocs.size(); ++i) {
        if (c->allocs[i].first == p) {
            c->mem_used -= c->allocs[i].second;
            c->allocs.erase(c->allocs.begin() + i);
            break;
        }
    }
    hipFree(p);
}

// ---------------------------------------------------------------- transfers
// Every host <-> device copy goes through two pinned staging buffers that the
// context allocates once: the runtime's path for pageable memory pins the
// caller's pages for large copies, which not every host allows, and pinned
// staging also keeps the copy engine at full PCIe rate.
#define NNM_STAGE_BYTES (size_t(16) << 20)

static int ensure_stage(nnm_ctx* c) {
    if (c->stage[0]) return NNM_OK;
    for (int i = 0; i < 2; ++i) {
        hipError_t e = hipHostMalloc(&c->stage[i], NNM_STAGE_BYTES, hipHostMallocDefault);
        if (e != hipSuccess) { c->stage[i] = nullptr; return nnm_hip_fail(c, e, "hipHostMalloc (staging)"); }
        e = hipEventCreateWithFlags(&c->stage_ev[i], hipEventDisableTiming);
        if (e != hipSuccess) return nnm_hip_fail(c, e, "hipEventCreate (staging)");
        // recorded once so the first wait on it returns at once
        NNM_HIP(c, hipEventRecord(c->stage_ev[i], c->stream));
    }
    return NNM_OK;
}

static void free_stage(nnm_ctx* c) {
    for (int i = 0; i < 2; ++i) {
        if (c->stage[i]) hipHostFree(c->stage[i]);
        if (c->stage_ev[i]) hipEventDestroy(c->stage_ev[i]);
        c->stage[i] = nullptr;
        c->stage_ev[i] = nullptr;
    }
}

int nnm_h2d(nnm_ctx* c, void* dst, const void* src, size_t bytes) {
    NNM_TRY(ensure_stage(c));
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    int k = 0;
    for (size_t off = 0; off < bytes; off += NNM_STAGE_BYTES, k ^= 1) {
        const size_t m = std::min(NNM_STAGE_BYTES, bytes - off);
        NNM_HIP(c, hipEventSynchronize(c->stage_ev[k]));  // buffer k is free again
        std::memcpy(c->stage[k], s + off, m);
        NNM_HIP(c, hipMemcpyAsync(d + off, c->stage[k], m, hipMemcpyHostToDevice, c->stream));
        NNM_HIP(c, hipEventRecord(c->stage_ev[k], c->stream));
    }
    return NNM_OK;
}

int nnm_h2d_rows(nnm_ctx* c, float* dst, uint32_t ld, const float* src, uint32_t d, uint64_t n) {
    NNM_TRY(ensure_stage(c));
    const uint64_t per = std::max<uint64_t>(1, NNM_STAGE_BYTES / ((size_t)ld * 4));
    if ((size_t)ld * 4 > NNM_STAGE_BYTES) return nnm_fail(c, NNM_ERR_INVALID, "row too long for staging");
    int k = 0;
    for (uint64_t r0 = 0; r0 < n; r0 += per, k ^= 1) {
        const uint64_t m = std::min(per, n - r0);
        NNM_HIP(c, hipEventSynchronize(c->stage_ev[k]));
        float* st = static_cast<float*>(c->stage[k]);
        for (uint64_t r = 0; r < m; ++r) {
            std::memcpy(st + r * ld, src + (r0 + r) * d, (size_t)d * 4);
            if (ld > d) std::memset(st + r * ld + d, 0, (size_t)(ld - d) * 4);
        }
        NNM_HIP(c, hipMemcpyAsync(dst + r0 * ld, st, m * ld * 4, hipMemcpyHostToDevice, c->stream));
        NNM_HIP(c, hipEventRecord(c->stage_ev[k], c->stream));
    }
    return NNM_OK;
}

int nnm_d2h(nnm_ctx* c, void* dst, const void* src, size_t bytes) {
    NNM_TRY(ensure_stage(c));
    const char* s = static_cast<const char*>(src);
    char* d = static_cast<char*>(dst);
    const size_t nch = (bytes + NNM_STAGE_BYTES - 1) / NNM_STAGE_BYTES;
    auto issue = [&](size_t i) -> int {
        const size_t off = i * NNM_STAGE_BYTES, m = std::min(NNM_STAGE_BYTES, bytes - off);
        NNM_HIP(c, hipMemcpyAsync(c->stage[i & 1], s + off, m, hipMemcpyDeviceToHost, c->stream));
        NNM_HIP(c, hipEventRecord(c->stage_ev[i & 1], c->stream));
        return NNM_OK;
    };
    if (nch) NNM_TRY(issue(0));
    for (size_t i = 0; i < nch; ++i) {
        if (i + 1 < nch) NNM_TRY(issue(i + 1));
        NNM_HIP(c, hipEventSynchronize(c->stage_ev[i & 1]));
        const size_t off = i * NNM_STAGE_BYTES;
        std::memcpy(d + off, c->stage[i & 1], std::min(NNM_STAGE_BYTES, bytes - off));
    }
    return NNM_OK;
}

template <typename T>
static int alloc_n(nnm_ctx* c, T** p, size_t n) {
    return nnm_alloc(c, reinterpret_cast<void**>(p), n * sizeof(T));
}


// Fill kernels are used instead of hipMemsetAsync in the iteration loop so the
// loop can be captured into a HIP graph.
__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

static uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap = 65535u * 16u) {
    uint64_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    if (g > cap) g = cap;
    return (uint32_t)g;
}

// Blocks of 256 threads (4 waves) for persistent wave loops: enough to fill
// every CU a few times over, never more than one wave per row.
static uint32_t wave_grid(nnm_ctx* c, uint64_t rows) {
    uint64_t want = (uint64_t)c->num_cus * 8;
    uint64_t need = (rows + 3) / 4;
    return (uint32_t)std::max<uint64_t>(1, std::min(want, need));
}

static int fill_u32(nnm_ctx* c, uint32_t* p, uint64_t n, uint32_t v) {
    if (!n) return NNM_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, c->stream, p, n, v);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}
static int fill_u64(nnm_ctx* c, uint64_t* p, uint64_t n, uint64_t v) {
    if (!n) return NNM_OK;
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n, 256, 8192)), dim3(256), 0, c->stream, p, n, v);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}
int nnm_fill_u64(nnm_ctx* c, uint64_t* p, uint64_t n, uint64_t v) { return fill_u64(c, p, n, v); }
int nnm_fill_u32(nnm_ctx* c, uint32_t* p, uint64_t n, uint32_t v) { return fill_u32(c, p, n, v); }

// ================================================================ context

extern "C" void nnm_default_params(nnm_params* p) {
    if (!p) return;
    p->K = 20;
    p->rho = 1.0f;
    p->rev_cap = 100;
    p->max_iters = 0;
    p->delta = 0.002f;
    p->rule = NNM_RULE_ALL;
    p->prec = NNM_PREC_FP32;
    p->rerank = NNM_RERANK_AUTO;
    p->offer_cap = 0;
    p->cut = 0;
    p->plain_tail = 1;
    p->flags = 0;
    p->learn_tau = NNM_LEARN_TAU_DEFAULT;
    p->learn_rho = 0.f;
    p->seed = 1;
}

extern "C" int nnm_create(nnm_ctx** out, int device) {
    if (!out) return NNM_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return NNM_ERR_HIP;
    if (device < 0 || device >= ndev) return NNM_ERR_INVALID;
    nnm_ctx* c = new nnm_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) { delete c; return NNM_ERR_HIP; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return NNM_ERR_HIP; }
    hipEventCreate(&c->ev0);
    hipEventCreate(&c->ev1);
    hipEventCreate(&c->evj0);
    hipEventCreate(&c->evj1);
    hipEventCreate(&c->evjm);
    hipEventCreate(&c->evd1);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
    if (alloc_n(c, &c->counters, CNT_N) != NNM_OK) { delete c; return NNM_ERR_NOMEM; }
    // the allocation-free brute force's scratch address on this device
    if (nnm_bf_scratch_init(c) != NNM_OK) { nnm_destroy(c); return NNM_ERR_HIP; }
    *out = c;
    return NNM_OK;
}

extern "C" int nnm_destroy(nnm_ctx* c) {
    if (!c) return NNM_ERR_INVALID;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    nnm_mlp_free(c);
    nnm_search_free(c);
    nnm_dpg_free(c);
    free_stage(c);
    for (auto& a : c->allocs) hipFree(a.first);
    c->allocs.clear();
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    hipEventDestroy(c->evj0);
    hipEventDestroy(c->evj1);
    hipEventDestroy(c->evjm);
    hipEventDestroy(c->evd1);
    hipStreamDestroy(c->stream);
    delete c;
    return NNM_OK;
}

extern "C" const char* nnm_last_error(const nnm_ctx* c) {
    if (!c) return "null context";
    return c->err.c_str();
}

extern "C" int nnm_set_mem_cap(nnm_ctx* c, uint64_t bytes) {
    if (!c) return NNM_ERR_INVALID;
    c->mem_cap = bytes;
    return NNM_OK;
}

extern "C" uint64_t nnm_mem_used(const nnm_ctx* c) { return c ? c->mem_used : 0; }

static void free_graph(nnm_ctx* c) {
    nnm_free(c, c->g_ids); nnm_free(c, c->g_dist); nnm_free(c, c->g_flag);
    c->g_ids = nullptr; c->g_dist = nullptr; c->g_flag = nullptr;
    nnm_free(c, c->s_ids); nnm_free(c, c->s_dist); nnm_free(c, c->s_flag);
    c->s_ids = nullptr; c->s_dist = nullptr; c->s_flag = nullptr;
    nnm_free(c, c->p_ids);
    c->p_ids = nullptr;
    nnm_free(c, c->p_dist);
    c->p_dist = nullptr;
    nnm_free(c, c->c_ids); nnm_free(c, c->c_dist); nnm_free(c, c->c_flag); nnm_free(c, c->has_cpn);
    c->c_ids = nullptr; c->c_dist = nullptr; c->c_flag = nullptr; c->has_cpn = nullptr;
    c->K = 0; c->cut = 0; c->cpn_live = 0;
}

static void free_iter(nnm_ctx* c) {
    void** ps[] = {(void**)&c->l_new, (void**)&c->l_old, (void**)&c->n_new, (void**)&c->n_old,
                   (void**)&c->newmask, (void**)&c->rcnt, (void**)&c->roff, (void**)&c->rcur,
                   (void**)&c->rbuf, (void**)&c->ocnt, (void**)&c->obuf, (void**)&c->spill,
                   (void**)&c->scan_tmp, (void**)&c->kth, (void**)&c->vmeta, (void**)&c->lrec};
    for (void** p : ps) { nnm_free(c, *p); *p = nullptr; }
    c->lrec_n = c->lrec_gen = 0;
    c->lrec_live = 0;
    nnm_offer_free(c);
    c->cap_new = c->cap_old = c->rev_cap = 0;
}

static void free_partition(nnm_ctx* c) {
    void** ps[] = {(void**)&c->label, (void**)&c->vsub, (void**)&c->sub_begin, (void**)&c->sub_grp,
                   (void**)&c->sub_built, (void**)&c->grp_begin, (void**)&c->grp_end,
                   (void**)&c->grp_flags, (void**)&c->grp_active};
    for (void** p : ps) { nnm_free(c, *p); *p = nullptr; }
    c->nsub = c->ngrp = 0;
    c->h_sub_begin.clear(); c->h_sub_group.clear(); c->h_sub_built.clear(); c->h_grp_first_sub.clear();
}

// ================================================================ data

__global__ void k_to_bf16(const float* __restrict__ X, uint16_t* __restrict__ Xb, float* __restrict__ norms,
                          uint64_t n, uint32_t ld) {
    // one wave per row: round to nearest even bf16, norm of the rounded row
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / NNM_WAVE) + threadIdx.x / NNM_WAVE;
    const uint32_t lane = threadIdx.x % NNM_WAVE;
    if (row >= n) return;
    float acc = 0.f;
    for (uint32_t k = lane; k < ld; k += NNM_WAVE) {
        float x = X[row * ld + k];
        uint32_t u = __float_as_uint(x);
        u += 0x7FFFu + ((u >> 16) & 1u);
        uint16_t h = (uint16_t)(u >> 16);
        Xb[row * ld + k] = h;
        float r = __uint_as_float((uint32_t)h << 16);
        acc = __builtin_fmaf(r, r, acc);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, NNM_WAVE);
    if (lane == 0) norms[row] = acc;
}

// int8 rows: value - shift for the d values, 0 past them; the squared norm
// as an exact integer.  One wave per row.
__global__ void k_to_i8(const float* __restrict__ X, int8_t* __restrict__ Xi8, float* __restrict__ norms, uint64_t n,
                        uint32_t ld, uint32_t d, float shift) {
    const uint64_t row = (uint64_t)blockIdx.x * (blockDim.x / NNM_WAVE) + threadIdx.x / NNM_WAVE;
    const uint32_t lane = threadIdx.x % NNM_WAVE;
    if (row >= n) return;
    int acc = 0;
    for (uint32_t k = lane; k < ld; k += NNM_WAVE) {
        const int q = k < d ? (int)(X[row * ld + k] - shift) : 0;
        Xi8[row * ld + k] = (int8_t)q;
        acc += q * q;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, NNM_WAVE);
    if (lane == 0) norms[row] = (float)acc;
}

int nnm_ensure_i8(nnm_ctx* c) {
    if (c->Xi8) return NNM_OK;
    if (!c->data_i8) return nnm_fail(c, NNM_ERR_STATE, "int8 rows need integer data spanning at most 255");
    NNM_TRY(alloc_n(c, &c->Xi8, c->n * c->ld));
    NNM_TRY(alloc_n(c, &c->norms_i8, c->n));
    hipLaunchKernelGGL(k_to_i8, dim3(grid_for(c->n, 4, 1u << 30)), dim3(256), 0, c->stream, c->X, c->Xi8, c->norms_i8,
                       c->n, c->ld, c->d, c->i8_shift);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

int nnm_ensure_bf16(nnm_ctx* c) {
    if (c->Xb) return NNM_OK;
    NNM_TRY(alloc_n(c, &c->Xb, c->n * c->ld));
    NNM_TRY(alloc_n(c, &c->norms, c->n));
    hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(c->n, 4, 1u << 30)), dim3(256), 0, c->stream,
                       c->X, c->Xb, c->norms, c->n, c->ld);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

// Squared row norms for the cosine metric, in the metric's own order.
__global__ void k_sumsq(const float* __restrict__ X, uint32_t ld, uint64_t n, float* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = nnm_sumsq_rows(X + i * ld, ld);
}

static int ensure_sumsq(nnm_ctx* c) {
    if (c->metric != NNM_METRIC_COSINE || !c->X || c->sumsq) return NNM_OK;
    NNM_TRY(alloc_n(c, &c->sumsq, c->n));
    hipLaunchKernelGGL(k_sumsq, dim3(grid_for(c->n, 256, 1u << 30)), dim3(256), 0, c->stream, c->X, c->ld, c->n,
                       c->sumsq);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

extern "C" int nnm_set_metric(nnm_ctx* c, uint32_t metric) {
    if (!c) return NNM_ERR_INVALID;
    if (metric > NNM_METRIC_CHI2) return nnm_fail(c, NNM_ERR_INVALID, "nnm_set_metric: unknown metric");
    NNM_HIP(c, hipSetDevice(c->device));
    c->metric = metric;
    return ensure_sumsq(c);
}

int nnm_require_l2(nnm_ctx* c, const char* what) {
    if (c->metric == NNM_METRIC_L2) return NNM_OK;
    return nnm_fail(c, NNM_ERR_UNSUPPORTED, std::string(what) + ": L2 metric only");
}

extern "C" int nnm_set_data(nnm_ctx* c, const float* X, uint64_t n, uint32_t d) {
    if (!c || !X || n < 2 || d == 0 || d > 1024) return nnm_fail(c, NNM_ERR_INVALID, "nnm_set_data: bad arguments");
    if (n >= 0xFFFFFFFFull) return nnm_fail(c, NNM_ERR_INVALID, "nnm_set_data: n must be < 2^32-1");
    NNM_HIP(c, hipSetDevice(c->device));
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    free_graph(c);
    free_iter(c);
    free_partition(c);
    nnm_mlp_invalidate_projection(c);
    nnm_search_free(c);
    nnm_dpg_free(c);
    nnm_free(c, c->X); c->X = nullptr;
    nnm_free(c, c->Xb); c->Xb = nullptr;
    nnm_free(c, c->Xi8); c->Xi8 = nullptr;
    nnm_free(c, c->norms_i8); c->norms_i8 = nullptr;
    nnm_free(c, c->norms); c->norms = nullptr;
    nnm_free(c, c->sumsq); c->sumsq = nullptr;
    c->n = n;
    c->d = d;
    c->ld = (d + 31) / 32 * 32;
    NNM_TRY(alloc_n(c, &c->X, n * c->ld));
    NNM_TRY(nnm_h2d_rows(c, c->X, c->ld, X, d, n));
    NNM_TRY(ensure_sumsq(c));
    // bf16 exactness: integers of magnitude <= 256 are exact in bf16; the
    // distances are then exact in FP32 in both forms when every partial sum
    // stays at most 2^24: d (max - min)^2 for the difference form's sums and
    // the distances, 2 d max|v|^2 for the norm form's norm sums
    int exact = (d <= 128) ? 1 : 0;
    const uint64_t total = n * (uint64_t)d;
    float vmin = X[0], vmax = X[0];
    for (uint64_t i = 0; i < total; ++i) {
        const float v = X[i];
        vmin = v < vmin ? v : vmin;
        vmax = v > vmax ? v : vmax;
        if (exact && (!(v >= -256.f && v <= 256.f) || v != std::nearbyint(v))) exact = 0;
    }
    if (exact) {
        const double span = (double)vmax - (double)vmin, amax = std::max(std::fabs((double)vmin), std::fabs((double)vmax));
        if ((double)d * span * span > 16777216.0 || 2.0 * d * amax * amax > 16777216.0) exact = 0;
    }
    c->data_exact_bf16 = exact;
    // int8 rows (value - (min + 128)) are exact when the values span at most
    // 255; the int8 join is built for rows of 128
    c->data_i8 = exact && (double)vmax - (double)vmin <= 255.0 && c->ld == 128;
    c->i8_shift = vmin + 128.f;
    c->data_min = vmin;
    c->data_max = vmax;
    // default partition: one built subgraph, one group
    uint32_t sb[2] = {0u, (uint32_t)n};
    uint8_t bu[1] = {1};
    uint32_t gr[1] = {0};
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    return nnm_set_partition(c, sb, bu, gr, 1);
}

// ================================================================ partition

__global__ void k_fill_labels(uint32_t* __restrict__ label, uint32_t* __restrict__ vsub,
                              const uint32_t* __restrict__ sub_begin, const uint8_t* __restrict__ built,
                              uint32_t nsub, uint64_t n) {
    uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t lo = 0, hi = nsub;  // find s with sub_begin[s] <= v < sub_begin[s+1]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2;
        if (sub_begin[mid] <= v) lo = mid; else hi = mid;
    }
    vsub[v] = lo;
    label[v] = built[lo] ? lo : NNM_RAW;
}

extern "C" int nnm_set_partition(nnm_ctx* c, const uint32_t* sub_begin, const uint8_t* sub_built,
                                 const uint32_t* sub_group, uint32_t nsub) {
    if (!c || !sub_begin || !sub_built || !sub_group || nsub == 0)
        return nnm_fail(c, NNM_ERR_INVALID, "nnm_set_partition: bad arguments");
    if (!c->X) return nnm_fail(c, NNM_ERR_STATE, "nnm_set_partition: no data");
    if (sub_begin[0] != 0 || sub_begin[nsub] > c->n)
        return nnm_fail(c, NNM_ERR_INVALID, "nnm_set_partition: sub_begin must run from 0 to at most n");
    for (uint32_t s = 0; s < nsub; ++s)
        if (sub_begin[s + 1] <= sub_begin[s])
            return nnm_fail(c, NNM_ERR_INVALID, "nnm_set_partition: empty or unordered subgraph");
    std::vector<uint32_t> gb, ge, first;
    for (uint32_t s = 0; s < nsub; ++s) {
        if (s == 0 || sub_group[s] != sub_group[s - 1]) {
            if (sub_group[s] != (uint32_t)gb.size())
                return nnm_fail(c, NNM_ERR_INVALID, "nnm_set_partition: groups must be numbered 0.. in row order");
            gb.push_back(sub_begin[s]);
            ge.push_back(sub_begin[s + 1]);
            first.push_back(s);
        } else {
            ge.back() = sub_begin[s + 1];
        }
    }
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    free_partition(c);
    c->nsub = nsub;
    c->ngrp = (uint32_t)gb.size();
    c->h_sub_begin.assign(sub_begin, sub_begin + nsub + 1);
    c->h_sub_built.assign(sub_built, sub_built + nsub);
    c->h_sub_group.assign(sub_group, sub_group + nsub);
    c->h_grp_first_sub = first;
    NNM_TRY(alloc_n(c, &c->label, c->n));
    NNM_TRY(alloc_n(c, &c->vsub, c->n));
    NNM_TRY(alloc_n(c, &c->sub_begin, nsub + 1));
    NNM_TRY(alloc_n(c, &c->sub_grp, nsub));
    NNM_TRY(alloc_n(c, &c->sub_built, nsub));
    NNM_TRY(alloc_n(c, &c->grp_begin, c->ngrp));
    NNM_TRY(alloc_n(c, &c->grp_end, c->ngrp));
    NNM_TRY(alloc_n(c, &c->grp_flags, c->ngrp));
    NNM_TRY(alloc_n(c, &c->grp_active, c->ngrp));
    std::vector<uint8_t> built01(nsub);
    for (uint32_t s = 0; s < nsub; ++s) built01[s] = sub_built[s] ? 1 : 0;
    NNM_TRY(nnm_h2d(c, c->sub_begin, sub_begin, (nsub + 1) * 4));
    NNM_TRY(nnm_h2d(c, c->sub_grp, sub_group, nsub * 4));
    NNM_TRY(nnm_h2d(c, c->sub_built, built01.data(), nsub));
    NNM_TRY(nnm_h2d(c, c->grp_begin, gb.data(), gb.size() * 4));
    NNM_TRY(nnm_h2d(c, c->grp_end, ge.data(), ge.size() * 4));
    c->n_act = sub_begin[nsub];
    hipLaunchKernelGGL(k_fill_labels, dim3(grid_for(c->n_act, 256, 1u << 30)), dim3(256), 0, c->stream, c->label,
                       c->vsub, c->sub_begin, c->sub_built, nsub, c->n_act);
    NNM_LAUNCH_CHECK(c);
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    return NNM_OK;
}

// ================================================================ graph kernels

// Persistent wave loop: wave w of the grid takes rows w, w + W, w + 2W, ...
#define NNM_WAVE_LOOP(v, n)                                                                       \
    for (uint64_t v = (uint64_t)blockIdx.x * (blockDim.x / NNM_WAVE) + threadIdx.x / NNM_WAVE; v < (n); \
         v += (uint64_t)gridDim.x * (blockDim.x / NNM_WAVE))

// One atomic per wave for a per-lane count.
__device__ inline void wave_count(uint64_t* ctr, uint64_t per_lane) {
    uint64_t t = per_lane;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, NNM_WAVE);
    if (nnm_lane() == 0 && t) atomicAdd((unsigned long long*)ctr, (unsigned long long)t);
}

// Write lanes [0, K) of a row: (key, flag) per lane, sorted by key first.
// Ids inside a row are distinct, so the flag of each sorted entry is found
// by id.
__device__ inline void wave_write_sorted_row(uint64_t key, uint32_t flag, uint32_t K,
                                             uint32_t* __restrict__ ids, float* __restrict__ dist,
                                             uint8_t* __restrict__ flg) {
    const uint32_t lane = nnm_lane();
    uint64_t s = nnm_wave_sort64(lane < K ? key : NNM_EMPTY_KEY);
    uint32_t sid = nnm_key_id(s);
    uint32_t myf = 1;
    for (uint32_t p = 0; p < K; ++p) {
        const uint32_t pid = wave_read32((uint32_t)key, p);
        const uint32_t pf = wave_read32(flag, p);
        if (pid == sid) myf = pf;
    }
    if (lane < K) {
        ids[lane] = sid;
        dist[lane] = nnm_key_dist(s);
        flg[lane] = (uint8_t)myf;
    }
}

// Two rows per wave (K <= 32): half h = lane / 32 holds row 2w + h, entry j
// in lane 32h + j.  Each half is sorted on its own by a bitonic network that
// moves a 32-bit payload (the entry's flag) with its key; the last stage
// uses block size 64 so that both halves come out ascending.
template <uint32_t KB, uint32_t J>
__device__ inline void bitonic_pass_kv(uint64_t& key, uint32_t& val) {
    const uint64_t ok = wave_xor64<J>(key);
    const uint32_t ov = wave_xor32<J>(val);
    const uint32_t lane = nnm_lane();
    const bool want_min = ((lane & KB) == 0) == ((lane & J) == 0);
    const bool take = want_min ? ok < key : key < ok;
    key = take ? ok : key;
    val = take ? ov : val;
    if constexpr (J > 1) bitonic_pass_kv<KB, J / 2>(key, val);
}

__device__ inline void half_sort_kv(uint64_t& key, uint32_t& val) {
    bitonic_pass_kv<2, 1>(key, val);
    bitonic_pass_kv<4, 2>(key, val);
    bitonic_pass_kv<8, 4>(key, val);
    bitonic_pass_kv<16, 8>(key, val);
    bitonic_pass_kv<64, 16>(key, val);
}

#define NNM_HALF_LOOP(v0, n)                                                                             \
    for (uint64_t v0 = 2 * ((uint64_t)blockIdx.x * (blockDim.x / NNM_WAVE) + threadIdx.x / NNM_WAVE); v0 < (n); \
         v0 += 2 * (uint64_t)gridDim.x * (blockDim.x / NNM_WAVE))

// Write the half's row (entries in half-lanes [0, K)) sorted by key; rows
// with ok false (past the end, or skipped) are left alone.
__device__ inline void half_write_sorted_row(uint64_t key, uint32_t flag, bool ok, uint32_t K,
                                             uint32_t* __restrict__ ids, float* __restrict__ dist,
                                             uint8_t* __restrict__ flg) {
    const uint32_t hl = nnm_lane() & 31u;
    uint64_t s = hl < K ? key : NNM_EMPTY_KEY;
    uint32_t f = flag;
    half_sort_kv(s, f);
    if (ok && hl < K) {
        ids[hl] = nnm_key_id(s);
        dist[hl] = nnm_key_dist(s);
        flg[hl] = (uint8_t)f;
    }
}

// K random distinct neighbours inside the row's group, self excluded.
// Draws K+1 values like getRndSds(rng, K0+1, seeds, N) (src/kgraph.cpp:71).
__device__ inline void wave_random_row(const float* __restrict__ X, const uint16_t* __restrict__ Xb, uint32_t ld,
                                       uint64_t v, uint32_t gb,
                                       uint32_t gsize, uint32_t K, uint64_t seed, uint32_t* __restrict__ ids,
                                       float* __restrict__ dist, uint8_t* __restrict__ flg, uint64_t* nd,
                                       const nnm_metric_args& mt) {
    const uint32_t lane = nnm_lane();
    uint32_t r = nnm_wave_rnd_sds(seed, v, K + 1, gsize);
    uint32_t id = gb + r;
    bool ok = lane < K + 1 && id != (uint32_t)v;
    uint64_t m = nnm_ballot(ok);
    uint32_t rank = nnm_mask_rank(m);
    bool take = ok && rank < K;
    // move the taken ids to lanes [0, K)
    uint32_t src_lane = 0;
    // lane t < K receives the id of the lane whose rank is t
    uint64_t mm = m;
    for (uint32_t t = 0; t < K; ++t) {
        uint32_t l = (uint32_t)__ffsll((long long)mm) - 1;
        mm &= mm - 1;
        if (lane == t) src_lane = l;
    }
    (void)take;
    uint32_t myid = (uint32_t)__shfl((int)id, (int)src_lane, NNM_WAVE);
    float dd = 0.f;
    if (lane < K) dd = nnm_row_dist(X, Xb, ld, v, myid, mt);
    uint64_t key = nnm_key(dd, myid);
    if (lane < K) *nd += 1;
    wave_write_sorted_row(key, 1u, K, ids, dist, flg);
}

// Distance of the list set-up kernels.  Where the metric is L2 the choice
// of rows is made at compile time, so those kernels carry no other metric's
// code or registers; each mode calls what nnm_row_dist would, so the bits
// are the same.  DM_ANY: the run-time metric (FP32 rows, or bf16 rows when
// given); DM_L2_BF16: L2 on bf16 rows (exact there); DM_L2_F32: L2 on FP32
// rows.
// DM_L2_I8: L2 on the int8 rows (passed as Xb; exact there).
enum : int { DM_ANY = 0, DM_L2_BF16 = 1, DM_L2_F32 = 2, DM_L2_I8 = 3 };
template <int DM>
__device__ inline float setup_dist(const float* __restrict__ X, const uint16_t* __restrict__ Xb, uint32_t ld,
                                   uint64_t a, uint64_t b, const nnm_metric_args& mt) {
    if constexpr (DM == DM_L2_I8) {
        const int8_t* q = reinterpret_cast<const int8_t*>(Xb);
        return nnm_l2_rows_i8(q + a * ld, q + b * ld, ld);
    } else if constexpr (DM == DM_L2_BF16) return nnm_l2_rows_bf16(Xb + a * ld, Xb + b * ld, ld);
    else if constexpr (DM == DM_L2_F32) return nnm_l2_rows(X + a * ld, X + b * ld, ld);
    else return nnm_row_dist(X, Xb, ld, a, b, mt);
}
static int setup_dm(const nnm_ctx* c, const uint16_t* xb) {
    return c->metric != NNM_METRIC_L2 ? DM_ANY : xb ? DM_L2_BF16 : DM_L2_F32;
}

// nnm_wave_rnd_sds for each 32-lane half on its own (k <= 32; seed, stream,
// k and M may differ between the halves): the same values in half-lanes
// [0, k) as the full-wave draw gives in lanes [0, k).
__device__ inline uint32_t half_rnd_sds(uint64_t seed, uint64_t stream, uint32_t k, uint32_t M) {
    const uint32_t hl = nnm_lane() & 31u;
    uint32_t r = 0xFFFFFFFFu;
    if (hl < k && M > k) r = (uint32_t)(nnm_rand(seed, stream, hl) % (uint64_t)(M - k));
    r = bitonic_pass32<2, 1>(r);
    r = bitonic_pass32<4, 2>(r);
    r = bitonic_pass32<8, 4>(r);
    r = bitonic_pass32<16, 8>(r);
    r = bitonic_pass32<64, 16>(r);  // both halves ascending
    // bump inside the half: a'[i] = i + max_{j<=i}(a[j] - j)
    int32_t t = hl < k ? (int32_t)r - (int32_t)hl : -0x7FFFFFFF;
#pragma unroll
    for (uint32_t o = 1; o < 32; o <<= 1) {
        const int32_t u = __shfl_up(t, o, NNM_WAVE);
        if (hl >= o) t = t > u ? t : u;
    }
    const uint32_t off = (uint32_t)(nnm_rand(seed, stream, 0xFFFFu) % M);
    if (hl >= k) return M;
    if (M == k) return hl;
    return (uint32_t)(((uint64_t)(t + (int32_t)hl) + off) % M);
}

// Bits of the lane's half of a wave mask, in the low 32 bits.
__device__ inline uint64_t half_bits(uint64_t m) { return (nnm_lane() & 32u) ? m >> 32 : m & 0xFFFFFFFFull; }

// wave_random_row for the half's row (K + 1 <= 32); ok false: no row.
template <int DM = DM_ANY>
__device__ inline void half_random_row(const float* __restrict__ X, const uint16_t* __restrict__ Xb, uint32_t ld,
                                       uint64_t v, bool ok, uint32_t gb, uint32_t gsize, uint32_t K, uint64_t seed,
                                       uint32_t* __restrict__ ids, float* __restrict__ dist,
                                       uint8_t* __restrict__ flg, uint64_t* nd, const nnm_metric_args& mt) {
    const uint32_t hl = nnm_lane() & 31u;
    const uint32_t r = half_rnd_sds(seed, v, K + 1, gsize);
    const uint32_t id = gb + r;
    const uint64_t m = half_bits(nnm_ballot(hl < K + 1 && id != (uint32_t)v));
    // half-lane t < K takes the id of the half-lane holding the t-th kept draw
    const uint32_t src = (nnm_lane() & 32u) + nth_set_bit(m, hl < K ? hl : 0u);
    const uint32_t myid = (uint32_t)__shfl((int)id, (int)src, NNM_WAVE);
    float dd = 0.f;
    if (ok && hl < K) {
        dd = setup_dist<DM>(X, Xb, ld, v, myid, mt);
        *nd += 1;
    }
    half_write_sorted_row(nnm_key(dd, myid), 1u, ok, K, ids, dist, flg);
}

// Random initial lists, two rows per wave (K <= 31).
__global__ void k_random_init_half(const float* __restrict__ X, uint32_t ld, uint64_t n, uint32_t K,
                                   const uint32_t* __restrict__ vsub, const uint32_t* __restrict__ sub_grp,
                                   const uint32_t* __restrict__ grp_begin, const uint32_t* __restrict__ grp_end,
                                   uint64_t seed, uint32_t* __restrict__ g_ids, float* __restrict__ g_dist,
                                   uint8_t* __restrict__ g_flag, uint64_t* __restrict__ counters,
                                   nnm_metric_args mt) {
    uint64_t nd = 0;
    NNM_HALF_LOOP(v0, n) {
        const uint64_t v = v0 + (nnm_lane() >> 5);
        const bool ok = v < n;
        const uint32_t g = ok ? sub_grp[vsub[v]] : 0u;
        const uint32_t gb = ok ? grp_begin[g] : 0u, gsize = ok ? grp_end[g] - gb : K + 1;
        half_random_row(X, nullptr, ld, v, ok, gb, gsize, K, seed, g_ids + v * K, g_dist + v * K, g_flag + v * K,
                        &nd, mt);
    }
    wave_count(counters + CNT_DISTS, nd);
}

__global__ void k_random_init(const float* __restrict__ X, uint32_t ld, uint64_t n, uint32_t K,
                              const uint32_t* __restrict__ vsub, const uint32_t* __restrict__ sub_grp,
                              const uint32_t* __restrict__ grp_begin, const uint32_t* __restrict__ grp_end,
                              uint64_t seed, uint32_t* __restrict__ g_ids, float* __restrict__ g_dist,
                              uint8_t* __restrict__ g_flag, uint64_t* __restrict__ counters, nnm_metric_args mt) {
    uint64_t nd = 0;
    NNM_WAVE_LOOP(v, n) {
        const uint32_t g = sub_grp[vsub[v]];
        const uint32_t gb = grp_begin[g], gsize = grp_end[g] - gb;
        wave_random_row(X, nullptr, ld, v, gb, gsize, K, seed, g_ids + v * K, g_dist + v * K, g_flag + v * K, &nd,
                        mt);
    }
    wave_count(counters + CNT_DISTS, nd);
}

// Upload fix-up: optional distance recompute, optional flags, sort rows.
// skip_raw (optional): rows labelled raw are left as they are.
template <int DM>
__global__ void k_set_graph(const float* __restrict__ X, const uint16_t* __restrict__ Xb, uint32_t ld, uint64_t n,
                            uint32_t K, int recompute, int have_flags, const uint32_t* __restrict__ skip_raw,
                            uint32_t* __restrict__ g_ids, float* __restrict__ g_dist,
                            uint8_t* __restrict__ g_flag, uint64_t* __restrict__ counters,
                            uint32_t* __restrict__ bad, nnm_metric_args mt) {
    const uint32_t hl = nnm_lane() & 31u;
    uint64_t nd = 0;
    NNM_HALF_LOOP(v0, n) {
        const uint64_t v = v0 + (nnm_lane() >> 5);
        const bool ok = v < n && !(skip_raw && skip_raw[v] == NNM_RAW);
        if (!__ballot(ok)) continue;
        uint32_t id = 0;
        float dd = 0.f;
        uint32_t f = 1;
        if (ok && hl < K) {
            id = g_ids[v * K + hl];
            if (id >= n || id == (uint32_t)v) { atomicAdd(bad, 1u); id = (uint32_t)((v + 1) % n); }
            dd = recompute ? setup_dist<DM>(X, Xb, ld, v, id, mt) : g_dist[v * K + hl];
            f = have_flags ? g_flag[v * K + hl] : 1u;
            nd += recompute ? 1 : 0;
        }
        half_write_sorted_row(nnm_key(dd, id), f, ok, K, g_ids + v * K, g_dist + v * K, g_flag + v * K);
    }
    wave_count(counters + CNT_DISTS, nd);
}


// Learned phase entry: every list distance becomes the learned estimate
// kscale * max(|p_v|^2 + |p_u|^2 - 2 p_v.p_u, 0) on the bf16 projections (the
// learned join's key), rows re-sorted, flags kept.  NC > 0: rows of NC
// 16-byte chunks (dout = 8 NC), every chunk loaded before the products;
// NC = 0: any dout (a multiple of 8).  The products run in k order either way.
// REC (NC = 4 only): rows and norms from the 128-byte learned records lrec
// (one line per entry instead of a row and a norm gather; the same values).
template <uint32_t NC, bool REC = false>
__global__ void k_lists_estimate(uint64_t n, uint32_t K, const uint16_t* __restrict__ P, const float* __restrict__ pn,
                                 uint32_t dout, float kscale, uint32_t* __restrict__ g_ids,
                                 float* __restrict__ g_dist, uint8_t* __restrict__ g_flag,
                                 const uint4* __restrict__ lrec) {
    const uint32_t hl = nnm_lane() & 31u;
    NNM_HALF_LOOP(v0, n) {
        const uint64_t v = v0 + (nnm_lane() >> 5);
        const bool ok = v < n;
        uint32_t id = 0xFFFFFFFFu, f = 0;
        float dd = 3.402823466e38f;
        if (ok && hl < K) {
            id = g_ids[v * K + hl];
            f = g_flag[v * K + hl];
            constexpr bool rec = NC == 4 && REC;
            const uint4* a = rec ? lrec + v * 8 : reinterpret_cast<const uint4*>(P + v * dout);
            const uint4* b = rec ? lrec + (uint64_t)id * 8 : reinterpret_cast<const uint4*>(P + (uint64_t)id * dout);
            const float nv = rec ? __uint_as_float(lrec[v * 8 + 4].z) : pn[v];
            const float nu = rec ? __uint_as_float(lrec[(uint64_t)id * 8 + 4].z) : pn[id];
            float dot = 0.f;
            auto step = [&](const uint4 x, const uint4 y) {
                const uint32_t xs[4] = {x.x, x.y, x.z, x.w}, ys[4] = {y.x, y.y, y.z, y.w};
#pragma unroll
                for (int h = 0; h < 4; ++h) {
                    dot = __builtin_fmaf(__uint_as_float(xs[h] << 16), __uint_as_float(ys[h] << 16), dot);
                    dot = __builtin_fmaf(__uint_as_float(xs[h] & 0xFFFF0000u), __uint_as_float(ys[h] & 0xFFFF0000u),
                                         dot);
                }
            };
            if constexpr (NC > 0) {
                uint4 x[NC], y[NC];
#pragma unroll
                for (uint32_t k = 0; k < NC; ++k) {
                    x[k] = a[k];
                    y[k] = b[k];
                }
#pragma unroll
                for (uint32_t k = 0; k < NC; ++k) step(x[k], y[k]);
            } else {
                for (uint32_t k = 0; k < dout / 8; ++k) step(a[k], b[k]);
            }
            dd = fmaxf((nv + nu) - 2.f * dot, 0.f) * kscale;
        }
        half_write_sorted_row(nnm_key(dd, id), f, ok, K, g_ids + v * K, g_dist + v * K, g_flag + v * K);
    }
}

// Learned phase entry with FP8 rows: every list distance becomes the FP8
// estimate kscale8 * max(|q_v|^2 + |q_u|^2 - 2 q_v.q_u, 0) on the e4m3 rows
// of d_out = 32 (the FP8 learned join's key), rows re-sorted, flags kept.
__device__ inline float e4m3_value(uint32_t b) {
    const uint32_t e = (b >> 3) & 15u, m = b & 7u;
    const float v = e ? __uint_as_float(((e + 120u) << 23) | (m << 20)) : (float)m * 0.001953125f;
    return (b & 0x80u) ? -v : v;
}
__global__ void k_lists_estimate_fp8(uint64_t n, uint32_t K, const uint8_t* __restrict__ P8,
                                     const float* __restrict__ pn8, float kscale, uint32_t* __restrict__ g_ids,
                                     float* __restrict__ g_dist, uint8_t* __restrict__ g_flag) {
    const uint32_t hl = nnm_lane() & 31u;
    NNM_HALF_LOOP(v0, n) {
        const uint64_t v = v0 + (nnm_lane() >> 5);
        const bool ok = v < n;
        uint32_t id = 0xFFFFFFFFu, f = 0;
        float dd = 3.402823466e38f;
        if (ok && hl < K) {
            id = g_ids[v * K + hl];
            f = g_flag[v * K + hl];
            const uint4* a = reinterpret_cast<const uint4*>(P8 + v * 32);
            const uint4* b = reinterpret_cast<const uint4*>(P8 + (uint64_t)id * 32);
            const uint4 x0 = a[0], x1 = a[1], y0 = b[0], y1 = b[1];
            const uint32_t xs[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
            const uint32_t ys[8] = {y0.x, y0.y, y0.z, y0.w, y1.x, y1.y, y1.z, y1.w};
            float dot = 0.f;
#pragma unroll
            for (int w = 0; w < 8; ++w)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    dot = __builtin_fmaf(e4m3_value((xs[w] >> (8 * j)) & 0xFFu), e4m3_value((ys[w] >> (8 * j)) & 0xFFu),
                                         dot);
            dd = fmaxf((pn8[v] + pn8[id]) - 2.f * dot, 0.f) * kscale;
        }
        half_write_sorted_row(nnm_key(dd, id), f, ok, K, g_ids + v * K, g_dist + v * K, g_flag + v * K);
    }
}

// Learned phase exit: exact FP32 distances for every list entry (from the
// bf16 rows where those are exact), rows re-sorted, flags kept.  Counted as
// exact distances outside the join.
// prior (optional): the lists when the learned phase started; an entry whose
// id is not in its row's prior list was found by the learned phase and is
// marked new.  prior_dist (optional, with prior): the exact distances of the
// prior lists; an entry that was there keeps its exact distance from them
// (the same function of the same rows, so the same bits) and only the
// entries the learned phase found are computed.
__global__ void k_lists_exact(uint64_t n, uint32_t K, const float* __restrict__ X, const uint16_t* __restrict__ Xb_exact,
                              const int8_t* __restrict__ Xi8, uint32_t ld, int renew, const uint32_t* __restrict__ prior,
                              const float* __restrict__ prior_dist, uint32_t* __restrict__ g_ids,
                              float* __restrict__ g_dist, uint8_t* __restrict__ g_flag,
                              uint64_t* __restrict__ counters) {
    const uint32_t lane = nnm_lane(), hl = lane & 31u, hbase = lane & 32u, wv = threadIdx.x / NNM_WAVE;
    // the wave's two prior rows (ids, distances) for the membership test
    __shared__ __align__(16) uint32_t s_pid[256 / NNM_WAVE][NNM_WAVE];
    __shared__ float s_pd[256 / NNM_WAVE][NNM_WAVE];
    uint64_t nd = 0;
    NNM_HALF_LOOP(v0, n) {
        const uint64_t v = v0 + (lane >> 5);
        const bool ok = v < n;
        const bool ent = ok && hl < K;
        uint32_t id = 0xFFFFFFFFu, f = 0, pr = 0xFFFFFFFFu;
        float dd = 3.402823466e38f, pdv = 0.f;
        if (ent) {
            id = g_ids[v * K + hl];
            f = renew ? 1u : g_flag[v * K + hl];
            if (prior) pr = prior[v * K + hl];
            if (prior_dist) pdv = prior_dist[v * K + hl];
        }
        bool found = false;
        float known = 0.f;
        if (prior) {
            // membership in the half's prior row: its ids four at a time from
            // LDS (entries past K hold ~0, which no entry's id equals)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();  // the previous rows' readers are done
            s_pid[wv][lane] = pr;
            s_pd[wv][lane] = pdv;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            uint32_t pos = NNM_WAVE;
            for (uint32_t p = 0; p < K; p += 4) {
                const uint4 q = *reinterpret_cast<const uint4*>(&s_pid[wv][hbase + p]);
                pos = q.x == id ? p : pos;
                pos = q.y == id ? p + 1 : pos;
                pos = q.z == id ? p + 2 : pos;
                pos = q.w == id ? p + 3 : pos;
            }
            found = ent && pos < NNM_WAVE;
            if (found) known = s_pd[wv][hbase + pos];
            if (!found) f = 1u;
        }
        if (ok && hl < K) {
            if (prior_dist && found) {
                dd = known;
            } else {
                dd = Xi8        ? nnm_l2_rows_i8(Xi8 + v * ld, Xi8 + (uint64_t)id * ld, ld)
                     : Xb_exact ? nnm_l2_rows_bf16(Xb_exact + v * ld, Xb_exact + (uint64_t)id * ld, ld)
                                : nnm_l2_rows(X + v * ld, X + (uint64_t)id * ld, ld);
                ++nd;
            }
        }
        half_write_sorted_row(nnm_key(dd, id), f, ok, K, g_ids + v * K, g_dist + v * K, g_flag + v * K);
    }
    wave_count(counters + CNT_RERANKED, nd);
}

// Merge set-up, one wave per row.
__global__ void k_merge_init(const float* __restrict__ X, const uint16_t* __restrict__ Xb, uint32_t ld, uint64_t n,
                             uint32_t K, uint32_t cut,
                             const uint32_t* __restrict__ vsub, const uint32_t* __restrict__ sub_begin,
                             const uint32_t* __restrict__ sub_grp, const uint8_t* __restrict__ sub_built,
                             const uint32_t* __restrict__ grp_begin, const uint32_t* __restrict__ grp_end,
                             uint64_t seed, uint32_t* __restrict__ g_ids, float* __restrict__ g_dist,
                             uint8_t* __restrict__ g_flag, uint32_t* __restrict__ c_ids,
                             float* __restrict__ c_dist, uint8_t* __restrict__ c_flag,
                             uint8_t* __restrict__ has_cpn, uint64_t* __restrict__ counters,
                             nnm_metric_args mt) {
    const uint32_t lane = nnm_lane();
    uint64_t nd = 0;
    NNM_WAVE_LOOP(v, n) {
        const uint32_t s = vsub[v];
        const uint32_t g = sub_grp[s];
        const uint32_t gb = grp_begin[g], gsize = grp_end[g] - gb;
        uint32_t* ids = g_ids + v * K;
        float* dist = g_dist + v * K;
        uint8_t* flg = g_flag + v * K;
        if (!sub_built[s]) {
            if (lane == 0) has_cpn[v] = 0;
            wave_random_row(X, Xb, ld, v, gb, gsize, K, seed ^ 0x5A5A5A5Aull, ids, dist, flg, &nd, mt);
        } else {
            const uint32_t sb = sub_begin[s], ssize = sub_begin[s + 1] - sb;
            const uint32_t remain = K - cut;
            uint32_t pid = 0, pf = 0;
            float pd = 0.f;
            if (lane < K) { pid = ids[lane]; pd = dist[lane]; pf = flg[lane]; }
            // park the far entries (cpnGraph)
            if (lane >= remain && lane < K) {
                c_ids[v * cut + (lane - remain)] = pid;
                c_dist[v * cut + (lane - remain)] = pd;
                c_flag[v * cut + (lane - remain)] = (uint8_t)pf;
            }
            if (lane == 0) has_cpn[v] = 1;
            const uint32_t others = gsize - ssize;
            uint32_t fresh_n = 0;
            uint32_t cand = 0;
            bool fresh = false;
            if (others > 0) {
                uint32_t draw = 2 * cut < others ? 2 * cut : others;
                uint32_t r = nnm_wave_rnd_sds(seed, v, draw, others);
                cand = r < (sb - gb) ? gb + r : gb + r + ssize;
                bool dup = false;
                for (uint32_t p = 0; p < K; ++p) {
                    const uint32_t q = wave_read32(pid, p);
                    dup |= (q == cand);
                }
                fresh = lane < draw && !dup;
                uint64_t m = nnm_ballot(fresh);
                uint32_t rank = nnm_mask_rank(m);
                fresh = fresh && rank < cut;
                fresh_n = (uint32_t)__popcll(m) < cut ? (uint32_t)__popcll(m) : cut;
                // route fresh candidate of rank j to lane remain + j
                uint32_t src = 0;
                uint64_t mm = m;
                for (uint32_t t = 0; t < fresh_n; ++t) {
                    uint32_t l = (uint32_t)__ffsll((long long)mm) - 1;
                    mm &= mm - 1;
                    if (lane == remain + t) src = l;
                }
                uint32_t got = (uint32_t)__shfl((int)cand, (int)src, NNM_WAVE);
                if (lane >= remain && lane < remain + fresh_n) {
                    pid = got;
                    pd = nnm_row_dist(X, Xb, ld, got, v, mt);
                    pf = 1;
                    nd += 1;
                }
            }
            (void)fresh;
            wave_write_sorted_row(nnm_key(pd, pid), pf, K, ids, dist, flg);
        }
    }
    wave_count(counters + CNT_DISTS, nd);
}

// k_merge_init with two rows per wave (K <= 31, so K + 1 draws fit a half):
// the same lists, parked entries and distance count.
template <int DM>
__global__ void k_merge_init_half(const float* __restrict__ X, const uint16_t* __restrict__ Xb, uint32_t ld,
                                  uint64_t n, uint32_t K, uint32_t cut,
                                  const uint32_t* __restrict__ vsub, const uint32_t* __restrict__ sub_begin,
                                  const uint32_t* __restrict__ sub_grp, const uint8_t* __restrict__ sub_built,
                                  const uint32_t* __restrict__ grp_begin, const uint32_t* __restrict__ grp_end,
                                  uint64_t seed, uint32_t* __restrict__ g_ids, float* __restrict__ g_dist,
                                  uint8_t* __restrict__ g_flag, uint32_t* __restrict__ c_ids,
                                  float* __restrict__ c_dist, uint8_t* __restrict__ c_flag,
                                  uint8_t* __restrict__ has_cpn, uint64_t* __restrict__ counters,
                                  nnm_metric_args mt) {
    const uint32_t lane = nnm_lane(), hl = lane & 31u, hbase = lane & 32u;
    uint64_t nd = 0;
    NNM_HALF_LOOP(v0, n) {
        const uint64_t v = v0 + (lane >> 5);
        const bool ok = v < n;
        const uint32_t s = ok ? vsub[v] : 0u;
        const uint32_t g = ok ? sub_grp[s] : 0u;
        const uint32_t gb = ok ? grp_begin[g] : 0u, gsize = ok ? grp_end[g] - gb : K + 1;
        const bool built = ok && sub_built[s];
        uint32_t* ids = g_ids + v * K;
        float* dist = g_dist + v * K;
        uint8_t* flg = g_flag + v * K;
        // the two halves may take different branches: every cross-lane step
        // below reads lanes of its own half only
        if (!built) {
            if (ok && hl == 0) has_cpn[v] = 0;
            half_random_row<DM>(X, Xb, ld, v, ok, gb, gsize, K, seed ^ 0x5A5A5A5Aull, ids, dist, flg, &nd, mt);
        } else {
            const uint32_t sb = sub_begin[s], ssize = sub_begin[s + 1] - sb;
            const uint32_t remain = K - cut;
            uint32_t pid = 0, pf = 0;
            float pd = 0.f;
            if (hl < K) { pid = ids[hl]; pd = dist[hl]; pf = flg[hl]; }
            // park the far entries (cpnGraph)
            if (hl >= remain && hl < K) {
                c_ids[v * cut + (hl - remain)] = pid;
                c_dist[v * cut + (hl - remain)] = pd;
                c_flag[v * cut + (hl - remain)] = (uint8_t)pf;
            }
            if (hl == 0) has_cpn[v] = 1;
            const uint32_t others = gsize - ssize;
            const uint32_t draw = 2 * cut < others ? 2 * cut : others;
            const uint32_t r = half_rnd_sds(seed, v, draw, others > 0 ? others : 1u);
            const uint32_t cand = r < (sb - gb) ? gb + r : gb + r + ssize;
            bool dup = false;
            for (uint32_t p = 0; p < K; ++p) {
                const uint32_t q0 = wave_read32(pid, p), q1 = wave_read32(pid, 32 + p);
                dup |= (hbase ? q1 : q0) == cand;
            }
            const bool fresh = hl < draw && !dup;
            const uint64_t m = half_bits(nnm_ballot(fresh));
            const uint32_t fresh_n = (uint32_t)__popcll(m) < cut ? (uint32_t)__popcll(m) : cut;
            // fresh candidate of rank j to half-lane remain + j
            const bool takes = hl >= remain && hl < remain + fresh_n;
            const uint32_t src = hbase + nth_set_bit(m, takes ? hl - remain : 0u);
            const uint32_t got = (uint32_t)__shfl((int)cand, (int)src, NNM_WAVE);
            if (takes) {
                pid = got;
                pd = setup_dist<DM>(X, Xb, ld, got, v, mt);
                pf = 1;
                nd += 1;
            }
            half_write_sorted_row(nnm_key(pd, pid), pf, true, K, ids, dist, flg);
        }
    }
    wave_count(counters + CNT_DISTS, nd);
}

// merge2one: pool (K, sorted) + parked (cut, sorted) -> first K by (dist, id).
__global__ void k_merge_finish(uint64_t n, uint32_t K, uint32_t cut, int dedup, const uint8_t* __restrict__ has_cpn,
                               uint32_t* __restrict__ g_ids, float* __restrict__ g_dist, uint8_t* __restrict__ g_flag,
                               const uint32_t* __restrict__ c_ids, const float* __restrict__ c_dist,
                               const uint8_t* __restrict__ c_flag) {
    const uint32_t lane = nnm_lane();
    NNM_WAVE_LOOP(v, n) {
        if (!has_cpn[v]) continue;
        uint64_t key = NNM_EMPTY_KEY;
        uint32_t f = 0;
        uint32_t id = 0xFFFFFFFFu;
        if (lane < K) {
            id = g_ids[v * K + lane];
            key = nnm_key(g_dist[v * K + lane], id);
            f = g_flag[v * K + lane];
        } else if (lane < K + cut) {
            id = c_ids[v * cut + lane - K];
            key = nnm_key(c_dist[v * cut + lane - K], id);
            f = c_flag[v * cut + lane - K];
        }
        if (dedup && lane >= K && lane < K + cut) {
            for (uint32_t p = 0; p < K; ++p) {
                // the list lanes are inactive here: read them through scalars
                const uint32_t q = wave_read32(id, p);
                if (q == id) key = NNM_EMPTY_KEY;
            }
        }
        uint64_t s = nnm_wave_sort64(key);
        uint32_t sid = nnm_key_id(s);
        uint32_t myf = 0;
        for (uint32_t p = 0; p < K + cut; ++p) {
            const uint32_t q = wave_read32(id, p);
            const uint32_t qf = wave_read32(f, p);
            if (q == sid) myf = qf;
        }
        if (lane < K) {
            g_ids[v * K + lane] = sid;
            g_dist[v * K + lane] = nnm_key_dist(s);
            g_flag[v * K + lane] = (uint8_t)myf;
        }
    }
}

// k_merge_finish with dedup and two rows per wave (K + cut <= 32): parked
// entries repeating a pool id are dropped first, so every kept key is
// distinct and each flag can travel with its key through the sort.
__global__ void k_merge_finish_half(uint64_t n, uint32_t K, uint32_t cut, const uint8_t* __restrict__ has_cpn,
                                    uint32_t* __restrict__ g_ids, float* __restrict__ g_dist,
                                    uint8_t* __restrict__ g_flag, const uint32_t* __restrict__ c_ids,
                                    const float* __restrict__ c_dist, const uint8_t* __restrict__ c_flag) {
    const uint32_t lane = nnm_lane(), hl = lane & 31u, hbase = lane & 32u;
    NNM_HALF_LOOP(v0, n) {
        const uint64_t v = v0 + (lane >> 5);
        const bool ok = v < n && has_cpn[v];
        if (!__ballot(ok)) continue;
        uint64_t key = NNM_EMPTY_KEY;
        uint32_t f = 0, id = 0xFFFFFFFFu;
        const bool parked = hl >= K && hl < K + cut;
        if (ok && hl < K) {
            id = g_ids[v * K + hl];
            key = nnm_key(g_dist[v * K + hl], id);
            f = g_flag[v * K + hl];
        } else if (ok && parked) {
            id = c_ids[v * cut + hl - K];
            key = nnm_key(c_dist[v * cut + hl - K], id);
            f = c_flag[v * cut + hl - K];
        }
        bool dup = false;
        for (uint32_t p = 0; p < K; ++p) {
            const uint32_t q0 = wave_read32(id, p), q1 = wave_read32(id, 32 + p);
            dup |= (hbase ? q1 : q0) == id;
        }
        if (parked && dup) key = NNM_EMPTY_KEY;
        half_sort_kv(key, f);
        if (ok && hl < K) {
            g_ids[v * K + hl] = nnm_key_id(key);
            g_dist[v * K + hl] = nnm_key_dist(key);
            g_flag[v * K + hl] = (uint8_t)f;
        }
    }
}

__global__ void k_pair_dists(const float* __restrict__ X, uint32_t ld, const uint32_t* __restrict__ a,
                             const uint32_t* __restrict__ b, uint64_t m, float* __restrict__ out,
                             nnm_metric_args mt) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    out[i] = nnm_row_dist(X, nullptr, ld, a[i], b[i], mt);
}

// ================================================================ graph API

static int ensure_graph(nnm_ctx* c, uint32_t K) {
    if (K == 0 || K > 32) return nnm_fail(c, NNM_ERR_INVALID, "K must be in [1, 32]");
    if (c->K == K && c->g_ids) return NNM_OK;
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    free_graph(c);
    free_iter(c);
    c->K = K;
    NNM_TRY(alloc_n(c, &c->g_ids, c->n * K));
    NNM_TRY(alloc_n(c, &c->g_dist, c->n * K));
    NNM_TRY(alloc_n(c, &c->g_flag, c->n * K));
    return NNM_OK;
}

static int check_groups(nnm_ctx* c, uint32_t K) {
    for (uint32_t g = 0; g < c->ngrp; ++g) {
        uint32_t s0 = c->h_grp_first_sub[g];
        uint32_t s1 = (g + 1 < c->ngrp) ? c->h_grp_first_sub[g + 1] : c->nsub;
        uint32_t size = c->h_sub_begin[s1] - c->h_sub_begin[s0];
        if (size < K + 1)
            return nnm_fail(c, NNM_ERR_INVALID, "every group needs at least K+1 rows");
    }
    return NNM_OK;
}

static void stats_reset(nnm_stats* st) {
    if (st) std::memset(st, 0, sizeof(*st));
}

static int read_counters(nnm_ctx* c, uint64_t* h) {
    NNM_TRY(nnm_d2h(c, h, c->counters, CNT_N * sizeof(uint64_t)));
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    return NNM_OK;
}

extern "C" int nnm_graph_random_init(nnm_ctx* c, uint32_t K, uint64_t seed, nnm_stats* st) {
    if (!c) return NNM_ERR_INVALID;
    if (!c->X) return nnm_fail(c, NNM_ERR_STATE, "no data");
    stats_reset(st);
    NNM_TRY(check_groups(c, K));
    NNM_TRY(ensure_graph(c, K));
    NNM_TRY(fill_u64(c, c->counters, CNT_N, 0));
    NNM_HIP(c, hipEventRecord(c->ev0, c->stream));
    if (K + 1 <= 32)  // two rows per wave
        hipLaunchKernelGGL(k_random_init_half, dim3(wave_grid(c, (c->n_act + 1) / 2)), dim3(256), 0, c->stream, c->X,
                           c->ld, c->n_act, K, c->vsub, c->sub_grp, c->grp_begin, c->grp_end, seed, c->g_ids,
                           c->g_dist, c->g_flag, c->counters, metric_args(c));
    else
        hipLaunchKernelGGL(k_random_init, dim3(wave_grid(c, c->n_act)), dim3(256), 0, c->stream, c->X, c->ld,
                           c->n_act, K, c->vsub, c->sub_grp, c->grp_begin, c->grp_end, seed, c->g_ids, c->g_dist,
                           c->g_flag, c->counters, metric_args(c));
    NNM_LAUNCH_CHECK(c);
    NNM_HIP(c, hipEventRecord(c->ev1, c->stream));
    uint64_t h[CNT_N];
    NNM_TRY(read_counters(c, h));
    if (st) {
        st->init_dists = h[CNT_DISTS];
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->ev0, c->ev1);
        st->secs = ms * 1e-3;
    }
    return NNM_OK;
}

extern "C" int nnm_set_graph(nnm_ctx* c, const uint32_t* ids, const float* dist, const uint8_t* flags, uint32_t K,
                             nnm_stats* st) {
    if (!c || !ids) return nnm_fail(c, NNM_ERR_INVALID, "nnm_set_graph: bad arguments");
    if (!c->X) return nnm_fail(c, NNM_ERR_STATE, "no data");
    stats_reset(st);
    NNM_TRY(ensure_graph(c, K));
    const uint64_t nk = c->n * K;
    NNM_TRY(nnm_h2d(c, c->g_ids, ids, nk * 4));
    if (dist) NNM_TRY(nnm_h2d(c, c->g_dist, dist, nk * 4));
    if (flags) NNM_TRY(nnm_h2d(c, c->g_flag, flags, nk));
    NNM_TRY(fill_u64(c, c->counters, CNT_N, 0));
    uint32_t* bad = reinterpret_cast<uint32_t*>(c->counters + CNT_TOTAL_REV);
    NNM_HIP(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(setup_dm(c, nullptr) == DM_ANY ? k_set_graph<DM_ANY> : k_set_graph<DM_L2_F32>,
                       dim3(wave_grid(c, (c->n + 1) / 2)), dim3(256), 0, c->stream, c->X, nullptr, c->ld, c->n, K,
                       dist ? 0 : 1, flags ? 1 : 0, nullptr, c->g_ids, c->g_dist, c->g_flag, c->counters, bad,
                       metric_args(c));
    NNM_LAUNCH_CHECK(c);
    NNM_HIP(c, hipEventRecord(c->ev1, c->stream));
    uint64_t h[CNT_N];
    NNM_TRY(read_counters(c, h));
    c->cpn_live = 0;
    if (st) {
        st->init_dists = h[CNT_DISTS];
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->ev0, c->ev1);
        st->secs = ms * 1e-3;
    }
    if ((uint32_t)h[CNT_TOTAL_REV] != 0)
        return nnm_fail(c, NNM_ERR_INVALID, "nnm_set_graph: ids out of range or self loops");
    return NNM_OK;
}

extern "C" int nnm_get_graph(nnm_ctx* c, uint32_t* ids, float* dist, uint8_t* flags) {
    if (!c) return NNM_ERR_INVALID;
    if (!c->g_ids) return nnm_fail(c, NNM_ERR_STATE, "no graph");
    const uint64_t nk = c->n * c->K;
    if (ids) NNM_TRY(nnm_d2h(c, ids, c->g_ids, nk * 4));
    if (dist) NNM_TRY(nnm_d2h(c, dist, c->g_dist, nk * 4));
    if (flags) NNM_TRY(nnm_d2h(c, flags, c->g_flag, nk));
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    return NNM_OK;
}

extern "C" int nnm_graph_save(nnm_ctx* c) {
    if (!c) return NNM_ERR_INVALID;
    if (!c->g_ids) return nnm_fail(c, NNM_ERR_STATE, "no graph");
    const uint64_t nk = c->n * c->K;
    if (!c->s_ids) {
        NNM_TRY(alloc_n(c, &c->s_ids, nk));
        NNM_TRY(alloc_n(c, &c->s_dist, nk));
        NNM_TRY(alloc_n(c, &c->s_flag, nk));
    }
    NNM_HIP(c, hipMemcpyAsync(c->s_ids, c->g_ids, nk * 4, hipMemcpyDeviceToDevice, c->stream));
    NNM_HIP(c, hipMemcpyAsync(c->s_dist, c->g_dist, nk * 4, hipMemcpyDeviceToDevice, c->stream));
    NNM_HIP(c, hipMemcpyAsync(c->s_flag, c->g_flag, nk, hipMemcpyDeviceToDevice, c->stream));
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    return NNM_OK;
}

extern "C" int nnm_graph_restore(nnm_ctx* c, int recompute, nnm_stats* st) {
    if (!c) return NNM_ERR_INVALID;
    if (!c->s_ids) return nnm_fail(c, NNM_ERR_STATE, "no saved graph");
    stats_reset(st);
    const uint64_t nk = c->n * c->K;
    NNM_TRY(fill_u64(c, c->counters, CNT_N, 0));
    NNM_HIP(c, hipEventRecord(c->ev0, c->stream));
    NNM_HIP(c, hipMemcpyAsync(c->g_ids, c->s_ids, nk * 4, hipMemcpyDeviceToDevice, c->stream));
    NNM_HIP(c, hipMemcpyAsync(c->g_dist, c->s_dist, nk * 4, hipMemcpyDeviceToDevice, c->stream));
    NNM_HIP(c, hipMemcpyAsync(c->g_flag, c->s_flag, nk, hipMemcpyDeviceToDevice, c->stream));
    if (recompute) {
        // InitSGraph (src/abstractmerge.cpp:55-76) recomputes the built
        // subgraphs' distances; raw rows are left to nnm_merge_init
        const uint16_t* xb = nullptr;
        if (recompute == 2 && c->data_exact_bf16 && c->metric == NNM_METRIC_L2) {
            NNM_TRY(nnm_ensure_bf16(c));
            xb = c->Xb;
        }
        uint32_t* bad = reinterpret_cast<uint32_t*>(c->counters + CNT_TOTAL_REV);
        int dm = setup_dm(c, xb);
        if (xb && c->data_i8) {  // the same exact distances from the int8 rows
            NNM_TRY(nnm_ensure_i8(c));
            xb = reinterpret_cast<const uint16_t*>(c->Xi8);
            dm = DM_L2_I8;
        }
        hipLaunchKernelGGL(dm == DM_L2_I8 ? k_set_graph<DM_L2_I8>
                           : dm == DM_L2_BF16 ? k_set_graph<DM_L2_BF16>
                                            : dm == DM_L2_F32 ? k_set_graph<DM_L2_F32> : k_set_graph<DM_ANY>,
                           dim3(wave_grid(c, (c->n_act + 1) / 2)), dim3(256), 0, c->stream, c->X, xb, c->ld, c->n_act,
                           c->K, 1, 0, c->label, c->g_ids, c->g_dist, c->g_flag, c->counters, bad, metric_args(c));
        NNM_LAUNCH_CHECK(c);
    }
    NNM_HIP(c, hipEventRecord(c->ev1, c->stream));
    uint64_t h[CNT_N];
    NNM_TRY(read_counters(c, h));
    c->cpn_live = 0;
    if (st) {
        st->init_dists = h[CNT_DISTS];
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->ev0, c->ev1);
        st->secs = ms * 1e-3;
    }
    return NNM_OK;
}

extern "C" int nnm_pair_dists(nnm_ctx* c, const uint32_t* a, const uint32_t* b, uint64_t m, float* out) {
    if (!c || !a || !b || !out) return nnm_fail(c, NNM_ERR_INVALID, "nnm_pair_dists: bad arguments");
    if (!c->X) return nnm_fail(c, NNM_ERR_STATE, "no data");
    if (m == 0) return NNM_OK;
    for (uint64_t i = 0; i < m; ++i)
        if (a[i] >= c->n || b[i] >= c->n) return nnm_fail(c, NNM_ERR_INVALID, "nnm_pair_dists: row out of range");
    uint32_t *da = nullptr, *db = nullptr;
    float* dout = nullptr;
    NNM_TRY(alloc_n(c, &da, m));
    NNM_TRY(alloc_n(c, &db, m));
    NNM_TRY(alloc_n(c, &dout, m));
    NNM_TRY(nnm_h2d(c, da, a, m * 4));
    NNM_TRY(nnm_h2d(c, db, b, m * 4));
    hipLaunchKernelGGL(k_pair_dists, dim3(grid_for(m, 256, 1u << 30)), dim3(256), 0, c->stream, c->X, c->ld, da,
                       db, m, dout, metric_args(c));
    NNM_LAUNCH_CHECK(c);
    NNM_TRY(nnm_d2h(c, out, dout, m * 4));
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    nnm_free(c, da);
    nnm_free(c, db);
    nnm_free(c, dout);
    return NNM_OK;
}

// ================================================================ merges

extern "C" int nnm_merge_init(nnm_ctx* c, const nnm_params* p, nnm_stats* st) {
    if (!c || !p) return NNM_ERR_INVALID;
    if (!c->X) return nnm_fail(c, NNM_ERR_STATE, "no data");
    stats_reset(st);
    const uint32_t K = p->K;
    const uint32_t cut = p->cut ? p->cut : K / 2;
    if (cut == 0 || cut >= K) return nnm_fail(c, NNM_ERR_INVALID, "cut must be in [1, K)");
    if (K + cut > NNM_WAVE) return nnm_fail(c, NNM_ERR_INVALID, "K + cut must be <= 64");
    NNM_TRY(check_groups(c, K));
    bool any_built = false;
    for (uint8_t b : c->h_sub_built) any_built |= (b != 0);
    if (any_built && (!c->g_ids || c->K != K))
        return nnm_fail(c, NNM_ERR_STATE, "nnm_merge_init: built subgraphs need a graph with the same K");
    NNM_TRY(ensure_graph(c, K));
    if (c->cut != cut || !c->c_ids) {
        nnm_free(c, c->c_ids); nnm_free(c, c->c_dist); nnm_free(c, c->c_flag); nnm_free(c, c->has_cpn);
        c->c_ids = nullptr; c->c_dist = nullptr; c->c_flag = nullptr; c->has_cpn = nullptr;
        c->cut = cut;
        NNM_TRY(alloc_n(c, &c->c_ids, c->n * cut));
        NNM_TRY(alloc_n(c, &c->c_dist, c->n * cut));
        NNM_TRY(alloc_n(c, &c->c_flag, c->n * cut));
        NNM_TRY(alloc_n(c, &c->has_cpn, c->n));
    }
    const uint16_t* xb = nullptr;  // bf16 rows where exact, for the bf16 / learned precisions
    if (p->prec != NNM_PREC_FP32 && c->data_exact_bf16 && c->metric == NNM_METRIC_L2) {
        NNM_TRY(nnm_ensure_bf16(c));
        xb = c->Xb;
    }
    NNM_TRY(fill_u64(c, c->counters, CNT_N, 0));
    NNM_HIP(c, hipEventRecord(c->ev0, c->stream));
    int dm = setup_dm(c, xb);
    if (xb && c->data_i8 && K + 1 <= 32) {
        // int8 rows (two rows per wave kernel): the same exact distances from
        // half the bytes
        NNM_TRY(nnm_ensure_i8(c));
        xb = reinterpret_cast<const uint16_t*>(c->Xi8);
        dm = DM_L2_I8;
    }
    if (K + 1 <= 32)  // two rows per wave
        hipLaunchKernelGGL(dm == DM_L2_I8 ? k_merge_init_half<DM_L2_I8>
                           : dm == DM_L2_BF16  ? k_merge_init_half<DM_L2_BF16>
                           : dm == DM_L2_F32 ? k_merge_init_half<DM_L2_F32>
                                             : k_merge_init_half<DM_ANY>,
                           dim3(wave_grid(c, (c->n_act + 1) / 2)), dim3(256), 0, c->stream, c->X,
                           xb, c->ld, c->n_act, K, cut, c->vsub, c->sub_begin, c->sub_grp, c->sub_built,
                           c->grp_begin, c->grp_end, p->seed * 0x9E3779B97F4A7C15ull + 17, c->g_ids, c->g_dist,
                           c->g_flag, c->c_ids, c->c_dist, c->c_flag, c->has_cpn, c->counters, metric_args(c));
    else
        hipLaunchKernelGGL(k_merge_init, dim3(wave_grid(c, c->n_act)), dim3(256), 0, c->stream, c->X, xb, c->ld,
                           c->n_act, K, cut, c->vsub, c->sub_begin, c->sub_grp, c->sub_built, c->grp_begin,
                           c->grp_end, p->seed * 0x9E3779B97F4A7C15ull + 17, c->g_ids, c->g_dist, c->g_flag,
                           c->c_ids, c->c_dist, c->c_flag, c->has_cpn, c->counters, metric_args(c));
    NNM_LAUNCH_CHECK(c);
    NNM_HIP(c, hipEventRecord(c->ev1, c->stream));
    uint64_t h[CNT_N];
    NNM_TRY(read_counters(c, h));
    c->cpn_live = 1;
    if (st) {
        st->init_dists = h[CNT_DISTS];
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->ev0, c->ev1);
        st->secs = ms * 1e-3;
    }
    return NNM_OK;
}

extern "C" int nnm_merge_finish(nnm_ctx* c, int dedup) {
    if (!c) return NNM_ERR_INVALID;
    if (!c->cpn_live) return nnm_fail(c, NNM_ERR_STATE, "nnm_merge_finish: no merge in progress");
    if (dedup && c->K + c->cut <= 32)  // two rows per wave
        hipLaunchKernelGGL(k_merge_finish_half, dim3(wave_grid(c, (c->n_act + 1) / 2)), dim3(256), 0, c->stream,
                           c->n_act, c->K, c->cut, c->has_cpn, c->g_ids, c->g_dist, c->g_flag, c->c_ids, c->c_dist,
                           c->c_flag);
    else
        hipLaunchKernelGGL(k_merge_finish, dim3(wave_grid(c, c->n_act)), dim3(256), 0, c->stream, c->n_act, c->K,
                           c->cut, dedup, c->has_cpn, c->g_ids, c->g_dist, c->g_flag, c->c_ids, c->c_dist,
                           c->c_flag);
    NNM_LAUNCH_CHECK(c);
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    c->cpn_live = 0;
    return NNM_OK;
}

// ================================================================ list build

// k-th distance of every list, compact (n floats stay in the last-level cache
// for the random lookups of the list build and the join).
// k-th list distance per vertex, and the join's packed per-vertex record
// {label, k-th distance x offer window, norm of the joined rows, 0}: one
// 16-byte random read per candidate-list entry in the join.
// lrec (optional): the record goes to bytes 64..79 of the vertex's 128-byte
// learned record instead (the learned join then reads it there).
__global__ void k_kth(uint64_t n, uint32_t K, const float* __restrict__ g_dist, const uint32_t* __restrict__ label,
                      const float* __restrict__ norms, float win, float* __restrict__ kth, uint4* __restrict__ vmeta,
                      uint4* __restrict__ lrec) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < n) {
        const float k = g_dist[v * K + K - 1];
        kth[v] = k;
        const uint4 r = make_uint4(label[v], __float_as_uint(k * win), norms ? __float_as_uint(norms[v]) : 0u, 0u);
        if (lrec) lrec[v * 8 + 4] = r;
        else vmeta[v] = r;
    }
}

// Learned records: the 64-byte projected row of every vertex into bytes
// 0..63 of its 128-byte record and its squared norm into the join record's
// norm field (bytes 72..75; k_kth rewrites the record with the same norm).
// A gathered row, its norm and its join record then share one 128-byte line.
// One thread per 16-byte chunk of the record's first 80 bytes.
__global__ void k_lrec_rows(uint64_t n, const uint4* __restrict__ P, const float* __restrict__ pn,
                            uint4* __restrict__ lrec) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n * 5; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = i / 5;
        const uint32_t c = (uint32_t)(i % 5);
        lrec[v * 8 + c] = c < 4 ? P[v * 4 + c] : make_uint4(0u, 0u, __float_as_uint(pn[v]), 0u);
    }
}

// List build, pass 1 (src/abstractmerge.cpp:94-134): one lane per list
// entry, G lanes per row.  Flagged entries become new samples (the first S
// in list order, flag cleared), unflagged ones old; an entry (v -> u) whose
// distance exceeds u's k-th distance also goes to u's reverse list (new or
// old): an atomic counter gives its position in u's fixed bucket of NNM_RB
// entries.  Entries past the bucket are collected by the overflow pass.
template <uint32_t G>
__global__ __launch_bounds__(256) void k_lb_build(uint64_t n, uint32_t K, uint32_t S, const uint32_t* __restrict__ g_ids,
                                                  const float* __restrict__ g_dist, const float* __restrict__ kthv,
                                                  uint8_t* __restrict__ g_flag, uint32_t* __restrict__ l_new,
                                                  uint32_t cap_new, uint32_t* __restrict__ l_old, uint32_t cap_old,
                                                  uint32_t* __restrict__ n_new, uint32_t* __restrict__ n_old,
                                                  uint64_t* __restrict__ newmask, uint32_t* __restrict__ rcnt,
                                                  uint32_t* __restrict__ rbuf, uint64_t* __restrict__ spill,
                                                  uint64_t* __restrict__ counters) {
    constexpr uint32_t GPW = NNM_WAVE / G;  // rows per wave
    const uint32_t lane = nnm_lane(), grp = lane / G, gl = lane % G;
    const uint64_t gmask = (G == NNM_WAVE) ? ~0ull : (((1ull << G) - 1) << (grp * G));
    const uint64_t below = (1ull << lane) - 1;
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / NNM_WAVE;
    const uint64_t nw = (uint64_t)gridDim.x * blockDim.x / NNM_WAVE;
    // Two row groups per step (v0 and v0 + nw * GPW), their loads, k-th
    // gathers and bucket atomics issued together: the step waits for each
    // round trip once for both.
    constexpr int R = 2;
    uint64_t kept_new = 0;  // new flags left in the lists (the delta test's base)
    for (uint64_t v0 = wid * GPW; v0 < n; v0 += (uint64_t)R * nw * GPW) {
        uint64_t v[R];
        bool act[R];
        uint32_t u[R];
        float de[R];
        uint8_t f[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            v[k] = v0 + (uint64_t)k * nw * GPW + grp;
            act[k] = v[k] < n && gl < K;
            u[k] = 0;
            de[k] = 0.f;
            f[k] = 0;
            if (act[k]) {
                u[k] = g_ids[v[k] * K + gl];
                de[k] = g_dist[v[k] * K + gl];
                f[k] = g_flag[v[k] * K + gl];
            }
        }
        bool rev[R];
        float ku[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            const uint64_t fm = nnm_ballot(act[k] && f[k]) & gmask;
            const uint64_t om = nnm_ballot(act[k] && !f[k]) & gmask;
            const uint32_t frank = (uint32_t)__popcll(fm & below);
            const bool is_new = act[k] && f[k] && frank < S;
            const bool is_old = act[k] && !f[k];
            if (is_new) {
                l_new[v[k] * cap_new + frank] = u[k];
                g_flag[v[k] * K + gl] = 0;
            }
            // new entries past the sample stay new; the mark of an entry
            // that entered in the last delivery (2) is stored as 1
            const bool stays = act[k] && f[k] && !is_new;
            if (stays && f[k] == 2) g_flag[v[k] * K + gl] = 1;
            kept_new += (uint32_t)__popcll(nnm_ballot(stays));
            if (is_old) l_old[v[k] * cap_old + (uint32_t)__popcll(om & below)] = u[k];
            const uint64_t nm = nnm_ballot(is_new) & gmask;
            if (gl == 0 && v[k] < n) {
                n_new[v[k]] = min(S, (uint32_t)__popcll(fm));
                n_old[v[k]] = (uint32_t)__popcll(om);
                newmask[v[k]] = (G == NNM_WAVE) ? nm : (nm >> (grp * G));
            }
            // bucket of the reverse entry: new entries' targets first, old second
            rev[k] = is_new || is_old;
            f[k] = is_new ? 1 : 0;
            ku[k] = rev[k] ? kthv[u[k]] : 0.f;
        }
        uint32_t pos[R];
        uint64_t b[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            rev[k] = rev[k] && ku[k] < de[k];
            b[k] = f[k] ? (uint64_t)u[k] : n + u[k];
            pos[k] = rev[k] ? atomicAdd(&rcnt[b[k]], 1u) : 0u;
        }
#pragma unroll
        for (int k = 0; k < R; ++k) {
            if (rev[k] && pos[k] < NNM_RB) rbuf[b[k] * NNM_RB + pos[k]] = (uint32_t)v[k];
            // entries past their bucket: a bit per list entry, one word per row
            const uint64_t sm = nnm_ballot(rev[k] && pos[k] >= NNM_RB) & gmask;
            if (gl == 0 && v[k] < n) spill[v[k]] = (G == NNM_WAVE) ? sm : (sm >> (grp * G));
        }
    }
    if (nnm_lane() == 0 && kept_new) atomicAdd((unsigned long long*)&counters[CNT_FLAGS], (unsigned long long)kept_new);
}

// Entries past each full bucket: their count (scanned into offsets into
// the overflow entries).
__global__ void k_lb_ovf_mark(uint64_t m, const uint32_t* __restrict__ rcnt, uint32_t* __restrict__ ocnt) {
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b < m; b += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t c = rcnt[b];
        ocnt[b] = c > NNM_RB ? c - NNM_RB : 0u;
    }
}

// List build, overflow pass: the entries past their bucket (marked per row
// by the build) grouped by bucket.  Order inside a bucket is free: the
// reverse cap picks by seeded key.
__global__ void k_lb_spill_fill(uint64_t n, uint32_t K, const uint32_t* __restrict__ g_ids,
                                const uint64_t* __restrict__ newmask, const uint64_t* __restrict__ spill,
                                const uint32_t* __restrict__ roff, uint32_t* __restrict__ rcur,
                                uint32_t* __restrict__ obuf) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t m = spill[v];
        if (!m) continue;
        const uint64_t nm = newmask[v];
        for (; m; m &= m - 1) {
            const uint32_t gl = (uint32_t)__builtin_ctzll(m);
            const uint32_t u = g_ids[v * K + gl];
            const uint64_t b = ((nm >> gl) & 1ull) ? (uint64_t)u : n + u;
            const uint32_t pos = atomicAdd(&rcur[b], 1u);
            obuf[roff[b] + pos] = (uint32_t)v;
        }
    }
}

// Append the reverse entries to the candidate lists.  A bucket's entries
// are its first NNM_RB in rbuf, then the rest in the overflow entries.  More
// than cap entries: keep the cap entries with the smallest random key (a
// seeded, order-free stand-in for random_shuffle + resize,
// src/abstractmerge.cpp:149-159).
__device__ inline uint32_t wave_append_reverse(const uint32_t* __restrict__ head, const uint32_t* __restrict__ rest,
                                               uint32_t cnt, uint32_t cap, uint32_t salt,
                                               uint32_t* __restrict__ dst) {
    const uint32_t lane = nnm_lane();
    auto at = [&](uint32_t i) { return i < NNM_RB ? head[i] : rest[i - NNM_RB]; };
    if (cnt <= cap) {
        for (uint32_t i = lane; i < cnt; i += NNM_WAVE) dst[i] = at(i);
        return cnt;
    }
    // find the cap-th smallest 64-bit key (hash << 32 | id), bit by bit
    uint64_t res = 0;
    for (int b = 63; b >= 0; --b) {
        const uint64_t trial = res + (1ull << b);
        uint32_t below = 0;
        for (uint32_t i = lane; i < cnt; i += NNM_WAVE) {
            const uint32_t id = at(i);
            const uint64_t key = ((uint64_t)nnm_hash32(id, salt) << 32) | id;
            below += key < trial ? 1u : 0u;
        }
        below = nnm_wave_sum32(below);
        if (below < cap) res = trial;
    }
    // res is the cap-th smallest key: keep keys <= res (exactly cap of them)
    uint32_t out = 0;
    for (uint32_t base = 0; base < cnt; base += NNM_WAVE) {
        const uint32_t i = base + lane;
        bool keep = false;
        uint32_t id = 0;
        if (i < cnt) {
            id = at(i);
            const uint64_t key = ((uint64_t)nnm_hash32(id, salt) << 32) | id;
            keep = key <= res;
        }
        const uint64_t m = nnm_ballot(keep);
        if (keep) dst[out + nnm_mask_rank(m)] = id;
        out += (uint32_t)__popcll(m);
    }
    return out;
}

// Four rows per wave: a quarter copies its row's reverse lists when both fit
// the cap (the row's counts and list sizes are read together: one round trip
// before the copy); a row with a list above the cap is then taken by the
// whole wave (the seeded selection).
__global__ void k_lb_select(uint64_t n, const uint32_t* __restrict__ rcnt, const uint32_t* __restrict__ roff,
                            const uint32_t* __restrict__ rbuf, const uint32_t* __restrict__ obuf, uint32_t cap,
                            uint64_t seed,
                            uint32_t* __restrict__ l_new, uint32_t cap_new, uint32_t* __restrict__ l_old,
                            uint32_t cap_old, uint32_t* __restrict__ n_new, uint32_t* __restrict__ n_old,
                            const uint64_t* __restrict__ counters) {
    const uint32_t lane = nnm_lane(), ql = lane & 15u;
    // sampling salt of this iteration (the iteration number lives on the device)
    const uint32_t salt = (uint32_t)(nnm_rand(seed, 0xC0FFEEull, counters[CNT_ITER]) >> 32);
    const uint64_t wpb = blockDim.x / NNM_WAVE;
    for (uint64_t v0 = 4 * ((uint64_t)blockIdx.x * wpb + threadIdx.x / NNM_WAVE); v0 < n;
         v0 += 4 * (uint64_t)gridDim.x * wpb) {
        const uint64_t v = v0 + (lane >> 4);
        const bool ok = v < n;
        const uint64_t vc = ok ? v : n - 1;  // in bounds: the loads need no branch
        const uint32_t cn = rcnt[vc], co = rcnt[n + vc], fn = n_new[vc], fo = n_old[vc];
        const bool big = ok && (cn > cap || co > cap);
        if (ok && !big) {
            const uint32_t* hn = rbuf + v * NNM_RB;
            const uint32_t* ho = rbuf + (n + v) * NNM_RB;
            const uint32_t* rn = cn > NNM_RB ? obuf + roff[v] : nullptr;
            const uint32_t* ro = co > NNM_RB ? obuf + roff[n + v] : nullptr;
            uint32_t* dn = l_new + v * cap_new + fn;
            uint32_t* dold = l_old + v * cap_old + fo;
            for (uint32_t i = ql; i < cn; i += 16) dn[i] = i < NNM_RB ? hn[i] : rn[i - NNM_RB];
            for (uint32_t i = ql; i < co; i += 16) dold[i] = i < NNM_RB ? ho[i] : ro[i - NNM_RB];
            if (ql == 0) {
                n_new[v] = fn + cn;
                n_old[v] = fo + co;
            }
        }
        for (uint64_t bb = nnm_ballot(big && ql == 0); bb; bb &= bb - 1) {
            const uint64_t w = v0 + ((uint32_t)__builtin_ctzll(bb) >> 4);
            const uint32_t wfn = n_new[w], wfo = n_old[w];
            const uint32_t wn = rcnt[w], wo = rcnt[n + w];
            const uint32_t* rn = wn > NNM_RB ? obuf + roff[w] : nullptr;
            const uint32_t* ro = wo > NNM_RB ? obuf + roff[n + w] : nullptr;
            const uint32_t an = wave_append_reverse(rbuf + w * NNM_RB, rn, wn, cap, nnm_hash32((uint32_t)w, salt),
                                                    l_new + w * cap_new + wfn);
            const uint32_t ao = wave_append_reverse(rbuf + (n + w) * NNM_RB, ro, wo, cap,
                                                    nnm_hash32((uint32_t)w, salt ^ 0xA511E9B3u),
                                                    l_old + w * cap_old + wfo);
            if (lane == 0) {
                n_new[w] = wfn + an;
                n_old[w] = wfo + ao;
            }
        }
    }
}

// Exclusive scan of 2n counts (three passes: tile sums, scan of sums, add).
#define SCAN_TILE 2048
__global__ void k_scan_tiles(const uint32_t* __restrict__ in, uint64_t m, uint32_t* __restrict__ tile_sum) {
    __shared__ uint32_t red[256];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    uint32_t s = 0;
    for (uint32_t i = threadIdx.x; i < SCAN_TILE; i += 256)
        if (base + i < m) s += in[base + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = red[0];
}

__global__ void k_scan_sums(uint32_t* __restrict__ tile_sum, uint32_t ntiles, uint32_t* __restrict__ total) {
    // single block of 1024 threads, sequential chunks
    __shared__ uint32_t buf[1024];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += 1024) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < ntiles ? tile_sum[i] : 0;
        buf[threadIdx.x] = v;
        __syncthreads();
        for (uint32_t o = 1; o < 1024; o <<= 1) {
            uint32_t t = threadIdx.x >= o ? buf[threadIdx.x - o] : 0;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < ntiles) tile_sum[i] = carry + buf[threadIdx.x] - v;
        carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ void k_scan_apply(const uint32_t* __restrict__ in, uint64_t m, const uint32_t* __restrict__ tile_off,
                             uint32_t* __restrict__ out) {
    __shared__ uint32_t wsum[256];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
    // each thread owns 8 consecutive elements
    const uint32_t t = threadIdx.x;
    uint32_t loc[8];
    uint32_t s = 0;
    for (int j = 0; j < 8; ++j) {
        uint64_t i = base + t * 8 + j;
        loc[j] = i < m ? in[i] : 0;
        s += loc[j];
    }
    wsum[t] = s;
    __syncthreads();
    for (uint32_t o = 1; o < 256; o <<= 1) {
        uint32_t x = t >= o ? wsum[t - o] : 0;
        __syncthreads();
        wsum[t] += x;
        __syncthreads();
    }
    uint32_t run = tile_off[blockIdx.x] + wsum[t] - s;
    for (int j = 0; j < 8; ++j) {
        uint64_t i = base + t * 8 + j;
        if (i < m) out[i] = run;
        run += loc[j];
    }
}

// ================================================================ FP64 error report

// FP64 squared L2 across the wave: lanes take elements lane, lane + 64, ...;
// the partial sums are combined in a fixed tree (same result every call).
__device__ inline double wave_l2_f64(const float* __restrict__ a, const float* __restrict__ b, uint32_t d) {
    double acc = 0.0;
    for (uint32_t i = nnm_lane(); i < d; i += NNM_WAVE) {
        const double t = (double)a[i] - (double)b[i];
        acc = __builtin_fma(t, t, acc);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, NNM_WAVE);
    return acc;
}

float nnm_near_tie_window(nnm_ctx* c, const nnm_params* p) {
    if (p->prec == NNM_PREC_LEARNED) return 1.f;  // learned estimates: no FP64 decision
    if (c->metric != NNM_METRIC_L2) return 1.f;  // the FP64 re-decision computes L2
    if (c->data_exact_bf16) return 1.f;  // integer rows: every distance is exact
    return 1.f + 2.f * (float)(c->d + 2) * 5.9604645e-8f;
}

// FP64 error of the stored FP32 list distances (per entry |d32 - d64|, and
// adjacent entries whose FP64 order disagrees with the stored order).
__global__ void k_dist_error(uint64_t n, uint32_t K, const float* __restrict__ X, uint32_t ld,
                             const uint32_t* __restrict__ g_ids, const float* __restrict__ g_dist,
                             unsigned long long* __restrict__ out) {
    const uint32_t lane = nnm_lane();
    double max_abs = 0.0, max_rel = 0.0, sum_rel = 0.0;
    uint64_t flips = 0;
    NNM_WAVE_LOOP(v, n) {
        double prev = -1.0;
        for (uint32_t j = 0; j < K; ++j) {
            const uint32_t id = g_ids[v * K + j];
            const double e = wave_l2_f64(X + v * ld, X + (uint64_t)id * ld, ld);
            const double err = fabs((double)g_dist[v * K + j] - e);
            max_abs = err > max_abs ? err : max_abs;
            const double rel = e > 0.0 ? err / e : (err > 0.0 ? 1.0 : 0.0);
            max_rel = rel > max_rel ? rel : max_rel;
            sum_rel += rel;
            if (e < prev) ++flips;
            prev = e;
        }
    }
    if (lane == 0) {
        // non-negative doubles order like their bit patterns
        atomicMax(&out[0], (unsigned long long)__double_as_longlong(max_abs));
        atomicMax(&out[1], (unsigned long long)__double_as_longlong(max_rel));
        atomicAdd(&out[2], (unsigned long long)(sum_rel * 1e12));
        atomicAdd(&out[3], (unsigned long long)flips);
    }
}

extern "C" int nnm_graph_dist_error(nnm_ctx* c, nnm_dist_error* e) {
    if (!c || !e) return NNM_ERR_INVALID;
    if (!c->X || !c->g_ids) return nnm_fail(c, NNM_ERR_STATE, "no data or graph");
    NNM_TRY(nnm_require_l2(c, "nnm_graph_dist_error"));
    unsigned long long* acc = nullptr;
    NNM_TRY(nnm_alloc(c, (void**)&acc, 4 * 8));
    int rc = fill_u64(c, reinterpret_cast<uint64_t*>(acc), 4, 0);
    if (rc == NNM_OK) {
        hipLaunchKernelGGL(k_dist_error, dim3(wave_grid(c, c->n_act)), dim3(256), 0, c->stream, c->n_act, c->K, c->X,
                           c->ld, c->g_ids, c->g_dist, acc);
        hipError_t he = hipGetLastError();
        if (he != hipSuccess) rc = nnm_hip_fail(c, he, "k_dist_error launch");
    }
    uint64_t h[4] = {0, 0, 0, 0};
    if (rc == NNM_OK) rc = nnm_d2h(c, h, acc, sizeof(h));
    nnm_free(c, acc);
    if (rc != NNM_OK) return rc;
    const double ent = (double)c->n_act * c->K;
    std::memcpy(&e->max_abs, &h[0], 8);
    std::memcpy(&e->max_rel, &h[1], 8);
    e->mean_rel = ent > 0 ? (double)h[2] * 1e-12 / ent : 0.0;
    e->entries = (uint64_t)ent;
    e->order_flips = h[3];
    return NNM_OK;
}

// ================================================================ NN-Descent

int nnm_ensure_iter_buffers(nnm_ctx* c, const nnm_params* p) {
    const uint32_t K = c->K;
    // the learned phase may sample more new entries than the plain iterations
    const float rho = p->prec == NNM_PREC_LEARNED ? std::max(p->rho, p->learn_rho) : p->rho;
    uint32_t S = (uint32_t)std::lround((double)rho * K);
    S = std::max<uint32_t>(1, std::min<uint32_t>(S, K));
    const uint32_t R = p->rev_cap;
    const uint32_t cap_new = S + R, cap_old = K + R;
    if (c->l_new && c->cap_new == cap_new && c->cap_old == cap_old) return nnm_offer_buffers(c, p);
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    free_iter(c);
    c->cap_new = cap_new;
    c->cap_old = cap_old;
    c->rev_cap = R;
    const uint64_t n = c->n;
    NNM_TRY(alloc_n(c, &c->l_new, n * cap_new));
    NNM_TRY(alloc_n(c, &c->l_old, n * cap_old));
    NNM_TRY(alloc_n(c, &c->n_new, n));
    NNM_TRY(alloc_n(c, &c->n_old, n));
    NNM_TRY(alloc_n(c, &c->newmask, n));
    NNM_TRY(alloc_n(c, &c->rcnt, 2 * n));
    NNM_TRY(alloc_n(c, &c->roff, 2 * n + 1));
    NNM_TRY(alloc_n(c, &c->rcur, 2 * n));
    NNM_TRY(alloc_n(c, &c->rbuf, 2 * n * NNM_RB));
    NNM_TRY(alloc_n(c, &c->ocnt, 2 * n));
    NNM_TRY(alloc_n(c, &c->obuf, n * K));
    NNM_TRY(alloc_n(c, &c->spill, n));
    NNM_TRY(alloc_n(c, &c->scan_tmp, (2 * n + SCAN_TILE - 1) / SCAN_TILE + 1));
    NNM_TRY(alloc_n(c, &c->kth, n));
    NNM_TRY(alloc_n(c, &c->vmeta, n));
    return nnm_offer_buffers(c, p);
}

static int validate_params(nnm_ctx* c, const nnm_params* p) {
    if (!p) return nnm_fail(c, NNM_ERR_INVALID, "null params");
    if (!c->X) return nnm_fail(c, NNM_ERR_STATE, "no data");
    if (!c->g_ids) return nnm_fail(c, NNM_ERR_STATE, "no graph");
    if (p->K != c->K) return nnm_fail(c, NNM_ERR_INVALID, "params K differs from the graph's K");
    if (!(p->rho > 0.f && p->rho <= 1.f)) return nnm_fail(c, NNM_ERR_INVALID, "rho must be in (0, 1]");
    if (!(p->learn_rho >= 0.f && p->learn_rho <= 1.f))
        return nnm_fail(c, NNM_ERR_INVALID, "learn_rho must be 0 (rho) or in (0, 1]");
    if (p->rev_cap == 0 || p->rev_cap > 4096) return nnm_fail(c, NNM_ERR_INVALID, "rev_cap must be in [1, 4096]");
    if (p->offer_cap != 0 && (p->offer_cap < 4 || p->offer_cap > 65536))
        return nnm_fail(c, NNM_ERR_INVALID, "offer_cap must be 0 (default) or in [4, 65536]");
    if (p->rule > NNM_RULE_MERGE) return nnm_fail(c, NNM_ERR_INVALID, "unknown rule");
    if (p->prec > NNM_PREC_LEARNED) return nnm_fail(c, NNM_ERR_INVALID, "unknown precision");
    if (p->rerank > NNM_RERANK_NEVER) return nnm_fail(c, NNM_ERR_INVALID, "unknown rerank policy");
    if (p->flags & ~(NNM_FLAG_GRAPH | NNM_FLAG_AUTOTUNE | NNM_FLAG_LEARNED_RENEW | NNM_FLAG_LEARNED_FP8 |
                     NNM_FLAG_LEARNED_MARK_FOUND))
        return nnm_fail(c, NNM_ERR_INVALID, "unknown flags");
    return NNM_OK;
}

int nnm_scan_u32(nnm_ctx* c, const uint32_t* in, uint64_t m, uint32_t* out, uint32_t* tmp) {
    const uint32_t ntiles = (uint32_t)((m + SCAN_TILE - 1) / SCAN_TILE);
    if (ntiles == 0) return nnm_fill_u32(c, out, 1, 0);
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(256), 0, c->stream, in, m, tmp);
    NNM_LAUNCH_CHECK(c);
    hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(1024), 0, c->stream, tmp, ntiles, out + m);
    NNM_LAUNCH_CHECK(c);
    hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(256), 0, c->stream, in, m, tmp, out);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

// Candidate lists of one iteration: forward samples, reverse lists, caps.
static int build_lists(nnm_ctx* c, const nnm_params* p) {
    const uint64_t n = c->n_act;
    const uint32_t K = c->K;
    uint32_t S = (uint32_t)std::lround((double)p->rho * K);
    S = std::max<uint32_t>(1, std::min<uint32_t>(S, K));
    NNM_TRY(fill_u32(c, c->rcnt, 2 * n, 0));
    NNM_TRY(fill_u32(c, c->rcur, 2 * n, 0));
    // row norms in the join's metadata: bf16 norms (norm form), learned
    // projection norms, or the cosine metric's squared norms
    const bool fp8 = p->prec == NNM_PREC_LEARNED && (p->flags & NNM_FLAG_LEARNED_FP8);
    const float* nrm = nnm_use_i8(c, p) ? c->norms_i8
                       : p->prec == NNM_PREC_BF16 ? c->norms
                       : fp8 ? nnm_mlp_norms8(c)
                       : p->prec == NNM_PREC_LEARNED ? nnm_mlp_norms(c)
                       : c->metric == NNM_METRIC_COSINE ? c->sumsq : nullptr;
    // learned join on 32-d bf16 rows: rows and join records in one 128-byte
    // record per vertex, prepared before the iterations (no allocation here:
    // an iteration may be captured into a graph)
    c->lrec_live = p->prec == NNM_PREC_LEARNED && !fp8 && nnm_mlp_dout(c) == 32 && c->lrec && c->lrec_n == c->n &&
                   c->lrec_gen == nnm_mlp_proj_gen(c) && nnm_mlp_proj_gen(c) != 0;
    hipLaunchKernelGGL(k_kth, dim3(grid_for(n, 256, 1u << 30)), dim3(256), 0, c->stream, n, K, c->g_dist, c->label,
                       nrm, nnm_near_tie_window(c, p), c->kth, c->vmeta, c->lrec_live ? c->lrec : nullptr);
    NNM_LAUNCH_CHECK(c);
    const uint32_t lb_grid = wave_grid(c, (n + 1) / 2);
    if (K <= 32) {
        hipLaunchKernelGGL(k_lb_build<32>, dim3(lb_grid), dim3(256), 0, c->stream, n, K, S, c->g_ids, c->g_dist,
                           c->kth, c->g_flag, c->l_new, c->cap_new, c->l_old, c->cap_old, c->n_new, c->n_old,
                           c->newmask, c->rcnt, c->rbuf, c->spill, c->counters);
    } else {
        hipLaunchKernelGGL(k_lb_build<64>, dim3(wave_grid(c, n)), dim3(256), 0, c->stream, n, K, S, c->g_ids,
                           c->g_dist, c->kth, c->g_flag, c->l_new, c->cap_new, c->l_old, c->cap_old, c->n_new,
                           c->n_old, c->newmask, c->rcnt, c->rbuf, c->spill, c->counters);
    }
    NNM_LAUNCH_CHECK(c);
    const uint64_t m = 2 * n;
    hipLaunchKernelGGL(k_lb_ovf_mark, dim3(grid_for(m, 256, 8192)), dim3(256), 0, c->stream, m, c->rcnt, c->ocnt);
    NNM_LAUNCH_CHECK(c);
    NNM_TRY(nnm_scan_u32(c, c->ocnt, m, c->roff, c->scan_tmp));
    hipLaunchKernelGGL(k_lb_spill_fill, dim3(grid_for(n, 256, 8192)), dim3(256), 0, c->stream, n, K, c->g_ids,
                       c->newmask, c->spill, c->roff, c->rcur, c->obuf);
    NNM_LAUNCH_CHECK(c);
    hipLaunchKernelGGL(k_lb_select, dim3(wave_grid(c, (n + 3) / 4)), dim3(256), 0, c->stream, n, c->rcnt, c->roff,
                       c->rbuf, c->obuf, c->rev_cap, p->seed, c->l_new, c->cap_new, c->l_old, c->cap_old, c->n_new,
                       c->n_old, c->counters);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

static int reset_queues(nnm_ctx* c) {
    NNM_TRY(fill_u64(c, c->counters + CNT_WORK, 1, 0));
    NNM_TRY(fill_u64(c, c->counters + CNT_WORK2, 2, 0));
    return NNM_OK;
}

// join_ms[0] += whole join, join_ms[1] += its short-list kernel (evj0..evjm),
// join_ms[2] += offer delivery (evj1..evd1)
static int add_join_times(nnm_ctx* c, double* join_ms) {
    NNM_HIP(c, hipEventSynchronize(c->evd1));
    float ms = 0.f, mm = 0.f, md = 0.f;
    hipEventElapsedTime(&ms, c->evj0, c->evj1);
    hipEventElapsedTime(&mm, c->evj0, c->evjm);
    hipEventElapsedTime(&md, c->evj1, c->evd1);
    join_ms[0] += ms;
    join_ms[1] += mm;
    join_ms[2] += md;
    return NNM_OK;
}

// One iteration; leaves the flag count in counters[CNT_FLAGS] (exact unless
// the join left vertices for a redo round, counters[CNT_REDO] > 0).  The
// iteration number is read from counters[CNT_ITER], which the caller sets.
// join_ms == nullptr: no events or waits, so the sequence can be captured.
static int run_iteration(nnm_ctx* c, const nnm_params* p, uint32_t js, uint32_t chunk, double* join_ms) {
    NNM_TRY(fill_u64(c, c->counters + CNT_FLAGS, 1, 0));
    NNM_TRY(reset_queues(c));
    NNM_TRY(nnm_offer_reset(c));
    NNM_TRY(build_lists(c, p));
    if (join_ms) NNM_HIP(c, hipEventRecord(c->evj0, c->stream));
    NNM_TRY(nnm_launch_join(c, p, js, chunk, join_ms ? c->evjm : nullptr));
    if (join_ms) NNM_HIP(c, hipEventRecord(c->evj1, c->stream));
    NNM_TRY(nnm_offer_deliver(c, p));
    if (join_ms) {
        NNM_HIP(c, hipEventRecord(c->evd1, c->stream));
        NNM_TRY(add_join_times(c, join_ms));
    }
    return NNM_OK;
}

// Redo rounds: join again the vertices whose offers did not fit the log
// (counters[CNT_REDO] after a round), deliver, until none is left.  The
// owners of the round that leaves none count the iteration's updates and
// flags.
static int redo_rounds(nnm_ctx* c, const nnm_params* p, uint32_t js, uint32_t chunk, uint64_t nwork,
                       uint32_t* rounds) {
    uint32_t round = 1;
    while (nwork) {
        NNM_TRY(reset_queues(c));
        NNM_TRY(nnm_offer_reset(c));
        NNM_TRY(nnm_launch_join(c, p, js, chunk, nullptr, round, nwork));
        NNM_TRY(nnm_offer_deliver(c, p));
        uint64_t left = 0;
        NNM_TRY(nnm_d2h(c, &left, c->counters + CNT_REDO, sizeof(left)));
        ++*rounds;
        if (left >= nwork)
            return nnm_fail(c, NNM_ERR_NOMEM, "offer log too small for the offers of single vertices; raise offer_cap");
        nwork = left;
        ++round;
    }
    return NNM_OK;
}

// One iteration that times every join launch shape on the lists it builds
// and keeps the fastest for this precision.  Every trial starts from an
// empty offer log, so the last launch leaves the log one launch writes; the
// work counters are reset before and the tallies restored after each trial.
static int autotune_iteration(nnm_ctx* c, const nnm_params* p, double* join_ms) {
    NNM_TRY(fill_u64(c, c->counters + CNT_FLAGS, 1, 0));
    NNM_TRY(reset_queues(c));
    NNM_TRY(build_lists(c, p));
    uint64_t h[CNT_N];
    NNM_TRY(read_counters(c, h));
    std::vector<std::pair<uint32_t, uint32_t>> space;
    nnm_join_tune_space(c, p->prec, &space);
    uint32_t best_js = 0, best_ch = 0;
    float best = 3.4e38f;
    for (int rep = 0; rep < 2; ++rep) {
        for (const auto& cand : space) {
            NNM_TRY(reset_queues(c));
            NNM_TRY(nnm_offer_reset(c));
            NNM_HIP(c, hipEventRecord(c->evj0, c->stream));
            NNM_TRY(nnm_launch_join(c, p, cand.first, cand.second));
            NNM_HIP(c, hipEventRecord(c->evj1, c->stream));
            NNM_HIP(c, hipEventSynchronize(c->evj1));
            float ms = 0.f;
            hipEventElapsedTime(&ms, c->evj0, c->evj1);
            if (rep == 1 && ms < best) { best = ms; best_js = cand.first; best_ch = cand.second; }
        }
    }
    c->tune_js[p->prec] = best_js;
    c->tune_chunk[p->prec] = best_ch;
    c->tune_ld[p->prec] = c->ld;
    for (int i : {CNT_DISTS, CNT_LIST_ENTRIES, CNT_OFFERS, CNT_LONG_DISTS, CNT_LONG_ENTRIES, CNT_LONG_OFFERS})
        NNM_TRY(fill_u64(c, c->counters + i, 1, h[i]));
    NNM_TRY(reset_queues(c));
    NNM_TRY(nnm_offer_reset(c));
    NNM_HIP(c, hipEventRecord(c->evj0, c->stream));
    NNM_TRY(nnm_launch_join(c, p, best_js, best_ch, c->evjm));
    NNM_HIP(c, hipEventRecord(c->evj1, c->stream));
    NNM_TRY(nnm_offer_deliver(c, p));
    NNM_HIP(c, hipEventRecord(c->evd1, c->stream));
    return add_join_times(c, join_ms);
}

// The marks of entries that entered in the last delivery (flag 2) stored as
// plain new flags (1), for readers after the loop.
__global__ void k_flag_norm(uint64_t m, uint8_t* __restrict__ g_flag) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        if (g_flag[i] == 2) g_flag[i] = 1;
}

// Instantiated iteration graphs of one descent call (one per join precision).
struct IterGraphs {
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    ~IterGraphs() {
        for (auto& e : exec)
            if (e) hipGraphExecDestroy(e);
    }
};

// The learned records' rows from the current projection (allocated and
// copied outside the iterations, once per projection).
static int prepare_lrec(nnm_ctx* c) {
    if (!c->lrec || c->lrec_n != c->n) {
        nnm_free(c, c->lrec);
        c->lrec = nullptr;
        c->lrec_n = c->lrec_gen = 0;
        NNM_TRY(alloc_n(c, &c->lrec, c->n * 8));
        c->lrec_n = c->n;
    }
    if (c->lrec_gen != nnm_mlp_proj_gen(c)) {
        hipLaunchKernelGGL(k_lrec_rows, dim3(grid_for(c->n * 5, 256, 8192)), dim3(256), 0, c->stream, c->n,
                           reinterpret_cast<const uint4*>(nnm_mlp_rows(c)), nnm_mlp_norms(c), c->lrec);
        NNM_LAUNCH_CHECK(c);
        c->lrec_gen = nnm_mlp_proj_gen(c);
    }
    return NNM_OK;
}

static int capture_iteration(nnm_ctx* c, const nnm_params* p, uint32_t js, uint32_t chunk, hipGraphExec_t* out) {
    NNM_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    const int rc = run_iteration(c, p, js, chunk, nullptr);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != NNM_OK) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return nnm_hip_fail(c, e, "hipStreamEndCapture");
    const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) {
        *out = nullptr;
        return nnm_hip_fail(c, ei, "hipGraphInstantiate");
    }
    return NNM_OK;
}

static int descent(nnm_ctx* c, const nnm_params* p, uint32_t only_iter, bool single, nnm_stats* st) {
    NNM_TRY(validate_params(c, p));
    if (p->prec != NNM_PREC_FP32) NNM_TRY(nnm_require_l2(c, "BF16 and learned joins"));
    NNM_TRY(nnm_ensure_iter_buffers(c, p));
    if (p->prec == NNM_PREC_LEARNED && c->cap_new + c->cap_old > nnm_join_learned_max_list())
        return nnm_fail(c, NNM_ERR_UNSUPPORTED,
                        "learned join: candidate lists (round(rho K) + K + 2 rev_cap) longer than 256 entries");
    // learned prefilter: needs a trained model; otherwise the plain BF16 join
    // runs every iteration (the fallback)
    bool learned = p->prec == NNM_PREC_LEARNED && nnm_mlp_trained(c);
    if (learned) NNM_TRY(nnm_mlp_ensure_projection(c, nullptr));
    const bool fp8 = learned && (p->flags & NNM_FLAG_LEARNED_FP8);
    const bool mark_found = learned && (p->flags & NNM_FLAG_LEARNED_MARK_FOUND) && !(p->flags & NNM_FLAG_LEARNED_RENEW);
    if (mark_found && !c->p_ids) NNM_TRY(alloc_n(c, &c->p_ids, c->n * c->K));
    // the exact distances of the lists at the learned phase's start, reused
    // at the switch for the entries still there (one call runs the whole
    // phase, so the lists it copies are exact)
    const bool reuse_dist = mark_found && !single;
    if (reuse_dist && !c->p_dist) NNM_TRY(alloc_n(c, &c->p_dist, c->n * c->K));
    if (fp8) {
        if (nnm_mlp_dout(c) != 32) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "FP8 learned join built for d_out = 32");
        NNM_TRY(nnm_mlp_ensure_fp8(c));
    }
    if (p->prec != NNM_PREC_FP32) NNM_TRY(nnm_ensure_bf16(c));
    // int8 rows for the plain joins on integer data (the learned run's tail too)
    if (p->prec != NNM_PREC_FP32 && c->data_i8 && c->metric == NNM_METRIC_L2) NNM_TRY(nnm_ensure_i8(c));
    if (learned && !fp8 && nnm_mlp_dout(c) == 32) NNM_TRY(prepare_lrec(c));
    stats_reset(st);
    NNM_TRY(fill_u64(c, c->counters, CNT_N, 0));
    NNM_HIP(c, hipEventRecord(c->ev0, c->stream));
    double join_ms[3] = {0.0, 0.0, 0.0};
    uint32_t redo_rounds_run = 0;
    uint64_t redo_rows = 0;
    uint32_t it = 0, n_learned = 0, n_plain = 0, n_graph = 0;
    uint32_t used_js = 0, used_chunk = 0;
    uint64_t prev_se = 0, prev_sd = 0, prev_so = 0, main_bytes = 0, main_flops = 0;
    uint64_t prev_dall = 0, learned_dists = 0;
    int converged = 0;
    double last_delta = 1.0;
    const bool autotune = (p->flags & NNM_FLAG_AUTOTUNE) != 0;
    const bool graph = (p->flags & NNM_FLAG_GRAPH) != 0;
    IterGraphs graphs;
    // Learned runs: the learned phase runs NN-Descent on the learned
    // estimates (the lists hold estimates, the join gathers 32-d projected
    // rows), then every list distance becomes exact once and the plain BF16
    // join finishes: the last plain_tail iterations of a capped run, and every
    // iteration after the learned phase meets the delta test.  The plain join
    // is the fallback and judges what the learned phase missed.
    bool tail = !learned;
    bool estimates = false;  // the lists hold learned estimates
    auto to_estimates = [&]() -> int {
        // the lists as the learned phase finds them, for marking its finds later
        if (mark_found)
            NNM_HIP(c, hipMemcpyAsync(c->p_ids, c->g_ids, (size_t)c->n_act * c->K * sizeof(uint32_t),
                                      hipMemcpyDeviceToDevice, c->stream));
        if (reuse_dist)
            NNM_HIP(c, hipMemcpyAsync(c->p_dist, c->g_dist, (size_t)c->n_act * c->K * sizeof(float),
                                      hipMemcpyDeviceToDevice, c->stream));
        const uint32_t dout = nnm_mlp_dout(c);
        if (fp8)
            hipLaunchKernelGGL(k_lists_estimate_fp8, dim3(wave_grid(c, (c->n_act + 1) / 2)), dim3(256), 0, c->stream,
                               c->n_act, c->K, nnm_mlp_rows8(c), nnm_mlp_norms8(c), nnm_mlp_scale8(c), c->g_ids,
                               c->g_dist, c->g_flag);
        else {
            const bool rec = dout == 32 && c->lrec && c->lrec_n == c->n && c->lrec_gen == nnm_mlp_proj_gen(c) &&
                             nnm_mlp_proj_gen(c) != 0;
            hipLaunchKernelGGL(rec ? (k_lists_estimate<4, true>) : dout == 32 ? k_lists_estimate<4> : k_lists_estimate<0>,
                               dim3(wave_grid(c, (c->n_act + 1) / 2)), dim3(256), 0, c->stream, c->n_act, c->K,
                               nnm_mlp_rows(c), nnm_mlp_norms(c), dout, nnm_mlp_scale(c), c->g_ids, c->g_dist,
                               c->g_flag, rec ? c->lrec : nullptr);
        }
        NNM_LAUNCH_CHECK(c);
        estimates = true;
        return NNM_OK;
    };
    auto to_exact = [&]() -> int {
        if (c->data_i8 && c->metric == NNM_METRIC_L2) NNM_TRY(nnm_ensure_i8(c));
        hipLaunchKernelGGL(k_lists_exact, dim3(wave_grid(c, (c->n_act + 1) / 2)), dim3(256), 0, c->stream, c->n_act, c->K, c->X,
                           (c->data_exact_bf16 && c->Xb) ? c->Xb : nullptr, c->data_i8 ? c->Xi8 : nullptr, c->ld,
                           (p->flags & NNM_FLAG_LEARNED_RENEW) ? 1 : 0, mark_found ? c->p_ids : nullptr,
                           reuse_dist ? c->p_dist : nullptr, c->g_ids,
                           c->g_dist, c->g_flag, c->counters);
        NNM_LAUNCH_CHECK(c);
        estimates = false;
        return NNM_OK;
    };
    for (;;) {
        const uint32_t iter = single ? only_iter : it;
        nnm_params q = *p;
        if (p->prec == NNM_PREC_LEARNED) {
            if (p->max_iters && iter + p->plain_tail >= p->max_iters) tail = true;
            q.prec = tail ? NNM_PREC_BF16 : NNM_PREC_LEARNED;
            if (!tail && p->learn_rho > 0.f) q.rho = p->learn_rho;
            if (q.prec == NNM_PREC_LEARNED && !estimates) NNM_TRY(to_estimates());
            if (q.prec != NNM_PREC_LEARNED && estimates) NNM_TRY(to_exact());
        }
        if (q.prec == NNM_PREC_LEARNED) ++n_learned; else ++n_plain;
        NNM_TRY(fill_u64(c, c->counters + CNT_ITER, 1, iter));
        const bool tuned = c->tune_ld[q.prec] == c->ld;
        const uint32_t js = (autotune && tuned) ? c->tune_js[q.prec] : 0;
        const uint32_t chunk = (autotune && tuned) ? c->tune_chunk[q.prec] : 0;
        if (autotune && !tuned) {
            NNM_TRY(autotune_iteration(c, &q, join_ms));
        } else if (graph) {
            hipGraphExec_t& ge = graphs.exec[q.prec];
            if (!ge) NNM_TRY(capture_iteration(c, &q, js, chunk, &ge));
            NNM_HIP(c, hipGraphLaunch(ge, c->stream));
            ++n_graph;
        } else {
            NNM_TRY(run_iteration(c, &q, js, chunk, join_ms));
        }
        used_js = autotune ? c->tune_js[q.prec] : js;
        used_chunk = autotune ? c->tune_chunk[q.prec] : chunk;
        uint64_t h[CNT_N];
        NNM_TRY(read_counters(c, h));
        if (h[CNT_REDO]) {
            // the offer log filled up: the left vertices join again
            redo_rows += h[CNT_REDO];
            NNM_TRY(redo_rounds(c, &q, used_js, used_chunk, h[CNT_REDO], &redo_rounds_run));
            NNM_TRY(read_counters(c, h));
        }
        {
            // algorithmic traffic and work of the short-list join kernel in
            // this iteration: per list entry one row, its id, label and
            // k-th distance (and norm for the norm form); per offer one
            // 12-byte log record; per vertex the two list sizes
            const uint64_t se = (h[CNT_LIST_ENTRIES] - h[CNT_LONG_ENTRIES]) - prev_se;
            const uint64_t sd = (h[CNT_DISTS] - h[CNT_LONG_DISTS]) - prev_sd;
            const uint64_t so = (h[CNT_OFFERS] - h[CNT_LONG_OFFERS]) - prev_so;
            prev_se += se;
            prev_sd += sd;
            prev_so += so;
            // pairs joined on learned estimates (projected distances)
            if (q.prec == NNM_PREC_LEARNED) learned_dists += h[CNT_DISTS] - prev_dall;
            prev_dall = h[CNT_DISTS];
            uint64_t row_bytes, flops_per_dist;
            if (q.prec == NNM_PREC_FP32) {
                row_bytes = (uint64_t)c->d * 4 + 12;
                flops_per_dist = 3ull * c->d;
            } else if (q.prec == NNM_PREC_LEARNED) {
                row_bytes = (uint64_t)nnm_mlp_dout(c) * (fp8 ? 1 : 2) + 16;
                flops_per_dist = 2ull * nnm_mlp_dout(c);
            } else {
                // bf16 rows, or int8 rows on integer data
                row_bytes = (uint64_t)c->d * (nnm_use_i8(c, &q) ? 1 : 2) + 16;
                flops_per_dist = 2ull * c->d;
            }
            main_bytes += se * row_bytes + so * 12 + c->n_act * 8;
            main_flops += sd * flops_per_dist;
        }
        ++it;
        last_delta = (double)h[CNT_FLAGS] / ((double)c->n_act * c->K);
        if (last_delta < p->delta) {
            if (q.prec == NNM_PREC_LEARNED && p->plain_tail > 0) {
                tail = true;
            } else {
                converged = 1;
                break;
            }
        }
        if (single) break;
        if (p->max_iters && it >= p->max_iters) break;
    }
    if (estimates) NNM_TRY(to_exact());  // a run that ended in the learned phase
    {
        const uint64_t m = c->n_act * c->K;
        hipLaunchKernelGGL(k_flag_norm, dim3(grid_for(m, 256, 8192)), dim3(256), 0, c->stream, m, c->g_flag);
        NNM_LAUNCH_CHECK(c);
    }
    NNM_HIP(c, hipEventRecord(c->ev1, c->stream));
    uint64_t h[CNT_N];
    NNM_TRY(read_counters(c, h));
    if (st) {
        st->iters = it;
        st->converged = converged;
        st->dists = h[CNT_DISTS];
        st->updates = h[CNT_UPDATES];
        st->reranked = h[CNT_RERANKED];
        st->list_entries = h[CNT_LIST_ENTRIES];
        st->offers = h[CNT_OFFERS];
        st->last_delta = last_delta;
        st->learned_iters = n_learned;
        st->plain_iters = n_plain;
        st->graph_iters = n_graph;
        st->tile_js = used_js ? used_js : nnm_join_default_js(c, p->prec == NNM_PREC_FP32 ? NNM_PREC_FP32 : NNM_PREC_BF16);
        st->tile_chunk = used_chunk ? used_chunk : nnm_join_default_chunk();
        float ms = 0.f;
        hipEventElapsedTime(&ms, c->ev0, c->ev1);
        st->secs = ms * 1e-3;
        st->join_secs = join_ms[0] * 1e-3;
        st->join_main_secs = join_ms[1] * 1e-3;
        st->long_dists = h[CNT_LONG_DISTS];
        st->long_entries = h[CNT_LONG_ENTRIES];
        st->long_offers = h[CNT_LONG_OFFERS];
        st->join_main_bytes = main_bytes;
        st->join_main_flops = main_flops;
        st->near_ties = h[CNT_NEAR_TIES];
        st->redo_rows = redo_rows;
        st->redo_rounds = redo_rounds_run;
        st->deliver_secs = join_ms[2] * 1e-3;
        st->learned_dists = learned_dists;
        st->tie_swaps = h[CNT_TIE_SWAPS];
    }
    return NNM_OK;
}

extern "C" int nnm_nndescent(nnm_ctx* c, const nnm_params* p, nnm_stats* st) {
    if (!c) return NNM_ERR_INVALID;
    return descent(c, p, 0, false, st);
}

extern "C" int nnm_local_join(nnm_ctx* c, const nnm_params* p, uint32_t iter, nnm_stats* st) {
    if (!c) return NNM_ERR_INVALID;
    return descent(c, p, iter, true, st);
}

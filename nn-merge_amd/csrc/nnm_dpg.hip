// DPG layer builder: LyNNDescent::dpg / DelaunayGraph
// (src/lynndescent.cpp:502-540, 640-710) on the context's k-NN lists.
//
// Phase 1 (occlusion prune, src/lynndescent.cpp:648-676), one wave per row:
// walk the row's k-NN list in order and keep a neighbour unless a kept entry
// is closer to it than the row is; kept lists are bounded sorted lists with
// the reference's insert2topk rule (src/lynndescent.cpp:548-623).  All the
// occlusion distances a row can need (kept-so-far x candidate) are computed
// by the wave up front into LDS, then the walk runs on them.
//
// Phase 2 (reverse edges, src/lynndescent.cpp:678-707): every kept edge
// i -> u offers i to u.  Offers are grouped by target (count, scan, fill);
// one wave per target applies them in increasing source order, against the
// target's list: while it has room the source is inserted, else only if no
// entry of the list is closer to the source than the target is.  The
// reference applies the same rule in a single pass over the rows, where a
// row's own list may already hold reverse entries when its turn comes; here
// every offer comes from the phase-1 lists, which makes the pass order-free
// (the oracle's orc_dpg_par states the same algorithm).
//
// Carry: the reference never clears LyNNDescent::graph between dpg calls, so
// rows present in the previous layer start from that layer's lists; with
// carry = 1 the previous nnm_dpg_build result is the starting point.

#include "nnm_internal.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

constexpr uint32_t DPG_MAXK = 64;   // longest k-NN list
constexpr uint32_t DPG_MAXCAP = 64; // longest DPG list

struct nnm_dpg_state {
    uint32_t rows = 0, cap = 0;
    uint32_t* cnt = nullptr;  // rows
    uint32_t* ids = nullptr;  // rows x cap, each list sorted by distance
    float* d = nullptr;
};

// insert2topk (src/lynndescent.cpp:548-623) on a list in LDS; run by one lane.
// r: optional per-entry payload moved with the entries.
__device__ inline void insert2topk(uint32_t* ids, float* ds, uint32_t* r, uint32_t& n, uint32_t nid, float dst,
                                   uint32_t rnew, uint32_t maxSize) {
    if (n == 0) {
        ids[0] = nid;
        ds[0] = dst;
        if (r) r[0] = rnew;
        n = 1;
        return;
    }
    const uint32_t topk = n;
    if (ds[topk - 1] < dst && topk == maxSize) return;
    if (topk < maxSize && ds[topk - 1] < dst) {
        ids[topk] = nid;
        ds[topk] = dst;
        if (r) r[topk] = rnew;
        n = topk + 1;
        return;
    }
    uint32_t i = topk;
    while (i > 0 && ds[i - 1] > dst) --i;
    for (uint32_t l = i; l > 0; --l) {
        if (ds[l - 1] < dst) break;
        if (ids[l - 1] == nid) return;
    }
    uint32_t j;
    if (topk < maxSize) {
        j = topk;
        n = topk + 1;
    } else {
        j = maxSize - 1;
    }
    while (j > i) {
        ids[j] = ids[j - 1];
        ds[j] = ds[j - 1];
        if (r) r[j] = r[j - 1];
        --j;
    }
    ids[j] = nid;
    ds[j] = dst;
    if (r) r[j] = rnew;
}

__device__ inline void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wave (block of 64) per row; dynamic LDS: D[(cap + K) * K] floats.
__global__ __launch_bounds__(64) void k_dpg_occlude(uint32_t n, uint32_t K, const uint32_t* __restrict__ knn,
                                                    const float* __restrict__ X, uint32_t ld, uint32_t M_,
                                                    uint32_t prev_rows, uint32_t prev_cap,
                                                    const uint32_t* __restrict__ prev_cnt,
                                                    const uint32_t* __restrict__ prev_ids,
                                                    const float* __restrict__ prev_d, uint32_t cap,
                                                    uint32_t* __restrict__ out_cnt, uint32_t* __restrict__ out_ids,
                                                    float* __restrict__ out_d, unsigned long long* __restrict__ dcount) {
    extern __shared__ float D[];
    __shared__ uint32_t rid[DPG_MAXCAP + DPG_MAXK];  // row ids of D's rows: prev entries, then candidates
    __shared__ float cdst[DPG_MAXK];
    __shared__ uint32_t tid_[DPG_MAXCAP], tr[DPG_MAXCAP];
    __shared__ float td[DPG_MAXCAP];
    const uint32_t lane = threadIdx.x;
    unsigned long long my_d = 0;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint32_t t = (prev_cnt && i < prev_rows) ? prev_cnt[i] : 0u;
        const float* xi = X + (uint64_t)i * ld;
        for (uint32_t l = lane; l < t; l += NNM_WAVE) {
            rid[l] = prev_ids[(uint64_t)i * prev_cap + l];
            tid_[l] = rid[l];
            td[l] = prev_d[(uint64_t)i * prev_cap + l];
            tr[l] = l;
        }
        for (uint32_t j = lane; j < K; j += NNM_WAVE) {
            const uint32_t nb = knn[(uint64_t)i * K + j];
            rid[t + j] = nb;
            cdst[j] = nnm_l2_rows(xi, X + (uint64_t)nb * ld, ld);
            ++my_d;
        }
        lds_sync();
        // D[r][j]: prev entry r (< t) or candidate r - t (< j) against candidate j
        const uint32_t npairs = t * K + K * (K - 1) / 2;
        for (uint32_t p = lane; p < npairs; p += NNM_WAVE) {
            uint32_t r, j;
            if (p < t * K) {
                r = p / K;
                j = p % K;
            } else {
                // q-th pair (a, b), a < b, ordered by b then a
                const uint32_t q = p - t * K;
                uint32_t b = 1;
                while ((b + 1) * b / 2 <= q) ++b;
                const uint32_t a = q - b * (b - 1) / 2;
                r = t + a;
                j = b;
            }
            D[r * K + j] = nnm_l2_rows(X + (uint64_t)rid[r] * ld, X + (uint64_t)rid[t + j] * ld, ld);
            ++my_d;
        }
        lds_sync();
        if (lane == 0) {
            uint32_t ts = t;
            for (uint32_t j = 0; j < K; ++j) {
                if (ts > M_) break;
                const float dst = cdst[j];
                bool good = true;
                for (uint32_t l = 0; l < ts; ++l)
                    if (D[tr[l] * K + j] < dst) good = false;
                if (good) insert2topk(tid_, td, tr, ts, rid[t + j], dst, t + j, M_);
            }
            out_cnt[i] = ts;
            for (uint32_t l = 0; l < ts; ++l) {
                out_ids[(uint64_t)i * cap + l] = tid_[l];
                out_d[(uint64_t)i * cap + l] = td[l];
            }
        }
        lds_sync();
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) my_d += __shfl_xor(my_d, o, NNM_WAVE);
    if (lane == 0) atomicAdd(dcount, my_d);
}

__global__ void k_dpg_rev_count(uint32_t n, uint32_t cap, const uint32_t* __restrict__ cnt,
                                const uint32_t* __restrict__ ids, uint32_t* __restrict__ rcnt) {
    const uint64_t tot = (uint64_t)n * cap;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(e / cap), l = (uint32_t)(e % cap);
        if (l < cnt[i]) atomicAdd(&rcnt[ids[e]], 1u);
    }
}

__global__ void k_dpg_rev_fill(uint32_t n, uint32_t cap, const uint32_t* __restrict__ cnt,
                               const uint32_t* __restrict__ ids, const float* __restrict__ ds,
                               const uint32_t* __restrict__ roff, uint32_t* __restrict__ rcur,
                               uint64_t* __restrict__ rkey) {
    const uint64_t tot = (uint64_t)n * cap;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(e / cap), l = (uint32_t)(e % cap);
        if (l < cnt[i]) {
            const uint32_t u = ids[e];
            const uint32_t pos = atomicAdd(&rcur[u], 1u);
            rkey[roff[u] + pos] = ((uint64_t)i << 32) | __float_as_uint(ds[e]);
        }
    }
}

// One wave per target: apply its reverse offers in increasing source order.
__global__ __launch_bounds__(64) void k_dpg_reverse(uint32_t n, uint32_t cap, uint32_t maxM, const float* __restrict__ X,
                                                    uint32_t ld, const uint32_t* __restrict__ p1_cnt,
                                                    const uint32_t* __restrict__ p1_ids,
                                                    const float* __restrict__ p1_d, const uint32_t* __restrict__ roff,
                                                    const uint64_t* __restrict__ rkey, uint32_t* __restrict__ out_cnt,
                                                    uint32_t* __restrict__ out_ids, float* __restrict__ out_d,
                                                    unsigned long long* __restrict__ dcount) {
    __shared__ uint32_t tid_[DPG_MAXCAP];
    __shared__ float td[DPG_MAXCAP];
    __shared__ uint32_t tsz;
    const uint32_t lane = threadIdx.x;
    unsigned long long my_d = 0;
    for (uint32_t u = blockIdx.x; u < n; u += gridDim.x) {
        uint32_t ts = p1_cnt[u];
        for (uint32_t l = lane; l < ts; l += NNM_WAVE) {
            tid_[l] = p1_ids[(uint64_t)u * cap + l];
            td[l] = p1_d[(uint64_t)u * cap + l];
        }
        if (lane == 0) tsz = ts;
        lds_sync();
        const uint32_t b = roff[u], c = roff[u + 1] - b;
        uint64_t last = 0;
        bool first = true;
        for (uint32_t it = 0; it < c; ++it) {
            // next offer: the smallest source above the last one
            uint64_t best = ~0ull;
            for (uint32_t j = lane; j < c; j += NNM_WAVE) {
                const uint64_t k = rkey[b + j];
                if ((first || k > last) && k < best) best = k;
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const uint64_t v = __shfl_xor(best, o, NNM_WAVE);
                best = v < best ? v : best;
            }
            first = false;
            last = best;
            const uint32_t s = (uint32_t)(best >> 32);
            const float dst = __uint_as_float((uint32_t)best);
            ts = tsz;
            bool ok = true;
            if (ts >= maxM) {
                bool closer = false;
                for (uint32_t l = lane; l < ts; l += NNM_WAVE) {
                    const float dd = nnm_l2_rows(X + (uint64_t)tid_[l] * ld, X + (uint64_t)s * ld, ld);
                    ++my_d;
                    closer = closer || dd < dst;
                }
                ok = nnm_ballot(closer) == 0;
            }
            if (ok && lane == 0) {
                insert2topk(tid_, td, nullptr, ts, s, dst, 0, maxM);
                tsz = ts;
            }
            lds_sync();
        }
        ts = tsz;
        if (lane == 0) out_cnt[u] = ts;
        for (uint32_t l = lane; l < ts; l += NNM_WAVE) {
            out_ids[(uint64_t)u * cap + l] = tid_[l];
            out_d[(uint64_t)u * cap + l] = td[l];
        }
        lds_sync();
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) my_d += __shfl_xor(my_d, o, NNM_WAVE);
    if (lane == 0) atomicAdd(dcount, my_d);
}

__global__ void k_dpg_compact(uint32_t n, uint32_t cap, const uint32_t* __restrict__ cnt,
                              const uint32_t* __restrict__ ids, const uint32_t* __restrict__ offs,
                              uint32_t* __restrict__ adj) {
    const uint64_t tot = (uint64_t)n * cap;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t i = (uint32_t)(e / cap), l = (uint32_t)(e % cap);
        if (l < cnt[i]) adj[offs[i] + l] = ids[e];
    }
}

nnm_dpg_state* dpg_of(nnm_ctx* c) { return static_cast<nnm_dpg_state*>(c->dpg); }

void free_state(nnm_ctx* c, nnm_dpg_state* S) {
    if (!S) return;
    nnm_free(c, S->cnt);
    nnm_free(c, S->ids);
    nnm_free(c, S->d);
    delete S;
}

uint32_t grid_1d(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 65536)); }

}  // namespace

void nnm_dpg_free(nnm_ctx* c) {
    free_state(c, dpg_of(c));
    c->dpg = nullptr;
}

extern "C" int nnm_dpg_build(nnm_ctx* c, uint32_t M, uint32_t maxM, int carry, nnm_stats* st) {
    if (!c) return NNM_ERR_INVALID;
    if (st) std::memset(st, 0, sizeof(*st));
    if (!c->X || !c->g_ids) return nnm_fail(c, NNM_ERR_STATE, "nnm_dpg_build needs data and a k-NN graph");
    NNM_TRY(nnm_require_l2(c, "nnm_dpg_build"));
    const uint32_t K = c->K, n = (uint32_t)c->n_act;
    if (K == 0 || K > DPG_MAXK) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "DPG supports K <= 64");
    if (M == 0 || maxM == 0 || M > DPG_MAXCAP || maxM > DPG_MAXCAP)
        return nnm_fail(c, NNM_ERR_INVALID, "M and maxM must be in [1, 64]");
    nnm_dpg_state* P = carry ? dpg_of(c) : nullptr;
    if (P && P->rows > n) return nnm_fail(c, NNM_ERR_INVALID, "carry: the previous layer has more rows");
    const uint32_t cap = std::max(std::max(M, maxM), P ? P->cap : 0u);
    if (cap > DPG_MAXCAP) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "DPG lists longer than 64");
    NNM_HIP(c, hipSetDevice(c->device));
    NNM_HIP(c, hipEventRecord(c->ev0, c->stream));
    nnm_dpg_state* S = new nnm_dpg_state();
    S->rows = n;
    S->cap = cap;
    uint32_t *p1c = nullptr, *p1i = nullptr, *rcnt = nullptr, *roff = nullptr, *rcur = nullptr, *tmp = nullptr;
    float* p1d = nullptr;
    uint64_t* rkey = nullptr;
    unsigned long long* dc = nullptr;
    int rc = NNM_OK;
    auto cleanup = [&]() {
        nnm_free(c, p1c); nnm_free(c, p1i); nnm_free(c, p1d); nnm_free(c, rcnt); nnm_free(c, roff);
        nnm_free(c, rcur); nnm_free(c, tmp); nnm_free(c, rkey); nnm_free(c, dc);
    };
    const uint64_t ne = (uint64_t)n * cap;
    if ((rc = nnm_alloc(c, (void**)&S->cnt, (uint64_t)n * 4)) || (rc = nnm_alloc(c, (void**)&S->ids, ne * 4)) ||
        (rc = nnm_alloc(c, (void**)&S->d, ne * 4)) || (rc = nnm_alloc(c, (void**)&p1c, (uint64_t)n * 4)) ||
        (rc = nnm_alloc(c, (void**)&p1i, ne * 4)) || (rc = nnm_alloc(c, (void**)&p1d, ne * 4)) ||
        (rc = nnm_alloc(c, (void**)&rcnt, (uint64_t)n * 4)) || (rc = nnm_alloc(c, (void**)&roff, ((uint64_t)n + 1) * 4)) ||
        (rc = nnm_alloc(c, (void**)&rcur, (uint64_t)n * 4)) ||
        (rc = nnm_alloc(c, (void**)&tmp, ((uint64_t)n + 2047) / 2048 * 4 + 8)) ||
        (rc = nnm_alloc(c, (void**)&rkey, std::max<uint64_t>(ne, 1) * 8)) || (rc = nnm_alloc(c, (void**)&dc, 8))) {
        cleanup();
        free_state(c, S);
        return rc;
    }
    auto body = [&]() -> int {
        NNM_TRY(nnm_fill_u64(c, reinterpret_cast<uint64_t*>(dc), 1, 0));
        NNM_TRY(nnm_fill_u32(c, rcnt, n, 0));
        NNM_TRY(nnm_fill_u32(c, rcur, n, 0));
        const uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)c->num_cus * 16, n);
        const size_t lds = (size_t)(cap + K) * K * sizeof(float);
        hipLaunchKernelGGL(k_dpg_occlude, dim3(grid), dim3(64), lds, c->stream, n, K, c->g_ids, c->X, c->ld, M,
                           P ? P->rows : 0u, P ? P->cap : 0u, P ? P->cnt : nullptr, P ? P->ids : nullptr,
                           P ? P->d : nullptr, cap, p1c, p1i, p1d, dc);
        NNM_LAUNCH_CHECK(c);
        hipLaunchKernelGGL(k_dpg_rev_count, dim3(grid_1d(ne)), dim3(256), 0, c->stream, n, cap, p1c, p1i, rcnt);
        NNM_LAUNCH_CHECK(c);
        NNM_TRY(nnm_scan_u32(c, rcnt, n, roff, tmp));
        hipLaunchKernelGGL(k_dpg_rev_fill, dim3(grid_1d(ne)), dim3(256), 0, c->stream, n, cap, p1c, p1i, p1d, roff,
                           rcur, rkey);
        NNM_LAUNCH_CHECK(c);
        hipLaunchKernelGGL(k_dpg_reverse, dim3(grid), dim3(64), 0, c->stream, n, cap, maxM, c->X, c->ld, p1c, p1i,
                           p1d, roff, rkey, S->cnt, S->ids, S->d, dc);
        NNM_LAUNCH_CHECK(c);
        NNM_HIP(c, hipEventRecord(c->ev1, c->stream));
        uint64_t h = 0;
        NNM_TRY(nnm_d2h(c, &h, dc, 8));
        if (st) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, c->ev0, c->ev1);
            st->secs = ms * 1e-3;
            st->dists = h;
        }
        return NNM_OK;
    };
    rc = body();
    cleanup();
    if (rc != NNM_OK) {
        free_state(c, S);
        return rc;
    }
    nnm_dpg_free(c);
    c->dpg = S;
    return NNM_OK;
}

extern "C" int nnm_dpg_get(nnm_ctx* c, uint32_t* offs, uint32_t* adj, float* dist) {
    if (!c) return NNM_ERR_INVALID;
    nnm_dpg_state* S = dpg_of(c);
    if (!S) return nnm_fail(c, NNM_ERR_STATE, "no DPG layer (nnm_dpg_build)");
    if (!offs) return nnm_fail(c, NNM_ERR_INVALID, "nnm_dpg_get: offs is required");
    std::vector<uint32_t> cnt(S->rows);
    NNM_TRY(nnm_d2h(c, cnt.data(), S->cnt, (uint64_t)S->rows * 4));
    offs[0] = 0;
    for (uint32_t i = 0; i < S->rows; ++i) offs[i + 1] = offs[i] + cnt[i];
    if (adj || dist) {
        const uint64_t ne = (uint64_t)S->rows * S->cap;
        std::vector<uint32_t> ids(adj ? ne : 0);
        std::vector<float> ds(dist ? ne : 0);
        if (adj) NNM_TRY(nnm_d2h(c, ids.data(), S->ids, ne * 4));
        if (dist) NNM_TRY(nnm_d2h(c, ds.data(), S->d, ne * 4));
        for (uint32_t i = 0; i < S->rows; ++i)
            for (uint32_t l = 0; l < cnt[i]; ++l) {
                if (adj) adj[offs[i] + l] = ids[(uint64_t)i * S->cap + l];
                if (dist) dist[offs[i] + l] = ds[(uint64_t)i * S->cap + l];
            }
    }
    return NNM_OK;
}

// The DPG layer (or, without one, the k-NN lists) as the single search layer.
extern "C" int nnm_search_use_graph(nnm_ctx* c, int source) {
    if (!c) return NNM_ERR_INVALID;
    uint32_t rows, cap;
    const uint32_t *cnt = nullptr, *ids;
    if (source == NNM_SEARCH_SRC_DPG) {
        nnm_dpg_state* S = dpg_of(c);
        if (!S) return nnm_fail(c, NNM_ERR_STATE, "no DPG layer (nnm_dpg_build)");
        rows = S->rows;
        cap = S->cap;
        cnt = S->cnt;
        ids = S->ids;
    } else if (source == NNM_SEARCH_SRC_KNN) {
        if (!c->g_ids) return nnm_fail(c, NNM_ERR_STATE, "no k-NN graph");
        rows = (uint32_t)c->n_act;
        cap = c->K;
        ids = c->g_ids;
    } else {
        return nnm_fail(c, NNM_ERR_INVALID, "unknown search graph source");
    }
    uint32_t *ccnt = nullptr, *offs = nullptr, *adj = nullptr, *tmp = nullptr;
    const uint64_t ne = (uint64_t)rows * cap;
    int rc = NNM_OK;
    if ((rc = nnm_alloc(c, (void**)&offs, ((uint64_t)rows + 1) * 4)) ||
        (rc = nnm_alloc(c, (void**)&adj, std::max<uint64_t>(ne, 1) * 4)) ||
        (rc = nnm_alloc(c, (void**)&tmp, ((uint64_t)rows + 2047) / 2048 * 4 + 8)) ||
        (!cnt && (rc = nnm_alloc(c, (void**)&ccnt, (uint64_t)rows * 4)))) {
        nnm_free(c, offs); nnm_free(c, adj); nnm_free(c, tmp); nnm_free(c, ccnt);
        return rc;
    }
    auto body = [&]() -> int {
        if (!cnt) {
            NNM_TRY(nnm_fill_u32(c, ccnt, rows, cap));
            cnt = ccnt;
        }
        NNM_TRY(nnm_scan_u32(c, cnt, rows, offs, tmp));
        hipLaunchKernelGGL(k_dpg_compact, dim3(grid_1d(ne)), dim3(256), 0, c->stream, rows, cap, cnt, ids, offs, adj);
        NNM_LAUNCH_CHECK(c);
        NNM_HIP(c, hipStreamSynchronize(c->stream));
        return NNM_OK;
    };
    rc = body();
    nnm_free(c, tmp);
    nnm_free(c, ccnt);
    if (rc != NNM_OK) {
        nnm_free(c, offs);
        nnm_free(c, adj);
        return rc;
    }
    return nnm_search_adopt_layer(c, rows, offs, adj);
}

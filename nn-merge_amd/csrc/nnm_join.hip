// Local join: NNList::join (src/pvetex.h:111-125) with the pair filters of
// Smerge::localJoin (src/smerge.cpp:212-235), Jmerge::localJoin
// (src/jmerge.cpp:178-203), LyNNDescent::localJoin (src/lynndescent.cpp:313-342)
// and KGraph::join (src/kgraph.cpp:191-210).
//
// Work unit: one vertex v, one wavefront.  Its candidate list is
//   list = new(v) ++ old(v)        (new first, as NNList::join walks them)
// and the pairs are (list[i], list[j]) with i < |new| and j > i: every
// new-new pair once plus every new-old pair.  The list is cut into windows of
// 64; a pair of windows (A = rows i, B = columns j) is staged into LDS 32
// dimensions (128 bytes of a row) at a time and every lane accumulates 2x4
// register micro-tiles of squared distances over the dimensions in order
// 0..d-1 with fused multiply-add, the same order as nnm_l2_rows, so a pair
// gives the same bits here as anywhere else in the library.
//
// A pair that passes the rule is counted as one distance (nCmps) and offered
// to both ends.  An offer (target t, id u, dist) whose distance is below t's
// k-th distance at the start of the iteration goes into slot
// hash(u) mod C of t by a 64-bit atomicMin of (dist bits << 32 | id): the
// slot keeps the best offer that hashes there, whatever the order of the
// atomics, so the result does not depend on scheduling.  k_update then merges
// each vertex's slots into its list (per-vertex ownership, no list locks).
#include <hip/hip_runtime.h>

#include "nnm_internal.hpp"

#define JW 64     // window length (list entries)
#define KC 32     // dimensions staged per pass (128 bytes of an FP32 row)
#define PITCH 36  // LDS floats per staged row (KC + 4 keeps 16-byte alignment)
#define TM 2      // micro-tile rows
#define TN 4      // micro-tile columns
#define MAXT ((JW / TM) * (JW / TN))  // 512 micro-tiles per window pair
#define MAXS (MAXT / NNM_WAVE)        // 8 micro-tiles per lane

struct JoinArgs {
    const float* X;
    uint32_t ld;
    uint64_t n;
    uint32_t K;
    uint32_t C;
    uint32_t rule;
    uint32_t salt;
    const uint32_t* l_new;
    uint32_t cap_new;
    const uint32_t* l_old;
    uint32_t cap_old;
    const uint32_t* n_new;
    const uint32_t* n_old;
    const uint32_t* label;
    const float* g_dist;
    uint64_t* slots;
    uint64_t* counters;
};

__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline void offer(uint64_t* __restrict__ slots, uint32_t C, uint32_t salt, uint32_t target,
                             uint32_t id, float dd) {
    const uint32_t s = nnm_hash32(id, salt) & (C - 1);
    atomicMin((unsigned long long*)&slots[(uint64_t)target * C + s], (unsigned long long)nnm_key(dd, id));
}

__global__ __launch_bounds__(64) void k_join_fp32(JoinArgs a) {
    __shared__ __align__(16) float rowsA[JW * PITCH];
    __shared__ __align__(16) float rowsB[JW * PITCH];
    __shared__ uint32_t metaA[3][JW];  // id, label, k-th distance bits
    __shared__ uint32_t metaB[3][JW];
    __shared__ uint16_t tiles[MAXT];

    const uint32_t lane = threadIdx.x;
    uint64_t my_dists = 0, my_entries = 0;

    for (;;) {
        uint32_t v = 0;
        if (lane == 0) v = (uint32_t)atomicAdd((unsigned long long*)&a.counters[CNT_WORK], 1ull);
        v = __shfl(v, 0, NNM_WAVE);
        if (v >= a.n) break;
        const uint32_t nn = a.n_new[v], no = a.n_old[v];
        const uint32_t m = nn + no;
        if (lane == 0) my_entries += m;
        if (nn == 0) continue;
        const uint32_t* Ln = a.l_new + (uint64_t)v * a.cap_new;
        const uint32_t* Lo = a.l_old + (uint64_t)v * a.cap_old;

        for (uint32_t a0 = 0; a0 < nn; a0 += JW) {
            const uint32_t na = min(JW, nn - a0);
            for (uint32_t b0 = a0; b0 < m; b0 += JW) {
                const uint32_t nb = min(JW, m - b0);
                const bool diag = (b0 == a0);
                wave_lds_sync();  // earlier readers of meta/tiles are done
                if (lane < nb) {
                    const uint32_t idx = b0 + lane;
                    const uint32_t id = idx < nn ? Ln[idx] : Lo[idx - nn];
                    metaB[0][lane] = id;
                    metaB[1][lane] = a.label[id];
                    metaB[2][lane] = __float_as_uint(a.g_dist[(uint64_t)id * a.K + a.K - 1]);
                }
                if (!diag && lane < na) {
                    const uint32_t id = Ln[a0 + lane];
                    metaA[0][lane] = id;
                    metaA[1][lane] = a.label[id];
                    metaA[2][lane] = __float_as_uint(a.g_dist[(uint64_t)id * a.K + a.K - 1]);
                }
                const float* rA = diag ? rowsB : rowsA;
                uint32_t(*mA)[JW] = diag ? metaB : metaA;

                // compact list of micro-tiles holding at least one pair
                const uint32_t RBn = (na + TM - 1) / TM, CBn = (nb + TN - 1) / TN;
                const uint32_t T = RBn * CBn;
                uint32_t ntile = 0;
                for (uint32_t base = 0; base < T; base += NNM_WAVE) {
                    const uint32_t t = base + lane;
                    bool use = false;
                    if (t < T) {
                        const uint32_t rb = t / CBn, cb = t % CBn;
                        const uint32_t jmax = min(cb * TN + TN - 1, nb - 1);
                        use = !diag || jmax > rb * TM;
                    }
                    const uint64_t msk = __ballot(use);
                    if (use) tiles[ntile + nnm_mask_rank(msk)] = (uint16_t)t;
                    ntile += (uint32_t)__popcll(msk);
                }
                const uint32_t my_n = ntile > lane ? (ntile - lane + NNM_WAVE - 1) / NNM_WAVE : 0;

                float acc[MAXS][TM * TN];
#pragma unroll
                for (int s = 0; s < MAXS; ++s)
#pragma unroll
                    for (int q = 0; q < TM * TN; ++q) acc[s][q] = 0.f;

                for (uint32_t kc = 0; kc < a.ld; kc += KC) {
                    wave_lds_sync();
                    {
                        const uint32_t q = lane & 7, r8 = lane >> 3;
                        for (uint32_t r0 = 0; r0 < nb; r0 += 8) {
                            const uint32_t r = r0 + r8;
                            if (r < nb) {
                                const uint32_t id = metaB[0][r];
                                const float4 val =
                                    *reinterpret_cast<const float4*>(a.X + (uint64_t)id * a.ld + kc + 4 * q);
                                *reinterpret_cast<float4*>(&rowsB[r * PITCH + 4 * q]) = val;
                            }
                        }
                        if (!diag) {
                            for (uint32_t r0 = 0; r0 < na; r0 += 8) {
                                const uint32_t r = r0 + r8;
                                if (r < na) {
                                    const uint32_t id = metaA[0][r];
                                    const float4 val =
                                        *reinterpret_cast<const float4*>(a.X + (uint64_t)id * a.ld + kc + 4 * q);
                                    *reinterpret_cast<float4*>(&rowsA[r * PITCH + 4 * q]) = val;
                                }
                            }
                        }
                    }
                    wave_lds_sync();
#pragma unroll
                    for (int s = 0; s < MAXS; ++s) {
                        if ((uint32_t)s < my_n) {
                            const uint32_t t = tiles[lane + NNM_WAVE * s];
                            const uint32_t rb = t / CBn, cb = t % CBn;
                            const float* pa = rA + rb * TM * PITCH;
                            const float* pb = rowsB + cb * TN * PITCH;
#pragma unroll
                            for (int k = 0; k < KC; k += 4) {
                                float4 x[TM], y[TN];
#pragma unroll
                                for (int i = 0; i < TM; ++i) x[i] = *reinterpret_cast<const float4*>(pa + i * PITCH + k);
#pragma unroll
                                for (int j = 0; j < TN; ++j) y[j] = *reinterpret_cast<const float4*>(pb + j * PITCH + k);
#pragma unroll
                                for (int i = 0; i < TM; ++i)
#pragma unroll
                                    for (int j = 0; j < TN; ++j) {
                                        float t0 = x[i].x - y[j].x;
                                        float r = __builtin_fmaf(t0, t0, acc[s][i * TN + j]);
                                        t0 = x[i].y - y[j].y;
                                        r = __builtin_fmaf(t0, t0, r);
                                        t0 = x[i].z - y[j].z;
                                        r = __builtin_fmaf(t0, t0, r);
                                        t0 = x[i].w - y[j].w;
                                        acc[s][i * TN + j] = __builtin_fmaf(t0, t0, r);
                                    }
                            }
                        }
                    }
                }

                // epilogue: rule, count, offers
#pragma unroll
                for (int s = 0; s < MAXS; ++s) {
                    if ((uint32_t)s < my_n) {
                        const uint32_t t = tiles[lane + NNM_WAVE * s];
                        const uint32_t rb = t / CBn, cb = t % CBn;
#pragma unroll
                        for (int i = 0; i < TM; ++i) {
                            const uint32_t ri = rb * TM + i;
                            if (ri >= na) continue;
                            const uint32_t aid = mA[0][ri], alab = mA[1][ri];
                            const float athr = __uint_as_float(mA[2][ri]);
#pragma unroll
                            for (int j = 0; j < TN; ++j) {
                                const uint32_t cj = cb * TN + j;
                                if (cj >= nb) continue;
                                if (diag && cj <= ri) continue;
                                const uint32_t bid = metaB[0][cj], blab = metaB[1][cj];
                                if (aid == bid) continue;
                                if (a.rule == NNM_RULE_MERGE && alab == blab && alab != NNM_RAW) continue;
                                const float dd = acc[s][i * TN + j];
                                ++my_dists;
                                if (dd < __uint_as_float(metaB[2][cj])) offer(a.slots, a.C, a.salt, bid, aid, dd);
                                if (dd < athr) offer(a.slots, a.C, a.salt, aid, bid, dd);
                            }
                        }
                    }
                }
            }
        }
    }
    // per-wave totals
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        my_dists += __shfl_xor(my_dists, o, NNM_WAVE);
        my_entries += __shfl_xor(my_entries, o, NNM_WAVE);
    }
    if (lane == 0) {
        atomicAdd((unsigned long long*)&a.counters[CNT_DISTS], (unsigned long long)my_dists);
        atomicAdd((unsigned long long*)&a.counters[CNT_LIST_ENTRIES], (unsigned long long)my_entries);
    }
}

int nnm_launch_join(nnm_ctx* c, const nnm_params* p, uint32_t iter, uint32_t salt) {
    (void)iter;
    JoinArgs a;
    a.X = c->X;
    a.ld = c->ld;
    a.n = c->n;
    a.K = c->K;
    a.C = c->slots_C;
    a.rule = p->rule;
    a.salt = salt;
    a.l_new = c->l_new;
    a.cap_new = c->cap_new;
    a.l_old = c->l_old;
    a.cap_old = c->cap_old;
    a.n_new = c->n_new;
    a.n_old = c->n_old;
    a.label = c->label;
    a.g_dist = c->g_dist;
    a.slots = c->slots;
    a.counters = c->counters;
    if (p->prec != NNM_PREC_FP32)
        return nnm_fail(c, NNM_ERR_UNSUPPORTED, "join precision not built yet");
    const uint32_t grid = (uint32_t)c->num_cus * 8;
    hipLaunchKernelGGL(k_join_fp32, dim3(grid), dim3(64), 0, c->stream, a);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

// Local join: NNList::join (src/pvetex.h:111-125) with the pair filters of
// Smerge::localJoin (src/smerge.cpp:212-235), Jmerge::localJoin
// (src/jmerge.cpp:178-203), LyNNDescent::localJoin (src/lynndescent.cpp:313-342)
// and KGraph::join (src/kgraph.cpp:191-210).
//
// Work unit: one vertex v, one wavefront.  Its candidate list is
//   list = new(v) ++ old(v)        (new first, as NNList::join walks them)
// and the pairs are (list[i], list[j]) with i < |new| and j > i: every
// new-new pair once plus every new-old pair.
//
// Fast path (the list fits one window of JS rows): the whole FP32 rows are
// copied into LDS by direct-to-LDS loads (several rows per instruction, the
// 16-byte chunks of row r stored at chunk index c ^ (r & 7), so the global
// read of a row stays contiguous and lanes reading different rows at one
// chunk hit different LDS banks).  The next vertex's list head is loaded
// while the current one computes.  Every lane accumulates 2x4 register
// micro-tiles of squared distances over the dimensions in order 0..d-1 with
// fused multiply-add, the order of nnm_l2_rows, so a pair gives the same bits
// here as anywhere else in the library.
//
// General path (longer lists): windows of 64 entries, staged 32 dimensions
// at a time.
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

#define JW 64     // general path: window length (list entries)
#define KC 32     // general path: dimensions staged per pass (128 bytes)
#define PITCH 36  // general path: LDS floats per staged row (KC + 4)
#define TM 2      // micro-tile rows
#define TN 4      // micro-tile columns
#define MAXT ((JW / TM) * (JW / TN))  // 512 micro-tiles per window pair
#define MAXS (MAXT / NNM_WAVE)        // 8 micro-tiles per lane
#define JOIN_CHUNK 16                 // rows taken from the work queue at a time
#define ROW_LDS_BYTES 24576           // LDS for staged rows (both paths)
#define RND 2                         // fast path: micro-tiles per lane per round

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

struct JoinLds {
    float rows[ROW_LDS_BYTES / 4];
    uint32_t metaA[3][JW];  // id, label, k-th distance bits
    uint32_t metaB[3][JW];
    uint16_t tiles[MAXT];
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

// Epilogue of one 2x4 micro-tile: pair rule, distance count, offers.
__device__ inline void tile_epilogue(const JoinArgs& a, const float (&acc)[TM * TN], uint32_t rb, uint32_t cb,
                                     uint32_t na, uint32_t nb, bool diag, const uint32_t (*mA)[JW],
                                     const uint32_t (*mB)[JW], uint64_t& my_dists, uint64_t& my_offers) {
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
            const uint32_t bid = mB[0][cj], blab = mB[1][cj];
            if (aid == bid) continue;
            if (a.rule == NNM_RULE_MERGE && alab == blab && alab != NNM_RAW) continue;
            const float dd = acc[i * TN + j];
            ++my_dists;
            if (dd < __uint_as_float(mB[2][cj])) {
                offer(a.slots, a.C, a.salt, bid, aid, dd);
                ++my_offers;
            }
            if (dd < athr) {
                offer(a.slots, a.C, a.salt, aid, bid, dd);
                ++my_offers;
            }
        }
    }
}

// Compact list of the micro-tiles of an na x nb window pair that hold at
// least one pair; returns the count.
__device__ inline uint32_t build_tiles(uint16_t* tiles, uint32_t na, uint32_t nb, bool diag, uint32_t lane) {
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
    return ntile;
}

// General path for one vertex: windows of 64, 32 dimensions per pass.
__device__ __attribute__((always_inline)) inline void join_vertex_windows(const JoinArgs& a, JoinLds& L, uint64_t vcur, uint32_t nn, uint32_t no,
                                    uint32_t lane, uint64_t& my_dists, uint64_t& my_offers) {
    float* rowsA = L.rows;
    float* rowsB = L.rows + JW * PITCH;
    const uint32_t m = nn + no;
    const uint32_t* Ln = a.l_new + vcur * a.cap_new;
    const uint32_t* Lo = a.l_old + vcur * a.cap_old;
    for (uint32_t a0 = 0; a0 < nn; a0 += JW) {
        const uint32_t na = min(JW, nn - a0);
        for (uint32_t b0 = a0; b0 < m; b0 += JW) {
            const uint32_t nb = min(JW, m - b0);
            const bool diag = (b0 == a0);
            wave_lds_sync();  // earlier readers of meta/tiles are done
            if (lane < nb) {
                const uint32_t idx = b0 + lane;
                const uint32_t id = idx < nn ? Ln[idx] : Lo[idx - nn];
                L.metaB[0][lane] = id;
                L.metaB[1][lane] = a.label[id];
                L.metaB[2][lane] = __float_as_uint(a.g_dist[(uint64_t)id * a.K + a.K - 1]);
            }
            if (!diag && lane < na) {
                const uint32_t id = Ln[a0 + lane];
                L.metaA[0][lane] = id;
                L.metaA[1][lane] = a.label[id];
                L.metaA[2][lane] = __float_as_uint(a.g_dist[(uint64_t)id * a.K + a.K - 1]);
            }
            const float* rA = diag ? rowsB : rowsA;
            const uint32_t(*mA)[JW] = diag ? L.metaB : L.metaA;
            wave_lds_sync();
            const uint32_t ntile = build_tiles(L.tiles, na, nb, diag, lane);
            const uint32_t CBn = (nb + TN - 1) / TN;
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
                            const uint32_t id = L.metaB[0][r];
                            const float4 val = *reinterpret_cast<const float4*>(a.X + (uint64_t)id * a.ld + kc + 4 * q);
                            *reinterpret_cast<float4*>(&rowsB[r * PITCH + 4 * q]) = val;
                        }
                    }
                    if (!diag) {
                        for (uint32_t r0 = 0; r0 < na; r0 += 8) {
                            const uint32_t r = r0 + r8;
                            if (r < na) {
                                const uint32_t id = L.metaA[0][r];
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
                        const uint32_t t = L.tiles[lane + NNM_WAVE * s];
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
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                if ((uint32_t)s < my_n) {
                    const uint32_t t = L.tiles[lane + NNM_WAVE * s];
                    tile_epilogue(a, acc[s], t / CBn, t % CBn, na, nb, diag, mA, L.metaB, my_dists, my_offers);
                }
            }
        }
    }
}

// Fast-path row address in LDS: row r, 16-byte chunk c of a row of CPR
// chunks, stored at chunk c ^ (r & 7).
template <int CPR>
__device__ inline const float4* lds_chunk(const float* rows, uint32_t r, uint32_t c) {
    return reinterpret_cast<const float4*>(rows) + r * CPR + (c ^ (r & 7u));
}

template <uint32_t JS>
struct FastLds {
    float rows[ROW_LDS_BYTES / 4];
    uint32_t metaB[3][JW];
    uint16_t tiles[((JS + TM - 1) / TM) * ((JS + TN - 1) / TN)];
};

template <int CPR>  // 16-byte chunks per row: ld = 4 * CPR
__global__ __launch_bounds__(64) void k_join_f32(JoinArgs a) {
    constexpr uint32_t LD = 4 * CPR;
    constexpr uint32_t JS_ROWS = ROW_LDS_BYTES / (LD * 4);
    constexpr uint32_t JS = JS_ROWS < 64 ? JS_ROWS : 64;   // fast-path window
    constexpr uint32_t RPI = NNM_WAVE / CPR;                // rows per direct-to-LDS instruction
    static_assert(CPR % 8 == 0 && RPI >= 1, "row length must be a multiple of 32 floats, at most 256");
    __shared__ __align__(16) FastLds<JS> L;

    const uint32_t lane = threadIdx.x;
    uint64_t my_dists = 0, my_entries = 0, my_offers = 0;

    // work queue: chunks of JOIN_CHUNK rows, one atomic per chunk
    uint64_t v = 0, vend = 0;
    {
        uint64_t base = 0;
        if (lane == 0) base = atomicAdd((unsigned long long*)&a.counters[CNT_WORK], (unsigned long long)JOIN_CHUNK);
        v = __shfl(base, 0, NNM_WAVE);
        vend = min(a.n, v + JOIN_CHUNK);
    }
    // prefetched list head of vertex v: counts and the first 64 ids of each list
    uint32_t p_nn = 0, p_no = 0, p_rn = 0, p_ro = 0;
    if (v < a.n) {
        p_nn = a.n_new[v];
        p_no = a.n_old[v];
        if (lane < a.cap_new) p_rn = a.l_new[v * a.cap_new + lane];
        if (lane < a.cap_old) p_ro = a.l_old[v * a.cap_old + lane];
    }
    while (v < a.n) {
        const uint64_t vcur = v;
        const uint32_t nn = p_nn, no = p_no, rn = p_rn, ro = p_ro;
        // advance and prefetch the next vertex's list head
        ++v;
        if (v >= vend) {
            uint64_t base = 0;
            if (lane == 0) base = atomicAdd((unsigned long long*)&a.counters[CNT_WORK], (unsigned long long)JOIN_CHUNK);
            v = __shfl(base, 0, NNM_WAVE);
            vend = min(a.n, v + JOIN_CHUNK);
        }
        if (v < a.n) {
            p_nn = a.n_new[v];
            p_no = a.n_old[v];
            if (lane < a.cap_new) p_rn = a.l_new[v * a.cap_new + lane];
            if (lane < a.cap_old) p_ro = a.l_old[v * a.cap_old + lane];
        }
        const uint32_t m = nn + no;
        if (m > JS) continue;  // long lists: k_join_f32_windows
        if (lane == 0) my_entries += m;
        if (nn == 0) continue;
        // window ids: lane j holds entry j of new ++ old
        const uint32_t src = lane < nn ? lane : lane - nn;
        const uint32_t from_new = (uint32_t)__shfl((int)rn, (int)(src & 63), NNM_WAVE);
        const uint32_t from_old = (uint32_t)__shfl((int)ro, (int)(src & 63), NNM_WAVE);
        const uint32_t wid = lane < nn ? from_new : from_old;
        wave_lds_sync();  // previous vertex's readers are done with LDS
        if (lane < m) {
            L.metaB[0][lane] = wid;
            L.metaB[1][lane] = a.label[wid];
            L.metaB[2][lane] = __float_as_uint(a.g_dist[(uint64_t)wid * a.K + a.K - 1]);
        }
        // rows -> LDS, RPI rows per instruction, chunk c of row r at c ^ (r & 7)
        {
            const uint32_t lrow = lane / CPR, slot = lane % CPR;
            for (uint32_t r0 = 0; r0 < m; r0 += RPI) {
                const uint32_t r = r0 + lrow;
                const uint32_t id = (uint32_t)__shfl((int)wid, (int)(r & 63), NNM_WAVE);
                if (lane < RPI * CPR && r < m) {
                    const uint32_t q = slot ^ (r & 7u);
                    const float* g = a.X + (uint64_t)id * LD + 4 * q;
                    __builtin_amdgcn_global_load_lds(
                        (const void*)g, (__attribute__((address_space(3))) void*)(L.rows + r0 * LD), 16, 0, 0);
                }
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        wave_lds_sync();
        const uint32_t ntile = build_tiles(L.tiles, nn, m, true, lane);
        wave_lds_sync();
        const uint32_t CBn = (m + TN - 1) / TN;
        for (uint32_t t0 = 0; t0 < ntile; t0 += NNM_WAVE * RND) {
            float acc[RND][TM * TN];
            uint32_t rb[RND], cb[RND];
            bool on[RND];
#pragma unroll
            for (int s = 0; s < RND; ++s) {
                const uint32_t ti = t0 + lane + NNM_WAVE * s;
                on[s] = ti < ntile;
                const uint32_t t = on[s] ? L.tiles[ti] : 0u;
                rb[s] = t / CBn;
                cb[s] = t % CBn;
#pragma unroll
                for (int q = 0; q < TM * TN; ++q) acc[s][q] = 0.f;
            }
#pragma unroll
            for (int s = 0; s < RND; ++s) {
                if (!on[s]) continue;
                const uint32_t ra = rb[s] * TM, rc = cb[s] * TN;
                for (uint32_t c = 0; c < CPR; ++c) {
                    float4 x[TM], y[TN];
#pragma unroll
                    for (int i = 0; i < TM; ++i) x[i] = *lds_chunk<CPR>(L.rows, ra + i, c);
#pragma unroll
                    for (int j = 0; j < TN; ++j) y[j] = *lds_chunk<CPR>(L.rows, rc + j, c);
#pragma unroll
                    for (int i = 0; i < TM; ++i)
#pragma unroll
                        for (int j = 0; j < TN; ++j) {
                            float t1 = x[i].x - y[j].x;
                            float r = __builtin_fmaf(t1, t1, acc[s][i * TN + j]);
                            t1 = x[i].y - y[j].y;
                            r = __builtin_fmaf(t1, t1, r);
                            t1 = x[i].z - y[j].z;
                            r = __builtin_fmaf(t1, t1, r);
                            t1 = x[i].w - y[j].w;
                            acc[s][i * TN + j] = __builtin_fmaf(t1, t1, r);
                        }
                }
            }
#pragma unroll
            for (int s = 0; s < RND; ++s)
                if (on[s]) tile_epilogue(a, acc[s], rb[s], cb[s], nn, m, true, L.metaB, L.metaB, my_dists, my_offers);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        my_dists += __shfl_xor(my_dists, o, NNM_WAVE);
        my_entries += __shfl_xor(my_entries, o, NNM_WAVE);
        my_offers += __shfl_xor(my_offers, o, NNM_WAVE);
    }
    if (lane == 0) {
        atomicAdd((unsigned long long*)&a.counters[CNT_DISTS], (unsigned long long)my_dists);
        atomicAdd((unsigned long long*)&a.counters[CNT_LIST_ENTRIES], (unsigned long long)my_entries);
        atomicAdd((unsigned long long*)&a.counters[CNT_OFFERS], (unsigned long long)my_offers);
    }
}

// Lists longer than min_m entries (every list when min_m = 0): windows of 64.
__global__ __launch_bounds__(64) void k_join_f32_windows(JoinArgs a, uint32_t min_m) {
    __shared__ __align__(16) JoinLds L;
    const uint32_t lane = threadIdx.x;
    uint64_t my_dists = 0, my_entries = 0, my_offers = 0;
    for (;;) {
        uint64_t base = 0;
        if (lane == 0)
            base = atomicAdd((unsigned long long*)&a.counters[CNT_WORK2], (unsigned long long)(JOIN_CHUNK * 16));
        base = __shfl(base, 0, NNM_WAVE);
        if (base >= a.n) break;
        const uint64_t end = min(a.n, base + JOIN_CHUNK * 16);
        for (uint64_t v = base; v < end; ++v) {
            const uint32_t nn = a.n_new[v], no = a.n_old[v];
            if (nn + no <= min_m) continue;
            if (lane == 0) my_entries += nn + no;
            if (nn) join_vertex_windows(a, L, v, nn, no, lane, my_dists, my_offers);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        my_dists += __shfl_xor(my_dists, o, NNM_WAVE);
        my_entries += __shfl_xor(my_entries, o, NNM_WAVE);
        my_offers += __shfl_xor(my_offers, o, NNM_WAVE);
    }
    if (lane == 0) {
        atomicAdd((unsigned long long*)&a.counters[CNT_DISTS], (unsigned long long)my_dists);
        atomicAdd((unsigned long long*)&a.counters[CNT_LIST_ENTRIES], (unsigned long long)my_entries);
        atomicAdd((unsigned long long*)&a.counters[CNT_OFFERS], (unsigned long long)my_offers);
    }
}

static uint32_t fast_window(uint32_t ld) {
    const uint32_t rows = ROW_LDS_BYTES / (ld * 4);
    return rows < 64 ? rows : 64;
}

template <typename Kern>
static uint32_t join_grid(nnm_ctx* c, Kern k) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, 0) != hipSuccess || per_cu < 1) per_cu = 4;
    return (uint32_t)c->num_cus * (uint32_t)per_cu;
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
    if (p->prec != NNM_PREC_FP32) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "join precision not built yet");
    // fast path for lists of at most js entries, windows kernel for the rest
    uint32_t js = 0;
#define NNM_FAST(CPR)                                                                                        \
    case 4 * CPR:                                                                                            \
        hipLaunchKernelGGL(k_join_f32<CPR>, dim3(join_grid(c, k_join_f32<CPR>)), dim3(64), 0, c->stream, a); \
        js = fast_window(4 * CPR);                                                                           \
        break;
    switch (c->ld) {
        NNM_FAST(8)
        NNM_FAST(16)
        NNM_FAST(24)
        NNM_FAST(32)
        NNM_FAST(40)
        NNM_FAST(48)
        NNM_FAST(56)
        NNM_FAST(64)
        default: break;
    }
#undef NNM_FAST
    NNM_LAUNCH_CHECK(c);
    hipLaunchKernelGGL(k_join_f32_windows, dim3(join_grid(c, k_join_f32_windows)), dim3(64), 0, c->stream, a, js);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

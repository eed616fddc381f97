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
// k-th distance at the start of the iteration is appended to the offer log
// (offer_emit: the wave's offers of one step are consecutive records of its
// log chunk).  No offer is dropped: the partition passes and owners of
// nnm_offers.hip merge every one of them into its target's list
// (UpdateKnnList, src/pvetex.h:40-85, for the batch of the iteration), so
// the result does not depend on scheduling and the lists have one writer.
//
// Work source: the rows [0, n) in order, or in a redo round the vertices a
// previous round could not finish (JoinArgs.work).
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
#define ROW_LDS_BYTES 24320           // LDS for staged rows (both paths)
#define RND 2                         // fast path: micro-tiles per lane per round

struct JoinArgs {
    const float* X;
    uint32_t ld;
    uint64_t n;       // work items: rows [0, n), or work[0, n)
    const uint32_t* work;  // redo round: the vertices to join (nullptr: every row)
    uint32_t K;
    uint32_t rule;
    const uint4* vmeta;  // per vertex {label, k-th distance x offer window, row norm, 0}
    uint32_t mpitch;     // uint4 from one vertex's record to the next (1: packed)
    uint32_t rpitch;     // bf16 joins: 16-byte chunks from one row to the next (0: packed rows)
    uint32_t js;      // short-list window rows (<= the kernel's compiled window)
    uint32_t chunk;   // rows per work-queue grab
    const uint32_t* l_new;
    uint32_t cap_new;
    const uint32_t* l_old;
    uint32_t cap_old;
    const uint32_t* n_new;
    const uint32_t* n_old;
    const uint32_t* label;
    const float* g_dist;
    const float* kth;
    OfferSink sink;
    uint64_t* counters;
};

// Vertex of work item i.
__device__ inline uint32_t work_vertex(const JoinArgs& a, uint64_t i) {
    return a.work ? a.work[i] : (uint32_t)i;
}

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

// Adds a wave's tallies (lanes' distances and offers, lane 0's list entries)
// to the counters; long = the long-list kernels, whose share is also kept apart.
__device__ inline void join_tally(const JoinArgs& a, uint64_t dists, uint64_t entries, uint64_t offers, bool lng) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        dists += __shfl_xor(dists, o, NNM_WAVE);
        entries += __shfl_xor(entries, o, NNM_WAVE);
        offers += __shfl_xor(offers, o, NNM_WAVE);
    }
    if (__lane_id() == 0) {
        atomicAdd((unsigned long long*)&a.counters[CNT_DISTS], (unsigned long long)dists);
        atomicAdd((unsigned long long*)&a.counters[CNT_LIST_ENTRIES], (unsigned long long)entries);
        atomicAdd((unsigned long long*)&a.counters[CNT_OFFERS], (unsigned long long)offers);
        if (lng) {
            atomicAdd((unsigned long long*)&a.counters[CNT_LONG_DISTS], (unsigned long long)dists);
            atomicAdd((unsigned long long*)&a.counters[CNT_LONG_ENTRIES], (unsigned long long)entries);
            atomicAdd((unsigned long long*)&a.counters[CNT_LONG_OFFERS], (unsigned long long)offers);
        }
    }
}

// Epilogue of one 2x4 micro-tile: pair rule, distance count, offers.  Under
// the cosine metric the accumulator is the dot product, finished here with
// the rows' squared norms (vmeta .z).  Called by the whole wave (the offers
// are appended by wave-wide steps); lanes with on = false take no pair.
template <uint32_t M = NNM_METRIC_L2>
__device__ inline void tile_epilogue(const JoinArgs& a, OfferWriter& w, bool on, const float (&acc)[TM * TN],
                                     uint32_t rb, uint32_t cb, uint32_t na, uint32_t nb, bool diag,
                                     const uint32_t (*mA)[JW], const uint32_t (*mB)[JW], uint64_t& my_dists,
                                     uint64_t& my_offers) {
#pragma unroll
    for (int i = 0; i < TM; ++i) {
        const uint32_t ri = rb * TM + i;
        const uint32_t aid = mA[0][ri], alab = mA[1][ri];
        const float athr = __uint_as_float(mA[2][ri]);
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const uint32_t cj = cb * TN + j;
            const uint32_t bid = mB[0][cj], blab = mB[1][cj];
            bool ok = on && ri < na && cj < nb && !(diag && cj <= ri) && aid != bid;
            ok = ok && !(a.rule == NNM_RULE_MERGE && alab == blab && alab != NNM_RAW);
            float dd = acc[i * TN + j];
            if constexpr (M == NNM_METRIC_COSINE) {
                const float sa = ok ? __uint_as_float(a.vmeta[aid].z) : 1.f;
                const float sb = ok ? __uint_as_float(a.vmeta[bid].z) : 1.f;
                dd = nnm_cos_finish(dd, sa, sb);
            }
            my_dists += ok ? 1u : 0u;
            const bool pb = ok && dd < __uint_as_float(mB[2][cj]);
            const bool pa = ok && dd < athr;
            my_offers += (pb ? 1u : 0u) + (pa ? 1u : 0u);
            offer_emit(a.sink, w, pb, bid, aid, dd);
            offer_emit(a.sink, w, pa, aid, bid, dd);
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
template <uint32_t M = NNM_METRIC_L2>
__device__ __attribute__((always_inline)) inline void join_vertex_windows(const JoinArgs& a, JoinLds& L, OfferWriter& w,
                                                                         uint64_t vcur, uint32_t nn, uint32_t no,
                                                                         uint32_t lane, uint64_t& my_dists,
                                                                         uint64_t& my_offers) {
    float* rowsA = L.rows;
    float* rowsB = L.rows + JW * PITCH;
    const uint32_t m = nn + no;
    const uint32_t* Ln = a.l_new + vcur * a.cap_new;
    const uint32_t* Lo = a.l_old + vcur * a.cap_old;
    for (uint32_t a0 = 0; a0 < nn; a0 += JW) {
        const uint32_t na = min(JW, nn - a0);
        for (uint32_t b0 = a0; b0 < m; b0 += JW) {
            if (w.full) return;  // the vertex is joined again in a redo round
            const uint32_t nb = min(JW, m - b0);
            const bool diag = (b0 == a0);
            wave_lds_sync();  // earlier readers of meta/tiles are done
            if (lane < nb) {
                const uint32_t idx = b0 + lane;
                const uint32_t id = idx < nn ? Ln[idx] : Lo[idx - nn];
                const uint4 mt = a.vmeta[id];
                L.metaB[0][lane] = id;
                L.metaB[1][lane] = mt.x;
                L.metaB[2][lane] = mt.y;
            }
            if (!diag && lane < na) {
                const uint32_t id = Ln[a0 + lane];
                const uint4 mt = a.vmeta[id];
                L.metaA[0][lane] = id;
                L.metaA[1][lane] = mt.x;
                L.metaA[2][lane] = mt.y;
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
                                    float r = metric_step<M>(acc[s][i * TN + j], x[i].x, y[j].x);
                                    r = metric_step<M>(r, x[i].y, y[j].y);
                                    r = metric_step<M>(r, x[i].z, y[j].z);
                                    acc[s][i * TN + j] = metric_step<M>(r, x[i].w, y[j].w);
                                }
                        }
                    }
                }
            }
#pragma unroll
            for (int s = 0; s < MAXS; ++s) {
                const bool on = (uint32_t)s < my_n;
                const uint32_t t = on ? L.tiles[lane + NNM_WAVE * s] : 0u;
                tile_epilogue<M>(a, w, on, acc[s], t / CBn, t % CBn, na, nb, diag, mA, L.metaB, my_dists,
                                 my_offers);
            }
        }
    }
}

// Fast path.  Consecutive short lists of the work queue are packed into one
// LDS window of at most js rows (their concatenated lists), so their
// micro-tiles fill the wave together.  Rows are copied into LDS by
// direct-to-LDS loads, one row per instruction, at a pitch of ld + 4 floats
// (row r starts 4r banks further, so lanes reading different rows at one
// offset hit different banks and every LDS read uses an immediate offset).
template <uint32_t JS>
struct FastLds {
    float rows[ROW_LDS_BYTES / 4];
    uint32_t meta[3][JW];                  // id, label, k-th distance bits per window row
    uint16_t tiles[((JS + TM - 1) / TM) * ((JS + TN - 1) / TN) + 64];
};

#define FB 4  // most lists packed into one window

// Joins the list heads of `nb_v` vertices packed in the window: window rows
// [off[b], off[b] + m[b]) hold vertex b's list, new entries first.
template <int CPR, uint32_t JS>
__device__ inline void join_packed(const JoinArgs& a, FastLds<JS>& L, OfferWriter& w, uint32_t lane, uint32_t nbv,
                                   const uint32_t (&off)[FB], const uint32_t (&nnv)[FB], const uint32_t (&mv)[FB],
                                   uint64_t& my_dists, uint64_t& my_offers) {
    constexpr uint32_t PF = 4 * CPR + 4;  // LDS pitch in floats
    // micro-tile list: (vertex, row block, column block) packed as b << 12 | rb << 6 | cb
    uint32_t ntile = 0;
    for (uint32_t b = 0; b < nbv; ++b) {
        const uint32_t RBn = (nnv[b] + TM - 1) / TM, CBn = (mv[b] + TN - 1) / TN;
        const uint32_t T = RBn * CBn;
        for (uint32_t base = 0; base < T; base += NNM_WAVE) {
            const uint32_t t = base + lane;
            bool use = false;
            uint32_t code = 0;
            if (t < T) {
                const uint32_t rb = t / CBn, cb = t % CBn;
                const uint32_t jmax = min(cb * TN + TN - 1, mv[b] - 1);
                use = jmax > rb * TM;
                code = (b << 12) | (rb << 6) | cb;
            }
            const uint64_t msk = __ballot(use);
            const uint32_t slot = ntile + nnm_mask_rank(msk);
            // packed lists hold at most ceil(js/2) * ceil(js/4) + 3 micro-tiles
            if (use && slot < sizeof(L.tiles) / sizeof(L.tiles[0])) L.tiles[slot] = (uint16_t)code;
            ntile += (uint32_t)__popcll(msk);
        }
    }
    ntile = min(ntile, (uint32_t)(sizeof(L.tiles) / sizeof(L.tiles[0])));
    wave_lds_sync();
    for (uint32_t t0 = 0; t0 < ntile; t0 += NNM_WAVE * RND) {
        if (w.full) return;  // the batch is joined again in a redo round
        float acc[RND][TM * TN];
        uint32_t code[RND];
        bool on[RND];
#pragma unroll
        for (int s = 0; s < RND; ++s) {
            const uint32_t ti = t0 + lane + NNM_WAVE * s;
            on[s] = ti < ntile;
            code[s] = on[s] ? L.tiles[ti] : 0u;
#pragma unroll
            for (int q = 0; q < TM * TN; ++q) acc[s][q] = 0.f;
        }
#pragma unroll
        for (int s = 0; s < RND; ++s) {
            if (!on[s]) continue;
            const uint32_t b = code[s] >> 12, rb = (code[s] >> 6) & 63u, cb = code[s] & 63u;
            const float* pa = L.rows + (off[b] + rb * TM) * PF;
            const float* pb = L.rows + (off[b] + cb * TN) * PF;
#pragma unroll
            for (uint32_t c = 0; c < (uint32_t)CPR; ++c) {
                float4 x[TM], y[TN];
#pragma unroll
                for (int i = 0; i < TM; ++i) x[i] = *reinterpret_cast<const float4*>(pa + i * PF + 4 * c);
#pragma unroll
                for (int j = 0; j < TN; ++j) y[j] = *reinterpret_cast<const float4*>(pb + j * PF + 4 * c);
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
        for (int s = 0; s < RND; ++s) {
            const uint32_t b = code[s] >> 12, rb = (code[s] >> 6) & 63u, cb = code[s] & 63u;
            tile_epilogue(a, w, on[s], acc[s], rb, cb, nnv[b], mv[b], true,
                          reinterpret_cast<const uint32_t(*)[JW]>(&L.meta[0][off[b]]),
                          reinterpret_cast<const uint32_t(*)[JW]>(&L.meta[0][off[b]]), my_dists, my_offers);
        }
    }
}

template <int CPR>  // 16-byte chunks per row: ld = 4 * CPR
__global__ __launch_bounds__(64) void k_join_f32(JoinArgs a) {
    constexpr uint32_t LD = 4 * CPR;
    constexpr uint32_t PF = LD + 4;
    constexpr uint32_t JS_ROWS = ROW_LDS_BYTES / (PF * 4);
    constexpr uint32_t JS = JS_ROWS < 64 ? JS_ROWS : 64;  // window rows
    static_assert(CPR % 8 == 0 && CPR <= 64, "row length must be a multiple of 32 floats, at most 256");
    __shared__ __align__(16) FastLds<JS> L;
    const uint32_t js = min(a.js, JS);

    const uint32_t lane = threadIdx.x;
    uint64_t my_dists = 0, my_entries = 0, my_offers = 0;
    OfferWriter w;

    for (;;) {
        uint64_t base = 0;
        if (lane == 0) base = atomicAdd((unsigned long long*)&a.counters[CNT_WORK], (unsigned long long)a.chunk);
        base = __shfl(base, 0, NNM_WAVE);
        if (base >= a.n) break;
        const uint32_t cnt = (uint32_t)min((uint64_t)a.chunk, a.n - base);
        // list sizes of the chunk: lane i < cnt holds work item base + i
        uint32_t vtx = 0, cn = 0, co = 0;
        if (lane < cnt) {
            vtx = work_vertex(a, base + lane);
            cn = a.n_new[vtx];
            co = a.n_old[vtx];
        }
        uint32_t i = 0;
        while (i < cnt) {
            // greedy packing of consecutive short lists
            uint32_t nbv = 0, rows = 0, be = 0;
            uint32_t off[FB], nnv[FB], mv[FB], vid[FB];
#pragma unroll
            for (int b = 0; b < FB; ++b) { off[b] = 0; nnv[b] = 0; mv[b] = 0; vid[b] = 0; }
            while (i < cnt && nbv < FB) {
                const uint32_t nn = (uint32_t)__shfl((int)cn, (int)i, NNM_WAVE);
                const uint32_t m = nn + (uint32_t)__shfl((int)co, (int)i, NNM_WAVE);
                if (m > js || nn == 0) {  // long lists: k_join_f32_windows; no new entries: no pairs
                    if (m <= js && lane == 0) my_entries += m;
                    ++i;
                    continue;
                }
                if (rows + m > js) break;
#pragma unroll
                for (int b = 0; b < FB; ++b)
                    if ((uint32_t)b == nbv) { off[b] = rows; nnv[b] = nn; mv[b] = m; vid[b] = i; }
                rows += m;
                ++nbv;
                ++i;
                be += m;
            }
            if (nbv == 0) continue;
            // vertex of packed list b in lane b (for the redo list)
            uint32_t bvtx = 0;
#pragma unroll
            for (int q = 0; q < FB; ++q) {
                const uint32_t t = (uint32_t)__shfl((int)vtx, (int)vid[q], NNM_WAVE);
                if (lane == (uint32_t)q) bvtx = t;
            }
            if (w.full) {  // the log is out of chunks: join the batch in a redo round
                offer_redo(a.sink, lane < nbv, bvtx);
                continue;
            }
            // window ids: lane j < rows holds window row j
            wave_lds_sync();  // previous batch's readers are done with LDS
            uint32_t wid = 0;
            {
                uint32_t b = 0;
#pragma unroll
                for (int q = 1; q < FB; ++q)
                    if ((uint32_t)q < nbv && lane >= off[q]) b = q;
                uint32_t ob = 0, nb_ = 0, vb = 0;
#pragma unroll
                for (int q = 0; q < FB; ++q)
                    if ((uint32_t)q == b) { ob = off[q]; nb_ = nnv[q]; vb = vid[q]; }
                const uint64_t v = (uint32_t)__shfl((int)vtx, (int)vb, NNM_WAVE);
                if (lane < rows) {
                    const uint32_t e = lane - ob;
                    wid = e < nb_ ? a.l_new[v * a.cap_new + e] : a.l_old[v * a.cap_old + (e - nb_)];
                }
            }
            // the rows' metadata is fetched with the rows (one wait for both)
            uint4 mt = make_uint4(0u, 0u, 0u, 0u);
            if (lane < rows) mt = a.vmeta[wid];
            // rows -> LDS, one row per direct-to-LDS instruction
            for (uint32_t r = 0; r < rows; ++r) {
                const uint32_t id = wave_read32(wid, r);
                if (lane < (uint32_t)CPR) {
                    const float* g = a.X + (uint64_t)id * LD + 4 * lane;
                    __builtin_amdgcn_global_load_lds(
                        (const void*)g, (__attribute__((address_space(3))) void*)(L.rows + r * PF), 16, 0, 0);
                }
            }
            if (lane < rows) {
                L.meta[0][lane] = wid;
                L.meta[1][lane] = mt.x;
                L.meta[2][lane] = mt.y;
            }
            __builtin_amdgcn_s_waitcnt(0);
            wave_lds_sync();
            uint64_t bd = 0, bo = 0;
            join_packed<CPR, JS>(a, L, w, lane, nbv, off, nnv, mv, bd, bo);
            if (w.full) {
                offer_redo(a.sink, lane < nbv, bvtx);
            } else {
                my_dists += bd;
                my_offers += bo;
                if (lane == 0) my_entries += be;
            }
        }
    }
    offer_close(a.sink, w);
    join_tally(a, my_dists, my_entries, my_offers, false);
}

// ---------------------------------------------------------------- BF16 MFMA join
//
// Chosen-precision form of the fast path: the Gram of a packed window
// (concatenated short lists) on MFMA (v_mfma_f32_16x16x16_bf16, FP32
// accumulation) and distances as n_i + n_j - 2 x_i.x_j with n the squared
// norms of the bf16 rows.  On integer rows of magnitude <= 256 with d <= 128
// every product and partial sum is an integer below 2^24, so the value equals
// the FP32 difference form exactly.  Otherwise the offered key is the lower
// bound d - eps, eps = s (2 sqrt(d) + s) + 1e-5 (n_i + n_j), s = 2^-8 (|x_i| +
// |x_j|) (bf16 rounding moves each element by at most 2^-8 of itself), and the
// update kernel recomputes every offered candidate in exact FP32.
//
// LDS: bf16 rows at a pitch of exactly one row, copied by 16-byte
// direct-to-LDS loads (several rows per instruction) into a chunk-swizzled
// layout (RowChunks).  Both join kernels read the Gram blocks in one order
// (gram_block), so a pair's value is the same whichever kernel joins it: the
// owners' de-dup of repeated offers relies on that for estimates, whose FP32
// sums depend on the order.
typedef short nnm_bf16x4 __attribute__((ext_vector_type(4)));
typedef float nnm_f32x4 __attribute__((ext_vector_type(4)));

#define BFB 8          // most lists packed into one bf16 window

#define KEY_EXACT 0    // bf16 norm form is exact (integer rows)
#define KEY_BOUND 1    // offer the lower bound d - eps; exact re-rank later
#define KEY_LEARN 2    // learned projection: estimate kscale * d_proj, margin tau
#define KEY_I8 3       // int8 rows (value - shift): exact integer Gram, exact key

#define BJS 64         // bf16 window rows (one meta entry per lane)

// One MFMA step of a 16x16 Gram block from one 8-byte unit per lane of each
// operand: four bf16 values (16x16x16 bf16 MFMA) or, for F8, eight OCP e4m3
// values (16x16x32 fp8 MFMA).  Lane (r, g) gives unit g of its block row r
// for A and of its block column r for B, so both operands take the same k
// positions and the sum over k is the dot product of the two rows.
template <bool F8>
__device__ inline nnm_f32x4 gram_step(const uint2 av, const uint2 bv, nnm_f32x4 acc) {
    if constexpr (F8) {
        const long A = (long)(((uint64_t)av.y << 32) | av.x);
        const long B = (long)(((uint64_t)bv.y << 32) | bv.x);
        return __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(A, B, acc, 0, 0, 0);
    } else {
        nnm_bf16x4 A, B;
        A[0] = (short)(av.x & 0xFFFFu); A[1] = (short)(av.x >> 16);
        A[2] = (short)(av.y & 0xFFFFu); A[3] = (short)(av.y >> 16);
        B[0] = (short)(bv.x & 0xFFFFu); B[1] = (short)(bv.x >> 16);
        B[2] = (short)(bv.y & 0xFFFFu); B[3] = (short)(bv.y >> 16);
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(A, B, acc, 0, 0, 0);
    }
}

template <int U>
struct Bf16Lds {
    uint2 rows[BJS * U];           // window rows x U units of 8 bytes
    uint4 rec[BJS];                // the rows' join records {label, k-th bits, norm bits, 0}
    uint32_t id[BJS];
};

// Window rows in LDS as 16-byte chunks (U / 2 per row, copied by 16-byte
// direct-to-LDS loads): chunk c of row r sits at chunk slot c ^ sw(r).  sw
// spreads the 16 rows of an MFMA operand over the eight 16-byte bank groups
// of a 128-byte LDS line (rows that share a line get different slots).
template <int U>
struct RowChunks {
    static constexpr uint32_t CH = U / 2;                                     // chunks per row
    static constexpr uint32_t S = U == 4 ? 2u : (U == 8 || U == 24) ? 1u : 0u;
    static constexpr uint32_t M = U == 4 ? 1u : (U == 8 || U == 24) ? 3u : 7u;
    static constexpr uint32_t RPI = NNM_WAVE / CH;                            // rows per load instruction
    __device__ static uint32_t sw(uint32_t r) { return (r >> S) & M; }
};

// Rows [0, rows) of a window into LDS (lane r of wcp: the id of row r): per
// instruction RPI rows, lane l fetching chunk (l % CH) ^ sw(r) of row r0 + l / CH.
// rp: 16-byte chunks from one row to the next in Xb.
template <int U>
__device__ inline void stage_rows(uint2* dst, const uint16_t* __restrict__ Xb, uint32_t rp, uint32_t wcp, uint32_t rows,
                                  uint32_t lane) {
    using C = RowChunks<U>;
    const uint32_t rr = lane / C::CH, c = lane % C::CH;
    const char* xb = reinterpret_cast<const char*>(Xb);
    for (uint32_t r0 = 0; r0 < rows; r0 += C::RPI) {
        const uint32_t r = r0 + rr;
        const uint32_t id = (uint32_t)__shfl((int)wcp, (int)min(r, NNM_WAVE - 1u), NNM_WAVE);
        if (rr < C::RPI && r < rows) {
            const char* g = xb + ((uint64_t)id * rp + (c ^ C::sw(r))) * 16u;
            __builtin_amdgcn_global_load_lds((const void*)g, (__attribute__((address_space(3))) void*)(dst + r0 * U),
                                             16, 0, 0);
        }
    }
}

// The 16x16 Gram block of window rows ra (lane row lr) and rc (lane column):
// lane group lg supplies chunks lg, lg + 4, ... (both halves of a chunk feed
// two MFMA steps), or for U = 4 unit lg.  Every k position is summed once;
// A and B take the same positions.  I8: int8 rows, one 16x16x64 integer
// MFMA per chunk (lane group lg's 16 values), the exact integer Gram.
template <int U, bool F8, bool I8 = false>
__device__ inline nnm_f32x4 gram_block(const uint2* rows, uint32_t ra, uint32_t rc, uint32_t lg) {
    using C = RowChunks<U>;
    nnm_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    if constexpr (I8) {
        static_assert(U >= 8, "int8 rows of at least 64 values");
        typedef int v4i __attribute__((ext_vector_type(4)));
        const uint4* r16 = reinterpret_cast<const uint4*>(rows);
        v4i s = {0, 0, 0, 0};
#pragma unroll
        for (uint32_t kc = 0; kc < (uint32_t)U / 8; ++kc) {
            const uint32_t c = kc * 4 + lg;
            const uint4 av = r16[ra * C::CH + (c ^ C::sw(ra))];
            const uint4 bv = r16[rc * C::CH + (c ^ C::sw(rc))];
            const v4i A = {(int)av.x, (int)av.y, (int)av.z, (int)av.w};
            const v4i B = {(int)bv.x, (int)bv.y, (int)bv.z, (int)bv.w};
            s = __builtin_amdgcn_mfma_i32_16x16x64_i8(A, B, s, 0, 0, 0);
        }
        acc[0] = (float)s[0];
        acc[1] = (float)s[1];
        acc[2] = (float)s[2];
        acc[3] = (float)s[3];
    } else if constexpr (U >= 8) {
        const uint4* r16 = reinterpret_cast<const uint4*>(rows);
#pragma unroll
        for (uint32_t kc = 0; kc < (uint32_t)U / 8; ++kc) {
            const uint32_t c = kc * 4 + lg;
            const uint4 av = r16[ra * C::CH + (c ^ C::sw(ra))];
            const uint4 bv = r16[rc * C::CH + (c ^ C::sw(rc))];
            acc = gram_step<F8>(make_uint2(av.x, av.y), make_uint2(bv.x, bv.y), acc);
            acc = gram_step<F8>(make_uint2(av.z, av.w), make_uint2(bv.z, bv.w), acc);
        }
    } else {
        const uint2 av = rows[ra * U + (lg ^ (C::sw(ra) << 1))];
        const uint2 bv = rows[rc * U + (lg ^ (C::sw(rc) << 1))];
        acc = gram_step<F8>(av, bv, acc);
    }
    return acc;
}

// Offer key of a pair from the squared norms and the Gram entry.
template <int MODE>
__device__ inline float offer_key(float an, float bn, float g, float kscale) {
    float dd = fmaxf((an + bn) - 2.f * g, 0.f);
    if constexpr (MODE == KEY_BOUND) {
        const float s = 0.00390625f * (sqrtf(an) + sqrtf(bn)) * 1.01f;
        const float eps = s * (2.f * sqrtf(dd) + s) + 1e-5f * (an + bn);
        dd = fmaxf(dd - eps, 0.f);
    } else if constexpr (MODE == KEY_LEARN) {
        dd *= kscale;
    }
    return dd;
}

// Rows of one 16-row block of a Gram block: lane (lr, lg) takes rows
// 4 lg + q, q < 4 (its entries G[4 lg + q][lr] of the MFMA result).
struct BlockRows {
    uint32_t id[4], lab[4];
    float thr[4], nrm[4];
    bool in[4];  // the row is one of the vertex's new entries
};

// Pairs of one 16x16 Gram block (column cj of lane (lr, lg), rows 4 lg + q):
// the pair rule, the distance count, and the offers below each end's
// threshold, appended to the wave's log chunk as one batch (one reservation
// per block, positions by ballot ranks).  diag: the block sits on the
// diagonal, where only the pairs above it (column > row) are joined.
// nd / no: wave-uniform distance and offer counts.
template <int MODE>
__device__ inline void block_offers(const JoinArgs& a, OfferWriter& w, const nnm_f32x4& acc, float kscale, float tau,
                                    const BlockRows& r, uint32_t rbase, bool colok, uint32_t cj, bool diag,
                                    uint32_t bid, uint32_t blab, float bthr, float bn, uint64_t& nd, uint64_t& no) {
    const bool merge = a.rule == NNM_RULE_MERGE;
    const float tb = tau * bthr;
    // the offer predicates are kept as wave masks (scalar registers) and
    // turned back into lane predicates by inverse ballots: no per-lane copies
    uint64_t mb[4], ma[4];
    float dd[4];
    uint32_t cnt = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        // bitwise: every term is cheap, branches would cost more
        bool ok = colok & r.in[q] & (!diag | (cj > rbase + (uint32_t)q)) & (r.id[q] != bid);
        ok = ok & !(merge & (r.lab[q] == blab) & (blab != NNM_RAW));
        dd[q] = offer_key<MODE>(r.nrm[q], bn, acc[q], kscale);
        // pairs outside the rule get an infinite key, so every ballot below
        // is of one comparison (a wave mask as it is)
        const float de = ok ? dd[q] : __builtin_inff();
        nd += (uint32_t)__popcll(__ballot(de < __builtin_inff()));
        mb[q] = __ballot(de < tb);
        ma[q] = __ballot(de < tau * r.thr[q]);
        cnt += (uint32_t)__popcll(mb[q]) + (uint32_t)__popcll(ma[q]);
    }
    if (!cnt) return;
    uint64_t first;
    if (!offer_reserve(a.sink, w, cnt, &first)) return;
    no += cnt;
    uint64_t* kp = a.sink.key + first;
    uint32_t* tp = a.sink.tgt + first;
    uint32_t off = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (__builtin_amdgcn_inverse_ballot_w64(mb[q])) {
            const uint32_t p = off + nnm_mbcnt(mb[q]);
            kp[p] = ((uint64_t)__float_as_uint(dd[q]) << 32) | r.id[q];
            tp[p] = bid;
        }
        off += (uint32_t)__popcll(mb[q]);
        if (__builtin_amdgcn_inverse_ballot_w64(ma[q])) {
            const uint32_t p = off + nnm_mbcnt(ma[q]);
            kp[p] = ((uint64_t)__float_as_uint(dd[q]) << 32) | bid;
            tp[p] = r.id[q];
        }
        off += (uint32_t)__popcll(ma[q]);
    }
}

// U: 8-byte units per row (bf16 rows of 4U elements; F8: e4m3 rows of 8U);
// MODE: the offer key (KEY_EXACT, KEY_BOUND, KEY_LEARN).
template <int U, bool F8, int MODE>
__global__ __launch_bounds__(64, 8) void k_join_bf16(JoinArgs a, const uint16_t* __restrict__ Xb, float kscale, float tau) {
    static_assert(U % 4 == 0 && U <= 32, "bf16 rows of 16..128 elements");
    __shared__ __align__(16) Bf16Lds<U> L;
    const uint32_t js = min(a.js, (uint32_t)BJS);
    const uint32_t lane = threadIdx.x;
    uint64_t my_dists = 0, my_entries = 0, my_offers = 0;  // wave-uniform
    OfferWriter w;

    for (;;) {
        uint64_t base = 0;
        if (lane == 0) base = atomicAdd((unsigned long long*)&a.counters[CNT_WORK], (unsigned long long)a.chunk);
        base = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(base >> 32)) << 32) |
               (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)base);
        if (base >= a.n) break;
        const uint32_t cnt = (uint32_t)min((uint64_t)a.chunk, a.n - base);
        uint32_t vtx = 0, cn = 0, co = 0;
        if (lane < cnt) {
            vtx = work_vertex(a, base + lane);
            cn = a.n_new[vtx];
            co = a.n_old[vtx];
        }
        uint32_t i = 0;
        while (i < cnt) {
            // pack up to BFB lists into one window: list b's start row, new
            // entries, length and work-chunk index sit in lane b; marks has
            // a bit at every list's start row
            uint32_t nbv = 0, rows = 0, be = 0;
            uint32_t l_off = 0, l_nn = 0, l_m = 0, l_vid = 0;
            uint64_t marks = 0;
            while (i < cnt && nbv < BFB) {
                const uint32_t nn = wave_read32(cn, i);
                const uint32_t m = nn + wave_read32(co, i);
                if (m > js || nn == 0) {
                    if (m <= js) my_entries += m;
                    ++i;
                    continue;
                }
                if (rows + m > js) break;
                if (lane == nbv) {
                    l_off = rows;
                    l_nn = nn;
                    l_m = m;
                    l_vid = i;
                }
                marks |= 1ull << rows;
                rows += m;
                ++nbv;
                ++i;
                be += m;
            }
            if (nbv == 0) continue;
            // vertex of packed list b in lane b (for the redo list)
            const uint32_t bvtx = (uint32_t)__shfl((int)vtx, (int)l_vid, NNM_WAVE);
            if (w.full) {  // the log is out of chunks: join the batch in a redo round
                offer_redo(a.sink, lane < nbv, bvtx);
                continue;
            }
            wave_lds_sync();
            uint32_t wid = 0;
            {
                // the list of window row `lane`: list starts at or below it, less one
                const uint32_t lb = (uint32_t)__popcll(marks & ((2ull << lane) - 1ull)) - 1u;
                const uint32_t ob = (uint32_t)__shfl((int)l_off, (int)lb, NNM_WAVE);
                const uint32_t nb_ = (uint32_t)__shfl((int)l_nn, (int)lb, NNM_WAVE);
                const uint32_t vb = (uint32_t)__shfl((int)l_vid, (int)lb, NNM_WAVE);
                const uint64_t v = (uint32_t)__shfl((int)vtx, (int)vb, NNM_WAVE);
                if (lane < rows) {
                    const uint32_t e = lane - ob;
                    wid = e < nb_ ? a.l_new[v * a.cap_new + e] : a.l_old[v * a.cap_old + (e - nb_)];
                }
            }
            // ids and join records -> LDS ahead of the rows (an LDS store after
            // a direct-to-LDS load would wait for it): the records and the
            // rows are in flight together, one wait for both
            // The row copy reads the ids from a VALU copy made in the same
            // block: a register written by a memory load and read after the
            // block would make the compiler wait for every outstanding load
            // (the records) first.
            uint32_t wcp = 0;
            if (lane < rows) {
                L.id[lane] = wid;
                __builtin_amdgcn_global_load_lds((const void*)(a.vmeta + (uint64_t)wid * a.mpitch),
                                                 (__attribute__((address_space(3))) void*)L.rec, 16, 0, 0);
                asm volatile("v_mov_b32 %0, %1" : "=v"(wcp) : "v"(wid));
            }
            stage_rows<U>(L.rows, Xb, a.rpitch ? a.rpitch : RowChunks<U>::CH, wcp, rows, lane);
            __builtin_amdgcn_s_waitcnt(0);
            wave_lds_sync();
            const uint32_t lr = lane & 15u, lg = lane >> 4;
            uint64_t bd = 0, bo = 0;
            for (uint32_t b = 0; b < nbv && !w.full; ++b) {
                const uint32_t ob = wave_read32(l_off, b), nb_ = wave_read32(l_nn, b), mb = wave_read32(l_m, b);
                const uint32_t RB = (nb_ + 15) / 16, CB = (mb + 15) / 16;
                for (uint32_t rb = 0; rb < RB && !w.full; ++rb) {
                    // the lane's four rows of the row block (rows past the
                    // window read its last row and are masked by in[])
                    BlockRows rw;
                    const uint32_t rbase = rb * 16 + 4 * lg;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const uint32_t ia = min(ob + rbase + q, (uint32_t)BJS - 1);
                        const uint4 ar = L.rec[ia];
                        rw.id[q] = L.id[ia];
                        rw.lab[q] = ar.x;
                        rw.thr[q] = __uint_as_float(ar.y);
                        rw.nrm[q] = __uint_as_float(ar.z);
                        rw.in[q] = rbase + q < nb_;
                    }
                    const uint32_t ra = min(ob + rb * 16 + lr, (uint32_t)BJS - 1);
                    for (uint32_t cb = rb; cb < CB; ++cb) {
                        const uint32_t rc = min(ob + cb * 16 + lr, (uint32_t)BJS - 1);
                        const nnm_f32x4 acc = gram_block<U, F8, MODE == KEY_I8>(L.rows, ra, rc, lg);
                        const uint32_t cj = cb * 16 + lr;
                        const bool colok = cj < mb;
                        const uint4 br = L.rec[rc];
                        block_offers<MODE>(a, w, acc, kscale, tau, rw, rbase, colok, cj, cb == rb, L.id[rc], br.x,
                                           __uint_as_float(br.y), __uint_as_float(br.z), bd, bo);
                        if (w.full) break;
                    }
                }
            }
            if (w.full) {
                offer_redo(a.sink, lane < nbv, bvtx);
            } else {
                my_dists += bd;
                my_offers += bo;
                my_entries += be;
            }
        }
    }
    offer_close(a.sink, w);
    join_tally(a, lane == 0 ? my_dists : 0, lane == 0 ? my_entries : 0, lane == 0 ? my_offers : 0, false);
}

// Long lists (BJS < m <= BLM) on BF16: one workgroup of 4 waves per vertex.
// The whole list is staged in LDS (one row per 4-byte direct-to-LDS
// instruction, same unit swizzle as the short-list kernel) and the 16x16
// MFMA blocks of the upper part of the (new x list) Gram are dealt to the
// four waves.
#define BLM 256

template <int U>
struct Bf16LongLds {
    uint2 rows[(BLM + 16) * U];
    uint32_t meta[4][BLM];
};

template <int U, bool F8, int MODE>
__global__ __launch_bounds__(256) void k_join_bf16_long(JoinArgs a, const uint16_t* __restrict__ Xb, float kscale,
                                                        float tau) {
    using C = RowChunks<U>;
    __shared__ __align__(16) Bf16LongLds<U> L;
    __shared__ uint32_t lq[256];    // long lists of the block's current work chunk
    __shared__ uint32_t wcnt[4];
    __shared__ uint64_t qbase;
    __shared__ uint32_t wfull;  // a wave of the block ran out of log chunks
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint64_t my_dists = 0, my_entries = 0, my_offers = 0;  // wave-uniform
    const char* xb = reinterpret_cast<const char*>(Xb);
    OfferWriter w;
    if (tid == 0) wfull = 0;
    for (;;) {
        // block-wide queue: 256 work items at a time, one per thread; the
        // long lists among them are collected in LDS and joined one by one
        if (tid == 0) qbase = atomicAdd((unsigned long long*)&a.counters[CNT_WORK2], 256ull);
        __syncthreads();
        const uint64_t b0 = qbase;
        if (b0 >= a.n) break;
        bool lng = false;
        uint32_t vi = 0;
        if (b0 + tid < a.n) {
            vi = work_vertex(a, b0 + tid);
            const uint32_t mm = a.n_new[vi] + a.n_old[vi];
            lng = mm > min(a.js, (uint32_t)BJS) && mm <= BLM;
        }
        const uint64_t bm = __ballot(lng);
        if (lane == 0) wcnt[wave] = (uint32_t)__popcll(bm);
        __syncthreads();
        uint32_t wo = 0;
        for (uint32_t q = 0; q < wave; ++q) wo += wcnt[q];
        const uint32_t nl = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
        if (lng) lq[wo + nnm_mbcnt(bm)] = vi;
        __syncthreads();
        for (uint32_t li = 0; li < nl; ++li) {
            const uint64_t v = lq[li];
            if (wfull) {  // the log is out of chunks: join the vertex in a redo round
                if (wave == 0) offer_redo(a.sink, lane == 0, (uint32_t)v);
                continue;
            }
            const uint32_t nn = a.n_new[v], no = a.n_old[v], m = nn + no;
            for (uint32_t j = tid; j < m; j += blockDim.x) {
                const uint32_t id = j < nn ? a.l_new[v * a.cap_new + j] : a.l_old[v * a.cap_old + (j - nn)];
                const uint4 mt = a.vmeta[(uint64_t)id * a.mpitch];
                L.meta[0][j] = id;
                L.meta[1][j] = mt.x;
                L.meta[2][j] = mt.y;
                L.meta[3][j] = mt.z;
            }
            __syncthreads();
            // rows in the short-list kernel's chunk layout: the waves take
            // turns by load instruction (RPI rows each)
            {
                const uint32_t rr = lane / C::CH, c = lane % C::CH;
                const uint32_t rp = a.rpitch ? a.rpitch : C::CH;
                for (uint32_t r0 = wave * C::RPI; r0 < m; r0 += 4 * C::RPI) {
                    const uint32_t r = r0 + rr;
                    if (rr < C::RPI && r < m) {
                        const char* g = xb + ((uint64_t)L.meta[0][r] * rp + (c ^ C::sw(r))) * 16u;
                        __builtin_amdgcn_global_load_lds((const void*)g,
                                                         (__attribute__((address_space(3))) void*)(L.rows + r0 * U), 16,
                                                         0, 0);
                    }
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            const uint32_t lr = lane & 15u, lg = lane >> 4;
            const uint32_t RB = (nn + 15) / 16, CB = (m + 15) / 16;
            uint32_t blk = 0;
            uint64_t bd = 0, bo = 0;
            for (uint32_t rb = 0; rb < RB; ++rb) {
                BlockRows rw;
                const uint32_t rbase = rb * 16 + 4 * lg;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t ri = rbase + q;  // < BLM: nn <= m <= BLM rows, RB * 16 <= BLM
                    rw.id[q] = L.meta[0][ri];
                    rw.lab[q] = L.meta[1][ri];
                    rw.thr[q] = __uint_as_float(L.meta[2][ri]);
                    rw.nrm[q] = __uint_as_float(L.meta[3][ri]);
                    rw.in[q] = ri < nn;
                }
                const uint32_t ra = rb * 16 + lr;
                for (uint32_t cb = rb; cb < CB; ++cb, ++blk) {
                    if ((blk & 3u) != wave || w.full) continue;
                    const uint32_t rc = cb * 16 + lr;
                    // same chunk order as the short-list kernel: a pair's
                    // value does not depend on the kernel that joins it
                    const nnm_f32x4 acc = gram_block<U, F8, MODE == KEY_I8>(L.rows, ra, rc, lg);
                    const uint32_t cj = cb * 16 + lr;
                    const bool colok = cj < m;
                    const uint32_t cm = colok ? cj : 0u;
                    block_offers<MODE>(a, w, acc, kscale, tau, rw, rbase, colok, cj, cb == rb, L.meta[0][cm], L.meta[1][cm],
                                       __uint_as_float(L.meta[2][cm]), __uint_as_float(L.meta[3][cm]), bd, bo);
                }
            }
            if (w.full && lane == 0) wfull = 1;
            __syncthreads();  // LDS reused by the next vertex; every wave's offers issued
            if (wfull) {
                if (wave == 0) offer_redo(a.sink, lane == 0, (uint32_t)v);
            } else {
                my_dists += bd;
                my_offers += bo;
                if (wave == 0) my_entries += m;
            }
        }
        __syncthreads();  // lq, wcnt and qbase are rewritten for the next chunk
    }
    offer_close(a.sink, w);
    join_tally(a, lane == 0 ? my_dists : 0, lane == 0 ? my_entries : 0, lane == 0 ? my_offers : 0, true);
}

// Lists longer than min_m entries (every list when min_m = 0): windows of 64.
template <uint32_t M>
__global__ __launch_bounds__(64) void k_join_f32_windows(JoinArgs a, uint32_t min_m) {
    __shared__ __align__(16) JoinLds L;
    const uint32_t lane = threadIdx.x;
    uint64_t my_dists = 0, my_entries = 0, my_offers = 0;
    OfferWriter w;
    for (;;) {
        uint64_t base = 0;
        if (lane == 0)
            base = atomicAdd((unsigned long long*)&a.counters[CNT_WORK3], (unsigned long long)(JOIN_CHUNK * 16));
        base = __shfl(base, 0, NNM_WAVE);
        if (base >= a.n) break;
        const uint64_t end = min(a.n, base + JOIN_CHUNK * 16);
        // the chunk's list sizes 64 at a time, one per lane; the lists this
        // kernel takes are then joined in order by the whole wave
        for (uint64_t i0 = base; i0 < end; i0 += NNM_WAVE) {
            uint32_t lv = 0, lnn = 0, lno = 0;
            if (i0 + lane < end) {
                lv = work_vertex(a, i0 + lane);
                lnn = a.n_new[lv];
                lno = a.n_old[lv];
            }
            for (uint64_t take = __ballot(lnn + lno > min_m); take; take &= take - 1) {
                const uint32_t src = (uint32_t)__builtin_ctzll(take);
                const uint32_t v = wave_read32(lv, src), nn = wave_read32(lnn, src), no = wave_read32(lno, src);
                if (!nn) {
                    if (lane == 0) my_entries += no;
                    continue;
                }
                if (w.full) {  // the log is out of chunks: join the vertex in a redo round
                    offer_redo(a.sink, lane == 0, v);
                    continue;
                }
                uint64_t bd = 0, bo = 0;
                join_vertex_windows<M>(a, L, w, v, nn, no, lane, bd, bo);
                if (w.full) {
                    offer_redo(a.sink, lane == 0, v);
                } else {
                    my_dists += bd;
                    my_offers += bo;
                    if (lane == 0) my_entries += nn + no;
                }
            }
        }
    }
    offer_close(a.sink, w);
    join_tally(a, my_dists, my_entries, my_offers, true);
}

static uint32_t fast_window(uint32_t ld) {
    const uint32_t rows = ROW_LDS_BYTES / ((ld + 4) * 4);
    return rows < 64 ? rows : 64;
}

template <typename Kern>
static uint32_t join_grid(nnm_ctx* c, Kern k) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 64, 0) != hipSuccess || per_cu < 1) per_cu = 4;
    return (uint32_t)c->num_cus * (uint32_t)per_cu;
}

template <typename Kern>
static uint32_t join_grid256(nnm_ctx* c, Kern k) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    return (uint32_t)c->num_cus * (uint32_t)per_cu;
}

uint32_t nnm_join_default_js(nnm_ctx* c, uint32_t prec) {
    return prec == NNM_PREC_FP32 ? fast_window(c->ld) : (uint32_t)BJS;
}
uint32_t nnm_join_default_chunk() { return JOIN_CHUNK; }

void nnm_join_tune_space(nnm_ctx* c, uint32_t prec, std::vector<std::pair<uint32_t, uint32_t>>* out) {
    out->clear();
    const uint32_t top = (prec == NNM_PREC_FP32) ? fast_window(c->ld) : (uint32_t)BJS;
    const uint32_t cand[3] = {top, top * 2 / 3, top / 2};
    // the learned phase keeps the full window: its first iteration (where the
    // tuning runs) has the longest new lists, and a narrower window chosen
    // there sends later iterations' lists to the long-list kernel (measured
    // at 5M: 24.2 ms of learned joins per merge at 64 rows, 26.1 ms at a
    // narrower choice)
    const uint32_t ncand = prec == NNM_PREC_LEARNED ? 1u : 3u;
    for (uint32_t i = 0; i < ncand; ++i)
        for (uint32_t ch : {4u, 16u, 64u})
            if (cand[i] >= 8) out->emplace_back(cand[i], ch);
}

uint32_t nnm_join_learned_max_list() { return BLM; }

int nnm_launch_join(nnm_ctx* c, const nnm_params* p, uint32_t js_req, uint32_t chunk, hipEvent_t mid, uint32_t round,
                    uint64_t nwork) {
    JoinArgs a;
    a.X = c->X;
    a.ld = c->ld;
    // round 0 joins every active row; a redo round the previous round's redo list
    a.n = round ? nwork : c->n_act;
    a.work = round ? c->work[(round - 1) & 1] : nullptr;
    a.K = c->K;
    a.rule = p->rule;
    a.chunk = chunk ? chunk : JOIN_CHUNK;
    a.vmeta = c->vmeta;
    a.mpitch = 1;
    a.rpitch = 0;
    a.l_new = c->l_new;
    a.cap_new = c->cap_new;
    a.l_old = c->l_old;
    a.cap_old = c->cap_old;
    a.n_new = c->n_new;
    a.n_old = c->n_old;
    a.label = c->label;
    a.g_dist = c->g_dist;
    a.kth = c->kth;
    a.sink = nnm_offer_sink(c, round);
    a.counters = c->counters;
    if (p->prec == NNM_PREC_BF16 || p->prec == NNM_PREC_LEARNED) {
        const uint16_t* rows = c->Xb;
        uint32_t rl = c->ld;
        int mode = c->data_exact_bf16 ? KEY_EXACT : KEY_BOUND;
        float kscale = 1.f, tau = 1.f;
        if (p->prec == NNM_PREC_LEARNED) {
            rows = nnm_mlp_rows(c);
            rl = nnm_mlp_dout(c);
            mode = KEY_LEARN;
            kscale = nnm_mlp_scale(c);
            // the lists hold estimates in the same metric during the learned
            // phase; pairs above learn_tau x the estimated k-th rarely enter
            tau = p->learn_tau > 0.f ? p->learn_tau : NNM_LEARN_TAU_DEFAULT;
            if (!rows) return nnm_fail(c, NNM_ERR_STATE, "learned join without a projection");
            if (p->flags & NNM_FLAG_LEARNED_FP8) {
                // FP8 prefilter: e4m3 rows of the projection (32 bytes at d_out 32)
                if (rl != 32) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "FP8 learned join built for d_out = 32");
                const uint16_t* rows8 = reinterpret_cast<const uint16_t*>(nnm_mlp_rows8(c));
                if (!rows8) return nnm_fail(c, NNM_ERR_STATE, "FP8 learned join without FP8 rows");
                a.js = js_req ? min(js_req, (uint32_t)BJS) : (uint32_t)BJS;
                const float ks8 = nnm_mlp_scale8(c);
                hipLaunchKernelGGL((k_join_bf16<4, true, KEY_LEARN>), dim3(join_grid(c, k_join_bf16<4, true, KEY_LEARN>)),
                                   dim3(64), 0, c->stream, a, rows8, ks8, tau);
                NNM_LAUNCH_CHECK(c);
                if (mid) NNM_HIP(c, hipEventRecord(mid, c->stream));
                hipLaunchKernelGGL((k_join_bf16_long<4, true, KEY_LEARN>),
                                   dim3(join_grid256(c, k_join_bf16_long<4, true, KEY_LEARN>)), dim3(256), 0, c->stream,
                                   a, rows8, ks8, tau);
                NNM_LAUNCH_CHECK(c);
                // lists longer than BLM are refused before the learned phase
                // (their FP32 join would offer exact distances against
                // estimated k-th thresholds)
                return NNM_OK;
            }
        }
        if (p->prec == NNM_PREC_LEARNED && rl == 32 && c->lrec_live) {
            // rows and join records from the 128-byte learned records: one
            // line per gathered row
            rows = reinterpret_cast<const uint16_t*>(c->lrec);
            a.vmeta = c->lrec + 4;
            a.mpitch = 8;
            a.rpitch = 8;
        }
        if (rl > 128) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "bf16 join supports d <= 128");
        if (nnm_use_i8(c, p)) {
            // int8 rows of 128 values (128 bytes: one line per gathered row;
            // made before the iterations, which may be captured into graphs)
            rows = reinterpret_cast<const uint16_t*>(c->Xi8);
            mode = KEY_I8;
        }
        a.js = js_req ? min(js_req, (uint32_t)BJS) : (uint32_t)BJS;
#define NNM_BFM(U, MODE)                                                                                            \
    hipLaunchKernelGGL((k_join_bf16<U, false, MODE>), dim3(join_grid(c, k_join_bf16<U, false, MODE>)), dim3(64), 0, \
                       c->stream, a, rows, kscale, tau);                                                           \
    NNM_LAUNCH_CHECK(c);                                                                                            \
    if (mid) NNM_HIP(c, hipEventRecord(mid, c->stream));                                                           \
    hipLaunchKernelGGL((k_join_bf16_long<U, false, MODE>), dim3(join_grid256(c, k_join_bf16_long<U, false, MODE>)),  \
                       dim3(256), 0, c->stream, a, rows, kscale, tau);
#define NNM_BF(U)                                                    \
    case 4 * U:                                                      \
        if (mode == KEY_EXACT) { NNM_BFM(U, KEY_EXACT) }             \
        else if (mode == KEY_BOUND) { NNM_BFM(U, KEY_BOUND) }        \
        else { NNM_BFM(U, KEY_LEARN) }                               \
        break;
        if (mode == KEY_I8) {
            NNM_BFM(16, KEY_I8)  // 128-byte rows: 16 units of 8 bytes
        } else {
            switch (rl) {
                NNM_BF(4)
                NNM_BF(8)
                NNM_BF(16)
                NNM_BF(24)
                NNM_BF(32)
                default: return nnm_fail(c, NNM_ERR_UNSUPPORTED, "bf16 join row length");
            }
        }
#undef NNM_BF
#undef NNM_BFM
        NNM_LAUNCH_CHECK(c);
        // lists beyond the long-list LDS window (rev_cap above 118) stay on
        // FP32 (plain BF16 joins only: a learned run with such lists is
        // refused before it starts)
        if (c->cap_new + c->cap_old > BLM && p->prec != NNM_PREC_LEARNED) {
            hipLaunchKernelGGL(k_join_f32_windows<NNM_METRIC_L2>, dim3(join_grid(c, k_join_f32_windows<NNM_METRIC_L2>)),
                               dim3(64), 0, c->stream,
                               a, (uint32_t)BLM);
            NNM_LAUNCH_CHECK(c);
        }
        return NNM_OK;
    }
    if (p->prec != NNM_PREC_FP32) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "join precision not built yet");
    if (c->metric != NNM_METRIC_L2) {
        // other metrics: the windows kernel joins every list (no packed fast path)
        a.js = 0;
        if (mid) NNM_HIP(c, hipEventRecord(mid, c->stream));
#define NNM_WIN(M)                                                                                           \
    case M:                                                                                                  \
        hipLaunchKernelGGL(k_join_f32_windows<M>, dim3(join_grid(c, k_join_f32_windows<M>)), dim3(64), 0,    \
                           c->stream, a, 0u);                                                                \
        break;
        switch (c->metric) {
            NNM_WIN(NNM_METRIC_L1)
            NNM_WIN(NNM_METRIC_COSINE)
            NNM_WIN(NNM_METRIC_CHI2)
            default: return nnm_fail(c, NNM_ERR_UNSUPPORTED, "join metric");
        }
#undef NNM_WIN
        NNM_LAUNCH_CHECK(c);
        return NNM_OK;
    }
    // fast path for lists of at most js entries, windows kernel for the rest
    uint32_t js = 0;
    a.js = js_req ? min(js_req, fast_window(c->ld)) : fast_window(c->ld);
#define NNM_FAST(CPR)                                                                                        \
    case 4 * CPR:                                                                                            \
        hipLaunchKernelGGL(k_join_f32<CPR>, dim3(join_grid(c, k_join_f32<CPR>)), dim3(64), 0, c->stream, a); \
        js = a.js;                                                                                           \
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
    if (mid) NNM_HIP(c, hipEventRecord(mid, c->stream));
    hipLaunchKernelGGL(k_join_f32_windows<NNM_METRIC_L2>, dim3(join_grid(c, k_join_f32_windows<NNM_METRIC_L2>)),
                       dim3(64), 0, c->stream, a, js);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

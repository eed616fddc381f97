// Batched graph search: LySearch::knnSearch (src/lysearch.cpp:363-437) with
// nndHNSWSrch (src/lysearch.cpp:236-322), and the seeded flat search
// nndFlatSrch (src/lysearch.cpp:439-544), for many queries in one launch.
//
// One wavefront per query, persistent waves on a query queue.  Per wave the
// LDS holds the query row, the result list (the reference's topLst: up to
// `beam` 64-bit keys dist_bits << 32 | id, sorted, with an expanded flag per
// entry; double buffered for merges), the candidate keys of one round and a
// visited hash.  The frontier (candis) is the set of unexpanded list
// entries: a candidate evicted from topLst can never be expanded in the
// reference either, because its distance exceeds lwbnd from then on.
//
// Each step expands the `width` best unexpanded entries (width = 1 is the
// reference's order): lanes take one neighbour each, skip visited ones, compute
// the distance (FP32 rows, or bf16 rows against the FP32 query), and the
// round's candidates are sorted by the wave and merged into the list by
// rank (merge path).  The visited hash is exact until half full; it is then
// cleared and refilled with the list's ids, so a list never holds an id
// twice and the search still ends (the worst list key only falls).
//
// Results: the first k list entries; in bf16 mode the first `rerank`
// entries are recomputed in FP32 and re-sorted first.
//
// Several beams in one launch: the queue holds (beam, query) items, the
// largest beam first so the long searches start early; each wave keeps its
// counters per beam and flushes them when its beam changes, and measures
// every item on the constant-rate wall clock, so the host can give each
// beam its share of the launch time.

#include "nnm_internal.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

constexpr uint32_t SR_L = 256;     // longest result list
constexpr uint32_t SR_HS = 4096;   // visited hash slots per wave
constexpr uint32_t SR_MAXLD = 256; // longest row (floats)
constexpr uint32_t SR_LAYERS = 8;
constexpr uint32_t SR_MAXW = 4;
constexpr uint32_t SR_MAXB = 32;   // beams per launch
constexpr uint64_t SR_NONE = ~0ull;

struct SearchLayer {
    const uint32_t* offs;
    const uint32_t* adj;
    uint32_t rows;
};

struct SearchArgs {
    const float* X;
    const uint16_t* Xb;
    uint32_t ld;
    uint64_t n;
    SearchLayer layer[SR_LAYERS];
    uint32_t nlayers;
    const float* Q;  // nq x ld, zero padded
    uint32_t nq;
    const uint32_t* seeds;  // flat mode: nq x beam, or null (drawn from seed)
    uint64_t seed;
    uint32_t flat;
    uint32_t k, width, rerank, entry;
    uint32_t max_steps;
    uint32_t nbeams;
    uint32_t beams[SR_MAXB];  // launch order: largest first
    uint32_t slot[SR_MAXB];   // output block (caller's order) of each launch-order beam
    uint32_t* out_ids;
    float* out_d;
    unsigned long long* counters;  // SC_N per launch-order beam; the queue in beam 0's block
};

enum { SC_DISTS = 0, SC_RERANKED = 1, SC_RESETS = 2, SC_QUEUE = 3, SC_CAPPED = 4, SC_STEPS = 5, SC_TICKS = 6,
       SC_N = 8 };

struct SearchLds {
    float q[SR_MAXLD];
    uint64_t beam[2][SR_L];
    uint8_t fl[2][SR_L];
    uint64_t cand[NNM_WAVE];
    uint32_t hash[SR_HS];
};

// LDS writes of the wave visible to its other lanes (one wave per block)
__device__ inline void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint64_t wave_min64(uint64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t u = __shfl_xor(v, o, NNM_WAVE);
        v = u < v ? u : v;
    }
    return v;
}

// Squared L2 of the LDS query against row `id`, elements in order with
// acc = fma(t, t, acc): the library's canonical FP32 order (nnm_l2_rows), so
// on rows whose values are exact in bf16 both modes give the same bits.
template <bool BF16>
__device__ inline float row_dist(const SearchArgs& a, const float* __restrict__ q, uint32_t id) {
    float acc = 0.f;
    if (BF16) {
        const uint4* r = reinterpret_cast<const uint4*>(a.Xb + (uint64_t)id * a.ld);
        for (uint32_t i = 0; i < a.ld / 8; ++i) {
            const uint4 v = r[i];
            const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int h = 0; h < 4; ++h) {
                float t = q[8 * i + 2 * h] - __uint_as_float(w[h] << 16);
                acc = __builtin_fmaf(t, t, acc);
                t = q[8 * i + 2 * h + 1] - __uint_as_float(w[h] & 0xFFFF0000u);
                acc = __builtin_fmaf(t, t, acc);
            }
        }
    } else {
        const float4* r = reinterpret_cast<const float4*>(a.X + (uint64_t)id * a.ld);
        for (uint32_t i = 0; i < a.ld / 4; ++i) {
            const float4 v = r[i];
            float t;
            t = q[4 * i] - v.x; acc = __builtin_fmaf(t, t, acc);
            t = q[4 * i + 1] - v.y; acc = __builtin_fmaf(t, t, acc);
            t = q[4 * i + 2] - v.z; acc = __builtin_fmaf(t, t, acc);
            t = q[4 * i + 3] - v.w; acc = __builtin_fmaf(t, t, acc);
        }
    }
    return acc;
}

// Insert id into the visited hash; true if it was not there.
__device__ inline bool hash_insert(uint32_t* H, uint32_t id) {
    const uint32_t key = id + 1u;
    uint32_t h = nnm_hash32(id, 0x5EA2C4u) & (SR_HS - 1);
    for (uint32_t probe = 0; probe < SR_HS; ++probe) {
        const uint32_t old = atomicCAS(&H[h], 0u, key);
        if (old == 0u) return true;
        if (old == key) return false;
        h = (h + 1) & (SR_HS - 1);
    }
    return false;  // unreachable: the table is kept at most half full
}

// Number of keys in s[0..cnt) below key (s sorted ascending).
__device__ inline uint32_t rank_in(const uint64_t* s, uint32_t cnt, uint64_t key) {
    uint32_t lo = 0, hi = cnt;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// Merge the wave's sorted candidate keys (lane order, c valid, all absent
// from the list) into list buffer cb of size bs; returns the new size (the
// list lives in buffer cb ^ 1 afterwards).
__device__ inline uint32_t merge_round(SearchLds& L, uint32_t cb, uint32_t bs, uint32_t lmax, uint64_t kk,
                                       uint32_t c, uint32_t lane) {
    L.cand[lane] = kk;
    wave_sync_lds();
    const uint32_t nb = cb ^ 1u;
    if (lane < c) {
        const uint32_t p = lane + rank_in(L.beam[cb], bs, kk);
        if (p < lmax) {
            L.beam[nb][p] = kk;
            L.fl[nb][p] = 0;
        }
    }
    for (uint32_t b = lane; b < bs; b += NNM_WAVE) {
        const uint64_t bk = L.beam[cb][b];
        const uint32_t p = b + rank_in(L.cand, c, bk);
        if (p < lmax) {
            L.beam[nb][p] = bk;
            L.fl[nb][p] = L.fl[cb][b];
        }
    }
    wave_sync_lds();
    return min(lmax, bs + c);
}

template <bool BF16>
__global__ __launch_bounds__(64) void k_search(SearchArgs a) {
    __shared__ __align__(16) SearchLds L;
    const uint32_t lane = threadIdx.x;
    unsigned long long my_d = 0, my_rr = 0, my_resets = 0, my_caps = 0, my_steps = 0, my_ticks = 0;
    const uint64_t items = (uint64_t)a.nq * a.nbeams;
    uint32_t cur_b = SR_MAXB;
    // the wave's counters into beam b's block (lane sums for the per-lane ones)
    auto flush = [&](uint32_t b) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            my_d += __shfl_xor(my_d, o, NNM_WAVE);
            my_rr += __shfl_xor(my_rr, o, NNM_WAVE);
        }
        if (lane == 0) {
            unsigned long long* cb = a.counters + (uint64_t)b * SC_N;
            atomicAdd(&cb[SC_DISTS], my_d);
            atomicAdd(&cb[SC_RERANKED], my_rr);
            atomicAdd(&cb[SC_RESETS], my_resets);
            atomicAdd(&cb[SC_CAPPED], my_caps);
            atomicAdd(&cb[SC_STEPS], my_steps);
            atomicAdd(&cb[SC_TICKS], my_ticks);
        }
        my_d = my_rr = my_resets = my_caps = my_steps = my_ticks = 0;
    };
    for (;;) {
        uint64_t item = 0;
        if (lane == 0) item = atomicAdd(&a.counters[SC_QUEUE], 1ull);
        item = __shfl(item, 0, NNM_WAVE);
        if (item >= items) break;
        const uint32_t bi = (uint32_t)(item / a.nq);
        const uint32_t qi = (uint32_t)(item - (uint64_t)bi * a.nq);
        if (bi != cur_b) {
            if (cur_b < SR_MAXB) flush(cur_b);
            cur_b = bi;
        }
        const uint32_t beam = a.beams[bi];
        const uint32_t lmax = beam;
        const uint64_t t_start = wall_clock64();
        for (uint32_t i = lane; i < a.ld; i += NNM_WAVE) L.q[i] = a.Q[(uint64_t)qi * a.ld + i];
        for (uint32_t i = lane; i < SR_HS; i += NNM_WAVE) L.hash[i] = 0u;
        wave_sync_lds();
        uint32_t cb = 0, bs = 0, nins = 0;
        if (!a.flat) {
            // greedy descent through the upper layers (src/lysearch.cpp:386-413)
            uint32_t cur = a.entry;
            float cd = row_dist<BF16>(a, L.q, cur);
            my_d += lane == 0 ? 1 : 0;
            for (uint32_t g = 0; g + 1 < a.nlayers; ++g) {
                const SearchLayer ly = a.layer[g];
                for (uint32_t moves = 0; moves < a.max_steps; ++moves) {
                    const uint64_t here = nnm_key(cd, cur);
                    uint64_t best = here;
                    const uint32_t b = ly.offs[cur], e = ly.offs[cur + 1];
                    for (uint32_t j0 = b; j0 < e; j0 += NNM_WAVE) {
                        uint64_t kk = SR_NONE;
                        if (j0 + lane < e) {
                            const uint32_t nb = ly.adj[j0 + lane];
                            kk = nnm_key(row_dist<BF16>(a, L.q, nb), nb);
                            ++my_d;
                        }
                        kk = wave_min64(kk);
                        best = kk < best ? kk : best;
                    }
                    if (best == here) break;
                    cur = nnm_key_id(best);
                    cd = nnm_key_dist(best);
                }
            }
            if (lane == 0) {
                hash_insert(L.hash, cur);
                L.beam[0][0] = nnm_key(cd, cur);
                L.fl[0][0] = 0;
            }
            wave_sync_lds();
            bs = 1;
            nins = 1;
        } else {
            // seeded start (src/lysearch.cpp:458-471): every seed enters the list
            for (uint32_t j0 = 0; j0 < beam; j0 += NNM_WAVE) {
                uint64_t kk = SR_NONE;
                bool ins = false;
                if (j0 + lane < beam) {
                    const uint32_t id = a.seeds ? a.seeds[(uint64_t)qi * beam + j0 + lane]
                                                : (uint32_t)(nnm_rand(a.seed, qi, j0 + lane) % a.n);
                    ins = hash_insert(L.hash, id);
                    if (ins) {
                        kk = nnm_key(row_dist<BF16>(a, L.q, id), id);
                        ++my_d;
                    }
                }
                nins += (uint32_t)__popcll(nnm_ballot(ins));
                kk = nnm_wave_sort64(kk);
                const uint32_t c = (uint32_t)__popcll(nnm_ballot(kk != SR_NONE));
                if (c) {
                    bs = merge_round(L, cb, bs, lmax, kk, c, lane);
                    cb ^= 1u;
                }
            }
        }
        // best-first over the list's unexpanded entries
        uint32_t steps = 0;
        for (;;) {
            // the `width` first unexpanded entries (fixed-index selects keep
            // the small arrays in registers)
            uint32_t pick[SR_MAXW] = {0, 0, 0, 0};
            uint32_t np = 0;
            for (uint32_t b0 = 0; b0 < bs && np < a.width; b0 += NNM_WAVE) {
                const uint32_t b = b0 + lane;
                uint64_t m = nnm_ballot(b < bs && L.fl[cb][b] == 0);
                while (m && np < a.width) {
                    const uint32_t t = (uint32_t)__builtin_ctzll(m);
                    m &= m - 1;
#pragma unroll
                    for (uint32_t w = 0; w < SR_MAXW; ++w)
                        if (w == np) pick[w] = b0 + t;
                    ++np;
                }
            }
            if (np == 0) break;
            if (++steps > a.max_steps) {
                my_caps += lane == 0 ? 1 : 0;
                break;
            }
            uint32_t nb0[SR_MAXW] = {0, 0, 0, 0}, deg[SR_MAXW] = {0, 0, 0, 0};
            uint32_t m = 0;
            const SearchLayer ly = a.layer[a.nlayers - 1];
#pragma unroll
            for (uint32_t w = 0; w < SR_MAXW; ++w) {
                if (w < np) {
                    const uint32_t node = nnm_key_id(L.beam[cb][pick[w]]);
                    nb0[w] = ly.offs[node];
                    deg[w] = ly.offs[node + 1] - nb0[w];
                    m += deg[w];
                }
            }
            wave_sync_lds();
            if (lane == 0) {
#pragma unroll
                for (uint32_t w = 0; w < SR_MAXW; ++w)
                    if (w < np) L.fl[cb][pick[w]] = 1;
            }
            wave_sync_lds();
            for (uint32_t base = 0; base < m; base += NNM_WAVE) {
                const uint32_t idx = base + lane;
                uint64_t kk = SR_NONE;
                bool ins = false;
                if (idx < m) {
                    uint32_t r = idx, src = 0;
                    bool found = false;
#pragma unroll
                    for (uint32_t w = 0; w < SR_MAXW; ++w) {
                        if (!found && w < np) {
                            if (r < deg[w] || w + 1 == np) {
                                src = nb0[w] + r;
                                found = true;
                            } else {
                                r -= deg[w];
                            }
                        }
                    }
                    const uint32_t nb = ly.adj[src];
                    ins = hash_insert(L.hash, nb);
                    if (ins) {
                        kk = nnm_key(row_dist<BF16>(a, L.q, nb), nb);
                        ++my_d;
                        if (bs >= lmax && kk > L.beam[cb][lmax - 1]) kk = SR_NONE;  // src/lysearch.cpp:287
                    }
                }
                nins += (uint32_t)__popcll(nnm_ballot(ins));
                kk = nnm_wave_sort64(kk);
                const uint32_t c = (uint32_t)__popcll(nnm_ballot(kk != SR_NONE));
                if (c) {
                    bs = merge_round(L, cb, bs, lmax, kk, c, lane);
                    cb ^= 1u;
                }
            }
            if (nins >= SR_HS / 2) {
                // keep the hash exact for the list: clear, re-mark the list's ids
                for (uint32_t i = lane; i < SR_HS; i += NNM_WAVE) L.hash[i] = 0u;
                wave_sync_lds();
                for (uint32_t b = lane; b < bs; b += NNM_WAVE) hash_insert(L.hash, nnm_key_id(L.beam[cb][b]));
                wave_sync_lds();
                nins = bs;
                my_resets += lane == 0 ? 1 : 0;
            }
        }
        my_steps += lane == 0 ? steps : 0;
        // results
        const uint32_t rr = a.rerank ? a.rerank : beam;
        const uint32_t R = BF16 ? min(bs, max(a.k, min(rr, (uint32_t)NNM_WAVE))) : min(bs, a.k);
        uint64_t kk = SR_NONE;
        if (lane < R) {
            const uint32_t id = nnm_key_id(L.beam[cb][lane]);
            if (BF16) {
                kk = nnm_key(row_dist<false>(a, L.q, id), id);
                ++my_rr;
            } else {
                kk = L.beam[cb][lane];
            }
        }
        if (BF16) kk = nnm_wave_sort64(kk);
        const uint64_t ob = ((uint64_t)a.slot[bi] * a.nq + qi) * a.k;
        if (lane < a.k) {
            a.out_ids[ob + lane] = kk == SR_NONE ? 0xFFFFFFFFu : nnm_key_id(kk);
            a.out_d[ob + lane] = kk == SR_NONE ? 3.402823466e38f : nnm_key_dist(kk);
        }
        wave_sync_lds();
        my_ticks += lane == 0 ? wall_clock64() - t_start : 0;
    }
    if (cur_b < SR_MAXB) flush(cur_b);
}

}  // namespace

// ------------------------------------------------------------------ host

struct nnm_search_state {
    uint32_t nlayers = 0;
    uint32_t rows[SR_LAYERS] = {};
    uint32_t* offs[SR_LAYERS] = {};
    uint32_t* adj[SR_LAYERS] = {};
    bool owned[SR_LAYERS] = {};
};

static nnm_search_state* search_of(nnm_ctx* c) { return static_cast<nnm_search_state*>(c->search); }

void nnm_search_free(nnm_ctx* c) {
    nnm_search_state* S = search_of(c);
    if (!S) return;
    for (uint32_t g = 0; g < SR_LAYERS; ++g)
        if (S->owned[g]) {
            nnm_free(c, S->offs[g]);
            nnm_free(c, S->adj[g]);
        }
    delete S;
    c->search = nullptr;
}

static nnm_search_state* search_reset(nnm_ctx* c) {
    nnm_search_free(c);
    nnm_search_state* S = new nnm_search_state();
    c->search = S;
    return S;
}

// Validate a CSR layer on the host: offsets non-decreasing, neighbours
// inside the layer's rows (upper layers are prefixes of the base rows).
static int check_layer(nnm_ctx* c, uint32_t rows, const uint32_t* offs, const uint32_t* adj) {
    if (offs[0] != 0) return nnm_fail(c, NNM_ERR_INVALID, "layer offsets must start at 0");
    for (uint32_t i = 0; i < rows; ++i)
        if (offs[i + 1] < offs[i]) return nnm_fail(c, NNM_ERR_INVALID, "layer offsets must not decrease");
    for (uint64_t e = 0; e < offs[rows]; ++e)
        if (adj[e] >= rows) return nnm_fail(c, NNM_ERR_INVALID, "layer neighbour id outside the layer");
    return NNM_OK;
}

extern "C" int nnm_search_set_layers(nnm_ctx* c, uint32_t nlayers, const uint32_t* rows, const uint32_t* const* offs,
                                     const uint32_t* const* adj) {
    if (!c) return NNM_ERR_INVALID;
    if (!c->X) return nnm_fail(c, NNM_ERR_STATE, "no data");
    if (nlayers == 0 || nlayers > SR_LAYERS || !rows || !offs || !adj)
        return nnm_fail(c, NNM_ERR_INVALID, "nnm_search_set_layers: 1..8 layers with rows, offs and adj");
    for (uint32_t g = 0; g < nlayers; ++g) {
        if (rows[g] == 0 || rows[g] > c->n || !offs[g] || !adj[g])
            return nnm_fail(c, NNM_ERR_INVALID, "layer rows must be in [1, n]");
        if (g > 0 && rows[g] < rows[g - 1]) return nnm_fail(c, NNM_ERR_INVALID, "layers must not shrink");
        NNM_TRY(check_layer(c, rows[g], offs[g], adj[g]));
    }
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    nnm_search_state* S = search_reset(c);
    S->nlayers = nlayers;
    for (uint32_t g = 0; g < nlayers; ++g) {
        S->rows[g] = rows[g];
        S->owned[g] = true;
        const uint64_t ne = offs[g][rows[g]];
        NNM_TRY(nnm_alloc(c, (void**)&S->offs[g], ((uint64_t)rows[g] + 1) * 4));
        NNM_TRY(nnm_alloc(c, (void**)&S->adj[g], std::max<uint64_t>(ne, 1) * 4));
        NNM_TRY(nnm_h2d(c, S->offs[g], offs[g], ((uint64_t)rows[g] + 1) * 4));
        if (ne) NNM_TRY(nnm_h2d(c, S->adj[g], adj[g], ne * 4));
    }
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    return NNM_OK;
}

// The base layer as device CSR built by another part of the library (the
// DPG builder or the k-NN lists): ownership passes to the search state.
int nnm_search_adopt_layer(nnm_ctx* c, uint32_t rows, uint32_t* offs, uint32_t* adj) {
    nnm_search_state* S = search_reset(c);
    S->nlayers = 1;
    S->rows[0] = rows;
    S->offs[0] = offs;
    S->adj[0] = adj;
    S->owned[0] = true;
    return NNM_OK;
}

// One launch over nbeams beams (nbeams >= 1); st holds nbeams records.
static int search_impl(nnm_ctx* c, const float* Q, uint32_t nq, const nnm_search_params* p, const uint32_t* beams,
                       uint32_t nbeams, const uint32_t* seeds, uint32_t* ids, float* dist, nnm_search_stats* st) {
    if (!c) return NNM_ERR_INVALID;
    if (st && beams && nbeams <= SR_MAXB) std::memset(st, 0, sizeof(*st) * nbeams);
    NNM_TRY(nnm_require_l2(c, "nnm_search"));
    nnm_search_state* S = search_of(c);
    if (!S || S->nlayers == 0) return nnm_fail(c, NNM_ERR_STATE, "no search graph (nnm_search_set_layers)");
    if (!p || !Q || !ids || !dist || !beams) return nnm_fail(c, NNM_ERR_INVALID, "nnm_search: null argument");
    if (nbeams == 0 || nbeams > SR_MAXB) return nnm_fail(c, NNM_ERR_INVALID, "1..32 beams per launch");
    if (nq == 0) return NNM_OK;
    uint32_t min_beam = SR_L;
    for (uint32_t b = 0; b < nbeams; ++b) {
        if (beams[b] == 0 || beams[b] > SR_L) return nnm_fail(c, NNM_ERR_INVALID, "beam must be in [1, 256]");
        min_beam = std::min(min_beam, beams[b]);
    }
    if (p->k == 0 || p->k > min_beam || p->k > NNM_WAVE)
        return nnm_fail(c, NNM_ERR_INVALID, "k must be in [1, min(beam, 64)]");
    if (p->width == 0 || p->width > SR_MAXW) return nnm_fail(c, NNM_ERR_INVALID, "width must be in [1, 4]");
    if (p->prec != NNM_PREC_FP32 && p->prec != NNM_PREC_BF16)
        return nnm_fail(c, NNM_ERR_INVALID, "search precision must be FP32 or BF16");
    if (c->ld > SR_MAXLD) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "search supports d <= 256");
    const bool flat = p->mode == NNM_SEARCH_FLAT;
    if (!flat && p->mode != NNM_SEARCH_LAYERED) return nnm_fail(c, NNM_ERR_INVALID, "unknown search mode");
    if (!flat && p->entry >= S->rows[0]) return nnm_fail(c, NNM_ERR_INVALID, "entry must be a row of layer 0");
    const uint32_t base_rows = S->rows[S->nlayers - 1];
    if (seeds) {
        if (!flat) return nnm_fail(c, NNM_ERR_INVALID, "seeds are for the flat mode");
        if (nbeams != 1) return nnm_fail(c, NNM_ERR_INVALID, "explicit seeds take one beam per launch");
        for (uint64_t i = 0; i < (uint64_t)nq * beams[0]; ++i)
            if (seeds[i] >= base_rows) return nnm_fail(c, NNM_ERR_INVALID, "seed outside the base layer");
    }
    NNM_HIP(c, hipSetDevice(c->device));
    if (p->prec == NNM_PREC_BF16) NNM_TRY(nnm_ensure_bf16(c));
    const uint64_t nout = (uint64_t)nbeams * nq * p->k;
    float* dQ = nullptr;
    uint32_t *dS = nullptr, *dI = nullptr;
    float* dD = nullptr;
    unsigned long long* cnt = nullptr;
    int rc = NNM_OK;
    auto cleanup = [&]() {
        nnm_free(c, dQ); nnm_free(c, dS); nnm_free(c, dI); nnm_free(c, dD); nnm_free(c, cnt);
    };
    if ((rc = nnm_alloc(c, (void**)&dQ, (uint64_t)nq * c->ld * 4)) ||
        (rc = nnm_alloc(c, (void**)&dI, nout * 4)) || (rc = nnm_alloc(c, (void**)&dD, nout * 4)) ||
        (rc = nnm_alloc(c, (void**)&cnt, (uint64_t)SC_N * nbeams * 8)) ||
        (seeds && (rc = nnm_alloc(c, (void**)&dS, (uint64_t)nq * beams[0] * 4)))) {
        cleanup();
        return rc;
    }
    if ((rc = nnm_h2d_rows(c, dQ, c->ld, Q, c->d, nq)) ||
        (seeds && (rc = nnm_h2d(c, dS, seeds, (uint64_t)nq * beams[0] * 4))) ||
        (rc = nnm_fill_u64(c, reinterpret_cast<uint64_t*>(cnt), (uint64_t)SC_N * nbeams, 0))) {
        cleanup();
        return rc;
    }
    SearchArgs a;
    std::memset(&a, 0, sizeof(a));
    a.X = c->X;
    a.Xb = c->Xb;
    a.ld = c->ld;
    a.n = base_rows;
    a.nlayers = flat ? 1 : S->nlayers;
    for (uint32_t g = 0; g < a.nlayers; ++g) {
        const uint32_t src = flat ? S->nlayers - 1 : g;
        a.layer[g] = {S->offs[src], S->adj[src], S->rows[src]};
    }
    a.Q = dQ;
    a.nq = nq;
    a.seeds = dS;
    a.seed = p->seed;
    a.flat = flat ? 1 : 0;
    a.k = p->k;
    a.width = p->width;
    a.rerank = p->rerank;
    a.entry = p->entry;
    a.max_steps = 1u << 22;
    // launch order: largest beam first (stable for equal beams)
    uint32_t order[SR_MAXB];
    for (uint32_t b = 0; b < nbeams; ++b) order[b] = b;
    std::stable_sort(order, order + nbeams, [&](uint32_t x, uint32_t y) { return beams[x] > beams[y]; });
    a.nbeams = nbeams;
    for (uint32_t b = 0; b < nbeams; ++b) {
        a.beams[b] = beams[order[b]];
        a.slot[b] = order[b];
    }
    a.out_ids = dI;
    a.out_d = dD;
    a.counters = cnt;
    int per_cu = 0;
    const bool bf16 = p->prec == NNM_PREC_BF16;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bf16 ? k_search<true> : k_search<false>, 64, 0) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 4;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((uint64_t)c->num_cus * per_cu, (uint64_t)nq * nbeams);
    NNM_HIP(c, hipEventRecord(c->ev0, c->stream));
    if (bf16)
        hipLaunchKernelGGL(k_search<true>, dim3(grid), dim3(64), 0, c->stream, a);
    else
        hipLaunchKernelGGL(k_search<false>, dim3(grid), dim3(64), 0, c->stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        cleanup();
        return nnm_hip_fail(c, e, "k_search launch");
    }
    NNM_HIP(c, hipEventRecord(c->ev1, c->stream));
    std::vector<uint64_t> h((size_t)SC_N * nbeams);
    if ((rc = nnm_d2h(c, ids, dI, nout * 4)) || (rc = nnm_d2h(c, dist, dD, nout * 4)) ||
        (rc = nnm_d2h(c, h.data(), cnt, h.size() * 8))) {
        cleanup();
        return rc;
    }
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    cleanup();
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device) != hipSuccess || khz <= 0)
        khz = 100000;
    uint64_t ticks = 0, capped = 0;
    for (uint32_t b = 0; b < nbeams; ++b) {
        ticks += h[(size_t)b * SC_N + SC_TICKS];
        capped += h[(size_t)b * SC_N + SC_CAPPED];
    }
    if (st)
        for (uint32_t b = 0; b < nbeams; ++b) {
            const uint64_t* hb = h.data() + (size_t)b * SC_N;
            nnm_search_stats& o = st[order[b]];
            o.dists = hb[SC_DISTS];
            o.reranked = hb[SC_RERANKED];
            o.hash_resets = hb[SC_RESETS];
            o.capped = hb[SC_CAPPED];
            o.steps = hb[SC_STEPS];
            o.wave_secs = (double)hb[SC_TICKS] / ((double)khz * 1e3);
            o.launch_secs = ms * 1e-3;
            o.secs = nbeams == 1 ? ms * 1e-3 : ms * 1e-3 * (ticks ? (double)hb[SC_TICKS] / (double)ticks : 0.0);
        }
    if (capped) return nnm_fail(c, NNM_ERR_STATE, "search step cap reached");
    return NNM_OK;
}

extern "C" int nnm_search(nnm_ctx* c, const float* Q, uint32_t nq, const nnm_search_params* p, const uint32_t* seeds,
                          uint32_t* ids, float* dist, nnm_search_stats* st) {
    if (!c) return NNM_ERR_INVALID;
    if (!p) return nnm_fail(c, NNM_ERR_INVALID, "nnm_search: null argument");
    const uint32_t beam = p->beam;
    return search_impl(c, Q, nq, p, &beam, 1, seeds, ids, dist, st);
}

extern "C" int nnm_search_beams(nnm_ctx* c, const float* Q, uint32_t nq, const nnm_search_params* p,
                                const uint32_t* beams, uint32_t nbeams, uint32_t* ids, float* dist,
                                nnm_search_stats* st) {
    return search_impl(c, Q, nq, p, beams, nbeams, nullptr, ids, dist, st);
}

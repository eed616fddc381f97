// Offer delivery: every offer of the local join reaches its target's list.
//
// The reference inserts each accepted pair into both lists at once
// (NNList::parallel_try_insert -> UpdateKnnList, src/pvetex.h:40-85, 99-109);
// no candidate is lost.  Here one iteration's offers form a batch:
//
//   join      each wave appends its offers {dist, id, target} to log chunks
//             it reserved (nnm_internal.hpp, offer_emit)
//   passes    one or two partition passes group the records by target.
//             The first pass counts each block's records per bin into a
//             bin-major table, scans it and scatters every tile through LDS
//             (ranks in LDS, runs written out; no atomics); a second pass, when
//             the targets need more digits, counts per-tile LDS histograms
//             (one atomic per nonempty bin) and scatters with one cursor
//             atomic per nonempty bin of a tile.  The last pass's bins are
//             blocks of 32 targets
//   owners    one workgroup per 32-target bin sorts the bin's records by
//             target in LDS and one wave per target with records merges its
//             candidates into its list: de-dup by id (candidates and list),
//             candidates below the k-th key enter at their rank by (dist,
//             id) -- UpdateKnnList's rule for the whole batch.  Only the
//             owner writes a list.
//
// The merged list of a target is the K smallest distinct (dist, id) keys of
// its list and every offer, whatever the order in which the offers were
// written, partitioned or split into chunks.  A vertex whose offers did not
// fit the log is joined again in a redo round (its earlier offers then
// arrive twice, and the de-dup by id absorbs them).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "nnm_internal.hpp"

#define PT_THREADS 1024u
#define PT_PER 8u                          // records per thread in a tile
#define PT_TILE (PT_THREADS * PT_PER)      // records per tile (one log chunk)
#define PT_BINS 2048u                      // LDS bins of a tile (relative to its smallest digit)
#define PT_MAXBITS 10u                     // digit bits of one pass
#define OWN_T (1u << NNM_OWN_BITS)         // targets per owner bin
#define OWN_CH 4096u                       // records an owner sorts in LDS at a time
#define NO_DIGIT 0xFFFFFFFFu
#define P1_SEGS 8u                         // second pass: segments of each first-pass bin

static_assert(PT_TILE == NNM_OCH, "first-pass tiles are log chunks");
static_assert(PT_BINS == 2 * PT_THREADS, "two bins per thread in the LDS scan");
static_assert(OWN_T <= 64, "owner bins of at most 64 targets (one LDS counter per lane)");

// ------------------------------------------------------------------ passes

struct PassIn {
    const uint64_t* key;         // record keys
    const uint32_t* tgt;         // record targets
    const uint32_t* chunk_used;  // first pass: records per log chunk (tile c = chunk c)
    const uint64_t* counters;    // first pass: CNT_OCHUNKS
    uint32_t nchunks;            // log capacity in chunks
    const uint32_t* total;       // later passes: record count of the input
    uint32_t nbins;              // bins of the pass (records of other digits are not moved)
};

// Records [*b, *b + *m) of tile t (m = 0: no tile).
__device__ inline void tile_range(const PassIn& in, uint64_t t, uint64_t* b, uint32_t* m) {
    *b = t * PT_TILE;
    *m = 0;
    if (in.chunk_used) {
        const uint64_t used = min(in.counters[CNT_OCHUNKS], (uint64_t)in.nchunks);
        if (t < used) *m = min(in.chunk_used[t], PT_TILE);
    } else {
        const uint64_t tot = *in.total;
        if (*b < tot) *m = (uint32_t)min((uint64_t)PT_TILE, tot - *b);
    }
}

// Tiles of the pass input: the reserved log chunks (first pass) or the
// record count in tiles.
__device__ inline uint64_t tile_count(const PassIn& in) {
    if (in.chunk_used) return min(in.counters[CNT_OCHUNKS], (uint64_t)in.nchunks);
    return ((uint64_t)*in.total + PT_TILE - 1) / PT_TILE;
}

// Smallest digit of the tile into *dmin_s (LDS; preset to NO_DIGIT).
__device__ inline uint32_t block_min_digit(const uint32_t (&d)[PT_PER], uint32_t* dmin_s) {
    uint32_t lm = NO_DIGIT;
#pragma unroll
    for (uint32_t k = 0; k < PT_PER; ++k) lm = min(lm, d[k]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) lm = min(lm, (uint32_t)__shfl_xor((int)lm, o, NNM_WAVE));
    if (__lane_id() == 0) atomicMin(dmin_s, lm);
    __syncthreads();
    return *dmin_s;
}

// Digits of thread tid's records of a tile (NO_DIGIT past the tile or the bins).
__device__ inline void tile_digits(const PassIn& in, uint64_t b, uint32_t m, uint32_t shift, uint32_t tid,
                                   uint32_t (&d)[PT_PER]) {
#pragma unroll
    for (uint32_t k = 0; k < PT_PER; ++k) {
        const uint32_t j = tid + k * PT_THREADS;
        d[k] = j < m ? in.tgt[b + j] >> shift : NO_DIGIT;
        if (d[k] >= in.nbins) d[k] = NO_DIGIT;
    }
}

// Exclusive prefix of the PT_BINS LDS counts hist[] into hist[] (thread tid
// owns bins 2 tid, 2 tid + 1); returns the total.  c2 gets the thread's counts.
__device__ inline uint32_t block_prefix_bins(uint32_t* hist, uint32_t* wsum, uint32_t (&c2)[2]) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    c2[0] = hist[2 * tid];
    c2[1] = hist[2 * tid + 1];
    const uint32_t s2 = c2[0] + c2[1];
    uint32_t x = s2;
#pragma unroll
    for (int o = 1; o < NNM_WAVE; o <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)x, (unsigned)o, NNM_WAVE);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == NNM_WAVE - 1) wsum[wave] = x;
    __syncthreads();
    uint32_t run = x - s2, tot = 0;
#pragma unroll
    for (uint32_t w = 0; w < PT_THREADS / NNM_WAVE; ++w) {
        run += w < wave ? wsum[w] : 0u;
        tot += wsum[w];
    }
    hist[2 * tid] = run;
    hist[2 * tid + 1] = run + c2[0];
    __syncthreads();
    return tot;
}

// Bin counts of one pass: cnt[target >> shift] += records (reads the targets
// only).  Passes after the first when the owner-bin digits need more than
// two passes.
__global__ __launch_bounds__(PT_THREADS) void k_part_count(PassIn in, uint32_t shift, uint32_t* __restrict__ cnt) {
    __shared__ uint32_t hist[PT_BINS];
    __shared__ uint32_t dmin_s;
    const uint32_t tid = threadIdx.x;
    const uint64_t nt = tile_count(in);
    for (uint64_t t = blockIdx.x; t < nt; t += gridDim.x) {
        uint64_t b;
        uint32_t m;
        tile_range(in, t, &b, &m);
        if (m == 0) continue;  // uniform over the block
        for (uint32_t i = tid; i < PT_BINS; i += PT_THREADS) hist[i] = 0;
        if (tid == 0) dmin_s = NO_DIGIT;
        __syncthreads();
        uint32_t d[PT_PER];
        tile_digits(in, b, m, shift, tid, d);
        const uint32_t dmin = block_min_digit(d, &dmin_s);
#pragma unroll
        for (uint32_t k = 0; k < PT_PER; ++k) {
            if (d[k] == NO_DIGIT) continue;
            const uint32_t r = d[k] - dmin;
            if (r < PT_BINS) atomicAdd(&hist[r], 1u);
            else atomicAdd(&cnt[d[k]], 1u);  // digits far above the tile's smallest
        }
        __syncthreads();
        for (uint32_t i = tid; i < PT_BINS; i += PT_THREADS)
            if (hist[i]) atomicAdd(&cnt[dmin + i], hist[i]);
        __syncthreads();  // hist and dmin_s are reset for the next tile
    }
}

// Scatter of one pass: every record to its bin at cur[digit]++ (cursors start
// at the scanned bin offsets).  The tile is first ordered by bin in LDS, so
// the writes go out as runs of consecutive records per bin, one cursor atomic
// per nonempty bin.
__global__ __launch_bounds__(PT_THREADS) void k_part_scatter(PassIn in, uint32_t shift, uint32_t* __restrict__ cur,
                                                             uint64_t* __restrict__ okey, uint32_t* __restrict__ otgt) {
    __shared__ uint64_t lkey[PT_TILE];
    __shared__ uint32_t ltgt[PT_TILE];
    __shared__ uint32_t hist[PT_BINS];   // counts, then their exclusive prefix
    __shared__ uint32_t basev[PT_BINS];  // global position of each bin's run
    __shared__ uint32_t wsum[PT_THREADS / NNM_WAVE];
    __shared__ uint32_t dmin_s;
    const uint32_t tid = threadIdx.x;
    const uint64_t nt = tile_count(in);
    for (uint64_t t = blockIdx.x; t < nt; t += gridDim.x) {
        uint64_t b;
        uint32_t m;
        tile_range(in, t, &b, &m);
        if (m == 0) continue;
        for (uint32_t i = tid; i < PT_BINS; i += PT_THREADS) hist[i] = 0;
        if (tid == 0) dmin_s = NO_DIGIT;
        __syncthreads();
        uint32_t d[PT_PER], tg[PT_PER], rk[PT_PER];
        uint64_t ky[PT_PER];
#pragma unroll
        for (uint32_t k = 0; k < PT_PER; ++k) {
            const uint32_t j = tid + k * PT_THREADS;
            tg[k] = 0;
            ky[k] = 0;
            d[k] = NO_DIGIT;
            if (j < m) {
                tg[k] = in.tgt[b + j];
                ky[k] = in.key[b + j];
                d[k] = tg[k] >> shift;
                if (d[k] >= in.nbins) d[k] = NO_DIGIT;
            }
        }
        const uint32_t dmin = block_min_digit(d, &dmin_s);
#pragma unroll
        for (uint32_t k = 0; k < PT_PER; ++k) {
            rk[k] = 0;
            if (d[k] == NO_DIGIT) continue;
            const uint32_t r = d[k] - dmin;
            if (r < PT_BINS) {
                rk[k] = atomicAdd(&hist[r], 1u);
            } else {
                // digits far above the tile's smallest: written one by one
                const uint32_t p = atomicAdd(&cur[d[k]], 1u);
                okey[p] = ky[k];
                otgt[p] = tg[k];
                d[k] = NO_DIGIT;
            }
        }
        __syncthreads();
        // global runs of the nonempty bins
        for (uint32_t i = tid; i < PT_BINS; i += PT_THREADS)
            if (hist[i]) basev[i] = atomicAdd(&cur[dmin + i], hist[i]);
        uint32_t c2[2];
        const uint32_t nl = block_prefix_bins(hist, wsum, c2);  // records staged in LDS
#pragma unroll
        for (uint32_t k = 0; k < PT_PER; ++k) {
            if (d[k] == NO_DIGIT) continue;
            const uint32_t lp = hist[d[k] - dmin] + rk[k];
            lkey[lp] = ky[k];
            ltgt[lp] = tg[k];
        }
        __syncthreads();
        for (uint32_t lp = tid; lp < nl; lp += PT_THREADS) {
            const uint32_t tgv = ltgt[lp];
            const uint32_t r = (tgv >> shift) - dmin;
            const uint32_t p = basev[r] + (lp - hist[r]);
            okey[p] = lkey[lp];
            otgt[p] = tgv;
        }
        __syncthreads();
    }
}

// First pass (at most PT_BINS bins): each block counts the records of the
// tiles it will scatter, per bin, over all its tiles; bcnt[bin * grid +
// block] (bin-major, so that one scan gives every block the start of its run
// in every bin: no cursor atomics, which would all meet on a few hundred
// addresses).
__global__ __launch_bounds__(PT_THREADS) void k_part0_count(PassIn in, uint32_t shift, uint32_t* __restrict__ bcnt) {
    __shared__ uint32_t hist[PT_BINS];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < PT_BINS; i += PT_THREADS) hist[i] = 0;
    __syncthreads();
    const uint64_t nt = tile_count(in);  // reserved log chunks (each PT_TILE records of capacity)
    for (uint64_t t = blockIdx.x; t < nt; t += gridDim.x) {
        // the whole chunk is read (no branches), records past its fill count are not counted
        const uint32_t m = min(in.chunk_used[t], PT_TILE);
        uint32_t tg[PT_PER];
#pragma unroll
        for (uint32_t k = 0; k < PT_PER; ++k) tg[k] = in.tgt[t * PT_TILE + tid + k * PT_THREADS];
#pragma unroll
        for (uint32_t k = 0; k < PT_PER; ++k) {
            const uint32_t d = tg[k] >> shift;
            if (tid + k * PT_THREADS < m && d < in.nbins) atomicAdd(&hist[d], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < in.nbins; i += PT_THREADS) bcnt[(uint64_t)i * gridDim.x + blockIdx.x] = hist[i];
}

// Records [b, b + m) of a tile into registers (thread tid's k-th: b + tid +
// k PT_THREADS; past m, a repeat of record b + m - 1).  The loads carry no
// branches, so they all go out before the first is waited for.
__device__ inline void tile_load(const uint64_t* __restrict__ key, const uint32_t* __restrict__ tgt, uint64_t b,
                                 uint32_t m, uint32_t tid, uint32_t (&tg)[PT_PER], uint64_t (&ky)[PT_PER]) {
    if (m == 0) return;
#pragma unroll
    for (uint32_t k = 0; k < PT_PER; ++k) {
        const uint32_t j = min(tid + k * PT_THREADS, m - 1);
        tg[k] = tgt[b + j];
        ky[k] = key[b + j];
    }
}

// A tile in registers, by digit d[k] (NO_DIGIT: not moved), out to its bins:
// ordered by digit in LDS, then written in runs from position curs[d] +
// (rank within the tile's digit d).  The cursors advance by the tile's
// counts.  Starts and ends at a block barrier.
__device__ inline void tile_scatter(const uint32_t (&d)[PT_PER], const uint32_t (&tg)[PT_PER],
                                    const uint64_t (&ky)[PT_PER], uint32_t shift, uint32_t base,
                                    uint64_t* lkey, uint32_t* ltgt, uint32_t* hist, uint32_t* curs, uint32_t* wsum,
                                    uint64_t* __restrict__ okey, uint32_t* __restrict__ otgt) {
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < PT_BINS; i += PT_THREADS) hist[i] = 0;
    __syncthreads();
    uint32_t rk[PT_PER];
#pragma unroll
    for (uint32_t k = 0; k < PT_PER; ++k) rk[k] = d[k] != NO_DIGIT ? atomicAdd(&hist[d[k]], 1u) : 0u;
    __syncthreads();
    uint32_t c2[2];
    const uint32_t nl = block_prefix_bins(hist, wsum, c2);
    // the bin starts first (unconditional reads), then the stores: the
    // reads are not held back behind the stores
    uint32_t lp[PT_PER];
#pragma unroll
    for (uint32_t k = 0; k < PT_PER; ++k) lp[k] = hist[min(d[k], PT_BINS - 1)] + rk[k];
#pragma unroll
    for (uint32_t k = 0; k < PT_PER; ++k) {
        if (d[k] == NO_DIGIT) continue;
        lkey[lp[k]] = ky[k];
        ltgt[lp[k]] = tg[k];
    }
    __syncthreads();
    if (nl) {
        // all LDS reads first (positions past nl read the last record and
        // store nothing): one LDS round trip per level, not one per record
        uint32_t tgv[PT_PER], p[PT_PER];
        uint64_t kv[PT_PER];
#pragma unroll
        for (uint32_t k = 0; k < PT_PER; ++k) {
            const uint32_t lp = min(tid + k * PT_THREADS, nl - 1);
            tgv[k] = ltgt[lp];
            kv[k] = lkey[lp];
        }
#pragma unroll
        for (uint32_t k = 0; k < PT_PER; ++k) {
            const uint32_t lp = min(tid + k * PT_THREADS, nl - 1);
            const uint32_t r = (tgv[k] >> shift) - base;
            p[k] = curs[r] + (lp - hist[r]);
        }
#pragma unroll
        for (uint32_t k = 0; k < PT_PER; ++k) {
            if (tid + k * PT_THREADS < nl) {
                okey[p[k]] = kv[k];
                otgt[p[k]] = tgv[k];
            }
        }
    }
    __syncthreads();
    curs[2 * tid] += c2[0];
    curs[2 * tid + 1] += c2[1];
    __syncthreads();
}

// First-pass scatter: the same blocks over the same tiles; block runs start
// at the scanned per-block counts boff[bin * grid + block].  The next tile
// (a whole log chunk, its fill count beside it) is read while the current one
// is ordered and written; records past the fill count are not moved.
__global__ __launch_bounds__(PT_THREADS) void k_part0_scatter(PassIn in, uint32_t shift,
                                                              const uint32_t* __restrict__ boff,
                                                              uint64_t* __restrict__ okey, uint32_t* __restrict__ otgt) {
    __shared__ uint64_t lkey[PT_TILE];
    __shared__ uint32_t ltgt[PT_TILE];
    __shared__ uint32_t hist[PT_BINS];   // per tile: counts, then their exclusive prefix
    __shared__ uint32_t curs[PT_BINS];   // the block's next global position in every bin
    __shared__ uint32_t wsum[PT_THREADS / NNM_WAVE];
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = tid; i < in.nbins; i += PT_THREADS) curs[i] = boff[(uint64_t)i * gridDim.x + blockIdx.x];
    const uint64_t nt = tile_count(in);  // reserved log chunks (each PT_TILE records of capacity)
    uint64_t t = blockIdx.x;
    uint32_t m = 0;
    uint32_t tg[PT_PER];
    uint64_t ky[PT_PER];
    if (t < nt) {
        m = min(in.chunk_used[t], PT_TILE);
        tile_load(in.key, in.tgt, t * PT_TILE, PT_TILE, tid, tg, ky);
    }
    // The first tile is waited for here: in the loop, the only loads in
    // flight are then the next tile's, and the current tile's first use
    // does not wait for them.
    __builtin_amdgcn_s_waitcnt(0);
    for (; t < nt; t += gridDim.x) {
        uint32_t d[PT_PER];
#pragma unroll
        for (uint32_t k = 0; k < PT_PER; ++k) {
            d[k] = tid + k * PT_THREADS < m ? tg[k] >> shift : NO_DIGIT;
            if (d[k] >= in.nbins) d[k] = NO_DIGIT;
        }
        // the next tile (past the last, the last again, not used): loads
        // without branches, waited for only after this tile is written
        const uint64_t tn = t + gridDim.x, tl = min(tn, nt - 1);
        const uint32_t mraw = in.chunk_used[tl];
        uint32_t ntg[PT_PER];
        uint64_t nky[PT_PER];
        tile_load(in.key, in.tgt, tl * PT_TILE, PT_TILE, tid, ntg, nky);
        if (m) tile_scatter(d, tg, ky, shift, 0u, lkey, ltgt, hist, curs, wsum, okey, otgt);
#pragma unroll
        for (uint32_t k = 0; k < PT_PER; ++k) {
            tg[k] = ntg[k];
            ky[k] = nky[k];
        }
        m = tn < nt ? min(mraw, PT_TILE) : 0u;
    }
}

// Second pass when each first-pass bin r holds nl = 2^b1 owner bins: the bin's
// records (contiguous in the first pass's output, [boff[r * G], boff[(r + 1) *
// G])) are split into P1_SEGS equal segments.  Counts per (bin, owner bin,
// segment) go to cnt[(r * nl + l) * P1_SEGS + seg]; their exclusive scan in
// that order is every segment's start in every owner bin, so the scatter
// needs no atomics and owner bin ob ends where entry (ob + 1) * P1_SEGS
// starts.
__device__ inline void p1_segment(const uint32_t* boff, uint32_t G, uint32_t r, uint32_t seg, uint32_t* s,
                                  uint32_t* e) {
    const uint32_t bs = boff[(uint64_t)r * G], be = boff[(uint64_t)(r + 1) * G];
    const uint32_t len = be - bs;
    *s = bs + (uint32_t)((uint64_t)len * seg / P1_SEGS);
    *e = bs + (uint32_t)((uint64_t)len * (seg + 1) / P1_SEGS);
}

__global__ __launch_bounds__(PT_THREADS) void k_part1_count(const uint32_t* __restrict__ tgt,
                                                            const uint32_t* __restrict__ boff, uint32_t G,
                                                            uint32_t nbins0, uint32_t b1, uint32_t* __restrict__ cnt) {
    __shared__ uint32_t hist[PT_BINS];
    const uint32_t tid = threadIdx.x, nl = 1u << b1;
    for (uint32_t u = blockIdx.x; u < nbins0 * P1_SEGS; u += gridDim.x) {
        const uint32_t r = u / P1_SEGS, seg = u % P1_SEGS, ob0 = r << b1;
        uint32_t s, e;
        p1_segment(boff, G, r, seg, &s, &e);
        for (uint32_t i = tid; i < nl; i += PT_THREADS) hist[i] = 0;
        __syncthreads();
        for (uint32_t i0 = s; i0 < e; i0 += PT_TILE) {
            uint32_t tg[PT_PER];
#pragma unroll
            for (uint32_t k = 0; k < PT_PER; ++k) tg[k] = tgt[min(i0 + tid + k * PT_THREADS, e - 1)];  // no branches
#pragma unroll
            for (uint32_t k = 0; k < PT_PER; ++k)
                if (i0 + tid + k * PT_THREADS < e) atomicAdd(&hist[(tg[k] >> NNM_OWN_BITS) - ob0], 1u);
        }
        __syncthreads();
        for (uint32_t i = tid; i < nl; i += PT_THREADS) cnt[((uint64_t)r * nl + i) * P1_SEGS + seg] = hist[i];
        __syncthreads();
    }
}

__global__ __launch_bounds__(PT_THREADS) void k_part1_scatter(const uint64_t* __restrict__ key,
                                                              const uint32_t* __restrict__ tgt,
                                                              const uint32_t* __restrict__ boff, uint32_t G,
                                                              uint32_t nbins0, uint32_t b1,
                                                              const uint32_t* __restrict__ off,
                                                              uint64_t* __restrict__ okey, uint32_t* __restrict__ otgt) {
    __shared__ uint64_t lkey[PT_TILE];
    __shared__ uint32_t ltgt[PT_TILE];
    __shared__ uint32_t hist[PT_BINS];   // per tile: counts, then their exclusive prefix
    __shared__ uint32_t curs[PT_BINS];   // the segment's next output position in every owner bin
    __shared__ uint32_t wsum[PT_THREADS / NNM_WAVE];
    const uint32_t tid = threadIdx.x, nl = 1u << b1;
    for (uint32_t u = blockIdx.x; u < nbins0 * P1_SEGS; u += gridDim.x) {
        const uint32_t r = u / P1_SEGS, seg = u % P1_SEGS, ob0 = r << b1;
        uint32_t s, e;
        p1_segment(boff, G, r, seg, &s, &e);
        uint32_t tg[PT_PER];
        uint64_t ky[PT_PER];
        tile_load(key, tgt, s, s < e ? min(PT_TILE, e - s) : 0u, tid, tg, ky);
        for (uint32_t i = tid; i < nl; i += PT_THREADS) curs[i] = off[((uint64_t)r * nl + i) * P1_SEGS + seg];
        __builtin_amdgcn_s_waitcnt(0);  // the segment's first tile (then only the next tile's loads are in flight)
        for (uint32_t t0 = s; t0 < e; t0 += PT_TILE) {
            const uint32_t m = min(PT_TILE, e - t0);
            uint32_t d[PT_PER];
#pragma unroll
            for (uint32_t k = 0; k < PT_PER; ++k)
                d[k] = tid + k * PT_THREADS < m ? (tg[k] >> NNM_OWN_BITS) - ob0 : NO_DIGIT;
            const uint32_t tn = t0 + PT_TILE;
            uint32_t ntg[PT_PER];
            uint64_t nky[PT_PER];
            tile_load(key, tgt, tn, tn < e ? min(PT_TILE, e - tn) : 0u, tid, ntg, nky);
            tile_scatter(d, tg, ky, NNM_OWN_BITS, ob0, lkey, ltgt, hist, curs, wsum, okey, otgt);
#pragma unroll
            for (uint32_t k = 0; k < PT_PER; ++k) {
                tg[k] = ntg[k];
                ky[k] = nky[k];
            }
        }
    }
}

// ------------------------------------------------------------------ owners

struct OwnArgs {
    uint64_t n;                 // active rows
    uint32_t K;
    const float* X;
    const uint16_t* Xb_exact;   // bf16 rows where they are exact (same bits as X)
    uint32_t ld;
    float rerank_tau;
    float tie_win;
    uint32_t* g_ids;
    float* g_dist;
    uint8_t* g_flag;
    const uint64_t* key;        // records grouped by owner bin: keys
    const uint32_t* tgt;        // ... and targets
    const uint32_t* bin_end;    // bin b holds records [bin_end[(b - 1) * stride], bin_end[b * stride]) (from 0 for b = 0)
    uint32_t bin_stride;
    uint32_t nbins;
    uint64_t* counters;
};

// A chunk's records [c0, c0 + m) into registers, thread tid's k-th at
// c0 + tid + k NT (past m: the last record again; nothing is loaded when m
// is 0).  No branches between the loads: all are in flight before the
// first wait.  tg gets the targets.
template <uint32_t PER, uint32_t NT>
__device__ inline void own_chunk_load(const OwnArgs& a, uint32_t c0, uint32_t m, uint64_t (&key)[PER],
                                      uint32_t (&tg)[PER]) {
    const uint32_t tid = threadIdx.x;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        key[k] = NNM_EMPTY_KEY;
        tg[k] = 0xFFFFFFFFu;
    }
    if (m == 0) return;
#pragma unroll
    for (uint32_t k = 0; k < PER; ++k) {
        const uint64_t j = (uint64_t)c0 + min(tid + NT * k, m - 1);
        key[k] = a.key[j];
        tg[k] = a.tgt[j];
    }
}

// FP64 squared L2 across the wave: lanes take elements lane, lane + 64, ...;
// the partial sums are combined in a fixed tree (same result every call).
__device__ inline double own_l2_f64(const float* __restrict__ a, const float* __restrict__ b, uint32_t d) {
    double acc = 0.0;
    for (uint32_t i = nnm_lane(); i < d; i += NNM_WAVE) {
        const double t = (double)a[i] - (double)b[i];
        acc = __builtin_fma(t, t, acc);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, NNM_WAVE);
    return acc;
}

// Orders the wave's LDS accesses across lanes: the stores before it are
// seen by the loads after it, and no load before it is moved past it (a
// store of another lane after it cannot clobber what an earlier load reads).
__device__ inline void own_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct OwnTally {
    uint64_t upd = 0, flags = 0, rr = 0, ties = 0, swaps = 0;  // wave-uniform
};

// Candidate count above which the ranks come from a sort of the candidates
// instead of a loop over them.
#define OWN_SORT_MIN 32u

// Bitonic sort of the keys in lanes [0, S) ascending (S = 2^lg, lg <= 6;
// the other lanes hold the empty key and stay out of the way).
__device__ inline uint64_t own_sort_prefix(uint64_t key, uint32_t lg) {
    if (lg >= 1) key = bitonic_pass64<2, 1>(key);
    if (lg >= 2) key = bitonic_pass64<4, 2>(key);
    if (lg >= 3) key = bitonic_pass64<8, 4>(key);
    if (lg >= 4) key = bitonic_pass64<16, 8>(key);
    if (lg >= 5) key = bitonic_pass64<32, 16>(key);
    if (lg >= 6) key = bitonic_pass64<64, 32>(key);
    return key;
}

// Keys of the lanes with keep set moved to lanes 0..popcount - 1 in lane
// order; the empty key above them.  perm: the wave's 64 words of LDS.
__device__ inline uint64_t own_compact(uint64_t key, bool keep, uint32_t* perm) {
    const uint32_t lane = nnm_lane();
    const uint64_t m = __ballot(keep);
    const uint32_t nc = (uint32_t)__popcll(m);
    if (m != (nc == NNM_WAVE ? ~0ull : (1ull << nc) - 1)) {
        own_wave_sync();  // earlier readers of perm are done
        if (keep) perm[nnm_mask_rank(m)] = lane;
        own_wave_sync();
        key = shfl64(key, lane < nc ? perm[lane] : lane);
    }
    return lane < nc ? key : NNM_EMPTY_KEY;
}

__device__ inline void own_cswap(uint64_t& x, uint64_t& y) {
    const uint64_t lo = x < y ? x : y, hi = x < y ? y : x;
    x = lo;
    y = hi;
}

// Merge `len` candidate keys (keys[0..len), any order, repeats allowed) into
// the list of row v: UpdateKnnList (src/pvetex.h:40-85) for the batch.  A
// candidate enters when its id is not listed and its (dist, id) key ranks
// among the K smallest; with RR its lower-bound key is first recomputed
// exactly; with TIE the entries ranked K-1 and K are re-decided on FP64
// distances when their FP32 ones are within rounding.
//
// The candidates come 256 at a time, four per lane, sorted inside each lane;
// round r merges every lane's r-th smallest.  After round 0 the list's k-th
// key has dropped to the best of the lanes' minima, so later rounds keep few
// candidates (and none at all once a round keeps none: each lane's keys
// only grow).  A round: filter by the k-th key, de-dup against the list by
// id, sort the survivors (a network sized to their count), drop repeats of
// one pair (equal keys), then merge with the list by a bitonic merge.
//
// `perm` is 64 words of LDS owned by the wave.  A target whose candidates
// span several chunks of its bin (or several redo rounds) is merged once per
// chunk.  Entries that entered in this iteration's delivery are stored with
// flag 2 (new, and entered now); the next list build stores them as 1.  Each
// merge adds its change of the row's new flags and of its entered entries to
// the tallies: summed over the delivery, the entries that entered and
// stayed (the iteration's updates, the same count in whatever order the
// candidates came) and the change of the new flags since the list build,
// which counted the rest (the delta test).
//
// IDD: one id may come with different keys (lower bounds from the bf16 join
// and exact distances from the FP32 long-list join, without re-rank), so
// the candidates of a round are de-duplicated by id first; otherwise a pair
// has the same key whichever window offered it and equal keys are repeats.
// A target's list in registers: entry j in lane j.
struct OwnList {
    uint32_t id = 0xFFFFFFFFu;
    float dist = 0.f;
    uint32_t flag = 0;
};

__device__ inline OwnList own_load(const OwnArgs& a, uint64_t v) {
    OwnList l;
    const uint32_t lane = nnm_lane();
    if (v < a.n && lane < a.K) {
        l.id = a.g_ids[v * a.K + lane];
        l.dist = a.g_dist[v * a.K + lane];
        l.flag = a.g_flag[v * a.K + lane];
    }
    return l;
}

template <bool RR, bool TIE, bool IDD>
__device__ inline void own_merge(const OwnArgs& a, uint64_t v, const OwnList& in, const uint64_t* keys, uint32_t len,
                                 uint32_t* perm, OwnTally& t) {
    const uint32_t lane = nnm_lane();
    const uint32_t K = a.K;
    const uint64_t lmask = K < NNM_WAVE ? (1ull << K) - 1 : ~0ull;  // list lanes
    uint32_t pid = in.id, pf = in.flag;
    float pd = in.dist;
    const uint32_t fresh0 = lane < K && pf == 2 ? 1u : 0u;
    uint32_t fresh = fresh0;  // the entry entered in this iteration's delivery
    pf = pf ? 1u : 0u;
    const uint32_t f0 = (uint32_t)__popcll(__ballot(lane < K && pf)), e0 = (uint32_t)__popcll(__ballot(fresh));
    bool changed = false;
    for (uint32_t g0 = 0; g0 < len; g0 += 4 * NNM_WAVE) {
        uint64_t kk[4];
#pragma unroll
        for (uint32_t r = 0; r < 4; ++r) {
            const uint32_t i = g0 + lane + NNM_WAVE * r;
            kk[r] = i < len ? keys[i] : NNM_EMPTY_KEY;
        }
        own_cswap(kk[0], kk[1]);
        own_cswap(kk[2], kk[3]);
        own_cswap(kk[0], kk[2]);
        own_cswap(kk[1], kk[3]);
        own_cswap(kk[1], kk[2]);
        const uint32_t rounds = min(4u, (len - g0 + NNM_WAVE - 1) / NNM_WAVE);
        for (uint32_t r = 0; r < rounds; ++r) {
            uint64_t ck = r == 0 ? kk[0] : r == 1 ? kk[1] : r == 2 ? kk[2] : kk[3];
            const uint64_t pkey = lane < K ? nnm_key(pd, pid) : NNM_EMPTY_KEY;
            const uint64_t kth_key = wave_read64(pkey, K - 1);
            const float kd = nnm_key_dist(kth_key);
            // cheap filter first: lower bounds against the margin, exact keys
            // against the k-th key (or its rounding window)
            bool valid;
            if constexpr (RR) valid = ck != NNM_EMPTY_KEY && nnm_key_dist(ck) <= kd * a.rerank_tau;
            else if constexpr (TIE) valid = ck != NNM_EMPTY_KEY && (ck < kth_key || nnm_key_dist(ck) <= kd * a.tie_win);
            else valid = ck < kth_key;  // the empty key is above every key
            if (!__ballot(valid)) break;  // the later rounds' keys are larger still
            if constexpr (IDD) {
                // de-dup by id: sort by (id, dist) and keep the first of each id
                // (its smallest key)
                uint64_t sk = valid ? (((uint64_t)nnm_key_id(ck) << 32) | (ck >> 32)) : NNM_EMPTY_KEY;
                sk = nnm_wave_sort64(sk);
                const uint64_t prev = shfl64(sk, lane ? lane - 1 : 0);
                valid = sk != NNM_EMPTY_KEY && (lane == 0 || (prev >> 32) != (sk >> 32));
                ck = valid ? ((sk << 32) | (sk >> 32)) : NNM_EMPTY_KEY;
            }
            const uint32_t cid = nnm_key_id(ck);
            uint64_t vb = __ballot(valid);
            // de-dup against the list: a loop over the fewer of candidates and
            // list entries
            if ((uint32_t)__popcll(vb) <= K) {
                for (uint64_t b = vb; b; b &= b - 1) {
                    const uint32_t cl = (uint32_t)__builtin_ctzll(b);
                    const uint32_t q = (uint32_t)__builtin_amdgcn_readlane((int)cid, (int)cl);
                    if (__ballot(pid == q) & lmask) vb &= ~(1ull << cl);
                }
                valid = (vb >> lane) & 1ull;
            } else {
                for (uint32_t p = 0; p < K; ++p)
                    valid = valid && (uint32_t)__builtin_amdgcn_readlane((int)pid, (int)p) != cid;
                vb = __ballot(valid);
            }
            if constexpr (RR) {
                // exact FP32 distances of the survivors, then the exact filter
                if (valid) {
                    const float e = a.Xb_exact
                                        ? nnm_l2_rows_bf16<1>(a.Xb_exact + v * a.ld, a.Xb_exact + (uint64_t)cid * a.ld, a.ld)
                                        : nnm_l2_rows<1>(a.X + v * a.ld, a.X + (uint64_t)cid * a.ld, a.ld);
                    ck = nnm_key(e, cid);
                }
                t.rr += (uint32_t)__popcll(vb);
                valid = valid && (ck < kth_key || (TIE && nnm_key_dist(ck) <= kd * a.tie_win));
                vb = __ballot(valid);
            }
            if (!vb) continue;
            // survivors to lanes 0..nv-1, sorted by a network sized to them;
            // the repeats of one pair (equal keys) are then adjacent: keep the
            // first of each
            const uint32_t nv = (uint32_t)__popcll(vb);
            ck = own_compact(ck, valid, perm);
            ck = own_sort_prefix(ck, nv > 1 ? 32u - (uint32_t)__builtin_clz(nv - 1) : 0u);
            {
                const uint64_t prev = shfl64(ck, lane ? lane - 1 : 0);
                ck = own_compact(ck, ck != NNM_EMPTY_KEY && (lane == 0 || prev != ck), perm);
            }
            // The candidates are now distinct from the list and from each
            // other, so every key is distinct.  The first 64 - K candidates
            // ascending and the list descending form one bitonic sequence of
            // 64 keys; six half-cleaner steps sort it, and the first K lanes
            // are the new list (at most K candidates can enter, and 64 - K >= K
            // for K <= 32).  A payload word carries each entry's flag and
            // whether it entered in this delivery.
            const uint32_t L = NNM_WAVE - K;
            const uint32_t src = lane >= L ? NNM_WAVE - 1 - lane : lane;
            const uint64_t lk = shfl64(pkey, src);
            const uint32_t lpay = (uint32_t)__shfl((int)(pf | (fresh << 1)), (int)src, NNM_WAVE);
            uint64_t sk = lane >= L ? lk : ck;
            uint32_t pay = lane >= L ? lpay : 3u;  // a candidate enters new (flag 1) and fresh
            auto half_clean = [&](auto J) {
                constexpr uint32_t j = decltype(J)::value;
                const uint64_t ok_ = wave_xor64<j>(sk);
                const uint32_t op = wave_xor32<j>(pay);
                const bool lower = (lane & j) == 0;
                const bool take = lower ? ok_ < sk : ok_ > sk;
                sk = take ? ok_ : sk;
                pay = take ? op : pay;
            };
            half_clean(std::integral_constant<uint32_t, 32>());
            half_clean(std::integral_constant<uint32_t, 16>());
            half_clean(std::integral_constant<uint32_t, 8>());
            half_clean(std::integral_constant<uint32_t, 4>());
            half_clean(std::integral_constant<uint32_t, 2>());
            half_clean(std::integral_constant<uint32_t, 1>());
            if constexpr (TIE) {
                // the entries ranked K-1 (last kept) and K (first dropped): when
                // their FP32 distances are within rounding of each other the
                // choice is re-made on FP64 distances (ties by id)
                const uint64_t ka = wave_read64(sk, K - 1), kb = wave_read64(sk, K);
                if (kb != NNM_EMPTY_KEY && nnm_key_dist(kb) <= nnm_key_dist(ka) * a.tie_win) {
                    ++t.ties;
                    const uint32_t ia = nnm_key_id(ka), ib = nnm_key_id(kb);
                    const double ea = own_l2_f64(a.X + v * a.ld, a.X + (uint64_t)ia * a.ld, a.ld);
                    const double eb = own_l2_f64(a.X + v * a.ld, a.X + (uint64_t)ib * a.ld, a.ld);
                    if (eb < ea || (eb == ea && ib < ia)) {
                        ++t.swaps;
                        const uint32_t pa_ = wave_read32(pay, K - 1), pb_ = wave_read32(pay, K);
                        if (lane == K - 1) {
                            sk = kb;
                            pay = pb_;
                        }
                        if (lane == K) {
                            sk = ka;
                            pay = pa_;
                        }
                    }
                }
            }
            if (!(__ballot(lane < K && sk != pkey))) continue;  // no candidate entered
            changed = true;
            if (lane < K) {
                pid = nnm_key_id(sk);
                pd = nnm_key_dist(sk);
                pf = pay & 1u;
                fresh = (pay >> 1) & 1u;
            }
        }
    }
    if (changed) {
        if (lane < K) {
            a.g_ids[v * K + lane] = pid;
            a.g_dist[v * K + lane] = pd;
            a.g_flag[v * K + lane] = (uint8_t)(fresh ? 2u : pf);
        }
        // wrap-around sums: the changes are added modulo 2^64
        t.upd += (uint64_t)__popcll(__ballot(lane < K && fresh)) - e0;
        t.flags += (uint64_t)__popcll(__ballot(lane < K && pf)) - f0;
    }
}

// One workgroup per bin of 32 targets (persistent over the bins): the bin's
// records in chunks of OWN_CH, each counting-sorted by target in LDS, then
// one wave per target with records merges its run.  Targets without records
// are not visited (the list build counted their new flags).
template <bool RR, bool TIE, bool IDD>
__global__ __launch_bounds__(256) void k_own(OwnArgs a) {
    __shared__ uint64_t skey[OWN_CH];
    __shared__ uint32_t hcnt[OWN_T];
    __shared__ uint32_t hoff[OWN_T];
    __shared__ uint32_t perm[4][NNM_WAVE];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    OwnTally t;
    for (uint32_t bin = blockIdx.x; bin < a.nbins; bin += gridDim.x) {
        const uint32_t s = bin ? a.bin_end[(uint64_t)(bin - 1) * a.bin_stride] : 0u;
        const uint32_t e = a.bin_end[(uint64_t)bin * a.bin_stride];
        const uint64_t t0 = (uint64_t)bin << NNM_OWN_BITS;
        uint32_t c0 = s;
        for (;;) {
            const uint32_t ce = min(e, c0 + OWN_CH), m = ce - c0;
            const bool last_chunk = ce >= e;
            // the chunk's records, all loads in flight before the first wait
            // (positions past m repeat the last record and are not counted)
            uint64_t key[OWN_CH / 256];
            uint32_t tl[OWN_CH / 256], rk[OWN_CH / 256];
            if (tid < OWN_T) hcnt[tid] = 0;
            own_chunk_load<OWN_CH / 256, 256>(a, c0, m, key, tl);
            __syncthreads();
#pragma unroll
            for (uint32_t k = 0; k < OWN_CH / 256; ++k) {
                // the partition puts only this bin's targets here
                const uint64_t tg = tl[k];
                tl[k] = tid + 256 * k < m && tg >= t0 && tg < t0 + OWN_T ? (uint32_t)(tg - t0) : OWN_T;
                if (tl[k] == OWN_T) key[k] = NNM_EMPTY_KEY;
                rk[k] = tl[k] < OWN_T ? atomicAdd(&hcnt[tl[k]], 1u) : 0u;
            }
            __syncthreads();
            if (wave == 0) {
                // exclusive scan of the bin's counts
                const uint32_t c = lane < OWN_T ? hcnt[lane] : 0u;
                uint32_t x = c;
#pragma unroll
                for (int o = 1; o < NNM_WAVE; o <<= 1) {
                    const uint32_t y = (uint32_t)__shfl_up((int)x, (unsigned)o, NNM_WAVE);
                    if (lane >= (uint32_t)o) x += y;
                }
                if (lane < OWN_T) hoff[lane] = x - c;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t k = 0; k < OWN_CH / 256; ++k)
                if (tl[k] < OWN_T) skey[hoff[tl[k]] + rk[k]] = key[k];
            __syncthreads();
            // the wave's targets q = wave, wave + 4, ...; the next one's list
            // is loaded while the current one merges
            auto need = [&](uint32_t q) { return t0 + q < a.n && hcnt[q] != 0; };
            uint32_t q = wave;
            while (q < OWN_T && !need(q)) q += 4;
            OwnList cur = q < OWN_T ? own_load(a, t0 + q) : OwnList();
            while (q < OWN_T) {
                uint32_t qn = q + 4;
                while (qn < OWN_T && !need(qn)) qn += 4;
                const OwnList nxt = qn < OWN_T ? own_load(a, t0 + qn) : OwnList();
                own_merge<RR, TIE, IDD>(a, t0 + q, cur, skey + hoff[q], hcnt[q], perm[wave], t);
                cur = nxt;
                q = qn;
            }
            __syncthreads();  // the chunk's LDS is reused
            if (last_chunk) break;
            c0 = ce;
        }
    }
    if (lane == 0) {
        if (t.upd) atomicAdd((unsigned long long*)&a.counters[CNT_UPDATES], (unsigned long long)t.upd);
        if (t.flags) atomicAdd((unsigned long long*)&a.counters[CNT_FLAGS], (unsigned long long)t.flags);
        if (t.rr) atomicAdd((unsigned long long*)&a.counters[CNT_RERANKED], (unsigned long long)t.rr);
        if (t.ties) atomicAdd((unsigned long long*)&a.counters[CNT_NEAR_TIES], (unsigned long long)t.ties);
        if (t.swaps) atomicAdd((unsigned long long*)&a.counters[CNT_TIE_SWAPS], (unsigned long long)t.swaps);
    }
}

// ------------------------------------------------------------ paired owners
//
// Lists of at most 32 entries with exact keys (no re-rank, no FP64 tie
// test, one key per pair): each wave merges two targets at once, one per
// 32-lane half.  A half keeps its target's list in LDS (keys and payloads,
// two buffers, empty keys past K), so the k-th key, the de-dup by id and
// the merge read it by broadcast.  Per group of 32 candidates: filter by
// the k-th key, drop ids the list holds, compact, sort (a network sized to
// the survivors), drop repeats of one pair (equal keys, adjacent), then a
// rank merge: every list entry and every candidate finds its rank in the
// other sorted run by binary search, and the entries of rank below K form
// the new list.  The result is UpdateKnnList's batch rule
// (src/pvetex.h:40-85), as in own_merge: the K smallest distinct keys of the
// list and the candidates.
#define OWN2_WAVES 8u

struct Own2Wave {
    uint64_t lkey[2][NNM_WAVE];   // list keys: lane 32 h + j holds entry j of half h
    uint32_t lpay[2][NNM_WAVE];   // payload: new flag | entered in this delivery << 1
    uint64_t cand[NNM_WAVE];      // candidates of each half (compacted, then sorted)
};

// Number of keys below `key` in a[0..32) (sorted ascending).
__device__ inline uint32_t own2_rank32(const uint64_t* a, uint64_t key) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t s = 16; s > 0; s >>= 1) pos += a[pos + s - 1] < key ? s : 0u;
    return pos + (a[pos] < key ? 1u : 0u);  // 31 below it after the steps: a[31] decides
}

// Bitonic sort of lanes [0, 2^lg) of each half ascending (lg <= 5); the
// lanes above hold the empty key.  The last stage of a 32-lane sort runs
// with block size 64 so that both halves sort upward.
__device__ inline uint64_t own2_sort(uint64_t key, uint32_t lg) {
    if (lg >= 1) key = bitonic_pass64<2, 1>(key);
    if (lg >= 2) key = bitonic_pass64<4, 2>(key);
    if (lg >= 3) key = bitonic_pass64<8, 4>(key);
    if (lg >= 4) key = bitonic_pass64<16, 8>(key);
    if (lg >= 5) key = bitonic_pass64<64, 16>(key);
    return key;
}

// Keys of the lanes with keep set moved to the first lanes of their half in
// lane order (the empty key above them); *nh: the half's count.
__device__ inline uint64_t own2_compact(uint64_t key, bool keep, uint64_t* buf, uint32_t* nh) {
    const uint32_t lane = nnm_lane(), hl = lane & 31u, half = lane >> 5;
    const uint64_t m = __ballot(keep);
    const uint32_t nlo = (uint32_t)__popc((uint32_t)m), nhi = (uint32_t)__popc((uint32_t)(m >> 32));
    *nh = half ? nhi : nlo;
    const uint32_t r = nnm_mbcnt(m) - (half ? nlo : 0u);
    own_wave_sync();  // earlier readers of buf are done
    if (keep) buf[half * 32u + r] = key;
    own_wave_sync();
    return hl < *nh ? buf[lane] : NNM_EMPTY_KEY;
}

struct Own2Target {
    uint64_t v = 0;      // target row
    bool has = false;    // the half holds a target
    uint32_t off = 0;    // its candidates: skey[off, off + len)
    uint32_t len = 0;
};

// Merge the candidates of two targets (one per half) into their lists.
__device__ inline void own2_merge(const OwnArgs& a, const Own2Target& tg, uint64_t lk0, uint32_t lp0, const uint64_t* skey,
                                  Own2Wave& W, OwnTally& t) {
    const uint32_t lane = nnm_lane(), hl = lane & 31u, half = lane >> 5;
    const uint32_t K = a.K;
    const uint64_t hmask = half ? 0xFFFFFFFF00000000ull : 0x00000000FFFFFFFFull;
    const uint64_t lmask = (K == 32 ? 0xFFFFFFFFull : ((1ull << K) - 1)) * 0x100000001ull;  // list lanes of both halves
    own_wave_sync();  // the previous pair's readers are done
    W.lkey[0][lane] = lk0;
    W.lpay[0][lane] = lp0;
    W.lkey[1][lane] = NNM_EMPTY_KEY;  // entries past K stay empty in both buffers
    own_wave_sync();
    const uint64_t f0 = __ballot(hl < K && (lp0 & 1u)), e0 = __ballot(hl < K && (lp0 & 2u));
    const uint32_t glen = max((uint32_t)__builtin_amdgcn_readlane((int)tg.len, 0),
                              (uint32_t)__builtin_amdgcn_readlane((int)tg.len, 32));
    uint32_t buf = 0;
    uint64_t entered = 0;  // lanes of the halves where a candidate entered
    for (uint32_t g0 = 0; g0 < glen; g0 += 32) {
        uint64_t ck = tg.has && g0 + hl < tg.len ? skey[tg.off + g0 + hl] : NNM_EMPTY_KEY;
        const uint64_t* lkb = &W.lkey[buf][half * 32u];
        bool valid = ck < lkb[K - 1];  // the empty key is above every key
        if (!__ballot(valid)) continue;
        // de-dup by id against the list (two entries per 16-byte read)
        const uint32_t cid = nnm_key_id(ck);
        bool dup = false;
        for (uint32_t p = 0; p < K; p += 2) {
            const uint4 q = *reinterpret_cast<const uint4*>(lkb + p);
            dup = dup | (cid == q.x) | (cid == q.z);  // an empty entry (id ~0) matches no id
        }
        valid = valid & !dup;
        if (!__ballot(valid)) continue;
        uint32_t nh;
        ck = own2_compact(ck, valid, W.cand, &nh);
        const uint32_t nmax = max((uint32_t)__builtin_amdgcn_readlane((int)nh, 0),
                                  (uint32_t)__builtin_amdgcn_readlane((int)nh, 32));
        ck = own2_sort(ck, nmax > 1 ? 32u - (uint32_t)__builtin_clz(nmax - 1) : 0u);
        {
            // repeats of one pair are adjacent: keep the first
            const uint64_t prev = shfl64(ck, hl ? lane - 1 : lane);
            const bool keep = ck != NNM_EMPTY_KEY && (hl == 0 || prev != ck);
            if (__ballot(keep) != __ballot(ck != NNM_EMPTY_KEY)) ck = own2_compact(ck, keep, W.cand, &nh);
        }
        // rank merge: candidate i (lane i < nh) and list entry j (lane j < K)
        // go to rank i + (list entries below) and j + (candidates below)
        own_wave_sync();
        W.cand[lane] = ck;  // sorted, empty past nh
        own_wave_sync();
        const uint64_t lk = W.lkey[buf][lane];
        const uint32_t lp = W.lpay[buf][lane];
        const uint32_t below_c = own2_rank32(&W.cand[half * 32u], lk);
        const uint32_t below_l = own2_rank32(lkb, ck);
        const uint32_t nb = buf ^ 1u;
        if (hl < K && hl + below_c < K) {
            W.lkey[nb][half * 32u + hl + below_c] = lk;
            W.lpay[nb][half * 32u + hl + below_c] = lp;
        }
        const bool in = hl < nh && hl + below_l < K;
        if (in) {
            W.lkey[nb][half * 32u + hl + below_l] = ck;
            W.lpay[nb][half * 32u + hl + below_l] = 3u;  // enters new and entered now
        }
        entered |= __ballot(in);
        buf = nb;
        own_wave_sync();
    }
    // halves where a candidate entered write their lists back
    const bool chg = (entered & hmask) != 0;
    const uint64_t nk = W.lkey[buf][lane];
    const uint32_t np = W.lpay[buf][lane];
    if (tg.has && chg && hl < K) {
        a.g_ids[tg.v * K + hl] = nnm_key_id(nk);
        a.g_dist[tg.v * K + hl] = nnm_key_dist(nk);
        a.g_flag[tg.v * K + hl] = (uint8_t)((np & 2u) ? 2u : (np & 1u));
    }
    // tallies: each changed half's change of entered entries and new flags
    // (wrap-around sums, modulo 2^64)
    const uint64_t f1 = __ballot(hl < K && (np & 1u)), e1 = __ballot(hl < K && (np & 2u));
#pragma unroll
    for (uint32_t h = 0; h < 2; ++h) {
        const uint64_t hm = h ? 0xFFFFFFFF00000000ull : 0x00000000FFFFFFFFull;
        if (entered & hm) {
            t.upd += (uint64_t)__popcll(e1 & hm & lmask) - (uint64_t)__popcll(e0 & hm & lmask);
            t.flags += (uint64_t)__popcll(f1 & hm & lmask) - (uint64_t)__popcll(f0 & hm & lmask);
        }
    }
}

// A half's list entry (lane j < K of the half) as key and payload.
__device__ inline void own2_load(const OwnArgs& a, const Own2Target& tg, uint64_t* key, uint32_t* pay) {
    const uint32_t hl = nnm_lane() & 31u;
    *key = NNM_EMPTY_KEY;
    *pay = 0;
    if (tg.has && hl < a.K) {
        const uint64_t i = tg.v * a.K + hl;
        const uint32_t f = a.g_flag[i];
        *key = nnm_key(a.g_dist[i], a.g_ids[i]);
        *pay = (f ? 1u : 0u) | (f == 2 ? 2u : 0u);
    }
}

// Owner bins as in k_own (counting sort of each chunk by target in LDS);
// the bin's targets with records are taken two at a time by the waves.
__global__ __launch_bounds__(OWN2_WAVES * NNM_WAVE, 6) void k_own2(OwnArgs a) {
    constexpr uint32_t NT = OWN2_WAVES * NNM_WAVE;
    constexpr uint32_t PER = OWN_CH / NT;
    __shared__ uint64_t skey[OWN_CH];
    __shared__ uint32_t hcnt[OWN_T];
    __shared__ uint32_t hoff[OWN_T];
    __shared__ __align__(16) Own2Wave W[OWN2_WAVES];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6, half = lane >> 5;
    OwnTally t;
    for (uint32_t bin = blockIdx.x; bin < a.nbins; bin += gridDim.x) {
        const uint32_t s = bin ? a.bin_end[(uint64_t)(bin - 1) * a.bin_stride] : 0u;
        const uint32_t e = a.bin_end[(uint64_t)bin * a.bin_stride];
        const uint64_t t0 = (uint64_t)bin << NNM_OWN_BITS;
        uint32_t c0 = s;
        for (;;) {
            const uint32_t ce = min(e, c0 + OWN_CH), m = ce - c0;
            const bool last_chunk = ce >= e;
            uint64_t key[PER];
            uint32_t tl[PER], rk[PER];
            if (tid < OWN_T) hcnt[tid] = 0;
            own_chunk_load<PER, NT>(a, c0, m, key, tl);
            __syncthreads();
#pragma unroll
            for (uint32_t k = 0; k < PER; ++k) {
                const uint64_t tg = tl[k];
                tl[k] = tid + NT * k < m && tg >= t0 && tg < t0 + OWN_T ? (uint32_t)(tg - t0) : OWN_T;
                if (tl[k] == OWN_T) key[k] = NNM_EMPTY_KEY;
                rk[k] = tl[k] < OWN_T ? atomicAdd(&hcnt[tl[k]], 1u) : 0u;
            }
            __syncthreads();
            if (wave == 0) {
                const uint32_t c = lane < OWN_T ? hcnt[lane] : 0u;
                uint32_t x = c;
#pragma unroll
                for (int o = 1; o < NNM_WAVE; o <<= 1) {
                    const uint32_t y = (uint32_t)__shfl_up((int)x, (unsigned)o, NNM_WAVE);
                    if (lane >= (uint32_t)o) x += y;
                }
                if (lane < OWN_T) hoff[lane] = x - c;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t k = 0; k < PER; ++k)
                if (tl[k] < OWN_T) skey[hoff[tl[k]] + rk[k]] = key[k];
            __syncthreads();
            // targets with records: pairs (2p, 2p + 1) of the set bits, pair p
            // to wave p % OWN2_WAVES; the next pair's lists load while the
            // current pair merges
            const uint64_t need = __ballot(lane < OWN_T && t0 + lane < a.n && hcnt[lane] != 0);
            const uint32_t nneed = (uint32_t)__popcll(need);
            auto target = [&](uint32_t p) {
                Own2Target tg;
                const uint32_t i = 2 * p + half;
                if (i < nneed) {
                    const uint32_t q = nth_set_bit(need, i);
                    tg.v = t0 + q;
                    tg.has = true;
                    tg.off = hoff[q];
                    tg.len = hcnt[q];
                }
                return tg;
            };
            uint32_t p = wave;
            Own2Target cur = target(p);
            uint64_t ck;
            uint32_t cp;
            own2_load(a, cur, &ck, &cp);
            while (2 * p < nneed) {
                const uint32_t pn = p + OWN2_WAVES;
                const Own2Target nxt = target(pn);
                uint64_t nk;
                uint32_t np;
                own2_load(a, nxt, &nk, &np);
                own2_merge(a, cur, ck, cp, skey, W[wave], t);
                cur = nxt;
                ck = nk;
                cp = np;
                p = pn;
            }
            __syncthreads();  // the chunk's LDS is reused
            if (last_chunk) break;
            c0 = ce;
        }
    }
    if (lane == 0) {
        if (t.upd) atomicAdd((unsigned long long*)&a.counters[CNT_UPDATES], (unsigned long long)t.upd);
        if (t.flags) atomicAdd((unsigned long long*)&a.counters[CNT_FLAGS], (unsigned long long)t.flags);
    }
}

// ------------------------------------------------------------------ host

namespace {

struct PassPlan {
    uint32_t npass = 1;
    uint32_t shift[3] = {0, 0, 0};
    uint32_t bins[3] = {0, 0, 0};
};

uint32_t ceil_log2(uint64_t x) {
    uint32_t b = 0;
    while ((1ull << b) < x) ++b;
    return b;
}

// Digits of every pass: the last pass's bins are the owner bins (target >>
// NNM_OWN_BITS); the bits above them are split into passes of at most
// PT_MAXBITS bits, so a tile's digits fit the LDS histogram.
PassPlan plan_passes(uint64_t n) {
    PassPlan p;
    const uint32_t bits = ceil_log2(std::max<uint64_t>(n, 2));
    const uint32_t r = bits > NNM_OWN_BITS ? bits - NNM_OWN_BITS : 0u;
    p.npass = std::max<uint32_t>(1, (r + PT_MAXBITS - 1) / PT_MAXBITS);
    const uint32_t per = p.npass ? (r + p.npass - 1) / p.npass : 0u;
    for (uint32_t k = 0; k < p.npass; ++k) {
        p.shift[k] = NNM_OWN_BITS + per * (p.npass - 1 - k);
        p.bins[k] = (uint32_t)(((n - 1) >> p.shift[k]) + 1);
    }
    return p;
}

int fill32(nnm_ctx* c, uint32_t* p, uint64_t n, uint32_t v) { return nnm_fill_u32(c, p, n, v); }

}  // namespace

static uint32_t pass_grid(nnm_ctx* c) { return (uint32_t)c->num_cus; }  // one tile workgroup per CU (LDS)
static uint64_t pass0_slots(nnm_ctx* c) { return (uint64_t)PT_BINS * pass_grid(c) + 1; }
// second-pass table entries for n rows: (first-pass bins x 2^b1) owner bins x segments
static uint64_t pass1_slots(uint64_t n) { return (((n >> NNM_OWN_BITS) + 2 + PT_BINS) * P1_SEGS) + 1; }

void nnm_offer_free(nnm_ctx* c) {
    void** ps[] = {(void**)&c->okey[0], (void**)&c->okey[1], (void**)&c->otgt[0], (void**)&c->otgt[1],
                   (void**)&c->ochunk_used, (void**)&c->opart_cnt,
                   (void**)&c->opart_cur[0], (void**)&c->opart_cur[1], (void**)&c->opart_cur[2],
                   (void**)&c->work[0], (void**)&c->work[1], (void**)&c->opart_blk, (void**)&c->opart_boff,
                   (void**)&c->opart_tmp, (void**)&c->opart_p1[0], (void**)&c->opart_p1[1]};
    for (void** p : ps) {
        nnm_free(c, *p);
        *p = nullptr;
    }
    c->ochunks = 0;
    c->ocap_row = 0;
}

int nnm_offer_buffers(nnm_ctx* c, const nnm_params* p) {
    const uint32_t cap_row = p->offer_cap ? p->offer_cap : 0u;
    if (c->okey[0] && c->ocap_row == cap_row) return NNM_OK;
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    nnm_offer_free(c);
    const uint64_t n = c->n;
    // the default: a chunk per NNM_OCH records at 256 records per row, plus room
    // for the partly filled last chunk of every join wave (at most 32 waves
    // per CU over the three join kernels); an explicit offer_cap sizes the
    // log to exactly offer_cap records per row
    const uint64_t per_row = cap_row ? cap_row : NNM_OFFER_CAP_DEFAULT;
    uint64_t chunks = std::max<uint64_t>(1, (n * per_row + NNM_OCH - 1) / NNM_OCH);
    if (!cap_row) chunks += (uint64_t)c->num_cus * 32;
    chunks = std::min<uint64_t>(chunks, (0xFFFFFFFFull / NNM_OCH) - 1);
    c->ochunks = (uint32_t)chunks;
    c->ocap_row = cap_row;  // 0: the default sizing
    const size_t recs = (size_t)chunks * NNM_OCH;
    for (int k = 0; k < 2; ++k) {
        NNM_TRY(nnm_alloc(c, (void**)&c->okey[k], recs * sizeof(uint64_t)));
        NNM_TRY(nnm_alloc(c, (void**)&c->otgt[k], recs * sizeof(uint32_t)));
    }
    NNM_TRY(nnm_alloc(c, (void**)&c->ochunk_used, (size_t)chunks * sizeof(uint32_t)));
    // bins of any pass over any active prefix: at most (n >> NNM_OWN_BITS) + 1
    const uint64_t maxbins = (n >> NNM_OWN_BITS) + 2;
    NNM_TRY(nnm_alloc(c, (void**)&c->opart_cnt, maxbins * sizeof(uint32_t)));
    for (int k = 0; k < 3; ++k) NNM_TRY(nnm_alloc(c, (void**)&c->opart_cur[k], (maxbins + 1) * sizeof(uint32_t)));
    NNM_TRY(nnm_alloc(c, (void**)&c->work[0], n * sizeof(uint32_t)));
    NNM_TRY(nnm_alloc(c, (void**)&c->work[1], n * sizeof(uint32_t)));
    // first pass: per-block counts and their scan; scan scratch for every pass
    NNM_TRY(nnm_alloc(c, (void**)&c->opart_blk, pass0_slots(c) * sizeof(uint32_t)));
    NNM_TRY(nnm_alloc(c, (void**)&c->opart_boff, (pass0_slots(c) + 1) * sizeof(uint32_t)));
    NNM_TRY(nnm_alloc(c, (void**)&c->opart_p1[0], pass1_slots(n) * sizeof(uint32_t)));
    NNM_TRY(nnm_alloc(c, (void**)&c->opart_p1[1], (pass1_slots(n) + 1) * sizeof(uint32_t)));
    const uint64_t scan_m = std::max<uint64_t>(std::max<uint64_t>(pass0_slots(c), maxbins), pass1_slots(n));
    NNM_TRY(nnm_alloc(c, (void**)&c->opart_tmp, ((scan_m + 2047) / 2048 + 2) * sizeof(uint32_t)));
    return NNM_OK;
}

OfferSink nnm_offer_sink(nnm_ctx* c, uint32_t round) {
    OfferSink s;
    s.key = c->okey[0];
    s.tgt = c->otgt[0];
    s.chunk_used = c->ochunk_used;
    s.counters = c->counters;
    s.redo = c->work[round & 1];
    s.nchunks = c->ochunks;
    return s;
}

int nnm_offer_reset(nnm_ctx* c) { return nnm_fill_u64(c, c->counters + CNT_OCHUNKS, 2, 0); }


int nnm_offer_deliver(nnm_ctx* c, const nnm_params* p) {
    PassPlan pl = plan_passes(c->n_act);
    // owner-bin digits of 11..20 bits: a first pass of the high half, then a
    // second pass over each first-pass bin's 2^b1 owner bins in segments
    // (k_part1_count, k_part1_scatter)
    const uint32_t lbits = ceil_log2(std::max<uint64_t>(c->n_act, 2));
    const uint32_t obits = lbits > NNM_OWN_BITS ? lbits - NNM_OWN_BITS : 0u;
    const bool local = obits > PT_MAXBITS && obits <= 2 * PT_MAXBITS;
    uint32_t b1 = 0;
    if (local) {
        b1 = (obits + 1) / 2;
        pl.npass = 1;
        pl.shift[0] = NNM_OWN_BITS + b1;
        pl.bins[0] = (uint32_t)(((c->n_act - 1) >> pl.shift[0]) + 1);
    }
    PassIn in;
    in.key = c->okey[0];
    in.tgt = c->otgt[0];
    in.chunk_used = c->ochunk_used;
    in.counters = c->counters;
    in.nchunks = c->ochunks;
    in.total = nullptr;
    const uint32_t G = pass_grid(c);
    const uint32_t* bin_end = nullptr;
    uint32_t bin_stride = 1;
    for (uint32_t k = 0; k < pl.npass; ++k) {
        const int o = (k + 1) & 1;
        const uint32_t bins = pl.bins[k];
        in.nbins = bins;
        if (k == 0) {
            // per-block counts, one scan, no cursor atomics
            hipLaunchKernelGGL(k_part0_count, dim3(G), dim3(PT_THREADS), 0, c->stream, in, pl.shift[k], c->opart_blk);
            NNM_LAUNCH_CHECK(c);
            NNM_TRY(nnm_scan_u32(c, c->opart_blk, (uint64_t)bins * G, c->opart_boff, c->opart_tmp));
            hipLaunchKernelGGL(k_part0_scatter, dim3(G), dim3(PT_THREADS), 0, c->stream, in, pl.shift[k], c->opart_boff,
                               c->okey[o], c->otgt[o]);
            NNM_LAUNCH_CHECK(c);
            in.total = c->opart_boff + (uint64_t)bins * G;  // the scan's total
            bin_end = c->opart_boff + G;  // bin r ends where bin r + 1 starts
            bin_stride = G;
        } else {
            NNM_TRY(fill32(c, c->opart_cnt, bins, 0));
            hipLaunchKernelGGL(k_part_count, dim3(G), dim3(PT_THREADS), 0, c->stream, in, pl.shift[k], c->opart_cnt);
            NNM_LAUNCH_CHECK(c);
            NNM_TRY(nnm_scan_u32(c, c->opart_cnt, bins, c->opart_cur[k], c->opart_tmp));
            hipLaunchKernelGGL(k_part_scatter, dim3(G), dim3(PT_THREADS), 0, c->stream, in, pl.shift[k],
                               c->opart_cur[k], c->okey[o], c->otgt[o]);
            NNM_LAUNCH_CHECK(c);
            in.total = c->opart_cur[k] + bins;  // the scan's total
            bin_end = c->opart_cur[k];  // the cursors end at their bins' ends
            bin_stride = 1;
        }
        in.key = c->okey[o];
        in.tgt = c->otgt[o];
        in.chunk_used = nullptr;
    }
    uint32_t nob = pl.bins[pl.npass - 1];
    if (local) {
        // the first pass wrote okey[1] / otgt[1]; the owner bins go to okey[0] / otgt[0]
        nob = (uint32_t)(((c->n_act - 1) >> NNM_OWN_BITS) + 1);
        const uint32_t bins0 = pl.bins[0];
        const uint64_t tab = ((uint64_t)bins0 << b1) * P1_SEGS;
        const uint32_t units = bins0 * P1_SEGS;
        const uint32_t grid = std::min<uint32_t>(units, (uint32_t)c->num_cus * 4);
        hipLaunchKernelGGL(k_part1_count, dim3(grid), dim3(PT_THREADS), 0, c->stream, c->otgt[1], c->opart_boff, G,
                           bins0, b1, c->opart_p1[0]);
        NNM_LAUNCH_CHECK(c);
        NNM_TRY(nnm_scan_u32(c, c->opart_p1[0], tab, c->opart_p1[1], c->opart_tmp));
        hipLaunchKernelGGL(k_part1_scatter, dim3(grid), dim3(PT_THREADS), 0, c->stream, c->okey[1], c->otgt[1],
                           c->opart_boff, G, bins0, b1, c->opart_p1[1], c->okey[0], c->otgt[0]);
        NNM_LAUNCH_CHECK(c);
        in.key = c->okey[0];
        in.tgt = c->otgt[0];
        bin_end = c->opart_p1[1] + P1_SEGS;  // owner bin ob ends where entry (ob + 1) x P1_SEGS starts
        bin_stride = P1_SEGS;
    }
    OwnArgs a;
    a.n = c->n_act;
    a.K = c->K;
    a.X = c->X;
    a.Xb_exact = (c->data_exact_bf16 && c->Xb) ? c->Xb : nullptr;
    a.ld = c->ld;
    a.rerank_tau = 1.f;
    a.tie_win = nnm_near_tie_window(c, p);
    a.g_ids = c->g_ids;
    a.g_dist = c->g_dist;
    a.g_flag = c->g_flag;
    a.key = in.key;
    a.tgt = in.tgt;
    a.bin_end = bin_end;
    a.bin_stride = bin_stride;
    a.nbins = nob;
    a.counters = c->counters;
    int rerank = 0;
    if (p->prec == NNM_PREC_BF16)
        rerank = (p->rerank == NNM_RERANK_ALWAYS) || (p->rerank == NNM_RERANK_AUTO && !c->data_exact_bf16);
    // learned phase: lists and offers are both estimates of one metric, so the
    // owners merge them as they are (the exact distances come at the switch)
    if (p->rerank == NNM_RERANK_NEVER || p->prec == NNM_PREC_LEARNED) rerank = 0;
    const bool tie = a.tie_win > 1.f;
    // lower-bound keys kept as they are next to exact keys of the FP32
    // long-list join: one id can carry two keys
    const bool idd = !rerank && p->prec == NNM_PREC_BF16 && !c->data_exact_bf16 &&
                     c->cap_new + c->cap_old > nnm_join_learned_max_list();
    void (*kern)(OwnArgs) = rerank ? (tie ? k_own<true, true, false> : k_own<true, false, false>)
                            : idd  ? (tie ? k_own<false, true, true> : k_own<false, false, true>)
                                   : (tie ? k_own<false, true, false> : k_own<false, false, false>);
    // exact keys and lists of at most 32 entries: two targets per wave
    const bool paired = !rerank && !tie && !idd && a.K <= 32;
    if (paired) kern = k_own2;
    const uint32_t threads = paired ? OWN2_WAVES * NNM_WAVE : 256u;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, 0) != hipSuccess || per_cu < 1) per_cu = 2;
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(a.nbins, (uint32_t)c->num_cus * per_cu));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), 0, c->stream, a);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

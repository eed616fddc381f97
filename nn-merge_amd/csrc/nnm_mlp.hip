// Learned candidate pruning: a small MLP projection f(x) = W2 relu(W1 a x + b1) + b2
// (d_in -> hidden -> d_out, FP32 master weights, BF16 MFMA or FP32 GEMMs),
// trained so that |f(x) - f(y)|^2 approximates a^2 |x - y|^2 on pairs that
// NN-Descent itself produces (list neighbours and two-hop candidates).  The
// projected rows (32-d bf16) feed the BF16 join kernels during the learned
// phase, whose lists hold the estimates kscale |f(x) - f(y)|^2; afterwards
// every list distance is recomputed exactly once and the plain BF16 join
// finishes the merge.  The reference has no learned component; the pair loss,
// gradients and optimizer are restated in FP64 by the CPU oracle for the
// tests.
//
// Loss over a batch of B pairs (a_i, b_i) with exact squared distance t_i and
// anchor k-th distance s_i (both scaled by a^2):
//   p_i = |f(x_a) - f(x_b)|^2,  r_i = (p_i - t_i) / (t_i + s_i),  L = mean r_i^2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include "nnm_internal.hpp"

typedef short mlp_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mlp_f32x4 __attribute__((ext_vector_type(4)));

struct nnm_mlp_state {
    uint32_t din = 0, hid = 0, dout = 0;
    float alpha = 1.f;  // input scale
    float tau = 1.f;    // keep-quantile of p / kth on true improvements (diagnostic)
    int trained = 0;
    // master weights and Adam moments (FP32), gradients
    float *W1 = nullptr, *b1 = nullptr, *W2 = nullptr, *b2 = nullptr;
    float *m = nullptr, *v = nullptr, *g = nullptr;  // packed [W1 b1 W2 b2]
    uint64_t nparam = 0;
    uint64_t step = 0;
    // batch scratch
    uint32_t cap_rows = 0;
    float *xs = nullptr, *pre = nullptr, *h = nullptr, *z = nullptr, *dz = nullptr, *dh = nullptr;
    float *t = nullptr, *s = nullptr, *loss = nullptr;
    uint32_t* pairs = nullptr;
    // gradient parts of the fused step: one nparam slice per block of 32 pairs
    float* part = nullptr;
    uint64_t part_cap = 0;
    // bf16 (RNE) copies of W1 and W2 for the projection, then the projection
    // of every row (bf16) and its squared norms
    uint16_t* Wb = nullptr;
    uint16_t* P = nullptr;
    float* pnorm = nullptr;
    uint64_t projected_n = 0;
    uint64_t proj_gen = 0;  // number of the current projection (unique in the process; 0: none)
    // FP8 copy of the projection (OCP e4m3, scaled by s8 = 448 / max |P|),
    // the squared norms of the encoded rows, and the projection it encodes
    uint8_t* P8 = nullptr;
    float* pnorm8 = nullptr;
    uint32_t* amax = nullptr;  // max |P| bits (one word)
    float s8 = 1.f;
    uint64_t fp8_rows = 0;  // rows P8 holds room for
    uint64_t fp8_gen = 0;   // proj_gen of the projection P8 encodes (0: none)
};

static nnm_mlp_state* mlp_of(nnm_ctx* c) { return reinterpret_cast<nnm_mlp_state*>(c->mlp); }

__device__ inline short mlp_to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// C[M x N] = op(A)[M x K] * op(B)[K x N] (+ bias[n]) with A(m,k) = A[m*am + k*ak],
// B(k,n) = B[k*bk + n*bn]; M, N, K multiples of 16.  One wave per 16x16 tile.
// BF16: operands rounded to bf16, MFMA with FP32 accumulation; otherwise FP32
// fused multiply-add in k order.
template <bool BF16>
__global__ void k_gemm(uint32_t M, uint32_t N, uint32_t K, const float* __restrict__ A, uint32_t am, uint32_t ak,
                       const float* __restrict__ B, uint32_t bk, uint32_t bn, const float* __restrict__ bias,
                       float* __restrict__ C, uint32_t cm) {
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) / NNM_WAVE;
    const uint32_t lane = threadIdx.x % NNM_WAVE;
    const uint32_t tiles_n = N / 16;
    if (wave >= (M / 16) * tiles_n) return;
    const uint32_t m0 = (wave / tiles_n) * 16, n0 = (wave % tiles_n) * 16;
    const uint32_t lr = lane & 15u, lg = lane >> 4;
    if (BF16) {
        mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (uint32_t k0 = 0; k0 < K; k0 += 16) {
            mlp_bf16x4 a, b;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t k = k0 + 4 * lg + q;
                a[q] = mlp_to_bf16(A[(uint64_t)(m0 + lr) * am + (uint64_t)k * ak]);
                b[q] = mlp_to_bf16(B[(uint64_t)k * bk + (uint64_t)(n0 + lr) * bn]);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t mm = m0 + 4 * lg + q, nn = n0 + lr;
            C[(uint64_t)mm * cm + nn] = acc[q] + (bias ? bias[nn] : 0.f);
        }
    } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t mm = m0 + 4 * lg + q, nn = n0 + lr;
            float acc = 0.f;
            for (uint32_t k = 0; k < K; ++k)
                acc = __builtin_fmaf(A[(uint64_t)mm * am + (uint64_t)k * ak], B[(uint64_t)k * bk + (uint64_t)nn * bn], acc);
            C[(uint64_t)mm * cm + nn] = acc + (bias ? bias[nn] : 0.f);
        }
    }
}

__global__ void k_relu(const float* __restrict__ pre, float* __restrict__ h, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) h[i] = fmaxf(pre[i], 0.f);
}

// Pair loss and dL/dz for B pairs stored as rows (2i, 2i+1) of z.
__global__ void k_pair_loss(uint32_t B, uint32_t dout, const float* __restrict__ z, const float* __restrict__ t,
                            const float* __restrict__ s, float* __restrict__ dz, float* __restrict__ loss) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const float* za = z + (uint64_t)(2 * i) * dout;
    const float* zb = za + dout;
    float p = 0.f;
    for (uint32_t j = 0; j < dout; ++j) {
        const float e = za[j] - zb[j];
        p = __builtin_fmaf(e, e, p);
    }
    const float den = t[i] + s[i];
    const float r = (p - t[i]) / den;
    loss[i] = r * r;
    const float g = 2.f * r / den / (float)B;  // dL/dp
    float* da = dz + (uint64_t)(2 * i) * dout;
    float* db = da + dout;
    for (uint32_t j = 0; j < dout; ++j) {
        const float e = 2.f * g * (za[j] - zb[j]);
        da[j] = e;
        db[j] = -e;
    }
}

__global__ void k_relu_grad(const float* __restrict__ pre, float* __restrict__ dh, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && pre[i] <= 0.f) dh[i] = 0.f;
}

// Column sums: out[j] = sum_r in[r * cols + j].  Block (64, 16): 64 columns,
// 16 row phases summed in a fixed order (deterministic).
__global__ __launch_bounds__(1024) void k_colsum(const float* __restrict__ in, uint32_t rows, uint32_t cols,
                                                 float* __restrict__ out) {
    __shared__ float part[16][64];
    const uint32_t tx = threadIdx.x, ty = threadIdx.y;
    const uint32_t j = blockIdx.x * 64 + tx;
    float acc = 0.f;
    if (j < cols)
        for (uint32_t r = ty; r < rows; r += 16) acc += in[(uint64_t)r * cols + j];
    part[ty][tx] = acc;
    __syncthreads();
    if (ty == 0 && j < cols) {
        float s = 0.f;
        for (int q = 0; q < 16; ++q) s += part[q][tx];
        out[j] = s;
    }
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       uint64_t n, float lr, float b1, float b2, float c1, float c2) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// Training pairs from the current lists: pair i has anchor a; even i pairs it
// with one of its list neighbours, odd i with a neighbour of a neighbour (a
// join candidate).  Rows are gathered scaled by alpha; t = alpha^2 |xa - xb|^2
// in the library's FP32 order, s = alpha^2 kth(a).
__global__ void k_mlp_pairs(uint32_t B, uint64_t n, uint32_t K, uint32_t ld, const float* __restrict__ X,
                            const uint32_t* __restrict__ ids, const float* __restrict__ g_dist,
                            const uint32_t* __restrict__ label, float alpha, uint64_t seed,
                            float* __restrict__ xs, float* __restrict__ t, float* __restrict__ s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const uint64_t r1 = nnm_rand(seed, i, 1), r2 = nnm_rand(seed, i, 2);
    // anchors from rows that carry a built list (raw rows of a merge have none yet)
    uint32_t a = (uint32_t)(nnm_rand(seed, i, 0) % n);
    for (uint32_t tr = 3; tr < 19 && label[a] == NNM_RAW; ++tr) a = (uint32_t)(nnm_rand(seed, i, tr) % n);
    uint32_t b = ids[(uint64_t)a * K + (uint32_t)(r1 % K)];
    if (i & 1u) {
        const uint32_t c = ids[(uint64_t)b * K + (uint32_t)(r2 % K)];
        if (c != a) b = c;
    }
    const float* xa = X + (uint64_t)a * ld;
    const float* xb = X + (uint64_t)b * ld;
    float* oa = xs + (uint64_t)(2 * i) * ld;
    float* ob = oa + ld;
    for (uint32_t k = 0; k < ld; ++k) {
        oa[k] = alpha * xa[k];
        ob[k] = alpha * xb[k];
    }
    t[i] = alpha * alpha * nnm_l2_rows(xa, xb, ld);
    s[i] = alpha * alpha * g_dist[(uint64_t)a * K + K - 1];
}

// ------------------------------------------------------------ fused step
//
// One training step in two launches instead of eleven.  k_mlp_step: a block
// of 256 threads takes 32 pairs (64 rows): gathers them (as k_mlp_pairs), or
// reads caller rows; forward (pre = x W1^T + b1, h = relu, z = h W2^T + b2)
// on 16x16x16 bf16 MFMA with the operands in LDS and the bf16 weights in
// global memory; the pair loss and dz; backward (dh = dz W2 masked by relu',
// then the block's gradient parts dz^T h, colsum dz, dh^T x, colsum dh), each
// part written to the block's slot of `part`.  k_mlp_reduce sums the parts
// over the blocks in a fixed order and either stores the gradient or takes an
// Adam step and repacks the bf16 weights for the next step.  Operands are
// rounded to bf16 exactly as k_gemm<true> rounds them; column sums and the
// loss are FP32, in a fixed order (deterministic).
#define MS_ROWS 64
#define MS_DMAX 256
#define MS_HMAX 256
#define MS_OMAX 64

struct MlpStepLds {
    uint16_t xb[MS_ROWS * (MS_DMAX + 4)];
    uint16_t hb[MS_ROWS * (MS_HMAX + 4)];
    uint16_t db[MS_ROWS * (MS_HMAX + 4)];
    float z[MS_ROWS * (MS_OMAX + 1)];
    float dz[MS_ROWS * (MS_OMAX + 1)];
    uint32_t row_id[MS_ROWS];
};

__device__ inline uint16_t ms_bf(float x) { return (uint16_t)mlp_to_bf16(x); }

// part: per block nparam floats [gW1 (H x D) | gb1 (H) | gW2 (O x H) | gb2 (O)]
__global__ __launch_bounds__(256) void k_mlp_step(uint32_t B, uint32_t D, uint32_t H, uint32_t O,
                                                  const uint16_t* __restrict__ W1b, const float* __restrict__ b1,
                                                  const uint16_t* __restrict__ W2b, const float* __restrict__ b2,
                                                  // pair source: gather (X != nullptr) or rows xs (2B x D)
                                                  const float* __restrict__ X, uint32_t ld, uint64_t n, uint32_t K,
                                                  const uint32_t* __restrict__ ids, const float* __restrict__ g_dist,
                                                  const uint32_t* __restrict__ label, float alpha, uint64_t seed,
                                                  const float* __restrict__ xs, float* __restrict__ t_io,
                                                  float* __restrict__ s_io, float* __restrict__ loss,
                                                  float* __restrict__ part, uint64_t nparam) {
    __shared__ __align__(16) MlpStepLds L;
    __shared__ float ts[MS_ROWS / 2], ss[MS_ROWS / 2];
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
    const uint32_t lr = lane & 15u, lg = lane >> 4;
    const uint32_t PD = D + 4, PH = H + 4, PO = O + 1;
    const uint32_t pair0 = blockIdx.x * (MS_ROWS / 2);
    // ---- pairs: ids, t, s (k_mlp_pairs' draws)
    if (tid < MS_ROWS / 2) {
        const uint32_t i = pair0 + tid;
        float t = 0.f, sv = 0.f;
        uint32_t a = 0, b = 0;
        if (i < B) {
            if (X) {
                const uint64_t r1 = nnm_rand(seed, i, 1), r2 = nnm_rand(seed, i, 2);
                a = (uint32_t)(nnm_rand(seed, i, 0) % n);
                for (uint32_t tr = 3; tr < 19 && label[a] == NNM_RAW; ++tr) a = (uint32_t)(nnm_rand(seed, i, tr) % n);
                b = ids[(uint64_t)a * K + (uint32_t)(r1 % K)];
                if (i & 1u) {
                    const uint32_t cc = ids[(uint64_t)b * K + (uint32_t)(r2 % K)];
                    if (cc != a) b = cc;
                }
                t = alpha * alpha * nnm_l2_rows(X + (uint64_t)a * ld, X + (uint64_t)b * ld, ld);
                sv = alpha * alpha * g_dist[(uint64_t)a * K + K - 1];
                t_io[i] = t;
                s_io[i] = sv;
            } else {
                t = t_io[i];
                sv = s_io[i];
            }
        }
        L.row_id[2 * tid] = a;
        L.row_id[2 * tid + 1] = b;
        ts[tid] = t;
        ss[tid] = sv;
    }
    __syncthreads();
    // ---- rows -> LDS (bf16 of alpha x, zero rows past the batch)
    for (uint32_t e = tid; e < MS_ROWS * D; e += 256) {
        const uint32_t r = e / D, k = e % D;
        const uint32_t i = pair0 + r / 2;
        float v = 0.f;
        if (i < B) v = X ? alpha * X[(uint64_t)L.row_id[r] * ld + k] : xs[(uint64_t)(2 * pair0 + r) * D + k];
        L.xb[r * PD + k] = ms_bf(v);
    }
    __syncthreads();
    // ---- pre = x W1^T + b1, h = relu(pre): wave w, column tiles w, w + 4, ...
    for (uint32_t ct = wv; ct < H / 16; ct += 4) {
        const uint32_t h0 = ct * 16;
        mlp_f32x4 acc[4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) acc[rt] = mlp_f32x4{0.f, 0.f, 0.f, 0.f};
        for (uint32_t k0 = 0; k0 < D; k0 += 16) {
            const uint2 bw = *reinterpret_cast<const uint2*>(W1b + (uint64_t)(h0 + lr) * D + k0 + 4 * lg);
            const mlp_bf16x4 bo = __builtin_bit_cast(mlp_bf16x4, bw);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) {
                const uint2 aw = *reinterpret_cast<const uint2*>(&L.xb[(rt * 16 + lr) * PD + k0 + 4 * lg]);
                acc[rt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(mlp_bf16x4, aw), bo, acc[rt],
                                                                    0, 0, 0);
            }
        }
        const float bb = b1[h0 + lr];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int q = 0; q < 4; ++q)
                L.hb[(rt * 16 + 4 * lg + q) * PH + h0 + lr] = ms_bf(fmaxf(acc[rt][q] + bb, 0.f));
    }
    __syncthreads();
    // ---- z = h W2^T + b2: wave w, row tile w, every column tile
    for (uint32_t ot = 0; ot < O / 16; ++ot) {
        const uint32_t o0 = ot * 16;
        mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (uint32_t k0 = 0; k0 < H; k0 += 16) {
            const uint2 aw = *reinterpret_cast<const uint2*>(&L.hb[(wv * 16 + lr) * PH + k0 + 4 * lg]);
            const uint2 bw = *reinterpret_cast<const uint2*>(W2b + (uint64_t)(o0 + lr) * H + k0 + 4 * lg);
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(mlp_bf16x4, aw),
                                                            __builtin_bit_cast(mlp_bf16x4, bw), acc, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) L.z[(wv * 16 + 4 * lg + q) * PO + o0 + lr] = acc[q] + b2[o0 + lr];
    }
    __syncthreads();
    // ---- pair loss and dz (k_pair_loss' arithmetic)
    if (tid < MS_ROWS / 2) {
        const uint32_t i = pair0 + tid;
        const float* za = &L.z[(2 * tid) * PO];
        const float* zb = za + PO;
        float* da = &L.dz[(2 * tid) * PO];
        float* db = da + PO;
        if (i < B) {
            float pp = 0.f;
            for (uint32_t j = 0; j < O; ++j) {
                const float e = za[j] - zb[j];
                pp = __builtin_fmaf(e, e, pp);
            }
            const float den = ts[tid] + ss[tid];
            const float r = (pp - ts[tid]) / den;
            loss[i] = r * r;
            const float gp = 2.f * r / den / (float)B;
            for (uint32_t j = 0; j < O; ++j) {
                const float e = 2.f * gp * (za[j] - zb[j]);
                da[j] = e;
                db[j] = -e;
            }
        } else {
            for (uint32_t j = 0; j < O; ++j) { da[j] = 0.f; db[j] = 0.f; }
        }
    }
    __syncthreads();
    float* pb = part + (uint64_t)blockIdx.x * nparam;
    float* pW1 = pb;
    float* pb1 = pW1 + (uint64_t)H * D;
    float* pW2 = pb1 + H;
    float* pb2 = pW2 + (uint64_t)O * H;
    // colsum dz -> gb2 part (rows in order)
    if (tid < O) {
        float acc = 0.f;
        for (uint32_t r = 0; r < MS_ROWS; ++r) acc += L.dz[r * PO + tid];
        pb2[tid] = acc;
    }
    // ---- dh = dz W2, masked by relu' (h > 0); colsum dh -> gb1 part.  Wave w
    // holds whole columns (tiles w, w + 4, ...), so its column sums are
    // complete: over its four row tiles, then over the lane groups.
    for (uint32_t ct = wv; ct < H / 16; ct += 4) {
        const uint32_t h0 = ct * 16;
        mlp_f32x4 acc[4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) acc[rt] = mlp_f32x4{0.f, 0.f, 0.f, 0.f};
        for (uint32_t k0 = 0; k0 < O; k0 += 16) {
            mlp_bf16x4 bo;
#pragma unroll
            for (int q = 0; q < 4; ++q) bo[q] = (short)W2b[(uint64_t)(k0 + 4 * lg + q) * H + h0 + lr];
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) {
                mlp_bf16x4 ao;
#pragma unroll
                for (int q = 0; q < 4; ++q) ao[q] = mlp_to_bf16(L.dz[(rt * 16 + lr) * PO + k0 + 4 * lg + q]);
                acc[rt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ao, bo, acc[rt], 0, 0, 0);
            }
        }
        float cs = 0.f;
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t r = rt * 16 + 4 * lg + q;
                const float dh = (int16_t)L.hb[r * PH + h0 + lr] > 0 ? acc[rt][q] : 0.f;  // h > 0
                cs += dh;
                L.db[r * PH + h0 + lr] = ms_bf(dh);
            }
        cs += __shfl_xor(cs, 16, NNM_WAVE);
        cs += __shfl_xor(cs, 32, NNM_WAVE);
        if (lg == 0) pb1[h0 + lr] = cs;
    }
    __syncthreads();
    // ---- gW2 part = dz^T h (O x H, k over the 64 rows)
    for (uint32_t tt = wv; tt < (O / 16) * (H / 16); tt += 4) {
        const uint32_t o0 = (tt / (H / 16)) * 16, h0 = (tt % (H / 16)) * 16;
        mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (uint32_t k0 = 0; k0 < MS_ROWS; k0 += 16) {
            mlp_bf16x4 ao, bo;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t r = k0 + 4 * lg + q;
                ao[q] = mlp_to_bf16(L.dz[r * PO + o0 + lr]);
                bo[q] = (short)L.hb[r * PH + h0 + lr];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ao, bo, acc, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) pW2[(uint64_t)(o0 + 4 * lg + q) * H + h0 + lr] = acc[q];
    }
    // ---- gW1 part = dh^T x (H x D, k over the 64 rows)
    for (uint32_t tt = wv; tt < (H / 16) * (D / 16); tt += 4) {
        const uint32_t h0 = (tt / (D / 16)) * 16, d0 = (tt % (D / 16)) * 16;
        mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
        for (uint32_t k0 = 0; k0 < MS_ROWS; k0 += 16) {
            mlp_bf16x4 ao, bo;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const uint32_t r = k0 + 4 * lg + q;
                ao[q] = (short)L.db[r * PH + h0 + lr];
                bo[q] = (short)L.xb[r * PD + d0 + lr];
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ao, bo, acc, 0, 0, 0);
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) pW1[(uint64_t)(h0 + 4 * lg + q) * D + d0 + lr] = acc[q];
    }
}

// Sum of the blocks' gradient parts (block order) per parameter; adam = 0:
// store it in g; otherwise an Adam step on w, m, v and the bf16 copies of
// W1 and W2 (Wb: W1 then W2) for the next step.
__global__ void k_mlp_reduce(uint64_t nparam, uint32_t nblk, const float* __restrict__ part, float* __restrict__ g,
                             int adam, float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, float lr,
                             float b1, float b2, float c1, float c2, uint64_t w1n, uint64_t b1n, uint64_t w2n,
                             uint16_t* __restrict__ Wb) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nparam) return;
    float gi = 0.f;
    for (uint32_t b = 0; b < nblk; ++b) gi += part[(uint64_t)b * nparam + i];
    if (!adam) {
        g[i] = gi;
        return;
    }
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    w[i] = wi;
    // bf16 copies: W1 at Wb[0, w1n), W2 after it; the biases have none
    if (i < w1n) Wb[i] = ms_bf(wi);
    else if (i >= w1n + b1n && i < w1n + b1n + w2n) Wb[w1n + (i - w1n - b1n)] = ms_bf(wi);
}

// Projection of every row: 16 rows per wave, hidden layer in LDS (bf16),
// output rows in bf16 plus their squared norms (of the bf16 values).
#define KMAX 256  // longest row (ld) the projection takes; nnm_mlp_init checks it

// W1 and W2 to bf16 once per projection (the MFMA operands; same rounding
// as an on-the-fly conversion).
__global__ void k_mlp_pack(const float* __restrict__ W, uint64_t m, uint16_t* __restrict__ Wb) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        Wb[i] = (uint16_t)mlp_to_bf16(W[i]);
}

// Register budget of 4 waves per SIMD (128 VGPRs, no spills): 4.90 -> 4.61 ms
// for 5M rows of 128 against 3 waves with the compiler's own budget.
template <uint32_t HID, uint32_t DOUT, uint32_t NC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_mlp_project(const float* __restrict__ X, uint64_t n, uint32_t ld,
                                                     float alpha, const uint16_t* __restrict__ W1b,
                                                     const float* __restrict__ b1, const uint16_t* __restrict__ W2b,
                                                     const float* __restrict__ b2, uint16_t* __restrict__ P,
                                                     float* __restrict__ pnorm) {
    __shared__ short hsh[4][16][HID + 4];
    const uint32_t wv = threadIdx.x / NNM_WAVE, lane = threadIdx.x % NNM_WAVE;
    const uint32_t lr = lane & 15u, lg = lane >> 4;
    const uint64_t nblk = (n + 15) / 16;
    for (uint64_t blk = (uint64_t)blockIdx.x * 4 + wv; blk < nblk; blk += (uint64_t)gridDim.x * 4) {
        const uint64_t r0 = blk * 16;
        const uint64_t myrow = min(r0 + lr, n - 1);
        // layer 1: H[16 x HID] = relu(a X W1^T + b1).  The row's operands
        // (bf16 of alpha x, elements k0 + 4 lg .. + 3 of each 16-chunk) are
        // loaded once; W1 comes by 16-byte loads (L2-resident).
        // NC chunks of 16 elements per row: a compile-time count lets every
        // load of a layer issue before its first MFMA; NC = 0 takes the
        // count from ld (a branch per chunk).
        constexpr uint32_t CM = NC ? NC : KMAX / 16;
        mlp_bf16x4 xa[CM];
#pragma unroll
        for (uint32_t c = 0; c < CM; ++c) {
            if (NC || c * 16 < ld) {
                const float4 x4 = *reinterpret_cast<const float4*>(X + myrow * ld + c * 16 + 4 * lg);
                xa[c][0] = mlp_to_bf16(alpha * x4.x);
                xa[c][1] = mlp_to_bf16(alpha * x4.y);
                xa[c][2] = mlp_to_bf16(alpha * x4.z);
                xa[c][3] = mlp_to_bf16(alpha * x4.w);
            }
        }
        for (uint32_t n0 = 0; n0 < HID; n0 += 16) {
            mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            const uint16_t* w1r = W1b + (uint64_t)(n0 + lr) * ld + 4 * lg;
#pragma unroll
            for (uint32_t c = 0; c < CM; ++c) {
                if (NC || c * 16 < ld) {
                    const uint2 w = *reinterpret_cast<const uint2*>(w1r + c * 16);
                    mlp_bf16x4 b;
                    b[0] = (short)(w.x & 0xFFFFu);
                    b[1] = (short)(w.x >> 16);
                    b[2] = (short)(w.y & 0xFFFFu);
                    b[3] = (short)(w.y >> 16);
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xa[c], b, acc, 0, 0, 0);
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) hsh[wv][4 * lg + q][n0 + lr] = mlp_to_bf16(fmaxf(acc[q] + b1[n0 + lr], 0.f));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // layer 2: Z[16 x DOUT] = H W2^T + b2
        float zq[DOUT / 16][4];
#pragma unroll
        for (uint32_t t = 0; t < DOUT / 16; ++t) {
            mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (uint32_t k0 = 0; k0 < HID; k0 += 16) {
                mlp_bf16x4 a, b;
                const uint2 w = *reinterpret_cast<const uint2*>(W2b + (uint64_t)(t * 16 + lr) * HID + k0 + 4 * lg);
#pragma unroll
                for (int q = 0; q < 4; ++q) a[q] = hsh[wv][lr][k0 + 4 * lg + q];
                b[0] = (short)(w.x & 0xFFFFu);
                b[1] = (short)(w.x >> 16);
                b[2] = (short)(w.y & 0xFFFFu);
                b[3] = (short)(w.y >> 16);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) zq[t][q] = acc[q] + b2[t * 16 + lr];
        }
        // lane holds Z[4 lg + q][t*16 + lr]; write bf16 and accumulate norms
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint64_t row = r0 + 4 * lg + q;
            float nrm = 0.f;
#pragma unroll
            for (uint32_t t = 0; t < DOUT / 16; ++t) {
                const short hb = mlp_to_bf16(zq[t][q]);
                const float r = __uint_as_float((uint32_t)(uint16_t)hb << 16);
                nrm = __builtin_fmaf(r, r, nrm);
                if (row < n) P[row * DOUT + t * 16 + lr] = (uint16_t)hb;
            }
            // sum over the 16 lanes of this row group
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) nrm += __shfl_xor(nrm, o, NNM_WAVE);
            if (lr == 0 && row < n) pnorm[row] = nrm;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Projection with both weight matrices in LDS (bf16, loaded once per block;
// rows padded by 4 elements, so the 16 weight rows that one MFMA operand
// reads fall in distinct banks).  PJ_WAVES waves, 16 data rows per wave at a
// time, the next 16 rows' loads issued before the current ones' MFMAs.  Same
// sums in the same order as k_mlp_project (bit-identical output).
#define PJ_WAVES 8u
template <uint32_t HID, uint32_t DOUT, uint32_t NC>
__global__ __launch_bounds__(PJ_WAVES * NNM_WAVE) void k_mlp_project_lds(
    const float* __restrict__ X, uint64_t n, float alpha, const uint16_t* __restrict__ W1b,
    const float* __restrict__ b1, const uint16_t* __restrict__ W2b, const float* __restrict__ b2,
    uint16_t* __restrict__ P, float* __restrict__ pnorm) {
    constexpr uint32_t D = NC * 16, P1 = D + 4, P2 = HID + 4;
    __shared__ __align__(16) short w1s[HID * P1];
    __shared__ __align__(16) short w2s[DOUT * P2];
    __shared__ __align__(16) short hsh[PJ_WAVES][16][HID + 4];
    for (uint32_t i = threadIdx.x; i < HID * D / 4; i += blockDim.x) {
        const uint32_t r = i / (D / 4), c = i % (D / 4);
        *reinterpret_cast<uint2*>(&w1s[r * P1 + 4 * c]) = *reinterpret_cast<const uint2*>(W1b + (uint64_t)r * D + 4 * c);
    }
    for (uint32_t i = threadIdx.x; i < DOUT * HID / 4; i += blockDim.x) {
        const uint32_t r = i / (HID / 4), c = i % (HID / 4);
        *reinterpret_cast<uint2*>(&w2s[r * P2 + 4 * c]) =
            *reinterpret_cast<const uint2*>(W2b + (uint64_t)r * HID + 4 * c);
    }
    __syncthreads();
    const uint32_t wv = threadIdx.x / NNM_WAVE, lane = threadIdx.x % NNM_WAVE;
    const uint32_t lr = lane & 15u, lg = lane >> 4;
    const uint64_t nblk = (n + 15) / 16, step = (uint64_t)gridDim.x * PJ_WAVES;
    uint64_t blk = (uint64_t)blockIdx.x * PJ_WAVES + wv;
    float4 xn[NC];
    auto load = [&](uint64_t b, float4(&x)[NC]) {
        const uint64_t row = min(b * 16 + lr, n - 1);
#pragma unroll
        for (uint32_t c = 0; c < NC; ++c) x[c] = *reinterpret_cast<const float4*>(X + row * D + c * 16 + 4 * lg);
    };
    if (blk < nblk) load(blk, xn);
    for (; blk < nblk; blk += step) {
        const uint64_t r0 = blk * 16;
        mlp_bf16x4 xa[NC];
#pragma unroll
        for (uint32_t c = 0; c < NC; ++c) {
            xa[c][0] = mlp_to_bf16(alpha * xn[c].x);
            xa[c][1] = mlp_to_bf16(alpha * xn[c].y);
            xa[c][2] = mlp_to_bf16(alpha * xn[c].z);
            xa[c][3] = mlp_to_bf16(alpha * xn[c].w);
        }
        if (blk + step < nblk) load(blk + step, xn);
        // layer 1: H[16 x HID] = relu(a X W1^T + b1)
        for (uint32_t n0 = 0; n0 < HID; n0 += 16) {
            mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            const short* w1r = &w1s[(n0 + lr) * P1 + 4 * lg];
#pragma unroll
            for (uint32_t c = 0; c < NC; ++c) {
                const uint2 w = *reinterpret_cast<const uint2*>(w1r + c * 16);
                mlp_bf16x4 b;
                b[0] = (short)(w.x & 0xFFFFu);
                b[1] = (short)(w.x >> 16);
                b[2] = (short)(w.y & 0xFFFFu);
                b[3] = (short)(w.y >> 16);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xa[c], b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) hsh[wv][4 * lg + q][n0 + lr] = mlp_to_bf16(fmaxf(acc[q] + b1[n0 + lr], 0.f));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // layer 2: Z[16 x DOUT] = H W2^T + b2
        float zq[DOUT / 16][4];
#pragma unroll
        for (uint32_t t = 0; t < DOUT / 16; ++t) {
            mlp_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (uint32_t k0 = 0; k0 < HID; k0 += 16) {
                mlp_bf16x4 a, b;
                const uint2 w = *reinterpret_cast<const uint2*>(&w2s[(t * 16 + lr) * P2 + k0 + 4 * lg]);
#pragma unroll
                for (int q = 0; q < 4; ++q) a[q] = hsh[wv][lr][k0 + 4 * lg + q];
                b[0] = (short)(w.x & 0xFFFFu);
                b[1] = (short)(w.x >> 16);
                b[2] = (short)(w.y & 0xFFFFu);
                b[3] = (short)(w.y >> 16);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) zq[t][q] = acc[q] + b2[t * 16 + lr];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint64_t row = r0 + 4 * lg + q;
            float nrm = 0.f;
#pragma unroll
            for (uint32_t t = 0; t < DOUT / 16; ++t) {
                const short hb = mlp_to_bf16(zq[t][q]);
                const float r = __uint_as_float((uint32_t)(uint16_t)hb << 16);
                nrm = __builtin_fmaf(r, r, nrm);
                if (row < n) P[row * DOUT + t * 16 + lr] = (uint16_t)hb;
            }
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) nrm += __shfl_xor(nrm, o, NNM_WAVE);
            if (lr == 0 && row < n) pnorm[row] = nrm;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---------------------------------------------------------------- host side

static uint32_t grid_for_mlp(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + 255) / 256); }

template <typename T>
static int mlp_alloc(nnm_ctx* c, T** p, size_t n) {
    return nnm_alloc(c, reinterpret_cast<void**>(p), n * sizeof(T));
}

void nnm_mlp_free(nnm_ctx* c) {
    nnm_mlp_state* M = mlp_of(c);
    if (!M) return;
    void* ps[] = {M->W1, M->m, M->v, M->g, M->xs, M->pre, M->h, M->z, M->dz, M->dh, M->t, M->s, M->loss,
                  M->pairs, M->Wb, M->P, M->pnorm, M->P8, M->pnorm8, M->amax, M->part};
    for (void* p : ps) nnm_free(c, p);
    delete M;
    c->mlp = nullptr;
}

static int gemm(nnm_ctx* c, bool bf16, uint32_t M, uint32_t N, uint32_t K, const float* A, uint32_t am, uint32_t ak,
                const float* B, uint32_t bk, uint32_t bn, const float* bias, float* C, uint32_t cm) {
    const uint32_t tiles = (M / 16) * (N / 16);
    const uint32_t blocks = (tiles + 3) / 4;
    if (bf16)
        hipLaunchKernelGGL(k_gemm<true>, dim3(blocks), dim3(256), 0, c->stream, M, N, K, A, am, ak, B, bk, bn, bias, C, cm);
    else
        hipLaunchKernelGGL(k_gemm<false>, dim3(blocks), dim3(256), 0, c->stream, M, N, K, A, am, ak, B, bk, bn, bias, C, cm);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

static int ensure_batch(nnm_ctx* c, uint32_t B) {
    nnm_mlp_state* M = mlp_of(c);
    const uint32_t rows = 2 * B;
    if (M->cap_rows >= rows) return NNM_OK;
    void* ps[] = {M->xs, M->pre, M->h, M->z, M->dz, M->dh, M->t, M->s, M->loss};
    for (void* p : ps) nnm_free(c, p);
    int r;
    if ((r = mlp_alloc(c, &M->xs, (size_t)rows * M->din)) || (r = mlp_alloc(c, &M->pre, (size_t)rows * M->hid)) ||
        (r = mlp_alloc(c, &M->h, (size_t)rows * M->hid)) || (r = mlp_alloc(c, &M->z, (size_t)rows * M->dout)) ||
        (r = mlp_alloc(c, &M->dz, (size_t)rows * M->dout)) || (r = mlp_alloc(c, &M->dh, (size_t)rows * M->hid)) ||
        (r = mlp_alloc(c, &M->t, B)) || (r = mlp_alloc(c, &M->s, B)) || (r = mlp_alloc(c, &M->loss, B)))
        return r;
    M->cap_rows = rows;
    return NNM_OK;
}

// forward + loss + backward on the batch already in xs / t / s; gradients in M->g
static int forward_backward(nnm_ctx* c, uint32_t B, bool bf16) {
    nnm_mlp_state* M = mlp_of(c);
    if (2 * B > M->cap_rows) return nnm_fail(c, NNM_ERR_STATE, "mlp batch larger than its buffers");
    const uint32_t R = 2 * B, D = M->din, H = M->hid, O = M->dout;
    float* gW1 = M->g;
    float* gb1 = gW1 + (size_t)H * D;
    float* gW2 = gb1 + H;
    float* gb2 = gW2 + (size_t)O * H;
    // pre = xs W1^T + b1 ; h = relu(pre) ; z = h W2^T + b2
    NNM_TRY_MLP(gemm(c, bf16, R, H, D, M->xs, D, 1, M->W1, 1, D, M->b1, M->pre, H));
    hipLaunchKernelGGL(k_relu, dim3((R * H + 255) / 256), dim3(256), 0, c->stream, M->pre, M->h, (uint64_t)R * H);
    NNM_TRY_MLP(gemm(c, bf16, R, O, H, M->h, H, 1, M->W2, 1, H, M->b2, M->z, O));
    hipLaunchKernelGGL(k_pair_loss, dim3((B + 255) / 256), dim3(256), 0, c->stream, B, O, M->z, M->t, M->s, M->dz,
                       M->loss);
    // gW2 = dz^T h ; gb2 = colsum dz ; dh = dz W2 (relu mask) ; gW1 = dh^T xs ; gb1 = colsum dh
    NNM_TRY_MLP(gemm(c, bf16, O, H, R, M->dz, 1, O, M->h, H, 1, nullptr, gW2, H));
    hipLaunchKernelGGL(k_colsum, dim3((O + 63) / 64), dim3(64, 16), 0, c->stream, M->dz, R, O, gb2);
    NNM_TRY_MLP(gemm(c, bf16, R, H, O, M->dz, O, 1, M->W2, H, 1, nullptr, M->dh, H));
    hipLaunchKernelGGL(k_relu_grad, dim3((R * H + 255) / 256), dim3(256), 0, c->stream, M->pre, M->dh, (uint64_t)R * H);
    NNM_TRY_MLP(gemm(c, bf16, H, D, R, M->dh, 1, H, M->xs, D, 1, nullptr, gW1, D));
    hipLaunchKernelGGL(k_colsum, dim3((H + 63) / 64), dim3(64, 16), 0, c->stream, M->dh, R, H, gb1);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

// The fused step takes rows of at most 256 values, 256 hidden units and 64
// outputs (the benchmark's 128 -> 256 -> 32 network).
static bool fused_ok(const nnm_mlp_state* M) {
    return M->din <= MS_DMAX && M->hid <= MS_HMAX && M->dout <= MS_OMAX;
}

// bf16 copies of W1 and W2 (Wb: W1 then W2) and room for the gradient parts
// of a batch of B pairs.
static int fused_prepare(nnm_ctx* c, uint32_t B) {
    nnm_mlp_state* M = mlp_of(c);
    const uint64_t w1n = (uint64_t)M->hid * M->din, w2n = (uint64_t)M->dout * M->hid;
    if (!M->Wb) NNM_TRY_MLP(mlp_alloc(c, &M->Wb, w1n + w2n));
    hipLaunchKernelGGL(k_mlp_pack, dim3(grid_for_mlp(w1n)), dim3(256), 0, c->stream, M->W1, w1n, M->Wb);
    hipLaunchKernelGGL(k_mlp_pack, dim3(grid_for_mlp(w2n)), dim3(256), 0, c->stream, M->W2, w2n, M->Wb + w1n);
    NNM_LAUNCH_CHECK(c);
    const uint32_t nblk = (B + MS_ROWS / 2 - 1) / (MS_ROWS / 2);
    if ((uint64_t)nblk * M->nparam > M->part_cap) {
        nnm_free(c, M->part);
        M->part = nullptr;
        M->part_cap = 0;
        NNM_TRY_MLP(mlp_alloc(c, &M->part, (size_t)nblk * M->nparam));
        M->part_cap = (uint64_t)nblk * M->nparam;
    }
    return NNM_OK;
}

// One fused step on B pairs: gathered from the lists (gather) or the rows in
// M->xs with M->t / M->s; then the gradient into M->g (lr < 0) or an Adam step.
static int fused_step(nnm_ctx* c, uint32_t B, bool gather, uint64_t seed, float lr) {
    nnm_mlp_state* M = mlp_of(c);
    const uint32_t nblk = (B + MS_ROWS / 2 - 1) / (MS_ROWS / 2);
    hipLaunchKernelGGL(k_mlp_step, dim3(nblk), dim3(256), 0, c->stream, B, M->din, M->hid, M->dout, M->Wb, M->b1,
                       M->Wb + (uint64_t)M->hid * M->din, M->b2, gather ? c->X : nullptr, c->ld, c->n_act,
                       c->K, c->g_ids, c->g_dist, c->label, M->alpha, seed, M->xs, M->t, M->s, M->loss, M->part,
                       M->nparam);
    NNM_LAUNCH_CHECK(c);
    const uint64_t w1n = (uint64_t)M->hid * M->din, w2n = (uint64_t)M->dout * M->hid;
    const bool adam = lr >= 0.f;
    float c1 = 1.f, c2 = 1.f;
    const float b1 = 0.9f, b2 = 0.999f;
    if (adam) {
        ++M->step;
        c1 = 1.f - std::pow(b1, (float)M->step);
        c2 = 1.f - std::pow(b2, (float)M->step);
    }
    hipLaunchKernelGGL(k_mlp_reduce, dim3(grid_for_mlp(M->nparam)), dim3(256), 0, c->stream, M->nparam, nblk, M->part,
                       M->g, adam ? 1 : 0, M->W1, M->m, M->v, lr, b1, b2, c1, c2, w1n, (uint64_t)M->hid, w2n, M->Wb);
    NNM_LAUNCH_CHECK(c);
    return NNM_OK;
}

extern "C" int nnm_mlp_init(nnm_ctx* c, uint32_t hidden, uint32_t d_out, uint64_t seed) {
    if (!c) return NNM_ERR_INVALID;
    if (!c->X) return nnm_fail(c, NNM_ERR_STATE, "no data");
    NNM_TRY(nnm_require_l2(c, "nnm_mlp_init"));
    if (c->ld > KMAX) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "the learned projection takes rows of at most 256 values");
    if (hidden == 0 || d_out == 0 || hidden % 16 || d_out % 16 || hidden > 1024 || d_out > 64)
        return nnm_fail(c, NNM_ERR_INVALID, "hidden and d_out must be multiples of 16 (d_out <= 64)");
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    nnm_mlp_free(c);
    nnm_mlp_state* M = new nnm_mlp_state();
    c->mlp = M;
    M->din = c->ld;
    M->hid = hidden;
    M->dout = d_out;
    M->nparam = (uint64_t)hidden * M->din + hidden + (uint64_t)d_out * hidden + d_out;
    int r;
    if ((r = mlp_alloc(c, &M->W1, M->nparam)) || (r = mlp_alloc(c, &M->m, M->nparam)) ||
        (r = mlp_alloc(c, &M->v, M->nparam)) || (r = mlp_alloc(c, &M->g, M->nparam)))
        return r;
    M->b1 = M->W1 + (size_t)hidden * M->din;
    M->W2 = M->b1 + hidden;
    M->b2 = M->W2 + (size_t)d_out * hidden;
    // He-normal init from the counter-based stream (Box-Muller), zero biases
    std::vector<float> w(M->nparam, 0.f);
    auto gauss = [&](uint64_t i) {
        const double u1 = ((nnm_rand(seed, i, 0) >> 11) + 1.0) / 9007199254740993.0;
        const double u2 = (nnm_rand(seed, i, 1) >> 11) / 9007199254740992.0;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    const double s1 = std::sqrt(2.0 / c->d), s2 = std::sqrt(1.0 / hidden);
    for (uint32_t o = 0; o < hidden; ++o)
        for (uint32_t k = 0; k < c->d; ++k) w[(size_t)o * M->din + k] = (float)(s1 * gauss((size_t)o * M->din + k));
    const size_t w2 = (size_t)hidden * M->din + hidden;
    for (size_t i = 0; i < (size_t)d_out * hidden; ++i) w[w2 + i] = (float)(s2 * gauss(w2 + i));
    NNM_TRY_MLP(nnm_h2d(c, M->W1, w.data(), M->nparam * 4));
    NNM_HIP(c, hipMemsetAsync(M->m, 0, M->nparam * 4, c->stream));
    NNM_HIP(c, hipMemsetAsync(M->v, 0, M->nparam * 4, c->stream));
    // input scale: a^2 mean |x|^2 = d over a sample of rows
    {
        const uint64_t ns = std::min<uint64_t>(c->n, 4096);
        std::vector<float> rows(ns * c->ld);
        double acc = 0.0;
        // the first rows of ns evenly spaced blocks of the matrix
        const uint64_t blk = 64, nb = (ns + blk - 1) / blk;
        for (uint64_t b = 0; b < nb; ++b) {
            const uint64_t r = (c->n - blk) * b / std::max<uint64_t>(1, nb - 1);
            const uint64_t m = std::min(blk, ns - b * blk);
            NNM_TRY_MLP(nnm_d2h(c, rows.data() + b * blk * c->ld, c->X + r * c->ld, m * c->ld * 4));
        }
        for (float x : rows) acc += (double)x * x;
        acc /= (double)ns;
        M->alpha = acc > 0 ? (float)std::sqrt((double)c->d / acc) : 1.f;
    }
    M->step = 0;
    M->trained = 0;
    M->tau = 1.f;
    return NNM_OK;
}

extern "C" int nnm_mlp_set_weights(nnm_ctx* c, const float* W1, const float* b1, const float* W2, const float* b2,
                                   float alpha) {
    if (!c || !mlp_of(c)) return nnm_fail(c, NNM_ERR_STATE, "no model");
    nnm_mlp_state* M = mlp_of(c);
    NNM_TRY_MLP(nnm_h2d(c, M->W1, W1, (size_t)M->hid * M->din * 4));
    NNM_TRY_MLP(nnm_h2d(c, M->b1, b1, (size_t)M->hid * 4));
    NNM_TRY_MLP(nnm_h2d(c, M->W2, W2, (size_t)M->dout * M->hid * 4));
    NNM_TRY_MLP(nnm_h2d(c, M->b2, b2, (size_t)M->dout * 4));
    if (alpha > 0.f) M->alpha = alpha;
    nnm_mlp_invalidate_projection(c);  // the cached rows came from the old weights
    return NNM_OK;
}

extern "C" int nnm_mlp_get_weights(nnm_ctx* c, float* W1, float* b1, float* W2, float* b2, float* alpha,
                                   float* tau) {
    if (!c || !mlp_of(c)) return nnm_fail(c, NNM_ERR_STATE, "no model");
    nnm_mlp_state* M = mlp_of(c);
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    if (W1) NNM_TRY_MLP(nnm_d2h(c, W1, M->W1, (size_t)M->hid * M->din * 4));
    if (b1) NNM_TRY_MLP(nnm_d2h(c, b1, M->b1, (size_t)M->hid * 4));
    if (W2) NNM_TRY_MLP(nnm_d2h(c, W2, M->W2, (size_t)M->dout * M->hid * 4));
    if (b2) NNM_TRY_MLP(nnm_d2h(c, b2, M->b2, (size_t)M->dout * 4));
    if (alpha) *alpha = M->alpha;
    if (tau) *tau = M->tau;
    return NNM_OK;
}

extern "C" int nnm_mlp_loss_grad(nnm_ctx* c, const float* xa, const float* xb, const float* t, const float* s,
                                 uint32_t B, int bf16, float* loss, float* gW1, float* gb1, float* gW2, float* gb2) {
    if (!c || !mlp_of(c)) return nnm_fail(c, NNM_ERR_STATE, "no model");
    if (!xa || !xb || !t || !s || B == 0 || B % 8) return nnm_fail(c, NNM_ERR_INVALID, "B must be a multiple of 8");
    nnm_mlp_state* M = mlp_of(c);
    NNM_TRY_MLP(ensure_batch(c, B));
    // rows 2i = xa[i], 2i+1 = xb[i] (already scaled by the caller); d_in columns, padded with zeros
    std::vector<float> xs((size_t)2 * B * M->din, 0.f);
    for (uint32_t i = 0; i < B; ++i)
        for (uint32_t k = 0; k < c->d; ++k) {
            xs[(size_t)(2 * i) * M->din + k] = xa[(size_t)i * c->d + k];
            xs[(size_t)(2 * i + 1) * M->din + k] = xb[(size_t)i * c->d + k];
        }
    NNM_TRY_MLP(nnm_h2d(c, M->xs, xs.data(), xs.size() * 4));
    NNM_TRY_MLP(nnm_h2d(c, M->t, t, B * 4));
    NNM_TRY_MLP(nnm_h2d(c, M->s, s, B * 4));
    if (bf16 && fused_ok(M)) {
        // the fused step that training runs, without its Adam update
        NNM_TRY_MLP(fused_prepare(c, B));
        NNM_TRY_MLP(fused_step(c, B, false, 0, -1.f));
    } else {
        NNM_TRY_MLP(forward_backward(c, B, bf16 != 0));
    }
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<float> lv(B);
    NNM_TRY_MLP(nnm_d2h(c, lv.data(), M->loss, B * 4));
    double acc = 0;
    for (float x : lv) acc += x;
    if (loss) *loss = (float)(acc / B);
    const float* g = M->g;
    const size_t nW1 = (size_t)M->hid * M->din, nW2 = (size_t)M->dout * M->hid;
    // report gW1 without the padded input columns
    if (gW1) {
        std::vector<float> tmp(nW1);
        NNM_TRY_MLP(nnm_d2h(c, tmp.data(), g, nW1 * 4));
        for (uint32_t o = 0; o < M->hid; ++o)
            std::memcpy(gW1 + (size_t)o * c->d, tmp.data() + (size_t)o * M->din, c->d * 4);
    }
    if (gb1) NNM_TRY_MLP(nnm_d2h(c, gb1, g + nW1, M->hid * 4));
    if (gW2) NNM_TRY_MLP(nnm_d2h(c, gW2, g + nW1 + M->hid, nW2 * 4));
    if (gb2) NNM_TRY_MLP(nnm_d2h(c, gb2, g + nW1 + M->hid + nW2, M->dout * 4));
    return NNM_OK;
}

// Ranking diagnostic of the trained model: on fresh pairs, the smallest tau
// such that a fraction `keep` of the pairs with t < s (true improvements near
// the k-th distance) has p < tau * s, computed on the host from a sample.  It
// is reported with the weights; the learned join's threshold is learn_tau.
static int calibrate_tau(nnm_ctx* c, uint32_t B, float keep, uint64_t seed) {
    nnm_mlp_state* M = mlp_of(c);
    NNM_TRY_MLP(ensure_batch(c, B));
    hipLaunchKernelGGL(k_mlp_pairs, dim3((B + 255) / 256), dim3(256), 0, c->stream, B, c->n_act, c->K, c->ld, c->X,
                       c->g_ids, c->g_dist, c->label, M->alpha, seed, M->xs, M->t, M->s);
    NNM_LAUNCH_CHECK(c);
    NNM_TRY_MLP(forward_backward(c, B, true));
    NNM_HIP(c, hipStreamSynchronize(c->stream));
    std::vector<float> z((size_t)2 * B * M->dout), t(B), s(B);
    NNM_TRY_MLP(nnm_d2h(c, z.data(), M->z, z.size() * 4));
    NNM_TRY_MLP(nnm_d2h(c, t.data(), M->t, B * 4));
    NNM_TRY_MLP(nnm_d2h(c, s.data(), M->s, B * 4));
    std::vector<float> ratio;
    for (uint32_t i = 0; i < B; ++i) {
        if (!(t[i] < s[i]) || s[i] <= 0.f) continue;
        double p = 0;
        for (uint32_t j = 0; j < M->dout; ++j) {
            const double e = (double)z[(size_t)(2 * i) * M->dout + j] - z[(size_t)(2 * i + 1) * M->dout + j];
            p += e * e;
        }
        ratio.push_back((float)(p / s[i]));
    }
    if (ratio.size() < 16) { M->tau = 2.f; return NNM_OK; }
    std::sort(ratio.begin(), ratio.end());
    const size_t q = std::min(ratio.size() - 1, (size_t)(keep * ratio.size()));
    M->tau = std::max(1.f, ratio[q]);
    return NNM_OK;
}

extern "C" int nnm_mlp_train(nnm_ctx* c, uint32_t steps, uint32_t batch, float lr, float keep, uint64_t seed,
                             float* loss_first, float* loss_last, double* secs) {
    if (!c || !mlp_of(c)) return nnm_fail(c, NNM_ERR_STATE, "no model");
    if (!c->g_ids) return nnm_fail(c, NNM_ERR_STATE, "training pairs come from a graph: none is set");
    if (batch == 0 || batch % 8 || steps == 0) return nnm_fail(c, NNM_ERR_INVALID, "bad batch or steps");
    nnm_mlp_state* M = mlp_of(c);
    NNM_TRY_MLP(ensure_batch(c, batch));
    NNM_HIP(c, hipEventRecord(c->ev0, c->stream));
    float lf = 0.f, ll = 0.f;
    std::vector<float> lv(batch);
    const bool fused = fused_ok(M);
    if (fused) NNM_TRY_MLP(fused_prepare(c, batch));
    for (uint32_t it = 0; it < steps; ++it) {
        if (2 * batch > M->cap_rows) return nnm_fail(c, NNM_ERR_STATE, "mlp batch larger than its buffers");
        if (fused) {
            // pairs, forward, backward and the Adam step in two launches
            NNM_TRY_MLP(fused_step(c, batch, true, nnm_rand(seed, it, 77), lr));
        } else {
            hipLaunchKernelGGL(k_mlp_pairs, dim3((batch + 255) / 256), dim3(256), 0, c->stream, batch, c->n_act,
                               c->K, c->ld, c->X, c->g_ids, c->g_dist, c->label, M->alpha, nnm_rand(seed, it, 77),
                               M->xs, M->t, M->s);
            NNM_LAUNCH_CHECK(c);
            NNM_TRY_MLP(forward_backward(c, batch, true));
            ++M->step;
            const float b1 = 0.9f, b2 = 0.999f;
            const float c1 = 1.f - std::pow(b1, (float)M->step), c2 = 1.f - std::pow(b2, (float)M->step);
            hipLaunchKernelGGL(k_adam, dim3((M->nparam + 255) / 256), dim3(256), 0, c->stream, M->W1, M->g, M->m,
                               M->v, M->nparam, lr, b1, b2, c1, c2);
            NNM_LAUNCH_CHECK(c);
        }
        if (it == 0 || it + 1 == steps) {
            NNM_TRY_MLP(nnm_d2h(c, lv.data(), M->loss, batch * 4));
            double acc = 0;
            for (float x : lv) acc += x;
            (it == 0 ? lf : ll) = (float)(acc / batch);
        }
    }
    NNM_HIP(c, hipEventRecord(c->ev1, c->stream));
    NNM_HIP(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    NNM_TRY_MLP(calibrate_tau(c, 8192, keep, nnm_rand(seed, 0xCA11B, 0)));
    M->trained = 1;
    nnm_mlp_invalidate_projection(c);  // the cached rows came from the old weights
    if (loss_first) *loss_first = lf;
    if (loss_last) *loss_last = ll;
    if (secs) *secs = ms * 1e-3;
    return NNM_OK;
}

int nnm_mlp_ensure_projection(nnm_ctx* c, double* secs) {
    nnm_mlp_state* M = mlp_of(c);
    if (!M) return nnm_fail(c, NNM_ERR_STATE, "learned join needs a model (nnm_mlp_init / nnm_mlp_train)");
    if (M->din != c->ld) return nnm_fail(c, NNM_ERR_STATE, "model was built for another row length");
    if (M->projected_n == c->n && M->P) {
        if (secs) *secs = 0.0;  // already current: no work
        return NNM_OK;
    }
    nnm_free(c, M->P);
    nnm_free(c, M->pnorm);
    M->P = nullptr;
    M->pnorm = nullptr;
    NNM_TRY_MLP(mlp_alloc(c, &M->P, (size_t)c->n * M->dout));
    NNM_TRY_MLP(mlp_alloc(c, &M->pnorm, c->n));
    const uint64_t w1n = (uint64_t)M->hid * M->din, w2n = (uint64_t)M->dout * M->hid;
    if (!M->Wb) NNM_TRY_MLP(mlp_alloc(c, &M->Wb, w1n + w2n));
    const uint16_t* W1b = M->Wb;
    const uint16_t* W2b = M->Wb + w1n;
    NNM_HIP(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_mlp_pack, dim3(64), dim3(256), 0, c->stream, M->W1, w1n, M->Wb);
    hipLaunchKernelGGL(k_mlp_pack, dim3(16), dim3(256), 0, c->stream, M->W2, w2n, M->Wb + w1n);
    NNM_LAUNCH_CHECK(c);
    const uint32_t grid = (uint32_t)std::min<uint64_t>((c->n + 63) / 64, (uint64_t)c->num_cus * 16);
#define NNM_PROJECT(H, D, NC)                                                                                   \
    hipLaunchKernelGGL((k_mlp_project<H, D, NC>), dim3(grid), dim3(256), 0, c->stream, c->X, c->n, c->ld, M->alpha, \
                       W1b, M->b1, W2b, M->b2, M->P, M->pnorm)
    if (M->hid == 256 && M->dout == 32 && c->ld == 128)
        hipLaunchKernelGGL((k_mlp_project_lds<256, 32, 8>), dim3((uint32_t)std::min<uint64_t>((c->n + 15) / 16, c->num_cus)),
                           dim3(PJ_WAVES * NNM_WAVE), 0, c->stream, c->X, c->n, M->alpha, W1b, M->b1, W2b, M->b2, M->P,
                           M->pnorm);
    else if (M->hid == 256 && M->dout == 32)
        NNM_PROJECT(256, 32, 0);
    else if (M->hid == 128 && M->dout == 32)
        NNM_PROJECT(128, 32, 0);
    else if (M->hid == 256 && M->dout == 16)
        NNM_PROJECT(256, 16, 0);
#undef NNM_PROJECT
    else
        return nnm_fail(c, NNM_ERR_UNSUPPORTED, "projection kernel built for (hidden, d_out) in {(256,32),(128,32),(256,16)}");
    NNM_LAUNCH_CHECK(c);
    NNM_HIP(c, hipEventRecord(c->ev1, c->stream));
    NNM_HIP(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    if (secs) *secs = ms * 1e-3;
    M->projected_n = c->n;
    // numbered across models and contexts: copies made from an earlier
    // model's projection never match a later one's number
    static std::atomic<uint64_t> proj_seq{0};
    M->proj_gen = ++proj_seq;
    return NNM_OK;
}

void nnm_mlp_invalidate_projection(nnm_ctx* c) {
    nnm_mlp_state* M = mlp_of(c);
    if (M) {
        M->projected_n = 0;
        M->fp8_gen = 0;
    }
}

// ---- FP8 projection rows (the learned join's FP8 prefilter)

// max |P| over the projection: bf16 magnitudes order like their bits, so an
// atomicMax on the bits (sign cleared) finds it.
__global__ void k_bf16_absmax(const uint16_t* __restrict__ P, uint64_t m, uint32_t* __restrict__ out) {
    uint32_t mx = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        mx = max(mx, (uint32_t)(P[i] & 0x7FFFu));
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, 64));
    if ((threadIdx.x & 63u) == 0 && mx) atomicMax(out, mx);
}

// OCP e4m3 (e4m3fn) value of a byte: bias 7, no infinities (the encoder
// below never produces the NaN codes).
__device__ inline float fp8_e4m3_value(uint32_t b) {
    const uint32_t e = (b >> 3) & 15u, m = b & 7u;
    const float v = e ? __uint_as_float(((e + 120u) << 23) | (m << 20)) : (float)m * 0.001953125f;  // m 2^-9
    return (b & 0x80u) ? -v : v;
}

// Row r: x = clamp(s P[r], +-448) rounded to e4m3 by the hardware convert
// (round to nearest even; the clamp keeps it off the NaN code), and the
// squared norm of the encoded row in element order (FMA), the norm form's
// |q|^2.  One thread per row, rows of dout (a multiple of 4) elements.
__global__ void k_mlp_to_fp8(const uint16_t* __restrict__ P, uint64_t n, uint32_t dout,
                             const uint32_t* __restrict__ amax, uint8_t* __restrict__ P8,
                             float* __restrict__ pnorm8) {
    const float mx = __uint_as_float(*amax << 16);
    const float s = mx > 0.f ? 448.f / mx : 1.f;
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (uint64_t)gridDim.x * blockDim.x) {
        const uint16_t* src = P + r * dout;
        uint32_t* dst = reinterpret_cast<uint32_t*>(P8 + r * dout);
        float nrm = 0.f;
        for (uint32_t k = 0; k < dout; k += 4) {
            float x[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                x[j] = fminf(fmaxf(__uint_as_float((uint32_t)src[k + j] << 16) * s, -448.f), 448.f);
            int w = __builtin_amdgcn_cvt_pk_fp8_f32(x[0], x[1], 0, false);
            w = __builtin_amdgcn_cvt_pk_fp8_f32(x[2], x[3], w, true);
            dst[k / 4] = (uint32_t)w;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float q = fp8_e4m3_value(((uint32_t)w >> (8 * j)) & 0xFFu);
                nrm = __builtin_fmaf(q, q, nrm);
            }
        }
        pnorm8[r] = nrm;
    }
}

int nnm_mlp_ensure_fp8(nnm_ctx* c) {
    nnm_mlp_state* M = mlp_of(c);
    NNM_TRY_MLP(nnm_mlp_ensure_projection(c, nullptr));
    if (M->dout % 4) return nnm_fail(c, NNM_ERR_UNSUPPORTED, "FP8 rows need d_out a multiple of 4");
    if (M->P8 && M->fp8_gen == M->proj_gen) return NNM_OK;
    if (!M->P8 || M->fp8_rows != c->n) {
        nnm_free(c, M->P8);
        nnm_free(c, M->pnorm8);
        M->P8 = nullptr;
        M->pnorm8 = nullptr;
        M->fp8_rows = 0;
        NNM_TRY_MLP(mlp_alloc(c, &M->P8, (size_t)c->n * M->dout));
        NNM_TRY_MLP(mlp_alloc(c, &M->pnorm8, c->n));
        M->fp8_rows = c->n;
    }
    if (!M->amax) NNM_TRY_MLP(mlp_alloc(c, &M->amax, 1));
    NNM_TRY_MLP(nnm_fill_u32(c, M->amax, 1, 0));
    const uint64_t m = (uint64_t)c->n * M->dout;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((m + 255) / 256, (uint64_t)c->num_cus * 8);
    hipLaunchKernelGGL(k_bf16_absmax, dim3(grid), dim3(256), 0, c->stream, M->P, m, M->amax);
    NNM_LAUNCH_CHECK(c);
    const uint32_t rgrid = (uint32_t)std::min<uint64_t>((c->n + 255) / 256, (uint64_t)c->num_cus * 16);
    hipLaunchKernelGGL(k_mlp_to_fp8, dim3(rgrid), dim3(256), 0, c->stream, M->P, c->n, M->dout, M->amax, M->P8,
                       M->pnorm8);
    NNM_LAUNCH_CHECK(c);
    uint32_t h = 0;
    NNM_TRY_MLP(nnm_d2h(c, &h, M->amax, 4));
    const uint32_t u = h << 16;
    float mx;
    std::memcpy(&mx, &u, 4);
    M->s8 = mx > 0.f ? 448.f / mx : 1.f;
    M->fp8_gen = M->proj_gen;
    return NNM_OK;
}

extern "C" int nnm_mlp_project_fp8(nnm_ctx* c, uint8_t* out, float* scale) {
    if (!c || !mlp_of(c)) return nnm_fail(c, NNM_ERR_STATE, "no model");
    NNM_TRY_MLP(nnm_mlp_ensure_fp8(c));
    nnm_mlp_state* M = mlp_of(c);
    if (out) NNM_TRY_MLP(nnm_d2h(c, out, M->P8, (size_t)c->n * M->dout));
    if (scale) *scale = M->s8;
    return NNM_OK;
}

extern "C" int nnm_mlp_project(nnm_ctx* c, float* out, double* secs) {
    if (!c || !mlp_of(c)) return nnm_fail(c, NNM_ERR_STATE, "no model");
    nnm_mlp_state* M = mlp_of(c);
    M->projected_n = 0;
    NNM_TRY_MLP(nnm_mlp_ensure_projection(c, secs));
    if (out) {
        std::vector<uint16_t> h((size_t)c->n * M->dout);
        NNM_TRY_MLP(nnm_d2h(c, h.data(), M->P, h.size() * 2));
        for (size_t i = 0; i < h.size(); ++i) {
            uint32_t u = (uint32_t)h[i] << 16;
            float f;
            std::memcpy(&f, &u, 4);
            out[i] = f;
        }
    }
    return NNM_OK;
}

// accessors for the join launcher
const uint16_t* nnm_mlp_rows(nnm_ctx* c) { return mlp_of(c) ? mlp_of(c)->P : nullptr; }
uint64_t nnm_mlp_proj_gen(nnm_ctx* c) { return mlp_of(c) && mlp_of(c)->P ? mlp_of(c)->proj_gen : 0; }
const float* nnm_mlp_norms(nnm_ctx* c) { return mlp_of(c) ? mlp_of(c)->pnorm : nullptr; }
uint32_t nnm_mlp_dout(nnm_ctx* c) { return mlp_of(c) ? mlp_of(c)->dout : 0; }
float nnm_mlp_scale(nnm_ctx* c) { return mlp_of(c) ? 1.f / (mlp_of(c)->alpha * mlp_of(c)->alpha) : 1.f; }
const uint8_t* nnm_mlp_rows8(nnm_ctx* c) { return mlp_of(c) ? mlp_of(c)->P8 : nullptr; }
const float* nnm_mlp_norms8(nnm_ctx* c) { return mlp_of(c) ? mlp_of(c)->pnorm8 : nullptr; }
// estimate scale of the FP8 rows: distances of s8-scaled rows back to a^2 units
float nnm_mlp_scale8(nnm_ctx* c) {
    return mlp_of(c) ? 1.f / (mlp_of(c)->alpha * mlp_of(c)->alpha * mlp_of(c)->s8 * mlp_of(c)->s8) : 1.f;
}
int nnm_mlp_trained(nnm_ctx* c) { return mlp_of(c) ? mlp_of(c)->trained : 0; }

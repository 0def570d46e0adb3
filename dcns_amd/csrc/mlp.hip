// mlp.hip -- the lambda/K surrogate network on gfx950: 8 inputs (match,
// mismatch, gap open, gap extend, long-gap open, GC fraction, log ref length,
// log query length) -> 4 x 256 GELU -> (log lambda, log K).
//
// BF16 operands on MFMA (v_mfma_f32_32x32x16_bf16) with FP32 accumulation;
// FP32 master weights and AdamW moments.  Shapes are padded to MFMA tiles:
// the 8 inputs to 32 (zero columns), the 2 outputs to 32 (zero rows), the
// batch to a multiple of 32.  No counterpart in the reference (which fits
// lambda and K by permutation only).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/dcns.h"
#include "ctx.h"

namespace dcns {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kIn = 8, kInPad = 32, kHid = 256, kOut = 2, kOutPad = 32, kLayers = 5;

__device__ __forceinline__ uint16_t bf16_bits(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_float(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

enum Epi { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_GELU_BACK = 3 };

// C[M x N] = sum_k A(m, k) B(k, n), A(m, k) = A[m*sam + k*sak], B(k, n) = B[k*sbk + n*sbn]
// (bf16).  One wave per 32 x 32 tile of C; each K step of 16 stages a
// 32 x 16 slice of A and a 16 x 32 slice of B in LDS, then every lane takes
// its MFMA fragment (row or column l % 32, k = 8*(l/32) .. +7).  The
// accumulator's element i of lane l is C(row0 + 8*(i/4) + 4*(l/32) + i%4,
// col0 + l%32).
template <int EPI>
__global__ __launch_bounds__(64) void k_gemm(const uint16_t *__restrict__ A, int64_t sam, int64_t sak,
                                             const uint16_t *__restrict__ B, int64_t sbk, int64_t sbn, int M, int N,
                                             int K, const float *__restrict__ bias, const float *__restrict__ pre_in,
                                             float *__restrict__ out_f, uint16_t *__restrict__ out_b, int ldo,
                                             int n_valid) {
    __shared__ uint16_t sa[32 * 16], sb[16 * 32];
    const int lane = threadIdx.x;
    const int tiles_n = N / 32;
    const int row0 = (blockIdx.x / tiles_n) * 32, col0 = (blockIdx.x % tiles_n) * 32;
    f32x16 acc = {};
    for (int k0 = 0; k0 < K; k0 += 16) {
        // 512 elements per operand slice, 8 per lane
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const int e = lane * 8 + t;
            const int r = e >> 4, kk = e & 15;  // A slice: row r, k kk
            sa[r * 16 + kk] = A[(int64_t)(row0 + r) * sam + (int64_t)(k0 + kk) * sak];
            const int kb = e >> 5, c = e & 31;  // B slice: k kb, col c
            sb[kb * 32 + c] = B[(int64_t)(k0 + kb) * sbk + (int64_t)(col0 + c) * sbn];
        }
        __syncthreads();
        bf16x8 fa, fb;
        const int r = lane & 31, kq = (lane >> 5) * 8;
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            fa[t] = __builtin_bit_cast(__bf16, sa[r * 16 + kq + t]);
            fb[t] = __builtin_bit_cast(__bf16, sb[(kq + t) * 32 + r]);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc, 0, 0, 0);
        __syncthreads();
    }
    const int col = col0 + (lane & 31);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int row = row0 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        const int64_t o = (int64_t)row * ldo + col;
        float v = acc[i];
        if (EPI == EPI_STORE) {
            out_f[o] = v;
        } else if (EPI == EPI_BIAS_GELU) {
            v += bias[col];
            out_f[o] = v;  // pre-activation, kept for the backward pass
            out_b[o] = bf16_bits(gelu(v));
        } else if (EPI == EPI_BIAS) {
            if (col < n_valid) out_f[(int64_t)row * n_valid + col] = v + bias[col];
        } else {  // EPI_GELU_BACK: v is dAct; dPre = dAct * gelu'(pre)
            const float d = v * gelu_grad(pre_in[o]);
            out_f[o] = d;
            out_b[o] = bf16_bits(d);
        }
    }
}

// The weights of the layers in MFMA fragment order: layer l (K = in_pad,
// N = out_pad), K step s (16 deep), column tile c (32 wide), lane q, element
// t holds Wt(32c + q % 32, 16s + 8 (q / 32) + t) -- the B fragment k_gemm
// builds in LDS, here one 16-byte read per lane.
__global__ void k_mlp_frag(const uint16_t *__restrict__ wt, int K, int N, uint16_t *__restrict__ frag) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (int64_t)K * N) return;
    const int t = (int)(e & 7), q = (int)((e >> 3) & 63);
    const int64_t rest = e >> 9;
    const int nt = N / 32;
    const int c = (int)(rest % nt), st = (int)(rest / nt);
    frag[e] = wt[(int64_t)(32 * c + (q & 31)) * K + 16 * st + 8 * (q >> 5) + t];
}

// Inference, all five layers in one launch.  A wave owns 32 rows: their
// activations stay in LDS (in place, layer after layer) and their 32 x 256
// outputs in 8 accumulators; a workgroup's 8 waves share each 16-deep slice
// of weight fragments through LDS (double-buffered, loaded while the
// previous slice multiplies).  Per output tile the MFMA instruction, operand
// fragments and K order are those of k_gemm, so the results are the same
// bits as the layer-by-layer path.
constexpr int kInferWaves = 8;
constexpr int kActStride = kHid + 8;  // bf16 per LDS activation row (16-byte pad against bank conflicts)
struct MlpOffsets {
    int64_t frag[kLayers];
    int64_t bias[kLayers];
};

__global__ __launch_bounds__(64 * kInferWaves) void k_mlp_infer(const uint16_t *__restrict__ x0, int64_t M,
                                                                const uint16_t *__restrict__ frag,
                                                                const float *__restrict__ p, MlpOffsets off,
                                                                float *__restrict__ out) {
    __shared__ uint16_t act[kInferWaves][32 * kActStride];
    __shared__ uint4 wbuf[2][8 * 64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t row0 = ((int64_t)blockIdx.x * kInferWaves + wave) * 32;
    const bool active = row0 < M;  // M is a multiple of 32
    uint16_t *A = act[wave];
    const int r = lane & 31, kq = (lane >> 5) * 8;
#pragma unroll 1
    for (int l = 0; l < kLayers; ++l) {
        const int S = (l == 0 ? kInPad : kHid) / 16;
        const int NT = l == kLayers - 1 ? 1 : kHid / 32;
        const uint4 *F = (const uint4 *)(frag + off.frag[l]);
        const int slice = NT * 64;
        f32x16 acc[kHid / 32];
#pragma unroll
        for (int c = 0; c < kHid / 32; ++c) acc[c] = f32x16{};
        if (tid < slice) wbuf[0][tid] = F[tid];
        __syncthreads();
#pragma unroll 1
        for (int st = 0; st < S; ++st) {
            const bool more = st + 1 < S && tid < slice;
            uint4 nxt;
            if (more) nxt = F[(int64_t)(st + 1) * slice + tid];
            if (active) {
                uint4 av;
                if (l == 0) av = *(const uint4 *)(x0 + (row0 + r) * kInPad + 16 * st + kq);
                else av = *(const uint4 *)(A + r * kActStride + 16 * st + kq);
                const bf16x8 fa = __builtin_bit_cast(bf16x8, av);
#pragma unroll
                for (int c = 0; c < kHid / 32; ++c)
                    if (c < NT)
                        acc[c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            fa, __builtin_bit_cast(bf16x8, wbuf[st & 1][c * 64 + lane]), acc[c], 0, 0, 0);
            }
            if (more) wbuf[(st + 1) & 1][tid] = nxt;
            __syncthreads();
        }
        if (!active) continue;
        const float *bias = p + off.bias[l];
        const int col_in = lane & 31;
        if (l < kLayers - 1) {
#pragma unroll
            for (int c = 0; c < kHid / 32; ++c)
#pragma unroll
                for (int i = 0; i < 16; ++i) {
                    const int row = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3), col = 32 * c + col_in;
                    A[row * kActStride + col] = bf16_bits(gelu(acc[c][i] + bias[col]));
                }
        } else if (col_in < kOut) {
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int row = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
                out[(row0 + row) * kOut + col_in] = acc[0][i] + bias[col_in];
            }
        }
    }
}

// Standardised inputs, padded to 16 columns, as bf16.
__global__ void k_prep_input(const float *__restrict__ X, const int32_t *__restrict__ idx, int64_t n, int64_t M,
                             const float *__restrict__ mean, const float *__restrict__ inv_std,
                             uint16_t *__restrict__ out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * kInPad) return;
    const int64_t m = t / kInPad;
    const int c = (int)(t % kInPad);
    float v = 0.f;
    if (m < n && c < kIn) {
        const int64_t src = idx ? idx[m] : m;
        v = (X[src * kIn + c] - mean[c]) * inv_std[c];
    }
    out[t] = bf16_bits(v);
}

// Mean squared error over the valid rows and both outputs, and its gradient
// (bf16, padded to 32 columns, zero outside the valid block).
__global__ void k_mse(const float *__restrict__ out, const float *__restrict__ Y, const int32_t *__restrict__ idx,
                      int64_t n, int64_t M, const float *__restrict__ ymean, const float *__restrict__ yinv,
                      uint16_t *__restrict__ grad, float *__restrict__ loss) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * kOutPad) return;
    const int64_t m = t / kOutPad;
    const int c = (int)(t % kOutPad);
    float g = 0.f;
    if (m < n && c < kOut) {
        const int64_t src = idx ? idx[m] : m;
        const float target = (Y[src * kOut + c] - ymean[c]) * yinv[c];
        const float d = out[m * kOut + c] - target;
        g = 2.f * d / (float)(n * kOut);
        atomicAdd(loss, d * d / (float)(n * kOut));
    }
    grad[t] = bf16_bits(g);
}

// Column sums over the valid rows: bias gradients.
__global__ void k_colsum(const float *__restrict__ D, int64_t rows, int ld, int cols, float *__restrict__ out) {
    const int c = blockIdx.x;
    float s = 0.f;
    for (int64_t r = threadIdx.x; r < rows; r += blockDim.x) s += D[r * ld + c];
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && c < cols) out[c] = red[0];
}

// AdamW on fp32 master parameters; refreshes the bf16 copies.  decay_mask
// selects the entries that take weight decay (weights, not biases).
__global__ void k_adamw(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                        float *__restrict__ m2, uint16_t *__restrict__ pb, const uint8_t *__restrict__ decay_mask,
                        int64_t n, float lr, float b1, float b2, float eps, float wd, float c1, float c2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = p[i];
    if (decay_mask[i]) x -= lr * wd * x;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    x -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = x;
    pb[i] = bf16_bits(x);
}

__global__ void k_to_bf16(const float *__restrict__ p, uint16_t *__restrict__ pb, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pb[i] = bf16_bits(p[i]);
}

unsigned blocks(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace
}  // namespace dcns

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
using namespace dcns;

struct dcns_mlp {
    dcns_ctx *ctx = nullptr;
    // padded layer shapes: Wt[l] is [out_pad x in_pad] (torch Linear layout), b[l] is [out_pad]
    int in_pad[kLayers] = {kInPad, kHid, kHid, kHid, kHid};
    int out_pad[kLayers] = {kHid, kHid, kHid, kHid, kOutPad};
    int in_real[kLayers] = {kIn, kHid, kHid, kHid, kHid};
    int out_real[kLayers] = {kHid, kHid, kHid, kHid, kOut};
    int64_t w_off[kLayers], b_off[kLayers], total = 0;
    DBuf p, g, m1, m2, pb, mask;  // fp32 params / grads / moments, bf16 params, decay mask
    DBuf scale;                   // in_mean[8], in_inv[8], out_mean[2], out_inv[2]
    float in_mean[kIn] = {0}, in_std[kIn] = {1, 1, 1, 1, 1, 1, 1, 1}, out_mean[kOut] = {0}, out_std[kOut] = {1, 1};
    int64_t adam_t = 0;
    // activations for a batch of M rows
    int64_t cap_rows = 0;
    DBuf x0, act[4], pre[4], outf, dout, dpre_b[4], dpre_f[4], dact, idx, loss, Xd, Yd;
    DBuf wfrag;  // bf16 weights in fragment order (k_mlp_infer)
};

namespace {

void mlp_layout(dcns_mlp *m) {
    int64_t o = 0;
    for (int l = 0; l < kLayers; ++l) {
        m->w_off[l] = o;
        o += (int64_t)m->out_pad[l] * m->in_pad[l];
        m->b_off[l] = o;
        o += m->out_pad[l];
    }
    m->total = o;
}

void upload_scale(dcns_mlp *m) {
    float s[2 * kIn + 2 * kOut];
    for (int i = 0; i < kIn; ++i) {
        s[i] = m->in_mean[i];
        s[kIn + i] = 1.f / m->in_std[i];
    }
    for (int i = 0; i < kOut; ++i) {
        s[2 * kIn + i] = m->out_mean[i];
        s[2 * kIn + kOut + i] = 1.f / m->out_std[i];
    }
    m->scale.ensure(sizeof(s));
    hip_check(hipMemcpyAsync(m->scale.p, s, sizeof(s), hipMemcpyHostToDevice, m->ctx->stream), "scale");
}

void ensure_rows(dcns_mlp *m, int64_t M) {
    if (M <= m->cap_rows) return;
    m->x0.ensure((size_t)M * kInPad * 2);
    for (int l = 0; l < 4; ++l) {
        m->act[l].ensure((size_t)M * kHid * 2);
        m->pre[l].ensure((size_t)M * kHid * 4);
        m->dpre_b[l].ensure((size_t)M * kHid * 2);
        m->dpre_f[l].ensure((size_t)M * kHid * 4);
    }
    m->dact.ensure((size_t)M * kHid * 4);
    m->outf.ensure((size_t)M * kOut * 4);
    m->dout.ensure((size_t)M * kOutPad * 2);
    m->idx.ensure((size_t)M * 4);
    m->loss.ensure(64);
    m->cap_rows = M;
}

template <int EPI>
void gemm(dcns_mlp *m, const uint16_t *A, int64_t sam, int64_t sak, const uint16_t *B, int64_t sbk, int64_t sbn,
          int M, int N, int K, const float *bias, const float *pre_in, float *out_f, uint16_t *out_b, int ldo,
          int n_valid) {
    dcns_ctx *c = m->ctx;
    const unsigned grid = (unsigned)((M / 32) * (N / 32));
    c->launch("mlp_gemm", [&] {
        hipLaunchKernelGGL(k_gemm<EPI>, dim3(grid), dim3(64), 0, c->stream, A, sam, sak, B, sbk, sbn, M, N, K, bias,
                           pre_in, out_f, out_b, ldo, n_valid);
        return hipGetLastError();
    });
}

// Forward pass over M padded rows (x0 already prepared).  Output rows are
// written unpadded ([M x 2]).
void forward(dcns_mlp *m, int64_t M) {
    const uint16_t *pb = m->pb.as<uint16_t>();
    const float *p = m->p.as<float>();
    const uint16_t *in = m->x0.as<uint16_t>();
    int64_t ld_in = kInPad;
    for (int l = 0; l < 4; ++l) {
        gemm<EPI_BIAS_GELU>(m, in, ld_in, 1, pb + m->w_off[l], 1, m->in_pad[l], (int)M, kHid, m->in_pad[l],
                            p + m->b_off[l], nullptr, m->pre[l].as<float>(), m->act[l].as<uint16_t>(), kHid, kHid);
        in = m->act[l].as<uint16_t>();
        ld_in = kHid;
    }
    gemm<EPI_BIAS>(m, in, kHid, 1, pb + m->w_off[4], 1, kHid, (int)M, kOutPad, kHid, p + m->b_off[4], nullptr,
                   m->outf.as<float>(), nullptr, 0, kOut);
}

// Inference over M padded rows (x0 prepared): the weights are put in
// fragment order (a few microseconds for 200 k parameters, so always current),
// then one k_mlp_infer launch writes the output rows unpadded ([M x 2]).
void infer(dcns_mlp *m, int64_t M) {
    dcns_ctx *c = m->ctx;
    MlpOffsets off{};
    int64_t fo = 0;
    for (int l = 0; l < kLayers; ++l) {
        off.frag[l] = fo;
        off.bias[l] = m->b_off[l];
        fo += (int64_t)m->in_pad[l] * m->out_pad[l];
    }
    m->wfrag.ensure((size_t)fo * 2);
    for (int l = 0; l < kLayers; ++l) {
        const int64_t n = (int64_t)m->in_pad[l] * m->out_pad[l];
        c->launch("mlp_frag", [&] {
            hipLaunchKernelGGL(k_mlp_frag, dim3(blocks(n, 256)), dim3(256), 0, c->stream, m->pb.as<uint16_t>() + m->w_off[l],
                               m->in_pad[l], m->out_pad[l], m->wfrag.as<uint16_t>() + off.frag[l]);
            return hipGetLastError();
        });
    }
    const unsigned grid = (unsigned)((M + 32 * kInferWaves - 1) / (32 * kInferWaves));
    c->launch("mlp_infer", [&] {
        hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(64 * kInferWaves), 0, c->stream, m->x0.as<uint16_t>(), M,
                           m->wfrag.as<uint16_t>(), m->p.as<float>(), off, m->outf.as<float>());
        return hipGetLastError();
    });
}

// Backward pass: gradient of the mean squared error into m->g (fp32).
void backward(dcns_mlp *m, int64_t M, int64_t n) {
    dcns_ctx *c = m->ctx;
    const uint16_t *pb = m->pb.as<uint16_t>();
    float *g = m->g.as<float>();
    // output layer: dW5t[32 x 256] = dout^T [32 x M] * act3 [M x 256]; db5 = colsum(dout)
    const uint16_t *dnext = m->dout.as<uint16_t>();
    int n_next = kOutPad;
    for (int l = 4; l >= 0; --l) {
        const uint16_t *a_in = l == 0 ? m->x0.as<uint16_t>() : m->act[l - 1].as<uint16_t>();
        const int in_pad = m->in_pad[l];
        // weight gradient: dWt(o, i) = sum_r dnext(r, o) * a_in(r, i)
        gemm<EPI_STORE>(m, dnext, 1, n_next, a_in, in_pad, 1, n_next, in_pad, (int)M, nullptr, nullptr,
                        g + m->w_off[l], nullptr, in_pad, 0);
        // bias gradient from the fp32 copy of dnext (output layer: from dout via its fp32 mirror)
        const float *dnext_f = l == 4 ? m->dact.as<float>() : m->dpre_f[l].as<float>();
        c->launch("mlp_colsum", [&] {
            hipLaunchKernelGGL(k_colsum, dim3(n_next), dim3(256), 0, c->stream, dnext_f, (int64_t)n, n_next,
                               m->out_real[l], g + m->b_off[l]);
            return hipGetLastError();
        });
        if (l == 0) break;
        // input gradient: dAct(r, i) = sum_o dnext(r, o) * Wt(o, i), then through GELU'
        gemm<EPI_GELU_BACK>(m, dnext, n_next, 1, pb + m->w_off[l], in_pad, 1, (int)M, in_pad, n_next, nullptr,
                            m->pre[l - 1].as<float>(), m->dpre_f[l - 1].as<float>(), m->dpre_b[l - 1].as<uint16_t>(),
                            kHid, 0);
        dnext = m->dpre_b[l - 1].as<uint16_t>();
        n_next = kHid;
    }
}

void prep_input(dcns_mlp *m, const float *Xd, const int32_t *idx, int64_t n, int64_t M) {
    dcns_ctx *c = m->ctx;
    const float *s = m->scale.as<float>();
    c->launch("mlp_prep", [&] {
        hipLaunchKernelGGL(k_prep_input, dim3(blocks(M * kInPad, 256)), dim3(256), 0, c->stream, Xd, idx, n, M, s,
                           s + kIn, m->x0.as<uint16_t>());
        return hipGetLastError();
    });
}

// loss + gradient of the output layer; the fp32 mirror of dout (for the
// bias gradient) goes to m->dact laid out [M x 32]
__global__ void k_dout_f32(const uint16_t *__restrict__ d, float *__restrict__ f, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = __uint_as_float((uint32_t)d[i] << 16);
}

float loss_and_grad(dcns_mlp *m, const float *Yd, const int32_t *idx, int64_t n, int64_t M, bool do_backward) {
    dcns_ctx *c = m->ctx;
    const float *s = m->scale.as<float>();
    hip_check(hipMemsetAsync(m->loss.p, 0, sizeof(float), c->stream), "hipMemsetAsync");
    c->launch("mlp_mse", [&] {
        hipLaunchKernelGGL(k_mse, dim3(blocks(M * kOutPad, 256)), dim3(256), 0, c->stream, m->outf.as<float>(), Yd,
                           idx, n, M, s + 2 * kIn, s + 2 * kIn + kOut, m->dout.as<uint16_t>(), m->loss.as<float>());
        return hipGetLastError();
    });
    if (do_backward) {
        c->launch("mlp_dout", [&] {
            hipLaunchKernelGGL(k_dout_f32, dim3(blocks(M * kOutPad, 256)), dim3(256), 0, c->stream,
                               m->dout.as<uint16_t>(), m->dact.as<float>(), M * kOutPad);
            return hipGetLastError();
        });
        backward(m, M, n);
    }
    float loss = 0;
    hip_check(hipMemcpyAsync(&loss, m->loss.p, sizeof(float), hipMemcpyDeviceToHost, c->stream), "loss");
    c->sync();
    return loss;
}

template <class F>
int mlp_guard(dcns_mlp *m, F &&f) {
    try {
        f();
        return DCNS_OK;
    } catch (const Error &e) {
        if (m && m->ctx) m->ctx->err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        if (m && m->ctx) m->ctx->err = e.what();
        return DCNS_ERR_INTERNAL;
    }
}

// logical (unpadded) parameter order: per layer W [out x in] row-major, then b [out]
template <class F>
void for_each_logical(const dcns_mlp *m, F &&f) {
    int64_t li = 0;
    for (int l = 0; l < kLayers; ++l) {
        for (int o = 0; o < m->out_real[l]; ++o)
            for (int i = 0; i < m->in_real[l]; ++i) f(li++, m->w_off[l] + (int64_t)o * m->in_pad[l] + i);
        for (int o = 0; o < m->out_real[l]; ++o) f(li++, m->b_off[l] + o);
    }
}

}  // namespace

extern "C" {

int dcns_mlp_create(dcns_ctx *ctx, uint64_t seed, dcns_mlp **out) {
    if (!ctx || !out) return DCNS_ERR_ARG;
    *out = nullptr;
    dcns_mlp *m = new (std::nothrow) dcns_mlp();
    if (!m) return DCNS_ERR_INTERNAL;
    m->ctx = ctx;
    int rc = mlp_guard(m, [&] {
        hip_check(hipSetDevice(ctx->device), "hipSetDevice");
        mlp_layout(m);
        // PyTorch Linear default init: U(-1/sqrt(in), 1/sqrt(in)) for weights and biases
        std::vector<float> h((size_t)m->total, 0.f);
        std::vector<uint8_t> mask((size_t)m->total, 0);
        std::mt19937_64 rng(seed);
        for (int l = 0; l < kLayers; ++l) {
            const float bound = 1.f / std::sqrt((float)m->in_real[l]);
            std::uniform_real_distribution<float> U(-bound, bound);
            for (int o = 0; o < m->out_real[l]; ++o) {
                for (int i = 0; i < m->in_real[l]; ++i) {
                    const int64_t k = m->w_off[l] + (int64_t)o * m->in_pad[l] + i;
                    h[(size_t)k] = U(rng);
                    mask[(size_t)k] = 1;
                }
            }
            for (int o = 0; o < m->out_real[l]; ++o) h[(size_t)(m->b_off[l] + o)] = U(rng);
        }
        const size_t n = (size_t)m->total;
        m->p.ensure(n * 4);
        m->g.ensure(n * 4);
        m->m1.ensure(n * 4);
        m->m2.ensure(n * 4);
        m->pb.ensure(n * 2);
        m->mask.ensure(n);
        hip_check(hipMemcpyAsync(m->p.p, h.data(), n * 4, hipMemcpyHostToDevice, ctx->stream), "params");
        hip_check(hipMemcpyAsync(m->mask.p, mask.data(), n, hipMemcpyHostToDevice, ctx->stream), "mask");
        hip_check(hipMemsetAsync(m->m1.p, 0, n * 4, ctx->stream), "m1");
        hip_check(hipMemsetAsync(m->m2.p, 0, n * 4, ctx->stream), "m2");
        hip_check(hipMemsetAsync(m->g.p, 0, n * 4, ctx->stream), "g");
        hipLaunchKernelGGL(k_to_bf16, dim3(blocks((int64_t)n, 256)), dim3(256), 0, ctx->stream, m->p.as<float>(),
                           m->pb.as<uint16_t>(), (int64_t)n);
        hip_check(hipGetLastError(), "k_to_bf16");
        upload_scale(m);
        ctx->sync();
    });
    if (rc != DCNS_OK) {
        delete m;
        return rc;
    }
    *out = m;
    return DCNS_OK;
}

void dcns_mlp_destroy(dcns_mlp *m) {
    if (m && m->ctx) (void)hipStreamSynchronize(m->ctx->stream);
    delete m;
}

int64_t dcns_mlp_num_params(const dcns_mlp *m) {
    if (!m) return -1;
    int64_t n = 0;
    for_each_logical(m, [&](int64_t, int64_t) { ++n; });
    return n;
}

int dcns_mlp_set_scaling(dcns_mlp *m, const float *in_mean, const float *in_std, const float *out_mean,
                         const float *out_std) {
    if (!m || !in_mean || !in_std || !out_mean || !out_std) return DCNS_ERR_ARG;
    return mlp_guard(m, [&] {
        for (int i = 0; i < kIn; ++i) {
            if (!(in_std[i] > 0)) throw Error(DCNS_ERR_ARG, "input scale must be positive");
            m->in_mean[i] = in_mean[i];
            m->in_std[i] = in_std[i];
        }
        for (int i = 0; i < kOut; ++i) {
            if (!(out_std[i] > 0)) throw Error(DCNS_ERR_ARG, "output scale must be positive");
            m->out_mean[i] = out_mean[i];
            m->out_std[i] = out_std[i];
        }
        upload_scale(m);
        m->ctx->sync();
    });
}

int dcns_mlp_get_params(const dcns_mlp *m, float *out, int64_t cap) {
    if (!m || !out || cap < dcns_mlp_num_params(m)) return DCNS_ERR_ARG;
    return mlp_guard(const_cast<dcns_mlp *>(m), [&] {
        std::vector<float> h((size_t)m->total);
        hip_check(hipMemcpyAsync(h.data(), m->p.p, h.size() * 4, hipMemcpyDeviceToHost, m->ctx->stream), "params");
        m->ctx->sync();
        for_each_logical(m, [&](int64_t li, int64_t pi) { out[li] = h[(size_t)pi]; });
    });
}

int dcns_mlp_set_params(dcns_mlp *m, const float *in, int64_t n) {
    if (!m || !in || n != dcns_mlp_num_params(m)) return DCNS_ERR_ARG;
    return mlp_guard(m, [&] {
        std::vector<float> h((size_t)m->total, 0.f);
        for_each_logical(m, [&](int64_t li, int64_t pi) { h[(size_t)pi] = in[li]; });
        hip_check(hipMemcpyAsync(m->p.p, h.data(), h.size() * 4, hipMemcpyHostToDevice, m->ctx->stream), "params");
        hipLaunchKernelGGL(k_to_bf16, dim3(blocks(m->total, 256)), dim3(256), 0, m->ctx->stream, m->p.as<float>(),
                           m->pb.as<uint16_t>(), m->total);
        hip_check(hipGetLastError(), "k_to_bf16");
        m->ctx->sync();
    });
}

// Predictions (log lambda, log K) for n raw input rows [n x 8] -> [n x 2].
int dcns_mlp_predict(dcns_mlp *m, const float *X, int64_t n, float *Y) {
    if (!m || (n > 0 && (!X || !Y)) || n < 0) return DCNS_ERR_ARG;
    return mlp_guard(m, [&] {
        if (n == 0) return;
        dcns_ctx *c = m->ctx;
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        const int64_t chunk = 65536;
        std::vector<float> o((size_t)chunk * kOut);
        for (int64_t s = 0; s < n; s += chunk) {
            const int64_t k = std::min(chunk, n - s), M = (k + 31) / 32 * 32;
            ensure_rows(m, M);
            m->Xd.ensure((size_t)k * kIn * 4);
            hip_check(hipMemcpyAsync(m->Xd.p, X + s * kIn, (size_t)k * kIn * 4, hipMemcpyHostToDevice, c->stream), "X");
            prep_input(m, m->Xd.as<float>(), nullptr, k, M);
            infer(m, M);
            hip_check(hipMemcpyAsync(o.data(), m->outf.p, (size_t)k * kOut * 4, hipMemcpyDeviceToHost, c->stream), "Y");
            c->sync();
            for (int64_t r = 0; r < k; ++r)
                for (int j = 0; j < kOut; ++j)
                    Y[(s + r) * kOut + j] = o[(size_t)(r * kOut + j)] * m->out_std[j] + m->out_mean[j];
        }
    });
}

// One forward + backward pass over n rows (no update): mean squared error of
// the standardised outputs and its gradient in the logical parameter order.
int dcns_mlp_loss_grad(dcns_mlp *m, const float *X, const float *Y, int64_t n, float *loss, float *grad,
                       int64_t grad_cap) {
    if (!m || !X || !Y || n <= 0 || (grad && grad_cap < dcns_mlp_num_params(m))) return DCNS_ERR_ARG;
    return mlp_guard(m, [&] {
        dcns_ctx *c = m->ctx;
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        const int64_t M = (n + 31) / 32 * 32;
        ensure_rows(m, M);
        m->Xd.ensure((size_t)n * kIn * 4);
        m->Yd.ensure((size_t)n * kOut * 4);
        hip_check(hipMemcpyAsync(m->Xd.p, X, (size_t)n * kIn * 4, hipMemcpyHostToDevice, c->stream), "X");
        hip_check(hipMemcpyAsync(m->Yd.p, Y, (size_t)n * kOut * 4, hipMemcpyHostToDevice, c->stream), "Y");
        prep_input(m, m->Xd.as<float>(), nullptr, n, M);
        forward(m, M);
        const float l = loss_and_grad(m, m->Yd.as<float>(), nullptr, n, M, true);
        if (loss) *loss = l;
        if (grad) {
            std::vector<float> h((size_t)m->total);
            hip_check(hipMemcpyAsync(h.data(), m->g.p, h.size() * 4, hipMemcpyDeviceToHost, c->stream), "grad");
            c->sync();
            for_each_logical(m, [&](int64_t li, int64_t pi) { grad[li] = h[(size_t)pi]; });
        }
    });
}

// AdamW training on n rows: `steps` minibatches of `batch` rows drawn with
// replacement by a seeded generator.  losses (optional) receives each step's
// minibatch loss.
int dcns_mlp_train(dcns_mlp *m, const float *X, const float *Y, int64_t n, int32_t steps, int32_t batch, float lr,
                   float weight_decay, uint64_t seed, float *losses) {
    if (!m || !X || !Y || n <= 0 || steps < 0 || batch <= 0) return DCNS_ERR_ARG;
    return mlp_guard(m, [&] {
        dcns_ctx *c = m->ctx;
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        const int64_t M = (std::min<int64_t>(batch, n) + 31) / 32 * 32;
        ensure_rows(m, M);
        m->Xd.ensure((size_t)n * kIn * 4);
        m->Yd.ensure((size_t)n * kOut * 4);
        hip_check(hipMemcpyAsync(m->Xd.p, X, (size_t)n * kIn * 4, hipMemcpyHostToDevice, c->stream), "X");
        hip_check(hipMemcpyAsync(m->Yd.p, Y, (size_t)n * kOut * 4, hipMemcpyHostToDevice, c->stream), "Y");
        std::mt19937_64 rng(seed);
        std::uniform_int_distribution<int64_t> pick(0, n - 1);
        std::vector<int32_t> idx((size_t)M, 0);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        // batch >= n: every step is one full-batch step over the rows in order
        const bool full = batch >= n;
        const int64_t rows = full ? n : batch;
        for (int32_t s = 0; s < steps; ++s) {
            const int32_t *d_idx = nullptr;
            if (!full) {
                for (int64_t r = 0; r < batch; ++r) idx[(size_t)r] = (int32_t)pick(rng);
                hip_check(hipMemcpyAsync(m->idx.p, idx.data(), (size_t)batch * 4, hipMemcpyHostToDevice, c->stream),
                          "idx");
                d_idx = m->idx.as<int32_t>();
            }
            prep_input(m, m->Xd.as<float>(), d_idx, rows, M);
            forward(m, M);
            const float l = loss_and_grad(m, m->Yd.as<float>(), d_idx, rows, M, true);
            if (losses) losses[s] = l;
            ++m->adam_t;
            const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
            c->launch("mlp_adamw", [&] {
                hipLaunchKernelGGL(k_adamw, dim3(blocks(m->total, 256)), dim3(256), 0, c->stream, m->p.as<float>(),
                                   m->g.as<float>(), m->m1.as<float>(), m->m2.as<float>(), m->pb.as<uint16_t>(),
                                   m->mask.as<uint8_t>(), m->total, lr, b1, b2, eps, weight_decay, c1, c2);
                return hipGetLastError();
            });
        }
        c->sync();
    });
}

}  // extern "C"

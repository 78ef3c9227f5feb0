// MLP of the learned warm start: 4 dense layers (n_in -> W -> W -> W -> n_out, tanh between),
// trained and evaluated with hand-written BF16 MFMA kernels (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation) on gfx950.  Master weights, Adam state, biases, losses and gradient
// reductions are FP32; activations and GEMM operands BF16.
//
// Layout: activations row-major [rows][cols] BF16; each layer keeps its weights twice in
// BF16, W^T [out][in] for the forward product and W [in][out] for the input-gradient product,
// so every MFMA operand fragment is one contiguous 16-byte load.  The weight gradient
// X^T * dZ reduces over rows; its tiles are transposed through LDS.  Every reduction runs in
// a fixed order, so training is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "device.hpp"
#include "fvm_kernels.hpp"

namespace gr {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 64;     // padded input width (n_in <= 64)
constexpr int kMlpOut = 32;    // padded output width (n_out <= 32)
constexpr int kMlpW = 256;     // hidden width
constexpr int kKC = 64;        // K chunk of the dense kernel staged in LDS
constexpr int kRowsPerBlock = 128;
constexpr int kWgradChunks = 64;   // row chunks (blocks along rows) of the weight-gradient kernels

__device__ inline bf16x8 zero8() {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = (bf16)0.0f;
    return z;
}

// Y[M x N] = act(X[M x K] * B + bias), B given as BT = B^T [N x K] row-major.  ACT 1: tanh.
// K % 32 == 0, N == 16*NT.  Block: 4 waves x 32 rows; each wave 2 row groups x NT column
// tiles of 16x16 MFMA accumulators.  In the 16x16x32 MFMA, lane l holds row/column (l mod 16)
// of its operand and eight consecutive k of group l/16; the accumulator element j of lane l
// is row 4*(l/16)+j, column l mod 16.
template <int ACT, typename OUT, int NT>
__global__ __launch_bounds__(kBlock) void k_dense(const bf16 *__restrict__ X, int ldx, int M, int K,
                                                  const bf16 *__restrict__ BT, const float *__restrict__ bias,
                                                  OUT *__restrict__ Y, int ldy) {
    constexpr int N = 16 * NT;
    __shared__ bf16 sB[N][kKC + 8];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int row0 = blockIdx.x * kRowsPerBlock + w * 32;
    const int lr = lane & 15, lg = lane >> 4;
    f32x4 acc[2][NT];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int kc = 0; kc < K; kc += kKC) {
        const int kn = min(kKC, K - kc);
        __syncthreads();
        for (int v = threadIdx.x; v < N * (kn / 8); v += blockDim.x) {
            const int n = v / (kn / 8), c8 = v - n * (kn / 8);
            *(bf16x8 *)&sB[n][c8 * 8] = *(const bf16x8 *)&BT[(size_t)n * K + kc + c8 * 8];
        }
        __syncthreads();
        for (int ks = 0; ks < kn; ks += 32) {
            bf16x8 a[2];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int row = row0 + r * 16 + lr;
                a[r] = row < M ? *(const bf16x8 *)&X[(size_t)row * ldx + kc + ks + 8 * lg] : zero8();
            }
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const bf16x8 b = *(const bf16x8 *)&sB[t * 16 + lr][ks + 8 * lg];
                acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b, acc[0][t], 0, 0, 0);
                acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b, acc[1][t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int col = t * 16 + lr;
            const float bb = bias ? bias[col] : 0.f;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = row0 + r * 16 + 4 * lg + j;
                if (row < M) {
                    float v = acc[r][t][j] + bb;
                    if (ACT == 1) v = tanhf(v);
                    Y[(size_t)row * ldy + col] = (OUT)v;
                }
            }
        }
}

// Partial weight gradient P[chunk][K][N] = X[rows of chunk]^T * D[rows of chunk] with
// X [M x K], D [M x N] (N == 16*NT) BF16; block (kt, chunk) computes rows kt*64..kt*64+63 of
// the K dimension.  Tiles of 64 rows are written transposed to LDS so that the MFMA
// fragments (eight consecutive rows) are contiguous.
template <int NT>
__global__ __launch_bounds__(kBlock) void k_wgrad(const bf16 *__restrict__ X, int ldx, const bf16 *__restrict__ D,
                                                  int ldd, int M, int K, int per, float *__restrict__ P) {
    constexpr int N = 16 * NT;
    __shared__ bf16 sX[64][64 + 8];   // [k_in local][row local]
    __shared__ bf16 sD[N][64 + 8];    // [n][row local]
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const int kt = blockIdx.x, chunk = blockIdx.y;
    const int r_begin = chunk * per, r_end = min(M, r_begin + per);
    f32x4 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int r0 = r_begin; r0 < r_end; r0 += 64) {
        __syncthreads();
        // X rows r0..r0+63, columns kt*64..+63 -> sX transposed (zeros past M)
        for (int v = threadIdx.x; v < 64 * 8; v += blockDim.x) {
            const int rr = v >> 3, c8 = v & 7;
            const int row = r0 + rr;
            bf16x8 x = row < r_end ? *(const bf16x8 *)&X[(size_t)row * ldx + kt * 64 + c8 * 8] : zero8();
#pragma unroll
            for (int j = 0; j < 8; ++j) sX[c8 * 8 + j][rr] = x[j];
        }
        for (int v = threadIdx.x; v < 64 * (N / 8); v += blockDim.x) {
            const int rr = v / (N / 8), c8 = v - rr * (N / 8);
            const int row = r0 + rr;
            bf16x8 d = row < r_end ? *(const bf16x8 *)&D[(size_t)row * ldd + c8 * 8] : zero8();
#pragma unroll
            for (int j = 0; j < 8; ++j) sD[c8 * 8 + j][rr] = d[j];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < 64; ks += 32) {
            const bf16x8 a = *(const bf16x8 *)&sX[w * 16 + lr][ks + 8 * lg];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const bf16x8 b = *(const bf16x8 *)&sD[t * 16 + lr][ks + 8 * lg];
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
            }
        }
    }
    float *Pc = P + (size_t)chunk * K * N;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = kt * 64 + w * 16 + 4 * lg + j;
            Pc[(size_t)k * N + t * 16 + lr] = acc[t][j];
        }
}

// out[e] = sum over chunks c (in order) of P[c][e]
__global__ void k_sum_chunks(const float *__restrict__ P, int nchunks, int n, float *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float s = 0.f;
    for (int c = 0; c < nchunks; ++c) s += P[(size_t)c * n + e];
    out[e] = s;
}

// Column sums of an FP32 [M x N] array over row chunks of `per` rows: P[chunk][N].  A block
// covers 64 columns (one per lane) of one chunk; its 4 waves take every fourth row and their sums
// are added in wave order through LDS (fixed order: deterministic).  1024+ blocks for the
// bias gradients of a 32,768-row batch, where one thread per column over a 512-row chunk (64
// blocks) took 123 us per launch, latency-bound.
__global__ __launch_bounds__(kBlock) void k_colsum(const float *__restrict__ Z, int ldz, int M, int N, int per,
                                                   float *__restrict__ P) {
    __shared__ float s_w[kBlock / 64][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + lane, chunk = blockIdx.y;
    const int r0 = chunk * per, r1 = min(M, r0 + per);
    float s = 0.f;
    if (n < N)
        for (int r = r0 + w; r < r1; r += kBlock / 64) s += Z[(size_t)r * ldz + n];
    s_w[w][lane] = s;
    __syncthreads();
    if (w == 0 && n < N) {
        float t = s_w[0][lane];
        for (int k = 1; k < kBlock / 64; ++k) t += s_w[k][lane];
        P[(size_t)chunk * N + n] = t;
    }
}

// FP32 host input [M x n_in] -> BF16 [M x kMlpIn], zero padded.
__global__ void k_pack_input(const float *__restrict__ x, int M, int n_in, bf16 *__restrict__ X) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)M * kMlpIn) return;
    const int r = (int)(e / kMlpIn), c = (int)(e % kMlpIn);
    X[e] = (bf16)(c < n_in ? x[(size_t)r * n_in + c] : 0.f);
}

// Mini-batch of a resident training set: row r of the batch is row (a*(pos0+r) + b) mod Mall of
// the set (a coprime to Mall, so one pass of pos over 0..Mall-1 visits every row once).
__global__ void k_gather_rows(const float *__restrict__ xall, const float *__restrict__ tall, uint64_t Mall, int n_in,
                              int n_out, uint64_t a, uint64_t b, uint64_t pos0, int M, bf16 *__restrict__ X,
                              float *__restrict__ T) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)M * kMlpIn) return;
    const int r = (int)(e / kMlpIn), c = (int)(e % kMlpIn);
    const uint64_t src = (a * ((pos0 + r) % Mall) + b) % Mall;
    X[e] = (bf16)(c < n_in ? xall[src * n_in + c] : 0.f);
    if (c < n_out) T[(size_t)r * n_out + c] = tall[src * n_out + c];
}

// Mean squared error: dOut = 2 (out - t) / (M n_out) (BF16 for the GEMMs, FP32 for the bias
// gradient), padded columns 0; per-block partial sums of the squared error.
__global__ __launch_bounds__(kBlock) void k_mse(const float *__restrict__ out, const float *__restrict__ t, int M,
                                                int n_out, bf16 *__restrict__ d16, float *__restrict__ d32,
                                                double *__restrict__ partials) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double se = 0;
    if (e < (size_t)M * kMlpOut) {
        const int r = (int)(e / kMlpOut), c = (int)(e % kMlpOut);
        float g = 0.f;
        if (c < n_out) {
            const float diff = out[e] - t[(size_t)r * n_out + c];
            se = (double)diff * (double)diff;
            g = 2.f * diff / ((float)M * (float)n_out);
        }
        d16[e] = (bf16)g;
        d32[e] = g;
    }
    const double bs = block_sum(se);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

// dst = (sum of the n partials, in a fixed order) / denom
__global__ __launch_bounds__(kBlock) void k_loss_final(const double *__restrict__ partials, int n, double denom,
                                                       double *__restrict__ dst) {
    double s = 0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) s += partials[i];
    const double t = block_sum(s);
    if (threadIdx.x == 0) *dst = t / denom;
}

// dZ = dH * (1 - H^2) for tanh layers (H the layer's BF16 output).
__global__ void k_dtanh(const float *__restrict__ dH, const bf16 *__restrict__ H, size_t n, bf16 *__restrict__ d16,
                        float *__restrict__ d32) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float h = (float)H[e];
    const float g = dH[e] * (1.f - h * h);
    d16[e] = (bf16)g;
    d32[e] = g;
}

// Adam on the FP32 master parameters.
__global__ void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, size_t n, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.f - b1) * gg;
    const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    p[e] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// BF16 copies of layer weights W [in][out] (FP32, padded to Kp x Np): WT [Np][Kp] and W [Kp][Np].
__global__ void k_pack_weights(const float *__restrict__ W, int Kp, int Np, bf16 *__restrict__ WT,
                               bf16 *__restrict__ Wb) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= Kp * Np) return;
    const int k = e / Np, n = e - k * Np;
    const bf16 v = (bf16)W[e];
    Wb[e] = v;
    WT[(size_t)n * Kp + k] = v;
}

constexpr int kMlpChunk = 1 << 18;   // rows per forward chunk of device-side evaluation

}  // namespace gr

using gr::guarded;
using gr::require;

struct gr_mlp {
    uint32_t n_in = 0, n_out = 0;
    // padded layer shapes: (in, out) = (64,256), (256,256), (256,256), (256,32)
    int Kp[4] = {gr::kMlpIn, gr::kMlpW, gr::kMlpW, gr::kMlpW};
    int Np[4] = {gr::kMlpW, gr::kMlpW, gr::kMlpW, gr::kMlpOut};
    size_t woff[4], boff[4], n_params = 0;   // offsets into the padded FP32 parameter vector
    hipStream_t stream = nullptr;
    gr::DevArray<float> p, m, v, g;
    gr::DevArray<gr::bf16> WT[4], Wb[4];
    int cap = 0;   // rows of the activation buffers
    gr::DevArray<float> xin, tgt, out, d32, dH, part;
    gr::DevArray<gr::bf16> X0, H[3], d16;
    gr::DevArray<double> lpart, loss;
    gr::DevArray<float> xall, tall;   // resident training set
    uint64_t n_rows = 0;
    long step = 0;

    ~gr_mlp() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
    }
    void ensure(int M) {
        if (M <= cap) return;
        cap = ((M + 4095) / 4096) * 4096;
        xin.alloc((size_t)cap * gr::kMlpIn);
        tgt.alloc((size_t)cap * gr::kMlpOut);
        out.alloc((size_t)cap * gr::kMlpOut);
        X0.alloc((size_t)cap * gr::kMlpIn);
        for (auto &h : H) h.alloc((size_t)cap * gr::kMlpW);
        d16.alloc((size_t)cap * gr::kMlpW);
        d32.alloc((size_t)cap * gr::kMlpW);
        dH.alloc((size_t)cap * gr::kMlpW);
        part.alloc((size_t)gr::kWgradChunks * gr::kMlpW * gr::kMlpW);
        lpart.alloc((size_t)cap * gr::kMlpOut / gr::kBlock + 1);
    }
    void pack(hipStream_t s) {
        for (int l = 0; l < 4; ++l)
            hipLaunchKernelGGL(gr::k_pack_weights, dim3(gr::blocks_for((size_t)Kp[l] * Np[l])), dim3(gr::kBlock), 0, s,
                               p.p + woff[l], Kp[l], Np[l], WT[l].p, Wb[l].p);
    }
    void upload_x(const float *x, int M) {
        GR_HIP(hipMemcpyAsync(xin.p, x, (size_t)M * n_in * sizeof(float), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(gr::k_pack_input, dim3(gr::blocks_for((size_t)M * gr::kMlpIn)), dim3(gr::kBlock), 0, stream,
                           xin.p, M, (int)n_in, X0.p);
    }
    void forward(int M, hipStream_t s) {
        const dim3 grid((M + gr::kRowsPerBlock - 1) / gr::kRowsPerBlock), blk(gr::kBlock);
        hipLaunchKernelGGL((gr::k_dense<1, gr::bf16, 16>), grid, blk, 0, s, X0.p, gr::kMlpIn, M, Kp[0], WT[0].p,
                           p.p + boff[0], H[0].p, gr::kMlpW);
        hipLaunchKernelGGL((gr::k_dense<1, gr::bf16, 16>), grid, blk, 0, s, H[0].p, gr::kMlpW, M, Kp[1], WT[1].p,
                           p.p + boff[1], H[1].p, gr::kMlpW);
        hipLaunchKernelGGL((gr::k_dense<1, gr::bf16, 16>), grid, blk, 0, s, H[1].p, gr::kMlpW, M, Kp[2], WT[2].p,
                           p.p + boff[2], H[2].p, gr::kMlpW);
        hipLaunchKernelGGL((gr::k_dense<0, float, 2>), grid, blk, 0, s, H[2].p, gr::kMlpW, M, Kp[3], WT[3].p,
                           p.p + boff[3], out.p, gr::kMlpOut);
    }
    // weight and bias gradients of layer l from its input A (BF16, lda) and d16/d32 (ld = Np)
    template <int NT>
    void wgrad(int l, const gr::bf16 *A, int lda, int M) {
        // row chunks of a multiple of 64 rows, at most kWgradChunks of them
        const int want = std::min(gr::kWgradChunks, (M + 63) / 64);
        const int per = (((M + want - 1) / want) + 63) / 64 * 64;
        const int chunks = (M + per - 1) / per;
        hipLaunchKernelGGL((gr::k_wgrad<NT>), dim3(Kp[l] / 64, chunks), dim3(gr::kBlock), 0, stream, A, lda, d16.p,
                           Np[l], M, Kp[l], per, part.p);
        const int n = Kp[l] * Np[l];
        hipLaunchKernelGGL(gr::k_sum_chunks, dim3(gr::blocks_for(n)), dim3(gr::kBlock), 0, stream, part.p, chunks, n,
                           g.p + woff[l]);
        // bias: column sums of the FP32 output gradient over chunks of 32 rows (more for batches
        // above 2^19 rows, so that the chunk partials fit the partial buffer)
        const int max_chunks = gr::kWgradChunks * gr::kMlpW * gr::kMlpW / gr::kMlpW;
        const int bper = std::max(32, (M + max_chunks - 1) / max_chunks), bchunks = (M + bper - 1) / bper;
        hipLaunchKernelGGL(gr::k_colsum, dim3((Np[l] + 63) / 64, bchunks), dim3(gr::kBlock), 0, stream, d32.p, Np[l], M,
                           Np[l], bper, part.p);
        hipLaunchKernelGGL(gr::k_colsum, dim3((Np[l] + 63) / 64, 1), dim3(gr::kBlock), 0, stream, (const float *)part.p,
                           Np[l], bchunks, Np[l], bchunks, g.p + boff[l]);
    }
    // loss (into *loss_dst on the device) and gradients for the targets in tgt; no host sync
    void backward(int M, double *loss_dst) {
        const size_t no = (size_t)M * gr::kMlpOut;
        const unsigned nb = gr::blocks_for(no);
        hipLaunchKernelGGL(gr::k_mse, dim3(nb), dim3(gr::kBlock), 0, stream, out.p, tgt.p, M, (int)n_out, d16.p, d32.p,
                           lpart.p);
        hipLaunchKernelGGL(gr::k_loss_final, dim3(1), dim3(gr::kBlock), 0, stream, lpart.p, (int)nb,
                           (double)M * n_out, loss_dst);
        const dim3 grid((M + gr::kRowsPerBlock - 1) / gr::kRowsPerBlock), blk(gr::kBlock);
        // layer 4 (linear): dW4 = H3^T dOut, dH3 = dOut W4^T
        wgrad<2>(3, H[2].p, gr::kMlpW, M);
        hipLaunchKernelGGL((gr::k_dense<0, float, 16>), grid, blk, 0, stream, d16.p, gr::kMlpOut, M, Np[3], Wb[3].p,
                           (const float *)nullptr, dH.p, gr::kMlpW);
        for (int l = 2; l >= 0; --l) {
            const size_t nh = (size_t)M * gr::kMlpW;
            hipLaunchKernelGGL(gr::k_dtanh, dim3(gr::blocks_for(nh)), dim3(gr::kBlock), 0, stream, dH.p, H[l].p, nh,
                               d16.p, d32.p);
            const gr::bf16 *A = l == 0 ? X0.p : H[l - 1].p;
            const int lda = l == 0 ? gr::kMlpIn : gr::kMlpW;
            wgrad<16>(l, A, lda, M);
            if (l > 0)
                hipLaunchKernelGGL((gr::k_dense<0, float, 16>), grid, blk, 0, stream, d16.p, gr::kMlpW, M, Np[l],
                                   Wb[l].p, (const float *)nullptr, dH.p, gr::kMlpW);
        }
        GR_HIP(hipGetLastError());
    }
    void adam(float lr) {
        step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
        hipLaunchKernelGGL(gr::k_adam, dim3(gr::blocks_for(n_params)), dim3(gr::kBlock), 0, stream, p.p, m.p, v.p, g.p,
                           n_params, lr, b1, b2, 1e-8f, c1, c2);
        pack(stream);
    }
    double host_loss() {
        double l = 0;
        GR_HIP(hipMemcpyAsync(&l, loss.p, sizeof l, hipMemcpyDeviceToHost, stream));
        GR_HIP(hipStreamSynchronize(stream));
        return l;
    }
};

namespace gr {
uint32_t mlp_n_in(const gr_mlp *n) { return n->n_in; }
uint32_t mlp_n_out(const gr_mlp *n) { return n->n_out; }

void mlp_run_device(gr_mlp *n, size_t M, hipStream_t s, const std::function<void(void *, size_t, int)> &fill,
                    const std::function<void(const float *, size_t, int)> &consume) {
    GR_HIP(hipStreamSynchronize(n->stream));
    n->ensure((int)std::min<size_t>(M, kMlpChunk));
    for (size_t r0 = 0; r0 < M; r0 += kMlpChunk) {
        const int m = (int)std::min<size_t>(kMlpChunk, M - r0);
        fill(n->X0.p, r0, m);
        n->forward(m, s);
        consume(n->out.p, r0, m);
        GR_HIP(hipGetLastError());
    }
}
}  // namespace gr

namespace {
void check_dims(gr_mlp *n, const float *x, int M) {
    require(n && x, "null argument");
    require(M > 0 && M <= (1 << 24), "row count must be in 1..2^24");
}

uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}
}  // namespace

extern "C" {

int gr_mlp_create(uint32_t n_in, uint32_t n_out, uint64_t seed, gr_mlp **out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        require(n_in >= 1 && n_in <= (uint32_t)gr::kMlpIn, "n_in must be 1..64");
        require(n_out >= 1 && n_out <= (uint32_t)gr::kMlpOut, "n_out must be 1..32");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw gr::Error(GR_ERR_DEVICE, "no HIP device available");
        auto *n = new gr_mlp();
        try {
            n->n_in = n_in;
            n->n_out = n_out;
            GR_HIP(hipStreamCreate(&n->stream));
            size_t off = 0;
            for (int l = 0; l < 4; ++l) {
                n->woff[l] = off;
                off += (size_t)n->Kp[l] * n->Np[l];
                n->boff[l] = off;
                off += n->Np[l];
            }
            n->n_params = off;
            // Xavier-uniform weights on the real (unpadded) block, zero elsewhere
            std::vector<float> h(off, 0.f);
            std::mt19937_64 rng(seed);
            const int real_in[4] = {(int)n_in, gr::kMlpW, gr::kMlpW, gr::kMlpW};
            const int real_out[4] = {gr::kMlpW, gr::kMlpW, gr::kMlpW, (int)n_out};
            for (int l = 0; l < 4; ++l) {
                const double a = std::sqrt(6.0 / (real_in[l] + real_out[l]));
                for (int k = 0; k < real_in[l]; ++k)
                    for (int c = 0; c < real_out[l]; ++c) {
                        const double u = (double)(rng() >> 11) * (1.0 / 9007199254740992.0);
                        h[n->woff[l] + (size_t)k * n->Np[l] + c] = (float)((2 * u - 1) * a);
                    }
            }
            for (auto *arr : {&n->p, &n->m, &n->v, &n->g}) arr->alloc(off);
            n->p.upload(h.data(), off, n->stream);
            GR_HIP(hipMemsetAsync(n->m.p, 0, off * sizeof(float), n->stream));
            GR_HIP(hipMemsetAsync(n->v.p, 0, off * sizeof(float), n->stream));
            GR_HIP(hipMemsetAsync(n->g.p, 0, off * sizeof(float), n->stream));
            for (int l = 0; l < 4; ++l) {
                n->WT[l].alloc((size_t)n->Kp[l] * n->Np[l]);
                n->Wb[l].alloc((size_t)n->Kp[l] * n->Np[l]);
            }
            n->loss.alloc(1);
            n->pack(n->stream);
            GR_HIP(hipStreamSynchronize(n->stream));
        } catch (...) {
            delete n;
            throw;
        }
        *out = n;
    });
}

void gr_mlp_free(gr_mlp *n) { delete n; }

int gr_mlp_n_params(const gr_mlp *n, size_t *count) {
    return guarded([&] {
        require(n && count, "null argument");
        *count = n->n_params;
    });
}

// Parameters in the padded layout: per layer W [in_pad][out_pad] then b [out_pad], with
// (in_pad, out_pad) = (64,256), (256,256), (256,256), (256,32).
int gr_mlp_get_params(const gr_mlp *n, float *out) {
    return guarded([&] {
        require(n && out, "null argument");
        n->p.download(out, n->n_params, n->stream);
        GR_HIP(hipStreamSynchronize(n->stream));
    });
}

int gr_mlp_set_params(gr_mlp *n, const float *in) {
    return guarded([&] {
        require(n && in, "null argument");
        n->p.upload(in, n->n_params, n->stream);
        n->pack(n->stream);
        GR_HIP(hipStreamSynchronize(n->stream));
    });
}

int gr_mlp_forward(gr_mlp *n, const float *x, int M, float *y) {
    return guarded([&] {
        check_dims(n, x, M);
        require(y != nullptr, "null output");
        std::vector<float> o;
        for (int r0 = 0; r0 < M; r0 += gr::kMlpChunk) {
            const int m = std::min(gr::kMlpChunk, M - r0);
            n->ensure(m);
            n->upload_x(x + (size_t)r0 * n->n_in, m);
            n->forward(m, n->stream);
            GR_HIP(hipGetLastError());
            o.resize((size_t)m * gr::kMlpOut);
            n->out.download(o.data(), o.size(), n->stream);
            GR_HIP(hipStreamSynchronize(n->stream));
            for (int r = 0; r < m; ++r)
                for (uint32_t c = 0; c < n->n_out; ++c) y[(size_t)(r0 + r) * n->n_out + c] = o[(size_t)r * gr::kMlpOut + c];
        }
    });
}

int gr_mlp_loss_grad(gr_mlp *n, const float *x, const float *t, int M, double *loss, float *grad) {
    return guarded([&] {
        check_dims(n, x, M);
        require(t && loss, "null argument");
        require(M <= (1 << 20), "at most 2^20 rows per gradient step");
        n->ensure(M);
        n->upload_x(x, M);
        GR_HIP(hipMemcpyAsync(n->tgt.p, t, (size_t)M * n->n_out * sizeof(float), hipMemcpyHostToDevice, n->stream));
        n->forward(M, n->stream);
        n->backward(M, n->loss.p);
        *loss = n->host_loss();
        if (grad) {
            n->g.download(grad, n->n_params, n->stream);
            GR_HIP(hipStreamSynchronize(n->stream));
        }
    });
}

int gr_mlp_train_step(gr_mlp *n, const float *x, const float *t, int M, float lr, double *loss) {
    return guarded([&] {
        check_dims(n, x, M);
        require(t && loss, "null argument");
        require(M <= (1 << 20), "at most 2^20 rows per training step");
        require(lr > 0, "learning rate must be positive");
        n->ensure(M);
        n->upload_x(x, M);
        GR_HIP(hipMemcpyAsync(n->tgt.p, t, (size_t)M * n->n_out * sizeof(float), hipMemcpyHostToDevice, n->stream));
        n->forward(M, n->stream);
        n->backward(M, n->loss.p);
        n->adam(lr);
        GR_HIP(hipGetLastError());
        *loss = n->host_loss();
    });
}

int gr_mlp_set_data(gr_mlp *n, const float *x, const float *t, uint64_t rows) {
    return guarded([&] {
        require(n && x && t, "null argument");
        require(rows >= 1 && rows <= (1ull << 31), "row count must be in 1..2^31");
        n->xall.alloc(rows * n->n_in);
        n->tall.alloc(rows * n->n_out);
        n->xall.upload(x, rows * n->n_in, n->stream);
        n->tall.upload(t, rows * n->n_out, n->stream);
        GR_HIP(hipStreamSynchronize(n->stream));
        n->n_rows = rows;
    });
}

int gr_mlp_train(gr_mlp *n, int batch, int steps, float lr, float lr_final, uint64_t seed, double *losses) {
    return guarded([&] {
        require(n != nullptr, "null argument");
        require(n->n_rows > 0, "no training set (gr_mlp_set_data)");
        require(batch >= 1 && batch <= (1 << 20) && (uint64_t)batch <= n->n_rows, "batch must be in 1..min(2^20, rows)");
        require(steps >= 1, "steps must be positive");
        require(lr > 0 && lr_final > 0, "learning rates must be positive");
        n->ensure(batch);
        gr::DevArray<double> ldev;
        ldev.alloc((size_t)steps);
        const uint64_t Mall = n->n_rows, per_epoch = Mall / (uint64_t)batch;
        uint64_t a = 1, b = 0;
        for (int s = 0; s < steps; ++s) {
            const uint64_t epoch = (uint64_t)s / per_epoch, within = (uint64_t)s % per_epoch;
            if (within == 0) {
                std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ull + epoch);
                do {
                    a = Mall > 1 ? 1 + rng() % (Mall - 1) : 1;
                } while (gcd64(a, Mall) != 1);
                b = rng() % Mall;
            }
            hipLaunchKernelGGL(gr::k_gather_rows, dim3(gr::blocks_for((size_t)batch * gr::kMlpIn)), dim3(gr::kBlock), 0,
                               n->stream, n->xall.p, n->tall.p, Mall, (int)n->n_in, (int)n->n_out, a, b,
                               within * (uint64_t)batch, batch, n->X0.p, n->tgt.p);
            n->forward(batch, n->stream);
            n->backward(batch, ldev.p + s);
            const double cosf_ = 0.5 * (1.0 + std::cos(M_PI * (double)s / (double)steps));
            n->adam((float)(lr_final + (lr - lr_final) * cosf_));
        }
        GR_HIP(hipGetLastError());
        if (losses) ldev.download(losses, (size_t)steps, n->stream);
        GR_HIP(hipStreamSynchronize(n->stream));
    });
}

}  // extern "C"

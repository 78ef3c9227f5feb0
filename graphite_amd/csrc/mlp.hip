// MLP of the learned warm start: 4 dense layers (n_in -> W -> W -> W -> n_out, tanh between),
// trained and evaluated with hand-written BF16 MFMA kernels (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation) on gfx950.  Master weights, Adam state, biases, losses and gradient
// reductions are FP32; activations and GEMM operands BF16.
//
// Layout: activations row-major [rows][cols] BF16; each layer keeps its weights twice in
// BF16, W^T [out][in] for the forward product and W [in][out] for the input-gradient product,
// so every MFMA operand fragment is one contiguous 16-byte load.  The weight gradient
// X^T * dZ reduces over rows: in training, activations and output gradients are also written
// in strips of eight rows per column (strip_index), where its fragments are contiguous too.
// The tanh derivative and the bias gradient's first stage run in the input-gradient product's
// epilogue.  Every reduction runs in a fixed order, so training is deterministic.
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
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 64;     // padded input width (n_in <= 64)
constexpr int kMlpOut = 32;    // padded output width (n_out <= 32)
constexpr int kMlpW = 256;     // hidden width
constexpr int kWgradChunks = 32;   // most row chunks (blocks along rows) of a weight-gradient product

// Strip layout of an activation or gradient matrix [M x C] (the weight-gradient operands): the
// eight rows 8b..8b+7 of column c are contiguous, at ((b * C + c) * 8 + r % 8).  A weight
// gradient reduces over rows, and the MFMA takes eight consecutive k per lane, so each of its
// operand fragments is one 16-byte load from this layout, with no transposition.
__device__ inline size_t strip_index(int r, int c, int C) { return ((size_t)(r >> 3) * C + c) * 8 + (r & 7); }

enum { kEpiTanh = 0, kEpiLinear = 1, kEpiDtanh = 2, kEpiMse = 3 };

// Targets and outputs of the output layer's training epilogue (kEpiMse of k_gemm).
struct MseArgs {
    const float *t;   // targets [M][n_out]
    int n_out;
    float *d32;       // dOut, FP32 row-major [M][kMlpOut]
    double *lpart;    // per-block partial sums of the squared error
};

// tanh for values about to be rounded to BF16, within about 3 ulp of binary32 (a correctly
// rounded tanh would still differ from it in the BF16 result of a few values in a million, as any
// binary32 tanh does): below |x| = 0.55 the odd Taylor series of tanh to x^15 (Horner in x^2;
// the next term is under 5e-8 |x| there), above it (1 - e) / (1 + e) with e = exp(-2|x|) <= 0.33
// (hardware exp2 and reciprocal), where 1 - e loses no digits.
__device__ inline float tanh_bf16(float x) {
    const float a = fabsf(x), s = a * a;
    float p = -929569.f / 638512875.f;
    p = fmaf(p, s, 21844.f / 6081075.f);
    p = fmaf(p, s, -1382.f / 155925.f);
    p = fmaf(p, s, 62.f / 2835.f);
    p = fmaf(p, s, -17.f / 315.f);
    p = fmaf(p, s, 2.f / 15.f);
    p = fmaf(p, s, -1.f / 3.f);
    const float poly = fmaf(a * s, p, a);
    const float e = __builtin_amdgcn_exp2f(-2.885390081777927f * a);   // 2 log2(e)
    const float t = a < 0.55f ? poly : (1.f - e) * __builtin_amdgcn_rcpf(1.f + e);
    return copysignf(t, x);
}

// Epilogue of one wave's 32 x 16*TN tile of a product's accumulators (rows row0.., columns
// n0..; accumulator element j of lane l is row 4*(l/16)+j, column l mod 16 of its 16x16 block):
//   kEpiTanh:   tanh(acc + bias) -> Y (BF16 row-major) and, if Ys is set, Ys (strip layout);
//   kEpiLinear: acc + bias -> Y32 (FP32 row-major);
//   kEpiDtanh:  acc * (1 - h^2) with h the layer output read from Hs (strip layout) -> Y (if
//               set) and Ys; column sums of the FP32 values over the wave's rows (< M) into
//               colpart[row0 / 32][N] (the bias gradient's first stage, fixed order).
// BF16 outputs are staged in the wave's LDS tiles sT and sS so that global stores are 16-byte
// rows and strips.  Rows at or past M (the products repeat row M-1 there, so their loads stay
// unconditional and pipeline) are not stored row-major; their strip entries (finite) are masked
// by every reader.  Every wave of the block calls it together (it synchronises the block).
template <int EPI, int TN>
__device__ __forceinline__ void gemm_epilogue(const f32x4 (&acc)[2][TN], int row0, int n0, int M, int N,
                                              const float *__restrict__ bias, const bf16 *__restrict__ Hs,
                                              bf16 *__restrict__ Y, bf16 *__restrict__ Ys, float *__restrict__ Y32,
                                              float *__restrict__ colpart, bf16 (*sT)[16 * TN + 8],
                                              bf16 (*sS)[16 * TN][8]) {
    constexpr int TC = 16 * TN;
    const int lane = threadIdx.x & 63, lr = lane & 15, lg = lane >> 4;
    if constexpr (EPI == kEpiLinear) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int t = 0; t < TN; ++t) {
                const int col = n0 + t * 16 + lr;
                const float bb = bias[col];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int row = row0 + r * 16 + 4 * lg + j;
                    if (row < M) Y32[(size_t)row * N + col] = acc[r][t][j] + bb;
                }
            }
    } else {
#pragma unroll
        for (int t = 0; t < TN; ++t) {
            const int col = n0 + t * 16 + lr;
            float csum = 0.f;
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                const int R = row0 + r * 16 + 4 * lg;   // first of this lane's 4 rows (a multiple of 4)
                float v[4];
                if constexpr (EPI == kEpiTanh) {
                    const float bb = bias[col];
#pragma unroll
                    for (int j = 0; j < 4; ++j) v[j] = tanh_bf16(acc[r][t][j] + bb);
                } else {
                    const bf16x4 h4 = *(const bf16x4 *)&Hs[strip_index(R, col, N)];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const float h = (float)h4[j];
                        v[j] = acc[r][t][j] * (1.f - h * h);
                        if (R + j < M) csum += v[j];
                    }
                }
                bf16x4 o;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    o[j] = (bf16)v[j];
                    sT[r * 16 + 4 * lg + j][t * 16 + lr] = o[j];
                }
                *(bf16x4 *)&sS[2 * r + (lg >> 1)][t * 16 + lr][4 * (lg & 1)] = o;
            }
            if constexpr (EPI == kEpiDtanh) {
                // the four lane groups' sums, (g0 + g1) + (g2 + g3) in every lane
                csum += __shfl_xor(csum, 16);
                csum += __shfl_xor(csum, 32);
                if (lg == 0) colpart[(size_t)(row0 / 32) * N + col] = csum;
            }
        }
        __syncthreads();
        if (Y) {
            for (int e = lane; e < 32 * (TC / 8); e += 64) {
                const int row = e / (TC / 8), c8 = e - row * (TC / 8);
                if (row0 + row < M)
                    *(bf16x8 *)&Y[(size_t)(row0 + row) * N + n0 + c8 * 8] = *(const bf16x8 *)&sT[row][c8 * 8];
            }
        }
        if (Ys) {
            for (int e = lane; e < 4 * TC; e += 64) {
                const int b = e / TC, c = e - b * TC;
                *(bf16x8 *)&Ys[strip_index(row0 + 8 * b, n0 + c, N)] = *(const bf16x8 *)&sS[b][c][0];
            }
        }
    }
}

// Y[M x N] = epilogue(X[M x K] * B), B given as BT = B^T [N x K] row-major, K % 32 == 0.  Wave
// tile 32 rows x 16*TN columns (2 x TN accumulators of the 16x16x32 MFMA: lane l holds
// row/column l mod 16 of its operand and eight consecutive k of group l/16), operands loaded
// straight from global memory (the weights stay in L2): the rows four k-steps ahead of the
// MFMAs, the weights one; a block's 4 waves take the N/(16*TN) column tiles of 4*16*TN/N row
// groups.  kEpiMse (the output layer in training, N = kMlpOut): o = acc + bias is compared with
// the targets, dOut = 2 (o - t) / (M n_out) (0 in padded columns) goes to Y (BF16 row-major),
// Ys (strips) and mse.d32, and the block's sum of (o - t)^2 to mse.lpart[blockIdx.x]; rows at or
// past M write nothing.  o is the binary32 value the inference epilogue stores, so the gradient
// equals the one formed from stored outputs.
template <int EPI, int TN, int K>
__global__ __launch_bounds__(kBlock) void k_gemm(const bf16 *__restrict__ X, int ldx, int M,
                                                 const bf16 *__restrict__ BT, int N, const float *__restrict__ bias,
                                                 const bf16 *__restrict__ Hs, bf16 *__restrict__ Y,
                                                 bf16 *__restrict__ Ys, float *__restrict__ Y32,
                                                 float *__restrict__ colpart, MseArgs mse) {
    constexpr int TC = 16 * TN;
    constexpr int S = K / 32, D = S < 4 ? S : 4;   // k-steps; k-steps of A in flight
    __shared__ __attribute__((aligned(16))) bf16 sT[kBlock / 64][32][TC + 8];   // row-major tile of each wave
    __shared__ __attribute__((aligned(16))) bf16 sS[kBlock / 64][4][TC][8];     // the same tile in strips
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const int CT = N / TC, RG = (kBlock / 64) / CT;
    const int row0 = (blockIdx.x * RG + w / CT) * 32, n0 = (w % CT) * TC;
    const bf16 *xa[2], *xb[TN];
#pragma unroll
    for (int r = 0; r < 2; ++r) xa[r] = X + (size_t)min(row0 + r * 16 + lr, M - 1) * ldx + 8 * lg;
#pragma unroll
    for (int t = 0; t < TN; ++t) xb[t] = BT + (size_t)(n0 + t * 16 + lr) * K + 8 * lg;
    f32x4 acc[2][TN];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int t = 0; t < TN; ++t) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // A fragments (from HBM) D k-steps ahead of the MFMAs, weight fragments (L2) one step ahead;
    // the scheduling barriers keep the loads where they are written instead of sinking them to
    // their uses
    bf16x8 fa[D][2], fb[2][TN];
#pragma unroll
    for (int s = 0; s < D; ++s)
#pragma unroll
        for (int r = 0; r < 2; ++r) fa[s][r] = *(const bf16x8 *)(xa[r] + 32 * s);
#pragma unroll
    for (int t = 0; t < TN; ++t) fb[0][t] = *(const bf16x8 *)xb[t];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        if (s + 1 < S) {
#pragma unroll
            for (int t = 0; t < TN; ++t) fb[(s + 1) & 1][t] = *(const bf16x8 *)(xb[t] + 32 * (s + 1));
        }
        __builtin_amdgcn_sched_barrier(0);
        const int q = s % D;
#pragma unroll
        for (int t = 0; t < TN; ++t) {
            acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[q][0], fb[s & 1][t], acc[0][t], 0, 0, 0);
            acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[q][1], fb[s & 1][t], acc[1][t], 0, 0, 0);
        }
        if (s + D < S) {
#pragma unroll
            for (int r = 0; r < 2; ++r) fa[q][r] = *(const bf16x8 *)(xa[r] + 32 * (s + D));
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (EPI == kEpiMse) {
        double se = 0;
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int t = 0; t < TN; ++t) {
                const int col = n0 + t * 16 + lr;
                const float bb = bias[col];
                const int R = row0 + r * 16 + 4 * lg;   // first of this lane's 4 rows (a multiple of 4)
                bf16x4 o;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    float g = 0.f;
                    if (R + j < M && col < mse.n_out) {
                        const float diff = (acc[r][t][j] + bb) - mse.t[(size_t)(R + j) * mse.n_out + col];
                        se += (double)diff * (double)diff;
                        g = 2.f * diff / ((float)M * (float)mse.n_out);
                    }
                    o[j] = (bf16)g;
                    if (R + j < M) {
                        Y[(size_t)(R + j) * N + col] = o[j];
                        mse.d32[(size_t)(R + j) * N + col] = g;
                    }
                }
                if (R < M) *(bf16x4 *)&Ys[strip_index(R, col, N)] = o;
            }
        const double bs = block_sum(se);
        if (threadIdx.x == 0) mse.lpart[blockIdx.x] = bs;
    } else {
        gemm_epilogue<EPI, TN>(acc, row0, n0, M, N, bias, Hs, Y, Ys, Y32, colpart, sT[w], sS[w]);
    }
}

// Y[M x 256] = epilogue(X[M x K] * B) (B given as BT = B^T [256 x K], rows ldb >= K apart: a
// product over the first K columns of wider rows of X and BT when the rest are zero) with the
// weights held in registers: wave w of a block owns
// the output columns 64w..64w+63 and loads their weight fragments for all of K once, then the
// block walks the 32-row tiles blockIdx.x, blockIdx.x + gridDim.x, ... (its four waves on the
// same rows), so a weight is fetched once per block instead of once per wave and k-step.  Row
// tiles go through LDS, double-buffered: each thread issues its share of the next tile's rows
// (16-byte chunks, coalesced along the rows) before the current tile's MFMAs and writes it to the
// other buffer after the epilogue, so the rows' latency hides under a whole tile of work (with
// fragments loaded two k-steps ahead from HBM it was exposed every tile).  The buffer holds the
// tile fragment-major -- chunk (k-step s, row block r, lane l) at (2 s + r) * 64 + l -- so a
// wave's fragment read is one contiguous 1 KB run.  The k-steps and MFMAs of each output element
// are those of k_gemm<EPI, 4, K>, in the same order: the same bits.
template <int EPI, int K>
__global__ __launch_bounds__(kBlock) void k_gemm_ws(const bf16 *__restrict__ X, int ldx, int M,
                                                    const bf16 *__restrict__ BT, int ldb, const float *__restrict__ bias,
                                                    const bf16 *__restrict__ Hs, bf16 *__restrict__ Y,
                                                    bf16 *__restrict__ Ys, float *__restrict__ colpart) {
    static_assert(EPI != kEpiLinear, "FP32 outputs go through k_gemm");
    constexpr int TN = 4, TC = 16 * TN, N = kMlpW;
    constexpr int S = K / 32;                          // k-steps
    constexpr int RC = K / 8;                          // 16-byte chunks per row
    constexpr int CH = 32 * RC;                        // chunks per row tile
    constexpr int CPT = (CH + kBlock - 1) / kBlock;    // chunks per thread
    static_assert((kBlock / 64) * TC == N, "the block's waves cover the columns");
    __shared__ bf16x8 sA[2][2 * S * 64];
    __shared__ __attribute__((aligned(16))) bf16 sT[kBlock / 64][32][TC + 8];
    __shared__ __attribute__((aligned(16))) bf16 sS[kBlock / 64][4][TC][8];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    const int n0 = w * TC, ntiles = (M + 31) / 32;
    int tile = blockIdx.x;
    if (tile >= ntiles) return;   // the whole block
    bf16x8 fb[S][TN];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
        for (int t = 0; t < TN; ++t) fb[s][t] = *(const bf16x8 *)(BT + (size_t)(n0 + t * 16 + lr) * ldb + 32 * s + 8 * lg);
    // this thread's chunks of a row tile: chunk q = threadIdx.x + kBlock i is row q / RC (rows at
    // or past M repeat row M-1), k-chunk q % RC; its place in the fragment-major buffer
    int dst[CPT];
#pragma unroll
    for (int i = 0; i < CPT; ++i) {
        const int q = (int)threadIdx.x + kBlock * i, R = q / RC, c = q % RC;
        dst[i] = (2 * (c >> 2) + (R >> 4)) * 64 + (R & 15) + 16 * (c & 3);
    }
    auto fetch = [&](int tl, bf16x8 (&v)[CPT]) {
#pragma unroll
        for (int i = 0; i < CPT; ++i) {
            const int q = (int)threadIdx.x + kBlock * i;
            if (CH % kBlock == 0 || q < CH)
                v[i] = *(const bf16x8 *)(X + (size_t)min(tl * 32 + q / RC, M - 1) * ldx + 8 * (q % RC));
        }
    };
    auto stage = [&](int b, const bf16x8 (&v)[CPT]) {
#pragma unroll
        for (int i = 0; i < CPT; ++i)
            if (CH % kBlock == 0 || (int)threadIdx.x + kBlock * i < CH) sA[b][dst[i]] = v[i];
    };
    bf16x8 pre[CPT];
    fetch(tile, pre);
    stage(0, pre);
    __syncthreads();
    int buf = 0;
    for (;;) {
        const int next = tile + (int)gridDim.x;
        const bool more = next < ntiles;
        fetch(more ? next : tile, pre);   // the last tile's own rows again when there is no next one
        f32x4 acc[2][TN];
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int t = 0; t < TN; ++t) acc[r][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const bf16x8 a0 = sA[buf][(2 * s) * 64 + lane], a1 = sA[buf][(2 * s + 1) * 64 + lane];
#pragma unroll
            for (int t = 0; t < TN; ++t) {
                acc[0][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, fb[s][t], acc[0][t], 0, 0, 0);
                acc[1][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, fb[s][t], acc[1][t], 0, 0, 0);
            }
        }
        gemm_epilogue<EPI, TN>(acc, tile * 32, n0, M, N, bias, Hs, Y, Ys, nullptr, colpart, sT[w], sS[w]);
        if (!more) break;
        stage(buf ^ 1, pre);
        __syncthreads();   // the next tile's rows are in place; the staging tiles are free again
        buf ^= 1;
        tile = next;
    }
}

// Fragments of one 32-row step of the weight gradient: this lane's strip (rows rs..rs+7) of
// four 16-column groups of A and TN of D; rows at or past r_end set to zero in both.
template <int TN>
struct WgradStep {
    bf16x8 a[4], b[TN];
    __device__ void load(const bf16 *__restrict__ As, int K, const bf16 *__restrict__ Ds, int N, int rs, int i0,
                         int n0, int lr) {
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] = *(const bf16x8 *)&As[strip_index(rs, i0 + u * 16 + lr, K)];
#pragma unroll
        for (int t = 0; t < TN; ++t) b[t] = *(const bf16x8 *)&Ds[strip_index(rs, n0 + t * 16 + lr, N)];
    }
    // r0: the step's first row (wave-uniform, so only the step that reaches r_end takes the
    // masking branch); rs: this lane's strip
    __device__ void mma(f32x4 (&acc)[4][TN], int r0, int rs, int r_end) {
        if (r0 + 32 > r_end) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const bool ok = rs + e < r_end;
#pragma unroll
                for (int u = 0; u < 4; ++u) a[u][e] = ok ? a[u][e] : (bf16)0.f;
#pragma unroll
                for (int t = 0; t < TN; ++t) b[t][e] = ok ? b[t][e] : (bf16)0.f;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int t = 0; t < TN; ++t) acc[u][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[u], b[t], acc[u][t], 0, 0, 0);
    }
};

// Partial weight gradient P[chunk][K][N] = A[rows of chunk]^T * D[rows of chunk] from the strip
// layouts As [M/8][K][8] and Ds [M/8][N][8] (K % 64 == 0, N % (16*TN) == 0, per % 32 == 0).
// Block (tile, chunk): a 64 x 16*TN tile of the gradient; its 4 waves take the chunk's 32-row
// steps w, w+4, ..., in order, each with the fragments of its next DQ steps in flight (a ring of
// DQ step buffers: a wave's steps are a chain of dependent HBM reads, so the depth of the ring
// divides the chain's latency), and add their tiles through LDS in wave order (fixed order:
// deterministic).  Rows at or past M are masked to zero in both operands.
template <int TN, int DQ>
__global__ __launch_bounds__(kBlock) void k_wgrad(const bf16 *__restrict__ As, int K, const bf16 *__restrict__ Ds,
                                                  int N, int M, int per, float *__restrict__ P) {
    constexpr int TC = 16 * TN, W = kBlock / 64;
    __shared__ f32x4 s_acc[W - 1][4][TN][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int lr = lane & 15, lg = lane >> 4;
    // (tile, chunk) in XCD order: the tiles of a chunk read the same strips, so they run on one
    // XCD, where the strips stay in its L2 (in dispatch order the four column tiles of a row of
    // tiles would land on four XCDs and fetch the strips four times from HBM)
    const uint32_t v = xcd_tile(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    const int tile = (int)(v % gridDim.x), chunk = (int)(v / gridDim.x);
    const int tiles_n = N / TC;
    const int i0 = (tile / tiles_n) * 64, n0 = (tile % tiles_n) * TC;
    const int r_begin = chunk * per, r_end = min(M, r_begin + per);
    f32x4 acc[4][TN];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int t = 0; t < TN; ++t) acc[u][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // step d of the ring holds rows r0 + 32 W d (row indices are wave-uniform); after its MFMAs
    // it takes the step DQ further on.  The loads are unconditional (a step past the chunk loads
    // the chunk's first rows again, in bounds, and is never multiplied), so the compiler can count
    // them and wait for the oldest step only.
    WgradStep<TN> st[DQ];
    const int r_first = r_begin + 32 * w;
    auto load_row = [&](int rr) { return (rr < r_end ? rr : r_begin) + 8 * lg; };
#pragma unroll
    for (int d = 0; d < DQ; ++d) st[d].load(As, K, Ds, N, load_row(r_first + 32 * W * d), i0, n0, lr);
    for (int r0 = r_first; r0 < r_end; r0 += 32 * W * DQ) {
#pragma unroll
        for (int d = 0; d < DQ; ++d) {
            const int rr = r0 + 32 * W * d;
            if (rr < r_end) st[d].mma(acc, rr, rr + 8 * lg, r_end);
            st[d].load(As, K, Ds, N, load_row(rr + 32 * W * DQ), i0, n0, lr);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    if (w > 0) {
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int t = 0; t < TN; ++t) s_acc[w - 1][u][t][lane] = acc[u][t];
    }
    __syncthreads();
    if (w == 0) {
        float *Pc = P + (size_t)chunk * K * N;
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int t = 0; t < TN; ++t) {
                f32x4 v = acc[u][t];
#pragma unroll
                for (int ww = 0; ww < W - 1; ++ww) v += s_acc[ww][u][t][lane];
#pragma unroll
                for (int j = 0; j < 4; ++j) Pc[(size_t)(i0 + u * 16 + 4 * lg + j) * N + n0 + t * 16 + lr] = v[j];
            }
    }
}

// Up to four independent sums of one kind in one launch (one per layer): blocks
// [start[j], start[j+1]) belong to job j.
constexpr int kMaxJobs = 4;
__device__ inline int job_of(const int *start, int n, int b) {
    int j = 0;
    while (j + 1 < n && b >= start[j + 1]) ++j;
    return j;
}

// out[e] = sum over chunks c (in order) of P[c][e], e < n: one thread per element.
struct ChunkSum {
    const float *P;
    float *out;
    int nchunks, n;
};
struct ChunkSums {
    ChunkSum job[kMaxJobs];
    int start[kMaxJobs + 1];
    int n;
};
__device__ inline void sum_chunks(const ChunkSums &J, int b) {
    const int j = job_of(J.start, J.n, b);
    const ChunkSum q = J.job[j];
    const int e = (b - J.start[j]) * blockDim.x + threadIdx.x;
    if (e >= q.n) return;
    float s = 0.f;
#pragma unroll 8
    for (int c = 0; c < q.nchunks; ++c) s += q.P[(size_t)c * q.n + e];
    q.out[e] = s;
}

// Column sums of an FP32 [M x N] array over row chunks of `per` rows: P[chunk][N].  A block
// covers 64 columns (one per lane) of one chunk; its 4 waves take every fourth row and their sums
// are added in wave order through LDS (fixed order: deterministic); eight loads in flight per
// thread.  Latency-bound: one thread per column over a 512-row chunk took 123 us per launch on a
// 32,768-row batch, and one block over 1,024 rows without the unrolling 60 us.
// Jobs: Z [M x N] (ld N) summed over chunks of `per` rows into P [chunks x N]; a job's blocks
// are its (64-column block, chunk) pairs, column blocks inner.
struct ColSum {
    const float *Z;
    float *P;
    int M, N, per;
};
struct ColSums {
    ColSum job[kMaxJobs];
    int start[kMaxJobs + 1];
    int n;
};
__device__ inline void colsum(const ColSums &J, int blk) {
    __shared__ float s_w[kBlock / 64][64];
    const int j = job_of(J.start, J.n, blk);
    const ColSum q = J.job[j];
    const float *__restrict__ Z = q.Z;
    float *__restrict__ P = q.P;
    const int ldz = q.N, M = q.M, N = q.N, per = q.per;
    const int nb = (N + 63) / 64, b = blk - J.start[j];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int n = (b % nb) * 64 + lane, chunk = b / nb;
    const int r0 = chunk * per, r1 = min(M, r0 + per);
    float s = 0.f;
    if (n < N) {
#pragma unroll 8
        for (int r = r0 + w; r < r1; r += kBlock / 64) s += Z[(size_t)r * ldz + n];
    }
    s_w[w][lane] = s;
    __syncthreads();
    if (w == 0 && n < N) {
        float t = s_w[0][lane];
        for (int k = 1; k < kBlock / 64; ++k) t += s_w[k][lane];
        P[(size_t)chunk * N + n] = t;
    }
}
__global__ __launch_bounds__(kBlock) void k_colsum(ColSums J) { colsum(J, (int)blockIdx.x); }

// The mean squared error of a batch: *dst = (sum of the n per-block partials, in a fixed order) /
// denom.
struct LossSum {
    const double *partials;
    int n;
    double denom;
    double *dst;
};
__device__ inline void loss_sum(const LossSum &L) {
    double s = 0;
    for (int i = threadIdx.x; i < L.n; i += blockDim.x) s += L.partials[i];
    const double t = block_sum(s);
    if (threadIdx.x == 0) *L.dst = t / L.denom;
}

// The sums that need only the backward pass, in one launch: the weight-gradient chunk partials
// of every layer (blocks 0 ..), the first pass of every layer's bias column sums (the next
// c1.start[c1.n] blocks) and the loss (the last block).
__global__ __launch_bounds__(kBlock) void k_grad_sums(ChunkSums cs, ColSums c1, LossSum L) {
    const int b = (int)blockIdx.x, nc = cs.start[cs.n], n1 = c1.start[c1.n];
    if (b < nc) sum_chunks(cs, b);
    else if (b < nc + n1) colsum(c1, b - nc);
    else loss_sum(L);
}

// FP32 host input [M x n_in] -> BF16 [M x kMlpIn], zero padded, and (if Xs is set) its strip
// layout.
__global__ void k_pack_input(const float *__restrict__ x, int M, int n_in, bf16 *__restrict__ X,
                             bf16 *__restrict__ Xs) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)M * kMlpIn) return;
    const int r = (int)(e / kMlpIn), c = (int)(e % kMlpIn);
    const bf16 v = (bf16)(c < n_in ? x[(size_t)r * n_in + c] : 0.f);
    X[e] = v;
    if (Xs) Xs[strip_index(r, c, kMlpIn)] = v;
}

// Mini-batch of a resident training set: row r of the batch is row (a*(pos0+r) + b) mod Mall of
// the set (a coprime to Mall, so one pass of pos over 0..Mall-1 visits every row once).  One
// thread per (row, group of 8 columns) for the G groups that hold inputs or targets: the source
// row (two 64-bit remainders) once per group, one 16-byte row-major store; input columns past
// the groups holding inputs are zero already and not written.
__global__ void k_gather_rows(const float *__restrict__ xall, const float *__restrict__ tall, uint64_t Mall, int n_in,
                              int n_out, uint64_t a, uint64_t b, uint64_t pos0, int M, int G, bf16 *__restrict__ X,
                              bf16 *__restrict__ Xs, float *__restrict__ T) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)M * G) return;
    const int r = (int)(e / G), c0 = 8 * (int)(e % G);
    const uint64_t src = (a * ((pos0 + r) % Mall) + b) % Mall;
    if (c0 < n_in) {
        bf16x8 v;
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = (bf16)(c0 + k < n_in ? xall[src * n_in + c0 + k] : 0.f);
        *(bf16x8 *)&X[(size_t)r * kMlpIn + c0] = v;
#pragma unroll
        for (int k = 0; k < 8; ++k) Xs[strip_index(r, c0 + k, kMlpIn)] = v[k];
    }
    for (int c = c0; c < c0 + 8 && c < n_out; ++c) T[(size_t)r * n_out + c] = tall[src * n_out + c];
}

// Where each layer's weights sit in the parameter vector and their BF16 copies.
struct PackTable {
    int woff[4], Kp[4], Np[4];
    bf16 *WT[4], *Wb[4];
};

// Adam on the FP32 master parameters, and the BF16 copies of each updated weight (W [in][out]
// and W^T [out][in], as k_pack_weights writes them).
__global__ __launch_bounds__(kBlock) void k_adam(float *__restrict__ p, float *__restrict__ m, float *__restrict__ v,
                                                 const float *__restrict__ g, size_t n, float lr, float b1, float b2,
                                                 float eps, float c1, float c2, PackTable T) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.f - b1) * gg;
    const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    const float pe = p[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    p[e] = pe;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
        const long o = (long)e - T.woff[l];
        if (o >= 0 && o < (long)T.Kp[l] * T.Np[l]) {
            const int k = (int)(o / T.Np[l]), c = (int)(o - (long)k * T.Np[l]);
            const bf16 b = (bf16)pe;
            T.Wb[l][o] = b;
            T.WT[l][(size_t)c * T.Kp[l] + k] = b;
        }
    }
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
    int cus = 256;   // compute units of the device
    gr::DevArray<float> p, m, v, g;
    gr::DevArray<gr::bf16> WT[4], Wb[4];
    // Row capacities (multiples of 4096): the forward pass's buffers (X0, H, out), the host
    // input staging (xin) and the training-only buffers (strips, targets, gradients, partials),
    // each grown only by what uses it, so that evaluating chunks of 2^18 rows does not grow the
    // training buffers to that size.
    int cap = 0, xcap = 0, tcap = 0;
    gr::DevArray<float> xin, tgt, out, d32, part, cpart, colpart;
    // activations row-major and in strips (X0s, Hs), output-gradient rows ping-pong (d16[2],
    // row-major and d16s[2] in strips)
    gr::DevArray<gr::bf16> X0, X0s, H[3], Hs[3], d16[2], d16s[2];
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
    static int rows_cap(int M) { return ((M + 4095) / 4096) * 4096; }
    void ensure(int M) {
        if (M <= cap) return;
        cap = rows_cap(M);
        out.alloc((size_t)cap * gr::kMlpOut);
        // input columns past the 8-column groups that hold inputs stay zero: cleared here, and
        // the writers (k_pack_input, k_gather_rows, the device-side fill) never write them non-zero
        X0.alloc((size_t)cap * gr::kMlpIn);
        GR_HIP(hipMemsetAsync(X0.p, 0, X0.n * sizeof(gr::bf16), stream));
        GR_HIP(hipStreamSynchronize(stream));   // the device-side fill writes X0 from another stream
        for (int l = 0; l < 3; ++l) H[l].alloc((size_t)cap * gr::kMlpW);
    }
    void ensure_host_input(int M) {
        if (M <= xcap) return;
        xcap = rows_cap(M);
        xin.alloc((size_t)xcap * gr::kMlpIn);
    }
    void ensure_train(int M) {
        ensure(M);
        if (M <= tcap) return;
        tcap = rows_cap(M);
        tgt.alloc((size_t)tcap * gr::kMlpOut);
        X0s.alloc((size_t)tcap * gr::kMlpIn);
        GR_HIP(hipMemsetAsync(X0s.p, 0, X0s.n * sizeof(gr::bf16), stream));
        for (int l = 0; l < 3; ++l) Hs[l].alloc((size_t)tcap * gr::kMlpW);
        for (int k = 0; k < 2; ++k) {
            d16[k].alloc((size_t)tcap * gr::kMlpW);
            d16s[k].alloc((size_t)tcap * gr::kMlpW);
        }
        d32.alloc((size_t)tcap * gr::kMlpOut);
        colpart.alloc((size_t)3 * (tcap / 32) * gr::kMlpW);
        part.alloc(4 * kPartLayer);
        cpart.alloc(4 * kCpartLayer);
        lpart.alloc((size_t)tcap * gr::kMlpOut / gr::kBlock + 1);
    }
    void pack(hipStream_t s) {
        for (int l = 0; l < 4; ++l)
            hipLaunchKernelGGL(gr::k_pack_weights, dim3(gr::blocks_for((size_t)Kp[l] * Np[l])), dim3(gr::kBlock), 0, s,
                               p.p + woff[l], Kp[l], Np[l], WT[l].p, Wb[l].p);
    }
    // strips: also write the input's strip layout (training)
    void upload_x(const float *x, int M, bool strips) {
        ensure_host_input(M);
        GR_HIP(hipMemcpyAsync(xin.p, x, (size_t)M * n_in * sizeof(float), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(gr::k_pack_input, dim3(gr::blocks_for((size_t)M * gr::kMlpIn)), dim3(gr::kBlock), 0, stream,
                           xin.p, M, (int)n_in, X0.p, strips ? X0s.p : nullptr);
    }
    static unsigned gemm_blocks(int M, int N, int TN) {
        const int rows_per_block = 32 * ((gr::kBlock / 64) / (N / (16 * TN)));
        return (unsigned)((M + rows_per_block - 1) / rows_per_block);
    }
    // blocks of a weight-stationary product: kWsBlocksPerCu per compute unit, at most one per
    // 32-row tile
    static constexpr int kWsBlocksPerCu = 2;
    unsigned ws_blocks(int M) const { return (unsigned)std::min((M + 31) / 32, kWsBlocksPerCu * cus); }
    // strips: also write the hidden layers' strip layouts (training)
    void forward(int M, hipStream_t s, bool strips) {
        const dim3 blk(gr::kBlock), gh(ws_blocks(M)), go(gemm_blocks(M, gr::kMlpOut, 2));
        const gr::bf16 *in[3] = {X0.p, H[0].p, H[1].p};
        const int ld[3] = {gr::kMlpIn, gr::kMlpW, gr::kMlpW};
        for (int l = 0; l < 3; ++l) {
            // layer 1 over the first 32 input columns when the inputs fit there (the padding
            // columns of X0 and rows of W1 are zero: the second k-step would add only zeros)
            const auto kern = l > 0 ? gr::k_gemm_ws<gr::kEpiTanh, gr::kMlpW>
                                    : n_in <= 32 ? gr::k_gemm_ws<gr::kEpiTanh, 32> : gr::k_gemm_ws<gr::kEpiTanh, gr::kMlpIn>;
            hipLaunchKernelGGL(kern, gh, blk, 0, s, in[l], ld[l], M, WT[l].p, Kp[l], p.p + boff[l],
                               (const gr::bf16 *)nullptr, H[l].p, strips ? Hs[l].p : nullptr, (float *)nullptr);
        }
        if (strips)   // training: the output layer's epilogue forms the loss gradient (k_gemm kEpiMse)
            hipLaunchKernelGGL((gr::k_gemm<gr::kEpiMse, 2, gr::kMlpW>), go, blk, 0, s, H[2].p, gr::kMlpW, M, WT[3].p,
                               Np[3], p.p + boff[3], (const gr::bf16 *)nullptr, d16[0].p, d16s[0].p, (float *)nullptr,
                               (float *)nullptr, gr::MseArgs{tgt.p, (int)n_out, d32.p, lpart.p});
        else
            hipLaunchKernelGGL((gr::k_gemm<gr::kEpiLinear, 2, gr::kMlpW>), go, blk, 0, s, H[2].p, gr::kMlpW, M, WT[3].p,
                               Np[3], p.p + boff[3], (const gr::bf16 *)nullptr, (gr::bf16 *)nullptr, (gr::bf16 *)nullptr,
                               out.p, (float *)nullptr, gr::MseArgs{});
    }
    // Partial buffers per layer l: weight-gradient chunk partials at part + l * kPartLayer,
    // first-pass bias column sums at cpart + l * kCpartLayer.
    static constexpr size_t kPartLayer = (size_t)gr::kWgradChunks * gr::kMlpW * gr::kMlpW;
    static constexpr size_t kCpartLayer = (size_t)1040 * gr::kMlpW;   // sqrt(2^20) + 16 chunks
    int wchunks[4] = {0, 0, 0, 0};
    // bias column sums of layer l from rows Z [rows x N]: about sqrt(rows) chunks of at least
    // 32 rows
    static int colsum_per(int rows) { return std::max(32, (int)std::ceil(std::sqrt((double)rows))); }
    // weight gradient of layer l from the strips of its input (As) and output gradient (Ds) into
    // its chunk partials: row chunks of multiples of 32 rows, at most kWgradChunks, up to 512
    // blocks (summed in order by sums())
    static constexpr int kWgradDepth = 3;   // row steps in flight per wave
    template <int TN>
    void wgrad(int l, const gr::bf16 *As, const gr::bf16 *Ds, int M) {
        const int tiles = (Kp[l] / 64) * (Np[l] / (16 * TN));
        const int want = std::max(1, std::min({(M + 127) / 128, 512 / tiles, gr::kWgradChunks}));
        const int per = ((M + want - 1) / want + 31) / 32 * 32;
        wchunks[l] = (M + per - 1) / per;
        hipLaunchKernelGGL((gr::k_wgrad<TN, kWgradDepth>), dim3(tiles, wchunks[l]), dim3(gr::kBlock), 0, stream, As,
                           Kp[l], Ds, Np[l], M, per, part.p + l * kPartLayer);
    }
    // The ordered sums of all four layers and the loss, two launches: the weight-gradient chunk
    // partials, the first pass of the bias column sums (layer 4 over the FP32 output gradient
    // d32, layers 1-3 over the per-32-row sums the input-gradient products left in colpart) and
    // the loss together, then the bias sums' second pass.
    void sums(int M, const gr::LossSum &loss) {
        gr::ChunkSums cs{};
        cs.n = 4;
        for (int l = 0, b = 0; l < 4; ++l) {
            const int n = Kp[l] * Np[l];
            cs.job[l] = gr::ChunkSum{part.p + l * kPartLayer, g.p + woff[l], wchunks[l], n};
            cs.start[l] = b;
            b += (int)gr::blocks_for(n);
            cs.start[l + 1] = b;
        }
        gr::ColSums c1{}, c2{};
        c1.n = c2.n = 4;
        const int hrows = (M + 31) / 32;
        for (int l = 0, b1 = 0, b2 = 0; l < 4; ++l) {
            const int N = Np[l], rows = l == 3 ? M : hrows, per = colsum_per(rows);
            const int chunks = (rows + per - 1) / per, nb = (N + 63) / 64;
            const float *Z = l == 3 ? d32.p : colpart.p + (size_t)l * (tcap / 32) * gr::kMlpW;
            float *P = cpart.p + l * kCpartLayer;
            c1.job[l] = gr::ColSum{Z, P, rows, N, per};
            c1.start[l] = b1;
            b1 += nb * chunks;
            c1.start[l + 1] = b1;
            c2.job[l] = gr::ColSum{P, g.p + boff[l], chunks, N, chunks};
            c2.start[l] = b2;
            b2 += nb;
            c2.start[l + 1] = b2;
        }
        hipLaunchKernelGGL(gr::k_grad_sums, dim3(cs.start[4] + c1.start[4] + 1), dim3(gr::kBlock), 0, stream, cs, c1,
                           loss);
        hipLaunchKernelGGL(gr::k_colsum, dim3(c2.start[4]), dim3(gr::kBlock), 0, stream, c2);
    }
    // loss (into *loss_dst on the device) and gradients for the targets in tgt (the training
    // forward pass must have run: it writes the strips and, in the output layer's epilogue, the
    // output gradient and the squared-error partials); no host sync
    void backward(int M, double *loss_dst) {
        const unsigned nb = gemm_blocks(M, gr::kMlpOut, 2);   // squared-error partials of the output layer
        const dim3 blk(gr::kBlock), gh(ws_blocks(M));
        // layer 4 (linear): dW4 = H3^T dOut
        wgrad<2>(3, Hs[2].p, d16s[0].p, M);
        // layers 3..1: dZ_l = (dZ_{l+1} W_{l+1}^T) * (1 - H_l^2) in the input-gradient product's
        // epilogue (with the first stage of the bias gradient), then dW_l = A_l^T dZ_l
        int cur = 0;
        for (int l = 2; l >= 0; --l) {
            const int nxt = 1 - cur;
            hipLaunchKernelGGL(l == 2 ? (gr::k_gemm_ws<gr::kEpiDtanh, gr::kMlpOut>) : (gr::k_gemm_ws<gr::kEpiDtanh, gr::kMlpW>),
                               gh, blk, 0, stream, d16[cur].p, Np[l + 1], M, Wb[l + 1].p, Np[l + 1], (const float *)nullptr,
                               (const gr::bf16 *)Hs[l].p, l > 0 ? d16[nxt].p : nullptr, d16s[nxt].p,
                               colpart.p + (size_t)l * (tcap / 32) * gr::kMlpW);
            wgrad<4>(l, l == 0 ? X0s.p : Hs[l - 1].p, d16s[nxt].p, M);
            cur = nxt;
        }
        sums(M, gr::LossSum{lpart.p, (int)nb, (double)M * n_out, loss_dst});
        GR_HIP(hipGetLastError());
    }
    // Adam step on every parameter; the BF16 weight copies are refreshed in the same launch
    void adam(float lr) {
        step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
        gr::PackTable T{};
        for (int l = 0; l < 4; ++l) {
            T.woff[l] = (int)woff[l];
            T.Kp[l] = Kp[l];
            T.Np[l] = Np[l];
            T.WT[l] = WT[l].p;
            T.Wb[l] = Wb[l].p;
        }
        hipLaunchKernelGGL(gr::k_adam, dim3(gr::blocks_for(n_params)), dim3(gr::kBlock), 0, stream, p.p, m.p, v.p, g.p,
                           n_params, lr, b1, b2, 1e-8f, c1, c2, T);
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
        n->forward(m, s, false);
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
            int dev = 0;
            GR_HIP(hipGetDevice(&dev));
            GR_HIP(hipDeviceGetAttribute(&n->cus, hipDeviceAttributeMultiprocessorCount, dev));
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
            n->upload_x(x + (size_t)r0 * n->n_in, m, false);
            n->forward(m, n->stream, false);
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
        n->ensure_train(M);
        n->upload_x(x, M, true);
        GR_HIP(hipMemcpyAsync(n->tgt.p, t, (size_t)M * n->n_out * sizeof(float), hipMemcpyHostToDevice, n->stream));
        n->forward(M, n->stream, true);
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
        n->ensure_train(M);
        n->upload_x(x, M, true);
        GR_HIP(hipMemcpyAsync(n->tgt.p, t, (size_t)M * n->n_out * sizeof(float), hipMemcpyHostToDevice, n->stream));
        n->forward(M, n->stream, true);
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
        n->ensure_train(batch);
        gr::DevArray<double> ldev;
        ldev.alloc((size_t)steps);
        const uint64_t Mall = n->n_rows, per_epoch = Mall / (uint64_t)batch;
        // 8-column groups holding inputs or targets (the gather's threads per row)
        const int G = (int)std::max((n->n_in + 7) / 8, (n->n_out + 7) / 8);
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
            hipLaunchKernelGGL(gr::k_gather_rows, dim3(gr::blocks_for((size_t)batch * G)), dim3(gr::kBlock), 0,
                               n->stream, n->xall.p, n->tall.p, Mall, (int)n->n_in, (int)n->n_out, a, b,
                               within * (uint64_t)batch, batch, G, n->X0.p, n->X0s.p, n->tgt.p);
            n->forward(batch, n->stream, true);
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

// Batched Euler solver: up to 64 flow instances (different far-field Mach vectors) on one
// shared mesh, advanced together.
//
// Layout is instance-inner: component c of instance k in cell i lives at c*stride + i*64 + k.
// One wavefront handles one cell, its 64 lanes the instances, so every state access is one
// fully coalesced 512-byte request and every geometry/connectivity access is uniform across
// the wave (scalar loads, amortised over the instances).  Each lane evaluates the same
// expressions in the same order as the single-instance kernels, so instance k of a batch
// equals a single solve with instance k's far field bit for bit in binary64.
// Convergence is per instance: an instance that reaches the tolerance is frozen while the
// others continue, and the launch sequence stops when all have stopped.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "fvm_tiles.hpp"

namespace gr {

constexpr int kBatchLanes = 64;
constexpr int kCellsPerBlock = kBlock / kBatchLanes;   // 4 (ghost kernels; default schedule)
constexpr int kFinalParts = 16;                       // partial-sum groups of the batch reduction
constexpr uint32_t kReduceBlocks = 256;                // first-stage blocks of the batch reduction

struct BatchCtl {
    double tol;
    int max_iter;
    int n_inst;
    int stop;        // all instances stopped
    int hist_cap;    // history entries per instance
    double ref[kBatchLanes];
    double last[kBatchLanes];
    double ratio[kBatchLanes];
    int iter[kBatchLanes];
    int done[kBatchLanes];
    int converged[kBatchLanes];
    int diverged[kBatchLanes];
};

template <typename S>
struct BatchArgs {
    const S *u;   // 5 component blocks of stride elements
    S *un;
    S *g;         // 15 blocks
    S *l;         // 5 blocks
    size_t stride;
    const S *frec;   // per-slot flux, gradient and ghost records (k_slot_geometry): the metrics
    const S *grec;   // every instance of a cell reads, cast as the kernels cast them, by slot
    const S *gnrm;   // index instead of gathered by neighbour and face
};

// Cell of this wave (wave-uniform, in a scalar register).  Blocks are mapped XCD-aware: the
// blocks one XCD runs (block index mod 8) take a contiguous range of cells, so the neighbour
// rows a cell re-reads are in that XCD's L2.
template <int CPB>
__device__ inline uint32_t batch_cell() {
    return __builtin_amdgcn_readfirstlane(xcd_tile(blockIdx.x, gridDim.x) * CPB + (threadIdx.x >> 6));
}

// Far-field region uses the instance's Mach vector: copy the region with it substituted.
__device__ inline BcRegion inst_region(const BcRegion &R, const double *inst_mach, uint32_t k) {
    BcRegion r = R;
    if (R.kind == GR_BC_FARFIELD) {
        r.mach[0] = inst_mach[3 * k];
        r.mach[1] = inst_mach[3 * k + 1];
        r.mach[2] = inst_mach[3 * k + 2];
    }
    return r;
}

// CPB cells (waves) per block: one of the schedules the batch's autotuner chooses among.
template <typename S, int CPB>
__global__ __launch_bounds__(CPB *kBatchLanes) void k_batch_gradlim(MeshView m, BatchArgs<S> a, BcTable bc,
                                                                    const double *__restrict__ inst_mach, int kind,
                                                                    int corr, const BatchCtl *ctl) {
    if (ctl->stop) return;
    const uint32_t i = batch_cell<CPB>();
    if (i >= m.N) return;
    const uint32_t k = threadIdx.x & 63;
    const size_t st = a.stride, me = (size_t)i * kBatchLanes + k;
    S ui[5], umin[5], umax[5], g[15];
#pragma unroll
    for (int c = 0; c < 5; ++c) ui[c] = umin[c] = umax[c] = a.u[c * st + me];
#pragma unroll
    for (int q = 0; q < 15; ++q) g[q] = S(0);
    const uint32_t s = m.cfs[i], e = m.cfs[i + 1];
    for (uint32_t q = s; q < e; ++q) {
        const uint32_t j = m.cc[q];
        // d = x_j - x_i from the slot's gradient record; the ghost's face normal from its record
        const S d[3] = {a.grec[(size_t)q * kGradRec], a.grec[(size_t)q * kGradRec + 1], a.grec[(size_t)q * kGradRec + 2]};
        S uj[5];
        const int r = j >= m.N ? bc_region(bc, j - m.N) : -1;
        if (r < 0) {
#pragma unroll
            for (int c = 0; c < 5; ++c) uj[c] = a.u[c * st + (size_t)j * kBatchLanes + k];
        } else {
            const S *gn = a.gnrm + (size_t)(j - m.N) * kGhostRec;
            const S n[3] = {gn[0], gn[1], gn[2]};
            ghost_state<S>(inst_region(bc.r[r], inst_mach, k), n, ui, uj);
        }
        const S w2 = fdiv(S(1), (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const S du = uj[c] - ui[c];
#pragma unroll
            for (int kk = 0; kk < 3; ++kk) g[3 * c + kk] += (d[kk] * du) * w2;
            umin[c] = fmin(umin[c], uj[c]);
            umax[c] = fmax(umax[c], uj[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        S out[3];
        lsq_apply<S>(m, i, &g[3 * c], out);
#pragma unroll
        for (int kk = 0; kk < 3; ++kk) g[3 * c + kk] = out[kk];
    }
    S lim[5] = {S(2), S(2), S(2), S(2), S(2)};
    for (uint32_t q = s; q < e; ++q) {
        const S *gq = a.grec + (size_t)q * kGradRec + 3;   // fc - x_i
        const S dx = gq[0], dy = gq[1], dz = gq[2];
#pragma unroll
        for (int c = 0; c < 5; ++c)
            lim[c] = fmin(lim[c], limiter_face<S>(ui[c], g[3 * c], g[3 * c + 1], g[3 * c + 2], umin[c], umax[c], dx,
                                                  dy, dz, kind));
    }
    // binary64: the gradient and the limiters; binary32: the gradient rows premultiplied by the
    // limiter the reconstruction applies to them (15 rows instead of 20, as the tile kernels)
    if constexpr (sizeof(S) == 8) {
#pragma unroll
        for (int q = 0; q < 15; ++q) store_nt(&a.g[q * st + me], g[q]);
#pragma unroll
        for (int c = 0; c < 5; ++c) store_nt(&a.l[c * st + me], lim[c]);
    } else {
#pragma unroll
        for (int q = 0; q < 15; ++q) store_nt(&a.g[q * st + me], g[q] * lim[row_limiter(q, corr)]);
    }
}

// The flux + update, CPB waves per block; block partials of the per-instance residual.
//
// G instance groups: the 64 instances are split into G groups of 64/G lanes, and a wave takes G
// consecutive cells for one group (lane l: cell G*(wave's first) + l/(64/G), instance
// h*(64/G) + l%(64/G)).  The grid is G sweeps over the cells, one per group, and each XCD's
// contiguous range of blocks lies within one sweep, so the bytes an XCD loads between a
// neighbour row's first load and its re-reads by the cell's other neighbours shrink G-fold: with
// G = 1 a binary64 neighbour costs 12.8 KB and the window of cells in flight on an XCD is larger
// than its 4 MB L2.  Every lane evaluates the same expressions as with G = 1, so the states are
// bitwise the same; only the order of the residual's per-instance sum changes.
template <int CPB, int G>
__device__ inline void batch_flux_slot(uint32_t N, uint32_t &i, uint32_t &k, uint32_t &h) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if constexpr (G == 1) {
        i = batch_cell<CPB>();
        k = lane;
        h = 0;
    } else {
        constexpr uint32_t LG = kBatchLanes / G;
        const uint32_t T = (N + CPB * G - 1) / (CPB * G);   // blocks per sweep
        const uint32_t v = xcd_tile(blockIdx.x, gridDim.x);   // gridDim.x == G * T
        h = v / T;
        const uint32_t t = v - h * T;
        i = (t * CPB + w) * G + lane / LG;
        k = h * LG + lane % LG;
    }
}

template <typename S, bool CORR, int CPB, int G>
__device__ __forceinline__ void batch_flux(const MeshView &m, const BatchArgs<S> &a, const BcTable &bc,
                                           const double *__restrict__ inst_mach, double cfl,
                                           double *__restrict__ partials, const BatchCtl *ctl) {
    if (ctl->stop) return;
    __shared__ double s_err[CPB][kBatchLanes];
    uint32_t i, k, h;
    batch_flux_slot<CPB, G>(m.N, i, k, h);
    const uint32_t w = threadIdx.x >> 6;
    double e_loc = 0;
    if (i < m.N) {
        const size_t st = a.stride, me = (size_t)i * kBatchLanes + k;
        S ui[5], gi[15], li[5] = {S(0), S(0), S(0), S(0), S(0)};
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            ui[c] = a.u[c * st + me];
            if constexpr (sizeof(S) == 8) li[c] = a.l[c * st + me];
        }
        if (ctl->done[k]) {   // frozen instance: carry its state over
#pragma unroll
            for (int c = 0; c < 5; ++c) store_nt(&a.un[c * st + me], ui[c]);
        } else {
#pragma unroll
            for (int q = 0; q < 15; ++q) gi[q] = a.g[q * st + me];
            const double V = m.cv[i];
            S acc[5] = {0, 0, 0, 0, 0}, dts = 0;
            const uint32_t s = m.cfs[i], e = m.cfs[i + 1];
            for (uint32_t q = s; q < e; ++q) {
                const uint32_t j = m.cc[q];
                // outward normal, A_f / V_i, fc - x_i, fc - x_j: the slot's flux record
                const S *fr = a.frec + (size_t)q * kFluxRec;
                const S n[3] = {fr[0], fr[1], fr[2]};
                const S ds = fr[3];
                const S d0[3] = {fr[4], fr[5], fr[6]};
                const S d1[3] = {fr[7], fr[8], fr[9]};
                S r0, ru0[3], e0, r1, ru1[3], e1;
                if constexpr (sizeof(S) == 8) reconstruct<S, CORR>(ui, gi, li, d0, r0, ru0, e0);
                else reconstruct_pre<S, CORR>(ui, gi, d0, r0, ru0, e0);
                if (j < m.N) {
                    const size_t nj = (size_t)j * kBatchLanes + k;
                    S uj[5], gj[15], lj[5];
#pragma unroll
                    for (int c = 0; c < 5; ++c) {
                        uj[c] = a.u[c * st + nj];
                        if constexpr (sizeof(S) == 8) lj[c] = a.l[c * st + nj];
                    }
#pragma unroll
                    for (int q2 = 0; q2 < 15; ++q2) gj[q2] = a.g[q2 * st + nj];
                    if constexpr (sizeof(S) == 8) reconstruct<S, CORR>(uj, gj, lj, d1, r1, ru1, e1);
                    else reconstruct_pre<S, CORR>(uj, gj, d1, r1, ru1, e1);
                } else {
                    S ug[5];
                    const int r = bc_region(bc, j - m.N);
                    if (r < 0) {
#pragma unroll
                        for (int c = 0; c < 5; ++c) ug[c] = a.u[c * st + (size_t)j * kBatchLanes + k];
                    } else {
                        ghost_state<S>(inst_region(bc.r[r], inst_mach, k), n, ui, ug);
                    }
                    if constexpr (sizeof(S) == 8) reconstruct<S, CORR>(ug, gi, li, d1, r1, ru1, e1);
                    else reconstruct_pre<S, CORR>(ug, gi, d1, r1, ru1, e1);
                }
                const S p0 = S(0.4) * (e0 - fdiv(S(0.5), r0) * ((ru0[0] * ru0[0] + ru0[1] * ru0[1]) + ru0[2] * ru0[2]));
                const S p1 = S(0.4) * (e1 - fdiv(S(0.5), r1) * ((ru1[0] * ru1[0] + ru1[1] * ru1[1]) + ru1[2] * ru1[2]));
                const S V0 = fdiv((ru0[0] * n[0] + ru0[1] * n[1]) + ru0[2] * n[2], r0);
                const S V1 = fdiv((ru1[0] * n[0] + ru1[1] * n[1]) + ru1[2] * n[2], r1);
                acc[0] += (((V0 * r0) + (V1 * r1)) * ds) * S(0.5);
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    const S fu0 = ru0[kk] * V0 + n[kk] * p0;
                    const S fu1 = ru1[kk] * V1 + n[kk] * p1;
                    acc[1 + kk] += ((fu0 + fu1) * ds) * S(0.5);
                }
                acc[4] += (((V0 * (e0 + p0)) + (V1 * (e1 + p1))) * ds) * S(0.5);
                const S v0[3] = {fdiv(ru0[0], r0), fdiv(ru0[1], r0), fdiv(ru0[2], r0)};
                const S v1[3] = {fdiv(ru1[0], r1), fdiv(ru1[1], r1), fdiv(ru1[2], r1)};
                const S cs0 = dsqrt(fmax(S(1e-12), fmin(S(1e30), fdiv(S(1.4) * p0, r0))));
                const S cs1 = dsqrt(fmax(S(1e-12), fmin(S(1e30), fdiv(S(1.4) * p1, r1))));
                const S eig = fmax(dsqrt((v0[0] * v0[0] + v0[1] * v0[1]) + v0[2] * v0[2]) + cs0,
                                   dsqrt((v1[0] * v1[0] + v1[1] * v1[1]) + v1[2] * v1[2]) + cs1);
                acc[0] -= ((eig * (r1 - r0)) * S(0.5)) * ds;
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) acc[1 + kk] -= (((ru1[kk] - ru0[kk]) * eig) * S(0.5)) * ds;
                acc[4] -= ((eig * (e1 - e0)) * S(0.5)) * ds;
                dts += eig * ds;
            }
            const S dt = fdiv((S)cfl, dts);
#pragma unroll
            for (int c = 0; c < 5; ++c) store_nt(&a.un[c * st + me], ui[c] + (-acc[c]) * dt);
            const double r = acc[0], u0 = acc[1], u1 = acc[2], u2 = acc[3], en = acc[4];
            e_loc = ((r * r + ((u0 * u0 + u1 * u1) + u2 * u2)) + en * en) * V;
        }
    }
    s_err[w][threadIdx.x & 63] = e_loc;
    __syncthreads();
    if (threadIdx.x < kBatchLanes) {
        // row of instance sums: instance kk of this block's group from lanes sub*LG + kk%LG of
        // every wave (waves outer, cells inner: a fixed order); the other groups' entries 0
        constexpr uint32_t LG = kBatchLanes / G;
        const uint32_t kk = threadIdx.x;
        double t = 0;
        if (kk / LG == h) {
#pragma unroll
            for (int ww = 0; ww < CPB; ++ww)
#pragma unroll
                for (int sub = 0; sub < G; ++sub) t += s_err[ww][sub * LG + kk % LG];
        }
        partials[(size_t)blockIdx.x * kBatchLanes + kk] = t;
    }
}

// Two occupancies of the same flux: left to the compiler (the binary64 kernel takes 174 VGPRs,
// 2 waves per SIMD) and capped at 3 waves per SIMD (168 VGPRs, 6 spilled: 360 -> 325 us on the
// 64-flow C-grid; 4 waves spill 45 and take 580 us).  The autotuner times both.
template <typename S, bool CORR, int CPB, int G>
__global__ __launch_bounds__(CPB *kBatchLanes) void k_batch_flux(MeshView m, BatchArgs<S> a, BcTable bc,
                                                                 const double *__restrict__ inst_mach, double cfl,
                                                                 double *__restrict__ partials, const BatchCtl *ctl) {
    batch_flux<S, CORR, CPB, G>(m, a, bc, inst_mach, cfl, partials, ctl);
}
template <typename S, bool CORR, int CPB, int G>
__global__ __launch_bounds__(CPB *kBatchLanes) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_batch_flux_w3(
    MeshView m, BatchArgs<S> a, BcTable bc, const double *__restrict__ inst_mach, double cfl,
    double *__restrict__ partials, const BatchCtl *ctl) {
    batch_flux<S, CORR, CPB, G>(m, a, bc, inst_mach, cfl, partials, ctl);
}

// First stage of the per-instance residual sums: block r adds the flux blocks' partial rows
// r*per .. r*per+per-1 (wave w takes every fourth row, then the 4 wave sums in order), so the
// 64-lane rows are read by many CUs at once.  Fixed order: deterministic.
__global__ __launch_bounds__(kBlock) void k_batch_reduce(const double *__restrict__ partials, uint32_t nblocks,
                                                         const BatchCtl *ctl, double *__restrict__ partials2) {
    if (ctl && ctl->stop) return;
    __shared__ double s[kBlock / kBatchLanes][kBatchLanes];
    const uint32_t k = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t per = (nblocks + kReduceBlocks - 1) / kReduceBlocks;
    const uint32_t b0 = blockIdx.x * per, b1 = min(nblocks, b0 + per);
    double acc = 0;
    for (uint32_t b = b0 + w; b < b1; b += kBlock / kBatchLanes) acc += partials[(size_t)b * kBatchLanes + k];
    s[w][k] = acc;
    __syncthreads();
    if (w == 0) partials2[(size_t)blockIdx.x * kBatchLanes + k] = ((s[0][k] + s[1][k]) + s[2][k]) + s[3][k];
}

// Per-instance residual sums (fixed order) and the per-instance stopping rule.
__global__ __launch_bounds__(kFinalParts *kBatchLanes) void k_batch_finalize(const double *__restrict__ partials,
                                                                           uint32_t nblocks, BatchCtl *ctl,
                                                                           double *hist, double *out) {
    if (ctl && ctl->stop) return;
    __shared__ double s_part[kFinalParts][kBatchLanes];
    __shared__ int s_alive;
    const uint32_t k = threadIdx.x & 63, p = threadIdx.x >> 6;
    double acc = 0;
    for (uint32_t b = p; b < nblocks; b += kFinalParts) acc += partials[(size_t)b * kBatchLanes + k];
    s_part[p][k] = acc;
    if (threadIdx.x == 0) s_alive = 0;
    __syncthreads();
    if (threadIdx.x < kBatchLanes) {
        double tot = 0;
        for (int q = 0; q < kFinalParts; ++q) tot += s_part[q][k];
        if (!ctl) {
            out[k] = tot;
        } else if ((int)k < ctl->n_inst && !ctl->done[k]) {
            const int it = ctl->iter[k];
            if (it < ctl->hist_cap) hist[(size_t)k * ctl->hist_cap + it] = tot;
            if (it == 0 && ctl->ref[k] <= 0) ctl->ref[k] = tot;
            ctl->last[k] = tot;
            ctl->ratio[k] = tot / ctl->ref[k];
            ctl->iter[k] = it + 1;
            if (!isfinite(tot)) {
                ctl->diverged[k] = 1;
                ctl->done[k] = 1;
            } else if (ctl->tol > 0 && ctl->ratio[k] < ctl->tol) {
                ctl->converged[k] = 1;
                ctl->done[k] = 1;
            }
            if (ctl->iter[k] >= ctl->max_iter) ctl->done[k] = 1;
            if (!ctl->done[k]) s_alive = 1;
        }
    }
    __syncthreads();
    if (ctl && threadIdx.x == 0 && !s_alive) ctl->stop = 1;
}

// Boundary-condition ghost states of all instances into the state array (for reading back).
template <typename S>
__global__ void k_batch_ghosts(MeshView m, S *u, size_t stride, BcTable bc, const double *__restrict__ inst_mach) {
    const uint32_t g = __builtin_amdgcn_readfirstlane(blockIdx.x * kCellsPerBlock + (threadIdx.x >> 6));
    if (g >= m.G) return;
    const uint32_t k = threadIdx.x & 63;
    const int r = bc_region(bc, g);
    if (r < 0) return;
    const uint32_t c0 = m.N + g, q = m.cfs[c0];
    const uint32_t i = m.cc[q], f = m.cf[q];
    S ui[5], ug[5], n[3];
#pragma unroll
    for (int c = 0; c < 5; ++c) ui[c] = u[c * stride + (size_t)i * kBatchLanes + k];
    outward_normal<S>(m, f, m.cx[i], m.cy[i], m.cz[i], n);
    ghost_state<S>(inst_region(bc.r[r], inst_mach, k), n, ui, ug);
#pragma unroll
    for (int c = 0; c < 5; ++c) u[c * stride + (size_t)c0 * kBatchLanes + k] = ug[c];
}

// ---- learned warm start: network rows are (cell i, instance k), row = i * n_inst + k ----

struct FarField {
    double rho, c, E;   // density, speed of sound, total energy per volume without the kinetic part
};

__device__ inline FarField far_field(double p_inf, double T_inf) {
    FarField f;
    f.rho = p_inf / T_inf;
    f.c = dsqrt((1.4 * p_inf) / f.rho);
    f.E = p_inf / 0.4;
    return f;
}

// Raw network inputs of rows row0..row0+rows-1, one thread per row: with nrm set, normalised and
// written as bfloat16 into the first 8 columns of X [rows][64] (the 6 inputs, then 2 zeros; the
// columns after them are zero already and left alone); else binary32 into X [rows][6].
template <bool NORM>
__global__ void k_warm_features(MeshView m, const double *__restrict__ wd, const double *__restrict__ inst_mach,
                                uint32_t n_inst, size_t row0, int rows, gr_warm_norm nrm, void *X) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)rows) return;
    const size_t r = row0 + e;
    const uint32_t i = (uint32_t)(r / n_inst), k = (uint32_t)(r % n_inst);
    const double mx = inst_mach[3 * k], my = inst_mach[3 * k + 1], mz = inst_mach[3 * k + 2];
    // centre coordinates compressed about the origin: (x - o) / (1 + |x - o|) resolves the
    // body's neighbourhood and keeps the far field bounded
    const double ex = m.cx[i] - nrm.origin[0], ey = m.cy[i] - nrm.origin[1], ez = m.cz[i] - nrm.origin[2];
    const double q = 1.0 + dsqrt((ex * ex + ey * ey) + ez * ez);
    const double f[6] = {ex / q, ey / q, ez / q, log(wd[i] + 1e-3), dsqrt((mx * mx + my * my) + mz * mz),
                         atan2(my, mx)};
    if constexpr (NORM) {
        typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
        bf16x8 v;
#pragma unroll
        for (int c = 0; c < 6; ++c) v[c] = (__bf16)(float)((f[c] - nrm.in_mean[c]) / nrm.in_std[c]);
        v[6] = v[7] = (__bf16)0.f;
        *(bf16x8 *)((__bf16 *)X + e * kMlpInPad) = v;
    } else {
#pragma unroll
        for (int c = 0; c < 6; ++c) ((float *)X)[e * 6 + c] = (float)f[c];
    }
}

// Raw targets t [rows][5] of every interior cell and instance from the state array.
template <typename S>
__global__ void k_warm_targets(const S *__restrict__ u, size_t stride, uint32_t N, uint32_t n_inst,
                               const double *__restrict__ inst_mach, double p_inf, double T_inf, float *__restrict__ t) {
    const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (size_t)N * n_inst) return;
    const uint32_t i = (uint32_t)(r / n_inst), k = (uint32_t)(r % n_inst);
    const FarField ff = far_field(p_inf, T_inf);
    const size_t o = (size_t)i * kBatchLanes + k;
    double vel[3], ke = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        vel[d] = inst_mach[3 * k + d] * ff.c;
        ke += vel[d] * vel[d];
    }
    const double Einf = ff.E + 0.5 * ff.rho * ke;
    t[5 * r + 0] = (float)((double)u[o] / ff.rho - 1.0);
#pragma unroll
    for (int d = 0; d < 3; ++d)
        t[5 * r + 1 + d] = (float)(((double)u[(1 + d) * stride + o] - ff.rho * vel[d]) / (ff.rho * ff.c));
    t[5 * r + 4] = (float)(((double)u[4 * stride + o] - Einf) / (ff.rho * ff.c * ff.c));
}

// Network outputs Y [rows][32] of rows row0.. -> conservative state in both state buffers.
template <typename S>
__global__ void k_warm_apply(const float *__restrict__ Y, size_t row0, int rows, uint32_t n_inst, gr_warm_norm nrm,
                             const double *__restrict__ wd, const double *__restrict__ inst_mach, double p_inf,
                             double T_inf, S *__restrict__ u0, S *__restrict__ u1, size_t stride) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= rows) return;
    const size_t r = row0 + q;
    const uint32_t i = (uint32_t)(r / n_inst), k = (uint32_t)(r % n_inst);
    double t[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) t[c] = (double)Y[(size_t)q * 32 + c] * nrm.out_std[c] + nrm.out_mean[c];
    // trusted region: the deviation in full up to wall distance trust[0], blended out
    // (smoothstep) by trust[1]
    if (nrm.trust[1] > 0) {
        const double d0 = nrm.trust[0], d1 = nrm.trust[1], d = wd[i];
        const double s = d <= d0 ? 0.0 : d >= d1 ? 1.0 : (d - d0) / (d1 - d0);
        const double wt = 1.0 - s * s * (3.0 - 2.0 * s);
#pragma unroll
        for (int c = 0; c < 5; ++c) t[c] *= wt;
    }
    const FarField ff = far_field(p_inf, T_inf);
    double vel[3], ke = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        vel[d] = inst_mach[3 * k + d] * ff.c;
        ke += vel[d] * vel[d];
    }
    double w[5];
    w[0] = ff.rho * (1.0 + t[0]);
#pragma unroll
    for (int d = 0; d < 3; ++d) w[1 + d] = ff.rho * vel[d] + ff.rho * ff.c * t[1 + d];
    w[4] = ff.E + 0.5 * ff.rho * ke + ff.rho * ff.c * ff.c * t[4];
    // a cell whose prediction is not a physical state (density or pressure below a quarter of
    // the far field's, or not finite) keeps the freestream
    const double pr = 0.4 * (w[4] - 0.5 * ((w[1] * w[1] + w[2] * w[2]) + w[3] * w[3]) / w[0]);
    const double p_inf_ = 0.4 * ff.E;
    if (!(w[0] > 0.25 * ff.rho) || !(pr > 0.25 * p_inf_) || !isfinite(w[1] + w[2] + w[3] + w[4])) {
        w[0] = ff.rho;
#pragma unroll
        for (int d = 0; d < 3; ++d) w[1 + d] = vel[d] * ff.rho;
        w[4] = ff.E + (0.5 * ff.rho) * ke;
    }
    S v[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) v[c] = (S)w[c];
    const size_t o = (size_t)i * kBatchLanes + k;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        u0[c * stride + o] = v[c];
        u1[c * stride + o] = v[c];
    }
}

// Lanes with reset[lane] set get vals[5*lane..] in every cell and ghost of both buffers.
template <typename S>
__global__ void k_lane_reset(S *__restrict__ u0, S *__restrict__ u1, size_t stride, size_t NG,
                             const int *__restrict__ reset, const double *__restrict__ vals) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NG * kBatchLanes) return;
    const uint32_t k = (uint32_t)(e & (kBatchLanes - 1));
    if (!reset[k]) return;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
        u0[c * stride + e] = (S)vals[5 * k + c];
        u1[c * stride + e] = (S)vals[5 * k + c];
    }
}

}  // namespace gr

using gr::guarded;
using gr::require;

struct gr_batch {
    gr_mesh *mesh = nullptr;
    gr_solver_config cfg{};
    uint32_t n_inst = 0;
    size_t stride = 0;   // elements per component block: pad64(N+G) * 64
    gr::DevArray<char> ubuf[2], grad, lim;
    gr::DevArray<char> slot_geo;      // per-slot flux and gradient records, ghost records
    size_t grec_off = 0, gnrm_off = 0;   // their offsets in values of the batch's precision
    gr::DevArray<double> inst_mach;   // 3 per lane
    std::vector<double> mach_h;       // 3 per instance
    double p_inf = 1.0, T_inf = 1.0 / 1.4;
    std::vector<int> region_kind;     // -1: no BC
    int cur = 0, ghost_src = 0;
    gr::DevArray<gr::BatchCtl> ctl;
    gr::DevArray<double> partials, partials2, hist, out;
    gr::DevArray<int> lane_flags;      // 64
    gr::DevArray<double> lane_vals;    // 5 per lane
    int hist_cap = 0;
    gr::BatchCtl *ctl_host = nullptr;   // pinned: [0] upload, [1],[2] polling
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int graph_iters = 10;
    int last_iters = 0;
    // launch schedule (the autotuner's choices): cells (waves) per block of the gradient and
    // flux kernels, whether the flux is capped at 3 waves per SIMD, and the flux's instance
    // groups (1, 2 or 4 sweeps over the cells, each with 64/groups lanes per cell)
    int cpb = 4;
    int flux_w3 = 1;
    int groups = 1;
    static constexpr int kMinCpb = 2;   // partial rows are allocated for the smallest choice
    static constexpr int kMaxGroups = 4;

    ~gr_batch() {
        if (mesh && mesh->stream) (void)hipStreamSynchronize(mesh->stream);
        drop_graphs();
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (ctl_host) (void)hipHostFree(ctl_host);
    }
    void drop_graphs() {
        for (auto &g : exec) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }
    uint32_t nblocks() const { return (mesh->h.n_cells + cpb - 1) / cpb; }
    // A group's row of a cell is 64/groups lanes of the state's width: kept at least one 128-byte
    // cache line (binary32 rows of 16 lanes half-fill the lines the other groups' sweeps fetch
    // again: measured 2.5x the flux traffic on the 64-flow C-grid).
    bool groups_ok(int g) const { return cfg.dtype == GR_F64 ? g <= 4 : g <= 2; }
    void get_schedule(int schedule[4]) const {
        schedule[0] = cpb;
        schedule[1] = flux_w3;
        schedule[2] = graph_iters;
        schedule[3] = groups;
    }
    // flux grid: groups sweeps of ceil(N / (cpb * groups)) blocks; its partial rows
    uint32_t flux_blocks() const {
        const uint32_t per = (uint32_t)(cpb * groups);
        return (uint32_t)groups * ((mesh->h.n_cells + per - 1) / per);
    }
    static size_t partial_rows(uint32_t n_cells) {   // the largest flux grid of any schedule
        return ((size_t)n_cells + kMinCpb - 1) / kMinCpb + kMaxGroups;
    }
    template <typename S> gr::BatchArgs<S> args(int b) const {
        const S *fr = (const S *)slot_geo.p;
        return gr::BatchArgs<S>{(const S *)ubuf[b].p, (S *)ubuf[1 - b].p, (S *)grad.p, (S *)lim.p, stride,
                                fr, fr + grec_off, fr + gnrm_off};
    }
    // Per-slot records of the tile kernels (k_slot_geometry), read by the batch kernels in place
    // of gathering the cell and face metrics face by face: the same values, cast the same way.
    template <typename S> void build_slot_geometry() {
        gr_mesh *m = mesh;
        const size_t slots = m->h.cell_faces_starts[m->h.n_cells];
        const size_t w = 16 / sizeof(S);
        grec_off = (slots * gr::kFluxRec + w - 1) / w * w;
        gnrm_off = grec_off + (slots * gr::kGradRec + w - 1) / w * w;
        slot_geo.alloc(std::max<size_t>(1, (gnrm_off + (size_t)m->h.n_ghosts * gr::kGhostRec) * sizeof(S)));
        if (m->h.n_cells) {
            S *fr = (S *)slot_geo.p;
            hipLaunchKernelGGL(gr::k_slot_geometry<S>, dim3(gr::blocks_for(m->h.n_cells)), dim3(gr::kBlock), 0, m->stream,
                               m->view(), fr, fr + grec_off, fr + gnrm_off);
        }
        GR_HIP(hipGetLastError());
    }
    gr::BcTable table() const {
        gr::BcTable t{};
        for (uint32_t r = 0; r < mesh->h.n_regions; ++r) {
            if (region_kind[r] < 0) continue;
            auto &R = t.r[t.n++];
            R.begin = mesh->h.ghost_cell_starts[r];
            R.end = mesh->h.ghost_cell_starts[r + 1];
            R.kind = region_kind[r];
            R.p = p_inf;
            R.T = T_inf;
        }
        return t;
    }
    // Stage 0: gradient+limiter, 1: flux+update, 2: per-instance reduction, of an iteration
    // reading buffer b.
    template <typename S, int CPB, int G> void launch_flux(int b, const gr::BatchCtl *cc) {
        gr_mesh *m = mesh;
        const dim3 grid(flux_blocks()), blk(CPB * gr::kBatchLanes);
#define GR_BATCH_FLUX(KERNEL, CORR)                                                                                      \
    hipLaunchKernelGGL((gr::KERNEL<S, CORR, CPB, G>), grid, blk, 0, m->stream, m->view(), args<S>(b), table(),          \
                       inst_mach.p, cfg.cfl, partials.p, cc)
        if (flux_w3) {
            if (cfg.corrected) GR_BATCH_FLUX(k_batch_flux_w3, true);
            else GR_BATCH_FLUX(k_batch_flux_w3, false);
        } else {
            if (cfg.corrected) GR_BATCH_FLUX(k_batch_flux, true);
            else GR_BATCH_FLUX(k_batch_flux, false);
        }
#undef GR_BATCH_FLUX
    }
    template <typename S, int CPB> void launch_cpb(int stage, int b, gr::BatchCtl *c) {
        gr_mesh *m = mesh;
        const gr::BatchCtl *cc = c;
        if (stage == 0) {
            hipLaunchKernelGGL((gr::k_batch_gradlim<S, CPB>), dim3(nblocks()), dim3(CPB * gr::kBatchLanes), 0, m->stream,
                               m->view(), args<S>(b), table(), inst_mach.p, cfg.limiter, cfg.corrected, cc);
            return;
        }
        if (groups == 2) {
            launch_flux<S, CPB, 2>(b, cc);
        } else if constexpr (sizeof(S) == 8) {
            if (groups == 4) launch_flux<S, CPB, 4>(b, cc);
            else launch_flux<S, CPB, 1>(b, cc);
        } else {
            launch_flux<S, CPB, 1>(b, cc);
        }
    }
    // Stage 0: gradient+limiter, 1: flux+update, 2: per-instance reduction, of an iteration
    // reading buffer b.
    template <typename S> void launch_stage(int stage, int b, gr::BatchCtl *c) {
        gr_mesh *m = mesh;
        if (stage < 2) {
            if (cpb == 2) launch_cpb<S, 2>(stage, b, c);
            else if (cpb == 8) launch_cpb<S, 8>(stage, b, c);
            else launch_cpb<S, 4>(stage, b, c);
        } else {
            hipLaunchKernelGGL(gr::k_batch_reduce, dim3(gr::kReduceBlocks), dim3(gr::kBlock), 0, m->stream, partials.p,
                               flux_blocks(), (const gr::BatchCtl *)c, partials2.p);
            hipLaunchKernelGGL(gr::k_batch_finalize, dim3(1), dim3(gr::kFinalParts * gr::kBatchLanes), 0, m->stream,
                               partials2.p, gr::kReduceBlocks, c, hist.p, out.p);
        }
    }
    void enqueue_stage(int stage, int b, gr::BatchCtl *c) {
        if (cfg.dtype == GR_F64) launch_stage<double>(stage, b, c);
        else launch_stage<float>(stage, b, c);
    }
    void enqueue(int b, gr::BatchCtl *c) {
        for (int stage = 0; stage < 3; ++stage) enqueue_stage(stage, b, c);
    }
    hipGraphExec_t graph(int b) {
        if (exec[b]) return exec[b];
        hipGraph_t g = nullptr;
        GR_HIP(hipStreamBeginCapture(mesh->stream, hipStreamCaptureModeRelaxed));
        try {
            for (int k = 0; k < graph_iters; ++k) enqueue((b + k) % 2, ctl.p);
        } catch (...) {
            (void)hipStreamEndCapture(mesh->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        GR_HIP(hipStreamEndCapture(mesh->stream, &g));
        hipError_t e = hipGraphInstantiate(&exec[b], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        GR_HIP(e);
        return exec[b];
    }
    // Algorithmic bytes per launch.  Geometry and connectivity are read once per cell for all
    // instances (wave-uniform loads); state traffic is per instance.
    void kernel_bytes(double out3[3]) const {
        const double N = mesh->h.n_cells, G = mesh->h.n_ghosts, F = mesh->h.n_faces;
        const double slots = mesh->h.cell_faces_starts[mesh->h.n_cells];
        const double su = (double)cfg.dtype, I = n_inst;
        out3[0] = I * su * (5 * N + 20 * N) + 8.0 * (3 * (N + G) + 6 * N + 3 * F + 3 * G) + 4.0 * (N + 1 + 2 * slots);
        out3[1] = I * su * (25 * N + 5 * N) + 8.0 * (3 * (N + G) + N + 7 * F) + 4.0 * (N + 1 + 2 * slots);
        out3[2] = 8.0 * 64 * (flux_blocks() + 2.0 * gr::kReduceBlocks);
    }
    // Conservative freestream state of lane k (padding lanes: instance 0).
    void freestream_values(uint32_t k, double val[5]) const {
        const double *mk = &mach_h[3 * (k < n_inst ? k : 0)];
        const double rho_v = p_inf / T_inf;
        const double c = std::sqrt((1.4 * p_inf) / rho_v);
        const double vel[3] = {mk[0] * c, mk[1] * c, mk[2] * c};
        const double e = p_inf / 0.4 + (0.5 * rho_v) * ((vel[0] * vel[0] + vel[1] * vel[1]) + vel[2] * vel[2]);
        val[0] = rho_v;
        for (int d = 0; d < 3; ++d) val[1 + d] = vel[d] * rho_v;
        val[4] = e;
    }
    // Lanes with flags[k] set (64 entries) back to their freestream, every cell and ghost of
    // both state buffers (padding lanes copy instance 0).
    template <typename S> void reset_lanes(const int *flags) {
        const size_t NG = (size_t)mesh->h.n_cells + mesh->h.n_ghosts;
        std::vector<double> vals(5 * gr::kBatchLanes);
        for (uint32_t k = 0; k < (uint32_t)gr::kBatchLanes; ++k) freestream_values(k, &vals[5 * k]);
        lane_flags.upload(flags, gr::kBatchLanes, mesh->stream);
        lane_vals.upload(vals.data(), vals.size(), mesh->stream);
        hipLaunchKernelGGL(gr::k_lane_reset<S>, dim3(gr::blocks_for(NG * gr::kBatchLanes)), dim3(gr::kBlock), 0,
                           mesh->stream, (S *)ubuf[0].p, (S *)ubuf[1].p, stride, NG, lane_flags.p, lane_vals.p);
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(mesh->stream));
    }
    template <typename S> void fill_freestream() {
        std::vector<int> all(gr::kBatchLanes, 1);
        reset_lanes<S>(all.data());
    }
    template <typename S> void set_instance(uint32_t k, const double *rho, const double *rhou, const double *rhoe) {
        const size_t NG = (size_t)mesh->h.n_cells + mesh->h.n_ghosts;
        std::vector<S> soa(5 * stride);
        // read-modify-write of the whole array keeps the other instances
        GR_HIP(hipMemcpyAsync(soa.data(), ubuf[cur].p, soa.size() * sizeof(S), hipMemcpyDeviceToHost, mesh->stream));
        GR_HIP(hipStreamSynchronize(mesh->stream));
        for (size_t i = 0; i < NG; ++i) {
            const size_t o = i * gr::kBatchLanes + k;
            soa[o] = (S)rho[i];
            for (int d = 0; d < 3; ++d) soa[(1 + d) * stride + o] = (S)rhou[3 * i + d];
            soa[4 * stride + o] = (S)rhoe[i];
        }
        for (int b = 0; b < 2; ++b)
            GR_HIP(hipMemcpyAsync(ubuf[b].p, soa.data(), soa.size() * sizeof(S), hipMemcpyHostToDevice, mesh->stream));
        GR_HIP(hipStreamSynchronize(mesh->stream));
        ghost_src = cur;
    }
    // Up to max_iter iterations of the active instances (all when active is null); inactive
    // instances keep their state.  Per-instance results go to the optional output arrays.
    void iterate(int max_iter, double tol, const double *reference, const int *active, int *iterations,
                 int *converged, int *diverged, int *launched) {
        gr_mesh *m = mesh;
        gr::BatchCtl &c0 = ctl_host[0];
        std::memset(&c0, 0, sizeof c0);
        c0.tol = tol;
        c0.max_iter = max_iter;
        c0.n_inst = (int)n_inst;
        c0.hist_cap = hist_cap;
        bool any = false;
        for (int k = 0; k < gr::kBatchLanes; ++k) {
            const bool on = k < (int)n_inst && (!active || active[k]);
            c0.done[k] = on ? 0 : 1;
            c0.ref[k] = (reference && k < (int)n_inst) ? reference[k] : 0.0;
            any = any || on;
        }
        if (!any) {
            for (uint32_t k = 0; k < n_inst; ++k) {
                if (iterations) iterations[k] = 0;
                if (converged) converged[k] = 0;
                if (diverged) diverged[k] = 0;
            }
            if (launched) *launched = 0;
            return;
        }
        if (cfg.use_graph) {
            graph(0);   // both parities captured before the first launch
            graph(1);
        }
        GR_HIP(hipMemcpyAsync(ctl.p, &c0, sizeof c0, hipMemcpyHostToDevice, m->stream));
        const int K = graph_iters;
        const int launches = (max_iter + K - 1) / K;
        int L = 0;
        for (; L < launches; ++L) {
            if (cfg.use_graph) {
                GR_HIP(hipGraphLaunch(graph(cur), m->stream));
            } else {
                for (int k = 0; k < K; ++k) enqueue((cur + k) % 2, ctl.p);
                GR_HIP(hipGetLastError());
            }
            gr::BatchCtl *poll = &ctl_host[1 + (L & 1)];
            GR_HIP(hipMemcpyAsync(poll, ctl.p, sizeof *poll, hipMemcpyDeviceToHost, m->stream));
            GR_HIP(hipEventRecord(ev[L & 1], m->stream));
            if (L >= 1) {
                GR_HIP(hipEventSynchronize(ev[(L - 1) & 1]));
                if (ctl_host[1 + ((L - 1) & 1)].stop) {
                    ++L;
                    break;
                }
            }
        }
        GR_HIP(hipStreamSynchronize(m->stream));
        gr::BatchCtl fin;
        GR_HIP(hipMemcpy(&fin, ctl.p, sizeof fin, hipMemcpyDeviceToHost));
        // buffer parity follows the longest-running instance; frozen lanes were copied along
        int ran = 0;
        for (uint32_t k = 0; k < n_inst; ++k) ran = std::max(ran, fin.iter[k]);
        if (ran > 0) {
            ghost_src = (cur + ran - 1) % 2;
            cur = (cur + ran) % 2;
        }
        last_iters = std::min(ran, hist_cap);
        for (uint32_t k = 0; k < n_inst; ++k) {
            if (iterations) iterations[k] = fin.iter[k];
            if (converged) converged[k] = fin.converged[k];
            if (diverged) diverged[k] = fin.diverged[k];
        }
        if (launched) *launched = std::max(ran, 0);
    }
    // Residual sum of each instance's current state (n_inst entries), state unchanged.
    void residual_norms(double *res) {
        gr_mesh *m = mesh;
        // a control block with nothing done and no stop, not recording (max_iter 0 marks it)
        gr::BatchCtl &c0 = ctl_host[0];
        std::memset(&c0, 0, sizeof c0);
        for (int k = 0; k < gr::kBatchLanes; ++k) c0.done[k] = k < (int)n_inst ? 0 : 1;
        GR_HIP(hipMemcpyAsync(ctl.p, &c0, sizeof c0, hipMemcpyHostToDevice, m->stream));
        // the iteration's gradient and flux launches (stages 0 and 1) on the current state
        for (int stage = 0; stage < 2; ++stage) enqueue_stage(stage, cur, ctl.p);
        hipLaunchKernelGGL(gr::k_batch_reduce, dim3(gr::kReduceBlocks), dim3(gr::kBlock), 0, m->stream, partials.p,
                           flux_blocks(), (const gr::BatchCtl *)nullptr, partials2.p);
        hipLaunchKernelGGL(gr::k_batch_finalize, dim3(1), dim3(gr::kFinalParts * gr::kBatchLanes), 0, m->stream,
                           partials2.p, gr::kReduceBlocks, (gr::BatchCtl *)nullptr, (double *)nullptr, out.p);
        GR_HIP(hipGetLastError());
        std::vector<double> o(gr::kBatchLanes);
        out.download(o.data(), o.size(), m->stream);
        GR_HIP(hipStreamSynchronize(m->stream));
        std::copy(o.begin(), o.begin() + n_inst, res);
        ghost_src = cur;
    }
    // All instances at once (one transfer each way): rho/rhoe [n_inst][N+G], rhou [n_inst][N+G][3].
    template <typename S> void set_all(const double *rho, const double *rhou, const double *rhoe) {
        const size_t NG = (size_t)mesh->h.n_cells + mesh->h.n_ghosts;
        std::vector<S> soa(5 * stride, S(0));
        for (uint32_t k = 0; k < (uint32_t)gr::kBatchLanes; ++k) {
            const size_t src = (size_t)(k < n_inst ? k : 0) * NG;   // padding lanes copy instance 0
            for (size_t i = 0; i < NG; ++i) {
                const size_t o = i * gr::kBatchLanes + k;
                soa[o] = (S)rho[src + i];
                for (int d = 0; d < 3; ++d) soa[(1 + d) * stride + o] = (S)rhou[3 * (src + i) + d];
                soa[4 * stride + o] = (S)rhoe[src + i];
            }
        }
        for (int b = 0; b < 2; ++b)
            GR_HIP(hipMemcpyAsync(ubuf[b].p, soa.data(), soa.size() * sizeof(S), hipMemcpyHostToDevice, mesh->stream));
        GR_HIP(hipStreamSynchronize(mesh->stream));
        ghost_src = cur;
    }
    template <typename S> void get_all(double *rho, double *rhou, double *rhoe) {
        gr_mesh *m = mesh;
        const size_t N = m->h.n_cells, NG = N + m->h.n_ghosts;
        if (ghost_src != cur && m->h.n_ghosts) {
            hipLaunchKernelGGL(gr::k_batch_ghosts<S>, dim3((m->h.n_ghosts + gr::kCellsPerBlock - 1) / gr::kCellsPerBlock),
                               dim3(gr::kBlock), 0, m->stream, m->view(), (S *)ubuf[ghost_src].p, stride, table(),
                               (const double *)inst_mach.p);
            GR_HIP(hipGetLastError());
        }
        std::vector<S> a(5 * stride), b(ghost_src != cur ? 5 * stride : 0);
        GR_HIP(hipMemcpyAsync(a.data(), ubuf[cur].p, a.size() * sizeof(S), hipMemcpyDeviceToHost, m->stream));
        if (!b.empty())
            GR_HIP(hipMemcpyAsync(b.data(), ubuf[ghost_src].p, b.size() * sizeof(S), hipMemcpyDeviceToHost, m->stream));
        GR_HIP(hipStreamSynchronize(m->stream));
        for (uint32_t k = 0; k < n_inst; ++k)
            for (size_t i = 0; i < NG; ++i) {
                const std::vector<S> &src = (i < N || b.empty()) ? a : b;
                const size_t o = i * gr::kBatchLanes + k, d0 = (size_t)k * NG + i;
                if (rho) rho[d0] = (double)src[o];
                if (rhou)
                    for (int d = 0; d < 3; ++d) rhou[3 * d0 + d] = (double)src[(1 + d) * stride + o];
                if (rhoe) rhoe[d0] = (double)src[4 * stride + o];
            }
    }
    template <typename S> void get_instance(uint32_t k, double *rho, double *rhou, double *rhoe) {
        gr_mesh *m = mesh;
        const size_t N = m->h.n_cells, NG = N + m->h.n_ghosts;
        if (ghost_src != cur && m->h.n_ghosts) {
            hipLaunchKernelGGL(gr::k_batch_ghosts<S>, dim3((m->h.n_ghosts + gr::kCellsPerBlock - 1) / gr::kCellsPerBlock),
                               dim3(gr::kBlock), 0, m->stream, m->view(), (S *)ubuf[ghost_src].p, stride, table(),
                               (const double *)inst_mach.p);
            GR_HIP(hipGetLastError());
        }
        std::vector<S> a(5 * stride), b(5 * stride);
        GR_HIP(hipMemcpyAsync(a.data(), ubuf[cur].p, a.size() * sizeof(S), hipMemcpyDeviceToHost, m->stream));
        GR_HIP(hipMemcpyAsync(b.data(), ubuf[ghost_src].p, b.size() * sizeof(S), hipMemcpyDeviceToHost, m->stream));
        GR_HIP(hipStreamSynchronize(m->stream));
        for (size_t i = 0; i < NG; ++i) {
            const std::vector<S> &src = i < N ? a : b;
            const size_t o = i * gr::kBatchLanes + k;
            if (rho) rho[i] = (double)src[o];
            if (rhou)
                for (int d = 0; d < 3; ++d) rhou[3 * i + d] = (double)src[(1 + d) * stride + o];
            if (rhoe) rhoe[i] = (double)src[4 * stride + o];
        }
    }
};

extern "C" {

int gr_batch_create(gr_mesh *m, uint32_t n_inst, const gr_solver_config *cfg, gr_batch **out) {
    return guarded([&] {
        require(m && cfg && out, "null argument");
        require(m->metrics_ready, "mesh metrics not computed (gr_mesh_compute_metrics)");
        require(n_inst >= 1 && n_inst <= (uint32_t)gr::kBatchLanes, "a batch holds 1..64 instances");
        require(cfg->dtype == GR_F64 || cfg->dtype == GR_F32, "batch dtype must be GR_F64 or GR_F32");
        require(cfg->cfl > 0 && cfg->limiter >= 0 && cfg->limiter <= 2, "bad cfl or limiter");
        require(m->h.n_regions <= (uint32_t)gr::kMaxRegions, "too many boundary regions");
        require(!m->h.one_dimensional, "the Euler solver needs a 3-D mesh");
        auto *b = new gr_batch();
        try {
            b->mesh = m;
            b->cfg = *cfg;
            b->n_inst = n_inst;
            b->stride = gr::pad64((size_t)m->h.n_cells + m->h.n_ghosts) * gr::kBatchLanes;
            const size_t blk = b->stride * (size_t)cfg->dtype;
            for (auto &u : b->ubuf) {
                u.alloc(5 * blk);
                GR_HIP(hipMemsetAsync(u.p, 0, u.n, m->stream));
            }
            b->grad.alloc(15 * blk);
            b->lim.alloc(5 * blk);
            GR_HIP(hipMemsetAsync(b->grad.p, 0, b->grad.n, m->stream));
            GR_HIP(hipMemsetAsync(b->lim.p, 0, b->lim.n, m->stream));
            if (cfg->dtype == GR_F64) b->build_slot_geometry<double>();
            else b->build_slot_geometry<float>();
            b->inst_mach.alloc(3 * gr::kBatchLanes);
            b->mach_h.assign(3 * n_inst, 0.0);
            b->region_kind.assign(m->h.n_regions, -1);
            b->ctl.alloc(1);
            b->partials.alloc(gr_batch::partial_rows(m->h.n_cells) * gr::kBatchLanes);
            b->partials2.alloc((size_t)gr::kReduceBlocks * gr::kBatchLanes);
            b->out.alloc(gr::kBatchLanes);
            b->lane_flags.alloc(gr::kBatchLanes);
            b->lane_vals.alloc(5 * gr::kBatchLanes);
            b->hist_cap = 1 << 14;
            b->hist.alloc((size_t)b->hist_cap * gr::kBatchLanes);
            GR_HIP(hipHostMalloc((void **)&b->ctl_host, 3 * sizeof(gr::BatchCtl)));
            for (auto &e : b->ev) GR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            GR_HIP(hipStreamSynchronize(m->stream));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

void gr_batch_free(gr_batch *b) { delete b; }

int gr_batch_set_bc(gr_batch *b, const char *region, int kind) {
    return guarded([&] {
        require(b && region, "null argument");
        require(kind == GR_BC_SLIP || kind == GR_BC_FARFIELD, "unknown boundary kind");
        uint32_t r = 0;
        for (; r < b->mesh->h.n_regions; ++r)
            if (b->mesh->h.region_names[r] == region) break;
        if (r == b->mesh->h.n_regions) throw gr::Error(GR_ERR_NOT_FOUND, std::string("region '") + region + "' not found in mesh");
        b->region_kind[r] = kind;
        b->drop_graphs();
    });
}

int gr_batch_set_freestream(gr_batch *b, const double *mach, double pressure, double temperature) {
    return guarded([&] {
        require(b && mach, "null argument");
        require(pressure > 0 && temperature > 0, "pressure and temperature must be positive");
        b->mach_h.assign(mach, mach + 3 * b->n_inst);
        b->p_inf = pressure;
        b->T_inf = temperature;
        std::vector<double> lanes(3 * gr::kBatchLanes);
        for (uint32_t k = 0; k < (uint32_t)gr::kBatchLanes; ++k)
            for (int d = 0; d < 3; ++d) lanes[3 * k + d] = b->mach_h[3 * (k < b->n_inst ? k : 0) + d];
        b->inst_mach.upload(lanes.data(), lanes.size(), b->mesh->stream);
        if (b->cfg.dtype == GR_F64) b->fill_freestream<double>();
        else b->fill_freestream<float>();
        b->ghost_src = b->cur;
        b->drop_graphs();
    });
}

int gr_batch_set_state(gr_batch *b, uint32_t inst, const double *rho, const double *rhou, const double *rhoe) {
    return guarded([&] {
        require(b && rho && rhou && rhoe, "null argument");
        require(inst < b->n_inst, "instance index out of range");
        if (b->cfg.dtype == GR_F64) b->set_instance<double>(inst, rho, rhou, rhoe);
        else b->set_instance<float>(inst, rho, rhou, rhoe);
    });
}

int gr_batch_get_state(gr_batch *b, uint32_t inst, double *rho, double *rhou, double *rhoe) {
    return guarded([&] {
        require(b != nullptr, "null argument");
        require(inst < b->n_inst, "instance index out of range");
        if (b->cfg.dtype == GR_F64) b->get_instance<double>(inst, rho, rhou, rhoe);
        else b->get_instance<float>(inst, rho, rhou, rhoe);
    });
}

int gr_batch_set_states(gr_batch *b, const double *rho, const double *rhou, const double *rhoe) {
    return guarded([&] {
        require(b && rho && rhou && rhoe, "null argument");
        if (b->cfg.dtype == GR_F64) b->set_all<double>(rho, rhou, rhoe);
        else b->set_all<float>(rho, rhou, rhoe);
    });
}

int gr_batch_get_states(gr_batch *b, double *rho, double *rhou, double *rhoe) {
    return guarded([&] {
        require(b != nullptr, "null argument");
        if (b->cfg.dtype == GR_F64) b->get_all<double>(rho, rhou, rhoe);
        else b->get_all<float>(rho, rhou, rhoe);
    });
}

int gr_batch_iterate(gr_batch *b, int max_iter, double tol, const double *reference, int *iterations,
                     int *converged, int *launched) {
    return guarded([&] {
        require(b != nullptr, "null argument");
        require(max_iter > 0, "max_iter must be positive");
        std::vector<int> div(b->n_inst, 0);
        b->iterate(max_iter, tol, reference, nullptr, iterations, converged, div.data(), launched);
        for (uint32_t k = 0; k < b->n_inst; ++k)
            if (div[k]) throw gr::Error(GR_ERR_DIVERGED, "instance " + std::to_string(k) + " diverged");
    });
}

int gr_batch_iterate_masked(gr_batch *b, int max_iter, double tol, const double *reference, const int *active,
                            int *iterations, int *converged, int *diverged, int *launched) {
    return guarded([&] {
        require(b != nullptr, "null argument");
        require(max_iter > 0, "max_iter must be positive");
        b->iterate(max_iter, tol, reference, active, iterations, converged, diverged, launched);
    });
}

int gr_batch_history(const gr_batch *b, uint32_t inst, double *hist, int n) {
    return guarded([&] {
        require(b && hist, "null argument");
        require(inst < b->n_inst, "instance index out of range");
        require(n >= 0 && n <= b->last_iters, "history request longer than the last run");
        if (n) GR_HIP(hipMemcpy(hist, b->hist.p + (size_t)inst * b->hist_cap, n * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int gr_batch_residual_norms(gr_batch *b, double *out) {
    return guarded([&] {
        require(b && out, "null argument");
        b->residual_norms(out);
    });
}

int gr_batch_profile(gr_batch *b, int n_iter, double ms[3]) {
    return guarded([&] {
        require(b && ms, "null argument");
        require(n_iter > 0 && n_iter <= 10000, "n_iter must be in 1..10000");
        gr_mesh *m = b->mesh;
        // Each stage n_iter times back to back on the current state, events around each run
        // only (per-launch events would add the cost of an event pair to every duration).  The
        // control block counts the reductions' iterations and is reset by the next iterate.
        hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
        struct Guard {
            hipEvent_t *e;
            ~Guard() {
                for (int k = 0; k < 4; ++k)
                    if (e[k]) (void)hipEventDestroy(e[k]);
            }
        } guard{ev};
        for (auto &e : ev) GR_HIP(hipEventCreate(&e));
        // the flux launches write buffer 1-cur, which holds the state the ghost values of a
        // download are built from: saved here and restored after the runs
        gr::DevArray<char> saved;
        saved.alloc(b->ubuf[1 - b->cur].n);
        GR_HIP(hipMemcpyAsync(saved.p, b->ubuf[1 - b->cur].p, saved.n, hipMemcpyDeviceToDevice, m->stream));
        gr::BatchCtl &c0 = b->ctl_host[0];
        std::memset(&c0, 0, sizeof c0);
        c0.max_iter = n_iter + 1;
        c0.n_inst = (int)b->n_inst;
        c0.hist_cap = b->hist_cap;
        for (int k = 0; k < gr::kBatchLanes; ++k) c0.done[k] = k < (int)b->n_inst ? 0 : 1;
        GR_HIP(hipMemcpyAsync(b->ctl.p, &c0, sizeof c0, hipMemcpyHostToDevice, m->stream));
        GR_HIP(hipEventRecord(ev[0], m->stream));
        for (int stage = 0; stage < 3; ++stage) {
            for (int k = 0; k < n_iter; ++k) b->enqueue_stage(stage, b->cur, b->ctl.p);
            GR_HIP(hipEventRecord(ev[stage + 1], m->stream));
        }
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(m->stream));
        for (int j = 0; j < 3; ++j) {
            float t = 0;
            GR_HIP(hipEventElapsedTime(&t, ev[j], ev[j + 1]));
            ms[j] = t / n_iter;
        }
        b->last_iters = 0;
        GR_HIP(hipMemcpyAsync(b->ubuf[1 - b->cur].p, saved.p, saved.n, hipMemcpyDeviceToDevice, m->stream));
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

int gr_batch_autotune(gr_batch *b, int n_iter, int schedule[4]) {
    return guarded([&] {
        require(b && schedule, "null argument");
        require(n_iter > 0 && n_iter <= 1000, "n_iter must be in 1..1000");
        gr_mesh *m = b->mesh;
        // The timed runs write both state buffers (and the gradient and limiter rows): both are
        // saved here and restored at the end, so the batch's state is unchanged.
        hipEvent_t ev[2] = {nullptr, nullptr};
        struct Guard {
            hipEvent_t *e;
            ~Guard() {
                for (int k = 0; k < 2; ++k)
                    if (e[k]) (void)hipEventDestroy(e[k]);
            }
        } guard{ev};
        for (auto &e : ev) GR_HIP(hipEventCreate(&e));
        gr::DevArray<char> saved[2];
        for (int k = 0; k < 2; ++k) {
            saved[k].alloc(b->ubuf[k].n);
            GR_HIP(hipMemcpyAsync(saved[k].p, b->ubuf[k].p, saved[k].n, hipMemcpyDeviceToDevice, m->stream));
        }
        gr::BatchCtl &c0 = b->ctl_host[0];
        auto reset_ctl = [&](int max_iter) {
            std::memset(&c0, 0, sizeof c0);
            c0.max_iter = max_iter;
            c0.n_inst = (int)b->n_inst;
            c0.hist_cap = b->hist_cap;
            for (int k = 0; k < gr::kBatchLanes; ++k) c0.done[k] = k < (int)b->n_inst ? 0 : 1;
            GR_HIP(hipMemcpyAsync(b->ctl.p, &c0, sizeof c0, hipMemcpyHostToDevice, m->stream));
        };
        auto elapsed = [&]() {
            GR_HIP(hipGetLastError());
            GR_HIP(hipEventSynchronize(ev[1]));
            float t = 0;
            GR_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
            return (double)t;
        };
        // 1. cells per block x flux occupancy x instance groups: gradient + flux launches back
        // to back
        reset_ctl(1 << 30);
        double best = 1e300;
        int best_cpb = b->cpb, best_w3 = b->flux_w3, best_groups = b->groups;
        for (int cpb : {2, 4, 8}) {
            for (int w3 : {1, 0}) {
                for (int groups : {1, 2, 4}) {
                    if (!b->groups_ok(groups)) continue;
                    b->cpb = cpb;
                    b->flux_w3 = w3;
                    b->groups = groups;
                    b->enqueue_stage(0, b->cur, b->ctl.p);   // first use of the variant, untimed
                    b->enqueue_stage(1, b->cur, b->ctl.p);
                    GR_HIP(hipEventRecord(ev[0], m->stream));
                    for (int k = 0; k < n_iter; ++k) {
                        b->enqueue_stage(0, b->cur, b->ctl.p);
                        b->enqueue_stage(1, b->cur, b->ctl.p);
                    }
                    GR_HIP(hipEventRecord(ev[1], m->stream));
                    const double t = elapsed();
                    if (t < best) {
                        best = t;
                        best_cpb = cpb;
                        best_w3 = w3;
                        best_groups = groups;
                    }
                }
            }
        }
        b->cpb = best_cpb;
        b->flux_w3 = best_w3;
        b->groups = best_groups;
        // 2. graph length: the same number of iterations replayed as the solve loop replays them
        // (graph, copy of the control block to the host, event)
        const int total = 40;
        double best_g = 1e300;
        int best_k = b->graph_iters;
        for (int K : {4, 10, 20, 40}) {   // even: a replay returns to the buffer it started from
            b->graph_iters = K;
            b->drop_graphs();
            hipGraphExec_t g = b->graph(b->cur);
            reset_ctl(1 << 30);
            GR_HIP(hipGraphLaunch(g, m->stream));   // untimed first replay
            GR_HIP(hipEventRecord(ev[0], m->stream));
            for (int r = 0; r < total / K; ++r) {
                GR_HIP(hipGraphLaunch(g, m->stream));
                GR_HIP(hipMemcpyAsync(&b->ctl_host[1], b->ctl.p, sizeof c0, hipMemcpyDeviceToHost, m->stream));
            }
            GR_HIP(hipEventRecord(ev[1], m->stream));
            const double t = elapsed();
            if (t < best_g) {
                best_g = t;
                best_k = K;
            }
        }
        b->graph_iters = best_k;
        b->drop_graphs();
        for (int k = 0; k < 2; ++k)
            GR_HIP(hipMemcpyAsync(b->ubuf[k].p, saved[k].p, saved[k].n, hipMemcpyDeviceToDevice, m->stream));
        GR_HIP(hipStreamSynchronize(m->stream));
        b->last_iters = 0;
        b->get_schedule(schedule);
    });
}

int gr_batch_get_schedule(const gr_batch *b, int schedule[4]) {
    return guarded([&] {
        require(b && schedule, "null argument");
        b->get_schedule(schedule);
    });
}

int gr_batch_set_schedule(gr_batch *b, const int schedule[4]) {
    return guarded([&] {
        require(b && schedule, "null argument");
        require(schedule[0] == 2 || schedule[0] == 4 || schedule[0] == 8, "cells per block must be 2, 4 or 8");
        require(schedule[1] == 0 || schedule[1] == 1, "flux occupancy flag must be 0 or 1");
        require(schedule[2] >= 2 && schedule[2] <= 200 && schedule[2] % 2 == 0, "graph length must be even, 2..200");
        require(schedule[3] == 1 || schedule[3] == 2 || schedule[3] == 4, "instance groups must be 1, 2 or 4");
        require(b->groups_ok(schedule[3]), "binary32 batches take 1 or 2 instance groups");
        b->cpb = schedule[0];
        b->flux_w3 = schedule[1];
        b->graph_iters = schedule[2];
        b->groups = schedule[3];
        b->drop_graphs();
    });
}

int gr_batch_kernel_bytes(const gr_batch *b, double bytes[3]) {
    return guarded([&] {
        require(b && bytes, "null argument");
        b->kernel_bytes(bytes);
    });
}

int gr_batch_force(gr_batch *b, const char *region, double *F) {
    return guarded([&] {
        require(b && region && F, "null argument");
        uint32_t r = 0;
        for (; r < b->mesh->h.n_regions; ++r)
            if (b->mesh->h.region_names[r] == region) break;
        if (r == b->mesh->h.n_regions) throw gr::Error(GR_ERR_NOT_FOUND, std::string("region '") + region + "' not found in mesh");
        for (uint32_t k = 0; k < b->n_inst; ++k) {
            int rc = gr_internal_wall_force(b->mesh, b->ubuf[b->cur].p, b->cfg.dtype, b->stride, gr::kBatchLanes, k, r,
                                            F + 3 * k);
            if (rc != GR_OK) throw gr::Error(rc, gr_last_error());
        }
    });
}

int gr_batch_set_graph_iters(gr_batch *b, int iters) {
    return guarded([&] {
        require(b != nullptr, "null argument");
        require(iters >= 2 && iters <= 200 && iters % 2 == 0, "iterations per graph must be even, 2..200");
        b->graph_iters = iters;
        b->drop_graphs();
    });
}

int gr_batch_warm_rows(gr_batch *b, const double *wall_distance, const float origin[3], float *x, float *t) {
    return guarded([&] {
        require(b && wall_distance && origin && x && t, "null argument");
        gr_mesh *m = b->mesh;
        const uint32_t N = m->h.n_cells;
        const size_t rows = (size_t)N * b->n_inst;
        gr::DevArray<double> wd;
        gr::DevArray<float> X, T;
        wd.alloc(N);
        X.alloc(rows * 6);
        T.alloc(rows * 5);
        wd.upload(wall_distance, N, m->stream);
        gr_warm_norm none{};
        for (int k = 0; k < 3; ++k) none.origin[k] = origin[k];
        hipLaunchKernelGGL(gr::k_warm_features<false>, dim3(gr::blocks_for(rows)), dim3(gr::kBlock), 0, m->stream,
                           m->view(), wd.p, (const double *)b->inst_mach.p, b->n_inst, (size_t)0, (int)rows, none,
                           (void *)X.p);
        if (b->cfg.dtype == GR_F64)
            hipLaunchKernelGGL(gr::k_warm_targets<double>, dim3(gr::blocks_for(rows)), dim3(gr::kBlock), 0, m->stream,
                               (const double *)b->ubuf[b->cur].p, b->stride, N, b->n_inst, (const double *)b->inst_mach.p,
                               b->p_inf, b->T_inf, T.p);
        else
            hipLaunchKernelGGL(gr::k_warm_targets<float>, dim3(gr::blocks_for(rows)), dim3(gr::kBlock), 0, m->stream,
                               (const float *)b->ubuf[b->cur].p, b->stride, N, b->n_inst, (const double *)b->inst_mach.p,
                               b->p_inf, b->T_inf, T.p);
        GR_HIP(hipGetLastError());
        X.download(x, rows * 6, m->stream);
        T.download(t, rows * 5, m->stream);
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

int gr_batch_warm_start(gr_batch *b, gr_mlp *n, const double *wall_distance, const gr_warm_norm *norm,
                        double max_ratio, double *fs_residual, int *used) {
    return guarded([&] {
        require(b && n && wall_distance && norm && fs_residual, "null argument");
        require(gr::mlp_n_in(n) == 6 && gr::mlp_n_out(n) == 5, "the warm-start network must have n_in 6 and n_out 5");
        for (int c = 0; c < 6; ++c) require(norm->in_std[c] > 0, "input scales must be positive");
        for (int c = 0; c < 5; ++c) require(norm->out_std[c] > 0, "output scales must be positive");
        require(norm->trust[1] <= 0 || norm->trust[1] > norm->trust[0], "trust radii must increase");
        require(max_ratio > 0, "max_ratio must be positive");
        require(!b->mach_h.empty(), "far field not set (gr_batch_set_freestream)");
        gr_mesh *m = b->mesh;
        const uint32_t N = m->h.n_cells, I = b->n_inst;
        // cold state and its residuals: the normalisation and the fallback
        if (b->cfg.dtype == GR_F64) b->fill_freestream<double>();
        else b->fill_freestream<float>();
        b->ghost_src = b->cur;
        b->residual_norms(fs_residual);
        // prediction of every (cell, instance) row
        gr::DevArray<double> wd;
        wd.alloc(N);
        wd.upload(wall_distance, N, m->stream);
        const gr_warm_norm nrm = *norm;
        const double *mach = (const double *)b->inst_mach.p;
        const auto fill = [&](void *X, size_t row0, int rows) {
            hipLaunchKernelGGL(gr::k_warm_features<true>, dim3(gr::blocks_for((size_t)rows)),
                               dim3(gr::kBlock), 0, m->stream, m->view(), wd.p, mach, I, row0, rows, nrm, X);
        };
        const auto consume = [&](const float *Y, size_t row0, int rows) {
            if (b->cfg.dtype == GR_F64)
                hipLaunchKernelGGL(gr::k_warm_apply<double>, dim3(gr::blocks_for(rows)), dim3(gr::kBlock), 0, m->stream,
                                   Y, row0, rows, I, nrm, wd.p, mach, b->p_inf, b->T_inf, (double *)b->ubuf[0].p,
                                   (double *)b->ubuf[1].p, b->stride);
            else
                hipLaunchKernelGGL(gr::k_warm_apply<float>, dim3(gr::blocks_for(rows)), dim3(gr::kBlock), 0, m->stream,
                                   Y, row0, rows, I, nrm, wd.p, mach, b->p_inf, b->T_inf, (float *)b->ubuf[0].p,
                                   (float *)b->ubuf[1].p, b->stride);
        };
        gr::mlp_run_device(n, (size_t)N * I, m->stream, fill, consume);
        // the plain solver judges the prediction
        std::vector<double> pred(I);
        b->residual_norms(pred.data());
        std::vector<int> reset(gr::kBatchLanes, 0);
        bool any = false;
        for (uint32_t k = 0; k < I; ++k) {
            const bool keep = std::isfinite(pred[k]) && pred[k] < max_ratio * fs_residual[k];
            if (used) used[k] = keep ? 1 : 0;
            reset[k] = keep ? 0 : 1;
            any = any || !keep;
        }
        if (any) {
            if (b->cfg.dtype == GR_F64) b->reset_lanes<double>(reset.data());
            else b->reset_lanes<float>(reset.data());
        }
        b->ghost_src = b->cur;
    });
}

int gr_batch_reset_freestream(gr_batch *b, const int *mask) {
    return guarded([&] {
        require(b && mask, "null argument");
        require(!b->mach_h.empty(), "far field not set (gr_batch_set_freestream)");
        std::vector<int> flags(gr::kBatchLanes, 0);
        for (uint32_t k = 0; k < b->n_inst; ++k) flags[k] = mask[k] ? 1 : 0;
        if (b->cfg.dtype == GR_F64) b->reset_lanes<double>(flags.data());
        else b->reset_lanes<float>(flags.data());
        b->ghost_src = b->cur;
    });
}

}  // extern "C"

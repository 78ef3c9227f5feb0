// Device kernels of the finite-volume residual path (gradient, limiter, boundary conditions,
// Euler flux + local-time-step update, reductions).  Templated on the state type S:
// double (binary64 throughout) or float (binary32 state/gradient/limiter/flux arithmetic
// with binary64 geometry and binary64 residual norms).
//
// The binary64 instantiations evaluate the reference's expressions (src/fvmops.cu:51-304,
// src/models/euler.cu:236-520) in the reference's order of operations; the library is built
// without floating-point contraction, so they agree with a host evaluation of the same
// expressions bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include "device.hpp"

namespace gr {

// Non-temporal store of a batched kernel's output (new state, gradient, limiter): the next
// kernel reads it long after the L2 has turned over, and keeping it out of the L2 leaves room
// for the neighbour gathers (the batched flux re-reads neighbours 192 cells back on the
// C-grid).  Measured: batched flux 375 -> 360 us (binary64), 131 -> 123 us (mixed); the
// single-flow tile kernels got slower with it (42 -> 45 us), so they store normally.
template <typename T>
__device__ inline void store_nt(T *p, T v) {
    __builtin_nontemporal_store(v, p);
}


// Device-side control block of the iteration loop (one per solver).
struct SolverCtl {
    double ref;        // normalisation; <= 0 takes the first residual
    double tol;        // stop when residual/ref < tol (tol <= 0: never)
    double last;       // last residual sum
    double ratio;      // last residual/ref
    int iter;          // iterations completed
    int max_iter;      // stop after this many
    int stop;          // set when converged, diverged or max_iter reached
    int converged;
    int diverged;
    int hist_cap;      // capacity of the history array
    int pending;       // set by a flux launch: its tile partials await the sum
};

template <typename S> struct Tol;
template <> struct Tol<double> { static constexpr double v = 1e-12; };   // FVM_TOL, include/datatypes.cuh:45
template <> struct Tol<float> { static constexpr float v = 1e-6f; };     // FVM_TOL, include/datatypes.cuh:39

// Division and square root of the state type.  binary64: correctly rounded (the parity
// path).  binary32 (mixed precision, no exactness claim): the hardware reciprocal and square
// root (1 ulp), which replace the multi-instruction IEEE sequences.
__device__ inline double dsqrt(double x) { return sqrt(x); }
__device__ inline float dsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ inline double fdiv(double a, double b) { return a / b; }
__device__ inline float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
// a/b of binary64 operands (geometry) rounded to the state type.
template <typename S> __device__ inline S geo_ratio(double a, double b);
template <> __device__ inline double geo_ratio<double>(double a, double b) { return a / b; }
template <> __device__ inline float geo_ratio<float>(double a, double b) { return fdiv((float)a, (float)b); }
template <typename S> __device__ inline S sgn_t(double x) { return (S)((double)(x > 0) - (double)(x < 0)); }

// LIMITER_FUNC of src/fvmops.cu:118-125.
template <typename S>
__device__ inline S limiter_func(S y, int kind) {
    if (kind == GR_LIM_BARTH_JESPERSEN) return fmax(S(0), fmin(S(1), y));
    if (kind == GR_LIM_VENKATAKRISHNAN) return fdiv(y * y + S(2) * y, y * y + y + S(2));
    return (y < S(2)) ? (y - (S(0.25) * y) * y) : S(1);
}

// limiter() of src/fvmops.cu:129-169 for one face offset d.
template <typename S>
__device__ inline S limiter_face(S u, S gx, S gy, S gz, S umin, S umax, S dx, S dy, S dz, int kind) {
    const S delta = (gx * dx + gy * dy) + gz * dz;
    const S d_u = umax - u;
    const S d_d = umin - u;
    const S dmm2 = (d_u - d_d) * (d_u - d_d);
    const S len = dsqrt((dx * dx + dy * dy) + dz * dz);
    S k3v = S(1) * len;
    k3v *= (k3v * k3v);
    S sig;
    if (dmm2 <= k3v) {
        sig = S(1);
    } else if (dmm2 <= S(2) * k3v) {
        const S y = fdiv(dmm2, k3v) - S(1);
        sig = ((S(2) * y) * y) * y - (S(3) * y) * y + S(1);
    } else {
        sig = S(0);
    }
    // the reference's two branches (delta > tol: d_u / delta, delta < -tol: d_d / delta) as one
    // division of the selected numerator: same values, and a wave whose lanes take both
    // branches divides and evaluates the limiter function once instead of twice
    S lim = S(1);
    if (delta > Tol<S>::v || delta < -Tol<S>::v) lim = limiter_func<S>(fdiv(delta > Tol<S>::v ? d_u : d_d, delta), kind);
    lim *= S(0.8);
    return sig + (S(1) - sig) * lim;
}

// tensor (columns) times vector with the symmetric LSQ matrix of cell i, include/datatypes.cuh:296-302.
template <typename S>
__device__ inline void lsq_apply(const MeshView &m, uint32_t i, const S g[3], S out[3]) {
    const size_t ms = m.mat_stride;
    const S xx = (S)m.mat[i], yy = (S)m.mat[ms + i], zz = (S)m.mat[2 * ms + i];
    const S xy = (S)m.mat[3 * ms + i], xz = (S)m.mat[4 * ms + i], yz = (S)m.mat[5 * ms + i];
    out[0] = (xx * g[0] + xy * g[1]) + xz * g[2];
    out[1] = (xy * g[0] + yy * g[1]) + yz * g[2];
    out[2] = (xz * g[0] + yz * g[1]) + zz * g[2];
}

template <typename S, int NC>
struct GradLimArgs {
    const S *u[NC];   // component values (N+G)
    S *g[3 * NC];     // gradients: component c, direction k at [3c+k]
    S *l[NC];         // limiters
};

// Fused LSQ gradient, neighbour min/max and limiter of NC components for interior cell i
// (Scalar/VectorVarComputeGradient + Scalar/VectorVarComputeLimiters, src/fvmops.cu:51-266).
// DO_GRAD=false reads the stored gradient; DO_LIM=false skips the limiter.
template <typename S, int NC, bool DO_GRAD, bool DO_LIM>
__global__ __launch_bounds__(kBlock) void k_grad_lim(MeshView m, GradLimArgs<S, NC> a, int kind,
                                                     const SolverCtl *ctl) {
    if (ctl && ctl->stop) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N) return;
    const uint32_t s = m.cfs[i], e = m.cfs[i + 1];
    const double xi = m.cx[i], yi = m.cy[i], zi = m.cz[i];
    S ui[NC], umin[NC], umax[NC], g[NC][3];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        ui[c] = a.u[c][i];
        umin[c] = ui[c];
        umax[c] = ui[c];
        g[c][0] = g[c][1] = g[c][2] = S(0);
    }
    for (uint32_t q = s; q < e; ++q) {
        const uint32_t j = m.cc[q];
        S d[3] = {(S)(m.cx[j] - xi), (S)(m.cy[j] - yi), (S)(m.cz[j] - zi)};
        const S w2 = S(1) / ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const S uj = a.u[c][j];
            if (DO_GRAD) {
                const S du = uj - ui[c];
#pragma unroll
                for (int k = 0; k < 3; ++k) g[c][k] += (d[k] * du) * w2;
            }
            umin[c] = fmin(umin[c], uj);
            umax[c] = fmax(umax[c], uj);
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        if (DO_GRAD) {
            S out[3];
            lsq_apply<S>(m, i, g[c], out);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                g[c][k] = out[k];
                a.g[3 * c + k][i] = out[k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 3; ++k) g[c][k] = a.g[3 * c + k][i];
        }
    }
    if (!DO_LIM) return;
    S lim[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) lim[c] = S(2);
    for (uint32_t q = s; q < e; ++q) {
        const uint32_t f = m.cf[q];
        const S dx = (S)(m.fcx[f] - xi), dy = (S)(m.fcy[f] - yi), dz = (S)(m.fcz[f] - zi);
#pragma unroll
        for (int c = 0; c < NC; ++c)
            lim[c] = fmin(lim[c], limiter_face<S>(ui[c], g[c][0], g[c][1], g[c][2], umin[c], umax[c], dx, dy, dz, kind));
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) a.l[c][i] = lim[c];
}

// fvmScalGradLimBc / fvmVecGradLimBc: ghost gradient and limiter = those of its cell.
template <typename S, int NC>
__global__ void k_gradlim_ghost(MeshView m, GradLimArgs<S, NC> a) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= m.G) return;
    const uint32_t c = m.N + g;
    const uint32_t o = m.cc[m.cfs[c]];
#pragma unroll
    for (int k = 0; k < 3 * NC; ++k) a.g[k][c] = a.g[k][o];
#pragma unroll
    for (int k = 0; k < NC; ++k) a.l[k][c] = a.l[k][o];
}

// Boundary regions applied by one launch.
struct BcRegion {
    uint32_t begin, end;   // ghost index range [begin, end) relative to the first ghost
    int kind;              // gr_bc_kind
    double mach[3], p, T;
};
struct BcTable {
    int n;
    BcRegion r[kMaxRegions];
};

template <typename S>
struct EulerArgs {
    S *u[5];          // rho, rhou x, y, z, rhoe (N+G); written by the BC kernel
    S *un[5];         // next state (Jacobi)
    const S *g[15];   // gradients [3c+k]
    const S *l[5];    // limiters
};

// euler_slip_bc / euler_inletoutlet_bc (src/models/euler.cu:412-520) for ghosts
// [g0, g0+count), each with the region of the table that contains it.
template <typename S>
__global__ void k_euler_bc(MeshView m, EulerArgs<S> a, BcTable t, uint32_t g0, uint32_t count,
                           const SolverCtl *ctl) {
    if (ctl && ctl->stop) return;
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= count) return;
    const uint32_t g = g0 + tid;
    int r = -1;
    for (int k = 0; k < t.n; ++k)
        if (g >= t.r[k].begin && g < t.r[k].end) r = k;
    if (r < 0) return;
    const BcRegion &R = t.r[r];
    const uint32_t c0 = m.N + g;
    const uint32_t s = m.cfs[c0];
    const uint32_t c1 = m.cc[s], f = m.cf[s];
    const double d1x = m.fcx[f] - m.cx[c1], d1y = m.fcy[f] - m.cy[c1], d1z = m.fcz[f] - m.cz[c1];
    const double sgd = (double)((m.fnx[f] * d1x + m.fny[f] * d1y) + m.fnz[f] * d1z > 0) -
                       (double)((m.fnx[f] * d1x + m.fny[f] * d1y) + m.fnz[f] * d1z < 0);
    const S n[3] = {(S)(m.fnx[f] * sgd), (S)(m.fny[f] * sgd), (S)(m.fnz[f] * sgd)};
    const S r1 = a.u[0][c1];
    const S ru1[3] = {a.u[1][c1], a.u[2][c1], a.u[3][c1]};
    const S e1 = a.u[4][c1];
    if (R.kind == GR_BC_SLIP) {
        const S un = (n[0] * ru1[0] + n[1] * ru1[1]) + n[2] * ru1[2];
#pragma unroll
        for (int k = 0; k < 3; ++k) a.u[1 + k][c0] = ru1[k] - (n[k] * S(2)) * un;
        a.u[0][c0] = r1;
        a.u[4][c0] = e1;
        return;
    }
    const S u[3] = {ru1[0] / r1, ru1[1] / r1, ru1[2] / r1};
    const S p = S(0.4) * (e1 - (S(0.5) / r1) * ((ru1[0] * ru1[0] + ru1[1] * ru1[1]) + ru1[2] * ru1[2]));
    const S tt = p / (S(1) * r1);
    const S mach_i = dsqrt(((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]) / ((S(1.4) * p) / r1));
    const S p_in = (S)R.p, T_in = (S)R.T;
    const S mach[3] = {(S)R.mach[0], (S)R.mach[1], (S)R.mach[2]};
    const S rho_in = p_in / (S(1) * T_in);
    const S cin = dsqrt((S(1.4) * p_in) / rho_in);
    const S vin[3] = {mach[0] * cin, mach[1] * cin, mach[2] * cin};
    const S e_in = p_in / S(0.4) + (S(0.5) * rho_in) * ((vin[0] * vin[0] + vin[1] * vin[1]) + vin[2] * vin[2]);
    S r0, ru0[3], e0;
    if ((u[0] * n[0] + u[1] * n[1]) + u[2] * n[2] < S(0)) {
        if (mach_i > S(1)) {
            r0 = rho_in;
#pragma unroll
            for (int k = 0; k < 3; ++k) ru0[k] = (mach[k] * rho_in) * cin;
            e0 = e_in;
        } else {
            r0 = p / (S(1) * T_in);
            const S c = dsqrt((S(1.4) * p) / r0);
            const S v[3] = {mach[0] * c, mach[1] * c, mach[2] * c};
#pragma unroll
            for (int k = 0; k < 3; ++k) ru0[k] = v[k] * r0;
            e0 = p / S(0.4) + (S(0.5) * r0) * ((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        }
    } else {
        if (mach_i > S(1)) {
            r0 = r1;
#pragma unroll
            for (int k = 0; k < 3; ++k) ru0[k] = ru1[k];
            e0 = e1;
        } else {
            r0 = p_in / (S(1) * tt);
#pragma unroll
            for (int k = 0; k < 3; ++k) ru0[k] = u[k] * r0;
            e0 = p_in / S(0.4) + (S(0.5) * r0) * ((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
        }
    }
    a.u[0][c0] = r0;
    a.u[1][c0] = ru0[0];
    a.u[2][c0] = ru0[1];
    a.u[3][c0] = ru0[2];
    a.u[4][c0] = e0;
}

// Deterministic block sum of one double per thread (fixed shuffle tree, then waves in order).
__device__ inline double block_sum(double v) {
    __shared__ double wsum[kBlock / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) wsum[wave] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += wsum[w];
    return t;
}

// Reconstructed conservative state at offset d of a cell (src/models/euler.cu:250-255).
template <typename S, bool CORR>
__device__ inline void reconstruct(const S u[5], const S g[15], const S l[5], const S d[3], S &r, S ru[3], S &re) {
    r = u[0] + ((g[0] * d[0] + g[1] * d[1]) + g[2] * d[2]) * l[0];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        S gd;
        if (CORR) gd = (g[3 * (1 + k)] * d[0] + g[3 * (1 + k) + 1] * d[1]) + g[3 * (1 + k) + 2] * d[2];
        else gd = (g[3 + k] * d[0] + g[6 + k] * d[1]) + g[9 + k] * d[2];
        ru[k] = u[1 + k] + gd * l[1 + k];
    }
    re = u[4] + ((g[12] * d[0] + g[13] * d[1]) + g[14] * d[2]) * l[4];
}

// fvmEulerSteadyOp + euler_steady_step (src/models/euler.cu:236-381) for interior cell i,
// Jacobi: reads u, writes un.  Optional outputs: err (V*|R|^2, field dtype), dt, resid
// (5 doubles per cell, AoS).  Each block writes the sum of its cells' V*|R|^2 (binary64) to
// partials[blockIdx.x].  copy_ghosts also copies ghost states to un.
template <typename S, bool CORR>
__global__ __launch_bounds__(kBlock) void k_euler_flux(MeshView m, EulerArgs<S> a, double cfl, S *err, S *dt_out,
                                                       double *resid, double *partials, int copy_ghosts,
                                                       const SolverCtl *ctl) {
    if (ctl && ctl->stop) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double e_loc = 0;
    if (i < m.N) {
        const double V = m.cv[i];
        const double xi = m.cx[i], yi = m.cy[i], zi = m.cz[i];
        S ui[5], gi[15], li[5];
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            ui[c] = a.u[c][i];
            li[c] = a.l[c][i];
        }
#pragma unroll
        for (int k = 0; k < 15; ++k) gi[k] = a.g[k][i];
        S fr = 0, fu[3] = {0, 0, 0}, fe = 0, dts = 0;
        const uint32_t s = m.cfs[i], e = m.cfs[i + 1];
        for (uint32_t q = s; q < e; ++q) {
            const uint32_t f = m.cf[q], j = m.cc[q];
            const double fcx = m.fcx[f], fcy = m.fcy[f], fcz = m.fcz[f];
            const double d0x = fcx - xi, d0y = fcy - yi, d0z = fcz - zi;
            const double nfx = m.fnx[f], nfy = m.fny[f], nfz = m.fnz[f];
            const double dn = (nfx * d0x + nfy * d0y) + nfz * d0z;
            const double sg = (double)(dn > 0) - (double)(dn < 0);
            const S n[3] = {(S)(nfx * sg), (S)(nfy * sg), (S)(nfz * sg)};
            const S ds = (S)(m.fa[f] / V);
            const S d0[3] = {(S)d0x, (S)d0y, (S)d0z};
            const S d1[3] = {(S)(fcx - m.cx[j]), (S)(fcy - m.cy[j]), (S)(fcz - m.cz[j])};
            S uj[5], gj[15], lj[5];
#pragma unroll
            for (int c = 0; c < 5; ++c) uj[c] = a.u[c][j];
            if (j < m.N) {
#pragma unroll
                for (int c = 0; c < 5; ++c) lj[c] = a.l[c][j];
#pragma unroll
                for (int k = 0; k < 15; ++k) gj[k] = a.g[k][j];
            } else {   // ghost: gradient and limiter of its cell, which is i
#pragma unroll
                for (int c = 0; c < 5; ++c) lj[c] = li[c];
#pragma unroll
                for (int k = 0; k < 15; ++k) gj[k] = gi[k];
            }
            S r0, ru0[3], e0, r1, ru1[3], e1;
            reconstruct<S, CORR>(ui, gi, li, d0, r0, ru0, e0);
            reconstruct<S, CORR>(uj, gj, lj, d1, r1, ru1, e1);
            const S p0 = S(0.4) * (e0 - (S(0.5) / r0) * ((ru0[0] * ru0[0] + ru0[1] * ru0[1]) + ru0[2] * ru0[2]));
            const S p1 = S(0.4) * (e1 - (S(0.5) / r1) * ((ru1[0] * ru1[0] + ru1[1] * ru1[1]) + ru1[2] * ru1[2]));
            const S V0 = ((ru0[0] * n[0] + ru0[1] * n[1]) + ru0[2] * n[2]) / r0;
            const S V1 = ((ru1[0] * n[0] + ru1[1] * n[1]) + ru1[2] * n[2]) / r1;
            const S fr0 = V0 * r0, fr1 = V1 * r1;
            const S fe0 = V0 * (e0 + p0), fe1 = V1 * (e1 + p1);
            fr += ((fr0 + fr1) * ds) * S(0.5);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const S fu0 = ru0[k] * V0 + n[k] * p0;
                const S fu1 = ru1[k] * V1 + n[k] * p1;
                fu[k] += ((fu0 + fu1) * ds) * S(0.5);
            }
            fe += ((fe0 + fe1) * ds) * S(0.5);
            const S v0[3] = {ru0[0] / r0, ru0[1] / r0, ru0[2] / r0};
            const S v1[3] = {ru1[0] / r1, ru1[1] / r1, ru1[2] / r1};
            const S c0 = dsqrt(fmax(S(1e-12), fmin(S(1e30), (S(1.4) * p0) / r0)));
            const S c1 = dsqrt(fmax(S(1e-12), fmin(S(1e30), (S(1.4) * p1) / r1)));
            const S eig = fmax(dsqrt((v0[0] * v0[0] + v0[1] * v0[1]) + v0[2] * v0[2]) + c0,
                               dsqrt((v1[0] * v1[0] + v1[1] * v1[1]) + v1[2] * v1[2]) + c1);
            fr -= ((eig * (r1 - r0)) * S(0.5)) * ds;
#pragma unroll
            for (int k = 0; k < 3; ++k) fu[k] -= (((ru1[k] - ru0[k]) * eig) * S(0.5)) * ds;
            fe -= ((eig * (e1 - e0)) * S(0.5)) * ds;
            dts += eig * ds;
        }
        const S dt = (S)cfl / dts;
        if (a.un[0]) {   // null when only the residual is wanted
            a.un[0][i] = ui[0] + (-fr) * dt;
#pragma unroll
            for (int k = 0; k < 3; ++k) a.un[1 + k][i] = ui[1 + k] + (-fu[k]) * dt;
            a.un[4][i] = ui[4] + (-fe) * dt;
        }
        const double dfr = fr, dfe = fe, dfu[3] = {fu[0], fu[1], fu[2]};
        e_loc = ((dfr * dfr + ((dfu[0] * dfu[0] + dfu[1] * dfu[1]) + dfu[2] * dfu[2])) + dfe * dfe) * V;
        if (err) err[i] = (S)e_loc;
        if (dt_out) dt_out[i] = dt;
        if (resid) {
            resid[5 * (size_t)i + 0] = fr;
            resid[5 * (size_t)i + 1] = fu[0];
            resid[5 * (size_t)i + 2] = fu[1];
            resid[5 * (size_t)i + 3] = fu[2];
            resid[5 * (size_t)i + 4] = fe;
        }
    } else if (copy_ghosts && i < m.N + m.G) {
#pragma unroll
        for (int c = 0; c < 5; ++c) a.un[c][i] = a.u[c][i];
    }
    const double bs = block_sum(e_loc);
    if (threadIdx.x == 0 && partials) partials[blockIdx.x] = bs;
}

// Sum of n partials by one block in a fixed order (valid in thread 0): eight independent loads
// in flight per thread (a single running sum waits on each load in turn), the eight sums then
// combined in a fixed order, then the block sum.
__device__ inline double block_sum_partials(const double *__restrict__ partials, uint32_t n) {
    double a8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const uint32_t st = blockDim.x;
    uint32_t k = threadIdx.x;
    for (; k + 7 * st < n; k += 8 * st)
#pragma unroll
        for (int u = 0; u < 8; ++u) a8[u] += partials[k + u * st];
    for (; k < n; k += st) a8[0] += partials[k];
    const double acc = ((a8[0] + a8[1]) + (a8[2] + a8[3])) + ((a8[4] + a8[5]) + (a8[6] + a8[7]));
    return block_sum(acc);
}

// Records an iteration's residual sum tot in ctl (one thread): history, normalisation (the
// first residual unless given), convergence and divergence, the iteration limit; clears pending.
__device__ inline void record_iteration(SolverCtl *ctl, double *hist, double tot) {
    const int it = ctl->iter;
    if (it < ctl->hist_cap) hist[it] = tot;
    if (it == 0 && ctl->ref <= 0) ctl->ref = tot;
    ctl->last = tot;
    ctl->ratio = tot / ctl->ref;
    ctl->iter = it + 1;
    ctl->pending = 0;
    if (!isfinite(tot)) {
        ctl->diverged = 1;
        ctl->stop = 1;
    } else if (ctl->tol > 0 && ctl->ratio < ctl->tol) {
        ctl->converged = 1;
        ctl->stop = 1;
    }
    if (ctl->iter >= ctl->max_iter) ctl->stop = 1;
}

// Sum of n partials by one block in a fixed order; result to out[0].  With ctl: records the
// iteration (history, normalisation, convergence, stop flags) instead of writing out, and only
// if a flux launch left partials pending.
__global__ __launch_bounds__(kBlock) void k_finalize(const double *partials, uint32_t n, double *out, SolverCtl *ctl,
                                                     double *hist);

// Sum over interior cells of a field component (FvmFieldReduceSum) into per-block partials.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_field_partial_sum(const S *v, uint32_t N, double *partials) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double x = i < N ? (double)v[i] : 0.0;
    const double bs = block_sum(x);
    if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

}  // namespace gr

// Incompressible pressure-correction solver (no arithmetic in the reference: src/models/nsi.cu is
// a stub that does not compile).  Collocated, cell-centred, constant density 1; per step:
//   1. ghosts (no-slip walls with the region's velocity, zero normal pressure gradient) and the
//      least-squares gradients of u and p (the Euler path's weighted LSQ, k_nsi_grad);
//   2. predictor u* = u + dt/V (-sum F_out u_up + nu sum a (u_j - u_i)) - dt grad p, with u_up
//      the second-order upwind value;
//   3. face fluxes F* = A (avg(u*).n + dt (avg(grad p).n - (p1 - p0)/D)) from the lower- to the
//      higher-numbered cell (momentum interpolation against pressure checkerboarding);
//   4. Poisson sum_f a (p'_i - p'_j) = -(1/dt) sum F*_out by Jacobi-preconditioned conjugate
//      gradients (binary64 dot products, fixed-order block reductions: deterministic), four
//      kernels per iteration (spmv with the direction update, alpha, vector update, beta),
//      replayed from a HIP graph of 16 iterations until the device reports convergence;
//   5. corrections F = F* - dt a (p'_1 - p'_0), u = u* - dt grad p', p += p'.
// Two precisions: all binary64 (S = T = double), or mixed (S = float, T = double): velocity,
// pressure and face fluxes (the carriers, T) stay binary64 and the predictor stores the increment
// du* = u* - u instead of u*, while gradients, increments and CG vectors (the bulk, S) are
// binary32.  A binary32 velocity alone holds the continuity residual of the 64^3 cavity at about
// 2e-4 of its start (the rounding of u, about 6e-8 |u|, enters every face flux); the increment
// form carries only the rounding of small increments.  Geometry, dot products and the continuity
// residual are binary64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "fvm_kernels.hpp"

namespace gr {

struct NsiWalls {
    uint32_t n;                        // regions
    uint32_t start[kMaxRegions + 1];   // ghost ranges (ghost index g, not cell index)
    double u[kMaxRegions][3];          // wall velocity per region (unset regions: at rest)
};

__device__ inline const double *nsi_wall(const NsiWalls &w, uint32_t g) {
    uint32_t r = 0;
    while (r + 1 < w.n && g >= w.start[r + 1]) ++r;
    return w.u[r];
}

struct NsiFaces {
    const uint32_t *c0, *c1;       // lower / higher cell of each face
    const double *nx, *ny, *nz;    // normal oriented from c0 to c1
    const double *D, *a;           // (x1 - x0).n > 0 and A / D
};

// K blocks of cell values passed by value (no pointer table in memory).
template <typename V, int K>
struct NsiVals {
    V *p[K];
};

// Ghost values: u_g = 2 u_wall - u_i, p_g = p_i (p may be null).
template <typename S>
__global__ void k_nsi_bc(MeshView m, S *ux, S *uy, S *uz, S *p, NsiWalls w) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= m.G) return;
    const uint32_t c = m.N + g, i = m.cc[m.cfs[c]];
    const double *uw = nsi_wall(w, g);
    if (ux) {
        ux[c] = (S)(2.0 * uw[0]) - ux[i];
        uy[c] = (S)(2.0 * uw[1]) - uy[i];
        uz[c] = (S)(2.0 * uw[2]) - uz[i];
    }
    if (p) p[c] = p[i];
}

// Weighted least-squares gradients of NC components of values T into S (the gradient part of the
// Euler path's k_grad_lim, src/fvmops.cu:51-111): differences u_j - u_i are formed in T before
// the cast, so binary64 carriers keep their accuracy in binary32 gradients.
template <typename S, typename T, int NC>
__global__ __launch_bounds__(kBlock) void k_nsi_grad(MeshView m, NsiVals<T, NC> u, NsiVals<S, 3 * NC> g) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N) return;
    const double xi = m.cx[i], yi = m.cy[i], zi = m.cz[i];
    T ui[NC];
    S acc[NC][3];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        ui[c] = u.p[c][i];
        acc[c][0] = acc[c][1] = acc[c][2] = S(0);
    }
    for (uint32_t q = m.cfs[i]; q < m.cfs[i + 1]; ++q) {
        const uint32_t j = m.cc[q];
        const S d[3] = {(S)(m.cx[j] - xi), (S)(m.cy[j] - yi), (S)(m.cz[j] - zi)};
        const S w2 = S(1) / ((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const S du = (S)(u.p[c][j] - ui[c]);
#pragma unroll
            for (int k = 0; k < 3; ++k) acc[c][k] += (d[k] * du) * w2;
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        S out[3];
        lsq_apply<S>(m, i, acc[c], out);
#pragma unroll
        for (int k = 0; k < 3; ++k) g.p[3 * c + k][i] = out[k];
    }
}

// Predictor for interior cell i: u* (all binary64) or the increment du* = u* - u (mixed).
template <typename S, typename T>
__global__ __launch_bounds__(kBlock) void k_nsi_predict(MeshView m, NsiVals<T, 3> u, const S *const *gu,
                                                        const S *const *gp, const T *__restrict__ F, NsiWalls w,
                                                        double nu, double dt, NsiFaces fc, NsiVals<S, 3> o) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N) return;
    S conv[3] = {S(0), S(0), S(0)}, diff[3] = {S(0), S(0), S(0)};
    const T ui[3] = {u.p[0][i], u.p[1][i], u.p[2][i]};
    for (uint32_t q = m.cfs[i]; q < m.cfs[i + 1]; ++q) {
        const uint32_t f = m.cf[q], j = m.cc[q];
        const S out = (S)(i < j ? F[f] : -F[f]);
        S uf[3];
        if (j >= m.N) {
            const double *uw = nsi_wall(w, j - m.N);
#pragma unroll
            for (int k = 0; k < 3; ++k) uf[k] = (S)uw[k];
        } else {
            const uint32_t up = out >= S(0) ? i : j;
            const S dd[3] = {(S)(m.fcx[f] - m.cx[up]), (S)(m.fcy[f] - m.cy[up]), (S)(m.fcz[f] - m.cz[up])};
#pragma unroll
            for (int k = 0; k < 3; ++k)
                uf[k] = (S)u.p[k][up] + ((gu[3 * k][up] * dd[0] + gu[3 * k + 1][up] * dd[1]) + gu[3 * k + 2][up] * dd[2]);
        }
        const S a = (S)fc.a[f];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            conv[k] += out * uf[k];
            diff[k] += a * (S)(u.p[k][j] - ui[k]);
        }
    }
    const S s = (S)(dt / m.cv[i]), sdt = (S)dt, snu = (S)nu;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if constexpr (std::is_same<S, T>::value) o.p[k][i] = ui[k] + s * (snu * diff[k] - conv[k]) - sdt * gp[k][i];
        else o.p[k][i] = s * (snu * diff[k] - conv[k]) - sdt * gp[k][i];
    }
}

// Component k of u* at interior cell c: stored (all binary64) or u + du* (mixed), in T.
template <typename S, typename T>
__device__ inline T nsi_ustar(const NsiVals<T, 3> &u, const NsiVals<S, 3> &us, int k, uint32_t c) {
    if constexpr (std::is_same<S, T>::value) return us.p[k][c];
    else return u.p[k][c] + (T)us.p[k][c];
}

// Predicted face fluxes, face-parallel, in T.
template <typename S, typename T>
__global__ void k_nsi_flux(MeshView m, NsiFaces fc, NsiVals<T, 3> u, NsiVals<S, 3> us, const S *const *gp,
                           const T *__restrict__ p, NsiWalls w, double dt, T *__restrict__ Fs) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const uint32_t c0 = fc.c0[f], c1 = fc.c1[f];
    const T n[3] = {(T)fc.nx[f], (T)fc.ny[f], (T)fc.nz[f]};
    const T A = (T)m.fa[f];
    if (c1 >= m.N) {
        const double *uw = nsi_wall(w, c1 - m.N);
        Fs[f] = A * (((T)uw[0] * n[0] + (T)uw[1] * n[1]) + (T)uw[2] * n[2]);
        return;
    }
    T un = T(0), gn = T(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        un += T(0.5) * (nsi_ustar<S, T>(u, us, k, c0) + nsi_ustar<S, T>(u, us, k, c1)) * n[k];
        gn += T(0.5) * ((T)gp[k][c0] + (T)gp[k][c1]) * n[k];
    }
    Fs[f] = A * (un + (T)dt * (gn - (p[c1] - p[c0]) / (T)fc.D[f]));
}

// Net outflow of the predicted fluxes: b_i = -(1/dt) sum F*_out (in T, stored as S); block
// partials of |sum| (the continuity residual) and of b.
template <typename S, typename T>
__global__ __launch_bounds__(kBlock) void k_nsi_div(MeshView m, const T *__restrict__ Fs, double dt, S *__restrict__ b,
                                                    double *__restrict__ part_abs, double *__restrict__ part_sum) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double ab = 0, sm = 0;
    if (i < m.N) {
        T s = T(0);
        for (uint32_t q = m.cfs[i]; q < m.cfs[i + 1]; ++q) {
            const uint32_t f = m.cf[q];
            s += i < m.cc[q] ? Fs[f] : -Fs[f];
        }
        const T bi = -s / (T)dt;
        b[i] = (S)bi;
        ab = fabs((double)s);
        sm = (double)(S)bi;
    }
    const double t1 = block_sum(ab);
    __syncthreads();
    const double t2 = block_sum(sm);
    if (threadIdx.x == 0) {
        part_abs[blockIdx.x] = t1;
        part_sum[blockIdx.x] = t2;
    }
}

// Scalars of the conjugate-gradient loop (device resident).
struct CgScalars {
    double bsum, bnorm, gam, del, rr, alpha, beta, xsum;   // the latest values (for the host)
    int done, iters;
    double gam_k[2], alpha_k[2];   // gamma and alpha of iteration k at [k & 1]
    int it_k[2];                   // k at [k & 1]: iterations done before iteration k
};

// Fixed-order sum of nb block partials (every thread of the block calls it; the result is valid
// in thread 0).  Blocks that all call it get the same bits.
__device__ inline double sum_parts(const double *__restrict__ p, int nb) {
    double s = 0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) s += p[k];
    return block_sum(s);
}

// The CG kernels run blocks of kCgBlock threads: a quarter of the partials a 256-thread grid
// leaves, which every block of the next CG launch reads again.
constexpr int kCgBlock = 1024;
constexpr int kCgRow = 6;   // operator row entries held in registers by the CG iteration

// Sums of three values over a block of kCgBlock threads (valid in thread 0): wave sums by a
// shuffle tree, then the wave sums in wave order (fixed order: equal inputs give equal bits).
__device__ inline void cg_block_sum3(double &a, double &b, double &c) {
    constexpr int NW = kCgBlock / 64;
    __shared__ double ws[3][NW];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        ws[0][w] = a;
        ws[1][w] = b;
        ws[2][w] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = b = c = 0;
        for (int k = 0; k < NW; ++k) {
            a += ws[0][k];
            b += ws[1][k];
            c += ws[2][k];
        }
    }
}

// Fixed-order sums of the three partial rows q[0..nb), q[nb..2nb), q[2nb..3nb) (valid in thread 0).
__device__ inline void cg_sum_parts3(const double *__restrict__ q, int nb, double &a, double &b, double &c) {
    a = b = c = 0;
    for (int k = threadIdx.x; k < nb; k += kCgBlock) {
        a += q[k];
        b += q[nb + k];
        c += q[2 * nb + k];
    }
    cg_block_sum3(a, b, c);
}

// Sum of partials in a fixed order into *out (one block).
__global__ __launch_bounds__(kBlock) void k_nsi_sum(const double *__restrict__ part, int n, double *out, int done_gate,
                                                    const CgScalars *cg) {
    if (done_gate && cg->done) return;
    double s = 0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) s += part[k];
    const double t = block_sum(s);
    if (threadIdx.x == 0) *out = t;
}

// The preconditioned residual of a cell: r / diag (0 where the cell has no interior neighbour).
template <typename S>
__device__ inline S cg_precond(S r, double diag) {
    return diag > 0 ? (S)((double)r / diag) : S(0);
}

// Jacobi-preconditioned conjugate gradients in the Chronopoulos-Gear form, whose two inner
// products (gamma = r.u, delta = w.u, with u = M^-1 r and w = A u) are formed together, so an
// iteration needs one grid-wide synchronisation instead of two:
//   beta_k = gamma_k / gamma_{k-1},  alpha_k = gamma_k / (delta_k - beta_k gamma_k / alpha_{k-1})
//   (k = 0: beta 0, alpha = gamma / delta),
//   p = u + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s,  u = M^-1 r,  w = A u.
// In exact arithmetic its iterates are those of the textbook algorithm.  The residual, w and s
// are double-buffered (slot k & 1 read, the other written): a cell's w = A u needs u of its
// neighbours after the update, which each block forms on the fly from their r, w and s of the
// previous iteration (cg_next_u), so no block reads a vector another block has already
// overwritten.

// u after this iteration's update, from the cell's r, w and s before it.
template <typename S>
__device__ inline S cg_next_u(S r, S w, S s_old, double diag, S alpha, S beta, bool first) {
    const S s = first ? w : w + beta * s_old;
    return cg_precond<S>(r - alpha * s, diag);
}

// Set-up: r = b - mean(b), x = 0, u = M^-1 r, w = A u (neighbours' u from b on the fly); block
// partials of gamma, delta and r.r into q[0 .. 3 nb).
template <typename S>
__global__ __launch_bounds__(kCgBlock) void k_cg_init(uint32_t N, const uint32_t *__restrict__ start,
                                                    const uint32_t *__restrict__ col, const S *__restrict__ coef,
                                                    const S *__restrict__ b, const double *__restrict__ diag,
                                                    const CgScalars *cg, S *__restrict__ x, S *__restrict__ r,
                                                    S *__restrict__ w, double *__restrict__ q, int nb) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const S mean = (S)(cg->bsum / (double)N);
    double g = 0, d = 0, rr = 0;
    if (i < N) {
        const S ri = b[i] - mean;
        const S ui = cg_precond<S>(ri, diag[i]);
        S wi = S(0);
        for (uint32_t k = start[i]; k < start[i + 1]; ++k) {
            const uint32_t j = col[k];
            wi += coef[k] * (ui - cg_precond<S>(b[j] - mean, diag[j]));
        }
        x[i] = S(0);
        r[i] = ri;
        w[i] = wi;
        g = (double)ri * (double)ui;
        d = (double)wi * (double)ui;
        rr = (double)ri * (double)ri;
    }
    cg_block_sum3(g, d, rr);
    if (threadIdx.x == 0) {
        q[blockIdx.x] = g;
        q[nb + blockIdx.x] = d;
        q[2 * nb + blockIdx.x] = rr;
    }
}

// After the set-up: |b| (= |r| of the set-up) and the start flags (one block).
__global__ __launch_bounds__(kBlock) void k_cg_start(const double *__restrict__ q, int nb, CgScalars *cg, int cg_max) {
    const double rr = sum_parts(q + 2 * nb, nb);
    if (threadIdx.x == 0) {
        cg->rr = rr;
        cg->bnorm = sqrt(rr);
        cg->iters = 0;
        cg->beta = 0.0;
        cg->gam_k[0] = cg->gam_k[1] = 0.0;
        cg->alpha_k[0] = cg->alpha_k[1] = 0.0;
        cg->it_k[0] = cg->it_k[1] = 0;
        cg->done = (cg->bnorm == 0.0 || cg_max == 0) ? 1 : 0;
    }
}

// Iteration k (slot = k & 1: static in a replayed graph of an even number of iterations).  Every
// block sums the previous launch's partials itself, in the same fixed order as every other
// block, so all blocks take the same decisions with no one-block reduction kernel and no
// device-scope fence (block 0 records the scalars): stop when |r| < tol |b| or k >= cg_max
// (k >= 1), or when the alpha denominator is 0.  Then per cell: p, s, x, r, u and w = A u (the
// neighbours' u by cg_next_u), and the partials of the next gamma, delta and r.r.  The cell's
// and its neighbours' values do not depend on this iteration's alpha and beta (each cell writes
// only its own p and x; r, w and s are read from the other buffer of their pair), so they are
// loaded before the partial sums, whose latency they then share: the first kCgRow entries of the
// row into registers (a hexahedron's 6 faces, the longest row of the operator), any further
// ones afterwards, in the same order.
template <typename S>
__global__ __launch_bounds__(kCgBlock) void k_cg_iter(uint32_t N, const uint32_t *__restrict__ start,
                                                    const uint32_t *__restrict__ col, const S *__restrict__ coef,
                                                    const double *__restrict__ diag, S *__restrict__ x,
                                                    S *__restrict__ p, const S *__restrict__ r_in,
                                                    const S *__restrict__ w_in, const S *__restrict__ s_in,
                                                    S *__restrict__ r_out, S *__restrict__ w_out,
                                                    S *__restrict__ s_out, const double *__restrict__ q_in,
                                                    double *__restrict__ q_out, int nb, CgScalars *cg, int slot,
                                                    double tol, int cg_max) {
    if (cg->done) return;
    __shared__ double s_alpha, s_beta;
    __shared__ int s_stop, s_first;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double di = 0, nd[kCgRow];
    S ri = S(0), wi = S(0), p_old = S(0), s_old = S(0), xi = S(0), nr[kCgRow], nw[kCgRow], ns[kCgRow], nc[kCgRow];
    uint32_t k0 = 0, k1 = 0;
    if (i < N) {
        di = diag[i];
        ri = r_in[i];
        wi = w_in[i];
        p_old = p[i];
        s_old = s_in[i];
        xi = x[i];
        k0 = start[i];
        k1 = start[i + 1];
        // the row's indices first, then the values at them: loads complete in order, so a wait
        // for one neighbour's index inside a per-neighbour loop would also wait for the values
        // of the neighbours before it.  Entries past the row's end repeat its last (never used).
        if (k1 > k0) {
            uint32_t jn[kCgRow];
#pragma unroll
            for (int e = 0; e < kCgRow; ++e) {
                const uint32_t ke = min(k0 + (uint32_t)e, k1 - 1);
                jn[e] = col[ke];
                nc[e] = coef[ke];
            }
#pragma unroll
            for (int e = 0; e < kCgRow; ++e) {
                const uint32_t j = jn[e];
                nr[e] = r_in[j];
                nw[e] = w_in[j];
                ns[e] = s_in[j];
                nd[e] = diag[j];
            }
        }
    }
    double gam, del, rr;
    cg_sum_parts3(q_in, nb, gam, del, rr);
    if (threadIdx.x == 0) {
        const int k = cg->it_k[slot];
        bool stop = k > 0 && ((tol > 0 && sqrt(rr) < tol * cg->bnorm) || k >= cg_max);
        double alpha = 0.0, beta = 0.0;
        if (!stop) {
            double den = del;
            if (k > 0) {
                const double gprev = cg->gam_k[slot ^ 1];
                beta = gprev != 0.0 ? gam / gprev : 0.0;
                den = del - beta * gam / cg->alpha_k[slot ^ 1];
            }
            if (den == 0.0) stop = true;
            else alpha = gam / den;
        }
        s_alpha = alpha;
        s_beta = beta;
        s_stop = stop ? 1 : 0;
        s_first = k == 0 ? 1 : 0;
        if (blockIdx.x == 0) {
            cg->gam_k[slot] = gam;
            cg->alpha_k[slot] = alpha;
            cg->it_k[slot ^ 1] = k + 1;
            cg->iters = k;
            cg->gam = gam;
            cg->del = del;
            cg->rr = rr;
            cg->alpha = alpha;
            cg->beta = beta;
            if (stop) cg->done = 1;
        }
    }
    __syncthreads();
    if (s_stop) return;
    const S al = (S)s_alpha, be = (S)s_beta;
    const bool first = s_first != 0;
    double g = 0, d = 0, rr1 = 0;
    if (i < N) {
        const S ui = cg_precond<S>(ri, di);
        const S pi = first ? ui : ui + be * p_old;
        const S si = first ? wi : wi + be * s_old;
        x[i] = xi + al * pi;
        const S r1 = ri - al * si;
        const S u1 = cg_precond<S>(r1, di);
        S w1 = S(0);
#pragma unroll
        for (int e = 0; e < kCgRow; ++e)
            if (k0 + e < k1) w1 += nc[e] * (u1 - cg_next_u<S>(nr[e], nw[e], first ? S(0) : ns[e], nd[e], al, be, first));
        for (uint32_t k = k0 + kCgRow; k < k1; ++k) {
            const uint32_t j = col[k];
            w1 += coef[k] * (u1 - cg_next_u<S>(r_in[j], w_in[j], first ? S(0) : s_in[j], diag[j], al, be, first));
        }
        p[i] = pi;
        s_out[i] = si;
        r_out[i] = r1;
        w_out[i] = w1;
        g = (double)r1 * (double)u1;
        d = (double)w1 * (double)u1;
        rr1 = (double)r1 * (double)r1;
    }
    cg_block_sum3(g, d, rr1);
    if (threadIdx.x == 0) {
        q_out[blockIdx.x] = g;
        q_out[nb + blockIdx.x] = d;
        q_out[2 * nb + blockIdx.x] = rr1;
    }
}

// x -= mean(x) and the ghost copy; p' ghosts = interior neighbour.
template <typename S>
__global__ void k_nsi_shift(uint32_t N, const CgScalars *cg, S *__restrict__ x) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) x[i] -= (S)(cg->xsum / (double)N);
}

template <typename S>
__global__ __launch_bounds__(kBlock) void k_nsi_partial_sum(uint32_t N, const S *__restrict__ x,
                                                            double *__restrict__ part) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double v = i < N ? (double)x[i] : 0.0;
    const double t = block_sum(v);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <typename S, typename T>
__global__ void k_nsi_correct_faces(MeshView m, NsiFaces fc, const T *__restrict__ Fs, const S *__restrict__ pp,
                                    double dt, T *__restrict__ F) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const uint32_t c0 = fc.c0[f], c1 = fc.c1[f];
    F[f] = c1 >= m.N ? Fs[f] : Fs[f] - (T)dt * (T)fc.a[f] * (T)(pp[c1] - pp[c0]);
}

template <typename S, typename T>
__global__ void k_nsi_correct_cells(uint32_t N, NsiVals<T, 3> u, NsiVals<S, 3> us, const S *const *gpp,
                                    const S *__restrict__ pp, double dt, T *__restrict__ p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const T sdt = (T)dt;
    T un[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) un[k] = nsi_ustar<S, T>(u, us, k, i) - sdt * (T)gpp[k][i];
#pragma unroll
    for (int k = 0; k < 3; ++k) u.p[k][i] = un[k];
    p[i] += (T)pp[i];
}

// Face tables: lower/higher cell from the CSR (the lower cell lists the face first).
__global__ void k_nsi_faces(MeshView m, const uint32_t *__restrict__ c0, const uint32_t *__restrict__ c1,
                            double *nx, double *ny, double *nz, double *D, double *a) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const uint32_t i = c0[f], j = c1[f];
    const double dx = m.cx[j] - m.cx[i], dy = m.cy[j] - m.cy[i], dz = m.cz[j] - m.cz[i];
    double d = (dx * m.fnx[f] + dy * m.fny[f]) + dz * m.fnz[f];
    const double sg = d < 0 ? -1.0 : 1.0;
    nx[f] = m.fnx[f] * sg;
    ny[f] = m.fny[f] * sg;
    nz[f] = m.fnz[f] * sg;
    d = d * sg;
    D[f] = d;
    a[f] = m.fa[f] / d;
}

// The CG operator's rows: for interior cell i, its interior neighbours in slot order with the
// face coefficient a_f in the bulk precision (the spmv then reads one index and one coefficient
// per entry instead of a face index, a neighbour index and a binary64 coefficient gathered by
// face; same entries, same order, same values).
template <typename S>
__global__ void k_nsi_lap_rows(MeshView m, const double *__restrict__ a, const uint32_t *__restrict__ start,
                               uint32_t *__restrict__ col, S *__restrict__ coef) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N) return;
    uint32_t o = start[i];
    for (uint32_t k = m.cfs[i]; k < m.cfs[i + 1]; ++k) {
        const uint32_t j = m.cc[k];
        if (j < m.N) {
            col[o] = j;
            coef[o] = (S)a[m.cf[k]];
            ++o;
        }
    }
}

__global__ void k_nsi_diag(MeshView m, const double *__restrict__ a, double *__restrict__ diag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N) return;
    double s = 0;
    for (uint32_t k = m.cfs[i]; k < m.cfs[i + 1]; ++k)
        if (m.cc[k] < m.N) s += a[m.cf[k]];
    diag[i] = s;
}

}  // namespace gr

using gr::guarded;
using gr::require;

struct gr_nsi {
    static constexpr int kCgChunk = 16;   // CG iterations per graph replay (even: d buffers return)
    gr_mesh *mesh = nullptr;
    gr_nsi_config cfg{};
    size_t cs = 0;                        // padded cell stride
    gr::NsiWalls walls{};
    gr::DevArray<char> st;                // bulk S: u* or du* (3 blocks), p', grad u (9), grad p (3), grad p' (3)
    gr::DevArray<double> carrier;         // T = binary64: u (3 blocks), p
    gr::DevArray<double> flux;            // T: F, F*
    gr::DevArray<char> cgv;               // b, r[2], w[2], s[2], p; x = p' (state block 3)
    gr::DevArray<uint32_t> c0, c1;
    gr::DevArray<double> fgeo;            // nx, ny, nz, D, a (face blocks)
    gr::DevArray<double> diag, part, part2, cont;   // CG partials of slots 0 and 1 (3 per block each)
    gr::DevArray<uint32_t> lap_start, lap_col;             // CG operator rows: interior neighbours
    gr::DevArray<char> lap_coef;                           // their face coefficients, bulk precision
    gr::DevArray<gr::CgScalars> cg;
    gr::CgScalars *cg_host = nullptr;     // pinned: [0], [1] CG polls, [2] end of step
    double *cont_host = nullptr;          // pinned
    gr::DevArray<void *> ptrs;            // gradient pointer tables
    hipGraphExec_t g_pre = nullptr, g_cg = nullptr, g_post = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<double> history;          // continuity residual per step
    std::vector<int> cg_iters;
    size_t history_start = 0;             // first step since creation or the last set_state

    ~gr_nsi() {
        drop_graphs();
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (cg_host) (void)hipHostFree(cg_host);
        if (cont_host) (void)hipHostFree(cont_host);
    }
    void drop_graphs() {
        for (hipGraphExec_t *g : {&g_pre, &g_cg, &g_post})
            if (*g) {
                (void)hipGraphExecDestroy(*g);
                *g = nullptr;
            }
    }
    size_t es() const { return (size_t)cfg.dtype; }
    static constexpr int kBulkBlocks = 19;
    template <typename S> S *blk(int b) const { return (S *)st.p + (size_t)b * cs; }
    // bulk blocks: 0-2 u* (binary64) or du* = u* - u (mixed), 3 p', 4-12 grad u, 13-15 grad p,
    // 16-18 grad p'
    double *cu(int k) const { return carrier.p + (size_t)k * cs; }   // carriers: 0-2 u, 3 p
    double *F() const { return flux.p; }
    double *Fs() const { return flux.p + gr::pad64(mesh->h.n_faces); }
    template <typename S> S *cgb(int k) const { return (S *)cgv.p + (size_t)k * cs; }
    gr::NsiFaces faces() const {
        const size_t fs = gr::pad64(mesh->h.n_faces);
        return gr::NsiFaces{c0.p, c1.p, fgeo.p, fgeo.p + fs, fgeo.p + 2 * fs, fgeo.p + 3 * fs, fgeo.p + 4 * fs};
    }
    unsigned nb() const { return gr::blocks_for(mesh->h.n_cells); }
    unsigned cg_nb() const { return (mesh->h.n_cells + gr::kCgBlock - 1) / gr::kCgBlock; }

    template <typename S, typename T, int NC> void gradient(T *const *vals, S *const *grads) {
        gr::NsiVals<T, NC> u{};
        gr::NsiVals<S, 3 * NC> g{};
        for (int c = 0; c < NC; ++c) u.p[c] = vals[c];
        for (int k = 0; k < 3 * NC; ++k) g.p[k] = grads[k];
        hipLaunchKernelGGL((gr::k_nsi_grad<S, T, NC>), dim3(nb()), dim3(gr::kBlock), 0, mesh->stream, mesh->view(), u, g);
    }
    gr::NsiVals<double, 3> uvals() const { return gr::NsiVals<double, 3>{{cu(0), cu(1), cu(2)}}; }
    template <typename S> gr::NsiVals<S, 3> usvals() const { return gr::NsiVals<S, 3>{{blk<S>(0), blk<S>(1), blk<S>(2)}}; }

    // Steps 1-3 and the CG set-up: ghosts, gradients, predictor, predicted fluxes, their
    // divergence (b and the continuity residual), r = b - mean(b), z, d, r.z, |b|.
    template <typename S> void enqueue_pre() {
        using T = double;
        gr_mesh *m = mesh;
        const gr::MeshView v = m->view();
        const uint32_t N = m->h.n_cells, G = m->h.n_ghosts, F = m->h.n_faces;
        const unsigned B = nb(), BG = gr::blocks_for(std::max<uint32_t>(G, 1)), BF = gr::blocks_for(F);
        T *p = cu(3);
        S **tab = (S **)ptrs.p;   // [0..8] grad u, [9..11] grad p, [12..14] grad p'
        S *gu[9], *gp[3];
        for (int k = 0; k < 9; ++k) gu[k] = blk<S>(4 + k);
        for (int k = 0; k < 3; ++k) gp[k] = blk<S>(13 + k);
        if (G) hipLaunchKernelGGL(gr::k_nsi_bc<T>, dim3(BG), dim3(gr::kBlock), 0, m->stream, v, cu(0), cu(1), cu(2), p, walls);
        T *uv[3] = {cu(0), cu(1), cu(2)};
        gradient<S, T, 3>(uv, gu);
        gradient<S, T, 1>(&p, gp);
        hipLaunchKernelGGL((gr::k_nsi_predict<S, T>), dim3(B), dim3(gr::kBlock), 0, m->stream, v, uvals(),
                           (const S *const *)tab, (const S *const *)(tab + 9), (const T *)this->F(), walls, cfg.nu,
                           cfg.dt, faces(), usvals<S>());
        hipLaunchKernelGGL((gr::k_nsi_flux<S, T>), dim3(BF), dim3(gr::kBlock), 0, m->stream, v, faces(), uvals(),
                           usvals<S>(), (const S *const *)(tab + 9), (const T *)p, walls, cfg.dt, Fs());
        hipLaunchKernelGGL((gr::k_nsi_div<S, T>), dim3(B), dim3(gr::kBlock), 0, m->stream, v, (const T *)Fs(), cfg.dt,
                           cgb<S>(0), part.p, part2.p);
        hipLaunchKernelGGL(gr::k_nsi_sum, dim3(1), dim3(gr::kBlock), 0, m->stream, (const double *)part.p, (int)B,
                           cont.p, 0, (const gr::CgScalars *)cg.p);
        // CG set-up: mean of b, r, x = 0, w = A M^-1 r, the first gamma, delta, r.r; |b|
        S *b = cgb<S>(0);
        gr::CgScalars *c = cg.p;
        double *pa = part.p;
        hipLaunchKernelGGL(gr::k_nsi_partial_sum<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, N, (const S *)b, pa);
        hipLaunchKernelGGL(gr::k_nsi_sum, dim3(1), dim3(gr::kBlock), 0, m->stream, (const double *)pa, (int)B,
                           &c->bsum, 0, (const gr::CgScalars *)c);
        const unsigned BC = cg_nb();
        hipLaunchKernelGGL(gr::k_cg_init<S>, dim3(BC), dim3(gr::kCgBlock), 0, m->stream, N,
                           (const uint32_t *)lap_start.p, (const uint32_t *)lap_col.p, (const S *)lap_coef.p,
                           (const S *)b, (const double *)diag.p, (const gr::CgScalars *)c, blk<S>(3), cg_r<S>(0),
                           cg_w<S>(0), part.p, (int)BC);
        hipLaunchKernelGGL(gr::k_cg_start, dim3(1), dim3(gr::kBlock), 0, m->stream, (const double *)part.p, (int)BC, c,
                           cfg.cg_max);
    }

    // CG vectors: r, w, s of slot k (iteration k reads slot k & 1 and writes the other), p
    template <typename S> S *cg_r(int k) const { return cgb<S>(1 + k); }
    template <typename S> S *cg_w(int k) const { return cgb<S>(3 + k); }
    template <typename S> S *cg_s(int k) const { return cgb<S>(5 + k); }
    template <typename S> S *cg_p() const { return cgb<S>(7); }
    double *cg_q(int k) const { return k ? part2.p : part.p; }

    // CG iteration of slot k (k = 0, 1)
    template <typename S> void launch_cg_iter(int k, double tol, int cg_max) {
        gr_mesh *m = mesh;
        const unsigned B = cg_nb();
        hipLaunchKernelGGL(gr::k_cg_iter<S>, dim3(B), dim3(gr::kCgBlock), 0, m->stream, m->h.n_cells,
                           (const uint32_t *)lap_start.p, (const uint32_t *)lap_col.p, (const S *)lap_coef.p,
                           (const double *)diag.p, blk<S>(3), cg_p<S>(), (const S *)cg_r<S>(k), (const S *)cg_w<S>(k),
                           (const S *)cg_s<S>(k), cg_r<S>(k ^ 1), cg_w<S>(k ^ 1), cg_s<S>(k ^ 1),
                           (const double *)cg_q(k), cg_q(k ^ 1), (int)B, cg.p, k, tol, cg_max);
    }

    // kCgChunk CG iterations (one kernel each); every kernel returns at once after convergence.
    template <typename S> void enqueue_cg() {
        for (int it = 0; it < kCgChunk; ++it) launch_cg_iter<S>(it & 1, cfg.cg_tol, cfg.cg_max);
    }

    // Zero mean of p', its ghosts and gradient, and step 5 (corrections of F, u, p; ghosts).
    template <typename S> void enqueue_post() {
        using T = double;
        gr_mesh *m = mesh;
        const gr::MeshView v = m->view();
        const uint32_t N = m->h.n_cells, G = m->h.n_ghosts, F = m->h.n_faces;
        const unsigned B = nb(), BG = gr::blocks_for(std::max<uint32_t>(G, 1)), BF = gr::blocks_for(F);
        S *pp = blk<S>(3);
        S **tab = (S **)ptrs.p;
        S *gpp[3];
        for (int k = 0; k < 3; ++k) gpp[k] = blk<S>(16 + k);
        gr::CgScalars *c = cg.p;
        double *pa = part.p;
        hipLaunchKernelGGL(gr::k_nsi_partial_sum<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, N, (const S *)pp, pa);
        hipLaunchKernelGGL(gr::k_nsi_sum, dim3(1), dim3(gr::kBlock), 0, m->stream, pa, (int)B, &c->xsum, 0,
                           (const gr::CgScalars *)c);
        hipLaunchKernelGGL(gr::k_nsi_shift<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, N, (const gr::CgScalars *)c, pp);
        if (G)
            hipLaunchKernelGGL(gr::k_nsi_bc<S>, dim3(BG), dim3(gr::kBlock), 0, m->stream, v, (S *)nullptr,
                               (S *)nullptr, (S *)nullptr, pp, walls);
        gradient<S, S, 1>(&pp, gpp);
        hipLaunchKernelGGL((gr::k_nsi_correct_faces<S, T>), dim3(BF), dim3(gr::kBlock), 0, m->stream, v, faces(),
                           (const T *)Fs(), (const S *)pp, cfg.dt, this->F());
        hipLaunchKernelGGL((gr::k_nsi_correct_cells<S, T>), dim3(B), dim3(gr::kBlock), 0, m->stream, N, uvals(),
                           usvals<S>(), (const S *const *)(tab + 12), (const S *)pp, cfg.dt, cu(3));
        if (G)
            hipLaunchKernelGGL(gr::k_nsi_bc<T>, dim3(BG), dim3(gr::kBlock), 0, m->stream, v, cu(0), cu(1), cu(2), cu(3),
                               walls);
    }

    template <typename Fn> hipGraphExec_t capture(Fn &&fn) {
        hipGraph_t g = nullptr;
        GR_HIP(hipStreamBeginCapture(mesh->stream, hipStreamCaptureModeRelaxed));
        try {
            fn();
        } catch (...) {
            (void)hipStreamEndCapture(mesh->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        GR_HIP(hipStreamEndCapture(mesh->stream, &g));
        hipGraphExec_t e = nullptr;
        const hipError_t r = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        GR_HIP(r);
        return e;
    }
    template <typename S> void ensure_graphs() {
        if (!g_pre) g_pre = capture([&] { enqueue_pre<S>(); });
        if (!g_cg) g_cg = capture([&] { enqueue_cg<S>(); });
        if (!g_post) g_post = capture([&] { enqueue_post<S>(); });
    }

    // One time step: the pre graph, CG graphs until the device reports convergence (two in
    // flight: the host waits for the previous replay's copy of the CG scalars while the next one
    // runs; a replay after convergence does no work), the post graph.
    template <typename S> void step(double *cont_out, int *its_out) {
        gr_mesh *m = mesh;
        ensure_graphs<S>();
        GR_HIP(hipGraphLaunch(g_pre, m->stream));
        // the spmv of iteration cg_max stops the loop, so one replay more than cg_max / 16
        const int chunks = cfg.cg_max / kCgChunk + 1;
        for (int L = 0; L < chunks; ++L) {
            GR_HIP(hipGraphLaunch(g_cg, m->stream));
            GR_HIP(hipMemcpyAsync(&cg_host[L & 1], cg.p, sizeof *cg_host, hipMemcpyDeviceToHost, m->stream));
            GR_HIP(hipEventRecord(ev[L & 1], m->stream));
            if (L >= 1) {
                GR_HIP(hipEventSynchronize(ev[(L - 1) & 1]));
                if (cg_host[(L - 1) & 1].done) break;
            }
        }
        GR_HIP(hipGraphLaunch(g_post, m->stream));
        GR_HIP(hipMemcpyAsync(cont_host, cont.p, sizeof *cont_host, hipMemcpyDeviceToHost, m->stream));
        GR_HIP(hipMemcpyAsync(&cg_host[2], cg.p, sizeof *cg_host, hipMemcpyDeviceToHost, m->stream));
        GR_HIP(hipStreamSynchronize(m->stream));
        if (cont_out) *cont_out = *cont_host;
        if (its_out) *its_out = cg_host[2].iters;
    }
};


namespace {

void upload_state(gr_nsi *s, const double *u, const double *p, const double *F) {
    gr_mesh *m = s->mesh;
    const size_t NG = (size_t)m->h.n_cells + m->h.n_ghosts, NF = m->h.n_faces;
    std::vector<double> blk(4 * s->cs, 0.0), fl(gr::pad64(NF), 0.0);
    for (size_t i = 0; i < NG; ++i) {
        for (int k = 0; k < 3; ++k) blk[k * s->cs + i] = u ? u[3 * i + k] : 0.0;
        blk[3 * s->cs + i] = p ? p[i] : 0.0;
    }
    for (size_t f = 0; f < NF; ++f) fl[f] = F ? F[f] : 0.0;
    GR_HIP(hipMemcpyAsync(s->carrier.p, blk.data(), blk.size() * sizeof(double), hipMemcpyHostToDevice, m->stream));
    GR_HIP(hipMemcpyAsync(s->flux.p, fl.data(), fl.size() * sizeof(double), hipMemcpyHostToDevice, m->stream));
    GR_HIP(hipStreamSynchronize(m->stream));
}

void download_state(gr_nsi *s, double *u, double *p, double *F) {
    gr_mesh *m = s->mesh;
    const size_t NG = (size_t)m->h.n_cells + m->h.n_ghosts, NF = m->h.n_faces;
    std::vector<double> blk(4 * s->cs), fl(NF ? NF : 1);
    GR_HIP(hipMemcpyAsync(blk.data(), s->carrier.p, blk.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    if (NF) GR_HIP(hipMemcpyAsync(fl.data(), s->flux.p, NF * sizeof(double), hipMemcpyDeviceToHost, m->stream));
    GR_HIP(hipStreamSynchronize(m->stream));
    for (size_t i = 0; i < NG; ++i) {
        if (u)
            for (int k = 0; k < 3; ++k) u[3 * i + k] = blk[k * s->cs + i];
        if (p) p[i] = blk[3 * s->cs + i];
    }
    if (F)
        for (size_t f = 0; f < NF; ++f) F[f] = fl[f];
}

template <typename S>
void fill_pointer_table(gr_nsi *s) {
    std::vector<void *> t(15);
    for (int k = 0; k < 9; ++k) t[k] = s->blk<S>(4 + k);
    for (int k = 0; k < 3; ++k) {
        t[9 + k] = s->blk<S>(13 + k);
        t[12 + k] = s->blk<S>(16 + k);
    }
    s->ptrs.upload(t.data(), t.size(), s->mesh->stream);
}

}  // namespace

extern "C" {

int gr_nsi_create(gr_mesh *m, const gr_nsi_config *cfg, gr_nsi **out) {
    return guarded([&] {
        require(m && cfg && out, "null argument");
        require(m->metrics_ready, "mesh metrics not computed (gr_mesh_compute_metrics)");
        require(cfg->dtype == GR_F64 || cfg->dtype == GR_F32, "dtype must be GR_F64 or GR_F32");
        require(cfg->nu >= 0 && cfg->dt > 0 && cfg->cg_max >= 0 && cfg->cg_tol >= 0, "bad nu, dt or CG settings");
        require(m->h.n_regions <= (uint32_t)gr::kMaxRegions, "too many boundary regions");
        require(!m->h.one_dimensional, "the incompressible solver needs a 3-D mesh");
        auto *s = new gr_nsi();
        try {
            s->mesh = m;
            s->cfg = *cfg;
            const gr::HostMesh &h = m->h;
            const uint32_t N = h.n_cells, G = h.n_ghosts, F = h.n_faces;
            s->cs = gr::pad64((size_t)N + G);
            s->st.alloc(gr_nsi::kBulkBlocks * s->cs * s->es());
            GR_HIP(hipMemsetAsync(s->st.p, 0, s->st.n, m->stream));
            s->carrier.alloc(4 * s->cs);
            GR_HIP(hipMemsetAsync(s->carrier.p, 0, s->carrier.n * sizeof(double), m->stream));
            s->flux.alloc(2 * gr::pad64(F));
            GR_HIP(hipMemsetAsync(s->flux.p, 0, s->flux.n * sizeof(double), m->stream));
            s->cgv.alloc(8 * s->cs * s->es());
            GR_HIP(hipMemsetAsync(s->cgv.p, 0, s->cgv.n, m->stream));
            // face tables: the lower-numbered cell lists the face first
            std::vector<uint32_t> a0(F, 0xFFFFFFFFu), a1(F, 0xFFFFFFFFu);
            for (uint32_t i = 0; i < N + G; ++i)
                for (uint32_t q = h.cell_faces_starts[i]; q < h.cell_faces_starts[i + 1]; ++q) {
                    const uint32_t f = h.cell_faces[q], j = h.cell_connects[q];
                    if (i < j) {
                        a0[f] = i;
                        a1[f] = j;
                    }
                }
            for (uint32_t f = 0; f < F; ++f) require(a0[f] != 0xFFFFFFFFu, "face without two cells");
            s->c0.alloc(F);
            s->c1.alloc(F);
            s->c0.upload(a0.data(), F, m->stream);
            s->c1.upload(a1.data(), F, m->stream);
            s->fgeo.alloc(5 * gr::pad64(F));
            const size_t fs = gr::pad64(F);
            hipLaunchKernelGGL(gr::k_nsi_faces, dim3(gr::blocks_for(F)), dim3(gr::kBlock), 0, m->stream, m->view(),
                               (const uint32_t *)s->c0.p, (const uint32_t *)s->c1.p, s->fgeo.p, s->fgeo.p + fs,
                               s->fgeo.p + 2 * fs, s->fgeo.p + 3 * fs, s->fgeo.p + 4 * fs);
            s->diag.alloc(std::max<uint32_t>(N, 1));
            hipLaunchKernelGGL(gr::k_nsi_diag, dim3(gr::blocks_for(N)), dim3(gr::kBlock), 0, m->stream, m->view(),
                               (const double *)(s->fgeo.p + 4 * fs), s->diag.p);
            {
                std::vector<uint32_t> start(N + 1, 0);
                for (uint32_t i = 0; i < N; ++i) {
                    uint32_t n_int = 0;
                    for (uint32_t q = h.cell_faces_starts[i]; q < h.cell_faces_starts[i + 1]; ++q)
                        if (h.cell_connects[q] < N) ++n_int;
                    start[i + 1] = start[i] + n_int;
                }
                s->lap_start.alloc(N + 1);
                s->lap_start.upload(start.data(), N + 1, m->stream);
                s->lap_col.alloc(std::max<uint32_t>(start[N], 1));
                s->lap_coef.alloc((size_t)std::max<uint32_t>(start[N], 1) * s->es());
                if (N) {
                    if (cfg->dtype == GR_F64)
                        hipLaunchKernelGGL(gr::k_nsi_lap_rows<double>, dim3(gr::blocks_for(N)), dim3(gr::kBlock), 0,
                                           m->stream, m->view(), (const double *)(s->fgeo.p + 4 * fs),
                                           (const uint32_t *)s->lap_start.p, s->lap_col.p, (double *)s->lap_coef.p);
                    else
                        hipLaunchKernelGGL(gr::k_nsi_lap_rows<float>, dim3(gr::blocks_for(N)), dim3(gr::kBlock), 0,
                                           m->stream, m->view(), (const double *)(s->fgeo.p + 4 * fs),
                                           (const uint32_t *)s->lap_start.p, s->lap_col.p, (float *)s->lap_coef.p);
                }
                GR_HIP(hipStreamSynchronize(m->stream));   // the host start array is freed at the end of this scope
            }
            s->part.alloc(3 * (size_t)gr::blocks_for(N) + 1);
            s->part2.alloc(3 * (size_t)gr::blocks_for(N) + 1);
            s->cont.alloc(1);
            s->cg.alloc(1);
            GR_HIP(hipMemsetAsync(s->cg.p, 0, sizeof(gr::CgScalars), m->stream));
            GR_HIP(hipHostMalloc((void **)&s->cg_host, 3 * sizeof(gr::CgScalars)));
            GR_HIP(hipHostMalloc((void **)&s->cont_host, sizeof(double)));
            for (auto &e : s->ev) GR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            s->ptrs.alloc(15);
            if (cfg->dtype == GR_F64) fill_pointer_table<double>(s);
            else fill_pointer_table<float>(s);
            s->walls.n = h.n_regions;
            for (uint32_t r = 0; r <= h.n_regions; ++r) s->walls.start[r] = h.ghost_cell_starts[r];
            GR_HIP(hipGetLastError());
            GR_HIP(hipStreamSynchronize(m->stream));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

void gr_nsi_free(gr_nsi *s) {
    if (s && s->mesh && s->mesh->stream) (void)hipStreamSynchronize(s->mesh->stream);
    delete s;
}

int gr_nsi_set_wall(gr_nsi *s, const char *region, const double u[3]) {
    return guarded([&] {
        require(s && region && u, "null argument");
        const gr::HostMesh &h = s->mesh->h;
        uint32_t r = 0;
        for (; r < h.n_regions; ++r)
            if (h.region_names[r] == region) break;
        if (r == h.n_regions) throw gr::Error(GR_ERR_NOT_FOUND, std::string("region '") + region + "' not found in mesh");
        for (int k = 0; k < 3; ++k) s->walls.u[r][k] = u[k];
        s->drop_graphs();   // the wall velocities are kernel arguments of the captured graphs
    });
}

int gr_nsi_set_state(gr_nsi *s, const double *u, const double *p, const double *F) {
    return guarded([&] {
        require(s != nullptr, "null argument");
        upload_state(s, u, p, F);
        s->history_start = s->history.size();
    });
}

int gr_nsi_get_state(gr_nsi *s, double *u, double *p, double *F) {
    return guarded([&] {
        require(s != nullptr, "null argument");
        download_state(s, u, p, F);
    });
}

int gr_nsi_step(gr_nsi *s, int n_steps, double *cont, int *cg_iters) {
    return guarded([&] {
        require(s != nullptr, "null argument");
        require(n_steps >= 1, "n_steps must be positive");
        for (int k = 0; k < n_steps; ++k) {
            double c = 0;
            int it = 0;
            if (s->cfg.dtype == GR_F64) s->step<double>(&c, &it);
            else s->step<float>(&c, &it);
            if (!std::isfinite(c)) throw gr::Error(GR_ERR_DIVERGED, "continuity residual is not finite (diverged)");
            s->history.push_back(c);
            s->cg_iters.push_back(it);
            if (cont) cont[k] = c;
            if (cg_iters) cg_iters[k] = it;
        }
    });
}

int gr_nsi_iterate(gr_nsi *s, int max_steps, double tol, int *steps, double *ratio) {
    return guarded([&] {
        require(s != nullptr, "null argument");
        require(max_steps >= 1 && tol >= 0, "bad max_steps or tol");
        double r = 1;
        int k = 0;
        while (k < max_steps) {
            double c = 0;
            int it = 0;
            if (s->cfg.dtype == GR_F64) s->step<double>(&c, &it);
            else s->step<float>(&c, &it);
            if (!std::isfinite(c)) throw gr::Error(GR_ERR_DIVERGED, "continuity residual is not finite (diverged)");
            s->history.push_back(c);
            s->cg_iters.push_back(it);
            ++k;
            const size_t done = s->history.size() - s->history_start;
            if (done < 5) continue;
            double ref = 0;
            for (size_t j = 0; j < 5; ++j) ref = std::max(ref, s->history[s->history_start + j]);
            r = ref > 0 ? c / ref : 0.0;
            if (tol > 0 && r < tol) break;
        }
        if (steps) *steps = k;
        if (ratio) *ratio = r;
    });
}

int gr_nsi_profile(gr_nsi *s, int n_iter, double ms[3]) {
    return guarded([&] {
        require(s && ms, "null argument");
        require(n_iter > 0 && n_iter <= 10000, "n_iter must be in 1..10000");
        gr_mesh *m = s->mesh;
        // Mean durations from HIP events around n_iter back-to-back launches on the library's
        // stream: [0] the predictor graph (ghosts, gradients, predictor, fluxes, divergence, CG
        // set-up: it writes scratch blocks only), [1] k_cg_iter (on the set-up's vectors, slot
        // 0 only, with the stopping test disabled: every launch does its work), [2] the
        // correction graph (velocity, pressure and face fluxes are saved before and restored
        // after).  The next step recomputes every scratch block, so the solve is unchanged.
        const bool f64 = s->cfg.dtype == GR_F64;
        if (f64) s->ensure_graphs<double>();
        else s->ensure_graphs<float>();
        hipEvent_t ev[4] = {};
        struct Guard {
            hipEvent_t *e;
            ~Guard() {
                for (int k = 0; k < 4; ++k)
                    if (e[k]) (void)hipEventDestroy(e[k]);
            }
        } guard{ev};
        for (auto &e : ev) GR_HIP(hipEventCreate(&e));
        gr::DevArray<double> saved_u, saved_f;
        saved_u.alloc(s->carrier.n);
        saved_f.alloc(s->flux.n);
        GR_HIP(hipMemcpyAsync(saved_u.p, s->carrier.p, s->carrier.n * sizeof(double), hipMemcpyDeviceToDevice, m->stream));
        GR_HIP(hipMemcpyAsync(saved_f.p, s->flux.p, s->flux.n * sizeof(double), hipMemcpyDeviceToDevice, m->stream));
        GR_HIP(hipEventRecord(ev[0], m->stream));
        for (int k = 0; k < n_iter; ++k) GR_HIP(hipGraphLaunch(s->g_pre, m->stream));
        GR_HIP(hipEventRecord(ev[1], m->stream));
        // tol 0 and an iteration limit no run reaches; slot 0 only keeps the counter at 0
        for (int k = 0; k < n_iter; ++k) {
            if (f64) s->launch_cg_iter<double>(0, 0.0, 1 << 30);
            else s->launch_cg_iter<float>(0, 0.0, 1 << 30);
        }
        GR_HIP(hipEventRecord(ev[2], m->stream));
        for (int k = 0; k < n_iter; ++k) GR_HIP(hipGraphLaunch(s->g_post, m->stream));
        GR_HIP(hipEventRecord(ev[3], m->stream));
        GR_HIP(hipMemcpyAsync(s->carrier.p, saved_u.p, s->carrier.n * sizeof(double), hipMemcpyDeviceToDevice, m->stream));
        GR_HIP(hipMemcpyAsync(s->flux.p, saved_f.p, s->flux.n * sizeof(double), hipMemcpyDeviceToDevice, m->stream));
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(m->stream));
        for (int k = 0; k < 3; ++k) {
            float t = 0;
            GR_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
            ms[k] = t / n_iter;
        }
    });
}

int gr_nsi_history(const gr_nsi *s, double *cont, int *cg_iters, int n) {
    return guarded([&] {
        require(s != nullptr, "null argument");
        require(n >= 0 && (size_t)n <= s->history.size(), "history request longer than the run");
        const size_t off = s->history.size() - n;
        for (int k = 0; k < n; ++k) {
            if (cont) cont[k] = s->history[off + k];
            if (cg_iters) cg_iters[k] = s->cg_iters[off + k];
        }
    });
}

}  // extern "C"

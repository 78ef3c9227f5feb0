// Incompressible pressure-correction solver (no arithmetic in the reference: src/models/nsi.cu is
// a stub that does not compile).  Collocated, cell-centred, constant density 1; per step:
//   1. ghosts (no-slip walls with the region's velocity, zero normal pressure gradient) and the
//      least-squares gradients of u and p (the Euler path's k_grad_lim);
//   2. predictor u* = u + dt/V (-sum F_out u_up + nu sum a (u_j - u_i)) - dt grad p, with u_up
//      the second-order upwind value;
//   3. face fluxes F* = A (avg(u*).n + dt (avg(grad p).n - (p1 - p0)/D)) from the lower- to the
//      higher-numbered cell (momentum interpolation against pressure checkerboarding);
//   4. Poisson sum_f a (p'_i - p'_j) = -(1/dt) sum F*_out by Jacobi-preconditioned conjugate
//      gradients (binary64 dot products, fixed-order block reductions: deterministic);
//   5. corrections F = F* - dt a (p'_1 - p'_0), u = u* - dt grad p', p += p'.
// State, fluxes and CG vectors are S (binary64 or binary32); geometry, dot products and the
// continuity residual are binary64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
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

// Predictor for interior cell i.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_nsi_predict(MeshView m, const S *__restrict__ ux, const S *__restrict__ uy,
                                                        const S *__restrict__ uz, const S *const *gu, const S *const *gp,
                                                        const S *__restrict__ F, NsiWalls w, double nu, double dt,
                                                        NsiFaces fc, S *__restrict__ sx, S *__restrict__ sy,
                                                        S *__restrict__ sz) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N) return;
    const S *u[3] = {ux, uy, uz};
    S conv[3] = {S(0), S(0), S(0)}, diff[3] = {S(0), S(0), S(0)};
    const S ui[3] = {ux[i], uy[i], uz[i]};
    for (uint32_t q = m.cfs[i]; q < m.cfs[i + 1]; ++q) {
        const uint32_t f = m.cf[q], j = m.cc[q];
        const S out = i < j ? F[f] : -F[f];
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
                uf[k] = u[k][up] + ((gu[3 * k][up] * dd[0] + gu[3 * k + 1][up] * dd[1]) + gu[3 * k + 2][up] * dd[2]);
        }
        const S a = (S)fc.a[f];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            conv[k] += out * uf[k];
            diff[k] += a * (u[k][j] - ui[k]);
        }
    }
    const S s = (S)(dt / m.cv[i]), sdt = (S)dt, snu = (S)nu;
    S *o[3] = {sx, sy, sz};
#pragma unroll
    for (int k = 0; k < 3; ++k) o[k][i] = ui[k] + s * (snu * diff[k] - conv[k]) - sdt * gp[k][i];
}

// Predicted face fluxes, face-parallel.
template <typename S>
__global__ void k_nsi_flux(MeshView m, NsiFaces fc, const S *__restrict__ sx, const S *__restrict__ sy,
                           const S *__restrict__ sz, const S *const *gp, const S *__restrict__ p, NsiWalls w,
                           double dt, S *__restrict__ Fs) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const uint32_t c0 = fc.c0[f], c1 = fc.c1[f];
    const S n[3] = {(S)fc.nx[f], (S)fc.ny[f], (S)fc.nz[f]};
    const S A = (S)m.fa[f];
    if (c1 >= m.N) {
        const double *uw = nsi_wall(w, c1 - m.N);
        Fs[f] = A * (((S)uw[0] * n[0] + (S)uw[1] * n[1]) + (S)uw[2] * n[2]);
        return;
    }
    const S *s[3] = {sx, sy, sz};
    S un = S(0), gn = S(0);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        un += S(0.5) * (s[k][c0] + s[k][c1]) * n[k];
        gn += S(0.5) * (gp[k][c0] + gp[k][c1]) * n[k];
    }
    Fs[f] = A * (un + (S)dt * (gn - (p[c1] - p[c0]) / (S)fc.D[f]));
}

// Net outflow of the predicted fluxes: b_i = -(1/dt) sum F*_out; block partials of |sum| and of b.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_nsi_div(MeshView m, const S *__restrict__ Fs, double dt, S *__restrict__ b,
                                                    double *__restrict__ part_abs, double *__restrict__ part_sum) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double ab = 0, sm = 0;
    if (i < m.N) {
        S s = S(0);
        for (uint32_t q = m.cfs[i]; q < m.cfs[i + 1]; ++q) {
            const uint32_t f = m.cf[q];
            s += i < m.cc[q] ? Fs[f] : -Fs[f];
        }
        const S bi = -s / (S)dt;
        b[i] = bi;
        ab = fabs((double)s);
        sm = (double)bi;
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
    double bsum, bnorm, rz, dq, rr, alpha, beta, xsum;
    int done, iters;
};

// Sum of partials in a fixed order into *out (one block).
__global__ __launch_bounds__(kBlock) void k_nsi_sum(const double *__restrict__ part, int n, double *out, int done_gate,
                                                    const CgScalars *cg) {
    if (done_gate && cg->done) return;
    double s = 0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) s += part[k];
    const double t = block_sum(s);
    if (threadIdx.x == 0) *out = t;
}

// r = b - mean(b), x = 0, z = r / diag, d = z; partials of r.z and r.r.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_cg_init(uint32_t N, const S *__restrict__ b, const double *__restrict__ diag,
                                                    CgScalars *cg, S *__restrict__ x, S *__restrict__ r,
                                                    S *__restrict__ d, double *__restrict__ p_rz,
                                                    double *__restrict__ p_rr) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0, rr = 0;
    if (i < N) {
        const S ri = b[i] - (S)(cg->bsum / (double)N);
        const S zi = diag[i] > 0 ? (S)((double)ri / diag[i]) : S(0);
        x[i] = S(0);
        r[i] = ri;
        d[i] = zi;
        rz = (double)ri * (double)zi;
        rr = (double)ri * (double)ri;
    }
    const double t1 = block_sum(rz);
    __syncthreads();
    const double t2 = block_sum(rr);
    if (threadIdx.x == 0) {
        p_rz[blockIdx.x] = t1;
        p_rr[blockIdx.x] = t2;
    }
}

// q = A d over interior faces; partials of d.q.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(MeshView m, const double *__restrict__ a, const S *__restrict__ d,
                                                    S *__restrict__ q, double *__restrict__ p_dq, const CgScalars *cg) {
    if (cg->done) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double dq = 0;
    if (i < m.N) {
        const S di = d[i];
        S s = S(0);
        for (uint32_t k = m.cfs[i]; k < m.cfs[i + 1]; ++k) {
            const uint32_t j = m.cc[k];
            if (j < m.N) s += (S)a[m.cf[k]] * (di - d[j]);
        }
        q[i] = s;
        dq = (double)di * (double)s;
    }
    const double t = block_sum(dq);
    if (threadIdx.x == 0) p_dq[blockIdx.x] = t;
}

// x += alpha d, r -= alpha q, z = r / diag; partials of r.z and r.r.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_cg_update(uint32_t N, const double *__restrict__ diag, CgScalars *cg,
                                                      S *__restrict__ x, S *__restrict__ r, S *__restrict__ z,
                                                      const S *__restrict__ d, const S *__restrict__ q,
                                                      double *__restrict__ p_rz, double *__restrict__ p_rr) {
    if (cg->done) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    double rz = 0, rr = 0;
    if (i < N) {
        const S al = (S)cg->alpha;
        x[i] += al * d[i];
        const S ri = r[i] - al * q[i];
        r[i] = ri;
        const S zi = diag[i] > 0 ? (S)((double)ri / diag[i]) : S(0);
        z[i] = zi;
        rz = (double)ri * (double)zi;
        rr = (double)ri * (double)ri;
    }
    const double t1 = block_sum(rz);
    __syncthreads();
    const double t2 = block_sum(rr);
    if (threadIdx.x == 0) {
        p_rz[blockIdx.x] = t1;
        p_rr[blockIdx.x] = t2;
    }
}

// d = z + beta d
template <typename S>
__global__ void k_cg_dir(uint32_t N, const CgScalars *cg, const S *__restrict__ z, S *__restrict__ d) {
    if (cg->done) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) d[i] = z[i] + (S)cg->beta * d[i];
}

// alpha = r.z / d.q from the spmv's block partials (one block; fixed-order sum).
__global__ __launch_bounds__(kBlock) void k_cg_alpha(const double *__restrict__ p_dq, int nb, CgScalars *cg) {
    if (cg->done) return;
    double s = 0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) s += p_dq[k];
    const double dq = block_sum(s);
    if (threadIdx.x == 0) {
        cg->dq = dq;
        if (dq == 0.0) cg->done = 1;
        else cg->alpha = cg->rz / dq;
    }
}

// New r.z and r.r from the update's partials; iteration count, convergence test and beta.
__global__ __launch_bounds__(kBlock) void k_cg_beta(const double *__restrict__ p_rz, const double *__restrict__ p_rr,
                                                    int nb, CgScalars *cg, double tol, int cg_max) {
    if (cg->done) return;
    double a = 0, b = 0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        a += p_rz[k];
        b += p_rr[k];
    }
    const double rz = block_sum(a);
    __syncthreads();
    const double rr = block_sum(b);
    if (threadIdx.x == 0) {
        const double rz_old = cg->rz;
        cg->iters += 1;
        cg->rz = rz;
        cg->rr = rr;
        if ((tol > 0 && sqrt(rr) < tol * cg->bnorm) || cg->iters >= cg_max) cg->done = 1;
        cg->beta = rz_old != 0.0 ? rz / rz_old : 0.0;
    }
}

// After init: r.z, |b| and the start flags (one block).
__global__ __launch_bounds__(kBlock) void k_cg_start(const double *__restrict__ p_rz, const double *__restrict__ p_rr,
                                                     int nb, CgScalars *cg, int cg_max) {
    double a = 0, b = 0;
    for (int k = threadIdx.x; k < nb; k += blockDim.x) {
        a += p_rz[k];
        b += p_rr[k];
    }
    const double rz = block_sum(a);
    __syncthreads();
    const double rr = block_sum(b);
    if (threadIdx.x == 0) {
        cg->rz = rz;
        cg->rr = rr;
        cg->bnorm = sqrt(rr);
        cg->iters = 0;
        cg->done = (cg->bnorm == 0.0 || cg_max == 0) ? 1 : 0;
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

template <typename S>
__global__ void k_nsi_correct_faces(MeshView m, NsiFaces fc, const S *__restrict__ Fs, const S *__restrict__ pp,
                                    double dt, S *__restrict__ F) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= m.F) return;
    const uint32_t c0 = fc.c0[f], c1 = fc.c1[f];
    F[f] = c1 >= m.N ? Fs[f] : Fs[f] - (S)dt * (S)fc.a[f] * (pp[c1] - pp[c0]);
}

template <typename S>
__global__ void k_nsi_correct_cells(uint32_t N, const S *__restrict__ sx, const S *__restrict__ sy,
                                    const S *__restrict__ sz, const S *const *gpp, const S *__restrict__ pp, double dt,
                                    S *__restrict__ ux, S *__restrict__ uy, S *__restrict__ uz, S *__restrict__ p) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const S sdt = (S)dt;
    ux[i] = sx[i] - sdt * gpp[0][i];
    uy[i] = sy[i] - sdt * gpp[1][i];
    uz[i] = sz[i] - sdt * gpp[2][i];
    p[i] += pp[i];
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
    gr_mesh *mesh = nullptr;
    gr_nsi_config cfg{};
    size_t cs = 0;                        // padded cell stride
    gr::NsiWalls walls{};
    gr::DevArray<char> st;                // u (3 blocks), p, u* (3), p', and gradient blocks
    gr::DevArray<char> flux;              // F, F*
    gr::DevArray<char> cgv;               // b, x (= p'), r, z, d, q
    gr::DevArray<uint32_t> c0, c1;
    gr::DevArray<double> fgeo;            // nx, ny, nz, D, a (face blocks)
    gr::DevArray<double> diag, part, part2, cont;
    gr::DevArray<gr::CgScalars> cg;
    gr::CgScalars *cg_host = nullptr;     // pinned
    gr::DevArray<void *> ptrs;            // gradient pointer tables
    int hist_n = 0;
    std::vector<double> history;          // continuity residual per step
    std::vector<int> cg_iters;

    ~gr_nsi() {
        if (cg_host) (void)hipHostFree(cg_host);
    }
    size_t es() const { return (size_t)cfg.dtype; }
    template <typename S> S *blk(int b) const { return (S *)st.p + (size_t)b * cs; }
    // blocks: 0-2 u, 3 p, 4-6 u*, 7 p', 8-16 grad u, 17-19 grad p, 20-22 grad p', 23-25 unused limiter
    template <typename S> S *F() const { return (S *)flux.p; }
    template <typename S> S *Fs() const { return (S *)flux.p + gr::pad64(mesh->h.n_faces); }
    template <typename S> S *cgb(int k) const { return (S *)cgv.p + (size_t)k * cs; }
    gr::NsiFaces faces() const {
        const size_t fs = gr::pad64(mesh->h.n_faces);
        return gr::NsiFaces{c0.p, c1.p, fgeo.p, fgeo.p + fs, fgeo.p + 2 * fs, fgeo.p + 3 * fs, fgeo.p + 4 * fs};
    }
    unsigned nb() const { return gr::blocks_for(mesh->h.n_cells); }

    template <typename S> void gradients(S *const *vals, int nc, S *const *grads, S *lim_scratch) {
        gr_mesh *m = mesh;
        if (nc == 3) {
            gr::GradLimArgs<S, 3> a{};
            for (int c = 0; c < 3; ++c) {
                a.u[c] = vals[c];
                a.l[c] = lim_scratch;
                for (int k = 0; k < 3; ++k) a.g[3 * c + k] = grads[3 * c + k];
            }
            hipLaunchKernelGGL((gr::k_grad_lim<S, 3, true, false>), dim3(nb()), dim3(gr::kBlock), 0, m->stream,
                               m->view(), a, 0, (const gr::SolverCtl *)nullptr);
        } else {
            gr::GradLimArgs<S, 1> a{};
            a.u[0] = vals[0];
            a.l[0] = lim_scratch;
            for (int k = 0; k < 3; ++k) a.g[k] = grads[k];
            hipLaunchKernelGGL((gr::k_grad_lim<S, 1, true, false>), dim3(nb()), dim3(gr::kBlock), 0, m->stream,
                               m->view(), a, 0, (const gr::SolverCtl *)nullptr);
        }
    }

    // CG on A x = b (x = block 7 of the state, p'); returns nothing, iterations in cg->iters.
    template <typename S> void solve_poisson() {
        gr_mesh *m = mesh;
        const uint32_t N = m->h.n_cells;
        const unsigned B = nb();
        S *b = cgb<S>(0), *r = cgb<S>(2), *z = cgb<S>(3), *d = cgb<S>(4), *q = cgb<S>(5);
        S *x = blk<S>(7);
        gr::CgScalars *c = cg.p;
        double *pa = part.p, *pb = part2.p;
        // mean of b, then init
        hipLaunchKernelGGL(gr::k_nsi_partial_sum<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, N, (const S *)b, pa);
        hipLaunchKernelGGL(gr::k_nsi_sum, dim3(1), dim3(gr::kBlock), 0, m->stream, (const double *)pa, (int)B,
                           &c->bsum, 0, (const gr::CgScalars *)c);
        hipLaunchKernelGGL(gr::k_cg_init<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, N, (const S *)b,
                           (const double *)diag.p, c, x, r, d, pa, pb);
        hipLaunchKernelGGL(gr::k_cg_start, dim3(1), dim3(gr::kBlock), 0, m->stream, (const double *)pa,
                           (const double *)pb, (int)B, c, cfg.cg_max);
        const double *a = fgeo.p + 4 * gr::pad64(m->h.n_faces);
        for (int it = 0; it < cfg.cg_max; ++it) {
            hipLaunchKernelGGL(gr::k_cg_spmv<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, m->view(), a, (const S *)d,
                               q, pa, (const gr::CgScalars *)c);
            hipLaunchKernelGGL(gr::k_cg_alpha, dim3(1), dim3(gr::kBlock), 0, m->stream, (const double *)pa, (int)B, c);
            hipLaunchKernelGGL(gr::k_cg_update<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, N, (const double *)diag.p,
                               c, x, r, z, (const S *)d, (const S *)q, pa, pb);
            hipLaunchKernelGGL(gr::k_cg_beta, dim3(1), dim3(gr::kBlock), 0, m->stream, (const double *)pa,
                               (const double *)pb, (int)B, c, cfg.cg_tol, cfg.cg_max);
            hipLaunchKernelGGL(gr::k_cg_dir<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, N, (const gr::CgScalars *)c,
                               (const S *)z, d);
            if ((it & 15) == 15) {
                // stop launching once the device reports convergence
                GR_HIP(hipMemcpyAsync(cg_host, c, sizeof *cg_host, hipMemcpyDeviceToHost, m->stream));
                GR_HIP(hipStreamSynchronize(m->stream));
                if (cg_host->done) break;
            }
        }
        // zero mean of x
        hipLaunchKernelGGL(gr::k_nsi_partial_sum<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, N, (const S *)x, pa);
        hipLaunchKernelGGL(gr::k_nsi_sum, dim3(1), dim3(gr::kBlock), 0, m->stream, pa, (int)B, &c->xsum, 0,
                           (const gr::CgScalars *)c);
        hipLaunchKernelGGL(gr::k_nsi_shift<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, N, (const gr::CgScalars *)c, x);
    }

    template <typename S> void step(double *cont_out, int *its_out) {
        gr_mesh *m = mesh;
        const gr::MeshView v = m->view();
        const uint32_t N = m->h.n_cells, G = m->h.n_ghosts, F = m->h.n_faces;
        const unsigned B = nb(), BG = gr::blocks_for(std::max<uint32_t>(G, 1)), BF = gr::blocks_for(F);
        S *ux = blk<S>(0), *uy = blk<S>(1), *uz = blk<S>(2), *p = blk<S>(3);
        S *sx = blk<S>(4), *sy = blk<S>(5), *sz = blk<S>(6), *pp = blk<S>(7);
        S *lim = blk<S>(23);
        S **tab = (S **)ptrs.p;   // [0..8] grad u, [9..11] grad p, [12..14] grad p'
        S *gu[9], *gp[3], *gpp[3];
        for (int k = 0; k < 9; ++k) gu[k] = blk<S>(8 + k);
        for (int k = 0; k < 3; ++k) {
            gp[k] = blk<S>(17 + k);
            gpp[k] = blk<S>(20 + k);
        }
        // 1. ghosts and gradients
        if (G) hipLaunchKernelGGL(gr::k_nsi_bc<S>, dim3(BG), dim3(gr::kBlock), 0, m->stream, v, ux, uy, uz, p, walls);
        S *uv[3] = {ux, uy, uz};
        gradients<S>(uv, 3, gu, lim);
        gradients<S>(&p, 1, gp, lim);
        // 2. predictor
        hipLaunchKernelGGL(gr::k_nsi_predict<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, v, (const S *)ux,
                           (const S *)uy, (const S *)uz, (const S *const *)tab, (const S *const *)(tab + 9),
                           (const S *)this->F<S>(), walls, cfg.nu, cfg.dt, faces(), sx, sy, sz);
        if (G)
            hipLaunchKernelGGL(gr::k_nsi_bc<S>, dim3(BG), dim3(gr::kBlock), 0, m->stream, v, sx, sy, sz, (S *)nullptr,
                               walls);
        // 3. face fluxes and divergence
        hipLaunchKernelGGL(gr::k_nsi_flux<S>, dim3(BF), dim3(gr::kBlock), 0, m->stream, v, faces(), (const S *)sx,
                           (const S *)sy, (const S *)sz, (const S *const *)(tab + 9), (const S *)p, walls, cfg.dt,
                           Fs<S>());
        hipLaunchKernelGGL(gr::k_nsi_div<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, v, (const S *)Fs<S>(), cfg.dt,
                           cgb<S>(0), part.p, part2.p);
        hipLaunchKernelGGL(gr::k_nsi_sum, dim3(1), dim3(gr::kBlock), 0, m->stream, (const double *)part.p, (int)B,
                           cont.p, 0, (const gr::CgScalars *)cg.p);
        // 4. pressure correction
        solve_poisson<S>();
        if (G)
            hipLaunchKernelGGL(gr::k_nsi_bc<S>, dim3(BG), dim3(gr::kBlock), 0, m->stream, v, (S *)nullptr,
                               (S *)nullptr, (S *)nullptr, pp, walls);
        gradients<S>(&pp, 1, gpp, lim);
        // 5. corrections
        hipLaunchKernelGGL(gr::k_nsi_correct_faces<S>, dim3(BF), dim3(gr::kBlock), 0, m->stream, v, faces(),
                           (const S *)Fs<S>(), (const S *)pp, cfg.dt, this->F<S>());
        hipLaunchKernelGGL(gr::k_nsi_correct_cells<S>, dim3(B), dim3(gr::kBlock), 0, m->stream, N, (const S *)sx,
                           (const S *)sy, (const S *)sz, (const S *const *)(tab + 12), (const S *)pp, cfg.dt, ux, uy,
                           uz, p);
        if (G) hipLaunchKernelGGL(gr::k_nsi_bc<S>, dim3(BG), dim3(gr::kBlock), 0, m->stream, v, ux, uy, uz, p, walls);
        GR_HIP(hipGetLastError());
        double c = 0;
        gr::CgScalars h;
        GR_HIP(hipMemcpyAsync(&c, cont.p, sizeof c, hipMemcpyDeviceToHost, m->stream));
        GR_HIP(hipMemcpyAsync(&h, cg.p, sizeof h, hipMemcpyDeviceToHost, m->stream));
        GR_HIP(hipStreamSynchronize(m->stream));
        if (cont_out) *cont_out = c;
        if (its_out) *its_out = h.iters;
    }
};


namespace {

template <typename S>
void upload_state(gr_nsi *s, const double *u, const double *p, const double *F) {
    gr_mesh *m = s->mesh;
    const size_t NG = (size_t)m->h.n_cells + m->h.n_ghosts, NF = m->h.n_faces;
    std::vector<S> blk(4 * s->cs, S(0)), fl(gr::pad64(NF), S(0));
    for (size_t i = 0; i < NG; ++i) {
        for (int k = 0; k < 3; ++k) blk[k * s->cs + i] = u ? (S)u[3 * i + k] : S(0);
        blk[3 * s->cs + i] = p ? (S)p[i] : S(0);
    }
    for (size_t f = 0; f < NF; ++f) fl[f] = F ? (S)F[f] : S(0);
    GR_HIP(hipMemcpyAsync(s->st.p, blk.data(), blk.size() * sizeof(S), hipMemcpyHostToDevice, m->stream));
    GR_HIP(hipMemcpyAsync(s->flux.p, fl.data(), fl.size() * sizeof(S), hipMemcpyHostToDevice, m->stream));
    GR_HIP(hipStreamSynchronize(m->stream));
}

template <typename S>
void download_state(gr_nsi *s, double *u, double *p, double *F) {
    gr_mesh *m = s->mesh;
    const size_t NG = (size_t)m->h.n_cells + m->h.n_ghosts, NF = m->h.n_faces;
    std::vector<S> blk(4 * s->cs), fl(NF ? NF : 1);
    GR_HIP(hipMemcpyAsync(blk.data(), s->st.p, blk.size() * sizeof(S), hipMemcpyDeviceToHost, m->stream));
    if (NF) GR_HIP(hipMemcpyAsync(fl.data(), s->flux.p, NF * sizeof(S), hipMemcpyDeviceToHost, m->stream));
    GR_HIP(hipStreamSynchronize(m->stream));
    for (size_t i = 0; i < NG; ++i) {
        if (u)
            for (int k = 0; k < 3; ++k) u[3 * i + k] = (double)blk[k * s->cs + i];
        if (p) p[i] = (double)blk[3 * s->cs + i];
    }
    if (F)
        for (size_t f = 0; f < NF; ++f) F[f] = (double)fl[f];
}

template <typename S>
void fill_pointer_table(gr_nsi *s) {
    std::vector<void *> t(15);
    for (int k = 0; k < 9; ++k) t[k] = s->blk<S>(8 + k);
    for (int k = 0; k < 3; ++k) {
        t[9 + k] = s->blk<S>(17 + k);
        t[12 + k] = s->blk<S>(20 + k);
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
            s->st.alloc(26 * s->cs * s->es());
            GR_HIP(hipMemsetAsync(s->st.p, 0, s->st.n, m->stream));
            s->flux.alloc(2 * gr::pad64(F) * s->es());
            GR_HIP(hipMemsetAsync(s->flux.p, 0, s->flux.n, m->stream));
            s->cgv.alloc(6 * s->cs * s->es());
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
            s->part.alloc(gr::blocks_for(N) + 1);
            s->part2.alloc(gr::blocks_for(N) + 1);
            s->cont.alloc(1);
            s->cg.alloc(1);
            GR_HIP(hipMemsetAsync(s->cg.p, 0, sizeof(gr::CgScalars), m->stream));
            GR_HIP(hipHostMalloc((void **)&s->cg_host, sizeof(gr::CgScalars)));
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
    });
}

int gr_nsi_set_state(gr_nsi *s, const double *u, const double *p, const double *F) {
    return guarded([&] {
        require(s != nullptr, "null argument");
        if (s->cfg.dtype == GR_F64) upload_state<double>(s, u, p, F);
        else upload_state<float>(s, u, p, F);
    });
}

int gr_nsi_get_state(gr_nsi *s, double *u, double *p, double *F) {
    return guarded([&] {
        require(s != nullptr, "null argument");
        if (s->cfg.dtype == GR_F64) download_state<double>(s, u, p, F);
        else download_state<float>(s, u, p, F);
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
        const size_t base = s->history.size();
        double ref = 0, r = 1;
        int k = 0;
        for (; k < max_steps; ++k) {
            double c = 0;
            int it = 0;
            if (s->cfg.dtype == GR_F64) s->step<double>(&c, &it);
            else s->step<float>(&c, &it);
            if (!std::isfinite(c)) throw gr::Error(GR_ERR_DIVERGED, "continuity residual is not finite (diverged)");
            s->history.push_back(c);
            s->cg_iters.push_back(it);
            if (k == 0) ref = c > 0 ? c : 1.0;
            r = c / ref;
            if (tol > 0 && k > 0 && r < tol) {
                ++k;
                break;
            }
        }
        (void)base;
        if (steps) *steps = k;
        if (ratio) *ratio = r;
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

/*
 * oracle_euler.c -- FV operators and the Euler residual path of the FP64 CPU oracle.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Every routine evaluates its expressions in the order the reference writes them
 * (src/fvmops.cu, src/models/euler.cu), in binary64 without contraction, so that a kernel
 * with the same evaluation order reproduces it exactly.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline double norm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double sgn(double x) { return (double)(x > 0) - (double)(x < 0); }
/* lsqrt, include/datatypes.cuh:375-377 */
static inline double lsqrt(double x) { return sqrt(fmax(1e-12, fmin(1e30, x))); }
/* tensor (columns u,v,w) times vector, include/datatypes.cuh:296-302 */
static inline void tmulv(const double *t, const double *x, double *r)
{
    for (int k = 0; k < 3; ++k) r[k] = (t[k] * x[0] + t[3 + k] * x[1]) + t[6 + k] * x[2];
}

#define FVM_TOL 1e-12

static inline uint32_t slot_begin(const orc_mesh *m, uint32_t c) { return m->cell_faces_starts[c]; }
static inline uint32_t slot_end(const orc_mesh *m, uint32_t c) { return m->cell_faces_starts[c + 1]; }

/* ---- gradients, src/fvmops.cu:51-111 -------------------------------------------------- */

void orc_scalar_gradient(const orc_mesh *m, const double *value, double *grad)
{
    const int64_t N = m->n_cells;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double g[3] = {0, 0, 0};
        const double *xi = &m->cell_centers[3 * i];
        for (uint32_t q = slot_begin(m, (uint32_t)i); q < slot_end(m, (uint32_t)i); ++q) {
            uint32_t j = m->cell_connects[q];
            const double *xj = &m->cell_centers[3 * j];
            double d[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
            double w2 = 1.0 / dot3(d, d);
            double du = value[j] - value[i];
            for (int k = 0; k < 3; ++k) g[k] += (d[k] * du) * w2;
        }
        tmulv(&m->cell_matrices[9 * i], g, &grad[3 * i]);
    }
}

void orc_vector_gradient(const orc_mesh *m, const double *value, double *grad)
{
    const int64_t N = m->n_cells;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double g[9] = {0};
        const double *xi = &m->cell_centers[3 * i];
        for (uint32_t q = slot_begin(m, (uint32_t)i); q < slot_end(m, (uint32_t)i); ++q) {
            uint32_t j = m->cell_connects[q];
            const double *xj = &m->cell_centers[3 * j];
            double d[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
            double w2 = 1.0 / dot3(d, d);
            for (int c = 0; c < 3; ++c) {
                double du = value[3 * j + c] - value[3 * i + c];
                for (int k = 0; k < 3; ++k) g[3 * c + k] += (d[k] * du) * w2;
            }
        }
        for (int c = 0; c < 3; ++c) tmulv(&m->cell_matrices[9 * i], &g[3 * c], &grad[9 * i + 3 * c]);
    }
}

/* ---- limiter, src/fvmops.cu:118-169 ----------------------------------------------------- */

static inline double limiter_func(double y, int kind)
{
    switch (kind) {
    case ORC_LIM_BARTH_JESPERSEN: return fmax(0.0, fmin(1.0, y));
    case ORC_LIM_VENKATAKRISHNAN: return (y * y + 2.0 * y) / (y * y + y + 2.0);
    default: return (y < 2.0) ? (y - (0.25 * y) * y) : 1.0;
    }
}

double orc_limiter_face(double u, const double grad[3], double umin, double umax,
                        const double d[3], int kind)
{
    double delta = dot3(grad, d);
    double d_u = umax - u;
    double d_d = umin - u;
    double dmm2 = (d_u - d_d) * (d_u - d_d);
    double dx = norm3(d);
    double k3v = 1.0 * dx;
    k3v *= (k3v * k3v);
    double sig;
    if (dmm2 <= k3v) {
        sig = 1.0;
    } else if (dmm2 <= 2 * k3v) {
        double y = dmm2 / k3v - 1.0;
        sig = ((2.0 * y) * y) * y - (3.0 * y) * y + 1.0;
    } else {
        sig = 0.0;
    }
    double lim = 1.0;
    if (delta > FVM_TOL) lim = limiter_func(d_u / delta, kind);
    else if (delta < -FVM_TOL) lim = limiter_func(d_d / delta, kind);
    lim *= 0.8;
    return sig + (1.0 - sig) * lim;
}

void orc_scalar_limiter(const orc_mesh *m, const double *value, const double *grad,
                        double *lim, int kind)
{
    const int64_t N = m->n_cells;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double vmin = value[i], vmax = value[i];
        uint32_t s = slot_begin(m, (uint32_t)i), e = slot_end(m, (uint32_t)i);
        for (uint32_t q = s; q < e; ++q) {
            double v = value[m->cell_connects[q]];
            vmin = fmin(vmin, v);
            vmax = fmax(vmax, v);
        }
        double l = 2.0;
        for (uint32_t q = s; q < e; ++q) {
            uint32_t f = m->cell_faces[q];
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = m->face_centers[3 * f + k] - m->cell_centers[3 * i + k];
            l = fmin(l, orc_limiter_face(value[i], &grad[3 * i], vmin, vmax, d, kind));
        }
        lim[i] = l;
    }
}

void orc_vector_limiter(const orc_mesh *m, const double *value, const double *grad,
                        double *lim, int kind)
{
    const int64_t N = m->n_cells;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < N; ++i) {
        double vmin[3], vmax[3];
        for (int c = 0; c < 3; ++c) vmin[c] = vmax[c] = value[3 * i + c];
        uint32_t s = slot_begin(m, (uint32_t)i), e = slot_end(m, (uint32_t)i);
        for (uint32_t q = s; q < e; ++q) {
            uint32_t j = m->cell_connects[q];
            for (int c = 0; c < 3; ++c) {
                vmin[c] = fmin(vmin[c], value[3 * j + c]);
                vmax[c] = fmax(vmax[c], value[3 * j + c]);
            }
        }
        double l[3] = {2.0, 2.0, 2.0};
        for (uint32_t q = s; q < e; ++q) {
            uint32_t f = m->cell_faces[q];
            double d[3];
            for (int k = 0; k < 3; ++k) d[k] = m->face_centers[3 * f + k] - m->cell_centers[3 * i + k];
            for (int c = 0; c < 3; ++c)
                l[c] = fmin(l[c], orc_limiter_face(value[3 * i + c], &grad[9 * i + 3 * c], vmin[c],
                                                   vmax[c], d, kind));
        }
        for (int c = 0; c < 3; ++c) lim[3 * i + c] = l[c];
    }
}

void orc_gradlim_bc(const orc_mesh *m, int ncomp, double *grad, double *lim)
{
    for (uint32_t g = 0; g < m->n_ghosts; ++g) {
        uint32_t c = m->ghost_cell_ids[g];
        uint32_t o = m->cell_connects[m->cell_faces_starts[c]];
        for (int k = 0; k < 3 * ncomp; ++k) grad[3 * ncomp * c + k] = grad[3 * ncomp * o + k];
        for (int k = 0; k < ncomp; ++k) lim[ncomp * c + k] = lim[ncomp * o + k];
    }
}

void orc_bc_fixed(const orc_mesh *m, uint32_t region, int ncomp, const double *v, double *value)
{
    for (uint32_t g = m->ghost_cell_starts[region]; g < m->ghost_cell_starts[region + 1]; ++g) {
        uint32_t c = m->ghost_cell_ids[g];
        for (int k = 0; k < ncomp; ++k) value[ncomp * c + k] = v[k];
    }
}

void orc_bc_zero_grad(const orc_mesh *m, uint32_t region, int ncomp, double *value)
{
    for (uint32_t g = m->ghost_cell_starts[region]; g < m->ghost_cell_starts[region + 1]; ++g) {
        uint32_t c = m->ghost_cell_ids[g];
        uint32_t o = m->cell_connects[m->cell_faces_starts[c]];
        for (int k = 0; k < ncomp; ++k) value[ncomp * c + k] = value[ncomp * o + k];
    }
}

double orc_reduce_sum(const double *v, uint32_t n)
{
    double s = 0.0, comp = 0.0;
    for (uint32_t i = 0; i < n; ++i) {
        double y = v[i] - comp;
        double t = s + y;
        comp = (t - s) - y;
        s = t;
    }
    return s;
}

/* ---- Euler model, src/models/euler.cu ----------------------------------------------------- */

void orc_euler_init(const orc_mesh *m, double *rho, double *rhou, double *rhoe,
                    const double mach[3], double p, double T)
{
    const uint32_t n = m->n_cells + m->n_ghosts;
    double rho_v = p / T;
    double c = sqrt((1.4 * p) / rho_v);
    double vel[3] = {mach[0] * c, mach[1] * c, mach[2] * c};
    double rhoe_v = p / 0.4 + (0.5 * rho_v) * dot3(vel, vel);
    for (uint32_t i = 0; i < n; ++i) {
        rho[i] = rho_v;
        for (int k = 0; k < 3; ++k) rhou[3 * i + k] = vel[k] * rho_v;
        rhoe[i] = rhoe_v;
    }
}

/* Outward (from the interior cell) normal of the ghost's face, src/models/euler.cu:430-434. */
static void ghost_frame(const orc_mesh *m, uint32_t g, uint32_t *c0, uint32_t *c1, double *n)
{
    *c0 = m->ghost_cell_ids[g];
    uint32_t s = m->cell_faces_starts[*c0];
    *c1 = m->cell_connects[s];
    uint32_t f = m->cell_faces[s];
    double d1[3];
    for (int k = 0; k < 3; ++k) d1[k] = m->face_centers[3 * f + k] - m->cell_centers[3 * *c1 + k];
    double sg = sgn(dot3(&m->face_normals[3 * f], d1));
    for (int k = 0; k < 3; ++k) n[k] = m->face_normals[3 * f + k] * sg;
}

void orc_euler_slip_bc(const orc_mesh *m, uint32_t region, double *rho, double *rhou,
                       double *rhoe)
{
    for (uint32_t g = m->ghost_cell_starts[region]; g < m->ghost_cell_starts[region + 1]; ++g) {
        uint32_t c0, c1;
        double n[3];
        ghost_frame(m, g, &c0, &c1, n);
        double un = dot3(n, &rhou[3 * c1]);
        for (int k = 0; k < 3; ++k) rhou[3 * c0 + k] = rhou[3 * c1 + k] - (n[k] * 2.0) * un;
        rho[c0] = rho[c1];
        rhoe[c0] = rhoe[c1];
    }
}

void orc_euler_farfield_bc(const orc_mesh *m, uint32_t region, double *rho, double *rhou,
                           double *rhoe, const double mach[3], double p_in, double T_in)
{
    for (uint32_t g = m->ghost_cell_starts[region]; g < m->ghost_cell_starts[region + 1]; ++g) {
        uint32_t c0, c1;
        double n[3];
        ghost_frame(m, g, &c0, &c1, n);
        const double r1 = rho[c1];
        const double *ru1 = &rhou[3 * c1];
        double u[3] = {ru1[0] / r1, ru1[1] / r1, ru1[2] / r1};
        double p = 0.4 * (rhoe[c1] - (0.5 / r1) * dot3(ru1, ru1));
        double t = p / (1.0 * r1);
        double mach_i = sqrt(dot3(u, u) / ((1.4 * p) / r1));
        double rho_in = p_in / (1.0 * T_in);
        double cin = sqrt((1.4 * p_in) / rho_in);
        double vin[3] = {mach[0] * cin, mach[1] * cin, mach[2] * cin};
        double e_in = p_in / 0.4 + (0.5 * rho_in) * dot3(vin, vin);
        if (dot3(u, n) < 0) {
            if (mach_i > 1.0) {
                rho[c0] = rho_in;
                for (int k = 0; k < 3; ++k) rhou[3 * c0 + k] = (mach[k] * rho_in) * cin;
                rhoe[c0] = e_in;
            } else {
                double r0 = p / (1.0 * T_in);
                rho[c0] = r0;
                double c = sqrt((1.4 * p) / r0);
                double v[3] = {mach[0] * c, mach[1] * c, mach[2] * c};
                for (int k = 0; k < 3; ++k) rhou[3 * c0 + k] = v[k] * r0;
                rhoe[c0] = p / 0.4 + (0.5 * r0) * dot3(v, v);
            }
        } else {
            if (mach_i > 1.0) {
                rho[c0] = r1;
                for (int k = 0; k < 3; ++k) rhou[3 * c0 + k] = ru1[k];
                rhoe[c0] = rhoe[c1];
            } else {
                double r0 = p_in / (1.0 * t);
                rho[c0] = r0;
                for (int k = 0; k < 3; ++k) rhou[3 * c0 + k] = u[k] * r0;
                rhoe[c0] = p_in / 0.4 + (0.5 * r0) * dot3(u, u);
            }
        }
    }
}

void orc_euler_step_prepare(const orc_mesh *m, orc_euler_state *s, int limiter_kind)
{
    orc_scalar_gradient(m, s->rho, s->grad_rho);
    orc_scalar_limiter(m, s->rho, s->grad_rho, s->lim_rho, limiter_kind);
    orc_gradlim_bc(m, 1, s->grad_rho, s->lim_rho);
    orc_vector_gradient(m, s->rhou, s->grad_rhou);
    orc_vector_limiter(m, s->rhou, s->grad_rhou, s->lim_rhou, limiter_kind);
    orc_gradlim_bc(m, 3, s->grad_rhou, s->lim_rhou);
    orc_scalar_gradient(m, s->rhoe, s->grad_rhoe);
    orc_scalar_limiter(m, s->rhoe, s->grad_rhoe, s->lim_rhoe, limiter_kind);
    orc_gradlim_bc(m, 1, s->grad_rhoe, s->lim_rhoe);
}

/* Reconstructed state of cell c at offset d (src/models/euler.cu:250-255). */
static void reconstruct(const orc_euler_state *s, uint32_t c, const double *d, int corrected,
                        double *r, double *ru, double *re)
{
    *r = s->rho[c] + dot3(&s->grad_rho[3 * c], d) * s->lim_rho[c];
    double gd[3];
    if (corrected) {
        for (int k = 0; k < 3; ++k) gd[k] = dot3(&s->grad_rhou[9 * c + 3 * k], d);
    } else {
        tmulv(&s->grad_rhou[9 * c], d, gd);
    }
    for (int k = 0; k < 3; ++k) ru[k] = s->rhou[3 * c + k] + gd[k] * s->lim_rhou[3 * c + k];
    *re = s->rhoe[c] + dot3(&s->grad_rhoe[3 * c], d) * s->lim_rhoe[c];
}

void orc_euler_step(const orc_mesh *m, const orc_euler_state *s, double cfl, int corrected,
                    double *rho_new, double *rhou_new, double *rhoe_new, double *err,
                    double *resid, double *dt_out)
{
    const int64_t N = m->n_cells;
    const uint32_t NG = m->n_cells + m->n_ghosts;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < N; ++ii) {
        const uint32_t c0 = (uint32_t)ii;
        const double v = m->cell_volumes[c0];
        double f_r = 0, f_ru[3] = {0, 0, 0}, f_re = 0, dt = 0.0;
        for (uint32_t q = slot_begin(m, c0); q < slot_end(m, c0); ++q) {
            uint32_t f = m->cell_faces[q];
            uint32_t c1 = m->cell_connects[q];
            double d0[3], d1[3], n[3];
            for (int k = 0; k < 3; ++k) {
                d0[k] = m->face_centers[3 * f + k] - m->cell_centers[3 * c0 + k];
                d1[k] = m->face_centers[3 * f + k] - m->cell_centers[3 * c1 + k];
            }
            double sg = sgn(dot3(&m->face_normals[3 * f], d0));
            for (int k = 0; k < 3; ++k) n[k] = m->face_normals[3 * f + k] * sg;
            double ds = m->face_areas[f] / v;

            double r0, ru0[3], re0, r1, ru1[3], re1;
            reconstruct(s, c0, d0, corrected, &r0, ru0, &re0);
            reconstruct(s, c1, d1, corrected, &r1, ru1, &re1);

            double p0 = 0.4 * (re0 - (0.5 / r0) * dot3(ru0, ru0));
            double p1 = 0.4 * (re1 - (0.5 / r1) * dot3(ru1, ru1));
            double V0 = dot3(ru0, n) / r0;
            double V1 = dot3(ru1, n) / r1;

            double fr0 = V0 * r0, fr1 = V1 * r1;
            double fu0[3], fu1[3];
            for (int k = 0; k < 3; ++k) {
                fu0[k] = ru0[k] * V0 + n[k] * p0;
                fu1[k] = ru1[k] * V1 + n[k] * p1;
            }
            double fe0 = V0 * (re0 + p0), fe1 = V1 * (re1 + p1);

            f_r += ((fr0 + fr1) * ds) * 0.5;
            for (int k = 0; k < 3; ++k) f_ru[k] += ((fu0[k] + fu1[k]) * ds) * 0.5;
            f_re += ((fe0 + fe1) * ds) * 0.5;

            double u0[3] = {ru0[0] / r0, ru0[1] / r0, ru0[2] / r0};
            double u1[3] = {ru1[0] / r1, ru1[1] / r1, ru1[2] / r1};
            double eig = fmax(norm3(u0) + lsqrt((1.4 * p0) / r0), norm3(u1) + lsqrt((1.4 * p1) / r1));

            f_r -= ((eig * (r1 - r0)) * 0.5) * ds;
            for (int k = 0; k < 3; ++k) f_ru[k] -= (((ru1[k] - ru0[k]) * eig) * 0.5) * ds;
            f_re -= ((eig * (re1 - re0)) * 0.5) * ds;

            dt += eig * ds;
        }
        dt = cfl / dt;
        rho_new[c0] = s->rho[c0] + (-f_r) * dt;
        for (int k = 0; k < 3; ++k) rhou_new[3 * c0 + k] = s->rhou[3 * c0 + k] + (-f_ru[k]) * dt;
        rhoe_new[c0] = s->rhoe[c0] + (-f_re) * dt;
        if (err) err[c0] = ((f_r * f_r + dot3(f_ru, f_ru)) + f_re * f_re) * v;
        if (resid) {
            resid[5 * c0 + 0] = f_r;
            for (int k = 0; k < 3; ++k) resid[5 * c0 + 1 + k] = f_ru[k];
            resid[5 * c0 + 4] = f_re;
        }
        if (dt_out) dt_out[c0] = dt;
    }
    for (uint32_t c = m->n_cells; c < NG; ++c) {
        rho_new[c] = s->rho[c];
        for (int k = 0; k < 3; ++k) rhou_new[3 * c + k] = s->rhou[3 * c + k];
        rhoe_new[c] = s->rhoe[c];
    }
}

int orc_euler_iterate(const orc_mesh *m, orc_euler_state *s, const orc_bc *bcs, int n_bcs,
                      double cfl, int limiter_kind, int corrected, int n_iter, double tol,
                      double ref_norm, double *history, int n_threads)
{
#ifdef _OPENMP
    int prev_threads = omp_get_max_threads();
    if (n_threads > 0) omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif
    const uint32_t N = m->n_cells, NG = m->n_cells + m->n_ghosts;
    double *nr = (double *)malloc(sizeof(double) * NG);
    double *nu = (double *)malloc(sizeof(double) * 3 * NG);
    double *ne = (double *)malloc(sizeof(double) * NG);
    double *err = (double *)malloc(sizeof(double) * (N ? N : 1));
    double ref = ref_norm;
    int it;
    for (it = 0; it < n_iter; ++it) {
        for (int b = 0; b < n_bcs; ++b) {
            if (bcs[b].kind == ORC_BC_SLIP)
                orc_euler_slip_bc(m, bcs[b].region, s->rho, s->rhou, s->rhoe);
            else
                orc_euler_farfield_bc(m, bcs[b].region, s->rho, s->rhou, s->rhoe, bcs[b].mach,
                                      bcs[b].p, bcs[b].T);
        }
        orc_euler_step_prepare(m, s, limiter_kind);
        orc_euler_step(m, s, cfl, corrected, nr, nu, ne, err, NULL, NULL);
        memcpy(s->rho, nr, sizeof(double) * NG);
        memcpy(s->rhou, nu, sizeof(double) * 3 * NG);
        memcpy(s->rhoe, ne, sizeof(double) * NG);
        double e = orc_reduce_sum(err, N);
        if (history) history[it] = e;
        if (it == 0 && ref <= 0) ref = e;
        if (tol > 0 && e / ref < tol) { ++it; break; }
    }
    free(err);
    free(ne);
    free(nu);
    free(nr);
#ifdef _OPENMP
    omp_set_num_threads(prev_threads);
#endif
    return it;
}

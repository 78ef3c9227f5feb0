/* FP64 CPU oracle of the incompressible pressure-correction step (test infrastructure only).
 *
 * The reference has no incompressible arithmetic: src/models/nsi.cu is a stub that does not
 * compile (it defines PhysicsEuler::step at :101 and uses undeclared FvmNSIInit, FvmNSISlipBc,
 * error and rho at :54,71,114,121,210).  This file states the scheme the device kernels
 * implement, so that they can be checked; parity with the reference is unpinned by
 * construction.  Scheme (collocated, cell-centred, constant density 1):
 *
 *   faces: c0 < c1 are the face's two cells (c1 a ghost on boundary faces); n^ = the face normal
 *   oriented from c0 to c1, D = (x_c1 - x_c0).n^ > 0, a = A/D.  F = volume flux c0 -> c1.
 *   ghosts: u_g = 2 u_wall - u_i, p_g = p_i (walls: no slip, zero normal pressure gradient).
 *   gradients: the least-squares gradient of the Euler path (orc_scalar/vector_gradient).
 *   1. predictor  u*_i = u_i + dt/V_i (-sum_f F_out u_f + nu sum_f a (u_j - u_i)) - dt grad p_i,
 *      u_f the second-order upwind value (u_up + grad u_up . (x_f - x_up); u_wall on walls).
 *   2. face flux  F*_f = A (avg(u*) . n^ + dt (avg(grad p) . n^ - (p_c1 - p_c0)/D)) (interior),
 *                 F*_f = A u_wall . n^ (boundary).
 *   3. Poisson    sum_f a (p'_i - p'_j) = -(1/dt) sum_f F*_out  (interior faces; Neumann walls),
 *      Jacobi-preconditioned conjugate gradients (Chronopoulos-Gear recurrences) from p' = 0,
 *      right-hand side and solution shifted to zero mean (the Neumann problem fixes p' up to a
 *      constant).
 *   4. correct    F = F* - dt a (p'_c1 - p'_c0),  u = u* - dt grad p',  p = p + p'.
 *   continuity residual = sum_i |sum_f F*_out| (the divergence the projection removes). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    uint32_t *c0, *c1;   /* F */
    double *a, *D, *nh;  /* F, F, 3F: A/D, D and n^ */
} faceinfo;

static faceinfo face_info(const orc_mesh *m) {
    const uint32_t N = m->n_cells, G = m->n_ghosts, F = m->n_faces;
    faceinfo fi;
    fi.c0 = malloc(F * sizeof(uint32_t));
    fi.c1 = malloc(F * sizeof(uint32_t));
    fi.a = malloc(F * sizeof(double));
    fi.D = malloc(F * sizeof(double));
    fi.nh = malloc(3 * (size_t)F * sizeof(double));
    for (uint32_t f = 0; f < F; ++f) fi.c0[f] = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < N + G; ++i)
        for (uint32_t q = m->cell_faces_starts[i]; q < m->cell_faces_starts[i + 1]; ++q) {
            const uint32_t f = m->cell_faces[q], j = m->cell_connects[q];
            if (i < j) {
                fi.c0[f] = i;
                fi.c1[f] = j;
            }
        }
    for (uint32_t f = 0; f < F; ++f) {
        const double *x0 = &m->cell_centers[3 * (size_t)fi.c0[f]], *x1 = &m->cell_centers[3 * (size_t)fi.c1[f]];
        const double *n = &m->face_normals[3 * (size_t)f];
        const double d[3] = {x1[0] - x0[0], x1[1] - x0[1], x1[2] - x0[2]};
        double D = (d[0] * n[0] + d[1] * n[1]) + d[2] * n[2];
        const double sg = D < 0 ? -1.0 : 1.0;
        for (int k = 0; k < 3; ++k) fi.nh[3 * f + k] = n[k] * sg;
        D = D * sg;
        fi.D[f] = D;
        fi.a[f] = m->face_areas[f] / D;
    }
    return fi;
}

static void face_free(faceinfo *fi) {
    free(fi->c0);
    free(fi->c1);
    free(fi->a);
    free(fi->D);
    free(fi->nh);
}

/* wall velocity of every ghost (region table); ghosts of regions not listed: at rest */
static double *ghost_walls(const orc_mesh *m, const orc_nsi_wall *walls, int n_walls) {
    const uint32_t G = m->n_ghosts;
    double *uw = calloc(3 * (size_t)G + 1, sizeof(double));
    for (int w = 0; w < n_walls; ++w) {
        const uint32_t r = walls[w].region;
        if (r >= m->n_regions) continue;
        for (uint32_t g = m->ghost_cell_starts[r]; g < m->ghost_cell_starts[r + 1]; ++g)
            for (int k = 0; k < 3; ++k) uw[3 * g + k] = walls[w].u[k];
    }
    return uw;
}

static void apply_bc(const orc_mesh *m, const double *uw, double *u, double *p) {
    const uint32_t N = m->n_cells, G = m->n_ghosts;
    for (uint32_t g = 0; g < G; ++g) {
        const uint32_t c = N + g, i = m->cell_connects[m->cell_faces_starts[c]];
        for (int k = 0; k < 3; ++k) u[3 * c + k] = 2.0 * uw[3 * g + k] - u[3 * i + k];
        if (p) p[c] = p[i];
    }
}

/* y = A x, A = -L on interior faces: (A x)_i = sum_f a_f (x_i - x_j) */
static void apply_A(const orc_mesh *m, const faceinfo *fi, const double *x, double *y) {
    const uint32_t N = m->n_cells;
    for (uint32_t i = 0; i < N; ++i) {
        double s = 0;
        for (uint32_t q = m->cell_faces_starts[i]; q < m->cell_faces_starts[i + 1]; ++q) {
            const uint32_t j = m->cell_connects[q];
            if (j >= N) continue;
            s += fi->a[m->cell_faces[q]] * (x[i] - x[j]);
        }
        y[i] = s;
    }
}

static double dotN(const double *a, const double *b, uint32_t n) {
    double s = 0;
    for (uint32_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static void remove_mean(double *x, uint32_t n) {
    double s = 0;
    for (uint32_t i = 0; i < n; ++i) s += x[i];
    s /= (double)n;
    for (uint32_t i = 0; i < n; ++i) x[i] -= s;
}

/* Jacobi-preconditioned CG in the Chronopoulos-Gear form (the device's k_cg_iter): with u = M^-1 r
 * and w = A u, gamma = r.u and delta = w.u are formed together after each update, and
 *   beta = gamma / gamma_prev,  alpha = gamma / (delta - beta gamma / alpha_prev)
 *   (first iteration: beta 0, alpha = gamma / delta),
 *   p = u + beta p,  s = w + beta s,  x += alpha p,  r -= alpha s.
 * An iteration is one x update; the loop stops before iteration k >= 1 once |r| < cg_tol |b| or
 * k = cg_max, or when the alpha denominator is 0.  Returns the iterations done. */
int orc_nsi_poisson(const orc_mesh *m, const double *b_in, double *x, int cg_max, double cg_tol) {
    const uint32_t N = m->n_cells;
    faceinfo fi = face_info(m);
    double *r = malloc(N * sizeof(double)), *u = calloc(N + 1, sizeof(double)), *w = malloc(N * sizeof(double));
    double *p = malloc(N * sizeof(double)), *s = malloc(N * sizeof(double)), *dg = malloc(N * sizeof(double));
    memcpy(r, b_in, N * sizeof(double));
    remove_mean(r, N);
    for (uint32_t i = 0; i < N; ++i) {
        double a = 0;
        for (uint32_t k = m->cell_faces_starts[i]; k < m->cell_faces_starts[i + 1]; ++k)
            if (m->cell_connects[k] < N) a += fi.a[m->cell_faces[k]];
        dg[i] = a;
        x[i] = 0;
        u[i] = a > 0 ? r[i] / a : 0.0;
    }
    apply_A(m, &fi, u, w);
    double gam = dotN(r, u, N), del = dotN(w, u, N), rr = dotN(r, r, N);
    const double bb = sqrt(rr);
    double gam_prev = 0, alpha_prev = 0;
    int it = 0;
    if (bb > 0 && cg_max > 0) {
        for (;;) {
            if (it > 0 && ((cg_tol > 0 && sqrt(rr) < cg_tol * bb) || it >= cg_max)) break;
            double beta = 0, den = del;
            if (it > 0) {
                beta = gam_prev != 0 ? gam / gam_prev : 0.0;
                den = del - beta * gam / alpha_prev;
            }
            if (den == 0) break;
            const double alpha = gam / den;
            for (uint32_t i = 0; i < N; ++i) {
                p[i] = it == 0 ? u[i] : u[i] + beta * p[i];
                s[i] = it == 0 ? w[i] : w[i] + beta * s[i];
                x[i] += alpha * p[i];
                r[i] -= alpha * s[i];
                u[i] = dg[i] > 0 ? r[i] / dg[i] : 0.0;
            }
            apply_A(m, &fi, u, w);
            gam_prev = gam;
            alpha_prev = alpha;
            gam = dotN(r, u, N);
            del = dotN(w, u, N);
            rr = dotN(r, r, N);
            ++it;
        }
    }
    remove_mean(x, N);
    free(r);
    free(u);
    free(w);
    free(p);
    free(s);
    free(dg);
    face_free(&fi);
    return it;
}

int orc_nsi_step(const orc_mesh *m, double *u, double *p, double *F, const orc_nsi_wall *walls, int n_walls,
                 const orc_nsi_params *prm, double *cont, int *cg_iters) {
    const uint32_t N = m->n_cells, G = m->n_ghosts, NF = m->n_faces;
    const double dt = prm->dt, nu = prm->nu;
    faceinfo fi = face_info(m);
    double *uw = ghost_walls(m, walls, n_walls);
    double *gu = calloc(9 * (size_t)(N + G), sizeof(double)), *gp = calloc(3 * (size_t)(N + G), sizeof(double));
    double *us = calloc(3 * (size_t)(N + G), sizeof(double)), *Fs = calloc(NF, sizeof(double));
    double *b = calloc(N, sizeof(double)), *pp = calloc(N + G, sizeof(double));
    /* 0. ghosts and gradients of the current state */
    apply_bc(m, uw, u, p);
    orc_vector_gradient(m, u, gu);
    orc_scalar_gradient(m, p, gp);
    /* 1. predictor */
    for (uint32_t i = 0; i < N; ++i) {
        double conv[3] = {0, 0, 0}, diff[3] = {0, 0, 0};
        for (uint32_t q = m->cell_faces_starts[i]; q < m->cell_faces_starts[i + 1]; ++q) {
            const uint32_t f = m->cell_faces[q], j = m->cell_connects[q];
            const double out = i < j ? F[f] : -F[f];
            double uf[3];
            if (j >= N) {
                for (int k = 0; k < 3; ++k) uf[k] = uw[3 * (j - N) + k];
            } else {
                const uint32_t up = out >= 0 ? i : j;
                const double *xf = &m->face_centers[3 * (size_t)f], *xu = &m->cell_centers[3 * (size_t)up];
                const double dd[3] = {xf[0] - xu[0], xf[1] - xu[1], xf[2] - xu[2]};
                for (int k = 0; k < 3; ++k) {
                    const double *g = &gu[9 * (size_t)up + 3 * k];
                    uf[k] = u[3 * up + k] + ((g[0] * dd[0] + g[1] * dd[1]) + g[2] * dd[2]);
                }
            }
            const double a = fi.a[f];
            for (int k = 0; k < 3; ++k) {
                conv[k] += out * uf[k];
                diff[k] += a * (u[3 * j + k] - u[3 * i + k]);
            }
        }
        const double s = dt / m->cell_volumes[i];
        for (int k = 0; k < 3; ++k) us[3 * i + k] = u[3 * i + k] + s * (nu * diff[k] - conv[k]) - dt * gp[3 * i + k];
    }
    apply_bc(m, uw, us, NULL);
    /* 2. predicted face fluxes */
    for (uint32_t f = 0; f < NF; ++f) {
        const uint32_t c0 = fi.c0[f], c1 = fi.c1[f];
        const double *nh = &fi.nh[3 * (size_t)f], A = m->face_areas[f];
        if (c1 >= N) {
            const double *w = &uw[3 * (size_t)(c1 - N)];
            Fs[f] = A * ((w[0] * nh[0] + w[1] * nh[1]) + w[2] * nh[2]);
            continue;
        }
        double un = 0, gn = 0;
        for (int k = 0; k < 3; ++k) {
            un += 0.5 * (us[3 * c0 + k] + us[3 * c1 + k]) * nh[k];
            gn += 0.5 * (gp[3 * c0 + k] + gp[3 * c1 + k]) * nh[k];
        }
        Fs[f] = A * (un + dt * (gn - (p[c1] - p[c0]) / fi.D[f]));
    }
    /* 3. pressure correction */
    double c = 0;
    for (uint32_t i = 0; i < N; ++i) {
        double s = 0;
        for (uint32_t q = m->cell_faces_starts[i]; q < m->cell_faces_starts[i + 1]; ++q) {
            const uint32_t f = m->cell_faces[q], j = m->cell_connects[q];
            s += i < j ? Fs[f] : -Fs[f];
        }
        c += fabs(s);
        b[i] = -s / dt;
    }
    *cont = c;
    const int its = orc_nsi_poisson(m, b, pp, prm->cg_max, prm->cg_tol);
    if (cg_iters) *cg_iters = its;
    /* 4. corrections */
    for (uint32_t g = 0; g < G; ++g) pp[N + g] = pp[m->cell_connects[m->cell_faces_starts[N + g]]];
    double *gpp = calloc(3 * (size_t)(N + G), sizeof(double));
    orc_scalar_gradient(m, pp, gpp);
    for (uint32_t f = 0; f < NF; ++f) {
        const uint32_t c0 = fi.c0[f], c1 = fi.c1[f];
        F[f] = c1 >= N ? Fs[f] : Fs[f] - dt * fi.a[f] * (pp[c1] - pp[c0]);
    }
    for (uint32_t i = 0; i < N; ++i) {
        for (int k = 0; k < 3; ++k) u[3 * i + k] = us[3 * i + k] - dt * gpp[3 * i + k];
        p[i] += pp[i];
    }
    apply_bc(m, uw, u, p);
    free(gpp);
    free(uw);
    free(gu);
    free(gp);
    free(us);
    free(Fs);
    free(b);
    free(pp);
    face_free(&fi);
    return 0;
}

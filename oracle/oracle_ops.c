/* FP64 CPU restatement of the reference's finite-volume operator library (test
 * infrastructure only): the per-cell face sums of include/fvmops.cuh:96-219, with the face
 * loop of fvmOpSetup (:37-51): for each face f of cell c0 with neighbour c1,
 *   d0 = x_f - x_c0, d1 = x_f - x_c1, n = outer_normal(n_f, d0), ds = A_f / V_c0.
 * Fields use the reference's AoS layouts (see oracle.h).  Tensors are written as three columns
 * (u, v, w) of three entries each; outer(a, b) has columns a * b.x, a * b.y, a * b.z
 * (include/datatypes.cuh:348-352). */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static inline double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline double norm3(const double *a) { return sqrt(dot3(a, a)); }
static inline double sgn(double x) { return (double)(x > 0) - (double)(x < 0); }

/* Evaluates operator `op` (ORC_OP_*) of cell c0; out receives 1, 3 or 9 values. */
static void op_cell(const orc_mesh *m, int op, uint32_t c0, const double *val, const double *grad, const double *lim,
                    const double *phi_c, const double *phi_f, double *out) {
    int nout = 1;
    if (op == ORC_OP_SCAL_GRAD_C || op == ORC_OP_SCAL_GRAD_C_CORRECT || op == ORC_OP_BOUND_SCAL_GRAD_C ||
        op == ORC_OP_CONV_VEC_U)
        nout = 3;
    if (op == ORC_OP_VEC_GRAD_C || op == ORC_OP_BOUND_VEC_GRAD_C) nout = 9;
    for (int k = 0; k < nout; ++k) out[k] = 0;
    const double v = m->cell_volumes[c0];
    for (uint32_t q = m->cell_faces_starts[c0]; q < m->cell_faces_starts[c0 + 1]; ++q) {
        const uint32_t f = m->cell_faces[q], c1 = m->cell_connects[q];
        const double *xf = &m->face_centers[3 * (size_t)f];
        double d0[3], d1[3], n[3];
        for (int k = 0; k < 3; ++k) {
            d0[k] = xf[k] - m->cell_centers[3 * (size_t)c0 + k];
            d1[k] = xf[k] - m->cell_centers[3 * (size_t)c1 + k];
        }
        const double *nf = &m->face_normals[3 * (size_t)f];
        const double sg = sgn(dot3(nf, d0));
        for (int k = 0; k < 3; ++k) n[k] = nf[k] * sg;
        const double ds = m->face_areas[f] / v;
        switch (op) {
        case ORC_OP_LAPLACIAN_C: { /* dot(grad1 + grad0, n) * 0.5 * ds */
            double g[3];
            for (int k = 0; k < 3; ++k) g[k] = grad[3 * (size_t)c1 + k] + grad[3 * (size_t)c0 + k];
            out[0] += dot3(g, n) * 0.5 * ds;
            break;
        }
        case ORC_OP_SCAL_GRAD_C: { /* phi = |d0| / |d0 - d1| */
            double dd[3] = {d0[0] - d1[0], d0[1] - d1[1], d0[2] - d1[2]};
            const double phi = norm3(d0) / norm3(dd);
            const double uf = val[c1] * phi + val[c0] * (1.0 - phi);
            for (int k = 0; k < 3; ++k) out[k] += n[k] * uf * ds;
            break;
        }
        case ORC_OP_SCAL_GRAD_C_CORRECT: { /* face centre minus a relative vector, as written */
            double dd[3] = {d0[0] - d1[0], d0[1] - d1[1], d0[2] - d1[2]};
            const double phi = norm3(d0) / norm3(dd);
            double err[3];
            for (int k = 0; k < 3; ++k) err[k] = xf[k] - dd[k] * phi;
            const double uc = dot3(&grad[3 * (size_t)c0], err);
            for (int k = 0; k < 3; ++k) out[k] += n[k] * uc * ds;
            break;
        }
        case ORC_OP_VEC_GRAD_C: { /* phi = |d0| / (|d0| + |d1|); outer(n, u_f) * ds */
            const double phi = norm3(d0) / (norm3(d0) + norm3(d1));
            for (int c = 0; c < 3; ++c) {
                const double w = val[3 * (size_t)c1 + c] * phi + val[3 * (size_t)c0 + c] * (1.0 - phi);
                for (int r = 0; r < 3; ++r) out[3 * c + r] += n[r] * w * ds;
            }
            break;
        }
        case ORC_OP_BOUND_SCAL_GRAD_C: {
            const double du = val[c1] - val[c0];
            for (int k = 0; k < 3; ++k) out[k] += n[k] * du * ds;
            break;
        }
        case ORC_OP_BOUND_VEC_GRAD_C: { /* outer(u1 - u0, n) * ds */
            double du[3];
            for (int r = 0; r < 3; ++r) du[r] = val[3 * (size_t)c1 + r] - val[3 * (size_t)c0 + r];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) out[3 * c + r] += du[r] * n[c] * ds;
            break;
        }
        case ORC_OP_DIV_C: {
            double s[3];
            for (int k = 0; k < 3; ++k) s[k] = val[3 * (size_t)c0 + k] + val[3 * (size_t)c1 + k];
            out[0] += dot3(s, n) * 0.5 * ds;
            break;
        }
        case ORC_OP_CONV_CONST_SCAL_U:
        case ORC_OP_CONV_CONST_SCAL_U2: {
            double u0 = val[c0], u1 = val[c1];
            if (op == ORC_OP_CONV_CONST_SCAL_U2) {
                u0 = val[c0] + dot3(&grad[3 * (size_t)c0], d0) * lim[c0];
                u1 = val[c1] + dot3(&grad[3 * (size_t)c1], d1) * lim[c1];
            }
            const double a0 = dot3(n, phi_c), a1 = dot3(n, phi_c);
            const double fl0 = a0 * u0, fl1 = a1 * u1;
            out[0] += ((fl0 + fl1) - fmax(fabs(a0), fabs(a1)) * (u1 - u0)) * 0.5 * ds;
            break;
        }
        case ORC_OP_CONV_SCAL_U: {
            const double a0 = dot3(n, &phi_f[3 * (size_t)c0]), a1 = dot3(n, &phi_f[3 * (size_t)c1]);
            const double fl0 = a0 * val[c0], fl1 = a1 * val[c1];
            out[0] += ((fl0 + fl1) - fmax(fabs(a0), fabs(a1)) * (val[c1] - val[c0])) * 0.5 * ds;
            break;
        }
        case ORC_OP_CONV_VEC_U: { /* fmax(a0, a1), not of the magnitudes, as written */
            const double a0 = dot3(n, &phi_f[3 * (size_t)c0]), a1 = dot3(n, &phi_f[3 * (size_t)c1]);
            const double am = fmax(a0, a1);
            for (int k = 0; k < 3; ++k) {
                const double fl0 = val[3 * (size_t)c0 + k] * a0, fl1 = val[3 * (size_t)c1 + k] * a1;
                out[k] += ((fl0 + fl1) - (val[3 * (size_t)c1 + k] - val[3 * (size_t)c0 + k]) * am) * 0.5 * ds;
            }
            break;
        }
        default:
            break;
        }
    }
}

int orc_fvm_op(const orc_mesh *m, int op, const double *val, const double *grad, const double *lim,
               const double *phi_const, const double *phi_field, double *out) {
    if (op < 0 || op > ORC_OP_CONV_VEC_U) return 1;
    int nout = 1;
    if (op == ORC_OP_SCAL_GRAD_C || op == ORC_OP_SCAL_GRAD_C_CORRECT || op == ORC_OP_BOUND_SCAL_GRAD_C ||
        op == ORC_OP_CONV_VEC_U)
        nout = 3;
    if (op == ORC_OP_VEC_GRAD_C || op == ORC_OP_BOUND_VEC_GRAD_C) nout = 9;
    for (uint32_t i = 0; i < m->n_cells; ++i)
        op_cell(m, op, i, val, grad, lim, phi_const, phi_field, &out[(size_t)nout * i]);
    return 0;
}

// Finite-volume operators as device functions for user HIP kernels (gfx950, hipcc).
//
// The reference lets a model author write new per-cell operators with its fvmOp / fvmOpSetup
// macros (include/fvmops.cuh:30-77) and ships eleven of them (:96-219).  This header is that
// device API over the meshes and fields of libgraphite_hip.so:
//
//   * for_each_face(mesh, c0, body) runs body(face) for every face of interior cell c0 in the
//     cell's face-slot order with the values fvmOpSetup defines (f, c1, d0 = x_f - x_c0,
//     d1 = x_f - x_c1, n = outer_normal(n_f, d0), ds = A_f / V_c0);
//   * the eleven operators, each a function of (mesh, c0, field[, velocity]).
//
// Views come from the C ABI: gr_mesh_get_view() and gr_field_get_view() fill plain structs of
// device pointers, which kernels take by value.  The library's own gr_field_operator evaluates
// these same functions, so a user kernel gets its results bit for bit.  Every function here is
// compiled without contraction of a*b+c into one rounding (the pragma below), whatever the
// including translation unit's -ffp-contract setting, because the binary64 results match the
// reference's operand order exactly only that way.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "../graphite_hip.h"

namespace graphite {
namespace fvm {

// One face of an interior cell as fvmOpSetup presents it (include/fvmops.cuh:37-51).
struct Face {
    uint32_t f;       // face id
    uint32_t c1;      // the other cell (a ghost for boundary faces)
    double xf[3];     // face centre
    double d0[3];     // x_f - x_c0
    double d1[3];     // x_f - x_c1
    double n[3];      // unit normal oriented by sign(n_f . d0) (outer_normal; 0 when n_f . d0 = 0)
    double ds;        // A_f / V_c0
};

// A field's component blocks: values, gradients (component c, direction k at 3c+k) and
// limiters, each n_cells + n_ghosts entries; S is double (GR_F64) or float (GR_F32).
template <typename S>
struct Field {
    const S *val[3];
    const S *grad[9];
    const S *lim[3];
};

// Field<S> of a gr_field_view (its dtype must be that of S: GR_F64 for double, GR_F32 for float).
template <typename S>
__host__ __device__ inline Field<S> field(const gr_field_view &v) {
    Field<S> f{};
    for (uint32_t c = 0; c < 3; ++c) {
        f.val[c] = c < v.n_comp ? (const S *)v.value[c] : nullptr;
        f.lim[c] = c < v.n_comp ? (const S *)v.limiter[c] : nullptr;
        for (int k = 0; k < 3; ++k) f.grad[3 * c + k] = c < v.n_comp ? (const S *)v.gradient[3 * c + k] : nullptr;
    }
    return f;
}

__device__ inline double dot3(const double a[3], const double b[3]) {
#pragma clang fp contract(off)
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

template <class Body>
__device__ inline void for_each_face(const gr_mesh_view &m, uint32_t c0, Body &&body) {
#pragma clang fp contract(off)
    const double v = m.cell_volume[c0];
    for (uint32_t q = m.cell_faces_starts[c0]; q < m.cell_faces_starts[c0 + 1]; ++q) {
        Face fc;
        fc.f = m.cell_faces[q];
        fc.c1 = m.cell_connects[q];
        const uint32_t f = fc.f, c1 = fc.c1;
        fc.xf[0] = m.face_center_x[f];
        fc.xf[1] = m.face_center_y[f];
        fc.xf[2] = m.face_center_z[f];
        fc.d0[0] = fc.xf[0] - m.cell_center_x[c0];
        fc.d0[1] = fc.xf[1] - m.cell_center_y[c0];
        fc.d0[2] = fc.xf[2] - m.cell_center_z[c0];
        fc.d1[0] = fc.xf[0] - m.cell_center_x[c1];
        fc.d1[1] = fc.xf[1] - m.cell_center_y[c1];
        fc.d1[2] = fc.xf[2] - m.cell_center_z[c1];
        const double nf[3] = {m.face_normal_x[f], m.face_normal_y[f], m.face_normal_z[f]};
        const double dn = dot3(nf, fc.d0);
        const double sg = (double)(dn > 0) - (double)(dn < 0);
        fc.n[0] = nf[0] * sg;
        fc.n[1] = nf[1] * sg;
        fc.n[2] = nf[2] * sg;
        fc.ds = m.face_area[f] / v;
        body(fc);
    }
}

// fvmLaplacianC (include/fvmops.cuh:96-102): sum of avg(grad u).n ds.
template <typename S>
__device__ inline double laplacian_c(const gr_mesh_view &m, uint32_t c0, const Field<S> &u) {
#pragma clang fp contract(off)
    double x = 0;
    for_each_face(m, c0, [&](const Face &fc) {
        double g[3];
        for (int k = 0; k < 3; ++k) g[k] = (double)u.grad[k][fc.c1] + (double)u.grad[k][c0];
        x += dot3(g, fc.n) * 0.5 * fc.ds;
    });
    return x;
}

// fvmScalGradC (:105-112): face value weighted by |d0| / |d0 - d1|.
template <typename S>
__device__ inline void scal_grad_c(const gr_mesh_view &m, uint32_t c0, const Field<S> &u, double out[3]) {
#pragma clang fp contract(off)
    out[0] = out[1] = out[2] = 0;
    for_each_face(m, c0, [&](const Face &fc) {
        const double dd[3] = {fc.d0[0] - fc.d1[0], fc.d0[1] - fc.d1[1], fc.d0[2] - fc.d1[2]};
        const double w = sqrt(dot3(fc.d0, fc.d0)) / sqrt(dot3(dd, dd));
        const double uf = (double)u.val[0][fc.c1] * w + (double)u.val[0][c0] * (1.0 - w);
        for (int k = 0; k < 3; ++k) out[k] += fc.n[k] * uf * fc.ds;
    });
}

// fvmScalGradCCorrect (:114-124): grad u(c0) . (x_f - (d0 - d1) w), the face centre taken as a
// point and the offset as a relative vector, as the reference writes it.
template <typename S>
__device__ inline void scal_grad_c_correct(const gr_mesh_view &m, uint32_t c0, const Field<S> &u, double out[3]) {
#pragma clang fp contract(off)
    out[0] = out[1] = out[2] = 0;
    for_each_face(m, c0, [&](const Face &fc) {
        const double dd[3] = {fc.d0[0] - fc.d1[0], fc.d0[1] - fc.d1[1], fc.d0[2] - fc.d1[2]};
        const double w = sqrt(dot3(fc.d0, fc.d0)) / sqrt(dot3(dd, dd));
        double err[3], g0[3];
        for (int k = 0; k < 3; ++k) {
            err[k] = fc.xf[k] - dd[k] * w;
            g0[k] = (double)u.grad[k][c0];
        }
        const double uc = dot3(g0, err);
        for (int k = 0; k < 3; ++k) out[k] += fc.n[k] * uc * fc.ds;
    });
}

// fvmVecGradC (:126-133): tensor columns n_r w_c ds, face value weighted by |d0| / (|d0| + |d1|).
template <typename S>
__device__ inline void vec_grad_c(const gr_mesh_view &m, uint32_t c0, const Field<S> &u, double out[9]) {
#pragma clang fp contract(off)
    for (int k = 0; k < 9; ++k) out[k] = 0;
    for_each_face(m, c0, [&](const Face &fc) {
        const double l0 = sqrt(dot3(fc.d0, fc.d0)), l1 = sqrt(dot3(fc.d1, fc.d1));
        const double w = l0 / (l0 + l1);
        for (int c = 0; c < 3; ++c) {
            const double uf = (double)u.val[c][fc.c1] * w + (double)u.val[c][c0] * (1.0 - w);
            for (int r = 0; r < 3; ++r) out[3 * c + r] += fc.n[r] * uf * fc.ds;
        }
    });
}

// fvmBoundScalGradC (:136-142): n (u1 - u0) ds.
template <typename S>
__device__ inline void bound_scal_grad_c(const gr_mesh_view &m, uint32_t c0, const Field<S> &u, double out[3]) {
#pragma clang fp contract(off)
    out[0] = out[1] = out[2] = 0;
    for_each_face(m, c0, [&](const Face &fc) {
        const double du = (double)u.val[0][fc.c1] - (double)u.val[0][c0];
        for (int k = 0; k < 3; ++k) out[k] += fc.n[k] * du * fc.ds;
    });
}

// fvmBoundVecGradC (:144-150): outer(u1 - u0, n) ds as tensor columns.
template <typename S>
__device__ inline void bound_vec_grad_c(const gr_mesh_view &m, uint32_t c0, const Field<S> &u, double out[9]) {
#pragma clang fp contract(off)
    for (int k = 0; k < 9; ++k) out[k] = 0;
    for_each_face(m, c0, [&](const Face &fc) {
        double du[3];
        for (int r = 0; r < 3; ++r) du[r] = (double)u.val[r][fc.c1] - (double)u.val[r][c0];
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) out[3 * c + r] += du[r] * fc.n[c] * fc.ds;
    });
}

// fvmDivC (:153-159): avg(u).n ds.
template <typename S>
__device__ inline double div_c(const gr_mesh_view &m, uint32_t c0, const Field<S> &u) {
#pragma clang fp contract(off)
    double x = 0;
    for_each_face(m, c0, [&](const Face &fc) {
        double s[3];
        for (int k = 0; k < 3; ++k) s[k] = (double)u.val[k][c0] + (double)u.val[k][fc.c1];
        x += dot3(s, fc.n) * 0.5 * fc.ds;
    });
    return x;
}

// Rusanov-type convective flux of a scalar with face speeds a0, a1 (dissipation max(|a0|, |a1|)).
__device__ inline double conv_face(double a0, double a1, double u0, double u1, double ds) {
#pragma clang fp contract(off)
    const double fl0 = a0 * u0, fl1 = a1 * u1;
    return ((fl0 + fl1) - fmax(fabs(a0), fabs(a1)) * (u1 - u0)) * 0.5 * ds;
}

// fvmConvConstScalU (:162-175): constant velocity phi.
template <typename S>
__device__ inline double conv_const_scal_u(const gr_mesh_view &m, uint32_t c0, const Field<S> &u, const double phi[3]) {
#pragma clang fp contract(off)
    double x = 0;
    for_each_face(m, c0, [&](const Face &fc) {
        const double a = dot3(fc.n, phi);
        x += conv_face(a, a, (double)u.val[0][c0], (double)u.val[0][fc.c1], fc.ds);
    });
    return x;
}

// fvmConvConstScalU2 (:178-191): as above with both sides reconstructed to the face, u + (grad u . d) lim.
template <typename S>
__device__ inline double conv_const_scal_u2(const gr_mesh_view &m, uint32_t c0, const Field<S> &u, const double phi[3]) {
#pragma clang fp contract(off)
    double x = 0;
    for_each_face(m, c0, [&](const Face &fc) {
        const uint32_t c1 = fc.c1;
        const double g0[3] = {(double)u.grad[0][c0], (double)u.grad[1][c0], (double)u.grad[2][c0]};
        const double g1[3] = {(double)u.grad[0][c1], (double)u.grad[1][c1], (double)u.grad[2][c1]};
        const double u0 = (double)u.val[0][c0] + dot3(g0, fc.d0) * (double)u.lim[0][c0];
        const double u1 = (double)u.val[0][c1] + dot3(g1, fc.d1) * (double)u.lim[0][c1];
        const double a = dot3(fc.n, phi);
        x += conv_face(a, a, u0, u1, fc.ds);
    });
    return x;
}

// fvmConvScalU (:194-206): scalar u convected by the vector field phi (face speeds phi_0.n, phi_1.n).
template <typename S>
__device__ inline double conv_scal_u(const gr_mesh_view &m, uint32_t c0, const Field<S> &u, const Field<S> &phi) {
#pragma clang fp contract(off)
    double x = 0;
    for_each_face(m, c0, [&](const Face &fc) {
        const uint32_t c1 = fc.c1;
        const double p0[3] = {(double)phi.val[0][c0], (double)phi.val[1][c0], (double)phi.val[2][c0]};
        const double p1[3] = {(double)phi.val[0][c1], (double)phi.val[1][c1], (double)phi.val[2][c1]};
        x += conv_face(dot3(fc.n, p0), dot3(fc.n, p1), (double)u.val[0][c0], (double)u.val[0][c1], fc.ds);
    });
    return x;
}

// fvmConvVecU (:208-219): vector u convected by phi; the dissipation speed is fmax(a0, a1) of the signed
// face speeds, as the reference writes it (not of their magnitudes).
template <typename S>
__device__ inline void conv_vec_u(const gr_mesh_view &m, uint32_t c0, const Field<S> &u, const Field<S> &phi,
                                  double out[3]) {
#pragma clang fp contract(off)
    out[0] = out[1] = out[2] = 0;
    for_each_face(m, c0, [&](const Face &fc) {
        const uint32_t c1 = fc.c1;
        const double p0[3] = {(double)phi.val[0][c0], (double)phi.val[1][c0], (double)phi.val[2][c0]};
        const double p1[3] = {(double)phi.val[0][c1], (double)phi.val[1][c1], (double)phi.val[2][c1]};
        const double a0 = dot3(fc.n, p0), a1 = dot3(fc.n, p1);
        const double am = fmax(a0, a1);
        for (int k = 0; k < 3; ++k) {
            const double u0 = (double)u.val[k][c0], u1 = (double)u.val[k][c1];
            const double fl0 = u0 * a0, fl1 = u1 * a1;
            out[k] += ((fl0 + fl1) - (u1 - u0) * am) * 0.5 * fc.ds;
        }
    });
}

// Number of values operator `op` (enum gr_fvm_op) writes per cell: 1, 3 or 9.
__host__ __device__ constexpr int op_outputs(int op) {
    return (op == GR_OP_VEC_GRAD_C || op == GR_OP_BOUND_VEC_GRAD_C) ? 9
           : (op == GR_OP_SCAL_GRAD_C || op == GR_OP_SCAL_GRAD_C_CORRECT || op == GR_OP_BOUND_SCAL_GRAD_C ||
              op == GR_OP_CONV_VEC_U)
               ? 3
               : 1;
}

// Operator `op` at interior cell c0 into out[op_outputs(op)]; phi_c is the constant velocity of
// the ConvConst operators, phi the velocity field of ConvScalU / ConvVecU.
template <typename S>
__device__ inline void evaluate(int op, const gr_mesh_view &m, uint32_t c0, const Field<S> &u, const double phi_c[3],
                                const Field<S> &phi, double *out) {
    switch (op) {
    case GR_OP_LAPLACIAN_C: out[0] = laplacian_c(m, c0, u); break;
    case GR_OP_SCAL_GRAD_C: scal_grad_c(m, c0, u, out); break;
    case GR_OP_SCAL_GRAD_C_CORRECT: scal_grad_c_correct(m, c0, u, out); break;
    case GR_OP_VEC_GRAD_C: vec_grad_c(m, c0, u, out); break;
    case GR_OP_BOUND_SCAL_GRAD_C: bound_scal_grad_c(m, c0, u, out); break;
    case GR_OP_BOUND_VEC_GRAD_C: bound_vec_grad_c(m, c0, u, out); break;
    case GR_OP_DIV_C: out[0] = div_c(m, c0, u); break;
    case GR_OP_CONV_CONST_SCAL_U: out[0] = conv_const_scal_u(m, c0, u, phi_c); break;
    case GR_OP_CONV_CONST_SCAL_U2: out[0] = conv_const_scal_u2(m, c0, u, phi_c); break;
    case GR_OP_CONV_SCAL_U: out[0] = conv_scal_u(m, c0, u, phi); break;
    case GR_OP_CONV_VEC_U: conv_vec_u(m, c0, u, phi, out); break;
    default: break;
    }
}

}  // namespace fvm
}  // namespace graphite

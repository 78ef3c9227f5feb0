// Finite-volume operator library: the reference's per-cell face sums (include/fvmops.cuh:96-219)
// as device functions over a field's SoA blocks, and a kernel that evaluates one operator for
// every interior cell.  The face loop is fvmOpSetup's (:37-51): for each face f of cell c0 with
// neighbour c1, d0 = x_f - x_c0, d1 = x_f - x_c1, n = outer_normal(n_f, d0), ds = A_f / V_c0.
// Expressions keep the reference's operand order (binary64 results are those of the oracle bit
// for bit) and its quirks: fvmScalGradC weights by |d0|/|d0 - d1| while fvmVecGradC uses
// |d0|/(|d0| + |d1|); fvmScalGradCCorrect subtracts a relative vector from the face centre;
// fvmConvVecU's dissipation uses fmax(a0, a1), not of the magnitudes.
#include <hip/hip_runtime.h>

#include <vector>

#include "fvm_kernels.hpp"

namespace gr {

enum FvmOp {
    kOpLaplacianC = 0,
    kOpScalGradC = 1,
    kOpScalGradCCorrect = 2,
    kOpVecGradC = 3,
    kOpBoundScalGradC = 4,
    kOpBoundVecGradC = 5,
    kOpDivC = 6,
    kOpConvConstScalU = 7,
    kOpConvConstScalU2 = 8,
    kOpConvScalU = 9,
    kOpConvVecU = 10
};

__host__ __device__ constexpr int op_outputs(int op) {
    return (op == kOpVecGradC || op == kOpBoundVecGradC) ? 9
           : (op == kOpScalGradC || op == kOpScalGradCCorrect || op == kOpBoundScalGradC || op == kOpConvVecU) ? 3
                                                                                                             : 1;
}

// A field as the operators read it: component values, gradients (component c, direction k at
// block 3c+k) and limiters, each a block of N+G entries.
template <typename S>
struct OpField {
    const S *val[3];
    const S *grad[9];
    const S *lim[3];
};

__device__ inline double odot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

template <typename S>
__device__ void fvm_op_cell(const MeshView &m, int op, uint32_t c0, const OpField<S> &u, const double *phi_c,
                            const OpField<S> &phi, double *out) {
    const int nout = op_outputs(op);
    for (int k = 0; k < nout; ++k) out[k] = 0;
    const double v = m.cv[c0];
    for (uint32_t q = m.cfs[c0]; q < m.cfs[c0 + 1]; ++q) {
        const uint32_t f = m.cf[q], c1 = m.cc[q];
        const double xf[3] = {m.fcx[f], m.fcy[f], m.fcz[f]};
        const double d0[3] = {xf[0] - m.cx[c0], xf[1] - m.cy[c0], xf[2] - m.cz[c0]};
        const double d1[3] = {xf[0] - m.cx[c1], xf[1] - m.cy[c1], xf[2] - m.cz[c1]};
        const double nf[3] = {m.fnx[f], m.fny[f], m.fnz[f]};
        const double dn = odot(nf, d0);
        const double sg = (double)(dn > 0) - (double)(dn < 0);
        const double n[3] = {nf[0] * sg, nf[1] * sg, nf[2] * sg};
        const double ds = m.fa[f] / v;
        switch (op) {
        case kOpLaplacianC: {
            double g[3];
            for (int k = 0; k < 3; ++k) g[k] = (double)u.grad[k][c1] + (double)u.grad[k][c0];
            out[0] += odot(g, n) * 0.5 * ds;
            break;
        }
        case kOpScalGradC: {
            const double dd[3] = {d0[0] - d1[0], d0[1] - d1[1], d0[2] - d1[2]};
            const double phi_w = sqrt(odot(d0, d0)) / sqrt(odot(dd, dd));
            const double uf = (double)u.val[0][c1] * phi_w + (double)u.val[0][c0] * (1.0 - phi_w);
            for (int k = 0; k < 3; ++k) out[k] += n[k] * uf * ds;
            break;
        }
        case kOpScalGradCCorrect: {
            const double dd[3] = {d0[0] - d1[0], d0[1] - d1[1], d0[2] - d1[2]};
            const double phi_w = sqrt(odot(d0, d0)) / sqrt(odot(dd, dd));
            double err[3], g0[3];
            for (int k = 0; k < 3; ++k) {
                err[k] = xf[k] - dd[k] * phi_w;
                g0[k] = (double)u.grad[k][c0];
            }
            const double uc = odot(g0, err);
            for (int k = 0; k < 3; ++k) out[k] += n[k] * uc * ds;
            break;
        }
        case kOpVecGradC: {
            const double l0 = sqrt(odot(d0, d0)), l1 = sqrt(odot(d1, d1));
            const double phi_w = l0 / (l0 + l1);
            for (int c = 0; c < 3; ++c) {
                const double w = (double)u.val[c][c1] * phi_w + (double)u.val[c][c0] * (1.0 - phi_w);
                for (int r = 0; r < 3; ++r) out[3 * c + r] += n[r] * w * ds;
            }
            break;
        }
        case kOpBoundScalGradC: {
            const double du = (double)u.val[0][c1] - (double)u.val[0][c0];
            for (int k = 0; k < 3; ++k) out[k] += n[k] * du * ds;
            break;
        }
        case kOpBoundVecGradC: {
            double du[3];
            for (int r = 0; r < 3; ++r) du[r] = (double)u.val[r][c1] - (double)u.val[r][c0];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) out[3 * c + r] += du[r] * n[c] * ds;
            break;
        }
        case kOpDivC: {
            double s[3];
            for (int k = 0; k < 3; ++k) s[k] = (double)u.val[k][c0] + (double)u.val[k][c1];
            out[0] += odot(s, n) * 0.5 * ds;
            break;
        }
        case kOpConvConstScalU:
        case kOpConvConstScalU2: {
            double u0 = u.val[0][c0], u1 = u.val[0][c1];
            if (op == kOpConvConstScalU2) {
                const double g0[3] = {u.grad[0][c0], u.grad[1][c0], u.grad[2][c0]};
                const double g1[3] = {u.grad[0][c1], u.grad[1][c1], u.grad[2][c1]};
                u0 = (double)u.val[0][c0] + odot(g0, d0) * (double)u.lim[0][c0];
                u1 = (double)u.val[0][c1] + odot(g1, d1) * (double)u.lim[0][c1];
            }
            const double a0 = odot(n, phi_c), a1 = odot(n, phi_c);
            const double fl0 = a0 * u0, fl1 = a1 * u1;
            out[0] += ((fl0 + fl1) - fmax(fabs(a0), fabs(a1)) * (u1 - u0)) * 0.5 * ds;
            break;
        }
        case kOpConvScalU: {
            const double p0[3] = {phi.val[0][c0], phi.val[1][c0], phi.val[2][c0]};
            const double p1[3] = {phi.val[0][c1], phi.val[1][c1], phi.val[2][c1]};
            const double a0 = odot(n, p0), a1 = odot(n, p1);
            const double u0 = u.val[0][c0], u1 = u.val[0][c1];
            const double fl0 = a0 * u0, fl1 = a1 * u1;
            out[0] += ((fl0 + fl1) - fmax(fabs(a0), fabs(a1)) * (u1 - u0)) * 0.5 * ds;
            break;
        }
        case kOpConvVecU: {
            const double p0[3] = {phi.val[0][c0], phi.val[1][c0], phi.val[2][c0]};
            const double p1[3] = {phi.val[0][c1], phi.val[1][c1], phi.val[2][c1]};
            const double a0 = odot(n, p0), a1 = odot(n, p1);
            const double am = fmax(a0, a1);
            for (int k = 0; k < 3; ++k) {
                const double u0 = u.val[k][c0], u1 = u.val[k][c1];
                const double fl0 = u0 * a0, fl1 = u1 * a1;
                out[k] += ((fl0 + fl1) - (u1 - u0) * am) * 0.5 * ds;
            }
            break;
        }
        default:
            break;
        }
    }
}

// out [N][op_outputs(op)]
template <typename S>
__global__ void k_fvm_op(MeshView m, int op, OpField<S> u, double px, double py, double pz, OpField<S> phi,
                         double *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.N) return;
    const double pc[3] = {px, py, pz};
    double o[9];
    fvm_op_cell<S>(m, op, i, u, pc, phi, o);
    const int nout = op_outputs(op);
    for (int k = 0; k < nout; ++k) out[(size_t)nout * i + k] = o[k];
}

}  // namespace gr

using gr::guarded;
using gr::require;

namespace {
template <typename S>
gr::OpField<S> op_field(const gr_field *f) {
    gr::OpField<S> v{};
    for (uint32_t c = 0; c < f->ncomp && c < 3; ++c) {
        v.val[c] = f->val<S>((int)c);
        if (f->grads) {
            v.lim[c] = f->lmt<S>((int)c);
            for (int k = 0; k < 3; ++k) v.grad[3 * c + k] = f->grd<S>((int)c, k);
        }
    }
    return v;
}
}  // namespace

extern "C" int gr_field_operator(const gr_field *u, int op, const double phi_const[3], const gr_field *phi_field,
                                 double *out) {
    return guarded([&] {
        require(u && out, "null argument");
        require(op >= 0 && op <= gr::kOpConvVecU, "unknown operator");
        const bool vec_in = op == gr::kOpVecGradC || op == gr::kOpBoundVecGradC || op == gr::kOpDivC ||
                            op == gr::kOpConvVecU;
        require(u->ncomp == (vec_in ? 3u : 1u), vec_in ? "operator needs a vector field" : "operator needs a scalar field");
        const bool needs_grad = op == gr::kOpLaplacianC || op == gr::kOpScalGradCCorrect || op == gr::kOpConvConstScalU2;
        require(!needs_grad || u->grads, "operator needs the field's gradient (and limiter)");
        const bool const_phi = op == gr::kOpConvConstScalU || op == gr::kOpConvConstScalU2;
        const bool field_phi = op == gr::kOpConvScalU || op == gr::kOpConvVecU;
        require(!const_phi || phi_const, "operator needs a constant velocity");
        if (field_phi) {
            require(phi_field != nullptr, "operator needs a velocity field");
            require(phi_field->ncomp == 3 && phi_field->mesh == u->mesh && phi_field->dtype == u->dtype,
                    "velocity field must be a vector field of the same mesh and dtype");
        }
        gr_mesh *m = u->mesh;
        const uint32_t N = m->h.n_cells;
        const int nout = gr::op_outputs(op);
        gr::DevArray<double> d;
        d.alloc((size_t)N * nout + 1);
        const double px = const_phi ? phi_const[0] : 0, py = const_phi ? phi_const[1] : 0, pz = const_phi ? phi_const[2] : 0;
        if (u->dtype == GR_F64) {
            const gr::OpField<double> pf = field_phi ? op_field<double>(phi_field) : gr::OpField<double>{};
            hipLaunchKernelGGL(gr::k_fvm_op<double>, dim3(gr::blocks_for(N)), dim3(gr::kBlock), 0, m->stream, m->view(),
                               op, op_field<double>(u), px, py, pz, pf, d.p);
        } else {
            const gr::OpField<float> pf = field_phi ? op_field<float>(phi_field) : gr::OpField<float>{};
            hipLaunchKernelGGL(gr::k_fvm_op<float>, dim3(gr::blocks_for(N)), dim3(gr::kBlock), 0, m->stream, m->view(),
                               op, op_field<float>(u), px, py, pz, pf, d.p);
        }
        GR_HIP(hipGetLastError());
        d.download(out, (size_t)N * nout, m->stream);
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

// Finite-volume operator library (include/fvmops.cuh:96-219): gr_field_operator evaluates one
// operator for every interior cell with the device functions of include/graphite/fvm_ops.hpp,
// the header user kernels include, so both get the same results bit for bit.  Expressions keep
// the reference's operand order (binary64 results are those of the oracle bit for bit) and its
// quirks: fvmScalGradC weights by |d0|/|d0 - d1| while fvmVecGradC uses |d0|/(|d0| + |d1|);
// fvmScalGradCCorrect subtracts a relative vector from the face centre; fvmConvVecU's
// dissipation uses fmax(a0, a1), not of the magnitudes.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/graphite/fvm_ops.hpp"
#include "fvm_kernels.hpp"

namespace gr {

// out [N][op_outputs(op)]: the operator functions of include/graphite/fvm_ops.hpp, which user
// kernels call as well.
template <typename S>
__global__ void k_fvm_op(gr_mesh_view m, int op, graphite::fvm::Field<S> u, double px, double py, double pz,
                         graphite::fvm::Field<S> phi, double *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_cells) return;
    const double pc[3] = {px, py, pz};
    double o[9];
    graphite::fvm::evaluate<S>(op, m, i, u, pc, phi, o);
    const int nout = graphite::fvm::op_outputs(op);
    for (int k = 0; k < nout; ++k) out[(size_t)nout * i + k] = o[k];
}

}  // namespace gr

using gr::guarded;
using gr::require;

namespace {
gr_field_view field_view(const gr_field *f) {
    gr_field_view v{};
    v.dtype = f->dtype;
    v.n_comp = f->ncomp;
    v.n = f->n;
    for (uint32_t c = 0; c < f->ncomp && c < 3; ++c) {
        v.value[c] = (void *)(f->value.p + (size_t)c * f->stride * f->esize());
        if (f->grads) {
            v.limiter[c] = (void *)(f->lim.p + (size_t)c * f->stride * f->esize());
            for (int k = 0; k < 3; ++k) v.gradient[3 * c + k] = (void *)(f->grad.p + (size_t)(3 * c + k) * f->stride * f->esize());
        }
    }
    return v;
}
}  // namespace

extern "C" int gr_field_get_view(const gr_field *f, gr_field_view *out) {
    return guarded([&] {
        require(f && out, "null argument");
        *out = field_view(f);
    });
}

extern "C" int gr_field_operator(const gr_field *u, int op, const double phi_const[3], const gr_field *phi_field,
                                 double *out) {
    return guarded([&] {
        require(u && out, "null argument");
        require(op >= GR_OP_LAPLACIAN_C && op <= GR_OP_CONV_VEC_U, "unknown operator");
        const bool vec_in = op == GR_OP_VEC_GRAD_C || op == GR_OP_BOUND_VEC_GRAD_C || op == GR_OP_DIV_C ||
                            op == GR_OP_CONV_VEC_U;
        require(u->ncomp == (vec_in ? 3u : 1u), vec_in ? "operator needs a vector field" : "operator needs a scalar field");
        const bool needs_grad = op == GR_OP_LAPLACIAN_C || op == GR_OP_SCAL_GRAD_C_CORRECT || op == GR_OP_CONV_CONST_SCAL_U2;
        require(!needs_grad || u->grads, "operator needs the field's gradient (and limiter)");
        const bool const_phi = op == GR_OP_CONV_CONST_SCAL_U || op == GR_OP_CONV_CONST_SCAL_U2;
        const bool field_phi = op == GR_OP_CONV_SCAL_U || op == GR_OP_CONV_VEC_U;
        require(!const_phi || phi_const, "operator needs a constant velocity");
        if (field_phi) {
            require(phi_field != nullptr, "operator needs a velocity field");
            require(phi_field->ncomp == 3 && phi_field->mesh == u->mesh && phi_field->dtype == u->dtype,
                    "velocity field must be a vector field of the same mesh and dtype");
        }
        gr_mesh *m = u->mesh;
        const uint32_t N = m->h.n_cells;
        const int nout = graphite::fvm::op_outputs(op);
        gr::DevArray<double> d;
        d.alloc((size_t)N * nout + 1);
        const double px = const_phi ? phi_const[0] : 0, py = const_phi ? phi_const[1] : 0, pz = const_phi ? phi_const[2] : 0;
        gr_mesh_view mv{};
        require(gr_mesh_get_view(m, &mv) == GR_OK, "mesh metrics not computed (gr_mesh_compute_metrics)");
        const gr_field_view uv = field_view(u), pv = field_phi ? field_view(phi_field) : gr_field_view{};
        if (u->dtype == GR_F64)
            hipLaunchKernelGGL(gr::k_fvm_op<double>, dim3(gr::blocks_for(N)), dim3(gr::kBlock), 0, m->stream, mv, op,
                               graphite::fvm::field<double>(uv), px, py, pz, graphite::fvm::field<double>(pv), d.p);
        else
            hipLaunchKernelGGL(gr::k_fvm_op<float>, dim3(gr::blocks_for(N)), dim3(gr::kBlock), 0, m->stream, mv, op,
                               graphite::fvm::field<float>(uv), px, py, pz, graphite::fvm::field<float>(pv), d.p);
        GR_HIP(hipGetLastError());
        d.download(out, (size_t)N * nout, m->stream);
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

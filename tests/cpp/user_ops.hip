// A user's HIP kernels written against include/graphite/fvm_ops.hpp, the device API of the
// operator library (the reference's fvmOp / fvmOpSetup, include/fvmops.cuh:30-77): a kernel that
// composes two library operators and a kernel with an operator of its own over for_each_face.
// Built as a shared library that links libgraphite_hip.so; tests/test_gpu_fvm_ops_api.py loads it.
#include <hip/hip_runtime.h>

#include <graphite/fvm_ops.hpp>

namespace fvm = graphite::fvm;

// laplacian(p) + div(v) per interior cell
__global__ void k_laplacian_plus_div(gr_mesh_view m, fvm::Field<double> p, fvm::Field<double> v, double *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_cells) return;
    out[i] = fvm::laplacian_c(m, i, p) + fvm::div_c(m, i, v);
}

// A new operator: sum of n ds (closure, 3 values) and of (d0 . n) ds (3 for a closed cell with
// planar faces: the divergence of the position vector), per interior cell.
__global__ void k_face_moments(gr_mesh_view m, double *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.n_cells) return;
    double s[4] = {0, 0, 0, 0};
    fvm::for_each_face(m, i, [&](const fvm::Face &f) {
        for (int k = 0; k < 3; ++k) s[k] += f.n[k] * f.ds;
        s[3] += fvm::dot3(f.d0, f.n) * f.ds;
    });
    for (int k = 0; k < 4; ++k) out[4 * (size_t)i + k] = s[k];
}

namespace {
int run(uint32_t n, size_t per_cell, double *host, void (*launch)(double *)) {
    double *d = nullptr;
    if (hipMalloc(&d, n * per_cell * sizeof(double) + 8) != hipSuccess) return 5;
    launch(d);
    int rc = hipGetLastError() == hipSuccess ? 0 : 5;
    if (!rc && hipMemcpy(host, d, n * per_cell * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) rc = 5;
    (void)hipFree(d);
    return rc;
}
}  // namespace

extern "C" int user_laplacian_plus_div(const gr_mesh *mesh, const gr_field *p, const gr_field *v, double *out) {
    static gr_mesh_view m;
    static gr_field_view pv, vv;
    int rc = gr_mesh_get_view(mesh, &m);
    if (!rc) rc = gr_field_get_view(p, &pv);
    if (!rc) rc = gr_field_get_view(v, &vv);
    if (rc) return rc;
    if (pv.dtype != GR_F64 || vv.dtype != GR_F64 || pv.n_comp != 1 || vv.n_comp != 3 || !pv.gradient[2]) return 1;
    return run(m.n_cells, 1, out, [](double *d) {
        hipLaunchKernelGGL(k_laplacian_plus_div, dim3((m.n_cells + 255) / 256), dim3(256), 0, 0, m,
                           fvm::field<double>(pv), fvm::field<double>(vv), d);
    });
}

extern "C" int user_face_moments(const gr_mesh *mesh, double *out) {
    static gr_mesh_view m;
    int rc = gr_mesh_get_view(mesh, &m);
    if (rc) return rc;
    return run(m.n_cells, 4, out, [](double *d) {
        hipLaunchKernelGGL(k_face_moments, dim3((m.n_cells + 255) / 256), dim3(256), 0, 0, m, d);
    });
}

// Fields, FV operators, the Euler model entry points and the device-resident Euler solver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>

#include "fvm_tiles.hpp"

namespace gr {

__global__ __launch_bounds__(kBlock) void k_finalize(const double *partials, uint32_t n, double *out, SolverCtl *ctl,
                                                     double *hist) {
    if (ctl && !ctl->pending) return;
    const double tot = block_sum_partials(partials, n);
    if (threadIdx.x != 0) return;
    if (!ctl) {
        if (out) out[0] = tot;
        return;
    }
    record_iteration(ctl, hist, tot);
}

template <typename S>
__global__ void k_fill(S *p, size_t n, S v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// fvmScalarBcFixed / fvmScalarBcZeroGrad / fvmVectorBcZeroGrad over ghosts [g0, g0+count).
template <typename S>
__global__ void k_bc_generic(MeshView m, S *val, size_t stride, int ncomp, uint32_t g0, uint32_t count, int zero_grad,
                             double v0, double v1, double v2) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const uint32_t c = m.N + g0 + t;
    const uint32_t o = m.cc[m.cfs[c]];
    const double v[3] = {v0, v1, v2};
    for (int k = 0; k < ncomp; ++k) val[k * stride + c] = zero_grad ? val[k * stride + o] : (S)v[k];
}

}  // namespace gr

using gr::guarded;
using gr::require;

void gr_field::ensure_next() {
    if (!next.p) {
        next.alloc(value.n);
        GR_HIP(hipMemsetAsync(next.p, 0, next.n, mesh->stream));
    }
}

namespace {

void check_ready(gr_mesh *m) {
    require(m != nullptr, "null mesh");
    require(m->metrics_ready, "mesh metrics not computed (gr_mesh_compute_metrics)");
}

void check_euler_fields(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, bool need_grads) {
    check_ready(m);
    require(rho && rhou && rhoe, "null field");
    require(rho->mesh == m && rhou->mesh == m && rhoe->mesh == m, "fields belong to another mesh");
    require(rho->ncomp == 1 && rhou->ncomp == 3 && rhoe->ncomp == 1, "rho, rhoe must be scalar and rhou vector fields");
    require(rho->dtype == rhou->dtype && rho->dtype == rhoe->dtype, "rho, rhou, rhoe must share a dtype");
    if (need_grads) require(rho->grads && rhou->grads && rhoe->grads, "fields need gradients and limiters");
}

template <typename S>
gr::EulerArgs<S> euler_args(gr_field *rho, gr_field *rhou, gr_field *rhoe) {
    gr::EulerArgs<S> a{};
    gr_field *f[5] = {rho, rhou, rhou, rhou, rhoe};
    int comp[5] = {0, 0, 1, 2, 0};
    for (int c = 0; c < 5; ++c) {
        a.u[c] = f[c]->val<S>(comp[c]);
        a.un[c] = f[c]->next.p ? f[c]->nxt<S>(comp[c]) : nullptr;
        if (f[c]->grads) {
            a.l[c] = f[c]->lmt<S>(comp[c]);
            for (int k = 0; k < 3; ++k) a.g[3 * c + k] = f[c]->grd<S>(comp[c], k);
        }
    }
    return a;
}

template <typename S>
gr::GradLimArgs<S, 5> euler_gradlim_args(gr_field *rho, gr_field *rhou, gr_field *rhoe) {
    gr::GradLimArgs<S, 5> a{};
    gr_field *f[5] = {rho, rhou, rhou, rhou, rhoe};
    int comp[5] = {0, 0, 1, 2, 0};
    for (int c = 0; c < 5; ++c) {
        a.u[c] = f[c]->val<S>(comp[c]);
        a.l[c] = f[c]->lmt<S>(comp[c]);
        for (int k = 0; k < 3; ++k) a.g[3 * c + k] = f[c]->grd<S>(comp[c], k);
    }
    return a;
}

template <typename S, int NC>
gr::GradLimArgs<S, NC> field_gradlim_args(gr_field *f) {
    gr::GradLimArgs<S, NC> a{};
    for (int c = 0; c < NC; ++c) {
        a.u[c] = f->val<S>(c);
        if (f->grads) {
            a.l[c] = f->lmt<S>(c);
            for (int k = 0; k < 3; ++k) a.g[3 * c + k] = f->grd<S>(c, k);
        }
    }
    return a;
}

template <typename S>
void fill(S *p, size_t n, S v, hipStream_t s) {
    if (n) hipLaunchKernelGGL(gr::k_fill<S>, dim3(gr::blocks_for(n)), dim3(gr::kBlock), 0, s, p, n, v);
}

// Freestream conservative state (euler_init, src/models/euler.cu:385-409), binary64.
void freestream(const double mach[3], double p, double T, double out[5]) {
    const double rho_v = p / T;
    const double c = std::sqrt((1.4 * p) / rho_v);
    const double vel[3] = {mach[0] * c, mach[1] * c, mach[2] * c};
    const double rhoe_v = p / 0.4 + (0.5 * rho_v) * ((vel[0] * vel[0] + vel[1] * vel[1]) + vel[2] * vel[2]);
    out[0] = rho_v;
    for (int k = 0; k < 3; ++k) out[1 + k] = vel[k] * rho_v;
    out[4] = rhoe_v;
}

gr::BcTable one_region_table(const gr_mesh *m, uint32_t region, int kind, const double *mach, double p, double T) {
    gr::BcTable t{};
    t.n = 1;
    t.r[0].begin = m->h.ghost_cell_starts[region];
    t.r[0].end = m->h.ghost_cell_starts[region + 1];
    t.r[0].kind = kind;
    for (int k = 0; k < 3; ++k) t.r[0].mach[k] = mach ? mach[k] : 0.0;
    t.r[0].p = p;
    t.r[0].T = T;
    return t;
}

uint32_t region_of(const gr_mesh *m, const char *name) {
    require(name != nullptr, "null region name");
    for (uint32_t r = 0; r < m->h.n_regions; ++r)
        if (m->h.region_names[r] == name) return r;
    throw gr::Error(GR_ERR_NOT_FOUND, std::string("region '") + name + "' not found in mesh");
}

template <typename S>
void launch_bc(gr_mesh *m, gr::EulerArgs<S> a, const gr::BcTable &t, uint32_t g0, uint32_t count,
               const gr::SolverCtl *ctl) {
    if (!count) return;
    hipLaunchKernelGGL(gr::k_euler_bc<S>, dim3(gr::blocks_for(count)), dim3(gr::kBlock), 0, m->stream, m->view(), a, t,
                       g0, count, ctl);
}

template <typename S>
void launch_flux(gr_mesh *m, int corrected, gr::EulerArgs<S> a, double cfl, S *err, S *dt, double *resid,
                 double *partials, int copy_ghosts, const gr::SolverCtl *ctl) {
    const size_t n = copy_ghosts ? (size_t)m->h.n_cells + m->h.n_ghosts : m->h.n_cells;
    if (corrected)
        hipLaunchKernelGGL((gr::k_euler_flux<S, true>), dim3(gr::blocks_for(n)), dim3(gr::kBlock), 0, m->stream,
                           m->view(), a, cfl, err, dt, resid, partials, copy_ghosts, ctl);
    else
        hipLaunchKernelGGL((gr::k_euler_flux<S, false>), dim3(gr::blocks_for(n)), dim3(gr::kBlock), 0, m->stream,
                           m->view(), a, cfl, err, dt, resid, partials, copy_ghosts, ctl);
}

template <typename S>
void launch_gradlim5(gr_mesh *m, const gr::GradLimArgs<S, 5> &a, int limiter, const gr::SolverCtl *ctl) {
    hipLaunchKernelGGL((gr::k_grad_lim<S, 5, true, true>), dim3(gr::blocks_for(m->h.n_cells)), dim3(gr::kBlock), 0,
                       m->stream, m->view(), a, limiter, ctl);
}

double finalize_sum(gr_mesh *m, uint32_t nparts, const double *partials = nullptr) {
    hipLaunchKernelGGL(gr::k_finalize, dim3(1), dim3(gr::kBlock), 0, m->stream, partials ? partials : m->partials.p, nparts,
                       m->scalar_out.p, (gr::SolverCtl *)nullptr, (double *)nullptr);
    GR_HIP(hipGetLastError());
    GR_HIP(hipMemcpyAsync(m->host_scalar, m->scalar_out.p, sizeof(double), hipMemcpyDeviceToHost, m->stream));
    GR_HIP(hipStreamSynchronize(m->stream));
    return *m->host_scalar;
}

}  // namespace

// ---- fields --------------------------------------------------------------------------------

extern "C" {

int gr_field_create(gr_mesh *m, uint32_t n_comp, int dtype, const char *name, int alloc_grads, gr_field **out) {
    return guarded([&] {
        require(m && out, "null argument");
        require(n_comp == 1 || n_comp == 3, "fields have 1 or 3 components");
        require(dtype == GR_F64 || dtype == GR_F32, "dtype must be GR_F64 or GR_F32");
        m->ensure_device();
        auto *f = new gr_field();
        try {
            f->mesh = m;
            f->ncomp = n_comp;
            f->dtype = dtype;
            f->name = name ? name : "";
            f->grads = alloc_grads != 0;
            f->n = (size_t)m->h.n_cells + m->h.n_ghosts;
            f->stride = gr::pad64(f->n);
            const size_t blk = f->stride * f->esize();
            f->value.alloc(n_comp * blk);
            GR_HIP(hipMemsetAsync(f->value.p, 0, f->value.n, m->stream));
            if (f->grads) {
                f->grad.alloc(3 * n_comp * blk);
                f->lim.alloc(n_comp * blk);
                GR_HIP(hipMemsetAsync(f->grad.p, 0, f->grad.n, m->stream));
                GR_HIP(hipMemsetAsync(f->lim.p, 0, f->lim.n, m->stream));
            }
            GR_HIP(hipStreamSynchronize(m->stream));
        } catch (...) {
            delete f;
            throw;
        }
        *out = f;
    });
}

void gr_field_free(gr_field *f) {
    if (f && f->mesh && f->mesh->stream) (void)hipStreamSynchronize(f->mesh->stream);
    delete f;
}

int gr_field_upload(gr_field *f, const double *value) {
    return guarded([&] {
        require(f && value, "null argument");
        hipStream_t s = f->mesh->stream;
        if (f->dtype == GR_F64) {
            std::vector<double> soa(f->ncomp * f->stride, 0.0);
            for (size_t i = 0; i < f->n; ++i)
                for (uint32_t c = 0; c < f->ncomp; ++c) soa[c * f->stride + i] = value[f->ncomp * i + c];
            GR_HIP(hipMemcpyAsync(f->value.p, soa.data(), soa.size() * 8, hipMemcpyHostToDevice, s));
            GR_HIP(hipStreamSynchronize(s));
        } else {
            std::vector<float> soa(f->ncomp * f->stride, 0.0f);
            for (size_t i = 0; i < f->n; ++i)
                for (uint32_t c = 0; c < f->ncomp; ++c) soa[c * f->stride + i] = (float)value[f->ncomp * i + c];
            GR_HIP(hipMemcpyAsync(f->value.p, soa.data(), soa.size() * 4, hipMemcpyHostToDevice, s));
            GR_HIP(hipStreamSynchronize(s));
        }
    });
}

int gr_field_download(const gr_field *f, double *value, double *gradient, double *limiter) {
    return guarded([&] {
        require(f != nullptr, "null argument");
        require(f->grads || (!gradient && !limiter), "field has no gradient/limiter arrays");
        hipStream_t s = f->mesh->stream;
        auto get = [&](const gr::DevArray<char> &arr, int blocks, double *dst, int per_cell) {
            if (!dst) return;
            std::vector<char> raw(arr.n);
            arr.download(raw.data(), raw.size(), s);
            GR_HIP(hipStreamSynchronize(s));
            for (size_t i = 0; i < f->n; ++i)
                for (int b = 0; b < blocks; ++b) {
                    size_t idx = (size_t)b * f->stride + i;
                    dst[(size_t)per_cell * i + b] =
                        f->dtype == GR_F64 ? ((const double *)raw.data())[idx] : (double)((const float *)raw.data())[idx];
                }
        };
        get(f->value, (int)f->ncomp, value, (int)f->ncomp);
        get(f->grad, 3 * (int)f->ncomp, gradient, 3 * (int)f->ncomp);
        get(f->lim, (int)f->ncomp, limiter, (int)f->ncomp);
    });
}

int gr_field_init(gr_field *f, const double *v) {
    return guarded([&] {
        require(f && v, "null argument");
        hipStream_t s = f->mesh->stream;
        for (uint32_t c = 0; c < f->ncomp; ++c) {
            if (f->dtype == GR_F64) {
                fill<double>(f->val<double>(c), f->n, v[c], s);
                if (f->grads) {
                    fill<double>(f->lmt<double>(c), f->n, 1.0, s);
                    for (int k = 0; k < 3; ++k) fill<double>(f->grd<double>(c, k), f->n, 0.0, s);
                }
            } else {
                fill<float>(f->val<float>(c), f->n, (float)v[c], s);
                if (f->grads) {
                    fill<float>(f->lmt<float>(c), f->n, 1.0f, s);
                    for (int k = 0; k < 3; ++k) fill<float>(f->grd<float>(c, k), f->n, 0.0f, s);
                }
            }
        }
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(s));
    });
}

int gr_field_reduce_sum(const gr_field *f, double *out) {
    return guarded([&] {
        require(f && out, "null argument");
        require(f->ncomp == 1, "reduce_sum is defined for scalar fields");
        gr_mesh *m = f->mesh;
        const uint32_t N = m->h.n_cells;
        const uint32_t nb = gr::blocks_for(N);
        require(nb <= m->partials.n, "reduction scratch too small");
        if (f->dtype == GR_F64)
            hipLaunchKernelGGL(gr::k_field_partial_sum<double>, dim3(nb), dim3(gr::kBlock), 0, m->stream,
                               f->val<double>(0), N, m->partials.p);
        else
            hipLaunchKernelGGL(gr::k_field_partial_sum<float>, dim3(nb), dim3(gr::kBlock), 0, m->stream,
                               f->val<float>(0), N, m->partials.p);
        *out = finalize_sum(m, nb);
    });
}

static int field_gradlim(gr_field *f, bool grad, bool lim, int kind) {
    return guarded([&] {
        require(f != nullptr, "null argument");
        check_ready(f->mesh);
        require(f->grads, "field has no gradient/limiter arrays");
        require(kind >= 0 && kind <= 2, "unknown limiter");
        gr_mesh *m = f->mesh;
        const dim3 grid(gr::blocks_for(m->h.n_cells)), blk(gr::kBlock);
        const gr::SolverCtl *nc = nullptr;
        if (!m->h.n_cells) return;
#define GR_GL(S, NC)                                                                                         \
    do {                                                                                                     \
        auto a = field_gradlim_args<S, NC>(f);                                                               \
        if (grad)                                                                                            \
            hipLaunchKernelGGL((gr::k_grad_lim<S, NC, true, false>), grid, blk, 0, m->stream, m->view(), a, kind, nc); \
        if (lim)                                                                                             \
            hipLaunchKernelGGL((gr::k_grad_lim<S, NC, false, true>), grid, blk, 0, m->stream, m->view(), a, kind, nc); \
    } while (0)
        if (f->dtype == GR_F64) {
            if (f->ncomp == 1) GR_GL(double, 1);
            else GR_GL(double, 3);
        } else {
            if (f->ncomp == 1) GR_GL(float, 1);
            else GR_GL(float, 3);
        }
#undef GR_GL
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

int gr_field_compute_gradient(gr_field *f) { return field_gradlim(f, true, false, 0); }
int gr_field_compute_limiter(gr_field *f, int limiter) { return field_gradlim(f, false, true, limiter); }

int gr_field_gradlim_bc(gr_field *f) {
    return guarded([&] {
        require(f != nullptr, "null argument");
        check_ready(f->mesh);
        require(f->grads, "field has no gradient/limiter arrays");
        gr_mesh *m = f->mesh;
        if (!m->h.n_ghosts) return;
        const dim3 grid(gr::blocks_for(m->h.n_ghosts)), blk(gr::kBlock);
        if (f->dtype == GR_F64) {
            if (f->ncomp == 1)
                hipLaunchKernelGGL((gr::k_gradlim_ghost<double, 1>), grid, blk, 0, m->stream, m->view(),
                                   field_gradlim_args<double, 1>(f));
            else
                hipLaunchKernelGGL((gr::k_gradlim_ghost<double, 3>), grid, blk, 0, m->stream, m->view(),
                                   field_gradlim_args<double, 3>(f));
        } else {
            if (f->ncomp == 1)
                hipLaunchKernelGGL((gr::k_gradlim_ghost<float, 1>), grid, blk, 0, m->stream, m->view(),
                                   field_gradlim_args<float, 1>(f));
            else
                hipLaunchKernelGGL((gr::k_gradlim_ghost<float, 3>), grid, blk, 0, m->stream, m->view(),
                                   field_gradlim_args<float, 3>(f));
        }
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

static int field_bc(gr_field *f, uint32_t region, const double *v, int zero_grad) {
    return guarded([&] {
        require(f != nullptr, "null argument");
        check_ready(f->mesh);
        gr_mesh *m = f->mesh;
        require(region < m->h.n_regions, "region index out of range");
        require(zero_grad || v, "null value");
        const uint32_t g0 = m->h.ghost_cell_starts[region], cnt = m->h.ghost_cell_starts[region + 1] - g0;
        if (!cnt) return;
        double vv[3] = {0, 0, 0};
        if (v)
            for (uint32_t k = 0; k < f->ncomp; ++k) vv[k] = v[k];
        if (f->dtype == GR_F64)
            hipLaunchKernelGGL(gr::k_bc_generic<double>, dim3(gr::blocks_for(cnt)), dim3(gr::kBlock), 0, m->stream,
                               m->view(), f->val<double>(0), f->stride, (int)f->ncomp, g0, cnt, zero_grad, vv[0],
                               vv[1], vv[2]);
        else
            hipLaunchKernelGGL(gr::k_bc_generic<float>, dim3(gr::blocks_for(cnt)), dim3(gr::kBlock), 0, m->stream,
                               m->view(), f->val<float>(0), f->stride, (int)f->ncomp, g0, cnt, zero_grad, vv[0],
                               vv[1], vv[2]);
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

int gr_field_bc_fixed(gr_field *f, uint32_t region, const double *v) { return field_bc(f, region, v, 0); }
int gr_field_bc_zero_grad(gr_field *f, uint32_t region) { return field_bc(f, region, nullptr, 1); }

// ---- Euler model --------------------------------------------------------------------------

int gr_euler_init(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, const double mach[3], double pressure,
                  double temperature) {
    return guarded([&] {
        require(mach != nullptr, "null mach");
        check_euler_fields(m, rho, rhou, rhoe, false);
        double fs[5];
        freestream(mach, pressure, temperature, fs);
        gr_field *f[5] = {rho, rhou, rhou, rhou, rhoe};
        int comp[5] = {0, 0, 1, 2, 0};
        for (int c = 0; c < 5; ++c) {
            if (rho->dtype == GR_F64) fill<double>(f[c]->val<double>(comp[c]), f[c]->n, fs[c], m->stream);
            else fill<float>(f[c]->val<float>(comp[c]), f[c]->n, (float)fs[c], m->stream);
        }
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

static int euler_bc(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, int kind, const double *mach, double p,
                    double T, const char *region) {
    return guarded([&] {
        check_euler_fields(m, rho, rhou, rhoe, false);
        if (kind == GR_BC_FARFIELD) require(mach != nullptr, "null mach");
        const uint32_t r = region_of(m, region);
        gr::BcTable t = one_region_table(m, r, kind, mach, p, T);
        const uint32_t g0 = t.r[0].begin, cnt = t.r[0].end - t.r[0].begin;
        if (rho->dtype == GR_F64) launch_bc<double>(m, euler_args<double>(rho, rhou, rhoe), t, g0, cnt, nullptr);
        else launch_bc<float>(m, euler_args<float>(rho, rhou, rhoe), t, g0, cnt, nullptr);
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

int gr_euler_bc_slip(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, const char *region) {
    return euler_bc(m, rho, rhou, rhoe, GR_BC_SLIP, nullptr, 0, 0, region);
}

int gr_euler_bc_farfield(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, const double mach[3],
                         double pressure, double temperature, const char *region) {
    return euler_bc(m, rho, rhou, rhoe, GR_BC_FARFIELD, mach, pressure, temperature, region);
}

int gr_euler_step_prepare(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, int limiter) {
    return guarded([&] {
        check_euler_fields(m, rho, rhou, rhoe, true);
        require(limiter >= 0 && limiter <= 2, "unknown limiter");
        if (!m->h.n_cells) return;
        auto ghost = [&](auto a, auto tag) {
            using S = decltype(tag);
            if (m->h.n_ghosts)
                hipLaunchKernelGGL((gr::k_gradlim_ghost<S, 5>), dim3(gr::blocks_for(m->h.n_ghosts)), dim3(gr::kBlock),
                                   0, m->stream, m->view(), a);
        };
        if (rho->dtype == GR_F64) {
            auto a = euler_gradlim_args<double>(rho, rhou, rhoe);
            launch_gradlim5<double>(m, a, limiter, nullptr);
            ghost(a, 0.0);
        } else {
            auto a = euler_gradlim_args<float>(rho, rhou, rhoe);
            launch_gradlim5<float>(m, a, limiter, nullptr);
            ghost(a, 0.0f);
        }
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

int gr_euler_step(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, double cfl, const char *time_scheme,
                  gr_field *dt, gr_field *error) {
    return guarded([&] {
        check_euler_fields(m, rho, rhou, rhoe, true);
        require(time_scheme != nullptr, "null time scheme");
        // The reference only implements "steady" (src/models/euler.cu:661); other schemes
        // leave the state unchanged there, and are rejected here.
        require(std::strcmp(time_scheme, "steady") == 0, "only the 'steady' time scheme exists");
        for (gr_field *o : {dt, error})
            if (o) require(o->mesh == m && o->ncomp == 1 && o->dtype == rho->dtype, "dt/error must be scalar fields of the state dtype");
        rho->ensure_next();
        rhou->ensure_next();
        rhoe->ensure_next();
        if (rho->dtype == GR_F64)
            launch_flux<double>(m, m->corrected, euler_args<double>(rho, rhou, rhoe), cfl,
                                error ? error->val<double>(0) : nullptr, dt ? dt->val<double>(0) : nullptr, nullptr,
                                nullptr, 1, nullptr);
        else
            launch_flux<float>(m, m->corrected, euler_args<float>(rho, rhou, rhoe), cfl,
                               error ? error->val<float>(0) : nullptr, dt ? dt->val<float>(0) : nullptr, nullptr,
                               nullptr, 1, nullptr);
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(m->stream));
        rho->swap_next();
        rhou->swap_next();
        rhoe->swap_next();
    });
}

int gr_euler_residual(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, double *resid) {
    return guarded([&] {
        check_euler_fields(m, rho, rhou, rhoe, true);
        require(resid != nullptr, "null output");
        rho->ensure_next();
        rhou->ensure_next();
        rhoe->ensure_next();
        gr::DevArray<double> r;
        r.alloc(5 * std::max<size_t>(1, m->h.n_cells));
        if (rho->dtype == GR_F64)
            launch_flux<double>(m, m->corrected, euler_args<double>(rho, rhou, rhoe), 1.0, nullptr, nullptr, r.p,
                                nullptr, 0, nullptr);
        else
            launch_flux<float>(m, m->corrected, euler_args<float>(rho, rhou, rhoe), 1.0, nullptr, nullptr, r.p,
                               nullptr, 0, nullptr);
        GR_HIP(hipGetLastError());
        r.download(resid, 5 * (size_t)m->h.n_cells, m->stream);
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

}  // extern "C"

// ---- residual without allocation -------------------------------------------------------------
// The staging buffer of gr_euler_residual_noalloc: a file-scope device array each GPU holds once
// for the life of the process (32 MiB), used through its device address, which is looked up once
// per device (hipGetSymbolAddress, that device current) and kept in g_scratch_ptr.  Calls that
// overlap on one GPU share it and race.
namespace {
constexpr size_t kScratchDoubles = size_t(1) << 22;
constexpr int kMaxDevices = 64;
static_assert(kScratchDoubles / 5 == GR_RESIDUAL_NOALLOC_MAX_CELLS, "header and scratch size disagree");
__device__ __align__(256) double scratch[kScratchDoubles];
double *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mutex;

double *scratch_ptr(int device) {
    require(device >= 0 && device < kMaxDevices, "device id out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        GR_HIP(hipGetDevice(&cur));
        require(cur == device, "the mesh's device must be the current device");
        void *p = nullptr;
        GR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "no device address for the residual scratch array");
        g_scratch_ptr[device] = (double *)p;
    }
    return g_scratch_ptr[device];
}
}  // namespace

extern "C" {

int gr_euler_residual_noalloc(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, double *resid) {
    return guarded([&] {
        check_euler_fields(m, rho, rhou, rhoe, true);
        require(resid != nullptr, "null output");
        const size_t N = m->h.n_cells;
        if (N > kScratchDoubles / 5)
            throw gr::Error(GR_ERR_INVALID, "gr_euler_residual_noalloc: " + std::to_string(N) +
                                                " cells exceed its maximum of " + std::to_string(kScratchDoubles / 5) +
                                                " (use gr_euler_residual)");
        double *r = scratch_ptr(m->device);
        // the kernel writes all 5 entries of every interior cell (nothing carries over between calls),
        // and no new state (un = null)
        if (rho->dtype == GR_F64) {
            auto a = euler_args<double>(rho, rhou, rhoe);
            for (auto &p : a.un) p = nullptr;
            launch_flux<double>(m, m->corrected, a, 1.0, nullptr, nullptr, r, nullptr, 0, nullptr);
        } else {
            auto a = euler_args<float>(rho, rhou, rhoe);
            for (auto &p : a.un) p = nullptr;
            launch_flux<float>(m, m->corrected, a, 1.0, nullptr, nullptr, r, nullptr, 0, nullptr);
        }
        GR_HIP(hipGetLastError());
        GR_HIP(hipMemcpyAsync(resid, r, 5 * N * sizeof(double), hipMemcpyDeviceToHost, m->stream));
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

int gr_euler_set_corrected_reconstruction(gr_mesh *m, int on) {
    return guarded([&] {
        require(m != nullptr, "null mesh");
        m->corrected = on ? 1 : 0;
    });
}

}  // extern "C"

// ---- solver --------------------------------------------------------------------------------

struct gr_solver {
    gr_mesh *mesh = nullptr;
    gr_solver_config cfg{};
    size_t stride = 0;
    gr::DevArray<char> ubuf[2], grad;   // grad: gradient + limiter records (kGL<S> per cell)
    gr::DevArray<char> slot_geo;        // flux, gradient and ghost geometry records
    gr::DevArray<float> lsq32;          // binary32 LSQ matrices (binary64 solvers use the mesh's)
    gr::DevArray<uint32_t> tiles;       // tile records, kTileRec words each
    uint32_t ntiles = 0, tile_cells = 0;
    gr::DevArray<double> partials;      // one V*|R|^2 partial per tile
    int cur = 0;         // buffer holding the current interior state
    int ghost_src = 0;   // buffer holding the latest ghost values
    struct Bc { bool set = false; int kind = 0; double mach[3] = {0, 0, 0}, p = 0, T = 0; };
    std::vector<Bc> bcs;
    gr::DevArray<gr::SolverCtl> ctl;
    gr::DevArray<double> hist;
    int hist_cap = 0;
    gr::SolverCtl *ctl_host = nullptr;   // pinned: [0] upload, [1],[2] polling
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int graph_iters = 20;   // iterations per captured graph (even)
    int last_iters = 0;

    ~gr_solver() {
        if (mesh && mesh->stream) (void)hipStreamSynchronize(mesh->stream);
        drop_graphs();
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (ctl_host) (void)hipHostFree(ctl_host);
    }
    void drop_graphs() {
        for (auto &g : exec) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }
    size_t es() const { return (size_t)cfg.dtype; }
    // State records (kUR values per cell and ghost, `stride` records) of buffer b.
    template <typename S> S *rec(int b) const { return (S *)ubuf[b].p; }

    // Values taken by the flux records, rounded up so the gradient records start 16-byte aligned.
    template <typename S> static size_t flux_span(size_t slots) {
        const size_t w = 16 / sizeof(S);
        return (slots * gr::kFluxRec + w - 1) / w * w;
    }
    // gradient records, rounded up to whole 16-byte words (the ghost records after them are
    // read with 16-byte loads)
    template <typename S> static size_t grad_span(size_t slots) {
        const size_t w = 16 / sizeof(S);
        return (slots * gr::kGradRec + w - 1) / w * w;
    }
    template <typename S> gr::SolverArgs<S> sargs(int b) const {
        gr::SolverArgs<S> a{};
        a.u = rec<S>(b);
        a.un = rec<S>(1 - b);
        a.gl = (S *)grad.p;
        const size_t slots = mesh->h.cell_faces_starts[mesh->h.n_cells];
        a.frec = (const S *)slot_geo.p;
        a.grec = a.frec + flux_span<S>(slots);
        a.gnrm = a.grec + grad_span<S>(slots);
        if constexpr (sizeof(S) == 8) a.lsq = (const S *)mesh->mat.p;
        else a.lsq = (const S *)lsq32.p;
        a.lsq_stride = std::max<size_t>(64, mesh->mstride());
        return a;
    }
    // Greedy cut of the interior cells into runs of at most kTileSlots face slots and max_cells
    // cells (one block of the tile kernels each).  Captured graphs hold the old tiles: dropped.
    void build_tiles(uint32_t max_cells) {
        gr_mesh *m = mesh;
        require(max_cells >= 1 && max_cells <= (uint32_t)gr::kTileCells, "tile cell count out of range");
        std::vector<uint32_t> t{0};
        const auto &st = m->h.cell_faces_starts;
        uint32_t first = 0;
        for (uint32_t c = 0; c < m->h.n_cells; ++c) {
            require(st[c + 1] - st[c] <= (uint32_t)gr::kTileSlots, "cell with more face slots than a tile holds");
            // the gradient kernel keeps a cell's results in its first three slot columns
            require(st[c + 1] - st[c] >= 3, "cell with fewer than three faces");
            if (st[c + 1] - st[first] > (uint32_t)gr::kTileSlots || c + 1 - first > max_cells) {
                t.push_back(c);
                first = c;
            }
        }
        if (t.back() != m->h.n_cells) t.push_back(m->h.n_cells);
        std::vector<uint32_t> rec((t.size() - 1) * gr::kTileRec, 0u);
        for (size_t k = 0; k + 1 < t.size(); ++k) {
            uint32_t *r = &rec[k * gr::kTileRec];
            const uint32_t c0 = t[k], c1 = t[k + 1], q0 = st[c0];
            r[0] = c0;
            r[1] = q0;
            r[2] = (c1 - c0) | (st[c1] - q0) << 16;
            uint8_t *b = reinterpret_cast<uint8_t *>(r + 3);
            for (uint32_t c = c0; c <= c1; ++c) b[c - c0] = (uint8_t)(st[c] - q0);
        }
        GR_HIP(hipStreamSynchronize(m->stream));   // no launch may still read the old tiles
        drop_graphs();
        ntiles = (uint32_t)t.size() - 1;
        tile_cells = max_cells;
        tiles.alloc(std::max<size_t>(1, rec.size()));
        tiles.upload(rec.data(), rec.size(), m->stream);
        partials.alloc(ntiles + 64);
        GR_HIP(hipStreamSynchronize(m->stream));
    }
    // Mean durations (ms) of n back-to-back launches of the gradient, the flux and the residual
    // sum kernels on the current state (not advanced), HIP events on the library's stream around
    // each run.
    void time_kernels(int n, double ms[3]) {
        gr_mesh *m = mesh;
        hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
        struct Guard {
            hipEvent_t *e;
            ~Guard() {
                for (int k = 0; k < 4; ++k)
                    if (e[k]) (void)hipEventDestroy(e[k]);
            }
        } guard{ev};
        for (auto &e : ev) GR_HIP(hipEventCreate(&e));
        const int b = cur;
        // the flux launches write the interior of buffer 1-b, which holds the state the ghost
        // values of a download are built from: saved here and restored after the runs
        gr::DevArray<char> saved;
        saved.alloc(ubuf[1 - b].n);
        GR_HIP(hipMemcpyAsync(saved.p, ubuf[1 - b].p, saved.n, hipMemcpyDeviceToDevice, m->stream));
        GR_HIP(hipEventRecord(ev[0], m->stream));
        for (int k = 0; k < n; ++k) {
            if (cfg.dtype == GR_F64) launch_gradlim<double>(b, nullptr);
            else launch_gradlim<float>(b, nullptr);
        }
        GR_HIP(hipEventRecord(ev[1], m->stream));
        for (int k = 0; k < n; ++k) {
            if (cfg.dtype == GR_F64) launch_flux<double>(b, nullptr);
            else launch_flux<float>(b, nullptr);
        }
        GR_HIP(hipEventRecord(ev[2], m->stream));
        for (int k = 0; k < n; ++k)
            hipLaunchKernelGGL(gr::k_finalize, dim3(1), dim3(gr::kBlock), 0, m->stream, partials.p, ntiles,
                               m->scalar_out.p, (gr::SolverCtl *)nullptr, (double *)nullptr);
        GR_HIP(hipEventRecord(ev[3], m->stream));
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(m->stream));
        for (int j = 0; j < 3; ++j) {
            float t = 0;
            GR_HIP(hipEventElapsedTime(&t, ev[j], ev[j + 1]));
            ms[j] = t / n;
        }
        GR_HIP(hipMemcpyAsync(ubuf[1 - b].p, saved.p, saved.n, hipMemcpyDeviceToDevice, m->stream));
        GR_HIP(hipStreamSynchronize(m->stream));
    }
    // Per-slot geometry records of the tile kernels, from the mesh metrics.
    template <typename S> void build_slot_geometry() {
        gr_mesh *m = mesh;
        const size_t slots = m->h.cell_faces_starts[m->h.n_cells];
        slot_geo.alloc(std::max<size_t>(1, (flux_span<S>(slots) + grad_span<S>(slots) + (size_t)m->h.n_ghosts * gr::kGhostRec) *
                                              sizeof(S)));
        if (sizeof(S) == 4) lsq32.alloc(6 * std::max<size_t>(64, m->mstride()));
        const gr::SolverArgs<S> a = sargs<S>(0);
        if (m->h.n_cells && sizeof(S) == 4)
            hipLaunchKernelGGL(gr::k_cast_lsq<float>, dim3(gr::blocks_for(m->h.n_cells)), dim3(gr::kBlock), 0,
                               m->stream, m->view(), lsq32.p, a.lsq_stride);
        if (m->h.n_cells)
            hipLaunchKernelGGL(gr::k_slot_geometry<S>, dim3(gr::blocks_for(m->h.n_cells)), dim3(gr::kBlock), 0, m->stream,
                               m->view(), (S *)a.frec, (S *)a.grec, (S *)a.gnrm);
        GR_HIP(hipGetLastError());
    }
    gr::BcTable table() const {
        gr::BcTable t{};
        for (uint32_t r = 0; r < mesh->h.n_regions; ++r) {
            if (!bcs[r].set) continue;
            auto &R = t.r[t.n++];
            R.begin = mesh->h.ghost_cell_starts[r];
            R.end = mesh->h.ghost_cell_starts[r + 1];
            R.kind = bcs[r].kind;
            for (int k = 0; k < 3; ++k) R.mach[k] = bcs[r].mach[k];
            R.p = bcs[r].p;
            R.T = bcs[r].T;
        }
        return t;
    }
    // An iteration reading buffer b: gradient+limiter, flux+update into 1-b, norm, with the
    // boundary conditions evaluated inside both kernels.
    template <typename S> void launch_gradlim(int b, gr::SolverCtl *c) {
        gr_mesh *m = mesh;
        hipLaunchKernelGGL(gr::k_gradlim_tile<S>, dim3(ntiles), dim3(gr::kBlock), 0, m->stream, m->view(), sargs<S>(b),
                           table(), tiles.p, ntiles, cfg.limiter, cfg.corrected, c, (const double *)partials.p, hist.p);
    }
    template <typename S> void launch_flux(int b, gr::SolverCtl *c) {
        gr_mesh *m = mesh;
        const dim3 grid(ntiles), blk(gr::kBlock);
        if (cfg.corrected)
            hipLaunchKernelGGL((gr::k_flux_tile<S, true>), grid, blk, 0, m->stream, m->view(), sargs<S>(b), table(),
                               tiles.p, ntiles, cfg.cfl, partials.p, c);
        else
            hipLaunchKernelGGL((gr::k_flux_tile<S, false>), grid, blk, 0, m->stream, m->view(), sargs<S>(b), table(),
                               tiles.p, ntiles, cfg.cfl, partials.p, c);
    }
    // One iteration reading buffer b; without a control block the caller reduces the partials.
    // With one, the iteration's residual is summed and recorded by the next iteration's gradient
    // launch (or by finish_pending after the last).
    void enqueue(int b, gr::SolverCtl *c) {
        if (cfg.dtype == GR_F64) {
            launch_gradlim<double>(b, c);
            launch_flux<double>(b, c);
        } else {
            launch_gradlim<float>(b, c);
            launch_flux<float>(b, c);
        }
    }
    // Sums and records the last iteration's residual if no gradient launch has (k_finalize
    // returns at once when nothing is pending).
    void finish_pending(gr::SolverCtl *c) {
        hipLaunchKernelGGL(gr::k_finalize, dim3(1), dim3(gr::kBlock), 0, mesh->stream, partials.p, ntiles,
                           (double *)nullptr, c, hist.p);
    }
    // Writes the boundary-condition ghost states of buffer b's interior into buffer b (the
    // iteration kernels evaluate them on the fly and never store them).
    void materialise_ghosts(int b) {
        if (cfg.dtype == GR_F64) ghost_states<double>(b);
        else ghost_states<float>(b);
        GR_HIP(hipGetLastError());
    }
    template <typename S> void ghost_states(int b) {
        gr_mesh *m = mesh;
        if (!m->h.n_ghosts) return;
        const gr::SolverArgs<S> a = sargs<S>(b);
        hipLaunchKernelGGL(gr::k_ghost_states<S>, dim3(gr::blocks_for(m->h.n_ghosts)), dim3(gr::kBlock), 0, m->stream,
                           m->view(), rec<S>(b), a.gnrm, table());
    }
    // Algorithmic bytes per launch: each array element a kernel reads or writes, once.
    void kernel_bytes(double out[4]) const {
        const double N = mesh->h.n_cells, G = mesh->h.n_ghosts, F = mesh->h.n_faces;
        const double slots = mesh->h.cell_faces_starts[mesh->h.n_cells];   // interior cells' face slots
        const double su = (double)cfg.dtype, sg = 8.0;
        out[0] = 0.0;   // boundary conditions are evaluated inside the two kernels below
        // gradient+limiter: U of cells, centres of cells and ghosts, LSQ matrix (6), face
        // centres, normals of ghost faces (fused BC), CSR (starts, faces, neighbours);
        // gradient (15) and limiter (5) written.
        out[1] = su * (5 * N) + sg * (3 * (N + G) + 6 * N + 3 * F + 3 * G) + 4.0 * (N + 1 + 2 * slots) +
                 su * (20 * N);
        // flux+update: U, gradient, limiter of cells, centres of cells and ghosts, volumes,
        // face centre+normal+area, CSR; new U written.
        out[2] = su * (25 * N) + sg * (3 * (N + G) + N + 7 * F) + 4.0 * (N + 1 + 2 * slots) + su * (5 * N);
        out[3] = 8.0 * ntiles;
    }
    // Bytes of the arrays each kernel reads or writes in this library's layout, each element
    // once: a gradient record (kGradRec) and a flux record (kFluxRec) per face slot in place of
    // the gathered cell and face metrics, state records padded to kUR values, kGL gradient +
    // limiter values per cell, the tile records.  HBM traffic above these is data read twice.
    void layout_bytes(double out[4]) const {
        const double N = mesh->h.n_cells;
        const double slots = mesh->h.cell_faces_starts[mesh->h.n_cells];
        double g_bc = 0, g_state = 0;   // ghosts of regions with a BC (ghost records) / without (state records)
        for (uint32_t r = 0; r < mesh->h.n_regions; ++r)
            (bcs[r].set ? g_bc : g_state) += mesh->h.ghost_cell_starts[r + 1] - mesh->h.ghost_cell_starts[r];
        const double su = (double)cfg.dtype, ur = su * gr::kUR;
        const double gl = su * (cfg.dtype == GR_F64 ? gr::kGL<double> : gr::kGL<float>);
        const double tiles = 4.0 * gr::kTileRec * ntiles;
        out[0] = 0.0;
        // tile records, neighbour ids, cell and ghost states, gradient records, ghost records, LSQ
        // matrices (6 in the solver's precision); gradient + limiter written
        out[1] = tiles + 4.0 * slots + ur * (N + g_state) + su * gr::kGradRec * slots + su * gr::kGhostRec * g_bc +
                 su * 6 * N + gl * N;
        // tile records, neighbour ids, flux records, cell and ghost states, gradient + limiter,
        // volumes; new state records and the tile partials written
        out[2] = tiles + 4.0 * slots + su * gr::kFluxRec * slots + ur * (N + g_state) + gl * N + 8.0 * N + ur * N +
                 8.0 * ntiles;
        out[3] = 8.0 * ntiles;
    }
    hipGraphExec_t graph(int b) {
        if (exec[b]) return exec[b];
        hipGraph_t g = nullptr;
        GR_HIP(hipStreamBeginCapture(mesh->stream, hipStreamCaptureModeRelaxed));
        try {
            for (int k = 0; k < graph_iters; ++k) enqueue((b + k) % 2, ctl.p);
        } catch (...) {
            (void)hipStreamEndCapture(mesh->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        GR_HIP(hipStreamEndCapture(mesh->stream, &g));
        hipError_t e = hipGraphInstantiate(&exec[b], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        GR_HIP(e);
        return exec[b];
    }
};

namespace {

// Writes the state into both Jacobi buffers: ghosts of regions without a boundary condition
// are read from memory by whichever buffer is current.
template <typename S>
void upload_state(gr_solver *s, const double *rho, const double *rhou, const double *rhoe) {
    const size_t n = (size_t)s->mesh->h.n_cells + s->mesh->h.n_ghosts;
    std::vector<S> rec(gr::kUR * s->stride, S(0));
    for (size_t i = 0; i < n; ++i) {
        S *r = &rec[gr::kUR * i];
        r[0] = (S)rho[i];
        for (int k = 0; k < 3; ++k) r[1 + k] = (S)rhou[3 * i + k];
        r[4] = (S)rhoe[i];
    }
    for (int b = 0; b < 2; ++b)
        GR_HIP(hipMemcpyAsync(s->ubuf[b].p, rec.data(), rec.size() * sizeof(S), hipMemcpyHostToDevice,
                              s->mesh->stream));
    GR_HIP(hipStreamSynchronize(s->mesh->stream));
}

// Interior cells from the current buffer; ghosts as the reference leaves them after a step:
// the boundary-condition values computed from the state the last iteration read.
template <typename S>
void download_state(gr_solver *s, double *rho, double *rhou, double *rhoe) {
    const gr_mesh *m = s->mesh;
    const size_t N = m->h.n_cells, n = N + m->h.n_ghosts;
    if (s->ghost_src != s->cur) s->materialise_ghosts(s->ghost_src);
    std::vector<S> a(gr::kUR * s->stride), b(gr::kUR * s->stride);
    GR_HIP(hipMemcpyAsync(a.data(), s->ubuf[s->cur].p, a.size() * sizeof(S), hipMemcpyDeviceToHost, m->stream));
    GR_HIP(hipMemcpyAsync(b.data(), s->ubuf[s->ghost_src].p, b.size() * sizeof(S), hipMemcpyDeviceToHost, m->stream));
    GR_HIP(hipStreamSynchronize(m->stream));
    for (size_t i = 0; i < n; ++i) {
        const S *r = &(i < N ? a : b)[gr::kUR * i];
        if (rho) rho[i] = (double)r[0];
        if (rhou)
            for (int k = 0; k < 3; ++k) rhou[3 * i + k] = (double)r[1 + k];
        if (rhoe) rhoe[i] = (double)r[4];
    }
}

}  // namespace

extern "C" {

int gr_solver_create(gr_mesh *m, const gr_solver_config *cfg, gr_solver **out) {
    return guarded([&] {
        require(m && cfg && out, "null argument");
        check_ready(m);
        require(cfg->dtype == GR_F64 || cfg->dtype == GR_F32, "solver dtype must be GR_F64 or GR_F32");
        require(cfg->cfl > 0, "cfl must be positive");
        require(cfg->limiter >= 0 && cfg->limiter <= 2, "unknown limiter");
        require(m->h.n_regions <= (uint32_t)gr::kMaxRegions, "too many boundary regions");
        require(!m->h.one_dimensional, "the Euler solver needs a 3-D mesh");
        auto *s = new gr_solver();
        try {
            s->mesh = m;
            s->cfg = *cfg;
            s->stride = gr::pad64((size_t)m->h.n_cells + m->h.n_ghosts);
            const size_t blk = s->stride * s->es();
            for (auto &b : s->ubuf) {
                b.alloc(gr::kUR * blk);
                GR_HIP(hipMemsetAsync(b.p, 0, b.n, m->stream));
            }
            s->grad.alloc((size_t)(cfg->dtype == GR_F64 ? gr::kGL<double> : gr::kGL<float>) * blk);
            GR_HIP(hipMemsetAsync(s->grad.p, 0, s->grad.n, m->stream));
            s->bcs.resize(m->h.n_regions);
            s->build_tiles(gr::kTileCells);
            if (cfg->dtype == GR_F64) s->build_slot_geometry<double>();
            else s->build_slot_geometry<float>();
            s->ctl.alloc(1);
            s->hist_cap = 1 << 17;
            s->hist.alloc(s->hist_cap);
            GR_HIP(hipHostMalloc((void **)&s->ctl_host, 3 * sizeof(gr::SolverCtl)));
            for (auto &e : s->ev) GR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            GR_HIP(hipStreamSynchronize(m->stream));
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

void gr_solver_free(gr_solver *s) { delete s; }

int gr_solver_set_bc(gr_solver *s, const char *region, int kind, const double mach[3], double pressure,
                     double temperature) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        require(kind == GR_BC_SLIP || kind == GR_BC_FARFIELD, "unknown boundary kind");
        if (kind == GR_BC_FARFIELD) require(mach != nullptr && pressure > 0 && temperature > 0, "farfield needs mach, p > 0, T > 0");
        const uint32_t r = region_of(s->mesh, region);
        auto &b = s->bcs[r];
        b.set = true;
        b.kind = kind;
        for (int k = 0; k < 3; ++k) b.mach[k] = mach ? mach[k] : 0.0;
        b.p = pressure;
        b.T = temperature;
        s->drop_graphs();   // the table is baked into captured launches
    });
}

int gr_solver_init_freestream(gr_solver *s, const double mach[3], double pressure, double temperature) {
    return guarded([&] {
        require(s && mach, "null argument");
        double fs[5];
        freestream(mach, pressure, temperature, fs);
        const size_t n = s->stride;
        const dim3 grid(gr::blocks_for(n * gr::kUR));
        for (int b = 0; b < 2; ++b) {
            if (s->cfg.dtype == GR_F64)
                hipLaunchKernelGGL(gr::k_fill_states<double>, grid, dim3(gr::kBlock), 0, s->mesh->stream,
                                   s->rec<double>(b), n, fs[0], fs[1], fs[2], fs[3], fs[4]);
            else
                hipLaunchKernelGGL(gr::k_fill_states<float>, grid, dim3(gr::kBlock), 0, s->mesh->stream,
                                   s->rec<float>(b), n, (float)fs[0], (float)fs[1], (float)fs[2], (float)fs[3],
                                   (float)fs[4]);
        }
        s->ghost_src = s->cur;
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(s->mesh->stream));
    });
}

int gr_solver_set_state(gr_solver *s, const double *rho, const double *rhou, const double *rhoe) {
    return guarded([&] {
        require(s && rho && rhou && rhoe, "null argument");
        if (s->cfg.dtype == GR_F64) upload_state<double>(s, rho, rhou, rhoe);
        else upload_state<float>(s, rho, rhou, rhoe);
        s->ghost_src = s->cur;
    });
}

int gr_solver_get_state(const gr_solver *s, double *rho, double *rhou, double *rhoe) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        // materialising the ghost states writes only ghost entries that the iteration never
        // reads, so the logical state is unchanged
        gr_solver *ms = const_cast<gr_solver *>(s);
        if (s->cfg.dtype == GR_F64) download_state<double>(ms, rho, rhou, rhoe);
        else download_state<float>(ms, rho, rhou, rhoe);
    });
}

int gr_solver_iterate(gr_solver *s, int max_iter, double tol, double reference, gr_solve_result *out) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        require(max_iter > 0, "max_iter must be positive");
        gr_mesh *m = s->mesh;
        gr::SolverCtl &c0 = s->ctl_host[0];
        std::memset(&c0, 0, sizeof c0);
        c0.ref = reference;
        c0.tol = tol;
        c0.max_iter = max_iter;
        c0.hist_cap = s->hist_cap;
        if (s->cfg.use_graph) {
            // both buffer parities captured before the first launch, so that no capture or
            // instantiation ever falls between the launches of a run
            s->graph(0);
            s->graph(1);
        }
        GR_HIP(hipMemcpyAsync(s->ctl.p, &c0, sizeof c0, hipMemcpyHostToDevice, m->stream));
        const int K = s->graph_iters;
        const int launches = (max_iter + K - 1) / K;
        for (int L = 0; L < launches; ++L) {
            if (s->cfg.use_graph) {
                GR_HIP(hipGraphLaunch(s->graph(s->cur), m->stream));
            } else {
                for (int k = 0; k < K; ++k) s->enqueue((s->cur + k) % 2, s->ctl.p);
                GR_HIP(hipGetLastError());
            }
            gr::SolverCtl *poll = &s->ctl_host[1 + (L & 1)];
            GR_HIP(hipMemcpyAsync(poll, s->ctl.p, sizeof *poll, hipMemcpyDeviceToHost, m->stream));
            GR_HIP(hipEventRecord(s->ev[L & 1], m->stream));
            if (L >= 1) {
                GR_HIP(hipEventSynchronize(s->ev[(L - 1) & 1]));
                if (s->ctl_host[1 + ((L - 1) & 1)].stop) break;
            }
        }
        s->finish_pending(s->ctl.p);
        GR_HIP(hipGetLastError());
        GR_HIP(hipStreamSynchronize(m->stream));
        gr::SolverCtl fin;
        GR_HIP(hipMemcpy(&fin, s->ctl.p, sizeof fin, hipMemcpyDeviceToHost));
        if (fin.iter > 0) {
            s->ghost_src = (s->cur + fin.iter - 1) % 2;
            s->cur = (s->cur + fin.iter) % 2;
        }
        s->last_iters = std::min(fin.iter, s->hist_cap);
        if (out) {
            out->iterations = fin.iter;
            out->converged = fin.converged;
            out->residual = fin.last;
            out->ratio = fin.ratio;
            out->reference = fin.ref;
        }
        if (fin.diverged) throw gr::Error(GR_ERR_DIVERGED, "residual is not finite (diverged)");
    });
}

int gr_solver_history(const gr_solver *s, double *hist, int n) {
    return guarded([&] {
        require(s && hist, "null argument");
        require(n >= 0 && n <= s->last_iters, "history request longer than the last run");
        if (n) GR_HIP(hipMemcpy(hist, s->hist.p, n * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int gr_solver_residual_norm(gr_solver *s, double *out) {
    return guarded([&] {
        require(s && out, "null argument");
        s->enqueue(s->cur, nullptr);   // gradient+limiter, flux partials; state not swapped
        GR_HIP(hipGetLastError());
        *out = finalize_sum(s->mesh, s->ntiles, s->partials.p);
        s->ghost_src = s->cur;
    });
}

int gr_solver_profile(gr_solver *s, int n_iter, double ms[4]) {
    return guarded([&] {
        require(s && ms, "null argument");
        require(n_iter > 0 && n_iter <= 100000, "n_iter must be in 1..100000");
        // Each kernel n_iter times back to back on the current state (its launches read the same
        // inputs and write the same outputs), with events around each run only: per-launch
        // events would add the cost of an event pair (about 5 us) to every duration.
        ms[0] = 0.0;   // boundary conditions are evaluated inside the two kernels
        s->time_kernels(n_iter, ms + 1);
    });
}

int gr_solver_autotune(gr_solver *s, int n_iter, uint32_t *tile_cells) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        require(n_iter > 0 && n_iter <= 10000, "n_iter must be in 1..10000");
        // Candidate tile sizes (cells per block, at most kTileCells, each within kTileSlots
        // slots): smaller tiles mean more, shorter blocks, which balances a small mesh's few
        // rounds of resident blocks over the CUs; larger ones amortise each block's set-up.
        const uint32_t cand[] = {(uint32_t)gr::kTileCells, 44, 40, 36, 32, 28, 24};
        uint32_t best = (uint32_t)gr::kTileCells;
        double best_ms = 1e300;
        for (uint32_t c : cand) {
            s->build_tiles(c);
            double ms[3];
            s->time_kernels(n_iter, ms);
            const double tot = ms[0] + ms[1] + ms[2];
            if (tot < best_ms) {
                best_ms = tot;
                best = c;
            }
        }
        s->build_tiles(best);
        if (tile_cells) *tile_cells = best;
    });
}

int gr_solver_force(gr_solver *s, const char *region, double F[3]) {
    return guarded([&] {
        require(s && region && F, "null argument");
        const uint32_t r = region_of(s->mesh, region);
        int rc = gr_internal_wall_force(s->mesh, s->ubuf[s->cur].p, s->cfg.dtype, 1, gr::kUR, 0, r, F);
        if (rc != GR_OK) throw gr::Error(rc, gr_last_error());
    });
}

int gr_solver_kernel_bytes(const gr_solver *s, double bytes[4]) {
    return guarded([&] {
        require(s && bytes, "null argument");
        s->kernel_bytes(bytes);
    });
}

int gr_solver_layout_bytes(const gr_solver *s, double bytes[4]) {
    return guarded([&] {
        require(s && bytes, "null argument");
        s->layout_bytes(bytes);
    });
}

}  // extern "C"

// graphite.hpp -- the reference's host call surface (Alexsimulation/graphite include/mesh.cuh,
// include/meshio.cuh, include/fvmfields.cuh, include/fvmops.cuh, include/models/euler.cuh)
// as inline C++ over the C ABI of libgraphite_hip.so.  Code written against those headers
// compiles against this one with `scalar` = double.
//
// Deliberate differences from the reference:
//   * errors throw graphite_error (the reference prints and calls exit());
//   * the explicit update is Jacobi (the reference updates in place while neighbours read);
//   * FvmFieldReduceSum reduces on the device;
//   * the reference's __global__ operators (ScalarVarComputeGradient<<<...>>>(mesh, field),
//     ...) are host functions with the same names taking (cuMesh*, FvmField*).
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../graphite_hip.h"

typedef double scalar;
struct vector {
    double x, y, z;
};
inline vector make_vector(double x, double y, double z) { return vector{x, y, z}; }

struct graphite_error : std::runtime_error {
    int code;
    graphite_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

struct cuMesh;

namespace graphite_detail {
inline void check(int rc) {
    if (rc != GR_OK) throw graphite_error(rc, gr_last_error());
}
struct Registry {
    std::vector<cuMesh *> meshes;
};
inline Registry &registry() {
    static Registry r;
    return r;
}
}  // namespace graphite_detail

// cuMesh (include/mesh.cuh:55-63): the handle plus host copies of the CSR arrays.
struct cuMesh {
    gr_mesh *h = nullptr;
    gr_mesh_sizes sizes{};
    std::vector<std::string> region_names;
    void refresh() {
        graphite_detail::check(gr_mesh_sizes_get(h, &sizes));
        region_names.clear();
        for (uint32_t r = 0; r < sizes.n_regions; ++r) {
            char name[128];
            graphite_detail::check(gr_mesh_region_name(h, r, name));
            region_names.emplace_back(name);
        }
    }
};

inline cuMesh *cuMeshNew() {
    auto *m = new cuMesh();
    graphite_detail::registry().meshes.push_back(m);
    return m;
}
inline void cuMeshFree(cuMesh *m) {
    auto &v = graphite_detail::registry().meshes;
    for (size_t i = 0; i < v.size(); ++i)
        if (v[i] == m) {
            v.erase(v.begin() + (long)i);
            break;
        }
    if (m && m->h) gr_mesh_free(m->h);
    delete m;
}
// cuMeshFromFile (src/meshio.cu:166-292): read + metric kernels.  scalar_bytes is the width of
// the node coordinates in the file (the reference's default build writes 4).
inline cuMesh *cuMeshFromFile(const char *filename, int scalar_bytes = 4) {
    cuMesh *m = cuMeshNew();
    try {
        graphite_detail::check(gr_mesh_read_fvm(filename, scalar_bytes, &m->h));
        graphite_detail::check(gr_mesh_compute_metrics(m->h));
        m->refresh();
    } catch (...) {
        cuMeshFree(m);
        throw;
    }
    return m;
}
// cuMeshFromSu2 (src/meshio.cu:817-1250): returns 0, or 1 on an inconsistent mesh.
inline int cuMeshFromSu2(const char *su2_filename, const char *out_filename, int scalar_bytes = 4) {
    return gr_mesh_from_su2(su2_filename, out_filename, scalar_bytes) == GR_OK ? 0 : 1;
}
// cuMeshMake1D (src/mesh.cu:540-634)
inline void cuMeshMake1D(cuMesh *mesh, size_t n, scalar x0, scalar x1) {
    graphite_detail::check(gr_mesh_make_1d(n, x0, x1, &mesh->h));
    graphite_detail::check(gr_mesh_compute_metrics(mesh->h));
    mesh->refresh();
}
inline size_t cuMeshGetVarSize(cuMesh *m) { return (size_t)m->sizes.n_cells + m->sizes.n_ghosts; }
inline size_t cuMeshGetnCells(cuMesh *m) { return m->sizes.n_cells; }
inline size_t cuMeshGetnGhosts(cuMesh *m) { return m->sizes.n_ghosts; }
inline size_t cuMeshGetnCellsAndGhosts(cuMesh *m) { return (size_t)m->sizes.n_cells + m->sizes.n_ghosts; }
inline size_t cuMeshGetnRegionCells(cuMesh *m, const unsigned &region_id) {
    uint32_t n = 0;
    graphite_detail::check(gr_mesh_region_size(m->h, region_id, &n));
    return n;
}
inline unsigned cuMeshGetRegionId(cuMesh *m, const char *name) {
    uint32_t id = 0;
    graphite_detail::check(gr_mesh_region_id(m->h, name, &id));   // throws for an unknown name
    return id;
}

// FvmField (include/fvmfields.cuh:16-30) with host mirrors filled by VarGpuToCpu.
enum FvmFieldType { SCALAR, VECTOR, TENSOR, FIELD_UNKNOWN };
struct FvmFieldBase {
    size_t elem_size = 0, size = 0;
    void *value = nullptr, *gradient = nullptr, *limiter = nullptr;
};
struct FvmField {
    gr_field *h = nullptr;        // created when the field is bound to a mesh
    cuMesh *mesh = nullptr;
    FvmFieldBase *_cpu = nullptr;
    std::vector<double> v, g, l;
    std::string name;
    unsigned ncomp = 1;
    size_t size = 0;
    int alloc_grads = 0;
};

namespace graphite_detail {
// The reference's fields are bare arrays sized cells+ghosts; the mesh comes with each call.  A
// field here is bound to the mesh of its first call that names one (and must be used with that
// mesh afterwards).  Calls without a mesh (VarGpuToCpu, VarCpuToGpu) on an unbound field bind it
// to the one registered mesh of its size, and refuse when several meshes have that size.
inline gr_field *bind(FvmField *f, cuMesh *m) {
    if (!f) throw graphite_error(GR_ERR_INVALID, "null field");
    if (f->mesh) {
        if (m && m != f->mesh)
            throw graphite_error(GR_ERR_INVALID, "field '" + f->name + "' is bound to another mesh");
        return f->h;
    }
    if (!m) {
        int count = 0;
        for (cuMesh *c : registry().meshes)
            if (c->h && (size_t)c->sizes.n_cells + c->sizes.n_ghosts == f->size) {
                m = c;
                ++count;
            }
        if (count == 0)
            throw graphite_error(GR_ERR_INVALID, "field '" + f->name + "': no mesh with this many cells and ghosts");
        if (count > 1)
            throw graphite_error(GR_ERR_INVALID, "field '" + f->name + "': " + std::to_string(count) +
                                                     " meshes have this many cells and ghosts; pass the field to a "
                                                     "call that names its mesh first");
    }
    if (!m->h) throw graphite_error(GR_ERR_INVALID, "mesh not loaded");
    if ((size_t)m->sizes.n_cells + m->sizes.n_ghosts != f->size)
        throw graphite_error(GR_ERR_INVALID, "field '" + f->name + "' is not sized for this mesh");
    check(gr_field_create(m->h, f->ncomp, GR_F64, f->name.c_str(), f->alloc_grads, &f->h));
    f->mesh = m;
    return f->h;
}
}  // namespace graphite_detail

// FvmFieldNew (src/fvmfields.cu:7-46): a field of `size` elements (cells + ghosts of the mesh it
// will be used with), zero-initialised on the device when it is bound.
inline FvmField *FvmFieldNew(size_t elem_size, size_t size, const char *name, int alloc_grads) {
    const unsigned nc = elem_size == sizeof(vector) ? 3 : 1;
    auto *f = new FvmField();
    f->ncomp = nc;
    f->name = name ? name : "";
    f->size = size;
    f->alloc_grads = alloc_grads;
    f->v.assign(size * nc, 0.0);
    if (alloc_grads) {
        f->g.assign(size * nc * 3, 0.0);
        f->l.assign(size * nc, 0.0);
    }
    f->_cpu = new FvmFieldBase{elem_size, size, f->v.data(), alloc_grads ? f->g.data() : nullptr,
                               alloc_grads ? f->l.data() : nullptr};
    return f;
}
inline void FvmFieldFree(FvmField *f) {
    if (!f) return;
    if (f->h) gr_field_free(f->h);
    delete f->_cpu;
    delete f;
}
inline FvmFieldType FvmFieldGetType(const FvmField *f) { return f->ncomp == 3 ? VECTOR : SCALAR; }
inline void *FvmFieldValueIndex(FvmField *f, const size_t &i) { return f->v.data() + f->ncomp * i; }
inline void VarGpuToCpu(FvmField *f) {
    graphite_detail::check(gr_field_download(graphite_detail::bind(f, nullptr), f->v.data(), nullptr, nullptr));
}
inline void VarAllGpuToCpu(FvmField *f) {
    graphite_detail::check(gr_field_download(graphite_detail::bind(f, nullptr), f->v.data(),
                                             f->g.empty() ? nullptr : f->g.data(), f->l.empty() ? nullptr : f->l.data()));
}
// Host -> device of the host mirror (the reference writes _hgpu arrays with cudaMemcpy).
inline void VarCpuToGpu(FvmField *f) {
    graphite_detail::check(gr_field_upload(graphite_detail::bind(f, nullptr), f->v.data()));
}

// fvmops (src/fvmops.cu) as host calls; the field is bound to (or must belong to) the mesh given
namespace graphite_detail {
inline gr_field *on(cuMesh *m, FvmField *f) {
    if (!m) throw graphite_error(GR_ERR_INVALID, "null mesh");
    return bind(f, m);
}
}  // namespace graphite_detail
inline void ScalarVarComputeGradient(cuMesh *m, FvmField *f) { graphite_detail::check(gr_field_compute_gradient(graphite_detail::on(m, f))); }
inline void VectorVarComputeGradient(cuMesh *m, FvmField *f) { graphite_detail::check(gr_field_compute_gradient(graphite_detail::on(m, f))); }
inline void ScalarVarComputeLimiters(cuMesh *m, FvmField *f) { graphite_detail::check(gr_field_compute_limiter(graphite_detail::on(m, f), GR_LIM_MICHALAK)); }
inline void VectorVarComputeLimiters(cuMesh *m, FvmField *f) { graphite_detail::check(gr_field_compute_limiter(graphite_detail::on(m, f), GR_LIM_MICHALAK)); }
inline void fvmScalGradLimBc(cuMesh *m, FvmField *f) { graphite_detail::check(gr_field_gradlim_bc(graphite_detail::on(m, f))); }
inline void fvmVecGradLimBc(cuMesh *m, FvmField *f) { graphite_detail::check(gr_field_gradlim_bc(graphite_detail::on(m, f))); }
inline void ScalarVarInit(cuMesh *m, FvmField *f, scalar value) { graphite_detail::check(gr_field_init(graphite_detail::on(m, f), &value)); }
inline void VectorVarInit(cuMesh *m, FvmField *f, vector value) {
    double v[3] = {value.x, value.y, value.z};
    graphite_detail::check(gr_field_init(graphite_detail::on(m, f), v));
}
inline void fvmScalarBcFixed(cuMesh *m, FvmField *f, scalar bc_value, unsigned region_id) {
    graphite_detail::check(gr_field_bc_fixed(graphite_detail::on(m, f), region_id, &bc_value));
}
inline void fvmScalarBcZeroGrad(cuMesh *m, FvmField *f, unsigned region_id) {
    graphite_detail::check(gr_field_bc_zero_grad(graphite_detail::on(m, f), region_id));
}
inline void fvmVectorBcZeroGrad(cuMesh *m, FvmField *f, unsigned region_id) {
    graphite_detail::check(gr_field_bc_zero_grad(graphite_detail::on(m, f), region_id));
}
inline scalar FvmFieldReduceSum(cuMesh *m, FvmField *f) {
    double s = 0;
    graphite_detail::check(gr_field_reduce_sum(graphite_detail::on(m, f), &s));
    return s;
}

// Euler model (include/models/euler.cuh:45-94); all return 0 as the reference does.
inline int FvmEulerInit(cuMesh *m, FvmField *rho, FvmField *rhou, FvmField *rhoe, const vector &mach,
                        const scalar &pressure, const scalar &temperature) {
    double mv[3] = {mach.x, mach.y, mach.z};
    graphite_detail::check(gr_euler_init(m->h, graphite_detail::on(m, rho), graphite_detail::on(m, rhou),
                                         graphite_detail::on(m, rhoe), mv, pressure, temperature));
    return 0;
}
inline int FvmEulerSlipBc(cuMesh *m, FvmField *rho, FvmField *rhou, FvmField *rhoe, const char *region) {
    graphite_detail::check(gr_euler_bc_slip(m->h, graphite_detail::on(m, rho), graphite_detail::on(m, rhou),
                                            graphite_detail::on(m, rhoe), region));
    return 0;
}
inline int FvmEulerInletOutletBc(cuMesh *m, FvmField *rho, FvmField *rhou, FvmField *rhoe, const vector &mach,
                                 const scalar &pressure, const scalar &temperature, const char *region) {
    double mv[3] = {mach.x, mach.y, mach.z};
    graphite_detail::check(gr_euler_bc_farfield(m->h, graphite_detail::on(m, rho), graphite_detail::on(m, rhou),
                                                graphite_detail::on(m, rhoe), mv, pressure, temperature, region));
    return 0;
}
inline int FvmEulerStepPrepare(cuMesh *m, FvmField *rho, FvmField *rhou, FvmField *rhoe) {
    graphite_detail::check(gr_euler_step_prepare(m->h, graphite_detail::on(m, rho), graphite_detail::on(m, rhou),
                                                 graphite_detail::on(m, rhoe), GR_LIM_MICHALAK));
    return 0;
}
inline int FvmEulerStep(cuMesh *m, FvmField *rho, FvmField *rhou, FvmField *rhoe, const scalar &cfl,
                        const char *time_scheme, FvmField *dt, FvmField *error) {
    graphite_detail::check(gr_euler_step(m->h, graphite_detail::on(m, rho), graphite_detail::on(m, rhou),
                                         graphite_detail::on(m, rhoe), cfl, time_scheme,
                                         dt ? graphite_detail::on(m, dt) : nullptr,
                                         error ? graphite_detail::on(m, error) : nullptr));
    return 0;
}

// Mesh handle, binary64 metric kernels and the mesh part of the C ABI.
//
// The metric kernels evaluate the reference's formulas (src/mesh.cu:44-219) in the same
// order of operations, with correctly rounded binary64 sqrt/div and no contraction, so their
// results are reproducible on the host bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <utility>

#include "device.hpp"

namespace gr {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }

// ---- metric kernels ---------------------------------------------------------------------

__device__ inline double sgn(double x) { return (double)(x > 0) - (double)(x < 0); }

// Fan triangle (centre c, nodes a, b) of a face, src/mesh.cu:67-93.
__device__ inline void fan_tri(const double c[3], const double a[3], const double b[3], double &area,
                               double nrm[3], double nc[3]) {
    double d0[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
    double d1[3] = {c[0] - b[0], c[1] - b[1], c[2] - b[2]};
    double cr[3] = {d0[1] * d1[2] - d0[2] * d1[1], d0[2] * d1[0] - d0[0] * d1[2], d0[0] * d1[1] - d0[1] * d1[0]};
    double len = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    area += len * 0.5;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double tc = ((a[k] + b[k]) + c[k]) / 3.0;
        nrm[k] += cr[k];
        nc[k] += (tc * len) * 0.5;
    }
}

// cuMeshComputeFaces: face centre (area-weighted fan centroid), unit normal, fan area.
__global__ void k_face_metrics(uint32_t F, const uint32_t *__restrict__ fns, const uint32_t *__restrict__ fn,
                               const double *__restrict__ nodes, double *__restrict__ geo, size_t fs) {
    uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const uint32_t s = fns[f], e = fns[f + 1];
    double c[3] = {0, 0, 0};
    for (uint32_t q = s; q < e; ++q) {
        const double *p = nodes + 3 * (size_t)fn[q];
        c[0] += p[0];
        c[1] += p[1];
        c[2] += p[2];
    }
    const double sz = (double)(float)(e - s);
    c[0] /= sz;
    c[1] /= sz;
    c[2] /= sz;
    double area = 0, nrm[3] = {0, 0, 0}, nc[3] = {0, 0, 0};
    for (uint32_t q = s; q + 1 < e; ++q) {
        double a[3], b[3];
        const double *pa = nodes + 3 * (size_t)fn[q], *pb = nodes + 3 * (size_t)fn[q + 1];
        for (int k = 0; k < 3; ++k) { a[k] = pa[k]; b[k] = pb[k]; }
        fan_tri(c, a, b, area, nrm, nc);
    }
    {
        double a[3], b[3];
        const double *pa = nodes + 3 * (size_t)fn[e - 1], *pb = nodes + 3 * (size_t)fn[s];
        for (int k = 0; k < 3; ++k) { a[k] = pa[k]; b[k] = pb[k]; }
        fan_tri(c, a, b, area, nrm, nc);
    }
    const double nn = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    geo[0 * fs + f] = nc[0] / area;
    geo[1 * fs + f] = nc[1] / area;
    geo[2 * fs + f] = nc[2] / area;
    geo[3 * fs + f] = nrm[0] / nn;
    geo[4 * fs + f] = nrm[1] / nn;
    geo[5 * fs + f] = nrm[2] / nn;
    geo[6 * fs + f] = area;
}

// cuMeshComputeCells: two-pass approximate centroid and divergence-theorem volume.
__global__ void k_cell_metrics(uint32_t N, const uint32_t *__restrict__ cfs, const uint32_t *__restrict__ cf,
                               const double *__restrict__ fgeo, size_t fs, double *__restrict__ cgeo, size_t cs) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const uint32_t s = cfs[i], e = cfs[i + 1];
    const double *fcx = fgeo, *fcy = fgeo + fs, *fcz = fgeo + 2 * fs;
    const double *fnx = fgeo + 3 * fs, *fny = fgeo + 4 * fs, *fnz = fgeo + 5 * fs, *fa = fgeo + 6 * fs;
    double c[3] = {0, 0, 0}, w = 0;
    for (uint32_t q = s; q < e; ++q) {
        uint32_t f = cf[q];
        c[0] += fcx[f] * fa[f];
        c[1] += fcy[f] * fa[f];
        c[2] += fcz[f] * fa[f];
        w += fa[f];
    }
    c[0] /= w;
    c[1] /= w;
    c[2] /= w;
    double nc[3] = {0, 0, 0};
    w = 0;
    for (uint32_t q = s; q < e; ++q) {
        uint32_t f = cf[q];
        double d[3] = {fcx[f] - c[0], fcy[f] - c[1], fcz[f] - c[2]};
        double wf = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * fa[f];
        for (int k = 0; k < 3; ++k) nc[k] += (c[k] + d[k] * 0.75) * wf;
        w += wf;
    }
    nc[0] /= w;
    nc[1] /= w;
    nc[2] /= w;
    double vol = 0;
    for (uint32_t q = s; q < e; ++q) {
        uint32_t f = cf[q];
        double d[3] = {fcx[f] - nc[0], fcy[f] - nc[1], fcz[f] - nc[2]};
        double sg = sgn(fnx[f] * d[0] + fny[f] * d[1] + fnz[f] * d[2]);
        double n[3] = {fnx[f] * sg, fny[f] * sg, fnz[f] * sg};
        vol += (d[0] * n[0] + d[1] * n[1] + d[2] * n[2]) * fa[f];
    }
    cgeo[0 * cs + i] = nc[0];
    cgeo[1 * cs + i] = nc[1];
    cgeo[2 * cs + i] = nc[2];
    cgeo[3 * cs + i] = vol / 3.0;
}

// cuMeshComputeGhostCells: ghost centre mirrored through the face along its normal.
__global__ void k_ghost_metrics(uint32_t N, uint32_t G, const uint32_t *__restrict__ cfs,
                                const uint32_t *__restrict__ cf, const uint32_t *__restrict__ cc,
                                const double *__restrict__ fgeo, size_t fs, double *__restrict__ cgeo, size_t cs) {
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= G) return;
    const uint32_t i = N + g;
    const uint32_t s = cfs[i];
    const uint32_t o = cc[s], f = cf[s];
    const double fc[3] = {fgeo[f], fgeo[fs + f], fgeo[2 * fs + f]};
    const double fnv[3] = {fgeo[3 * fs + f], fgeo[4 * fs + f], fgeo[5 * fs + f]};
    const double a = fgeo[6 * fs + f];
    double d[3] = {fc[0] - cgeo[o], fc[1] - cgeo[cs + o], fc[2] - cgeo[2 * cs + o]};
    double dx = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double sg = sgn(fnv[0] * d[0] + fnv[1] * d[1] + fnv[2] * d[2]);
    for (int k = 0; k < 3; ++k) cgeo[k * cs + i] = fc[k] + (fnv[k] * sg) * dx;
    cgeo[3 * cs + i] = (a * 2) * dx;
}

// cuMeshComputeCellMatrices: inverse of sum_j d d^T / |d|^2 via adjugate and determinant.
// The inverse of the symmetric matrix is symmetric term by term, so 6 entries are stored.
__global__ void k_lsq_matrices(uint32_t N, const uint32_t *__restrict__ cfs, const uint32_t *__restrict__ cc,
                               const double *__restrict__ cgeo, size_t cs, double *__restrict__ mat, size_t ms) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};   // A[col][row]
    const double xi[3] = {cgeo[i], cgeo[cs + i], cgeo[2 * cs + i]};
    for (uint32_t q = cfs[i]; q < cfs[i + 1]; ++q) {
        uint32_t j = cc[q];
        double d[3] = {cgeo[j] - xi[0], cgeo[cs + j] - xi[1], cgeo[2 * cs + j] - xi[2]};
        double w = 1.0 / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
        for (int col = 0; col < 3; ++col)
#pragma unroll
            for (int row = 0; row < 3; ++row) A[col][row] += ((d[row] * d[col]) * w) * w;
    }
    const double m00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const double m01 = A[2][1] * A[0][2] - A[2][2] * A[0][1];   // column 0, row 1
    const double m02 = A[0][1] * A[1][2] - A[0][2] * A[1][1];   // column 0, row 2
    const double m11 = A[0][0] * A[2][2] - A[0][2] * A[2][0];
    const double m12 = A[1][0] * A[0][2] - A[1][2] * A[0][0];   // column 1, row 2
    const double m22 = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    const double det = (A[0][0] * m00 + A[1][0] * m01) + A[2][0] * m02;
    mat[0 * ms + i] = m00 / det;
    mat[1 * ms + i] = m11 / det;
    mat[2 * ms + i] = m22 / det;
    mat[3 * ms + i] = m01 / det;
    mat[4 * ms + i] = m02 / det;
    mat[5 * ms + i] = m12 / det;
}

// Mesh validity after the metric kernels (counters: [0] interior cells with a non-positive or
// non-finite volume, [1] face slots of interior cells whose face does not separate the two
// cell centres (n.(x_f - x_i) and n.(x_f - x_j) not of strictly opposite signs: a face with
// n.d = 0, where the reference's binary64 sign() returns 0 and drops the face from the flux,
// or a tangled cell), [2] interior cells with a non-finite least-squares matrix, [3] first bad
// cell; [4] faces with a non-positive or non-finite area or a non-finite normal, [5] first bad
// face).
__global__ void k_mesh_check_cells(uint32_t N, const uint32_t *__restrict__ cfs, const uint32_t *__restrict__ cf,
                                   const uint32_t *__restrict__ cc, const double *__restrict__ fgeo, size_t fs,
                                   const double *__restrict__ cgeo, size_t cs, const double *__restrict__ mat,
                                   size_t ms, int check_mat, uint32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    bool bad = false;
    const double v = cgeo[3 * cs + i];
    if (!(v > 0.0) || !isfinite(v)) {
        atomicAdd(&out[0], 1u);
        bad = true;
    }
    const double xi[3] = {cgeo[i], cgeo[cs + i], cgeo[2 * cs + i]};
    for (uint32_t q = cfs[i]; q < cfs[i + 1]; ++q) {
        const uint32_t f = cf[q], j = cc[q];
        const double fc[3] = {fgeo[f], fgeo[fs + f], fgeo[2 * fs + f]};
        const double n[3] = {fgeo[3 * fs + f], fgeo[4 * fs + f], fgeo[5 * fs + f]};
        const double si = (n[0] * (fc[0] - xi[0]) + n[1] * (fc[1] - xi[1])) + n[2] * (fc[2] - xi[2]);
        const double sj = (n[0] * (fc[0] - cgeo[j]) + n[1] * (fc[1] - cgeo[cs + j])) + n[2] * (fc[2] - cgeo[2 * cs + j]);
        if (!((si > 0.0 && sj < 0.0) || (si < 0.0 && sj > 0.0))) {
            atomicAdd(&out[1], 1u);
            bad = true;
        }
    }
    if (check_mat) {
        bool fin = true;
        for (int k = 0; k < 6; ++k) fin = fin && isfinite(mat[k * ms + i]);
        if (!fin) {
            atomicAdd(&out[2], 1u);
            bad = true;
        }
    }
    if (bad) atomicMin(&out[3], i);
}

__global__ void k_mesh_check_faces(uint32_t F, const double *__restrict__ fgeo, size_t fs, uint32_t *__restrict__ out) {
    const uint32_t f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const double a = fgeo[6 * fs + f];
    const bool fin = isfinite(fgeo[3 * fs + f]) && isfinite(fgeo[4 * fs + f]) && isfinite(fgeo[5 * fs + f]);
    if (!(a > 0.0) || !isfinite(a) || !fin) {
        atomicAdd(&out[4], 1u);
        atomicMin(&out[5], f);
    }
}

}  // namespace gr

// ---- gr_mesh ------------------------------------------------------------------------------

gr_mesh::~gr_mesh() {
    if (host_scalar) (void)hipHostFree(host_scalar);
    if (stream) (void)hipStreamDestroy(stream);
}

void gr_mesh::ensure_device() {
    if (on_device) return;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        throw gr::Error(GR_ERR_DEVICE, "no HIP device available");
    GR_HIP(hipGetDevice(&device));
    if (!stream) GR_HIP(hipStreamCreate(&stream));
    if (!host_scalar) GR_HIP(hipHostMalloc((void **)&host_scalar, sizeof(double)));
    const size_t NG = (size_t)h.n_cells + h.n_ghosts;
    fns.alloc(h.face_nodes_starts.size());
    fn.alloc(std::max<size_t>(1, h.face_nodes.size()));
    cfs.alloc(h.cell_faces_starts.size());
    cf.alloc(std::max<size_t>(1, h.cell_faces.size()));
    cc.alloc(std::max<size_t>(1, h.cell_connects.size()));
    nodes.alloc(std::max<size_t>(1, h.nodes.size()));
    face_geo.alloc(7 * fstride());
    cell_geo.alloc(4 * cstride());
    mat.alloc(6 * std::max<size_t>(64, mstride()));
    partials.alloc(gr::blocks_for(std::max<size_t>(NG, h.n_faces)) + 64);
    scalar_out.alloc(1);
    fns.upload(h.face_nodes_starts.data(), h.face_nodes_starts.size(), stream);
    fn.upload(h.face_nodes.data(), h.face_nodes.size(), stream);
    cfs.upload(h.cell_faces_starts.data(), h.cell_faces_starts.size(), stream);
    cf.upload(h.cell_faces.data(), h.cell_faces.size(), stream);
    cc.upload(h.cell_connects.data(), h.cell_connects.size(), stream);
    nodes.upload(h.nodes.data(), h.nodes.size(), stream);
    GR_HIP(hipStreamSynchronize(stream));
    on_device = true;
}

namespace gr {
const HostMesh &host_mesh(const gr_mesh *m) { return m->h; }
}  // namespace gr

gr::MeshView gr_mesh::view() const {
    gr::MeshView v;
    v.N = h.n_cells;
    v.G = h.n_ghosts;
    v.F = h.n_faces;
    v.cfs = cfs.p;
    v.cf = cf.p;
    v.cc = cc.p;
    const size_t fs = fstride(), cs = cstride();
    v.fcx = face_geo.p;
    v.fcy = face_geo.p + fs;
    v.fcz = face_geo.p + 2 * fs;
    v.fnx = face_geo.p + 3 * fs;
    v.fny = face_geo.p + 4 * fs;
    v.fnz = face_geo.p + 5 * fs;
    v.fa = face_geo.p + 6 * fs;
    v.cx = cell_geo.p;
    v.cy = cell_geo.p + cs;
    v.cz = cell_geo.p + 2 * cs;
    v.cv = cell_geo.p + 3 * cs;
    v.mat = mat.p;
    v.mat_stride = std::max<size_t>(64, mstride());
    return v;
}

using gr::guarded;
using gr::require;

static gr_mesh *wrap(gr::HostMesh &&h) {
    auto *m = new gr_mesh();
    m->h = std::move(h);
    return m;
}

extern "C" {

const char *gr_last_error(void) { return gr::g_last_error.c_str(); }
int gr_abi_version(void) { return GR_ABI_VERSION; }
int gr_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int gr_mesh_build(const gr_elements *e, gr_mesh **out) {
    return guarded([&] {
        require(e && out, "null argument");
        *out = wrap(gr::build_mesh(*e));
    });
}

int gr_mesh_from_su2(const char *su2_path, const char *fvm_path, int scalar_bytes) {
    return guarded([&] {
        require(su2_path && fvm_path, "null argument");
        gr_elements e;
        gr::read_su2(su2_path, e);
        gr::HostMesh m = gr::build_mesh(e);
        gr::write_fvm(m, fvm_path, scalar_bytes);
    });
}

int gr_mesh_read_fvm(const char *path, int scalar_bytes, gr_mesh **out) {
    return guarded([&] {
        require(path && out, "null argument");
        gr::HostMesh h;
        gr::read_fvm(path, scalar_bytes, h);
        *out = wrap(std::move(h));
    });
}

int gr_mesh_write_fvm(const gr_mesh *m, const char *path, int scalar_bytes) {
    return guarded([&] {
        require(m && path, "null argument");
        gr::write_fvm(m->h, path, scalar_bytes);
    });
}

int gr_mesh_make_1d(size_t n, double x0, double x1, gr_mesh **out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        *out = wrap(gr::make_1d(n, x0, x1));
    });
}

void gr_mesh_free(gr_mesh *m) { delete m; }

int gr_mesh_sizes_get(const gr_mesh *m, gr_mesh_sizes *out) {
    return guarded([&] {
        require(m && out, "null argument");
        out->n_nodes = m->h.n_nodes;
        out->n_faces = m->h.n_faces;
        out->n_cells = m->h.n_cells;
        out->n_ghosts = m->h.n_ghosts;
        out->n_regions = m->h.n_regions;
        out->n_face_nodes = m->h.face_nodes_starts.back();
        out->n_cell_slots = m->h.cell_faces_starts.back();
    });
}

int gr_mesh_get_connectivity(const gr_mesh *m, double *nodes, uint32_t *face_nodes_starts,
                             uint32_t *face_nodes, uint32_t *cell_faces_starts, uint32_t *cell_faces,
                             uint32_t *cell_connects, uint32_t *ghost_cell_ids, uint32_t *ghost_cell_starts) {
    return guarded([&] {
        require(m != nullptr, "null argument");
        const auto &h = m->h;
        auto cp = [](const auto &v, auto *dst) {
            if (dst) std::copy(v.begin(), v.end(), dst);
        };
        cp(h.nodes, nodes);
        cp(h.face_nodes_starts, face_nodes_starts);
        cp(h.face_nodes, face_nodes);
        cp(h.cell_faces_starts, cell_faces_starts);
        cp(h.cell_faces, cell_faces);
        cp(h.cell_connects, cell_connects);
        cp(h.ghost_cell_ids, ghost_cell_ids);
        cp(h.ghost_cell_starts, ghost_cell_starts);
    });
}

int gr_mesh_region_id(const gr_mesh *m, const char *name, uint32_t *id) {
    return guarded([&] {
        require(m && name && id, "null argument");
        for (uint32_t r = 0; r < m->h.n_regions; ++r)
            if (m->h.region_names[r] == name) {
                *id = r;
                return;
            }
        throw gr::Error(GR_ERR_NOT_FOUND, std::string("region '") + name + "' not found in mesh");
    });
}

int gr_mesh_region_name(const gr_mesh *m, uint32_t id, char name[128]) {
    return guarded([&] {
        require(m && name, "null argument");
        require(id < m->h.n_regions, "region index out of range");
        std::memset(name, 0, 128);
        std::strncpy(name, m->h.region_names[id].c_str(), 127);
    });
}

int gr_mesh_region_size(const gr_mesh *m, uint32_t id, uint32_t *n) {
    return guarded([&] {
        require(m && n, "null argument");
        require(id < m->h.n_regions, "region index out of range");
        *n = m->h.ghost_cell_starts[id + 1] - m->h.ghost_cell_starts[id];
    });
}

int gr_mesh_renumber_rcm(gr_mesh *m, uint32_t *new2old_cells, uint32_t *new2old_faces) {
    return guarded([&] {
        require(m != nullptr, "null argument");
        require(!m->on_device, "renumbering must precede gr_mesh_compute_metrics");
        require(!m->h.one_dimensional, "renumbering of the 1-D channel mesh is not supported");
        std::vector<uint32_t> c2o, f2o;
        gr::renumber_rcm(m->h, c2o, f2o);
        m->h.validate();
        if (new2old_cells) std::copy(c2o.begin(), c2o.end(), new2old_cells);
        if (new2old_faces) std::copy(f2o.begin(), f2o.end(), new2old_faces);
    });
}

int gr_mesh_renumber_hilbert(gr_mesh *m, uint32_t *new2old_cells, uint32_t *new2old_faces) {
    return guarded([&] {
        require(m != nullptr, "null argument");
        require(!m->on_device, "renumbering must precede gr_mesh_compute_metrics");
        require(!m->h.one_dimensional, "renumbering of the 1-D channel mesh is not supported");
        std::vector<uint32_t> c2o, f2o;
        gr::renumber_hilbert(m->h, c2o, f2o);
        m->h.validate();
        if (new2old_cells) std::copy(c2o.begin(), c2o.end(), new2old_cells);
        if (new2old_faces) std::copy(f2o.begin(), f2o.end(), new2old_faces);
    });
}

int gr_mesh_face_colors(const gr_mesh *m, uint32_t *colors, uint32_t *n_colors) {
    return guarded([&] {
        require(m && n_colors, "null argument");
        auto c = gr::face_colors(m->h, *n_colors);
        if (colors) std::copy(c.begin(), c.end(), colors);
    });
}

int gr_mesh_bandwidth(const gr_mesh *m, uint32_t *bw) {
    return guarded([&] {
        require(m && bw, "null argument");
        *bw = gr::bandwidth(m->h);
    });
}

int gr_mesh_compute_metrics(gr_mesh *m) {
    return guarded([&] {
        require(m != nullptr, "null argument");
        m->ensure_device();
        const auto &h = m->h;
        const size_t fs = m->fstride(), cs = m->cstride();
        if (h.n_faces)
            hipLaunchKernelGGL(gr::k_face_metrics, dim3(gr::blocks_for(h.n_faces)), dim3(gr::kBlock), 0, m->stream,
                               h.n_faces, m->fns.p, m->fn.p, m->nodes.p, m->face_geo.p, fs);
        if (h.n_cells)
            hipLaunchKernelGGL(gr::k_cell_metrics, dim3(gr::blocks_for(h.n_cells)), dim3(gr::kBlock), 0, m->stream,
                               h.n_cells, m->cfs.p, m->cf.p, m->face_geo.p, fs, m->cell_geo.p, cs);
        if (h.n_ghosts)
            hipLaunchKernelGGL(gr::k_ghost_metrics, dim3(gr::blocks_for(h.n_ghosts)), dim3(gr::kBlock), 0,
                               m->stream, h.n_cells, h.n_ghosts, m->cfs.p, m->cf.p, m->cc.p, m->face_geo.p, fs,
                               m->cell_geo.p, cs);
        const size_t ms = std::max<size_t>(64, m->mstride());
        if (h.n_cells && !h.one_dimensional)
            hipLaunchKernelGGL(gr::k_lsq_matrices, dim3(gr::blocks_for(h.n_cells)), dim3(gr::kBlock), 0, m->stream,
                               h.n_cells, m->cfs.p, m->cc.p, m->cell_geo.p, cs, m->mat.p, ms);
        else
            GR_HIP(hipMemsetAsync(m->mat.p, 0, m->mat.n * sizeof(double), m->stream));
        // reject meshes the scheme cannot run on (the reference runs them silently)
        gr::DevArray<uint32_t> chk;
        chk.alloc(6);
        const uint32_t init[6] = {0u, 0u, 0u, 0xFFFFFFFFu, 0u, 0xFFFFFFFFu};
        chk.upload(init, 6, m->stream);
        if (h.n_cells)
            hipLaunchKernelGGL(gr::k_mesh_check_cells, dim3(gr::blocks_for(h.n_cells)), dim3(gr::kBlock), 0, m->stream,
                               h.n_cells, m->cfs.p, m->cf.p, m->cc.p, m->face_geo.p, fs, m->cell_geo.p, cs, m->mat.p, ms,
                               h.one_dimensional ? 0 : 1, chk.p);
        if (h.n_faces)
            hipLaunchKernelGGL(gr::k_mesh_check_faces, dim3(gr::blocks_for(h.n_faces)), dim3(gr::kBlock), 0, m->stream,
                               h.n_faces, m->face_geo.p, fs, chk.p);
        GR_HIP(hipGetLastError());
        uint32_t c[6];
        chk.download(c, 6, m->stream);
        GR_HIP(hipStreamSynchronize(m->stream));
        m->metrics_ready = false;
        if (c[4])
            throw gr::Error(GR_ERR_TOPOLOGY, std::to_string(c[4]) + " face(s) with a non-positive or non-finite area "
                                             "(first: face " + std::to_string(c[5]) + ")");
        if (c[0] || c[1] || c[2])
            throw gr::Error(GR_ERR_TOPOLOGY,
                            "invalid mesh: " + std::to_string(c[0]) + " cell(s) with a non-positive volume, " +
                                std::to_string(c[1]) + " face slot(s) whose face does not separate the two cell "
                                "centres (n.d = 0 or a tangled cell), " + std::to_string(c[2]) +
                                " singular least-squares matrix(es); first: cell " + std::to_string(c[3]));
        m->metrics_ready = true;
    });
}

int gr_mesh_get_view(const gr_mesh *m, gr_mesh_view *out) {
    return guarded([&] {
        require(m && out, "null argument");
        require(m->metrics_ready, "metrics not computed (gr_mesh_compute_metrics)");
        const gr::MeshView v = m->view();
        gr_mesh_view o{};
        o.n_cells = v.N;
        o.n_ghosts = v.G;
        o.n_faces = v.F;
        o.cell_faces_starts = v.cfs;
        o.cell_faces = v.cf;
        o.cell_connects = v.cc;
        o.face_center_x = v.fcx;
        o.face_center_y = v.fcy;
        o.face_center_z = v.fcz;
        o.face_normal_x = v.fnx;
        o.face_normal_y = v.fny;
        o.face_normal_z = v.fnz;
        o.face_area = v.fa;
        o.cell_center_x = v.cx;
        o.cell_center_y = v.cy;
        o.cell_center_z = v.cz;
        o.cell_volume = v.cv;
        o.lsq_matrix = v.mat;
        o.lsq_stride = v.mat_stride;
        *out = o;
    });
}

int gr_mesh_get_metrics(const gr_mesh *m, double *face_centers, double *face_normals, double *face_areas,
                        double *cell_centers, double *cell_volumes, double *cell_matrices) {
    return guarded([&] {
        require(m != nullptr, "null argument");
        require(m->metrics_ready, "metrics not computed (gr_mesh_compute_metrics)");
        const auto &h = m->h;
        const size_t F = h.n_faces, NG = (size_t)h.n_cells + h.n_ghosts, N = h.n_cells;
        std::vector<double> fg(m->face_geo.n), cg(m->cell_geo.n), mt(m->mat.n);
        m->face_geo.download(fg.data(), fg.size(), m->stream);
        m->cell_geo.download(cg.data(), cg.size(), m->stream);
        m->mat.download(mt.data(), mt.size(), m->stream);
        GR_HIP(hipStreamSynchronize(m->stream));
        const size_t fs = m->fstride(), cs = m->cstride(), ms = std::max<size_t>(64, m->mstride());
        for (size_t f = 0; f < F; ++f)
            for (int k = 0; k < 3; ++k) {
                if (face_centers) face_centers[3 * f + k] = fg[k * fs + f];
                if (face_normals) face_normals[3 * f + k] = fg[(3 + k) * fs + f];
            }
        if (face_areas)
            for (size_t f = 0; f < F; ++f) face_areas[f] = fg[6 * fs + f];
        for (size_t c = 0; c < NG; ++c) {
            if (cell_centers)
                for (int k = 0; k < 3; ++k) cell_centers[3 * c + k] = cg[k * cs + c];
            if (cell_volumes) cell_volumes[c] = cg[3 * cs + c];
        }
        if (cell_matrices)
            for (size_t c = 0; c < N; ++c) {
                // tensor columns (u,v,w) each xyz; symmetric
                const double xx = mt[c], yy = mt[ms + c], zz = mt[2 * ms + c];
                const double xy = mt[3 * ms + c], xz = mt[4 * ms + c], yz = mt[5 * ms + c];
                const double t[9] = {xx, xy, xz, xy, yy, yz, xz, yz, zz};
                std::copy(t, t + 9, cell_matrices + 9 * c);
            }
    });
}

}  // extern "C"

// ---- bandwidth probe -------------------------------------------------------------------------

namespace gr {
// 16-byte vector copy, grid-stride.  Measured on the MI355X over 1 GiB: one element per thread
// per pass with 65,536 blocks of 256 threads reaches 5.46 TB/s; four loads in flight per thread
// with 16,384 blocks 4.54, non-temporal loads and stores 5.3, a contiguous chunk per block 5.4.
__global__ __launch_bounds__(kBlock) void k_copy_probe(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                       size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}
}  // namespace gr

extern "C" int gr_stream_copy_probe(size_t bytes, int reps, double *gbps) {
    return gr::guarded([&] {
        gr::require(gbps != nullptr, "null argument");
        gr::require(bytes >= 16 && bytes % 16 == 0 && bytes <= (size_t(64) << 30), "bytes must be a multiple of 16, <= 64 GiB");
        gr::require(reps > 0 && reps <= 1000, "reps must be in 1..1000");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw gr::Error(GR_ERR_DEVICE, "no HIP device available");
        gr::DevArray<double2> a, b;
        const size_t n = bytes / 16;
        a.alloc(n);
        b.alloc(n);
        hipStream_t s;
        GR_HIP(hipStreamCreate(&s));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        struct Cleanup {
            hipStream_t &s;
            hipEvent_t &e0, &e1;
            ~Cleanup() {
                if (e0) (void)hipEventDestroy(e0);
                if (e1) (void)hipEventDestroy(e1);
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
        } cleanup{s, e0, e1};
        GR_HIP(hipMemsetAsync(a.p, 0, bytes, s));
        GR_HIP(hipEventCreate(&e0));
        GR_HIP(hipEventCreate(&e1));
        const unsigned grid = (unsigned)std::min<size_t>((n + gr::kBlock - 1) / gr::kBlock, 65536);
        hipLaunchKernelGGL(gr::k_copy_probe, dim3(grid), dim3(gr::kBlock), 0, s, a.p, b.p, n);   // warm-up
        GR_HIP(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(gr::k_copy_probe, dim3(grid), dim3(gr::kBlock), 0, s, a.p, b.p, n);
        GR_HIP(hipEventRecord(e1, s));
        GR_HIP(hipEventSynchronize(e1));
        GR_HIP(hipGetLastError());
        float ms = 0;
        GR_HIP(hipEventElapsedTime(&ms, e0, e1));
        *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    });
}

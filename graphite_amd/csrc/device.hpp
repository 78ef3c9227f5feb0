// Device-side types of libgraphite_hip.so: HIP error handling, device arrays, the mesh and
// field handles and the plain-pointer views passed to kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <string>

#include "gr_internal.hpp"

#define GR_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t gr_e_ = (call);                                                           \
        if (gr_e_ != hipSuccess)                                                             \
            throw gr::Error(GR_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(gr_e_)); \
    } while (0)

namespace gr {

constexpr int kBlock = 256;   // threads per block of the per-cell kernels (4 wave64)

inline unsigned blocks_for(size_t n, unsigned b = kBlock) { return (unsigned)((n + b - 1) / b); }

// Bijection workgroup id -> work item (0..T-1) such that XCD x (= w mod 8: the MI355X hands
// workgroups to its 8 XCDs in turn) gets a contiguous range of items, whose shared inputs then
// stay in that XCD's own L2.
__device__ inline uint32_t xcd_tile(uint32_t w, uint32_t T) {
    const uint32_t x = w & 7u, k = w >> 3;
    return x * (T >> 3) + min(x, T & 7u) + k;
}

// Owning device allocation.
template <typename T>
struct DevArray {
    T *p = nullptr;
    size_t n = 0;
    DevArray() = default;
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    ~DevArray() { release(); }
    void alloc(size_t count) {
        release();
        if (count) GR_HIP(hipMalloc((void **)&p, count * sizeof(T)));
        n = count;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void upload(const T *h, size_t count, hipStream_t s) {
        require(count <= n, "device upload larger than allocation");
        if (count) GR_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *h, size_t count, hipStream_t s) const {
        require(count <= n, "device download larger than allocation");
        if (count) GR_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
    }
};

inline size_t pad64(size_t n) { return (n + 63) & ~size_t(63); }

// Mesh arrays as kernels see them.  Geometry is binary64, structure-of-arrays.
struct MeshView {
    uint32_t N, G, F;
    const uint32_t *cfs;          // cell_faces_starts, N+G+1
    const uint32_t *cf;           // cell_faces
    const uint32_t *cc;           // cell_connects
    const double *fcx, *fcy, *fcz;  // face centres
    const double *fnx, *fny, *fnz;  // unit face normals (orientation as computed)
    const double *fa;               // face areas
    const double *cx, *cy, *cz;     // cell centres (N+G)
    const double *cv;               // cell volumes (N+G)
    const double *mat;              // LSQ matrices, 6 SoA blocks of N: xx,yy,zz,xy,xz,yz
    size_t mat_stride;              // block stride of mat
};

constexpr int kMaxRegions = 32;

}  // namespace gr

struct gr_mesh {
    gr::HostMesh h;
    hipStream_t stream = nullptr;
    bool on_device = false;
    bool metrics_ready = false;
    int corrected = 0;   // momentum reconstruction of gr_euler_step
    int device = -1;     // HIP device current when the mesh went to the device
    gr::DevArray<uint32_t> fns, fn, cfs, cf, cc;
    gr::DevArray<double> nodes;
    gr::DevArray<double> face_geo;   // 7 blocks of pad64(F): cx,cy,cz,nx,ny,nz,a
    gr::DevArray<double> cell_geo;   // 4 blocks of pad64(N+G): cx,cy,cz,v
    gr::DevArray<double> mat;        // 6 blocks of pad64(N)
    gr::DevArray<double> partials;   // reduction scratch, one per block of the largest launch
    gr::DevArray<double> scalar_out; // one double
    double *host_scalar = nullptr;   // pinned, one double
    ~gr_mesh();
    void ensure_device();            // stream + connectivity upload
    gr::MeshView view() const;
    size_t fstride() const { return gr::pad64(h.n_faces); }
    size_t cstride() const { return gr::pad64((size_t)h.n_cells + h.n_ghosts); }
    size_t mstride() const { return gr::pad64(h.n_cells); }
};

// Pressure force on wall region `region` from a state array whose element (c, cell, lane) is
// at c*stride + cell*lanes + lane (post.hip).  Returns a gr_status.
int gr_internal_wall_force(gr_mesh *m, const void *u, int dtype, size_t stride, uint32_t lanes, uint32_t lane,
                           uint32_t region, double F[3]);

// Device-side evaluation of the warm-start network (mlp.hip): rows are processed in chunks on
// stream s; fill(X, row0, m) writes the bfloat16 inputs of rows row0..row0+m-1 into X
// ([m][kMlpInPad]: the 8-column groups that hold inputs, padding within them zero; the columns
// after those groups are zero on entry and must stay zero), consume(Y, row0, m) reads the
// binary32 outputs ([m][kMlpOutPad]).
struct gr_mlp;
namespace gr {
constexpr int kMlpInPad = 64;
constexpr int kMlpOutPad = 32;
uint32_t mlp_n_in(const gr_mlp *n);
uint32_t mlp_n_out(const gr_mlp *n);
void mlp_run_device(gr_mlp *n, size_t M, hipStream_t s, const std::function<void(void *, size_t, int)> &fill,
                    const std::function<void(const float *, size_t, int)> &consume);
}  // namespace gr

struct gr_field {
    gr_mesh *mesh = nullptr;
    uint32_t ncomp = 1;
    int dtype = GR_F64;
    std::string name;
    bool grads = false;
    size_t n = 0;        // cells + ghosts
    size_t stride = 0;   // SoA block stride (elements)
    gr::DevArray<char> value, next, grad, lim;
    size_t esize() const { return (size_t)dtype; }
    template <typename S>
    S *val(int c) const { return (S *)value.p + (size_t)c * stride; }
    template <typename S>
    S *nxt(int c) const { return (S *)next.p + (size_t)c * stride; }
    template <typename S>
    S *grd(int c, int k) const { return (S *)grad.p + (size_t)(3 * c + k) * stride; }
    template <typename S>
    S *lmt(int c) const { return (S *)lim.p + (size_t)c * stride; }
    void ensure_next();
    void swap_next() { std::swap(value.p, next.p); }
};

// Post-processing on the device: pressure force on wall regions (lift/drag) and distance to
// the nearest wall face (a feature of the learned warm start).  Neither exists in the
// reference; both serve the batched sweeps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "fvm_tiles.hpp"

namespace gr {

// Force on the body from ghosts [g0, g0+count): sum over wall faces of p * n_out * A, with p
// the pressure of the wall cell and n_out the face normal pointing out of that cell (into the
// body).  State element (c, i, lane) at c*stride + i*lanes + lane.
template <typename S>
__global__ __launch_bounds__(kBlock) void k_wall_force(MeshView m, const S *u, size_t stride, uint32_t lanes,
                                                       uint32_t lane, uint32_t g0, uint32_t count, double *partials) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    double f[3] = {0, 0, 0};
    if (t < count) {
        const uint32_t c0 = m.N + g0 + t, q = m.cfs[c0];
        const uint32_t i = m.cc[q], fc = m.cf[q];
        const size_t o = (size_t)i * lanes + lane;
        const double r = u[o], ru[3] = {u[stride + o], u[2 * stride + o], u[3 * stride + o]}, e = u[4 * stride + o];
        const double p = 0.4 * (e - (0.5 / r) * ((ru[0] * ru[0] + ru[1] * ru[1]) + ru[2] * ru[2]));
        double n[3];
        outward_normal<double>(m, fc, m.cx[i], m.cy[i], m.cz[i], n);
        for (int k = 0; k < 3; ++k) f[k] = p * n[k] * m.fa[fc];
    }
    for (int k = 0; k < 3; ++k) {
        const double bs = block_sum(f[k]);
        if (threadIdx.x == 0) partials[3 * blockIdx.x + k] = bs;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBlock) void k_force_final(const double *partials, uint32_t n, double *out) {
    for (int k = 0; k < 3; ++k) {
        double acc = 0;
        for (uint32_t b = threadIdx.x; b < n; b += blockDim.x) acc += partials[3 * b + k];
        const double tot = block_sum(acc);
        if (threadIdx.x == 0) out[k] = tot;
        __syncthreads();
    }
}

// Distance from each interior cell centre to the nearest of W wall-face centres, wall
// centres streamed through LDS in chunks.
constexpr int kWallChunk = 1024;
__global__ __launch_bounds__(kBlock) void k_wall_distance(MeshView m, const double *__restrict__ wc, uint32_t W,
                                                          double *out) {
    __shared__ double s[3][kWallChunk];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const double x = i < m.N ? m.cx[i] : 0, y = i < m.N ? m.cy[i] : 0, z = i < m.N ? m.cz[i] : 0;
    double best = 1e300;
    for (uint32_t w0 = 0; w0 < W; w0 += kWallChunk) {
        const uint32_t nw = min((uint32_t)kWallChunk, W - w0);
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < nw; k += blockDim.x) {
            s[0][k] = wc[3 * (w0 + k)];
            s[1][k] = wc[3 * (w0 + k) + 1];
            s[2][k] = wc[3 * (w0 + k) + 2];
        }
        __syncthreads();
        for (uint32_t k = 0; k < nw; ++k) {
            const double dx = x - s[0][k], dy = y - s[1][k], dz = z - s[2][k];
            best = fmin(best, (dx * dx + dy * dy) + dz * dz);
        }
    }
    if (i < m.N) out[i] = sqrt(best);
}

}  // namespace gr

using gr::guarded;
using gr::require;

// Force helper shared by the single and batched solvers (declared in device.hpp users).
int gr_internal_wall_force(gr_mesh *m, const void *u, int dtype, size_t stride, uint32_t lanes, uint32_t lane,
                           uint32_t region, double F[3]) {
    return guarded([&] {
        require(region < m->h.n_regions, "region index out of range");
        const uint32_t g0 = m->h.ghost_cell_starts[region], cnt = m->h.ghost_cell_starts[region + 1] - g0;
        if (!cnt) {
            F[0] = F[1] = F[2] = 0;
            return;
        }
        const uint32_t nb = gr::blocks_for(cnt);
        gr::DevArray<double> part, out;
        part.alloc(3 * (size_t)nb);
        out.alloc(3);
        if (dtype == GR_F64)
            hipLaunchKernelGGL(gr::k_wall_force<double>, dim3(nb), dim3(gr::kBlock), 0, m->stream, m->view(),
                               (const double *)u, stride, lanes, lane, g0, cnt, part.p);
        else
            hipLaunchKernelGGL(gr::k_wall_force<float>, dim3(nb), dim3(gr::kBlock), 0, m->stream, m->view(),
                               (const float *)u, stride, lanes, lane, g0, cnt, part.p);
        hipLaunchKernelGGL(gr::k_force_final, dim3(1), dim3(gr::kBlock), 0, m->stream, part.p, nb, out.p);
        GR_HIP(hipGetLastError());
        out.download(F, 3, m->stream);
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

extern "C" int gr_mesh_wall_distance(gr_mesh *m, const char *const *regions, uint32_t n_regions, double *out) {
    return guarded([&] {
        require(m && regions && out, "null argument");
        require(m->metrics_ready, "mesh metrics not computed (gr_mesh_compute_metrics)");
        std::vector<uint32_t> faces;
        for (uint32_t k = 0; k < n_regions; ++k) {
            uint32_t r = 0;
            for (; r < m->h.n_regions; ++r)
                if (m->h.region_names[r] == regions[k]) break;
            if (r == m->h.n_regions)
                throw gr::Error(GR_ERR_NOT_FOUND, std::string("region '") + regions[k] + "' not found in mesh");
            for (uint32_t g = m->h.ghost_cell_starts[r]; g < m->h.ghost_cell_starts[r + 1]; ++g)
                faces.push_back(m->h.cell_faces[m->h.cell_faces_starts[m->h.n_cells + g]]);
        }
        require(!faces.empty(), "no wall faces in the given regions");
        // face centres from the device metrics
        const size_t fs = m->fstride();
        std::vector<double> fg(3 * fs);
        GR_HIP(hipMemcpyAsync(fg.data(), m->face_geo.p, fg.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream));
        GR_HIP(hipStreamSynchronize(m->stream));
        std::vector<double> wc(3 * faces.size());
        for (size_t k = 0; k < faces.size(); ++k)
            for (int d = 0; d < 3; ++d) wc[3 * k + d] = fg[d * fs + faces[k]];
        gr::DevArray<double> dwc, dout;
        dwc.alloc(wc.size());
        dout.alloc(std::max<size_t>(1, m->h.n_cells));
        dwc.upload(wc.data(), wc.size(), m->stream);
        hipLaunchKernelGGL(gr::k_wall_distance, dim3(gr::blocks_for(m->h.n_cells)), dim3(gr::kBlock), 0, m->stream,
                           m->view(), dwc.p, (uint32_t)faces.size(), dout.p);
        GR_HIP(hipGetLastError());
        dout.download(out, m->h.n_cells, m->stream);
        GR_HIP(hipStreamSynchronize(m->stream));
    });
}

// Internal types of libgraphite_hip.so (host side).  Not part of the C ABI.
#pragma once

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/graphite_hip.h"

namespace gr {

// Error raised inside the library and turned into a status code at the C boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

void set_last_error(const std::string &msg);
void clear_last_error();

template <typename F>
int guarded(F &&fn) {
    try {
        clear_last_error();
        fn();
        return GR_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host out of memory");
        return GR_ERR_INVALID;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return GR_ERR_INVALID;
    }
}

inline void require(bool ok, const std::string &msg, int code = GR_ERR_INVALID) {
    if (!ok) throw Error(code, msg);
}

// Local face tables of the su2 element kinds (src/meshio.cu:297-475).
struct ElemKind {
    int n_nodes;
    int n_faces;
    int8_t f[6][4];   // -1 padded
};
bool elem_kind(uint32_t kind, ElemKind &out);

struct Marker {
    std::string name;
    std::vector<uint32_t> kinds, starts, nodes;   // starts has count+1 entries
};

}  // namespace gr

struct gr_elements {
    std::vector<double> xyz;             // 3 per node
    std::vector<uint32_t> kinds, starts, nodes;
    std::vector<gr::Marker> markers;
    uint32_t n_nodes() const { return (uint32_t)(xyz.size() / 3); }
    uint32_t n_elems() const { return (uint32_t)kinds.size(); }
};

namespace gr {

// Reference CSR mesh on the host (cuMeshBase _cpu, include/mesh.cuh:18-52).
struct HostMesh {
    uint32_t n_nodes = 0, n_faces = 0, n_cells = 0, n_ghosts = 0, n_regions = 0;
    std::vector<double> nodes;   // 3 per node
    std::vector<uint32_t> face_nodes_starts, face_nodes;
    std::vector<uint32_t> cell_faces_starts, cell_faces, cell_connects;
    std::vector<uint32_t> ghost_cell_ids, ghost_cell_starts;
    std::vector<std::string> region_names;
    // cuMeshMake1D meshes: two faces per cell, no LSQ matrices (src/mesh.cu:625-627)
    bool one_dimensional = false;
    // Checks sizes, index ranges and the ghost layout the kernels rely on.
    void validate() const;
};

HostMesh build_mesh(const gr_elements &e);
HostMesh make_1d(size_t n, double x0, double x1);
void read_fvm(const char *path, int scalar_bytes, HostMesh &m);
void write_fvm(const HostMesh &m, const char *path, int scalar_bytes);
void renumber_rcm(HostMesh &m, std::vector<uint32_t> &cell_new2old,
                  std::vector<uint32_t> &face_new2old);
void renumber_hilbert(HostMesh &m, std::vector<uint32_t> &cell_new2old,
                      std::vector<uint32_t> &face_new2old);
std::vector<uint32_t> face_colors(const HostMesh &m, uint32_t &n_colors);
uint32_t bandwidth(const HostMesh &m);

void read_su2(const char *path, gr_elements &e);
void write_su2(const gr_elements &e, const char *path);
void gen_box(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz, int tets,
             double jitter, uint64_t seed, gr_elements &e);
void gen_ogrid(uint32_t nt, uint32_t nr, uint32_t nz, double r0, double r1, double ratio,
               double dz, gr_elements &e);
void gen_cgrid(uint32_t ni, uint32_t nj, uint32_t nwake, double thick, double radius, double h0, double dz,
               gr_elements &e);
void gen_wing(const gr_wing_params &p, gr_elements &e);
// Host connectivity of a mesh handle (mesh_device.hip).
const HostMesh &host_mesh(const gr_mesh *m);

}  // namespace gr

/*
 * graphite_hip.h -- C ABI of libgraphite_hip.so, the MI355X (gfx950) implementation of
 * GRAPHITE's second-order unstructured finite-volume residual path.
 *
 * Every entry point returns GR_OK (0) or an error code and never exits the process; the
 * message of the last error on the calling thread is gr_last_error().  Handles are opaque and
 * owned by the caller (free with the matching *_free).  Vectors are passed as double[3].
 * All device work is enqueued on the library's stream; entry points that return data to the
 * host synchronise that stream before returning.
 *
 * Citations are Alexsimulation/graphite file:line of the interface each entry point replaces.
 */
#ifndef GRAPHITE_HIP_H
#define GRAPHITE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GR_ABI_VERSION 1

enum gr_status {
    GR_OK = 0,
    GR_ERR_INVALID = 1,      /* bad argument or state */
    GR_ERR_TOPOLOGY = 2,     /* inconsistent mesh (face not shared by exactly two cells, ...) */
    GR_ERR_IO = 3,           /* file cannot be opened, read or written */
    GR_ERR_NOT_FOUND = 4,    /* unknown region name */
    GR_ERR_DEVICE = 5,       /* HIP runtime error (no device, out of memory, launch failure) */
    GR_ERR_DIVERGED = 6      /* non-finite residual */
};

/* Data types of fields; mesh metrics are always binary64. */
enum gr_dtype { GR_F32 = 4, GR_F64 = 8 };

/* su2 element kinds, src/meshio.cu:297-313 */
enum gr_elem_kind { GR_TRI = 5, GR_QUAD = 9, GR_TET = 10, GR_HEX = 12, GR_WEDGE = 13, GR_PYRAMID = 14 };

/* Limiter functions, src/fvmops.cu:118-127 (MICHALAK is the reference's active one). */
enum gr_limiter { GR_LIM_MICHALAK = 0, GR_LIM_BARTH_JESPERSEN = 1, GR_LIM_VENKATAKRISHNAN = 2 };

/* Boundary kinds of PhysicsEuler's "bcs" entries, src/models/euler.cu:103-135. */
enum gr_bc_kind { GR_BC_SLIP = 0, GR_BC_FARFIELD = 1 };

const char *gr_last_error(void);
int gr_abi_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int gr_device_count(void);

/* ---- element lists (su2 form) ------------------------------------------------------------ */
typedef struct gr_elements gr_elements;

typedef struct gr_elements_info {
    uint32_t n_nodes, n_elems, n_elem_nodes, n_markers;
} gr_elements_info;

/* Volume elements: kinds GR_TET/GR_HEX/GR_WEDGE/GR_PYRAMID, node lists concatenated with
 * starts[n_elems+1].  Nodes are xyz triples.  (su2ReadElements, src/meshio.cu:771-811) */
int gr_elements_create(uint32_t n_nodes, const double *xyz, uint32_t n_elems,
                       const uint32_t *kinds, const uint32_t *starts, const uint32_t *nodes,
                       gr_elements **out);
/* Boundary marker (MARKER_TAG) of GR_TRI/GR_QUAD elements, src/meshio.cu:877-896. */
int gr_elements_add_marker(gr_elements *e, const char *name, uint32_t n, const uint32_t *kinds,
                           const uint32_t *starts, const uint32_t *nodes);
/* su2 text reader (NDIME/NPOIN/NELEM/NMARK), src/meshio.cu:817-897, and writer. */
int gr_elements_read_su2(const char *path, gr_elements **out);
int gr_elements_write_su2(const gr_elements *e, const char *path);
int gr_elements_info_get(const gr_elements *e, gr_elements_info *out);
int gr_elements_get(const gr_elements *e, double *xyz, uint32_t *kinds, uint32_t *starts,
                    uint32_t *nodes);
/* Marker m: name (up to 127 chars + NUL into name[128]) and element count; arrays may be
 * NULL to query the sizes first (n_nodes receives the node-list length). */
int gr_elements_get_marker(const gr_elements *e, uint32_t m, char name[128], uint32_t *n,
                           uint32_t *n_nodes, uint32_t *kinds, uint32_t *starts, uint32_t *nodes);
void gr_elements_free(gr_elements *e);

/* Synthetic meshes (no reference counterpart; they stand in for the gmsh .su2 inputs).
 * Box [0,lx]x[0,ly]x[0,lz] of nx*ny*nz hexes, each split into 6 tets along the main diagonal
 * when tets != 0; interior nodes moved by U[-jitter*h, +jitter*h] per axis (seeded
 * mt19937_64); markers xmin,xmax,ymin,ymax,zmin,zmax. */
int gr_gen_box(uint32_t nx, uint32_t ny, uint32_t nz, double lx, double ly, double lz,
               int tets, double jitter, uint64_t seed, gr_elements **out);
/* O-grid hexes around a cylinder of radius r0 along z: n_theta x n_r x n_z cells, radial
 * spacing geometric with ratio from r0 to r1, span n_z*dz; markers wall, farfield, zmin,
 * zmax. */
int gr_gen_ogrid(uint32_t n_theta, uint32_t n_r, uint32_t n_z, double r0, double r1,
                 double ratio, double dz, gr_elements **out);
/* C-grid hexes around a symmetric NACA 4-digit airfoil (chord 1, leading edge at the origin,
 * closed trailing edge) extruded one cell of depth dz: n_i cells along the C (n_wake in each
 * wake part, the rest cosine-spaced around the airfoil), n_j cells outward with geometric
 * growth from first_height to an outer C of the given radius around the trailing edge; the
 * wake cut is merged.  Markers airfoil, farfield, zmin, zmax. */
int gr_gen_cgrid_naca(uint32_t n_i, uint32_t n_j, uint32_t n_wake, double thickness, double radius,
                      double first_height, double dz, gr_elements **out);
/* Swept, tapered wing with a symmetric NACA 4-digit section (root chord 1, root leading edge
 * at the origin, span along +z from the symmetry plane z = 0) in a half-sphere far field
 * centred on the root mid-chord.  n_prism wall layers of prisms (first_height, geometric
 * growth), then n_tet layers of tetrahedra out to the sphere; n_span stations along the wing
 * and n_cap stations above the tip, over which the section tapers to a zero-thickness plate
 * (upper and lower nodes merged; above it the cells are prisms).  Every length is multiplied
 * by scale at the end (the reference's limiter threshold is dimensional, so the scale changes
 * which cells it limits).  Markers wall, farfield, symmetry.  Defaults giving 1,069,112 cells:
 * 88, 20, 70, 20, 16, 0.12, 30, 0.56, 1.5, 3e-3, 1.2, 25, 1. */
typedef struct gr_wing_params {
    uint32_t n_around, n_prism, n_tet, n_span, n_cap;
    double thickness, sweep_deg, taper, semispan, first_height, growth, radius, scale;
} gr_wing_params;
int gr_gen_wing(const gr_wing_params *p, gr_elements **out);

/* ---- mesh ------------------------------------------------------------------------------ */
typedef struct gr_mesh gr_mesh;

typedef struct gr_mesh_sizes {
    uint32_t n_nodes, n_faces, n_cells, n_ghosts, n_regions;
    uint32_t n_face_nodes;   /* face_nodes_starts[n_faces] */
    uint32_t n_cell_slots;   /* cell_faces_starts[n_cells+n_ghosts] */
} gr_mesh_sizes;

/* Unique faces, cell/face/ghost CSR, in cuMeshFromSu2's numbering (src/meshio.cu:902-1138).
 * GR_ERR_TOPOLOGY when a face is not shared by exactly two cells or a marker face is not a
 * face of the volume mesh. */
int gr_mesh_build(const gr_elements *e, gr_mesh **out);
/* cuMeshFromSu2 (src/meshio.cu:817-1250): su2 file -> binary .fvm file. */
int gr_mesh_from_su2(const char *su2_path, const char *fvm_path, int scalar_bytes);
/* cuMeshFromFile (src/meshio.cu:166-292): read a .fvm file whose node coordinates are
 * scalar_bytes wide (4 as the reference's default FP32 build writes, or 8).  Metrics are
 * computed by gr_mesh_compute_metrics. */
int gr_mesh_read_fvm(const char *path, int scalar_bytes, gr_mesh **out);
int gr_mesh_write_fvm(const gr_mesh *m, const char *path, int scalar_bytes);
/* cuMeshMake1D (src/mesh.cu:540-634): n cells on [x0,x1] with unit cross-section. */
int gr_mesh_make_1d(size_t n, double x0, double x1, gr_mesh **out);
void gr_mesh_free(gr_mesh *m);   /* cuMeshFree, src/mesh.cu:401-411 */

/* cuMeshGetnCells / GetnGhosts / GetVarSize ..., src/mesh.cu:5-31 */
int gr_mesh_sizes_get(const gr_mesh *m, gr_mesh_sizes *out);
/* Host copies of the CSR arrays (any pointer may be NULL).  nodes: 3*n_nodes. */
int gr_mesh_get_connectivity(const gr_mesh *m, double *nodes, uint32_t *face_nodes_starts,
                             uint32_t *face_nodes, uint32_t *cell_faces_starts,
                             uint32_t *cell_faces, uint32_t *cell_connects,
                             uint32_t *ghost_cell_ids, uint32_t *ghost_cell_starts);
/* Region names (cuMeshGetRegionId src/mesh.cu:17-23): GR_ERR_NOT_FOUND for an unknown name
 * (the reference prints and falls back to region 0). */
int gr_mesh_region_id(const gr_mesh *m, const char *name, uint32_t *id);
int gr_mesh_region_name(const gr_mesh *m, uint32_t id, char name[128]);
/* cuMeshGetnRegionCells, src/mesh.cu:13-15 */
int gr_mesh_region_size(const gr_mesh *m, uint32_t id, uint32_t *n);

/* Reverse Cuthill-McKee renumbering of interior cells, ghosts sorted by their interior cell
 * within each region, faces ordered by their lower-numbered cell.  Each cell keeps its face
 * slot order, so every per-cell sum runs in the same order as before.  new2old_cells
 * (n_cells+n_ghosts) and new2old_faces (n_faces) may be NULL.  Must precede
 * gr_mesh_compute_metrics. */
int gr_mesh_renumber_rcm(gr_mesh *m, uint32_t *new2old_cells, uint32_t *new2old_faces);
/* As gr_mesh_renumber_rcm with the interior cells in Hilbert-curve order of their centres: a
 * contiguous range of cells (what one XCD's blocks have in flight) is a compact region with few
 * neighbours outside it.  No reference counterpart (the reference keeps the su2 order). */
int gr_mesh_renumber_hilbert(gr_mesh *m, uint32_t *new2old_cells, uint32_t *new2old_faces);
/* Greedy face colouring: no two faces of one colour share a cell (interior or ghost).
 * colors[n_faces] may be NULL; n_colors receives the count. */
int gr_mesh_face_colors(const gr_mesh *m, uint32_t *colors, uint32_t *n_colors);
/* Matrix bandwidth max |i-j| over interior neighbour pairs. */
int gr_mesh_bandwidth(const gr_mesh *m, uint32_t *bw);

/* cuMeshPassCpuToGpu + the four metric kernels cuMeshComputeFaces/Cells/GhostCells/
 * CellMatrices (src/meshio.cu:277-286, src/mesh.cu:44-219), binary64 on the device. */
int gr_mesh_compute_metrics(gr_mesh *m);
/* cuMeshPassGpuToCpu of the metric arrays (src/mesh.cu:467-517), reference AoS layouts:
 * face_centers/face_normals 3*n_faces, face_areas n_faces, cell_centers 3*(N+G),
 * cell_volumes N+G, cell_matrices 9*N (tensor columns u,v,w).  Pointers may be NULL. */
int gr_mesh_get_metrics(const gr_mesh *m, double *face_centers, double *face_normals,
                        double *face_areas, double *cell_centers, double *cell_volumes,
                        double *cell_matrices);
/* The mesh as a device kernel reads it (the reference's device struct cuMeshBase* that its
 * kernels take, include/mesh.cuh:18-52, kernelInput(x) = x->_gpu, include/datatypes.cuh:76):
 * device pointers to the CSR arrays and the binary64 metric blocks (centres, unit normals,
 * areas; cell centres and volumes over cells and ghosts; LSQ matrices as six blocks xx, yy, zz,
 * xy, xz, yz of stride lsq_stride).  Valid until gr_mesh_free, and until gr_mesh_renumber_rcm
 * or gr_mesh_compute_metrics run again.  include/graphite/fvm_ops.hpp builds on it. */
typedef struct gr_mesh_view {
    uint32_t n_cells, n_ghosts, n_faces;
    const uint32_t *cell_faces_starts, *cell_faces, *cell_connects;
    const double *face_center_x, *face_center_y, *face_center_z;
    const double *face_normal_x, *face_normal_y, *face_normal_z;
    const double *face_area;
    const double *cell_center_x, *cell_center_y, *cell_center_z;
    const double *cell_volume;
    const double *lsq_matrix;
    size_t lsq_stride;
} gr_mesh_view;
int gr_mesh_get_view(const gr_mesh *m, gr_mesh_view *out);

/* ---- fields (FvmField, include/fvmfields.cuh:16-43) ---------------------------------------- */
typedef struct gr_field gr_field;

/* FvmFieldNew (src/fvmfields.cu:7-46): n_comp 1 (scalar) or 3 (vector) over the mesh's
 * cells and ghosts, values of dtype GR_F64/GR_F32; gradients and limiters when alloc_grads. */
int gr_field_create(gr_mesh *m, uint32_t n_comp, int dtype, const char *name, int alloc_grads,
                    gr_field **out);
void gr_field_free(gr_field *f);   /* FvmFieldFree, src/fvmfields.cu:49-64 */
/* Host <-> device in the reference's AoS layout (value n*n_comp; gradient n*n_comp*3 with
 * vector gradients as tensor columns; limiter n*n_comp).  VarGpuToCpu / VarAllGpuToCpu
 * (src/fvmfields.cu:85-101). */
int gr_field_upload(gr_field *f, const double *value);
int gr_field_download(const gr_field *f, double *value, double *gradient, double *limiter);
/* Finite-volume operators of include/fvmops.cuh:96-219 (fvmLaplacianC, fvmScalGradC,
 * fvmScalGradCCorrect, fvmVecGradC, fvmBoundScalGradC, fvmBoundVecGradC, fvmDivC,
 * fvmConvConstScalU, fvmConvConstScalU2, fvmConvScalU, fvmConvVecU), each evaluated for every
 * interior cell of field u (ghost entries as they stand, e.g. after gr_field_gradlim_bc).
 * out: n_cells x (1 | 3 | 9) values (scalar, vector, or tensor as three columns), binary64.
 * phi_const: the constant velocity of the ConvConst operators; phi_field: the vector field of
 * ConvScalU / ConvVecU.  The operators reading gradients or limiters need alloc_grads. */
enum gr_fvm_op {
    GR_OP_LAPLACIAN_C = 0,
    GR_OP_SCAL_GRAD_C = 1,
    GR_OP_SCAL_GRAD_C_CORRECT = 2,
    GR_OP_VEC_GRAD_C = 3,
    GR_OP_BOUND_SCAL_GRAD_C = 4,
    GR_OP_BOUND_VEC_GRAD_C = 5,
    GR_OP_DIV_C = 6,
    GR_OP_CONV_CONST_SCAL_U = 7,
    GR_OP_CONV_CONST_SCAL_U2 = 8,
    GR_OP_CONV_SCAL_U = 9,
    GR_OP_CONV_VEC_U = 10
};
int gr_field_operator(const gr_field *u, int op, const double phi_const[3], const gr_field *phi_field, double *out);
/* A field as a device kernel reads it (FvmFieldBase's device arrays, include/fvmfields.cuh:16-30):
 * per component a block of n = cells + ghosts values, gradients (component c, direction k at
 * gradient[3c+k]) and limiters (NULL without alloc_grads), of dtype GR_F64 or GR_F32.  Valid
 * until the field is freed or gr_euler_step writes it (that call swaps its value buffer). */
typedef struct gr_field_view {
    int dtype;
    uint32_t n_comp;
    size_t n;
    void *value[3];
    void *gradient[9];
    void *limiter[3];
} gr_field_view;
int gr_field_get_view(const gr_field *f, gr_field_view *out);
/* ScalarVarInit / VectorVarInit (src/fvmops.cu:8-43): value = v, gradient 0, limiter 1. */
int gr_field_init(gr_field *f, const double *v);
/* FvmFieldReduceSum (src/fvmops.cu:369-383): sum over interior cells, device reduction in
 * binary64 (one 8-byte copy to the host). */
int gr_field_reduce_sum(const gr_field *f, double *out);

/* Scalar/VectorVarComputeGradient (src/fvmops.cu:51-111). */
int gr_field_compute_gradient(gr_field *f);
/* Scalar/VectorVarComputeLimiters (src/fvmops.cu:175-266). */
int gr_field_compute_limiter(gr_field *f, int limiter);
/* fvmScalGradLimBc / fvmVecGradLimBc (src/fvmops.cu:272-304). */
int gr_field_gradlim_bc(gr_field *f);
/* fvmScalarBcFixed (src/fvmops.cu:308-323); v has n_comp entries. */
int gr_field_bc_fixed(gr_field *f, uint32_t region, const double *v);
/* fvmScalarBcZeroGrad / fvmVectorBcZeroGrad (src/fvmops.cu:326-365). */
int gr_field_bc_zero_grad(gr_field *f, uint32_t region);

/* ---- Euler model (src/models/euler.cu, include/models/euler.cuh:45-94) ------------------- */

/* FvmEulerInit (src/models/euler.cu:529-549). */
int gr_euler_init(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe,
                  const double mach[3], double pressure, double temperature);
/* FvmEulerSlipBc (src/models/euler.cu:552-568). */
int gr_euler_bc_slip(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe,
                     const char *region);
/* FvmEulerInletOutletBc (src/models/euler.cu:571-593). */
int gr_euler_bc_farfield(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe,
                         const double mach[3], double pressure, double temperature,
                         const char *region);
/* FvmEulerStepPrepare (src/models/euler.cu:597-647): gradients, limiters and their ghost
 * copies of rho, rhou, rhoe (one fused kernel plus the ghost copy). */
int gr_euler_step_prepare(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe,
                          int limiter);
/* FvmEulerStep (src/models/euler.cu:649-673) with time_scheme "steady": flux residual and
 * local-time-step update.  Jacobi: the update reads the old state of every cell (the
 * reference updates in place while neighbours read).  dt (scalar field, may be NULL)
 * receives cfl/sum(eig*ds); error (scalar field, may be NULL) receives V*|R|^2 per cell. */
int gr_euler_step(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, double cfl,
                  const char *time_scheme, gr_field *dt, gr_field *error);
/* Per-cell flux residual R (5 per interior cell: rho, rhou xyz, rhoe) of the current state,
 * without updating it (the sums fvmEulerSteadyOp forms, src/models/euler.cu:236-339). */
int gr_euler_residual(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, double *resid);
/* gr_euler_residual without any device allocation, for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  It has no
 * counterpart in the reference.  The residual is staged in a 32 MiB device array that the
 * library holds on each GPU it runs on until the process ends, so meshes of at most
 * GR_RESIDUAL_NOALLOC_MAX_CELLS interior cells are accepted (larger: GR_ERR_INVALID).  Calls
 * that overlap on one GPU (several host threads) share that array and race; calls from one
 * thread are ordered (each synchronises the library's stream before returning).  The current
 * device must be the mesh's. */
#define GR_RESIDUAL_NOALLOC_MAX_CELLS 838860
int gr_euler_residual_noalloc(gr_mesh *m, gr_field *rho, gr_field *rhou, gr_field *rhoe, double *resid);
/* Momentum reconstruction: 0 = the reference's transposed product (default), 1 = grad(u).d
 * per component. */
int gr_euler_set_corrected_reconstruction(gr_mesh *m, int on);

/* ---- Euler solver: the whole iteration on the device ----------------------------------------
 * PhysicsEuler::step (src/models/euler.cu:99-175) as one device pipeline per iteration:
 * all boundary regions in one launch, fused gradient+limiter, fused flux+update+residual,
 * device reduction of the residual norm, captured in a HIP graph; convergence is decided on
 * the device so the host does not synchronise per iteration. */
typedef struct gr_solver gr_solver;

typedef struct gr_solver_config {
    int dtype;         /* GR_F64: binary64 throughout; GR_F32: binary32 state, gradient,
                          limiter and flux with binary64 metrics and norms */
    double cfl;
    int limiter;       /* gr_limiter */
    int corrected;     /* momentum reconstruction, see gr_euler_set_corrected_reconstruction */
    int use_graph;     /* capture iterations in a HIP graph */
} gr_solver_config;

typedef struct gr_solve_result {
    int iterations;        /* iterations run by this call */
    int converged;         /* residual ratio fell below tol */
    double residual;       /* last un-normalised residual sum */
    double ratio;          /* last residual / reference */
    double reference;      /* normalisation used */
} gr_solve_result;

int gr_solver_create(gr_mesh *m, const gr_solver_config *cfg, gr_solver **out);
void gr_solver_free(gr_solver *s);
/* Boundary condition of a region (slip, or farfield with mach/p/T). */
int gr_solver_set_bc(gr_solver *s, const char *region, int kind, const double mach[3],
                     double pressure, double temperature);
/* Uniform state over cells and ghosts (euler_init). */
int gr_solver_init_freestream(gr_solver *s, const double mach[3], double pressure,
                              double temperature);
/* Set/get the conservative state, reference layouts: rho (N+G), rhou 3(N+G), rhoe (N+G). */
int gr_solver_set_state(gr_solver *s, const double *rho, const double *rhou, const double *rhoe);
int gr_solver_get_state(const gr_solver *s, double *rho, double *rhou, double *rhoe);
/* Runs up to max_iter iterations; stops after the iteration whose residual / reference
 * falls below tol (tol <= 0: never).  reference <= 0 uses the residual of the first
 * iteration of this call, as the reference's err0 (src/models/euler.cu:158-169). */
int gr_solver_iterate(gr_solver *s, int max_iter, double tol, double reference,
                      gr_solve_result *out);
/* Residual sums of the iterations of the last gr_solver_iterate call (n <= iterations). */
int gr_solver_history(const gr_solver *s, double *hist, int n);
/* Residual of the current state without updating it (one iteration's BC+prepare+flux). */
int gr_solver_residual_norm(gr_solver *s, double *out);
/* Mean launch duration in milliseconds of each kernel of an iteration, from n_iter back-to-back
 * launches of it on the current state (the state is not advanced; HIP events on the library's
 * stream around each run of launches): [0] boundary conditions (0: evaluated inside [1] and
 * [2]), [1] gradient+limiter, [2] flux+update, [3] residual reduction. */
int gr_solver_profile(gr_solver *s, int n_iter, double ms[4]);
/* Tile-size autotuner (no counterpart in the reference): for each candidate tile size (cells per
 * block of the gradient and flux kernels) times n_iter back-to-back launches of the iteration's
 * kernels and keeps the fastest for this solver; tile_cells (optional) receives it.  The state
 * is not advanced, and cell results do not depend on the tiling (residual sums add the same
 * cell terms in another grouping). */
int gr_solver_autotune(gr_solver *s, int n_iter, uint32_t *tile_cells);
/* Algorithmic bytes one launch of each kernel moves (every array element it reads or writes
 * counted once), same order as gr_solver_profile. */
int gr_solver_kernel_bytes(const gr_solver *s, double bytes[4]);
/* Bytes of the arrays one launch of each kernel reads or writes as this library lays them out
 * (per-slot gradient and flux records in place of gathered metrics, state records padded to 6
 * values, gradient + limiter records of 20 values in binary64 and 16 in binary32, tile records),
 * each element counted once; same order as gr_solver_profile.  Measured HBM traffic above these
 * is data fetched twice. */
int gr_solver_layout_bytes(const gr_solver *s, double bytes[4]);

/* ---- batched solver: up to 64 flow instances on one mesh (no reference counterpart) ---------
 * All instances share the mesh, the boundary kinds, p and T; each has its own far-field Mach
 * vector.  State layout is instance-inner (one wavefront per cell, lanes = instances); in
 * binary64 each instance evolves bit for bit as a gr_solver with the same far field would.
 * Convergence is per instance: an instance that reaches the tolerance stops changing while
 * the others continue. */
typedef struct gr_batch gr_batch;

int gr_batch_create(gr_mesh *m, uint32_t n_inst, const gr_solver_config *cfg, gr_batch **out);
void gr_batch_free(gr_batch *b);
/* Boundary kind of a region; far-field regions use each instance's Mach vector. */
int gr_batch_set_bc(gr_batch *b, const char *region, int kind);
/* Far field of every instance (mach: 3 per instance) and freestream initial state. */
int gr_batch_set_freestream(gr_batch *b, const double *mach, double pressure, double temperature);
/* State of one instance, reference layouts (rho N+G, rhou 3(N+G), rhoe N+G). */
int gr_batch_set_state(gr_batch *b, uint32_t inst, const double *rho, const double *rhou, const double *rhoe);
int gr_batch_get_state(gr_batch *b, uint32_t inst, double *rho, double *rhou, double *rhoe);
/* Up to max_iter iterations; instance k stops after the iteration whose residual / reference[k]
 * falls below tol (reference NULL or entries <= 0: that instance's first residual).
 * iterations/converged (n_inst entries, may be NULL) receive per-instance results; launched
 * receives the number of batch iterations that did work. */
int gr_batch_iterate(gr_batch *b, int max_iter, double tol, const double *reference, int *iterations,
                     int *converged, int *launched);
/* As gr_batch_iterate for the instances with active[k] != 0 (active NULL: all); the others keep
 * their state.  A diverged instance is reported in diverged[k] (may be NULL) and frozen, and is
 * not an error here. */
int gr_batch_iterate_masked(gr_batch *b, int max_iter, double tol, const double *reference, const int *active,
                            int *iterations, int *converged, int *diverged, int *launched);
/* Instances with mask[k] != 0 back to their freestream state. */
int gr_batch_reset_freestream(gr_batch *b, const int *mask);
int gr_batch_history(const gr_batch *b, uint32_t inst, double *hist, int n);
/* Residual sum of each instance's current state (n_inst entries), state unchanged. */
int gr_batch_residual_norms(gr_batch *b, double *out);
/* Mean launch durations (ms) from n_iter back-to-back launches of each stage on the current
 * state (not advanced): [0] gradient+limiter, [1] flux+update, [2] per-instance residual
 * reduction; and the algorithmic bytes of one launch of each. */
int gr_batch_profile(gr_batch *b, int n_iter, double ms[3]);
int gr_batch_kernel_bytes(const gr_batch *b, double bytes[3]);
/* Iterations captured per HIP graph (even, 2..200; default 10). */
int gr_batch_set_graph_iters(gr_batch *b, int iters);
/* Launch schedule of the batched kernels: schedule[0] cells (waves) per block of the gradient and
 * flux kernels (2, 4 or 8; default 4), [1] 1 if the flux kernel is capped at 3 waves per SIMD
 * (default 1), [2] iterations per HIP graph (as gr_batch_set_graph_iters), [3] instance groups of
 * the flux kernel (1, 2 or 4 in binary64, 1 or 2 in binary32; default 1): the flux sweeps the
 * cells once per group of 64/groups instances, a wave taking that many consecutive cells, which
 * shrinks the neighbour rows an XCD must keep in L2.  Results do not depend on it (per-cell arithmetic is the same; only the block
 * grouping of the residual partial sums changes, in the last bits of the residual norms).  The
 * reference launches a fixed CKW of 256 threads (include/datatypes.cuh:19,
 * src/models/euler.cu:662). */
int gr_batch_get_schedule(const gr_batch *b, int schedule[4]);
int gr_batch_set_schedule(gr_batch *b, const int schedule[4]);
/* Chooses the schedule: each (cells per block, flux occupancy, instance groups) choice times
 * n_iter gradient + flux launches back to back, the fastest is kept; then graph lengths 4, 10,
 * 20, 40 time 40 iterations replayed as gr_batch_iterate replays them.  The batch's state is
 * saved and restored.  schedule: the choice, as gr_batch_get_schedule. */
int gr_batch_autotune(gr_batch *b, int n_iter, int schedule[4]);

/* ---- post-processing (no reference counterpart) ---------------------------------------------
 * Pressure force on a wall region: F = sum over its faces of p_cell * n * A, n pointing out
 * of the fluid cell (into the body); p from the current state (first-order wall pressure).
 * Lift and drag are its projections.  gr_batch_force writes 3 entries per instance. */
int gr_solver_force(gr_solver *s, const char *region, double F[3]);
int gr_batch_force(gr_batch *b, const char *region, double *F);
/* Distance from each interior cell centre to the nearest face centre of the given wall
 * regions (out: n_cells entries). */
int gr_mesh_wall_distance(gr_mesh *m, const char *const *regions, uint32_t n_regions, double *out);

/* Bandwidth probe (no reference counterpart): a plain copy kernel over two device buffers of
 * `bytes` each, run `reps` times; gbps receives (read + written bytes) / mean kernel time.
 * Calibrates hardware byte counters and measures the achievable HBM rate. */
int gr_stream_copy_probe(size_t bytes, int reps, double *gbps);

/* ---- case driver and field output -----------------------------------------------------------
 * ASCII VTU of the interior cells as polyhedra (VTK type 42: sorted unique nodes, face stream,
 * face offsets) with cell data arrays, FieldsWriteVtu (src/fvmfieldsio.cu:5-287).  Field k has
 * n_comp[k] (1, 3 or 9) values per cell in values[k] (at least n_cells*n_comp entries; ghost
 * entries after them are ignored — the reference writes them past NumberOfCells). */
int gr_write_vtu(const gr_mesh *m, const char *path, uint32_t n_fields, const char *const *names,
                 const uint32_t *n_comp, const double *const *values);
/* Runs a JSON case file as PhysicsEuler does (src/models/euler.cu:15-221): "mesh" (su2
 * converted to .fvm beside it), "init" {mach{x,y,z}, pressure, temperature}, "bcs" {region:
 * {type slip | farfield (+ mach, pressure, temperature)}}, "controls" {cfl, time "steady",
 * tolerance; optional max_iterations, precision "f64" | "f32", renumber}, optional "save"
 * {file} (VTU).  Every key is checked before any device work; gr_case_check only checks. */
int gr_case_check(const char *case_path);
int gr_case_run(const char *case_path, int verbose, gr_solve_result *result);

/* ---- incompressible pressure-correction solver ----------------------------------------------
 * Replaces the reference's PhysicsNSI (include/models/nsi.cuh:8-40), a stub whose
 * src/models/nsi.cu does not compile and holds no arithmetic.  Collocated cell-centred
 * scheme, constant density 1: ghosts (no-slip walls moving with the region's velocity, zero
 * normal pressure gradient), least-squares gradients, predictor (second-order upwind
 * convection, two-point diffusion, explicit pressure gradient), momentum-interpolated face
 * fluxes, a pressure-correction Poisson equation solved by Jacobi-preconditioned conjugate
 * gradients, and flux/velocity/pressure corrections.  State: velocity (N+G)x3 AoS, pressure
 * N+G, face fluxes n_faces (volume flux from the face's lower- to its higher-numbered cell). */
typedef struct gr_nsi gr_nsi;
typedef struct gr_nsi_config {
    int dtype;       /* GR_F64 (all binary64), or GR_F32 (mixed): binary64 velocity, pressure and
                        face fluxes, binary32 gradients, predictor increments u* - u and CG
                        vectors; binary64 dot products and continuity residual in both */
    double nu, dt;   /* kinematic viscosity, time step */
    int cg_max;      /* CG iterations per step at most */
    double cg_tol;   /* stop CG when |r| < cg_tol |b| (0: always cg_max iterations) */
} gr_nsi_config;

int gr_nsi_create(gr_mesh *m, const gr_nsi_config *cfg, gr_nsi **out);
void gr_nsi_free(gr_nsi *s);
/* Wall velocity of a boundary region (default: at rest). */
int gr_nsi_set_wall(gr_nsi *s, const char *region, const double u[3]);
/* Any of u, p, F may be NULL (zero). */
int gr_nsi_set_state(gr_nsi *s, const double *u, const double *p, const double *F);
int gr_nsi_get_state(gr_nsi *s, double *u, double *p, double *F);
/* n_steps steps; cont / cg_iters (n_steps entries, may be NULL) receive each step's continuity
 * residual (sum over cells of |net outflow of the predicted fluxes|) and CG iterations. */
int gr_nsi_step(gr_nsi *s, int n_steps, double *cont, int *cg_iters);
/* Steps until the continuity ratio falls below tol (or max_steps): the continuity residual of
 * the step over the largest of the solver's first five steps since creation or the last
 * gr_nsi_set_state (the normalisation is fixed once five steps exist; before that the ratio is
 * not tested).  steps: steps run by this call; ratio: the last step's ratio. */
int gr_nsi_iterate(gr_nsi *s, int max_steps, double tol, int *steps, double *ratio);
/* The last n steps' continuity residuals and CG iterations. */
int gr_nsi_history(const gr_nsi *s, double *cont, int *cg_iters, int n);
/* Mean durations (ms) from HIP events around n_iter back-to-back launches: [0] the predictor
 * graph (ghosts, gradients, predictor, fluxes, divergence, CG set-up), [1] k_cg_iter (one CG
 * iteration), [2] the correction graph.  The solve is left unchanged (velocity, pressure and
 * face fluxes are restored). */
int gr_nsi_profile(gr_nsi *s, int n_iter, double ms[3]);

/* ---- warm-start network (no reference counterpart) -------------------------------------------
 * A 4-layer MLP n_in -> 256 -> 256 -> 256 -> n_out (tanh on the hidden layers) trained and run
 * with bfloat16 matrix-core products accumulated in binary32.  Parameters are binary32 in a
 * padded layout: per layer W [in_pad][out_pad] (row-major, y = x W + b) then b [out_pad], with
 * (in_pad, out_pad) = (64,256), (256,256), (256,256), (256,32).  Rows are samples, row-major. */
typedef struct gr_mlp gr_mlp;

/* n_in 1..64, n_out 1..32; Xavier-uniform weights from seed, zero biases. */
int gr_mlp_create(uint32_t n_in, uint32_t n_out, uint64_t seed, gr_mlp **out);
void gr_mlp_free(gr_mlp *n);
int gr_mlp_n_params(const gr_mlp *n, size_t *count);
int gr_mlp_get_params(const gr_mlp *n, float *out);
int gr_mlp_set_params(gr_mlp *n, const float *in);
/* y [M][n_out] = MLP(x [M][n_in]), M up to 2^24 rows. */
int gr_mlp_forward(gr_mlp *n, const float *x, int M, float *y);
/* Mean squared error against t [M][n_out] and its gradient (padded layout; grad may be NULL),
 * parameters unchanged.  M up to 2^20. */
int gr_mlp_loss_grad(gr_mlp *n, const float *x, const float *t, int M, double *loss, float *grad);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on the mean squared error of one mini-batch;
 * loss receives the loss before the step.  Deterministic: the same calls give the same bits. */
int gr_mlp_train_step(gr_mlp *n, const float *x, const float *t, int M, float lr, double *loss);
/* Keeps a training set resident on the device (x [rows][n_in], t [rows][n_out]). */
int gr_mlp_set_data(gr_mlp *n, const float *x, const float *t, uint64_t rows);
/* `steps` Adam steps on mini-batches of `batch` rows of the resident set: each pass over the set
 * visits its rows in a seeded permuted order; the learning rate decays from lr to lr_final on a
 * cosine.  losses (steps entries, may be NULL) receives each step's mini-batch loss. */
int gr_mlp_train(gr_mlp *n, int batch, int steps, float lr, float lr_final, uint64_t seed, double *losses);

/* ---- learned warm start of the batched solver (no reference counterpart) --------------------
 * Network rows are (cell, instance) pairs, row = cell * n_inst + instance, over interior cells.
 * Raw inputs (n_in = 6): the cell centre x relative to an origin o, compressed as
 * (x - o) / (1 + |x - o|) (3 entries), log(d + 1e-3) with d the wall distance, the instance's
 * Mach number |M| and angle of attack atan2(M_y, M_x) in radians.  Raw targets (n_out = 5):
 * rho/rho_inf - 1, (rho u - rho_inf u_inf)/(rho_inf c_inf) (3 entries) and
 * (rho E - rho_inf E_inf)/(rho_inf c_inf^2), with the far-field state of the instance. */
typedef struct gr_warm_norm {
    float origin[3];                 /* origin of the compressed centre coordinates */
    float in_mean[6], in_std[6];     /* network input = (raw - mean) / std */
    float out_mean[5], out_std[5];   /* raw target = network output * std + mean */
    float trust[2];                  /* the predicted deviation from freestream is applied in full
                                        up to wall distance trust[0] and blended out by trust[1]
                                        (trust[1] <= 0: everywhere) */
} gr_warm_norm;

/* All instances' states at once: rho/rhoe [n_inst][N+G], rhou [n_inst][N+G][3]. */
int gr_batch_set_states(gr_batch *b, const double *rho, const double *rhou, const double *rhoe);
int gr_batch_get_states(gr_batch *b, double *rho, double *rhou, double *rhoe);
/* Raw inputs x [n_inst*N][6] and targets t [n_inst*N][5] of the batch's current states
 * (a training set); wall_distance has n_cells entries. */
int gr_batch_warm_rows(gr_batch *b, const double *wall_distance, const float origin[3], float *x, float *t);
/* Resets every instance to its freestream (gr_batch_set_freestream must have run), records
 * the freestream residuals (fs_residual, n_inst entries: the normalisation of the solve that
 * follows), replaces the states by the network's prediction and then judges each prediction
 * by its residual: an instance whose predicted residual is not finite or not below
 * max_ratio * its freestream residual goes back to freestream.  used[k] (may be NULL) = 1 where
 * the prediction was kept.  The network must have n_in = 6 and n_out = 5. */
int gr_batch_warm_start(gr_batch *b, gr_mlp *n, const double *wall_distance, const gr_warm_norm *norm,
                        double max_ratio, double *fs_residual, int *used);

#ifdef __cplusplus
}
#endif
#endif

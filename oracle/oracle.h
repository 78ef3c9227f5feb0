/*
 * oracle.h -- FP64 serial CPU restatement of GRAPHITE's finite-volume residual path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the smoke check in
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  The product
 * (graphite_amd/, libgraphite_hip.so) never links, loads or calls it.
 *
 * Semantics follow the reference (citations are Alexsimulation/graphite file:line) with the
 * decisions listed below; every routine is written fresh in plain C from the reference's
 * behaviour, not copied.
 *   - Arithmetic: IEEE-754 binary64, correctly rounded sqrt/div, no FMA contraction
 *     (built with -ffp-contract=off), operations evaluated in the same order as the
 *     reference's expressions so that a GPU kernel with the same order can match bit for bit.
 *   - norm(double3) uses a double sqrt (the reference's double build calls sqrtf,
 *     include/datatypes.cuh:194-196).
 *   - The explicit update is Jacobi: U^{n+1} is written to separate arrays (the reference
 *     updates in place while neighbours read, src/models/euler.cu:376-378).
 *   - Ghost cells are cells n_cells .. n_cells+n_ghosts-1 (as cuMeshFromSu2 writes them).
 *
 * Field layouts mirror the reference's FvmField AoS layout (include/fvmfields.cuh:16-22):
 *   scalar value[n], gradient vector[n] (xyz), limiter scalar[n];
 *   vector value[n] (xyz), gradient tensor[n] = columns (u,v,w) each xyz, i.e. 9 doubles
 *   where entries 0..2 are grad(value.x), 3..5 grad(value.y), 6..8 grad(value.z)
 *   (src/fvmops.cu:101-110), limiter vector[n].
 *
 * Parity status per routine is recorded in the header comment of each function below.
 */
#ifndef GRAPHITE_ORACLE_H
#define GRAPHITE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- connectivity (src/meshio.cu:297-475, 523-764, 902-1138) ---------------------------- */

/* Mesh in the reference's CSR form plus metric arrays.  All arrays are malloc'd by the
 * oracle and released by orc_mesh_free. */
typedef struct orc_mesh {
    uint32_t n_nodes, n_faces, n_cells, n_ghosts, n_regions;
    double *nodes;                 /* 3*n_nodes */
    uint32_t *face_nodes_starts;   /* n_faces+1 */
    uint32_t *face_nodes;
    uint32_t *cell_faces_starts;   /* n_cells+n_ghosts+1 */
    uint32_t *cell_faces;
    uint32_t *cell_connects;
    uint32_t *ghost_cell_ids;      /* n_ghosts */
    uint32_t *ghost_cell_starts;   /* n_regions+1 */
    /* metrics (filled by orc_mesh_metrics) */
    double *face_centers;          /* 3*n_faces */
    double *face_normals;          /* 3*n_faces */
    double *face_areas;            /* n_faces */
    double *cell_centers;          /* 3*(n_cells+n_ghosts) */
    double *cell_volumes;          /* n_cells+n_ghosts */
    double *cell_matrices;         /* 9*n_cells, column-major tensor (u,v,w) */
} orc_mesh;

/* Hash of a face key, src/meshio.cu:551-563: product over the sorted key of
 * (998244353 + 2*k) in uint32 arithmetic, halved, modulo the table size. */
uint32_t orc_face_hash(const uint32_t *key, uint32_t n, uint32_t table_size);

/* Unique-face extraction and CSR assembly as cuMeshFromSu2 does it
 * (src/meshio.cu:713-764 su2ComputeFaces, :902-1138 face/ghost/cell arrays), with a literal
 * chained hash table of n_elems+1 buckets (:718, FlexArray size :780).
 * Markers: n_markers regions, marker m has counts[m] boundary elements; their types, node
 * starts (per marker, local, counts[m]+1 entries) and nodes are concatenated over markers.
 * Returns 0 and a mesh on success; 1 on an inconsistent topology (a face not shared by
 * exactly two cells, a marker face that is not a face of the volume mesh, an unknown element
 * type), as the reference returns 1 (:990-999, :1106-1128). */
int orc_build_mesh(uint32_t n_nodes, const double *nodes,
                   uint32_t n_elems, const uint32_t *elem_types,
                   const uint32_t *elem_starts, const uint32_t *elem_nodes,
                   uint32_t n_markers, const uint32_t *marker_counts,
                   const uint32_t *marker_types, const uint32_t *marker_starts,
                   const uint32_t *marker_nodes, orc_mesh **out);

void orc_mesh_free(orc_mesh *m);

/* ---- metrics (src/mesh.cu:44-219) --------------------------------------------------------- */

/* All four metric kernels in order: faces, cells, ghost cells, LSQ matrices. */
void orc_mesh_metrics(orc_mesh *m);
/* Validity counts after orc_mesh_metrics (non-positive volumes, faces that do not separate
 * their two cell centres, singular least-squares matrices, degenerate faces). */
void orc_mesh_check(const orc_mesh *m, uint32_t counts[4]);

/* ---- FV operators (src/fvmops.cu) -------------------------------------------------------- */

enum { ORC_LIM_MICHALAK = 0, ORC_LIM_BARTH_JESPERSEN = 1, ORC_LIM_VENKATAKRISHNAN = 2 };

/* Limiter function of one face, src/fvmops.cu:129-169 (kind selects LIMITER_FUNC :118-125). */
double orc_limiter_face(double u, const double grad[3], double umin, double umax,
                        const double d[3], int kind);

/* Scalar/VectorVarComputeGradient, src/fvmops.cu:51-111 (interior cells only). */
void orc_scalar_gradient(const orc_mesh *m, const double *value, double *grad);
void orc_vector_gradient(const orc_mesh *m, const double *value, double *grad);
/* Scalar/VectorVarComputeLimiters, src/fvmops.cu:175-266 (interior cells only). */
void orc_scalar_limiter(const orc_mesh *m, const double *value, const double *grad,
                        double *lim, int kind);
void orc_vector_limiter(const orc_mesh *m, const double *value, const double *grad,
                        double *lim, int kind);
/* fvmScalGradLimBc / fvmVecGradLimBc, src/fvmops.cu:272-304. */
void orc_gradlim_bc(const orc_mesh *m, int ncomp, double *grad, double *lim);
/* fvmScalarBcFixed / fvmScalarBcZeroGrad / fvmVectorBcZeroGrad, src/fvmops.cu:308-365. */
void orc_bc_fixed(const orc_mesh *m, uint32_t region, int ncomp, const double *v, double *value);
void orc_bc_zero_grad(const orc_mesh *m, uint32_t region, int ncomp, double *value);
/* FvmFieldReduceSum, src/fvmops.cu:369-383, summed over interior cells with Kahan
 * compensation in FP64. */
double orc_reduce_sum(const double *v, uint32_t n);

/* ---- Euler model (src/models/euler.cu) --------------------------------------------------- */

/* euler_init, src/models/euler.cu:385-409, over cells and ghosts. */
void orc_euler_init(const orc_mesh *m, double *rho, double *rhou, double *rhoe,
                    const double mach[3], double p, double T);
/* euler_slip_bc, src/models/euler.cu:412-442. */
void orc_euler_slip_bc(const orc_mesh *m, uint32_t region, double *rho, double *rhou,
                       double *rhoe);
/* euler_inletoutlet_bc, src/models/euler.cu:446-520. */
void orc_euler_farfield_bc(const orc_mesh *m, uint32_t region, double *rho, double *rhou,
                           double *rhoe, const double mach[3], double p, double T);

/* Euler state with gradients and limiters, reference AoS layouts (see top of file). */
typedef struct orc_euler_state {
    double *rho, *rhou, *rhoe;           /* (N+G), 3(N+G), (N+G) */
    double *grad_rho, *grad_rhou, *grad_rhoe;  /* 3(N+G), 9(N+G), 3(N+G) */
    double *lim_rho, *lim_rhou, *lim_rhoe;     /* (N+G), 3(N+G), (N+G) */
} orc_euler_state;

/* FvmEulerStepPrepare, src/models/euler.cu:597-647: gradient, limiter, ghost copy for rho,
 * rhou, rhoe. */
void orc_euler_step_prepare(const orc_mesh *m, orc_euler_state *s, int limiter_kind);

/* FvmEulerStep 'steady', src/models/euler.cu:649-673 with euler_steady_step :348-381 and
 * fvmEulerSteadyOp :236-339, Jacobi form: reads s, writes the updated interior values to
 * rho_new/rhou_new/rhoe_new (ghost entries are copied unchanged), err[i] to err (may be
 * NULL), the per-cell flux residual R_i (5 per cell: rho, rhou xyz, rhoe) to resid (may be
 * NULL) and the local time step cfl/sum(eig*ds) to dt (may be NULL).
 * corrected != 0 reconstructs momentum with grad(u)*d per component instead of the
 * reference's transposed product (src/models/euler.cu:252-253, include/datatypes.cuh:296). */
void orc_euler_step(const orc_mesh *m, const orc_euler_state *s, double cfl, int corrected,
                    double *rho_new, double *rhou_new, double *rhoe_new, double *err,
                    double *resid, double *dt);

/* Boundary condition of one region, applied in the order given each iteration. */
enum { ORC_BC_SLIP = 0, ORC_BC_FARFIELD = 1 };
typedef struct orc_bc {
    uint32_t region;
    int kind;
    double mach[3];
    double p, T;
} orc_bc;

/* n_iter iterations of PhysicsEuler::step (src/models/euler.cu:99-175): BCs, prepare,
 * Jacobi step, residual sum.  history[k] receives the un-normalised residual sum of
 * iteration k.  Stops early when history[k]/ref < tol (ref = history[0] if ref_norm <= 0,
 * else ref_norm); returns the number of iterations run.  The state in s is advanced in
 * place (the solution after the last iteration).  Uses n_threads OpenMP threads
 * (results do not depend on the thread count). */
int orc_euler_iterate(const orc_mesh *m, orc_euler_state *s, const orc_bc *bcs, int n_bcs,
                      double cfl, int limiter_kind, int corrected, int n_iter, double tol,
                      double ref_norm, double *history, int n_threads);

/* Warm-start MLP (oracle_mlp.c): padded parameter layout of the device library. */
size_t orc_mlp_n_params(void);
double orc_bf16_round(double v);
int orc_mlp_forward(const double *params, int n_in, int n_out, const double *x, int M, int emulate_bf16,
                    double *y);
int orc_mlp_loss_grad(const double *params, int n_in, int n_out, const double *x, const double *t, int M,
                      int emulate_bf16, double *loss, double *grad);

/* Finite-volume operator library (oracle_ops.c), include/fvmops.cuh:96-219.  Inputs in the
 * AoS layouts above (val/grad/lim of the input field over N+G; phi_const a vector; phi_field a
 * vector field over N+G); out: 1, 3 or 9 values per interior cell. */
enum {
    ORC_OP_LAPLACIAN_C = 0,
    ORC_OP_SCAL_GRAD_C = 1,
    ORC_OP_SCAL_GRAD_C_CORRECT = 2,
    ORC_OP_VEC_GRAD_C = 3,
    ORC_OP_BOUND_SCAL_GRAD_C = 4,
    ORC_OP_BOUND_VEC_GRAD_C = 5,
    ORC_OP_DIV_C = 6,
    ORC_OP_CONV_CONST_SCAL_U = 7,
    ORC_OP_CONV_CONST_SCAL_U2 = 8,
    ORC_OP_CONV_SCAL_U = 9,
    ORC_OP_CONV_VEC_U = 10
};
int orc_fvm_op(const orc_mesh *m, int op, const double *val, const double *grad, const double *lim,
               const double *phi_const, const double *phi_field, double *out);

/* Incompressible pressure-correction step (oracle_nsi.c); no reference arithmetic exists.
 * u: 3(N+G) AoS velocity, p: N+G pressure, F: n_faces volume fluxes from the face's
 * lower-numbered cell to the higher.  Walls: no slip with the region's velocity (regions not
 * listed: at rest).  cont receives sum_i |net outflow of the predicted fluxes|. */
typedef struct orc_nsi_wall {
    uint32_t region;
    double u[3];
} orc_nsi_wall;
typedef struct orc_nsi_params {
    double nu, dt;
    int cg_max;       /* conjugate-gradient iterations at most */
    double cg_tol;    /* stop when |r| < cg_tol |b| (0: run cg_max iterations) */
} orc_nsi_params;
int orc_nsi_step(const orc_mesh *m, double *u, double *p, double *F, const orc_nsi_wall *walls, int n_walls,
                 const orc_nsi_params *prm, double *cont, int *cg_iters);
/* Solves sum_f a_f (x_i - x_j) = b_i (interior faces, a = A/D) by Jacobi-preconditioned CG from
 * x = 0 with b and x shifted to zero mean; returns the iterations run. */
int orc_nsi_poisson(const orc_mesh *m, const double *b, double *x, int cg_max, double cg_tol);

#ifdef __cplusplus
}
#endif
#endif

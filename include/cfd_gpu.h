/*
 * cfd_gpu.h — C-ABI of the MI355X (gfx950) residual + JFNK-GMRES library.
 *
 * The reference solver (AmberMond/mycode) has no FFI layer: its seam is the
 * set of `Solver` member functions listed beside each entry point below.  A
 * host that today calls those members calls these functions instead; every
 * argument is a plain pointer or size, no C++ or torch type crosses the
 * boundary.
 *
 * Conventions
 *  - A "state vector" passed through a host pointer is the reference's
 *    std::vector<State> memory: N cells x 6 doubles (rho, rhou, rhov, rhoE,
 *    rhok, rhow), 48-byte stride, in the reference's cell order (Gmsh element
 *    order).  With nvar == 4 the last two slots are ignored on input and
 *    written as 0 on output.
 *  - Device-handle variants (cfd_dev_*) take device pointers to the library's
 *    internal layout: structure-of-arrays, nvar rows of `cfd_pitch(ctx)`
 *    elements, cells in the context's RCM order.  They are asynchronous on
 *    the context stream.
 *  - Every function returns 0 on success and a negative cfd_status on
 *    failure; cfd_last_error(ctx) gives a message.  Nothing throws or exits
 *    across the boundary.  A context is not thread-safe.
 */
#ifndef CFD_GPU_H
#define CFD_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum cfd_status {
  CFD_OK = 0,
  CFD_ERR_ARG = -1,      /* bad argument (null pointer, size, enum value)   */
  CFD_ERR_HIP = -2,      /* a HIP runtime call failed                       */
  CFD_ERR_MESH = -3,     /* inconsistent mesh (non-manifold edge, bad node) */
  CFD_ERR_STATE = -4,    /* call order (e.g. Jv before a local time step)   */
  CFD_ERR_NUMERIC = -5,  /* NaN / non-positive density or pressure seen     */
  CFD_ERR_IO = -6        /* file could not be read                          */
};

/* ---- scheme selectors: values follow Types.hpp:38-49 ------------------- */
enum cfd_flux { CFD_FLUX_LLF = 0, CFD_FLUX_ROE = 1, CFD_FLUX_HLLC = 2, CFD_FLUX_AUSM = 3 };
enum cfd_spatial {
  CFD_FIRST_ORDER = 0,
  CFD_SECOND_ORDER_CENTRAL = 1,
  CFD_SECOND_ORDER_LIMITED = 2
};
/* Boundary-condition sets.  The reference selects one of these by
 * commenting code in and out of Solver::getBoundaryGhostState:
 *   BUMP       Solver.cpp:806-863  (the active one: P0/T0 inlet, static-p
 *                                    outlet, slip wall)
 *   SUPERSONIC Solver.cpp:973-1011 (frozen inflow, extrapolated outflow)
 *   PLATE      Solver.cpp:1088-1139 (frozen farfield 1/5, extrapolated
 *                                    outflow, slip 3, no-slip 4)
 *   SWBLI      PLATE plus tag 6 = frozen post-shock state of an oblique
 *              shock (no reference counterpart). */
enum cfd_bcset { CFD_BC_BUMP = 0, CFD_BC_SUPERSONIC = 1, CFD_BC_PLATE = 2, CFD_BC_SWBLI = 3 };

/* ---- physical / numerical parameters ----------------------------------- */
typedef struct cfd_params {
  double gamma;          /* ratio of specific heats (Solver.hpp:88, Flux.hpp:9) */
  double R_gas;          /* gas constant (Solver.hpp:89)                        */
  double M_inf;          /* freestream Mach (Solver.hpp:29)                     */
  double P_total_inlet;  /* Solver.hpp:93                                       */
  double T_total_inlet;  /* Solver.hpp:94                                       */
  double P_static_exit;  /* Solver.hpp:95                                       */
  double nu;             /* kinematic viscosity (Solver.hpp:31, default 0.01)   */
  double Pr;             /* Prandtl number, 0.72 (Solver.cpp:541)               */
  int flux;              /* enum cfd_flux                                       */
  int spatial;           /* enum cfd_spatial                                    */
  int bcset;             /* enum cfd_bcset                                      */
  int viscous;           /* 0: total flux = inviscid flux (Solver.cpp:792, as
                            shipped); 1: inviscid - viscous (Solver.cpp:787)    */
  int nvar;              /* 4 (Euler / laminar) or 6 (SST k-omega)              */
  int turb;              /* 0 none, 1 SST k-omega (no reference counterpart)    */
  int flux_alt;          /* flux used on faces flagged by cfd_set_face_flux    */
  int turb_first_order;  /* SST only: 0 = k and omega reconstructed (MUSCL) like
                            the mean flow; 1 = k and omega convected first-order
                            upwind (the cell values), the mean flow unchanged  */
  /* post-shock state for the SWBLI tag-6 boundary (primitive) */
  double shock_rho, shock_u, shock_v, shock_p;
  /* SST k-omega: freestream specific k and omega of the frozen inflow */
  double turb_k_inf, turb_w_inf;
  double reserved[6];
} cfd_params;

/* Helpers that fill the freestream fields exactly as the reference does. */
int cfd_params_default(cfd_params* p);                               /* Solver.hpp:27-40 */
int cfd_params_flow_conditions(cfd_params* p, double M, double P, double T);   /* Solver.cpp:12-32 */
int cfd_params_freestream_static(cfd_params* p, double M, double P, double T); /* Solver.cpp:2017-2025 */
int cfd_params_kinematic_viscosity(cfd_params* p, double Re, double L);       /* Solver.cpp:34-43 */
/* Uniform initial state as Solver::initialize (Solver.cpp:46-64), 6 doubles. */
int cfd_params_initial_state(const cfd_params* p, double* U6_out);

/* ---- mesh --------------------------------------------------------------- */
/* A mesh description in the form Mesh::readGmsh leaves it (Mesh.cpp:30-97):
 * nodes with 1-based ids (x[0], y[0] unused), cells as CSR lists of 1-based
 * node ids in file order, boundary edges as (n1, n2, physical tag) triples. */
typedef struct cfd_mesh_desc {
  int n_nodes;
  const double* x;       /* n_nodes + 1 */
  const double* y;       /* n_nodes + 1 */
  int n_cells;
  const int* cell_ptr;   /* n_cells + 1 */
  const int* cell_nodes; /* cell_ptr[n_cells], 3 or 4 per cell */
  int n_bedges;
  const int* bedges;     /* 3 * n_bedges */
} cfd_mesh_desc;

typedef struct cfd_mesh cfd_mesh;

/* Connectivity + geometry on the host; replaces Mesh::load's
 * buildConnectivity (Mesh.cpp:102-143) and computeGeometry
 * (Mesh.cpp:148-202).  Face order, owner/neighbor and every double are
 * identical to the reference.  Also computes the RCM cell permutation and
 * the face edge coloring used by the device kernels. */
int cfd_mesh_build(const cfd_mesh_desc* desc, cfd_mesh** out);
/* Mesh::load (Mesh.cpp:10-26) from a Gmsh v2.2 ASCII file. */
int cfd_mesh_read_gmsh(const char* path, cfd_mesh** out);
/* Seeded synthetic meshes of the benchmark configurations (1..5), scaled
 * by `scale` in each direction (1.0 = full size).  The description arrays
 * stay owned by the mesh and are valid until cfd_mesh_destroy. */
int cfd_mesh_generate(int config, double scale, cfd_mesh** out);
int cfd_mesh_destroy(cfd_mesh* m);
int cfd_mesh_sizes(const cfd_mesh* m, int* n_nodes, int* n_cells, int* n_faces, int* n_colors);
int cfd_mesh_get_desc(const cfd_mesh* m, cfd_mesh_desc* out);
/* Faces in reference order (Types.hpp:16-27 field for field). */
int cfd_mesh_get_faces(const cfd_mesh* m, int* node1, int* node2, int* owner, int* neighbor,
                       int* bc_tag, double* nx, double* ny, double* area, double* x_mid,
                       double* y_mid);
int cfd_mesh_get_cells(const cfd_mesh* m, double* volume, double* x_c, double* y_c);
/* perm[i] = RCM position of reference cell i; color[f] = color of reference face f. */
int cfd_mesh_get_reorder(const cfd_mesh* m, int* perm, int* face_color);
/* FNV-1a hash of the mesh arrays, logged by benchmarks. */
uint64_t cfd_mesh_hash(const cfd_mesh* m);
/* Move the nodes (n_nodes + 1 coordinates, index 0 unused), keeping the
 * topology, face order and colors; geometry is recomputed exactly as
 * Mesh::computeGeometry.  Used for design sensitivities. */
int cfd_mesh_set_nodes(cfd_mesh* m, const double* x, const double* y);
/* Write Gmsh v2.2 ASCII with %.17g coordinates. */
int cfd_mesh_write_gmsh(const cfd_mesh* m, const char* path);
/* Line segments of the line-implicit preconditioner (no reference
 * counterpart): chains of strongly coupled cells (coupling A_f / |dx . n_f|,
 * cells whose largest coupling is at least 4x their smallest), at most 64
 * cells each, every other cell a segment of its own; segments listed color
 * by color (adjacent segments differ in color).  seg_ptr (n_segments + 1),
 * seg_cells (N, reference cell ids in chain order) and segcol_ptr
 * (n_colors + 1) may be NULL to query the sizes. */
int cfd_mesh_get_lines(const cfd_mesh* m, int* n_segments, int* n_colors, int* seg_ptr, int* seg_cells,
                       int* segcol_ptr);
/* Cell colors of the multicolor block symmetric Gauss-Seidel preconditioner
 * (no reference counterpart): a distance-1 coloring, adjacent cells differ.
 * color (N, reference cell order) may be NULL to query n_colors. */
int cfd_mesh_get_cell_colors(const cfd_mesh* m, int* n_colors, int* color);
/* main.cpp saveResults (main.cpp:8-51): legacy ASCII VTK of the mesh with the
 * cell fields Mach, Pressure and Density of U6 (reference cell order),
 * pressure with the given gamma (the reference uses Flux.hpp's 1.4), numbers
 * at the reference stream's default 6 significant digits. */
int cfd_mesh_write_vtk(const cfd_mesh* m, const double* U6, double gamma, const char* path);

/* ---- context ------------------------------------------------------------ */
typedef struct cfd_ctx cfd_ctx;

/* Solver::Solver (Solver.cpp:7-10) + setNumericalMethod (Solver.hpp:33-36):
 * uploads the SoA mesh, RCM permutation and colors to the current device.
 * The mesh is borrowed and must outlive the context, as the reference's
 * Solver holds a const Mesh& (Solver.hpp:71). */
int cfd_ctx_create(const cfd_mesh* m, const cfd_params* p, cfd_ctx** out);
int cfd_ctx_destroy(cfd_ctx* ctx);
/* Re-upload the geometry of `m` (same topology as the context's mesh, e.g.
 * after cfd_mesh_set_nodes); the local time step must be recomputed. */
int cfd_ctx_update_geometry(cfd_ctx* ctx, const cfd_mesh* m);
const char* cfd_last_error(const cfd_ctx* ctx);
int cfd_ctx_set_params(cfd_ctx* ctx, const cfd_params* p);
int cfd_ctx_get_params(const cfd_ctx* ctx, cfd_params* p);
/* Elements per SoA row of device vectors. */
int cfd_pitch(const cfd_ctx* ctx);
int cfd_ncells(const cfd_ctx* ctx);
/* hipStream_t the context launches and copies on, as void*. */
void* cfd_stream(const cfd_ctx* ctx);
/* The stream the context created for itself (its default), as void*. */
void* cfd_ctx_own_stream(const cfd_ctx* ctx);
/* Launch and copy on the caller's hipStream_t `stream` (any stream of the
 * context's device, including the null stream 0) from now on, so that work
 * the caller queues on that stream before and after a cfd_dev_* call is
 * ordered with it without device-wide synchronisation.  Work already queued
 * on the previous stream is waited for first.  The caller keeps ownership;
 * cfd_ctx_set_stream(ctx, cfd_ctx_own_stream(ctx)) returns to the context's
 * own stream, which it does before the caller destroys theirs.  No reference
 * counterpart (the reference is single-threaded CPU code). */
int cfd_ctx_set_stream(cfd_ctx* ctx, void* stream);
/* Per-face flux override (0 = params.flux, 1 = params.flux_alt), reference face order. */
int cfd_set_face_flux(cfd_ctx* ctx, const unsigned char* use_alt);
/* Residual assembly strategy: 0 = edge-colored scatter (faces laid out color
 * by color), 1 = face-flux buffer + per-cell gather in reference face order
 * (faces laid out by owner; the default).  Changing it re-lays the faces. */
int cfd_set_assembly(cfd_ctx* ctx, int mode);

/* ---- host-pointer drop-ins (reference order, AoS stride 6) -------------- */
/* Solver::computeResiduals(U_in, R_out, update_limiters) — Solver.cpp:529-801 */
int cfd_residual(cfd_ctx* ctx, const double* U6, double* R6, int update_limiters);
/* Solver::cellGradients after the last residual, 10 doubles per cell in
 * CellGradient order (drhodx, drhody, dudx, dudy, dvdx, dvdy, dpdx, dpdy,
 * dTdx, dTdy) — Solver.hpp:13-24, Solver.cpp:81-149 */
int cfd_get_gradients(cfd_ctx* ctx, double* grad10);
/* Solver::limiters — Solver.cpp:152-253 */
int cfd_get_limiters(cfd_ctx* ctx, double* phi);
int cfd_set_limiters(cfd_ctx* ctx, const double* phi);
/* Solver::computeLocalTimeStep(CFL) with U_current = U6 — Solver.cpp:1388-1422 */
int cfd_local_timestep(cfd_ctx* ctx, const double* U6, double cfl);
int cfd_get_dt(cfd_ctx* ctx, double* dt);
/* Solver::computeDiagonalPreconditioner with U_current = U6 — Solver.cpp:1424-1465 */
int cfd_diag_preconditioner(cfd_ctx* ctx, const double* U6, double* D);
/* Solver::applyJacobian(U_base, R_base, v, Jv) — Solver.cpp:1331-1386 */
int cfd_jv_fd(cfd_ctx* ctx, const double* U6, const double* R6, const double* v6, double* Jv6);
/* Forward-mode tangent J v = dR/dU v + (V/dtau) v with the limiter frozen —
 * the exact derivative of what Solver::applyJacobian (Solver.cpp:1331-1386)
 * approximates by a finite difference.  No reference counterpart.
 * precision: 64 = FP64 tangent; 32 = FP32 tangent (values and derivatives). */
int cfd_jv_tangent(cfd_ctx* ctx, const double* U6, const double* v6, double* Jv6, int precision);
/* Reverse-mode product J^T w = (dR/dU)^T w + (V/dtau) w with the limiter
 * frozen: the transpose of cfd_jv_tangent, for discrete-adjoint solves.
 * Gradients are recomputed, not stored.  No reference counterpart. */
int cfd_vjp_adjoint(cfd_ctx* ctx, const double* U6, const double* w6, double* JTw6);
/* Block-Jacobi preconditioner (no reference counterpart; the reference uses
 * the scalar diagonal of Solver.cpp:1424-1465): per cell the nv x nv diagonal
 * block B of the first-order Jacobian plus V/dtau and a viscous diagonal,
 * and its inverse stored in `precision`.  B and Binv (N x nv x nv, row
 * major, reference cell order, may be NULL) are copied out for inspection. */
int cfd_block_jacobi(cfd_ctx* ctx, const double* U6, int precision, double* B, double* Binv);
/* Multicolor block symmetric Gauss-Seidel preconditioner (no reference
 * counterpart): M = (D + L) D^-1 (D + U) of the same first-order Jacobian,
 * D the blocks of cfd_block_jacobi and L, U the couplings to the face
 * neighbors of lower and higher color of a distance-1 cell coloring.
 * z6 = M^-1 x6 (M^-T x6 when transpose != 0) if both are given.  Aoff
 * (N x 4 x nv x nv: per cell, its faces in ascending face index, row-major
 * coupling block dR_c/dU_nbr, zero where the face has no neighbor), nbr
 * (N x 4 neighbor cell or -1) and color (N) may be NULL. */
int cfd_block_sgs(cfd_ctx* ctx, const double* U6, int precision, int transpose, const double* x6, double* z6,
                  double* Aoff, int* nbr, int* color);
/* Line-implicit symmetric block Gauss-Seidel (no reference counterpart): the
 * same first-order blocks, with the cells of each line segment
 * (cfd_mesh_get_lines) solved together by a block-tridiagonal elimination and
 * the segments swept color by color, ascending then descending.
 * z6 = M^-1 x6 at U6 (for inspection; GMRES uses it with precond = 3). */
int cfd_line_precond(cfd_ctx* ctx, const double* U6, int precision, const double* x6, double* z6);
/* Solver::norm_L2, dot_product, norm_L2_VolumeWeighted — Solver.cpp:1250-1303 */
int cfd_norm_l2(cfd_ctx* ctx, const double* v6, double* out);
int cfd_dot(cfd_ctx* ctx, const double* a6, const double* b6, double* out);
int cfd_norm_vw(cfd_ctx* ctx, const double* R6, double* out);

/* GMRES options. */
typedef struct cfd_gmres_opts {
  int m;            /* Krylov dimension, 1..250 (reference: 40, Solver.cpp:1479)  */
  int restarts;     /* extra restart cycles (reference: 0)                         */
  double tol;       /* relative tolerance on |g_{j+1}| (reference: 1e-2, 1587)     */
  int jv;           /* 0 finite difference (reference), 1 tangent, 2 transpose:
                       solve J^T x = RHS with the reverse product (adjoint)      */
  int precision;    /* 64: FP64 basis/matvec; 32: FP32 basis, matvec, precond
                       (jv 1 and 2; with jv 2 the reverse product itself runs in
                       FP64 between FP32 basis vectors)                         */
  int precond;      /* 0 scalar diagonal (reference, 1424-1465), 1 block-Jacobi,   *
                     * 2 multicolor block symmetric Gauss-Seidel, 3 line-implicit *
                     * symmetric Gauss-Seidel (jv 0 and 1 only)                    */
  int scaling;      /* 0: unscaled 6-component inner product (reference);
                       1: rows scaled by the freestream state, so every equation
                          weighs alike in the Krylov norm (no reference counterpart) */
} cfd_gmres_opts;
int cfd_gmres_opts_default(cfd_gmres_opts* o);

/* Solver::solveGMRES(U_base, R_spatial_base, RHS, x) — Solver.cpp:1472-1618.
 * Uses dt from the last cfd_local_timestep.  iters_out = k + 1. */
int cfd_gmres(cfd_ctx* ctx, const double* U6, const double* R6, const double* RHS6, double* x6,
              const cfd_gmres_opts* opts, int* iters_out);

/* Newton / pseudo-transient continuation options (Solver.cpp:1696-1799). */
typedef struct cfd_newton_opts {
  double tolerance;     /* absolute VM-norm tolerance (reference)                 */
  double rel_drop;      /* stop when VM_k / VM_0 <= rel_drop (0 = off)             */
  int max_steps;
  int log;              /* 1: one JSON line per step to stdout; k > 1: one line
                           every k steps to stderr (progress of long solves)     */
  double cfl0;          /* 0.1                                                     */
  cfd_gmres_opts gmres;
  int controller;       /* 0: the reference CFL controller (Solver.cpp:1747-1796);
                           1: switched evolution relaxation, CFL = cfl0 (VM_0/VM_k)^1
                              clamped to [cfl_min, cfl_max] (no reference counterpart);
                           2: adaptive: after a step that lowered VM by 1 % or more,
                              x2 (GMRES at its tolerance) or x1.25 (GMRES missed it);
                              after a smaller change up to a 20 % rise, x1.1, or x0.7
                              when GMRES missed its tolerance; x0.5 after a larger rise;
                              a rise above 10x (or NaN) is rejected: the state is
                              restored and CFL x0.1; CFL stays in [0.01 cfl0, cfl_max]
                              (no reference counterpart) */
  int pad;
  double cfl_max;       /* SER upper bound (0 = 1e12)                              */
} cfd_newton_opts;
int cfd_newton_opts_default(cfd_newton_opts* o);

typedef struct cfd_newton_result {
  int steps;
  int converged;
  double res_vm0;
  double res_vm;
  double seconds;
  int total_gmres_iters;
  int pad;
} cfd_newton_result;

/* Solver::solveImplicit(tolerance, maxNewtonstep) on U6 (in/out).
 * hist_vm / hist_iters / hist_cfl (each max_steps long, may be NULL)
 * receive the per-step VM norm, GMRES iteration count and CFL. */
int cfd_solve_implicit(cfd_ctx* ctx, double* U6, const cfd_newton_opts* opts,
                       cfd_newton_result* res, double* hist_vm, int* hist_iters,
                       double* hist_cfl);

/* Solver::solveUnsteady(dt_phys, maxTimeSteps) — Solver.cpp:1801-1864: BDF1
 * dual time stepping.  The Jacobian's time term becomes vol / dt_phys; per
 * physical step, Newton on F = (U - U_old) vol / dt_phys + R(U) with GMRES
 * (gopts, NULL = the reference's FD Jv + scalar diagonal, m = 40, tol 1e-2),
 * stopping when VM||F|| over rho, rho u, rho v, rho E < tolerance after at
 * least one update, U += relax dU without clips.  max_newton <= 0, tolerance
 * <= 0 and relax <= 0 select the reference's 15, 1e-6 and 0.7; max_newton <= 0
 * with a non-NULL hist_vm is an error (CFD_ERR_ARG), since the rows of hist_vm
 * are max_newton long.  hist_vm (n_steps x max_newton, may be NULL) receives
 * VM||F|| per Newton step, hist_newton (n_steps, may be NULL) the updates
 * taken per physical step. */
int cfd_solve_unsteady(cfd_ctx* ctx, double* U6, double dt_phys, int n_steps, const cfd_gmres_opts* gopts,
                       int max_newton, double tolerance, double relax, double* hist_vm, int* hist_newton);
/* Solver::timeStepExplicit(dt) — Solver.cpp:1219-1243 (forward Euler). */
int cfd_explicit_step(cfd_ctx* ctx, double* U6, double dt);
/* SSP-RK3 (Shu-Osher) built from three residual calls (no reference counterpart). */
int cfd_rk3_step(cfd_ctx* ctx, double* U6, double dt);

/* Wall outputs: Cp per wall face (tags 3 and 4), Cf per tag-4 face
 * (Solver::extractSkinFriction, Solver.cpp:1899-1979), both in reference face
 * order (0 elsewhere), and J = sum over tag-3 faces of p_owner * area, the
 * wall pressure integral used as the design functional (no reference
 * counterpart; calculateDrag below is the reference's projected form). */
int cfd_wall_outputs(cfd_ctx* ctx, const double* U6, double* Cp, double* Cf, double* J);
/* Solver::calculateDrag (Solver.cpp:1307-1325) at U6: the sum over the
 * tag-3 faces, in face order, of -p_owner * area * n_x.  The reference
 * prints it; here it is returned. */
int cfd_drag(cfd_ctx* ctx, const double* U6, double* drag);
/* Solver::extractWallPressure (Solver.cpp:2027-2070) at U6: one row
 * (x_mid * x_scale, p_owner / p_divisor) per tag-3 face whose owner centroid
 * has y_c < y_max, ascending in x (rows of equal x by pressure).  The
 * reference writes the rows to a CSV file with y_max = 0.2, x_scale = 12
 * (ft -> in) and p_divisor = 144 (psf -> psi); here they fill rows2 (up to
 * `capacity` rows) and *n_rows receives their number (an error if it
 * exceeds the capacity). */
int cfd_wall_pressure(cfd_ctx* ctx, const double* U6, double y_max, double x_scale, double p_divisor, int capacity,
                      double* rows2, int* n_rows);

/* Solver::calculateEntropyError (Solver.cpp:1620-1646): the volume-weighted L2
 * norm of (p / rho^gamma - S_ref) / S_ref over the cells with p, rho > 0, with
 * S_ref from the freestream implied by the total conditions and M_inf.  The
 * reference prints it; here it is returned. */
int cfd_entropy_error(cfd_ctx* ctx, const double* U6, double* err);
/* Solver::extractBoundaryLayerProfile (Solver.cpp:1866-1897): one row
 * (y_c, u / u_inf, v / u_inf, eta) per cell with |x_c - x_target| < 0.005, in
 * reference cell order; eta = y_c sqrt(u_inf / (nu max(x_target, 1e-6))).
 * The reference writes the rows to a CSV file; here they fill rows4 (up to
 * `capacity` rows) and *n_rows receives their number (an error if it exceeds
 * the capacity). */
int cfd_bl_profile(cfd_ctx* ctx, const double* U6, double x_target, int capacity, double* rows4, int* n_rows);

/* ---- device-handle variants (internal SoA layout, RCM order, async) ----- */
/* Reference-order AoS host/device buffer <-> internal SoA device vector. */
int cfd_dev_pack(cfd_ctx* ctx, const double* U6_dev, double* Usoa_dev);
int cfd_dev_unpack(cfd_ctx* ctx, const double* Usoa_dev, double* U6_dev);
int cfd_dev_residual(cfd_ctx* ctx, const double* U, double* R, int update_limiters);
int cfd_dev_jv_fd(cfd_ctx* ctx, const double* U, const double* Rb, const double* v, double* Jv);
int cfd_dev_local_timestep(cfd_ctx* ctx, const double* U, double cfl);
/* v and Jv are FP64 (precision 64) or FP32 (precision 32) SoA vectors. */
int cfd_dev_jv_tangent(cfd_ctx* ctx, const double* U, const void* v, void* Jv, int precision);
int cfd_dev_vjp_adjoint(cfd_ctx* ctx, const double* U, const double* w, double* JTw);
/* cfd_gmres on device SoA vectors (x = solution); returns when the solve has
 * finished.  rhs and x must not alias U or R. */
int cfd_dev_gmres(cfd_ctx* ctx, const double* U, const double* R, const double* rhs, double* x,
                  const cfd_gmres_opts* opts, int* iters_out);
int cfd_dev_synchronize(cfd_ctx* ctx);
/* cfd_norm_vw (Solver::norm_L2_VolumeWeighted, Solver.cpp:1280-1303) of the
 * device SoA vector R, written to out_dev[0] (device memory) on the context
 * stream, allocating nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops).  Its scratch (the reduction's first-stage partial sums and totals,
 * 16.1 KiB) is a static device array the library holds on each GPU until the
 * process ends; all calls on one GPU share it, so calls that overlap on one
 * GPU (several streams or threads) race.  The sums are the same as
 * cfd_norm_vw's (same grid, same order).  Vectors of nvar x pitch beyond
 * 2^31 - 1 elements are rejected (CFD_ERR_ARG). */
int cfd_dev_norm_vw_noalloc(cfd_ctx* ctx, const double* R, double* out_dev);

/* ---- measurement -------------------------------------------------------- */
/* Runs `reps` residual evaluations on device vectors U -> R and records HIP
 * events on the context stream around each stage.  ms_out[0..3] receive the
 * average milliseconds per residual of: gradients+limiter, face flux (all
 * colors), assembly (fill or gather), and the whole residual.  launches_out
 * (may be NULL) receives the number of face-flux launches per residual. */
int cfd_dev_time_residual(cfd_ctx* ctx, const double* U, double* R, int reps, int update_limiters,
                          double* ms_out, int* launches_out);

#ifdef __cplusplus
}
#endif
#endif /* CFD_GPU_H */

/* ORACLE — test infrastructure only (see cfd_oracle.cpp).
 * C entry points of the CPU restatement, loaded by tests/ through ctypes.
 * State vectors are AoS, 6 doubles per cell, reference cell order. */
#ifndef CFD_ORACLE_H
#define CFD_ORACLE_H
#include "../include/cfd_gpu.h"
#ifdef __cplusplus
extern "C" {
#endif
typedef struct orc_mesh orc_mesh;
typedef struct orc_solver orc_solver;

orc_mesh* orc_mesh_from_desc(const cfd_mesh_desc* d);
orc_mesh* orc_mesh_read_gmsh(const char* path);
void orc_mesh_destroy(orc_mesh* m);
void orc_mesh_sizes(const orc_mesh* m, int* nn, int* nc, int* nf);
void orc_mesh_faces(const orc_mesh* m, int* n1, int* n2, int* own, int* nei, int* tag, double* nx, double* ny,
                    double* area, double* xm, double* ym);
void orc_mesh_cells(const orc_mesh* m, double* vol, double* xc, double* yc);

orc_solver* orc_solver_create(const orc_mesh* m, const cfd_params* p);
void orc_solver_destroy(orc_solver* s);
void orc_set_params(orc_solver* s, const cfd_params* p);
void orc_set_face_flux(orc_solver* s, const unsigned char* alt);
void orc_wall_distance(orc_solver* s, double* d);
void orc_residual(orc_solver* s, const double* U6, double* R6, int update);
/* extension (measurement): FP64 operation counts of one residual (add, mul,
 * div, sqrt, trans; then the same for its cell stage) */
void orc_residual_opcount(orc_solver* s, const double* U6, int update, unsigned long long* counts, double* R6);
void orc_residual_cs(orc_solver* s, const double* U6, const double* v6, double h, double* Rre, double* dR,
                     int update);
void orc_face_flux(orc_solver* s, const double* U6, double* F6, int update);
void orc_get_gradients(orc_solver* s, double* g10);
void orc_get_limiters(orc_solver* s, double* phi);
void orc_set_limiters(orc_solver* s, const double* phi);
void orc_local_timestep(orc_solver* s, const double* U6, double cfl);
void orc_get_dt(orc_solver* s, double* dt);
void orc_set_dt(orc_solver* s, const double* dt);
void orc_diag_precond(orc_solver* s, const double* U6, double* D);
void orc_jv_fd(orc_solver* s, const double* U6, const double* R6, const double* v6, double* Jv6);
double orc_norm_l2(orc_solver* s, const double* v6);
double orc_dot(orc_solver* s, const double* a6, const double* b6);
double orc_norm_vw(orc_solver* s, const double* R6);
int orc_gmres(orc_solver* s, const double* U6, const double* R6, const double* RHS6, double* x6, int m, double tol);
int orc_solve_implicit(orc_solver* s, double* U6, double tol, int maxsteps, double rel_drop, int m, double* hvm,
                       int* hit, double* hcfl, int* converged);
void orc_explicit_step(orc_solver* s, double* U6, double dt);
void orc_solve_unsteady(orc_solver* s, double* U6, double dtp, int nsteps, double* hvm, int* hnewton);
double orc_entropy_error(orc_solver* s, const double* U6);
int orc_bl_profile(orc_solver* s, const double* U6, double x_target, double* rows);
void orc_rk3_step(orc_solver* s, double* U6, double dt);
void orc_wall_outputs(orc_solver* s, const double* U6, double* Cp, double* Cf, double* J);
#ifdef __cplusplus
}
#endif
#endif

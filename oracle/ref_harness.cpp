// ORACLE — test infrastructure only.
//
// C entry points around the reference solver itself, compiled from its own
// sources where they lie (Mesh.cpp, Solver.cpp; see oracle/Makefile) into
// oracle/_ref/libref.so.  Used in this build container to pin the CPU
// restatement in cfd_oracle.cpp and to write the fixtures in tests/golden.
// Private members of Solver are reached by compiling this translation unit
// with `private` mapped to `public`; the class layout is unchanged.
#define private public
#include "Mesh.hpp"
#include "Solver.hpp"
#undef private

#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

// main.cpp's VTK writer, compiled from main.cpp with its main() renamed
// (oracle/Makefile); it writes results/sol_<step>.vtk under the working directory.
void saveResults(const Mesh& mesh, const std::vector<State>& U, int step);

extern "C" {

void* ref_mesh_load(const char* path) {
  Mesh* m = new Mesh();
  m->load(path);
  return m;
}
void ref_mesh_destroy(void* m) { delete static_cast<Mesh*>(m); }
void ref_mesh_sizes(void* mp, int* nn, int* nc, int* nf) {
  Mesh* m = static_cast<Mesh*>(mp);
  *nn = m->getNumNodes() - 1;
  *nc = m->getNumCells();
  *nf = m->getNumFaces();
}
void ref_mesh_faces(void* mp, int* n1, int* n2, int* own, int* nei, int* tag, double* nx, double* ny, double* area,
                    double* xm, double* ym) {
  Mesh* m = static_cast<Mesh*>(mp);
  for (size_t f = 0; f < m->faces.size(); ++f) {
    const Face& F = m->faces[f];
    n1[f] = F.node1; n2[f] = F.node2; own[f] = F.owner; nei[f] = F.neighbor; tag[f] = F.bcTag;
    nx[f] = F.nx; ny[f] = F.ny; area[f] = F.area; xm[f] = F.x_mid; ym[f] = F.y_mid;
  }
}
void ref_mesh_cells(void* mp, double* vol, double* xc, double* yc) {
  Mesh* m = static_cast<Mesh*>(mp);
  for (size_t c = 0; c < m->cells.size(); ++c) {
    vol[c] = m->cells[c].volume; xc[c] = m->cells[c].x_c; yc[c] = m->cells[c].y_c;
  }
}

// mode 0: setFlowConditions(M, P, T); mode 1: setFreestreamStatic(M, P, T).
// nu <= 0 keeps the default.  flux / spatial follow Types.hpp:38-49.
void* ref_solver_create(void* mp, int mode, double M, double P, double T, double nu, int flux, int spatial) {
  Solver* s = new Solver(*static_cast<Mesh*>(mp));
  if (mode == 0) s->setFlowConditions(M, P, T);
  else s->setFreestreamStatic(M, P, T);
  if (nu > 0) s->nu = nu;
  s->setNumericalMethod(static_cast<FluxScheme>(flux), static_cast<SpatialScheme>(spatial));
  return s;
}
void ref_solver_destroy(void* s) { delete static_cast<Solver*>(s); }
void ref_set_method(void* sp, int flux, int spatial) {
  static_cast<Solver*>(sp)->setNumericalMethod(static_cast<FluxScheme>(flux), static_cast<SpatialScheme>(spatial));
}
void ref_get_freestream(void* sp, double* out4) {
  Solver* s = static_cast<Solver*>(sp);
  out4[0] = s->M_inf; out4[1] = s->P_total_inlet; out4[2] = s->T_total_inlet; out4[3] = s->P_static_exit;
}

static std::vector<State> to_states(const double* U6, size_t n) {
  std::vector<State> v(n);
  std::memcpy(v.data(), U6, n * sizeof(State));
  return v;
}
static void from_states(const std::vector<State>& v, double* U6) { std::memcpy(U6, v.data(), v.size() * sizeof(State)); }

void ref_initialize(void* sp, double* U6) {
  Solver* s = static_cast<Solver*>(sp);
  s->initialize();
  from_states(s->U_current, U6);
}
void ref_residual(void* sp, const double* U6, double* R6, int update) {
  Solver* s = static_cast<Solver*>(sp);
  size_t n = s->mesh.cells.size();
  std::vector<State> U = to_states(U6, n), R(n);
  s->computeResiduals(U, R, update != 0);
  from_states(R, R6);
}
void ref_get_gradients(void* sp, double* g10) {
  Solver* s = static_cast<Solver*>(sp);
  for (size_t c = 0; c < s->cellGradients.size(); ++c) {
    const CellGradient& g = s->cellGradients[c];
    double v[10] = {g.drhodx, g.drhody, g.dudx, g.dudy, g.dvdx, g.dvdy, g.dpdx, g.dpdy, g.dTdx, g.dTdy};
    std::memcpy(g10 + 10 * c, v, sizeof v);
  }
}
void ref_get_limiters(void* sp, double* phi) {
  Solver* s = static_cast<Solver*>(sp);
  for (size_t c = 0; c < s->limiters.size(); ++c) phi[c] = s->limiters[c];
}
void ref_local_timestep(void* sp, const double* U6, double cfl) {
  Solver* s = static_cast<Solver*>(sp);
  s->U_current = to_states(U6, s->mesh.cells.size());
  s->computeLocalTimeStep(cfl);
}
void ref_get_dt(void* sp, double* dt) {
  Solver* s = static_cast<Solver*>(sp);
  std::memcpy(dt, s->dt_local.data(), s->dt_local.size() * sizeof(double));
}
void ref_diag_precond(void* sp, const double* U6, double* D) {
  Solver* s = static_cast<Solver*>(sp);
  s->U_current = to_states(U6, s->mesh.cells.size());
  std::vector<double> d;
  s->computeDiagonalPreconditioner(d);
  std::memcpy(D, d.data(), d.size() * sizeof(double));
}
void ref_jv(void* sp, const double* U6, const double* R6, const double* v6, double* Jv6) {
  Solver* s = static_cast<Solver*>(sp);
  size_t n = s->mesh.cells.size();
  std::vector<State> U = to_states(U6, n), R = to_states(R6, n), v = to_states(v6, n), Jv(n);
  s->applyJacobian(U, R, v, Jv);
  from_states(Jv, Jv6);
}
int ref_gmres(void* sp, const double* U6, const double* R6, const double* RHS6, double* x6) {
  Solver* s = static_cast<Solver*>(sp);
  size_t n = s->mesh.cells.size();
  s->U_current = to_states(U6, n);
  std::vector<State> U = to_states(U6, n), R = to_states(R6, n), b = to_states(RHS6, n), x(n);
  int it = s->solveGMRES(U, R, b, x);
  from_states(x, x6);
  return it;
}
double ref_norm_l2(void* sp, const double* v6) {
  Solver* s = static_cast<Solver*>(sp);
  return s->norm_L2(to_states(v6, s->mesh.cells.size()));
}
double ref_norm_vw(void* sp, const double* v6) {
  Solver* s = static_cast<Solver*>(sp);
  return s->norm_L2_VolumeWeighted(to_states(v6, s->mesh.cells.size()));
}
// Solver::solveImplicit on U6 (in/out).  The reference prints one line per
// step and writes results/residual_history.csv when that folder exists.
void ref_solve_implicit(void* sp, double* U6, double tol, int maxsteps) {
  Solver* s = static_cast<Solver*>(sp);
  s->U_current = to_states(U6, s->mesh.cells.size());
  s->solveImplicit(tol, maxsteps);
  from_states(s->U_current, U6);
}
void ref_explicit_step(void* sp, double* U6, double dt) {
  Solver* s = static_cast<Solver*>(sp);
  s->U_current = to_states(U6, s->mesh.cells.size());
  s->timeStepExplicit(dt);
  from_states(s->U_current, U6);
}
// Solver::solveUnsteady on U6 (in/out); the reference prints per physical step.
void ref_solve_unsteady(void* sp, double* U6, double dt_phys, int nsteps) {
  Solver* s = static_cast<Solver*>(sp);
  s->U_current = to_states(U6, s->mesh.cells.size());
  s->solveUnsteady(dt_phys, nsteps);
  from_states(s->U_current, U6);
}
// Solver::calculateEntropyError prints its result; the value is read back
// from what it prints (default stream precision, 6 significant digits).
double ref_entropy_error(void* sp, const double* U6) {
  Solver* s = static_cast<Solver*>(sp);
  s->U_current = to_states(U6, s->mesh.cells.size());
  std::ostringstream cap;
  std::streambuf* old = std::cout.rdbuf(cap.rdbuf());
  s->calculateEntropyError();
  std::cout.rdbuf(old);
  const std::string t = cap.str();
  const std::string key = "L2 Entropy Error:";
  const size_t k = t.find(key);
  return k == std::string::npos ? -1.0 : std::strtod(t.c_str() + k + key.size(), nullptr);
}
// Solver::extractBoundaryLayerProfile(x, path) on U6
void ref_bl_profile(void* sp, const double* U6, double x_target, const char* path) {
  Solver* s = static_cast<Solver*>(sp);
  s->U_current = to_states(U6, s->mesh.cells.size());
  std::ostringstream cap;
  std::streambuf* old = std::cout.rdbuf(cap.rdbuf());
  s->extractBoundaryLayerProfile(x_target, path);
  std::cout.rdbuf(old);
}
// main.cpp saveResults(mesh, U6, step)
void ref_save_vtk(void* mp, const double* U6, int step) {
  const Mesh* m = static_cast<Mesh*>(mp);
  saveResults(*m, to_states(U6, m->cells.size()), step);
}
// Solver::extractSkinFriction(path) on U6: the reference writes x, Cf and
// the Blasius value per no-slip face to a CSV file (6 significant digits).
void ref_skin_friction(void* sp, const double* U6, const char* path) {
  Solver* s = static_cast<Solver*>(sp);
  s->U_current = to_states(U6, s->mesh.cells.size());
  std::ostringstream cap;
  std::streambuf* old = std::cout.rdbuf(cap.rdbuf());
  s->extractSkinFriction(path);
  std::cout.rdbuf(old);
}
// Solver::extractWallPressure(path) on U6 (CSV rows x_in, p_psi).
void ref_wall_pressure(void* sp, const double* U6, const char* path) {
  Solver* s = static_cast<Solver*>(sp);
  s->U_current = to_states(U6, s->mesh.cells.size());
  std::ostringstream cap;
  std::streambuf* old = std::cout.rdbuf(cap.rdbuf());
  s->extractWallPressure(path);
  std::cout.rdbuf(old);
}
// Solver::calculateDrag on U6 prints " Drag :<value>"; the value is read
// back from what it prints.  The stream precision is raised to 17 digits
// around the call, so the printed value is the double the reference summed.
double ref_drag(void* sp, const double* U6) {
  Solver* s = static_cast<Solver*>(sp);
  s->U_current = to_states(U6, s->mesh.cells.size());
  std::ostringstream cap;
  cap.precision(17);
  std::streambuf* old = std::cout.rdbuf(cap.rdbuf());
  const std::streamsize prec = std::cout.precision(17);
  s->calculateDrag();
  std::cout.precision(prec);
  std::cout.rdbuf(old);
  const std::string t = cap.str();
  const std::string key = "Drag :";
  const size_t k = t.find(key);
  return k == std::string::npos ? -1.0 : std::strtod(t.c_str() + k + key.size(), nullptr);
}
void ref_set_nu(void* sp, double nu) { static_cast<Solver*>(sp)->nu = nu; }
void ref_kinematic_viscosity(void* sp, double Re, double L) { static_cast<Solver*>(sp)->setKinematicViscosity(Re, L); }
double ref_get_nu(void* sp) { return static_cast<Solver*>(sp)->nu; }

}  // extern "C"

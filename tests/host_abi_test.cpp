// A C++ host that calls the library the way the reference's Solver would
// after the edits shown in the integration notes: only include/cfd_gpu.h,
// a std::vector of a 48-byte State (Types.hpp:55-82), host pointers.
//
// Usage: host_abi_test OUT_DIR
//   Builds the laminar flat plate (configuration 2) at 0.1 scale, sets the
//   flow conditions and scheme as Solver::setFlowConditions /
//   setKinematicViscosity / setNumericalMethod do, perturbs the initial state
//   by a fixed pattern and writes, as raw little-endian doubles (N x 6 each):
//     U.bin   the state passed in
//     R.bin   Solver::computeResiduals  -> cfd_residual
//     Jv.bin  Solver::applyJacobian     -> cfd_jv_fd, after
//             Solver::computeLocalTimeStep(2.0) -> cfd_local_timestep
//   and prints one line "host_abi_test ok N=<cells>".  Exit status 0 on
//   success, 1 with the library's message otherwise.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../include/cfd_gpu.h"

struct State {  // Types.hpp:55-82, field for field
  double rho, rhou, rhov, rhoE, rhok, rhow;
};
static_assert(sizeof(State) == 48, "the C-ABI takes std::vector<State> memory with a 48-byte stride");

static int check(int rc, cfd_ctx* ctx, const char* what) {
  if (rc != CFD_OK) {
    std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, ctx ? cfd_last_error(ctx) : "");
    std::exit(1);
  }
  return rc;
}

static void write(const std::string& path, const std::vector<State>& v) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f || std::fwrite(v.data(), sizeof(State), v.size(), f) != v.size()) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(1);
  }
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: %s OUT_DIR\n", argv[0]);
    return 1;
  }
  const std::string out = argv[1];
  cfd_mesh* mesh = nullptr;
  check(cfd_mesh_generate(2, 0.1, &mesh), nullptr, "cfd_mesh_generate");
  int nn = 0, nc = 0, nf = 0, ncol = 0;
  check(cfd_mesh_sizes(mesh, &nn, &nc, &nf, &ncol), nullptr, "cfd_mesh_sizes");

  cfd_params p;
  cfd_params_default(&p);
  cfd_params_flow_conditions(&p, 0.2, 101325.0, 300.0);  // Solver::setFlowConditions
  cfd_params_kinematic_viscosity(&p, 1e5, 1.0);          // Solver::setKinematicViscosity
  p.flux = CFD_FLUX_ROE;                                  // Solver::setNumericalMethod
  p.spatial = CFD_SECOND_ORDER_CENTRAL;
  p.bcset = CFD_BC_PLATE;
  p.viscous = 1;
  p.nvar = 4;

  cfd_ctx* ctx = nullptr;
  check(cfd_ctx_create(mesh, &p, &ctx), ctx, "cfd_ctx_create");

  // freestream of the plate set's frozen inflow, with a fixed perturbation
  State u0{};
  check(cfd_params_initial_state(&p, &u0.rho), ctx, "cfd_params_initial_state");
  std::vector<State> U(nc, u0), R(nc), v(nc), Jv(nc);
  for (int i = 0; i < nc; ++i) {
    const double s = 1.0 + 1e-3 * ((i * 7919) % 101 - 50) / 50.0;
    U[i].rho *= s;
    U[i].rhou *= s;
    U[i].rhov = 1e-3 * U[i].rhou * (((i * 104729) % 13) - 6) / 6.0;
    U[i].rhoE *= 1.0 + 2e-4 * ((i * 31) % 17 - 8) / 8.0;
    v[i] = State{1e-3 * U[i].rho, -2e-3 * U[i].rhou, 3e-3 * U[i].rhou, 1e-3 * U[i].rhoE, 0.0, 0.0};
  }
  check(cfd_residual(ctx, &U[0].rho, &R[0].rho, 1), ctx, "cfd_residual");
  check(cfd_local_timestep(ctx, &U[0].rho, 2.0), ctx, "cfd_local_timestep");
  check(cfd_jv_fd(ctx, &U[0].rho, &R[0].rho, &v[0].rho, &Jv[0].rho), ctx, "cfd_jv_fd");
  write(out + "/U.bin", U);
  write(out + "/R.bin", R);
  write(out + "/Jv.bin", Jv);
  cfd_ctx_destroy(ctx);
  cfd_mesh_destroy(mesh);
  std::printf("host_abi_test ok N=%d\n", nc);
  return 0;
}

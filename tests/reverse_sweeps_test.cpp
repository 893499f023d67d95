// Host check of the hand-written reverse sweeps used by the structured
// reverse product (mycode_amd/csrc/riemann_vjp.hpp, visc.hpp, physics.hpp:
// turb_cell_vjp) against forward-mode duals of the functions they transpose,
// on seeded random inputs chosen to take every branch (sub- and supersonic
// faces both ways, the HLLC star sides, the Roe entropy fix, the LLF wave
// choice, the SST min / max switches).  For each input x_j it compares the
// sweep's adjoint with sum_k b_k dF_k/dx_j from a one-direction dual; the
// error is taken relative to the largest such component (x_j-scaled for the
// SST terms, whose inputs span many decades).  Exit status 0 when every case
// is within 1e-11.
#include <cmath>
#include <cstdio>
#include <random>

#include "../mycode_amd/csrc/riemann_vjp.hpp"
#include "../mycode_amd/csrc/visc.hpp"

using namespace cfd;
typedef DualV<double, 1> D1;

static Phys gas() {
  Phys P{};
  P.gamma = 1.4;
  P.R = 287.0;
  P.nu = 1.5e-5;
  P.Pr = 0.72;
  P.Cp = P.gamma * P.R / (P.gamma - 1.0);
  return P;
}

static double riemann_case(std::mt19937_64& rng, int scheme) {
  std::uniform_real_distribution<double> u01(0.0, 1.0), s11(-1.0, 1.0);
  const Phys P = gas();
  const double th = 3.2 * s11(rng), nx = cos(th), ny = sin(th);
  const double rho = 0.5 + u01(rng), p = 1e5 * (0.5 + u01(rng)), c = sqrt(1.4 * p / rho);
  double mach = 3.0 * s11(rng), tang = 0.3 * s11(rng);
  if (u01(rng) < 0.2) mach = 1.0 + 0.02 * s11(rng);  // near the sonic point (Roe's entropy fix)
  const double u = c * (mach * nx - tang * ny), v = c * (mach * ny + tang * nx);
  const S6<double> L = prim2cons(P, rho, u, v, p);
  const S6<double> R = prim2cons(P, rho * (1.0 + 0.2 * s11(rng)), u + 0.2 * c * s11(rng), v + 0.2 * c * s11(rng),
                                 p * (1.0 + 0.3 * s11(rng)));
  double Fb[4];
  for (double& x : Fb) x = s11(rng);
  double Lb[4] = {0, 0, 0, 0}, Rb[4] = {0, 0, 0, 0};
  riemann_vjp(P, L, R, nx, ny, scheme, Fb, Lb, Rb);
  double fwd[8], big = 0.0, err = 0.0;
  for (int j = 0; j < 8; ++j) {
    S6<D1> Ld, Rd;
    for (int k = 0; k < 6; ++k) {
      Ld[k] = D1(L[k]);
      Ld[k].d[0] = j == k ? 1.0 : 0.0;
      Rd[k] = D1(R[k]);
      Rd[k].d[0] = j == 4 + k ? 1.0 : 0.0;
    }
    const S6<D1> F = riemann(P, Ld, Rd, nx, ny, scheme);
    fwd[j] = 0.0;
    for (int k = 0; k < 4; ++k) fwd[j] += Fb[k] * F[k].d[0];
    big = fmax(big, fabs(fwd[j]));
  }
  for (int j = 0; j < 8; ++j) err = fmax(err, fabs(fwd[j] - (j < 4 ? Lb[j] : Rb[j - 4])) / big);
  return err;
}

template <bool TURB>
static double visc_case(std::mt19937_64& rng) {
  std::uniform_real_distribution<double> u01(0.5, 1.5), s11(-1.0, 1.0);
  const Phys P = gas();
  FaceGeo g{};
  const double th = 3.0 * s11(rng);
  g.nx = cos(th);
  g.ny = sin(th);
  g.A = u01(rng);
  g.xco = s11(rng);
  g.yco = s11(rng);
  g.xcn = g.xco + s11(rng);
  g.ycn = g.yco + s11(rng);
  S6<double> L, R;
  for (int k = 0; k < 6; ++k) {
    L[k] = u01(rng) * (k == 3 ? 2e5 : 1.0) + ((k == 1 || k == 2) ? 50.0 * s11(rng) : 0.0);
    R[k] = L[k] * (1.0 + 0.1 * s11(rng));
  }
  double gb[6], tq[4], Vb[6];
  for (double& x : gb) x = 100.0 * s11(rng);
  for (int q = 0; q < 4; ++q) tq[q] = (q < 2 ? 1e-4 : 0.5) * u01(rng);
  for (double& x : Vb) x = s11(rng);
  double Lb[6], Rb[6], gbb[6], tqb[4] = {0, 0, 0, 0};
  visc_interior_vjp<TURB>(P, g, L, R, gb, tq, Vb, Lb, Rb, gbb, tqb);
  const int nv = TURB ? 6 : 4;
  double fwd[22], big = 0.0, err = 0.0;
  for (int j = 0; j < 22; ++j) {
    S6<D1> Ld, Rd;
    D1 gd[6], td[4];
    for (int k = 0; k < 6; ++k) {
      Ld[k] = D1(L[k]);
      Ld[k].d[0] = j == k ? 1.0 : 0.0;
      Rd[k] = D1(R[k]);
      Rd[k].d[0] = j == 6 + k ? 1.0 : 0.0;
    }
    for (int q = 0; q < 6; ++q) {
      gd[q] = D1(gb[q]);
      gd[q].d[0] = j == 12 + q ? 1.0 : 0.0;
    }
    for (int q = 0; q < 4; ++q) {
      td[q] = D1(tq[q]);
      td[q].d[0] = j == 18 + q ? 1.0 : 0.0;
    }
    const S6<D1> V = visc_interior<TURB>(P, g, Ld, Rd, gd, td);
    fwd[j] = 0.0;
    for (int k = 1; k < nv; ++k) fwd[j] += Vb[k] * V[k].d[0];
    big = fmax(big, fabs(fwd[j]));
  }
  for (int j = 0; j < 22; ++j) {
    const bool used = TURB || !((j >= 4 && j < 6) || (j >= 10 && j < 12) || j >= 18);
    if (!used) continue;
    const double rev = j < 6 ? Lb[j] : j < 12 ? Rb[j - 6] : j < 18 ? gbb[j - 12] : tqb[j - 18];
    err = fmax(err, fabs(fwd[j] - rev) / big);
  }
  return err;
}

static double turb_case(std::mt19937_64& rng) {
  std::uniform_real_distribution<double> u01(0.0, 1.0), s11(-1.0, 1.0);
  const Phys P = gas();
  const double rho = 0.8 + 0.4 * u01(rng);
  const double k = pow(10.0, -6.0 + 5.0 * u01(rng)), w = pow(10.0, 1.0 + 8.0 * u01(rng));
  const S6<double> U(rho, rho * 60.0 * s11(rng), rho * 5.0 * s11(rng), 2.5e5, rho * k, rho * w);
  double g[14];
  for (double& x : g) x = s11(rng) * pow(10.0, 1.0 + 5.0 * u01(rng));
  const double d = pow(10.0, -6.0 + 5.0 * u01(rng));
  const double lk = s11(rng), lw = 1e-3 * s11(rng), am = 1e3 * s11(rng), aF = s11(rng);
  double Ub[6] = {0, 0, 0, 0, 0, 0}, gbar[14] = {0};
  turb_cell_vjp(P, U, g, d, lk, lw, am, aF, Ub, gbar);
  double fwd[20], big = 0.0, err = 0.0;
  for (int j = 0; j < 20; ++j) {
    S6<D1> Ud;
    D1 gd[14];
    for (int q = 0; q < 6; ++q) {
      Ud[q] = D1(U[q]);
      Ud[q].d[0] = j == q ? 1.0 : 0.0;
    }
    for (int q = 0; q < 14; ++q) {
      gd[q] = D1(g[q]);
      gd[q].d[0] = j == 6 + q ? 1.0 : 0.0;
    }
    const TurbQ<D1> t = turb_cell(P, Ud, gd, d);
    fwd[j] = lk * t.Sk.d[0] + lw * t.Sw.d[0] + am * t.mut.d[0] + aF * t.F1.d[0];
    big = fmax(big, fabs(fwd[j]) * (j < 6 ? fabs(U[j]) : fabs(g[j - 6])));
  }
  for (int j = 0; j < 20; ++j) {
    const double size = j < 6 ? fabs(U[j]) : fabs(g[j - 6]);
    err = fmax(err, fabs(fwd[j] - (j < 6 ? Ub[j] : gbar[j - 6])) * size / big);
  }
  return err;
}

int main() {
  std::mt19937_64 rng(20261204);
  const char* names[] = {"LLF", "Roe", "HLLC"};
  const int schemes[] = {CFD_FLUX_LLF, CFD_FLUX_ROE, CFD_FLUX_HLLC};
  int fails = 0;
  for (int s = 0; s < 3; ++s) {
    double worst = 0.0;
    for (int t = 0; t < 4000; ++t) worst = fmax(worst, riemann_case(rng, schemes[s]));
    printf("%s flux: worst %.3e\n", names[s], worst);
    fails += worst > 1e-11;
  }
  double wl = 0.0, wt = 0.0, ws = 0.0;
  for (int t = 0; t < 2000; ++t) {
    wl = fmax(wl, visc_case<false>(rng));
    wt = fmax(wt, visc_case<true>(rng));
  }
  for (int t = 0; t < 20000; ++t) ws = fmax(ws, turb_case(rng));
  printf("viscous terms: worst %.3e (laminar), %.3e (SST)\nSST cell terms: worst %.3e\n", wl, wt, ws);
  fails += (wl > 1e-11) + (wt > 1e-11) + (ws > 1e-11);
  return fails ? 1 : 0;
}

// Viscous terms of an interior face as the structured reverse product sees
// them: the forward function and its hand-written reverse sweep.  Host- and
// device-callable (HD), so a host program can check one against the other.
#pragma once

#include "physics.hpp"

namespace cfd {

// Geometry and frozen quantities of one face, as the flux core sees them.
struct FaceGeo {
  double nx, ny, A, xm, ym, xco, yco, xcn, ycn, phL, phR;
  double wdo, wdn, dnw;  // SST: wall distances of the two cells, owner normal distance
  int tg, scheme;
  bool boundary;
};

// The viscous terms flux_core subtracts on an interior face, written as a
// function of the two cell states, the six averaged gradient components
// (u_x, u_y, v_x, v_y, T_x, T_y), which is how they depend on the gradients
// there, and with SST (TURB) of the two cells' eddy viscosity and F1
// (tq = mu_t,o, mu_t,n, F1_o, F1_n): the same operations as flux_core's
// viscous block.  The structured adjoint differentiates it in these inputs.
template <bool TURB, class T>
HD S6<T> visc_interior(const Phys& P, const FaceGeo& g, const S6<T>& L, const S6<T>& R,
                                               const T (&gb)[6], const T* tq) {
  const double nx = g.nx, ny = g.ny;
  const T rf = 0.5 * (L[0] + R[0]);
  const T uf = 0.5 * (dv(L[1], L[0]) + dv(R[1], R[0]));
  const T vf = 0.5 * (dv(L[2], L[0]) + dv(R[2], R[0]));
  T ux = gb[0], uy = gb[1], vx = gb[2], vy = gb[3];
  const T Tx = gb[4], Ty = gb[5];
  const double dxLR = g.xcn - g.xco, dyLR = g.ycn - g.yco;
  const double d2 = dxLR * dxLR + dyLR * dyLR;
  if (d2 > 1e-14) {
    const T du = dv(R[1], R[0]) - dv(L[1], L[0]);
    const T dvv = dv(R[2], R[0]) - dv(L[2], L[0]);
    const T ux0 = ux, uy0 = uy, vx0 = vx, vy0 = vy;
    ux = ux0 - dv((ux0 * dxLR + uy0 * dyLR - du) * dxLR, d2);
    uy = uy0 - dv((ux0 * dxLR + uy0 * dyLR - du) * dyLR, d2);
    vx = vx0 - dv((vx0 * dxLR + vy0 * dyLR - dvv) * dxLR, d2);
    vy = vy0 - dv((vx0 * dxLR + vy0 * dyLR - dvv) * dyLR, d2);
  }
  T mu = rf * P.nu;
  T kc = dv(mu * P.Cp, P.Pr);
  T mut = T(0.0), F1f = T(0.0);
  if constexpr (TURB) {
    mut = 0.5 * (tq[0] + tq[1]);
    F1f = 0.5 * (tq[2] + tq[3]);
    kc = P.Cp * (dv(mu, P.Pr) + dv(mut, sst::Prt));
    mu = mu + mut;
  }
  const T div = ux + vy;
  const T txx = mu * (2.0 * ux - (2.0 / 3.0) * div);
  const T tyy = mu * (2.0 * vy - (2.0 / 3.0) * div);
  const T txy = mu * (uy + vx);
  const T qx = -kc * Tx, qy = -kc * Ty;
  const T wx = uf * txx + vf * txy - qx;
  const T wy = uf * txy + vf * tyy - qy;
  S6<T> V(T(0.0), txx * nx + txy * ny, txy * nx + tyy * ny, wx * nx + wy * ny);
  if constexpr (TURB) {  // two-point normal differences of k and omega
    const T kcl = dv(L[4], L[0]), wcl = dv(L[5], L[0]);
    double dn = fabs((g.xcn - g.xco) * nx + (g.ycn - g.yco) * ny);
    if (dn < 1e-300) dn = 1e-300;
    const T kn = dv(dv(R[4], R[0]) - kcl, dn);
    const T wn = dv(dv(R[5], R[0]) - wcl, dn);
    const T mul = rf * P.nu;
    const T sgk = F1f * sst::sk1 + (1.0 - F1f) * sst::sk2, sgw = F1f * sst::sw1 + (1.0 - F1f) * sst::sw2;
    V[4] = (mul + sgk * mut) * kn;
    V[5] = (mul + sgw * mut) * wn;
  }
  return V;
}

// Reverse sweep of visc_interior<TURB> for double inputs: given the adjoints
// Vb[1..5] of its outputs (Vb[0] is unused: the mass flux has no viscous
// part), the adjoints of its inputs, L and R (6 each), the six averaged
// gradient components gb and, with TURB, tq = (mu_t,o, mu_t,n, F1_o, F1_n).
// Each step below undoes one line of the forward function, last line first;
// the intermediate values are recomputed in the same order as there.
template <bool TURB>
HD void visc_interior_vjp(const Phys& P, const FaceGeo& g, const S6<double>& L, const S6<double>& R,
                          const double (&gb)[6], const double* tq, const double (&Vb)[6], double (&Lb)[6],
                          double (&Rb)[6], double (&gbb)[6], double* tqb) {
  const double nx = g.nx, ny = g.ny;
  // ---- forward, keeping what the reverse sweep reads
  const double rf = 0.5 * (L[0] + R[0]);
  const double uL = dv(L[1], L[0]), uR = dv(R[1], R[0]), vL = dv(L[2], L[0]), vR = dv(R[2], R[0]);
  const double uf = 0.5 * (uL + uR);
  const double vf = 0.5 * (vL + vR);
  double ux = gb[0], uy = gb[1], vx = gb[2], vy = gb[3];
  const double Tx = gb[4], Ty = gb[5];
  const double dxLR = g.xcn - g.xco, dyLR = g.ycn - g.yco;
  const double d2 = dxLR * dxLR + dyLR * dyLR;
  const bool corr = d2 > 1e-14;
  if (corr) {
    const double du = uR - uL, dvv = vR - vL;
    const double eu = ux * dxLR + uy * dyLR - du, ev = vx * dxLR + vy * dyLR - dvv;
    const double ux0 = ux, uy0 = uy, vx0 = vx, vy0 = vy;
    ux = ux0 - dv(eu * dxLR, d2);
    uy = uy0 - dv(eu * dyLR, d2);
    vx = vx0 - dv(ev * dxLR, d2);
    vy = vy0 - dv(ev * dyLR, d2);
  }
  const double mu0 = rf * P.nu;
  double mu = mu0, kc = dv(mu0 * P.Cp, P.Pr), mut = 0.0, F1f = 0.0;
  if (TURB) {
    mut = 0.5 * (tq[0] + tq[1]);
    F1f = 0.5 * (tq[2] + tq[3]);
    kc = P.Cp * (dv(mu0, P.Pr) + dv(mut, sst::Prt));
    mu = mu0 + mut;
  }
  const double div = ux + vy;
  const double txx = mu * (2.0 * ux - (2.0 / 3.0) * div);
  const double tyy = mu * (2.0 * vy - (2.0 / 3.0) * div);
  const double txy = mu * (uy + vx);
  // ---- reverse
  double rfb = 0.0, mutb = 0.0, F1b = 0.0, uLb = 0.0, uRb = 0.0, vLb = 0.0, vRb = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) Lb[k] = Rb[k] = 0.0;
  if (TURB) {  // V4 = (mul + sgk mut) kn, V5 = (mul + sgw mut) wn, mul = rf nu
    const double kcl = dv(L[4], L[0]), wcl = dv(L[5], L[0]);
    const double kcr = dv(R[4], R[0]), wcr = dv(R[5], R[0]);
    double dn = fabs((g.xcn - g.xco) * nx + (g.ycn - g.yco) * ny);
    if (dn < 1e-300) dn = 1e-300;
    const double kn = dv(kcr - kcl, dn), wn = dv(wcr - wcl, dn);
    const double mul = rf * P.nu;
    const double sgk = F1f * sst::sk1 + (1.0 - F1f) * sst::sk2, sgw = F1f * sst::sw1 + (1.0 - F1f) * sst::sw2;
    const double knb = Vb[4] * (mul + sgk * mut), wnb = Vb[5] * (mul + sgw * mut);
    const double mulb = Vb[4] * kn + Vb[5] * wn;
    const double sgkb = Vb[4] * mut * kn, sgwb = Vb[5] * mut * wn;
    mutb += Vb[4] * sgk * kn + Vb[5] * sgw * wn;
    F1b += sgkb * (sst::sk1 - sst::sk2) + sgwb * (sst::sw1 - sst::sw2);
    rfb += mulb * P.nu;
    const double kcrb = dv(knb, dn), wcrb = dv(wnb, dn);
    Rb[4] += dv(kcrb, R[0]);
    Rb[5] += dv(wcrb, R[0]);
    Rb[0] -= dv(kcrb * kcr + wcrb * wcr, R[0]);
    Lb[4] -= dv(kcrb, L[0]);
    Lb[5] -= dv(wcrb, L[0]);
    Lb[0] += dv(kcrb * kcl + wcrb * wcl, L[0]);
  }
  // V1 = txx nx + txy ny, V2 = txy nx + tyy ny, V3 = wx nx + wy ny,
  // wx = uf txx + vf txy - qx, wy = uf txy + vf tyy - qy, q = -kc grad T
  const double wxb = Vb[3] * nx, wyb = Vb[3] * ny;
  const double txxb = Vb[1] * nx + wxb * uf;
  const double txyb = Vb[1] * ny + Vb[2] * nx + wxb * vf + wyb * uf;
  const double tyyb = Vb[2] * ny + wyb * vf;
  const double ufb = wxb * txx + wyb * txy, vfb = wxb * txy + wyb * tyy;
  const double kcb = wxb * Tx + wyb * Ty;
  gbb[4] = wxb * kc;
  gbb[5] = wyb * kc;
  // txx = mu (2 ux - 2/3 div), tyy = mu (2 vy - 2/3 div), txy = mu (uy + vx)
  const double mub = txxb * (2.0 * ux - (2.0 / 3.0) * div) + tyyb * (2.0 * vy - (2.0 / 3.0) * div) + txyb * (uy + vx);
  const double divb = -(2.0 / 3.0) * mu * (txxb + tyyb);
  double uxb = 2.0 * mu * txxb + divb, vyb = 2.0 * mu * tyyb + divb;
  double uyb = mu * txyb, vxb = mu * txyb;
  double mu0b;
  if (TURB) {  // mu = mu0 + mut, kc = Cp (mu0 / Pr + mut / Prt)
    mu0b = mub + kcb * dv(P.Cp, P.Pr);
    mutb += mub + kcb * dv(P.Cp, sst::Prt);
    tqb[0] = 0.5 * mutb;
    tqb[1] = 0.5 * mutb;
    tqb[2] = 0.5 * F1b;
    tqb[3] = 0.5 * F1b;
  } else {  // kc = mu0 Cp / Pr
    mu0b = mub + kcb * dv(P.Cp, P.Pr);
  }
  rfb += mu0b * P.nu;
  if (corr) {  // u_x = u_x0 - (e_u dx) / d2, e_u = u_x0 dx + u_y0 dy - (u_R - u_L), same for v
    const double eub = -dv(uxb * dxLR + uyb * dyLR, d2), evb = -dv(vxb * dxLR + vyb * dyLR, d2);
    uxb += eub * dxLR;
    uyb += eub * dyLR;
    vxb += evb * dxLR;
    vyb += evb * dyLR;
    uRb -= eub;
    uLb += eub;
    vRb -= evb;
    vLb += evb;
  }
  gbb[0] = uxb;
  gbb[1] = uyb;
  gbb[2] = vxb;
  gbb[3] = vyb;
  uLb += 0.5 * ufb;
  uRb += 0.5 * ufb;
  vLb += 0.5 * vfb;
  vRb += 0.5 * vfb;
  Lb[1] += dv(uLb, L[0]);
  Lb[2] += dv(vLb, L[0]);
  Lb[0] += 0.5 * rfb - dv(uLb * uL + vLb * vL, L[0]);
  Rb[1] += dv(uRb, R[0]);
  Rb[2] += dv(vRb, R[0]);
  Rb[0] += 0.5 * rfb - dv(uRb * uR + vRb * vR, R[0]);
}

}  // namespace cfd

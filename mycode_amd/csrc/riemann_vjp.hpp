// Reverse sweeps of the Riemann fluxes (physics.hpp: normal_flux, flux_llf,
// flux_roe, flux_hllc) and of prim2cons, for the structured reverse product.
// Each takes the adjoints Fb[0..3] of the flux's mass, momentum and energy
// components and adds the adjoints of the conservative states L, R (4 each)
// to Lb, Rb; the forward values are recomputed in the forward functions' own
// order and every branch is taken on the same comparisons, so the sweep is the
// transpose of the derivative along the branch the residual takes.  The k and
// omega components are not covered (the structured product carries them by
// hand).  Host- and device-callable, so a host program can check each sweep
// against forward duals.
#pragma once

#include "physics.hpp"

namespace cfd {

// p = (gamma-1)(U3 - U0 (U1^2 + U2^2) / (2 U0^2)), u = U1/U0, v = U2/U0:
// adds pb dp/dU + ub du/dU + vb dv/dU to Ub
HD void prim_uvp_vjp(const Phys& P, const S6<double>& U, double u, double v, double pb, double ub, double vb,
                     double (&Ub)[4]) {
  const double gm1 = P.gamma - 1.0;
  Ub[0] += pb * gm1 * 0.5 * (u * u + v * v) - dv(ub * u + vb * v, U[0]);
  Ub[1] += dv(ub, U[0]) - pb * gm1 * u;
  Ub[2] += dv(vb, U[0]) - pb * gm1 * v;
  Ub[3] += pb * gm1;
}

// F = (U0 un, U1 un + p nx, U2 un + p ny, (U3 + p) un)
HD void normal_flux_vjp(const Phys& P, const S6<double>& U, double nx, double ny, const double (&Fb)[4],
                        double (&Ub)[4]) {
  const double p = pressure(P, U);
  const double u = dv(U[1], U[0]), v = dv(U[2], U[0]);
  const double un = u * nx + v * ny;
  const double unb = Fb[0] * U[0] + Fb[1] * U[1] + Fb[2] * U[2] + Fb[3] * (U[3] + p);
  const double pb = Fb[1] * nx + Fb[2] * ny + Fb[3] * un;
  Ub[0] += Fb[0] * un;
  Ub[1] += Fb[1] * un;
  Ub[2] += Fb[2] * un;
  Ub[3] += Fb[3] * un;
  prim_uvp_vjp(P, U, u, v, pb, unb * nx, unb * ny, Ub);
}

// |x| as ab() takes it: -x when x < 0
HD double ab_sign(double x) { return x < 0 ? -1.0 : 1.0; }

HD void flux_llf_vjp(const Phys& P, const S6<double>& L, const S6<double>& R, double nx, double ny,
                     const double (&Fb)[4], double (&Lb)[4], double (&Rb)[4]) {
  const double g = P.gamma;
  const double pL = pressure(P, L), uL = dv(L[1], L[0]), vL = dv(L[2], L[0]);
  const double unL = uL * nx + vL * ny, cL = sqrt(dv(g * pL, L[0]));
  const double pR = pressure(P, R), uR = dv(R[1], R[0]), vR = dv(R[2], R[0]);
  const double unR = uR * nx + vR * ny, cR = sqrt(dv(g * pR, R[0]));
  const double aL = (unL < 0 ? -unL : unL) + cL, aR = (unR < 0 ? -unR : unR) + cR;
  const bool right = aL < aR;  // mx(aL, aR)
  const double a = right ? aR : aL;
  // F = 0.5 (FL + FR) - (R - L) (0.5 a)
  const double hF[4] = {0.5 * Fb[0], 0.5 * Fb[1], 0.5 * Fb[2], 0.5 * Fb[3]};
  normal_flux_vjp(P, L, nx, ny, hF, Lb);
  normal_flux_vjp(P, R, nx, ny, hF, Rb);
  double abar = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    Rb[k] -= Fb[k] * (0.5 * a);
    Lb[k] += Fb[k] * (0.5 * a);
    abar -= 0.5 * Fb[k] * (R[k] - L[k]);
  }
  // a = |un| + c of the chosen side, c = sqrt(g p / rho)
  const S6<double>& K = right ? R : L;
  double(&Kb)[4] = right ? Rb : Lb;
  const double unK = right ? unR : unL, cK = right ? cR : cL, pK = right ? pR : pL;
  const double uK = right ? uR : uL, vK = right ? vR : vL;
  const double unb = abar * ab_sign(unK);
  const double qb = dv(abar, 2.0 * cK);  // adjoint of g p / rho
  Kb[0] -= qb * dv(g * pK, K[0] * K[0]);
  prim_uvp_vjp(P, K, uK, vK, qb * dv(g, K[0]), unb * nx, unb * ny, Kb);
}

HD void flux_roe_vjp(const Phys& P, const S6<double>& L, const S6<double>& R, double nx, double ny,
                     const double (&Fb)[4], double (&Lb)[4], double (&Rb)[4]) {
  const double g = P.gamma;
  // ---- forward (flux_roe's order)
  const double pL = pressure(P, L), uL = dv(L[1], L[0]), vL = dv(L[2], L[0]), hL = dv(L[3] + pL, L[0]);
  const double pR = pressure(P, R), uR = dv(R[1], R[0]), vR = dv(R[2], R[0]), hR = dv(R[3] + pR, R[0]);
  const double sL = sqrt(L[0]), sR = sqrt(R[0]);
  const double inv = dv(1.0, sL + sR);
  const double ut = (sL * uL + sR * uR) * inv, vt = (sL * vL + sR * vR) * inv, ht = (sL * hL + sR * hR) * inv;
  const double V2 = ut * ut + vt * vt;
  const double c2 = (g - 1.0) * (ht - 0.5 * V2);
  const bool c2clip = c2 < 1e-8;
  const double ct = sqrt(c2clip ? 1e-8 : c2);
  const double unt = ut * nx + vt * ny, utt = -ut * ny + vt * nx;
  const double raw[4] = {unt - ct, unt, unt, unt + ct};
  const double eps = 0.1 * ct;
  double lam[4], fixd[4];
  bool fixed[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double l = raw[i];
    fixed[i] = (l < 0 ? -l : l) < eps;
    if (fixed[i]) l = 0.5 * (dv(l * l, eps) + eps);
    fixd[i] = l;
    lam[i] = l < 0 ? -l : l;
  }
  const double drho = R[0] - L[0], du = uR - uL, dvl = vR - vL, dp = pR - pL;
  const double dun = du * nx + dvl * ny, dut = -du * ny + dvl * nx;
  const double rt = sL * sR;
  const double hc = dv(0.5, c2);
  const double a0 = hc * (dp - rt * ct * dun), a1 = drho - dv(dp, c2), a2 = rt * dut, a3 = hc * (dp + rt * ct * dun);
  const double w1 = lam[0] * a0, w2 = lam[1] * a1, w3 = lam[2] * a2, w4 = lam[3] * a3;
  // ---- reverse: F = 0.5 (FL + FR) - 0.5 D
  const double hF[4] = {0.5 * Fb[0], 0.5 * Fb[1], 0.5 * Fb[2], 0.5 * Fb[3]};
  normal_flux_vjp(P, L, nx, ny, hF, Lb);
  normal_flux_vjp(P, R, nx, ny, hF, Rb);
  const double Db0 = -0.5 * Fb[0], Db1 = -0.5 * Fb[1], Db2 = -0.5 * Fb[2], Db3 = -0.5 * Fb[3];
  const double w1b = Db0 + Db1 * (ut - ct * nx) + Db2 * (vt - ct * ny) + Db3 * (ht - unt * ct);
  const double w2b = Db0 + Db1 * ut + Db2 * vt + Db3 * (0.5 * V2);
  const double w3b = Db1 * (-ny) + Db2 * nx + Db3 * utt;
  const double w4b = Db0 + Db1 * (ut + ct * nx) + Db2 * (vt + ct * ny) + Db3 * (ht + unt * ct);
  double utb = (Db1 + 0.0) * (w1 + w2 + w4), vtb = Db2 * (w1 + w2 + w4), htb = Db3 * (w1 + w4);
  double ctb = (Db1 * nx + Db2 * ny + Db3 * unt) * (w4 - w1);
  double untb = Db3 * ct * (w4 - w1);
  double V2b = Db3 * 0.5 * w2;
  const double uttb = Db3 * w3;
  const double lamb[4] = {w1b * a0, w2b * a1, w3b * a2, w4b * a3};
  const double a0b = w1b * lam[0], a1b = w2b * lam[1], a2b = w3b * lam[2], a3b = w4b * lam[3];
  // a0 = hc (dp - rt ct dun), a3 = hc (dp + rt ct dun), a1 = drho - dp / c2, a2 = rt dut
  const double hcb = a0b * (dp - rt * ct * dun) + a3b * (dp + rt * ct * dun);
  double dpb = hc * (a0b + a3b) - dv(a1b, c2);
  const double dm = hc * (a3b - a0b);
  double rtb = dm * ct * dun + a2b * dut;
  ctb += dm * rt * dun;
  const double dunb = dm * rt * ct, dutb = a2b * rt;
  const double drhob = a1b;
  double c2b = dv(a1b * dp, c2 * c2) - dv(hcb * 0.5, c2 * c2);
  // dun = du nx + dvl ny, dut = -du ny + dvl nx
  const double dub = dunb * nx - dutb * ny, dvlb = dunb * ny + dutb * nx;
  double uLb = -dub, uRb = dub, vLb = -dvlb, vRb = dvlb, pLb = -dpb, pRb = dpb;
  Lb[0] -= drhob;
  Rb[0] += drhob;
  double sLb = rtb * sR, sRb = rtb * sL;
  // entropy fix and |.| of the eigenvalues
  double epsb = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double lb = lamb[i] * (fixd[i] < 0 ? -1.0 : 1.0);
    if (fixed[i]) {
      epsb += lb * 0.5 * (1.0 - dv(raw[i] * raw[i], eps * eps));
      lb = lb * dv(raw[i], eps);
    }
    untb += lb;
    if (i == 0) ctb -= lb;
    if (i == 3) ctb += lb;
  }
  ctb += 0.1 * epsb;
  // utt = -ut ny + vt nx, unt = ut nx + vt ny
  utb += untb * nx - uttb * ny;
  vtb += untb * ny + uttb * nx;
  // ct = sqrt(max(c2, 1e-8)), c2 = (g-1)(ht - V2 / 2), V2 = ut^2 + vt^2
  if (!c2clip) c2b += dv(ctb, 2.0 * ct);
  htb += (g - 1.0) * c2b;
  V2b -= 0.5 * (g - 1.0) * c2b;
  utb += 2.0 * ut * V2b;
  vtb += 2.0 * vt * V2b;
  // Roe averages: x_t = (sL xL + sR xR) inv, inv = 1 / (sL + sR)
  const double invb = utb * (sL * uL + sR * uR) + vtb * (sL * vL + sR * vR) + htb * (sL * hL + sR * hR);
  sLb += (utb * uL + vtb * vL + htb * hL) * inv - invb * inv * inv;
  sRb += (utb * uR + vtb * vR + htb * hR) * inv - invb * inv * inv;
  uLb += utb * sL * inv;
  uRb += utb * sR * inv;
  vLb += vtb * sL * inv;
  vRb += vtb * sR * inv;
  const double hLb = htb * sL * inv, hRb = htb * sR * inv;
  // s = sqrt(rho), h = (U3 + p) / rho
  Lb[0] += dv(sLb, 2.0 * sL) - dv(hLb * hL, L[0]);
  Rb[0] += dv(sRb, 2.0 * sR) - dv(hRb * hR, R[0]);
  Lb[3] += dv(hLb, L[0]);
  Rb[3] += dv(hRb, R[0]);
  pLb += dv(hLb, L[0]);
  pRb += dv(hRb, R[0]);
  prim_uvp_vjp(P, L, uL, vL, pLb, uLb, vLb, Lb);
  prim_uvp_vjp(P, R, uR, vR, pRb, uRb, vRb, Rb);
}

HD void flux_hllc_vjp(const Phys& P, const S6<double>& L, const S6<double>& R, double nx, double ny,
                      const double (&Fb)[4], double (&Lb)[4], double (&Rb)[4]) {
  const double g = P.gamma;
  // ---- forward (flux_hllc's order)
  const double pL = pressure(P, L), pR = pressure(P, R);
  const double uL = dv(L[1], L[0]), uR = dv(R[1], R[0]), vL = dv(L[2], L[0]), vR = dv(R[2], R[0]);
  const double unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
  const double cL = sqrt(dv(g * pL, L[0])), cR = sqrt(dv(g * pR, R[0]));
  const double hL = dv(L[3] + pL, L[0]), hR = dv(R[3] + pR, R[0]);
  const double sL = sqrt(L[0]), sR = sqrt(R[0]);
  const double ss = sL + sR;
  const double ut = dv(sL * uL + sR * uR, ss), vt = dv(sL * vL + sR * vR, ss), ht = dv(sL * hL + sR * hR, ss);
  const double unt = ut * nx + vt * ny;
  const double ct = sqrt((g - 1.0) * (ht - 0.5 * (ut * ut + vt * vt)));
  const double SLa = unL - cL, SLb_ = unt - ct, SRa = unR + cR, SRb_ = unt + ct;
  const bool SLroe = SLb_ < SLa;   // mn(SLa, SLb_)
  const bool SRroe = SRa < SRb_;   // mx(SRa, SRb_)
  const double SL = SLroe ? SLb_ : SLa, SR = SRroe ? SRb_ : SRa;
  if (SL >= 0.0) {
    normal_flux_vjp(P, L, nx, ny, Fb, Lb);
    return;
  }
  if (SR <= 0.0) {
    normal_flux_vjp(P, R, nx, ny, Fb, Rb);
    return;
  }
  const double num = pR - pL + L[0] * unL * (SL - unL) - R[0] * unR * (SR - unR);
  const double den = L[0] * (SL - unL) - R[0] * (SR - unR);
  const double Ss = dv(num, den);
  const bool left = Ss >= 0.0;
  const S6<double>& K = left ? L : R;
  double(&Kb)[4] = left ? Lb : Rb;
  const double S = left ? SL : SR, pK = left ? pL : pR, uK = left ? uL : uR, vK = left ? vL : vR;
  const double unK = left ? unL : unR;
  const double y = S - Ss, sk = S - unK;
  const double coef = dv(K[0] * sk, y);
  const double e = Ss - unK;
  const double z = K[0] * sk;
  const double t = Ss + dv(pK, z);
  const double Us[4] = {coef, coef * (uK + e * nx), coef * (vK + e * ny), coef * (dv(K[3], K[0]) + e * t)};
  // ---- reverse: F = normal_flux(K) + S (Us - K)
  normal_flux_vjp(P, K, nx, ny, Fb, Kb);
  double Sb = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    Sb += Fb[k] * (Us[k] - K[k]);
    Kb[k] -= Fb[k] * S;
  }
  const double Usb0 = Fb[0] * S, Usb1 = Fb[1] * S, Usb2 = Fb[2] * S, Usb3 = Fb[3] * S;
  // Us3 = coef (K3/K0 + e t), t = Ss + pK / z, z = K0 (S - unK)
  const double mk = dv(K[3], K[0]) + e * t;
  double coefb = Usb0 + Usb1 * (uK + e * nx) + Usb2 * (vK + e * ny) + Usb3 * mk;
  const double mb = Usb3 * coef;
  Kb[3] += dv(mb, K[0]);
  double K0b = -dv(mb * K[3], K[0] * K[0]);
  double eb = mb * t + Usb1 * coef * nx + Usb2 * coef * ny;
  const double tb = mb * e;
  double Ssb = tb;
  double pKb = dv(tb, z);
  const double zb = -dv(tb * pK, z * z);
  double skb = zb * K[0];
  K0b += zb * sk;
  double uKb = Usb1 * coef, vKb = Usb2 * coef;
  // e = Ss - unK
  Ssb += eb;
  double unKb = -eb;
  // coef = K0 sk / y, y = S - Ss, sk = S - unK
  K0b += dv(coefb * sk, y);
  skb += dv(coefb * K[0], y);
  const double yb = -dv(coefb * coef, y);
  Sb += yb + skb;
  Ssb -= yb;
  unKb -= skb;
  // Ss = num / den
  const double numb = dv(Ssb, den), denb = -dv(Ssb * Ss, den);
  // the star side's own adjoints, then num and den on both sides
  double SLbar = 0.0, SRbar = 0.0, unLb = 0.0, unRb = 0.0, pLb = 0.0, pRb = 0.0;
  double uLb = 0.0, vLb = 0.0, uRb = 0.0, vRb = 0.0;
  if (left) {
    SLbar += Sb;
    unLb += unKb;
    pLb += pKb;
    uLb += uKb;
    vLb += vKb;
  } else {
    SRbar += Sb;
    unRb += unKb;
    pRb += pKb;
    uRb += uKb;
    vRb += vKb;
  }
  Kb[0] += K0b;
  pRb += numb;
  pLb -= numb;
  Lb[0] += numb * unL * (SL - unL) + denb * (SL - unL);
  unLb += numb * L[0] * (SL - 2.0 * unL) - denb * L[0];
  SLbar += numb * L[0] * unL + denb * L[0];
  Rb[0] -= numb * unR * (SR - unR) + denb * (SR - unR);
  unRb -= numb * R[0] * (SR - 2.0 * unR) - denb * R[0];
  SRbar -= numb * R[0] * unR + denb * R[0];
  // SL = min(unL - cL, unt - ct), SR = max(unR + cR, unt + ct)
  double untb = 0.0, ctb = 0.0, cLb = 0.0, cRb = 0.0;
  if (SLroe) {
    untb += SLbar;
    ctb -= SLbar;
  } else {
    unLb += SLbar;
    cLb -= SLbar;
  }
  if (SRroe) {
    untb += SRbar;
    ctb += SRbar;
  } else {
    unRb += SRbar;
    cRb += SRbar;
  }
  // ct = sqrt((g-1)(ht - (ut^2 + vt^2)/2)), unt = ut nx + vt ny
  const double qb = dv(ctb, 2.0 * ct) * (g - 1.0);
  double utb = untb * nx - qb * ut, vtb = untb * ny - qb * vt;
  const double htb = qb;
  // x_t = (sL xL + sR xR) / ss
  const double ssb = -dv(utb * ut + vtb * vt + htb * ht, ss);
  const double sLb = dv(utb * uL + vtb * vL + htb * hL, ss) + ssb, sRb = dv(utb * uR + vtb * vR + htb * hR, ss) + ssb;
  uLb += dv(utb * sL, ss);
  uRb += dv(utb * sR, ss);
  vLb += dv(vtb * sL, ss);
  vRb += dv(vtb * sR, ss);
  const double hLb = dv(htb * sL, ss), hRb = dv(htb * sR, ss);
  // h = (U3 + p) / rho, s = sqrt(rho), c = sqrt(g p / rho), un = u nx + v ny
  Lb[0] += dv(sLb, 2.0 * sL) - dv(hLb * hL, L[0]);
  Rb[0] += dv(sRb, 2.0 * sR) - dv(hRb * hR, R[0]);
  Lb[3] += dv(hLb, L[0]);
  Rb[3] += dv(hRb, R[0]);
  pLb += dv(hLb, L[0]);
  pRb += dv(hRb, R[0]);
  const double qLb = dv(cLb, 2.0 * cL), qRb = dv(cRb, 2.0 * cR);
  Lb[0] -= qLb * dv(g * pL, L[0] * L[0]);
  Rb[0] -= qRb * dv(g * pR, R[0] * R[0]);
  pLb += qLb * dv(g, L[0]);
  pRb += qRb * dv(g, R[0]);
  uLb += unLb * nx;
  vLb += unLb * ny;
  uRb += unRb * nx;
  vRb += unRb * ny;
  prim_uvp_vjp(P, L, uL, vL, pLb, uLb, vLb, Lb);
  prim_uvp_vjp(P, R, uR, vR, pRb, uRb, vRb, Rb);
}

// riemann() by its scheme switch
HD void riemann_vjp(const Phys& P, const S6<double>& L, const S6<double>& R, double nx, double ny, int scheme,
                    const double (&Fb)[4], double (&Lb)[4], double (&Rb)[4]) {
  if (scheme == CFD_FLUX_ROE) flux_roe_vjp(P, L, R, nx, ny, Fb, Lb, Rb);
  else if (scheme == CFD_FLUX_HLLC) flux_hllc_vjp(P, L, R, nx, ny, Fb, Lb, Rb);
  else flux_llf_vjp(P, L, R, nx, ny, Fb, Lb, Rb);
}

// U = prim2cons(rho, u, v, p): adds Ub^T dU/d(rho, u, v, p) to Wb
HD void prim2cons_vjp(const Phys& P, double rho, double u, double v, const double (&Ub)[4], double (&Wb)[4]) {
  Wb[0] += Ub[0] + Ub[1] * u + Ub[2] * v + Ub[3] * 0.5 * (u * u + v * v);
  Wb[1] += Ub[1] * rho + Ub[3] * rho * u;
  Wb[2] += Ub[2] * rho + Ub[3] * rho * v;
  Wb[3] += dv(Ub[3], P.gamma - 1.0);
}

}  // namespace cfd

// HIP kernels (gfx950) of the residual + JFNK-GMRES path and their launchers.
//
// Residual evaluation (Solver::computeResiduals, Solver.cpp:529-801) runs as
//   k_gradlim  one thread per cell: Green-Gauss gradients (Solver.cpp:81-149)
//              and the Barth-Jespersen limiter (Solver.cpp:152-253), gathered
//              over the cell's faces in reference face order, so the sums are
//              formed in the reference's order;
//   k_flux     one thread per face: MUSCL reconstruction, Riemann flux,
//              wall pressure flux, viscous flux.  Assembly mode 0 launches it
//              once per edge color and scatters +-F*A straight into R (no two
//              faces of a color share a cell, so no atomics); mode 1 writes
//              F*A to a face buffer and
//   k_gather   sums it per cell in reference face order.
// Every block index goes through xcd_block(), which gives each of the eight
// XCDs a contiguous range of cells / faces so neighbouring gathers hit its L2.
#include "ctx.hpp"
#include "riemann_vjp.hpp"
#include "visc.hpp"

namespace cfd {

constexpr int BS = 256;

// Blocks are dispatched round-robin over the 8 XCDs: remap so XCD x walks a
// contiguous run of logical blocks.
__device__ __forceinline__ int xcd_block() {
  int b = blockIdx.x, nb = gridDim.x;
  int q = nb >> 3, r = nb & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// Neighbouring lanes of a wavefront by DPP wave shifts (a VALU move, no LDS
// round trip): lane_prev gives lane l the value of lane l-1, lane_next that of
// lane l+1 (lane 0, resp. 63, receives 0).  Every lane of the wave must be
// active.
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <class T>
__device__ __forceinline__ T lane_prev(T v) { return dpp_move<DPP_WAVE_SHR1>(v); }
template <class T>
__device__ __forceinline__ T lane_next(T v) { return dpp_move<DPP_WAVE_SHL1>(v); }

inline int nblocks(size_t n) { return (int)((n + BS - 1) / BS); }
inline int nblocks(size_t n, int bs) { return (int)((n + bs - 1) / bs); }

// Block sizes of the residual's register-heavy kernels (2 and 3 waves per
// SIMD).  Smaller blocks let a CU refill its wave slots one wave at a time;
// on the 1.32M-cell plate the face flux takes 148 us in 64-thread blocks
// against 156 us in 256-thread ones, and the gradients 120.6 against 122.7
// in 128-thread blocks (the gather, at 8 waves per SIMD, gains nothing).
constexpr int FLUX_BS = 64;
constexpr int GRAD_BS = 128;
// The line preconditioner's per-cell pass: a GMRES(40) iteration on that
// plate, 2.052 / 1.270 ms (FD + FP64 / tangent + FP32) in 64-thread blocks
// against 2.080 / 1.274 ms in 256-thread ones.
constexpr int LINE_BS = 64;
// Basis vectors per block of the Arnoldi reduction: every block also reads
// the candidate u and the product w, so larger groups read them fewer times.
constexpr int DOTS_GROUP = 16;

template <int NV, class T>
__device__ __forceinline__ S6<double> ld(const T* __restrict__ U, int p, int c) {
  S6<double> s;
  s[0] = (double)U[c];
  s[1] = (double)U[p + c];
  s[2] = (double)U[2 * p + c];
  s[3] = (double)U[3 * p + c];
  if (NV == 6) {
    s[4] = (double)U[4 * p + c];
    s[5] = (double)U[5 * p + c];
  }
  return s;
}

// ---- scalar-type plumbing: T = double (residual) or Dual<F> (tangent) -----------
// A state is loaded from U (FP64) and, for the tangent, the direction v (type F).
template <class T> struct Scal {
  typedef double F;
  static __device__ __forceinline__ T mk(double u, F) { return u; }
  static __device__ __forceinline__ double dpart(T x) { return x; }
  static __device__ __forceinline__ void put(F* v, F*, size_t i, T x) { v[i] = x; }
  static __device__ __forceinline__ T get(const F* v, const F*, size_t i) { return v[i]; }
};
template <class F_> struct Scal<Dual<F_>> {
  typedef F_ F;
  static __device__ __forceinline__ Dual<F> mk(double u, F d) { return Dual<F>(F(u), d); }
  static __device__ __forceinline__ F dpart(Dual<F> x) { return x.d; }
  static __device__ __forceinline__ void put(F* v, F* d, size_t i, Dual<F> x) {
    v[i] = x.v;
    d[i] = x.d;
  }
  static __device__ __forceinline__ Dual<F> get(const F* v, const F* d, size_t i) { return Dual<F>(v[i], d[i]); }
};

template <int NV, class T>
__device__ __forceinline__ S6<T> ldT(const double* __restrict__ U, const typename Scal<T>::F* __restrict__ v, int p,
                                     int c) {
  S6<T> s;
#pragma unroll
  for (int k = 0; k < NV; ++k) s[k] = Scal<T>::mk(U[k * p + c], v ? v[k * p + c] : typename Scal<T>::F(0));
  return s;
}

// ---- gradients + limiter ------------------------------------------------------
// Gv / Gd: 10 x pitch values (and, for the tangent, derivatives) of the
// Green-Gauss gradients in CellGradient order.  The limiter is only formed in
// the FP64 residual; the tangent keeps it frozen.
// With SST the cell's eddy viscosity, blending F1 and V * (k, omega sources)
// are formed here once per cell (rows 0..3 of Qv / Qd) for the face fluxes.
// The FP64 residual's four-variable instance runs at 4 waves per SIMD (100
// VGPRs); the SST one at 3 (134 VGPRs): held to 128 it spills six registers
// and took 115.9 against 110.5 us on the 1.32M-cell plate.  The FP32
// tangent's (135 VGPRs, five spilled at 128) also runs at 3.
template <int NV, class T>
__global__ __launch_bounds__(GRAD_BS) __attribute__((amdgpu_waves_per_eu(std::is_same<T, double>::value && NV == 4 ? 4 : 1, 8)))
void k_gradlim(DMesh m, Phys P, const double* __restrict__ U,
                                                 const typename Scal<T>::F* __restrict__ v,
                                                 typename Scal<T>::F* __restrict__ Gv,
                                                 typename Scal<T>::F* __restrict__ Gd, double* __restrict__ phi,
                                                 int do_lim, typename Scal<T>::F* __restrict__ Qv = nullptr,
                                                 typename Scal<T>::F* __restrict__ Qd = nullptr) {
  typedef Scal<T> Sc;
  constexpr int NQM = NV == 6 ? 7 : 5;  // gradient variables: rho u v p T (+ k omega with SST)
  const int c = xcd_block() * GRAD_BS + threadIdx.x;
  if (c >= m.N) return;
  const int p = m.pitch;
  const double Rg = P.R;
  const bool turb = NV == 6 && P.turb;
  const int nq = turb ? 7 : 5;
  const S6<T> Uc = ldT<NV, T>(U, v, p, c);
  const T uc = dv(Uc[1], Uc[0]), vc = dv(Uc[2], Uc[0]), pc = pressure(P, Uc), Tc = dv(pc, Uc[0] * Rg);
  T g[2 * NQM];
#pragma unroll
  for (int q = 0; q < 2 * NQM; ++q) g[q] = T(0.0);
  double pmax = val(pc), pmin = val(pc);
  const int4 s4 = m.slots[c], n4 = m.cnb[c];
  const int sl[4] = {s4.x, s4.y, s4.z, s4.w}, on[4] = {n4.x, n4.y, n4.z, n4.w};
  bool bnd = false;  // the cell has a boundary face
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int e = sl[k];
    if (e < 0) continue;
    const int f = e >> 1;
    const bool as_nei = e & 1;
    const int other = on[k];
    bnd = bnd || other < 0;
    const double nx = m.nx[f], ny = m.ny[f], A = m.area[f];
    const S6<T> Uo = other >= 0 ? ldT<NV, T>(U, v, p, other)
                                : ghost(P, Uc, m.tag[f], nx, ny, turb ? m.dnw[f] : 0.0);
    const T uo = dv(Uo[1], Uo[0]), vo = dv(Uo[2], Uo[0]), po = pressure(P, Uo), To = dv(po, Uo[0] * Rg);
    T fv[NQM];
    fv[0] = 0.5 * (Uc[0] + Uo[0]);
    fv[1] = 0.5 * (uc + uo);
    fv[2] = 0.5 * (vc + vo);
    fv[3] = 0.5 * (pc + po);
    fv[4] = 0.5 * (Tc + To);
    if constexpr (NQM == 7) {
      if (turb) {
        fv[5] = 0.5 * (dv(Uc[4], Uc[0]) + dv(Uo[4], Uo[0]));
        fv[6] = 0.5 * (dv(Uc[5], Uc[0]) + dv(Uo[5], Uo[0]));
      } else {
        fv[5] = fv[6] = T(0.0);
      }
    }
    if (!as_nei) {
#pragma unroll
      for (int q = 0; q < NQM; ++q) {
        g[2 * q] = g[2 * q] + fv[q] * nx * A;
        g[2 * q + 1] = g[2 * q + 1] + fv[q] * ny * A;
      }
    } else {
#pragma unroll
      for (int q = 0; q < NQM; ++q) {
        g[2 * q] = g[2 * q] - fv[q] * nx * A;
        g[2 * q + 1] = g[2 * q + 1] - fv[q] * ny * A;
      }
    }
    pmax = mx(pmax, (double)val(po));
    pmin = mn(pmin, (double)val(po));
  }
  const double vol = m.vol[c];
  // flags bit 1: the k and omega rows (10-13) are read only on boundary faces
  // (k and omega convected with the cell values), so only cells with one
  // write them
  const bool wkw = !(do_lim & 2) || bnd;
#pragma unroll
  for (int q = 0; q < 2 * NQM; ++q) {
    if (q < 2 * nq) {
      g[q] = dv(g[q], vol);
      if (q < 10 || wkw) Sc::put(Gv, Gd, (size_t)q * p + c, g[q]);
    }
  }
  if constexpr (NQM == 7) {
    if (turb && Qv) {
      const TurbQ<T> tq = turb_cell(P, Uc, g, m.wd[c]);
      Sc::put(Qv, Qd, (size_t)c, tq.mut);
      Sc::put(Qv, Qd, (size_t)p + c, tq.F1);
      Sc::put(Qv, Qd, (size_t)2 * p + c, vol * tq.Sk);
      Sc::put(Qv, Qd, (size_t)3 * p + c, vol * tq.Sw);
    }
  }
  if constexpr (std::is_same<T, double>::value) {
    if (!(do_lim & 1)) return;
    const double xc = m.xc[c], yc = m.yc[c];
    double ph = 1.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int e = sl[k];
      if (e < 0) continue;
      const int f = e >> 1;
      const double dx = m.xm[f] - xc, dy = m.ym[f] - yc;
      const double Tf = Tc + 1.0 * (g[8] * dx + g[9] * dy);
      const double Pf = Uc[0] * Rg * Tf;
      const double d = Pf - pc;
      double a = 1.0;
      if (d > 0.0) a = mn(1.0, (pmax - pc) / (d + 1e-12));
      else if (d < 0.0) a = mn(1.0, (pmin - pc) / (d - 1e-12));
      ph = mn(ph, a);
    }
    phi[c] = ph;
  }
}

// SST sources: R_c(k, omega) -= V_c S_c, with V_c S_c formed by k_gradlim
template <class F>
__global__ __launch_bounds__(BS) void k_turb_src(int N, int p, const F* __restrict__ Q, F* __restrict__ out) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= N) return;
  out[4 * p + c] = out[4 * p + c] - Q[(size_t)2 * p + c];
  out[5 * p + c] = out[5 * p + c] - Q[(size_t)3 * p + c];
}

// ---- face flux ------------------------------------------------------------------------

// Every load is unconditional apart from kernel-uniform switches, so the
// loads of one face go out together: a boundary face (n < 0) reads its owner
// in place of the missing neighbour and the value is not used.
__device__ __forceinline__ FaceGeo face_geo(const DMesh& m, const Phys& P, const double* __restrict__ phi, int f,
                                            int o, int n, bool need_cells) {
  FaceGeo g;
  const int nn = n >= 0 ? n : o;
  g.nx = m.nx[f];
  g.ny = m.ny[f];
  g.A = m.area[f];
  g.tg = m.tag[f];
  g.boundary = n < 0;
  g.scheme = (m.alt && m.alt[f]) ? P.flux_alt : P.flux;
  g.xm = g.ym = g.xco = g.yco = g.xcn = g.ycn = 0.0;
  g.phL = g.phR = 0.5;
  g.wdo = g.wdn = g.dnw = 1.0;
  if (P.turb) {
    g.wdo = m.wd[o];
    const double wn = m.wd[nn];
    g.wdn = n >= 0 ? wn : g.wdo;
    g.dnw = n < 0 ? m.dnw[f] : 1.0;  // read only by the ghost state and the wall terms
  }
  if (need_cells) {
    g.xm = m.xm[f];
    g.ym = m.ym[f];
    g.xco = m.xc[o];
    g.yco = m.yc[o];
    const double xn = m.xc[nn], yn = m.yc[nn];
    if (n >= 0) {
      g.xcn = xn;
      g.ycn = yn;
    }
    if (P.spatial == CFD_SECOND_ORDER_LIMITED) {
      g.phL = phi[o];
      const double pn = phi[nn];
      g.phR = n >= 0 ? pn : 0.5;
    }
  }
  return g;
}

// F * A of one face from its owner state L, right state R (neighbor or
// ghost), and the two cells' gradients (Solver.cpp:543-799): MUSCL
// reconstruction with the reconstruction clips, wall pressure flux or Riemann
// flux, viscous flux with the centroid-line correction.
template <int VISC, class T>
__device__ __forceinline__ S6<T> flux_core(const Phys& P, const FaceGeo& g, const S6<T>& L, const S6<T>& R,
                                           const T* gL, const T* gR, const TurbQ<T>* tqo = nullptr,
                                           const TurbQ<T>* tqn = nullptr) {
  const double nx = g.nx, ny = g.ny;
  S6<T> Lr = L, Rr = R;
  if (P.spatial != CFD_FIRST_ORDER && !g.boundary) {
    const double dxL = g.xm - g.xco, dyL = g.ym - g.yco, dxR = g.xm - g.xcn, dyR = g.ym - g.ycn;
    const double phL = g.phL, phR = g.phR;
    const T uL = dv(L[1], L[0]), vL = dv(L[2], L[0]), pL = pressure(P, L);
    T rLf = L[0] + phL * (gL[0] * dxL + gL[1] * dyL);
    T pLf = pL + phL * (gL[6] * dxL + gL[7] * dyL);
    const T uLf = uL + phL * (gL[2] * dxL + gL[3] * dyL);
    const T vLf = vL + phL * (gL[4] * dxL + gL[5] * dyL);
    if (val(rLf) < 1e-4) rLf = L[0];
    if (val(pLf) < 1e-4) pLf = pL;
    Lr = prim2cons(P, rLf, uLf, vLf, pLf);
    const T uR = dv(R[1], R[0]), vR = dv(R[2], R[0]), pR = pressure(P, R);
    T rRf = R[0] + phR * (gR[0] * dxR + gR[1] * dyR);
    T pRf = pR + phR * (gR[6] * dxR + gR[7] * dyR);
    const T uRf = uR + phR * (gR[2] * dxR + gR[3] * dyR);
    const T vRf = vR + phR * (gR[4] * dxR + gR[5] * dyR);
    if (val(rRf) < 1e-4) rRf = R[0];
    if (val(pRf) < 1e-4) pRf = pR;
    Rr = prim2cons(P, rRf, uRf, vRf, pRf);
    if (P.turb) {  // MUSCL on specific k and omega, clipped at zero; or the cell values
      const T kL = dv(L[4], L[0]), wL = dv(L[5], L[0]), kR = dv(R[4], R[0]), wR = dv(R[5], R[0]);
      T kLf = kL, wLf = wL, kRf = kR, wRf = wR;
      if (!P.turb1) {
        kLf = kL + phL * (gL[10] * dxL + gL[11] * dyL);
        wLf = wL + phL * (gL[12] * dxL + gL[13] * dyL);
        kRf = kR + phR * (gR[10] * dxR + gR[11] * dyR);
        wRf = wR + phR * (gR[12] * dxR + gR[13] * dyR);
        if (val(kLf) < 0.0) kLf = kL;
        if (val(wLf) < 0.0) wLf = wL;
        if (val(kRf) < 0.0) kRf = kR;
        if (val(wRf) < 0.0) wRf = wR;
      }
      Lr[4] = rLf * kLf;
      Lr[5] = rLf * wLf;
      Rr[4] = rRf * kRf;
      Rr[5] = rRf * wRf;
    }
  }
  S6<T> Fl;
  if (g.boundary && (g.tg == 3 || g.tg == 4)) {
    const T pw_ = pressure(P, Lr);
    Fl[0] = T(0.0);
    Fl[1] = pw_ * nx;
    Fl[2] = pw_ * ny;
    Fl[3] = T(0.0);
  } else {
    Fl = riemann(P, Lr, Rr, nx, ny, g.scheme);
  }
  if (P.turb) {  // k and omega carried by the mass flux, upwinded
    const T mf = Fl[0];
    const T mp = val(mf) > 0.0 ? mf : T(0.0), mm = val(mf) > 0.0 ? T(0.0) : mf;
    Fl[4] = mp * dv(Lr[4], Lr[0]) + mm * dv(Rr[4], Rr[0]);
    Fl[5] = mp * dv(Lr[5], Lr[0]) + mm * dv(Rr[5], Rr[0]);
  }
  if (VISC) {
    const T rf = 0.5 * (L[0] + R[0]);
    const T uf = 0.5 * (dv(L[1], L[0]) + dv(R[1], R[0]));
    const T vf = 0.5 * (dv(L[2], L[0]) + dv(R[2], R[0]));
    T ux, uy, vx, vy, Tx, Ty;
    if (g.boundary && g.tg == 4) {
      const double dx = g.xm - g.xco, dy = g.ym - g.yco;
      const double d2 = dx * dx + dy * dy;
      const T uc = dv(L[1], L[0]), vc = dv(L[2], L[0]);
      ux = dv(-uc * dx, d2);
      uy = dv(-uc * dy, d2);
      vx = dv(-vc * dx, d2);
      vy = dv(-vc * dy, d2);
      Tx = T(0.0);
      Ty = T(0.0);
    } else {
      ux = 0.5 * (gL[2] + gR[2]);
      uy = 0.5 * (gL[3] + gR[3]);
      vx = 0.5 * (gL[4] + gR[4]);
      vy = 0.5 * (gL[5] + gR[5]);
      Tx = 0.5 * (gL[8] + gR[8]);
      Ty = 0.5 * (gL[9] + gR[9]);
      if (!g.boundary) {
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
      }
    }
    T mu = rf * P.nu;
    T kc = dv(mu * P.Cp, P.Pr);
    T mut = T(0.0), F1f = T(0.0);
    if (P.turb) {  // eddy viscosity in stress and heat flux
      const TurbQ<T> qo = tqo ? *tqo : turb_cell(P, L, gL, g.wdo);
      if (!g.boundary) {
        const TurbQ<T> qn = tqn ? *tqn : turb_cell(P, R, gR, g.wdn);
        mut = 0.5 * (qo.mut + qn.mut);
        F1f = 0.5 * (qo.F1 + qn.F1);
      } else {
        mut = g.tg == 4 ? T(0.0) : qo.mut;
        F1f = qo.F1;
      }
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
    Fl[0] = Fl[0] - 0.0;
    Fl[1] = Fl[1] - (txx * nx + txy * ny);
    Fl[2] = Fl[2] - (txy * nx + tyy * ny);
    Fl[3] = Fl[3] - (wx * nx + wy * ny);
    if (P.turb) {  // (mu + sigma mu_t) dk/dn, domega/dn by two-point normal differences
      // wall: to k = 0 and omega_wall over the owner's normal distance;
      // interior: between the centroids over their normal separation (a
      // positive coupling on the stretched split cells, where the averaged
      // gradients decouple the two triangles of a quad); averaged gradients on
      // the other boundary faces
      T kn, wn;
      const T kcl = dv(L[4], L[0]), wcl = dv(L[5], L[0]);
      if (g.boundary && g.tg == 4) {
        const double dn = g.dnw;
        const double ww = omega_wall(P, g.dnw);
        kn = dv(0.0 - kcl, dn);
        wn = dv(ww - wcl, dn);
      } else if (!g.boundary) {
        double dn = fabs((g.xcn - g.xco) * nx + (g.ycn - g.yco) * ny);
        if (dn < 1e-300) dn = 1e-300;
        kn = dv(dv(R[4], R[0]) - kcl, dn);
        wn = dv(dv(R[5], R[0]) - wcl, dn);
      } else {
        kn = 0.5 * (gL[10] + gR[10]) * nx + 0.5 * (gL[11] + gR[11]) * ny;
        wn = 0.5 * (gL[12] + gR[12]) * nx + 0.5 * (gL[13] + gR[13]) * ny;
      }
      const T mul = rf * P.nu;
      const T sgk = F1f * sst::sk1 + (1.0 - F1f) * sst::sk2, sgw = F1f * sst::sw1 + (1.0 - F1f) * sst::sw2;
      Fl[4] = Fl[4] - (mul + sgk * mut) * kn;
      Fl[5] = Fl[5] - (mul + sgw * mut) * wn;
    } else {
      Fl[4] = Fl[4] - 0.0;
      Fl[5] = Fl[5] - 0.0;
    }
  }
  return scale(Fl, T(g.A));
}

template <int NV, int VISC, class T>
__device__ __forceinline__ S6<T> face_flux(const DMesh& m, const Phys& P, const double* __restrict__ U,
                                           const typename Scal<T>::F* __restrict__ v,
                                           const typename Scal<T>::F* __restrict__ Gv,
                                           const typename Scal<T>::F* __restrict__ Gd,
                                           const double* __restrict__ phi, int f, int& o, int& n,
                                           const typename Scal<T>::F* __restrict__ Qv,
                                           const typename Scal<T>::F* __restrict__ Qd) {
  typedef Scal<T> Sc;
  const int p = m.pitch;
  o = m.own[f];
  n = m.nei[f];
  // All of the face's cell data is loaded before any of it is used (a
  // boundary face reads its owner twice), so the loads overlap: the face's
  // cells, then everything indexed by them, in two dependent rounds.
  const int nn = n >= 0 ? n : o;
  const bool useg = VISC || P.spatial != CFD_FIRST_ORDER;
  const FaceGeo g = face_geo(m, P, phi, f, o, n, useg);
  constexpr int NG = NV == 6 ? 14 : 10;
  const int ng = (NV == 6 && P.turb) ? 14 : 10;
  const S6<T> L = ldT<NV, T>(U, v, p, o);
  const S6<T> Rn = ldT<NV, T>(U, v, p, nn);
  T gL[NG], gR[NG];
  if (useg) {
#pragma unroll
    for (int q = 0; q < NG; ++q) {
      // the k and omega rows (10-13) only where they are read: reconstructed
      // k and omega, or a boundary face's averaged gradients
      const bool rd = q < ng && (q < 10 || !P.turb1 || n < 0);
      gL[q] = rd ? Sc::get(Gv, Gd, (size_t)q * p + o) : T(0.0);
      gR[q] = rd ? Sc::get(Gv, Gd, (size_t)q * p + nn) : T(0.0);
    }
  }
  TurbQ<T> qo, qn;
  const bool tq = NV == 6 && VISC && P.turb;
  if (tq) {
    qo.mut = Sc::get(Qv, Qd, (size_t)o);
    qo.F1 = Sc::get(Qv, Qd, (size_t)p + o);
    qn.mut = Sc::get(Qv, Qd, (size_t)nn);
    qn.F1 = Sc::get(Qv, Qd, (size_t)p + nn);
  }
  const S6<T> R = n >= 0 ? Rn : ghost(P, L, g.tg, g.nx, g.ny, g.dnw);
  if (useg && n < 0) {  // boundary faces use the owner's gradients on both sides
#pragma unroll
    for (int q = 0; q < NG; ++q) gR[q] = gL[q];
  }
  if (tq) {
    if (n < 0) qn = qo;
    return flux_core<VISC, T>(P, g, L, R, gL, gR, &qo, &qn);
  }
  return flux_core<VISC, T>(P, g, L, R, gL, gR);
}

// The face buffer holds F * A of face f in the row Fm[f * NV + k] (k the
// variable): a face's values are adjacent, so they move as 16-byte (FP64, or
// FP32 with NV = 4) or 8-byte (FP32, NV = 6) vectors.  Rows start on that
// alignment because the buffer does (device allocations are 256-byte aligned).
template <int NV, class F> constexpr size_t row_align() { return (sizeof(F) * NV) % 16 == 0 ? 16 : 8; }

// MODE 0: scatter into `out` (the residual) for faces [f0, f1) of one color.
// MODE 1: write the face buffer.  The tangent writes derivative parts only.
template <int NV, int VISC, int MODE, class T>
__global__ __launch_bounds__(FLUX_BS) void k_flux(DMesh m, Phys P, const double* __restrict__ U,
                                              const typename Scal<T>::F* __restrict__ v,
                                              const typename Scal<T>::F* __restrict__ Gv,
                                              const typename Scal<T>::F* __restrict__ Gd,
                                              const double* __restrict__ phi, typename Scal<T>::F* __restrict__ out,
                                              int f0, int f1, const typename Scal<T>::F* __restrict__ Qv,
                                              const typename Scal<T>::F* __restrict__ Qd) {
  typedef Scal<T> Sc;
  const int i = f0 + xcd_block() * FLUX_BS + threadIdx.x;
  if (i >= f1) return;
  const int f = MODE == 0 ? m.corder[i] : i;
  int o, n;
  const S6<T> TA = face_flux<NV, VISC, T>(m, P, U, v, Gv, Gd, phi, f, o, n, Qv, Qd);
  if (MODE == 0) {
    const int p = m.pitch;
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k * p + o] = out[k * p + o] + Sc::dpart(TA[k]);
    if (n >= 0) {
#pragma unroll
      for (int k = 0; k < NV; ++k) out[k * p + n] = out[k * p + n] - Sc::dpart(TA[k]);
    }
  } else {  // face-major row of NV values, written with wide stores
    typedef typename Scal<T>::F F;
    F* row = (F*)__builtin_assume_aligned(out + (size_t)f * NV, row_align<NV, F>());
#pragma unroll
    for (int k = 0; k < NV; ++k) row[k] = Sc::dpart(TA[k]);
  }
}

template <int NV, class F>
__global__ __launch_bounds__(BS) void k_gather(DMesh m, const F* __restrict__ Fm, F* __restrict__ R,
                                                const F* __restrict__ Q = nullptr, const F* __restrict__ v = nullptr,
                                                const double* __restrict__ dt = nullptr) {
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= m.N) return;
  const int4 s4 = m.slots[c];
  const int sl[4] = {s4.x, s4.y, s4.z, s4.w};
  F r[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) r[k] = F(0);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int e = sl[j];
    if (e < 0) continue;
    const F* row = (const F*)__builtin_assume_aligned(Fm + (size_t)(e >> 1) * NV, row_align<NV, F>());
    F fr[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) fr[k] = row[k];
    if (e & 1) {
#pragma unroll
      for (int k = 0; k < NV; ++k) r[k] = r[k] - fr[k];
    } else {
#pragma unroll
      for (int k = 0; k < NV; ++k) r[k] = r[k] + fr[k];
    }
  }
  if constexpr (NV == 6) {
    if (Q) {  // SST sources: R(k, omega) -= V S (rows 2 and 3 of the per-cell turbulence quantities)
      r[4] = r[4] - Q[(size_t)2 * m.pitch + c];
      r[5] = r[5] - Q[(size_t)3 * m.pitch + c];
    }
  }
  if (v) {  // tangent: + (vol / dt) v  (Solver.cpp:1381-1384 time term)
    const double tc = m.vol[c] / (dt[c] + 1e-30);
#pragma unroll
    for (int k = 0; k < NV; ++k) R[k * m.pitch + c] = (F)((double)r[k] + (double)v[k * m.pitch + c] * tc);
  } else {
#pragma unroll
    for (int k = 0; k < NV; ++k) R[k * m.pitch + c] = r[k];
  }
}

// Tangent Jv: J v = dR/dU v + (vol / dt) v  (Solver.cpp:1381-1384 time term)
template <int NV, class F>
__global__ __launch_bounds__(BS) void k_tcombine(int N, int p, F* __restrict__ Jv, const F* __restrict__ v,
                                                  const double* __restrict__ vol, const double* __restrict__ dt) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= N) return;
  const double tc = vol[c] / (dt[c] + 1e-30);
#pragma unroll
  for (int k = 0; k < NV; ++k) Jv[k * p + c] = (F)((double)Jv[k * p + c] + (double)v[k * p + c] * tc);
}

// ---- reverse product J^T w ------------------------------------------------------------
// With the limiter frozen, R_c = sum_f s_cf F_f A_f(U_o, U_n, G_o, G_n) and
// G_c = (1/V_c) sum_f s_cf A_f n_f q_f(U_o, U_n), q = face mean of (rho, u, v,
// p, T).  The transpose product is formed in two colored face sweeps:
//   k_vjp_face  w^T dF/d(U_o, U_n, G_o, G_n) per face, lambda = w_o - w_n,
//               by VK-direction forward duals over the face's inputs
//               (gradients recomputed, never stored per face);
//   k_vjp_gg    the Green-Gauss transpose: aG_c / V_c back onto the face means
//               and through the primitive (and ghost) maps onto U.
// Directions per pass.  Two directions keep the SST instance nearly free of
// spills (4: 208 of 512 registers spilled, 2: 6) and cost least overall:
// 13.8 / 10.1 / 15.1 ms per product with 4 / 2 / 1 on the 1.32M-cell SST
// plate with k and omega reconstructed (this kernel's whole product), and
// 1.068 -> 1.040 ms with the boundary faces of the plate's product.
constexpr int VK_GEN = 2;
// passes of VK directions that cover a face's inputs (both states, both gradients)
template <int NV, int VK> __host__ __device__ constexpr int vjp_passes() {
  return (2 * NV + 2 * (NV == 6 ? 14 : 10) + VK - 1) / VK;
}

// One pass of VK directions of face f: w^T dF/d(inputs base .. base + VK - 1),
// added to the inputs' adjoint rows.
template <int NV, int VISC, int VK>
__device__ __forceinline__ void vjp_face_pass(const DMesh& m, const Phys& P, const double* __restrict__ U,
                                              const double* __restrict__ G, const double* __restrict__ phi,
                                              const double* __restrict__ w, double* __restrict__ aU,
                                              double* __restrict__ aG, int f, int pass) {
  typedef DualV<double, VK> T;
  constexpr int NGM = NV == 6 ? 14 : 10;
  const int p = m.pitch;
  const int o = m.own[f], n = m.nei[f];
  const bool interior = n >= 0;
  const int ng = (NV == 6 && P.turb) ? 14 : 10;
  const bool need = (P.spatial != CFD_FIRST_ORDER && interior) || VISC;
  const FaceGeo g = face_geo(m, P, phi, f, o, n, need);
  double Uo[NV], Un[NV], lam[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    Uo[k] = U[k * p + o];
    Un[k] = interior ? U[k * p + n] : 0.0;
    lam[k] = w[k * p + o] - (interior ? w[k * p + n] : 0.0);
  }
  const int nU = interior ? 2 * NV : NV;
  const int nin = nU + (need ? (interior ? 2 * ng : ng) : 0);
  for (int base = pass * VK; base < nin && base < (pass + 1) * VK; base += VK) {
    S6<T> L, R;
    T gL[NGM], gR[NGM];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      L[j] = T(Uo[j]);
#pragma unroll
      for (int d = 0; d < VK; ++d) L[j].d[d] = (j == base + d) ? 1.0 : 0.0;
    }
    if (interior) {
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        R[j] = T(Un[j]);
#pragma unroll
        for (int d = 0; d < VK; ++d) R[j].d[d] = (NV + j == base + d) ? 1.0 : 0.0;
      }
    } else {
      R = ghost(P, L, g.tg, g.nx, g.ny, g.dnw);
    }
    if (need) {
#pragma unroll
      for (int q = 0; q < NGM; ++q) {
        gL[q] = T(q < ng ? G[(size_t)q * p + o] : 0.0);
#pragma unroll
        for (int d = 0; d < VK; ++d) gL[q].d[d] = (q < ng && nU + q == base + d) ? 1.0 : 0.0;
      }
      if (interior) {
#pragma unroll
        for (int q = 0; q < NGM; ++q) {
          gR[q] = T(q < ng ? G[(size_t)q * p + n] : 0.0);
#pragma unroll
          for (int d = 0; d < VK; ++d) gR[q].d[d] = (q < ng && nU + ng + q == base + d) ? 1.0 : 0.0;
        }
      } else {
#pragma unroll
        for (int q = 0; q < NGM; ++q) gR[q] = gL[q];
      }
    }
    const S6<T> F = flux_core<VISC, T>(P, g, L, R, gL, gR);
#pragma unroll
    for (int d = 0; d < VK; ++d) {
      const int idx = base + d;
      if (idx >= nin) break;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < NV; ++k) s += lam[k] * F[k].d[d];
      double* dst;
      if (idx < NV) dst = aU + (size_t)idx * p + o;
      else if (idx < nU) dst = aU + (size_t)(idx - NV) * p + n;
      else if (idx < nU + ng) dst = aG + (size_t)(idx - nU) * p + o;
      else dst = aG + (size_t)(idx - nU - ng) * p + n;
      *dst = *dst + s;
    }
  }
}

// One thread per (face, pass) over the faces corder[f0 .. f1) of one color
// (no two share a cell): the passes of a face run on neighbouring lanes and
// write disjoint inputs' adjoints.
template <int NV, int VISC>
__global__ __launch_bounds__(FLUX_BS) void k_vjp_face(DMesh m, Phys P, const double* __restrict__ U,
                                                  const double* __restrict__ G, const double* __restrict__ phi,
                                                  const double* __restrict__ w, double* __restrict__ aU,
                                                  double* __restrict__ aG, int f0, int f1) {
  const int t = xcd_block() * FLUX_BS + threadIdx.x;
  const int i = f0 + t / vjp_passes<NV, VK_GEN>(), pass = t % vjp_passes<NV, VK_GEN>();
  if (i >= f1) return;
  vjp_face_pass<NV, VISC, VK_GEN>(m, P, U, G, phi, w, aU, aG, m.corder[i], pass);
}

// The boundary faces in one launch: one thread per (cell with boundary faces,
// pass), which adds its cell's boundary faces one after the other in the
// order of their colors, as the per-color launches did (the sums are the
// same to the bit).  Each launch of the general kernel lasts about as long
// as one thread's dual passes whatever its size (28 us on the 1.32M-cell SST
// plate), so three color launches cost three times one.
template <int NV, int VISC>
__global__ __launch_bounds__(FLUX_BS) void k_vjp_face_bnd(DMesh m, Phys P, const double* __restrict__ U,
                                                      const double* __restrict__ G, const double* __restrict__ phi,
                                                      const double* __restrict__ w, double* __restrict__ aU,
                                                      double* __restrict__ aG, const int4* __restrict__ grp,
                                                      int ngrp) {
  const int t = xcd_block() * FLUX_BS + threadIdx.x;
  const int gi = t / vjp_passes<NV, VK_GEN>(), pass = t % vjp_passes<NV, VK_GEN>();
  if (gi >= ngrp) return;
  const int4 g4 = grp[gi];
  const int fs[4] = {g4.x, g4.y, g4.z, g4.w};
  for (int k = 0; k < 4; ++k) {
    if (fs[k] < 0) break;
    vjp_face_pass<NV, VISC, VK_GEN>(m, P, U, G, phi, w, aU, aG, fs[k], pass);
  }
}


// Interior faces of a second-order residual with four variables (Euler or
// laminar), or with SST and k, omega convected first-order upwind (NV = 6):
// k_vjp_face's product, using the face's structure, one thread per face over
// all faces in one launch.  Such a face's flux F A depends on the two cells'
// gradients only through the reconstructed face primitives (rho, u, v, p) of
// each side, linearly (weights phi dx, phi dy, none where a reconstruction
// clip took the cell value), and on the cell states through their
// primitives.  So lambda^T d(F A) is formed by the hand-written reverse sweeps
// of the Riemann flux and of prim2cons (riemann_vjp.hpp) over the 8 face
// primitives and carried back by hand onto the gradient adjoints and, through
// the primitive map's Jacobian, onto the states.  The viscous terms (VISC)
// depend on the two states and on the averages of six gradient components
// (and with SST on the cells' mu_t and F1, read from Q rows 0 and 1): their
// adjoints come from the reverse sweep visc_interior_vjp (visc.hpp).  Each
// sweep costs a few forward evaluations; forward duals took one pass of 8
// directions through the Riemann flux and two (laminar) or four (SST) passes
// of 6-7 directions through the viscous terms, and held the whole register
// file.  With SST the k and omega fluxes are the mass flux times the upwind
// cell's specific k and omega, carried back by hand.
//
// The results go to the face's row of the adjoint face buffer (layout below),
// not to the cells: k_vjp_gather sums each cell's rows.  Against one launch
// per edge color with a read-modify-write of both cells' adjoint rows, this
// halves the face pass on the 1.32M-cell SST plate (1.75 -> 0.84 ms per
// product): the faces are read in owner order and written as contiguous
// rows, and there is one launch instead of six tails.  Spreading a face over
// four lanes of 2-direction duals (a quad of wavefronts exchanging partial
// results through LDS) was measured too: it recomputes the primal flux in
// every lane and ran at 1.70 ms.
//
// Face buffer: two arrays of rows, side 0 (the owner's) then side 1 (the
// neighbour's), each F rows of vjp_row doubles: the side's NV state adjoints
// and, with viscous terms, its gradient slots 0-9 (the averaged components'
// adjoints go half to each side); without them, the adjoints of the side's
// four reconstructed face primitives, whose gradient slots k_vjp_gather forms
// as a (wx, wy) from the face midpoint (12 doubles a row had the slots
// themselves: the bump's face pass and sums moved 148 MB more); then, with
// SST, the sides' mu_t and F1 adjoints, (side, face) pairs of doubles.  A cell reads only its
// own side's rows, and the owner side's rows of a cell's faces are adjacent
// (faces in owner order).  A boundary face's rows are zero: the general
// kernel adds its product straight into the cells' rows.
// One side's row: NV state and 8 (Euler) or 10 (viscous) gradient
// adjoints.  With SST that is 16 doubles, one 128-byte cache line (the buffer
// is 256-byte aligned), and the mu_t and F1 adjoints go to a third array of
// two doubles per side and face: a 144-byte row straddled two lines and the
// per-cell sums took 271 instead of 233 us on the 1.32M-cell plate.  Padding
// the Euler rows (12 doubles) to 16 made the bump's sums slower (137 against
// 105 us): the padding is read too.
template <int NV, int VISC> __host__ __device__ constexpr int vjp_row() { return NV + (VISC ? 10 : 4); }
static_assert(vjp_row<6, 1>() == 16 && 2 * 16 + 2 * 2 <= VJP_ROW_MAX, "adjoint face buffer rows");
// Held to two waves per SIMD: the SST instance needs 270 registers for one
// wave and spills 18 at two; both run the plate's face pass in the same time,
// and the two-wave form leaves room for the four-variable instances (214-238).
template <int NV, int VISC>
__global__ __launch_bounds__(FLUX_BS) __attribute__((amdgpu_waves_per_eu(2, 8))) void k_vjp_face_inv(DMesh m, Phys P, const double* __restrict__ U,
                                                      const double* __restrict__ G, const double* __restrict__ phi,
                                                      const double* __restrict__ Q, const double* __restrict__ w,
                                                      double* __restrict__ buf) {
  static_assert(NV == 4 || (NV == 6 && VISC), "four variables, or SST with viscous terms");
  constexpr bool SST = NV == 6;
  constexpr int NGS = VISC ? 10 : 4;
  constexpr int ROW = vjp_row<NV, VISC>();
  const int f = xcd_block() * FLUX_BS + threadIdx.x;
  if (f >= m.F) return;
  const int p = m.pitch;
  const int o = m.own[f], n = m.nei[f];
  double* rw[2] = {buf + (size_t)f * ROW, buf + ((size_t)m.F + f) * ROW};
  double* rq[2] = {buf + (size_t)2 * m.F * ROW + (size_t)2 * f, buf + (size_t)2 * m.F * ROW + (size_t)2 * (m.F + f)};
  if (n < 0) {
#pragma unroll
    for (int q = 0; q < NV + NGS; ++q) rw[0][q] = rw[1][q] = 0.0;
    if (SST) rq[0][0] = rq[0][1] = rq[1][0] = rq[1][1] = 0.0;
    if constexpr (VISC == 0) {
      // Euler boundary face (no reconstruction there): F A of the owner's
      // state alone, the wall pressure flux or the Riemann flux against the
      // ghost state, whose Jacobian comes from 4-direction duals; lambda = w_o
      const double nx = m.nx[f], ny = m.ny[f], A = m.area[f];
      const int tg = m.tag[f];
      S6<double> L;
#pragma unroll
      for (int k = 0; k < NV; ++k) L[k] = U[k * p + o];
      double Fb[4], Lb[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 4; ++k) Fb[k] = w[k * p + o] * A;
      if (tg == 3 || tg == 4) {  // (0, p nx, p ny, 0)
        prim_uvp_vjp(P, L, dv(L[1], L[0]), dv(L[2], L[0]), Fb[1] * nx + Fb[2] * ny, 0.0, 0.0, Lb);
      } else {
        typedef DualV<double, 4> T4;
        S6<T4> Ld;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          Ld[k] = T4(L[k]);
#pragma unroll
          for (int d = 0; d < 4; ++d) Ld[k].d[d] = k == d ? 1.0 : 0.0;
        }
        const S6<T4> Rg = ghost(P, Ld, tg, nx, ny, 0.0);
        S6<double> R;
#pragma unroll
        for (int k = 0; k < 4; ++k) R[k] = Rg[k].v;
        double Rb[4] = {0.0, 0.0, 0.0, 0.0};
        riemann_vjp(P, L, R, nx, ny, (m.alt && m.alt[f]) ? P.flux_alt : P.flux, Fb, Lb, Rb);
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int k = 0; k < 4; ++k) Lb[j] += Rb[k] * Rg[k].d[j];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) rw[0][k] = Lb[k];
    }
    return;
  }
  const FaceGeo g = face_geo(m, P, phi, f, o, n, true);
  const double dx[2] = {g.xm - g.xco, g.xm - g.xcn}, dy[2] = {g.ym - g.yco, g.ym - g.ycn};
  const double ph[2] = {g.phL, g.phR};
  double lam[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) lam[k] = w[k * p + o] - w[k * p + n];
  // cell primitives and reconstructed face primitives, as flux_core forms them
  double rho[2], uc[2], vc[2], Wf[2][4], Uc[2][NV], kw[2][2];
  bool clip_r[2], clip_p[2];
#pragma unroll
  for (int sd = 0; sd < 2; ++sd) {
    const int c = sd ? n : o;
    const S6<double> Us(U[c], U[p + c], U[2 * p + c], U[3 * p + c], SST ? U[4 * p + c] : 0.0,
                        SST ? U[5 * p + c] : 0.0);
#pragma unroll
    for (int k = 0; k < NV; ++k) Uc[sd][k] = Us[k];
    kw[sd][0] = SST ? dv(Us[4], Us[0]) : 0.0;  // the upwind values of the k and omega fluxes
    kw[sd][1] = SST ? dv(Us[5], Us[0]) : 0.0;
    double gq[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) gq[q] = G[(size_t)q * p + c];
    rho[sd] = Us[0];
    uc[sd] = dv(Us[1], Us[0]);
    vc[sd] = dv(Us[2], Us[0]);
    const double pc = pressure(P, Us);
    double rf = Us[0] + ph[sd] * (gq[0] * dx[sd] + gq[1] * dy[sd]);
    double pf = pc + ph[sd] * (gq[6] * dx[sd] + gq[7] * dy[sd]);
    const double uf = uc[sd] + ph[sd] * (gq[2] * dx[sd] + gq[3] * dy[sd]);
    const double vf = vc[sd] + ph[sd] * (gq[4] * dx[sd] + gq[5] * dy[sd]);
    clip_r[sd] = rf < 1e-4;
    clip_p[sd] = pf < 1e-4;
    if (clip_r[sd]) rf = Us[0];
    if (clip_p[sd]) pf = pc;
    Wf[sd][0] = rf;
    Wf[sd][1] = uf;
    Wf[sd][2] = vf;
    Wf[sd][3] = pf;
  }
  // lambda^T d(F A) / d(face primitives) by the reverse sweep of the Riemann
  // flux and of prim2cons; with SST the mass flux also carries k and omega
  // (its adjoint gains lambda_k k_up + lambda_w w_up, and its sign picks the
  // upwind cell)
  double aW[2][4] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}}, mflux = 0.0;
  {
    const S6<double> Lr = prim2cons(P, Wf[0][0], Wf[0][1], Wf[0][2], Wf[0][3]);
    const S6<double> Rr = prim2cons(P, Wf[1][0], Wf[1][1], Wf[1][2], Wf[1][3]);
    double lkw = 0.0;
    if (SST) {
      mflux = riemann(P, Lr, Rr, g.nx, g.ny, g.scheme)[0];
      const int up = mflux > 0.0 ? 0 : 1;
      lkw = lam[4] * kw[up][0] + lam[5] * kw[up][1];
    }
    const double Fb[4] = {(lam[0] + lkw) * g.A, lam[1] * g.A, lam[2] * g.A, lam[3] * g.A};
    double Lb[4] = {0.0, 0.0, 0.0, 0.0}, Rb[4] = {0.0, 0.0, 0.0, 0.0};
    riemann_vjp(P, Lr, Rr, g.nx, g.ny, g.scheme, Fb, Lb, Rb);
    prim2cons_vjp(P, Wf[0][0], Wf[0][1], Wf[0][2], Lb, aW[0]);
    prim2cons_vjp(P, Wf[1][0], Wf[1][1], Wf[1][2], Rb, aW[1]);
  }
  // back through the reconstruction and the primitive map
  const double gm1 = P.gamma - 1.0;
  double ru[2][NV], gsl[2][4];  // state rows; gradient slots 2-5 (the viscous part adds to them)
#pragma unroll
  for (int sd = 0; sd < 2; ++sd) {
    double* rs = rw[sd];
    const double wx = ph[sd] * dx[sd], wy = ph[sd] * dy[sd];
    const double ar = aW[sd][0], au = aW[sd][1], av = aW[sd][2], ap = aW[sd][3];
    if constexpr (VISC != 0) {
      rs[NV + 0] = clip_r[sd] ? 0.0 : ar * wx;
      rs[NV + 1] = clip_r[sd] ? 0.0 : ar * wy;
      gsl[sd][0] = au * wx;
      gsl[sd][1] = au * wy;
      gsl[sd][2] = av * wx;
      gsl[sd][3] = av * wy;
      rs[NV + 6] = clip_p[sd] ? 0.0 : ap * wx;
      rs[NV + 7] = clip_p[sd] ? 0.0 : ap * wy;
    } else {  // the face primitives' adjoints: k_vjp_gather forms the slots a (wx, wy)
      rs[NV + 0] = clip_r[sd] ? 0.0 : ar;
      rs[NV + 1] = au;
      rs[NV + 2] = av;
      rs[NV + 3] = clip_p[sd] ? 0.0 : ap;
    }
    const double u = uc[sd], v = vc[sd], ir = 1.0 / rho[sd];
    ru[sd][0] = ar - au * u * ir - av * v * ir + ap * gm1 * 0.5 * (u * u + v * v);
    ru[sd][1] = au * ir - ap * gm1 * u;
    ru[sd][2] = av * ir - ap * gm1 * v;
    ru[sd][3] = ap * gm1;
  }
  if constexpr (SST) {  // the upwind specific k and omega: d(U4 / U0), d(U5 / U0)
    const double mp = mflux > 0.0 ? mflux : 0.0, mm = mflux > 0.0 ? 0.0 : mflux;
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
      const double mf = (sd == 0 ? mp : mm) * g.A, ir = 1.0 / rho[sd];
      const double ak = lam[4] * mf, aw = lam[5] * mf;
      ru[sd][0] = ru[sd][0] - (ak * kw[sd][0] + aw * kw[sd][1]) * ir;
      ru[sd][4] = ak * ir;
      ru[sd][5] = aw * ir;
    }
  }
  if constexpr (VISC != 0) {
    // the viscous terms' adjoints of (L, R, averaged gradients[, mu_t and F1
    // of both cells]) by their reverse sweep, seeded with -lambda A
    constexpr int NIN = 2 * NV + 6 + (SST ? 4 : 0);
    const int gq[6] = {2, 3, 4, 5, 8, 9};
    double gbv[6], tqv[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 6; ++q) gbv[q] = 0.5 * (G[(size_t)gq[q] * p + o] + G[(size_t)gq[q] * p + n]);
    if (SST) {
      tqv[0] = Q[o];
      tqv[1] = Q[n];
      tqv[2] = Q[p + o];
      tqv[3] = Q[p + n];
    }
    S6<double> Ls, Rs;
    double Vb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      Ls[k] = Uc[0][k];
      Rs[k] = Uc[1][k];
      if (k > 0) Vb[k] = -lam[k] * g.A;
    }
    double Lb[6], Rb[6], gbb[6], tqb[4] = {0.0, 0.0, 0.0, 0.0};
    visc_interior_vjp<SST>(P, g, Ls, Rs, gbv, tqv, Vb, Lb, Rb, gbb, tqb);
    double ain[NIN];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      ain[k] = Lb[k];
      ain[NV + k] = Rb[k];
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) ain[2 * NV + q] = gbb[q];
    if (SST) {
#pragma unroll
      for (int q = 0; q < 4; ++q) ain[2 * NV + 6 + q] = tqb[q];
    }
    // each side: its state rows, half of each averaged component (slots 2-5
    // add to the reconstruction's; 8 and 9 are the viscous terms' alone) and,
    // with SST, its own cell's mu_t and F1
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
      double* rs = rw[sd];
#pragma unroll
      for (int k = 0; k < NV; ++k) ru[sd][k] += ain[NV * sd + k];
#pragma unroll
      for (int q = 0; q < 4; ++q) rs[NV + 2 + q] = gsl[sd][q] + 0.5 * ain[2 * NV + q];
      rs[NV + 8] = 0.5 * ain[2 * NV + 4];
      rs[NV + 9] = 0.5 * ain[2 * NV + 5];
      if (SST) {
        rq[sd][0] = ain[2 * NV + 6 + sd];
        rq[sd][1] = ain[2 * NV + 8 + sd];
      }
    }
  }
#pragma unroll
  for (int sd = 0; sd < 2; ++sd)
#pragma unroll
    for (int k = 0; k < NV; ++k) rw[sd][k] = ru[sd][k];
}


// Per cell: the sums of its own side's rows of the adjoint face buffer over
// its faces, in the cell's slot order (zeros past the last cell, up to the
// pitch), and with SST the cell terms' part (the reverse sweep of turb_cell:
// the cell's source adjoint -V w and the summed mu_t and F1 adjoints, which
// are not stored).  They are the first values of the cell's state adjoint
// (aU) and its 14 gradient adjoints (aG): the kernels that follow add to them.
template <int NV, int VISC>
__global__ __launch_bounds__(BS) void k_vjp_gather(DMesh m, Phys P, const double* __restrict__ buf,
                                                   const double* __restrict__ U, const double* __restrict__ G,
                                                   const double* __restrict__ phi, const double* __restrict__ w,
                                                   double* __restrict__ aU, double* __restrict__ aG) {
  constexpr bool SST = NV == 6;
  constexpr int NGS = VISC ? 10 : 4;
  constexpr int ROW = vjp_row<NV, VISC>();
  const int c = xcd_block() * BS + threadIdx.x;
  const int p = m.pitch;
  if (c >= p) return;
  double xc = 0.0, yc = 0.0, ph = 0.5;  // the reconstruction's vector: ph (x_m - x_c), as face_geo forms it
  if (VISC == 0 && c < m.N) {
    xc = m.xc[c];
    yc = m.yc[c];
    if (P.spatial == CFD_SECOND_ORDER_LIMITED) ph = phi[c];
  }
  const int4 s4 = c < m.N ? m.slots[c] : make_int4(-1, -1, -1, -1);  // the padding cells get zeros
  const int sl[4] = {s4.x, s4.y, s4.z, s4.w};
  double su[NV], sg[14], sq[2] = {0.0, 0.0};
#pragma unroll
  for (int k = 0; k < NV; ++k) su[k] = 0.0;
#pragma unroll
  for (int q = 0; q < 14; ++q) sg[q] = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int e = sl[j];
    if (e < 0) continue;
    const size_t sf = (size_t)(e & 1) * m.F + (e >> 1);
    const double* rs = buf + sf * ROW;
#pragma unroll
    for (int k = 0; k < NV; ++k) su[k] += rs[k];
    if constexpr (VISC != 0) {
#pragma unroll
      for (int q = 0; q < NGS; ++q) sg[q] += rs[NV + q];
    } else {
      const int f = e >> 1;
      const double wx = ph * (m.xm[f] - xc), wy = ph * (m.ym[f] - yc);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        sg[2 * q] += rs[NV + q] * wx;
        sg[2 * q + 1] += rs[NV + q] * wy;
      }
    }
    if constexpr (SST) {
      const double* rq = buf + (size_t)2 * m.F * ROW + 2 * sf;
      sq[0] += rq[0];
      sq[1] += rq[1];
    }
  }
  if constexpr (SST) {
    if (P.turb && c < m.N) {
      S6<double> Uc;
#pragma unroll
      for (int j = 0; j < NV; ++j) Uc[j] = U[j * p + c];
      double g[14];
#pragma unroll
      for (int q = 0; q < 14; ++q) g[q] = G[(size_t)q * p + c];
      const double V = m.vol[c];
      double Ub[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
      turb_cell_vjp(P, Uc, g, m.wd[c], -V * w[4 * p + c], -V * w[5 * p + c], sq[0], sq[1], Ub, sg);
      su[0] += Ub[0];
      su[4] += Ub[4];
      su[5] += Ub[5];
    }
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) aU[k * p + c] = su[k];
#pragma unroll
  for (int q = 0; q < 14; ++q)  // the k and omega rows (10-13) exist with SST only
    if (q < 10 || (SST && P.turb)) aG[(size_t)q * p + c] = sg[q];
}

// SST cell terms: w_c^T d(-V S)/d(U_c, G_c), plus the face sweep's adjoints
// of the cell's mu_t and F1 (aQ rows 0 and 1) times d(mu_t, F1)/d(U_c, G_c),
// onto aU_c and aG_c (one thread per cell), by the reverse sweep of
// turb_cell (five passes of 4-direction duals took 199 us per product on the
// 1.32M-cell plate).
template <int NV>
__global__ __launch_bounds__(BS) void k_vjp_src(DMesh m, Phys P, const double* __restrict__ U,
                                                 const double* __restrict__ G, const double* __restrict__ w,
                                                 const double* __restrict__ aQ, double* __restrict__ aU,
                                                 double* __restrict__ aG) {
  static_assert(NV == 6, "SST only");
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= m.N) return;
  const int p = m.pitch;
  const double V = m.vol[c];
  S6<double> Uc;
#pragma unroll
  for (int j = 0; j < NV; ++j) Uc[j] = U[j * p + c];
  double g[14];
#pragma unroll
  for (int q = 0; q < 14; ++q) g[q] = G[(size_t)q * p + c];
  double Ub[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, gb[14];
#pragma unroll
  for (int q = 0; q < 14; ++q) gb[q] = 0.0;
  turb_cell_vjp(P, Uc, g, m.wd[c], -V * w[4 * p + c], -V * w[5 * p + c], aQ[c], aQ[p + c], Ub, gb);
  aU[c] += Ub[0];
  aU[(size_t)4 * p + c] += Ub[4];
  aU[(size_t)5 * p + c] += Ub[5];
  const int rows[8] = {2, 3, 4, 5, 10, 11, 12, 13};
#pragma unroll
  for (int q = 0; q < 8; ++q) aG[(size_t)rows[q] * p + c] += gb[rows[q]];
}

template <class T> __device__ __forceinline__ void prims(const Phys& P, const S6<T>& U, T* q, bool turb) {
  q[0] = U[0];
  q[1] = U[1] / U[0];
  q[2] = U[2] / U[0];
  q[3] = pressure(P, U);
  q[4] = q[3] / (U[0] * P.R);
  if (turb) {
    q[5] = U[4] / U[0];
    q[6] = U[5] / U[0];
  }
}

// Green-Gauss transpose, one thread per cell: the gradients are face means
// of the primitives q over the cell's faces, so with the face mean's adjoint
// aq_f = A (aG_o . n / V_o - aG_n . n / V_n) (aG holds the 2 slots of each
// primitive), aU_c += 1/2 (sum_f aq_f)^T dq(U_c)/dU_c, written out for
// q = (rho, u, v, p, T[, k, omega]); no write to another cell, so a single
// launch.  The product's time term (vol / dt) w (Solver.cpp:1381-1384) is
// added here.  A boundary face's mean also holds the ghost state's
// primitives: k_vjp_gg_ghost adds that part.
// The cell's own share, sum_f s_f A_f n_f . aG_c / V_c over all its faces
// (s_f = +1 as owner, -1 as neighbour), is left out: the outward normals
// times the areas of a closed cell sum to zero, so the share is rounding
// (about 1e-16 of the other terms), and the kernel need not read the cell's
// own gradient adjoints (products: 1.049 -> 1.039 ms on the 1.32M-cell SST
// plate, 0.428 -> 0.424 ms on the 824k-cell bump, 2.53 -> 2.50 ms on the
// 4.4M-cell SWBLI).
template <int NV>
__global__ __launch_bounds__(BS) void k_vjp_gg(DMesh m, Phys P, const double* __restrict__ U,
                                                const double* __restrict__ aG, double* __restrict__ aU,
                                                const double* __restrict__ w, const double* __restrict__ dt) {
  constexpr int NQM = NV == 6 ? 7 : 5;
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= m.N) return;
  const int p = m.pitch;
  const bool turb = NV == 6 && P.turb;
  const int nq = turb ? 7 : 5;
  const double vc = m.vol[c];
  double S[NQM];
#pragma unroll
  for (int k = 0; k < NQM; ++k) S[k] = 0.0;
  const int4 s4 = m.slots[c], n4 = m.cnb[c];
  const int sl[4] = {s4.x, s4.y, s4.z, s4.w}, on[4] = {n4.x, n4.y, n4.z, n4.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = sl[q];
    if (e < 0) continue;
    const int f = e >> 1;
    const bool as_nei = e & 1;
    const int other = on[q];
    const double nx = m.nx[f], ny = m.ny[f], A = m.area[f];
    const double vb = other >= 0 ? m.vol[other] : 1.0;
#pragma unroll
    for (int k = 0; k < NQM; ++k) {
      if (k >= nq || other < 0) continue;
      const double oth_part =
          A * dv(aG[(size_t)(2 * k) * p + other] * nx + aG[(size_t)(2 * k + 1) * p + other] * ny, vb);
      // aq = A (aG_o . n / V_o - aG_n . n / V_n), the cell's own term left out
      S[k] += as_nei ? oth_part : -oth_part;
    }
  }
  // a = S / 2 through dq/dU: u = U1/rho, v = U2/rho, p = (gamma-1)(U3 -
  // (U1^2 + U2^2)/(2 rho)), T = p / (rho R), k = U4/rho, omega = U5/rho
  const double r = U[c], u = dv(U[p + c], r), v = dv(U[2 * p + c], r);
  const double pr = pressure(P, S6<double>(r, U[p + c], U[2 * p + c], U[3 * p + c]));
  const double gm1 = P.gamma - 1.0, rR = r * P.R, Tc = dv(pr, rR);
  const double ar = 0.5 * S[0], au = 0.5 * S[1], av = 0.5 * S[2], aT = 0.5 * S[4];
  const double ap = 0.5 * S[3] + dv(aT, rR);  // p's adjoint, with T's share
  double wk = 0.0, k_ = 0.0, w_ = 0.0;
  if (NV == 6 && turb) {
    k_ = dv(U[4 * p + c], r);
    w_ = dv(U[5 * p + c], r);
    wk = 0.5 * S[5] * k_ + 0.5 * S[6] * w_;
  }
  double s[NV];
  s[0] = ar - dv(au * u + av * v + wk + aT * Tc, r) + ap * gm1 * 0.5 * (u * u + v * v);
  s[1] = dv(au, r) - ap * gm1 * u;
  s[2] = dv(av, r) - ap * gm1 * v;
  s[3] = ap * gm1;
  if (NV == 6) {
    s[4] = turb ? dv(0.5 * S[5], r) : 0.0;
    s[5] = turb ? dv(0.5 * S[6], r) : 0.0;
  }
  const double tc = vc / (dt[c] + 1e-30);
#pragma unroll
  for (int j = 0; j < NV; ++j) aU[(size_t)j * p + c] = (aU[(size_t)j * p + c] + s[j]) + w[(size_t)j * p + c] * tc;
}

// The ghost state's part of k_vjp_gg for the boundary faces:
// aU_o += 1/2 aq_f^T dq(ghost(U_o))/dU_o, aq_f = A aG_o . n / V_o.  One thread
// per (cell with boundary faces, state row j), over its boundary faces in the
// order of their colors: a 1-direction dual along U_o[j] gives row j, so each
// thread writes its own row, and one launch serves every boundary face.
template <int NV>
__device__ __forceinline__ void vjp_ghost_row(const DMesh& m, const Phys& P, const double* __restrict__ U,
                                              const double* __restrict__ aG, double* __restrict__ aU, int f, int j) {
  typedef DualV<double, 1> T;
  constexpr int NQM = NV == 6 ? 7 : 5;
  const int c = m.own[f];
  const int p = m.pitch;
  const bool turb = NV == 6 && P.turb;
  const int nq = turb ? 7 : 5;
  const double nx = m.nx[f], ny = m.ny[f], A = m.area[f], vc = m.vol[c];
  S6<T> Uc;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    Uc[k] = T(U[k * p + c]);
    Uc[k].d[0] = k == j ? 1.0 : 0.0;
  }
  T qg[NQM];
  prims(P, ghost(P, Uc, m.tag[f], nx, ny, turb ? m.dnw[f] : 0.0), qg, turb);
  double ag = 0.0;
#pragma unroll
  for (int k = 0; k < NQM; ++k) {
    if (k >= nq) continue;
    const double aq = A * dv(aG[(size_t)(2 * k) * p + c] * nx + aG[(size_t)(2 * k + 1) * p + c] * ny, vc);
    ag += aq * 0.5 * qg[k].d[0];
  }
  aU[(size_t)j * p + c] += ag;
}
template <int NV>
__global__ __launch_bounds__(BS) void k_vjp_gg_ghost(DMesh m, Phys P, const double* __restrict__ U,
                                                      const double* __restrict__ aG, double* __restrict__ aU,
                                                      const int4* __restrict__ grp, int ngrp) {
  const int t = blockIdx.x * BS + threadIdx.x;
  const int gi = t / NV, j = t % NV;
  if (gi >= ngrp) return;
  const int4 g4 = grp[gi];
  const int fs[4] = {g4.x, g4.y, g4.z, g4.w};
  for (int k = 0; k < 4; ++k) {
    if (fs[k] < 0) break;
    vjp_ghost_row<NV>(m, P, U, aG, aU, fs[k], j);
  }
}

template <class F>
__global__ void k_fill(F* __restrict__ x, F v, size_t n) {
  size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) x[i] = v;
}

// ---- wall distance (Solver.cpp:1981-2015) -------------------------------------------
// min over the no-slip face midpoints of the centroid distance; wall points are
// staged through LDS in tiles shared by the block.
__global__ __launch_bounds__(BS) void k_walldist(int N, const double* __restrict__ xc, const double* __restrict__ yc,
                                                  const double* __restrict__ wxy, int nw, double* __restrict__ wd) {
  __shared__ double sx[BS], sy[BS];
  const int c = blockIdx.x * BS + threadIdx.x;
  const double cx = c < N ? xc[c] : 0.0, cy = c < N ? yc[c] : 0.0;
  double md = 1e20;
  for (int t0 = 0; t0 < nw; t0 += BS) {
    __syncthreads();
    if (t0 + (int)threadIdx.x < nw) {
      sx[threadIdx.x] = wxy[2 * (t0 + threadIdx.x)];
      sy[threadIdx.x] = wxy[2 * (t0 + threadIdx.x) + 1];
    }
    __syncthreads();
    const int cnt = min(BS, nw - t0);
    for (int j = 0; j < cnt; ++j) {
      const double dx = cx - sx[j], dy = cy - sy[j];
      const double d2 = dx * dx + dy * dy;
      if (d2 < md) md = d2;
    }
  }
  if (c < N) wd[c] = nw > 0 ? sqrt(md) : 1e10;
}

// ---- local time step (Solver.cpp:1388-1422) ----------------------------------------
template <int NV>
__global__ __launch_bounds__(BS) void k_timestep(DMesh m, Phys P, const double* __restrict__ U, double cfl,
                                                  double* __restrict__ dt) {
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= m.N) return;
  const S6<double> Uc = ld<NV>(U, m.pitch, c);
  const double pc = pressure(P, Uc);
  const double cc = sqrt(P.gamma * pc / Uc[0]);
  double lam = 1e-12;
  const int4 s4 = m.slots[c];
  const int sl[4] = {s4.x, s4.y, s4.z, s4.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int e = sl[k];
    if (e < 0) continue;
    const int f = e >> 1;
    double nx = m.nx[f], ny = m.ny[f];
    if (e & 1) {
      nx = -nx;
      ny = -ny;
    }
    const double un = (Uc[1] * nx + Uc[2] * ny) / Uc[0];
    lam += (fabs(un) + cc) * m.area[f];
  }
  dt[c] = cfl * m.vol[c] / lam;
}

// ---- scalar diagonal preconditioner (Solver.cpp:1424-1465) --------------------------
template <int NV>
__global__ __launch_bounds__(BS) void k_diagprec(DMesh m, Phys P, const double* __restrict__ U,
                                                  const double* __restrict__ dt, double* __restrict__ D) {
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= m.N) return;
  const int p = m.pitch;
  const double g = P.gamma;
  double acc = 0.0;
  const int4 s4 = m.slots[c];
  const int sl[4] = {s4.x, s4.y, s4.z, s4.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int e = sl[k];
    if (e < 0) continue;
    const int f = e >> 1;
    const int o = m.own[f], n = m.nei[f];
    const double nx = m.nx[f], ny = m.ny[f];
    const S6<double> L = ld<NV>(U, p, o);
    const S6<double> R = n >= 0 ? ld<NV>(U, p, n) : ghost(P, L, m.tag[f], nx, ny);
    const double cL = sqrt(g * pressure(P, L) / L[0]);
    const double pRr = pressure(P, R);
    const double cR = sqrt(g * (pRr > 0 ? pRr : 1e-3) / (R[0] > 0 ? R[0] : 1e-3));
    const double cf = 0.5 * (cL + cR);
    const double uf = 0.5 * (L[1] / L[0] + R[1] / R[0]);
    const double vf = 0.5 * (L[2] / L[0] + R[2] / R[0]);
    const double vn = fabs(uf * nx + vf * ny);
    const double li = vn + cf;
    const double dx = sqrt(m.vol[o]);
    const double mu = 0.5 * (L[0] + R[0]) * P.nu;
    const double lv = 2.0 * mu / (0.5 * (L[0] + R[0]) * dx + 1e-12);
    const double lt = (li + lv) * m.area[f];
    acc += lt * 0.5;
  }
  double d = (m.vol[c] / (dt[c] + 1e-30)) + acc;
  if (d < 1e-12) d = 1e-12;
  D[c] = d;
}

// ---- block-Jacobi preconditioner ------------------------------------------------------
// Diagonal block of the first-order Jacobian of cell c:
//   B_c = (V/dtau) I + sum_f (+-) A dF(U_L, U_R)/dU_c + nu A / d I,
// each column k formed by a forward-mode dual with direction e_k on U_c (the
// ghost state of a boundary face is differentiated with it), d the normal
// distance between the centroids (twice the centroid-face distance at a
// boundary).  B_c is inverted by Gauss-Jordan with partial pivoting in FP64
// and stored as TV, NV*NV rows of pitch.  With Aoff the coupling blocks of
// the same first-order Jacobian to the interior neighbors are stored too,
// A_{c,q} = (+-) A dF/dU_other - nu A / d I for slot q of cell c, in rows
// (q * NV + i) * NV + j of pitch (slots without a neighbor are not written).
// Face Jacobians of the first-order inviscid flux (times A): dF/dU_L and, on
// interior faces, dF/dU_R, one forward-dual pass per direction.  Entry
// e = s * NV^2 + i * NV + k (s = 0 for L, 1 for R) of face f sits at
// jf_index: faces in tiles of 64, entry-major inside a tile, so that a wave's
// stores are contiguous and the cell kernel's gather of a face stays within
// one tile.  On a boundary face the ghost state is differentiated with U_L.
// Each face is differentiated once; the cell kernel below assembles both of
// its cells.
template <int NV> __device__ __forceinline__ size_t jf_index(int f, int e) {
  return ((size_t)(f >> 6) * (2 * NV * NV) + e) * 64 + (f & 63);
}
// TV is the precision of the blocks: FP32 blocks take FP32 duals (the
// preconditioner of the FP32 Krylov path needs no more), FP64 blocks FP64 ones.
template <int NV, class TV>
__global__ __launch_bounds__(BS) void k_face_jac(DMesh m, Phys P, const double* __restrict__ U,
                                                  TV* __restrict__ Jf) {
  typedef Dual<TV> D;
  const int f = xcd_block() * BS + threadIdx.x;
  if (f >= m.F) return;
  const int p = m.pitch;
  Phys P1 = P;
  P1.spatial = CFD_FIRST_ORDER;
  const int o = m.own[f], n = m.nei[f];
  const FaceGeo g = face_geo(m, P1, nullptr, f, o, n, false);
  const S6<double> UL = ld<NV>(U, p, o);
  const S6<double> UR = n >= 0 ? ld<NV>(U, p, n) : UL;
  const int nside = n >= 0 ? 2 : 1;
#pragma unroll 1
  for (int sk = 0; sk < NV * nside; ++sk) {
    const int side = sk / NV, k = sk - side * NV;
    S6<D> Ld, Rd;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      Ld[i] = D(TV(UL[i]), TV(side == 0 && i == k ? 1.0 : 0.0));
      Rd[i] = D(TV(UR[i]), TV(side == 1 && i == k ? 1.0 : 0.0));
    }
    if (n < 0) Rd = ghost(P1, Ld, g.tg, g.nx, g.ny, g.dnw);
    const S6<D> Fl = flux_core<0, D>(P1, g, Ld, Rd, nullptr, nullptr);
#pragma unroll
    for (int i = 0; i < NV; ++i) Jf[jf_index<NV>(f, side * NV * NV + i * NV + k)] = Fl[i].d;
  }
}

template <int NV, class TV>
__global__ __launch_bounds__(BS) void k_block(DMesh m, Phys P, const double* __restrict__ U,
                                               const double* __restrict__ G, const double* __restrict__ dt,
                                               const TV* __restrict__ Jf, TV* __restrict__ Binv,
                                               double* __restrict__ Bout, TV* __restrict__ Aoff,
                                               const double* __restrict__ Q, const int* __restrict__ pos) {
  typedef Dual<TV> D;  // SST source Jacobian in the precision of the blocks
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= m.N) return;
  const int p = m.pitch;
  const int oc = pos ? pos[c] : c;  // output column: color-major position for the block SGS
  double B[NV][NV];
#pragma unroll
  for (int i = 0; i < NV; ++i)
#pragma unroll
    for (int j = 0; j < NV; ++j) B[i][j] = 0.0;
  const S6<double> Uc = ld<NV>(U, p, c);
  const int4 s4 = m.slots[c];
  const int sl[4] = {s4.x, s4.y, s4.z, s4.w};
  double visc[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) visc[i] = 0.0;
  for (int q = 0; q < 4; ++q) {
    const int e = sl[q];
    if (e < 0) continue;
    const int f = e >> 1;
    const bool as_nei = e & 1;
    const int o = m.own[f], n = m.nei[f];
    const int other = as_nei ? o : n;
    // this cell's side of the face Jacobian, with the sign of its residual
    // (R_o += F A, R_n -= F A); the other side couples to the neighbor
    const double sg = as_nei ? -1.0 : 1.0;
    const int es = as_nei ? NV * NV : 0;
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int k = 0; k < NV; ++k) B[i][k] += sg * (double)Jf[jf_index<NV>(f, es + i * NV + k)];
    // viscous diagonal: the two-point diffusion coefficient of each row,
    // (nu + nu_t) for momentum, gamma (nu / Pr + nu_t / Pr_t) for energy,
    // (nu + sigma nu_t) for k and omega, times A over the normal distance
    // (interior: between the centroids; no-slip wall: centroid to face)
    double vf[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) vf[i] = 0.0;
    if (P.viscous && (other >= 0 || m.tag[f] == 4)) {
      const double nx = m.nx[f], ny = m.ny[f];
      double d, nut = 0.0, F1 = 0.0;
      if (other >= 0) {
        d = fabs((m.xc[other] - m.xc[c]) * nx + (m.yc[other] - m.yc[c]) * ny);
        if (Q) {
          nut = (Q[c] + Q[other]) / (Uc[0] + U[other]);
          F1 = 0.5 * (Q[p + c] + Q[p + other]);
        }
      } else {
        d = fabs((m.xm[f] - m.xc[c]) * nx + (m.ym[f] - m.yc[c]) * ny);
        if (Q) F1 = Q[p + c];
      }
      const double a = m.area[f] / fmax(d, 1e-300);
      vf[1] = vf[2] = (P.nu + nut) * a;
      vf[3] = P.gamma * (P.nu / P.Pr + nut / sst::Prt) * a;
      if constexpr (NV == 6) {
        vf[4] = (P.nu + (F1 * sst::sk1 + (1.0 - F1) * sst::sk2) * nut) * a;
        vf[5] = (P.nu + (F1 * sst::sw1 + (1.0 - F1) * sst::sw2) * nut) * a;
      }
#pragma unroll
      for (int i = 0; i < NV; ++i) visc[i] += vf[i];
    }
    if (Aoff && other >= 0) {
      const int eo = as_nei ? 0 : NV * NV;
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int k = 0; k < NV; ++k)
          Aoff[(size_t)((q * NV + i) * NV + k) * p + oc] =
              (TV)(sg * (double)Jf[jf_index<NV>(f, eo + i * NV + k)] - (i == k ? vf[i] : 0.0));
    }
  }
  if constexpr (NV == 6) {
    if (P.turb) {  // source Jacobian with the gradients frozen
      D gd[14];
#pragma unroll
      for (int q = 0; q < 14; ++q) gd[q] = D(TV(G[(size_t)q * p + c]), TV(0));
      const double V = m.vol[c];
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        S6<D> Ud;
#pragma unroll
        for (int i = 0; i < 6; ++i) Ud[i] = D(TV(Uc[i]), TV(i == k ? 1.0 : 0.0));
        const TurbQ<D> tq = turb_cell(P, Ud, gd, m.wd[c]);
        B[4][k] -= V * (double)tq.Sk.d;
        B[5][k] -= V * (double)tq.Sw.d;
      }
    }
  }
  const double tc = m.vol[c] / (dt[c] + 1e-30);
#pragma unroll
  for (int i = 0; i < NV; ++i) B[i][i] += tc + visc[i];
  if (Bout) {
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int j = 0; j < NV; ++j) Bout[(size_t)(i * NV + j) * p + oc] = B[i][j];
  }
  // Gauss-Jordan inverse
  double Inv[NV][NV];
#pragma unroll
  for (int i = 0; i < NV; ++i)
#pragma unroll
    for (int j = 0; j < NV; ++j) Inv[i][j] = i == j ? 1.0 : 0.0;
  // fully unrolled with the row swap done by compare-and-select, so that
  // every index is static and both matrices stay in registers
#pragma unroll
  for (int col = 0; col < NV; ++col) {
    int piv = col;
    double best = fabs(B[col][col]);
#pragma unroll
    for (int r = col + 1; r < NV; ++r)
      if (fabs(B[r][col]) > best) {
        best = fabs(B[r][col]);
        piv = r;
      }
#pragma unroll
    for (int r = col + 1; r < NV; ++r) {
      if (piv == r) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          double t = B[col][j];
          B[col][j] = B[r][j];
          B[r][j] = t;
          t = Inv[col][j];
          Inv[col][j] = Inv[r][j];
          Inv[r][j] = t;
        }
      }
    }
    const double ip = 1.0 / B[col][col];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      B[col][j] *= ip;
      Inv[col][j] *= ip;
    }
#pragma unroll
    for (int r = 0; r < NV; ++r) {
      if (r == col) continue;
      const double fct = B[r][col];
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        B[r][j] -= fct * B[col][j];
        Inv[r][j] -= fct * Inv[col][j];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NV; ++i)
#pragma unroll
    for (int j = 0; j < NV; ++j) Binv[(size_t)(i * NV + j) * p + oc] = (TV)Inv[i][j];
}

// z = Binv_c x_c (or Binv_c^T x_c for the transposed system) per cell
template <int NV, class TV, class TZ, bool TR>
__global__ __launch_bounds__(BS) void k_block_apply(int N, int p, const TV* __restrict__ Binv,
                                                     const TV* __restrict__ x, TZ* __restrict__ z) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= N) return;
  double xv[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) xv[j] = (double)x[j * p + c];
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NV; ++j) s += (double)Binv[(size_t)(TR ? j * NV + i : i * NV + j) * p + c] * xv[j];
    z[i * p + c] = (TZ)s;
  }
}

// x = Binv (sum_{i<=k} y_i V_i), or Binv^T for the transposed system
template <int NV, class TV, bool TR>
__global__ __launch_bounds__(BS) void k_solution_bj(int N, int p, size_t vstride, const TV* __restrict__ V,
                                                     const double* __restrict__ y, int nvec,
                                                     const TV* __restrict__ Binv, double* __restrict__ x) {
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= N) return;
  double s[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double a = 0.0;
    for (int i = 0; i < nvec; ++i) a = a + (double)V[(size_t)i * vstride + k * p + c] * y[i];
    s[k] = a;
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < NV; ++j) a += (double)Binv[(size_t)(TR ? j * NV + i : i * NV + j) * p + c] * s[j];
    x[i * p + c] = a;
  }
}

// ---- multicolor block symmetric Gauss-Seidel --------------------------------------
// M = (D + L) D^-1 (D + U) of the first-order Jacobian, L and U the couplings
// to neighbors of lower and higher cell color (no two neighbors share a
// color).  M z = x is two sweeps over the colors, each a launch per color of
//   z_c = D_c^-1 (x_c - sum_q A_{c,q} z_{nbr(c,q)}),
// ascending then descending (the last color of the ascending sweep is final
// and not repeated); z starts at zero.  M^T z = x (the transposed system)
// runs the same sweeps with D_c^-T and A_{nbr,q'}^T, q' the slot of the
// neighbor that points back to c.  nq[c] holds (neighbor << 2) | q' per slot,
// -1 where the slot has no interior neighbor.

// The blocks are stored color-major: the blocks of the cell at position t of
// the color-grouped cell list are in column t, so that a color's launch reads
// its own contiguous range instead of every cache line of the arrays.

// first launch: z = D^-1 x on color 0, zero elsewhere
template <int NV, class TB, class TX, class TZ, bool TR>
__global__ __launch_bounds__(BS) void k_sgs_first(int N, int p, const unsigned char* __restrict__ col,
                                                   const int* __restrict__ pos, const TB* __restrict__ Binv,
                                                   const TX* __restrict__ x, TZ* __restrict__ z) {
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= N) return;
  if (col[c] != 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) z[i * p + c] = (TZ)0.0;
    return;
  }
  const int b = pos[c];
  double xv[NV];
#pragma unroll
  for (int j = 0; j < NV; ++j) xv[j] = (double)x[j * p + c];
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NV; ++j) s += (double)Binv[(size_t)(TR ? j * NV + i : i * NV + j) * p + b] * xv[j];
    z[i * p + c] = (TZ)s;
  }
}

// cells [b0, b0 + n) of the color-grouped list; couplings to cells of a color
// above `above` are skipped (their z is zero during the forward sweep)
template <int NV, class TB, class TX, class TZ, bool TR>
__global__ __launch_bounds__(BS) void k_sgs_color(int n, int b0, const int* __restrict__ cells, int p,
                                                   const int4* __restrict__ nq, const int* __restrict__ pos,
                                                   const unsigned char* __restrict__ ccol, int above,
                                                   const TB* __restrict__ Binv, const TB* __restrict__ Aoff,
                                                   const TX* __restrict__ x, TZ* __restrict__ z) {
  const int t = xcd_block() * BS + threadIdx.x;
  if (t >= n) return;
  const int b = b0 + t;
  const int c = cells[b];
  double s[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) s[i] = (double)x[i * p + c];
  const int4 e4 = nq[c];
  const int es[4] = {e4.x, e4.y, e4.z, e4.w};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int e = es[q];
    if (e < 0) continue;
    const int o = e >> 2;
    if (ccol[o] > above) continue;  // forward sweep: z of a later color is still zero
    double zo[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) zo[j] = (double)z[j * p + o];
    if (TR) {
      const int r = e & 3, bo = pos[o];
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) s[i] -= (double)Aoff[(size_t)((r * NV + j) * NV + i) * p + bo] * zo[j];
    } else {
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) s[i] -= (double)Aoff[(size_t)((q * NV + i) * NV + j) * p + b] * zo[j];
    }
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < NV; ++j) a += (double)Binv[(size_t)(TR ? j * NV + i : i * NV + j) * p + b] * s[j];
    z[i * p + c] = (TZ)a;
  }
}

// ---- line-implicit block Gauss-Seidel ----------------------------------------------
// Cells are grouped into segments (chains of strongly coupled cells, at most
// 64, or single cells); segments of one color share no face, and no face joins
// two non-consecutive cells of a segment.  The blocks are stored by line
// position t (the cell's index in the segment list, so a color's cells are a
// contiguous range).  Per color, a sweep solves every segment's
// block-tridiagonal system exactly with the couplings to other segments taken
// from the current z:
//   w_t = Dp_t^-1 (x_t - sum_{off-segment q} A_{t,q} z_nbr),
//   y_t = w_t - P_t y_{t-1},   z_t = y_t - G_t z_{t+1}   (k_line_sweep, one launch per color)
// with Dp_t = B_t - A_{t,prev} G_{t-1}, P_t = Dp_t^-1 A_{t,prev} and
// G_t = Dp_t^-1 A_{t,next} from k_line_factor.  Colors are swept ascending,
// then descending.  A single-cell segment has Dp = B, whose inverse the block
// kernel has already written at its position, and z = w.
//
// Multi-cell segments are packed into bundles of 64 slots, one wavefront
// each: a color's segments, longest first, fill a bundle while they fit, one
// slot per cell in chain order (lb_t[m]: line position of slot m, -1 for
// padding; lb_kl[m] = k | L << 8, the cell's place k in its segment of length
// L).  The recurrences run across the lanes: at step s the lanes holding the
// s-th cell of their segment update, taking the previous cell's value from the
// lane below (the next cell's from the lane above) by a DPP wave shift.  Every
// load is issued before the first step, so a step costs one block product and
// a few lane moves rather than a chain of dependent memory loads.  P and G are kept
// by slot (nv^2 consecutive values per slot).


// Gauss-Jordan inverse with partial pivoting by selects (all indices
// compile-time, the matrix stays in registers); M is destroyed.
template <int NV>
__device__ __forceinline__ void line_inverse(double (&M)[NV][NV], double (&Inv)[NV][NV]) {
#pragma unroll
  for (int i = 0; i < NV; ++i)
#pragma unroll
    for (int j = 0; j < NV; ++j) Inv[i][j] = i == j ? 1.0 : 0.0;
#pragma unroll
  for (int col = 0; col < NV; ++col) {
    int piv = col;
    double best = fabs(M[col][col]);
#pragma unroll
    for (int r = col + 1; r < NV; ++r)
      if (fabs(M[r][col]) > best) {
        best = fabs(M[r][col]);
        piv = r;
      }
#pragma unroll
    for (int r = col + 1; r < NV; ++r) {
      if (piv == r) {
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          double tt = M[col][j];
          M[col][j] = M[r][j];
          M[r][j] = tt;
          tt = Inv[col][j];
          Inv[col][j] = Inv[r][j];
          Inv[r][j] = tt;
        }
      }
    }
    const double ip = 1.0 / M[col][col];
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      M[col][j] *= ip;
      Inv[col][j] *= ip;
    }
#pragma unroll
    for (int r = 0; r < NV; ++r) {
      if (r == col) continue;
      const double f = M[r][col];
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        M[r][j] -= f * M[col][j];
        Inv[r][j] -= f * Inv[col][j];
      }
    }
  }
}

// One wavefront per bundle.  The inputs of every lane (B in FP64, A_prev and
// A_next) are staged in LDS at the start, lane-major so that each lane reads
// only its own column; G of the previous cell arrives by a DPP wave shift.
template <int NV, class TV>
__global__ __launch_bounds__(64) void k_line_factor(const int* __restrict__ lb_t, const int* __restrict__ lb_kl,
                                                    const char2* __restrict__ lslot, int p,
                                                    const int* __restrict__ seg_cells,
                                                    const double* __restrict__ Bd, const TV* __restrict__ Aoff,
                                                    TV* __restrict__ Dpinv, TV* __restrict__ Pm,
                                                    TV* __restrict__ G) {
  constexpr int NN = NV * NV;
  __shared__ double sB[NN][64];
  __shared__ TV sA[2][NN][64];
  const int l = threadIdx.x;
  const size_t bb = (size_t)blockIdx.x * 64, m = bb + l;
  const int t = lb_t[m];
  const int kl = lb_kl[m];
  const int k = kl & 255, L = kl >> 8;
  const int steps = __builtin_amdgcn_readfirstlane(lb_kl[bb] >> 8);  // slot 0 holds a longest segment
  {  // loads from clamped, in-bounds positions (a padding slot reads position 0,
     // a missing neighbour slot 0) so that none is guarded; unused values are
     // never read back
    const int tc = t >= 0 ? t : 0;
    const char2 qs = lslot[seg_cells[tc]];
    const int qp = qs.x >= 0 ? qs.x : 0, qn = qs.y >= 0 ? qs.y : 0;
#pragma unroll
    for (int e = 0; e < NN; ++e) {
      sB[e][l] = Bd[(size_t)e * p + tc];
      sA[0][e][l] = Aoff[(size_t)(qp * NN + e) * p + tc];
      sA[1][e][l] = Aoff[(size_t)(qn * NN + e) * p + tc];
    }
  }
  TV g[NN];
#pragma unroll
  for (int e = 0; e < NN; ++e) g[e] = TV(0);
  for (int s = 0; s < steps; ++s) {
    TV gp[NN];
#pragma unroll
    for (int e = 0; e < NN; ++e) gp[e] = lane_prev(g[e]);
    if (t >= 0 && k == s) {
      double M[NV][NV];
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) M[i][j] = sB[i * NV + j][l];
      if (k > 0) {  // Dp = B - A_prev G_{t-1}
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
          for (int j = 0; j < NV; ++j) {
            double a = 0.0;
#pragma unroll
            for (int q = 0; q < NV; ++q) a = fma((double)sA[0][i * NV + q][l], (double)gp[q * NV + j], a);
            M[i][j] -= a;
          }
      }
      double Inv[NV][NV];
      line_inverse<NV>(M, Inv);
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) Dpinv[(size_t)(i * NV + j) * p + t] = (TV)Inv[i][j];
#pragma unroll
      for (int side = 0; side < 2; ++side) {  // P = Dp^-1 A_prev, G = Dp^-1 A_next
        if (side == 0 ? k == 0 : k == L - 1) continue;
        TV* out = side == 0 ? Pm : G;
#pragma unroll
        for (int i = 0; i < NV; ++i)
#pragma unroll
          for (int j = 0; j < NV; ++j) {
            double a = 0.0;
#pragma unroll
            for (int q = 0; q < NV; ++q) a = fma(Inv[i][q], (double)sA[side][q * NV + j][l], a);
            out[m * NN + i * NV + j] = (TV)a;
            if (side == 1) g[i * NV + j] = (TV)a;
          }
      }
    }
  }
}

// The two recurrences of a bundle, y_t = w_t - P_t y_{t-1} ascending and then
// z_t = y_t - G_t z_{t+1} descending, as scans of affine maps across the
// lanes.  Each lane holds the map y -> b + M y of its cell (b = w, M = -P; the
// first cell of a segment has no map to compose with).  At distance
// d = 1, 2, 4, ... every lane at least d cells into its segment composes its
// map with that of the lane d below: (b, M) <- (b + M b', M M').  After
// ceil(log2 L) levels b is y_t: that many dependent steps instead of L - 1.
// The descending pass mirrors this with M = -G and the lane d above.  reach is
// the number of cells of the segment before (ASC) or after the lane's cell;
// steps, the longest segment of the bundle, bounds the distances.  y holds w
// on entry and z on exit.
template <int NV, bool ASC, class TB, class TC>
__device__ __forceinline__ void line_scan(int steps, int reach, const TB (&A)[NV * NV], TC (&y)[NV]) {
  constexpr int NN = NV * NV;
  TC M[NN];
#pragma unroll
  for (int e = 0; e < NN; ++e) M[e] = -(TC)A[e];
  for (int d = 1; d < steps; d <<= 1) {
    const bool last = 2 * d >= steps;  // the maps are not needed after the last level
    TC bo[NV];
#pragma unroll
    for (int i = 0; i < NV; ++i) bo[i] = ASC ? __shfl_up(y[i], d) : __shfl_down(y[i], d);
    if (last) {
      if (reach >= d) {
        TC bn[NV];
#pragma unroll
        for (int i = 0; i < NV; ++i) {
          TC a = y[i];
#pragma unroll
          for (int j = 0; j < NV; ++j) a = fma(M[i * NV + j], bo[j], a);
          bn[i] = a;
        }
#pragma unroll
        for (int i = 0; i < NV; ++i) y[i] = bn[i];
      }
      break;
    }
    TC Mo[NN];
#pragma unroll
    for (int e = 0; e < NN; ++e) Mo[e] = ASC ? __shfl_up(M[e], d) : __shfl_down(M[e], d);
    if (reach >= d) {
      TC bn[NV], Mn[NN];
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        TC a = y[i];
#pragma unroll
        for (int j = 0; j < NV; ++j) a = fma(M[i * NV + j], bo[j], a);
        bn[i] = a;
      }
#pragma unroll
      for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int j = 0; j < NV; ++j) {
          TC a = TC(0);
#pragma unroll
          for (int q = 0; q < NV; ++q) a = fma(M[i * NV + q], Mo[q * NV + j], a);
          Mn[i * NV + j] = a;
        }
#pragma unroll
      for (int i = 0; i < NV; ++i) y[i] = bn[i];
#pragma unroll
      for (int e = 0; e < NN; ++e) M[e] = Mn[e];
    }
  }
}
// Both passes for a lane at place k of a segment of length L (live: not a
// padding slot)
template <int NV, class TB, class TC>
__device__ __forceinline__ void line_recur(int steps, int k, int L, bool live, const TB (&pr)[NV * NV],
                                           const TB (&gr)[NV * NV], TC (&y)[NV]) {
  line_scan<NV, true, TB, TC>(steps, live ? k : 0, pr, y);
  line_scan<NV, false, TB, TC>(steps, live ? L - 1 - k : 0, gr, y);
}

// Sweep passes.  A color is visited once ascending (forward) and, below the
// top color, once descending (backward).  During the forward pass only
// neighbours of lower colors carry a nonzero z, and during the backward pass
// the lower ones still hold exactly the values the forward pass used, so
//   forward:  s = x - sum_{lower q} A_{t,q} z_nbr   (kept in sL below the top color)
//   backward: s = sL - sum_{higher q} A_{t,q} z_nbr
// reads every coupling block once per application instead of one and a half
// times.  lmask[c] marks the off-segment neighbour slots of lower (bits 0-3)
// and higher (bits 4-7) segment color.
enum LinePass { LINE_FWD = 0, LINE_TOP = 1, LINE_BWD = 2 };

// w_t = Dp_t^-1 s for the cell at line position t (pass as above); keep
// false suppresses the store of s (a padding lane standing in for a cell).
// x is indexed by cell; the sweep's own vectors, the kept sums sL and the
// result z, are indexed by line position (so are the neighbour positions of
// lnc), which keeps their loads and stores coalesced along the segments.
// (Gathering x into line order first, in a kernel of its own, cost more than
// it saved: 42 / 62 us per application in FP32 / FP64 on the 1.32M-cell
// plate, against 60 / 102 us per color launch.)  The off-segment
// couplings come from the packed slots of Ac: the forward pass takes slots
// 0, 1, ... (lower colors), the backward pass slots 3, 2, ... (higher), so
// that every lane of a wavefront reads the same slot, in the cells'
// ascending slot order as before.  FP32 blocks load Dp^-1 first, so that
// everything but s and the neighbours' z is in flight after one dependent
// load (FP64 blocks load it last: held from the start they cost more
// occupancy than the early load gains).
// s -= the n couplings of packed slots r0, r0 + dr, ... (dr = +-1, uniform)
template <int NV, class TB, class TZ>
__device__ __forceinline__ void line_couple(int t, int n, int r0, int dr, int4 e4, int p, const TB* __restrict__ Ac,
                                            const TZ* __restrict__ z, double (&s)[NV]) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k >= n) continue;
    const int r = r0 + dr * k;
    const int o = r == 0 ? e4.x : r == 1 ? e4.y : r == 2 ? e4.z : e4.w;
    const TB* __restrict__ A = Ac + (size_t)r * NV * NV * p + t;
    double zo[NV];
#pragma unroll
    for (int j = 0; j < NV; ++j) zo[j] = (double)z[j * p + o];
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
      for (int j = 0; j < NV; ++j) s[i] -= (double)A[(size_t)(i * NV + j) * p] * zo[j];
  }
}
template <int NV, class TB, class TX, class TZ>
__device__ __forceinline__ void line_w(int t, int c, int pass, bool keep, const int4* __restrict__ lnc,
                                       const unsigned char* __restrict__ lmask, int p,
                                       const TB* __restrict__ Dpinv, const TB* __restrict__ Ac,
                                       const TX* __restrict__ x, TZ* __restrict__ sL, const TZ* __restrict__ z,
                                       double (&w)[NV]) {
  constexpr bool early = sizeof(TB) == 4;
  TB dp[NV * NV];
  if (early) {
#pragma unroll
    for (int e = 0; e < NV * NV; ++e) dp[e] = Dpinv[(size_t)e * p + t];
  }
  const unsigned mk = lmask[t];
  double s[NV];
  if (pass == LINE_BWD) {
#pragma unroll
    for (int i = 0; i < NV; ++i) s[i] = (double)sL[i * p + t];
  } else {
#pragma unroll
    for (int i = 0; i < NV; ++i) s[i] = (double)x[i * p + c];
  }
  const int4 e4 = lnc[t];
  const bool bwd = pass == LINE_BWD;
  line_couple<NV, TB, TZ>(t, __popc(bwd ? mk >> 4 : mk & 15u), bwd ? 3 : 0, bwd ? -1 : 1, e4, p, Ac, z, s);
  if (pass == LINE_FWD && keep) {
#pragma unroll
    for (int i = 0; i < NV; ++i) sL[i * p + t] = (TZ)s[i];
  }
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double a = 0.0;
#pragma unroll
    for (int j = 0; j < NV; ++j) a += (double)(early ? dp[i * NV + j] : Dpinv[(size_t)(i * NV + j) * p + t]) * s[j];
    w[i] = a;
  }
}

// The off-segment coupling blocks of every line position into its packed
// slots (lqm: the cell slot each packed slot takes, -1 for none), once per
// factorisation.
template <int NV, class TB>
__global__ __launch_bounds__(BS) void k_line_pack(int N, int p, const char4* __restrict__ lqm,
                                                  const TB* __restrict__ Aoff, TB* __restrict__ Ac) {
  constexpr int NN = NV * NV;
  const int t = xcd_block() * BS + threadIdx.x;
  if (t >= N) return;
  const char4 q4 = lqm[t];
  const int qs[4] = {q4.x, q4.y, q4.z, q4.w};
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int q = qs[r];
    if (q < 0) continue;
#pragma unroll
    for (int e = 0; e < NN; ++e) Ac[(size_t)(r * NN + e) * p + t] = Aoff[(size_t)(q * NN + e) * p + t];
  }
}

// All cells of one color (positions [t0, t1)) for large meshes: z = w for a
// single-cell segment, w into its bundle slot (FP64) for a cell of a longer
// one, whose recurrence k_line_solve then runs.  This bandwidth-bound part runs
// at the occupancy its own registers allow.
template <int NV, class TB, class TX, class TZ>
__global__ __launch_bounds__(LINE_BS) void k_line_cells(int t0, int t1, int pass, const int* __restrict__ seg_cells,
                                                   const int4* __restrict__ lnc, const int* __restrict__ lcomp,
                                                   const unsigned char* __restrict__ lmask, int p,
                                                   const TB* __restrict__ Dpinv, const TB* __restrict__ Ac,
                                                   const TX* __restrict__ x, TZ* __restrict__ sL, TZ* __restrict__ z,
                                                   double* __restrict__ W) {
  const int t = t0 + xcd_block() * LINE_BS + threadIdx.x;
  if (t >= t1) return;
  const int c = seg_cells[t];
  const int m = lcomp[t];
  double w[NV];
  line_w<NV, TB, TX, TZ>(t, c, pass, true, lnc, lmask, p, Dpinv, Ac, x, sL, z, w);
  if (m >= 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) W[(size_t)m * NV + i] = w[i];
  } else {
#pragma unroll
    for (int i = 0; i < NV; ++i) z[i * p + t] = (TZ)w[i];
  }
}

// One wavefront per bundle b0 + blockIdx.x after k_line_cells: forward and
// backward recurrences across the lanes (in the precision of z), then every
// cell's z.  Every slot is allocated, so the loads are unconditional (a guarded
// load becomes a branch that waits for it); values a lane does not own are
// never used.
template <int NV, class TB, class TZ>
__global__ __launch_bounds__(64) void k_line_solve(int b0, const int* __restrict__ lb_t,
                                                   const int* __restrict__ lb_kl, const int* __restrict__ seg_cells,
                                                   int p, const TB* __restrict__ Pm, const TB* __restrict__ G,
                                                   const double* __restrict__ W, TZ* __restrict__ z) {
  constexpr int NN = NV * NV;
  typedef TZ TC;
  const size_t bb = (size_t)(b0 + blockIdx.x) * 64, m = bb + threadIdx.x;
  const int t = lb_t[m];
  const int kl = lb_kl[m];
  const int k = kl & 255, L = kl >> 8;
  const int steps = __builtin_amdgcn_readfirstlane(lb_kl[bb] >> 8);  // slot 0 holds a longest segment
  TC y[NV];
  TB pr[NN], gr[NN];
#pragma unroll
  for (int i = 0; i < NV; ++i) y[i] = (TC)W[m * NV + i];
#pragma unroll
  for (int e = 0; e < NN; ++e) {
    pr[e] = Pm[m * NN + e];
    gr[e] = G[m * NN + e];
  }
  line_recur<NV, TB, TC>(steps, k, L, t >= 0, pr, gr, y);
  if (t >= 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) z[i * p + t] = (TZ)y[i];
  }
}

// One color of a sweep in one launch (64-thread blocks).  Blocks [0, nb) are
// the color's bundles b0 .. b0 + nb - 1, first so that their recurrences start
// early: each lane forms w of its cell, keeps it in registers, and the
// recurrences run across the lanes.  Blocks from nb on take the line positions
// [t0, t1) of the color's single-cell segments (which follow its longer ones)
// 64 at a time and finish them (z = w).  Segments of
// one color share no face, so no block reads a z that another block of the
// launch writes.  The recurrences run in the precision of z (TZ).
template <int NV, class TB, class TX, class TZ>
__global__ __launch_bounds__(64) void k_line_sweep(int b0, int nb, int t0, int t1, int pass,
                                                   const int* __restrict__ lb_t, const int* __restrict__ lb_kl,
                                                   const int* __restrict__ seg_cells, const int4* __restrict__ lnc,
                                                   const unsigned char* __restrict__ lmask, int p,
                                                   const TB* __restrict__ Dpinv, const TB* __restrict__ Ac,
                                                   const TB* __restrict__ Pm, const TB* __restrict__ G,
                                                   const TX* __restrict__ x, TZ* __restrict__ sL,
                                                   TZ* __restrict__ z) {
  constexpr int NN = NV * NV;
  typedef TZ TC;  // recurrence arithmetic
  const int blk = blockIdx.x;
  if (blk >= nb) {  // single-cell segments (uniform per block)
    const int t = t0 + (blk - nb) * 64 + threadIdx.x;
    if (t >= t1) return;
    const int c = seg_cells[t];
    double w[NV];
    line_w<NV, TB, TX, TZ>(t, c, pass, true, lnc, lmask, p, Dpinv, Ac, x, sL, z, w);
#pragma unroll
    for (int i = 0; i < NV; ++i) z[i * p + t] = (TZ)w[i];
    return;
  }
  const size_t bb = (size_t)(b0 + blk) * 64, m = bb + threadIdx.x;
  const int t = lb_t[m];
  const int kl = lb_kl[m];
  const int k = kl & 255, L = kl >> 8;
  const int steps = __builtin_amdgcn_readfirstlane(lb_kl[bb] >> 8);  // slot 0 holds a longest segment
  // every lane stays active for the lane moves: a padding slot works on slot
  // 0's cell and writes nothing; every slot of P and G is allocated, so their
  // loads are unconditional (a guarded load becomes a branch that waits for it)
  const int tc = t >= 0 ? t : lb_t[bb];
  const int c = seg_cells[tc];
  TB pr[NN], gr[NN];
#pragma unroll
  for (int e = 0; e < NN; ++e) {
    pr[e] = Pm[m * NN + e];
    gr[e] = G[m * NN + e];
  }
  double w[NV];
  line_w<NV, TB, TX, TZ>(tc, c, pass, t >= 0, lnc, lmask, p, Dpinv, Ac, x, sL, z, w);
  TC y[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) y[i] = (TC)w[i];
  line_recur<NV, TB, TC>(steps, k, L, t >= 0, pr, gr, y);
  if (t >= 0) {
#pragma unroll
    for (int i = 0; i < NV; ++i) z[i * p + t] = (TZ)y[i];
  }
}

// z in cell order from the sweep's result by line position (reads gathered,
// writes coalesced)
template <int NV, class TZ>
__global__ __launch_bounds__(BS) void k_line_gather(int N, int p, const int* __restrict__ lpos,
                                                    const TZ* __restrict__ zt, TZ* __restrict__ z) {
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= N) return;
  const int t = lpos[c];
#pragma unroll
  for (int i = 0; i < NV; ++i) z[i * p + c] = zt[i * p + t];
}

// ---- reductions (deterministic: fixed grid, fixed tree; RED_BLOCKS in ctx.hpp) -------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// block-wide sum; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < BS / 64; ++i) r += sh[i];
  }
  return r;
}

template <int NV, int KIND, class TA, class TB>
__global__ __launch_bounds__(BS) void k_red(int N, int p, const TA* __restrict__ a, const TB* __restrict__ b,
                                             const double* __restrict__ vol, double* __restrict__ part) {
  __shared__ double sh[BS / 64];
  double s0 = 0.0, s1 = 0.0;
  for (int c = blockIdx.x * BS + threadIdx.x; c < N; c += gridDim.x * BS) {
    if (KIND == RED_L2) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < NV; ++k) {
        const double x = (double)a[k * p + c];
        t += x * x;
      }
      s0 += t;
    } else if (KIND == RED_DOT) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < NV; ++k) t += (double)a[k * p + c] * (double)b[k * p + c];
      s0 += t;
    } else {  // RED_VW over all rows, RED_VW4 over rho, rho u, rho v, rho E only
      const double v = vol[c];
      constexpr int NK = KIND == RED_VW4 ? 4 : NV;
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        const double r = (double)a[k * p + c] / v;
        t += r * r;
      }
      s0 += t * v;
      s1 += v;
    }
  }
  s0 = block_sum(s0, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s0;
  if (KIND == RED_VW || KIND == RED_VW4) {
    s1 = block_sum(s1, sh);
    if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = s1;
  }
}

// sums nsum groups of nb partials; group i -> out[i]
__global__ __launch_bounds__(BS) void k_red_final(const double* __restrict__ part, int nb, int nsum,
                                                   double* __restrict__ out, const double* __restrict__ gate = nullptr) {
  // one block per sum (launched with nsum blocks); each block adds its nb
  // partials in a fixed order, so the result is deterministic
  if (gate && *gate == 0.0) return;
  __shared__ double sh[BS / 64];
  const int i = blockIdx.x;
  if (i >= nsum) return;
  double s = 0.0;
  for (int k = threadIdx.x; k < nb; k += BS) s += part[(size_t)i * nb + k];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) out[i] = s;
}

// ---- FD Jacobian-vector product (Solver.cpp:1331-1386) -------------------------------
__device__ __forceinline__ double fd_eps(const double* scal, int su, int sv) {
  return 1e-7 * (1.0 + sqrt(scal[su])) / (sqrt(scal[sv]) + 1e-30);
}

template <int NV>
__global__ __launch_bounds__(BS) void k_perturb(int N, int p, const double* __restrict__ U,
                                                 const double* __restrict__ v, const double* __restrict__ scal,
                                                 int su, int sv, double* __restrict__ Up) {
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= N) return;
  const double eps = fd_eps(scal, su, sv);
#pragma unroll
  for (int k = 0; k < NV; ++k) Up[k * p + c] = U[k * p + c] + v[k * p + c] * eps;
}

template <int NV>
__global__ __launch_bounds__(BS) void k_fd_combine(int N, int p, const double* __restrict__ Rp,
                                                    const double* __restrict__ Rb, const double* __restrict__ v,
                                                    const double* __restrict__ vol, const double* __restrict__ dt,
                                                    const double* __restrict__ scal, int su, int sv,
                                                    double* __restrict__ Jv) {
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= N) return;
  const double eps = fd_eps(scal, su, sv);
  const double tc = vol[c] / (dt[c] + 1e-30);
#pragma unroll
  for (int k = 0; k < NV; ++k) Jv[k * p + c] = (Rp[k * p + c] - Rb[k * p + c]) * (1.0 / eps) + v[k * p + c] * tc;
}

template <class A, class B>
__global__ void k_convert(const A* __restrict__ a, B* __restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) b[i] = (B)a[i];
}

// ---- pack / unpack between reference-order AoS and RCM-order SoA ---------------------
template <int NV>
__global__ __launch_bounds__(BS) void k_pack(int N, int p, const int* __restrict__ iperm,
                                              const double* __restrict__ U6, double* __restrict__ U) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= N) return;
  const size_t o = (size_t)iperm[i] * 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) U[k * p + i] = U6[o + k];
}
template <int NV>
__global__ __launch_bounds__(BS) void k_unpack(int N, int p, const int* __restrict__ iperm,
                                                const double* __restrict__ U, double* __restrict__ U6) {
  const int i = blockIdx.x * BS + threadIdx.x;
  if (i >= N) return;
  const size_t o = (size_t)iperm[i] * 6;
#pragma unroll
  for (int k = 0; k < 6; ++k) U6[o + k] = k < NV ? U[k * p + i] : 0.0;
}

// ---- vector kernels ----------------------------------------------------------------------
__global__ void k_axpby(double* __restrict__ y, double a, const double* __restrict__ x, double b, size_t n) {
  size_t i = (size_t)blockIdx.x * BS + threadIdx.x;
  if (i < n) y[i] = a * y[i] + b * x[i];
}

// Newton update with the positivity clips of Solver.cpp:1760-1778
template <int NV>
__global__ __launch_bounds__(BS) void k_newton_update(int N, int p, Phys P, double* __restrict__ U,
                                                       const double* __restrict__ dU, double relax) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= N) return;
  double u[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) u[k] = U[k * p + c] + dU[k * p + c] * relax;
  const double ke = 0.5 * (u[1] * u[1] + u[2] * u[2]) / (u[0] + 1e-12);
  double pn = (P.gamma - 1.0) * (u[3] - ke);
  if (u[0] < 1e-5) u[0] = 1e-5;
  if (pn < 1e-5) {
    pn = 1e-5;
    u[3] = pn / (P.gamma - 1.0) + ke;
  }
  if constexpr (NV == 6) {
    if (P.turb) {  // SST: keep k and omega positive
      if (u[4] < u[0] * P.kmin) u[4] = u[0] * P.kmin;
      if (u[5] < u[0] * P.wmin) u[5] = u[0] * P.wmin;
    }
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) U[k * p + c] = u[k];
}

// Entropy error partials (Solver.cpp:1620-1646): per block, sum of
// ((p / rho^gamma - S_ref) / S_ref)^2 V and of V over the cells with p, rho > 0.
template <int NV>
__global__ __launch_bounds__(BS) void k_entropy(int N, int p, Phys P, const double* __restrict__ vol,
                                                 const double* __restrict__ U, double S_ref,
                                                 double* __restrict__ part) {
  __shared__ double sh[BS / 64];
  double s0 = 0.0, s1 = 0.0;
  for (int c = blockIdx.x * BS + threadIdx.x; c < N; c += gridDim.x * BS) {
    const S6<double> u = ld<NV>(U, p, c);
    const double pr = pressure(P, u), rho = u[0];
    if (pr <= 0 || rho <= 0) continue;
    const double S = pr / pow(rho, P.gamma);
    const double e = (S - S_ref) / S_ref;
    s0 += (e * e) * vol[c];
    s1 += vol[c];
  }
  s0 = block_sum(s0, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s0;
  s1 = block_sum(s1, sh);
  if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = s1;
}

// Boundary-layer profile (Solver.cpp:1866-1897): for cells with
// |x_c - x_target| < 0.005, (y_c, u / u_inf, v / u_inf, eta) with
// eta = y_c sqrt(u_inf / (nu max(x_target, 1e-6))); flag marks them.
template <int NV>
__global__ __launch_bounds__(BS) void k_bl_profile(int N, int p, const double* __restrict__ xc,
                                                    const double* __restrict__ yc, const double* __restrict__ U,
                                                    double x_target, double u_inf, double nu,
                                                    double* __restrict__ out4, unsigned char* __restrict__ flag) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= N) return;
  const double x = xc[c], y = yc[c];
  const bool in = fabs(x - x_target) < 0.005;
  flag[c] = in ? 1 : 0;
  if (!in) return;
  const double u = U[p + c] / U[c], v = U[2 * p + c] / U[c];
  const double x_safe = fmax(x_target, 1e-6);
  out4[c] = y;
  out4[p + c] = u / u_inf;
  out4[2 * p + c] = v / u_inf;
  out4[3 * p + c] = y * sqrt(u_inf / (nu * x_safe));
}

// Dual time stepping (Solver.cpp:1831-1843): the unsteady residual
// F = (U - U_old) vol / dt_phys + R, stored negated as the GMRES right-hand side.
template <int NV>
__global__ __launch_bounds__(BS) void k_unsteady_rhs(int N, int p, const double* __restrict__ vol,
                                                      const double* __restrict__ U, const double* __restrict__ Uold,
                                                      const double* __restrict__ R, double dtp,
                                                      double* __restrict__ rhs) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= N) return;
  const double f = vol[c] / dtp;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double F = (U[k * p + c] - Uold[k * p + c]) * f + R[k * p + c];
    rhs[k * p + c] = F * (-1.0);
  }
}

// Explicit stage: Unew = a0 U0 + a1 (U1 - R dt / vol), with the clips of
// Solver.cpp:1233-1241.  a0 == 0 gives Solver::timeStepExplicit.
template <int NV>
__global__ __launch_bounds__(BS) void k_explicit(int N, int p, const double* __restrict__ vol,
                                                  double* __restrict__ Unew, const double* __restrict__ U0, double a0,
                                                  const double* __restrict__ U1, double a1,
                                                  const double* __restrict__ R, double dtc) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= N) return;
  const double f = dtc / vol[c];
  double u[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double e = U1[k * p + c] - R[k * p + c] * f;
    u[k] = a0 == 0.0 ? e : U0[k * p + c] * a0 + e * a1;
  }
  if (u[0] < 1e-3) u[0] = 1e-3;
  const double ke = 0.5 * (u[1] * u[1] + u[2] * u[2]) / u[0];
  if (u[3] - ke < 1e-3) u[3] = ke + 1e-3;
#pragma unroll
  for (int k = 0; k < NV; ++k) Unew[k * p + c] = u[k];
}

// Wall outputs per device face: Cp on tags 3/4, Cf on tag 4
// (Solver.cpp:1919-1967), p A on tag 3.
template <int NV>
__global__ __launch_bounds__(BS) void k_wall(DMesh m, Phys P, const double* __restrict__ U, double qinf,
                                              double* __restrict__ out3) {
  const int f = blockIdx.x * BS + threadIdx.x;
  if (f >= m.F) return;
  double cp = 0.0, cf = 0.0, pw = 0.0;
  const int tg = m.tag[f];
  if (m.nei[f] < 0 && (tg == 3 || tg == 4)) {
    const int o = m.own[f];
    const S6<double> L = ld<NV>(U, m.pitch, o);
    const double pr = pressure(P, L);
    cp = (pr - P.pexit) / qinf;
    pw = pr;
    if (tg == 4) {
      const double uc = L[1] / L[0], vc = L[2] / L[0];
      double tx = -m.ny[f], ty = m.nx[f];
      if (tx < 0) {
        tx = -tx;
        ty = -ty;
      }
      const double ut = uc * tx + vc * ty;
      const double dx = m.xm[f] - m.xc[o], dy = m.ym[f] - m.yc[o];
      double dn = fabs(dx * m.nx[f] + dy * m.ny[f]);
      if (dn < 1e-12) dn = 1e-12;
      const double mu = L[0] * P.nu;
      cf = mu * (ut / dn) / qinf;
    }
  }
  out3[f] = cp;
  out3[m.F + f] = cf;
  out3[2 * m.F + f] = pw;
}

// ---- Krylov vector kernels (basis element type TV: double or float) -------------------
template <int NV, class TV>
__global__ __launch_bounds__(BS) void k_kinit(int N, int p, TV* __restrict__ V0, const double* __restrict__ r,
                                               double sign, const double* __restrict__ scal, int slot) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= N) return;
  const double ib = 1.0 / sqrt(scal[slot]);
#pragma unroll
  for (int k = 0; k < NV; ++k) V0[k * p + c] = (TV)(sign * (r[k * p + c] * ib));
}

template <int NV, class TV, class TZ>
__global__ __launch_bounds__(BS) void k_precond(int N, int p, TZ* __restrict__ z, const TV* __restrict__ Vj,
                                                 const double* __restrict__ D) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= N) return;
  const double id = 1.0 / D[c];
#pragma unroll
  for (int k = 0; k < NV; ++k) z[k * p + c] = (TZ)((double)Vj[k * p + c] * id);
}

// row scaling of a vector: w_k *= s_k (six factors passed by value)
struct Scale6 {
  double s[6];
};
template <int NV, class TV>
__global__ __launch_bounds__(BS) void k_rowscale(int N, int p, TV* __restrict__ w, Scale6 s) {
  const int c = blockIdx.x * BS + threadIdx.x;
  if (c >= N) return;
#pragma unroll
  for (int k = 0; k < NV; ++k) w[k * p + c] = (TV)((double)w[k * p + c] * s.s[k]);
}

// ---- Arnoldi with delayed reorthogonalisation (DCGS2) ---------------------------------
// Step j multiplies the candidate u_j = V[j], orthogonalised once against the
// final basis Q = V[0 .. j-1] but not yet a second time, and then finishes it
// together with the next candidate, reading Q twice per step (one reduction,
// one update) where two-pass Gram-Schmidt read it four times:
//   reduction   s = Q^T u, alpha = u.u, t = Q^T w, beta = u.w  (w = V[j+1] = A M^-1 u)
//   r = sqrt(alpha - s.s);  column j-1 of H: h_prev + s, H[j][j-1] = r
//   q_j = (u - Q s) / r,    delta = q_j . w = (beta - s.t) / r
//   u_{j+1} = (w - Q t - delta q_j) / r   (w projected off [Q, q_j] once, scaled)
//   next partial column: h_prev[i < j] = (t_i - (H s)_i) / r,
//                        h_prev[j] = (delta - r s_{j-1}) / r
// because A M^-1 Q s = Q (H s) + q_j r s_{j-1} from the finished columns.  A
// last reduction of u_m (no product) finishes column m-1.

// Projections of u and of w on V_0 .. V_{nq-1} and on u itself, 8 vectors per
// grid row: R[i] = V_i . u (R[nq] = u . u) and, with w, R[nq + 1 + i] =
// V_i . w (R[2 nq + 1] = u . w); per-block partials into part.
template <int NV, class TV, int DG>
__global__ __launch_bounds__(BS) void k_dots2(int N, int p, size_t vstride, const TV* __restrict__ V, int nq,
                                               const TV* __restrict__ u, const TV* __restrict__ w,
                                               double* __restrict__ part) {
  __shared__ double sh[BS / 64];
  const int i0 = blockIdx.y * DG;
  const int cnt = min(DG, nq + 1 - i0);
  const bool two = w != nullptr;
  double au[DG], aw[DG];
#pragma unroll
  for (int i = 0; i < DG; ++i) au[i] = aw[i] = 0.0;
  for (int c = blockIdx.x * BS + threadIdx.x; c < N; c += gridDim.x * BS) {
    double uv[NV], wv[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      uv[k] = (double)u[k * p + c];
      wv[k] = two ? (double)w[k * p + c] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < DG; ++i) {
      if (i < cnt) {
        const TV* Vi = i0 + i < nq ? V + (size_t)(i0 + i) * vstride : u;
        double tu = 0.0, tw = 0.0;
#pragma unroll
        for (int k = 0; k < NV; ++k) {
          const double vi = (double)Vi[k * p + c];
          tu += uv[k] * vi;
          tw += wv[k] * vi;
        }
        au[i] += tu;
        aw[i] += tw;
      }
    }
  }
  for (int i = 0; i < cnt; ++i) {
    double sv = block_sum(au[i], sh);
    if (threadIdx.x == 0) part[(size_t)(i0 + i) * gridDim.x + blockIdx.x] = sv;
    if (two) {
      sv = block_sum(aw[i], sh);
      if (threadIdx.x == 0) part[(size_t)(nq + 1 + i0 + i) * gridDim.x + blockIdx.x] = sv;
    }
  }
}

// The step's scalars from the reduction R (one wave; each lane takes the
// basis indices i = lane, lane + 64, ...: up to 64 vectors a lane's partial
// sums hold one term each, as when the kernel took one index per lane):
// finishes column j-1 of H (device copy Hd, column-major with leading
// dimension ldh, and hout, pinned host memory), forms the next partial column
// hp and the update's coefficients coef: s at [0, j), t at [ldh, ldh + j),
// delta at 2 ldh, 1/r at 2 ldh + 1.  Step 0 has Q empty and r = 1.  A norm r
// at or below 1e-12 (breakdown) keeps the vectors unscaled, as the reference
// normalises only above it (Solver.cpp:1550-1556).
__global__ __launch_bounds__(64) void k_dcgs_coef(const double* __restrict__ R, int j, int last,
                                                   double* __restrict__ Hd, int ldh, double* __restrict__ hp,
                                                   double* __restrict__ coef, double* __restrict__ hout) {
  const int l = threadIdx.x;
  double ssq = 0.0;
  for (int i = l; i < j; i += 64) ssq += R[i] * R[i];
  const double ss = __shfl(wave_sum(ssq), 0, 64);
  const double r = j > 0 ? sqrt(fmax(R[j] - ss, 0.0)) : 1.0;
  if (j > 0) {
    for (int i = l; i < j; i += 64) {
      const double hv = hp[i] + R[i];
      Hd[(size_t)(j - 1) * ldh + i] = hv;
      if (hout) hout[i] = hv;
    }
    if (l == (j & 63)) {
      Hd[(size_t)(j - 1) * ldh + j] = r;
      if (hout) hout[j] = r;
    }
  }
  if (last) return;
  __syncthreads();
  const double ir = r > 1e-12 ? 1.0 / r : 1.0;
  double sst = 0.0;
  for (int i = l; i < j; i += 64) sst += R[i] * R[j + 1 + i];
  const double st = __shfl(wave_sum(sst), 0, 64);
  const double delta = (R[2 * j + 1] - st) * ir;
  const double gamma = j > 0 ? r * R[j - 1] : 0.0;
  for (int i = l; i < j; i += 64) {
    double gi = 0.0;
    for (int k = i > 0 ? i - 1 : 0; k < j; ++k) gi += Hd[(size_t)k * ldh + i] * R[k];
    const double ti = R[j + 1 + i];
    hp[i] = (ti - gi) * ir;
    coef[i] = R[i];
    coef[ldh + i] = ti;
  }
  if (l == (j & 63)) hp[j] = (delta - gamma) * ir;
  if (l == 0) {
    coef[2 * ldh] = delta;
    coef[2 * ldh + 1] = ir;
  }
}

// q_j = (u - Q s) / r into V[j] and u_{j+1} = (w - Q t - delta q_j) / r into
// V[j+1], reading Q once; u_{j+1} uses q_j as stored.
// LDC > 0: the coefficients' stride as a constant (t at coef + LDC, delta and
// 1/r at coef + 2 LDC); 0: the runtime stride ldc.  The two forms compile to
// different loops: on the 1.32M-cell SST plate the FP32 update took 139 us
// with the constant stride of 64 and 191 us with a runtime one, the FP64
// update 293.5 and 277.8 us, so the launcher takes the constant form for FP32
// bases (leading dimension 64, 128 or 256) and the runtime form for FP64.
template <int NV, class TV, int LDC>
__global__ __launch_bounds__(BS) void k_dcgs_update(int N, int p, size_t vstride, TV* __restrict__ V, int j,
                                                     const double* __restrict__ coef, int ldc) {
  const int ld = LDC > 0 ? LDC : ldc;
  const double* __restrict__ cs = coef;
  const double* __restrict__ ct = coef + ld;
  const double* __restrict__ cd = coef + 2 * ld;
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= N) return;
  TV* u = V + (size_t)j * vstride;
  TV* w = u + vstride;
  double as[NV], aw[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    as[k] = (double)u[k * p + c];
    aw[k] = (double)w[k * p + c];
  }
#pragma unroll 4
  for (int i = 0; i < j; ++i) {
    const double si = cs[i], ti = ct[i];
    const TV* Vi = V + (size_t)i * vstride + c;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const double vi = (double)Vi[k * p];
      as[k] -= si * vi;
      aw[k] -= ti * vi;
    }
  }
  const double delta = cd[0], ir = cd[1];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const TV q = (TV)(as[k] * ir);
    u[k * p + c] = q;
    w[k * p + c] = (TV)((aw[k] - delta * (double)q) * ir);
  }
}

// x = sum_{i<=k} y_i V_i / D, accumulated in i order (Solver.cpp:1610-1615)
template <int NV, class TV>
__global__ __launch_bounds__(BS) void k_solution(int N, int p, size_t vstride, const TV* __restrict__ V,
                                                  const double* __restrict__ y, int nvec,
                                                  const double* __restrict__ D, double* __restrict__ x) {
  const int c = xcd_block() * BS + threadIdx.x;
  if (c >= N) return;
  const double id = D ? 1.0 / D[c] : 1.0;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int i = 0; i < nvec; ++i) s = s + (double)V[(size_t)i * vstride + k * p + c] * y[i] * id;
    x[k * p + c] = s;
  }
}

// ---- launchers --------------------------------------------------------------------------------
#define NV_DISPATCH(nv, ...) \
  do {                       \
    if ((nv) == 6) {         \
      constexpr int NVt = 6; \
      __VA_ARGS__;           \
    } else {                 \
      constexpr int NVt = 4; \
      __VA_ARGS__;           \
    }                        \
  } while (0)

void launch_fill(cfd_ctx* c, double* x, double v, size_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill<double>, dim3(nblocks(n)), dim3(BS), 0, c->stream, x, v, n);
}

template <int NV, int VISC>
static void residual_nv(cfd_ctx* c, const double* U, double* R, bool upd, hipEvent_t* ev) {
  const DMesh& m = c->dm;
  const Phys& P = c->phys;
  const bool lim = P.spatial == CFD_SECOND_ORDER_LIMITED;
  const int do_lim = lim && (upd || !c->phi_valid);
  const int kw_bnd = NV == 6 && P.turb && P.turb1 ? 2 : 0;  // k, omega rows only where a boundary face reads them
  if (ev) (void)hipEventRecord(ev[0], c->stream);
  hipLaunchKernelGGL((k_gradlim<NV, double>), dim3(nblocks(m.N, GRAD_BS)), dim3(GRAD_BS), 0, c->stream, m, P, U,
                     (const double*)nullptr, c->grad, (double*)nullptr, c->phi, do_lim | kw_bnd, c->tq,
                     (double*)nullptr);
  if (do_lim) c->phi_valid = true;
  if (ev) (void)hipEventRecord(ev[1], c->stream);
  if (c->assembly == 0) {
    launch_fill(c, R, 0.0, (size_t)NV * m.pitch);
    if (ev) (void)hipEventRecord(ev[2], c->stream);
    for (int col = 0; col < c->ncolors; ++col) {
      const int f0 = c->color_ptr[col], f1 = c->color_ptr[col + 1];
      if (f1 > f0)
        hipLaunchKernelGGL((k_flux<NV, VISC, 0, double>), dim3(nblocks(f1 - f0, FLUX_BS)), dim3(FLUX_BS), 0, c->stream, m, P, U,
                           (const double*)nullptr, c->grad, (const double*)nullptr, c->phi, R, f0, f1,
                           (const double*)c->tq, (const double*)nullptr);
    }
    if (ev) (void)hipEventRecord(ev[3], c->stream);
  } else {
    if (ev) (void)hipEventRecord(ev[2], c->stream);
    hipLaunchKernelGGL((k_flux<NV, VISC, 1, double>), dim3(nblocks(m.F, FLUX_BS)), dim3(FLUX_BS), 0, c->stream, m, P, U,
                       (const double*)nullptr, c->grad, (const double*)nullptr, c->phi, c->Fm, 0, m.F,
                       (const double*)c->tq, (const double*)nullptr);
    if (ev) (void)hipEventRecord(ev[3], c->stream);
    hipLaunchKernelGGL((k_gather<NV, double>), dim3(nblocks(m.N)), dim3(BS), 0, c->stream, m, c->Fm, R,
                       NV == 6 && P.turb ? (const double*)c->tq : nullptr, (const double*)nullptr,
                       (const double*)nullptr);
  }
  if (c->assembly == 0 && NV == 6 && P.turb)
    hipLaunchKernelGGL((k_turb_src<double>), dim3(nblocks(m.N)), dim3(BS), 0, c->stream, m.N, m.pitch,
                       (const double*)c->tq, R);
  if (ev) (void)hipEventRecord(ev[4], c->stream);
}

// Forward tangent of the residual with the limiter frozen, element type F.
template <int NV, int VISC, class F>
static void tangent_nv(cfd_ctx* c, const double* U, const F* v, F* Jv) {
  typedef Dual<F> T;
  const DMesh& m = c->dm;
  const Phys& P = c->phys;
  F* Gv = (F*)c->gradT;
  F* Gd = Gv + (size_t)14 * m.pitch;
  F* Qv = (F*)c->tq;
  F* Qd = Qv + (size_t)4 * m.pitch;
  hipLaunchKernelGGL((k_gradlim<NV, T>), dim3(nblocks(m.N, GRAD_BS)), dim3(GRAD_BS), 0, c->stream, m, P, U, v, Gv, Gd, c->phi, 0,
                     Qv, Qd);
  if (c->assembly == 0) {
    hipLaunchKernelGGL(k_fill<F>, dim3(nblocks((size_t)NV * m.pitch)), dim3(BS), 0, c->stream, Jv, F(0),
                       (size_t)NV * m.pitch);
    for (int col = 0; col < c->ncolors; ++col) {
      const int f0 = c->color_ptr[col], f1 = c->color_ptr[col + 1];
      if (f1 > f0)
        hipLaunchKernelGGL((k_flux<NV, VISC, 0, T>), dim3(nblocks(f1 - f0, FLUX_BS)), dim3(FLUX_BS), 0, c->stream, m, P, U, v,
                           (const F*)Gv, (const F*)Gd, c->phi, Jv, f0, f1, (const F*)Qv, (const F*)Qd);
    }
  } else {
    F* Fm = (F*)c->Fm;
    hipLaunchKernelGGL((k_flux<NV, VISC, 1, T>), dim3(nblocks(m.F, FLUX_BS)), dim3(FLUX_BS), 0, c->stream, m, P, U, v,
                       (const F*)Gv, (const F*)Gd, c->phi, Fm, 0, m.F, (const F*)Qv, (const F*)Qd);
    hipLaunchKernelGGL((k_gather<NV, F>), dim3(nblocks(m.N)), dim3(BS), 0, c->stream, m, (const F*)Fm, Jv,
                       NV == 6 && P.turb ? (const F*)Qd : nullptr, v, (const double*)c->dt);
    return;
  }
  if (NV == 6 && P.turb)
    hipLaunchKernelGGL((k_turb_src<F>), dim3(nblocks(m.N)), dim3(BS), 0, c->stream, m.N, m.pitch, (const F*)Qd, Jv);
  hipLaunchKernelGGL((k_tcombine<NV, F>), dim3(nblocks(m.N)), dim3(BS), 0, c->stream, m.N, m.pitch, Jv, v, c->d_vol,
                     c->dt);
}

template <int NV, int VISC>
static void vjp_nv(cfd_ctx* c, const double* U, const double* w, double* JTw, bool grads_ready) {
  const DMesh& m = c->dm;
  const Phys& P = c->phys;
  double* aG = c->adj;                         // 14 rows: adjoints of the gradients
  double* aQ = c->adj + (size_t)14 * m.pitch;  // 2 rows: adjoints of mu_t and F1 (SST)
  if (!grads_ready)  // the gradients (and SST cell quantities) of U, limiter frozen
    hipLaunchKernelGGL((k_gradlim<NV, double>), dim3(nblocks(m.N, GRAD_BS)), dim3(GRAD_BS), 0, c->stream, m, P, U,
                       (const double*)nullptr, c->grad, (double*)nullptr, c->phi, 0, c->tq, (double*)nullptr);
  // a second-order residual with four variables (Euler or laminar), or with
  // SST and first-order k and omega convection, takes its interior faces
  // through the structured product (the adjoint face buffer, then per-cell
  // sums, which initialise the adjoint rows); the boundary faces (each
  // color's trailing run), and every face of the other residuals, through
  // the general one
  constexpr bool has_inv = NV == 4 || VISC != 0;
  const bool inv = has_inv && P.spatial != CFD_FIRST_ORDER && (NV == 4 || (P.turb && P.turb1));
  if constexpr (has_inv) {
    if (inv) {
      hipLaunchKernelGGL((k_vjp_face_inv<NV, VISC>), dim3(nblocks(m.F, FLUX_BS)), dim3(FLUX_BS), 0, c->stream,
                         m, P, U, (const double*)c->grad, (const double*)c->phi, (const double*)c->tq, w, c->adjF);
      hipLaunchKernelGGL((k_vjp_gather<NV, VISC>), dim3(nblocks(m.pitch)), dim3(BS), 0, c->stream, m, P,
                         (const double*)c->adjF, U, (const double*)c->grad, (const double*)c->phi, w, JTw, aG);
    }
  }
  if (!inv) {
    launch_fill(c, JTw, 0.0, (size_t)NV * m.pitch);
    launch_fill(c, aG, 0.0, (size_t)16 * m.pitch);
  }
  // (with Euler the structured pass also covers the boundary faces)
  if (inv) {
    if (VISC != 0 && c->nbgrp > 0)
      hipLaunchKernelGGL((k_vjp_face_bnd<NV, VISC>), dim3(nblocks((size_t)c->nbgrp * vjp_passes<NV, VK_GEN>(), FLUX_BS)),
                         dim3(FLUX_BS), 0, c->stream, m, P, U, (const double*)c->grad, (const double*)c->phi, w,
                         JTw, aG, (const int4*)c->d_bgrp, c->nbgrp);
  } else {
    for (int col = 0; col < c->ncolors; ++col) {
      const int f0 = c->color_ptr[col], f1 = c->color_ptr[col + 1];
      if (f1 > f0)
        hipLaunchKernelGGL((k_vjp_face<NV, VISC>), dim3(nblocks((size_t)(f1 - f0) * vjp_passes<NV, VK_GEN>(), FLUX_BS)),
                           dim3(FLUX_BS), 0, c->stream, m, P, U,
                           (const double*)c->grad, (const double*)c->phi, w, JTw, aG, f0, f1);
    }
  }
  if constexpr (NV == 6) {  // (the structured path forms the SST cell terms in k_vjp_gather)
    if (P.turb && !inv)
      hipLaunchKernelGGL((k_vjp_src<NV>), dim3(nblocks(m.N)), dim3(BS), 0, c->stream, m, P, U,
                         (const double*)c->grad, w, (const double*)aQ, JTw, aG);
  }
  hipLaunchKernelGGL((k_vjp_gg<NV>), dim3(nblocks(m.N)), dim3(BS), 0, c->stream, m, P, U, (const double*)aG, JTw, w,
                     (const double*)c->dt);
  if (c->nbgrp > 0)  // the ghost states' part
    hipLaunchKernelGGL((k_vjp_gg_ghost<NV>), dim3(nblocks((size_t)c->nbgrp * NV)), dim3(BS), 0, c->stream, m, P, U,
                       (const double*)aG, JTw, (const int4*)c->d_bgrp, c->nbgrp);
}

void launch_vjp(cfd_ctx* c, const double* U, const double* w, double* JTw, bool grads_ready) {
  const bool visc = c->phys.viscous != 0;
  if (c->P.nvar == 6) {
    if (visc) vjp_nv<6, 1>(c, U, w, JTw, grads_ready);
    else vjp_nv<6, 0>(c, U, w, JTw, grads_ready);
  } else {
    if (visc) vjp_nv<4, 1>(c, U, w, JTw, grads_ready);
    else vjp_nv<4, 0>(c, U, w, JTw, grads_ready);
  }
}

void launch_vjp_gradients(cfd_ctx* c, const double* U) {
  const DMesh& m = c->dm;
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL((k_gradlim<NVt, double>), dim3(nblocks(m.N, GRAD_BS)), dim3(GRAD_BS), 0,
                                            c->stream, m, c->phys, U, (const double*)nullptr, c->grad,
                                            (double*)nullptr, c->phi, 0, c->tq, (double*)nullptr));
}

template <class F> static void tangent_t(cfd_ctx* c, const double* U, const F* v, F* Jv) {
  const bool visc = c->phys.viscous != 0;
  if (c->P.nvar == 6) {
    if (visc) tangent_nv<6, 1, F>(c, U, v, Jv);
    else tangent_nv<6, 0, F>(c, U, v, Jv);
  } else {
    if (visc) tangent_nv<4, 1, F>(c, U, v, Jv);
    else tangent_nv<4, 0, F>(c, U, v, Jv);
  }
}

void launch_jv_tangent(cfd_ctx* c, const double* U, const void* v, void* Jv, int prec) {
  if (prec == 32) tangent_t<float>(c, U, (const float*)v, (float*)Jv);
  else tangent_t<double>(c, U, (const double*)v, (double*)Jv);
}

void launch_residual_timed(cfd_ctx* c, const double* U, double* R, bool upd, hipEvent_t* ev) {
  const bool visc = c->phys.viscous != 0;
  if (c->P.nvar == 6) {
    if (visc) residual_nv<6, 1>(c, U, R, upd, ev);
    else residual_nv<6, 0>(c, U, R, upd, ev);
  } else {
    if (visc) residual_nv<4, 1>(c, U, R, upd, ev);
    else residual_nv<4, 0>(c, U, R, upd, ev);
  }
}

void launch_residual(cfd_ctx* c, const double* U, double* R, bool upd) { launch_residual_timed(c, U, R, upd, nullptr); }

void launch_pack(cfd_ctx* c, const double* U6, double* U) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_pack<NVt>, dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->N,
                                            c->pitch, c->d_iperm, U6, U));
}
void launch_unpack(cfd_ctx* c, const double* U, double* U6) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_unpack<NVt>, dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->N,
                                            c->pitch, c->d_iperm, U, U6));
}
void launch_timestep(cfd_ctx* c, const double* U, double cfl) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_timestep<NVt>, dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->dm,
                                            c->phys, U, cfl, c->dt));
  c->dt_valid = true;
}
void launch_diag_precond(cfd_ctx* c, const double* U, double* D) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_diagprec<NVt>, dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->dm,
                                            c->phys, U, c->dt, D));
}

static int red_grid(int N) { return std::min(RED_BLOCKS, std::max(1, nblocks(N))); }

template <class TA, class TB>
static void reduce_t(cfd_ctx* c, int kind, const TA* a, const TB* b, int slot) {
  const int nb = red_grid(c->N);
  NV_DISPATCH(c->P.nvar, {
    if (kind == RED_L2)
      hipLaunchKernelGGL((k_red<NVt, RED_L2, TA, TB>), dim3(nb), dim3(BS), 0, c->stream, c->N, c->pitch, a, b,
                         c->d_vol, c->red);
    else if (kind == RED_DOT)
      hipLaunchKernelGGL((k_red<NVt, RED_DOT, TA, TB>), dim3(nb), dim3(BS), 0, c->stream, c->N, c->pitch, a, b,
                         c->d_vol, c->red);
    else if (kind == RED_VW)
      hipLaunchKernelGGL((k_red<NVt, RED_VW, TA, TB>), dim3(nb), dim3(BS), 0, c->stream, c->N, c->pitch, a, b,
                         c->d_vol, c->red);
    else
      hipLaunchKernelGGL((k_red<NVt, RED_VW4, TA, TB>), dim3(nb), dim3(BS), 0, c->stream, c->N, c->pitch, a, b,
                         c->d_vol, c->red);
  });
  const int nsum = kind == RED_VW || kind == RED_VW4 ? 2 : 1;
  hipLaunchKernelGGL(k_red_final, dim3(nsum), dim3(BS), 0, c->stream, c->red, nb, nsum, c->scal + slot);
}

__global__ void k_vw_finish(const double* __restrict__ sums, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[0] = sums[1] == 0.0 ? 0.0 : sqrt(sums[0] / sums[1]);
}

void launch_norm_vw_into(cfd_ctx* c, const double* a, double* part, double* sums, double* out) {
  const int nb = red_grid(c->N);
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL((k_red<NVt, RED_VW, double, double>), dim3(nb), dim3(BS), 0,
                                            c->stream, c->N, c->pitch, a, a, c->d_vol, part));
  hipLaunchKernelGGL(k_red_final, dim3(2), dim3(BS), 0, c->stream, part, nb, 2, sums);
  hipLaunchKernelGGL(k_vw_finish, dim3(1), dim3(64), 0, c->stream, sums, out);
}

void launch_reduce(cfd_ctx* c, int kind, const double* a, const double* b, int slot) {
  reduce_t<double, double>(c, kind, a, b ? b : a, slot);
}

void launch_jv_fd(cfd_ctx* c, const double* U, const double* Rb, const double* v, double* Jv, int su, int sv,
                  double* tmpU, double* tmpR) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_perturb<NVt>, dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->N,
                                            c->pitch, U, v, c->scal, su, sv, tmpU));
  launch_residual(c, tmpU, tmpR, false);
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_fd_combine<NVt>, dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->N,
                                            c->pitch, tmpR, Rb, v, c->d_vol, c->dt, c->scal, su, sv, Jv));
}

void launch_d2f(cfd_ctx* c, const double* a, float* b, size_t n) {
  if (n) hipLaunchKernelGGL((k_convert<double, float>), dim3(nblocks(n)), dim3(BS), 0, c->stream, a, b, n);
}
void launch_f2d(cfd_ctx* c, const float* a, double* b, size_t n) {
  if (n) hipLaunchKernelGGL((k_convert<float, double>), dim3(nblocks(n)), dim3(BS), 0, c->stream, a, b, n);
}

void launch_rowscale(cfd_ctx* c, void* w, const double* s, int prec) {
  Scale6 sc;
  for (int i = 0; i < 6; ++i) sc.s[i] = s[i];
  if (prec == 32)
    NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL((k_rowscale<NVt, float>), dim3(nblocks(c->N)), dim3(BS), 0, c->stream,
                                              c->N, c->pitch, (float*)w, sc));
  else
    NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL((k_rowscale<NVt, double>), dim3(nblocks(c->N)), dim3(BS), 0,
                                              c->stream, c->N, c->pitch, (double*)w, sc));
}

void launch_axpby(cfd_ctx* c, double* y, double a, const double* x, double b, size_t n) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_axpby, dim3(nblocks(n)), dim3(BS), 0, c->stream, y, a, x, b, n);
}

void launch_newton_update(cfd_ctx* c, double* U, const double* dU, double relax) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_newton_update<NVt>, dim3(nblocks(c->N)), dim3(BS), 0, c->stream,
                                            c->N, c->pitch, c->phys, U, dU, relax));
}

void launch_explicit(cfd_ctx* c, double* Unew, const double* U0, double a0, const double* U1, double a1,
                     const double* R, double dtc) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_explicit<NVt>, dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->N,
                                            c->pitch, c->d_vol, Unew, U0, a0, U1, a1, R, dtc));
}

void launch_unsteady_rhs(cfd_ctx* c, const double* U, const double* Uold, const double* R, double dtp,
                         double* rhs) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_unsteady_rhs<NVt>, dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->N,
                                            c->pitch, c->d_vol, U, Uold, R, dtp, rhs));
}

void launch_entropy(cfd_ctx* c, const double* U, double S_ref, int slot) {
  const int nb = red_grid(c->N);
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_entropy<NVt>, dim3(nb), dim3(BS), 0, c->stream, c->N, c->pitch,
                                            c->phys, c->d_vol, U, S_ref, c->red));
  hipLaunchKernelGGL(k_red_final, dim3(2), dim3(BS), 0, c->stream, c->red, nb, 2, c->scal + slot,
                     (const double*)nullptr);
}

void launch_bl_profile(cfd_ctx* c, const double* U, double x_target, double u_inf, double nu, double* out4,
                       unsigned char* flag) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_bl_profile<NVt>, dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->N,
                                            c->pitch, c->d_xc, c->d_yc, U, x_target, u_inf, nu, out4, flag));
}

void launch_wall_distance(cfd_ctx* c, const double* wxy, int nw, double* wd) {
  hipLaunchKernelGGL(k_walldist, dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->N, c->d_xc, c->d_yc, wxy, nw, wd);
}

void launch_wall(cfd_ctx* c, const double* U, double* out3) {
  const Phys& P = c->phys;
  const double g = P.gamma;
  const double fT = 1.0 + 0.5 * (g - 1.0) * c->P.M_inf * c->P.M_inf;
  const double Ts = c->P.T_total_inlet / fT;
  const double cs = std::sqrt(g * c->P.R_gas * Ts);
  const double uinf = c->P.M_inf * cs;
  const double rinf = c->P.P_static_exit / (c->P.R_gas * Ts);
  const double qinf = 0.5 * rinf * uinf * uinf;
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL(k_wall<NVt>, dim3(nblocks(c->F)), dim3(BS), 0, c->stream, c->dm,
                                            c->phys, U, qinf, out3));
}

// ---- Krylov launchers ---------------------------------------------------------------------
template <class TV> static void kinit_t(cfd_ctx* c, TV* V0, const double* rhs, double sign, int slot) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL((k_kinit<NVt, TV>), dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->N,
                                            c->pitch, V0, rhs, sign, c->scal, slot));
}
void launch_krylov_init(cfd_ctx* c, void* V0, const double* rhs, double sign, int slot, int prec) {
  if (prec == 32) kinit_t<float>(c, (float*)V0, rhs, sign, slot);
  else kinit_t<double>(c, (double*)V0, rhs, sign, slot);
}

template <class TV, class TZ> static void precond_t(cfd_ctx* c, TZ* z, const TV* Vj, int slot_norm) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL((k_precond<NVt, TV, TZ>), dim3(nblocks(c->N)), dim3(BS), 0, c->stream,
                                            c->N, c->pitch, z, Vj, c->kry.D));
  if (slot_norm >= 0) reduce_t<TZ, TZ>(c, RED_L2, z, z, slot_norm);
}
void launch_precond_apply(cfd_ctx* c, void* z, const void* Vj, int prec, int slot_norm) {
  if (prec == 32) precond_t<float, float>(c, (float*)z, (const float*)Vj, slot_norm);
  else precond_t<double, double>(c, (double*)z, (const double*)Vj, slot_norm);
}

// One DCGS2 step j (or, with last, the reduction that finishes column j-1
// only); column j-1 of H goes to hout.  Device scalars: R at kry.h[0, 2 j + 2),
// H, the partial column and the coefficients in kry.dc (leading dimension
// kry.ld).
template <class TV> static void dcgs_t(cfd_ctx* c, TV* V, int j, bool last, double* hout) {
  const int N = c->N, p = c->pitch, nv = c->P.nvar;
  const size_t vs = (size_t)nv * p;
  Krylov& k = c->kry;
  const int nb = std::min(RED_BLOCKS, std::max(1, nblocks(N)));
  const int nsum = last ? j + 1 : 2 * (j + 1);
  TV* u = V + (size_t)j * vs;
  const int ld = k.ld;
  double* Hd = k.dc;
  double* hp = k.dc + (size_t)ld * ld;
  double* coef = hp + ld;
  NV_DISPATCH(nv, {
    hipLaunchKernelGGL((k_dots2<NVt, TV, DOTS_GROUP>), dim3(nb, (j + DOTS_GROUP) / DOTS_GROUP), dim3(BS), 0,
                       c->stream, N, p, vs, (const TV*)V, j, (const TV*)u,
                       last ? (const TV*)nullptr : (const TV*)(u + vs), k.part);
    hipLaunchKernelGGL(k_red_final, dim3(nsum), dim3(BS), 0, c->stream, (const double*)k.part, nb, nsum, k.h,
                       (const double*)nullptr);
    hipLaunchKernelGGL(k_dcgs_coef, dim3(1), dim3(64), 0, c->stream, (const double*)k.h, j, last ? 1 : 0, Hd, ld,
                       hp, coef, hout);
    if (!last) {
      if constexpr (std::is_same<TV, float>::value) {
        if (ld == 64)
          hipLaunchKernelGGL((k_dcgs_update<NVt, TV, 64>), dim3(nblocks(N)), dim3(BS), 0, c->stream, N, p, vs, V,
                             j, (const double*)coef, ld);
        else if (ld == 128)
          hipLaunchKernelGGL((k_dcgs_update<NVt, TV, 128>), dim3(nblocks(N)), dim3(BS), 0, c->stream, N, p, vs, V,
                             j, (const double*)coef, ld);
        else
          hipLaunchKernelGGL((k_dcgs_update<NVt, TV, 256>), dim3(nblocks(N)), dim3(BS), 0, c->stream, N, p, vs, V,
                             j, (const double*)coef, ld);
      } else {
        hipLaunchKernelGGL((k_dcgs_update<NVt, TV, 0>), dim3(nblocks(N)), dim3(BS), 0, c->stream, N, p, vs, V, j,
                           (const double*)coef, ld);
      }
    }
  });
}
void launch_dcgs(cfd_ctx* c, void* V, int j, bool last, double* hout, int prec) {
  if (prec == 32) dcgs_t<float>(c, (float*)V, j, last, hout);
  else dcgs_t<double>(c, (double*)V, j, last, hout);
}

template <class TV>
static void block_t(cfd_ctx* c, const double* U, TV* Binv, double* Bout, TV* Aoff, const int* pos) {
  NV_DISPATCH(c->P.nvar, {
    const bool turb = NVt == 6 && c->P.turb;
    hipLaunchKernelGGL((k_gradlim<NVt, double>), dim3(nblocks(c->N, GRAD_BS)), dim3(GRAD_BS), 0, c->stream, c->dm, c->phys, U,
                       (const double*)nullptr, c->grad, (double*)nullptr, c->phi, 0,
                       turb ? c->tq : (double*)nullptr, (double*)nullptr);
    hipLaunchKernelGGL((k_face_jac<NVt, TV>), dim3(nblocks(c->F)), dim3(BS), 0, c->stream, c->dm, c->phys, U,
                       (TV*)c->kry.fjac);
    hipLaunchKernelGGL((k_block<NVt, TV>), dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->dm, c->phys, U,
                       (const double*)c->grad, c->dt, (const TV*)c->kry.fjac, Binv, Bout, Aoff,
                       turb ? (const double*)c->tq : nullptr, pos);
  });
}
void launch_block_precond(cfd_ctx* c, const double* U, void* Binv, int prec, double* Bout, void* Aoff,
                          const int* pos) {
  if (prec == 32) block_t<float>(c, U, (float*)Binv, Bout, (float*)Aoff, pos);
  else block_t<double>(c, U, (double*)Binv, Bout, (double*)Aoff, pos);
}
template <class TV>
static void line_factor_t(cfd_ctx* c, const double* Bd, const TV* Aoff, TV* Dpinv, TV* Pm, TV* G, TV* Ac) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL((k_line_pack<NVt, TV>), dim3(nblocks(c->N)), dim3(BS), 0, c->stream,
                                            c->N, c->pitch, c->d_lqm, Aoff, Ac));
  if (c->nlb == 0) return;  // single-cell segments only: Dp^-1 = B^-1 from the block kernel
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL((k_line_factor<NVt, TV>), dim3(c->nlb), dim3(64), 0, c->stream,
                                            c->d_lb_t, c->d_lb_kl, c->d_lslot, c->pitch, c->d_seg_cells, Bd, Aoff,
                                            Dpinv, Pm, G));
}
void launch_line_factor(cfd_ctx* c, const double* Bd, const void* Aoff, void* Dpinv, void* Pm, void* G, void* Ac,
                        int prec) {
  if (prec == 32) line_factor_t<float>(c, Bd, (const float*)Aoff, (float*)Dpinv, (float*)Pm, (float*)G, (float*)Ac);
  else line_factor_t<double>(c, Bd, (const double*)Aoff, (double*)Dpinv, (double*)Pm, (double*)G, (double*)Ac);
}
template <class TB, class TX, class TZ>
static void line_t(cfd_ctx* c, const TB* Dpinv, const TB* Pm, const TB* G, const TB* Ac, const TX* x, TZ* zc) {
  const int nc = (int)c->segcol_ptr.size() - 1;
  TZ* sL = (TZ*)c->kry.lineS;
  TZ* z = (TZ*)c->kry.lineZ;  // the sweeps' z by line position; zc (by cell) is formed at the end
  // z needs no clearing: with a proper segment coloring every z a cell reads
  // was written earlier in the same application (lower colors going up,
  // higher colors coming down), and only cells' entries are ever read
  NV_DISPATCH(c->P.nvar, {
    for (int s = 0; s < 2 * nc - 1; ++s) {
      const int col = s < nc ? s : 2 * nc - 2 - s;
      const int s0 = c->segcol_ptr[col], s1 = c->segcol_ptr[col + 1];
      if (s1 <= s0) continue;
      const int t0 = c->seg_ptr[s0], t1 = c->seg_ptr[s1];
      const int b0 = c->lbcol_ptr[col], nb = c->lbcol_ptr[col + 1] - b0;
      const int pass = s >= nc ? LINE_BWD : col == nc - 1 ? LINE_TOP : LINE_FWD;
      // small meshes are bound by launches: one launch per color; large ones
      // form w over all the color's cells in a higher-occupancy kernel first
      const bool fused = c->N < c->line_fused_max;
      if (fused) {
        const int ts = t0 + c->lmcol_cells[col];  // first single-cell position of the color
        const int nsb = (t1 - ts + 63) / 64;
        if (nb + nsb > 0)
          hipLaunchKernelGGL((k_line_sweep<NVt, TB, TX, TZ>), dim3(nb + nsb), dim3(64), 0, c->stream, b0, nb, ts, t1,
                             pass, c->d_lb_t, c->d_lb_kl, c->d_seg_cells, c->d_lnc, c->d_lmask, c->pitch, Dpinv,
                             Ac, Pm, G, x, sL, z);
      } else {
        hipLaunchKernelGGL((k_line_cells<NVt, TB, TX, TZ>), dim3(nblocks(t1 - t0, LINE_BS)), dim3(LINE_BS), 0,
                           c->stream, t0, t1, pass, c->d_seg_cells, c->d_lnc, c->d_lcomp, c->d_lmask, c->pitch, Dpinv,
                           Ac, x, sL, z, c->kry.lineW);
        if (nb > 0)
          hipLaunchKernelGGL((k_line_solve<NVt, TB, TZ>), dim3(nb), dim3(64), 0, c->stream, b0, c->d_lb_t,
                             c->d_lb_kl, c->d_seg_cells, c->pitch, Pm, G, (const double*)c->kry.lineW, z);
      }
    }
    hipLaunchKernelGGL((k_line_gather<NVt, TZ>), dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->N, c->pitch,
                       c->d_lpos, (const TZ*)z, zc);
  });
}
void launch_line_apply(cfd_ctx* c, const void* Dpinv, const void* Pm, const void* G, const void* Ac, const void* x,
                       int xprec, void* z, int zprec, int bprec) {
  if (bprec == 32) {
    if (xprec == 32 && zprec == 32)
      line_t<float, float, float>(c, (const float*)Dpinv, (const float*)Pm, (const float*)G, (const float*)Ac,
                                  (const float*)x, (float*)z);
    else
      line_t<float, double, double>(c, (const float*)Dpinv, (const float*)Pm, (const float*)G, (const float*)Ac,
                                    (const double*)x, (double*)z);
  } else {
    line_t<double, double, double>(c, (const double*)Dpinv, (const double*)Pm, (const double*)G,
                                   (const double*)Ac, (const double*)x, (double*)z);
  }
}
template <class TB, class TX, class TZ, bool TR>
static void sgs_t(cfd_ctx* c, const TB* Binv, const TB* Aoff, const TX* x, TZ* z) {
  const int nc = (int)c->ccell_ptr.size() - 1;
  NV_DISPATCH(c->P.nvar, {
    hipLaunchKernelGGL((k_sgs_first<NVt, TB, TX, TZ, TR>), dim3(nblocks(c->N)), dim3(BS), 0, c->stream, c->N,
                       c->pitch, c->d_ccol, c->d_cpos, Binv, x, z);
    for (int s = 1; s < 2 * nc - 1; ++s) {
      const int col = s < nc ? s : 2 * nc - 2 - s;
      const int b = c->ccell_ptr[col], n = c->ccell_ptr[col + 1] - b;
      if (n == 0) continue;
      hipLaunchKernelGGL((k_sgs_color<NVt, TB, TX, TZ, TR>), dim3(nblocks(n)), dim3(BS), 0, c->stream, n, b,
                         c->d_ccells, c->pitch, c->d_nq, c->d_cpos, c->d_ccol, s < nc ? col : 255, Binv, Aoff,
                         x, z);
    }
  });
}
void launch_sgs_apply(cfd_ctx* c, const void* Binv, const void* Aoff, const void* x, int xprec, void* z,
                      int zprec, int bprec, bool transpose) {
#define SGS_CASE(TB, TX, TZ)                                                                              \
  if (transpose) sgs_t<TB, TX, TZ, true>(c, (const TB*)Binv, (const TB*)Aoff, (const TX*)x, (TZ*)z);    \
  else sgs_t<TB, TX, TZ, false>(c, (const TB*)Binv, (const TB*)Aoff, (const TX*)x, (TZ*)z);
  if (bprec == 32) {
    if (xprec == 32 && zprec == 32) { SGS_CASE(float, float, float) }
    else { SGS_CASE(float, double, double) }
  } else {
    SGS_CASE(double, double, double)
  }
#undef SGS_CASE
}
template <class TV, bool TR> static void bapply_t(cfd_ctx* c, const TV* Binv, const TV* x, TV* z) {
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL((k_block_apply<NVt, TV, TV, TR>), dim3(nblocks(c->N)), dim3(BS), 0,
                                            c->stream, c->N, c->pitch, Binv, x, z));
}
void launch_block_apply(cfd_ctx* c, const void* Binv, const void* x, void* z, int prec, bool transpose) {
  if (prec == 32) {
    if (transpose) bapply_t<float, true>(c, (const float*)Binv, (const float*)x, (float*)z);
    else bapply_t<float, false>(c, (const float*)Binv, (const float*)x, (float*)z);
  } else {
    if (transpose) bapply_t<double, true>(c, (const double*)Binv, (const double*)x, (double*)z);
    else bapply_t<double, false>(c, (const double*)Binv, (const double*)x, (double*)z);
  }
}
template <class TV, bool TR>
static void solbj_t(cfd_ctx* c, double* x, const TV* V, const double* y, int k, const TV* B) {
  const size_t vs = (size_t)c->P.nvar * c->pitch;
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL((k_solution_bj<NVt, TV, TR>), dim3(nblocks(c->N)), dim3(BS), 0,
                                            c->stream, c->N, c->pitch, vs, V, y, k + 1, B, x));
}
void launch_solution_bj(cfd_ctx* c, double* x, const void* V, const double* y, int k, const void* Binv, int prec,
                        bool transpose) {
  if (prec == 32) {
    if (transpose) solbj_t<float, true>(c, x, (const float*)V, y, k, (const float*)Binv);
    else solbj_t<float, false>(c, x, (const float*)V, y, k, (const float*)Binv);
  } else {
    if (transpose) solbj_t<double, true>(c, x, (const double*)V, y, k, (const double*)Binv);
    else solbj_t<double, false>(c, x, (const double*)V, y, k, (const double*)Binv);
  }
}

template <class TV>
static void solution_t(cfd_ctx* c, double* x, const TV* V, const double* y, int k, const double* D) {
  const size_t vs = (size_t)c->P.nvar * c->pitch;
  NV_DISPATCH(c->P.nvar, hipLaunchKernelGGL((k_solution<NVt, TV>), dim3(nblocks(c->N)), dim3(BS), 0, c->stream,
                                            c->N, c->pitch, vs, V, y, k + 1, D, x));
}
void launch_solution(cfd_ctx* c, double* x, const void* V, const double* y, int k, int prec, bool diag) {
  const double* D = diag ? c->kry.D : nullptr;
  if (prec == 32) solution_t<float>(c, x, (const float*)V, y, k, D);
  else solution_t<double>(c, x, (const double*)V, y, k, D);
}

}  // namespace cfd

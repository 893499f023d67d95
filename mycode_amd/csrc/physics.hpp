// Pointwise numerics of the residual, written once for any scalar type T
// (double for the residual, float for the FP32 path, Dual<> for the forward
// tangent).  Every expression keeps the operand order of the reference so
// that, built without multiply-add contraction, the FP64 results agree with
// the reference bit for bit.
//
//   pressure / normal flux      Flux.hpp:12-35
//   HLLC                        Flux.hpp:36-74
//   LLF (Rusanov)               Flux.hpp:81-110
//   Roe + Harten entropy fix    Flux.hpp:112-213
//   ghost states (per BC set)   Solver.cpp:806-863, 973-1011, 1088-1139
//   primitive -> conservative   Solver.cpp:498-501
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "../../include/cfd_gpu.h"

#define HD __host__ __device__ __forceinline__

namespace cfd {

// ---- FP64 quotients that share their denominator's reciprocal -----------------
// gfx950 forms an IEEE FP64 quotient a / b as: r = v_rcp_f64(b) refined by two
// Newton steps r += r (1 - b r); q = a r; q += (a - b q) r; plus operand
// scaling (v_div_scale / v_div_fmas) that only acts when an exponent nears the
// ends of the double range, and a fix-up (v_div_fixup) for zero, infinite and
// NaN operands.  r depends on b alone, so `dv` writes the same steps without
// the scaling and fix-up: each quotient is bit-identical to a / b for finite,
// nonzero b of ordinary magnitude (densities, pressures, lengths, coefficients
// here), and the compiler merges the r of quotients with one denominator
// (u, v, k, omega by rho; the 14 gradients by the volume), which saves seven
// of the ten instructions of every further quotient.  Host code divides.
HD double dv(double a, double b) {
#ifdef __HIP_DEVICE_COMPILE__
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  const double q = a * r;
  return __builtin_fma(__builtin_fma(-b, q, a), r, q);
#else
  return a / b;
#endif
}
template <class A, class B> HD auto dv(A a, B b) -> decltype(a / b) { return a / b; }

// ---- dual numbers for forward-mode differentiation -------------------------
template <class F> struct Dual {
  F v, d;
  HD Dual() : v(0), d(0) {}
  HD Dual(F a) : v(a), d(0) {}
  HD Dual(F a, F b) : v(a), d(b) {}
};
template <class F> HD Dual<F> operator+(Dual<F> a, Dual<F> b) { return Dual<F>(a.v + b.v, a.d + b.d); }
template <class F> HD Dual<F> operator-(Dual<F> a, Dual<F> b) { return Dual<F>(a.v - b.v, a.d - b.d); }
template <class F> HD Dual<F> operator-(Dual<F> a) { return Dual<F>(-a.v, -a.d); }
template <class F> HD Dual<F> operator*(Dual<F> a, Dual<F> b) { return Dual<F>(a.v * b.v, a.d * b.v + a.v * b.d); }
template <class F> HD Dual<F> operator/(Dual<F> a, Dual<F> b) {
  F q = a.v / b.v;
  return Dual<F>(q, (a.d - q * b.d) / b.v);
}
// mixed operations with plain arithmetic constants: the constant is first
// converted to the dual's component type, so an FP32 dual stays in FP32
#define CFD_ARITH template <class F, class A, class = typename std::enable_if<std::is_arithmetic<A>::value>::type>
CFD_ARITH HD Dual<F> operator+(Dual<F> a, A b) { return Dual<F>(a.v + F(b), a.d); }
CFD_ARITH HD Dual<F> operator+(A a, Dual<F> b) { return Dual<F>(F(a) + b.v, b.d); }
CFD_ARITH HD Dual<F> operator-(Dual<F> a, A b) { return Dual<F>(a.v - F(b), a.d); }
CFD_ARITH HD Dual<F> operator-(A a, Dual<F> b) { return Dual<F>(F(a) - b.v, -b.d); }
CFD_ARITH HD Dual<F> operator*(Dual<F> a, A b) { return Dual<F>(a.v * F(b), a.d * F(b)); }
CFD_ARITH HD Dual<F> operator*(A a, Dual<F> b) { return Dual<F>(F(a) * b.v, F(a) * b.d); }
CFD_ARITH HD Dual<F> operator/(Dual<F> a, A b) { return Dual<F>(a.v / F(b), a.d / F(b)); }
CFD_ARITH HD Dual<F> operator/(A a, Dual<F> b) {
  F q = F(a) / b.v;
  return Dual<F>(q, -q * b.d / b.v);
}

// FP64 dual quotients: both parts divide by b.v, so they share its reciprocal
// (same values as the generic quotients above).
HD Dual<double> operator/(Dual<double> a, Dual<double> b) {
  const double q = dv(a.v, b.v);
  return Dual<double>(q, dv(a.d - q * b.d, b.v));
}
HD Dual<double> operator/(Dual<double> a, double b) { return Dual<double>(dv(a.v, b), dv(a.d, b)); }
HD Dual<double> operator/(double a, Dual<double> b) {
  const double q = dv(a, b.v);
  return Dual<double>(q, dv(-q * b.d, b.v));
}

// FP32 duals serve only the FP32 tangent, which keeps the derivative parts;
// their value parts (which steer the branches) are FP32 approximations of the
// FP64 state anyway.  Quotients and square roots therefore use the hardware
// reciprocal and square root (about 1 ulp) instead of the correctly rounded
// sequences, with one reciprocal per dual quotient instead of two divisions.
HD float fast_rcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
HD float fast_sqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_sqrtf(x);
#else
  return std::sqrt(x);
#endif
}
HD Dual<float> operator/(Dual<float> a, Dual<float> b) {
  const float r = fast_rcp(b.v), q = a.v * r;
  return Dual<float>(q, (a.d - q * b.d) * r);
}
HD Dual<float> operator/(Dual<float> a, double b) {
  const float r = fast_rcp(float(b));
  return Dual<float>(a.v * r, a.d * r);
}
HD Dual<float> operator/(double a, Dual<float> b) {
  const float r = fast_rcp(b.v), q = float(a) * r;
  return Dual<float>(q, -q * b.d * r);
}

// K-direction forward dual: value plus K directional derivatives, used for
// face-local Jacobian-transpose products (one pass covers K inputs).
template <class F, int K> struct DualV {
  F v;
  F d[K];
  HD DualV() : v(0) {
#pragma unroll
    for (int i = 0; i < K; ++i) d[i] = 0;
  }
  HD DualV(F a) : v(a) {
#pragma unroll
    for (int i = 0; i < K; ++i) d[i] = 0;
  }
};
#define DV_BIN(op, dexpr)                                                      \
  template <class F, int K> HD DualV<F, K> operator op(const DualV<F, K>& a, const DualV<F, K>& b) { \
    DualV<F, K> r;                                                             \
    r.v = a.v op b.v;                                                          \
    _Pragma("unroll") for (int i = 0; i < K; ++i) r.d[i] = dexpr;              \
    return r;                                                                  \
  }
DV_BIN(+, a.d[i] + b.d[i])
DV_BIN(-, a.d[i] - b.d[i])
DV_BIN(*, a.d[i] * b.v + a.v * b.d[i])
#undef DV_BIN
// The quotients of a K-direction dual all divide by b.v: written with dv
// (FP64), the K + 1 quotients share one reciprocal instead of running K + 1
// divisions, with the same values.
template <class F, int K> HD DualV<F, K> operator/(const DualV<F, K>& a, const DualV<F, K>& b) {
  DualV<F, K> r;
  r.v = dv(a.v, b.v);
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = dv(a.d[i] - r.v * b.d[i], b.v);
  return r;
}
template <class F, int K> HD DualV<F, K> operator-(const DualV<F, K>& a) {
  DualV<F, K> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = -a.d[i];
  return r;
}
#define CFD_ARITHV template <class F, int K, class A, class = typename std::enable_if<std::is_arithmetic<A>::value>::type>
CFD_ARITHV HD DualV<F, K> operator+(const DualV<F, K>& a, A b) { DualV<F, K> r = a; r.v = a.v + F(b); return r; }
CFD_ARITHV HD DualV<F, K> operator+(A b, const DualV<F, K>& a) { DualV<F, K> r = a; r.v = F(b) + a.v; return r; }
CFD_ARITHV HD DualV<F, K> operator-(const DualV<F, K>& a, A b) { DualV<F, K> r = a; r.v = a.v - F(b); return r; }
CFD_ARITHV HD DualV<F, K> operator-(A b, const DualV<F, K>& a) {
  DualV<F, K> r;
  r.v = F(b) - a.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = -a.d[i];
  return r;
}
CFD_ARITHV HD DualV<F, K> operator*(const DualV<F, K>& a, A b) {
  DualV<F, K> r;
  r.v = a.v * F(b);
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = a.d[i] * F(b);
  return r;
}
CFD_ARITHV HD DualV<F, K> operator*(A b, const DualV<F, K>& a) { return a * b; }
CFD_ARITHV HD DualV<F, K> operator/(const DualV<F, K>& a, A b) {
  DualV<F, K> r;
  r.v = dv(a.v, F(b));
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = dv(a.d[i], F(b));
  return r;
}
CFD_ARITHV HD DualV<F, K> operator/(A b, const DualV<F, K>& a) {
  DualV<F, K> r;
  r.v = dv(F(b), a.v);
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = dv(-r.v * a.d[i], a.v);
  return r;
}
template <class F, int K> HD F val(const DualV<F, K>& x) { return x.v; }
template <class F, int K> HD DualV<F, K> sq_rt(const DualV<F, K>& x) {
  DualV<F, K> r;
  r.v = sqrt(x.v);
  const F h = F(2) * r.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = dv(x.d[i], h);
  return r;
}
template <class F, int K, class A> HD DualV<F, K> pw(const DualV<F, K>& a, A bb) {
  const F b = F(bb);
  DualV<F, K> r;
  r.v = pow(a.v, b);
  const F dp = b * pow(a.v, b - F(1));
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = dp * a.d[i];
  return r;
}

template <class T> struct Real { typedef T type; };
template <class F> struct Real<Dual<F>> { typedef F type; };

HD double val(double x) { return x; }
HD float val(float x) { return x; }
template <class F> HD F val(Dual<F> x) { return x.v; }

HD double sq_rt(double x) { return sqrt(x); }
HD float sq_rt(float x) { return sqrtf(x); }
template <class F> HD Dual<F> sq_rt(Dual<F> x) {
  F s = sq_rt(x.v);
  return Dual<F>(s, x.d / (F(2) * s));
}
HD Dual<float> sq_rt(Dual<float> x) {
  const float s = fast_sqrt(x.v);
  return Dual<float>(s, x.d * (0.5f * fast_rcp(s)));
}
HD double pw(double a, double b) { return pow(a, b); }
HD float pw(float a, float b) { return powf(a, b); }
template <class F, class A> HD Dual<F> pw(Dual<F> a, A bb) {
  F b = F(bb);
  F p = pw(a.v, b);
  return Dual<F>(p, b * pw(a.v, b - F(1)) * a.d);
}
// |x|, std::min, std::max with the reference's tie behaviour
template <class T> HD T ab(T x) { return val(x) < 0 ? T(-x) : x; }
template <class T> HD T mn(T a, T b) { return val(b) < val(a) ? b : a; }
template <class T> HD T mx(T a, T b) { return val(a) < val(b) ? b : a; }

// ---- state ---------------------------------------------------------------------
template <class T> struct S6 {
  T q[6];
  HD S6() {
#pragma unroll
    for (int i = 0; i < 6; ++i) q[i] = T(0);
  }
  HD S6(T a, T b, T c, T d, T e = T(0), T f = T(0)) {
    q[0] = a; q[1] = b; q[2] = c; q[3] = d; q[4] = e; q[5] = f;
  }
  HD T& operator[](int i) { return q[i]; }
  HD const T& operator[](int i) const { return q[i]; }
};
template <class T> HD S6<T> add(const S6<T>& a, const S6<T>& b) {
  S6<T> s;
#pragma unroll
  for (int i = 0; i < 6; ++i) s.q[i] = a.q[i] + b.q[i];
  return s;
}
template <class T> HD S6<T> sub(const S6<T>& a, const S6<T>& b) {
  S6<T> s;
#pragma unroll
  for (int i = 0; i < 6; ++i) s.q[i] = a.q[i] - b.q[i];
  return s;
}
template <class T> HD S6<T> scale(const S6<T>& a, T k) {
  S6<T> s;
#pragma unroll
  for (int i = 0; i < 6; ++i) s.q[i] = a.q[i] * k;
  return s;
}

// ---- parameters as the kernels see them ---------------------------------------------
struct Phys {
  double gamma, R, P0, T0, pexit, nu, Pr, Cp;
  double fs[6];  // frozen freestream (SUPERSONIC tag 1, PLATE tags 1 and 5)
  double sh[6];  // frozen post-shock state (SWBLI tag 6)
  double kmin, wmin;  // SST positivity floors of specific k and omega
  int flux, flux_alt, spatial, bcset, viscous, nvar, turb;
  int turb1;  // SST: first-order upwind convection of k and omega
};

template <class T> HD T pressure(const Phys& P, const S6<T>& U) {
  T V2 = dv(U[1] * U[1] + U[2] * U[2], U[0] * U[0]);
  return (P.gamma - 1.0) * (U[3] - 0.5 * U[0] * V2);
}
template <class T> HD S6<T> prim2cons(const Phys& P, T rho, T u, T v, T p) {
  T rE = dv(p, P.gamma - 1.0) + 0.5 * rho * (u * u + v * v);
  return S6<T>(rho, rho * u, rho * v, rE, rho * T(0.0), rho * T(0.0));
}
template <class T> HD S6<T> prim2cons(const Phys& P, T rho, T u, T v, T p, T k, T w) {
  T rE = dv(p, P.gamma - 1.0) + 0.5 * rho * (u * u + v * v);
  return S6<T>(rho, rho * u, rho * v, rE, rho * k, rho * w);
}

// ---- SST k-omega (no reference counterpart: Menter SST-2003 as defined by
// the NASA Turbulence Modeling Resource; the reference reserves the rhok, rhow
// slots, Types.hpp:61-64, and the wall distance, Solver.cpp:1981-2015) -------
namespace sst {
constexpr double sk1 = 0.85, sw1 = 0.5, b1 = 0.075, sk2 = 1.0, sw2 = 0.856, b2 = 0.0828, bs = 0.09, a1 = 0.31,
                 kap = 0.41, Prt = 0.9;
HD double g1() { return b1 / bs - sw1 * kap * kap / sqrt(bs); }
HD double g2() { return b2 / bs - sw2 * kap * kap / sqrt(bs); }
}  // namespace sst

HD double th(double x) { return tanh(x); }
HD float th(float x) { return tanhf(x); }
template <class F> HD Dual<F> th(Dual<F> x) {
  F t = th(x.v);
  return Dual<F>(t, (F(1) - t * t) * x.d);
}
template <class F, int K> HD DualV<F, K> th(const DualV<F, K>& x) {
  DualV<F, K> r;
  r.v = tanh(x.v);
  const F dt = F(1) - r.v * r.v;
#pragma unroll
  for (int i = 0; i < K; ++i) r.d[i] = dt * x.d[i];
  return r;
}

// Eddy viscosity, blending F1 and the k / omega sources (per unit volume) of
// a cell from its state, its 14 gradients (rho u v p T k w, x then y) and its
// wall distance d.
template <class T> struct TurbQ {
  T mut, F1, Sk, Sw;
};
template <class T> HD TurbQ<T> turb_cell(const Phys& P, const S6<T>& U, const T* g, double d) {
  using namespace sst;
  const double nu = P.nu;
  const T rho = U[0];
  const T k = mx(dv(U[4], rho), T(1e-20)), w = mx(dv(U[5], rho), T(1e-20));
  const T ux = g[2], uy = g[3], vx = g[4], vy = g[5];
  const T S2 = 2.0 * ux * ux + 2.0 * vy * vy + (uy + vx) * (uy + vx);
  const T S = sq_rt(S2 + 1e-20);
  const T gkw = g[10] * g[12] + g[11] * g[13];
  const T CD = mx(dv(2.0 * rho * sw2 * gkw, w), T(1e-10));
  const T sk = sq_rt(k);
  const T arg1 = mn(mx(dv(sk, bs * w * d), dv(500.0 * nu, d * d * w)), dv(4.0 * rho * sw2 * k, CD * d * d));
  const T F1 = th(arg1 * arg1 * arg1 * arg1);
  const T arg2 = mx(dv(2.0 * sk, bs * w * d), dv(500.0 * nu, d * d * w));
  const T F2 = th(arg2 * arg2);
  const T mut = dv(rho * a1 * k, mx(a1 * w, S * F2));
  const T Pk = mn(mut * S2, 10.0 * bs * rho * w * k);
  const T beta = F1 * b1 + (1.0 - F1) * b2;
  const T gam = F1 * g1() + (1.0 - F1) * g2();
  TurbQ<T> q;
  q.mut = mut;
  q.F1 = F1;
  q.Sk = Pk - bs * rho * w * k;
  q.Sw = gam * rho * S2 - beta * rho * w * w + dv(2.0 * (1.0 - F1) * rho * sw2 * gkw, w);
  return q;
}

// Reverse sweep of turb_cell for double inputs: adds to Ub (rows 0, 4, 5) and
// gb (slots 2-5, 10-13) the gradient of lk Sk + lw Sw + amut mu_t + aF1 F1
// with respect to the cell state and gradients.  The forward values are
// recomputed in turb_cell's order and every min / max takes the branch
// turb_cell takes.
HD void turb_cell_vjp(const Phys& P, const S6<double>& U, const double* g, double d, double lk, double lw,
                      double amut, double aF1, double (&Ub)[6], double (&gb)[14]) {
  using namespace sst;
  const double nu = P.nu;
  // ---- forward
  const double rho = U[0];
  const double k0 = dv(U[4], rho), w0 = dv(U[5], rho);
  const bool kcl = k0 < 1e-20, wcl = w0 < 1e-20;
  const double k = kcl ? 1e-20 : k0, w = wcl ? 1e-20 : w0;
  const double ux = g[2], uy = g[3], vx = g[4], vy = g[5];
  const double S2 = 2.0 * ux * ux + 2.0 * vy * vy + (uy + vx) * (uy + vx);
  const double S = sqrt(S2 + 1e-20);
  const double gkw = g[10] * g[12] + g[11] * g[13];
  const double CD0 = dv(2.0 * rho * sw2 * gkw, w);
  const bool cdcl = CD0 < 1e-10;
  const double CD = cdcl ? 1e-10 : CD0;
  const double sk = sqrt(k);
  const double A1 = dv(sk, bs * w * d), A2 = dv(500.0 * nu, d * d * w), A3 = dv(4.0 * rho * sw2 * k, CD * d * d);
  const bool a2big = A1 < A2;
  const double mA = a2big ? A2 : A1;
  const bool a3min = A3 < mA;
  const double arg1 = a3min ? A3 : mA;
  const double F1 = tanh(arg1 * arg1 * arg1 * arg1);
  const double B1 = dv(2.0 * sk, bs * w * d);
  const bool b2big = B1 < A2;
  const double arg2 = b2big ? A2 : B1;
  const double F2 = tanh(arg2 * arg2);
  const bool dsf = a1 * w < S * F2;
  const double den = dsf ? S * F2 : a1 * w;
  const double mut = dv(rho * a1 * k, den);
  const double P1 = mut * S2, P2 = 10.0 * bs * rho * w * k;
  const bool p2min = P2 < P1;
  const double beta = F1 * b1 + (1.0 - F1) * b2;
  const double gam = F1 * g1() + (1.0 - F1) * g2();
  // ---- reverse
  double rhob = 0.0, kb = 0.0, wb = 0.0, S2b = 0.0, gkwb = 0.0, F1b = aF1, mutb = amut;
  // Sw = gam rho S2 - beta rho w^2 + 2 (1 - F1) rho sw2 gkw / w
  const double cw = dv(2.0 * rho * sw2 * gkw, w);
  const double gamb = lw * rho * S2, betab = -lw * rho * w * w;
  rhob += lw * (gam * S2 - beta * w * w + dv(2.0 * (1.0 - F1) * sw2 * gkw, w));
  S2b += lw * gam * rho;
  wb += lw * (-2.0 * beta * rho * w - (1.0 - F1) * dv(cw, w));
  F1b += -lw * cw;
  gkwb += lw * dv(2.0 * (1.0 - F1) * rho * sw2, w);
  // Sk = Pk - bs rho w k
  const double Pkb = lk;
  rhob -= lk * bs * w * k;
  wb -= lk * bs * rho * k;
  kb -= lk * bs * rho * w;
  F1b += betab * (b1 - b2) + gamb * (g1() - g2());
  // Pk = min(mut S2, 10 bs rho w k)
  if (p2min) {
    rhob += Pkb * 10.0 * bs * w * k;
    wb += Pkb * 10.0 * bs * rho * k;
    kb += Pkb * 10.0 * bs * rho * w;
  } else {
    mutb += Pkb * S2;
    S2b += Pkb * mut;
  }
  // mut = rho a1 k / den, den = max(a1 w, S F2)
  rhob += dv(mutb * a1 * k, den);
  kb += dv(mutb * rho * a1, den);
  const double denb = -dv(mutb * mut, den);
  double Sb = 0.0, F2b = 0.0;
  if (dsf) {
    Sb += denb * F2;
    F2b += denb * S;
  } else {
    wb += denb * a1;
  }
  // F2 = tanh(arg2^2), arg2 = max(B1, A2)
  const double arg2b = F2b * (1.0 - F2 * F2) * 2.0 * arg2;
  double A2b = 0.0, B1b = 0.0, A1b = 0.0, A3b = 0.0;
  if (b2big) A2b += arg2b;
  else B1b += arg2b;
  // F1 = tanh(arg1^4), arg1 = min(max(A1, A2), A3)
  const double arg1b = F1b * (1.0 - F1 * F1) * 4.0 * arg1 * arg1 * arg1;
  if (a3min) {
    A3b += arg1b;
  } else if (a2big) {
    A2b += arg1b;
  } else {
    A1b += arg1b;
  }
  // A1 = sk / (bs w d), B1 = 2 sk / (bs w d), A2 = 500 nu / (d^2 w), A3 = 4 rho sw2 k / (CD d^2)
  const double skb = dv(A1b + 2.0 * B1b, bs * w * d);
  wb -= dv(A1b * A1 + B1b * B1 + A2b * A2, w);
  const double a3c = dv(4.0 * sw2, CD * d * d);
  rhob += A3b * a3c * k;
  kb += A3b * a3c * rho;
  const double CDb = -dv(A3b * A3, CD);
  kb += dv(skb, 2.0 * sk);
  // CD = max(2 rho sw2 gkw / w, 1e-10)
  if (!cdcl) {
    rhob += CDb * dv(2.0 * sw2 * gkw, w);
    gkwb += CDb * dv(2.0 * rho * sw2, w);
    wb -= dv(CDb * CD0, w);
  }
  // S = sqrt(S2 + 1e-20)
  S2b += dv(Sb, 2.0 * S);
  // S2 = 2 ux^2 + 2 vy^2 + (uy + vx)^2, gkw = kx wx + ky wy
  gb[2] += S2b * 4.0 * ux;
  gb[5] += S2b * 4.0 * vy;
  gb[3] += S2b * 2.0 * (uy + vx);
  gb[4] += S2b * 2.0 * (uy + vx);
  gb[10] += gkwb * g[12];
  gb[12] += gkwb * g[10];
  gb[11] += gkwb * g[13];
  gb[13] += gkwb * g[11];
  // k = max(U4 / rho, 1e-20), w = max(U5 / rho, 1e-20), rho = U0
  const double k0b = kcl ? 0.0 : kb, w0b = wcl ? 0.0 : wb;
  Ub[4] += dv(k0b, rho);
  Ub[5] += dv(w0b, rho);
  Ub[0] += rhob - dv(k0b * k0 + w0b * w0, rho);
}
// omega at a no-slip wall: 10 * 6 nu / (beta1 dn^2), dn the owner centroid's normal distance
HD double omega_wall(const Phys& P, double dn) { return 60.0 * P.nu / (sst::b1 * dn * dn); }
template <class T> HD S6<T> frozen(const double* s) {
  return S6<T>(T(s[0]), T(s[1]), T(s[2]), T(s[3]), T(s[4]), T(s[5]));
}
template <class T> HD S6<T> slip_ghost(const S6<T>& U, double nx, double ny) {
  T u = dv(U[1], U[0]), v = dv(U[2], U[0]);
  T vn = u * nx + v * ny;
  T ug = u - 2.0 * vn * nx, vg = v - 2.0 * vn * ny;
  S6<T> G = U;
  G[1] = U[0] * ug;
  G[2] = U[0] * vg;
  return G;
}

// Ghost state of a boundary face with tag `tag` seen from interior state U.
template <class T> HD S6<T> ghost(const Phys& P, const S6<T>& U, int tag, double nx, double ny, double dn = 0.0) {
  if (P.bcset == CFD_BC_BUMP) {  // Solver.cpp:806-863
    if (tag == 1) {
      T pin = pressure(P, U);
      T puse = mn(pin, T(P.P0 - 1e-4));
      T Tg = P.T0 * pw(dv(puse, P.P0), dv(P.gamma - 1.0, P.gamma));
      T rg = dv(puse, P.R * Tg);
      T V2 = 2.0 * P.Cp * (P.T0 - Tg);
      T V = val(V2) > 0 ? sq_rt(V2) : T(0.0);
      return prim2cons(P, rg, V, T(0.0), puse);
    }
    if (tag == 2) {
      T rho = U[0], u = dv(U[1], U[0]), v = dv(U[2], U[0]);
      return prim2cons(P, rho, u, v, T(P.pexit));
    }
    if (tag == 3) return slip_ghost(U, nx, ny);
    return U;
  }
  if (P.bcset == CFD_BC_SUPERSONIC) {  // Solver.cpp:973-1011
    if (tag == 1) return frozen<T>(P.fs);
    if (tag == 3) return slip_ghost(U, nx, ny);
    return U;  // tag 2 extrapolates; untagged faces mirror the interior
  }
  // PLATE (Solver.cpp:1088-1139) and SWBLI (PLATE + tag 6)
  if (tag == 1 || tag == 5) return frozen<T>(P.fs);
  if (tag == 3) return slip_ghost(U, nx, ny);
  if (tag == 4) {
    T ug = -dv(U[1], U[0]), vg = -dv(U[2], U[0]);
    T ek = dv(0.5 * (U[1] * U[1] + U[2] * U[2]), U[0]);
    T pin = (P.gamma - 1.0) * (U[3] - ek);
    T Tin = dv(pin, U[0] * P.R);
    T rg = dv(pin, P.R * Tin);
    if (P.turb) {  // k = 0 and omega = omega_wall on the wall face
      T kg = -dv(U[4], U[0]);
      T wg = 2.0 * omega_wall(P, dn) - dv(U[5], U[0]);
      return prim2cons(P, rg, ug, vg, pin, kg, wg);
    }
    return prim2cons(P, rg, ug, vg, pin);
  }
  if (tag == 6 && P.bcset == CFD_BC_SWBLI) return frozen<T>(P.sh);
  return U;
}

template <class T> HD S6<T> normal_flux(const Phys& P, const S6<T>& U, double nx, double ny) {
  T p = pressure(P, U);
  T u = dv(U[1], U[0]), v = dv(U[2], U[0]);
  T un = u * nx + v * ny;
  return S6<T>(U[0] * un, U[1] * un + p * nx, U[2] * un + p * ny, (U[3] + p) * un, T(0.0), T(0.0));
}

template <class T> HD S6<T> flux_hllc(const Phys& P, const S6<T>& L, const S6<T>& R, double nx, double ny) {
  const double g = P.gamma;
  T pL = pressure(P, L), pR = pressure(P, R);
  T uL = dv(L[1], L[0]), uR = dv(R[1], R[0]);
  T vL = dv(L[2], L[0]), vR = dv(R[2], R[0]);
  T unL = uL * nx + vL * ny, unR = uR * nx + vR * ny;
  T cL = sq_rt(dv(g * pL, L[0])), cR = sq_rt(dv(g * pR, R[0]));
  T hL = dv(L[3] + pL, L[0]), hR = dv(R[3] + pR, R[0]);
  T sL = sq_rt(L[0]), sR = sq_rt(R[0]);
  T ut = dv(sL * uL + sR * uR, sL + sR);
  T vt = dv(sL * vL + sR * vR, sL + sR);
  T ht = dv(sL * hL + sR * hR, sL + sR);
  T unt = ut * nx + vt * ny;
  T ct = sq_rt((g - 1.0) * (ht - 0.5 * (ut * ut + vt * vt)));
  T SL = mn(unL - cL, unt - ct);
  T SR = mx(unR + cR, unt + ct);
  T Ss = dv(pR - pL + L[0] * unL * (SL - unL) - R[0] * unR * (SR - unR), L[0] * (SL - unL) - R[0] * (SR - unR));
  if (val(SL) >= 0.0) return normal_flux(P, L, nx, ny);
  if (val(SR) <= 0.0) return normal_flux(P, R, nx, ny);
  // star region: upwind side K with wave speed S
  const bool left = val(Ss) >= 0.0;
  const S6<T>& K = left ? L : R;
  T S = left ? SL : SR;
  T pK = left ? pL : pR, uK = left ? uL : uR, vK = left ? vL : vR, unK = left ? unL : unR;
  T coef = dv(K[0] * (S - unK), S - Ss);
  S6<T> Us(coef, coef * (uK + (Ss - unK) * nx), coef * (vK + (Ss - unK) * ny),
           coef * (dv(K[3], K[0]) + (Ss - unK) * (Ss + dv(pK, K[0] * (S - unK)))), T(0.0), T(0.0));
  return add(normal_flux(P, K, nx, ny), scale(sub(Us, K), S));
}

template <class T> HD S6<T> flux_llf(const Phys& P, const S6<T>& L, const S6<T>& R, double nx, double ny) {
  const double g = P.gamma;
  S6<T> FL = normal_flux(P, L, nx, ny), FR = normal_flux(P, R, nx, ny);
  T pL = pressure(P, L);
  T uL = dv(L[1], L[0]), vL = dv(L[2], L[0]);
  T unL = uL * nx + vL * ny;
  T cL = sq_rt(dv(g * pL, L[0]));
  T pR = pressure(P, R);
  T uR = dv(R[1], R[0]), vR = dv(R[2], R[0]);
  T unR = uR * nx + vR * ny;
  T cR = sq_rt(dv(g * pR, R[0]));
  T a = mx(ab(unL) + cL, ab(unR) + cR);
  S6<T> D = scale(sub(R, L), T(0.5 * a));
  S6<T> A = scale(add(FL, FR), T(0.5));
  return sub(A, D);
}

template <class T> HD S6<T> flux_roe(const Phys& P, const S6<T>& L, const S6<T>& R, double nx, double ny) {
  const double g = P.gamma;
  T pL = pressure(P, L);
  T uL = dv(L[1], L[0]), vL = dv(L[2], L[0]);
  T hL = dv(L[3] + pL, L[0]);
  T pR = pressure(P, R);
  T uR = dv(R[1], R[0]), vR = dv(R[2], R[0]);
  T hR = dv(R[3] + pR, R[0]);
  T sL = sq_rt(L[0]), sR = sq_rt(R[0]);
  T inv = dv(1.0, sL + sR);
  T ut = (sL * uL + sR * uR) * inv;
  T vt = (sL * vL + sR * vR) * inv;
  T ht = (sL * hL + sR * hR) * inv;
  T V2 = ut * ut + vt * vt;
  T c2 = (g - 1.0) * (ht - 0.5 * V2);
  T ct = sq_rt(mx(c2, T(1e-8)));
  T unt = ut * nx + vt * ny;
  T utt = -ut * ny + vt * nx;
  T lam[4] = {unt - ct, unt, unt, unt + ct};
  T eps = 0.1 * ct;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (val(ab(lam[i])) < val(eps)) lam[i] = 0.5 * (dv(lam[i] * lam[i], eps) + eps);
    lam[i] = ab(lam[i]);
  }
  T drho = R[0] - L[0], du = uR - uL, dvl = vR - vL, dp = pR - pL;
  T dun = du * nx + dvl * ny;
  T dut = -du * ny + dvl * nx;
  T rt = sL * sR;
  T a0 = dv(0.5, c2) * (dp - rt * ct * dun);
  T a1 = drho - dv(dp, c2);
  T a2 = rt * dut;
  T a3 = dv(0.5, c2) * (dp + rt * ct * dun);
  T w1 = lam[0] * a0, w2 = lam[1] * a1, w3 = lam[2] * a2, w4 = lam[3] * a3;
  S6<T> D;
  D[0] = w1 * 1.0 + w2 * 1.0 + w3 * 0.0 + w4 * 1.0;
  D[1] = w1 * (ut - ct * nx) + w2 * ut + w3 * (-ny) + w4 * (ut + ct * nx);
  D[2] = w1 * (vt - ct * ny) + w2 * vt + w3 * nx + w4 * (vt + ct * ny);
  D[3] = w1 * (ht - unt * ct) + w2 * (0.5 * V2) + w3 * utt + w4 * (ht + unt * ct);
  D[4] = ab(unt) * (dv(R[4], R[0]) - dv(L[4], L[0])) * rt;
  D[5] = ab(unt) * (dv(R[5], R[0]) - dv(L[5], L[0])) * rt;
  S6<T> FL = normal_flux(P, L, nx, ny), FR = normal_flux(P, R, nx, ny);
  return sub(scale(add(FL, FR), T(0.5)), scale(D, T(0.5)));
}

template <class T> HD S6<T> riemann(const Phys& P, const S6<T>& L, const S6<T>& R, double nx, double ny, int scheme) {
  if (scheme == CFD_FLUX_ROE) return flux_roe(P, L, R, nx, ny);
  if (scheme == CFD_FLUX_HLLC) return flux_hllc(P, L, R, nx, ny);
  return flux_llf(P, L, R, nx, ny);  // LLF, and AUSM falls back to LLF (Flux.hpp:224-225)
}

}  // namespace cfd

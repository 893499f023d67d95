// C-ABI entry points (include/cfd_gpu.h): context set-up, the host-pointer
// drop-ins for the reference Solver members, and the host side of GMRES and
// the Newton driver.  Only scalars (norms, one Hessenberg column per Arnoldi
// step) cross PCIe inside a solve; every vector stays in HBM.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ctx.hpp"

using namespace cfd;

namespace {

// scal[] slots
enum { S_UNORM = 0, S_VNORM = 1, S_BETA = 2, S_VW = 3, S_VWVOL = 4, S_L2 = 5, S_DOT = 6, S_W = 7 };

#define HIP_TRY(ctx, expr)                                                          \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess) {                                                         \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);               \
      return CFD_ERR_HIP;                                                           \
    }                                                                               \
  } while (0)

int fail(cfd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

// The products hold the limiter frozen (Solver.cpp:1366 evaluates with
// update_limiters = false), so with SECOND_ORDER_LIMITED one must exist.
int need_limiter(cfd_ctx* c) {
  if (c->P.spatial == CFD_SECOND_ORDER_LIMITED && !c->phi_valid)
    return fail(c, CFD_ERR_STATE, "limiter not formed: evaluate the residual first");
  return CFD_OK;
}

template <class T> hipError_t dalloc(T** p, size_t n) { return hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)); }

// The structured reverse product's face buffer, allocated on first use (a
// context that never forms a reverse product does not hold it).
int ensure_adjoint(cfd_ctx* c) {
  if (c->adjF) return CFD_OK;
  HIP_TRY(c, dalloc(&c->adjF, (size_t)cfd::VJP_ROW_MAX * c->F));
  return CFD_OK;
}

void fill_phys(cfd_ctx* c) {
  const cfd_params& p = c->P;
  Phys& q = c->phys;
  q.gamma = p.gamma;
  q.R = p.R_gas;
  q.P0 = p.P_total_inlet;
  q.T0 = p.T_total_inlet;
  q.pexit = p.P_static_exit;
  q.nu = p.nu;
  q.Pr = p.Pr;
  q.Cp = p.gamma * p.R_gas / (p.gamma - 1.0);
  // frozen freestream, Solver.cpp:1094-1100
  const double g = p.gamma;
  const double fT = 1.0 + 0.5 * (g - 1.0) * p.M_inf * p.M_inf;
  const double Ts = p.T_total_inlet / fT;
  const double rho = p.P_static_exit / (p.R_gas * Ts);
  const double cs = std::sqrt(g * p.R_gas * Ts);
  const double u = p.M_inf * cs;
  const double v = 0.0;
  q.fs[0] = rho;
  q.fs[1] = rho * u;
  q.fs[2] = rho * v;
  q.fs[3] = p.P_static_exit / (g - 1.0) + 0.5 * rho * (u * u + v * v);
  q.fs[4] = rho * 0.0;
  q.fs[5] = rho * 0.0;
  q.sh[0] = p.shock_rho;
  q.sh[1] = p.shock_rho * p.shock_u;
  q.sh[2] = p.shock_rho * p.shock_v;
  q.sh[3] = p.shock_p / (g - 1.0) + 0.5 * p.shock_rho * (p.shock_u * p.shock_u + p.shock_v * p.shock_v);
  q.sh[4] = p.shock_rho * 0.0;
  q.sh[5] = p.shock_rho * 0.0;
  if (p.turb) {  // frozen freestream carries k_inf, omega_inf
    q.fs[4] = rho * p.turb_k_inf;
    q.fs[5] = rho * p.turb_w_inf;
  }
  q.kmin = 1e-8 * p.turb_k_inf;
  q.wmin = 1e-3 * p.turb_w_inf;
  q.flux = p.flux;
  q.flux_alt = p.flux_alt;
  q.spatial = p.spatial;
  q.bcset = p.bcset;
  q.viscous = p.viscous;
  q.nvar = p.nvar;
  q.turb = p.turb;
  q.turb1 = p.turb && p.turb_first_order;
}

int check_params(const cfd_params* p) {
  if (!p) return CFD_ERR_ARG;
  if (p->nvar != 4 && p->nvar != 6) return CFD_ERR_ARG;
  if (p->flux < 0 || p->flux > 3 || p->flux_alt < 0 || p->flux_alt > 3) return CFD_ERR_ARG;
  if (p->spatial < 0 || p->spatial > 2 || p->bcset < 0 || p->bcset > 3) return CFD_ERR_ARG;
  if (p->turb && (p->nvar != 6 || !p->viscous || !(p->turb_k_inf > 0) || !(p->turb_w_inf > 0))) return CFD_ERR_ARG;
  if (p->turb_first_order != 0 && p->turb_first_order != 1) return CFD_ERR_ARG;
  return CFD_OK;
}

void free_krylov(cfd_ctx* c) {
  Krylov& k = c->kry;
  (void)hipFree(k.V);
  (void)hipFree(k.D);
  (void)hipFree(k.blk);
  (void)hipFree(k.aoff);
  (void)hipFree(k.fjac);
  (void)hipFree(k.bd);
  (void)hipFree(k.lineG);
  (void)hipFree(k.lineP);
  (void)hipFree(k.lineA);
  (void)hipFree(k.lineS);
  (void)hipFree(k.lineZ);
  (void)hipFree(k.lineW);
  (void)hipFree(k.xt);
  (void)hipFree(k.z);
  (void)hipFree(k.w);
  (void)hipFree(k.rr);
  (void)hipFree(k.w2);
  (void)hipFree(k.h);
  (void)hipFree(k.part);
  (void)hipFree(k.dc);
  (void)hipHostFree(k.hpin);
  (void)hipHostFree(k.hcol);
  for (hipEvent_t& e : k.col_ev)
    if (e) (void)hipEventDestroy(e);
  k = Krylov();
}

int ensure_krylov(cfd_ctx* c, int m, int prec) {
  Krylov& k = c->kry;
  if (k.V && k.m >= m && k.prec == prec) return CFD_OK;
  free_krylov(c);
  const size_t es = prec == 32 ? 4 : 8;
  const size_t vs = (size_t)c->P.nvar * c->pitch;
  HIP_TRY(c, hipMalloc(&k.V, (size_t)(m + 1) * vs * es));
  HIP_TRY(c, hipMalloc((void**)&k.D, (size_t)c->pitch * 8));
  HIP_TRY(c, hipMalloc(&k.blk, (size_t)c->P.nvar * c->P.nvar * c->pitch * es));
  HIP_TRY(c, hipMalloc((void**)&k.fjac, (size_t)2 * c->P.nvar * c->P.nvar * ((c->F + 63) / 64 * 64) * 8));
  HIP_TRY(c, hipMalloc((void**)&k.z, vs * 8));
  HIP_TRY(c, hipMemset(k.z, 0, vs * 8));  // padding entries stay zero (the line sweep writes cells only)
  HIP_TRY(c, hipMalloc((void**)&k.w, vs * 8));
  HIP_TRY(c, hipMalloc((void**)&k.rr, vs * 8));
  HIP_TRY(c, hipMalloc((void**)&k.w2, vs * 8));
  // the step's scalars: R (2 (j + 1) sums) at h[0..), y at h[yoff..); H
  // columns and the coefficients with leading dimension ld: 64, 128 or 256,
  // the first that holds m + 2 (the FP32 update kernel is compiled for each)
  k.ld = m + 2 <= 64 ? 64 : m + 2 <= 128 ? 128 : 256;
  k.yoff = std::max(128, 2 * (m + 2));
  const size_t ld = (size_t)k.ld;
  HIP_TRY(c, hipMalloc((void**)&k.h, ((size_t)k.yoff + m + 8) * 8));
  HIP_TRY(c, hipMalloc((void**)&k.part, (size_t)2 * (m + 2) * 1024 * 8));
  HIP_TRY(c, hipMalloc((void**)&k.dc, (ld * ld + 3 * ld + 8) * 8));
  HIP_TRY(c, hipHostMalloc((void**)&k.hpin, std::max<size_t>(256, m + 8) * 8, hipHostMallocDefault));
  HIP_TRY(c, hipHostMalloc((void**)&k.hcol, ld * std::max(64, m + 1) * 8, hipHostMallocDefault));
  for (hipEvent_t& e : k.col_ev) HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  k.m = m;
  k.prec = prec;
  return CFD_OK;
}

// coupling blocks and the solution staging vector of the block SGS preconditioner
int ensure_sgs(cfd_ctx* c) {
  Krylov& k = c->kry;
  if (k.aoff) return CFD_OK;
  const size_t es = k.prec == 32 ? 4 : 8;
  HIP_TRY(c, hipMalloc(&k.aoff, (size_t)4 * c->P.nvar * c->P.nvar * c->pitch * es));
  HIP_TRY(c, hipMalloc((void**)&k.xt, (size_t)c->P.nvar * c->pitch * 8));
  return CFD_OK;
}

// line preconditioner: the SGS buffers plus FP64 diagonal blocks and the line factors
int ensure_line(cfd_ctx* c) {
  int rc = ensure_sgs(c);
  if (rc) return rc;
  Krylov& k = c->kry;
  if (k.bd) return CFD_OK;
  const size_t es = k.prec == 32 ? 4 : 8;
  const size_t nv = c->P.nvar, nb = nv * nv * c->pitch, slots = std::max((size_t)c->nlb * 64, (size_t)1);
  HIP_TRY(c, hipMalloc((void**)&k.bd, nb * 8));
  HIP_TRY(c, hipMalloc(&k.lineG, nv * nv * slots * es));
  HIP_TRY(c, hipMalloc(&k.lineP, nv * nv * slots * es));
  HIP_TRY(c, hipMalloc(&k.lineA, 4 * nb * es));
  HIP_TRY(c, hipMalloc(&k.lineS, nv * c->pitch * 8));
  HIP_TRY(c, hipMalloc(&k.lineZ, nv * c->pitch * 8));
  HIP_TRY(c, hipMalloc((void**)&k.lineW, nv * slots * 8));
  return CFD_OK;
}

// upload a host AoS state vector (reference order) into SoA vector dst
int upload_state(cfd_ctx* c, const double* U6, double* dst) {
  HIP_TRY(c, hipMemcpyAsync(c->io6, U6, (size_t)c->N * 6 * 8, hipMemcpyHostToDevice, c->stream));
  launch_pack(c, c->io6, dst);
  return CFD_OK;
}
int download_state(cfd_ctx* c, const double* src, double* U6) {
  launch_unpack(c, src, c->io6);
  HIP_TRY(c, hipMemcpyAsync(U6, c->io6, (size_t)c->N * 6 * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return CFD_OK;
}
int sync(cfd_ctx* c) {
  HIP_TRY(c, hipGetLastError());
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  return CFD_OK;
}
// copy n device scalars scal[slot..] to host
int read_scal(cfd_ctx* c, int slot, int n, double* out) {
  HIP_TRY(c, hipMemcpyAsync(c->scal_host, c->scal + slot, n * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  std::memcpy(out, c->scal_host, n * 8);
  return CFD_OK;
}
// per-cell scalar (RCM order) -> host array in reference order
int download_cellscalar(cfd_ctx* c, const double* d, double* host) {
  std::vector<double> tmp(c->N);
  HIP_TRY(c, hipMemcpyAsync(tmp.data(), d, (size_t)c->N * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->N; ++i) host[c->iperm[i]] = tmp[i];
  return CFD_OK;
}

// Geometry of mesh m (reference order) into the context's device arrays (RCM /
// color order), and the wall distances computed on the device.
int upload_geometry(cfd_ctx* c, const cfd_mesh* m) {
  const int N = c->N, F = c->F, P_ = c->pitch;
  std::vector<double> vol(P_, 1.0), xc(P_, 0.0), yc(P_, 0.0);
  for (int i = 0; i < N; ++i) {
    const int o = m->iperm[i];
    vol[i] = m->vol[o];
    xc[i] = m->xc[o];
    yc[i] = m->yc[o];
  }
  std::vector<double> nx(F), ny(F), area(F), xm(F), ym(F), dnw(F);
  std::vector<double> wxy;  // no-slip face midpoints (Solver.cpp:1986-1991)
  for (int d = 0; d < F; ++d) {
    const int f = c->face_order[d];
    nx[d] = m->nx[f];
    ny[d] = m->ny[f];
    area[d] = m->area[f];
    xm[d] = m->xm[f];
    ym[d] = m->ym[f];
    const int o = m->own[f];
    dnw[d] = std::abs((m->xm[f] - m->xc[o]) * m->nx[f] + (m->ym[f] - m->yc[o]) * m->ny[f]);
  }
  for (int f = 0; f < F; ++f)
    if (m->nei[f] < 0 && m->tag[f] == 4) {
      wxy.push_back(m->xm[f]);
      wxy.push_back(m->ym[f]);
    }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipMemcpy(c->d_vol, vol.data(), P_ * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_xc, xc.data(), P_ * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_yc, yc.data(), P_ * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_nx, nx.data(), F * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_ny, ny.data(), F * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_area, area.data(), F * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_xm, xm.data(), F * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_ym, ym.data(), F * 8, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_dnw, dnw.data(), F * 8, hipMemcpyHostToDevice));
  double* dw = nullptr;
  HIP_TRY(c, hipMalloc((void**)&dw, std::max<size_t>(wxy.size(), 2) * 8));
  if (!wxy.empty()) HIP_TRY(c, hipMemcpy(dw, wxy.data(), wxy.size() * 8, hipMemcpyHostToDevice));
  launch_wall_distance(c, dw, (int)wxy.size() / 2, c->d_wd);
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  HIP_TRY(c, hipFree(dw));
  c->dt_valid = false;
  return CFD_OK;
}

// Device face order for the context's assembly mode: edge-color-major (then
// owner position) for the colored scatter, owner position for the face
// buffer, so that in buffer mode the faces of a cell are processed together
// and its gathered data is reused from cache.  corder lists the device faces
// color by color for the kernels that need color ranges in either mode.
int layout_faces(cfd_ctx* c) {
  const cfd_mesh* m = c->mesh;
  const int N = c->N, F = c->F, P_ = c->pitch;
  if (c->assembly == 0) {
    c->face_order = m->face_order;
  } else {
    c->face_order.resize(F);
    for (int f = 0; f < F; ++f) c->face_order[f] = f;
    std::stable_sort(c->face_order.begin(), c->face_order.end(),
                     [&](int a, int b) { return m->perm[m->own[a]] < m->perm[m->own[b]]; });
  }
  std::vector<int> fpos(F);  // reference face -> device face
  for (int d = 0; d < F; ++d) fpos[c->face_order[d]] = d;
  // within each color the interior faces come first (the order inside a
  // color does not change any result), so that kernels for interior faces
  // only take a color's leading run [color_ptr, color_bptr)
  std::vector<int> corder(F);
  c->color_bptr.assign(c->color_ptr.size() > 0 ? c->color_ptr.size() - 1 : 0, 0);
  for (size_t col = 0; col + 1 < c->color_ptr.size(); ++col) {
    int k = c->color_ptr[col];
    for (int pass = 0; pass < 2; ++pass) {
      if (pass == 1) c->color_bptr[col] = k;
      for (int i = c->color_ptr[col]; i < c->color_ptr[col + 1]; ++i)
        if ((m->nei[m->face_order[i]] >= 0) == (pass == 0)) corder[k++] = fpos[m->face_order[i]];
    }
  }
  std::vector<int> own(F), nei(F), tag(F);
  for (int d = 0; d < F; ++d) {
    const int f = c->face_order[d];
    own[d] = m->perm[m->own[f]];
    nei[d] = m->nei[f] >= 0 ? m->perm[m->nei[f]] : -1;
    tag[d] = m->tag[f];
  }
  // the boundary faces grouped by owner, each group in corder order (ascending color),
  // so that one thread per group can add them in the order the per-color launches did
  std::vector<int4> grp;
  {
    std::vector<int> gidx(N, -1);
    for (size_t col = 0; col + 1 < c->color_ptr.size(); ++col)
      for (int i = c->color_bptr[col]; i < c->color_ptr[col + 1]; ++i) {
        const int d = corder[i], o = own[d];
        if (gidx[o] < 0) {
          gidx[o] = (int)grp.size();
          grp.push_back(make_int4(-1, -1, -1, -1));
        }
        int* g = &grp[gidx[o]].x;
        int k = 0;
        while (k < 4 && g[k] >= 0) ++k;
        if (k == 4) return fail(c, CFD_ERR_MESH, "cell with more than 4 boundary faces");
        g[k] = d;
      }
  }
  // per-cell face slots in ascending reference face index
  // and, slot by slot, the cell across each face (so a kernel reaches the
  // neighbour's data without first loading the face's owner and neighbour)
  std::vector<int4> slots(P_, make_int4(-1, -1, -1, -1)), cnb(P_, make_int4(-1, -1, -1, -1));
  std::vector<int> cnt(N, 0);
  for (int f = 0; f < F; ++f) {
    const int d = fpos[f];
    const int co = m->perm[m->own[f]];
    const int cn = m->nei[f] >= 0 ? m->perm[m->nei[f]] : -1;
    if (cnt[co] >= 4) return fail(c, CFD_ERR_MESH, "cell with more than 4 faces");
    (&cnb[co].x)[cnt[co]] = cn;
    (&slots[co].x)[cnt[co]++] = d << 1;
    if (cn >= 0) {
      if (cnt[cn] >= 4) return fail(c, CFD_ERR_MESH, "cell with more than 4 faces");
      (&cnb[cn].x)[cnt[cn]] = co;
      (&slots[cn].x)[cnt[cn]++] = (d << 1) | 1;
    }
  }
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (!c->d_bgrp) HIP_TRY(c, dalloc(&c->d_bgrp, std::max<size_t>(grp.size(), 1)));
  c->nbgrp = (int)grp.size();
  if (!grp.empty())
    HIP_TRY(c, hipMemcpy(c->d_bgrp, grp.data(), grp.size() * sizeof(int4), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_slots, slots.data(), P_ * sizeof(int4), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_cnb, cnb.data(), P_ * sizeof(int4), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_own, own.data(), F * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_nei, nei.data(), F * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_tag, tag.data(), F * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_facemap, c->face_order.data(), F * 4, hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(c->d_corder, corder.data(), F * 4, hipMemcpyHostToDevice));
  if (!c->alt_ref.empty()) {
    std::vector<unsigned char> al(F);
    for (int d = 0; d < F; ++d) al[d] = c->alt_ref[c->face_order[d]];
    HIP_TRY(c, hipMemcpy(c->d_alt, al.data(), F, hipMemcpyHostToDevice));
  }
  return upload_geometry(c, m);
}

}  // namespace

// ============================================================================
extern "C" {

int cfd_params_default(cfd_params* p) {
  if (!p) return CFD_ERR_ARG;
  std::memset(p, 0, sizeof *p);
  p->gamma = 1.4;
  p->R_gas = 287.0;
  p->nu = 0.01;
  p->Pr = 0.72;
  p->flux = CFD_FLUX_HLLC;
  p->flux_alt = CFD_FLUX_LLF;
  p->spatial = CFD_SECOND_ORDER_LIMITED;
  p->bcset = CFD_BC_BUMP;
  p->viscous = 0;
  p->nvar = 4;
  return CFD_OK;
}

int cfd_params_flow_conditions(cfd_params* p, double M, double P, double T) {
  if (!p) return CFD_ERR_ARG;
  const double g = p->gamma;
  p->M_inf = M;
  const double fT = 1.0 + 0.5 * (g - 1.0) * M * M;
  const double fP = std::pow(fT, g / (g - 1.0));
  p->T_total_inlet = T * fT;
  p->P_total_inlet = P * fP;
  p->P_static_exit = P;
  return CFD_OK;
}

int cfd_params_freestream_static(cfd_params* p, double M, double P, double T) {
  return cfd_params_flow_conditions(p, M, P, T);  // Solver.cpp:2017-2025 computes the same fields
}

int cfd_params_kinematic_viscosity(cfd_params* p, double Re, double L) {
  if (!p || !(Re > 0)) return CFD_ERR_ARG;
  const double g = p->gamma;
  const double fT = 1.0 + 0.5 * (g - 1.0) * p->M_inf * p->M_inf;
  const double Ts = p->T_total_inlet / fT;
  const double c = std::sqrt(g * p->R_gas * Ts);
  const double u = p->M_inf * c;
  p->nu = (u * L) / Re;
  return CFD_OK;
}

int cfd_params_initial_state(const cfd_params* p, double* U6) {
  if (!p || !U6) return CFD_ERR_ARG;
  const double g = p->gamma, R = p->R_gas;
  const double pr = p->P_static_exit;
  const double fT = std::pow(p->P_total_inlet / pr, (g - 1.0) / g);
  const double T = p->T_total_inlet / fT;
  const double rho = pr / (R * T);
  const double V = std::sqrt(2.0 * (g * R / (g - 1.0)) * (p->T_total_inlet - T));
  const double u = V, v = 0.0;
  U6[0] = rho;
  U6[1] = rho * u;
  U6[2] = rho * v;
  U6[3] = pr / (g - 1.0) + 0.5 * rho * (u * u + v * v);
  U6[4] = rho * 0.0;
  U6[5] = rho * 0.0;
  return CFD_OK;
}

int cfd_gmres_opts_default(cfd_gmres_opts* o) {
  if (!o) return CFD_ERR_ARG;
  std::memset(o, 0, sizeof *o);
  o->m = 40;
  o->restarts = 0;
  o->tol = 1e-2;
  o->jv = 0;
  o->precision = 64;
  o->precond = 0;
  return CFD_OK;
}

int cfd_newton_opts_default(cfd_newton_opts* o) {
  if (!o) return CFD_ERR_ARG;
  std::memset(o, 0, sizeof *o);
  o->tolerance = 1e-6;
  o->rel_drop = 0.0;
  o->max_steps = 50;
  o->log = 0;
  o->cfl0 = 0.1;
  cfd_gmres_opts_default(&o->gmres);
  return CFD_OK;
}

int cfd_ctx_create(const cfd_mesh* m, const cfd_params* p, cfd_ctx** out) {
  if (!m || !out || check_params(p)) return CFD_ERR_ARG;
  cfd_ctx* c = new cfd_ctx();
  *out = c;
  c->P = *p;
  fill_phys(c);
  HIP_TRY(c, hipGetDevice(&c->device));
  if (int rc = noalloc_init(c)) return rc;
  HIP_TRY(c, hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
  c->stream = c->own_stream;
  const int N = m->ncells(), F = m->nfaces();
  c->N = N;
  c->F = F;
  c->pitch = (N + 63) / 64 * 64;
  c->ncolors = m->ncolors;
  c->color_ptr = m->color_ptr;
  c->face_order = m->face_order;
  c->perm = m->perm;
  c->iperm = m->iperm;
  const int P_ = c->pitch;
  c->mesh = m;
  HIP_TRY(c, dalloc(&c->d_vol, P_));
  HIP_TRY(c, dalloc(&c->d_xc, P_));
  HIP_TRY(c, dalloc(&c->d_yc, P_));
  HIP_TRY(c, dalloc(&c->d_slots, P_));
  HIP_TRY(c, dalloc(&c->d_cnb, P_));
  HIP_TRY(c, dalloc(&c->d_own, F));
  HIP_TRY(c, dalloc(&c->d_nei, F));
  HIP_TRY(c, dalloc(&c->d_tag, F));
  HIP_TRY(c, dalloc(&c->d_nx, F));
  HIP_TRY(c, dalloc(&c->d_ny, F));
  HIP_TRY(c, dalloc(&c->d_area, F));
  HIP_TRY(c, dalloc(&c->d_xm, F));
  HIP_TRY(c, dalloc(&c->d_ym, F));
  HIP_TRY(c, dalloc(&c->d_iperm, N));
  HIP_TRY(c, dalloc(&c->d_facemap, F));
  HIP_TRY(c, dalloc(&c->d_corder, F));
  HIP_TRY(c, dalloc(&c->d_wd, P_));
  HIP_TRY(c, dalloc(&c->d_dnw, F));
  HIP_TRY(c, hipMemcpy(c->d_iperm, m->iperm.data(), N * 4, hipMemcpyHostToDevice));
  DMesh& d = c->dm;
  d.N = N;
  d.F = F;
  d.pitch = P_;
  d.vol = c->d_vol;
  d.xc = c->d_xc;
  d.yc = c->d_yc;
  d.slots = c->d_slots;
  d.cnb = c->d_cnb;
  d.own = c->d_own;
  d.nei = c->d_nei;
  d.tag = c->d_tag;
  d.nx = c->d_nx;
  d.ny = c->d_ny;
  d.area = c->d_area;
  d.xm = c->d_xm;
  d.ym = c->d_ym;
  d.alt = nullptr;
  d.wd = c->d_wd;
  d.dnw = c->d_dnw;
  d.corder = c->d_corder;
  c->assembly = 1;
  {
    int rc = layout_faces(c);
    if (rc) return rc;
  }
  {  // cell coloring and the neighbor/back-slot table of the block SGS sweeps
    std::vector<int4> nq(P_, make_int4(-1, -1, -1, -1));
    std::vector<int> fl((size_t)4 * N, -1), cnt(N, 0);
    for (int f = 0; f < F; ++f) {
      const int co = m->perm[m->own[f]];
      fl[(size_t)4 * co + cnt[co]++] = f;
      if (m->nei[f] >= 0) {
        const int cn = m->perm[m->nei[f]];
        fl[(size_t)4 * cn + cnt[cn]++] = f;
      }
    }
    for (int i = 0; i < N; ++i)
      for (int q = 0; q < cnt[i]; ++q) {
        const int f = fl[(size_t)4 * i + q];
        if (m->nei[f] < 0) continue;
        const int co = m->perm[m->own[f]], cn = m->perm[m->nei[f]];
        const int o = co == i ? cn : co;
        int r = 0;
        while (fl[(size_t)4 * o + r] != f) ++r;
        (&nq[i].x)[q] = (o << 2) | r;
      }
    if (m->ccell_ptr.size() > 256) return fail(c, CFD_ERR_MESH, "more than 255 cell colors");
    std::vector<unsigned char> col(P_, 0);
    for (int i = 0; i < N; ++i) col[i] = (unsigned char)m->cell_color[i];
    c->ccell_ptr = m->ccell_ptr;
    HIP_TRY(c, dalloc(&c->d_ccol, P_));
    HIP_TRY(c, dalloc(&c->d_ccells, std::max(N, 1)));
    HIP_TRY(c, dalloc(&c->d_nq, P_));
    std::vector<int> cpos(P_, 0);
    for (int t = 0; t < N; ++t) cpos[m->ccells[t]] = t;
    HIP_TRY(c, dalloc(&c->d_cpos, P_));
    HIP_TRY(c, hipMemcpy(c->d_cpos, cpos.data(), (size_t)P_ * 4, hipMemcpyHostToDevice));
    // line segments, checked before any kernel indexes with them: a permutation
    // of the cells, at most 64 per segment, and the previous / next slots of each
    // cell lead to its chain neighbours
    {
      std::vector<char> seen(N, 0);
      if ((int)m->seg_cells.size() != N || m->seg_ptr.empty() || m->seg_ptr.back() != N)
        return fail(c, CFD_ERR_MESH, "line segments do not cover the cells");
      for (size_t k = 0; k + 1 < m->seg_ptr.size(); ++k) {
        const int b = m->seg_ptr[k], e = m->seg_ptr[k + 1];
        if (e <= b || e - b > 64) return fail(c, CFD_ERR_MESH, "bad line segment length");
        for (int t = b; t < e; ++t) {
          const int cc = m->seg_cells[t];
          if (cc < 0 || cc >= N || seen[cc]) return fail(c, CFD_ERR_MESH, "line segments are not a permutation");
          seen[cc] = 1;
          const int qp = m->lslot_prev[cc], qn = m->lslot_next[cc];
          const bool ok_p = t == b ? qp == -1 : (qp >= 0 && qp < 4 && (&nq[cc].x)[qp] >= 0 &&
                                                  ((&nq[cc].x)[qp] >> 2) == m->seg_cells[t - 1]);
          const bool ok_n = t == e - 1 ? qn == -1 : (qn >= 0 && qn < 4 && (&nq[cc].x)[qn] >= 0 &&
                                                      ((&nq[cc].x)[qn] >> 2) == m->seg_cells[t + 1]);
          if (!ok_p || !ok_n) return fail(c, CFD_ERR_MESH, "line segment links are inconsistent");
        }
      }
      // the sweeps read off-segment neighbours while the segment's own cells
      // are rewritten: no face may join two non-consecutive cells of a segment
      std::vector<int> sof(N), tpos(N);
      for (size_t k = 0; k + 1 < m->seg_ptr.size(); ++k)
        for (int t = m->seg_ptr[k]; t < m->seg_ptr[k + 1]; ++t) {
          sof[m->seg_cells[t]] = (int)k;
          tpos[m->seg_cells[t]] = t;
        }
      for (int i = 0; i < N; ++i)
        for (int q = 0; q < 4; ++q) {
          const int en = (&nq[i].x)[q];
          if (en < 0) continue;
          const int o = en >> 2;
          if (sof[o] == sof[i] && std::abs(tpos[o] - tpos[i]) != 1)
            return fail(c, CFD_ERR_MESH, "line segment touches itself");
        }
    }
    c->seg_ptr = m->seg_ptr;
    c->segcol_ptr = m->segcol_ptr;
    std::vector<int> lpos(P_, 0);
    for (int t = 0; t < N; ++t) lpos[m->seg_cells[t]] = t;
    std::vector<char2> ls(P_, make_char2(-1, -1));
    for (int i = 0; i < N; ++i) ls[i] = make_char2((signed char)m->lslot_prev[i], (signed char)m->lslot_next[i]);
    HIP_TRY(c, dalloc(&c->d_seg_ptr, m->seg_ptr.size()));
    HIP_TRY(c, dalloc(&c->d_seg_cells, std::max(N, 1)));
    HIP_TRY(c, dalloc(&c->d_lpos, P_));
    HIP_TRY(c, dalloc(&c->d_lslot, P_));
    HIP_TRY(c, hipMemcpy(c->d_seg_ptr, m->seg_ptr.data(), m->seg_ptr.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_seg_cells, m->seg_cells.data(), (size_t)N * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_lpos, lpos.data(), (size_t)P_ * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_lslot, ls.data(), (size_t)P_ * sizeof(char2), hipMemcpyHostToDevice));
    // bundles: per color, the multi-cell segments longest first, each bundle
    // filled while the next segment fits in its 64 slots
    {
      const int ncol = (int)m->segcol_ptr.size() - 1;
      std::vector<int> lb_t, lb_kl, lcomp(P_, -1), segs;
      c->lbcol_ptr.assign(1, 0);
      c->lmcol_cells.assign(ncol, 0);
      for (int col = 0; col < ncol; ++col) {
        segs.clear();
        for (int k = m->segcol_ptr[col]; k < m->segcol_ptr[col + 1]; ++k)
          if (m->seg_ptr[k + 1] - m->seg_ptr[k] > 1) segs.push_back(k);
        std::stable_sort(segs.begin(), segs.end(), [&](int a, int b) {
          return m->seg_ptr[a + 1] - m->seg_ptr[a] > m->seg_ptr[b + 1] - m->seg_ptr[b];
        });
        int fill = 64;
        for (int k : segs) {
          const int b = m->seg_ptr[k], L = m->seg_ptr[k + 1] - b;
          if (fill + L > 64) {
            lb_t.insert(lb_t.end(), 64, -1);
            lb_kl.insert(lb_kl.end(), 64, 0);
            fill = 0;
          }
          if (b != m->seg_ptr[m->segcol_ptr[col]] + c->lmcol_cells[col])
            return fail(c, CFD_ERR_MESH, "a color's multi-cell segments do not lead its line positions");
          c->lmcol_cells[col] += L;
          const size_t base = lb_t.size() - 64 + fill;
          for (int i = 0; i < L; ++i) {
            lb_t[base + i] = b + i;
            lb_kl[base + i] = i | L << 8;
            lcomp[b + i] = (int)(base + i);
          }
          fill += L;
        }
        c->lbcol_ptr.push_back((int)(lb_t.size() / 64));
      }
      c->nlb = (int)(lb_t.size() / 64);
      if (const char* e = std::getenv("CFD_LINE_FUSED_MAX_CELLS")) c->line_fused_max = std::atoi(e);
      if (ncol > 255) return fail(c, CFD_ERR_MESH, "more than 255 segment colors");
      // per cell: off-segment neighbour slots of a lower (bits 0-3) and of a
      // higher (bits 4-7) segment color
      std::vector<int> lcol(N, 0);
      for (int col = 0; col < ncol; ++col)
        for (int k = m->segcol_ptr[col]; k < m->segcol_ptr[col + 1]; ++k)
          for (int t = m->seg_ptr[k]; t < m->seg_ptr[k + 1]; ++t) lcol[m->seg_cells[t]] = col;
      // all indexed by line position t.  The off-segment neighbours go to
      // packed slots: those of lower colors to slots 0, 1, ... and those of
      // higher colors to slots 3, 2, ..., each group in ascending cell-slot
      // order (the order the sweep subtracts them in).  A sweep pass then
      // reads the same packed slot in every lane of a wavefront, where the
      // cell slots of the couplings it needs differ from cell to cell.
      std::vector<unsigned char> lmask(P_, 0);
      std::vector<int4> lnc(P_, make_int4(-1, -1, -1, -1));
      std::vector<char4> lqm(P_, make_char4(-1, -1, -1, -1));
      for (int t = 0; t < N; ++t) {
        const int i = m->seg_cells[t];
        int lo = 0, hi = 3;
        for (int q = 0; q < 4; ++q) {
          const int en = (&nq[i].x)[q];
          if (en < 0 || q == m->lslot_prev[i] || q == m->lslot_next[i]) continue;
          const int o = en >> 2;
          if (lcol[o] == lcol[i]) return fail(c, CFD_ERR_MESH, "neighbouring segments share a color");
          const bool lower = lcol[o] < lcol[i];
          lmask[t] |= (unsigned char)(lower ? 1u << q : 16u << q);
          const int r = lower ? lo++ : hi--;
          (&lnc[t].x)[r] = lpos[o];
          (&lqm[t].x)[r] = (signed char)q;
        }
      }
      HIP_TRY(c, dalloc(&c->d_lmask, P_));
      HIP_TRY(c, dalloc(&c->d_lnc, P_));
      HIP_TRY(c, dalloc(&c->d_lqm, P_));
      HIP_TRY(c, hipMemcpy(c->d_lmask, lmask.data(), P_, hipMemcpyHostToDevice));
      HIP_TRY(c, hipMemcpy(c->d_lnc, lnc.data(), (size_t)P_ * sizeof(int4), hipMemcpyHostToDevice));
      HIP_TRY(c, hipMemcpy(c->d_lqm, lqm.data(), (size_t)P_ * sizeof(char4), hipMemcpyHostToDevice));
      const size_t ns = std::max(lb_t.size(), (size_t)1);
      HIP_TRY(c, dalloc(&c->d_lb_t, ns));
      HIP_TRY(c, dalloc(&c->d_lb_kl, ns));
      HIP_TRY(c, dalloc(&c->d_lcomp, P_));
      if (!lb_t.empty()) {
        HIP_TRY(c, hipMemcpy(c->d_lb_t, lb_t.data(), lb_t.size() * 4, hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(c->d_lb_kl, lb_kl.data(), lb_kl.size() * 4, hipMemcpyHostToDevice));
      }
      HIP_TRY(c, hipMemcpy(c->d_lcomp, lcomp.data(), (size_t)P_ * 4, hipMemcpyHostToDevice));
    }
    HIP_TRY(c, hipMemcpy(c->d_ccol, col.data(), P_, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_ccells, m->ccells.data(), (size_t)N * 4, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(c->d_nq, nq.data(), (size_t)P_ * sizeof(int4), hipMemcpyHostToDevice));
  }
  // solver buffers
  HIP_TRY(c, dalloc(&c->grad, (size_t)14 * P_));
  HIP_TRY(c, dalloc(&c->phi, P_));
  HIP_TRY(c, dalloc(&c->dt, P_));
  HIP_TRY(c, dalloc(&c->Fm, (size_t)6 * F));
  HIP_TRY(c, dalloc(&c->red, 65536));
  HIP_TRY(c, dalloc(&c->scal, 64));
  HIP_TRY(c, dalloc(&c->gradT, (size_t)28 * P_));
  HIP_TRY(c, dalloc(&c->adj, (size_t)16 * P_));
  HIP_TRY(c, dalloc(&c->tq, (size_t)8 * P_));
  HIP_TRY(c, hipHostMalloc((void**)&c->scal_host, 64 * 8, hipHostMallocDefault));
  HIP_TRY(c, dalloc(&c->io6, (size_t)6 * N));
  for (int i = 0; i < 6; ++i) {
    HIP_TRY(c, dalloc(&c->vec[i], (size_t)6 * P_));
    HIP_TRY(c, hipMemset(c->vec[i], 0, (size_t)6 * P_ * 8));
  }
  HIP_TRY(c, hipMemset(c->phi, 0, (size_t)P_ * 8));
  HIP_TRY(c, hipDeviceSynchronize());
  return CFD_OK;
}

int cfd_ctx_update_geometry(cfd_ctx* c, const cfd_mesh* m) {
  if (!c || !m || m->ncells() != c->N || m->nfaces() != c->F) return CFD_ERR_ARG;
  if (m != c->mesh) return fail(c, CFD_ERR_ARG, "geometry must come from the context's own mesh");
  c->phi_valid = false;  // the limiter depends on the geometry
  return upload_geometry(c, m);
}

int cfd_ctx_destroy(cfd_ctx* c) {
  if (!c) return CFD_OK;
  (void)hipStreamSynchronize(c->stream);
  free_krylov(c);
  void* bufs[] = {c->d_vol, c->d_xc,  c->d_yc,  c->d_slots, c->d_cnb, c->d_own, c->d_nei, c->d_tag,  c->d_nx,
                  c->d_ny,  c->d_area, c->d_xm, c->d_ym,    c->d_iperm, c->d_facemap, c->d_alt, c->grad, c->d_wd, c->d_dnw, c->d_corder, c->d_bgrp, c->d_ccol, c->d_ccells, c->d_nq, c->d_cpos, c->d_seg_ptr, c->d_seg_cells, c->d_lpos, c->d_lslot, c->d_lb_t, c->d_lb_kl, c->d_lcomp, c->d_lmask, c->d_lnc, c->d_lqm,
                  c->phi,   c->dt,    c->Fm,    c->red,     c->scal, c->gradT, c->adj, c->adjF, c->tq, c->io6};
  for (void* b : bufs) (void)hipFree(b);
  for (int i = 0; i < 6; ++i) (void)hipFree(c->vec[i]);
  (void)hipFree(c->aux);
  (void)hipHostFree(c->scal_host);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return CFD_OK;
}

const char* cfd_last_error(const cfd_ctx* c) { return c ? c->err.c_str() : "null context"; }

int cfd_ctx_set_params(cfd_ctx* c, const cfd_params* p) {
  if (!c || check_params(p)) return CFD_ERR_ARG;
  if (p->nvar != c->P.nvar) return fail(c, CFD_ERR_ARG, "nvar cannot change on a live context");
  // the stored limiter belongs to the scheme and physics it was formed with
  if (p->turb != c->P.turb || p->spatial != c->P.spatial || p->bcset != c->P.bcset) c->phi_valid = false;
  c->P = *p;
  fill_phys(c);
  return CFD_OK;
}
int cfd_ctx_get_params(const cfd_ctx* c, cfd_params* p) {
  if (!c || !p) return CFD_ERR_ARG;
  *p = c->P;
  return CFD_OK;
}
int cfd_pitch(const cfd_ctx* c) { return c ? c->pitch : -1; }
int cfd_ncells(const cfd_ctx* c) { return c ? c->N : -1; }
void* cfd_stream(const cfd_ctx* c) { return c ? (void*)c->stream : nullptr; }
void* cfd_ctx_own_stream(const cfd_ctx* c) { return c ? (void*)c->own_stream : nullptr; }

int cfd_ctx_set_stream(cfd_ctx* c, void* stream) {
  if (!c) return CFD_ERR_ARG;
  hipStream_t s = (hipStream_t)stream;
  if (s == c->stream) return CFD_OK;
  // work already queued on the old stream finishes before the new one is used
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (s) {
    int dev = -1;
    HIP_TRY(c, hipStreamGetDevice(s, &dev));
    if (dev != c->device) return fail(c, CFD_ERR_ARG, "stream belongs to another device than the context");
  }
  c->stream = s;
  return CFD_OK;
}

int cfd_set_face_flux(cfd_ctx* c, const unsigned char* use_alt) {
  if (!c) return CFD_ERR_ARG;
  if (!use_alt) {
    c->dm.alt = nullptr;
    c->alt_ref.clear();
    return CFD_OK;
  }
  c->alt_ref.assign(use_alt, use_alt + c->F);
  for (auto& v : c->alt_ref) v = v ? 1 : 0;
  std::vector<unsigned char> a(c->F);
  for (int d = 0; d < c->F; ++d) a[d] = c->alt_ref[c->face_order[d]];
  if (!c->d_alt) HIP_TRY(c, dalloc(&c->d_alt, c->F));
  HIP_TRY(c, hipMemcpy(c->d_alt, a.data(), c->F, hipMemcpyHostToDevice));
  c->dm.alt = c->d_alt;
  return CFD_OK;
}

int cfd_set_assembly(cfd_ctx* c, int mode) {
  if (!c || (mode != 0 && mode != 1)) return CFD_ERR_ARG;
  if (mode == c->assembly) return CFD_OK;
  c->assembly = mode;
  return layout_faces(c);
}

// ---- host-pointer drop-ins -----------------------------------------------------
int cfd_residual(cfd_ctx* c, const double* U6, double* R6, int update_limiters) {
  if (!c || !U6 || !R6) return CFD_ERR_ARG;
  int rc = upload_state(c, U6, c->vec[0]);
  if (rc) return rc;
  launch_residual(c, c->vec[0], c->vec[1], update_limiters != 0);
  return download_state(c, c->vec[1], R6);
}

int cfd_get_gradients(cfd_ctx* c, double* g10) {
  if (!c || !g10) return CFD_ERR_ARG;
  std::vector<double> tmp((size_t)10 * c->pitch);
  HIP_TRY(c, hipMemcpyAsync(tmp.data(), c->grad, tmp.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->N; ++i)
    for (int q = 0; q < 10; ++q) g10[(size_t)c->iperm[i] * 10 + q] = tmp[(size_t)q * c->pitch + i];
  return CFD_OK;
}

int cfd_get_limiters(cfd_ctx* c, double* phi) {
  if (!c || !phi) return CFD_ERR_ARG;
  return download_cellscalar(c, c->phi, phi);
}

int cfd_set_limiters(cfd_ctx* c, const double* phi) {
  if (!c || !phi) return CFD_ERR_ARG;
  std::vector<double> tmp(c->pitch, 1.0);
  for (int i = 0; i < c->N; ++i) tmp[i] = phi[c->iperm[i]];
  HIP_TRY(c, hipMemcpy(c->phi, tmp.data(), (size_t)c->pitch * 8, hipMemcpyHostToDevice));
  c->phi_valid = true;
  return CFD_OK;
}

int cfd_local_timestep(cfd_ctx* c, const double* U6, double cfl) {
  if (!c || !U6) return CFD_ERR_ARG;
  int rc = upload_state(c, U6, c->vec[0]);
  if (rc) return rc;
  launch_timestep(c, c->vec[0], cfl);
  return sync(c);
}

int cfd_get_dt(cfd_ctx* c, double* dt) {
  if (!c || !dt) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  return download_cellscalar(c, c->dt, dt);
}

int cfd_diag_preconditioner(cfd_ctx* c, const double* U6, double* D) {
  if (!c || !U6 || !D) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  int rc = upload_state(c, U6, c->vec[0]);
  if (rc) return rc;
  rc = ensure_krylov(c, 1, 64);
  if (rc) return rc;
  launch_diag_precond(c, c->vec[0], c->kry.D);
  return download_cellscalar(c, c->kry.D, D);
}

int cfd_jv_fd(cfd_ctx* c, const double* U6, const double* R6, const double* v6, double* Jv6) {
  if (!c || !U6 || !R6 || !v6 || !Jv6) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  int rc;
  if ((rc = upload_state(c, U6, c->vec[0]))) return rc;
  if ((rc = upload_state(c, R6, c->vec[1]))) return rc;
  if ((rc = upload_state(c, v6, c->vec[2]))) return rc;
  launch_reduce(c, RED_L2, c->vec[0], nullptr, S_UNORM);
  launch_reduce(c, RED_L2, c->vec[2], nullptr, S_VNORM);
  launch_jv_fd(c, c->vec[0], c->vec[1], c->vec[2], c->vec[3], S_UNORM, S_VNORM, c->vec[4], c->vec[5]);
  return download_state(c, c->vec[3], Jv6);
}

int cfd_jv_tangent(cfd_ctx* c, const double* U6, const double* v6, double* Jv6, int precision) {
  if (!c || !U6 || !v6 || !Jv6 || (precision != 64 && precision != 32)) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  if (int rc = need_limiter(c)) return rc;
  int rc;
  if ((rc = upload_state(c, U6, c->vec[0]))) return rc;
  if ((rc = upload_state(c, v6, c->vec[1]))) return rc;
  const size_t n = (size_t)c->P.nvar * c->pitch;
  if (precision == 64) {
    launch_jv_tangent(c, c->vec[0], c->vec[1], c->vec[2], 64);
  } else {
    float* vf = (float*)c->vec[3];
    float* jf = (float*)c->vec[4];
    launch_d2f(c, c->vec[1], vf, n);
    launch_jv_tangent(c, c->vec[0], vf, jf, 32);
    launch_f2d(c, jf, c->vec[2], n);
  }
  return download_state(c, c->vec[2], Jv6);
}

int cfd_vjp_adjoint(cfd_ctx* c, const double* U6, const double* w6, double* JTw6) {
  if (!c || !U6 || !w6 || !JTw6) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  if (int rc = need_limiter(c)) return rc;
  if (int rc = ensure_adjoint(c)) return rc;
  int rc;
  if ((rc = upload_state(c, U6, c->vec[0]))) return rc;
  if ((rc = upload_state(c, w6, c->vec[1]))) return rc;
  launch_vjp(c, c->vec[0], c->vec[1], c->vec[2]);
  return download_state(c, c->vec[2], JTw6);
}

int cfd_block_jacobi(cfd_ctx* c, const double* U6, int precision, double* B, double* Binv) {
  if (!c || !U6 || (precision != 64 && precision != 32)) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  int rc;
  if ((rc = upload_state(c, U6, c->vec[0]))) return rc;
  if ((rc = ensure_krylov(c, c->kry.m > 0 ? c->kry.m : 1, precision))) return rc;
  const int nv = c->P.nvar;
  const size_t nb = (size_t)nv * nv * c->pitch;
  double* Bd = nullptr;
  HIP_TRY(c, hipMalloc((void**)&Bd, nb * 8));
  launch_block_precond(c, c->vec[0], c->kry.blk, precision, Bd);
  std::vector<double> hB(nb), hI(nb);
  std::vector<float> hIf;
  hipError_t e = hipMemcpyAsync(hB.data(), Bd, nb * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) {
    if (precision == 64) {
      e = hipMemcpyAsync(hI.data(), c->kry.blk, nb * 8, hipMemcpyDeviceToHost, c->stream);
    } else {
      hIf.resize(nb);
      e = hipMemcpyAsync(hIf.data(), c->kry.blk, nb * 4, hipMemcpyDeviceToHost, c->stream);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(Bd);
  if (e != hipSuccess) return fail(c, CFD_ERR_HIP, hipGetErrorString(e));
  if (precision == 32)
    for (size_t i = 0; i < nb; ++i) hI[i] = hIf[i];
  for (int i = 0; i < c->N; ++i) {
    const size_t o = (size_t)c->iperm[i] * nv * nv;
    for (int q = 0; q < nv * nv; ++q) {
      if (B) B[o + q] = hB[(size_t)q * c->pitch + i];
      if (Binv) Binv[o + q] = hI[(size_t)q * c->pitch + i];
    }
  }
  return CFD_OK;
}

int cfd_block_sgs(cfd_ctx* c, const double* U6, int precision, int transpose, const double* x6, double* z6,
                  double* Aoff, int* nbr, int* color) {
  if (!c || !U6 || (precision != 64 && precision != 32) || (!x6) != (!z6)) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  int rc;
  if ((rc = upload_state(c, U6, c->vec[0]))) return rc;
  if ((rc = ensure_krylov(c, c->kry.m > 0 ? c->kry.m : 1, precision))) return rc;
  if ((rc = ensure_sgs(c))) return rc;
  const int nv = c->P.nvar, P_ = c->pitch;
  launch_block_precond(c, c->vec[0], c->kry.blk, precision, nullptr, c->kry.aoff, c->d_cpos);
  if (x6) {
    if ((rc = upload_state(c, x6, c->vec[1]))) return rc;
    launch_sgs_apply(c, c->kry.blk, c->kry.aoff, c->vec[1], 64, c->vec[2], 64, precision, transpose != 0);
    if ((rc = download_state(c, c->vec[2], z6))) return rc;
  }
  if (Aoff || nbr) {
    const size_t na = (size_t)4 * nv * nv * P_;
    std::vector<double> hA(Aoff ? na : 0);
    std::vector<float> hAf;
    std::vector<int4> hq(P_);
    hipError_t e = hipMemcpyAsync(hq.data(), c->d_nq, (size_t)P_ * sizeof(int4), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && Aoff) {
      if (precision == 64) {
        e = hipMemcpyAsync(hA.data(), c->kry.aoff, na * 8, hipMemcpyDeviceToHost, c->stream);
      } else {
        hAf.resize(na);
        e = hipMemcpyAsync(hAf.data(), c->kry.aoff, na * 4, hipMemcpyDeviceToHost, c->stream);
      }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, CFD_ERR_HIP, hipGetErrorString(e));
    if (precision == 32)
      for (size_t i = 0; i < hAf.size(); ++i) hA[i] = hAf[i];
    std::vector<int> cpos(c->N);
    for (int t = 0; t < c->N; ++t) cpos[c->mesh->ccells[t]] = t;
    for (int i = 0; i < c->N; ++i) {
      const int r = c->iperm[i], b = cpos[i];  // blocks are stored color-major
      for (int q = 0; q < 4; ++q) {
        const int e4 = (&hq[i].x)[q];
        if (nbr) nbr[(size_t)r * 4 + q] = e4 < 0 ? -1 : c->iperm[e4 >> 2];
        if (!Aoff) continue;
        for (int k = 0; k < nv * nv; ++k)
          Aoff[((size_t)r * 4 + q) * nv * nv + k] = e4 < 0 ? 0.0 : hA[(size_t)(q * nv * nv + k) * P_ + b];
      }
    }
  }
  if (color)
    for (int i = 0; i < c->N; ++i) color[c->iperm[i]] = c->mesh->cell_color[i];
  return sync(c);
}

int cfd_line_precond(cfd_ctx* c, const double* U6, int precision, const double* x6, double* z6) {
  if (!c || !U6 || !x6 || !z6 || (precision != 64 && precision != 32)) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  int rc;
  if ((rc = upload_state(c, U6, c->vec[0]))) return rc;
  if ((rc = ensure_krylov(c, c->kry.m > 0 ? c->kry.m : 1, precision))) return rc;
  if ((rc = ensure_line(c))) return rc;
  Krylov& k = c->kry;
  launch_block_precond(c, c->vec[0], k.blk, precision, k.bd, k.aoff, c->d_lpos);
  launch_line_factor(c, k.bd, k.aoff, k.blk, k.lineP, k.lineG, k.lineA, precision);
  if ((rc = upload_state(c, x6, c->vec[1]))) return rc;
  launch_line_apply(c, k.blk, k.lineP, k.lineG, k.lineA, c->vec[1], 64, c->vec[2], 64, precision);
  return download_state(c, c->vec[2], z6);
}

int cfd_norm_l2(cfd_ctx* c, const double* v6, double* out) {
  if (!c || !v6 || !out) return CFD_ERR_ARG;
  int rc = upload_state(c, v6, c->vec[0]);
  if (rc) return rc;
  launch_reduce(c, RED_L2, c->vec[0], nullptr, S_L2);
  double s;
  if ((rc = read_scal(c, S_L2, 1, &s))) return rc;
  *out = std::sqrt(s);
  return CFD_OK;
}

int cfd_dot(cfd_ctx* c, const double* a6, const double* b6, double* out) {
  if (!c || !a6 || !b6 || !out) return CFD_ERR_ARG;
  int rc;
  if ((rc = upload_state(c, a6, c->vec[0]))) return rc;
  if ((rc = upload_state(c, b6, c->vec[1]))) return rc;
  launch_reduce(c, RED_DOT, c->vec[0], c->vec[1], S_DOT);
  return read_scal(c, S_DOT, 1, out);
}

int cfd_norm_vw(cfd_ctx* c, const double* R6, double* out) {
  if (!c || !R6 || !out) return CFD_ERR_ARG;
  int rc = upload_state(c, R6, c->vec[0]);
  if (rc) return rc;
  launch_reduce(c, RED_VW, c->vec[0], nullptr, S_VW);
  double s[2];
  if ((rc = read_scal(c, S_VW, 2, s))) return rc;
  *out = s[1] == 0 ? 0.0 : std::sqrt(s[0] / s[1]);
  return CFD_OK;
}

// Row scale factors 1 / |reference magnitude| of each conserved equation,
// from the frozen freestream (momentum uses rho (|u| + a)).
static void row_scales(const cfd_ctx* c, double* s) {
  const cfd_params& p = c->P;
  const double g = p.gamma;
  const double fT = 1.0 + 0.5 * (g - 1.0) * p.M_inf * p.M_inf;
  const double Ts = p.T_total_inlet / fT;
  const double rho = p.P_static_exit / (p.R_gas * Ts);
  const double a = std::sqrt(g * p.R_gas * Ts);
  const double u = p.M_inf * a;
  const double E = p.P_static_exit / (g - 1.0) + 0.5 * rho * u * u;
  s[0] = 1.0 / rho;
  s[1] = s[2] = 1.0 / (rho * (u + a));
  s[3] = 1.0 / E;
  s[4] = p.turb ? 1.0 / (rho * p.turb_k_inf) : 1.0;
  s[5] = p.turb ? 1.0 / (rho * p.turb_w_inf) : 1.0;
}

// ---- GMRES (Solver.cpp:1472-1618) on device vectors -------------------------------
// One GMRES(m) cycle from x = 0 on the right-hand side sign * r, written to x.
// Returns the number of Arnoldi steps (k + 1) and the cycle's final
// least-squares residual |g_{k+1}| in *res_out.
static int gmres_cycle(cfd_ctx* c, const double* U, const double* Rb, const double* r, double sign, double* x,
                       const cfd_gmres_opts* o, double tol_abs, int* iters, double* res_out, double* tmpU,
                       double* tmpR) {
  const int m = o->m, prec = o->precision;
  Krylov& k = c->kry;
  const size_t es = prec == 32 ? 4 : 8;
  const size_t vs = (size_t)c->P.nvar * c->pitch;
  char* V = (char*)k.V;
  const bool bj = o->precond == 1, sgs = o->precond == 2, line = o->precond == 3;
  const bool tr = o->jv == 2;
  double sc[6];
  row_scales(c, sc);
  int rc;
  if (o->scaling) {  // scaled copy of the right-hand side
    HIP_TRY(c, hipMemcpyAsync(k.w2, r, vs * 8, hipMemcpyDeviceToDevice, c->stream));
    launch_rowscale(c, k.w2, sc, 64);
    r = k.w2;
  }
  launch_reduce(c, RED_L2, r, nullptr, S_BETA);
  double beta2;
  if ((rc = read_scal(c, S_BETA, 1, &beta2))) return rc;
  const double beta = std::sqrt(beta2);
  launch_fill(c, x, 0.0, vs);
  *res_out = beta;
  if (beta < 1e-12) {
    *iters = 0;
    return sync(c);
  }
  launch_krylov_init(c, V, r, sign, S_BETA, prec);
  std::vector<std::vector<double>> H(m + 1, std::vector<double>(m, 0.0));
  std::vector<double> sn(m, 0.0), cs(m, 0.0), g(m + 1, 0.0);
  g[0] = beta;
  // Arnoldi step j (Gram-Schmidt with delayed reorthogonalisation): the
  // preconditioner and the product act on V[j], the candidate orthogonalised
  // once; the step then finishes V[j] (the basis vector q_j) and column j-1 of
  // H, which goes to pinned slot j-1 (event col_ev[(j-1) & 1]), and leaves the
  // next candidate in V[j+1].  Step m only finishes column m-1.
  auto step = [&](int j) -> int {
    double* hslot = j > 0 ? k.hcol + (size_t)k.ld * (j - 1) : nullptr;
    if (j < m) {
      void* Vj = V + (size_t)j * vs * es;
      void* Vj1 = V + (size_t)(j + 1) * vs * es;
      if (bj) launch_block_apply(c, k.blk, Vj, k.z, prec, tr);
      else if (sgs) launch_sgs_apply(c, k.blk, k.aoff, Vj, prec, k.z, prec, prec, tr);
      else if (line) launch_line_apply(c, k.blk, k.lineP, k.lineG, k.lineA, Vj, prec, k.z, prec, prec);
      else launch_precond_apply(c, k.z, Vj, prec, o->jv == 0 ? S_VNORM : -1);
      if (o->jv == 0) {
        if (bj || sgs || line) launch_reduce(c, RED_L2, (const double*)k.z, nullptr, S_VNORM);
        launch_jv_fd(c, U, Rb, (const double*)k.z, (double*)Vj1, S_UNORM, S_VNORM, tmpU, tmpR);
      } else if (o->jv == 1) {
        launch_jv_tangent(c, U, k.z, Vj1, prec);
      } else if (prec == 32) {  // the reverse product runs in FP64 on FP32 basis vectors
        launch_f2d(c, (const float*)k.z, tmpU, vs);
        launch_vjp(c, U, tmpU, tmpR, true);
        launch_d2f(c, tmpR, (float*)Vj1, vs);
      } else {
        launch_vjp(c, U, (const double*)k.z, (double*)Vj1, true);
      }
      if (o->scaling) launch_rowscale(c, Vj1, sc, prec);
      launch_dcgs(c, V, j, false, hslot, prec);
    } else {
      launch_dcgs(c, V, j, true, hslot, prec);
    }
    if (j > 0) HIP_TRY(c, hipEventRecord(k.col_ev[(j - 1) & 1], c->stream));
    return CFD_OK;
  };
  // Column col is final after step col + 1.  The device runs up to two steps
  // ahead of the host's rotations; the steps past the column that meets the
  // tolerance were extra work that nothing reads.  Column col's event is
  // waited for before step col + 3 records the same event again.
  int kk = 0;
  if ((rc = step(0))) return rc;
  if ((rc = step(1))) return rc;
  for (int j = 0; j < m; ++j) {
    kk = j;
    if (j + 2 <= m && (rc = step(j + 2))) return rc;
    HIP_TRY(c, hipEventSynchronize(k.col_ev[j & 1]));
    const double* col = k.hcol + (size_t)k.ld * j;
    for (int i = 0; i <= j + 1; ++i) H[i][j] = col[i];
    for (int i = 0; i < j; ++i) {
      const double t = cs[i] * H[i][j] + sn[i] * H[i + 1][j];
      H[i + 1][j] = -sn[i] * H[i][j] + cs[i] * H[i + 1][j];
      H[i][j] = t;
    }
    const double a = H[j][j], b = H[j + 1][j];
    if (std::abs(b) < 1e-14) {
      cs[j] = 1.0;
      sn[j] = 0.0;
    } else {
      const double t = std::sqrt(a * a + b * b);
      cs[j] = a / t;
      sn[j] = b / t;
    }
    H[j][j] = cs[j] * a + sn[j] * b;
    H[j + 1][j] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    if (std::abs(g[j + 1]) < tol_abs) break;
  }
  std::vector<double> y(kk + 1);
  for (int i = kk; i >= 0; --i) {
    y[i] = g[i];
    for (int l = i + 1; l <= kk; ++l) y[i] -= H[i][l] * y[l];
    y[i] /= H[i][i];
  }
  std::memcpy(k.hpin, y.data(), (kk + 1) * 8);
  double* yd = k.h + k.yoff;
  HIP_TRY(c, hipMemcpyAsync(yd, k.hpin, (kk + 1) * 8, hipMemcpyHostToDevice, c->stream));
  if (bj) {
    launch_solution_bj(c, x, V, yd, kk, k.blk, prec, tr);
  } else if (sgs) {
    launch_solution(c, k.xt, V, yd, kk, prec, false);
    launch_sgs_apply(c, k.blk, k.aoff, k.xt, 64, x, 64, prec, tr);
  } else if (line) {
    launch_solution(c, k.xt, V, yd, kk, prec, false);
    launch_line_apply(c, k.blk, k.lineP, k.lineG, k.lineA, k.xt, 64, x, 64, prec);
  } else {
    launch_solution(c, x, V, yd, kk, prec);
  }
  *iters = kk + 1;
  *res_out = std::abs(g[kk + 1]);
  return sync(c);
}

// GMRES(m) with o->restarts extra cycles; rhs = sign * r.  Restart cycles
// solve for the correction on the true residual rhs - A x.
static int gmres_dev(cfd_ctx* c, const double* U, const double* Rb, const double* r, double sign, double* x,
                     const cfd_gmres_opts* o, int* iters, double* tmpU, double* tmpR) {
  const int m = o->m, prec = o->precision;
  if (m < 1 || m > cfd::GMRES_M_MAX) return fail(c, CFD_ERR_ARG, "GMRES m must be in [1, 250]");
  if (prec != 64 && prec != 32) return fail(c, CFD_ERR_ARG, "precision must be 64 or 32");
  if (o->jv < 0 || o->jv > 2) return fail(c, CFD_ERR_ARG, "jv must be 0 (FD), 1 (tangent) or 2 (transpose)");
  if (o->jv == 0 && prec != 64) return fail(c, CFD_ERR_ARG, "the FD Jv runs in FP64 only");
  if (o->precond < 0 || o->precond > 3) return fail(c, CFD_ERR_ARG, "unknown preconditioner");
  if (o->precond == 3 && o->jv == 2)
    return fail(c, CFD_ERR_ARG, "the line preconditioner has no transposed form (use precond 2 for jv 2)");
  if (o->restarts < 0) return fail(c, CFD_ERR_ARG, "restarts must be >= 0");
  int rc = ensure_krylov(c, m, prec);
  if (rc) return rc;
  if (o->jv == 2 && (rc = ensure_adjoint(c))) return rc;
  Krylov& k = c->kry;
  const size_t vs = (size_t)c->P.nvar * c->pitch;
  if (o->precond == 1) {
    launch_block_precond(c, U, k.blk, prec);
  } else if (o->precond == 2) {
    if ((rc = ensure_sgs(c))) return rc;
    launch_block_precond(c, U, k.blk, prec, nullptr, k.aoff, c->d_cpos);
  } else if (o->precond == 3) {
    if ((rc = ensure_line(c))) return rc;
    launch_block_precond(c, U, k.blk, prec, k.bd, k.aoff, c->d_lpos);
    launch_line_factor(c, k.bd, k.aoff, k.blk, k.lineP, k.lineG, k.lineA, prec);
  } else {
    launch_diag_precond(c, U, k.D);
  }
  // U is fixed for the whole solve: the reverse products share its gradients,
  // formed once here (nothing in the solve writes them with another state)
  if (o->jv == 2) launch_vjp_gradients(c, U);
  launch_reduce(c, RED_L2, U, nullptr, S_UNORM);
  double b2;
  if (o->scaling) {
    double sc[6];
    row_scales(c, sc);
    HIP_TRY(c, hipMemcpyAsync(k.w2, r, vs * 8, hipMemcpyDeviceToDevice, c->stream));
    launch_rowscale(c, k.w2, sc, 64);
    launch_reduce(c, RED_L2, k.w2, nullptr, S_BETA);
  } else {
    launch_reduce(c, RED_L2, r, nullptr, S_BETA);
  }
  if ((rc = read_scal(c, S_BETA, 1, &b2))) return rc;
  const double tol_abs = o->tol * std::sqrt(b2);
  int total = 0, it = 0;
  double res = 0.0;
  if ((rc = gmres_cycle(c, U, Rb, r, sign, x, o, tol_abs, &it, &res, tmpU, tmpR))) return rc;
  total += it;
  if (o->restarts > 0) {
    double* xs = k.w;  // accumulated solution
    HIP_TRY(c, hipMemcpyAsync(xs, x, vs * 8, hipMemcpyDeviceToDevice, c->stream));
    for (int cyc = 0; cyc < o->restarts && res >= tol_abs && it > 0; ++cyc) {
      // true residual: t = sign * r - A xs  (into tmpR, via the unpreconditioned operator)
      double* Ax = k.rr;
      if (o->jv == 0) {
        launch_reduce(c, RED_L2, xs, nullptr, S_VNORM);
        launch_jv_fd(c, U, Rb, xs, Ax, S_UNORM, S_VNORM, tmpU, tmpR);
      } else if (o->jv == 1) {
        if (prec == 32) {
          float* xf = (float*)k.z;
          float* af = (float*)tmpU;
          launch_d2f(c, xs, xf, vs);
          launch_jv_tangent(c, U, xf, af, 32);
          launch_f2d(c, af, Ax, vs);
        } else {
          launch_jv_tangent(c, U, xs, Ax, 64);
        }
      } else {
        launch_vjp(c, U, xs, Ax, true);
      }
      launch_axpby(c, Ax, -1.0, r, sign, vs);  // Ax <- sign r - A xs
      if ((rc = gmres_cycle(c, U, Rb, Ax, 1.0, x, o, tol_abs, &it, &res, tmpU, tmpR))) return rc;
      total += it;
      launch_axpby(c, xs, 1.0, x, 1.0, vs);
    }
    HIP_TRY(c, hipMemcpyAsync(x, xs, vs * 8, hipMemcpyDeviceToDevice, c->stream));
  }
  *iters = total;
  c->kry.last_rel = tol_abs > 0 ? res / (tol_abs / o->tol) : 0.0;
  return sync(c);
}

int cfd_gmres(cfd_ctx* c, const double* U6, const double* R6, const double* RHS6, double* x6,
              const cfd_gmres_opts* opts, int* iters_out) {
  if (!c || !U6 || !R6 || !RHS6 || !x6) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  cfd_gmres_opts o;
  if (opts) o = *opts;
  else cfd_gmres_opts_default(&o);
  int rc;
  if ((rc = upload_state(c, U6, c->vec[0]))) return rc;
  if ((rc = upload_state(c, R6, c->vec[1]))) return rc;
  if ((rc = upload_state(c, RHS6, c->vec[2]))) return rc;
  int it = 0;
  if ((rc = gmres_dev(c, c->vec[0], c->vec[1], c->vec[2], 1.0, c->vec[3], &o, &it, c->vec[4], c->vec[5]))) return rc;
  if (iters_out) *iters_out = it;
  return download_state(c, c->vec[3], x6);
}

// ---- Newton / pseudo-transient driver (Solver.cpp:1696-1799) ---------------------
int cfd_solve_implicit(cfd_ctx* c, double* U6, const cfd_newton_opts* opts, cfd_newton_result* res,
                       double* hist_vm, int* hist_iters, double* hist_cfl) {
  if (!c || !U6) return CFD_ERR_ARG;
  cfd_newton_opts o;
  if (opts) o = *opts;
  else cfd_newton_opts_default(&o);
  const auto t0 = std::chrono::steady_clock::now();
  double* U = c->vec[0];
  double* R = c->vec[1];
  double* dU = c->vec[2];
  int rc = upload_state(c, U6, U);
  if (rc) return rc;
  if (o.controller < 0 || o.controller > 2) return fail(c, CFD_ERR_ARG, "controller must be 0, 1 or 2");
  double cfl = o.cfl0 > 0 ? o.cfl0 : 0.1;
  double vm0 = -1.0, vm = 0.0, vm_prev = -1.0;
  bool gmres_ok = true;
  double* Ubak = c->vec[3];  // adaptive controller: state before the last update
  const size_t vs = (size_t)c->P.nvar * c->pitch;
  int step = 0, conv = 0, total_it = 0;
  for (step = 0; step < o.max_steps; ++step) {
    launch_timestep(c, U, cfl);
    const bool upd = !(step > 150 || cfl > 10.0);
    launch_residual(c, U, R, upd);
    launch_reduce(c, RED_VW, R, nullptr, S_VW);
    double s[2];
    if ((rc = read_scal(c, S_VW, 2, s))) return rc;
    vm = s[1] == 0 ? 0.0 : std::sqrt(s[0] / s[1]);
    if (o.controller == 2 && vm_prev > 0) {
      const double cmin = 1e-2 * (o.cfl0 > 0 ? o.cfl0 : 0.1);
      const double cmax = o.cfl_max > 0 ? o.cfl_max : 1e12;
      if (!(vm == vm) || vm > 10.0 * vm_prev) {  // reject: back to the previous state, smaller step
        HIP_TRY(c, hipMemcpyAsync(U, Ubak, vs * 8, hipMemcpyDeviceToDevice, c->stream));
        cfl *= 0.1;
        launch_timestep(c, U, std::max(cfl, cmin));
        launch_residual(c, U, R, upd);
        vm = vm_prev;
      } else if (vm > 1.2 * vm_prev) {
        cfl *= 0.5;
      } else if (vm <= 0.99 * vm_prev) {
        cfl *= gmres_ok ? 2.0 : 1.25;
      } else if (!gmres_ok) {  // stalled with an unconverged linear solve: an easier system
        cfl *= 0.7;
      } else {  // a stall or small rise with GMRES at its tolerance: mild growth, so that a
        cfl *= 1.1;  // transient that must pass through higher residuals is not frozen
      }
      cfl = std::min(std::max(cfl, cmin), cmax);
      launch_timestep(c, U, cfl);
    }
    if (vm0 < 0) vm0 = vm;
    if (hist_vm) hist_vm[step] = vm;
    if (hist_cfl) hist_cfl[step] = cfl;
    if (!(vm == vm)) return fail(c, CFD_ERR_NUMERIC, "residual norm is NaN");
    if (vm < o.tolerance || (o.rel_drop > 0 && vm <= o.rel_drop * vm0)) {
      conv = 1;
      if (hist_iters) hist_iters[step] = 0;
      break;
    }
    int it = 0;
    if ((rc = gmres_dev(c, U, R, R, -1.0, dU, &o.gmres, &it, c->vec[4], c->vec[5]))) return rc;
    total_it += it;
    if (hist_iters) hist_iters[step] = it;
    if (o.log == 1) std::printf("{\"step\": %d, \"vm\": %.6e, \"cfl\": %.4g, \"gmres\": %d}\n", step, vm, cfl, it);
    if (o.log > 1 && step % o.log == 0) {
      std::fprintf(stderr, "{\"step\": %d, \"vm\": %.6e, \"cfl\": %.4g, \"gmres\": %d}\n", step, vm, cfl, it);
      std::fflush(stderr);
    }
    if (o.controller == 0) {  // Solver.cpp:1747-1796
      if (it >= 29) {
        cfl *= 0.5;
        if (cfl < 0.05) cfl = 0.05;
      }
      const double relax = vm > 1e6 ? 0.5 : 1.0;
      launch_newton_update(c, U, dU, relax);
      if (step > 5) {
        cfl *= vm > 1e6 ? 1.2 : 1.5;
        if (it >= 25) cfl *= 0.5;
        if (cfl > 10000) cfl = 10000;
        if (cfl < 0.1) cfl = 0.1;
      }
    } else if (o.controller == 2) {
      gmres_ok = c->kry.last_rel <= 1.5 * o.gmres.tol;
      vm_prev = vm;
      HIP_TRY(c, hipMemcpyAsync(Ubak, U, vs * 8, hipMemcpyDeviceToDevice, c->stream));
      launch_newton_update(c, U, dU, 1.0);
    } else {  // switched evolution relaxation, backed off when GMRES misses its tolerance
      launch_newton_update(c, U, dU, 1.0);
      const double cmax = o.cfl_max > 0 ? o.cfl_max : 1e12;
      const double c0 = o.cfl0 > 0 ? o.cfl0 : 0.1;
      double next = c0 * (vm0 / vm);
      if (next > 2.0 * cfl) next = 2.0 * cfl;   // at most doubling per step
      if (next < 0.1 * cfl) next = 0.1 * cfl;
      if (c->kry.last_rel > 1.5 * o.gmres.tol) next = 0.5 * cfl;
      if (next < 1e-3 * c0) next = 1e-3 * c0;
      if (next > cmax) next = cmax;
      cfl = next;
    }
  }
  if ((rc = download_state(c, U, U6))) return rc;
  if (res) {
    res->steps = step;
    res->converged = conv;
    res->res_vm0 = vm0;
    res->res_vm = vm;
    res->total_gmres_iters = total_it;
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return CFD_OK;
}

// ---- dual time stepping (Solver.cpp:1801-1864) ---------------------------------
int cfd_solve_unsteady(cfd_ctx* c, double* U6, double dt_phys, int n_steps, const cfd_gmres_opts* gopts,
                       int max_newton, double tolerance, double relax, double* hist_vm, int* hist_newton) {
  if (!c || !U6 || !(dt_phys > 0) || n_steps < 0) return CFD_ERR_ARG;
  // hist_vm rows are max_newton long as the caller sized them: a default
  // chosen here would overrun a buffer sized for the value passed
  if (max_newton <= 0 && hist_vm)
    return fail(c, CFD_ERR_ARG, "max_newton must be given (> 0) when hist_vm is requested");
  cfd_gmres_opts g;
  if (gopts) g = *gopts;
  else cfd_gmres_opts_default(&g);
  if (max_newton <= 0) max_newton = 15;  // Solver.cpp:1819-1820
  if (!(tolerance > 0)) tolerance = 1e-6;
  if (!(relax > 0)) relax = 0.7;         // Solver.cpp:1858
  const size_t vs = (size_t)c->P.nvar * c->pitch;
  if (!c->aux) HIP_TRY(c, hipMalloc((void**)&c->aux, (size_t)6 * c->pitch * 8));
  double *U = c->vec[0], *R = c->vec[1], *dU = c->vec[2], *Uold = c->vec[3], *rhs = c->aux;
  int rc = upload_state(c, U6, U);
  if (rc) return rc;
  // the pseudo-time term of the Jacobian becomes vol / dt_phys (Solver.cpp:1808-1809)
  launch_fill(c, c->dt, dt_phys, (size_t)c->pitch);
  c->dt_valid = true;
  for (int t = 0; t < n_steps; ++t) {
    HIP_TRY(c, hipMemcpyAsync(Uold, U, vs * 8, hipMemcpyDeviceToDevice, c->stream));
    int step = 0;
    for (step = 0; step < max_newton; ++step) {
      launch_residual(c, U, R, true);
      launch_unsteady_rhs(c, U, Uold, R, dt_phys, rhs);
      launch_reduce(c, RED_VW4, rhs, nullptr, S_VW);
      double sv[2];
      if ((rc = read_scal(c, S_VW, 2, sv))) return rc;
      const double vm = sv[1] == 0 ? 0.0 : std::sqrt(sv[0] / sv[1]);
      if (hist_vm) hist_vm[(size_t)t * max_newton + step] = vm;
      if (!(vm == vm)) return fail(c, CFD_ERR_NUMERIC, "unsteady residual norm is NaN");
      if (vm < tolerance && step > 0) break;
      int it = 0;
      if ((rc = gmres_dev(c, U, R, rhs, 1.0, dU, &g, &it, c->vec[4], c->vec[5]))) return rc;
      launch_axpby(c, U, 1.0, dU, relax, vs);  // U <- U + relax dU, no clips (Solver.cpp:1856-1859)
    }
    if (hist_newton) hist_newton[t] = step;
  }
  return download_state(c, U, U6);
}

// ---- verification outputs (Solver.cpp:1620-1646, 1866-1897) -----------------------
int cfd_entropy_error(cfd_ctx* c, const double* U6, double* err) {
  if (!c || !U6 || !err) return CFD_ERR_ARG;
  int rc = upload_state(c, U6, c->vec[0]);
  if (rc) return rc;
  const cfd_params& P = c->P;
  const double f = 1.0 + 0.5 * (P.gamma - 1.0) * P.M_inf * P.M_inf;
  const double T_inf = P.T_total_inlet / f;
  const double P_inf = P.P_total_inlet / std::pow(f, P.gamma / (P.gamma - 1.0));
  const double rho_inf = P_inf / (P.R_gas * T_inf);
  const double S_ref = P_inf / std::pow(rho_inf, P.gamma);
  launch_entropy(c, c->vec[0], S_ref, S_L2);
  double s[2];
  if ((rc = read_scal(c, S_L2, 2, s))) return rc;
  *err = std::sqrt(s[0] / s[1]);
  return CFD_OK;
}

int cfd_bl_profile(cfd_ctx* c, const double* U6, double x_target, int capacity, double* rows4, int* n_rows) {
  if (!c || !U6 || !n_rows || capacity < 0 || (capacity > 0 && !rows4)) return CFD_ERR_ARG;
  int rc = upload_state(c, U6, c->vec[0]);
  if (rc) return rc;
  const cfd_params& P = c->P;
  const double T_static = P.T_total_inlet / (1.0 + 0.5 * (P.gamma - 1.0) * P.M_inf * P.M_inf);
  const double u_inf = P.M_inf * std::sqrt(P.gamma * P.R_gas * T_static);
  double* out4 = c->vec[1];
  unsigned char* flag = (unsigned char*)c->vec[2];
  launch_bl_profile(c, c->vec[0], x_target, u_inf, P.nu, out4, flag);
  const int P_ = c->pitch;
  std::vector<double> h((size_t)4 * P_);
  std::vector<unsigned char> fl(c->N);
  HIP_TRY(c, hipMemcpyAsync(h.data(), out4, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipMemcpyAsync(fl.data(), flag, c->N, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  int n = 0;
  for (int i = 0; i < c->N; ++i) {  // rows in reference cell order
    const int d = c->perm[i];
    if (!fl[d]) continue;
    if (n < capacity)
      for (int k = 0; k < 4; ++k) rows4[(size_t)n * 4 + k] = h[(size_t)k * P_ + d];
    ++n;
  }
  *n_rows = n;
  return n > capacity ? fail(c, CFD_ERR_ARG, "boundary-layer profile has more rows than the capacity") : CFD_OK;
}

int cfd_explicit_step(cfd_ctx* c, double* U6, double dt) {
  if (!c || !U6) return CFD_ERR_ARG;
  int rc = upload_state(c, U6, c->vec[0]);
  if (rc) return rc;
  launch_residual(c, c->vec[0], c->vec[1], true);
  launch_explicit(c, c->vec[0], c->vec[0], 0.0, c->vec[0], 1.0, c->vec[1], dt);
  return download_state(c, c->vec[0], U6);
}

int cfd_rk3_step(cfd_ctx* c, double* U6, double dt) {
  if (!c || !U6) return CFD_ERR_ARG;
  int rc = upload_state(c, U6, c->vec[0]);
  if (rc) return rc;
  double *U0 = c->vec[0], *R = c->vec[1], *U1 = c->vec[2], *U2 = c->vec[3];
  launch_residual(c, U0, R, true);
  launch_explicit(c, U1, U0, 0.0, U0, 1.0, R, dt);
  launch_residual(c, U1, R, true);
  launch_explicit(c, U2, U0, 0.75, U1, 0.25, R, dt);
  launch_residual(c, U2, R, true);
  launch_explicit(c, U0, U0, 1.0 / 3.0, U2, 2.0 / 3.0, R, dt);
  return download_state(c, U0, U6);
}

// Per reference face: Cp and Cf (k_wall), and the owner pressure of the
// wall faces (tags 3 and 4, 0 elsewhere); host vectors in reference order.
static int wall_faces(cfd_ctx* c, const double* U6, std::vector<double>& cp, std::vector<double>& cf,
                      std::vector<double>& pw) {
  int rc = upload_state(c, U6, c->vec[0]);
  if (rc) return rc;
  double* out3 = c->Fm;  // 3 x F scratch (face buffer is free between residuals)
  launch_wall(c, c->vec[0], out3);
  std::vector<double> h((size_t)3 * c->F);
  HIP_TRY(c, hipMemcpyAsync(h.data(), out3, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(c, hipStreamSynchronize(c->stream));
  cp.assign(c->F, 0.0);
  cf.assign(c->F, 0.0);
  pw.assign(c->F, 0.0);
  for (int d = 0; d < c->F; ++d) {
    const int f = c->face_order[d];
    cp[f] = h[d];
    cf[f] = h[c->F + d];
    pw[f] = h[2 * c->F + d];
  }
  return CFD_OK;
}

int cfd_wall_outputs(cfd_ctx* c, const double* U6, double* Cp, double* Cf, double* J) {
  if (!c || !U6) return CFD_ERR_ARG;
  std::vector<double> cp, cf, pw;
  if (int rc = wall_faces(c, U6, cp, cf, pw)) return rc;
  const cfd_mesh& m = *c->mesh;
  double j = 0.0;
  for (int f = 0; f < c->F; ++f) {
    if (Cp) Cp[f] = cp[f];
    if (Cf) Cf[f] = cf[f];
    if (m.nei[f] < 0 && m.tag[f] == 3) j += pw[f] * m.area[f];
  }
  if (J) *J = j;
  return CFD_OK;
}

int cfd_drag(cfd_ctx* c, const double* U6, double* drag) {
  if (!c || !U6 || !drag) return CFD_ERR_ARG;
  std::vector<double> cp, cf, pw;
  if (int rc = wall_faces(c, U6, cp, cf, pw)) return rc;
  const cfd_mesh& m = *c->mesh;
  // Solver.cpp:1310-1321: over the faces in order, Fx = -p A n_x on tag 3
  double d = 0.0;
  for (int f = 0; f < c->F; ++f)
    if (m.tag[f] == 3) {
      const double fx = -pw[f] * m.area[f] * m.nx[f];
      d += fx;
    }
  *drag = d;
  return CFD_OK;
}

int cfd_wall_pressure(cfd_ctx* c, const double* U6, double y_max, double x_scale, double p_divisor, int capacity,
                      double* rows2, int* n_rows) {
  if (!c || !U6 || !n_rows || capacity < 0 || (capacity > 0 && !rows2) || p_divisor == 0.0) return CFD_ERR_ARG;
  std::vector<double> cp, cf, pw;
  if (int rc = wall_faces(c, U6, cp, cf, pw)) return rc;
  const cfd_mesh& m = *c->mesh;
  // Solver.cpp:2043-2062: tag-3 faces whose owner centroid lies below y_max
  std::vector<std::pair<double, double>> s;
  for (int f = 0; f < c->F; ++f)
    if (m.tag[f] == 3 && m.yc[m.own[f]] < y_max) s.emplace_back(m.xm[f] * x_scale, pw[f] / p_divisor);
  // ascending x; equal x (the reference's order is unspecified there) by pressure
  std::sort(s.begin(), s.end());
  *n_rows = (int)s.size();
  if ((int)s.size() > capacity) return fail(c, CFD_ERR_ARG, "wall pressure rows exceed the capacity");
  for (size_t i = 0; i < s.size(); ++i) {
    rows2[2 * i] = s[i].first;
    rows2[2 * i + 1] = s[i].second;
  }
  return CFD_OK;
}

// ---- device-handle variants --------------------------------------------------------
int cfd_dev_pack(cfd_ctx* c, const double* U6_dev, double* U) {
  if (!c || !U6_dev || !U) return CFD_ERR_ARG;
  launch_pack(c, U6_dev, U);
  HIP_TRY(c, hipGetLastError());
  return CFD_OK;
}
int cfd_dev_unpack(cfd_ctx* c, const double* U, double* U6_dev) {
  if (!c || !U6_dev || !U) return CFD_ERR_ARG;
  launch_unpack(c, U, U6_dev);
  HIP_TRY(c, hipGetLastError());
  return CFD_OK;
}
int cfd_dev_residual(cfd_ctx* c, const double* U, double* R, int update_limiters) {
  if (!c || !U || !R) return CFD_ERR_ARG;
  launch_residual(c, U, R, update_limiters != 0);
  HIP_TRY(c, hipGetLastError());
  return CFD_OK;
}
int cfd_dev_local_timestep(cfd_ctx* c, const double* U, double cfl) {
  if (!c || !U) return CFD_ERR_ARG;
  launch_timestep(c, U, cfl);
  HIP_TRY(c, hipGetLastError());
  return CFD_OK;
}
int cfd_dev_jv_fd(cfd_ctx* c, const double* U, const double* Rb, const double* v, double* Jv) {
  if (!c || !U || !Rb || !v || !Jv) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  launch_reduce(c, RED_L2, U, nullptr, S_UNORM);
  launch_reduce(c, RED_L2, v, nullptr, S_VNORM);
  launch_jv_fd(c, U, Rb, v, Jv, S_UNORM, S_VNORM, c->vec[4], c->vec[5]);
  HIP_TRY(c, hipGetLastError());
  return CFD_OK;
}
int cfd_dev_jv_tangent(cfd_ctx* c, const double* U, const void* v, void* Jv, int precision) {
  if (!c || !U || !v || !Jv || (precision != 64 && precision != 32)) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  if (int rc = need_limiter(c)) return rc;
  launch_jv_tangent(c, U, v, Jv, precision);
  HIP_TRY(c, hipGetLastError());
  return CFD_OK;
}

int cfd_dev_vjp_adjoint(cfd_ctx* c, const double* U, const double* w, double* JTw) {
  if (!c || !U || !w || !JTw) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  if (int rc = need_limiter(c)) return rc;
  if (int rc = ensure_adjoint(c)) return rc;
  launch_vjp(c, U, w, JTw);
  HIP_TRY(c, hipGetLastError());
  return CFD_OK;
}

int cfd_dev_gmres(cfd_ctx* c, const double* U, const double* R, const double* rhs, double* x,
                  const cfd_gmres_opts* opts, int* iters_out) {
  if (!c || !U || !R || !rhs || !x) return CFD_ERR_ARG;
  if (!c->dt_valid) return fail(c, CFD_ERR_STATE, "no local time step computed");
  if (int rc = need_limiter(c)) return rc;
  cfd_gmres_opts o;
  if (opts) o = *opts;
  else cfd_gmres_opts_default(&o);
  int it = 0;
  const int rc = gmres_dev(c, U, R, rhs, 1.0, x, &o, &it, c->vec[4], c->vec[5]);
  if (iters_out) *iters_out = it;
  return rc;
}

int cfd_dev_synchronize(cfd_ctx* c) {
  if (!c) return CFD_ERR_ARG;
  return sync(c);
}

int cfd_dev_time_residual(cfd_ctx* c, const double* U, double* R, int reps, int update_limiters, double* ms_out,
                          int* launches_out) {
  if (!c || !U || !R || reps < 1 || reps > 100000 || !ms_out) return CFD_ERR_ARG;
  // five events per residual, recorded back to back on the context stream;
  // the host waits once, after the last residual
  std::vector<hipEvent_t> ev((size_t)5 * reps, nullptr);
  int rc = CFD_OK;
  for (auto& e : ev) {
    hipError_t er = hipEventCreate(&e);
    if (er != hipSuccess) {
      rc = fail(c, CFD_ERR_HIP, hipGetErrorString(er));
      break;
    }
  }
  if (rc == CFD_OK) {
    for (int r = 0; r < reps; ++r) launch_residual_timed(c, U, R, update_limiters != 0, &ev[(size_t)5 * r]);
    hipError_t er = hipEventSynchronize(ev.back());
    if (er != hipSuccess) rc = fail(c, CFD_ERR_HIP, hipGetErrorString(er));
  }
  double acc[4] = {0, 0, 0, 0};
  if (rc == CFD_OK) {
    for (int r = 0; r < reps; ++r) {
      hipEvent_t* e = &ev[(size_t)5 * r];
      float t01 = 0, t12 = 0, t23 = 0, t34 = 0, t04 = 0;
      (void)hipEventElapsedTime(&t01, e[0], e[1]);
      (void)hipEventElapsedTime(&t12, e[1], e[2]);
      (void)hipEventElapsedTime(&t23, e[2], e[3]);
      (void)hipEventElapsedTime(&t34, e[3], e[4]);
      (void)hipEventElapsedTime(&t04, e[0], e[4]);
      acc[0] += t01;
      acc[1] += t23;
      acc[2] += t12 + t34;
      acc[3] += t04;
    }
  }
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  for (int i = 0; i < 4; ++i) ms_out[i] = acc[i] / reps;
  if (launches_out) *launches_out = c->assembly == 0 ? c->ncolors : 1;
  return rc;
}

}  // extern "C"

// Device context: SoA mesh arrays in RCM order, per-context buffers, and the
// launchers the C-ABI entry points are built from.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mesh.hpp"
#include "physics.hpp"

namespace cfd {

// Device view of the mesh.  Cells are in RCM order; faces are grouped by
// edge color (color-major, then owner position).  slots[c] lists the faces of
// cell c in ascending reference face index, each as (device face << 1) | side
// with side 1 when c is the neighbor; -1 pads triangles.
struct DMesh {
  int N, F, pitch;
  const double *vol, *xc, *yc;
  const int4* slots;
  const int4* cnb;  // per cell: the other cell of each slot's face, -1 for a boundary face or an empty slot
  const int *own, *nei, *tag;
  const double *nx, *ny, *area, *xm, *ym;
  const unsigned char* alt;  // per device face: 1 = use Phys::flux_alt (may be null)
  const double* wd;          // per cell: distance to the nearest no-slip face midpoint
  const double* dnw;         // per face: normal distance of the owner centroid to the face
  const int* corder;         // device faces listed color by color (ranges: cfd_ctx::color_ptr)
};

// Workspace of one GMRES solve.
struct Krylov {
  int m = 0, prec = 64;        // basis size, element precision (64 or 32)
  int ld = 64;                 // leading dimension of the Arnoldi step's H columns and coefficients: 64, or m + 2
  int yoff = 128;              // offset of the least-squares solution y in h
  void* V = nullptr;           // (m + 1) vectors, nv x pitch each
  double* D = nullptr;         // scalar preconditioner
  void* blk = nullptr;         // block-Jacobi inverses (nv x nv per cell)
  void* aoff = nullptr;        // neighbor coupling blocks (4 x nv x nv per cell), block SGS only
  double* xt = nullptr;        // basis combination before the block SGS solve (FP64)
  double* fjac = nullptr;      // face Jacobians of the first-order flux (2 nv^2 x F, block precision)
  double* bd = nullptr;        // FP64 diagonal blocks in segment order (nv^2 x pitch), line preconditioner
  void* lineG = nullptr;       // line factors G_t = Dp_t^-1 A_{t,next} (nv^2 per bundle slot, basis precision)
  void* lineP = nullptr;       // line factors P_t = Dp_t^-1 A_{t,prev} (nv^2 per bundle slot, basis precision)
  void* lineA = nullptr;       // off-segment coupling blocks by line position in packed slots: those of lower
                               // segment colors from slot 0 up, of higher colors from slot 3 down (4 nv^2 x pitch,
                               // block precision)
  void* lineS = nullptr;       // forward-pass partial sums of the line sweep by line position (nv x pitch, up to FP64)
  void* lineZ = nullptr;       // the line sweep's z by line position (nv x pitch, up to FP64)
  double* lineW = nullptr;     // w of the cells of multi-cell segments (nv per bundle slot, FP64)
  double* z = nullptr;         // preconditioned direction (FP64)
  double* w = nullptr;         // accumulated solution of restarted GMRES (FP64)
  double* rr = nullptr;        // true residual of a restart cycle (FP64)
  double* w2 = nullptr;        // row-scaled right-hand side (FP64)
  double last_rel = 0.0;       // relative residual reached by the last solve
  double* h = nullptr;         // device scalars: an Arnoldi step's reduction sums [0, 124), solution weights y [128, 192)
  double* part = nullptr;      // per-block partial sums of the Arnoldi reductions (2 (m + 2) x 1024)
  double* dc = nullptr;        // device Hessenberg matrix (64 x 64, column-major), partial column, step coefficients
  double* hpin = nullptr;      // pinned host staging of the solution weights y
  double* hcol = nullptr;      // pinned host slots of the Hessenberg columns, 64 doubles per Arnoldi step
  hipEvent_t col_ev[2] = {nullptr, nullptr};  // column j copied to its slot (ping-pong by j parity)
  size_t bytes = 0;
};

}  // namespace cfd

struct cfd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;      // the stream every launch and copy of the context goes to
  hipStream_t own_stream = nullptr;  // created with the context; `stream` unless cfd_ctx_set_stream chose another
  cfd_params P{};
  cfd::Phys phys{};
  int N = 0, F = 0, pitch = 0, ncolors = 0;
  std::vector<int> color_ptr;
  std::vector<int> color_bptr;  // per color: first boundary face in corder (interior faces lead each color)
  int4* d_bgrp = nullptr;       // per cell with boundary faces: those faces (device index, corder order; -1 pads)
  int nbgrp = 0;
  std::vector<int> face_order, perm, iperm;  // host copies
  std::vector<int> ccell_ptr;                 // cell color ranges into d_ccells
  cfd::DMesh dm{};
  int assembly = 1;  // 0 colored scatter, 1 face buffer + gather
  const cfd_mesh* mesh = nullptr;   // borrowed: must outlive the context (Solver holds const Mesh&)
  std::vector<unsigned char> alt_ref;  // per reference face flux choice
  // device mesh storage
  double *d_vol = nullptr, *d_xc = nullptr, *d_yc = nullptr;
  int4* d_slots = nullptr;
  int4* d_cnb = nullptr;
  int *d_own = nullptr, *d_nei = nullptr, *d_tag = nullptr, *d_iperm = nullptr, *d_facemap = nullptr;
  int* d_corder = nullptr;
  double *d_nx = nullptr, *d_ny = nullptr, *d_area = nullptr, *d_xm = nullptr, *d_ym = nullptr;
  double *d_wd = nullptr, *d_dnw = nullptr;
  unsigned char* d_alt = nullptr;
  unsigned char* d_ccol = nullptr;  // per cell: color of the distance-1 cell coloring
  int* d_ccells = nullptr;          // cells grouped by color
  int* d_cpos = nullptr;            // per cell: its position in d_ccells (column of its SGS blocks)
  std::vector<int> seg_ptr, segcol_ptr;  // line segments (host copies of the ranges)
  int* d_seg_ptr = nullptr;         // n_segments + 1
  int* d_seg_cells = nullptr;       // cells in segment order
  int* d_lpos = nullptr;            // per cell: position in d_seg_cells (column of its line blocks)
  char2* d_lslot = nullptr;         // per cell: slots of the previous / next cell of its segment (-1: none)
  // multi-cell segments packed into 64-slot bundles (one wavefront each), per color longest first
  int nlb = 0;                      // number of bundles
  std::vector<int> lbcol_ptr;       // per color: range of its bundles
  std::vector<int> lmcol_cells;     // per color: cells in its multi-cell segments (they lead its positions)
  int* d_lb_t = nullptr;            // per slot: line position of its cell (-1: padding)
  int* d_lb_kl = nullptr;           // per slot: place in the segment | segment length << 8
  int* d_lcomp = nullptr;           // per line position: bundle slot (-1: single-cell segment)
  unsigned char* d_lmask = nullptr; // per line position: off-segment neighbour slots of lower (bits 0-3) / higher
                                    // (4-7) color
  int4* d_lnc = nullptr;            // per line position: the line positions of the off-segment neighbours in the
                                    // packed slot order of Krylov::lineA (-1: none)
  char4* d_lqm = nullptr;           // per line position: the cell slot each packed slot takes its block from
  int line_fused_max = 262144;      // below this many cells a line sweep color is one launch
  int4* d_nq = nullptr;             // per cell slot: (neighbor << 2) | back slot, or -1
  // solver state
  double* grad = nullptr;   // 14 x pitch (rho u v p T gradients; k, omega with SST)
  double* phi = nullptr;    // pitch
  double* dt = nullptr;     // pitch
  double* Fm = nullptr;     // nv x F face buffer (assembly 1)
  double* red = nullptr;    // reduction partials
  double* scal = nullptr;   // device scalars (norms, eps, ...)
  double* scal_host = nullptr;  // pinned
  double* gradT = nullptr;  // tangent gradients (2 x 14 x pitch: values, derivatives)
  double* adj = nullptr;    // adjoints of the gradients (14 x pitch), of SST mu_t and F1 (2 x pitch)
  double* adjF = nullptr;   // adjoint face buffer (cfd::VJP_ROW_MAX x F), allocated by ensure_adjoint before the first reverse product
  double* tq = nullptr;     // SST per cell: mu_t, F1, V S_k, V S_w (values; + derivatives in the tangent)
  bool phi_valid = false, dt_valid = false;
  // host-pointer staging: AoS device copy + SoA vectors
  double* io6 = nullptr;     // 6 x N (AoS)
  double* vec[6] = {};       // 6 SoA vectors of nv x pitch
  double* aux = nullptr;     // one more SoA vector (6 x pitch), allocated by the unsteady driver
  cfd::Krylov kry;
  std::string err;
  int nv() const { return P.nvar; }
  size_t vbytes() const { return (size_t)P.nvar * pitch * sizeof(double); }
};

namespace cfd {
// Most blocks of a reduction's first stage (one partial sum per block and sum).
constexpr int RED_BLOCKS = 1024;
constexpr int GMRES_M_MAX = 250;  // largest Krylov dimension a GMRES cycle takes
// Doubles per face of the adjoint face buffer: two sides of a 16-double row
// (state and gradient adjoints) and of 2 mu_t / F1 adjoints (SST).
constexpr int VJP_ROW_MAX = 36;
// Per-device address of the library's static scratch array (noalloc.hip);
// called by cfd_ctx_create with the context's device current.
int noalloc_init(cfd_ctx* c);
// ---- launchers (kernels.hip) -------------------------------------------------
// All launch on ctx->stream.  Vectors are SoA (nv x pitch, RCM order).
void launch_residual(cfd_ctx* c, const double* U, double* R, bool update_limiters);
// same, recording ev[0..4] around: gradients | fill-or-nothing | flux | gather-or-nothing
void launch_residual_timed(cfd_ctx* c, const double* U, double* R, bool update_limiters, hipEvent_t* ev);
void launch_pack(cfd_ctx* c, const double* U6, double* U);
void launch_unpack(cfd_ctx* c, const double* U, double* U6);
void launch_timestep(cfd_ctx* c, const double* U, double cfl);
void launch_diag_precond(cfd_ctx* c, const double* U, double* D);
// reductions: result written to c->scal[slot] (device)
// RED_VW: sum (a_k / vol)^2 vol over all rows and sum vol (two results);
// RED_VW4: the same over rows 0-3 (rho, rho u, rho v, rho E)
enum RedKind { RED_L2 = 0, RED_DOT = 1, RED_VW = 2, RED_VW4 = 3 };
void launch_reduce(cfd_ctx* c, int kind, const double* a, const double* b, int slot);
// RED_VW of the FP64 SoA vector a with its first-stage partials in part
// (2 RED_BLOCKS doubles) and the two sums in sums[0..1] (the same grid and
// order as launch_reduce), then out[0] = sqrt(sums[0] / sums[1]) (0 if the
// volume sum is 0); nothing is allocated
void launch_norm_vw_into(cfd_ctx* c, const double* a, double* part, double* sums, double* out);
void launch_fill(cfd_ctx* c, double* x, double v, size_t n);
// FD Jv: Jv = (R(U + eps v) - Rb) / eps + (vol/dt) v, eps from scal[su] (|U|^2) and scal[sv] (|v|^2)
void launch_jv_fd(cfd_ctx* c, const double* U, const double* Rb, const double* v, double* Jv, int su, int sv,
                  double* tmpU, double* tmpR);
// forward tangent (limiter frozen): Jv = dR/dU v + (vol/dt) v; prec 64 or 32; v, Jv double or float
void launch_jv_tangent(cfd_ctx* c, const double* U, const void* v, void* Jv, int prec);
// reverse product (limiter frozen): JTw = (dR/dU)^T w + (vol/dt) w
// J^T w; grads_ready: the context's gradient rows already hold U's gradients
// (launch_vjp_gradients since the last kernel that wrote them), so the product
// does not form them again
void launch_vjp(cfd_ctx* c, const double* U, const double* w, double* JTw, bool grads_ready = false);
void launch_vjp_gradients(cfd_ctx* c, const double* U);
// vector kernels
void launch_axpby(cfd_ctx* c, double* y, double a, const double* x, double b, size_t n);
// w_k *= s_k for the nv rows of an SoA vector of the basis type
void launch_rowscale(cfd_ctx* c, void* w, const double* s, int prec);
void launch_scale_by_scal(cfd_ctx* c, double* y, const double* x, int slot, int mode);
void launch_newton_update(cfd_ctx* c, double* U, const double* dU, double relax);
void launch_explicit(cfd_ctx* c, double* Unew, const double* U0, double a0, const double* U1, double a1,
                     const double* R, double dtc);
void launch_wall(cfd_ctx* c, const double* U, double* out3);
// entropy-error sums (sum e^2 V, sum V) into scal[slot], scal[slot + 1]
void launch_entropy(cfd_ctx* c, const double* U, double S_ref, int slot);
// boundary-layer profile rows (4 x pitch) and per-cell selection flags
void launch_bl_profile(cfd_ctx* c, const double* U, double x_target, double u_inf, double nu, double* out4,
                       unsigned char* flag);
// rhs = -((U - Uold) vol / dtp + R)
void launch_unsteady_rhs(cfd_ctx* c, const double* U, const double* Uold, const double* R, double dtp,
                         double* rhs);
// wall distance of every cell to the nw no-slip face midpoints (xy interleaved)
void launch_wall_distance(cfd_ctx* c, const double* wxy, int nw, double* wd);
// Krylov (precision-generic basis)
void launch_krylov_init(cfd_ctx* c, void* V0, const double* rhs, double sign, int slot, int prec);
void launch_precond_apply(cfd_ctx* c, void* z, const void* Vj, int prec, int slot_norm);
void launch_d2f(cfd_ctx* c, const double* a, float* b, size_t n);
void launch_f2d(cfd_ctx* c, const float* a, double* b, size_t n);
// Arnoldi step j with delayed reorthogonalisation (or, with last, only the
// reduction that finishes column j-1); column j-1 of H goes to hout
void launch_dcgs(cfd_ctx* c, void* V, int j, bool last, double* hout, int prec);
// x = D^-1 sum_{i<=k} y_i V_i (diag) or the plain combination
void launch_solution(cfd_ctx* c, double* x, const void* V, const double* y, int k, int prec, bool diag = true);
// block-Jacobi: assemble + invert (Binv: nv*nv x pitch of the basis type), apply, solution
// Binv, Bout and Aoff are stored in column pos[cell] (pos = nullptr: the cell itself); the block
// SGS passes d_cpos (color-major), the line preconditioner d_lpos (segment order)
void launch_block_precond(cfd_ctx* c, const double* U, void* Binv, int prec, double* Bout = nullptr,
                          void* Aoff = nullptr, const int* pos = nullptr);
// line preconditioner: factorisation of the segments (Bd: diagonal blocks in FP64, segment order;
// Aoff: the coupling blocks by cell slot; Ac receives the off-segment ones in packed slots) and one
// symmetric line Gauss-Seidel application z = M^-1 x (same precision rules as the block SGS)
void launch_line_factor(cfd_ctx* c, const double* Bd, const void* Aoff, void* Dpinv, void* Pm, void* G, void* Ac,
                        int prec);
void launch_line_apply(cfd_ctx* c, const void* Dpinv, const void* Pm, const void* G, const void* Ac, const void* x,
                       int xprec, void* z, int zprec, int bprec);
// multicolor block SGS: z = M^-1 x (M^-T with transpose); blocks of precision bprec, x/z of xprec/zprec
// (bprec 64 needs FP64 x and z; bprec 32 takes FP32 x and z or FP64 x and z)
void launch_sgs_apply(cfd_ctx* c, const void* Binv, const void* Aoff, const void* x, int xprec, void* z,
                      int zprec, int bprec, bool transpose);
void launch_block_apply(cfd_ctx* c, const void* Binv, const void* x, void* z, int prec, bool transpose = false);
void launch_solution_bj(cfd_ctx* c, double* x, const void* V, const double* y, int k, const void* Binv, int prec,
                        bool transpose = false);
}  // namespace cfd

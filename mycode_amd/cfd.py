"""ctypes binding of ``libcfdgpu.so`` (include/cfd_gpu.h).

``Solver`` mirrors the reference's ``Solver`` class (include/Solver.hpp): the
method names follow the reference members they replace, arrays are
``numpy`` float64 of shape (N, 6) in reference cell order, exactly the memory
of ``std::vector<State>``.  Every call goes to the HIP library; there is no
CPU fallback — if the library is missing or no GPU is present the call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcfdgpu.so")

# enum values (Types.hpp:38-49 and include/cfd_gpu.h)
LLF, ROE, HLLC, AUSM = 0, 1, 2, 3
FIRST_ORDER, SECOND_ORDER_CENTRAL, SECOND_ORDER_LIMITED = 0, 1, 2
BC_BUMP, BC_SUPERSONIC, BC_PLATE, BC_SWBLI = 0, 1, 2, 3


class CfdError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("R_gas", C.c_double), ("M_inf", C.c_double),
        ("P_total_inlet", C.c_double), ("T_total_inlet", C.c_double),
        ("P_static_exit", C.c_double), ("nu", C.c_double), ("Pr", C.c_double),
        ("flux", C.c_int), ("spatial", C.c_int), ("bcset", C.c_int), ("viscous", C.c_int),
        ("nvar", C.c_int), ("turb", C.c_int), ("flux_alt", C.c_int), ("turb_first_order", C.c_int),
        ("shock_rho", C.c_double), ("shock_u", C.c_double), ("shock_v", C.c_double),
        ("shock_p", C.c_double), ("turb_k_inf", C.c_double), ("turb_w_inf", C.c_double),
        ("reserved", C.c_double * 6),
    ]

    def copy(self) -> "Params":
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int), ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
        ("n_cells", C.c_int), ("cell_ptr", C.POINTER(C.c_int)), ("cell_nodes", C.POINTER(C.c_int)),
        ("n_bedges", C.c_int), ("bedges", C.POINTER(C.c_int)),
    ]


class GmresOpts(C.Structure):
    _fields_ = [("m", C.c_int), ("restarts", C.c_int), ("tol", C.c_double), ("jv", C.c_int),
                ("precision", C.c_int), ("precond", C.c_int), ("scaling", C.c_int)]


class NewtonOpts(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("rel_drop", C.c_double), ("max_steps", C.c_int),
                ("log", C.c_int), ("cfl0", C.c_double), ("gmres", GmresOpts), ("controller", C.c_int),
                ("pad", C.c_int), ("cfl_max", C.c_double)]


class NewtonResult(C.Structure):
    _fields_ = [("steps", C.c_int), ("converged", C.c_int), ("res_vm0", C.c_double),
                ("res_vm", C.c_double), ("seconds", C.c_double), ("total_gmres_iters", C.c_int),
                ("pad", C.c_int)]


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
_VP = C.c_void_p

# (name, restype, argtypes) of every exported function in include/cfd_gpu.h
SIGNATURES = [
    ("cfd_params_default", C.c_int, [C.POINTER(Params)]),
    ("cfd_params_flow_conditions", C.c_int, [C.POINTER(Params), C.c_double, C.c_double, C.c_double]),
    ("cfd_params_freestream_static", C.c_int, [C.POINTER(Params), C.c_double, C.c_double, C.c_double]),
    ("cfd_params_kinematic_viscosity", C.c_int, [C.POINTER(Params), C.c_double, C.c_double]),
    ("cfd_params_initial_state", C.c_int, [C.POINTER(Params), _DP]),
    ("cfd_mesh_build", C.c_int, [C.POINTER(MeshDesc), C.POINTER(_VP)]),
    ("cfd_mesh_read_gmsh", C.c_int, [C.c_char_p, C.POINTER(_VP)]),
    ("cfd_mesh_generate", C.c_int, [C.c_int, C.c_double, C.POINTER(_VP)]),
    ("cfd_mesh_destroy", C.c_int, [_VP]),
    ("cfd_mesh_sizes", C.c_int, [_VP, _IP, _IP, _IP, _IP]),
    ("cfd_mesh_get_desc", C.c_int, [_VP, C.POINTER(MeshDesc)]),
    ("cfd_mesh_get_faces", C.c_int, [_VP, _IP, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _DP, _DP]),
    ("cfd_mesh_get_cells", C.c_int, [_VP, _DP, _DP, _DP]),
    ("cfd_mesh_get_reorder", C.c_int, [_VP, _IP, _IP]),
    ("cfd_mesh_hash", C.c_uint64, [_VP]),
    ("cfd_mesh_write_gmsh", C.c_int, [_VP, C.c_char_p]),
    ("cfd_mesh_set_nodes", C.c_int, [_VP, _DP, _DP]),
    ("cfd_ctx_update_geometry", C.c_int, [_VP, _VP]),
    ("cfd_ctx_create", C.c_int, [_VP, C.POINTER(Params), C.POINTER(_VP)]),
    ("cfd_ctx_destroy", C.c_int, [_VP]),
    ("cfd_last_error", C.c_char_p, [_VP]),
    ("cfd_ctx_set_params", C.c_int, [_VP, C.POINTER(Params)]),
    ("cfd_ctx_get_params", C.c_int, [_VP, C.POINTER(Params)]),
    ("cfd_pitch", C.c_int, [_VP]),
    ("cfd_ncells", C.c_int, [_VP]),
    ("cfd_stream", C.c_void_p, [_VP]),
    ("cfd_ctx_own_stream", C.c_void_p, [_VP]),
    ("cfd_ctx_set_stream", C.c_int, [_VP, C.c_void_p]),
    ("cfd_drag", C.c_int, [_VP, _DP, _DP]),
    ("cfd_wall_pressure", C.c_int, [_VP, _DP, C.c_double, C.c_double, C.c_double, C.c_int, _DP, _IP]),
    ("cfd_set_face_flux", C.c_int, [_VP, C.POINTER(C.c_ubyte)]),
    ("cfd_set_assembly", C.c_int, [_VP, C.c_int]),
    ("cfd_residual", C.c_int, [_VP, _DP, _DP, C.c_int]),
    ("cfd_get_gradients", C.c_int, [_VP, _DP]),
    ("cfd_get_limiters", C.c_int, [_VP, _DP]),
    ("cfd_set_limiters", C.c_int, [_VP, _DP]),
    ("cfd_local_timestep", C.c_int, [_VP, _DP, C.c_double]),
    ("cfd_get_dt", C.c_int, [_VP, _DP]),
    ("cfd_diag_preconditioner", C.c_int, [_VP, _DP, _DP]),
    ("cfd_jv_fd", C.c_int, [_VP, _DP, _DP, _DP, _DP]),
    ("cfd_jv_tangent", C.c_int, [_VP, _DP, _DP, _DP, C.c_int]),
    ("cfd_vjp_adjoint", C.c_int, [_VP, _DP, _DP, _DP]),
    ("cfd_block_jacobi", C.c_int, [_VP, _DP, C.c_int, _DP, _DP]),
    ("cfd_block_sgs", C.c_int, [_VP, _DP, C.c_int, C.c_int, _DP, _DP, _DP, _IP, _IP]),
    ("cfd_dev_gmres", C.c_int, [_VP, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(GmresOpts), _IP]),
    ("cfd_entropy_error", C.c_int, [_VP, _DP, _DP]),
    ("cfd_bl_profile", C.c_int, [_VP, _DP, C.c_double, C.c_int, _DP, _IP]),
    ("cfd_mesh_write_vtk", C.c_int, [_VP, _DP, C.c_double, C.c_char_p]),
    ("cfd_mesh_get_lines", C.c_int, [_VP, _IP, _IP, _IP, _IP, _IP]),
    ("cfd_mesh_get_cell_colors", C.c_int, [_VP, _IP, _IP]),
    ("cfd_line_precond", C.c_int, [_VP, _DP, C.c_int, _DP, _DP]),
    ("cfd_solve_unsteady", C.c_int, [_VP, _DP, C.c_double, C.c_int, C.POINTER(GmresOpts), C.c_int, C.c_double,
                                     C.c_double, _DP, _IP]),
    ("cfd_norm_l2", C.c_int, [_VP, _DP, _DP]),
    ("cfd_dot", C.c_int, [_VP, _DP, _DP, _DP]),
    ("cfd_norm_vw", C.c_int, [_VP, _DP, _DP]),
    ("cfd_gmres_opts_default", C.c_int, [C.POINTER(GmresOpts)]),
    ("cfd_gmres", C.c_int, [_VP, _DP, _DP, _DP, _DP, C.POINTER(GmresOpts), _IP]),
    ("cfd_newton_opts_default", C.c_int, [C.POINTER(NewtonOpts)]),
    ("cfd_solve_implicit", C.c_int, [_VP, _DP, C.POINTER(NewtonOpts), C.POINTER(NewtonResult), _DP, _IP, _DP]),
    ("cfd_explicit_step", C.c_int, [_VP, _DP, C.c_double]),
    ("cfd_rk3_step", C.c_int, [_VP, _DP, C.c_double]),
    ("cfd_wall_outputs", C.c_int, [_VP, _DP, _DP, _DP, _DP]),
    ("cfd_dev_pack", C.c_int, [_VP, _VP, _VP]),
    ("cfd_dev_unpack", C.c_int, [_VP, _VP, _VP]),
    ("cfd_dev_residual", C.c_int, [_VP, _VP, _VP, C.c_int]),
    ("cfd_dev_jv_fd", C.c_int, [_VP, _VP, _VP, _VP, _VP]),
    ("cfd_dev_jv_tangent", C.c_int, [_VP, _VP, _VP, _VP, C.c_int]),
    ("cfd_dev_local_timestep", C.c_int, [_VP, _VP, C.c_double]),
    ("cfd_dev_vjp_adjoint", C.c_int, [_VP, _VP, _VP, _VP]),
    ("cfd_dev_synchronize", C.c_int, [_VP]),
    ("cfd_dev_norm_vw_noalloc", C.c_int, [_VP, _VP, _VP]),
    ("cfd_dev_time_residual", C.c_int, [_VP, _VP, _VP, C.c_int, C.c_int, _DP, _IP]),
]

_lib = None


def load_library(path: str = LIB_PATH):
    """Load the HIP library; raises if it has not been built.

    PyTorch ships its own HIP runtime.  When torch is installed it is imported
    first, so that the library binds to the runtime torch loads and device
    pointers from torch tensors are valid in both (loading the library first
    leaves torch without a usable device)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise CfdError(f"{path} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


def _check(rc, ctx=None, what=""):
    if rc != 0:
        msg = ""
        if ctx is not None:
            m = load_library().cfd_last_error(ctx)
            msg = m.decode() if m else ""
        raise CfdError(f"{what} failed with status {rc}: {msg}")


# ---- parameters ----------------------------------------------------------------
def default_params() -> Params:
    p = Params()
    _check(load_library().cfd_params_default(C.byref(p)), what="cfd_params_default")
    return p


def flow_conditions(p: Params, M, P, T):
    _check(load_library().cfd_params_flow_conditions(C.byref(p), M, P, T))


def kinematic_viscosity(p: Params, Re, L):
    _check(load_library().cfd_params_kinematic_viscosity(C.byref(p), Re, L))


def initial_state(p: Params) -> np.ndarray:
    u = np.zeros(6)
    _check(load_library().cfd_params_initial_state(C.byref(p), _dp(u)))
    return u


# ---- mesh --------------------------------------------------------------------------
class Mesh:
    """Mesh::load equivalent: connectivity + geometry + RCM + edge colors (host)."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self.lib = load_library()
        nn, nc, nf, ncol = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _check(self.lib.cfd_mesh_sizes(self._h, C.byref(nn), C.byref(nc), C.byref(nf), C.byref(ncol)))
        self.n_nodes, self.n_cells, self.n_faces, self.n_colors = nn.value, nc.value, nf.value, ncol.value

    @classmethod
    def generate(cls, config: int, scale: float = 1.0) -> "Mesh":
        h = C.c_void_p()
        _check(load_library().cfd_mesh_generate(config, scale, C.byref(h)), what="cfd_mesh_generate")
        m = cls(h)
        m.config = config
        return m

    @classmethod
    def read_gmsh(cls, path: str) -> "Mesh":
        h = C.c_void_p()
        _check(load_library().cfd_mesh_read_gmsh(path.encode(), C.byref(h)), what="cfd_mesh_read_gmsh")
        return cls(h)

    @classmethod
    def from_arrays(cls, x, y, cell_ptr, cell_nodes, bedges) -> "Mesh":
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        cp = np.ascontiguousarray(cell_ptr, np.int32)
        cn = np.ascontiguousarray(cell_nodes, np.int32)
        be = np.ascontiguousarray(bedges, np.int32).reshape(-1)
        d = MeshDesc(len(x) - 1, _dp(x), _dp(y), len(cp) - 1, _ip(cp), _ip(cn), len(be) // 3, _ip(be))
        h = C.c_void_p()
        _check(load_library().cfd_mesh_build(C.byref(d), C.byref(h)), what="cfd_mesh_build")
        return cls(h)

    def __del__(self):
        try:
            if self._h:
                self.lib.cfd_mesh_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def desc(self):
        """(x, y, cell_ptr, cell_nodes, bedges) copies of the description."""
        d = MeshDesc()
        _check(self.lib.cfd_mesh_get_desc(self._h, C.byref(d)))
        x = np.ctypeslib.as_array(d.x, shape=(d.n_nodes + 1,)).copy()
        y = np.ctypeslib.as_array(d.y, shape=(d.n_nodes + 1,)).copy()
        cp = np.ctypeslib.as_array(d.cell_ptr, shape=(d.n_cells + 1,)).copy()
        cn = np.ctypeslib.as_array(d.cell_nodes, shape=(int(cp[-1]),)).copy()
        be = (np.ctypeslib.as_array(d.bedges, shape=(3 * d.n_bedges,)).copy().reshape(-1, 3)
              if d.n_bedges else np.zeros((0, 3), np.int32))
        return x, y, cp, cn, be

    def faces(self) -> dict:
        F = self.n_faces
        ints = {k: np.zeros(F, np.int32) for k in ("node1", "node2", "owner", "neighbor", "bc_tag")}
        dbl = {k: np.zeros(F) for k in ("nx", "ny", "area", "x_mid", "y_mid")}
        _check(self.lib.cfd_mesh_get_faces(self._h, *[_ip(ints[k]) for k in ints], *[_dp(dbl[k]) for k in dbl]))
        return {**ints, **dbl}

    def cells(self) -> dict:
        N = self.n_cells
        out = {k: np.zeros(N) for k in ("volume", "x_c", "y_c")}
        _check(self.lib.cfd_mesh_get_cells(self._h, _dp(out["volume"]), _dp(out["x_c"]), _dp(out["y_c"])))
        return out

    def reorder(self):
        perm = np.zeros(self.n_cells, np.int32)
        color = np.zeros(self.n_faces, np.int32)
        _check(self.lib.cfd_mesh_get_reorder(self._h, _ip(perm), _ip(color)))
        return perm, color

    def hash(self) -> int:
        return int(self.lib.cfd_mesh_hash(self._h))

    def set_nodes(self, x, y):
        """Move the nodes (arrays of n_nodes + 1, index 0 unused); geometry is recomputed."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        if x.shape != (self.n_nodes + 1,) or y.shape != (self.n_nodes + 1,):
            raise ValueError("node arrays must have n_nodes + 1 entries")
        _check(self.lib.cfd_mesh_set_nodes(self._h, _dp(x), _dp(y)), what="cfd_mesh_set_nodes")

    def write_gmsh(self, path: str):
        _check(self.lib.cfd_mesh_write_gmsh(self._h, path.encode()))

    def lines(self):
        """(seg_ptr, seg_cells, segcol_ptr) of the line-implicit preconditioner's segments."""
        ns, nc = C.c_int(), C.c_int()
        _check(self.lib.cfd_mesh_get_lines(self._h, C.byref(ns), C.byref(nc), None, None, None))
        sp = np.zeros(ns.value + 1, np.int32)
        sc = np.zeros(self.n_cells, np.int32)
        cp = np.zeros(nc.value + 1, np.int32)
        _check(self.lib.cfd_mesh_get_lines(self._h, C.byref(ns), C.byref(nc), _ip(sp), _ip(sc), _ip(cp)))
        return sp, sc, cp

    def cell_colors(self):
        """(n_colors, color per cell) of the block SGS preconditioner's cell coloring."""
        nc = C.c_int()
        color = np.zeros(self.n_cells, np.int32)
        _check(self.lib.cfd_mesh_get_cell_colors(self._h, C.byref(nc), _ip(color)))
        return nc.value, color

    def write_vtk(self, U, path: str, gamma: float = 1.4):
        """main.cpp saveResults: legacy VTK with Mach, Pressure, Density cell fields."""
        U = np.ascontiguousarray(U, np.float64)
        if U.shape != (self.n_cells, 6):
            raise ValueError(f"state must have shape ({self.n_cells}, 6)")
        _check(self.lib.cfd_mesh_write_vtk(self._h, _dp(U), gamma, path.encode()))


# ---- solver ------------------------------------------------------------------------
def _state(a, N):
    a = np.ascontiguousarray(a, np.float64)
    if a.shape != (N, 6):
        raise ValueError(f"state arrays must have shape ({N}, 6), got {a.shape}")
    return a


class Solver:
    """The reference Solver's residual / Jacobian / Krylov members on the GPU."""

    def __init__(self, mesh: Mesh, params: Params, assembly: int = 1):
        self.lib = load_library()
        self.mesh = mesh
        self.N = mesh.n_cells
        self._h = C.c_void_p()
        rc = self.lib.cfd_ctx_create(mesh.handle, C.byref(params), C.byref(self._h))
        if rc != 0:
            msg = self.lib.cfd_last_error(self._h) if self._h else b""
            if self._h:
                self.lib.cfd_ctx_destroy(self._h)
            self._h = None
            raise CfdError(f"cfd_ctx_create failed ({rc}): {msg.decode() if msg else ''}")
        self.set_assembly(assembly)

    def __del__(self):
        try:
            if self._h:
                self.lib.cfd_ctx_destroy(self._h)
                self._h = None
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def _c(self, rc, what):
        _check(rc, self._h, what)

    def set_params(self, p: Params):
        self._c(self.lib.cfd_ctx_set_params(self._h, C.byref(p)), "cfd_ctx_set_params")

    def params(self) -> Params:
        p = Params()
        self._c(self.lib.cfd_ctx_get_params(self._h, C.byref(p)), "cfd_ctx_get_params")
        return p

    def update_geometry(self, mesh: "Mesh"):
        """Re-upload geometry after mesh.set_nodes (same topology)."""
        self._c(self.lib.cfd_ctx_update_geometry(self._h, mesh.handle), "cfd_ctx_update_geometry")

    def set_assembly(self, mode: int):
        self._c(self.lib.cfd_set_assembly(self._h, mode), "cfd_set_assembly")

    def set_face_flux(self, use_alt):
        a = np.ascontiguousarray(use_alt, np.uint8)
        self._c(self.lib.cfd_set_face_flux(self._h, a.ctypes.data_as(C.POINTER(C.c_ubyte))), "cfd_set_face_flux")

    # Solver::computeResiduals
    def compute_residuals(self, U, update_limiters: bool = True) -> np.ndarray:
        U = _state(U, self.N)
        R = np.zeros((self.N, 6))
        self._c(self.lib.cfd_residual(self._h, _dp(U), _dp(R), int(update_limiters)), "cfd_residual")
        return R

    def gradients(self) -> np.ndarray:
        g = np.zeros((self.N, 10))
        self._c(self.lib.cfd_get_gradients(self._h, _dp(g)), "cfd_get_gradients")
        return g

    def limiters(self) -> np.ndarray:
        p = np.zeros(self.N)
        self._c(self.lib.cfd_get_limiters(self._h, _dp(p)), "cfd_get_limiters")
        return p

    def set_limiters(self, phi):
        phi = np.ascontiguousarray(phi, np.float64)
        self._c(self.lib.cfd_set_limiters(self._h, _dp(phi)), "cfd_set_limiters")

    # Solver::computeLocalTimeStep (with U_current = U)
    def compute_local_time_step(self, U, cfl: float):
        U = _state(U, self.N)
        self._c(self.lib.cfd_local_timestep(self._h, _dp(U), cfl), "cfd_local_timestep")

    def dt_local(self) -> np.ndarray:
        d = np.zeros(self.N)
        self._c(self.lib.cfd_get_dt(self._h, _dp(d)), "cfd_get_dt")
        return d

    # Solver::computeDiagonalPreconditioner
    def diagonal_preconditioner(self, U) -> np.ndarray:
        U = _state(U, self.N)
        D = np.zeros(self.N)
        self._c(self.lib.cfd_diag_preconditioner(self._h, _dp(U), _dp(D)), "cfd_diag_preconditioner")
        return D

    # Solver::applyJacobian
    def apply_jacobian(self, U, R, v) -> np.ndarray:
        U, R, v = (_state(a, self.N) for a in (U, R, v))
        J = np.zeros((self.N, 6))
        self._c(self.lib.cfd_jv_fd(self._h, _dp(U), _dp(R), _dp(v), _dp(J)), "cfd_jv_fd")
        return J

    def jv_tangent(self, U, v, precision: int = 64) -> np.ndarray:
        U, v = (_state(a, self.N) for a in (U, v))
        J = np.zeros((self.N, 6))
        self._c(self.lib.cfd_jv_tangent(self._h, _dp(U), _dp(v), _dp(J), precision), "cfd_jv_tangent")
        return J

    def vjp_adjoint(self, U, w) -> np.ndarray:
        U, w = (_state(a, self.N) for a in (U, w))
        J = np.zeros((self.N, 6))
        self._c(self.lib.cfd_vjp_adjoint(self._h, _dp(U), _dp(w), _dp(J)), "cfd_vjp_adjoint")
        return J

    def block_jacobi(self, U, precision: int = 64):
        """(B, Binv): per-cell nv x nv blocks of the block-Jacobi preconditioner."""
        U = _state(U, self.N)
        nv = self.params().nvar
        B = np.zeros((self.N, nv, nv))
        Bi = np.zeros((self.N, nv, nv))
        self._c(self.lib.cfd_block_jacobi(self._h, _dp(U), precision, _dp(B), _dp(Bi)), "cfd_block_jacobi")
        return B, Bi

    def block_sgs(self, U, x=None, precision: int = 64, transpose: bool = False, blocks: bool = False):
        """Multicolor block SGS preconditioner at U.

        Returns (z, Aoff, nbr, color): z = M^-1 x (M^-T x with transpose) or
        None; with blocks=True the coupling blocks (N x 4 x nv x nv, slots in
        ascending face index), neighbor cells (N x 4, -1 for none) and the cell
        colors (N), else None for each."""
        U = _state(U, self.N)
        nv = self.params().nvar
        xz = _state(x, self.N) if x is not None else None
        z = np.zeros((self.N, 6)) if x is not None else None
        A = np.zeros((self.N, 4, nv, nv)) if blocks else None
        nb = np.zeros((self.N, 4), dtype=np.int32) if blocks else None
        col = np.zeros(self.N, dtype=np.int32) if blocks else None
        dp = lambda a: _dp(a) if a is not None else None
        ip = lambda a: _ip(a) if a is not None else None
        self._c(self.lib.cfd_block_sgs(self._h, _dp(U), precision, int(transpose), dp(xz), dp(z), dp(A),
                                       ip(nb), ip(col)), "cfd_block_sgs")
        return z, A, nb, col

    def line_precond(self, U, x, precision: int = 64) -> np.ndarray:
        """z = M^-1 x of the line-implicit preconditioner at U."""
        U, x = _state(U, self.N), _state(x, self.N)
        z = np.zeros((self.N, 6))
        self._c(self.lib.cfd_line_precond(self._h, _dp(U), precision, _dp(x), _dp(z)), "cfd_line_precond")
        return z

    def norm_L2(self, v) -> float:
        v = _state(v, self.N)
        out = C.c_double()
        self._c(self.lib.cfd_norm_l2(self._h, _dp(v), C.byref(out)), "cfd_norm_l2")
        return out.value

    def dot_product(self, a, b) -> float:
        a, b = _state(a, self.N), _state(b, self.N)
        out = C.c_double()
        self._c(self.lib.cfd_dot(self._h, _dp(a), _dp(b), C.byref(out)), "cfd_dot")
        return out.value

    def norm_L2_volume_weighted(self, R) -> float:
        R = _state(R, self.N)
        out = C.c_double()
        self._c(self.lib.cfd_norm_vw(self._h, _dp(R), C.byref(out)), "cfd_norm_vw")
        return out.value

    @staticmethod
    def gmres_opts(**kw) -> GmresOpts:
        o = GmresOpts()
        load_library().cfd_gmres_opts_default(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    @staticmethod
    def newton_opts(gmres: GmresOpts | None = None, **kw) -> NewtonOpts:
        o = NewtonOpts()
        load_library().cfd_newton_opts_default(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        if gmres is not None:
            o.gmres = gmres
        return o

    # Solver::solveGMRES
    def solve_gmres(self, U, R, RHS, opts: GmresOpts | None = None):
        U, R, RHS = (_state(a, self.N) for a in (U, R, RHS))
        x = np.zeros((self.N, 6))
        it = C.c_int()
        o = opts if opts is not None else self.gmres_opts()
        self._c(self.lib.cfd_gmres(self._h, _dp(U), _dp(R), _dp(RHS), _dp(x), C.byref(o), C.byref(it)),
                "cfd_gmres")
        return x, it.value

    # Solver::solveImplicit
    def solve_implicit(self, U, opts: NewtonOpts | None = None):
        U = _state(U, self.N).copy()
        o = opts if opts is not None else self.newton_opts()
        n = max(1, o.max_steps)
        hvm, hit, hcfl = np.zeros(n), np.zeros(n, np.int32), np.zeros(n)
        res = NewtonResult()
        self._c(self.lib.cfd_solve_implicit(self._h, _dp(U), C.byref(o), C.byref(res), _dp(hvm), _ip(hit), _dp(hcfl)),
                "cfd_solve_implicit")
        k = res.steps + (1 if res.converged else 0)
        return U, res, {"vm": hvm[:k], "iters": hit[:k], "cfl": hcfl[:k]}

    # Solver::timeStepExplicit
    def time_step_explicit(self, U, dt: float) -> np.ndarray:
        U = _state(U, self.N).copy()
        self._c(self.lib.cfd_explicit_step(self._h, _dp(U), dt), "cfd_explicit_step")
        return U

    def rk3_step(self, U, dt: float) -> np.ndarray:
        U = _state(U, self.N).copy()
        self._c(self.lib.cfd_rk3_step(self._h, _dp(U), dt), "cfd_rk3_step")
        return U

    def wall_outputs(self, U):
        U = _state(U, self.N)
        F = self.mesh.n_faces
        Cp, Cf, J = np.zeros(F), np.zeros(F), C.c_double()
        self._c(self.lib.cfd_wall_outputs(self._h, _dp(U), _dp(Cp), _dp(Cf), C.byref(J)), "cfd_wall_outputs")
        return Cp, Cf, J.value

    def calculate_drag(self, U) -> float:
        """Solver::calculateDrag: sum over tag-3 faces of -p_owner A n_x."""
        out = C.c_double()
        self._c(self.lib.cfd_drag(self._h, _dp(_state(U, self.N)), C.byref(out)), "cfd_drag")
        return out.value

    def extract_wall_pressure(self, U, y_max: float = 0.2, x_scale: float = 12.0, p_divisor: float = 144.0):
        """Solver::extractWallPressure rows (x_mid * x_scale, p_owner / p_divisor)
        of the tag-3 faces with owner y_c < y_max, ascending in x."""
        U = _state(U, self.N)
        n = C.c_int()
        cap = self.mesh.n_faces
        rows = np.zeros((max(cap, 1), 2))
        self._c(self.lib.cfd_wall_pressure(self._h, _dp(U), y_max, x_scale, p_divisor, cap, _dp(rows), C.byref(n)),
                "cfd_wall_pressure")
        return rows[:n.value]

    # device-handle variants (raw device pointers, e.g. torch tensors' data_ptr()).
    # They run on the context's stream (its own unless set_stream chose
    # another): torch work that produces or clears a buffer passed here must
    # be ordered before the call, either by running the context on torch's
    # current stream (set_stream(torch.cuda.current_stream().cuda_stream)) or
    # by torch.cuda.synchronize().
    def dev_residual(self, U_ptr: int, R_ptr: int, update_limiters: bool = True):
        self._c(self.lib.cfd_dev_residual(self._h, C.c_void_p(U_ptr), C.c_void_p(R_ptr), int(update_limiters)),
                "cfd_dev_residual")

    def dev_pack(self, U6_ptr: int, U_ptr: int):
        self._c(self.lib.cfd_dev_pack(self._h, C.c_void_p(U6_ptr), C.c_void_p(U_ptr)), "cfd_dev_pack")

    def dev_unpack(self, U_ptr: int, U6_ptr: int):
        self._c(self.lib.cfd_dev_unpack(self._h, C.c_void_p(U_ptr), C.c_void_p(U6_ptr)), "cfd_dev_unpack")

    def dev_local_timestep(self, U_ptr: int, cfl: float):
        self._c(self.lib.cfd_dev_local_timestep(self._h, C.c_void_p(U_ptr), cfl), "cfd_dev_local_timestep")

    def dev_jv_fd(self, U_ptr, Rb_ptr, v_ptr, Jv_ptr):
        self._c(self.lib.cfd_dev_jv_fd(self._h, C.c_void_p(U_ptr), C.c_void_p(Rb_ptr), C.c_void_p(v_ptr),
                                       C.c_void_p(Jv_ptr)), "cfd_dev_jv_fd")

    def dev_jv_tangent(self, U_ptr, v_ptr, Jv_ptr, precision: int = 64):
        """Forward tangent J v on device SoA vectors (v, Jv FP64 or FP32 by precision)."""
        self._c(self.lib.cfd_dev_jv_tangent(self._h, C.c_void_p(U_ptr), C.c_void_p(v_ptr), C.c_void_p(Jv_ptr),
                                            int(precision)), "cfd_dev_jv_tangent")

    def dev_vjp_adjoint(self, U_ptr, w_ptr, JTw_ptr):
        """Reverse product J^T w (with the time term) on device SoA FP64 vectors."""
        self._c(self.lib.cfd_dev_vjp_adjoint(self._h, C.c_void_p(U_ptr), C.c_void_p(w_ptr), C.c_void_p(JTw_ptr)),
                "cfd_dev_vjp_adjoint")

    def dev_time_residual(self, U_ptr: int, R_ptr: int, reps: int, update_limiters: bool = False):
        """Average ms per residual of (gradients, face flux, assembly, total),
        timed with HIP events on the context stream, and flux launches per residual."""
        ms = np.zeros(4)
        nl = C.c_int()
        self._c(self.lib.cfd_dev_time_residual(self._h, C.c_void_p(U_ptr), C.c_void_p(R_ptr), reps,
                                               int(update_limiters), _dp(ms), C.byref(nl)), "cfd_dev_time_residual")
        return ms, nl.value

    def calculate_entropy_error(self, U) -> float:
        """Solver::calculateEntropyError: volume-weighted L2 entropy error."""
        out = C.c_double()
        self._c(self.lib.cfd_entropy_error(self._h, _dp(_state(U, self.N)), C.byref(out)), "cfd_entropy_error")
        return out.value

    def extract_boundary_layer_profile(self, U, x_target: float) -> np.ndarray:
        """Solver::extractBoundaryLayerProfile rows (y, u/u_inf, v/u_inf, eta), reference cell order."""
        U = _state(U, self.N)
        n = C.c_int()
        rows = np.zeros((self.N, 4))
        self._c(self.lib.cfd_bl_profile(self._h, _dp(U), x_target, self.N, _dp(rows), C.byref(n)), "cfd_bl_profile")
        return rows[:n.value]

    def solve_unsteady(self, U, dt_phys: float, n_steps: int, gmres: GmresOpts | None = None, max_newton: int = 15,
                       tolerance: float = 1e-6, relax: float = 0.7):
        """Solver::solveUnsteady (BDF1 dual time stepping): (U, VM||F|| per Newton
        step (n_steps x max_newton, zero after the stop), updates per step)."""
        U = _state(U, self.N).copy()
        if max_newton <= 0:
            max_newton = 15  # the reference's default (Solver.cpp:1819-1820)
        hvm = np.zeros((max(1, n_steps), max_newton))
        hn = np.zeros(max(1, n_steps), dtype=np.int32)
        g = C.byref(gmres) if gmres is not None else None
        self._c(self.lib.cfd_solve_unsteady(self._h, _dp(U), dt_phys, n_steps, g, max_newton, tolerance, relax,
                                            _dp(hvm), _ip(hn)), "cfd_solve_unsteady")
        return U, hvm[:n_steps], hn[:n_steps]

    def dev_gmres(self, U_ptr, R_ptr, rhs_ptr, x_ptr, opts: GmresOpts) -> int:
        """GMRES on device SoA vectors; returns the iteration count."""
        it = C.c_int()
        self._c(self.lib.cfd_dev_gmres(self._h, C.c_void_p(U_ptr), C.c_void_p(R_ptr), C.c_void_p(rhs_ptr),
                                       C.c_void_p(x_ptr), C.byref(opts), C.byref(it)), "cfd_dev_gmres")
        return it.value

    def dev_norm_vw_noalloc(self, R_ptr: int, out_ptr: int):
        """Volume-weighted norm of a device SoA vector into out_ptr[0] (device),
        with no allocation (static per-GPU scratch shared by all calls)."""
        self._c(self.lib.cfd_dev_norm_vw_noalloc(self._h, C.c_void_p(R_ptr), C.c_void_p(out_ptr)),
                "cfd_dev_norm_vw_noalloc")

    def synchronize(self):
        self._c(self.lib.cfd_dev_synchronize(self._h), "cfd_dev_synchronize")

    @property
    def pitch(self) -> int:
        return int(self.lib.cfd_pitch(self._h))

    @property
    def stream(self) -> int:
        return int(self.lib.cfd_stream(self._h) or 0)

    def set_stream(self, stream: int | None):
        """Launch on the given hipStream_t handle (e.g. torch's
        current_stream().cuda_stream); None returns to the context's own."""
        h = self.lib.cfd_ctx_own_stream(self._h) if stream is None else stream
        self._c(self.lib.cfd_ctx_set_stream(self._h, C.c_void_p(h or 0)), "cfd_ctx_set_stream")

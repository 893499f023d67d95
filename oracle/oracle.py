"""ORACLE — test infrastructure only.

ctypes bindings of the CPU restatement (``liboracle.so``) and, where it has
been built in this container, of the reference solver itself
(``_ref/libref.so``).  Only tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")
# variants of the reference built by `make ref-variants` (oracle/Makefile):
# "visc" viscous flux + bump BCs, "plate" viscous + flat-plate BCs,
# "super" inviscid + supersonic BCs
REF_VARIANTS = ("visc", "plate", "super")
REFERENCE = "/root/reference"

_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
_VP = C.c_void_p


def _dp(a):
    return a.ctypes.data_as(_DP)


def _ip(a):
    return a.ctypes.data_as(_IP)


def build(ref: bool = True):
    """make liboracle.so (always) and, when the reference is present, its
    compiled builds _ref/libref.so and _ref/libref_<variant>.so."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if ref and os.path.isdir(os.path.join(REFERENCE, "src")):
        subprocess.run(["make", "-s", "-C", HERE, "ref", "ref-variants"], check=True)


_orc = None
_ref = {}


def oracle_lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_LIB):
            build(ref=False)
        from mycode_amd.cfd import MeshDesc, Params  # struct layouts only
        lib = C.CDLL(ORACLE_LIB)
        P = C.POINTER(Params)
        sig = {
            "orc_mesh_from_desc": (_VP, [C.POINTER(MeshDesc)]),
            "orc_mesh_read_gmsh": (_VP, [C.c_char_p]),
            "orc_mesh_destroy": (None, [_VP]),
            "orc_mesh_sizes": (None, [_VP, _IP, _IP, _IP]),
            "orc_mesh_faces": (None, [_VP, _IP, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _DP, _DP]),
            "orc_mesh_cells": (None, [_VP, _DP, _DP, _DP]),
            "orc_solver_create": (_VP, [_VP, P]),
            "orc_solver_destroy": (None, [_VP]),
            "orc_set_params": (None, [_VP, P]),
            "orc_set_face_flux": (None, [_VP, C.POINTER(C.c_ubyte)]),
            "orc_wall_distance": (None, [_VP, _DP]),
            "orc_residual": (None, [_VP, _DP, _DP, C.c_int]),
            "orc_residual_cs": (None, [_VP, _DP, _DP, C.c_double, _DP, _DP, C.c_int]),
            "orc_residual_opcount": (None, [_VP, _DP, C.c_int, C.POINTER(C.c_ulonglong), _DP]),
            "orc_face_flux": (None, [_VP, _DP, _DP, C.c_int]),
            "orc_get_gradients": (None, [_VP, _DP]),
            "orc_get_gradients14": (None, [_VP, _DP]),
            "orc_turb_cells": (None, [_VP, _DP, _DP]),
            "orc_get_limiters": (None, [_VP, _DP]),
            "orc_set_limiters": (None, [_VP, _DP]),
            "orc_local_timestep": (None, [_VP, _DP, C.c_double]),
            "orc_get_dt": (None, [_VP, _DP]),
            "orc_set_dt": (None, [_VP, _DP]),
            "orc_diag_precond": (None, [_VP, _DP, _DP]),
            "orc_jv_fd": (None, [_VP, _DP, _DP, _DP, _DP]),
            "orc_norm_l2": (C.c_double, [_VP, _DP]),
            "orc_dot": (C.c_double, [_VP, _DP, _DP]),
            "orc_norm_vw": (C.c_double, [_VP, _DP]),
            "orc_gmres": (C.c_int, [_VP, _DP, _DP, _DP, _DP, C.c_int, C.c_double]),
            "orc_solve_implicit": (C.c_int, [_VP, _DP, C.c_double, C.c_int, C.c_double, C.c_int, _DP, _IP, _DP, _IP]),
            "orc_explicit_step": (None, [_VP, _DP, C.c_double]),
            "orc_solve_unsteady": (None, [_VP, _DP, C.c_double, C.c_int, _DP, _IP]),
            "orc_entropy_error": (C.c_double, [_VP, _DP]),
            "orc_bl_profile": (C.c_int, [_VP, _DP, C.c_double, _DP]),
            "orc_rk3_step": (None, [_VP, _DP, C.c_double]),
            "orc_wall_outputs": (None, [_VP, _DP, _DP, _DP, _DP]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _orc = lib
    return _orc


def ref_lib_path(variant: str | None = None) -> str:
    return REF_LIB if variant is None else os.path.join(HERE, "_ref", f"libref_{variant}.so")


def ref_available(variant: str | None = None) -> bool:
    return os.path.exists(ref_lib_path(variant))


def ref_lib(variant: str | None = None):
    if variant not in _ref:
        path = ref_lib_path(variant)
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path)
        sig = {
            "ref_mesh_load": (_VP, [C.c_char_p]),
            "ref_mesh_destroy": (None, [_VP]),
            "ref_mesh_sizes": (None, [_VP, _IP, _IP, _IP]),
            "ref_mesh_faces": (None, [_VP, _IP, _IP, _IP, _IP, _IP, _DP, _DP, _DP, _DP, _DP]),
            "ref_mesh_cells": (None, [_VP, _DP, _DP, _DP]),
            "ref_solver_create": (_VP, [_VP, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                        C.c_int]),
            "ref_solver_destroy": (None, [_VP]),
            "ref_set_method": (None, [_VP, C.c_int, C.c_int]),
            "ref_get_freestream": (None, [_VP, _DP]),
            "ref_initialize": (None, [_VP, _DP]),
            "ref_residual": (None, [_VP, _DP, _DP, C.c_int]),
            "ref_get_gradients": (None, [_VP, _DP]),
            "ref_get_limiters": (None, [_VP, _DP]),
            "ref_local_timestep": (None, [_VP, _DP, C.c_double]),
            "ref_get_dt": (None, [_VP, _DP]),
            "ref_diag_precond": (None, [_VP, _DP, _DP]),
            "ref_jv": (None, [_VP, _DP, _DP, _DP, _DP]),
            "ref_gmres": (C.c_int, [_VP, _DP, _DP, _DP, _DP]),
            "ref_norm_l2": (C.c_double, [_VP, _DP]),
            "ref_norm_vw": (C.c_double, [_VP, _DP]),
            "ref_solve_implicit": (None, [_VP, _DP, C.c_double, C.c_int]),
            "ref_explicit_step": (None, [_VP, _DP, C.c_double]),
            "ref_solve_unsteady": (None, [_VP, _DP, C.c_double, C.c_int]),
            "ref_entropy_error": (C.c_double, [_VP, _DP]),
            "ref_bl_profile": (None, [_VP, _DP, C.c_double, C.c_char_p]),
            "ref_save_vtk": (None, [_VP, _DP, C.c_int]),
            "ref_set_nu": (None, [_VP, C.c_double]),
            "ref_kinematic_viscosity": (None, [_VP, C.c_double, C.c_double]),
            "ref_get_nu": (C.c_double, [_VP]),
            "ref_skin_friction": (None, [_VP, _DP, C.c_char_p]),
            "ref_wall_pressure": (None, [_VP, _DP, C.c_char_p]),
            "ref_drag": (C.c_double, [_VP, _DP]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _ref[variant] = lib
    return _ref[variant]


def _faces_from(lib_fn, h, F):
    ints = {k: np.zeros(F, np.int32) for k in ("node1", "node2", "owner", "neighbor", "bc_tag")}
    dbl = {k: np.zeros(F) for k in ("nx", "ny", "area", "x_mid", "y_mid")}
    lib_fn(h, *[_ip(ints[k]) for k in ints], *[_dp(dbl[k]) for k in dbl])
    return {**ints, **dbl}


class OracleMesh:
    def __init__(self, x=None, y=None, cell_ptr=None, cell_nodes=None, bedges=None, gmsh: str | None = None):
        lib = oracle_lib()
        self.lib = lib
        if gmsh is not None:
            self.h = lib.orc_mesh_read_gmsh(gmsh.encode())
            if not self.h:
                raise FileNotFoundError(gmsh)
        else:
            from mycode_amd.cfd import MeshDesc
            self._keep = [np.ascontiguousarray(a, t) for a, t in
                          ((x, np.float64), (y, np.float64), (cell_ptr, np.int32), (cell_nodes, np.int32),
                           (np.asarray(bedges).reshape(-1), np.int32))]
            xx, yy, cp, cn, be = self._keep
            d = MeshDesc(len(xx) - 1, _dp(xx), _dp(yy), len(cp) - 1, _ip(cp), _ip(cn), len(be) // 3, _ip(be))
            self.h = lib.orc_mesh_from_desc(C.byref(d))
        nn, nc, nf = C.c_int(), C.c_int(), C.c_int()
        lib.orc_mesh_sizes(self.h, C.byref(nn), C.byref(nc), C.byref(nf))
        self.n_nodes, self.n_cells, self.n_faces = nn.value, nc.value, nf.value

    @classmethod
    def from_mesh(cls, mesh):
        """Oracle mesh built from the same description as a product Mesh."""
        x, y, cp, cn, be = mesh.desc()
        return cls(x, y, cp, cn, be)

    def faces(self):
        return _faces_from(self.lib.orc_mesh_faces, self.h, self.n_faces)

    def cells(self):
        out = {k: np.zeros(self.n_cells) for k in ("volume", "x_c", "y_c")}
        self.lib.orc_mesh_cells(self.h, _dp(out["volume"]), _dp(out["x_c"]), _dp(out["y_c"]))
        return out

    def __del__(self):
        try:
            self.lib.orc_mesh_destroy(self.h)
        except Exception:
            pass


class OracleSolver:
    """The CPU restatement with the reference Solver's member names."""

    def __init__(self, mesh: OracleMesh, params):
        self.lib = oracle_lib()
        self.mesh = mesh
        self.N = mesh.n_cells
        self.params = params.copy()
        self.h = self.lib.orc_solver_create(mesh.h, C.byref(self.params))

    def __del__(self):
        try:
            self.lib.orc_solver_destroy(self.h)
        except Exception:
            pass

    def set_params(self, p):
        self.params = p.copy()
        self.lib.orc_set_params(self.h, C.byref(self.params))

    def set_face_flux(self, alt):
        a = np.ascontiguousarray(alt, np.uint8)
        self.lib.orc_set_face_flux(self.h, a.ctypes.data_as(C.POINTER(C.c_ubyte)))

    def wall_distance(self):
        d = np.zeros(self.N)
        self.lib.orc_wall_distance(self.h, _dp(d))
        return d

    def compute_residuals(self, U, update_limiters=True):
        U = np.ascontiguousarray(U, np.float64)
        R = np.zeros((self.N, 6))
        self.lib.orc_residual(self.h, _dp(U), _dp(R), int(update_limiters))
        return R

    OPS = ("add", "mul", "div", "sqrt", "trans")

    def residual_opcount(self, U, update_limiters=True):
        """FP64 operations of one residual at U, counted by running the
        discretisation on an op-counting scalar: (whole residual, its cell
        stage: gradients, limiter, SST cell terms), each a dict of add, mul,
        div, sqrt and trans (pow, tanh) counts plus flops = add + mul + div +
        sqrt; and the residual itself (bit-identical to compute_residuals)."""
        U = np.ascontiguousarray(U, np.float64)
        R = np.zeros((self.N, 6))
        c = (C.c_ulonglong * 10)()
        self.lib.orc_residual_opcount(self.h, _dp(U), int(update_limiters), c, _dp(R))
        out = []
        for base in (0, 5):
            d = {k: int(c[base + i]) for i, k in enumerate(self.OPS)}
            d["flops"] = d["add"] + d["mul"] + d["div"] + d["sqrt"]
            out.append(d)
        return out[0], out[1], R

    def residual_complex_step(self, U, v, h=1e-30, update_limiters=True):
        U = np.ascontiguousarray(U, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        R, dR = np.zeros((self.N, 6)), np.zeros((self.N, 6))
        self.lib.orc_residual_cs(self.h, _dp(U), _dp(v), h, _dp(R), _dp(dR), int(update_limiters))
        return R, dR

    def face_fluxes(self, U, update_limiters=True):
        """Per-face F*A (reference face order), shape (F, 6)."""
        U = np.ascontiguousarray(U, np.float64)
        Fm = np.zeros((self.mesh.n_faces, 6))
        self.lib.orc_face_flux(self.h, _dp(U), _dp(Fm), int(update_limiters))
        return Fm

    def residual_scale(self, U, update_limiters=True):
        """S_i = sum over the faces of cell i of |F_f A_f| (per component)."""
        Fm = np.abs(self.face_fluxes(U, update_limiters))
        f = self.mesh.faces()
        S = np.zeros((self.N, 6))
        np.add.at(S, f["owner"], Fm)
        nb = f["neighbor"] >= 0
        np.add.at(S, f["neighbor"][nb], Fm[nb])
        return S

    def gradients(self):
        g = np.zeros((self.N, 10))
        self.lib.orc_get_gradients(self.h, _dp(g))
        return g

    def limiters(self):
        p = np.zeros(self.N)
        self.lib.orc_get_limiters(self.h, _dp(p))
        return p

    def gradients14(self):
        g = np.zeros((self.N, 14))
        self.lib.orc_get_gradients14(self.h, _dp(g))
        return g

    def turb_cells(self, U):
        """SST per cell at U: (mu_t, F1, S_k, S_omega) with S per unit volume."""
        out = np.zeros((self.N, 4))
        self.lib.orc_turb_cells(self.h, _dp(np.ascontiguousarray(U, np.float64)), _dp(out))
        return out

    def set_limiters(self, phi):
        phi = np.ascontiguousarray(phi, np.float64)
        self.lib.orc_set_limiters(self.h, _dp(phi))

    def compute_local_time_step(self, U, cfl):
        U = np.ascontiguousarray(U, np.float64)
        self.lib.orc_local_timestep(self.h, _dp(U), cfl)

    def dt_local(self):
        d = np.zeros(self.N)
        self.lib.orc_get_dt(self.h, _dp(d))
        return d

    def set_dt(self, dt):
        dt = np.ascontiguousarray(dt, np.float64)
        self.lib.orc_set_dt(self.h, _dp(dt))

    def diagonal_preconditioner(self, U):
        U = np.ascontiguousarray(U, np.float64)
        D = np.zeros(self.N)
        self.lib.orc_diag_precond(self.h, _dp(U), _dp(D))
        return D

    def apply_jacobian(self, U, R, v):
        U, R, v = (np.ascontiguousarray(a, np.float64) for a in (U, R, v))
        J = np.zeros((self.N, 6))
        self.lib.orc_jv_fd(self.h, _dp(U), _dp(R), _dp(v), _dp(J))
        return J

    def norm_L2(self, v):
        return self.lib.orc_norm_l2(self.h, _dp(np.ascontiguousarray(v, np.float64)))

    def dot_product(self, a, b):
        return self.lib.orc_dot(self.h, _dp(np.ascontiguousarray(a, np.float64)),
                                _dp(np.ascontiguousarray(b, np.float64)))

    def norm_L2_volume_weighted(self, R):
        return self.lib.orc_norm_vw(self.h, _dp(np.ascontiguousarray(R, np.float64)))

    def solve_gmres(self, U, R, RHS, m=40, tol=1e-2):
        U, R, RHS = (np.ascontiguousarray(a, np.float64) for a in (U, R, RHS))
        x = np.zeros((self.N, 6))
        it = self.lib.orc_gmres(self.h, _dp(U), _dp(R), _dp(RHS), _dp(x), m, tol)
        return x, it

    def solve_implicit(self, U, tol=1e-6, max_steps=50, rel_drop=0.0, m=40):
        U = np.ascontiguousarray(U, np.float64).copy()
        n = max(1, max_steps)
        hvm, hit, hcfl, conv = np.zeros(n), np.zeros(n, np.int32), np.zeros(n), C.c_int()
        steps = self.lib.orc_solve_implicit(self.h, _dp(U), tol, max_steps, rel_drop, m, _dp(hvm), _ip(hit),
                                            _dp(hcfl), C.byref(conv))
        k = steps + (1 if conv.value else 0)
        return U, {"steps": steps, "converged": conv.value, "vm": hvm[:k], "iters": hit[:k], "cfl": hcfl[:k]}

    def time_step_explicit(self, U, dt):
        U = np.ascontiguousarray(U, np.float64).copy()
        self.lib.orc_explicit_step(self.h, _dp(U), dt)
        return U

    def entropy_error(self, U):
        return self.lib.orc_entropy_error(self.h, _dp(np.ascontiguousarray(U, np.float64)))

    def bl_profile(self, U, x_target):
        U = np.ascontiguousarray(U, np.float64)
        n = self.lib.orc_bl_profile(self.h, _dp(U), x_target, None)
        rows = np.zeros((max(n, 1), 4))
        self.lib.orc_bl_profile(self.h, _dp(U), x_target, _dp(rows))
        return rows[:n]

    def solve_unsteady(self, U, dt_phys, n_steps):
        """Solver::solveUnsteady: (U, VM||F|| per Newton step (n_steps x 15), updates per step)."""
        U = np.ascontiguousarray(U, np.float64).copy()
        hvm = np.zeros((max(1, n_steps), 15))
        hn = np.zeros(max(1, n_steps), np.int32)
        self.lib.orc_solve_unsteady(self.h, _dp(U), dt_phys, n_steps, _dp(hvm), _ip(hn))
        return U, hvm[:n_steps], hn[:n_steps]

    def rk3_step(self, U, dt):
        U = np.ascontiguousarray(U, np.float64).copy()
        self.lib.orc_rk3_step(self.h, _dp(U), dt)
        return U

    def wall_outputs(self, U):
        U = np.ascontiguousarray(U, np.float64)
        F = self.mesh.n_faces
        Cp, Cf, J = np.zeros(F), np.zeros(F), C.c_double()
        self.lib.orc_wall_outputs(self.h, _dp(U), _dp(Cp), _dp(Cf), C.byref(J))
        return Cp, Cf, J.value


class RefMesh:
    """Mesh::load of the compiled reference (this container only)."""

    def __init__(self, path: str, variant: str | None = None):
        self.variant = variant
        self.lib = ref_lib(variant)
        self.h = self.lib.ref_mesh_load(path.encode())
        nn, nc, nf = C.c_int(), C.c_int(), C.c_int()
        self.lib.ref_mesh_sizes(self.h, C.byref(nn), C.byref(nc), C.byref(nf))
        self.n_nodes, self.n_cells, self.n_faces = nn.value, nc.value, nf.value

    def faces(self):
        return _faces_from(self.lib.ref_mesh_faces, self.h, self.n_faces)

    def save_vtk(self, U, step):
        """main.cpp saveResults: writes results/sol_<step>.vtk under the working directory."""
        self.lib.ref_save_vtk(self.h, _dp(np.ascontiguousarray(U, np.float64)), step)

    def cells(self):
        out = {k: np.zeros(self.n_cells) for k in ("volume", "x_c", "y_c")}
        self.lib.ref_mesh_cells(self.h, _dp(out["volume"]), _dp(out["x_c"]), _dp(out["y_c"]))
        return out


class RefSolver:
    """The reference Solver: as shipped (inviscid, bump boundary conditions)
    or, with the mesh's variant, one of the commented-in variants."""

    def __init__(self, mesh: RefMesh, M, P, T, flux, spatial, nu=-1.0, mode=0):
        self.lib = ref_lib(mesh.variant)
        self.N = mesh.n_cells
        self.mesh = mesh
        self.h = self.lib.ref_solver_create(mesh.h, mode, M, P, T, nu, flux, spatial)

    def freestream(self):
        o = np.zeros(4)
        self.lib.ref_get_freestream(self.h, _dp(o))
        return o

    def initialize(self):
        U = np.zeros((self.N, 6))
        self.lib.ref_initialize(self.h, _dp(U))
        return U

    def compute_residuals(self, U, update_limiters=True):
        U = np.ascontiguousarray(U, np.float64)
        R = np.zeros((self.N, 6))
        self.lib.ref_residual(self.h, _dp(U), _dp(R), int(update_limiters))
        return R

    def gradients(self):
        g = np.zeros((self.N, 10))
        self.lib.ref_get_gradients(self.h, _dp(g))
        return g

    def limiters(self):
        p = np.zeros(self.N)
        self.lib.ref_get_limiters(self.h, _dp(p))
        return p

    def compute_local_time_step(self, U, cfl):
        self.lib.ref_local_timestep(self.h, _dp(np.ascontiguousarray(U, np.float64)), cfl)

    def dt_local(self):
        d = np.zeros(self.N)
        self.lib.ref_get_dt(self.h, _dp(d))
        return d

    def diagonal_preconditioner(self, U):
        D = np.zeros(self.N)
        self.lib.ref_diag_precond(self.h, _dp(np.ascontiguousarray(U, np.float64)), _dp(D))
        return D

    def apply_jacobian(self, U, R, v):
        U, R, v = (np.ascontiguousarray(a, np.float64) for a in (U, R, v))
        J = np.zeros((self.N, 6))
        self.lib.ref_jv(self.h, _dp(U), _dp(R), _dp(v), _dp(J))
        return J

    def solve_gmres(self, U, R, RHS):
        U, R, RHS = (np.ascontiguousarray(a, np.float64) for a in (U, R, RHS))
        x = np.zeros((self.N, 6))
        it = self.lib.ref_gmres(self.h, _dp(U), _dp(R), _dp(RHS), _dp(x))
        return x, it

    def norm_L2(self, v):
        return self.lib.ref_norm_l2(self.h, _dp(np.ascontiguousarray(v, np.float64)))

    def norm_L2_volume_weighted(self, v):
        return self.lib.ref_norm_vw(self.h, _dp(np.ascontiguousarray(v, np.float64)))

    def solve_implicit(self, U, tol, maxsteps):
        U = np.ascontiguousarray(U, np.float64).copy()
        self.lib.ref_solve_implicit(self.h, _dp(U), tol, maxsteps)
        return U

    def time_step_explicit(self, U, dt):
        U = np.ascontiguousarray(U, np.float64).copy()
        self.lib.ref_explicit_step(self.h, _dp(U), dt)
        return U

    def solve_unsteady(self, U, dt_phys, n_steps):
        U = np.ascontiguousarray(U, np.float64).copy()
        self.lib.ref_solve_unsteady(self.h, _dp(U), dt_phys, n_steps)
        return U

    def entropy_error(self, U):
        """The value calculateEntropyError prints (6 significant digits)."""
        return self.lib.ref_entropy_error(self.h, _dp(np.ascontiguousarray(U, np.float64)))

    def bl_profile_csv(self, U, x_target, path):
        self.lib.ref_bl_profile(self.h, _dp(np.ascontiguousarray(U, np.float64)), x_target, path.encode())

    def set_nu(self, nu):
        self.lib.ref_set_nu(self.h, nu)

    def skin_friction_csv(self, U, path):
        """Solver::extractSkinFriction: writes x,Cf,Cf_theory rows to path."""
        self.lib.ref_skin_friction(self.h, _dp(np.ascontiguousarray(U, np.float64)), path.encode())

    def wall_pressure_csv(self, U, path):
        """Solver::extractWallPressure: writes x_in, p_psi rows to path."""
        self.lib.ref_wall_pressure(self.h, _dp(np.ascontiguousarray(U, np.float64)), path.encode())

    def drag(self, U):
        """The value Solver::calculateDrag prints (read at 17 digits)."""
        return self.lib.ref_drag(self.h, _dp(np.ascontiguousarray(U, np.float64)))

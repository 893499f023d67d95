"""The C-ABI library loads and exports every function include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

from mycode_amd import cfd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(cfd_\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    assert {"cfd_residual", "cfd_jv_fd", "cfd_gmres", "cfd_solve_implicit", "cfd_ctx_create"} <= names


def test_library_exports_every_declared_symbol():
    lib = cfd.load_library()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    bound = {n for n, _, _ in cfd.SIGNATURES}
    assert declared_functions() <= bound, declared_functions() - bound


def test_struct_sizes():
    assert ctypes.sizeof(cfd.Params) == 192
    assert ctypes.sizeof(cfd.GmresOpts) == 32
    assert ctypes.sizeof(cfd.NewtonOpts) == 80
    assert ctypes.sizeof(cfd.NewtonResult) == 40


def test_params_helpers_match_reference_formulas():
    p = cfd.default_params()
    assert (p.gamma, p.R_gas, p.nu, p.Pr) == (1.4, 287.0, 0.01, 0.72)
    cfd.flow_conditions(p, 0.5, 101325.0, 300.0)
    fT = 1.0 + 0.5 * (1.4 - 1.0) * 0.5 * 0.5
    assert p.T_total_inlet == 300.0 * fT
    assert p.P_total_inlet == 101325.0 * fT ** (1.4 / 0.4)
    U = cfd.initial_state(p)
    assert abs(U[1] / U[0] / (1.4 * 287 * 300) ** 0.5 - 0.5) < 1e-12

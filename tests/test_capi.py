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


def test_host_program_compiles_and_links_against_the_header(tmp_path):
    """A C++ host using only include/cfd_gpu.h and a 48-byte State vector
    (tests/host_abi_test.cpp) compiles warning-free and links: every entry
    point it calls resolves in the library."""
    import subprocess

    exe = tmp_path / "host_abi_test"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{os.path.join(ROOT, 'include')}",
           os.path.join(ROOT, "tests", "host_abi_test.cpp"), f"-L{os.path.join(ROOT, 'mycode_amd')}", "-lcfdgpu",
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert exe.exists()


def test_header_compiles_as_c(tmp_path):
    import subprocess

    src = tmp_path / "h.c"
    src.write_text('#include "cfd_gpu.h"\nint main(void) { cfd_params p; return cfd_params_default(&p); }\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{os.path.join(ROOT, 'include')}", "-c", str(src),
                        "-o", str(tmp_path / "h.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr

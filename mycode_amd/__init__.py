"""MI355X (gfx950) residual + JFNK-GMRES library for the 2D unstructured
finite-volume solver AmberMond/mycode.

The compute path is ``libcfdgpu.so`` (hand-written HIP, built by
``mycode_amd.build``); ``mycode_amd.cfd`` binds its C-ABI and mirrors the
reference ``Solver`` interface.
"""
from . import cfd  # noqa: F401
from .cfd import Mesh, Solver, Params, CfdError  # noqa: F401

__all__ = ["cfd", "Mesh", "Solver", "Params", "CfdError"]

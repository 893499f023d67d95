"""Residual throughput and JFNK time-to-converge of the MI355X library on the
turbulent flat plate (the configuration the headline metric is quoted on).

Workload: turbulent flat plate, Mach 0.2, Re = 5e6, 1200 x 1000 stretched
quads with a seeded 10 % split into triangles (1,320,146 cells, 2,522,346
faces), Roe + viscous + SST k-omega, second-order central, six conserved
variables, plate boundary conditions.  One "step" is one full FP64 residual
evaluation on device-resident state.

The JSON line carries:
  value              Mcell-residuals/s (cells x residuals / wall second) over
                     the timed region: `steps` residuals launched back to back;
  roofline           the dominant residual kernel: algorithmic bytes of its own
                     inputs and outputs per residual over its HIP-event time,
                     from a second pass of `steps` residuals right after the
                     timed one with HIP events between the stages on the
                     context stream (the events add about 25 us per residual
                     of wall time, which is why the timed region has none);
  roofline_residual  the whole residual against B_res = N(16 n_v + 24) + 48 F
                     over the timed region's wall time per residual;
  tangent_products   forward tangent products J v per second on the same
                     state (FP32 and FP64 directions) and reverse products
                     J^T w (discrete adjoint, FP64), beside the residual;
  jfnk               FD Jv + FP64 Krylov against the tangent Jv with FP32
                     matvec, preconditioner and basis (line-implicit block
                     Gauss-Seidel in both): milliseconds per GMRES iteration
                     on this workload (and the FP32 / FD solution
                     differences), the wall time to a 1e-10 drop of the
                     volume-weighted residual norm on the laminar flat plate
                     (GMRES(40)) and on this workload (GMRES(60), one
                     restart), with the largest Cp and Cf differences
                     between the two converged solutions;
  cpu_baseline       the CPU restatement of the reference residual (g++ -O2,
                     one core) on the same mesh and state.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CONFIG = 3
WORKLOADS = {
    2: "laminar flat plate M0.2 Re_L=1e5, 400x160 stretched quads 20% split, Roe + viscous, 2nd-order central",
    3: "turbulent flat plate M0.2 Re=5e6, 1200x1000 stretched quads 10% split, Roe + viscous + SST k-omega, "
       "2nd-order central",
}


def stage_bytes(N, F, nv, assembly, limiter, turb):
    """Compulsory bytes of each residual stage's own inputs and outputs.

    gradients: U, V and slots per cell, face geometry and indices, the
    gradients (and limiter) written, and with SST the wall distance read and
    mu_t, F1, V S_k, V S_w written; flux: both cells' U, gradients (limiter,
    mu_t and F1) and the centroids, face geometry, and the face buffer (or the
    scatter's read-modify-write of R); assembly: the face buffer, slots and R
    (and the two SST source rows)."""
    ng = 14 if turb else 10
    grad = N * (8 * nv + 16 + 8 * ng + (8 + 16 if limiter else 0)) + F * (8 + 24 + (16 if limiter else 0))
    if turb:
        grad += N * (8 + 32)
    flux = N * (8 * nv + 8 * ng + 16 + (8 if limiter else 0) + (8 if turb else 0)) + F * (12 + 40)
    if assembly == 0:
        flux += N * 16 * nv
        asm = N * 8 * nv
    else:
        flux += F * 8 * nv
        asm = F * 8 * nv + N * (16 + 8 * nv)
        if turb:
            asm += N * 16
    return grad, flux, asm


def cpu_baseline(mesh, p, U, budget_s=25.0):
    from oracle import oracle
    o = oracle.OracleSolver(oracle.OracleMesh.from_mesh(mesh), p)
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 and time.perf_counter() - t_start < budget_s:
        t0 = time.perf_counter()
        o.compute_residuals(U)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": mesh.n_cells / t / 1e6, "unit": "Mcell-residuals/s", "cores": 1, "kind": "port",
            "sample": f"{len(times)} full residual evaluations of the same {mesh.n_cells}-cell mesh and state "
                      f"(median {t:.3f} s), oracle/cfd_oracle.cpp built g++ -O2"}


LEGS = (("fd_fp64", 0, 64), ("tangent_fp64", 1, 64), ("tangent_fp32", 1, 32))


def gmres_iteration_cost(solver, Ud, Rd, nv, pitch, cfl=50.0, m=40, precond=3, legs=LEGS):
    """Per-iteration cost of GMRES(m) at the bench state on the bench mesh:
    FD Jv + FP64 Krylov vs the tangent with FP32 matvec, preconditioner and
    basis (and the FP64 tangent, to separate the FP32 error from the FD
    error), both with the line-implicit preconditioner.  The tolerance is
    unreachable, so every variant runs all m iterations."""
    import torch
    solver.dev_local_timestep(Ud.data_ptr(), cfl)
    solver.dev_residual(Ud.data_ptr(), Rd.data_ptr(), False)
    solver.synchronize()
    rhs = -Rd
    out, xs = {}, {}
    for name, jv, prec in legs:
        x = torch.zeros_like(Rd)
        o = solver.gmres_opts(m=m, tol=1e-30, jv=jv, precision=prec, precond=precond, restarts=0)
        torch.cuda.synchronize()  # torch's kernels and the library's run on different streams
        solver.dev_gmres(Ud.data_ptr(), Rd.data_ptr(), rhs.data_ptr(), x.data_ptr(), o)  # warm-up, allocations
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        it = solver.dev_gmres(Ud.data_ptr(), Rd.data_ptr(), rhs.data_ptr(), x.data_ptr(), o)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[name] = {"ms_per_iteration": 1e3 * dt / it, "iterations": it}
        xs[name] = x.view(nv, pitch)[:, :mesh_cells(solver)].double()
    out["speedup_tangent_fp32_vs_fd_fp64"] = out["fd_fp64"]["ms_per_iteration"] / out["tangent_fp32"]["ms_per_iteration"]
    if "tangent_fp64" in xs:
        ref = xs["tangent_fp64"]
        rel = lambda a: float(torch.linalg.norm(a - ref) / torch.linalg.norm(ref))
        out["rel_diff_fp32_vs_fp64_solution"] = rel(xs["tangent_fp32"])
        out["rel_diff_fd_vs_tangent_solution"] = rel(xs["fd_fp64"])
    out["setup"] = f"GMRES({m}), line-implicit block Gauss-Seidel, CFL {cfl}, at the benchmark state"
    return out


def tangent_throughput(solver, Ud, nv, pitch, N, reps=20, cfl=50.0):
    """Forward tangent products J v (the full second-order residual's
    tangent, as GMRES uses it) per second on the workload's state, FP32 and
    FP64 directions, and reverse products J^T w (the discrete adjoint, FP64),
    against the FP64 residual: wall time of `reps` device products after a
    warm-up, on the context stream."""
    import torch
    solver.dev_local_timestep(Ud.data_ptr(), cfl)
    out = {}
    w = torch.randn(nv * pitch, dtype=torch.float64, device=Ud.device,
                    generator=torch.Generator(device=Ud.device).manual_seed(8))
    JTw = torch.zeros_like(w)
    torch.cuda.synchronize()
    solver.dev_vjp_adjoint(Ud.data_ptr(), w.data_ptr(), JTw.data_ptr())
    solver.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        solver.dev_vjp_adjoint(Ud.data_ptr(), w.data_ptr(), JTw.data_ptr())
    solver.synchronize()
    dt_s = (time.perf_counter() - t0) / reps
    out["adjoint_fp64"] = {"ms_per_product": 1e3 * dt_s, "Mcell_products_per_s": N / dt_s / 1e6}
    for prec, dt in ((32, torch.float32), (64, torch.float64)):
        v = torch.randn(nv * pitch, dtype=torch.float64, device=Ud.device, generator=torch.Generator(
            device=Ud.device).manual_seed(7)).to(dt)
        Jv = torch.zeros_like(v)
        torch.cuda.synchronize()  # torch's kernels and the library's run on different streams
        solver.dev_jv_tangent(Ud.data_ptr(), v.data_ptr(), Jv.data_ptr(), prec)
        solver.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            solver.dev_jv_tangent(Ud.data_ptr(), v.data_ptr(), Jv.data_ptr(), prec)
        solver.synchronize()
        dt_s = (time.perf_counter() - t0) / reps
        out[f"fp{prec}"] = {"ms_per_product": 1e3 * dt_s, "Mcell_products_per_s": N / dt_s / 1e6}
    return out


def mesh_cells(solver):
    return solver.N


def jfnk_converge(config=2, scale=1.0, rel_drop=1e-10, max_steps=600, m=40, restarts=0):
    """Time to a rel_drop drop of the volume-weighted residual norm, FD Jv +
    FP64 Krylov against tangent Jv + FP32 Krylov, same preconditioner
    (line-implicit block Gauss-Seidel), GMRES(m) with `restarts` restarts at
    tolerance 1e-2 and the adaptive CFL controller, from the configuration's
    initial state: the laminar flat plate (76,790 cells) with GMRES(40), the
    turbulent flat plate (1,320,146 cells) with GMRES(60) and one restart."""
    from mycode_amd import cfd, configs
    mesh, p, U0 = configs.setup(config, scale)
    out, sol = {}, {}
    for name, jv, prec in (("fd_fp64", 0, 64), ("tangent_fp32", 1, 32)):
        s = cfd.Solver(mesh, p)
        g = s.gmres_opts(m=m, tol=1e-2, jv=jv, precision=prec, precond=3, restarts=restarts)
        o = s.newton_opts(gmres=g, tolerance=0.0, rel_drop=rel_drop, max_steps=max_steps, controller=2, cfl0=1.0,
                          cfl_max=1e12)
        U, res, hist = s.solve_implicit(U0, o)
        sol[name] = s.wall_outputs(U)
        out[name] = {"seconds": res.seconds, "newton_steps": res.steps, "converged": bool(res.converged),
                     "drop": float(hist["vm"][-1] / hist["vm"][0]), "gmres_iterations": res.total_gmres_iters}
        print(f"bench: {WORKLOADS[config]}, {name}: {res.steps} Newton steps, {res.seconds:.1f} s, "
              f"drop {out[name]['drop']:.2e}", file=sys.stderr, flush=True)
    out["speedup_tangent_fp32_vs_fd_fp64"] = out["fd_fp64"]["seconds"] / out["tangent_fp32"]["seconds"]
    out["max_abs_Cp_diff"] = float(np.max(np.abs(sol["fd_fp64"][0] - sol["tangent_fp32"][0])))
    out["max_abs_Cf_diff"] = float(np.max(np.abs(sol["fd_fp64"][1] - sol["tangent_fp32"][1])))
    out["workload"] = WORKLOADS[config] + f" ({mesh.n_cells} cells)"
    out["gmres"] = f"GMRES({m}), {restarts} restart(s), tolerance 1e-2"
    return out


def copy_calibration():
    import torch
    n = 32 * 1024 * 1024  # doubles: 256 MiB read and 256 MiB written per copy
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 8, "copies": 3}), flush=True)


def measured_traffic(kernel, mesh_hash):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    the same mesh (FETCH_SIZE and WRITE_SIZE scaled by a plain copy), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except OSError:
        return None
    for r in rec.get("kernels", []):
        if r.get("stage") == kernel and r.get("mesh_hash") == mesh_hash:
            return r["bytes_per_launch"]
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--assembly", type=int, default=1)
    ap.add_argument("--config", type=int, default=CONFIG)
    ap.add_argument("--scale", type=float, default=1.0, help="mesh coarsening (1.0 = the named configuration)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-jfnk", action="store_true", help="skip the JFNK legs")
    ap.add_argument("--gmres-only", action="store_true",
                    help="only the GMRES(40) legs FD + FP64 and tangent + FP32 on the workload (for a kernel trace)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy three times (scales the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU layout is design-only",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration()
        return

    import torch
    from mycode_amd import cfd, configs

    mesh, p, U0 = configs.setup(args.config, args.scale)
    rng = np.random.default_rng(12345)
    U = U0.copy()
    U[:, :4] *= 1.0 + 1e-3 * rng.standard_normal((U.shape[0], 4))
    N, F, nv = mesh.n_cells, mesh.n_faces, p.nvar
    solver = cfd.Solver(mesh, p, assembly=args.assembly)
    pitch = solver.pitch
    dev = torch.device("cuda:0")
    U6 = torch.from_numpy(U).to(dev)
    Ud = torch.zeros(nv * pitch, dtype=torch.float64, device=dev)
    Rd = torch.zeros_like(Ud)
    torch.cuda.synchronize()
    solver.dev_pack(U6.data_ptr(), Ud.data_ptr())
    if args.gmres_only:
        legs = tuple(l for l in LEGS if l[0] != "tangent_fp64")
        print(json.dumps({"gmres": gmres_iteration_cost(solver, Ud, Rd, nv, pitch, legs=legs)}), flush=True)
        return
    for _ in range(args.warmup):
        solver.dev_residual(Ud.data_ptr(), Rd.data_ptr(), False)
    solver.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        solver.dev_residual(Ud.data_ptr(), Rd.data_ptr(), False)
    solver.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms, launches = solver.dev_time_residual(Ud.data_ptr(), Rd.data_ptr(), args.steps, False)
    solver.synchronize()
    R6 = torch.zeros(N * 6, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()  # the zero fill (torch's stream) must finish before the unpack (context stream)
    solver.dev_unpack(Rd.data_ptr(), R6.data_ptr())
    solver.synchronize()
    R_dev = R6.cpu().numpy().reshape(N, 6)
    R_host = solver.compute_residuals(U, False)
    assert np.array_equal(R_dev, R_host), "device and host residual paths disagree"

    limiter = p.spatial == cfd.SECOND_ORDER_LIMITED
    bg, bf, ba = stage_bytes(N, F, nv, args.assembly, limiter, bool(p.turb))
    stages = {"gradients": (bg, ms[0]), "face_flux": (bf, ms[1]), "assembly": (ba, ms[2])}
    dom = max(stages, key=lambda k: stages[k][1])
    b_dom, t_dom = stages[dom]
    achieved = b_dom / (t_dom * 1e-3) / 1e9
    B_res = N * (16 * nv + 24) + 48 * F
    ach_res = B_res / (wall / args.steps) / 1e9
    wl = WORKLOADS.get(args.config, f"configuration {args.config}") + ", FP64 residual"
    if args.scale != 1.0:
        wl += f" (mesh scaled by {args.scale})"
    out = {
        "metric": "Mcell-residuals/s",
        "value": N * args.steps / wall / 1e6,
        "unit": "Mcell-residuals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded mesh and state)",
        "config": {"workload": wl, "cells": N, "faces": F, "nvar": nv,
                   "assembly": ["colored-scatter", "face-buffer-gather"][args.assembly],
                   "flux_launches_per_residual": launches, "mesh_hash": hex(mesh.hash())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(dom, hex(mesh.hash())),
                     "kernel": dom,
                     "algorithmic_bytes_per_residual": b_dom, "ms_per_residual": t_dom},
        "roofline_residual": {"bound": "hbm", "achieved": ach_res, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": ach_res / HBM_PEAK_GBS, "B_res": B_res,
                              "ms_wall": wall / args.steps * 1e3, "ms_event": ms[3]},
        "stage_ms": {"gradients": ms[0], "face_flux": ms[1], "assembly": ms[2], "residual": ms[3]},
    }
    out["tangent_products"] = tangent_throughput(solver, Ud, nv, pitch, N)
    if not args.no_jfnk:
        out["jfnk"] = {"per_iteration_on_this_workload": gmres_iteration_cost(solver, Ud, Rd, nv, pitch),
                       "time_to_1e-10": jfnk_converge()}
        if args.config == CONFIG and args.scale == 1.0:
            out["jfnk"]["time_to_1e-10_this_workload"] = jfnk_converge(CONFIG, 1.0, rel_drop=1e-10, max_steps=800,
                                                                       m=60, restarts=1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, p, U)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

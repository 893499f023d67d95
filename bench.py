"""Residual throughput and JFNK time-to-converge of the MI355X library on the
turbulent flat plate (the configuration the headline metric is quoted on).

Workload: turbulent flat plate, Mach 0.2, Re = 5e6, 1200 x 1000 stretched
quads with a seeded 10 % split into triangles (1,320,146 cells, 2,522,346
faces), Roe + viscous + SST k-omega with k and omega convected first-order
upwind, mean flow second-order central, six conserved variables, plate
boundary conditions.  One "step" is one full FP64 residual evaluation on
device-resident state.

The JSON line carries:
  value              Mcell-residuals/s (cells x residuals / wall second) over
                     the timed region: `steps` residuals launched back to back;
  roofline           the dominant residual kernel against the compulsory
                     bytes of one residual, B_res = N(16 n_v + 24) + 48 F
                     (U read and R written once, cell and face geometry),
                     over that kernel's HIP-event time per launch, from a
                     second pass of `steps` residuals with HIP events between
                     the stages on the context stream (the events add about
                     25 us per residual of wall time, which is why the timed
                     region has none); `traffic` is that kernel's HBM bytes
                     per launch from the committed counter summary;
  roofline_residual  the whole residual against B_res over the timed
                     region's wall time per residual (both also against the
                     achievable line, a stream triad measured in the run);
  stage_traffic      each residual kernel against the compulsory bytes of its
                     own inputs and outputs (gradients and face buffer
                     included), the view that locates wasted re-reads;
  tangent_products   forward tangent products J v (FP32 and FP64) and reverse
                     products J^T w (FP64) on the same state, beside the
                     residual;
  other_configs      the same residual, roofline and product measurements on
                     configuration 4 (bump, 824,378 cells) and configuration 5
                     (shock / boundary-layer interaction, 4,399,865 cells);
  jfnk               FD Jv + FP64 Krylov against the tangent Jv with FP32
                     matvec, preconditioner and basis: per GMRES iteration on
                     this workload; the wall time to a 1e-10 drop of the
                     volume-weighted residual norm on this workload and on the
                     laminar flat plate, with the per-equation drops and the
                     Cp / Cf differences of the two solutions; and the
                     laminar plate as configuration 2 states it (GMRES(30),
                     block-Jacobi, FD against the tangent, all FP64);
  design_gradient    configuration 4 at full size: steady state, discrete
                     adjoint and the gradient of the bump wall-pressure
                     integral for 64 bump-node displacements, checked against
                     central differences of re-converged flows (step 1e-6);
  cpu_baseline       the CPU restatement of the reference (g++ -O2, one core)
                     on the same mesh and state: residual and FD Jv times, and
                     the reference algorithm's Newton-step time extrapolated
                     from them; host core count and CPU model; the FP64
                     operation count of one residual (op-counting scalar);
  roofline_fp64      that count over the timed residual's wall time, and per
                     stage over the HIP-event times of its kernels, against
                     the 78.6 TFLOP/s FP64 vector peak (with cpu_baseline).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

T_START = time.perf_counter()  # the optional checks stop once the run has used RUN_BUDGET_S
RUN_BUDGET_S = 430.0

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 vector
CONFIG = 3
WORKLOADS = {
    2: "laminar flat plate M0.2 Re_L=1e5, 400x160 stretched quads 20% split, Roe + viscous, 2nd-order central",
    3: "turbulent flat plate M0.2 Re=5e6, 1200x1000 stretched quads 10% split, Roe + viscous + SST k-omega, "
       "mean flow 2nd-order central, k and omega convected 1st-order upwind",
    4: "inviscid bump M0.7, 3x1 channel, 5% circular arc, 1280x560 quads 15% split, HLLC, 2nd-order central",
    5: "shock / boundary-layer interaction M2.0 Re_x=3e5, 8-degree incident shock, 2000x2000 stretched quads "
       "10% split, LLF in the shock band and HLLC elsewhere, viscous, Barth-Jespersen limited",
}


def stage_bytes(N, F, nv, assembly, limiter, turb, kw_recon=False):
    """Compulsory bytes of each residual stage's own inputs and outputs.

    gradients: U, V and slots per cell, face geometry and indices, the
    gradients (and limiter) written, and with SST the wall distance read and
    mu_t, F1, V S_k, V S_w written; flux: both cells' U, gradients (limiter,
    mu_t and F1) and the centroids, face geometry, and the face buffer (or the
    scatter's read-modify-write of R); assembly: the face buffer, slots and R
    (and the two SST source rows).  The k and omega gradient rows count only
    when k and omega are reconstructed (kw_recon); convected with the cell
    values they are written and read on the boundary alone."""
    ng = 14 if turb and kw_recon else 10
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


def host_info():
    """Host cores (those this process may run on) and CPU model."""
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        ncores = os.cpu_count()
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return ncores, model


def cpu_baseline(mesh, p, U, budget_s=25.0):
    """The oracle (the CPU restatement of the reference, g++ -O2, one core) on
    the workload's mesh and state: the residual (median of up to 3) and the FD
    Jv (median of up to 2, Solver::applyJacobian), and from them the time of
    one Newton step of the reference algorithm (Solver.cpp:1696-1799: a local
    time step, a residual and GMRES(40) with one FD Jv per iteration, taken
    at its 40-iteration cap) and the extrapolated time of `steps` such steps.
    Also the FP64 operation count of one residual on the same state (the
    residual run on an op-counting scalar), which the FP64 roofline uses."""
    from oracle import oracle
    o = oracle.OracleSolver(oracle.OracleMesh.from_mesh(mesh), p)
    t_start = time.perf_counter()
    res_t = []
    while len(res_t) < 3 and time.perf_counter() - t_start < budget_s:
        t0 = time.perf_counter()
        R = o.compute_residuals(U)
        res_t.append(time.perf_counter() - t0)
    o.compute_local_time_step(U, 10.0)
    v = np.random.default_rng(3).standard_normal(U.shape) * np.abs(U).mean(0)
    v[:, p.nvar:] = 0.0
    jv_t = []
    while len(jv_t) < 2 and time.perf_counter() - t_start < 2 * budget_s:
        t0 = time.perf_counter()
        o.apply_jacobian(U, R, v)
        jv_t.append(time.perf_counter() - t0)
    t_res, t_jv = float(np.median(res_t)), float(np.median(jv_t))
    ops, cell_ops, _ = o.residual_opcount(U)
    ncores, model = host_info()
    step = 2 * t_res + 40 * t_jv
    return {"value": mesh.n_cells / t_res / 1e6, "unit": "Mcell-residuals/s", "cores": 1, "kind": "port",
            "sample": f"{len(res_t)} full residual evaluations and {len(jv_t)} FD Jv products of the same "
                      f"{mesh.n_cells}-cell mesh and state (medians {t_res:.3f} s and {t_jv:.3f} s), "
                      "oracle/cfd_oracle.cpp built g++ -O2, one thread",
            "residual_s": t_res, "fd_jv_s": t_jv,
            "newton_step_s_estimate": step,
            "newton_step_note": "reference algorithm: local time step + residual + GMRES(40) at its cap, one FD Jv "
                                "per iteration; estimated from the two measured times",
            "host_cores": ncores, "host_cpu": model,
            "fp64_ops_per_residual": ops, "fp64_ops_cell_stage": cell_ops,
            "fp64_ops_note": "the oracle's residual run once on an op-counting scalar: + - * / and square root one "
                             "flop each as written, pow and tanh apart (trans); the cell stage is the gradients, "
                             "limiter and SST cell terms, the rest the face fluxes and their sums"}


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


def vm_per_equation(mesh, R, nv):
    """Volume-weighted norm of each equation's residual (their squares sum to
    the square of the reference's VM norm)."""
    vol = mesh.cells()["volume"]
    return np.sqrt(np.sum(R[:, :nv] ** 2 / vol[:, None], 0) / np.sum(vol))


EQ_NAMES = ("rho", "rho_u", "rho_v", "rho_E", "rho_k", "rho_omega")


CF_REF = os.path.join(ROOT, "mycode_amd", "data", "sst_plate_cf_ref.npz")


def load_cf_reference(mesh):
    """Cf of the tightly converged FP64 solution of the turbulent flat plate
    (mycode_amd/data/make_sst_plate_ref.py), if it was made for this mesh."""
    try:
        with np.load(CF_REF, allow_pickle=False) as z:
            if int(z["mesh_hash"]) != mesh.hash():
                return None
            return {k: z[k] for k in z.files}
    except OSError:
        return None


def jfnk_converge(config=2, scale=1.0, rel_drop=1e-10, max_steps=600, m=40, restarts=0, precond=3,
                  legs=(("fd_fp64", 0, 64), ("tangent_fp32", 1, 32)), tol=1e-2, label=None, cf_ref=None,
                  cf_tol=1e-6, extra_steps=400, chunk=10):
    """Time to a rel_drop drop of the volume-weighted residual norm, FD Jv +
    FP64 Krylov against tangent Jv + FP32 Krylov, same preconditioner, GMRES(m)
    with `restarts` restarts at tolerance `tol` and the adaptive CFL
    controller, from the configuration's initial state.  Reports each leg's
    per-equation drops and the largest Cp and Cf differences between the
    legs' solutions.  With `cf_ref` (Cf of a tightly converged FP64 solution
    on the no-slip faces) each leg then continues, `chunk` Newton steps at a
    time, until its Cf is within `cf_tol` of it (at most `extra_steps` more
    steps), and the solver time to that point is reported beside the time to
    the drop."""
    from mycode_amd import cfd, configs
    mesh, p, U0 = configs.setup(config, scale)
    out, sol = {}, {}
    for name, jv, prec in legs:
        s = cfd.Solver(mesh, p)
        g = s.gmres_opts(m=m, tol=tol, jv=jv, precision=prec, precond=precond, restarts=restarts)
        o = s.newton_opts(gmres=g, tolerance=0.0, rel_drop=rel_drop, max_steps=max_steps, controller=2, cfl0=1.0,
                          cfl_max=1e12, log=50)  # progress to stderr every 50 Newton steps
        U, res, hist = s.solve_implicit(U0, o)
        sol[name] = s.wall_outputs(U)
        d0 = vm_per_equation(mesh, s.compute_residuals(U0), p.nvar)
        d1 = vm_per_equation(mesh, s.compute_residuals(U), p.nvar)
        out[name] = {"seconds": res.seconds, "newton_steps": res.steps, "converged": bool(res.converged),
                     "drop": float(hist["vm"][-1] / hist["vm"][0]), "gmres_iterations": res.total_gmres_iters,
                     "drop_per_equation": {EQ_NAMES[k]: float(d1[k] / d0[k]) for k in range(p.nvar)}}
        print(f"bench: {label or WORKLOADS[config]}, {name}: {res.steps} Newton steps, {res.seconds:.1f} s, "
              f"drop {out[name]['drop']:.2e}, converged {bool(res.converged)}", file=sys.stderr, flush=True)
        if cf_ref is not None:
            f = mesh.faces()
            wall = (f["neighbor"] < 0) & (f["bc_tag"] == 4)
            err = float(np.max(np.abs(sol[name][1][wall] - cf_ref)))
            t_extra, n_extra, its_extra = 0.0, 0, 0
            cfl = float(hist["cfl"][-1]) if len(hist["cfl"]) else 1.0
            while err > cf_tol and n_extra < extra_steps:
                o2 = s.newton_opts(gmres=g, tolerance=0.0, rel_drop=0.0, max_steps=chunk, controller=2, cfl0=cfl,
                                   cfl_max=1e12)
                U, r2, h2 = s.solve_implicit(U, o2)
                t_extra += r2.seconds
                n_extra += r2.steps
                its_extra += r2.total_gmres_iters
                cfl = float(h2["cfl"][-1])
                sol[name] = s.wall_outputs(U)
                err = float(np.max(np.abs(sol[name][1][wall] - cf_ref)))
            d2 = vm_per_equation(mesh, s.compute_residuals(U), p.nvar)
            out[name]["to_cf_within_tol"] = {
                "seconds": res.seconds + t_extra, "newton_steps": res.steps + n_extra,
                "gmres_iterations": res.total_gmres_iters + its_extra, "max_abs_Cf_err": err, "met": err <= cf_tol,
                "drop_per_equation": {EQ_NAMES[k]: float(d2[k] / d0[k]) for k in range(p.nvar)}}
            print(f"bench: {name}: Cf within {err:.2e} of the tight FP64 solution after {res.steps + n_extra} steps, "
                  f"{res.seconds + t_extra:.1f} s", file=sys.stderr, flush=True)
    a, b = legs[0][0], legs[-1][0]
    out[f"speedup_{b}_vs_{a}"] = out[a]["seconds"] / out[b]["seconds"]
    if cf_ref is not None:
        out[f"speedup_{b}_vs_{a}_to_cf_within_tol"] = (out[a]["to_cf_within_tol"]["seconds"]
                                                       / out[b]["to_cf_within_tol"]["seconds"])
        out["cf_tol"] = cf_tol
    out["max_abs_Cp_diff"] = float(np.max(np.abs(sol[a][0] - sol[b][0])))
    out["max_abs_Cf_diff"] = float(np.max(np.abs(sol[a][1] - sol[b][1])))
    out["workload"] = (label or WORKLOADS[config]) + f" ({mesh.n_cells} cells)"
    out["gmres"] = f"GMRES({m}), {restarts} restart(s), tolerance {tol:g}, preconditioner " + \
        {1: "block-Jacobi", 2: "block SGS", 3: "line-implicit block Gauss-Seidel"}[precond]
    return out


def other_config(config, steps, warmup):
    """Residual throughput, its roofline (B_res over the dominant kernel's
    HIP-event time and over the whole residual's wall time), and the tangent
    and adjoint products of another configuration at full size."""
    import torch
    from mycode_amd import cfd, configs
    t0 = time.perf_counter()
    mesh, p, U0 = configs.setup(config, 1.0)
    U = U0.copy()
    U[:, :4] *= 1.0 + 1e-3 * np.random.default_rng(12345).standard_normal((U.shape[0], 4))
    N, F, nv = mesh.n_cells, mesh.n_faces, p.nvar
    s = cfd.Solver(mesh, p)
    if config == 5:
        s.set_face_flux(configs.shock_band(mesh))
    s.set_stream(torch.cuda.current_stream().cuda_stream)
    dev = torch.device("cuda:0")
    Ud = torch.zeros(nv * s.pitch, dtype=torch.float64, device=dev)
    Rd = torch.zeros_like(Ud)
    s.dev_pack(torch.from_numpy(U).to(dev).data_ptr(), Ud.data_ptr())
    s.dev_residual(Ud.data_ptr(), Rd.data_ptr(), True)  # forms the limiter (configuration 5)
    for _ in range(warmup):
        s.dev_residual(Ud.data_ptr(), Rd.data_ptr(), False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        s.dev_residual(Ud.data_ptr(), Rd.data_ptr(), False)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t1) / steps
    ms, _ = s.dev_time_residual(Ud.data_ptr(), Rd.data_ptr(), steps, False)
    B_res = N * (16 * nv + 24) + 48 * F
    stages = {"gradients": ms[0], "face_flux": ms[1], "assembly": ms[2]}
    dom = max(stages, key=stages.get)
    prods = tangent_throughput(s, Ud, nv, s.pitch, N)
    s.set_stream(None)
    return {"workload": WORKLOADS[config], "cells": N, "faces": F, "nvar": nv, "mesh_hash": hex(mesh.hash()),
            "Mcell_residuals_per_s": N / wall / 1e6, "ms_per_residual": wall * 1e3,
            "stage_ms": {k: v for k, v in stages.items()},
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": B_res / (stages[dom] * 1e-3) / 1e9,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": B_res / (stages[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS, "B_res": B_res},
            "roofline_residual": {"achieved": B_res / wall / 1e9, "frac": B_res / wall / 1e9 / HBM_PEAK_GBS},
            "tangent_products": prods, "seconds": time.perf_counter() - t0}


def design_gradient_leg(h=1e-6, fd_modes=(5, 20, 31, 47), deadline=None):
    """Configuration 4 at full size: the steady state (tangent + FP32 Krylov,
    block SGS, adaptive CFL), the discrete adjoint of the bump wall-pressure
    integral and its gradient for 64 bump-node displacements, and central
    differences of re-converged flows (step h) for `fd_modes` while each
    next check is expected to end before `deadline` (a perf_counter value)."""
    from mycode_amd import adjoint, cfd, configs
    t0 = time.perf_counter()
    mesh, p, U0 = configs.setup(4, 1.0)
    s = cfd.Solver(mesh, p)
    U, res, hist = adjoint.converge(s, U0, rel_drop=1e-12, log_every=50)
    t_conv = time.perf_counter() - t0
    print(f"bench: design: steady state in {res.steps} Newton steps, {t_conv:.1f} s", file=sys.stderr, flush=True)
    modes = adjoint.design_modes(mesh, 64)
    t1 = time.perf_counter()
    psi, its, outer, defect = adjoint.adjoint_solve_fgmres(
        s, U, adjoint.functional_dU(mesh, p, U),
        log=lambda msg: print("bench: design: " + msg, file=sys.stderr, flush=True))
    t_adj = time.perf_counter() - t1
    t2 = time.perf_counter()
    g = adjoint.gradient_from_adjoint(mesh, s, p, U, modes, psi)
    t_grad = time.perf_counter() - t2
    print(f"bench: design: adjoint in {outer} outer iterations, {t_adj:.1f} s; 64-mode gradient in {t_grad:.1f} s",
          file=sys.stderr, flush=True)
    out = {"workload": WORKLOADS[4] + f" ({mesh.n_cells} cells)", "converged": bool(res.converged),
           "newton_steps": res.steps, "drop": float(hist["vm"][-1] / hist["vm"][0]), "converge_s": t_conv,
           "steady_solver": "pseudo-time JFNK (tangent, FP32 Krylov, block SGS) to a 1e-4 drop, then Newton "
                            "steps solved by FGMRES(100) over GMRES(60) at CFL 1e6",
           "adjoint_s": t_adj, "adjoint_solver": "FGMRES(200), inner transposed GMRES(60) at CFL 1e6",
           "adjoint_outer_iterations": outer, "adjoint_gmres_iterations": its,
           "adjoint_relative_residual": defect, "gradient_64_modes_s": t_grad,
           "functional": adjoint.functional(mesh, p, U), "gradient": [float(x) for x in g], "fd_check": []}
    x0, y0, _, _, _ = mesh.desc()
    last = 0.0
    for k in fd_modes:
        if deadline is not None and time.perf_counter() + last > deadline:
            out["fd_check_note"] = "stopped at the run's time budget"
            break
        tk = time.perf_counter()
        J = []
        for sgn in (1.0, -1.0):
            m2 = cfd.Mesh.generate(4, 1.0)
            m2.set_nodes(x0, y0 + sgn * h * modes[k])
            s2 = cfd.Solver(m2, p)
            U2, r2, _ = adjoint.converge(s2, U, rel_drop=0.0, tolerance=2.0 * res.res_vm, cfl0=1e4, log_every=50,
                                         newton_below=1.0)
            J.append(adjoint.functional(m2, p, U2))
        fd = (J[0] - J[1]) / (2 * h)
        out["fd_check"].append({"mode": k, "adjoint": float(g[k]), "central_fd": fd, "step": h,
                                "rel_diff": abs(g[k] - fd) / abs(fd)})
        print(f"bench: design mode {k}: adjoint {g[k]:.8e} FD {fd:.8e}", file=sys.stderr, flush=True)
        last = time.perf_counter() - tk
    out["seconds"] = time.perf_counter() - t0
    return out


def triad_bandwidth(n=64 * 1024 * 1024, reps=20):
    """The achievable HBM line: a stream triad a = b + s c over three FP64
    arrays of n values (512 MiB each by default, far beyond the caches; torch's
    elementwise kernel), 3 x 8 n bytes per pass, timed by events on torch's
    stream over `reps` passes after three untimed ones.  GB/s."""
    import torch
    b = torch.rand(n, dtype=torch.float64, device="cuda")
    c = torch.rand_like(b)
    a = torch.empty_like(b)
    for _ in range(3):
        torch.add(b, c, alpha=0.5, out=a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        torch.add(b, c, alpha=0.5, out=a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b, c
    return 3 * 8 * n / (ms * 1e-3) / 1e9


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
    ap.add_argument("--no-others", action="store_true", help="skip configurations 4 and 5")
    ap.add_argument("--no-design", action="store_true", help="skip the configuration-4 design gradient")
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
    bg, bf, ba = stage_bytes(N, F, nv, args.assembly, limiter, bool(p.turb), bool(p.turb) and not p.turb_first_order)
    stages = {"gradients": (bg, ms[0]), "face_flux": (bf, ms[1]), "assembly": (ba, ms[2])}
    dom = max(stages, key=lambda k: stages[k][1])
    b_dom, t_dom = stages[dom]
    B_res = N * (16 * nv + 24) + 48 * F
    achieved = B_res / (t_dom * 1e-3) / 1e9  # compulsory bytes of one residual over the dominant kernel's time
    ach_res = B_res / (wall / args.steps) / 1e9
    triad = triad_bandwidth()
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
                     "kernel": dom, "algorithmic_bytes_per_launch": B_res, "ms_per_launch": t_dom,
                     "achievable": triad, "frac_of_achievable": achieved / triad,
                     "achievable_rule": "stream triad a = b + s c, 3 x 512 MiB FP64, torch's elementwise kernel",
                     "bytes_rule": "B_res = N (16 n_v + 24) + 48 F: U read and R written once, cell volume and "
                                   "centroid, face indices and geometry"},
        "roofline_residual": {"bound": "hbm", "achieved": ach_res, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": ach_res / HBM_PEAK_GBS, "B_res": B_res,
                              "frac_of_achievable": ach_res / triad,
                              "ms_wall": wall / args.steps * 1e3, "ms_event": ms[3]},
        "stage_traffic": {k: {"bytes": b, "ms": t, "GB_per_s": b / (t * 1e-3) / 1e9,
                              "frac": b / (t * 1e-3) / 1e9 / HBM_PEAK_GBS}
                          for k, (b, t) in stages.items()},
        "stage_ms": {"gradients": ms[0], "face_flux": ms[1], "assembly": ms[2], "residual": ms[3]},
    }
    out["tangent_products"] = tangent_throughput(solver, Ud, nv, pitch, N)
    if not args.no_others and args.config == CONFIG and args.scale == 1.0:
        out["other_configs"] = {}
        for cfg in (4, 5):
            out["other_configs"][str(cfg)] = other_config(cfg, args.steps, args.warmup)
            print(f"bench: configuration {cfg}: {out['other_configs'][str(cfg)]['ms_per_residual']:.3f} ms per "
                  "residual", file=sys.stderr, flush=True)
    if not args.no_jfnk:
        out["jfnk"] = {"per_iteration_on_this_workload": gmres_iteration_cost(solver, Ud, Rd, nv, pitch),
                       "time_to_1e-10": jfnk_converge()}
        # configuration 2 as it is stated: GMRES(30), block-Jacobi, FD against the tangent, all FP64; with
        # two restarts per solve (without them both legs stall near a 1.5e-4 drop)
        out["jfnk"]["configuration_2_as_stated"] = jfnk_converge(
            2, 1.0, rel_drop=1e-10, max_steps=1500, m=30, restarts=2, precond=1,
            legs=(("fd_fp64", 0, 64), ("tangent_fp64", 1, 64)),
            label=WORKLOADS[2] + ", GMRES(30) with two restarts, block-Jacobi, FD vs tangent, all FP64")
        if args.config == CONFIG and args.scale == 1.0:
            ref = load_cf_reference(mesh)
            out["jfnk"]["time_to_1e-10_this_workload"] = jfnk_converge(
                CONFIG, 1.0, rel_drop=1e-10, max_steps=800, m=60, restarts=1,
                cf_ref=None if ref is None else ref["Cf"])
            if ref is not None:
                out["jfnk"]["time_to_1e-10_this_workload"]["cf_reference"] = (
                    f"tangent + FP64 Krylov to a {float(ref['drop']):.1e} drop in {int(ref['steps'])} Newton steps "
                    "(mycode_amd/data/make_sst_plate_ref.py)")
    if not args.no_design and args.config == CONFIG and args.scale == 1.0:
        out["design_gradient"] = design_gradient_leg(deadline=T_START + RUN_BUDGET_S)
    if not args.no_cpu:
        out["cpu_baseline"] = cb = cpu_baseline(mesh, p, U)
        f_all, f_cell = cb["fp64_ops_per_residual"]["flops"], cb["fp64_ops_cell_stage"]["flops"]

        def rate(f, t_ms):
            return {"flops": f, "ms": t_ms, "achieved": f / (t_ms * 1e-3) / 1e12,
                    "frac": f / (t_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}
        out["roofline_fp64"] = dict(
            bound="fp64", unit="TFLOP/s", peak=FP64_PEAK_TFLOPS, flops_per_residual=f_all,
            flops_per_cell=f_all / N, **{k: v for k, v in rate(f_all, wall / args.steps * 1e3).items()
                                         if k in ("achieved", "frac")},
            stages={"cells (gradients kernel)": rate(f_cell, ms[0]),
                    "faces (flux and assembly kernels)": rate(f_all - f_cell, ms[1] + ms[2])},
            count_rule="the reference algorithm's FP64 operations of one residual, counted by the CPU "
                       "restatement on an op-counting scalar (+ - * / sqrt one flop each); the GPU's "
                       "executed FP64 instructions are higher (division and square root expand)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

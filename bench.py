"""Residual throughput of the MI355X library on the laminar flat-plate case.

Workload (one "step" = one full second-order residual evaluation
R(U) = sum of Roe + viscous face fluxes, FP64, on device-resident state):
laminar flat plate, Mach 0.2, Re_L = 1e5, 400 x 160 stretched quads with a
seeded 20 % split into triangles (76,790 cells), plate boundary conditions.

Prints one JSON line:  value = Mcell-residuals/s (cells x residuals / s,
wall clock around K residuals, inputs already in HBM), plus
  roofline      the dominant kernel (face flux), algorithmic bytes of its
                inputs/outputs per launch over its HIP-event duration;
  roofline_residual   the whole residual against B_res = N(16 n_v + 24) + 48 F;
  cpu_baseline  the CPU restatement of the reference residual (g++ -O2,
                one core) on the same mesh and state.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--assembly 0|1]
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
CONFIG = 2


def stage_bytes(N, F, nv, assembly, limiter):
    """Compulsory bytes of each residual stage's own inputs and outputs."""
    grad = N * (8 * nv + 16 + 80 + (8 + 16 if limiter else 0)) + F * (8 + 24 + (16 if limiter else 0))
    flux = N * (8 * nv + 80 + 16 + (8 if limiter else 0)) + F * (12 + 40)
    if assembly == 0:
        flux += N * 16 * nv  # read-modify-write of R
        asm = N * 8 * nv      # zero fill
    else:
        flux += F * 8 * nv    # face buffer write
        asm = F * 8 * nv + N * (16 + 8 * nv)
    return grad, flux, asm


def cpu_baseline(mesh, p, U, budget_s=20.0):
    from oracle import oracle
    o = oracle.OracleSolver(oracle.OracleMesh.from_mesh(mesh), p)
    o.compute_residuals(U)  # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 5 and time.perf_counter() - t_start < budget_s:
        t0 = time.perf_counter()
        o.compute_residuals(U)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": mesh.n_cells / t / 1e6, "unit": "Mcell-residuals/s", "cores": 1, "kind": "port",
            "sample": f"{len(times)} residual evaluations of the same {mesh.n_cells}-cell mesh and state "
                      f"(median {t * 1e3:.2f} ms), oracle/cfd_oracle.cpp built g++ -O2"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--assembly", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", type=int, default=CONFIG, help="benchmark configuration to run (1-5)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU layout is design-only",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from mycode_amd import cfd, configs

    mesh, p, U0 = configs.setup(args.config)
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
    for _ in range(args.warmup):
        solver.dev_residual(Ud.data_ptr(), Rd.data_ptr(), False)
    solver.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ms, launches = solver.dev_time_residual(Ud.data_ptr(), Rd.data_ptr(), args.steps, False)
    solver.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # sanity: the device residual equals the host-pointer path
    R6 = torch.zeros(N * 6, dtype=torch.float64, device=dev)
    solver.dev_unpack(Rd.data_ptr(), R6.data_ptr())
    solver.synchronize()
    R_dev = R6.cpu().numpy().reshape(N, 6)
    R_host = solver.compute_residuals(U, False)
    assert np.array_equal(R_dev, R_host), "device and host residual paths disagree"

    limiter = p.spatial == cfd.SECOND_ORDER_LIMITED
    bg, bf, ba = stage_bytes(N, F, nv, args.assembly, limiter)
    stages = {"gradients": (bg, ms[0]), "face_flux": (bf, ms[1]), "assembly": (ba, ms[2])}
    dom = max(stages, key=lambda k: stages[k][1])
    b_dom, t_dom = stages[dom]
    achieved = b_dom / (t_dom * 1e-3) / 1e9
    B_res = N * (16 * nv + 24) + 48 * F
    ach_res = B_res / (ms[3] * 1e-3) / 1e9
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
        "config": {"workload": "laminar flat plate M0.2 Re_L=1e5, 400x160 stretched quads 20% split, "
                               "Roe + viscous, 2nd-order central, FP64 residual",
                   "cells": N, "faces": F, "nvar": nv, "assembly": ["colored-scatter", "face-buffer-gather"][
                       args.assembly], "flux_launches_per_residual": launches, "mesh_hash": hex(mesh.hash())},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                     "algorithmic_bytes_per_residual": b_dom, "ms_per_residual": t_dom},
        "roofline_residual": {"bound": "hbm", "achieved": ach_res, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": ach_res / HBM_PEAK_GBS, "B_res": B_res, "ms_event": ms[3]},
        "stage_ms": {"gradients": ms[0], "face_flux": ms[1], "assembly": ms[2], "residual": ms[3]},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, p, U)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

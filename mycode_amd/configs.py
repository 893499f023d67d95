"""The five benchmark configurations: mesh, parameters and initial state.

Each follows the reference's own set-up calls — ``setFlowConditions``
(Solver.cpp:12-32), ``setKinematicViscosity`` (Solver.cpp:34-43),
``setNumericalMethod`` (Solver.hpp:33-36) and the boundary-condition set the
case needs — applied to a seeded synthetic mesh (``cfd_mesh_generate``).

  1  Euler, Mach 2.5 over a 15-degree wedge, 160 x 80 quads all split,
     HLLC + Barth-Jespersen, supersonic BC set, SSP-RK3.
  2  laminar flat plate, Mach 0.2, Re_L = 1e5, 400 x 160 (20 % split),
     Roe + viscous, second-order central, plate BC set, JFNK-GMRES(30).
  3  turbulent flat plate, Mach 0.2, Re = 5e6, 1200 x 1000 (10 % split),
     Roe + viscous + SST k-omega (six conserved variables), freestream
     k = 9e-9 a^2 and omega = 1e-6 a^2 / nu (mu_t / mu = 0.009, the NASA
     Turbulence Modeling Resource's flat-plate values).
  4  Mach 0.7 inviscid flow over a 5 % circular-arc bump in a 3 x 1 channel,
     1280 x 560 (15 % split), HLLC central, bump BC set.
  5  Mach 2 laminar shock / boundary-layer interaction, 2000 x 2000 (10 %
     split), LLF in the incident-shock band and HLLC elsewhere, BJ limiter.
"""
from __future__ import annotations

import math

import numpy as np

from . import cfd

SEED_BASE = 0x5EED0000
MASK64 = (1 << 64) - 1


def splitmix64_stream(seed: int, n: int) -> np.ndarray:
    """First n outputs of SplitMix64 (Vigna) started at `seed`, as uint64."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform_stream(seed: int, n: int) -> np.ndarray:
    return (splitmix64_stream(seed, n) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def oblique_shock(M: float, theta_deg: float, gamma: float = 1.4):
    """Weak-branch oblique shock: (beta, p2/p1, rho2/rho1, M2)."""
    th = math.radians(theta_deg)

    def f(b):
        return 2.0 / math.tan(b) * (M * M * math.sin(b) ** 2 - 1.0) / (M * M * (gamma + math.cos(2 * b)) + 2.0) \
            - math.tan(th)

    lo, hi = math.asin(1.0 / M) + 1e-9, math.pi / 2
    grid = [lo + (hi - lo) * k / 2000.0 for k in range(2001)]
    bmax = max(grid, key=f)
    a, b = lo, bmax
    for _ in range(200):
        mid = 0.5 * (a + b)
        if f(mid) < 0:
            a = mid
        else:
            b = mid
    beta = 0.5 * (a + b)
    Mn1 = M * math.sin(beta)
    p21 = 1.0 + 2.0 * gamma / (gamma + 1.0) * (Mn1 * Mn1 - 1.0)
    r21 = (gamma + 1.0) * Mn1 * Mn1 / ((gamma - 1.0) * Mn1 * Mn1 + 2.0)
    Mn2 = math.sqrt((1.0 + 0.5 * (gamma - 1.0) * Mn1 * Mn1) / (gamma * Mn1 * Mn1 - 0.5 * (gamma - 1.0)))
    M2 = Mn2 / math.sin(beta - th)
    return beta, p21, r21, M2


def freestream_state(p: cfd.Params) -> np.ndarray:
    """Conservative freestream of the frozen inflow (Solver.cpp:1094-1100)."""
    g = p.gamma
    fT = 1.0 + 0.5 * (g - 1.0) * p.M_inf * p.M_inf
    Ts = p.T_total_inlet / fT
    rho = p.P_static_exit / (p.R_gas * Ts)
    u = p.M_inf * math.sqrt(g * p.R_gas * Ts)
    k, w = (p.turb_k_inf, p.turb_w_inf) if p.turb else (0.0, 0.0)
    return np.array([rho, rho * u, 0.0, p.P_static_exit / (g - 1.0) + 0.5 * rho * u * u, rho * k, rho * w])


def params_for(config: int) -> cfd.Params:
    p = cfd.default_params()
    if config == 1:
        cfd.flow_conditions(p, 2.5, 101325.0, 300.0)
        p.bcset, p.flux, p.spatial, p.viscous, p.nvar = cfd.BC_SUPERSONIC, cfd.HLLC, cfd.SECOND_ORDER_LIMITED, 0, 4
    elif config == 2:
        cfd.flow_conditions(p, 0.2, 101325.0, 300.0)
        cfd.kinematic_viscosity(p, 1e5, 1.0)
        p.bcset, p.flux, p.spatial, p.viscous, p.nvar = cfd.BC_PLATE, cfd.ROE, cfd.SECOND_ORDER_CENTRAL, 1, 4
    elif config == 3:
        cfd.flow_conditions(p, 0.2, 101325.0, 300.0)
        cfd.kinematic_viscosity(p, 5e6, 1.0)
        p.bcset, p.flux, p.spatial, p.viscous, p.nvar = cfd.BC_PLATE, cfd.ROE, cfd.SECOND_ORDER_CENTRAL, 1, 6
        a2 = p.gamma * p.R_gas * p.T_total_inlet / (1.0 + 0.5 * (p.gamma - 1.0) * p.M_inf ** 2)
        p.turb, p.turb_k_inf, p.turb_w_inf = 1, 9e-9 * a2, 1e-6 * a2 / p.nu
        # k and omega convected first-order upwind: the second-order
        # reconstruction carries the wall omega (~1e8 1/s) into the upwind
        # faces of the cells just ahead of the leading edge, pins their omega
        # at its floor and stalls the Newton iteration near a 3e-6 drop
        p.turb_first_order = 1
    elif config == 4:
        cfd.flow_conditions(p, 0.7, 101325.0, 300.0)
        p.bcset, p.flux, p.spatial, p.viscous, p.nvar = cfd.BC_BUMP, cfd.HLLC, cfd.SECOND_ORDER_CENTRAL, 0, 4
    elif config == 5:
        cfd.flow_conditions(p, 2.0, 101325.0, 300.0)
        cfd.kinematic_viscosity(p, 3e5, 0.5)
        p.bcset, p.flux, p.flux_alt, p.spatial, p.viscous, p.nvar = (
            cfd.BC_SWBLI, cfd.HLLC, cfd.LLF, cfd.SECOND_ORDER_LIMITED, 1, 4)
        fs = freestream_state(p)
        rho1, u1 = fs[0], fs[1] / fs[0]
        beta, p21, r21, M2 = oblique_shock(2.0, 8.0, p.gamma)
        p2, rho2 = p.P_static_exit * p21, rho1 * r21
        c2 = math.sqrt(p.gamma * p2 / rho2)
        V2 = M2 * c2
        th = math.radians(8.0)
        p.shock_rho, p.shock_u, p.shock_v, p.shock_p = rho2, V2 * math.cos(th), -V2 * math.sin(th), p2
    else:
        raise ValueError(f"unknown configuration {config}")
    return p


def initial_state(config: int, mesh: cfd.Mesh, p: cfd.Params) -> np.ndarray:
    """Seeded initial state in reference cell order, shape (N, 6)."""
    N = mesh.n_cells
    if config == 4:
        U = np.tile(cfd.initial_state(p), (N, 1))  # Solver::initialize (Solver.cpp:46-64)
        return U
    fs = freestream_state(p)
    U = np.tile(fs, (N, 1))
    if config == 1:  # 1 % uniform-random density perturbation, stream 2
        u = uniform_stream(SEED_BASE + 16 * config + 2, N)
        g = p.gamma
        rho = fs[0] * (1.0 + 0.01 * (2.0 * u - 1.0))
        vel = fs[1] / fs[0]
        U[:, 0] = rho
        U[:, 1] = rho * vel
        U[:, 2] = 0.0
        U[:, 3] = p.P_static_exit / (g - 1.0) + 0.5 * rho * vel * vel
    return U


def shock_band(mesh: cfd.Mesh, width: float = 0.03) -> np.ndarray:
    """Configuration 5: faces within `width` of the incident shock line use LLF."""
    f = mesh.faces()
    beta = oblique_shock(2.0, 8.0)[0]
    # shock line through (0.5, 0) rising upstream at angle beta
    nx, ny = math.sin(beta), math.cos(beta)
    d = (f["x_mid"] - 0.5) * nx + f["y_mid"] * ny
    return (np.abs(d) < width).astype(np.uint8)


def setup(config: int, scale: float = 1.0):
    """(mesh, params, U0) of a configuration, optionally coarsened by `scale`."""
    mesh = cfd.Mesh.generate(config, scale)
    p = params_for(config)
    return mesh, p, initial_state(config, mesh, p)

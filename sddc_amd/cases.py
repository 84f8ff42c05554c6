"""The workloads of this path, built from the mesh generators.

  cylinder(...)     compressible MHD flow past an adiabatic cylinder on an
                    O-grid (WALL at r0, INLE/OUTL far field)
  orszag_tang(...)  periodic Orszag-Tang vortex on [0,2pi]^2

The reference has no CFL logic; dt is fixed (readinput.f90:23-24).  Here
it is computed once on the host from the initial state as
dt = CFL * h_min / ((|u| + c_f)_max * (2p+1)), with h_min the shortest
cell edge and c_f the fast magnetosonic speed, and then kept fixed.
"""
from __future__ import annotations

import math

import numpy as np

from . import mesh as M


def fast_speed_max(Q: np.ndarray, gam: float = M.GAM) -> float:
    """max over points of |u| + c_f (fast magnetosonic speed)."""
    rho = Q[0]
    u = Q[1] / rho; v = Q[2] / rho; w = Q[3] / rho
    B2 = Q[5] ** 2 + Q[6] ** 2 + Q[7] ** 2
    p = (gam - 1) * (Q[4] - 0.5 * rho * (u * u + v * v + w * w) - 0.5 * B2 - 0.5 * Q[8] ** 2)
    a2 = gam * p / rho
    b2 = B2 / rho
    cf = np.sqrt(a2 + b2)          # upper bound of c_f in any direction
    return float(np.max(np.sqrt(u * u + v * v + w * w) + cf))


def min_edge(m: M.Mesh) -> float:
    c = m.cells - 1
    h = np.inf
    for a, b in M.IVFACE:
        d = np.hypot(m.xv[c[:, b]] - m.xv[c[:, a]], m.yv[c[:, b]] - m.yv[c[:, a]])
        h = min(h, float(d.min()))
    return h


def cfl_dt(case: M.Case, Q: np.ndarray, cfl: float) -> float:
    p = case.N - 1
    return cfl * min_edge(case.mesh) / (fast_speed_max(Q) * (2 * p + 1))


def cylinder(nr: int, ntheta: int, N: int, r1: float = 15.0, mach: float = 1.5,
             alfven_mach: float | None = None, b_angle_deg: float = 0.0,
             bmag: float = 0.3, rho_pert: float = 0.0, seed: int = 0,
             cfl: float = 0.3, k_stage: int = 3, farfield: str = "split",
             ramp: float | None = None):
    """Cylinder case.  With alfven_mach given, |B| = mach * sqrt(rho) /
    alfven_mach (c_s = 1) at angle b_angle_deg; otherwise |B| = bmag along x.
    Returns (case, Q0)."""
    gam = M.GAM
    rinf, pinf = 1.0, 1.0 / gam
    u = mach * math.sqrt(gam * pinf / rinf)
    if alfven_mach is not None:
        bmag = u * math.sqrt(rinf) / alfven_mach
    th = math.radians(b_angle_deg)
    p = M.RunParams(N=N, k_stage=k_stage, rinf=rinf, pinf=pinf, uinf=u, vinf=0.0,
                    bxinf=bmag * math.cos(th), byinf=bmag * math.sin(th))
    m = M.ogrid(nr, ntheta, 0.5, r1, farfield=farfield)
    case = M.build_case(m, p)
    Q0 = M.freestream_state(case, rho_pert=rho_pert, seed=seed,
                            wall_ramp=None if ramp is None else (0.5, ramp))
    case.params.dt = cfl_dt(case, Q0, cfl)
    return case, Q0


# An impulsive start (uniform Mach 1.5 free stream against the wall)
# drives the rear stagnation point towards vacuum within ~5e-4 time units;
# the unlimited p=3/p=4 scheme then undershoots to negative pressure and the
# state turns NaN within a few steps at these resolutions (reference and
# this path alike).  The workloads therefore start at rest on the wall with
# the velocity rising over one cylinder radius (RAMP).
RAMP = 0.5


def c1_cylinder(**kw):
    """32 x 64 O-grid, r in [0.5, 15], p=3, Mach 1.5, B=(0.3,0,0)."""
    kw.setdefault("ramp", RAMP)
    return cylinder(32, 64, 4, **kw)


def c2_cylinder(**kw):
    """512 x 1024 O-grid, r in [0.5, 30], p=4, Mach 1.5, Alfven Mach 4 at 30
    degrees, seeded 1e-3 uniform density perturbation."""
    kw.setdefault("seed", 2)
    kw.setdefault("ramp", RAMP)
    return cylinder(512, 1024, 5, r1=30.0, alfven_mach=4.0, b_angle_deg=30.0,
                    rho_pert=1e-3, **kw)


def orszag_tang(n: int, N: int = 4, rho_pert: float = 1e-4, seed: int = 3,
                cfl: float = 0.3, k_stage: int = 3):
    """Periodic Orszag-Tang vortex on n x n quads (C3 uses n=1024, p=3)."""
    m = M.periodic_square(n, n)
    p = M.RunParams(N=N, k_stage=k_stage, dxcycl=m.dxcycl, dycycl=m.dycycl,
                    rinf=25.0 / 9.0, pinf=5.0 / 3.0, uinf=0.0, vinf=0.0,
                    bxinf=0.0, byinf=0.0)
    case = M.build_case(m, p)
    Q0 = M.orszag_tang_state(case, rho_pert=rho_pert, seed=seed)
    case.params.dt = cfl_dt(case, Q0, cfl)
    return case, Q0


def c4_ensemble(nmem: int = 64, seed: int = 4, r1: float = 15.0, ramp: float = 2.0,
                cfl: float = 0.15):
    """The C4 ensemble: nmem cylinder runs with Nr x Ntheta drawn from
    {32x64, 64x128, 128x256, 256x512}, p from {2,3,4} (N = p+1), free-stream
    Mach from U[1.2, 3.0], Alfven Mach from U[2, 8] and field angle from
    U[0, 90] degrees, all from one seed.  Every member starts at rest on the
    wall with the velocity rising over `ramp` radii and steps at CFL `cfl`.
    In FP64 all 64 members of seed 4 stay finite over 500 steps at CFL 0.15
    (and at 0.1 and 0.075, and with a ramp over 4 radii); in the mixed
    precision form 8 to 11 of the draws with the strongest field and the
    highest Mach number (Mach 1.6-2.9, Alfven Mach 2.2-3.4) turn non-finite
    within 125 to 500 steps, at each of those CFL numbers and ramps.
    Meshes and metrics are shared between members of equal size and order.
    Returns a list of (case, Q0, draw) with draw the drawn parameters."""
    import dataclasses
    rng = np.random.default_rng(seed)
    sizes = [(32, 64), (64, 128), (128, 256), (256, 512)]
    base = {}
    out = []
    for i in range(nmem):
        nr, nt = sizes[int(rng.integers(0, 4))]
        N = int(rng.integers(2, 5)) + 1
        mach = float(rng.uniform(1.2, 3.0))
        ma = float(rng.uniform(2.0, 8.0))
        ang = float(rng.uniform(0.0, 90.0))
        key = (nr, nt, N)
        if key not in base:
            base[key] = M.build_case(M.ogrid(nr, nt, 0.5, r1), M.RunParams(N=N))
        gam = M.GAM
        rinf, pinf = 1.0, 1.0 / gam
        u = mach * math.sqrt(gam * pinf / rinf)
        bmag = u * math.sqrt(rinf) / ma
        th = math.radians(ang)
        p = M.RunParams(N=N, rinf=rinf, pinf=pinf, uinf=u, vinf=0.0,
                        bxinf=bmag * math.cos(th), byinf=bmag * math.sin(th))
        case = dataclasses.replace(base[key], params=p)
        Q0 = M.freestream_state(case, wall_ramp=(0.5, ramp))
        case.params.dt = cfl_dt(case, Q0, cfl)
        out.append((case, Q0, {"nr": nr, "ntheta": nt, "N": N, "mach": mach,
                               "alfven_mach": ma, "b_angle_deg": ang}))
    return out


def dof(case: M.Case) -> int:
    return case.ncell * case.N * case.N


def alg_bytes_per_dof_stage(N: int, k_stage: int = 3, metric_bytes: int = 8) -> float:
    """Compulsory HBM bytes per DOF-stage of a fused stage: the RK state
    traffic (RK_S5P3: 13 state passes of 72 B over 5 stages; FiveStageRk:
    18 + the stored residual) plus Jac and the face-normal metric vectors
    S1(1,:), S2(2,:) read once per stage."""
    state = 13 * 72 / 5 if k_stage == 3 else 18 * 72 / 5
    metrics = metric_bytes * (N * N + 4 * N * (N + 1)) / (N * N)
    return state + metrics

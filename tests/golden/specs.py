"""Recipes of the golden fixtures: how each fixture's case is built.

Shared by make_golden.py (which runs the reference on them) and the tests
(which rebuild the same cases and compare against the stored outputs).
"""
from __future__ import annotations

import numpy as np

from sddc_amd import mesh as M

# name: (mesh recipe, run parameters, initial state recipe, staged steps,
#        whole RK steps afterwards, store full stage arrays)
SPECS = {
    "ogrid4x8_N4": (dict(kind="ogrid", nr=4, nt=8, r0=0.5, r1=15.0),
                    dict(N=4, dt=1e-3), dict(kind="smooth", amp=0.05, seed=1), 2, 3),
    "channel5x4_N3": (dict(kind="channel", nx=5, ny=4, seed=3, free_left=2),
                      dict(N=3, dt=1e-3), dict(kind="smooth", amp=0.05, seed=1), 2, 3),
    "periodic4x4_N5": (dict(kind="periodic", nx=4, ny=4),
                       dict(N=5, dt=1e-3), dict(kind="smooth", amp=0.05, seed=1), 2, 3),
    "channel7x6_N5": (dict(kind="channel", nx=7, ny=6, seed=11, free_left=3,
                           types=["WALL", "INLE", "SYMP", "OUTL"]),
                      dict(N=5, dt=5e-4), dict(kind="smooth", amp=0.05, seed=1), 1, 2),
    "viscous_channel6x5_N4": (dict(kind="channel", nx=6, ny=5, seed=7, free_left=2),
                              dict(N=4, dt=5e-4, mu=0.01, eta0=0.02),
                              dict(kind="smooth", amp=0.05, seed=2), 1, 2),
    # the viscous/resistive path on a periodic mesh: BR1 gradient averaging
    # across the CYCL pairs (cyclocflux_all.f90:114-204); 5 x 4 cells so that
    # the x- and y-pairs differ in number
    "viscous_periodic5x4_N3": (dict(kind="periodic", nx=5, ny=4),
                               dict(N=3, dt=5e-4, mu=0.01, eta0=0.02),
                               dict(kind="smooth", amp=0.05, seed=3), 2, 2),
    "fsrk_ogrid4x8_N4": (dict(kind="ogrid", nr=4, nt=8, r0=0.5, r1=15.0),
                         dict(N=4, dt=1e-3, k_stage=5), dict(kind="smooth", amp=0.05, seed=4), 0, 3),
    # 8-node cells (CURVE_WALL = 1, mapder.f90:91-253): an O-grid whose arc
    # edges follow their circles through a midside node; the two cells of a
    # face form its normal in different orders, so the normals differ by
    # rounding and the device takes the left cell's (interfaceflux_all.f90:47)
    "ogrid_curved4x8_N4": (dict(kind="ogrid", nr=4, nt=8, r0=0.5, r1=15.0, curved=True),
                           dict(N=4, dt=1e-3, curve_wall=1),
                           dict(kind="smooth", amp=0.05, seed=6), 2, 3),
    "viscous_ogrid_curved3x10_N3": (dict(kind="ogrid", nr=3, nt=10, r0=0.5, r1=4.0, curved=True),
                                    dict(N=3, dt=5e-4, mu=0.01, eta0=0.02, curve_wall=1),
                                    dict(kind="smooth", amp=0.05, seed=7), 1, 2),
    "c1_cylinder": (dict(kind="ogrid", nr=32, nt=64, r0=0.5, r1=15.0),
                    dict(N=4, dt=5e-4),
                    dict(kind="freestream", uv_pert=0.02, seed=5, ramp=0.5), 1, 2),
}


def make_mesh(spec: dict) -> M.Mesh:
    s = dict(spec)
    kind = s.pop("kind")
    if kind == "ogrid":
        return M.ogrid(s["nr"], s["nt"], s["r0"], s["r1"], curved=s.get("curved", False))
    if kind == "channel":
        if "types" in s:
            s["types"] = tuple(s["types"])
        return M.channel(**s)
    if kind == "periodic":
        return M.periodic_square(s["nx"], s["ny"])
    raise ValueError(kind)


def make_state(case: M.Case, spec: dict) -> np.ndarray:
    if spec["kind"] == "smooth":
        return M.smooth_perturbed_state(case, spec["amp"], seed=spec["seed"])
    if spec["kind"] == "freestream":
        ramp = spec.get("ramp")
        return M.freestream_state(case, uv_pert=spec.get("uv_pert", 0.0),
                                  rho_pert=spec.get("rho_pert", 0.0), seed=spec["seed"],
                                  wall_ramp=None if ramp is None else (0.5, ramp))
    raise ValueError(spec["kind"])


def build(name: str):
    mspec, pspec, ispec, _, _ = SPECS[name]
    m = make_mesh(mspec)
    p = M.RunParams(dxcycl=m.dxcycl, dycycl=m.dycycl, **pspec)
    case = M.build_case(m, p)
    return case, make_state(case, ispec)


def load(name: str):
    """(case, Q0, fixture dict) of a golden fixture."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), name + ".npz")
    fx = dict(np.load(path, allow_pickle=False))
    case, Q0 = build(name)
    return case, Q0, fx

"""Shared helpers of the test-suite: error norms and the standard cases."""
from __future__ import annotations

import numpy as np

from sddc_amd import mesh as M


def per_var_error(a, b, nvar_axis=0):
    """Per-variable normwise relative error ||a-b||_inf / ||b||_inf, with
    the reference's small variables (||b_v|| < 1e-10 max_v ||b_v||) measured
    against the largest variable instead.  Returns (errors, scales)."""
    a = np.moveaxis(np.asarray(a), nvar_axis, 0).reshape(np.shape(a)[nvar_axis], -1)
    b = np.moveaxis(np.asarray(b), nvar_axis, 0).reshape(a.shape[0], -1)
    sc = np.abs(b).max(axis=1)
    big = sc.max()
    den = np.where(sc < 1e-10 * big, big, sc)
    err = np.abs(a - b).max(axis=1) / np.where(den == 0, 1.0, den)
    return err, sc


CASES = {
    # name: (mesh factory, RunParams kwargs, IC kind)
    "ogrid4x8_N4": (lambda: M.ogrid(4, 8, 0.5, 15.0), dict(N=4, dt=1e-3), "smooth"),
    "channel5x4_N3": (lambda: M.channel(5, 4, seed=3, free_left=2), dict(N=3, dt=1e-3), "smooth"),
    "periodic4x4_N5": (lambda: M.periodic_square(4, 4), dict(N=5, dt=1e-3), "smooth"),
    "channel7x6_N5": (lambda: M.channel(7, 6, seed=11, free_left=3,
                                        types=("WALL", "INLE", "SYMP", "OUTL")),
                      dict(N=5, dt=5e-4), "smooth"),
    "ogrid8x16_N3": (lambda: M.ogrid(8, 16, 0.5, 15.0), dict(N=3, dt=1e-3), "smooth"),
    "periodic6x5_N4": (lambda: M.periodic_square(6, 5, 2.0, 1.5), dict(N=4, dt=1e-3), "smooth"),
    # rows of 193, 241 and 321 cells: strips padded to the mesh rows (the
    # tiles of one row line up with the next, padding slots inside the mesh)
    "ogrid2x193_N5": (lambda: M.ogrid(2, 193, 5.0, 6.0), dict(N=5, dt=5e-4), "smooth"),
    "periodic241x2_N4": (lambda: M.periodic_square(241, 2, 24.1, 0.2), dict(N=4, dt=1e-3), "smooth"),
    "ogrid2x321_N3": (lambda: M.ogrid(2, 321, 5.0, 5.6), dict(N=3, dt=5e-4), "smooth"),
    # 8-node cells (CURVE_WALL = 1): the two cells of a face form its normal
    # in different orders, so the device takes the left cell's
    "ogrid_curved4x8_N4": (lambda: M.ogrid(4, 8, 0.5, 15.0, curved=True),
                           dict(N=4, dt=1e-3, curve_wall=1), "smooth"),
    "ogrid_curved5x12_N5": (lambda: M.ogrid(5, 12, 0.5, 6.0, curved=True),
                            dict(N=5, dt=5e-4, curve_wall=1), "smooth"),
}


# device-test order: lowest order first
ORDERED = sorted(CASES, key=lambda n: (int(n.split("_N")[1]), n))


def make_case(name, **over):
    mk, kw, ic = CASES[name]
    m = mk()
    kw = dict(kw)
    kw.update(over)
    p = M.RunParams(dxcycl=m.dxcycl, dycycl=m.dycycl, **kw)
    case = M.build_case(m, p)
    Q0 = M.smooth_perturbed_state(case, 0.05, seed=1)
    return case, Q0

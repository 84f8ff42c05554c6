"""The mixed-precision stepping path (SD2D_MIXED): interpolation, fluxes,
Rusanov and derivatives in FP32; state, Jacobian scaling, sources, RK
combination, norms and AV sensor in FP64.

Bars (stated, not bit-level: FP32 rounding of the fluxes is amplified by the
derivative operators by ~N^2 relative to the flux size):
  * three RK steps against the FP64 oracle: per variable
    ||dQ_mixed - dQ_oracle||_inf / ||dQ_oracle||_inf <= 5e-5 for the state
    change, running maxima to 1e-6 relative;
  * a smooth periodic Orszag-Tang vortex (16 x 16, N=4) over 200 steps:
    relative L2 difference of every variable from the FP64 device run
    <= 1e-6 (absolute against the largest variable for w, Bz, psi, which
    start at zero).
"""
import numpy as np
import pytest

from sddc_amd import abi, cases
from tests.util import ORDERED, make_case, per_var_error

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("k_stage", [3, 5])
@pytest.mark.parametrize("name", ORDERED)
def test_mixed_steps_against_oracle(name, k_stage):
    from sddc_amd.solver import Sd2d
    from oracle.oracle import Oracle
    print(f"[gpu-mixed] steps {name} k_stage={k_stage}", flush=True)
    case, Q0 = make_case(name, k_stage=k_stage)
    g = Sd2d(case, precision=abi.SD2D_MIXED)
    o = Oracle(case)
    g.set_state(Q0, [0, 0, 0, 0]); o.set_state(Q0, [0, 0, 0, 0])
    g.step(3, 2); o.step(3, 2)
    Qg, mg = g.get_state()
    Qo, mo = o.get_state()
    e = per_var_error(Qg - Q0, Qo - Q0)[0]
    assert e.max() <= 5e-5, (name, e)
    np.testing.assert_allclose(mg, mo, rtol=1e-6)
    dg, do = g.diag(), o.diag()
    np.testing.assert_allclose([dg["resid_rms"], dg["divb_rms"]],
                               [do["resid_rms"], do["divb_rms"]], rtol=1e-4)


@pytest.mark.parametrize("name", ["ogrid4x8_N4", "channel7x6_N5", "periodic6x5_N4"])
def test_mixed_av_sensor(name):
    """The mixed kernel's AV sensor (FP64 rho and Jz^2 handed between the
    lanes of one wave through LDS): its non-finite cells are the oracle's,
    and it is bitwise the same in a second run and with another tile shape
    (an LDS hand-over that raced would not repeat).  Its values follow the
    mixed state, whose FP32 fluxes move small smoothness differences, so
    they are compared with the oracle's only loosely (log10 within 4)."""
    from sddc_amd.solver import Sd2d
    from oracle.oracle import Oracle
    print(f"[gpu-mixed] AV sensor {name}", flush=True)
    case, Q0 = make_case(name)
    res = []
    for shape in (abi.SD2D_TILE_ROWPAIRS, abi.SD2D_TILE_ROWPAIRS, abi.SD2D_TILE_STRIPS):
        g = Sd2d(case, precision=abi.SD2D_MIXED, tile_shape=shape)
        g.set_state(Q0, [0, 0, 0, 0])
        g.step(3, 2)
        res.append(np.asarray(g.get_av()[0]))
        g.close()
    for other in res[1:]:
        assert np.array_equal(res[0], other, equal_nan=True), name
    o = Oracle(case)
    o.set_state(Q0, [0, 0, 0, 0])
    o.step(3, 2)
    so = np.asarray(o.get_av()[0])
    o.close()
    fin = np.isfinite(so)
    assert np.array_equal(np.isfinite(res[0]), fin), name
    assert np.abs(res[0][fin] - so[fin]).max() <= 4.0, name


def _l2_rel(a, b):
    """Per-variable relative L2 difference; w, Bz and psi (zero initially
    and O(rounding) throughout in 2-D) against the largest variable's norm."""
    a = a.reshape(9, -1, order="F"); b = b.reshape(9, -1, order="F")
    nb = np.sqrt((b * b).sum(axis=1))
    den = nb.copy()
    den[[3, 7, 8]] = nb.max()
    return np.sqrt(((a - b) ** 2).sum(axis=1)) / den


def test_mixed_drift_orszag_tang():
    from sddc_amd.solver import Sd2d
    print("[gpu-mixed] Orszag-Tang 16x16 N=4, 200 steps", flush=True)
    case, Q0 = cases.orszag_tang(16, N=4)
    out = {}
    for prec in (abi.SD2D_FP64, abi.SD2D_MIXED):
        g = Sd2d(case, precision=prec)
        g.set_state(Q0, [0, 0, 0, 0])
        g.step(200)
        out[prec] = g.get_state()[0]
        g.close()
    d = _l2_rel(out[abi.SD2D_MIXED], out[abi.SD2D_FP64])
    print(f"[gpu-mixed] relative L2 drift per variable: {np.array2string(d, precision=2)}",
          flush=True)
    assert np.all(np.isfinite(out[abi.SD2D_MIXED]))
    assert d.max() <= 1e-6, d

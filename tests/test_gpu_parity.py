"""Device path against the oracle (the CPU restatement of the reference,
bit-exact to it on the golden fixtures) on the same seeded inputs.

Bar (FP64): per variable ||r_gpu - r_oracle||_inf / ||r_oracle||_inf
<= 1e-12 for the residual and nablaQs of every stage; running maxima and
c_h to 1e-14 relative (the device forms FMAs, so not bit for bit).
"""
import sys

import numpy as np
import pytest

from tests.util import ORDERED, make_case, per_var_error

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _gpu():
    from sddc_amd.solver import Sd2d
    return Sd2d


def _oracle():
    from oracle.oracle import Oracle
    return Oracle


@pytest.mark.parametrize("name", ORDERED)
def test_compresid_stage_parity(name):
    print(f"[gpu-parity] compresid {name}", flush=True)
    case, Q0 = make_case(name)
    g = _gpu()(case)
    o = _oracle()(case)
    g.set_state(Q0, [0, 0, 0, 0]); o.set_state(Q0, [0, 0, 0, 0])
    for it in range(3):          # c_h = 0 on the first call, lagged after
        rg, ng = g.compresid()
        ro, no = o.compresid()
        e, _ = per_var_error(rg, ro)
        assert e.max() <= TOL, (name, it, e)
        e2, _ = per_var_error(ng.reshape(18, -1, order="F"), no.reshape(18, -1, order="F"))
        assert e2.max() <= TOL, (name, it, e2)
        _, mg = g.get_state()
        _, mo = o.get_state()
        assert np.allclose(mg, mo, rtol=1e-14, atol=0), (mg, mo)


@pytest.mark.parametrize("k_stage", [3, 5])
@pytest.mark.parametrize("name", ORDERED)
def test_rk_steps_parity(name, k_stage):
    print(f"[gpu-parity] steps {name} k_stage={k_stage}", flush=True)
    case, Q0 = make_case(name, k_stage=k_stage)
    g = _gpu()(case)
    o = _oracle()(case)
    g.set_state(Q0, [0, 0, 0, 0]); o.set_state(Q0, [0, 0, 0, 0])
    g.step(3); o.step(3)
    Qg, mg = g.get_state()
    Qo, mo = o.get_state()
    # the state itself: increments of 3 steps relative to the state change
    d = per_var_error(Qg - Q0, Qo - Q0)[0]
    assert d.max() <= 1e-11, (name, d)
    assert np.allclose(mg, mo, rtol=1e-13, atol=0), (mg, mo)


# The AV smoothness sensor divides a squared small difference by a squared
# value (AV.f90:65-81), so rounding differences of 1e-15 in its inputs move
# log10(Se) by up to ~1e-8; it is a diagnostic with no effect on the flow.
SM_TOL = 1e-6


@pytest.mark.parametrize("name", ORDERED)
def test_av_sensor_parity(name):
    """SM_IND, muAV and epsilon0 after a compresid call and after whole
    steps (the sensor of the last stage) match the oracle's."""
    print(f"[gpu-parity] AV sensor {name}", flush=True)
    case, Q0 = make_case(name)
    g = _gpu()(case)
    o = _oracle()(case)
    g.set_state(Q0, [0, 0, 0, 0]); o.set_state(Q0, [0, 0, 0, 0])
    for what in ("compresid", "steps"):
        if what == "compresid":
            g.compresid(); o.compresid()
        else:
            g.step(2); o.step(2)
        sg, mg, eg = g.get_av()
        so, mo = o.get_av()
        so = np.asarray(so); mo = np.asarray(mo)
        fin = np.isfinite(so)
        assert np.array_equal(np.isfinite(sg), fin), (name, what)
        assert np.abs(sg[fin] - so[fin]).max() <= SM_TOL, (name, what)
        np.testing.assert_allclose(mg, mo, rtol=1e-6, atol=1e-12 * max(np.abs(mo).max(), 1e-300))
        np.testing.assert_allclose(eg, o.diag()["epsilon0"], rtol=1e-14)


@pytest.mark.parametrize("name", ["ogrid4x8_N4", "channel7x6_N5"])
def test_norms_parity(name):
    """Residual and divB RMS norms of iterations.f90:34-62 from sd2d_step
    (steps with mod(iter,10) == 1) match the oracle's."""
    print(f"[gpu-parity] norms {name}", flush=True)
    case, Q0 = make_case(name)
    g = _gpu()(case)
    o = _oracle()(case)
    g.set_state(Q0, [0, 0, 0, 0]); o.set_state(Q0, [0, 0, 0, 0])
    g.step(11, 10); o.step(11, 10)
    dg, do = g.diag(), o.diag()
    assert dg["norms_valid"] == 1 and dg["iter"] == 11
    np.testing.assert_allclose([dg["resid_rms"], dg["divb_rms"]],
                               [do["resid_rms"], do["divb_rms"]], rtol=1e-10)
    t_o, it_o = o.get_time()
    assert it_o == 11
    np.testing.assert_allclose(dg["ctime"], t_o, rtol=1e-15)


def test_determinism():
    """Two runs of the same steps give bitwise the same state, maxima and
    norms (maxima by atomicMax on bit patterns, norms summed in a fixed
    order)."""
    print("[gpu-parity] determinism", flush=True)
    case, Q0 = make_case("channel7x6_N5")
    res = []
    for _ in range(2):
        g = _gpu()(case)
        g.set_state(Q0, [0, 0, 0, 0])
        g.step(3, 2)
        Q, mx = g.get_state()
        d = g.diag()
        res.append((Q, mx, d["resid_rms"], d["divb_rms"]))
        g.close()
    assert np.array_equal(res[0][0], res[1][0])
    assert np.array_equal(res[0][1], res[1][1])
    assert res[0][2:] == res[1][2:]


@pytest.mark.parametrize("name", ["ogrid8x16_N3", "periodic6x5_N4", "channel7x6_N5"])
def test_tile_shapes_bitwise_equal(name):
    """The tile shape only decides which cells share a workgroup: the
    residual, the AV sensor, the state after whole steps and the maxima are
    bitwise the same for mesh-order strips and compact blocks (a face point
    is evaluated from the same operands in the same roles either way), and
    for row pairs (two rows of the mesh per workgroup)."""
    print(f"[gpu-parity] tile shapes {name}", flush=True)
    from sddc_amd import abi
    case, Q0 = make_case(name)
    res = []
    for shape in (abi.SD2D_TILE_STRIPS, abi.SD2D_TILE_BLOCKS, abi.SD2D_TILE_ROWPAIRS):
        g = _gpu()(case, tile_shape=shape)
        g.set_state(Q0, [0, 0, 0, 0])
        r, nq = g.compresid()
        sm = g.get_av()[0]
        g.set_state(Q0, [0, 0, 0, 0])
        g.step(3, 0)
        Q, mx = g.get_state()
        res.append((r, nq, sm, Q, mx))
        g.close()
    for other in res[1:]:
        for a, b in zip(res[0], other):
            assert np.array_equal(np.asarray(a), np.asarray(b), equal_nan=True)

"""Viscous and resistive path on the device (vismode = 1, mu and eta0
non-zero): BR1 gradients, gradient traces and ghosts (adiabatic wall
included), the viscous/resistive flux and its divergence.

Against the reference's own outputs on the viscous golden fixture (a
channel with all boundary types, mu = 0.01, eta0 = 0.02), and against the
oracle on every standard mesh.  Bar as for the inviscid path: per variable
||r_gpu - r_ref||_inf / ||r_ref||_inf <= 1e-12 for resid and nablaQs; the
state increment of three whole RK steps to 1e-11.
"""
import json

import numpy as np
import pytest

from tests.golden import specs
from tests.util import ORDERED, make_case, per_var_error

pytestmark = pytest.mark.gpu

TOL = 1e-12
MU, ETA0 = 0.01, 0.02


def _gpu():
    from sddc_amd.solver import Sd2d
    return Sd2d


def test_viscous_fixture_against_reference():
    """First residual evaluation and the state after all steps of the
    reference's viscous run (tests/golden/viscous_channel6x5_N4.npz)."""
    print("[gpu-viscous] fixture viscous_channel6x5_N4", flush=True)
    case, Q0, fx = specs.load("viscous_channel6x5_N4")
    assert case.params.mu == MU and case.params.eta0 == ETA0
    g = _gpu()(case)
    g.set_state(Q0, [0, 0, 0, 0])
    r, nq = g.compresid()
    N, nc = case.N, case.ncell
    e, _ = per_var_error(r, fx["stage00_resid"].reshape(9, N, N, nc, order="F"))
    assert e.max() <= TOL, e
    e2, _ = per_var_error(nq.reshape(18, -1, order="F"),
                          fx["stage00_nablaQs"].reshape(18, -1, order="F"))
    assert e2.max() <= TOL, e2
    meta = json.loads(str(fx["meta"]))
    g.set_state(Q0, [0, 0, 0, 0])
    g.step(meta["nstaged"] + meta["nrk"])
    Q, mx = g.get_state()
    Qr = fx["final_Q"].reshape(Q.shape, order="F")
    d = per_var_error(Q - Q0, Qr - Q0)[0]
    assert d.max() <= 1e-11, d
    np.testing.assert_allclose(mx, fx["final_maxima"][:4], rtol=1e-13)


@pytest.mark.parametrize("name", ORDERED)
def test_viscous_compresid_parity(name):
    print(f"[gpu-viscous] compresid {name}", flush=True)
    from oracle.oracle import Oracle
    case, Q0 = make_case(name, mu=MU, eta0=ETA0)
    g = _gpu()(case)
    o = Oracle(case)
    g.set_state(Q0, [0, 0, 0, 0]); o.set_state(Q0, [0, 0, 0, 0])
    for it in range(2):
        rg, ng = g.compresid()
        ro, no = o.compresid()
        e, _ = per_var_error(rg, ro)
        assert e.max() <= TOL, (name, it, e)
        e2, _ = per_var_error(ng.reshape(18, -1, order="F"), no.reshape(18, -1, order="F"))
        assert e2.max() <= TOL, (name, it, e2)


@pytest.mark.parametrize("k_stage", [3, 5])
@pytest.mark.parametrize("name", ORDERED)
def test_viscous_steps_parity(name, k_stage):
    print(f"[gpu-viscous] steps {name} k_stage={k_stage}", flush=True)
    from oracle.oracle import Oracle
    case, Q0 = make_case(name, k_stage=k_stage, mu=MU, eta0=ETA0)
    g = _gpu()(case)
    o = Oracle(case)
    g.set_state(Q0, [0, 0, 0, 0]); o.set_state(Q0, [0, 0, 0, 0])
    g.step(3); o.step(3)
    Qg, mg = g.get_state()
    Qo, mo = o.get_state()
    d = per_var_error(Qg - Q0, Qo - Q0)[0]
    assert d.max() <= 1e-11, (name, d)
    assert np.allclose(mg, mo, rtol=1e-13, atol=0), (mg, mo)

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

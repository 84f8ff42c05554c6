"""sd2d_rk_stage: the stages of a step run one call at a time give bitwise
the state, running maxima, ctime and iter of sd2d_step, for RK_S5P3 and
FiveStageRk; stages out of order, and sd2d_step in the middle of a step,
are refused; a new state abandons the step."""
import numpy as np
import pytest

from tests.util import make_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,k_stage", [("ogrid4x8_N4", 3), ("channel7x6_N5", 3),
                                          ("periodic4x4_N5", 5)])
def test_stages_equal_steps(name, k_stage):
    from sddc_amd.solver import Sd2d
    print(f"[gpu-rk-stage] {name} k_stage={k_stage}", flush=True)
    case, Q0 = make_case(name, k_stage=k_stage)
    a, b = Sd2d(case), Sd2d(case)
    try:
        for g in (a, b):
            g.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
            g.set_time(0.25, 3)
        a.step(2, 0)
        for _ in range(2):
            for s in range(1, 6):
                b.rk_stage(s)
        qa, ma = a.get_state()
        qb, mb = b.get_state()
        assert np.array_equal(qa, qb)
        assert np.array_equal(ma, mb)
        da, db = a.diag(), b.diag()
        assert (da["iter"], da["ctime"]) == (db["iter"], db["ctime"]) == (5, da["ctime"])
        assert da["c_h"] == db["c_h"]
    finally:
        a.close()
        b.close()


def test_stage_order_enforced():
    """Stages out of order and, between stage 1 and stage 5 of a step
    (RK_S5P3 stage 4 writes its output in place), every call that reads or
    writes the state or evaluates the residual outside the sequence are
    refused; the running maxima alone can be read; a new state abandons
    the step."""
    from sddc_amd.solver import Sd2d, Sd2dError
    print("[gpu-rk-stage] order", flush=True)
    case, Q0 = make_case("ogrid4x8_N4")
    g = Sd2d(case)
    try:
        g.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
        with pytest.raises(Sd2dError):
            g.rk_stage(2)
        g.rk_stage(1)
        g.rk_stage(2)
        with pytest.raises(Sd2dError):
            g.step(1)
        with pytest.raises(Sd2dError):
            g.rk_stage(4)
        g.rk_stage(3)
        g.rk_stage(4)
        for call in (g.get_state, g.compresid, lambda: g.write_restart("/nonexistent/x.dat"),
                     lambda: g.compresid_vjp(np.zeros_like(Q0)), lambda: g.fwd(1, 1)):
            with pytest.raises(Sd2dError, match="in progress"):
                call()
        mx = g.maxima()                                          # allowed mid-step
        assert np.all(np.isfinite(mx)) and mx[2] > 0.0
        g.rk_stage(5)
        q_end, _ = g.get_state()
        assert not np.array_equal(q_end, np.asfortranarray(Q0))
        g.set_state(Q0)                                          # a new state ...
        g.rk_stage(1)
        g.set_state(Q0)                                          # ... abandons the step
        g.step(1)
        g.rk_stage(1)
    finally:
        g.close()


@pytest.mark.parametrize("k_stage", [3, 5])
def test_diag_c_h_is_the_last_evaluations(k_stage):
    """sd2d_get_diag's c_h after sd2d_step / sd2d_rk_stage is the one the
    last stage used: sqrt(eigvmax_g (eigvmax_g - umax_g)) of the maxima
    before that stage (compflux.f90:6), and equal to the oracle's."""
    from sddc_amd.solver import Sd2d
    from oracle.oracle import Oracle
    print(f"[gpu-rk-stage] diag c_h k_stage={k_stage}", flush=True)
    case, Q0 = make_case("channel7x6_N5", k_stage=k_stage)
    g = Sd2d(case)
    o = Oracle(case)
    try:
        g.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
        o.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
        assert g.diag()["c_h"] == 0.0
        for s in range(1, 5):
            g.rk_stage(s)
        mx = g.maxima()
        g.rk_stage(5)
        ch = np.sqrt(mx[2] * (mx[2] - mx[3]))
        assert abs(g.diag()["c_h"] - ch) <= 1e-15 * ch, (g.diag()["c_h"], ch)
        g.step(2)
        o.step(3)
        ch_g, ch_o = g.diag()["c_h"], o.diag()["c_h"]
        assert ch_g > 0.0
        assert abs(ch_g - ch_o) <= 1e-14 * ch_o, (ch_g, ch_o)
    finally:
        g.close()
        o.close()

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
        q_mid, _ = g.get_state()
        assert np.array_equal(q_mid, np.asfortranarray(Q0))   # the step is not complete
        g.set_state(Q0)                                         # abandons the step
        g.step(1)
        g.rk_stage(1)
    finally:
        g.close()

"""The opt-in finiteness check (sd2d_set_check_finite,
sd2d_ens_set_check_finite): a state holding a NaN is reported with
SD2D_ENONFINITE after the step (one run), and by member (ensembles); with
the check off the NaN propagates silently as in the reference, and finite
runs are not flagged.  The NaN enters through the state, as the reference's
would through a restart file; every kernel runs normally on it."""
import numpy as np
import pytest

from tests.util import make_case

pytestmark = pytest.mark.gpu


def test_single_run_nonfinite_check():
    from sddc_amd.solver import Sd2d, Sd2dNonFinite
    print("[gpu-nonfinite] single run", flush=True)
    case, Q0 = make_case("ogrid4x8_N4")
    Qn = Q0.copy(order="F")
    Qn[4, 1, 2, 5] = np.nan
    g = Sd2d(case)
    try:
        g.set_check_finite(2)
        g.set_state(Q0, [0, 0, 0, 0])
        g.step(3)                                     # finite: no error
        assert g.diag()["iter"] == 3
        g.set_state(Qn, [0, 0, 0, 0])
        g.set_time(0.0, 0)
        with pytest.raises(Sd2dNonFinite, match="after step 2"):
            g.step(5)                                 # checked after step 2
        assert g.diag()["iter"] == 2                  # the step is kept
        assert np.isnan(g.get_state()[0]).any()
        g.set_check_finite(0)                         # off: silent, as the reference
        g.set_state(Qn, [0, 0, 0, 0])
        g.step(2)
        g.set_check_finite(1)
        g.set_state(Qn, [0, 0, 0, 0])
        for s in range(1, 5):
            g.rk_stage(s)
        with pytest.raises(Sd2dNonFinite):
            g.rk_stage(5)
    finally:
        g.close()


def test_ensemble_nonfinite_members():
    from sddc_amd.solver import Sd2dEnsemble, Sd2dNonFinite
    print("[gpu-nonfinite] ensemble", flush=True)
    mem = [make_case(n) for n in ("ogrid8x16_N3", "ogrid4x8_N4", "channel7x6_N5",
                                  "ogrid4x8_N4")]
    ens = Sd2dEnsemble([c for c, _ in mem])
    try:
        ens.set_check_finite(1)
        for i, (c, Q0) in enumerate(mem):
            ens.set_state(i, Q0, [0, 0, 0, 0])
        ens.step(2)
        assert list(ens.nonfinite()) == [0, 0, 0, 0]
        bad = mem[3][1].copy(order="F")
        bad[0, 0, 0, 1] = np.inf
        ens.set_state(3, bad, [0, 0, 0, 0])
        with pytest.raises(Sd2dNonFinite, match="member"):
            ens.step(1)
        assert list(ens.nonfinite()) == [0, 0, 0, 1]
        # the other members stepped on undisturbed
        assert all(np.all(np.isfinite(ens.get_state(i)[0])) for i in range(3))
    finally:
        ens.close()

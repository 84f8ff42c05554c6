"""Ensembles (sd2d_ens_*): members of different meshes, orders, free
streams and time steps stepped together in shared launches and a replayed
HIP graph give bitwise the states, maxima and times of the same runs made
one by one through sd2d_step, in FP64 and in mixed precision, before and
after the tile-shape autotuner has rebuilt the groups."""
import numpy as np
import pytest

from tests.util import make_case

pytestmark = pytest.mark.gpu

MEMBERS = [
    ("ogrid8x16_N3", dict(dt=1e-3)),
    ("channel5x4_N3", dict(dt=7e-4, uinf=1.2, bxinf=0.2, byinf=0.1)),
    ("ogrid4x8_N4", dict(dt=1e-3)),
    ("periodic6x5_N4", dict(dt=8e-4)),
    ("channel7x6_N5", dict(dt=5e-4, uinf=2.0, vinf=0.3, bxinf=0.1)),
    ("ogrid4x8_N4", dict(dt=6e-4, uinf=1.8, byinf=0.2)),
]


def _members():
    return [make_case(n, **kw) for n, kw in MEMBERS]


def _single(case, Q0, nsteps, precision):
    from sddc_amd.solver import Sd2d
    g = Sd2d(case, precision=precision)
    g.set_state(Q0, [0, 0, 0, 0])
    g.step(nsteps)
    Q, mx = g.get_state()
    t = g.diag()["ctime"]
    g.close()
    return Q, mx, t


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_equals_single_runs(precision):
    print(f"[gpu-ens] precision {precision}", flush=True)
    from sddc_amd.solver import Sd2dEnsemble
    mem = _members()
    ens = Sd2dEnsemble([c for c, _ in mem], precision=precision)
    for i, (c, Q0) in enumerate(mem):
        ens.set_state(i, Q0, [0, 0, 0, 0])
    for k in range(1, 4):            # three graph replays, eager reads between
        ens.step(1)
        for i, (c, Q0) in enumerate(mem):
            Qe, mxe = ens.get_state(i)
            Qs, mxs, ts = _single(c, Q0, k, precision)
            assert np.array_equal(Qe, Qs), (k, i)
            assert np.array_equal(mxe, mxs), (k, i, mxe, mxs)
            te, ite = ens.get_time(i)
            assert ite == k and te == ts, (k, i, te, ts)
    ens.close()


def test_ensemble_tune_keeps_states():
    print("[gpu-ens] tune", flush=True)
    from sddc_amd.solver import Sd2dEnsemble
    mem = _members()
    ens = Sd2dEnsemble([c for c, _ in mem])
    for i, (c, Q0) in enumerate(mem):
        ens.set_state(i, Q0, [0, 0, 0, 0])
    shapes = ens.tune(2)
    assert set(shapes) == {3, 4, 5} and set(shapes.values()) <= {1, 2, 3}, shapes
    ens.step(2)
    for i, (c, Q0) in enumerate(mem):
        Qe, mxe = ens.get_state(i)
        Qs, mxs, _ = _single(c, Q0, 2, 0)
        assert np.array_equal(Qe, Qs), i
        assert np.array_equal(mxe, mxs), i
    ens.close()

"""The drop-in: the reference program itself (its main, time loop and
RK_S5P3 update, built from the reference sources by oracle/ref.mk) with
only `compresid` replaced by sddc_amd/fortran/compresid_sd2d.f90, which
calls this library through the ISO_C_BINDING module sddc_amd/fortran/sd2d_c.f90.

On the C1 cylinder (32 x 64 O-grid, N=4) for 11 steps its resid.out norms
and restart states are compared with those of the unmodified reference
program, stored in tests/golden/c1_cylinder.npz.  Bar: the state change
over 1 and 11 steps per variable to 1e-10 relative, the RMS norms to
1e-10 relative (the device forms FMAs, so not bit for bit).
"""
import os

import numpy as np
import pytest

from tests.golden import specs

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "oracle", "_ref", "sd2d_MHD_dropin")
STRIDE = 13


def _increment_error(q, q0, sub_ref):
    """Per-variable max error of the state change on the stored subsample
    (every STRIDE-th element of Q in Fortran order)."""
    d_got = (q.ravel(order="F") - q0.ravel(order="F"))[::STRIDE]
    d_ref = sub_ref - q0.ravel(order="F")[::STRIDE]
    var = np.arange(q.size)[::STRIDE] % 9
    errs = []
    for v in range(9):
        m = var == v
        scale = max(np.abs(d_ref[m]).max(), 1e-10 * np.abs(d_ref).max())
        errs.append(np.abs(d_got[m] - d_ref[m]).max() / scale)
    return np.array(errs)


def test_dropin_reference_program_c1():
    if not os.path.exists(EXE):
        pytest.skip("drop-in program not built (oracle/ref.mk dropin needs the reference sources)")
    from tests.golden import refrun
    print("[gpu-dropin] reference program + device compresid, C1, 11 steps", flush=True)
    case, Q0, fx = specs.load("c1_cylinder")
    out = refrun.run_program(case, Q0, maxiter=11, exe="sd2d_MHD_dropin")
    rows = out["resid_out"]
    ref = fx["prog_resid_out"]
    assert rows.shape == ref.shape and list(rows[:, 0]) == [1, 11]
    np.testing.assert_allclose(rows[:, 2:], ref[:, 2:], rtol=1e-10)
    q1 = out["restart.00000001.dat"][0]
    q11 = out["restart.00000011.dat"][0]
    e1 = _increment_error(q1, Q0, fx["prog_q_iter1__sub"])
    e11 = _increment_error(q11, Q0, fx["prog_q_iter11__sub"])
    assert e1.max() <= 1e-10, e1
    assert e11.max() <= 1e-10, e11
    np.testing.assert_allclose(fx["prog_ctime"], [out["restart.00000001.dat"][1],
                                                  out["restart.00000011.dat"][1]], rtol=1e-15)

"""The run driver (python -m sddc_amd.run) on the reference's own input
deck for C1 (32 x 64 O-grid, N=4, restart from the seeded initial state,
MAXITER 11): its resid.out rows and restart files against those of the
unmodified reference program on the same deck (tests/golden/c1_cylinder.npz).
Bar as for the drop-in: the state change over 1 and 11 steps per variable
to 1e-10 relative, the RMS norms to 1e-10 relative, ctime to 1e-15."""
import numpy as np
import pytest

from sddc_amd import mesh as M
from sddc_amd import run as R
from tests.golden import refrun, specs
from tests.test_gpu_dropin import _increment_error

pytestmark = pytest.mark.gpu

def test_run_driver_reference_deck_c1(tmp_path):
    print("[gpu-run] run driver on the C1 reference deck, 11 steps", flush=True)
    case, Q0, fx = specs.load("c1_cylinder")
    refrun.write_case(str(tmp_path), case, Q0, 0.0, maxiter=11, nwrite=21)
    out = R.run_dir(str(tmp_path))
    rows = np.array(out["resid_out"])
    ref = fx["prog_resid_out"]
    assert rows.shape == ref.shape and list(rows[:, 0]) == [1, 11]
    np.testing.assert_allclose(rows[:, 2:], ref[:, 2:], rtol=1e-10)
    assert out["restarts"] == ["restart.00000001.dat", "restart.00000011.dat"]
    q1, t1, i1 = M.read_restart(str(tmp_path / "restart.00000001.dat"), case.N)
    q11, t11, i11 = M.read_restart(str(tmp_path / "restart.00000011.dat"), case.N)
    assert (i1, i11) == (1, 11)
    e1 = _increment_error(q1, Q0, fx["prog_q_iter1__sub"])
    e11 = _increment_error(q11, Q0, fx["prog_q_iter11__sub"])
    assert e1.max() <= 1e-10, e1
    assert e11.max() <= 1e-10, e11
    np.testing.assert_allclose(fx["prog_ctime"], [t1, t11], rtol=1e-15)
    # resid.out as written parses like the reference's (list-directed columns)
    lines = (tmp_path / "resid.out").read_text().splitlines()
    assert len(lines) == 3
    np.testing.assert_allclose([float(x) for x in lines[2].split()[1:]], rows[1, 1:], rtol=1e-15)


def test_run_driver_resumes_bitwise(tmp_path):
    """The deck run for 11 steps, against the same deck run for 6 steps and
    then restarted from the restart file it wrote at step 6 (with its
    running-maxima sidecar) up to 11: the final restart files are byte for
    byte the same, and so are the resid.out rows after the restart."""
    import os
    import shutil
    print("[gpu-run] run driver resume from a checkpoint", flush=True)
    case, Q0, fx = specs.load("c1_cylinder")
    full, part = tmp_path / "full", tmp_path / "part"
    for d, mi in ((full, 11), (part, 6)):
        d.mkdir()
        refrun.write_case(str(d), case, Q0, 0.0, maxiter=mi, nwrite=100)
    R.run_dir(str(full))
    R.run_dir(str(part))
    assert os.path.exists(part / "restart.00000006.dat.max")
    # restart = 1 from restart.00000006.dat, MAXITER 11
    shutil.copy(part / "restart.00000006.dat", part / "ic.dat")
    shutil.copy(part / "restart.00000006.dat.max", part / "ic.dat.max")
    lines = (part / "QUAD.INP").read_text().split("\n")
    lines[17] = "11"                                  # MAXITER (readinput.f90:25)
    (part / "QUAD.INP").write_text("\n".join(lines))
    out = R.run_dir(str(part))
    assert out["iter"] == 11
    assert (part / "restart.00000011.dat").read_bytes() == (full / "restart.00000011.dat").read_bytes()
    assert (part / "restart.00000011.dat.max").read_bytes() == \
        (full / "restart.00000011.dat.max").read_bytes()
    r_full = (full / "resid.out").read_text().splitlines()
    r_part = (part / "resid.out").read_text().splitlines()
    assert r_part[-1] == r_full[-1]                   # the step-11 row

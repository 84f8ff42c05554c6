"""Reference-format restart files from and into the device state
(write_data.f90:44-74, read_data_restart.f90:35-67).

sddc_amd.mesh.write_restart / read_restart are the format as the tests pin
it against the reference itself (the reference program reads the initial
states they write, and its own restart files are read back in
test_golden_oracle / test_gpu_dropin), so the library's files are compared
with them byte for byte.
"""
import os

import numpy as np
import pytest

from sddc_amd import mesh as M
from tests.util import make_case

pytestmark = pytest.mark.gpu


def test_restart_write_matches_format(tmp_path):
    from sddc_amd.solver import Sd2d
    print("[gpu-io] write restart", flush=True)
    case, Q0 = make_case("channel7x6_N5")
    g = Sd2d(case)
    g.set_state(Q0, [0.1, 0.2, 0.3, 0.05])
    g.set_time(0.125, 17)
    g.write_restart(tmp_path / "dev.dat")
    M.write_restart(str(tmp_path / "ref.dat"), Q0, ctime=0.125, it=17)
    assert (tmp_path / "dev.dat").read_bytes() == (tmp_path / "ref.dat").read_bytes()
    # global cell numbers are written as given
    icg = np.arange(case.ncell, dtype=np.int32)[::-1] + 1
    g.write_restart(tmp_path / "perm.dat", icg)
    Qr, t, it = M.read_restart(str(tmp_path / "perm.dat"), case.N)
    assert np.array_equal(Qr, Q0[..., ::-1]) and t == 0.125 and it == 17


def test_restart_round_trip(tmp_path):
    from sddc_amd.solver import Sd2d, Sd2dError
    print("[gpu-io] read restart", flush=True)
    case, Q0 = make_case("ogrid4x8_N4")
    # a file as the reference writes it, cells in scrambled record order
    rng = np.random.default_rng(4)
    order = rng.permutation(case.ncell)
    N = case.N
    with open(tmp_path / "scr.dat", "wb") as fh:
        fh.write(np.int32(40).tobytes()); fh.write(np.float64(2.5).tobytes())
        for ic in order:
            fh.write(np.int32(ic + 1).tobytes())
            fh.write(np.asfortranarray(Q0[..., ic]).ravel(order="F").tobytes())
    g = Sd2d(case)
    g.set_state(np.zeros_like(Q0), [1, 1, 1, 1])
    g.read_restart(tmp_path / "scr.dat")
    Q, mx = g.get_state()
    assert np.array_equal(Q, Q0)
    assert np.array_equal(mx, np.zeros(4))          # initsetup.f90:166-167
    d = g.diag()
    assert d["iter"] == 40 and d["ctime"] == 2.5
    # stepping continues from the file's clock
    g.step(1)
    assert g.diag()["iter"] == 41
    # malformed files are refused
    (tmp_path / "short.dat").write_bytes((tmp_path / "scr.dat").read_bytes()[:-8])
    with pytest.raises(Sd2dError):
        g.read_restart(tmp_path / "short.dat")
    bad = bytearray((tmp_path / "scr.dat").read_bytes())
    bad[12:16] = np.int32(case.ncell + 5).tobytes()
    (tmp_path / "bad.dat").write_bytes(bytes(bad))
    with pytest.raises(Sd2dError):
        g.read_restart(tmp_path / "bad.dat")

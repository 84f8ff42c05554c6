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


@pytest.mark.parametrize("name", ["ogrid4x8_N4", "periodic6x5_N4", "channel7x6_N5"])
def test_checkpoint_resumes_bitwise(tmp_path, name):
    """A run stopped after 3 steps, checkpointed (restart file plus the
    running-maxima sidecar) and resumed in a fresh context for 3 more steps
    equals the continuous 6-step run bit for bit.  The restart file alone
    (sd2d_read_restart, the reference's behaviour: maxima reset to 0,
    initsetup.f90:154-167) gives a different state, which is why the sidecar
    exists; the restart file is byte-for-byte the one sd2d_write_restart
    writes, so the reference program reads it as well."""
    from sddc_amd.solver import Sd2d, Sd2dError
    print(f"[gpu-io] checkpoint resume {name}", flush=True)
    case, Q0 = make_case(name)
    a = Sd2d(case)
    a.set_state(Q0, [0, 0, 0, 0])
    a.step(6, 2)
    Qc, mc = a.get_state()
    dc = a.diag()
    a.close()
    b = Sd2d(case)
    b.set_state(Q0, [0, 0, 0, 0])
    b.step(3, 2)
    b.write_checkpoint(tmp_path / "ck.dat")
    b.write_restart(tmp_path / "plain.dat")
    assert (tmp_path / "ck.dat").read_bytes() == (tmp_path / "plain.dat").read_bytes()
    assert len((tmp_path / "ck.dat.max").read_bytes()) == 8 + 4 + 8 + 32
    b.close()
    r = Sd2d(case)
    r.read_checkpoint(tmp_path / "ck.dat")
    d3 = r.diag()
    assert d3["iter"] == 3
    r.step(3, 2)
    Qr, mr = r.get_state()
    dr = r.diag()
    assert np.array_equal(Qr, Qc)
    assert np.array_equal(mr, mc)
    assert dr["iter"] == dc["iter"] == 6 and dr["ctime"] == dc["ctime"]
    # the reference's restart alone does not resume the same run
    r.read_restart(tmp_path / "ck.dat")
    r.step(3, 2)
    assert not np.array_equal(r.get_state()[0], Qc)
    # a sidecar of another step is refused
    os.replace(tmp_path / "ck.dat.max", tmp_path / "keep.max")
    r.write_checkpoint(tmp_path / "other.dat")
    os.replace(tmp_path / "other.dat.max", tmp_path / "ck.dat.max")
    with pytest.raises(Sd2dError, match="another step"):
        r.read_checkpoint(tmp_path / "ck.dat")
    os.remove(tmp_path / "ck.dat.max")
    with pytest.raises(Sd2dError, match="sidecar"):
        r.read_checkpoint(tmp_path / "ck.dat")
    r.close()


def test_ensemble_checkpoint_resumes_bitwise(tmp_path):
    """An ensemble member checkpointed after 2 steps and loaded into another
    ensemble (as a different member slot of the same mesh) steps on to
    bitwise the continuous run's state, maxima and time."""
    from sddc_amd.solver import Sd2dEnsemble
    print("[gpu-io] ensemble checkpoint", flush=True)
    mem = [make_case(n) for n in ("ogrid8x16_N3", "ogrid4x8_N4", "channel7x6_N5")]
    e = Sd2dEnsemble([c for c, _ in mem])
    for i, (c, Q0) in enumerate(mem):
        e.set_state(i, Q0, [0, 0, 0, 0])
    e.step(4)
    cont = [(e.get_state(i), e.get_time(i)) for i in range(3)]
    e.close()
    e = Sd2dEnsemble([c for c, _ in mem])
    for i, (c, Q0) in enumerate(mem):
        e.set_state(i, Q0, [0, 0, 0, 0])
    e.step(2)
    for i in range(3):
        e.write_checkpoint(i, tmp_path / f"m{i}.dat")
    e.close()
    # a fresh ensemble with the members in another order
    perm = [2, 0, 1]
    f = Sd2dEnsemble([mem[k][0] for k in perm])
    for j, k in enumerate(perm):
        f.read_checkpoint(j, tmp_path / f"m{k}.dat")
        assert f.get_time(j)[1] == 2
    f.step(2)
    for j, k in enumerate(perm):
        (Qc, mc), (tc, itc) = cont[k]
        Qf, mf = f.get_state(j)
        assert np.array_equal(Qf, Qc), k
        assert np.array_equal(mf, mc), k
        assert f.get_time(j) == (tc, itc)
    f.close()

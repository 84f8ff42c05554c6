"""The stage kernels' indexing, checked on the CPU.

tests/emu/stage_emu.cpp compiles the device stage kernels
(sddc_amd/csrc/sd2d_stage.hpp and sd2d_stage_mx.hpp, unchanged) with g++
against the CPU stand-ins of tests/emu/hip/hip_runtime.h, under
AddressSanitizer and UBSan, and launches them on host threads with the
library's grids, LDS sizes, arguments and buffer sizes: 128 threads per
workgroup with real barriers, exact-size buffers that start as NaN (as
uninitialised device memory may), LDS poisoned before every workgroup with
a canary beyond each launch's size.  An out-of-bounds access, a barrier that
not every work-item reaches, an LDS write past the launch size or a signed
overflow aborts the program; a read of memory nothing wrote turns the
results into NaN.  The results are then held to the device parity bars
against the oracle: sd2d_compresid's resid and nablaQs (two calls, c_h = 0
then lagged) to 1e-12 per variable, three RK_S5P3 / FiveStageRk steps to
1e-11 on the increment, the mixed-precision kernel's steps to 1e-6.

Meshes: every face type, N = 3, 4, 5, rows padded to the tile so padding
slots lie inside the mesh, periodic pairs, 8-node curved cells, and the
three tile shapes (strips, blocks, row pairs)."""
import hashlib
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle
from sddc_amd import mesh as M
from tests.util import make_case, per_var_error

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
EMU = os.path.join(HERE, "emu")
CSRC = os.path.join(REPO, "sddc_amd", "csrc")
BUILD = os.path.join(EMU, "_build")
SOURCES = [os.path.join(EMU, "stage_emu.cpp"), os.path.join(EMU, "hip", "hip_runtime.h"),
           os.path.join(REPO, "include", "sd2d.h")] + \
          [os.path.join(CSRC, f) for f in ("sd2d_device.hpp", "sd2d_stage.hpp", "sd2d_stage_mx.hpp",
                                           "sd2d_host.hpp", "sd2d_rkseq.hpp")]


@pytest.fixture(scope="module")
def emu_exe():
    if shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    h = hashlib.sha256(b"".join(open(p, "rb").read() for p in SOURCES)).hexdigest()[:16]
    exe = os.path.join(BUILD, f"stage_emu_{h}")
    if not os.path.exists(exe):
        os.makedirs(BUILD, exist_ok=True)
        cmd = ["g++", "-std=c++20", "-O1", "-g", "-fno-omit-frame-pointer",
               "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-pthread",
               "-Wno-unknown-pragmas", "-I", EMU, "-I", CSRC,
               os.path.join(EMU, "stage_emu.cpp"), "-o", exe + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-4000:]
        os.replace(exe + ".tmp", exe)
    return exe


def _run(exe, tmp_path, case, Q, job, nsteps=0, precision=0, tile_shape=0):
    """Write the emulator's input, run it, return the float64 output."""
    p, t, N, nc = case.params, case.topo, case.N, case.ncell
    L = t.lists
    lists = [t.ibfinter, L["INLE"], L["OUTL"], L["WALL"], L["SYMP"], L["FREE"]]
    hdr = np.array([N, nc, t.if2c.shape[0], *[len(x) for x in lists], t.cyc2fpair.shape[0],
                    p.k_stage, p.artiv, job, nsteps, precision, tile_shape], dtype="<i4")
    prm = np.array([p.mu, p.eta0, p.alpha, p.rinf, p.pinf, p.uinf, p.vinf, p.bxinf, p.byinf,
                    p.dxcycl, p.dycycl, p.dt, 0.0, 0.0, 0.0, 0.0], dtype="<f8")
    fin = tmp_path / "in.bin"
    fout = tmp_path / "out.bin"
    with open(fin, "wb") as fh:
        fh.write(hdr.tobytes())
        fh.write(prm.tobytes())
        fh.write(np.asarray(t.if2c, dtype="<i4").ravel(order="F").tobytes())
        for x in lists:
            fh.write(np.asarray(x, dtype="<i4").tobytes())
        fh.write(np.asarray(t.cyc2fpair, dtype="<i4").ravel(order="F").tobytes())
        for a in (case.lmat, case.mmat, case.xs, case.xs_low, case.jac, case.s1, case.s2, Q):
            fh.write(np.asarray(a, dtype="<f8").ravel(order="F").tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, str(fin), str(fout)], capture_output=True, text=True, env=env,
                       timeout=600)
    assert r.returncode == 0, f"emulated kernels failed ({r.returncode}):\n{r.stderr[-6000:]}"
    return np.fromfile(fout, dtype="<f8")


# rows of 23 (N=5, tiles of 12) and 31 (N=4, tiles of 16) cells are padded
# to 24 and 32 slots, so padding slots lie between the rows of the mesh
PADDED = {"ogrid2x23_N5": (lambda: M.ogrid(2, 23, 5.0, 6.0), dict(N=5, dt=5e-4)),
          "periodic31x2_N4": (lambda: M.periodic_square(31, 2, 3.1, 0.2), dict(N=4, dt=1e-3))}


def _case(name, **over):
    if name not in PADDED:
        return make_case(name, **over)
    mk, kw = PADDED[name]
    m = mk()
    p = M.RunParams(dxcycl=m.dxcycl, dycycl=m.dycycl, **{**kw, **over})
    case = M.build_case(m, p)
    return case, M.smooth_perturbed_state(case, 0.05, seed=1)


CASES = ["channel5x4_N3", "ogrid4x8_N4", "periodic6x5_N4", "channel7x6_N5", "ogrid2x23_N5",
         "periodic31x2_N4", "ogrid_curved4x8_N4"]


def test_padded_meshes_have_padding_slots():
    """The two padded meshes really hold padding slots inside the mesh (the
    host's row-aligned strips), so the kernels see them."""
    for name in PADDED:
        case, _ = _case(name)
        N = case.N
        cpw = {5: 12, 4: 16, 3: 20}[N]
        R = case.ncell // 2
        assert R % cpw != 0 and 16 * (-R % cpw) <= R, name


@pytest.mark.parametrize("name", CASES)
def test_emulated_compresid(emu_exe, tmp_path, name):
    case, Q0 = _case(name)
    n = 9 * case.N ** 2 * case.ncell
    out = _run(emu_exe, tmp_path, case, Q0, job=0)
    nb = 3 * n + 2 + 2 * case.ncell
    assert out.size == 2 * nb
    o = Oracle(case)
    o.set_state(Q0, [0, 0, 0, 0])
    for call in range(2):
        blk = out[call * nb:(call + 1) * nb]
        ro, no = o.compresid()
        r = blk[:n].reshape(ro.shape, order="F")
        g = blk[n:3 * n].reshape(no.shape, order="F")
        e, _ = per_var_error(r, ro)
        assert e.max() <= 1e-12, (name, call, e)
        e2, _ = per_var_error(g.reshape(18, -1, order="F"), no.reshape(18, -1, order="F"))
        assert e2.max() <= 1e-12, (name, call, e2)
        _, mo = o.get_state()
        np.testing.assert_allclose(blk[3 * n:3 * n + 2], np.asarray(mo)[2:4], rtol=1e-14, atol=0)
        # the AV smoothness sensor, to the device test's bar (log10 of a
        # squared small difference over a squared value: 1e-15 in its
        # inputs moves it by up to ~1e-8)
        sm = blk[3 * n + 2:].reshape(2, -1, order="F")
        so = np.asarray(o.get_av()[0]).reshape(2, -1, order="F")
        fin = np.isfinite(so)
        assert np.array_equal(np.isfinite(sm), fin), (name, call)
        assert np.abs(sm[fin] - so[fin]).max() <= 1e-6, (name, call)
    o.close()


@pytest.mark.parametrize("name,k_stage,tile_shape", [("channel5x4_N3", 3, 0), ("ogrid4x8_N4", 5, 0),
                                                     ("channel7x6_N5", 3, 2), ("ogrid2x23_N5", 5, 0),
                                                     ("periodic31x2_N4", 3, 0),
                                                     ("periodic31x2_N4", 3, 2),
                                                     ("periodic31x2_N4", 3, 3),
                                                     ("ogrid2x23_N5", 3, 3),
                                                     ("ogrid_curved4x8_N4", 3, 0)])
def test_emulated_rk_steps(emu_exe, tmp_path, name, k_stage, tile_shape):
    case, Q0 = _case(name, k_stage=k_stage)
    n = 9 * case.N ** 2 * case.ncell
    out = _run(emu_exe, tmp_path, case, Q0, job=1, nsteps=3, tile_shape=tile_shape)
    o = Oracle(case)
    o.set_state(Q0, [0, 0, 0, 0])
    o.step(3)
    Qo, mo = o.get_state()
    o.close()
    Q = out[:n].reshape(Qo.shape, order="F")
    d = per_var_error(Q - Q0, Qo - Q0)[0]
    assert d.max() <= 1e-11, (name, d)
    np.testing.assert_allclose(out[n:], np.asarray(mo)[2:4], rtol=1e-13, atol=0)


@pytest.mark.parametrize("name", ["channel5x4_N3", "periodic6x5_N4", "channel7x6_N5"])
def test_emulated_mixed_steps(emu_exe, tmp_path, name):
    case, Q0 = _case(name)
    n = 9 * case.N ** 2 * case.ncell
    out = _run(emu_exe, tmp_path, case, Q0, job=1, nsteps=3, precision=1)
    o = Oracle(case)
    o.set_state(Q0, [0, 0, 0, 0])
    o.step(3)
    Qo, _ = o.get_state()
    o.close()
    Q = out[:n].reshape(Qo.shape, order="F")
    assert np.isfinite(Q).all()
    d = per_var_error(Q, Qo)[0]
    assert d.max() <= 1e-6, (name, d)

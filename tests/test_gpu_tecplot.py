"""Tecplot output from the device (sd2d_plot_nodes, the run driver's TEC/
files).

* The device's nodal values equal the oracle's restatement of tecplotter2d
  (tecplotter.f90:17-60) bit for bit on the same state, on O-grid, channel
  and periodic meshes, N = 3, 4, 5 and NRE = 1, 2, 3.
* The run driver on the C1 deck (one step, NRE = 1 and 3) and on a deck of
  8-node cells (CURVE_WALL = 1, NRE = 3) writes the file the reference
  program wrote at that step (tests/golden/c1_tecplot.npz,
  tests/golden/ogrid_curved_tecplot.npz):
  the same name, size, header, coordinates, AV fields and connectivity
  byte for byte; each of the 9 flow variables to one float32 spacing of
  the reference's value plus 1e-11 of that variable's largest magnitude,
  and the header's min/max doubles to 1e-11 of that magnitude.  The device
  state agrees with the reference's to about 1e-13 of each variable's
  norm: below float32 resolution for most values, but a value lying on a
  rounding boundary rounds either way, and a value near zero (v on the
  symmetry line, say) carries that absolute difference over several of its
  own float32 spacings.  The byte-for-byte pin of the whole pipeline with
  the reference's own state is the CPU test tests/test_tecplot.py.
"""
import os

import numpy as np
import pytest

from sddc_amd import tecplot as TP
from tests.golden import refrun, specs
from tests.util import make_case

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name,nre", [("ogrid4x8_N4", 1), ("channel5x4_N3", 3),
                                      ("periodic4x4_N5", 2), ("ogrid8x16_N3", 3),
                                      ("ogrid_curved5x12_N5", 3)])
def test_plot_nodes_bitwise(name, nre):
    from oracle.tecplot_oracle import nodal_values
    from sddc_amd.solver import Sd2d
    print(f"[gpu-tecplot] plot nodes {name} NRE {nre}", flush=True)
    case, Q0 = make_case(name)
    g = Sd2d(case)
    try:
        g.set_state(Q0, [0, 0, 0, 0])
        g.step(2)
        Q, _ = g.get_state()
        _, muav, _ = g.get_av()
        topo = TP.plot_topology(case, nre)
        q, av = g.plot_nodes(topo, muav)
        rq, rav = nodal_values(Q, muav, topo, case.xs)
        assert np.array_equal(q, rq), np.abs(q - rq).max()
        assert np.array_equal(av, rav)
        q2, av2 = g.plot_nodes(topo)
        assert np.array_equal(q2, rq) and av2 is None
    finally:
        g.close()


def _split(raw: bytes, nn: int, ncp: int):
    head = raw[:392]                                  # up to the min/max block
    mm = np.frombuffer(raw[392:600], dtype="<f8")
    data = np.frombuffer(raw[600:600 + 13 * 4 * nn], dtype="<f4").reshape(13, nn)
    conn = raw[600 + 13 * 4 * nn:]
    return head, mm, data, conn


@pytest.mark.parametrize("fn,spec,nre", [("c1_tecplot.npz", "c1_cylinder", 1),
                                         ("c1_tecplot.npz", "c1_cylinder", 3),
                                         ("ogrid_curved_tecplot.npz", "ogrid_curved4x8_N4", 3)])
def test_run_driver_writes_reference_tecplot(tmp_path, fn, spec, nre):
    from sddc_amd import run as R
    print(f"[gpu-tecplot] run driver {spec} NRE {nre}", flush=True)
    fx = np.load(os.path.join(HERE, "golden", fn))
    case, Q0 = specs.build(spec)
    case.params.nre = nre
    refrun.write_case(str(tmp_path), case, Q0, 0.0, maxiter=1, nwrite=100000)
    out = R.run_dir(str(tmp_path))
    name = str(fx[f"nre{nre}_name"])
    assert out["tecplot"] == [name]
    raw = (tmp_path / "TEC" / name).read_bytes()
    ref = fx[f"nre{nre}_plt"].tobytes()
    assert len(raw) == len(ref)
    topo = TP.plot_topology(case, nre)
    nn, ncp = topo.nnode, topo.conn.shape[0]
    h, mm, d, cn = _split(raw, nn, ncp)
    rh, rmm, rd, rcn = _split(ref, nn, ncp)
    assert h == rh                                    # header up to min/max
    assert cn == rcn                                  # connectivity
    assert np.array_equal(d[:2], rd[:2])              # X, Y
    assert np.array_equal(d[11:], rd[11:])            # muAV, etaAV
    dv = d[2:11].astype(np.float64)
    rv = rd[2:11].astype(np.float64)
    scale = np.abs(rv).max(axis=1, keepdims=True)
    tol = np.spacing(np.abs(rd[2:11])).astype(np.float64) + 1e-11 * scale
    print(f"[gpu-tecplot] flow values not bitwise equal: {int((dv != rv).sum())} of {dv.size}; "
          f"worst |diff|/tol {float((np.abs(dv - rv) / tol).max()):.3g}", flush=True)
    assert (np.abs(dv - rv) <= tol).all()
    mscale = np.abs(rmm.reshape(13, 2)).max(axis=1).repeat(2)
    assert (np.abs(mm - rmm) <= 1e-11 * mscale).all(), np.abs(mm - rmm).max()

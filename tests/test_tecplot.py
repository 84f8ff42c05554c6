"""The Tecplot output pipeline on the CPU (tecplotter2dsetup.f90,
tecplotter.f90): the plot mesh of sddc_amd.tecplot, the nodal values of the
oracle's restatement of tecplotter2d and the #!TDV111 writer, from the
oracle's state after one step of a deck (bit for bit the reference
program's state), against the file the reference program itself wrote at
that step: byte for byte.  Decks: C1 (tests/golden/c1_tecplot.npz, NRE = 1
and 3) and the 8-node O-grid of the ogrid_curved4x8_N4 fixture, whose plot
nodes lie on the cells' quadratic edges
(tests/golden/ogrid_curved_tecplot.npz, NRE = 3)."""
import os

import numpy as np
import pytest

from sddc_amd import mesh as M
from sddc_amd import tecplot as TP
from tests.golden import specs

HERE = os.path.dirname(os.path.abspath(__file__))


def _fixture(fn="c1_tecplot.npz"):
    return np.load(os.path.join(HERE, "golden", fn))


@pytest.mark.parametrize("nre", [1, 3])
def test_plot_mesh_counts(nre):
    case, _ = specs.build("c1_cylinder")
    topo = TP.plot_topology(case, nre)
    nc, nv = case.ncell, len(case.mesh.xv)
    nface = case.topo.if2c.shape[0]
    assert topo.nnode == nv + nface * (nre - 1) + nc * (nre - 1) ** 2
    assert topo.conn.shape == (nc * nre * nre, 4)
    assert topo.factor.sum() == nc * (nre + 1) ** 2
    # every plot cell is a counter-clockwise quadrilateral of positive area
    x, y = topo.xy[0][topo.conn], topo.xy[1][topo.conn]
    area = 0.5 * ((x * np.roll(y, -1, axis=1)) - (np.roll(x, -1, axis=1) * y)).sum(axis=1)
    assert np.all(area > 0)


@pytest.mark.parametrize("fn,spec,nre", [("c1_tecplot.npz", "c1_cylinder", 1),
                                         ("c1_tecplot.npz", "c1_cylinder", 3),
                                         ("ogrid_curved_tecplot.npz", "ogrid_curved4x8_N4", 3)])
def test_plt_matches_reference_program(tmp_path, fn, spec, nre):
    from oracle.oracle import Oracle
    from oracle.tecplot_oracle import nodal_values
    fx = _fixture(fn)
    case, Q0 = specs.build(spec)
    o = Oracle(case)
    o.set_state(Q0, [0, 0, 0, 0])
    o.set_time(0.0, 0)
    o.step(1)
    Q, _ = o.get_state()
    ctime, _ = o.get_time()
    _, muav = o.get_av()
    o.close()
    topo = TP.plot_topology(case, nre)
    plotq, plotav = nodal_values(Q, muav, topo, case.xs)
    name = TP.tec_filename(ctime)
    assert name == str(fx[f"nre{nre}_name"])
    TP.write_plt(str(tmp_path / name), ctime, topo, plotq, plotav)
    got = np.frombuffer((tmp_path / name).read_bytes(), dtype=np.uint8)
    ref = fx[f"nre{nre}_plt"]
    assert got.size == ref.size
    assert np.array_equal(got, ref), np.nonzero(got != ref)[0][:10]


def test_tec_filename():
    """f6.3 of the fraction, characters 3-6 (tecplotter.f90:176-177)."""
    assert TP.tec_filename(0.0) == "TEC0000000.000.plt"
    assert TP.tec_filename(12.3456) == "TEC0000012.346.plt"
    assert TP.tec_filename(0.0625) == "TEC0000000.062.plt"      # nearest, ties to even
    assert TP.tec_filename(3.9996) == "TEC0000003.000.plt"      # the reference's own quirk

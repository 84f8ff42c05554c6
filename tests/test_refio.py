"""Reading the reference's input deck: QUAD.INP (readinput.f90:6-43) and
the .cel/.vrt/.bnd mesh files (readGRIDinfo.f90:33,78,269), against the
writers that produce the decks the reference program itself is run on
(tests/golden/refrun.py), and the report schedule of the run driver
(iterations.f90:34-84)."""
import dataclasses
import os

import numpy as np
import pytest

from sddc_amd import cases
from sddc_amd import mesh as M
from sddc_amd import run as R


def test_quad_inp_round_trip(tmp_path):
    case, _ = cases.c1_cylinder()
    p = dataclasses.replace(case.params, maxiter=37, nwrite=5, mu=1.25e-3, eta0=0.0,
                            dt=1.2345678901234567e-4, k_stage=5)
    M.write_quad_inp(str(tmp_path / "QUAD.INP"), p)
    qi = M.read_quad_inp(str(tmp_path / "QUAD.INP"))
    assert qi.params == p
    assert (qi.namecel, qi.namevrt, qi.namebnd, qi.namerestart) == \
        ("mesh.cel", "mesh.vrt", "mesh.bnd", "ic.dat")
    assert qi.curve_wall == 0 and qi.restart_ord == p.N


def test_quad_inp_list_directed_forms(tmp_path):
    """Quoted names, trailing text after the item, commas and D exponents,
    as a list-directed READ accepts them."""
    recs = ["SD2D input", "files", "'grid a.cel'  cell file", "grid.vrt, vertices",
            "grid.vrt8", '"grid.bnd"', "grid.mts", "ic.dat", "N CURVE vismode mu eta0",
            "5", "0", "1", "0.0D0", "1.5d-2 eta0", "k_stage dt MAXITER", "3",
            "2.5D-04", "200", "cyclic", "6.283185307179586D0", "6.283185307179586", "AV",
            "1", "NRE nwrite", "1", "50", "restart", "1", "5", "free stream", "1.0",
            "0.6", "1.5D0", "0.0", "0.3", "0.0"]
    assert len(recs) == 36
    (tmp_path / "QUAD.INP").write_text("\n".join(recs) + "\n")
    qi = M.read_quad_inp(str(tmp_path / "QUAD.INP"))
    assert qi.namecel == "grid a.cel" and qi.namevrt == "grid.vrt" and qi.namebnd == "grid.bnd"
    p = qi.params
    assert (p.N, p.vismode, p.k_stage, p.maxiter, p.nwrite, p.restart) == (5, 1, 3, 200, 50, 1)
    assert p.eta0 == 1.5e-2 and p.dt == 2.5e-4 and p.uinf == 1.5 and p.pinf == 0.6
    assert p.dxcycl == 6.283185307179586 and qi.restart_ord == 5


def test_quad_inp_too_short(tmp_path):
    (tmp_path / "QUAD.INP").write_text("a\nb\nc\n")
    with pytest.raises(ValueError):
        M.read_quad_inp(str(tmp_path / "QUAD.INP"))


@pytest.mark.parametrize("which", ["ogrid", "channel", "periodic"])
def test_mesh_files_round_trip(tmp_path, which):
    """The mesh read back is the mesh written, and the case built from it
    (topology, operators, metrics) is bitwise the same."""
    if which == "ogrid":
        m = M.ogrid(6, 12, 0.5, 15.0)
    elif which == "channel":
        m = M.channel(7, 6)
    else:
        m = M.periodic_square(5, 4)
    M.write_mesh_files(str(tmp_path), m)
    r = M.read_mesh_files(str(tmp_path / "mesh.cel"), str(tmp_path / "mesh.vrt"),
                          str(tmp_path / "mesh.bnd"), m.dxcycl, m.dycycl)
    assert np.array_equal(r.cells, m.cells)
    assert np.array_equal(r.xv, m.xv) and np.array_equal(r.yv, m.yv)
    assert r.bnd == [tuple(b) for b in m.bnd]
    p = M.RunParams(N=4, dxcycl=m.dxcycl, dycycl=m.dycycl)
    a, b = M.build_case(m, p), M.build_case(r, p)
    assert np.array_equal(a.cells, b.cells)
    for f in ("jac", "s1", "s2", "lmat", "mmat"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    for f in ("if2c", "if2v", "ic2f", "ibfinter", "cyc2fpair"):
        x, y = getattr(a.topo, f), getattr(b.topo, f)
        assert (x is None and y is None) or np.array_equal(x, y), f
    assert a.topo.lists.keys() == b.topo.lists.keys()
    for k in a.topo.lists:
        assert np.array_equal(a.topo.lists[k], b.topo.lists[k]), k


def test_mesh_files_bad_input(tmp_path):
    (tmp_path / "m.cel").write_text("1 1 2 3 9 1 1\n")
    (tmp_path / "m.vrt").write_text("1 0 0\n2 1 0\n3 1 1\n4 0 1\n")
    (tmp_path / "m.bnd").write_text("1 1 2 1 1 WALL\n")
    with pytest.raises(ValueError):
        M.read_mesh_files(str(tmp_path / "m.cel"), str(tmp_path / "m.vrt"), str(tmp_path / "m.bnd"))
    (tmp_path / "m.cel").write_text("1 1 2 3 4 1 1\n")
    (tmp_path / "m.bnd").write_text("1 1 2 1 1 WALX\n")
    with pytest.raises(ValueError):
        M.read_mesh_files(str(tmp_path / "m.cel"), str(tmp_path / "m.vrt"), str(tmp_path / "m.bnd"))


def test_report_schedule():
    """Norms on mod(iter,10) == 1, restarts on mod(iter,nwrite) == 1."""
    steps = [R._next(it, 10) for it in (0, 1, 5, 10, 11, 20)]
    assert steps == [1, 11, 11, 11, 21, 21]
    assert R._next(3, 4) == 5 and R._next(5, 4) == 9
    assert R._next(7, 1) > 10 ** 9      # mod(iter, 1) is never 1


def _pairs_all(m, topo, dxc, dyc):
    """All-pairs form of the reference's midpoint test (map_cyc_loc.f90:15-35)."""
    cyc = topo.lists["CYCL"]
    v = topo.if2v[cyc - 1] - 1
    xm = (m.xv[v[:, 0]] + m.xv[v[:, 1]]) / 2
    ym = (m.yv[v[:, 0]] + m.yv[v[:, 1]]) / 2
    out = []
    for i in range(len(cyc)):
        for j in range(i + 1, len(cyc)):
            dx, dy = abs(xm[i] - xm[j]), abs(ym[i] - ym[j])
            if (abs(dx - dxc) < M.TOLCYC and dy < M.TOLCYC) or \
               (abs(dy - dyc) < M.TOLCYC and dx < M.TOLCYC):
                out.append((cyc[i], cyc[j]))
    return np.array(out, dtype=np.int32)


@pytest.mark.parametrize("nx,ny", [(4, 4), (5, 3), (16, 7), (1, 6)])
def test_cyclic_pairs_match_all_pairs_test(nx, ny):
    """The hashed pairing returns exactly the pairs, in the same order, as
    the all-pairs test, also on meshes with one cell across."""
    m = M.periodic_square(nx, ny)
    topo = M.build_topology(m, m.dxcycl, m.dycycl)
    ref = _pairs_all(m, topo, m.dxcycl, m.dycycl)
    assert np.array_equal(topo.cyc2fpair, ref)

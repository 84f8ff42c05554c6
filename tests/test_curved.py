"""8-node (curved) cells, CURVE_WALL = 1, on the CPU: the metrics of
MAPDER_8_NODE + calcjacob (mapder.f90:91-253, calcjacob.f90:26-53), the
.vrt8 file, and what the set-up refuses.

The bit-for-bit pins against the reference program's own set-up arrays
and residuals are the ogrid_curved4x8_N4 and viscous_ogrid_curved3x10_N3
fixtures of tests/test_golden_oracle.py; the checks here pin the
restatement by the mathematics instead, so that a wrong sign, term or
node order in the shape-function derivatives fails on its own."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from sddc_amd import mesh as M

# the 8 nodes in the reference's local coordinates (mapder.f90:107-116)
NODES = np.array([[0, 0], [1, 0], [1, 1], [0, 1], [0.5, 0], [1, 0.5], [0.5, 1], [0, 0.5]], float)


def _quad_map(rng):
    """A random map in the span of the 8-node serendipity space
    {1, xi, eta, xi^2, xi*eta, eta^2, xi^2*eta, xi*eta^2} and its exact
    derivatives."""
    a = rng.normal(size=(2, 8)) * np.array([1, 1, 1, 0.2, 0.2, 0.2, 0.1, 0.1])
    a[0, 1] += 2.0                       # keep the map orientation-preserving
    a[1, 2] += 2.0

    def f(x, e):
        b = np.array([1, x, e, x * x, x * e, e * e, x * x * e, x * e * e])
        return a @ b

    def d(x, e):
        bx = np.array([0, 1, 0, 2 * x, e, 0, 2 * x * e, e * e])
        be = np.array([0, 0, 1, 0, x, 2 * e, x * x, 2 * x * e])
        return a @ bx, a @ be            # (dx/dxi, dy/dxi), (dx/deta, dy/deta)
    return f, d


@pytest.mark.parametrize("N", [3, 4, 5])
def test_metrics8_exact_for_serendipity_maps(N):
    """For a map in the serendipity space, sampled at the 8 nodes, Jac, S1
    and S2 equal the map's exact derivatives at every solution and flux
    point."""
    rng = np.random.default_rng(N)
    maps = [_quad_map(rng) for _ in range(3)]
    x8 = np.array([[f(*n) for n in NODES] for f, _ in maps])        # (3, 8, 2)
    cells = np.array([[1, 2, 3, 4]] * 3)
    jac, s1, s2 = M.metrics(cells, None, None, N, x8)
    xs, xf = M.setsandpoints(N)
    for c, (_, d) in enumerate(maps):
        for js in range(N):
            for is_ in range(N):
                (xp, yp), (xe, ye) = d(xs[is_], xs[js])
                assert abs(jac[is_, js, c] - (xp * ye - xe * yp)) < 1e-13
        for jfp in range(N):
            for ifp in range(N + 1):
                (xp, yp), (xe, ye) = d(xf[ifp], xs[jfp])
                np.testing.assert_allclose(s1[:, :, ifp, jfp, c], [[ye, -xe], [-yp, xp]],
                                           atol=1e-13)
        for jfp in range(N + 1):
            for ifp in range(N):
                (xp, yp), (xe, ye) = d(xs[ifp], xf[jfp])
                np.testing.assert_allclose(s2[:, :, ifp, jfp, c], [[ye, -xe], [-yp, xp]],
                                           atol=1e-13)


def test_straight_8node_cells_equal_bilinear():
    """An 8-node cell whose midside nodes are the midpoints of straight
    edges is the bilinear cell: the same metrics and coordinates to
    rounding."""
    m = M.channel(5, 4, seed=3, free_left=2)
    cells = M.orient_ccw(m)
    X = m.xv[cells - 1]; Y = m.yv[cells - 1]
    x8 = np.empty((len(cells), 8, 2))
    x8[:, :4, 0] = X; x8[:, :4, 1] = Y
    for k, (a, b) in enumerate(M.IVFACE):
        x8[:, 4 + k, 0] = 0.5 * (X[:, a] + X[:, b])
        x8[:, 4 + k, 1] = 0.5 * (Y[:, a] + Y[:, b])
    j4, a4, b4 = M.metrics(cells, m.xv, m.yv, 4)
    j8, a8, b8 = M.metrics(cells, m.xv, m.yv, 4, x8)
    for u, v in ((j4, j8), (a4, a8), (b4, b8)):
        np.testing.assert_allclose(v, u, rtol=0, atol=1e-14 * np.abs(u).max())
    c4 = M.solution_coords(cells, m.xv, m.yv, 4)
    c8 = M.solution_coords(cells, m.xv, m.yv, 4, x8)
    np.testing.assert_allclose(c8, c4, rtol=0, atol=1e-14)


def test_curved_ogrid_follows_the_circles():
    """The curved O-grid's edge nodes lie on their circles, its Jacobians
    are positive, and the mesh area (sum of Jac over Gauss-Legendre points
    of the cells) approaches the annulus area far closer than the
    polygonal mesh's."""
    r0, r1 = 0.5, 4.0
    mc = M.ogrid(6, 16, r0, r1, curved=True)
    mp = M.ogrid(6, 16, r0, r1)
    cc = M.build_case(mc, M.RunParams(N=4, curve_wall=1))
    cp = M.build_case(mp, M.RunParams(N=4))
    assert (cc.jac > 0).all()
    r = np.hypot(mc.x8[:, 7, 0], mc.x8[:, 7, 1])            # inner arc midpoints
    ring0 = np.arange(16)
    np.testing.assert_allclose(r[ring0], r0, rtol=1e-15)
    # area with a tensor Gauss-Legendre rule on [0,1]^2 (exact for the
    # polynomial Jacobian of either map at 4 points)
    g, w = np.polynomial.legendre.leggauss(4)
    g = 0.5 * (g + 1); w = 0.5 * w
    area = {}
    for name, m, x8 in (("c", mc, mc.x8), ("p", mp, None)):
        cells = M.orient_ccw(m)
        tot = 0.0
        for i, xi in enumerate(g):
            for j, et in enumerate(g):
                d = M._dmdx8(xi, et) if x8 is not None else M._dmdx(xi, et)
                XV = x8[:, :, 0] if x8 is not None else m.xv[cells - 1]
                YV = x8[:, :, 1] if x8 is not None else m.yv[cells - 1]
                xp = sum(d[k][0] * XV[:, k] for k in range(len(d)))
                yp = sum(d[k][0] * YV[:, k] for k in range(len(d)))
                xe = sum(d[k][1] * XV[:, k] for k in range(len(d)))
                ye = sum(d[k][1] * YV[:, k] for k in range(len(d)))
                tot += w[i] * w[j] * (xp * ye - xe * yp).sum()
        area[name] = tot
    exact = np.pi * (r1 ** 2 - r0 ** 2)
    assert abs(area["c"] - exact) < 0.01 * abs(area["p"] - exact)
    assert (cp.jac > 0).all()


def test_freestream_preserved_on_8node_cells():
    """A uniform state is a steady solution on the curved O-grid (far
    field all INLE, the wall made INLE too): the metric identities hold for
    the quadratic map to rounding, so the residual vanishes to round-off."""
    m = M.ogrid(6, 12, 0.5, 15.0, farfield="inle", curved=True)
    m.bnd = [(a, b, "INLE" if t == "WALL" else t) for a, b, t in m.bnd]
    p = M.RunParams(N=4, bxinf=0.3, byinf=0.1, curve_wall=1)
    case = M.build_case(m, p)
    Q = M.freestream_state(case)
    o = Oracle(case)
    o.set_state(Q, [0, 0, 0, 0])
    for _ in range(2):
        r, g = o.compresid()
        scale = np.abs(Q).max() * p.uinf / np.abs(case.jac).min() ** 0.5
        assert np.abs(r).max() < 1e-11 * scale, np.abs(r).max()
        assert np.abs(g).max() < 1e-10
    o.close()


def test_vrt8_and_deck_round_trip(tmp_path):
    m = M.ogrid(3, 10, 0.5, 4.0, curved=True)
    p = M.RunParams(N=3, curve_wall=1)
    M.write_quad_inp(str(tmp_path / "QUAD.INP"), p)
    M.write_mesh_files(str(tmp_path), m)
    qi = M.read_quad_inp(str(tmp_path / "QUAD.INP"))
    assert qi.curve_wall == 1 and qi.params.curve_wall == 1
    x8 = M.read_vrt8(str(tmp_path / qi.namevrt8), m.ncell)
    assert np.array_equal(x8, m.x8)
    with pytest.raises(ValueError):
        M.read_vrt8(str(tmp_path / qi.namevrt8), m.ncell + 1)


def test_refusals(tmp_path):
    m = M.ogrid(3, 10, 0.5, 4.0, curved=True)
    with pytest.raises(ValueError):              # CURVE_WALL = 1 without the 8 nodes
        M.build_case(M.ogrid(3, 10, 0.5, 4.0), M.RunParams(N=3, curve_wall=1))
    with pytest.raises(ValueError):
        M.build_case(m, M.RunParams(N=3, curve_wall=2))
    flipped = M.Mesh(m.cells[:, [1, 0, 3, 2]], m.xv, m.yv, m.bnd, x8=m.x8)
    with pytest.raises(ValueError):              # a clockwise cell would be mirrored
        M.build_case(flipped, M.RunParams(N=3, curve_wall=1))
    # a curved deck must start from a restart file
    from sddc_amd import run as R
    p = M.RunParams(N=3, curve_wall=1, restart=0)
    M.write_quad_inp(str(tmp_path / "QUAD.INP"), p)
    M.write_mesh_files(str(tmp_path), m)
    with pytest.raises(ValueError):
        R.run_dir(str(tmp_path))

"""Host-side mesh, topology, SD operators and metrics for the SD2D-MHD path.

Everything here runs once per mesh on the host and produces the tables the
C ABI (include/sd2d.h, struct sd2d_mesh) takes, in the reference's own
Fortran layouts and 1-based conventions:

* mesh generators: O-grid around a cylinder, periodic square, and a
  channel with every boundary type and shuffled local vertex order;
* writers/readers for the reference's input formats (QUAD.INP, .cel, .vrt,
  .bnd, restart stream file) so the same case can be run by the reference;
* topology (face numbering, left/right cells, boundary lists, periodic
  pairs) with the same results as CONNECTIVITY / CONNECT_BDRY / MAPCYCLOC;
* solution/flux point sets and the Lmat/Mmat operators (initsetup.f90);
* bilinear metrics Jac, S1, S2 (mapder.f90, calcjacob.f90).

The arithmetic of the operators and metrics follows the reference term by
term (same operation order, no fused multiply-add), so the arrays are
bitwise identical to what INIT_SETUP computes.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

GAM = 5.0 / 3.0          # setup2d.f90:67
PI = math.atan(1.0) * 4.0  # setup2d.f90:66
TOLCYC = 1e-5            # setup2d.f90:62

BC_TYPES = ("INLE", "OUTL", "SYMP", "WALL", "CYCL", "FREE")


# --------------------------------------------------------------------------
# Mesh container and generators
# --------------------------------------------------------------------------
@dataclass
class Mesh:
    """Unstructured quad mesh in the reference's global numbering.

    cells: (NC,4) 1-based vertex ids (the .cel columns IV1..IV4)
    xv, yv: (NV,) vertex coordinates (.vrt)
    bnd: list of (iv1, iv2, type) boundary edges, 1-based (.bnd)
    x8: (NC,8,2) or None: the eight nodes of every cell (.vrt8, read when
        CURVE_WALL = 1, readGRIDinfo.f90:86-117): the corners in the
        cell's vertex order, then the midside nodes of its edges 1-2, 2-3,
        3-4 and 4-1 (the node numbering drawn at mapder.f90:107-116)
    """
    cells: np.ndarray
    xv: np.ndarray
    yv: np.ndarray
    bnd: list
    dxcycl: float = 0.0
    dycycl: float = 0.0
    name: str = "mesh"
    x8: np.ndarray | None = None

    @property
    def ncell(self) -> int:
        return int(self.cells.shape[0])


def ogrid(nr: int, ntheta: int, r0: float = 0.5, r1: float = 15.0,
          farfield: str = "split", curved: bool = False) -> Mesh:
    """O-grid around a cylinder of radius r0 out to r1.

    Radial spacing is geometric, r_i = r0*q**i with q = (r1/r0)**(1/nr).
    Cells are numbered ring by ring (theta fastest); vertex order is CCW so
    local face 4 lies on the inner circle and face 2 on the outer one.  The
    seam at theta=0 shares its vertices, so it is an ordinary interior face.
    The inner circle is WALL.  The outer circle is INLE where the face
    midpoint has x < 0 and OUTL elsewhere (farfield="split"), or INLE
    everywhere (farfield="inle").

    curved=True adds the 8-node description of every cell (Mesh.x8, for
    CURVE_WALL = 1): the midside node of an arc edge lies on its circle at
    the middle angle, that of a radial edge at the mean radius of its ends,
    so every cell follows the circles exactly to second order.  Nodes
    shared by two cells are computed once, so both cells see the same bits.
    """
    q = (r1 / r0) ** (1.0 / nr)
    r = np.array([r0 * q ** i for i in range(nr + 1)])
    r[-1] = r1
    th = 2.0 * np.pi * np.arange(ntheta) / ntheta
    X = (r[:, None] * np.cos(th)[None, :]).ravel()
    Y = (r[:, None] * np.sin(th)[None, :]).ravel()

    def vid(i, j):
        return i * ntheta + (j % ntheta) + 1

    cells = np.empty((nr * ntheta, 4), dtype=np.int64)
    ii, jj = np.meshgrid(np.arange(nr), np.arange(ntheta), indexing="ij")
    ii = ii.ravel(); jj = jj.ravel()
    cells[:, 0] = vid(ii, jj)
    cells[:, 1] = vid(ii + 1, jj)
    cells[:, 2] = vid(ii + 1, jj + 1)
    cells[:, 3] = vid(ii, jj + 1)
    bnd = []
    for j in range(ntheta):
        bnd.append((vid(0, j + 1), vid(0, j), "WALL"))
    for j in range(ntheta):
        a, b = vid(nr, j), vid(nr, j + 1)
        xm = 0.5 * (X[a - 1] + X[b - 1])
        t = "INLE" if (farfield == "inle" or xm < 0.0) else "OUTL"
        bnd.append((a, b, t))
    if not curved:
        return Mesh(cells, X, Y, bnd, name=f"ogrid{nr}x{ntheta}")
    thm = 2.0 * np.pi * (np.arange(ntheta) + 0.5) / ntheta      # middle angle of column j
    rm = 0.5 * (r[:-1] + r[1:])                                  # mean radius of ring i
    x8 = np.empty((nr * ntheta, 8, 2))
    for k in range(4):
        x8[:, k, 0] = X[cells[:, k] - 1]
        x8[:, k, 1] = Y[cells[:, k] - 1]
    j1 = (jj + 1) % ntheta
    x8[:, 4, 0] = rm[ii] * np.cos(th[jj]); x8[:, 4, 1] = rm[ii] * np.sin(th[jj])          # 1-2
    x8[:, 5, 0] = r[ii + 1] * np.cos(thm[jj]); x8[:, 5, 1] = r[ii + 1] * np.sin(thm[jj])  # 2-3
    x8[:, 6, 0] = rm[ii] * np.cos(th[j1]); x8[:, 6, 1] = rm[ii] * np.sin(th[j1])          # 3-4
    x8[:, 7, 0] = r[ii] * np.cos(thm[jj]); x8[:, 7, 1] = r[ii] * np.sin(thm[jj])          # 4-1
    return Mesh(cells, X, Y, bnd, name=f"ogrid_curved{nr}x{ntheta}", x8=x8)


def periodic_square(nx: int, ny: int, lx: float = 2 * math.pi,
                    ly: float = 2 * math.pi) -> Mesh:
    """Uniform nx*ny quads on [0,lx]x[0,ly]; all boundary faces CYCL."""
    xs = np.array([lx * i / nx for i in range(nx + 1)])
    ys = np.array([ly * j / ny for j in range(ny + 1)])
    X = np.tile(xs, ny + 1)
    Y = np.repeat(ys, nx + 1)

    def vid(i, j):
        return j * (nx + 1) + i + 1

    jj, ii = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ii = ii.ravel(); jj = jj.ravel()
    cells = np.stack([vid(ii, jj), vid(ii + 1, jj), vid(ii + 1, jj + 1),
                      vid(ii, jj + 1)], axis=1).astype(np.int64)
    bnd = []
    for i in range(nx):
        bnd.append((vid(i, 0), vid(i + 1, 0), "CYCL"))
        bnd.append((vid(i + 1, ny), vid(i, ny), "CYCL"))
    for j in range(ny):
        bnd.append((vid(0, j + 1), vid(0, j), "CYCL"))
        bnd.append((vid(nx, j), vid(nx, j + 1), "CYCL"))
    return Mesh(cells, X, Y, bnd, dxcycl=lx, dycycl=ly,
                name=f"periodic{nx}x{ny}")


def channel(nx: int, ny: int, lx: float = 2.0, ly: float = 1.0,
            jitter: float = 0.15, seed: int = 0,
            types=("SYMP", "OUTL", "WALL", "INLE"), free_left: int = 0,
            rotate: bool = True) -> Mesh:
    """Rectangle with a different boundary type on each side.

    types gives the boundary type of (bottom, right, top, left).  Interior
    vertices are jittered by up to `jitter` of a cell size so the metrics
    vary from cell to cell, and (rotate=True) the local vertex order of
    every cell is cyclically shifted by a seeded random amount, so
    neighbouring cells meet through every pair of local faces.  The first
    `free_left` left-side faces are FREE instead of types[3].
    """
    rng = np.random.default_rng(seed)
    dx, dy = lx / nx, ly / ny
    X = np.zeros((ny + 1, nx + 1)); Y = np.zeros((ny + 1, nx + 1))
    for j in range(ny + 1):
        for i in range(nx + 1):
            x, y = i * dx, j * dy
            if 0 < i < nx and 0 < j < ny:
                x += jitter * dx * (rng.random() - 0.5)
                y += jitter * dy * (rng.random() - 0.5)
            X[j, i], Y[j, i] = x, y
    X = X.ravel(); Y = Y.ravel()

    def vid(i, j):
        return j * (nx + 1) + i + 1

    cells = []
    for j in range(ny):
        for i in range(nx):
            c = [vid(i, j), vid(i + 1, j), vid(i + 1, j + 1), vid(i, j + 1)]
            if rotate:
                s = int(rng.integers(0, 4))
                c = c[s:] + c[:s]
            cells.append(c)
    cells = np.array(cells, dtype=np.int64)
    bnd = []
    for i in range(nx):
        bnd.append((vid(i, 0), vid(i + 1, 0), types[0]))
    for j in range(ny):
        bnd.append((vid(nx, j), vid(nx, j + 1), types[1]))
    for i in range(nx):
        bnd.append((vid(i + 1, ny), vid(i, ny), types[2]))
    for j in range(ny):
        t = "FREE" if j < free_left else types[3]
        bnd.append((vid(0, j + 1), vid(0, j), t))
    return Mesh(cells, X, Y, bnd, name=f"channel{nx}x{ny}")


# --------------------------------------------------------------------------
# Reference input formats (readinput.f90, readGRIDinfo.f90,
# read_data_restart.f90, write_data.f90)
# --------------------------------------------------------------------------
@dataclass
class RunParams:
    """The QUAD.INP records of readinput.f90:6-43 that matter here."""
    N: int = 4
    vismode: int = 1
    mu: float = 0.0
    eta0: float = 0.0
    k_stage: int = 3
    dt: float = 1e-3
    maxiter: int = 1
    dxcycl: float = 0.0
    dycycl: float = 0.0
    artiv: int = 1
    nre: int = 1
    nwrite: int = 100000
    restart: int = 1
    rinf: float = 1.0
    pinf: float = 1.0 / GAM
    uinf: float = 1.5
    vinf: float = 0.0
    bxinf: float = 0.3
    byinf: float = 0.0
    alpha: float = 2.0  # initsetup.f90:165
    curve_wall: int = 0  # CURVE_WALL: 1 = 8-node cells from the .vrt8 file


def _f(x: float) -> str:
    return repr(float(x)).replace("e", "E") if "e" in repr(float(x)) else repr(float(x))


def write_quad_inp(path: str, p: RunParams, names=("mesh.cel", "mesh.vrt",
                   "mesh.vrt8", "mesh.bnd", "mesh.mts", "ic.dat")) -> None:
    """QUAD.INP in the positional order of readinput.f90:8-43."""
    L = ["SD2D-MHD input", "files", *names, "N CURVE vismode mu eta0",
         str(p.N), str(p.curve_wall), str(p.vismode), _f(p.mu), _f(p.eta0),
         "k_stage dt MAXITER", str(p.k_stage), _f(p.dt), str(p.maxiter),
         "DXCYCL DYCYCL", _f(p.dxcycl), _f(p.dycycl), "ARTIV", str(p.artiv),
         "NRE nwrite", str(p.nre), str(p.nwrite), "restart",
         str(p.restart), str(p.N), "free stream", _f(p.rinf), _f(p.pinf),
         _f(p.uinf), _f(p.vinf), _f(p.bxinf), _f(p.byinf)]
    assert len(L) == 36
    with open(path, "w") as fh:
        fh.write("\n".join(L) + "\n")


def write_mesh_files(dirpath: str, m: Mesh, stem: str = "mesh") -> None:
    """.cel/.vrt/.bnd files (readGRIDinfo.f90:33,78,269)."""
    with open(os.path.join(dirpath, stem + ".cel"), "w") as fh:
        for i, c in enumerate(m.cells):
            fh.write(f"{i+1} {c[0]} {c[1]} {c[2]} {c[3]} 1 1\n")
    with open(os.path.join(dirpath, stem + ".vrt"), "w") as fh:
        for i in range(len(m.xv)):
            fh.write(f"{i+1} {m.xv[i]:.17E} {m.yv[i]:.17E}\n")
    with open(os.path.join(dirpath, stem + ".bnd"), "w") as fh:
        for i, (a, b, t) in enumerate(m.bnd):
            fh.write(f"{i+1} {a} {b} 1 1 {t}\n")
    with open(os.path.join(dirpath, stem + ".vrt8"), "w") as fh:
        if m.x8 is not None:     # eight "x y" records per cell (readGRIDinfo.f90:97-101)
            for c in range(m.ncell):
                for k in range(8):
                    fh.write(f"{m.x8[c, k, 0]:.17E} {m.x8[c, k, 1]:.17E}\n")
    open(os.path.join(dirpath, stem + ".mts"), "w").close()


def read_vrt8(path: str, ncell: int) -> np.ndarray:
    """The .vrt8 file as READ_VRT8_DATA reads it (readGRIDinfo.f90:94-113):
    eight list-directed (x, y) records per cell, cells in global order.
    Returns x8 (ncell, 8, 2)."""
    rec = _records(path)
    if len(rec) < 8 * ncell:
        raise ValueError(f"{path}: {len(rec)} records, {8 * ncell} needed for {ncell} cells")
    v = np.array([[_ld_float(r[0]), _ld_float(r[1])] for r in rec[:8 * ncell]])
    return v.reshape(ncell, 8, 2)


def write_restart(path: str, Q: np.ndarray, ctime: float = 0.0,
                  it: int = 0) -> None:
    """Restart stream file: int32 iter, float64 ctime, then per cell
    int32 ICG + float64 Q(9,N,N) (write_data.f90:44-74)."""
    N = Q.shape[1]
    nc = Q.shape[3]
    rec = np.dtype([("icg", "<i4"), ("q", "<f8", (9 * N * N,))])
    body = np.empty(nc, dtype=rec)
    body["icg"] = np.arange(1, nc + 1, dtype=np.int32)
    body["q"] = np.asfortranarray(Q).reshape(9 * N * N, nc, order="F").T
    with open(path, "wb") as fh:
        fh.write(np.int32(it).tobytes())
        fh.write(np.float64(ctime).tobytes())
        fh.write(body.tobytes())


def read_restart(path: str, N: int):
    """Inverse of write_restart (read_data_restart.f90:35-67)."""
    raw = open(path, "rb").read()
    it = int(np.frombuffer(raw[:4], "<i4")[0])
    ctime = float(np.frombuffer(raw[4:12], "<f8")[0])
    rec = np.dtype([("icg", "<i4"), ("q", "<f8", (9 * N * N,))])
    body = np.frombuffer(raw[12:], dtype=rec)
    nc = body.shape[0]
    Q = np.zeros((9, N, N, nc), order="F")
    flat = Q.reshape(9 * N * N, nc, order="F")
    flat[:, body["icg"] - 1] = body["q"].T
    return Q, ctime, it


def _ld_first(line: str) -> str:
    """First item of a Fortran list-directed record: a quoted string, or the
    text up to the first blank, comma or slash."""
    t = line.strip()
    if t[:1] in ("'", '"'):
        q = t[0]
        j = t.find(q, 1)
        return t[1:j] if j > 0 else t[1:]
    for k, ch in enumerate(t):
        if ch in " \t,/":
            return t[:k]
    return t


def _ld_float(tok: str) -> float:
    """A list-directed real: D and Q exponent letters read as E."""
    return float(tok.translate(str.maketrans("dDqQ", "eEeE")))


@dataclass
class QuadInput:
    """QUAD.INP as readinput.f90:6-43 reads it."""
    params: RunParams
    namecel: str
    namevrt: str
    namevrt8: str
    namebnd: str
    metis_cell: str
    namerestart: str
    curve_wall: int
    restart_ord: int


def read_quad_inp(path: str) -> QuadInput:
    """Read QUAD.INP record by record in the order of readinput.f90:8-43
    (36 records, one item each; the comment records 1, 2, 9, 15, 19, 22,
    24, 27 and 30 are skipped)."""
    with open(path) as fh:
        recs = fh.read().splitlines()
    if len(recs) < 36:
        raise ValueError(f"{path}: {len(recs)} records, QUAD.INP has 36")
    it = iter(enumerate(recs, start=1))

    def item(kind):
        k, line = next(it)
        tok = _ld_first(line)
        if tok == "":
            raise ValueError(f"{path}: record {k} is empty")
        return tok if kind is str else (int(tok) if kind is int else _ld_float(tok))

    def skip():
        next(it)

    skip(); skip()
    names = [item(str) for _ in range(6)]
    skip()
    N, curve, vismode, mu, eta0 = item(int), item(int), item(int), item(float), item(float)
    skip()
    k_stage, dt, maxiter = item(int), item(float), item(int)
    skip()
    dxcycl, dycycl = item(float), item(float)
    skip()
    artiv = item(int)
    skip()
    nre, nwrite = item(int), item(int)
    skip()
    restart, restart_ord = item(int), item(int)
    skip()
    rinf, pinf, uinf, vinf, bxinf, byinf = (item(float) for _ in range(6))
    p = RunParams(N=N, vismode=vismode, mu=mu, eta0=eta0, k_stage=k_stage, dt=dt,
                  maxiter=maxiter, dxcycl=dxcycl, dycycl=dycycl, artiv=artiv, nre=nre,
                  nwrite=nwrite, restart=restart, rinf=rinf, pinf=pinf, uinf=uinf,
                  vinf=vinf, bxinf=bxinf, byinf=byinf, curve_wall=curve)
    return QuadInput(p, *names, curve_wall=curve, restart_ord=restart_ord)


def _records(path: str):
    with open(path) as fh:
        return [ln.replace(",", " ").split() for ln in fh if ln.strip()]


def read_mesh_files(cel: str, vrt: str, bnd: str, dxcycl: float = 0.0,
                    dycycl: float = 0.0, name: str = "mesh") -> Mesh:
    """The .cel, .vrt and .bnd files as READ_CELL_DATA, READ_VRT_DATA and
    READ_BND_DATA read them (readGRIDinfo.f90:33,78,269): the i-th record
    of each file is the i-th cell / vertex / boundary edge; the leading
    number of a record is not used."""
    rc = _records(cel)
    cells = np.array([[int(r[1]), int(r[2]), int(r[3]), int(r[4])] for r in rc], dtype=np.int64)
    rv = _records(vrt)
    xv = np.array([_ld_float(r[1]) for r in rv])
    yv = np.array([_ld_float(r[2]) for r in rv])
    bd = []
    for r in _records(bnd):
        t = r[5].strip("'\"")[:4]
        if t not in BC_TYPES:
            raise ValueError(f"{bnd}: boundary type {r[5]!r} is not one of {BC_TYPES}")
        bd.append((int(r[1]), int(r[2]), t))
    if cells.size and (cells.min() < 1 or cells.max() > len(xv)):
        raise ValueError(f"{cel}: vertex number outside 1..{len(xv)}")
    return Mesh(cells, xv, yv, bd, dxcycl=dxcycl, dycycl=dycycl, name=name)


# --------------------------------------------------------------------------
# Topology (connectivity.f90, connect_bdry.f90, map_cyc_loc.f90)
# --------------------------------------------------------------------------
IVFACE = ((0, 1), (1, 2), (2, 3), (3, 0))  # connectivity.f90:73-83 (0-based)


@dataclass
class Topology:
    if2c: np.ndarray            # (NFACE,4) int32, 1-based, Fortran order
    if2v: np.ndarray            # (NFACE,2) 1-based vertex ids
    ic2f: np.ndarray            # (NC,4) 1-based face ids
    ibfinter: np.ndarray
    lists: dict = field(default_factory=dict)   # type -> face ids (1-based)
    cyc2fpair: np.ndarray = None                 # (NPAIR,2)


def orient_ccw(m: Mesh) -> np.ndarray:
    """Return cells with every cell counter-clockwise, as CHECK_MESH
    (connectivity.f90:253-296) does: a cell whose edge sum
    sum (x2-x1)(y1+y2) is positive gets vertex order (2,1,4,3)."""
    c = m.cells.copy()
    xs = m.xv[c - 1]; ys = m.yv[c - 1]
    s = np.zeros(len(c))
    for a, b in IVFACE:
        s = s + (xs[:, b] - xs[:, a]) * (ys[:, a] + ys[:, b])
    flip = s > 0
    c[flip] = c[flip][:, [1, 0, 3, 2]]
    return c


def build_topology(m: Mesh, dxcycl: float = None, dycycl: float = None
                   ) -> Topology:
    """Faces, boundary lists and periodic pairs with the numbering of the
    reference on one rank.

    Face ids follow CONNECTIVITY (connectivity.f90:96-150): cells are
    visited in order, local faces 1..4, and a face gets the next id when it
    is first met, so the left cell IF2C(:,1) is the lower-numbered cell and
    IF2C(:,3)/(:,4) are the local faces in the left/right cell
    (connectivity.f90:166-216).  Boundary lists keep the .bnd order
    (connect_bdry.f90:126-223), IBFINTER is in face order
    (connect_bdry.f90:228-244), and periodic pairs are found as in
    MAPCYCLOC (map_cyc_loc.f90:15-35).
    """
    dxcycl = m.dxcycl if dxcycl is None else dxcycl
    dycycl = m.dycycl if dycycl is None else dycycl
    cells = orient_ccw(m)
    nc = len(cells)
    a = np.stack([cells[:, i] for i, _ in IVFACE], axis=1)  # (nc,4)
    b = np.stack([cells[:, j] for _, j in IVFACE], axis=1)
    lo = np.minimum(a, b).ravel(); hi = np.maximum(a, b).ravel()
    key = lo.astype(np.int64) * (int(hi.max()) + 1) + hi
    uniq, first, inv, cnt = np.unique(key, return_index=True,
                                      return_inverse=True, return_counts=True)
    if cnt.max() > 2:
        raise ValueError("edge shared by more than two cells")
    # face id = order of first appearance in (cell, local face) order
    order = np.argsort(first, kind="stable")
    fid_of_uniq = np.empty_like(order); fid_of_uniq[order] = np.arange(len(order))
    fid = fid_of_uniq[inv]                       # (nc*4,) 0-based face id
    nface = len(uniq)
    slot = np.arange(nc * 4)
    left_slot = first[order]                     # per face id, first slot
    # second slot for shared faces
    right_slot = np.full(nface, -1, dtype=np.int64)
    is_second = slot != first[inv]
    right_slot[fid[is_second]] = slot[is_second]
    if2c = np.zeros((nface, 4), dtype=np.int32)
    if2c[:, 0] = left_slot // 4 + 1
    if2c[:, 2] = left_slot % 4 + 1
    sh = right_slot >= 0
    if2c[sh, 1] = right_slot[sh] // 4 + 1
    if2c[sh, 3] = right_slot[sh] % 4 + 1
    if2v = np.stack([a.ravel()[left_slot], b.ravel()[left_slot]], axis=1)
    ic2f = (fid.reshape(nc, 4) + 1).astype(np.int32)

    # boundary lists in .bnd order
    lists = {t: [] for t in BC_TYPES}
    lut = {int(k): int(f) for k, f in zip(uniq, fid_of_uniq)}
    mult = int(hi.max()) + 1
    bface = np.zeros(nface, dtype=bool)
    for (v1, v2, t) in m.bnd:
        k = min(v1, v2) * mult + max(v1, v2)
        f = lut.get(k)
        if f is None:
            raise ValueError(f"boundary edge {v1}-{v2} is not a cell edge")
        lists[t].append(f + 1)
        bface[f] = True
    ibfinter = (np.nonzero(if2c[:, 1] != 0)[0] + 1).astype(np.int32)
    unlisted = np.nonzero((if2c[:, 1] == 0) & ~bface)[0]
    if len(unlisted):
        raise ValueError(f"{len(unlisted)} boundary faces missing from .bnd "
                         "(they would become processor faces)")
    topo = Topology(if2c=if2c, if2v=if2v, ic2f=ic2f, ibfinter=ibfinter,
                    lists={t: np.array(v, dtype=np.int32)
                           for t, v in lists.items()})
    topo.cyc2fpair = _cyclic_pairs(m, topo, dxcycl, dycycl)
    return topo


def _cyclic_pairs(m: Mesh, topo: Topology, dxc: float, dyc: float):
    """Local periodic face pairs (connect_bdry.f90:281-308 classifies,
    map_cyc_loc.f90:15-35 pairs) with the reference's midpoint test:
    faces i, j pair when |xm_i - xm_j| is within TOLCYC of DXCYCL and
    |ym_i - ym_j| < TOLCYC, or the same with x and y exchanged.  Candidates
    come from a hash of the midpoints on a grid of spacing TOLCYC (a
    partner lies in one of the 3 x 3 cells around the point shifted by a
    period), so the cost is O(n) instead of the reference's all-pairs loop;
    every candidate is then checked with the reference's test."""
    cyc = topo.lists["CYCL"]
    if len(cyc) == 0:
        return np.zeros((0, 2), dtype=np.int32)
    v = topo.if2v[cyc - 1] - 1
    xm = (m.xv[v[:, 0]] + m.xv[v[:, 1]]) / 2
    ym = (m.yv[v[:, 0]] + m.yv[v[:, 1]]) / 2
    n = len(cyc)
    h = TOLCYC
    grid = {}
    for i in range(n):
        grid.setdefault((int(math.floor(xm[i] / h)), int(math.floor(ym[i] / h))), []).append(i)
    pairs = set()
    matched = np.zeros(n, dtype=bool)
    for i in range(n):
        for sx, sy in ((dxc, 0.0), (-dxc, 0.0), (0.0, dyc), (0.0, -dyc)):
            if sx == 0.0 and sy == 0.0:
                continue
            cx = int(math.floor((xm[i] + sx) / h))
            cy = int(math.floor((ym[i] + sy) / h))
            for gx in (cx - 1, cx, cx + 1):
                for gy in (cy - 1, cy, cy + 1):
                    for j in grid.get((gx, gy), ()):
                        if j == i:
                            continue
                        dx = abs(xm[i] - xm[j])
                        dy = abs(ym[i] - ym[j])
                        if (abs(dx - dxc) < TOLCYC and dy < TOLCYC) or \
                           (abs(dy - dyc) < TOLCYC and dx < TOLCYC):
                            matched[i] = True
                            pairs.add((min(i, j), max(i, j)))
    if not matched.all():
        raise ValueError("periodic faces without a local partner")
    if len(pairs) != n // 2:
        raise ValueError("error in matching CYCLOC faces")
    i1, i2 = np.array(sorted(pairs)).T
    return np.stack([cyc[i1], cyc[i2]], axis=1).astype(np.int32)


# --------------------------------------------------------------------------
# SD point sets and operators (initsetup.f90:171-344)
# --------------------------------------------------------------------------
def setsandpoints(n: int):
    """Xs = Chebyshev-Gauss points on [0,1], Xf = {0, Gauss-Legendre(n-1),
    1} with the closed forms of initsetup.f90:178-206 (n <= 6)."""
    xs = [0.5 * (1.0 - math.cos(float(2 * i - 1) * PI / (2 * float(n))))
          for i in range(1, n + 1)]
    s = math.sqrt
    if n == 1:
        xf = [0.0, 1.0]
    elif n == 2:
        xf = [0.0, 0.5, 1.0]
    elif n == 3:
        xf = [0.0, (1.0 - s(3.0) / 3.0) / 2.0, (1.0 + s(3.0) / 3.0) / 2.0, 1.0]
    elif n == 4:
        xf = [0.0, (1.0 - s(15.0) / 5.0) / 2.0, 0.5,
              (1.0 + s(15.0) / 5.0) / 2.0, 1.0]
    elif n == 5:
        a = s(525.0 + 70.0 * s(30.0)) / 35.0
        b = s(525.0 - 70.0 * s(30.0)) / 35.0
        xf = [0.0, (1.0 - a) / 2.0, (1.0 - b) / 2.0, (1.0 + b) / 2.0,
              (1.0 + a) / 2.0, 1.0]
    elif n == 6:
        a = s(245.0 + 14.0 * s(70.0)) / 21.0
        b = s(245.0 - 14.0 * s(70.0)) / 21.0
        xf = [0.0, (1.0 - a) / 2.0, (1.0 - b) / 2.0, 0.5, (1.0 + b) / 2.0,
              (1.0 + a) / 2.0, 1.0]
    else:
        raise ValueError("N outside 1..6 is not tabulated")
    return np.array(xs), np.array(xf)


def _lagrange(nodes, i, x):
    """hhval (initsetup.f90:322-344): Lagrange basis i of `nodes` at x, with
    the numerator and denominator products formed separately."""
    num = 1.0; den = 1.0
    for s in range(len(nodes)):
        if s != i:
            num = num * (x - nodes[s])
            den = den * (nodes[i] - nodes[s])
    return num / den


def lmat_mmat(N: int):
    """Lmat(ifp,is) = l_is(Xf(ifp)); Mmat(ifp,is) = h'_ifp(Xs(is))
    (computeLandM, initsetup.f90:262-305)."""
    xs, xf = setsandpoints(N)
    L = np.zeros((N + 1, N)); M = np.zeros((N + 1, N))
    for ifp in range(N + 1):
        for is_ in range(N):
            L[ifp, is_] = _lagrange(xs, is_, xf[ifp])
    for ifp in range(N + 1):
        for is_ in range(N):
            xval = xs[is_]
            tot = 0.0
            for k in range(N + 1):
                if k != ifp:
                    num = 1.0; den = 1.0
                    for s in range(N + 1):
                        if s != ifp and s != k:
                            num = num * (xval - xf[s])
                        if s != ifp:
                            den = den * (xf[ifp] - xf[s])
                    tot = tot + num / den
            M[ifp, is_] = tot
    return xs, xf, L, M


def av_projectors(N: int):
    """The two interpolation matrices of the AV sensor (AV.f90:18-63):
    Plow[isl,is] = hhval(is, Xs_low(isl)) and Phigh[is,isl] =
    hhval_low(isl, Xs(is))."""
    xs, _ = setsandpoints(N)
    xl, _ = setsandpoints(N - 1)
    Plow = np.array([[_lagrange(xs, i, xl[a]) for i in range(N)]
                     for a in range(N - 1)])
    Phigh = np.array([[_lagrange(xl, a, xs[i]) for a in range(N - 1)]
                      for i in range(N)])
    return xl, Plow, Phigh


# --------------------------------------------------------------------------
# Metrics (mapder.f90:1-76, calcjacob.f90:1-131), bilinear cells
# --------------------------------------------------------------------------
def _dmdx(psi, eta):
    """Bilinear shape-function derivatives (mapder.f90:18-28): rows k=1..4,
    columns d/dpsi, d/deta."""
    return ((eta - 1.0, psi - 1.0), (1.0 - eta, -psi), (eta, psi),
            (-eta, 1.0 - psi))


def _dmdx8(psi, eta):
    """Serendipity 8-node shape-function derivatives (MAPDER_8_NODE,
    mapder.f90:117-253, the same expressions at solution and flux points):
    rows k=1..8, columns d/dpsi, d/deta."""
    return (((1.0 - eta) * (4.0 * psi + 2.0 * eta - 3.0), (1.0 - psi) * (4.0 * eta + 2.0 * psi - 3.0)),
            ((1.0 - eta) * (4.0 * psi - 2.0 * eta - 1.0), psi * (4.0 * eta - 2.0 * psi - 1.0)),
            (eta * (4.0 * psi + 2.0 * eta - 3.0), psi * (4.0 * eta + 2.0 * psi - 3.0)),
            (eta * (4.0 * psi - 2.0 * eta - 1.0), (1.0 - psi) * (4.0 * eta - 2.0 * psi - 1.0)),
            (4.0 * (1.0 - eta) * (1.0 - 2.0 * psi), 4.0 * psi * (psi - 1.0)),
            (4.0 * eta * (1.0 - eta), 4.0 * psi * (1.0 - 2.0 * eta)),
            (4.0 * eta * (1.0 - 2.0 * psi), 4.0 * psi * (1.0 - psi)),
            (4.0 * eta * (eta - 1.0), 4.0 * (1.0 - psi) * (1.0 - 2.0 * eta)))


def metrics(cells: np.ndarray, xv: np.ndarray, yv: np.ndarray, N: int,
            x8: np.ndarray | None = None):
    """Jac(N,N,NC), S1(2,2,N+1,N,NC), S2(2,2,N,N+1,NC) in Fortran order.

    Bilinear cells (x8 None): cells must already be CCW (orient_ccw) and the
    sums run over the four vertices k=1..4 from 0.0 as in
    calcjacob.f90:39-46.  8-node cells (CURVE_WALL = 1): the sums run over
    the eight nodes x8 (Mesh.x8) with the derivatives of MAPDER_8_NODE
    (calcjacob.f90:26-29,47-53).  Either way the values are bitwise those of
    the reference (compiled without FMA contraction)."""
    xs, xf = setsandpoints(N)
    nc = len(cells)
    if x8 is None:
        XV = xv[cells - 1]; YV = yv[cells - 1]   # (nc,4)
        dmdx, K = _dmdx, 4
    else:
        XV = np.ascontiguousarray(x8[:, :, 0]); YV = np.ascontiguousarray(x8[:, :, 1])
        dmdx, K = _dmdx8, 8

    def derivs(psi, eta):
        d = dmdx(psi, eta)
        xp = np.zeros(nc); yp = np.zeros(nc); xe = np.zeros(nc); ye = np.zeros(nc)
        for k in range(K):
            xp = xp + d[k][0] * XV[:, k]
            yp = yp + d[k][0] * YV[:, k]
            xe = xe + d[k][1] * XV[:, k]
            ye = ye + d[k][1] * YV[:, k]
        return xp, yp, xe, ye

    jac = np.zeros((N, N, nc), order="F")
    for js in range(N):
        for is_ in range(N):
            xp, yp, xe, ye = derivs(xs[is_], xs[js])
            jac[is_, js] = xp * ye - xe * yp
    s1 = np.zeros((2, 2, N + 1, N, nc), order="F")
    for jfp in range(N):
        for ifp in range(N + 1):
            xp, yp, xe, ye = derivs(xf[ifp], xs[jfp])
            s1[0, 0, ifp, jfp] = ye; s1[0, 1, ifp, jfp] = -xe
            s1[1, 0, ifp, jfp] = -yp; s1[1, 1, ifp, jfp] = xp
    s2 = np.zeros((2, 2, N, N + 1, nc), order="F")
    for jfp in range(N + 1):
        for ifp in range(N):
            xp, yp, xe, ye = derivs(xs[ifp], xf[jfp])
            s2[0, 0, ifp, jfp] = ye; s2[0, 1, ifp, jfp] = -xe
            s2[1, 0, ifp, jfp] = -yp; s2[1, 1, ifp, jfp] = xp
    return jac, s1, s2


def base_functions(xi, eta, nodes: int = 4):
    """MapBaseFunc (mapder.f90:79-87) or MapBaseFunc8 (mapder.f90:91-102)
    at (xi, eta): the 4 or 8 node weights of xyCoor_atGps."""
    if nodes == 4:
        return ((1.0 - xi) * (1.0 - eta), (1.0 - eta) * xi, xi * eta, (1.0 - xi) * eta)
    return ((1.0 - xi) * (1.0 - eta) * (-2.0 * xi - 2.0 * eta + 1.0),
            xi * (1.0 - eta) * (2.0 * xi - 2.0 * eta - 1.0),
            xi * eta * (2.0 * xi + 2.0 * eta - 3.0),
            (1.0 - xi) * eta * (-2.0 * xi + 2.0 * eta - 1.0),
            4.0 * (1.0 - eta) * (1.0 - xi) * xi,
            4.0 * eta * (1.0 - eta) * xi,
            4.0 * (1.0 - xi) * xi * eta,
            4.0 * (1.0 - eta) * (1.0 - xi) * eta)


def solution_coords(cells: np.ndarray, xv: np.ndarray, yv: np.ndarray,
                    N: int, x8: np.ndarray | None = None):
    """Physical coordinates of the solution points, (2,N,N,NC) as XXsolu
    (initsetup.f90:91-106, xyCoor_atGps mapder.f90:255-273), from the four
    vertices or, for 8-node cells, from x8 (Mesh.x8).  (The reference
    leaves the node array of this loop unset for 8-node cells,
    initsetup.f90:88-89, so its restart = 0 start on such a mesh reads
    undefined coordinates; only its restart = 1 start is defined there.)"""
    xs, _ = setsandpoints(N)
    if x8 is None:
        XV = xv[cells - 1]; YV = yv[cells - 1]; K = 4
    else:
        XV = x8[:, :, 0]; YV = x8[:, :, 1]; K = 8
    out = np.zeros((2, N, N, len(cells)), order="F")
    for js in range(N):
        for is_ in range(N):
            f = base_functions(xs[is_], xs[js], K)
            px = np.zeros(len(cells)); py = np.zeros(len(cells))
            for k in range(K):
                px = px + XV[:, k] * f[k]
                py = py + YV[:, k] * f[k]
            out[0, is_, js] = px; out[1, is_, js] = py
    return out


def case_coords(case: "Case") -> np.ndarray:
    """solution_coords of a case, with its 8-node cells when it has them."""
    return solution_coords(case.cells, case.mesh.xv, case.mesh.yv, case.N,
                           case.mesh.x8 if case.params.curve_wall == 1 else None)


# --------------------------------------------------------------------------
# Assembled case
# --------------------------------------------------------------------------
@dataclass
class Case:
    """Everything sd2d_create needs, plus the mesh it came from."""
    mesh: Mesh
    params: RunParams
    cells: np.ndarray       # CCW cells (1-based vertices)
    topo: Topology
    xs: np.ndarray
    xf: np.ndarray
    lmat: np.ndarray
    mmat: np.ndarray
    xs_low: np.ndarray
    jac: np.ndarray
    s1: np.ndarray
    s2: np.ndarray

    @property
    def N(self) -> int:
        return self.params.N

    @property
    def ncell(self) -> int:
        return len(self.cells)


def build_case(m: Mesh, p: RunParams) -> Case:
    N = p.N
    topo = build_topology(m, p.dxcycl, p.dycycl)
    cells = orient_ccw(m)
    x8 = None
    if p.curve_wall == 1:
        if m.x8 is None or m.x8.shape != (m.ncell, 8, 2):
            raise ValueError("CURVE_WALL = 1 needs the 8 nodes of every cell (Mesh.x8)")
        if not np.array_equal(cells, np.asarray(m.cells)):
            # CHECK_MESH turns a clockwise cell's vertices round but leaves its
            # .vrt8 nodes as read, so the 8-node map of that cell would be
            # mirrored (negative Jacobian)
            raise ValueError("CURVE_WALL = 1 needs every cell counter-clockwise in the .cel file")
        x8 = m.x8
    elif p.curve_wall != 0:
        raise ValueError(f"CURVE_WALL must be 0 or 1, not {p.curve_wall}")
    xs, xf, L, M = lmat_mmat(N)
    xl, _ = setsandpoints(N - 1) if N > 1 else (np.zeros(0), None)
    jac, s1, s2 = metrics(cells, m.xv, m.yv, N, x8)
    return Case(m, p, cells, topo, xs, xf, L, M, np.asarray(xl), jac, s1, s2)


# --------------------------------------------------------------------------
# Initial states (conserved variables, Fortran layout Q(9,N,N,NC))
# --------------------------------------------------------------------------
def conserved(rho, u, v, w, p, bx, by, bz, psi, gam=GAM):
    """Q from primitives, energy including 1/2 psi^2 as the reference's IC
    (setinitialcond.f90:31-32)."""
    e = p / (gam - 1) + 0.5 * rho * (u * u + v * v + w * w) + \
        0.5 * (bx * bx + by * by + bz * bz) + 0.5 * psi * psi
    return np.stack([rho, rho * u, rho * v, rho * w, e, bx, by, bz, psi])


def freestream_state(case: Case, rho_pert: float = 0.0, uv_pert: float = 0.0,
                     seed: int = 0, wall_ramp: tuple[float, float] | None = None
                     ) -> np.ndarray:
    """Free stream of the run parameters, optionally with a seeded uniform
    density perturbation rho*(1+rho_pert*U[-1,1]) at fixed pressure and a
    seeded 0..uv_pert perturbation of u and v (the reference's own
    perturbation, setinitialcond.f90:17-18, with a seed).

    wall_ramp=(r0, delta) starts the flow at rest on the circle r=r0 and
    lets the velocity rise smoothly to the free stream,
    u = u_inf*(1 - exp(-((r - r0)/delta)^2)), instead of the impulsive
    start of a uniform free stream against the wall."""
    p = case.params; N = p.N; nc = case.ncell
    rng = np.random.default_rng(seed)
    shp = (N, N, nc)
    rho = np.full(shp, p.rinf)
    if rho_pert:
        rho = rho * (1.0 + rho_pert * rng.uniform(-1.0, 1.0, size=shp))
    u = np.full(shp, p.uinf); v = np.full(shp, p.vinf)
    if wall_ramp is not None:
        xy = case_coords(case)
        r = np.hypot(xy[0], xy[1])
        s = 1.0 - np.exp(-(np.maximum(r - wall_ramp[0], 0.0) / wall_ramp[1]) ** 2)
        u = u * s; v = v * s
    if uv_pert:
        u = u + uv_pert * rng.random(shp)
        v = v + uv_pert * rng.random(shp)
    z = np.zeros(shp)
    Q = conserved(rho, u, v, z, np.full(shp, p.pinf), np.full(shp, p.bxinf),
                  np.full(shp, p.byinf), z, z)
    return np.asfortranarray(Q)


def orszag_tang_state(case: Case, rho_pert: float = 1e-4, seed: int = 0
                      ) -> np.ndarray:
    """Orszag-Tang vortex on [0,2pi]^2: rho=25/9, p=5/3, u=(-sin y, sin x),
    B=(-sin y, sin 2x), plus a seeded normal perturbation of rho."""
    xy = case_coords(case)
    x, y = xy[0], xy[1]
    rng = np.random.default_rng(seed)
    rho = np.full(x.shape, 25.0 / 9.0)
    if rho_pert:
        rho = rho + rho_pert * rng.standard_normal(x.shape)
    z = np.zeros(x.shape)
    Q = conserved(rho, -np.sin(y), np.sin(x), z, np.full(x.shape, 5.0 / 3.0),
                  -np.sin(y), np.sin(2 * x), z, z)
    return np.asfortranarray(Q)


def smooth_perturbed_state(case: Case, amp: float = 0.05, seed: int = 0
                           ) -> np.ndarray:
    """Free stream plus smooth seeded perturbations of every variable
    (including w, Bz and psi), so that every term of the residual is
    exercised and no residual component cancels to round-off."""
    p = case.params
    xy = case_coords(case)
    x, y = xy[0], xy[1]
    rng = np.random.default_rng(seed)
    k = rng.uniform(0.5, 2.0, size=(9, 2)); ph = rng.uniform(0, 2 * np.pi, 9)

    def wave(i):
        return np.sin(k[i, 0] * x + ph[i]) * np.cos(k[i, 1] * y + 0.3 * ph[i])

    rho = p.rinf * (1 + amp * wave(0))
    u = p.uinf + amp * wave(1); v = p.vinf + amp * wave(2); w = amp * wave(3)
    pr = p.pinf * (1 + amp * wave(4))
    bx = p.bxinf + amp * wave(5); by = p.byinf + amp * wave(6)
    bz = amp * wave(7); psi = amp * 0.5 * wave(8)
    return np.asfortranarray(conserved(rho, u, v, w, pr, bx, by, bz, psi))

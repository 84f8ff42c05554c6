"""Tecplot output of the reference solver (tecplotter2dsetup.f90,
tecplotter.f90): the plot mesh, the nodal values and the binary #!TDV111
file, as the reference program writes them into TEC/ at its output steps
(iterations.f90:76-81).

    topo = plot_topology(case, nre)           # once per mesh and NRE
    plotq, plotav = solver.plot_nodes(topo)   # on the device, per output step
    write_plt(os.path.join(rundir, "TEC", tec_filename(ctime)), ctime, topo, plotq, plotav)

The plot mesh refines every cell into NRE x NRE quadrilaterals.  Its nodes
are numbered as tecplotter2dsetup.f90:32-152 numbers them: the mesh
vertices first, in the order in which the cells name them (the local
numbering of readGRIDinfo.f90:180-206); then, cell by cell, the NRE-1 nodes of each
face seen for the first time (faces k = 1..4 in order) and the (NRE-1)^2
interior nodes (ii outer, jj inner); a face's nodes seen again from the
other cell are found by distance (at most 1e-5).  A node's value is the
average over the cells that share it of the cell's solution polynomial at
the node (tecplotter.f90:17-60), converted to rho, u, v, w, p, Bx, By, Bz,
psi; muAV and etaAV are averaged likewise.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

from . import mesh as M

TOLPLOT = 1e-5                    # tecplotter2dsetup.f90:18
NVAR = 13                         # X, Y, 9 variables, muAV, etaAV
VARNAMES = ("X", "Y", "RHO", "U", "V", "W", "P", "Bx", "By", "Bz", "Az", "muAV", "etaAV")


@dataclass
class PlotTopology:
    """The plot mesh of one mesh and NRE (0-based node numbers)."""
    nre: int
    ncell: int
    xy: np.ndarray          # (2, nnode) node coordinates
    gnum: np.ndarray        # (ncell, nre+1, nre+1) node of (cell, ii, jj), ii/jj 0-based
    factor: np.ndarray      # (nnode,) cells sharing each node
    conn: np.ndarray        # (nplotcell, 4) node numbers of every plot cell
    xadj: np.ndarray        # (nnode+1,) CSR of the (cell, ii, jj) occurrences of each node,
    contrib: np.ndarray     #   ascending in cell: cell*(nre+1)^2 + ii + (nre+1)*jj

    @property
    def nnode(self) -> int:
        return self.xy.shape[1]


def _xy_at(xn, yn, xsi, eta):
    """xyCoor_atGps (mapder.f90:255-273) over the last axis of xn/yn: four
    vertices with the bilinear base functions or eight nodes with the
    8-node ones (mapder.f90:79-102), the node terms summed in order from 0."""
    K = xn.shape[-1]
    f = M.base_functions(xsi, eta, K)
    x = 0.0 * xn[..., 0]
    y = 0.0 * yn[..., 0]
    for k in range(K):
        x = x + xn[..., k] * f[k]
        y = y + yn[..., k] * f[k]
    return x, y


def plot_topology(case: M.Case, nre: int) -> PlotTopology:
    """The plot mesh of tecplotter2dsetup.f90:32-152: the cell vertices,
    then the nodes of the refined faces and cell interiors, placed with the
    four vertices (CURVE_WALL = 0) or with the eight nodes of the .vrt8
    file (CURVE_WALL = 1, tecplotter2dsetup.f90:58-66)."""
    if nre < 1:
        raise ValueError("NRE must be at least 1")
    mesh = case.mesh
    cells = np.asarray(case.cells, dtype=np.int64) - 1          # (nc, 4) 0-based vertices
    ic2f = np.asarray(case.topo.ic2f, dtype=np.int64) - 1       # (nc, 4) 0-based faces
    nc, nv = cells.shape[0], len(mesh.xv)
    Np = nre
    # the vertices numbered as GLOB2LOCAL numbers them (readGRIDinfo.f90:
    # 180-206): in the order of first appearance, cell by cell, k = 1..4
    if len(np.unique(cells)) != nv:
        raise ValueError("a mesh vertex belongs to no cell (tecplotter2dsetup.f90:58-60)")
    _, first = np.unique(cells.ravel(), return_index=True)
    glob = np.unique(cells.ravel())[np.argsort(first)]          # local -> mesh vertex
    local = np.empty(nv, dtype=np.int64)
    local[glob] = np.arange(nv)
    cells = local[cells]
    xv = np.asarray(mesh.xv, dtype=np.float64)[glob]
    yv = np.asarray(mesh.yv, dtype=np.float64)[glob]
    gnum = np.full((nc, Np + 1, Np + 1), -1, dtype=np.int64)
    # the four vertices (tecplotter2dsetup.f90:32-56)
    gnum[:, 0, 0] = cells[:, 0]
    gnum[:, Np, 0] = cells[:, 1]
    gnum[:, Np, Np] = cells[:, 2]
    gnum[:, 0, Np] = cells[:, 3]
    factor_v = np.bincount(cells.ravel(), minlength=nv)
    xy_parts = [np.vstack([xv, yv])]
    factor_parts = [factor_v]
    if Np > 1:
        ne = Np - 1
        # node positions of face k, enode = 1..Np-1 (tecplotter2dsetup.f90:76-91)
        e = np.arange(1, Np)
        face_ij = {0: (e + 1, np.ones_like(e)), 1: (np.full_like(e, Np + 1), e + 1),
                   2: (Np + 1 - e, np.full_like(e, Np + 1)), 3: (np.ones_like(e), Np + 1 - e)}
        if case.params.curve_wall == 1:
            X4 = np.ascontiguousarray(mesh.x8[:, :, 0])           # (nc, 8)
            Y4 = np.ascontiguousarray(mesh.x8[:, :, 1])
        else:
            X4, Y4 = xv[cells], yv[cells]                        # (nc, 4)
        # the first (cell, k) visit of every face creates its nodes
        occ = ic2f.ravel()                                      # visit order cell*4 + k
        order = np.arange(occ.size)
        first_idx = np.full(ic2f.max() + 1, occ.size, dtype=np.int64)
        np.minimum.at(first_idx, occ, order)
        creates = first_idx[occ] == order                        # (nc*4,)
        new_per_cell = creates.reshape(nc, 4).sum(axis=1) * ne + ne * ne
        base = nv + np.concatenate([[0], np.cumsum(new_per_cell)[:-1]])
        nnode = nv + int(new_per_cell.sum())
        xy = np.zeros((2, nnode))
        xy[:, :nv] = xy_parts[0]
        factor = np.zeros(nnode, dtype=np.int64)
        factor[:nv] = factor_v
        face_nodes = np.full((ic2f.max() + 1, ne), -1, dtype=np.int64)   # gfacenode
        # creating visits, in cell order: node numbers and coordinates
        cr = creates.reshape(nc, 4)
        slot = np.cumsum(cr, axis=1) - cr                        # created faces before k
        for k in range(4):
            ii, jj = face_ij[k]
            xsi = (ii - 1.0) / float(Np)
            eta = (jj - 1.0) / float(Np)
            cs = np.nonzero(cr[:, k])[0]
            ids = base[cs, None] + slot[cs, k, None] * ne + np.arange(ne)[None, :]
            x, y = _xy_at(X4[cs, None, :], Y4[cs, None, :], xsi[None, :], eta[None, :])
            xy[0, ids] = x
            xy[1, ids] = y
            factor[ids] += 1
            face_nodes[ic2f[cs, k]] = ids
            gnum[cs[:, None], ii[None, :] - 1, jj[None, :] - 1] = ids
        # visits of a face already nodalised: match by distance
        for k in range(4):
            ii, jj = face_ij[k]
            xsi = (ii - 1.0) / float(Np)
            eta = (jj - 1.0) / float(Np)
            cs = np.nonzero(~cr[:, k])[0]
            if cs.size == 0:
                continue
            x, y = _xy_at(X4[cs, None, :], Y4[cs, None, :], xsi[None, :], eta[None, :])
            cand = face_nodes[ic2f[cs, k]]                      # (m, ne) existing nodes
            dx = x[:, :, None] - xy[0, cand][:, None, :]
            dy = y[:, :, None] - xy[1, cand][:, None, :]
            dist = np.abs(np.sqrt(dx * dx + dy * dy))            # distcalc
            hit = dist <= TOLPLOT
            if not np.all(hit.any(axis=2)):
                raise ValueError("a face node has no match on the neighbouring cell "
                                 "(tecplotter2dsetup.f90:109-116)")
            # the last matching existing node is taken, every match counts
            last = ne - 1 - np.argmax(hit[:, :, ::-1], axis=2)
            ids = np.take_along_axis(cand, last, axis=1)
            gnum[cs[:, None], ii[None, :] - 1, jj[None, :] - 1] = ids
            np.add.at(factor, np.broadcast_to(cand[:, None, :], hit.shape)[hit], 1)
        # interior nodes, ii outer and jj inner (tecplotter2dsetup.f90:131-141)
        iin, jjn = np.meshgrid(np.arange(2, Np + 1), np.arange(2, Np + 1), indexing="ij")
        iin, jjn = iin.ravel(), jjn.ravel()
        xsi = (iin - 1.0) / float(Np)
        eta = (jjn - 1.0) / float(Np)
        ids = base[:, None] + cr.sum(axis=1)[:, None] * ne + np.arange(ne * ne)[None, :]
        x, y = _xy_at(X4[:, None, :], Y4[:, None, :], xsi[None, :], eta[None, :])
        xy[0, ids] = x
        xy[1, ids] = y
        factor[ids] += 1
        gnum[:, iin - 1, jjn - 1] = ids
    else:
        xy = xy_parts[0]
        factor = factor_parts[0]
    # plot cells, ii outer and jj inner (tecplotter2dsetup.f90:143-152)
    I, J = np.meshgrid(np.arange(Np), np.arange(Np), indexing="ij")
    I, J = I.ravel(), J.ravel()
    conn = np.stack([gnum[:, I, J], gnum[:, I + 1, J], gnum[:, I + 1, J + 1], gnum[:, I, J + 1]],
                    axis=2).reshape(-1, 4)
    # every node's occurrences, ascending in cell (the order of tecplotter.f90:20-43)
    occ_node = gnum.transpose(0, 2, 1).reshape(-1)               # cell, jj, ii
    cell_idx = np.repeat(np.arange(nc), (Np + 1) ** 2)
    jj_idx = np.tile(np.repeat(np.arange(Np + 1), Np + 1), nc)
    ii_idx = np.tile(np.tile(np.arange(Np + 1), Np + 1), nc)
    code = cell_idx * (Np + 1) ** 2 + ii_idx + (Np + 1) * jj_idx
    srt = np.lexsort((code, occ_node))                           # by node, then cell
    nn = xy.shape[1]
    xadj = np.zeros(nn + 1, dtype=np.int64)
    np.add.at(xadj, occ_node + 1, 1)
    xadj = np.cumsum(xadj)
    if not np.array_equal(np.diff(xadj), factor):
        raise ValueError("internal error: node occurrences differ from the node factors")
    return PlotTopology(nre=nre, ncell=nc, xy=xy, gnum=gnum, factor=factor,
                        conn=conn.astype(np.int64), xadj=xadj.astype(np.int32),
                        contrib=code[srt].astype(np.int32))


def tec_filename(ctime: float) -> str:
    """'TEC' // i7.7 of int(ctime) // characters 3-6 of f6.3 of its
    fraction // '.plt' (tecplotter.f90:176-177)."""
    whole = int(ctime)
    dec = f"{ctime - whole:6.3f}"
    return f"TEC{whole:07d}{dec[2:6]}.plt"


def _tec_string(s: str) -> bytes:
    return np.array([ord(ch) for ch in s] + [0], dtype="<i4").tobytes()


def write_plt(path: str, ctime: float, topo: PlotTopology, plotq: np.ndarray,
              plotav: np.ndarray) -> None:
    """The binary file of tecpost_parallel (tecplotter.f90:90-403) for one
    rank: header, the 13 variables as float32 blocks, the connectivity
    (0-based).  plotq is (9, nnode) (rho, u, v, w, p, Bx, By, Bz, psi) and
    plotav (2, nnode)."""
    nn = topo.nnode
    cols = [topo.xy[0], topo.xy[1]] + [plotq[v] for v in range(9)] + [plotav[0], plotav[1]]
    minmax = np.empty((2, NVAR))
    for i, c in enumerate(cols):
        minmax[0, i] = np.min(c)
        minmax[1, i] = np.max(c)
    out = bytearray()
    out += b"#!TDV111"
    out += np.array([1, 0], dtype="<i4").tobytes()
    out += _tec_string("step data")
    out += np.array([NVAR], dtype="<i4").tobytes()
    for name in VARNAMES:
        out += _tec_string(name)
    out += np.array([299.0], dtype="<f4").tobytes()           # zone marker
    out += _tec_string("ZONE 001")
    out += np.array([-1, 0], dtype="<i4").tobytes()            # parent zone, strand id
    out += struct.pack("<d", ctime)                            # solution time
    out += np.array([-1, 3, 0, 0, 0, 0, nn, topo.conn.shape[0], 0, 0, 0, 0],
                    dtype="<i4").tobytes()
    out += np.array([357.0, 299.0], dtype="<f4").tobytes()     # end of header, zone
    out += np.array([1] * NVAR + [0, 0, -1], dtype="<i4").tobytes()
    out += minmax.ravel(order="F").astype("<f8").tobytes()     # minmax(2,nvar)
    for c in cols:
        out += np.asarray(c, dtype=np.float64).astype("<f4").tobytes()
    out += np.ascontiguousarray(topo.conn, dtype="<i4").tobytes()
    with open(path, "wb") as fh:
        fh.write(bytes(out))

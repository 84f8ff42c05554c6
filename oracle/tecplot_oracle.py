"""The reference's nodal values for the Tecplot output, restated in numpy
(tecplotter.f90:1-84, subroutine tecplotter2d and its function hval).
Test infrastructure: the checker of the device's sd2d_plot_nodes and of
the output pipeline (tests/); nothing in the product calls it.

Plain IEEE double arithmetic in the reference's order of operations, so
that the values are bit for bit those of the reference program:
  hval(i, x)  = prod_{s != i} (x - Xs(s)) / prod_{s != i} (Xs(i) - Xs(s)),
                both products from 1.0 with s ascending;
  Qs          = sum over js (outer), is (inner) of (Q(:,is,js) * hval(is, xsi)) * hval(js, eta),
                from 0.0, at xsi = (ii-1)/NRE, eta = (jj-1)/NRE;
  plotQ(node) = the Qs of every cell holding the node added in ascending cell
                order from 0.0, then divided by the number of those cells;
  rho, u = m_x/rho, v, w, p = (E - 0.5*rho*|u|^2 - 0.5*|B|^2)*(gam-1),
  Bx, By, Bz, psi; plotAV likewise from muAV(1:2, cell).
"""
from __future__ import annotations

import numpy as np

GAM = 5.0 / 3       # setup2d.f90:67


def hval(xs: np.ndarray, i: int, x: float) -> float:
    """tecplotter.f90:62-80."""
    hn = 1.0
    hd = 1.0
    for s in range(len(xs)):
        if s != i:
            hn = hn * (x - xs[s])
            hd = hd * (xs[i] - xs[s])
    return hn / hd


def nodal_values(Q: np.ndarray, muav: np.ndarray, topo, xs: np.ndarray):
    """plotQ (9, nnode) and plotAV (2, nnode) of tecplotter2d for the state
    Q(9,N,N,NCELL) and muAV(2,NCELL) on the plot mesh `topo`
    (sddc_amd.tecplot.plot_topology)."""
    N = Q.shape[1]
    Np = topo.nre
    h = np.array([[hval(xs, i, (k - 0.0) / float(Np)) for i in range(N)] for k in range(Np + 1)])
    code = topo.contrib.astype(np.int64)
    cell = code // (Np + 1) ** 2
    loc = code % (Np + 1) ** 2
    ii = loc % (Np + 1)
    jj = loc // (Np + 1)
    Qc = np.asarray(Q)[:, :, :, cell]                    # (9, N, N, nocc)
    qs = np.zeros((9, code.size))
    for js in range(N):
        for is_ in range(N):
            qs = qs + (Qc[:, is_, js, :] * h[ii, is_]) * h[jj, js]
    av = np.asarray(muav)[:, cell]                       # (2, nocc)
    nn = topo.nnode
    start = topo.xadj[:-1].astype(np.int64)
    cnt = np.diff(topo.xadj).astype(np.int64)
    sq = np.zeros((9, nn))
    sa = np.zeros((2, nn))
    for k in range(int(cnt.max())):
        m = cnt > k
        sq[:, m] = sq[:, m] + qs[:, start[m] + k]
        sa[:, m] = sa[:, m] + av[:, start[m] + k]
    f = topo.factor.astype(np.float64)
    Qs = sq / f
    plotq = np.empty((9, nn))
    plotq[0] = Qs[0]
    plotq[1] = Qs[1] / Qs[0]
    plotq[2] = Qs[2] / Qs[0]
    plotq[3] = Qs[3] / Qs[0]
    plotq[5:9] = Qs[5:9]
    nu = plotq[1] ** 2 + plotq[2] ** 2 + plotq[3] ** 2
    nb = plotq[5] ** 2 + plotq[6] ** 2 + plotq[7] ** 2
    plotq[4] = (Qs[4] - 0.5 * Qs[0] * nu - 0.5 * nb) * (GAM - 1)
    return plotq, sa / f

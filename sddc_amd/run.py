"""Run a reference input deck on the device: the time loop of the SD2D-MHD
program (main.f90, iterations.f90:1-84) with the device-resident stepping
of the C ABI in place of the host loop.

    python -m sddc_amd.run RUNDIR [--precision fp64|mixed] [--maxiter N]

RUNDIR holds what the reference program reads from its working directory:
QUAD.INP (readinput.f90:6-43), the .cel/.vrt/.bnd mesh files named there
(readGRIDinfo.f90) and, with restart = 1, the restart file
(read_data_restart.f90:35-67).  Written into RUNDIR, as the reference
writes them:

* resid.out: a header line, then `iter ctime RESIDNORM divB` on every step
  with mod(iter, 10) == 1 (iterations.f90:34-72);
* restart.%08d.dat on every step with mod(iter, nwrite) == 1 and after the
  last step (write_data.f90:44-74), each with its running-maxima sidecar
  restart.%08d.dat.max (sd2d_write_checkpoint).  A restart = 1 start from a
  file that has such a sidecar resumes bit for bit (the maxima restored);
  from a file without one it resets the maxima to 0 as the reference does
  (initsetup.f90:154-167).

* TEC/TEC%07d.%03d.plt (the whole part and the fraction of ctime), the
  Tecplot binary file of tecplotter.f90, on the same steps as the restart
  files: the plot mesh refined NRE x NRE per cell, the nodal values formed
  on the device (sd2d_plot_nodes) and the #!TDV111 writer of
  sddc_amd/tecplot.py.  Only with ARTIV = 1: the reference allocates muAV,
  which the file holds, only then (initsetup.f90:36-45).

The reference's
restart = 0 start (setinitialcond.f90:17-18) perturbs u and v by
0.02*ran1() from an unseeded generator; here the perturbation is drawn
from a seeded generator (--seed), so such a start is reproducible but not
the reference's own draw.  CURVE_WALL = 1 reads the 8 nodes of every cell
from the .vrt8 file named in QUAD.INP (readGRIDinfo.f90:86-117) and builds
the curved metrics from them; such a deck must start from a restart file,
as the reference's restart = 0 start reads coordinates it never sets for
8-node cells (initsetup.f90:82-89).
"""
from __future__ import annotations

import argparse
import os
import sys

from . import abi
from . import mesh as M
from . import tecplot as TP


def _next(it: int, every: int) -> int:
    """The first step after `it` with mod(step, every) == 1 (none for
    every <= 1, as mod(step, 1) is 0)."""
    if every <= 1:
        return sys.maxsize
    return it + 1 + ((1 - (it + 1)) % every)


def run_dir(d: str, precision: int = abi.SD2D_FP64, maxiter: int | None = None,
            seed: int = 0, log=None) -> dict:
    """Run the input deck in directory d; returns the resid.out rows and
    the names of the restart files written."""
    from .solver import Sd2d

    qi = M.read_quad_inp(os.path.join(d, "QUAD.INP"))
    p = qi.params
    if maxiter is not None:
        p.maxiter = int(maxiter)
    if qi.curve_wall not in (0, 1):
        raise ValueError(f"CURVE_WALL must be 0 or 1, not {qi.curve_wall}")
    if qi.curve_wall == 1 and p.restart != 1:
        raise ValueError("CURVE_WALL = 1 needs restart = 1: the reference's restart = 0 start "
                         "reads coordinates it leaves unset for 8-node cells")
    if p.restart == 1 and qi.restart_ord != p.N:
        # initsetup.f90:157-161 prints and leaves Q unset; refuse instead
        raise ValueError(f"restart_ord {qi.restart_ord} differs from N {p.N}")
    mesh = M.read_mesh_files(os.path.join(d, qi.namecel), os.path.join(d, qi.namevrt),
                             os.path.join(d, qi.namebnd), p.dxcycl, p.dycycl)
    if qi.curve_wall == 1:
        mesh.x8 = M.read_vrt8(os.path.join(d, qi.namevrt8), mesh.ncell)
    case = M.build_case(mesh, p)
    g = Sd2d(case, precision=precision)
    try:
        if p.restart == 1:
            rpath = os.path.join(d, qi.namerestart)
            if os.path.exists(rpath + ".max"):
                g.read_checkpoint(rpath)
            else:
                g.read_restart(rpath)
        else:
            g.set_state(M.freestream_state(case, uv_pert=0.02, seed=seed))
        dg = g.diag()
        it = int(dg["iter"])
        rows, written, tec = [], [], []
        topo = None
        with open(os.path.join(d, "resid.out"), "w") as fh:
            fh.write(" ITER    ctime     RESIDNORM   divB\n")
        nwrite = int(p.nwrite)
        while it < p.maxiter:
            nxt = min(_next(it, 10), _next(it, nwrite), p.maxiter)
            g.step(nxt - it, 10)
            it = nxt
            dg = g.diag()
            if it % 10 == 1:
                row = (it, dg["ctime"], dg["resid_rms"], dg["divb_rms"])
                rows.append(row)
                with open(os.path.join(d, "resid.out"), "a") as fh:
                    fh.write(f"{row[0]:12d} {row[1]:25.17E} {row[2]:25.17E} {row[3]:25.17E}\n")
                if log:
                    log(f"iter {it} ctime {row[1]:.6e} resid {row[2]:.6e} divB {row[3]:.6e} "
                        f"eigvmax {dg['eigvmax_g']:.6e}")
            if (nwrite > 1 and it % nwrite == 1) or it >= p.maxiter:
                if p.artiv == 1:
                    # tecplotter2dsetup, tecplotter2d, tecpost_parallel (iterations.f90:76-81)
                    if topo is None:
                        topo = TP.plot_topology(case, int(p.nre))
                    _, muav, _ = g.get_av()
                    plotq, plotav = g.plot_nodes(topo, muav)
                    os.makedirs(os.path.join(d, "TEC"), exist_ok=True)
                    tn = TP.tec_filename(dg["ctime"])
                    TP.write_plt(os.path.join(d, "TEC", tn), dg["ctime"], topo, plotq, plotav)
                    tec.append(tn)
                fn = f"restart.{it:08d}.dat"
                g.write_checkpoint(os.path.join(d, fn))
                written.append(fn)
        return {"resid_out": rows, "restarts": written, "tecplot": tec, "iter": it,
                "ctime": g.diag()["ctime"]}
    finally:
        g.close()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("rundir", nargs="?", default=".")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--maxiter", type=int, default=None,
                    help="override MAXITER of QUAD.INP")
    ap.add_argument("--seed", type=int, default=0,
                    help="seed of the u, v perturbation of a restart = 0 start")
    a = ap.parse_args(argv)
    prec = abi.SD2D_FP64 if a.precision == "fp64" else abi.SD2D_MIXED
    out = run_dir(a.rundir, prec, a.maxiter, a.seed, log=lambda m: print(m, flush=True))
    print(f"done: iter {out['iter']}, ctime {out['ctime']:.17e}, "
          f"restart files {', '.join(out['restarts'])}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Times the reference Fortran program itself (oracle/_ref/sd2d_MHD_O3, built
by oracle/ref.mk from the reference's sources with flang -O3) on a workload
of this path: the CPU baseline of kind "reference".

    python -m oracle.time_reference --case c1 --ranks 1 --steps 1000
    python -m oracle.time_reference --case c2 --ranks 16 --steps 2

The program times itself only as a whole (mpi_wtime around set-up,
stepping and output, main.f90:14,60-65), so it runs twice on the same deck,
for K and for 2K steps, and the time per step is (T(2K) - T(K)) / K: the
set-up, the restart read and the output written at the last step
(iterations.f90:76-81) cancel.  With more than one rank the cells are split
into contiguous ranges (the rings of an O-grid numbered ring by ring; the
reference reads one rank number per cell, readGRIDinfo.f90:134-139) and the
program runs under the MPICH of /opt/conda.  Measurement infrastructure:
imported only by bench.py's CPU-baseline leg.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "_ref", "sd2d_MHD_O3")
MPIRUN = "/opt/conda/bin/mpirun"


def available() -> bool:
    return os.path.exists(EXE)


def write_deck(d: str, case, Q0: np.ndarray, ranks: int) -> None:
    """Reference input deck: QUAD.INP, mesh files, the initial state as a
    restart file (restart = 1) and, for several ranks, the partition file."""
    from sddc_amd import mesh as M
    p = case.params
    p.restart = 1
    p.nwrite = 10 ** 8                 # no restart files during the run
    M.write_quad_inp(os.path.join(d, "QUAD.INP"), p)
    M.write_mesh_files(d, case.mesh)
    M.write_restart(os.path.join(d, "ic.dat"), Q0, ctime=0.0, it=0)
    os.makedirs(os.path.join(d, "TEC"), exist_ok=True)
    nc = case.ncell
    part = (np.arange(nc, dtype=np.int64) * ranks) // nc
    with open(os.path.join(d, "mesh.mts"), "w") as fh:
        fh.write("\n".join(str(int(r)) for r in part) + "\n")


def _set_maxiter(d: str, k: int) -> None:
    path = os.path.join(d, "QUAD.INP")
    lines = open(path).read().split("\n")
    lines[17] = str(k)                 # MAXITER (readinput.f90:25)
    open(path, "w").write("\n".join(lines))


def run_once(d: str, ranks: int, steps: int, timeout: float) -> float:
    """Run the program for `steps` steps; returns its own elapsed seconds."""
    _set_maxiter(d, steps)
    for fn in os.listdir(d):
        if fn.startswith("restart.") and fn.endswith(".dat"):
            os.remove(os.path.join(d, fn))
    cmd = [EXE] if ranks == 1 else [MPIRUN, "-np", str(ranks), EXE]
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run(cmd, cwd=d, capture_output=True, text=True, timeout=timeout, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"reference program failed ({r.returncode}):\n{r.stdout[-2000:]}\n"
                           f"{r.stderr[-2000:]}")
    m = re.search(r"Elapsed time\s+([0-9.Ee+-]+)", r.stdout)
    if not m:
        raise RuntimeError("no elapsed time in the reference program's output")
    return float(m.group(1))


def time_reference(case, Q0: np.ndarray, ranks: int, steps: int,
                   timeout: float = 1800.0, workdir: str | None = None) -> dict:
    """Per-step time of the reference program on (case, Q0) with `ranks`
    MPI ranks, from runs of `steps` and 2*`steps` steps."""
    if not available():
        raise RuntimeError(f"{EXE} is not built (make -f oracle/ref.mk)")
    d = tempfile.mkdtemp(prefix="sd2dref_", dir=workdir)
    try:
        t_w = time.perf_counter()
        write_deck(d, case, Q0, ranks)
        t_w = time.perf_counter() - t_w
        t1 = run_once(d, ranks, steps, timeout)
        t2 = run_once(d, ranks, 2 * steps, timeout)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    per_step = (t2 - t1) / steps
    ndof = case.ncell * case.N * case.N
    stages = 5                          # RK_S5P3 and FiveStageRk alike
    return {"ranks": ranks, "steps": [steps, 2 * steps], "elapsed_s": [t1, t2],
            "s_per_step": per_step, "dof": ndof,
            "dof_stage_per_s": ndof * stages / per_step if per_step > 0 else float("nan"),
            "deck_write_s": t_w}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--case", choices=("c1", "c2"), default="c1")
    ap.add_argument("--ranks", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    a = ap.parse_args(argv)
    from sddc_amd import cases
    case, Q0 = cases.c1_cylinder() if a.case == "c1" else cases.c2_cylinder()
    r = time_reference(case, Q0, a.ranks, a.steps)
    r["case"] = a.case
    print(json.dumps(r))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())

#!/usr/bin/env python3
"""Generate the golden fixtures of tests/golden/ from the reference solver.

Runs only where the reference sources are (it builds them with
oracle/ref.mk into oracle/_ref/ first).  Every fixture stores how its input
is made (mesh generator, run parameters, initial-state recipe and seed; the
tests rebuild the same case with sddc_amd.mesh) and what the reference
produced:

  setup_*    the reference's set-up arrays after INIT_SETUP (IF2C, face
             lists, Lmat, Mmat, Jac, S1, S2, Xs, Xf, Xs_low)
  stage_*    per residual evaluation: resid, nablaQs, maxima before/after,
             c_h, SM_IND, muAV, epsilon0 (from ref_driver.f90 calling the
             reference's compresid stage by stage)
  final_*    state and maxima after whole RK_S5P3 / FiveStageRk steps
  prog_*     the unmodified program's resid.out rows and restart state

    python tests/golden/make_golden.py [name ...]     # default: all fixtures
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

import refrun  # noqa: E402
from specs import SPECS, build  # noqa: E402

STRIDE = 13   # subsampling of the large arrays of the C1 fixture


def shrink(key: str, v: np.ndarray, big: bool, out: dict) -> None:
    """Store v under key; for large arrays of the C1 case store every
    STRIDE-th element and the per-variable 2- and inf-norms instead."""
    if not big or v.size < 4096:
        out[key] = v
        return
    out[key + "__sub"] = v[::STRIDE]
    w = v.reshape(9, -1, order="F") if v.size % 9 == 0 else v.reshape(1, -1)
    out[key + "__l2"] = np.sqrt((w * w).sum(axis=1))
    out[key + "__linf"] = np.abs(w).max(axis=1)


def main(names: list[str]) -> None:
    refrun.build_reference()
    for name, (mspec, pspec, ispec, nstaged, nrk) in SPECS.items():
        if names and name not in names:
            continue
        case, Q0 = build(name)
        res = refrun.run_driver(case, Q0, nstaged=nstaged, nrk=nrk)
        out = {"meta": json.dumps({"mesh": mspec, "params": pspec, "state": ispec,
                                   "nstaged": nstaged, "nrk": nrk})}
        big = case.ncell > 200
        for k, v in res["setup"].items():
            if big and k in ("Jac", "S1", "S2", "XXsolu", "Q0"):
                continue          # rebuilt and compared on the small cases
            out["setup_" + k] = v
        for i, st in enumerate(res["stages"]):
            for k, v in st.items():
                if k == "Qin" and i > 0:
                    continue
                shrink(f"stage{i:02d}_{k}", v, big, out)
        for i, rp in enumerate(res["replay"]):
            out[f"replay{i:02d}_bitequal"] = rp["bitequal"]
        for k, v in res["final"].items():
            shrink("final_" + k, v, big, out)
        if name == "c1_cylinder":
            prog = refrun.run_program(case, Q0, maxiter=11)
            out["prog_resid_out"] = prog["resid_out"]
            q1, t1, it1 = prog["restart.00000001.dat"]
            q11, t11, it11 = prog["restart.00000011.dat"]
            shrink("prog_q_iter1", q1.ravel(order="F"), True, out)
            shrink("prog_q_iter11", q11.ravel(order="F"), True, out)
            out["prog_ctime"] = np.array([t1, t11])
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{name}: {os.path.getsize(path) / 1e3:.0f} kB, replay bit-equal "
              f"{[int(r['bitequal'][0]) for r in res['replay']]}")


if __name__ == "__main__":
    main(sys.argv[1:])

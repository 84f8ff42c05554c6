#!/usr/bin/env python3
"""Capture the reference program's Tecplot output for the C1 case into
tests/golden/c1_tecplot.npz: the unmodified program (oracle/_ref/sd2d_MHD,
built by oracle/ref.mk) runs one step of the C1 deck with NRE = 1 and with
NRE = 3 and writes TEC/*.plt at iteration 1 (iterations.f90:76-81); the
fixture holds those files' bytes and names.  Runs only where the reference
sources are.

    python tests/golden/make_tecplot_golden.py
"""
from __future__ import annotations

import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

import refrun  # noqa: E402
from specs import build  # noqa: E402


def main() -> None:
    refrun.build_reference()
    out = {}
    for nre in (1, 3):
        case, Q0 = build("c1_cylinder")
        case.params.nre = nre
        d = tempfile.mkdtemp(prefix="sd2dtec_")
        try:
            refrun.run_program(case, Q0, maxiter=1, keep=d)
            files = sorted(os.listdir(os.path.join(d, "TEC")))
            assert len(files) == 1, files
            out[f"nre{nre}_name"] = np.array(files[0])
            out[f"nre{nre}_plt"] = np.frombuffer(
                open(os.path.join(d, "TEC", files[0]), "rb").read(), dtype=np.uint8)
            print(f"NRE {nre}: {files[0]}, {out[f'nre{nre}_plt'].size} bytes")
        finally:
            shutil.rmtree(d, ignore_errors=True)
    np.savez_compressed(os.path.join(HERE, "c1_tecplot.npz"), **out)


if __name__ == "__main__":
    main()

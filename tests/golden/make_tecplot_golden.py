#!/usr/bin/env python3
"""Capture the reference program's Tecplot output: the unmodified program
(oracle/_ref/sd2d_MHD, built by oracle/ref.mk) runs one step of a deck and
writes TEC/*.plt at iteration 1 (iterations.f90:76-81); each fixture holds
those files' bytes and names.
  tests/golden/c1_tecplot.npz              the C1 deck, NRE = 1 and 3
  tests/golden/ogrid_curved_tecplot.npz    the 8-node O-grid of the
                                           ogrid_curved4x8_N4 fixture, NRE = 3
Runs only where the reference sources are.

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


FIXTURES = (("c1_tecplot.npz", "c1_cylinder", (1, 3)),
            ("ogrid_curved_tecplot.npz", "ogrid_curved4x8_N4", (3,)))


def capture(spec: str, nres) -> dict:
    out = {}
    for nre in nres:
        case, Q0 = build(spec)
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
    return out


def main() -> None:
    refrun.build_reference()
    for fn, spec, nres in FIXTURES:
        np.savez_compressed(os.path.join(HERE, fn), **capture(spec, nres))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the SD2D-MHD residual + RK path on one MI355X.

Workload: the 512 x 1024-quad cylinder O-grid, r in [0.5, 30], p=4 (N=5,
25 solution points per cell, 13.1M solution points x 9 variables), FP64,
RK_S5P3 (5 residual evaluations per step), AV sensor on, residual/divB
norms every 10 steps as the reference's time loop does.  One step is one
RK_S5P3 step over the whole mesh with the state resident in HBM.

Prints one JSON line:
  value       DOF-stage updates/s = NDOF * 5 * steps / wall time
  roofline    the fused stage kernel against HBM bandwidth: algorithmic
              bytes per launch (NDOF * B_alg) / average launch time from
              HIP events on the library's stream
  cpu_baseline  the reference Fortran program itself (oracle/_ref/sd2d_MHD_O3,
              flang -O3, MPI ranks on this host's cores) on a sample of
              the same workload: (T(2K) - T(K)) / K per step from two runs
              (kind "reference"); cpu_baseline_port: the C restatement of
              the reference (OpenMP) on a few steps of the same workload
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
METRIC = "DOF-stage updates/s, FP64 plain vs mixed at stated L2 error; % HBM peak"


def build_workload(name: str):
    from sddc_amd import cases
    if name == "c2":
        return cases.c2_cylinder(), "cylinder O-grid 512x1024 quads, r 0.5-30, p=4 (N=5), FP64, RK_S5P3"
    if name == "c1":
        return cases.c1_cylinder(), "cylinder O-grid 32x64 quads, r 0.5-15, p=3 (N=4), FP64, RK_S5P3"
    if name == "c3":
        return cases.orszag_tang(1024, N=4), ("periodic Orszag-Tang vortex 1024x1024 quads on "
                                              "[0,2pi]^2, p=3 (N=4), RK_S5P3")
    raise SystemExit(f"unknown workload {name}")


def mixed_to_time(case, Q0, t_end: float, desc: str) -> dict:
    """FP64 and mixed runs of the same workload from t=0 to t_end: the time
    of each and the per-variable relative L2 difference at t_end (w, Bz and
    psi, which start at zero, against the largest variable's norm)."""
    import numpy as np
    from sddc_amd import abi, cases
    from sddc_amd.solver import Sd2d
    nsteps = int(np.ceil(t_end / case.params.dt - 1e-9))
    ndof = cases.dof(case)
    out, times = {}, {}
    for prec in (abi.SD2D_FP64, abi.SD2D_MIXED):
        g = Sd2d(case, precision=prec)
        g.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
        g.step(1, 0)
        t0 = time.perf_counter()
        g.step(nsteps - 1, 0)
        times[prec] = time.perf_counter() - t0
        out[prec] = g.get_state()[0]
        ctime = g.diag()["ctime"]
        g.close()
        print(f"[bench] precision {prec}: {nsteps} steps to t={ctime:.4f} in "
              f"{times[prec]:.1f}s", file=sys.stderr, flush=True)
    a = out[abi.SD2D_MIXED].reshape(9, -1, order="F")
    b = out[abi.SD2D_FP64].reshape(9, -1, order="F")
    nb = np.sqrt((b * b).sum(axis=1))
    den = nb.copy()
    den[[3, 7, 8]] = nb.max()
    l2 = np.sqrt(((a - b) ** 2).sum(axis=1)) / den
    v64 = ndof * 5 * (nsteps - 1) / times[abi.SD2D_FP64]
    vmx = ndof * 5 * (nsteps - 1) / times[abi.SD2D_MIXED]
    return {"metric": METRIC, "workload": desc, "t_end": t_end, "steps": nsteps,
            "dt": case.params.dt, "dof": ndof,
            "fp64_value": v64, "mixed_value": vmx, "unit": "DOF-stage/s",
            "speedup": vmx / v64, "l2_rel_mixed_vs_fp64": [float(x) for x in l2],
            "all_finite": bool(np.all(np.isfinite(a)))}


def rounding_sensitivity(case, Q0, checkpoints, desc: str) -> dict:
    """How far rounding-level changes carry on this workload: four runs
    stepped side by side,
      fp64            FP64 from Q0,
      fp64_q0_f32     FP64 from Q0 rounded once to float32,
      fp64_q0_1e-15   FP64 from Q0 * (1 + 1e-15 * seeded U[-1, 1]),
      mixed           the mixed form from Q0,
    and at every checkpoint step the per-variable L2 difference of the last
    three from the first, against each variable's own norm and against the
    norm of rho*u (the momentum scale of the flow)."""
    import numpy as np
    from sddc_amd import abi
    from sddc_amd.solver import Sd2d
    rng = np.random.default_rng(7)
    starts = {"fp64": (Q0, abi.SD2D_FP64),
              "fp64_q0_f32": (Q0.astype(np.float32).astype(np.float64), abi.SD2D_FP64),
              "fp64_q0_1e-15": (Q0 * (1.0 + 1e-15 * rng.uniform(-1, 1, Q0.shape)), abi.SD2D_FP64),
              "mixed": (Q0, abi.SD2D_MIXED)}
    runs = {}
    for name, (q, prec) in starts.items():
        g = Sd2d(case, precision=prec)
        g.set_state(q, [0.0, 0.0, 0.0, 0.0])
        runs[name] = g
    rows, done = [], 0
    try:
        for k in checkpoints:
            for g in runs.values():
                g.step(k - done, 0)
            done = k
            st = {n: g.get_state()[0].reshape(9, -1, order="F") for n, g in runs.items()}
            a = st["fp64"]
            na = np.sqrt((a * a).sum(axis=1))
            row = {"step": k, "norm": [float(x) for x in na]}
            for n in ("fp64_q0_f32", "fp64_q0_1e-15", "mixed"):
                nd = np.sqrt(((st[n] - a) ** 2).sum(axis=1))
                row[n + "_vs_own"] = [float(x) for x in nd / np.where(na > 0, na, 1.0)]
                row[n + "_vs_rhou"] = [float(x) for x in nd / na[1]]
            rows.append(row)
            print(f"[bench] sensitivity step {k}: rho*v vs own norm  f32 start "
                  f"{row['fp64_q0_f32_vs_own'][2]:.2e}  1e-15 start "
                  f"{row['fp64_q0_1e-15_vs_own'][2]:.2e}  mixed {row['mixed_vs_own'][2]:.2e}",
                  file=sys.stderr, flush=True)
    finally:
        for g in runs.values():
            g.close()
    return {"workload": desc, "dt": case.params.dt, "checkpoints": rows,
            "definition": "||q_run - q_fp64||_2 per variable over all points, divided by "
                          "the FP64 run's own norm of that variable (vs_own) or by its "
                          "norm of rho*u (vs_rhou)"}


def run_c4(steps: int, warmup: int, tune_steps: int = 2) -> dict:
    """The C4 ensemble (64 cylinder runs of uneven size and order) stepped
    as one sd2d_ens ensemble in mixed precision after the tile-shape
    autotuner, against the same ensemble in FP64 and against the 64 runs
    stepped one after another through sd2d_step (mixed)."""
    import numpy as np
    from sddc_amd import abi, cases
    from sddc_amd.solver import Sd2d, Sd2dEnsemble
    t = time.perf_counter()
    mem = cases.c4_ensemble()
    ndof = sum(cases.dof(c) for c, _, _ in mem)
    print(f"[bench] c4: {len(mem)} members, {ndof / 1e6:.1f} M DOF, built in "
          f"{time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    res, states = {}, {}
    for prec in (abi.SD2D_MIXED, abi.SD2D_FP64):
        ens = Sd2dEnsemble([c for c, _, _ in mem], precision=prec)
        for i, (c, Q0, _) in enumerate(mem):
            ens.set_state(i, Q0, [0.0, 0.0, 0.0, 0.0])
        tt = time.perf_counter()
        shapes = ens.tune(tune_steps)
        tt = time.perf_counter() - tt
        ens.step(warmup)
        t0 = time.perf_counter()
        ens.step(steps)
        wall = time.perf_counter() - t0
        states[prec] = [ens.get_state(i)[0] for i in range(len(mem))]
        ens.close()
        fin = [bool(np.all(np.isfinite(q))) for q in states[prec]]
        dof_fin = sum(cases.dof(c) for (c, _, _), f in zip(mem, fin) if f)
        res[prec] = {"value": ndof * 5 * steps / wall, "wall_s": wall,
                     "finite_members": int(sum(fin)),
                     "value_finite_members": dof_fin * 5 * steps / wall,
                     "tile_shapes": {str(n): {abi.SD2D_TILE_STRIPS: "strips",
                                              abi.SD2D_TILE_BLOCKS: "blocks",
                                              abi.SD2D_TILE_ROWPAIRS: "rowpairs"}.get(v, str(v))
                                     for n, v in shapes.items()},
                     "tune_s": tt}
        print(f"[bench] c4 precision {prec}: {steps} steps in {wall:.2f}s, shapes {shapes}",
              file=sys.stderr, flush=True)
    # the same runs one after another (mixed), same number of steps
    seq = 0.0
    for c, Q0, _ in mem:
        g = Sd2d(c, precision=abi.SD2D_MIXED)
        g.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
        g.step(warmup)
        t0 = time.perf_counter()
        g.step(steps)
        seq += time.perf_counter() - t0
        g.close()
    print(f"[bench] c4 one by one: {seq:.2f}s", file=sys.stderr, flush=True)
    # mixed vs FP64 after warmup + steps, per variable, worst member
    l2 = np.zeros(9)
    nfin = 0
    for a, b in zip(states[abi.SD2D_MIXED], states[abi.SD2D_FP64]):
        if not (np.all(np.isfinite(a)) and np.all(np.isfinite(b))):
            continue
        nfin += 1
        a = a.reshape(9, -1, order="F"); b = b.reshape(9, -1, order="F")
        nb = np.sqrt((b * b).sum(axis=1))
        den = nb.copy()
        den[[3, 7, 8]] = nb.max()
        den[den == 0] = nb.max()
        l2 = np.maximum(l2, np.sqrt(((a - b) ** 2).sum(axis=1)) / den)
    vm, v64 = res[abi.SD2D_MIXED]["value"], res[abi.SD2D_FP64]["value"]
    return {"metric": METRIC, "workload": "C4 ensemble: 64 cylinder runs, Nr x Ntheta in "
            "{32x64..256x512}, p in {2,3,4}, Mach U[1.2,3], Alfven Mach U[2,8], field angle "
            "U[0,90] deg (seed 4), RK_S5P3 at CFL 0.15",
            "members": len(mem), "dof": ndof, "steps": steps, "warmup": warmup,
            "mixed_value": vm, "fp64_value": v64, "unit": "DOF-stage/s",
            "speedup_mixed_vs_fp64": vm / v64,
            "one_by_one_mixed_value": ndof * 5 * steps / seq,
            "speedup_ensemble_vs_one_by_one": (ndof * 5 * steps / res[abi.SD2D_MIXED]["wall_s"])
            / (ndof * 5 * steps / seq),
            "tile_shapes_mixed": res[abi.SD2D_MIXED]["tile_shapes"],
            "tile_shapes_fp64": res[abi.SD2D_FP64]["tile_shapes"],
            "finite_members_fp64": res[abi.SD2D_FP64]["finite_members"],
            "finite_members_mixed": res[abi.SD2D_MIXED]["finite_members"],
            "mixed_value_finite_members": res[abi.SD2D_MIXED]["value_finite_members"],
            "fp64_value_finite_members": res[abi.SD2D_FP64]["value_finite_members"],
            "finite_in_both": nfin,
            "l2_rel_mixed_vs_fp64_worst_member": [float(x) for x in l2]}


def run_c5(nsteps: int = 200, ckpt: int = 20) -> dict:
    """The C5 adjoint workload: 128 x 256 O-grid cylinder, p=3, FP64,
    RK_S5P3 for 200 steps with checkpoints every 20; objective the
    time-averaged wall drag.  Gradients with respect to the free-stream
    Mach number, the field angle and the AV coefficient, and the VJP of a
    seeded N(0,1) cotangent on the final state, each against central
    differences with h = 1e-6."""
    import numpy as np
    from sddc_amd.adjoint import CylinderObjective, CylinderSpec
    from sddc_amd.solver import Sd2d
    spec = CylinderSpec(128, 256, 4, mach=1.5, b_angle_deg=10.0)
    t = time.perf_counter()
    o = CylinderObjective(spec)
    print(f"[bench] c5 built in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    mach, ang = spec.mach, spec.b_angle_deg
    o.value(mach, ang, 2, ckpt)                       # warm-up (first launches)
    t0 = time.perf_counter()
    J = o.value(mach, ang, nsteps, ckpt)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    J2, gr = o.gradients(mach, ang, nsteps, ckpt)
    t_grad = time.perf_counter() - t0
    h = 1e-6                          # Mach, and the angle in radians
    hd = math.degrees(h)
    fdm = (o.value(mach + h, ang, nsteps, ckpt) - o.value(mach - h, ang, nsteps, ckpt)) / (2 * h)
    fda = (o.value(mach, ang + hd, nsteps, ckpt) - o.value(mach, ang - hd, nsteps, ckpt)) / (2 * h)
    print(f"[bench] c5 J={J:.10e} dJ/dM {gr['mach']:.8e} (fd {fdm:.8e}), dJ/dtheta "
          f"{gr['angle_rad']:.8e} (fd {fda:.8e})", file=sys.stderr, flush=True)
    # VJP of a seeded N(0,1) cotangent on the final state, checked along
    # seeded N(0,1) directions of the initial state: one supported on the 32
    # innermost rings (r < 1.17) and one on the whole mesh
    c, Q0 = o.case(mach, ang)
    rng = np.random.default_rng(20)
    Cb = np.asfortranarray(rng.standard_normal(Q0.shape))
    V = rng.standard_normal(Q0.shape) * np.abs(Q0).max(axis=(1, 2, 3))[:, None, None, None]
    Vnear = V.copy()
    Vnear[..., 32 * spec.ntheta:] = 0.0      # cells are numbered ring by ring
    g = Sd2d(c)
    g.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
    g.fwd(nsteps, ckpt)
    t0 = time.perf_counter()
    q0b, _, _ = g.vjp(Cb, 0.0)
    t_vjp = time.perf_counter() - t0

    def final(Q):
        g.set_state(Q, [0.0, 0.0, 0.0, 0.0])
        g.fwd(nsteps, ckpt)
        return float(np.sum(Cb * g.get_state()[0]))

    f0 = final(Q0)
    fp, fm = final(Q0 + h * Vnear), final(Q0 - h * Vnear)
    fdv = (fp - fm) / (2 * h)
    adv = float(np.sum(q0b * Vnear))
    gp, gmn = final(Q0 + h * V), final(Q0 - h * V)
    g.close()
    rel = lambda a, b: abs(a - b) / abs(b) if b != 0 else abs(a)
    return {"metric": METRIC, "workload": "C5 adjoint: cylinder O-grid 128x256, r 0.5-15, p=3 "
            "(N=4), Mach 1.5, |B| 0.3 at 10 deg, FP64, RK_S5P3, 200 steps, checkpoints every 20",
            "objective": "time-averaged wall drag", "J": J, "steps": nsteps, "ckpt": ckpt,
            "dJ_dMach": {"adjoint": gr["mach"], "fd": fdm, "rel": rel(gr["mach"], fdm)},
            "dJ_dtheta_rad": {"adjoint": gr["angle_rad"], "fd": fda,
                              "rel": rel(gr["angle_rad"], fda)},
            "dJ_dAV": {"adjoint": 0.0, "fd": 0.0, "note": "the AV sensor does not act on the state"},
            "state_vjp_dot": {"adjoint": adv, "fd": fdv, "rel": rel(adv, fdv),
                              "cotangent": "N(0,1) seed 20 on the final state",
                              "direction": "N(0,1) (x variable scale) on the 32 innermost rings",
                              "fd_forward": (fp - f0) / h, "fd_backward": (f0 - fm) / h},
            "state_vjp_dot_whole_mesh": {
                "adjoint": float(np.sum(q0b * V)), "fd": (gp - gmn) / (2 * h),
                "fd_forward": (gp - f0) / h, "fd_backward": (f0 - gmn) / h,
                "note": "J is not differentiable along this direction: the running maxima "
                        "behind c_h are attained with exact ties across the uniform free "
                        "stream (|u| = 1.5, equal face wave speeds), which a direction that "
                        "perturbs each point differently splits; the one-sided differences "
                        "disagree"},
            "fd_step": h, "forward_s": t_fwd, "gradient_s": t_grad, "vjp_s": t_vjp,
            "vjp_over_forward": t_vjp / t_fwd}


def traffic_from_profiles(kernel_tag: str):
    """HBM bytes per launch of the stage kernel measured with rocprofv3
    --pmc (FETCH_SIZE, WRITE_SIZE), scaled by the copy-kernel calibration,
    as committed under profiles/; None when not measured."""
    p = os.path.join(HERE, "profiles", "pmc_traffic.json")
    if not os.path.exists(p):
        return None
    try:
        d = json.load(open(p))
        return d.get(kernel_tag, {}).get("bytes_per_launch")
    except Exception:
        return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"],
                    help="c4: the 64-member ensemble; c5: the adjoint; each reported on a "
                         "line of its own")
    ap.add_argument("--t-end", type=float, default=None,
                    help="instead of the throughput run: FP64 and mixed runs to this "
                         "time, reporting the mixed form's L2 difference")
    ap.add_argument("--sensitivity", default=None, metavar="STEPS",
                    help="instead of the throughput run: FP64 from the workload's state, "
                         "from it rounded to float32 and from it perturbed by 1e-15, and "
                         "the mixed form, compared at these comma-separated steps")
    ap.add_argument("--cpu-steps", type=int, default=3,
                    help="steps of the CPU baseline sample (0 = skip); 3 C2 steps are "
                         "about 10 s of the oracle on 16 host threads")
    ap.add_argument("--cpu-ref-steps", type=int, default=2,
                    help="K of the reference program's (T(2K) - T(K)) / K timing (0 = skip)")
    ap.add_argument("--cpu-ref-ranks", type=int, default=0,
                    help="MPI ranks of the reference program (0: this host's CPU share, "
                         "OMP_NUM_THREADS or the core count, at most 16)")
    ap.add_argument("--tile-shape", default="rowpairs", choices=["strips", "blocks", "rowpairs"],
                    help="cells of one workgroup: two mesh rows (default; strips where the "
                         "mesh is not numbered row by row), mesh-order strips, or compact blocks")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.stderr.write("bench.py measures one GPU only: this path runs the whole mesh on "
                         "one MI355X (multi-GPU sharding is a design, not built); "
                         "run with --gpus 1\n")
        return 2

    from sddc_amd import abi, cases
    from sddc_amd.solver import Sd2d

    if args.workload == "c4":
        print(json.dumps(run_c4(args.steps, args.warmup)), flush=True)
        return 0
    if args.workload == "c5":
        print(json.dumps(run_c5()), flush=True)
        return 0
    t = time.perf_counter()
    (case, Q0), desc = build_workload(args.workload)
    print(f"[bench] workload {args.workload}: {case.ncell} cells, N={case.N}, "
          f"built in {time.perf_counter() - t:.1f}s", file=sys.stderr, flush=True)
    if args.t_end is not None:
        print(json.dumps(mixed_to_time(case, Q0, args.t_end, desc)), flush=True)
        return 0
    if args.sensitivity is not None:
        ck = sorted({int(x) for x in args.sensitivity.split(",") if x.strip()})
        if not ck or ck[0] < 1:
            sys.stderr.write("--sensitivity takes positive step numbers\n")
            return 2
        print(json.dumps(rounding_sensitivity(case, Q0, ck, desc)), flush=True)
        return 0
    N = case.N
    ndof = cases.dof(case)
    shape = {"strips": abi.SD2D_TILE_STRIPS, "blocks": abi.SD2D_TILE_BLOCKS,
             "rowpairs": abi.SD2D_TILE_ROWPAIRS}[args.tile_shape]
    g = Sd2d(case, tile_shape=shape)
    g.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
    g.step(args.warmup, 10)
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    g.step(args.steps, 10)                 # returns after the device work is done
    t1 = time.perf_counter()
    wall = t1 - t0
    value = ndof * 5 * args.steps / wall
    print(f"[bench] timed {args.steps} steps in {wall:.4f}s", file=sys.stderr, flush=True)

    # stage-kernel time from HIP events on the library's stream
    g.set_profiling(True)
    g.step(args.steps, 10)
    ms, nlaunch = g.get_profile()
    g.set_profiling(False)
    avg_s = ms / 1e3 / max(nlaunch, 1)
    balg = cases.alg_bytes_per_dof_stage(N, case.params.k_stage)
    achieved = ndof * balg / avg_s / 1e9
    d = g.diag()
    nsteps_done = args.warmup + 2 * args.steps
    q64 = g.get_state()[0] if not args.no_mixed else None

    res = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: free stream (Mach 1.5, Alfven Mach 4 at 30 degrees) with a seeded "
                "1e-3 uniform density perturbation, started at rest on the wall with the "
                "velocity rising to the free stream over one radius (an impulsive start turns "
                "NaN within three steps, in the reference too)",
        "config": {"workload": desc, "ncell": case.ncell, "N": N, "dof": ndof,
                   "dt": case.params.dt, "rk": "RK_S5P3", "av": True,
                   "norm_every": 10, "tile_shape": args.tile_shape},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_from_profiles("stage_kernel_N5_rk"),
                     "bytes_per_dof_stage": balg, "stage_kernel_ms": avg_s * 1e3,
                     "launches": nlaunch},
        "diag": {"resid_rms": d["resid_rms"], "divb_rms": d["divb_rms"],
                 "eigvmax_g": d["eigvmax_g"]},
    }
    g.close()

    if not args.no_mixed:
        # the same workload with SD2D_MIXED: throughput, stage-kernel time and
        # the relative L2 difference from the FP64 state after the same steps
        import numpy as np
        from sddc_amd import abi
        gm = Sd2d(case, precision=abi.SD2D_MIXED, tile_shape=shape)
        gm.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
        gm.step(args.warmup, 10)
        tm0 = time.perf_counter()
        gm.step(args.steps, 10)
        wall_m = time.perf_counter() - tm0
        gm.set_profiling(True)
        gm.step(args.steps, 10)
        ms_m, n_m = gm.get_profile()
        gm.set_profiling(False)
        qm = gm.get_state()[0]
        gm.close()
        a = qm.reshape(9, -1, order="F"); b = q64.reshape(9, -1, order="F")
        nb = np.sqrt((b * b).sum(axis=1))
        den = nb.copy()
        den[[3, 7, 8]] = nb.max()      # w, Bz, psi: against the largest variable
        l2 = np.sqrt(((a - b) ** 2).sum(axis=1)) / den
        avg_m = ms_m / 1e3 / max(n_m, 1)
        balg_m = cases.alg_bytes_per_dof_stage(N, case.params.k_stage, metric_bytes=4)
        vm = ndof * 5 * args.steps / wall_m
        res["mixed"] = {
            "value": vm, "unit": "DOF-stage/s", "ms_per_step": wall_m / args.steps * 1e3,
            "speedup_vs_fp64": vm / value, "stage_kernel_ms": avg_m * 1e3,
            "bytes_per_dof_stage": balg_m,
            "roofline_frac": ndof * balg_m / avg_m / 1e9 / HBM_PEAK_GBS,
            "l2_rel_vs_fp64": [float(x) for x in l2], "steps_compared": nsteps_done,
            "l2_definition": "per variable ||q_mixed - q_fp64||_2 / ||q_fp64||_2; for w, Bz "
                             "and psi the denominator is the largest variable's norm",
            "note": "interpolation, fluxes, Rusanov and derivatives in FP32; state, "
                    "pressure recovery, sources, RK combination, norms and AV sensor in FP64"}
        print(f"[bench] mixed: {vm / 1e9:.2f} G DOF-stage/s, x{vm / value:.2f}, "
              f"max L2 rel {l2.max():.1e}", file=sys.stderr, flush=True)

    if args.cpu_steps > 0:
        from oracle.oracle import Oracle, threads
        o = Oracle(case)
        o.set_state(Q0, [0.0, 0.0, 0.0, 0.0])
        tc = time.perf_counter()
        o.step(args.cpu_steps)
        tcpu = time.perf_counter() - tc
        res["cpu_baseline"] = {"value": ndof * 5 * args.cpu_steps / tcpu,
                               "unit": "DOF-stage/s", "cores": threads(), "kind": "port",
                               "sample": f"{args.cpu_steps} RK_S5P3 step(s) of the same "
                                         f"workload ({5 * args.cpu_steps} residual "
                                         f"evaluations), {tcpu:.1f}s"}
        o.close()
    if args.cpu_ref_steps > 0:
        from oracle import time_reference as TR
        if TR.available():
            ranks = args.cpu_ref_ranks or min(16, int(os.environ.get("OMP_NUM_THREADS", "0"))
                                              or os.cpu_count() or 1)
            print(f"[bench] reference program: {ranks} MPI ranks, {args.cpu_ref_steps} and "
                  f"{2 * args.cpu_ref_steps} steps", file=sys.stderr, flush=True)
            r = TR.time_reference(case, Q0, ranks, args.cpu_ref_steps)
            if "cpu_baseline" in res:
                res["cpu_baseline_port"] = res["cpu_baseline"]
            res["cpu_baseline"] = {
                "value": r["dof_stage_per_s"], "unit": "DOF-stage/s", "cores": ranks,
                "kind": "reference",
                "sample": f"the reference program (flang -O3, {ranks} MPI ranks, cells split "
                          f"into contiguous ring ranges) on the same workload for "
                          f"{r['steps'][0]} and {r['steps'][1]} RK_S5P3 steps: "
                          f"{r['elapsed_s'][0]:.1f}s and {r['elapsed_s'][1]:.1f}s, "
                          f"{r['s_per_step']:.2f}s per step"}
            print(f"[bench] reference program: {r['s_per_step']:.2f} s per step", file=sys.stderr,
                  flush=True)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

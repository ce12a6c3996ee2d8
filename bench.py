#!/usr/bin/env python3
"""Benchmark of the line-implicit steady Euler iteration on one MI355X.

Workload (default): O-grid around a NACA 0012 thickness-formula section,
512 x 256 cells whose outer 128 layers are split into triangles (65,536 quads
+ 131,072 triangles = 196,608 cells; radial lines run through both triangles
of each split cell, the 65,536 faces left over are orphan faces), MUSCL
(order 3) with the reference's smooth limiter, M = 0.5, alpha = 2 deg, CFL =
10, msweep = 1, synthetic freestream start.  --tri-layers 0 gives the all-quad
variant.  A "step" is one stepSolution ADI iteration: computeRHS("implicit")
+ ADI (two chain phases) + computeLinearRHS + update.  FAST contexts are
autotuned (hamstr_autotune) before the untimed warm-up, outside the clock.

Prints one JSON line: cell-iterations/s (value), the roofline of the dominant
kernel from HIP events over the timed region, the time to nonlinear
L2 <= 1e-10 from freestream (FAST, EXACT, FP32-LHS), FAST against EXACT beyond
1e-10, the explicit euler / rk3 steps of the same mesh, the 1024x384 WENO
transonic case, the mixed-batch sweep with the learned warm start and its
256-case held-out evaluation, and the CPU restatement timed on this host on a
bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def algorithmic_bytes(n_cell, n_proc, phase_faces, b, msweep=1, phase_cells=None):
    """Bytes per cell-iteration of the iteration's dataflow when every
    materialised array (q, r0, sigma, ff, r, dq, face fluxes) is read and
    written once per phase and stencil data is reused on chip.

    phi = processed faces per cell, b = bytes per LHS scalar (8 FP64, 4 FP32).
    computeRHS + Jacobians: 72 + phi*(24 + 32b); ADI phase k: 72 per cell its
    chains cover (sigma, r read, r written) + 32*b per face of the phase;
    computeLinearRHS + update: 168 + phi*(8 + 32b); extra sweeps:
    312 + phi*(8 + 64b) each.
    """
    phi = n_proc / n_cell
    stages = {"rhs": 72 + phi * (24 + 32 * b)}
    for k, nf in enumerate(phase_faces):
        cells = n_cell if phase_cells is None else phase_cells[k]
        stages[f"adi_phase{k}"] = (72 * cells + 32 * b * nf) / n_cell
    stages["linear"] = 168 + phi * (8 + 32 * b)
    total = sum(stages.values()) + (msweep - 1) * (312 + phi * (8 + 64 * b))
    return total, stages


def phase_face_counts(mesh):
    """Processed faces per ADI phase for an O-grid (loops, then lines), in
    all, and the cells each phase's chains cover; orphan faces count with the
    last phase, whose rows take their diagonal blocks."""
    import numpy as np
    start = np.asarray(mesh.chain_start)
    conn = np.asarray(mesh.chain_conn)
    per, cells = [], []
    for c in range(len(start) - 1):
        f1, f2 = start[c], start[c + 1]
        per.append((f2 - f1) - (1 if conn[f1] == conn[f2 - 1] else 0))
        cells.append((f2 - f1) - 1)   # a closed chain repeats its first face
    nl = mesh.chains_per_color[0]
    no = int(len(mesh.orphans))
    return ([int(sum(per[:nl])), int(sum(per[nl:])) + no], int(sum(per)) + no,
            [int(sum(cells[:nl])), int(sum(cells[nl:]))])


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform's name."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


JAC_NAMES = ("stored", "q", "states", "states_lines")   # HAMSTR_JAC_STORED, _FROM_Q, _FROM_STATES, _STATES_LINES


def measured_traffic(stage, nc, nr, order, lhs, adi, tri=0, jsrc=0):
    """HBM bytes per launch of the stage's kernel from the committed rocprofv3
    counter summary of this workload and Jacobian source (FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy of known size), or None for another
    workload or source."""
    tag = f"_tri{tri}" if tri else ""
    src = JAC_NAMES[jsrc]
    path = os.path.join(ROOT, "profiles", f"traffic_{lhs}_naca{nc}x{nr}{tag}_o{order}_{src}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    w = t.get("workload", {})
    if (w.get("nc"), w.get("nr"), w.get("order"), w.get("lhs"), w.get("adi"), w.get("tri_layers", 0),
            w.get("jacobian_source", "stored")) != (nc, nr, order, lhs, adi, tri, src):
        return None
    st = t["stages"].get(stage)
    return None if st is None else st["traffic_bytes"]


def counter_summary(args, fetch_dir, write_dir, trace_dir, timed=20):
    """The traffic summary measured_traffic reads, from three rocprofv3 runs of
    this command with the Jacobian source fixed (--jacobians) and
    --copy-probe: `--pmc FETCH_SIZE`, `--pmc WRITE_SIZE` and `--kernel-trace`
    (each `-f csv -o run`).  The copy of 256 MiB scales each counter to bytes.
    Per stage: the eager region's last `timed` dispatches of its kernel
    (`stages`, the region `roofline` and `stage_ms` come from), and the
    timed graph region's (`graph_stages`): the `timed` ADI dispatches before
    those and the fused kernels' last timed - 1.  Writes
    profiles/traffic_*.json and returns its path."""
    import csv
    import re

    def short(name):
        return re.split(r"\(", name.replace("(anonymous namespace)::", "").replace("void ", ""))[0].strip()

    def load(d, value):
        out = {}
        f = os.path.join(d, "run_counter_collection.csv" if value else "run_kernel_trace.csv")
        with open(f) as fh:
            for r in csv.DictReader(fh):
                v = float(r["Counter_Value"]) if value else \
                    (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
                out.setdefault(short(r["Kernel_Name"]), []).append((int(r.get("Dispatch_Id", 0)), v))
        return {k: [v for _, v in sorted(x)] for k, x in out.items()}

    fetch, write, trace = load(fetch_dir, True), load(write_dir, True), load(trace_dir, False)
    copy = "__amd_rocclr_copyBuffer"
    fk, wk = max(fetch[copy]), max(write[copy])
    fs, ws = (256 << 10) / fk, (256 << 10) / wk
    src = args.jacobians
    # the ADI kernels compute the Jacobians (JQ) for a computed source; with
    # states_lines phase 0 reads stored ones
    jq1 = "true" if src in ("states", "q", "states_lines") else "false"
    jq0 = "true" if src in ("states", "q") else "false"
    o, t = args.order, "double"  # the face kernels staged (the default on this mesh)
    eager = {"adi_phase1": f"k_adi_fast_half2<{t}, 384, {jq1}>", "adi_phase0": f"k_adi_fast<{t}, 1, 512, {jq0}>",
             "lin_faces": f"k_lin_faces<{o}, {t}, true, true, true>",
             "rhs_faces": f"k_rhs_faces<{o}, {t}, true, true, true>",
             "lin_cells": "k_lin_cells<true>", "rhs_cells": "k_rhs_cells<true>", "norms": "k_norm_slots"}
    graph = {"adi_phase1": (eager["adi_phase1"], 2 * timed, timed), "adi_phase0": (eager["adi_phase0"], 2 * timed, timed),
             "lin_rhs_faces": (f"k_lin_rhs_faces<{o}, {t}, true, true>", timed - 1, 0),
             "lin_rhs_cells": ("k_lin_rhs_cells", timed - 1, 0)}

    def stage(name, a, b):
        sl = slice(-a, -b if b else None)
        f, w, d = fetch[name][sl], write[name][sl], trace[name][sl]
        fm, wm = sum(f) / len(f), sum(w) / len(w)
        return {"kernels": [name], "avg_us": sum(d) / len(d), "traffic_bytes": (fm * fs + wm * ws) * 1024,
                "fetch_kib": fm, "write_kib": wm}

    res = {"workload": {"nc": args.nc, "nr": args.nr, "tri_layers": args.tri_layers, "order": args.order,
                        "lhs": "fp64", "adi": args.adi, "jacobian_source": src},
           "timed_iterations": timed, "fetch_scale": fs, "write_scale": ws,
           "calibration": {"copy_bytes": 256 << 20, "fetch_kib": fk, "write_kib": wk},
           "stages": {k: stage(n, timed, 0) for k, n in eager.items() if n in fetch},
           "graph_stages": {k: stage(n, a, b) for k, (n, a, b) in graph.items() if n in fetch}}
    tag = f"_tri{args.tri_layers}" if args.tri_layers else ""
    path = os.path.join(ROOT, "profiles", f"traffic_fp64_naca{args.nc}x{args.nr}{tag}_o{args.order}_{src}.json")
    with open(path, "w") as fh:
        json.dump(res, fh, indent=1)
    return path


def cpu_baseline(mesh, mach, alpha, order, cfl, budget_s=12.0, max_iters=30):
    """The reference's own code (oracle/_ref, compiled from its sources) on a
    bounded sample: `n` iterations of the same case from freestream, 1 core.
    The oracle (a leak-free C++ restatement) instead when the reference build
    is absent, or when the mesh has orphan faces: the reference computes chain
    faces only, so a mesh with triangles is outside what it can run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    from hamstr_oop_amd import meshgen
    if oracle_py.Reference.available("O2") and not len(mesh.orphans):
        ref = oracle_py.Reference("O2")
        with tempfile.TemporaryDirectory() as d:
            meshgen.write_reference_case(mesh, d, mach, alpha, order, 1, cfl)
            rc = ref.load(d)
            _, _, t1 = rc.step(1)
            n = int(max(1, min(max_iters, budget_s / max(t1, 1e-6))))
            _, _, sec = rc.step(n)
        kind = "reference"
        what = "reference SOLVER (computeRHS+ADI) compiled -O2 from its sources"
    else:
        orc = oracle_py.Oracle()
        oc = orc.case(mesh, mach, alpha, order, cfl)
        _, _, _, t1 = oc.step(1)
        n = int(max(1, min(max_iters, budget_s / max(t1, 1e-6))))
        _, _, _, sec = oc.step(n)
        kind = "port"
        what = "oracle C++ restatement, -O2"
    return {
        "value": mesh.n_cell * n / sec, "unit": "cell-iters/s", "cores": 1, "kind": kind,
        "sample": f"{n} iterations (after 1 warm-up) of the same case from freestream; {what}; "
                  f"{sec:.2f} s wall on {cpu_model()}, nproc={os.cpu_count()}",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--nc", type=int, default=512)
    ap.add_argument("--nr", type=int, default=256)
    ap.add_argument("--tri-layers", type=int, default=128, help="outer layers split into triangles (0: all quads)")
    ap.add_argument("--order", type=int, default=3)
    ap.add_argument("--mach", type=float, default=0.5)
    ap.add_argument("--alpha", type=float, default=2.0)
    ap.add_argument("--cfl", type=float, default=10.0)
    ap.add_argument("--adi", choices=["exact", "fast"], default="fast")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--max-iters", type=int, default=20000)
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes (FAST)")
    ap.add_argument("--jacobians", choices=["auto", "stored", "states", "q", "states_lines"], default="auto",
                    help="where the FAST ADI kernels take the chain faces' Jacobians (auto: the autotuner's "
                         "choice; a fixed source keeps profiled runs on the same kernels)")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-exact", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-explicit", action="store_true", help="skip the euler / rk3 stepSolution leg")
    ap.add_argument("--no-weno", action="store_true", help="skip the 1024x384 WENO FP32-LHS case")
    ap.add_argument("--sweep-cases", type=int, default=256)
    ap.add_argument("--train-cases", type=int, default=384)
    ap.add_argument("--train-steps", type=int, default=20000)
    ap.add_argument("--stall-window", type=int, default=2000)
    ap.add_argument("--heldout", type=int, default=256,
                    help="held-out 256x128 cases for the warm start's own evaluation (0: skip)")
    ap.add_argument("--no-deep", action="store_true",
                    help="skip the FAST vs EXACT comparison beyond L2 1e-10")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one plain 256 MiB device-to-device copy (counter calibration under rocprofv3)")
    ap.add_argument("--counter-summary", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "TRACE_DIR"),
                    help="write profiles/traffic_*.json for this workload and --jacobians source from the csv "
                         "output of three rocprofv3 runs of the same command (no GPU needed)")
    ap.add_argument("--weno-profile", type=int, default=0, metavar="N",
                    help="only the WENO transonic case: N iterations per left-hand side, with stage times "
                         "(a short command for rocprofv3)")
    args = ap.parse_args()

    if args.counter_summary:
        print(counter_summary(args, *args.counter_summary), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU only; multi-GPU runs are not part of it "
                                   "(cases would shard whole meshes per GPU)", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.weno_profile > 0:
        print(json.dumps(weno_profile(args, args.weno_profile)), flush=True)
        return

    from hamstr_oop_amd import hip, meshgen
    mesh = meshgen.naca_ogrid(args.nc, args.nr, t=0.12, tri_layers=args.tri_layers)
    phase_faces, n_proc, phase_cells = phase_face_counts(mesh)
    adi_mode = hip.ADI_FAST if args.adi == "fast" else hip.ADI_EXACT

    def run(lhs, mode=None, steps=None):
        """Timed region 1: `steps` iterations replaying captured graphs (the
        product path), no events.  Timed region 2: the same number of
        iterations launched eagerly with a HIP event between kernels on the
        context's stream, for per-kernel times."""
        mode = adi_mode if mode is None else mode
        steps = args.steps if steps is None else steps
        s = hip.Solver(mesh, args.mach, args.alpha, args.order, args.cfl, lhs=lhs, adi_mode=mode)
        tune_s = tune(s)
        # the timed region's graph is captured and uploaded first; the warm-up
        # steps then run right before the clock starts (a GPU left idle while
        # the host captures starts the timed replay at lower clocks)
        s.prepare_graph(steps)
        s.step(args.warmup, check_every=steps)
        s.sync()
        t0 = time.perf_counter()
        n, l2r, _ = s.step(steps, check_every=steps)
        s.sync()
        dt = time.perf_counter() - t0
        s.set_timing(True)
        s.sync()
        t1 = time.perf_counter()
        s.step(steps, check_every=steps)
        s.sync()
        dt_ev = time.perf_counter() - t1
        stages = s.last_timings()
        jsrc = hip.jacobian_source(s)
        s.close()
        return n, dt, stages, float(l2r[-1]), dt_ev, (tune_s, jsrc)

    def tune(s):
        """The autotuner (FAST: launch shapes, kernels, stored or recomputed
        Jacobians) on the freestream state, outside the clock like context
        creation and graph capture; its seconds are reported."""
        if args.adi != "fast" or args.no_autotune:
            return 0.0
        t0 = time.perf_counter()
        s.computeRHS("implicit")
        hip.autotune(s)
        if args.jacobians != "auto":
            hip.set_jacobian_source(s, {"stored": hip.JAC_STORED, "states": hip.JAC_FROM_STATES,
                                        "q": hip.JAC_FROM_Q, "states_lines": hip.JAC_STATES_LINES}[args.jacobians])
        s.sync()
        return time.perf_counter() - t0

    def converge(lhs, mode=None):
        mode = adi_mode if mode is None else mode
        s = hip.Solver(mesh, args.mach, args.alpha, args.order, args.cfl, lhs=lhs, adi_mode=mode)
        tune(s)
        s.prepare_graph(50)
        s.sync()
        t0 = time.perf_counter()
        n, l2r, _ = s.step(args.max_iters, tol=args.tol, check_every=50)
        s.sync()
        dt = time.perf_counter() - t0
        q = s.get("q")
        s.close()
        return {"iterations": n, "seconds": dt, "final_l2": float(l2r[-1]), "converged": bool(l2r[-1] <= args.tol)}, q

    n, dt, stages, l2, dt_ev, tune_s = run(hip.LHS_FP64)
    value = mesh.n_cell * n / dt
    total_b, stage_b = algorithmic_bytes(mesh.n_cell, n_proc, phase_faces, 8, phase_cells=phase_cells)
    # dominant kernel by event time; its algorithmic bytes per launch (a phase's
    # bytes are split over the kernels of that phase in proportion to time)
    group = {"rhs_faces": "rhs", "rhs_cells": "rhs", "accum_dq": "linear", "lin_faces": "linear",
             "lin_cells": "linear", "norms": "linear"}
    dom = max(stages, key=stages.get)
    g = group.get(dom, dom)
    g_time = sum(v for k, v in stages.items() if group.get(k, k) == g)
    dom_bytes = stage_b[g] * mesh.n_cell * stages[dom] / g_time
    achieved = dom_bytes / (stages[dom] * 1e-3) / 1e9
    out = {
        "metric": "cell-iters/s (FP64 line-implicit ADI iteration)",
        "value": value,
        "unit": "cell-iters/s",
        "n_gpus": 1,
        "steps": n,
        "warmup": args.warmup,
        "ms_per_step": dt / n * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: generated O-grid, freestream initial state",
        "config": {"workload": f"NACA0012 O-grid {args.nc}x{args.nr}, " +
                               (f"outer {args.tri_layers} layers split into triangles" if args.tri_layers
                                else "all-quad") +
                               f", order {args.order} (MUSCL), M={args.mach}, alpha={args.alpha}, CFL={args.cfl}, "
                               f"msweep=1",
                   "cells": mesh.n_cell, "faces": n_proc, "orphan_faces": int(len(mesh.orphans)), "lhs": "fp64",
                   "adi": args.adi + (" (parallel cyclic reduction per chain: residual, norms and update "
                                      "bit-identical to the reference; block solves equal up to rounding)"
                                      if args.adi == "fast" else " (the reference's sequential block Thomas, "
                                      "bit-identical)")},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(dom, args.nc, args.nr, args.order, "fp64", args.adi,
                                                 args.tri_layers, tune_s[1]),
                     "algorithmic_bytes_per_launch": dom_bytes, "avg_launch_us": stages[dom] * 1e3},
        "iteration_roofline": {"bytes_per_cell_iter": total_b,
                               "achieved_gbs": value * total_b / 1e9,
                               "frac": value * total_b / 1e9 / HBM_PEAK_GBS},
        "stage_ms": {k: round(v, 5) for k, v in stages.items()},
        "ms_per_step_eager_with_events": dt_ev / n * 1e3,
        "l2_after_timed_steps": l2,
        "autotune_seconds": tune_s[0],
        "jacobian_source": ["stored", "from q", "from stored face states",
                            "stored for phase 0, from stored face states for the others"][tune_s[1]],
    }
    if not args.no_converge:
        c64, q64 = converge(hip.LHS_FP64)
        out["time_to_tol_fp64"] = c64
    if args.adi == "fast" and not args.no_exact:
        ne, dte, ste, _, _, _ = run(hip.LHS_FP64, hip.ADI_EXACT, steps=min(args.steps, 30))
        out["exact_adi_fp64"] = {"value": mesh.n_cell * ne / dte, "ms_per_step": dte / ne * 1e3,
                                 "stage_ms": {k: round(v, 5) for k, v in ste.items()}}
        if not args.no_converge:
            import numpy as np
            ce, qe = converge(hip.LHS_FP64, hip.ADI_EXACT)
            scale = np.abs(qe.reshape(-1, 4)).max(axis=0)
            ce["max_field_diff_vs_fast_rel"] = float((np.abs(q64 - qe).reshape(-1, 4) / scale).max())
            out["exact_adi_fp64"]["time_to_tol"] = ce
    if not args.no_fp32:
        n32, dt32, st32, _, _, _ = run(hip.LHS_FP32)
        v32 = mesh.n_cell * n32 / dt32
        b32, _ = algorithmic_bytes(mesh.n_cell, n_proc, phase_faces, 4, phase_cells=phase_cells)
        out["fp32_lhs"] = {"value": v32, "ms_per_step": dt32 / n32 * 1e3,
                           "iteration_roofline": {"bytes_per_cell_iter": b32, "achieved_gbs": v32 * b32 / 1e9,
                                                  "frac": v32 * b32 / 1e9 / HBM_PEAK_GBS},
                           "stage_ms": {k: round(v, 5) for k, v in st32.items()}}
        if not args.no_converge:
            import numpy as np
            c32, q32 = converge(hip.LHS_FP32)
            scale = np.abs(q64.reshape(-1, 4)).max(axis=0)
            c32["max_field_diff_vs_fp64_rel"] = float((np.abs(q32 - q64).reshape(-1, 4) / scale).max())
            out["fp32_lhs"]["time_to_tol"] = c32
    if args.adi == "fast" and not args.no_exact and not args.no_deep and not args.no_converge:
        out["fast_vs_exact_deep"] = deep_section(args, mesh)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(mesh, args.mach, args.alpha, args.order, args.cfl, args.cpu_budget)
    if not args.no_explicit:
        out["explicit_schemes"] = explicit_section(args, mesh, n_proc)
    if not args.no_weno:
        out["weno_transonic_fp32_lhs"] = weno_section(args)
    if args.copy_probe:
        out["copy_probe_bytes"] = copy_probe(256 << 20)
    if not args.no_sweep and args.sweep_cases > 0:
        out["sweep"] = sweep_section(args, with_cpu=not args.no_cpu)
    print(json.dumps(out), flush=True)


def deep_section(args, mesh, steps=8000):
    """FAST against EXACT beyond L2 = 1e-10.

    On the bench mesh the scheme settles into a limit cycle (its norm stays
    near 3.2e-11 from about step 7,000 on, in the oracle as on the GPU), so
    both run `steps` iterations and the band of the norm over the last 1,000
    is reported with the field difference at the end.  The cycle belongs to
    the iteration at CFL 10: continued from there with the sweep's CFL
    back-off (CFL 5, then 2.5, each pass under the sweep's stall rule), both
    reach L2 <= 1e-13 on the bench mesh itself, and those converged fields
    are compared.  The all-quad variant of the same mesh converges to the
    roundoff floor at CFL 10: there both run to L2 <= 1e-13 and the
    converged fields are compared."""
    import numpy as np
    from hamstr_oop_amd import hip, meshgen, sweep

    def two(m, **kw):
        qs, recs = {}, {}
        for name, mode in (("fast", hip.ADI_FAST), ("exact", hip.ADI_EXACT)):
            s = hip.Solver(m, args.mach, args.alpha, args.order, args.cfl, adi_mode=mode)
            t0 = time.perf_counter()
            n, l2, _ = s.step(kw["n"], tol=kw.get("tol", 0.0), check_every=50)
            s.sync()
            recs[name] = {"iterations": int(n), "seconds": time.perf_counter() - t0, "final_l2": float(l2[-1])}
            if kw.get("tol", 0.0) == 0.0:
                tail = l2[-1000:]
                recs[name]["l2_band_last_1000"] = [float(tail.min()), float(tail.max())]
            qs[name] = s.get("q").reshape(-1, 4)
            s.close()
        scale = np.abs(qs["exact"]).max(axis=0)
        return recs, float((np.abs(qs["fast"] - qs["exact"]) / scale).max()), qs

    def backoff(m, qs, tol):
        """Both orderings continued from their states with the CFL back-off
        until L2 <= tol; per pass (cfl, iterations, final L2)."""
        recs, qe = {}, {}
        for name, mode in (("fast", hip.ADI_FAST), ("exact", hip.ADI_EXACT)):
            q, passes, t = qs[name], [], 0.0
            for cf in sweep.BACKOFF_CFL:
                s = hip.Solver(m, args.mach, args.alpha, args.order, cf, adi_mode=mode)
                s.set("q", q)
                s.set_retire(True, args.stall_window)
                t0 = time.perf_counter()
                n, l2, _ = s.step(args.max_iters, tol=tol, check_every=50)
                s.sync()
                t += time.perf_counter() - t0
                passes.append({"cfl": cf, "iterations": int(n), "final_l2": float(l2[-1])})
                q = s.get("q").reshape(-1, 4).copy()
                s.close()
                if l2[-1] <= tol:
                    break
            recs[name] = {"passes": passes, "seconds": t, "converged": bool(passes[-1]["final_l2"] <= tol)}
            qe[name] = q
        scale = np.abs(qe["exact"]).max(axis=0)
        d = np.abs(qe["fast"] - qe["exact"]) / scale
        return recs, float(d.max()), float(np.median(d.max(axis=1)))

    out = {}
    recs, d, qs = two(mesh, n=steps)
    out["bench_mesh_limit_cycle"] = {"steps": steps, **recs, "max_field_diff_rel": d}
    recs, d, dm = backoff(mesh, qs, 1e-13)
    out["bench_mesh_to_1e-13_cfl_backoff"] = {"tol": 1e-13, **recs, "max_field_diff_rel": d,
                                              "median_field_diff_rel": dm}
    quad = meshgen.naca_ogrid(args.nc, args.nr, t=0.12, tri_layers=0)
    recs, d, _ = two(quad, n=args.max_iters, tol=1e-13)
    out["all_quad_to_1e-13"] = {"tol": 1e-13, **recs, "max_field_diff_rel": d, "cells": quad.n_cell}
    return out


def explicit_bytes(n_cell, n_proc, scheme):
    """Bytes per cell-iteration of an explicit step's dataflow when every
    materialised array is read or written once per stage (face fluxes kept on
    chip): per stage q gathered (32), r and sigma written (40), the update's
    source read unless it is the gathered q (32) and its destination written
    (32), face normals read (16 per face).  euler: 104 + 16 phi.  rk3: three
    stages of 136 + 16 phi (stage 1 writes qt and q from the gathered q,
    stages 2 and 3 read qt and write q)."""
    phi = n_proc / n_cell
    if scheme == "euler":
        return 104 + 16 * phi
    return 3 * (136 + 16 * phi)


def explicit_section(args, mesh, n_proc, cfl=0.8, steps=200):
    """stepSolution's euler and rk3 branches on the bench mesh (CFL 0.8, the
    explicit schemes' stable range): graph-replayed steps timed like the ADI
    value, per-kernel event times, and the oracle (or the reference build when
    present) on a bounded CPU sample."""
    from hamstr_oop_amd import hip
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    out = {}
    for scheme in ("euler", "rk3"):
        s = hip.Solver(mesh, args.mach, args.alpha, args.order, cfl, adi_mode=hip.ADI_FAST)
        s.set_scheme(scheme)
        s.step(args.warmup, check_every=max(args.warmup, 1))
        s.prepare_graph(steps)
        s.sync()
        t0 = time.perf_counter()
        n, _, printed = s.step(steps, check_every=steps)
        s.sync()
        dt = time.perf_counter() - t0
        s.set_timing(True)
        s.step(20, check_every=20)
        stages = s.last_timings()
        s.close()
        value = mesh.n_cell * n / dt
        b = explicit_bytes(mesh.n_cell, n_proc, scheme)
        rec = {"value": value, "unit": "cell-iters/s", "ms_per_step": dt / n * 1e3, "cfl": cfl,
               "stage_ms": {k: round(v, 5) for k, v in stages.items()},
               "iteration_roofline": {"bytes_per_cell_iter": b, "achieved_gbs": value * b / 1e9,
                                      "frac": value * b / 1e9 / HBM_PEAK_GBS},
               "printed_l2_last": float(printed[-1])}
        if not args.no_cpu:
            rec["cpu_baseline"] = explicit_cpu_baseline(mesh, args.mach, args.alpha, args.order, cfl, scheme)
        out[scheme] = rec
    return out


def explicit_cpu_baseline(mesh, mach, alpha, order, cfl, scheme, budget_s=4.0, max_iters=30):
    """One core, a bounded number of explicit steps of the same case from
    freestream: the reference's stepSolution when its build is present, else
    the oracle's restatement."""
    import oracle_py
    from hamstr_oop_amd import meshgen
    if oracle_py.Reference.available("O2") and not len(mesh.orphans):
        ref = oracle_py.Reference("O2")
        with tempfile.TemporaryDirectory() as d:
            meshgen.write_reference_case(mesh, d, mach, alpha, order, 1, cfl, scheme=scheme)
            rc = ref.load(d)
            _, _, t1 = rc.step_solution(1)
            n = int(max(1, min(max_iters, budget_s / max(t1, 1e-6))))
            _, _, sec = rc.step_solution(n)
        kind, what = "reference", "reference SOLVER::stepSolution compiled -O2 from its sources"
    else:
        oc = oracle_py.Oracle().case(mesh, mach, alpha, order, cfl)
        _, _, _, t1 = oc.step_scheme(scheme, 1)
        n = int(max(1, min(max_iters, budget_s / max(t1, 1e-6))))
        _, _, _, sec = oc.step_scheme(scheme, n)
        kind, what = "port", "oracle C++ restatement, -O2"
    return {"value": mesh.n_cell * n / sec, "unit": "cell-iters/s", "cores": 1, "kind": kind,
            "sample": f"{n} {scheme} steps (after 1 warm-up) from freestream; {what}; {sec:.2f} s wall on "
                      f"{cpu_model()}"}


def weno_section(args, nc=1024, nr=412, tri=28, mach=0.8, alpha=1.25):
    """WENO (order 5), transonic: NACA0012 O-grid of 1024 x 384 quads plus a
    triangulated far-field ring of 28 layers (1024 x 2 x 28 triangles; 450,560
    cells), M=0.8, alpha=1.25 deg.  FP64 and FP32-LHS runs from the freestream
    to L2 <= tol under the sweep's stall rule; the FP32-LHS error is stated
    against the FP64 field.  Each context is autotuned first, outside the
    clock like context creation."""
    import numpy as np
    from hamstr_oop_amd import hip, meshgen
    mesh = meshgen.naca_ogrid(nc, nr, t=0.12, tri_layers=tri)
    res, qs = {}, {}
    for name, lhs in (("fp64", hip.LHS_FP64), ("fp32_lhs", hip.LHS_FP32)):
        s = hip.Solver(mesh, mach, alpha, 5, args.cfl, lhs=lhs, adi_mode=hip.ADI_FAST)
        s.computeRHS("implicit")
        hip.autotune(s)
        s.set_retire(True, args.stall_window)
        s.sync()
        t0 = time.perf_counter()
        n, l2, _ = s.step(args.max_iters, tol=args.tol, check_every=50)
        s.sync()
        dt = time.perf_counter() - t0
        st, it = s.case_status()
        res[name] = {"status": ["running", "converged", "stalled", "diverged"][int(st[0])], "iterations": int(n),
                     "seconds": dt, "cell_iters_per_s": mesh.n_cell * n / dt, "final_l2": float(l2[-1]),
                     "min_l2": float(l2.min())}
        qs[name] = s.get("q").reshape(-1, 4)
        s.close()
    scale = np.abs(qs["fp64"]).max(axis=0)
    # both stop in the scheme's limit cycle, so this is the spread of the cycle
    res["fp32_lhs"]["max_field_diff_vs_fp64_rel"] = float((np.abs(qs["fp32_lhs"] - qs["fp64"]) / scale).max())
    res["workload"] = f"NACA0012 O-grid {nc}x{nr - tri} quads + {tri} triangulated far-field layers " \
                      f"({mesh.n_cell} cells, {len(mesh.orphans)} orphan faces), order 5 (WENO), M={mach}, " \
                      f"alpha={alpha}, CFL={args.cfl}"
    # The FP32 left-hand side's error where the iteration converges: the same
    # mesh and scheme at M=0.5 (no shock).  Both precisions run to L2 <= 1e-12,
    # WENO's noise floor here (about 5e-13), and the fields are compared; the
    # FP64 field's own change over 1000 more steps at that floor is the scale
    # to read the difference against (WENO's weights keep switching near the
    # leading edge, moving a few cells there by ~1e-3).
    sub, qsub = {"mach": 0.5, "tol": 1e-12}, {}
    for name, lhs in (("fp64", hip.LHS_FP64), ("fp32_lhs", hip.LHS_FP32)):
        s = hip.Solver(mesh, 0.5, alpha, 5, args.cfl, lhs=lhs, adi_mode=hip.ADI_FAST)
        s.computeRHS("implicit")
        hip.autotune(s)
        s.sync()
        t0 = time.perf_counter()
        n, l2, _ = s.step(40000, tol=1e-12, check_every=100)
        s.sync()
        sub[name] = {"iterations": int(n), "seconds": time.perf_counter() - t0, "final_l2": float(l2[-1]),
                     "converged": bool(l2[-1] <= 1e-12)}
        qsub[name] = s.get("q").reshape(-1, 4).copy()
        if name == "fp64":
            s.step(1000, check_every=100)
            sc = np.abs(qsub[name]).max(axis=0)
            dself = (np.abs(s.get("q").reshape(-1, 4) - qsub[name]) / sc).max(axis=1)
            sub[name]["self_change_next_1000_steps_rel"] = {"max": float(dself.max()),
                                                            "median": float(np.median(dself))}
        s.close()
    scale = np.abs(qsub["fp64"]).max(axis=0)
    d = (np.abs(qsub["fp32_lhs"] - qsub["fp64"]) / scale).max(axis=1)
    sub["fp32_lhs"]["max_field_diff_vs_fp64_rel"] = float(d.max())
    sub["fp32_lhs"]["median_field_diff_vs_fp64_rel"] = float(np.median(d))
    res["subsonic_fixed_point"] = sub
    return res


def weno_profile(args, steps, nc=1024, nr=412, tri=28, mach=0.8, alpha=1.25):
    """The WENO transonic case of weno_section, `steps` iterations per
    left-hand side from the freestream after the autotuner: `steps` untimed,
    `steps` timed replaying a captured graph, then `steps` with events between
    kernels for stage times."""
    from hamstr_oop_amd import hip, meshgen
    mesh = meshgen.naca_ogrid(nc, nr, t=0.12, tri_layers=tri)
    out = {"workload": f"NACA0012 O-grid {nc}x{nr - tri} quads + {tri} triangulated far-field layers "
                       f"({mesh.n_cell} cells), order 5 (WENO), M={mach}, alpha={alpha}, CFL={args.cfl}",
           "steps": steps}
    for name, lhs in (("fp64", hip.LHS_FP64), ("fp32_lhs", hip.LHS_FP32)):
        s = hip.Solver(mesh, mach, alpha, 5, args.cfl, lhs=lhs, adi_mode=hip.ADI_FAST)
        s.computeRHS("implicit")
        hip.autotune(s)
        s.prepare_graph(steps)
        s.step(steps, check_every=steps)  # the graph's first replay carries a one-time cost
        s.sync()
        t0 = time.perf_counter()
        s.step(steps, check_every=steps)
        s.sync()
        dt = time.perf_counter() - t0
        s.set_timing(True)
        s.step(steps, check_every=steps)
        out[name] = {"ms_per_step": dt / steps * 1e3, "cell_iters_per_s": mesh.n_cell * steps / dt,
                     "stage_ms": {k: round(v, 5) for k, v in s.last_timings().items()}}
        s.close()
    return out


def copy_probe(nbytes):
    """One plain device-to-device copy of nbytes through the HIP runtime (its
    blit kernel reads nbytes and writes nbytes): the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes as C
    rt = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    ok = rt.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0 and rt.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    if ok:
        ok = rt.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
        ok = ok and rt.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0   # hipMemcpyDeviceToDevice
        ok = ok and rt.hipDeviceSynchronize() == 0
    for p in (src, dst):
        if p.value:
            rt.hipFree(p)
    return {"read": nbytes, "written": nbytes, "ok": bool(ok)}


def sweep_section(args, with_cpu=True):
    """Mixed batch of flow cases to L2 <= tol: plain FP64 one case at a time
    from the freestream, against one batched FP32-LHS context (autotuned,
    converged cases retired) from the freestream and from the learned warm
    start.  Cases that settle into a limit cycle count as stalled in every run."""
    import numpy as np
    from hamstr_oop_amd import hip, sweep
    log = lambda m: print(f"[sweep] {m}", file=sys.stderr, flush=True)  # noqa: E731
    specs = sweep.draw_table(args.sweep_cases)
    meshes = [c.mesh() for c in specs]
    ncell = np.array([m.n_cell for m in meshes])
    tol, mx, sw = args.tol, args.max_iters, args.stall_window
    res = {"cases": len(specs), "seed": sweep.SEED, "tol": tol, "stall_window": sw,
           "cells": int(ncell.sum()),
           "mesh_mix": {f"{a}x{b}": int(sum(1 for c in specs if (c.nc, c.nr) == (a, b))) for a, b in sweep.SIZES}}
    log(f"{len(specs)} cases, {ncell.sum()} cells")
    cold64 = sweep.run_sequential(specs, meshes, tol, mx, sw, lhs=hip.LHS_FP64, progress=log)
    res["sequential_fp64_cold"] = cold64.summary()
    log(f"sequential FP64 cold: {cold64.seconds:.2f} s")
    b64, o64, _ = sweep.run_batched(specs, meshes, tol, mx, sw, lhs=hip.LHS_FP64)
    o64.close()
    res["batched_fp64_cold"] = b64.summary()
    b32, o32, i32 = sweep.run_batched(specs, meshes, tol, mx, sw, lhs=hip.LHS_FP32)
    res["batched_fp32_cold"] = {**b32.summary(), **i32}
    log(f"batched cold: FP64 {b64.seconds:.2f} s, FP32-LHS {b32.seconds:.2f} s ({i32})")
    if args.train_cases > 0:
        from hamstr_oop_amd import warmstart as ws
        mlp, norm, info = sweep.train_network(args.train_cases, steps=args.train_steps, progress=log)
        res["warm_start_training"] = info
        feats = [ws.cell_features(m, c.mach, c.alpha, c.thickness) for c, m in zip(specs, meshes)]
        fb = []
        bw, ow, iw = sweep.run_batched(specs, meshes, tol, mx, sw, lhs=hip.LHS_FP32,
                                       init=sweep.warm_init(mlp, norm, feats, fb))
        if args.heldout > 0:
            res["heldout_256x128"] = heldout_section(args, mlp, norm, log)
        mlp.close()
        s = {**bw.summary(), **iw}
        s["fallback_cells_first_pass"] = int(fb[0]) if fb else None
        both = (bw.status == 1) & (b32.status == 1)
        s["iterations_vs_cold_fp32"] = float(bw.iters[both].sum() / max(b32.iters[both].sum(), 1))
        # the fixed point does not depend on the start: converged fields of both runs
        diffs = []
        for k in np.nonzero(both)[0]:
            qc, qw = o32.case_q(k), ow.case_q(k)
            diffs.append(float((np.abs(qw - qc) / np.abs(qc).max(axis=0)).max()))
        s["max_field_diff_vs_cold_rel"] = max(diffs) if diffs else None
        ow.close()
        res["batched_fp32_warm"] = s
        res["speedup_batched_fp32_warm_vs_sequential_fp64_cold"] = cold64.seconds / bw.seconds
        log(f"batched FP32-LHS warm: {bw.seconds:.2f} s ({iw})")
    o32.close()
    res["speedup_batched_fp32_cold_vs_sequential_fp64_cold"] = cold64.seconds / b32.seconds
    # the same comparison over the cases every run converges, without the
    # stall-detection time of cases that never converge (a batched run's time
    # is when the last of them converged)
    runs = (("batched_fp64_cold", b64), ("batched_fp32_cold", b32)) + \
        ((("batched_fp32_warm", bw),) if args.train_cases > 0 else ())
    conv = [k for k in range(len(specs)) if cold64.status[k] == 1 and all(r.status[k] == 1 for _, r in runs)]
    t_seq = cold64.time_to_converge(conv)
    cmp = {"cases": len(conv), "sequential_fp64_cold_s": t_seq}
    for name, r in runs:
        t = r.time_to_converge(conv)
        cmp[name + "_s"] = t
        cmp["speedup_" + name] = (t_seq / t) if t else None
    res["converged_set"] = cmp
    if with_cpu:
        res["cpu_baseline"] = sweep_cpu_baseline(specs, meshes, cold64)
    return res


def heldout_section(args, mlp, norm, log):
    """The warm start's own evaluation: held-out cases on 256x128 meshes (a
    seed of their own, the sweep's parameter ranges), batched FP32-LHS with the
    FP64 fallback from the freestream and from the network; iterations and
    wall-clock to L2 <= tol."""
    from hamstr_oop_amd import hip, sweep
    from hamstr_oop_amd import warmstart as ws
    specs = sweep.draw_table(args.heldout, seed=sweep.SEED + 2, sizes=((256, 128),))
    meshes = [c.mesh() for c in specs]
    tol, mx, sw = args.tol, args.max_iters, args.stall_window
    cold, oc, ic = sweep.run_batched(specs, meshes, tol, mx, sw, lhs=hip.LHS_FP32)
    oc.close()
    feats = [ws.cell_features(m, c.mach, c.alpha, c.thickness) for c, m in zip(specs, meshes)]
    warm, ow, iw = sweep.run_batched(specs, meshes, tol, mx, sw, lhs=hip.LHS_FP32,
                                     init=sweep.warm_init(mlp, norm, feats))
    ow.close()
    both = (cold.status == 1) & (warm.status == 1)
    out = {"cases": len(specs), "seed": sweep.SEED + 2, "mesh": "256x128",
           "cold": {**cold.summary(), **ic}, "warm": {**warm.summary(), **iw},
           "converged_both": int(both.sum()),
           "iterations_warm_over_cold": float(warm.iters[both].sum() / max(cold.iters[both].sum(), 1)),
           "wallclock_warm_over_cold": warm.seconds / cold.seconds,
           "time_to_converge_both_s": {"cold": cold.time_to_converge(list(both.nonzero()[0])),
                                       "warm": warm.time_to_converge(list(both.nonzero()[0]))}}
    log(f"held-out 256x128: cold {cold.seconds:.2f} s, warm {warm.seconds:.2f} s, "
        f"iterations x{out['iterations_warm_over_cold']:.3f}")
    return out


def sweep_cpu_baseline(specs, meshes, cold64, iters=40):
    """The oracle's C++ restatement on the host cores, one case per core, on a
    bounded sample (the first `iters` iterations of 2 cases per mesh size);
    time to tol extrapolated from its per-core rate and the GPU run's iteration
    counts.  Not the reference binary: one process per core needs the
    leak-free restatement."""
    import concurrent.futures as cf
    import numpy as np
    from hamstr_oop_amd import sweep
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    cores = min(os.cpu_count() or 1, 16)
    picks = []
    for size in sweep.SIZES:
        picks += [k for k, c in enumerate(specs) if (c.nc, c.nr) == size][:2]
    orc = oracle_py.Oracle()

    def one(k):
        c = specs[k]
        oc = orc.case(meshes[k], c.mach, c.alpha, 3, 10.0)
        _, _, _, sec = oc.step(iters)
        return meshes[k].n_cell * iters / sec

    with cf.ThreadPoolExecutor(max_workers=min(cores, len(picks))) as ex:
        rates = list(ex.map(one, picks))
    rate = float(np.mean(rates))   # cell-iters/s per core with min(cores, samples) running at once
    need = float(sum(m.n_cell * (it if st > 0 else 0) for m, it, st in zip(meshes, cold64.iters, cold64.status)))
    return {"value": rate * cores, "unit": "cell-iters/s", "cores": cores, "kind": "port",
            "per_core": rate, "time_to_tol_extrapolated_s": need / (rate * cores),
            "sample": f"{iters} iterations of {len(picks)} cases (2 per mesh size), one case per thread on "
                      f"{cpu_model()}; time to tol = sum of cells x FP64 GPU iterations / (per-core rate x {cores})"}


if __name__ == "__main__":
    main()

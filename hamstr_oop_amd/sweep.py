"""Mixed-batch sweep: many flow cases to a residual tolerance, run one case at
a time in plain FP64 from the freestream, and as one batched context (FP32 LHS,
autotuned persistent launches, retirement of converged cases, optionally a
learned warm start).

Cases are drawn from a seed: Mach ~ U[0.2, 0.75], alpha ~ U[-4, 6] deg,
thickness ~ U[0.08, 0.16], mesh drawn from 128x64, 256x128 and 384x192
NACA-00xx O-grids (radial lines of 64 to 192 cells).  MUSCL (order 3), CFL 10.

A case converges at the first iteration whose nonlinear L2 = sqrt(sum r_rho^2)
/ nCell is <= tol.  Some cases of this scheme settle into a limit cycle just
above 1e-10 (the MUSCL limiter's eps = 1e-10 makes it non-smooth at extrema;
computeRHS.cpp:268-271): a case whose norm has not halved within stall_window
iterations is called stalled, in every run alike, and counted separately.  A
few cases blow up from the freestream at CFL 10 (a NaN sound speed, where the
reference exits): they count as diverged.

Most of those limit cycles belong to the iteration at CFL 10, not to the
discrete equations: continued at CFL 5, nearly every stalled case converges
within a few hundred iterations, to the same fixed point (the CFL number
scales the implicit operator and the pseudo-time step, not the residual).
Every run therefore ends with the same CFL back-off (BACKOFF_CFL): cases
still stalled continue from the state they reached, and diverged ones restart
from their initial state, at CFL 5 and then 2.5, plain FP64 ADI each time,
under the same stall rule.  A case converged in any pass counts as converged.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import hip, meshgen

SIZES = ((128, 64), (256, 128), (384, 192))
SEED = 20261203
# CFL numbers of the passes that continue stalled cases and restart diverged
# ones (after the run at its own CFL); () turns the back-off off
BACKOFF_CFL = (5.0, 2.5)


@dataclass
class CaseSpec:
    mach: float
    alpha: float
    thickness: float
    nc: int
    nr: int

    def mesh(self) -> meshgen.Mesh:
        return meshgen.naca_ogrid(self.nc, self.nr, t=self.thickness)


def draw_table(n: int = 256, seed: int = SEED, sizes: Sequence = SIZES) -> List[CaseSpec]:
    """The case table: per case Mach, alpha, thickness, then the mesh index."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        m = float(rng.uniform(0.2, 0.75))
        a = float(rng.uniform(-4.0, 6.0))
        t = float(rng.uniform(0.08, 0.16))
        nc, nr = sizes[int(rng.integers(len(sizes)))]
        out.append(CaseSpec(m, a, t, nc, nr))
    return out


@dataclass
class RunResult:
    seconds: float
    iters: np.ndarray       # iteration of convergence (or stall detection) per case
    status: np.ndarray      # 1 converged, 2 stalled, 3 diverged, 0 none of these within max_iters
    cell_iters: float       # sum over cases of cells x iterations run
    case_seconds: Optional[np.ndarray] = None  # sequential: each case's solve time
    conv_time: Optional[np.ndarray] = None     # batched: wall time from the start at which the case converged

    def time_to_converge(self, cases) -> Optional[float]:
        """Wall time to converge every case in `cases` (None if one of them did
        not): the sum of their solve times for a sequential run, the time the
        last of them converged for a batched one."""
        cases = list(cases)
        if not all(self.status[k] == 1 for k in cases):
            return None
        if self.case_seconds is not None:
            return float(sum(self.case_seconds[k] for k in cases))
        return float(max(self.conv_time[k] for k in cases)) if cases else 0.0

    def summary(self) -> dict:
        return {"seconds": round(self.seconds, 4), "converged": int((self.status == 1).sum()),
                "stalled": int((self.status == 2).sum()), "diverged": int((self.status == 3).sum()),
                "unfinished": int((self.status == 0).sum()),
                "iterations_sum": int(self.iters[self.status > 0].sum()), "cell_iters": float(self.cell_iters)}


def run_sequential(specs: Sequence[CaseSpec], meshes, tol: float, max_iters: int, stall_window: int,
                   lhs: int = hip.LHS_FP64, check_every: int = 50, order: int = 3, cfl: float = 10.0,
                   progress=None, backoff: Sequence[float] = BACKOFF_CFL) -> RunResult:
    """One single-case context per case, freestream start; the clock covers the
    iterations only (context set-up excluded).  A case that stalls or
    diverges goes through the back-off passes (a context of its own per CFL
    number, FP64 left-hand side), their iterations added to its own."""
    iters = np.full(len(specs), -1, np.int64)
    status = np.zeros(len(specs), np.int32)
    secs = np.zeros(len(specs))
    total = 0.0
    cell_iters = 0.0
    for k, (c, mesh) in enumerate(zip(specs, meshes)):
        done = 0
        q = None
        for p, cf in enumerate((cfl,) + tuple(backoff)):
            if p > 0 and status[k] not in (2, 3):
                break
            s = hip.Solver(mesh, c.mach, c.alpha, order, cf, lhs=lhs if p == 0 else hip.LHS_FP64,
                           adi_mode=hip.ADI_FAST)
            s.set_retire(True, stall_window)
            s.prepare_graph(check_every)   # capture outside the clock, like context creation
            s.sync()
            t0 = time.perf_counter()
            if q is not None:
                s.set("q", q)   # a stalled case continues from its state
            n, _, _ = s.step(max_iters, tol=tol, check_every=check_every)
            s.sync()
            secs[k] += time.perf_counter() - t0
            st, it = s.case_status()
            ran = int(it[0]) if st[0] > 0 else n   # iterations until the case retired
            status[k] = st[0]
            iters[k] = done + ran
            done += ran
            cell_iters += mesh.n_cell * ran
            q = s.get("q") if st[0] == 2 else None
            s.close()
        total += secs[k]
        if progress and (k + 1) % 32 == 0:
            progress(f"sequential: {k + 1}/{len(specs)} cases, {total:.1f} s")
    return RunResult(total, iters, status, cell_iters, case_seconds=secs)


class BatchOutcome:
    """Final states of a batched run, per case: from the last of the run's
    contexts (the first pass, the FP64 fallback, the back-off passes) that
    ran the case."""

    def __init__(self, main):
        self.main = main
        self.layers = []   # (context, {case: index in that context}), in pass order

    def add(self, ctx, idx):
        self.layers.append((ctx, {k: j for j, k in enumerate(idx)}))

    def case_q(self, k: int) -> np.ndarray:
        for ctx, where in reversed(self.layers):
            if k in where:
                return ctx.case_q(where[k])
        return self.main.case_q(k)

    def close(self):
        for b in [self.main] + [ctx for ctx, _ in self.layers]:
            b.close()


def run_batched(specs: Sequence[CaseSpec], meshes, tol: float, max_iters: int, stall_window: int,
                lhs: int = hip.LHS_FP32, check_every: int = 50, order: int = 3, cfl: float = 10.0,
                init=None, autotune: bool = True, fallback: bool = True,
                backoff: Sequence[float] = BACKOFF_CFL):
    """All cases in one batched context.  init(batch, case_indices) sets the
    initial state (a warm start); it and the autotuner are timed with the
    iterations.  With an FP32 LHS, cases that stall or diverge are re-run with
    the FP64 LHS (the plain ADI iteration is the fallback, and decides convergence): a
    stalled case from the state it reached, a diverged one from its initial
    state.  Then (with fallback) the cases still stalled or diverged go through
    the back-off passes the same way, one batched FP64 context per CFL number.
    Every pass applies the stall rule of the sequential runs (a shorter
    window for the FP32 pass hands cases that are merely on a plateau to the
    FP64 pass and costs more than it saves).
    Returns (RunResult, BatchOutcome, info)."""
    n = len(specs)
    b = hip.Batch([(m, c.mach, c.alpha) for c, m in zip(specs, meshes)], order=order, cfl=cfl, lhs=lhs,
                  adi_mode=hip.ADI_FAST)
    b.set_retire(True, stall_window)
    b.sync()
    t0 = time.perf_counter()
    if autotune:
        b.computeRHS("implicit")   # a valid state for the timing launches (r and dq are restored)
        hip.autotune(b)
    if init is not None:
        init(b, list(range(n)))
    conv_time = np.full(n, np.nan)

    def stepped(ctx, idx, limit):
        """Read-back-sized chunks until every case retired (or limit), noting
        when each converged."""
        ran = 0
        while ran < limit:
            done, _, _ = ctx.batch_step(min(check_every, limit - ran), tol=tol, check_every=check_every)
            ran += done
            stc, _ = ctx.case_status()
            now = time.perf_counter() - t0
            for j, k in enumerate(idx):
                if stc[j] == 1 and np.isnan(conv_time[k]):
                    conv_time[k] = now
            if (stc != 0).all() or done == 0:
                break

    stepped(b, list(range(n)), max_iters)
    t_first = time.perf_counter() - t0
    st, it = b.case_status()
    status, iters = st.copy(), it.astype(np.int64)
    ncell = np.diff(b.cell_off).astype(np.float64)
    cell_iters = float((ncell * np.where(st > 0, it, max_iters)).sum())
    info = {"stalled_first_pass": int((st == 2).sum()), "diverged_first_pass": int((st == 3).sum()),
            "fallback_cases": 0}
    outcome = BatchOutcome(b)
    passes = []
    if fallback:
        if lhs == hip.LHS_FP32:
            passes.append(("fallback", cfl))
        passes += [(f"backoff_cfl{cf:g}", cf) for cf in backoff]
    for name, cf in passes:
        redo = [k for k in range(n) if status[k] in (2, 3)]
        if not redo:
            break
        # a stalled case continues from the state it reached (in the last
        # context that ran it), a diverged one restarts from its initial state
        prev = {k: outcome.case_q(k) for k in redo if status[k] == 2}
        sub = hip.Batch([(meshes[k], specs[k].mach, specs[k].alpha) for k in redo], order=order, cfl=cf,
                        lhs=hip.LHS_FP64, adi_mode=hip.ADI_FAST)
        outcome.add(sub, redo)   # closed with the outcome
        if init is not None:
            init(sub, redo)
        for j, k in enumerate(redo):
            if k in prev:
                sub.set_case_q(j, prev[k])
        sub.set_retire(True, stall_window)
        stepped(sub, redo, max_iters)
        st2, it2 = sub.case_status()
        for j, k in enumerate(redo):
            run2 = it2[j] if st2[j] > 0 else max_iters
            status[k] = st2[j]
            iters[k] += run2
            cell_iters += ncell[k] * run2
        if name == "fallback":
            info["fallback_cases"] = len(redo)
            info["fallback_converged"] = int((st2 == 1).sum())
        else:
            info[name + "_cases"] = len(redo)
            info[name + "_converged"] = int((st2 == 1).sum())
    b.sync()
    sec = time.perf_counter() - t0
    info["first_pass_s"] = round(t_first, 3)
    info["fallback_s"] = round(sec - t_first, 3)
    return RunResult(sec, iters, status, cell_iters, conv_time=conv_time), outcome, info


def train_network(n_train: int, seed: int = SEED + 1, size=(256, 128), tol: float = 1e-11, max_iters: int = 20000,
                  stall_window: int = 2000, per_case: int = 8192, chunk: int = 250, steps: int = 20000,
                  batch: int = 4096, lr: float = 2e-3, lr_final: float = 1e-5, mlp_seed: int = 2, progress=None):
    """Solves n_train cases (drawn like the sweep, on one mesh size) with the
    batched FP32-LHS path to tol, samples per_case cells of each converged field,
    and trains the warm-start MLP on the GPU.

    The depth of the training fields matters more than their number: fields at
    L2 <= 1e-8 still carry a circulation several per cent off, the slowest
    mode of the iteration, and a network trained on them leaves it to the
    solver; trained on fields at 1e-11 it removes about half of the remaining
    iterations (warm/cold iteration ratio 0.63 -> 0.28 on 64 held-out cases).
    Cases that stall in a limit cycle above tol are left out of the set.
    Returns (mlp, normaliser, info)."""
    from . import warmstart as ws
    specs = draw_table(n_train, seed, sizes=(size,))
    rng = np.random.default_rng(seed)
    xs, ts = [], []
    t_solve = 0.0
    n_used = 0
    for c0 in range(0, n_train, chunk):
        part = specs[c0:c0 + chunk]
        meshes = [c.mesh() for c in part]
        res, out, _ = run_batched(part, meshes, tol, max_iters, stall_window, lhs=hip.LHS_FP32, autotune=False,
                                  fallback=False)
        t_solve += res.seconds
        for k in range(len(part)):
            if res.status[k] != 1:
                continue
            c = ws.Case(meshes[k], part[k].mach, part[k].alpha, part[k].thickness)
            x, t = ws.case_samples(c, out.case_q(k), per_case, rng)
            xs.append(x)
            ts.append(t)
            n_used += 1
        out.close()
        if progress:
            progress(f"training set: {c0 + len(part)}/{n_train} cases solved ({t_solve:.1f} s)")
    x = np.concatenate(xs)
    t = np.concatenate(ts)
    norm = ws.Normalizer.fit(x, t)
    mlp = ws.MLP(seed=mlp_seed)
    mlp.set_data(norm.nx(x), norm.nt(t))
    t0 = time.perf_counter()
    losses = mlp.train(steps, batch=batch, lr=lr, lr_final=lr_final, seed=seed)
    t_train = time.perf_counter() - t0
    info = {"cases": n_train, "cases_converged": n_used, "mesh": f"{size[0]}x{size[1]}", "tol": tol,
            "rows": int(x.shape[0]), "solve_s": round(t_solve, 3), "train_s": round(t_train, 3), "steps": steps,
            "batch": batch, "loss_first": float(losses[:10].mean()), "loss_last": float(losses[-100:].mean())}
    return mlp, norm, info


def warm_init(mlp, norm, feats: Sequence[np.ndarray], fallback: Optional[list] = None):
    """init(batch, case_indices) for run_batched: the MLP's prediction from the
    per-case feature arrays feats[k] as initial state."""
    from . import warmstart as ws

    def init(b, idx):
        n = ws.warm_start(b, mlp, np.concatenate([feats[k] for k in idx]), norm)
        if fallback is not None:
            fallback.append(n)
    return init


def time_batch_iteration(n_cases: int = 256, lhs: int = hip.LHS_FP32, warmup: int = 20, iters: int = 50,
                         autotune: bool = True, seed: int = SEED) -> dict:
    """ms per batched iteration of the drawn case table, every case stepping
    (no tolerance, no retirement): first `iters` iterations replaying captured
    graphs, then `iters` launched eagerly with events between kernels for the
    per-stage times.  The launch groups are autotuned first unless
    autotune=False."""
    specs = draw_table(n_cases, seed)
    b = hip.Batch([(c.mesh(), c.mach, c.alpha) for c in specs], order=3, cfl=10.0, lhs=lhs,
                  adi_mode=hip.ADI_FAST)
    try:
        b.computeRHS("implicit")
        groups = hip.autotune(b)[0] if autotune else hip.launch_groups(b)
        b.batch_step(warmup, check_every=max(warmup, 1))
        b.sync()
        t0 = time.perf_counter()
        b.batch_step(iters, check_every=iters)
        b.sync()
        ms = (time.perf_counter() - t0) / iters * 1e3
        b.set_timing(True)
        b.batch_step(iters, check_every=iters)
        stages = b.last_timings()
        b.set_timing(False)
        return {"cases": n_cases, "cells": int(b.cell_off[-1]), "lhs": "fp32" if lhs == hip.LHS_FP32 else "fp64",
                "ms_per_iter": round(ms, 4), "stage_ms": {k: round(v, 4) for k, v in stages.items()},
                "jacobian_source": ["stored", "from q", "from stored face states",
                                    "stored for phase 0, from stored face states for the others"][
                                        hip.jacobian_source(b)],
                "groups": [{k: g[k] for k in ("phase", "kernel", "final", "chains", "max_rows", "grid", "cr_levels")}
                           for g in groups]}
    finally:
        b.close()


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="Time the batched iteration of the seeded case table (one GPU).")
    ap.add_argument("--cases", type=int, default=256)
    ap.add_argument("--lhs", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--no-autotune", action="store_true")
    a = ap.parse_args()
    print(json.dumps(time_batch_iteration(a.cases, hip.LHS_FP32 if a.lhs == "fp32" else hip.LHS_FP64, a.warmup,
                                          a.iters, not a.no_autotune)), flush=True)

"""Writes the oracle's own long runs, for the GPU tests to compare with at
sizes the oracle takes minutes to hours on:

  c2_oracle.npz    configuration 2: the 512x256 mixed tri/quad NACA0012
                   O-grid (MUSCL, M=0.5, alpha=2, CFL 10), 10,000 steps from the
                   freestream (about 45 minutes on one core)
  weno_cycle_oracle.npz  a reduced transonic WENO case that settles into a
                   limit cycle: the 128x48 NACA0012 O-grid with 4 triangulated
                   far-field layers, order 5, M=0.8, alpha=1.25, CFL 10, 6,000
                   steps (a minute)

The reference's own code cannot compute a mesh with triangles (its chains
cannot cover them), so these fixtures are the oracle's -- the CPU
restatement that the other fixtures in this directory pin bit for bit to the
reference -- running the orphan-face extension.

Stored per case (no state is kept whole):
  l2rhs, l2lin      per-step nonlinear and printed (linear) L2 norms
  q_sha256_<n>      SHA-256 of q's bytes after n steps (n in the case's list)
  q_sub_<n>         q of every 97th cell after n steps
  seconds           the oracle's time for the whole run, one core

Usage: make_oracle_runs.py [c2] [weno_cycle]   (no argument: both)
"""
from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from hamstr_oop_amd import meshgen  # noqa: E402
import oracle_py  # noqa: E402

CASES = {
    "c2": dict(nc=512, nr=256, tri_layers=128, mach=0.5, alpha=2.0, order=3, cfl=10.0, steps=10000,
               kept=(1000, 4200, 10000)),
    "weno_cycle": dict(nc=128, nr=52, tri_layers=4, mach=0.8, alpha=1.25, order=5, cfl=10.0, steps=6000,
                       kept=(1000, 6000)),
}
SUB = 97


def mesh(case):
    c = CASES[case]
    return meshgen.naca_ogrid(c["nc"], c["nr"], tri_layers=c["tri_layers"])


def digest(q: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(q, np.float64).tobytes()).hexdigest()


def run(case) -> None:
    c = CASES[case]
    o = oracle_py.Oracle().case(mesh(case), c["mach"], c["alpha"], c["order"], c["cfl"])
    out = {k: np.array(v) for k, v in c.items()}
    l2r, l2l = [], []
    done = 0
    t0 = time.time()
    for n in c["kept"]:
        while done < n:
            k = min(500, n - done)
            rc, a, b, _ = o.step(k)
            if rc != 0:
                raise SystemExit(f"oracle step failed with {rc} at step {done}")
            l2r.append(a)
            l2l.append(b)
            done += k
            print(f"{case} step {done}: l2 {a[-1]:.4e} ({time.time() - t0:.0f} s)", flush=True)
        q = o.get("q")
        out[f"q_sha256_{n}"] = np.array(digest(q))
        out[f"q_sub_{n}"] = q.reshape(-1, 4)[::SUB].copy()
    out["l2rhs"] = np.concatenate(l2r)
    out["l2lin"] = np.concatenate(l2l)
    out["seconds"] = np.array(time.time() - t0)
    np.savez_compressed(os.path.join(HERE, f"{case}_oracle.npz"), **out)
    print(f"wrote {case}_oracle.npz", flush=True)


if __name__ == "__main__":
    for name in (sys.argv[1:] or list(CASES)):
        run(name)

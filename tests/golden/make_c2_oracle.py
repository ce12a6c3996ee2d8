"""Writes c2_oracle.npz: the oracle's run of the 512x256 mixed tri/quad NACA0012
O-grid (MUSCL, M=0.5, alpha=2, CFL 10) from the freestream, to compare the
GPU with it at full size.

The reference's own code cannot compute a mesh with triangles (its chains
cannot cover them), so this fixture is the oracle's -- the CPU restatement
that the other fixtures in this directory pin bit for bit to the reference --
running the orphan-face extension.  It takes about 20 minutes on one core,
so the result is kept here instead of being recomputed by every test run.

Stored (no state is kept whole: q has 786,432 values):
  l2rhs, l2lin      per-step nonlinear and printed (linear) L2 norms
  q_sha256_<n>      SHA-256 of q's bytes after n steps (n in STEPS_KEPT)
  q_sub_<n>         q of every 97th cell after n steps
  seconds           the oracle's time for the whole run, one core
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

C2 = dict(nc=512, nr=256, tri_layers=128, mach=0.5, alpha=2.0, order=3, cfl=10.0, steps=10000)
STEPS_KEPT = (1000, 4200, 10000)
SUB = 97


def mesh():
    return meshgen.naca_ogrid(C2["nc"], C2["nr"], tri_layers=C2["tri_layers"])


def digest(q: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(q, np.float64).tobytes()).hexdigest()


def main() -> None:
    m = mesh()
    o = oracle_py.Oracle().case(m, C2["mach"], C2["alpha"], C2["order"], C2["cfl"])
    out = {k: np.array(v) for k, v in C2.items()}
    l2r, l2l = [], []
    done = 0
    t0 = time.time()
    for n in STEPS_KEPT:
        while done < n:
            k = min(500, n - done)
            rc, a, b, _ = o.step(k)
            if rc != 0:
                raise SystemExit(f"oracle step failed with {rc} at step {done}")
            l2r.append(a)
            l2l.append(b)
            done += k
            print(f"step {done}: l2 {a[-1]:.4e} ({time.time() - t0:.0f} s)", flush=True)
        q = o.get("q")
        out[f"q_sha256_{n}"] = np.array(digest(q))
        out[f"q_sub_{n}"] = q.reshape(-1, 4)[::SUB].copy()
    out["l2rhs"] = np.concatenate(l2r)
    out["l2lin"] = np.concatenate(l2l)
    out["seconds"] = np.array(time.time() - t0)
    np.savez_compressed(os.path.join(HERE, "c2_oracle.npz"), **out)
    print("wrote c2_oracle.npz", flush=True)


if __name__ == "__main__":
    main()

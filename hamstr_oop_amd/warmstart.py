"""Learned warm start for the ADI iteration.

An MLP (6 -> 256 x 4 SiLU -> 4, BF16 matrix-core kernels with FP32
accumulation, trained on the GPU by hamstr_mlp_train) maps per-cell features
(Mach, alpha in degrees, thickness, centroid x, centroid y, wall distance) to
the primitive state (rho, u, v, p) of the converged flow, as its deviation
from the case's freestream.  The centroid enters radially compressed,
(x, y) * asinh(r) / r, and the wall distance as its logarithm, so that the
cells near the section, where the field varies fastest, are spread over the
input range.  The network is trained on converged fields the solver itself
produces and only supplies the initial state of a new solve: the ADI
iteration still has to bring the nonlinear residual to the tolerance, and
cells whose prediction is not a valid state fall back to the freestream
(hamstr_warm_start).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

from . import hip, meshgen

_fp = C.POINTER(C.c_float)


def _f(a):
    return a.ctypes.data_as(_fp)


def _chk(rc):
    if rc != hip.OK:
        raise hip.HamstrError(rc, hip.lib().hamstr_mlp_last_error().decode())


def cell_features(mesh: meshgen.Mesh, mach: float, alpha: float, thickness: float) -> np.ndarray:
    """(nCell, 6) float32: Mach, alpha (deg), thickness, centroid x, y, wall distance."""
    from scipy.spatial import cKDTree
    cen = mesh.x[mesh.conn].mean(axis=1)
    # wall faces: boundary faces within the reference's wall radius (preprocess.cpp:169-191)
    f = mesh.faces
    bnd = f[:, 4] < 0
    mid = 0.5 * (mesh.x[f[bnd, 0]] + mesh.x[f[bnd, 1]])
    wall = mid[np.hypot(mid[:, 0], mid[:, 1]) < 3.0]
    # sample the wall densely (face midpoints and end nodes)
    wn = np.concatenate([wall, mesh.x[f[bnd, 0]][np.hypot(*mesh.x[f[bnd, 0]].T) < 3.0]])
    d, _ = cKDTree(wn).query(cen)
    out = np.empty((mesh.n_cell, 6), np.float32)
    out[:, 0] = mach
    out[:, 1] = alpha
    out[:, 2] = thickness
    r = np.maximum(np.hypot(cen[:, 0], cen[:, 1]), 1e-12)
    out[:, 3] = cen[:, 0] * np.arcsinh(r) / r
    out[:, 4] = cen[:, 1] * np.arcsinh(r) / r
    out[:, 5] = np.log(d + 1e-4)
    return out


def primitive(q: np.ndarray, gamma: float = 1.4) -> np.ndarray:
    q = q.reshape(-1, 4)
    rho = q[:, 0]
    u = q[:, 1] / rho
    v = q[:, 2] / rho
    p = (gamma - 1.0) * (q[:, 3] - 0.5 * rho * (u * u + v * v))
    return np.stack([rho, u, v, p], axis=1)


def freestream_primitive(mach: float, alpha: float, gamma: float = 1.4) -> np.ndarray:
    """(rho, u, v, p) of the freestream initFlow sets (initFlow.cpp:107-142)."""
    r, u, v, _ = meshgen.freestream(mach, alpha, gamma)
    return np.array([r, u, v, 1.0 / gamma])


@dataclass
class Normalizer:
    fmean: np.ndarray
    fstd: np.ndarray
    tmean: np.ndarray
    tstd: np.ndarray

    @staticmethod
    def fit(x: np.ndarray, t: np.ndarray) -> "Normalizer":
        return Normalizer(x.mean(0), x.std(0) + 1e-6, t.mean(0), t.std(0) + 1e-6)

    def packed(self) -> np.ndarray:
        return np.concatenate([self.fmean, self.fstd, self.tmean, self.tstd]).astype(np.float32)

    def nx(self, x):
        return ((x - self.fmean) / self.fstd).astype(np.float32)

    def nt(self, t):
        return ((t - self.tmean) / self.tstd).astype(np.float32)

    def dt(self, y):
        return y * self.tstd + self.tmean


class MLP:
    """The device MLP (hamstr_mlp_*)."""

    N_PARAMS = 200196

    def __init__(self, seed: int = 0):
        L = hip.lib()
        L.hamstr_mlp_last_error.restype = C.c_char_p
        h = C.c_void_p()
        _chk(L.hamstr_mlp_create(C.c_uint(seed), C.byref(h)))
        self.h = h
        assert L.hamstr_mlp_num_params() == self.N_PARAMS

    def close(self):
        if getattr(self, "h", None):
            hip.lib().hamstr_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graph(self, on: bool):
        """Replay one captured training step per step (default) or launch eagerly."""
        _chk(hip.lib().hamstr_mlp_set_graph(self.h, int(on)))

    def params(self) -> np.ndarray:
        p = np.zeros(self.N_PARAMS, np.float32)
        _chk(hip.lib().hamstr_mlp_get_params(self.h, _f(p)))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        _chk(hip.lib().hamstr_mlp_set_params(self.h, _f(p)))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], 4), np.float32)
        _chk(hip.lib().hamstr_mlp_forward(self.h, x.shape[0], _f(x), _f(y)))
        return y

    def loss_grad(self, x: np.ndarray, t: np.ndarray) -> Tuple[float, np.ndarray]:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        g = np.zeros(self.N_PARAMS, np.float32)
        loss = C.c_float()
        _chk(hip.lib().hamstr_mlp_loss_grad(self.h, x.shape[0], _f(x), _f(t), C.byref(loss), _f(g)))
        return loss.value, g

    def set_data(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        _chk(hip.lib().hamstr_mlp_set_data(self.h, C.c_long(x.shape[0]), _f(x), _f(t)))

    def train(self, steps: int, batch: int = 4096, lr: float = 1e-3, lr_final: float = 1e-5,
              seed: int = 1) -> np.ndarray:
        losses = np.zeros(max(steps, 1), np.float32)
        _chk(hip.lib().hamstr_mlp_train(self.h, int(steps), int(batch), C.c_float(lr), C.c_float(lr_final),
                                        C.c_uint(seed), _f(losses)))
        return losses[:steps]


def warm_start(solver: hip.Solver, mlp: MLP, feat: np.ndarray, norm: Normalizer) -> int:
    """Sets the solver's q from the MLP; returns the number of fallback cells."""
    feat = np.ascontiguousarray(feat, np.float32)
    nrm = norm.packed()
    nfb = C.c_int()
    _chk(hip.lib().hamstr_warm_start(solver.h, mlp.h, _f(feat), _f(nrm), C.byref(nfb)))
    return nfb.value


@dataclass
class Case:
    mesh: meshgen.Mesh
    mach: float
    alpha: float
    thickness: float


def draw_cases(n: int, rng: np.random.Generator, sizes: Sequence[Tuple[int, int]] = ((256, 128),),
               mach=(0.2, 0.75), alpha=(-4.0, 6.0), thick=(0.08, 0.16)) -> List[Case]:
    out = []
    for _ in range(n):
        m = float(rng.uniform(*mach))
        a = float(rng.uniform(*alpha))
        t = float(rng.uniform(*thick))
        nc, nr = sizes[int(rng.integers(len(sizes)))]
        out.append(Case(meshgen.naca_ogrid(nc, nr, t=t), m, a, t))
    return out


def solve_batch(cases: Sequence[Case], order: int, cfl: float, tol: float, max_iters: int,
                lhs: int = hip.LHS_FP32, check_every: int = 50, init=None):
    """Runs the cases as one batch to tol; returns (batch, iterations per case, steps, seconds)."""
    import time
    b = hip.Batch([(c.mesh, c.mach, c.alpha) for c in cases], order=order, cfl=cfl, lhs=lhs,
                  adi_mode=hip.ADI_FAST)
    if init is not None:
        init(b)
    b.sync()
    t0 = time.perf_counter()
    n, iters, norms = b.batch_step(max_iters, tol=tol, check_every=check_every)
    b.sync()
    return b, iters, n, time.perf_counter() - t0


def dataset(cases: Sequence[Case], batch: hip.Batch, per_case: int, rng: np.random.Generator, only=None):
    """Features and targets (primitive state minus the case's freestream) of a
    random subset of each case's cells (of the cases listed in only, if given)."""
    q = batch.get("q").reshape(-1, 4)
    xs, ts = [], []
    for k, c in enumerate(cases):
        if only is not None and k not in only:
            continue
        f = cell_features(c.mesh, c.mach, c.alpha, c.thickness)
        qk = q[batch.cell_off[k]:batch.cell_off[k + 1]]
        idx = rng.choice(c.mesh.n_cell, size=min(per_case, c.mesh.n_cell), replace=False)
        xs.append(f[idx])
        ts.append((primitive(qk[idx]) - freestream_primitive(c.mach, c.alpha)).astype(np.float32))
    return np.concatenate(xs), np.concatenate(ts)


def case_samples(case: Case, q: np.ndarray, per_case: int, rng: np.random.Generator):
    """Features and targets of per_case random cells of one converged field q."""
    f = cell_features(case.mesh, case.mach, case.alpha, case.thickness)
    q = np.asarray(q).reshape(-1, 4)
    idx = rng.choice(case.mesh.n_cell, size=min(per_case, case.mesh.n_cell), replace=False)
    t = (primitive(q[idx]) - freestream_primitive(case.mach, case.alpha)).astype(np.float32)
    return f[idx], t


def batch_features(cases: Sequence[Case]) -> np.ndarray:
    return np.concatenate([cell_features(c.mesh, c.mach, c.alpha, c.thickness) for c in cases])


def time_training(steps: int = 500, batch: int = 4096, graph: bool = True, n: int = 200000,
                  seed: int = 1) -> float:
    """Milliseconds per training step at `batch` on a synthetic data set of n
    rows (6 normal features, 4 smooth targets); 20 untimed steps first."""
    import time
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 6)).astype(np.float32)
    t = np.tanh(0.7 * x[:, :4]).astype(np.float32)
    mlp = MLP(seed=2)
    mlp.set_data(x, t)
    mlp.set_graph(graph)
    mlp.train(20, batch=batch, lr=1e-3, lr_final=1e-4, seed=1)
    t0 = time.perf_counter()
    mlp.train(steps, batch=batch, lr=1e-3, lr_final=1e-4, seed=3)
    ms = (time.perf_counter() - t0) / max(steps, 1) * 1e3
    mlp.close()
    return ms


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="time the warm-start network's training step")
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--eager", action="store_true", help="launch each step's kernels instead of a graph")
    a = ap.parse_args()
    print(f"{a.steps} steps at batch {a.batch}: {time_training(a.steps, a.batch, not a.eager):.3f} ms/step",
          flush=True)

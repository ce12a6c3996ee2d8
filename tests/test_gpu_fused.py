"""Iterations fused across their boundary: inside a captured block of
iterations, computeLinearRHS of one iteration (with the update q += dq) and
computeRHS("implicit") of the next run as one face kernel and one cell kernel
(k_lin_rhs_faces, k_lin_rhs_cells) for a single case with one sweep.  Each
value is computed by the same operations as unfused, so a run replaying
graphs (fused) equals the same run launched one iteration at a time
(unfused) bit for bit: the state, the residual of the last computeRHS, the
spectral radii and both norm histories; and EXACT's state stays bit-identical
to the oracle's over a fused block."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def hip():
    from hamstr_oop_amd import hip as H
    H.lib()
    return H


def _eq(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        raise AssertionError(f"{what}: {np.count_nonzero(a != b)} of {a.size} differ; max |diff| {d.max():.3e}")


def _pair(hip, mesh, order, lhs, mode, jsrc, steps, mach=0.5, alpha=2.0):
    out = []
    for graph in (True, False):
        s = hip.Solver(mesh, mach, alpha, order, 10.0, lhs=lhs, adi_mode=mode)
        if jsrc is not None:
            hip.set_jacobian_source(s, jsrc)
        s.set_graph(graph)
        n, l2r, l2l = s.step(steps, check_every=steps)
        assert n == steps
        out.append((s.get("q"), s.get("sigma"), s.get("r0"), l2r, l2l, s.get("r")))
        s.close()
    return out


@pytest.mark.parametrize("nc,nr", [(64, 24), (128, 40)])
@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("lhs,mode", [("fp64", "fast"), ("fp32", "fast"), ("fp64", "exact")])
def test_fused_block_equals_single_iterations(hip, order, lhs, mode, nc, nr):
    """O-grid with triangulated outer layers (orphan faces; phase 0's launch
    carries the other cells' residual sums), 24 iterations in one block; the
    loops' cells fill 4 blocks of the cell kernel (all blocks then sum the
    residuals too) or 16 (a multiple of 8: those blocks lead the launch)."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(nc, nr, tri_layers=8)
    print(f"fused {nc}x{nr} order={order} {lhs} {mode}", flush=True)
    lm = hip.LHS_FP64 if lhs == "fp64" else hip.LHS_FP32
    am = hip.ADI_FAST if mode == "fast" else hip.ADI_EXACT
    (qa, sa, ra, l2a, lla, xa), (qb, sb, rb, l2b, llb, xb) = _pair(hip, mesh, order, lm, am, None, 24)
    _eq(qa, qb, "q")
    _eq(xa, xb, "linear residual r")
    _eq(sa, sb, "sigma")
    _eq(ra, rb, "r0")
    _eq(l2a, l2b, "nonlinear L2 history")
    _eq(lla, llb, "linear L2 history")


@pytest.mark.parametrize("src", ["q", "states"])
def test_fused_block_with_computed_jacobians(hip, src):
    """The fused tail stores what each Jacobian source needs (orphan and
    chain-end faces' Jacobians, or the face states) for the next ADI."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(128, 48, tri_layers=16)
    js = {"q": hip.JAC_FROM_Q, "states": hip.JAC_FROM_STATES}[src]
    (qa, _, _, l2a, lla, _), (qb, _, _, l2b, llb, _) = _pair(hip, mesh, 3, hip.LHS_FP64, hip.ADI_FAST, js, 16)
    _eq(qa, qb, "q")
    _eq(l2a, l2b, "nonlinear L2 history")
    _eq(lla, llb, "linear L2 history")


def test_fused_block_without_cell_split(hip):
    """A mesh whose phase 0 is not one launch of the one-row kernel over the
    leading cells: the fused cell kernel sums every cell's residual."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.circle_ogrid(48, 301)   # 301-row lines: two per workgroup
    (qa, sa, _, l2a, lla, _), (qb, sb, _, l2b, llb, _) = _pair(hip, mesh, 3, hip.LHS_FP64, hip.ADI_FAST, None, 12)
    _eq(qa, qb, "q")
    _eq(sa, sb, "sigma")
    _eq(l2a, l2b, "nonlinear L2 history")
    _eq(lla, llb, "linear L2 history")


def test_fused_exact_block_matches_the_oracle(hip):
    """EXACT over a fused block of 20 iterations against the oracle: q bit for
    bit, both norm histories to 1e-12 (the GPU sums the squares by blocks)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(32, 16, tri_layers=4)
    g = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_EXACT)
    n, l2g, llg = g.step(20, check_every=20)
    o = oracle_py.Oracle().case(mesh, 0.5, 2.0, 3, 10.0)
    _, l2o, llo, _ = o.step(20)
    _eq(g.get("q"), o.get("q"), "q")
    np.testing.assert_allclose(l2g, l2o, rtol=1e-12, atol=0)
    np.testing.assert_allclose(llg, llo, rtol=1e-12, atol=0)
    g.close()


def test_batches_and_sweeps_stay_unfused(hip):
    """Batches and msweep > 1 run unfused (equal results either way)."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.circle_ogrid(32, 16)
    out = []
    for graph in (True, False):
        s = hip.Solver(mesh, 0.3, 0.0, 3, 10.0, msweep=2, adi_mode=hip.ADI_FAST)
        s.set_graph(graph)
        s.step(8, check_every=8)
        out.append(s.get("q"))
        s.close()
    _eq(out[0], out[1], "msweep 2")

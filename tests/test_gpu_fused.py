"""Iterations fused across their boundary: inside a captured block of
iterations, computeLinearRHS of one iteration (with the update q += dq) and
computeRHS("implicit") of the next run as one face kernel and one cell kernel
(k_lin_rhs_faces, k_lin_rhs_cells) for a single case or a batch with one
sweep.  Each
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


def _batch_cases():
    from hamstr_oop_amd import meshgen
    return [(meshgen.circle_ogrid(32, 12), 0.3, 0.0), (meshgen.naca_ogrid(40, 16, tri_layers=6), 0.6, 1.0),
            (meshgen.naca_ogrid(64, 16), 0.5, -2.0), (meshgen.naca_ogrid(48, 20, t=0.10), 0.7, 3.0)]


@pytest.mark.parametrize("lhs", ["fp64", "fp32"])
def test_fused_batch_equals_single_iterations(hip, lhs):
    """A batch of uneven cases: blocks of 10 iterations replayed from graphs
    (fused) against the same iterations launched one at a time, to a tolerance
    some cases reach: states, iterations to tolerance and final per-case norms
    bit for bit (each stage's per-case norms taken before the fused tail
    overwrites its residuals)."""
    lm = hip.LHS_FP64 if lhs == "fp64" else hip.LHS_FP32
    out = []
    for graph in (True, False):
        b = hip.Batch(_batch_cases(), 3, 10.0, lhs=lm, adi_mode=hip.ADI_FAST)
        b.set_graph(graph)
        n, it, nrm = b.batch_step(200, tol=1e-6, check_every=10)
        out.append((n, it.copy(), nrm.copy(), [b.case_q(k).copy() for k in range(b.n_case)]))
        b.close()
    (na, ia, ra, qa), (nb, ib, rb, qb) = out
    assert na == nb
    _eq(ia, ib, "iterations to tolerance")
    assert (ia > 0).any()  # some case reaches the tolerance
    _eq(ra, rb, "final norms")
    for k in range(len(qa)):
        _eq(qa[k], qb[k], f"case {k}")


def test_fused_batch_with_retirement(hip):
    """Retired cases drop out of the fused kernels as of every other one: a
    batch retiring converged and stalled cases, graphs against one iteration
    at a time."""
    out = []
    for graph in (True, False):
        b = hip.Batch(_batch_cases(), 3, 10.0, adi_mode=hip.ADI_FAST)
        b.set_graph(graph)
        b.set_retire(True, 40)
        n, it, nrm = b.batch_step(300, tol=1e-7, check_every=10)
        st, its = b.case_status()
        out.append((n, it.copy(), nrm.copy(), st.copy(), its.copy(), [b.case_q(k).copy() for k in range(b.n_case)]))
        b.close()
    (na, ia, ra, sa, ta, qa), (nb, ib, rb, sb, tb, qb) = out
    assert na == nb
    _eq(ia, ib, "iterations to tolerance")
    _eq(sa, sb, "case status")
    _eq(ta, tb, "retirement iterations")
    _eq(ra, rb, "final norms")
    for k in range(len(qa)):
        _eq(qa[k], qb[k], f"case {k}")


def test_batches_and_sweeps_stay_unfused(hip):
    """msweep > 1 runs unfused (equal results either way)."""
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


def _run(hip, mesh, order, lhs, staged, steps=24, graph=True):
    s = hip.Solver(mesh, 0.5, 2.0, order, 10.0, lhs=lhs, adi_mode=hip.ADI_FAST)
    if staged is not None:
        hip.set_face_staging(s, staged)
    s.set_graph(graph)
    n, l2r, l2l = s.step(steps, check_every=steps)
    assert n == steps
    out = (s.get("q"), s.get("r"), s.get("sigma"), s.get("r0"), l2r, l2l)
    s.close()
    return out


@pytest.mark.parametrize("graph", [True, False])
@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("lhs", ["fp64", "fp32"])
def test_staged_faces_equal_gathered(hip, order, lhs, graph):
    """The face kernels (fused in graphs, k_rhs_faces and k_lin_faces one
    iteration at a time) with each workgroup's stencil rows staged in LDS
    (the default on an O-grid) and with every face gathering its own give
    the same iterations bit for bit."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(64, 24, tri_layers=8)
    lm = hip.LHS_FP64 if lhs == "fp64" else hip.LHS_FP32
    s = hip.Solver(mesh, 0.5, 2.0, order, 10.0, lhs=lm, adi_mode=hip.ADI_FAST)
    assert hip.face_staging(s)
    s.close()
    print(f"staged/gathered order={order} {lhs} graph={graph}", flush=True)
    a = _run(hip, mesh, order, lm, True, graph=graph)
    b = _run(hip, mesh, order, lm, False, graph=graph)
    for x, y, what in zip(a, b, ("q", "r", "sigma", "r0", "L2 history", "linear L2 history")):
        _eq(np.asarray(x), np.asarray(y), what)


def test_staged_faces_equal_gathered_batch(hip):
    from hamstr_oop_amd import meshgen
    cases = [(meshgen.naca_ogrid(32, 12), 0.8, 1.25), (meshgen.naca_ogrid(64, 16, t=0.14, tri_layers=4), 0.6, -3.0)]
    qs = []
    for staged in (True, False):
        b = hip.Batch(cases, order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
        hip.set_face_staging(b, staged)
        n, _, norms = b.batch_step(16, check_every=16)
        assert n == 16
        qs.append((b.get("q"), norms))
        b.close()
    _eq(qs[0][0], qs[1][0], "q")
    _eq(qs[0][1], qs[1][1], "per-case norms")


def test_staging_refused_where_the_slots_do_not_fit(hip):
    """Radial lines of 3 cells at order 5: a workgroup of 256 line faces
    spans 576 slots (64 lines of 9), beyond the 520 the kernel stages; the
    fused iteration then gathers, still equal to the unfused one."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(128, 3)
    s = hip.Solver(mesh, 0.5, 2.0, 5, 10.0, adi_mode=hip.ADI_FAST)
    assert not hip.face_staging(s)
    with pytest.raises(hip.HamstrError):
        hip.set_face_staging(s, True)
    s.close()
    s = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_FAST)
    assert hip.face_staging(s)
    s.close()
    a = _run(hip, mesh, 5, hip.LHS_FP64, None, steps=16, graph=True)
    b = _run(hip, mesh, 5, hip.LHS_FP64, None, steps=16, graph=False)
    for x, y, what in zip(a, b, ("q", "r", "sigma", "r0", "L2 history", "linear L2 history")):
        _eq(np.asarray(x), np.asarray(y), what)


def test_staged_faces_with_retired_cases(hip):
    """Cases of a batch that converge retire: workgroups whose faces all
    belong to them stage nothing and exit, the others stage their whole slot
    range; every case's iterations and state equal the gathered run's."""
    from hamstr_oop_amd import meshgen
    cases = [(meshgen.naca_ogrid(32, 12), 0.8, 1.25), (meshgen.naca_ogrid(64, 16, t=0.14), 0.6, -3.0),
             (meshgen.naca_ogrid(48, 16, tri_layers=4), 0.3, 1.0)]
    out = []
    for staged in (True, False):
        b = hip.Batch(cases, order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
        hip.set_face_staging(b, staged)
        n, iters, _ = b.batch_step(3000, tol=1e-9, check_every=20)
        out.append((iters.copy(), b.get("q")))
        b.close()
    assert (out[0][0] > 0).all() and len(set(out[0][0].tolist())) > 1, out[0][0]
    _eq(out[0][0], out[1][0], "iterations to tolerance")
    _eq(out[0][1], out[1][1], "q")

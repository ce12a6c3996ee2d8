"""Batched contexts: many flow cases in one set of launches.

A case inside a batch runs the same arithmetic as alone, so its state must be
bit-identical to a single-case context (EXACT and FAST alike), and the
per-case norms and convergence iterations must agree.  The autotuner only
changes launch shapes, so results must not change either.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from hamstr_oop_amd import hip as H
    H.lib()
    return H


def _cases():
    from hamstr_oop_amd import meshgen
    return [
        (meshgen.circle_ogrid(16, 8), 0.3, 0.0),
        (meshgen.naca_ogrid(32, 12), 0.8, 1.25),
        (meshgen.naca_ogrid(24, 10, t=0.10), 0.5, 2.0),   # 24-row loops: odd final subsystems
        (meshgen.naca_ogrid(64, 16, t=0.14), 0.6, -3.0),
    ]


@pytest.mark.parametrize("mode", ["exact", "fast", "fast32"])
def test_batch_matches_single_cases(hip, mode):
    adi = hip.ADI_EXACT if mode == "exact" else hip.ADI_FAST
    lhs = hip.LHS_FP32 if mode == "fast32" else hip.LHS_FP64
    cases = _cases()
    b = hip.Batch(cases, order=3, cfl=10.0, lhs=lhs, adi_mode=adi)
    n, iters, norms = b.batch_step(12, check_every=12)
    assert n == 12
    for k, (mesh, mach, alpha) in enumerate(cases):
        s = hip.Solver(mesh, mach, alpha, 3, 10.0, lhs=lhs, adi_mode=adi)
        _, l2, l2l = s.step(12, check_every=12)
        assert np.array_equal(b.case_q(k).ravel(), s.get("q")), f"case {k}"
        assert norms[k, 0] == pytest.approx(l2[-1], rel=1e-12)
        assert norms[k, 2] == pytest.approx(l2l[-1], rel=1e-12)
        s.close()


def test_batch_convergence_iterations(hip):
    cases = _cases()
    tol = 1e-9
    b = hip.Batch(cases, order=1, cfl=10.0, adi_mode=hip.ADI_FAST)
    n, iters, _ = b.batch_step(4000, tol=tol, check_every=20)
    assert (iters > 0).all()
    for k, (mesh, mach, alpha) in enumerate(cases):
        s = hip.Solver(mesh, mach, alpha, 1, 10.0, adi_mode=hip.ADI_FAST)
        _, l2, _ = s.step(int(iters[k]) + 5, check_every=10)
        first = int(np.argmax(l2 <= tol)) + 1
        assert l2[first - 1] <= tol and first == iters[k], f"case {k}: {first} vs {iters[k]}"
        s.close()


def test_autotune_keeps_results(hip):
    """The autotuner changes persistent grids (the same arithmetic) and the
    cyclic-reduction levels of the FAST solve (the same systems, reduced in
    another order): the states agree to rounding."""
    cases = _cases()
    b1 = hip.Batch(cases, order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
    b2 = hip.Batch(cases, order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
    b2.step(2)  # a state to tune on
    groups, best = hip.autotune(b2)
    assert all(g["grid"] >= 1 and 0 <= g["cr_levels"] <= 3 for g in groups if g["fast"])
    b2.initFlow()
    b1.batch_step(8, check_every=8)
    b2.batch_step(8, check_every=8)
    q1, q2 = b1.get("q").reshape(-1, 4), b2.get("q").reshape(-1, 4)
    assert (np.abs(q1 - q2) / np.abs(q1).max(axis=0)).max() < 1e-11
    # grids alone do not change the arithmetic
    b3 = hip.Batch(cases, order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
    for k, g in enumerate(hip.launch_groups(b3)):
        if g["fast"]:
            hip.set_launch(b3, k, max(1, g["chains"] // 3), g["cr_levels"])
    b3.batch_step(8, check_every=8)
    assert np.array_equal(b3.get("q"), q1.ravel())


@pytest.mark.parametrize("cr", [0, 1, 2, 3])
def test_cyclic_reduction_levels_match_exact(hip, cr):
    """Every number of cyclic-reduction levels before the PCR solves the same
    block systems as the reference's sequential order, to rounding: open
    chains, two-row cycles (16 and 64 rows) and odd cycles (24 rows)."""
    cases = _cases()
    fast = hip.Batch(cases, order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
    exact = hip.Batch(cases, order=3, cfl=10.0, adi_mode=hip.ADI_EXACT)
    for k, g in enumerate(hip.launch_groups(fast)):
        if g["fast"]:
            hip.set_launch(fast, k, g["grid"], cr)
    for b in (fast, exact):
        b.computeRHS("implicit")
        b.ADI()
    xe, xf = exact.get("dq").reshape(-1, 4), fast.get("dq").reshape(-1, 4)
    scale = np.abs(xe).max(axis=0)
    assert (np.abs(xf - xe) / scale).max() < 1e-11


@pytest.mark.parametrize("mode", ["exact", "fast", "fast32"])
def test_graph_replays_match_eager_launches(hip, mode):
    """Captured graphs replay the same kernels in the same order: every replay,
    with eager work and synchronisations between them, is bit-identical."""
    adi = hip.ADI_EXACT if mode == "exact" else hip.ADI_FAST
    lhs = hip.LHS_FP32 if mode == "fast32" else hip.LHS_FP64
    cases = _cases()
    eager = hip.Batch(cases, order=3, cfl=10.0, lhs=lhs, adi_mode=adi)
    eager.set_graph(False)
    graph = hip.Batch(cases, order=3, cfl=10.0, lhs=lhs, adi_mode=adi)
    for rep in range(4):
        ne, _, norms_e = eager.batch_step(30, check_every=10)
        ng, _, norms_g = graph.batch_step(30, check_every=10)   # 3 replays of a 10-iteration graph
        assert ne == ng == 30
        assert np.array_equal(eager.get("q"), graph.get("q")), f"replay block {rep}"
        assert np.array_equal(norms_e, norms_g)
        graph.get("r")   # eager work on the stream between replays
    # single case, graph of 50 iterations with an eager tail of 7
    mesh, mach, alpha = cases[1]
    a = hip.Solver(mesh, mach, alpha, 3, 10.0, lhs=lhs, adi_mode=adi)
    a.set_graph(False)
    b = hip.Solver(mesh, mach, alpha, 3, 10.0, lhs=lhs, adi_mode=adi)
    _, l2a, la = a.step(107, check_every=50)
    _, l2b, lb = b.step(107, check_every=50)
    assert np.array_equal(l2a, l2b) and np.array_equal(la, lb)
    assert np.array_equal(a.get("q"), b.get("q"))


def test_retired_cases_freeze_at_their_read_back(hip):
    """A converged case drops out of the launches at the read-back where it
    converged: its state equals a single-case run that stops there."""
    cases = _cases()
    tol, every = 1e-8, 25
    b = hip.Batch(cases, order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
    b.set_retire(True)
    n, iters, _ = b.batch_step(6000, tol=tol, check_every=every)
    st, it = b.case_status()
    assert (st == 1).all(), st
    assert (it == iters).all()
    assert n == max(int(np.ceil(i / every)) * every for i in iters)
    for k, (mesh, mach, alpha) in enumerate(cases):
        s = hip.Solver(mesh, mach, alpha, 3, 10.0, adi_mode=hip.ADI_FAST)
        stop = int(np.ceil(iters[k] / every)) * every
        s.step(stop, check_every=every)
        assert np.array_equal(b.case_q(k).ravel(), s.get("q")), f"case {k}"
        s.close()
    # the frozen state stays frozen, and set_retire makes every case live again
    q0 = b.get("q")
    n2, _, _ = b.batch_step(50, tol=tol, check_every=every)
    assert n2 == every and np.array_equal(b.get("q"), q0)
    b.set_retire(False)
    b.batch_step(5, check_every=5)
    assert not np.array_equal(b.get("q"), q0)


def test_stalled_cases_retire(hip):
    """A tolerance no case can reach: each case's norm stops halving at the
    rounding floor, it is found stalled and the call returns early."""
    from hamstr_oop_amd import meshgen
    cases = [(meshgen.circle_ogrid(16, 8), 0.3, 0.0), (meshgen.naca_ogrid(24, 10, t=0.10), 0.5, 2.0)]
    b = hip.Batch(cases, order=1, cfl=10.0, adi_mode=hip.ADI_FAST)
    b.set_retire(True, stall_window=300)
    n, iters, norms = b.batch_step(20000, tol=1e-40, check_every=50)
    st, it = b.case_status()
    assert (st == 2).all(), st
    assert n < 20000 and (it <= n).all()
    assert (norms[:, 0] < 1e-12).all()


def test_diverged_case_retires_without_stopping_the_batch(hip):
    """A case whose state has no real sound speed (negative pressure) makes a
    NaN spectral radius: with retirement on, only that case retires (diverged);
    the other case runs on exactly as alone.  Without retirement the call
    fails with HAMSTR_ERR_NAN, as the reference exits."""
    cases = _cases()[:2]
    b = hip.Batch(cases, order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
    q1 = b.case_q(1)
    q1[:, 3] = -1.0     # E < kinetic energy: p < 0
    b.set_case_q(1, q1)
    b.set_retire(True)
    n, iters, _ = b.batch_step(40, tol=1e-30, check_every=10)
    st, it = b.case_status()
    assert st[1] == 3 and it[1] == 10
    assert st[0] == 0 and n == 40
    mesh, mach, alpha = cases[0]
    s = hip.Solver(mesh, mach, alpha, 3, 10.0, adi_mode=hip.ADI_FAST)
    s.step(40, check_every=10)
    assert np.array_equal(b.case_q(0).ravel(), s.get("q"))
    b2 = hip.Batch(cases, order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
    b2.set_case_q(1, q1)
    with pytest.raises(hip.HamstrError) as e:
        b2.batch_step(10, check_every=10)
    assert e.value.code == hip.ERR_NAN


def test_fp32_lhs_reaches_the_fp64_fixed_point(hip):
    """The converged state depends only on the FP64 residual: driven to a deep
    tolerance, FP32-LHS and FP64-LHS runs agree to far below the FP32 rounding
    of the LHS itself."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(64, 32, t=0.12)
    qs = {}
    for lhs in (hip.LHS_FP64, hip.LHS_FP32):
        s = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, lhs=lhs, adi_mode=hip.ADI_FAST)
        n, l2, _ = s.step(40000, tol=1e-13, check_every=100)
        assert l2[-1] <= 1e-13 or (l2 <= 1e-13).any(), l2[-1]
        qs[lhs] = s.get("q").reshape(-1, 4)
        s.close()
    scale = np.abs(qs[hip.LHS_FP64]).max(axis=0)
    d = (np.abs(qs[hip.LHS_FP32] - qs[hip.LHS_FP64]) / scale).max()
    print(f"FP32-LHS vs FP64-LHS at L2 <= 1e-13: max rel diff {d:.2e}", flush=True)
    assert d < 1e-6


def test_held_case_resumes_exactly(hip):
    """A case held out of the launches keeps its state; made live again it
    continues exactly as if it had never been held, and the other cases run
    on meanwhile as alone."""
    cases = _cases()[:2]
    b = hip.Batch(cases, order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
    b.set_retire(True)
    b.set_case_live(1, False)
    b.batch_step(20, tol=1e-30, check_every=10)
    st, _ = b.case_status()
    assert list(st) == [0, 4]
    b.set_case_live(1, True)
    b.batch_step(20, tol=1e-30, check_every=10)
    assert b.iterations() == 40
    for k, steps in ((0, 40), (1, 20)):
        mesh, mach, alpha = cases[k]
        s = hip.Solver(mesh, mach, alpha, 3, 10.0, adi_mode=hip.ADI_FAST)
        s.step(steps, check_every=10)
        assert np.array_equal(b.case_q(k).ravel(), s.get("q")), k
        s.close()


@pytest.mark.parametrize("lhs", ["fp32", "fp64"])
def test_time_batch_iteration_runs(hip, lhs):
    """The batch timing entry (python -m hamstr_oop_amd.sweep) on a few drawn
    cases: every stage timed, one launch group per ADI phase at least."""
    from hamstr_oop_amd import sweep
    out = sweep.time_batch_iteration(6, hip.LHS_FP32 if lhs == "fp32" else hip.LHS_FP64, warmup=2, iters=3)
    assert out["cases"] == 6 and out["lhs"] == lhs and out["ms_per_iter"] > 0
    assert {"rhs_faces", "adi_phase0", "adi_phase1", "lin_faces"} <= set(out["stage_ms"])
    assert {g["phase"] for g in out["groups"]} == {0, 1}


@pytest.mark.parametrize("lhs", ["fp64", "fp32"])
def test_large_batch_matches_single_cases(hip, lhs):
    """72 cases of a 256x128 mesh (2.4M cells, 4.7M faces, 18,432 radial
    lines): cases stay bit-identical to their single-case contexts at a size
    where every launch of the iteration spans the whole chip many times."""
    from hamstr_oop_amd import meshgen
    lh = hip.LHS_FP32 if lhs == "fp32" else hip.LHS_FP64
    mesh = meshgen.naca_ogrid(256, 128, t=0.12)
    machs = np.linspace(0.3, 0.6, 72)
    b = hip.Batch([(mesh, float(m), 1.0) for m in machs], order=3, cfl=10.0, lhs=lh, adi_mode=hip.ADI_FAST)
    assert int(b.cell_off[-1]) > 8192 * 256
    n, _, norms = b.batch_step(6, check_every=6)
    assert n == 6
    for k in (0, 37, 71):
        s = hip.Solver(mesh, float(machs[k]), 1.0, 3, 10.0, lhs=lh, adi_mode=hip.ADI_FAST)
        _, l2, _ = s.step(6, check_every=6)
        assert np.array_equal(b.case_q(k).ravel(), s.get("q")), k
        assert norms[k, 0] == pytest.approx(l2[-1], rel=1e-12), k  # per-case and block-partial sums
        s.close()
    b.close()


def test_batch_step_rejects_negative_steps(hip):
    """A negative step count returns HAMSTR_ERR_ARG before any launch, with the
    per-case outputs still written (no case converged, nothing run), and the
    context keeps working afterwards."""
    import ctypes as C
    b = hip.Batch(_cases()[:2], order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
    iters = np.full(b.n_case, 7, np.int32)
    norms = np.zeros((b.n_case, 4))
    done = C.c_int(5)
    rc = hip.lib().hamstr_batch_step(b.h, -1, C.c_double(0.0), 1, hip._i(iters), hip._d(norms), C.byref(done))
    assert rc == hip.ERR_ARG
    assert done.value == 0 and np.array_equal(iters, np.full(b.n_case, -1))
    n, it, _ = b.batch_step(3, check_every=3)
    assert n == 3 and np.array_equal(it, np.full(b.n_case, -1))
    b.close()

"""GPU parity of stepSolution's explicit branches and of updateSolution
(march.cpp:73-108, 142-164) against the golden vectors the reference wrote and
the CPU oracle.  q, qt, r and sigma are bit-identical; the density-residual
norms are summed in another order on the device and compared to 1e-12
relative.  The explicit schemes compute no Jacobians, so the LHS precision and
the ADI mode must not change any of their results.
"""
import glob
import os

import numpy as np
import pytest

from cases import SMALL_CASES, GoldenMesh, perturbed_state

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CASE_FILES = sorted(glob.glob(os.path.join(HERE, "golden", "case_*.npz")))


@pytest.fixture(scope="module")
def hip():
    from hamstr_oop_amd import hip as H
    H.lib()
    return H


def _eq(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        i = int(np.argmax(d))
        raise AssertionError(f"{what}: {np.count_nonzero(a != b)} of {a.size} differ; max |diff| {d.flat[i]:.3e}")


@pytest.mark.parametrize("check_every", [1, 5])
@pytest.mark.parametrize("scheme", ["euler", "rk3"])
@pytest.mark.parametrize("path", CASE_FILES, ids=[os.path.basename(p)[5:-4] for p in CASE_FILES])
def test_explicit_golden(hip, path, scheme, check_every):
    """5 steps from the fixture's perturbed state; check_every 5 replays a
    captured graph, 1 launches each kernel from the host."""
    d = np.load(path)
    print(f"explicit golden {os.path.basename(path)} {scheme} check_every={check_every}", flush=True)
    g = hip.Solver(GoldenMesh(d), float(d["mach"]), float(d["alpha"]), int(d["order"]), float(d["explicit_cfl"]),
                   adi_mode=hip.ADI_FAST)
    g.set_scheme(scheme)
    assert g.scheme == scheme
    g.set("q", d["q0"])
    n, first, printed = g.step(5, check_every=check_every)
    assert n == 5
    np.testing.assert_allclose(printed, d[f"{scheme}5_l2"], rtol=1e-12)
    for k in ("q", "qt", "r", "sigma"):
        _eq(g.get(k), d[f"{scheme}5_{k}"], k)
    if scheme == "euler":
        np.testing.assert_array_equal(first, printed)
    g.close()


@pytest.mark.parametrize("lhs", [0, 1])
@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("scheme", ["euler", "rk3"])
def test_explicit_matches_oracle(hip, oracle, scheme, order, lhs):
    """40 steps from freestream on a NACA O-grid at CFL 0.8, graph-replayed;
    the FP32 LHS setting must not matter."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(48, 16)
    print(f"explicit vs oracle {scheme} order={order} lhs={lhs}", flush=True)
    g = hip.Solver(mesh, 0.6, 2.0, order, 0.8, lhs=lhs, adi_mode=hip.ADI_FAST)
    g.set_scheme(scheme)
    o = oracle.case(mesh, 0.6, 2.0, order, 0.8)
    n, first, printed = g.step(40, check_every=20)
    assert n == 40
    _, f0, p0, _ = o.step_scheme(scheme, 40)
    np.testing.assert_allclose(first, f0, rtol=1e-12)
    np.testing.assert_allclose(printed, p0, rtol=1e-12)
    for k in ("q", "qt", "r", "sigma"):
        _eq(g.get(k), o.get(k), f"{k} after 40 steps")
    g.close()


def test_update_solution(hip, oracle):
    """updateSolution on its own after computeRHS("explicit"), between q and
    qt in every direction."""
    make, mach, alpha = SMALL_CASES["naca32x12"]
    mesh = make()
    g = hip.Solver(mesh, mach, alpha, 3, 1.3)
    o = oracle.case(mesh, mach, alpha, 3, 1.3)
    rng = np.random.default_rng(11)
    q = perturbed_state(o, rng)
    g.set("q", q)
    o.set("q", q)
    g.computeRHS("explicit")
    o.compute_rhs(False)
    for src, dst, coef in (("q", "qt", 0.25), ("q", "q", 8.0 / 15), ("qt", "q", 5.0 / 12), ("qt", "qt", 0.75)):
        g.updateSolution(src, dst, coef)
        o.update_solution(src, dst, coef)
        for k in ("q", "qt"):
            _eq(g.get(k), o.get(k), f"{k} after updateSolution({src}, {dst}, {coef})")
    with pytest.raises(hip.HamstrError):
        g.updateSolution("q", "r", 1.0)
    g.close()


def test_scheme_switch(hip, oracle):
    """rk3 steps then ADI steps in one context (each with its own captured
    graph), the same sequence as the oracle, EXACT mode: bit-identical."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.circle_ogrid(32, 12)
    g = hip.Solver(mesh, 0.4, 1.0, 3, 0.8, adi_mode=hip.ADI_EXACT)
    o = oracle.case(mesh, 0.4, 1.0, 3, 0.8)
    for scheme, n in (("rk3", 10), ("ADI", 10), ("euler", 5), ("rk3", 5)):
        g.set_scheme(scheme)
        done, first, printed = g.step(n, check_every=5)
        assert done == n
        _, f0, p0, _ = o.step_scheme(scheme, n)
        np.testing.assert_allclose(first, f0, rtol=1e-12)
        np.testing.assert_allclose(printed, p0, rtol=1e-12)
        _eq(g.get("q"), o.get("q"), f"q after {scheme}")
    g.close()


@pytest.mark.parametrize("scheme", ["euler", "rk3"])
def test_batch_explicit(hip, oracle, scheme):
    """Three cases of different meshes and freestreams in one batch, 12 steps
    (graph of 4 between read-backs): each case bit-identical to its own oracle
    run, per-case norms to 1e-12."""
    specs = list(SMALL_CASES.values())
    meshes = [mk() for mk, _, _ in specs]
    b = hip.Batch([(m, mach, alpha) for m, (_, mach, alpha) in zip(meshes, specs)], order=3, cfl=0.8,
                  adi_mode=hip.ADI_FAST)
    b.set_scheme(scheme)
    done, _, norms = b.batch_step(12, check_every=4)
    assert done == 12
    for k, (m, (_, mach, alpha)) in enumerate(zip(meshes, specs)):
        o = oracle.case(m, mach, alpha, 3, 0.8)
        _, f0, p0, _ = o.step_scheme(scheme, 12)
        _eq(b.case_q(k).reshape(-1), o.get("q"), f"case {k} q")
        assert norms[k, 0] == pytest.approx(f0[-1], rel=1e-12)
        assert norms[k, 2] == pytest.approx(p0[-1], rel=1e-12)
    b.close()


@pytest.mark.parametrize("scheme", ["euler", "rk3"])
def test_cpp_driver_explicit(hip, oracle, tmp_path, scheme):
    """The C++ drop-in driver reading scheme= from input.hamstr: final state
    bit-identical to the oracle and the printed L_2 column equal to the
    reference's (the last computeRHS's norm)."""
    import subprocess
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(32, 12)
    meshgen.write_reference_case(mesh, str(tmp_path), 0.5, 1.0, 3, 8, 0.8, scheme=scheme)
    exe = os.path.join(os.path.dirname(HERE), "hamstr_oop_amd", "hamstr_mi355x")
    out = subprocess.run([exe, "--dump", "q.bin"], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = [ln for ln in out.stdout.splitlines() if ln.strip().startswith("|") and "Iter" not in ln]
    l2_printed = np.array([float(ln.split("|")[2]) for ln in rows])
    q = np.fromfile(str(tmp_path / "q.bin"), dtype=np.float64)
    o = oracle.case(mesh, 0.5, 1.0, 3, 0.8)
    _, _, printed, _ = o.step_scheme(scheme, 8)
    _eq(q, o.get("q"), "driver q after 8 steps")
    np.testing.assert_allclose(l2_printed, np.array([float(f"{v:e}") for v in printed]), rtol=1e-6)

"""FAST ADI on chains longer than a workgroup (more than 512 rows in FP64,
1024 in FP32): two rows per thread, the first reduction level's odd rows in
global scratch.  Against EXACT (the reference's block Thomas order) to
rounding, on closed loops of every final-solve type and on open lines."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from hamstr_oop_amd import hip as H
    H.lib()
    return H


CASES = {
    # name: (mesh factory, lhs, order)
    "fp64_loops600_odd": (lambda m: m.circle_ogrid(600, 5), 0, 3),
    "fp64_loops1024_pair": (lambda m: m.circle_ogrid(1024, 4), 0, 3),
    "fp64_lines520_open": (lambda m: m.circle_ogrid(12, 520, ratio=1.004), 0, 3),
    "fp64_weno_loops640": (lambda m: m.naca_ogrid(640, 8, tri_layers=2), 0, 5),
    "fp32_loops1100_odd": (lambda m: m.circle_ogrid(1100, 4), 1, 3),
}


@pytest.mark.parametrize("name", list(CASES))
def test_long_chains_fast_matches_exact(hip, name):
    from hamstr_oop_amd import meshgen
    mk, lhs, order = CASES[name]
    mesh = mk(meshgen)
    print(f"long chains {name}: {mesh.n_cell} cells", flush=True)
    s = hip.Solver(mesh, 0.4, 1.0, order, 10.0, lhs=lhs, adi_mode=hip.ADI_FAST)
    groups = hip.launch_groups(s)
    assert any(g["fast"] and g["max_rows"] > (1024 if lhs else 512) for g in groups), groups
    e = hip.Solver(mesh, 0.4, 1.0, order, 10.0, lhs=lhs, adi_mode=hip.ADI_EXACT)
    _, l2f, _ = s.step(12, check_every=4)
    _, l2e, _ = e.step(12, check_every=4)
    qf, qe = s.get("q").reshape(-1, 4), e.get("q").reshape(-1, 4)
    scale = np.abs(qe).max(axis=0)
    tol = 1e-5 if lhs else 1e-10
    assert (np.abs(qf - qe) / scale).max() < tol
    np.testing.assert_allclose(l2f, l2e, rtol=1e-3 if lhs else 1e-8)
    # the cyclic-reduction depth of the long groups does not change the answer beyond rounding
    for k, g in enumerate(groups):
        if g["fast"] and g["max_rows"] > (1024 if lhs else 512):
            hip.set_launch(s, k, g["grid"], 3)
    _, l2g, _ = s.step(4, check_every=4)
    _, l2h, _ = e.step(4, check_every=4)
    np.testing.assert_allclose(l2g, l2h, rtol=1e-3 if lhs else 1e-8)
    s.close()
    e.close()

"""GPU parity: the HIP path through the C ABI against the CPU oracle.

FP64 results must be bit-identical to the oracle (which is itself
bit-identical to the reference, tests/test_oracle_golden.py), except the
residual norms, whose sums run in a different order on the GPU: those are
compared to 1e-12 relative.  ``np.array_equal`` treats +0 and -0 as equal,
the only difference the device's dropped exact-zero terms can make.
"""
import os

import numpy as np
import pytest

from cases import SMALL_CASES, perturbed_state, random_block_system, random_normals, random_states

pytestmark = pytest.mark.gpu

THIRD = 0.3333333333333333


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
        rel = d.flat[i] / max(abs(b.flat[i]), 1e-300)
        raise AssertionError(f"{what}: {np.count_nonzero(a != b)} of {a.size} differ; "
                             f"max |diff| {d.flat[i]:.3e} (rel {rel:.3e}) at {i}")


def test_flux_roe2d_bitwise(hip, oracle):
    rng = np.random.default_rng(1)
    n = 4096
    nv = random_normals(rng, n)
    ql = random_states(rng, n)
    qr = random_states(rng, n)
    flux, sr = hip.flux_roe2d(nv, ql, qr)
    for i in range(0, n, 97):
        f0, s0 = oracle.flux_roe2d(nv[i], ql[i], qr[i])
        _eq(flux[i], f0, f"flux[{i}]")
        _eq(sr[i], s0, f"sr[{i}]")


def test_jacobian_roe2d_bitwise_fp64(hip, oracle):
    rng = np.random.default_rng(2)
    n = 2048
    nv = random_normals(rng, n)
    ql = random_states(rng, n)
    qr = random_states(rng, n)
    lm, rm = hip.jacobian_roe2d(nv, ql, qr)
    for i in range(0, n, 41):
        l0, r0 = oracle.jacobian_roe2d(nv[i], ql[i], qr[i])
        _eq(lm[i], l0, f"lmat[{i}]")
        _eq(rm[i], r0, f"rmat[{i}]")


def test_jacobian_roe2d_fp32_close(hip, oracle):
    rng = np.random.default_rng(3)
    n = 1024
    nv = random_normals(rng, n)
    ql = random_states(rng, n)
    qr = random_states(rng, n)
    lm, rm = hip.jacobian_roe2d(nv, ql, qr, lhs=hip.LHS_FP32)
    for i in range(0, n, 17):
        l0, r0 = oracle.jacobian_roe2d(nv[i], ql[i], qr[i])
        scale = max(np.abs(l0).max(), np.abs(r0).max())
        assert np.abs(lm[i] - l0).max() <= 2e-5 * scale
        assert np.abs(rm[i] - r0).max() <= 2e-5 * scale


@pytest.mark.parametrize("kind", ["muscl1", "muscl3", "weno", "muscl_deriv1", "muscl_deriv3", "weno_deriv"])
def test_reconstruction_bitwise(hip, oracle, kind):
    rng = np.random.default_rng(4)
    nch, imax = 7, 23
    f = rng.uniform(0.5, 2.0, (nch, imax, 4))
    dq = rng.uniform(-0.1, 0.1, (nch, imax, 4))
    if kind.startswith("muscl"):
        qt = 0.0 if kind.endswith("1") else 0.25
        is_, ie = 1, imax - 1
        if kind.startswith("muscl_deriv"):
            ql, qr = hip.muscl_deriv(f, dq, is_, ie, THIRD, qt)
        else:
            ql, qr = hip.muscl(f, is_, ie, THIRD, qt)
        rows = slice(is_, ie + 1) if qt == 0.0 else slice(is_, ie)
    else:
        is_, ie = 2, imax - 1
        if kind == "weno_deriv":
            ql, qr = hip.weno_deriv(f, dq, is_, ie)
        else:
            ql, qr = hip.weno(f, is_, ie)
        rows = slice(is_, imax - 2)
    for c in range(nch):
        if kind.startswith("muscl_deriv"):
            l0, r0 = oracle.muscl_deriv(f[c], dq[c], is_, ie, THIRD, qt, 1e-10)
        elif kind.startswith("muscl"):
            l0, r0 = oracle.muscl(f[c], is_, ie, THIRD, qt, 1e-10)
        elif kind == "weno_deriv":
            l0, r0 = oracle.weno_deriv(f[c], dq[c], is_, ie, 1e-10)
        else:
            l0, r0 = oracle.weno(f[c], is_, ie, 1e-10)
        _eq(ql[c, rows], l0[rows], f"{kind} ql chain {c}")
        _eq(qr[c, rows], r0[rows], f"{kind} qr chain {c}")


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("N", [3, 7, 64])
def test_block_tridag_bitwise(hip, oracle, periodic, N):
    rng = np.random.default_rng(5 + N)
    nsys = 5
    a, b, c, f = random_block_system(rng, nsys, N)
    x = hip.block_tridag4(a, b, c, f, periodic=periodic)
    for s in range(nsys):
        x0 = oracle.block_tridag4(a[s], b[s], c[s], f[s], periodic=periodic)
        _eq(x[s], x0, f"system {s}")


def _pair(hip, oracle, name, order, lhs=0):
    make, mach, alpha = SMALL_CASES[name]
    mesh = make()
    g = hip.Solver(mesh, mach, alpha, order, 10.0, lhs=lhs)
    o = oracle.case(mesh, mach, alpha, order, 10.0)
    return g, o


@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("name", list(SMALL_CASES))
def test_compute_rhs_bitwise(hip, oracle, name, order):
    g, o = _pair(hip, oracle, name, order)
    rng = np.random.default_rng(6)
    q = perturbed_state(o, rng)
    g.set("q", q)
    o.set("q", q)
    l2g, lig = g.computeRHS("implicit")
    _, l2o, lio = o.compute_rhs(True)
    _eq(g.get("r"), o.get("r"), "r")
    _eq(g.get("sigma"), o.get("sigma"), "sigma")
    _eq(g.get("ff"), o.get("ff"), "ff")
    assert l2g == pytest.approx(l2o, rel=1e-12)
    assert lig == lio


@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("name", list(SMALL_CASES))
def test_api_iterations_bitwise(hip, oracle, name, order):
    """computeRHS + ADI through the API, 6 iterations from a perturbed state."""
    g, o = _pair(hip, oracle, name, order)
    rng = np.random.default_rng(7)
    q = perturbed_state(o, rng)
    g.set("q", q)
    o.set("q", q)
    for it in range(6):
        l2g, _ = g.computeRHS("implicit")
        _, l2o, _ = o.compute_rhs(True)
        assert l2g == pytest.approx(l2o, rel=1e-12), f"nonlinear L2 at iteration {it}"
        lg, _ = g.ADI()
        lo, _ = o.adi()
        assert lg == pytest.approx(lo, rel=1e-12), f"linear L2 at iteration {it}"
        for fld in ("q", "dq", "r0", "r"):
            _eq(g.get(fld), o.get(fld), f"{fld} after iteration {it}")


@pytest.mark.parametrize("order", [1, 3])
def test_step_matches_oracle(hip, oracle, order):
    """The fused iteration (hamstr_step) from freestream, 20 steps, config-1 O-grid topology."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.circle_ogrid(32, 16)
    g = hip.Solver(mesh, 0.3, 0.0, order, 10.0)
    o = oracle.case(mesh, 0.3, 0.0, order, 10.0)
    n, l2g, lling = g.step(20)
    assert n == 20
    _, l2o, llino, _ = o.step(20)
    np.testing.assert_allclose(l2g, l2o, rtol=1e-12)
    np.testing.assert_allclose(lling, llino, rtol=1e-12)
    _eq(g.get("q"), o.get("q"), "q after 20 steps")


def test_compute_linear_rhs_bitwise(hip, oracle):
    g, o = _pair(hip, oracle, "naca32x12", 3)
    rng = np.random.default_rng(8)
    q = perturbed_state(o, rng)
    for s in (g, o):
        s.set("q", q)
    g.computeRHS("implicit")
    o.compute_rhs(True)
    dq = rng.uniform(-1e-3, 1e-3, 4 * o.n_cell)
    r0 = rng.uniform(-1e-3, 1e-3, 4 * o.n_cell)
    for s in (g, o):
        s.set("dq", dq)
        s.set("r0", r0)
    lg, _ = g.computeLinearRHS()
    lo, _ = o.compute_linear_rhs()
    assert lg == pytest.approx(lo, rel=1e-12)
    _eq(g.get("r"), o.get("r"), "linear residual")


# ---- FAST ADI ordering: equal to the exact ordering up to rounding ----------

def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("name", list(SMALL_CASES))
def test_fast_adi_matches_exact(hip, oracle, name, order):
    """One ADI call: FP64 FAST (parallel cyclic reduction) vs the exact order;
    FP32-LHS FAST vs FP64.  naca24x10 has 24-row loops, whose reduction ends in
    odd (3-row) cyclic subsystems; circle16x8 ends in two-row pairs."""
    make, mach, alpha = SMALL_CASES[name]
    mesh = make()
    rng = np.random.default_rng(9)
    o = oracle.case(mesh, mach, alpha, order, 10.0)
    q = perturbed_state(o, rng)
    out = {}
    for key, lhs, mode in (("exact", hip.LHS_FP64, hip.ADI_EXACT), ("fast", hip.LHS_FP64, hip.ADI_FAST),
                           ("fast32", hip.LHS_FP32, hip.ADI_FAST)):
        g = hip.Solver(mesh, mach, alpha, order, 10.0, lhs=lhs, adi_mode=mode)
        g.set("q", q)
        g.computeRHS("implicit")
        g.ADI()
        out[key] = g.get("dq")
        g.close()
    assert _rel(out["fast"], out["exact"]) < 1e-11
    assert _rel(out["fast32"], out["exact"]) < 1e-4


def test_fast_adi_trajectory(hip):
    """100 fused steps from freestream: FAST and EXACT states agree to 1e-10."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(48, 24)
    res = {}
    for mode in (hip.ADI_EXACT, hip.ADI_FAST):
        g = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, adi_mode=mode)
        n, l2, _ = g.step(100, check_every=100)
        res[mode] = (g.get("q"), l2)
        g.close()
    q0 = res[hip.ADI_EXACT][0].reshape(-1, 4)
    q1 = res[hip.ADI_FAST][0].reshape(-1, 4)
    scale = np.abs(q0).max(axis=0)
    assert (np.abs(q1 - q0) / scale).max() < 1e-10
    np.testing.assert_allclose(res[hip.ADI_FAST][1], res[hip.ADI_EXACT][1], rtol=1e-8)


def test_cpp_driver_matches_oracle(hip, oracle, tmp_path):
    """The C++ drop-in driver (reference file formats, SOLVER surface) in the
    EXACT mode: final state bit-identical to the oracle after 8 steps, and the
    printed L_2 column equal to the reference's printed (linear) norm."""
    import subprocess
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(32, 12)
    meshgen.write_reference_case(mesh, str(tmp_path), 0.8, 1.25, 3, 8, 10.0)
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hamstr_oop_amd", "hamstr_mi355x")
    out = subprocess.run([exe, "--exact", "--dump", "q.bin"], cwd=str(tmp_path), capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = [ln for ln in out.stdout.splitlines() if ln.strip().startswith("|") and "Iter" not in ln]
    l2_printed = np.array([float(ln.split("|")[2]) for ln in rows])
    q = np.fromfile(str(tmp_path / "q.bin"), dtype=np.float64)
    o = oracle.case(mesh, 0.8, 1.25, 3, 10.0)
    _, _, l2lin, _ = o.step(8)
    _eq(q, o.get("q"), "driver q after 8 steps")
    assert np.array_equal(l2_printed, np.array([float(f"{v:e}") for v in l2lin]))

"""The transonic WENO limit cycle belongs to the scheme, not to a kernel.

Configuration 3 (WENO, M=0.8, a shock) never reaches 1e-10: its nonlinear
norm settles into a limit cycle.  tests/golden/weno_cycle_oracle.npz holds
the oracle's own run of a reduced case of the same kind -- the 128x48
NACA0012 O-grid with 4 triangulated far-field layers, order 5, M=0.8,
alpha=1.25, CFL 10, 6,000 steps (tests/golden/make_oracle_runs.py).  The
oracle's norm cycles between 8.4e-9 and 2.1e-8 with a period of 23 steps
from step 2,000 on.

* EXACT is bit-identical to the oracle through the whole run, cycle included.
* FAST settles into the same band with the same period, with the FP64 and
  with the FP32 left-hand side.
"""
import hashlib
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def wc():
    from hamstr_oop_amd import meshgen
    d = np.load(os.path.join(GOLD, "weno_cycle_oracle.npz"))
    mesh = meshgen.naca_ogrid(int(d["nc"]), int(d["nr"]), tri_layers=int(d["tri_layers"]))
    return d, mesh


@pytest.fixture(scope="module")
def hip():
    from hamstr_oop_amd import hip as H
    H.lib()
    return H


def period(series):
    """The lag (5..400) of the strongest autocorrelation of log(series)."""
    x = np.log(series) - np.log(series).mean()
    ac = np.correlate(x, x, "full")[len(x) - 1:]
    ac = ac / ac[0]
    k = int(np.argmax(ac[5:400])) + 5
    return k, float(ac[k])


def _solver(hip, d, mesh, lhs, mode):
    return hip.Solver(mesh, float(d["mach"]), float(d["alpha"]), int(d["order"]), float(d["cfl"]), lhs=lhs,
                      adi_mode=mode)


def test_oracle_cycle(wc):
    d, _ = wc
    tail = d["l2rhs"][2000:]
    assert 8e-9 < tail.min() and tail.max() < 2.2e-8
    k, strength = period(tail)
    assert k == 23 and strength > 0.9


def test_exact_is_the_oracle_through_the_cycle(hip, wc):
    d, mesh = wc
    print("WENO cycle, EXACT, 6000 steps", flush=True)
    g = _solver(hip, d, mesh, hip.LHS_FP64, hip.ADI_EXACT)
    n, l2a, lla = g.step(1000, check_every=100)
    assert n == 1000
    assert hashlib.sha256(g.get("q").tobytes()).hexdigest() == str(d["q_sha256_1000"])
    n, l2b, llb = g.step(5000, check_every=100)
    assert n == 5000
    assert hashlib.sha256(g.get("q").tobytes()).hexdigest() == str(d["q_sha256_6000"])
    np.testing.assert_allclose(np.concatenate([l2a, l2b]), d["l2rhs"], rtol=1e-12, atol=0)
    np.testing.assert_allclose(np.concatenate([lla, llb]), d["l2lin"], rtol=1e-12, atol=0)
    g.close()


@pytest.mark.parametrize("lhs", ["fp64", "fp32"])
def test_fast_settles_into_the_same_cycle(hip, wc, lhs):
    d, mesh = wc
    g = _solver(hip, d, mesh, hip.LHS_FP32 if lhs == "fp32" else hip.LHS_FP64, hip.ADI_FAST)
    n, l2, _ = g.step(6000, check_every=100)
    assert n == 6000
    ref = d["l2rhs"][2000:]
    tail = l2[2000:]
    k, strength = period(tail)
    print(f"WENO cycle FAST {lhs}: band [{tail.min():.4e}, {tail.max():.4e}] (oracle [{ref.min():.4e}, "
          f"{ref.max():.4e}]), period {k} ({strength:.2f})", flush=True)
    assert 0.97 * ref.min() <= tail.min() <= 1.03 * ref.min()
    assert 0.97 * ref.max() <= tail.max() <= 1.03 * ref.max()
    assert k == period(ref)[0] and strength > 0.9
    g.close()

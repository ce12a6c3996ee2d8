"""The oracle's 3D / strand building blocks (5 conserved variables):
fluxRoe3D (interfaceFlux.cpp:176-319), jacobianRoe3D (jacobianRoe.cpp:92-620,
IMODE 1) and the 5x5 block solvers blockTridag / blockTridagPeriodic
(triSolvers.cpp:443-766).

Pinned three ways:
* bit for bit against tests/golden/unit3d.npz, written by the reference's own
  2D build (-O2 and -O0 agree), and against the live reference harness where
  it is built;
* against what the mathematics fixes: the flux of equal states is the
  physical flux through the unit normal (no area factor, unlike fluxRoe2D);
  the flux is antisymmetric under swapping the states and the normal; the
  Jacobians of equal states sum to the physical flux Jacobian and differ by
  the Roe dissipation matrix; the block solvers match a dense (cyclic) solve;
* against the 2D path, which the golden cases pin independently: with w = 0
  and n_z = 0 the 3D flux times the face length is fluxRoe2D bit for bit, and
  the 3D Jacobians' rows and columns {0,1,2,4} times the face length are
  jacobianRoe2D's golden vectors.
"""
import os

import numpy as np
import pytest

from cases import random_block_system5, random_normals3, random_states5

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
G = 1.4


@pytest.fixture(scope="module")
def u3():
    return np.load(os.path.join(GOLD, "unit3d.npz"))


def phys_flux5(q, n):
    r, ru, rv, rw, E = q
    u, v, w = ru / r, rv / r, rw / r
    p = (G - 1) * (E - 0.5 * r * (u * u + v * v + w * w))
    nh = n / np.linalg.norm(n)
    un = u * nh[0] + v * nh[1] + w * nh[2]
    return np.array([r * un, ru * un + p * nh[0], rv * un + p * nh[1], rw * un + p * nh[2], (E + p) * un])


def phys_jac5(q, n, h=1e-6):
    J = np.zeros((5, 5))
    for k in range(5):
        e = np.zeros(5)
        e[k] = h * max(1.0, abs(q[k]))
        J[:, k] = (phys_flux5(q + e, n) - phys_flux5(q - e, n)) / (2 * e[k])
    return J


def sound5(q):
    r, ru, rv, rw, E = q
    p = (G - 1) * (E - 0.5 * (ru * ru + rv * rv + rw * rw) / r)
    return np.sqrt(G * p / r)


def test_golden_flux_and_jacobian(oracle, u3):
    for i in range(u3["flux_n"].shape[0]):
        f, s = oracle.flux_roe3d(u3["flux_n"][i], u3["flux_ql"][i], u3["flux_qr"][i])
        assert np.array_equal(f, u3["flux"][i]) and s == u3["flux_sr"][i], i
        lm, rm = oracle.jacobian_roe3d(u3["flux_n"][i], u3["flux_ql"][i], u3["flux_qr"][i])
        assert np.array_equal(lm, u3["jac_l"][i]) and np.array_equal(rm, u3["jac_r"][i]), i


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("N", [3, 6, 11])
def test_golden_block_tridag5(oracle, u3, periodic, N):
    tag = "per" if periodic else "tri"
    a, b, c, f, x = (u3[f"{tag}{N}_{k}"] for k in "abcfx")
    for s in range(a.shape[0]):
        assert np.array_equal(oracle.block_tridag5(a[s], b[s], c[s], f[s], periodic=periodic), x[s])


def test_flux3d_consistency_and_antisymmetry(oracle):
    rng = np.random.default_rng(31)
    for q, ql, qr, n in zip(random_states5(rng, 40), random_states5(rng, 40), random_states5(rng, 40),
                            random_normals3(rng, 40)):
        f, sr = oracle.flux_roe3d(n, q, q)
        np.testing.assert_allclose(f, phys_flux5(q, n), rtol=1e-13, atol=1e-13)
        un = (q[1:4] @ (n / np.linalg.norm(n))) / q[0]
        assert sr == pytest.approx(abs(un) + sound5(q), rel=1e-13)
        f1, s1 = oracle.flux_roe3d(n, ql, qr)
        f2, s2 = oracle.flux_roe3d(-n, qr, ql)
        np.testing.assert_allclose(f1, -f2, rtol=1e-12, atol=1e-13)
        assert s1 == pytest.approx(s2, rel=1e-13)


def test_flux3d_upwinds_when_supersonic(oracle):
    rng = np.random.default_rng(32)
    for _ in range(20):
        n = random_normals3(rng, 1)[0]
        nh = n / np.linalg.norm(n)
        ql = random_states5(rng, 1, vel=0.0)[0]
        # add a normal velocity above the Roe-average sound speed to both states
        for q in (ql,):
            r = q[0]
            q[1:4] += r * 3.0 * nh
            q[4] += 0.5 * r * 9.0
        qr = ql * (1 + 0.01 * rng.uniform(-1, 1, 5))
        f, _ = oracle.flux_roe3d(n, ql, qr)
        np.testing.assert_allclose(f, phys_flux5(ql, n), rtol=1e-12, atol=1e-12)


def test_flux3d_reduces_to_flux2d(oracle):
    """w = 0, n_z = 0: flux3d x face length is fluxRoe2D bit for bit."""
    rng = np.random.default_rng(33)
    from cases import random_normals, random_states
    for q2l, q2r, n2 in zip(random_states(rng, 40), random_states(rng, 40), random_normals(rng, 40)):
        f2, s2 = oracle.flux_roe2d(n2, q2l, q2r)
        n3 = np.array([n2[0], n2[1], 0.0])
        f3, s3 = oracle.flux_roe3d(n3, np.insert(q2l, 3, 0.0), np.insert(q2r, 3, 0.0))
        area = np.sqrt(n2[0] * n2[0] + n2[1] * n2[1])
        assert np.array_equal(f3[[0, 1, 2, 4]] * area, f2)
        assert f3[3] == 0.0 and s3 * area == s2


def test_jacobian3d_equal_states(oracle):
    """lmat + rmat = A(q) and lmat - rmat = |A| (Roe dissipation) at qL = qR."""
    rng = np.random.default_rng(34)
    for q, n in zip(random_states5(rng, 30), random_normals3(rng, 30)):
        lm, rm = oracle.jacobian_roe3d(n, q, q)
        np.testing.assert_allclose(lm + rm, phys_jac5(q, n), rtol=1e-7, atol=1e-7)
        A = phys_jac5(q, n)
        w, V = np.linalg.eig(A)
        absA = (V @ np.diag(np.abs(w)) @ np.linalg.inv(V)).real
        np.testing.assert_allclose(lm - rm, absA, rtol=1e-6, atol=1e-6)


def test_jacobian3d_embeds_the_2d_golden(oracle):
    """Rows and columns {0,1,2,4} of the 3D Jacobians times the face length
    are the reference's own jacobianRoe2D outputs (tests/golden/unit.npz)."""
    u = np.load(os.path.join(GOLD, "unit.npz"))
    idx = [0, 1, 2, 4]
    for i in range(u["flux_n"].shape[0]):
        n2 = u["flux_n"][i]
        lm, rm = oracle.jacobian_roe3d(np.array([n2[0], n2[1], 0.0]), np.insert(u["flux_ql"][i], 3, 0.0),
                                       np.insert(u["flux_qr"][i], 3, 0.0))
        area = np.sqrt(n2[0] * n2[0] + n2[1] * n2[1])
        assert np.array_equal(lm[np.ix_(idx, idx)] * area, u["jac_l"][i]), i
        assert np.array_equal(rm[np.ix_(idx, idx)] * area, u["jac_r"][i]), i


@pytest.mark.parametrize("periodic", [False, True])
def test_block_tridag5_dense(oracle, periodic):
    rng = np.random.default_rng(35)
    for N in (3, 4, 9, 17):
        a, b, c, f = random_block_system5(rng, 1, N)
        a, b, c, f = a[0], b[0], c[0], f[0]
        M = np.zeros((5 * N, 5 * N))
        for i in range(N):
            M[5 * i:5 * i + 5, 5 * i:5 * i + 5] = b[i]
            im, ip = i - 1, i + 1
            if im >= 0 or periodic:
                M[5 * i:5 * i + 5, 5 * (im % N):5 * (im % N) + 5] += a[i]
            if ip < N or periodic:
                M[5 * i:5 * i + 5, 5 * (ip % N):5 * (ip % N) + 5] += c[i]
        x = np.linalg.solve(M, f.reshape(-1)).reshape(N, 5)
        got = oracle.block_tridag5(a, b, c, f, periodic=periodic)
        np.testing.assert_allclose(got, x, rtol=1e-11, atol=1e-12)


def _reference(opt):
    import oracle_py
    if not oracle_py.Reference.available(opt):
        pytest.skip("reference harness not built (needs the reference tree; make -C oracle ref)")
    return oracle_py.Reference(opt)


def test_vs_reference_harness(oracle):
    """More inputs straight against the compiled reference, bit for bit."""
    ref = _reference("O2")
    rng = np.random.default_rng(36)
    for q1, q2, n in zip(random_states5(rng, 50), random_states5(rng, 50), random_normals3(rng, 50)):
        assert all(np.array_equal(x, y) for x, y in zip(oracle.flux_roe3d(n, q1, q2), ref.flux_roe3d(n, q1, q2)))
        lo, ro = oracle.jacobian_roe3d(n, q1, q2)
        lr, rr = ref.jacobian_roe3d(n, q1, q2)
        assert np.array_equal(lo, lr) and np.array_equal(ro, rr)
    for periodic in (False, True):
        a, b, c, f = random_block_system5(rng, 1, 24)
        assert np.array_equal(oracle.block_tridag5(a[0], b[0], c[0], f[0], periodic),
                              ref.block_tridag5(a[0], b[0], c[0], f[0], periodic))

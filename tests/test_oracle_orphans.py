"""The oracle's orphan-face extension for meshes with triangles.

The reference cannot run these meshes (a chain covers exactly two faces of
each cell it visits, so a triangle leaves one face uncovered).  The extension
gives each orphan face (interior, in no chain) a first-order Roe flux of its
two cell states, puts its Jacobian blocks on the diagonal of the ADI rows of
the cell's last chain, and its full linearised flux into computeLinearRHS.

What pins it here, independently of the oracle's chain code:
  - at order 1 every face, chain or orphan, carries the first-order flux of
    its two cell states (or the cell state and its ghost), so a face-by-face
    restatement of the residual and of the linear residual must agree with
    the oracle to rounding (the summation order differs);
  - with no orphans the extension is the reference (the golden tests);
  - the iteration converges to that pinned residual's zero.
"""
import numpy as np
import pytest

from cases import perturbed_state

from hamstr_oop_amd import meshgen


def _ghost(o, q_left, f, rc, qinf):
    if rc == -1:
        return qinf.copy()
    R = o.refmtx[f]
    g = q_left.copy()
    g[1] = R[0] * q_left[1] + R[1] * q_left[2]
    g[2] = R[2] * q_left[1] + R[3] * q_left[2]
    return g


def _facewise_rhs(oracle, o, q):
    q = q.reshape(-1, 4)
    r = np.zeros_like(q)
    sig = np.zeros(q.shape[0])
    qinf = o.freestream()
    for f, (n1, n2, cl, el, cr, er) in enumerate(o.faces):
        ql = q[cl]
        qr = q[cr] if cr >= 0 else _ghost(o, ql, f, cr, qinf)
        fl, sr = oracle.flux_roe2d(o.normvec[f], ql, qr)
        r[cl] -= fl
        sig[cl] += sr
        if cr >= 0:
            r[cr] += fl
            sig[cr] += sr
    return r.reshape(-1), sig


@pytest.mark.parametrize("mk", [lambda: meshgen.circle_ogrid(16, 8, tri_layers=3),
                                lambda: meshgen.naca_ogrid(24, 10, tri_layers=4)])
def test_orphan_rhs_facewise_order1(oracle, mk):
    mesh = mk()
    assert len(mesh.orphans) > 0
    o = oracle.case(mesh, 0.5, 2.0, 1, 5.0)
    q = perturbed_state(o, np.random.default_rng(3))
    o.set("q", q)
    o.compute_rhs(True)
    r, sig = _facewise_rhs(oracle, o, q)
    scale = np.abs(r).max()
    assert np.abs(o.get("r") - r).max() <= 1e-13 * scale
    np.testing.assert_allclose(o.get("sigma"), sig, rtol=1e-14)
    # orphan Jacobians are the frozen-Roe blocks of the two cell states
    ff = o.get("ff").reshape(-1, 32)
    qq = q.reshape(-1, 4)
    for f in mesh.orphans[:5]:
        lm, rm = oracle.jacobian_roe2d(o.normvec[f], qq[o.faces[f, 2]], qq[o.faces[f, 4]])
        assert np.array_equal(ff[f, :16], lm.reshape(-1)) and np.array_equal(ff[f, 16:], rm.reshape(-1))


def test_orphan_linear_rhs_facewise_order1(oracle):
    mesh = meshgen.circle_ogrid(16, 8, tri_layers=3)
    cfl = 5.0
    o = oracle.case(mesh, 0.5, 2.0, 1, cfl)
    rng = np.random.default_rng(4)
    q = perturbed_state(o, rng)
    o.set("q", q)
    o.compute_rhs(True)
    n = mesh.n_cell
    r0 = rng.uniform(-1e-3, 1e-3, 4 * n)
    dq = rng.uniform(-1e-3, 1e-3, 4 * n)
    o.set("r0", r0)
    o.set("dq", dq)
    o.compute_linear_rhs()
    ff = o.get("ff").reshape(-1, 2, 4, 4)
    sig = o.get("sigma")
    d = dq.reshape(-1, 4)
    want = (r0 - dq).reshape(-1, 4)
    for f, (n1, n2, cl, el, cr, er) in enumerate(o.faces):
        if cr >= 0:
            dr = d[cr]
        elif cr == -1:
            dr = np.zeros(4)
        else:
            dr = _ghost(o, d[cl], f, cr, np.zeros(4))
        df = ff[f, 0] @ d[cl] + ff[f, 1] @ dr
        want[cl] -= df * cfl / sig[cl]
        if cr >= 0:
            want[cr] += df * cfl / sig[cr]
    got = o.get("r").reshape(-1, 4)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("order", [1, 3])
def test_triangulated_mesh_converges(oracle, order):
    """ADI with orphan diagonals drives the residual of a triangulated O-grid
    to 1e-10 from freestream; at order 1 the face-by-face residual of the
    final state is at that level too."""
    mesh = meshgen.naca_ogrid(24, 10, tri_layers=4)
    o = oracle.case(mesh, 0.5, 2.0, order, 10.0)
    rc, l2r, _, _ = o.step(800)
    assert rc == 0
    k = int(np.argmax(l2r <= 1e-10)) if (l2r <= 1e-10).any() else -1
    assert k > 0, f"min nonlinear L2 {l2r.min():.3e}"
    if order == 1:
        r, _ = _facewise_rhs(oracle, o, o.get("q"))
        assert np.sqrt((r.reshape(-1, 4)[:, 0] ** 2).sum()) / mesh.n_cell < 1e-9


def test_empty_orphan_list_is_the_reference_path(oracle):
    """The same all-quad mesh with and without an (empty) orphan array: identical."""
    mesh = meshgen.naca_ogrid(24, 10)
    assert mesh.orphans.size == 0
    o1 = oracle.case(mesh, 0.6, 1.0, 3, 10.0)
    _, a, b, _ = o1.step(5)
    mesh.orphans = np.zeros(0, np.int32)
    o2 = oracle.case(mesh, 0.6, 1.0, 3, 10.0)
    _, c, d, _ = o2.step(5)
    assert np.array_equal(a, c) and np.array_equal(b, d) and np.array_equal(o1.get("q"), o2.get("q"))

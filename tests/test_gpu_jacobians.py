"""The fused FAST iteration with the chain faces' Jacobians computed again in
the ADI kernels (hamstr_set_jacobian_source: from the face states k_rhs_faces
stores, or from q; k_rhs_faces then stores the Jacobians of orphan faces and
chain-end faces only) against the same iteration storing every face's
Jacobians and reading them back (HAMSTR_JAC_STORED): face_jacobians is opaque
to the code around it, so the recomputed blocks are the stored ones bit for
bit, and the states after a run of iterations are equal (up to the sign of
zero entries, which numpy's comparison ignores).  Covered: orders 1, 3 and 5, FP64 and FP32
left-hand sides, the one-row kernel (loops with two-row and odd final cycles,
open lines), both half-LDS kernels, triangulated meshes with orphan faces,
msweep 2 and a batch.  The reference always stores them (computeRHS.cpp:362-369)
and ADI reads them (ADI.cpp:96-97)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from hamstr_oop_amd import hip as H
    H.lib()
    return H


def _eq(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        raise AssertionError(f"{what}: {np.count_nonzero(a != b)} of {a.size} differ; max |diff| {d.max():.3e}")


MESHES = {
    # loops of 40 rows (two-row cycles), lines of 10 quads + 6 triangle pairs
    "naca40x16_tri6": lambda m: m.naca_ogrid(40, 16, tri_layers=6),
    # loops of 24 rows (odd three-row cycles), lines of 10
    "circle24x10": lambda m: m.circle_ogrid(24, 10),
    # lines of 136 quads + 128 triangle pairs = 264 rows: two per workgroup (FP64)
    "naca48x200_tri64": lambda m: m.naca_ogrid(48, 200, tri_layers=64),
}


def _pair(hip, mesh, order, lhs, msweep=1):
    kw = dict(lhs=lhs, adi_mode=hip.ADI_FAST, msweep=msweep)
    return (hip.Solver(mesh, 0.5, 2.0, order, 10.0, **kw), hip.Solver(mesh, 0.5, 2.0, order, 10.0, **kw))


SOURCES = {"states": 2, "q": 1, "states_lines": 3}


@pytest.mark.parametrize("src", list(SOURCES))
@pytest.mark.parametrize("lhs", ["fp64", "fp32"])
@pytest.mark.parametrize("order", [1, 3, 5])
@pytest.mark.parametrize("name", list(MESHES))
def test_recomputed_jacobians_give_the_same_iteration(hip, name, order, lhs, src):
    from hamstr_oop_amd import meshgen
    mesh = MESHES[name](meshgen)
    L = hip.LHS_FP32 if lhs == "fp32" else hip.LHS_FP64
    print(f"recomputed Jacobians {name} order={order} {lhs} from {src}", flush=True)
    a, b = _pair(hip, mesh, order, L)
    assert hip.jacobian_source(a) == hip.JAC_STORED
    odd32 = lhs == "fp32" and any(g["final"] == 2 for g in hip.launch_groups(b))
    if odd32:
        # FP32 odd cycles read their Jacobians from ff
        with pytest.raises(hip.HamstrError):
            hip.set_jacobian_source(b, SOURCES[src])
        a.close()
        b.close()
        return
    hip.set_jacobian_source(b, SOURCES[src])
    assert hip.jacobian_source(b) == SOURCES[src]
    na, l2a, lla = a.step(12, check_every=4)
    nb, l2b, llb = b.step(12, check_every=4)
    assert na == nb == 12
    _eq(a.get("q"), b.get("q"), "q")
    _eq(a.get("dq"), b.get("dq"), "dq")
    _eq(l2a, l2b, "nonlinear L2")
    _eq(lla, llb, "linear L2")
    a.close()
    b.close()


def test_half_lds_kernels_and_msweep(hip):
    """Both half-LDS kernels (FP64; the lines, the odd-cycle loops keeping the
    one-row kernel) and two sweeps per iteration."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(48, 200, tri_layers=64)
    for kernel, src in ((hip.KERNEL_ROW, hip.JAC_FROM_STATES), (hip.KERNEL_HALF, hip.JAC_FROM_STATES),
                        (hip.KERNEL_HALF2, hip.JAC_FROM_STATES), (hip.KERNEL_HALF2, hip.JAC_FROM_Q)):
        for msweep in (1, 2):
            print(f"recomputed Jacobians kernel={kernel} source={src} msweep={msweep}", flush=True)
            a, b = _pair(hip, mesh, 3, hip.LHS_FP64, msweep)
            for s in (a, b):
                for k, g in enumerate(hip.launch_groups(s)):
                    kk = kernel
                    if kk == hip.KERNEL_HALF2 and (g["final"] != 0 or g["max_rows"] <= 256):
                        kk = hip.KERNEL_HALF
                    try:
                        hip.set_adi_kernel(s, k, kk)
                    except hip.HamstrError:
                        assert kk != hip.KERNEL_ROW  # odd cycles keep the one-row kernel
            hip.set_jacobian_source(b, src)
            a.step(8, check_every=8)
            b.step(8, check_every=8)
            _eq(a.get("q"), b.get("q"), f"q, kernel {kernel}, msweep {msweep}")
            a.close()
            b.close()


def test_batch_and_autotuner(hip):
    """A batch of uneven meshes with the Jacobians recomputed, and the
    autotuner's choice (whichever it is) giving the same states as both
    settings under the launch shapes it picked."""
    from hamstr_oop_amd import meshgen
    cases = [(meshgen.circle_ogrid(32, 12), 0.3, 0.0), (meshgen.naca_ogrid(40, 16, tri_layers=6), 0.6, 1.0),
             (meshgen.naca_ogrid(64, 16), 0.5, -2.0)]
    ref = hip.Batch(cases, 3, 10.0, adi_mode=hip.ADI_FAST)
    on = hip.Batch(cases, 3, 10.0, adi_mode=hip.ADI_FAST)
    onq = hip.Batch(cases, 3, 10.0, adi_mode=hip.ADI_FAST)
    tuned = hip.Batch(cases, 3, 10.0, adi_mode=hip.ADI_FAST)
    tuned.computeRHS("implicit")
    hip.autotune(tuned)
    print(f"autotuner chose Jacobian source {hip.jacobian_source(tuned)}", flush=True)
    # the tuned kernels, grids and depths on the other three (the depths change
    # the rounding of the block solves)
    for k, g in enumerate(hip.launch_groups(tuned)):
        for b in (ref, on, onq):
            if g["kernel"] in (hip.KERNEL_ROW, hip.KERNEL_HALF, hip.KERNEL_HALF2):
                hip.set_adi_kernel(b, k, g["kernel"])
            hip.set_launch(b, k, g["grid"], g["cr_levels"])
            if g["thomas"] >= 0:
                hip.set_thomas(b, k, g["thomas"])
    hip.set_jacobian_source(on, hip.JAC_FROM_STATES)
    hip.set_jacobian_source(onq, hip.JAC_FROM_Q)
    for b in (ref, on, onq, tuned):
        b.batch_step(20, check_every=10)
    for k in range(len(cases)):
        _eq(ref.case_q(k), on.case_q(k), f"case {k} from states")
        _eq(ref.case_q(k), onq.case_q(k), f"case {k} from q")
        _eq(ref.case_q(k), tuned.case_q(k), f"case {k} autotuned")
    for b in (ref, on, onq, tuned):
        b.close()


def test_refused_where_a_kernel_cannot_compute_them(hip):
    from hamstr_oop_amd import meshgen
    s = hip.Solver(meshgen.naca_ogrid(40, 16, tri_layers=6), 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_EXACT)
    for src in (hip.JAC_FROM_Q, hip.JAC_FROM_STATES, 7):
        with pytest.raises(hip.HamstrError):
            hip.set_jacobian_source(s, src)
    hip.set_jacobian_source(s, hip.JAC_STORED)
    assert hip.jacobian_source(s) == hip.JAC_STORED
    s.close()


@pytest.mark.parametrize("batch", [False, True])
def test_autotuner_restores_the_state(hip, batch):
    """The autotuner times the fused iteration's steady state, which steps the
    solution: q, r, r0, dq, sigma and (single case) the stored Jacobians come
    back as they were (bitwise), and the next steps equal those of an untuned
    context given the tuned launch shapes and source."""
    from hamstr_oop_amd import meshgen
    cases = [(meshgen.naca_ogrid(64, 24, tri_layers=8), 0.5, 2.0), (meshgen.circle_ogrid(32, 12), 0.3, 0.0)]
    if batch:
        mk = lambda: hip.Batch(cases, 3, 10.0, adi_mode=hip.ADI_FAST)  # noqa: E731
    else:
        mk = lambda: hip.Solver(cases[0][0], cases[0][1], cases[0][2], 3, 10.0, adi_mode=hip.ADI_FAST)  # noqa: E731
    a, ref = mk(), mk()
    for s in (a, ref):
        s.step(5, check_every=5)
        s.computeRHS("implicit")
    fields = ("q", "r", "r0", "dq", "sigma") + (() if batch else ("ff",))   # a batch has no per-face read-back
    before = {f: a.get(f).copy() for f in fields}
    print(f"autotuner state restore, batch={batch}", flush=True)
    hip.autotune(a)
    for f, v in before.items():
        _eq(a.get(f), v, f"{f} after the autotuner")
    for k, g in enumerate(hip.launch_groups(a)):
        if g["kernel"] in (hip.KERNEL_ROW, hip.KERNEL_HALF, hip.KERNEL_HALF2):
            hip.set_adi_kernel(ref, k, g["kernel"])
        if g["fast"]:
            hip.set_launch(ref, k, g["grid"], g["cr_levels"])
        if g["thomas"] >= 0:
            hip.set_thomas(ref, k, g["thomas"])
    if hip.jacobian_source(a) != hip.JAC_STORED:
        hip.set_jacobian_source(ref, hip.jacobian_source(a))
    for s in (a, ref):
        if batch:
            s.batch_step(12, check_every=6)
        else:
            s.step(12, check_every=6)
    _eq(a.get("q"), ref.get("q"), "q 12 steps after tuning")
    a.close()
    ref.close()

"""The Thomas finish of the FAST ADI kernels (hamstr_set_thomas): open chains
stop the parallel cyclic reduction after a few levels and solve the
interleaved subsystems left with the block Thomas algorithm, one quad of
lanes per subsystem.

* Every kernel (one row per thread, half the rows in LDS, two lines per
  workgroup) runs the same operations for the same setting, so their states
  after a run of iterations are equal bit for bit (up to the sign of zero).
* Against the reference's sequential block Thomas order (EXACT) the ADI
  increment agrees to rounding, for every setting and cyclic-reduction depth.
"""
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


def _open_groups(hip, s):
    return [k for k, g in enumerate(hip.launch_groups(s))
            if g["final"] == 0 and g["kernel"] in (hip.KERNEL_ROW, hip.KERNEL_HALF, hip.KERNEL_HALF2)]


def test_thomas_setting_api(hip):
    """Defaults, bounds and the groups that take a Thomas finish."""
    from hamstr_oop_amd import meshgen
    s = hip.Solver(meshgen.naca_ogrid(64, 24), 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_FAST)
    gs = hip.launch_groups(s)
    lines = _open_groups(hip, s)
    loops = [k for k, g in enumerate(gs) if g["final"] != 0]
    assert lines and loops
    for k in lines:
        hip.set_thomas(s, k, 3)
        assert hip.thomas_levels(s, k) == 3
        for bad in (-2, 5):
            with pytest.raises(hip.HamstrError):
                hip.set_thomas(s, k, bad)
    with pytest.raises(hip.HamstrError):   # closed loops have no Thomas finish
        hip.set_thomas(s, loops[0], 3)
    s.close()
    e = hip.Solver(meshgen.naca_ogrid(64, 24), 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_EXACT)
    with pytest.raises(hip.HamstrError):   # EXACT groups have no reduction levels
        hip.set_thomas(e, 0, 3)
    e.close()


@pytest.mark.parametrize("tpcr", [2, 3, 4])
@pytest.mark.parametrize("name,mk,ncr", [
    ("naca512x256_tri128", lambda m: m.naca_ogrid(512, 256, tri_layers=128), 2),  # 384-row lines
    ("circle96x45", lambda m: m.circle_ogrid(96, 45), 1),                         # odd line length
    ("circle48x301", lambda m: m.circle_ogrid(48, 301), 3),                       # odd, two per workgroup
])
def test_thomas_kernels_agree(hip, name, mk, ncr, tpcr):
    from hamstr_oop_amd import meshgen
    mesh = mk(meshgen)
    print(f"thomas {name} ncr={ncr} tpcr={tpcr}", flush=True)
    out = {}
    for kernel in (hip.KERNEL_ROW, hip.KERNEL_HALF, hip.KERNEL_HALF2):
        s = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_FAST)
        gs = hip.launch_groups(s)
        for k in _open_groups(hip, s):
            try:
                hip.set_adi_kernel(s, k, kernel)
            except hip.HamstrError:
                assert kernel == hip.KERNEL_HALF2  # lines of at most 256 rows
                hip.set_adi_kernel(s, k, hip.KERNEL_HALF)
            hip.set_launch(s, k, gs[k]["grid"], ncr)
            hip.set_thomas(s, k, tpcr)
        s.step(12, check_every=12)
        out[kernel] = s.get("q")
        s.close()
    for kernel in (hip.KERNEL_HALF, hip.KERNEL_HALF2):
        _eq(out[kernel], out[hip.KERNEL_ROW], f"kernel {kernel}: q after 12 steps")


def test_thomas_fp32_half2_equals_row_kernel(hip):
    """FP32 left-hand side, 384-row lines: two lines per workgroup against the
    one-row kernel with the same Thomas finish."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(512, 256, tri_layers=128)
    out = {}
    for kernel in (hip.KERNEL_ROW, hip.KERNEL_HALF2):
        s = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, lhs=hip.LHS_FP32, adi_mode=hip.ADI_FAST)
        for k in _open_groups(hip, s):
            hip.set_adi_kernel(s, k, kernel)
            hip.set_thomas(s, k, 3)
        s.step(12, check_every=12)
        out[kernel] = s.get("q")
        s.close()
    _eq(out[hip.KERNEL_HALF2], out[hip.KERNEL_ROW], "FP32 q after 12 steps")


def _cases():
    from hamstr_oop_amd import meshgen
    return [
        (meshgen.circle_ogrid(16, 8), 0.3, 0.0),
        (meshgen.naca_ogrid(32, 12), 0.8, 1.25),
        (meshgen.naca_ogrid(64, 40, t=0.14, tri_layers=8), 0.6, -3.0),
        (meshgen.circle_ogrid(40, 130), 0.5, 2.0),
    ]


@pytest.mark.parametrize("lhs,tol", [("fp64", 1e-11), ("fp32", 2e-4)])
@pytest.mark.parametrize("tpcr", [0, 2, 3, 4])
@pytest.mark.parametrize("cr", [0, 1, 2])
def test_thomas_matches_exact(hip, cr, tpcr, lhs, tol):
    """One computeRHS + ADI: the ADI increment with the Thomas finish against
    the reference's sequential order (EXACT, the same left-hand-side
    precision), scaled per component; lines of 8 to 130 rows with orphan
    faces, FP64 to 1e-11, FP32 to its rounding."""
    lm = hip.LHS_FP64 if lhs == "fp64" else hip.LHS_FP32
    cases = _cases()
    fast = hip.Batch(cases, order=3, cfl=10.0, lhs=lm, adi_mode=hip.ADI_FAST)
    exact = hip.Batch(cases, order=3, cfl=10.0, lhs=lm, adi_mode=hip.ADI_EXACT)
    gs = hip.launch_groups(fast)
    for k in _open_groups(hip, fast):
        hip.set_launch(fast, k, gs[k]["grid"], cr)
        hip.set_thomas(fast, k, tpcr)
    for b in (fast, exact):
        b.computeRHS("implicit")
        b.ADI()
    xe, xf = exact.get("dq").reshape(-1, 4), fast.get("dq").reshape(-1, 4)
    scale = np.abs(xe).max(axis=0)
    assert (np.abs(xf - xe) / scale).max() < tol
    fast.close()
    exact.close()


def test_autotune_with_thomas_keeps_results(hip):
    """The autotuner's Thomas coordinate changes the increment only to
    rounding: 20 tuned FAST steps of the 384-row-line mesh against 20 untuned
    ones, to 1e-10 of the freestream scale."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(512, 256, tri_layers=128)
    a = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_FAST)
    b = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_FAST)
    hip.autotune(b)
    lv = [hip.thomas_levels(b, k) for k in _open_groups(hip, b)]
    print(f"tuned Thomas levels {lv}", flush=True)
    a.step(20, check_every=20)
    b.step(20, check_every=20)
    qa, qb = a.get("q").reshape(-1, 4), b.get("q").reshape(-1, 4)
    scale = np.array([1.0, 0.5, 0.5, 1.0 / (1.4 * 0.4) + 0.125])
    assert (np.abs(qa - qb) / scale).max() < 1e-10
    a.close()
    b.close()

"""The half-LDS FAST ADI kernels (HAMSTR_KERNEL_HALF: one row per thread, the
rows the first cyclic-reduction level eliminates in global scratch;
HAMSTR_KERNEL_HALF2: the same for two open chains per workgroup) against the
one-row kernel with every row in LDS (HAMSTR_KERNEL_ROW, whose parallel cyclic
reduction runs by quads of lanes where the workgroup allows): the same
operations in the same order, so the states after a run of iterations are
equal (up to the sign of zero entries, which numpy's comparison ignores), for
open lines, two-row-cycle loops, odd chain lengths, an odd number of chains,
batches and every cyclic-reduction depth."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from hamstr_oop_amd import hip as H
    H.lib()
    return H


def _kernels(hip):
    return (hip.KERNEL_ROW, hip.KERNEL_HALF, hip.KERNEL_HALF2)


def _set(hip, s, k, g, kernel):
    """`kernel` on a one-row group whose chains fit the half-LDS kernels (two
    chains per workgroup on open groups only)"""
    if g["kernel"] in (hip.KERNEL_ROW, hip.KERNEL_HALF, hip.KERNEL_HALF2):
        if kernel == hip.KERNEL_HALF2 and g["final"] != 0:
            kernel = hip.KERNEL_HALF
        try:
            hip.set_adi_kernel(s, k, kernel)
        except hip.HamstrError:
            assert kernel != hip.KERNEL_ROW  # the group's chains do not fit the half-LDS kernels


def _run(hip, s, kernel, ncr, steps):
    """`kernel` on the groups that start with a half-LDS kernel, the same
    cyclic-reduction depth on every one-row group"""
    for k, g in enumerate(hip.launch_groups(s)):
        if g["kernel"] in _kernels(hip):
            _set(hip, s, k, g, kernel)
            hip.set_launch(s, k, g["grid"], ncr)
    n, l2, _ = s.step(steps, check_every=steps)
    return s.get("q"), l2


def _eq(a, b, what):
    if not np.array_equal(a, b):
        d = np.abs(a - b)
        raise AssertionError(f"{what}: {np.count_nonzero(a != b)} of {a.size} differ; max |diff| {d.max():.3e}")


def test_default_groups(hip):
    """FP64 contexts pick two lines per workgroup (half-LDS) for open lines of
    257 to 384 rows, the one-row kernel elsewhere; the half-LDS kernels stay
    available where the chains fit them (two-row-cycle loops and open lines of
    up to 384 rows); FP32 contexts and longer chains keep theirs."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(256, 128)  # loops of 256 rows, lines of 128
    s = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_FAST)
    ks = {g["max_rows"]: g["kernel"] for g in hip.launch_groups(s)}
    assert ks == {256: hip.KERNEL_ROW, 128: hip.KERNEL_ROW}, ks
    for k in range(2):
        hip.set_adi_kernel(s, k, hip.KERNEL_HALF)
    lines = [g["max_rows"] for g in hip.launch_groups(s)].index(128)
    with pytest.raises(hip.HamstrError):   # two per workgroup: lines of more than 256 rows only
        hip.set_adi_kernel(s, lines, hip.KERNEL_HALF2)
    loops = [g["max_rows"] for g in hip.launch_groups(s)].index(256)
    with pytest.raises(hip.HamstrError):
        hip.set_adi_kernel(s, loops, hip.KERNEL_HALF2)
    s.close()
    s = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, lhs=hip.LHS_FP32, adi_mode=hip.ADI_FAST)
    assert all(g["kernel"] == hip.KERNEL_ROW for g in hip.launch_groups(s))
    s.close()
    big = meshgen.naca_ogrid(512, 256, tri_layers=128)  # loops of 512 rows, lines of 384
    s = hip.Solver(big, 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_FAST)
    ks = {g["max_rows"]: g["kernel"] for g in hip.launch_groups(s)}
    assert ks == {512: hip.KERNEL_ROW, 384: hip.KERNEL_HALF2}, ks
    with pytest.raises(hip.HamstrError):
        hip.set_adi_kernel(s, [g["max_rows"] for g in hip.launch_groups(s)].index(512), hip.KERNEL_HALF)
    s.close()


@pytest.mark.parametrize("ncr", [1, 2, 3])
@pytest.mark.parametrize("name,mk", [
    ("naca512x256_tri128", lambda m: m.naca_ogrid(512, 256, tri_layers=128)),
    ("naca256x128", lambda m: m.naca_ogrid(256, 128)),
    ("circle96x45", lambda m: m.circle_ogrid(96, 45)),       # odd line length
    ("circle48x301", lambda m: m.circle_ogrid(48, 301)),     # odd, long enough for two per workgroup
])
def test_half_equals_row_kernel(hip, name, mk, ncr):
    from hamstr_oop_amd import meshgen
    mesh = mk(meshgen)
    print(f"half vs row {name} ncr={ncr}", flush=True)
    out = {}
    for kernel in _kernels(hip):
        s = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, adi_mode=hip.ADI_FAST)
        out[kernel] = _run(hip, s, kernel, ncr, 12)
        s.close()
    for kernel in (hip.KERNEL_HALF, hip.KERNEL_HALF2):
        _eq(out[kernel][0], out[hip.KERNEL_ROW][0], f"kernel {kernel}: q after 12 steps")
        np.testing.assert_array_equal(out[kernel][1], out[hip.KERNEL_ROW][1])


def test_half_in_batch_and_small_grid(hip):
    """A batch of meshes of different sizes (persistent launches walking
    several chains per workgroup), half against row kernel; lines of 300 and
    270 rows share the two-per-workgroup group."""
    from hamstr_oop_amd import meshgen
    meshes = [meshgen.naca_ogrid(128, 64), meshgen.circle_ogrid(48, 300), meshgen.circle_ogrid(40, 270)]
    flows = [(0.3, 0.0), (0.6, 1.0), (0.5, 2.0)]
    out = {}
    for kernel in _kernels(hip):
        b = hip.Batch([(m, ma, al) for m, (ma, al) in zip(meshes, flows)], order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
        for k, g in enumerate(hip.launch_groups(b)):
            if g["kernel"] in _kernels(hip):
                _set(hip, b, k, g, kernel)
                hip.set_launch(b, k, max(1, g["chains"] // 5), 2)
        b.batch_step(10, check_every=5)
        out[kernel] = [b.case_q(k).copy() for k in range(len(meshes))]
        b.close()
    for kernel in (hip.KERNEL_HALF, hip.KERNEL_HALF2):
        for k in range(len(meshes)):
            _eq(out[kernel][k], out[hip.KERNEL_ROW][k], f"kernel {kernel}, case {k}")


@pytest.mark.parametrize("retire", [False, True])
def test_half2_mixed_pairs(hip, retire):
    """Two chains per workgroup where a pair holds lines of different cases and
    lengths (63 lines of 301 rows, then 64 of 291): both live, or the first
    case retired, whose lines then drop out of their pairs; the results equal
    the one-row kernel's."""
    from hamstr_oop_amd import meshgen
    meshes = [meshgen.circle_ogrid(63, 301), meshgen.circle_ogrid(64, 291)]
    out = {}
    for kernel in (hip.KERNEL_ROW, hip.KERNEL_HALF2):
        b = hip.Batch([(m, 0.4, 1.0) for m in meshes], order=3, cfl=10.0, adi_mode=hip.ADI_FAST)
        for k, g in enumerate(hip.launch_groups(b)):
            if g["kernel"] in _kernels(hip):
                if g["final"] == 0:
                    hip.set_adi_kernel(b, k, kernel)
                hip.set_launch(b, k, g["grid"], 2)
        if retire:
            b.set_retire(True, 0)
            b.set_case_live(0, False)
        b.batch_step(6, check_every=3)
        out[kernel] = [b.case_q(k).copy() for k in range(2)]
        b.close()
    for k in range(2):
        _eq(out[hip.KERNEL_HALF2][k], out[hip.KERNEL_ROW][k], f"case {k}")


@pytest.mark.parametrize("lhs", ["fp64", "fp32"])
@pytest.mark.parametrize("order", [1, 3, 5])
def test_fused_step_equals_api_calls(hip, order, lhs):
    """The fused iteration's linear residual computes each face's Jacobians
    again from q instead of reading the stored ones: FAST iterations through
    hamstr_step equal computeRHS("implicit") + ADI() through the API (which
    reads them), bit for bit, on a mesh with orphan faces."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(64, 24, tri_layers=8)
    lm = hip.LHS_FP64 if lhs == "fp64" else hip.LHS_FP32
    a = hip.Solver(mesh, 0.5, 2.0, order, 10.0, lhs=lm, adi_mode=hip.ADI_FAST)
    b = hip.Solver(mesh, 0.5, 2.0, order, 10.0, lhs=lm, adi_mode=hip.ADI_FAST)
    a.step(5, check_every=5)
    for _ in range(5):
        b.computeRHS("implicit")
        b.ADI()
    _eq(a.get("q"), b.get("q"), "q after 5 iterations")
    _eq(a.get("r"), b.get("r"), "linear residual")
    a.close()
    b.close()


def test_half2_fp32_equals_row_kernel(hip):
    """FP32 left-hand side: two lines per workgroup (an autotuner option there)
    against the one-row kernel, 12 steps on the mesh with 384-row lines."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(512, 256, tri_layers=128)
    out = {}
    for kernel in (hip.KERNEL_ROW, hip.KERNEL_HALF2):
        s = hip.Solver(mesh, 0.5, 2.0, 3, 10.0, lhs=hip.LHS_FP32, adi_mode=hip.ADI_FAST)
        for k, g in enumerate(hip.launch_groups(s)):
            if g["final"] == 0 and g["max_rows"] == 384:
                hip.set_adi_kernel(s, k, kernel)
        with pytest.raises(hip.HamstrError):
            hip.set_adi_kernel(s, [g["final"] for g in hip.launch_groups(s)].index(0), hip.KERNEL_HALF)
        s.step(12, check_every=12)
        out[kernel] = s.get("q")
        s.close()
    _eq(out[hip.KERNEL_HALF2], out[hip.KERNEL_ROW], "FP32 q after 12 steps")



def test_fp32_odd_cycle_kernels_agree(hip):
    """FP32 odd cycles run the 128-VGPR kernel when a launch has more
    workgroups than CUs and the unbounded one when it fits one round: the
    same arithmetic, so a batch stepped with either grid ends bit-identical."""
    from hamstr_oop_amd import meshgen
    meshes = [meshgen.circle_ogrid(24, 10) for _ in range(40)]  # 400 loops of 24 rows: odd 3-cycles
    out = {}
    for grid in (64, 4096):  # one round on any MI355X (256 CUs), and many
        b = hip.Batch([(m, 0.3 + 0.005 * k, 1.0) for k, m in enumerate(meshes)], order=3, cfl=10.0,
                      lhs=hip.LHS_FP32, adi_mode=hip.ADI_FAST)
        for k, g in enumerate(hip.launch_groups(b)):
            if g["final"] == 2:
                assert g["chains"] == 400
                hip.set_launch(b, k, grid, 1)
        b.batch_step(8, check_every=4)
        out[grid] = [b.case_q(k).copy() for k in range(len(meshes))]
        b.close()
    a, c = out.values()
    for k in range(len(meshes)):
        _eq(a[k], c[k], f"case {k}")

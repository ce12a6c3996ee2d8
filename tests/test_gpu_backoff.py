"""The sweep's CFL back-off (hamstr_oop_amd/sweep.py): cases of the seeded
256-case table that settle into a limit cycle at CFL 10 (coarse 128x64 meshes
at low Mach) converge when continued at CFL 5, in the batched and the
sequential runs alike, and the converged state is the same fixed point."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# table indices: three cases that stall at CFL 10 and two that converge there
STALL = (16, 20, 31)
PLAIN = (0, 3)


@pytest.fixture(scope="module")
def hip():
    from hamstr_oop_amd import hip as H
    H.lib()
    return H


@pytest.fixture(scope="module")
def table():
    from hamstr_oop_amd import sweep
    full = sweep.draw_table(256)
    specs = [full[k] for k in STALL + PLAIN]
    for c in specs[:len(STALL)]:
        assert (c.nc, c.nr) == (128, 64)
    return specs, [c.mesh() for c in specs]


def test_batched_backoff_converges_every_case(hip, table):
    from hamstr_oop_amd import sweep
    specs, meshes = table
    print("batched, FP32 LHS, fallback and back-off", flush=True)
    res, out, info = sweep.run_batched(specs, meshes, 1e-10, 20000, 2000, lhs=hip.LHS_FP32, autotune=False)
    try:
        assert (res.status == 1).all(), (res.status, info)
        assert info["backoff_cfl5_converged"] >= 1, info
        # without the back-off the same cases end stalled
        res0, out0, info0 = sweep.run_batched(specs, meshes, 1e-10, 20000, 2000, lhs=hip.LHS_FP32,
                                              autotune=False, backoff=())
        try:
            assert (res0.status[:len(STALL)] == 2).all(), (res0.status, info0)
            assert (res0.status[len(STALL):] == 1).all()
            # the cases converged at CFL 10 are untouched by the back-off
            for k in range(len(STALL), len(specs)):
                np.testing.assert_array_equal(out.case_q(k), out0.case_q(k))
                assert res.iters[k] == res0.iters[k]
        finally:
            out0.close()
    finally:
        out.close()


def test_sequential_backoff_matches_batched(hip, table):
    from hamstr_oop_amd import sweep
    specs, meshes = table
    specs, meshes = specs[:2], meshes[:2]
    print("sequential FP64 with back-off against batched FP64", flush=True)
    seq = sweep.run_sequential(specs, meshes, 1e-10, 20000, 2000, lhs=hip.LHS_FP64)
    bat, out, info = sweep.run_batched(specs, meshes, 1e-10, 20000, 2000, lhs=hip.LHS_FP64, autotune=False)
    try:
        assert (seq.status == 1).all() and (bat.status == 1).all(), (seq.status, bat.status, info)
        # a batch steps each case as its own context would: the same
        # iterations in every pass
        np.testing.assert_array_equal(seq.iters, bat.iters)
    finally:
        out.close()

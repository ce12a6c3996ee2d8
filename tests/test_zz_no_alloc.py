"""hamstr_block_tridag4_noalloc, the block solve that allocates nothing (its
scratch rows live in a library-owned device array), against the allocating
entry points hamstr_block_tridag4 / hamstr_block_tridag_periodic4 on the same
systems: bit-identical, over two calls in a row on one GPU (the second one
smaller, so it runs on scratch the first one left behind), and sizes beyond
the array's rows are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from cases import random_block_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    """The library, and the HIP runtime it is linked with (libamdhip64.so.7,
    already loaded by it), for the caller's own device buffers."""
    from hamstr_oop_amd import hip
    hip.lib()
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    n = C.c_int(0)
    if rt.hipGetDeviceCount(C.byref(n)) != 0 or n.value < 1:
        pytest.skip("no GPU")
    return hip, rt


def _noalloc(hip, rt, periodic, a, b, c, f):
    """The caller's device buffers, the solve through the C ABI on the null
    stream, the solution back to the host."""
    host = [np.ascontiguousarray(v, np.float64) for v in (a, b, c, f)]
    dev = []
    try:
        for h in host:
            p = C.c_void_p()
            assert rt.hipMalloc(C.byref(p), h.nbytes) == 0
            dev.append(p)
            assert rt.hipMemcpy(p, h.ctypes.data, h.nbytes, 1) == 0          # host to device
        nsys, N = a.shape[0], a.shape[1]
        rc = hip.lib().hamstr_block_tridag4_noalloc(int(periodic), nsys, N, dev[0], dev[1], dev[2], dev[3], None)
        assert rt.hipDeviceSynchronize() == 0
        out = np.empty_like(host[3])
        assert rt.hipMemcpy(out.ctypes.data, dev[3], out.nbytes, 2) == 0     # device to host
        return rc, out
    finally:
        for p in dev:
            rt.hipFree(p)


@pytest.mark.parametrize("periodic", [False, True])
def test_noalloc_equals_allocating_entry_point(env, periodic):
    hip, rt = env
    rng = np.random.default_rng(77 + periodic)
    for nsys, N in ((48, 160), (7, 23)):   # 7680 block rows, then a smaller call
        print(f"noalloc periodic={periodic} nsys={nsys} N={N}", flush=True)
        a, b, c, f = random_block_system(rng, nsys, N)
        want = hip.block_tridag4(a, b, c, f, periodic=periodic)
        rc, got = _noalloc(hip, rt, periodic, a, b, c, f)
        assert rc == hip.OK, hip.lib().hamstr_last_error()
        assert np.array_equal(got, want), (nsys, N, np.abs(got - want).max())


def test_noalloc_refuses_sizes_beyond_its_scratch(env):
    hip, rt = env
    rng = np.random.default_rng(5)
    a, b, c, f = random_block_system(rng, 2, 8193)   # 16386 rows > 16384
    rc, got = _noalloc(hip, rt, False, a, b, c, f)
    assert rc == hip.ERR_ARG
    assert b"exceed" in hip.lib().hamstr_last_error()
    assert np.array_equal(got, f)   # nothing ran
    rc, _ = _noalloc(hip, rt, True, a[:, :2], b[:, :2], c[:, :2], f[:, :2])   # periodic needs N >= 3
    assert rc == hip.ERR_ARG

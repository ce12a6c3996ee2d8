"""AddressSanitizer and UndefinedBehaviorSanitizer over the CPU-side code.

* The oracle (oracle/_asan/liboracle.so, `make -C oracle asan`) runs its whole
  pinning suite: the golden vectors the reference wrote, the closed forms,
  the orphan-face restatement and the 3D building blocks.  The reference's
  own undefined behaviour (fluxRoe2D's fm[4] write and qr[4] read,
  interfaceFlux.cpp:138-139; WENO's write past f0, reconstruction.cpp:143,150)
  is what this guards the restatement against.
* The library's host code (hamstr_oop_amd/_asan/libhamstr_hip.so, host code
  instrumented, `python -m hamstr_oop_amd.build --sanitize`) builds its chain
  slot tables, accumulation lists, orphan tables, ADI phases and launch
  groups for generated, builder-chained and broken meshes and batches, and
  preprocesses meshes.  Without a device every hamstr_create still runs all
  of that before it reports HAMSTR_ERR_NODEV.
* The drop-in driver (sanitized build) reads the reference's input deck and
  mesh files, orphans.dat included.

Each run is a child process with ROCm clang's ASan runtime first in
LD_PRELOAD (the instrumented libraries are loaded into an uninstrumented
Python); any report fails it (halt on error, UBSan not recoverable).
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_ASAN = os.path.join(ROOT, "oracle", "_asan", "liboracle.so")
HIP_ASAN = os.path.join(ROOT, "hamstr_oop_amd", "_asan", "libhamstr_hip.so")
DRIVER_ASAN = os.path.join(ROOT, "hamstr_oop_amd", "_asan", "hamstr_mi355x")


def _runtime():
    rts = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not rts:
        pytest.skip("ROCm clang's ASan runtime is not installed")
    return rts[-1]


def _stale(target, sources):
    return not os.path.exists(target) or any(os.path.getmtime(f) > os.path.getmtime(target) for f in sources)


@pytest.fixture(scope="module")
def san_env():
    rt = _runtime()
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)  # incremental
    srcs = glob.glob(os.path.join(ROOT, "hamstr_oop_amd", "csrc", "*")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
    if _stale(HIP_ASAN, srcs) or _stale(DRIVER_ASAN, srcs):
        sys.path.insert(0, ROOT)
        from hamstr_oop_amd import build
        build.build(sanitize=True)
    env = dict(os.environ)
    pre = env.get("LD_PRELOAD")
    env["LD_PRELOAD"] = rt + (" " + pre if pre else "")
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0:exitcode=97"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1:exitcode=98"
    env["HAMSTR_ORACLE_LIB"] = ORACLE_ASAN
    env["HAMSTR_HIP_LIB"] = HIP_ASAN
    return env


def _pytest(env, *files):
    # the first check proves the runtime is live in the child
    probe = ("import ctypes, sys; "
             "assert hasattr(ctypes.CDLL(None), '__asan_report_load8'), 'ASan runtime not loaded'; "
             "import pytest; sys.exit(pytest.main(['-q', '-x', '-p', 'no:cacheprovider', *sys.argv[1:]]))")
    out = subprocess.run([sys.executable, "-c", probe, *files], cwd=ROOT, env=env, capture_output=True,
                         text=True, timeout=900)
    tail = (out.stdout + out.stderr)[-3000:]
    assert out.returncode == 0, tail
    assert "AddressSanitizer" not in tail and "runtime error" not in tail, tail


def test_oracle_suite_under_sanitizers(san_env):
    _pytest(san_env, "tests/test_oracle_golden.py", "tests/test_oracle_math.py", "tests/test_oracle_orphans.py",
            "tests/test_oracle_strand.py")


def test_library_host_code_under_sanitizers(san_env):
    _pytest(san_env, "tests/test_cpu_tables.py", "tests/test_cpu_lib.py")


def test_driver_readers_under_sanitizers(san_env, tmp_path):
    sys.path.insert(0, ROOT)
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(40, 16, tri_layers=6)
    meshgen.write_reference_case(mesh, str(tmp_path), 0.5, 2.0, 3, 3, 10.0)
    out = subprocess.run([DRIVER_ASAN, "--exact"], cwd=str(tmp_path), env=san_env, capture_output=True, text=True,
                         timeout=300)
    text = out.stdout + out.stderr
    assert "AddressSanitizer" not in text and "runtime error" not in text, text[-3000:]
    # without a device the driver stops at hamstr_create, after reading every file
    assert out.returncode == 0 or "no HIP device" in text, text[-3000:]

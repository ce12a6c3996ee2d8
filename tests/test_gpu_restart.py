"""Solution files (the nwrite / outform intent of solnBlock.h:23,26): a run
written out and restarted in a fresh context continues bit for bit, for every
scheme and for batches; files of another mesh, truncated files and foreign
files are refused; the C++ driver writes every nwrite steps and restarts."""
import os
import subprocess

import numpy as np
import pytest

from cases import SMALL_CASES

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def hip():
    from hamstr_oop_amd import hip as H
    H.lib()
    return H


@pytest.mark.parametrize("scheme,cfl", [("ADI", 10.0), ("euler", 0.8), ("rk3", 0.8)])
def test_restart_continues_bitwise(hip, tmp_path, scheme, cfl):
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(64, 24)
    path = str(tmp_path / "sol.bin")
    print(f"restart {scheme}", flush=True)
    a = hip.Solver(mesh, 0.6, 1.5, 3, cfl, adi_mode=hip.ADI_FAST)
    a.set_scheme(scheme)
    a.step(20, check_every=10)
    a.write_solution(path, 20)
    _, l2a, _ = a.step(15, check_every=5)
    qa = a.get("q")
    a.close()
    b = hip.Solver(mesh, 0.6, 1.5, 3, cfl, adi_mode=hip.ADI_FAST)
    b.set_scheme(scheme)
    assert b.read_solution(path) == 20
    _, l2b, _ = b.step(15, check_every=5)
    assert np.array_equal(b.get("q"), qa)
    assert np.array_equal(l2a, l2b)
    b.close()


def test_batch_restart_and_file_layout(hip, tmp_path):
    specs = list(SMALL_CASES.values())
    meshes = [mk() for mk, _, _ in specs]
    cases = [(m, mach, alpha) for m, (_, mach, alpha) in zip(meshes, specs)]
    path = str(tmp_path / "batch.bin")
    a = hip.Batch(cases, order=3, cfl=10.0)
    a.batch_step(12, check_every=4)
    a.write_solution(path, 12)
    q12 = a.get("q")
    a.batch_step(8, check_every=4)
    qa = a.get("q")
    b = hip.Batch(cases, order=3, cfl=10.0)
    assert b.read_solution(path) == 12
    b.batch_step(8, check_every=4)
    assert np.array_equal(b.get("q"), qa)
    # the documented layout
    raw = open(path, "rb").read()
    assert raw[:8] == b"HMSTRSOL"
    version, ncase = np.frombuffer(raw, np.int32, 2, 8)
    ncell, step = np.frombuffer(raw, np.int64, 2, 16)
    order, scheme = np.frombuffer(raw, np.int32, 2, 40)
    cfl = np.frombuffer(raw, np.float64, 1, 48)[0]
    assert (version, ncase, ncell, step, order, scheme, cfl) == (1, 3, a.n_cell, 12, 3, 0, 10.0)
    rec = np.frombuffer(raw, np.dtype([("mach", "<f8"), ("alpha", "<f8"), ("first", "<i8"), ("count", "<i8")]),
                        3, 64)
    assert np.allclose(rec["mach"], [s[1] for s in specs]) and np.allclose(rec["alpha"], [s[2] for s in specs])
    assert np.array_equal(rec["first"], a.cell_off[:-1]) and np.array_equal(rec["count"], np.diff(a.cell_off))
    q = np.frombuffer(raw, np.float64, 4 * a.n_cell, 64 + 32 * 3)
    assert np.array_equal(q, q12)
    a.close()
    b.close()


def test_foreign_files_are_refused(hip, tmp_path):
    from hamstr_oop_amd import meshgen
    a = hip.Solver(meshgen.naca_ogrid(32, 12), 0.5, 1.0, 3, 10.0)
    b = hip.Solver(meshgen.naca_ogrid(32, 12, t=0.10), 0.5, 1.0, 3, 10.0)  # same sizes, other coordinates
    c = hip.Solver(meshgen.naca_ogrid(48, 8), 0.5, 1.0, 3, 10.0)          # same cell count, other faces
    path = str(tmp_path / "a.bin")
    a.write_solution(path, 3)
    assert b.read_solution(path) == 3   # the fingerprint covers topology, not coordinates
    with pytest.raises(hip.HamstrError) as e:
        c.read_solution(path)
    assert e.value.code == hip.ERR_ARG and "another mesh" in str(e.value)
    raw = open(path, "rb").read()
    short = str(tmp_path / "short.bin")
    open(short, "wb").write(raw[:-8])
    with pytest.raises(hip.HamstrError, match="truncated"):
        a.read_solution(short)
    junk = str(tmp_path / "junk.bin")
    open(junk, "wb").write(b"\0" * 256)
    with pytest.raises(hip.HamstrError, match="not a solution file"):
        a.read_solution(junk)
    with pytest.raises(hip.HamstrError):
        a.read_solution(str(tmp_path / "missing.bin"))
    for s in (a, b, c):
        s.close()


def test_driver_nwrite_and_restart(hip, oracle, tmp_path):
    """input.hamstr nwrite=5, nsteps=10: the driver leaves the 10-step state
    in its solution file; a second run restarted from it for 5 more steps
    ends where 15 uninterrupted oracle steps end (EXACT mode)."""
    from hamstr_oop_amd import meshgen
    mesh = meshgen.naca_ogrid(32, 12)
    exe = os.path.join(ROOT, "hamstr_oop_amd", "hamstr_mi355x")
    d1 = tmp_path / "run1"
    meshgen.write_reference_case(mesh, str(d1), 0.6, 1.0, 3, 10, 10.0, nwrite=5)
    out = subprocess.run([exe, "--exact", "--solution", "s.bin"], cwd=str(d1), capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    d2 = tmp_path / "run2"
    meshgen.write_reference_case(mesh, str(d2), 0.6, 1.0, 3, 5, 10.0)
    out = subprocess.run([exe, "--exact", "--restart", str(d1 / "s.bin"), "--dump", "q.bin"], cwd=str(d2),
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    rows = [ln for ln in out.stdout.splitlines() if ln.strip().startswith("|") and "Iter" not in ln]
    assert [int(ln.split("|")[1]) for ln in rows] == [10, 11, 12, 13, 14]
    o = oracle.case(mesh, 0.6, 1.0, 3, 10.0)
    o.step(15)
    assert np.array_equal(np.fromfile(str(d2 / "q.bin"), dtype=np.float64), o.get("q"))

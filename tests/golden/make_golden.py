"""Writes the golden fixtures in this directory from the reference's own code.

Runs only where the reference tree is present (it is absent where the GPU tests run):
the reference harness in oracle/_ref/ is compiled from the reference's
sources by ``make -C oracle ref``.  Two builds exist, -O2 and -O0; order-5
(WENO) cases and the WENO unit vectors use -O0 because the reference's -O2 build crashes there (WENO
writes one element past its stack array f0, reconstruction.cpp:143,150), and
for orders 1 and 3 the script checks that -O0 and -O2 agree before writing.

Fixtures (numpy .npz, inputs and outputs only):
  unit.npz       per-function vectors: fluxRoe2D, jacobianRoe2D, MUSCL, WENO,
                 MUSCL_DERIV, WENO_DERIV, blockTridag4, blockTridagPeriodic4
  case_<name>.npz  a mesh (coordinates, faces, chains), its preprocess
                 output, a perturbed initial state, and computeRHS / ADI /
                 5-step results of the reference SOLVER; 5 steps of the
                 explicit schemes (stepType "euler" and "rk3", CFL 0.8) from
                 the same state
  unit3d.npz     the 3D / strand building blocks (compiled in the reference's
                 2D build): fluxRoe3D, jacobianRoe3D (5x5), blockTridag and
                 blockTridagPeriodic (5x5 blocks)
  msweep.npz     msweep = 2 and 3 (the sweep loop, ADI.cpp:54-252): one ADI
                 call and 5 steps from a perturbed state, circle12x6 and
                 naca16x8 at order 3 (and order 5 on naca16x8, -O0)
  c1.npz         the 128x64 circle O-grid (first order, M=0.3, alpha=0,
                 CFL 10): 500 steps of marchInTime from the freestream --
                 per-step nonlinear and printed (linear) L2 / Linf, q after
                 100 and 500 steps, and the reference's own seconds

Usage: make_golden.py [unit] [unit3d] [cases] [msweep] [c1]   (no argument: all)
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

from hamstr_oop_amd import meshgen  # noqa: E402
import oracle_py  # noqa: E402
from cases import (MSWEEP_CASES, mesh_digest, random_block_system, random_block_system5,  # noqa: E402
                   random_normals, random_normals3, random_states, random_states5)

THIRD = 0.3333333333333333
EXPLICIT_CFL = 0.8

CASES = [
    # name, mesh factory, Mach, alpha, order
    ("circle12x6_o1", lambda: meshgen.circle_ogrid(12, 6), 0.3, 0.0, 1),
    ("circle12x6_o3", lambda: meshgen.circle_ogrid(12, 6), 0.3, 0.0, 3),
    ("circle12x6_o5", lambda: meshgen.circle_ogrid(12, 6), 0.3, 0.0, 5),
    ("naca16x8_o3", lambda: meshgen.naca_ogrid(16, 8), 0.8, 1.25, 3),
    ("naca16x8_o5", lambda: meshgen.naca_ogrid(16, 8), 0.5, 2.0, 5),
]


def unit_vectors(ref: oracle_py.Reference, ref_weno: oracle_py.Reference) -> dict:
    rng = np.random.default_rng(20261203)
    out = {}
    n = 64
    nv = random_normals(rng, n)
    ql = random_states(rng, n)
    qr = random_states(rng, n)
    ql[:8] = qr[:8]  # equal states: the consistency case
    flux = np.zeros((n, 4))
    sr = np.zeros(n)
    lm = np.zeros((n, 4, 4))
    rm = np.zeros((n, 4, 4))
    for i in range(n):
        flux[i], sr[i] = ref.flux_roe2d(nv[i], ql[i], qr[i])
        lm[i], rm[i] = ref.jacobian_roe2d(nv[i], ql[i], qr[i])
    out.update(flux_n=nv, flux_ql=ql, flux_qr=qr, flux=flux, flux_sr=sr, jac_l=lm, jac_r=rm)

    imax = 19
    f = rng.uniform(0.5, 2.0, (4, imax, 4))
    dq = rng.uniform(-0.1, 0.1, (4, imax, 4))
    out["recon_f"] = f
    out["recon_dq"] = dq
    for qt in (0.0, 0.25):
        tag = "1" if qt == 0.0 else "3"
        L = np.zeros_like(f)
        R = np.zeros_like(f)
        DL = np.zeros_like(f)
        DR = np.zeros_like(f)
        for c in range(f.shape[0]):
            L[c], R[c] = ref.muscl(f[c], 1, imax - 1, THIRD, qt, 1e-10)
            DL[c], DR[c] = ref.muscl_deriv(f[c], dq[c], 1, imax - 1, THIRD, qt, 1e-10)
        out[f"muscl{tag}_l"], out[f"muscl{tag}_r"] = L, R
        out[f"muscld{tag}_l"], out[f"muscld{tag}_r"] = DL, DR
    L = np.zeros_like(f)
    R = np.zeros_like(f)
    DL = np.zeros_like(f)
    DR = np.zeros_like(f)
    for c in range(f.shape[0]):
        L[c], R[c] = ref_weno.weno(f[c], 2, imax - 1, 1e-10)
        DL[c], DR[c] = ref_weno.weno_deriv(f[c], dq[c], 2, imax - 1, 1e-10)
    out["weno_l"], out["weno_r"], out["wenod_l"], out["wenod_r"] = L, R, DL, DR

    for periodic in (False, True):
        tag = "per" if periodic else "tri"
        for N in (3, 6, 11):
            a, b, c, rhs = random_block_system(rng, 2, N)
            x = np.stack([ref.block_tridag4(a[s], b[s], c[s], rhs[s], periodic=periodic) for s in range(2)])
            out[f"{tag}{N}_a"], out[f"{tag}{N}_b"], out[f"{tag}{N}_c"] = a, b, c
            out[f"{tag}{N}_f"], out[f"{tag}{N}_x"] = rhs, x
    return out


def unit3d_vectors(ref: oracle_py.Reference) -> dict:
    rng = np.random.default_rng(20261204)
    out = {}
    n = 64
    nv = random_normals3(rng, n)
    ql = random_states5(rng, n)
    qr = random_states5(rng, n)
    ql[:8] = qr[:8]  # equal states: the consistency case
    flux = np.zeros((n, 5))
    sr = np.zeros(n)
    lm = np.zeros((n, 5, 5))
    rm = np.zeros((n, 5, 5))
    for i in range(n):
        flux[i], sr[i] = ref.flux_roe3d(nv[i], ql[i], qr[i])
        lm[i], rm[i] = ref.jacobian_roe3d(nv[i], ql[i], qr[i])
    out.update(flux_n=nv, flux_ql=ql, flux_qr=qr, flux=flux, flux_sr=sr, jac_l=lm, jac_r=rm)
    for periodic in (False, True):
        tag = "per" if periodic else "tri"
        for N in (3, 6, 11):
            a, b, c, rhs = random_block_system5(rng, 2, N)
            x = np.stack([ref.block_tridag5(a[s], b[s], c[s], rhs[s], periodic=periodic) for s in range(2)])
            out[f"{tag}{N}_a"], out[f"{tag}{N}_b"], out[f"{tag}{N}_c"] = a, b, c
            out[f"{tag}{N}_f"], out[f"{tag}{N}_x"] = rhs, x
    return out


def run_case(ref: oracle_py.Reference, mesh, mach, alpha, order, workdir) -> dict:
    d = tempfile.mkdtemp(dir=workdir)
    meshgen.write_reference_case(mesh, d, mach, alpha, order, 5, 10.0)
    rc = ref.load(d)
    out = {
        "x": mesh.x, "faces_in": mesh.faces, "chain_start": mesh.chain_start, "chain_conn": mesh.chain_conn,
        "mach": mach, "alpha": alpha, "order": order, "cfl": 10.0,
        "faces": rc.get_ints("faces").reshape(-1, 6), "normvec": rc.get("normvec").reshape(-1, 2),
        "refmtx": rc.get("refmtx").reshape(-1, 4), "q_freestream": rc.get("q"),
    }
    rng = np.random.default_rng(order * 101 + mesh.n_cell)
    q = rc.get("q").reshape(-1, 4).copy()
    n = q.shape[0]
    q[:, 0] *= 1 + 0.02 * rng.uniform(-1, 1, n)
    q[:, 1] += 0.02 * rng.uniform(-1, 1, n)
    q[:, 2] += 0.02 * rng.uniform(-1, 1, n)
    q[:, 3] *= 1 + 0.02 * rng.uniform(-1, 1, n)
    q = q.reshape(-1)
    rc.set_q(q)
    out["q0"] = q
    l2, li = rc.compute_rhs(True)
    out.update(rhs_r=rc.get("r"), rhs_sigma=rc.get("sigma"), rhs_ff=rc.get("ff"), rhs_l2=l2, rhs_linf=li)
    l2, li = rc.adi()
    out.update(adi_q=rc.get("q"), adi_dq=rc.get("dq"), adi_r=rc.get("r"), adi_r0=rc.get("r0"),
               adi_l2=l2, adi_linf=li)
    l2r, l2l, _ = rc.step(5)
    out.update(step5_q=rc.get("q"), step5_l2rhs=l2r, step5_l2lin=l2l)
    shutil.rmtree(d, ignore_errors=True)
    for scheme in ("euler", "rk3"):
        d = tempfile.mkdtemp(dir=workdir)
        meshgen.write_reference_case(mesh, d, mach, alpha, order, 5, EXPLICIT_CFL, scheme=scheme)
        rc = ref.load(d)
        rc.set_q(q)
        l2, li, _ = rc.step_solution(5)
        out.update({f"{scheme}5_q": rc.get("q"), f"{scheme}5_qt": rc.get("qt"), f"{scheme}5_r": rc.get("r"),
                    f"{scheme}5_sigma": rc.get("sigma"), f"{scheme}5_l2": l2, f"{scheme}5_linf": li})
        shutil.rmtree(d, ignore_errors=True)
    out["explicit_cfl"] = EXPLICIT_CFL
    return out


def msweep_case(ref: oracle_py.Reference, mesh, mach, alpha, order, msweep, workdir) -> dict:
    """One ADI call and 5 steps at msweep sweeps from a perturbed state."""
    d = tempfile.mkdtemp(dir=workdir)
    meshgen.write_reference_case(mesh, d, mach, alpha, order, 5, 10.0, msweep=msweep)
    rc = ref.load(d)
    assert rc.msweep == msweep
    rng = np.random.default_rng(1000 * msweep + order * 101 + mesh.n_cell)
    q = rc.get("q").reshape(-1, 4).copy()
    n = q.shape[0]
    q[:, 0] *= 1 + 0.02 * rng.uniform(-1, 1, n)
    q[:, 1] += 0.02 * rng.uniform(-1, 1, n)
    q[:, 2] += 0.02 * rng.uniform(-1, 1, n)
    q[:, 3] *= 1 + 0.02 * rng.uniform(-1, 1, n)
    q = q.reshape(-1)
    rc.set_q(q)
    out = {"q0": q}
    l2, li = rc.compute_rhs(True)
    out.update(rhs_l2=l2, rhs_linf=li)
    l2, li = rc.adi()
    out.update(adi_q=rc.get("q"), adi_dq=rc.get("dq"), adi_r=rc.get("r"), adi_r0=rc.get("r0"),
               adi_l2=l2, adi_linf=li)
    l2r, l2l, _ = rc.step(5)
    out.update(step5_q=rc.get("q"), step5_dq=rc.get("dq"), step5_l2rhs=l2r, step5_l2lin=l2l)
    shutil.rmtree(d, ignore_errors=True)
    return out


def msweep_fixture(ref2, ref0, work) -> dict:
    out = {}
    for name, (make, mach, alpha, order) in MSWEEP_CASES.items():
        mesh = make()
        for msweep in (2, 3):
            data0 = msweep_case(ref0, mesh, mach, alpha, order, msweep, work)
            if order != 5:
                data2 = msweep_case(ref2, mesh, mach, alpha, order, msweep, work)
                for k in data0:
                    if not np.array_equal(data0[k], data2[k]):
                        raise SystemExit(f"msweep {name} {msweep}: reference -O2 and -O0 disagree on {k}")
            for k, v in data0.items():
                out[f"{name}_s{msweep}_{k}"] = v
            print("msweep", name, msweep, flush=True)
    return out


C1 = dict(nc=128, nr=64, mach=0.3, alpha=0.0, order=1, cfl=10.0, steps=500)


def c1_run(opt: str, out_path: str) -> None:
    """Config 1 through the reference build `opt`, in a process of its own: the
    reference's Jacobians leak about 12 GB over 500 steps of this mesh."""
    ref = oracle_py.Reference(opt)
    mesh = meshgen.circle_ogrid(C1["nc"], C1["nr"])
    work = tempfile.mkdtemp()
    meshgen.write_reference_case(mesh, work, C1["mach"], C1["alpha"], C1["order"], C1["steps"], C1["cfl"])
    rc = ref.load(work)
    l2r1, l2l1, s1 = rc.step(100)
    q100 = rc.get("q")
    l2r2, l2l2, s2 = rc.step(C1["steps"] - 100)
    np.savez_compressed(out_path, q100=q100, q500=rc.get("q"), l2rhs=np.concatenate([l2r1, l2r2]),
                        l2lin=np.concatenate([l2l1, l2l2]), seconds=s1 + s2)
    shutil.rmtree(work, ignore_errors=True)


def c1_fixture() -> dict:
    work = tempfile.mkdtemp()
    res = {}
    for opt in ("O2", "O0"):
        path = os.path.join(work, f"c1_{opt}.npz")
        subprocess.run([sys.executable, os.path.abspath(__file__), "_c1run", opt, path], check=True)
        res[opt] = dict(np.load(path))
    for k in ("q100", "q500", "l2rhs", "l2lin"):
        if not np.array_equal(res["O2"][k], res["O0"][k]):
            raise SystemExit(f"c1: reference -O2 and -O0 disagree on {k}")
    mesh = meshgen.circle_ogrid(C1["nc"], C1["nr"])
    out = {k: res["O2"][k] for k in ("q100", "q500", "l2rhs", "l2lin")}
    out.update({k: np.array(v) for k, v in C1.items()})
    out["mesh_sha256"] = np.array(mesh_digest(mesh))
    out["ref_seconds_O2"] = res["O2"]["seconds"]
    shutil.rmtree(work, ignore_errors=True)
    print(f"c1: reference -O2 took {float(res['O2']['seconds']):.2f} s for {C1['steps']} steps", flush=True)
    return out


def main(what):
    if not (oracle_py.Reference.available("O2") and oracle_py.Reference.available("O0")):
        oracle_py.build(reference=True)
    ref2 = oracle_py.Reference("O2")
    ref0 = oracle_py.Reference("O0")
    work = tempfile.mkdtemp()
    if "unit" in what:
        u2 = unit_vectors(ref2, ref0)
        u0 = unit_vectors(ref0, ref0)
        for k in u2:
            if not np.array_equal(u2[k], u0[k]):
                raise SystemExit(f"reference -O2 and -O0 disagree on {k}")
        np.savez_compressed(os.path.join(HERE, "unit.npz"), **u2)
    if "unit3d" in what:
        u2 = unit3d_vectors(ref2)
        u0 = unit3d_vectors(ref0)
        for k in u2:
            if not np.array_equal(u2[k], u0[k]):
                raise SystemExit(f"reference -O2 and -O0 disagree on {k} (3D)")
        np.savez_compressed(os.path.join(HERE, "unit3d.npz"), **u2)
    if "cases" in what:
        for name, make, mach, alpha, order in CASES:
            mesh = make()
            data0 = run_case(ref0, mesh, mach, alpha, order, work)
            if order != 5:
                data2 = run_case(ref2, mesh, mach, alpha, order, work)
                for k in data0:
                    if not np.array_equal(data0[k], data2[k]):
                        raise SystemExit(f"{name}: reference -O2 and -O0 disagree on {k}")
            np.savez_compressed(os.path.join(HERE, f"case_{name}.npz"), **data0)
            print("wrote", name, flush=True)
    if "msweep" in what:
        np.savez_compressed(os.path.join(HERE, "msweep.npz"), **msweep_fixture(ref2, ref0, work))
    if "c1" in what:
        np.savez_compressed(os.path.join(HERE, "c1.npz"), **c1_fixture())
    shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "_c1run":
        c1_run(sys.argv[2], sys.argv[3])
    else:
        main(set(sys.argv[1:]) or {"unit", "unit3d", "cases", "msweep", "c1"})

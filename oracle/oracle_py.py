"""ctypes bindings for the CPU oracle (liboracle.so) and the reference harness.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the oracle.  The GPU product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)

FIELD_IDS = {"q": 0, "r": 1, "r0": 2, "dq": 3, "sigma": 4, "ff": 5, "qt": 6}
SCHEMES = {"ADI": 0, "euler": 1, "rk3": 2}


def _d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    return a.ctypes.data_as(_ip)


def build(reference: bool = False) -> None:
    """Compile liboracle.so (and the reference harness when asked and present)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if reference and os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _load(path: str) -> C.CDLL:
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run make -C oracle)")
    return C.CDLL(path)


class _Unit:
    """Per-function entry points shared by the oracle and the reference harness."""

    prefix = ""

    def __init__(self, lib: C.CDLL, prefix: str):
        self.lib = lib
        self.prefix = prefix

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def flux_roe2d(self, n, ql, qr, gamma=1.4):
        n = np.ascontiguousarray(n, np.float64)
        ql = np.ascontiguousarray(ql, np.float64)
        qr = np.ascontiguousarray(qr, np.float64)
        flux = np.zeros(4)
        sr = C.c_double(0.0)
        self.fn("flux_roe2d")(_d(n), _d(ql), _d(qr), _d(flux), C.byref(sr), C.c_double(gamma))
        return flux, sr.value

    def jacobian_roe2d(self, n, ql, qr, gamma=1.4):
        n = np.ascontiguousarray(n, np.float64)
        ql = np.ascontiguousarray(ql, np.float64)
        qr = np.ascontiguousarray(qr, np.float64)
        lm = np.zeros(16)
        rm = np.zeros(16)
        self.fn("jacobian_roe2d")(_d(n), _d(ql), _d(qr), _d(lm), _d(rm), C.c_double(gamma))
        return lm.reshape(4, 4), rm.reshape(4, 4)

    def muscl(self, f, is_, ie, th, qt, eps):
        f = np.ascontiguousarray(f, np.float64)
        imax = f.shape[0]
        ql = np.zeros_like(f)
        qr = np.zeros_like(f)
        self.fn("muscl")(_d(f), _d(ql), _d(qr), is_, ie, C.c_double(th), C.c_double(qt), C.c_double(eps), imax)
        return ql, qr

    def weno(self, f, is_, ie, eps):
        f = np.ascontiguousarray(f, np.float64)
        imax = f.shape[0]
        ql = np.zeros_like(f)
        qr = np.zeros_like(f)
        self.fn("weno")(_d(f), _d(ql), _d(qr), is_, ie, C.c_double(eps), imax)
        return ql, qr

    def muscl_deriv(self, f, dq, is_, ie, th, qt, eps):
        f = np.ascontiguousarray(f, np.float64)
        dq = np.ascontiguousarray(dq, np.float64)
        imax = f.shape[0]
        ql = np.zeros_like(f)
        qr = np.zeros_like(f)
        self.fn("muscl_deriv")(_d(f), _d(dq), _d(ql), _d(qr), is_, ie, C.c_double(th), C.c_double(qt),
                               C.c_double(eps), imax)
        return ql, qr

    def weno_deriv(self, f, dq, is_, ie, eps):
        f = np.ascontiguousarray(f, np.float64)
        dq = np.ascontiguousarray(dq, np.float64)
        imax = f.shape[0]
        ql = np.zeros_like(f)
        qr = np.zeros_like(f)
        self.fn("weno_deriv")(_d(f), _d(dq), _d(ql), _d(qr), is_, ie, C.c_double(eps), imax)
        return ql, qr

    def block_tridag4(self, a, b, c, f, periodic=False):
        a = np.array(a, np.float64, copy=True, order="C")
        b = np.array(b, np.float64, copy=True, order="C")
        c = np.array(c, np.float64, copy=True, order="C")
        f = np.array(f, np.float64, copy=True, order="C")
        n = f.shape[0]
        name = "block_tridag_periodic4" if periodic else "block_tridag4"
        self.fn(name)(_d(a), _d(b), _d(c), _d(f), n)
        return f


    # ---- the 3D / strand building blocks (5 conserved variables) ----
    def flux_roe3d(self, n, ql, qr, gamma=1.4):
        n = np.ascontiguousarray(n, np.float64)
        ql = np.ascontiguousarray(ql, np.float64)
        qr = np.ascontiguousarray(qr, np.float64)
        flux = np.zeros(5)
        sr = C.c_double(0.0)
        self.fn("flux_roe3d")(_d(n), _d(ql), _d(qr), _d(flux), C.byref(sr), C.c_double(gamma))
        return flux, sr.value

    def jacobian_roe3d(self, n, ql, qr, gamma=1.4):
        n = np.ascontiguousarray(n, np.float64)
        ql = np.ascontiguousarray(ql, np.float64)
        qr = np.ascontiguousarray(qr, np.float64)
        lm = np.zeros(25)
        rm = np.zeros(25)
        self.fn("jacobian_roe3d")(_d(n), _d(ql), _d(qr), _d(lm), _d(rm), C.c_double(gamma))
        return lm.reshape(5, 5), rm.reshape(5, 5)

    def block_tridag5(self, a, b, c, f, periodic=False):
        a = np.array(a, np.float64, copy=True, order="C")
        b = np.array(b, np.float64, copy=True, order="C")
        c = np.array(c, np.float64, copy=True, order="C")
        f = np.array(f, np.float64, copy=True, order="C")
        name = "block_tridag_periodic5" if periodic else "block_tridag5"
        self.fn(name)(_d(a), _d(b), _d(c), _d(f), f.shape[0])
        return f


class Oracle(_Unit):
    """The CPU restatement (liboracle.so)."""

    def __init__(self):
        # HAMSTR_ORACLE_LIB selects another build (the sanitizer build, _asan/)
        lib = _load(os.environ.get("HAMSTR_ORACLE_LIB") or os.path.join(HERE, "liboracle.so"))
        lib.orc_create.restype = C.c_void_p
        for name in ("orc_destroy", "orc_get", "orc_set", "orc_compute_rhs", "orc_adi",
                     "orc_compute_linear_rhs", "orc_step", "orc_freestream"):
            getattr(lib, name).argtypes = None
        super().__init__(lib, "orc_")

    def preprocess(self, x, faces, sol_visc=0):
        x = np.ascontiguousarray(x, np.float64)
        faces = np.array(faces, np.int32, copy=True, order="C")
        nf = faces.shape[0]
        nv = np.zeros((nf, 2))
        rm = np.zeros((nf, 4))
        self.lib.orc_preprocess(x.shape[0], _d(x), nf, _i(faces), _d(nv), _d(rm), sol_visc)
        return faces, nv, rm

    def case(self, mesh, mach, alpha, order, cfl, msweep=1, sol_visc=0):
        return OracleCase(self, mesh, mach, alpha, order, cfl, msweep, sol_visc)


class OracleCase:
    def __init__(self, orc: Oracle, mesh, mach, alpha, order, cfl, msweep=1, sol_visc=0):
        self.orc = orc
        self.lib = orc.lib
        faces, nv, rm = orc.preprocess(mesh.x, mesh.faces, sol_visc)
        self.faces, self.normvec, self.refmtx = faces, nv, rm
        self.n_cell = mesh.n_cell
        self.n_face = mesh.n_face
        start = np.ascontiguousarray(mesh.chain_start, np.int32)
        cc = np.ascontiguousarray(mesh.chain_conn, np.int32)
        orph = np.ascontiguousarray(getattr(mesh, "orphans", np.zeros(0, np.int32)), np.int32)
        if orph.size and ((faces[orph, 2] < 0).any() or (faces[orph, 4] < 0).any()):
            raise ValueError("orphan faces must be interior faces")
        self._keep = (start, cc, orph)
        self.h = C.c_void_p(self.lib.orc_create(
            mesh.n_cell, mesh.n_face, mesh.n_chain, _i(faces), _i(start), _i(cc), _d(nv), _d(rm),
            C.c_double(mach), C.c_double(alpha), order, C.c_double(cfl), msweep, int(orph.size),
            _i(orph) if orph.size else None))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def _size(self, name):
        return {"sigma": self.n_cell, "ff": 32 * self.n_face}.get(name, 4 * self.n_cell)

    def get(self, name):
        out = np.zeros(self._size(name))
        self.lib.orc_get(self.h, FIELD_IDS[name], _d(out))
        return out

    def set(self, name, val):
        val = np.ascontiguousarray(val, np.float64).reshape(-1)
        assert val.size == self._size(name)
        self.lib.orc_set(self.h, FIELD_IDS[name], _d(val))

    def freestream(self):
        out = np.zeros(4)
        self.lib.orc_freestream(self.h, _d(out))
        return out

    def compute_rhs(self, implicit=True):
        l2, li = C.c_double(), C.c_double()
        rc = self.lib.orc_compute_rhs(self.h, int(implicit), C.byref(l2), C.byref(li))
        return rc, l2.value, li.value

    def adi(self):
        l2, li = C.c_double(), C.c_double()
        self.lib.orc_adi(self.h, C.byref(l2), C.byref(li))
        return l2.value, li.value

    def compute_linear_rhs(self):
        l2, li = C.c_double(), C.c_double()
        self.lib.orc_compute_linear_rhs(self.h, C.byref(l2), C.byref(li))
        return l2.value, li.value

    def step(self, n):
        l2r = np.zeros(max(n, 1))
        l2l = np.zeros(max(n, 1))
        sec = C.c_double()
        rc = self.lib.orc_step(self.h, n, _d(l2r), _d(l2l), C.byref(sec))
        return rc, l2r[:n], l2l[:n], sec.value

    def step_scheme(self, scheme, n):
        """n x stepSolution with stepType scheme ("ADI", "euler", "rk3"):
        (rc, first-computeRHS L2 per step, printed L2 per step, seconds)."""
        first = np.zeros(max(n, 1))
        printed = np.zeros(max(n, 1))
        sec = C.c_double()
        rc = self.lib.orc_step_scheme(self.h, SCHEMES[scheme], n, _d(first), _d(printed), C.byref(sec))
        return rc, first[:n], printed[:n], sec.value

    def update_solution(self, src, dst, coef):
        rc = self.lib.orc_update_solution(self.h, FIELD_IDS[src], FIELD_IDS[dst], C.c_double(coef))
        assert rc == 0


class Reference(_Unit):
    """The reference's own code, compiled from its sources by oracle/Makefile."""

    def __init__(self, opt: str = "O2"):
        lib = _load(os.path.join(REF_DIR, f"libhamstr_ref_{opt}.so"))
        lib.ref_load.restype = C.c_void_p
        lib.ref_load.argtypes = [C.c_char_p]
        super().__init__(lib, "ref_")

    @staticmethod
    def available(opt: str = "O2") -> bool:
        return os.path.exists(os.path.join(REF_DIR, f"libhamstr_ref_{opt}.so"))

    def load(self, case_dir: str):
        return ReferenceCase(self, case_dir)


class ReferenceCase:
    def __init__(self, ref: Reference, case_dir: str):
        self.lib = ref.lib
        h = self.lib.ref_load(case_dir.encode())
        if not h:
            raise RuntimeError(f"reference failed to load {case_dir}")
        self.h = C.c_void_p(h)
        cnt = np.zeros(9, np.int32)
        self.lib.ref_counts(self.h, _i(cnt))
        (self.n_cell, self.n_face, self.n_chain, self.n_chain_face, self.n_max_chain,
         self.n_node, self.order, self.msweep, self.nsteps) = [int(v) for v in cnt]

    def params(self):
        out = np.zeros(8)
        self.lib.ref_params(self.h, _d(out))
        return dict(zip(["mach", "alpha", "cfl", "gamma", "rinf", "uinf", "vinf", "einf"], out))

    def _size(self, fid):
        return {4: self.n_cell, 5: 32 * self.n_face, 6: 2 * self.n_face, 7: 4 * self.n_face}.get(fid, 4 * self.n_cell)

    def get(self, name):
        fid = {**FIELD_IDS, "normvec": 6, "refmtx": 7, "qt": 8}[name]
        out = np.zeros(self._size(fid))
        self.lib.ref_get(self.h, fid, _d(out))
        return out

    def get_ints(self, name):
        fid = {"faces": 0, "chain_start": 1, "chain_conn": 2}[name]
        n = {0: 6 * self.n_face, 1: self.n_chain + 1, 2: self.n_chain_face}[fid]
        out = np.zeros(n, np.int32)
        self.lib.ref_get_ints(self.h, fid, _i(out))
        return out

    def set_q(self, q):
        q = np.ascontiguousarray(q, np.float64).reshape(-1)
        assert q.size == 4 * self.n_cell
        self.lib.ref_set_q(self.h, _d(q))

    def compute_rhs(self, implicit=True):
        l2, li = C.c_double(), C.c_double()
        self.lib.ref_compute_rhs(self.h, int(implicit), C.byref(l2), C.byref(li))
        return l2.value, li.value

    def adi(self):
        l2, li = C.c_double(), C.c_double()
        self.lib.ref_adi(self.h, C.byref(l2), C.byref(li))
        return l2.value, li.value

    def compute_linear_rhs(self):
        l2, li = C.c_double(), C.c_double()
        self.lib.ref_compute_linear_rhs(self.h, C.byref(l2), C.byref(li))
        return l2.value, li.value

    def step(self, n):
        l2r = np.zeros(max(n, 1))
        l2l = np.zeros(max(n, 1))
        sec = C.c_double()
        self.lib.ref_step(self.h, n, _d(l2r), _d(l2l), C.byref(sec))
        return l2r[:n], l2l[:n], sec.value

    def set_scheme(self, name):
        assert self.lib.ref_set_scheme(self.h, name.encode()) == 0

    def step_solution(self, n):
        """n x SOLVER::stepSolution with the current scheme: (printed L2,
        printed Linf, seconds) per step."""
        l2 = np.zeros(max(n, 1))
        li = np.zeros(max(n, 1))
        sec = C.c_double()
        self.lib.ref_step_solution(self.h, n, _d(l2), _d(li), C.byref(sec))
        return l2[:n], li[:n], sec.value

    def update_solution(self, src, dst, coef):
        ids = {"q": 0, "qt": 8}
        assert self.lib.ref_update_solution(self.h, ids[src], ids[dst], C.c_double(coef)) == 0

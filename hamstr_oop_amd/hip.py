"""ctypes binding of the C ABI in include/hamstr_hip.h (libhamstr_hip.so).

``Solver`` mirrors the reference's SOLVER call surface (solver.h:14-55):
``initFlow``, ``computeRHS``, ``ADI``, ``computeLinearRHS``, ``stepSolution``,
``marchInTime``, plus field access.  The library must be built
(``python -m hamstr_oop_amd.build``); loading fails loudly otherwise -- there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# HAMSTR_HIP_LIB selects another build of the same library (the sanitizer
# build of the host code, _asan/libhamstr_hip.so, in tests/test_sanitizers.py)
LIB_PATH = os.environ.get("HAMSTR_HIP_LIB") or os.path.join(HERE, "libhamstr_hip.so")

OK, ERR_ARG, ERR_HIP, ERR_NAN, ERR_MESH, ERR_NODEV = range(6)
LHS_FP64, LHS_FP32 = 0, 1
ADI_EXACT, ADI_FAST = 0, 1
# ADI kernels of a launch group (HAMSTR_KERNEL_*)
KERNEL_EXACT, KERNEL_ROW, KERNEL_HALF, KERNEL_TWO, KERNEL_HALF2 = range(5)
FIELDS = {"q": 0, "r": 1, "r0": 2, "dq": 3, "sigma": 4, "ff": 5, "qt": 6}
# SOLVER::stepType values (the input deck's scheme= line)
SCHEMES = {"ADI": 0, "euler": 1, "rk3": 2}

# every entry point declared in include/hamstr_hip.h
EXPORTS = [
    "hamstr_preprocess", "hamstr_create", "hamstr_destroy", "hamstr_init_flow", "hamstr_freestream",
    "hamstr_set_field", "hamstr_get_field", "hamstr_compute_rhs", "hamstr_adi",
    "hamstr_compute_linear_rhs", "hamstr_step", "hamstr_last_norms", "hamstr_sync", "hamstr_set_timing",
    "hamstr_last_timings", "hamstr_last_timing_names", "hamstr_flux_roe2d", "hamstr_jacobian_roe2d", "hamstr_muscl", "hamstr_weno",
    "hamstr_muscl_deriv", "hamstr_weno_deriv", "hamstr_block_tridag4", "hamstr_block_tridag_periodic4",
    "hamstr_last_error", "hamstr_batch_create", "hamstr_batch_info", "hamstr_batch_step", "hamstr_autotune",
    "hamstr_launch_groups", "hamstr_case_freestream", "hamstr_set_retire", "hamstr_case_status",
    "hamstr_set_graph", "hamstr_get_cells", "hamstr_set_cells", "hamstr_prepare_graph",
    "hamstr_set_launch", "hamstr_set_adi_kernel", "hamstr_set_case_live", "hamstr_iterations",
    "hamstr_mlp_create", "hamstr_mlp_destroy", "hamstr_mlp_num_params", "hamstr_mlp_get_params",
    "hamstr_mlp_set_params", "hamstr_mlp_forward", "hamstr_mlp_loss_grad", "hamstr_mlp_set_data",
    "hamstr_mlp_train", "hamstr_warm_start", "hamstr_mlp_last_error", "hamstr_mlp_set_graph",
    "hamstr_update_solution", "hamstr_set_scheme", "hamstr_get_scheme",
    "hamstr_write_solution", "hamstr_read_solution", "hamstr_block_tridag4_noalloc",
    "hamstr_flux_roe3d", "hamstr_jacobian_roe3d", "hamstr_block_tridag5", "hamstr_block_tridag_periodic5",
    "hamstr_set_jacobian_source", "hamstr_jacobian_source", "hamstr_set_thomas", "hamstr_thomas",
    "hamstr_set_face_staging", "hamstr_face_staging",
]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class HamstrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hamstr error {code}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [("nCell", C.c_int), ("nFace", C.c_int), ("nChain", C.c_int),
                ("faces", _ip), ("chainStart", _ip), ("chainConn", _ip),
                ("normVec", _dp), ("refMtx", _dp), ("nOrphan", C.c_int), ("orphans", _ip)]


class Params(C.Structure):
    _fields_ = [("mach", C.c_double), ("alpha", C.c_double), ("cfl", C.c_double),
                ("order", C.c_int), ("msweep", C.c_int), ("lhs", C.c_int), ("adi_mode", C.c_int)]


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is not built: run python -m hamstr_oop_amd.build")
        L = C.CDLL(LIB_PATH)
        L.hamstr_last_error.restype = C.c_char_p
        L.hamstr_iterations.restype = C.c_long
        L.hamstr_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(C.c_void_p)]
        L.hamstr_batch_create.argtypes = [C.c_int, C.POINTER(MeshDesc), C.POINTER(Params), C.POINTER(C.c_void_p)]
        for name in EXPORTS:
            getattr(L, name)  # raises if missing
        _lib = L
    return _lib


def _chk(rc: int) -> None:
    if rc != OK:
        raise HamstrError(rc, lib().hamstr_last_error().decode())


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _orphans(mesh) -> np.ndarray:
    """The mesh's orphan faces (interior faces on no chain), or an empty array."""
    return np.ascontiguousarray(getattr(mesh, "orphans", np.zeros(0, np.int32)), np.int32)


def preprocess(x: np.ndarray, faces: np.ndarray, sol_visc: int = 0):
    """MESHBLOCK::preprocess: returns (faces with wall tags, normVec, refMtx)."""
    x = np.ascontiguousarray(x, np.float64)
    faces = np.array(faces, np.int32, copy=True, order="C")
    nf = faces.shape[0]
    nv = np.zeros((nf, 2))
    rm = np.zeros((nf, 4))
    _chk(lib().hamstr_preprocess(x.shape[0], _d(x), nf, _i(faces), _d(nv), _d(rm), sol_visc))
    return faces, nv, rm


class Solver:
    """A device context for one mesh/flow case (SOLVER + SOLNBLOCK state)."""

    def __init__(self, mesh, mach: float, alpha: float, order: int, cfl: float, msweep: int = 1,
                 lhs: int = LHS_FP64, adi_mode: int = ADI_EXACT, sol_visc: int = 0):
        self.mesh = mesh
        faces, nv, rm = preprocess(mesh.x, mesh.faces, sol_visc)
        self.faces, self.normvec, self.refmtx = faces, nv, rm
        self.n_cell = mesh.n_cell
        self.n_face = mesh.n_face
        orph = _orphans(mesh)
        self._keep = [faces, nv, rm, np.ascontiguousarray(mesh.chain_start, np.int32),
                      np.ascontiguousarray(mesh.chain_conn, np.int32), orph]
        md = MeshDesc(mesh.n_cell, mesh.n_face, mesh.n_chain, _i(faces), _i(self._keep[3]),
                      _i(self._keep[4]), _d(nv), _d(rm), int(orph.size), _i(orph) if orph.size else None)
        prm = Params(float(mach), float(alpha), float(cfl), int(order), int(msweep), int(lhs), int(adi_mode))
        h = C.c_void_p()
        _chk(lib().hamstr_create(C.byref(md), C.byref(prm), C.byref(h)))
        self.h = h
        self.order, self.cfl, self.msweep, self.lhs = order, cfl, msweep, lhs

    def close(self):
        if getattr(self, "h", None):
            lib().hamstr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- fields -------------------------------------------------------------
    def _size(self, name):
        return {"sigma": self.n_cell, "ff": 32 * self.n_face}.get(name, 4 * self.n_cell)

    def get(self, name: str) -> np.ndarray:
        out = np.zeros(self._size(name))
        _chk(lib().hamstr_get_field(self.h, FIELDS[name], _d(out)))
        return out

    def set(self, name: str, val) -> None:
        val = np.ascontiguousarray(val, np.float64).reshape(-1)
        if val.size != self._size(name):
            raise ValueError(f"{name}: expected {self._size(name)} values, got {val.size}")
        _chk(lib().hamstr_set_field(self.h, FIELDS[name], _d(val)))

    def freestream(self) -> np.ndarray:
        out = np.zeros(4)
        _chk(lib().hamstr_freestream(self.h, _d(out)))
        return out

    def write_solution(self, path: str, step: int = 0) -> None:
        """Solution file of q (every case), the step count and a mesh fingerprint."""
        _chk(lib().hamstr_write_solution(self.h, os.fsencode(path), C.c_long(step)))

    def read_solution(self, path: str) -> int:
        """Loads q from a solution file of this mesh (or batch); returns its step."""
        step = C.c_long()
        _chk(lib().hamstr_read_solution(self.h, os.fsencode(path), C.byref(step)))
        return step.value

    # ---- SOLVER surface -----------------------------------------------------
    def initFlow(self):
        _chk(lib().hamstr_init_flow(self.h))

    def computeRHS(self, option: str = "implicit"):
        if option not in ("implicit", "explicit"):
            raise ValueError("option must be 'implicit' or 'explicit'")
        l2, li = C.c_double(), C.c_double()
        _chk(lib().hamstr_compute_rhs(self.h, int(option == "implicit"), C.byref(l2), C.byref(li)))
        return l2.value, li.value

    def ADI(self):
        l2, li = C.c_double(), C.c_double()
        _chk(lib().hamstr_adi(self.h, C.byref(l2), C.byref(li)))
        return l2.value, li.value

    def computeLinearRHS(self):
        l2, li = C.c_double(), C.c_double()
        _chk(lib().hamstr_compute_linear_rhs(self.h, C.byref(l2), C.byref(li)))
        return l2.value, li.value

    def updateSolution(self, qold: str, qnew: str, coef: float):
        """SOLVER::updateSolution: qnew = qold + (coef*CFL/sigma) r, with qold
        and qnew each "q" or "qt"."""
        _chk(lib().hamstr_update_solution(self.h, FIELDS[qold], FIELDS[qnew], C.c_double(coef)))

    def set_scheme(self, scheme: str):
        """stepType: "ADI" (default), "euler" or "rk3"."""
        if scheme not in SCHEMES:
            raise ValueError(f"scheme must be one of {sorted(SCHEMES)}")
        _chk(lib().hamstr_set_scheme(self.h, SCHEMES[scheme]))

    @property
    def scheme(self) -> str:
        v = lib().hamstr_get_scheme(self.h)
        return {i: k for k, i in SCHEMES.items()}[v]

    def step(self, nsteps: int, tol: float = 0.0, check_every: int = 1):
        """nsteps x stepSolution with the current scheme; returns (steps run,
        l2rhs[] of each step's first computeRHS, the printed l2[])."""
        l2r = np.zeros(max(nsteps, 1))
        l2l = np.zeros(max(nsteps, 1))
        done = C.c_int(0)
        rc = lib().hamstr_step(self.h, int(nsteps), C.c_double(tol), int(check_every), _d(l2r), _d(l2l),
                               C.byref(done))
        _chk(rc)
        n = done.value
        return n, l2r[:n], l2l[:n]

    def stepSolution(self):
        return self.step(1)

    def marchInTime(self, nsteps: int):
        return self.step(nsteps)

    def last_norms(self) -> dict:
        out = np.zeros(4)
        _chk(lib().hamstr_last_norms(self.h, _d(out)))
        return dict(zip(("l2_rhs", "linf_rhs", "l2_lin", "linf_lin"), out))

    def sync(self):
        _chk(lib().hamstr_sync(self.h))

    def set_graph(self, on: bool):
        """Replay captured HIP graphs between norm read-backs (default on)."""
        _chk(lib().hamstr_set_graph(self.h, int(on)))

    def prepare_graph(self, check_every: int):
        """Capture the graph step(…, check_every) replays, outside any timed region."""
        _chk(lib().hamstr_prepare_graph(self.h, int(check_every)))

    def set_retire(self, on: bool = True, stall_window: int = 0):
        """Retire converged (and, with stall_window > 0, stalled) cases."""
        _chk(lib().hamstr_set_retire(self.h, int(on), int(stall_window)))

    def set_case_live(self, k: int, live: bool):
        """Hold case k out of the launches, or make it live again (retirement on)."""
        _chk(lib().hamstr_set_case_live(self.h, int(k), int(live)))

    def iterations(self) -> int:
        """Iterations run since set_retire."""
        return int(lib().hamstr_iterations(self.h))

    def case_status(self):
        """(status[k]: 0 live, 1 converged, 2 stalled; iteration of that event or -1)."""
        n = getattr(self, "n_case", 1)
        st = np.zeros(n, np.int32)
        it = np.zeros(n, np.int64)
        _chk(lib().hamstr_case_status(self.h, _i(st), it.ctypes.data_as(C.POINTER(C.c_long))))
        return st, it

    def set_timing(self, on: bool):
        _chk(lib().hamstr_set_timing(self.h, int(on)))

    def last_timings(self) -> dict:
        """Average per-step ms of each stage of the last step() with timing on."""
        buf = np.zeros(64)
        n = C.c_int()
        _chk(lib().hamstr_last_timings(self.h, _d(buf), 64, C.byref(n)))
        names = C.create_string_buffer(1024)
        _chk(lib().hamstr_last_timing_names(self.h, names, 1024))
        keys = names.value.decode().split(",") if names.value else []
        out = {}
        for k, v in zip(keys, buf[: n.value]):
            out[k] = out.get(k, 0.0) + float(v)
        return out


# ---- per-function entry points ---------------------------------------------

def flux_roe2d(nvec, ql, qr, gamma=1.4):
    nvec = np.ascontiguousarray(nvec, np.float64).reshape(-1, 2)
    ql = np.ascontiguousarray(ql, np.float64).reshape(-1, 4)
    qr = np.ascontiguousarray(qr, np.float64).reshape(-1, 4)
    n = nvec.shape[0]
    flux = np.zeros((n, 4))
    sr = np.zeros(n)
    _chk(lib().hamstr_flux_roe2d(n, _d(nvec), _d(ql), _d(qr), C.c_double(gamma), _d(flux), _d(sr)))
    return flux, sr


def jacobian_roe2d(nvec, ql, qr, gamma=1.4, lhs=LHS_FP64):
    nvec = np.ascontiguousarray(nvec, np.float64).reshape(-1, 2)
    ql = np.ascontiguousarray(ql, np.float64).reshape(-1, 4)
    qr = np.ascontiguousarray(qr, np.float64).reshape(-1, 4)
    n = nvec.shape[0]
    lm = np.zeros((n, 4, 4))
    rm = np.zeros((n, 4, 4))
    _chk(lib().hamstr_jacobian_roe2d(n, _d(nvec), _d(ql), _d(qr), C.c_double(gamma), int(lhs), _d(lm), _d(rm)))
    return lm, rm


def _recon(name, f, dq, is_, ie, th=None, qt=None, eps=1e-10):
    f = np.ascontiguousarray(f, np.float64)
    if f.ndim == 2:
        f = f[None]
    nch, imax, _ = f.shape
    ql = np.zeros_like(f)
    qr = np.zeros_like(f)
    L = lib()
    if name == "muscl":
        _chk(L.hamstr_muscl(nch, imax, is_, ie, C.c_double(th), C.c_double(qt), C.c_double(eps), _d(f), _d(ql), _d(qr)))
    elif name == "weno":
        _chk(L.hamstr_weno(nch, imax, is_, ie, C.c_double(eps), _d(f), _d(ql), _d(qr)))
    else:
        dq = np.ascontiguousarray(dq, np.float64).reshape(f.shape)
        if name == "muscl_deriv":
            _chk(L.hamstr_muscl_deriv(nch, imax, is_, ie, C.c_double(th), C.c_double(qt), C.c_double(eps), _d(f),
                                      _d(dq), _d(ql), _d(qr)))
        else:
            _chk(L.hamstr_weno_deriv(nch, imax, is_, ie, C.c_double(eps), _d(f), _d(dq), _d(ql), _d(qr)))
    return ql, qr


def muscl(f, is_, ie, th, qt, eps=1e-10):
    return _recon("muscl", f, None, is_, ie, th, qt, eps)


def weno(f, is_, ie, eps=1e-10):
    return _recon("weno", f, None, is_, ie, eps=eps)


def muscl_deriv(f, dq, is_, ie, th, qt, eps=1e-10):
    return _recon("muscl_deriv", f, dq, is_, ie, th, qt, eps)


def weno_deriv(f, dq, is_, ie, eps=1e-10):
    return _recon("weno_deriv", f, dq, is_, ie, eps=eps)


def block_tridag4(a, b, c, f, periodic=False):
    """a, b, c: (nsys, N, 4, 4); f: (nsys, N, 4).  Returns the solution."""
    a = np.ascontiguousarray(a, np.float64)
    if a.ndim == 3:
        a, b, c, f = a[None], np.asarray(b)[None], np.asarray(c)[None], np.asarray(f)[None]
    b = np.ascontiguousarray(b, np.float64)
    c = np.ascontiguousarray(c, np.float64)
    x = np.array(f, np.float64, copy=True, order="C")
    nsys, N = a.shape[0], a.shape[1]
    fn = lib().hamstr_block_tridag_periodic4 if periodic else lib().hamstr_block_tridag4
    _chk(fn(nsys, N, _d(a), _d(b), _d(c), _d(x)))
    return x


# ---- the 3D / strand building blocks (5 conserved variables) ----------------

def flux_roe3d(nxyz, ql, qr, gamma=1.4):
    """fluxRoe3D on n faces: (flux (n, 5), spectral radius (n,)), unit normal."""
    nxyz = np.ascontiguousarray(nxyz, np.float64).reshape(-1, 3)
    ql = np.ascontiguousarray(ql, np.float64).reshape(-1, 5)
    qr = np.ascontiguousarray(qr, np.float64).reshape(-1, 5)
    n = nxyz.shape[0]
    flux = np.zeros((n, 5))
    sr = np.zeros(n)
    _chk(lib().hamstr_flux_roe3d(n, _d(nxyz), _d(ql), _d(qr), C.c_double(gamma), _d(flux), _d(sr)))
    return flux, sr


def jacobian_roe3d(nxyz, ql, qr, gamma=1.4):
    """jacobianRoe3D on n faces: (lmat (n, 5, 5), rmat (n, 5, 5))."""
    nxyz = np.ascontiguousarray(nxyz, np.float64).reshape(-1, 3)
    ql = np.ascontiguousarray(ql, np.float64).reshape(-1, 5)
    qr = np.ascontiguousarray(qr, np.float64).reshape(-1, 5)
    n = nxyz.shape[0]
    lm = np.zeros((n, 5, 5))
    rm = np.zeros((n, 5, 5))
    _chk(lib().hamstr_jacobian_roe3d(n, _d(nxyz), _d(ql), _d(qr), C.c_double(gamma), _d(lm), _d(rm)))
    return lm, rm


def block_tridag5(a, b, c, f, periodic=False):
    """a, b, c: (nsys, N, 5, 5); f: (nsys, N, 5).  Returns the solution."""
    a = np.ascontiguousarray(a, np.float64)
    if a.ndim == 3:
        a, b, c, f = a[None], np.asarray(b)[None], np.asarray(c)[None], np.asarray(f)[None]
    b = np.ascontiguousarray(b, np.float64)
    c = np.ascontiguousarray(c, np.float64)
    x = np.array(f, np.float64, copy=True, order="C")
    nsys, N = a.shape[0], a.shape[1]
    fn = lib().hamstr_block_tridag_periodic5 if periodic else lib().hamstr_block_tridag5
    _chk(fn(nsys, N, _d(a), _d(b), _d(c), _d(x)))
    return x


class Batch(Solver):
    """Many flow cases in one context (hamstr_batch_create); every launch
    covers all cases.  cases: list of (mesh, mach, alpha)."""

    def __init__(self, cases, order: int, cfl: float, msweep: int = 1, lhs: int = LHS_FP64,
                 adi_mode: int = ADI_FAST, sol_visc: int = 0):
        n = len(cases)
        descs = (MeshDesc * n)()
        prms = (Params * n)()
        self._keep = []
        self.meshes = []
        for k, (mesh, mach, alpha) in enumerate(cases):
            faces, nv, rm = preprocess(mesh.x, mesh.faces, sol_visc)
            st = np.ascontiguousarray(mesh.chain_start, np.int32)
            cc = np.ascontiguousarray(mesh.chain_conn, np.int32)
            orph = _orphans(mesh)
            self._keep += [faces, nv, rm, st, cc, orph]
            descs[k] = MeshDesc(mesh.n_cell, mesh.n_face, mesh.n_chain, _i(faces), _i(st), _i(cc), _d(nv), _d(rm),
                                int(orph.size), _i(orph) if orph.size else None)
            prms[k] = Params(float(mach), float(alpha), float(cfl), int(order), int(msweep), int(lhs), int(adi_mode))
            self.meshes.append(mesh)
        h = C.c_void_p()
        _chk(lib().hamstr_batch_create(n, descs, prms, C.byref(h)))
        self.h = h
        self.n_case = n
        off = np.zeros(n + 1, np.int32)
        nc = C.c_int()
        _chk(lib().hamstr_batch_info(self.h, C.byref(nc), _i(off)))
        self.cell_off = off
        self.n_cell = int(off[-1])
        self.n_face = sum(m.n_face for m in self.meshes)
        self.order, self.cfl, self.msweep, self.lhs = order, cfl, msweep, lhs

    def get(self, name: str) -> np.ndarray:
        if name == "ff":
            raise ValueError("per-face Jacobians of a batch: use a single-case Solver")
        return super().get(name)

    def case_q(self, k: int) -> np.ndarray:
        c0, c1 = int(self.cell_off[k]), int(self.cell_off[k + 1])
        out = np.zeros((c1 - c0, 4))
        _chk(lib().hamstr_get_cells(self.h, FIELDS["q"], c0, c1 - c0, _d(out)))
        return out

    def set_case_q(self, k: int, q) -> None:
        c0, c1 = int(self.cell_off[k]), int(self.cell_off[k + 1])
        q = np.ascontiguousarray(q, np.float64).reshape(c1 - c0, 4)
        _chk(lib().hamstr_set_cells(self.h, FIELDS["q"], c0, c1 - c0, _d(q)))

    def batch_step(self, nsteps: int, tol: float = 0.0, check_every: int = 50):
        """Returns (steps run, iterations-to-tol per case (-1: not reached), norms (ncase, 4))."""
        iters = np.full(self.n_case, -1, np.int32)
        norms = np.zeros((self.n_case, 4))
        done = C.c_int()
        _chk(lib().hamstr_batch_step(self.h, int(nsteps), C.c_double(tol), int(check_every), _i(iters), _d(norms),
                                     C.byref(done)))
        return done.value, iters, norms


def autotune(solver: Solver, reps: int = 3):
    """Runs the launch-shape autotuner; returns the chosen launch groups."""
    best = np.zeros(64)
    n = C.c_int()
    _chk(lib().hamstr_autotune(solver.h, int(reps), _d(best), 64, C.byref(n)))
    return launch_groups(solver), best[: n.value]


def set_launch(solver: Solver, group: int, grid: int, cr_levels: int):
    """Persistent grid and cyclic-reduction levels of a FAST launch group."""
    _chk(lib().hamstr_set_launch(solver.h, int(group), int(grid), int(cr_levels)))


def set_thomas(solver: Solver, group: int, pcr_levels: int):
    """PCR levels (0..4) before the block-Thomas finish of a FAST launch group
    of open chains; -1: parallel cyclic reduction to the end."""
    _chk(lib().hamstr_set_thomas(solver.h, int(group), int(pcr_levels)))


def thomas_levels(solver: Solver, group: int) -> int:
    """The group's PCR levels before the Thomas finish (-1: none)."""
    v = C.c_int()
    _chk(lib().hamstr_thomas(solver.h, int(group), C.byref(v)))
    return v.value


def set_adi_kernel(solver: Solver, group: int, kernel: int):
    """ADI kernel of a FAST launch group: KERNEL_ROW, or KERNEL_HALF /
    KERNEL_HALF2 (open chains) where the group's chains fit them."""
    _chk(lib().hamstr_set_adi_kernel(solver.h, int(group), int(kernel)))


JAC_STORED, JAC_FROM_Q, JAC_FROM_STATES, JAC_STATES_LINES = 0, 1, 2, 3


def set_jacobian_source(solver: Solver, source: int):
    """Fused FAST iteration: the ADI kernels read the chain faces' stored
    Jacobians (JAC_STORED, the default) or compute them again from stored face
    states (JAC_FROM_STATES) or from q (JAC_FROM_Q); JAC_STATES_LINES: phase 0
    reads stored Jacobians and the other phases use stored states (where phase
    0's chains come first)."""
    _chk(lib().hamstr_set_jacobian_source(solver.h, int(source)))


def jacobian_source(solver: Solver) -> int:
    src = C.c_int()
    _chk(lib().hamstr_jacobian_source(solver.h, C.byref(src)))
    return src.value


def set_face_staging(solver: Solver, on: bool):
    """Fused face kernel: stage each workgroup's stencil rows in LDS (the
    default where they fit; HamstrError where they do not) or gather them
    directly.  The same results either way."""
    _chk(lib().hamstr_set_face_staging(solver.h, int(bool(on))))


def face_staging(solver: Solver) -> bool:
    on = C.c_int()
    _chk(lib().hamstr_face_staging(solver.h, C.byref(on)))
    return bool(on.value)


def launch_groups(solver: Solver):
    """Launch groups as dicts: phase, kernel (KERNEL_*; "fast" is true for the
    FAST kernels), final solve, chains, max_rows, grid, block, cr_levels, and
    thomas (PCR levels before the Thomas finish of open chains, -1: none)."""
    out = np.zeros(8 * 64, np.int32)
    n = C.c_int()
    _chk(lib().hamstr_launch_groups(solver.h, _i(out), 64, C.byref(n)))
    keys = ("phase", "kernel", "final", "chains", "max_rows", "grid", "block", "cr_levels")
    groups = [dict(zip(keys, map(int, out[8 * k:8 * k + 8]))) for k in range(n.value)]
    for k, g in enumerate(groups):
        g["fast"] = int(g["kernel"] != KERNEL_EXACT)
        v = C.c_int()
        _chk(lib().hamstr_thomas(solver.h, k, C.byref(v)))
        g["thomas"] = v.value
    return groups

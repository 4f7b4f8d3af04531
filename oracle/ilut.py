"""ctypes front end of oracle/ilut_bicgstab.c: BiCGSTAB with a dual-threshold ILUT
preconditioner, the reference's default linear solve (src/solver.cpp:53-80: Eigen's
BiCGSTAB + IncompleteLUT, fill factor 10, drop tolerance 1e-4, tolerance 1e-10, at most
10000 iterations, x0 = 0), restated in C from the published algorithms.  Test
infrastructure: used by tests/ and by bench.py's CPU baseline only.

One deliberate difference: Eigen orders the ILUT with AMD on the pattern of A + A^T; this
restatement takes reverse Cuthill-McKee (scipy.sparse.csgraph), because AMD is not
available here.  The ordering changes the incomplete factors and the iteration count, not
the solution the solve converges to.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
import time

import numpy as np
import scipy.sparse as sp
from scipy.sparse.csgraph import reverse_cuthill_mckee

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "liboracle_ilut.so")
_lib = None


def build() -> str:
    """Compile the C restatement (oracle/Makefile) into oracle/_lib."""
    r = subprocess.run(["make", "-C", HERE, "-s"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building {LIB_PATH} failed:\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp, i64, i32, dbl = C.c_void_p, C.c_int64, C.c_int, C.c_double
    lib.efo_ilut_factor.argtypes = [i64, vp, vp, vp, vp, dbl, i32, dbl, C.POINTER(vp), C.POINTER(i64),
                                    C.POINTER(dbl)]
    lib.efo_ilut_factor.restype = i32
    lib.efo_ilut_nnz.argtypes = [vp]
    lib.efo_ilut_nnz.restype = i64
    lib.efo_ilut_free.argtypes = [vp]
    lib.efo_ilut_free.restype = None
    lib.efo_ilut_apply.argtypes = [vp, vp, vp]
    lib.efo_ilut_apply.restype = i32
    lib.efo_bicgstab.argtypes = [i64, vp, vp, vp, vp, vp, vp, dbl, i32, dbl, C.POINTER(i32), C.POINTER(dbl),
                                 C.POINTER(dbl)]
    lib.efo_bicgstab.restype = i32
    _lib = lib
    return lib


def _csr_arrays(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    rp = np.ascontiguousarray(A.indptr.astype(np.int64))
    ci = np.ascontiguousarray(A.indices.astype(np.int32))
    v = np.ascontiguousarray(A.data.astype(np.complex128))
    return A.shape[0], rp, ci, v


def rcm_order(A) -> np.ndarray:
    """Reverse Cuthill-McKee of the pattern of A + A^T (new -> old)."""
    P = sp.csr_matrix(A, copy=True)
    P.data = np.ones_like(P.data, dtype=np.int8)
    S = (P + P.T).tocsr()
    return np.ascontiguousarray(reverse_cuthill_mckee(S, symmetric_mode=True).astype(np.int32))


class ILUTTimeLimit(RuntimeError):
    """The factorization passed its time limit; rows_done / seconds give its rate."""

    def __init__(self, rows_done, n, seconds):
        super().__init__(f"ILUT stopped at its time limit after {rows_done} of {n} rows ({seconds:.1f} s)")
        self.rows_done, self.n, self.seconds = rows_done, n, seconds


class ILUT:
    """Incomplete LU of P A P^T with dual thresholds (drop tolerance, fill factor).
    time_limit > 0 stops the row loop after that many seconds (ILUTTimeLimit)."""

    def __init__(self, A, droptol: float = 1e-4, fillfactor: int = 10, perm="rcm", time_limit: float = 0.0):
        lib = load()
        n, rp, ci, v = _csr_arrays(A)
        if isinstance(perm, str) and perm == "rcm":
            perm = rcm_order(A)
        pp = None if perm is None else np.ascontiguousarray(np.asarray(perm, np.int32))
        h = C.c_void_p()
        secs = C.c_double()
        rows = C.c_int64()
        rc = lib.efo_ilut_factor(n, rp.ctypes.data, ci.ctypes.data, v.ctypes.data,
                                 None if pp is None else pp.ctypes.data, float(droptol), int(fillfactor),
                                 float(time_limit), C.byref(h), C.byref(rows), C.byref(secs))
        if rc == 3:
            raise ILUTTimeLimit(rows.value, n, secs.value)
        if rc != 0:
            raise RuntimeError({1: "ILUT: matrix has an empty row", 2: "ILUT: out of memory"}.get(rc, f"ILUT rc {rc}"))
        self._h = h
        self.n = n
        self.seconds = secs.value
        self.nnz = int(lib.efo_ilut_nnz(h))

    def close(self):
        if getattr(self, "_h", None):
            load().efo_ilut_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def apply(M: ILUT, v):
    """M^-1 v = P^T U^-1 L^-1 P v."""
    vv = np.ascontiguousarray(np.asarray(v, np.complex128))
    y = np.zeros_like(vv)
    if load().efo_ilut_apply(M._h, vv.ctypes.data, y.ctypes.data) != 0:
        raise RuntimeError("ILUT apply: out of memory")
    return y


def bicgstab(A, b, M: ILUT = None, tol: float = 1e-10, max_iterations: int = 10000, time_limit: float = 0.0):
    """x0 = 0 BiCGSTAB on A x = b, right-preconditioned by M (None: unpreconditioned).
    Returns (x, info) with info = {status: 0 converged / 1 not converged / 2 time limit,
    iters, resid (recursive relative residual), seconds}."""
    lib = load()
    n, rp, ci, v = _csr_arrays(A)
    bb = np.ascontiguousarray(np.asarray(b, np.complex128))
    x = np.zeros(n, np.complex128)
    it = C.c_int()
    res = C.c_double()
    secs = C.c_double()
    st = lib.efo_bicgstab(n, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, None if M is None else M._h,
                          bb.ctypes.data, x.ctypes.data, float(tol), int(max_iterations), float(time_limit),
                          C.byref(it), C.byref(res), C.byref(secs))
    if st == 3:
        raise RuntimeError("BiCGSTAB: out of memory")
    return x, dict(status=int(st), iters=it.value, resid=res.value, seconds=secs.value)


def solve_linear(A, b, tol: float = 1e-10, max_iterations: int = 10000, droptol: float = 1e-4,
                 fillfactor: int = 10, time_limit: float = 0.0):
    """The reference's default solve_linear: ILUT set-up, then preconditioned BiCGSTAB
    (src/solver.cpp:53-80).  info adds ilut_seconds and ilut_nnz."""
    t0 = time.perf_counter()
    M = ILUT(A, droptol, fillfactor)
    t_f = time.perf_counter() - t0
    x, info = bicgstab(A, b, M, tol, max_iterations, time_limit)
    info.update(ilut_seconds=t_f, ilut_nnz=M.nnz)
    M.close()
    return x, info

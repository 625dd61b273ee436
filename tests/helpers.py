"""Test-side helpers: tiled split-plane SoA layout converters, ctypes bindings
for the CPU oracle (oracle/liboracle.so) and, where it was built, the
reference's own code (oracle/_ref/libref*.so), plus the comparison metrics
used by the parity tests.

Layout (utils.cpp:96, 153 of the reference): matrix m = b*T + t,
element (r, c) at b*T*n*n + (r + c*n)*T + t, vector entry r at b*T*n + r*T + t.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent
ORACLE_DIR = REPO / "oracle"
ORACLE_SO = ORACLE_DIR / "liboracle.so"
REF_SO = ORACLE_DIR / "_ref" / "libref.so"
REF_EXACT_SO = ORACLE_DIR / "_ref" / "libref_exact.so"
REF_V4_SO = ORACLE_DIR / "_ref" / "libref_v4.so"

# oracle behaviour switches (oracle/eig_oracle.h)
KEEP_TRASH = 1
LAPACK_SIGN = 2
LAPACK_TST = 4
NO_T_RESCALE = 8
ZGEMV_FIX = 16
EIG_ONLY = 32

_P = ctypes.POINTER(ctypes.c_double)
_I32P = ctypes.POINTER(ctypes.c_int32)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


# --------------------------------------------------------------------------
# layout
# --------------------------------------------------------------------------
def to_soa(A: np.ndarray, thread_cnt: int):
    """(N, n, n) complex -> (re, im) planes in the tiled SoA layout."""
    N, n, _ = A.shape
    assert N % thread_cnt == 0
    bc = N // thread_cnt
    X = A.reshape(bc, thread_cnt, n, n)          # [b, t, r, c]
    X = np.transpose(X, (0, 3, 2, 1))            # [b, c, r, t]
    X = np.ascontiguousarray(X).reshape(-1)
    return np.ascontiguousarray(X.real), np.ascontiguousarray(X.imag)


def from_soa(re: np.ndarray, im: np.ndarray, n: int, thread_cnt: int) -> np.ndarray:
    bc = re.size // (n * n * thread_cnt)
    X = (re + 1j * im).reshape(bc, n, n, thread_cnt)   # [b, c, r, t]
    return np.transpose(X, (0, 3, 2, 1)).reshape(bc * thread_cnt, n, n)


def vec_to_soa(E: np.ndarray, thread_cnt: int):
    N, n = E.shape
    bc = N // thread_cnt
    X = np.transpose(E.reshape(bc, thread_cnt, n), (0, 2, 1)).reshape(-1)
    return np.ascontiguousarray(X.real), np.ascontiguousarray(X.imag)


def vec_from_soa(re: np.ndarray, im: np.ndarray, n: int, thread_cnt: int, mult: int = 1):
    bc = re.size // (n * thread_cnt * mult)
    X = (re + 1j * im).reshape(bc, mult * n, thread_cnt)
    return np.transpose(X, (0, 2, 1)).reshape(bc * thread_cnt, mult * n)


def gaussian(N: int, n: int, seed: int) -> np.ndarray:
    """Complex matrices with real and imaginary parts iid N(0, 1)."""
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, n, n)) + 1j * rng.standard_normal((N, n, n))


# --------------------------------------------------------------------------
# ctypes bindings (oracle and reference share argument conventions)
# --------------------------------------------------------------------------
class _Stats(ctypes.Structure):
    _fields_ = [("info", _I32P), ("sweeps", _I32P), ("steps", _I32P)]


def load_oracle():
    lib = ctypes.CDLL(str(ORACLE_SO))
    i, d = ctypes.c_int, ctypes.c_double
    lib.oracle_eigensolver.argtypes = [i, i, i] + [_P] * 6 + [i, ctypes.POINTER(_Stats)]
    lib.oracle_zgehrd.argtypes = [i, i, i] + [_P] * 4 + [i]
    lib.oracle_zunghr.argtypes = [i, i, i] + [_P] * 4 + [i]
    lib.oracle_hessenberg.argtypes = [i, i, i] + [_P] * 4 + [i]
    lib.oracle_zlahqr.argtypes = [i, i, i] + [_P] * 6 + [i, ctypes.POINTER(_Stats)]
    lib.oracle_call_zlahqr_new.argtypes = [i, i, i] + [_P] * 6 + [i]
    lib.oracle_ztrevc.argtypes = [i, i, i] + [_P] * 4 + [i]
    lib.oracle_zgemv.argtypes = [i, i, i, i] + [_P] * 6 + [d, d, i]
    lib.oracle_backsubstitution.argtypes = [i, i, i] + [_P] * 4
    lib.oracle_set_threads.argtypes = [i]
    if hasattr(lib, "oracle_get_threads"):
        lib.oracle_get_threads.restype = i
    lib.oracle_zlahqr_eig.argtypes = [i, i, i] + [_P] * 4 + [i, ctypes.POINTER(_Stats)]
    lib.oracle_companion.argtypes = [i, i, i] + [_P] * 4
    lib.oracle_zlahqr_roots.argtypes = [i, i, i] + [_P] * 8 + [i, ctypes.POINTER(_Stats)]
    return lib


def load_ref(exact: bool = True, isa: str = "v3"):
    """The reference checker build: exact (FMA contraction off, x86-64-v3)
    or as the reference builds itself (contraction on), for x86-64-v3 or,
    with isa="v4", x86-64-v4 (AVX-512)."""
    path = REF_EXACT_SO if exact else (REF_V4_SO if isa == "v4" and REF_V4_SO.exists() else REF_SO)
    if not path.exists():
        return None
    lib = ctypes.CDLL(str(path))
    i, d = ctypes.c_int, ctypes.c_double
    for name in ("ref_eigensolver", "ref_eigensolver_clean", "ref_eigensolver_shipped"):
        if hasattr(lib, name):
            getattr(lib, name).argtypes = [i, i, i] + [_P] * 6
    for name in ("ref_zgehrd", "ref_zunghr", "ref_hessenberg", "ref_ztrevc", "ref_backsubstitution"):
        getattr(lib, name).argtypes = [i, i, i] + [_P] * 4
    lib.ref_zlahqr_new.argtypes = [i, i, i] + [_P] * 6
    lib.ref_call_zlahqr_new.argtypes = [i, i, i] + [_P] * 6
    lib.ref_zgemv.argtypes = [i, i, i, i] + [_P] * 6 + [d, d]
    if hasattr(lib, "ref_set_threads"):
        lib.ref_set_threads.argtypes = [i]
        lib.ref_set_threads.restype = i
    if hasattr(lib, "ref_get_threads"):
        lib.ref_get_threads.restype = i
    if hasattr(lib, "ref_zlahqr_eig"):
        lib.ref_zlahqr_eig.argtypes = [i, i, i] + [_P] * 4
        lib.ref_zlahqr_roots.argtypes = [i, i, i] + [_P] * 8
    return lib


def oracle_eig(A: np.ndarray, thread_cnt: int = 1, flags: int = 0, lib=None, stats: bool = False):
    """Run the oracle's eigensolver on (N,n,n) complex input.

    Returns (E, V, T_out) in matrix form, plus (info, sweeps, steps) if stats."""
    lib = lib or load_oracle()
    N, n, _ = A.shape
    rA, iA = to_soa(A, thread_cnt)
    rE = np.zeros(N * n); iE = np.zeros(N * n)
    rV = np.zeros(N * n * n); iV = np.zeros(N * n * n)
    info = np.zeros(N, np.int32); sw = np.zeros(N, np.int32); st = np.zeros(N, np.int32)
    S = _Stats(info.ctypes.data_as(_I32P), sw.ctypes.data_as(_I32P), st.ctypes.data_as(_I32P))
    lib.oracle_eigensolver(n, N // thread_cnt, thread_cnt, dptr(rA), dptr(iA), dptr(rE), dptr(iE),
                           dptr(rV), dptr(iV), flags, ctypes.byref(S))
    out = (vec_from_soa(rE, iE, n, thread_cnt), from_soa(rV, iV, n, thread_cnt),
           from_soa(rA, iA, n, thread_cnt))
    if stats:
        return out + ((info, sw, st),)
    return out


def ref_eig(A: np.ndarray, thread_cnt: int = 1, clean: bool = True, lib=None):
    lib = lib or load_ref()
    N, n, _ = A.shape
    rA, iA = to_soa(A, thread_cnt)
    rE = np.zeros(N * n); iE = np.zeros(N * n)
    rV = np.zeros(N * n * n); iV = np.zeros(N * n * n)
    # the as-shipped sequence with zeroed tau/work: the shipped call itself
    # reads its uninitialised workspace (0 * y in zgemv), see ref_harness.cpp
    shipped = lib.ref_eigensolver_shipped if hasattr(lib, "ref_eigensolver_shipped") else lib.ref_eigensolver
    fn = lib.ref_eigensolver_clean if clean else shipped
    fn(n, N // thread_cnt, thread_cnt, dptr(rA), dptr(iA), dptr(rE), dptr(iE), dptr(rV), dptr(iV))
    return (vec_from_soa(rE, iE, n, thread_cnt), from_soa(rV, iV, n, thread_cnt),
            from_soa(rA, iA, n, thread_cnt))


def hessenberg_input(N: int, n: int, seed: int, thread_cnt: int = 1, lib=None) -> np.ndarray:
    """Upper Hessenberg matrices with a real subdiagonal: the oracle's
    reduction of Gaussian matrices, entries below the subdiagonal cleared."""
    lib = lib or load_oracle()
    A = gaussian(N, n, seed)
    rA, iA = to_soa(A, thread_cnt)
    rQ = np.zeros_like(rA); iQ = np.zeros_like(iA)
    lib.oracle_hessenberg(n, N // thread_cnt, thread_cnt, dptr(rA), dptr(iA), dptr(rQ), dptr(iQ), 0)
    return np.triu(from_soa(rA, iA, n, thread_cnt), -1)


def eig_only(H: np.ndarray, thread_cnt: int = 1, lib=None, ref: bool = False):
    """Eigenvalues of Hessenberg H by the oracle's (or the reference's)
    eigenvalues-only QR.  Returns (E, H after the iteration, info)."""
    N, n, _ = H.shape
    rH, iH = to_soa(H, thread_cnt)
    rE = np.zeros(N * n); iE = np.zeros(N * n)
    info = np.zeros(N, np.int32)
    if ref:
        lib.ref_zlahqr_eig(n, N // thread_cnt, thread_cnt, dptr(rE), dptr(iE), dptr(rH), dptr(iH))
    else:
        lib = lib or load_oracle()
        S = _Stats(info.ctypes.data_as(_I32P), None, None)
        lib.oracle_zlahqr_eig(n, N // thread_cnt, thread_cnt, dptr(rE), dptr(iE), dptr(rH), dptr(iH), 0,
                              ctypes.byref(S))
    return vec_from_soa(rE, iE, n, thread_cnt), from_soa(rH, iH, n, thread_cnt), info


def roots_run(Cf: np.ndarray, thread_cnt: int = 1, lib=None, ref: bool = False):
    """Polynomial roots by the oracle's (or the reference's) roots kernel.
    Cf: (N, n+1) complex coefficients c_0..c_n.  Returns (E, T, Z)."""
    N, n1 = Cf.shape
    n = n1 - 1
    rC, iC = vec_to_soa(Cf, thread_cnt)
    rE = np.zeros(N * n); iE = np.zeros(N * n)
    rV = np.zeros(N * n * n); iV = np.zeros(N * n * n)
    rZ = np.zeros(N * n * n); iZ = np.zeros(N * n * n)
    args = [dptr(rC), dptr(iC), dptr(rE), dptr(iE), dptr(rV), dptr(iV), dptr(rZ), dptr(iZ)]
    if ref:
        lib.ref_zlahqr_roots(n, N // thread_cnt, thread_cnt, *args)
    else:
        lib = lib or load_oracle()
        lib.oracle_zlahqr_roots(n, N // thread_cnt, thread_cnt, *args, 0, None)
    return (vec_from_soa(rE, iE, n, thread_cnt), from_soa(rV, iV, n, thread_cnt),
            from_soa(rZ, iZ, n, thread_cnt))


# --------------------------------------------------------------------------
# comparison metrics
# --------------------------------------------------------------------------
def backward_error(A: np.ndarray, E: np.ndarray, V: np.ndarray) -> np.ndarray:
    """||A V - V diag(E)||_F / (||A||_F ||V||_F) per matrix."""
    R = A @ V - V * E[:, None, :]
    return (np.linalg.norm(R, axis=(1, 2))
            / (np.linalg.norm(A, axis=(1, 2)) * np.linalg.norm(V, axis=(1, 2))))


def match_eigenvalues(E1: np.ndarray, E2: np.ndarray):
    """Per matrix, a permutation p with E2[p] closest to E1 (Hungarian
    assignment on |E1_i - E2_j|).  Returns (perm, max abs difference)."""
    from scipy.optimize import linear_sum_assignment
    N, n = E1.shape
    perm = np.empty((N, n), np.int64)
    err = np.empty(N)
    for k in range(N):
        C = np.abs(E1[k][:, None] - E2[k][None, :])
        r, c = linear_sum_assignment(C)
        perm[k, r] = c
        err[k] = C[r, c].max() if n else 0.0
    return perm, err


def phase_normalise(V: np.ndarray) -> np.ndarray:
    """Scale each column to unit 2-norm and rotate it so that its
    largest-magnitude entry is real and positive."""
    nrm = np.linalg.norm(V, axis=-2, keepdims=True)
    W = V / np.where(nrm == 0, 1, nrm)
    idx = np.argmax(np.abs(W), axis=-2)
    piv = np.take_along_axis(W, idx[..., None, :], axis=-2)
    ph = piv / np.abs(np.where(piv == 0, 1, piv))
    return W / np.where(ph == 0, 1, ph)


def eigvec_distance(V1: np.ndarray, V2: np.ndarray, perm: np.ndarray) -> np.ndarray:
    """max over columns of || pn(V1[:,i]) - pn(V2[:,perm[i]]) ||_2, per matrix,
    where pn is phase_normalise.  The distance is phase invariant."""
    W1 = phase_normalise(V1)
    W2 = phase_normalise(np.take_along_axis(V2, perm[:, None, :], axis=2))
    # guard against the pivot picking different (equal-magnitude) entries:
    # align W2 to W1 by the optimal phase instead of the pivot phase
    ip = np.sum(np.conj(W2) * W1, axis=1, keepdims=True)
    ph = ip / np.where(np.abs(ip) == 0, 1, np.abs(ip))
    return np.max(np.linalg.norm(W1 - W2 * ph, axis=1), axis=1)


class single_threaded:
    """Context: the oracle's and the reference build's OpenMP loops on one
    thread, so that a bit-for-bit comparison of the two compares their
    arithmetic and not how the batch was scheduled over threads."""

    def __init__(self, orc, ref):
        self.orc, self.ref, self.prev = orc, ref, None

    def __enter__(self):
        self.orc.oracle_set_threads(1)
        if self.ref is not None and hasattr(self.ref, "ref_set_threads"):
            self.prev = self.ref.ref_set_threads(1)
        return self

    def __exit__(self, *exc):
        self.orc.oracle_set_threads(0)
        if self.prev is not None:
            self.ref.ref_set_threads(self.prev)
        return False

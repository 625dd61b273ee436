"""ctypes binding of libveig.so (include/veig.h) and device-tensor helpers.

The library is the product: every compute call here goes to the gfx950
kernels.  There is no CPU fallback; if libveig.so is missing or the GPU is
not usable the calls raise.

Device entry points take torch tensors on the GPU (float64, contiguous, in
the tiled split-plane layout of the reference, utils.cpp:85-157) and enqueue
on torch's current stream.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
LIB_PATH = PKG / "libveig.so"
HEADER = REPO / "include" / "veig.h"

VEIG_OK = 0
VEIG_WRITE_SCHUR = 1
MAX_DIM = 64          # fused entries (veig_zgeev, veig_zgeev_fp32, veig_zgeev_mixed)
MAX_DIM_STAGES = 32   # stage entries and varsize batches


class VeigError(RuntimeError):
    pass


_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.c_int64
_i = ctypes.c_int
_vp = ctypes.c_void_p

# argument types of every function declared in include/veig.h
_SIGS = {
    "veig_status_string": ([_i], ctypes.c_char_p),
    "veig_last_error": ([], ctypes.c_char_p),
    "veig_version": ([], ctypes.c_char_p),
    "veig_zgeev": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32,
                    ctypes.c_uint32, _vp], _i),
    "veig_zgeev_fp32": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp], _i),
    "veig_zgeev_mixed": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                          ctypes.c_double, ctypes.c_int32, _vp, ctypes.c_size_t, _vp], _i),
    "veig_mixed_workspace_size": ([_i64, _i64, _i64], ctypes.c_size_t),
    "veig_zgehrd": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp], _i),
    "veig_zunghr": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp], _i),
    "veig_hessenberg": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp], _i),
    "veig_zlahqr": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp], _i),
    "veig_ztrevc": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp], _i),
    "veig_zgemv": ([_i, _i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_double,
                    ctypes.c_double, _vp], _i),
    "veig_backsubstitution": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp], _i),
    "veig_zlahqr_eig": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp], _i),
    "veig_eigvals": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32, _vp], _i),
    "veig_zlahqr_roots": ([_i64, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                           ctypes.c_int32, _vp], _i),
    "veig_rearrange_matrices": ([_i64, _i64, _i64, _vp, _vp, _vp], _i),
    "veig_arrange_matrices_aos": ([_i64, _i64, _i64, _vp, _vp, _vp], _i),
    "veig_rep_matrix_soa": ([_i64, _i64, _i64, _vp, _vp, _vp], _i),
    "veig_rep_vector_soa": ([_i64, _i64, _i64, _vp, _vp, _vp], _i),
    "veig_host_eigensolver": ([_i, _i, _i, _D, _D, _D, _D, _D, _D], _i),
    "veig_host_zgehrd": ([_i, _i, _i, _D, _D, _D, _D], _i),
    "veig_host_zunghr": ([_i, _i, _i, _D, _D, _D, _D], _i),
    "veig_host_hessenberg": ([_i, _i, _i, _D, _D, _D, _D], _i),
    "veig_host_zlahqr_new": ([_i, _i, _i, _D, _D, _D, _D, _D, _D], _i),
    "veig_host_ztrevc": ([_i, _i, _i, _D, _D, _D, _D], _i),
    "veig_host_zgemv": ([_i, _i, _i, _i, _D, _D, _D, _D, _D, _D, ctypes.c_double, ctypes.c_double], _i),
    "veig_host_backsubstitution": ([_i, _i, _i, _D, _D, _D, _D], _i),
    "veig_varsize_workspace_size": ([_i64], ctypes.c_size_t),
    "veig_zgeev_varsize": ([_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32,
                            _vp, _vp, _vp, ctypes.c_size_t, _vp], _i),
    "veig_zgeev_varsize_noalloc": ([_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                    ctypes.c_int32, _vp, _vp, _vp], _i),
    "veig_varsize_graph_create": ([_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                   ctypes.c_int32, _vp, _vp, _vp, ctypes.c_size_t,
                                   ctypes.POINTER(ctypes.c_void_p)], _i),
    "veig_varsize_autotune": ([_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int32,
                               ctypes.POINTER(ctypes.c_int8), ctypes.POINTER(ctypes.c_int32), _vp,
                               ctypes.c_size_t, _vp], _i),
    "veig_graph_launch": ([_vp, _vp], _i),
    "veig_graph_destroy": ([_vp], _i),
}


def header_functions() -> list[str]:
    """Names of the functions declared in include/veig.h."""
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(veig_[a-z0-9_]+)\s*\(", txt)))


def lib():
    """Load libveig.so (once).  Raises VeigError if it is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise VeigError(f"{LIB_PATH} is not built; run vectorized-eigensolver_amd/build.py")
        # torch ships its own HIP runtime under the same soname as the system
        # one that libveig.so is linked against: whichever loads first serves
        # the whole process, and torch does not find the GPU on a runtime
        # other than its own.  Load torch's first (where torch is installed),
        # so that the library and torch share it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (argt, rest) in _SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.argtypes = argt
                f.restype = rest
        _lib = L
    return _lib


def check(status: int, what: str = "") -> None:
    if status != VEIG_OK:
        L = lib()
        raise VeigError(f"{what}: {L.veig_status_string(status).decode()} "
                        f"({L.veig_last_error().decode()})")


# --------------------------------------------------------------------------
# device-tensor helpers (torch)
# --------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def _ptr(t):
    """Device address of a GPU tensor (no shape check: callers that know the
    geometry use _arg)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise VeigError("device entry points need GPU tensors")
    if not t.is_contiguous():
        raise VeigError("tensors must be contiguous")
    return ctypes.c_void_p(t.data_ptr())


def _arg(t, name: str, dtype: str, numel: int, optional: bool = False):
    """`t` after checking what the kernels will assume of it: dtype (float64
    planes, int32 info/fallback/dims, int64 offsets, int8 variant table, uint8
    workspace), at least `numel` elements, contiguous.  An undersized or
    mistyped tensor would be read or written out of bounds by the kernels, so
    it is refused here; _call then checks that the tensors are on the GPU."""
    if t is None:
        if optional:
            return None
        raise VeigError(f"{name}: a tensor is required")
    torch = _torch()
    want = getattr(torch, dtype)
    if t.dtype != want:
        raise VeigError(f"{name}: dtype {t.dtype}, expected {want}")
    if t.numel() < numel:
        raise VeigError(f"{name}: {t.numel()} elements, the batch needs at least {numel}")
    if not t.is_contiguous():
        raise VeigError(f"{name}: tensors must be contiguous")
    return t


class _StreamRef:
    """A stream argument, resolved (torch's current stream if None) by _call
    once the tensors are known to be on the GPU."""

    def __init__(self, stream):
        self.stream = stream


def _call(f, *args):
    """f(*args) with every tensor argument (checked by _arg) passed as its
    device address; all of them must be GPU tensors."""
    if any(hasattr(a, "data_ptr") and not a.is_cuda for a in args):
        raise VeigError("device entry points need GPU tensors")
    conv = []
    for a in args:
        if hasattr(a, "data_ptr"):
            a = ctypes.c_void_p(a.data_ptr())
        elif isinstance(a, _StreamRef):
            a = _stream(a.stream)
        conv.append(a)
    return f(*conv)


def _geom(dim: int, block_cnt: int, thread_cnt: int):
    """(matrices, matrix-plane elements, vector-plane elements) of a batch."""
    if dim < 0 or block_cnt < 0 or thread_cnt < 1:
        raise VeigError("invalid batch geometry (dim < 0, block_cnt < 0 or thread_cnt < 1)")
    nm = block_cnt * thread_cnt
    return nm, nm * dim * dim, nm * dim


def _stream(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def zgeev(rA, iA, dim: int, block_cnt: int, thread_cnt: int, out=None, info=None,
          itmax: int = 30, write_schur: bool = False, precision: str = "fp64", stream=None):
    """Fused batched eigendecomposition on device tensors.

    precision: "fp64" (veig_zgeev) or "fp32" (veig_zgeev_fp32, no refinement).
    Returns (rE, iE, rV, iV)."""
    torch = _torch()
    nm, na, ne = _geom(dim, block_cnt, thread_cnt)
    dev = rA.device
    if out is None:
        rE = torch.empty(ne, dtype=torch.float64, device=dev)
        iE = torch.empty_like(rE)
        rV = torch.empty(na, dtype=torch.float64, device=dev)
        iV = torch.empty_like(rV)
    else:
        rE, iE, rV, iV = out
    if precision not in ("fp64", "fp32"):
        raise ValueError(precision)
    args = (_arg(rA, "rA", "float64", na), _arg(iA, "iA", "float64", na),
            _arg(rE, "rE", "float64", ne), _arg(iE, "iE", "float64", ne),
            _arg(rV, "rV", "float64", na), _arg(iV, "iV", "float64", na),
            _arg(info, "info", "int32", nm, optional=True))
    L = lib()
    if precision == "fp64":
        st = _call(L.veig_zgeev, dim, block_cnt, thread_cnt, *args, itmax,
                          VEIG_WRITE_SCHUR if write_schur else 0, _StreamRef(stream))
    else:
        st = _call(L.veig_zgeev_fp32, dim, block_cnt, thread_cnt, *args, itmax, _StreamRef(stream))
    check(st, "veig_zgeev")
    return rE, iE, rV, iV


def eigvals(rA, iA, dim: int, block_cnt: int, thread_cnt: int, out=None, info=None,
            itmax: int = 30, stream=None):
    """Eigenvalues only of general matrices (veig_eigvals).  Returns (rE, iE)."""
    torch = _torch()
    nm, na, ne = _geom(dim, block_cnt, thread_cnt)
    if out is None:
        rE = torch.empty(ne, dtype=torch.float64, device=rA.device)
        iE = torch.empty_like(rE)
    else:
        rE, iE = out
    st = _call(lib().veig_eigvals, dim, block_cnt, thread_cnt, _arg(rA, "rA", "float64", na),
                            _arg(iA, "iA", "float64", na), _arg(rE, "rE", "float64", ne),
                            _arg(iE, "iE", "float64", ne), _arg(info, "info", "int32", nm, True),
                            itmax, _StreamRef(stream))
    check(st, "veig_eigvals")
    return rE, iE


def zlahqr_eig(rH, iH, dim: int, block_cnt: int, thread_cnt: int, out=None, info=None,
               itmax: int = 30, stream=None):
    """The reference's eigenvalues-only zlahqr_ on Hessenberg matrices
    (veig_zlahqr_eig; H updated in place).  Returns (rE, iE)."""
    torch = _torch()
    nm, na, ne = _geom(dim, block_cnt, thread_cnt)
    if out is None:
        rE = torch.empty(ne, dtype=torch.float64, device=rH.device)
        iE = torch.empty_like(rE)
    else:
        rE, iE = out
    st = _call(lib().veig_zlahqr_eig, dim, block_cnt, thread_cnt, _arg(rE, "rE", "float64", ne),
                               _arg(iE, "iE", "float64", ne), _arg(rH, "rH", "float64", na),
                               _arg(iH, "iH", "float64", na), _arg(info, "info", "int32", nm, True),
                               itmax, _StreamRef(stream))
    check(st, "veig_zlahqr_eig")
    return rE, iE


def roots(rC, iC, dim: int, block_cnt: int, thread_cnt: int, schur: bool = False, info=None,
          itmax: int = 30, stream=None):
    """Roots of the polynomials c_0 + ... + c_dim x^dim (coefficients in the
    vector layout, dim+1 per polynomial) via their companion matrices
    (veig_zlahqr_roots).  Returns (rE, iE) or, with schur, (rE, iE, rT, iT,
    rZ, iZ)."""
    torch = _torch()
    nm, na, ne = _geom(dim, block_cnt, thread_cnt)
    dev = rC.device
    rE = torch.empty(ne, dtype=torch.float64, device=dev)
    iE = torch.empty_like(rE)
    mats = [None] * 4
    if schur:
        mats = [torch.empty(na, dtype=torch.float64, device=dev) for _ in range(4)]
    nc = nm * (dim + 1)
    st = _call(lib().veig_zlahqr_roots, dim, block_cnt, thread_cnt, _arg(rC, "rC", "float64", nc),
                                 _arg(iC, "iC", "float64", nc), _ptr(rE), _ptr(iE),
                                 *[_ptr(m) for m in mats], _arg(info, "info", "int32", nm, True),
                                 itmax, _StreamRef(stream))
    check(st, "veig_zlahqr_roots")
    return (rE, iE, *mats) if schur else (rE, iE)


def mixed_workspace(dim: int, block_cnt: int, thread_cnt: int, device="cuda"):
    torch = _torch()
    nbytes = lib().veig_mixed_workspace_size(dim, block_cnt, thread_cnt)
    return torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=device)


def zgeev_mixed(rA, iA, dim: int, block_cnt: int, thread_cnt: int, out=None, info=None,
                berr=None, fallback=None, tol: float = 1e-12, itmax: int = 30, workspace=None,
                stream=None):
    """Mixed precision: complex64 reduction/QR/eigenvectors, then FP64
    residual-driven refinement of every eigenpair to backward error <= tol,
    with an FP64 re-solve of the matrices that do not get there.
    Returns (rE, iE, rV, iV)."""
    torch = _torch()
    nm, na, ne = _geom(dim, block_cnt, thread_cnt)
    dev = rA.device
    if out is None:
        rE = torch.empty(ne, dtype=torch.float64, device=dev)
        iE = torch.empty_like(rE)
        rV = torch.empty(na, dtype=torch.float64, device=dev)
        iV = torch.empty_like(rV)
    else:
        rE, iE, rV, iV = out
    if workspace is None:
        workspace = mixed_workspace(dim, block_cnt, thread_cnt, dev)
    need = int(lib().veig_mixed_workspace_size(dim, block_cnt, thread_cnt))
    ws = _arg(workspace, "workspace", "uint8", need)
    st = _call(lib().veig_zgeev_mixed, dim, block_cnt, thread_cnt, _arg(rA, "rA", "float64", na),
                                _arg(iA, "iA", "float64", na), _arg(rE, "rE", "float64", ne),
                                _arg(iE, "iE", "float64", ne), _arg(rV, "rV", "float64", na),
                                _arg(iV, "iV", "float64", na), _arg(info, "info", "int32", nm, True),
                                _arg(fallback, "fallback", "int32", nm, True),
                                _arg(berr, "berr", "float64", nm, True), tol, itmax, ws,
                                workspace.numel() * workspace.element_size(), _StreamRef(stream))
    check(st, "veig_zgeev_mixed")
    return rE, iE, rV, iV


def _stage(name, dim, bc, tc, *planes, stream=None):
    """A stage entry (name, geometry, then its float64 planes as
    (tensor, label, 'm' for an n x n plane or 'v' for a vector plane))."""
    _, na, ne = _geom(dim, bc, tc)
    conv = [_arg(t, label, "float64", na if kind == "m" else ne) for t, label, kind in planes]
    f = getattr(lib(), name)
    check(_call(f, dim, bc, tc, *conv, _StreamRef(stream)), name)


def zgehrd(dim, bc, tc, rA, iA, rtau, itau, stream=None):
    _stage("veig_zgehrd", dim, bc, tc, (rA, "rA", "m"), (iA, "iA", "m"), (rtau, "rtau", "v"),
           (itau, "itau", "v"), stream=stream)


def zunghr(dim, bc, tc, rA, iA, rtau, itau, stream=None):
    _stage("veig_zunghr", dim, bc, tc, (rA, "rA", "m"), (iA, "iA", "m"), (rtau, "rtau", "v"),
           (itau, "itau", "v"), stream=stream)


def hessenberg(dim, bc, tc, rA, iA, rQ, iQ, stream=None):
    _stage("veig_hessenberg", dim, bc, tc, (rA, "rA", "m"), (iA, "iA", "m"), (rQ, "rQ", "m"),
           (iQ, "iQ", "m"), stream=stream)


def zlahqr(dim, bc, tc, rE, iE, rH, iH, rZ, iZ, info=None, itmax=30, stream=None):
    nm, na, ne = _geom(dim, bc, tc)
    f = lib().veig_zlahqr
    check(_call(f, dim, bc, tc, _arg(rE, "rE", "float64", ne), _arg(iE, "iE", "float64", ne),
            _arg(rH, "rH", "float64", na), _arg(iH, "iH", "float64", na),
            _arg(rZ, "rZ", "float64", na), _arg(iZ, "iZ", "float64", na),
            _arg(info, "info", "int32", nm, True), itmax, _StreamRef(stream)), "veig_zlahqr")


def ztrevc(dim, bc, tc, rT, iT, rR, iR, stream=None):
    _stage("veig_ztrevc", dim, bc, tc, (rT, "rT", "m"), (iT, "iT", "m"), (rR, "rR", "m"),
           (iR, "iR", "m"), stream=stream)


def layout(kind: str, dim, bc, tc, src, dst, stream=None):
    """The device layout converters: "soa" (AoS -> tiled SoA), "aos" (tiled
    SoA -> AoS), "repm" (one n x n matrix into every slot), "repv" (one
    vector of n into every slot); one float64 plane each."""
    name = {"soa": "veig_rearrange_matrices", "aos": "veig_arrange_matrices_aos",
            "repm": "veig_rep_matrix_soa", "repv": "veig_rep_vector_soa"}[kind]
    _, na, ne = _geom(dim, bc, tc)
    n_src = {"soa": na, "aos": na, "repm": dim * dim, "repv": dim}[kind]
    n_dst = ne if kind == "repv" else na
    f = getattr(lib(), name)
    check(_call(f, dim, bc, tc, _arg(src, "src", "float64", n_src), _arg(dst, "dst", "float64", n_dst),
            _StreamRef(stream)), name)


# --------------------------------------------------------------------------
# batches of different orders
# --------------------------------------------------------------------------
class VarBatch:
    """Device buffers for a batch of matrices of orders dims (1-D int tensor):
    A/V planes in the per-matrix column-major layout, E planes, offsets,
    info and workspace.  Fill rA/iA, then call run() or graph()."""

    def __init__(self, dims, device="cuda"):
        torch = _torch()
        d = torch.as_tensor(dims, dtype=torch.int64)
        self.nmat = int(d.numel())
        sq = d * d
        a_off = torch.zeros_like(d)
        e_off = torch.zeros_like(d)
        if self.nmat > 1:
            a_off[1:] = torch.cumsum(sq, 0)[:-1]
            e_off[1:] = torch.cumsum(d, 0)[:-1]
        self.total_a = int(sq.sum())
        self.total_e = int(d.sum())
        dev = torch.device(device)
        self.dims = d.to(torch.int32).to(dev)
        self.a_off = a_off.to(dev)
        self.e_off = e_off.to(dev)
        f64 = dict(dtype=torch.float64, device=dev)
        self.rA = torch.zeros(self.total_a, **f64)
        self.iA = torch.zeros(self.total_a, **f64)
        self.rV = torch.zeros(self.total_a, **f64)
        self.iV = torch.zeros(self.total_a, **f64)
        self.rE = torch.zeros(self.total_e, **f64)
        self.iE = torch.zeros(self.total_e, **f64)
        self.info = torch.zeros(self.nmat, dtype=torch.int32, device=dev)
        nbytes = lib().veig_varsize_workspace_size(self.nmat)
        self.ws = torch.empty(max(int(nbytes), 1), dtype=torch.uint8, device=dev)
        self.variant_table = None  # device int8[33]: kernel variant per order (None: defaults)
        self.grid = None           # 5 ints: persistent grid per variant (None / 0: defaults)

    def _args(self, itmax):
        grid = None
        if self.grid is not None:   # host array, read when the launches are enqueued
            self._grid_host = (ctypes.c_int32 * 5)(*[int(x) for x in self.grid])
            grid = ctypes.cast(self._grid_host, ctypes.c_void_p)
        n, ta, te = self.nmat, self.total_a, self.total_e
        return (n, _arg(self.dims, "dims", "int32", n), _arg(self.a_off, "a_off", "int64", n),
                _arg(self.e_off, "e_off", "int64", n), _arg(self.rA, "rA", "float64", ta),
                _arg(self.iA, "iA", "float64", ta), _arg(self.rE, "rE", "float64", te),
                _arg(self.iE, "iE", "float64", te), _arg(self.rV, "rV", "float64", ta),
                _arg(self.iV, "iV", "float64", ta), _arg(self.info, "info", "int32", n),
                itmax, _arg(self.variant_table, "variant_table", "int8", 33, optional=True), grid,
                _arg(self.ws, "workspace", "uint8", int(lib().veig_varsize_workspace_size(n))),
                self.ws.numel())

    def run(self, itmax: int = 30, stream=None):
        check(_call(lib().veig_zgeev_varsize, *self._args(itmax), _StreamRef(stream)), "veig_zgeev_varsize")

    def run_noalloc(self, itmax: int = 30, stream=None):
        """The same solve through veig_zgeev_varsize_noalloc (the library's
        static scratch instead of self.ws; at most 4,194,304 matrices)."""
        args = self._args(itmax)
        check(_call(lib().veig_zgeev_varsize_noalloc, *args[:14], _StreamRef(stream)),
              "veig_zgeev_varsize_noalloc")

    def graph(self, itmax: int = 30):
        """An instantiated hipGraph of the whole sequence (binning + solves)."""
        h = ctypes.c_void_p()
        check(_call(lib().veig_varsize_graph_create, *self._args(itmax), ctypes.byref(h)),
              "veig_varsize_graph_create")
        return Graph(h)

    def autotune(self, itmax: int = 30, stream=None):
        """Run the library's autotuner on this batch (overwrites the outputs)
        and keep its choice for run()/graph(); returns (variant_of_n, grid)."""
        torch = _torch()
        tab = (ctypes.c_int8 * 33)()
        grid = (ctypes.c_int32 * 5)()
        full = self._args(itmax)
        check(_call(lib().veig_varsize_autotune, *full[:12], tab, grid, full[14], self.ws.numel(),
                                          _StreamRef(stream)),
              "veig_varsize_autotune")
        dev = self.dims.device
        self.variant_table = torch.tensor(list(tab), dtype=torch.int8, device=dev)
        self.grid = list(grid)
        return list(tab), list(grid)


class Graph:
    def __init__(self, handle):
        self.h = handle

    def launch(self, stream=None):
        check(lib().veig_graph_launch(self.h, _stream(stream)), "veig_graph_launch")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.veig_graph_destroy(self.h)
            self.h = None

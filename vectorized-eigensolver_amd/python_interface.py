"""python_interface -- drop-in for the reference's Cython module
(python_interface.pyx:1-220): same function names, argument order and
in-place semantics on 1-D float64 numpy arrays in the tiled split-plane
layout, with the compute running on the gfx950 kernels through libveig.so.

    import sys; sys.path.insert(0, "<repo>/vectorized-eigensolver_amd")
    import python_interface as pi
    pi.my_zgehrd(dim, block_cnt, thread_cnt, rA, iA, rtau, itau)

Layout helpers (my_repMatrix, my_repVector, arrangeMatrices,
arrangeMatricesAoS) are host loops in the reference (utils.cpp:85-157): here
numpy on numpy arrays, and the device converters (veig.layout) on GPU tensors.  The test_* helpers draw their random matrices from an optional
numpy Generator (the reference's are unseeded, python_interface.pyx:98-220)
and return the same quantities.
"""
from __future__ import annotations

import ctypes

import numpy as np

import veig

_D = ctypes.POINTER(ctypes.c_double)


def _p(a: np.ndarray):
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.ndim == 1
            and a.flags.c_contiguous):
        raise TypeError("expected a contiguous 1-D float64 numpy array")
    return a.ctypes.data_as(_D)


# ---------------------------------------------------------------------------
# layout helpers (utils.cpp:85-157, python_interface.pyx:31-39, 54-57, 92-96)
# ---------------------------------------------------------------------------
def _on_gpu(*arrays) -> bool:
    """True if every array is a GPU tensor (the device converters run), False
    if none is (numpy); a mix is refused."""
    dev = [hasattr(a, "is_cuda") and a.is_cuda for a in arrays]
    if any(dev) and not all(dev):
        raise TypeError("give either numpy arrays or GPU tensors, not both")
    return all(dev)


def my_repMatrix(dim, block_cnt, thread_cnt, matrix, out):
    """repMatrixSoA: copy one column-major dim x dim plane into every slot
    (veig_rep_matrix_soa for GPU tensors)."""
    if _on_gpu(matrix, out):
        veig.layout("repm", dim, block_cnt, thread_cnt, matrix, out)
        return
    m = np.asarray(matrix, dtype=np.float64)[: dim * dim]
    v = out.reshape(block_cnt, dim * dim, thread_cnt)
    v[...] = m[None, :, None]


def my_repVector(dim, block_cnt, thread_cnt, vector, out):
    """repVectorSoA (veig_rep_vector_soa for GPU tensors)."""
    if _on_gpu(vector, out):
        veig.layout("repv", dim, block_cnt, thread_cnt, vector, out)
        return
    x = np.asarray(vector, dtype=np.float64)[:dim]
    v = out.reshape(block_cnt, dim, thread_cnt)
    v[...] = x[None, :, None]


def arrangeMatrices(dim, block_cnt, thread_cnt, inputMat, outputMat):
    """rearrangeMatrices: AoS (one column-major matrix after another) -> tiled
    SoA (veig_rearrange_matrices for GPU tensors)."""
    if _on_gpu(inputMat, outputMat):
        veig.layout("soa", dim, block_cnt, thread_cnt, inputMat, outputMat)
        return
    src = inputMat.reshape(block_cnt, thread_cnt, dim * dim)
    outputMat.reshape(block_cnt, dim * dim, thread_cnt)[...] = np.transpose(src, (0, 2, 1))


def arrangeMatricesAoS(dim, block_cnt, thread_cnt, inputMat, outputMat):
    """arrangeMatricesAoS (utils.cpp:105): tiled SoA -> AoS
    (veig_arrange_matrices_aos for GPU tensors)."""
    if _on_gpu(inputMat, outputMat):
        veig.layout("aos", dim, block_cnt, thread_cnt, inputMat, outputMat)
        return
    src = inputMat.reshape(block_cnt, dim * dim, thread_cnt)
    outputMat.reshape(block_cnt, thread_cnt, dim * dim)[...] = np.transpose(src, (0, 2, 1))


def extractMatrix(dim, thread_cnt, matrix):
    return matrix[::thread_cnt][: dim * dim].reshape(dim, dim).T


def extractVector(dim, thread_cnt, vector):
    return vector[::thread_cnt][:dim]


# ---------------------------------------------------------------------------
# compute wrappers (python_interface.pyx:41-90), all on the GPU
# ---------------------------------------------------------------------------
def _call(name, *args):
    L = veig.lib()
    veig.check(getattr(L, name)(*args), name)


def my_zlahqr(dim, block_cnt, thread_cnt, rE, iE, rM, iM, rS, iS):
    _call("veig_host_zlahqr_new", dim, block_cnt, thread_cnt, _p(rE), _p(iE), _p(rM), _p(iM),
          _p(rS), _p(iS))


def my_backsubs(dim, block_cnt, thread_cnt, rT, iT, rb, ib):
    _call("veig_host_backsubstitution", dim, block_cnt, thread_cnt, _p(rT), _p(iT), _p(rb), _p(ib))


def my_ztrevec(dim, block_cnt, thread_cnt, rT, iT, rR, iR):
    _call("veig_host_ztrevc", dim, block_cnt, thread_cnt, _p(rT), _p(iT), _p(rR), _p(iR))


def my_zgemv(hermitian, dim, block_cnt, thread_cnt, rA, iA, rx, ix, rbeta, ibeta, ry, iy):
    _call("veig_host_zgemv", int(bool(hermitian)), dim, block_cnt, thread_cnt, _p(rA), _p(iA),
          _p(rx), _p(ix), _p(ry), _p(iy), float(rbeta), float(ibeta))


def my_zgehrd(dim, block_cnt, thread_cnt, rA, iA, rtau, itau):
    _call("veig_host_zgehrd", dim, block_cnt, thread_cnt, _p(rA), _p(iA), _p(rtau), _p(itau))


def my_zunghr(dim, block_cnt, thread_cnt, rA, iA, rtau, itau):
    _call("veig_host_zunghr", dim, block_cnt, thread_cnt, _p(rA), _p(iA), _p(rtau), _p(itau))


def my_hessenberg(dim, block_cnt, thread_cnt, rA, iA, rQ, iQ):
    _call("veig_host_hessenberg", dim, block_cnt, thread_cnt, _p(rA), _p(iA), _p(rQ), _p(iQ))


def my_eigensolver(dim, block_cnt, thread_cnt, rA, iA, rE, iE, rV, iV):
    """eigensolver(dim, block_cnt, thread_cnt, rA, iA, rE, iE, rV, iV)
    (utils.cpp:578; the reference has no Python binding for it).  A is
    overwritten with the Schur form."""
    _call("veig_host_eigensolver", dim, block_cnt, thread_cnt, _p(rA), _p(iA), _p(rE), _p(iE),
          _p(rV), _p(iV))


# ---------------------------------------------------------------------------
# test helpers (python_interface.pyx:98-220); same return values
# ---------------------------------------------------------------------------
def _rng(rng):
    return rng if rng is not None else np.random.default_rng()


def test_backsubs(dim, block_cnt, thread_cnt, rng=None):
    g = _rng(rng)
    a = np.triu(g.random((dim, dim)))
    ra = np.zeros(dim * dim * thread_cnt * block_cnt)
    ia = np.zeros_like(ra)
    my_repMatrix(dim, block_cnt, thread_cnt, a.T.reshape(dim * dim), ra)
    x = g.random(dim)
    rx = np.zeros(dim * thread_cnt * block_cnt)
    ix = np.zeros_like(rx)
    my_repVector(dim, block_cnt, thread_cnt, x, rx)
    my_backsubs(dim, block_cnt, thread_cnt, ra, ia, rx, ix)
    sol = extractVector(dim, thread_cnt, rx)
    return a.dot(sol) - x


def test_zgehrd(dim, block_cnt, thread_cnt, rng=None):
    g = _rng(rng)
    a = g.random((dim, dim))
    ra = np.zeros(dim * dim * thread_cnt * block_cnt)
    ia = np.zeros_like(ra)
    my_repMatrix(dim, block_cnt, thread_cnt, a.T.reshape(dim * dim), ra)
    rtau = np.zeros(dim * thread_cnt * block_cnt)
    itau = np.zeros_like(rtau)
    my_zgehrd(dim, block_cnt, thread_cnt, ra, ia, rtau, itau)
    erg = extractMatrix(dim, thread_cnt, ra) + 1.0j * extractMatrix(dim, thread_cnt, ia)
    return erg, a


def test_hessenberg(dim, block_cnt, thread_cnt, rng=None):
    g = _rng(rng)
    a = g.random((dim, dim))
    ra = np.zeros(dim * dim * thread_cnt * block_cnt)
    ia = np.zeros_like(ra)
    my_repMatrix(dim, block_cnt, thread_cnt, a.T.reshape(dim * dim), ra)
    rq = np.zeros_like(ra)
    iq = np.zeros_like(rq)
    my_hessenberg(dim, block_cnt, thread_cnt, ra, ia, rq, iq)
    h = extractMatrix(dim, thread_cnt, ra) + 1.0j * extractMatrix(dim, thread_cnt, ia)
    q = extractMatrix(dim, thread_cnt, rq) + 1.0j * extractMatrix(dim, thread_cnt, iq)
    return h, q, a


def test_zgemv(dim, block_cnt, thread_cnt, hermitian=False, rng=None):
    g = _rng(rng)
    a = g.random((dim, dim))
    ra = np.zeros(dim * dim * thread_cnt * block_cnt)
    ia = np.zeros_like(ra)
    my_repMatrix(dim, block_cnt, thread_cnt, a.T.reshape(dim * dim), ra)
    y = g.random(dim)
    ry = np.zeros(dim * thread_cnt * block_cnt)
    iy = np.zeros_like(ry)
    my_repVector(dim, block_cnt, thread_cnt, y, ry)
    x = g.random(dim)
    rx = np.zeros_like(iy)
    ix = np.zeros_like(rx)
    my_repVector(dim, block_cnt, thread_cnt, x, rx)
    my_zgemv(hermitian, dim, block_cnt, thread_cnt, ra, ia, rx, ix, 1.0, 0.0, ry, iy)
    axpy = extractVector(dim, thread_cnt, ry)
    sol = (np.conj(a.T).dot(x) + y) if hermitian else (a.dot(x) + y)
    return np.abs(axpy - sol)


def test_zlahqr(dim, block_cnt, thread_cnt, rng=None):
    g = _rng(rng)
    dim2 = dim * dim
    size = dim2 * block_cnt * thread_cnt
    sizev = dim * block_cnt * thread_cnt
    a = np.triu(g.random((dim, dim)), -1)
    ra = np.zeros(size)
    ia = np.zeros_like(ra)
    my_repMatrix(dim, block_cnt, thread_cnt, a.T.reshape(dim2), ra)
    s = np.eye(dim)
    rschur = np.zeros(size)
    ischur = np.zeros(size)
    my_repMatrix(dim, block_cnt, thread_cnt, s.T.reshape(dim2), rschur)
    re = np.zeros(sizev)
    ie = np.zeros(sizev)
    my_zlahqr(dim, block_cnt, thread_cnt, re, ie, ra, ia, rschur, ischur)
    eigs = extractVector(dim, thread_cnt, re) + 1.0j * extractVector(dim, thread_cnt, ie)
    schurv = extractMatrix(dim, thread_cnt, rschur) + 1.0j * extractMatrix(dim, thread_cnt, ischur)
    t = extractMatrix(dim, thread_cnt, ra) + 1.0j * extractMatrix(dim, thread_cnt, ia)
    return (eigs, schurv, t, a)


def test_eigenvectors(dim, block_cnt, thread_cnt, rng=None):
    g = _rng(rng)
    dim2 = dim * dim
    size = dim2 * block_cnt * thread_cnt
    sizev = dim * block_cnt * thread_cnt
    a = np.triu(g.random((dim, dim)), -1)
    ra = np.zeros(size)
    ia = np.zeros_like(ra)
    my_repMatrix(dim, block_cnt, thread_cnt, a.T.reshape(dim2), ra)
    s = np.eye(dim)
    rsch = np.zeros(size)
    isch = np.zeros(size)
    my_repMatrix(dim, block_cnt, thread_cnt, s.T.reshape(dim2), rsch)
    re = np.zeros(sizev)
    ie = np.zeros(sizev)
    my_zlahqr(dim, block_cnt, thread_cnt, re, ie, ra, ia, rsch, isch)
    my_ztrevec(dim, block_cnt, thread_cnt, ra, ia, rsch, isch)
    eigs = extractVector(dim, thread_cnt, re) + 1.0j * extractVector(dim, thread_cnt, ie)
    vecs = extractMatrix(dim, thread_cnt, rsch) + 1.0j * extractMatrix(dim, thread_cnt, isch)
    return (eigs, vecs, a)

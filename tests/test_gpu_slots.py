"""The per-matrix entries of the header library in veig_compat.hpp --
zgehrd (Toolbox.hpp:1315), zunghr (1492), zlahqr_schur (690) and ztrevc
(1133) -- called on ONE slot of a tiled batch through pointers pre-offset to
it, as the reference's batch loops pass them (Toolbox.hpp:2150-2177): each
must give exactly what the matching batch entry (test_zgehrd, test_zunghr,
zlahqr_new, test_ztrevc) gives for that slot, and leave every other slot
untouched."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import helpers as H
import veig

pytestmark = pytest.mark.gpu

DP = ctypes.POINTER(ctypes.c_double)
BC, TC = 2, 4
SLOTS = [(0, 0), (1, 2), (1, 3)]          # (block, thread)
# orders on the lane-group stage kernels (1, 7, 16) and on the one-workgroup
# kernels (40)
DIMS = [1, 7, 16, 40]


def _fn(name, argtypes):
    f = getattr(veig.lib(), name)
    f.argtypes = argtypes
    f.restype = None
    return f


def _at(a: np.ndarray, off: int):
    """a pointer into a, pre-offset by off doubles"""
    return ctypes.cast(a.ctypes.data + 8 * off, DP)


def _moff(b, t, DIM):
    return t + b * TC * DIM * DIM


def _voff(b, t, DIM):
    return t + b * TC * DIM


def _slot_mask(size_per, b, t):
    """boolean mask of one slot's elements in a tiled plane of size_per entries per matrix"""
    m = np.zeros(BC * TC * size_per, dtype=bool)
    m[(b * TC * size_per + t)::TC][:size_per] = True
    return m


def _planes(A):
    return [np.ascontiguousarray(x) for x in H.to_soa(A, TC)]


def _vec(N, DIM, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(N * DIM), rng.standard_normal(N * DIM)


@pytest.mark.parametrize("DIM", DIMS)
@pytest.mark.parametrize("b,t", SLOTS)
def test_zgehrd_zunghr_slot(b, t, DIM):
    N = BC * TC
    A = H.gaussian(N, DIM, 2101)
    rA0, iA0 = _planes(A)
    rt0, it0 = _vec(N, DIM, 2102)                 # tau planes: entries past the reflectors stay
    batch_g = _fn("_Z11test_zgehrdPdS_S_S_S_S_iii", [DP] * 6 + [ctypes.c_int] * 3)
    batch_q = _fn("_Z11test_zunghrPdS_S_S_S_S_iii", [DP] * 6 + [ctypes.c_int] * 3)
    slot_g = _fn("_Z6zgehrdPdS_S_S_S_S_iii", [DP] * 6 + [ctypes.c_int] * 3)
    slot_q = _fn("_Z6zunghrPdS_S_S_S_S_iii", [DP] * 6 + [ctypes.c_int] * 3)
    # batch
    rA, iA, rt, it = rA0.copy(), iA0.copy(), rt0.copy(), it0.copy()
    batch_g(*(H.dptr(x) for x in (rA, iA, rt, it)), None, None, DIM, BC, TC)
    rH, iH, rtH, itH = rA.copy(), iA.copy(), rt.copy(), it.copy()
    batch_q(*(H.dptr(x) for x in (rA, iA, rt, it)), None, None, DIM, BC, TC)
    # one slot
    sA, siA, st, sit = rA0.copy(), iA0.copy(), rt0.copy(), it0.copy()
    mo, vo = _moff(b, t, DIM), _voff(b, t, DIM)
    slot_g(_at(sA, mo), _at(siA, mo), _at(st, vo), _at(sit, vo), None, None, DIM, BC, TC)
    mm, mv = _slot_mask(DIM * DIM, b, t), _slot_mask(DIM, b, t)
    assert np.array_equal(sA[mm], rH[mm]) and np.array_equal(siA[mm], iH[mm])
    assert np.array_equal(st[mv], rtH[mv]) and np.array_equal(sit[mv], itH[mv])
    assert np.array_equal(sA[~mm], rA0[~mm]) and np.array_equal(st[~mv], rt0[~mv])
    slot_q(_at(sA, mo), _at(siA, mo), _at(st, vo), _at(sit, vo), None, None, DIM, BC, TC)
    assert np.array_equal(sA[mm], rA[mm]) and np.array_equal(siA[mm], iA[mm])
    assert np.array_equal(sA[~mm], rA0[~mm]) and np.array_equal(siA[~mm], iA0[~mm])
    Q = H.from_soa(sA, siA, DIM, TC)[b * TC + t]
    assert np.allclose(Q.conj().T @ Q, np.eye(DIM), atol=1e-14)


@pytest.mark.parametrize("DIM", DIMS)
@pytest.mark.parametrize("b,t", SLOTS)
@pytest.mark.parametrize("want_z", [True, False])
def test_zlahqr_schur_slot(b, t, want_z, DIM):
    N = BC * TC
    Hm = H.hessenberg_input(N, DIM, 2103)
    rng = np.random.default_rng(2104)
    Z0 = rng.standard_normal((N, DIM, DIM)) + 1j * rng.standard_normal((N, DIM, DIM))
    rV0, iV0 = _planes(Hm)
    rZ0, iZ0 = _planes(Z0)
    rE0, iE0 = _vec(N, DIM, 2105)
    batch = _fn("_Z10zlahqr_newPdS_S_S_S_S_iiii", [DP] * 6 + [ctypes.c_int] * 4)
    slot = _fn("_Z12zlahqr_schurbbiPdS_S_S_S_S_ii", [ctypes.c_bool, ctypes.c_bool, ctypes.c_int] + [DP] * 6 +
               [ctypes.c_int] * 2)
    rV, iV, rZ, iZ, rE, iE = (x.copy() for x in (rV0, iV0, rZ0, iZ0, rE0, iE0))
    batch(*(H.dptr(x) for x in (rE, iE, rV, iV, rZ, iZ)), DIM, DIM * DIM, BC, TC)
    sV, siV, sZ, siZ, sE, siE = (x.copy() for x in (rV0, iV0, rZ0, iZ0, rE0, iE0))
    mo, vo = _moff(b, t, DIM), _voff(b, t, DIM)
    slot(True, want_z, DIM, _at(sV, mo), _at(siV, mo), _at(sE, vo), _at(siE, vo), _at(sZ, mo), _at(siZ, mo), BC, TC)
    mm, mv = _slot_mask(DIM * DIM, b, t), _slot_mask(DIM, b, t)
    assert np.array_equal(sV[mm], rV[mm]) and np.array_equal(siV[mm], iV[mm])
    assert np.array_equal(sE[mv], rE[mv]) and np.array_equal(siE[mv], iE[mv])
    if want_z:
        assert np.array_equal(sZ[mm], rZ[mm]) and np.array_equal(siZ[mm], iZ[mm])
    else:
        assert np.array_equal(sZ, rZ0) and np.array_equal(siZ, iZ0)
    assert np.array_equal(sV[~mm], rV0[~mm]) and np.array_equal(sE[~mv], rE0[~mv])


@pytest.mark.parametrize("DIM", DIMS)
@pytest.mark.parametrize("b,t", SLOTS)
def test_ztrevc_slot(b, t, DIM):
    N = BC * TC
    A = H.gaussian(N, DIM, 2106)
    E, V, T = H.oracle_eig(A, TC)
    # a Schur form and Schur vectors: T from the oracle's zlahqr, Z unitary
    rng = np.random.default_rng(2107)
    Zr = np.linalg.qr(rng.standard_normal((N, DIM, DIM)) + 1j * rng.standard_normal((N, DIM, DIM)))[0]
    rT0, iT0 = _planes(np.triu(T))
    rR0, iR0 = _planes(Zr)
    batch = _fn("_Z11test_ztrevcPdS_S_S_S_S_iiii", [DP] * 6 + [ctypes.c_int] * 4)
    slot = _fn("_Z6ztrevcPdS_S_S_S_S_S_S_iii", [DP] * 8 + [ctypes.c_int] * 3)
    rT, iT, rR, iR = (x.copy() for x in (rT0, iT0, rR0, iR0))
    batch(*(H.dptr(x) for x in (rT, iT, rR, iR)), None, None, DIM, DIM * DIM, BC, TC)
    sT, siT, sR, siR = (x.copy() for x in (rT0, iT0, rR0, iR0))
    mo = _moff(b, t, DIM)
    slot(_at(sT, mo), _at(siT, mo), None, None, _at(sR, mo), _at(siR, mo), None, None, DIM, BC, TC)
    mm = _slot_mask(DIM * DIM, b, t)
    assert np.array_equal(sR[mm], rR[mm]) and np.array_equal(siR[mm], iR[mm])
    assert np.array_equal(sR[~mm], rR0[~mm]) and np.array_equal(sT, rT0) and np.array_equal(siT, iT0)

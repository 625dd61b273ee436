"""The exported scalar complex API, include/veig_complex.hpp (the drop-in for
ComplexDouble.hpp:110-488), on the host and compiled for gfx950, checked
against what the mathematics fixes:

* cinv and cdiv (Smith's method) within a few ulp of the exact quotient,
  computed in extended precision, and finite where a naive |b|^2 overflows or
  underflows (|b| ~ 1e+-200);
* csqrt on the principal branch (Re >= 0, the cut on the negative reals
  approached from above: sqrt(-4 + 0i) = 2i), squaring back to its argument;
* dabs2 / dabs3 as scaled norms: exact-to-ulp hypot and 3-norm, finite at
  1e200 where the unscaled cabs overflows (the reference's documented limit,
  ComplexDouble.hpp:439-449);
* the ComplexDouble compound operators as the arithmetic they spell.

The probe sources (tests/probe/) are built by build(); the device test runs
the same element function in a kernel.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import pytest

PROBE = Path(__file__).resolve().parent / "probe"
HOST_LIB = PROBE / "libcomplex_probe_host.so"
DEV_LIB = PROBE / "libcomplex_probe.so"

OPS = dict(cinv=0, cdiv=1, csqrt=2, dabs2=3, dabs3=4, cmul=5, cmulc=6, cabs=7, cd_ops=8,
           cd_inv_sqrt=9, cdiv_real=10)


def _host():
    if not HOST_LIB.exists():
        pytest.fail(f"{HOST_LIB} is not built; run __graft_entry__.build()")
    L = ctypes.CDLL(str(HOST_LIB))
    L.probe_host.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    return L


def run_host(op, a):
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 4)
    out = np.zeros((a.shape[0], 2))
    _host().probe_host(OPS[op], a.shape[0], a.ctypes.data, out.ctypes.data)
    return out


def run_device(op, a):
    import torch
    L = ctypes.CDLL(str(DEV_LIB))
    L.probe_device.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    ta = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 4)).cuda()
    out = torch.zeros((ta.shape[0], 2), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert L.probe_device(OPS[op], ta.shape[0], ta.data_ptr(), out.data_ptr()) == 0
    return out.cpu().numpy()


def samples(n=4000, seed=5):
    rng = np.random.default_rng(seed)
    mag = 10.0 ** rng.uniform(-8, 8, size=(n, 4))
    a = rng.standard_normal((n, 4)) * mag
    # exact zeros and equal magnitudes in some components (Smith's branches)
    a[::13, 2] = a[::13, 3]
    a[::7, 1] = 0.0
    a[::11, 3] = 0.0
    return a


def cx(a, k):
    return a[:, 2 * k].astype(np.longdouble) + 1j * a[:, 2 * k + 1].astype(np.longdouble)


def ulp_err(got, want):
    """max over components of |got - want| / ulp(|want|) (want in long double)."""
    w = np.abs(want).astype(np.float64)
    scale = np.spacing(np.maximum(w, np.finfo(float).tiny))
    g = got[:, 0] + 1j * got[:, 1]
    return np.abs((g.astype(np.clongdouble) - want)).astype(np.float64) / scale


BACKENDS = [pytest.param(run_host, id="host"),
            pytest.param(run_device, id="gfx950", marks=pytest.mark.gpu)]


@pytest.mark.parametrize("run", BACKENDS)
def test_division_and_inverse(run):
    a = samples()
    want = cx(a, 0) / cx(a, 1)
    assert np.max(ulp_err(run("cdiv", a), want)) <= 4
    want_inv = 1 / cx(a, 0)
    assert np.max(ulp_err(run("cinv", a), want_inv)) <= 4
    # |b|^2 would overflow / underflow; Smith's scaling keeps the quotient finite and exact
    big = np.array([[3e200, -4e200, 1e200, 2e200], [1.0, 1.0, 3e-200, 4e-200],
                    [5e-200, 0.0, 0.0, 2e-200], [7.0, -1.0, 1e300, 1e300]])
    got = run("cdiv", big)
    assert np.all(np.isfinite(got))
    assert np.max(ulp_err(got, cx(big, 0) / cx(big, 1))) <= 4
    inv = run("cinv", np.hstack([big[:, 2:], big[:, :2]]))  # 1 / b (the probe inverts its first argument)
    assert np.all(np.isfinite(inv))
    assert np.max(ulp_err(inv, 1 / cx(big, 1))) <= 4
    # division by a real
    d = run("cdiv_real", a)
    assert np.max(ulp_err(d, cx(a, 0) / a[:, 2].astype(np.longdouble))) <= 1


@pytest.mark.parametrize("run", BACKENDS)
def test_principal_sqrt(run):
    a = samples(seed=6)
    s = run("csqrt", a)
    assert np.all(s[:, 0] >= 0)
    z = s[:, 0] + 1j * s[:, 1]
    x = a[:, 0] + 1j * a[:, 1]
    assert np.allclose(z * z, x, rtol=1e-14, atol=0)
    # numpy's principal square root off the cut; sign of Im follows Im(x)
    assert np.allclose(z, np.sqrt(x), rtol=1e-15, atol=0)
    cut = np.array([[-4.0, 0.0, 0, 0], [-4.0, -0.0, 0, 0], [-1e-100, 0.0, 0, 0], [9.0, 0.0, 0, 0],
                    [0.0, 0.0, 0, 0], [0.0, 2.0, 0, 0], [0.0, -2.0, 0, 0]])
    c = run("csqrt", cut)
    assert np.array_equal(c[0], [0.0, 2.0])     # approached from above: +2i
    assert np.array_equal(c[1], [0.0, 2.0])     # -0 imaginary part: the reference tests Im < 0, so +2i
    assert c[2, 0] == 0.0 and np.isclose(c[2, 1], 1e-50, rtol=1e-15)
    assert np.array_equal(c[3], [3.0, 0.0])
    assert np.array_equal(c[4], [0.0, 0.0])
    assert np.allclose(c[5], [1.0, 1.0], rtol=1e-15) and np.allclose(c[6], [1.0, -1.0], rtol=1e-15)


@pytest.mark.parametrize("run", BACKENDS)
def test_scaled_norms(run):
    a = samples(seed=7)
    h = run("dabs2", a)[:, 0]
    want = np.hypot(a[:, 0].astype(np.longdouble), a[:, 1].astype(np.longdouble))
    assert np.max(np.abs(h - want) / np.spacing(want.astype(float))) <= 2
    n3 = run("dabs3", a)[:, 0]
    w3 = np.sqrt(np.sum(a[:, :3].astype(np.longdouble) ** 2, axis=1))
    assert np.max(np.abs(n3 - w3) / np.spacing(w3.astype(float))) <= 3
    huge = np.array([[1e200, 1e200, 1e200, 0.0], [3e300, 4e300, 0.0, 0.0], [0.0, 0.0, 0.0, 0.0],
                     [1e-300, 1e-300, 1e-300, 0.0]])
    h = run("dabs2", huge)[:, 0]
    assert np.isclose(h[0], np.sqrt(2) * 1e200, rtol=1e-15) and np.isclose(h[1], 5e300, rtol=1e-15)
    assert h[2] == 0.0 and np.isclose(h[3], np.sqrt(2) * 1e-300, rtol=1e-15)
    n3 = run("dabs3", huge)[:, 0]
    assert np.isclose(n3[0], np.sqrt(3) * 1e200, rtol=1e-15) and n3[2] == 0.0
    # cabs1 and the unscaled cabs (which overflows at 1e200, as the reference's does)
    c = run("cabs", np.vstack([a[:5], huge[:1]]))
    assert np.allclose(c[:5, 0], np.abs(a[:5, 0]) + np.abs(a[:5, 1]), rtol=1e-15)
    assert np.allclose(c[:5, 1], np.hypot(a[:5, 0], a[:5, 1]), rtol=1e-15)
    assert np.isinf(c[5, 1])


@pytest.mark.parametrize("run", BACKENDS)
def test_products_and_value_type(run):
    rng = np.random.default_rng(8)
    a = rng.standard_normal((500, 4))
    x, y = a[:, 0] + 1j * a[:, 1], a[:, 2] + 1j * a[:, 3]
    m = run("cmul", a)
    assert np.allclose(m[:, 0] + 1j * m[:, 1], x * y, rtol=1e-15, atol=1e-15)
    mc = run("cmulc", a)
    assert np.allclose(mc[:, 0] + 1j * mc[:, 1], x * np.conj(y), rtol=1e-15, atol=1e-15)
    v = run("cd_ops", a)
    assert np.allclose(v[:, 0] + 1j * v[:, 1], ((x + y) * y - x) / y / 2, rtol=1e-13, atol=1e-13)
    w = run("cd_inv_sqrt", a)
    assert np.allclose(w[:, 0] + 1j * w[:, 1], np.sqrt(1 / x), rtol=1e-14)


def _scale_rc_host(rV, iV, zr, zi, I, I1, I2, dim, n):
    L = _host()
    L.probe_scale_rc_host.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double] + \
        [ctypes.c_int] * 5
    L.probe_scale_rc_host(rV.ctypes.data, iV.ctypes.data, zr, zi, I, I1, I2, dim, n)
    return rV, iV


def _scale_rc_device(rV, iV, zr, zi, I, I1, I2, dim, n):
    import torch
    L = ctypes.CDLL(str(DEV_LIB))
    L.probe_scale_rc_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double] + \
        [ctypes.c_int] * 5
    tr, ti = torch.from_numpy(rV).cuda(), torch.from_numpy(iV).cuda()
    torch.cuda.synchronize()
    assert L.probe_scale_rc_device(tr.data_ptr(), ti.data_ptr(), zr, zi, I, I1, I2, dim, n) == 0
    return tr.cpu().numpy(), ti.cpu().numpy()


@pytest.mark.parametrize("run", [pytest.param(_scale_rc_host, id="host"),
                                 pytest.param(_scale_rc_device, id="gfx950", marks=pytest.mark.gpu)])
@pytest.mark.parametrize("dim,n,I,I1,I2", [(6, 1, 3, 1, 6), (6, 4, 4, 2, 5), (9, 3, 1, 1, 9), (9, 3, 9, 1, 9),
                                           (5, 2, 2, 2, 2)])
def test_scale_rc(run, dim, n, I, I1, I2):
    """scale_rc (Toolbox.hpp:14-29) in the strided layout: row I (1-based)
    over columns I+1..I2 times z, column I over rows I1..I-1 times conj(z),
    every other element untouched -- including the other slots between the
    matrix's elements."""
    rng = np.random.default_rng(dim * 100 + I)
    M = rng.standard_normal((dim * dim * n,)) + 1j * rng.standard_normal((dim * dim * n,))
    z = complex(0.6, -0.8)
    rV, iV = run(M.real.copy(), M.imag.copy(), z.real, z.imag, I, I1, I2, dim, n)
    want = M.copy()
    idx = lambda r, c: ((r - 1) + (c - 1) * dim) * n  # 1-based (r, c) of the slot at offset 0
    for c in range(I + 1, I2 + 1):
        want[idx(I, c)] *= z
    for r in range(I1, I):
        want[idx(r, I)] *= np.conj(z)
    got = rV + 1j * iV
    assert np.allclose(got, want, rtol=5e-16, atol=0)  # one rounding of a complex product
    changed = np.flatnonzero(got != M)
    touched = {idx(I, c) for c in range(I + 1, I2 + 1)} | {idx(r, I) for r in range(I1, I)}
    assert set(changed.tolist()) <= touched

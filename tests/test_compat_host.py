"""The C++ drop-in shim (include/veig_compat.hpp): libveig.so exports the
reference's host API under its C++ names (utils.h:37-65, Toolbox-def.hpp:5-22),
and its host-side layout helpers behave like the reference's
(utils.cpp:85-157) -- checked against the reference build where present,
else against the layout formula.  No GPU needed."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import helpers as H
import veig

# Itanium-mangled names of the reference's functions (same signatures)
MANGLED = {
    "eigensolver": "_Z11eigensolveriiiPdS_S_S_S_S_",
    "call_zgehrd": "_Z11call_zgehrdiiiPdS_S_S_",
    "call_zunghr": "_Z11call_zunghriiiPdS_S_S_",
    "call_zlahqr_new": "_Z15call_zlahqr_newiiiPdS_S_S_S_S_",
    "call_ztrevc": "_Z11call_ztrevciiiPdS_S_S_",
    "call_zgemv": "_Z10call_zgemvbiiiPdS_S_S_S_S_dd",
    "call_backsubstitution": "_Z21call_backsubstitutioniiiPdS_S_S_",
    "hessenberg": "_Z10hessenbergiiiPdS_S_S_",
    "repMatrixSoA": "_Z12repMatrixSoAiiiPdS_",
    "repVectorSoA": "_Z12repVectorSoAiiiPdS_",
    "rearrangeMatrices": "_Z17rearrangeMatricesiiiPdS_",
    "arrangeMatricesAoS": "_Z18arrangeMatricesAoSiiiPdS_",
    "test_zgehrd": "_Z11test_zgehrdPdS_S_S_S_S_iii",
    "test_zunghr": "_Z11test_zunghrPdS_S_S_S_S_iii",
    "zlahqr_new": "_Z10zlahqr_newPdS_S_S_S_S_iiii",
    "test_ztrevc": "_Z11test_ztrevcPdS_S_S_S_S_iiii",
    "test_zgemv": "_Z10test_zgemvbPdS_S_S_ddS_S_iii",
    "test_ident": "_Z10test_identPdS_iii",
    "test_diag": "_Z9test_diagPdS_ddiii",
    "test_backsubstitution": "_Z21test_backsubstitutionPdS_S_S_iii",
    "zlahqr": "_Z6zlahqrPdS_S_S_S_S_S_S_iiiii",
    "loadPackages": "_Z12loadPackagesiiiP13ComplexDouble",
}


def sym(name):
    return getattr(veig.lib(), MANGLED[name])


def test_reference_names_exported():
    L = veig.lib()
    missing = [n for n, m in MANGLED.items() if not hasattr(L, m)]
    assert not missing, missing


def _layout_fn(name):
    f = sym(name)
    f.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_double)] * 2
    f.restype = None
    return f


@pytest.mark.parametrize("dim,bc,tc", [(3, 2, 5), (8, 4, 4), (1, 3, 1)])
def test_layout_helpers(dim, bc, tc):
    N = bc * tc
    rng = np.random.default_rng(dim * 100 + tc)
    aos = rng.standard_normal(N * dim * dim)
    soa = np.zeros_like(aos)
    _layout_fn("rearrangeMatrices")(dim, bc, tc, H.dptr(aos), H.dptr(soa))
    A = aos.reshape(N, dim, dim).transpose(0, 2, 1)          # column-major per matrix
    assert np.array_equal(soa, H.to_soa(A.astype(complex), tc)[0])
    back = np.zeros_like(aos)
    _layout_fn("arrangeMatricesAoS")(dim, bc, tc, H.dptr(soa), H.dptr(back))
    assert np.array_equal(back, aos)
    m = rng.standard_normal(dim * dim)
    rep = np.zeros(N * dim * dim)
    _layout_fn("repMatrixSoA")(dim, bc, tc, H.dptr(m), H.dptr(rep))
    Mt = np.broadcast_to(m.reshape(dim, dim).T, (N, dim, dim)).astype(complex)
    assert np.array_equal(rep, H.to_soa(Mt, tc)[0])
    v = rng.standard_normal(dim)
    repv = np.zeros(N * dim)
    _layout_fn("repVectorSoA")(dim, bc, tc, H.dptr(v), H.dptr(repv))
    assert np.array_equal(repv, H.vec_to_soa(np.tile(v, (N, 1)).astype(complex), tc)[0])


def test_diag_and_ident():
    dim, bc, tc = 4, 2, 3
    N = bc * tc
    f = sym("test_diag")
    f.argtypes = [ctypes.POINTER(ctypes.c_double)] * 2 + [ctypes.c_double] * 2 + [ctypes.c_int] * 3
    r = np.zeros(N * dim * dim)
    i = np.zeros(N * dim * dim)
    f(H.dptr(r), H.dptr(i), 2.0, -1.0, dim, bc, tc)
    M = H.from_soa(r, i, dim, tc)
    assert np.array_equal(M, np.broadcast_to((2.0 - 1.0j) * np.eye(dim), M.shape))
    g = sym("test_ident")
    g.argtypes = [ctypes.POINTER(ctypes.c_double)] * 2 + [ctypes.c_int] * 3
    g(H.dptr(r), H.dptr(i), dim, bc, tc)
    assert np.array_equal(H.from_soa(r, i, dim, tc), np.broadcast_to(np.eye(dim), M.shape))


def test_load_packages_reads_reference_csv(tmp_path, monkeypatch):
    # utils.cpp:17-50: testcases/f_coeffs_<k>.csv, one polynomial per line,
    # re,im pairs of c_0..c_{n-1}; C[coeff + line*n + package*matrices*n]
    rng = np.random.default_rng(3)
    packages, matrices, n = 2, 3, 4
    want = np.zeros(packages * matrices * n, complex)
    (tmp_path / "testcases").mkdir()
    for k in range(packages):
        lines = []
        for row in range(matrices):
            c = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            want[k * matrices * n + row * n: k * matrices * n + (row + 1) * n] = c
            lines.append(",".join(f"{v:.17g}" for z in c for v in (z.real, z.imag)))
        (tmp_path / "testcases" / f"f_coeffs_{k + 1}.csv").write_text("\n".join(lines) + "\n")
    monkeypatch.chdir(tmp_path)
    out = np.zeros(packages * matrices * n, complex)
    f = sym("loadPackages")
    f.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p]
    f.restype = None
    f(packages, matrices, n, out.ctypes.data)
    assert np.array_equal(out, want)

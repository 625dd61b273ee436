"""Pin the CPU oracle by mathematics, independently of the reference build:
closed-form invariants of each stage, special cases with known answers,
numpy/LAPACK (zgeev, eigh), companion matrices with known roots, and planted
mistakes (the oracle's deviation switches) that these checks must catch.
"""
from __future__ import annotations

import numpy as np
import pytest

import helpers as H

EPS = np.finfo(float).eps


@pytest.fixture(scope="module")
def orc():
    return H.load_oracle()


def hessenberg(orc, A, tc=1, flags=0):
    N, n, _ = A.shape
    rA, iA = H.to_soa(A, tc)
    rQ = np.zeros(N * n * n); iQ = np.zeros(N * n * n)
    orc.oracle_hessenberg(n, N // tc, tc, H.dptr(rA), H.dptr(iA), H.dptr(rQ), H.dptr(iQ), flags)
    return H.from_soa(rA, iA, n, tc), H.from_soa(rQ, iQ, n, tc)


def lahqr(orc, Hm, Z, tc=1):
    N, n, _ = Hm.shape
    rH, iH = H.to_soa(Hm, tc)
    rZ, iZ = H.to_soa(Z, tc)
    rE = np.zeros(N * n); iE = np.zeros(N * n)
    orc.oracle_zlahqr(n, N // tc, tc, H.dptr(rE), H.dptr(iE), H.dptr(rH), H.dptr(iH), H.dptr(rZ),
                      H.dptr(iZ), 0, None)
    return H.vec_from_soa(rE, iE, n, tc), H.from_soa(rH, iH, n, tc), H.from_soa(rZ, iZ, n, tc)


def herm(X):
    return np.conj(np.transpose(X, (0, 2, 1)))


@pytest.mark.parametrize("flags", [0, H.LAPACK_SIGN])
@pytest.mark.parametrize("n", [2, 5, 9, 16, 27])
def test_hessenberg_invariants(orc, n, flags):
    A = H.gaussian(16, n, 100 + n)
    Hm, Q = hessenberg(orc, A, 4 if n % 2 else 1, flags) if False else hessenberg(orc, A, 1, flags)
    Hh = np.triu(Hm, -1)
    I = np.eye(n)
    assert np.max(np.abs(herm(Q) @ Q - I)) < 50 * n * EPS
    nrm = np.linalg.norm(A, axis=(1, 2))[:, None, None]
    assert np.max(np.abs(Q @ Hh @ herm(Q) - A) / nrm) < 50 * n * EPS
    # subdiagonal real (zlarfg returns a real beta)
    assert np.all(np.diagonal(Hh, -1, 1, 2).imag == 0)
    # planted mistake: keeping the reflector entries (the reference's
    # eigensolver feeds them to QR) breaks the similarity
    if n > 2:
        assert np.max(np.abs(Q @ Hm @ herm(Q) - A) / nrm) > 1e-3


@pytest.mark.parametrize("n", [1, 3, 8, 16, 24])
def test_lahqr_invariants(orc, n):
    A = H.gaussian(16, n, 200 + n)
    Hm, Q = hessenberg(orc, A)
    E, T, Z = lahqr(orc, np.triu(Hm, -1), Q)
    nrm = np.linalg.norm(A, axis=(1, 2))[:, None, None]
    assert np.max(np.abs(herm(Z) @ Z - np.eye(n))) < 100 * n * EPS
    assert np.max(np.abs(np.tril(T, -1))) == 0.0          # exactly triangular
    assert np.array_equal(E, np.diagonal(T, 0, 1, 2))
    assert np.max(np.abs(Z @ T @ herm(Z) - A) / nrm) < 100 * n * EPS


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32])
def test_eigensolver_invariants(orc, n):
    A = H.gaussian(32, n, 300 + n)
    E, V, _ = H.oracle_eig(A, 1, 0, orc)
    assert np.max(H.backward_error(A, E, V)) < 10 * n * EPS
    # the reference's scaling: V = Z X with X unit upper triangular, so
    # Z^H V is unit upper triangular (Toolbox.hpp:1272-1275)
    Hm, Q = hessenberg(orc, A)
    _, T, Z = lahqr(orc, np.triu(Hm, -1), Q)
    X = herm(Z) @ V
    assert np.max(np.abs(np.tril(X, -1))) < 1e-10 * max(1.0, np.abs(X).max())
    assert np.allclose(np.diagonal(X, 0, 1, 2), 1.0, atol=1e-10)
    # planted mistake: the reference's eigensolver() as shipped
    E2, V2, _ = H.oracle_eig(A, 1, H.KEEP_TRASH, orc)
    if n > 2:
        assert np.max(H.backward_error(A, E2, V2)) > 1e-6


@pytest.mark.parametrize("n", [3, 7, 16, 32])
def test_against_lapack_zgeev(orc, n):
    A = H.gaussian(32, n, 400 + n)
    E, V, _ = H.oracle_eig(A, 1, 0, orc)
    w = np.linalg.eigvals(A)
    _, err = H.match_eigenvalues(w, E)
    assert np.all(err < 1e-12 * np.linalg.norm(A, axis=(1, 2)))


def test_order_one_and_two(orc):
    rng = np.random.default_rng(5)
    A1 = (rng.standard_normal((8, 1, 1)) + 1j * rng.standard_normal((8, 1, 1)))
    E, V, _ = H.oracle_eig(A1, 1, 0, orc)
    assert np.array_equal(E[:, 0], A1[:, 0, 0]) and np.all(V == 1)
    A2 = rng.standard_normal((64, 2, 2)) + 1j * rng.standard_normal((64, 2, 2))
    E, V, _ = H.oracle_eig(A2, 1, 0, orc)
    tr = A2[:, 0, 0] + A2[:, 1, 1]
    det = A2[:, 0, 0] * A2[:, 1, 1] - A2[:, 0, 1] * A2[:, 1, 0]
    disc = np.sqrt(tr * tr - 4 * det)
    closed = np.stack([(tr + disc) / 2, (tr - disc) / 2], 1)
    _, err = H.match_eigenvalues(closed, E)
    assert err.max() < 1e-13


@pytest.mark.parametrize("n", [4, 12, 20])
def test_hermitian_real_spectrum(orc, n):
    B = H.gaussian(16, n, 500 + n)
    A = 0.5 * (B + herm(B))
    E, V, _ = H.oracle_eig(A, 1, 0, orc)
    w = np.linalg.eigvalsh(A)
    assert np.max(np.abs(np.sort(E.real, 1) - w)) < 1e-12 * np.abs(w).max()
    assert np.max(np.abs(E.imag)) < 1e-12 * np.abs(w).max()


@pytest.mark.parametrize("n", [3, 6, 10])
def test_companion_roots(orc, n):
    # the original use of the reference (polynomial roots via companion
    # matrices, Toolbox.hpp:607-643): eigenvalues are the prescribed roots
    rng = np.random.default_rng(600 + n)
    roots = rng.standard_normal((16, n)) + 1j * rng.standard_normal((16, n))
    C = np.zeros((16, n, n), complex)
    for k in range(16):
        c = np.poly(roots[k])                      # monic, highest power first
        C[k, 0, :] = -c[1:]
        C[k, 1:, :-1] = np.eye(n - 1)
    # companion matrices are already upper Hessenberg: the reference's
    # polynomial-root kernel runs QR on them directly (Toolbox.hpp:1894-1957)
    Z0 = np.broadcast_to(np.eye(n, dtype=complex), C.shape).copy()
    E, T, Z = lahqr(orc, C, Z0)
    _, err = H.match_eigenvalues(roots, E)
    assert err.max() < 1e-8 * np.abs(roots).max()
    # through the full chain the reference's zgehrd meets x = 0 columns
    # (1/0, Toolbox.hpp:1366); with LAPACK's zlarfg the chain gives the roots
    E, V, _ = H.oracle_eig(C, 1, 0, orc)
    assert not np.all(np.isfinite(E))
    E, V, _ = H.oracle_eig(C, 1, H.LAPACK_SIGN, orc)
    _, err = H.match_eigenvalues(roots, E)
    assert err.max() < 1e-8 * np.abs(roots).max()


def test_triangular_input_reference_sign(orc):
    # upper-triangular input: every column below the subdiagonal is zero.  The
    # reference's zgehrd has no ||x|| = 0 guard and the opposite sign
    # (Toolbox.hpp:1366-1373), so a real positive subdiagonal entry gives
    # 1/0; LAPACK's zlarfg (ORACLE_LAPACK_SIGN) returns the diagonal.
    rng = np.random.default_rng(7)
    A = np.triu(rng.standard_normal((4, 6, 6)) + 1j * rng.standard_normal((4, 6, 6)))
    A[:, 1, 0] = 1.0  # real positive alpha, x = 0
    E, V, _ = H.oracle_eig(A, 1, 0, orc)
    assert not np.all(np.isfinite(E)) or not np.all(np.isfinite(V))
    E, V, _ = H.oracle_eig(A, 1, H.LAPACK_SIGN, orc)
    assert np.all(np.isfinite(E)) and np.max(H.backward_error(A, E, V)) < 1e-14
    w = np.linalg.eigvals(A)
    _, err = H.match_eigenvalues(w, E)
    assert err.max() < 1e-13


def test_sign_convention_changes_h_not_spectrum(orc):
    # the two reflector signs give different H (fixtures pin the reference's)
    # but the same eigenpairs up to column phase
    A = H.gaussian(16, 10, 800)
    H1, _ = hessenberg(orc, A, 1, 0)
    H2, _ = hessenberg(orc, A, 1, H.LAPACK_SIGN)
    assert not np.allclose(np.triu(H1, -1), np.triu(H2, -1))
    E1, V1, _ = H.oracle_eig(A, 1, 0, orc)
    E2, V2, _ = H.oracle_eig(A, 1, H.LAPACK_SIGN, orc)
    perm, err = H.match_eigenvalues(E1, E2)
    assert err.max() < 1e-12 * np.linalg.norm(A, axis=(1, 2)).max()
    assert H.eigvec_distance(V2, V1, np.argsort(perm, axis=1)).max() < 1e-10


def test_sweep_statistics(orc):
    # measured sweeps per eigenvalue for Gaussian inputs (used to state the
    # actual flop count beside the nominal s = 2 model)
    A = H.gaussian(256, 16, 900)
    _, _, _, (info, sweeps, steps) = H.oracle_eig(A, 1, 0, orc, stats=True)
    assert np.all(info == 0)
    s = sweeps.sum() / (256 * 16)
    assert 2.0 < s < 5.0

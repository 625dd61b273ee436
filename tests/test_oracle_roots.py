"""The oracle's eigenvalues-only QR (zlahqr_, Toolbox.hpp:1663-1890) and
polynomial-roots path (create_companion_matrices + zlahqr kernel,
Toolbox.hpp:605-643, 1894-1957), pinned three ways:
  * bit for bit against the reference's own code built from its sources
    (oracle/_ref/libref_exact.so, FMA contraction off on both sides);
  * eigenvalues-only E equal, bit for bit, to the full Schur iteration's E on
    the same Hessenberg input (the active block evolves identically; only the
    updates outside it and the Schur vectors are skipped);
  * mathematics: companion structure (subdiagonal ones, last column
    -c_i/c_n, characteristic polynomial = c/c_n), roots equal to numpy's
    LAPACK roots, A Z = Z T with Z unitary and T upper triangular.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import helpers as H


@pytest.fixture(scope="module")
def orc():
    return H.load_oracle()


ref_lib = H.load_ref(exact=True)
needs_ref = pytest.mark.skipif(ref_lib is None or not hasattr(ref_lib, "ref_zlahqr_eig"),
                               reason="oracle/_ref not built (needs the reference tree)")


def coeffs(N, n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, n + 1)) + 1j * rng.standard_normal((N, n + 1))


@needs_ref
@pytest.mark.parametrize("n,tc", [(4, 1), (8, 4), (16, 2), (24, 1)])
def test_eig_only_equals_reference_build(orc, n, tc):
    Hm = H.hessenberg_input(8, n, 100 + n, tc, orc)
    with H.single_threaded(orc, ref_lib):
        Eo, Ho, _ = H.eig_only(Hm, tc, orc)
        Er, Hr, _ = H.eig_only(Hm, tc, ref_lib, ref=True)
    assert np.array_equal(Eo, Er)
    assert np.array_equal(Ho, Hr)


@needs_ref
@pytest.mark.parametrize("n,tc", [(3, 1), (8, 2), (12, 4), (20, 1)])
def test_roots_equal_reference_build(orc, n, tc):
    Cf = coeffs(8, n, 200 + n)
    with H.single_threaded(orc, ref_lib):
        Eo, To, Zo = H.roots_run(Cf, tc, orc)
        Er, Tr, Zr = H.roots_run(Cf, tc, ref_lib, ref=True)
    assert np.array_equal(Eo, Er) and np.array_equal(To, Tr) and np.array_equal(Zo, Zr)


@pytest.mark.parametrize("n,tc", [(5, 1), (16, 4), (32, 2)])
def test_eig_only_matches_full_schur_bitwise(orc, n, tc):
    Hm = H.hessenberg_input(8, n, 300 + n, tc, orc)
    E1, _, info = H.eig_only(Hm, tc, orc)
    rH, iH = H.to_soa(Hm, tc)
    rZ = np.zeros_like(rH); iZ = np.zeros_like(iH)
    N = Hm.shape[0]
    for k in range(N):  # Z = I
        b, t = divmod(k, tc)
        for j in range(n):
            rZ[b * tc * n * n + (j + j * n) * tc + t] = 1.0
    rE = np.zeros(N * n); iE = np.zeros(N * n)
    orc.oracle_zlahqr(n, N // tc, tc, H.dptr(rE), H.dptr(iE), H.dptr(rH), H.dptr(iH), H.dptr(rZ),
                      H.dptr(iZ), 0, None)
    E2 = H.vec_from_soa(rE, iE, n, tc)
    assert np.all(info == 0)
    assert np.array_equal(E1, E2)
    # and they are the eigenvalues of H (LAPACK)
    _, err = H.match_eigenvalues(np.linalg.eigvals(Hm), E1)
    assert err.max() < 1e-11 * np.abs(Hm).max() * n


@pytest.mark.parametrize("n", [1, 2, 6, 16])
def test_companion_structure(orc, n):
    N, tc = 6, 2
    Cf = coeffs(N, n, 500 + n)
    rC, iC = H.vec_to_soa(Cf, tc)
    rM = np.zeros(N * n * n); iM = np.zeros(N * n * n)
    orc.oracle_companion(n, N // tc, tc, H.dptr(rC), H.dptr(iC), H.dptr(rM), H.dptr(iM))
    M = H.from_soa(rM, iM, n, tc)
    lead = Cf[:, n]
    assert np.allclose(M[:, :, n - 1], -Cf[:, :n] / lead[:, None], rtol=1e-15, atol=0)
    sub = np.eye(n, k=-1)
    assert np.array_equal(M[:, :, : n - 1], np.broadcast_to(sub[:, : n - 1], (N, n, n - 1)))
    for k in range(N):  # characteristic polynomial = monic p
        want = (Cf[k, ::-1] / lead[k])
        assert np.allclose(np.poly(M[k]), want, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("n", [2, 7, 16, 30])
def test_roots_against_numpy(orc, n):
    Cf = coeffs(16, n, 600 + n)
    E, T, Z = H.roots_run(Cf, 4, orc)
    for k in range(16):
        want = np.roots(Cf[k, ::-1])
        _, err = H.match_eigenvalues(want[None], E[k][None])
        assert err[0] < 1e-8 * max(1.0, np.abs(want).max())
    # Schur form of the companion matrix
    rC, iC = H.vec_to_soa(Cf, 4)
    rM = np.zeros(16 * n * n); iM = np.zeros(16 * n * n)
    orc.oracle_companion(n, 4, 4, H.dptr(rC), H.dptr(iC), H.dptr(rM), H.dptr(iM))
    M = H.from_soa(rM, iM, n, 4)
    eye = np.eye(n)
    assert np.all(np.tril(T, -1) == 0)
    assert np.allclose(np.diagonal(T, 0, 1, 2), E, rtol=0, atol=0)
    assert np.max(np.abs(np.conj(np.transpose(Z, (0, 2, 1))) @ Z - eye)) < 1e-13 * n
    nrm = np.linalg.norm(M, axis=(1, 2))
    assert np.max(np.linalg.norm(M @ Z - Z @ T, axis=(1, 2)) / nrm) < 1e-13 * n

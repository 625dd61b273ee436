"""GPU parity: the fused gfx950 eigensolver (veig_zgeev) against the CPU
oracle (the reference's zgehrd -> zunghr -> zlahqr -> ztrevc chain with the
reflector entries cleared before QR, oracle/eig_oracle.c) on the same seeded
inputs.

Tolerances (FP64, tolerance-based parity; the path is floating point):
  * eigenvalues, matched by assignment:  |dE| <= 1e-11 * ||A||_F
  * eigenvectors, unit-norm and phase normalised, per column:
        dist <= 1e-12 * ||A||_F / gap_k + 1e-11      (gap_k = distance from E_k
        to the nearest other eigenvalue; first-order perturbation bound)
  * backward error ||AV - VE||_F / (||A||_F ||V||_F) <= 64 * n * eps
  * where the Schur (deflation) order agrees with the oracle's, the
    unnormalised column norms of V agree to 1e-8 relative (the reference's
    scaling, column k = Z [x; 1; 0]).
"""
from __future__ import annotations

import numpy as np
import pytest

from helpers import (backward_error, eigvec_distance, gaussian, load_oracle, match_eigenvalues,
                     oracle_eig)

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def hermitian(N, n, seed):
    B = gaussian(N, n, seed)
    return 0.5 * (B + np.conj(np.transpose(B, (0, 2, 1))))


def constructed(N, n, seed, cond=1e2, cluster=None):
    """A = V diag(lam) V^{-1}, V = U diag(sigma) W^H with Haar U, W and sigma
    log-spaced over [1, 1/cond]; lam iid CN(0,1) or in clusters spaced `cluster`."""
    rng = np.random.default_rng(seed)

    def haar(k):
        Z = rng.standard_normal((N, k, k)) + 1j * rng.standard_normal((N, k, k))
        Q, R = np.linalg.qr(Z)
        d = np.diagonal(R, axis1=1, axis2=2)
        return Q * (d / np.abs(d))[:, None, :]

    U, W = haar(n), haar(n)
    sig = np.logspace(0, -np.log10(cond), n)
    V = (U * sig[None, None, :]) @ np.conj(np.transpose(W, (0, 2, 1)))
    if cluster is None:
        lam = (rng.standard_normal((N, n)) + 1j * rng.standard_normal((N, n))) / np.sqrt(2)
    else:
        ncl = max(1, n // 3)
        centers = (rng.standard_normal((N, ncl)) + 1j * rng.standard_normal((N, ncl))) / np.sqrt(2)
        idx = np.arange(n) % ncl
        lam = centers[:, idx] + cluster * (np.arange(n) // ncl)[None, :]
    return V @ (lam[:, :, None] * np.linalg.inv(V))


def gaps(E):
    D = np.abs(E[:, :, None] - E[:, None, :])
    D[:, np.arange(E.shape[1]), np.arange(E.shape[1])] = np.inf
    return D.min(axis=2)


def check_parity(A, E, V, info, Eo, Vo, vec_check=True):
    N, n, _ = A.shape
    nrmA = np.linalg.norm(A, axis=(1, 2))
    assert np.all(np.isfinite(E)) and np.all(np.isfinite(V)), "non-finite output"
    assert np.all(info == 0), f"unconverged eigenvalues reported: {np.unique(info)}"
    perm, eerr = match_eigenvalues(Eo, E)
    assert np.all(eerr <= 1e-11 * np.maximum(nrmA, 1e-300)), f"eigenvalue mismatch {np.max(eerr / nrmA):.2e}"
    be = backward_error(A, E, V)
    assert np.all(be <= 64 * n * EPS), f"backward error {be.max():.2e}"
    if vec_check and n > 1:
        # perm maps oracle column i to GPU column perm[i]; compare GPU column
        # j with the oracle column that matched it
        inv = np.argsort(perm, axis=1)
        W2 = np.take_along_axis(Vo, inv[:, None, :], axis=2)
        from helpers import phase_normalise
        P1, P2 = phase_normalise(V), phase_normalise(W2)
        ip = np.sum(np.conj(P2) * P1, axis=1, keepdims=True)
        ph = ip / np.where(np.abs(ip) == 0, 1, np.abs(ip))
        dist = np.linalg.norm(P1 - P2 * ph, axis=1)           # (N, n)
        tol = 1e-12 * nrmA[:, None] / gaps(E) + 1e-11
        assert np.all(dist <= tol), f"eigenvector mismatch, worst ratio {np.max(dist / tol):.2e}"
        # reference scaling where the Schur order agrees
        same = np.all(np.abs(E - Eo) <= 1e-9 * nrmA[:, None], axis=1)
        if same.any():
            n1 = np.linalg.norm(V[same], axis=1)
            n2 = np.linalg.norm(Vo[same], axis=1)
            assert np.allclose(n1, n2, rtol=1e-8), "unnormalised eigenvector scaling differs"
    return be


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 9, 12, 15, 16, 17, 24, 31, 32, 33, 40, 48, 63, 64])
def test_gaussian_parity(gpu, n):
    from gpu_util import gpu_eig
    A = gaussian(64, n, 1000 + n)
    Eo, Vo, _ = oracle_eig(A, 1)
    E, V, info = gpu_eig(A, 1)
    check_parity(A, E, V, info, Eo, Vo)


@pytest.mark.parametrize("thread_cnt", [1, 2, 4, 32, 64])
def test_tile_layouts(gpu, thread_cnt):
    from gpu_util import gpu_eig
    A = gaussian(128, 16, 7)
    Eo, Vo, _ = oracle_eig(A, thread_cnt)
    E, V, info = gpu_eig(A, thread_cnt)
    check_parity(A, E, V, info, Eo, Vo)


@pytest.mark.parametrize("n", [4, 8, 16, 24, 32, 64])
def test_hermitian_parity(gpu, n):
    from gpu_util import gpu_eig
    A = hermitian(64, n, 2000 + n)
    Eo, Vo, _ = oracle_eig(A, 1)
    E, V, info = gpu_eig(A, 1)
    be = check_parity(A, E, V, info, Eo, Vo, vec_check=False)
    # eigenvalues real to working accuracy and equal to eigh's
    w = np.sort(np.linalg.eigvalsh(A), axis=1)
    assert np.max(np.abs(np.sort(E.real, axis=1) - w)) <= 1e-11 * np.linalg.norm(A, axis=(1, 2)).max()
    assert np.max(np.abs(E.imag)) <= 1e-12 * np.linalg.norm(A, axis=(1, 2)).max()


@pytest.mark.parametrize("n", [8, 16, 24, 64])
def test_constructed_parity(gpu, n):
    from gpu_util import gpu_eig
    A = constructed(64, n, 3000 + n, cond=1e2)
    Eo, Vo, _ = oracle_eig(A, 1)
    E, V, info = gpu_eig(A, 1)
    check_parity(A, E, V, info, Eo, Vo)


@pytest.mark.parametrize("n", [8, 16, 32])
def test_clustered_backward_error(gpu, n):
    # clustered eigenvalues (1e-6 apart): vectors are ill-conditioned, so
    # parity is judged by eigenvalues and backward error only
    from gpu_util import gpu_eig
    A = constructed(64, n, 4000 + n, cond=1e2, cluster=1e-6)
    Eo, Vo, _ = oracle_eig(A, 1)
    E, V, info = gpu_eig(A, 1)
    check_parity(A, E, V, info, Eo, Vo, vec_check=False)


@pytest.mark.parametrize("thread_cnt", [1, 4])
def test_order_64_tiles(gpu, thread_cnt):
    # one matrix per wavefront, eigenvector solves in LDS
    from gpu_util import gpu_eig
    A = gaussian(32, 64, 77)
    Eo, Vo, _ = oracle_eig(A, thread_cnt)
    E, V, info = gpu_eig(A, thread_cnt)
    check_parity(A, E, V, info, Eo, Vo)


def unit_upper_eigvecs(T):
    """X with column k = [x; 1; 0], (T - T_kk I) X(:, k) = 0 (numpy solve)."""
    N, n, _ = T.shape
    X = np.zeros_like(T)
    for b in range(N):
        for k in range(n):
            X[b, k, k] = 1
            if k:
                M = T[b, :k, :k] - T[b, k, k] * np.eye(k)
                X[b, :k, k] = np.linalg.solve(M, -T[b, :k, k])
    return X


@pytest.mark.parametrize("n", [16, 40, 64])
def test_schur_form_written(gpu, n):
    # VEIG_WRITE_SCHUR: the input planes receive T; Z = V X^{-1} must be
    # unitary with A Z = Z T, T upper triangular with diag(T) = E
    from gpu_util import gpu_eig
    A = gaussian(8, n, 5000 + n)
    E, V, info, T = gpu_eig(A, 4, write_schur=True)
    assert np.all(info == 0)
    assert np.all(np.tril(T, -1) == 0)
    assert np.array_equal(np.diagonal(T, 0, 1, 2), E)
    Z = V @ np.linalg.inv(unit_upper_eigvecs(T))
    eye = np.eye(n)
    nrm = np.linalg.norm(A, axis=(1, 2))
    assert np.max(np.linalg.norm(np.conj(np.transpose(Z, (0, 2, 1))) @ Z - eye, axis=(1, 2))) < 1e-12 * n
    assert np.max(np.linalg.norm(A @ Z - Z @ T, axis=(1, 2)) / nrm) < 1e-13 * n

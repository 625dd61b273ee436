"""GPU parity of the two-kernel FP64 paths (Hessenberg reduction in compact
form, then the QR iteration and eigenvectors on a banded tile): orders 9..16
(four matrices per wavefront), 17..24 and 25..28 (slots of 24 and 28 on 32
lanes), 29..32 (two per wavefront) and 33..64 (one matrix per workgroup).

Cases the split makes special: batches that leave lane groups or tiles
partly empty (the reduction and the QR kernel must agree on which matrices
exist), the compact form's reflectors with tau = 0 (columns already zero
below the subdiagonal), the Schur form written back to the input planes from
the banded tile, and the iteration cap.  Each result is checked against the
CPU oracle (eigenvalues, eigenvectors up to phase, the reference's scaling)
and by its backward error.
"""
from __future__ import annotations

import numpy as np
import pytest

import helpers as H
from gpu_util import gpu_eig
from helpers import oracle_eig
from test_gpu_parity import check_parity, unit_upper_eigvecs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,N,T", [(9, 5, 1), (13, 7, 1), (16, 13, 1), (11, 12, 4), (16, 6, 2),
                                   (10, 3, 3), (17, 3, 1), (20, 5, 1), (24, 6, 2), (25, 5, 1),
                                   (28, 4, 2), (29, 3, 1), (32, 3, 3), (37, 3, 1), (64, 2, 2)])
def test_ragged_batches(gpu, n, N, T):
    A = H.gaussian(N, n, 4000 + 10 * n + N)
    Eo, Vo, _ = oracle_eig(A, 1)
    E, V, info = gpu_eig(A, T)
    check_parity(A, E, V, info, Eo, Vo)


@pytest.mark.parametrize("n", [9, 12, 16, 20, 28, 36])
def test_reflectors_with_zero_tau(gpu, n):
    # upper Hessenberg input with some zero subdiagonal entries and a
    # triangular input: every or some reflectors of the reduction are the
    # identity (tau = 0), which the QR kernel must rebuild Q from unchanged
    Hs = np.triu(H.gaussian(8, n, 5000 + n), -1)
    Hs[:, 3, 2] = 0.0
    Hs[::2, n - 1, n - 2] = 0.0
    # (the reference's zgehrd divides by the zero norm of such columns, so
    # the eigenvalues are checked against LAPACK's and, for the triangular
    # input, against its diagonal)
    Tr = np.triu(H.gaussian(8, n, 5100 + n))
    for A, want in ((Hs, np.linalg.eigvals(Hs)), (Tr, np.diagonal(Tr, 0, 1, 2))):
        E, V, info = gpu_eig(A, 4)
        assert np.all(info == 0) and np.all(np.isfinite(V))
        nrmA = np.linalg.norm(A, axis=(1, 2))
        _, err = H.match_eigenvalues(want, E)
        assert np.all(err <= 1e-9 * nrmA)
        assert np.max(H.backward_error(A, E, V)) < 64 * n * np.finfo(float).eps


@pytest.mark.parametrize("n,N,T", [(9, 5, 1), (14, 8, 4), (22, 3, 1), (30, 4, 2), (33, 3, 1)])
def test_schur_form_from_banded_tile(gpu, n, N, T):
    # VEIG_WRITE_SCHUR on the two-kernel path: T (from the banded tile) into
    # the input planes, exact zeros below the diagonal, diag(T) = E, and
    # Z = V X^{-1} unitary with A Z = Z T
    A = H.gaussian(N, n, 5200 + n)
    E, V, info, Tm = gpu_eig(A, T, write_schur=True)
    assert np.all(info == 0)
    assert np.all(np.tril(Tm, -1) == 0.0)
    assert np.array_equal(np.diagonal(Tm, 0, 1, 2), E)
    Z = V @ np.linalg.inv(unit_upper_eigvecs(Tm))
    nrmA = np.linalg.norm(A, axis=(1, 2))
    eye = np.eye(n)
    assert np.max(np.linalg.norm(np.conj(np.transpose(Z, (0, 2, 1))) @ Z - eye, axis=(1, 2))) < 1e-12 * n
    assert np.max(np.linalg.norm(A @ Z - Z @ Tm, axis=(1, 2)) / nrmA) < 1e-13 * n


@pytest.mark.parametrize("n", [12, 40])
def test_iteration_cap_banded(gpu, n):
    A = H.gaussian(32, n, 5300 + n)
    _, _, info = gpu_eig(A, 4 if n <= 16 else 1, itmax=1)
    assert np.any(info > 0)
    E, V, info = gpu_eig(A, 4 if n <= 16 else 1, itmax=30)
    assert np.all(info == 0) and np.max(H.backward_error(A, E, V)) < 64 * n * np.finfo(float).eps

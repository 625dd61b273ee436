"""GPU parity where the QR iteration leaves its usual path: the exceptional
shift of iteration 10 (zlahqr_schur, Toolbox.hpp:896-910), the Wilkinson
shift's u^2 = 0 case and the deflation test's zero diagonal pair
(Toolbox.hpp:829-843: the stalled sweeps keep the diagonal at zero), in
wavefronts whose other matrices take the usual path in the same sweeps.

The cyclic permutation matrix is unitary and already upper Hessenberg with a
zero diagonal: H(I-1, I) = 0 makes u^2 = 0, the shift is H(I, I) = 0, and a
QR step with a zero shift maps the matrix to itself, so the first eigenvalue
stalls until the exceptional shift at iteration 10 (the oracle takes 16 to
17 sweeps for it at n = 12..24, against 2 to 3 per eigenvalue otherwise).
Its eigenvalues are the n-th roots of unity, well separated, so eigenvalues
and eigenvectors are compared with the oracle as in the parity tests.

The oracle runs with the LAPACK sign and deflation-test conventions that the
HIP path follows: the reference's own conventions, taken literally, do not
converge on this matrix (NaN eigenvalues and info = n - 1 from the oracle
with flags 0, the reference defect in its deflation test's K+1 term).
"""
from __future__ import annotations

import numpy as np
import pytest

import helpers as H
from gpu_util import gpu_eig
from test_gpu_parity import check_parity

pytestmark = pytest.mark.gpu


def cyclic(n: int, scale: float = 1.0) -> np.ndarray:
    P = np.zeros((n, n), complex)
    P[np.arange(1, n), np.arange(n - 1)] = scale
    P[0, n - 1] = scale
    return P


@pytest.mark.parametrize("n,T", [(12, 4), (16, 4), (16, 1), (24, 2), (40, 1)])
def test_exceptional_shift_beside_usual_sweeps(gpu, n, T):
    # matrices 1 and 6 stall; the others share their wavefronts (four order-16
    # matrices or two order-24 matrices per wavefront) and their sweeps
    A = H.gaussian(8, n, 7100 + n)
    A[1] = cyclic(n)
    A[6] = cyclic(n, 2.0)
    Eo, Vo, _, (info_o, sweeps, _) = H.oracle_eig(A, 1, H.LAPACK_SIGN | H.LAPACK_TST, stats=True)
    assert np.all(info_o == 0)
    assert sweeps[1] >= 11 + 2 * (n - 2), "the cyclic matrix must need the exceptional shift"
    E, V, info = gpu_eig(A, T)
    check_parity(A, E, V, info, Eo, Vo)
    for m, r in ((1, 1.0), (6, 2.0)):
        roots = np.sort(np.angle(E[m] / r))
        assert np.allclose(np.abs(E[m]), r, atol=1e-13 * r)
        assert np.allclose(np.diff(roots), 2 * np.pi / n, atol=1e-12)


@pytest.mark.parametrize("n", [12, 16])
def test_exceptional_shift_mixed_precision(gpu, n):
    A = H.gaussian(8, n, 7200 + n)
    A[2] = cyclic(n)
    E, V, info, berr = gpu_eig(A, 4, precision="mixed")
    assert np.all(info == 0)
    assert np.max(berr) <= 1e-12
    assert np.max(H.backward_error(A, E, V)) <= 1e-12
    assert np.allclose(np.abs(E[2]), 1.0, atol=1e-12)

"""GPU edge cases and full-size properties of the fused path.

Edge cases the reference mishandles or the layout makes special: empty
batches, n = 1 and 2, identity / diagonal / triangular inputs (zero columns
below the subdiagonal, where the reference's zgehrd divides by zero), NaN
input (must terminate; reported through info or NaN output), an iteration cap
of one sweep (info > 0 reported instead of the reference's silent accept).
At the headline size (1,048,576 matrices of order 16): size-independent
properties -- every matrix converged, every output finite, sum of eigenvalues
= trace for every matrix, backward error on a sample.
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import helpers as H
import veig
from gpu_util import gpu_eig

pytestmark = pytest.mark.gpu


def test_empty_batches(gpu):
    dev = torch.device("cuda:0")
    z = torch.empty(0, dtype=torch.float64, device=dev)
    veig.zgeev(z, z, 16, 0, 64, out=(z, z, z, z))
    veig.zgeev(z, z, 0, 4, 4, out=(z, z, z, z))


def test_order_one_and_two(gpu):
    A1 = H.gaussian(64, 1, 1)
    E, V, info = gpu_eig(A1, 4)
    assert np.array_equal(E[:, 0], A1[:, 0, 0]) and np.all(V == 1) and np.all(info == 0)
    A2 = H.gaussian(64, 2, 2)
    E, V, info = gpu_eig(A2, 4)
    tr = A2[:, 0, 0] + A2[:, 1, 1]
    det = A2[:, 0, 0] * A2[:, 1, 1] - A2[:, 0, 1] * A2[:, 1, 0]
    disc = np.sqrt(tr * tr - 4 * det)
    _, err = H.match_eigenvalues(np.stack([(tr + disc) / 2, (tr - disc) / 2], 1), E)
    assert err.max() < 1e-13 and np.max(H.backward_error(A2, E, V)) < 1e-14


@pytest.mark.parametrize("n", [5, 16, 32, 48, 64])
def test_structured_inputs(gpu, n):
    rng = np.random.default_rng(n)
    eye = np.broadcast_to(np.eye(n, dtype=complex), (8, n, n)).copy()
    diag = np.zeros((8, n, n), complex)
    d = rng.standard_normal((8, n)) + 1j * rng.standard_normal((8, n))
    diag[:, np.arange(n), np.arange(n)] = d
    tri = np.triu(H.gaussian(8, n, n + 100))
    for A, want in ((eye, np.ones((8, n))), (diag, d), (tri, np.diagonal(tri, 0, 1, 2))):
        E, V, info = gpu_eig(A, 4)
        assert np.all(np.isfinite(E)) and np.all(np.isfinite(V)) and np.all(info == 0)
        _, err = H.match_eigenvalues(want, E)
        assert err.max() < 1e-12 * max(1.0, np.abs(want).max())
        assert np.max(H.backward_error(A, E, V)) < 1e-13


def test_companion_matrices(gpu):
    # polynomial roots, the reference's original use (Toolbox.hpp:607-643):
    # Hessenberg input whose columns are zero below the subdiagonal
    rng = np.random.default_rng(9)
    n = 12
    roots = rng.standard_normal((16, n)) + 1j * rng.standard_normal((16, n))
    C = np.zeros((16, n, n), complex)
    for k in range(16):
        C[k, 0, :] = -np.poly(roots[k])[1:]
        C[k, 1:, :-1] = np.eye(n - 1)
    E, V, info = gpu_eig(C, 4)
    _, err = H.match_eigenvalues(roots, E)
    assert err.max() < 1e-8 * np.abs(roots).max()


def test_nan_input_terminates(gpu):
    A = H.gaussian(8, 16, 3)
    A[2, 5, 7] = np.nan
    E, V, info = gpu_eig(A, 4)
    ok = [k for k in range(8) if k != 2]
    assert np.all(info[ok] == 0) and np.max(H.backward_error(A[ok], E[ok], V[ok])) < 1e-13
    assert info[2] > 0 or not np.all(np.isfinite(E[2]))


def test_iteration_cap_reported(gpu):
    A = H.gaussian(128, 16, 4)
    E, V, info = gpu_eig(A, 4, itmax=1)
    assert np.any(info > 0), "a 2-sweep cap must leave some eigenvalues unconverged"
    E, V, info = gpu_eig(A, 4, itmax=30)
    assert np.all(info == 0)


def test_headline_size_properties(gpu):
    n, N, T = 16, 1 << 20, 64
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(2024)
    rA = torch.randn(N * n * n, dtype=torch.float64, device=dev, generator=g)
    iA = torch.randn(N * n * n, dtype=torch.float64, device=dev, generator=g)
    info = torch.full((N,), -1, dtype=torch.int32, device=dev)
    rE, iE, rV, iV = veig.zgeev(rA, iA, n, N // T, T, info=info)
    torch.cuda.synchronize()
    assert int((info != 0).sum()) == 0
    assert bool(torch.isfinite(rV).all() and torch.isfinite(iV).all() and torch.isfinite(rE).all())
    # trace = sum of eigenvalues, every matrix
    bc = N // T
    Ar = rA.view(bc, n, n, T)          # [b, c, r, t]
    Ai = iA.view(bc, n, n, T)
    idx = torch.arange(n, device=dev)
    trr = Ar[:, idx, idx, :].sum(1).reshape(-1)
    tri = Ai[:, idx, idx, :].sum(1).reshape(-1)
    er = rE.view(bc, n, T).sum(1).reshape(-1)
    ei = iE.view(bc, n, T).sum(1).reshape(-1)
    nrm = torch.sqrt((Ar ** 2 + Ai ** 2).sum((1, 2))).reshape(-1)
    dev_tr = torch.sqrt((trr - er) ** 2 + (tri - ei) ** 2) / nrm
    assert float(dev_tr.max()) < 1e-13
    # backward error on a spread sample of 2048 matrices
    pick = torch.randint(0, bc, (32,), device=dev, generator=g)
    A = torch.complex(Ar[pick], Ai[pick]).permute(0, 3, 2, 1).reshape(-1, n, n)
    V = torch.complex(rV.view(bc, n, n, T)[pick], iV.view(bc, n, n, T)[pick]).permute(0, 3, 2, 1).reshape(-1, n, n)
    E = torch.complex(rE.view(bc, n, T)[pick], iE.view(bc, n, T)[pick]).permute(0, 2, 1).reshape(-1, n)
    R = A @ V - V * E[:, None, :]
    be = torch.linalg.norm(R, dim=(1, 2)) / (torch.linalg.norm(A, dim=(1, 2)) * torch.linalg.norm(V, dim=(1, 2)))
    assert float(be.max()) < 1e-13


@pytest.mark.parametrize("n", [40, 64])
def test_nan_input_terminates_large_orders(gpu, n):
    # the two-wavefront kernel: a NaN stalls only its own matrix, which the
    # iteration cap ends (info > 0 or non-finite output); the others are exact
    A = H.gaussian(4, n, 30 + n)
    A[1, 3, 5] = np.nan
    E, V, info = gpu_eig(A, 1)
    ok = [0, 2, 3]
    assert np.all(info[ok] == 0) and np.max(H.backward_error(A[ok], E[ok], V[ok])) < 1e-13
    assert info[1] > 0 or not np.all(np.isfinite(E[1]))


def test_iteration_cap_reported_order_64(gpu):
    A = H.gaussian(8, 64, 44)
    _, _, info = gpu_eig(A, 1, itmax=1)
    assert np.any(info > 0)
    _, _, info = gpu_eig(A, 1, itmax=30)
    assert np.all(info == 0)


@pytest.mark.parametrize("n", [16, 48])
def test_mixed_nan_input(gpu, n):
    # complex64 QR fails on the NaN matrix -> FP64 fallback, which reports it
    A = H.gaussian(8, n, 50 + n)
    A[6, 0, 1] = np.nan
    rA, iA = H.to_soa(A, 1)
    dev = torch.device("cuda:0")
    info = torch.full((8,), -7, dtype=torch.int32, device=dev)
    fb = torch.full((8,), -7, dtype=torch.int32, device=dev)
    be = torch.empty(8, dtype=torch.float64, device=dev)
    rE, iE, rV, iV = veig.zgeev_mixed(torch.from_numpy(rA).to(dev), torch.from_numpy(iA).to(dev), n, 8, 1,
                                      info=info, fallback=fb, berr=be)
    torch.cuda.synchronize()
    info, fb, be = info.cpu().numpy(), fb.cpu().numpy(), be.cpu().numpy()
    E = H.vec_from_soa(rE.cpu().numpy(), iE.cpu().numpy(), n, 1)
    ok = [k for k in range(8) if k != 6]
    assert np.all(info[ok] == 0) and np.all(be[ok] <= 1e-12)
    assert fb[6] == 1 and (info[6] > 0 or not np.all(np.isfinite(E[6])))


def test_roots_zero_leading_coefficient(gpu):
    # c_n = 0: the companion matrix holds inf/NaN; its iteration ends at the
    # cap and the other polynomials are unaffected
    rng = np.random.default_rng(61)
    n, N = 8, 4
    Cf = rng.standard_normal((N, n + 1)) + 1j * rng.standard_normal((N, n + 1))
    Cf[2, n] = 0
    rC, iC = H.vec_to_soa(Cf, 1)
    dev = torch.device("cuda:0")
    info = torch.full((N,), -7, dtype=torch.int32, device=dev)
    rE, iE = veig.roots(torch.from_numpy(rC).to(dev), torch.from_numpy(iC).to(dev), n, N, 1, info=info)
    torch.cuda.synchronize()
    E = H.vec_from_soa(rE.cpu().numpy(), iE.cpu().numpy(), n, 1)
    info = info.cpu().numpy()
    for k in (0, 1, 3):
        want = np.roots(Cf[k, ::-1])
        _, err = H.match_eigenvalues(want[None], E[k][None])
        assert info[k] == 0 and err[0] < 1e-9 * max(1.0, np.abs(want).max())
    assert info[2] > 0 or not np.all(np.isfinite(E[2]))

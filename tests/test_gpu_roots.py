"""GPU tests of the eigenvalues-only and polynomial-roots entries
(veig_eigvals, veig_zlahqr_eig, veig_zlahqr_roots), against the CPU oracle's
restatement of the reference's zlahqr_ (Toolbox.hpp:1663) and zlahqr kernel
(Toolbox.hpp:1894), which the CPU suite pins bit for bit to the reference's
own build.

Bars: eigenvalues-only E equal to the fused eigensolver's E on the same
input to 1e-12 ||A||_F (the active block evolves identically; the kernels are
compiled separately and round differently, ~1e-14); E within
1e-12 * ||H||_F of the oracle's (matched by assignment); roots within
1e-9 * max|root| of numpy's; Schur outputs with T upper triangular,
diag(T) = E, Z unitary to 1e-13 n and ||C Z - Z T|| <= 1e-13 n ||C||.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import helpers as H
import veig

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to(DEV)


def coeffs(N, n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, n + 1)) + 1j * rng.standard_normal((N, n + 1))


@pytest.mark.parametrize("n", [3, 8, 16, 24, 40, 64])
def test_eigvals_equal_fused_eigenvalues(gpu, n):
    tc = 4
    A = H.gaussian(32, n, 7000 + n)
    rA, iA = H.to_soa(A, tc)
    N = A.shape[0]
    info = torch.full((N,), -1, dtype=torch.int32, device=DEV)
    rE, iE = veig.eigvals(_dev(rA), _dev(iA), n, N // tc, tc, info=info)
    rE2, iE2, _, _ = veig.zgeev(_dev(rA), _dev(iA), n, N // tc, tc)
    torch.cuda.synchronize()
    assert int((info != 0).sum()) == 0
    # the same iteration with and without Schur vectors (the active block
    # evolves identically; separately compiled kernels may round differently)
    E = H.vec_from_soa(rE.cpu().numpy(), iE.cpu().numpy(), n, tc)
    E2 = H.vec_from_soa(rE2.cpu().numpy(), iE2.cpu().numpy(), n, tc)
    _, err = H.match_eigenvalues(E2, E)
    assert np.all(err <= 1e-12 * np.linalg.norm(A, axis=(1, 2)))


@pytest.mark.parametrize("n,tc", [(4, 1), (16, 4), (32, 2), (64, 1)])
def test_zlahqr_eig_against_oracle(gpu, n, tc):
    Hm = H.hessenberg_input(16, n, 7100 + n, tc)
    Eo, _, info_o = H.eig_only(Hm, tc)
    rH, iH = H.to_soa(Hm, tc)
    N = Hm.shape[0]
    tH, tI = _dev(rH), _dev(iH)
    info = torch.full((N,), -1, dtype=torch.int32, device=DEV)
    rE, iE = veig.zlahqr_eig(tH, tI, n, N // tc, tc, info=info)
    torch.cuda.synchronize()
    E = H.vec_from_soa(rE.cpu().numpy(), iE.cpu().numpy(), n, tc)
    nrm = np.linalg.norm(Hm, axis=(1, 2))
    _, err = H.match_eigenvalues(Eo, E)
    assert np.all(info.cpu().numpy() == info_o)
    assert np.all(err <= 1e-12 * nrm)
    # H was updated in place inside the active blocks: its diagonal is E
    Hout = H.from_soa(tH.cpu().numpy(), tI.cpu().numpy(), n, tc)
    assert np.array_equal(np.diagonal(Hout, 0, 1, 2), E)


@pytest.mark.parametrize("n,tc", [(1, 1), (2, 4), (7, 2), (16, 4), (30, 1), (48, 2)])
def test_roots_against_oracle_and_numpy(gpu, n, tc):
    Cf = coeffs(16, n, 7200 + n)
    Eo, _, _ = H.roots_run(Cf, tc)
    rC, iC = H.vec_to_soa(Cf, tc)
    N = Cf.shape[0]
    info = torch.full((N,), -1, dtype=torch.int32, device=DEV)
    rE, iE = veig.roots(_dev(rC), _dev(iC), n, N // tc, tc, info=info)
    torch.cuda.synchronize()
    E = H.vec_from_soa(rE.cpu().numpy(), iE.cpu().numpy(), n, tc)
    assert np.all(info.cpu().numpy() == 0)
    _, err = H.match_eigenvalues(Eo, E)
    scale = np.maximum(1.0, np.abs(Eo).max(axis=1))
    assert np.all(err <= 1e-10 * scale)
    for k in range(N):
        want = np.roots(Cf[k, ::-1])
        _, e2 = H.match_eigenvalues(want[None], E[k][None])
        assert e2[0] <= 1e-9 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("n", [5, 16, 24])
def test_roots_schur_outputs(gpu, n):
    tc = 4
    Cf = coeffs(16, n, 7300 + n)
    rC, iC = H.vec_to_soa(Cf, tc)
    N = Cf.shape[0]
    rE, iE, rT, iT, rZ, iZ = veig.roots(_dev(rC), _dev(iC), n, N // tc, tc, schur=True)
    torch.cuda.synchronize()
    E = H.vec_from_soa(rE.cpu().numpy(), iE.cpu().numpy(), n, tc)
    T = H.from_soa(rT.cpu().numpy(), iT.cpu().numpy(), n, tc)
    Z = H.from_soa(rZ.cpu().numpy(), iZ.cpu().numpy(), n, tc)
    orc = H.load_oracle()
    rM = np.zeros(N * n * n); iM = np.zeros(N * n * n)
    orc.oracle_companion(n, N // tc, tc, H.dptr(rC), H.dptr(iC), H.dptr(rM), H.dptr(iM))
    C = H.from_soa(rM, iM, n, tc)
    assert np.all(np.tril(T, -1) == 0)
    assert np.array_equal(np.diagonal(T, 0, 1, 2), E)
    eye = np.eye(n)
    assert np.max(np.abs(np.conj(np.transpose(Z, (0, 2, 1))) @ Z - eye)) < 1e-13 * n
    nrm = np.linalg.norm(C, axis=(1, 2))
    assert np.max(np.linalg.norm(C @ Z - Z @ T, axis=(1, 2)) / nrm) < 1e-13 * n
    # the roots-only launch gives the same roots (to rounding: a separately
    # compiled iteration on the same active blocks)
    rE2, iE2 = veig.roots(_dev(rC), _dev(iC), n, N // tc, tc)
    torch.cuda.synchronize()
    E2 = H.vec_from_soa(rE2.cpu().numpy(), iE2.cpu().numpy(), n, tc)
    _, err = H.match_eigenvalues(E, E2)
    assert np.all(err <= 1e-12 * np.maximum(1.0, np.abs(E).max(axis=1)))


def test_roots_known_polynomials(gpu):
    # polynomials built from known roots (scaled leading coefficient)
    rng = np.random.default_rng(74)
    n, N = 10, 8
    want = rng.standard_normal((N, n)) + 1j * rng.standard_normal((N, n))
    Cf = np.stack([np.poly(w)[::-1] * (2.0 - 1.5j) for w in want])
    rC, iC = H.vec_to_soa(Cf, 1)
    rE, iE = veig.roots(_dev(rC), _dev(iC), n, N, 1)
    torch.cuda.synchronize()
    E = H.vec_from_soa(rE.cpu().numpy(), iE.cpu().numpy(), n, 1)
    _, err = H.match_eigenvalues(want, E)
    assert err.max() < 1e-8


def test_compat_roots_kernel_host_mirror(gpu):
    # zlahqr(rC, iC, rE, iE, rV, iV, rZ, iZ, dim, dim1, dim2, bc, tc) on host arrays
    n, bc, tc = 6, 2, 4
    N = bc * tc
    Cf = coeffs(N, n, 7400)
    rC, iC = H.vec_to_soa(Cf, tc)
    out = [np.zeros(N * n) for _ in range(2)] + [np.zeros(N * n * n) for _ in range(4)]
    f = getattr(veig.lib(), "_Z6zlahqrPdS_S_S_S_S_S_S_iiiii")
    f.argtypes = [ctypes.POINTER(ctypes.c_double)] * 8 + [ctypes.c_int] * 5
    f.restype = None
    f(H.dptr(rC), H.dptr(iC), *[H.dptr(o) for o in out], n, n + 1, n * n, bc, tc)
    rE, iE, rT, iT, rZ, iZ = veig.roots(_dev(rC), _dev(iC), n, bc, tc, schur=True)
    torch.cuda.synchronize()
    for host, dev in zip(out, (rE, iE, rT, iT, rZ, iZ)):
        assert np.array_equal(host, dev.cpu().numpy())

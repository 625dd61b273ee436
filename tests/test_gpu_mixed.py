"""GPU tests of the mixed-precision path (veig_zgeev_mixed): complex64
solve + FP64 Newton refinement + FP64 fallback, against the CPU oracle.

Bar: every matrix ends with backward error ||AV - VE||_F/(||A||_F ||V||_F)
<= 1e-12 (measured here independently of the kernel's own figure, which must
agree), eigenvalues equal to the FP64 oracle's within 1e-10 * ||A||_F *
cond, where cond is the eigenvector-matrix condition bound of the input
family (1 for Gaussian/Hermitian, the construction's cond(V) otherwise).
"""
from __future__ import annotations

import numpy as np
import pytest
import torch

import veig
from helpers import backward_error, gaussian, match_eigenvalues, oracle_eig, to_soa, from_soa, vec_from_soa
from test_gpu_parity import constructed, hermitian

pytestmark = pytest.mark.gpu


def run_mixed(A, tc=1, tol=1e-12):
    N, n, _ = A.shape
    rA, iA = to_soa(A, tc)
    dev = torch.device("cuda:0")
    trA, tiA = torch.from_numpy(rA).to(dev), torch.from_numpy(iA).to(dev)
    info = torch.full((N,), -7, dtype=torch.int32, device=dev)
    fb = torch.full((N,), -7, dtype=torch.int32, device=dev)
    be = torch.full((N,), -1.0, dtype=torch.float64, device=dev)
    rE, iE, rV, iV = veig.zgeev_mixed(trA, tiA, n, N // tc, tc, info=info, berr=be, fallback=fb, tol=tol)
    torch.cuda.synchronize()
    E = vec_from_soa(rE.cpu().numpy(), iE.cpu().numpy(), n, tc)
    V = from_soa(rV.cpu().numpy(), iV.cpu().numpy(), n, tc)
    return E, V, info.cpu().numpy(), fb.cpu().numpy(), be.cpu().numpy()


def check(A, E, V, info, fb, be, tol, cond=1.0):
    nrm = np.linalg.norm(A, axis=(1, 2))
    my_be = backward_error(A, E, V)
    assert np.all(np.isin(fb, [0, 1])), "fallback flags not written"
    assert np.all(info == 0), f"info {np.unique(info)}"
    assert np.all(my_be <= tol), f"backward error {my_be.max():.2e} > {tol}"
    assert np.allclose(be, my_be, rtol=1e-3, atol=1e-16), "kernel backward error disagrees"
    Eo, _, _ = oracle_eig(A, 1)
    _, err = match_eigenvalues(Eo, E)
    assert np.all(err <= 1e-10 * nrm * cond), f"eigenvalue error {np.max(err / nrm):.2e}"
    return my_be


@pytest.mark.parametrize("n", [4, 8, 16, 24, 32, 40, 64])
def test_mixed_gaussian(gpu, n):
    A = gaussian(128, n, 5000 + n)
    E, V, info, fb, be = run_mixed(A)
    check(A, E, V, info, fb, be, 1e-12)
    assert fb.mean() < 0.05, f"fallback rate {fb.mean():.3f}"


@pytest.mark.parametrize("n", [8, 16])
def test_mixed_hermitian(gpu, n):
    A = hermitian(128, n, 5100 + n)
    E, V, info, fb, be = run_mixed(A)
    check(A, E, V, info, fb, be, 1e-12)


@pytest.mark.parametrize("n,cond", [(24, 1e1), (24, 1e3), (24, 1e6), (16, 1e3), (16, 1e6), (8, 1e6),
                                    (64, 1e3), (48, 1e6)])
def test_mixed_conditioned(gpu, n, cond):
    # configuration of the mixed-precision workload: A = V diag(lam) V^-1
    # with cond(V) up to 1e6; matrices the refinement cannot bring to the
    # tolerance fall back to FP64, so the bar holds for all of them (n <= 16
    # refines from the complex64 Schur factors, n > 16 from V and W; n > 32
    # on four wavefronts per matrix)
    A = constructed(128, n, 5200 + 10 * n + int(np.log10(cond)), cond=cond)
    E, V, info, fb, be = run_mixed(A)
    check(A, E, V, info, fb, be, 1e-12, cond=cond)


def test_mixed_tolerance_and_tiles(gpu):
    A = gaussian(256, 16, 5300)
    for tc in (1, 4, 64):
        E, V, info, fb, be = run_mixed(A, tc=tc, tol=1e-14)
        check(A, E, V, info, fb, be, 1e-14)


@pytest.mark.parametrize("n", [16, 64])
def test_fp32_solve_accuracy(gpu, n):
    # the complex64 building block alone: single-precision accuracy
    from gpu_util import gpu_eig
    A = gaussian(128, n, 5400 + n)
    E, V, info = gpu_eig(A, 1, precision="fp32")
    be = backward_error(A, E, V)
    assert np.all(info == 0)
    assert be.max() < 1e-5 and be.max() > 1e-9, f"fp32 backward error {be.max():.2e}"


def test_mixed_large_order_tiles(gpu):
    A = gaussian(64, 64, 5500)
    for tc in (1, 4):
        E, V, info, fb, be = run_mixed(A, tc=tc, tol=1e-13)
        check(A, E, V, info, fb, be, 1e-13)


@pytest.mark.parametrize("n,tc", [(16, 1), (40, 1), (12, 4), (24, 4)])
def test_mixed_all_fall_back(gpu, n, tc):
    # an unreachable tolerance sends every matrix to the FP64 solve of the
    # list (and its backward-error pass): FP64 results, flagged, info 0
    # (orders 9..32: the banded kernels walking the list, tiles of tc)
    A = gaussian(64, n, 5600 + n)
    E, V, info, fb, be = run_mixed(A, tc=tc, tol=1e-300)
    assert np.all(fb == 1) and np.all(info == 0)
    my_be = backward_error(A, E, V)
    assert my_be.max() < 64 * n * np.finfo(np.float64).eps
    assert np.allclose(be, my_be, rtol=1e-3, atol=1e-16)
    Eo, _, _ = oracle_eig(A, 1)
    _, err = match_eigenvalues(Eo, E)
    assert np.all(err <= 1e-11 * np.linalg.norm(A, axis=(1, 2)))


def test_mixed_workspace_checks(gpu):
    A = gaussian(8, 8, 1)
    rA, iA = to_soa(A, 1)
    dev = torch.device("cuda:0")
    trA, tiA = torch.from_numpy(rA).to(dev), torch.from_numpy(iA).to(dev)
    small = torch.empty(16, dtype=torch.uint8, device=dev)
    # the binding refuses it before the call ...
    with pytest.raises(veig.VeigError, match="workspace: 16 elements"):
        veig.zgeev_mixed(trA, tiA, 8, 8, 1, workspace=small)
    # ... and so does the library itself
    out = [torch.empty(64 * 8, dtype=torch.float64, device=dev) for _ in range(2)] + \
          [torch.empty(64 * 64, dtype=torch.float64, device=dev) for _ in range(2)]
    st = veig.lib().veig_zgeev_mixed(8, 8, 1, veig._ptr(trA), veig._ptr(tiA), *[veig._ptr(t) for t in out],
                                     None, None, None, 1e-12, 30, veig._ptr(small), 16, veig._stream(None))
    assert st == 1 and b"workspace too small" in veig.lib().veig_last_error()

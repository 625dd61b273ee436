"""Orders beyond the lane-group kernels (veig_big.hip): the fused
eigendecomposition at 64 < n <= 96 (the reference's own driver sweeps
n = 10..79, zlahqr.cpp:206-249; its eigensolver() takes any order,
utils.cpp:578) and the stage entries at 32 < n <= 96, against the oracle on
the same seeded inputs at the bar of test_gpu_parity."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import helpers as H
import python_interface as pi
import veig
from gpu_util import gpu_eig
from test_gpu_parity import check_parity, constructed, hermitian
from test_gpu_stages import check_schur_pair

pytestmark = pytest.mark.gpu
EPS = np.finfo(float).eps


def herm(X):
    return np.conj(np.swapaxes(X, -1, -2))


@pytest.mark.parametrize("n,tc", [(65, 1), (79, 3), (96, 2)])
def test_fused_large_orders_vs_oracle(gpu, n, tc):
    print(f"[big] fused n={n} tc={tc}", flush=True)
    A = H.gaussian(6, n, 700 + n)
    E, V, info = gpu_eig(A, tc)
    Eo, Vo, _ = H.oracle_eig(A, tc)
    check_parity(A, E, V, info, Eo, Vo)


@pytest.mark.parametrize("kind", ["hermitian", "constructed"])
def test_fused_large_order_other_families(gpu, kind):
    n = 72
    A = hermitian(4, n, 71) if kind == "hermitian" else constructed(4, n, 72, cond=1e2)
    E, V, info = gpu_eig(A, 1)
    Eo, Vo, _ = H.oracle_eig(A, 1)
    check_parity(A, E, V, info, Eo, Vo, vec_check=kind == "hermitian")


def test_fused_large_order_schur_form(gpu):
    n = 70
    A = H.gaussian(4, n, 770)
    E, V, info, T = gpu_eig(A, 2, write_schur=True)
    assert np.all(info == 0)
    assert np.max(np.abs(np.tril(T, -1))) == 0.0
    assert np.allclose(np.diagonal(T, axis1=1, axis2=2), E, atol=0)
    be = H.backward_error(A, E, V)
    assert np.all(be <= 64 * n * EPS)


def test_host_eigensolver_large_order(gpu):
    # the reference's host call (utils.cpp:578) through the python_interface mirror
    n, bc, tc = 80, 2, 2
    A = H.gaussian(bc * tc, n, 880)
    rA, iA = H.to_soa(A, tc)
    rE, iE = np.zeros(bc * tc * n), np.zeros(bc * tc * n)
    rV, iV = np.zeros_like(rA), np.zeros_like(iA)
    pi.my_eigensolver(n, bc, tc, rA, iA, rE, iE, rV, iV)
    E = H.vec_from_soa(rE, iE, n, tc)
    V = H.from_soa(rV, iV, n, tc)
    assert np.all(H.backward_error(A, E, V) <= 64 * n * EPS)
    Eo, _, _ = H.oracle_eig(A, tc)
    _, err = H.match_eigenvalues(Eo, E)
    assert np.all(err <= 1e-11 * np.linalg.norm(A, axis=(1, 2)))


@pytest.mark.parametrize("n,tc", [(48, 1), (79, 2)])
def test_stage_entries_large_orders(gpu, n, tc):
    print(f"[big] stages n={n} tc={tc}", flush=True)
    N = 4
    A = H.gaussian(N, n, 900 + n)
    orc = H.load_oracle()
    nrm = np.linalg.norm(A, axis=(1, 2))[:, None, None]
    # zgehrd + zunghr against the oracle with LAPACK's sign, and Q H Q^H = A
    rA, iA = H.to_soa(A, tc)
    rt, it = np.zeros(N * n), np.zeros(N * n)
    pi.my_zgehrd(n, N // tc, tc, rA, iA, rt, it)
    Hm = H.from_soa(rA, iA, n, tc)
    rA2, iA2 = H.to_soa(A, tc)
    rt2, it2 = np.zeros(N * n), np.zeros(N * n)
    orc.oracle_zgehrd(n, N // tc, tc, H.dptr(rA2), H.dptr(iA2), H.dptr(rt2), H.dptr(it2), H.LAPACK_SIGN)
    assert np.allclose(Hm, H.from_soa(rA2, iA2, n, tc), atol=1e-12 * nrm.max())
    assert np.allclose(H.vec_from_soa(rt, it, n, tc)[:, : n - 1], H.vec_from_soa(rt2, it2, n, tc)[:, : n - 1],
                       atol=1e-12)
    rQ, iQ = rA.copy(), iA.copy()
    pi.my_zunghr(n, N // tc, tc, rQ, iQ, rt, it)
    Q = H.from_soa(rQ, iQ, n, tc)
    assert np.max(np.abs(herm(Q) @ Q - np.eye(n))) < 50 * n * EPS
    assert np.max(np.abs(Q @ np.triu(Hm, -1) @ herm(Q) - A) / nrm) < 50 * n * EPS
    # hessenberg: H and Q together, as the oracle's
    rA3, iA3 = H.to_soa(A, tc)
    rQ3, iQ3 = np.zeros_like(rA3), np.zeros_like(iA3)
    pi.my_hessenberg(n, N // tc, tc, rA3, iA3, rQ3, iQ3)
    rA4, iA4 = H.to_soa(A, tc)
    rQ4, iQ4 = np.zeros_like(rA4), np.zeros_like(iA4)
    orc.oracle_hessenberg(n, N // tc, tc, H.dptr(rA4), H.dptr(iA4), H.dptr(rQ4), H.dptr(iQ4), H.LAPACK_SIGN)
    assert np.allclose(H.from_soa(rA3, iA3, n, tc), H.from_soa(rA4, iA4, n, tc), atol=1e-12 * nrm.max())
    assert np.allclose(H.from_soa(rQ3, iQ3, n, tc), H.from_soa(rQ4, iQ4, n, tc), atol=1e-12)
    # zlahqr on the clean Hessenberg form with Z = Q, against the oracle's
    Hc = np.triu(Hm, -1)
    rH, iH = H.to_soa(Hc, tc)
    rZ, iZ = H.to_soa(Q, tc)
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(x).to(dev) for x in (rH, iH, rZ, iZ)]
    rE = torch.zeros(N * n, dtype=torch.float64, device=dev)
    iE = torch.zeros_like(rE)
    info = torch.full((N,), -1, dtype=torch.int32, device=dev)
    veig.zlahqr(n, N // tc, tc, rE, iE, t[0], t[1], t[2], t[3], info=info)
    torch.cuda.synchronize()
    E = H.vec_from_soa(rE.cpu().numpy(), iE.cpu().numpy(), n, tc)
    T = H.from_soa(t[0].cpu().numpy(), t[1].cpu().numpy(), n, tc)
    Z = H.from_soa(t[2].cpu().numpy(), t[3].cpu().numpy(), n, tc)
    assert np.all(info.cpu().numpy() == 0)
    assert np.max(np.abs(np.tril(T, -1))) == 0.0
    rH5, iH5 = H.to_soa(Hc, tc)
    rZ5, iZ5 = H.to_soa(Q, tc)
    rE5, iE5 = np.zeros(N * n), np.zeros(N * n)
    orc.oracle_zlahqr(n, N // tc, tc, H.dptr(rE5), H.dptr(iE5), H.dptr(rH5), H.dptr(iH5), H.dptr(rZ5),
                      H.dptr(iZ5), 0, None)
    Eo = H.vec_from_soa(rE5, iE5, n, tc)
    _, err = H.match_eigenvalues(Eo, E)
    assert np.all(err <= 1e-11 * nrm[:, 0, 0])
    # Z unitary, A = (QZ... ) : Z here is Q Z_s, so A = Z T Z^H
    assert np.max(np.abs(herm(Z) @ Z - np.eye(n))) < 100 * n * EPS
    assert np.max(np.abs(Z @ T @ herm(Z) - A) / nrm) < 100 * n * EPS
    if np.allclose(E, Eo, atol=1e-11 * nrm.max()):   # same Schur order: same T and Z up to the gauge
        check_schur_pair(T, Z, H.from_soa(rH5, iH5, n, tc), H.from_soa(rZ5, iZ5, n, tc), nrm.max())
    # ztrevc on that Schur form: R <- Z X, the eigenvectors of A, as the oracle's
    rT, iT = H.to_soa(T, tc)
    rR, iR = H.to_soa(Z, tc)
    pi.my_ztrevec(n, N // tc, tc, rT, iT, rR, iR)
    V = H.from_soa(rR, iR, n, tc)
    rT6, iT6 = H.to_soa(T, tc)
    rR6, iR6 = H.to_soa(Z, tc)
    orc.oracle_ztrevc(n, N // tc, tc, H.dptr(rT6), H.dptr(iT6), H.dptr(rR6), H.dptr(iR6), 0)
    Vo = H.from_soa(rR6, iR6, n, tc)
    assert np.allclose(V, Vo, atol=1e-9 * np.abs(Vo).max())
    assert np.all(H.backward_error(A, E, V) <= 64 * n * EPS)


def test_orders_beyond_the_limit_are_refused(gpu):
    n = 97
    rA = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    with pytest.raises(veig.VeigError, match="exceeds"):
        veig.zgeev(rA, rA.clone(), n, 1, 1)

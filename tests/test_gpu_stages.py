"""GPU tests of the stage entries, the python_interface mirror and the C++
shim against the reference's golden fixtures (tests/golden/), the oracle and
closed-form invariants."""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import pytest

import helpers as H
import python_interface as pi
import veig

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
EPS = np.finfo(float).eps


def herm(X):
    return np.conj(np.swapaxes(X, -1, -2))


def schur_gauge(Z_ref, Z):
    """Unit phases d with Z ~ Z_ref diag(d).

    A Schur form with a fixed eigenvalue order is unique only up to
    T -> D^H T D, Z -> Z D with D diagonal unitary.  In the QR iteration D is
    set by the phases of subdiagonal entries at the rounding level (each is
    rotated to be real as it converges), so two correct implementations that
    round differently agree on E and on |T|, but T and Z agree only after
    this alignment: relative perturbations of 1e-15 of the input already
    change the oracle's own T by O(1) through D."""
    c = np.einsum("nij,nij->nj", np.conj(Z_ref), Z)
    return c / np.abs(c)


def check_schur_pair(T, Z, T_ref, Z_ref, scale):
    d = schur_gauge(Z_ref, Z)
    assert np.allclose(Z, Z_ref * d[:, None, :], atol=1e-10)
    T_al = np.conj(d)[:, :, None] * np.triu(T_ref) * d[:, None, :]
    assert np.allclose(np.triu(T), T_al, atol=1e-10 * scale)


@pytest.mark.parametrize("path", sorted(GOLD.glob("stages_*.npz")), ids=lambda p: p.stem)
def test_hessenberg_stage(gpu, path):
    d = np.load(path)
    A, tc = d["A"], int(d["tc"])
    N, n, _ = A.shape
    rA, iA = H.to_soa(A, tc)
    rQ, iQ = np.zeros_like(rA), np.zeros_like(iA)
    pi.my_hessenberg(n, N // tc, tc, rA, iA, rQ, iQ)
    Hm = H.from_soa(rA, iA, n, tc)
    Q = H.from_soa(rQ, iQ, n, tc)
    Hh = np.triu(Hm, -1)
    nrm = np.linalg.norm(A, axis=(1, 2))[:, None, None]
    assert np.max(np.abs(herm(Q) @ Q - np.eye(n))) < 50 * n * EPS
    assert np.max(np.abs(Q @ Hh @ herm(Q) - A) / nrm) < 50 * n * EPS
    # same reflectors as the oracle with LAPACK's sign (the reference's sign
    # differs: |H| and |tau| entries agree, phases may not)
    rA2, iA2 = H.to_soa(A, tc)
    rQ2, iQ2 = np.zeros_like(rA2), np.zeros_like(iA2)
    H.load_oracle().oracle_hessenberg(n, N // tc, tc, H.dptr(rA2), H.dptr(iA2), H.dptr(rQ2), H.dptr(iQ2),
                                      H.LAPACK_SIGN)
    assert np.allclose(Hm, H.from_soa(rA2, iA2, n, tc), atol=1e-12 * nrm.max())
    assert np.allclose(Q, H.from_soa(rQ2, iQ2, n, tc), atol=1e-12)
    assert np.allclose(np.abs(Hh), np.abs(np.triu(d["H"], -1)), atol=1e-12 * nrm.max())


@pytest.mark.parametrize("path", sorted(GOLD.glob("stages_*.npz")), ids=lambda p: p.stem)
def test_zgehrd_zunghr_stages(gpu, path):
    d = np.load(path)
    A, tc = d["A"], int(d["tc"])
    N, n, _ = A.shape
    rA, iA = H.to_soa(A, tc)
    rt, it = np.zeros(N * n), np.zeros(N * n)
    pi.my_zgehrd(n, N // tc, tc, rA, iA, rt, it)
    Hm = H.from_soa(rA, iA, n, tc)
    rQ, iQ = rA.copy(), iA.copy()
    pi.my_zunghr(n, N // tc, tc, rQ, iQ, rt, it)
    Q = H.from_soa(rQ, iQ, n, tc)
    nrm = np.linalg.norm(A, axis=(1, 2))[:, None, None]
    assert np.max(np.abs(herm(Q) @ Q - np.eye(n))) < 50 * n * EPS
    assert np.max(np.abs(Q @ np.triu(Hm, -1) @ herm(Q) - A) / nrm) < 50 * n * EPS
    # H (with the reflector vectors below the subdiagonal) and tau equal the
    # oracle's with LAPACK's sign
    rA2, iA2 = H.to_soa(A, tc)
    rt2, it2 = np.zeros(N * n), np.zeros(N * n)
    H.load_oracle().oracle_zgehrd(n, N // tc, tc, H.dptr(rA2), H.dptr(iA2), H.dptr(rt2), H.dptr(it2),
                                  H.LAPACK_SIGN)
    tau = H.vec_from_soa(rt, it, n, tc)[:, : n - 1]
    assert np.allclose(tau, H.vec_from_soa(rt2, it2, n, tc)[:, : n - 1], atol=1e-12)
    assert np.allclose(Hm, H.from_soa(rA2, iA2, n, tc), atol=1e-12 * nrm.max())


@pytest.mark.parametrize("path", sorted(GOLD.glob("stages_*.npz")), ids=lambda p: p.stem)
def test_zlahqr_stage_vs_reference(gpu, path):
    # zlahqr_new on the clean Hessenberg form with Z = Q: same algorithm and
    # same shifts as the reference, so T, Z and E agree to rounding (same order)
    d = np.load(path)
    tc = int(d["tc"])
    Hc, Q = d["Hclean"], d["Q"]
    N, n, _ = Hc.shape
    rH, iH = H.to_soa(Hc, tc)
    rZ, iZ = H.to_soa(Q, tc)
    import torch
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
    scale = np.abs(d["E"]).max()
    assert np.allclose(E, d["E"], atol=1e-11 * scale), "eigenvalues / Schur order differ"
    check_schur_pair(T, Z, d["T"], d["Z"], scale)
    assert np.max(np.abs(np.tril(T, -1))) == 0.0


@pytest.mark.parametrize("path", sorted(GOLD.glob("stages_*.npz")), ids=lambda p: p.stem)
def test_call_zlahqr_new_and_ztrevc_vs_reference(gpu, path):
    d = np.load(path)
    tc = int(d["tc"])
    a = d["cz_in"]
    N, n, _ = a.shape
    rM, iM = H.to_soa(a, tc)
    rS, iS = np.zeros_like(rM), np.zeros_like(iM)
    rE, iE = np.zeros(N * n), np.zeros(N * n)
    pi.my_zlahqr(n, N // tc, tc, rE, iE, rM, iM, rS, iS)   # Z starts as i*I (utils.cpp:169)
    E = H.vec_from_soa(rE, iE, n, tc)
    S = H.from_soa(rS, iS, n, tc)
    scale = max(1.0, np.abs(d["cz_E"]).max())
    assert np.allclose(E, d["cz_E"], atol=1e-11 * scale)
    check_schur_pair(H.from_soa(rM, iM, n, tc), S, d["cz_T"], d["cz_S"], scale)
    # ztrevc on the reference's Schur form: eigenvectors equal the reference's
    rT, iT = H.to_soa(d["Tin"], tc)
    rR, iR = H.to_soa(d["Z"], tc)
    pi.my_ztrevec(n, N // tc, tc, rT, iT, rR, iR)
    V = H.from_soa(rR, iR, n, tc)
    assert np.allclose(V, d["V"], atol=1e-10 * np.abs(d["V"]).max())


def test_zgemv_and_backsubstitution(gpu):
    d = np.load(GOLD / "stages_n8.npz")
    A, x, y, tc = d["A"], d["gemv_x"], d["gemv_y"], int(d["tc"])
    N, n, _ = A.shape
    for herm_flag in (False, True):
        rM, iM = H.to_soa(A, tc)
        rx, ix = H.vec_to_soa(x, tc)
        ry, iy = H.vec_to_soa(y, tc)
        pi.my_zgemv(herm_flag, n, N // tc, tc, rM, iM, rx, ix, 1.0, 0.0, ry, iy)
        got = H.vec_from_soa(ry, iy, n, tc)
        M = np.conj(np.swapaxes(A, 1, 2)) if herm_flag else A
        want = np.einsum("nij,nj->ni", M, x) + y     # LAPACK semantics (the reference's Hermitian branch is off)
        assert np.allclose(got, want, atol=1e-12)
        if not herm_flag:
            assert np.allclose(got, d["gemv_out_h0"], atol=1e-12)
    rT, iT = H.to_soa(d["bs_T"], tc)
    rb, ib = H.vec_to_soa(y, tc)
    pi.my_backsubs(n, N // tc, tc, rT, iT, rb, ib)
    xs = H.vec_from_soa(rb, ib, n, tc)
    assert np.allclose(xs, d["bs_x"], atol=1e-13)
    assert np.allclose(np.einsum("nij,nj->ni", d["bs_T"], xs), y, atol=1e-12)


def test_python_test_helpers(gpu):
    rng = np.random.default_rng(3)
    r = pi.test_backsubs(8, 2, 4, rng=rng)
    assert np.max(np.abs(r)) < 1e-12
    h, q, a = pi.test_hessenberg(10, 2, 4, rng=rng)
    assert np.allclose(q @ np.triu(h, -1) @ np.conj(q.T), a, atol=1e-13)
    assert np.max(pi.test_zgemv(9, 1, 4, rng=rng)) < 1e-12
    assert np.max(pi.test_zgemv(9, 1, 4, hermitian=True, rng=rng)) < 1e-12
    eigs, schurv, t, a = pi.test_zlahqr(12, 2, 4, rng=rng)
    Z = schurv / 1j                                          # S started as i*I
    assert np.allclose(Z @ t @ np.conj(Z.T), a, atol=1e-12)
    eigs, vecs, a = pi.test_eigenvectors(12, 2, 4, rng=rng)
    assert np.max(np.abs(a @ vecs - vecs * eigs)) < 1e-12 * np.abs(vecs).max()


def test_fused_golden_through_gpu(gpu):
    # the reference's fused results (cleared reflectors) reproduced by the GPU
    from gpu_util import gpu_eig
    from test_gpu_parity import check_parity
    for path in sorted(GOLD.glob("eig_*.npz")):
        d = np.load(path)
        A, tc = d["A"], int(d["tc"])
        E, V, info = gpu_eig(A, tc)
        check_parity(A, E, V, info, d["E"], d["V"], vec_check="clus" not in path.stem)


def test_host_eigensolver_and_cpp_shim(gpu):
    A = H.gaussian(16, 12, 9)
    tc = 4
    outs = []
    for fn in ("my_eigensolver", "cpp"):
        rA, iA = H.to_soa(A, tc)
        rE, iE = np.zeros(16 * 12), np.zeros(16 * 12)
        rV, iV = np.zeros(16 * 144), np.zeros(16 * 144)
        if fn == "cpp":
            f = getattr(veig.lib(), "_Z11eigensolveriiiPdS_S_S_S_S_")
            f.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_double)] * 6
            f.restype = None
            f(12, 4, tc, *(H.dptr(x) for x in (rA, iA, rE, iE, rV, iV)))
        else:
            pi.my_eigensolver(12, 4, tc, rA, iA, rE, iE, rV, iV)
        outs.append((H.vec_from_soa(rE, iE, 12, tc), H.from_soa(rV, iV, 12, tc), H.from_soa(rA, iA, 12, tc)))
    (E1, V1, T1), (E2, V2, T2) = outs
    assert np.array_equal(E1, E2) and np.array_equal(V1, V2) and np.array_equal(T1, T2)
    assert np.max(H.backward_error(A, E1, V1)) < 1e-13
    assert np.max(np.abs(np.tril(T1, -1))) == 0.0            # A receives the Schur form
    assert np.allclose(np.diagonal(T1, 0, 1, 2), E1)

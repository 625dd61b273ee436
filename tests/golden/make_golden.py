"""Write the golden fixtures of tests/golden/ from the reference's own code.

Runs oracle/_ref/libref_exact.so -- the reference's utils.cpp (with
Toolbox.hpp and ComplexDouble.hpp) compiled from the reference tree by
oracle/Makefile with FMA contraction off -- on small seeded inputs and stores
inputs and outputs as .npz (data only).  Needs the reference tree, so it runs
in the build container, never on the GPU host:

    make -C oracle ref && python tests/golden/make_golden.py

Fixtures (complex arrays are (N, n, n) / (N, n), matrix index first):
  eig_<fam>_n<n>.npz     eigensolver() with the reflector entries cleared
                         before QR: A, E, V, T (A on exit); thread_cnt = tc
  trash_n<n>.npz         eigensolver() exactly as shipped (utils.cpp:578):
                         A, E, V, T -- documents the corrupted first sweep
  stages_n<n>.npz        call_zgehrd (H, tau), hessenberg (Q), zlahqr_new on
                         the clean Hessenberg form (T, Z, E), call_ztrevc (V),
                         call_zlahqr_new on a real triu(a,-1) input as the
                         Cython tests use it (S, E, T), call_zgemv both ways
                         with beta = 1, call_backsubstitution
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import helpers as H  # noqa: E402

N_PER = 8


def fam_input(fam: str, N: int, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((N, n, n)) + 1j * rng.standard_normal((N, n, n))
    if fam == "gauss":
        return B
    if fam == "herm":
        return 0.5 * (B + np.conj(np.transpose(B, (0, 2, 1))))
    if fam == "clus":  # V diag(lam) V^-1, lam in clusters 1e-6 apart, cond(V) ~ 1e2
        Q1, _ = np.linalg.qr(rng.standard_normal((N, n, n)) + 1j * rng.standard_normal((N, n, n)))
        Q2, _ = np.linalg.qr(rng.standard_normal((N, n, n)) + 1j * rng.standard_normal((N, n, n)))
        V = (Q1 * np.logspace(0, -2, n)[None, None, :]) @ np.conj(np.transpose(Q2, (0, 2, 1)))
        ncl = max(1, n // 3)
        c = rng.standard_normal((N, ncl)) + 1j * rng.standard_normal((N, ncl))
        lam = c[:, np.arange(n) % ncl] + 1e-6 * (np.arange(n) // ncl)[None, :]
        return V @ (lam[:, :, None] * np.linalg.inv(V))
    raise ValueError(fam)


def ref_eig(lib, A, tc, clean):
    E, V, T = H.ref_eig(A, tc, clean=clean, lib=lib)
    return E, V, T


def stages(lib, A, tc):
    N, n, _ = A.shape
    bc = N // tc
    out = {}
    rA, iA = H.to_soa(A, tc)
    rt = np.zeros(N * n); it = np.zeros(N * n)
    lib.ref_zgehrd(n, bc, tc, H.dptr(rA), H.dptr(iA), H.dptr(rt), H.dptr(it))
    out["H"] = H.from_soa(rA, iA, n, tc)
    out["tau"] = H.vec_from_soa(rt, it, n, tc)
    rA, iA = H.to_soa(A, tc)
    rQ = np.zeros(N * n * n); iQ = np.zeros(N * n * n)
    lib.ref_hessenberg(n, bc, tc, H.dptr(rA), H.dptr(iA), H.dptr(rQ), H.dptr(iQ))
    Q = H.from_soa(rQ, iQ, n, tc)
    out["Q"] = Q
    # QR on the clean Hessenberg form with Z = Q
    Hc = np.triu(H.from_soa(rA, iA, n, tc), -1)
    out["Hclean"] = Hc
    rH, iH = H.to_soa(Hc, tc)
    rZ, iZ = H.to_soa(Q, tc)
    rE = np.zeros(N * n); iE = np.zeros(N * n)
    lib.ref_zlahqr_new(n, bc, tc, H.dptr(rE), H.dptr(iE), H.dptr(rH), H.dptr(iH), H.dptr(rZ), H.dptr(iZ))
    out["T"] = H.from_soa(rH, iH, n, tc)
    out["Z"] = H.from_soa(rZ, iZ, n, tc)
    out["E"] = H.vec_from_soa(rE, iE, n, tc)
    # ztrevc on that Schur form
    T = np.triu(out["T"])
    out["Tin"] = T
    rT, iT = H.to_soa(T, tc)
    rR, iR = H.to_soa(out["Z"], tc)
    lib.ref_ztrevc(n, bc, tc, H.dptr(rT), H.dptr(iT), H.dptr(rR), H.dptr(iR))
    out["V"] = H.from_soa(rR, iR, n, tc)
    out["Tout"] = H.from_soa(rT, iT, n, tc)
    # call_zlahqr_new as the Cython tests use it: real triu(a,-1), S set inside
    rng = np.random.default_rng(n + 77)
    a = np.triu(rng.random((N, n, n)), -1).astype(complex)
    out["cz_in"] = a
    rM, iM = H.to_soa(a, tc)
    rS = np.zeros(N * n * n); iS = np.zeros(N * n * n)
    rE = np.zeros(N * n); iE = np.zeros(N * n)
    lib.ref_call_zlahqr_new(n, bc, tc, H.dptr(rE), H.dptr(iE), H.dptr(rM), H.dptr(iM), H.dptr(rS), H.dptr(iS))
    out["cz_E"] = H.vec_from_soa(rE, iE, n, tc)
    out["cz_T"] = H.from_soa(rM, iM, n, tc)
    out["cz_S"] = H.from_soa(rS, iS, n, tc)
    # zgemv, beta = 1 (the Hermitian branch's known defect, Toolbox.hpp:1125)
    x = rng.standard_normal((N, n)) + 1j * rng.standard_normal((N, n))
    y = rng.standard_normal((N, n)) + 1j * rng.standard_normal((N, n))
    out["gemv_x"], out["gemv_y"] = x, y
    for herm in (0, 1):
        rM, iM = H.to_soa(A, tc)
        rx, ix = H.vec_to_soa(x, tc)
        ry, iy = H.vec_to_soa(y, tc)
        lib.ref_zgemv(herm, n, bc, tc, H.dptr(rM), H.dptr(iM), H.dptr(rx), H.dptr(ix), H.dptr(ry),
                      H.dptr(iy), 1.0, 0.0)
        out[f"gemv_out_h{herm}"] = H.vec_from_soa(ry, iy, n, tc)
    # backsubstitution on triu(A) + n*I (well conditioned)
    Tb = np.triu(A) + n * np.eye(n)[None]
    out["bs_T"] = Tb
    rT, iT = H.to_soa(Tb, tc)
    rb, ib = H.vec_to_soa(y, tc)
    lib.ref_backsubstitution(n, bc, tc, H.dptr(rT), H.dptr(iT), H.dptr(rb), H.dptr(ib))
    out["bs_x"] = H.vec_from_soa(rb, ib, n, tc)
    return out


def main() -> None:
    lib = H.load_ref(exact=True)
    if lib is None:
        raise SystemExit("oracle/_ref/libref_exact.so is not built (make -C oracle ref)")
    written = []
    for fam, sizes in (("gauss", (1, 2, 3, 4, 5, 8, 10, 16, 24, 32)), ("herm", (4, 8, 16)),
                       ("clus", (6, 12, 16))):
        for n in sizes:
            N = N_PER if n <= 16 else 4
            tc = 4 if N % 4 == 0 else 1
            A = fam_input(fam, N, n, 10_000 + 100 * n + len(fam))
            E, V, T = ref_eig(lib, A, tc, True)
            p = HERE / f"eig_{fam}_n{n}.npz"
            np.savez_compressed(p, A=A, E=E, V=V, T=T, tc=tc)
            written.append(p)
    for n in (4, 8, 16):
        A = fam_input("gauss", N_PER, n, 20_000 + n)
        E, V, T = ref_eig(lib, A, 4, False)
        p = HERE / f"trash_n{n}.npz"
        np.savez_compressed(p, A=A, E=E, V=V, T=T, tc=4)
        written.append(p)
        A = fam_input("gauss", N_PER, n, 30_000 + n)
        st = stages(lib, A, 4)
        p = HERE / f"stages_n{n}.npz"
        np.savez_compressed(p, A=A, tc=4, **st)
        written.append(p)
    total = sum(p.stat().st_size for p in written)
    print(f"wrote {len(written)} fixtures, {total / 1024:.0f} KiB")


if __name__ == "__main__":
    main()

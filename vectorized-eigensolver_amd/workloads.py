"""Synthetic inputs of the five measured configurations, generated on the GPU
with torch (fixed seeds), in the layouts the entry points take.

  c1  4,096 complex128 8x8, re/im iid N(0,1)                 tiled SoA
  c2  1,048,576 complex128 16x16, re/im iid N(0,1)           tiled SoA (headline)
  c3  262,144 matrices, n uniform in [4, 32]: 50% Gaussian, 25% Hermitian
      (B + B^H)/2, 25% V diag(lam) V^-1 with lam in clusters spaced 1e-6 and
      cond(V) = 1e2                                          per-matrix offsets
  c4  524,288 complex128 24x24, A = V diag(lam) V^-1, lam iid CN(0,1),
      cond(V) log-uniform in [1e1, 1e6]                      tiled SoA
  c5  65,536 complex128 64x64, re/im iid N(0,1)              tile 1 (one matrix
      after another: a wavefront per matrix reads and writes whole columns)

V is built as U diag(sigma) W^H with Haar-random unitary U, W (QR of a complex
Gaussian matrix with the phase fix) and sigma log-spaced from 1 down to
1/cond; V^{-1} = W diag(1/sigma) U^H.
"""
from __future__ import annotations

import math

import torch

CONFIGS = {
    "c1": dict(n=8, batch=4096, seed=1001),
    "c2": dict(n=16, batch=1 << 20, seed=1002),
    "c3": dict(nmin=4, nmax=32, batch=262144, seed=1003),
    "c4": dict(n=24, batch=524288, seed=1004),
    "c5": dict(n=64, batch=65536, seed=1005, tile=1),
    # polynomial roots (the reference's companion-matrix kernel): degree 16,
    # complex Gaussian coefficients c_0..c_16
    "roots": dict(n=16, batch=1 << 20, seed=1006),
    # the reference driver's sweep (zlahqr.cpp:196-252): 13,312 matrices per order
    "sweep": dict(nmin=10, nmax=79, batch=13312, seed=1100),
    # the layout converters on one float64 plane of the c2 batch
    "layout": dict(n=16, batch=1 << 20, seed=1107),
}


def _gen(seed, device):
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    return g


def gaussian_soa(N: int, n: int, tile: int, seed: int, device="cuda"):
    """re/im planes (N*n*n float64 each) in the tiled SoA layout with the given
    tile; entries iid N(0,1).  Layout order is irrelevant for iid data, so
    the planes are drawn directly."""
    g = _gen(seed, device)
    rA = torch.randn(N * n * n, dtype=torch.float64, device=device, generator=g)
    iA = torch.randn(N * n * n, dtype=torch.float64, device=device, generator=g)
    return rA, iA


def coeff_soa(N: int, n: int, tile: int, seed: int, device="cuda"):
    """Coefficients c_0..c_n of N polynomials (re/im iid N(0,1)) in the
    vector layout with n+1 entries per polynomial, tile `tile`."""
    g = _gen(seed, device)
    re = torch.randn(N * (n + 1), dtype=torch.float64, device=device, generator=g)
    im = torch.randn(N * (n + 1), dtype=torch.float64, device=device, generator=g)
    return re, im


def haar_unitary(N: int, n: int, g, device):
    """Haar-distributed unitary matrices: the Q of the QR factorisation of a
    complex Gaussian matrix whose R has a positive real diagonal (Mezzadri,
    "How to generate random matrices from the classical compact groups",
    2007).  Q comes from Gram-Schmidt with reorthogonalisation (each column
    projected out twice), batched over the N matrices -- its R diagonal is
    the positive column norm, so no phase correction is needed -- rather than
    a batched library QR, which is slow for many small matrices on the GPU."""
    Z = torch.complex(torch.randn(N, n, n, dtype=torch.float64, device=device, generator=g),
                      torch.randn(N, n, n, dtype=torch.float64, device=device, generator=g))
    Q = torch.empty_like(Z)
    for j in range(n):
        v = Z[:, :, j:j + 1]
        if j:
            P = Q[:, :, :j]
            for _ in range(2):
                v = v - P @ (P.conj().transpose(-2, -1) @ v)
        Q[:, :, j:j + 1] = v / torch.linalg.vector_norm(v, dim=1, keepdim=True)
    return Q


def conditioned(N: int, n: int, cond, g, device, lam=None):
    """A = V diag(lam) V^{-1} with V = U diag(sigma) W^H, cond(V) = cond
    (scalar or per-matrix tensor); V^{-1} = W diag(1/sigma) U^H is formed
    directly, so no solve is needed."""
    U, W = haar_unitary(N, n, g, device), haar_unitary(N, n, g, device)
    c = torch.as_tensor(cond, dtype=torch.float64, device=device).reshape(-1, 1)
    t = torch.linspace(0, 1, n, dtype=torch.float64, device=device).reshape(1, -1)
    sig = torch.exp(-t * torch.log(c)).expand(N, n)             # 1 .. 1/cond
    if lam is None:
        lam = torch.complex(torch.randn(N, n, dtype=torch.float64, device=device, generator=g),
                            torch.randn(N, n, dtype=torch.float64, device=device, generator=g)) / math.sqrt(2)
    WH = W.conj().transpose(-2, -1).contiguous()
    M = (WH * lam.unsqueeze(-2)) @ W                            # W^H diag(lam) W
    left = (U * sig.unsqueeze(1).to(U.dtype)).contiguous()      # U diag(sigma)
    right = (U.conj().transpose(-2, -1) / sig.unsqueeze(2).to(U.dtype)).contiguous()  # diag(1/sigma) U^H
    return left @ M @ right


def clustered_lam(N: int, n: int, g, device, spacing=1e-6):
    ncl = max(1, n // 3)
    c = torch.complex(torch.randn(N, ncl, dtype=torch.float64, device=device, generator=g),
                      torch.randn(N, ncl, dtype=torch.float64, device=device, generator=g)) / math.sqrt(2)
    idx = torch.arange(n, device=device) % ncl
    off = spacing * (torch.arange(n, device=device) // ncl).to(torch.float64)
    return c[:, idx] + off.unsqueeze(0)


def to_soa_planes(A: torch.Tensor, tile: int):
    """(N, n, n) complex -> tiled SoA (re, im) planes."""
    N, n, _ = A.shape
    X = A.reshape(N // tile, tile, n, n).permute(0, 3, 2, 1).contiguous().reshape(-1)
    return X.real.contiguous(), X.imag.contiguous()


def c4_soa(N: int, n: int, tile: int, seed: int, device="cuda", chunk=65536):
    g = _gen(seed, device)
    rs, is_ = [], []
    for s in range(0, N, chunk):
        m = min(chunk, N - s)
        cond = torch.pow(10.0, 1 + 5 * torch.rand(m, dtype=torch.float64, device=device, generator=g))
        A = conditioned(m, n, cond, g, device)
        r, i = to_soa_planes(A, tile)
        rs.append(r)
        is_.append(i)
    return torch.cat(rs), torch.cat(is_)


def c3_dims(N: int, nmin: int, nmax: int, seed: int, device="cuda"):
    g = _gen(seed, device)
    dims = torch.randint(nmin, nmax + 1, (N,), device=device, generator=g, dtype=torch.int64)
    kind = torch.randint(0, 4, (N,), device=device, generator=g)   # 0,1 Gaussian, 2 Hermitian, 3 clustered
    return dims, kind


def fill_c3(vb, dims: torch.Tensor, kind: torch.Tensor, seed: int, device="cuda"):
    """Fill a veig.VarBatch (per-matrix column-major, offsets) with the c3 mix."""
    g = _gen(seed + 1, device)
    a_off = vb.a_off
    for n in range(int(dims.min()), int(dims.max()) + 1):
        for k in range(4):
            sel = torch.nonzero((dims == n) & (kind == k)).flatten()
            m = int(sel.numel())
            if m == 0:
                continue
            B = torch.complex(torch.randn(m, n, n, dtype=torch.float64, device=device, generator=g),
                              torch.randn(m, n, n, dtype=torch.float64, device=device, generator=g))
            if k == 2:
                A = 0.5 * (B + B.conj().transpose(-2, -1))
            elif k == 3:
                A = conditioned(m, n, 1e2, g, device, lam=clustered_lam(m, n, g, device))
            else:
                A = B
            # column-major per matrix: element (r, c) at a_off + r + c*n
            Acm = A.transpose(-2, -1).reshape(m, n * n)
            idx = a_off[sel].unsqueeze(1) + torch.arange(n * n, device=device).unsqueeze(0)
            vb.rA[idx.reshape(-1)] = Acm.real.reshape(-1)
            vb.iA[idx.reshape(-1)] = Acm.imag.reshape(-1)

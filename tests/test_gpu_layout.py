"""The device layout converters (veig_rearrange_matrices, veig_arrange_matrices_aos,
veig_rep_matrix_soa, veig_rep_vector_soa; the reference's host loops
utils.cpp:85-157 behind python_interface.pyx:31-39, 92-96) against the
library's host versions of those loops and numpy, bit for bit: they only move
doubles.  One batch has planes past 2^31 bytes, where 32-bit offsets would
wrap; it is checked on the device against the index formula of the layout."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import helpers as H
import python_interface as pi
import veig

pytestmark = pytest.mark.gpu

MANGLED = {
    "rearrangeMatrices": "_Z17rearrangeMatricesiiiPdS_",
    "arrangeMatricesAoS": "_Z18arrangeMatricesAoSiiiPdS_",
    "repMatrixSoA": "_Z12repMatrixSoAiiiPdS_",
    "repVectorSoA": "_Z12repVectorSoAiiiPdS_",
}


def host(name, dim, bc, tc, src, dst):
    f = getattr(veig.lib(), MANGLED[name])
    f.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_double)] * 2
    f.restype = None
    f(dim, bc, tc, H.dptr(src), H.dptr(dst))


def dev(a):
    return torch.from_numpy(a).cuda()


@pytest.mark.parametrize("dim,bc,tc", [(3, 2, 5), (8, 4, 4), (1, 3, 1), (16, 3, 64), (5, 2, 100),
                                       (24, 2, 130), (7, 1, 1), (2, 5, 63)])
def test_converters_match_host_loops(gpu, dim, bc, tc):
    N = bc * tc
    rng = np.random.default_rng(1000 * dim + tc)
    aos = rng.standard_normal(N * dim * dim)
    soa_h = np.zeros_like(aos)
    host("rearrangeMatrices", dim, bc, tc, aos, soa_h)
    soa_d = torch.zeros(N * dim * dim, dtype=torch.float64, device=gpu)
    pi.arrangeMatrices(dim, bc, tc, dev(aos), soa_d)
    back_d = torch.zeros_like(soa_d)
    pi.arrangeMatricesAoS(dim, bc, tc, soa_d, back_d)
    torch.cuda.synchronize()
    assert np.array_equal(soa_d.cpu().numpy(), soa_h)
    assert np.array_equal(back_d.cpu().numpy(), aos)
    # numpy path of the same helpers gives the same planes
    soa_n = np.zeros_like(aos)
    pi.arrangeMatrices(dim, bc, tc, aos, soa_n)
    assert np.array_equal(soa_n, soa_h)

    m = rng.standard_normal(dim * dim)
    rep_h = np.zeros(N * dim * dim)
    host("repMatrixSoA", dim, bc, tc, m, rep_h)
    rep_d = torch.full((N * dim * dim,), np.nan, dtype=torch.float64, device=gpu)
    pi.my_repMatrix(dim, bc, tc, dev(m), rep_d)
    v = rng.standard_normal(dim)
    repv_h = np.zeros(N * dim)
    host("repVectorSoA", dim, bc, tc, v, repv_h)
    repv_d = torch.full((N * dim,), np.nan, dtype=torch.float64, device=gpu)
    pi.my_repVector(dim, bc, tc, dev(v), repv_d)
    torch.cuda.synchronize()
    assert np.array_equal(rep_d.cpu().numpy(), rep_h)
    assert np.array_equal(repv_d.cpu().numpy(), repv_h)


def test_converters_past_2gib(gpu):
    # 16,400 tiles of 64 order-16 matrices: 2.15e9 bytes per plane
    dim, bc, tc = 16, 16400, 64
    n2 = dim * dim
    total = bc * tc * n2
    assert total * 8 > 2**31
    print(f"[layout] dim={dim} bc={bc} tc={tc} plane={total * 8 / 2**30:.2f} GiB", flush=True)
    aos = torch.arange(total, dtype=torch.float64, device=gpu)  # value = AoS index (exact below 2^53)
    soa = torch.empty_like(aos)
    veig.layout("soa", dim, bc, tc, aos, soa)
    # SoA index k = b T n^2 + e T + t holds AoS element (b T + t) n^2 + e
    k = torch.arange(total, dtype=torch.int64, device=gpu)
    b = k // (tc * n2)
    rem = k - b * (tc * n2)
    e = rem // tc
    t = rem - e * tc
    want = (b * tc + t) * n2 + e
    del k, b, rem, e, t
    assert torch.equal(soa.to(torch.int64), want)
    del want
    back = torch.empty_like(aos)
    veig.layout("aos", dim, bc, tc, soa, back)
    assert torch.equal(back, aos)
    del back
    # replication into the same large plane: slot value = the matrix's entry e
    m = torch.arange(n2, dtype=torch.float64, device=gpu)
    veig.layout("repm", dim, bc, tc, m, soa)
    k = torch.arange(total, dtype=torch.int64, device=gpu)
    assert torch.equal(soa.to(torch.int64), (k % (tc * n2)) // tc)

"""GPU tests of batches of different orders (veig_zgeev_varsize and its
captured hipGraph) against the CPU oracle, matrix by matrix."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import veig
from helpers import backward_error, gaussian, match_eigenvalues, oracle_eig
from test_gpu_parity import check_parity, constructed, hermitian

pytestmark = pytest.mark.gpu


def build(dims, seed=0):
    """VarBatch with the mix of the uneven-order workload (262,144 matrices,
    n in [4, 32]): half Gaussian, a quarter Hermitian, a quarter
    V diag(lam) V^-1 with cond(V) = 1e2 and lam in clusters 1e-6 apart;
    returns (vb, mats)."""
    rng = np.random.default_rng(seed)
    vb = veig.VarBatch(dims)
    mats = []
    a_off = vb.a_off.cpu().numpy()
    rA = np.zeros(vb.total_a)
    iA = np.zeros(vb.total_a)
    for m, n in enumerate(dims):
        if n <= 0 or n > 32:
            mats.append(None)
            continue
        kind = m % 4
        s = int(rng.integers(1 << 30))
        A = (gaussian(1, n, s) if kind < 2 else hermitian(1, n, s) if kind == 2
             else constructed(1, n, s, cond=1e2, cluster=1e-6))[0]
        mats.append(A)
        cm = A.T.reshape(-1)
        rA[a_off[m]:a_off[m] + n * n] = cm.real
        iA[a_off[m]:a_off[m] + n * n] = cm.imag
    vb.rA.copy_(torch.from_numpy(rA))
    vb.iA.copy_(torch.from_numpy(iA))
    return vb, mats


def fetch(vb, dims):
    a_off = vb.a_off.cpu().numpy()
    e_off = vb.e_off.cpu().numpy()
    rV, iV = vb.rV.cpu().numpy(), vb.iV.cpu().numpy()
    rE, iE = vb.rE.cpu().numpy(), vb.iE.cpu().numpy()
    out = []
    for m, n in enumerate(dims):
        if n <= 0 or n > 32:
            out.append(None)
            continue
        V = (rV[a_off[m]:a_off[m] + n * n] + 1j * iV[a_off[m]:a_off[m] + n * n]).reshape(n, n).T
        E = rE[e_off[m]:e_off[m] + n] + 1j * iE[e_off[m]:e_off[m] + n]
        out.append((E, V))
    return out


def check_all(mats, res, info, dims):
    for n in sorted(set(int(d) for d in dims if 0 < d <= 32)):
        idx = [m for m, d in enumerate(dims) if d == n]
        A = np.stack([mats[m] for m in idx])
        E = np.stack([res[m][0] for m in idx])
        V = np.stack([res[m][1] for m in idx])
        Eo, Vo, _ = oracle_eig(A, 1)
        check_parity(A, E, V, info[idx], Eo, Vo, vec_check=False)


def test_varsize_parity(gpu):
    rng = np.random.default_rng(3)
    dims = rng.integers(1, 33, 700).tolist()
    vb, mats = build(dims, 1)
    vb.run()
    torch.cuda.synchronize()
    check_all(mats, fetch(vb, dims), vb.info.cpu().numpy(), dims)


def test_varsize_edge_orders(gpu):
    dims = [0, 1, 2, 33, 32, 40, 7, 0, 16, 1]
    vb, mats = build(dims, 2)
    vb.run()
    torch.cuda.synchronize()
    info = vb.info.cpu().numpy()
    assert info[3] == -1 and info[5] == -1          # orders beyond 32 rejected per matrix
    assert info[0] == 0 and info[7] == 0            # empty matrices: nothing to do
    res = fetch(vb, dims)
    for m, n in enumerate(dims):
        if 0 < n <= 32:
            E, V = res[m]
            assert backward_error(mats[m][None], E[None], V[None])[0] < 1e-13


def test_varsize_graph_replays(gpu):
    rng = np.random.default_rng(4)
    dims = rng.integers(4, 33, 300).tolist()
    vb, mats = build(dims, 5)
    vb.run()
    torch.cuda.synchronize()
    ref = [r for r in fetch(vb, dims)]
    g = vb.graph()
    for rep in range(3):
        vb.rV.zero_()
        vb.rE.zero_()
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        got = fetch(vb, dims)
        for a, b in zip(ref, got):
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), f"replay {rep} differs"
        _ = torch.ones(1000, device="cuda").sum().item()   # eager work between replays
    # the graph reads the inputs at launch time
    vb.rA.mul_(2.0)
    vb.iA.mul_(2.0)
    g.launch()
    torch.cuda.synchronize()
    got = fetch(vb, dims)
    for m in range(len(dims)):
        _, err = match_eigenvalues(2 * ref[m][0][None], got[m][0][None])
        assert err[0] <= 1e-11 * np.abs(ref[m][0]).max() * 2


def test_varsize_variant_table(gpu):
    # other kernel variants per order (lane-group width and lanes per matrix,
    # the autotuner's knob) give the same eigenpairs
    rng = np.random.default_rng(6)
    dims = rng.integers(2, 17, 200).tolist()
    vb, mats = build(dims, 7)
    vb.run()
    torch.cuda.synchronize()
    base = fetch(vb, dims)
    for variants in ({n: 4 for n in range(1, 17)},                                  # groups of 32
                     {**{n: 1 for n in range(1, 9)}, **{n: 3 for n in range(9, 17)}}):  # 2 indices per lane
        table = np.full(33, -1, np.int8)
        for n, v in variants.items():
            table[n] = v
        vb.variant_table = torch.from_numpy(table).cuda()
        vb.grid = [256, 512, 768, 1024, 1280]
        vb.rV.zero_()
        vb.run()
        torch.cuda.synchronize()
        alt = fetch(vb, dims)
        for m, n in enumerate(dims):
            _, err = match_eigenvalues(base[m][0][None], alt[m][0][None])
            assert err[0] < 1e-12 * max(1.0, np.abs(base[m][0]).max())
            assert backward_error(mats[m][None], alt[m][0][None], alt[m][1][None])[0] < 1e-13


def test_varsize_autotune(gpu):
    rng = np.random.default_rng(8)
    dims = rng.integers(1, 33, 600).tolist()
    vb, mats = build(dims, 9)
    tab, grid = vb.autotune()
    assert len(tab) == 33 and len(grid) == 5
    for n in range(1, 33):
        v = tab[n]
        width = 8 if v < 2 else (16 if v < 4 else 32)
        assert 0 <= v <= 4 and width >= n, (n, v)
    vb.rV.zero_()
    vb.run()
    torch.cuda.synchronize()
    got = fetch(vb, dims)
    assert np.all(vb.info.cpu().numpy() == 0)
    for m, n in enumerate(dims):
        assert backward_error(mats[m][None], got[m][0][None], got[m][1][None])[0] < 1e-13

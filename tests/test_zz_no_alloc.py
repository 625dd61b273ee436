"""The allocation-free entry veig_zgeev_varsize_noalloc (the library's static
scratch in place of the caller's workspace) against veig_zgeev_varsize on the
same batches, bit for bit: the same kernels on the same inputs.  Two calls in
a row, the second on a smaller batch, so that the second finds the first's
binning state in the shared array and must clear it."""
from __future__ import annotations

import numpy as np
import pytest
import torch

import veig
from test_gpu_varsize import build, fetch

pytestmark = pytest.mark.gpu


def _same(a, b):
    for x, y in zip(a, b):
        if x is None:
            assert y is None
            continue
        assert np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1])


@pytest.mark.parametrize("sizes", [(700, 150)])
def test_noalloc_matches_workspace_entry(gpu, sizes):
    rng = np.random.default_rng(11)
    for call, nmat in enumerate(sizes):
        dims = rng.integers(0, 33, nmat).tolist()
        print(f"[noalloc] call {call}: nmat={nmat}", flush=True)
        vb, _ = build(dims, 20 + call)
        vb.run()
        torch.cuda.synchronize()
        ref = fetch(vb, dims)
        ref_info = vb.info.cpu().numpy().copy()
        vb.rV.fill_(np.nan)
        vb.iV.fill_(np.nan)
        vb.rE.fill_(np.nan)
        vb.iE.fill_(np.nan)
        vb.info.fill_(-7)
        vb.run_noalloc()
        torch.cuda.synchronize()
        _same(ref, fetch(vb, dims))
        assert np.array_equal(ref_info, vb.info.cpu().numpy())


def test_noalloc_rejects_oversized_batch(gpu):
    vb = veig.VarBatch([4, 5])
    vb.nmat = (1 << 22) + 1   # beyond the static scratch; refused before any launch
    with pytest.raises(veig.VeigError):
        veig.check(veig.lib().veig_zgeev_varsize_noalloc(
            vb.nmat, veig._ptr(vb.dims), veig._ptr(vb.a_off), veig._ptr(vb.e_off), veig._ptr(vb.rA),
            veig._ptr(vb.iA), veig._ptr(vb.rE), veig._ptr(vb.iE), veig._ptr(vb.rV), veig._ptr(vb.iV),
            veig._ptr(vb.info), 30, None, None, veig._stream(None)), "noalloc")

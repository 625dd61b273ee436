"""Helpers for the GPU tests: run the device entry points on numpy inputs
(copied to the GPU in the tiled layout) and bring the results back."""
from __future__ import annotations

import numpy as np
import torch

import veig
from helpers import from_soa, to_soa, vec_from_soa


def gpu_eig(A: np.ndarray, thread_cnt: int = 1, precision: str = "fp64", itmax: int = 30,
            write_schur: bool = False):
    """Returns (E, V, info[, T]) for (N,n,n) complex input via veig_zgeev*."""
    N, n, _ = A.shape
    rA, iA = to_soa(A, thread_cnt)
    dev = torch.device("cuda:0")
    trA = torch.from_numpy(rA).to(dev)
    tiA = torch.from_numpy(iA).to(dev)
    info = torch.full((N,), -1, dtype=torch.int32, device=dev)
    if precision == "mixed":
        berr = torch.empty(N, dtype=torch.float64, device=dev)
        rE, iE, rV, iV = veig.zgeev_mixed(trA, tiA, n, N // thread_cnt, thread_cnt, info=info,
                                          berr=berr, itmax=itmax)
    else:
        rE, iE, rV, iV = veig.zgeev(trA, tiA, n, N // thread_cnt, thread_cnt, info=info,
                                    itmax=itmax, write_schur=write_schur, precision=precision)
    torch.cuda.synchronize()
    E = vec_from_soa(rE.cpu().numpy(), iE.cpu().numpy(), n, thread_cnt)
    V = from_soa(rV.cpu().numpy(), iV.cpu().numpy(), n, thread_cnt)
    out = (E, V, info.cpu().numpy())
    if write_schur:
        out = out + (from_soa(trA.cpu().numpy(), tiA.cpu().numpy(), n, thread_cnt),)
    if precision == "mixed":
        out = out + (berr.cpu().numpy(),)
    return out

"""The fused FP64 and mixed-precision entries inside a captured HIP graph
(torch.cuda.CUDAGraph over the launch stream): every launch they make is a
kernel or a device-to-device copy (no memset, no allocation, no host
synchronisation), so the sequence can be replayed.  The graph is replayed
several times with eager work and synchronisations between the replays, and
every replay is compared with the eager result bit for bit."""
from __future__ import annotations

import pytest
import torch

import helpers as H
import veig

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,precision", [(16, "fp64"), (24, "fp64"), (48, "fp64"), (16, "mixed"), (40, "mixed")])
def test_graph_replays_match_eager(gpu, n, precision):
    dev = torch.device("cuda:0")
    tc, N = 4, 64
    A = H.gaussian(N, n, 900 + n)
    rA_np, iA_np = H.to_soa(A, tc)
    rA, iA = torch.from_numpy(rA_np).to(dev), torch.from_numpy(iA_np).to(dev)
    f64 = dict(dtype=torch.float64, device=dev)
    outs = [torch.empty(N * n, **f64), torch.empty(N * n, **f64), torch.empty(N * n * n, **f64),
            torch.empty(N * n * n, **f64)]
    info = torch.empty(N, dtype=torch.int32, device=dev)
    ws = veig.mixed_workspace(n, N // tc, tc, dev)
    berr = torch.empty(N, **f64)

    def run(stream):
        if precision == "fp64":
            veig.zgeev(rA, iA, n, N // tc, tc, out=tuple(outs), info=info, stream=stream)
        else:
            veig.zgeev_mixed(rA, iA, n, N // tc, tc, out=tuple(outs), info=info, berr=berr, workspace=ws,
                             stream=stream)

    run(torch.cuda.current_stream())
    torch.cuda.synchronize()
    eager = [o.clone() for o in outs] + [info.clone()]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        run(side)  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        run(torch.cuda.current_stream())
    for rep in range(3):
        for o in outs:
            o.fill_(float("nan"))
        info.fill_(-9)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = outs + [info]
        for a, b in zip(eager, got):
            assert torch.equal(a, b), f"replay {rep} differs"
        _ = (rA * 2).sum().item()  # eager work between replays

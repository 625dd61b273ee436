"""Benchmark: batched complex eigendecomposition throughput on one MI355X.

Workload (default): 1,048,576 complex128 16x16 matrices, real and imaginary
parts iid N(0,1) from a fixed seed, in the tiled split-plane layout with tile
64 (block_cnt 16384, thread_cnt 64); one step = one fused eigendecomposition
(eigenvalues + right eigenvectors) of the whole batch, inputs resident in HBM.

Prints ONE JSON line.  `value` is the FP64 path's matrices/s; the same line
carries the mixed-precision path's rate and backward error, the roofline of
the dominant kernel (HIP events on the launch stream), and the reference
OpenMP CPU path timed on a bounded sample of the same inputs.

    python bench.py [--steps K] [--warmup W] [--n 16] [--batch 1048576] [--tile 64]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "vectorized-eigensolver_amd"))
sys.path.insert(0, str(REPO / "tests"))

PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def nominal_flops(n: int, sweeps_per_eig: float = 2.0) -> float:
    """Algorithmic FP64 flops per matrix: gehrd 40/3 n^3, unghr 16/3 n^3,
    trevc 16/3 n^3 (exact sums) + QR 20(2n+3) per bulge step over
    s*n(n-1)/2 steps (complex FMA = 8 flops)."""
    gehrd = sum(16 * n * k + 16 * k * k for k in range(1, n))
    unghr = sum(16 * m * (m - 1) for m in range(1, n))
    trevc = sum(4 * (k - 1) ** 2 + 8 * n * (k - 1) for k in range(2, n + 1))
    qr = 20 * (2 * n + 3) * sweeps_per_eig * n * (n - 1) / 2
    return gehrd + unghr + trevc + qr


def alg_bytes(n: int) -> int:
    """Read A, write V and E (complex128)."""
    return 32 * n * n + 16 * n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=16)
    ap.add_argument("--batch", type=int, default=1 << 20)
    ap.add_argument("--tile", type=int, default=64)
    ap.add_argument("--seed", type=int, default=1002)
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the measurement, run a plain 1 GiB device copy (known bytes) so "
                         "that hardware byte counters can be scaled against it")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the batch path is data parallel over "
              "independent matrices and this build runs on a single MI355X "
              "(see the sharding note in DESIGN.md); use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import veig

    n, N, T = a.n, a.batch, a.tile
    assert N % T == 0
    bc = N // T
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(a.seed)
    rA = torch.randn(N * n * n, dtype=torch.float64, device=dev, generator=g)
    iA = torch.randn(N * n * n, dtype=torch.float64, device=dev, generator=g)
    rE = torch.empty(N * n, dtype=torch.float64, device=dev)
    iE = torch.empty_like(rE)
    rV = torch.empty(N * n * n, dtype=torch.float64, device=dev)
    iV = torch.empty_like(rV)
    info = torch.empty(N, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def step_fp64():
        veig.zgeev(rA, iA, n, bc, T, out=(rE, iE, rV, iV), info=info, stream=stream)

    for _ in range(a.warmup):
        step_fp64()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(a.steps):
        step_fp64()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = ev0.elapsed_time(ev1) / a.steps  # one launch per step
    value = N / (ms_per_step * 1e-3)

    # correctness of the timed outputs on a sample (backward error, info)
    nfail = int((info != 0).sum())
    S = min(N, 4096)
    A_s = _gather(rA, iA, n, T, S)
    E_s = _gather_vec(rE, iE, n, T, S)
    V_s = _gather(rV, iV, n, T, S)
    R = A_s @ V_s - V_s * E_s[:, None, :]
    berr = (torch.linalg.norm(R, dim=(1, 2)) /
            (torch.linalg.norm(A_s, dim=(1, 2)) * torch.linalg.norm(V_s, dim=(1, 2))))
    max_berr = float(berr.max())

    flops = nominal_flops(n)
    achieved_tf = N * flops / (ms_per_step * 1e-3) / 1e12
    roofline = {
        "bound": "mfma", "achieved": round(achieved_tf, 3), "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s", "frac": round(achieved_tf / PEAK_FP64_TFLOPS, 4), "traffic": None,
        "kernel": f"k_zgeev<double,{16 if n > 8 else 8}>",
        "flops_per_matrix": flops,
        "hbm_alg_GBs": round(N * alg_bytes(n) / (ms_per_step * 1e-3) / 1e9, 1),
    }
    tf = Path(REPO / "profiles" / f"traffic_zgeev_fp64_n{n}.json")
    if tf.exists():
        try:
            roofline["traffic"] = json.loads(tf.read_text()).get("bytes_per_launch")
        except Exception:
            pass

    mixed = None
    if not a.no_mixed and hasattr(veig.lib(), "veig_zgeev_mixed"):
        mixed = _bench_mixed(veig, torch, rA, iA, n, bc, T, a, stream, value)

    cpu = None
    if not a.no_cpu:
        cpu = _cpu_baseline(rA, iA, n, T, N, a.cpu_seconds)

    out = {
        "metric": f"eigendecomp matrices/s, n={n} c128: FP64 vs mixed-prec at backward err <=1e-12",
        "value": round(value, 1),
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{N} complex128 {n}x{n} matrices, re/im iid N(0,1) seed {a.seed}, "
                               f"tiled SoA tile {T} (block_cnt {bc}, thread_cnt {T}); "
                               "eigenvalues + right eigenvectors",
                   "batch": N, "n": n, "thread_cnt": T, "block_cnt": bc},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed": mixed,
        "check": {"sample": S, "max_backward_err": max_berr, "unconverged_matrices": nfail},
        "wall_s_timed": round(wall, 3),
    }
    print(json.dumps(out), flush=True)

    if a.calibrate_copy:
        # 2^27 doubles = 1 GiB read + 1 GiB written per copy
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": 1 << 30, "copy_calibration_bytes_written": 1 << 30}),
              flush=True)


def _gather(rX, iX, n, T, S):
    """First S matrices of a tiled-SoA pair of planes as an (S,n,n) complex tensor."""
    import torch
    bcs = S // T
    X = torch.complex(rX[: bcs * T * n * n], iX[: bcs * T * n * n]).view(bcs, n, n, T)  # [b,c,r,t]
    return X.permute(0, 3, 2, 1).reshape(bcs * T, n, n).contiguous()


def _gather_vec(rX, iX, n, T, S):
    import torch
    bcs = S // T
    X = torch.complex(rX[: bcs * T * n], iX[: bcs * T * n]).view(bcs, n, T)
    return X.permute(0, 2, 1).reshape(bcs * T, n).contiguous()


def _bench_mixed(veig, torch, rA, iA, n, bc, T, a, stream, fp64_value):
    N = bc * T
    dev = rA.device
    rE = torch.empty(N * n, dtype=torch.float64, device=dev)
    iE = torch.empty_like(rE)
    rV = torch.empty(N * n * n, dtype=torch.float64, device=dev)
    iV = torch.empty_like(rV)
    info = torch.empty(N, dtype=torch.int32, device=dev)
    berr = torch.empty(N, dtype=torch.float64, device=dev)
    fb = torch.empty(N, dtype=torch.int32, device=dev)
    ws = veig.mixed_workspace(n, bc, T, dev)

    def step():
        veig.zgeev_mixed(rA, iA, n, bc, T, out=(rE, iE, rV, iV), info=info, berr=berr,
                         fallback=fb, tol=1e-12, workspace=ws, stream=stream)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(a.steps):
        step()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / a.steps
    v = N / (ms * 1e-3)
    return {"value": round(v, 1), "unit": "matrices/s", "ms_per_step": round(ms, 4),
            "ratio_vs_fp64": round(v / fp64_value, 3),
            "max_backward_err": float(berr.max()),
            "fp64_fallback_matrices": int(fb.sum()),
            "unconverged_matrices": int((info != 0).sum()),
            "dtype": "c64 reduction/QR + f64 refinement"}


def _cpu_baseline(rA, iA, n, T, N, target_s):
    """The reference's own OpenMP code (oracle/_ref/libref.so, four stages with
    the reflector entries cleared before QR) on a bounded sample of the same
    inputs; falls back to the oracle restatement when the reference build is
    absent."""
    import ctypes
    import numpy as np
    import helpers

    lib = helpers.load_ref(exact=False)
    kind = "reference"
    if lib is None:
        lib = helpers.load_oracle()
        kind = "port"
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def run(S):
        bcs = S // T
        a = rA[: bcs * T * n * n].cpu().numpy().copy()
        b = iA[: bcs * T * n * n].cpu().numpy().copy()
        rE = np.zeros(S * n); iE = np.zeros(S * n)
        rV = np.zeros(S * n * n); iV = np.zeros(S * n * n)
        P = helpers.dptr
        t = time.perf_counter()
        if kind == "reference":
            lib.ref_eigensolver_clean(n, bcs, T, P(a), P(b), P(rE), P(iE), P(rV), P(iV))
        else:
            lib.oracle_eigensolver(n, bcs, T, P(a), P(b), P(rE), P(iE), P(rV), P(iV), 0, None)
        return time.perf_counter() - t

    S0 = max(T, (4096 // T) * T)
    run(S0)  # warm the thread pool
    t0 = run(S0)
    S = int(min(N, max(S0, target_s / max(t0, 1e-6) * S0)))
    S = max(T, (S // T) * T)
    dt = run(S)
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(S / dt, 1), "unit": "matrices/s", "cores": cores, "kind": kind,
            "sample": f"first {S} matrices of the same batch (tile {T}), four stages timed once "
                      f"after a warm-up call, {dt:.1f} s",
            "cpu": cpu_model}


if __name__ == "__main__":
    main()

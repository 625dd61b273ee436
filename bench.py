"""Benchmark: batched complex eigendecomposition throughput on one MI355X.

Default workload (--config c2, the headline): 1,048,576 complex128 16x16
matrices, real and imaginary parts iid N(0,1) from a fixed seed, in the tiled
split-plane layout with tile 64 (block_cnt 16384, thread_cnt 64).  One step =
one eigendecomposition (eigenvalues + right eigenvectors) of the whole batch,
inputs resident in HBM.

Other configurations (--config c1..c5) are described in
vectorized-eigensolver_amd/workloads.py; c4 is the mixed-precision workload
(its `value` is the mixed path's rate), c3 the batch of orders 4..32 run as
one captured graph.

Prints ONE JSON line: `value` (matrices/s), the roofline of the dominant
kernel timed with HIP events on its launch stream, the mixed-precision rate
beside the FP64 one, and the reference's OpenMP CPU path timed on a bounded
sample of the same inputs.
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
METRIC = "eigendecomp matrices/s, n=16 c128: FP64 vs mixed-prec at backward err <=1e-12"


def nominal_flops(n: int, sweeps_per_eig: float = 2.0) -> float:
    """Algorithmic FP64 flops per matrix: gehrd 40/3 n^3, unghr 16/3 n^3,
    trevc 16/3 n^3 (exact sums) + QR 20(2n+3) per bulge step over
    s*n(n-1)/2 steps (complex FMA = 8 flops)."""
    gehrd = sum(16 * n * k + 16 * k * k for k in range(1, n))
    unghr = sum(16 * m * (m - 1) for m in range(1, n))
    trevc = sum(4 * (k - 1) ** 2 + 8 * n * (k - 1) for k in range(2, n + 1))
    qr = 20 * (2 * n + 3) * sweeps_per_eig * n * (n - 1) / 2
    return gehrd + unghr + trevc + qr


def nominal_flops_eig(n: int, sweeps_per_eig: float = 2.0) -> float:
    """Eigenvalues-only QR (no Hessenberg reduction: companion input): the
    bulge step in an active block of order m touches about m + 3 entries of
    each of two rows/columns, 20 flops each, over s (m - 1) steps per
    deflation, m = n .. 2."""
    return sum(20.0 * (m + 3) * sweeps_per_eig * (m - 1) for m in range(2, n + 1))


def alg_bytes(n: int) -> int:
    """Read A, write V and E (complex128)."""
    return 32 * n * n + 16 * n


def progress(msg):
    """Phase boundaries on stderr (the JSON line stays the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, time.perf_counter() - t0


def kernel_sources_sha() -> str:
    """Hash of the kernel sources (csrc/*.hip, *.hpp; the host-only C++ shim
    veig_compat.cpp holds no kernels): traffic figures measured with
    rocprofv3 are reported only for the build they were measured on."""
    import hashlib
    h = hashlib.sha256()
    for f in sorted((REPO / "vectorized-eigensolver_amd" / "csrc").iterdir()):
        if f.suffix in (".hip", ".hpp"):
            h.update(f.name.encode())
            h.update(f.read_bytes())
    return h.hexdigest()[:16]


def roofline(flops_per_launch, bytes_per_launch, kernel_ms, kernel, traffic_key, bound="fp64-valu"):
    """The FP64 kernels run on the vector ALUs (no matrix-core instructions);
    MI355X's FP64 vector peak equals its FP64 matrix peak, 78.6 TFLOP/s."""
    tf = flops_per_launch / (kernel_ms * 1e-3) / 1e12
    r = {"bound": bound, "achieved": round(tf, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
         "frac": round(tf / PEAK_FP64_TFLOPS, 4), "traffic": None, "kernel": kernel,
         "kernel_ms": round(kernel_ms, 4),
         "alg_bytes_per_launch": bytes_per_launch,
         "hbm_alg_GBs": round(bytes_per_launch / (kernel_ms * 1e-3) / 1e9, 1)}
    p = REPO / "profiles" / "traffic.json"
    if p.exists():
        try:
            rec = json.loads(p.read_text()).get(traffic_key, {})
            if rec.get("sources_sha") == kernel_sources_sha():
                r["traffic"] = rec.get("bytes_per_launch")
            else:
                r["traffic_note"] = "no rocprofv3 traffic measured on this build of the kernels"
        except Exception:
            pass
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "roots", "sweep", "layout"])
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=0, help="override the configuration's batch size")
    ap.add_argument("--tile", type=int, default=0, help="thread_cnt of the layout (default: 64, c5: 1)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="target CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the measurement, run a plain 1 GiB device copy (known bytes) so "
                         "that hardware byte counters can be scaled against it")
    a = ap.parse_args()

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the batch path is data parallel over "
              "independent matrices, sharded across GPUs only by splitting the batch "
              "into contiguous tile ranges, which this library leaves to the caller; use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch

    import veig
    import workloads as W

    cfg = dict(W.CONFIGS[a.config])
    if a.batch:
        cfg["batch"] = a.batch
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    if a.config == "c3":
        out = bench_varsize(a, cfg, veig, W, torch, dev, stream)
    elif a.config == "roots":
        out = bench_roots(a, cfg, veig, W, torch, dev, stream)
    elif a.config == "sweep":
        out = bench_sweep(a, cfg, veig, W, torch, dev, stream)
    elif a.config == "layout":
        out = bench_layout(a, cfg, veig, torch, dev, stream)
    else:
        out = bench_uniform(a, cfg, veig, W, torch, dev, stream)
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


def bench_uniform(a, cfg, veig, W, torch, dev, stream):
    n, N = cfg["n"], cfg["batch"]
    T = min(a.tile or cfg.get("tile", 64), N)
    assert N % T == 0
    bc = N // T
    progress(f"{a.config}: generating {N} inputs of order {n}")
    if a.config == "c4":
        rA, iA = W.c4_soa(N, n, T, cfg["seed"], dev)
        workload = (f"{N} complex128 {n}x{n} matrices A = V diag(lam) V^-1, lam iid CN(0,1), cond(V) "
                    f"log-uniform in [1e1, 1e6], seed {cfg['seed']}, tiled SoA tile {T}")
    else:
        rA, iA = W.gaussian_soa(N, n, T, cfg["seed"], dev)
        workload = (f"{N} complex128 {n}x{n} matrices, re/im iid N(0,1) seed {cfg['seed']}, "
                    f"tiled SoA tile {T} (block_cnt {bc}, thread_cnt {T})")
    f64 = dict(dtype=torch.float64, device=dev)
    rE, iE = torch.empty(N * n, **f64), torch.empty(N * n, **f64)
    rV, iV = torch.empty(N * n * n, **f64), torch.empty(N * n * n, **f64)
    info = torch.empty(N, dtype=torch.int32, device=dev)

    def step_fp64():
        veig.zgeev(rA, iA, n, bc, T, out=(rE, iE, rV, iV), info=info, stream=stream)

    progress("timing the FP64 path")
    ms64, wall64 = timed(step_fp64, a.steps, a.warmup, stream)
    v64 = N / (ms64 * 1e-3)
    chk64 = sample_check(rA, iA, rE, iE, rV, iV, n, T, N, info, torch)

    mixed = None
    if not a.no_mixed:
        progress("timing the mixed-precision path")
        mixed = bench_mixed(veig, torch, rA, iA, n, bc, T, a, stream, v64, (rE, iE))

    flops = N * nominal_flops(n)
    nbytes = N * alg_bytes(n)
    if a.config == "c4" and mixed is not None:
        value, ms, dtype = mixed["value"], mixed["ms_per_step"], "c64+f64"
        rf = roofline(flops, nbytes, mixed["fp32_kernel_ms"], f"k_zgeev<float, {nb(n)}, 1, 0>", f"zgeev_fp32_n{n}")
    else:
        value, ms, dtype = v64, ms64, "f64"
        kname = ("k_gehrd64 + k_qr64 (banded QR, half-row Schur vectors)" if n > 32 else
                 f"k_gehrd_c<{nb(n)}> + k_qr_band<{nb(n)}> (banded QR)" if n > 8 else f"k_zgeev<double, {nb(n)}, 1, 0>")
        rf = roofline(flops, nbytes, ms64, kname, f"zgeev_fp64_n{n}")
    rf["flops_per_matrix"] = nominal_flops(n)

    if not a.no_cpu:
        progress("timing the reference CPU path on a sample")
    cpu = None if a.no_cpu else cpu_baseline_uniform(rA, iA, n, T, N, a.cpu_seconds)
    return {
        "metric": METRIC if a.config == "c2" else f"eigendecomp matrices/s, {a.config}: {workload}",
        "value": round(value, 1), "unit": "matrices/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": {"workload": workload + "; eigenvalues + right eigenvectors", "name": a.config,
                   "batch": N, "n": n, "thread_cnt": T, "block_cnt": bc},
        "roofline": rf, "cpu_baseline": cpu,
        "fp64": {"value": round(v64, 1), "ms_per_step": round(ms64, 4), **chk64},
        "mixed": mixed, "wall_s_timed": round(wall64, 3),
    }


def nb(n):
    return 8 if n <= 8 else (16 if n <= 16 else (32 if n <= 32 else 64))


def sample_check(rA, iA, rE, iE, rV, iV, n, T, N, info, torch):
    """Backward error of the timed outputs on the first matrices; info."""
    S = min(N, 4096)
    A_s = gather(rA, iA, n, T, S)
    E_s = gather_vec(rE, iE, n, T, S)
    V_s = gather(rV, iV, n, T, S)
    R = A_s @ V_s - V_s * E_s[:, None, :]
    berr = (torch.linalg.norm(R, dim=(1, 2)) /
            (torch.linalg.norm(A_s, dim=(1, 2)) * torch.linalg.norm(V_s, dim=(1, 2))))
    return {"check_sample": S, "max_backward_err": float(berr.max()),
            "unconverged_matrices": int((info != 0).sum())}


def gather(rX, iX, n, T, S):
    """First S matrices of a tiled-SoA pair of planes as an (S,n,n) complex tensor."""
    import torch
    bcs = max(1, S // T)
    X = torch.complex(rX[: bcs * T * n * n], iX[: bcs * T * n * n]).view(bcs, n, n, T)  # [b,c,r,t]
    return X.permute(0, 3, 2, 1).reshape(bcs * T, n, n).contiguous()


def gather_vec(rX, iX, n, T, S):
    import torch
    bcs = max(1, S // T)
    X = torch.complex(rX[: bcs * T * n], iX[: bcs * T * n]).view(bcs, n, T)
    return X.permute(0, 2, 1).reshape(bcs * T, n).contiguous()


def eig_err_vs_fp64(rE, iE, rE64, iE64, rA, iA, n, T, N, torch):
    """Eigenvalue error of the mixed path against the FP64 path on the whole
    batch: each mixed eigenvalue against the nearest FP64 eigenvalue of its
    matrix, relative to that eigenvalue and to ||A||_F."""
    chunk = max(T, (1 << 16) // T * T)
    worst_rel = worst_nrm = 0.0
    for s0 in range(0, N, chunk):
        m = min(chunk, N - s0)
        E = gather_vec(rE[s0 * n:], iE[s0 * n:], n, T, m)
        E64 = gather_vec(rE64[s0 * n:], iE64[s0 * n:], n, T, m)
        D = torch.abs(E[:, :, None] - E64[:, None, :])
        d, j = D.min(dim=2)
        lam = torch.gather(E64, 1, j)
        A = torch.complex(rA[s0 * n * n: (s0 + m) * n * n], iA[s0 * n * n: (s0 + m) * n * n])
        nrm = torch.linalg.vector_norm(A.view(m // T, n * n, T), dim=1).reshape(-1)  # ||A||_F per matrix
        worst_rel = max(worst_rel, float((d / torch.abs(lam).clamp_min(1e-300)).max()))
        worst_nrm = max(worst_nrm, float((d.max(dim=1).values / nrm).max()))
    return {"max_rel_eig_err_vs_fp64": worst_rel, "max_eig_err_vs_fp64_over_normA": worst_nrm,
            "eig_err_note": "every matrix of the batch; each mixed eigenvalue against the nearest "
                            "eigenvalue of the FP64 path's result"}


def bench_mixed(veig, torch, rA, iA, n, bc, T, a, stream, fp64_value, fp64_E=None):
    N = bc * T
    dev = rA.device
    f64 = dict(dtype=torch.float64, device=dev)
    rE, iE = torch.empty(N * n, **f64), torch.empty(N * n, **f64)
    rV, iV = torch.empty(N * n * n, **f64), torch.empty(N * n * n, **f64)
    info = torch.empty(N, dtype=torch.int32, device=dev)
    berr = torch.empty(N, **f64)
    fb = torch.empty(N, dtype=torch.int32, device=dev)
    ws = veig.mixed_workspace(n, bc, T, dev)

    def step():
        veig.zgeev_mixed(rA, iA, n, bc, T, out=(rE, iE, rV, iV), info=info, berr=berr,
                         fallback=fb, tol=1e-12, workspace=ws, stream=stream)

    ms, _ = timed(step, a.steps, a.warmup, stream)
    v = N / (ms * 1e-3)
    # the complex64 solve alone (the dominant kernel of the mixed path)
    ms32, _ = timed(lambda: veig.zgeev(rA, iA, n, bc, T, out=(rE, iE, rV, iV), info=info,
                                       precision="fp32", stream=stream), a.steps, 1, stream)
    step()
    torch.cuda.synchronize()
    eerr = {} if fp64_E is None else eig_err_vs_fp64(rE, iE, fp64_E[0], fp64_E[1], rA, iA, n, T, N, torch)
    return {"value": round(v, 1), "unit": "matrices/s", "ms_per_step": round(ms, 4),
            "ratio_vs_fp64": round(v / fp64_value, 3),
            "max_backward_err": float(berr.max()), **eerr,
            "fp64_fallback_matrices": int(fb.sum()),
            "unconverged_matrices": int((info != 0).sum()),
            "fp32_kernel_ms": round(ms32, 4),
            "dtype": "c64 reduction/QR/eigenvectors + f64 Newton refinement (tol 1e-12)"}


def _cpu_lib():
    """The reference's own code (oracle/_ref), built for the widest x86-64
    level this CPU runs (x86-64-v4 = AVX-512 where present, else v3), with
    its OpenMP threads bound close to the cores; the oracle restatement where
    the reference build is absent.  Returns (lib, kind, threads)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import helpers
    lib = helpers.load_ref(exact=False, isa="v4" if _cpu_has_v4() else "v3")
    kind = "reference"
    if lib is None:
        lib, kind = helpers.load_oracle(), "port"
    threads = 0
    if kind == "reference" and hasattr(lib, "ref_get_threads"):
        threads = int(lib.ref_get_threads())
    elif kind == "port" and hasattr(lib, "oracle_get_threads"):
        threads = int(lib.oracle_get_threads())
    return lib, kind, threads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def _cpu_has_v4() -> bool:
    try:
        flags = next(line for line in open("/proc/cpuinfo") if line.startswith("flags")).split()
    except (OSError, StopIteration):
        return False
    return all(f in flags for f in ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"))


def _median_runs(run, S, reps=5):
    """Median wall time of `reps` timed runs of run(S), after one warm-up."""
    import statistics
    run(S)
    return statistics.median(run(S) for _ in range(reps))


def _run_cpu(lib, kind, n, tc, rA_np, iA_np):
    import numpy as np
    import helpers
    S = rA_np.size // (n * n)
    rE = np.zeros(S * n); iE = np.zeros(S * n)
    rV = np.zeros(S * n * n); iV = np.zeros(S * n * n)
    P = helpers.dptr
    t = time.perf_counter()
    if kind == "reference":
        lib.ref_eigensolver_clean(n, S // tc, tc, P(rA_np), P(iA_np), P(rE), P(iE), P(rV), P(iV))
    else:
        lib.oracle_eigensolver(n, S // tc, tc, P(rA_np), P(iA_np), P(rE), P(iE), P(rV), P(iV), 0, None)
    return time.perf_counter() - t


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cpu_baseline_uniform(rA, iA, n, T, N, target_s):
    """The reference's own OpenMP code (oracle/_ref/libref*.so: its four
    stages with the reflector entries cleared before QR) on a bounded sample
    of the same inputs, sized so that one run takes about target_s / 5; the
    median of five runs.  The oracle restatement where the reference build is
    absent."""
    lib, kind, cores = _cpu_lib()

    def run(S):
        bcs = S // T
        a = rA[: bcs * T * n * n].cpu().numpy().copy()
        b = iA[: bcs * T * n * n].cpu().numpy().copy()
        return _run_cpu(lib, kind, n, T, a, b)

    S0 = max(T, (min(N, 4096) // T) * T)
    run(S0)  # warm the thread pool
    t0 = run(S0)
    S = int(min(N, max(S0, target_s / 5 / max(t0, 1e-6) * S0)))
    S = max(T, (S // T) * T)
    dt = _median_runs(run, S)
    return {"value": round(S / dt, 1), "unit": "matrices/s", "cores": cores, "kind": kind,
            "sample": f"first {S} matrices of the same batch (tile {T}), four stages, median of 5 runs "
                      f"after a warm-up run, {dt:.2f} s per run; OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}",
            "cpu": _cpu_model()}


def bench_roots(a, cfg, veig, W, torch, dev, stream):
    """Polynomial roots through companion matrices (veig_zlahqr_roots):
    roots only (the eigenvalues-only iteration) is `value`; the Schur form
    and vectors, which the reference's kernel also returns, are timed beside."""
    import numpy as np
    n, N = cfg["n"], cfg["batch"]
    T = min(a.tile or 64, N)
    bc = N // T
    progress(f"roots: generating {N} polynomials of degree {n}")
    rC, iC = W.coeff_soa(N, n, T, cfg["seed"], dev)
    f64 = dict(dtype=torch.float64, device=dev)
    rE, iE = torch.empty(N * n, **f64), torch.empty(N * n, **f64)
    info = torch.empty(N, dtype=torch.int32, device=dev)
    L = veig.lib()

    def step_roots():
        veig.check(L.veig_zlahqr_roots(n, bc, T, veig._ptr(rC), veig._ptr(iC), veig._ptr(rE), veig._ptr(iE),
                                       None, None, None, None, veig._ptr(info), 30, veig._stream(stream)),
                   "veig_zlahqr_roots")

    progress("timing roots only")
    ms, wall = timed(step_roots, a.steps, a.warmup, stream)
    value = N / (ms * 1e-3)
    # accuracy on a sample: against numpy's roots of the same coefficients
    S = 256
    cr = rC[: (S // T) * T * (n + 1) if S >= T else T * (n + 1)].cpu().numpy()
    ci = iC[: cr.size].cpu().numpy()
    Sm = cr.size // (n + 1)
    import helpers
    Cf = helpers.vec_from_soa(cr, ci, n + 1, T)
    E = helpers.vec_from_soa(rE[: Sm * n].cpu().numpy(), iE[: Sm * n].cpu().numpy(), n, T)
    worst = 0.0
    for k in range(Sm):
        want = np.roots(Cf[k, ::-1])
        _, err = helpers.match_eigenvalues(want[None], E[k][None])
        worst = max(worst, float(err[0] / max(1.0, np.abs(want).max())))
    schur = None
    if not a.no_mixed:
        mats = [torch.empty(N * n * n, **f64) for _ in range(4)]

        def step_schur():
            veig.check(L.veig_zlahqr_roots(n, bc, T, veig._ptr(rC), veig._ptr(iC), veig._ptr(rE),
                                           veig._ptr(iE), *[veig._ptr(m) for m in mats], veig._ptr(info), 30,
                                           veig._stream(stream)), "veig_zlahqr_roots")

        progress("timing roots with Schur form and vectors")
        ms2, _ = timed(step_schur, a.steps, a.warmup, stream)
        schur = {"value": round(N / (ms2 * 1e-3), 1), "ms_per_step": round(ms2, 4)}
        del mats
    flops = N * nominal_flops_eig(n)
    nbytes = N * (16 * (n + 1) + 16 * n)
    rf = roofline(flops, nbytes, ms, f"k_eig<{nb(n)}, 1, SRC_POLY, false>", f"roots_n{n}")
    rf["flops_per_matrix"] = nominal_flops_eig(n)
    cpu = None
    if not a.no_cpu:
        progress("timing the reference CPU roots kernel on a sample")
        cpu = cpu_baseline_roots(rC, iC, n, T, N, a.cpu_seconds)
    return {
        "metric": f"polynomial roots/s, degree {n}, complex128 (companion matrix QR)",
        "value": round(value, 1), "unit": "polynomials/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"{N} polynomials c_0..c_{n}, re/im iid N(0,1) seed {cfg['seed']}, "
                               f"vector layout tile {T}; roots only", "name": "roots", "batch": N, "n": n},
        "roofline": rf, "cpu_baseline": cpu,
        "roots": {"value": round(value, 1), "max_rel_root_err_vs_numpy": worst, "sample": Sm,
                  "unconverged": int((info != 0).sum())},
        "roots_with_schur": schur, "wall_s_timed": round(wall, 3)}


def bench_sweep(a, cfg, veig, W, torch, dev, stream):
    """The reference driver's sweep (zlahqr.cpp:196-252: 13,312 matrices per
    order, one batch per order, orders 10..79): FP64 matrices/s per order;
    `value` = all matrices / all time."""
    N, T = cfg["batch"], 64
    rates, total_ms = {}, 0.0
    for n in range(cfg["nmin"], cfg["nmax"] + 1):
        rA, iA = W.gaussian_soa(N, n, T, cfg["seed"] + n, dev)
        f64 = dict(dtype=torch.float64, device=dev)
        rE, iE = torch.empty(N * n, **f64), torch.empty(N * n, **f64)
        rV, iV = torch.empty(N * n * n, **f64), torch.empty(N * n * n, **f64)
        info = torch.empty(N, dtype=torch.int32, device=dev)
        ms, _ = timed(lambda: veig.zgeev(rA, iA, n, N // T, T, out=(rE, iE, rV, iV), info=info, stream=stream),
                      a.steps, a.warmup, stream)
        rates[n] = round(N / (ms * 1e-3), 1)
        total_ms += ms
        progress(f"sweep n={n}: {rates[n]:.0f} matrices/s")
    n_dims = len(rates)
    value = N * n_dims / (total_ms * 1e-3)
    cpu = None
    if not a.no_cpu:
        # the reference's own OpenMP code on the same 13,312-matrix batches at
        # four of the orders (whole batch, timed once after a warm-up call)
        lib, kind, cores = _cpu_lib()
        per = {}
        for n in (16, 32, 48, 64, 79):
            if not cfg["nmin"] <= n <= cfg["nmax"]:
                continue
            rA, iA = W.gaussian_soa(N, n, T, cfg["seed"] + n, dev)
            a_np, b_np = rA.cpu().numpy(), iA.cpu().numpy()
            _run_cpu(lib, kind, n, T, a_np[: T * n * n].copy(), b_np[: T * n * n].copy())
            dt = _run_cpu(lib, kind, n, T, a_np.copy(), b_np.copy())
            per[n] = round(N / dt, 1)
            progress(f"sweep CPU reference n={n}: {per[n]:.0f} matrices/s")
        cpu = {"per_order": per, "unit": "matrices/s", "kind": kind, "cores": cores,
               "sample": f"the full {N}-matrix batch at orders {sorted(per)}", "cpu": _cpu_model()}
    return {"metric": f"eigendecomp matrices/s, orders {cfg['nmin']}..{cfg['nmax']}, {N} matrices per order",
            "value": round(value, 1), "unit": "matrices/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
            "ms_per_step": round(total_ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{N} complex128 Gaussian matrices per order, tile {T}, one launch per order "
                                   f"(the reference driver's sweep sizes)", "name": "sweep", "batch": N},
            "roofline": None, "cpu_baseline": cpu, "per_order": rates}


def bench_layout(a, cfg, veig, torch, dev, stream):
    """The device layout converters (the reference's rearrange_matrices,
    arrange_matrices_aos, rep_matrix_soa, rep_vector_soa; utils.cpp) on one
    float64 plane of the c2 batch: HBM bytes moved per second (a read and a
    write of every element; the replicators read one matrix or vector)."""
    n, N, T = cfg["n"], cfg["batch"], 64
    bc = N // T
    na, ne = N * n * n, N * n
    f64 = dict(dtype=torch.float64, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(cfg["seed"])
    src = torch.randn(na, generator=g, **f64)
    dst = torch.empty(na, **f64)
    one = torch.randn(n * n, generator=g, **f64)
    res = {}
    for kind, s_, d_, nbytes in (("soa", src, dst, 16 * na), ("aos", src, dst, 16 * na),
                                 ("repm", one, dst, 8 * na), ("repv", one[:n], dst[:ne], 8 * ne)):
        ms, _ = timed(lambda: veig.layout(kind, n, bc, T, s_, d_, stream=stream), a.steps, a.warmup, stream)
        res[kind] = {"ms": round(ms, 4), "GB/s": round(nbytes / (ms * 1e-3) / 1e9, 1), "bytes": nbytes}
        progress(f"layout {kind}: {res[kind]['GB/s']} GB/s")
    # parity: the AoS -> SoA -> AoS round trip is the identity
    veig.layout("soa", n, bc, T, src, dst, stream=stream)
    back = torch.empty_like(src)
    veig.layout("aos", n, bc, T, dst, back, stream=stream)
    torch.cuda.synchronize()
    exact = bool(torch.equal(back, src))
    v = res["soa"]["GB/s"]
    return {"metric": "layout conversion GB/s (AoS -> tiled SoA, one float64 plane)", "value": v, "unit": "GB/s",
            "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": res["soa"]["ms"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": f"{N} order-{n} matrices, one float64 plane ({8 * na} bytes), tile {T}",
                       "name": "layout", "batch": N, "n": n},
            "roofline": {"bound": "hbm", "achieved": v, "peak": 8000.0, "unit": "GB/s", "frac": round(v / 8000.0, 4),
                         "traffic": None, "kernel": "k_transpose (64 x 64 LDS tiles)"},
            "cpu_baseline": None, "kinds": res, "round_trip_exact": exact}


def cpu_baseline_roots(rC, iC, n, T, N, target_s):
    """The reference's roots kernel (oracle/_ref/libref.so; companion + Schur
    form and vectors, OpenMP) on a bounded sample; the oracle restatement
    where the reference build is absent."""
    import numpy as np
    import helpers
    lib, kind, cores = _cpu_lib()

    def run(S):
        bcs = S // T
        c = rC[: bcs * T * (n + 1)].cpu().numpy().copy()
        d = iC[: bcs * T * (n + 1)].cpu().numpy().copy()
        out = [np.zeros(S * n), np.zeros(S * n)] + [np.zeros(S * n * n) for _ in range(4)]
        P = helpers.dptr
        t = time.perf_counter()
        if kind == "reference":
            lib.ref_zlahqr_roots(n, bcs, T, P(c), P(d), *[P(o) for o in out])
        else:
            lib.oracle_zlahqr_roots(n, bcs, T, P(c), P(d), *[P(o) for o in out], 0, None)
        return time.perf_counter() - t

    S0 = max(T, (min(N, 4096) // T) * T)
    run(S0)
    t0 = run(S0)
    S = int(min(N, max(S0, target_s / 5 / max(t0, 1e-6) * S0)))
    S = max(T, (S // T) * T)
    dt = _median_runs(run, S)
    return {"value": round(S / dt, 1), "unit": "polynomials/s", "cores": cores, "kind": kind,
            "sample": f"first {S} polynomials of the same batch (tile {T}), the reference's zlahqr roots "
                      f"kernel (Schur form and vectors included), median of 5 runs, {dt:.2f} s per run",
            "cpu": _cpu_model()}


def bench_varsize(a, cfg, veig, W, torch, dev, stream):
    N = cfg["batch"]
    progress(f"c3: generating {N} inputs")
    dims, kind = W.c3_dims(N, cfg["nmin"], cfg["nmax"], cfg["seed"], dev)
    vb = veig.VarBatch(dims.cpu(), device=dev)
    W.fill_c3(vb, dims, kind, cfg["seed"], dev)
    torch.cuda.synchronize()
    progress("autotuning kernel variant per order and grids (untimed)")
    variants, grids = vb.autotune()
    W.fill_c3(vb, dims, kind, cfg["seed"], dev)  # the tuner overwrote the outputs only; refill anyway
    torch.cuda.synchronize()
    g = vb.graph()
    progress("timing the captured graph")
    ms, wall = timed(lambda: g.launch(stream), a.steps, a.warmup, stream)
    value = N / (ms * 1e-3)
    u, c = torch.unique(dims.cpu(), return_counts=True)
    flops = float(sum(nominal_flops(int(n)) * int(k) for n, k in zip(u, c)))
    nbytes = int(sum(alg_bytes(int(n)) * int(k) for n, k in zip(u, c)))
    rf = roofline(flops, nbytes, ms, "graph: bin kernels + per-variant kernels (k_zgeev_var<8,P>, "
                  "k_zgeev_var<16,2>, k_reduce_var<NB> + k_qr_var<NB> for NB = 16, 32)", "varsize_c3")
    rf["note"] = "whole graph timed (binning + three persistent bin kernels)"
    info = vb.info
    # backward error on a sample of every order
    chk = varsize_check(vb, dims, torch)
    if not a.no_cpu:
        progress("timing the reference CPU path on a sample")
    cpu = None if a.no_cpu else cpu_baseline_varsize(vb, dims, a.cpu_seconds)
    workload = (f"{N} complex128 matrices, n uniform in [{cfg['nmin']},{cfg['nmax']}]: 50% Gaussian, "
                f"25% Hermitian, 25% V diag(lam) V^-1 (clusters 1e-6 apart, cond(V)=1e2), seed "
                f"{cfg['seed']}, per-matrix column-major with offsets; one captured graph per step")
    return {
        "metric": f"eigendecomp matrices/s, c3: {workload}", "value": round(value, 1),
        "unit": "matrices/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": workload, "name": "c3", "batch": N,
                   "autotuned": {"variant_of_n": variants[1:], "grid_per_variant": grids,
                                 "variants": "0: 8 lanes/matrix (width 8), 1: 4 (width 8, 2 per lane), "
                                             "2: 16 (width 16), 3: 8 (width 16, 2 per lane), 4: 32"}},
        "roofline": rf, "cpu_baseline": cpu, "mixed": None,
        "fp64": {"value": round(value, 1), "ms_per_step": round(ms, 4),
                 "unconverged_matrices": int((info > 0).sum()), **chk},
        "wall_s_timed": round(wall, 3),
    }


def varsize_check(vb, dims, torch):
    worst = 0.0
    a_off, e_off = vb.a_off, vb.e_off
    for n in torch.unique(dims).tolist():
        sel = torch.nonzero(dims == n).flatten()[:256]
        idx = a_off[sel].unsqueeze(1) + torch.arange(n * n, device=dims.device).unsqueeze(0)
        A = torch.complex(vb.rA[idx], vb.iA[idx]).view(-1, n, n).transpose(-2, -1)
        V = torch.complex(vb.rV[idx], vb.iV[idx]).view(-1, n, n).transpose(-2, -1)
        eidx = e_off[sel].unsqueeze(1) + torch.arange(n, device=dims.device).unsqueeze(0)
        E = torch.complex(vb.rE[eidx], vb.iE[eidx])
        R = A @ V - V * E[:, None, :]
        be = torch.linalg.norm(R, dim=(1, 2)) / (torch.linalg.norm(A, dim=(1, 2)) * torch.linalg.norm(V, dim=(1, 2)))
        worst = max(worst, float(be.max()))
    return {"check_sample": "256 per order", "max_backward_err": worst}


def cpu_baseline_varsize(vb, dims, target_s):
    """The reference's code on a sample of the batch: it only takes
    equal-size batches, so the sample is grouped by order (thread_cnt 1)."""
    import numpy as np
    lib, kind, cores = _cpu_lib()
    d = dims.cpu().numpy()
    a_off = vb.a_off.cpu().numpy()
    rA = vb.rA.cpu().numpy()
    iA = vb.iA.cpu().numpy()

    def run(S):
        tot = 0.0
        for n in np.unique(d[:S]):
            sel = np.nonzero(d[:S] == n)[0]
            idx = (a_off[sel][:, None] + np.arange(n * n)[None, :]).reshape(-1)
            tot += _run_cpu(lib, kind, int(n), 1, np.ascontiguousarray(rA[idx]), np.ascontiguousarray(iA[idx]))
        return tot

    S0 = min(len(d), 4096)
    run(S0)
    t0 = run(S0)
    S = int(min(len(d), max(S0, target_s / 5 / max(t0, 1e-6) * S0)))
    dt = _median_runs(run, S)
    return {"value": round(S / dt, 1), "unit": "matrices/s", "cores": cores, "kind": kind,
            "sample": f"first {S} matrices of the batch, grouped by order (thread_cnt 1), median of 5 runs, "
                      f"{dt:.2f} s per run",
            "cpu": _cpu_model()}


if __name__ == "__main__":
    main()

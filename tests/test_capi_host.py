"""Host-side tests (no GPU): the C-ABI library loads and exports every
function include/veig.h declares; argument validation answers without
touching a device; the python_interface mirror's layout helpers follow the
reference's layout functions (utils.cpp:85-157, python_interface.pyx:92-96);
bench.py's work model and its one-GPU contract."""
from __future__ import annotations

import ctypes
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import helpers as H
import veig

REPO = Path(__file__).resolve().parent.parent


def test_library_exports_header():
    L = veig.lib()
    names = veig.header_functions()
    assert "veig_zgeev" in names and "veig_host_eigensolver" in names
    missing = [f for f in names if not hasattr(L, f)]
    assert not missing, f"libveig.so lacks {missing}"
    # every declared function has a ctypes signature in the binding
    assert not [f for f in names if f not in veig._SIGS]
    assert L.veig_version().startswith(b"veig")


def test_argument_validation_without_device():
    L = veig.lib()
    z = [None] * 6
    assert L.veig_zgeev(-1, 1, 1, *z, None, 30, 0, None) == 1
    assert L.veig_zgeev(8, 1, 0, *z, None, 30, 0, None) == 1
    assert L.veig_zgeev(97, 1, 1, *z, None, 30, 0, None) == 2
    assert b"exceeds 96" in L.veig_last_error()
    assert L.veig_zgehrd(97, 1, 1, None, None, None, None, None) == 2   # stage entries: n <= 96
    assert L.veig_zgehrd(33, 1, 1, None, None, None, None, None) == 1   # ... then the NULL pointers
    assert L.veig_zgeev_fp32(65, 1, 1, *z, None, 30, None) == 2         # complex64: n <= 64
    assert L.veig_zgeev(8, 2, 2, *z, None, 30, 0, None) == 1
    assert b"NULL pointer" in L.veig_last_error()
    # empty batches are a successful no-op (nothing to do, no device needed)
    assert L.veig_zgeev(8, 0, 4, *z, None, 30, 0, None) == 0
    assert L.veig_zgeev(0, 5, 4, *z, None, 30, 0, None) == 0
    assert L.veig_status_string(2) == b"unsupported order"
    with pytest.raises(veig.VeigError, match="invalid argument"):
        veig.check(L.veig_zgehrd(-3, 1, 1, None, None, None, None, None), "veig_zgehrd")


def test_roots_and_eigvals_validation_without_device():
    L = veig.lib()
    d = [ctypes.c_void_p(8)] * 4  # never dereferenced: validation fails first
    # Schur outputs: all four or none
    st = L.veig_zlahqr_roots(6, 1, 1, *d, ctypes.c_void_p(8), None, None, None, None, 30, None)
    assert st == 1 and b"all four" in L.veig_last_error()
    assert L.veig_zlahqr_roots(65, 1, 1, *d, None, None, None, None, None, 30, None) == 2
    assert L.veig_eigvals(8, 1, 1, None, None, None, None, None, 30, None) == 1
    assert L.veig_zlahqr_eig(0, 3, 2, None, None, None, None, None, 30, None) == 0


def test_mixed_workspace_size():
    L = veig.lib()
    n, bc, tc = 16, 16384, 64
    N = bc * tc
    size = L.veig_mixed_workspace_size(n, bc, tc)
    assert size >= 2 * N * n * n * 4 + 2 * N * 4
    assert size % 256 == 0
    assert L.veig_mixed_workspace_size(-1, 1, 1) == 0


def test_layout_helpers_match_reference_layout():
    import python_interface as pi
    n, bc, tc = 5, 3, 4
    N = bc * tc
    A = H.gaussian(N, n, 11)
    # AoS: one column-major matrix after another
    aos = np.ascontiguousarray(np.transpose(A.real, (0, 2, 1))).reshape(-1)
    soa = np.zeros_like(aos)
    pi.arrangeMatrices(n, bc, tc, aos, soa)
    rA, _ = H.to_soa(A, tc)
    assert np.array_equal(soa, rA)
    back = np.zeros_like(aos)
    pi.arrangeMatricesAoS(n, bc, tc, soa, back)
    assert np.array_equal(back, aos)
    # repMatrixSoA / extractMatrix round trip (python_interface.pyx:92-93)
    m = np.arange(n * n, dtype=float)
    out = np.zeros(n * n * N)
    pi.my_repMatrix(n, bc, tc, m, out)
    for t in range(tc):
        assert np.array_equal(pi.extractMatrix(n, tc, out[t:]), m.reshape(n, n).T)
    v = np.arange(n, dtype=float)
    outv = np.zeros(n * N)
    pi.my_repVector(n, bc, tc, v, outv)
    assert np.array_equal(pi.extractVector(n, tc, outv), v)
    assert np.array_equal(H.vec_from_soa(outv, np.zeros_like(outv), n, tc).real, np.tile(v, (N, 1)))


def test_python_interface_names():
    import python_interface as pi
    # the names of the reference's Cython module (python_interface.pyx:31-220)
    for name in ("my_repMatrix", "arrangeMatrices", "my_zlahqr", "my_backsubs", "my_repVector",
                 "my_ztrevec", "my_zgemv", "my_zgehrd", "my_zunghr", "my_hessenberg",
                 "extractMatrix", "extractVector", "test_backsubs", "test_zgehrd",
                 "test_hessenberg", "test_zgemv", "test_zlahqr", "test_eigenvectors"):
        assert callable(getattr(pi, name)), name
    with pytest.raises(TypeError):
        pi.my_zgehrd(4, 1, 1, np.zeros(16, np.float32), np.zeros(16), np.zeros(4), np.zeros(4))


def test_bench_flop_model():
    sys.path.insert(0, str(REPO))
    import bench
    # leading terms 40/3, 16/3, 16/3 n^3 and 20(2n+3) s n(n-1)/2 (s = 2)
    assert bench.nominal_flops(16) == pytest.approx(256800, rel=0.01)
    assert bench.alg_bytes(16) == 32 * 256 + 16 * 16
    big = bench.nominal_flops(256)
    assert big / 256 ** 3 == pytest.approx(40 / 3 + 16 / 3 + 16 / 3 + 40, rel=0.03)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr

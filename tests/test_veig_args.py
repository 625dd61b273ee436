"""Argument checks of the device-tensor binding (veig.py), on the CPU.

The kernels trust the geometry: an undersized or mistyped plane would be read
or written out of bounds on the GPU.  The binding therefore refuses such
tensors before anything reaches the library; these tests give it CPU tensors,
so a correctly shaped call ends at the "needs GPU tensors" check and nothing
is launched.
"""
import pytest
import torch

import veig

n, bc, tc = 5, 3, 4
nm = bc * tc


def planes(na=nm * n * n, ne=nm * n, dtype=torch.float64):
    z = lambda k: torch.zeros(k, dtype=dtype)  # noqa: E731
    return z(na), z(na), (z(ne), z(ne), z(na), z(na))


def test_zgeev_well_formed_reaches_device_check():
    rA, iA, out = planes()
    with pytest.raises(veig.VeigError, match="GPU tensors"):
        veig.zgeev(rA, iA, n, bc, tc, out=out, info=torch.zeros(nm, dtype=torch.int32))


@pytest.mark.parametrize("which", ["rA", "iA", "rE", "iE", "rV", "iV"])
def test_zgeev_undersized_plane(which):
    rA, iA, (rE, iE, rV, iV) = planes()
    t = dict(rA=rA, iA=iA, rE=rE, iE=iE, rV=rV, iV=iV)
    t[which] = t[which][:-1].clone()
    with pytest.raises(veig.VeigError, match=f"{which}: .* elements"):
        veig.zgeev(t["rA"], t["iA"], n, bc, tc, out=(t["rE"], t["iE"], t["rV"], t["iV"]))


def test_zgeev_wrong_dtypes():
    rA, iA, out = planes()
    with pytest.raises(veig.VeigError, match="rA: dtype"):
        veig.zgeev(rA.to(torch.int64), iA, n, bc, tc, out=out)
    with pytest.raises(veig.VeigError, match="rA: dtype"):
        veig.zgeev(rA.float(), iA, n, bc, tc, out=out)
    with pytest.raises(veig.VeigError, match="info: dtype"):
        veig.zgeev(rA, iA, n, bc, tc, out=out, info=torch.zeros(nm, dtype=torch.int64))
    with pytest.raises(veig.VeigError, match="info: .* elements"):
        veig.zgeev(rA, iA, n, bc, tc, out=out, info=torch.zeros(nm - 1, dtype=torch.int32))


def test_geometry_rejected():
    rA, iA, out = planes()
    with pytest.raises(veig.VeigError, match="geometry"):
        veig.zgeev(rA, iA, -1, bc, tc, out=out)
    with pytest.raises(veig.VeigError, match="geometry"):
        veig.zgeev(rA, iA, n, bc, 0, out=out)


def test_mixed_workspace_and_outputs():
    rA, iA, out = planes()
    need = int(veig.lib().veig_mixed_workspace_size(n, bc, tc))
    ok = torch.zeros(need, dtype=torch.uint8)
    with pytest.raises(veig.VeigError, match="workspace: dtype"):
        veig.zgeev_mixed(rA, iA, n, bc, tc, out=out, workspace=torch.zeros(need, dtype=torch.float64))
    with pytest.raises(veig.VeigError, match="workspace: .* elements"):
        veig.zgeev_mixed(rA, iA, n, bc, tc, out=out, workspace=ok[:need - 1].clone())
    with pytest.raises(veig.VeigError, match="berr: dtype"):
        veig.zgeev_mixed(rA, iA, n, bc, tc, out=out, workspace=ok, berr=torch.zeros(nm, dtype=torch.float32))
    with pytest.raises(veig.VeigError, match="fallback: .* elements"):
        veig.zgeev_mixed(rA, iA, n, bc, tc, out=out, workspace=ok,
                         fallback=torch.zeros(nm - 1, dtype=torch.int32))
    with pytest.raises(veig.VeigError, match="GPU tensors"):
        veig.zgeev_mixed(rA, iA, n, bc, tc, out=out, workspace=ok)


def test_stage_and_layout_sizes():
    m = torch.zeros(nm * n * n, dtype=torch.float64)
    v = torch.zeros(nm * n, dtype=torch.float64)
    with pytest.raises(veig.VeigError, match="rtau: .* elements"):
        veig.zgehrd(n, bc, tc, m, m.clone(), v[:-1].clone(), v.clone())
    with pytest.raises(veig.VeigError, match="GPU tensors"):
        veig.zgehrd(n, bc, tc, m, m.clone(), v, v.clone())
    with pytest.raises(veig.VeigError, match="rZ: .* elements"):
        veig.zlahqr(n, bc, tc, v, v.clone(), m, m.clone(), m[:-1].clone(), m.clone())
    with pytest.raises(veig.VeigError, match="dst: .* elements"):
        veig.layout("soa", n, bc, tc, m, m[:-1].clone())
    with pytest.raises(veig.VeigError, match="src: .* elements"):
        veig.layout("repm", n, bc, tc, torch.zeros(n * n - 1, dtype=torch.float64), m)
    with pytest.raises(veig.VeigError, match="dst: .* elements"):
        veig.layout("repv", n, bc, tc, torch.zeros(n, dtype=torch.float64), v[:-1].clone())
    with pytest.raises(veig.VeigError, match="GPU tensors"):
        veig.layout("repv", n, bc, tc, torch.zeros(n, dtype=torch.float64), v)


def test_roots_coefficients_size():
    c = torch.zeros(nm * (n + 1), dtype=torch.float64)
    with pytest.raises(veig.VeigError, match="rC: .* elements"):
        veig.roots(c[:-1].clone(), c.clone(), n, bc, tc)


def test_varsize_variant_table_dtype():
    vb = veig.VarBatch([3, 7, 12], device="cpu")
    vb.variant_table = torch.zeros(33, dtype=torch.int32)
    with pytest.raises(veig.VeigError, match="variant_table: dtype"):
        vb.run()
    vb.variant_table = torch.zeros(32, dtype=torch.int8)
    with pytest.raises(veig.VeigError, match="variant_table: .* elements"):
        vb.run()
    vb.variant_table = None
    with pytest.raises(veig.VeigError, match="GPU tensors"):
        vb.run()

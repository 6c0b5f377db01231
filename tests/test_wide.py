"""GPU parity of the 256 x 128 split-BF16 kernel (k_gett_wide) against the
128 x 128 kernel (k_gett_lp) and against binary64 products, and the
split-F16 form at the same shapes.

Split-BF16 products large enough to give every CU a 256 x 128 tile run the
wide kernel by default; TAPP_AMD_ATTR_TILES = 2 runs the same call on the
128 x 128 kernel.  Cases cover the four vector load layouts (A with unit M
or unit K stride, B with unit N or unit K stride), edge tiles in M and N
and a K that is not a multiple of the 32-deep K tile.

  * split-BF16 x3: both kernels run the same products in the same order per
    accumulator, so D is bitwise equal; and within 1e-5 normwise of the
    binary64 product (torch float64 on the device, the operands' exact
    values).
  * split-F16 x3 runs k_gett_lp at these shapes: within 1e-6 normwise of the
    binary64 product.  A case with operands far below binary16's range
    takes the rescaled launch, one with a NaN the FP32 fallback, which is
    bitwise the plain FP32 kernel's result.
"""
import numpy as np
import pytest

import tapp_cases as T

tc = T.tc
pytestmark = pytest.mark.gpu

ATTR_TILES = tc.ATTR_TILES


def _mk(torch, rows, cols, contig_rows, seed, scale=1.0):
    """rows x cols FP32 matrix on the device with unit stride along rows
    (contig_rows) or along cols; returns (tensor view [rows, cols],
    strides in elements)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    if contig_rows:
        base = torch.rand((cols, rows), generator=g, device="cuda", dtype=torch.float32) * 2 - 1
        t = base.t()
        strides = [1, rows]
    else:
        base = torch.rand((rows, cols), generator=g, device="cuda", dtype=torch.float32) * 2 - 1
        t = base
        strides = [cols, 1]
    if scale != 1.0:
        base.mul_(scale)
    return base, t, strides


def _run(lib, torch, A, sa, B, sb, M, N, K, split, tiles):
    """D[m, n] = sum_k A[m, k] B[k, n] through TAPP_execute_product; returns
    (D [M, N] as float64 numpy, launched kernel name)."""
    rc, ex = lib.create_executor()
    assert rc == 0
    assert lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, split) == 0
    assert lib.attr_set_int(ex, ATTR_TILES, tiles) == 0
    m, n, k = (ord(x) for x in "mnk")
    infos = [lib.create_info(tc.F32, [M, K], sa)[1], lib.create_info(tc.F32, [K, N], sb)[1],
             lib.create_info(tc.F32, [M, N], [1, M])[1], lib.create_info(tc.F32, [M, N], [1, M])[1]]
    rc, plan = lib.create_plan([0] * 4, infos, [[m, k], [k, n], [m, n], [m, n]], tc.DEFAULT_PREC)
    assert rc == 0
    D = torch.zeros((N, M), device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    one, zero = np.ones(1, np.float32), np.zeros(1, np.float32)
    rc = lib.execute(plan, ex, one.ctypes.data, A.data_ptr(), B.data_ptr(), zero.ctypes.data, None, D.data_ptr())
    assert rc == 0, lib.explain(rc)
    rc, name = lib.last_kernel_name(ex)
    assert rc == 0
    out = D.t().double().cpu().numpy()
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)
    lib.destroy_executor(ex)
    return out, name


def _exact(torch, a_view, b_view):
    return (a_view.double() @ b_view.double()).cpu().numpy()


def _normwise(x, y):
    return float(np.abs(x - y).max() / np.abs(y).max())


SHAPES = [(4096, 2048, 512), (4136, 2072, 272)]      # exact tiles; edge tiles with a K tail


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    return torch


@pytest.mark.parametrize("a_rows", [True, False], ids=["A_unitM", "A_unitK"])
@pytest.mark.parametrize("b_rows", [True, False], ids=["B_unitN", "B_unitK"])
@pytest.mark.parametrize("shape", SHAPES, ids=["exact", "edges"])
def test_wide_split_bf16_bitwise_equal_to_128_tiles(lib, torch_mod, a_rows, b_rows, shape):
    torch = torch_mod
    M, N, K = shape
    a_base, a_view, sa = _mk(torch, M, K, a_rows, 11)
    b_base, b_view, sb = _mk(torch, K, N, not b_rows, 12)
    # B's "rows" in the kernel are N: unit N stride is unit stride along cols of [K, N]
    wide, name = _run(lib, torch, a_base, sa, b_base, sb, M, N, K, 1, 0)
    assert name.startswith("k_gett_wide<"), name
    small, name2 = _run(lib, torch, a_base, sa, b_base, sb, M, N, K, 1, 2)
    assert name2.startswith("k_gett_lp<"), name2
    assert np.array_equal(wide, small)
    assert _normwise(wide, _exact(torch, a_view, b_view)) <= 1e-5


@pytest.mark.parametrize("a_rows", [True, False], ids=["A_unitM", "A_unitK"])
@pytest.mark.parametrize("b_rows", [True, False], ids=["B_unitN", "B_unitK"])
@pytest.mark.parametrize("shape", SHAPES, ids=["exact", "edges"])
def test_split_f16_at_wide_shapes_against_fp64(lib, torch_mod, a_rows, b_rows, shape):
    torch = torch_mod
    M, N, K = shape
    a_base, a_view, sa = _mk(torch, M, K, a_rows, 21)
    b_base, b_view, sb = _mk(torch, K, N, not b_rows, 22)
    got, name = _run(lib, torch, a_base, sa, b_base, sb, M, N, K, 3, 0)
    assert name.startswith("k_gett_lp<") and "SPLIT_F16" in name, name
    assert _normwise(got, _exact(torch, a_view, b_view)) <= 1e-6


def test_split_f16_rescaled_launch(lib, torch_mod):
    """max |A| = 1e-4 is below binary16's accepted range [2^-3, 2^15): the
    product reruns with A scaled by a power of two."""
    torch = torch_mod
    M, N, K = 4096, 2048, 256
    a_base, a_view, sa = _mk(torch, M, K, True, 31, scale=1e-4)
    b_base, b_view, sb = _mk(torch, K, N, False, 32)
    got, name = _run(lib, torch, a_base, sa, b_base, sb, M, N, K, 3, 0)
    assert "SPLIT_F16" in name, name
    assert _normwise(got, _exact(torch, a_view, b_view)) <= 1e-6


def test_split_f16_nan_falls_back_to_fp32(lib, torch_mod):
    torch = torch_mod
    M, N, K = 4096, 2048, 256
    a_base, a_view, sa = _mk(torch, M, K, True, 41)
    b_base, b_view, sb = _mk(torch, K, N, False, 42)
    a_base[7, 100] = float("nan")
    got, name = _run(lib, torch, a_base, sa, b_base, sb, M, N, K, 3, 0)
    assert name.startswith("k_gett_f32<"), name
    plain, _ = _run(lib, torch, a_base, sa, b_base, sb, M, N, K, 0, 0)
    assert np.array_equal(np.isnan(got), np.isnan(plain))
    ok = ~np.isnan(plain)
    assert np.array_equal(got[ok], plain[ok])


def test_tiles_attribute_values(lib):
    rc, ex = lib.create_executor()
    assert rc == 0
    for v in (0, 1, 2, 3):
        assert lib.attr_set_int(ex, ATTR_TILES, v) == 0
        assert lib.attr_get_int(ex, ATTR_TILES) == (0, v)
    assert lib.attr_set_int(ex, ATTR_TILES, 4) != 0
    assert lib.attr_get_int(ex, ATTR_TILES) == (0, 3)
    lib.destroy_executor(ex)

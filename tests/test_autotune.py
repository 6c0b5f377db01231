"""Cached autotuner (executor attribute TAPP_AMD_ATTR_AUTOTUNE): the chosen
configuration per shape class is measured once, written to the cache file
and reused; results stay within the parity tolerances whatever is chosen."""
import os
import tempfile

import numpy as np
import pytest

import tapp_cases as T
from test_gpu_parity import _gemm_case, assert_parity, run_both

tc = T.tc

# the library reads the cache path on first use: set it before any test runs
CACHE = os.path.join(tempfile.mkdtemp(prefix="tapp_tune_"), "tune.txt")
os.environ["TAPP_AMD_TUNE_CACHE"] = CACHE


def _lines(kind):
    if not os.path.exists(CACHE):
        return []
    with open(CACHE) as f:
        return [ln.split() for ln in f if ln.split()[2:3] == [kind]]


def test_autotune_attribute_values(lib):
    rc, ex = lib.create_executor()
    assert rc == 0
    assert lib.attr_get_int(ex, tc.ATTR_AUTOTUNE) == (0, 0)
    for v in (1, 2, 0):
        assert lib.attr_set_int(ex, tc.ATTR_AUTOTUNE, v) == 0
        assert lib.attr_get_int(ex, tc.ATTR_AUTOTUNE) == (0, v)
    assert lib.attr_set_int(ex, tc.ATTR_AUTOTUNE, 3) == tc.AMD_ERR_INVALID
    lib.destroy_executor(ex)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tc.F32, tc.F64, tc.C64])
def test_autotuned_products_match_and_are_cached(lib, oracle, dtype):
    shapes = [(200, 130, 257), (64, 64, 64), (40, 1200, 24)]
    _, tune = lib.create_executor()
    assert lib.attr_set_int(tune, tc.ATTR_AUTOTUNE, 1) == 0
    _, cached = lib.create_executor()
    assert lib.attr_set_int(cached, tc.ATTR_AUTOTUNE, 2) == 0
    for i, (M, N, K) in enumerate(shapes):
        c = _gemm_case(dtype, M, N, K, 300 + i, transA=i == 1, ldpad=i % 2)
        before = len(_lines("gett"))
        rc, got, want = run_both(lib, tune, oracle, c)
        assert rc == 0
        assert_parity(got, want, c)
        entries = _lines("gett")
        assert len(entries) == before + 1, "one measured class per new shape"
        chosen = int(entries[-1][1]) & 15
        assert lib.attr_get_int(tune, tc.ATTR_LAST_KERNEL)[1] == chosen
        # the same class again, cache only: same kernel, nothing measured
        rc, got2, _ = run_both(lib, cached, oracle, c)
        assert rc == 0
        assert_parity(got2, want, c)
        assert lib.attr_get_int(cached, tc.ATTR_LAST_KERNEL)[1] == chosen
        assert len(_lines("gett")) == before + 1
    lib.destroy_executor(tune)
    lib.destroy_executor(cached)


@pytest.mark.gpu
def test_autotune_skips_in_place_beta_c(lib, oracle):
    """C and D the same memory with beta != 0: measuring would feed D back
    into C, so the class is not measured and the rules decide."""
    import torch
    _, ex = lib.create_executor()
    assert lib.attr_set_int(ex, tc.ATTR_AUTOTUNE, 1) == 0
    c = _gemm_case(tc.F32, 96, 72, 150, 77, beta=0.5)
    c.D.buf[:] = c.C.buf                       # D starts as C; the product runs in place on one buffer
    rc_o, want = oracle.run(c)
    assert rc_o == 0
    buf = torch.from_numpy(c.D.buf.copy()).cuda()
    infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in c.operands()]
    _, plan = lib.create_plan([0, 0, 0, 0], infos, [o.idx for o in c.operands()], c.prec)
    a = torch.from_numpy(c.A.buf.copy()).cuda()
    b = torch.from_numpy(c.B.buf.copy()).cuda()
    al, be = T.scalar_bytes(c.alpha, c.D.type), T.scalar_bytes(c.beta, c.D.type)
    before = len(_lines("gett"))
    rc = lib.execute(plan, ex, al.ctypes.data, a.data_ptr() + 4 * c.A.offset, b.data_ptr() + 4 * c.B.offset,
                     be.ctypes.data, buf.data_ptr() + 4 * c.C.offset, buf.data_ptr() + 4 * c.D.offset)
    assert rc == 0
    assert len(_lines("gett")) == before
    assert_parity(buf.cpu().numpy(), want, c)
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)
    lib.destroy_executor(ex)


@pytest.mark.gpu
def test_autotuned_group_matches_single_calls(lib, gpu_exec):
    from test_group import DevCase, make_group
    cases = T.batching_mix(5, 48, tucker_range=(8, 48), r_range=(2, 16), c0_range=(4, 9))
    dcs = [DevCase(lib, c) for c in cases]
    want = []
    for d in dcs:
        assert d.run_single(gpu_exec) == 0
        want.append(d.d().copy())
        d.reset_d()
    _, ex = lib.create_executor()
    assert lib.attr_set_int(ex, tc.ATTR_AUTOTUNE, 1) == 0
    rc, g = make_group(lib, ex, dcs)
    assert rc == 0
    assert len(_lines("group")) >= 1, "thin-capable classes measured at creation"
    for rep in range(2):
        for d in dcs:
            d.reset_d()
        assert lib.execute_group(g) == 0
        for d, w in zip(dcs, want):
            assert_parity(d.d(), w, d.case, tol=1e-5)
    lib.destroy_group(g)
    lib.destroy_executor(ex)
    for d in dcs:
        d.close()

"""GPU parity on the precision ladder, D[a,b,i,j] = sum_{c,d} A[a,b,c,d]
B[c,i,d,j]: the layout the benchmark times (B permuted, so its K group is
read gathered across d), with the decaying N(0, 1) entries of that
workload.

Small extents (a..d = 24, i, j = 8; K = 576) are compared with the oracle
in full; the full size (a..d = 128, i, j = 32; K = 16384) on D-slices the
oracle computes at the full contraction length.  Tolerances:
  * plain FP32 against the oracle (the reference's binary32 loop): 1e-5
    normwise;
  * split forms against the FP64 oracle: split-BF16 x3 1e-5, three-piece
    split-BF16 x6 and split-F16 x3 1e-6 (normwise);
  * BF16BF16_ACCUM_F32 / F16F16_ACCUM_F32 against the FP64 oracle on the
    unrounded inputs, elementwise |D - D_exact| <= (2 u_in + gamma_K(u32))
    (sum_k |alpha a b|), u_in = 2^-8 / 2^-11, and against the oracle's GPU
    form (same input rounding, binary32 products) at 1e-5.
"""
import ctypes

import numpy as np
import pytest

import tapp_cases as T

tc = T.tc
pytestmark = pytest.mark.gpu

SPLIT = {"split_bf16x3": (1, 1e-5), "split_bf16x6": (2, 1e-6), "split_f16x3": (3, 1e-6)}
U_IN = {tc.BF16BF16_ACCUM_F32: 2.0 ** -8, tc.F16F16_ACCUM_F32: 2.0 ** -11}


def _executor(lib, split=0):
    rc, ex = lib.create_executor()
    assert rc == 0
    assert lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, split) == 0
    return ex


def _k_bound(prec, K):
    u32 = 2.0 ** -24
    return 2 * U_IN[prec] + K * u32 / (1 - K * u32)


def _magnitude(case):
    """sum_{c,d} |alpha A[a,b,c,d] B[c,i,d,j]| + |beta C| in binary64."""
    c = case.copy()
    for o in c.operands():
        o.buf = np.abs(o.buf).astype(np.float64)
        o.type = tc.F64
    c.alpha, c.beta = abs(c.alpha), abs(c.beta)
    c.prec = tc.F64F64_ACCUM_F64
    return c


@pytest.mark.parametrize("beta", [0.0, 0.5])
def test_ladder_fp32_against_oracle(lib, gpu_exec, oracle, beta):
    c = T.ladder_case(seed=3, beta=beta)
    rc, got = T.run_lib(lib, c, gpu_exec)
    assert rc == 0
    assert lib.attr_get_int(gpu_exec, tc.ATTR_LAST_KERNEL)[1] in (tc.KERNEL_MFMA, tc.KERNEL_MFMA_SMALL)
    _, want = oracle.run(c)
    assert T.normwise_error(got, want, T.written_mask(c.D)) <= 1e-5


@pytest.mark.parametrize("form", sorted(SPLIT))
def test_ladder_split_forms_against_fp64_oracle(lib, oracle, form):
    split, tol = SPLIT[form]
    ex = _executor(lib, split)
    c = T.ladder_case(seed=4)
    rc, got = T.run_lib(lib, c, ex)
    assert rc == 0 and lib.attr_get_int(ex, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_MFMA
    _, exact = oracle.run(c, f64=True)
    err = T.normwise_error(got, exact, T.written_mask(c.D))
    assert err <= tol, (form, err)
    lib.destroy_executor(ex)


@pytest.mark.parametrize("prec", [tc.BF16BF16_ACCUM_F32, tc.F16F16_ACCUM_F32])
def test_ladder_sixteen_bit_precisions(lib, gpu_exec, oracle, prec):
    c = T.ladder_case(seed=5, prec=prec)
    rc, got = T.run_lib(lib, c, gpu_exec)
    assert rc == 0
    mask = T.written_mask(c.D)
    _, gpu_form = oracle.run(c, gpu=True)
    assert T.normwise_error(got, gpu_form, mask) <= 1e-5
    _, exact = oracle.run(c, f64=True)
    _, mag = oracle.run(_magnitude(c), f64=True)
    K = 24 * 24
    diff = np.abs(got[mask].astype(np.float64) - exact[mask])
    assert np.all(diff <= _k_bound(prec, K) * mag[mask])


# --- full size: D-slices at the full contraction length ------------------------

V, O = 128, 32


@pytest.fixture(scope="module")
def ladder_full():
    """The benchmark's operands (generated on the device), host copies, and a
    D buffer for each form."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(1234)
    idx = torch.arange(V, device="cuda", dtype=torch.float32)
    decay = torch.exp(-(idx[:, None] - idx[None, :]).abs() / 16.0)
    A = torch.randn((V, V, V, V), generator=g, device="cuda", dtype=torch.float32)   # dims (d, c, b, a)
    A.mul_(decay.view(1, V, 1, V))
    B = torch.randn((O, V, O, V), generator=g, device="cuda", dtype=torch.float32)   # dims (j, d, i, c)
    B.mul_(decay.view(1, V, 1, V))
    for t in (A, B):
        m = t.abs().max().item()
        if m > 8.0:
            t.mul_(8.0 / m)
    torch.cuda.synchronize()
    yield torch, A, B, A.cpu().numpy().reshape(-1), B.cpu().numpy().reshape(-1)


def _run_full(lib, torch, A, B, prec, split):
    ex = _executor(lib, split)
    a, b, cc, d, i, j = (ord(x) for x in "abcdij")
    infos = [lib.create_info(tc.F32, [V, V, V, V], [1, V, V * V, V ** 3])[1],
             lib.create_info(tc.F32, [V, O, V, O], [1, V, V * O, V * O * V])[1],
             lib.create_info(tc.F32, [V, V, O, O], [1, V, V * V, V * V * O])[1],
             lib.create_info(tc.F32, [V, V, O, O], [1, V, V * V, V * V * O])[1]]
    rc, plan = lib.create_plan([0] * 4, infos, [[a, b, cc, d], [cc, i, d, j], [a, b, i, j], [a, b, i, j]], prec)
    assert rc == 0
    D = torch.zeros((O, O, V, V), device="cuda", dtype=torch.float32)
    torch.cuda.synchronize()
    one, zero = np.ones(1, np.float32), np.zeros(1, np.float32)
    rc = lib.execute(plan, ex, one.ctypes.data, A.data_ptr(), B.data_ptr(), zero.ctypes.data, None, D.data_ptr())
    assert rc == 0, lib.explain(rc)
    assert lib.attr_get_int(ex, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_MFMA
    out = D.cpu().numpy().reshape(-1)
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)
    lib.destroy_executor(ex)
    return out


def _oracle_slice(oracle, a_host, b_host, i0, j0, f64=False):
    """D[:, :, i0, j0] by the oracle at the full K = 16384: B and D as the
    subtensors (c, i0, d, j0) and (a, b, i0, j0) of the full layouts."""
    A = T.Operand(tc.F32, [V, V, V, V], [1, V, V * V, V ** 3], T.labels("abcd"), 0, a_host, 0)
    B = T.Operand(tc.F32, [V, 1, V, 1], [1, V, V * O, V * O * V], T.labels("cidj"), 0, b_host,
                  i0 * V + j0 * V * O * V)
    D = T.Operand(tc.F64 if f64 else tc.F32, [V, V, 1, 1], [1, V, V * V, V * V * O], T.labels("abij"), 0,
                  np.zeros(V * V, np.float64 if f64 else np.float32), 0)
    keep = []
    ts = [ctypes.byref(oracle._t(o, keep)) for o in (A, B, D, D)]
    one = np.ones(1, np.float64 if f64 else np.float32)
    zero = np.zeros(1, one.dtype)
    if f64:
        rc = oracle.lib.orc_execute_f64(*ts, tc.DEFAULT_PREC, one.ctypes.data, A.ptr(), B.ptr(), zero.ctypes.data,
                                        None, D.ptr(), 0)
    else:
        kind = ctypes.c_int(1)
        rc = oracle.lib.orc_execute_ex(*ts, tc.DEFAULT_PREC, one.ctypes.data, A.ptr(), B.ptr(), zero.ctypes.data,
                                       None, D.ptr(), ctypes.byref(kind), 0)
    assert rc == 0
    return D.buf


def _slice_of(d_full, i0, j0):
    return d_full[i0 * V * V + j0 * V * V * O:][:V * V]


def test_ladder_full_size_fp32_slices_against_oracle(lib, oracle, ladder_full):
    torch, A, B, a_host, b_host = ladder_full
    got = _run_full(lib, torch, A, B, tc.DEFAULT_PREC, 0)
    for i0, j0 in ((0, 0), (31, 17)):
        want = _oracle_slice(oracle, a_host, b_host, i0, j0)
        err = T.normwise_error(_slice_of(got, i0, j0), want)
        assert err <= 1e-5, (i0, j0, err)


def test_ladder_full_size_forms_against_fp64_oracle(lib, oracle, ladder_full):
    """Every form at the full size on the slice i = 5, j = 9, against the
    FP64 oracle: the split forms at their normwise bounds, the 16-bit
    precisions elementwise at 2 u_in + gamma_K(u32) of sum_k |a b| (the
    magnitudes summed in binary64 on the device)."""
    torch, A, B, a_host, b_host = ladder_full
    i0, j0 = 5, 9
    exact = _oracle_slice(oracle, a_host, b_host, i0, j0, f64=True)
    bvec = B[j0, :, i0, :].abs().double().reshape(-1)            # (d, c) flattened as d * V + c
    mag = (A.abs().double().view(V * V, V * V).t() @ bvec).cpu().numpy()   # rows (d, c), cols (b, a)
    for form, (split, tol) in sorted(SPLIT.items()):
        got = _slice_of(_run_full(lib, torch, A, B, tc.DEFAULT_PREC, split), i0, j0)
        err = T.normwise_error(got, exact)
        assert err <= tol, (form, err)
    for prec in (tc.BF16BF16_ACCUM_F32, tc.F16F16_ACCUM_F32):
        got = _slice_of(_run_full(lib, torch, A, B, prec, 0), i0, j0).astype(np.float64)
        assert np.all(np.abs(got - exact) <= _k_bound(prec, V * V) * mag), prec

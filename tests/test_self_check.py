"""Self-checking executor (the reference's executor kind 12,
product.c:229-249, 424-448; comparison rule tblis_bind.cpp:577-640).

Kind 12 computes each product twice with two kernel families and returns
137 when the results differ by the reference's rule.  Both results are
reproduced here from the same library with the kernel family forced, and
the rule is applied in numpy to predict the outcome: the return code, the
number of differing elements printed, and that D holds exactly what the
executor without the check writes.
"""
import ctypes

import numpy as np
import pytest

import tapp_cases as T

tc = T.tc


def differing(x, y, dtype):
    """Number of positions where the reference's compare_tensors_ reports a
    difference (tblis_bind.cpp:577-640): |(a - b) / max(a, b)| > 5e-5 in the
    storage precision, signed max, NaN quotients not counted; complex per
    component; binary64 complex also needs an absolute difference > 1e-12."""
    def rel(a, b):
        with np.errstate(all="ignore"):
            q = (a - b) / np.where(a > b, a, b)
            return np.abs(q).astype(np.float64) > 0.00005

    if dtype in (tc.F32, tc.F64):
        return int(np.count_nonzero(rel(x, y)))
    if dtype in (tc.C32, tc.C64):
        real = np.float32 if dtype == tc.C32 else np.float64
        a, b = x.view(real).reshape(-1, 2), y.view(real).reshape(-1, 2)
        bad = rel(a[:, 0], b[:, 0]) | rel(a[:, 1], b[:, 1])
        if dtype == tc.C64:
            bad &= (np.abs(a[:, 0] - b[:, 0]) > 1e-12) | (np.abs(a[:, 1] - b[:, 1]) > 1e-12)
        return int(np.count_nonzero(bad))
    if dtype == tc.F16:
        return int(np.count_nonzero(rel(x.astype(np.float32), y.astype(np.float32))))
    if dtype == tc.BF16:
        f = lambda v: (v.astype(np.uint32) << 16).view(np.float32)  # noqa: E731
        return int(np.count_nonzero(rel(f(x), f(y))))
    raise ValueError(dtype)


def test_attribute_writes_the_kind(lib):
    """TAPP_AMD_ATTR_SELF_CHECK sets the int the executor points at to 12
    (or back to 1), the value a reference caller writes for kind 12."""
    rc, ex = lib.create_executor()
    assert rc == 0
    kind = ctypes.c_int.from_address(ex)
    assert kind.value == 1
    assert lib.attr_set_int(ex, tc.ATTR_SELF_CHECK, 1) == 0
    assert kind.value == 12 and lib.attr_get_int(ex, tc.ATTR_SELF_CHECK) == (0, 1)
    assert lib.attr_set_int(ex, tc.ATTR_SELF_CHECK, 0) == 0
    assert kind.value == 1
    kind.value = 12
    assert lib.attr_get_int(ex, tc.ATTR_SELF_CHECK) == (0, 1)
    lib.destroy_executor(ex)


# --------------------------------------------------------------------------
# GPU: outcome predicted from the two kernel families run separately
# --------------------------------------------------------------------------
def _family_results(lib, case):
    """(primary family, D of the plain run, D of the check family's run)."""
    rc, ex = lib.create_executor()
    rc, d_plain = T.run_lib(lib, case, ex)
    fam = lib.attr_get_int(ex, tc.ATTR_LAST_KERNEL)[1]
    lib.destroy_executor(ex)
    assert rc == 0, lib.explain(rc)
    check = tc.KERNEL_SIMPLE if fam != tc.KERNEL_SIMPLE else tc.KERNEL_MFMA
    rc, ex = lib.create_executor()
    lib.attr_set_int(ex, tc.ATTR_KERNEL, check)
    rc, d_check = T.run_lib(lib, case, ex)
    lib.destroy_executor(ex)
    assert rc == 0, lib.explain(rc)
    return fam, d_plain, d_check


def _run_checked(lib, case, capfd, via_int=False):
    rc, ex = lib.create_executor()
    if via_int:
        ctypes.c_int.from_address(ex).value = 12
    else:
        lib.attr_set_int(ex, tc.ATTR_SELF_CHECK, 1)
    capfd.readouterr()
    rc, d = T.run_lib(lib, case, ex)
    out = capfd.readouterr().out
    lib.destroy_executor(ex)
    return rc, d, out


def _assert_outcome(lib, case, capfd, via_int=False, want_diff=None):
    fam, d_plain, d_check = _family_results(lib, case)
    n = differing(d_plain, d_check, case.D.type)
    if want_diff is not None:
        assert (n > 0) == want_diff, (case.name, n)
    rc, d, out = _run_checked(lib, case, capfd, via_int)
    assert rc == (137 if n else 0), (case.name, fam, n, rc, out)
    assert np.array_equal(d.view(np.uint8), d_plain.view(np.uint8)), case.name
    if n:
        assert f"{n} of {len(d_plain)} elements of D differ" in out, out
    else:
        assert "self-check" not in out
    return n


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tc.F32, tc.F64, tc.C32, tc.C64])
@pytest.mark.parametrize("scenario", sorted(T.SCENARIOS))
def test_random_products(lib, gpu_exec, capfd, dtype, scenario):
    for seed in range(2):
        case = T.random_case(seed + 900, scenario, dtype, max_ext=6)
        _assert_outcome(lib, case, capfd)


def _cancelling_case(seed, K2=256, M=64, N=64):
    """D[m, n] = sum_k X[m, k] Y[k, n] - X[m, k] (Y[k, n] + e[k, n]) with
    |X|, |Y| ~ 10 and |e| ~ 1e-3: partial sums near 1e3 cancel down to about
    0.1, so FP32 rounding in two summation orders leaves relative
    differences far above 5e-5."""
    rng = np.random.default_rng(seed)
    X = (10 * rng.standard_normal((M, K2))).astype(np.float32)
    Y = (10 * rng.standard_normal((K2, N))).astype(np.float32)
    e = (1e-3 * rng.standard_normal((K2, N))).astype(np.float32)
    Am = np.concatenate([X, -X], axis=1)
    Bm = np.concatenate([Y, Y + e], axis=0)
    A = T.make_operand(tc.F32, [M, 2 * K2], "mk", 1)
    B = T.make_operand(tc.F32, [2 * K2, N], "kn", 2)
    C = T.make_operand(tc.F32, [M, N], "mn", 3)
    D = T.make_operand(tc.F32, [M, N], "mn", 4)
    A.buf[:] = Am.ravel(order="F")
    B.buf[:] = Bm.ravel(order="F")
    return T.Case(A, B, C, D, alpha=1.0, beta=0.0, name=f"cancelling/{seed}")


def _positive_case(seed, beta=0.0, M=48, N=40, K=96):
    """Positive operands: no cancellation, both families agree to ~1e-7."""
    A = T.make_operand(tc.F32, [M, K], "mk", seed * 4 + 1)
    B = T.make_operand(tc.F32, [K, N], "kn", seed * 4 + 2)
    C = T.make_operand(tc.F32, [M, N], "mn", seed * 4 + 3)
    D = T.make_operand(tc.F32, [M, N], "mn", seed * 4 + 4)
    for o in (A, B, C, D):
        o.buf[:] = np.abs(o.buf) + 0.5
    return T.Case(A, B, C, D, alpha=1.0, beta=beta, name=f"positive/{seed}")


@pytest.mark.gpu
def test_detects_differing_results(lib, gpu_exec, capfd):
    n = _assert_outcome(lib, _cancelling_case(1), capfd, want_diff=True)
    assert n > 0


@pytest.mark.gpu
def test_kind_int_written_by_caller(lib, gpu_exec, capfd):
    _assert_outcome(lib, _positive_case(2), capfd, via_int=True, want_diff=False)
    _assert_outcome(lib, _cancelling_case(3), capfd, via_int=True, want_diff=True)


def _run_in_place(lib, case, ex):
    """D = alpha A B + beta D with C and D the same host array."""
    D = case.D.copy()
    infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in (case.A, case.B, D, D)]
    _, plan = lib.create_plan([0, 0, 0, 0], infos, [case.A.idx, case.B.idx, D.idx, D.idx], case.prec)
    al = T.scalar_bytes(case.alpha, tc.F32)
    be = T.scalar_bytes(case.beta, tc.F32)
    rc = lib.execute(plan, ex, al.ctypes.data, case.A.ptr(), case.B.ptr(), be.ctypes.data, D.ptr(), D.ptr())
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)
    return rc, D.buf


@pytest.mark.gpu
def test_c_and_d_the_same_array(lib, gpu_exec, capfd):
    """D = A B + D in place: both runs must read the original D as C (a
    check run reading the product's output would differ by about |D|)."""
    case = _positive_case(4, beta=1.0)
    rc, d_plain = _run_in_place(lib, case, gpu_exec)
    assert rc == 0
    assert not np.array_equal(d_plain, case.D.buf)
    rc, ex = lib.create_executor()
    lib.attr_set_int(ex, tc.ATTR_SELF_CHECK, 1)
    capfd.readouterr()
    rc, d = _run_in_place(lib, case, ex)
    out = capfd.readouterr().out
    lib.destroy_executor(ex)
    assert rc == 0, out
    assert np.array_equal(d, d_plain)


@pytest.mark.gpu
def test_device_operands(lib, gpu_exec, capfd):
    """Operands in device memory (torch allocations), D included."""
    import torch
    for case in (_positive_case(5), _cancelling_case(6)):
        _, d_plain, d_check = _family_results(lib, case)
        want_rc = 137 if differing(d_plain, d_check, tc.F32) else 0
        bufs = [torch.from_numpy(o.buf.copy()).cuda() for o in case.operands()]
        ptr = [b.data_ptr() + o.offset * o.buf.itemsize for b, o in zip(bufs, case.operands())]
        torch.cuda.synchronize()
        infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in case.operands()]
        _, plan = lib.create_plan([0, 0, 0, 0], infos, [o.idx for o in case.operands()], case.prec)
        rc, ex = lib.create_executor()
        lib.attr_set_int(ex, tc.ATTR_SELF_CHECK, 1)
        al = T.scalar_bytes(case.alpha, tc.F32)
        be = T.scalar_bytes(case.beta, tc.F32)
        capfd.readouterr()
        rc = lib.execute(plan, ex, al.ctypes.data, ptr[0], ptr[1], be.ctypes.data, ptr[2], ptr[3])
        out = capfd.readouterr().out
        assert rc == want_rc, (case.name, rc, out)
        assert np.array_equal(bufs[3].cpu().numpy(), d_plain), case.name
        lib.destroy_executor(ex)
        lib.destroy_plan(plan)
        for h in infos:
            lib.destroy_info(h)


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [tc.F16, tc.BF16])
def test_sixteen_bit_storage(lib, gpu_exec, capfd, dtype):
    """D stored as binary16 or bfloat16, compared as binary32 values."""
    for seed in range(3):
        case = T.random_case(960 + seed, "contraction", tc.F32, max_ext=8, min_ext=4, beta_zero=False)
        for o in case.operands():
            lo, hi = o.span()
            o.type = dtype
            o.buf = T.fill(dtype, max(1, hi - lo + 1), 970 + seed)
        case.prec = tc.F32F32_ACCUM_F32
        _assert_outcome(lib, case, capfd)

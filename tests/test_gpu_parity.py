"""GPU parity: the library's TAPP_execute_product on an MI355X against the
oracle (itself pinned bit-exact to the compiled reference) on the same
seeded inputs.

Tolerances (normwise, max|D - D_ref| / max|D_ref| over the written
elements): FP32/C32 1e-5, FP64/C64 1e-12.  Elements of D the product does
not write must be bit-identical to their input values.
"""
import ctypes

import numpy as np
import pytest

import tapp_cases as T

tc = T.tc
pytestmark = pytest.mark.gpu

TOL = {tc.F32: 1e-5, tc.C32: 1e-5, tc.F64: 1e-12, tc.C64: 1e-12, tc.F16: 2e-3, tc.BF16: 1.6e-2}


def tol_for(case):
    t = TOL[case.D.type]
    if case.prec in (tc.F32F32_ACCUM_F32,):
        t = max(t, 1e-5)
    if case.prec == tc.F64F64_ACCUM_F64 and case.D.type in (tc.F32, tc.C32):
        t = 1e-6   # result stored in binary32
    return t


def assert_parity(got, want, case, tol=None):
    mask = T.written_mask(case.D)
    tol = tol_for(case) if tol is None else tol
    err = T.normwise_error(got, want, mask)
    assert err <= tol, f"{case.name}: normwise error {err:.3e} > {tol:.1e}"
    untouched = ~mask
    assert np.array_equal(np.asarray(got).view(np.uint8).reshape(len(mask), -1)[untouched],
                          case.D.buf.view(np.uint8).reshape(len(mask), -1)[untouched]), f"{case.name}: wrote outside D"
    return err


def run_both(lib, ex, oracle, case, f64_oracle=False):
    rc_o, d_o = oracle.run(case, f64=f64_oracle, gpu=True)
    rc_g, d_g = T.run_lib(lib, case, ex)
    assert rc_g == rc_o, (case.name, rc_g, rc_o, lib.explain(rc_g))
    return rc_o, d_g, d_o


def run_device(lib, ex, case):
    """Same call with every operand in device memory (torch allocations)."""
    import torch
    case = case.copy()
    dev = {}
    ptrs = {}
    for name, o in zip("ABCD", case.operands()):
        t = torch.from_numpy(o.buf.view(np.uint8).copy()).cuda()
        dev[name] = t
        ptrs[name] = t.data_ptr() + o.offset * o.buf.itemsize
    al = T.scalar_bytes(case.alpha, case.D.type)
    be = T.scalar_bytes(case.beta, case.D.type)
    infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in case.operands()]
    _, plan = lib.create_plan([o.op for o in case.operands()], infos, [o.idx for o in case.operands()], case.prec)
    torch.cuda.synchronize()
    rc = lib.execute(plan, ex, al.ctypes.data, ptrs["A"], ptrs["B"], be.ctypes.data,
                     None if case.c_null else ptrs["C"], ptrs["D"])
    out = dev["D"].cpu().numpy().view(case.D.buf.dtype)
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)
    return rc, out


@pytest.mark.parametrize("dtype", [tc.F32, tc.F64, tc.C32, tc.C64])
@pytest.mark.parametrize("prec", [tc.DEFAULT_PREC, tc.F32F32_ACCUM_F32, tc.F64F64_ACCUM_F64])
@pytest.mark.parametrize("scenario", sorted(T.SCENARIOS))
def test_random_scenarios(lib, gpu_exec, oracle, dtype, prec, scenario):
    for seed in range(2):
        c = T.random_case(2000 + seed * 31, scenario, dtype, prec=prec, max_ext=5, repeated_D=(seed == 1))
        rc, got, want = run_both(lib, gpu_exec, oracle, c)
        if rc == 0:
            assert_parity(got, want, c)


@pytest.mark.parametrize("dtype", [tc.F32, tc.F64, tc.C32, tc.C64])
def test_layout_mix_multi_tile(lib, gpu_exec, oracle, dtype):
    """Multi-tile products with random label orders, storage orders, padded
    and negative strides, a Hadamard mode and conj(B): the auto kernel
    choice and the MFMA kernel forced, each against the oracle."""
    _, exm = lib.create_executor()
    assert lib.attr_set_int(exm, tc.ATTR_KERNEL, tc.KERNEL_MFMA) == 0
    for seed in range(6):
        c = T.layout_mix_case(1000 * dtype + seed, dtype)
        rc, got, want = run_both(lib, gpu_exec if seed % 2 == 0 else exm, oracle, c)
        assert rc == 0
        assert_parity(got, want, c)
    lib.destroy_executor(exm)


@pytest.mark.parametrize("mode", ["split_bf16", "split6_bf16", "split_f16"])
def test_layout_mix_split_forms(lib, oracle, mode):
    """The split forms on the same layouts, against the FP64 oracle."""
    _, ex = lib.create_executor()
    assert lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, SPLIT_ATTR[mode]) == 0
    for seed in range(4):
        c = T.layout_mix_case(7000 + 100 * SPLIT_ATTR[mode] + seed, tc.F32)
        rc, got, exact = run_both(lib, ex, oracle, c, f64_oracle=True)
        assert rc == 0
        assert T.normwise_error(got, exact, T.written_mask(c.D)) < SPLIT_TOL[mode], c.name
    lib.destroy_executor(ex)


def test_reference_program_known_answers(lib, gpu_exec):
    from test_oracle import _case_from_kat, load_kats
    for k in load_kats():
        c = _case_from_kat(k)
        rc, got = T.run_lib(lib, c, gpu_exec)
        assert rc == k["rc"]
        want = np.array(k["D_out"]["data"], dtype=np.float64)
        g = np.asarray(got)
        if g.dtype.kind == "c":
            g = np.stack([g.real, g.imag], axis=-1).reshape(-1)
        assert np.allclose(g.astype(np.float64), want, rtol=1e-6, atol=1e-5), (k["program"], k["call"])


def test_four_index16_against_reference_golden(lib, gpu_exec):
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_four_index16.npz"))
    c = T.four_index16_case(seed=int(g["seed"]))
    rc, got = T.run_lib(lib, c, gpu_exec)
    assert rc == 0
    assert_parity(got, g["D"], c)


def test_c64_feature_against_reference_golden_sample(lib, gpu_exec):
    """The complex-double feature product on the GPU against the compiled reference's D at 4096
    sampled positions (FP64 tolerance, normwise over the sample)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_c64_feature_sample.npz"))
    c = T.c64_feature_case(seed=int(g["seed"]))
    rc, got = T.run_lib(lib, c, gpu_exec)
    assert rc == 0
    want = g["D"]
    err = np.abs(got[g["idx"]] - want).max() / np.abs(want).max()
    assert err < 1e-12, err


def test_small_tile_kernel_by_rule_and_batches(lib, gpu_exec, oracle):
    """Few outputs and a moderate K take the 16 x 16 small-tile kernel by
    rule (the 16^6 four-index product itself; a Hadamard batch; ragged edges; K not a
    multiple of 4; beta != 0 with a transposed A), each against the oracle,
    and the same product run twice gives the same bits."""
    cases = [T.four_index16_case(seed=3),
             T.four_index_case(5, (9, 7, 13, 6, 11, 5)),
             _gemm_case(tc.F32, 100, 90, 301, 21, transA=True),
             _gemm_case(tc.F32, 17, 250, 1000, 22, transB=True, beta=0.0)]
    A = T.make_operand(tc.F32, [40, 37, 5], "mkl", 31)
    B = T.make_operand(tc.F32, [37, 24, 5], "knl", 32)
    C = T.make_operand(tc.F32, [40, 24, 5], "mnl", 33)
    D = T.make_operand(tc.F32, [40, 24, 5], "mnl", 34, pad=1)
    cases.append(T.Case(A, B, C, D, tc.DEFAULT_PREC, 0.5, -1.25, name="small_batch"))
    for c in cases:
        rc, got, want = run_both(lib, gpu_exec, oracle, c)
        assert rc == 0
        assert lib.attr_get_int(gpu_exec, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_MFMA_SMALL, c.name
        assert_parity(got, want, c)
        rc, again = T.run_lib(lib, c, gpu_exec)
        assert rc == 0 and np.array_equal(again.view(np.uint32), got.view(np.uint32)), c.name


def test_c64_feature_full_size(lib, gpu_exec, oracle):
    c = T.c64_feature_case(seed=1)
    rc, got, want = run_both(lib, gpu_exec, oracle, c)
    assert rc == 0
    assert_parity(got, want, c)


@pytest.mark.parametrize("scenario", ["contraction", "everything", "hadamard_contraction", "unary"])
@pytest.mark.parametrize("dtype", [tc.F32, tc.C64])
def test_device_pointers_equal_host_pointers(lib, gpu_exec, oracle, scenario, dtype):
    c = T.random_case(77, scenario, dtype, max_ext=6)
    rc_h, d_h, want = run_both(lib, gpu_exec, oracle, c)
    rc_d, d_d = run_device(lib, gpu_exec, c)
    assert rc_d == rc_h
    if rc_h == 0:
        assert_parity(d_d, want, c)


def _gemm_case(dtype, M, N, K, seed, transA=False, transB=False, ldpad=0, conj=(0, 0), beta=0.5):
    a_order = [1, 0] if transA else None
    b_order = [1, 0] if transB else None
    A = T.make_operand(dtype, [M, K], "mk", seed, order=a_order, pad=ldpad, op=conj[0])
    B = T.make_operand(dtype, [K, N], "kn", seed + 1, order=b_order, pad=ldpad, op=conj[1])
    C = T.make_operand(dtype, [M, N], "mn", seed + 2)
    D = T.make_operand(dtype, [M, N], "mn", seed + 3)
    alpha = 0.75 - 0.25j if dtype in T.COMPLEX_TYPES else 0.75
    return T.Case(A, B, C, D, tc.DEFAULT_PREC, alpha, beta, name=f"gemm{M}x{N}x{K}")


@pytest.mark.parametrize("dtype", [tc.F32, tc.F64, tc.C32, tc.C64])
@pytest.mark.parametrize("kernel", [tc.KERNEL_SIMPLE, tc.KERNEL_MFMA, tc.KERNEL_SPLITK, tc.KERNEL_MFMA_SMALL])
def test_forced_kernels_on_gemm_shapes(lib, oracle, dtype, kernel):
    _, ex = lib.create_executor()
    assert lib.attr_set_int(ex, tc.ATTR_KERNEL, kernel) == 0
    shapes = [(1, 1, 1), (17, 33, 5), (64, 64, 64), (129, 70, 37), (200, 130, 257)]
    for i, (M, N, K) in enumerate(shapes):
        if kernel == tc.KERNEL_SPLITK and M * N > 4096:
            continue
        for tA, tB in ((False, False), (True, True), (True, False)):
            c = _gemm_case(dtype, M, N, K, 10 * i + 1, tA, tB, ldpad=i % 2,
                           conj=(1, 0) if dtype in T.COMPLEX_TYPES else (0, 0))
            rc, got, want = run_both(lib, ex, oracle, c)
            assert rc == 0
            assert_parity(got, want, c)
            ran = kernel if kernel != tc.KERNEL_MFMA_SMALL or dtype == tc.F32 else tc.KERNEL_MFMA
            assert lib.attr_get_int(ex, tc.ATTR_LAST_KERNEL)[1] == ran
    lib.destroy_executor(ex)


SPLIT_ATTR = {"split_bf16": 1, "split6_bf16": 2, "split_f16": 3}
SPLIT_TOL = {"split_bf16": 1e-5, "split6_bf16": 1e-6, "split_f16": 1e-6}


@pytest.mark.parametrize("mode", ["bf16", "f16", "split_bf16", "split6_bf16", "split_f16"])
def test_low_precision_kernel_load_modes(lib, oracle, mode):
    """The 16-bit-input GETT on GEMM shapes whose layouts reach every pair of
    load modes (rows contiguous, k contiguous, gathered), with ragged edges
    and beta != 0: same input rounding as the oracle, so only the FP32
    summation order differs; split-BF16 against the FP64 oracle."""
    prec = {"bf16": tc.BF16BF16_ACCUM_F32, "f16": tc.F16F16_ACCUM_F32}.get(mode, tc.DEFAULT_PREC)
    _, ex = lib.create_executor()
    assert lib.attr_set_int(ex, tc.ATTR_KERNEL, tc.KERNEL_MFMA) == 0
    if mode.startswith("split"):
        assert lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, SPLIT_ATTR[mode]) == 0
    shapes = [(200, 130, 257), (129, 70, 37), (256, 384, 96), (64, 64, 64)]
    for i, (M, N, K) in enumerate(shapes):
        for tA, tB, pad in ((False, False, 0), (True, True, 0), (True, False, 0), (False, True, 1)):
            c = _gemm_case(tc.F32, M, N, K, 100 * i + 7, tA, tB, ldpad=pad)
            c.prec = prec
            rc, got = T.run_lib(lib, c, ex)
            assert rc == 0
            assert lib.attr_get_int(ex, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_MFMA
            if mode.startswith("split"):
                _, exact = oracle.run(c, f64=True)
                assert T.normwise_error(got, exact, T.written_mask(c.D)) < SPLIT_TOL[mode], (M, N, K, tA, tB, pad)
            else:
                _, want = oracle.run(c, gpu=True)
                assert_parity(got, want, c, tol=1e-5)
    lib.destroy_executor(ex)


def test_split6_bf16_below_fp32_error_at_long_k(lib, oracle):
    """Three-piece split-BF16 (six products, compensated sums of the FP32
    K-tile sums) against FP64 at K = 16384, where FP32 accumulation itself
    errs by about 1e-5: below 1e-6, and below the plain FP32 kernel's error.
    Also through the one-lane-per-output kernel on a thin shape."""
    _, ex6 = lib.create_executor()
    assert lib.attr_set_int(ex6, tc.ATTR_SPLIT_BF16, 2) == 0
    assert lib.attr_get_int(ex6, tc.ATTR_SPLIT_BF16) == (0, 2)
    _, ex32 = lib.create_executor()
    c = _gemm_case(tc.F32, 160, 136, 16384, 910, beta=0.0)
    _, exact = oracle.run(c, f64=True)
    mask = T.written_mask(c.D)
    rc, got6 = T.run_lib(lib, c, ex6)
    assert rc == 0 and lib.attr_get_int(ex6, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_MFMA
    rc, got32 = T.run_lib(lib, c, ex32)
    assert rc == 0
    err6, err32 = T.normwise_error(got6, exact, mask), T.normwise_error(got32, exact, mask)
    assert err6 < 1e-6, (err6, err32)
    assert err6 < err32, (err6, err32)
    thin = _gemm_case(tc.F32, 3, 700, 96, 911, beta=0.25)
    rc, got = T.run_lib(lib, thin, ex6)
    assert rc == 0 and lib.attr_get_int(ex6, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_SIMPLE
    _, exact = oracle.run(thin, f64=True)
    assert T.normwise_error(got, exact, T.written_mask(thin.D)) < 1e-6
    lib.destroy_executor(ex6)
    lib.destroy_executor(ex32)


def _with_split(lib, value):
    _, ex = lib.create_executor()
    assert lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, value) == 0
    return ex


def test_split_f16_below_1e6_at_long_k(lib, oracle):
    """Split-F16 (x = hi + lo in binary16, three products, compensated sums
    of the FP32 K-tile sums) against FP64 at K = 16384: below 1e-6 and below
    the plain FP32 kernel's own error."""
    ex16, ex32 = _with_split(lib, 3), _with_split(lib, 0)
    assert lib.attr_get_int(ex16, tc.ATTR_SPLIT_BF16) == (0, 3)
    c = _gemm_case(tc.F32, 160, 136, 16384, 930, beta=0.0)
    _, exact = oracle.run(c, f64=True)
    mask = T.written_mask(c.D)
    rc, got16 = T.run_lib(lib, c, ex16)
    assert rc == 0 and lib.attr_get_int(ex16, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_MFMA
    rc, got32 = T.run_lib(lib, c, ex32)
    err16, err32 = T.normwise_error(got16, exact, mask), T.normwise_error(got32, exact, mask)
    assert err16 < 1e-6 and err16 < err32, (err16, err32)
    lib.destroy_executor(ex16)
    lib.destroy_executor(ex32)


@pytest.mark.parametrize("sa,sb", [(1e6, 1.0), (1.0, 1e-7), (3e9, 2e-9), (1e-12, 1e-9), (4e4, 3.0)])
def test_split_f16_rescales_out_of_range_operands(lib, oracle, sa, sb):
    """Operands whose max |x| leaves [2^-3, 2^15) (binary16 overflow above,
    subnormal pieces below) are split after a power-of-two scaling: the
    error against FP64 stays below 1e-6 at every magnitude."""
    ex = _with_split(lib, 3)
    c = _gemm_case(tc.F32, 136, 72, 2048, 940, beta=0.0)
    c.A.buf *= np.float32(sa)
    c.B.buf *= np.float32(sb)
    c.alpha = 1.0
    _, exact = oracle.run(c, f64=True)
    rc, got = T.run_lib(lib, c, ex)
    assert rc == 0
    assert np.isfinite(got).all()
    assert T.normwise_error(got, exact, T.written_mask(c.D)) < 1e-6, (sa, sb)
    lib.destroy_executor(ex)


def test_split_f16_falls_back_bitwise(lib):
    """Where split-F16 does not apply, D is bit-identical to the form that
    runs instead: plain FP32 for inf / NaN inputs, for magnitudes below
    2^-100 and for products small enough for the one-lane kernel; the
    three-piece split-BF16 for D = A B + beta D in place (a second launch
    would read the overwritten D)."""
    ex16, ex32, ex6 = _with_split(lib, 3), _with_split(lib, 0), _with_split(lib, 2)

    def same(case, other):
        rc, a = T.run_lib(lib, case, ex16)
        rc2, b = T.run_lib(lib, case, other)
        assert rc == 0 and rc2 == 0
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), case.name

    # the fallback is the FP32 MFMA GETT (plain FP32 by rule would take the
    # small-tile kernel at these sizes)
    ex32m = _with_split(lib, 0)
    assert lib.attr_set_int(ex32m, tc.ATTR_KERNEL, tc.KERNEL_MFMA) == 0
    c = _gemm_case(tc.F32, 136, 72, 512, 950)
    c.A.buf[c.A.offset + 5] = np.inf
    same(c, ex32m)
    c = _gemm_case(tc.F32, 136, 72, 512, 951)
    c.B.buf[c.B.offset + 7] = np.nan
    same(c, ex32m)
    c = _gemm_case(tc.F32, 136, 72, 512, 952)
    c.A.buf *= np.float32(1e-32)
    same(c, ex32m)
    lib.destroy_executor(ex32m)
    c = _gemm_case(tc.F32, 3, 700, 96, 953, beta=0.25)
    same(c, ex32)
    assert lib.attr_get_int(ex16, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_SIMPLE
    c = _gemm_case(tc.C32, 96, 80, 64, 955, conj=(1, 0))   # complex: the C32 MFMA kernel in FP32
    rc, a = T.run_lib(lib, c, ex16)
    rc2, b = T.run_lib(lib, c, ex32)
    assert rc == 0 and rc2 == 0 and np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert lib.attr_get_int(ex16, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_MFMA

    def in_place(case, ex):
        D = case.D.copy()
        infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in (case.A, case.B, D, D)]
        _, plan = lib.create_plan([0, 0, 0, 0], infos, [case.A.idx, case.B.idx, D.idx, D.idx], case.prec)
        al, be = T.scalar_bytes(case.alpha, tc.F32), T.scalar_bytes(case.beta, tc.F32)
        rc = lib.execute(plan, ex, al.ctypes.data, case.A.ptr(), case.B.ptr(), be.ctypes.data, D.ptr(), D.ptr())
        lib.destroy_plan(plan)
        for h in infos:
            lib.destroy_info(h)
        assert rc == 0
        return D.buf

    c = _gemm_case(tc.F32, 136, 72, 512, 954, beta=0.5)
    assert np.array_equal(in_place(c, ex16).view(np.uint32), in_place(c, ex6).view(np.uint32))
    for ex in (ex16, ex32, ex6):
        lib.destroy_executor(ex)


@pytest.mark.parametrize("dtype,M", [(tc.F64, 512), (tc.C32, 512), (tc.C64, 256)])
def test_large_tile_variants(lib, gpu_exec, oracle, dtype, M):
    """Batched products with at least 4 large-tile workgroups per CU (1024
    of them on 256 CUs) run the large-tile MFMA kernels; the small-tile ones
    are what the shapes of the other tests reach."""
    L, K = 64, 12
    A = T.make_operand(dtype, [M, K, L], "mkl", 81)
    B = T.make_operand(dtype, [L, K, M], "lkn", 82, op=tc.CONJUGATE if dtype in T.COMPLEX_TYPES else 0)
    C = T.make_operand(dtype, [M, M, L], "mnl", 83)
    D = T.make_operand(dtype, [M, M, L], "mnl", 84)
    alpha, beta = (0.5 - 0.25j, 0.75 + 0.5j) if dtype in T.COMPLEX_TYPES else (0.5, 0.75)
    c = T.Case(A, B, C, D, tc.DEFAULT_PREC, alpha, beta, name=f"large_tiles_{dtype}")
    rc, got, want = run_both(lib, gpu_exec, oracle, c)
    assert rc == 0
    assert lib.attr_get_int(gpu_exec, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_MFMA
    assert_parity(got, want, c)


@pytest.mark.parametrize("dtype", [tc.F64, tc.C32, tc.C64])
def test_tile_sizes_give_the_same_bits(lib, oracle, dtype):
    """F64 and complex products on each MFMA tile size (TAPP_AMD_ATTR_TILES
    1 large, 2 small, 3 tall: the small tile's columns and twice its rows),
    with ragged M, N and K, a Hadamard batch and conj(B): every size gives
    the oracle's result within the compute type's bound and the same bits as
    the others."""
    M, N, K, L = 75, 41, 19, 3
    A = T.make_operand(dtype, [M, K, L], "mkl", 91)
    B = T.make_operand(dtype, [K, N, L], "knl", 92, op=tc.CONJUGATE if dtype in T.COMPLEX_TYPES else 0)
    C = T.make_operand(dtype, [M, N, L], "mnl", 93)
    D = T.make_operand(dtype, [M, N, L], "mnl", 94, pad=1)
    alpha, beta = (0.5 - 0.25j, 0.75 + 0.5j) if dtype in T.COMPLEX_TYPES else (0.5, 0.75)
    c = T.Case(A, B, C, D, tc.DEFAULT_PREC, alpha, beta, name=f"tile_sizes_{dtype}")
    _, want = oracle.run(c, gpu=True)
    outs, names = [], []
    for v in (1, 2, 3):
        _, ex = lib.create_executor()
        assert lib.attr_set_int(ex, tc.ATTR_KERNEL, tc.KERNEL_MFMA) == 0
        assert lib.attr_set_int(ex, tc.ATTR_TILES, v) == 0
        rc, got = T.run_lib(lib, c, ex)
        assert rc == 0
        names.append(lib.last_kernel_name(ex)[1])
        lib.destroy_executor(ex)
        assert_parity(got, want, c)
        outs.append(got)
    assert len(set(names)) == 3, names
    for got in outs[1:]:
        assert np.array_equal(got.view(np.uint8), outs[0].view(np.uint8))


@pytest.mark.parametrize("mode", ["plain", "bf16", "f16", "split_bf16", "split6_bf16", "split_f16"])
def test_split_k_few_tiles(lib, oracle, mode):
    """Few 128 x 128 output tiles and long K: the FP32 MFMA kernel splits K
    over workgroups and a second kernel sums the partials in split order.
    Ragged K (not a multiple of the 32-deep K tile), a Hadamard mode,
    beta != 0 and a transposed operand are covered, and each product run
    again gives the same bits."""
    prec = {"bf16": tc.BF16BF16_ACCUM_F32, "f16": tc.F16F16_ACCUM_F32}.get(mode, tc.DEFAULT_PREC)
    _, ex = lib.create_executor()
    # forced: by rule these shapes take the small-tile kernel in plain FP32
    assert lib.attr_set_int(ex, tc.ATTR_KERNEL, tc.KERNEL_MFMA) == 0
    if mode.startswith("split"):
        assert lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, SPLIT_ATTR[mode]) == 0
    for i, (M, N, K, L) in enumerate([(100, 90, 301, 3), (256, 256, 256, 1), (40, 200, 1000, 2)]):
        A = T.make_operand(tc.F32, [M, K, L], "mkl", 40 + i, order=[1, 0, 2] if i == 1 else None)
        B = T.make_operand(tc.F32, [K, N, L], "knl", 50 + i)
        C = T.make_operand(tc.F32, [M, N, L], "mnl", 60 + i)
        D = T.make_operand(tc.F32, [M, N, L], "mnl", 70 + i, pad=1)
        c = T.Case(A, B, C, D, prec, 0.75, -0.5, name=f"splitk{M}x{N}x{K}x{L}_{mode}")
        rc, got, want = run_both(lib, ex, oracle, c)
        assert rc == 0
        assert lib.attr_get_int(ex, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_MFMA
        rc, again = T.run_lib(lib, c, ex)
        assert rc == 0 and np.array_equal(again.view(np.uint32), got.view(np.uint32)), c.name
        if mode.startswith("split"):
            _, exact = oracle.run(c, f64=True)
            assert T.normwise_error(got, exact, T.written_mask(c.D)) < SPLIT_TOL[mode]
        else:
            assert_parity(got, want, c)
    lib.destroy_executor(ex)


def test_mixed_storage_types(lib, gpu_exec, oracle):
    combos = [(tc.F32, tc.F64, tc.F32, tc.F64), (tc.F32, tc.C32, tc.C32, tc.C32), (tc.C64, tc.F64, tc.C64, tc.C32),
              (tc.F64, tc.F64, tc.F32, tc.F32), (tc.F16, tc.F16, tc.F32, tc.F32), (tc.BF16, tc.F32, tc.F32, tc.BF16)]
    for i, (ta, tb, tcc, td) in enumerate(combos):
        c = T.random_case(900 + i, "everything", tc.F32, max_ext=4, beta_zero=False)
        for o, t in zip(c.operands(), (ta, tb, tcc, td)):
            lo, hi = o.span()
            o.type = t
            o.buf = T.fill(t, max(1, hi - lo + 1), 900 + i + t)
            o.op = 1 if t in T.COMPLEX_TYPES else 0
        c.alpha, c.beta = (0.5 + 0.25j, -0.75 + 0.5j) if td in T.COMPLEX_TYPES else (0.5, -0.75)
        rc, got, want = run_both(lib, gpu_exec, oracle, c)
        assert rc == 0
        tol = TOL[td] if td in (tc.F16, tc.BF16) else (1e-5 if td in (tc.F32, tc.C32) else 1e-6)
        assert_parity(got, want, c, tol=tol)


@pytest.mark.parametrize("prec", [tc.F16F16_ACCUM_F32, tc.BF16BF16_ACCUM_F32, tc.F16F16_ACCUM_F16])
def test_low_precision_forms(lib, gpu_exec, oracle, prec):
    for seed in range(3):
        c = T.random_case(60 + seed, "contraction", tc.F32, prec=prec, max_ext=8, min_ext=4)
        rc, got, want = run_both(lib, gpu_exec, oracle, c)
        assert rc == 0
        # same input rounding as the oracle; only the FP32 summation order differs
        assert_parity(got, want, c, tol=1e-5)
        # against FP64 on the unrounded inputs: 2 u_in + gamma_K(u_f32)
        rc64, exact = oracle.run(c, f64=True)
        K = int(np.prod([e for e, l in zip(c.A.extents, c.A.idx) if l in c.B.idx and l not in c.D.idx]))
        u_in = 2.0 ** -11 if prec != tc.BF16BF16_ACCUM_F32 else 2.0 ** -8
        u32 = 2.0 ** -24
        bound = 2 * u_in + K * u32 / (1 - K * u32)
        mask = T.written_mask(c.D)
        # normwise with the magnitude sum as denominator scale
        assert T.normwise_error(got, exact, mask) <= 8 * bound


def _abs_case(case):
    """The case with |alpha|, |beta| and every element replaced by its
    magnitude (complex: modulus) in binary64: its FP64 result is
    |beta| |C| + sum_k |alpha| |sum_A| |sum_B| elementwise, the scale of a
    componentwise error bound."""
    c = case.copy()
    for o in c.operands():
        v = np.abs(o.buf.astype(np.float32) if o.type == tc.F16 else
                   ((o.buf.astype(np.uint32) << 16).view(np.float32) if o.type == tc.BF16 else o.buf))
        o.buf = v.astype(np.float64)
        o.type = tc.F64
        o.op = 0
    c.alpha, c.beta = abs(c.alpha), abs(c.beta)
    c.prec = tc.F64F64_ACCUM_F64
    return c


@pytest.mark.parametrize("prec", [tc.F16F16_ACCUM_F32, tc.F16F16_ACCUM_F16])
@pytest.mark.parametrize("scenario", ["contraction", "unary", "everything", "hadamard_contraction"])
def test_f16_precisions_against_reference_semantics(lib, gpu_exec, oracle, prec, scenario):
    """TAPP_F16F16_ACCUM_F32 / _F16 on the GPU against the oracle's
    restatement of the reference (bit-exact to its clang F16 build), which
    rounds every alpha * sum_A * sum_B to binary16 and, for ACCUM_F16,
    accumulates in binary16.  The GPU rounds the inputs and the unary sums
    exactly as the reference does (checked at 1e-5 against the oracle's GPU
    form) but keeps products and sums in binary32, which datatype.h:85-87
    allows.  Elementwise bound: |D_gpu - D_ref| <= 2 (2 u16 + gamma_K(u_acc)
    + gamma_K(u32)) (|beta C| + sum_k |alpha sum_A sum_B|), u_acc the
    reference's accumulator unit."""
    u16, u32 = 2.0 ** -11, 2.0 ** -24
    for seed in range(3):
        c = T.random_case(7100 + 17 * seed, scenario, tc.F32, prec=prec, max_ext=6, min_ext=2, beta_zero=seed == 2)
        rc, got = T.run_lib(lib, c, gpu_exec)
        assert rc == 0
        _, gpu_form = oracle.run(c, gpu=True)
        _, ref_form = oracle.run(c)
        mask = T.written_mask(c.D)
        assert T.normwise_error(got, gpu_form, mask) <= 1e-5, c.name
        _, mag = oracle.run(_abs_case(c), f64=True)
        K = max(1, int(np.prod([e for e, l in zip(c.A.extents, c.A.idx) if l in c.B.idx and l not in c.D.idx]
                                or [1])))
        gam = lambda u: K * u / (1 - K * u)
        u_acc = u16 if prec == tc.F16F16_ACCUM_F16 else u32
        bound = 2 * (2 * u16 + gam(u_acc) + gam(u32)) * np.asarray(mag)[mask] + 1e-30
        diff = np.abs(np.asarray(got)[mask].astype(np.float64) - np.asarray(ref_form)[mask].astype(np.float64))
        assert np.all(diff <= bound), (c.name, float(np.max(diff / bound)))


def test_split_bf16_accuracy(lib, oracle):
    _, ex = lib.create_executor()
    assert lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, 1) == 0
    for seed in range(3):
        c = _gemm_case(tc.F32, 96, 80, 200, 700 + seed, beta=0.0)
        rc, got = T.run_lib(lib, c, ex)
        _, exact = oracle.run(c, f64=True)
        assert rc == 0
        err = T.normwise_error(got, exact, T.written_mask(c.D))
        assert err < 1e-5, err
    lib.destroy_executor(ex)


def test_batched_product_equals_single_calls(lib, gpu_exec, oracle):
    cases = [T.random_case(5, "hadamard_contraction", tc.F64, max_ext=5, beta_zero=False, layouts=False)]
    base = cases[0]
    for s in range(1, 3):
        c = base.copy()
        for o in c.operands():
            o.buf = T.fill(o.type, len(o.buf), 5000 + s * 7 + o.type)
        cases.append(c)
    bufs = [c.copy() for c in cases]
    infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in base.operands()]
    _, plan = lib.create_plan([o.op for o in base.operands()], infos, [o.idx for o in base.operands()], base.prec)
    al = T.scalar_bytes(base.alpha, base.D.type)
    be = T.scalar_bytes(base.beta, base.D.type)
    rc = lib.execute_batched(plan, gpu_exec, al.ctypes.data, [b.A.ptr() for b in bufs], [b.B.ptr() for b in bufs],
                             be.ctypes.data, [b.C.ptr() for b in bufs], [b.D.ptr() for b in bufs])
    assert rc == 0
    for b, c in zip(bufs, cases):
        _, want = oracle.run(c, gpu=True)
        assert_parity(b.D.buf, want, c)
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)


def test_zero_extents_and_pinned_output(lib, gpu_exec, oracle):
    # contracted extent 0: D = beta * C
    c = _gemm_case(tc.F32, 5, 4, 0, 3)
    rc, got, want = run_both(lib, gpu_exec, oracle, c)
    assert rc == 0
    assert_parity(got, want, c)
    # D extent 0: nothing written
    c = _gemm_case(tc.F32, 0, 4, 3, 4)
    rc, got, want = run_both(lib, gpu_exec, oracle, c)
    assert rc == 0 and np.array_equal(got, c.D.buf)
    # one-mode D with stride 0: the last coordinate's value remains
    A = T.make_operand(tc.F64, [6, 3], "ak", 1)
    B = T.make_operand(tc.F64, [3], "k", 2)
    C = T.Operand(tc.F64, [6], [0], T.labels("a"), 0, T.fill(tc.F64, 1, 3), 0)
    D = T.Operand(tc.F64, [6], [0], T.labels("a"), 0, T.fill(tc.F64, 1, 4), 0)
    c = T.Case(A, B, C, D, tc.DEFAULT_PREC, 1.5, 0.25, name="pinned")
    rc, got, want = run_both(lib, gpu_exec, oracle, c)
    assert rc == 0
    assert np.allclose(got, want, rtol=1e-12, atol=0)
    # unary extent 0: every product term is zero
    A = T.make_operand(tc.F32, [4, 3, 0], "akx", 5)
    B = T.make_operand(tc.F32, [3, 2], "kn", 6)
    C = T.make_operand(tc.F32, [4, 2], "an", 7)
    D = T.make_operand(tc.F32, [4, 2], "an", 8)
    c = T.Case(A, B, C, D, tc.DEFAULT_PREC, 1.0, 2.0, name="unary0")
    rc, got, want = run_both(lib, gpu_exec, oracle, c)
    assert rc == 0
    assert_parity(got, want, c)


def test_layout_change_after_plan_creation(lib, gpu_exec, oracle):
    """Layouts are held by handle: TAPP_set_strides after plan creation
    changes the next execute (reference product.c:86-107)."""
    c = _gemm_case(tc.F32, 40, 30, 20, 9)
    infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in c.operands()]
    _, plan = lib.create_plan([o.op for o in c.operands()], infos, [o.idx for o in c.operands()], c.prec)
    al = T.scalar_bytes(c.alpha, c.D.type)
    be = T.scalar_bytes(c.beta, c.D.type)
    for transposed in (False, True):
        cc = c.copy()
        if transposed:
            # store A transposed in the same buffer and tell the plan through its info
            M, K = c.A.extents
            a2 = c.A.buf[:M * K].reshape(K, M).T.copy().reshape(-1)
            cc.A.buf = a2.astype(np.float32)
            cc.A.strides = [K, 1]
            s = (ctypes.c_int64 * 2)(K, 1)
            assert lib.lib.TAPP_set_strides(infos[0], s) == 0
        rc = lib.execute(plan, gpu_exec, al.ctypes.data, cc.A.ptr(), cc.B.ptr(), be.ctypes.data, cc.C.ptr(),
                         cc.D.ptr())
        assert rc == 0
        _, want = oracle.run(cc, gpu=True)
        assert_parity(cc.D.buf, want, cc)
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)


def test_reference_style_executor_int(lib, oracle):
    """A caller's bare int executor (the reference's representation,
    executor.c:9-18) runs on the default device context; other ints are 33."""
    c = T.random_case(3, "contraction", tc.F32)
    kind = ctypes.c_int(1)
    rc, got = T.run_lib(lib, c, ctypes.addressof(kind))
    _, want = oracle.run(c, gpu=True)
    assert rc == 0
    assert_parity(got, want, c)
    bad = ctypes.c_int(5)
    rc, _ = T.run_lib(lib, c, ctypes.addressof(bad))
    assert rc == 33


def _nan_c(case):
    """beta = 0 and C full of NaN: the reference skips C entirely when beta
    is zero (product.c:351-360), so no NaN may reach D."""
    case.beta = 0.0
    nan = complex(float("nan"), float("nan")) if case.C.type in T.COMPLEX_TYPES else float("nan")
    case.C.buf[:] = nan
    return case


@pytest.mark.parametrize("dtype,prec,M,N,K,kernel", [
    (tc.F32, tc.DEFAULT_PREC, 3, 70, 40, tc.KERNEL_SIMPLE),        # thin: one lane per output
    (tc.F32, tc.DEFAULT_PREC, 4, 4, 2048, tc.KERNEL_SPLITK),       # few outputs, long K
    (tc.F32, tc.DEFAULT_PREC, 200, 130, 257, tc.KERNEL_MFMA),      # split-K over the FP32 MFMA kernel
    (tc.F32, tc.DEFAULT_PREC, 640, 520, 48, tc.KERNEL_MFMA),       # FP32 MFMA kernel without split-K
    (tc.F32, tc.BF16BF16_ACCUM_F32, 200, 130, 257, tc.KERNEL_MFMA),
    (tc.F64, tc.DEFAULT_PREC, 96, 80, 33, tc.KERNEL_MFMA),
    (tc.C64, tc.DEFAULT_PREC, 64, 48, 20, tc.KERNEL_MFMA),
])
def test_beta_zero_ignores_nan_in_c(lib, oracle, dtype, prec, M, N, K, kernel):
    _, ex = lib.create_executor()
    assert lib.attr_set_int(ex, tc.ATTR_KERNEL, kernel) == 0
    c = _nan_c(_gemm_case(dtype, M, N, K, 1234))
    c.prec = prec
    rc, got, want = run_both(lib, ex, oracle, c)
    assert rc == 0
    mask = T.written_mask(c.D)
    assert np.all(np.isfinite(np.asarray(got)[mask])), "NaN from C reached D"
    assert np.all(np.isfinite(np.asarray(want)[mask]))
    assert_parity(got, want, c, tol=1e-5 if dtype in (tc.F32, tc.C32) else 1e-12)
    assert lib.attr_get_int(ex, tc.ATTR_LAST_KERNEL)[1] == kernel
    lib.destroy_executor(ex)


def test_beta_zero_ignores_nan_in_c_in_group(lib, gpu_exec, oracle):
    from test_group import DevCase, make_group
    cases = [_nan_c(T.tucker_case(40 + i, 12, 9, 7, 3)) for i in range(3)]
    cases += [_nan_c(T.four_index_case(50 + i, (6, 7, 8, 5, 9, 4))) for i in range(3)]
    dcs = [DevCase(lib, c) for c in cases]
    rc, g = make_group(lib, gpu_exec, dcs)
    assert rc == 0
    assert lib.execute_group(g) == 0
    for d in dcs:
        _, want = oracle.run(d.case, gpu=True)
        got = d.d()
        assert np.all(np.isfinite(got[T.written_mask(d.case.D)]))
        assert_parity(got, want, d.case)
    lib.destroy_group(g)
    for d in dcs:
        d.close()

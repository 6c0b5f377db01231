"""Pins the oracle (oracle/tapp_oracle.c) before it is trusted as the
checker: against the compiled reference library (bit-exact), the
reference's exported plan helpers, the known answers of the reference's
own programs, library rounding routines and closed forms."""
import ctypes
import json
import os

import numpy as np
import pytest

import tapp_cases as T

tc = T.tc
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.mark.parametrize("dtype", [tc.F32, tc.F64, tc.C32, tc.C64])
@pytest.mark.parametrize("prec", [tc.DEFAULT_PREC, tc.F32F32_ACCUM_F32, tc.F64F64_ACCUM_F64])
@pytest.mark.parametrize("scenario", sorted(T.SCENARIOS))
def test_oracle_bitexact_vs_reference(oracle, ref, dtype, prec, scenario):
    for seed in range(3):
        c = T.random_case(1000 + seed * 17, scenario, dtype, prec=prec, repeated_D=(seed == 2))
        rc_r, d_r = T.run_lib(ref, c)
        rc_o, d_o = oracle.run(c)
        assert rc_r == rc_o
        if rc_r == 0:
            assert same_bits(d_r, d_o), (c.name, np.abs(d_r - d_o).max())


def test_oracle_four_index16_golden(oracle):
    g = np.load(os.path.join(GOLDEN, "ref_four_index16.npz"))
    rc, d = oracle.run(T.four_index16_case(seed=int(g["seed"])))
    assert rc == 0
    assert same_bits(d, g["D"])


def test_oracle_c64_feature_golden_sample(oracle):
    """The complex-double feature product (a diagonal, unary labels on both sides,
    conj(B), negative and padded strides): the oracle's D equals the
    compiled reference's, bit for bit, at 4096 sampled positions."""
    g = np.load(os.path.join(GOLDEN, "ref_c64_feature_sample.npz"))
    rc, d = oracle.run(T.c64_feature_case(seed=int(g["seed"])))
    assert rc == 0
    assert same_bits(d[g["idx"]], g["D"])


def _case_from_kat(k):
    ops = []
    for name in ("A", "B", "C", "D_in"):
        o = k[name]
        dt = o["type"]
        vals = np.array(o["data"] if o["data"] is not None else [0.0], dtype=np.float64)
        if dt in (tc.C32, tc.C64):
            vals = vals[0::2] + 1j * vals[1::2]
        buf = vals.astype(T.NP_DTYPE[dt])
        ops.append(T.Operand(dt, o["extents"], o["strides"], o["idx"], o["op"], buf, -o["offset"]))
    dt = k["D_in"]["type"]

    def scal(v):
        return complex(v[0], v[1]) if dt in (tc.C32, tc.C64) else v[0]
    return T.Case(*ops, k["prec"], scal(k["alpha"]), scal(k["beta"]), c_null=k["C"]["data"] is None)


def load_kats():
    with open(os.path.join(GOLDEN, "kat_reference_programs.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("k", load_kats(), ids=lambda k: f"{k['program']}-{k['call']}")
def test_oracle_reproduces_reference_programs(oracle, k):
    c = _case_from_kat(k)
    rc, d = oracle.run(c)
    assert rc == k["rc"]
    want = np.array(k["D_out"]["data"], dtype=np.float64)
    got = np.asarray(d)
    if got.dtype.kind == "c":
        got = np.stack([got.real, got.imag], axis=-1).reshape(-1)
    assert np.array_equal(got.astype(np.float64), want)


def test_exercise_documented_answer():
    # exercise_contraction_answers.c:229-236: 53.090/61.090 grid
    k = [k for k in load_kats() if k["program"] == "exercise_contraction_answers"][-1]
    d = k["D_out"]
    v = np.array(d["data"])
    grid = [[v[i * d["strides"][0] + j * d["strides"][1] - d["offset"]] for j in range(3)] for i in range(3)]
    assert np.allclose(grid, k["documented_expected_rows"], atol=5e-4)


# --- validation codes ----------------------------------------------------
def negative_cases():
    """(name, mutate) pairs, each producing one of the reference's error codes."""
    base = lambda: T.random_case(5, "hadamard_contraction", tc.F32, layouts=False, min_ext=2, beta_zero=False)  # noqa

    def ext(o, i, v):
        o.extents[i] = v
        return o

    out = []
    c = base(); c.D.idx[0] = 999; out.append(("code4", c))
    c = base(); c.A.idx.append(c.A.idx[0]); c.A.extents.append(c.A.extents[0] + 1); c.A.strides.append(1); out.append(("code9", c))
    c = base(); c.B.idx.append(c.B.idx[0]); c.B.extents.append(c.B.extents[0] + 1); c.B.strides.append(1); out.append(("code10", c))
    c = base()
    shared = [l for l in c.A.idx if l in c.B.idx]
    i = c.A.idx.index(shared[0]); ext(c.A, i, c.A.extents[i] + 1); out.append(("code1or2", c))
    c = base()
    j = [n for n, l in enumerate(c.B.idx) if l in c.D.idx and l not in c.A.idx][0]
    ext(c.B, j, c.B.extents[j] + 1); out.append(("code3", c))
    c = base(); c.C.idx = c.C.idx[:-1]; c.C.extents = c.C.extents[:-1]; c.C.strides = c.C.strides[:-1]; out.append(("code5", c))
    c = base(); c.C.idx = c.C.idx[::-1]; out.append(("code6", c))
    c = base(); c.C.extents[0] += 1; out.append(("code7", c))
    c = base(); c.D.strides[1] = 0; out.append(("code8_zero", c))
    c = base(); c.D.strides = [1] * len(c.D.strides); out.append(("code8_overlap", c))
    c = base(); c.c_null = True; out.append(("code12", c))
    c = base(); c.D.strides = [1 << 32] + c.D.strides[1:]; out.append(("int32_truncation", c))
    return out


@pytest.mark.parametrize("name,case", negative_cases(), ids=lambda x: x if isinstance(x, str) else "")
def test_oracle_validation_codes_match_reference(oracle, ref, name, case):
    lib = ref
    infos = []
    for o in case.operands():
        rc, h = lib.create_info(o.type, o.extents, o.strides)
        infos.append(h)
    _, plan = lib.create_plan([o.op for o in case.operands()], infos, [o.idx for o in case.operands()], case.prec)
    _, ex = lib.create_executor()
    # D never written on a validation failure: pass no data pointers at all
    be = T.scalar_bytes(case.beta, case.D.type)
    want = None
    if name != "int32_truncation":
        want = lib.execute(plan, ex, None, None, None, be.ctypes.data, None if case.c_null else 1, None)
    lib.destroy_executor(ex)
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)
    got = oracle.check(case)
    if want is not None:
        assert got == want and got != 0, (name, got, want)
    else:
        assert got == 0   # the reference accepts it: low 32 bits of 2^32 are 0 < 1*e
    assert oracle.check(case, exec_kind=None) in (got, 33) if got == 0 else True


def test_oracle_executor_codes(oracle):
    c = T.random_case(3, "contraction", tc.F32)
    assert oracle.check(c, exec_kind=None) == 33
    assert oracle.check(c, exec_kind=7) == 33
    for k in (1, 2, 12):
        assert oracle.check(c, exec_kind=k) == 0


# --- plan stage vs the reference's own exported helpers ---------------------
def test_oracle_classification_vs_reference_helpers(oracle):
    if not os.path.exists(T.REF_LIB):
        pytest.skip("compiled reference not built")
    R = ctypes.CDLL(T.REF_LIB, mode=os.RTLD_LOCAL)
    P64 = ctypes.POINTER(ctypes.c_int64)
    PP64 = ctypes.POINTER(P64)
    R.compress_repeated_indices.argtypes = [ctypes.POINTER(ctypes.c_int), PP64, PP64, PP64]
    R.extract_binary_contractions_indices.argtypes = [ctypes.c_int] * 3 + [P64] * 3 + [PP64]
    R.extract_unary_contracted_indices.argtypes = [ctypes.c_int, P64, ctypes.c_int, P64, ctypes.c_int, P64, PP64]
    R.extract_free_strides.argtypes = [ctypes.c_int, P64, P64, ctypes.c_int, P64, PP64]
    R.extract_contracted_strides.argtypes = [ctypes.c_int, P64, P64, ctypes.c_int, P64, PP64]
    libc = ctypes.CDLL(None)
    libc.malloc.restype = ctypes.c_void_p
    libc.malloc.argtypes = [ctypes.c_size_t]

    def heap(vals):
        p = ctypes.cast(libc.malloc(8 * max(1, len(vals))), P64)
        for i, v in enumerate(vals):
            p[i] = v
        return p

    def compress(o):
        n = ctypes.c_int(o.nmode)
        idx, ext, st = heap(o.idx), heap(o.extents), heap(o.strides)
        R.compress_repeated_indices(ctypes.byref(n), ctypes.byref(idx), ctypes.byref(ext), ctypes.byref(st))
        k = n.value
        return k, idx, [idx[i] for i in range(k)], ext, st

    for scen in sorted(T.SCENARIOS):
        for seed in range(5):
            c = T.random_case(seed + 77, scen, tc.F64, repeated_D=(seed % 2 == 0))
            if oracle.check(c) != 0:
                continue
            nA, iA, _, eA, sA = compress(c.A)
            nB, iB, _, eB, sB = compress(c.B)
            nD, iD, _, eD, sD = compress(c.D)
            con = P64()
            ncon = R.extract_binary_contractions_indices(nA, nB, nD, iA, iB, iD, ctypes.byref(con))
            ua = P64()
            nua = R.extract_unary_contracted_indices(nA, iA, nB, iB, nD, iD, ctypes.byref(ua))
            ub = P64()
            nub = R.extract_unary_contracted_indices(nB, iB, nA, iA, nD, iD, ctypes.byref(ub))
            fa, fb, ka, kb = P64(), P64(), P64(), P64()
            R.extract_free_strides(nA, iA, sA, nD, iD, ctypes.byref(fa))
            R.extract_free_strides(nB, iB, sB, nD, iD, ctypes.byref(fb))
            R.extract_contracted_strides(nA, iA, sA, ncon, con, ctypes.byref(ka))
            R.extract_contracted_strides(nB, iB, sB, ncon, con, ctypes.byref(kb))
            rc, o = oracle.classify(c)
            assert rc == 0
            assert o.n_contracted == ncon and o.n_unary_A == nua and o.n_unary_B == nub and o.nmode_D == nD
            assert [o.contracted[i] for i in range(ncon)] == [con[i] for i in range(ncon)]
            assert [o.unary_A[i] for i in range(nua)] == [ua[i] for i in range(nua)]
            assert [o.unary_B[i] for i in range(nub)] == [ub[i] for i in range(nub)]
            assert [o.free_strides_A[i] for i in range(nD)] == [fa[i] for i in range(nD)]
            assert [o.free_strides_B[i] for i in range(nD)] == [fb[i] for i in range(nD)]
            assert [o.contracted_strides_A[i] for i in range(ncon)] == [ka[i] for i in range(ncon)]
            assert [o.contracted_strides_B[i] for i in range(ncon)] == [kb[i] for i in range(ncon)]
            assert [o.strides_D[i] for i in range(nD)] == [sD[i] for i in range(nD)]


# --- rounding known answers ---------------------------------------------------
def test_round_f16_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-9, 5, 20000),
                        np.array([0.0, -0.0, 65504.0, 65519.0, 65520.0, 6.1e-5, 5.96e-8, 2.98e-8, 1.0 + 2 ** -11,
                                  1.0 + 3 * 2 ** -11], dtype=np.float32)]).astype(np.float32)
    got = np.array([oracle.lib.orc_round_f16(float(v)) for v in x], dtype=np.float32)
    want = x.astype(np.float16).astype(np.float32)
    assert np.array_equal(got, want)


def test_round_bf16_matches_torch(oracle):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000)).astype(np.float32)
    x = np.concatenate([x, np.array([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 3.4e38], dtype=np.float32)])
    got = np.array([oracle.lib.orc_round_bf16(float(v)) for v in x], dtype=np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(got, want)


# --- FP64 accuracy oracle ------------------------------------------------------
def test_f64_oracle_equals_f64_path_on_f64_data(oracle):
    for seed in range(6):
        c = T.random_case(seed + 300, "everything", tc.F64, prec=tc.F64F64_ACCUM_F64)
        rc1, d1 = oracle.run(c)
        rc2, d2 = oracle.run(c, f64=True)
        assert rc1 == rc2 == 0
        assert same_bits(d1, d2)


def test_f64_oracle_closed_form_rank1(oracle):
    # A[m,k] = x_m y_k, B[k,n] = y_k z_n  =>  D[m,n] = alpha x_m z_n sum_k y_k^2
    M, N, K = 7, 5, 33
    x, y, z = T.uniform(1, M), T.uniform(2, K), T.uniform(3, N)
    A = T.Operand(tc.F64, [M, K], [1, M], T.labels("mk"), 0, np.outer(x, y).reshape(-1, order="F").copy(), 0)
    B = T.Operand(tc.F64, [K, N], [1, K], T.labels("kn"), 0, np.outer(y, z).reshape(-1, order="F").copy(), 0)
    C = T.Operand(tc.F64, [M, N], [1, M], T.labels("mn"), 0, np.zeros(M * N), 0)
    D = T.Operand(tc.F64, [M, N], [1, M], T.labels("mn"), 0, np.zeros(M * N), 0)
    c = T.Case(A, B, C, D, tc.F64F64_ACCUM_F64, 1.25, 0.0)
    rc, d = oracle.run(c, f64=True)
    want = (1.25 * np.outer(x, z) * np.dot(y, y)).reshape(-1, order="F")
    assert rc == 0 and np.allclose(d, want, rtol=1e-14, atol=0)


@pytest.fixture(scope="module")
def ref16():
    if not os.path.exists(T.REF_F16_LIB):
        pytest.skip("F16 reference build missing")
    return tc.TappLib(T.REF_F16_LIB)


F16_PINNED = [(tc.F16F16_ACCUM_F32, d) for d in (tc.F32, tc.F64, tc.C32, tc.C64, tc.F16)] + \
             [(tc.F16F16_ACCUM_F16, d) for d in (tc.F32, tc.F64, tc.F16)] + [(tc.DEFAULT_PREC, tc.F16)]


@pytest.mark.parametrize("prec,dtype", F16_PINNED)
def test_oracle_f16_semantics_bit_exact_vs_reference_f16_build(oracle, ref16, prec, dtype):
    """The binary16 branches (product.c:1094, 1956-1970, 3272-3366,
    10412-10615; DEFAULT_PREC with binary16 storage): every scenario class,
    with unary labels, diagonals and beta != 0, gives the same bits as the
    reference compiled by clang with TAPP_REFERENCE_ENABLE_F16."""
    for scen in sorted(T.SCENARIOS):
        for seed in range(4):
            c = T.random_case(seed * 13 + 7, scen, dtype, prec=prec, max_ext=5, repeated_D=(seed == 3))
            rc1, d1 = T.run_lib(ref16, c)
            rc2, d2 = oracle.run(c)
            assert rc1 == rc2, (scen, seed)
            assert same_bits(d1, d2), (scen, seed, T.normwise_error(d2, d1, T.written_mask(c.D)))


def _mixed_case(i, seed, types, prec):
    c = T.random_case(900 + i * 5 + seed, "everything", tc.F32, max_ext=4, beta_zero=(seed == 2), prec=prec)
    for o, t in zip(c.operands(), types):
        lo, hi = o.span()
        o.type = t
        o.buf = T.fill(t, max(1, hi - lo + 1), 900 + i + t + seed)
        o.op = 1 if t in T.COMPLEX_TYPES else 0
    c.alpha, c.beta = (0.5 + 0.25j, -0.75 + 0.5j) if types[3] in T.COMPLEX_TYPES else (0.5, -0.75)
    if seed == 2:
        c.beta = 0.0
    return c


MIXED16 = [(tc.F16, tc.F16, tc.F32, tc.F32), (tc.F16, tc.F32, tc.F32, tc.F16), (tc.F32, tc.F16, tc.F16, tc.F64),
           (tc.F64, tc.F16, tc.F32, tc.F16), (tc.F16, tc.C32, tc.C32, tc.C32), (tc.C64, tc.F16, tc.C64, tc.C32),
           (tc.F16, tc.F16, tc.F16, tc.F16)]


@pytest.mark.parametrize("prec", [tc.DEFAULT_PREC, tc.F16F16_ACCUM_F32, tc.F16F16_ACCUM_F16, tc.F32F32_ACCUM_F32,
                                  tc.F64F64_ACCUM_F64])
def test_oracle_mixed_storage_with_binary16_vs_reference_f16_build(oracle, ref16, prec):
    """Mixed storage types with binary16 operands: the slot types of every
    value (sum_A in A's type under DEFAULT_PREC, binary16 under the F16
    precisions, ...) give the reference's bits.  Left out: explicit
    precisions with one real and one complex operand (the reference zeroes
    sum_B with is_complex_A, product.c:381, reading or writing past the
    allocation: undefined) and F16F16_ACCUM_F16 with complex D (a 2-byte
    accumulator written as complex, product.c:1094)."""
    for i, types in enumerate(MIXED16):
        cplx = [t in T.COMPLEX_TYPES for t in types]
        if prec != tc.DEFAULT_PREC and cplx[0] != cplx[1]:
            continue
        if prec == tc.F16F16_ACCUM_F16 and cplx[3]:
            continue
        for seed in range(3):
            c = _mixed_case(i, seed, types, prec)
            rc1, d1 = T.run_lib(ref16, c)
            rc2, d2 = oracle.run(c)
            assert rc1 == rc2 == 0, (types, seed)
            assert same_bits(d1, d2), (types, seed, T.normwise_error(d2, d1, T.written_mask(c.D)))


def test_f16_build_deviates_on_complex_double_D_with_real_C(oracle, ref, ref16):
    """One path where the clang build with TAPP_REFERENCE_ENABLE_F16 differs
    from the plain build on the same source: DEFAULT_PREC, complex binary64
    D and a real C.  The F16 build's beta * C carries an error of about
    1e-12 in its real part that the plain build (and plain arithmetic) does
    not have; the oracle follows the plain build there, bit for bit."""
    c = _mixed_case(9, 0, (tc.F32, tc.F32, tc.F32, tc.C64), tc.DEFAULT_PREC)
    c.beta = -0.75 + 0.5j
    _, d_plain = T.run_lib(ref, c)
    _, d_16 = T.run_lib(ref16, c)
    _, d_o = oracle.run(c)
    assert same_bits(d_plain, d_o)
    assert not same_bits(d_plain, d_16)
    assert T.normwise_error(d_16, d_plain, T.written_mask(c.D)) < 1e-10


def test_oracle_generic_loop_equals_typed_loops(oracle):
    """The slot-precision loop (used whenever a value is binary16 or
    bfloat16) on binary32 / binary64 / complex data gives the same bits as
    the per-type loops, which are pinned to the compiled reference."""
    for dtype in (tc.F32, tc.F64, tc.C32, tc.C64):
        for prec in (tc.DEFAULT_PREC, tc.F32F32_ACCUM_F32, tc.F64F64_ACCUM_F64):
            for scen in sorted(T.SCENARIOS):
                for seed in range(2):
                    c = T.random_case(seed * 7 + 3, scen, dtype, prec=prec, max_ext=5, repeated_D=seed == 1)
                    r1, d1 = oracle.run(c)
                    r2, d2 = oracle.run(c, flags=oracle.GENERIC)
                    assert r1 == r2 and same_bits(d1, d2), (dtype, prec, scen, seed)


TYPES4 = (tc.F32, tc.F64, tc.C32, tc.C64)


@pytest.mark.parametrize("prec", [tc.DEFAULT_PREC, tc.F32F32_ACCUM_F32, tc.F64F64_ACCUM_F64])
def test_oracle_mixed_storage_types_bit_exact_vs_reference(oracle, ref, prec):
    """Every combination of F32/F64/C32/C64 storage over A, B, C and D (not
    all equal): each value keeps the reference's slot type (sum_A in A's
    type under DEFAULT_PREC, binary32 under F32F32_ACCUM_F32, ...), and the
    result has the compiled reference's bits.  Explicit precisions with one
    real and one complex operand are left out (product.c:381 zeroes sum_B
    with is_complex_A: undefined)."""
    import itertools
    for i, types in enumerate(itertools.product(TYPES4, repeat=4)):
        if len(set(types)) == 1:
            continue
        cplx = [t in T.COMPLEX_TYPES for t in types]
        if prec != tc.DEFAULT_PREC and cplx[0] != cplx[1]:
            continue
        c = _mixed_case(i % 50, i % 3, types, prec)
        rc1, d1 = T.run_lib(ref, c)
        rc2, d2 = oracle.run(c)
        assert rc1 == rc2 == 0 and same_bits(d1, d2), (types, T.normwise_error(d2, d1, T.written_mask(c.D)))


def _rounded_inputs(case, rnd):
    """The case with A, B, C, alpha and beta rounded by `rnd` (binary32 in,
    binary32 out) and the binary32 precision: what the GPU's 16-bit forms
    compute when no unary labels are present."""
    c = case.copy()
    for o in (c.A, c.B, c.C):
        if o.type == tc.F32:
            o.buf = rnd(o.buf).astype(np.float32)
    c.alpha = float(rnd(np.array([c.alpha], dtype=np.float32))[0])
    c.beta = float(rnd(np.array([c.beta], dtype=np.float32))[0])
    c.prec = tc.F32F32_ACCUM_F32
    return c


@pytest.mark.parametrize("prec", [tc.F16F16_ACCUM_F32, tc.F16F16_ACCUM_F16, tc.BF16BF16_ACCUM_F32])
def test_gpu_products_variant(oracle, ref16, prec):
    """ORC_GPU_PRODUCTS (the GPU library's 16-bit semantics) without unary
    labels is exactly the binary32 oracle on rounded inputs -- so it is
    pinned through the binary32 path -- and it differs from the
    reference's F16 semantics (per-product binary16 rounding), which the
    faithful oracle reproduces."""
    rnd = (lambda x: np.asarray(x, dtype=np.float32).astype(np.float16).astype(np.float32)) \
        if prec != tc.BF16BF16_ACCUM_F32 else T.bf16_round
    differ = 0
    for seed in range(6):
        c = T.random_case(seed + 40, "hadamard_contraction" if seed % 2 else "contraction", tc.F32, prec=prec,
                          max_ext=6, beta_zero=seed == 4)
        _, gp = oracle.run(c, flags=oracle.GPU_PRODUCTS)
        _, f32 = oracle.run(_rounded_inputs(c, rnd))
        assert same_bits(gp, f32), seed
        if prec != tc.BF16BF16_ACCUM_F32:
            _, faithful = oracle.run(c)
            _, r16 = T.run_lib(ref16, c)
            assert same_bits(faithful, r16)
            differ += not same_bits(gp, faithful)
    assert prec == tc.BF16BF16_ACCUM_F32 or differ == 6


def test_f16_accum_f16_complex_D_is_not_written(oracle, ref16):
    """TAPP_F16F16_ACCUM_F16 with complex D: the reference's assign_D stores
    nothing (product.c:11096-11099), and neither does the oracle."""
    c = T.random_case(5, "contraction", tc.C32, prec=tc.F16F16_ACCUM_F16, max_ext=4)
    rc1, d1 = T.run_lib(ref16, c)
    rc2, d2 = oracle.run(c)
    assert rc1 == rc2 == 0
    assert same_bits(d1, c.D.buf) and same_bits(d2, c.D.buf)

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


def test_oracle_configs0_golden(oracle):
    g = np.load(os.path.join(GOLDEN, "ref_configs0.npz"))
    rc, d = oracle.run(T.configs0_case(seed=int(g["seed"])))
    assert rc == 0
    assert same_bits(d, g["D"])


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


def test_oracle_f16_precision_vs_reference_f16_build(oracle):
    if not os.path.exists(T.REF_F16_LIB):
        pytest.skip("F16 reference build missing")
    ref16 = tc.TappLib(T.REF_F16_LIB)
    for seed in range(3):
        c = T.random_case(seed + 40, "contraction", tc.F32, prec=tc.F16F16_ACCUM_F32, max_ext=6)
        rc1, d1 = T.run_lib(ref16, c)
        rc2, d2 = oracle.run(c)
        assert rc1 == rc2 == 0
        # both round inputs to binary16 and accumulate in binary32; the
        # reference's _Float16 products may keep excess precision
        assert T.normwise_error(d2, d1, T.written_mask(c.D)) < 2e-3

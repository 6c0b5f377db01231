"""Test-side model of a TAPP product: operands with layouts and data,
a portable seeded generator, and runners for

  * any library exporting the TAPP API (this package's GPU library, or the
    compiled reference in oracle/_ref) through tapp_ctypes, and
  * the oracle (oracle/build/libtapp_oracle.so).

Data generator: SplitMix64 (Steele, Lea, Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014); u = (x >> 11) * 2^-53 in
[0, 1), value = 2u - 1, rounded once to the storage type.
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
PKG = os.path.join(ROOT, "reference-implementation_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)
import tapp_ctypes as tc  # noqa: E402

ORACLE_LIB = os.path.join(ROOT, "oracle", "build", "libtapp_oracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libtapp-reference.so")
REF_F16_LIB = os.path.join(ROOT, "oracle", "_ref", "libtapp-reference-f16.so")

NP_DTYPE = {tc.F32: np.float32, tc.F64: np.float64, tc.C32: np.complex64, tc.C64: np.complex128,
            tc.F16: np.float16, tc.BF16: np.uint16}
REAL_OF = {tc.F32: np.float32, tc.F64: np.float64, tc.C32: np.float32, tc.C64: np.float64,
           tc.F16: np.float16, tc.BF16: np.float32}
COMPLEX_TYPES = (tc.C32, tc.C64)


# --------------------------------------------------------------------------
# generator
# --------------------------------------------------------------------------
def splitmix64(seed, n):
    """n successive SplitMix64 outputs from state `seed` (uint64 array)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n):
    """n values in [-1, 1) as float64."""
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)
    return 2.0 * u - 1.0


def bf16_round(x):
    """float32 -> nearest-even bfloat16 value, as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) & np.uint64(0xFFFF0000)
    return r.astype(np.uint32).view(np.float32)


def fill(dtype, n, seed):
    if dtype in COMPLEX_TYPES:
        v = uniform(seed, 2 * n)
        return (v[0::2] + 1j * v[1::2]).astype(NP_DTYPE[dtype])
    v = uniform(seed, n)
    if dtype == tc.BF16:
        return (bf16_round(v.astype(np.float32)).view(np.uint32) >> 16).astype(np.uint16)
    return v.astype(NP_DTYPE[dtype])


# --------------------------------------------------------------------------
# operands
# --------------------------------------------------------------------------
class Operand:
    """A tensor as the API sees it: storage type, layout, labels, element op,
    and a flat storage buffer with element (0, ..., 0) at buf[offset]."""

    def __init__(self, dtype, extents, strides, idx, op=0, buf=None, offset=0):
        self.type = dtype
        self.extents = [int(e) for e in extents]
        self.strides = [int(s) for s in strides]
        self.idx = [int(i) for i in idx]
        self.op = op
        self.buf = buf
        self.offset = offset

    @property
    def nmode(self):
        return len(self.extents)

    def span(self):
        lo = hi = 0
        for e, s in zip(self.extents, self.strides):
            if e == 0:
                return 0, -1
            d = (e - 1) * s
            if d < 0:
                lo += d
            else:
                hi += d
        return lo, hi

    def ptr(self):
        if self.buf is None:
            return None
        return self.buf.ctypes.data + self.offset * self.buf.itemsize

    def copy(self):
        return Operand(self.type, self.extents, self.strides, self.idx, self.op,
                       None if self.buf is None else self.buf.copy(), self.offset)


def labels(s):
    return [ord(c) for c in s]


def layout(extents, order=None, pad=0, negative=(), dense_stride=1):
    """Strides for `extents` stored with mode `order` (slowest last), each
    stride padded by `pad` elements, modes in `negative` stored backwards.
    Returns (strides, offset_of_element0, buffer_length)."""
    n = len(extents)
    order = list(range(n)) if order is None else list(order)
    strides = [0] * n
    s = dense_stride
    for m in order:
        strides[m] = s
        s = s * max(1, extents[m]) + pad
    for m in negative:
        strides[m] = -strides[m]
    lo = hi = 0
    for e, st in zip(extents, strides):
        if e == 0:
            return strides, 0, 1
        d = (e - 1) * st
        if d < 0:
            lo += d
        else:
            hi += d
    return strides, -lo, hi - lo + 1


def make_operand(dtype, extents, idx, seed, order=None, pad=0, negative=(), op=0, strides=None):
    if strides is None:
        strides, off, n = layout(extents, order, pad, negative)
    else:
        o = Operand(dtype, extents, strides, idx if not isinstance(idx, str) else labels(idx))
        lo, hi = o.span()
        off, n = -lo, max(1, hi - lo + 1)
    buf = fill(dtype, n, seed)
    return Operand(dtype, extents, strides, idx if not isinstance(idx, str) else labels(idx), op, buf, off)


class Case:
    def __init__(self, A, B, C, D, prec=tc.DEFAULT_PREC, alpha=1.0, beta=0.0, c_null=False, name=""):
        self.A, self.B, self.C, self.D = A, B, C, D
        self.prec = prec
        self.alpha, self.beta = alpha, beta
        self.c_null = c_null
        self.name = name

    def operands(self):
        return [self.A, self.B, self.C, self.D]

    def copy(self):
        return Case(self.A.copy(), self.B.copy(), self.C.copy(), self.D.copy(), self.prec, self.alpha, self.beta,
                    self.c_null, self.name)


def scalar_bytes(value, dtype):
    """The scalar as the library reads it: stored as type_D."""
    if dtype in COMPLEX_TYPES:
        arr = np.array([complex(value)], dtype=NP_DTYPE[dtype])
    elif dtype == tc.BF16:
        arr = (bf16_round(np.array([value], dtype=np.float32)).view(np.uint32) >> 16).astype(np.uint16)
    else:
        arr = np.array([complex(value).real], dtype=NP_DTYPE[dtype])
    return arr


# --------------------------------------------------------------------------
# runners
# --------------------------------------------------------------------------
def run_lib(lib, case, exec_=None, own_exec=True):
    """Runs the case through a TAPP library with host pointers.  Returns
    (rc, D buffer after the call)."""
    case = case.copy()
    infos = []
    for o in case.operands():
        rc, h = lib.create_info(o.type, o.extents, o.strides)
        assert rc == 0, rc
        infos.append(h)
    rc, plan = lib.create_plan([o.op for o in case.operands()], infos, [o.idx for o in case.operands()], case.prec)
    assert rc == 0
    made = False
    if exec_ is None:
        rc, exec_ = lib.create_executor()
        made = True
    al = scalar_bytes(case.alpha, case.D.type)
    be = scalar_bytes(case.beta, case.D.type)
    c_ptr = None if case.c_null else case.C.ptr()
    rc = lib.execute(plan, exec_, al.ctypes.data, case.A.ptr(), case.B.ptr(), be.ctypes.data, c_ptr, case.D.ptr())
    if made:
        lib.destroy_executor(exec_)
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)
    return rc, case.D.buf


class _OrcTensor(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("nmode", ctypes.c_int), ("extents", ctypes.POINTER(ctypes.c_int64)),
                ("strides", ctypes.POINTER(ctypes.c_int64)), ("idx", ctypes.POINTER(ctypes.c_int64)),
                ("op", ctypes.c_int)]


class _OrcClasses(ctypes.Structure):
    _fields_ = [("n_contracted", ctypes.c_int), ("n_unary_A", ctypes.c_int), ("n_unary_B", ctypes.c_int),
                ("nmode_D", ctypes.c_int)] + [
        (name, ctypes.c_int64 * 64) for name in (
            "contracted", "unary_A", "unary_B", "ext_contracted", "free_strides_A", "free_strides_B",
            "contracted_strides_A", "contracted_strides_B", "unary_strides_A", "unary_strides_B",
            "strides_D", "strides_C", "extents_D")]


class Oracle:
    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built: run `make -C oracle`")
        self.lib = ctypes.CDLL(path, mode=os.RTLD_LOCAL | os.RTLD_NOW)
        P = ctypes.POINTER(_OrcTensor)
        self.lib.orc_execute.argtypes = [P, P, P, P, ctypes.c_int] + [ctypes.c_void_p] * 6 + [
            ctypes.POINTER(ctypes.c_int)]
        self.lib.orc_execute_ex.argtypes = [P, P, P, P, ctypes.c_int] + [ctypes.c_void_p] * 6 + [
            ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        self.lib.orc_execute_f64.argtypes = [P, P, P, P, ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_int]
        self.lib.orc_check.argtypes = [P, P, P, P, ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        self.lib.orc_classify.argtypes = [P, P, P, P, ctypes.POINTER(_OrcClasses)]
        self.lib.orc_round_bf16.argtypes = [ctypes.c_float]
        self.lib.orc_round_bf16.restype = ctypes.c_float
        self.lib.orc_round_f16.argtypes = [ctypes.c_float]
        self.lib.orc_round_f16.restype = ctypes.c_float

    @staticmethod
    def _t(o, keep):
        e = (ctypes.c_int64 * max(1, o.nmode))(*o.extents)
        s = (ctypes.c_int64 * max(1, o.nmode))(*o.strides)
        i = (ctypes.c_int64 * max(1, o.nmode))(*o.idx)
        keep += [e, s, i]
        return _OrcTensor(o.type, o.nmode, e, s, i, o.op)

    def tensors(self, case, keep):
        return [ctypes.byref(self._t(o, keep)) for o in case.operands()]

    GENERIC = 1        # ORC_GENERIC: the slot-precision loop for every precision
    GPU_PRODUCTS = 2   # ORC_GPU_PRODUCTS: the GPU library's 16-bit semantics (binary32 products)

    def run(self, case, exec_kind=1, f64=False, round_inputs=0, flags=0, gpu=False):
        """gpu=True: the GPU library's semantics for products that involve
        16-bit values (its one deviation from the reference there: products
        and accumulation in binary32, see ORC_GPU_PRODUCTS); the same as
        gpu=False for every other product."""
        if gpu and not f64:
            flags |= gpu_semantics_flags(case)
        case = case.copy()
        keep = []
        ts = self.tensors(case, keep)
        al = scalar_bytes(case.alpha, case.D.type)
        be = scalar_bytes(case.beta, case.D.type)
        c_ptr = None if case.c_null else case.C.ptr()
        if f64:
            rc = self.lib.orc_execute_f64(*ts, case.prec, al.ctypes.data, case.A.ptr(), case.B.ptr(), be.ctypes.data,
                                          c_ptr, case.D.ptr(), round_inputs)
        else:
            kind = None if exec_kind is None else ctypes.byref(ctypes.c_int(exec_kind))
            rc = self.lib.orc_execute_ex(*ts, case.prec, al.ctypes.data, case.A.ptr(), case.B.ptr(), be.ctypes.data,
                                         c_ptr, case.D.ptr(), kind, flags)
        return rc, case.D.buf

    def check(self, case, exec_kind=1):
        keep = []
        ts = self.tensors(case, keep)
        be = scalar_bytes(case.beta, case.D.type)
        c_ptr = None if case.c_null else case.C.ptr()
        kind = None if exec_kind is None else ctypes.byref(ctypes.c_int(exec_kind))
        return self.lib.orc_check(*ts, be.ctypes.data, c_ptr, kind)

    def classify(self, case):
        keep = []
        out = _OrcClasses()
        rc = self.lib.orc_classify(*self.tensors(case, keep), ctypes.byref(out))
        return rc, out


LOWP_PRECS = (tc.F16F16_ACCUM_F32, tc.F16F16_ACCUM_F16, tc.BF16BF16_ACCUM_F32)


def gpu_semantics_flags(case):
    """Oracle flags for what the GPU computes: binary32 products and
    accumulation wherever the reference would keep them in 16 bits."""
    lowp = case.prec in LOWP_PRECS or any(o.type in (tc.F16, tc.BF16) for o in case.operands())
    return Oracle.GPU_PRODUCTS if lowp else 0


# --------------------------------------------------------------------------
# comparisons
# --------------------------------------------------------------------------
def written_mask(D):
    """Boolean mask over D.buf of the elements the product writes (the
    compressed diagonal of repeated labels counts once)."""
    mask = np.zeros(len(D.buf), dtype=bool)
    # unique labels: first occurrence keeps its extent; stride summed
    seen = {}
    for e, s, l in zip(D.extents, D.strides, D.idx):
        if l in seen:
            seen[l][1] += s
        else:
            seen[l] = [e, s]
    off = np.zeros(1, dtype=np.int64)
    for e, s in seen.values():
        if e == 0:
            return mask
        off = (off[:, None] + np.arange(e, dtype=np.int64)[None, :] * s).reshape(-1)
    mask[off + D.offset] = True
    return mask


def normwise_error(got, ref, mask=None):
    g = np.asarray(got)
    r = np.asarray(ref)
    if mask is not None:
        g, r = g[mask], r[mask]
    if g.dtype == np.uint16:
        g = (g.astype(np.uint32) << 16).view(np.float32)
        r = (r.astype(np.uint32) << 16).view(np.float32)
    g = g.astype(np.complex128)
    r = r.astype(np.complex128)
    if r.size == 0:
        return 0.0
    den = np.max(np.abs(r))
    num = np.max(np.abs(g - r))
    if den == 0:
        return float(num)
    return float(num / den)


# --------------------------------------------------------------------------
# the benchmark configurations as cases
# --------------------------------------------------------------------------
def four_index16_case(seed=1, dtype=tc.F32):
    """D[a,b,c,d] = 1.5 sum_{e,f} A[a,e,b,f] B[d,f,c,e] + 0.5 C[a,b,c,d],
    all extents 16, dense column-major, values uniform [-1, 1)."""
    e = [16] * 4
    A = make_operand(dtype, e, "aebf", seed * 4 + 0)
    B = make_operand(dtype, e, "dfce", seed * 4 + 1)
    C = make_operand(dtype, e, "abcd", seed * 4 + 2)
    D = make_operand(dtype, e, "abcd", seed * 4 + 3)
    return Case(A, B, C, D, tc.DEFAULT_PREC, 1.5, 0.5, name="four_index16")


def c64_feature_case(seed=1, scale=1):
    """Complex double feature coverage:
    D[a,b,h,i] = alpha sum_{c,u,v} A[a,c,c,h,u] conj(B[h,b,c,i,v]) + beta C[a,b,h,i]
    h=32 Hadamard, c=24 contracted and repeated in A, a=40, b=40, i=32 free,
    u=8 unary in A, v=6 unary in B; A stored with a negative stride on a, B
    with padded strides.  `scale` divides every extent (for small tests)."""
    a, b, c, h, i, u, v = [max(1, x // scale) for x in (40, 40, 24, 32, 32, 8, 6)]
    A = make_operand(tc.C64, [a, c, c, h, u], "acchu", seed * 4 + 0, negative=(0,))
    B = make_operand(tc.C64, [h, b, c, i, v], "hbciv", seed * 4 + 1, pad=1, op=tc.CONJUGATE)
    C = make_operand(tc.C64, [a, b, h, i], "abhi", seed * 4 + 2)
    D = make_operand(tc.C64, [a, b, h, i], "abhi", seed * 4 + 3)
    return Case(A, B, C, D, tc.F64F64_ACCUM_F64, 0.7 - 0.3j, 1.1 + 0.2j, name="c64_feature")


def ladder_case(seed=1, v=24, o=8, prec=tc.DEFAULT_PREC, beta=0.0):
    """The precision-ladder contraction D[a,b,i,j] = alpha sum_{c,d}
    A[a,b,c,d] B[c,i,d,j] + beta C[a,b,i,j] (a CCSD-like particle-particle
    ladder; B permuted so that c and d are not adjacent), dense column-major,
    virtual extents a..d = v and occupied i, j = o.  Entries are N(0, 1)
    exp(-|a - c| / 16) in A and N(0, 1) exp(-|c - d| / 16) in B, each scaled
    so that max |x| <= 8."""
    rng = np.random.default_rng(seed)
    idx = np.arange(v, dtype=np.float64)
    decay = np.exp(-np.abs(idx[:, None] - idx[None, :]) / 16.0)
    a = rng.standard_normal((v, v, v, v)) * decay[:, None, :, None]          # [a, b, c, d]: decay over (a, c)
    b = rng.standard_normal((v, o, v, o)) * decay[:, None, :, None]          # [c, i, d, j]: decay over (c, d)
    ops = []
    for x in (a, b):
        m = np.abs(x).max()
        if m > 8.0:
            x = x * (8.0 / m)
        ops.append(x.astype(np.float32))
    a, b = ops
    A = Operand(tc.F32, [v, v, v, v], layout([v] * 4)[0], labels("abcd"), 0, a.reshape(-1, order="F").copy(), 0)
    B = Operand(tc.F32, [v, o, v, o], layout([v, o, v, o])[0], labels("cidj"), 0, b.reshape(-1, order="F").copy(), 0)
    C = make_operand(tc.F32, [v, v, o, o], "abij", seed * 4 + 2)
    D = make_operand(tc.F32, [v, v, o, o], "abij", seed * 4 + 3)
    return Case(A, B, C, D, prec, 1.0, beta, name=f"ladder{v}x{o}")


def run_reference(path, case):
    lib = tc.TappLib(path)
    rc, d = run_lib(lib, case)
    assert rc == 0, rc
    return d


# --------------------------------------------------------------------------
# random scenarios (the scenario classes of the reference's test.cpp:27-56)
# --------------------------------------------------------------------------
SCENARIOS = {
    # name: (free_A, free_B, hadamard, contracted, unary_A, unary_B, repeated_A, repeated_B)
    "contraction": (2, 2, 0, 2, 0, 0, 0, 0),
    "hadamard": (0, 0, 2, 0, 0, 0, 0, 0),
    "outer_product": (2, 2, 0, 0, 0, 0, 0, 0),
    "full_contraction": (0, 0, 0, 3, 0, 0, 0, 0),
    "zero_dim_A": (0, 2, 0, 0, 0, 0, 0, 0),
    "one_dim": (1, 0, 0, 1, 0, 0, 0, 0),
    "hadamard_free": (1, 1, 1, 0, 0, 0, 0, 0),
    "hadamard_contraction": (1, 1, 1, 1, 0, 0, 0, 0),
    "unary": (1, 1, 0, 1, 1, 1, 0, 0),
    "repeated": (1, 1, 0, 1, 0, 0, 1, 1),
    "everything": (1, 1, 1, 1, 1, 1, 1, 1),
}


def random_case(seed, scenario="contraction", dtype=tc.F32, max_ext=4, prec=tc.DEFAULT_PREC, layouts=True,
                repeated_D=False, beta_zero=None, min_ext=1):
    rng = np.random.default_rng(seed)
    fa, fb, ha, co, ua, ub, ra, rb = SCENARIOS[scenario]
    next_label = [100]

    def new(n):
        out = list(range(next_label[0], next_label[0] + n))
        next_label[0] += n
        return out

    FA, FB, HA, CO, UA, UB = new(fa), new(fb), new(ha), new(co), new(ua), new(ub)
    ext = {l: int(rng.integers(min_ext, max_ext + 1)) for l in FA + FB + HA + CO + UA + UB}
    A_lab = FA + HA + CO + UA
    B_lab = FB + HA + CO + UB
    # repeated labels: duplicate some labels that already occur (diagonals)
    for _ in range(ra):
        if A_lab:
            A_lab.append(A_lab[int(rng.integers(len(A_lab)))])
    for _ in range(rb):
        if B_lab:
            B_lab.append(B_lab[int(rng.integers(len(B_lab)))])
    D_lab = FA + FB + HA
    if repeated_D and D_lab:
        D_lab.append(D_lab[int(rng.integers(len(D_lab)))])
    A_lab = list(rng.permutation(A_lab)) if A_lab else []
    B_lab = list(rng.permutation(B_lab)) if B_lab else []
    D_lab = list(rng.permutation(D_lab)) if D_lab else []

    def op_for(dt):
        return int(rng.integers(0, 2)) if dt in COMPLEX_TYPES else 0

    def mk(labs, s, out=False):
        e = [ext[l] for l in labs]
        if layouts and not out:
            order = list(rng.permutation(len(e))) if e else None
            pad = int(rng.choice([0, 0, 1, 3]))
            neg = tuple(i for i in range(len(e)) if rng.random() < 0.25)
        elif layouts and out:
            order = list(rng.permutation(len(e))) if e else None
            pad = int(rng.choice([0, 1]))
            neg = tuple(i for i in range(len(e)) if rng.random() < 0.25)
        else:
            order, pad, neg = None, 0, ()
        return make_operand(dtype, e, labs, s, order=order, pad=pad, negative=neg, op=op_for(dtype))

    A = mk(A_lab, seed * 8 + 1)
    B = mk(B_lab, seed * 8 + 2)
    C = mk(D_lab, seed * 8 + 3)
    D = mk(D_lab, seed * 8 + 4, out=True)
    if dtype in COMPLEX_TYPES:
        alpha = complex(*uniform(seed * 8 + 5, 2))
        beta = complex(*uniform(seed * 8 + 6, 2))
    else:
        alpha = float(uniform(seed * 8 + 5, 1)[0])
        beta = float(uniform(seed * 8 + 6, 1)[0])
    if beta_zero is None:
        beta_zero = rng.random() < 0.25
    c_null = False
    if beta_zero:
        beta = 0.0
        c_null = bool(rng.random() < 0.5)
    return Case(A, B, C, D, prec, alpha, beta, c_null, name=f"{scenario}/{seed}")


# --------------------------------------------------------------------------
# the batching workload (uneven products)
# --------------------------------------------------------------------------
def layout_mix_case(seed, dtype=tc.F32, prec=tc.DEFAULT_PREC, max_size=3e7):
    """A multi-tile binary contraction with every layout freedom at once:
    D[a,b,c,d,(h)] = alpha sum_{e,f} A[a,e,b,f,(h)] op(B[e,c,f,d,(h)]) + beta C,
    extents a..d ~ U{2..48}, e, f ~ U{2..40}, an optional Hadamard mode h,
    and each operand with its own label order, storage order, padding and
    sometimes one negative stride; D's storage order differs from C's.
    Complex cases conjugate B half the time; beta is zero 30 % of the time."""
    rng = np.random.default_rng(seed)
    while True:
        ext = {c: int(rng.integers(2, 49)) for c in "abcd"}
        ext.update({c: int(rng.integers(2, 41)) for c in "ef"})
        ext["h"] = int(rng.integers(1, 5)) if rng.random() < 0.5 else 0
        total = np.prod([ext[c] for c in "abcdef"]) * max(1, ext["h"])
        if 2e5 < total < max_size:
            break

    def operand(labs, s):
        labs = [x for x in labs if x != "h" or ext["h"]]
        perm = [str(x) for x in rng.permutation(labs)]             # label order
        order = [int(x) for x in rng.permutation(len(perm))]     # storage order of the modes
        neg = (int(rng.integers(len(perm))),) if rng.random() < 0.25 else ()
        pad = int(rng.choice([0, 0, 1, 3]))
        return make_operand(dtype, [ext[x] for x in perm], "".join(perm), s, order=order, pad=pad, negative=neg)

    A = operand("aebfh", seed * 4 + 1)
    B = operand("ecfdh", seed * 4 + 2)
    C = operand("abcdh", seed * 4 + 3)
    D = make_operand(dtype, C.extents, "".join(chr(x) for x in C.idx), seed * 4 + 4,
                     order=[int(x) for x in rng.permutation(len(C.extents))], pad=int(rng.choice([0, 2])))
    if dtype in COMPLEX_TYPES:
        B.op = int(rng.random() < 0.5)
        alpha = complex(rng.normal(), rng.normal())
        beta = 0.0 if rng.random() < 0.3 else complex(rng.normal(), rng.normal())
    else:
        alpha = float(rng.normal())
        beta = 0.0 if rng.random() < 0.3 else float(rng.normal())
    return Case(A, B, C, D, prec, alpha, beta, name=f"layout_mix{seed}_{dtype}")


def tucker_case(seed, i, j, k, r, dtype=tc.F32, prec=tc.DEFAULT_PREC):
    """D[r,j,k] = sum_i A[i,j,k] U[r,i] (a mode-1 Tucker product), dense."""
    A = make_operand(dtype, [i, j, k], "ijk", seed * 4 + 0)
    U = make_operand(dtype, [r, i], "ri", seed * 4 + 1)
    C = make_operand(dtype, [r, j, k], "rjk", seed * 4 + 2)
    D = make_operand(dtype, [r, j, k], "rjk", seed * 4 + 3)
    return Case(A, U, C, D, prec, 1.0, 0.0, name=f"tucker{i}x{j}x{k}r{r}")


def four_index_case(seed, ext, dtype=tc.F32, prec=tc.DEFAULT_PREC):
    """Four-index form D[a,b,c,d] = sum_{e,f} A[a,e,b,f] B[d,f,c,e] with
    per-label extents ext = (a, b, c, d, e, f)."""
    a, b, c, d, e, f = ext
    A = make_operand(dtype, [a, e, b, f], "aebf", seed * 4 + 0)
    B = make_operand(dtype, [d, f, c, e], "dfce", seed * 4 + 1)
    C = make_operand(dtype, [a, b, c, d], "abcd", seed * 4 + 2)
    D = make_operand(dtype, [a, b, c, d], "abcd", seed * 4 + 3)
    return Case(A, B, C, D, prec, 1.5, 0.5, name=f"four_index{ext}")


def batching_mix(seed, count, tucker_range=(4, 64), r_range=(2, 16), c0_range=(6, 24), tucker_frac=0.8,
                 prec=tc.DEFAULT_PREC, per_index=True):
    """The uneven batching workload: `tucker_frac` Tucker mode products with
    i, j, k ~ U{tucker_range} and r ~ U{r_range}, the rest four-index
    products with each extent ~ U{c0_range} (drawn per index)."""
    rng = np.random.default_rng(seed)
    cases = []
    for n in range(count):
        if rng.random() < tucker_frac:
            i, j, k = (int(x) for x in rng.integers(tucker_range[0], tucker_range[1] + 1, 3))
            r = int(rng.integers(r_range[0], r_range[1] + 1))
            cases.append(tucker_case(seed * 100003 + n, i, j, k, r, prec=prec))
        else:
            if per_index:
                ext = tuple(int(x) for x in rng.integers(c0_range[0], c0_range[1] + 1, 6))
            else:
                ext = (int(rng.integers(c0_range[0], c0_range[1] + 1)),) * 6
            cases.append(four_index_case(seed * 100003 + n, ext, prec=prec))
    return cases

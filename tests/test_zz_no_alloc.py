"""TAPP_AMD_execute_product_noalloc: the allocation-free variant of
TAPP_execute_product, whose tables, pre-pass buffers and split-K partial
sums live in one static device array per GPU.  Each product is run through
both entry points on the same device operands, and D must be bit-identical
(the same kernels on the same inputs); the noalloc calls come largest first,
so later calls reuse the scratch a larger one wrote.  Each is also checked
against the oracle."""
import ctypes

import numpy as np
import pytest

import tapp_cases as T

tc = T.tc
pytestmark = pytest.mark.gpu


class Dev:
    """A case with operands in device memory and a plan."""

    def __init__(self, lib, case):
        import torch
        self.lib, self.case = lib, case
        self.t = {n: torch.from_numpy(o.buf.view(np.uint8).copy()).cuda() for n, o in zip("ABCD", case.operands())}
        self.ptr = {n: self.t[n].data_ptr() + o.offset * o.buf.itemsize for n, o in zip("ABCD", case.operands())}
        self.infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in case.operands()]
        _, self.plan = lib.create_plan([o.op for o in case.operands()], self.infos,
                                       [o.idx for o in case.operands()], case.prec)
        self.al = T.scalar_bytes(case.alpha, case.D.type)
        self.be = T.scalar_bytes(case.beta, case.D.type)
        torch.cuda.synchronize()

    def args(self):
        return (self.al.ctypes.data, self.ptr["A"], self.ptr["B"], self.be.ctypes.data,
                None if self.case.c_null else self.ptr["C"], self.ptr["D"])

    def run(self, ex, noalloc):
        import torch
        self.t["D"].copy_(torch.from_numpy(self.case.D.buf.view(np.uint8).copy()).cuda())
        torch.cuda.synchronize()
        f = self.lib.execute_noalloc if noalloc else self.lib.execute
        rc = f(self.plan, ex, *self.args())
        return rc, self.t["D"].cpu().numpy().view(self.case.D.buf.dtype).copy()

    def close(self):
        self.lib.destroy_plan(self.plan)
        for h in self.infos:
            self.lib.destroy_info(h)


def _gemm(M, N, K, seed, beta=0.5):
    A = T.make_operand(tc.F32, [M, K], "mk", seed, order=[1, 0])
    B = T.make_operand(tc.F32, [K, N], "kn", seed + 1)
    C = T.make_operand(tc.F32, [M, N], "mn", seed + 2)
    D = T.make_operand(tc.F32, [M, N], "mn", seed + 3)
    return T.Case(A, B, C, D, tc.DEFAULT_PREC, 0.75, beta, name=f"gemm{M}x{N}x{K}")


def _products():
    """Largest scratch use first: split-K partials of a few-tile long-K
    product; the complex-double feature product at half size (pre-passes of both operands,
    C64); a layout mix; the unary/Hadamard/diagonal scenario in C32."""
    return [
        ("splitk", _gemm(256, 200, 6000, 11), 0),
        ("split_f16", _gemm(136, 72, 2048, 13, beta=0.0), 3),
        ("c64_feature_half", T.c64_feature_case(seed=2, scale=2), 0),
        ("layout_mix", T.layout_mix_case(4242, tc.F32), 0),
        ("everything_c32", T.random_case(77, "everything", tc.C32, max_ext=5, beta_zero=False), 0),
        ("small", T.random_case(78, "contraction", tc.F32, max_ext=4), 0),
    ]


def test_noalloc_equals_execute_product(lib, oracle):
    prods = _products()
    execs = {}
    for s in sorted({p[2] for p in prods}):
        rc, ex = lib.create_executor()
        assert rc == 0 and lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, s) == 0
        execs[s] = ex
    for name, case, split in prods:
        d = Dev(lib, case)
        ex = execs[split]
        rc1, want_bits = d.run(ex, noalloc=False)
        rc2, got = d.run(ex, noalloc=True)
        assert rc1 == 0 and rc2 == 0, (name, rc2, lib.last_error_message())
        assert np.array_equal(got.view(np.uint8), want_bits.view(np.uint8)), name
        # a second noalloc call on the same scratch gives the same bits
        rc3, again = d.run(ex, noalloc=True)
        assert rc3 == 0 and np.array_equal(again.view(np.uint8), got.view(np.uint8)), name
        if split:
            _, exact = oracle.run(case, f64=True)
            assert T.normwise_error(got, exact, T.written_mask(case.D)) <= 1e-6, name
        else:
            _, ref = oracle.run(case, gpu=True)
            tol = 1e-12 if case.D.type in (tc.F64, tc.C64) else 1e-5
            assert T.normwise_error(got, ref, T.written_mask(case.D)) <= tol, name
        d.close()
    for ex in execs.values():
        lib.destroy_executor(ex)


def test_noalloc_rejections(lib, gpu_exec, oracle):
    """Host operands (104), a product whose buffers exceed the scratch (108,
    with the size in the last-error text), a bare int executor (104), the
    self-checking kind (101); validation codes come first, as in
    TAPP_execute_product."""
    import torch
    c = T.random_case(90, "contraction", tc.F32, max_ext=5, beta_zero=False)
    infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in c.operands()]
    _, plan = lib.create_plan([o.op for o in c.operands()], infos, [o.idx for o in c.operands()], c.prec)
    al, be = T.scalar_bytes(c.alpha, c.D.type), T.scalar_bytes(c.beta, c.D.type)
    rc = lib.execute_noalloc(plan, gpu_exec, al.ctypes.data, c.A.ptr(), c.B.ptr(), be.ctypes.data, c.C.ptr(),
                             c.D.ptr())
    assert rc == tc.AMD_ERR_INVALID and "host memory" in lib.last_error_message()
    kind = ctypes.c_int(1)
    d = Dev(lib, c)
    assert lib.execute_noalloc(d.plan, ctypes.addressof(kind), *d.args()) == tc.AMD_ERR_INVALID
    assert lib.execute_noalloc(d.plan, 0, *d.args()) == 33
    rc, ex12 = lib.create_executor()
    assert lib.attr_set_int(ex12, tc.ATTR_SELF_CHECK, 1) == 0
    assert lib.execute_noalloc(d.plan, ex12, *d.args()) == tc.AMD_ERR_UNSUPPORTED
    lib.destroy_executor(ex12)
    d.close()
    bad = T.random_case(91, "contraction", tc.F32)
    bad.D.idx[0] = 999
    db = Dev(lib, bad)
    assert lib.execute_noalloc(db.plan, gpu_exec, *db.args()) == 4
    db.close()
    # unary labels on both operands: dense pre-pass buffers of 32 MiB each
    M = N = 2048
    K = 4096
    A = T.Operand(tc.F32, [M, K, 2], [1, M, M * K], T.labels("mku"), 0, None, 0)
    B = T.Operand(tc.F32, [K, N, 2], [1, K, K * N], T.labels("knv"), 0, None, 0)
    D = T.Operand(tc.F32, [M, N], [1, M], T.labels("mn"), 0, None, 0)
    big = [A, B, D, D]
    infos2 = [lib.create_info(o.type, o.extents, o.strides)[1] for o in big]
    _, plan2 = lib.create_plan([0] * 4, infos2, [o.idx for o in big], tc.DEFAULT_PREC)
    ta = torch.zeros(M * K * 2, device="cuda")
    tb = torch.zeros(K * N * 2, device="cuda")
    td = torch.zeros(M * N, device="cuda")
    torch.cuda.synchronize()
    one, zero = np.ones(1, np.float32), np.zeros(1, np.float32)
    rc = lib.execute_noalloc(plan2, gpu_exec, one.ctypes.data, ta.data_ptr(), tb.data_ptr(), zero.ctypes.data, None,
                             td.data_ptr())
    assert rc == tc.AMD_ERR_SCRATCH, rc
    assert "needs" in lib.last_error_message()
    lib.destroy_plan(plan2)
    for h in infos2:
        lib.destroy_info(h)
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)

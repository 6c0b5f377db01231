"""Persistent launch over uneven products (product groups) and the one-plan
TAPP_execute_batched_product, on device operands."""
import numpy as np
import pytest

import tapp_cases as T
from test_gpu_parity import assert_parity

tc = T.tc
pytestmark = pytest.mark.gpu


class DevCase:
    """A case with operands copied to device memory and a plan."""

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

    def c_ptr(self):
        return None if self.case.c_null else self.ptr["C"]

    def run_single(self, ex):
        return self.lib.execute(self.plan, ex, self.al.ctypes.data, self.ptr["A"], self.ptr["B"], self.be.ctypes.data,
                                self.c_ptr(), self.ptr["D"])

    def d(self):
        return self.t["D"].cpu().numpy().view(self.case.D.buf.dtype)

    def reset_d(self):
        import torch
        self.t["D"].copy_(torch.from_numpy(self.case.D.buf.view(np.uint8).copy()).cuda())
        torch.cuda.synchronize()   # torch's stream; the library launches on its own

    def close(self):
        self.lib.destroy_plan(self.plan)
        for h in self.infos:
            self.lib.destroy_info(h)


def make_group(lib, ex, dcs):
    return lib.create_group(ex, [d.plan for d in dcs], [d.al.ctypes.data for d in dcs], [d.ptr["A"] for d in dcs],
                            [d.ptr["B"] for d in dcs], [d.be.ctypes.data for d in dcs], [d.c_ptr() for d in dcs],
                            [d.ptr["D"] for d in dcs])


def mixed_cases(seed):
    cases = T.batching_mix(seed, 24, tucker_range=(4, 24), r_range=(1, 16), c0_range=(2, 9))
    # products outside the persistent kernel (run one by one inside the group)
    cases.append(T.random_case(seed + 1, "unary", tc.F32, max_ext=4))
    cases.append(T.random_case(seed + 2, "everything", tc.C64, max_ext=3))
    # edge shapes: K = 0, S = 1, long K over several LDS chunks, Hadamard batch
    cases.append(T.tucker_case(seed + 3, 200, 7, 5, 3))
    cases.append(T.tucker_case(seed + 4, 5, 9, 300, 1))
    cases.append(T.random_case(seed + 5, "hadamard_contraction", tc.F32, max_ext=6, min_ext=3, layouts=True))
    cases.append(T.random_case(seed + 6, "contraction", tc.F32, max_ext=12, min_ext=8, layouts=True))
    c = T.four_index_case(seed + 7, (3, 4, 5, 6, 0, 2))
    cases.append(c)
    return cases


def test_group_matches_single_calls_and_oracle(lib, gpu_exec, oracle):
    cases = mixed_cases(11)
    dcs = [DevCase(lib, c) for c in cases]
    rc, g = make_group(lib, gpu_exec, dcs)
    assert rc == 0, lib.explain(rc)
    rc, info = lib.group_info(g)
    assert rc == 0 and info[0] > 0 and info[1] >= 2, info
    assert lib.execute_group(g) == 0
    grouped = [d.d().copy() for d in dcs]
    for d, got in zip(dcs, grouped):
        d.reset_d()
        assert d.run_single(gpu_exec) == 0
        single = d.d()
        _, want = oracle.run(d.case, gpu=True)
        assert_parity(got, want, d.case)
        assert_parity(single, want, d.case)
    lib.destroy_group(g)
    for d in dcs:
        d.close()


def test_group_graph_replays(lib, gpu_exec, oracle):
    """Replays of the captured graph, with eager work and new input values
    between them; every replay checked against the oracle."""
    import torch
    cases = T.batching_mix(5, 12, tucker_range=(4, 20), r_range=(2, 8), c0_range=(3, 8))
    dcs = [DevCase(lib, c) for c in cases]
    rc, g = make_group(lib, gpu_exec, dcs)
    assert rc == 0
    for replay in range(3):
        for d in dcs:   # eager work: new values for A in place
            new = T.fill(d.case.A.type, len(d.case.A.buf), 1000 * replay + len(d.case.A.buf))
            d.case.A.buf = new
            d.t["A"].copy_(torch.from_numpy(new.view(np.uint8).copy()).cuda())
            torch.cuda.synchronize()
        x = torch.ones(1 << 20, device="cuda")
        x.mul_(2.0).sum().item()
        torch.cuda.synchronize()
        assert lib.execute_group(g) == 0
        for d in dcs:
            _, want = oracle.run(d.case, gpu=True)
            assert_parity(d.d(), want, d.case)
    lib.destroy_group(g)
    for d in dcs:
        d.close()


def test_group_bf16_products(lib, gpu_exec, oracle):
    cases = T.batching_mix(7, 10, tucker_range=(4, 16), r_range=(2, 8), c0_range=(3, 7),
                           prec=tc.BF16BF16_ACCUM_F32)
    dcs = [DevCase(lib, c) for c in cases]
    rc, g = make_group(lib, gpu_exec, dcs)
    assert rc == 0
    assert lib.execute_group(g) == 0
    for d in dcs:
        _, want = oracle.run(d.case, gpu=True)
        assert_parity(d.d(), want, d.case, tol=1e-5)
    lib.destroy_group(g)
    for d in dcs:
        d.close()


def test_group_reports_validation_codes(lib, gpu_exec):
    from test_oracle import negative_cases
    good = DevCase(lib, T.tucker_case(1, 4, 5, 6, 3))
    for name, case in negative_cases():
        if name == "int32_truncation":
            continue
        bad = DevCase(lib, case)
        rc, g = make_group(lib, gpu_exec, [good, bad])
        want = lib.execute(bad.plan, gpu_exec, None, None, None, bad.be.ctypes.data, bad.c_ptr(), None)
        assert rc == want != 0, (name, rc, want)
        bad.close()
    good.close()


def test_batched_product_device_pointers(lib, gpu_exec, oracle):
    base = T.four_index_case(3, (5, 6, 7, 4, 3, 5))
    sets = []
    for s in range(6):
        c = base.copy()
        for o in c.operands():
            o.buf = T.fill(o.type, len(o.buf), 77 + 13 * s + o.type)
        sets.append(DevCase(lib, c))
    b = sets[0]
    rc = lib.execute_batched(b.plan, gpu_exec, b.al.ctypes.data, [d.ptr["A"] for d in sets],
                             [d.ptr["B"] for d in sets], b.be.ctypes.data, [d.ptr["C"] for d in sets],
                             [d.ptr["D"] for d in sets])
    assert rc == 0
    assert lib.attr_get_int(gpu_exec, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_GROUP
    for d in sets:
        _, want = oracle.run(d.case, gpu=True)
        assert_parity(d.d(), want, d.case)
    for d in sets:
        d.close()


@pytest.mark.parametrize("split", [1, 2, 3])
def test_group_with_split_forms(lib, oracle, split):
    """An executor with a split form: the group runs those FP32 products one
    by one inside the launch (the persistent kernel has no split form), so
    each D equals a single call on the same executor bit for bit, and stays
    within 1e-5 of the FP64 oracle."""
    _, ex = lib.create_executor()
    assert lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, split) == 0
    cases = T.batching_mix(31, 8, tucker_range=(4, 16), r_range=(2, 8), c0_range=(3, 7))
    cases.append(T.random_case(32, "contraction", tc.F32, max_ext=12, min_ext=8, layouts=True))
    dcs = [DevCase(lib, c) for c in cases]
    rc, g = make_group(lib, ex, dcs)
    assert rc == 0
    assert lib.execute_group(g) == 0
    grouped = [d.d().copy() for d in dcs]
    for d, got in zip(dcs, grouped):
        d.reset_d()
        assert d.run_single(ex) == 0
        assert np.array_equal(got.view(np.uint32), d.d().view(np.uint32)), d.case.name
        _, exact = oracle.run(d.case, f64=True)
        assert T.normwise_error(got, exact, T.written_mask(d.case.D)) < 1e-5, d.case.name
    lib.destroy_group(g)
    for d in dcs:
        d.close()
    lib.destroy_executor(ex)


class BatchSets:
    """`n` operand sets of one layout on the device, executed through one
    plan with TAPP_execute_batched_product."""

    def __init__(self, lib, base, n, seed):
        self.lib = lib
        self.sets = []
        for s in range(n):
            c = base.copy()
            for o in c.operands():
                o.buf = T.fill(o.type, len(o.buf), seed + 13 * s + o.type)
            self.sets.append(DevCase(lib, c))
        self.plan = self.sets[0].plan

    def call(self, ex, alpha, beta, pick=None, c_null=False, n=None, a_ptr=None):
        """D of set i from A, B of set pick[i] and C of set i; returns
        (rc, reused, expected cases)."""
        n = len(self.sets) if n is None else n
        pick = list(range(n)) if pick is None else pick
        d_type = self.sets[0].case.D.type
        al, be = T.scalar_bytes(alpha, d_type), T.scalar_bytes(beta, d_type)
        for d in self.sets[:n]:
            d.reset_d()
        a = [self.sets[pick[i]].ptr["A"] for i in range(n)]
        if a_ptr:
            for i, p in a_ptr.items():
                a[i] = p
        rc = self.lib.execute_batched(self.plan, ex, al.ctypes.data, a, [self.sets[pick[i]].ptr["B"] for i in range(n)],
                                      be.ctypes.data, None if c_null else [d.ptr["C"] for d in self.sets[:n]],
                                      [d.ptr["D"] for d in self.sets[:n]])
        want = []
        for i in range(n):
            c = self.sets[i].case.copy()
            c.A, c.B = self.sets[pick[i]].case.A.copy(), self.sets[pick[i]].case.B.copy()
            c.alpha, c.beta, c.c_null = alpha, beta, c_null
            want.append(c)
        return rc, self.lib.attr_get_int(ex, tc.ATTR_LAST_REUSED)[1], want

    def check(self, oracle, want):
        for d, c in zip(self.sets, want):
            _, w = oracle.run(c, gpu=True)
            assert_parity(d.d(), w, c)

    def close(self):
        for d in self.sets:
            d.close()


@pytest.mark.parametrize("shape", ["four_index", "tucker"])
def test_batched_product_reuses_prepared_batch(lib, oracle, shape):
    """Repeated TAPP_execute_batched_product calls on one plan: the second
    call with the same batch count replays the prepared batch (descriptors
    patched for changed pointers, scalars refreshed) and every D still
    equals the oracle; beta switching to 0, a host operand and another batch
    count build the batch again."""
    import torch
    _, ex = lib.create_executor()
    base = T.four_index_case(5, (6, 5, 7, 4, 3, 5)) if shape == "four_index" else T.tucker_case(6, 24, 9, 7, 5)
    bs = BatchSets(lib, base, 6, 300)
    rc, reused, want = bs.call(ex, 1.5, 0.5)
    assert (rc, reused) == (0, 0)
    assert lib.attr_get_int(ex, tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_GROUP
    bs.check(oracle, want)
    # same pointers, new data in A of set 2: read, not cached
    s2 = bs.sets[2]
    s2.case.A.buf[:] = T.fill(s2.case.A.type, len(s2.case.A.buf), 999)
    s2.t["A"].copy_(torch.from_numpy(s2.case.A.buf.view(np.uint8).copy()).cuda())
    torch.cuda.synchronize()
    rc, reused, want = bs.call(ex, 1.5, 0.5)
    assert (rc, reused) == (0, 1)
    bs.check(oracle, want)
    # operand sets permuted, new scalars
    rc, reused, want = bs.call(ex, -0.75, 2.0, pick=[5, 3, 4, 0, 1, 2])
    assert (rc, reused) == (0, 1)
    bs.check(oracle, want)
    # beta = 0 without C: built again, then reused
    for expect in (0, 1):
        rc, reused, want = bs.call(ex, 2.0, 0.0, c_null=True)
        assert (rc, reused) == (0, expect)
        bs.check(oracle, want)
    # a host operand: staged through the general path, not reused
    host_a = bs.sets[1].case.A.ptr()
    rc, reused, want = bs.call(ex, 2.0, 0.0, c_null=True, a_ptr={1: host_a})
    assert (rc, reused) == (0, 0)
    bs.check(oracle, want)
    # another batch count
    rc, reused, want = bs.call(ex, 1.0, 1.0, n=4)
    assert (rc, reused) == (0, 0)
    bs.check(oracle, want[:4])
    rc, reused, want = bs.call(ex, 1.0, 1.0, n=4, pick=[1, 0, 3, 2])
    assert (rc, reused) == (0, 1)
    bs.check(oracle, want[:4])
    # a NULL operand is reported as on the first call
    d_type = bs.sets[0].case.D.type
    al = T.scalar_bytes(1.0, d_type)
    ptrs = [d.ptr["A"] for d in bs.sets[:4]]
    ptrs[2] = None
    rc = lib.execute_batched(bs.plan, ex, al.ctypes.data, ptrs, [d.ptr["B"] for d in bs.sets[:4]], al.ctypes.data,
                             [d.ptr["C"] for d in bs.sets[:4]], [d.ptr["D"] for d in bs.sets[:4]])
    rc_single = lib.execute(bs.plan, ex, al.ctypes.data, None, bs.sets[0].ptr["B"], al.ctypes.data,
                            bs.sets[0].ptr["C"], bs.sets[0].ptr["D"])
    assert rc == rc_single != 0
    bs.close()   # frees the prepared batch with the plan
    lib.destroy_executor(ex)


def test_batched_product_reuse_follows_layout_change(lib, oracle):
    """A layout changed through TAPP_set_strides after a batched call: the
    next call rebuilds (the plan's offset tables changed) and reads the new
    layout."""
    import ctypes
    from test_gpu_parity import _gemm_case
    _, ex = lib.create_executor()
    base = _gemm_case(tc.F32, 40, 30, 20, 9)
    bs = BatchSets(lib, base, 3, 500)
    rc, reused, want = bs.call(ex, 1.0, 0.5)
    assert (rc, reused) == (0, 0)
    bs.check(oracle, want)
    rc, reused, _ = bs.call(ex, 1.0, 0.5)
    assert (rc, reused) == (0, 1)
    # A stored transposed in every set, announced through the plan's info
    import torch
    M, K = base.A.extents
    for d in bs.sets:
        a2 = d.case.A.buf[:M * K].reshape(K, M).T.copy().reshape(-1).astype(np.float32)
        d.case.A.buf = a2
        d.case.A.strides = [K, 1]
        d.case.A.offset = 0
        d.t["A"] = torch.from_numpy(a2.view(np.uint8).copy()).cuda()
        d.ptr["A"] = d.t["A"].data_ptr()
    torch.cuda.synchronize()
    s = (ctypes.c_int64 * 2)(K, 1)
    assert lib.lib.TAPP_set_strides(bs.sets[0].infos[0], s) == 0
    rc, reused, want = bs.call(ex, 1.0, 0.5)
    assert (rc, reused) == (0, 0)
    bs.check(oracle, want)
    bs.close()
    lib.destroy_executor(ex)


def test_group_full_distribution_sample(lib, gpu_exec, oracle):
    """1024 products drawn from the uneven batch's full distribution (80 %
    Tucker i, j, k ~ U{4..64}, r ~ U{2..16}; 20 % four-index products with
    each extent ~ U{6..24}) in one product group, both persistent launches:
    every Tucker product and 16 four-index products (drawn with a seed
    among those of at most 3e7 MACs) equal the oracle's."""
    cases = T.batching_mix(2, 1024)
    dcs = [DevCase(lib, c) for c in cases]
    rc, g = make_group(lib, gpu_exec, dcs)
    assert rc == 0
    assert lib.group_info(g)[1][0] == len(dcs)
    assert lib.execute_group(g) == 0
    tucker = [i for i, c in enumerate(cases) if c.name.startswith("tucker")]
    four = [i for i, c in enumerate(cases) if c.name.startswith("four_index")
            and np.prod([float(e) for e in c.A.extents]) * c.D.extents[2] * c.D.extents[3] <= 3e7]
    assert len(tucker) > 700 and len(four) >= 16
    pick = sorted(np.random.default_rng(5).choice(four, 16, replace=False).tolist())
    for i in tucker + pick:
        _, want = oracle.run(cases[i], gpu=True)
        assert_parity(dcs[i].d(), want, cases[i])
    assert lib.destroy_group(g) == 0
    for d in dcs:
        d.close()


def test_group_thin_tiles_affine_and_table_layouts(lib, gpu_exec, oracle):
    """Thin tiles over a dense Y take the strided path when every other
    table of the product is index * stride (a mode's stride is the previous
    one times its extent), and the table path otherwise.  Mode products
    D[r,j,k] = alpha sum_i A[i,j,k] U[r,i] + beta C with: dense operands;
    U stored transposed; D and C with a padded but affine column stride;
    D and C stored backwards along r; D with j and k swapped in storage
    (table path); A padded (not dense: table path); a Hadamard label k in
    U and D.  Each D equals the oracle's."""
    i, j, k, r = 37, 11, 29, 9
    mk = T.make_operand
    f32 = tc.F32

    def case(n, A, U, C, D, beta=0.5):
        return T.Case(A, U, C, D, tc.DEFAULT_PREC, 1.25, beta, name=f"thin_layout{n}")

    def dense_A(seed, pad=0):
        return mk(f32, [i, j, k], "ijk", seed, pad=pad)

    cases = [
        case(0, dense_A(1), mk(f32, [r, i], "ri", 2), mk(f32, [r, j, k], "rjk", 3), mk(f32, [r, j, k], "rjk", 4)),
        case(1, dense_A(5), mk(f32, [r, i], "ri", 6, order=[1, 0]), mk(f32, [r, j, k], "rjk", 7),
             mk(f32, [r, j, k], "rjk", 8), beta=0.0),
        case(2, dense_A(9), mk(f32, [r, i], "ri", 10), mk(f32, [r, j, k], "rjk", 11, strides=[1, r + 3, (r + 3) * j]),
             mk(f32, [r, j, k], "rjk", 12, strides=[1, r + 5, (r + 5) * j])),
        case(3, dense_A(13), mk(f32, [r, i], "ri", 14), mk(f32, [r, j, k], "rjk", 15, negative=(0,)),
             mk(f32, [r, j, k], "rjk", 16, negative=(0,))),
        case(4, dense_A(17), mk(f32, [r, i], "ri", 18), mk(f32, [r, j, k], "rjk", 19),
             mk(f32, [r, j, k], "rjk", 20, order=[0, 2, 1])),
        case(5, dense_A(21, pad=2), mk(f32, [r, i], "ri", 22), mk(f32, [r, j, k], "rjk", 23),
             mk(f32, [r, j, k], "rjk", 24)),
        case(6, dense_A(25), mk(f32, [r, i, k], "rik", 26), mk(f32, [r, j, k], "rjk", 27),
             mk(f32, [r, j, k], "rjk", 28)),
    ]
    dcs = [DevCase(lib, c) for c in cases]
    rc, g = make_group(lib, gpu_exec, dcs)
    assert rc == 0
    assert lib.group_info(g)[1][0] == len(dcs)          # every product in the persistent launch
    assert lib.execute_group(g) == 0
    for d in dcs:
        _, want = oracle.run(d.case, gpu=True)
        assert_parity(d.d(), want, d.case)
    assert lib.destroy_group(g) == 0
    for d in dcs:
        d.close()


def test_group_dense_thin_tiles_across_k_passes(lib, gpu_exec, oracle):
    """Tucker products on thin tiles across K chunk and pass edges: K = 1,
    31, 32, 33 (one row past two 16-deep chunks), 64, 65, 97, with column
    counts across the 256-column tile edge and 1 to 16 rows.  Each D equals
    the oracle's."""
    shapes = [(1, 5, 7, 3), (31, 17, 16, 4), (32, 16, 16, 16), (33, 9, 30, 5), (64, 20, 13, 8), (65, 16, 17, 1),
              (97, 7, 40, 12), (33, 1, 1, 2), (40, 300, 1, 16)]
    cases = [T.tucker_case(500 + n, i, j, k, r) for n, (i, j, k, r) in enumerate(shapes)]
    dcs = [DevCase(lib, c) for c in cases]
    rc, g = make_group(lib, gpu_exec, dcs)
    assert rc == 0
    assert lib.group_info(g)[1][0] == len(dcs)
    assert lib.execute_group(g) == 0
    for d in dcs:
        _, want = oracle.run(d.case, gpu=True)
        assert_parity(d.d(), want, d.case)
    assert lib.destroy_group(g) == 0
    for d in dcs:
        d.close()

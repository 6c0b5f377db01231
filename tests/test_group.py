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

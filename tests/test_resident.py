"""Resident tensors: handle-scoped device copies of host buffers, read by
TAPP_execute_product, product groups and the serving loop in place of
staging the host operand."""

import numpy as np
import pytest

import tapp_cases as T

tc = T.tc


class HostPlan:
    """A case with host operands and a plan created with `handle`."""

    def __init__(self, lib, case, handle, copy=True):
        self.lib, self.case = lib, (case.copy() if copy else case)
        self.infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in self.case.operands()]
        rc, self.plan = lib.create_plan([o.op for o in self.case.operands()], self.infos,
                                        [o.idx for o in self.case.operands()], self.case.prec, handle=handle)
        assert rc == 0
        self.al = T.scalar_bytes(self.case.alpha, self.case.D.type)
        self.be = T.scalar_bytes(self.case.beta, self.case.D.type)

    def args(self):
        c = self.case
        return (self.al.ctypes.data, c.A.ptr(), c.B.ptr(), self.be.ctypes.data, None if c.c_null else c.C.ptr(),
                c.D.ptr())

    def run(self, ex):
        self.case.D.buf[:] = self.d0
        return self.lib.execute(self.plan, ex, *self.args())

    def close(self):
        self.lib.destroy_plan(self.plan)
        for h in self.infos:
            self.lib.destroy_info(h)


def _nbytes(o):
    return o.buf.nbytes


def test_resident_api_errors_without_device(lib):
    """Argument checks that need no GPU: unknown handle, NULL range, empty range."""
    buf = np.zeros(16, np.float32)
    assert lib.make_resident(12345, buf.ctypes.data, buf.nbytes) == tc.AMD_ERR_INVALID
    rc, h = lib.create_handle()
    assert rc == 0
    assert lib.resident_info(h) == (0, [0, 0])
    assert lib.make_resident(h, None, 64) == tc.AMD_ERR_INVALID
    assert lib.make_resident(h, buf.ctypes.data, 0) == tc.AMD_ERR_INVALID
    assert lib.update_resident(h, buf.ctypes.data) == tc.AMD_ERR_INVALID
    assert lib.release_resident(h, buf.ctypes.data) == tc.AMD_ERR_INVALID
    assert lib.resident_info(777)[0] == tc.AMD_ERR_INVALID
    assert lib.destroy_handle(h) == 0


@pytest.mark.gpu
def test_resident_operand_is_read_from_its_copy(lib, gpu_exec, oracle):
    """A resident A is read from the device copy: host writes after
    make_resident are not seen until update_resident; after release the
    host buffer is staged again."""
    _, h = lib.create_handle()
    case = T.random_case(11, "contraction", tc.F32, max_ext=6, min_ext=3, beta_zero=False)
    hp = HostPlan(lib, case, h)
    hp.d0 = hp.case.D.buf.copy()
    A = hp.case.A
    assert lib.make_resident(h, A.buf.ctypes.data, _nbytes(A)) == 0
    assert lib.resident_info(h) == (0, [1, _nbytes(A)])
    # overlapping range on the same handle
    assert lib.make_resident(h, A.buf.ctypes.data + 4, 8) == tc.AMD_ERR_INVALID

    _, want_old = oracle.run(hp.case, gpu=True)
    assert hp.run(gpu_exec) == 0
    got = hp.case.D.buf.copy()
    assert T.normwise_error(got, want_old, T.written_mask(hp.case.D)) <= 1e-5

    old_a = A.buf.copy()
    A.buf[:] = T.fill(A.type, A.buf.size, 99)            # host change, not uploaded
    assert hp.run(gpu_exec) == 0
    assert np.array_equal(hp.case.D.buf, got), "resident copy must be read, not the host buffer"

    assert lib.update_resident(h, A.buf.ctypes.data) == 0  # now the copy follows
    _, want_new = oracle.run(hp.case, gpu=True)
    assert hp.run(gpu_exec) == 0
    new = hp.case.D.buf.copy()
    assert T.normwise_error(new, want_new, T.written_mask(hp.case.D)) <= 1e-5
    assert not np.allclose(new, got)

    assert lib.release_resident(h, A.buf.ctypes.data) == 0
    assert lib.resident_info(h) == (0, [0, 0])
    A.buf[:] = old_a                                      # staged again: host data is read
    assert hp.run(gpu_exec) == 0
    assert T.normwise_error(hp.case.D.buf, want_old, T.written_mask(hp.case.D)) <= 1e-5
    hp.close()
    assert lib.destroy_handle(h) == 0


@pytest.mark.gpu
def test_resident_sub_range_and_plan_without_handle(lib, gpu_exec, oracle):
    """An operand inside a larger resident buffer (offset base pointer) is
    translated; a plan created without the handle stages the host data."""
    _, h = lib.create_handle()
    case = T.random_case(12, "everything", tc.C64, max_ext=4, beta_zero=False)
    big = np.zeros(case.B.buf.size + 64, dtype=case.B.buf.dtype)
    big[32:32 + case.B.buf.size] = case.B.buf
    case.B.buf = big[32:32 + case.B.buf.size]           # a view inside the resident range
    assert lib.make_resident(h, big.ctypes.data, big.nbytes) == 0
    for handle in (h, 0):
        hp = HostPlan(lib, case, handle, copy=False)
        hp.d0 = hp.case.D.buf.copy()
        rc_o, want = oracle.run(hp.case, gpu=True)
        assert hp.run(gpu_exec) == rc_o == 0
        assert T.normwise_error(hp.case.D.buf, want, T.written_mask(hp.case.D)) <= 1e-12
        hp.close()
    assert lib.destroy_handle(h) == 0                     # frees the copy


def _resident_tucker_plans(lib, torch, h):
    cases = [T.tucker_case(20 + i, 16, 12, 10, 4) for i in range(4)]
    hps, dts = [], []
    for c in cases:
        hp = HostPlan(lib, c, h)
        for o in (hp.case.A, hp.case.B):
            assert lib.make_resident(h, o.buf.ctypes.data, _nbytes(o)) == 0
        dt = torch.from_numpy(hp.case.D.buf.view(np.uint8).copy()).cuda()
        hps.append(hp)
        dts.append(dt)
    dptr = [dt.data_ptr() + hp.case.D.offset * hp.case.D.buf.itemsize for hp, dt in zip(hps, dts)]
    return hps, dts, dptr


@pytest.mark.gpu
def test_resident_operands_in_group(lib, gpu_exec, oracle):
    """Product groups run resident host operands in the persistent launch."""
    import torch
    _, h = lib.create_handle()
    hps, dts, dptr = _resident_tucker_plans(lib, torch, h)
    rc, g = lib.create_group(gpu_exec, [hp.plan for hp in hps], [hp.al.ctypes.data for hp in hps],
                             [hp.case.A.ptr() for hp in hps], [hp.case.B.ptr() for hp in hps],
                             [hp.be.ctypes.data for hp in hps],
                             [None if hp.case.c_null else hp.case.C.ptr() for hp in hps], dptr)
    assert rc == 0
    assert lib.group_info(g)[1][0] == len(hps), "resident operands run in the persistent launch"
    assert lib.execute_group(g) == 0
    for hp, dt in zip(hps, dts):
        _, want = oracle.run(hp.case, gpu=True)
        got = dt.cpu().numpy().view(hp.case.D.buf.dtype)
        assert T.normwise_error(got, want, T.written_mask(hp.case.D)) <= 1e-5
    assert lib.destroy_group(g) == 0
    for hp in hps:
        hp.close()
    assert lib.destroy_handle(h) == 0


@pytest.mark.gpu
def test_resident_operands_in_group_and_server(lib, gpu_exec, oracle):
    """Product groups run resident host operands in the persistent launch;
    the serving loop does not charge them to the HBM budget."""
    import torch
    _, h = lib.create_handle()
    cases = [T.tucker_case(20 + i, 16, 12, 10, 4) for i in range(4)]
    hps, dts = [], []
    for c in cases:
        hp = HostPlan(lib, c, h)
        for o in (hp.case.A, hp.case.B):
            assert lib.make_resident(h, o.buf.ctypes.data, _nbytes(o)) == 0
        dt = torch.from_numpy(hp.case.D.buf.view(np.uint8).copy()).cuda()
        hps.append(hp)
        dts.append(dt)
    dptr = [dt.data_ptr() + hp.case.D.offset * hp.case.D.buf.itemsize for hp, dt in zip(hps, dts)]
    rc, g = lib.create_group(gpu_exec, [hp.plan for hp in hps], [hp.al.ctypes.data for hp in hps],
                             [hp.case.A.ptr() for hp in hps], [hp.case.B.ptr() for hp in hps],
                             [hp.be.ctypes.data for hp in hps],
                             [None if hp.case.c_null else hp.case.C.ptr() for hp in hps], dptr)
    assert rc == 0
    assert lib.group_info(g)[1][0] == len(hps), "resident operands run in the persistent launch"
    assert lib.execute_group(g) == 0
    for hp, dt in zip(hps, dts):
        _, want = oracle.run(hp.case, gpu=True)
        got = dt.cpu().numpy().view(hp.case.D.buf.dtype)
        assert T.normwise_error(got, want, T.written_mask(hp.case.D)) <= 1e-5
    assert lib.destroy_group(g) == 0

    # server: a budget below one A's bytes still admits requests whose A/B are resident
    smallest = min(_nbytes(hp.case.A) for hp in hps)
    rc, srv = lib.create_server(gpu_exec, smallest // 2, slots=2)
    assert rc == 0
    for dt in dts:
        dt.zero_()
    torch.cuda.synchronize()   # the fills run on torch's stream, the server on its own
    sts = []
    for hp, dt, d in zip(hps, dts, dptr):
        a = hp.args()
        rc, st = lib.submit(srv, hp.plan, a[0], a[1], a[2], a[3], a[4], d)
        assert rc == 0, lib.explain(rc)
        sts.append(st)
    for st in sts:
        assert lib.wait(srv, st) == (0, 0)
        lib.destroy_status(st)
    for hp, dt in zip(hps, dts):
        _, want = oracle.run(hp.case, gpu=True)
        got = dt.cpu().numpy().view(hp.case.D.buf.dtype)
        assert T.normwise_error(got, want, T.written_mask(hp.case.D)) <= 1e-5
    assert lib.destroy_server(srv) == 0
    for hp in hps:
        hp.close()
    assert lib.destroy_handle(h) == 0

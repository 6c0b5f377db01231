"""Serving loop: admission by HBM budget, cancellation, continuous batching,
and every result the same alone and in the mix."""
import numpy as np
import pytest

import tapp_cases as T

tc = T.tc


def _plan(lib, case):
    infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in case.operands()]
    _, plan = lib.create_plan([o.op for o in case.operands()], infos, [o.idx for o in case.operands()], case.prec)
    return plan, infos


def _free(lib, plan, infos):
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)


# --- queue and admission logic (no kernel runs: the server is paused) ----------
def test_budget_rejection_and_cancel_on_paused_server(lib):
    rc, ex = lib.create_executor()
    assert rc == 0
    case = T.tucker_case(1, 8, 9, 10, 4)            # host operands: footprint = their spans
    footprint = sum(len(o.buf) * o.buf.itemsize for o in (case.A, case.B, case.D))
    plan, infos = _plan(lib, case)
    al = T.scalar_bytes(case.alpha, case.D.type)
    be = T.scalar_bytes(case.beta, case.D.type)
    # budget below one request: rejected at submit, nothing queued
    rc, small = lib.create_server(ex, footprint - 1, 2, 8, lib.SERVER_PAUSED)
    assert rc == 0
    rc, st = lib.submit(small, plan, al.ctypes.data, case.A.ptr(), case.B.ptr(), be.ctypes.data, None, case.D.ptr())
    assert rc == 106
    assert lib.server_stats(small)[1]["queued"] == 0
    lib.destroy_server(small)
    # enough budget: queued while paused; cancel one; validation still applies
    rc, srv = lib.create_server(ex, 10 * footprint, 2, 8, lib.SERVER_PAUSED)
    sts = []
    for _ in range(3):
        rc, st = lib.submit(srv, plan, al.ctypes.data, case.A.ptr(), case.B.ptr(), be.ctypes.data, None,
                            case.D.ptr())
        assert rc == 0
        sts.append(st)
    assert lib.server_stats(srv)[1]["queued"] == 3
    assert lib.cancel(srv, sts[1]) == 0
    assert lib.wait(srv, sts[1]) == (0, 104)          # a cancelled request reports an invalid-argument code
    lib.destroy_status(sts[1])
    bad = T.random_case(4, "contraction", tc.F32)
    bad.D.idx[0] = 999
    bplan, binfos = _plan(lib, bad)
    rc, _ = lib.submit(srv, bplan, al.ctypes.data, bad.A.ptr(), bad.B.ptr(), be.ctypes.data, None, bad.D.ptr())
    assert rc == 4
    lib.destroy_server(srv)                             # drops the never-run requests
    for st in (sts[0], sts[2]):
        lib.destroy_status(st)
    _free(lib, plan, infos)
    _free(lib, bplan, binfos)
    lib.destroy_executor(ex)


# --- on the GPU ----------------------------------------------------------------
@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix(lib, gpu_exec, oracle):
    from test_group import DevCase, mixed_cases
    from test_gpu_parity import assert_parity
    # plus the ladder form D[a,b,i,j] = sum A[a,b,c,d] B[c,i,d,j] of the
    # serving workload's large class, at a..d = 24, i, j = 8
    cases = mixed_cases(31) + [T.four_index16_case(2), T.ladder_case(seed=3, v=24, o=8)]
    dcs = [DevCase(lib, c) for c in cases]
    alone = []
    for d in dcs:
        assert d.run_single(gpu_exec) == 0
        alone.append(d.d().copy())
        d.reset_d()
    rc, srv = lib.create_server(gpu_exec, 1 << 30, 4, 16)
    assert rc == 0
    # all submitted back to back: they land in several batches on 4 streams
    failed, rcs, late = tc.trace_replay(lib, srv, [d.plan for d in dcs], [d.al.ctypes.data for d in dcs],
                                        [d.ptr["A"] for d in dcs], [d.ptr["B"] for d in dcs],
                                        [d.be.ctypes.data for d in dcs], [d.c_ptr() for d in dcs],
                                        [d.ptr["D"] for d in dcs], [0.0] * len(dcs))
    assert failed == 0 and all(r == 0 for r in rcs), rcs
    rc, stats = lib.server_stats(srv)
    assert stats["completed"] == len(dcs) and stats["persistent_products"] > 0, stats
    for d, a in zip(dcs, alone):
        _, want = oracle.run(d.case, gpu=True)
        assert_parity(d.d(), want, d.case)
        assert_parity(d.d(), a, d.case)
    lib.destroy_server(srv)
    for d in dcs:
        d.close()


@pytest.mark.gpu
def test_budget_stalls_then_completes(lib, gpu_exec, oracle):
    """Host-operand requests whose staging footprints do not all fit the
    budget at once: they wait (stalls) and all complete correctly."""
    from test_gpu_parity import assert_parity
    cases = [T.tucker_case(100 + i, 24, 20, 16, 6) for i in range(8)]
    fp = max(sum(len(o.buf) * o.buf.itemsize for o in (c.A, c.B, c.D)) for c in cases)
    rc, srv = lib.create_server(gpu_exec, int(2.5 * fp), 2, 4, lib.SERVER_PAUSED)
    assert rc == 0
    keep, sts = [], []
    for c in cases:
        c2 = c.copy()
        plan, infos = _plan(lib, c2)
        al = T.scalar_bytes(c2.alpha, c2.D.type)
        be = T.scalar_bytes(c2.beta, c2.D.type)
        rc, st = lib.submit(srv, plan, al.ctypes.data, c2.A.ptr(), c2.B.ptr(), be.ctypes.data, None, c2.D.ptr())
        assert rc == 0
        keep.append((c2, plan, infos, al, be))
        sts.append(st)
    assert lib.server_stats(srv)[1]["queued"] == 8
    assert lib.set_paused(srv, 0) == 0          # all 8 queued: at most 2 fit the budget at a time
    for st in sts:
        assert lib.wait(srv, st) == (0, 0)
        lib.destroy_status(st)
    rc, stats = lib.server_stats(srv)
    assert stats["completed"] == 8 and stats["stalls"] > 0, stats
    for c2, plan, infos, al, be in keep:
        _, want = oracle.run(c2, gpu=True)
        assert_parity(c2.D.buf, want, c2)
        _free(lib, plan, infos)
    lib.destroy_server(srv)


@pytest.mark.gpu
def test_split_f16_requests_run_the_three_piece_split(lib, gpu_exec, oracle):
    """The serving loop runs requests asynchronously, where split-F16 cannot
    check its range and run again: an MFMA product takes the three-piece
    split-BF16, a product for the one-lane kernel plain FP32.  Each D equals
    a single call in that form, bit for bit."""
    from test_group import DevCase
    cases = [T.random_case(60 + i, "contraction", tc.F32, max_ext=12, min_ext=8, layouts=True) for i in range(3)]
    cases += [T.random_case(70 + i, "contraction", tc.F32, max_ext=3) for i in range(3)]
    dcs = [DevCase(lib, c) for c in cases]
    exs = {}
    for v in (0, 2, 3):
        _, exs[v] = lib.create_executor()
        assert lib.attr_set_int(exs[v], tc.ATTR_SPLIT_BF16, v) == 0
    want = []
    for d in dcs:
        assert d.run_single(exs[3]) == 0
        mfma = lib.attr_get_int(exs[3], tc.ATTR_LAST_KERNEL)[1] == tc.KERNEL_MFMA
        d.reset_d()
        assert d.run_single(exs[2 if mfma else 0]) == 0
        want.append(d.d().copy())
        d.reset_d()
    assert any(lib.attr_get_int(exs[3], tc.ATTR_LAST_KERNEL)[1] == k for k in (tc.KERNEL_SIMPLE, tc.KERNEL_MFMA))
    rc, srv = lib.create_server(exs[3], 1 << 30, 4, 16)
    assert rc == 0
    failed, rcs, late = tc.trace_replay(lib, srv, [d.plan for d in dcs], [d.al.ctypes.data for d in dcs],
                                        [d.ptr["A"] for d in dcs], [d.ptr["B"] for d in dcs],
                                        [d.be.ctypes.data for d in dcs], [d.c_ptr() for d in dcs],
                                        [d.ptr["D"] for d in dcs], [0.0] * len(dcs))
    assert failed == 0 and all(r == 0 for r in rcs), rcs
    for d, w in zip(dcs, want):
        assert np.array_equal(d.d().view(np.uint32), w.view(np.uint32)), d.case.name
    lib.destroy_server(srv)
    for d in dcs:
        d.close()
    for ex in exs.values():
        lib.destroy_executor(ex)


@pytest.mark.gpu
@pytest.mark.parametrize("slots", [1, 2, 4])
def test_heavy_requests_leave_a_stream_to_small_ones(lib, gpu_exec, oracle, slots):
    """Requests with a footprint above 4 MiB (here: a staged host A) are
    heavy: while every other slot holds one, admission passes over them and
    takes the small requests queued behind them (with one slot, heavy
    requests are admitted as they come).  One request per batch, heavy
    requests queued first: every request completes, each D as the oracle's."""
    from test_gpu_parity import assert_parity
    heavy = [T.tucker_case(200 + i, 128, 96, 96, 4) for i in range(3)]     # A alone is 4.5 MiB
    small = [T.tucker_case(210 + i, 12, 10, 8, 3) for i in range(6)]
    cases = [heavy[0], heavy[1], small[0], small[1], heavy[2]] + small[2:]
    rc, srv = lib.create_server(gpu_exec, 1 << 30, slots, 1, lib.SERVER_PAUSED)
    assert rc == 0
    keep, sts = [], []
    for c in cases:
        c2 = c.copy()
        plan, infos = _plan(lib, c2)
        al = T.scalar_bytes(c2.alpha, c2.D.type)
        be = T.scalar_bytes(c2.beta, c2.D.type)
        rc, st = lib.submit(srv, plan, al.ctypes.data, c2.A.ptr(), c2.B.ptr(), be.ctypes.data, None, c2.D.ptr())
        assert rc == 0
        keep.append((c2, plan, infos, al, be))
        sts.append(st)
    assert lib.set_paused(srv, 0) == 0
    for st in sts:
        assert lib.wait(srv, st) == (0, 0)
        lib.destroy_status(st)
    rc, stats = lib.server_stats(srv)
    assert stats["completed"] == len(cases) and stats["batches"] == len(cases), stats
    for c2, plan, infos, al, be in keep:
        _, want = oracle.run(c2, gpu=True)
        assert_parity(c2.D.buf, want, c2)
        _free(lib, plan, infos)
    lib.destroy_server(srv)


@pytest.mark.gpu
def test_concurrent_client_threads(lib, gpu_exec, oracle):
    """Four client threads submit their own mixes at the same time (ctypes
    calls release the GIL, so submits, waits and the dispatcher interleave),
    twice each; every request completes with 0 and every D matches the same
    product run alone."""
    import threading
    from test_group import DevCase, mixed_cases
    from test_gpu_parity import assert_parity
    sets = [[DevCase(lib, c) for c in mixed_cases(41 + 7 * t)] for t in range(4)]
    alone = []
    for dcs in sets:
        row = []
        for d in dcs:
            assert d.run_single(gpu_exec) == 0
            row.append(d.d().copy())
            d.reset_d()
        alone.append(row)
    rc, srv = lib.create_server(gpu_exec, 1 << 30, 4, 8)
    assert rc == 0
    errors = []

    def client(dcs):
        try:
            for _ in range(2):
                sts = []
                for d in dcs:
                    rc, st = lib.submit(srv, d.plan, d.al.ctypes.data, d.ptr["A"], d.ptr["B"], d.be.ctypes.data,
                                        d.c_ptr(), d.ptr["D"])
                    assert rc == 0, rc
                    sts.append(st)
                for st in sts:
                    assert lib.wait(srv, st) == (0, 0)
                    lib.destroy_status(st)
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=client, args=(dcs,)) for dcs in sets]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads), "client threads did not finish"
    assert not errors, errors
    rc, stats = lib.server_stats(srv)
    assert stats["completed"] == 2 * sum(len(s) for s in sets), stats
    lib.destroy_server(srv)
    for dcs, row in zip(sets, alone):
        for d, a in zip(dcs, row):
            assert_parity(d.d(), a, d.case)
            d.close()


@pytest.mark.gpu
def test_budget_counts_workspace_and_split_k_partials(lib, gpu_exec, oracle):
    """Requests on the general path (a unary label v on B keeps them out of
    the persistent kernel): D[m,n] = sum_{k,v} A[m,k] B[k,n,v] with a staged
    host A (1 MiB), B's pre-reduction (1 MiB) and split-K partial sums (64
    slices of one 128 x 128 tile: 4 MiB): 6 MiB of workspace each.  With an
    8 MiB budget only one runs at a time although three streams are free:
    the others stall, idle workspaces are reclaimed, the server never holds
    more than the budget, and every D is right."""
    from test_gpu_parity import assert_parity

    def unary_gemm(seed, M=128, N=128, K=2048, V=2):
        A = T.make_operand(tc.F32, [M, K], "mk", seed)
        B = T.make_operand(tc.F32, [K, N, V], "knv", seed + 1)
        C = T.make_operand(tc.F32, [M, N], "mn", seed + 2)
        D = T.make_operand(tc.F32, [M, N], "mn", seed + 3)
        return T.Case(A, B, C, D, tc.DEFAULT_PREC, 0.75, 0.5, name=f"unary_gemm{seed}")

    cases = [unary_gemm(40 + 4 * i) for i in range(6)]
    plan, infos = _plan(lib, cases[0])
    budget = 8 << 20
    rc, srv = lib.create_server(gpu_exec, budget, 3, 4, lib.SERVER_PAUSED)
    assert rc == 0
    import torch
    keep, sts = [], []
    for c in cases:
        dev = {n: torch.from_numpy(o.buf.copy()).cuda() for n, o in zip("BCD", (c.B, c.C, c.D))}
        torch.cuda.synchronize()
        al = T.scalar_bytes(c.alpha, c.D.type)
        be = T.scalar_bytes(c.beta, c.D.type)
        rc, st = lib.submit(srv, plan, al.ctypes.data, c.A.ptr(), dev["B"].data_ptr(), be.ctypes.data,
                            dev["C"].data_ptr(), dev["D"].data_ptr())
        assert rc == 0, lib.explain(rc)
        keep.append((c, dev, al, be))
        sts.append(st)
    assert lib.set_paused(srv, 0) == 0
    for st in sts:
        assert lib.wait(srv, st) == (0, 0)
        lib.destroy_status(st)
    rc, stats = lib.server_stats(srv)
    assert stats["completed"] == 6 and stats["stalls"] > 0, stats
    assert 5 << 20 <= stats["max_held_bytes"] <= budget, stats
    assert stats["persistent_products"] == 0, stats
    for c, dev, _, _ in keep:
        _, want = oracle.run(c, gpu=True)
        assert_parity(dev["D"].cpu().numpy(), want, c)
    lib.destroy_server(srv)
    _free(lib, plan, infos)


@pytest.mark.gpu
def test_host_operand_requests_join_the_persistent_launch(lib, gpu_exec, oracle):
    """Small FP32 products whose A (and for some B, or C with beta != 0) is
    host memory that is not resident -- pinned or pageable -- and whose D is
    device memory: the batch copies those operands into its workspace and
    runs every request in its persistent launch.  Each D equals the
    oracle's; the server held workspace for the staged operands."""
    import torch
    from test_gpu_parity import assert_parity
    cases = [T.tucker_case(700 + i, 20 + i, 9, 13, 3 + i) for i in range(6)]
    cases += [T.four_index_case(720 + i, (6 + i, 5, 7, 4, 3 + i, 5)) for i in range(4)]
    for c in cases[::3]:
        c.beta = 0.5                                     # C read (and staged where host)
    rc, srv = lib.create_server(gpu_exec, 1 << 30, 2, 16, lib.SERVER_PAUSED)
    assert rc == 0
    keep, sts = [], []
    for n, c in enumerate(cases):
        plan, infos = _plan(lib, c)
        dev = {x: torch.from_numpy(o.buf.view(np.uint8).copy()).cuda() for x, o in zip("ABCD", c.operands())}
        host = {}
        if n % 2:                                        # pinned host A
            host["A"] = torch.from_numpy(c.A.buf.view(np.uint8).copy()).pin_memory()
        ptr = {x: dev[x].data_ptr() + o.offset * o.buf.itemsize for x, o in zip("ABCD", c.operands())}
        ptr["A"] = (host["A"].data_ptr() + c.A.offset * c.A.buf.itemsize) if n % 2 else c.A.ptr()
        if n % 3 == 1:
            ptr["B"] = c.B.ptr()                          # pageable host B
        if n % 3 == 0:
            ptr["C"] = c.C.ptr()                          # pageable host C, read with beta 0.5
        torch.cuda.synchronize()
        al = T.scalar_bytes(c.alpha, c.D.type)
        be = T.scalar_bytes(c.beta, c.D.type)
        rc, st = lib.submit(srv, plan, al.ctypes.data, ptr["A"], ptr["B"], be.ctypes.data, ptr["C"], ptr["D"])
        assert rc == 0, lib.explain(rc)
        keep.append((c, plan, infos, dev, host, al, be))
        sts.append(st)
    assert lib.set_paused(srv, 0) == 0
    for st in sts:
        assert lib.wait(srv, st) == (0, 0)
        lib.destroy_status(st)
    rc, stats = lib.server_stats(srv)
    assert stats["completed"] == len(cases), stats
    assert stats["persistent_products"] == len(cases), stats
    assert stats["max_held_bytes"] > 0, stats
    for c, plan, infos, dev, host, al, be in keep:
        _, want = oracle.run(c, gpu=True)
        assert_parity(dev["D"].cpu().numpy().view(c.D.buf.dtype), want, c)
        _free(lib, plan, infos)
    lib.destroy_server(srv)

"""The C ABI on the GPU: validation codes returned by TAPP_execute_product
itself, a C caller executing products, the threading and stream-ordering
contract of executors, and the scalar copies of product groups."""
import ctypes
import os
import shutil
import subprocess
import threading

import numpy as np
import pytest

import tapp_cases as T
from test_oracle import _case_from_kat, load_kats, negative_cases

tc = T.tc
pytestmark = pytest.mark.gpu


def _plan(lib, case):
    infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in case.operands()]
    _, plan = lib.create_plan([o.op for o in case.operands()], infos, [o.idx for o in case.operands()], case.prec)
    return plan, infos


def _free(lib, plan, infos):
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)


@pytest.mark.parametrize("name,case", [nc for nc in negative_cases() if nc[0] != "int32_truncation"],
                         ids=lambda x: x if isinstance(x, str) else "")
def test_execute_returns_the_oracles_validation_code(lib, gpu_exec, oracle, name, case):
    """Each of the reference's error codes comes back from
    TAPP_execute_product (product.c:197-222) on this library's executor, on a
    caller's bare int executor and with a NULL executor (33 unless an earlier
    check fails), equal to the oracle's code; D is left untouched.  (The
    one negative case the reference accepts, a D stride of 2^32, spans 16 GiB
    and is checked through the plan stage only.)"""
    want = oracle.check(case)
    assert want != 0
    plan, infos = _plan(lib, case)
    al = T.scalar_bytes(case.alpha, case.D.type)
    be = T.scalar_bytes(case.beta, case.D.type)
    d_before = case.D.buf.copy()
    kind = ctypes.c_int(1)
    for ex, expect in ((gpu_exec, want), (ctypes.addressof(kind), want), (0, want)):
        rc = lib.execute(plan, ex, al.ctypes.data, case.A.ptr(), case.B.ptr(), be.ctypes.data,
                         None if case.c_null else case.C.ptr(), case.D.ptr())
        assert rc == expect, (name, rc, expect)
        assert np.array_equal(case.D.buf.view(np.uint8), d_before.view(np.uint8))
    _free(lib, plan, infos)


def _c_array(name, ctype, values):
    body = ", ".join(repr(float(v)) for v in values) or "0"
    return f"static {ctype} {name}[] = {{{body}}};\n"


def _c_program(kats):
    """A C99 caller of the public API (tapp.h only): for each known-answer
    product it builds infos, a plan and an executor, runs
    TAPP_execute_product on host buffers and prints D."""
    src = ['#include <stdio.h>\n#include <stdint.h>\n#include "tapp.h"\n\n']
    calls = []
    for n, k in enumerate(kats):
        dt = k["D_in"]["type"]
        cplx = dt in (tc.C32, tc.C64)
        for op in ("A", "B", "C", "D_in"):
            o = k[op]
            ctype = "double" if o["type"] in (tc.F64, tc.C64) else "float"
            if o["data"] is not None:
                src.append(_c_array(f"d{n}{op[0]}", ctype, o["data"]))
            nm = max(1, len(o["extents"]))
            src.append(f"static int64_t e{n}{op[0]}[{nm}] = {{{', '.join(map(str, o['extents'])) or '0'}}};\n")
            src.append(f"static int64_t s{n}{op[0]}[{nm}] = {{{', '.join(map(str, o['strides'])) or '0'}}};\n")
            src.append(f"static int64_t i{n}{op[0]}[{nm}] = {{{', '.join(map(str, o['idx'])) or '0'}}};\n")
        ctype = "double" if dt in (tc.F64, tc.C64) else "float"
        src.append(_c_array(f"al{n}", ctype, k["alpha"][:2] if cplx else k["alpha"][:1]))
        src.append(_c_array(f"be{n}", ctype, k["beta"][:2] if cplx else k["beta"][:1]))
        calls.append((n, k, cplx))
    src.append("\nint main(void)\n{\n    TAPP_executor ex;\n    if (TAPP_create_executor(&ex)) return 90;\n")
    for n, k, cplx in calls:
        w = 2 if cplx else 1
        src.append("    {\n        TAPP_tensor_info t[4];\n")
        for x, op in enumerate("ABCD"):
            o = k[op if op != "D" else "D_in"]
            src.append(f"        if (TAPP_create_tensor_info(&t[{x}], {o['type']}, {len(o['extents'])}, e{n}{op}, "
                       f"s{n}{op})) return 91;\n")
        ops = [k[o]["op"] for o in ("A", "B", "C", "D_in")]
        src.append(f"        TAPP_tensor_product p;\n"
                   f"        if (TAPP_create_tensor_product(&p, (TAPP_handle)0, {ops[0]}, t[0], i{n}A, {ops[1]}, t[1], "
                   f"i{n}B, {ops[2]}, t[2], i{n}C, {ops[3]}, t[3], i{n}D, {k['prec']})) return 92;\n")
        offs = {op: -k[op if op != "D" else "D_in"]["offset"] * w for op in "ABCD"}
        cptr = "NULL" if k["C"]["data"] is None else f"d{n}C + {offs['C']}"
        src.append(f"        int rc = TAPP_execute_product(p, ex, NULL, al{n}, d{n}A + {offs['A']}, "
                   f"d{n}B + {offs['B']}, be{n}, {cptr}, d{n}D + {offs['D']});\n")
        src.append(f'        printf("call {n} rc %d\\n", rc);\n')
        src.append(f"        for (size_t i = 0; i < sizeof d{n}D / sizeof d{n}D[0]; i++) "
                   f'printf("%.9g\\n", (double)d{n}D[i]);\n')
        src.append("        TAPP_destroy_tensor_product(p);\n"
                   "        for (int x = 0; x < 4; x++) TAPP_destroy_tensor_info(t[x]);\n    }\n")
    src.append("    TAPP_destroy_executor(ex);\n    return 0;\n}\n")
    return "".join(src)


def test_c_caller_executes_known_answers(tmp_path):
    """A C program of this project's own, compiled against include/ and
    linked to the library by name, executes the reference programs'
    products (demo.c, driver.c, the contraction exercise) on the GPU through
    TAPP_execute_product and prints D; every value equals the reference's
    printed answer."""
    cc = shutil.which("gcc")
    if not cc:
        pytest.skip("gcc not found")
    kats = load_kats()
    src = tmp_path / "kat_caller.c"
    src.write_text(_c_program(kats))
    exe = tmp_path / "kat_caller"
    libdir = os.path.join(T.PKG, "lib")
    subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-I", os.path.join(T.ROOT, "include"), str(src), "-L", libdir,
                    "-ltapp_amd", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=120).stdout.split("\n")
    pos = 0
    for n, k in enumerate(kats):
        assert out[pos] == f"call {n} rc {k['rc']}", (n, out[pos])
        pos += 1
        want = np.array(k["D_out"]["data"], dtype=np.float64)
        got = np.array([float(v) for v in out[pos:pos + len(want)]])
        pos += len(want)
        assert np.allclose(got, want, rtol=1e-6, atol=1e-5), (k["program"], k["call"])


def test_threads_share_the_default_context(lib, oracle):
    """Eight threads run products at the same time through a bare int
    executor (the reference's representation), which all share the
    process-wide default context: each D equals the oracle's (the context's
    workspace and stream are taken in turn)."""
    cases = [T.random_case(8100 + t, "everything" if t % 2 else "contraction", tc.F32, max_ext=7, min_ext=3,
                           beta_zero=False) for t in range(8)]
    kind = ctypes.c_int(1)
    results = [None] * len(cases)

    def work(t):
        rcs, ds = [], []
        for _ in range(4):
            rc, d = T.run_lib(lib, cases[t], ctypes.addressof(kind))
            rcs.append(rc)
            ds.append(d)
        results[t] = (rcs, ds)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(len(cases))]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    for c, (rcs, ds) in zip(cases, results):
        _, want = oracle.run(c)
        for rc, d in zip(rcs, ds):
            assert rc == 0
            assert T.normwise_error(d, want, T.written_mask(c.D)) <= 1e-5, c.name


def test_threads_share_one_plan(lib, oracle):
    """Four threads, each with its own executor, execute the same plan at
    once (its offset tables are built once, under the plan's lock)."""
    c = T.random_case(8200, "hadamard_contraction", tc.F32, max_ext=9, min_ext=4, beta_zero=False)
    plan, infos = _plan(lib, c)
    al = T.scalar_bytes(c.alpha, c.D.type)
    be = T.scalar_bytes(c.beta, c.D.type)
    outs = [None] * 4

    def work(t):
        _, ex = lib.create_executor()
        d = c.D.copy()
        rcs = []
        for _ in range(3):
            d.buf[:] = c.D.buf
            rcs.append(lib.execute(plan, ex, al.ctypes.data, c.A.ptr(), c.B.ptr(), be.ctypes.data, c.C.ptr(),
                                   d.ptr()))
        lib.destroy_executor(ex)
        outs[t] = (rcs, d.buf)

    threads = [threading.Thread(target=work, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    _, want = oracle.run(c)
    for rcs, d in outs:
        assert rcs == [0, 0, 0]
        assert T.normwise_error(d, want, T.written_mask(c.D)) <= 1e-5
    _free(lib, plan, infos)


def test_null_stream_binding_orders_after_torch_work(lib, oracle):
    """Binding the null stream (0) is a valid binding: products run ordered
    after work queued there, here a device fill of A that is not
    synchronised before the call."""
    import torch
    c = T.random_case(8300, "contraction", tc.F32, max_ext=12, min_ext=8, beta_zero=True, layouts=False)
    _, ex = lib.create_executor()
    null = ctypes.c_void_p(0)
    assert lib.lib.TAPP_attr_set(ex, tc.ATTR_STREAM, ctypes.byref(null)) == 0
    got = ctypes.c_void_p(1)
    assert lib.lib.TAPP_attr_get(ex, tc.ATTR_STREAM, ctypes.byref(got)) == 0 and not got.value
    plan, infos = _plan(lib, c)
    al = T.scalar_bytes(c.alpha, c.D.type)
    be = T.scalar_bytes(0.0, c.D.type)
    with torch.cuda.stream(torch.cuda.default_stream()):
        a_dev = torch.zeros(c.A.buf.size, dtype=torch.float32, device="cuda")
        b_dev = torch.from_numpy(c.B.buf).cuda()
        d_dev = torch.zeros(c.D.buf.size, dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        a_src = torch.from_numpy(c.A.buf).cuda()
        torch.cuda.synchronize()
        # queued on the null stream, not waited for by the host
        for _ in range(20):
            a_dev.copy_(a_src * 3.0)
        a_dev.copy_(a_src)
        rc = lib.execute(plan, ex, al.ctypes.data, a_dev.data_ptr(), b_dev.data_ptr(), be.ctypes.data, None,
                         d_dev.data_ptr())
    assert rc == 0, lib.explain(rc)
    _, want = oracle.run(c)
    assert T.normwise_error(d_dev.cpu().numpy(), want, T.written_mask(c.D)) <= 1e-5
    _free(lib, plan, infos)
    lib.destroy_executor(ex)


def test_group_keeps_copies_of_alpha_and_beta(lib, gpu_exec, oracle):
    """alpha and beta are read at group creation: the caller's scalar
    buffers are overwritten and freed before the group runs, for products in
    the persistent launch and for products run one by one (here FP64 and
    host operands)."""
    import torch
    from test_group import DevCase
    dev_cases = [T.tucker_case(8400 + i, 12, 9, 7, 3) for i in range(3)]
    for c in dev_cases:
        c.beta = 0.5
    dcs = [DevCase(lib, c) for c in dev_cases]
    host = [T.random_case(8410 + i, "contraction", tc.F64, max_ext=5, beta_zero=False) for i in range(2)]
    hplans = [_plan(lib, c) for c in host]
    scal = [np.array([d.case.alpha], np.float32) for d in dcs] + [np.array([c.alpha], np.float64) for c in host]
    betas = [np.array([d.case.beta], np.float32) for d in dcs] + [np.array([c.beta], np.float64) for c in host]
    hbufs = [c.copy() for c in host]
    rc, g = lib.create_group(gpu_exec, [d.plan for d in dcs] + [p for p, _ in hplans],
                             [s.ctypes.data for s in scal], [d.ptr["A"] for d in dcs] + [c.A.ptr() for c in hbufs],
                             [d.ptr["B"] for d in dcs] + [c.B.ptr() for c in hbufs], [b.ctypes.data for b in betas],
                             [d.ptr["C"] for d in dcs] + [c.C.ptr() for c in hbufs],
                             [d.ptr["D"] for d in dcs] + [c.D.ptr() for c in hbufs])
    assert rc == 0
    assert lib.group_info(g)[1][:2] == [3, 2]
    for s in scal + betas:
        s[:] = np.nan
    del scal, betas
    torch.cuda.synchronize()
    assert lib.execute_group(g) == 0
    for d in dcs:
        _, want = oracle.run(d.case)
        assert T.normwise_error(d.d(), want, T.written_mask(d.case.D)) <= 1e-5
    for c, b in zip(host, hbufs):
        _, want = oracle.run(c)
        assert T.normwise_error(b.D.buf, want, T.written_mask(c.D)) <= 1e-12
    lib.destroy_group(g)
    for d in dcs:
        d.close()
    for p, infos in hplans:
        _free(lib, p, infos)

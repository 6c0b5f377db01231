"""CPU checks of the GPU library's boundary: it loads, exports every symbol
the headers declare, and its host-side behaviour (object API, error text,
validation codes, plan classification) matches the reference without
launching anything on a device."""
import os
import re

import numpy as np
import pytest

import tapp_cases as T
from test_oracle import negative_cases

tc = T.tc
INCLUDE = os.path.join(T.ROOT, "include")


def declared_symbols():
    names = set()
    for root, _, files in os.walk(INCLUDE):
        for f in files:
            if f.endswith(".h"):
                text = open(os.path.join(root, f)).read()
                names |= set(re.findall(r"TAPP_EXPORT[^;{]*?\b(TAPP_\w+)\s*\(", text, re.S))
    return sorted(names)


def test_headers_declare_the_reference_api():
    names = set(declared_symbols())
    for n in tc.TappLib.API + tc.TappLib.EXTENSIONS:
        assert n in names, n


@pytest.mark.parametrize("name", declared_symbols())
def test_library_exports_declared_symbol(lib, name):
    assert hasattr(lib.lib, name), name


def test_explain_error_matches_reference(lib, ref):
    for code in range(0, 15):
        assert lib.explain(code) == ref.explain(code)
    for code in (15, 33, 137, 101, 102, 103, 104):
        assert lib.explain(code)            # defined text, no undefined pointer
    # maxlen semantics
    import ctypes
    buf = ctypes.create_string_buffer(8)
    n = lib.lib.TAPP_explain_error(4, 8, buf)
    assert n == 7 and buf.value == lib.explain(4)[:7].encode()
    assert lib.lib.TAPP_explain_error(4, 0, None) == len(lib.explain(4))


def test_tensor_info_codes_match_reference(lib, ref):
    for L in (lib, ref):
        assert L.create_info(tc.F32, [2, -1], [1, 2])[0] == 15
        rc, h = L.create_info(tc.F32, [2, 3], [1, 2])
        assert rc == 0
        assert L.lib.TAPP_set_nmodes(h, -1) == 14
        arr = (T.tc.ctypes.c_int64 * 2)(4, -5)
        assert L.lib.TAPP_set_extents(h, arr) == 15
        L.destroy_info(h)
    import ctypes
    assert lib.lib.TAPP_create_tensor_info(ctypes.byref(ctypes.c_ssize_t()), 0, -1, None, None) == 14


def test_tensor_info_roundtrip(lib):
    import ctypes
    rc, h = lib.create_info(tc.C64, [3, 4, 5], [1, -3, 40])
    assert rc == 0 and lib.lib.TAPP_get_nmodes(h) == 3
    e = (ctypes.c_int64 * 3)()
    s = (ctypes.c_int64 * 3)()
    lib.lib.TAPP_get_extents(h, e)
    lib.lib.TAPP_get_strides(h, s)
    assert list(e) == [3, 4, 5] and list(s) == [1, -3, 40]
    assert lib.lib.TAPP_set_nmodes(h, 4) == 0 and lib.lib.TAPP_get_nmodes(h) == 4
    lib.destroy_info(h)


def _plan(lib, case):
    infos = [lib.create_info(o.type, o.extents, o.strides)[1] for o in case.operands()]
    _, plan = lib.create_plan([o.op for o in case.operands()], infos, [o.idx for o in case.operands()], case.prec)
    return plan, infos


@pytest.mark.parametrize("name,case", negative_cases(), ids=lambda x: x if isinstance(x, str) else "")
def test_validation_codes_match_oracle(lib, oracle, name, case):
    plan, infos = _plan(lib, case)
    _, ex = lib.create_executor()
    be = T.scalar_bytes(case.beta, case.D.type)
    rc, _, _ = lib.plan_info(plan, ex, be.ctypes.data, None if case.c_null else 1)
    assert rc == oracle.check(case)
    rc33, _, _ = lib.plan_info(plan, 0, be.ctypes.data, None if case.c_null else 1)
    assert rc33 == (rc if rc else 33)
    lib.destroy_executor(ex)
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)


@pytest.mark.parametrize("dtype", [tc.F32, tc.C64])
@pytest.mark.parametrize("scenario", sorted(T.SCENARIOS))
def test_validation_and_groups_on_random_cases(lib, oracle, dtype, scenario):
    for seed in range(4):
        c = T.random_case(seed + 500, scenario, dtype, repeated_D=(seed == 3))
        plan, infos = _plan(lib, c)
        _, ex = lib.create_executor()
        be = T.scalar_bytes(c.beta, c.D.type)
        rc, sizes, ct = lib.plan_info(plan, ex, be.ctypes.data, None if c.c_null else 1)
        assert rc == oracle.check(c)
        if rc == 0:
            orc_rc, cl = oracle.classify(c)
            # group sizes agree with the reference's classification
            ext_k = [cl.ext_contracted[i] for i in range(cl.n_contracted)]
            nK = int(np.prod(ext_k)) if ext_k else 1
            assert sizes[3] == nK or (sizes[3] == 0 and (sizes[4] == 0 or sizes[5] == 0))
            nD = int(np.prod([cl.extents_D[i] for i in range(cl.nmode_D)])) if cl.nmode_D else 1
            pinned = int(np.prod([cl.extents_D[i] for i in range(cl.nmode_D) if cl.strides_D[i] == 0] or [1]))
            assert sizes[0] * sizes[1] * sizes[2] * pinned == nD
        lib.destroy_executor(ex)
        lib.destroy_plan(plan)
        for h in infos:
            lib.destroy_info(h)


def test_groups_of_benchmark_products(lib):
    cases = {
        "c0": (T.four_index16_case(1), [1, 256, 256, 256, 1, 1], 0),
        "c1": (T.c64_feature_case(1), [32, 40, 40 * 32, 24, 8, 6], 3),
    }
    for name, (c, want, ct_want) in cases.items():
        plan, infos = _plan(lib, c)
        _, ex = lib.create_executor()
        be = T.scalar_bytes(c.beta, c.D.type)
        rc, sizes, ct = lib.plan_info(plan, ex, be.ctypes.data, 1)
        assert rc == 0 and sizes == want and ct == ct_want, (name, sizes, ct)
        lib.destroy_executor(ex)
        lib.destroy_plan(plan)
        for h in infos:
            lib.destroy_info(h)


def test_unsupported_precision_is_an_error_code(lib):
    c = T.random_case(1, "contraction", tc.F32, prec=42)
    plan, infos = _plan(lib, c)
    _, ex = lib.create_executor()
    be = T.scalar_bytes(0.0, c.D.type)
    rc, _, _ = lib.plan_info(plan, ex, be.ctypes.data, 1)
    assert rc == 101
    lib.destroy_executor(ex)
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)


C_PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "tapp.h"
#include "tapp_amd.h"

int main(void)
{
    /* D[a,b] = sum_c A[a,c] B[c,b]: the object API a reference caller uses */
    int64_t ea[2] = {4, 3}, sa[2] = {1, 4}, eb[2] = {3, 5}, sb[2] = {1, 3}, ed[2] = {4, 5}, sd[2] = {1, 4};
    int64_t ia[2] = {'a', 'c'}, ib[2] = {'c', 'b'}, id[2] = {'a', 'b'};
    TAPP_tensor_info A, B, C, D;
    if (TAPP_create_tensor_info(&A, TAPP_F32, 2, ea, sa) || TAPP_create_tensor_info(&B, TAPP_F32, 2, eb, sb) ||
        TAPP_create_tensor_info(&C, TAPP_F32, 2, ed, sd) || TAPP_create_tensor_info(&D, TAPP_F32, 2, ed, sd))
        return 1;
    if (TAPP_get_nmodes(A) != 2) return 2;
    TAPP_tensor_product plan;
    if (TAPP_create_tensor_product(&plan, (TAPP_handle)0, TAPP_IDENTITY, A, ia, TAPP_IDENTITY, B, ib,
                                   TAPP_IDENTITY, C, id, TAPP_IDENTITY, D, id, TAPP_DEFAULT_PREC))
        return 3;
    char msg[128];
    size_t n = TAPP_explain_error(0, sizeof msg, msg);
    printf("version=%s explain0=%s n=%zu\n", TAPP_AMD_version(), msg, n);
    if (TAPP_destroy_tensor_product(plan)) return 4;
    TAPP_destroy_tensor_info(A); TAPP_destroy_tensor_info(B);
    TAPP_destroy_tensor_info(C); TAPP_destroy_tensor_info(D);
    return 0;
}
"""


def test_c_program_builds_against_the_headers_and_links(tmp_path):
    """A reference caller's C code compiles against include/ (C99, -Wall
    -Werror), links to the library by name, and runs its object API calls
    without a GPU; the headers also compile as C++ (extern "C" guards)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc")
    cxx = shutil.which("g++")
    if not cc or not cxx:
        pytest.skip("gcc / g++ not found")
    libdir = os.path.join(T.PKG, "lib")
    src = tmp_path / "caller.c"
    src.write_text(C_PROGRAM)
    exe = tmp_path / "caller"
    subprocess.run([cc, "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", INCLUDE, str(src), "-L", libdir,
                    "-ltapp_amd", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    assert out.startswith("version=") and "n=" in out, out
    hdr = tmp_path / "headers.cpp"
    hdr.write_text('#include "tapp.h"\n#include "tapp_amd.h"\nint main() { return TAPP_AMD_ERR_BUSY == 107 ? 0 : 1; }\n')
    subprocess.run([cxx, "-std=c++11", "-Wall", "-Werror", "-I", INCLUDE, "-fsyntax-only", str(hdr)], check=True)

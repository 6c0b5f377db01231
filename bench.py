#!/usr/bin/env python3
"""Benchmark of TAPP_execute_product on one MI355X.

Workload (the precision ladder, a CCSD-like contraction):
    D[a,b,i,j] = sum_{c,d} A[a,b,c,d] B[c,i,d,j]
    a,b,c,d = 128, i,j = 32, dense column-major FP32, beta = 0
    => M = a*b = 16384, N = i*j = 1024, K = c*d = 16384, 5.498e11 flop
A entries N(0,1) exp(-|a-c|/16), B entries N(0,1) exp(-|c-d|/16), both
scaled to max|x| <= 8, seeded, generated on the device.

One step = one TAPP_execute_product call on device-resident operands
(synchronous, as the API is).  `value` is the plain FP32 path; the
split-BF16, BF16 and F16 forms are measured on the same operands and
reported with their max relative error against the library's FP64 path
(TAPP_F64F64_ACCUM_F64).  The CPU baseline is the compiled reference
library (oracle/_ref) on a D-slice with the full contraction length.
"""
import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "reference-implementation_amd"))
import tapp_ctypes as tc  # noqa: E402

PEAK_FP32_TFLOPS = 157.3     # MI355X dense FP32 matrix
PEAK_BF16_TFLOPS = 2500.0    # MI355X dense BF16/FP16 matrix
PEAK_HBM_GBS = 8000.0

V, O = 128, 32


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def ladder_operands(torch, seed=1234):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    idx = torch.arange(V, device="cuda", dtype=torch.float32)
    decay = torch.exp(-(idx[:, None] - idx[None, :]).abs() / 16.0)          # [x, y] -> exp(-|x-y|/16)
    A = torch.randn((V, V, V, V), generator=g, device="cuda", dtype=torch.float32)  # torch dims (d, c, b, a)
    A.mul_(decay.view(1, V, 1, V))                                          # exp(-|a-c|/16): dims c and a
    B = torch.randn((O, V, O, V), generator=g, device="cuda", dtype=torch.float32)  # torch dims (j, d, i, c)
    B.mul_(decay.view(1, V, 1, V))                                          # exp(-|c-d|/16): dims d and c
    for t in (A, B):
        m = t.abs().max().item()
        if m > 8.0:
            t.mul_(8.0 / m)
    D = torch.zeros((O, O, V, V), device="cuda", dtype=torch.float32)       # torch dims (j, i, b, a)
    return A, B, D


class Product:
    """A plan over the ladder layouts (column-major, labels a..j)."""

    def __init__(self, lib, prec, d_type=tc.F32, ext_ij=(O, O)):
        self.lib = lib
        a, b, c, d, i, j = (ord(x) for x in "abcdij")
        ei, ej = ext_ij
        self.infos = [
            lib.create_info(tc.F32, [V, V, V, V], [1, V, V * V, V * V * V])[1],
            lib.create_info(tc.F32, [V, ei, V, ej], [1, V, V * O, V * O * V])[1],
            lib.create_info(d_type, [V, V, ei, ej], [1, V, V * V, V * V * O])[1],
            lib.create_info(d_type, [V, V, ei, ej], [1, V, V * V, V * V * O])[1],
        ]
        idx = [[a, b, c, d], [c, i, d, j], [a, b, i, j], [a, b, i, j]]
        rc, self.plan = lib.create_plan([0, 0, 0, 0], self.infos, idx, prec)
        assert rc == 0
        self.d_type = d_type
        self.alpha = np.array([1.0], dtype=np.float64 if d_type == tc.F64 else np.float32)
        self.beta = np.zeros(1, dtype=self.alpha.dtype)

    def run(self, ex, A, B, D):
        return self.lib.execute(self.plan, ex, self.alpha.ctypes.data, A, B, self.beta.ctypes.data, None, D)

    def close(self):
        self.lib.destroy_plan(self.plan)
        for h in self.infos:
            self.lib.destroy_info(h)


def time_form(lib, torch, prec, split, A, B, D, steps, warmup, tiles=0):
    rc, ex = lib.create_executor()
    assert rc == 0
    lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, int(split))
    lib.attr_set_int(ex, tc.ATTR_TILES, int(tiles))
    lib.attr_set_int(ex, tc.ATTR_TIMING, 1)
    p = Product(lib, prec)
    for _ in range(warmup):
        rc = p.run(ex, A.data_ptr(), B.data_ptr(), D.data_ptr())
        assert rc == 0, lib.explain(rc)
    torch.cuda.synchronize()
    dev_us = []
    t0 = time.perf_counter()
    for _ in range(steps):
        rc = p.run(ex, A.data_ptr(), B.data_ptr(), D.data_ptr())
        assert rc == 0, lib.explain(rc)
        dev_us.append(lib.attr_get_double(ex, tc.ATTR_LAST_DEVICE_US)[1])
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    kern = lib.attr_get_int(ex, tc.ATTR_LAST_KERNEL)[1]
    name = lib.last_kernel_name(ex)[1]
    p.close()
    lib.destroy_executor(ex)
    return wall, float(np.mean(dev_us)), (kern, name)


def fp64_reference(lib, torch, A, B):
    rc, ex = lib.create_executor()
    p = Product(lib, tc.F64F64_ACCUM_F64, d_type=tc.F64)
    D64 = torch.zeros((O, O, V, V), device="cuda", dtype=torch.float64)
    torch.cuda.synchronize()   # torch's fill before the library's stream writes D64
    rc = p.run(ex, A.data_ptr(), B.data_ptr(), D64.data_ptr())
    assert rc == 0, lib.explain(rc)
    p.close()
    lib.destroy_executor(ex)
    return D64


def cpu_baseline(A, B, D_gpu, sample_j=2):
    """The compiled reference (single-threaded, reentrant) on D-slices at full
    K: one thread on the slice i = 0, j < sample_j, then one thread per core,
    each with its own plan on the disjoint slice i = t, j < sample_j."""
    path = os.path.join(ROOT, "oracle", "_ref", "libtapp-reference.so")
    kind = "reference"
    if not os.path.exists(path):
        return None
    import threading
    ref = tc.TappLib(path)
    a_h = A.cpu().numpy()
    b_h = B.cpu().numpy()
    d_h = np.zeros((O, O, V, V), dtype=np.float32)
    cores = max(1, min(16, len(os.sched_getaffinity(0)), O))

    def run_slice(i, out):
        p = Product(ref, tc.DEFAULT_PREC, ext_ij=(1, sample_j))
        _, ex = ref.create_executor()
        t0 = time.perf_counter()
        rc = p.run(ex, a_h.ctypes.data, b_h.ctypes.data + 4 * V * i, d_h.ctypes.data + 4 * V * V * i)
        out[i] = (rc, time.perf_counter() - t0)
        ref.destroy_executor(ex)
        p.close()

    res = {}
    run_slice(0, res)
    rc, dt1 = res[0]
    assert rc == 0
    threads = [threading.Thread(target=run_slice, args=(i, res)) for i in range(cores)]
    t0 = time.perf_counter()
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    dtn = time.perf_counter() - t0
    assert all(res[i][0] == 0 for i in range(cores))
    flops = 2.0 * V * V * 1 * sample_j * V * V
    got = D_gpu.cpu().numpy()[:sample_j, :cores]
    want = d_h[:sample_j, :cores]
    err = float(np.abs(got - want).max() / np.abs(want).max())
    return {"value": cores * flops / dtn / 1e9, "unit": "GFLOP/s", "cores": cores, "kind": kind,
            "sample": f"{cores} threads, thread t on the D-slice i=t, j<{sample_j} (1/{O * O // sample_j} of D each) "
                      f"at full K=16384: {cores * flops:.3e} flop in {dtn:.1f} s",
            "one_core_value": flops / dt1 / 1e9,
            "gpu_vs_reference_slices_normwise_err": err}


def batching_problems(seed=0, count=8192):
    """The uneven batching workload, drawn per problem: 80 % Tucker mode
    products D[r,j,k] = sum_i A[i,j,k] U[r,i] with i,j,k ~ U{4..64},
    r ~ U{2..16}; 20 % four-index products (D[a,b,c,d] = sum A[a,e,b,f] B[d,f,c,e]) with each of the six extents
    ~ U{6..24} (drawn per index).  Returns (kind, extents) tuples."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        if rng.random() < 0.8:
            i, j, k = (int(x) for x in rng.integers(4, 65, 3))
            out.append(("tucker", (i, j, k, int(rng.integers(2, 17)))))
        else:
            out.append(("c0", tuple(int(x) for x in rng.integers(6, 25, 6))))
    return out


def cpu_reference_batch(sample=16):
    """The compiled reference (oracle/_ref) on one host core for the batching
    and serving shape classes: the first `sample` Tucker and four-index
    products of the batching draw, timed one by one and extrapolated per
    class to the 8192-product batch; and one virtual-extent-64 ladder-form
    serving request on a D-slice (i = j = 0, 1/1024 of D) at its full K,
    extrapolated linearly."""
    path = os.path.join(ROOT, "oracle", "_ref", "libtapp-reference.so")
    if not os.path.exists(path):
        return None
    ref = tc.TappLib(path)
    _, ex = ref.create_executor()
    rng = np.random.default_rng(11)
    L = lambda t: [ord(c) for c in t]  # noqa: E731

    def dense(ext):
        st, acc = [], 1
        for x in ext:
            st.append(acc)
            acc *= x
        return st, acc

    def run(ops, alpha, beta, strides=None):
        infos, bufs = [], []
        for n, (ext, lab) in enumerate(ops):
            st, size = dense(ext)
            if strides and strides[n]:
                st, size = strides[n]
            infos.append(ref.create_info(tc.F32, ext, st)[1])
            bufs.append((rng.random(size, dtype=np.float32) * 2 - 1).astype(np.float32))
        _, plan = ref.create_plan([0, 0, 0, 0], infos, [L(x[1]) for x in ops], tc.DEFAULT_PREC)
        al, be = np.array([alpha], np.float32), np.array([beta], np.float32)
        t0 = time.perf_counter()
        rc = ref.execute(plan, ex, al.ctypes.data, bufs[0].ctypes.data, bufs[1].ctypes.data, be.ctypes.data,
                         bufs[2].ctypes.data, bufs[3].ctypes.data)
        dt = time.perf_counter() - t0
        assert rc == 0, rc
        ref.destroy_plan(plan)
        for h in infos:
            ref.destroy_info(h)
        return dt

    probs = batching_problems()
    counts = {"tucker": 0, "c0": 0}
    times = {"tucker": [], "c0": []}
    for kind, e in probs:
        counts[kind] += 1
        if len(times[kind]) >= sample:
            continue
        if kind == "tucker":
            i, j, k, r = e
            times[kind].append(run([([i, j, k], "ijk"), ([r, i], "ri"), ([r, j, k], "rjk"), ([r, j, k], "rjk")],
                                   1.0, 0.0))
        else:
            a, b, c, d, ee, ff = e
            times[kind].append(run([([a, ee, b, ff], "aebf"), ([d, ff, c, ee], "dfce"), ([a, b, c, d], "abcd"),
                                    ([a, b, c, d], "abcd")], 1.5, 0.5))
    mean = {k: float(np.mean(v)) for k, v in times.items()}
    batch_s = sum(counts[k] * mean[k] for k in counts)
    # one ladder-form serving request at virtual extent 64, on the D-slice i = j = 0
    v, o = 64, 32
    slice_s = run([([v, v, v, v], "abcd"), ([v, 1, v, 1], "cidj"), ([v, v, 1, 1], "abij"), ([v, v, 1, 1], "abij")],
                  1.0, 0.0)
    ref.destroy_executor(ex)
    return {"batch_8192_s_extrapolated": batch_s,
            "batch_sample": f"first {sample} Tucker and {sample} four-index products of the seed-0 draw, "
                            f"one core; extrapolated per class to {counts['tucker']} + {counts['c0']}",
            "tucker_ms_mean": 1e3 * mean["tucker"], "four_index_form_ms_mean": 1e3 * mean["c0"],
            "serving_ladder64_request_s_extrapolated": slice_s * (o * o),
            "serving_ladder64_sample": "D-slice i = j = 0 (1/1024 of D) at full K = 4096, one core"}


def batching_bench(lib, torch, steps, prec=tc.DEFAULT_PREC, full=True):
    """The uneven 8192-product batch: one persistent group launch vs one TAPP_execute_product per
    product, same device operands, same GPU.  full=False (profiler runs):
    only the group's launches, no one-call-per-product loop and no autotuned
    group."""
    probs = batching_problems()
    L = lambda s: [ord(c) for c in s]  # noqa: E731
    specs, sizes = [], []
    flops = bytes_ = ideal = 0.0
    for kind, e in probs:
        if kind == "tucker":
            i, j, k, r = e
            ops = [([i, j, k], "ijk"), ([r, i], "ri"), ([r, j, k], "rjk"), ([r, j, k], "rjk")]
            f = 2.0 * r * j * k * i
            b = 4.0 * (i * j * k + r * i + r * j * k)
            alpha, beta = 1.0, 0.0
        else:
            a, bb, c, d, ee, ff = e
            ops = [([a, ee, bb, ff], "aebf"), ([d, ff, c, ee], "dfce"), ([a, bb, c, d], "abcd"), ([a, bb, c, d], "abcd")]
            f = 2.0 * a * bb * c * d * ee * ff
            b = 4.0 * (a * ee * bb * ff + d * ff * c * ee + 2 * a * bb * c * d)
            alpha, beta = 1.5, 0.5
        specs.append((ops, alpha, beta))
        sizes.append([int(np.prod(x[0])) for x in ops])
        flops += f
        bytes_ += b
        ideal += max(f / (PEAK_FP32_TFLOPS * 1e12), b / (PEAK_HBM_GBS * 1e9))
    total = sum(sum(x) for x in sizes)
    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    pool = torch.rand(total, generator=g, device="cuda", dtype=torch.float32).mul_(2.0).sub_(1.0)
    torch.cuda.synchronize()   # operands written on torch's stream before the library reads them
    base = pool.data_ptr()
    off = 0
    plans, infos, ptrs, scal = [], [], [], []
    for (ops, alpha, beta), sz in zip(specs, sizes):
        hs = []
        p4 = []
        for (ext, lab), n in zip(ops, sz):
            st, acc = [], 1
            for x in ext:
                st.append(acc)
                acc *= x
            hs.append(lib.create_info(tc.F32, ext, st)[1])
            p4.append(base + 4 * off)
            off += n
        _, plan = lib.create_plan([0, 0, 0, 0], hs, [L(x[1]) for x in ops], prec)
        plans.append(plan)
        infos.append(hs)
        ptrs.append(p4)
        scal.append((np.array([alpha], dtype=np.float32), np.array([beta], dtype=np.float32)))
    rc, ex = lib.create_executor()
    lib.attr_set_int(ex, tc.ATTR_TIMING, 1)
    if not full:
        rc, grp = lib.create_group(ex, plans, [sc[0].ctypes.data for sc in scal], [p[0] for p in ptrs],
                                   [p[1] for p in ptrs], [sc[1].ctypes.data for sc in scal], [p[2] for p in ptrs],
                                   [p[3] for p in ptrs])
        assert rc == 0, lib.explain(rc)
        assert lib.execute_group(grp) == 0
        dev_us = []
        for _ in range(steps):
            assert lib.execute_group(grp) == 0
            dev_us.append(lib.attr_get_double(ex, tc.ATTR_LAST_DEVICE_US)[1])
        torch.cuda.synchronize()
        lib.destroy_group(grp)
        lib.destroy_executor(ex)
        for plan, hs in zip(plans, infos):
            lib.destroy_plan(plan)
            for h in hs:
                lib.destroy_info(h)
        kernel_s = float(np.mean(dev_us)) * 1e-6
        return {"group_kernel_ms": kernel_s * 1e3, "roofline_frac": ideal / kernel_s, "ideal_ms": ideal * 1e3}
    # one call per product
    for i in range(len(plans)):
        rc = lib.execute(plans[i], ex, scal[i][0].ctypes.data, ptrs[i][0], ptrs[i][1], scal[i][1].ctypes.data,
                         ptrs[i][2], ptrs[i][3])
        assert rc == 0, lib.explain(rc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(len(plans)):
        lib.execute(plans[i], ex, scal[i][0].ctypes.data, ptrs[i][0], ptrs[i][1], scal[i][1].ctypes.data,
                    ptrs[i][2], ptrs[i][3])
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    D_loop = [pool[(p[3] - base) // 4:(p[3] - base) // 4 + sz[3]].clone() for p, sz in zip(ptrs, sizes)]
    torch.cuda.synchronize()   # the copies run on torch's stream; the library's launches use their own
    # one persistent group
    t0 = time.perf_counter()
    rc, grp = lib.create_group(ex, plans, [sc[0].ctypes.data for sc in scal], [p[0] for p in ptrs],
                               [p[1] for p in ptrs], [sc[1].ctypes.data for sc in scal], [p[2] for p in ptrs],
                               [p[3] for p in ptrs])
    create_s = time.perf_counter() - t0
    assert rc == 0, lib.explain(rc)
    _, info = lib.group_info(grp)
    assert lib.execute_group(grp) == 0
    torch.cuda.synchronize()
    dev_us = []
    t0 = time.perf_counter()
    for _ in range(steps):
        assert lib.execute_group(grp) == 0
        dev_us.append(lib.attr_get_double(ex, tc.ATTR_LAST_DEVICE_US)[1])
    torch.cuda.synchronize()
    group_s = (time.perf_counter() - t0) / steps
    def snapshot():
        out = [pool[(p[3] - base) // 4:(p[3] - base) // 4 + sz[3]].clone() for p, sz in zip(ptrs, sizes)]
        torch.cuda.synchronize()
        return out

    def max_rel_err(got, want):
        e = 0.0
        for g_, w_ in zip(got, want):
            den = w_.abs().max().item()
            if den > 0:
                e = max(e, (g_ - w_).abs().max().item() / den)
        return e

    D_group = snapshot()
    err = max_rel_err(D_group, D_loop)
    lib.destroy_group(grp)
    # the same group with the autotuner: thin vs MFMA tiles measured per shape
    # class at creation (fresh cache file per bench run)
    _, tex = lib.create_executor()
    lib.attr_set_int(tex, tc.ATTR_TIMING, 1)
    lib.attr_set_int(tex, tc.ATTR_AUTOTUNE, 1)
    t0 = time.perf_counter()
    rc, tgrp = lib.create_group(tex, plans, [sc[0].ctypes.data for sc in scal], [p[0] for p in ptrs],
                                [p[1] for p in ptrs], [sc[1].ctypes.data for sc in scal], [p[2] for p in ptrs],
                                [p[3] for p in ptrs])
    tuned_create_s = time.perf_counter() - t0
    assert rc == 0, lib.explain(rc)
    assert lib.execute_group(tgrp) == 0
    torch.cuda.synchronize()
    tuned_us = []
    t0 = time.perf_counter()
    for _ in range(steps):
        assert lib.execute_group(tgrp) == 0
        tuned_us.append(lib.attr_get_double(tex, tc.ATTR_LAST_DEVICE_US)[1])
    torch.cuda.synchronize()
    tuned_s = (time.perf_counter() - t0) / steps
    D_tuned = snapshot()
    tuned_err = max_rel_err(D_tuned, D_loop)
    # with the rule's tile classes kept, the tuned group computes every output
    # the same way as the default one: same bits
    repeatable = all(torch.equal(a, b) for a, b in zip(D_tuned, D_group))
    lib.destroy_group(tgrp)
    lib.destroy_executor(tex)
    lib.destroy_executor(ex)
    for plan, hs in zip(plans, infos):
        lib.destroy_plan(plan)
        for h in hs:
            lib.destroy_info(h)
    kernel_s = float(np.mean(dev_us)) * 1e-6
    return {"workload": "8192 uneven products: 80% Tucker D[r,j,k]=sum_i A[i,j,k]U[r,i] (i,j,k~U{4..64}, "
                        "r~U{2..16}), 20% four-index form D[a,b,c,d]=sum A[a,e,b,f]B[d,f,c,e] (extents ~U{6..24} per index), seed 0",
            "precision": "FP32" if prec == tc.DEFAULT_PREC else "BF16BF16_ACCUM_F32",
            "flop": flops, "bytes": bytes_, "group_gflops": flops / group_s / 1e9,
            "group_ms": group_s * 1e3, "group_kernel_ms": kernel_s * 1e3,
            "loop_gflops": flops / loop_s / 1e9, "loop_ms": loop_s * 1e3,
            "speedup_vs_loop": loop_s / group_s, "create_ms": create_s * 1e3,
            "persistent_products": info[0], "one_by_one_products": info[1], "tiles": info[2],
            "roofline_frac": ideal / kernel_s, "ideal_ms": ideal * 1e3,
            "max_rel_err_group_vs_loop": err, "tuned_group_bitwise_equal_to_default": repeatable,
            "autotuned": {"group_ms": tuned_s * 1e3, "group_kernel_ms": float(np.mean(tuned_us)) * 1e-3,
                          "group_gflops": flops / tuned_s / 1e9, "speedup_vs_loop": loop_s / tuned_s,
                          "create_ms_with_tuning": tuned_create_s * 1e3, "max_rel_err_group_vs_loop": tuned_err}}


def batched_product_bench(lib, torch, steps, count=8192, shape=(32, 32, 32, 8)):
    """TAPP_execute_batched_product on one plan -- the Tucker product
    D[r,j,k] = sum_i A[i,j,k] U[r,i], i = j = k = 32, r = 8, FP32 -- over
    `count` operand sets: the first call (builds the batch and keeps it on
    the plan), repeated calls with the same pointers, calls alternating
    between two operand pools (every descriptor patched and uploaded), and
    the same products as a product group replayed."""
    i, j, k, r = shape
    L = lambda t: [ord(c) for c in t]  # noqa: E731
    ops = [([i, j, k], "ijk"), ([r, i], "ri"), ([r, j, k], "rjk"), ([r, j, k], "rjk")]
    sizes = [int(np.prod(e)) for e, _ in ops]
    per = sum(sizes)
    g = torch.Generator(device="cuda")
    g.manual_seed(17)
    pool = torch.rand(2 * count * per, generator=g, device="cuda", dtype=torch.float32).mul_(2.0).sub_(1.0)
    torch.cuda.synchronize()   # written on torch's stream before the library reads it
    base = pool.data_ptr()
    sets = []
    for q in range(2):
        cols = [[], [], [], []]
        for n in range(count):
            off = (q * count + n) * per
            for x in range(4):
                cols[x].append(base + 4 * off)
                off += sizes[x]
        sets.append(cols)
    infos = []
    for ext, _ in ops:
        st, acc = [], 1
        for x in ext:
            st.append(acc)
            acc *= x
        infos.append(lib.create_info(tc.F32, ext, st)[1])
    _, plan = lib.create_plan([0, 0, 0, 0], infos, [L(x[1]) for x in ops], tc.DEFAULT_PREC)
    al, be = np.ones(1, np.float32), np.zeros(1, np.float32)
    _, ex = lib.create_executor()
    lib.attr_set_int(ex, tc.ATTR_TIMING, 1)

    arrays = [[lib.pointer_array(col) for col in cols] for cols in sets]   # built once, as a C caller would

    def call(q):
        a, b, _, d = arrays[q]
        rc = lib.execute_batched(plan, ex, al.ctypes.data, a, b, be.ctypes.data, None, d)
        assert rc == 0, lib.explain(rc)
        return lib.attr_get_int(ex, tc.ATTR_LAST_REUSED)[1], lib.attr_get_double(ex, tc.ATTR_LAST_DEVICE_US)[1]

    t0 = time.perf_counter()
    reused_first, _ = call(0)
    first_s = time.perf_counter() - t0
    out = {"first_call_ms": first_s * 1e3, "first_call_reused": reused_first}
    for label, seq in (("same_pointers", [0] * steps), ("alternating_pools", [1, 0] * max(1, steps // 2))):
        call(seq[0])
        torch.cuda.synchronize()
        reused, dev = [], []
        t0 = time.perf_counter()
        for q in seq:
            rz, du = call(q)
            reused.append(rz)
            dev.append(du)
        wall = (time.perf_counter() - t0) / len(seq)
        out[label] = {"call_ms": wall * 1e3, "kernel_ms": float(np.mean(dev)) * 1e-3, "all_reused": all(reused)}
    D_batched = [pool[(p - base) // 4:(p - base) // 4 + sizes[3]].clone() for p in sets[0][3]]
    torch.cuda.synchronize()
    # the same products as a product group, replayed
    rc, grp = lib.create_group(ex, [plan] * count, [al.ctypes.data] * count, sets[0][0], sets[0][1],
                               [be.ctypes.data] * count, [None] * count, sets[0][3])
    assert rc == 0, lib.explain(rc)
    assert lib.execute_group(grp) == 0
    dev = []
    t0 = time.perf_counter()
    for _ in range(steps):
        assert lib.execute_group(grp) == 0
        dev.append(lib.attr_get_double(ex, tc.ATTR_LAST_DEVICE_US)[1])
    group_s = (time.perf_counter() - t0) / steps
    D_group = [pool[(p - base) // 4:(p - base) // 4 + sizes[3]].clone() for p in sets[0][3]]
    torch.cuda.synchronize()
    lib.destroy_group(grp)
    lib.destroy_executor(ex)
    lib.destroy_plan(plan)
    for h in infos:
        lib.destroy_info(h)
    del pool
    flop = 2.0 * i * j * k * r * count
    out.update({"workload": f"{count} operand sets of D[r,j,k] = sum_i A[i,j,k] U[r,i], i=j=k={i}, r={r}, FP32, "
                            "beta 0, one plan",
                "group_replay": {"call_ms": group_s * 1e3, "kernel_ms": float(np.mean(dev)) * 1e-3},
                "batched_equals_group_bitwise": all(torch.equal(x, y) for x, y in zip(D_batched, D_group)),
                "gflops_repeated_call": flop / (out["same_pointers"]["call_ms"] * 1e-3) / 1e9})
    return out


def serving_bench(lib, torch, rates, nreq, split=False, seed=3, budget=64 << 30, sequential=False, small_only=False):
    """The serving workload: open-loop Poisson arrivals into the serving loop.
    Requests: 5 % large ladder-form products (virtual extent 64:
    D[a,b,i,j] = sum A[a,b,c,d] B[c,i,d,j], a..d = 64, i,j = 32), 20 %
    four-index products (extents 16), 75 % drawn from 256 shapes of the batch mix.
    With `small_only` the large class is left out (the four-index and mix
    shapes keep their 20 : 75 proportion).
    70 % of requests pass a host A that is resident on the plans' handle
    (TAPP_AMD_make_resident: read from its HBM copy, no transfer); 30 % pass
    another pinned host A that the server stages.  B, C and D are device
    tensors (D from a ring per shape).  The server's device memory --
    workspaces (staging, pre-pass buffers, split-K partials), descriptor
    buffers, tables being built -- is admitted against `budget`.  With
    `sequential`, the last rate's trace is also run as one
    TAPP_execute_product call per request, back to back on one executor."""
    rng = np.random.default_rng(seed)
    L = lambda t: [ord(c) for c in t]  # noqa: E731

    def dense(ext):
        st, acc = [], 1
        for x in ext:
            st.append(acc)
            acc *= x
        return st, acc

    shapes = [("ladder64", [([64, 64, 64, 64], "abcd"), ([64, 32, 64, 32], "cidj"), ([64, 64, 32, 32], "abij"),
                            ([64, 64, 32, 32], "abij")], 1.0, 0.0, 2.0 * 64 ** 4 * 32 * 32),
              ("c0", [([16] * 4, "aebf"), ([16] * 4, "dfce"), ([16] * 4, "abcd"), ([16] * 4, "abcd")], 1.5, 0.5,
               2.0 * 16 ** 6)]
    for kind, e in batching_problems(seed=1, count=256):
        if kind == "tucker":
            i, j, k, r = e
            shapes.append(("tucker", [([i, j, k], "ijk"), ([r, i], "ri"), ([r, j, k], "rjk"), ([r, j, k], "rjk")],
                           1.0, 0.0, 2.0 * r * j * k * i))
        else:
            a, b, c, d, ee, ff = e
            shapes.append(("c0mix", [([a, ee, b, ff], "aebf"), ([d, ff, c, ee], "dfce"), ([a, b, c, d], "abcd"),
                                     ([a, b, c, d], "abcd")], 1.5, 0.5, 2.0 * a * b * c * d * ee * ff))
    ring = 16
    cat = []
    _, handle = lib.create_handle()
    for name, ops, alpha, beta, fl in shapes:
        infos = []
        for ext, lab in ops:
            st, _ = dense(ext)
            infos.append(lib.create_info(tc.F32, ext, st)[1])
        _, plan = lib.create_plan([0, 0, 0, 0], infos, [L(x[1]) for x in ops], tc.DEFAULT_PREC, handle=handle)
        n = [dense(x[0])[1] for x in ops]
        A = torch.rand(n[0]) * 2 - 1                       # resident on the handle
        assert lib.make_resident(handle, A.data_ptr(), A.numel() * 4) == 0
        Ah = (torch.rand(n[0]) * 2 - 1).pin_memory()       # staged per request
        B = torch.rand(n[1], device="cuda") * 2 - 1
        C = torch.rand(n[2], device="cuda") * 2 - 1
        Ds = torch.zeros((ring, n[3]), device="cuda")
        cat.append({"plan": plan, "infos": infos, "A": A, "Ah": Ah, "B": B, "C": C, "D": Ds, "flop": fl,
                    "al": np.array([alpha], np.float32), "be": np.array([beta], np.float32), "next": 0})
    torch.cuda.synchronize()
    rc, ex = lib.create_executor()
    if split:
        lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, 1)
    rc, srv = lib.create_server(ex, budget, 4, 64)
    assert rc == 0
    # warm-up: one request of every shape (resident and staged A), so that
    # no kernel is launched for the first time inside a measured rate
    wplans, wal, wA, wB, wbe, wC, wD = [], [], [], [], [], [], []
    for c in cat:
        for slot, a_ptr in enumerate((c["A"].data_ptr(), c["Ah"].data_ptr())):
            wplans.append(c["plan"]); wal.append(c["al"].ctypes.data); wbe.append(c["be"].ctypes.data)
            wA.append(a_ptr); wB.append(c["B"].data_ptr()); wC.append(c["C"].data_ptr())
            wD.append(c["D"][slot].data_ptr())
    failed, rcs, _ = tc.trace_replay(lib, srv, wplans, wal, wA, wB, wbe, wC, wD, [0.0] * len(wplans))
    assert failed == 0 and all(r == 0 for r in rcs)

    def draw(n):
        u = rng.random(n)
        if small_only:
            u = 0.05 + 0.95 * u
        idx = np.where(u < 0.05, 0, np.where(u < 0.25, 1, 2 + rng.integers(0, 256, n)))
        host_a = rng.random(n) >= 0.7
        return idx, host_a

    results = []
    for rate in rates:
        idx, host_a = draw(nreq)
        arrivals = np.cumsum(rng.exponential(1e6 / rate, nreq))
        plans, als, As, Bs, bes, Cs, Ds = [], [], [], [], [], [], []
        flop = 0.0
        for i, h in zip(idx, host_a):
            c = cat[int(i)]
            slot = c["next"] % ring
            c["next"] += 1
            plans.append(c["plan"]); als.append(c["al"].ctypes.data); bes.append(c["be"].ctypes.data)
            As.append(c["Ah"].data_ptr() if h else c["A"].data_ptr())
            Bs.append(c["B"].data_ptr()); Cs.append(c["C"].data_ptr()); Ds.append(c["D"][slot].data_ptr())
            flop += c["flop"]
        lib.server_stats(srv, reset=True)
        # replayed in segments of about a minute of arrivals (one segment at
        # the default request count); the server's statistics span them all
        seg = max(1, int(rate * 60))
        wall, late, failed, rcs = 0.0, 0.0, 0, []
        for s0 in range(0, nreq, seg):
            s1 = min(nreq, s0 + seg)
            base = arrivals[s0 - 1] if s0 else 0.0
            t0 = time.perf_counter()
            f, r, lt = tc.trace_replay(lib, srv, plans[s0:s1], als[s0:s1], As[s0:s1], Bs[s0:s1], bes[s0:s1],
                                       Cs[s0:s1], Ds[s0:s1], list(arrivals[s0:s1] - base))
            wall += time.perf_counter() - t0
            failed += f
            late += lt
            rcs += list(r)
            if s1 < nreq:
                log(f"bench: serving {rate} req/s: {s1} of {nreq} requests done")
        _, st = lib.server_stats(srv)
        assert failed == 0 and all(r == 0 for r in rcs), (failed, [r for r in rcs if r][:5])
        row = {"rate_req_s": rate, "requests": nreq, "p50_us": st["p50_us"], "p99_us": st["p99_us"],
               "mean_us": st["mean_us"], "sustained_gflops": st["gflops"],
               "sustained_req_s": nreq / wall, "offered_gflops": flop / (arrivals[-1] * 1e-6) / 1e9,
               "admission_stalls": st["stalls"], "workspace_reclaims": st["reclaims"],
               "max_device_bytes_held": st["max_held_bytes"], "budget_bytes": budget,
               "max_in_flight": st["max_inflight"], "batches": st["batches"],
               "persistent_products": st["persistent_products"], "wall_s": wall,
               "mean_submit_late_us": late / nreq}
        if sequential and rate == rates[-1]:
            # the same requests, one TAPP_execute_product call at a time
            seq_ex = lib.create_executor()[1]
            f, r, el = tc.trace_sequential(lib, seq_ex, plans, als, As, Bs, bes, Cs, Ds)
            assert f == 0, [x for x in r if x][:5]
            row["one_call_at_a_time"] = {"wall_s": el, "req_s": nreq / el, "gflops": flop / el / 1e9}
            row["served_vs_one_call_at_a_time_req_s"] = (nreq / wall) / (nreq / el)
            row["served_vs_one_call_at_a_time_gflops"] = st["gflops"] / (flop / el / 1e9)
            large = [i for i, c in enumerate(idx) if c == 0]
            if large:
                # the large requests' share of the one-call-at-a-time time:
                # served throughput beats one call at a time by about
                # 1 / share at most, unless serving also runs the large
                # requests faster (two of them overlapping on the device)
                pick = lambda v: [v[i] for i in large]  # noqa: E731
                f, r, el_large = tc.trace_sequential(lib, seq_ex, pick(plans), pick(als), pick(As), pick(Bs),
                                                     pick(bes), pick(Cs), pick(Ds))
                assert f == 0, [x for x in r if x][:5]
                row["one_call_at_a_time"]["large_requests_wall_s"] = el_large
                row["one_call_at_a_time"]["large_requests_share"] = el_large / el
            lib.destroy_executor(seq_ex)
        results.append(row)
        log(f"bench: serving {'split-BF16' if split else 'FP32'} {rate} req/s: p50 {st['p50_us']:.0f} us "
            f"p99 {st['p99_us']:.0f} us, {st['gflops']:.0f} GFLOP/s, stalls {st['stalls']:.0f}")
    lib.destroy_server(srv)
    lib.destroy_executor(ex)
    for c in cat:
        lib.destroy_plan(c["plan"])
        for h in c["infos"]:
            lib.destroy_info(h)
    lib.destroy_handle(handle)                              # frees the resident copies
    return results


def config_cases(lib, torch, steps):
    """Per-call device time and rate of the four-index product (FP32, extents 16) and
    the complex-double feature product (C64 with a diagonal, unary labels, conj(B), a negative and
    padded strides, TAPP_F64F64_ACCUM_F64) on device operands."""
    L = lambda t: [ord(c) for c in t]  # noqa: E731
    out = {}
    operands, specs_by_name = {}, {}

    def span(ext, st):
        lo = sum(min(0, (e - 1) * s) for e, s in zip(ext, st))
        hi = sum(max(0, (e - 1) * s) for e, s in zip(ext, st))
        return -lo, hi - lo + 1

    def run(name, specs, prec, alpha, beta, ctype, flop, bytes_):
        specs_by_name[name] = (specs, prec, alpha, beta, ctype, flop)
        dt = torch.float32 if ctype == tc.F32 else torch.complex128
        infos, bufs, ptrs = [], [], []
        for ext, st, lab, dtype in specs:
            infos.append(lib.create_info(dtype, ext, st)[1])
            off, n = span(ext, st)
            t = (torch.rand(n, device="cuda", dtype=torch.float64) * 2 - 1)
            if dt == torch.complex128:
                t = torch.complex(t, torch.rand(n, device="cuda", dtype=torch.float64) * 2 - 1)
            else:
                t = t.float()
            bufs.append(t)
            ptrs.append(t.data_ptr() + off * t.element_size())
        ops = [0, 1 if name == "c64_feature" else 0, 0, 0]
        torch.cuda.synchronize()   # operands written on torch's stream before the library reads them
        _, plan = lib.create_plan(ops, infos, [L(x[2]) for x in specs], prec)
        npd = np.complex128 if ctype == tc.C64 else np.float32
        al = np.array([alpha], dtype=npd)
        be = np.array([beta], dtype=npd)
        _, ex = lib.create_executor()
        lib.attr_set_int(ex, tc.ATTR_TIMING, 1)
        rc = lib.execute(plan, ex, al.ctypes.data, ptrs[0], ptrs[1], be.ctypes.data, ptrs[2], ptrs[3])
        assert rc == 0, lib.explain(rc)
        torch.cuda.synchronize()
        dev = []
        t0 = time.perf_counter()
        for _ in range(steps):
            lib.execute(plan, ex, al.ctypes.data, ptrs[0], ptrs[1], be.ctypes.data, ptrs[2], ptrs[3])
            dev.append(lib.attr_get_double(ex, tc.ATTR_LAST_DEVICE_US)[1])
        wall = (time.perf_counter() - t0) / steps
        d_us = float(np.mean(dev))
        kname = lib.last_kernel_name(ex)[1]
        torch.cuda.synchronize()
        operands[name] = [b.cpu() for b in bufs]     # D holds the GPU result
        lib.destroy_executor(ex)
        lib.destroy_plan(plan)
        for h in infos:
            lib.destroy_info(h)
        ideal = max(flop / (PEAK_FP32_TFLOPS * 1e12 if ctype == tc.F32 else 78.6e12), bytes_ / (PEAK_HBM_GBS * 1e9))
        out[name] = {"call_us": wall * 1e6, "device_us": d_us, "kernel": kname,
                     "gflops": flop / (d_us * 1e-6) / 1e9,
                     "algorithmic_bytes": bytes_, "achieved_gbs": bytes_ / (d_us * 1e-6) / 1e9,
                     "roofline_frac": ideal / (d_us * 1e-6), "ideal_us": ideal * 1e6}

    e = [16] * 4
    dense = [1, 16, 256, 4096]
    run("four_index16", [(e, dense, "aebf", tc.F32), (e, dense, "dfce", tc.F32), (e, dense, "abcd", tc.F32),
                     (e, dense, "abcd", tc.F32)], tc.DEFAULT_PREC, 1.5, 0.5, tc.F32,
        2.0 * 16 ** 6, 4.0 * 4 * 16 ** 4)
    a, b, c, h, i, u, v = 40, 40, 24, 32, 32, 8, 6
    sA = [-1, 40, 40 * 24, 40 * 24 * 24, 40 * 24 * 24 * 32]
    sB, acc = [], 1
    for x in (h, b, c, i, v):
        sB.append(acc)
        acc = acc * x + 1
    sD = [1, a, a * b, a * b * h]
    run("c64_feature", [([a, c, c, h, u], sA, "acchu", tc.C64), ([h, b, c, i, v], sB, "hbciv", tc.C64),
                     ([a, b, h, i], sD, "abhi", tc.C64), ([a, b, h, i], sD, "abhi", tc.C64)],
        tc.F64F64_ACCUM_F64, 0.7 - 0.3j, 1.1 + 0.2j, tc.C64,
        8.0 * h * a * (b * i) * c, 16.0 * (a * c * h * u + h * b * c * i * v + 2 * a * b * h * i))
    return out, (operands, specs_by_name)


def cpu_reference_products(cases):
    """The four-index and complex-double feature products run whole on one host core through the
    compiled reference (oracle/_ref), on the operands the GPU used; the GPU's
    D against the reference's."""
    path = os.path.join(ROOT, "oracle", "_ref", "libtapp-reference.so")
    if not os.path.exists(path):
        return None
    operands, specs_by_name = cases
    ref = tc.TappLib(path)
    L = lambda t: [ord(c) for c in t]  # noqa: E731
    out = {}
    for name, (specs, prec, alpha, beta, ctype, flop) in specs_by_name.items():
        host = [t.numpy().copy() for t in operands[name]]
        gpu_d = host[3].copy()
        infos = [ref.create_info(dtype, ext, st)[1] for ext, st, lab, dtype in specs]
        _, plan = ref.create_plan([0, 1 if name == "c64_feature" else 0, 0, 0], infos, [L(x[2]) for x in specs], prec)
        npd = np.complex128 if ctype == tc.C64 else np.float32
        al, be = np.array([alpha], dtype=npd), np.array([beta], dtype=npd)
        offs = []
        for (ext, st, lab, dtype), h in zip(specs, host):
            lo = sum(min(0, (e - 1) * x) for e, x in zip(ext, st))
            offs.append(h.ctypes.data - lo * h.itemsize)
        _, ex = ref.create_executor()
        t0 = time.perf_counter()
        rc = ref.execute(plan, ex, al.ctypes.data, offs[0], offs[1], be.ctypes.data, offs[2], offs[3])
        dt = time.perf_counter() - t0
        assert rc == 0, rc
        err = float(np.abs(gpu_d - host[3]).max() / np.abs(host[3]).max())
        out[name] = {"one_core_s": dt, "one_core_gflops": flop / dt / 1e9, "gpu_vs_reference_normwise_err": err}
        ref.destroy_executor(ex)
        ref.destroy_plan(plan)
        for hnd in infos:
            ref.destroy_info(hnd)
    return out


def measured_traffic():
    """HBM bytes per launch of the FP32 ladder kernel, read from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by the copy-kernel
    calibration factors recorded beside them): not measured by this run.
    Returns (bytes or None, where the figure comes from)."""
    path = os.path.join(ROOT, "profiles", "ladder_counters.json")
    try:
        with open(path) as f:
            c = json.load(f)["ladder_fp32"]
        b = c["fetch_kib"] * 1024 * c["fetch_factor"] + c["write_kib"] * 1024 * c["write_factor"]
        return b, (f"read from profiles/ladder_counters.json (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE passes over "
                   f"{c['kernel']}; see {c['source']}), not measured in this run")
    except (OSError, KeyError, ValueError):
        return None, "profiles/ladder_counters.json not found"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-forms", action="store_true", help="skip the BF16/F16/split-BF16 forms")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU reference baseline")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven-batch (product group) measurement")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop sweep")
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests per arrival rate in the serving sweep (the full workload is 200000)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--only-form", choices=["fp32", "bf16", "f16", "split_bf16x3", "split_bf16x6", "split_f16x3"],
                    help="only time one form of the ladder (for profiler runs)")
    ap.add_argument("--only-batch", choices=["fp32", "bf16"],
                    help="only time the uneven batch's group launches (for profiler runs)")
    ap.add_argument("--only-serve", action="store_true",
                    help="only run the serving sweeps (full mix with one call at a time, small requests)")
    ap.add_argument("--tiles", type=int, choices=[0, 1, 2, 3], default=0,
                    help="with --only-form: MFMA tile size (TAPP_AMD_ATTR_TILES: 0 by shape, 1 large, 2 small, 3 tall)")
    args = ap.parse_args()
    # autotuned measurements start from an empty cache on every run
    os.environ.setdefault("TAPP_AMD_TUNE_CACHE", os.path.join(tempfile.mkdtemp(prefix="tapp_bench_tune_"), "tune.txt"))
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (products shard over D tiles; "
              "see the sharding design).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    # The serving loop runs its 4 slots on 4 streams besides the executor's
    # and torch's.  HIP's default of 4 hardware queues per process makes
    # slot streams share an in-order queue, where a small request's kernel
    # waits behind a large product's; 8 gives every stream its own.  Read by
    # the HIP runtime when it starts, so set before torch loads it.
    os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

    import torch
    assert torch.cuda.is_available()
    if args.copy_calibration:
        src = torch.ones(1 << 28, device="cuda", dtype=torch.float32)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 4, "dispatches": 3}))
        return
    lib = tc.TappLib()
    if args.only_batch:
        prec = tc.DEFAULT_PREC if args.only_batch == "fp32" else tc.BF16BF16_ACCUM_F32
        print(json.dumps({"batch": args.only_batch, **batching_bench(lib, torch, args.steps, prec, full=False)}))
        return
    if args.only_serve:
        out = {"fp32": serving_bench(lib, torch, [50000], args.serve_requests, sequential=True),
               "fp32_small_requests": serving_bench(lib, torch, [500000], args.serve_requests, sequential=True,
                                                    small_only=True)}
        print(json.dumps({"serving": out}))
        return
    flops = 2.0 * (V * V) * (O * O) * (V * V)
    log("bench: generating ladder operands")
    A, B, D = ladder_operands(torch)
    torch.cuda.synchronize()
    if args.only_form:
        prec, split = {"fp32": (tc.DEFAULT_PREC, 0), "bf16": (tc.BF16BF16_ACCUM_F32, 0),
                       "f16": (tc.F16F16_ACCUM_F32, 0), "split_bf16x3": (tc.DEFAULT_PREC, 1),
                       "split_bf16x6": (tc.DEFAULT_PREC, 2), "split_f16x3": (tc.DEFAULT_PREC, 3)}[args.only_form]
        w, du, k = time_form(lib, torch, prec, split, A, B, D, args.steps, args.warmup, args.tiles)
        print(json.dumps({"form": args.only_form, "kernel_us": du, "gflops": flops / w / 1e9, "kernel_family": k[0],
                          "kernel": k[1]}))
        return

    log(f"bench: FP32 plain, warmup {args.warmup}, steps {args.steps}")
    wall, dev_us, kern = time_form(lib, torch, tc.DEFAULT_PREC, False, A, B, D, args.steps, args.warmup)
    D_fp32 = D.clone()
    torch.cuda.synchronize()   # before the next products overwrite D from the library's stream
    achieved = flops / (dev_us * 1e-6) / 1e12
    value = flops / wall / 1e9

    forms = {}
    if not args.no_forms:
        log("bench: FP64 reference result")
        D64 = fp64_reference(lib, torch, A, B)
        den = D64.abs().max().item()
        keep = D64.abs() >= 1e-3 * den        # elementwise errors where |D| is not tiny

        def elementwise(Dx):
            e = ((Dx.double() - D64).abs() / D64.abs())[keep]
            return {"median": e.median().item(), "max": e.max().item(), "floor": "|D| >= 1e-3 max|D|"}

        forms["fp32"] = {"gflops": value, "kernel_us": dev_us,
                         "max_rel_err_vs_fp64": (D_fp32.double() - D64).abs().max().item() / den,
                         "elementwise_rel_err_vs_fp64": elementwise(D_fp32),
                         "roofline_frac": flops / (dev_us * 1e-6) / (PEAK_FP32_TFLOPS * 1e12),
                         "peak_unit_tflops": PEAK_FP32_TFLOPS}
        peak_unit = {"split_bf16x3": PEAK_BF16_TFLOPS / 3, "split_bf16x6": PEAK_BF16_TFLOPS / 6,
                     "split_f16x3": PEAK_BF16_TFLOPS / 3, "bf16": PEAK_BF16_TFLOPS, "f16": PEAK_BF16_TFLOPS}
        for name, prec, split in (("split_bf16x3", tc.DEFAULT_PREC, 1), ("split_bf16x6", tc.DEFAULT_PREC, 2),
                                  ("split_f16x3", tc.DEFAULT_PREC, 3),
                                  ("bf16", tc.BF16BF16_ACCUM_F32, 0), ("f16", tc.F16F16_ACCUM_F32, 0)):
            log(f"bench: form {name}")
            w, du, k = time_form(lib, torch, prec, split, A, B, D, max(1, args.steps // 4), 1)
            err = (D.double() - D64).abs().max().item() / den
            forms[name] = {"gflops": flops / w / 1e9, "kernel_us": du, "kernel_family": k[0], "kernel": k[1],
                           "max_rel_err_vs_fp64": err,
                           "elementwise_rel_err_vs_fp64": elementwise(D),
                           "roofline_frac": flops / (du * 1e-6) / (peak_unit[name] * 1e12),
                           "peak_unit_tflops": peak_unit[name],
                           "speedup_vs_fp32": (flops / w / 1e9) / value}
        del D64

    cpu = None
    if not args.no_cpu:
        log("bench: CPU reference baseline on a D-slice")
        cpu = cpu_baseline(A, B, D_fp32)

    log("bench: four-index and complex-double feature products per call")
    others, config_operands = config_cases(lib, torch, max(5, args.steps))
    if cpu is not None:
        log("bench: four-index and complex-double feature products on one host core through the reference")
        cpu["products"] = cpu_reference_products(config_operands)
        log("bench: batching and serving shape classes on one host core through the reference (sampled)")
        cpu["classes"] = cpu_reference_batch()
    del config_operands

    batched = {}
    if not args.no_batch:
        del A, B, D, D_fp32
        torch.cuda.empty_cache()
        log("bench: uneven batch, one persistent launch vs one call per product")
        batched["fp32"] = batching_bench(lib, torch, max(2, args.steps // 2))
        batched["bf16"] = batching_bench(lib, torch, max(2, args.steps // 2), prec=tc.BF16BF16_ACCUM_F32)
        torch.cuda.empty_cache()
        log("bench: TAPP_execute_batched_product, repeated calls on one plan")
        batched["batched_product"] = batched_product_bench(lib, torch, max(4, args.steps // 2))

    serving = {}
    if not args.no_serve:
        torch.cuda.empty_cache()
        log("bench: serving loop, open-loop Poisson arrivals")
        rates = [1000, 5000, 20000, 50000]
        serving["requests_per_rate"] = args.serve_requests
        serving["requests_note"] = (f"{args.serve_requests} requests per arrival rate (this run's --serve-requests; "
                                    "the workload's full trace is 200000 requests)")
        serving["fp32"] = serving_bench(lib, torch, rates, args.serve_requests, sequential=True)
        serving["split_bf16"] = serving_bench(lib, torch, [5000], args.serve_requests, split=True)
        log("bench: serving loop, the mix without its large ladder-form class")
        serving["fp32_small_requests"] = serving_bench(lib, torch, [500000], args.serve_requests, sequential=True,
                                                       small_only=True)
        log("bench: serving loop with a 128 MiB device-memory budget")
        serving["fp32_budget_128MiB"] = serving_bench(lib, torch, [20000], args.serve_requests, budget=128 << 20)

    traffic, traffic_src = measured_traffic()
    out = {
        "metric": "GFLOP/s + % MFMA/HBM roofline, split-BF16 vs FP32 GPU; p50/p99 req latency",
        "value": value,
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "precision ladder D[a,b,i,j] = sum_{c,d} A[a,b,c,d] B[c,i,d,j], a..d=128, i,j=32, "
                               "FP32, TAPP_DEFAULT_PREC",
                   "M": V * V, "N": O * O, "K": V * V, "flop_per_step": flops},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": kern[1], "kernel_us": dev_us, "kernel_family": kern[0]},
        "cpu_baseline": cpu,
        "forms": forms,
        "products": others,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

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


def time_form(lib, torch, prec, split, A, B, D, steps, warmup):
    rc, ex = lib.create_executor()
    assert rc == 0
    lib.attr_set_int(ex, tc.ATTR_SPLIT_BF16, 1 if split else 0)
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
    p.close()
    lib.destroy_executor(ex)
    return wall, float(np.mean(dev_us)), kern


def fp64_reference(lib, torch, A, B):
    rc, ex = lib.create_executor()
    p = Product(lib, tc.F64F64_ACCUM_F64, d_type=tc.F64)
    D64 = torch.zeros((O, O, V, V), device="cuda", dtype=torch.float64)
    rc = p.run(ex, A.data_ptr(), B.data_ptr(), D64.data_ptr())
    assert rc == 0, lib.explain(rc)
    p.close()
    lib.destroy_executor(ex)
    return D64


def cpu_baseline(A, B, D_gpu, sample_j=2):
    """The compiled reference on the D-slice i = 0, j < sample_j at full K."""
    path = os.path.join(ROOT, "oracle", "_ref", "libtapp-reference.so")
    kind = "reference"
    if not os.path.exists(path):
        return None
    ref = tc.TappLib(path)
    a_h = A.cpu().numpy()
    b_h = B.cpu().numpy()
    d_h = np.zeros((O, O, V, V), dtype=np.float32)
    p = Product(ref, tc.DEFAULT_PREC, ext_ij=(1, sample_j))
    _, ex = ref.create_executor()
    t0 = time.perf_counter()
    rc = p.run(ex, a_h.ctypes.data, b_h.ctypes.data, d_h.ctypes.data)
    dt = time.perf_counter() - t0
    assert rc == 0
    ref.destroy_executor(ex)
    p.close()
    flops = 2.0 * V * V * 1 * sample_j * V * V
    got = D_gpu.cpu().numpy()[:sample_j, :1]
    want = d_h[:sample_j, :1]
    err = float(np.abs(got - want).max() / np.abs(want).max())
    return {"value": flops / dt / 1e9, "unit": "GFLOP/s", "cores": 1, "kind": kind,
            "sample": f"D-slice i=0, j<{sample_j} (1/{O * O // sample_j} of D) at full K=16384, "
                      f"{flops:.3e} flop in {dt:.1f} s",
            "gpu_vs_reference_slice_normwise_err": err}


def measured_traffic():
    """HBM bytes per launch of the FP32 ladder kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by the copy-kernel
    calibration factors recorded beside them)."""
    path = os.path.join(ROOT, "profiles", "ladder_counters.json")
    try:
        with open(path) as f:
            c = json.load(f)["ladder_fp32"]
        return c["fetch_kib"] * 1024 * c["fetch_factor"] + c["write_kib"] * 1024 * c["write_factor"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-forms", action="store_true", help="skip the BF16/F16/split-BF16 forms")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU reference baseline")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (products shard over D tiles; "
              "see the sharding design).  Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

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
    flops = 2.0 * (V * V) * (O * O) * (V * V)
    log("bench: generating ladder operands")
    A, B, D = ladder_operands(torch)
    torch.cuda.synchronize()

    log(f"bench: FP32 plain, warmup {args.warmup}, steps {args.steps}")
    wall, dev_us, kern = time_form(lib, torch, tc.DEFAULT_PREC, False, A, B, D, args.steps, args.warmup)
    D_fp32 = D.clone()
    achieved = flops / (dev_us * 1e-6) / 1e12
    value = flops / wall / 1e9

    forms = {}
    if not args.no_forms:
        log("bench: FP64 reference result")
        D64 = fp64_reference(lib, torch, A, B)
        den = D64.abs().max().item()
        forms["fp32"] = {"gflops": value, "kernel_us": dev_us,
                         "max_rel_err_vs_fp64": (D_fp32.double() - D64).abs().max().item() / den}
        for name, prec, split in (("split_bf16x3", tc.DEFAULT_PREC, True), ("bf16", tc.BF16BF16_ACCUM_F32, False),
                                  ("f16", tc.F16F16_ACCUM_F32, False)):
            log(f"bench: form {name}")
            w, du, k = time_form(lib, torch, prec, split, A, B, D, max(1, args.steps // 4), 1)
            err = (D.double() - D64).abs().max().item() / den
            forms[name] = {"gflops": flops / w / 1e9, "kernel_us": du, "kernel": k, "max_rel_err_vs_fp64": err,
                           "speedup_vs_fp32": (flops / w / 1e9) / value}
        del D64

    cpu = None
    if not args.no_cpu:
        log("bench: CPU reference baseline on a D-slice")
        cpu = cpu_baseline(A, B, D_fp32)

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
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": measured_traffic(),
                     "kernel": "k_gett_f32<CONTIG_ROW, CONTIG_K> (v_mfma_f32_32x32x2f32)", "kernel_us": dev_us, "kernel_family": kern},
        "cpu_baseline": cpu,
        "forms": forms,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

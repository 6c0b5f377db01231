#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ (run in a container that
has the reference sources at /root/reference and oracle/_ref built).

1. kat_reference_programs.json: every TAPP_execute_product call made by the
   reference's deterministic programs (test/demo.c, examples/driver/driver.c,
   examples/exercise_contraction/answers/exercise_contraction_answers.c),
   captured by linking them against kat_recorder.c, which forwards each call
   to the compiled reference library.  The exercise's documented expected
   output (exercise_contraction_answers.c:229-236) is checked here against
   the captured D before the file is written.
2. ref_four_index16.npz: the four-index contraction (D[a,b,c,d] = 1.5 sum_{e,f}
   A[a,e,b,f] B[d,f,c,e] + 0.5 C, extents 16, FP32, seed 1) computed by the
   compiled reference, full D.
3. ref_c64_feature_sample.npz: the complex-double feature contraction
   (TAPP_F64F64_ACCUM_F64: D[a,b,h,i] = alpha sum_{c,u,v} A[a,c,c,h,u]
   conj(B[h,b,c,i,v]) + beta C, seed 1) computed by the compiled reference;
   its D storage (26 MB) is sampled at 4096 positions drawn with seed 11,
   which are stored with their values.

The fixtures are data only (inputs and outputs).  Nothing from the reference
sources is stored.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
REF = os.environ.get("TAPP_REFERENCE_DIR", "/root/reference")
REFLIB = os.path.join(ROOT, "oracle", "_ref", "libtapp-reference.so")

PROGRAMS = {
    "demo": [f"{REF}/test/demo.c", f"{REF}/test/helpers.c"],
    "driver": [f"{REF}/examples/driver/driver.c", f"{REF}/test/helpers.c"],
    "exercise_contraction_answers": [
        f"{REF}/examples/exercise_contraction/answers/exercise_contraction_answers.c",
        f"{REF}/test/helpers.c"],
}


def capture_programs(tmp):
    rec = os.path.join(tmp, "librecorder.so")
    subprocess.check_call(["gcc", "-std=c11", "-O1", "-fPIC", "-shared", "-w",
                           os.path.join(HERE, "kat_recorder.c"), "-ldl", "-o", rec])
    cases = []
    for name, srcs in PROGRAMS.items():
        exe = os.path.join(tmp, name)
        subprocess.check_call(["gcc", "-std=c99", "-O1", "-w", f"-I{REF}/api/include", f"-I{REF}/test",
                               *srcs, rec, f"-Wl,-rpath,{tmp}", "-lm", "-o", exe])
        log = os.path.join(tmp, name + ".jsonl")
        env = dict(os.environ, TAPP_REF_LIB=REFLIB, TAPP_KAT_OUT=log)
        stdout = subprocess.run([exe], env=env, check=True, capture_output=True, text=True).stdout
        with open(log) as f:
            for n, line in enumerate(f):
                c = json.loads(line)
                c["program"] = name
                c["call"] = n
                cases.append(c)
        print(f"{name}: {n + 1} calls; stdout {len(stdout)} bytes")
    return cases


def check_exercise(cases):
    # exercise_contraction_answers.c:229-236 documents the printed D as a 3x3
    # grid, rows 53.090 53.090 61.090 / 53.090 61.090 61.090 / 61.090 x3.
    ex = [c for c in cases if c["program"] == "exercise_contraction_answers"][-1]
    d = ex["D_out"]
    vals = np.array(d["data"], dtype=np.float64)
    ext, st, off = d["extents"], d["strides"], d["offset"]
    assert len(ext) == 2, ext
    grid = np.array([[vals[i * st[0] + j * st[1] - off] for j in range(ext[1])] for i in range(ext[0])])
    documented = np.array([[53.09, 53.09, 61.09], [53.09, 61.09, 61.09], [61.09, 61.09, 61.09]])
    assert np.allclose(grid, documented, atol=5e-4), grid
    ex["documented_expected_rows"] = documented.tolist()


def four_index16(tmp):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import tapp_cases as tc  # noqa: E402
    case = tc.four_index16_case(seed=1)
    out = tc.run_reference(REFLIB, case)
    np.savez_compressed(os.path.join(HERE, "ref_four_index16.npz"), D=out, seed=np.int64(1))
    print("four_index16: D", out.shape, float(np.abs(out).max()))


def c64_feature_sample():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import tapp_cases as tc  # noqa: E402
    case = tc.c64_feature_case(seed=1)
    out = tc.run_reference(REFLIB, case)
    idx = np.sort(np.random.default_rng(11).choice(out.size, 4096, replace=False)).astype(np.int64)
    np.savez_compressed(os.path.join(HERE, "ref_c64_feature_sample.npz"), idx=idx, D=out[idx], seed=np.int64(1))
    print("c64_feature: D sample", idx.size, "of", out.size, float(np.abs(out[idx]).max()))


def main():
    with tempfile.TemporaryDirectory() as tmp:
        cases = capture_programs(tmp)
        check_exercise(cases)
        with open(os.path.join(HERE, "kat_reference_programs.json"), "w") as f:
            json.dump({"source": "reference programs run against the compiled reference library",
                       "cases": cases}, f, indent=None, separators=(",", ":"))
        four_index16(tmp)
    c64_feature_sample()


if __name__ == "__main__":
    if sys.argv[1:] == ["c64_feature"]:
        c64_feature_sample()
    else:
        main()

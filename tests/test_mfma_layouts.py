"""Checks the lane layouts the GETT kernels assume for the MFMA builtins
(operand lane -> (row, k), accumulator (lane, register) -> (row, col))
against what the hardware computes, with one-hot inputs on one wave."""
import ctypes
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PROBE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "probe", "libtk_probe.so")

# (probe id, block size S, registers NR, A/B lane -> (row, k), acc (lane, r) -> (row, col))
LAYOUTS = {
    "f64_16x16x4": (0, 16, 4,
                    lambda p: (p % 16, p // 16),
                    lambda lane, r: (lane // 16 + 4 * r, lane % 16)),
    "f32_32x32x2": (1, 32, 16,
                    lambda p: (p % 32, p // 32),
                    lambda lane, r: (8 * (r // 4) + 4 * (lane // 32) + (r % 4), lane % 32)),
    "bf16_32x32x16_item0": (2, 32, 16,
                            lambda p: (p % 32, 8 * (p // 32)),
                            lambda lane, r: (8 * (r // 4) + 4 * (lane // 32) + (r % 4), lane % 32)),
}


def run_probe(which, nr):
    import torch  # noqa: F401  (one HIP runtime in the process)
    lib = ctypes.CDLL(PROBE, mode=os.RTLD_LOCAL)
    lib.tk_probe.argtypes = [ctypes.c_int, ctypes.c_void_p]
    out = np.zeros(128 * 64 * nr, dtype=np.float64)
    assert lib.tk_probe(which, out.ctypes.data) == 0
    return out.reshape(128, 64, nr)


def decode(out, S, nr):
    """Observed layout: for A-lane p the output positions of its row and the
    B-lane feeding each of them; for B-lane q the positions of its column."""
    rows = {p: {(l, r) for l in range(64) for r in range(nr) if out[p, l, r] != 0} for p in range(64)}
    feed = {p: {(l, r): int(out[p, l, r]) - 1 for (l, r) in rows[p]} for p in range(64)}
    cols = {q: {(l, r) for l in range(64) for r in range(nr) if out[64 + q, l, r] != 0} for q in range(64)}
    return rows, feed, cols


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_mfma_lane_layout(name):
    which, S, nr, opnd, acc = LAYOUTS[name]
    out = run_probe(which, nr)
    rows, feed, cols = decode(out, S, nr)
    bad = []
    for p in range(64):
        if name.startswith("bf16"):
            pass
        i, k = opnd(p)
        want_rows = {(l, r) for l in range(64) for r in range(nr) if acc(l, r)[0] == i}
        if rows[p] != want_rows:
            bad.append(f"A-lane {p}: row positions {sorted(rows[p])[:6]}... expected row {i}")
            continue
        for (l, r), q in feed[p].items():
            iq, kq = opnd(q)          # B uses the same lane -> (col, k) map
            if kq != k or iq != acc(l, r)[1]:
                bad.append(f"A-lane {p} at (lane {l}, reg {r}): fed by B-lane {q}")
                break
    for q in range(64):
        j, k = opnd(q)
        want_cols = {(l, r) for l in range(64) for r in range(nr) if acc(l, r)[1] == j}
        if name.startswith("bf16") and q // 32 > 0:
            # B-lane q holds k = 8 * (q // 32) + item; with one-hot item 0 on
            # A only for a = 1 everywhere, every column still appears
            pass
        if cols[q] != want_cols:
            bad.append(f"B-lane {q}: col positions {sorted(cols[q])[:6]}... expected col {j}")
    if bad:
        # print the observed map so a wrong assumption can be read off the log
        obs = {}
        for p in range(64):
            for (l, r) in rows[p]:
                obs.setdefault((l, r), [None, None])[0] = p
        for q in range(64):
            for (l, r) in cols[q]:
                obs.setdefault((l, r), [None, None])[1] = q
        lines = [f"(lane {l:2d}, reg {r:2d}): row of A-lane {v[0]}, col of B-lane {v[1]}"
                 for (l, r), v in sorted(obs.items())[:64 * nr]]
        print("\n".join(lines))
    assert not bad, "\n".join(bad[:20])

"""GEMM routing tuner: writes hifde3d_amd/gemm_tune.txt (runs on the GPU).

Every batched product of a factorization goes to one of three kernels --
0: FMA 64 x 64 tiles, 1: matrix-core 128 x 128 tiles, 2: matrix-core 64 x 64
tiles.  Without a table a fixed rule decides (depth 64 or less and thin
outputs on the FMA kernel, 64 x 64 matrix-core tiles where 128 x 128 ones
would waste over a quarter of their output, 128 x 128 otherwise).  This
script replaces the rule, class by class, with measurements:

1. factor the recorded workloads with the table cleared and option
   "gemm_record" on: every descriptor (element size, transA, lower, m, n, k,
   descriptors in its call, kind chosen) is logged;
2. group the descriptors by routing class (the library's key: element size,
   transA, lower, depth bin, bin of the smaller output side, bin of the
   call's descriptor count);
3. per class, take the calls that carry the most of the class's flops (up to
   --samples), and time that part of each call on every candidate kernel
   (hifde_gemm_time: packed operands, the launch repeated --reps times);
4. keep the fastest candidate when it beats the fixed rule's choice for the
   class by more than --margin (the rule's choice otherwise), and write one
   line per class.

Usage: python -m hifde3d_amd.tune [--out PATH] [--quick]
"""
import argparse
import os
import sys
import time
from collections import defaultdict


HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import hifde3d_amd as H  # noqa: E402
from hifde3d_amd import problems as pb  # noqa: E402

# the library's bins (gemm_tune_key in dense_batched.hip)
K_EDGES = (16, 32, 64, 128, 256, 512, 1024)
MN_EDGES = (31, 63, 127, 255, 511)
NB_EDGES = (1, 7, 63, 511)


def _bin(v, edges):
    b = 0
    while b < len(edges) and v > edges[b]:
        b += 1
    return b


def class_of(esize, transA, lower, m, n, k, nb):
    return (esize, transA, lower, _bin(k, K_EDGES), _bin(min(m, n), MN_EDGES), _bin(nb, NB_EDGES))


def workloads(quick):
    """(name, N, P, eps, precision, coefficient builder)"""
    out = [("n128_lognormal_mixed", 127, 4, 1e-4, "mixed", lambda n: pb.lognormal_coeff(n, 2)),
           ("n128_lognormal_fp64", 127, 4, 1e-4, "fp64", lambda n: pb.lognormal_coeff(n, 2))]
    if not quick:
        for NC in (24, 32, 40, 64, 96):
            P = pb.C4_LEAF[NC]
            out.append((f"batch_nc{NC}_mixed", NC - 1, P, 1e-4, "mixed",
                        lambda n: pb.pow10_coeff(n, 7, -2.0, 2.0)))
        out.append(("n128_leaf8_mixed", 127, 8, 1e-3, "mixed",
                    lambda n: pb.lognormal_coeff(n, 4)))
    return out


def record(name, N, P, eps, precision, coeff):
    NC = N + 1
    f = H.HIFFactor(N, P, eps, precision=precision)
    f.set_option("gemm_record", 1)
    f.set_stencil(coeff(NC), pb.zeros_coeff(NC), True)
    assert f.Factor() == 0
    log = f.gemm_log().copy()
    f.close()
    print(f"[tune] {name}: {len(log)} descriptors in {int(log[:, 0].max()) + 1 if len(log) else 0}"
          " calls", flush=True)
    return log


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "gemm_tune.txt"))
    ap.add_argument("--samples", type=int, default=3)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--margin", type=float, default=0.05)
    ap.add_argument("--quick", action="store_true")
    args = ap.parse_args()
    H.gemm_table_clear()
    t0 = time.time()
    # class -> list of (flops, esize, transA, lower, [(m, n, k)], default kinds)
    cls = defaultdict(list)
    for wl in workloads(args.quick):
        log = record(*wl)
        calls = defaultdict(list)
        for row in log:
            calls[(wl[0], int(row[0]))].append(row)
        for key, rows in calls.items():
            per = defaultdict(list)
            for r in rows:
                _, es, ta, lo, m, n, k, nb, kind = (int(x) for x in r)
                if m < 1 or n < 1 or k < 1:
                    continue  # empty products (a scaling of C at most)
                per[class_of(es, ta, lo, m, n, k, nb)].append((m, n, k, kind))
            for c, items in per.items():
                fl = sum(2.0 * m * n * k * (0.5 if c[2] else 1.0) for m, n, k, _ in items)
                cls[c].append((fl, items))
    lines = []
    gain_total = 0.0
    for c in sorted(cls):
        es, ta, lo = c[0], c[1], c[2]
        prec = "mixed" if es == 4 else "fp64"
        samples = sorted(cls[c], key=lambda x: -x[0])[:args.samples]
        core_ok = all(m >= 32 and n >= 32 and k > 0 for _, it in samples for m, n, k, _ in it)
        kinds = [0, 1, 2] if core_ok else [0]
        # the fixed rule's choice for the class: its most common kind
        dk = defaultdict(int)
        for _, it in samples:
            for *_, kind in it:
                dk[kind] += 1
        default = max(dk, key=dk.get)
        t = {}
        for kind in kinds:
            tot = 0.0
            for _, it in samples:
                shapes = [(m, n, k) for m, n, k, _ in it]
                tot += H.gemm_time(prec, ta, lo, shapes, kind, args.reps)
            t[kind] = tot
        best = min(t, key=t.get)
        if default not in t:
            default = 0
        choice = best if t[best] < (1.0 - args.margin) * t[default] else default
        gain_total += t[default] - t[choice]
        times = " ".join(f"k{k}={t[k] * 1e6:.1f}us" for k in kinds)
        lines.append(f"{es} {ta} {lo} {c[3]} {c[4]} {c[5]} {choice}  # rule {default}; {times}")
        print(f"[tune] class {c}: rule {default} -> {choice} ({times})", flush=True)
    with open(args.out, "w") as fh:
        fh.write("# GEMM routing table written by hifde3d_amd/tune.py (python -m hifde3d_amd.tune)\n")
        fh.write("# esize transA lower kbin mnbin nbbin kind   (kind 0 FMA 64x64, 1 matrix-core "
                 "128x128, 2 matrix-core 64x64)\n")
        fh.write(f"# {len(lines)} classes from the recorded workloads; sampled launch time saved "
                 f"against the fixed rule: {gain_total * 1e3:.3f} ms\n")
        for ln in lines:
            fh.write(ln + "\n")
    print(f"[tune] wrote {args.out}: {len(lines)} classes in {time.time() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()

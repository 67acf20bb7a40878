#!/usr/bin/env python3
"""Autotuner for the batched-mode tile table (run on an MI355X).

For every row of the table -- call class (reduce of 4-byte elements,
reduce of 8-byte elements, fcollect, alltoall, broadcast) x size bucket
(destination bytes per PE: <64 KiB, <1 MiB, <16 MiB, >=16 MiB) -- it builds
a batch of calls of that class and size on a 64-PE world, then tunes by
coordinate descent: the reduce kernel shape (wide 8x2 or narrow 1x8), the
tile length (block iterations or KiB; 0 = each call alone through the
per-call kernel), and the grid cap of the family launch.  Each candidate is timed as the median of 3 graph launches with
HIP events.  The winners are written to tuning/mi355x.tune (next to lib/),
which osc_init loads.

    python osss-ucx_amd/autotune.py [--out PATH] [--report PATH]
"""
import argparse
import ctypes as C
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import osc  # noqa: E402
import osc_calls  # noqa: E402

BUCKETS = [(4 << 10, 64 << 10), (64 << 10, 1 << 20), (1 << 20, 16 << 20),
           (16 << 20, 48 << 20)]
CLASSES = ["reduce4", "reduce8", "fcollect", "alltoall", "broadcast"]
WORLD = 64
TARGET_BYTES = 3e9  # algorithmic bytes per timed batch


def make_calls(rng, world, cls, bucket):
    lo, hi = BUCKETS[bucket]
    calls, total = [], 0.0
    while total < TARGET_BYTES and len(calls) < 2048:
        P = int(rng.integers(2, 65))
        start = int(rng.integers(0, WORLD - P + 1))
        per_pe = osc_calls.log_uniform_int(rng, lo, hi - 1)
        if cls in (0, 1):
            op, dt = [("sum", "f32"), ("max", "f32")][len(calls) % 2] \
                if cls == 0 else [("sum", "f64"), ("xor", "u64")][len(calls) % 2]
            n = max(1, min(per_pe // osc_calls.ESZ[dt], world.max_reduce))
            c = osc_calls.reduce_call(world, op, dt, P, start, n)
            total += 2.0 * P * n * osc_calls.ESZ[dt]
        elif cls == 4:
            block = max(16, per_pe // 16 * 16)
            c = osc_calls.move_call(world, osc.CALL_BROADCAST, P, start,
                                    block)
            total += (P + 1.0) * block
        else:
            block = max(16, per_pe // P // 16 * 16)
            kind = osc.CALL_FCOLLECT if cls == 2 else osc.CALL_ALLTOALL
            c = osc_calls.move_call(world, kind, P, start, block)
            total += (P + P * P if cls == 2 else 2.0 * P * P) * block
        calls.append(c)
    return calls, total


def time_batch(calls, stream):
    b = osc.Batch(calls)
    t = osc.Timer(stream)
    try:
        b.launch()
        osc.lib().osc_stream_sync(stream)
        ms = []
        for _ in range(3):
            t.start()
            b.launch()
            t.stop()
            ms.append(t.ms())
    finally:
        t.close()
        b.close()
    return sorted(ms)[1]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tuning",
                                                  "mi355x.tune"))
    ap.add_argument("--report", default=None)
    args = ap.parse_args()
    L = osc.lib()
    world = osc_calls.CallWorld(npes=WORLD, max_reduce=4 << 20,
                                move_bytes=64 << 20, dest_bytes=640 << 20)
    stream = L.osc_default_stream()
    rng = np.random.default_rng(2024)
    report = ["| class | bucket | shape | tile | grid | ms | GB/s | "
              "default ms | gain |", "|---|---|---|---|---|---|---|---|---|"]
    t0 = time.time()
    for cls in range(5):
        for bucket in range(4):
            calls, total = make_calls(rng, world, cls, bucket)
            calls, nw = osc.plan(calls, world.dest_off, world.dest_bytes)
            s, t, g = C.c_int(), C.c_int(), C.c_int()
            L.osc_tune_get(cls, bucket, C.byref(s), C.byref(t), C.byref(g))
            best = (s.value, t.value, g.value)
            L.osc_tune_set(cls, bucket, *best)
            default_ms = best_ms = time_batch(calls, stream)

            def trial(cand):
                L.osc_tune_set(cls, bucket, *cand)
                return time_batch(calls, stream)

            shapes = [0, 3] if cls < 2 else [0]
            # tile 0: every call of the row runs alone through the per-call
            # kernel (a launch of its own inside the batch)
            tiles = [0, 1, 2, 4, 8] if cls < 2 else [0, 16, 32, 64, 128, 256]
            grids = [512, 1024, 2048, 4096, 8192]
            for sh in shapes:
                if sh == best[0]:
                    continue
                ms = trial((sh, best[1], best[2]))
                if ms < best_ms:
                    best, best_ms = (sh, best[1], best[2]), ms
            for tl in tiles:
                if tl == best[1]:
                    continue
                ms = trial((best[0], tl, best[2]))
                if ms < best_ms:
                    best, best_ms = (best[0], tl, best[2]), ms
            for gr in grids:
                if gr == best[2]:
                    continue
                ms = trial((best[0], best[1], gr))
                if ms < best_ms:
                    best, best_ms = (best[0], best[1], gr), ms
            L.osc_tune_set(cls, bucket, *best)
            line = "| %s | %d | %d | %d | %d | %.3f | %.0f | %.3f | %.2fx |" % (
                CLASSES[cls], bucket, best[0], best[1], best[2], best_ms,
                total / (best_ms / 1e3) / 1e9, default_ms,
                default_ms / best_ms)
            report.append(line)
            print(line, "(%d calls, %d waves, %.0fs)" % (
                len(calls), nw, time.time() - t0), flush=True)
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    osc.check(L.osc_tune_save(args.out.encode()), "osc_tune_save")
    if args.report:
        with open(args.report, "w") as f:
            f.write("\n".join(report) + "\n")
    world.close()
    print("saved", args.out)


if __name__ == "__main__":
    main()

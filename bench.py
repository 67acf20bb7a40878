#!/usr/bin/env python3
"""Benchmark of the collective hot path on one MI355X.

Workload (the headline metric's configuration): the team reduction
op/type matrix at 64 emulated PEs x 4,194,304 elements per PE --
and/or/xor on uint64 (random bits), min/max/sum on int64 (uniform in
[-2^40, 2^40]), prod on int64 ({-1, 1, 2}), sum/min/max on float32 and
float64 (N(0, 1)), prod on float32 and float64 (uniform [0.999, 1.001]).
Inputs are generated on the device and resident in HBM before timing.  A
step is one pass over the 15 cells (15 reductions on SHMEM_TEAM_WORLD).

value = algorithmic bytes (2 * P * n * sizeof(T) per reduction: every
source slice read once, every destination slice written once) summed over
the timed steps / elapsed seconds, in GB/s (1e9 B/s).  Each timed step is
bracketed by two HIP events on the library stream (none between its 15
reductions: an event between two calls leaves the GPU idle for a few
microseconds).  The per-cell figures and the roofline object (the kernels of the
cell with the largest total time) come from a second pass of the same
steps with an event before every reduction; value_with_cell_events is the
aggregate of that pass.

    python bench.py [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "osss-ucx_amd"))

import ctypes as C  # noqa: E402

import osc  # noqa: E402

PEAK_HBM_GBPS = 8000.0  # MI355X HBM3E, AMD published
_T0 = time.perf_counter()


def log(msg):
    """Progress line on stderr (stdout carries only the JSON line)."""
    print("[bench %7.1fs] %s" % (time.perf_counter() - _T0, msg),
          file=sys.stderr, flush=True)
P, N = 64, 4194304
CELLS = [  # (op, dtype, dist, a, b)
    ("and", "u64", "bits", 0, 0), ("or", "u64", "bits", 0, 0),
    ("xor", "u64", "bits", 0, 0),
    ("min", "i64", "int_range", -2.0 ** 40, 2.0 ** 40),
    ("max", "i64", "int_range", -2.0 ** 40, 2.0 ** 40),
    ("sum", "i64", "int_range", -2.0 ** 40, 2.0 ** 40),
    ("prod", "i64", "set_m1_1_2", 0, 0),
    ("sum", "f32", "normal", 0.0, 1.0), ("min", "f32", "normal", 0.0, 1.0),
    ("max", "f32", "normal", 0.0, 1.0),
    ("prod", "f32", "real_range", 0.999, 1.001),
    ("sum", "f64", "normal", 0.0, 1.0), ("min", "f64", "normal", 0.0, 1.0),
    ("max", "f64", "normal", 0.0, 1.0),
    ("prod", "f64", "real_range", 0.999, 1.001),
]
ESIZE = {"u64": 8, "i64": 8, "f32": 4, "f64": 8}


def max_rel_err_exact(got, src, op):
    """Largest |got - exact| / |exact| over the columns of src[P, k]; the
    exact sum or product comes from rational arithmetic on the inputs."""
    from fractions import Fraction
    worst = 0.0
    for j in range(src.shape[1]):
        col = [Fraction(float(x)) for x in src[:, j]]
        ex = col[0]
        for x in col[1:]:
            ex = ex + x if op == "sum" else ex * x
        if ex != 0:
            worst = max(worst, float(abs(Fraction(float(got[j])) - ex) / abs(ex)))
    return worst


def accuracy(L, h, src, dest, windows=4, width=256, seed=9):
    """Float sum/prod cells of the matrix: max relative error vs the exact
    value on sampled columns (windows of `width` consecutive elements), in
    the default accurate mode and in the reference's recursive-doubling
    order in the element type's own arithmetic (bit-identical to the
    reference's results), for every member's copy of PE 0 and PE 63."""
    import numpy as np
    rng = np.random.default_rng(seed)
    starts = sorted(int(x) * width for x in
                    rng.choice(N // width, windows, replace=False))
    res = {}
    for op, dt, dist, a, b in CELLS:
        if dt not in ("f32", "f64") or op not in ("sum", "prod"):
            continue
        npt = np.float32 if dt == "f32" else np.float64
        es = ESIZE[dt]
        s = src[(dt, dist, a, b)]
        cols = np.concatenate([np.stack([h.get(s + o * es, width, npt, pe)
                                         for pe in range(P)])
                               for o in starts], axis=1)
        cell = {}
        for name, mode in (("accurate", osc.FP_ACCURATE),
                           ("reference_order", osc.FP_REFERENCE)):
            osc.check(L.osc_set_fp_mode(mode), "fp mode")
            osc.check(L.osc_reduce(h.world, osc.OPS[op], osc.DTYPES[dt], dest,
                                   s, N, None), "osc_reduce")
            h.sync()
            err = 0.0
            for pe in (0, P - 1):
                got = np.concatenate([h.get(dest + o * es, width, npt, pe)
                                      for o in starts])
                err = max(err, max_rel_err_exact(got, cols, op))
            cell[name + "_max_rel_err"] = err
        res[op + "_" + dt] = cell
    res["sample"] = "%d columns (%d windows of %d), all %d PEs' inputs" % (
        windows * width, windows, width, P)
    return res


def last_kernels(L):
    """The kernels the last reduce call launched, as rocprof names them."""
    buf = C.create_string_buffer(4096)
    L.osc_last_reduce_kernels(buf, len(buf))
    return [k for k in buf.value.decode().split(";") if k]


HELPERS = ("k_wait_ticket", "__amd_rocclr", "k_fill", "k_sink_clear",
           "k_xor_read", "k_fill64")


def _pmc_pass(counter, cell, mode):
    """One rocprofv3 counter pass over `--traffic-probe`: per-dispatch
    values of `counter` in order, as (kernel name, value) pairs."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        raise RuntimeError("rocprofv3 not found")
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d,
               "-o", "pmc", "--", sys.executable, os.path.abspath(__file__),
               "--traffic-probe", str(cell), "--fp-mode", mode]
        env = dict(os.environ, TMPDIR="/tmp")
        pr = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE,
                              stderr=subprocess.PIPE, text=True,
                              start_new_session=True)
        try:
            out, err = pr.communicate(timeout=120)
        except subprocess.TimeoutExpired:
            os.killpg(pr.pid, signal.SIGKILL)
            pr.communicate()
            raise RuntimeError("rocprofv3 %s pass timed out" % counter)
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"),
                          recursive=True)
        if not files:
            raise RuntimeError("rocprofv3 %s pass wrote no counter file "
                               "(rc %d): %s" % (counter, pr.returncode,
                                                err[-500:]))
        rows = []
        for fn in files:
            with open(fn) as f:
                for r in csv.DictReader(f):
                    if r.get("Counter_Name") == counter:
                        rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"],
                                     float(r["Counter_Value"])))
        rows.sort()
        return [(k, v) for _, k, v in rows]


def measured_traffic(cell, mode):
    """HBM bytes per call of CELLS[cell], measured now: two rocprofv3
    passes (FETCH_SIZE, then WRITE_SIZE) over a child process that runs a
    warm-up call, three 1 GiB copies and three calls of the cell.  Each
    counter is scaled by what it reads for the copies (1 GiB read, 1 GiB
    written); the cell's dispatches after the last copy are summed and
    divided by three.  Returns (bytes, detail)."""
    det = {}
    total = 0.0
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        rows = _pmc_pass(counter, cell, mode)
        copies = [i for i, (k, _) in enumerate(rows) if "k_copy<" in k]
        if len(copies) < 3:
            raise RuntimeError("%s: calibration copies not found" % counter)
        scale = float(1 << 30) / (sum(rows[i][1] for i in copies[-3:]) / 3.0)
        cell_rows = [(k, v) for k, v in rows[copies[-1] + 1:]
                     if not any(h in k for h in HELPERS)]
        per_call = sum(v for _, v in cell_rows) * scale / 3.0
        det[counter] = {"scale_to_bytes": round(scale, 4),
                        "bytes_per_call": round(per_call),
                        "dispatches": len(cell_rows)}
        total += per_call
    return total, det


CPU_CELLS = [("xor", "u64"), ("sum", "i64"), ("sum", "f32"), ("sum", "f64"),
             ("max", "f64")]


def cpu_baseline():
    """The reference's algorithms on the host cores, one forked process per
    PE over one MAP_SHARED heap (the oracle's restatement: rec_dbl with its
    pSync handshakes and the wrapper's two WORLD barriers, Bruck fcollect
    with the byte-granular rotation, shift-exchange alltoall).  Runs in a
    child process that never loads the HIP library, so no PE process holds
    a device handle.  Main value: 5 of the 15 matrix cells at the full
    64 PEs x 4,194,304 (one warm-up and one timed call each)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    total_b = total_s = 0.0
    cells = {}
    for op, dt in CPU_CELLS:
        sec = O.cpu_bench_procs("reduce", P, N, 1, op, dt)
        b = 2.0 * P * N * ESIZE[dt]
        total_b += b
        total_s += sec
        cells[f"{op}_{dt}_GBps"] = round(b / sec / 1e9, 3)
    other = {"cells": cells,
             "int_sum_8pe_64K_call_us": round(1e6 * O.cpu_bench_procs(
                 "reduce", 8, 65536, 51, "sum", "i32"), 1)}
    sec = O.cpu_bench_procs("alltoall", 64, 65536, 5)
    other["alltoall_64pe_64KiB_GBps"] = round(2 * 64 * 64 * 65536 / sec / 1e9,
                                              2)
    # fcollect at 256 PEs x 1 MiB needs 256 x (1 + 256) MiB = 64.25 GiB
    need = 256 * 257 * (1 << 20)
    avail = 0
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemAvailable:"):
                    avail = int(line.split()[1]) * 1024
    except OSError:
        pass
    if avail >= need * 1.25:
        sec = O.cpu_bench_procs("fcollect", 256, 1 << 20, 1)
        other["fcollect_256pe_1MiB_GBps"] = round(
            (256 + 256 * 256) * (1 << 20) / sec / 1e9, 3)
    else:
        other["fcollect_256pe_1MiB"] = "skipped (RAM: %.1f GiB available, " \
            "%.2f GiB needed)" % (avail / 2 ** 30, need / 2 ** 30)
        sec = O.cpu_bench_procs("fcollect", 64, 65536, 3)
        other["fcollect_64pe_64KiB_GBps"] = round(
            (64 + 64 * 64) * 65536 / sec / 1e9, 3)
    return {"value": round(total_b / total_s / 1e9, 3), "unit": "GB/s",
            "cores": cores, "kind": "port",
            "sample": "rec_dbl restatement (oracle/orc.c), one forked process "
                      "per PE over a MAP_SHARED heap on %d cores, 64 PEs x "
                      "4194304 (full size), cells %s, one timed call each "
                      "after a warm-up" % (cores, ", ".join(
                          f"{o} {d}" for o, d in CPU_CELLS)),
            "other": other,
            "other_sample": "one process per PE: int sum 8 PEs x 65536 "
                            "(median of 51 calls), shift-exchange alltoall "
                            "64 PEs x 64 KiB (median of 5), Bruck fcollect "
                            "with rotation " + (
                                "256 PEs x 1 MiB (one call)"
                                if "fcollect_256pe_1MiB_GBps" in other else
                                "64 PEs x 64 KiB (median of 3; 256 PEs x "
                                "1 MiB skipped for host RAM)")}


def cpu_baseline_subprocess():
    import subprocess
    r = subprocess.run([sys.executable, os.path.abspath(__file__),
                        "--cpu-baseline"], capture_output=True, text=True,
                       timeout=900)
    if r.returncode != 0:
        raise RuntimeError("CPU baseline failed: " + r.stderr[-2000:])
    return json.loads(r.stdout.strip().splitlines()[-1])


def _median_event_ms(L, stream, fn, reps=10):
    t = osc.Timer(stream)
    fn()
    L.osc_stream_sync(stream)
    ms = []
    for _ in range(reps):
        t.start()
        fn()
        t.stop()
        ms.append(t.ms())
    t.close()
    return sorted(ms)[len(ms) // 2]


def bench_movement():
    """The data-movement workloads at full size: shmem_int64_alltoall with 65,536-byte
    blocks on 64 PEs (2*P^2*b bytes) and shmem_int32_fcollect of 1 MiB per
    PE on 256 PEs (P*m + P^2*m bytes), inputs resident, HIP-event medians.
    Also the latency of the small plain reduce: shmem_int_sum_reduce, 8 PEs x 65,536 int32,
    kernel time and the blocking call's wall time, and of shmem_long_sum_reduce and
    shmem_double_sum_reduce on 16 elements at 64 and 128 PEs (the small-call
    kernels)."""
    import numpy as np
    L = osc.lib()
    out = {}
    P, blk = 64, 8192
    with osc.Heap(npes=P, heap_bytes=2 * P * blk * 8 + (1 << 20)) as h:
        s, d = h.malloc(P * blk * 8), h.malloc(P * blk * 8)
        h.fill(None, s, P * blk, "i64", "a2a_pattern", blk, 0)
        st = L.osc_default_stream()
        ms = _median_event_ms(L, st, lambda: L.osc_alltoall(
            h.world, d, s, blk * 8, st))
        b = 2.0 * P * P * blk * 8
        out["alltoall_64pe_64KiB"] = {"ms": round(ms, 4),
                                      "GBps": round(b / ms / 1e6, 1),
                                      "pct_peak": round(b / ms / 1e6 /
                                                        PEAK_HBM_GBPS * 100, 1)}
    P, n = 256, 262144
    with osc.Heap(npes=P, heap_bytes=(P + 1) * n * 4 + (1 << 20)) as h:
        s, d = h.malloc(n * 4), h.malloc(P * n * 4)
        h.fill(None, s, n, "i32", "fcol_pattern")
        st = L.osc_default_stream()
        ms = _median_event_ms(L, st, lambda: L.osc_fcollect(
            h.world, d, s, n * 4, st), reps=5)
        b = P * n * 4.0 + P * P * n * 4.0
        # write-only ceiling over the same footprint: every member's
        # destination region (P x 256 MiB) filled, timed as one group
        sl = L.osc_slice_bytes()
        t = osc.Timer(st)

        def fill_all():
            for j in range(P):
                osc.check(L.osc_device_fill64(d + j * sl, P * n * 4, j, st),
                          "fill")
        fill_all()
        h.sync()
        t.start()
        fill_all()
        t.stop()
        wr = P * P * n * 4.0 / (t.ms() / 1e3) / 1e9
        t.close()
        out["fcollect_256pe_1MiB"] = {"ms": round(ms, 3),
                                      "GBps": round(b / ms / 1e6, 1),
                                      "pct_peak": round(b / ms / 1e6 /
                                                        PEAK_HBM_GBPS * 100, 1),
                                      "write_ceiling_64GiB_GBps": round(wr, 1),
                                      "frac_of_write_ceiling": round(
                                          b / ms / 1e6 / wr, 4)}
    P, n = 8, 65536
    with osc.Heap(npes=P, heap_bytes=4 << 20) as h:
        s, d = h.malloc(n * 4), h.malloc(n * 4)
        h.fill(None, s, n, "i32", "int_range", -1000, 1000, seed=0)
        st = L.osc_default_stream()
        kern = _median_event_ms(L, st, lambda: L.osc_reduce(
            h.world, osc.OPS["sum"], osc.DTYPES["i32"], d, s, n, st), reps=50)
        fn = osc.typed("shmem_int_sum_reduce")
        walls = []
        for _ in range(200):
            t0 = time.perf_counter()
            fn(h.world, d, s, n)
            walls.append((time.perf_counter() - t0) * 1e6)
        walls.sort()
        out["int_sum_8pe_64K"] = {"kernel_us": round(kern * 1e3, 2),
                                     "python_call_p50_us": round(walls[100], 2),
                                     "python_call_p99_us": round(walls[198], 2)}
    # small calls on large teams, 16 elements per PE: int64 (the cross-lane
    # butterfly kernel) and double (the gather kernel, team-order fold)
    n = 16
    for P in (64, 128):
        with osc.Heap(npes=P, heap_bytes=1 << 20) as h:
            s, d = h.malloc(n * 8), h.malloc(n * 8)
            st = L.osc_default_stream()
            for tn, dt, dist, a, b in (("long", "i64", "int_range", -1000, 1000),
                                       ("double", "f64", "normal", 0.0, 1.0)):
                h.fill(None, s, n, dt, dist, a, b, seed=2)
                kern = _median_event_ms(L, st, lambda: L.osc_reduce(
                    h.world, osc.OPS["sum"], osc.DTYPES[dt], d, s, n, st),
                    reps=50)
                fn = osc.typed(f"shmem_{tn}_sum_reduce")
                walls = []
                for _ in range(200):
                    t0 = time.perf_counter()
                    fn(h.world, d, s, n)
                    walls.append((time.perf_counter() - t0) * 1e6)
                walls.sort()
                out[f"{tn}_sum_{P}pe_16"] = {
                    "kernel_us": round(kern * 1e3, 2),
                    "python_call_p50_us": round(walls[100], 2),
                    "python_call_p99_us": round(walls[198], 2)}
    # the same blocking calls timed from C (bin/latency: 2000 calls each,
    # CLOCK_MONOTONIC), as a C program makes them: the Python loops above
    # add the ctypes call (about 2 us)
    import subprocess
    exe = os.path.join(ROOT, "osss-ucx_amd", "bin", "latency")
    try:
        r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-500:])
        lat = json.loads(r.stdout.strip().splitlines()[-1])
        for k, v in lat.items():
            out.setdefault(k, {}).update({"call_p50_us": v["p50_us"],
                                          "call_p90_us": v["p90_us"],
                                          "call_p99_us": v["p99_us"]})
        out["call_latency_source"] = "bin/latency (osss-ucx_amd/examples/" \
                                     "latency.c), 2000 calls from C"
    except (OSError, RuntimeError, ValueError, subprocess.TimeoutExpired) as e:
        out["call_latency_source"] = "C timing failed: %s" % e
    return out


def bench_batched(seed=4, ncalls=4096):
    """The batched workload: ncalls independent calls (osc_calls.batched_mix) in one
    captured graph vs. the same calls issued one per kernel, on a 128-PE
    world; every call's result is verified batched == one-by-one by
    per-call digests, wave by wave."""
    import numpy as np
    import osc_calls
    world = osc_calls.CallWorld(npes=128, max_reduce=1048576,
                                move_bytes=32 << 20, dest_bytes=1 << 30)
    try:
        rng = np.random.default_rng(seed)
        calls = osc_calls.batched_mix(rng, world, ncalls)
        calls, nwaves = osc.plan(calls, world.dest_off, world.dest_bytes)
        L = world.h.L
        stream = L.osc_default_stream()
        # correctness, wave by wave (destinations of one wave are disjoint)
        match = True
        for w in range(nwaves):
            bw = osc.Batch([c for c in calls if c.wave == w])
            bw.launch_serial()
            d1 = bw.digests()
            bw.launch()
            d2 = bw.digests()
            match = match and bool(np.array_equal(d1, d2))
            bw.close()
        b = osc.Batch(calls)
        info = b.info()
        t = osc.Timer(stream)
        b.launch()  # capture + warm
        b.launch_serial()
        world.h.sync()
        serial, batched = [], []
        for _ in range(3):
            t.start()
            b.launch_serial()
            t.stop()
            serial.append(t.ms())
            t.start()
            b.launch()
            t.stop()
            batched.append(t.ms())
        t.close()
        b.close()
        ms_b, ms_s = sorted(batched)[1], sorted(serial)[1]
        gb = info.algorithmic_bytes
        return {"calls": ncalls, "world_pes": 128, "waves": nwaves,
                "launches_in_graph": info.nlaunches,
                "tiles": int(info.ntiles),
                "algorithmic_GB": round(gb / 1e9, 3),
                "batched_ms": round(ms_b, 3), "serial_ms": round(ms_s, 3),
                "batched_GBps": round(gb / (ms_b / 1e3) / 1e9, 1),
                "serial_GBps": round(gb / (ms_s / 1e3) / 1e9, 1),
                "speedup": round(ms_s / ms_b, 3),
                "digests_match": match}
    finally:
        world.close()


# served admission order: shortest first, with one second of waiting worth
# 500 GB of algorithmic bytes (about 9% of a second of work at the measured
# rate), so large requests are not starved.  A sweep of the weight (0 to
# 1e13 B/s, three request streams) gave the lowest p99 at 5e11 on every
# stream; lower weights cut p50 further but let p99 and the maximum grow.
AGE_BYTES_PER_S = 5e11


def bench_served(seed=5, nreq=2000, rate=20000.0, verify=96):
    """The served workload: 16 teams of 8..128 PEs in a 128-PE world send Poisson
    requests at `rate` (60% f32/f64/i64 sum|max reduce with nreduce
    log-uniform 256..4M, 20% fcollect, 20% alltoall, blocks 1 KiB..1 MiB).
    Heap (inputs + destination arena) is sized to a 200 GB budget; the
    admission budget is what the inputs leave of it.  Runs continuous
    batching and the per-call baseline on the same request stream, then a
    verification run with digests: sampled requests must give identical
    results alone, and the f32 sums' error vs the exact sum is measured
    on sampled columns."""
    import numpy as np
    import osc_calls
    budget_total = 200e9
    npes = 128
    inputs = (4 + 8 + 8) * 4194304 + (128 << 20)   # f32, f64, i64 + moves
    arena = int((budget_total / npes - inputs) // (2 << 20)) * (2 << 20) \
        - (64 << 20)
    scratch = 128 << 20
    # the arena is most of the GPU's memory: memory a process that just
    # ended still holds can take a moment to come back, so a failed
    # allocation is retried a few times
    for attempt in range(5):
        try:
            world = osc_calls.CallWorld(npes=npes, max_reduce=4194304,
                                        move_bytes=128 << 20,
                                        dest_bytes=arena + scratch,
                                        keys=["f32", "f64", "i64"], seed=seed)
            break
        except osc.OscError as e:
            if e.code != -5 or attempt == 4:  # OSC_ERR_NOMEM
                raise
            log("served: arena allocation failed, retrying in 2 s")
            time.sleep(2.0)
    try:
        rng = np.random.default_rng(seed)
        teams = osc_calls.served_teams(rng, world, 16, 8, 128)
        reqs = osc_calls.served_mix(rng, world, teams, nreq, rate)

        def cfg(mode, digests, order=osc.ORDER_SIZE):
            c = osc.ServeCfg()
            c.mode, c.max_inflight, c.nstreams, c.max_group = mode, 256, 4, 64
            c.order, c.age_bytes_per_s = order, AGE_BYTES_PER_S
            c.dest_base, c.dest_bytes = world.dest_off, arena
            c.page_bytes = 256 << 10
            c.budget_bytes = budget_total - npes * inputs
            c.digests = digests
            return c

        def summary(st):
            return {"p50_ms": round(st.p50_ms, 3), "p90_ms": round(st.p90_ms, 3),
                    "p99_ms": round(st.p99_ms, 3), "max_ms": round(st.max_ms, 3),
                    "mean_ms": round(st.mean_ms, 3),
                    "wall_s": round(st.wall_s, 4), "GBps": round(st.gbps, 1),
                    "groups": st.groups,
                    "max_inflight": st.max_inflight_seen,
                    "peak_inflight_GB": round(st.peak_inflight_bytes / 1e9, 2),
                    "completed": st.completed, "rejected": st.rejected}

        osc.serve(reqs[:64], cfg(osc.SERVE_BATCHED, 0))  # warm-up
        _, st_b = osc.serve(reqs, cfg(osc.SERVE_BATCHED, 0))
        _, st_a = osc.serve(reqs, cfg(osc.SERVE_BATCHED, 0, osc.ORDER_ARRIVAL))
        _, st_p = osc.serve(reqs, cfg(osc.SERVE_PER_CALL, 0, osc.ORDER_ARRIVAL))
        # verification: same stream with digests, then sampled requests alone
        out, _ = osc.serve(reqs, cfg(osc.SERVE_BATCHED, 1))
        pick = rng.choice(len(out), size=min(verify, len(out)), replace=False)
        same = all(osc.request_alone(out[i].call, world.dest_off + arena) ==
                   out[i].digest for i in pick)
        # FP32 sum error vs the exact sum on sampled columns
        h, base0 = world.h, world.h.L.osc_heap_base()
        max_err = 0.0
        f32_sums = [r for r in out if r.call.kind == osc.CALL_REDUCE and
                    r.call.dtype == osc.DTYPES["f32"] and
                    r.call.op == osc.OPS["sum"] and
                    r.call.count <= 262144][:8]
        for r in f32_sums:
            c = r.call
            osc.request_alone(c, world.dest_off + arena)
            cols = rng.choice(c.count, size=min(512, c.count), replace=False)
            pes = [c.team_start + k for k in range(c.team_npes)]
            src = np.stack([h.get(base0 + c.src_off, c.count, np.float32,
                                  p)[cols] for p in pes])
            got = h.get(base0 + world.dest_off + arena, c.count, np.float32,
                        pes[0])[cols]
            max_err = max(max_err, max_rel_err_exact(got, src, "sum"))
        return {"requests": nreq, "rate_per_s": rate, "teams": 16,
                "world_pes": npes,
                "hbm_budget_GB": budget_total / 1e9,
                "admission_budget_GB": round((budget_total - npes * inputs)
                                             / 1e9, 1),
                "offered_GB": round(st_b.algorithmic_bytes / 1e9, 1),
                "order": "size-aware, %.3g B/s aging" % AGE_BYTES_PER_S,
                "served": summary(st_b), "per_call": summary(st_p),
                "served_arrival_order": summary(st_a),
                "p50_speedup": round(st_p.p50_ms / st_b.p50_ms, 3),
                "p99_speedup": round(st_p.p99_ms / st_b.p99_ms, 3),
                "GBps_speedup": round(st_b.gbps / st_p.gbps, 3),
                "same_alone_and_in_mix": bool(same),
                "verified_requests": int(len(pick)),
                "max_fp32_sum_rel_err_vs_exact": max_err}
    finally:
        world.close()


def matrix_heap(L, only=None):
    """The workload's heap: 64 PEs, one source array per (type,
    distribution) of the matrix, filled on the device, one destination of
    4M 8-byte elements and a 16-byte sink.  only: fill just that cell's
    source (the layout stays the same)."""
    dists = sorted({(dt, dist, a, b) for _, dt, dist, a, b in CELLS})
    per_pe = sum(N * ESIZE[d[0]] for d in dists) + N * 8 + (8 << 20)
    h = osc.Heap(npes=P, heap_bytes=per_pe)
    src = {}
    for k, (dt, dist, a, b) in enumerate(dists):
        src[(dt, dist, a, b)] = h.malloc(N * ESIZE[dt])
        if only is None or only == (dt, dist, a, b):
            h.fill(None, src[(dt, dist, a, b)], N, dt, dist, a, b, seed=k + 1)
    dest = h.malloc(N * 8)
    sink = h.malloc(16)
    h.sync()
    return h, src, dest, sink


def traffic_probe(L, cell, mode):
    """Child process of measured_traffic (run under rocprofv3 --pmc): ten
    warm-up calls of CELLS[cell] (past the replicate layout timing), three
    1 GiB copies (the counters' calibration), three calls of the cell."""
    op, dt, dist, a, b = CELLS[cell]
    h, src, dest, _ = matrix_heap(L, only=(dt, dist, a, b))
    osc.check(L.osc_set_fp_mode(mode), "fp mode")
    st = L.osc_default_stream()

    def call():
        osc.check(L.osc_reduce(h.world, osc.OPS[op], osc.DTYPES[dt], dest,
                               src[(dt, dist, a, b)], N, st), "osc_reduce")
    for _ in range(9):  # past the replicate layout timing
        call()
    h.sync()
    call()
    h.sync()
    base = L.osc_heap_base()
    half = (P // 2) * L.osc_slice_bytes()
    for _ in range(3):
        osc.check(L.osc_device_copy(base + half, base, 1 << 30, st), "copy")
    h.sync()
    for _ in range(3):
        call()
    h.sync()
    h.close()


def ceilings(L, h, sink, stream, reps=5):
    """Plain-copy, read-only and write-only rates on the workload's arena
    (after the timed region: they overwrite slices 32..63), 2 GiB each way
    like one 8-byte cell, and the bound they set for a kernel that reads
    and writes equal bytes in two phases (read phase, then write phase)."""
    base = L.osc_heap_base()
    half = (P // 2) * L.osc_slice_bytes()
    nb = 1 << 31
    t = osc.Timer(stream)

    def rate(fn, nbytes):
        fn()
        h.sync()
        t.start()
        for _ in range(reps):
            fn()
        t.stop()
        return reps * nbytes / (t.ms() / 1e3) / 1e9
    # 1 GiB read + 1 GiB written, as in the earlier measurements
    copy = rate(lambda: L.osc_device_copy(base + half, base, nb // 2, stream),
                nb)
    copy2 = rate(lambda: L.osc_device_copy(base + half, base, nb, stream),
                 2 * nb)
    rd = rate(lambda: L.osc_device_xor_read(base, nb, sink, stream), nb)
    wr = rate(lambda: L.osc_device_fill64(base + half, nb, 0x0123456789abcdef,
                                          stream), nb)
    t.close()
    return {"copy_1GiB_GBps": round(copy, 1), "copy_2GiB_GBps": round(copy2, 1),
            "read_2GiB_GBps": round(rd, 1), "write_2GiB_GBps": round(wr, 1),
            "read_then_write_GBps": round(2.0 / (1.0 / rd + 1.0 / wr), 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--fp-mode", choices=["accurate", "reference"],
                    default="accurate")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-movement", action="store_true")
    ap.add_argument("--cpu-baseline", action="store_true",
                    help="only the CPU baseline (run in a child process)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter passes")
    ap.add_argument("--traffic-probe", type=int, default=None,
                    help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline:
        print(json.dumps(cpu_baseline()))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the emulated PEs all live in one "
              "MI355X's HBM; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    L = osc.lib()
    mode = 0 if args.fp_mode == "accurate" else 1
    if args.traffic_probe is not None:
        traffic_probe(L, args.traffic_probe, mode)
        return
    log("matrix: heap and inputs")
    h, src, dest, sink = matrix_heap(L)
    osc.check(L.osc_set_fp_mode(mode), "fp mode")
    stream = L.osc_default_stream()
    evs = []
    for _ in range(len(CELLS) + 1):
        e = C.c_void_p()
        osc.check(L.osc_event_create(C.byref(e)), "event")
        evs.append(e)

    def step(record):
        """One pass over the 15 cells.  record "step": an event before the
        first call and after the last; "cells": an event before every
        call and after the last; None: no events."""
        for i, (op, dt, dist, a, b) in enumerate(CELLS):
            if record == "cells" or (record == "step" and i == 0):
                L.osc_event_record(evs[i], stream)
            rc = L.osc_reduce(h.world, osc.OPS[op], osc.DTYPES[dt], dest,
                              src[(dt, dist, a, b)], N, stream)
            osc.check(rc, "osc_reduce")
        if record:
            L.osc_event_record(evs[-1], stream)

    # warm-up (the first calls of each size class also time the replicate
    # layouts: at least 9 calls and a completed stream), then the kernels
    # each cell launches from then on
    for _ in range(max(args.warmup, 3)):
        step(None)
    h.sync()
    cell_kernels = []
    for i, (op, dt, dist, a, b) in enumerate(CELLS):
        osc.check(L.osc_reduce(h.world, osc.OPS[op], osc.DTYPES[dt], dest,
                               src[(dt, dist, a, b)], N, stream), "osc_reduce")
        cell_kernels.append(last_kernels(L))
    h.sync()
    # the timed steps: an event at each step's start and end only (an
    # event between two calls leaves the GPU idle for a few microseconds)
    log("matrix: %d timed steps" % args.steps)
    step_s = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step("step")
        h.sync()
        ms = C.c_float()
        L.osc_event_elapsed_ms(evs[0], evs[-1], C.byref(ms))
        step_s += ms.value / 1e3
    wall = time.perf_counter() - t0
    # the same steps again with an event before every call: the per-cell
    # times (and the roofline cell)
    log("matrix: %d steps timed per cell" % args.steps)
    cell_ms = [0.0] * len(CELLS)
    for _ in range(args.steps):
        step("cells")
        h.sync()
        ms = C.c_float()
        for i in range(len(CELLS)):
            L.osc_event_elapsed_ms(evs[i], evs[i + 1], C.byref(ms))
            cell_ms[i] += ms.value
    kern_s = sum(cell_ms) / 1e3

    step_bytes = sum(2.0 * P * N * ESIZE[dt] for _, dt, _, _, _ in CELLS)
    value = step_bytes * args.steps / step_s / 1e9
    cells = []
    for i, (op, dt, _, _, _) in enumerate(CELLS):
        avg = cell_ms[i] / args.steps
        b = 2.0 * P * N * ESIZE[dt]
        cells.append({"op": op, "type": dt, "ms": round(avg, 4),
                      "GBps": round(b / (avg / 1e3) / 1e9, 1),
                      "kernels": cell_kernels[i]})
    dom = max(range(len(CELLS)), key=lambda i: cell_ms[i])
    dom_avg = cell_ms[dom] / args.steps / 1e3
    dom_b = 2.0 * P * N * ESIZE[CELLS[dom][1]]
    achieved = dom_b / dom_avg / 1e9

    log("matrix: accuracy")
    acc = accuracy(L, h, src, dest)
    osc.check(L.osc_set_fp_mode(mode), "fp mode")

    log("ceilings")
    ceil = ceilings(L, h, sink, stream)
    traffic, traffic_note = None, None
    if not args.no_traffic:
        log("traffic: rocprofv3 counter passes")
        try:
            traffic, traffic_note = measured_traffic(dom, args.fp_mode)
        except (RuntimeError, OSError, ValueError, KeyError) as e:
            traffic_note = "not measured: %s" % e

    out = {
        "metric": "reduce HBM GB/s (% peak) at 64 PEs x 4M; p50/p99 latency "
                  "served vs per-call",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3 / args.steps, 4),
        "wall_ms_per_step": round(wall * 1e3 / args.steps, 4),
        "value_with_cell_events": round(step_bytes * args.steps / kern_s
                                        / 1e9, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "pct_peak": round(100.0 * value / PEAK_HBM_GBPS, 2),
        "dtype": "u64/i64/f32/f64",
        "data": "synthetic (seeded SplitMix64 on device)",
        "config": {"workload": "team reduce op/type matrix, 15 cells, "
                               "SHMEM_TEAM_WORLD of 64 PEs x 4194304 "
                               "elements per PE",
                   "npes": P, "nreduce": N, "fp_mode": args.fp_mode,
                   "parallelism": "1 GPU, 64 emulated PEs"},
        "roofline": {"bound": "hbm",
                     "kernel": " + ".join(cell_kernels[dom]),
                     "cell": "%s %s" % (CELLS[dom][0], CELLS[dom][1]),
                     "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBPS, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "traffic_ratio": None if traffic is None else
                     round(traffic / dom_b, 4),
                     "traffic_detail": traffic_note,
                     "algorithmic_bytes_per_launch": dom_b,
                     "avg_launch_us": round(dom_avg * 1e6, 1),
                     "launch": "one reduce call (its kernels back to back on "
                               "the library stream), HIP events around it"},
        "copy_ceiling_GBps": ceil["copy_1GiB_GBps"],
        "ceilings": ceil,
        "frac_of_read_then_write_bound": round(
            value / ceil["read_then_write_GBps"], 4),
        "cells": cells,
        "accuracy": acc,
    }
    h.close()
    log("movement")
    out["movement"] = None if args.no_movement else bench_movement()
    log("batched")
    out["batched"] = None if args.no_batched else bench_batched()
    log("served")
    out["served"] = None if args.no_served else bench_served()
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline_subprocess()
    else:
        out["cpu_baseline"] = None
    log("done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()

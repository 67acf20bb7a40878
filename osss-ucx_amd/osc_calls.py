"""Synthetic call mixes for the batched and served modes (used by bench.py,
the autotuner and the tests).

The batched mix: team sizes uniform in {2..128}; reduce nreduce
log-uniform in [64, 1,048,576] with (op, type) uniform over the 15 cells of
the reduce matrix (and/or/xor u64; min/max/sum/prod i64; sum/min/max/prod
f32 and f64); 20% data movement (half alltoall, half fcollect) with block
sizes log-uniform in [256 B, 256 KiB].  Sources are read-only input areas
shared by all calls of a type; destinations are placed by osc_batch_plan.
"""
from __future__ import annotations

import numpy as np

import osc

MATRIX = [("and", "u64"), ("or", "u64"), ("xor", "u64"),
          ("min", "i64"), ("max", "i64"), ("sum", "i64"), ("prod", "i64"),
          ("sum", "f32"), ("min", "f32"), ("max", "f32"), ("prod", "f32"),
          ("sum", "f64"), ("min", "f64"), ("max", "f64"), ("prod", "f64")]
# input distribution per (type, op) as in the reduce-matrix configuration
DIST = {"u64": ("bits", 0.0, 0.0),
        "i64": ("int_range", -2.0 ** 40, 2.0 ** 40),
        "i64prod": ("set_m1_1_2", 0.0, 0.0),
        "f32": ("normal", 0.0, 1.0), "f32prod": ("real_range", 0.999, 1.001),
        "f64": ("normal", 0.0, 1.0), "f64prod": ("real_range", 0.999, 1.001)}
ESZ = {"u64": 8, "i64": 8, "f32": 4, "f64": 8}


def dist_key(op, dt):
    return dt + "prod" if (op == "prod" and dt != "u64") else dt


class CallWorld:
    """A library world with shared read-only inputs and a destination area.

    max_reduce: largest nreduce; move_bytes: per-PE bytes of the
    data-movement input area; dest_bytes: per-PE destination area."""

    def __init__(self, npes, max_reduce, move_bytes, dest_bytes, seed=1,
                 keys=None):
        self.npes = npes
        self.max_reduce = max_reduce
        keys = sorted(DIST) if keys is None else sorted(keys)
        self.in_bytes = {k: max_reduce * ESZ[k[:3]] for k in keys}
        heap = sum(self.in_bytes.values()) + move_bytes + dest_bytes + (1 << 20)
        self.h = osc.Heap(npes=npes, heap_bytes=heap)
        base0 = self.h.L.osc_heap_base()
        self.off = {}
        for k in keys:
            p = self.h.malloc(self.in_bytes[k])
            dist, a, b = DIST[k]
            self.h.fill(None, p, max_reduce, k[:3], dist, a, b, seed=seed)
            self.off[k] = p - base0
            seed += 1
        p = self.h.malloc(move_bytes)
        self.h.fill(None, p, move_bytes // 8, "u64", "bits", seed=seed)
        self.move_off = p - base0
        self.move_bytes = move_bytes
        p = self.h.malloc(dest_bytes)
        self.dest_off = p - base0
        self.dest_bytes = dest_bytes
        self.h.sync()

    def close(self):
        self.h.close()


def reduce_call(world, op, dt, P, start, n):
    c = osc.Call()
    c.kind = osc.CALL_REDUCE
    c.op, c.dtype = osc.OPS[op], osc.DTYPES[dt]
    c.team_start, c.team_stride, c.team_npes = start, 1, P
    c.count = n
    c.src_off = world.off[dist_key(op, dt)]
    return c


def move_call(world, kind, P, start, block, root=0):
    c = osc.Call()
    c.kind = kind
    c.team_start, c.team_stride, c.team_npes = start, 1, P
    c.count = block
    c.root = root
    c.src_off = world.move_off
    return c


def log_uniform_int(rng, lo, hi):
    return int(np.exp(rng.uniform(np.log(lo), np.log(hi + 1))))


def batched_mix(rng, world, ncalls, pmin=2, pmax=128, nmin=64,
                nmax=1048576, bmin=256, bmax=262144, move_frac=0.2):
    """The batched configuration's call mix, in a world of >= pmax PEs."""
    calls = []
    pmax = min(pmax, world.npes)
    for _ in range(ncalls):
        P = int(rng.integers(pmin, pmax + 1))
        start = int(rng.integers(0, world.npes - P + 1))
        if rng.random() < move_frac:
            kind = osc.CALL_ALLTOALL if rng.random() < 0.5 else \
                osc.CALL_FCOLLECT
            block = log_uniform_int(rng, bmin, bmax) // 16 * 16 or 16
            if kind == osc.CALL_ALLTOALL:
                block = min(block, world.move_bytes // P // 16 * 16)
            calls.append(move_call(world, kind, P, start, block))
        else:
            op, dt = MATRIX[int(rng.integers(len(MATRIX)))]
            n = min(log_uniform_int(rng, nmin, nmax), world.max_reduce)
            calls.append(reduce_call(world, op, dt, P, start, n))
    return calls


# ---- the served configuration ---------------------------------------------
SERVED_REDUCE = [("sum", "f32"), ("max", "f32"), ("sum", "f64"),
                 ("max", "f64"), ("sum", "i64"), ("max", "i64")]


def served_teams(rng, world, nteams=16, pmin=8, pmax=128):
    """nteams contiguous teams of uniform size in [pmin, pmax] at random
    positions of the world (teams may overlap, as split teams can)."""
    teams = []
    for _ in range(nteams):
        P = int(rng.integers(pmin, min(pmax, world.npes) + 1))
        teams.append((int(rng.integers(0, world.npes - P + 1)), P))
    return teams


def served_mix(rng, world, teams, nreq, rate_hz, nmin=256, nmax=4194304,
               bmin=1024, bmax=1048576):
    """Poisson arrivals at rate_hz from uniformly chosen teams: 60% reduce
    (f32/f64/i64 sum or max, nreduce log-uniform [nmin, nmax]), 20%
    fcollect, 20% alltoall (block log-uniform [bmin, bmax] bytes, a
    multiple of 16)."""
    reqs = []
    t = 0.0
    for _ in range(nreq):
        t += rng.exponential(1.0 / rate_hz)
        start, P = teams[int(rng.integers(len(teams)))]
        u = rng.random()
        if u < 0.6:
            op, dt = SERVED_REDUCE[int(rng.integers(len(SERVED_REDUCE)))]
            n = min(log_uniform_int(rng, nmin, nmax), world.max_reduce)
            c = reduce_call(world, op, dt, P, start, n)
        else:
            kind = osc.CALL_FCOLLECT if u < 0.8 else osc.CALL_ALLTOALL
            block = max(16, log_uniform_int(rng, bmin, bmax) // 16 * 16)
            if kind == osc.CALL_ALLTOALL:
                block = min(block, world.move_bytes // P // 16 * 16)
            c = move_call(world, kind, P, start, block)
        r = osc.Request()
        r.call = c
        r.arrival_s = t
        reqs.append(r)
    return reqs

"""The benchmark configurations at their own size, checked against the
oracle.

* Reduce matrix: all 15 (op, type) cells at 64 PEs x 4,194,304 elements
  per PE, inputs generated on the device, every member's result read on
  eight windows of 256 columns (the first and last columns of the call,
  span boundaries of every kernel shape, and random places).  Integer,
  bitwise and min/max cells are bit-exact against the oracle's recursive
  doubling (src/shcoll/src/reduce.c:182-293); float sums and products are
  within one rounding of the exact value in the accurate mode (f32 bound
  2^-24 < 1e-6) and bit-exact against the recursive doubling in the
  reference's order.
* Batched: the 4096-call mix on a 128-PE world (teams 2..128, nreduce
  64..1M, 20% alltoall/fcollect with 256 B..256 KiB blocks), wave by wave:
  batched digests equal the one-by-one path's, and calls of every kind on
  teams of 65..128 PEs equal the oracle.
* Served: 16 teams of 8..128 PEs on a 128-PE world, Poisson requests at
  20k req/s with nreduce up to 4M and blocks up to 1 MiB: results are the
  same alone and in the mix, in both admission orders and in the per-call
  baseline, and f32 sums are within 1e-6 of the exact sum.
"""
import zlib

import numpy as np
import pytest

import osc
import osc_calls
import oracle as O
from gpu_util import bit_equal, max_rel_err

pytestmark = pytest.mark.gpu

P, N = 64, 4194304
WIN = 256
TYPENAME = {"u64": "uint64", "i64": "int64", "f32": "float", "f64": "double"}
# one rounding of the exact value, plus the accumulator's own error (FP64
# accumulation for f32; Sum2 / double-double for f64), far below 1e-6
TOL = {"f32": 2.0 ** -24 * 1.001, "f64": 2.0 ** -53 * 1.01}


def windows(rng):
    """Start columns of 8 windows: both ends, the span boundaries of the
    8-wide (2048 units) and narrow shapes, and random places."""
    fixed = [0, N - WIN, 4096 - WIN // 2, 8192 * 3 - 8, 2048 * 7 - 3]
    rand = [int(x) for x in rng.integers(WIN, N - 2 * WIN, 3)]
    return sorted(set(fixed + rand))


def read_windows(h, sym, npt, starts, pes):
    es = np.dtype(npt).itemsize
    return np.stack([np.concatenate([h.get(sym + s * es, WIN, npt, pe)
                                     for s in starts]) for pe in pes])


@pytest.fixture(scope="module")
def matrix_heap():
    h = osc.Heap(npes=P, heap_bytes=2 * N * 8 + (4 << 20))
    yield h
    h.close()


@pytest.mark.parametrize("op,dt", osc_calls.MATRIX)
def test_reduce_matrix_cell_full_size(matrix_heap, op, dt):
    h = matrix_heap
    rng = np.random.default_rng(zlib.crc32((op + dt).encode()))
    npt = O.NP_DTYPE[dt]
    s, d = h.malloc(N * 8), h.malloc(N * 8)
    try:
        dist, a, b = osc_calls.DIST[osc_calls.dist_key(op, dt)]
        h.fill(None, s, N, dt, dist, a, b,
               seed=osc_calls.MATRIX.index((op, dt)) + 1)
        starts = windows(rng)
        src = read_windows(h, s, npt, starts, range(P))
        fn = osc.typed(f"shmem_{TYPENAME[dt]}_{op}_reduce")
        modes = [osc.FP_ACCURATE]
        if dt[0] == "f" and op in ("sum", "prod"):
            modes.append(osc.FP_REFERENCE)
        for mode in modes:
            osc.check(h.L.osc_set_fp_mode(mode), "fp mode")
            h.L.osc_device_copy(d, s, N * 8, None)   # stale bytes first
            assert fn(h.world, d, s, N) == 0
            got = read_windows(h, d, npt, starts, range(P))
            # every member holds the same result
            assert bit_equal(got, np.broadcast_to(got[:1], got.shape), dt)
            exact_order_free = dt[0] in "ui" or op in ("min", "max")
            if exact_order_free or mode == osc.FP_REFERENCE:
                assert bit_equal(got, O.rec_dbl(op, dt, src), dt), mode
            else:
                assert max_rel_err(got[:2], src, op, dt) <= TOL[dt]
    finally:
        h.L.osc_set_fp_mode(0)
        h.free(s)
        h.free(d)


# ---- batched: the 4096-call mix on 128 PEs ---------------------------------

def _check_call(h, c):
    """One call's result (already run) against the oracle."""
    base0 = h.L.osc_heap_base()
    pes = [c.team_start + k * c.team_stride for k in range(c.team_npes)]
    Pc, m = c.team_npes, c.count
    if c.kind == osc.CALL_REDUCE:
        dt = {v: k for k, v in osc.DTYPES.items()}[c.dtype]
        op = {v: k for k, v in osc.OPS.items()}[c.op]
        npt = O.NP_DTYPE[dt]
        src = np.stack([h.get(base0 + c.src_off, m, npt, p) for p in pes])
        got = np.stack([h.get(base0 + c.dest_off, m, npt, p) for p in pes])
        if dt[0] in "ui" or op in ("min", "max"):
            assert bit_equal(got, O.rec_dbl(op, dt, src), dt), (dt, op, Pc)
        else:
            assert max_rel_err(got, src, op, dt) <= TOL[dt], (dt, op, Pc)
    else:
        slen = Pc * m if c.kind == osc.CALL_ALLTOALL else m
        src = np.stack([h.get(base0 + c.src_off, slen, np.uint8, p)
                        for p in pes])
        got = np.stack([h.get(base0 + c.dest_off, Pc * m, np.uint8, p)
                        for p in pes])
        want = O.alltoall(src) if c.kind == osc.CALL_ALLTOALL else \
            O.fcollect(src)
        assert np.array_equal(got, want), (c.kind, Pc, m)


def test_batched_4096_calls_on_128_pes():
    world = osc_calls.CallWorld(npes=128, max_reduce=1048576,
                                move_bytes=32 << 20, dest_bytes=1 << 30)
    try:
        rng = np.random.default_rng(4)
        calls = osc_calls.batched_mix(rng, world, 4096)
        assert max(c.team_npes for c in calls) > 120
        calls, nwaves = osc.plan(calls, world.dest_off, world.dest_bytes)
        checked = {"reduce": 0, "move": 0}
        for w in range(nwaves):
            wave = [c for c in calls if c.wave == w]
            b = osc.Batch(wave)
            b.launch_serial()
            d_serial = b.digests()
            # stale bytes in the destinations, then the captured graph
            world.h.L.osc_device_copy(world.h.L.osc_heap_base() +
                                      world.dest_off,
                                      world.h.L.osc_heap_base(),
                                      256 << 20, None)
            b.launch()
            d_batch = b.digests()
            assert np.array_equal(d_batch, d_serial), w
            world.h.sync()
            # oracle: a few calls of this wave on teams of 65..128 PEs
            red = [c for c in wave if c.team_npes >= 65 and
                   c.kind == osc.CALL_REDUCE and
                   c.count * c.team_npes <= (1 << 22)]
            mov = [c for c in wave if c.team_npes >= 65 and
                   c.kind != osc.CALL_REDUCE and
                   c.count * c.team_npes ** 2 <= (1 << 25)]
            for c in red[:4] + mov[:2]:
                _check_call(world.h, c)
                checked["reduce" if c.kind == osc.CALL_REDUCE else
                        "move"] += 1
            b.close()
        assert checked["reduce"] >= 10 and checked["move"] >= 2, checked
    finally:
        world.close()


@pytest.mark.parametrize("phases", [1, 0])
def test_batch_phases_split_reductions_and_movers(phases):
    """Waves of split reductions (1-4 MiB per member on 32-64 members, whole
    and partial last units, float min/max among them) and large fcollect /
    alltoall calls: with the wave in two phases (walks, then replicates with
    the data movement) or in one, the captured graph gives the one-by-one
    digests on every replay, with every member's destinations poisoned
    before each replay."""
    world = osc_calls.CallWorld(npes=64, max_reduce=1 << 19,
                                move_bytes=8 << 20, dest_bytes=256 << 20)
    h = world.h
    osc.check(h.L.osc_set_batch_phases(phases), "phases")
    try:
        rng = np.random.default_rng(77 + phases)
        calls = []
        for k in range(24):
            op, dt = osc_calls.MATRIX[k % len(osc_calls.MATRIX)]
            P = int(rng.integers(32, 65))
            n = int(rng.integers(1 << 17, 1 << 19)) | (k & 1)
            calls.append(osc_calls.reduce_call(world, op, dt, P,
                                               int(rng.integers(0, 65 - P)),
                                               n))
        for k in range(6):
            P = int(rng.integers(16, 65))
            kind = osc.CALL_ALLTOALL if k % 2 else osc.CALL_FCOLLECT
            blk = min(96 << 10, (8 << 20) // P // 16 * 16)
            calls.append(osc_calls.move_call(world, kind, P,
                                             int(rng.integers(0, 65 - P)),
                                             blk))
        calls, nw = osc.plan(calls, world.dest_off, world.dest_bytes)
        dest = h.L.osc_heap_base() + world.dest_off
        for wv in range(nw):
            wave = [c for c in calls if c.wave == wv]
            b = osc.Batch(wave)
            b.launch_serial()
            d_serial = b.digests()
            for rep in range(2):
                h.fill(None, dest, world.dest_bytes // 8, "u64", "bits",
                       seed=300 + rep)
                b.launch()
                assert np.array_equal(b.digests(), d_serial), (wv, rep)
            b.close()
    finally:
        h.L.osc_set_batch_phases(1)
        world.close()
    assert h.L.osc_set_batch_phases(2) != 0


# ---- served: 16 teams of 8..128 PEs ------------------------------------------

def test_served_16_teams_of_8_to_128_pes():
    npes, arena, scratch = 128, 384 << 20, 128 << 20
    world = osc_calls.CallWorld(npes=npes, max_reduce=4194304,
                                move_bytes=128 << 20,
                                dest_bytes=arena + scratch,
                                keys=["f32", "f64", "i64"], seed=5)
    try:
        rng = np.random.default_rng(5)
        teams = osc_calls.served_teams(rng, world, 16, 8, 128)
        reqs = osc_calls.served_mix(rng, world, teams, 400, 20000.0)
        assert max(r.call.count for r in reqs
                   if r.call.kind == osc.CALL_REDUCE) > (1 << 20)

        def cfg(mode, order):
            c = osc.ServeCfg()
            c.mode, c.max_inflight, c.nstreams, c.max_group = mode, 256, 4, 64
            c.order, c.age_bytes_per_s = order, 5e11
            c.dest_base, c.dest_bytes = world.dest_off, arena
            c.page_bytes = 256 << 10
            c.budget_bytes = float(npes * arena)
            c.digests = 1
            return c

        out, st = osc.serve(reqs, cfg(osc.SERVE_BATCHED, osc.ORDER_SIZE))
        assert st.completed == len(reqs) and st.rejected == 0
        assert st.groups < len(reqs) and st.max_inflight_seen > 1
        digests = [r.digest for r in out]
        for mode, order in ((osc.SERVE_BATCHED, osc.ORDER_ARRIVAL),
                            (osc.SERVE_PER_CALL, osc.ORDER_ARRIVAL)):
            o2, s2 = osc.serve(reqs, cfg(mode, order))
            assert s2.completed == len(reqs)
            assert [r.digest for r in o2] == digests, (mode, order)
        pick = rng.choice(len(out), size=64, replace=False)
        for i in pick:
            assert osc.request_alone(out[i].call, world.dest_off + arena) \
                == out[i].digest, i
        # f32 sums: within 1e-6 of the exact sum on sampled columns
        h, base0 = world.h, world.h.L.osc_heap_base()
        f32 = [r.call for r in out if r.call.kind == osc.CALL_REDUCE and
               r.call.dtype == osc.DTYPES["f32"] and
               r.call.op == osc.OPS["sum"]][:6]
        assert f32
        for c in f32:
            osc.request_alone(c, world.dest_off + arena)
            w = min(256, c.count)               # one window of columns
            lo = int(rng.integers(0, c.count - w + 1))
            pes = [c.team_start + k for k in range(c.team_npes)]
            src = np.stack([h.get(base0 + c.src_off + 4 * lo, w, np.float32,
                                  p) for p in pes])
            got = np.stack([h.get(base0 + world.dest_off + arena + 4 * lo, w,
                                  np.float32, p) for p in (pes[0], pes[-1])])
            assert max_rel_err(got, src, "sum", "f32") <= TOL["f32"]
    finally:
        world.close()

"""Batched execution on the GPU: thousands of uneven calls in one captured
graph must give bit-identical results to the same calls issued one by one
(per-call digests of every member's destination), replay identically, and
match the oracle on spot-checked calls."""
import numpy as np
import pytest

import osc
import oracle as O
from gpu_util import TYPENAME, bit_equal, max_rel_err

pytestmark = pytest.mark.gpu

WORLD = 32
SRC_REGION = 4 << 20     # per-PE inputs
DEST_REGION = 12 << 20   # per-PE destinations (planner packs into waves)
RED_TYPES = [("u64", "and"), ("u64", "xor"), ("i64", "min"), ("i64", "max"),
             ("i64", "sum"), ("i64", "prod"), ("f32", "sum"), ("f32", "min"),
             ("f32", "max"), ("f32", "prod"), ("f64", "sum"), ("f64", "max"),
             ("f64", "prod"), ("i32", "sum"), ("u8", "xor"), ("i16", "max"),
             ("c32", "sum"), ("c64", "prod")]
ESZ = {"u8": 1, "i16": 2, "i32": 4, "u64": 8, "i64": 8, "f32": 4, "f64": 8,
       "c32": 8, "c64": 16}
DIST = {"u64": ("bits", 0, 0), "i64": ("int_range", -2.0 ** 40, 2.0 ** 40),
        "f32": ("normal", 0.0, 1.0), "f64": ("normal", 0.0, 1.0),
        "i32": ("bits", 0, 0), "u8": ("bits", 0, 0), "i16": ("bits", 0, 0),
        "c32": ("normal", 0.0, 1.0), "c64": ("real_range", 0.999, 1.001)}


class World:
    def __init__(self):
        self.h = osc.Heap(npes=WORLD, heap_bytes=SRC_REGION + DEST_REGION)
        self.src_base = self.h.malloc(SRC_REGION)
        self.dest_base = self.h.malloc(DEST_REGION)
        base0 = self.h.L.osc_heap_base()
        self.src_off = self.src_base - base0
        self.dest_off = self.dest_base - base0
        # per-type input areas of 256 KiB, data-movement bytes after them
        self.type_off = {}
        for k, dt in enumerate(ESZ):
            off = self.src_base + k * (256 << 10)
            self.type_off[dt] = off - base0
            dist, a, b = DIST[dt]
            self.h.fill(None, off, (256 << 10) // ESZ[dt], dt, dist, a, b,
                        seed=k + 11)
        self.move_off = self.src_off + (3 << 20)
        self.h.fill(None, self.src_base + (3 << 20), 1 << 17, "u64", "bits",
                    seed=99)
        self.h.sync()


_W = []


@pytest.fixture
def world():
    if not _W or not _W[0].h.alive:
        _W[:] = [World()]
    return _W[0]


def make_calls(rng, w, n):
    calls = []
    for _ in range(n):
        c = osc.Call()
        r = rng.random()
        P = int(rng.integers(1, WORLD + 1))
        stride = 2 if (P <= WORLD // 2 and rng.random() < 0.2) else 1
        c.team_npes, c.team_stride = P, stride
        c.team_start = int(rng.integers(0, WORLD - (P - 1) * stride))
        if r < 0.6:
            dt, op = RED_TYPES[int(rng.integers(len(RED_TYPES)))]
            c.kind = osc.CALL_REDUCE
            c.dtype, c.op = osc.DTYPES[dt], osc.OPS[op]
            c.count = int(np.exp(rng.uniform(0, np.log((256 << 10) // ESZ[dt]))))
            c.src_off = w.type_off[dt]
        else:
            c.kind = [osc.CALL_FCOLLECT, osc.CALL_ALLTOALL,
                      osc.CALL_BROADCAST][int(rng.integers(3))]
            room = (1 << 20) - 64  # the movement-input area, minus the skew
            cap = room if c.kind == osc.CALL_BROADCAST else room // P
            c.count = max(1, int(np.exp(rng.uniform(0, np.log(max(2, cap))))))
            c.src_off = w.move_off + 16 * int(rng.integers(0, 4))
            c.root = int(rng.integers(P))
        calls.append(c)
    return calls


def pes_of(c):
    return [c.team_start + k * c.team_stride for k in range(c.team_npes)]


@pytest.mark.parametrize("mode", [0, 1])
def test_batched_equals_one_by_one(world, mode):
    h = world.h
    h.L.osc_set_fp_mode(mode)
    try:
        rng = np.random.default_rng(100 + mode)
        calls, nw = osc.plan(make_calls(rng, world, 600), world.dest_off,
                             DEST_REGION)
        assert nw >= 1
        for wv in range(nw):
            wave = [c for c in calls if c.wave == wv]
            b = osc.Batch(wave)
            b.launch_serial()
            d_serial = b.digests()
            # poison the destinations, then replay the graph three times
            h.L.osc_device_copy(world.dest_base, world.src_base,
                                DEST_REGION // 4, None)
            for rep in range(3):
                b.launch()
                d_batch = b.digests()
                assert np.array_equal(d_batch, d_serial), (wv, rep)
            inf = b.info()
            assert inf.ncalls == len(wave) and inf.nlaunches >= 1
            b.close()
    finally:
        h.L.osc_set_fp_mode(0)


def test_direct_rows_mixed_with_tiled_rows(world):
    """Table rows with tile 0 run their calls through the per-call kernel
    inside the batch (and its graph); the rest stay tiled.  Results must
    still equal the one-by-one path, on every replay."""
    import ctypes as C
    h = world.h
    L = h.L
    saved = {}
    for cls in range(5):
        for bkt in range(4):
            s, t, g = C.c_int(), C.c_int(), C.c_int()
            L.osc_tune_get(cls, bkt, C.byref(s), C.byref(t), C.byref(g))
            saved[(cls, bkt)] = (s.value, t.value, g.value)
    try:
        for cls in range(5):          # direct: the odd buckets of each class
            for bkt in (1, 3):
                s0, _, g0 = saved[(cls, bkt)]
                assert L.osc_tune_set(cls, bkt, s0, 0, g0) == 0
        rng = np.random.default_rng(31)
        calls, nw = osc.plan(make_calls(rng, world, 300), world.dest_off,
                             DEST_REGION)
        for wv in range(nw):
            wave = [c for c in calls if c.wave == wv]
            b = osc.Batch(wave)
            b.launch_serial()
            d_serial = b.digests()
            h.L.osc_device_copy(world.dest_base, world.src_base,
                                DEST_REGION // 4, None)
            for rep in range(2):
                b.launch()
                assert np.array_equal(b.digests(), d_serial), (wv, rep)
            b.close()
    finally:
        for (cls, bkt), row in saved.items():
            L.osc_tune_set(cls, bkt, *row)


def test_batched_results_match_oracle(world):
    h = world.h
    rng = np.random.default_rng(7)
    calls, nw = osc.plan(make_calls(rng, world, 200), world.dest_off,
                         DEST_REGION)
    wave = [c for c in calls if c.wave == 0]
    b = osc.Batch(wave)
    b.launch()
    h.sync()
    base0 = h.L.osc_heap_base()
    checked = 0
    for c in wave:
        pes = pes_of(c)
        if c.kind == osc.CALL_REDUCE and c.count <= 20000:
            dt = {v: k for k, v in osc.DTYPES.items()}[c.dtype]
            op = {v: k for k, v in osc.OPS.items()}[c.op]
            npt = O.NP_DTYPE[dt]
            src = np.stack([h.get(base0 + c.src_off, c.count, npt, p)
                            for p in pes])
            got = np.stack([h.get(base0 + c.dest_off, c.count, npt, p)
                            for p in pes])
            if dt[0] in "ui" or op in ("min", "max"):
                assert bit_equal(got, O.rec_dbl(op, dt, src), dt), (dt, op)
            else:
                tol = 1.2e-7 if dt in ("f32", "c32") else 2.3e-16
                assert max_rel_err(got, src, op, dt) <= tol, (dt, op)
            checked += 1
        elif c.kind != osc.CALL_REDUCE and c.count * c.team_npes <= 200000:
            P, m = c.team_npes, c.count
            slen = P * m if c.kind == osc.CALL_ALLTOALL else m
            dlen = m if c.kind == osc.CALL_BROADCAST else P * m
            src = np.stack([h.get(base0 + c.src_off, slen, np.uint8, p)
                            for p in pes])
            got = np.stack([h.get(base0 + c.dest_off, dlen, np.uint8, p)
                            for p in pes])
            want = {osc.CALL_FCOLLECT: lambda: O.fcollect(src),
                    osc.CALL_ALLTOALL: lambda: O.alltoall(src),
                    osc.CALL_BROADCAST: lambda: O.broadcast(src, c.root)}[
                        c.kind]()
            assert np.array_equal(got, want), c.kind
            checked += 1
    b.close()
    assert checked >= 20


def test_batch_rejects_bad_calls(world):
    c = osc.Call()
    c.kind, c.op, c.dtype = osc.CALL_REDUCE, osc.OPS["xor"], osc.DTYPES["f32"]
    c.team_start, c.team_stride, c.team_npes = 0, 1, 4
    c.count, c.src_off, c.dest_off = 16, world.type_off["f32"], world.dest_off
    with pytest.raises(osc.OscError):
        osc.Batch([c])
    c.op = osc.OPS["sum"]
    c.src_off += 4  # misaligned for the batched reduce
    with pytest.raises(osc.OscError):
        osc.Batch([c])


def test_batch_rejects_wrapping_team_and_count(world):
    """A stride whose last member wraps 32 bits and a count whose byte
    count wraps 64 bits are errors in batch creation and in a request run
    alone; nothing is launched for them."""
    c = osc.Call()
    c.kind, c.op, c.dtype = osc.CALL_FCOLLECT, 0, 0
    c.team_start, c.team_stride, c.team_npes = 0, 1 << 30, 3
    c.count, c.src_off, c.dest_off = 16, world.type_off["f32"], world.dest_off
    for run in (lambda: osc.Batch([c]), lambda: osc.request_alone(c, world.dest_off)):
        with pytest.raises(osc.OscError) as e:
            run()
        assert e.value.code == -3  # OSC_ERR_TEAM
    c.team_stride, c.team_npes, c.count = 1, 4, 1 << 62
    for run in (lambda: osc.Batch([c]), lambda: osc.request_alone(c, world.dest_off)):
        with pytest.raises(osc.OscError):
            run()


@pytest.mark.parametrize("branches", [1, 2, 8])
def test_parallel_branches_give_the_same_results(world, branches):
    """The captured graph spreads each wave's launches over parallel
    branches (4 by default); every branch count gives the digests of the
    one-by-one path, on every replay."""
    h = world.h
    osc.check(h.L.osc_set_batch_branches(branches), "branches")
    try:
        rng = np.random.default_rng(500 + branches)
        calls, nw = osc.plan(make_calls(rng, world, 400), world.dest_off,
                             DEST_REGION)
        for wv in range(nw):
            wave = [c for c in calls if c.wave == wv]
            b = osc.Batch(wave)
            b.launch_serial()
            d_serial = b.digests()
            h.L.osc_device_copy(world.dest_base, world.src_base,
                                DEST_REGION // 4, None)
            for rep in range(2):
                b.launch()
                assert np.array_equal(b.digests(), d_serial), (wv, rep)
            b.close()
    finally:
        h.L.osc_set_batch_branches(4)
    assert h.L.osc_set_batch_branches(0) != 0
    assert h.L.osc_set_batch_branches(9) != 0

"""Per-call records (OSC_LOG=1): every collective writes one JSON line with
its arguments and what it cost -- algorithmic bytes (the same formulas as
the benchmark's), elapsed microseconds between two HIP events around the
launch, GB/s and percent of the 8 TB/s peak -- to OSC_LOG_FILE."""
import ctypes as C
import json

import numpy as np
import pytest

import osc

pytestmark = pytest.mark.gpu


def test_records_of_each_collective(monkeypatch, tmp_path):
    log = tmp_path / "calls.jsonl"
    monkeypatch.setenv("OSC_LOG", "1")
    monkeypatch.setenv("OSC_LOG_FILE", str(log))
    P, n = 16, 100000
    with osc.Heap(npes=P, heap_bytes=32 << 20) as h:
        s, d = h.malloc(n * 8), h.malloc(P * n * 8)
        h.fill(None, s, n, "f32", "normal", 0.0, 1.0, seed=1)
        L = h.L
        assert osc.typed("shmem_float_sum_reduce")(h.world, d, s, n) == 0
        h.L.osc_set_fp_mode(osc.FP_LINEAR)
        assert osc.typed("shmem_long_max_reduce")(h.world, d, s, 1000) == 0
        h.L.osc_set_fp_mode(0)
        team = h.split(2, 3, 5)
        assert L.shmem_fcollectmem(team, d, s, 4096) == 0
        assert L.shmem_alltoallmem(h.world, d, s, 1024) == 0
        assert L.shmem_broadcastmem(h.world, d, s, 8192, 3) == 0
        sizes = (C.c_size_t * P)(*[64 * (k + 1) for k in range(P)])
        assert L.osc_collect(h.world, d, s, sizes, None) == 0
        h.sync()
    recs = [json.loads(x) for x in log.read_text().splitlines()]
    calls = [r["call"] for r in recs]
    assert calls == ["reduce", "reduce", "fcollect", "alltoall", "broadcast",
                     "collect"]
    r = recs[0]
    assert (r["op"], r["type"], r["algo"]) == ("sum", "f32", "hip")
    assert r["team"] == [0, 1, P] and r["count"] == n
    assert r["bytes"] == 2 * P * n * 4
    assert recs[1]["algo"] == "linear" and recs[1]["type"] == "i64"
    assert recs[2]["team"] == [2, 3, 5]
    assert recs[2]["bytes"] == 4096 * 5 + 5 * 5 * 4096
    assert recs[3]["bytes"] == 2 * P * P * 1024
    assert recs[4]["bytes"] == 8192 * (1 + P)
    tot = sum(64 * (k + 1) for k in range(P))
    assert recs[5]["count"] == tot and recs[5]["bytes"] == tot * (P + 1)
    for r in recs:
        assert r["event"] == "COLLECTIVES" and r["us"] > 0
        assert abs(r["GBps"] - r["bytes"] / (r["us"] * 1e3)) <= \
            0.05 + 1e-3 * r["GBps"]
        assert abs(r["pct_peak"] - 100 * r["GBps"] / 8000) <= 0.01


def test_no_records_without_osc_log(monkeypatch, tmp_path):
    log = tmp_path / "calls.jsonl"
    monkeypatch.delenv("OSC_LOG", raising=False)
    monkeypatch.setenv("OSC_LOG_FILE", str(log))
    with osc.Heap(npes=4, heap_bytes=1 << 20) as h:
        s, d = h.malloc(4096), h.malloc(4096)
        assert osc.typed("shmem_int_sum_reduce")(h.world, d, s, 16) == 0
        h.sync()
    assert not log.exists()

"""The batched-mode wave planner (osc_batch_plan_world): pure host logic.

Within a wave no two calls may write the same byte of the same PE; calls
keep their order across waves; every destination fits its region."""
import numpy as np
import pytest

import osc


def _calls(rng, n, world, max_bytes):
    out = []
    for _ in range(n):
        P = int(rng.integers(1, world + 1))
        start = int(rng.integers(0, world - P + 1))
        kind = int(rng.choice([0, 1, 2, 3]))
        c = osc.Call()
        c.kind = kind
        c.op, c.dtype = 5, 9  # sum, f64
        c.team_start, c.team_stride, c.team_npes = start, 1, P
        if kind == 0:
            c.count = int(rng.integers(1, max_bytes // 8))
        elif kind in (1, 2):
            c.count = int(rng.integers(1, max(2, max_bytes // P)))
        else:
            c.count = int(rng.integers(1, max_bytes))
        out.append(c)
    return out


def _dest_bytes(c):
    if c.kind == 0:
        return c.count * 8
    if c.kind in (1, 2):
        return c.count * c.team_npes
    return c.count


@pytest.mark.parametrize("seed", range(6))
def test_waves_never_overlap(seed):
    rng = np.random.default_rng(seed)
    world, region, base = 40, 1 << 20, 4096
    calls = _calls(rng, 300, world, region)
    planned, nw = osc.plan(calls, base, region, world_npes=world)
    assert nw >= 1
    prev_wave = 0
    for w in range(nw):
        used = [[] for _ in range(world)]
        for c in planned:
            if c.wave != w:
                continue
            lo, hi = c.dest_off, c.dest_off + _dest_bytes(c)
            assert base <= lo and hi <= base + region
            for p in range(c.team_start, c.team_start + c.team_npes):
                for a, b in used[p]:
                    assert hi <= a or b <= lo, (w, p)
                used[p].append((lo, hi))
    for c in planned:
        assert c.wave >= prev_wave  # call order is kept
        prev_wave = c.wave


def test_disjoint_teams_share_a_wave_at_offset_zero():
    calls = []
    for start in (0, 8, 16):
        c = osc.Call()
        c.kind, c.op, c.dtype = 0, 5, 9
        c.team_start, c.team_stride, c.team_npes = start, 1, 8
        c.count = 1000
        calls.append(c)
    planned, nw = osc.plan(calls, 0, 1 << 20, world_npes=24)
    assert nw == 1
    assert [c.dest_off for c in planned] == [0, 0, 0]


def test_call_larger_than_region_is_rejected():
    c = osc.Call()
    c.kind, c.op, c.dtype = 0, 5, 9
    c.team_start, c.team_stride, c.team_npes = 0, 1, 4
    c.count = 1 << 20
    with pytest.raises(osc.OscError) as e:
        osc.plan([c], 0, 1 << 20, world_npes=4)
    assert e.value.code == -9  # OSC_ERR_BUDGET


def test_team_outside_world_is_rejected():
    c = osc.Call()
    c.kind = 1
    c.team_start, c.team_stride, c.team_npes = 5, 2, 4  # last PE 11
    c.count = 16
    with pytest.raises(osc.OscError):
        osc.plan([c], 0, 1 << 20, world_npes=10)


def test_stride_that_wraps_32_bits_is_rejected():
    """start 0, stride 2^30, 3 members: the last member's PE, 2^31, does
    not fit an int; it is outside the world, not a negative PE."""
    c = osc.Call()
    c.kind = 1
    c.team_start, c.team_stride, c.team_npes = 0, 1 << 30, 3
    c.count = 16
    with pytest.raises(osc.OscError) as e:
        osc.plan([c], 0, 1 << 20, world_npes=10)
    assert e.value.code == -3  # OSC_ERR_TEAM


def test_count_whose_byte_count_wraps_is_rejected():
    """fcollect of 2^62 bytes per PE on 4 PEs: 4 x 2^62 wraps to 0 in 64
    bits; the call is rejected, not planned as an empty one."""
    c = osc.Call()
    c.kind = 1
    c.team_start, c.team_stride, c.team_npes = 0, 1, 4
    c.count = 1 << 62
    with pytest.raises(osc.OscError) as e:
        osc.plan([c], 0, 1 << 20, world_npes=4)
    assert e.value.code == -9  # OSC_ERR_BUDGET


def test_tuning_table_roundtrip(tmp_path):
    import ctypes as C
    L = osc.lib()
    s0, t0, g0 = C.c_int(), C.c_int(), C.c_int()
    assert L.osc_tune_get(3, 2, C.byref(s0), C.byref(t0), C.byref(g0)) == 0
    assert L.osc_tune_set(3, 2, 0, 96, 1024) == 0
    p = str(tmp_path / "t.tune").encode()
    assert L.osc_tune_save(p) == 0
    assert L.osc_tune_set(3, 2, 0, 8, 64) == 0
    assert L.osc_tune_load(p) == 0
    s, t, g = C.c_int(), C.c_int(), C.c_int()
    assert L.osc_tune_get(3, 2, C.byref(s), C.byref(t), C.byref(g)) == 0
    assert (t.value, g.value) == (96, 1024)
    assert L.osc_tune_set(9, 0, 0, 1, 1) != 0
    assert L.osc_tune_set(3, 2, 0, -1, 64) != 0   # tile 0 = direct; < 0 bad
    assert L.osc_tune_set(3, 2, 0, 0, 64) == 0
    assert L.osc_tune_set(3, 2, s0.value, t0.value, g0.value) == 0

"""Served mode on the GPU: requests from several teams arrive over time,
are admitted against an HBM budget and batched continuously.  Every
result must be identical alone and in the mix (per-request digests), the
per-call baseline must produce the same results, and requests that can
never fit the budget are rejected rather than blocking the queue."""
import numpy as np
import pytest

import osc
import osc_calls

pytestmark = pytest.mark.gpu

WORLD = 32
ARENA = 16 << 20
SCRATCH = 12 << 20


class ServeWorld:
    def __init__(self):
        self.w = osc_calls.CallWorld(npes=WORLD, max_reduce=1 << 16,
                                     move_bytes=4 << 20,
                                     dest_bytes=ARENA + SCRATCH,
                                     keys=["f32", "f64", "i64"])
        self.arena = self.w.dest_off
        self.scratch = self.w.dest_off + ARENA

    def cfg(self, mode, digests=1, budget=None, max_inflight=64):
        c = osc.ServeCfg()
        c.mode = mode
        c.max_inflight = max_inflight
        c.nstreams = 3
        c.max_group = 16
        c.dest_base = self.arena
        c.dest_bytes = ARENA
        c.page_bytes = 64 << 10
        c.budget_bytes = float(budget if budget is not None
                               else WORLD * ARENA)
        c.digests = digests
        return c


_S = []


@pytest.fixture
def sw():
    if not _S or not _S[0].w.h.alive:
        _S[:] = [ServeWorld()]
    return _S[0]


def requests(sw, seed, n, rate):
    rng = np.random.default_rng(seed)
    teams = osc_calls.served_teams(rng, sw.w, nteams=6, pmin=2, pmax=16)
    return osc_calls.served_mix(rng, sw.w, teams, n, rate, nmin=16,
                                nmax=1 << 16, bmin=16, bmax=64 << 10)


def test_results_same_alone_and_in_the_mix(sw):
    reqs = requests(sw, 1, 160, 2e5)  # all arrive within about 1 ms
    out, st = osc.serve(reqs, sw.cfg(osc.SERVE_BATCHED))
    assert st.completed == len(reqs) and st.rejected == 0
    assert st.groups < len(reqs)          # requests were batched together
    assert st.max_inflight_seen > 1       # ... and overlapped
    groups = {}
    for r in out:
        assert r.status == 0
        assert r.arrival_s <= r.admit_s <= r.launch_s <= r.done_s
        groups.setdefault(r.group, []).append(r)
    assert max(len(g) for g in groups.values()) > 1
    for r in out:
        alone = osc.request_alone(r.call, sw.scratch)
        assert alone == r.digest, (r.call.kind, r.call.team_npes,
                                   r.call.count)


def test_per_call_baseline_gives_the_same_results(sw):
    reqs = requests(sw, 2, 80, 1e5)
    b, st_b = osc.serve(reqs, sw.cfg(osc.SERVE_BATCHED))
    p, st_p = osc.serve(reqs, sw.cfg(osc.SERVE_PER_CALL))
    assert st_b.completed == st_p.completed == len(reqs)
    assert st_p.groups == len(reqs)
    assert [r.digest for r in b] == [r.digest for r in p]


def test_budget_limits_what_is_in_flight(sw):
    reqs = requests(sw, 3, 60, 1e6)
    areas = []
    for r in reqs:
        c = r.call
        db = c.count * {8: 4, 9: 8, 7: 8}.get(c.dtype, 1) \
            if c.kind == osc.CALL_REDUCE else c.count * c.team_npes
        areas.append(db * c.team_npes)
    budget = max(areas) * 1.01            # room for the largest, alone
    out, st = osc.serve(reqs, sw.cfg(osc.SERVE_BATCHED, budget=budget))
    assert st.completed == len(reqs)
    assert st.peak_inflight_bytes <= budget
    # a request larger than the budget can never run: rejected, not stuck
    big = requests(sw, 4, 1, 1e3)
    big[0].call.kind = osc.CALL_ALLTOALL
    big[0].call.count = 64 << 10
    big[0].call.team_npes = 16
    big[0].call.team_start = 0
    out, st = osc.serve(big, sw.cfg(osc.SERVE_BATCHED, budget=1 << 20))
    assert st.rejected == 1 and out[0].status == -9


def test_arrivals_must_be_ordered(sw):
    reqs = requests(sw, 5, 3, 1e3)
    reqs[1].arrival_s, reqs[2].arrival_s = reqs[2].arrival_s, reqs[1].arrival_s
    with pytest.raises(osc.OscError):
        osc.serve(reqs, sw.cfg(osc.SERVE_BATCHED))


def _alg_bytes(c):
    """Algorithmic bytes of a call: 2*P*n*s (reduce), P*b + P*P*b
    (fcollect), 2*P*P*b (alltoall)."""
    P = c.team_npes
    if c.kind == osc.CALL_REDUCE:
        names = {v: k for k, v in osc.DTYPES.items()}
        return 2 * P * c.count * np.dtype(osc.NP_DTYPE[names[c.dtype]]).itemsize
    if c.kind == osc.CALL_FCOLLECT:
        return P * c.count + P * P * c.count
    return 2 * P * P * c.count


def _at_once(reqs):
    for r in reqs:
        r.arrival_s = 0.0
    return reqs


@pytest.mark.parametrize("mode", [osc.SERVE_BATCHED, osc.SERVE_PER_CALL])
def test_size_order_admits_shortest_first(sw, mode):
    """Everything arrives at once and one request runs at a time, so the
    launch order is the queue order: ascending algorithmic bytes (ties by
    index) with aging off, index order in arrival order."""
    reqs = _at_once(requests(sw, 6, 40, 1e3))
    cfg = sw.cfg(mode, digests=0, max_inflight=1)
    cfg.order, cfg.age_bytes_per_s = osc.ORDER_SIZE, 0.0
    out, st = osc.serve(reqs, cfg)
    assert st.completed == len(reqs)
    launched = sorted(range(len(out)), key=lambda i: out[i].group)
    want = sorted(range(len(out)), key=lambda i: (_alg_bytes(out[i].call), i))
    assert launched == want
    cfg.order = osc.ORDER_ARRIVAL
    out, _ = osc.serve(_at_once(requests(sw, 6, 40, 1e3)), cfg)
    assert sorted(range(len(out)), key=lambda i: out[i].group) == \
        list(range(len(out)))


def test_size_order_ages_waiting_requests(sw):
    """Aging: a request that arrived earlier goes first unless the later one
    is smaller by more than age_bytes_per_s times the arrival gap.  The
    arrivals are picoseconds apart (all queued together); a weight that
    makes each picosecond worth 1 GB gives arrival order, weight 0 with
    simultaneous arrivals gives size order."""
    reqs = requests(sw, 7, 30, 1e3)
    for i, r in enumerate(reqs):
        r.arrival_s = i * 1e-12
    cfg = sw.cfg(osc.SERVE_BATCHED, digests=0, max_inflight=1)
    cfg.order, cfg.age_bytes_per_s = osc.ORDER_SIZE, 1e21
    assert max(_alg_bytes(r.call) for r in reqs) < 1e9
    out, _ = osc.serve(reqs, cfg)
    assert sorted(range(len(out)), key=lambda i: out[i].group) == \
        list(range(len(out)))
    cfg.age_bytes_per_s = 0.0
    out, _ = osc.serve(_at_once(reqs), cfg)
    assert sorted(range(len(out)), key=lambda i: out[i].group) == \
        sorted(range(len(out)), key=lambda i: (_alg_bytes(out[i].call), i))


def test_size_order_same_results_and_bad_order_rejected(sw):
    reqs = requests(sw, 8, 120, 2e5)
    cfg = sw.cfg(osc.SERVE_BATCHED)
    cfg.order, cfg.age_bytes_per_s = osc.ORDER_SIZE, 1e11
    out, st = osc.serve(reqs, cfg)
    assert st.completed == len(reqs) and st.rejected == 0
    for r in out:
        assert r.arrival_s <= r.admit_s <= r.launch_s <= r.done_s
        assert osc.request_alone(r.call, sw.scratch) == r.digest
    bad = sw.cfg(osc.SERVE_BATCHED)
    bad.order = 2
    with pytest.raises(osc.OscError):
        osc.serve(reqs[:2], bad)
    bad.order, bad.age_bytes_per_s = osc.ORDER_SIZE, -1.0
    with pytest.raises(osc.OscError):
        osc.serve(reqs[:2], bad)


def test_request_needing_more_whole_pages_than_the_arena_is_rejected(sw):
    """dest_bytes that is not a multiple of the page holds floor(dest_bytes
    / page) whole pages; a request that fits in bytes but needs one page
    more can never be placed: rejected, not left waiting."""
    reqs = requests(sw, 9, 1, 1e3)
    c = reqs[0].call
    c.kind, c.team_start, c.team_npes, c.root = osc.CALL_FCOLLECT, 0, 16, 0
    c.count = 1048448                      # dest 16 x 1048448 = 16 MiB - 2 KiB
    cfg = sw.cfg(osc.SERVE_BATCHED)
    cfg.dest_bytes = ARENA - 1000          # 255 whole 64 KiB pages
    assert 16 * c.count <= cfg.dest_bytes
    assert -(-16 * c.count // cfg.page_bytes) > cfg.dest_bytes // cfg.page_bytes
    out, st = osc.serve(reqs, cfg)
    assert st.rejected == 1 and out[0].status == -9


def test_request_reading_from_the_destination_arena_is_rejected(sw):
    """The server hands out the arena's pages to requests in flight, so a
    request whose source lies in the arena could read another request's
    partial results: it is rejected at arrival; the others complete."""
    reqs = requests(sw, 10, 12, 1e5)
    bad = reqs[5].call
    bad.src_off = sw.arena + ARENA - 256   # its last 256 bytes, or more
    out, st = osc.serve(reqs, sw.cfg(osc.SERVE_BATCHED))
    assert st.rejected == 1 and out[5].status == -1
    assert st.completed == len(reqs) - 1
    assert all(r.status == 0 for k, r in enumerate(out) if k != 5)


def test_wrapping_or_empty_destination_arena_is_an_argument_error(sw):
    """dest_base + dest_bytes must not wrap in 64 bits, and the arena must
    hold at least one page: a wrapped or empty arena is OSC_ERR_ARG before
    any request is looked at, not a loop that never places anything."""
    reqs = requests(sw, 11, 2, 1e4)
    for base, nbytes in [(sw.arena, 2 ** 64 - 256),
                         (2 ** 64 - 256, 512),
                         (sw.arena, 1000)]:        # less than one page
        cfg = sw.cfg(osc.SERVE_BATCHED)
        cfg.dest_base, cfg.dest_bytes = base, nbytes
        with pytest.raises(osc.OscError) as e:
            osc.serve(reqs, cfg)
        assert e.value.code == -1

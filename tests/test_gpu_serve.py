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

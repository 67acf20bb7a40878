"""The library's tuning switches reject values outside their range and take
the valid ones without a device (they only set host-side state)."""
import osc


def test_move_interleave_switch():
    L = osc.lib()
    assert L.osc_set_move_interleave(2) != 0
    assert L.osc_set_move_interleave(-1) != 0
    assert L.osc_set_move_interleave(0) == 0
    assert L.osc_set_move_interleave(1) == 0


def test_reduce_split_switch():
    L = osc.lib()
    for bad in (3, 5, 32, -2):
        assert L.osc_set_reduce_split(bad) != 0
    for good in (0, 1, 2, 4, 8, 16, -1):
        assert L.osc_set_reduce_split(good) == 0


def test_batch_phases_switch():
    L = osc.lib()
    for bad in (2, -1):
        assert L.osc_set_batch_phases(bad) != 0
    assert L.osc_set_batch_phases(0) == 0
    assert L.osc_set_batch_phases(1) == 0

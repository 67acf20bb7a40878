"""Pins of the oracle's collect, alltoalls, active-set broadcast and strided
copy against independent formulations: numpy's roll for the rotation, the
concatenation closed form for Bruck collect, a fancy-index gather for the
strided alltoall, and numpy slicing for the strided copy.  CPU only."""
import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("size,dist", [(1, 0), (7, 3), (64, 64), (100, 37),
                                       (4096, 1000), (12, 0)])
def test_rotate_is_roll(size, dist):
    buf = np.random.default_rng(size).integers(0, 256, size, dtype=np.uint8)
    assert np.array_equal(O.rotate(buf, dist), np.roll(buf, dist))
    assert np.array_equal(O.rotate(buf, dist), O.rotate_inplace(buf, dist))


@pytest.mark.parametrize("P", [1, 2, 3, 5, 8, 13, 17])
def test_collect_bruck_is_concatenation(P):
    rng = np.random.default_rng(P)
    for trial in range(6):
        sizes = rng.integers(0, 40, P)
        if trial == 0:
            sizes[:] = 0
        elif trial == 1:
            sizes[:] = 9                      # the fcollect case
        elif trial == 2:
            sizes[rng.integers(0, P)] = 0     # one empty contributor
        rows = [rng.integers(0, 256, int(k), dtype=np.uint8) for k in sizes]
        got = O.collect(rows)
        want = O.collect_closed_form(rows)
        assert np.array_equal(got, want), (P, sizes)


def test_collect_uniform_equals_fcollect():
    rng = np.random.default_rng(4)
    src = rng.integers(0, 256, (6, 24), dtype=np.uint8)
    assert np.array_equal(O.collect(list(src)), O.fcollect(src))


@pytest.mark.parametrize("P,es,dst,sst,n", [
    (1, 4, 1, 1, 3), (2, 8, 1, 1, 5), (3, 4, 2, 1, 4), (5, 1, 1, 3, 7),
    (4, 8, 3, 2, 2), (7, 2, 1, 1, 1), (6, 16, 2, 2, 3), (3, 3, 1, 1, 2)])
def test_alltoalls_shift_is_gather(P, es, dst, sst, n):
    rng = np.random.default_rng(P * 100 + es)
    sbytes = ((P * n - 1) * sst + 1) * es
    dbytes = ((P * n - 1) * dst + 1) * es
    src = rng.integers(0, 256, (P, sbytes), dtype=np.uint8)
    got = O.alltoalls(src, es, dst, sst, n, dbytes)
    want = O.alltoalls_closed_form(src, es, dst, sst, n, dbytes)
    assert np.array_equal(got, want)


def test_alltoalls_unit_strides_is_alltoall():
    rng = np.random.default_rng(9)
    P, n, es = 5, 3, 8
    src = rng.integers(0, 256, (P, P * n * es), dtype=np.uint8)
    got = O.alltoalls(src, es, 1, 1, n, P * n * es)
    assert np.array_equal(got, O.alltoall(src))
    assert np.array_equal(got, O.alltoall_closed_form(src))


def test_alltoalls_leaves_gaps_untouched():
    rng = np.random.default_rng(10)
    P, n, es, dst = 3, 2, 4, 3
    dbytes = ((P * n - 1) * dst + 1) * es
    src = rng.integers(0, 256, (P, P * n * es), dtype=np.uint8)
    init = np.full((P, dbytes), 0xEE, np.uint8)
    got = O.alltoalls(src, es, dst, 1, n, dbytes, dest_init=init)
    mask = np.zeros(dbytes, bool)
    for e in range(P * n):
        mask[e * dst * es:e * dst * es + es] = True
    assert np.all(got[:, ~mask] == 0xEE)


@pytest.mark.parametrize("P,root", [(1, 0), (4, 0), (5, 3), (8, 7)])
def test_broadcast_active_set_skips_root(P, root):
    rng = np.random.default_rng(P + root)
    src = rng.integers(0, 256, (P, 33), dtype=np.uint8)
    init = rng.integers(0, 256, (P, 33), dtype=np.uint8)
    got = O.broadcast_active_set(src, root, init)
    for j in range(P):
        want = init[root] if j == root else src[root]
        assert np.array_equal(got[j], want)
    # the team form writes the root too
    assert np.array_equal(O.broadcast(src, root)[root], src[root])


@pytest.mark.parametrize("es,tst,sst,n", [(4, 1, 1, 10), (8, 3, 2, 7),
                                          (2, 1, 5, 9), (1, 4, 1, 16)])
def test_strided_copy_is_slicing(es, tst, sst, n):
    rng = np.random.default_rng(es * 7 + n)
    src = rng.integers(0, 256, (n * sst) * es, dtype=np.uint8)
    dst0 = rng.integers(0, 256, (n * tst) * es, dtype=np.uint8)
    got = O.strided_copy(dst0, src, es, tst, sst, n)
    want = dst0.copy().reshape(-1, es)
    want[::tst][:n] = src.reshape(-1, es)[::sst][:n]
    assert np.array_equal(got, want.reshape(-1))


def test_strided_copy_negative_and_zero_strides():
    rng = np.random.default_rng(3)
    es, n = 4, 6
    src = rng.integers(0, 256, n * es, dtype=np.uint8)
    # negative source stride: element 0 is the last one in memory
    got = O.strided_copy(np.zeros(n * es, np.uint8), src, es, 1, -1, n,
                         src_base=(n - 1) * es)
    want = src.reshape(n, es)[::-1].reshape(-1)
    assert np.array_equal(got, want)
    # zero target stride: elements land in order, the last one stays
    got = O.strided_copy(np.zeros(es, np.uint8), src, es, 0, 1, n)
    assert np.array_equal(got, src[(n - 1) * es:])


def test_cpu_timing_harness_runs_every_collective():
    """The CPU baseline's timing entry points run and return a positive
    median time at tiny sizes."""
    assert O.cpu_bench_procs("reduce", 3, 1000, 3, "sum", "i32") > 0
    assert O.cpu_bench_procs("alltoall", 5, 256, 3) > 0
    assert O.cpu_bench_procs("fcollect", 5, 256, 3) > 0


def test_process_per_pe_baseline_reproduces_the_oracle():
    """The CPU baseline's process-per-PE run (fork, one MAP_SHARED heap,
    pSync handshakes, memcpy puts) leaves the oracle's results in every
    PE's destination: recursive doubling (power-of-two and folded teams),
    Bruck fcollect with its rotation, shift-exchange alltoall."""
    rng = np.random.default_rng(8)
    for P in (1, 3, 6, 8):
        x = rng.standard_normal((P, 777)).astype(np.float32)
        _, d = O.cpu_bench_procs("reduce", P, 777, 2, "sum", "f32", src=x,
                                 want_dest=True)
        assert np.array_equal(d.view(np.uint32),
                              O.rec_dbl("sum", "f32", x).view(np.uint32))
        y = rng.integers(-2 ** 40, 2 ** 40, (P, 333))
        _, d = O.cpu_bench_procs("reduce", P, 333, 1, "max", "i64", src=y,
                                 want_dest=True)
        assert np.array_equal(d.view(np.int64), O.rec_dbl("max", "i64", y))
    b = rng.integers(0, 256, (7, 40), dtype=np.uint8)
    _, d = O.cpu_bench_procs("fcollect", 7, 40, 2, src=b, want_dest=True)
    assert np.array_equal(d, O.fcollect(b))
    b = rng.integers(0, 256, (5, 5 * 24), dtype=np.uint8)
    _, d = O.cpu_bench_procs("alltoall", 5, 24, 2, src=b, want_dest=True)
    assert np.array_equal(d, O.alltoall(b))

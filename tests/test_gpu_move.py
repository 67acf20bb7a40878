"""GPU parity of fcollect, alltoall, broadcast and put/get through the C ABI:
bit-exact against the oracle's step-by-step restatements of the reference
(Bruck in-place fcollect with its byte rotation, shift-exchange alltoall,
linear broadcast) and against the closed-form self-describing patterns of
the data-movement workloads."""
import ctypes as C

import numpy as np
import pytest

import osc
import oracle as O
from gpu_util import move

pytestmark = pytest.mark.gpu

SIZES = [1, 3, 8, 16, 17, 24, 64, 65, 100, 256, 1000, 4096, 4104, 16392]
TEAMS = [1, 2, 3, 5, 8, 13, 64]


_HEAP = []


@pytest.fixture
def heap():
    """64 PEs x 80 MiB, re-created if another test re-initialised the
    library with a different world."""
    if not _HEAP or not _HEAP[0].alive:
        _HEAP[:] = [osc.Heap(npes=64, heap_bytes=80 << 20)]
    return _HEAP[0]


def team_of(h, P):
    return h.world if P == h.npes else h.split(0, 1, P)


@pytest.mark.parametrize("nb", SIZES)
def test_fcollect(heap, nb):
    rng = np.random.default_rng(nb)
    for P in TEAMS:
        if P * nb > (16 << 20):
            continue
        src = rng.integers(0, 256, (P, nb), dtype=np.uint8)
        got = move(heap, "fcollect", src, nb, team=team_of(heap, P))
        assert np.array_equal(got, O.fcollect(src)), P


@pytest.mark.parametrize("il", [1, 0])
@pytest.mark.parametrize("nb", [(64 << 10) - 16, 64 << 10, (64 << 10) + 16,
                                (128 << 10) - 16, (128 << 10) + 48,
                                (256 << 10) + 16, (512 << 10) - 16, 512 << 10,
                                (512 << 10) + 16, (1 << 20) + 48, 3 << 19])
def test_fcollect_interleaved_blocks(heap, nb, il):
    """Blocks of 64 KiB and more (16-byte units) take the interleaved
    gather, 2, 4, 8 or 16 units per thread by block size: whole
    super-chunks without bounds tests, then the ragged rest; with the
    interleave off, the destination-major chunks."""
    rng = np.random.default_rng(nb)
    osc.check(heap.L.osc_set_move_interleave(il), "move order")
    try:
        for P in (2, 5, 13):
            src = rng.integers(0, 256, (P, nb), dtype=np.uint8)
            got = move(heap, "fcollect", src, nb, team=team_of(heap, P))
            assert np.array_equal(got, O.fcollect(src)), P
    finally:
        heap.L.osc_set_move_interleave(1)


@pytest.mark.parametrize("root", [0, 6])
@pytest.mark.parametrize("nb", [(96 << 10) + 16, 512 << 10, (1 << 20) + 16])
def test_broadcast_interleaved_blocks(heap, root, nb):
    rng = np.random.default_rng(root + nb)
    P = 9
    src = rng.integers(0, 256, (P, nb), dtype=np.uint8)
    got = move(heap, "broadcast", src, nb, team=team_of(heap, P), root=root)
    assert np.array_equal(got, O.broadcast(src, root))


@pytest.mark.parametrize("il", [1, 0])
@pytest.mark.parametrize("nb", [(16 << 10) - 16, 16 << 10, (20 << 10) + 48,
                                (24 << 10) - 16, 24 << 10, (40 << 10) + 16,
                                (64 << 10) - 16, 64 << 10, (64 << 10) + 16,
                                (96 << 10) + 48, (192 << 10) + 16, 256 << 10])
def test_alltoall_interleaved_blocks(heap, nb, il):
    """Blocks of 64 KiB and more (16-byte units) take the interleaved
    alltoall: whole super-chunks without bounds tests, then the ragged
    rest.  P^2 x block above 256 MiB (40 members x 192 or 256 KiB) takes
    streaming loads, below it cached loads; with the interleave off, and
    for smaller blocks at these footprints, the pair-major chunks (below
    16 KiB, several pairs per workgroup)."""
    rng = np.random.default_rng(7 + nb)
    osc.check(heap.L.osc_set_move_interleave(il), "move order")
    try:
        for P in (2, 5, 13, 40):
            if P * nb > (12 << 20):
                continue
            src = rng.integers(0, 256, (P, P * nb), dtype=np.uint8)
            got = move(heap, "alltoall", src, nb, team=team_of(heap, P))
            assert np.array_equal(got, O.alltoall(src)), P
    finally:
        heap.L.osc_set_move_interleave(1)


@pytest.mark.parametrize("nb", [23536, 24576, 24592, (32 << 10) + 48])
def test_alltoall_streaming_mid_blocks(nb):
    """112 members, blocks around 24 KiB and 32 KiB: P^2 x block above
    256 MiB, so blocks of 24 KiB to 64 KiB take the interleaved alltoall
    with streaming loads (one ragged super-chunk per pair) and 23,536
    bytes the pair-major chunks."""
    P = 112
    rng = np.random.default_rng(nb)
    src = rng.integers(0, 256, (P, P * nb), dtype=np.uint8)
    with osc.Heap(npes=P, heap_bytes=12 << 20) as h:
        got = move(h, "alltoall", src, nb, team=h.world)
    assert np.array_equal(got, O.alltoall(src))


@pytest.mark.parametrize("nb", SIZES)
def test_alltoall(heap, nb):
    rng = np.random.default_rng(1000 + nb)
    for P in TEAMS:
        if P * nb > (16 << 20):
            continue
        src = rng.integers(0, 256, (P, P * nb), dtype=np.uint8)
        got = move(heap, "alltoall", src, nb, team=team_of(heap, P))
        assert np.array_equal(got, O.alltoall(src)), P


@pytest.mark.parametrize("root", [0, 4, 12])
@pytest.mark.parametrize("nb", [1, 17, 4096, 100003])
def test_broadcast(heap, root, nb):
    rng = np.random.default_rng(root * 7 + nb)
    P = 13
    src = rng.integers(0, 256, (P, nb), dtype=np.uint8)
    got = move(heap, "broadcast", src, nb, team=team_of(heap, P), root=root)
    assert np.array_equal(got, O.broadcast(src, root))


def test_typed_entry_points(heap):
    """shmem_int64_alltoall / shmem_int32_fcollect count elements, not
    bytes; strided team."""
    team = heap.split(1, 3, 7)
    pes = [1 + 3 * k for k in range(7)]
    n = 33
    rng = np.random.default_rng(5)
    src = rng.integers(-2 ** 62, 2 ** 62, (7, 7 * n), dtype=np.int64)
    s, d = heap.malloc(src[0].nbytes), heap.malloc(src[0].nbytes)
    for k, pe in enumerate(pes):
        heap.put(s, src[k], pe)
    assert osc.typed("shmem_int64_alltoall")(team, d, s, n) == 0
    got = np.stack([heap.get(d, 7 * n, np.int64, pe) for pe in pes])
    want = O.alltoall(src.view(np.uint8).reshape(7, -1)).view(np.int64)
    assert np.array_equal(got, want)
    src32 = rng.integers(-2 ** 31, 2 ** 31, (7, n), dtype=np.int32)
    for k, pe in enumerate(pes):
        heap.put(s, src32[k], pe)
    assert osc.typed("shmem_int32_fcollect")(team, d, s, n) == 0
    got = np.stack([heap.get(d, 7 * n, np.int32, pe) for pe in pes])
    assert np.array_equal(got, np.tile(src32.reshape(-1), (7, 1)))
    heap.free(s)
    heap.free(d)


def test_overlap_rejected(heap):
    s = heap.malloc(4096)
    assert heap.L.shmem_fcollectmem(heap.world, s, s, 8) != 0
    assert heap.L.shmem_alltoallmem(heap.world, s, s + 8, 8) != 0
    heap.free(s)


def test_alltoall_64pe_64KiB_full_size():
    """alltoall, 64 PEs, 65,536-byte blocks of int64
    src_i[j*8192 + idx] = i<<40 | j<<20 | idx; every dest must read
    i<<40 | j<<20 | idx in block i of PE j."""
    P, blk = 64, 8192
    with osc.Heap(npes=P, heap_bytes=2 * P * blk * 8 + (1 << 20)) as h:
        s, d = h.malloc(P * blk * 8), h.malloc(P * blk * 8)
        h.fill(None, s, P * blk, "i64", "a2a_pattern", blk, 0)
        assert osc.typed("shmem_int64_alltoall")(h.world, d, s, blk) == 0
        idx = np.arange(blk, dtype=np.int64)
        i = np.arange(P, dtype=np.int64)[:, None]
        for j in range(P):
            got = h.get(d, P * blk, np.int64, j).reshape(P, blk)
            want = (i << 40) | (j << 20) | idx[None, :]
            assert np.array_equal(got, want), j


def test_fcollect_256pe_1MiB_full_size():
    """fcollect, 256 PEs x 1 MiB int32 (pe<<24 | idx); dest of sampled PEs
    is the PE-ordered concatenation."""
    P, n = 256, 262144
    with osc.Heap(npes=P, heap_bytes=(P + 1) * n * 4 + (1 << 20)) as h:
        s, d = h.malloc(n * 4), h.malloc(P * n * 4)
        h.fill(None, s, n, "i32", "fcol_pattern")
        assert osc.typed("shmem_int32_fcollect")(h.world, d, s, n) == 0
        idx = np.arange(n, dtype=np.uint32)
        want = ((np.arange(P, dtype=np.uint32)[:, None] << np.uint32(24)) |
                idx[None, :]).astype(np.uint32).view(np.int32).reshape(-1)
        for pe in (0, 77, 255):
            assert np.array_equal(h.get(d, P * n, np.int32, pe), want), pe


def test_put_get_between_pes(heap):
    L = heap.L
    a = heap.malloc(4096)
    b = heap.malloc(4096)
    data = np.arange(1024, dtype=np.int32)
    heap.put(a, data, 0)
    L.shmem_putmem(b, a, 4096, 9)           # device -> PE 9's copy of b
    assert np.array_equal(heap.get(b, 1024, np.int32, 9), data)
    out = np.zeros(1024, np.int32)
    L.shmem_getmem(out.ctypes.data, b, 4096, 9)   # PE 9's b -> host
    assert np.array_equal(out, data)
    heap.free(a)
    heap.free(b)


def test_heap_allocator(heap):
    L = heap.L
    a = L.shmem_malloc(1000)
    b = L.shmem_align(1 << 16, 10)
    assert b % (1 << 16) == 0
    c = L.shmem_calloc(100, 8)
    assert np.all(heap.get(c, 100, np.int64, 5) == 0)
    heap.put(a, np.arange(250, dtype=np.int32), 3)
    a2 = L.shmem_realloc(a, 1 << 20)
    assert np.array_equal(heap.get(a2, 250, np.int32, 3), np.arange(250))
    assert L.shmem_ptr(a2, 3) == L.shmem_ptr(a2, 0) + 3 * \
        heap.L.osc_slice_bytes()
    for p in (a2, b, c):
        L.shmem_free(p)
    assert L.shmem_malloc(1 << 40) is None    # exhaustion -> NULL
    # all freed memory is reusable in one piece
    big = L.shmem_malloc(70 << 20)
    assert big
    L.shmem_free(big)

"""Point-to-point synchronization on symmetric variables through the C ABI
(shmem_<T>_test, wait_until and the _all/_any/_some/_vector forms):

* the test forms on values placed with puts: every comparison operator
  against Python's, status masks, per-element comparison values, unsigned
  compares above 2^63 (as unsigned), nothing included, bad arguments;
* the wait forms whose condition already holds return at once;
* SPMD mode: PE threads block in wait_until, wait_until_any,
  wait_until_some and wait_until_all_vector until puts from other PE
  threads satisfy them."""
import ctypes as C
import operator

import numpy as np
import pytest

import osc

pytestmark = pytest.mark.gpu

EQ, NE, GT, LE, LT, GE = range(6)
PYOP = [operator.eq, operator.ne, operator.gt, operator.le, operator.lt,
        operator.ge]
SIZE_MAX = 2 ** 64 - 1
SZ, VP, I, IP = C.c_size_t, C.c_void_p, C.c_int, C.POINTER(C.c_int)


def fn(name, res, args):
    f = getattr(osc.lib(), name)
    f.restype, f.argtypes = res, args
    return f


def long_fns():
    L = C.c_long
    return dict(
        test=fn("shmem_long_test", I, [VP, I, L]),
        wait_until=fn("shmem_long_wait_until", None, [VP, I, L]),
        test_all=fn("shmem_long_test_all", I, [VP, SZ, IP, I, L]),
        test_any=fn("shmem_long_test_any", SZ, [VP, SZ, IP, I, L]),
        test_some=fn("shmem_long_test_some", SZ, [VP, SZ, VP, IP, I, L]),
        wait_all=fn("shmem_long_wait_until_all", None, [VP, SZ, IP, I, L]),
        wait_any=fn("shmem_long_wait_until_any", SZ, [VP, SZ, IP, I, L]),
        wait_some=fn("shmem_long_wait_until_some", SZ,
                     [VP, SZ, VP, IP, I, L]),
        test_all_v=fn("shmem_long_test_all_vector", I, [VP, SZ, IP, I, VP]),
        test_any_v=fn("shmem_long_test_any_vector", SZ, [VP, SZ, IP, I, VP]),
        test_some_v=fn("shmem_long_test_some_vector", SZ,
                       [VP, SZ, VP, IP, I, VP]),
        wait_all_v=fn("shmem_long_wait_until_all_vector", None,
                      [VP, SZ, IP, I, VP]),
        wait_some_v=fn("shmem_long_wait_until_some_vector", SZ,
                       [VP, SZ, VP, IP, I, VP]))


def mask(bits):
    return (C.c_int * len(bits))(*bits)


def test_test_forms_against_python():
    if osc.Heap._current is not None:
        osc.Heap._current.close()
    F = long_fns()
    vals = np.array([-5, 0, 3, 3, 7, -1, 2 ** 40, 9], np.int64)
    n = len(vals)
    with osc.Heap(npes=4, heap_bytes=1 << 20) as h:
        L = h.L
        L.shmem_ptr.restype, L.shmem_ptr.argtypes = VP, [VP, I]
        v = h.malloc(8 * n)
        h.put(v, vals, 2)
        p = L.shmem_ptr(v, 2)                  # PE 2's copy
        for cmp in range(6):
            for x in (-5, 0, 3, 8):
                want = [PYOP[cmp](int(a), x) for a in vals]
                for k in range(n):
                    assert F["test"](p + 8 * k, cmp, x) == want[k]
                assert F["test_all"](p, n, None, cmp, x) == all(want)
                hit = [k for k in range(n) if want[k]]
                assert F["test_any"](p, n, None, cmp, x) == \
                    (hit[0] if hit else SIZE_MAX)
                idx = (C.c_size_t * n)()
                assert F["test_some"](p, n, idx, None, cmp, x) == len(hit)
                assert list(idx[:len(hit)]) == hit
        # status: nonzero excludes; test_all is over the included ones
        st = mask([0, 1, 0, 0, 1, 1, 0, 1])    # include 0, 2, 3, 6
        assert F["test_all"](p, n, st, NE, 0) == 1    # 0 is excluded
        assert F["test_any"](p, n, st, GT, 100) == 6
        idx = (C.c_size_t * n)()
        assert F["test_some"](p, n, idx, st, EQ, 3) == 2
        assert list(idx[:2]) == [2, 3]
        # per-element comparison values
        cv = (C.c_long * n)(*[int(a) for a in vals])
        assert F["test_all_v"](p, n, None, EQ, cv) == 1
        cv[4] = 8
        assert F["test_all_v"](p, n, None, EQ, cv) == 0
        assert F["test_any_v"](p, n, None, LT, cv) == 4
        assert F["test_some_v"](p, n, idx, None, GE, cv) == n - 1
        # nothing included
        none = mask([1] * n)
        assert F["test_all"](p, n, none, EQ, 12345) == 1
        assert F["test_any"](p, n, none, EQ, 0) == SIZE_MAX
        assert F["test_some"](p, n, idx, none, EQ, 0) == 0
        assert F["test_any"](p, 0, None, EQ, 0) == SIZE_MAX
        # waits whose condition holds return at once
        F["wait_until"](p + 8 * 4, GE, 7)
        F["wait_all"](p, n, st, GT, -10)
        assert F["wait_any"](p, n, None, EQ, 9) == 7
        assert F["wait_some"](p, n, idx, None, LT, 0) == 2
        assert list(idx[:2]) == [0, 5]
        F["wait_all_v"](p, n, None, LE, cv)
        assert F["wait_some_v"](p, n, idx, st, EQ, cv) == 4
        assert list(idx[:4]) == [0, 2, 3, 6]
        # bad arguments: error recorded, "nothing satisfied" returned
        L.osc_clear_error()
        assert F["test_any"](p, n, None, 6, 0) == SIZE_MAX      # no such cmp
        assert L.osc_last_error_code() != 0
        L.osc_clear_error()
        host = (C.c_long * n)()
        assert F["test"](C.addressof(host), EQ, 0) == 0       # not symmetric
        assert L.osc_last_error_code() != 0
        F["wait_until"](C.addressof(host), EQ, 0)              # returns
        assert F["test_some"](p, n, None, None, EQ, 3) == 0    # NULL indices
        L.osc_clear_error()


def test_unsigned_compares_as_unsigned():
    if osc.Heap._current is not None:
        osc.Heap._current.close()
    f = fn("shmem_uint64_test", I, [VP, I, C.c_uint64])
    any_ = fn("shmem_ulonglong_test_any", SZ,
              [VP, SZ, IP, I, C.c_ulonglong])
    with osc.Heap(npes=2, heap_bytes=1 << 20) as h:
        v = h.malloc(16)
        h.put(v, np.array([2 ** 63 + 5, 5], np.uint64), 0)
        assert f(v, GT, 5) == 1          # signed compare would say "less"
        assert f(v, LT, 2 ** 63) == 0
        assert any_(v, 2, None, GE, 2 ** 63) == 0


PE_FN = C.CFUNCTYPE(C.c_int, C.c_int, C.c_void_p)


def test_spmd_waits_block_until_puts_arrive():
    """PE 0 waits in four forms; each condition becomes true only through
    puts that other PE threads issue after their own barrier, so PE 0 must
    block and then see them."""
    if osc.Heap._current is not None:
        osc.Heap._current.close()
    npes = 4
    F = long_fns()
    L = osc.lib()
    L.osc_spmd_run.restype = C.c_int
    L.osc_spmd_run.argtypes = [C.c_int, C.c_size_t, C.c_int, PE_FN, VP]
    L.shmem_long_p.restype, L.shmem_long_p.argtypes = None, [VP, C.c_long, I]
    got, errors = {}, []

    def body(pe, _arg):
        try:
            flags = L.shmem_calloc(8, 8)       # zeroed on every PE
            L.shmem_barrier_all()
            if pe == 0:
                F["wait_until"](flags, GE, 1)          # set by PE 1
                first = F["wait_any"](flags + 8, 3, None, NE, 0)
                idx = (C.c_size_t * 8)()
                some = F["wait_some"](flags, 8, idx, mask([1, 1, 1, 1, 0, 0,
                                                           1, 1]), EQ, 5)
                cv = (C.c_long * 2)(6, 7)
                F["wait_all_v"](flags + 48, 2, None, GE, cv)
                got[0] = (first, some, list(idx[:some]))
            else:
                # PE k writes into PE 0's copy once PE 0 is waiting or soon
                if pe == 1:
                    L.shmem_long_p(flags, 1, 0)
                    L.shmem_long_p(flags + 16, 3, 0)       # index 2 of 1..3
                if pe == 2:
                    L.shmem_long_p(flags + 40, 5, 0)       # index 5
                    L.shmem_long_p(flags + 48, 6, 0)
                if pe == 3:
                    L.shmem_long_p(flags + 56, 9, 0)
            L.shmem_barrier_all()
            L.shmem_free(flags)
            return 0
        except Exception as e:  # reported by the test, not lost in a thread
            errors.append((pe, repr(e)))
            return 1
    cb = PE_FN(body)
    assert L.osc_spmd_run(npes, 1 << 20, 0, cb, None) == 0
    assert not errors, errors
    first, some, idx = got[0]
    assert first == 1          # flags[1 + 1] is the first nonzero of [1..3]
    assert some == 1 and idx == [5]

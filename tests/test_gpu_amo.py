"""Atomic memory operations on symmetric words through the C ABI:

* every operation of the spec's tables against a Python model on 4- and
  8-byte types (two's-complement wrap-around, float/double bit patterns
  including -0.0 and a NaN payload), the _nbi forms delivering to host
  memory and to a symmetric address, the deprecated names;
* errors: a PE outside the world, a non-symmetric or misaligned target;
* SPMD mode: PE threads fetch_inc one counter (the fetched values are a
  permutation, the final count exact), add into one word, and guard a
  non-atomic read-modify-write with a compare_swap lock."""
import ctypes as C
import struct

import numpy as np
import pytest

import osc

pytestmark = pytest.mark.gpu

VP, I = C.c_void_p, C.c_int


def fn(name, res, args):
    f = getattr(osc.lib(), name)
    f.restype, f.argtypes = res, args
    return f


def amo_fns(tn, ct):
    F = {}
    F["fetch"] = fn(f"shmem_{tn}_atomic_fetch", ct, [VP, I])
    F["set"] = fn(f"shmem_{tn}_atomic_set", None, [VP, ct, I])
    F["swap"] = fn(f"shmem_{tn}_atomic_swap", ct, [VP, ct, I])
    F["fetch_nbi"] = fn(f"shmem_{tn}_atomic_fetch_nbi", None, [VP, VP, I])
    F["swap_nbi"] = fn(f"shmem_{tn}_atomic_swap_nbi", None, [VP, VP, ct, I])
    if ct in (C.c_float, C.c_double):
        return F
    F["cswap"] = fn(f"shmem_{tn}_atomic_compare_swap", ct, [VP, ct, ct, I])
    F["finc"] = fn(f"shmem_{tn}_atomic_fetch_inc", ct, [VP, I])
    F["inc"] = fn(f"shmem_{tn}_atomic_inc", None, [VP, I])
    F["fadd"] = fn(f"shmem_{tn}_atomic_fetch_add", ct, [VP, ct, I])
    F["add"] = fn(f"shmem_{tn}_atomic_add", None, [VP, ct, I])
    F["cswap_nbi"] = fn(f"shmem_{tn}_atomic_compare_swap_nbi", None,
                        [VP, VP, ct, ct, I])
    F["finc_nbi"] = fn(f"shmem_{tn}_atomic_fetch_inc_nbi", None, [VP, VP, I])
    F["fadd_nbi"] = fn(f"shmem_{tn}_atomic_fetch_add_nbi", None,
                       [VP, VP, ct, I])
    return F


def bit_fns(tn, ct):
    F = {}
    for op in ("and", "or", "xor"):
        F["f" + op] = fn(f"shmem_{tn}_atomic_fetch_{op}", ct, [VP, ct, I])
        F[op] = fn(f"shmem_{tn}_atomic_{op}", None, [VP, ct, I])
        F["f" + op + "_nbi"] = fn(f"shmem_{tn}_atomic_fetch_{op}_nbi", None,
                                  [VP, VP, ct, I])
    F["xor_nbi"] = fn(f"shmem_{tn}_atomic_xor_nbi", None, [VP, VP, ct, I])
    return F


def wrap(v, bits, signed):
    v &= (1 << bits) - 1
    if signed and v >> (bits - 1):
        v -= 1 << bits
    return v


INT_CASES = [("int", C.c_int, 32, True), ("long", C.c_long, 64, True),
             ("uint32", C.c_uint32, 32, False),
             ("uint64", C.c_uint64, 64, False),
             ("size", C.c_size_t, 64, False),
             ("ptrdiff", C.c_ssize_t, 64, True)]


@pytest.fixture
def heap():
    if osc.Heap._current is not None:
        osc.Heap._current.close()
    with osc.Heap(npes=4, heap_bytes=1 << 20) as h:
        yield h


@pytest.mark.parametrize("tn,ct,bits,signed", INT_CASES)
def test_integer_operations_against_a_model(heap, tn, ct, bits, signed):
    h, pe = heap, 2
    F = amo_fns(tn, ct)
    w = osc.lib().shmem_calloc(8, 8)       # zero on every PE
    sym_fetch = h.malloc(64)
    host = (ct * 1)()
    top = (1 << (bits - 1)) - 1 if signed else (1 << bits) - 1
    model = top                       # next add wraps around
    F["set"](w, model, pe)
    assert F["fetch"](w, pe) == model
    assert F["finc"](w, pe) == model
    model = wrap(model + 1, bits, signed)
    assert F["fetch"](w, pe) == model
    F["inc"](w, pe)
    model = wrap(model + 1, bits, signed)
    assert F["fadd"](w, wrap(-5, bits, signed), pe) == model
    model = wrap(model - 5, bits, signed)
    F["add"](w, 12, pe)
    model = wrap(model + 12, bits, signed)
    assert F["swap"](w, 77, pe) == model
    model = 77
    assert F["cswap"](w, 76, 1, pe) == 77          # no match: unchanged
    assert F["cswap"](w, 77, 9, pe) == 77          # match: now 9
    model = 9
    F["fetch_nbi"](C.addressof(host), w, pe)
    assert host[0] == model
    F["swap_nbi"](C.addressof(host), w, 3, pe)
    assert host[0] == model
    model = 3
    F["cswap_nbi"](sym_fetch, w, 3, 4, pe)         # fetch into PE 0's copy
    got = h.get(sym_fetch, 1, np.dtype(ct), 0)[0]
    assert got == 3
    model = 4
    F["finc_nbi"](C.addressof(host), w, pe)
    assert host[0] == model
    model += 1
    F["fadd_nbi"](C.addressof(host), w, 10, pe)
    assert host[0] == model
    model += 10
    # the word on the other PEs is untouched, and it is one word only
    assert h.get(w, 2, np.dtype(ct), pe)[1] == 0
    for other in (0, 1, 3):
        assert h.get(w, 1, np.dtype(ct), other)[0] == 0
    assert h.get(w, 1, np.dtype(ct), pe)[0] == model


@pytest.mark.parametrize("tn,ct,bits", [("uint", C.c_uint, 32),
                                        ("int64", C.c_int64, 64),
                                        ("ulonglong", C.c_ulonglong, 64)])
def test_bitwise_operations_against_a_model(heap, tn, ct, bits):
    h, pe = heap, 1
    A, F = amo_fns(tn, ct), bit_fns(tn, ct)
    rng = np.random.default_rng(bits)
    w = h.malloc(64)
    host = (ct * 1)()
    signed = tn == "int64"
    model = wrap(int(rng.integers(0, 2 ** 63)) * 2 + 1, bits, signed)
    A["set"](w, model, pe)
    for _ in range(4):
        for op, f in (("and", lambda a, b: a & b), ("or", lambda a, b: a | b),
                      ("xor", lambda a, b: a ^ b)):
            v = wrap(int(rng.integers(0, 2 ** 63)) * 3, bits, signed)
            assert F["f" + op](w, v, pe) == model
            model = wrap(f(model, v), bits, signed)
            v = wrap(int(rng.integers(0, 2 ** 63)), bits, signed)
            F[op](w, v, pe)
            model = wrap(f(model, v), bits, signed)
            v = wrap(int(rng.integers(0, 2 ** 63)), bits, signed)
            F["f" + op + "_nbi"](C.addressof(host), w, v, pe)
            assert host[0] == model
            model = wrap(f(model, v), bits, signed)
    F["xor_nbi"](C.addressof(host), w, 1, pe)
    assert host[0] == model
    model = wrap(model ^ 1, bits, signed)
    assert A["fetch"](w, pe) == model


@pytest.mark.parametrize("tn,ct,fmt", [("float", C.c_float, "<f"),
                                       ("double", C.c_double, "<d")])
def test_float_fetch_set_swap_keep_bits(heap, tn, ct, fmt):
    h, pe = heap, 3
    F = amo_fns(tn, ct)
    w = h.malloc(64)
    ubits = "<I" if fmt == "<f" else "<Q"
    nan = struct.unpack(fmt, struct.pack(ubits, (0x7fc0beef if fmt == "<f"
                                                  else 0x7ff800000000beef)))[0]
    F["set"](w, -0.0, pe)
    got = F["fetch"](w, pe)
    assert got == 0.0 and struct.pack(fmt, got) == struct.pack(fmt, -0.0)
    assert struct.pack(fmt, F["swap"](w, 1.5, pe)) == struct.pack(fmt, -0.0)
    assert F["fetch"](w, pe) == 1.5
    # a NaN payload through the device word, read back as raw bits
    F["set"](w, nan, pe)
    raw = h.get(w, 1, np.uint32 if fmt == "<f" else np.uint64, pe)[0]
    assert int(raw) == struct.unpack(ubits, struct.pack(fmt, nan))[0]
    host = (ct * 1)()
    F["swap_nbi"](C.addressof(host), w, 2.25, pe)
    assert struct.pack(fmt, host[0]) == struct.pack(fmt, nan)
    F["fetch_nbi"](C.addressof(host), w, pe)
    assert host[0] == 2.25


def test_deprecated_names(heap):
    h, pe = heap, 0
    w = h.malloc(64)
    L = osc.lib()
    fn("shmem_long_set", None, [VP, C.c_long, I])(w, 40, pe)
    assert fn("shmem_long_finc", C.c_long, [VP, I])(w, pe) == 40
    fn("shmem_long_add", None, [VP, C.c_long, I])(w, 2, pe)
    assert fn("shmem_long_fadd", C.c_long, [VP, C.c_long, I])(w, 1, pe) == 43
    assert fn("shmem_long_cswap", C.c_long, [VP, C.c_long, C.c_long, I])(
        w, 44, 7, pe) == 44
    fn("shmem_long_inc", None, [VP, I])(w, pe)
    assert fn("shmem_long_fetch", C.c_long, [VP, I])(w, pe) == 8
    assert fn("shmem_long_swap", C.c_long, [VP, C.c_long, I])(w, 0, pe) == 8
    assert fn("shmem_double_swap", C.c_double, [VP, C.c_double, I])(
        w, 0.5, pe) == 0.0
    assert fn("shmem_double_fetch", C.c_double, [VP, I])(w, pe) == 0.5
    L.osc_clear_error()


def test_errors_leave_words_alone(heap):
    h = heap
    F = amo_fns("long", C.c_long)
    L = osc.lib()
    w = h.malloc(64)
    F["set"](w, 5, 0)
    L.osc_clear_error()
    assert F["finc"](w, 4) == 0 and L.osc_last_error_code() != 0   # no PE 4
    L.osc_clear_error()
    host = (C.c_long * 2)()
    assert F["fadd"](C.addressof(host), 1, 0) == 0                 # host word
    assert L.osc_last_error_code() != 0
    L.osc_clear_error()
    F["add"](w + 4, 1, 0)                                          # misaligned
    assert L.osc_last_error_code() != 0
    L.osc_clear_error()
    assert F["fetch"](w, 0) == 5


def test_negative_pe_is_an_error(heap):
    """A negative PE number records an error and moves nothing, for the
    atomics and for every RMA form on the default context."""
    h = heap
    F = amo_fns("long", C.c_long)
    L = osc.lib()
    w, src, sig = h.malloc(64), h.malloc(64), h.malloc(8)
    F["set"](w, 5, 1)
    h.put(src, np.full(8, 9, np.int64), 0)
    calls = [
        lambda: F["finc"](w, -1),
        lambda: F["add"](w, 1, -1),
        lambda: F["fetch_nbi"](w, src, -1),
        lambda: fn("shmem_putmem", None, [VP, VP, C.c_size_t, I])(w, src, 8, -1),
        lambda: fn("shmem_getmem", None, [VP, VP, C.c_size_t, I])(src, w, 8, -1),
        lambda: fn("shmem_long_p", None, [VP, C.c_long, I])(w, 3, -1),
        lambda: fn("shmem_putmem_signal", None,
                   [VP, VP, C.c_size_t, VP, C.c_uint64, I, I])(w, src, 8, sig, 1, 0, -1),
        lambda: fn("shmem_long_iput", None,
                   [VP, VP, C.c_ssize_t, C.c_ssize_t, C.c_size_t, I])(w, src, 1, 1, 2, -1),
    ]
    for k, call in enumerate(calls):
        L.osc_clear_error()
        call()
        assert L.osc_last_error_code() != 0, k
    L.osc_clear_error()
    assert F["fetch"](w, 1) == 5
    assert h.get(src, 8, np.int64, 0).tolist() == [9] * 8


def test_size_overflow_is_an_error(heap):
    """An element count whose byte count overflows size_t (2^61 + 1 longs
    would wrap to 8 bytes) records an error and moves nothing."""
    h = heap
    L = osc.lib()
    w, src, sig = h.malloc(64), h.malloc(64), h.malloc(8)
    for pe in range(h.npes):
        h.put(w, np.full(8, 4, np.int64), pe)
    h.put(src, np.full(8, 9, np.int64), 0)
    big = (1 << 61) + 1
    calls = [
        lambda: osc.typed("shmem_long_put")(w, src, big, 1),
        lambda: osc.typed("shmem_long_get")(w, src, big, 1),
        lambda: fn("shmem_put64_nbi", None, [VP, VP, C.c_size_t, I])(w, src, big, 1),
        lambda: osc.typed("shmem_long_put_signal")(w, src, big, sig, 1, 0, 1),
        lambda: osc.typed("shmem_long_fcollect")(h.world, w, src, big),
    ]
    for k, call in enumerate(calls):
        L.osc_clear_error()
        call()
        assert L.osc_last_error_code() != 0, k
    L.osc_clear_error()
    for pe in range(h.npes):
        assert h.get(w, 8, np.int64, pe).tolist() == [4] * 8, pe


PE_FN = C.CFUNCTYPE(C.c_int, C.c_int, C.c_void_p)


def _spmd(npes, body):
    if osc.Heap._current is not None:
        osc.Heap._current.close()
    L = osc.lib()
    L.osc_spmd_run.restype = C.c_int
    L.osc_spmd_run.argtypes = [C.c_int, C.c_size_t, C.c_int, PE_FN, VP]
    errors = []

    def run(pe, _arg):
        try:
            return body(pe)
        except Exception as e:  # reported by the test, not lost in a thread
            errors.append((pe, repr(e)))
            return 1
    cb = PE_FN(run)
    rc = L.osc_spmd_run(npes, 1 << 20, 0, cb, None)
    assert not errors, errors
    return rc


def test_spmd_fetch_inc_hands_out_every_ticket_once():
    npes, per = 8, 100
    F = amo_fns("long", C.c_long)
    L = osc.lib()
    got, final = {}, {}

    def body(pe):
        ctr = L.shmem_calloc(2, 8)
        L.shmem_barrier_all()
        got[pe] = [F["finc"](ctr, 0) for _ in range(per)]
        F["add"](ctr + 8, pe + 1, 0)                 # one add per PE
        L.shmem_barrier_all()
        final[pe] = (F["fetch"](ctr, 0), F["fetch"](ctr + 8, 0))
        L.shmem_barrier_all()
        L.shmem_free(ctr)
        return 0
    assert _spmd(npes, body) == 0
    tickets = sorted(t for pe in got for t in got[pe])
    assert tickets == list(range(npes * per))
    for pe in range(npes):
        assert got[pe] == sorted(got[pe])           # each PE's own order
        assert final[pe] == (npes * per, npes * (npes + 1) // 2)


def test_spmd_compare_swap_lock_guards_a_plain_update():
    npes, per = 6, 20
    F = amo_fns("int", C.c_int)
    L = osc.lib()
    L.shmem_int_g.restype, L.shmem_int_g.argtypes = C.c_int, [VP, I]
    L.shmem_int_p.restype, L.shmem_int_p.argtypes = None, [VP, C.c_int, I]
    out = {}

    def body(pe):
        lock = L.shmem_calloc(2, 4)       # [lock word, shared counter]
        L.shmem_barrier_all()
        for _ in range(per):
            while F["cswap"](lock, 0, pe + 1, 0) != 0:
                pass
            v = L.shmem_int_g(lock + 4, 0)          # not atomic on its own
            L.shmem_int_p(lock + 4, v + 1, 0)
            assert F["swap"](lock, 0, 0) == pe + 1  # release: it was ours
        L.shmem_barrier_all()
        out[pe] = L.shmem_int_g(lock + 4, 0)
        L.shmem_barrier_all()
        L.shmem_free(lock)
        return 0
    assert _spmd(npes, body) == 0
    assert all(v == npes * per for v in out.values()), out


def test_spmd_locks_guard_a_plain_update_and_test_lock_reports_holders():
    npes, per = 5, 15
    L = osc.lib()
    L.shmem_set_lock.restype, L.shmem_set_lock.argtypes = None, [VP]
    L.shmem_clear_lock.restype, L.shmem_clear_lock.argtypes = None, [VP]
    L.shmem_test_lock.restype, L.shmem_test_lock.argtypes = I, [VP]
    L.shmem_long_g.restype, L.shmem_long_g.argtypes = C.c_long, [VP, I]
    L.shmem_long_p.restype, L.shmem_long_p.argtypes = None, [VP, C.c_long, I]
    out = {}

    def body(pe):
        lock = L.shmem_calloc(1, 8)
        ctr = L.shmem_calloc(1, 8)
        L.shmem_barrier_all()
        for _ in range(per):
            L.shmem_set_lock(lock)
            v = L.shmem_long_g(ctr, 0)
            L.shmem_long_p(ctr, v + 1, 0)
            L.shmem_clear_lock(lock)
        L.shmem_barrier_all()
        held = None
        if pe == 0:                        # PE 0 takes it; the others try
            assert L.shmem_test_lock(lock) == 0
        L.shmem_barrier_all()
        if pe != 0:
            held = L.shmem_test_lock(lock)
        L.shmem_barrier_all()
        if pe == 0:
            L.shmem_clear_lock(lock)
        L.shmem_barrier_all()
        out[pe] = (L.shmem_long_g(ctr, 0), held)
        L.shmem_barrier_all()
        L.shmem_free(ctr)
        L.shmem_free(lock)
        return 0
    assert _spmd(npes, body) == 0
    for pe in range(npes):
        assert out[pe][0] == npes * per
        assert out[pe][1] == (None if pe == 0 else 1)

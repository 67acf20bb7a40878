"""ctypes binding of libosc.so, the MI355X OpenSHMEM collective library.

This is the Python-side binding of the C ABI declared in include/shmem.h
and include/osc.h: the OpenSHMEM typed entry points (shmem_<T>_<op>_reduce,
shmem_<T>_fcollect, shmem_<T>_alltoall, shmem_<T>_broadcast, the *mem
forms, heap and team calls) plus the osc_* launch layer.  It loads the
in-tree build (osss-ucx_amd/lib/libosc.so) and raises if it is missing: there
is no fallback implementation.

All device work happens inside the library; numpy arrays are only used to
move data between the host and one PE's copy of a symmetric object.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libosc.so")

# numbering of include/osc.h
OPS = {"and": 0, "or": 1, "xor": 2, "min": 3, "max": 4, "sum": 5, "prod": 6}
DTYPES = {"u8": 0, "u16": 1, "u32": 2, "u64": 3, "i8": 4, "i16": 5, "i32": 6,
          "i64": 7, "f32": 8, "f64": 9, "c32": 10, "c64": 11}
NP_DTYPE = {"u8": np.uint8, "u16": np.uint16, "u32": np.uint32,
            "u64": np.uint64, "i8": np.int8, "i16": np.int16,
            "i32": np.int32, "i64": np.int64, "f32": np.float32,
            "f64": np.float64, "c32": np.complex64, "c64": np.complex128}
DIST = {"bits": 0, "int_range": 1, "set_m1_1_2": 2, "normal": 3,
        "real_range": 4, "a2a_pattern": 5, "fcol_pattern": 6}
FP_ACCURATE, FP_REFERENCE = 0, 1
FP_LINEAR, FP_BINOMIAL, FP_RABENSEIFNER = 2, 3, 4

# OpenSHMEM typename -> kernel dtype (include/shmem.h type tables, LP64)
TYPENAME_DTYPE = {
    "char": "i8", "schar": "i8", "short": "i16", "int": "i32", "long": "i64",
    "longlong": "i64", "ptrdiff": "i64", "uchar": "u8", "ushort": "u16",
    "uint": "u32", "ulong": "u64", "ulonglong": "u64", "int8": "i8",
    "int16": "i16", "int32": "i32", "int64": "i64", "uint8": "u8",
    "uint16": "u16", "uint32": "u32", "uint64": "u64", "size": "u64",
    "float": "f32", "double": "f64", "complexf": "c32", "complexd": "c64",
}

ERRORS = {0: "OK", -1: "ARG", -2: "NOTINIT", -3: "TEAM", -4: "PTR",
          -5: "NOMEM", -6: "HIP", -7: "UNSUPPORTED", -8: "ALGO",
          -9: "BUDGET", -10: "STATE"}


class OscError(RuntimeError):
    def __init__(self, code: int, what: str, msg: str):
        super().__init__(f"{what}: {ERRORS.get(code, code)} ({code}): {msg}")
        self.code = code


CALL_REDUCE, CALL_FCOLLECT, CALL_ALLTOALL, CALL_BROADCAST = 0, 1, 2, 3


class Call(C.Structure):
    """osc_call of include/osc.h."""
    _fields_ = [("kind", C.c_int32), ("op", C.c_int32), ("dtype", C.c_int32),
                ("root", C.c_int32), ("team_start", C.c_int32),
                ("team_stride", C.c_int32), ("team_npes", C.c_int32),
                ("wave", C.c_int32), ("count", C.c_uint64),
                ("src_off", C.c_uint64), ("dest_off", C.c_uint64)]


class BatchInfo(C.Structure):
    _fields_ = [("ncalls", C.c_int), ("nwaves", C.c_int),
                ("nlaunches", C.c_int), ("ntiles", C.c_uint64),
                ("algorithmic_bytes", C.c_double)]


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """Load libosc.so (once).  Raises if the build is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} not found: build it first (make -C osss-ucx_amd)")
    L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    vp, sz, i, u64, d = C.c_void_p, C.c_size_t, C.c_int, C.c_uint64, C.c_double
    pd = C.c_ssize_t  # ptrdiff_t (LP64)
    sig = {
        "osc_init": (i, [i, sz, i]),
        "osc_finalize": (i, []),
        "osc_initialized": (i, []),
        "osc_npes": (i, []),
        "osc_slice_bytes": (sz, []),
        "osc_heap_bytes": (sz, []),
        "osc_heap_base": (vp, []),
        "osc_default_stream": (vp, []),
        "osc_set_fp_mode": (i, [i]),
        "osc_set_reduce_shape": (i, [i]),
        "osc_get_reduce_shape": (i, []),
        "osc_set_batch_branches": (i, [i]),
        "osc_set_batch_phases": (i, [i]),
        "osc_set_reduce_split": (i, [i]),
        "osc_set_move_interleave": (i, [i]),
        "osc_reduce_split_group": (i, [i, sz, C.POINTER(C.c_int)]),
        "osc_device_xor_read": (i, [vp, sz, vp, vp]),
        "osc_device_fill64": (i, [vp, sz, u64, vp]),
        "osc_last_reduce_kernels": (i, [C.c_char_p, sz]),
        "osc_get_fp_mode": (i, []),
        "osc_strerror": (C.c_char_p, [i]),
        "osc_last_error": (C.c_char_p, []),
        "osc_set_my_pe": (i, [i]),
        "osc_sym_offset": (i, [vp, C.POINTER(i), C.POINTER(sz)]),
        "osc_pe_ptr": (vp, [vp, i]),
        "osc_put_host": (i, [vp, vp, sz, i]),
        "osc_get_host": (i, [vp, vp, sz, i]),
        "osc_reduce": (i, [vp, i, i, vp, vp, sz, vp]),
        "osc_fcollect": (i, [vp, vp, vp, sz, vp]),
        "osc_alltoall": (i, [vp, vp, vp, sz, vp]),
        "osc_broadcast": (i, [vp, vp, vp, sz, i, vp]),
        "osc_stream_sync": (i, [vp]),
        "osc_fill": (i, [vp, vp, sz, i, i, d, d, u64, vp]),
        "osc_event_create": (i, [C.POINTER(vp)]),
        "osc_event_destroy": (i, [vp]),
        "osc_event_record": (i, [vp, vp]),
        "osc_event_elapsed_ms": (i, [vp, vp, C.POINTER(C.c_float)]),
        "osc_parse_size": (i, [C.c_char_p, C.POINTER(sz)]),
        "osc_algo_resolve": (i, [C.c_char_p, C.c_char_p, C.POINTER(i)]),
        "osc_device_copy": (i, [vp, vp, sz, vp]),
        "osc_batch_plan": (i, [C.POINTER(Call), i, u64, u64, C.POINTER(i)]),
        "osc_batch_plan_world": (i, [C.POINTER(Call), i, i, u64, u64,
                                     C.POINTER(i)]),
        "osc_batch_create": (i, [C.POINTER(Call), i, C.POINTER(vp)]),
        "osc_batch_launch": (i, [vp, vp]),
        "osc_batch_launch_serial": (i, [vp, vp]),
        "osc_batch_digests": (i, [vp, C.POINTER(C.c_uint64), vp]),
        "osc_batch_get_info": (i, [vp, C.POINTER(BatchInfo)]),
        "osc_batch_destroy": (i, [vp]),
        "osc_tune_load": (i, [C.c_char_p]),
        "osc_tune_save": (i, [C.c_char_p]),
        "osc_tune_set": (i, [i, i, i, i, i]),
        "osc_tune_get": (i, [i, i, C.POINTER(i), C.POINTER(i), C.POINTER(i)]),
        "shmem_init": (None, []),
        "shmem_finalize": (None, []),
        "shmem_my_pe": (i, []),
        "shmem_n_pes": (i, []),
        "shmem_malloc": (vp, [sz]),
        "shmem_calloc": (vp, [sz, sz]),
        "shmem_align": (vp, [sz, sz]),
        "shmem_realloc": (vp, [vp, sz]),
        "shmem_free": (None, [vp]),
        "shmem_ptr": (vp, [vp, i]),
        "shmem_addr_accessible": (i, [vp, i]),
        "shmem_barrier_all": (None, []),
        "shmem_team_my_pe": (i, [vp]),
        "shmem_team_n_pes": (i, [vp]),
        "shmem_team_translate_pe": (i, [vp, i, vp]),
        "shmem_team_split_strided": (i, [vp, i, i, i, vp, C.c_long,
                                         C.POINTER(vp)]),
        "shmem_team_split_2d": (i, [vp, i, vp, C.c_long, C.POINTER(vp), vp,
                                    C.c_long, C.POINTER(vp)]),
        "shmem_team_destroy": (None, [vp]),
        "shmem_team_sync": (i, [vp]),
        "shmem_putmem": (None, [vp, vp, sz, i]),
        "shmem_getmem": (None, [vp, vp, sz, i]),
        "shmem_putmem_nbi": (None, [vp, vp, sz, i]),
        "shmem_getmem_nbi": (None, [vp, vp, sz, i]),
        "shmem_fcollectmem": (i, [vp, vp, vp, sz]),
        "shmem_alltoallmem": (i, [vp, vp, vp, sz]),
        "shmem_broadcastmem": (i, [vp, vp, vp, sz, i]),
        "osc_last_error_code": (i, []),
        "osc_clear_error": (None, []),
        "osc_collect": (i, [vp, vp, vp, C.POINTER(sz), vp]),
        "osc_collect_noalloc": (i, [vp, vp, vp, C.POINTER(sz), vp]),
        "osc_alltoalls": (i, [vp, vp, vp, pd, pd, sz, sz, vp]),
        "shmem_collectmem": (i, [vp, vp, vp, sz]),
        "shmem_alltoallsmem": (i, [vp, vp, vp, pd, pd, sz]),
        "shmem_putmem_signal": (None, [vp, vp, sz, vp, u64, i, i]),
        "shmem_putmem_signal_nbi": (None, [vp, vp, sz, vp, u64, i, i]),
        "shmem_signal_fetch": (u64, [vp]),
        "shmem_signal_wait_until": (u64, [vp, i, u64]),
        "shmem_quiet": (None, []),
        "shmem_fence": (None, []),
        "shmem_sync_all": (None, []),
        "shmem_barrier": (None, [i, i, i, vp]),
        "shmem_sync_deprecated": (None, [i, i, i, vp]),
    }
    for size in (8, 16, 32, 64, 128):
        for base, args in (("put%d", [vp, vp, sz, i]),
                           ("put%d_nbi", [vp, vp, sz, i]),
                           ("get%d", [vp, vp, sz, i]),
                           ("get%d_nbi", [vp, vp, sz, i]),
                           ("iput%d", [vp, vp, pd, pd, sz, i]),
                           ("iget%d", [vp, vp, pd, pd, sz, i]),
                           ("put%d_signal", [vp, vp, sz, vp, u64, i, i]),
                           ("put%d_signal_nbi", [vp, vp, sz, vp, u64, i, i])):
            sig["shmem_" + base % size] = (None, args)
    for size in (32, 64):
        sig[f"shmem_fcollect{size}"] = (None, [vp, vp, sz, i, i, i, vp])
        sig[f"shmem_collect{size}"] = (None, [vp, vp, sz, i, i, i, vp])
        sig[f"shmem_alltoall{size}"] = (None, [vp, vp, sz, i, i, i, vp])
        sig[f"shmem_alltoalls{size}"] = (None, [vp, vp, pd, pd, sz, i, i, i,
                                                vp])
        sig[f"shmem_broadcast{size}"] = (None, [vp, vp, sz, i, i, i, i, vp])
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def team_world() -> int:
    return C.c_void_p.in_dll(lib(), "SHMEM_TEAM_WORLD").value


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise OscError(rc, what, lib().osc_last_error().decode())


# C types of the typed entry points' element arguments (shmem_<T>_p/_g)
CTYPE = {"float": C.c_float, "double": C.c_double, "char": C.c_char,
         "schar": C.c_byte, "short": C.c_short, "int": C.c_int,
         "long": C.c_long, "longlong": C.c_longlong, "uchar": C.c_ubyte,
         "ushort": C.c_ushort, "uint": C.c_uint, "ulong": C.c_ulong,
         "ulonglong": C.c_ulonglong, "int8": C.c_int8, "int16": C.c_int16,
         "int32": C.c_int32, "int64": C.c_int64, "uint8": C.c_uint8,
         "uint16": C.c_uint16, "uint32": C.c_uint32, "uint64": C.c_uint64,
         "size": C.c_size_t, "ptrdiff": C.c_ssize_t}


def typed(name: str) -> C._CFuncPtr:
    """An OpenSHMEM typed entry point, e.g. typed('shmem_int_sum_reduce'),
    typed('shmem_long_put'), typed('shmem_int_max_to_all')."""
    fn = getattr(lib(), name)
    vp, sz, i, pd = C.c_void_p, C.c_size_t, C.c_int, C.c_ssize_t
    tn = name[len("shmem_"):].split("_")[0]
    fn.restype = C.c_int
    if name.endswith("_broadcast"):
        fn.argtypes = [vp, vp, vp, sz, i]
    elif name.endswith("_alltoalls"):
        fn.argtypes = [vp, vp, vp, pd, pd, sz]
    elif name.endswith("_to_all"):
        fn.restype = None
        fn.argtypes = [vp, vp, i, i, i, i, vp, vp]
    elif name.endswith(("_put", "_get", "_put_nbi", "_get_nbi")):
        fn.restype = None
        fn.argtypes = [vp, vp, sz, i]
    elif name.endswith(("_iput", "_iget")):
        fn.restype = None
        fn.argtypes = [vp, vp, pd, pd, sz, i]
    elif name.endswith(("_put_signal", "_put_signal_nbi")):
        fn.restype = None
        fn.argtypes = [vp, vp, sz, vp, C.c_uint64, i, i]
    elif name.endswith("_p"):
        fn.restype = None
        fn.argtypes = [vp, CTYPE[tn], i]
    elif name.endswith("_g"):
        fn.restype = CTYPE[tn]
        fn.argtypes = [vp, i]
    else:
        fn.argtypes = [vp, vp, vp, sz]
    return fn


class Heap:
    """Convenience wrapper: an initialised library with npes emulated PEs."""

    _current = None  # the Heap whose osc_init is live

    def __init__(self, npes: int, heap_bytes: int, device: int = 0):
        L = lib()
        if L.osc_initialized():
            L.osc_finalize()
        check(L.osc_init(npes, heap_bytes, device), "osc_init")
        self.L = L
        self.npes = npes
        self.heap_bytes = heap_bytes
        self.world = team_world()
        Heap._current = self

    @property
    def alive(self) -> bool:
        return Heap._current is self and bool(self.L.osc_initialized())

    def close(self) -> None:
        if Heap._current is self:
            self.L.osc_finalize()
            Heap._current = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def malloc(self, nbytes: int) -> int:
        p = self.L.shmem_malloc(nbytes)
        if not p:
            raise OscError(-5, "shmem_malloc",
                           self.L.osc_last_error().decode())
        return p

    def free(self, p: int) -> None:
        self.L.shmem_free(p)

    def put(self, sym: int, arr: np.ndarray, pe: int) -> None:
        a = np.ascontiguousarray(arr)
        check(self.L.osc_put_host(sym, a.ctypes.data, a.nbytes, pe),
              "osc_put_host")

    def get(self, sym: int, count: int, dtype, pe: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        check(self.L.osc_get_host(out.ctypes.data, sym, out.nbytes, pe),
              "osc_get_host")
        return out

    def put_all(self, sym: int, arrs, team=None, start=0, stride=1) -> None:
        for k, a in enumerate(arrs):
            self.put(sym, a, start + k * stride)

    def get_all(self, sym: int, count: int, dtype, pes) -> np.ndarray:
        return np.stack([self.get(sym, count, dtype, pe) for pe in pes])

    def fill(self, team, sym, nelems, dtype: str, dist: str, a=0.0, b=0.0,
             seed=0) -> None:
        check(self.L.osc_fill(team or self.world, sym, nelems, DTYPES[dtype],
                              DIST[dist], float(a), float(b), seed, None),
              "osc_fill")

    def sync(self) -> None:
        check(self.L.osc_stream_sync(None), "osc_stream_sync")

    def split(self, start: int, stride: int, size: int, parent=None) -> int:
        out = C.c_void_p()
        rc = self.L.shmem_team_split_strided(parent or self.world, start,
                                             stride, size, None, 0,
                                             C.byref(out))
        check(rc, "shmem_team_split_strided")
        return out.value


class Timer:
    """HIP-event timing on the library stream (or a given stream)."""

    def __init__(self, stream=None):
        L = lib()
        self.L = L
        self.stream = stream
        self.a, self.b = C.c_void_p(), C.c_void_p()
        check(L.osc_event_create(C.byref(self.a)), "event")
        check(L.osc_event_create(C.byref(self.b)), "event")

    def start(self):
        check(self.L.osc_event_record(self.a, self.stream), "record")

    def stop(self):
        check(self.L.osc_event_record(self.b, self.stream), "record")

    def ms(self) -> float:
        self.L.osc_stream_sync(self.stream)
        out = C.c_float()
        check(self.L.osc_event_elapsed_ms(self.a, self.b, C.byref(out)),
              "elapsed")
        return out.value

    def close(self):
        self.L.osc_event_destroy(self.a)
        self.L.osc_event_destroy(self.b)


class Batch:
    """A created osc_batch over a list of Call records."""

    def __init__(self, calls):
        L = lib()
        self.L = L
        self.n = len(calls)
        self.calls = (Call * self.n)(*calls)
        h = C.c_void_p()
        check(L.osc_batch_create(self.calls, self.n, C.byref(h)),
              "osc_batch_create")
        self.h = h

    def launch(self, stream=None):
        check(self.L.osc_batch_launch(self.h, stream), "osc_batch_launch")

    def launch_serial(self, stream=None):
        check(self.L.osc_batch_launch_serial(self.h, stream),
              "osc_batch_launch_serial")

    def digests(self, stream=None) -> np.ndarray:
        out = np.zeros(self.n, np.uint64)
        check(self.L.osc_batch_digests(
            self.h, out.ctypes.data_as(C.POINTER(C.c_uint64)), stream),
            "osc_batch_digests")
        return out

    def info(self) -> BatchInfo:
        inf = BatchInfo()
        check(self.L.osc_batch_get_info(self.h, C.byref(inf)), "info")
        return inf

    def close(self):
        if self.h:
            self.L.osc_batch_destroy(self.h)
            self.h = None


def plan(calls, dest_base, dest_bytes, world_npes=None):
    """osc_batch_plan (or _world): fills dest_off / wave in place."""
    arr = (Call * len(calls))(*calls)
    nw = C.c_int()
    L = lib()
    if world_npes is None:
        rc = L.osc_batch_plan(arr, len(calls), dest_base, dest_bytes,
                              C.byref(nw))
    else:
        rc = L.osc_batch_plan_world(arr, len(calls), world_npes, dest_base,
                                    dest_bytes, C.byref(nw))
    check(rc, "osc_batch_plan")
    return list(arr), nw.value


SERVE_BATCHED, SERVE_PER_CALL = 0, 1
ORDER_ARRIVAL, ORDER_SIZE = 0, 1


class Request(C.Structure):
    """osc_request of include/osc.h."""
    _fields_ = [("call", Call), ("arrival_s", C.c_double),
                ("admit_s", C.c_double), ("launch_s", C.c_double),
                ("done_s", C.c_double), ("digest", C.c_uint64),
                ("group", C.c_int32), ("status", C.c_int32)]


class ServeCfg(C.Structure):
    _fields_ = [("mode", C.c_int32), ("max_inflight", C.c_int32),
                ("nstreams", C.c_int32), ("max_group", C.c_int32),
                ("dest_base", C.c_uint64), ("dest_bytes", C.c_uint64),
                ("page_bytes", C.c_uint64), ("budget_bytes", C.c_double),
                ("digests", C.c_int32), ("order", C.c_int32),
                ("age_bytes_per_s", C.c_double)]


class ServeStats(C.Structure):
    _fields_ = [("completed", C.c_int32), ("rejected", C.c_int32),
                ("groups", C.c_int32), ("max_inflight_seen", C.c_int32),
                ("wall_s", C.c_double), ("mean_ms", C.c_double),
                ("p50_ms", C.c_double), ("p90_ms", C.c_double),
                ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("algorithmic_bytes", C.c_double), ("gbps", C.c_double),
                ("peak_inflight_bytes", C.c_double)]


def serve(reqs, cfg):
    """osc_serve_run over a list of Request; returns (requests, stats)."""
    L = lib()
    fn = L.osc_serve_run
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(Request), C.c_int, C.POINTER(ServeCfg),
                   C.POINTER(ServeStats)]
    arr = (Request * len(reqs))(*reqs)
    st = ServeStats()
    check(fn(arr, len(reqs), C.byref(cfg), C.byref(st)), "osc_serve_run")
    return list(arr), st


def request_alone(call, dest_off):
    """osc_request_alone: run one call by itself at dest_off; its digest."""
    L = lib()
    fn = L.osc_request_alone
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(Call), C.c_uint64, C.POINTER(C.c_uint64)]
    d = C.c_uint64()
    check(fn(C.byref(call), dest_off, C.byref(d)), "osc_request_alone")
    return d.value

"""The C-ABI library loads and exports every function include/*.h declares,
and the pure host logic (size parsing, algorithm registry, argument checks
that run before any HIP call) behaves.  No GPU needed.
"""
import ctypes as C
import os
import re
import subprocess

import pytest

import osc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    """Function names declared by include/*.h after macro expansion."""
    names = set()
    for h in sorted(f for f in os.listdir(INCLUDE) if f.endswith(".h")):
        out = subprocess.run(["gcc", "-E", "-P", "-I", INCLUDE,
                              os.path.join(INCLUDE, h)],
                             check=True, capture_output=True, text=True).stdout
        for decl in out.split(";"):
            decl = " ".join(decl.split())
            if "typedef" in decl or "(" not in decl or "{" in decl:
                continue
            m = re.search(r"[\s\*](\w+)\s*\(", decl)
            if m and not m.group(1).startswith("_"):
                names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = osc.lib()
    names = declared_functions()
    assert len(names) > 250, len(names)  # the typed tables expanded
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing[:20]


def test_team_handles_exported():
    lib = osc.lib()
    for n in ("SHMEM_TEAM_WORLD", "SHMEM_TEAM_SHARED", "SHMEM_TEAM_INVALID"):
        C.c_void_p.in_dll(lib, n)


def test_typed_reduce_symbol_set_matches_spec_tables():
    """Spot-check the cross product: bitwise ops only on integer types,
    min/max also on float/double, sum/prod also on complex."""
    lib = osc.lib()
    assert hasattr(lib, "shmem_uint64_xor_reduce")
    assert hasattr(lib, "shmem_float_max_reduce")
    assert hasattr(lib, "shmem_complexd_prod_reduce")
    assert hasattr(lib, "shmem_int_sum_reduce")
    names = declared_functions()
    assert "shmem_float_xor_reduce" not in names
    assert "shmem_complexf_max_reduce" not in names


@pytest.mark.parametrize("text,val", [("32M", 32 << 20), ("1G", 1 << 30),
                                      ("4096", 4096), ("512k", 512 << 10),
                                      ("2GB", 2 << 30), ("1.5M", 3 << 19)])
def test_parse_size(text, val):
    out = C.c_size_t()
    assert osc.lib().osc_parse_size(text.encode(), C.byref(out)) == 0
    assert out.value == val


@pytest.mark.parametrize("text", ["", "abc", "12Q", "-5", "3MX"])
def test_parse_size_rejects(text):
    out = C.c_size_t()
    assert osc.lib().osc_parse_size(text.encode(), C.byref(out)) != 0


@pytest.mark.parametrize("kind,spec,ok,mode", [
    ("reduce", "hip", True, 0),
    ("reduce", "rec_dbl", True, 1),
    ("reduce", "rec_dbl:float", True, 1),
    ("reduce", "hip:int", True, 0),
    ("reduce", "bogus", False, None),
    ("reduce", "rec_dbl:notatype", False, None),
    ("reduce", "linear", True, 2),
    ("reduce", "binomial:double", True, 3),
    ("reduce", "rabenseifner", True, 4),
    ("reduce", "rabenseifner2:long", True, 4),
    ("reduce", "rabenseifner3", False, None),
    ("fcollect", "bruck_inplace", True, None),
    ("fcollect", "hip", True, None),
    ("alltoall", "shift_exchange_barrier", True, None),
    ("alltoall", "nope", False, None),
    ("broadcast", "binomial_tree", True, None),
])
def test_algo_resolve(kind, spec, ok, mode):
    m = C.c_int(-1)
    rc = osc.lib().osc_algo_resolve(kind.encode(), spec.encode(), C.byref(m))
    assert (rc == 0) == ok
    if ok and mode is not None:
        assert m.value == mode


def test_init_argument_errors_before_any_device_call():
    lib = osc.lib()
    assert lib.osc_init(0, 1 << 20, 0) == -1          # OSC_ERR_ARG
    assert lib.osc_init(100000, 1 << 20, 0) == -1
    assert lib.osc_init(4, 0, 0) == -1
    assert lib.osc_initialized() == 0


def test_bad_algorithm_env_is_an_error(monkeypatch):
    monkeypatch.setenv("SHMEM_SUM_REDUCE_ALGO", "bogus")
    lib = osc.lib()
    assert lib.osc_init(4, 1 << 20, 0) == -8          # OSC_ERR_ALGO
    assert b"SHMEM_SUM_REDUCE_ALGO" in lib.osc_last_error()
    assert lib.osc_initialized() == 0


@pytest.mark.parametrize("var,val", [
    ("SHMEM_SUM_TO_ALL_ALGO", "bogus"), ("SHMEM_COLLECT_ALGO", "nope"),
    ("SHMEM_ALLTOALLS_ALGO", "bruck2"), ("SHMEM_BARRIER_ALGO", "fast"),
    ("SHMEM_FCOLLECT_SIZE_ALGO", "x")])
def test_other_algorithm_variables_are_checked(monkeypatch, var, val):
    """The reference's other per-collective variables (src/shmemc/env.c:
    187-281) are read too: an unknown algorithm name fails init."""
    monkeypatch.setenv(var, val)
    lib = osc.lib()
    assert lib.osc_init(4, 1 << 20, 0) == -8          # OSC_ERR_ALGO
    assert var.encode() in lib.osc_last_error()
    assert lib.osc_initialized() == 0


def test_calls_before_init_fail_cleanly():
    lib = osc.lib()
    assert lib.osc_reduce(None, 5, 6, None, None, 10, None) == -2
    assert lib.shmem_malloc(100) is None
    assert lib.shmem_my_pe() == -1
    f = osc.typed("shmem_int_sum_reduce")
    assert f(osc.team_world(), None, None, 1) != 0
    g = osc.typed("shmem_longdouble_sum_reduce")
    assert g(osc.team_world(), None, None, 1) != 0


def test_strerror_names_every_code():
    lib = osc.lib()
    for code in range(0, -11, -1):
        assert lib.osc_strerror(code) != b"unknown error"


def test_wtime_is_monotonic_seconds():
    lib = osc.lib()
    lib.shmemx_wtime.restype = C.c_double
    a = lib.shmemx_wtime()
    b = lib.shmemx_wtime()
    assert 0 < a <= b < a + 1.0


def test_version_queries():
    lib = osc.lib()
    v = C.c_int()
    lib.shmem_info_get_version_number(C.byref(v))
    assert v.value == 105
    a, b, c = C.c_int(), C.c_int(), C.c_int()
    lib.shmem_info_get_version(C.byref(a), C.byref(b))
    assert (a.value, b.value) == (1, 5)
    lib.shmem_info_get_vendor_version(C.byref(a), C.byref(b), C.byref(c))
    lib.shmem_info_get_vendor_version_number(C.byref(v))
    assert v.value == a.value * 10000 + b.value * 100 + c.value

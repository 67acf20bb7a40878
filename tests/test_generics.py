"""The C11 type-generic routines of include/shmem.h (shmem_sum_reduce,
shmem_fcollect, shmem_put, shmem_sum_to_all, ...), on the CPU:

* tests/c/generics.c uses every generic with element types of its table
  and is built with -Werror, so a selection whose parameter types do not
  match its arguments fails the build; it is linked against libosc.so and
  run to compare the selected routines by address (canonical C types and
  the <stdint.h>/<stddef.h> typedefs) -- no routine is called;
* element types outside a routine's table must not compile."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "osss-ucx_amd", "lib")


def _gcc(src, out, extra=()):
    return subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror",
                           "-I", INC, src, "-o", out, *extra],
                          capture_output=True, text=True)


def test_generic_selection(tmp_path):
    exe = str(tmp_path / "generics")
    r = _gcc(os.path.join(ROOT, "tests", "c", "generics.c"), exe,
             ["-L", LIBDIR, "-losc", "-Wl,-rpath," + LIBDIR])
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "generics ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("stmt", [
    "shmem_and_reduce(t, (float *)0, (const float *)0, 4);",
    "shmem_xor_reduce(t, (char *)0, (const char *)0, 4);",
    "shmem_min_reduce(t, (float _Complex *)0, (const float _Complex *)0, 4);",
    "shmem_sum_to_all((unsigned *)0, (const unsigned *)0, 4, 0, 0, 2,"
    " (unsigned *)0, ps);",
    "shmem_and_to_all((float *)0, (const float *)0, 4, 0, 0, 2,"
    " (float *)0, ps);",
    "shmem_fcollect(t, (void *)0, (const void *)0, 4);",
    "shmem_sum_reduce(t, (long *)0, (const int *)0, 4);",
    "shmem_test_any((short *)0, 4, 0, SHMEM_CMP_EQ, (short)0);",
    "shmem_wait_until((double *)0, SHMEM_CMP_EQ, 0.0);",
    "shmem_atomic_fetch_add((float *)0, 1.0f, 0);",
    "shmem_atomic_fetch_and((short *)0, (short)1, 0);",
    "shmem_atomic_inc((double *)0, 0);",
])
def test_types_outside_the_tables_do_not_compile(tmp_path, stmt):
    src = tmp_path / "bad.c"
    src.write_text("#include <shmem.h>\n"
                   "void f(shmem_team_t t, long *ps) { %s }\n" % stmt)
    r = _gcc(str(src), str(tmp_path / "bad.o"), ["-c"])
    assert r.returncode != 0, stmt


def test_header_still_compiles_as_cxx_and_c99(tmp_path):
    """The generics are C11-only; C++ and C99 translation units see the
    typed routines alone."""
    src = tmp_path / "plain.c"
    src.write_text("#include <shmem.h>\n"
                   "int f(shmem_team_t t, long *d, const long *s)"
                   " { return shmem_long_sum_reduce(t, d, s, 4); }\n")
    for cmd in (["gcc", "-std=c99", "-Wall", "-Werror", "-x", "c"],
                ["g++", "-std=c++17", "-Wall", "-Werror", "-x", "c++"]):
        r = subprocess.run(cmd + ["-I", INC, "-c", str(src), "-o",
                                  str(tmp_path / "plain.o")],
                           capture_output=True, text=True)
        assert r.returncode == 0, (cmd, r.stderr)

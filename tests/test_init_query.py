"""Setup and query routines around the collectives: shmem_init_thread,
shmem_query_thread, shmem_initialized/finalized, shmem_pe_accessible,
shmem_global_exit, shmem_malloc_with_hints, shmem_team_ptr.

The CPU tests touch no device: the state before any init, a rejected
thread level, and global_exit's status and flushed output from a small C
program.  The GPU tests start the library through shmem_init_thread."""
import ctypes as C
import os
import subprocess

import pytest

import osc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "osss-ucx_amd", "lib")
SINGLE, FUNNELED, SERIALIZED, MULTIPLE = 0, 1, 2, 3


def _lib():
    L = osc.lib()
    for name, res, args in [
            ("shmem_init_thread", C.c_int, [C.c_int, C.POINTER(C.c_int)]),
            ("shmem_query_thread", None, [C.POINTER(C.c_int)]),
            ("shmem_initialized", C.c_int, []),
            ("shmem_finalized", C.c_int, []),
            ("shmem_pe_accessible", C.c_int, [C.c_int]),
            ("shmem_malloc_with_hints", C.c_void_p, [C.c_size_t, C.c_long]),
            ("shmem_team_ptr", C.c_void_p, [C.c_void_p, C.c_void_p, C.c_int]),
            ("shmem_ptr", C.c_void_p, [C.c_void_p, C.c_int])]:
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    return L


def test_state_before_init_and_bad_thread_level():
    L = _lib()
    if L.osc_initialized():
        pytest.skip("a library instance is live in this process")
    assert L.shmem_initialized() == 0
    assert L.shmem_pe_accessible(0) == 0
    got = C.c_int(-1)
    L.shmem_query_thread(C.byref(got))
    assert got.value == MULTIPLE
    assert L.shmem_init_thread(7, C.byref(got)) != 0   # no such level
    assert L.shmem_init_thread(-1, None) != 0
    assert L.shmem_initialized() == 0


def test_global_exit_flushes_and_ends_with_status(tmp_path):
    src = tmp_path / "gexit.c"
    src.write_text('#include <shmem.h>\n#include <stdio.h>\n'
                   'int main(void) { printf("before exit"); '
                   'shmem_global_exit(3); printf("not reached"); return 0; }\n')
    exe = str(tmp_path / "gexit")
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", INC,
                        str(src), "-o", exe, "-L", LIBDIR, "-losc",
                        "-Wl,-rpath," + LIBDIR], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 3
    assert r.stdout == "before exit"


@pytest.mark.gpu
def test_init_thread_query_and_pointers(monkeypatch):
    if osc.Heap._current is not None:
        osc.Heap._current.close()
    monkeypatch.setenv("OSC_NPES", "6")
    monkeypatch.setenv("SHMEM_SYMMETRIC_SIZE", "4M")
    L = _lib()
    got = C.c_int(-1)
    assert L.shmem_init_thread(SERIALIZED, C.byref(got)) == 0
    try:
        assert got.value == SERIALIZED
        assert L.shmem_initialized() == 1 and L.shmem_finalized() == 0
        assert L.shmem_n_pes() == 6
        assert [L.shmem_pe_accessible(p) for p in (-1, 0, 5, 6)] == [0, 1, 1, 0]
        p = L.shmem_malloc_with_hints(4096, 1)   # SHMEM_MALLOC_SIGNAL_REMOTE
        assert p
        # team index -> global PE: the odd PEs 1, 3, 5
        odd = C.c_void_p()
        assert L.shmem_team_split_strided(osc.team_world(), 1, 2, 3, None, 0,
                                          C.byref(odd)) == 0
        for k in range(3):
            assert L.shmem_team_ptr(odd.value, p, k) == L.shmem_ptr(p, 1 + 2 * k)
        assert L.shmem_team_ptr(odd.value, p, 3) is None
        assert L.shmem_team_ptr(None, p, 0) is None
        L.shmem_team_destroy(odd.value)
        L.shmem_free(p)
    finally:
        L.shmem_finalize()
    assert L.shmem_initialized() == 0 and L.shmem_finalized() == 1


def test_heap_names_and_interoperability_without_a_device():
    """Names get indices in first-lookup order (the table is process-wide,
    so the test works relative to what earlier lookups took); only index 0
    allocates; no interoperability is offered."""
    L = osc.lib()
    L.shmemx_name_to_index.restype = C.c_int
    L.shmemx_name_to_index.argtypes = [C.c_char_p]
    L.shmemx_index_to_name.restype = C.c_char_p
    L.shmemx_index_to_name.argtypes = [C.c_int]
    L.shmemx_malloc_by_index.restype = C.c_void_p
    L.shmemx_malloc_by_index.argtypes = [C.c_int, C.c_size_t]
    a = L.shmemx_name_to_index(b"osc-test-heap-a")
    b = L.shmemx_name_to_index(b"osc-test-heap-b")
    assert a >= 0 and b == a + 1
    assert L.shmemx_name_to_index(b"osc-test-heap-a") == a
    assert L.shmemx_index_to_name(b) == b"osc-test-heap-b"
    assert L.shmemx_index_to_name(10 ** 6) is None
    L.osc_clear_error()
    assert L.shmemx_malloc_by_index(1, 64) is None      # one heap: index 0
    assert L.osc_last_error_code() != 0
    L.osc_clear_error()
    assert all(L.shmemx_query_interoperability(p) == 0 for p in range(4))


@pytest.mark.gpu
def test_heap_index_zero_and_completion_tests(tmp_path):
    """A C11 program: shmemx_malloc with index 0 and with the name that
    owns index 0, the C11 generic forms, quiet/fence tests after a put."""
    src = tmp_path / "xheap.c"
    src.write_text(r'''
#include <shmemx.h>
#include <stdio.h>
int main(void) {
  long v[4] = {1, 2, 3, 4}, w[4] = {0, 0, 0, 0};
  int bad = 0;
  shmem_init();
  const char *first = shmemx_index_to_name(0);
  shmemx_heap_index_t i0 = first ? 0 : shmemx_name_to_index("main");
  long *a = (long *)shmemx_malloc(i0, 4 * sizeof(long));
  long *b = (long *)shmemx_malloc(shmemx_index_to_name(0), 4 * sizeof(long));
  long *c = (long *)shmemx_calloc_by_index(0, 4, sizeof(long));
  if (!a || !b || !c) bad |= 1;
  shmem_putmem(a, v, sizeof v, 1);
  if (!shmemx_fence_test()) bad |= 2;
  shmem_quiet();
  if (!shmemx_quiet_test()) bad |= 4;
  shmem_getmem(w, a, sizeof w, 1);
  if (w[3] != 4) bad |= 8;
  if (shmemx_lookup_remote_addr(a, 1) != shmem_ptr(a, 1)) bad |= 16;
  if (shmemx_malloc_by_index(1, 8) != NULL) bad |= 32;
  shmemx_free(i0, c);
  shmemx_free(shmemx_index_to_name(0), b);
  shmemx_free_by_index(0, a);
  shmem_finalize();
  printf("bad=%d\n", bad);
  return bad;
}
''')
    exe = str(tmp_path / "xheap")
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", INC,
                        str(src), "-o", exe, "-L", LIBDIR, "-losc",
                        "-Wl,-rpath," + LIBDIR], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, OSC_NPES="2", SHMEM_SYMMETRIC_SIZE="4M")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120,
                       env=env)
    assert r.returncode == 0, r.stdout + r.stderr

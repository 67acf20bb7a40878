"""SPMD calling mode on the GPU: one host thread per emulated PE, every PE
calls each collective with its own arguments (the reference's execution
model), and the last member to arrive launches once for the team.

* An SPMD program written the OpenSHMEM way (osss-ucx_amd/examples/
  spmd_collectives.c: reduce, fcollect, variable-size collect, alltoall,
  broadcast, team split, an active-set form, a put ring and a
  put-with-signal ring), run under the SPMD main() at several PE counts.
* From Python through osc_spmd_run: per-PE variable collect sizes against
  the oracle's Bruck restatement, 2-d team splits, and members that
  disagree on a call's arguments all getting an error instead of a launch.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import osc
import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "osss-ucx_amd", "bin", "spmd_collectives")


@pytest.mark.parametrize("npes", [1, 3, 8])
def test_spmd_program(npes):
    env = dict(os.environ, OSC_NPES=str(npes), SHMEM_SYMMETRIC_SIZE="8M")
    print(f"spmd_collectives with {npes} PEs", flush=True)
    r = subprocess.run([EXE], env=env, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = sorted(ln for ln in r.stdout.splitlines() if ln.startswith("PE "))
    assert len(lines) == npes, r.stdout
    assert all(ln.endswith("PASS") for ln in lines), r.stdout


PE_FN = C.CFUNCTYPE(C.c_int, C.c_int, C.c_void_p)


def _spmd(npes, fn, heap=4 << 20):
    """Run fn(pe) on npes PE threads; returns osc_spmd_run's code."""
    if osc.Heap._current is not None:
        osc.Heap._current.close()
    L = osc.lib()
    L.osc_spmd_run.restype = C.c_int
    L.osc_spmd_run.argtypes = [C.c_int, C.c_size_t, C.c_int, PE_FN, C.c_void_p]
    errors = []

    def body(pe, _arg):
        try:
            return fn(pe)
        except Exception as e:  # reported by the test, not lost in a thread
            errors.append((pe, repr(e)))
            return 1
    cb = PE_FN(body)
    rc = L.osc_spmd_run(npes, heap, 0, cb, None)
    assert not errors, errors
    return rc


def test_spmd_variable_collect_matches_oracle():
    npes = 6
    rng = np.random.default_rng(11)
    sizes = [int(x) for x in rng.integers(0, 40, npes)]
    sizes[2] = 0
    rows = [rng.integers(-2 ** 62, 2 ** 62, k, dtype=np.int64) for k in sizes]
    want = O.collect([r.view(np.uint8) for r in rows])[0].view(np.int64)
    results = {}
    L = osc.lib()
    coll = osc.typed("shmem_int64_collect")

    def pe_main(pe):
        src = L.shmem_malloc(8 * 64)
        dst = L.shmem_malloc(8 * 64 * npes)
        if sizes[pe]:
            L.shmem_putmem(src, rows[pe].ctypes.data, 8 * sizes[pe], pe)
        rc = coll(osc.team_world(), dst, src, sizes[pe])
        out = np.zeros(sum(sizes), np.int64)
        L.shmem_getmem(out.ctypes.data, dst, out.nbytes, pe)
        results[pe] = (rc, out)
        L.shmem_free(dst)
        L.shmem_free(src)
        return 0
    assert _spmd(npes, pe_main) == 0
    for pe in range(npes):
        rc, out = results[pe]
        assert rc == 0
        assert np.array_equal(out, want), pe


def test_spmd_disagreement_fails_every_member():
    npes = 4
    codes = {}
    L = osc.lib()
    red = osc.typed("shmem_long_sum_reduce")

    def pe_main(pe):
        src = L.shmem_malloc(8 * 32)
        dst = L.shmem_malloc(8 * 32)
        n = 16 if pe == 2 else 32            # PE 2 disagrees on nreduce
        codes[pe] = red(osc.team_world(), dst, src, n)
        ok = red(osc.team_world(), dst, src, 32)   # the next call works
        codes[pe] = (codes[pe], ok)
        L.shmem_free(dst)
        L.shmem_free(src)
        return 0
    assert _spmd(npes, pe_main) == 0
    assert all(codes[pe] == (-1, 0) for pe in range(npes)), codes


def test_spmd_split_2d_and_team_reduce():
    npes, xr = 6, 3
    got = {}
    L = osc.lib()
    red = osc.typed("shmem_int_sum_reduce")

    def pe_main(pe):
        xt, yt = C.c_void_p(), C.c_void_p()
        rc = L.shmem_team_split_2d(osc.team_world(), xr, None, 0,
                                   C.byref(xt), None, 0, C.byref(yt))
        src = L.shmem_malloc(4 * 8)
        dst = L.shmem_malloc(4 * 8)
        v = np.full(8, pe, np.int32)
        L.shmem_putmem(src, v.ctypes.data, v.nbytes, pe)
        rx = red(xt.value, dst, src, 8)
        ox = np.zeros(8, np.int32)
        L.shmem_getmem(ox.ctypes.data, dst, ox.nbytes, pe)
        ry = red(yt.value, dst, src, 8)
        oy = np.zeros(8, np.int32)
        L.shmem_getmem(oy.ctypes.data, dst, oy.nbytes, pe)
        got[pe] = (rc, rx, ry, int(ox[0]), int(oy[0]),
                   L.shmem_team_my_pe(xt.value), L.shmem_team_my_pe(yt.value))
        L.shmem_team_destroy(xt.value)
        L.shmem_team_destroy(yt.value)
        L.shmem_free(dst)
        L.shmem_free(src)
        return 0
    assert _spmd(npes, pe_main) == 0
    for pe in range(npes):
        row, col = pe // xr, pe % xr
        row_sum = sum(range(row * xr, row * xr + xr))
        col_sum = sum(range(col, npes, xr))
        assert got[pe] == (0, 0, 0, row_sum, col_sum, col, row), (pe, got[pe])


def test_spmd_split_2d_bad_xrange_fails_every_member():
    """xrange < 1 is an argument error on every member (no team is made)
    and the members go on to the next call."""
    npes = 4
    got = {}
    L = osc.lib()

    def pe_main(pe):
        xt, yt = C.c_void_p(1), C.c_void_p(1)
        rc = L.shmem_team_split_2d(osc.team_world(), 0, None, 0,
                                   C.byref(xt), None, 0, C.byref(yt))
        x2, y2 = C.c_void_p(), C.c_void_p()
        ok = L.shmem_team_split_2d(osc.team_world(), 2, None, 0,
                                   C.byref(x2), None, 0, C.byref(y2))
        got[pe] = (rc, xt.value, yt.value, ok)
        L.shmem_team_destroy(x2.value)
        L.shmem_team_destroy(y2.value)
        return 0
    assert _spmd(npes, pe_main) == 0
    for pe in range(npes):
        rc, xt, yt, ok = got[pe]
        assert rc != 0 and not xt and not yt and ok == 0, (pe, got[pe])

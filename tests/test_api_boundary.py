"""The drop-in boundary, pinned to the reference's own headers.

tests/golden/reference_api.json (written by tests/golden/
make_api_fixture.py from the reference's include/shmem/api.h and
include/pshmem.h) records every routine the reference declares, with its
types.  include/shmem.h must declare exactly that set with identical
return and parameter types, libosc.so must export all of it, and the
profiling names ("p" + each routine, reference --enable-pshmem) must be
declared in include/pshmem.h and exported as aliases of the routines.
No GPU needed."""
import ctypes as C
import json
import os
import subprocess

import pytest

import c_protos as cp
import gen_pshmem_h
import osc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
FIX = json.load(open(os.path.join(ROOT, "tests", "golden",
                                  "reference_api.json")))


def ours(defines=()):
    return cp.prototypes(cp.preprocess(os.path.join(INC, "shmem.h"), [INC],
                                       defines))


@pytest.mark.parametrize("key,defines", [("api", ()),
                                         ("api_pr463", ("PR463",))])
def test_header_declares_the_reference_api_exactly(key, defines):
    ref, mine = FIX[key], ours(defines)
    assert len(ref) > 1600
    missing = sorted(set(ref) - set(mine))
    extra = sorted(set(mine) - set(ref))
    differ = sorted((n, ref[n], mine[n]) for n in ref
                    if n in mine and ref[n] != mine[n])
    assert not missing, missing
    assert not extra, extra
    assert not differ, differ[:10]


def _addr(L, name):
    return C.cast(getattr(L, name), C.c_void_p).value


def test_library_exports_every_reference_routine_and_its_profiling_name():
    L = osc.lib()
    for name in sorted(FIX["api_pr463"]):
        a = _addr(L, name)
        assert a, name
        # the profiling name is the same code (an alias, not a wrapper)
        assert _addr(L, "p" + name) == a, name


def test_profiling_header_is_current_and_matches_the_routines():
    with open(os.path.join(INC, "pshmem.h")) as f:
        assert f.read() == gen_pshmem_h.render(), \
            "include/pshmem.h is stale: run osss-ucx_amd/gen_pshmem_h.py"
    for key, defines in (("api", ()), ("api_pr463", ("PR463",))):
        prof = cp.prototypes(cp.preprocess(os.path.join(INC, "pshmem.h"),
                                           [INC], defines), r"^p")
        prof = {n: t for n, t in prof.items() if n[1:] in FIX[key]}
        assert sorted(n[1:] for n in prof) == sorted(FIX[key])
        assert all(prof["p" + n] == FIX[key][n] for n in FIX[key])


def test_reference_profiling_declarations_agree():
    """The reference's own include/pshmem.h declares 120 routines by hand;
    those whose routine api.h declares have the same types here (the rest,
    pshmem_ctx_int_add and its kin, name routines the reference defines
    nowhere)."""
    prof = cp.prototypes(cp.preprocess(os.path.join(INC, "pshmem.h"), [INC]),
                         r"^p")
    known = {n: t for n, t in FIX["pshmem"].items() if n[1:] in FIX["api"]}
    assert len(known) >= 90
    assert {n: prof.get(n) for n in known} == known


def test_a_profiling_tool_interposes_through_the_p_names(tmp_path):
    """A tool defines shmem_info_get_version, counts the call and forwards
    to pshmem_info_get_version; the program's call reaches the tool first,
    then the library (no device needed for this query)."""
    src = tmp_path / "tool.c"
    src.write_text(r'''
#include <stdio.h>
#include <pshmem.h>
static int calls;
void shmem_info_get_version(int *major, int *minor) {
  ++calls;
  pshmem_info_get_version(major, minor);
}
int main(void) {
  int ma = 0, mi = 0;
  shmem_info_get_version(&ma, &mi);
  printf("%d %d.%d\n", calls, ma, mi);
  return 0;
}
''')
    lib = os.path.join(ROOT, "osss-ucx_amd", "lib")
    exe = str(tmp_path / "tool")
    r = subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I", INC,
                        str(src), "-L", lib, "-losc", "-Wl,-rpath," + lib,
                        "-Wl,-rpath-link,/opt/rocm/lib", "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.split() == ["1", "1.5"]

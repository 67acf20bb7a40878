"""Writes tests/golden/reference_api.json: every function prototype the
reference declares in its public headers, as data.

  python tests/golden/make_api_fixture.py [/root/reference]

* include/shmem/api.h is preprocessed with gcc -E (it includes only
  headers that exist in the tree: defs.h, depr.h, teams.h, api_types.h,
  generics.h, cache.h), once as is and once with -DPR463 (api.h:321-331
  declares three version queries under that macro).
* include/pshmem.h includes the configure-generated shmem/defs_subst.h, so
  it is read as text instead: comments and preprocessor lines removed and
  its hand-written declarations parsed.  22 of them (the pshmem_ctx_<T>_add
  family and its kin) name routines the reference defines nowhere; the
  profiling names the library provides are "p" + every api.h routine.

Only names and types are recorded (parameter names dropped), e.g.
"shmem_int_sum_reduce": "int(shmem_team_t, int *, const int *, size_t)".
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "osss-ucx_amd"))
import c_protos as cp  # noqa: E402


def pshmem_text(path):
    t = open(path).read()
    t = re.sub(r"/\*.*?\*/", " ", t, flags=re.S)
    t = re.sub(r"//[^\n]*", " ", t)
    # drop preprocessor lines, including continued #define bodies
    lines, cont = [], False
    for ln in t.splitlines():
        s = ln.strip()
        if cont or s.startswith("#"):
            cont = s.endswith("\\")
            lines.append("")
            continue
        lines.append(ln)
    return "\n".join(lines)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    inc = os.path.join(ref, "include")
    api = os.path.join(inc, "shmem", "api.h")
    out = {
        "source": "include/shmem/api.h and include/pshmem.h of "
                  "BTMichalowicz/osss-ucx",
        "api": cp.prototypes(cp.preprocess(api, [inc])),
        "api_pr463": cp.prototypes(cp.preprocess(api, [inc], ["PR463"])),
    }
    out["pshmem"] = cp.prototypes(pshmem_text(os.path.join(inc, "pshmem.h")),
                                  r"^p")
    dst = os.path.join(HERE, "reference_api.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0, sort_keys=True)
        f.write("\n")
    print(dst, {k: len(v) for k, v in out.items() if k != "source"})


if __name__ == "__main__":
    main()

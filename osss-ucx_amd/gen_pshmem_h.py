#!/usr/bin/env python3
"""Writes include/pshmem.h: the profiling interface's names, "p" + every
routine declared in include/shmem.h (pshmem_int_sum_reduce,
pshmalloc, pstart_pes, p_my_pe, ...), with the same types.

The library exports each of them as an alias of the routine itself
(Makefile: pshmem.ld), as the reference's --enable-pshmem build does with
"#pragma weak shmem_X = pshmem_X" (src/api/collectives/shcoll-shim.c:
146-177, config/m4/confargs.m4:49-59).  A profiling tool defines shmem_X,
does its own work, and calls pshmem_X to reach the library.

    python osss-ucx_amd/gen_pshmem_h.py      # run after changing shmem.h
"""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
INC = os.path.join(HERE, "..", "include")
sys.path.insert(0, HERE)
import c_protos as cp  # noqa: E402


def decl(name, proto):
    ret, params = proto.split("(", 1)
    return "%s p%s(%s;" % (ret, name, params)


def render():
    hdr = os.path.join(INC, "shmem.h")
    base = cp.prototypes(cp.preprocess(hdr, [INC]))
    extra = cp.prototypes(cp.preprocess(hdr, [INC], ["PR463"]))
    lines = [
        "/*",
        " * pshmem.h - the OpenSHMEM profiling interface of this library: for",
        " * every routine of shmem.h the same routine under its \"p\" name",
        " * (pshmem_X for shmem_X, pshmalloc, pstart_pes, p_my_pe, ...).  The",
        " * library exports each p name as an alias of the routine, so a",
        " * profiling tool can define shmem_X itself and call pshmem_X to reach",
        " * the library (reference --enable-pshmem, include/pshmem.h).",
        " *",
        " * Generated by osss-ucx_amd/gen_pshmem_h.py from shmem.h; do not edit.",
        " */",
        "#ifndef OSC_PSHMEM_H",
        "#define OSC_PSHMEM_H 1",
        "",
        "#include \"shmem.h\"",
        "",
        "#ifdef __cplusplus",
        "extern \"C\" {",
        "#endif",
        "",
    ]
    lines += [decl(n, base[n]) for n in sorted(base)]
    pr = sorted(set(extra) - set(base))
    if pr:
        lines += ["", "#ifdef PR463"] + [decl(n, extra[n]) for n in pr] + \
                 ["#endif"]
    lines += ["", "#ifdef __cplusplus", "}", "#endif", "",
              "#endif /* OSC_PSHMEM_H */", ""]
    return "\n".join(lines)


def main():
    out = os.path.join(INC, "pshmem.h")
    text = render()
    with open(out, "w") as f:
        f.write(text)
    print(out, text.count(";") )


if __name__ == "__main__":
    main()

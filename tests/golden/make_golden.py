"""Writes tests/golden/golden.json: fixed vectors for the seeded generator
and the reference reduce on the 8-PE int sum workload, produced by the CPU
oracle (oracle/liborc.so) and the numpy generator, which tests/ check
against each other and against the GPU.

The reference ships no golden vectors of its own (src/shcoll/tests holds
only timing/debug headers), so these fixtures pin the oracle against
regressions; its correctness is pinned independently in tests/test_oracle.py.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402


def main():
    out = {}
    out["hash3"] = [[s, p, i, int(O.lib().orc_hash3(s, p, i))]
                    for s in (0, 1, 12345) for p in (0, 5, 255)
                    for i in (0, 1, 65535, 4194303)]
    # plain int sum: 8 PEs x 65536 int32 uniform in [-1000, 1000], seed 0
    src = O.gen_int_range(0, range(8), 65536, -1000, 1000, "i32")
    dest = O.rec_dbl("sum", "i32", src)
    out["int_sum_8pe"] = {
        "npes": 8, "n": 65536, "seed": 0, "lo": -1000, "hi": 1000,
        "src_sha256": hashlib.sha256(src.tobytes()).hexdigest(),
        "dest_sha256": hashlib.sha256(dest.tobytes()).hexdigest(),
        "dest_pe0_head": dest[0, :16].tolist(),
        "src_pe0_head": src[0, :16].tolist(),
    }
    # reference-order float32 sums for non-power-of-two teams, inputs from
    # the integer generator (exactly reproducible): x = k / 2^20, k uniform
    # in [-2^23, 2^23]
    small = {}
    for P in (3, 5, 6, 7, 12):
        k = O.gen_int_range(9, range(P), 64, -2 ** 23, 2 ** 23, "i64")
        x = (k.astype(np.float64) / 2 ** 20).astype(np.float32)
        y = O.rec_dbl("sum", "f32", x)
        small[str(P)] = {
            "src_u32": x.view(np.uint32).tolist(),
            "dest_u32": y.view(np.uint32).tolist(),
        }
    out["rec_dbl_f32_sum"] = small
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=None, separators=(",", ":"))
        f.write("\n")


if __name__ == "__main__":
    main()

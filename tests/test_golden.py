"""The committed fixtures in tests/golden/golden.json agree with the oracle
and the numpy generator today (CPU only)."""
import hashlib
import json
import os

import numpy as np

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    with open(os.path.join(HERE, "golden", "golden.json")) as f:
        return json.load(f)


def test_hash_fixture():
    for seed, pe, idx, val in load()["hash3"]:
        assert int(O.lib().orc_hash3(seed, pe, idx)) == val
        assert int(O.hash3(seed, pe, idx)) == val


def test_int_sum_8pe_fixture():
    g = load()["int_sum_8pe"]
    src = O.gen_int_range(g["seed"], range(g["npes"]), g["n"], g["lo"],
                          g["hi"], "i32")
    assert src.min() >= -1000 and src.max() <= 1000
    assert src[0, :16].tolist() == g["src_pe0_head"]
    assert hashlib.sha256(src.tobytes()).hexdigest() == g["src_sha256"]
    dest = O.rec_dbl("sum", "i32", src)
    assert np.array_equal(dest[0], src.sum(axis=0, dtype=np.int32))
    assert hashlib.sha256(dest.tobytes()).hexdigest() == g["dest_sha256"]


def test_small_float_fixture_matches_closed_form():
    for P, v in load()["rec_dbl_f32_sum"].items():
        x = np.array(v["src_u32"], dtype=np.uint32).view(np.float32)
        y = np.array(v["dest_u32"], dtype=np.uint32)
        assert np.array_equal(O.rec_dbl("sum", "f32", x).view(np.uint32), y)
        assert np.array_equal(
            O.reduce_closed_form("sum", x).view(np.uint32), y)

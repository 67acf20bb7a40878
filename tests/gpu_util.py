"""Helpers for the GPU parity tests: run one collective through the C ABI of
libosc on given per-PE inputs and read every member's result back."""
import ctypes as C

import numpy as np

import osc
import oracle as O

# kernel dtype -> an OpenSHMEM typename with that element type
TYPENAME = {"u8": "uint8", "u16": "uint16", "u32": "uint32", "u64": "uint64",
            "i8": "int8", "i16": "int16", "i32": "int32", "i64": "int64",
            "f32": "float", "f64": "double", "c32": "complexf",
            "c64": "complexd"}
UINT_VIEW = {"f32": np.uint32, "f64": np.uint64, "c32": np.uint32,
             "c64": np.uint64}


def inputs(rng, op, dtype, P, n):
    """Per-PE inputs [P, n] of the kinds the reduce workloads use."""
    npt = O.NP_DTYPE[dtype]
    if dtype in ("f32", "f64"):
        if op == "prod":
            return rng.uniform(0.999, 1.001, (P, n)).astype(npt)
        return rng.standard_normal((P, n)).astype(npt)
    if dtype in ("c32", "c64"):
        if op == "prod":
            re = rng.uniform(0.999, 1.001, (P, n))
            im = rng.uniform(-0.001, 0.001, (P, n))
        else:
            re, im = rng.standard_normal((P, n)), rng.standard_normal((P, n))
        return (re + 1j * im).astype(npt)
    info = np.iinfo(npt)
    if op == "prod":
        return rng.choice(np.array([-1, 1, 2]) if info.min < 0 else
                          np.array([1, 2, 3]), size=(P, n)).astype(npt)
    return rng.integers(info.min, info.max, (P, n), endpoint=True, dtype=npt)


def run_reduce(h, op, dtype, src, team=None, pes=None, mode=0,
               inplace=False, misalign=0, dest_init=None):
    """Reduce src[P, n] over `team` (members = global PEs `pes`) through
    shmem_<T>_<op>_reduce; returns dest[P, n] read back from every member."""
    P, n = src.shape
    team = team or h.world
    pes = list(range(P)) if pes is None else list(pes)
    es = src.dtype.itemsize
    nbytes = n * es
    sbuf = h.malloc(nbytes + 64)
    dbuf = sbuf if inplace else h.malloc(nbytes + 64)
    s = sbuf + misalign * es
    d = s if inplace else dbuf + misalign * es
    try:
        for k, pe in enumerate(pes):
            h.put(s, src[k], pe)
            if not inplace and dest_init is not None:
                h.put(d, dest_init[k], pe)
        osc.check(h.L.osc_set_fp_mode(mode), "fp mode")
        fn = osc.typed(f"shmem_{TYPENAME[dtype]}_{op}_reduce")
        rc = fn(team, d, s, n)
        osc.check(rc, f"shmem_{TYPENAME[dtype]}_{op}_reduce")
        out = np.stack([h.get(d, n, src.dtype, pe) for pe in pes])
    finally:
        h.L.osc_set_fp_mode(0)
        h.free(sbuf)
        if not inplace:
            h.free(dbuf)
    return out


def bit_equal(a, b, dtype):
    """Bitwise equality, NaN payloads excepted (NaN-ness must agree)."""
    if dtype in UINT_VIEW:
        fa = a.view(a.real.dtype) if np.iscomplexobj(a) else a
        fb = b.view(b.real.dtype) if np.iscomplexobj(b) else b
        na, nb = np.isnan(fa), np.isnan(fb)
        if not np.array_equal(na, nb):
            return False
        ua, ub = fa.view(UINT_VIEW[dtype]), fb.view(UINT_VIEW[dtype])
        return bool(np.array_equal(ua[~na], ub[~nb]))
    return bool(np.array_equal(a, b))


def max_rel_err(got, src, op, dtype):
    """Max elementwise relative error of got (every member) vs the
    high-precision oracle (exact sum / double-double product)."""
    hi, lo = O.accurate(op, dtype, src)
    if dtype in ("c32", "c64"):
        g = got.view(got.real.dtype).astype(np.float64)
    else:
        g = got.astype(np.float64)
    err = np.abs((g - hi) - lo)
    den = np.abs(hi)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.where(den > 0, err / den, err)
    return float(rel.max()) if rel.size else 0.0


def put_bytes(h, sym, rows, pes):
    for k, pe in enumerate(pes):
        h.put(sym, rows[k], pe)


def get_bytes(h, sym, nbytes, pes):
    return np.stack([h.get(sym, nbytes, np.uint8, pe) for pe in pes])


def move(h, kind, src_rows, block_bytes, team=None, pes=None, root=0):
    """fcollect / alltoall / broadcast of byte rows through the *mem entry
    points; returns every member's dest bytes."""
    P = src_rows.shape[0]
    team = team or h.world
    pes = list(range(P)) if pes is None else list(pes)
    src_len = src_rows.shape[1]
    dest_len = {"fcollect": P * block_bytes, "alltoall": P * block_bytes,
                "broadcast": block_bytes}[kind]
    s = h.malloc(max(src_len, 1))
    d = h.malloc(max(dest_len, 1))
    try:
        put_bytes(h, s, src_rows, pes)
        if kind == "fcollect":
            rc = h.L.shmem_fcollectmem(team, d, s, block_bytes)
        elif kind == "alltoall":
            rc = h.L.shmem_alltoallmem(team, d, s, block_bytes)
        else:
            rc = h.L.shmem_broadcastmem(team, d, s, block_bytes, root)
        osc.check(rc, kind)
        return get_bytes(h, d, dest_len, pes)
    finally:
        h.free(s)
        h.free(d)

"""Writes tests/golden/rtl_gather_kat.json: known-answer vectors for the pull gather.

Source of the cases: the reference's RTL unit-test fixtures and stimulus policies --
  fixtures        03_capi_integration/accelerator_verification/rtl/unit/algorithms/tb/
                  graph_fixture_pkg.sv:37-107 (13 hand-written edge lists, with weights)
  inverse rows    graph_fixture_pkg.sv:130-163 (k-th in-edge of v = k-th edge in list order
                  whose destination is v)
  PageRank values pagerank_kernel_tb.sv:242-265 (contribution classes SMALL/ZERO/MAX/MIXED,
                  per in-edge k of row v with in-neighbour nb)
  PageRank golden pagerank_kernel_tb.sv:303-314 (stream order; float: the first value starts
                  the accumulator, binary32 adds; FixedPoint: 8-B values, Quantized: 4-B values
                  zero-extended, both summed modulo 2^64)
  SpMV values     spmv_kernel_tb.sv:252-295 (x by class, coefficient = fixture weight / 2 in
                  float or the raw weight in fixed point)
  SpMV golden     spmv_kernel_tb.sv:298-311, 349-363 (float: binary32 product then stream-order
                  binary32 add; fixed: (u32)((x*w) >> 16) summed modulo 2^64)
  run list        pagerank_kernel_tb.sv:617-634, spmv_kernel_tb.sv:683-697

The expected values are computed here with exact rational arithmetic (fractions.Fraction)
and an explicit round-to-nearest-even to binary32 -- not with IEEE float adds -- and with
Python integers for the modular sums, so they check the C oracle and the GPU kernels
independently. Values are stored as bit patterns (float32 as u32, integers as decimal
strings).

    python tests/golden/make_rtl_kat.py
"""
from __future__ import annotations

import json
import os
import struct
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "rtl_gather_kat.json")

# (name, num_vertices, edges[(src, dst)], weights[])  -- graph_fixture_pkg.sv:37-107
FIXTURES = [
    ("empty", 0, [], []),
    ("single-vertex", 1, [], []),
    ("directed-chain", 4, [(0, 1), (1, 2), (2, 3)], [1, 2, 3]),
    ("star", 5, [(0, 1), (0, 2), (0, 3), (0, 4)], [1, 2, 3, 4]),
    ("directed-cycle", 3, [(0, 1), (1, 2), (2, 0)], [2, 3, 4]),
    ("disconnected", 4, [(0, 1), (2, 3)], [5, 6]),
    ("self-loop", 2, [(0, 0), (0, 1)], [7, 8]),
    ("duplicate-edge", 2, [(0, 1), (0, 1)], [3, 3]),
    ("one-triangle", 3, [(0, 1), (1, 0), (1, 2), (2, 1), (0, 2), (2, 0)], [1] * 6),
    ("shared-edge-triangles", 4,
     [(0, 1), (1, 0), (0, 2), (2, 0), (1, 2), (2, 1), (1, 3), (3, 1), (2, 3), (3, 2)], [1] * 10),
    ("k4", 4, [(0, 1), (1, 0), (0, 2), (2, 0), (0, 3), (3, 0), (1, 2), (2, 1), (1, 3), (3, 1),
               (2, 3), (3, 2)], [1] * 12),
    ("weighted-toy-matrix", 3, [(0, 0), (0, 1), (1, 2), (2, 0), (2, 2)], [65535, 3, 100, 7, 2]),
    ("dangling-sink", 4, [(0, 1), (1, 2)], [1, 1]),
]

MASK64 = (1 << 64) - 1

# ------------------------------------------------------------------ binary32, exactly

F32_MAX = Fraction((1 << 24) - 1) * Fraction(2) ** (127 - 23)


def f32_value(bits: int) -> Fraction | str:
    """Exact value of a binary32 bit pattern ('inf'/'-inf' for infinities)."""
    sign = -1 if bits >> 31 else 1
    e = (bits >> 23) & 0xFF
    m = bits & 0x7FFFFF
    if e == 0xFF:
        if m:
            raise ValueError("NaN not expected")
        return "inf" if sign > 0 else "-inf"
    if e == 0:
        return sign * Fraction(m) * Fraction(2) ** (-149)
    return sign * Fraction((1 << 23) | m) * Fraction(2) ** (e - 150)


def f32_round(x: Fraction, neg_zero: bool = False) -> int:
    """Round an exact rational to binary32 with round-to-nearest-even; returns bits."""
    if x == 0:
        return 0x80000000 if neg_zero else 0
    sign = 0x80000000 if x < 0 else 0
    a = -x if x < 0 else x
    # overflow: at or beyond the midpoint between F32_MAX and 2^128
    if a >= F32_MAX + Fraction(2) ** (127 - 24):
        return sign | 0x7F800000
    # find exponent e with 2^e <= a < 2^(e+1)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    e = max(e, -126)
    ulp = Fraction(2) ** (e - 23)
    q, r = divmod(a, ulp)  # q integer part count of ulps
    q = int(q)
    half = ulp / 2
    if r > half or (r == half and (q & 1)):
        q += 1
    # q may have overflowed into the next binade; reconstruct via value
    val = q * ulp
    if val >= F32_MAX + Fraction(2) ** (127 - 24):
        return sign | 0x7F800000
    if e == -126 and q < (1 << 23):
        return sign | q  # subnormal
    if q == (1 << 24):
        e += 1
        q >>= 1
    if e > 127:
        return sign | 0x7F800000
    return sign | ((e + 127) << 23) | (q & 0x7FFFFF)


def f32_add(a_bits: int, b_bits: int) -> int:
    a, b = f32_value(a_bits), f32_value(b_bits)
    if isinstance(a, str) or isinstance(b, str):
        if isinstance(a, str) and isinstance(b, str) and a != b:
            raise ValueError("inf - inf")
        s = a if isinstance(a, str) else b
        return 0x7F800000 if s == "inf" else 0xFF800000
    s = a + b
    neg_zero = s == 0 and (a_bits >> 31) and (b_bits >> 31)
    return f32_round(s, bool(neg_zero))


def f32_mul(a_bits: int, b_bits: int) -> int:
    a, b = f32_value(a_bits), f32_value(b_bits)
    if isinstance(a, str) or isinstance(b, str):
        raise ValueError("inf operand not expected in products")
    p = a * b
    neg_zero = p == 0 and ((a_bits ^ b_bits) >> 31)
    return f32_round(p, bool(neg_zero))


def f32_ratio(num: int, den: int) -> int:
    """fp32_from_ratio (models/fp32_reference.cpp:46-49): float(num)/float(den), one rounding
    (num, den are small integers, exactly representable)."""
    return f32_round(Fraction(num, den))


# ------------------------------------------------------------------ fixtures

def inverse_rows(n, edges):
    """Inverse CSR in the fixture's own order (graph_fixture_pkg.sv:130-163)."""
    rows = [[] for _ in range(n)]
    for e, (s, d) in enumerate(edges):
        rows[d].append(e)
    return rows


def pr_value(kind: str, cls: str, nb: int, k: int) -> int:
    """pagerank_kernel_tb.sv:242-265"""
    if kind == "float":
        if cls == "zero":
            return 0
        if cls == "max":
            return 0x7F7FFFFF
        if cls == "mixed":
            return f32_ratio(nb + 1, 3) if k % 2 == 0 else (0x80000000 | f32_ratio(k + 1, 7))
        return f32_ratio(nb + 1, 3)
    width = 64 if kind == "fixed64" else 32
    mask = (1 << width) - 1
    if cls == "zero":
        return 0
    if cls == "max":
        return mask
    if cls == "mixed":
        return (0xFFFFFFFF00000001 & mask) if k % 2 == 0 else ((nb + 1) * 7) & mask
    return ((nb + 1) * 7 + k) & mask


def spmv_x(kind: str, cls: str, nb: int, k: int) -> int:
    """spmv_kernel_tb.sv:256-279"""
    if kind == "float":
        return pr_value("float", cls, nb, k)
    if cls == "zero":
        return 0
    if cls == "max":
        return 0xFFFFFFFF
    if cls == "mixed":
        return 0xFFFF0001 if k % 2 == 0 else ((nb + 1) * 7) & 0xFFFFFFFF
    return ((nb + 1) * 7 + k) & 0xFFFFFFFF


def spmv_w(kind: str, cls: str, weight: int) -> int:
    """spmv_kernel_tb.sv:281-295"""
    if kind == "float":
        if cls == "zero":
            return 0
        if cls == "max":
            return 0x7F7FFFFF
        return f32_ratio(weight, 2)
    if cls == "zero":
        return 0
    if cls == "max":
        return 0xFFFFFFFF
    return weight & 0xFFFFFFFF


PR_RUNS = [("single-vertex", "small"), ("directed-chain", "small"), ("star", "small"),
           ("directed-cycle", "small"), ("disconnected", "small"), ("self-loop", "small"),
           ("duplicate-edge", "small"), ("dangling-sink", "zero"), ("k4", "mixed"),
           ("k4", "max"), ("shared-edge-triangles", "small"), ("directed-chain", "mixed"),
           ("empty", "small")]
SPMV_RUNS = [("single-vertex", "small"), ("directed-chain", "small"), ("star", "small"),
             ("directed-cycle", "small"), ("disconnected", "small"), ("self-loop", "small"),
             ("duplicate-edge", "small"), ("dangling-sink", "zero"), ("k4", "mixed"),
             ("k4", "max"), ("weighted-toy-matrix", "small"), ("weighted-toy-matrix", "max"),
             ("directed-chain", "mixed"), ("empty", "small")]


def case_common(name):
    fx = {f[0]: f for f in FIXTURES}[name]
    _, n, edges, weights = fx
    rows = inverse_rows(n, edges)
    idx, deg, nbs, eids = [], [], [], []
    pos = 0
    for v in range(n):
        idx.append(pos)
        deg.append(len(rows[v]))
        for e in rows[v]:
            nbs.append(edges[e][0])
            eids.append(e)
            pos += 1
    return n, edges, weights, rows, idx, deg, nbs, eids


def make_pr_cases():
    cases = []
    for kind in ("float", "fixed64", "quant32"):
        for name, cls in PR_RUNS:
            n, edges, weights, rows, idx, deg, nbs, eids = case_common(name)
            values, expected = [], []
            for v in range(n):
                acc = 0
                for k in range(deg[v]):
                    val = pr_value(kind, cls, nbs[idx[v] + k], k)
                    values.append(val)
                    if kind == "float":
                        acc = val if k == 0 else f32_add(acc, val)
                    else:
                        acc = (acc + val) & MASK64
                expected.append(acc)
            cases.append(dict(op="pagerank_gather", kind=kind, fixture=name, cls=cls, n=n,
                              idx=idx, deg=deg, neighbour=nbs,
                              values=[str(x) for x in values],
                              expected=[str(x) for x in expected]))
    return cases


def make_spmv_cases():
    cases = []
    for kind in ("float", "fixed"):
        for name, cls in SPMV_RUNS:
            n, edges, weights, rows, idx, deg, nbs, eids = case_common(name)
            xs, ws, expected = [], [], []
            for v in range(n):
                acc = 0
                for k in range(deg[v]):
                    e = eids[idx[v] + k]
                    x = spmv_x(kind, cls, nbs[idx[v] + k], k)
                    w = spmv_w(kind, cls, weights[e])
                    xs.append(x)
                    ws.append(w)
                    if kind == "float":
                        p = f32_mul(x, w)
                        acc = p if k == 0 else f32_add(acc, p)
                    else:
                        p = ((x * w) >> 16) & 0xFFFFFFFF
                        acc = (acc + p) & MASK64
                expected.append(acc)
            cases.append(dict(op="spmv_gather", kind=kind, fixture=name, cls=cls, n=n, idx=idx,
                              deg=deg, neighbour=nbs, x=[str(v) for v in xs],
                              w=[str(v) for v in ws], expected=[str(x) for x in expected]))
    return cases


def main():
    # self-checks of the exact rounding against struct's (IEEE) conversion of doubles
    for num, den in [(1, 3), (2, 3), (5, 7), (65535, 2), (1, 10)]:
        assert f32_ratio(num, den) == struct.unpack("<I", struct.pack("<f", num / den))[0]
    assert f32_add(0x7F7FFFFF, 0x7F7FFFFF) == 0x7F800000
    data = dict(source="03_capi_integration/accelerator_verification/rtl/unit/algorithms",
                generator="tests/golden/make_rtl_kat.py",
                cases=make_pr_cases() + make_spmv_cases())
    with open(OUT, "w") as f:
        json.dump(data, f, indent=0, separators=(",", ":"))
    print(f"wrote {len(data['cases'])} cases to {OUT}")


if __name__ == "__main__":
    main()

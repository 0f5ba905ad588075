"""Writers and readers of combined_sieve's unknown-file lines, restated in
Python for the tests (the checker side; the product's C++ reader lives in
prime-gap_amd/csrc/unknowns.cpp).

Line formats (combined_sieve.cpp:1261-1445, gap_common.cpp:1186-1359):
  plain (0):          "m : -A +B | -p1 -p2 ... | +n1 +n2 ...\\n"
  rle (1):            "m : -A +B | <2 bytes per delta> | <2 bytes per delta>\\n",
                      delta from the previous distance (first from 0) as
                      chr(48 + (d >> 7)) + chr(48 + (d & 127))
  bitcompressed (2):  "m : <unknowns> <bytes> || <raw>\\n", one bit per x in
                      [-SL, SL] with gcd(K, x) = 1 and gcd(m*K + x, d) = 1,
                      bit 1 = composite, 7 bits per byte, high bit set.
"""
from __future__ import annotations

import math


def plain_line(m, prev, nxt) -> bytes:
    s = f"{m} : -{len(prev)} +{len(nxt)} |" + "".join(f" -{int(d)}" for d in prev) + " |" + \
        "".join(f" +{int(d)}" for d in nxt) + "\n"
    return s.encode()


def _rle(dists) -> bytes:
    out = bytearray()
    last = 0
    for d in dists:
        delta = int(d) - last
        last = int(d)
        assert 0 <= delta < 128 * 128
        out += bytes((48 + (delta >> 7), 48 + (delta & 127)))
    return bytes(out)


def rle_line(m, prev, nxt) -> bytes:
    return f"{m} : -{len(prev)} +{len(nxt)} | ".encode() + _rle(prev) + b" | " + _rle(nxt) + b"\n"


def coprime_x(K: int, SL: int):
    return [x for x in range(-SL, SL + 1) if math.gcd(K, x) == 1]


def bitcompressed_line(m, prev, nxt, K: int, D: int, SL: int, cop=None) -> bytes:
    cop = coprime_x(K, SL) if cop is None else cop
    unknown = set(-int(d) for d in prev) | set(int(d) for d in nxt)
    raw = bytearray()
    b = 1 << 7
    bit = 0
    for x in cop:
        if math.gcd(m * K + x, D) != 1:
            continue
        comp = 0 if x in unknown else 1
        b |= comp << bit
        bit += 1
        if bit == 7:
            raw.append(b)
            b = 1 << 7
            bit = 0
    if bit:
        raw.append(b)
    return f"{m} : {len(prev) + len(nxt)} {len(raw)} || ".encode() + bytes(raw) + b"\n"


def parse_line(line: bytes, K: int = 0, D: int = 0, SL: int = 0, cop=None):
    """-> (m, prev distances, next distances); prev/next outward (increasing)."""
    head, _, rest = line.partition(b" : ")
    m = int(head)
    if b" || " in line[:40]:
        counts, _, raw = rest.partition(b" || ")
        total, nbytes = map(int, counts.split())
        raw = raw[:nbytes]
        cop = coprime_x(K, SL) if cop is None else cop
        prev, nxt = [], []
        i = 0
        for x in cop:
            if math.gcd(m * K + x, D) != 1:
                continue
            if ((raw[i // 7] >> (i % 7)) & 1) == 0:
                (prev if x < 0 else nxt).append(abs(x))
            i += 1
        assert len(prev) + len(nxt) == total
        return m, sorted(prev), nxt
    hdr, _, body = rest.partition(b" | ")
    a, b = hdr.split()
    na, nb = -int(a), int(b)
    left, _, right = body.partition(b" | ")
    right = right.rstrip(b"\n")
    if na and b" " not in left[:20] and not left.startswith(b"-"):
        def un(s, n):
            vals, acc = [], 0
            for k in range(n):
                acc += 128 * (s[2 * k] - 48) + (s[2 * k + 1] - 48)
                vals.append(acc)
            return vals
        return m, un(left, na), un(right, nb)
    prev = [-int(t) for t in left.split()] if left.strip() else []
    nxt = [int(t) for t in right.split()] if right.strip() else []
    assert len(prev) == na and len(nxt) == nb
    return m, prev, nxt

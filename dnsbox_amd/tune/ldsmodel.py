"""Bank-conflict model of the x pass's LDS accesses (box.hpp, XPass / RowFFT).

The LDS serves a wave's access in lane groups of BANKS*4 bytes; a group costs as many cycles
as the most-used bank in it.  The model replays every LDS access of one block of the x pass
(staging of the input lines, c2r pre-processing, each Stockham stage of the row FFTs in both
directions, the pointwise op, r2c post-processing) for a row layout `hp` (row stride, complex
elements) and `xpd` (element j at j + j // xpd), and returns cycles / conflict-free cycles.
It is evaluated for 32 and 64 banks; the autotuner keeps the layouts good under both.
"""
from __future__ import annotations


def _cycles(addrs, esz, banks):
    words = esz // 4
    per = banks * 4 // esz
    tot = 0
    for g0 in range(0, len(addrs), per):
        cnt = {}
        for a in addrs[g0:g0 + per]:
            if a is None:
                continue
            for w in range(words):
                b = (a * words + w) % banks
                cnt[b] = cnt.get(b, 0) + 1
        if cnt:
            tot += max(cnt.values())
    return tot


def _stages(radices):
    out, p = [], 1
    for r in radices:
        out.append((r, p))
        p *= r
    return p, out


def xpass_ratio(n, esz, rb, ntx, radices, hp, xpd, fi=6, fo=3, banks=32):
    m = n // 2
    h = m + 1
    length, stages = _stages(radices)
    assert length == m
    pad = (lambda j: j + j // xpd) if xpd else (lambda j: j)
    at = lambda row, j: row * hp + pad(j)
    cost = ideal = 0

    def acc(fn, count):
        nonlocal cost, ideal
        for w0 in range(0, ntx, 64):
            for q in range(count):
                addrs = [fn(t, q) for t in range(w0, w0 + 64)]
                if any(a is not None for a in addrs):
                    cost += _cycles(addrs, esz, banks)
                    ideal += _cycles(list(range(64)), esz, banks)

    def ceil(a, b):
        return -(-a // b)

    nl = fi * rb * h
    acc(lambda t, q: None if t + q * ntx >= nl else
        at(((t + q * ntx) // (rb * h)) * rb + ((t + q * ntx) % (rb * h)) // h, ((t + q * ntx) % (rb * h)) % h),
        ceil(nl, ntx))
    ni = fi * rb * m
    acc(lambda t, q: None if t + q * ntx >= ni else at((t + q * ntx) // m, (t + q * ntx) % m), ceil(ni, ntx))
    acc(lambda t, q: None if t + q * ntx >= ni else at((t + q * ntx) // m, m - (t + q * ntx) % m), ceil(ni, ntx))
    acc(lambda t, q: None if t + q * ntx >= ni else at((t + q * ntx) // m, (t + q * ntx) % m), ceil(ni, ntx))
    for rows in (fi * rb, fo * rb):
        for r_, p in stages:
            nbf = rows * (length // r_)
            for r in range(r_):
                def rd(t, q, r=r, r_=r_):
                    b = t + q * ntx
                    if b >= nbf:
                        return None
                    s, i = divmod(b, length // r_)
                    return at(s, i + r * (length // r_))

                def wr(t, q, r=r, r_=r_, p=p):
                    b = t + q * ntx
                    if b >= nbf:
                        return None
                    s, i = divmod(b, length // r_)
                    k = i % p
                    return at(s, (i - k) * r_ + k + r * p)
                acc(rd, ceil(nbf, ntx))
                acc(wr, ceil(nbf, ntx))
    acc(lambda t, q: None if t + q * ntx >= rb * m else at((t + q * ntx) // m, (t + q * ntx) % m), ceil(rb * m, ntx))
    no = fo * rb * h

    def post(t, q, mirror):
        e = t + q * ntx
        if e >= no:
            return None
        f, rk = divmod(e, rb * h)
        r, k = divmod(rk, h)
        j = (0 if k == 0 else m - k) if mirror else (0 if k == m else k)
        return at(f * rb + r, j)
    acc(lambda t, q: post(t, q, False), ceil(no, ntx))
    acc(lambda t, q: post(t, q, True), ceil(no, ntx))
    return cost / ideal


def best_layouts(n, prec, rb, ntx, radices, top=2, periods=(0, 8, 16, 32), extra=16):
    """The `top` (xpd, xpad, hp) layouts by the sum of the 32- and 64-bank ratios."""
    esz = 8 if prec == 32 else 16
    h = n // 2 + 1
    rows = []
    for xpd in periods:
        base = h + (h // xpd + 1 if xpd else 0)
        for xpad in range(extra):
            hp = base + xpad
            r = sum(xpass_ratio(n, esz, rb, ntx, radices, hp, xpd, banks=b) for b in (32, 64))
            rows.append((r, hp, xpd, xpad))
    rows.sort()
    out, seen = [], set()
    for r, hp, xpd, xpad in rows:
        if (xpd, xpad) in seen:
            continue
        seen.add((xpd, xpad))
        out.append((xpd, xpad, hp, r / 2))
        if len(out) == top:
            break
    return out

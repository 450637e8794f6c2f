"""Writes lds_layout.hpp: the LDS placement of the line FFTs' exchanges for every line shape
the kernels instantiate (frame sides 16 … 512; the row, pass-C, column and pass-B shapes of float and
double), chosen to minimise LDS bank conflicts.

    python ptyrs_amd/csrc/gen_lds_layout.py

The exchanges of a mixed-radix Stockham line FFT (fft.hpp) write element
`(jb / NS)·NS·R + jb mod NS + r·NS` and read element `jl + TL·e` of each line. A layout maps
(line, element i) to an LDS element (a float or a double: one or two 4-byte banks), either one
region per line, `line·LD + i + (i >> SH)·MU`, or interleaved across the workgroup's L lines,
`(i + (i >> SH)·MU)·L + line`. For each shape and for both lane maps (lanes along a line: row
passes; lanes across lines: column passes) this script counts, per LDS instruction of a wave,
the most lanes that fall on one bank, under three bank models (32 banks over 32-lane halves,
64 banks over the whole wave, 64 banks over halves), and keeps the layout with the smallest
worst-model mean, then the smallest mean over models, then the fewest elements (at most 5/4 of
the plain padded layout's).
"""

import pathlib

import numpy as np

OUT = pathlib.Path(__file__).with_name("lds_layout.hpp")
SIDES = (16, 24, 32, 48, 64, 96, 128, 256, 512)
THREADS = 256
MODELS = ((32, True), (64, False), (64, True))


def pick_radix(rem, e):
    for r in (16, 8, 4, 2, 3, 5):
        if r <= e and e % r == 0 and rem % r == 0:
            return r
    return 0


def n_stages(n, e):
    rem, s = n, 0
    while rem > 1:
        r = pick_radix(rem, e)
        if r == 0:
            return -1
        rem //= r
        s += 1
    return s


def choose_e(n, pref):
    if n % pref == 0 and n_stages(n, pref) > 0 and n // pref <= 1024:
        return pref
    for e in (8, 16, 12, 24, 4, 6, 10, 20, 32, 40, 48, 2):
        if n % e == 0 and n_stages(n, e) > 0 and n // e <= 256:
            return e
    return 0


def lines_per_block(n, tl, maxt, maxl):
    best = 1
    for l in range(1, min(n, maxl) + 1):
        if n % l == 0 and l * tl <= maxt:
            best = l
    return best


def shapes():
    """(N, E, L, element bytes) of RowShape, RowShapeC, ColShape and BShape (kernels.hpp) for
    float and double."""
    out = set()
    for size in (4, 8):
        for n in SIDES:
            e = choose_e(n, 16 if size == 4 else 8)
            out.add((n, e, lines_per_block(n, n // e, THREADS, n), size))
            cg = (8 if size == 4 else 4) // (2 if n >= 512 else 1)
            out.add((n, e, lines_per_block(n, n // e, 1024, cg), size))
            e = choose_e(n, 8 if size == 4 else 4)
            out.add((n, e, lines_per_block(n, n // e, THREADS, n), size))
            out.add((n, e, lines_per_block(n, n // e, 512, 8 if size == 4 else 4), size))
    return sorted(out)


def instructions(n, e, l, along):
    """(line, element) of every lane of every LDS instruction: [instr, lanes, 2]."""
    tl = n // e
    lane = np.arange(tl * l)
    line, jl = (lane // tl, lane % tl) if along else (lane % l, lane // l)
    rows = []
    ns = 1
    rem = n
    while rem > 1:
        r_ = pick_radix(rem, e)
        if ns * r_ == n:
            break
        for u in range(e // r_):
            jb = jl + tl * u
            base = (jb // ns) * ns * r_ + jb % ns
            for r in range(r_):
                rows.append(np.stack([line, base + r * ns], axis=1))
        for k in range(e):
            rows.append(np.stack([line, jl + tl * k], axis=1))
        ns *= r_
        rem //= r_
    return np.array(rows) if rows else None


def score(addr, banks, half, words=1):
    """Mean over instructions and lane groups of the most lanes on one bank; an element of
    `words` 4-byte words occupies that many consecutive banks."""
    nin, nl = addr.shape
    group = 32 if half else 64
    ng = (nl + group - 1) // group
    pad = ng * group - nl
    a = np.pad(addr, ((0, 0), (0, pad)), constant_values=-1).reshape(nin * ng, group)
    valid = a >= 0
    total = np.zeros((nin * ng, banks), dtype=np.int64)
    for w in range(words):
        b = (a * words + w) % banks
        key = np.where(valid, np.arange(nin * ng)[:, None] * banks + b, nin * ng * banks)
        total += np.bincount(key.ravel(), minlength=nin * ng * banks + 1)[:-1].reshape(
            nin * ng, banks)
    has = valid.any(axis=1)
    return float(total.max(axis=1)[has].mean())


def best_layout(n, e, l, along, size):
    words = size // 4
    ins = instructions(n, e, l, along)
    plain = l * (n + (n >> 4) + 1)
    if ins is None:
        return (0, 4, 1, n + (n >> 4) + 1, plain), None
    line, i = ins[..., 0], ins[..., 1]
    cands = []
    for sh in range(1, 7):
        for mu in range(0, 4):
            span = (n - 1) + ((n - 1) >> sh) * mu + 1
            if span * l <= plain * 5 // 4:
                cands.append((1, sh, mu, span, span * l))
            for ld in range(span, span + 33):
                if ld * l <= plain * 5 // 4:
                    cands.append((0, sh, mu, ld, ld * l))
    best = None
    for c in cands:
        inter, sh, mu, ld, elems = c
        p = i + (i >> sh) * mu
        addr = p * l + line if inter else line * ld + p
        s = [score(addr, b, h, words) for b, h in MODELS]
        key = (round(max(s), 6), round(sum(s), 6), elems, c)
        if best is None or key < best[0]:
            best = (key, c, s)
    cur = line * (n + (n >> 4) + 1) + i + (i >> 4)
    return best[1], ([score(cur, b, h, words) for b, h in MODELS], best[2])


def main():
    lines = []
    report = []
    for (n, e, l, size) in shapes():
        for along in (True, False):
            c, s = best_layout(n, e, l, along, size)
            lines.append(f"  if (n == {n} && e == {e} && l == {l} && along == {int(along)} && "
                         f"bytes == {size}) "
                         f"return LdsLay{{{c[0]}, {c[1]}, {c[2]}, {c[3]}, {c[4]}}};")
            if s:
                report.append(f"// N={n} E={e} L={l} {size}-byte {'along' if along else 'across'}: "
                              f"plain {[round(x, 2) for x in s[0]]} -> "
                              f"{[round(x, 2) for x in s[1]]}")
    text = f'''// Generated by gen_lds_layout.py (same folder); do not edit by hand.
// LDS placement of the line FFTs' exchanges per line shape (N values per line, E per thread,
// L lines per workgroup, element bytes) and lane map (along: lanes run along a line; across:
// lanes run across the lines): element i of line `line` sits at
//   interleaved: (i + (i >> shift)·mul)·L + line,   else: line·ld + i + (i >> shift)·mul,
// in an array of `words` elements. Most lanes on one LDS bank per instruction (mean; models:
// 32 banks per half-wave, 64 per wave, 64 per half-wave), the plain padded layout -> this one:
{chr(10).join(report)}
#pragma once

namespace ptyrs {{

struct LdsLay {{
  int inter, shift, mul, ld, words;
}};

constexpr LdsLay lds_layout(int n, int e, int l, bool along, int bytes) {{
{chr(10).join(lines)}
  return LdsLay{{0, 4, 1, n + (n >> 4) + 1, l * (n + (n >> 4) + 1)}};
}}

}}  // namespace ptyrs
'''
    OUT.write_text(text)
    print("wrote", OUT)


if __name__ == "__main__":
    main()

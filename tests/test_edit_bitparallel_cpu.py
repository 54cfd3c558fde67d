"""CPU check of the recurrence the edit verify kernel runs (k_edit_verify_bp
in rnng_amd/csrc/edit_eps.hip): Myers/Hyyro bit-parallel Levenshtein
restricted to a (2k+1)-row window that slides down the diagonal, with +1
surrogate deltas at the window's edges.  Restated here in Python and
compared with the oracle's full dynamic programme: equal whenever the
distance is <= k, above k otherwise."""
import random

import oracle


def banded_bit_parallel(p: bytes, t: bytes, k: int) -> int:
    m, n = len(p), len(t)
    if abs(m - n) > k:
        return k + 1
    w = 2 * k + 1
    W = (1 << w) - 1
    Pv, Mv, T, R = W, 0, 0, 1
    for j in range(1, n + 1):
        if j > k + 1:
            T += 1 + (Pv & 1) - (Mv & 1)
            Pv = (Pv >> 1) | (1 << (w - 1))
            Mv >>= 1
            R += 1
        else:
            T += 1
        Eq = 0
        for b in range(w):
            r = R + b
            if r <= m and p[r - 1] == t[j - 1]:
                Eq |= 1 << b
        Xv = Eq | Mv
        Xh = ((((Eq & Pv) + Pv) & W) ^ Pv) | Eq
        Ph = (Mv | ~(Xh | Pv)) & W
        Mh = Pv & Xh
        Ph = ((Ph << 1) | 1) & W
        Mh = (Mh << 1) & W
        Pv = (Mh | ~(Xv | Ph)) & W
        Mv = Ph & Xv
    nb = m - R + 1
    mask = (1 << nb) - 1 if nb > 0 else 0
    return T + bin(Pv & mask).count("1") - bin(Mv & mask).count("1")


def _mutate(rng, s, alpha, edits):
    t = bytearray(s)
    for _ in range(edits):
        op = rng.random()
        if op < 0.33 and t:
            del t[rng.randrange(len(t))]
        elif op < 0.66:
            t.insert(rng.randint(0, len(t)), rng.choice(alpha))
        elif t:
            t[rng.randrange(len(t))] = rng.choice(alpha)
    return bytes(t)


def test_banded_bit_parallel_equals_full_dp():
    rng = random.Random(5)
    checked = within = 0
    for _ in range(4000):
        alpha = rng.choice([b"AC", b"ACGT", b"abcdxyz", b"a"])
        s = bytes(rng.choice(alpha) for _ in range(rng.randint(0, 45)))
        t = _mutate(rng, s, alpha, rng.randint(0, 12)) if rng.random() < 0.8 else \
            bytes(rng.choice(alpha) for _ in range(rng.randint(0, 45)))
        k = rng.randint(0, 14)
        d = oracle.edit_distance(s, t)
        b = banded_bit_parallel(s, t, k)
        if d <= k:
            assert b == d, (s, t, k, d, b)
            within += 1
        else:
            assert b > k, (s, t, k, d, b)
        checked += 1
    assert within > 1000


M64 = (1 << 64) - 1


def multiword_bit_parallel(p: bytes, t: bytes, k: int, words: int, past_end: bytes) -> int:
    """The recurrence of k_edit_verify_myers (radii above 32): the same column
    step over every row of the pattern, held in `words` 64-bit words with the
    carries of the addition and of the shifts passed word to word, bits past
    the pattern's end taken from `past_end` (codes read past the string), the
    score following row m, and the early exit after each group of 32 columns."""
    m, n = len(p), len(t)
    if m == 0:
        return n
    pp = p + past_end[: 64 * words - m]
    Pv, Mv = [M64] * words, [0] * words
    wl, bl = (m - 1) >> 6, (m - 1) & 63
    score = m
    for g in range((n + 31) // 32):
        cols = min(32, n - 32 * g)
        for u in range(cols):
            c = t[32 * g + u]
            cy, ph_in, mh_in = 0, 1, 0
            for w in range(words):
                Eq = sum(1 << b for b in range(64) if pp[64 * w + b] == c)
                Xv = Eq | Mv[w]
                a = Eq & Pv[w]
                s1 = (a + Pv[w]) & M64
                s2 = (s1 + cy) & M64
                cy = int(s1 < a or s2 < s1)
                Xh = (s2 ^ Pv[w]) | Eq
                Ph = (Mv[w] | ~(Xh | Pv[w])) & M64
                Mh = Pv[w] & Xh
                if w == wl:
                    score += ((Ph >> bl) & 1) - ((Mh >> bl) & 1)
                Phs = ((Ph << 1) | ph_in) & M64
                Mhs = ((Mh << 1) | mh_in) & M64
                ph_in, mh_in = Ph >> 63, Mh >> 63
                Pv[w] = (Mhs | ~(Xv | Phs)) & M64
                Mv[w] = Phs & Xv
        if score - (n - 32 * g - cols) > k:
            return k + 1
    return score


def test_multiword_bit_parallel_equals_full_dp():
    """Equal to the full dynamic programme whenever the distance is <= k,
    above k otherwise, for patterns filling 1, 2 or 3 words (carries across
    word boundaries), garbage codes past the pattern's end and radii above
    32 (the early exit fires on the far pairs)."""
    rng = random.Random(7)
    within = 0
    for _ in range(500):
        alpha = rng.choice([b"AC", b"ACGT", b"abcdefghijklmnopqrstu"])
        words = rng.choice([1, 2, 3])
        m = rng.randint(0, 64 * words)
        s = bytes(rng.choice(alpha) for _ in range(m))
        if rng.random() < 0.6:
            t = _mutate(rng, s, alpha, rng.randint(0, 50))
        else:
            t = bytes(rng.choice(alpha) for _ in range(rng.randint(0, 64 * words + 30)))
        if len(t) < len(s):
            s, t = t, s                     # the shorter string is the pattern
        k = rng.choice([33, 40, 64, 10 ** 9])
        past_end = bytes(rng.choice(alpha) for _ in range(64 * words))
        d = oracle.edit_distance(s, t)
        b = multiword_bit_parallel(s, t, k, words, past_end)
        if d <= k:
            assert b == d, (s, t, k, d, b)
            within += 1
        else:
            assert b > k, (s, t, k, d, b)
    assert within > 150


def wavefront_bit_parallel(p: bytes, t: bytes, k: int, lanes: int, past_end: bytes) -> int:
    """The recurrence of k_edit_verify_long (shorter string above 512 atoms),
    with `lanes` lanes per wave (64 on the GPU): bands of `lanes` pattern
    words; at step s lane w of a band works on text column s - w with the
    carries lane w - 1 produced at step s - 1; lane 0 takes them from the top
    row (band 0) or from the carries the band below's last lane left per
    column; row m's lane keeps the score and the wave stops after every 32nd
    step once D[m][j] - (n - j) > k."""
    m, n = len(p), len(t)
    if m == 0:
        return n
    W = (m + 63) >> 6
    nb = (W + lanes - 1) // lanes
    pp = p + past_end[: 64 * W - m]
    wl, bl = (m - 1) >> 6, (m - 1) & 63
    score = m
    below = [0] * n                         # carries of the band below, per column
    for b in range(nb):
        last = min(lanes - 1, W - 1 - lanes * b)
        top = b == nb - 1
        wlane = wl - lanes * b
        eq = [None] * lanes
        for lane in range(last + 1):
            w = lanes * b + lane
            eq[lane] = {c: sum(1 << r for r in range(64) if pp[64 * w + r] == c) for c in set(t)}
        Pv, Mv = [M64] * lanes, [0] * lanes
        cin = [0] * lanes
        above = [0] * n
        for s in range(n + last):
            outs = [0] * lanes
            for lane in range(last + 1):
                col = s - lane
                if not 0 <= col < n:
                    continue
                inn = cin[lane] if lane else (2 if b == 0 else below[col])
                Eq = eq[lane][t[col]]
                Xv = Eq | Mv[lane]
                a = Eq & Pv[lane]
                s1 = (a + Pv[lane]) & M64
                s2 = (s1 + (inn & 1)) & M64
                cyo = int(s1 < a or s2 < s1)
                Xh = (s2 ^ Pv[lane]) | Eq
                Ph = (Mv[lane] | ~(Xh | Pv[lane])) & M64
                Mh = Pv[lane] & Xh
                if top and lane == wlane:
                    score += ((Ph >> bl) & 1) - ((Mh >> bl) & 1)
                Phs = ((Ph << 1) | ((inn >> 1) & 1)) & M64
                Mhs = ((Mh << 1) | ((inn >> 2) & 1)) & M64
                Pv[lane] = (Mhs | ~(Xv | Phs)) & M64
                Mv[lane] = Phs & Xv
                outs[lane] = cyo | ((Ph >> 63) << 1) | ((Mh >> 63) << 2)
                if not top and lane == lanes - 1:
                    above[col] = outs[lane]
            cin = [0] + outs[:-1]
            if top and s % 32 == 31:
                cw = s - wlane
                if 0 <= cw < n and score - (n - 1 - cw) > k:
                    return k + 1
        below = above
    return score


def test_wavefront_bit_parallel_equals_full_dp():
    """Equal to the full dynamic programme whenever the distance is <= k,
    above k otherwise: one band and several (4-lane waves: bands of 256
    rows, patterns up to 900 atoms), and two 64-lane bands (4,200 atoms),
    with garbage codes past the pattern's end."""
    rng = random.Random(11)
    within = 0
    cases = [(4, rng.randint(60, 900)) for _ in range(40)] + [(64, 4200)]
    for lanes, m in cases:
        alpha = rng.choice([b"ACGT", b"abcdefghijklmnopqrstu"]) if lanes == 4 else b"ACGT"
        s = bytes(rng.choice(alpha) for _ in range(m))
        t = _mutate(rng, s, alpha, rng.randint(0, 80))
        if len(t) < len(s):
            s, t = t, s
        d = oracle.edit_distance(s, t)
        past_end = bytes(rng.choice(alpha) for _ in range(64))
        for k in (max(0, d - 1), d, 10 ** 9):
            b = wavefront_bit_parallel(s, t, k, lanes, past_end)
            if d <= k:
                assert b == d, (len(s), len(t), k, d, b)
                within += 1
            else:
                assert b > k, (len(s), len(t), k, d, b)
    assert within >= 80

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

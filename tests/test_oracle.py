"""Pins the CPU oracle (oracle/) against known answers, closed forms and
independent brute force, so that a plausible mistake (wrong summation order,
fused multiply-add, off-by-one, transposed operand, wrong tie rule) fails.

The reference ships no tests or golden vectors for this path, so these checks
are what fixes the oracle; see oracle/oracle.h for which check pins what.
"""
import itertools
import math
from functools import lru_cache

import numpy as np
import pytest

F = np.float32


# ---------------------------------------------------------------------------
# L2 distance in the reference order (drivers/systolic.cpp:253-269)
# ---------------------------------------------------------------------------

def test_l2_known_integer_answers(orc):
    p = np.array([3, 0, 0, 1], F)
    q = np.array([0, 4, 0, 1], F)
    assert orc.l2sq(p, q) == F(25)
    assert orc.l2(p, q) == F(5)
    assert orc.l2(p, p) == F(0)
    # d = 128, integer data in [0, 255]: the FP32 sum is exact
    rng = np.random.default_rng(0)
    a = rng.integers(0, 256, 128).astype(F)
    b = rng.integers(0, 256, 128).astype(F)
    exact = int(((a.astype(np.int64) - b.astype(np.int64)) ** 2).sum())
    assert orc.l2sq(a, b) == F(exact)


def test_l2_sequential_order_is_pinned(orc):
    # squares [2^24, 1, 1]: sequential gives 2^24 (each +1 rounds away),
    # any other association (pairwise, reversed) gives 2^24 + 2.
    p = np.array([4096, 1, 1], F)
    z = np.zeros(3, F)
    assert orc.l2sq(p, z) == F(16777216)
    assert orc.l2sq(p[::-1].copy(), z) == F(16777218)


def _seq_nonfused(p, q):
    acc = F(0)
    for a, b in zip(p, q):
        t = F(a - b)
        acc = F(acc + F(t * t))
    return acc


def _seq_fused(p, q):
    acc = F(0)
    for a, b in zip(p, q):
        t = F(a - b)
        acc = F(np.float64(acc) + np.float64(t) * np.float64(t))   # one rounding
    return acc


def test_l2_is_not_fused(orc):
    # find vectors where a fused a*b+c gives a different float than separate
    # rounding; the oracle must give the separately rounded one
    rng = np.random.default_rng(7)
    found = 0
    for _ in range(2000):
        p = rng.standard_normal(5).astype(F)
        q = rng.standard_normal(5).astype(F)
        nf, fu = _seq_nonfused(p, q), _seq_fused(p, q)
        if nf != fu:
            assert orc.l2sq(p, q) == nf
            found += 1
    assert found > 20


def test_l2_symmetric_bitwise(orc):
    rng = np.random.default_rng(1)
    for _ in range(200):
        p = rng.standard_normal(37).astype(F)
        q = rng.standard_normal(37).astype(F)
        assert orc.l2sq(p, q).tobytes() == orc.l2sq(q, p).tobytes()


def test_l2_close_to_double(orc):
    rng = np.random.default_rng(2)
    for d in (1, 3, 16, 96, 128):
        p = rng.standard_normal(d).astype(F)
        q = rng.standard_normal(d).astype(F)
        ex = math.sqrt(float(((p.astype(np.float64) - q.astype(np.float64)) ** 2).sum()))
        assert abs(float(orc.l2(p, q)) - ex) <= (d + 3) * 2.0 ** -24 * ex + 1e-30


# ---------------------------------------------------------------------------
# sq_threshold: sqrtf(v) <= eps  <=>  v <= T(eps)
# ---------------------------------------------------------------------------

def _check_threshold(orc, eps):
    T = orc.sq_threshold(eps)
    eps = F(eps)
    if T < 0:
        assert not (np.sqrt(F(0)) <= eps)
        return
    assert np.sqrt(T) <= eps
    with np.errstate(over="ignore"):
        up = np.nextafter(T, F(np.inf))
    assert not (np.sqrt(up) <= eps)
    # a neighbourhood of T: the equivalence holds value by value
    v = T
    for _ in range(8):
        v = np.nextafter(v, F(0))
        assert (np.sqrt(v) <= eps) == (v <= T)


def test_sq_threshold_property(orc):
    rng = np.random.default_rng(3)
    cases = [0.0, 1e-30, 1e-20, 1e-3, 0.08, 0.5, 1.0, 1.5, 2.0, 255.9, 256.2889, 1e3, 1e15, 1e19,
             3e19, -1.0]
    cases += list(rng.uniform(0, 10, 200)) + list(np.exp(rng.uniform(-30, 40, 200)))
    for e in cases:
        _check_threshold(orc, e)
    assert orc.sq_threshold(-0.5) < 0
    assert orc.sq_threshold(1.5) == F(2.25)


# ---------------------------------------------------------------------------
# Brute-force epsilon pairs (include/search.hpp:389-407 semantics)
# ---------------------------------------------------------------------------

def _np_pairs(X, Y, eps, self_join):
    """Independent enumeration: numpy float32, one op at a time, per pair in
    sequential dimension order."""
    nx, ny = X.shape[0], Y.shape[0]
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    I, J = I.ravel(), J.ravel()
    if self_join:
        keep = I < J
        I, J = I[keep], J[keep]
    acc = np.zeros(I.size, F)
    for k in range(X.shape[1]):
        t = (X[I, k] - Y[J, k]).astype(F)
        acc = (acc + (t * t).astype(F)).astype(F)
    w = np.sqrt(acc)
    m = w <= F(eps)
    return np.stack([I[m], J[m]], 1).astype(np.int64), w[m]


@pytest.mark.parametrize("n,d", [(1, 4), (2, 3), (37, 3), (150, 16), (257, 17), (300, 128)])
def test_eps_pairs_matches_numpy(orc, n, d):
    rng = np.random.default_rng(n * 1000 + d)
    X = rng.standard_normal((n, d)).astype(F)
    eps = F(np.sqrt(d) * 1.1)
    ij, w = orc.l2_eps_pairs(X, eps)
    ij2, w2 = _np_pairs(X, X, eps, True)
    assert np.array_equal(ij, ij2)
    assert w.tobytes() == w2.tobytes()
    # bipartite against another set
    Y = rng.standard_normal((max(1, n // 2), d)).astype(F)
    ij, w = orc.l2_eps_pairs(X, eps, Y)
    ij2, w2 = _np_pairs(X, Y, eps, False)
    assert np.array_equal(ij, ij2)
    assert w.tobytes() == w2.tobytes()


def test_eps_pairs_symmetry_monotone_duplicates(orc):
    rng = np.random.default_rng(11)
    X = rng.standard_normal((200, 5)).astype(F)
    X[50] = X[10]   # duplicate point: distance exactly 0
    ij, w = orc.l2_eps_pairs(X, 1.0)
    full, fw = orc.l2_eps_pairs(X, 1.0, X)
    # bipartite X x X = self pairs both ways + the diagonal
    off = full[full[:, 0] != full[:, 1]]
    und = {tuple(sorted(p)) for p in off}
    assert und == {tuple(p) for p in ij}
    assert np.sum(full[:, 0] == full[:, 1]) == 200
    assert (10, 50) in {tuple(p) for p in ij}
    e0 = {tuple(p) for p in orc.l2_eps_pairs(X, 0.0)[0]}
    assert e0 == {(10, 50)}
    small = {tuple(p) for p in orc.l2_eps_pairs(X, 0.7)[0]}
    assert small <= {tuple(p) for p in ij}
    assert orc.l2_eps_pairs(X, -1.0)[0].shape[0] == 0


def test_eps_pairs_cap_and_count(orc):
    rng = np.random.default_rng(12)
    X = rng.standard_normal((120, 4)).astype(F)
    n = orc.l2_eps_pairs_count(X, 2.0)
    ij, _ = orc.l2_eps_pairs(X, 2.0)
    assert n == ij.shape[0]


# ---------------------------------------------------------------------------
# Edit distance (drivers/systolic.cpp:271-299)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("s,t,d", [
    (b"kitten", b"sitting", 3), (b"flaw", b"lawn", 2), (b"", b"abc", 3), (b"abc", b"", 3),
    (b"", b"", 0), (b"intention", b"execution", 5), (b"sunday", b"saturday", 3),
    (b"ACGT", b"ACGT", 0), (b"ACGTACGT", b"TGCATGCA", 6), (b"AAAA", b"AAAAAAA", 3)])
def test_edit_known_answers(orc, s, t, d):
    assert orc.edit_distance(s, t) == d
    assert orc.edit_distance(t, s) == d


def _lev_recursive(s, t):
    @lru_cache(None)
    def f(i, j):
        if i == 0:
            return j
        if j == 0:
            return i
        return min(f(i - 1, j) + 1, f(i, j - 1) + 1, f(i - 1, j - 1) + (s[i - 1] != t[j - 1]))
    return f(len(s), len(t))


def _lev_matrix(s, t):
    """Full-matrix Wagner-Fischer (a different formulation from the oracle's
    two-row loop)."""
    D = [[0] * (len(t) + 1) for _ in range(len(s) + 1)]
    for i in range(len(s) + 1):
        D[i][0] = i
    for j in range(len(t) + 1):
        D[0][j] = j
    for i in range(1, len(s) + 1):
        for j in range(1, len(t) + 1):
            D[i][j] = min(D[i - 1][j] + 1, D[i][j - 1] + 1, D[i - 1][j - 1] + (s[i - 1] != t[j - 1]))
    return D[-1][-1]


def test_edit_matches_recursive_definition(orc):
    rng = np.random.default_rng(5)
    for _ in range(300):
        s = bytes(rng.choice(list(b"ACGT"), rng.integers(0, 9)))
        t = bytes(rng.choice(list(b"ACGT"), rng.integers(0, 9)))
        assert orc.edit_distance(s, t) == _lev_recursive(s, t)


def test_edit_substitution_only_is_hamming(orc):
    rng = np.random.default_rng(6)
    for _ in range(50):
        s = rng.choice(list(b"ACGT"), 40)
        t = s.copy()
        pos = rng.choice(40, 3, replace=False)
        t[pos] = [b"ACGT"[(b"ACGT".index(int(c)) + 1) % 4] for c in t[pos]]
        # Hamming distance bounds Levenshtein from above; exactly one
        # substitution is at distance exactly 1
        assert 1 <= orc.edit_distance(bytes(s), bytes(t)) <= 3
        u = s.copy()
        u[pos[0]] = t[pos[0]]
        assert orc.edit_distance(bytes(s), bytes(u)) == 1


def test_edit_eps_pairs_matches_pairwise(orc):
    rng = np.random.default_rng(8)
    strs = []
    base = bytes(rng.choice(list(b"ACGT"), 30))
    for _ in range(40):
        s = bytearray(base)
        for _ in range(rng.integers(0, 6)):
            op = rng.integers(0, 3)
            pos = rng.integers(0, len(s))
            if op == 0:
                s[pos] = b"ACGT"[rng.integers(0, 4)]
            elif op == 1:
                s.insert(pos, b"ACGT"[rng.integers(0, 4)])
            elif len(s) > 1:
                del s[pos]
        strs.append(bytes(s))
    strs.append(b"")
    atoms = np.frombuffer(b"".join(strs), np.uint8)
    offs = np.zeros(len(strs) + 1, np.int64)
    offs[1:] = np.cumsum([len(s) for s in strs])
    allp = [(i, j, _lev_matrix(strs[i], strs[j])) for i, j in
            itertools.combinations(range(len(strs)), 2)]
    for eps in (0, 2, 4, 8):
        ij, dist = orc.edit_eps_pairs(atoms, offs, eps)
        exp = [(i, j, d) for i, j, d in allp if d <= eps]
        got = [(int(i), int(j), int(d)) for (i, j), d in zip(ij, dist)]
        assert got == exp


# ---------------------------------------------------------------------------
# Voronoi assignment (include/point.hpp:535-559) and ghosts (:894-918)
# ---------------------------------------------------------------------------

def test_voronoi_assign_argmin_later_wins(orc):
    rng = np.random.default_rng(9)
    pts = rng.standard_normal((300, 6)).astype(F)
    cen = pts[rng.choice(300, 12, replace=False)].copy()
    cen = np.concatenate([cen, cen[3:4]])     # planted exact tie: center 12 == center 3
    cell, dist = orc.voronoi_assign(pts, cen)
    D = np.array([[orc.l2(c, p) for c in cen] for p in pts])
    for i in range(300):
        m = D[i].min()
        assert dist[i] == m
        assert cell[i] == np.nonzero(D[i] == m)[0].max()   # the later center wins ties
    assert 3 not in set(cell.tolist())


def _landmark_pairs(orc, pts, eps, ncent, seed):
    rng = np.random.default_rng(seed)
    cidx = np.sort(rng.choice(pts.shape[0], ncent, replace=False))
    cen = pts[cidx]
    cell, delta = orc.voronoi_assign(pts, cen)
    gh = orc.ghosts(pts, cell, delta, cen, eps)
    pairs = set()
    for c in range(ncent):
        mem = np.nonzero(cell == c)[0]
        ghosts = gh[gh[:, 0] == c][:, 1]
        for a_i, a in enumerate(mem):
            for b in mem[a_i + 1:]:
                if orc.l2(pts[a], pts[b]) <= F(eps):
                    pairs.add((min(a, b), max(a, b)))
            for g in ghosts:
                if orc.l2(pts[a], pts[g]) <= F(eps):
                    pairs.add((min(a, g), max(a, g)))
    return pairs, cell, delta, cen, gh


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_ghosts_cover_every_cross_cell_pair(orc, seed):
    rng = np.random.default_rng(100 + seed)
    pts = rng.standard_normal((400, 4)).astype(F)
    eps = F(0.6)
    pairs, cell, delta, cen, gh = _landmark_pairs(orc, pts, eps, 9, seed)
    ij, _ = orc.l2_eps_pairs(pts, eps)
    assert pairs == {tuple(map(int, p)) for p in ij}
    # every reported ghost satisfies the criterion literally
    for c, q in gh:
        assert cell[q] != c
        assert orc.l2(cen[c], pts[q]) <= F(delta[q] + F(2) * eps)
        mem = np.nonzero(cell == c)[0]
        assert any(orc.l2(pts[m], pts[q]) <= eps for m in mem)


# ---------------------------------------------------------------------------
# Simplex ids (include/simplex.hpp:5-19) and the Rips filtration
# ---------------------------------------------------------------------------

def test_binom_closed_form(orc):
    for n in range(0, 40):
        for k in range(0, 6):
            assert orc.binom(n, k) == (math.comb(n, k) if k <= n else 0)
    assert orc.binom(300000, 3) == math.comb(300000, 3)
    assert orc.binom(3, 5) == 0


def test_simplex_id_is_combinatorial_rank(orc):
    for n in (5, 9):
        for k in (1, 2, 3, 4):
            ids = sorted(orc.simplex_id(list(c)) for c in itertools.combinations(range(n), k))
            tag = (k - 1) << 60
            assert ids == [tag + r for r in range(math.comb(n, k))]
    assert orc.simplex_id([7, 2, 5]) == orc.simplex_id([2, 5, 7])
    assert orc.simplex_id([4]) == 4
    assert orc.simplex_id([1, 3]) == (1 << 60) + 1 + 3


def test_rips_complete_graph_counts(orc):
    rng = np.random.default_rng(13)
    pts = (rng.standard_normal((12, 3)) * 0.01).astype(F)
    val, ids = orc.rips_filtration(pts, 1.0, maxdim=3)
    dims = ids >> 60
    assert [(dims == p).sum() for p in range(4)] == [12, math.comb(12, 2), math.comb(12, 3),
                                                     math.comb(12, 4)]


def test_rips_matches_brute_force_triples(orc):
    rng = np.random.default_rng(14)
    pts = rng.uniform(0, 1, (45, 3)).astype(F)
    eps = F(0.35)
    val, ids = orc.rips_filtration(pts, eps, maxdim=2)
    n = pts.shape[0]
    W = {}
    for i, j in itertools.combinations(range(n), 2):
        w = orc.l2(pts[i], pts[j])
        if w <= eps:
            W[(i, j)] = w
    exp = [(F(0), v) for v in range(n)]
    exp += [(w, i + math.comb(j, 2) + (1 << 60)) for (i, j), w in W.items()]
    for i, j, k in itertools.combinations(range(n), 3):
        if (i, j) in W and (i, k) in W and (j, k) in W:
            exp.append((max(W[(i, j)], W[(i, k)], W[(j, k)]),
                        i + math.comb(j, 2) + math.comb(k, 3) + (2 << 60)))
    exp.sort()
    assert [(F(v), int(i)) for v, i in zip(val, ids)] == exp

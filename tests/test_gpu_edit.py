"""GPU parity of the edit-distance epsilon graph against the CPU oracle's
reference-order Levenshtein (bit-exact edge lists and distances)."""
import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu


def _edit():
    from rnng_amd import edit
    return edit


@pytest.mark.parametrize("eps", [0, 1, 3, 8, 8.5, 12, 20])
def test_dna_mutants_subset_bit_exact(orc, eps):
    from rnng_amd import synth
    atoms, offs = synth.dna_mutants(n_seeds=30, per_seed=40)
    flush_print(f"dna 1200 strings eps={eps}")
    ij_ref, d_ref = orc.edit_eps_pairs(atoms, offs, eps)
    ij, d = _edit().eps_pairs(atoms, offs, eps)
    assert np.array_equal(ij, ij_ref)
    assert np.array_equal(d, d_ref)
    if eps >= 8:
        assert ij.shape[0] > 1000


def test_random_alphabet_and_lengths(orc):
    rng = np.random.default_rng(3)
    strs = []
    for _ in range(400):
        L = int(rng.integers(0, 60))
        strs.append(bytes(rng.choice(list(b"abcdxyz"), L)))
    # near duplicates with edits
    for s in list(strs[:100]):
        t = bytearray(s)
        for _ in range(rng.integers(0, 5)):
            if t and rng.random() < 0.5:
                del t[rng.integers(0, len(t))]
            else:
                t.insert(int(rng.integers(0, len(t) + 1)), int(rng.choice(list(b"abcdxyz"))))
        strs.append(bytes(t))
    atoms, offs = _edit().as_csr(strs)
    for eps in (2, 6, 15, 32):
        flush_print(f"random alphabet eps={eps}")
        ij_ref, d_ref = orc.edit_eps_pairs(atoms, offs, eps)
        ij, d = _edit().eps_pairs(atoms, offs, eps)
        assert np.array_equal(ij, ij_ref)
        assert np.array_equal(d, d_ref)


def test_edge_cases(orc):
    edit = _edit()
    atoms, offs = edit.as_csr([b"", b"", b"A", b"AC", b"ACGT" * 40, b"ACGT" * 40])
    ij, d = edit.eps_pairs(atoms, offs, 1)
    assert [tuple(p) for p in ij] == [(0, 1), (0, 2), (1, 2), (2, 3), (4, 5)]
    assert list(d) == [0, 1, 1, 1, 0]
    assert edit.eps_pairs(atoms, offs, -1)[0].shape[0] == 0
    a1, o1 = edit.as_csr([b"ACGT"])
    assert edit.eps_pairs(a1, o1, 5)[0].shape[0] == 0
    # radii above 32 verify with the unbanded bit-parallel kernel
    ij_ref, d_ref = orc.edit_eps_pairs(atoms, offs, 40)
    ij, d = edit.eps_pairs(atoms, offs, 40)
    assert np.array_equal(ij, ij_ref) and np.array_equal(d, d_ref)
    # ... and strings longer than 512 atoms with the one-wave-per-pair kernel
    a2, o2 = edit.as_csr([b"A" * 513, b"A" * 500, b"A" * 540 + b"C"])
    ij, d = edit.eps_pairs(a2, o2, 40)
    assert [tuple(p) for p in ij] == [(0, 1), (0, 2)] and list(d) == [13, 28]   # (1, 2) is at 41
    assert edit.eps_pairs(a2, o2, 20)[0].shape[0] == 1


@pytest.mark.parametrize("lens,eps", [((480, 900), 40), ((480, 900), float("inf")), ((3900, 4400), 60),
                                      ((4000, 9000), 120), ((8100, 8300), float("inf"))])
def test_long_strings_radius_above_32(orc, lens, eps):
    """Strings longer than 512 atoms at radii above 32: pairs whose shorter
    string has more than 512 atoms run one wave per pair (bands of 4,096
    pattern rows, the carries between bands kept per column), the others the
    thread-per-pair kernel; pairs and distances equal the oracle's full DP.
    An infinite radius gives every pair with its exact distance."""
    lo, hi = lens
    rng = np.random.default_rng(lo + hi)
    alphabet = list(b"ACGT")
    seeds = [rng.choice(alphabet, int(rng.integers(lo, hi + 1))).astype(np.uint8).tobytes() for _ in range(3)]
    strs = []
    for s in seeds:
        strs.append(s)
        for _ in range(5):
            t = bytearray(s)
            for _ in range(int(rng.integers(0, max(2, int(eps) if eps != float("inf") else 200)))):
                r = rng.random()
                if t and r < 0.4:
                    t[int(rng.integers(0, len(t)))] = int(rng.choice(alphabet))
                elif t and r < 0.7:
                    del t[int(rng.integers(0, len(t)))]
                else:
                    t.insert(int(rng.integers(0, len(t) + 1)), int(rng.choice(alphabet)))
            strs.append(bytes(t))
    atoms, offs = _edit().as_csr(strs)
    flush_print(f"long strings {lo}-{hi} eps {eps}: {len(strs)} strings")
    ij_ref, d_ref = orc.edit_eps_pairs(atoms, offs, eps)
    ij, d = _edit().eps_pairs(atoms, offs, eps)
    assert np.array_equal(ij, ij_ref)
    assert np.array_equal(d, d_ref)
    assert ij.shape[0] > 0
    if eps == float("inf"):
        assert ij.shape[0] == len(strs) * (len(strs) - 1) // 2


@pytest.mark.parametrize("alphabet,maxlen", [(b"ACGT", 150), (b"ACGT", 300), (b"abcdefghijklmnopqrstu", 64),
                                             (b"abcdefghijklmnopqrstu", 512)])
def test_radius_above_32(orc, alphabet, maxlen):
    """Radii above 32 (beyond the banded kernels): the multiword bit-parallel
    verify gives the oracle's pairs and distances, with 1, 2, 4 or 8 words and
    2 or 5 code bit-planes; an infinite radius links every pair."""
    rng = np.random.default_rng(maxlen)
    strs = [rng.choice(list(alphabet), int(rng.integers(0, maxlen + 1))).astype(np.uint8).tobytes()
            for _ in range(120)]
    for s in list(strs[:60]):
        t = bytearray(s)
        for _ in range(int(rng.integers(0, 60))):
            r = rng.random()
            if t and r < 0.33:
                t[int(rng.integers(0, len(t)))] = int(rng.choice(list(alphabet)))
            elif t and r < 0.66:
                del t[int(rng.integers(0, len(t)))]
            elif len(t) < maxlen:
                t.insert(int(rng.integers(0, len(t) + 1)), int(rng.choice(list(alphabet))))
        strs.append(bytes(t))
    atoms, offs = _edit().as_csr(strs)
    for eps in (33, 47.5, 90, float("inf")):
        flush_print(f"radius above 32: alphabet {len(alphabet)} maxlen {maxlen} eps {eps}")
        ij_ref, d_ref = orc.edit_eps_pairs(atoms, offs, eps)
        ij, d = _edit().eps_pairs(atoms, offs, eps)
        assert np.array_equal(ij, ij_ref)
        assert np.array_equal(d, d_ref)
        assert ij.shape[0] > 0
    assert ij.shape[0] == len(strs) * (len(strs) - 1) // 2


def _clamped_profiles(atoms, offs):
    """64-bin 3-gram counts (2-bit codes (c >> 1) & 3, as the filter hashes
    them), clamped at 4."""
    code = (atoms.astype(np.int64) >> 1) & 3
    n = offs.shape[0] - 1
    P = np.zeros((n, 64), np.int64)
    for i in range(n):
        c = code[offs[i]:offs[i + 1]]
        if c.shape[0] >= 3:
            P[i] = np.bincount((c[:-2] << 4) | (c[1:-1] << 2) | c[2:], minlength=64)
    return np.minimum(P, 4)


@pytest.mark.parametrize("eps", [0, 3, 8, 12])
def test_filter_keeps_exactly_the_clamped_qgram_pairs(eps):
    """The MFMA filter's survivors are exactly the pairs inside the length
    window whose clamped 3-gram profiles are within L1 distance 6k: the
    candidate count equals that count computed on the CPU (a wrong operand
    layout or accumulator start would change it)."""
    from rnng_amd import Stats, synth
    atoms, offs = synth.dna_mutants(n_seeds=25, per_seed=80)
    rng = np.random.default_rng(4)
    extra = [bytes(rng.choice(list(b"ACGTN"), int(rng.integers(90, 160)))) for _ in range(300)]
    a2, o2 = _edit().as_csr(extra)
    atoms = np.concatenate([atoms, a2])
    offs = np.concatenate([offs, o2[1:] + offs[-1]])
    n = offs.shape[0] - 1
    Q = _clamped_profiles(atoms, offs)
    L = np.diff(offs)
    k = int(np.floor(eps))
    want = 0
    for i in range(n - 1):
        ok = np.abs(L[i + 1:] - L[i]) <= k
        want += int((np.abs(Q[i + 1:][ok] - Q[i]).sum(1) <= 6 * k).sum())
    st = Stats()
    flush_print(f"filter count n={n} eps={eps}")
    _edit().eps_pairs(atoms, offs, eps, stats=st)
    assert st.candidates == want
    assert st.candidates >= st.edges


def test_full_size_dna_properties():
    """200,000 strings, eps = 8: every reported edge is verified by the
    oracle's Levenshtein on a sample, the list is canonical, and most edges
    join mutants of the same seed."""
    import oracle
    from rnng_amd import Stats, synth
    atoms, offs = synth.dna_mutants()
    st = Stats()
    flush_print("dna 200000 strings eps=8")
    ij, d = _edit().eps_pairs(atoms, offs, 8, stats=st)
    flush_print(f"stats {st.as_dict()}")
    assert ij.shape[0] > 1_000_000
    assert np.all(ij[:, 0] < ij[:, 1])
    key = ij[:, 0] * (1 << 32) + ij[:, 1]
    assert np.all(np.diff(key) > 0)
    assert d.min() >= 0 and d.max() <= 8
    rng = np.random.default_rng(0)
    for e in rng.choice(ij.shape[0], 300, replace=False):
        i, j = ij[e]
        s = atoms[offs[i]:offs[i + 1]].tobytes()
        t = atoms[offs[j]:offs[j + 1]].tobytes()
        assert oracle.edit_distance(s, t) == d[e]


def _bipartite_ref(orc, A, B, eps):
    out = []
    for i, s in enumerate(A):
        for j, t in enumerate(B):
            if abs(len(s) - len(t)) <= eps:
                d = orc.edit_distance(s, t)
                if d <= eps:
                    out.append((i, j, d))
    if not out:
        return np.zeros((0, 2), np.int64), np.zeros(0, np.int32)
    a = np.array(out, np.int64)
    return a[:, :2], a[:, 2].astype(np.int32)


@pytest.mark.parametrize("eps", [0, 3, 8, 31, 32, 40])
def test_bipartite_matches_oracle(orc, eps):
    """Rows of one string set against columns of another (cell x ghosts, a
    single query): every (i, j) within eps, including pairs whose row string
    is the longer one, and two sets with different byte alphabets."""
    from rnng_amd import synth
    edit = _edit()
    atoms, offs = synth.dna_mutants(n_seeds=8, per_seed=30)
    S = [atoms[offs[i]:offs[i + 1]].tobytes() for i in range(offs.shape[0] - 1)]
    rng = np.random.default_rng(eps)
    A = [S[i] for i in rng.choice(len(S), 90, replace=False)]
    B = [S[i] for i in rng.choice(len(S), 70, replace=False)]
    # a few strings in another alphabet on one side only, near a DNA string
    for _ in range(6):
        t = bytearray(B[int(rng.integers(0, len(B)))])
        for _ in range(3):
            t[int(rng.integers(0, len(t)))] = ord("N")
        A.append(bytes(t))
    flush_print(f"edit bipartite {len(A)} x {len(B)} eps={eps}")
    SA, SB = edit.StringSet(*edit.as_csr(A)), edit.StringSet(*edit.as_csr(B))
    g = SA.bipartite(SB, eps)
    ij, w = g.edges(sorted=True)
    g.close()
    SA.close()
    SB.close()
    ij_ref, d_ref = _bipartite_ref(orc, A, B, eps)
    assert np.array_equal(ij, ij_ref)
    assert np.array_equal(w.astype(np.int32), d_ref)
    if eps >= 8:
        assert ij.shape[0] > 50

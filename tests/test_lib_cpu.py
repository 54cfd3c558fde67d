"""CPU-side checks of the product library: it loads, it exports every symbol
the C header declares, host-only helpers agree with the oracle, error paths
return codes instead of crashing, and the seeded generators are deterministic.
No GPU compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    syms = set()
    for name in os.listdir(os.path.join(ROOT, "include")):
        if name.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", name)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            syms |= set(re.findall(r"\b(rnng_[a-z0-9_]+)\s*\(", txt))
    return syms


def test_library_exports_every_declared_symbol():
    from rnng_amd import _lib
    L = _lib.lib()
    syms = _header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    # the ctypes table covers the header exactly
    assert set(_lib.declared_symbols()) == syms


def test_library_has_gfx950_code_object():
    from rnng_amd import _lib
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_sq_threshold_matches_oracle(orc):
    from rnng_amd import l2
    rng = np.random.default_rng(0)
    for e in list(rng.uniform(0, 300, 300)) + [0.0, 0.08, 1.5, 256.2889, -1.0, 1e19]:
        assert l2.sq_threshold(e) == orc.sq_threshold(e)


def test_no_device_is_an_error_code_not_a_crash():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    from rnng_amd import _lib
    L = _lib.lib()
    x = np.zeros((4, 3), np.float32)
    h = C.c_void_p()
    rc = L.rnng_dataset_create_f32(x.ctypes.data_as(C.c_void_p), 4, 3, C.byref(h))
    assert rc == _lib.RNNG_ERR_NODEV
    assert L.rnng_last_error()


def test_bad_arguments_are_rejected():
    from rnng_amd import _lib
    L = _lib.lib()
    h = C.c_void_p()
    assert L.rnng_dataset_create_f32(None, 4, 3, C.byref(h)) == _lib.RNNG_ERR_ARG
    assert L.rnng_dataset_create_f32(None, 0, 0, C.byref(h)) == _lib.RNNG_ERR_ARG
    assert L.rnng_graph_num_edges(None, None) == _lib.RNNG_ERR_ARG
    assert L.rnng_set_initial_capacity(-1, 0) == _lib.RNNG_ERR_ARG
    assert L.rnng_set_initial_capacity(0, 0) == _lib.RNNG_OK


def test_generators_are_deterministic():
    from rnng_amd import synth
    a1, e1 = synth.points_gmm_small()
    a2, e2 = synth.points_gmm_small()
    assert a1.tobytes() == a2.tobytes() and e1 == e2
    assert a1.shape == (10_000, 16) and a1.dtype == np.float32
    p, e = synth.points_gmm_int(n=5000)
    assert p.min() >= 0 and p.max() <= 255 and np.all(p == np.rint(p))
    atoms, offs = synth.dna_mutants(n_seeds=20, per_seed=10)
    assert offs.shape == (201,) and set(np.unique(atoms).tobytes()) <= set(b"ACGT")
    lens = np.diff(offs)
    assert lens.min() >= 80 and lens.max() <= 170
    r, eps = synth.points_rips(n=1000)
    assert r.shape == (1000, 3) and eps == np.float32(0.08)


def test_xcd_permute_pads_with_empty_items_and_keeps_every_item():
    """The XCD-ordered work list (host code, no device): every input item once,
    the length a multiple of 8, every padding item empty (first column tile
    not below the end tile, which each filter kernel skips before touching
    memory), and XCD x's workgroups (every 8th, from x) a contiguous run of the
    input in order."""
    from rnng_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(3)
    for n in list(range(0, 20)) + [63, 64, 65, 1000, 4097]:
        rb = rng.integers(0, 50, n)
        lo = rng.integers(0, 100, n)
        items = np.stack([rb, lo, lo + rng.integers(1, 9, n), np.zeros(n, np.int64)], 1).astype(np.int32)
        items = np.ascontiguousarray(items)
        cap = 8 * ((n + 7) // 8)
        out = np.full((max(cap, 1), 4), -9, np.int32)
        n_out = C.c_int64()
        assert L.rnng_debug_xcd_permute(items.ctypes.data_as(C.c_void_p), n, out.ctypes.data_as(C.c_void_p), cap,
                                        C.byref(n_out)) == 0
        m = n_out.value
        assert m == cap and m % 8 == 0
        out = out[:m]
        real = out[:, 1] < out[:, 2]
        assert int(real.sum()) == n
        assert (out[~real, 1] >= out[~real, 2]).all()
        assert sorted(map(tuple, out[real].tolist())) == sorted(map(tuple, items.tolist()))
        per = m // 8
        for x in range(8):
            share = out[x::8]
            share = share[share[:, 1] < share[:, 2]]
            assert np.array_equal(share, items[x * per: x * per + share.shape[0]])
        if n:
            assert L.rnng_debug_xcd_permute(items.ctypes.data_as(C.c_void_p), n, out.ctypes.data_as(C.c_void_p),
                                            m - 1, C.byref(n_out)) == _lib.RNNG_ERR_CAPACITY


def cell_work(L, cnt, adj, perm, bt, sched, rb_cand=None, device=False):
    """The cell filter's work lists of one layout: (start, items, ctl, covered)
    from the host builder or the device builder."""
    m = cnt.shape[0]
    start = np.zeros(m, np.int32)
    cap = 1 << 22
    items = np.zeros((cap, 4), np.int32)
    ctl = np.zeros(cap, np.int32)
    ni, nc, cov = C.c_int64(), C.c_int64(), C.c_int64()
    fn = L.rnng_debug_cell_work_device if device else L.rnng_debug_cell_work
    rbc = None if rb_cand is None else np.ascontiguousarray(rb_cand, np.float32)
    rc = fn(cnt.ctypes.data_as(C.c_void_p), m, adj.ctypes.data_as(C.c_void_p), perm.ctypes.data_as(C.c_void_p), bt,
            sched, None if rbc is None else rbc.ctypes.data_as(C.c_void_p), start.ctypes.data_as(C.c_void_p),
            items.ctypes.data_as(C.c_void_p), cap, ctl.ctypes.data_as(C.c_void_p), cap,
            C.byref(ni), C.byref(nc), C.byref(cov))
    assert rc == 0, L.rnng_last_error()
    return start, items[:ni.value], ctl[:nc.value], cov.value


def random_cells(m, seed, density=0.15):
    """m cells (sizes 0..399, one empty) and a symmetric random adjacency
    bitset without the empty cells."""
    rng = np.random.default_rng(seed)
    cnt = rng.integers(0, 400, m).astype(np.int32)
    cnt[rng.integers(0, m)] = 0   # an empty cell
    words = (m + 31) // 32
    A = rng.random((m, m)) < density
    A = A | A.T
    np.fill_diagonal(A, False)
    A[cnt == 0, :] = False
    A[:, cnt == 0] = False
    adj = np.zeros((m, words), np.uint32)
    for a in range(m):
        for b in np.nonzero(A[a])[0]:
            adj[a, b >> 5] |= np.uint32(1) << np.uint32(b & 31)
    return cnt, adj, rng


_cell_work = cell_work


def needed_tiles(cnt, adj, perm, bt):
    """(row block, column tile) pairs the cell filter must run, from the
    definition: row block rb holds tiles rb*bt .. rb*bt+bt-1 of the layout
    (cells in position order perm, each padded to 32 rows); it needs every
    tile from its own first tile on whose cell is one of its cells or adjacent
    to one of them (every position between its first and last cell counts as
    its own)."""
    m = cnt.shape[0]
    pad = (cnt + 31) // 32 * 32
    rs = np.concatenate([[0], np.cumsum(pad[perm])])
    ntiles = int(rs[-1] // 32)
    tcell = np.zeros(max(ntiles, 1), np.int64)
    for k in range(m):
        tcell[rs[k] // 32: rs[k + 1] // 32] = k
    pos = np.empty(m, np.int64)
    pos[perm] = np.arange(m)
    nbr = [set(pos[c] for c in range(m) if (adj[a, c >> 5] >> (c & 31)) & 1) for a in range(m)]
    out = set()
    for rb in range((ntiles + bt - 1) // bt):
        t0, t1 = rb * bt, min(ntiles, rb * bt + bt)
        k0, k1 = int(tcell[t0]), int(tcell[t1 - 1])
        need = set(range(k0, k1 + 1))
        for k in range(k0, k1 + 1):
            need |= nbr[perm[k]]
        for t in range(t0, ntiles):
            if tcell[t] in need:
                out.add((rb, t))
    return out


@pytest.mark.parametrize("m,bt,seed", [(7, 8, 0), (130, 8, 1), (300, 24, 2), (64, 8, 3)])
def test_cell_tile_lists_cover_the_runs_pairs(m, bt, seed):
    """The cell filter's tile-list work (host code, no device) covers exactly
    the (row block, column tile) pairs of the runs work it replaces, every
    list entry is a tile of the layout and every item indexes its row block's
    list; the layout puts each cell at a 32-aligned start in position order."""
    from rnng_amd import _lib
    L = _lib.lib()
    cnt, adj, rng = random_cells(m, seed)
    for perm in (np.arange(m, dtype=np.int32), rng.permutation(m).astype(np.int32)):
        start, runs, _, cov0 = _cell_work(L, cnt, adj, perm, bt, 0)
        pad = (cnt + 31) // 32 * 32
        rs = np.concatenate([[0], np.cumsum(pad[perm])])
        assert np.array_equal(start[perm], rs[:-1])
        ntiles = rs[-1] // 32
        want = set()
        for rb, lo, hi, _ in runs:
            assert 0 <= lo < hi <= ntiles
            want |= {(int(rb), t) for t in range(lo, hi)}
        assert want == needed_tiles(cnt, adj, perm, bt)
        for sched in (1, 2):
            _, items, ctl, cov = _cell_work(L, cnt, adj, perm, bt, sched)
            assert cov == cov0
            assert ((ctl >= 0) & (ctl < ntiles)).all()
            got = []
            for rb, p0, p1, _ in items:
                assert 0 <= p0 < p1 <= ctl.shape[0] and p1 - p0 <= 256
                assert rb * bt * 32 < rs[-1]
                got += [(int(rb), int(t)) for t in ctl[p0:p1]]
            assert len(got) == len(set(got)) and set(got) == want
            if sched == 2:
                sizes = items[:, 2] - items[:, 1]
                assert (np.diff(sizes) <= 0).all()

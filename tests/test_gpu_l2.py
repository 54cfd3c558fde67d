"""GPU parity of the L2 epsilon-graph path: the HIP kernels (plain FP32 and
BF16 MFMA filter + FP32 recheck), called through the C ABI, against the CPU
oracle on the same seeded inputs.  Integer/index results are bit-exact: the
sorted (i, j) lists are equal and the float weights are equal bit for bit."""
import ctypes as C

import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu
F = np.float32


def _l2():
    from rnng_amd import l2
    return l2


def _check_same(ij, w, ij_ref, w_ref):
    assert ij.shape == ij_ref.shape, (ij.shape, ij_ref.shape)
    assert np.array_equal(ij, ij_ref)
    assert w.tobytes() == w_ref.tobytes()


def test_mfma_output_layout():
    rows, cols = _l2().mfma_layout()
    lane = np.arange(64)[:, None]
    r = np.arange(16)[None, :]
    assert np.array_equal(cols, np.broadcast_to(lane % 32, (64, 16)))
    assert np.array_equal(rows, 8 * (r // 4) + 4 * (lane // 32) + r % 4)


def test_device_sqrt_is_correctly_rounded():
    """The edge weight is sqrtf of the FP32 squared distance; the device root
    must be the correctly rounded one (numpy's float32 sqrt is IEEE)."""
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 0x7f800000, size=4_000_000, dtype=np.uint32)
    x = bits.view(np.float32)
    dense = np.arange(65000.0, 66500.0, 1.0 / 64, dtype=np.float32)   # squared radii near 256
    small = np.nextafter(np.float32(0), np.float32(1)) * np.arange(1, 5000, dtype=np.float32)
    x = np.concatenate([x, dense, small, np.float32([0.0, 1.0, 2.0, 2.25, 16777216.0])])
    got = _l2().device_sqrt(x)
    assert np.array_equal(got.view(np.uint32), np.sqrt(x).view(np.uint32))


@pytest.mark.parametrize("mode", ["plain", "mfma"])
def test_small_gmm_config_bit_exact(orc, mode):
    from rnng_amd import synth
    pts, eps = synth.points_gmm_small()
    flush_print(f"gmm_small 10000x16 eps={eps} mode={mode}")
    ij_ref, w_ref = orc.l2_eps_pairs(pts, eps)
    ij, w = _l2().eps_pairs(pts, eps, mode)
    _check_same(ij, w, ij_ref, w_ref)
    assert ij.shape[0] > 50_000


@pytest.mark.parametrize("n,d", [(1, 3), (2, 3), (33, 3), (1000, 16), (3001, 17), (2500, 37),
                                 (2000, 96), (1800, 128), (700, 130), (1, 64), (2, 64), (33, 64), (65, 128)])
@pytest.mark.parametrize("mode", ["plain", "mfma"])
def test_random_shapes_bit_exact(orc, n, d, mode):
    rng = np.random.default_rng(n * 7 + d)
    pts = (rng.standard_normal((n, d)) * 3 + 1).astype(F)
    eps = F(np.sqrt(2 * d) * 3 * 0.8)
    flush_print(f"random n={n} d={d} mode={mode}")
    ij_ref, w_ref = orc.l2_eps_pairs(pts, eps)
    ij, w = _l2().eps_pairs(pts, eps, mode)
    _check_same(ij, w, ij_ref, w_ref)


@pytest.mark.parametrize("n", [1, 2, 31, 33, 64, 513])
@pytest.mark.parametrize("d", [64, 128])
def test_int8_small_and_ragged_sizes(orc, n, d):
    """The int8 filter (16x16 forms at an even k-step count) at sizes below,
    at and past one 32-point tile, self join and bipartite against a ragged
    second set."""
    rng = np.random.default_rng(n * 3 + d)
    X = rng.integers(0, 60, size=(n, d)).astype(F)
    Y = rng.integers(0, 60, size=(n + 7, d)).astype(F)
    if n > 1:
        X[n - 1] = X[0]
    eps = F(np.sqrt(600.0 * d) * 0.95)   # about a tenth of the pairs
    flush_print(f"int8 small n={n} d={d}")
    ij_ref, w_ref = orc.l2_eps_pairs(X, eps)
    ij, w = _l2().eps_pairs(X, eps, "mfma")
    _check_same(ij, w, ij_ref, w_ref)
    ijb_ref, wb_ref = orc.l2_eps_pairs(X, eps, Y)
    ijb, wb = _l2().eps_bipartite(X, Y, eps, "mfma")
    _check_same(ijb, wb, ijb_ref, wb_ref)


@pytest.mark.parametrize("mode", ["plain", "mfma"])
def test_integer_data_and_band(orc, mode):
    from rnng_amd import synth
    pts, eps = synth.points_gmm_int(n=20_000)
    flush_print(f"gmm_int 20000x128 eps={eps} mode={mode}")
    ij_ref, w_ref = orc.l2_eps_pairs(pts, eps)
    from rnng_amd import Stats
    st = Stats()
    ij, w = _l2().eps_pairs(pts, eps, mode, stats=st)
    _check_same(ij, w, ij_ref, w_ref)
    if mode == "mfma":
        assert st.candidates >= st.edges
        flush_print(f"stats {st.as_dict()}")


@pytest.mark.parametrize("mode", ["plain", "mfma"])
def test_bipartite_bit_exact(orc, mode):
    rng = np.random.default_rng(21)
    A = rng.standard_normal((1500, 24)).astype(F)
    B = rng.standard_normal((2300, 24)).astype(F)
    eps = F(5.2)
    flush_print(f"bipartite 1500x2300 d=24 mode={mode}")
    ij_ref, w_ref = orc.l2_eps_pairs(A, eps, B)
    ij, w = _l2().eps_bipartite(A, B, eps, mode)
    _check_same(ij, w, ij_ref, w_ref)


@pytest.mark.parametrize("mode", ["plain", "mfma"])
def test_edge_cases(orc, mode):
    l2 = _l2()
    rng = np.random.default_rng(5)
    X = rng.standard_normal((300, 8)).astype(F)
    X[7] = X[3]
    X[250] = X[3]
    flush_print(f"edge cases mode={mode}")
    # eps = 0: only exact duplicates
    ij, w = l2.eps_pairs(X, 0.0, mode)
    assert [tuple(p) for p in ij] == [(3, 7), (3, 250), (7, 250)]
    assert np.all(w == 0)
    # negative eps: nothing
    assert l2.eps_pairs(X, -1.0, mode)[0].shape[0] == 0
    # huge eps: every pair
    ij, w = l2.eps_pairs(X[:100], 1e6, mode)
    assert ij.shape[0] == 100 * 99 // 2
    # empty and single-point sets
    assert l2.eps_pairs(np.zeros((0, 8), F), 1.0, mode)[0].shape[0] == 0
    assert l2.eps_pairs(X[:1], 1.0, mode)[0].shape[0] == 0


@pytest.mark.parametrize("mode", ["plain", "mfma"])
def test_count_then_fill_refill(orc, mode):
    from rnng_amd import Stats, _lib
    l2 = _l2()
    rng = np.random.default_rng(6)
    X = rng.standard_normal((2000, 16)).astype(F)
    eps = F(4.0)
    ij_ref, w_ref = orc.l2_eps_pairs(X, eps)
    assert ij_ref.shape[0] > 5000
    flush_print(f"refill mode={mode}")
    _lib.check(_lib.lib().rnng_set_initial_capacity(100, 100))
    try:
        st = Stats()
        ij, w = l2.eps_pairs(X, eps, mode, stats=st)
        assert st.refills >= 1
        _check_same(ij, w, ij_ref, w_ref)
    finally:
        _lib.check(_lib.lib().rnng_set_initial_capacity(0, 0))
    # the one-shot host entry point reports the size when cap is too small
    ij2, w2 = l2.eps_pairs_oneshot(X, eps, mode)
    _check_same(ij2, w2, ij_ref, w_ref)
    L = _lib.lib()
    needed = C.c_int64()
    buf = np.empty((10, 2), np.int64)
    rc = L.rnng_l2_eps_pairs_f32(X.ctypes.data_as(C.c_void_p), 2000, 16, eps, l2._mode(mode),
                                 buf.ctypes.data_as(C.c_void_p), None, 10, C.byref(needed), None)
    assert rc == _lib.RNNG_ERR_CAPACITY and needed.value == ij_ref.shape[0]


def test_full_size_int_config_mfma_equals_plain():
    """1,000,000 x 128 integer workload: the MFMA path's edge set equals the
    plain FP32 kernel's and edges are canonical (i < j, sorted, unique)."""
    from rnng_amd import Stats, synth
    l2 = _l2()
    pts, eps = synth.points_gmm_int()
    flush_print(f"gmm_int 1Mx128 eps={eps}")
    ds = l2.Dataset(pts)
    s1, s2 = Stats(), Stats()
    g1 = ds.graph(eps, "mfma", s1)
    g2 = ds.graph(eps, "plain", s2)
    ij1, w1 = g1.edges()
    ij2, w2 = g2.edges()
    flush_print(f"mfma {s1.as_dict()}\nplain {s2.as_dict()}")
    _check_same(ij1, w1, ij2, w2)
    assert np.all(ij1[:, 0] < ij1[:, 1])
    key = ij1[:, 0] * (1 << 32) + ij1[:, 1]
    assert np.all(np.diff(key) > 0)
    assert 5_000_000 < ij1.shape[0] < 50_000_000
    assert s1.candidates >= s1.edges
    # spot-check a sample of edges with the oracle's distance
    import oracle
    rng = np.random.default_rng(0)
    for e in rng.choice(ij1.shape[0], 200, replace=False):
        i, j = ij1[e]
        assert oracle.l2(pts[i], pts[j]) == w1[e]


@pytest.mark.parametrize("loops", [False, True])
def test_symmetric_csr(orc, loops):
    """rnng_graph_copy_csr: rows = vertices, both orientations of every edge
    (plus the zero loops if asked), columns ascending, weights with them;
    expanded it equals the directed copy's list."""
    from rnng_amd import _lib
    rng = np.random.default_rng(31)
    X = rng.standard_normal((2500, 8)).astype(F)
    X[100] = X[7]
    eps = F(2.2)
    ij_ref, w_ref = orc.l2_eps_pairs(X, eps)
    n = X.shape[0]
    u = np.concatenate([ij_ref[:, 0], ij_ref[:, 1]] + ([np.arange(n)] if loops else []))
    v = np.concatenate([ij_ref[:, 1], ij_ref[:, 0]] + ([np.arange(n)] if loops else []))
    w = np.concatenate([w_ref, w_ref] + ([np.zeros(n, F)] if loops else []))
    o = np.lexsort((v, u))
    ds = _l2().Dataset(X)
    g = ds.graph(eps, "mfma")
    rowptr, col, ww = g.csr(n, loops)
    assert rowptr[0] == 0 and rowptr[-1] == u.shape[0]
    assert np.array_equal(np.repeat(np.arange(n), np.diff(rowptr)), u[o])
    assert np.array_equal(col, v[o])
    assert ww.tobytes() == w[o].tobytes()
    # an edge beyond n_vertices is refused
    with pytest.raises(_lib.RnngError) as ei:
        g.csr(50, loops)
    assert ei.value.code == _lib.RNNG_ERR_ARG
    g.close()
    ds.close()


@pytest.mark.parametrize("d,lo,hi", [(3, 0, 255), (17, -100, 155), (100, 0, 255), (128, 0, 255),
                                     (144, 7, 262), (64, 0, 256), (33, 0, 40)])
def test_int8_filter_bit_exact(orc, d, lo, hi):
    """Integer data spanning at most 255 per dimension runs the exact int8
    filter (int8_filter counted); a span of 256 falls back to BF16; both give
    the oracle's edges, also with the int8 path switched off."""
    from rnng_amd import Stats, _lib
    rng = np.random.default_rng(d * 13 + hi)
    n = 3000
    X = rng.integers(lo, hi + 1, size=(n, d)).astype(F)
    X[:, 0] = lo
    X[1, 0] = hi
    X[77] = X[12]
    D2 = ((X[:40, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(2)
    eps = F(np.sqrt(np.quantile(D2[D2 > 0], 0.002)))
    flush_print(f"int8 filter d={d} range [{lo},{hi}] eps={eps}")
    ij_ref, w_ref = orc.l2_eps_pairs(X, eps)
    assert ij_ref.shape[0] > n // 10
    eligible = hi - lo <= 255
    st = Stats()
    ij, w = _l2().eps_pairs(X, eps, "mfma", stats=st)
    _check_same(ij, w, ij_ref, w_ref)
    assert st.int8_filter == (1 if eligible else 0)
    Y = X[: n // 2] + 0
    ijb_ref, wb_ref = orc.l2_eps_pairs(X, eps, Y)
    stb = Stats()
    ijb, wb = _l2().eps_bipartite(X, Y, eps, "mfma", stats=stb)
    _check_same(ijb, wb, ijb_ref, wb_ref)
    assert stb.int8_filter == (1 if eligible else 0)
    _lib.check(_lib.lib().rnng_set_int8_filter(0))
    try:
        st2 = Stats()
        ij2, w2 = _l2().eps_pairs(X, eps, "mfma", stats=st2)
        _check_same(ij2, w2, ij_ref, w_ref)
        assert st2.int8_filter == 0
    finally:
        _lib.check(_lib.lib().rnng_set_int8_filter(1))


def test_every_filter_kernel_shape(orc):
    """Every selectable filter kernel (int8 codes 0-12, BF16 variants 0-19, set
    through rnng_set_filter_kernels) gives the oracle's edges bit for bit, on
    integer data (the int8 path) and float data (the BF16 path), for self joins
    and a bipartite join, at d = 17 (scalar loads, partial k-steps), 64 and 128.
    The 16x16 kernels (int8 codes 9-12; BF16 variants 18, 19) run on the
    16-row fragment layout where the k-step count is even (int8 at d = 64 and
    128, BF16 at every d here) and as their 32x32 forms otherwise."""
    from rnng_amd import Stats, _lib
    lib = _lib.lib()
    code0, var0 = C.c_int32(), C.c_int32()
    _lib.check(lib.rnng_get_filter_kernels(C.byref(code0), C.byref(var0)))
    rng = np.random.default_rng(77)

    def case(X):
        D2 = ((X[:40, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(2)
        eps = F(np.sqrt(np.quantile(D2[D2 > 0], 0.003)))
        Y = X[:1100] + 0
        return X, Y, eps, orc.l2_eps_pairs(X, eps), orc.l2_eps_pairs(X, eps, Y)

    ints, floats = [], []
    for d in (17, 64, 128):
        Xi = rng.integers(0, 200, size=(2600, d)).astype(F)
        Xi[5] = Xi[9]
        ints.append(case(Xi))
        floats.append(case(rng.standard_normal((2600, d)).astype(F)))

    def check(cases, want_int8):
        for X, Y, eps, (ij_ref, w_ref), (ijb_ref, wb_ref) in cases:
            st = Stats()
            ij, w = _l2().eps_pairs(X, eps, "mfma", stats=st)
            _check_same(ij, w, ij_ref, w_ref)
            assert st.int8_filter == want_int8
            ijb, wb = _l2().eps_bipartite(X, Y, eps, "mfma")
            _check_same(ijb, wb, ijb_ref, wb_ref)

    try:
        for code in range(13):
            flush_print(f"int8 filter kernel code {code}")
            _lib.check(lib.rnng_set_filter_kernels(code, -1))
            check(ints, 1)
        for var in range(20):
            flush_print(f"BF16 filter variant {var}")
            _lib.check(lib.rnng_set_filter_kernels(-1, var))
            check(floats, 0)
        assert lib.rnng_set_filter_kernels(13, -1) != 0 and lib.rnng_set_filter_kernels(-1, 20) != 0
    finally:
        _lib.check(lib.rnng_set_filter_kernels(code0.value, var0.value))


def test_int8_filter_eps_extremes(orc):
    from rnng_amd import Stats
    rng = np.random.default_rng(3)
    X = rng.integers(0, 256, size=(500, 16)).astype(F)
    X[10] = X[3]
    X[400] = X[3]
    for eps, want in ((0.0, [(3, 10), (3, 400), (10, 400)]), (-1.0, [])):
        st = Stats()
        ij, w = _l2().eps_pairs(X, eps, "mfma", stats=st)
        assert [tuple(p) for p in ij] == want and st.int8_filter == 1
    st = Stats()
    ij, w = _l2().eps_pairs(X[:200], 1e9, "mfma", stats=st)
    assert ij.shape[0] == 200 * 199 // 2 and st.int8_filter == 1
    ij_ref, w_ref = orc.l2_eps_pairs(X[:200], F(1e9))
    _check_same(ij, w, ij_ref, w_ref)


def test_nonfinite_and_huge_coordinates(orc):
    """The MFMA filters bound distances from centered BF16 or int8 forms, which
    needs finite coordinates whose squared differences stay far from FP32
    overflow.  Other inputs (NaN, infinities, a span of 1e20) run the plain
    kernel, which follows the reference's FP32 arithmetic for any input: a NaN
    distance is never an edge, an infinite one only for eps = inf.  The cell
    path refuses such inputs with an error."""
    from rnng_amd import Stats, _lib
    from rnng_amd.l2 import Dataset
    from rnng_amd.voronoi import Voronoi
    rng = np.random.default_rng(21)

    def eps_for(M):
        S = M[:200].astype(np.float64)
        D2 = ((S[:, None, :] - S[None, :, :]) ** 2).sum(2)
        return F(np.sqrt(np.quantile(D2[D2 > 0], 0.01)))

    X = rng.standard_normal((2000, 16)).astype(F)
    Xi = rng.integers(0, 256, size=(2000, 24)).astype(F)
    eps_x, eps_i = eps_for(X), eps_for(Xi)
    A = X.copy()
    A[5, 0] = np.nan
    A[17, 3] = A[18, 3] = np.inf
    A[40, 1] = -np.inf
    Bi = Xi.copy()
    Bi[7, 2] = Bi[9, 2] = np.inf          # integer-valued except two infinities
    H = X.copy()
    H[:50] *= F(1e20)                      # squares of this span overflow FP32
    for name, M, eps in (("nan/inf", A, eps_x), ("int+inf", Bi, eps_i), ("huge", H, eps_x),
                         ("inf eps", A[:300], F(np.inf))):
        flush_print(f"non-finite coordinates: {name}")
        st = Stats()
        ij, w = _l2().eps_pairs(M, eps, "mfma", stats=st)
        ij_ref, w_ref = orc.l2_eps_pairs(M, eps)
        _check_same(ij, w, ij_ref, w_ref)
        assert st.plain_fallbacks == 1 and st.int8_filter == 0 and st.candidates == 0
        assert ij.shape[0] > 0
    assert not np.isin([5, 17, 18, 40], _l2().eps_pairs(A, eps_x, "mfma")[0]).any()
    st = Stats()
    ij, w = _l2().eps_bipartite(A[:1000], A[1000:], eps_x, "mfma", stats=st)
    ij_ref, w_ref = orc.l2_eps_pairs(A[:1000], eps_x, A[1000:])
    _check_same(ij, w, ij_ref, w_ref)
    assert st.plain_fallbacks == 1
    st = Stats()
    ij, w = _l2().eps_pairs(X, eps_x, "mfma", stats=st)
    assert st.plain_fallbacks == 0 and st.candidates > 0
    flush_print("non-finite coordinates: cell path")
    ds = Dataset(A)
    v = Voronoi(ds, fps=16)
    try:
        with pytest.raises(_lib.RnngError):
            v.graph(eps_x).close()
    finally:
        v.close()
        ds.close()


def test_mfma_mode_above_d144_runs_plain(orc):
    """The filters hold row fragments in registers up to d = 144; MFMA mode
    on wider points (d = 200 here, like GIST-style vectors) runs the plain
    kernel and gives the oracle's edges."""
    from rnng_amd import Stats
    rng = np.random.default_rng(200)
    X = rng.standard_normal((1500, 200)).astype(F)
    X[9] = X[4]
    D2 = ((X[:40, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(2)
    eps = F(np.sqrt(np.quantile(D2[D2 > 0], 0.01)))
    flush_print(f"d=200 mfma mode eps={eps}")
    st = Stats()
    ij, w = _l2().eps_pairs(X, eps, "mfma", stats=st)
    ij_ref, w_ref = orc.l2_eps_pairs(X, eps)
    _check_same(ij, w, ij_ref, w_ref)
    assert st.plain_fallbacks == 1 and ij.shape[0] > 100

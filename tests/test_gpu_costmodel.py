"""The cost model's own kernels against the float64 oracle (forward,
mean-squared-error gradient), training on the GPU, and what the trained model
gains over count-then-fill in the Voronoi cell build: one filter pass instead
of two, with the identical edge set."""
import numpy as np
import pytest

from conftest import flush_print

pytestmark = pytest.mark.gpu


def _cm():
    from rnng_amd import costmodel
    return costmodel


def test_infer_matches_oracle(orc):
    cm = _cm()
    rng = np.random.default_rng(1)
    p = cm.init_params(3)
    p[256:320] = rng.standard_normal(64) * 0.1     # non-zero biases
    x = rng.uniform(0, 1, size=(3000, 4)).astype(np.float32)
    y = cm.infer(p, x)
    ref = orc.mlp_forward(p.astype(np.float64), x.astype(np.float64))
    assert np.allclose(y, ref, rtol=1e-4, atol=1e-5)


@pytest.mark.parametrize("batch,under", [(1, 4.0), (31, 4.0), (32, 4.0), (1000, 4.0), (4133, 4.0), (1000, 1.0),
                                         (4133, 2.5)])
def test_loss_grad_matches_oracle(orc, batch, under):
    cm = _cm()
    rng = np.random.default_rng(batch)
    p = cm.init_params(batch)
    p += (rng.standard_normal(p.size) * 0.05).astype(np.float32)
    x = rng.uniform(0, 1, size=(batch, 4)).astype(np.float32)
    t = rng.uniform(0, 1, size=(batch, 2)).astype(np.float32)
    w0 = cm.under_weight()
    try:
        cm.set_under_weight(under)
        loss, g = cm.loss_grad(p, x, t)
    finally:
        cm.set_under_weight(w0)
    loss_ref, g_ref = orc.mlp_loss_grad(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64),
                                        under=under)
    assert np.isclose(loss, loss_ref, rtol=1e-4)
    # FP32 sums over the batch against float64: tolerance scaled by the gradient size
    tol = 1e-4 * np.abs(g_ref).max() + 1e-7
    assert np.max(np.abs(g - g_ref)) <= tol, np.max(np.abs(g - g_ref)) / np.abs(g_ref).max()


def test_train_fits_a_smooth_target():
    cm = _cm()
    rng = np.random.default_rng(7)
    x = rng.uniform(0, 1, size=(4000, 4)).astype(np.float32)
    t = np.stack([0.5 * x[:, 0] + 0.3 * x[:, 1] * x[:, 2], np.sqrt(x[:, 3]) * 0.4 + 0.1], 1).astype(np.float32)
    p0 = cm.init_params(0)
    l0, _ = cm.loss_grad(p0, x, t)
    p, l1 = cm.train(p0, x, t, iters=800, lr=3e-3)
    flush_print(f"train: loss {l0:.3e} -> {l1:.3e}")
    assert l1 < 0.02 * l0 and l1 < 2e-4
    # the returned loss is the loss of the returned parameters
    l2, _ = cm.loss_grad(p, x, t)
    assert np.isclose(l1, l2, rtol=1e-5)


def _build(pts, eps, m, seed, model):
    from rnng_amd import l2, voronoi
    cm = _cm()
    cm.set_cost_model(model)
    ds = l2.Dataset(pts)
    v = voronoi.Voronoi(ds, centers=voronoi.selection_sample(pts.shape[0], m, seed))
    g = v.graph(eps)
    ij, w = g.edges(sorted=True)
    st = g.stats
    g.close()
    v.close()
    ds.close()
    cm.set_cost_model(None)
    return ij, w, st


def test_cost_model_saves_the_counting_pass():
    """Trained on per-cell counts of 48 synthetic runs, the model sizes the
    candidate buffer of held-out runs: the build runs the filter once where
    count-then-fill runs it twice, and the edge set is identical."""
    cm = _cm()
    X, T = cm.collect(runs=48, seed=11)
    p, loss = cm.train(cm.init_params(2), X, T, iters=2500)
    flush_print(f"collected {X.shape[0]} cells; training loss {loss:.3e}")
    rng = np.random.default_rng(99)
    saved = 0
    ratios = []
    for r in range(6):
        pts, eps, m, _ = cm.synthetic_run(rng)
        ij0, w0, s0 = _build(pts, eps, m, r, None)
        ij1, w1, s1 = _build(pts, eps, m, r, p)
        assert np.array_equal(ij0, ij1) and w0.tobytes() == w1.tobytes()
        assert s0.filter_passes == 2 and s0.predicted_candidates == 0
        ratio = s1.predicted_candidates / max(1, s1.candidates)
        ratios.append(ratio)
        flush_print(f"held-out {r}: n={pts.shape[0]} d={pts.shape[1]} m={m} candidates={s1.candidates} "
                    f"predicted={s1.predicted_candidates} passes={s1.filter_passes}")
        saved += s1.filter_passes == 1
    assert saved >= 4, ratios
    assert np.median(ratios) < 4.0, ratios


def test_shipped_model_sizes_held_out_runs():
    """The committed model (trained on 256 runs) on fresh synthetic runs of the
    same families, including 100k+ point runs."""
    cm = _cm()
    p = cm.load()
    if p is None:
        pytest.skip("no shipped cost model")
    rng = np.random.default_rng(2024)
    saved = 0
    for r in range(5):
        pts, eps, m, _ = cm.synthetic_run(rng, big=0.4)
        ij0, w0, s0 = _build(pts, eps, m, r, None)
        ij1, w1, s1 = _build(pts, eps, m, r, p)
        assert np.array_equal(ij0, ij1) and w0.tobytes() == w1.tobytes()
        flush_print(f"shipped model, run {r}: n={pts.shape[0]} m={m} candidates={s1.candidates} "
                    f"predicted={s1.predicted_candidates} edges={s1.edges} "
                    f"predicted_edges={s1.predicted_edges} passes={s1.filter_passes}")
        saved += s1.filter_passes == 1
    assert saved >= 4


def test_wild_model_outputs_keep_the_edge_set():
    """A model whose outputs are NaN sizes nothing (count-then-fill, two
    passes); one whose outputs sit at the clamp (2^36 candidates per cell) is
    capped by the pairs the work list covers (one pass).  Either way the edges
    equal the plain kernel's."""
    from rnng_amd import l2
    cm = _cm()
    rng = np.random.default_rng(5)
    pts = (rng.standard_normal((6000, 24)) + 4 * rng.integers(0, 6, size=(6000, 1))).astype(np.float32)
    eps = np.float32(2.2)
    ij_ref, w_ref = l2.eps_pairs(pts, eps, "plain")
    nan_model = np.full(cm.NPARAMS, np.nan, np.float32)
    huge_model = np.zeros(cm.NPARAMS, np.float32)
    huge_model[-2:] = 10.0   # output biases: every cell at the clamp
    for name, model, passes in (("nan", nan_model, 2), ("huge", huge_model, 1)):
        flush_print(f"cost model outputs: {name}")
        ij, w, st = _build(pts, eps, 24, 3, model)
        assert np.array_equal(ij, ij_ref) and w.tobytes() == w_ref.tobytes(), name
        assert st.filter_passes == passes, (name, st.filter_passes)
        if name == "nan":
            assert st.predicted_candidates == 0
        else:
            assert 0 < st.predicted_candidates and st.refills == 0


def _features_ref(X, C, cell, delta, cptr, cids, eps, T):
    """The four per-cell features restated in numpy: FP32 sums over the
    dimensions in order, the triangle-inequality cell adjacency in double,
    the sampled within-cell counts against the FP32 squared threshold."""
    m, d = C.shape
    cnt = np.diff(cptr).astype(np.int64)
    R = np.zeros(m, np.float32)
    np.maximum.at(R, cell, delta)
    acc = np.zeros((m, m), np.float32)
    for k in range(d):
        t = C[:, None, k] - C[None, :, k]
        acc = acc + t * t
    D = np.sqrt(acc).astype(np.float64)
    R64 = R.astype(np.float64)
    bound = (R64[:, None] + R64[None, :] + np.float64(np.float32(eps))) * (1.0 + 1e-3) + 1e-15
    adj = (D <= bound) & (cnt[:, None] > 0) & (cnt[None, :] > 0)
    np.fill_diagonal(adj, False)
    col = np.triu(adj, 1).astype(np.int64) @ cnt
    nb = np.zeros(m, np.int64)
    for a in range(m):
        mem = cids[cptr[a]:cptr[a + 1]]
        smp = mem[:min(32, mem.size)]
        s2 = np.zeros((mem.size, smp.size), np.float32)
        for k in range(d):
            t = X[smp, k][None, :] - X[mem, k][:, None]
            s2 = s2 + t * t
        nb[a] = int(np.count_nonzero(s2 <= T))
    n = cnt.astype(np.float64)
    S = np.minimum(32, cnt)
    kk = np.where(S > 0, nb / np.maximum(S, 1), 0.0)
    e = max(float(np.float32(eps)), 1e-30)
    rr = np.where(cnt > 0, R64, 0.0)
    f = np.stack([np.log2(1 + n) / 16, np.log2(1 + 0.5 * n + col) / 24,
                  np.log2(1 + np.minimum(rr / e, 1e30)) / 8, np.log2(1 + n * kk) / 24], 1)
    return f.astype(np.float32)


@pytest.mark.parametrize("d", [20, 19, 144])
def test_cell_features_match_the_restatement(orc, d):
    """The cost model's inputs for every cell (members, paired points of
    adjacent later cells, radius, sampled neighbour count) against the numpy
    restatement, on both load paths of the count kernel (d % 4 == 0 or not)
    and at the cell graph's largest d."""
    from rnng_amd import l2, synth, voronoi
    cm = _cm()
    rng = np.random.default_rng(d)
    k = 12
    mu = rng.uniform(0, 10, size=(k, d)).astype(np.float32)
    X = (mu[rng.integers(0, k, 4000)] + 0.7 * rng.standard_normal((4000, d))).astype(np.float32)
    eps = synth.radius_for_degree(X, 12.0)
    ds = l2.Dataset(X)
    v = voronoi.Voronoi(ds, fps=48)
    try:
        g = v.graph(eps)
        f, _ = cm.cell_stats(v)
        g.close()
        cell, delta = v.assignment()
        cptr, cids = v.cells()
        C = X[v.centers()]
        ref = _features_ref(X, C, cell, delta, cptr, cids, eps, orc.sq_threshold(eps))
    finally:
        v.close()
        ds.close()
    flush_print(f"d={d}: features max |diff| {np.abs(f - ref).max():.3g}")
    assert np.allclose(f, ref, rtol=0, atol=1e-6), np.argwhere(~np.isclose(f, ref, rtol=0, atol=1e-6))[:5]

"""Learned cost model of the Voronoi cell query.

A 4 -> 64 -> 64 -> 2 ReLU perceptron (FP32) maps per-cell features (members,
paired column points, radius / eps, a sampled within-cell neighbour count;
log-scaled) to the cell's log-scaled candidate and edge counts.  The library's own kernels run
its forward, backward (squared error, 4x weight below the target), Adam and
inference passes
(rnng_mlp_*).  Once set (set_cost_model), rnng_voronoi_graph_build sizes its
candidate buffer from the summed per-cell estimates and runs the filter once;
without it the build counts first and fills in a second pass.

`python -m rnng_amd.costmodel OUT.npz` collects per-cell counts from 256
synthetic cell builds of 3k-300k points and 32 of 0.8M-2M points (random or
farthest-point landmarks at every size) on the GPU and
trains the network on them (12,000 full-batch Adam steps, lr 3e-3): the shipped
cost_model.npz.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import time

import numpy as np

from . import _lib
from ._lib import check

NPARAMS = 4610
DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cost_model.npz")


def _vp(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _f32(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 2 or a.shape[1] != cols:
        raise ValueError(f"expected [batch, {cols}]")
    return a


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised weights, zero biases, packed as the C ABI expects."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_out, fan_in in [(64, 4), (64, 64), (2, 64)]:
        parts.append((rng.standard_normal((fan_out, fan_in)) * np.sqrt(2.0 / fan_in)).ravel())
        parts.append(np.zeros(fan_out))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == NPARAMS
    return p


def infer(params, x) -> np.ndarray:
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = _f32(x, 4)
    y = np.empty((x.shape[0], 2), np.float32)
    check(_lib.lib().rnng_mlp_infer(_vp(p), _vp(x), x.shape[0], _vp(y)))
    return y


def loss_grad(params, x, t):
    p = np.ascontiguousarray(params, dtype=np.float32)
    x, t = _f32(x, 4), _f32(t, 2)
    loss = C.c_float()
    g = np.empty(NPARAMS, np.float32)
    check(_lib.lib().rnng_mlp_loss_grad(_vp(p), _vp(x), _vp(t), x.shape[0], C.byref(loss), _vp(g)))
    return float(loss.value), g


def set_under_weight(w: float) -> None:
    """Loss weight of an estimate below its target (default 4), for
    loss_grad and train."""
    check(_lib.lib().rnng_mlp_set_under_weight(C.c_float(w)))


def under_weight() -> float:
    return float(_lib.lib().rnng_mlp_under_weight())


def train(params, x, t, iters: int = 3000, lr: float = 3e-3):
    """Full-batch Adam on the GPU.  Returns (trained params, final loss)."""
    p = np.array(params, dtype=np.float32, copy=True)
    x, t = _f32(x, 4), _f32(t, 2)
    loss = C.c_float()
    check(_lib.lib().rnng_mlp_train(_vp(p), _vp(x), _vp(t), x.shape[0], int(iters), np.float32(lr),
                                    C.byref(loss)))
    return p, float(loss.value)


def set_cost_model(params) -> None:
    """Use `params` for cell graph builds (None: count-then-fill)."""
    if params is None:
        check(_lib.lib().rnng_set_cost_model(None))
        return
    p = np.ascontiguousarray(params, dtype=np.float32)
    if p.size != NPARAMS:
        raise ValueError("cost model needs 4610 parameters")
    check(_lib.lib().rnng_set_cost_model(_vp(p)))


def load(path: str = DEFAULT_PATH) -> np.ndarray | None:
    if not os.path.exists(path):
        return None
    with np.load(path, allow_pickle=False) as z:
        return np.asarray(z["params"], dtype=np.float32)


def use_default() -> bool:
    """Set the shipped model if present; returns whether one was set."""
    p = load()
    if p is None:
        return False
    set_cost_model(p)
    return True


def cell_stats(vor) -> tuple[np.ndarray, np.ndarray]:
    """(features [m,4], targets [m,2]) of the last graph build of `vor`."""
    f = np.empty((vor.m, 4), np.float32)
    t = np.empty((vor.m, 2), np.float32)
    check(_lib.lib().rnng_voronoi_cell_stats(vor._h, _vp(f), _vp(t)))
    return f, t


def decode(y: np.ndarray) -> np.ndarray:
    """Counts from log-scaled outputs."""
    return np.exp2(24.0 * np.clip(y, 0.0, 1.5)) - 1.0


def synthetic_run(rng: np.random.Generator, big: float = 0.0, huge: float = 0.0):
    """One varied cell workload: (points, eps, m, fps?).  With probability
    `big` the run has 100k-300k points and up to 2,048 cells, with probability
    `huge` 0.8M-2M points in 2,048 or 4,096 cells."""
    from . import synth
    u = rng.random()
    very = u < huge
    large = very or u < huge + big
    if very:
        n = int(rng.integers(800_000, 2_000_001))
    elif large:
        n = int(rng.integers(100_000, 300_001))
    else:
        n = int(rng.integers(3000, 30001))
    d = int(rng.choice([8, 16, 32, 64, 96, 128]))
    kind = int(rng.integers(0, 4))
    if kind == 3:       # the vMF-like unit-vector mixture of the 2M x 96 cell workload, smaller
        pts, _ = synth.points_vmf(n=n, d=96, k=int(rng.integers(64, 1025)), kappa=float(rng.uniform(20, 100)),
                                  seed=int(rng.integers(1 << 30)), degree=5.0)
        d = 96
    if kind == 0:       # Gaussian mixture
        k = int(rng.integers(4, 64))
        mu = rng.uniform(0, 10, size=(k, d)).astype(np.float32)
        pts = mu[rng.integers(0, k, n)] + rng.uniform(0.2, 1.5) * rng.standard_normal((n, d)).astype(np.float32)
    elif kind == 1:     # unit vectors around a few directions
        k = int(rng.integers(4, 256))
        mu = rng.standard_normal((k, d))
        mu /= np.linalg.norm(mu, axis=1, keepdims=True)
        pts = mu[rng.integers(0, k, n)] + rng.uniform(0.05, 0.5) * rng.standard_normal((n, d))
        pts /= np.linalg.norm(pts, axis=1, keepdims=True)
    elif kind == 2:     # integer-valued mixture in [0, 255]
        k = int(rng.integers(4, 64))
        mu = rng.uniform(30, 225, size=(k, d))
        pts = np.clip(np.rint(mu[rng.integers(0, k, n)] + rng.uniform(5, 30) * rng.standard_normal((n, d))), 0, 255)
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    degree = float(rng.uniform(2.0, 60.0))
    eps = synth.radius_for_degree(pts, degree, sample=min(n, 1500), seed=int(rng.integers(1 << 30)))
    sizes = [2048, 4096] if very else ([256, 512, 1024, 2048] if large else [8, 16, 32, 64, 128, 256, 512])
    m = int(min(n // 8, rng.choice(sizes)))
    return pts, eps, max(m, 2), bool(rng.integers(0, 2))


def collect(runs: int = 256, seed: int = 1, log=None, big: float = 0.0, huge: float = 0.0):
    """Per-cell (features, targets) from `runs` synthetic cell builds, each
    run without a model (count-then-fill)."""
    from . import l2, voronoi
    set_cost_model(None)
    rng = np.random.default_rng(seed)
    X, T = [], []
    for r in range(runs):
        pts, eps, m, use_fps = synthetic_run(rng, big, huge)
        ds = l2.Dataset(pts)
        if use_fps:
            v = voronoi.Voronoi(ds, fps=m, first=int(rng.integers(0, pts.shape[0])))
        else:
            v = voronoi.Voronoi(ds, centers=voronoi.selection_sample(pts.shape[0], m, int(rng.integers(1 << 30))))
        g = v.graph(eps)
        f, t = cell_stats(v)
        X.append(f)
        T.append(t)
        if log:
            log(f"run {r}: n={pts.shape[0]} d={pts.shape[1]} m={m} eps={eps:.4g} edges={g.num_edges} "
                f"candidates={g.stats.candidates}")
        g.close()
        v.close()
        ds.close()
    return np.concatenate(X), np.concatenate(T)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="train the cell-query cost model on the GPU")
    ap.add_argument("out")
    ap.add_argument("--runs", type=int, default=256, help="runs of 3k-300k points")
    ap.add_argument("--huge-runs", type=int, default=32, help="additional runs of 0.8M-2M points")
    ap.add_argument("--iters", type=int, default=12000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--big", type=float, default=0.25, help="share of 100k-300k point runs")
    ap.add_argument("--under", type=float, default=4.0, help="loss weight of an estimate below its count")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    t0 = time.perf_counter()
    X, T = collect(a.runs, a.seed, log, big=a.big)
    if a.huge_runs:
        Xh, Th = collect(a.huge_runs, a.seed + 4, log, huge=1.0)
        X, T = np.concatenate([X, Xh]), np.concatenate([T, Th])
    log(f"collected {X.shape[0]} cells from {a.runs + a.huge_runs} runs in {time.perf_counter() - t0:.1f}s")
    set_under_weight(a.under)
    p, loss = train(init_params(a.seed), X, T, iters=a.iters, lr=a.lr)
    log(f"trained: loss {loss:.3e}")
    np.savez(a.out, params=p, features=X, targets=T)


if __name__ == "__main__":
    main()

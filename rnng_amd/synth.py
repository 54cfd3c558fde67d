"""Seeded synthetic inputs for the five workloads of this path.

Everything is generated with numpy's PCG64 (np.random.default_rng(seed)), so
the bytes are identical on every machine with the same numpy.  The radius of
the L2 workloads is set from a 2,000-point sample: the k-th smallest pairwise
distance, k chosen so that the expected mean degree over the full set is the
target.

  points_gmm_small   10,000 x 16, 20 Gaussians, centers U[0,10]^16, sigma 0.5,
                     mean degree ~20 (the reference's CPU-runnable size)
  points_gmm_int     1,000,000 x 128, 256 Gaussians (sigma 20), clipped to
                     [0,255] and rounded, mean degree ~30
  dna_mutants        200,000 strings over ACGT: 2,000 seeds of length
                     U[100,150], 100 mutants each (sub/ins/del 1%/0.5%/0.5%)
  points_vmf         2,000,000 x 96 unit vectors, 1,024 directions, kappa 50,
                     mean degree ~25
  points_rips        300,000 x 3: 50% torus (R=3, r=1), 30% unit sphere at
                     (6,0,0), 20% uniform in the bounding box, noise sigma 0.02
"""
from __future__ import annotations

import numpy as np

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def radius_for_degree(points: np.ndarray, degree: float, sample: int = 2000,
                      seed: int = 12345) -> np.float32:
    """Radius whose expected mean degree over `points` is `degree`, from the
    pairwise distances (float64) of a `sample`-point random subset."""
    n = points.shape[0]
    rng = np.random.default_rng(seed)
    s = min(sample, n)
    idx = np.sort(rng.choice(n, size=s, replace=False))
    P = points[idx].astype(np.float64)
    sq = (P * P).sum(1)
    D2 = sq[:, None] + sq[None, :] - 2.0 * (P @ P.T)
    iu = np.triu_indices(s, 1)
    d = np.sqrt(np.maximum(D2[iu], 0.0))
    k = int(round(degree / max(n - 1, 1) * d.size))
    k = min(max(k, 1), d.size)
    return np.float32(np.partition(d, k - 1)[k - 1])


def points_gmm_small(n: int = 10_000, d: int = 16, k: int = 20, sigma: float = 0.5,
                     seed: int = 1):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0.0, 10.0, size=(k, d))
    lab = rng.integers(0, k, size=n)
    pts = (centers[lab] + sigma * rng.standard_normal((n, d))).astype(np.float32)
    return pts, radius_for_degree(pts, 20.0)


def points_gmm_int(n: int = 1_000_000, d: int = 128, k: int = 256, sigma: float = 20.0,
                   seed: int = 2, degree: float = 30.0):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0.0, 255.0, size=(k, d)).astype(np.float32)
    lab = rng.integers(0, k, size=n)
    pts = np.empty((n, d), np.float32)
    step = 1 << 16
    for s in range(0, n, step):
        e = min(n, s + step)
        blk = centers[lab[s:e]] + np.float32(sigma) * rng.standard_normal((e - s, d), dtype=np.float32)
        np.clip(np.rint(blk, out=blk), 0.0, 255.0, out=blk)
        pts[s:e] = blk
    return pts, radius_for_degree(pts, degree)


def dna_mutants(n_seeds: int = 2000, per_seed: int = 100, lmin: int = 100, lmax: int = 150,
                p_sub: float = 0.01, p_ins: float = 0.005, p_del: float = 0.005, seed: int = 3):
    """Returns (atoms uint8 [total], offsets int64 [n+1]) in shuffled order."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(lmin, lmax + 1, size=n_seeds)
    seed_off = np.zeros(n_seeds + 1, np.int64)
    np.cumsum(lens, out=seed_off[1:])
    seeds = rng.integers(0, 4, size=int(seed_off[-1])).astype(np.uint8)
    # every mutant copies its seed's bases
    m_seed = np.repeat(np.arange(n_seeds), per_seed)
    m_len = lens[m_seed]
    starts = np.repeat(seed_off[m_seed], m_len)
    within = np.arange(int(m_len.sum())) - np.repeat(np.cumsum(m_len) - m_len, m_len)
    base = seeds[starts + within]
    u = rng.random(base.size)
    sub = u < p_sub
    ins = (u >= p_sub) & (u < p_sub + p_ins)
    dele = (u >= p_sub + p_ins) & (u < p_sub + p_ins + p_del)
    base = base.copy()
    base[sub] = (base[sub] + rng.integers(1, 4, size=int(sub.sum()))) % 4
    # deleted bases are dropped; an insertion adds one random base after this one
    reps = np.where(dele, 0, np.where(ins, 2, 1))
    out = np.repeat(base, reps)
    owner = np.repeat(np.repeat(np.arange(m_seed.size), m_len), reps)
    first = np.concatenate(([True], np.diff(np.repeat(np.arange(base.size), reps)) != 0))
    inserted = ~first
    out[inserted] = rng.integers(0, 4, size=int(inserted.sum())).astype(np.uint8)
    counts = np.bincount(owner, minlength=m_seed.size)
    # shuffle the strings
    perm = rng.permutation(m_seed.size)
    offs = np.zeros(m_seed.size + 1, np.int64)
    np.cumsum(counts, out=offs[1:])
    new_counts = counts[perm]
    new_offs = np.zeros(m_seed.size + 1, np.int64)
    np.cumsum(new_counts, out=new_offs[1:])
    idx = np.repeat(offs[perm], new_counts) + (np.arange(int(new_offs[-1])) -
                                                np.repeat(new_offs[:-1], new_counts))
    atoms = ACGT[out[idx]]
    return atoms, new_offs


def points_vmf(n: int = 2_000_000, d: int = 96, k: int = 1024, kappa: float = 50.0,
               seed: int = 4, degree: float = 25.0):
    """Unit vectors x = normalize(mu + z / sqrt(kappa)), z ~ N(0, I)."""
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal((k, d))
    mu /= np.linalg.norm(mu, axis=1, keepdims=True)
    mu = mu.astype(np.float32)
    lab = rng.integers(0, k, size=n)
    pts = np.empty((n, d), np.float32)
    step = 1 << 16
    sc = np.float32(1.0 / np.sqrt(kappa))
    for s in range(0, n, step):
        e = min(n, s + step)
        blk = mu[lab[s:e]] + sc * rng.standard_normal((e - s, d), dtype=np.float32)
        blk /= np.linalg.norm(blk, axis=1, keepdims=True)
        pts[s:e] = blk
    return pts, radius_for_degree(pts, degree)


def points_rips(n: int = 300_000, noise: float = 0.02, seed: int = 5):
    rng = np.random.default_rng(seed)
    n_t, n_s = n // 2, (3 * n) // 10
    n_u = n - n_t - n_s
    th = rng.uniform(0, 2 * np.pi, n_t)
    ph = rng.uniform(0, 2 * np.pi, n_t)
    R, r = 3.0, 1.0
    torus = np.stack([(R + r * np.cos(th)) * np.cos(ph), (R + r * np.cos(th)) * np.sin(ph),
                      r * np.sin(th)], 1)
    v = rng.standard_normal((n_s, 3))
    sphere = v / np.linalg.norm(v, axis=1, keepdims=True) + np.array([6.0, 0.0, 0.0])
    both = np.concatenate([torus, sphere])
    lo, hi = both.min(0), both.max(0)
    unif = rng.uniform(lo, hi, size=(n_u, 3))
    pts = np.concatenate([torus, sphere, unif]) + noise * rng.standard_normal((n, 3))
    pts = pts[rng.permutation(n)].astype(np.float32)
    return pts, np.float32(0.08)

"""Benchmark of the epsilon-graph hot path on one MI355X.

Workload: 1,000,000 points x 128 dims, float32, integer-valued in [0,255]
(256-component Gaussian mixture, sigma 20, clipped and rounded, seeded),
eps set for mean degree ~30.  One step = one full epsilon-graph build over
the resident points: fragment preparation, the MFMA candidate filter over all
499,999,500,000 unordered pairs, the FP32 reference-order recheck and FP64
band check.  The plain FP32 kernel (reference order, every pair) is
timed once beside it.

Prints one JSON line (rank 0).  The rates are pair-evaluations per second
(value) and edges per second; roofline is for the MFMA filter kernel, whose
per-launch duration is measured with HIP events on the library's stream.  On
this integer-valued data (every dimension spans at most 255) the filter runs
exactly on int8 MFMA and its roofline is the dense int8 peak; the BF16 filter
is timed beside it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_BF16_TFLOPS = 2500.0
PEAK_INT8_TOPS = 5000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(pts, eps, sample):
    """The CPU oracle (own C restatement of the reference's L2 functor and
    brute-force range query) on a bounded sample: self join of the first
    `sample` points.  Returns the JSON object."""
    import oracle
    cores = min(len(os.sched_getaffinity(0)), 16)
    sub = pts[:sample]
    t0 = time.perf_counter()
    edges = oracle.l2_eps_pairs_count(sub, eps, nthreads=cores)
    dt = time.perf_counter() - t0
    pairs = sample * (sample - 1) // 2
    return {"value": pairs / dt, "unit": "pair-evals/s", "cores": cores, "kind": "port",
            "sample": f"self join of the first {sample} points ({pairs} pairs, {edges} edges) "
                      f"in {dt:.2f} s; edges/s {edges / dt:.4g}"}


def reference_baseline(pts, eps, sample):
    """The reference's own systolic driver (built from its sources into
    oracle/_ref by oracle/Makefile) on the first `sample` points, one MPI
    rank per core: its cover-tree build plus the ring of radius queries, as
    its own timers report them (drivers/systolic.cpp:149-156).  None when
    the binary or MPICH's launcher is absent."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "systolic")
    mpiexec = os.path.join(os.environ.get("RNNG_MPI_HOME", "/opt/conda"), "bin", "mpiexec.hydra")
    if not (os.path.exists(exe) and os.path.exists(mpiexec)):
        return None
    ranks = min(len(os.sched_getaffinity(0)), 16)
    sub = np.ascontiguousarray(pts[:sample], np.float32)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "sample.fvecs")
        rec = np.empty((sub.shape[0], sub.shape[1] + 1), np.float32)
        rec[:, 0] = np.array([sub.shape[1]], np.int32).view(np.float32)[0]
        rec[:, 1:] = sub
        rec.tofile(path)
        t0 = time.perf_counter()
        r = subprocess.run([mpiexec, "-n", str(ranks), exe, "-i", path, "-r", "%.9g" % float(eps),
                            "-o", os.path.join(tmp, "out.txt")], capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
    if r.returncode != 0:
        log(f"reference driver failed ({r.returncode}): {r.stderr[-500:]}")
        return None
    timers = {}
    for line in (r.stdout + r.stderr).splitlines():
        if line.startswith("[time="):
            t, _, what = line[6:].partition("] ")
            timers[what.split(" [")[0].split(" '")[0]] = float(t)
    search = timers.get("built cover tree", 0.0) + timers.get("found neighbors", 0.0)
    if search <= 0:
        return None
    pairs = sample * (sample - 1) // 2
    return {"value": pairs / search, "unit": "pair-evals/s", "cores": ranks, "kind": "reference",
            "sample": f"the reference's systolic driver (cover trees, {ranks} MPI ranks) on the first {sample} "
                      f"points: tree build {timers.get('built cover tree', 0):.3f} s + ring of radius queries "
                      f"{timers.get('found neighbors', 0):.3f} s ({pairs} pairs, all-pairs equivalent); "
                      f"whole run with file I/O {wall:.2f} s"}


def edge_digest(g):
    """Edge count and SHA-256 of the (i, j)-sorted pairs (int64) and of their
    weights (float32): the form of the oracle's frozen full-size digests.
    The sort runs on the device at copy time, outside any timed step."""
    import hashlib
    ij, w = g.edges(sorted=True)
    return {"edges": int(ij.shape[0]), "sha256_ij_int64": hashlib.sha256(ij.tobytes()).hexdigest(),
            "sha256_w_float32": hashlib.sha256(w.tobytes()).hexdigest()}


def same_edges(a, b):
    return all(a[k] == b[k] for k in ("edges", "sha256_ij_int64", "sha256_w_float32"))


def oracle_digest(name, pts, eps):
    """The committed oracle digest of this exact input (None if absent or for
    other points or another eps)."""
    import hashlib
    path = os.path.join(ROOT, "tests", "golden", name)
    if not os.path.exists(path):
        return None
    ref = json.load(open(path))
    if (ref["n"], ref["d"]) != pts.shape or ref["eps_bits"] != int(np.float32(eps).view(np.uint32)):
        return None
    if hashlib.sha256(np.ascontiguousarray(pts, np.float32).tobytes()).hexdigest() != ref["sha256_points_float32"]:
        return None
    return ref


def measured_traffic(profile, n):
    """HBM bytes per filter launch from the committed rocprofv3 summary (two
    --pmc passes, FETCH_SIZE and WRITE_SIZE, scaled by a copy kernel of known
    size); None when the summary is absent or for another size."""
    path = os.path.join(ROOT, "profiles", profile)
    if n != 1_000_000 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def bench_cells(args):
    """Voronoi-cell workload: 2,000,000 x 96 unit vectors (vMF-like mixture of
    1,024 directions, kappa 50), 4,096 farthest-point landmarks, eps for mean
    degree ~25 (--cells-data int: the 1,000,000 x 128 integer-valued GMM
    instead, on the exact int8 filter).  One step = ghost test + batched
    per-cell query + dedup over the resident cells; the landmark selection and
    the cell grouping are timed once.  The all-pairs MFMA self join runs on the
    same points beside it, and both edge sets are compared by digest."""
    from rnng_amd import Stats, costmodel, l2, synth, voronoi
    have_model = costmodel.use_default()
    log(f"cost model: {'shipped model' if have_model else 'none (count-then-fill)'}")
    t0 = time.perf_counter()
    if args.cells_data == "int":
        pts, eps = synth.points_gmm_int(n=args.n)
        data = "synthetic (seeded 256-component Gaussian mixture, integer-valued)"
        what = "fp32 integer-valued [0,255] (mean degree ~30)"
    else:
        n = args.n if args.n != 1_000_000 else 2_000_000
        pts, eps = synth.points_vmf(n=n)
        data = "synthetic (seeded vMF-like mixture)"
        what = "unit-norm fp32 (mean degree ~25)"
    n, d = pts.shape
    log(f"generated {n}x{d} eps={eps} in {time.perf_counter() - t0:.1f}s")
    ds = l2.Dataset(pts)
    t0 = time.perf_counter()
    v = voronoi.Voronoi(ds, fps=args.cells, first=0)
    t_lm = time.perf_counter() - t0
    ptr, _ = v.cells()
    sizes = np.diff(ptr)
    log(f"fps landmarks + cells: {t_lm * 1e3:.1f} ms; cell sizes min {sizes.min()} "
        f"median {int(np.median(sizes))} max {sizes.max()}")
    if args.schedule:
        order, form = (int(x) for x in args.schedule.split(","))
        v.set_schedule(order, form)
    for w in range(args.warmup):
        g = v.graph(eps)
        log(f"warmup {w}: {g.num_edges} edges, stats {g.stats.as_dict()}")
        g.close()
    st = Stats()
    t0 = time.perf_counter()
    ne = 0
    g = None
    for _ in range(args.steps):
        if g is not None:
            g.close()
        g = v.graph(eps, "mfma", st)
        ne = g.num_edges
    dt = time.perf_counter() - t0
    K = args.steps
    log(f"cells: {K} steps {dt / K * 1e3:.2f} ms/step, {ne} edges, stats {st.as_dict()}")
    dig = edge_digest(g)
    g.close()
    tv, tx, tms = v.tuning()
    order, form, cfgs = v.schedule()
    int8 = st.int8_filter == K
    P_all = n * (n - 1) // 2
    out = {
        "metric": f"ε-graph edges/s via Voronoi cells + ghosts, {n // 1000000}M×{d}",
        "value": ne / (dt / K), "unit": "edges/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": dt / K * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("int8 mfma filter (int32 accumulate, exact on this integer data) + fp32 reference-order recheck"
                  if int8 else "bf16 mfma filter (fp32 accumulate) + fp32 reference-order recheck"),
        "data": data,
        "config": {"workload": f"eps-graph through {args.cells} Voronoi cells, {n} points x {d} dims {what}, "
                               f"eps={float(eps):.6g}",
                   "n": n, "d": d, "eps": float(eps), "cells": args.cells},
        "edges": ne,
        "edge_set": dict(dig, note="sorted on the device at copy time, outside the timed step"),
        "all_pairs_equiv_per_s": P_all / (dt / K),
        "cell_pair_evals": st.pair_evals // K,
        "ghost_tests": st.ghost_tests // K, "ghost_candidates": st.ghost_candidates // K,
        "ghosts": st.ghosts // K, "candidates": st.candidates // K,
        "band_pairs": st.band_pairs // K, "band_flips": st.band_flips // K, "refills": st.refills,
        "cost_model": {"used": have_model, "filter_passes_per_step": st.filter_passes / K,
                       "predicted_candidates": st.predicted_candidates // K,
                       "predicted_over_actual": (st.predicted_candidates / max(1, st.candidates)),
                       "predicted_edges": st.predicted_edges // K},
        "landmarks_and_cells_ms": t_lm * 1e3,
        "schedule": {"row_order": ["cells in landmark order", "cells along the center chain",
                                   "flat (points in their own order, every pair)"][order],
                     "work_form": ["runs of column tiles", "tile lists", "tile lists on a persistent queue"][form],
                     "forced": bool(args.schedule), "filter_variant": None if int8 else tv,
                     "tuned_best_shape": (min(cfgs, key=lambda c: c[3])[4] if cfgs else None),
                     "xcd_order": tx,
                     "tuned_ms_on_an_eighth": {f"order{o}/form{f}/xcd{x}/shape{k}": round(t, 3)
                                               for o, f, x, t, k in cfgs}},
        "cell_sizes": {"min": int(sizes.min()), "median": int(np.median(sizes)), "max": int(sizes.max())},
        "stage_ms": {"ghost_test": st.ghost_ms / K, "prep": st.prep_ms / K, "filter": st.filter_ms / K,
                     "recheck_dedup": st.recheck_ms / K, "device_total": st.total_ms / K},
    }
    # the filter's work: 2 d ops per pair of the padded 32 x 32 tiles it runs
    # (cell_pair_evals; on the vMF data every cell pair is adjacent, so that
    # is every unordered pair plus padding)
    peak = PEAK_INT8_TOPS if int8 else PEAK_BF16_TFLOPS
    ach = 2.0 * d * (st.pair_evals / K) / (st.filter_ms / K * 1e-3) / 1e12
    kname = ("k_l2_mfma_filter_ring<int8>" if int8 else
             {12: "k_l2_mfma_filter_ring<bf16, G=2, S=3>", 13: "k_l2_mfma_filter_ring<bf16, G=2, S=2>",
              14: "k_l2_mfma_filter_ring<bf16, G=3, S=2>",
              17: "k_l2_mfma_filter_ring<bf16, G=8, S=2, 12 waves>",
              18: "k_l2_mfma_filter_ring<bf16 16x16, G=8, S=2, 12 waves>",
              19: "k_l2_mfma_filter_ring<bf16 16x16, G=2, S=2>"}.get(tv, "k_l2_mfma_filter_ring"))
    if form > 0:
        kname += ", tile lists" + (", persistent" if form == 2 else "")
    out["roofline"] = {"kernel": kname, "bound": "mfma", "achieved": ach, "peak": peak,
                       "unit": "TOP/s (int8)" if int8 else "TFLOP/s", "frac": ach / peak, "traffic": None,
                       "work": "2 d ops per pair of the tiles the filter runs, fill-pass filter time from HIP events"}
    v.close()
    if not args.no_plain:
        # the all-pairs MFMA self join on the same resident points
        sa = Stats()
        for rep in range(2):
            t1 = time.perf_counter()
            g = ds.graph(eps, "mfma", sa)
            ta = time.perf_counter() - t1
            if rep == 1:
                adig = edge_digest(g)
            g.close()
        out["all_pairs"] = {"ms": ta * 1e3, "filter_ms": sa.filter_ms / 2, "edges": adig["edges"],
                            "same_edge_set": same_edges(adig, dig),
                            "cells_speedup": ta / (dt / K)}
        log(f"all pairs: {ta * 1e3:.1f} ms, {adig['edges']} edges, same set {same_edges(adig, dig)}")
    ds.close()
    print(json.dumps(out), flush=True)


def _secondary_line(metric, value, unit, K, warmup, ms, workload, extra):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": K, "warmup": warmup,
           "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "data": "synthetic (seeded)", "config": {"workload": workload}}
    out.update(extra)
    return out


# int32 VALU lane-operations per second: 256 CUs x 128 lanes x 2.4 GHz, the
# rate behind the 157.3 TFLOP/s FP32 figure (an FMA counted as 2 flops); the
# verify kernel's counters show 66 lane-ops per CU-clock, so the VALU issues
# more than 64 lanes per clock
PEAK_VALU_INT32_TOPS = 78.64
# VALU instructions per text column in the banded bit-parallel verify
# (k <= 15: 32-bit band words, two code bit-planes), counted in its gfx950 code
VERIFY_OPS_PER_COLUMN = 39


def _verify_roofline(st, K):
    """The banded bit-parallel verify (k_edit_verify_bp): its text columns
    (counted by the kernel over the candidates inside the length window) at
    VERIFY_OPS_PER_COLUMN lane-operations each, over its HIP-event time."""
    if st.kernel_units <= 0 or st.kernel_ms <= 0:
        return None
    ops = st.kernel_units / K * VERIFY_OPS_PER_COLUMN
    ach = ops / (st.kernel_ms / K * 1e-3) / 1e12
    return {"kernel": "k_edit_verify_bp<uint32, 2 planes>", "bound": "valu (int32)", "achieved": ach,
            "peak": PEAK_VALU_INT32_TOPS, "unit": "Tops/s (int32 lane-ops)", "frac": ach / PEAK_VALU_INT32_TOPS,
            "traffic": None, "columns": st.kernel_units // K, "ms": st.kernel_ms / K,
            "work": f"{VERIFY_OPS_PER_COLUMN} VALU lane-ops per text column of each candidate inside the "
                    "length window"}


def bench_edit(args):
    """Edit-distance workload: 200,000 DNA-like strings (2,000 random seeds of
    length 100-150, 100 mutants each at 1%/0.5%/0.5% substitution/insertion/
    deletion), integer eps = 8.  One step = the full graph build over the
    resident strings."""
    from rnng_amd import Stats, edit, synth
    t0 = time.perf_counter()
    atoms, offs = synth.dna_mutants()
    n = offs.shape[0] - 1
    log(f"generated {n} strings in {time.perf_counter() - t0:.1f}s")
    S = edit.StringSet(atoms, offs)
    eps = 8.0
    for w in range(args.warmup):
        g = S.graph(eps)
        log(f"warmup {w}: {g.num_edges} edges, stats {g.stats.as_dict()}")
        g.close()
    st = Stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g = S.graph(eps, st)
        ne = g.num_edges
        g.close()
    dt = (time.perf_counter() - t0) / args.steps
    K = args.steps
    # algorithmic work of the filter: 256 int8 multiply-adds (64 bins x 4
    # thermometer levels) per pair inside the length window |m - n| <= k
    lens = np.sort(np.diff(offs))
    win = int((np.searchsorted(lens, lens + int(eps), side="right") - np.arange(1, n + 1)).sum())
    filt_ms = st.filter_ms / K
    achieved = 2.0 * 256 * win / (filt_ms * 1e-3) / 1e12
    out = _secondary_line(
        "edit-distance ε-graph pair-evals/s and edges/s, 200k strings", st.pair_evals / K / dt, "pair-evals/s",
        K, args.warmup, dt * 1e3, f"edit-distance eps-graph, {n} DNA-like strings (length 100-150), eps=8",
        {"dtype": "int32", "edges": ne, "edges_per_s": ne / dt, "pair_evals": st.pair_evals // K,
         "length_window_pairs": win, "candidates": st.candidates // K,
         "stage_ms": {"filter": filt_ms, "verify": st.recheck_ms / K, "device_total": st.total_ms / K},
         "roofline": {"kernel": "k_edit_filter", "bound": "mfma", "achieved": achieved, "peak": 5000.0,
                      "unit": "TOP/s (int8)", "frac": achieved / 5000.0, "traffic": None,
                      "work": "2 x 256 int8 ops per length-window pair, filter time from HIP events"},
         "verify_roofline": _verify_roofline(st, K)})
    if not args.no_cpu:
        import oracle
        cores = min(len(os.sched_getaffinity(0)), 16)
        m = 4000
        t1 = time.perf_counter()
        ij, _ = oracle.edit_eps_pairs(atoms[:offs[m]], offs[:m + 1], eps, nthreads=cores)
        tc = time.perf_counter() - t1
        pairs = m * (m - 1) // 2
        out["cpu_baseline"] = {"value": pairs / tc, "unit": "pair-evals/s", "cores": cores, "kind": "port",
                               "sample": f"first {m} strings ({pairs} pairs, {ij.shape[0]} edges) in {tc:.2f} s"}
    S.close()
    print(json.dumps(out), flush=True)


def bench_rips(args):
    """Rips workload: 300,000 points in 3 dims (torus, sphere, uniform box,
    noise 0.02), eps 0.08, 2-skeleton.  One step = epsilon graph + cliques +
    the (value, id)-sorted filtration, left in HBM; copying it to the host is
    timed separately."""
    from rnng_amd import Stats, l2, rips, synth
    pts, eps = synth.points_rips()
    ds = l2.Dataset(pts)
    for w in range(args.warmup):
        with rips.Filtration(ds, eps, 2) as f:
            log(f"warmup {w}: {f.size()[0]} simplices")
    st = Stats()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rips.Filtration(ds, eps, 2, st).close()
    dt = (time.perf_counter() - t0) / args.steps
    K = args.steps
    with rips.Filtration(ds, eps, 2) as f:
        total, per = f.size()
        t1 = time.perf_counter()
        f.copy()
        copy_s = time.perf_counter() - t1
    # the triangle fill (k_triangles<true>): per triangle 12 bytes written
    # (u32 value group + i64 id); the lower CSR read once (per edge 8 bytes of
    # (group, vertex) + 4 of row + 8 of output offset, 8 per row pointer)
    n, E, T = pts.shape[0], int(per[1]), int(per[2])
    fill_bytes = 12.0 * T + 20.0 * E + 8.0 * (n + 1)
    fill_ms = st.kernel_ms / K
    ach = fill_bytes / (fill_ms * 1e-3) / 1e9 if fill_ms > 0 else None
    out = _secondary_line(
        "Rips 2-skeleton simplices/s, 300k x 3", total / dt, "simplices/s", K, args.warmup, dt * 1e3,
        "Vietoris-Rips 2-skeleton, 300000 points x 3 dims (torus/sphere/box), eps=0.08",
        {"dtype": "fp32 distances, int64 simplex ids", "simplices": [int(x) for x in per[:3]],
         "graph_pair_evals": st.pair_evals // K, "edges": st.edges // K,
         "stage_ms": {"graph_grid": st.filter_ms / K, "cliques": st.recheck_ms / K,
                      "order": st.prep_ms / K, "device_total": st.total_ms / K},
         "copy_to_host_ms": copy_s * 1e3,
         "simplices_per_s_with_copy": total / (dt + copy_s),
         "roofline": None if ach is None else {
             "kernel": "k_triangles<true> (triangle fill)", "bound": "hbm", "achieved": ach, "peak": 8000.0,
             "unit": "GB/s", "frac": ach / 8000.0, "traffic": None, "ms": fill_ms,
             "work": "12 B per triangle written + 20 B per lower edge and 8 B per row pointer read, "
                     "fill time from HIP events"}})
    if not args.no_cpu:
        import oracle
        cores = min(len(os.sched_getaffinity(0)), 16)
        m = 60000
        t1 = time.perf_counter()
        vv, _ = oracle.rips_filtration(pts[:m], eps, 2, nthreads=cores)
        tc = time.perf_counter() - t1
        out["cpu_baseline"] = {"value": vv.shape[0] / tc, "unit": "simplices/s", "cores": cores, "kind": "port",
                               "sample": f"first {m} points ({vv.shape[0]} simplices) in {tc:.2f} s"}
    ds.close()
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=150_000,
                    help="points of the oracle's brute-force baseline (without the reference driver)")
    ap.add_argument("--ref-sample", type=int, default=250_000,
                    help="points of the reference driver's baseline")
    ap.add_argument("--port-baseline", action="store_true",
                    help="time the oracle's brute force instead of the reference's driver")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first run a plain device copy of MIB MiB (counter calibration)")
    ap.add_argument("--workload", choices=["l2", "cells", "edit", "rips"], default="l2",
                    help="l2: 1M x 128 all-pairs (default, the headline); cells: 2M x 96 through "
                         "Voronoi cells; edit: 200k strings; rips: 300k x 3 2-skeleton")
    ap.add_argument("--cells", type=int, default=4096, help="landmarks of the cells workload")
    ap.add_argument("--cells-data", choices=["vmf", "int"], default="vmf",
                    help="cells workload points: 2M x 96 unit vectors, or the 1M x 128 integer GMM")
    ap.add_argument("--schedule", default="", metavar="ORDER,FORM",
                    help="force the row order (0 cells in landmark order, 1 cells along the center chain, "
                         "2 flat) and work form (0 runs, 1 tile lists, 2 persistent queue; flat: 0 only) "
                         "instead of the autotuner's choice")
    ap.add_argument("--bf16", action="store_true",
                    help="run the timed steps on the BF16 filter even where the exact int8 one applies")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "this build has no multi-GPU path", file=sys.stderr)
        sys.exit(2)

    if args.workload != "l2":
        {"cells": bench_cells, "edit": bench_edit, "rips": bench_rips}[args.workload](args)
        return

    from rnng_amd import Stats, l2, synth
    from rnng_amd import _lib

    t0 = time.perf_counter()
    pts, eps = synth.points_gmm_int(n=args.n)
    n, d = pts.shape
    P = n * (n - 1) // 2
    log(f"generated {n}x{d} eps={eps} in {time.perf_counter() - t0:.1f}s")
    ds = l2.Dataset(pts)
    if args.copy_calibration:
        import ctypes
        ms = ctypes.c_double()
        nb = args.copy_calibration << 20
        _lib.check(_lib.lib().rnng_debug_copy(nb, ctypes.byref(ms)))
        log(f"copy calibration: {nb} bytes read + written in {ms.value:.3f} ms "
            f"({2 * nb / ms.value / 1e6:.0f} GB/s)")

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731

    if args.bf16:
        _lib.check(_lib.lib().rnng_set_int8_filter(0))
    for w in range(args.warmup):
        g = ds.graph(eps, "mfma")
        log(f"warmup {w}: {g.num_edges} edges, {g.stats.total_ms:.1f} ms device")
        g.close()

    st = Stats()
    sync()
    t0 = time.perf_counter()
    ne = 0
    g = None
    for _ in range(args.steps):
        if g is not None:
            g.close()
        g = ds.graph(eps, "mfma", st)
        ne = g.num_edges
    sync()
    dt = time.perf_counter() - t0
    K = args.steps
    ms = dt / K * 1e3
    log(f"mfma: {K} steps {ms:.2f} ms/step, {ne} edges, stats {st.as_dict()}")
    # the last step's edges against the oracle's frozen digest of this input
    dig = edge_digest(g)
    g.close()
    ref = oracle_digest("full_gmm_int_1Mx128.json", pts, eps)
    log(f"edge digest {dig}; oracle digest {'present' if ref else 'absent'}")

    filt_s = st.filter_ms / K / 1e3
    flops = 2.0 * d * P
    achieved = flops / filt_s / 1e12
    # integer data spanning <= 255 per dimension runs the exact int8 filter
    # (v_mfma_i32_32x32x32_i8): its roofline is the dense int8 peak
    int8 = st.int8_filter == K
    peak = PEAK_INT8_TOPS if int8 else PEAK_BF16_TFLOPS
    out = {
        "metric": "ε-graph edges/s + pair-evals/s at 1M×128 fp32, MFMA-filter vs plain, % peak",
        "value": P / (dt / K),
        "unit": "pair-evals/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("int8 mfma filter (int32 accumulate) + exact integer recheck: exact on this integer data"
                  if int8 else "bf16 mfma filter (fp32 accumulate) + fp32 reference-order recheck"),
        "data": "synthetic (seeded 256-component Gaussian mixture, integer-valued)",
        "config": {"workload": f"eps-graph self join, {n} points x {d} dims fp32 integer-valued "
                               f"[0,255], eps={float(eps):.6g} (mean degree ~30)",
                   "n": n, "d": d, "eps": float(eps), "mode": "mfma"},
        "edges": ne,
        "edges_per_s": ne / (dt / K),
        "edge_set": dict(dig, equals_oracle_digest=(same_edges(dig, ref) if ref else None),
                         note="the (i, j) sort of the reference's Graph constructor runs on the device "
                              "when the edges are copied out, outside the timed step"),
        "candidates": st.candidates // K,
        "band_pairs": st.band_pairs // K,
        "band_flips": st.band_flips // K,
        "stage_ms": {"prep": st.prep_ms / K, "filter": st.filter_ms / K,
                     "recheck": st.recheck_ms / K, "device_total": st.total_ms / K},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak,
                     "unit": "TOP/s (int8)" if int8 else "TFLOP/s", "frac": achieved / peak,
                     "traffic": measured_traffic("l2_mfma_filter_i8_gmm_int_1Mx128.md" if int8
                                                 else "l2_mfma_filter_gmm_int_1Mx128.md", n)},
    }

    if int8 and not args.no_plain:
        # the BF16 filter on the same points, for comparison (same edges)
        _lib.check(_lib.lib().rnng_set_int8_filter(0))
        try:
            sb = Stats()
            for rep in range(2):
                g = ds.graph(eps, "mfma", sb)
                be = g.num_edges
                if rep == 1:
                    bdig = edge_digest(g)
                g.close()
            bf = sb.filter_ms / 2
            out["bf16_filter"] = {"filter_ms": bf, "achieved_tflops": flops / (bf / 1e3) / 1e12,
                                  "frac_of_bf16_peak": flops / (bf / 1e3) / 1e12 / PEAK_BF16_TFLOPS,
                                  "same_edge_set": same_edges(bdig, dig)}
            log(f"bf16 filter: {bf:.2f} ms, {be} edges")
        finally:
            _lib.check(_lib.lib().rnng_set_int8_filter(1))

    if not args.no_plain:
        sp = Stats()
        sync()
        t1 = time.perf_counter()
        g = ds.graph(eps, "plain", sp)
        pe = g.num_edges
        sync()
        tp = time.perf_counter() - t1
        pdig = edge_digest(g)
        g.close()
        log(f"plain: {tp * 1e3:.1f} ms, {pe} edges, stats {sp.as_dict()}")
        out["plain"] = {"ms": tp * 1e3, "pair_evals_per_s": P / tp, "edges_per_s": pe / tp,
                        "same_edge_set": same_edges(pdig, dig),
                        # 3 non-fused VALU ops per dim per pair against the FP32
                        # VALU op rate (256 CU x 128 lanes/clk x 2.4 GHz = 7.86e13/s;
                        # the 157.3 TFLOP/s peak counts an FMA as 2 flops)
                        "valu_frac": (3.0 * d * P / (sp.filter_ms / 1e3)) / 7.86e13}
        out["speedup_vs_plain"] = tp / (dt / K)

    if not args.no_cpu:
        cb = reference_baseline(pts, eps, min(args.ref_sample, n)) if not args.port_baseline else None
        if cb is None:
            cb = cpu_baseline(pts, eps, min(args.cpu_sample, n))
        log(f"cpu baseline: {cb}")
        out["cpu_baseline"] = cb
    ds.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

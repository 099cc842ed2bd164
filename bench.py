"""Benchmark of the t-SNE gradient path on one MI355X.

Workload (default): N=1,000,000 points, D=50 float32 from a seeded mixture of 100
Gaussians (Dirichlet(1) weights, per-cluster std U[0.5, 2], centres N(0, 5^2)); the exact
k=90 graph from the library's own device search (tsne_knn_exact), joint P on the device,
perplexity 30, 2D, h2_tol 1e-2, rebuild every 3 (10 in early exaggeration), leaf 100,
explicit seeded N(0, 1e-4^2) initial embedding.

A step is one optimisation iteration of the library's loop on device-resident state.
--steps K --warmup W: one job of the full 1000-iteration schedule whose first 1000 - K
iterations run untimed (tsne_run_range; at least W of them), then the last K iterations
[1000 - K, 1000) are timed, bracketed by device synchronisation (an extra untimed
W-iteration job first when 1000 - K < W).  The metric is iterations/s of the
mixed-precision H2 path over that window; the plain FP32 path is timed the same way and
reported beside it, each with a whole-schedule timing and a per-phase split.  Prints ONE
JSON line.  Other modes: --serving (jobs/s, p50/p99), --knn (exact k-NN graph seconds),
--pipeline (k-NN + joint P + optimisation + quality, seconds per stage).
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "tsne-h2pack-python_amd"), os.path.join(ROOT, "oracle"),
           os.path.join(ROOT, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
FP32_VALU_PEAK_TFLOPS = 157.3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def knn_csr(be, X, k):
    """Exact k-NN graph on the GPU (tsne_knn_exact) in the CSR form the joint P takes:
    columns sorted within each row (the reference's neighbour indexing,
    _backend.py:651-653), float64 distances."""
    n = X.shape[0]
    idx, dist, _ = be.knn_exact(X, k)
    order = np.argsort(idx, axis=1, kind="stable")
    cols = np.ascontiguousarray(np.take_along_axis(idx, order, 1)).ravel().astype(np.int32)
    dd = np.ascontiguousarray(np.take_along_axis(dist, order, 1), dtype=np.float64).ravel()
    return dd, cols, np.arange(0, (n + 1) * k, k, dtype=np.int32)


def make_inputs(n, dim_in, n_clusters, k, perplexity, seed=0):
    """Seeded Gaussian mixture -> exact k-NN graph (the library's device search) -> joint P
    on the GPU (the library's perplexity kernel, symmetrisation and normalisation, equal
    bit for bit to the reference's scipy arithmetic).
    Returns (row_ptr, col_idx, values) int32/int32/float64 of the joint P."""
    from tsne_h2pack._backend import get_backend
    be = get_backend()
    X = mixture_points(n, dim_in, n_clusters, seed)
    t0 = time.time()
    dd, cols, indptr = knn_csr(be, X, k)
    del X
    log(f"k-NN graph (device, exact): {time.time() - t0:.1f}s")
    t0 = time.time()
    P = be.joint_P_device(dd, cols, indptr, perplexity)
    csr = (P.indptr.astype(np.int32), P.indices.astype(np.int32), P.data.astype(np.float64))
    del P
    log(f"joint P on the GPU (library kernels) nnz={csr[1].size}: {time.time() - t0:.2f}s")
    return csr


def make_workspace(be, csr, n, dim, precision, n_iter, leaf, tol, seed=0):
    """Workspace of the benchmark job: the reference's TSNE defaults for this shape
    (learning rate max(N/12, 200), rebuild every 3, 10 during early exaggeration) and an
    explicit seeded N(0, 1e-4^2) initial embedding (numpy default_rng(seed))."""
    from scipy.sparse import csr_matrix
    p = be.create_default_params()
    p.verbose = 0
    p.n_iter = n_iter
    p.h2_tol = tol
    p.h2_rebuild_freq = 3
    p.h2_rebuild_freq_early = 10
    p.h2_leaf_size = leaf
    p.learning_rate = max(n / 12.0, 200.0)
    p.random_seed = seed
    P = csr_matrix((csr[2], csr[1], csr[0]), shape=(n, n))
    Y0 = np.random.default_rng(seed).normal(scale=1e-4, size=(n, dim))
    ws = be.init_workspace(P, dim, p, init_Y=Y0)
    be.lib.tsne_set_h2_precision(ws, 1 if precision == "mixed" else 0)
    return ws


def rebuilds_in(lo, hi, freq=3, freq_early=10, exag_iter=250):
    """Rebuild iterations of the schedule in [lo, hi) (the reference's cadence,
    src/tsne_h2pack.c:1765-1776, with force reuse)."""
    def reb(it):
        eff = 1 if it < 20 else freq
        if 20 <= it < exag_iter and freq_early > eff:
            eff = freq_early
        return it > 0 and it % eff == 0
    return sum(reb(it) for it in range(lo, hi))


def _timings(lib, ws):
    a, r, pr, o = (ctypes.c_double() for _ in range(4))
    nr = ctypes.c_int()
    lib.tsne_run_timings(ws, ctypes.byref(a), ctypes.byref(r), ctypes.byref(pr),
                         ctypes.byref(o), ctypes.byref(nr))
    return dict(attr_ms=a.value, rebuild_ms=r.value, product_ms=pr.value, other_ms=o.value,
                n_rebuild=nr.value)


def time_path(be, csr, n, dim, precision, warmup, steps, leaf, tol, n_iter=1000,
              full_schedule=True, phase_split=True):
    """Steady-state window of one job: a job of n_iter iterations whose first
    n_iter - K iterations run untimed (tsne_run_range), then the last K iterations
    [n_iter - K, n_iter) are timed, bracketed by device synchronisation (K = steps).
    When the untimed prefix is shorter than `warmup`, an untimed warm-up job of `warmup`
    iterations runs first.  Also: the gradient error at the window's end; a fresh
    n_iter-iteration job timed whole (full_schedule); a third job with per-phase HIP
    events (synchronising every iteration, so kept out of both timed runs) for the phase
    split of the whole schedule and of the window."""
    lib = be.lib
    K = max(1, min(steps, n_iter))
    lo = n_iter - K
    if lo < warmup:
        ws = make_workspace(be, csr, n, dim, precision, warmup, leaf, tol)
        try:
            assert lib.tsne_run(ws) == 0
            lib.tsne_synchronize(ws)
        finally:
            lib.tsne_destroy(ws)
    ws = make_workspace(be, csr, n, dim, precision, n_iter, leaf, tol)
    try:
        if lo > 0:
            assert lib.tsne_run_range(ws, 0, lo) == 0
        # mode 2: HIP events on the job's stream around every fused attractive+update
        # launch of the window, read after the run (no per-iteration synchronisation)
        lib.tsne_enable_timing(ws, 2)
        lib.tsne_synchronize(ws)
        t0 = time.perf_counter()
        rc = lib.tsne_run_range(ws, lo, n_iter)
        lib.tsne_synchronize(ws)
        dt = time.perf_counter() - t0
        assert rc == 0
        ka, kn = ctypes.c_double(), ctypes.c_int()
        lib.tsne_attr_kernel_time(ws, ctypes.byref(ka), ctypes.byref(kn))
        attr_kernel_us = 1e3 * ka.value / max(kn.value, 1)
        kl = float(lib.compute_kl_divergence(ws))
        st = None
        from tsne_h2pack._backend import H2Stats
        s = H2Stats()
        if lib.tsne_h2_stats(ws, ctypes.byref(s)) == 0:
            st = {f: getattr(s, f) for f, _ in H2Stats._fields_}
        err = gradient_error(be, ws, n, dim)
        probe = attr_pattern_probe(lib, ws)
    finally:
        lib.tsne_destroy(ws)
    out = dict(window=[lo, n_iter], seconds=dt, its=K / dt, kl=kl,
               attr_kernel_us=attr_kernel_us, attr_kernel_launches=kn.value,
               window_rebuilds=rebuilds_in(lo, n_iter), h2=st, grad_err_final=err,
               attr_pattern_us=probe)
    if full_schedule:
        ws = make_workspace(be, csr, n, dim, precision, n_iter, leaf, tol)
        try:
            lib.tsne_synchronize(ws)
            t0 = time.perf_counter()
            assert lib.tsne_run(ws) == 0
            lib.tsne_synchronize(ws)
            tf = time.perf_counter() - t0
            out["full_schedule"] = dict(iterations=n_iter, seconds=tf, its=n_iter / tf,
                                        kl=float(lib.compute_kl_divergence(ws)))
        finally:
            lib.tsne_destroy(ws)
    if phase_split:
        ws = make_workspace(be, csr, n, dim, precision, n_iter, leaf, tol)
        try:
            lib.tsne_enable_timing(ws, 1)
            assert lib.tsne_run_range(ws, 0, lo) == 0
            t1 = _timings(lib, ws)
            assert lib.tsne_run_range(ws, lo, n_iter) == 0
            t2 = _timings(lib, ws)
            from tsne_h2pack._backend import H2Stats
            s = H2Stats()
            idr = None
            if lib.tsne_h2_stats(ws, ctypes.byref(s)) == 0 and s.id_ms > 0:
                idr = id_roofline(s.id_flops, s.id_ms, precision)
        finally:
            lib.tsne_destroy(ws)
        out["phases_full_schedule"] = t2
        out["phases_window"] = {k: t2[k] - t1[k] for k in t2}
        out["id_roofline"] = idr
    return out


def attr_pattern_probe(lib, ws, reps=20):
    """Average launch time (us) of the attractive pass's memory pattern over the job's
    final P and embedding: P streamed alone, P with the neighbour gathers, and the
    attractive sums without the update (tsne_attr_probe modes 0-2).  The fused kernel's
    time against the second one is its distance from its own access pattern's ceiling."""
    out = {}
    for mode, name in ((0, "p_stream"), (1, "p_and_gathers"), (2, "attractive_sums")):
        ms = ctypes.c_double()
        if lib.tsne_attr_probe(ws, mode, reps, ctypes.byref(ms)) != 0:
            return None
        out[name] = 1e3 * ms.value
    return out


def id_roofline(flops, ms, precision):
    """The last rebuild's interpolative decompositions (proxy-point pivoted Gram-Schmidt
    and interpolation solve, fused with the upward pass) against the FP64 vector peak:
    flops = sum over nodes of (3d+3) m p + 4 m p r (kernel evaluation of m points against
    p proxies, then the pivoted Gram-Schmidt to rank r), from tsne_h2_stats; time from HIP
    events around the ID launches of that build.  The mixed path runs the Gram-Schmidt in
    FP32 (peak 157.3 TFLOP/s)."""
    peak = 157.3 if precision == "mixed" else 78.6
    ach = flops / (ms * 1e-3) / 1e12
    return {"bound": "fp64" if precision != "mixed" else "fp32", "achieved": ach,
            "peak": peak, "unit": "TFLOP/s", "frac": ach / peak, "flops": flops,
            "id_ms": ms}


def gradient_error(be, ws, n, dim, n_rows=10000, seed=1):
    """Relative error of the product-path gradient against the FP64 exact gradient at
    exaggeration 1, ||g - g64|| / ||g64||, on all rows for n <= 100000 and on n_rows
    sampled rows above.  Also the repulsive-only error and both Z."""
    lib = be.lib
    g1 = np.zeros((dim, n))
    g0 = np.zeros((dim, n))
    Z = ctypes.c_double()
    if lib.tsne_compute_gradient(ws, 1.0, 0, g1, ctypes.byref(Z)) != 0:
        return None
    if lib.tsne_compute_gradient(ws, 0.0, 0, g0, ctypes.byref(Z)) != 0:
        return None
    rng = np.random.default_rng(seed)
    rows = (np.arange(n, dtype=np.int32) if n <= 100000 else
            np.sort(rng.choice(n, size=n_rows, replace=False)).astype(np.int32))
    rep64 = np.zeros((rows.size, dim))
    Ze = ctypes.c_double()
    if lib.tsne_exact_repulsive_rows(ws, rows, rows.size, rep64, ctypes.byref(Ze)) != 0:
        return None
    rep = -g0[:, rows].T
    attr = (g1 - g0)[:, rows].T
    g = attr - rep
    g64 = attr - rep64
    return dict(rel_l2=float(np.linalg.norm(g - g64) / np.linalg.norm(g64)),
                rep_rel_l2=float(np.linalg.norm(rep - rep64) / np.linalg.norm(rep64)),
                max_row_rel=float((np.linalg.norm(g - g64, axis=1) /
                                   np.maximum(np.linalg.norm(g64, axis=1), 1e-300)).max()),
                Z=Z.value, Z64=Ze.value, rows=int(rows.size))


def cpu_baseline(be, csr, n, dim, tol, leaf, n_iter=1000, window=20):
    """The oracle's FP64 C/OpenMP restatement of the reference's H2 path
    (oracle/tsne_oracle_h2.c: Morton tree, admissible and near lists, proxy-point ID with
    nested bases, upward / coupling / downward passes and the exact near field; pinned to
    the exact sum within h2_tol by tests/test_oracle_h2.py) timed on the host cores over the
    last `window` iterations [n_iter - window, n_iter) of the benchmark job.  The embedding
    at iteration n_iter - window comes from the library (the same job run on the device up
    to there, untimed); from it the oracle runs the window on the reference's schedule: a
    fresh CPU H2 build and product at the window's first iteration and at every rebuild
    iteration, the force reused between them, the attractive pass, the update and the
    centring every iteration, the KL at the last.  Momentum and gains start at 0 and 1
    (their values do not change the work).  The attractive pass and the update are also
    timed alone over all N."""
    import oracle
    from scipy.sparse import csr_matrix
    lo = n_iter - window
    ws = make_workspace(be, csr, n, dim, "plain", n_iter, leaf, tol)
    try:
        assert be.lib.tsne_run_range(ws, 0, lo) == 0
        Y = np.zeros((dim, n))
        be.lib.tsne_get_embedding(ws, Y)
    finally:
        be.lib.tsne_destroy(ws)
    P = csr_matrix((csr[2], csr[1], csr[0]), shape=(n, n))
    t0 = time.perf_counter()
    _, kl, builds = oracle.run_h2(P, Y, n_iter=n_iter, it_begin=lo, h2_tol=tol, leaf=leaf)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = oracle.attractive(P, Y, 1.0)
    t_attr = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.update(Y, np.zeros_like(Y), np.ones_like(Y), g, 0.8, max(n / 12.0, 200.0))
    t_upd = time.perf_counter() - t0
    t_h2 = max(dt - window * (t_attr + t_upd), 0.0) / max(builds, 1)
    return {"value": window / dt, "unit": "iters/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": (f"oracle FP64 CPU restatement of the reference's H2 path "
                       f"(oracle/tsne_oracle_h2.c), iterations [{lo}, {n_iter}) of the "
                       f"benchmark job from the library's embedding at iteration {lo}: "
                       f"{builds} H2 builds + products, {window} attractive passes and "
                       f"updates in {dt:.1f}s"),
            "window": [lo, n_iter], "seconds": dt, "h2_builds": builds, "kl": kl,
            "phases_s": {"attractive_per_iteration": t_attr, "update_per_iteration": t_upd,
                         "h2_build_and_product_per_build": t_h2},
            "reference_published_h2_its_12_threads_m5max": 36.0}


def attr_algorithmic_bytes(n, dim, nnz):
    """Bytes one attractive+update launch must move at least: P (int32 column + fp32
    value per nonzero, int32 row pointers), per row its own fp32 gather record, its fp32
    repulsive force (dim floats), FP64 Y/uY/gains read and written, and the next gather
    record written.  The neighbour gathers are not counted (they hit the caches)."""
    gath = 8 if dim == 2 else 16          # float2 / float4 gather record
    return 8 * nnz + 4 * (n + 1) + n * (2 * gath + 4 * dim + 3 * 2 * 8 * dim)


def attr_roofline(n, dim, nnz, head):
    """Roofline of the dominant kernel, the fused attractive + update pass
    (attr_update_kernel), from its launches inside the timed window: algorithmic bytes
    per launch / the average launch duration from HIP events on the job's stream.
    traffic: HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes over the same window of the same workload (None when no pass matches)."""
    kernel_us = head["attr_kernel_us"]
    by = attr_algorithmic_bytes(n, dim, nnz)
    ach = by / (kernel_us * 1e-6) / 1e9
    roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS, "traffic": None,
            "kernel": "attr_update_kernel", "kernel_avg_us": kernel_us,
            "algorithmic_bytes": by, "launches": head["attr_kernel_launches"],
            "window": head["window"]}
    t = measured_traffic(n, dim, nnz, head["window"])
    if t is not None:
        roof["traffic"] = t["bytes"]
        roof["traffic_source"] = t["source"]
    return roof


def measured_traffic(n, dim, nnz, window, kernel="attr_update_kernel"):
    """HBM bytes per attractive launch from the committed rocprofv3 counter passes
    (FETCH_SIZE and WRITE_SIZE, KiB per dispatch, averaged over the fused launches of
    the same timed window), each scaled by the ratio of a plain copy's known bytes to
    what the same counter read for it; None when no pass matches this workload."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        m = json.load(f)
    w = m.get("workload", {})
    if (w.get("n"), w.get("dim"), w.get("nnz"), w.get("window")) != (n, dim, nnz, list(window)):
        return None
    cf = m["copy_bytes"] / (m["copy_fetch_kib"] * 1024.0)
    cw = m["copy_bytes"] / (m["copy_write_kib"] * 1024.0)
    b = m["attr_fetch_kib"] * 1024.0 * cf + m["attr_write_kib"] * 1024.0 * cw
    return {"bytes": b, "source": os.path.relpath(path, ROOT)}


def copy_probe(mib=1024, reps=5):
    """A plain device copy of known size (torch; run under rocprofv3 --pmc to calibrate
    FETCH_SIZE / WRITE_SIZE): `reps` copies of `mib` MiB."""
    import torch
    a = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": mib * (1 << 20), "reps": reps}), flush=True)


def pmc_summary(paths, last=0):
    """Per-kernel mean of every counter over its dispatches (in dispatch order; with
    last > 0 over the last `last` dispatches of each kernel only, e.g. the fused launches
    of the timed window), from rocprofv3 counter_collection.csv files."""
    import csv
    import collections
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for p in paths:
        with open(p) as f:
            for row in csv.DictReader(f):
                acc[row["Kernel_Name"]][row["Counter_Name"]].append(
                    (int(row["Dispatch_Id"]), float(row["Counter_Value"])))
    for k in sorted(acc):
        for c, v in sorted(acc[k].items()):
            v = [x for _, x in sorted(v)]
            if last > 0:
                v = v[-last:]
            print(f"{c:12s} dispatches {len(v):6d} mean {np.mean(v):14.6g} KiB-or-count  {k[:90]}")


# ----------------------------------------------------------------------------- serving

def _serving_specs(n_jobs, seed=0):
    """Job shapes of the serving workload: N log-uniform in [2e3, 2e5], 5-50 clusters,
    every fourth job 3D (2D:3D = 3:1)."""
    rng = np.random.default_rng(seed)
    specs = []
    for j in range(n_jobs):
        n = int(round(np.exp(rng.uniform(np.log(2000), np.log(200000)))))
        specs.append({"n": n, "clusters": int(rng.integers(5, 51)), "dim": 3 if j % 4 == 3 else 2,
                      "seed": 1000 + j})
    return specs


def serving_inputs(specs, tmp):
    """Every job's exact k-NN graph (tsne_knn_exact) and joint P (device), saved to tmp."""
    from tsne_h2pack._backend import get_backend
    be = get_backend()
    t0 = time.time()
    for j, sp in enumerate(specs):
        dd, cols, indptr = knn_csr(be, mixture_points(sp["n"], 50, sp["clusters"], sp["seed"]), 90)
        P = be.joint_P_device(dd, cols, indptr, 30.0)
        np.savez(os.path.join(tmp, f"j{j}_P.npz"), ptr=P.indptr.astype(np.int32),
                 idx=P.indices.astype(np.int32), val=P.data)
    log(f"serving: k-NN graphs and joint P of {len(specs)} jobs on the GPU in "
        f"{time.time() - t0:.1f}s")


def _load_job_P(tmp, j, n):
    from scipy.sparse import csr_matrix
    z = np.load(os.path.join(tmp, f"j{j}_P.npz"))
    return csr_matrix((z["val"], z["idx"], z["ptr"]), shape=(n, n))


def _job_init(sp):
    return np.random.default_rng(sp["seed"]).normal(scale=1e-2, size=(sp["n"], sp["dim"]))


def _job_params(p, sp):
    """The TSNE front end's settings for a job of this dimension (h2_tol 1e-2 in 2D, 3e-2
    in 3D, rebuild every 3, 10 during early exaggeration, lr max(N/12, 200)), on a
    tsne_params_t from the library defaults."""
    p.verbose = 0
    p.learning_rate = max(sp["n"] / 12.0, 200.0)
    p.h2_tol = 1e-2 if sp["dim"] == 2 else 3e-2
    p.h2_rebuild_freq = 3
    p.h2_rebuild_freq_early = 10
    p.h2_leaf_size = 100
    return p


def serving_bench(n_jobs, slots_list, seq_sample, load=0.7, precision="plain", autotune=True):
    """Serving workload: capacity (all jobs queued at once), then Poisson arrivals at
    `load` x capacity with per-job latency (arrival -> result); baseline: one job at a
    time through tsne_init_with_P + tsne_run (sequential plain runs) on a sample.
    precision: the H2 precision of the served jobs (the baseline stays plain).
    autotune: first tune the launch shapes of every size bucket of the mix (untimed; the
    table then serves the baseline and the served jobs alike)."""
    import tempfile
    import shutil
    from tsne_h2pack._backend import get_backend
    from tsne_h2pack.serving import TSNEServer
    specs = _serving_specs(n_jobs)
    tmp = tempfile.mkdtemp(prefix="tsne_serving_")
    try:
        serving_inputs(specs, tmp)
        Ps = [_load_job_P(tmp, j, sp["n"]) for j, sp in enumerate(specs)]
        inits = [_job_init(sp) for sp in specs]
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    be = get_backend()
    lib = be.lib
    tuned = {}
    if autotune:
        from tsne_h2pack.serving import autotune_sizes
        t0 = time.perf_counter()
        dims = sorted({sp["dim"] for sp in specs})
        tuned = {f"{d}d_b{b}": v for (d, b), v in
                 autotune_sizes([sp["n"] for sp in specs], dims=dims).items()}
        log(f"serving: autotuned {len(tuned)} (dimension, size bucket) launch shapes in "
            f"{time.perf_counter() - t0:.1f}s: {tuned}")
    # sequential baseline on every (n_jobs // seq_sample)-th job
    sample = sorted(set(np.linspace(0, n_jobs - 1, min(seq_sample, n_jobs)).astype(int).tolist()))
    t_seq, kl_seq = 0.0, {}
    for j in sample:
        sp = specs[j]
        p = _job_params(be.create_default_params(), sp)
        t0 = time.perf_counter()
        ws = be.init_workspace(Ps[j], sp["dim"], p, init_Y=inits[j])
        assert lib.tsne_run(ws) == 0
        Y = np.zeros((sp["dim"], sp["n"]))
        lib.tsne_get_embedding(ws, Y)
        kl_seq[j] = float(lib.compute_kl_divergence(ws))
        lib.tsne_destroy(ws)
        t_seq += time.perf_counter() - t0
        log(f"serving: alone job {j} n={sp['n']} d={sp['dim']} {time.perf_counter() - t0:.2f}s")
    seq_rate = len(sample) / t_seq
    log(f"serving: sequential plain runs {seq_rate:.2f} jobs/s ({len(sample)} jobs, {t_seq:.1f}s)")

    slots = slots_list[0]

    def run(arrivals):
        with TSNEServer(slots=slots, h2_precision=precision) as srv:
            ids = []
            t0 = time.perf_counter()
            for j, sp in enumerate(specs):
                if arrivals is not None:
                    dt = arrivals[j] - (time.perf_counter() - t0)
                    if dt > 0:
                        time.sleep(dt)
                ids.append(srv.submit(Ps[j], sp["dim"],
                                      _job_params(srv.default_params(sp["n"]), sp),
                                      init_Y=inits[j]))
            infos = []
            for k, (i, sp) in enumerate(zip(ids, specs)):
                infos.append(srv.wait(i, sp["n"], sp["dim"])[1])
                if (k + 1) % 32 == 0:
                    log(f"serving: {k + 1}/{n_jobs} jobs done ({time.perf_counter() - t0:.1f}s)")
            wall = time.perf_counter() - t0
        return infos, wall

    caps = {}
    for sl in slots_list:
        slots = sl
        infos_s, wall_s = run(None)
        caps[sl] = n_jobs / wall_s
        log(f"serving: capacity {caps[sl]:.2f} jobs/s with {sl} slots "
            f"({n_jobs} jobs queued at once, {wall_s:.1f}s)")
        if caps[sl] >= max(caps.values()):
            infos_c, best = infos_s, sl
    slots = best
    cap = caps[best]
    rng = np.random.default_rng(1)
    rate = load * cap
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_jobs))
    infos_p, wall_p = run(arrivals)
    lat = np.array([i["latency_s"] for i in infos_p])
    done = np.array([i["end_s"] for i in infos_p])
    rate_p = n_jobs / (done.max() - min(i["submit_s"] for i in infos_p))
    kl_mix = {j: infos_c[j]["kl"] for j in sample}
    kl_dev = max(abs(kl_mix[j] - kl_seq[j]) / abs(kl_seq[j]) for j in sample)
    log(f"serving: Poisson arrivals at {rate:.2f} jobs/s: p50 {np.percentile(lat, 50):.2f}s "
        f"p99 {np.percentile(lat, 99):.2f}s")
    return {
        "metric": "jobs/s + p99 (serving: 256 t-SNE jobs, N log-uniform 2e3..2e5, 2D:3D 3:1)",
        "value": cap, "unit": "jobs/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic (seeded Gaussian mixtures, exact k-NN on GPU)",
        "config": {"workload": f"serving {n_jobs} jobs, 1000 iterations each, slots {slots}, "
                               f"{precision} H2 products (sequential baseline plain)",
                   "precision": precision,
                   "jobs": n_jobs, "slots": slots, "n_total": int(sum(s["n"] for s in specs)),
                   "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "default")},
        "capacity_jobs_per_s": cap, "capacity_by_slots": caps,
        "sequential_jobs_per_s": seq_rate, "sequential_sample": len(sample),
        "speedup_vs_sequential": cap / seq_rate,
        "poisson": {"offered_jobs_per_s": rate, "achieved_jobs_per_s": rate_p,
                    "p50_s": float(np.percentile(lat, 50)),
                    "p99_s": float(np.percentile(lat, 99)), "max_s": float(lat.max())},
        "kl_max_rel_diff_mix_vs_alone": kl_dev,  # 0 for plain; mixed vs plain otherwise
        "launch_shapes": tuned or "built-in defaults (no autotune)",
    }


def mixture_points(n, dim_in, n_clusters, seed):
    """The seeded Gaussian mixture of the k-NN inputs (same draws as _torch_knn)."""
    rng = np.random.default_rng(seed)
    centers = rng.normal(scale=5.0, size=(n_clusters, dim_in))
    w = rng.dirichlet(np.ones(n_clusters))
    std = rng.uniform(0.5, 2.0, size=n_clusters)
    labels = rng.choice(n_clusters, size=n, p=w)
    X = centers[labels] + rng.normal(size=(n, dim_in)) * std[labels, None]
    return X.astype(np.float32).astype(np.float64)  # float32 data, as the configs state


def knn_bench(n, dim_in, k, clusters, reps, cpu_rows):
    """Exact k-NN graph on the device (tsne_knn_exact): float64 mixture points, k
    neighbours.  value = seconds of device work per graph (HIP events from the first
    preparation kernel to the last kernel, inputs resident); the call's wall time (host
    validation, PCIe copies both ways) is reported beside it.  Roofline: the MFMA
    screening kernel, algorithmic flops 2 * 3 * n^2 * D (three BF16 products per pair)."""
    from tsne_h2pack._backend import get_backend
    be = get_backend()
    X = mixture_points(n, dim_in, clusters, 0)
    be.knn_exact(X[:4096], k)  # load the code objects
    dev, wall, scr = [], [], []
    for r in range(reps):
        t0 = time.perf_counter()
        idx, dist, st = be.knn_exact(X, k)
        wall.append(time.perf_counter() - t0)
        dev.append(st.device_ms * 1e-3)
        scr.append(st.screen_ms * 1e-3)
        log(f"knn rep {r}: device {st.device_ms:.1f} ms (screening {st.screen_ms:.1f} ms), "
            f"wall {wall[-1]:.2f} s, uncertain rows {st.uncertain_rows}")
    d_s, s_s = float(np.median(dev)), float(np.median(scr))
    alg = 2.0 * 3.0 * float(n) * float(n) * dim_in
    ach = alg / s_s / 1e12
    # spot check of the result against the oracle on a few rows
    import oracle
    rows = np.random.default_rng(1).choice(n, size=8, replace=False)
    ok = True
    for q in rows:
        d2 = ((X - X[q]) ** 2).sum(1)
        order = np.lexsort((np.arange(n), d2))[:k + 1]
        want = [c for c in order if c != q][:k]
        ok &= list(idx[q]) == want
    cpu = None
    if cpu_rows > 0:
        sub = np.ascontiguousarray(X[:cpu_rows])
        t0 = time.perf_counter()
        oracle.knn_exact(sub, k)
        t = time.perf_counter() - t0
        # the oracle is O(n^2 D + n^2 log n): scale the n_s x n_s run by (n / n_s)^2
        est = t * (n / cpu_rows) ** 2
        cpu = {"value": est, "unit": "s", "cores": oracle.num_threads(), "kind": "port",
               "sample": f"oracle FP64 brute force on the first {cpu_rows} points "
                         f"({t:.1f}s), scaled by (N/n)^2"}
    return {
        "metric": f"exact k-NN graph seconds at N={n} D={dim_in} k={k}",
        "value": d_s, "unit": "s", "n_gpus": 1, "steps": reps, "warmup": 1,
        "ms_per_step": d_s * 1e3, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16x3/f64",
        "data": "synthetic (seeded Gaussian mixture)",
        "config": {"workload": f"knn N={n} D={dim_in} k={k} clusters={clusters}", "n": n,
                   "dim_in": dim_in, "k": k},
        "wall_s": float(np.median(wall)), "screen_s": s_s,
        "uncertain_rows": int(st.uncertain_rows), "survivors": int(st.survivors),
        "spot_check_vs_bruteforce": bool(ok),
        "roofline": {"bound": "mfma", "achieved": ach, "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": ach / 2500.0, "traffic": None, "kernel": "knn_screen_kernel",
                     "algorithmic_flops": alg, "executed_flops": int(st.screen_flops)},
        "cpu_baseline": cpu,
    }


def pipeline_bench(n, dim, k, perplexity, clusters, cpu_quality):
    """Whole path on the device for one dataset: exact k-NN graph, the reference's
    neighbour indexing and the joint P in one call (tsne_joint_P_from_points), the
    1000-iteration optimisation (plain), and the quality score of the result
    (tsne_quality_score, the reference's scorer with k=10, 5000 queries, 1000 anchors;
    its high-dimensional neighbour rows come from a k=11 search, not counted in the
    total).  Seconds per stage, wall clock around each call."""
    from tsne_h2pack._backend import get_backend
    from tsne_h2pack.quality import score
    be = get_backend()
    X = mixture_points(n, 50, clusters, 0)
    be.knn_exact(X[:4096], k)  # load the code objects
    stages = {}
    t0 = time.perf_counter()
    P, kst = be.joint_P_from_points(X, k, perplexity)  # k-NN + indexing + joint P
    stages["knn_and_joint_P_s"] = time.perf_counter() - t0
    stages["knn_device_s"] = kst.device_ms * 1e-3
    t0 = time.perf_counter()
    idx, _, _ = be.knn_exact(X, 10 + 1)  # the scorer's high-dimensional neighbours
    stages["quality_knn_s"] = time.perf_counter() - t0
    csr = (P.indptr.astype(np.int32), P.indices.astype(np.int32), P.data)
    t0 = time.perf_counter()
    ws = make_workspace(be, csr, n, dim, "plain", 1000, 100, 1e-2)
    stages["init_s"] = time.perf_counter() - t0
    try:
        t0 = time.perf_counter()
        assert be.lib.tsne_run(ws) == 0
        be.lib.tsne_synchronize(ws)
        stages["optimise_s"] = time.perf_counter() - t0
        Y = np.zeros((dim, n))
        be.lib.tsne_get_embedding(ws, Y)
        kl_lib = float(be.lib.compute_kl_divergence(ws))
    finally:
        be.lib.tsne_destroy(ws)
    Y = Y.T.copy()
    t0 = time.perf_counter()
    q = score(Y, idx[:, :10], P, k=10)
    # (the scorer's hd rows: the graph's first 10 columns in distance order, which the
    # k=11 search gives exactly since ties are ordered by index in both)
    stages["quality_s"] = time.perf_counter() - t0
    total = sum(v for key, v in stages.items() if key not in ("knn_device_s", "quality_knn_s"))
    cpu = None
    if cpu_quality:
        import oracle
        r = np.random.default_rng(0)
        qs = r.choice(n, size=5000, replace=False)
        an = r.choice(n, size=1000, replace=False)
        t0 = time.perf_counter()
        qo = oracle.quality_score(Y, idx[:, :10], qs, P, an, k=10)
        cpu = {"value": time.perf_counter() - t0, "unit": "s", "cores": 1, "kind": "port",
               "sample": "the oracle's restatement of the reference's scorer (numpy + sklearn "
                         "KDTree) on the same embedding, for the quality stage only",
               "agrees": bool(qo["knn_preservation"] == q["knn_preservation"]
                              and abs(qo["kl"] - q["kl"]) < 1e-9 * abs(qo["kl"]))}
    log(f"pipeline stages {json.dumps(stages)} quality {q}")
    return {
        "metric": f"end-to-end seconds at N={n} {dim}D (exact k-NN k={k} + joint P + 1000 "
                  f"iterations + quality score)",
        "value": total, "unit": "s", "n_gpus": 1, "steps": 1000, "warmup": 0,
        "ms_per_step": None, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32/f64",
        "data": "synthetic (seeded Gaussian mixture, 50 features)",
        "config": {"workload": f"pipeline N={n} D=50 {dim}D k={k} perplexity={perplexity:g} "
                               f"clusters={clusters}", "n": n, "dim": dim, "k": k},
        "stages": stages, "quality": q, "kl_library": kl_lib,
        "knn_uncertain_rows": int(kst.uncertain_rows), "nnz_P": int(P.nnz),
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=250)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--dim", type=int, default=2)
    ap.add_argument("--leaf", type=int, default=100)
    ap.add_argument("--leaf-mixed", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-2)
    ap.add_argument("--k", type=int, default=90, help="neighbours of the exact k-NN graph")
    ap.add_argument("--perplexity", type=float, default=30.0)
    ap.add_argument("--clusters", type=int, default=0,
                    help="Gaussian clusters (default: 100 at N >= 1M, else 20)")
    ap.add_argument("--paths", default="mixed,plain")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-full-schedule", action="store_true",
                    help="skip the whole-schedule timing job of each path")
    ap.add_argument("--no-phase-split", action="store_true",
                    help="skip the per-phase instrumented job of each path")
    ap.add_argument("--save-inputs", default="", help="write the prepared P here and exit")
    ap.add_argument("--load-inputs", default="", help="read a P written by --save-inputs")
    ap.add_argument("--serving", action="store_true",
                    help="serving workload (jobs/s, p50/p99) instead of the 1M iteration rate")
    ap.add_argument("--jobs", type=int, default=256)
    ap.add_argument("--slots", default="12", help="concurrent job slots (comma list: each "
                    "is measured, the best one is used for the arrival run; 12 measured best "
                    "of 8, 12, 16 and 24 on the 256-job mix)")
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="GPU_MAX_HW_QUEUES for this process (serving; 0 leaves the default)")
    ap.add_argument("--seq-sample", type=int, default=32)
    ap.add_argument("--no-autotune", action="store_true",
                    help="serving: keep the built-in launch shapes (no tsne_autotune first)")
    ap.add_argument("--serving-precision", default="plain", choices=("plain", "mixed"),
                    help="H2 precision of the served jobs (the sequential baseline is plain)")
    ap.add_argument("--blocking-sync", action="store_true",
                    help="serving: host threads block (instead of spinning) while they wait "
                         "for the GPU (hipDeviceScheduleBlockingSync, set before HIP starts)")
    ap.add_argument("--knn", action="store_true",
                    help="exact k-NN graph on the GPU (seconds per graph) instead of the "
                         "iteration rate")
    ap.add_argument("--pipeline", action="store_true",
                    help="k-NN + joint P + optimisation + quality score, seconds per stage")
    ap.add_argument("--knn-cpu-rows", type=int, default=20000,
                    help="points of the oracle's CPU baseline run for --knn (0: none)")
    ap.add_argument("--tile-rows", type=int, default=-1,
                    help="rows per attractive tile (0: per-row kernel only; default: library's)")
    a = ap.parse_args()
    if a.tile_rows >= 0:
        os.environ["TSNE_ATTR_TILE_ROWS"] = str(a.tile_rows)
    if a.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the reference is "
              "single-process; jobs shard across GPUs by job, one server per device)",
              file=sys.stderr)
        sys.exit(2)
    if a.pipeline:
        clusters = a.clusters if a.clusters > 0 else (100 if a.n >= 1_000_000 else 20)
        print(json.dumps(pipeline_bench(a.n, a.dim, a.k, a.perplexity, clusters,
                                        not a.no_cpu_baseline)), flush=True)
        return
    if a.knn:
        clusters = a.clusters if a.clusters > 0 else (100 if a.n >= 1_000_000 else 20)
        print(json.dumps(knn_bench(a.n, 50, a.k, clusters, max(1, min(a.steps, 5)),
                                   0 if a.no_cpu_baseline else a.knn_cpu_rows)), flush=True)
        return
    if a.serving:
        if a.blocking_sync:
            hip = ctypes.CDLL("libamdhip64.so")
            rc = hip.hipSetDeviceFlags(ctypes.c_uint(4))  # hipDeviceScheduleBlockingSync
            log(f"serving: hipSetDeviceFlags(BlockingSync) -> {rc}")
        if a.hw_queues:
            os.environ["GPU_MAX_HW_QUEUES"] = str(min(max(a.hw_queues, 1), 16))
        slots = [int(x) for x in str(a.slots).split(",")]
        print(json.dumps(serving_bench(a.jobs, slots, a.seq_sample,
                                       precision=a.serving_precision,
                                       autotune=not a.no_autotune)), flush=True)
        return
    from tsne_h2pack._backend import get_backend
    be = get_backend()
    n, dim = a.n, a.dim
    clusters = a.clusters if a.clusters > 0 else (100 if n >= 1_000_000 else 20)
    if a.load_inputs:
        z = np.load(a.load_inputs)
        csr = (z["row_ptr"], z["col_idx"], z["values"])
        assert csr[0].size == n + 1
    else:
        csr = make_inputs(n, 50, clusters, a.k, a.perplexity)
    if a.save_inputs:
        np.savez(a.save_inputs, row_ptr=csr[0], col_idx=csr[1], values=csr[2])
        return
    res = {}
    n_iter = 1000
    for path in a.paths.split(","):
        leaf = a.leaf_mixed if (path == "mixed" and a.leaf_mixed) else a.leaf
        log(f"timing {path} (leaf {leaf}) warmup={a.warmup} steps={a.steps}")
        res[path] = time_path(be, csr, n, dim, path, a.warmup, a.steps, leaf, a.tol,
                              n_iter=n_iter, full_schedule=not a.no_full_schedule,
                              phase_split=not a.no_phase_split)
        log(f"{path}: {res[path]['its']:.1f} it/s {json.dumps(res[path])}")
    head = res.get("mixed") or res[next(iter(res))]
    nnz = int(csr[1].size)
    roof = attr_roofline(n, dim, nnz, head)
    cpu = None if a.no_cpu_baseline else cpu_baseline(be, csr, n, dim, a.tol, a.leaf)
    lo, hi = head["window"]
    out = {
        "metric": "iters/s at N=1M 2D, mixed vs plain FP32 GPU @<=1e-2 grad err; jobs/s + p99",
        "value": head["its"], "unit": "iters/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": 1e3 / head["its"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Gaussian mixture, float32 points, exact k-NN on GPU)",
        "config": {"workload": f"tsne-{dim}d N={n} D=50 clusters={clusters} k={a.k} "
                               f"perplexity={a.perplexity:g} h2_tol={a.tol} rebuild 3/10 "
                               f"leaf {a.leaf}; timed window = iterations [{lo}, {hi}) of "
                               f"a {n_iter}-iteration run ({head['window_rebuilds']} H2 "
                               f"rebuilds in the window), the {lo} iterations before it "
                               f"untimed",
                   "n": n, "dim": dim, "nnz_P": nnz, "paths": a.paths,
                   "window": [lo, hi]},
        "precision": {"plain": "near field, far field and attractive force FP32; ID, "
                               "perplexity, KL, Z and the update state FP64",
                      "mixed": "as plain, but the far-field coupling as BF16-in/FP32-acc "
                               "MFMA (nodes whose error estimate exceeds h2_tol redone in "
                               "FP32) and the ID's pivoted Gram-Schmidt in FP32 at ID "
                               "tolerance >= 1e-3 (R entries, solve and interpolation "
                               "coefficients FP64)",
                      "reciprocal": "1-ulp hardware reciprocal in the FP32 kernels"},
        "paths": res,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if "plain" in res and "mixed" in res:
        out["mixed_over_plain"] = res["mixed"]["its"] / res["plain"]["its"]
        if "full_schedule" in res["mixed"] and "full_schedule" in res["plain"]:
            out["full_schedule_its"] = {k: res[k]["full_schedule"]["its"] for k in res}
            out["mixed_over_plain_full_schedule"] = (res["mixed"]["full_schedule"]["its"] /
                                                     res["plain"]["full_schedule"]["its"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--copy-probe":
        copy_probe()
    elif len(sys.argv) > 1 and sys.argv[1] == "--pmc-summary":
        args = sys.argv[2:]
        last = 0
        if len(args) >= 2 and args[0] == "--last":
            last, args = int(args[1]), args[2:]
        pmc_summary(args, last)
    else:
        main()

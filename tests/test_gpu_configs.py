"""The benchmark configurations, each exercised on the GPU through the library.

* 5K, 2D, PCA init (the CPU-runnable configuration): the reference's front end
  TSNE(init='pca', knn_method='exact') on a 10-cluster mixture (5,000 x 50, centres
  N(0, 4^2), unit variance), k = 90, perplexity 30, 1000 iterations.  The final KL is
  within 1% of the oracle's run (exact repulsion, the same schedule) from the same P and
  the same PCA initial embedding.
* 1M, 2D (the headline): 100-cluster float32 mixture, k = 90, perplexity 30, h2_tol 1e-2,
  joint P from tsne_joint_P_from_points.  Mixed and plain paths: after the 250
  early-exaggeration iterations the gradient error against the FP64 exact gradient on
  10,000 sampled rows is <= 1e-2 (bar: h2_tol; the repulsive force alone <= 2 h2_tol, the
  bar of the other H2 tests); 200 of those rows of the FP64 exact
  repulsion are checked against the CPU oracle at the same Z (1e-9); then both run to
  iteration 1000 and their final KLs agree within 1%.
* 200K, 2D, the whole schedule: the H2 runs (mixed and plain) end within 1% of the final
  KL of the same run with the exact repulsion (direct O(N^2) sum) in place of every H2
  product, from the same start.
* 500K, 3D: the mixed path on the same check as the plain one (test_gpu_h2.py): 20,000
  sampled rows within 3e-2 of FP64 after 250 iterations.
* Serving: a seeded sample of the 256-job serving mix (8 jobs with N from 2,000 to
  200,000, log-spaced, 2D:3D = 3:1, 5-50 clusters, k = 90, perplexity 30, 1000 iterations
  each).  Each job's embedding and KL are the same alone and in the mix (1e-9), and the
  mix runs its jobs concurrently.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from datasets import gaussian_mixture
from tsne_h2pack._backend import H2Stats

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_config_5k_pca_front_end_kl_within_one_percent(backend):
    """TSNE(method='h2pack', init='pca', knn_method='exact') end to end vs the oracle."""
    from scipy.sparse import csr_matrix
    from tsne_h2pack import TSNE
    _say("5K 2D PCA init, front end")
    X, _ = gaussian_mixture(5000, dim=50, n_clusters=10, seed=0, spread=4.0)
    tsne = TSNE(n_components=2, perplexity=30.0, init='pca', knn_method='exact',
                random_state=0, method='h2pack')
    Y = tsne.fit_transform(X)
    kl = tsne.kl_divergence_
    # the oracle's run from the same P and the same initial embedding
    Y0 = tsne._initial_embedding(X, X.shape[0])
    from tsne_h2pack._knn import get_knn_method
    G = get_knn_method('exact').compute(X, k=90)
    P = backend.joint_P(G, 30.0)
    assert isinstance(P, csr_matrix)
    _, kl_o = oracle.run(P, np.ascontiguousarray(Y0.T), n_iter=1000)
    _say(f"  KL front end {kl:.5f} oracle {kl_o:.5f}")
    assert Y.shape == (5000, 2) and np.isfinite(Y).all()
    assert abs(kl - kl_o) / kl_o < 0.01


def _grad_check(backend, ws, n, dim, rows):
    g1 = np.zeros((dim, n))
    g0 = np.zeros((dim, n))
    Z = ctypes.c_double()
    assert backend.lib.tsne_compute_gradient(ws, 1.0, 0, g1, ctypes.byref(Z)) == 0
    assert backend.lib.tsne_compute_gradient(ws, 0.0, 0, g0, ctypes.byref(Z)) == 0
    rep64 = np.zeros((rows.size, dim))
    Ze = ctypes.c_double()
    assert backend.lib.tsne_exact_repulsive_rows(ws, rows, rows.size, rep64,
                                                 ctypes.byref(Ze)) == 0
    attr = (g1 - g0).T[rows]
    g = attr + g0.T[rows]
    g64 = attr - rep64
    err = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    rep_err = np.linalg.norm(-g0.T[rows] - rep64) / np.linalg.norm(rep64)
    return err, rep_err, Z.value, Ze.value, rep64


def _headline_run(backend, P, precision, rows):
    """One 1000-iteration job at the headline shape: the gradient check after 250
    iterations, then the rest of the schedule.  Returns the figures the test checks."""
    n, dim = P.shape[0], 2
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = 1000
    p.h2_tol = 1e-2
    p.h2_rebuild_freq = 3
    p.h2_rebuild_freq_early = 10
    p.learning_rate = n / 12.0
    Y0 = np.random.default_rng(0).normal(scale=1e-4, size=(n, dim))
    ws = backend.init_workspace(P, dim, p, init_Y=Y0)
    try:
        assert backend.lib.tsne_set_h2_precision(ws, 1 if precision == "mixed" else 0) == 0
        assert backend.lib.tsne_run_range(ws, 0, 250) == 0
        err, rep_err, Z, Ze, rep64 = _grad_check(backend, ws, n, dim, rows)
        st = H2Stats()
        assert backend.lib.tsne_h2_stats(ws, ctypes.byref(st)) == 0
        _say(f"  {precision}: grad err {err:.2e} rep err {rep_err:.2e} Z {Z:.6e}/{Ze:.6e} "
             f"nodes {st.n_nodes} levels {st.levels} rank max {st.rank_max} mixed fallbacks "
             f"{st.mixed_fallback_nodes} rank-capped builds {st.rank_capped_builds}")
        # the FP64 check kernel against the CPU oracle on 200 of the rows, at the same Z
        Y = np.zeros((dim, n))
        backend.lib.tsne_get_embedding(ws, Y)
        sub = rows[::50]
        rep_cpu = oracle.repulsive_rows(Y, Ze, sub)
        cpu_rel = np.abs(rep_cpu - rep64[::50]).max() / np.abs(rep_cpu).max()
        _say(f"  {precision}: FP64 device rows vs CPU oracle ({sub.size} rows): {cpu_rel:.2e}")
        assert backend.lib.tsne_run_range(ws, 250, 1000) == 0
        kl = backend.lib.compute_kl_divergence(ws)
        nr, nn = ctypes.c_int(), ctypes.c_int()
        assert backend.lib.tsne_run_state(ws, ctypes.byref(nr), ctypes.byref(nn), None, None) == 0
    finally:
        backend.lib.tsne_destroy(ws)
    _say(f"  {precision}: final KL {kl:.5f}; rebuilds {nr.value}, renumberings {nn.value}")
    return dict(err=err, rep_err=rep_err, Z=Z, Ze=Ze, cpu_rel=cpu_rel, kl=kl,
                n_rebuild=nr.value, n_renumber=nn.value)


@pytest.fixture(scope="module")
def p_1m(backend):
    """The headline's joint P: 1M points of a seeded 100-cluster float32 mixture, D = 50,
    exact k = 90 graph and perplexity 30 on the device."""
    n = 1_000_000
    X, _ = gaussian_mixture(n, dim=50, n_clusters=100, seed=0, spread=5.0, dirichlet=True,
                            std_range=(0.5, 2.0), dtype=np.float32)
    P, _ = backend.joint_P_from_points(X.astype(np.float64), 90, 30.0)
    return P


def test_config_1m_2d_mixed_and_plain(backend, p_1m):
    _say("1M 2D: gradient on 10000 sampled rows after 250 iterations, mixed and plain")
    n = 1_000_000
    P = p_1m
    rows = np.sort(np.random.default_rng(1).choice(n, size=10_000, replace=False)).astype(np.int32)
    res = {pr: _headline_run(backend, P, pr, rows) for pr in ("mixed", "plain")}
    for r in res.values():
        assert r["err"] <= 1e-2
        assert r["rep_err"] <= 2e-2  # the repulsive force alone: 2 h2_tol, as at every scale
        assert abs(r["Z"] / r["Ze"] - 1) <= 1e-2
        assert r["cpu_rel"] <= 1e-9
        assert r["n_rebuild"] == 292 and r["n_renumber"] == 6
    assert abs(res["mixed"]["kl"] - res["plain"]["kl"]) / res["plain"]["kl"] < 0.01


def _full_run_kl(backend, P, precision, mode, Y0):
    """mode 'h2': the H2 path; 'exact': method='exact' (direct sum every iteration);
    'exact_reuse': the direct sum on the H2 path's rebuild cadence and force reuse."""
    import os
    n, dim = P.shape[0], 2
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = 1000
    p.h2_tol = 1e-2
    p.h2_rebuild_freq = 999999 if mode == "exact" else 3  # > 1e5: method='exact'
    p.h2_rebuild_freq_early = 10
    p.learning_rate = n / 12.0
    if mode == "exact_reuse":
        os.environ["TSNE_DIRECT_REUSE"] = "1"
    try:
        ws = backend.init_workspace(P, dim, p, init_Y=Y0)
    finally:
        os.environ.pop("TSNE_DIRECT_REUSE", None)
    try:
        assert backend.lib.tsne_set_h2_precision(ws, 1 if precision == "mixed" else 0) == 0
        assert backend.lib.tsne_run(ws) == 0
        kl = backend.lib.compute_kl_divergence(ws)
        nr = ctypes.c_int()
        assert backend.lib.tsne_run_state(ws, ctypes.byref(nr), None, None, None) == 0
    finally:
        backend.lib.tsne_destroy(ws)
    return kl, nr.value


def test_config_200k_final_kl_against_exact_repulsion(backend):
    """The whole schedule at N = 200,000 (2D, the headline's data model, k = 90,
    perplexity 30, h2_tol 1e-2).  Anchor: the same run with the exact repulsion (direct
    O(N^2) sum, FP32 products) in place of every H2 product, on the same rebuild cadence
    and force reuse, from the same initial embedding; the H2 runs, mixed and plain, end
    within 1% of its final KL.  method='exact' (a fresh direct sum every iteration, no
    reuse, as the reference's exact path) is reported beside it: it differs from both by
    the reuse of the repulsive force between rebuilds as well."""
    n = 200_000
    X, _ = gaussian_mixture(n, dim=50, n_clusters=40, seed=5, spread=5.0, dirichlet=True,
                            std_range=(0.5, 2.0), dtype=np.float32)
    P, _ = backend.joint_P_from_points(X.astype(np.float64), 90, 30.0)
    del X
    Y0 = np.random.default_rng(2).normal(scale=1e-4, size=(n, 2))
    _say("200K 2D: exact repulsion on the H2 schedule")
    kl_ref, nr_ref = _full_run_kl(backend, P, "plain", "exact_reuse", Y0)
    kl_exact, _ = _full_run_kl(backend, P, "plain", "exact", Y0)
    _say(f"  exact on the H2 schedule: final KL {kl_ref:.5f}; method='exact': {kl_exact:.5f}")
    assert nr_ref == 0  # no H2 structure was built
    for precision in ("plain", "mixed"):
        kl, nr = _full_run_kl(backend, P, precision, "h2", Y0)
        _say(f"  H2 {precision}: final KL {kl:.5f} ({(kl / kl_ref - 1) * 100:+.3f}% vs exact on "
             f"the schedule, {(kl / kl_exact - 1) * 100:+.3f}% vs method='exact'), rebuilds {nr}")
        assert nr == 292
        assert abs(kl - kl_ref) / kl_ref < 0.01


def test_config_1m_final_kl_against_exact_repulsion(backend, p_1m):
    """The headline's final KL anchored to exact repulsion: the whole 1M 2D schedule with
    the direct O(N^2) sum (FP32 products, ~0.2 s each on the device) in place of every H2
    product, on the same rebuild cadence and force reuse, from the same initial embedding;
    the H2 runs, mixed and plain, end within 1% of its final KL."""
    n = 1_000_000
    Y0 = np.random.default_rng(0).normal(scale=1e-4, size=(n, 2))
    _say("1M 2D: exact repulsion on the H2 schedule (293 direct sums)")
    kl_ref, nr_ref = _full_run_kl(backend, p_1m, "plain", "exact_reuse", Y0)
    _say(f"  exact on the H2 schedule: final KL {kl_ref:.5f}")
    assert nr_ref == 0  # no H2 structure was built
    for precision in ("plain", "mixed"):
        kl, nr = _full_run_kl(backend, p_1m, precision, "h2", Y0)
        _say(f"  H2 {precision}: final KL {kl:.5f} ({(kl / kl_ref - 1) * 100:+.3f}%), rebuilds {nr}")
        assert nr == 292
        assert abs(kl - kl_ref) / kl_ref < 0.01


def test_config_500k_3d_mixed_sampled_rows(backend):
    _say("500K 3D mixed: gradient on 20000 sampled rows after 250 iterations")
    n, dim = 500_000, 3
    X, _ = gaussian_mixture(n, dim=50, n_clusters=50, seed=1, spread=5.0, dirichlet=True,
                            std_range=(0.5, 2.0), dtype=np.float32)
    P, _ = backend.joint_P_from_points(X.astype(np.float64), 150, 50.0)
    del X
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = 1000
    p.h2_tol = 3e-2
    p.h2_rebuild_freq = 3
    p.h2_rebuild_freq_early = 10
    p.learning_rate = n / 12.0
    Y0 = np.random.default_rng(1).normal(scale=1e-4, size=(n, dim))
    ws = backend.init_workspace(P, dim, p, init_Y=Y0)
    rows = np.sort(np.random.default_rng(2).choice(n, size=20_000, replace=False)).astype(np.int32)
    try:
        assert backend.lib.tsne_set_h2_precision(ws, 1) == 0
        assert backend.lib.tsne_run_range(ws, 0, 250) == 0
        err, rep_err, Z, Ze, _ = _grad_check(backend, ws, n, dim, rows)
        st = H2Stats()
        assert backend.lib.tsne_h2_stats(ws, ctypes.byref(st)) == 0
    finally:
        backend.lib.tsne_destroy(ws)
    _say(f"  grad err {err:.2e} rep err {rep_err:.2e} Z {Z:.6e}/{Ze:.6e} mixed fallback nodes "
         f"{st.mixed_fallback_nodes} of {st.n_nodes}")
    assert err <= 3e-2
    assert rep_err <= 3e-2
    assert abs(Z / Ze - 1) <= 3e-2


def _serving_sample():
    """8 jobs of the serving mix: N log-spaced in [2000, 200000], every fourth one 3D,
    clusters drawn from [5, 50], seeded."""
    rng = np.random.default_rng(7)
    ns = np.round(np.exp(np.linspace(np.log(2000), np.log(200000), 8))).astype(int)
    return [dict(n=int(n), dim=3 if j % 4 == 3 else 2, clusters=int(rng.integers(5, 51)),
                 seed=100 + j) for j, n in enumerate(ns)]


def test_config_serving_sample_alone_equals_mix(backend):
    from tsne_h2pack.serving import TSNEServer
    specs = _serving_sample()
    jobs = []
    for sp in specs:
        X, _ = gaussian_mixture(sp["n"], dim=50, n_clusters=sp["clusters"], seed=sp["seed"],
                                spread=5.0, dirichlet=True, std_range=(0.5, 2.0),
                                dtype=np.float32)
        P, _ = backend.joint_P_from_points(X.astype(np.float64), 90, 30.0)
        Y0 = np.random.default_rng(sp["seed"]).normal(scale=1e-4, size=(sp["n"], sp["dim"]))
        jobs.append((P, sp["dim"], Y0))
    alone = []
    for P, dim, Y0 in jobs:
        n = P.shape[0]
        _say(f"alone n={n} dim={dim}")
        p = backend.create_default_params()
        p.verbose = 0
        p.learning_rate = max(n / 12.0, 200.0)
        p.h2_tol = 1e-2 if dim == 2 else 3e-2
        p.h2_rebuild_freq = 3
        p.h2_rebuild_freq_early = 10
        ws = backend.init_workspace(P, dim, p, init_Y=Y0)
        try:
            assert backend.lib.tsne_run(ws) == 0
            Y = np.zeros((dim, n))
            backend.lib.tsne_get_embedding(ws, Y)
            alone.append((Y.T.copy(), backend.lib.compute_kl_divergence(ws)))
        finally:
            backend.lib.tsne_destroy(ws)
    with TSNEServer(slots=8) as srv:
        ids = [srv.submit(P, dim, srv.default_params(P.shape[0], h2_tol=1e-2 if dim == 2 else 3e-2,
                                                      h2_rebuild_freq=3, h2_rebuild_freq_early=10),
                          init_Y=Y0) for P, dim, Y0 in jobs]
        got = [srv.wait(j, P.shape[0], dim) for j, (P, dim, _) in zip(ids, jobs)]
    infos = [g[1] for g in got]
    overlap = sum(1 for a in infos for b in infos
                  if a is not b and a["start_s"] < b["end_s"] and b["start_s"] < a["end_s"]) // 2
    _say(f"overlapping job pairs: {overlap}")
    assert overlap > 0
    for (Ya, kla), (Ym, info) in zip(alone, got):
        rel = np.abs(Ya - Ym).max() / np.abs(Ya).max()
        _say(f"  n={info['n_point']} d={info['low_dim']} rel diff {rel:.2e} KL {kla:.5f}/{info['kl']:.5f}")
        assert np.isfinite(Ym).all()
        assert rel <= 1e-9
        assert abs(kla - info["kl"]) <= 1e-9 * abs(kla)

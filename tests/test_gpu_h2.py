"""H2 repulsive product on the GPU against the FP64 oracle (all rows), through the C ABI.

Bar (the h2_tol bound): the gradient error ||g - g64|| / ||g64|| <= h2_tol (1e-2 in 2D)
at exaggeration 1 and 12, on embeddings of several scales (cluster spread from 0.05 to
200 units, which puts the node boxes on both sides of the kernel's unit length).  The
repulsive force alone must be within 2 * h2_tol.  Determinism: two builds of the same
embedding give bit-identical forces.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from datasets import gaussian_mixture, init_normal, joint_P_oracle, knn_graph
from tsne_h2pack._backend import H2Stats

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def problem():
    X, labels = gaussian_mixture(6000, dim=20, n_clusters=12, seed=3)
    P = joint_P_oracle(knn_graph(X, 45), 15.0)
    return X, labels, P


def _embedding(n, dim, scale, seed=0):
    rng = np.random.default_rng(seed)
    cent = rng.normal(scale=scale, size=(30, dim))
    lab = rng.integers(0, 30, n)
    Y = cent[lab] + rng.normal(size=(n, dim)) * rng.uniform(0.5, 2.0, 30)[lab, None] * scale / 20
    return Y.T.copy()


def _ws(backend, P, Y0, dim, tol, leaf=100):
    p = backend.create_default_params()
    p.verbose = 0
    p.h2_tol = tol
    p.h2_rebuild_freq = 3
    p.h2_leaf_size = leaf
    return backend.init_workspace(P, dim, p, init_Y=Y0.T.copy())


def _grad(backend, ws, dim, n, exag, exact):
    g = np.zeros((dim, n))
    Z = ctypes.c_double()
    assert backend.lib.tsne_compute_gradient(ws, exag, exact, g, ctypes.byref(Z)) == 0
    return g, Z.value


@pytest.mark.parametrize("scale", [0.05, 2.0, 8.0, 40.0, 200.0])
def test_h2_gradient_within_tol_2d(backend, problem, scale):
    _say(f"h2 2d scale={scale}")
    _, _, P = problem
    n = P.shape[0]
    Y0 = _embedding(n, 2, scale)
    ws = _ws(backend, P, Y0, 2, 1e-2)
    try:
        st = H2Stats()
        for exag in (1.0, 12.0):
            g, Z = _grad(backend, ws, 2, n, exag, 0)
            rep64, Z64, _, _ = oracle.repulsive_exact(Y0)
            g64 = oracle.attractive(P, Y0, exag) - rep64
            err = np.linalg.norm(g - g64) / np.linalg.norm(g64)
            g0, _ = _grad(backend, ws, 2, n, 0.0, 0)
            rep_err = np.linalg.norm(-g0 - rep64) / np.linalg.norm(rep64)
            assert backend.lib.tsne_h2_stats(ws, ctypes.byref(st)) == 0
            _say(f"  exag {exag}: grad err {err:.2e} rep err {rep_err:.2e} Z {Z:.6e}/{Z64:.6e} "
                 f"levels {st.levels} nodes {st.n_nodes} rank max {st.rank_max} "
                 f"mean {st.rank_mean:.1f} proxies {st.proxy_max}")
            assert err <= 1e-2
            assert rep_err <= 2e-2
            assert abs(Z / Z64 - 1) <= 1e-2
        assert st.levels >= 3 and st.n_adm_pairs > 0
    finally:
        backend.lib.tsne_destroy(ws)


def test_h2_product_is_deterministic(backend, problem):
    _, _, P = problem
    n = P.shape[0]
    Y0 = _embedding(n, 2, 20.0, seed=5)
    ws = _ws(backend, P, Y0, 2, 1e-2)
    try:
        g1, Z1 = _grad(backend, ws, 2, n, 1.0, 0)
        g2, Z2 = _grad(backend, ws, 2, n, 1.0, 0)
    finally:
        backend.lib.tsne_destroy(ws)
    assert Z1 == Z2
    assert np.array_equal(g1, g2)


@pytest.mark.parametrize("leaf", [16, 64, 128])
def test_h2_leaf_sizes(backend, problem, leaf):
    """h2_leaf_size changes speed and memory only (reference manifold.py:150-152)."""
    _, _, P = problem
    n = P.shape[0]
    Y0 = _embedding(n, 2, 30.0, seed=7)
    ws = _ws(backend, P, Y0, 2, 1e-2, leaf=leaf)
    try:
        g, _ = _grad(backend, ws, 2, n, 1.0, 0)
    finally:
        backend.lib.tsne_destroy(ws)
    rep64, _, _, _ = oracle.repulsive_exact(Y0)
    g64 = oracle.attractive(P, Y0, 1.0) - rep64
    assert np.linalg.norm(g - g64) / np.linalg.norm(g64) <= 1e-2


def test_h2_run_kl_within_one_percent(backend, problem):
    """Full 1000-iteration run (H2 + force reuse, rebuild 3/10) ends within 1% of the
    oracle's KL (exact repulsion, same schedule) from the same init."""
    _say("h2 full run N=6000")
    _, labels, P = problem
    n = P.shape[0]
    Y0 = init_normal(n, 2, 11)
    p = backend.create_default_params()
    p.verbose = 0
    p.h2_tol = 1e-2
    p.h2_rebuild_freq = 3
    p.h2_rebuild_freq_early = 10
    p.learning_rate = max(n / 12.0, 200.0)
    ws = backend.init_workspace(P, 2, p, init_Y=Y0.T.copy())
    try:
        assert backend.lib.tsne_run(ws) == 0
        kl = backend.lib.compute_kl_divergence(ws)
        Y = np.zeros((2, n))
        backend.lib.tsne_get_embedding(ws, Y)
    finally:
        backend.lib.tsne_destroy(ws)
    _, kl_o = oracle.run(P, Y0, n_iter=1000)
    _say(f"  KL gpu {kl:.5f} oracle {kl_o:.5f}")
    assert abs(kl - kl_o) / kl_o < 0.01
    assert np.isfinite(Y).all()


def _run(backend, P, Y0, n_iter, reorder):
    import os
    old = os.environ.get("TSNE_REORDER")
    os.environ["TSNE_REORDER"] = "1" if reorder else "0"
    try:
        p = backend.create_default_params()
        p.verbose = 0
        p.n_iter = n_iter
        p.h2_tol = 1e-2
        p.h2_rebuild_freq = 3
        p.h2_rebuild_freq_early = 10
        p.learning_rate = max(P.shape[0] / 12.0, 200.0)
        ws = backend.init_workspace(P, 2, p, init_Y=Y0.T.copy())
    finally:
        if old is None:
            del os.environ["TSNE_REORDER"]
        else:
            os.environ["TSNE_REORDER"] = old
    try:
        assert backend.lib.tsne_run(ws) == 0
        Y = np.zeros((2, P.shape[0]))
        backend.lib.tsne_get_embedding(ws, Y)
    finally:
        backend.lib.tsne_destroy(ws)
    return Y


def test_renumbering_is_transparent(backend, problem):
    """Reference tests/test_optimizations.py:77-95: the first renumbering happens at
    iteration 30; the embedding comes back in the caller's order and matches a run
    without renumbering up to float32 summation-order effects (a mis-tracked permutation
    gives an O(1) difference)."""
    _say("renumbering transparency")
    _, _, P = problem
    Y0 = init_normal(P.shape[0], 2, 21)
    ref = _run(backend, P, Y0, 34, False)
    ren = _run(backend, P, Y0, 34, True)
    rel = np.abs(ref - ren).max() / np.abs(ref).max()
    _say(f"  rel diff {rel:.2e}")
    assert rel < 1e-3


def test_renumbering_keeps_structure(backend, problem):
    """After all six renumberings the rows of the output still line up with the input:
    10-NN overlap with the high-dimensional neighbours (reference
    tests/test_optimizations.py:98-111)."""
    from sklearn.neighbors import NearestNeighbors
    X, _, P = problem
    Y0 = init_normal(P.shape[0], 2, 22)
    Y = _run(backend, P, Y0, 1000, True).T
    k = 10
    hd = NearestNeighbors(n_neighbors=k).fit(X).kneighbors(return_distance=False)
    ld = NearestNeighbors(n_neighbors=k).fit(Y).kneighbors(return_distance=False)
    overlap = np.mean([len(set(a) & set(b)) for a, b in zip(hd, ld)]) / k
    _say(f"  10-NN overlap {overlap:.3f}")
    assert overlap > 0.15


def test_default_tol_100k_early_iterations(backend):
    """Reference defaults (h2_tol 1e-3, rebuild every 10) on N=100,000 from a compressed
    start (init scale 1e-2): the U pool grows past its first size, ranks reach the cap,
    and the interaction lists fill several rebuilds of list and sort buffers.  Regression
    test for pair lists written past their buffers after the sort scratch was exchanged
    with a smaller list."""
    from scipy.sparse import csr_matrix
    _say("default tol N=100000")
    n, k = 100_000, 30
    X, _ = gaussian_mixture(n, dim=10, n_clusters=30, seed=8, dirichlet=True,
                            std_range=(0.5, 2.0), dtype=np.float32)
    G = knn_graph(X, k)
    P = joint_P_oracle(G, 10.0)
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = 40
    p.learning_rate = n / 12.0
    Y0 = np.random.default_rng(8).normal(scale=1e-2, size=(n, 2))
    ws = backend.init_workspace(P, 2, p, init_Y=Y0)
    try:
        assert backend.lib.tsne_run(ws) == 0
        st = H2Stats()
        assert backend.lib.tsne_h2_stats(ws, ctypes.byref(st)) == 0
        Y = np.zeros((2, n))
        backend.lib.tsne_get_embedding(ws, Y)
    finally:
        backend.lib.tsne_destroy(ws)
    _say(f"  nodes {st.n_nodes} far pairs {st.n_adm_pairs} near pairs {st.n_near_pairs} "
         f"rank max {st.rank_max}")
    assert np.isfinite(Y).all()


@pytest.mark.parametrize("scale", [0.05, 3.0, 30.0])
def test_h2_gradient_within_tol_3d(backend, problem, scale):
    """3D embeddings at h2_tol 3e-2 (the 3D bound): gradient error <= 3e-2 against the
    FP64 exact gradient on all rows, exaggeration 1 and 12."""
    _say(f"h2 3d scale={scale}")
    _, _, P = problem
    n = P.shape[0]
    Y0 = _embedding(n, 3, scale, seed=4)
    ws = _ws(backend, P, Y0, 3, 3e-2)
    try:
        st = H2Stats()
        for exag in (1.0, 12.0):
            g, Z = _grad(backend, ws, 3, n, exag, 0)
            rep64, Z64, _, _ = oracle.repulsive_exact(Y0)
            g64 = oracle.attractive(P, Y0, exag) - rep64
            err = np.linalg.norm(g - g64) / np.linalg.norm(g64)
            assert backend.lib.tsne_h2_stats(ws, ctypes.byref(st)) == 0
            _say(f"  exag {exag}: grad err {err:.2e} Z {Z:.6e}/{Z64:.6e} levels {st.levels} "
                 f"rank max {st.rank_max} mean {st.rank_mean:.1f} proxies {st.proxy_max}")
            assert err <= 3e-2
            assert abs(Z / Z64 - 1) <= 3e-2
        assert st.levels >= 3 and st.n_adm_pairs > 0
    finally:
        backend.lib.tsne_destroy(ws)


def test_plain_gradient_all_rows_100k(backend):
    """N=100,000 (20 Gaussians, Dirichlet weights, per-cluster std U[0.5,2], centres
    N(0,5^2), D=50 float32; exact k=90 graph; perplexity 30; 2D; h2_tol 1e-2): after the
    250 early-exaggeration iterations the plain FP32 gradient is within 1e-2 of the FP64
    exact gradient on all rows (the exact repulsion comes from the FP64 all-pairs sum on
    the device)."""
    _say("plain gradient all rows N=100000")
    n = 100_000
    X, _ = gaussian_mixture(n, dim=50, n_clusters=20, seed=0, spread=5.0, dirichlet=True,
                            std_range=(0.5, 2.0), dtype=np.float32)
    P = joint_P_oracle(knn_graph(X, 90), 30.0)
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = 1000
    p.h2_tol = 1e-2
    p.h2_rebuild_freq = 3
    p.h2_rebuild_freq_early = 10
    p.learning_rate = n / 12.0
    Y0 = np.random.default_rng(0).normal(scale=1e-4, size=(n, 2))
    ws = backend.init_workspace(P, 2, p, init_Y=Y0)
    try:
        assert backend.lib.tsne_run_range(ws, 0, 250) == 0
        g1 = np.zeros((2, n))
        g0 = np.zeros((2, n))
        Z = ctypes.c_double()
        assert backend.lib.tsne_compute_gradient(ws, 1.0, 0, g1, ctypes.byref(Z)) == 0
        assert backend.lib.tsne_compute_gradient(ws, 0.0, 0, g0, ctypes.byref(Z)) == 0
        rows = np.arange(n, dtype=np.int32)
        rep64 = np.zeros((n, 2))
        Ze = ctypes.c_double()
        assert backend.lib.tsne_exact_repulsive_rows(ws, rows, n, rep64, ctypes.byref(Ze)) == 0
    finally:
        backend.lib.tsne_destroy(ws)
    attr = (g1 - g0).T
    g = attr + g0.T
    g64 = attr - rep64
    err = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    rep_err = np.linalg.norm(-g0.T - rep64) / np.linalg.norm(rep64)
    _say(f"  grad err {err:.2e} rep err {rep_err:.2e} Z {Z.value:.6e}/{Ze.value:.6e}")
    assert err <= 1e-2
    assert rep_err <= 1e-2


def test_plain_gradient_3d_500k_sampled_rows(backend):
    """N=500,000, 3D (50 Gaussians in D=50, Dirichlet weights, per-cluster std U[0.5,2],
    centres N(0,5^2); k=150 graph and joint P from tsne_joint_P_from_points, perplexity 50;
    h2_tol 3e-2): after the 250 early-exaggeration iterations the plain gradient is within
    3e-2 of the FP64 exact gradient on 20,000 sampled rows.  This is the size of the 500K 3D
    benchmark, whose widest levels (thousands of nodes) run their IDs split over the 64-,
    128- and 384-thread classes."""
    _say("plain gradient 3D N=500000 (20000 sampled rows)")
    n = 500_000
    X, _ = gaussian_mixture(n, dim=50, n_clusters=50, seed=1, spread=5.0, dirichlet=True,
                            std_range=(0.5, 2.0))
    P, _ = backend.joint_P_from_points(X, 150, 50.0)
    del X
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = 1000
    p.h2_tol = 3e-2
    p.h2_rebuild_freq = 3
    p.h2_rebuild_freq_early = 10
    p.learning_rate = n / 12.0
    Y0 = np.random.default_rng(1).normal(scale=1e-4, size=(n, 3))
    ws = backend.init_workspace(P, 3, p, init_Y=Y0)
    rows = np.sort(np.random.default_rng(2).choice(n, size=20_000, replace=False)).astype(np.int32)
    try:
        assert backend.lib.tsne_run_range(ws, 0, 250) == 0
        g1 = np.zeros((3, n))
        g0 = np.zeros((3, n))
        Z = ctypes.c_double()
        assert backend.lib.tsne_compute_gradient(ws, 1.0, 0, g1, ctypes.byref(Z)) == 0
        assert backend.lib.tsne_compute_gradient(ws, 0.0, 0, g0, ctypes.byref(Z)) == 0
        rep64 = np.zeros((rows.size, 3))
        Ze = ctypes.c_double()
        assert backend.lib.tsne_exact_repulsive_rows(ws, rows, rows.size, rep64,
                                                     ctypes.byref(Ze)) == 0
        st = H2Stats()
        assert backend.lib.tsne_h2_stats(ws, ctypes.byref(st)) == 0
    finally:
        backend.lib.tsne_destroy(ws)
    attr = (g1 - g0).T[rows]
    g = attr + g0.T[rows]
    g64 = attr - rep64
    err = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    rep_err = np.linalg.norm(-g0.T[rows] - rep64) / np.linalg.norm(rep64)
    _say(f"  {st.n_nodes} nodes ({st.n_leaves} leaves, {st.levels} levels): grad err {err:.2e} "
         f"rep err {rep_err:.2e} Z {Z.value:.6e}/{Ze.value:.6e}")
    assert st.n_nodes > 10_000  # a tree of the benchmark's size (tens of thousands of nodes)
    assert err <= 3e-2
    assert rep_err <= 3e-2
    assert abs(Z.value / Ze.value - 1) <= 3e-2


def test_overlapped_attractive_pass_matches_fused(backend, problem):
    """On rebuild iterations the attractive sums run on a second stream next to the
    rebuild and the product, and a separate kernel applies the update with the fused
    kernel's order of centring sums: the result equals the fused single-stream pass bit
    for bit, whether the split pass is queued beside the build's front or after it, and
    whatever its grid (one workgroup per logical block, the default 7/8 of the CUs, or 16
    workgroups walking the 376 logical blocks of this problem)."""
    import os
    _say("overlap vs fused")
    _, _, P = problem
    Y0 = init_normal(P.shape[0], 2, 23)
    variants = {
        "fused": {"TSNE_OVERLAP": "0"},
        "beside-front": {"TSNE_OVERLAP": "1", "TSNE_ATTR_AFTER_FRONT": "0",
                         "TSNE_ATTR_SPLIT_GRID": "0"},
        "after-front": {"TSNE_OVERLAP": "1"},
        "after-front-16": {"TSNE_OVERLAP": "1", "TSNE_ATTR_AFTER_FRONT": "1",
                           "TSNE_ATTR_SPLIT_GRID": "16"},
    }
    keys = ("TSNE_OVERLAP", "TSNE_ATTR_AFTER_FRONT", "TSNE_ATTR_SPLIT_GRID")
    out = {}
    for name, env in variants.items():
        old = {k: os.environ.get(k) for k in keys}
        for k in keys:
            os.environ.pop(k, None)
        os.environ.update(env)
        try:
            p = backend.create_default_params()
            p.verbose = 0
            p.n_iter = 12  # rebuilds at iterations 1..11 (every iteration below 20)
            p.h2_tol = 1e-2
            p.h2_rebuild_freq = 3
            p.learning_rate = max(P.shape[0] / 12.0, 200.0)
            ws = backend.init_workspace(P, 2, p, init_Y=Y0.T.copy())
        finally:
            for k, v in old.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v
        try:
            assert backend.lib.tsne_run(ws) == 0
            Y = np.zeros((2, P.shape[0]))
            backend.lib.tsne_get_embedding(ws, Y)
            out[name] = (Y, backend.lib.compute_kl_divergence(ws))
        finally:
            backend.lib.tsne_destroy(ws)
    ref = out["fused"]
    for name, (Y, kl) in out.items():
        rel = np.abs(ref[0] - Y).max() / np.abs(ref[0]).max()
        _say(f"  {name}: rel diff {rel:.2e}, KL {ref[1]:.8f} / {kl:.8f}")
        assert np.array_equal(ref[0], Y) and ref[1] == kl, name


def test_oversized_deepest_leaf_3d(backend, problem):
    """A 3D embedding whose extent (outliers at 1e4) leaves the deepest octree level
    with boxes of ~0.03 while 300 points sit within 0.01 of each other: that leaf holds
    more points than one chunk of the leaf kernel.  The product must still use the H2
    structure (no direct-sum fallback) and meet the 3D bound against FP64."""
    _say("oversized deepest leaf 3d")
    _, _, P = problem
    n = P.shape[0]
    rng = np.random.default_rng(8)
    Y0 = _embedding(n, 3, 3.0, seed=8).T.copy()
    Y0[:300] = 5.0 + rng.uniform(0.0, 0.01, size=(300, 3))
    Y0[300:304] = np.array([[1e4, 0, 0], [-1e4, 0, 0], [0, 1e4, 0], [0, 0, -1e4]])
    Y0 = Y0.T.copy()
    ws = _ws(backend, P, Y0, 3, 3e-2)
    try:
        g, Z = _grad(backend, ws, 3, n, 1.0, 0)
        st = H2Stats()
        assert backend.lib.tsne_h2_stats(ws, ctypes.byref(st)) == 0
        rep64, Z64, _, _ = oracle.repulsive_exact(Y0)
        g64 = oracle.attractive(P, Y0, 1.0) - rep64
        err = np.linalg.norm(g - g64) / np.linalg.norm(g64)
        _say(f"  levels {st.levels} direct fallbacks {st.direct_fallbacks} grad err {err:.2e} "
             f"Z {Z:.6e}/{Z64:.6e}")
        assert st.direct_fallbacks == 0
        assert st.levels >= 20
        assert err <= 3e-2
    finally:
        backend.lib.tsne_destroy(ws)

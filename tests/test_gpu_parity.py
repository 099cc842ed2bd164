"""Parity of the HIP path with the oracle (GPU).  Every call goes through the C ABI
of libtsne_h2pack.so; the oracle (oracle/tsne_oracle.c, FP64) is the checker.

Tolerances (stated per test): perplexity rows 1e-6 relative (FP64 on both sides);
CSR/neighbour indexing bit-exact; FP32 direct repulsion and attractive force 1e-4
relative to the FP64 gradient; FP64 check path 1e-10.
"""
import ctypes
import sys

import numpy as np
import pytest
from scipy.sparse import csr_matrix

import oracle
from datasets import gaussian_mixture, init_normal, joint_P_oracle, knn_graph

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def small_problem():
    X, labels = gaussian_mixture(1500, dim=20, n_clusters=6, seed=0)
    G = knn_graph(X, 45)
    P = joint_P_oracle(G, 15.0)
    return X, labels, G, P


def _workspace(backend, P, Y0, dim=2, **kw):
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = kw.pop("n_iter", 1000)
    p.h2_tol = kw.pop("h2_tol", 1e-2)
    p.h2_rebuild_freq = kw.pop("freq", 3)
    p.h2_rebuild_freq_early = kw.pop("freq_early", 10)
    p.learning_rate = kw.pop("lr", max(P.shape[0] / 12.0, 200.0))
    p.h2_leaf_size = kw.pop("leaf", 100)
    return backend.init_workspace(P, dim, p, init_Y=Y0.T.copy()), p


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


# ---------------------------------------------------------------- perplexity

@pytest.mark.parametrize("scale", [1e-3, 1.0, 1e3, 1e6])
def test_perplexity_kernel_matches_oracle(backend, scale):
    _say(f"perplexity scale={scale}")
    d = np.load(__import__("os").path.join(__import__("os").path.dirname(__file__),
                                           "golden", "golden_perplexity.npz"))["dist"] * scale
    n, k = d.shape
    indptr = np.arange(0, (n + 1) * k, k, dtype=np.int32)
    indices = np.tile(np.arange(k, dtype=np.int32), n)
    P = backend.conditional_P(np.ascontiguousarray(d.ravel()), indices, indptr, 30.0)
    p_or, _ = oracle.perplexity_rows(d.ravel(), indptr, 30.0)
    np.testing.assert_allclose(P.data, p_or, rtol=1e-6, atol=0)
    assert np.array_equal(P.indices, indices) and np.array_equal(P.indptr, indptr)


def test_perplexity_kernel_ragged_rows(backend, small_problem):
    _, _, G, _ = small_problem
    data, indices, indptr = backend.prepare_knn_csr(G)
    # make rows ragged: drop a varying number of trailing neighbours
    rng = np.random.default_rng(0)
    keep = np.ones(data.size, bool)
    for i in range(G.shape[0]):
        drop = rng.integers(0, 10)
        if drop:
            keep[indptr[i + 1] - drop:indptr[i + 1]] = False
    rows = np.repeat(np.arange(G.shape[0]), np.diff(indptr))
    data, indices = data[keep], indices[keep]
    indptr = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=G.shape[0]))]).astype(np.int32)
    P = backend.conditional_P(np.ascontiguousarray(data), np.ascontiguousarray(indices), indptr, 15.0)
    p_or, _ = oracle.perplexity_rows(data, indptr, 15.0)
    np.testing.assert_allclose(P.data, p_or, rtol=1e-6)


def test_joint_P_matches_oracle_pipeline(backend, small_problem):
    """Symmetrised P from the GPU bisection vs the oracle's: same sparsity pattern
    (bit-exact indexing), values to 1e-6."""
    _, _, G, P_or = small_problem
    P = backend.joint_P(G, 15.0)
    assert np.array_equal(P.indptr, P_or.indptr) and np.array_equal(P.indices, P_or.indices)
    np.testing.assert_allclose(P.data, P_or.data, rtol=1e-6)


# ---------------------------------------------------------------- gradient

@pytest.mark.parametrize("dim", [2, 3])
def test_direct_gradient_matches_oracle(backend, small_problem, dim):
    """N < 2000: the FP32 direct repulsion + FP32 attractive force vs the FP64 oracle."""
    _say(f"direct gradient dim={dim}")
    _, _, _, P = small_problem
    rng = np.random.default_rng(1)
    Y0 = rng.normal(scale=5.0, size=(dim, P.shape[0]))
    ws, _ = _workspace(backend, P, Y0, dim)
    try:
        for exag in (12.0, 1.0):
            g = np.zeros((dim, P.shape[0]))
            Z = ctypes.c_double()
            assert backend.lib.tsne_compute_gradient(ws, exag, 0, g, ctypes.byref(Z)) == 0
            rep, Zo, _, _ = oracle.repulsive_exact(Y0)
            ref = oracle.attractive(P, Y0, exag) - rep
            assert abs(Z.value - Zo) / Zo < 1e-5
            assert _rel(g, ref) < 1e-4
            # FP64 check path
            assert backend.lib.tsne_compute_gradient(ws, exag, 1, g, ctypes.byref(Z)) == 0
            assert abs(Z.value - Zo) / Zo < 1e-11
            assert _rel(g, ref) < 1e-10
    finally:
        backend.lib.tsne_destroy(ws)


def test_exact_repulsive_rows(backend, small_problem):
    _, _, _, P = small_problem
    rng = np.random.default_rng(2)
    Y0 = rng.normal(scale=3.0, size=(2, P.shape[0]))
    ws, _ = _workspace(backend, P, Y0, 2)
    try:
        rows = np.array([0, 5, 1499, 700], dtype=np.int32)
        out = np.zeros((rows.size, 2))
        Z = ctypes.c_double()
        assert backend.lib.tsne_exact_repulsive_rows(ws, rows, rows.size, out, ctypes.byref(Z)) == 0
        rep, Zo, _, _ = oracle.repulsive_exact(Y0)
        np.testing.assert_allclose(out, rep[:, rows].T, rtol=1e-10, atol=1e-14)
        assert abs(Z.value - Zo) < 1e-11 * Zo
    finally:
        backend.lib.tsne_destroy(ws)


# ---------------------------------------------------------------- optimisation loop

def _run_gpu(backend, P, Y0, n_iter, dim=2, **kw):
    ws, p = _workspace(backend, P, Y0, dim, n_iter=n_iter, **kw)
    try:
        assert backend.lib.tsne_run(ws) == 0
        Y = np.zeros((dim, P.shape[0]))
        backend.lib.tsne_get_embedding(ws, Y)
        kl = backend.lib.compute_kl_divergence(ws)
    finally:
        backend.lib.tsne_destroy(ws)
    return Y, kl


def test_first_iterations_track_oracle(backend, small_problem):
    """10 iterations of the fused attractive + update loop against the FP64 oracle.
    t-SNE amplifies rounding: the gains rule switches discretely on the sign of
    uY * grad, so a rounding difference near zero moves a point by O(step), and the
    centring spreads that over all points.  The bound is therefore on the typical point
    (median distance 1e-3 of the embedding's extent, 99th percentile 1e-2), which an O(1)
    error in the force, the update or the centring fails."""
    _say("10-iteration trajectory")
    _, _, _, P = small_problem
    Y0 = init_normal(P.shape[0], 2, 42) * 1e2
    Y, _ = _run_gpu(backend, P, Y0, 10)
    Yo, _ = oracle.run(P, Y0, n_iter=10)
    d = np.linalg.norm(Y - Yo, axis=0) / np.abs(Yo).max()
    _say(f"  per-point distance: median {np.median(d):.2e}, 99th percentile "
         f"{np.percentile(d, 99):.2e}, max {d.max():.2e}")
    assert np.median(d) < 1e-3
    assert np.percentile(d, 99) < 1e-2


def _seeded_init(backend, P, seed):
    p = backend.create_default_params()
    p.verbose = 0
    p.random_seed = seed
    ws = backend.init_workspace(P, 2, p, init_Y=None)
    try:
        Y = np.zeros((2, P.shape[0]))
        backend.lib.tsne_get_embedding(ws, Y)
    finally:
        backend.lib.tsne_destroy(ws)
    return Y


def test_random_init_is_seeded_normal(backend, small_problem):
    """init_method 0 (the reference's srand/rand Box-Muller, src/tsne_utils.c:1129-1148).
    Parity unpinned: no reference vector pins glibc's rand() sequence, and every other
    test and the benchmark pass an explicit init_Y.  Held to what the method fixes:
    N(0, 1e-4^2) moments; consecutive storage-order pairs are Box-Muller pairs, so
    (a^2 + b^2) / (2e-8) is Exp(1) and atan2(b, a) is uniform (Kolmogorov-Smirnov); the
    draw is a function of the seed."""
    from scipy import stats
    _, _, _, P = small_problem
    Y = _seeded_init(backend, P, 77)
    v = Y.ravel()  # storage order of the dim x n array
    assert abs(v.std() / 1e-4 - 1) < 0.05 and abs(v.mean()) < 1e-5
    a, b = v[0::2], v[1::2]
    assert stats.kstest((a * a + b * b) / 2e-8, "expon").pvalue > 1e-3
    assert stats.kstest(np.arctan2(b, a), stats.uniform(-np.pi, 2 * np.pi).cdf).pvalue > 1e-3
    assert np.array_equal(Y, _seeded_init(backend, P, 77))
    assert not np.array_equal(Y, _seeded_init(backend, P, 78))


def test_full_run_kl_within_one_percent(backend, small_problem):
    """1000 iterations (N=1500: direct repulsion every iteration) end within 1% of the
    oracle's final KL from the same init."""
    _say("full run N=1500")
    _, labels, _, P = small_problem
    Y0 = init_normal(P.shape[0], 2, 3)
    Y, kl = _run_gpu(backend, P, Y0, 1000)
    _, kl_o = oracle.run(P, Y0, n_iter=1000)
    assert abs(kl - kl_o) / kl_o < 0.01, (kl, kl_o)
    assert np.isfinite(Y).all()

"""Pins of the CPU oracle (oracle/tsne_oracle.c) against independent implementations,
the reference's own test fixtures and closed forms.  CPU only."""
import os

import numpy as np
import pytest
from scipy.sparse import csr_matrix

import oracle

GOLD = __import__("os").path.join(__import__("os").path.dirname(__file__), "golden")


def _load(name):
    return np.load(f"{GOLD}/{name}")


# ---------------------------------------------------------------- perplexity (A1)

def _row_perplexity(P):
    H = -np.sum(np.where(P > 0, P * np.log2(np.maximum(P, 1e-300)), 0.0), axis=1)
    return 2.0 ** H


def _calibrate(dist2d, perplexity=30.0):
    n, k = dist2d.shape
    indptr = np.arange(0, (n + 1) * k, k, dtype=np.int32)
    p, steps = oracle.perplexity_rows(dist2d.ravel(), indptr, perplexity)
    return p.reshape(n, k), steps


@pytest.mark.parametrize("scale", [1e-3, 1.0, 1e3, 1e6])
def test_rows_hit_requested_perplexity(scale):
    """Reference tests/test_perplexity_calibration.py:48-52 on its own fixture."""
    d = _load("golden_perplexity.npz")["dist"]
    P, _ = _calibrate(d * scale)
    np.testing.assert_allclose(P.sum(axis=1), 1.0, rtol=1e-10)
    np.testing.assert_allclose(_row_perplexity(P), 30.0, rtol=1e-3)


@pytest.mark.parametrize("scale", [1e-3, 1e3, 1e6])
def test_calibration_is_scale_invariant(scale):
    """Reference tests/test_perplexity_calibration.py:55-59."""
    d = _load("golden_perplexity.npz")["dist"]
    P_ref, _ = _calibrate(d)
    P_s, _ = _calibrate(d * scale)
    np.testing.assert_allclose(P_s, P_ref, rtol=1e-3, atol=1e-9)


def test_beta_matches_entropy_root():
    """The stopping rule |H - ln 30| < 1e-5 nats: P equals the exact-root P to the
    accuracy that tolerance allows (dP/dH bounded on this fixture)."""
    g = _load("golden_perplexity.npz")
    d, beta = g["dist"], g["beta"]
    P, _ = _calibrate(d)
    sq = d ** 2 - (d ** 2).min(axis=1, keepdims=True)
    P_exact = np.exp(-beta[:, None] * sq)
    P_exact /= P_exact.sum(axis=1, keepdims=True)
    np.testing.assert_allclose(P, P_exact, rtol=2e-4, atol=1e-8)


def test_matches_sklearn_binary_search():
    """sklearn's bisection (different start and step rule) agrees to its own tolerance."""
    g = _load("golden_perplexity.npz")
    P, _ = _calibrate(g["dist"])
    np.testing.assert_allclose(P, g["P_sklearn"], rtol=5e-3, atol=1e-6)


def test_uniform_row_is_uniform():
    d = np.full((3, 40), 2.5)
    P, steps = _calibrate(d, perplexity=10.0)
    np.testing.assert_allclose(P, 1.0 / 40, rtol=1e-14)


def test_ln_not_log2():
    """A plausible mistake (log2 target) would give perplexity 30*ln2-ish; the row
    entropy in nats must equal ln(perp)."""
    d = _load("golden_perplexity.npz")["dist"][:20]
    P, _ = _calibrate(d, perplexity=12.0)
    H = -np.sum(P * np.log(P), axis=1)
    np.testing.assert_allclose(H, np.log(12.0), atol=1.1e-5)


def test_ragged_and_empty_rows():
    dist = np.array([1.0, 2.0, 3.0, 0.5, 0.7], dtype=np.float64)
    indptr = np.array([0, 3, 3, 5], dtype=np.int32)
    p, steps = oracle.perplexity_rows(dist, indptr, 1.5)
    np.testing.assert_allclose([p[:3].sum(), p[3:].sum()], 1.0, rtol=1e-12)
    assert steps[1] == 0


# ---------------------------------------------------------------- gradient (A7, A12, A14)

@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("exag", [1, 12])
def test_gradient_and_kl_match_sklearn_exact(dim, exag):
    """attractive - repulsive == sklearn's exact t-SNE gradient; KL == sklearn's KL of
    the exaggerated P."""
    g = _load("golden_gradient.npz")
    Y = g[f"Y{dim}"].T.copy()
    P = csr_matrix(g[f"P{dim}"])
    att = oracle.attractive(P, Y, float(exag))
    rep, Z, s, v = oracle.repulsive_exact(Y)
    grad = (att - rep).T
    ref = g[f"grad{dim}_x{exag}"]
    assert np.abs(grad - ref).max() <= 1e-12 * np.abs(ref).max()
    kl = oracle.kl(P, Y, Z, float(exag))
    assert abs(kl - float(g[f"kl{dim}_x{exag}"])) <= 1e-9 * max(1.0, abs(kl))


def test_repulsive_closed_form_two_points():
    """Two points at distance r: K = (1+r^2)^-2, Z = 2/(1+r^2), force (4/Z) K (y_i - y_j)."""
    Y = np.array([[0.0, 3.0], [0.0, 4.0]])  # (dim, n): points (0,0) and (3,4), r = 5
    rep, Z, s, v = oracle.repulsive_exact(Y)
    assert abs(Z - 2.0 / 26.0) < 1e-15
    K = 1.0 / 26.0 ** 2
    np.testing.assert_allclose(rep[:, 0], 4.0 / Z * K * np.array([-3.0, -4.0]), rtol=1e-14)
    np.testing.assert_allclose(rep[:, 1], -rep[:, 0], rtol=1e-14)


def test_repulsive_translation_invariant_and_balanced():
    rng = np.random.default_rng(3)
    Y = rng.normal(size=(2, 300)) * 5
    r1, Z1, _, _ = oracle.repulsive_exact(Y)
    r2, Z2, _, _ = oracle.repulsive_exact(Y + np.array([[40.0], [-25.0]]))
    assert abs(Z1 - Z2) < 1e-9 * Z1
    np.testing.assert_allclose(r1, r2, atol=1e-9 * np.abs(r1).max())
    assert np.abs(r1.sum(axis=1)).max() < 1e-10 * np.abs(r1).max()


def test_repulsive_rows_matches_full():
    rng = np.random.default_rng(4)
    Y = rng.normal(size=(3, 200)) * 4
    rep, Z, _, _ = oracle.repulsive_exact(Y)
    rows = np.array([0, 17, 199, 5], dtype=np.int32)
    out = oracle.repulsive_rows(Y, Z, rows)
    np.testing.assert_allclose(out, rep[:, rows].T, rtol=1e-10, atol=1e-14)


# ---------------------------------------------------------------- update (A15), init (A4)

def test_update_matches_sklearn_gradient_descent():
    g = _load("golden_update.npz")
    p0, g0, p2 = g["p0"], g["g0"], g["p2"]
    n = p0.size // 2
    Y = p0.reshape(n, 2).T.copy()
    uY = np.zeros_like(Y)
    gains = np.ones_like(Y)
    for _ in range(2):
        grad = g0.reshape(n, 2).T.copy()
        oracle.update(Y, uY, gains, grad, float(g["momentum"]), float(g["lr"]))
    ref = p2.reshape(n, 2).T
    ref = ref - ref.mean(axis=1, keepdims=True)
    np.testing.assert_allclose(Y, ref, rtol=1e-12, atol=1e-10)


def test_update_centres():
    rng = np.random.default_rng(5)
    Y = rng.normal(size=(2, 100)) + 3.0
    uY = np.zeros_like(Y); gains = np.ones_like(Y); grad = rng.normal(size=Y.shape)
    oracle.update(Y, uY, gains, grad, 0.5, 10.0)
    assert np.abs(Y.mean(axis=1)).max() < 1e-12


# ---------------------------------------------------------------- schedule (A6)

def test_schedule_counts_292_rebuilds():
    """1000 iterations, rebuild every 3 (10 in early exaggeration): 292 rebuilds plus the
    initial build, i.e. 293 products (counted from the reference's profile CSV)."""
    n = sum(oracle.is_rebuild_iter(it, 3, 10, 250, True) for it in range(1000))
    assert n == 292


def test_schedule_exact_method_never_rebuilds_late():
    assert not any(oracle.is_rebuild_iter(it, 999999, 999999, 250, True) for it in range(20, 1000))


def test_run_small_decreases_kl():
    from datasets import gaussian_mixture, init_normal, joint_P_oracle, knn_graph
    X, _ = gaussian_mixture(400, dim=10, n_clusters=4, seed=1)
    P = joint_P_oracle(knn_graph(X, 30), 10.0)
    Y0 = init_normal(400, 2, 1)
    _, kl_short = oracle.run(P, Y0, n_iter=300)
    _, kl_long = oracle.run(P, Y0, n_iter=600)
    assert kl_long < kl_short


# ---------------------------------------------------------------- exact k-NN graph

def _knn_golden():
    return np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_knn.npz"))


def test_knn_matches_sklearn_brute_fixture():
    """oracle_knn_exact == the reference's ExactKNN call (sklearn brute kneighbors_graph,
    float32 data) on a 50-D mixture: same neighbours in the same order, same float32
    distances."""
    z = _knn_golden()
    idx, dist = oracle.knn_exact(z["mix_X"], z["mix_idx"].shape[1])
    assert np.array_equal(idx, z["mix_idx"])
    assert np.array_equal(dist, z["mix_dist"])


def test_knn_duplicates_match_sklearn_up_to_tie_order():
    """With duplicate points sklearn orders equal distances arbitrarily; the oracle orders
    them by index.  Distances agree exactly, every listed neighbour is at its listed
    distance, no row lists itself, and rows whose boundary is not tied hold the same set."""
    z = _knn_golden()
    X, gi, gd = z["dup_X"], z["dup_idx"], z["dup_dist"]
    n, k = gi.shape
    idx, dist = oracle.knn_exact(X, k)
    assert np.array_equal(dist, gd)
    true = np.sqrt(((X[:, None, :] - X[idx]) ** 2).sum(-1)).astype(np.float32)
    assert np.allclose(true, dist, rtol=1e-6, atol=1e-7)
    assert not (idx == np.arange(n)[:, None]).any()
    full = np.sort(np.sqrt(((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)), axis=1)
    untied = full[:, k] != full[:, k + 1]  # column 0 is the point itself (distance 0)
    for i in np.nonzero(untied)[0]:
        assert set(idx[i]) == set(gi[i])


def test_knn_brute_force_tiny():
    """Pure-Python brute force on 13 points (ties by index, self removed)."""
    rng = np.random.default_rng(3)
    X = rng.integers(0, 4, size=(13, 3)).astype(np.float64)  # many exact ties
    k = 5
    idx, dist = oracle.knn_exact(X, k)
    for q in range(13):
        cand = sorted((sum((X[q, f] - X[c, f]) ** 2 for f in range(3)), c) for c in range(13))
        cand = cand[:k + 1]
        drop = next((r for r, (_, c) in enumerate(cand) if c == q), 0)
        want = [c for r, (_, c) in enumerate(cand) if r != drop]
        assert idx[q].tolist() == want
        assert np.allclose(dist[q], [np.sqrt(d) for r, (d, _) in enumerate(cand) if r != drop])


# ---------------------------------------------------------------- quality score

def _quality_problem(n=600, seed=2):
    from datasets import gaussian_mixture, joint_P_oracle, knn_graph
    X, _ = gaussian_mixture(n, dim=10, n_clusters=4, seed=seed)
    P = joint_P_oracle(knn_graph(X, 30), 10.0)
    hd, _ = oracle.knn_exact(X, 10)
    Y = np.random.default_rng(seed).normal(scale=5.0, size=(n, 2))
    return X, P, hd, Y


def test_quality_with_every_anchor_is_the_exact_kl():
    """Anchors = every row: Z is the exact sum_{i != j} w_ij (oracle_repulsive_exact) and
    the KL is the exact KL(P || Q) (oracle_kl, pinned to sklearn), since sum p = 1."""
    X, P, hd, Y = _quality_problem()
    n = Y.shape[0]
    q = oracle.quality_score(Y, hd, np.arange(50), P, np.arange(n))
    _, Z, _, _ = oracle.repulsive_exact(Y.T.copy())
    assert abs(q["Z"] / Z - 1) < 1e-12
    assert abs(q["kl"] - oracle.kl(P, Y.T.copy(), Z)) < 1e-10 * max(1.0, abs(q["kl"]))


def test_quality_preservation_limits():
    """The data itself as the embedding (2-D data) preserves every neighbour; a random
    relabelling preserves about k/n."""
    rng = np.random.default_rng(4)
    X = rng.normal(size=(500, 2))
    hd, _ = oracle.knn_exact(X, 10)
    P = csr_matrix(np.full((500, 500), 1.0 / 500 ** 2))
    q = oracle.quality_score(X, hd, np.arange(500), P, np.arange(10))
    assert q["knn_preservation"] == 1.0
    q2 = oracle.quality_score(X[rng.permutation(500)], hd, np.arange(500), P, np.arange(10))
    assert q2["knn_preservation"] < 0.1


def test_quality_kdtree_neighbours_are_exact():
    """The KDTree query of the scorer equals the brute-force k+1 nearest (no ties)."""
    from sklearn.neighbors import KDTree
    Y = np.random.default_rng(9).normal(size=(400, 2))
    ld = KDTree(Y).query(Y, k=11, return_distance=False)
    idx, _ = oracle.knn_exact(Y, 10)
    assert np.array_equal(ld[:, 0], np.arange(400))
    for i in range(400):
        assert set(ld[i, 1:]) == set(idx[i])

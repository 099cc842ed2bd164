"""Exact k-NN graph on the device (tsne_knn_exact) against the oracle's FP64 brute force
(oracle_knn_exact, pinned to the reference's sklearn call in test_oracle.py) and the
sklearn fixture, through the C ABI.

Bar: bit-exact.  The device ranks its survivors with the same FP64 squared distances
(differences, fma in feature order) and the same order (distance, then index) as the
oracle, so indices and float32 distances must be identical -- including ties between
duplicate points, and including rows that the BF16 screening cannot settle (ill-conditioned
data), which take the FP64 fallback.
"""
import os
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mixture(n, dim, n_clusters, seed, spread=5.0):
    rng = np.random.default_rng(seed)
    centers = rng.normal(scale=spread, size=(n_clusters, dim))
    lab = rng.integers(0, n_clusters, n)
    std = rng.uniform(0.5, 2.0, n_clusters)
    return centers[lab] + rng.normal(size=(n, dim)) * std[lab, None]


def _check(backend, X, k):
    idx, dist, st = backend.knn_exact(X, k)
    ridx, rdist = oracle.knn_exact(X, k)
    bad = np.nonzero((idx != ridx).any(1) | (dist != rdist).any(1))[0]
    _say(f"  n={X.shape[0]} D={X.shape[1]} k={k}: survivors {st.survivors}, "
         f"uncertain rows {st.uncertain_rows}, overflow rows {st.overflow_rows}, "
         f"rows differing from the oracle {bad.size}")
    assert bad.size == 0, (bad[:5], idx[bad[:1]], ridx[bad[:1]])
    return st


def test_knn_matches_sklearn_fixture(backend):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_knn.npz"))
    _say("knn: sklearn fixture (n=1000, D=50, k=60)")
    idx, dist, _ = backend.knn_exact(z["mix_X"], 60)
    assert np.array_equal(idx, z["mix_idx"])
    assert np.array_equal(dist, z["mix_dist"])


@pytest.mark.parametrize("n,dim,k", [(777, 3, 10), (2049, 50, 90), (3000, 100, 150),
                                     (1500, 300, 30), (6000, 50, 90), (300, 64, 299),
                                     (10, 1, 3), (65, 2, 64), (1100, 1030, 20)])
def test_knn_matches_oracle(backend, n, dim, k):
    _say(f"knn: mixture n={n} D={dim} k={k}")
    X = _mixture(n, dim, 6, seed=n + dim)
    _check(backend, X, k)


def test_knn_duplicates_ordered_by_index(backend):
    _say("knn: duplicate points")
    rng = np.random.default_rng(5)
    X = np.vstack([np.repeat(rng.normal(size=(40, 8)), 10, axis=0), rng.normal(size=(77, 8))])
    _check(backend, X, 6)   # more duplicates than k + 1: the nearest entry is dropped
    _check(backend, X, 25)


def test_knn_ill_conditioned_rows_take_the_fp64_fallback(backend):
    """Two clusters 2000 apart with spread 1e-3: after centring |x|^2 ~ 1e6 while
    neighbour d^2 ~ 1e-5, far below the screening's error bound, so rows are re-ranked in
    FP64 (collected within their refined bound, or fully sorted when too many points lie
    inside it)."""
    _say("knn: ill-conditioned data")
    rng = np.random.default_rng(11)
    c = rng.normal(size=(2, 10)) * 1000.0
    lab = rng.integers(0, 2, 3000)
    X = c[lab] + rng.normal(size=(3000, 10)) * 1e-3
    st = _check(backend, X, 15)
    assert st.uncertain_rows > 0


def test_knn_overflowing_fallback_rows(backend):
    _say("knn: fallback rows beyond the collection capacity")
    rng = np.random.default_rng(12)
    c = rng.normal(size=(2, 6)) * 1000.0
    lab = np.repeat([0, 1], 10000)
    X = c[lab] + rng.normal(size=(20000, 6)) * 1e-3
    st = _check(backend, X, 40)
    assert st.overflow_rows > 0


def test_knn_provider_matches_exact_provider(backend):
    """knn_method='gpu' returns the CSR graph of ExactKNN (float32 distances, rows in
    distance order)."""
    from tsne_h2pack._knn import get_knn_method
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_knn.npz"))
    _say("knn: provider")
    G = get_knn_method("gpu").compute(z["mix_X"], 60)
    assert G.shape == (1000, 1000) and G.nnz == 60000 and G.data.dtype == np.float32
    assert np.array_equal(G.indices.reshape(1000, 60), z["mix_idx"])
    assert np.array_equal(G.data.reshape(1000, 60), z["mix_dist"])


def test_knn_rejects_invalid_input(backend):
    _say("knn: invalid input")
    X = np.zeros((10, 3))
    with pytest.raises(RuntimeError):
        backend.knn_exact(X, 10)          # k >= n
    X[3, 1] = np.nan
    with pytest.raises(RuntimeError):
        backend.knn_exact(X, 3)           # not finite

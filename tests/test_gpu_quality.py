"""Quality score on the device (tsne_quality_score) against the oracle's restatement of
the reference's scorer (benchmarks/quality.py:56-81), through the C ABI.

Bar: kNN-preservation identical (an integer count over the queries), Z and KL within 1e-12
relative (FP64 sums in another order), span identical.
"""
import sys

import numpy as np
import pytest

import oracle
from datasets import gaussian_mixture, joint_P_oracle, knn_graph

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n,dim,k", [(3000, 2, 10), (5000, 3, 10), (2500, 2, 5)])
def test_quality_matches_oracle(backend, n, dim, k):
    _say(f"quality: n={n} dim={dim} k={k}")
    X, _ = gaussian_mixture(n, dim=20, n_clusters=6, seed=n)
    P = joint_P_oracle(knn_graph(X, 45), 15.0)
    hd, _ = oracle.knn_exact(X, 12)
    rng = np.random.default_rng(n + dim)
    Y = rng.normal(scale=8.0, size=(n, dim)) + np.repeat(rng.normal(scale=30, size=(6, dim)),
                                                         -(-n // 6), axis=0)[:n]
    r = np.random.default_rng(0)
    queries = r.choice(n, size=min(1000, n), replace=False)
    anchors = r.choice(n, size=min(300, n), replace=False)
    want = oracle.quality_score(Y, hd, queries, P, anchors, k=k)
    got = backend.quality_score(Y, hd, queries, P, anchors, k=k)
    _say(f"  device {got}\n  oracle {want}")
    assert got["knn_preservation"] == want["knn_preservation"]
    assert abs(got["Z"] / want["Z"] - 1) < 1e-12
    assert abs(got["kl"] - want["kl"]) < 1e-12 * max(1.0, abs(want["kl"]))
    assert got["span"] == want["span"]


def test_quality_python_mirror_draws_like_the_reference(backend):
    """tsne_h2pack.quality.score draws queries, then anchors, from default_rng(0) as the
    reference's score() does."""
    from tsne_h2pack.quality import score
    _say("quality: python mirror")
    n = 2000
    X, _ = gaussian_mixture(n, dim=10, n_clusters=3, seed=8)
    P = joint_P_oracle(knn_graph(X, 30), 10.0)
    hd, _ = oracle.knn_exact(X, 10)
    Y = np.random.default_rng(1).normal(size=(n, 2))
    got = score(Y, hd, P, k=10, n_eval=700, n_anchors=150)
    r = np.random.default_rng(0)
    q = r.choice(n, size=700, replace=False)
    a = r.choice(n, size=150, replace=False)
    want = oracle.quality_score(Y, hd, q, P, a, k=10)
    assert got["knn_preservation"] == want["knn_preservation"]
    assert abs(got["kl"] - want["kl"]) < 1e-12 * max(1.0, abs(want["kl"]))


def test_quality_rejects_invalid_input(backend):
    _say("quality: invalid input")
    n = 100
    Y = np.random.default_rng(0).normal(size=(n, 2))
    hd = np.zeros((n, 10), dtype=np.int32)
    P = joint_P_oracle(knn_graph(np.random.default_rng(1).normal(size=(n, 5)), 10), 5.0)
    with pytest.raises(RuntimeError):
        backend.quality_score(Y, hd, np.array([n]), P, np.arange(5))     # query out of range
    with pytest.raises(RuntimeError):
        backend.quality_score(Y, hd, np.arange(5), P, np.arange(5), k=16)  # k too large

"""The oracle's CPU H2 product (oracle/tsne_oracle_h2.c), the restatement of the reference's
H2 repulsion that bench.py times as the CPU baseline.  Pinned against the exact O(N^2) sum
(oracle_repulsive_exact, itself pinned to sklearn): the repulsive force and Z within h2_tol
in 2D and 3D on clustered and on freshly initialised embeddings, convergence to the exact
sum as the tolerance shrinks (so a dropped term, a wrong sign or index in the tree, lists,
ID or passes cannot hide), translation invariance, and a whole run on the reference's
schedule whose final KL is within 1% of the exact-repulsion run."""
import numpy as np
import pytest

import oracle
from datasets import gaussian_mixture, init_normal, joint_P_oracle, knn_graph


def _blobs(n, dim, seed, clusters=12, spread=40.0, width=3.0):
    """A clustered embedding like a late t-SNE state: Gaussian blobs, (dim, n)."""
    rng = np.random.default_rng(seed)
    c = rng.normal(scale=spread, size=(clusters, dim))
    lab = rng.integers(0, clusters, n)
    return (c[lab] + rng.normal(scale=width, size=(n, dim))).T.copy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("dim,n,tol", [(2, 6000, 1e-2), (2, 6000, 1e-3), (3, 3000, 3e-2)])
def test_h2_repulsion_within_tolerance(dim, n, tol):
    for Y in (_blobs(n, dim, 1), init_normal(n, dim, 2) * 3e3):
        rep, Z, st = oracle.repulsive_h2(Y, h2_tol=tol)
        repx, Zx, _, _ = oracle.repulsive_exact(Y)
        assert _rel(rep, repx) <= tol, (_rel(rep, repx), st)
        assert abs(Z - Zx) / Zx <= tol
        assert st["n_far"] > 0 and st["n_leaves"] > 1  # the H2 path, not a direct sum


def test_h2_converges_to_the_exact_sum():
    """Error falls with the tolerance and reaches 1e-6 at h2_tol = 1e-6."""
    Y = _blobs(4000, 2, 3)
    repx, Zx, _, _ = oracle.repulsive_exact(Y)
    errs = [_rel(oracle.repulsive_h2(Y, h2_tol=t)[0], repx) for t in (1e-2, 1e-4, 1e-6)]
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-6
    rep, Z, _ = oracle.repulsive_h2(Y, h2_tol=1e-6)
    assert abs(Z - Zx) / Zx < 1e-7


def test_h2_translation_invariant():
    Y = _blobs(5000, 2, 4)
    rep0, Z0, _ = oracle.repulsive_h2(Y, h2_tol=1e-2)
    rep1, Z1, _ = oracle.repulsive_h2(Y + np.array([[123.25], [-77.5]]), h2_tol=1e-2)
    assert _rel(rep1, rep0) < 1e-9 and abs(Z1 - Z0) / Z0 < 1e-9


def test_h2_ranks_like_the_reference():
    """The reference reports a maximum H2 rank of 17 in 2D at h2_tol 1e-2
    (tsne_h2pack/manifold.py:32-37); the restatement's nested bases stay in that range."""
    _, _, st = oracle.repulsive_h2(_blobs(20000, 2, 5), h2_tol=1e-2)
    assert 8 <= st["rank_max"] <= 24


def test_h2_run_kl_within_one_percent_of_exact():
    """The reference's schedule (force reuse, rebuild every 3, 10 during exaggeration) with
    the CPU H2 product against the same run with the exact sum."""
    X, _ = gaussian_mixture(2500, dim=16, n_clusters=6, seed=7)
    P = joint_P_oracle(knn_graph(X, 45), 15.0)
    Y0 = init_normal(2500, 2, 8)
    _, kl_exact = oracle.run(P, Y0, n_iter=300)
    _, kl_h2, builds = oracle.run_h2(P, Y0, n_iter=300, h2_tol=1e-2)
    expected = 1 + sum(oracle.is_rebuild_iter(it) for it in range(1, 300))
    assert builds == expected
    assert abs(kl_h2 - kl_exact) / kl_exact < 0.01, (kl_h2, kl_exact)

"""Writes the golden fixtures of tests/golden/ from independent implementations
(sklearn's exact t-SNE routines and scipy root finding), not from the oracle or the
library.  Run from the repository root: python tests/golden/make_golden.py

  golden_gradient.npz   Y, dense joint P, sklearn _kl_divergence gradient and KL
                        (2D and 3D, exaggeration 1 and 12).
  golden_update.npz     one step of sklearn _gradient_descent (gains, momentum).
  golden_perplexity.npz the reference's calibration fixture
                        (default_rng(0).uniform(1, 3, (200, 90)), sorted rows;
                        reference tests/test_perplexity_calibration.py:42-45), the beta
                        that solves H(beta) = ln(30) per row (brentq, 1e-14), and
                        sklearn's _binary_search_perplexity P.
  golden_knn.npz        the reference's exact k-NN graph call (NearestNeighbors(
                        algorithm='brute').kneighbors_graph(mode='distance'), float32
                        data; reference _knn_methods.py:71-119) on a seeded 50-D
                        Gaussian mixture (n=1000, k=60) and on data with duplicate
                        points (n=240, k=6).
"""
import os

import numpy as np
from scipy.optimize import brentq
from scipy.spatial.distance import squareform
from sklearn.manifold import _t_sne
from sklearn.manifold._utils import _binary_search_perplexity

HERE = os.path.dirname(os.path.abspath(__file__))


def gradient_fixture():
    out = {}
    rng = np.random.default_rng(7)
    for dim in (2, 3):
        n = 90
        Y = rng.normal(scale=3.0, size=(n, dim))
        A = rng.random((n, n)) * (rng.random((n, n)) < 0.15)
        A = A + A.T
        np.fill_diagonal(A, 0.0)
        A /= A.sum()
        for exag in (1.0, 12.0):
            Pc = squareform(A * exag, checks=False)
            kl, g = _t_sne._kl_divergence(Y.ravel().copy(), Pc, 1.0, n, dim)
            out[f"Y{dim}"] = Y
            out[f"P{dim}"] = A
            out[f"grad{dim}_x{int(exag)}"] = g.reshape(n, dim)
            out[f"kl{dim}_x{int(exag)}"] = np.array(kl)
    np.savez_compressed(os.path.join(HERE, "golden_gradient.npz"), **out)


def update_fixture():
    rng = np.random.default_rng(11)
    n, dim = 50, 2
    p0 = rng.normal(size=n * dim)
    g0 = rng.normal(size=n * dim)

    def objective(p, *args, **kwargs):
        return 0.0, g0.copy()

    # two steps so that the second sees a nonzero update and exercises the gains rule
    p_after, _, _ = _t_sne._gradient_descent(objective, p0.copy(), 0, 2, n_iter_check=10**9,
                                             momentum=0.5, learning_rate=200.0,
                                             min_gain=0.01, min_grad_norm=0.0)
    np.savez_compressed(os.path.join(HERE, "golden_update.npz"), p0=p0, g0=g0, p2=p_after,
                        momentum=0.5, lr=200.0)


def perplexity_fixture():
    rng = np.random.default_rng(0)
    d = np.sort(rng.uniform(1.0, 3.0, size=(200, 90)), axis=1)
    target = np.log(30.0)
    betas = np.zeros(200)
    for i in range(200):
        sq = d[i] ** 2 - (d[i] ** 2).min()

        def H(b):
            w = np.exp(-b * sq)
            s = w.sum()
            return np.log(s) + b * (sq * w).sum() / s - target
        betas[i] = brentq(H, 1e-8, 1e4, xtol=1e-14, rtol=1e-14)
    P_sk = _binary_search_perplexity((d ** 2).astype(np.float32), 30.0, 0)
    np.savez_compressed(os.path.join(HERE, "golden_perplexity.npz"), dist=d, beta=betas,
                        P_sklearn=P_sk)


def knn_fixture():
    from sklearn.neighbors import NearestNeighbors
    rng = np.random.default_rng(7)
    centers = rng.normal(scale=5.0, size=(8, 50))
    lab = rng.integers(0, 8, 1000)
    X = centers[lab] + rng.normal(size=(1000, 50)) * rng.uniform(0.5, 2.0, 8)[lab, None]
    Xd = np.vstack([np.repeat(rng.normal(size=(50, 5)), 4, axis=0), rng.normal(size=(40, 5))])
    out = {}
    for name, A, k in (("mix", X, 60), ("dup", Xd, 6)):
        G = NearestNeighbors(n_neighbors=k, algorithm='brute').fit(A).kneighbors_graph(
            mode='distance')
        out[f"{name}_X"] = A
        out[f"{name}_idx"] = G.indices.reshape(A.shape[0], k).astype(np.int32)
        out[f"{name}_dist"] = G.data.reshape(A.shape[0], k).astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "golden_knn.npz"), **out)


if __name__ == "__main__":
    knn_fixture()
    gradient_fixture()
    update_fixture()
    perplexity_fixture()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))

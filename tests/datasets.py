"""Synthetic inputs shared by the tests and bench.py: seeded Gaussian mixtures, exact
k-NN graphs and the symmetric P of the reference pipeline.  P is formed with the
oracle's bisection (tests) or the library's (product runs) and scipy symmetrisation,
as the reference does in Python (reference _backend.py:709-718)."""
import numpy as np
from scipy.sparse import csr_matrix


def gaussian_mixture(n, dim=50, n_clusters=10, seed=0, spread=4.0, dirichlet=False,
                     std_range=None, dtype=np.float64):
    rng = np.random.default_rng(seed)
    centers = rng.normal(scale=spread, size=(n_clusters, dim))
    if dirichlet:
        w = rng.dirichlet(np.ones(n_clusters))
    else:
        w = np.full(n_clusters, 1.0 / n_clusters)
    labels = rng.choice(n_clusters, size=n, p=w)
    std = np.ones(n_clusters) if std_range is None else rng.uniform(*std_range, size=n_clusters)
    X = centers[labels] + rng.normal(size=(n, dim)) * std[labels, None]
    return X.astype(dtype), labels


def knn_graph(X, k):
    """Exact k-NN (self excluded) as CSR of float32 L2 distances, columns sorted."""
    from sklearn.neighbors import NearestNeighbors
    nn = NearestNeighbors(n_neighbors=k + 1, algorithm='brute').fit(X)
    dist, idx = nn.kneighbors(X)
    n = X.shape[0]
    rows = np.arange(n)[:, None]
    keep = idx != rows
    keep[~(idx == rows).any(axis=1), -1] = False
    idx = idx[keep].reshape(n, k)
    dist = dist[keep].reshape(n, k)
    G = csr_matrix((dist.ravel().astype(np.float32), idx.ravel(),
                    np.arange(0, (n + 1) * k, k)), shape=(n, n))
    G.sort_indices()
    return G


def joint_P_oracle(G, perplexity):
    """Conditional P by the oracle's bisection, then scipy symmetrisation."""
    import oracle
    data = np.ascontiguousarray(G.data, dtype=np.float64)
    p, _ = oracle.perplexity_rows(data, G.indptr.astype(np.int32), perplexity)
    Pc = csr_matrix((p, G.indices.astype(np.int32), G.indptr.astype(np.int32)), shape=G.shape)
    P = Pc + Pc.T
    P.data /= P.sum()
    return P


def init_normal(n, dim, seed, scale=1e-4):
    """Explicit initial embedding (dim, n): numpy default_rng(seed) N(0, scale^2), the
    distribution of the reference's random init (src/tsne_utils.c:1129-1148), passed to
    the library as init_Y so that no result depends on glibc's rand() sequence."""
    return np.random.default_rng(seed).normal(scale=scale, size=(dim, n))

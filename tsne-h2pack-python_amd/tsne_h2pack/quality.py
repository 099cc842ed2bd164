"""Embedding quality on a common footing, computed on the GPU.

Mirror of the reference's benchmark scorer (benchmarks/quality.py:56-81, score()): every
embedding of one dataset is scored against ONE joint P and one high-dimensional k-NN graph,
with Z estimated the same way for all:

    KL = sum p log p + sum p log(1 + d^2) + log Z,   Z = n * mean_{anchors a} sum_{j != a} w_aj

plus kNN-preservation@k: the share of each query point's k high-dimensional neighbours that
are among its k nearest neighbours in the embedding.  The random query and anchor rows are
drawn exactly as the reference draws them (numpy default_rng(seed): queries first, then
anchors, both without replacement).
"""
import numpy as np

from ._backend import get_backend


def graph_neighbors(knn_graph, k):
    """First k neighbour columns of each row of a k-NN distance CSR (rows in distance order,
    as the k-NN providers return them), self entries removed: an (n, k) int32 array."""
    G = knn_graph
    n = G.shape[0]
    out = np.empty((n, k), dtype=np.int32)
    for i in range(n):
        s, e = G.indptr[i], G.indptr[i + 1]
        cols, d = G.indices[s:e], G.data[s:e]
        order = np.argsort(d, kind="stable")
        cols = cols[order]
        cols = cols[cols != i][:k]
        if cols.size < k:
            raise ValueError(f"row {i} has fewer than {k} neighbours")
        out[i] = cols
    return out


def score(Y, hd, P, k=10, n_eval=5000, n_anchors=1000, seed=0):
    """dict(kl, knn_preservation, Z, span); NaNs when Y is not finite (as the reference).
    Y (n, dim); hd (n, >= k) high-dimensional neighbours in distance order; P the common
    joint P (scipy CSR summing to 1)."""
    Y = np.asarray(Y, dtype=np.float64)
    n = Y.shape[0]
    if not np.isfinite(Y).all():
        return dict(kl=float("nan"), knn_preservation=float("nan"), Z=float("nan"),
                    span=float("nan"))
    rng = np.random.default_rng(seed)
    queries = rng.choice(n, size=min(n_eval, n), replace=False)
    anchors = rng.choice(n, size=min(n_anchors, n), replace=False)
    return get_backend().quality_score(Y, hd, queries, P, anchors, k=k)

"""k-NN providers for TSNE(metric='euclidean').

The k-NN search is upstream of the gradient path; this module offers the exact search
the reference falls back to (reference tsne_h2pack/_knn_methods.py: ExactKNN :71-119,
PyNNDescent's exact branch below 10K points :228-233) and returns the same CSR of L2
distances with the point itself dropped.  'pynndescent' is served by pynndescent when
it is importable and by the exact search otherwise.
"""
import time

import numpy as np


def _drop_self(indices, distances):
    """Remove each row's own index (or its farthest extra neighbour) so that k remain
    (reference _knn_methods.py:21-34)."""
    n, kp1 = indices.shape
    rows = np.arange(n)[:, None]
    is_self = indices == rows
    has_self = is_self.any(axis=1)
    keep = ~is_self
    keep[~has_self, -1] = False
    k = kp1 - 1
    return indices[keep].reshape(n, k), distances[keep].reshape(n, k)


def _to_csr(indices, distances, n):
    from scipy.sparse import csr_matrix
    k = indices.shape[1]
    indptr = np.arange(0, (n + 1) * k, k, dtype=np.int64)
    return csr_matrix((distances.ravel().astype(np.float32), indices.ravel(), indptr),
                      shape=(n, n))


class ExactKNN:
    """Brute-force Euclidean k-NN (sklearn), float32 distances."""

    def compute(self, X, k, n_jobs=1, verbose=False, **kwargs):
        from sklearn.neighbors import NearestNeighbors
        t0 = time.time()
        nn = NearestNeighbors(n_neighbors=k + 1, algorithm='brute', n_jobs=n_jobs).fit(X)
        dist, idx = nn.kneighbors(X)
        idx, dist = _drop_self(idx, dist)
        G = _to_csr(idx, dist, X.shape[0])
        G.knn_time = time.time() - t0
        return G


class PyNNDescentKNN(ExactKNN):
    def compute(self, X, k, n_jobs=1, verbose=False, **kwargs):
        try:
            import pynndescent  # noqa: F401
        except ImportError:
            return super().compute(X, k, n_jobs=n_jobs, verbose=verbose)
        if X.shape[0] < 10000:
            return super().compute(X, k, n_jobs=n_jobs, verbose=verbose)
        from pynndescent import NNDescent
        t0 = time.time()
        index = NNDescent(X, n_neighbors=k + 1, n_jobs=n_jobs, **kwargs)
        idx, dist = index.neighbor_graph
        idx, dist = _drop_self(idx, dist)
        G = _to_csr(idx, dist, X.shape[0])
        G.knn_time = time.time() - t0
        return G


class GPUExactKNN:
    """Exact Euclidean k-NN on the GPU (tsne_knn_exact): the same graph as ExactKNN --
    sklearn's brute-force kneighbors_graph(mode='distance') with float32 distances
    (reference _knn_methods.py:71-119) -- with neighbour sets exact in FP64."""

    def compute(self, X, k, n_jobs=1, verbose=False, **kwargs):
        from scipy.sparse import csr_matrix
        from ._backend import get_backend
        t0 = time.time()
        X = np.asarray(X, dtype=np.float64)
        n = X.shape[0]
        idx, dist, st = get_backend().knn_exact(X, k)
        indptr = np.arange(0, (n + 1) * k, k, dtype=np.int64)
        G = csr_matrix((dist.ravel(), idx.ravel(), indptr), shape=(n, n))
        G.knn_time = time.time() - t0
        if verbose:
            print(f"[t-SNE kNN] GPU exact search: {G.knn_time:.3f}s "
                  f"({st.uncertain_rows} rows re-ranked in FP64)")
        return G


_METHODS = {'exact': ExactKNN, 'pynndescent': PyNNDescentKNN, 'balltree': ExactKNN,
            'gpu': GPUExactKNN, 'exact_gpu': GPUExactKNN}


def get_knn_method(name):
    try:
        return _METHODS[name]()
    except KeyError:
        raise ValueError(f"Unknown knn_method '{name}'. Available: {sorted(_METHODS)}")

"""sklearn-compatible TSNE whose optimisation runs on an MI355X.

`TSNE(method='h2pack')` is a drop-in for the reference's estimator
(tsne_h2pack/manifold.py of xinxing02/tsne-h2pack-python 3.3.1): same constructor
signature and defaults, same fitted attributes, and the error-message fragments its
tests match.  One keyword is added: `h2_precision` ('plain' or 'mixed'; default from the
environment variable TSNE_H2_PRECISION, else 'plain') picks the precision of the
far-field H2 products on the GPU.

Structure of this module:
  * `_RULES`: declarative parameter checks, evaluated at construction and before use;
  * `_RunPlan`: the run parameters of one fit (schedule constants, learning rate, H2
    settings), written onto the library's `tsne_params_t`;
  * `_scaled_pca`: the 'pca' initial embedding;
  * `TSNE.fit_transform`: a k-NN graph (from the features, or the precomputed one), then
    one call into the backend.
"""
import os
import warnings
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Tuple, Union

import numpy as np

from ._backend import NP_DTYPE, TSNEParams, get_backend

# H2 settings per embedding dimension (the reference's defaults, manifold.py:44-47).
H2_DEFAULTS = {
    2: dict(h2_tol=1e-2, h2_rebuild_freq=3, h2_rebuild_freq_early=10, h2_leaf_size=100),
    3: dict(h2_tol=3e-2, h2_rebuild_freq=3, h2_rebuild_freq_early=10, h2_leaf_size=100),
}


def h2_defaults(n_components: int) -> dict:
    """H2 settings for an embedding of `n_components` dimensions (2D below 3)."""
    return dict(H2_DEFAULTS[3 if n_components >= 3 else 2])


def _worker_count(n_jobs: Optional[int]) -> int:
    """joblib's convention: None or 0 -> 1, negative -> cpu_count + 1 + n_jobs."""
    if not n_jobs:
        return 1
    return n_jobs if n_jobs > 0 else max(1, (os.cpu_count() or 1) + 1 + n_jobs)


# ----------------------------------------------------------------------------- checks

def _init_problem(init) -> Optional[str]:
    if isinstance(init, str):
        if init in ('random', 'pca'):
            return None
        return (f"init={init!r} not supported: use 'random', 'pca' or an array of shape "
                "(n_samples, n_components)")
    return None


def _init_array_problem(est) -> Optional[str]:
    if isinstance(est.init, str):
        return None
    arr = np.asarray(est.init)
    if arr.ndim != 2 or arr.shape[1] != est.n_components:
        got = f"shape {arr.shape}" if arr.ndim else repr(est.init)
        return (f"an init array needs n_components={est.n_components} columns "
                f"(shape (n_samples, n_components)); got {got}")
    if not np.issubdtype(arr.dtype, np.number) or not np.all(np.isfinite(arr)):
        return "the init array must hold finite numbers only"
    return None


def _learning_rate_problem(lr) -> Optional[str]:
    if isinstance(lr, str):
        return None if lr == 'auto' else f"learning_rate={lr!r}: give 'auto' or a number > 0"
    return None if lr > 0 else f"learning_rate must be > 0 (got {lr})"


# (check, message) pairs; a check returns a message (a problem) or None.
_Rule = Callable[["TSNE"], Optional[str]]
_RULES: List[_Rule] = [
    lambda e: None if e.metric in ('euclidean', 'precomputed') else
    f"metric={e.metric!r} not supported: use 'euclidean' or 'precomputed'",
    lambda e: _init_problem(e.init),
    lambda e: _init_array_problem(e),
    lambda e: None if e.method in ('h2pack', 'exact') else
    f"method={e.method!r} not supported: use 'h2pack' or 'exact'",
    lambda e: None if e.n_components >= 1 else f"n_components must be >= 1 (got {e.n_components})",
    lambda e: None if e.perplexity > 0 else f"perplexity must be > 0 (got {e.perplexity})",
    lambda e: _learning_rate_problem(e.learning_rate),
    lambda e: None if e.max_iter >= 1 else f"max_iter must be >= 1 (got {e.max_iter})",
    lambda e: None if min(e.h2_rebuild_freq, e.h2_rebuild_freq_early) >= 1 else
    "h2_rebuild_freq and h2_rebuild_freq_early must be >= 1",
    lambda e: None if e.h2_leaf_size >= 1 else f"h2_leaf_size must be >= 1 (got {e.h2_leaf_size})",
    lambda e: None if e.h2_precision in ('plain', 'mixed') else
    f"h2_precision={e.h2_precision!r}: use 'plain' or 'mixed'",
]


# ----------------------------------------------------------------------------- run plan

# Schedule of the reference's sklearn-compatible mode: exaggeration and the momentum
# switch at iteration 250 (reference src/tsne_h2pack.c:1736-1761).
_SCHEDULE = dict(early_exag_iter=250, initial_momentum=0.5, momentum=0.8,
                 momentum_switch_iter=250, print_freq=50, min_iter=250,
                 sklearn_compatible=1)
# method='exact': the H2 structure is never used after the first iterations and every
# iteration takes the direct O(N^2) sum (reference src/tsne_h2pack.c:404).
_EXACT_H2 = dict(h2_rebuild_freq=999999, h2_tol=1e-4)


@dataclass
class _RunPlan:
    fields: dict = field(default_factory=dict)
    keep_alive: Tuple = ()

    def apply(self, params: TSNEParams) -> TSNEParams:
        for name, value in self.fields.items():
            setattr(params, name, value)
        return params


def _auto_learning_rate(n_samples: Optional[int]) -> float:
    return max((n_samples or 0) / 12.0, 200.0)


# ----------------------------------------------------------------------------- init

def _scaled_pca(X, n_components: int, random_state) -> np.ndarray:
    """sklearn's randomized PCA of X, scaled so its first column has std 1e-4 (the
    reference's 'pca' init, manifold.py:380-424)."""
    from sklearn.decomposition import PCA
    n, d = X.shape
    if n_components > min(n, d):
        raise ValueError(f"init='pca' needs n_components={n_components} <= "
                         f"min(n_samples, n_features) = {min(n, d)}")
    with np.errstate(divide='ignore', over='ignore', invalid='ignore'):
        Y = PCA(n_components=n_components, svd_solver='randomized',
                random_state=random_state).fit_transform(np.asarray(X, dtype=np.float64))
    scale = np.std(Y[:, 0]) if Y.size else 0.0
    if not np.all(np.isfinite(Y)) or not np.isfinite(scale) or scale == 0:
        raise ValueError("init='pca' failed: the principal components of X are not finite "
                         "or have no spread; check X or use init='random'")
    return Y * (1e-4 / scale)


# ----------------------------------------------------------------------------- estimator

class TSNE:
    """t-SNE with the H2-accelerated repulsive force computed on an MI355X.

    Parameters follow sklearn.manifold.TSNE and the reference (manifold.py:209-264);
    `h2_precision` is added.  Fitted attributes: embedding_, kl_divergence_, n_iter_,
    knn_time_."""

    def __init__(
        self,
        n_components: int = 2,
        perplexity: float = 30.0,
        early_exaggeration: float = 12.0,
        learning_rate: Union[float, str] = 'auto',
        max_iter: int = 1000,
        n_iter_without_progress: int = 300,
        min_grad_norm: float = 1e-7,
        metric: str = 'euclidean',
        init: Union[str, np.ndarray] = 'pca',
        verbose: int = 0,
        random_state: Optional[int] = None,
        method: str = 'h2pack',
        n_jobs: Optional[int] = None,
        h2_tol: Optional[float] = None,
        h2_rebuild_freq: Optional[int] = None,
        enable_profiling: bool = False,
        profiling_output_file: Optional[str] = None,
        knn_method: str = 'pynndescent',
        knn_params: Optional[dict] = None,
        h2_rebuild_freq_early: Optional[int] = None,
        h2_leaf_size: Optional[int] = None,
        h2_precision: Optional[str] = None,
    ):
        self.n_components = n_components
        self.perplexity = perplexity
        self.early_exaggeration = early_exaggeration
        self.learning_rate = learning_rate
        self.max_iter = max_iter
        self.n_iter_without_progress = n_iter_without_progress
        self.min_grad_norm = min_grad_norm
        self.metric = metric
        self.init = init
        self.verbose = verbose
        self.random_state = random_state
        self.method = method
        self.n_jobs = _worker_count(n_jobs)
        self.enable_profiling = enable_profiling
        self.profiling_output_file = profiling_output_file
        self.knn_method = knn_method
        self.knn_params = dict(knn_params or {})
        # unset H2 settings take the defaults of the embedding dimension
        dims = n_components if isinstance(n_components, int) and n_components > 0 else 2
        given = dict(h2_tol=h2_tol, h2_rebuild_freq=h2_rebuild_freq,
                     h2_rebuild_freq_early=h2_rebuild_freq_early, h2_leaf_size=h2_leaf_size)
        for name, default in h2_defaults(dims).items():
            setattr(self, name, default if given[name] is None else given[name])
        self.h2_precision = h2_precision or os.environ.get("TSNE_H2_PRECISION", "plain")
        self.embedding_: Optional[np.ndarray] = None
        self.kl_divergence_: Optional[float] = None
        self.n_iter_: Optional[int] = None
        self._validate_params()

    # ---- checks

    def _validate_params(self):
        for rule in _RULES:
            problem = rule(self)
            if problem:
                raise ValueError(problem)

    def _validate_data(self, X) -> np.ndarray:
        """X as a float64 (n_samples, n_features) array with n_samples >= 2, finite, and
        more samples than the perplexity."""
        X = np.asarray(X, dtype=NP_DTYPE)
        if X.ndim != 2:
            raise ValueError(f"X must be 2-dimensional (n_samples, n_features); it has "
                             f"{X.ndim} dimensions")
        n = X.shape[0]
        if n < 2:
            raise ValueError(f"n_samples={n} should be >= 2")
        if self.perplexity >= n:
            raise ValueError(f"perplexity={self.perplexity} must be less than n_samples={n}")
        if not np.all(np.isfinite(X)):
            raise ValueError("X contains NaN or inf values")
        return X

    # ---- run parameters

    def _run_plan(self, n_samples: Optional[int]) -> _RunPlan:
        f = dict(_SCHEDULE)
        f.update(n_iter=self.max_iter, perplexity=self.perplexity,
                 early_exag=self.early_exaggeration,
                 learning_rate=(_auto_learning_rate(n_samples) if self.learning_rate == 'auto'
                                else self.learning_rate),
                 verbose=int(self.verbose > 0), num_threads=self.n_jobs,
                 convergence_patience=self.n_iter_without_progress,
                 h2_leaf_size=self.h2_leaf_size,
                 h2_rebuild_freq_early=max(self.h2_rebuild_freq_early, self.h2_rebuild_freq),
                 enable_profiling=int(bool(self.enable_profiling)))
        f.update(_EXACT_H2 if self.method == 'exact' else
                 dict(h2_rebuild_freq=self.h2_rebuild_freq, h2_tol=self.h2_tol))
        keep = ()
        if self.profiling_output_file is not None:
            name = self.profiling_output_file.encode('utf-8')
            f["profiling_output_file"] = name
            keep = (name,)
        if self.random_state is not None:
            f["random_seed"] = self.random_state
        return _RunPlan(f, keep)

    def _map_params_to_c_struct(self, n_samples: Optional[int] = None) -> TSNEParams:
        """The library's tsne_params_t for a fit on n_samples points (the reference's
        mapping, manifold.py:319-378)."""
        plan = self._run_plan(n_samples)
        self._plan_refs = plan.keep_alive  # the C struct points into these bytes
        return plan.apply(get_backend().create_default_params())

    # ---- initial embedding

    def _initial_embedding(self, X, n_samples: int) -> Optional[np.ndarray]:
        """(n_samples, n_components) start, or None for the library's seeded random draw
        (init='random', and init='pca' without features)."""
        self._validate_params()
        if isinstance(self.init, str):
            if self.init == 'random':
                return None
            if X is None:
                warnings.warn("init='pca' needs the features, which metric='precomputed' does "
                              "not give; falling back to init='random' (pass init='random' or "
                              "an array to silence this)", UserWarning, stacklevel=3)
                return None
            return _scaled_pca(X, self.n_components, self.random_state)
        Y = np.array(self.init, dtype=np.float64)
        if Y.shape != (n_samples, self.n_components):
            raise ValueError(f"the init array has shape {Y.shape}, expected "
                             f"(n_samples, n_components) = {(n_samples, self.n_components)}")
        return Y

    # ---- k-NN graphs

    def _graph_from_precomputed(self, G):
        """The precomputed k-NN distance graph as the reference takes it: square CSR,
        no self entries (nor zero distances), columns sorted, finite non-negative
        distances, more neighbours per row than the perplexity."""
        from scipy.sparse import csr_matrix, issparse
        if not issparse(G):
            raise ValueError("metric='precomputed' takes a sparse matrix of k-NN distances, "
                             "e.g. kneighbors_graph(X, n_neighbors=90, mode='distance')")
        if G.shape[0] != G.shape[1]:
            raise ValueError(f"the precomputed k-NN graph must be square, got {G.shape}")
        graph = csr_matrix(G, copy=True)
        graph.setdiag(0)
        graph.eliminate_zeros()
        graph.sort_indices()
        if not np.all(np.isfinite(graph.data)) or np.any(graph.data < 0):
            raise ValueError("precomputed distances must be finite and non-negative")
        fewest = int(np.diff(graph.indptr).min()) if graph.shape[0] else 0
        if fewest <= self.perplexity:
            raise ValueError(f"every row needs more than perplexity={self.perplexity} "
                             f"neighbors; one row has {fewest}")
        return graph

    def _graph_from_features(self, X):
        from ._knn import get_knn_method
        k = min(X.shape[0] - 1, int(3 * self.perplexity))
        if self.verbose:
            print(f"Computing {k} nearest neighbors using {self.knn_method}...")
        return get_knn_method(self.knn_method).compute(X, k=k, n_jobs=self.n_jobs,
                                                       verbose=self.verbose > 0,
                                                       **self.knn_params)

    # ---- fit

    def fit(self, X, y=None):
        raise NotImplementedError("fit() is not implemented. Use fit_transform() instead.")

    def fit_transform(self, X, y=None) -> np.ndarray:
        if self.metric == 'precomputed':
            features, graph = None, self._graph_from_precomputed(X)
        else:
            features = self._validate_data(X)
            graph = None
        n = graph.shape[0] if graph is not None else features.shape[0]
        init_Y = self._initial_embedding(features, n)
        backend = get_backend()
        if self.verbose:
            print(f"Running t-SNE with backend: {backend.backend_type}")
        if graph is None:
            graph = self._graph_from_features(features)
        self.knn_time_ = float(getattr(graph, 'knn_time', 0.0))
        params = self._map_params_to_c_struct(n_samples=n)
        Y, info = backend.fit_transform_with_knn_distances(
            features, graph, self.n_components, params, init_Y=init_Y,
            h2_precision=self.h2_precision)
        self.embedding_ = Y
        self.n_iter_ = int(info.get('n_iter', params.n_iter))
        self.kl_divergence_ = float(info.get('kl_divergence', 0.0))
        return self.embedding_

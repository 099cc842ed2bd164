"""ctypes binding of libtsne_h2pack.so (the MI355X HIP library) and the host side of
the P pipeline.

Mirrors the reference's binding layer (tsne_h2pack/_backend.py of
xinxing02/tsne-h2pack-python 3.3.1): the same struct mirrors (TSNEParams 160 bytes,
SparseMatrix 40 bytes), the same thirteen C entry points with the same argument
types, and the same k-NN-CSR -> conditional P -> symmetric P -> optimisation flow
(_backend.py:611-781).  The native library is the GPU one; if it is missing the
loader raises instead of falling back to anything else.
"""
import ctypes
import os
import time
from pathlib import Path
from typing import Optional, Tuple

import numpy as np

DTYPE = ctypes.c_double
NP_DTYPE = np.float64


class TSNEParams(ctypes.Structure):
    """tsne_params_t (include/tsne_h2pack.h), 160 bytes; explicit padding fields keep the
    offsets of the reference's mirror (_backend.py:53-95)."""
    _fields_ = [
        ('n_iter', ctypes.c_int),
        ('_pad1', ctypes.c_int),
        ('perplexity', DTYPE),
        ('learning_rate', DTYPE),
        ('momentum', DTYPE),
        ('initial_momentum', DTYPE),
        ('momentum_switch_iter', ctypes.c_int),
        ('_pad2', ctypes.c_int),
        ('early_exag', DTYPE),
        ('early_exag_iter', ctypes.c_int),
        ('use_vptree', ctypes.c_int),
        ('h2_tol', DTYPE),
        ('h2_rebuild_freq', ctypes.c_int),
        ('h2pack_print_level', ctypes.c_int),
        ('init_method', ctypes.c_int),
        ('random_seed', ctypes.c_int),
        ('verbose', ctypes.c_int),
        ('print_freq', ctypes.c_int),
        ('enable_early_stopping', ctypes.c_int),
        ('_pad4', ctypes.c_int),
        ('convergence_tol', DTYPE),
        ('convergence_patience', ctypes.c_int),
        ('min_iter', ctypes.c_int),
        ('sklearn_compatible', ctypes.c_int),
        ('enable_profiling', ctypes.c_int),
        ('profiling_output_file', ctypes.c_char_p),
        ('num_threads', ctypes.c_int),
        ('h2_leaf_size', ctypes.c_int),
        ('h2_rebuild_freq_early', ctypes.c_int),
        ('_pad5', ctypes.c_int),
        ('init_Y', ctypes.POINTER(DTYPE)),
    ]


class SparseMatrix(ctypes.Structure):
    """sparse_matrix_t: CSR with int32 indices and float64 values (40 bytes)."""
    _fields_ = [
        ('n_rows', ctypes.c_int),
        ('n_cols', ctypes.c_int),
        ('nnz', ctypes.c_int),
        ('row_ptr', ctypes.POINTER(ctypes.c_int)),
        ('col_idx', ctypes.POINTER(ctypes.c_int)),
        ('values', ctypes.POINTER(DTYPE)),
    ]


class TSNEWorkspace(ctypes.Structure):
    """Opaque tsne_workspace_t."""


class H2Stats(ctypes.Structure):
    """tsne_h2_stats_t."""
    _fields_ = [
        ('levels', ctypes.c_int), ('n_nodes', ctypes.c_int), ('n_leaves', ctypes.c_int),
        ('leaf_points_max', ctypes.c_int),
        ('n_adm_pairs', ctypes.c_longlong), ('n_near_pairs', ctypes.c_longlong),
        ('near_point_pairs', ctypes.c_longlong),
        ('rank_max', ctypes.c_int), ('rank_mean', ctypes.c_double),
        ('proxy_max', ctypes.c_int), ('far_flops', ctypes.c_double),
        ('near_flops', ctypes.c_double), ('mixed_fallback_nodes', ctypes.c_int),
        ('direct_fallbacks', ctypes.c_int),
        ('rank_capped_builds', ctypes.c_int),
        ('id_flops', ctypes.c_double), ('id_ms', ctypes.c_double),
    ]


class KnnStats(ctypes.Structure):
    """tsne_knn_stats_t."""
    _fields_ = [
        ('survivors', ctypes.c_int), ('uncertain_rows', ctypes.c_int),
        ('overflow_rows', ctypes.c_int), ('eps', ctypes.c_double),
        ('device_ms', ctypes.c_double), ('screen_ms', ctypes.c_double),
        ('screen_flops', ctypes.c_longlong),
    ]


class Quality(ctypes.Structure):
    """tsne_quality_t."""
    _fields_ = [
        ('kl', ctypes.c_double), ('knn_preservation', ctypes.c_double),
        ('Z', ctypes.c_double), ('span', ctypes.c_double),
    ]


class Tuning(ctypes.Structure):
    """tsne_tuning_t: launch shapes of a job (none changes a result)."""
    _fields_ = [
        ('overlap', ctypes.c_int), ('level_batch', ctypes.c_int), ('graph', ctypes.c_int),
        ('one_class_level', ctypes.c_int), ('ms', ctypes.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class JobInfo(ctypes.Structure):
    """tsne_job_info_t."""
    _fields_ = [
        ('job_id', ctypes.c_int), ('status', ctypes.c_int), ('n_point', ctypes.c_int),
        ('low_dim', ctypes.c_int), ('kl', ctypes.c_double), ('submit_s', ctypes.c_double),
        ('start_s', ctypes.c_double), ('end_s', ctypes.c_double),
        ('footprint_bytes', ctypes.c_longlong),
    ]


def _optional_ndpointer(**kwargs):
    """ndpointer that also accepts None (NULL)."""
    base = np.ctypeslib.ndpointer(**kwargs)

    def from_param(cls, obj):
        return None if obj is None else base.from_param(obj)

    return type(base.__name__, (base,), {'from_param': classmethod(from_param)})


def library_path() -> Path:
    """The in-tree HIP library (tsne_h2pack/_lib/libtsne_h2pack.so), or the file named by
    TSNE_H2PACK_LIB."""
    env = os.environ.get("TSNE_H2PACK_LIB")
    if env:
        return Path(env)
    return Path(__file__).parent / '_lib' / 'libtsne_h2pack.so'


class TSNEBackend:
    """Loads libtsne_h2pack.so and drives it (reference _backend.py:129-781)."""

    def __init__(self):
        self._lib = self._load_native_library()
        self._backend_type = 'hip'

    @staticmethod
    def _load_native_library():
        path = library_path()
        if not path.exists():
            raise RuntimeError(
                f"Native library not found at {path}. Build it with "
                "`python tsne-h2pack-python_amd/build_lib.py` (hipcc, gfx950).")
        try:
            lib = ctypes.CDLL(str(path))
        except OSError as e:
            raise RuntimeError(f"Failed to load native library from {path}: {e}")
        configure_signatures(lib)
        return lib

    @property
    def lib(self):
        return self._lib

    @property
    def backend_type(self) -> str:
        return self._backend_type

    def is_available(self) -> bool:
        return True

    def create_default_params(self) -> TSNEParams:
        """Defaults as in the reference (_backend.py:450-502)."""
        p = TSNEParams()
        p.n_iter = 1000
        p.perplexity = 30.0
        p.learning_rate = -1.0
        p.momentum = 0.8
        p.initial_momentum = 0.5
        p.momentum_switch_iter = 250
        p.early_exag = 12.0
        p.early_exag_iter = 250
        p.use_vptree = 2
        p.h2_tol = 1e-3
        p.h2_rebuild_freq = 10
        p.h2pack_print_level = 0
        p.init_method = 0
        p.random_seed = -1
        p.verbose = 1
        p.print_freq = 50
        p.enable_early_stopping = 0
        p.convergence_tol = 1e-5
        p.convergence_patience = 10
        p.min_iter = 250
        p.sklearn_compatible = 1
        p.enable_profiling = 0
        p.profiling_output_file = None
        p.num_threads = 1
        p.h2_leaf_size = 100
        p.h2_rebuild_freq_early = 10
        p.init_Y = None
        return p

    def fit_transform(self, X: np.ndarray, n_components: int,
                      params: TSNEParams) -> Tuple[np.ndarray, dict]:
        """Legacy features path (reference _backend.py:504-551).  The library has no
        built-in k-NN search, so this raises; use fit_transform_with_knn_distances."""
        n, f = X.shape
        Xc = np.ascontiguousarray(np.asarray(X, dtype=NP_DTYPE).T)
        Y = np.zeros((n_components, n), dtype=NP_DTYPE)
        ret = self._lib.tsne_fit_transform(Xc, n, f, n_components, ctypes.byref(params), Y)
        if ret != 0:
            raise RuntimeError(f"tsne_fit_transform failed with code {ret}")
        return np.ascontiguousarray(Y.T), {'kl_divergence': 0.0, 'n_iter': params.n_iter}

    def fit_transform_with_P(self, X, P_scipy, n_components: int,
                             params: TSNEParams) -> Tuple[np.ndarray, dict]:
        """Embed with a precomputed symmetric P (reference _backend.py:553-609).  P is
        copied into library-owned memory (the library takes ownership)."""
        from scipy.sparse import csr_matrix
        if not isinstance(P_scipy, csr_matrix):
            raise TypeError(f"P must be scipy CSR matrix, got {type(P_scipy)}")
        n = P_scipy.shape[0]
        n_features = X.shape[1] if X is not None else 0
        P_c = scipy_csr_to_c_sparse_matrix(P_scipy, self._lib)
        Y = np.zeros((n_components, n), dtype=NP_DTYPE)
        ret = self._lib.tsne_fit_transform_with_P(None, n, n_features, n_components,
                                                  ctypes.byref(params), P_c, Y)
        if ret != 0:
            raise RuntimeError(f"tsne_fit_transform_with_P failed with code {ret}")
        return np.ascontiguousarray(Y.T), {'kl_divergence': 0.0, 'n_iter': params.n_iter}

    def prepare_knn_csr(self, knn_distances_csr):
        """Neighbour indexing (reference _backend.py:647-674): columns sorted within each
        row, self entries removed; returns (float64 data, int32 indices, int32 indptr)."""
        from scipy.sparse import csr_matrix
        G = knn_distances_csr
        n = G.shape[0]
        if not G.has_sorted_indices:
            G = G.copy()
            G.sort_indices()
        rows = np.repeat(np.arange(n), np.diff(G.indptr))
        if np.any(G.indices == rows):
            keep = G.indices != rows
            G = csr_matrix((G.data[keep], G.indices[keep],
                            np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n))])),
                           shape=G.shape)
        return (np.ascontiguousarray(G.data, dtype=np.float64),
                np.ascontiguousarray(G.indices, dtype=np.int32),
                np.ascontiguousarray(G.indptr, dtype=np.int32))

    def conditional_P(self, data, indices, indptr, perplexity, num_threads=1):
        """compute_P_from_knn_distances on the GPU; returns scipy CSR of P_j|i."""
        from scipy.sparse import csr_matrix
        n = indptr.size - 1
        ptr = self._lib.compute_P_from_knn_distances(
            data.ctypes.data_as(ctypes.POINTER(DTYPE)),
            indices.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            indptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            n, int(indptr[-1]), ctypes.c_double(perplexity), int(num_threads))
        if not ptr:
            raise RuntimeError("compute_P_from_knn_distances returned NULL")
        m = ptr.contents
        rp = np.ctypeslib.as_array(m.row_ptr, shape=(m.n_rows + 1,)).copy()
        ci = np.ctypeslib.as_array(m.col_idx, shape=(max(m.nnz, 1),))[:m.nnz].copy()
        va = np.ctypeslib.as_array(m.values, shape=(max(m.nnz, 1),))[:m.nnz].copy()
        self._lib.sparse_matrix_destroy(ptr)
        return csr_matrix((va, ci, rp), shape=(n, n))

    @staticmethod
    def symmetrize(P_cond):
        """P = (P + P^T) / sum on the host (reference _backend.py:709-718, scipy
        arithmetic); joint_P does the same on the device."""
        P = P_cond + P_cond.T
        total = P.sum()
        if total > 0:
            P.data /= total
        return P

    def joint_P_device(self, data, indices, indptr, perplexity):
        """tsne_joint_P_from_knn_distances: conditional P, symmetrisation and
        normalisation on the GPU, equal bit for bit to symmetrize(conditional_P(...))."""
        from scipy.sparse import csr_matrix
        n = indptr.size - 1
        ptr = self._lib.tsne_joint_P_from_knn_distances(
            data.ctypes.data_as(ctypes.POINTER(DTYPE)),
            indices.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            indptr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
            n, int(indptr[-1]), ctypes.c_double(perplexity))
        if not ptr:
            raise RuntimeError("tsne_joint_P_from_knn_distances returned NULL")
        m = ptr.contents
        rp = np.ctypeslib.as_array(m.row_ptr, shape=(m.n_rows + 1,)).copy()
        ci = np.ctypeslib.as_array(m.col_idx, shape=(max(m.nnz, 1),))[:m.nnz].copy()
        va = np.ctypeslib.as_array(m.values, shape=(max(m.nnz, 1),))[:m.nnz].copy()
        self._lib.sparse_matrix_destroy(ptr)
        return csr_matrix((va, ci, rp), shape=(n, n))

    def knn_exact(self, X, k):
        """tsne_knn_exact: exact Euclidean k-NN on the GPU.  Returns (indices (n, k) int32,
        distances (n, k) float32, KnnStats); rows in increasing distance, the point itself
        removed as sklearn's kneighbors_graph removes it."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        if X.ndim != 2:
            raise ValueError("X must be 2-D (n_samples, n_features)")
        n, d = X.shape
        idx = np.empty((n, k), dtype=np.int32)
        dist = np.empty((n, k), dtype=np.float32)
        st = KnnStats()
        rc = self._lib.tsne_knn_exact(X, n, d, int(k), idx, dist, ctypes.byref(st))
        if rc != 0:
            raise RuntimeError("tsne_knn_exact failed (see stderr)")
        return idx, dist, st

    def quality_score(self, Y, hd, queries, P, anchors, k=10):
        """tsne_quality_score: the reference's common-footing quality score on the GPU
        (benchmarks/quality.py:56-81) with its random draws passed in.  Y (n, dim); hd (n,
        >= k) high-dimensional neighbours in distance order; P scipy CSR.  Returns dict(kl,
        knn_preservation, Z, span)."""
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        n, dim = Y.shape
        hd = np.ascontiguousarray(hd, dtype=np.int32)
        q = np.ascontiguousarray(queries, dtype=np.int32)
        an = np.ascontiguousarray(anchors, dtype=np.int32)
        rp = np.ascontiguousarray(P.indptr, dtype=np.int32)
        ci = np.ascontiguousarray(P.indices, dtype=np.int32)
        va = np.ascontiguousarray(P.data, dtype=np.float64)
        out = Quality()
        i32 = ctypes.POINTER(ctypes.c_int)
        rc = self._lib.tsne_quality_score(
            Y.ctypes.data_as(ctypes.POINTER(DTYPE)), n, dim, hd.ctypes.data_as(i32), hd.shape[1],
            int(k), q.ctypes.data_as(i32), q.size, rp.ctypes.data_as(i32), ci.ctypes.data_as(i32),
            va.ctypes.data_as(ctypes.POINTER(DTYPE)), an.ctypes.data_as(i32), an.size,
            ctypes.byref(out))
        if rc != 0:
            raise RuntimeError("tsne_quality_score failed (see stderr)")
        return dict(kl=out.kl, knn_preservation=out.knn_preservation, Z=out.Z, span=out.span)

    def joint_P_from_points(self, X, k, perplexity):
        """tsne_joint_P_from_points: exact k-NN graph, neighbour indexing and joint P, all on
        the device; equal bit for bit to joint_P(knn graph of ExactKNN).  Returns (scipy CSR,
        KnnStats)."""
        from scipy.sparse import csr_matrix
        X = np.ascontiguousarray(X, dtype=np.float64)
        n, d = X.shape
        st = KnnStats()
        ptr = self._lib.tsne_joint_P_from_points(X, n, d, int(k), ctypes.c_double(perplexity),
                                                 ctypes.byref(st))
        if not ptr:
            raise RuntimeError("tsne_joint_P_from_points returned NULL")
        m = ptr.contents
        rp = np.ctypeslib.as_array(m.row_ptr, shape=(m.n_rows + 1,)).copy()
        ci = np.ctypeslib.as_array(m.col_idx, shape=(max(m.nnz, 1),))[:m.nnz].copy()
        va = np.ctypeslib.as_array(m.values, shape=(max(m.nnz, 1),))[:m.nnz].copy()
        self._lib.sparse_matrix_destroy(ptr)
        return csr_matrix((va, ci, rp), shape=(n, n)), st

    def tuning_lookup(self, low_dim, n_point):
        """The device tuning table's launch shapes for a job of this size: (dict, whether
        the table holds an entry for its size bucket)."""
        t = Tuning()
        rc = self._lib.tsne_tuning_lookup(int(low_dim), int(n_point), ctypes.byref(t))
        if rc < 0:
            raise ValueError("tsne_tuning_lookup: low_dim must be 2 or 3 and n_point >= 1")
        return t.as_dict(), rc == 1

    def autotune(self, low_dim, n_point, n_iter=400):
        """tsne_autotune: times a synthetic job of n_point points under each launch shape,
        checks that every shape gives the same embedding bit for bit, and stores the
        fastest in the device's persistent tuning table.  Returns the chosen shapes."""
        t = Tuning()
        if self._lib.tsne_autotune(int(low_dim), int(n_point), int(n_iter), ctypes.byref(t)) != 0:
            raise RuntimeError("tsne_autotune failed")
        return t.as_dict()

    def tuning_path(self):
        """Path of the device's tuning table file (None when the cache is off)."""
        buf = ctypes.create_string_buffer(4096)
        return buf.value.decode() if self._lib.tsne_tuning_path(buf, 4096) == 0 else None

    def set_tuning(self, ws, overlap=-1, level_batch=0, graph=-1, one_class_level=-1):
        """tsne_set_tuning on a workspace (out-of-range fields keep the current choice)."""
        t = Tuning(int(overlap), int(level_batch), int(graph), int(one_class_level), 0.0)
        if self._lib.tsne_set_tuning(ws, ctypes.byref(t)) != 0:
            raise RuntimeError("tsne_set_tuning failed")

    def joint_P(self, knn_distances_csr, perplexity, num_threads=1):
        """Neighbour indexing on the host, then the joint P on the device."""
        data, indices, indptr = self.prepare_knn_csr(knn_distances_csr)
        return self.joint_P_device(data, indices, indptr, perplexity)

    def init_workspace(self, P, n_components, params, init_Y=None, n_features=0):
        """tsne_init_with_P with a library-owned copy of P; returns the workspace."""
        n = P.shape[0]
        P_c = scipy_csr_to_c_sparse_matrix(P, self._lib)
        keep = None
        if init_Y is not None:
            init_Y = np.asarray(init_Y)
            if init_Y.shape != (n, n_components):
                raise ValueError(
                    f"init_Y must have shape {(n, n_components)}, got {init_Y.shape}")
            keep = np.ascontiguousarray(init_Y.T, dtype=NP_DTYPE)
            params.init_method = 1
            params.init_Y = keep.ctypes.data_as(ctypes.POINTER(DTYPE))
        else:
            params.init_method = 0
            params.init_Y = None
        ws = self._lib.tsne_init_with_P(None, n, n_features, n_components,
                                        ctypes.byref(params), P_c)
        params.init_Y = None
        del keep
        if not ws:
            self._lib.sparse_matrix_destroy(P_c)
            raise RuntimeError("tsne_init_with_P returned NULL")
        return ws

    def fit_transform_with_knn_distances(self, X, knn_distances_csr, n_components: int,
                                         params: TSNEParams,
                                         init_Y: Optional[np.ndarray] = None,
                                         h2_precision: Optional[str] = None
                                         ) -> Tuple[np.ndarray, dict]:
        """k-NN CSR distances (L2, not squared) -> P -> embedding (reference
        _backend.py:611-781)."""
        from scipy.sparse import csr_matrix
        if not isinstance(knn_distances_csr, csr_matrix):
            raise TypeError(
                f"knn_distances must be scipy CSR matrix, got {type(knn_distances_csr)}")
        n = knn_distances_csr.shape[0]
        n_features = X.shape[1] if X is not None else 0
        t0 = time.time()
        P = self.joint_P(knn_distances_csr, params.perplexity, params.num_threads)
        p_time = time.time() - t0
        knn_time = getattr(knn_distances_csr, 'knn_time', 0.0)
        ws = self.init_workspace(P, n_components, params, init_Y, n_features)
        try:
            if h2_precision is not None:
                mode = {'plain': 0, 'mixed': 1}[h2_precision]
                self._lib.tsne_set_h2_precision(ws, mode)
            if params.enable_profiling:
                self._lib.tsne_set_init_profiling_times(ws, ctypes.c_double(knn_time),
                                                        ctypes.c_double(p_time))
            ret = self._lib.tsne_run(ws)
            if ret != 0:
                raise RuntimeError(f"tsne_run failed with code {ret}")
            Y = np.zeros((n_components, n), dtype=NP_DTYPE)
            self._lib.tsne_get_embedding(ws, Y)
            kl = float(self._lib.compute_kl_divergence(ws))
        finally:
            self._lib.tsne_destroy(ws)
        return np.ascontiguousarray(Y.T), {'kl_divergence': kl, 'n_iter': params.n_iter}


def configure_signatures(lib):
    """argtypes / restype of every entry point (reference _backend.py:345-448 for the
    thirteen it binds, plus the additive ones)."""
    c_ws = ctypes.POINTER(TSNEWorkspace)
    c_sm = ctypes.POINTER(SparseMatrix)
    c_pp = ctypes.POINTER(TSNEParams)
    f64c = np.ctypeslib.ndpointer(dtype=NP_DTYPE, flags='C_CONTIGUOUS')
    lib.tsne_create_default_params.argtypes = []
    lib.tsne_create_default_params.restype = TSNEParams
    lib.tsne_init.argtypes = [f64c, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_pp]
    lib.tsne_init.restype = c_ws
    lib.tsne_init_with_P.argtypes = [
        _optional_ndpointer(dtype=NP_DTYPE, flags='C_CONTIGUOUS', ndim=2),
        ctypes.c_int, ctypes.c_int, ctypes.c_int, c_pp, c_sm]
    lib.tsne_init_with_P.restype = c_ws
    lib.tsne_run.argtypes = [c_ws]
    lib.tsne_run.restype = ctypes.c_int
    lib.tsne_get_embedding.argtypes = [c_ws, f64c]
    lib.tsne_get_embedding.restype = None
    lib.tsne_destroy.argtypes = [c_ws]
    lib.tsne_destroy.restype = None
    lib.tsne_fit_transform.argtypes = [f64c, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_pp, f64c]
    lib.tsne_fit_transform.restype = ctypes.c_int
    lib.tsne_fit_transform_with_P.argtypes = [
        _optional_ndpointer(dtype=NP_DTYPE, flags='C_CONTIGUOUS', ndim=2),
        ctypes.c_int, ctypes.c_int, ctypes.c_int, c_pp, c_sm, f64c]
    lib.tsne_fit_transform_with_P.restype = ctypes.c_int
    lib.compute_P_from_knn_distances.argtypes = [
        ctypes.POINTER(DTYPE), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
        ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    lib.compute_P_from_knn_distances.restype = c_sm
    lib.sparse_matrix_destroy.argtypes = [c_sm]
    lib.sparse_matrix_destroy.restype = None
    lib.sparse_matrix_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.sparse_matrix_create.restype = c_sm
    lib.tsne_set_init_profiling_times.argtypes = [c_ws, ctypes.c_double, ctypes.c_double]
    lib.tsne_set_init_profiling_times.restype = None
    lib.compute_kl_divergence.argtypes = [c_ws]
    lib.compute_kl_divergence.restype = DTYPE
    # additive entry points
    ip = ctypes.POINTER(ctypes.c_int)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.tsne_set_h2_precision.argtypes = [c_ws, ctypes.c_int]
    lib.tsne_set_h2_precision.restype = ctypes.c_int
    lib.tsne_workspace_info.argtypes = [c_ws, ip, ip, dp]
    lib.tsne_workspace_info.restype = ctypes.c_int
    lib.tsne_set_embedding.argtypes = [c_ws, f64c]
    lib.tsne_set_embedding.restype = ctypes.c_int
    lib.tsne_compute_gradient.argtypes = [c_ws, ctypes.c_double, ctypes.c_int, f64c, dp]
    lib.tsne_compute_gradient.restype = ctypes.c_int
    lib.tsne_exact_repulsive_rows.argtypes = [
        c_ws, np.ctypeslib.ndpointer(dtype=np.int32, flags='C_CONTIGUOUS'), ctypes.c_int,
        f64c, dp]
    lib.tsne_exact_repulsive_rows.restype = ctypes.c_int
    lib.tsne_exact_repulsive_rows_noalloc.argtypes = lib.tsne_exact_repulsive_rows.argtypes
    lib.tsne_exact_repulsive_rows_noalloc.restype = ctypes.c_int
    lib.tsne_h2_stats.argtypes = [c_ws, ctypes.POINTER(H2Stats)]
    lib.tsne_h2_stats.restype = ctypes.c_int
    lib.tsne_run_timings.argtypes = [c_ws, dp, dp, dp, dp, ip]
    lib.tsne_run_timings.restype = ctypes.c_int
    lib.tsne_set_graph_mode.argtypes = [c_ws, ctypes.c_int]
    lib.tsne_set_graph_mode.restype = ctypes.c_int
    lib.tsne_graph_launches.argtypes = [c_ws]
    lib.tsne_graph_launches.restype = ctypes.c_int
    lib.tsne_set_tuning.argtypes = [c_ws, ctypes.POINTER(Tuning)]
    lib.tsne_set_tuning.restype = ctypes.c_int
    lib.tsne_tuning_lookup.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(Tuning)]
    lib.tsne_tuning_lookup.restype = ctypes.c_int
    lib.tsne_autotune.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Tuning)]
    lib.tsne_autotune.restype = ctypes.c_int
    lib.tsne_tuning_path.argtypes = [ctypes.c_char_p, ctypes.c_int]
    lib.tsne_tuning_path.restype = ctypes.c_int
    lib.tsne_attr_probe.argtypes = [c_ws, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.tsne_attr_probe.restype = ctypes.c_int
    lib.tsne_run_state.argtypes = [c_ws, ip, ip, dp, dp]
    lib.tsne_run_state.restype = ctypes.c_int
    lib.tsne_run_range.argtypes = [c_ws, ctypes.c_int, ctypes.c_int]
    lib.tsne_run_range.restype = ctypes.c_int
    lib.tsne_synchronize.argtypes = [c_ws]
    lib.tsne_synchronize.restype = ctypes.c_int
    lib.tsne_enable_timing.argtypes = [c_ws, ctypes.c_int]
    lib.tsne_enable_timing.restype = ctypes.c_int
    lib.tsne_attr_kernel_time.argtypes = [c_ws, dp, ip]
    lib.tsne_attr_kernel_time.restype = ctypes.c_int
    lib.tsne_attr_kernel_time_tiled.argtypes = [c_ws, dp, ip]
    lib.tsne_attr_kernel_time_tiled.restype = ctypes.c_int
    lib.tsne_joint_P_from_knn_distances.argtypes = [
        ctypes.POINTER(DTYPE), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
        ctypes.c_int, ctypes.c_int, ctypes.c_double]
    lib.tsne_joint_P_from_knn_distances.restype = c_sm
    lib.tsne_numpy_sum.argtypes = [ctypes.POINTER(DTYPE), ctypes.c_longlong]
    lib.tsne_numpy_sum.restype = ctypes.c_double
    lib.tsne_attr_tile_stats.argtypes = [c_ws, ip, ip, ip, ctypes.POINTER(ctypes.c_longlong), ip]
    lib.tsne_attr_tile_stats.restype = ctypes.c_int
    lib.tsne_knn_exact.argtypes = [
        f64c, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        np.ctypeslib.ndpointer(dtype=np.int32, flags='C_CONTIGUOUS'),
        np.ctypeslib.ndpointer(dtype=np.float32, flags='C_CONTIGUOUS'),
        ctypes.POINTER(KnnStats)]
    lib.tsne_knn_exact.restype = ctypes.c_int
    lib.tsne_joint_P_from_points.argtypes = [f64c, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double, ctypes.POINTER(KnnStats)]
    lib.tsne_joint_P_from_points.restype = c_sm
    i32p = ctypes.POINTER(ctypes.c_int)
    lib.tsne_quality_score.argtypes = [
        ctypes.POINTER(DTYPE), ctypes.c_int, ctypes.c_int, i32p, ctypes.c_int, ctypes.c_int,
        i32p, ctypes.c_int, i32p, i32p, ctypes.POINTER(DTYPE), i32p, ctypes.c_int,
        ctypes.POINTER(Quality)]
    lib.tsne_quality_score.restype = ctypes.c_int
    # job server
    c_srv = ctypes.c_void_p
    lp = ctypes.POINTER(ctypes.c_longlong)
    f64n = _optional_ndpointer(dtype=NP_DTYPE, flags='C_CONTIGUOUS')
    lib.tsne_job_footprint.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_longlong]
    lib.tsne_job_footprint.restype = ctypes.c_longlong
    lib.tsne_server_create.argtypes = [ctypes.c_int, ctypes.c_longlong]
    lib.tsne_server_create.restype = c_srv
    lib.tsne_server_submit.argtypes = [c_srv, ctypes.c_int, ctypes.c_int, c_pp, c_sm, f64n]
    lib.tsne_server_submit.restype = ctypes.c_int
    lib.tsne_server_set_h2_precision.argtypes = [c_srv, ctypes.c_int]
    lib.tsne_server_set_h2_precision.restype = ctypes.c_int
    lib.tsne_server_poll.argtypes = [c_srv, ctypes.c_int, ctypes.POINTER(JobInfo)]
    lib.tsne_server_poll.restype = ctypes.c_int
    lib.tsne_server_wait.argtypes = [c_srv, ctypes.c_int, f64n, ctypes.POINTER(JobInfo)]
    lib.tsne_server_wait.restype = ctypes.c_int
    lib.tsne_server_status.argtypes = [c_srv, ip, ip, lp]
    lib.tsne_server_status.restype = ctypes.c_int
    lib.tsne_server_destroy.argtypes = [c_srv]
    lib.tsne_server_destroy.restype = None


def scipy_csr_to_sparse_matrix(P_scipy):
    """Borrowing SparseMatrix view of a scipy CSR (reference _backend.py:784-839); the
    returned dict keeps the arrays alive."""
    from scipy.sparse import csr_matrix
    if not isinstance(P_scipy, csr_matrix):
        raise TypeError(f"Expected scipy CSR matrix, got {type(P_scipy)}")
    rp = np.ascontiguousarray(P_scipy.indptr, dtype=np.int32)
    ci = np.ascontiguousarray(P_scipy.indices, dtype=np.int32)
    va = np.ascontiguousarray(P_scipy.data, dtype=np.float64)
    m = SparseMatrix()
    m.n_rows, m.n_cols = P_scipy.shape
    m.nnz = P_scipy.nnz
    m.row_ptr = rp.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    m.col_idx = ci.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    m.values = va.ctypes.data_as(ctypes.POINTER(DTYPE))
    return m, {'row_ptr': rp, 'col_idx': ci, 'values': va}


def scipy_csr_to_c_sparse_matrix(P_scipy, lib):
    """Library-allocated copy of a scipy CSR (reference _backend.py:842-882), for the
    entry points that take ownership."""
    from scipy.sparse import csr_matrix
    if not isinstance(P_scipy, csr_matrix):
        raise TypeError(f"Expected scipy CSR matrix, got {type(P_scipy)}")
    n_rows, n_cols = P_scipy.shape
    nnz = P_scipy.nnz
    ptr = lib.sparse_matrix_create(n_rows, n_cols, nnz)
    if not ptr:
        raise MemoryError("Failed to allocate sparse matrix in C")
    m = ptr.contents
    rp = np.ascontiguousarray(P_scipy.indptr, dtype=np.int32)
    ci = np.ascontiguousarray(P_scipy.indices, dtype=np.int32)
    va = np.ascontiguousarray(P_scipy.data, dtype=np.float64)
    ctypes.memmove(m.row_ptr, rp.ctypes.data, (n_rows + 1) * 4)
    if nnz:
        ctypes.memmove(m.col_idx, ci.ctypes.data, nnz * 4)
        ctypes.memmove(m.values, va.ctypes.data, nnz * 8)
    return ptr


_backend: Optional[TSNEBackend] = None


def get_backend() -> TSNEBackend:
    """Process-wide backend (loads the library on first use)."""
    global _backend
    if _backend is None:
        _backend = TSNEBackend()
    return _backend

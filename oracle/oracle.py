"""ctypes wrapper of the CPU oracle (oracle/tsne_oracle.c, oracle/tsne_oracle_h2.c).

TEST INFRASTRUCTURE ONLY.  This module is the checker: tests/, the smoke() of
__graft_entry__ and the cpu_baseline leg of bench.py import it.  The product
library (tsne_h2pack + libtsne_h2pack.so) never imports or links it.

Arrays follow the reference's conventions: Y is (dim, n) dimension-major
("d*n+i", include/tsne_h2pack.h:107-117 of the reference), FP64, and P is a CSR
triple (row_ptr int32, col_idx int32, values float64).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_perplexity_rows.argtypes = [_dp, _ip, ctypes.c_int, ctypes.c_double, _dp,
                                             ctypes.c_void_p]
        L.oracle_perplexity_rows.restype = None
        L.oracle_attractive.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, _dp,
                                        ctypes.c_double, _dp]
        L.oracle_attractive.restype = None
        L.oracle_repulsive_exact.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_repulsive_exact.restype = ctypes.c_double
        L.oracle_repulsive_rows.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_double,
                                            _ip, ctypes.c_int, _dp]
        L.oracle_repulsive_rows.restype = None
        L.oracle_kl.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, _dp,
                                ctypes.c_double, ctypes.c_double]
        L.oracle_kl.restype = ctypes.c_double
        L.oracle_update.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp,
                                    ctypes.c_double, ctypes.c_double]
        L.oracle_update.restype = None
        L.oracle_is_rebuild_iter.argtypes = [ctypes.c_int] * 5
        L.oracle_is_rebuild_iter.restype = ctypes.c_int
        L.oracle_run.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, _dp, _dp, _dp,
                                 ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_run.restype = ctypes.c_double
        L.oracle_knn_exact.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip,
                                       np.ctypeslib.ndpointer(dtype=np.float32,
                                                              flags="C_CONTIGUOUS")]
        L.oracle_knn_exact.restype = None
        L.oracle_num_threads.argtypes = []
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_h2_build.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                      ctypes.c_double]
        L.oracle_h2_build.restype = ctypes.c_void_p
        L.oracle_h2_repulsive.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oracle_h2_repulsive.restype = ctypes.c_double
        L.oracle_h2_free.argtypes = [ctypes.c_void_p]
        L.oracle_h2_free.restype = None
        ip_ = ctypes.POINTER(ctypes.c_int)
        lp_ = ctypes.POINTER(ctypes.c_longlong)
        L.oracle_h2_stats.argtypes = [ctypes.c_void_p, ip_, ip_, lp_, lp_, ip_,
                                      ctypes.POINTER(ctypes.c_double)]
        L.oracle_h2_stats.restype = None
        L.oracle_run_h2.argtypes = L.oracle_run.argtypes + [ctypes.c_int, ctypes.c_double, ip_]
        L.oracle_run_h2.restype = ctypes.c_double
        _lib = L
    return _lib


def _csr(P):
    return (np.ascontiguousarray(P.indptr, dtype=np.int32),
            np.ascontiguousarray(P.indices, dtype=np.int32),
            np.ascontiguousarray(P.data, dtype=np.float64))


def perplexity_rows(dist, indptr, perplexity):
    """Conditional P for CSR k-NN distances (L2, not squared); returns (p, steps)."""
    dist = np.ascontiguousarray(dist, dtype=np.float64)
    indptr = np.ascontiguousarray(indptr, dtype=np.int32)
    n = indptr.size - 1
    p = np.zeros_like(dist)
    steps = np.zeros(n, dtype=np.int32)
    lib().oracle_perplexity_rows(dist, indptr, n, float(perplexity), p,
                                 steps.ctypes.data_as(ctypes.c_void_p))
    return p, steps


def attractive(P, Y, exag):
    """4*exag*sum_j p_ij w_ij (y_i - y_j); Y is (dim, n). Returns (dim, n)."""
    dim, n = Y.shape
    rp, ci, va = _csr(P)
    Yc = np.ascontiguousarray(Y, dtype=np.float64)
    g = np.zeros((dim, n), dtype=np.float64)
    lib().oracle_attractive(n, dim, rp, ci, va, Yc, float(exag), g)
    return g


def repulsive_exact(Y):
    """(rep, Z, s, v): the reference's direct O(N^2) repulsive branch in FP64."""
    dim, n = Y.shape
    Yc = np.ascontiguousarray(Y, dtype=np.float64)
    rep = np.zeros((dim, n))
    s = np.zeros(n)
    v = np.zeros((dim, n))
    Z = lib().oracle_repulsive_exact(n, dim, Yc, rep.ctypes.data_as(ctypes.c_void_p),
                                     s.ctypes.data_as(ctypes.c_void_p),
                                     v.ctypes.data_as(ctypes.c_void_p))
    return rep, Z, s, v


def repulsive_rows(Y, Z, rows):
    """Exact repulsive force (4/Z) sum_j K_ij (y_i-y_j) on selected rows: (len(rows), dim)."""
    dim, n = Y.shape
    Yc = np.ascontiguousarray(Y, dtype=np.float64)
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    out = np.zeros((rows.size, dim))
    lib().oracle_repulsive_rows(n, dim, Yc, float(Z), rows, rows.size, out)
    return out


def kl(P, Y, Z, exag=1.0):
    dim, n = Y.shape
    rp, ci, va = _csr(P)
    return lib().oracle_kl(n, dim, rp, ci, va, np.ascontiguousarray(Y, dtype=np.float64),
                           float(Z), float(exag))


def update(Y, uY, gains, grad, momentum, lr):
    """In-place gains/momentum/position update plus centring (all (dim, n) float64)."""
    dim, n = Y.shape
    lib().oracle_update(n, dim, Y, uY, gains, grad, float(momentum), float(lr))


def is_rebuild_iter(it, freq=3, freq_early=10, early_exag_iter=250, reuse=True):
    return bool(lib().oracle_is_rebuild_iter(it, freq, freq_early, early_exag_iter, int(reuse)))


def run(P, Y0, n_iter=1000, lr=None, mom0=0.5, mom1=0.8, mom_switch=250, exag=12.0,
        exag_iter=250, freq=3, freq_early=10):
    """The reference's tsne_run loop with exact repulsion; returns (Y (dim,n), KL)."""
    dim, n = Y0.shape
    if lr is None:
        lr = max(n / 12.0, 200.0)
    rp, ci, va = _csr(P)
    Y = np.array(Y0, dtype=np.float64, order="C")
    uY = np.zeros_like(Y)
    gains = np.ones_like(Y)
    klv = lib().oracle_run(n, dim, rp, ci, va, Y, uY, gains, int(n_iter), float(lr),
                           float(mom0), float(mom1), int(mom_switch), float(exag),
                           int(exag_iter), int(freq), int(max(freq_early, freq)), 0)
    return Y, klv


def repulsive_h2(Y, h2_tol=1e-2, leaf=100):
    """The repulsive force from a fresh CPU H2 build and product (ID tolerance
    0.5 h2_tol, as the reference's); returns (rep (dim, n), Z, stats dict)."""
    Y = np.ascontiguousarray(Y, dtype=np.float64)
    dim, n = Y.shape
    L = lib()
    h = L.oracle_h2_build(n, dim, Y, int(leaf), 0.5 * float(h2_tol))
    if not h:
        raise ValueError("oracle_h2_build: dim must be 2 or 3 and n >= 2")
    try:
        rep = np.zeros_like(Y)
        Z = L.oracle_h2_repulsive(h, Y, rep)
        nn, nl, rm = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        nf, nr = ctypes.c_longlong(), ctypes.c_longlong()
        rmean = ctypes.c_double()
        L.oracle_h2_stats(h, ctypes.byref(nn), ctypes.byref(nl), ctypes.byref(nf),
                          ctypes.byref(nr), ctypes.byref(rm), ctypes.byref(rmean))
        st = dict(n_nodes=nn.value, n_leaves=nl.value, n_far=nf.value, n_near=nr.value,
                  rank_max=rm.value, rank_mean=rmean.value)
    finally:
        L.oracle_h2_free(h)
    return rep, Z, st


def run_h2(P, Y0, n_iter=1000, lr=None, mom0=0.5, mom1=0.8, mom_switch=250, exag=12.0,
           exag_iter=250, freq=3, freq_early=10, it_begin=0, h2_tol=1e-2, leaf=100,
           uY=None, gains=None):
    """The reference's tsne_run loop with the CPU H2 product on rebuild iterations and the
    force reused between them; returns (Y (dim, n), KL, builds).  it_begin > 0 continues a
    run from the state (Y0, uY, gains) at that iteration."""
    dim, n = Y0.shape
    if lr is None:
        lr = max(n / 12.0, 200.0)
    rp, ci, va = _csr(P)
    Y = np.array(Y0, dtype=np.float64, order="C")
    uY = np.zeros_like(Y) if uY is None else np.array(uY, dtype=np.float64, order="C")
    gains = np.ones_like(Y) if gains is None else np.array(gains, dtype=np.float64, order="C")
    nb = ctypes.c_int()
    klv = lib().oracle_run_h2(n, dim, rp, ci, va, Y, uY, gains, int(n_iter), float(lr),
                              float(mom0), float(mom1), int(mom_switch), float(exag),
                              int(exag_iter), int(freq), int(max(freq_early, freq)),
                              int(it_begin), int(leaf), float(h2_tol), ctypes.byref(nb))
    return Y, klv, nb.value


def num_threads():
    return lib().oracle_num_threads()


def knn_exact(X, k):
    """Exact k-NN graph rows (indices (n, k) int32, float32 distances), sklearn's
    kneighbors_graph(mode='distance') order: increasing distance, ties by index."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, d = X.shape
    idx = np.empty((n, k), dtype=np.int32)
    dist = np.empty((n, k), dtype=np.float32)
    lib().oracle_knn_exact(X, n, d, k, idx, dist)
    return idx, dist


def quality_score(Y, hd, queries, P, anchors, k=10):
    """The reference's common-footing score (benchmarks/quality.py:56-81) with its random
    draws passed in: kNN-preservation@k against the high-dimensional neighbour rows hd
    (distance order), sklearn KDTree in the embedding (k+1 nearest, first column dropped);
    KL = sum p log p + sum p log1p(d^2) + log Z with Z = n * mean anchor row sum.
    Returns dict(kl, knn_preservation, Z, span)."""
    from sklearn.neighbors import KDTree
    Y = np.asarray(Y, dtype=np.float64)
    n = Y.shape[0]
    ld = KDTree(Y).query(Y[queries], k=k + 1, return_distance=False)[:, 1:]
    preserved = np.mean([len(set(hd[i, :k]) & set(ld[r])) for r, i in enumerate(queries)]) / k
    indptr, indices, p = P.indptr, P.indices, P.data
    rows = np.repeat(np.arange(n), np.diff(indptr))
    d2 = np.sum((Y[rows] - Y[indices]) ** 2, axis=1)
    row_sums = np.empty(len(anchors))
    for a, i in enumerate(anchors):
        w = 1.0 / (1.0 + np.sum((Y - Y[i]) ** 2, axis=1))
        w[i] = 0.0
        row_sums[a] = w.sum()
    Z = n * row_sums.mean()
    kl = float(np.sum(p * np.log(np.maximum(p, 1e-300))) + np.sum(p * np.log1p(d2)) + np.log(Z))
    return dict(kl=kl, knn_preservation=float(preserved), Z=float(Z),
                span=float(np.ptp(Y, axis=0).max()))

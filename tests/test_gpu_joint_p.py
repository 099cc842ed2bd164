"""Joint P on the device (tsne_joint_P_from_knn_distances) against the reference's host
pipeline: the library's conditional P, then scipy's P = Pc + Pc.T and P.data /= P.sum()
(reference tsne_h2pack/_backend.py:695-718).  Bar: bit-exact -- same row pointers,
columns and float64 values -- including rows whose conditional probabilities underflow
to exactly 0 (scipy drops entries that sum to 0), empty rows, and N above one 8192-row
block of numpy's sum.
"""
import sys

import numpy as np
import pytest
from scipy.sparse import csr_matrix

from datasets import gaussian_mixture, knn_graph

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _host(backend, data, indices, indptr, perplexity):
    Pc = backend.conditional_P(data, indices, indptr, perplexity)
    return backend.symmetrize(Pc)


def _check(backend, G, perplexity):
    data, indices, indptr = backend.prepare_knn_csr(G)
    ref = _host(backend, data, indices, indptr, perplexity)
    dev = backend.joint_P_device(data, indices, indptr, perplexity)
    _say(f"  n={G.shape[0]} nnz conditional {indices.size} joint {ref.nnz} / {dev.nnz}")
    assert np.array_equal(ref.indptr, dev.indptr)
    assert np.array_equal(ref.indices, dev.indices)
    assert np.array_equal(ref.data.view(np.uint64), dev.data.view(np.uint64))
    return ref


@pytest.mark.parametrize("n,k,perp", [(3000, 30, 10.0), (20000, 90, 30.0)])
def test_joint_P_bit_exact(backend, n, k, perp):
    _say(f"joint P n={n} k={k}")
    X, _ = gaussian_mixture(n, dim=20, n_clusters=8, seed=n)
    _check(backend, knn_graph(X, k), perp)


def test_joint_P_underflow_and_empty_rows(backend):
    """Rows with distances spread over 1e3 (far neighbours' probabilities underflow to
    exactly 0, so symmetric pairs of zeros are dropped), some rows empty, ragged rows."""
    _say("joint P underflow / empty rows")
    rng = np.random.default_rng(5)
    n = 4000
    rows, cols, vals = [], [], []
    for i in range(n):
        if i % 97 == 0:
            continue  # empty row
        k = int(rng.integers(1, 40))
        c = np.sort(rng.choice(np.delete(np.arange(n), i), size=k, replace=False))
        d = np.sort(rng.uniform(0.0, 1.0, size=k))
        if i % 3 == 0:
            d = d * 1e3  # far neighbours: exp(-beta d^2) underflows
        rows += [i] * k
        cols += list(c)
        vals += list(d)
    G = csr_matrix((np.array(vals), (np.array(rows), np.array(cols))), shape=(n, n))
    G.sort_indices()
    P = _check(backend, G, 5.0)
    Pc = backend.conditional_P(*backend.prepare_knn_csr(G), 5.0)
    assert (Pc.data == 0).any()  # the case is exercised


def test_joint_P_feeds_the_optimisation(backend):
    """The mirror's TSNE path (joint P on the device) gives the same embedding as the
    host-symmetrised P for a short run."""
    import ctypes
    _say("joint P in the optimisation")
    X, _ = gaussian_mixture(2500, dim=16, n_clusters=5, seed=2)
    G = knn_graph(X, 30)
    data, indices, indptr = backend.prepare_knn_csr(G)
    P_host = _host(backend, data, indices, indptr, 10.0)
    P_dev = backend.joint_P_device(data, indices, indptr, 10.0)
    Y0 = np.random.default_rng(3).normal(scale=1e-2, size=(2500, 2))
    out = []
    for P in (P_host, P_dev):
        p = backend.create_default_params()
        p.verbose = 0
        p.n_iter = 50
        p.learning_rate = 208.0
        ws = backend.init_workspace(P, 2, p, init_Y=Y0)
        try:
            assert backend.lib.tsne_run(ws) == 0
            Y = np.zeros((2, 2500))
            backend.lib.tsne_get_embedding(ws, Y)
            out.append(Y)
        finally:
            backend.lib.tsne_destroy(ws)
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("n,dim,k,perp", [(3000, 20, 30, 10.0), (12000, 50, 90, 30.0)])
def test_joint_P_from_points_equals_host_pipeline(backend, n, dim, k, perp):
    """tsne_joint_P_from_points (k-NN graph, neighbour indexing and joint P on the device)
    equals the reference's flow on the host -- sklearn's exact graph (the ExactKNN call),
    prepare_knn_csr, conditional P, scipy symmetrisation -- bit for bit."""
    _say(f"joint P from points n={n} D={dim} k={k}")
    X, _ = gaussian_mixture(n, dim=dim, n_clusters=5, seed=n)
    ref = _check(backend, knn_graph(X, k), perp)
    P, st = backend.joint_P_from_points(X, k, perp)
    _say(f"  device P nnz {P.nnz}, k-NN uncertain rows {st.uncertain_rows}")
    assert np.array_equal(ref.indptr, P.indptr)
    assert np.array_equal(ref.indices, P.indices)
    assert np.array_equal(ref.data.view(np.uint64), P.data.view(np.uint64))

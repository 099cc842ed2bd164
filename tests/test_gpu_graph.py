"""Graph mode (tsne_set_graph_mode): the force-reusing iterations replayed from captured
HIP graphs give the eager launches' results bit for bit.  Each graph is replayed hundreds of
times between eager rebuild iterations, renumberings (which drop the graphs) and host
synchronisations; the embedding is compared at several points of the run."""
import ctypes
import sys

import numpy as np
import pytest

from datasets import gaussian_mixture, init_normal, joint_P_oracle, knn_graph

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def problem():
    X, _ = gaussian_mixture(6000, dim=20, n_clusters=12, seed=3)
    return joint_P_oracle(knn_graph(X, 45), 15.0)


def _checkpoints(backend, P, dim, graph, cuts):
    n = P.shape[0]
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = 1000
    p.h2_tol = 1e-2 if dim == 2 else 3e-2
    p.h2_rebuild_freq = 3
    p.h2_rebuild_freq_early = 10
    p.learning_rate = max(n / 12.0, 200.0)
    ws = backend.init_workspace(P, dim, p, init_Y=init_normal(n, dim, 31).T.copy())
    out = []
    try:
        assert backend.lib.tsne_set_graph_mode(ws, 1 if graph else 0) == 0
        lo = 0
        for hi in cuts:
            assert backend.lib.tsne_run_range(ws, lo, hi) == 0
            Y = np.zeros((dim, n))
            backend.lib.tsne_get_embedding(ws, Y)
            out.append(Y)
            lo = hi
        kl = backend.lib.compute_kl_divergence(ws)
        launches = backend.lib.tsne_graph_launches(ws)
    finally:
        backend.lib.tsne_destroy(ws)
    return out, kl, launches


@pytest.mark.parametrize("dim", [2, 3])
def test_graph_mode_bit_identical(backend, problem, dim):
    cuts = [40, 250, 610, 1000]
    eager, kl_e, n_e = _checkpoints(backend, problem, dim, False, cuts)
    graph, kl_g, n_g = _checkpoints(backend, problem, dim, True, cuts)
    _say(f"{dim}D: graph launches {n_g} (eager run {n_e}), KL {kl_e:.8f} / {kl_g:.8f}")
    assert n_e == 0 and n_g > 600
    for a, b in zip(eager, graph):
        assert np.array_equal(a, b)
    assert kl_e == kl_g

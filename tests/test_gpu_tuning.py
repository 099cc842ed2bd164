"""Launch shapes and their autotuner (tsne_set_tuning, tsne_autotune, tsne_tuning_lookup).

The tuned choices -- the second stream of the attractive pass on rebuild iterations, tree
levels per host check, graph replay, and which workgroup shape runs each node's ID -- must
leave every result unchanged, bit for bit; the autotuner checks that on each of its trials,
and these tests check it on whole runs with renumbering.  So must the tree levels' grid
hints from the previous build (TSNE_LEVEL_HINTS, off in one run and on in the other)."""
import os

import numpy as np
import pytest

from datasets import gaussian_mixture, init_normal, joint_P_oracle, knn_graph

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _run(backend, P, dim, Y0, shapes, n_iter=300, hints=None):
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = n_iter
    p.h2_tol = 1e-2 if dim == 2 else 3e-2
    p.h2_rebuild_freq = 3
    p.learning_rate = max(P.shape[0] / 12.0, 200.0)
    old = os.environ.get("TSNE_LEVEL_HINTS")
    if hints is not None:
        os.environ["TSNE_LEVEL_HINTS"] = hints  # read when the workspace's H2 is made
    try:
        ws = backend.init_workspace(P, dim, p, init_Y=Y0.T.copy())
    finally:
        if old is None:
            os.environ.pop("TSNE_LEVEL_HINTS", None)
        else:
            os.environ["TSNE_LEVEL_HINTS"] = old
    try:
        backend.set_tuning(ws, **shapes)
        assert backend.lib.tsne_run(ws) == 0
        Y = np.zeros((dim, P.shape[0]))
        backend.lib.tsne_get_embedding(ws, Y)
        return Y, backend.lib.compute_kl_divergence(ws)
    finally:
        backend.lib.tsne_destroy(ws)


@pytest.mark.parametrize("dim,n", [(2, 6000), (3, 4000)])
def test_launch_shapes_do_not_change_results(backend, dim, n):
    """Two opposite launch-shape sets over 300 iterations (rebuilds, three renumberings,
    graph replay in one of them, level grid hints in the other): the same embedding and KL,
    bit for bit."""
    _say(f"launch shapes {dim}D n={n}")
    X, _ = gaussian_mixture(n, dim=12, n_clusters=6, seed=31)
    P = joint_P_oracle(knn_graph(X, 60), 20.0)
    Y0 = init_normal(n, dim, 32)
    a = dict(overlap=0, level_batch=1, graph=0, one_class_level=0)
    b = dict(overlap=1, level_batch=8, graph=1, one_class_level=1_000_000)
    Ya, kla = _run(backend, P, dim, Y0, a, hints="0")
    Yb, klb = _run(backend, P, dim, Y0, b, hints="1")
    _say(f"  KL {kla:.10f} / {klb:.10f}, max |dY| {np.abs(Ya - Yb).max():.3e}")
    assert np.isfinite(Ya).all()
    assert np.array_equal(Ya, Yb) and kla == klb


def test_autotune_stores_and_applies_its_choice(backend):
    """The autotuner returns valid shapes, persists them in the device's table file, and a
    later lookup for the same size bucket returns them."""
    _say("autotune 2D n=3000")
    best = backend.autotune(2, 3000, n_iter=100)
    _say(f"  chosen {best}")
    assert best["overlap"] in (0, 1) and best["graph"] in (0, 1)
    assert 1 <= best["level_batch"] <= 8 and best["one_class_level"] >= 0 and best["ms"] > 0
    shapes, found = backend.tuning_lookup(2, 3000)
    assert found and shapes == best
    assert backend.tuning_lookup(2, 2900)[1]  # same bucket round(log2 N) = 12
    path = backend.tuning_path()
    assert path
    lines = open(path).read().splitlines()
    assert lines[0].startswith("tsne_h2pack tuning v1 ")
    assert any(l.split()[:2] == ["2", "12"] for l in lines[1:])

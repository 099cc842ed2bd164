"""Row tiles of the attractive pass (LDS-staged neighbours, 16-bit local indices) against
the per-row kernel that gathers every neighbour from HBM, through the C ABI.

The tiles exist once the points are renumbered along the tree (first at iteration 30),
so a run of 31 iterations differs from a per-row run only in its last attractive pass.
Both kernels do the same FP32 arithmetic per nonzero; only the order of the FP32
partial sums differs, so the embeddings agree to ~1e-7 relative after one tiled pass
(a wrong local index or a dropped nonzero gives an O(1) difference).  Over more
iterations the gains rule (a sign test on uY * grad) turns such differences into
macroscopic ones, so longer runs are compared through their KL divergence.  The reference's attractive force (src/tsne_h2pack.c:199-356) is
pinned against the FP64 oracle by test_gpu_parity.py on the per-row kernel.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle
from datasets import gaussian_mixture, init_normal, joint_P_oracle, knn_graph

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def problem2d():
    X, _ = gaussian_mixture(6000, dim=20, n_clusters=12, seed=3)
    return joint_P_oracle(knn_graph(X, 45), 15.0)


@pytest.fixture(scope="module")
def problem3d():
    X, _ = gaussian_mixture(5000, dim=20, n_clusters=8, seed=4)
    return joint_P_oracle(knn_graph(X, 60), 20.0)


def _run(backend, P, dim, Y0, n_iter, tile_rows):
    old = os.environ.get("TSNE_ATTR_TILE_ROWS")
    os.environ["TSNE_ATTR_TILE_ROWS"] = str(tile_rows)
    try:
        p = backend.create_default_params()
        p.verbose = 0
        p.n_iter = n_iter
        p.h2_tol = 1e-2 if dim == 2 else 3e-2
        p.h2_rebuild_freq = 3
        p.h2_rebuild_freq_early = 10
        p.learning_rate = max(P.shape[0] / 12.0, 200.0)
        ws = backend.init_workspace(P, dim, p, init_Y=Y0.T.copy())
    finally:
        if old is None:
            del os.environ["TSNE_ATTR_TILE_ROWS"]
        else:
            os.environ["TSNE_ATTR_TILE_ROWS"] = old
    try:
        assert backend.lib.tsne_run(ws) == 0
        Y = np.zeros((dim, P.shape[0]))
        backend.lib.tsne_get_embedding(ws, Y)
        st = [ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong(), ctypes.c_int()]
        assert backend.lib.tsne_attr_tile_stats(ws, *[ctypes.byref(v) for v in st]) == 0
        stats = dict(zip(["rows", "ntiles", "umax", "ucount", "n_global"], [v.value for v in st]))
    finally:
        backend.lib.tsne_destroy(ws)
    return Y, stats


def _compare(backend, P, dim, seed, tile_rows, n_iter=31):
    Y0 = init_normal(P.shape[0], dim, seed) * 1e2
    ref, st0 = _run(backend, P, dim, Y0, n_iter, 0)
    til, st = _run(backend, P, dim, Y0, n_iter, tile_rows)
    rel = np.abs(ref - til).max() / np.abs(ref).max()
    _say(f"  d={dim} rows={tile_rows} stats {st}  rel diff {rel:.2e}")
    assert st0["ntiles"] == 0
    assert np.isfinite(til).all()
    return rel, st


def test_tiles_match_row_kernel_2d(backend, problem2d):
    _say("tiles 2d")
    n = problem2d.shape[0]
    rel, st = _compare(backend, problem2d, 2, 31, 256)
    assert st["rows"] == 256 and st["ntiles"] == (n + 255) // 256
    assert st["n_global"] == 0 and 0 < st["umax"] <= 8192
    assert st["ucount"] >= n  # every point is some tile's neighbour or its own row's
    assert rel < 1e-5


def test_tiles_match_row_kernel_3d(backend, problem3d):
    _say("tiles 3d")
    rel, st = _compare(backend, problem3d, 3, 32, 128)
    assert st["ntiles"] == (problem3d.shape[0] + 127) // 128
    assert rel < 1e-5


def test_tiles_over_lds_capacity_gather_from_hbm(backend, problem3d):
    """Tiles of 2048 rows in 3D reference more distinct points than the 4096 records the
    LDS holds: those tiles take the HBM-gather branch of the same kernel."""
    _say("tiles 3d over capacity")
    rel, st = _compare(backend, problem3d, 3, 33, 2048)
    assert st["n_global"] >= 1
    assert rel < 1e-5


def test_tiled_full_run_kl_matches_row_kernel(backend, problem2d):
    """1000 iterations (six renumberings, each rebuilding the tiles): final KL of the tiled
    run within 0.5% of the per-row run."""
    _say("tiles full run")
    Y0 = init_normal(problem2d.shape[0], 2, 34)
    kls = []
    for rows in (0, 256):
        old = os.environ.get("TSNE_ATTR_TILE_ROWS")
        os.environ["TSNE_ATTR_TILE_ROWS"] = str(rows)
        try:
            p = backend.create_default_params()
            p.verbose = 0
            p.h2_tol = 1e-2
            p.h2_rebuild_freq = 3
            p.learning_rate = max(problem2d.shape[0] / 12.0, 200.0)
            ws = backend.init_workspace(problem2d, 2, p, init_Y=Y0.T.copy())
        finally:
            if old is None:
                del os.environ["TSNE_ATTR_TILE_ROWS"]
            else:
                os.environ["TSNE_ATTR_TILE_ROWS"] = old
        try:
            assert backend.lib.tsne_run(ws) == 0
            kls.append(backend.lib.compute_kl_divergence(ws))
        finally:
            backend.lib.tsne_destroy(ws)
    _say(f"  KL per-row {kls[0]:.6f} tiled {kls[1]:.6f}")
    assert abs(kls[1] - kls[0]) <= 5e-3 * abs(kls[0])

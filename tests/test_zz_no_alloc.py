"""tsne_exact_repulsive_rows_noalloc (the entry point that allocates nothing) against
tsne_exact_repulsive_rows, through the C ABI: two calls in a row, the second on a
smaller input (fewer points and fewer rows), each bit-identical to the allocating entry
point on the same workspace and rows; a request beyond the scratch maximum is refused."""
import ctypes
import sys

import numpy as np
import pytest

from datasets import gaussian_mixture, init_normal, joint_P_oracle, knn_graph

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ws(backend, n, dim, seed):
    X, _ = gaussian_mixture(n, dim=16, n_clusters=6, seed=seed)
    P = joint_P_oracle(knn_graph(X, 30), 10.0)
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = 1000
    p.h2_tol = 1e-2 if dim == 2 else 3e-2
    p.learning_rate = max(n / 12.0, 200.0)
    ws = backend.init_workspace(P, dim, p, init_Y=init_normal(n, dim, seed).T.copy())
    # a few iterations past the first renumbering (iteration 30), so the workspace's
    # current point order differs from the caller's
    assert backend.lib.tsne_run_range(ws, 0, 40) == 0
    return ws


@pytest.mark.parametrize("dim", [2, 3])
def test_noalloc_rows_equal_allocating_entry_point(backend, dim):
    lib = backend.lib
    cases = [(6000, 1500, 11), (2500, 300, 12)]  # the second call: a smaller input
    for n, m, seed in cases:
        _say(f"noalloc exact rows: {dim}D n={n} rows={m}")
        ws = _ws(backend, n, dim, seed)
        try:
            rows = np.sort(np.random.default_rng(seed).choice(n, size=m, replace=False)).astype(np.int32)
            ref = np.zeros((m, dim))
            got = np.full((m, dim), np.nan)
            Zr, Zg = ctypes.c_double(), ctypes.c_double()
            assert lib.tsne_exact_repulsive_rows(ws, rows, m, ref, ctypes.byref(Zr)) == 0
            assert lib.tsne_exact_repulsive_rows_noalloc(ws, rows, m, got, ctypes.byref(Zg)) == 0
        finally:
            lib.tsne_destroy(ws)
        assert Zg.value == Zr.value
        assert np.array_equal(got, ref)


def test_noalloc_rejects_requests_beyond_its_scratch(backend, capfd):
    lib = backend.lib
    n = 3000
    ws = _ws(backend, n, 2, 13)
    try:
        # 48 MiB of scratch: 2.5 million rows need 2.5e6 * 32 bytes = 80 MB
        m = 2_500_000
        rows = np.zeros(m, dtype=np.int32)
        out = np.zeros((m, 2))
        assert lib.tsne_exact_repulsive_rows_noalloc(ws, rows, m, out, None) == -1
        assert "scratch" in capfd.readouterr().err
        # and the entry point still works afterwards
        rows = np.arange(0, n, 7, dtype=np.int32)
        ref = np.zeros((rows.size, 2))
        got = np.zeros((rows.size, 2))
        assert lib.tsne_exact_repulsive_rows(ws, rows, rows.size, ref, None) == 0
        assert lib.tsne_exact_repulsive_rows_noalloc(ws, rows, rows.size, got, None) == 0
        assert np.array_equal(got, ref)
    finally:
        lib.tsne_destroy(ws)

"""The rebuild's own radix sort (radix.hip) against the library sort it replaced: the
Morton-key sort and the pair-list sorts are stable sorts over the same bits, so the tree,
the interaction lists and every H2 product must be bit-identical whichever sorts them.
TSNE_SORT=rocprim selects the library sorts (rocprim onesweep / default) at run time."""
import ctypes
import os
import sys

import numpy as np
import pytest

from datasets import gaussian_mixture
from tsne_h2pack._backend import H2Stats

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _gradient(backend, ws, dim, n, sort):
    old = os.environ.get("TSNE_SORT")
    os.environ["TSNE_SORT"] = sort
    try:
        g = np.zeros((dim, n))
        Z = ctypes.c_double()
        assert backend.lib.tsne_compute_gradient(ws, 1.0, 0, g, ctypes.byref(Z)) == 0
        st = H2Stats()
        assert backend.lib.tsne_h2_stats(ws, ctypes.byref(st)) == 0
    finally:
        if old is None:
            del os.environ["TSNE_SORT"]
        else:
            os.environ["TSNE_SORT"] = old
    return g, Z.value, {f: getattr(st, f) for f, _ in H2Stats._fields_}


@pytest.mark.parametrize("n,dim", [(200_000, 2), (120_000, 3), (40_000, 2)])
def test_own_sort_equals_library_sort(backend, n, dim):
    X, _ = gaussian_mixture(n, dim=20, n_clusters=25, seed=n % 97, dirichlet=True,
                            std_range=(0.5, 2.0))
    P, _ = backend.joint_P_from_points(X, 30, 10.0)
    del X
    rng = np.random.default_rng(dim)
    cent = rng.normal(scale=40.0, size=(25, dim))
    Y0 = cent[rng.integers(0, 25, n)] + rng.normal(size=(n, dim)) * rng.uniform(0.2, 3.0, (n, 1))
    p = backend.create_default_params()
    p.verbose = 0
    p.h2_tol = 1e-2 if dim == 2 else 3e-2
    ws = backend.init_workspace(P, dim, p, init_Y=Y0)
    try:
        g1, Z1, s1 = _gradient(backend, ws, dim, n, "own")
        g2, Z2, s2 = _gradient(backend, ws, dim, n, "rocprim")
        g3, Z3, s3 = _gradient(backend, ws, dim, n, "own")
    finally:
        backend.lib.tsne_destroy(ws)
    _say(f"n={n} {dim}D: nodes {s1['n_nodes']} near pairs {s1['n_near_pairs']} "
         f"far pairs {s1['n_adm_pairs']}")
    assert s1 == s2 == s3
    assert Z1 == Z2 == Z3
    assert np.array_equal(g1, g2) and np.array_equal(g1, g3)

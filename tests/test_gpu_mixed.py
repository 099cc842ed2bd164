"""Mixed-precision H2 path (coupling on MFMA with BF16 operands and FP32 accumulation,
FP32 fallback for nodes whose rounding-error estimate is too large; the ID's pivoted
Gram-Schmidt in FP32 at ID tolerances >= 1e-3, as here) against the FP64 exact gradient
and against the plain path, through the C ABI.

Bar: the same h2_tol bound as the plain path (gradient error <= 1e-2 in 2D, 3e-2 in 3D,
all rows, exaggeration 1 and 12), the mixed repulsion within h2_tol / 4 of the plain
one, bit-identical repeats, and a full run whose final KL is within 1% of the plain run.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from datasets import gaussian_mixture, init_normal, joint_P_oracle, knn_graph
from tsne_h2pack._backend import H2Stats

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def problem():
    X, _ = gaussian_mixture(6000, dim=20, n_clusters=12, seed=3)
    return joint_P_oracle(knn_graph(X, 45), 15.0)


def _embedding(n, dim, scale, seed=0):
    rng = np.random.default_rng(seed)
    cent = rng.normal(scale=scale, size=(30, dim))
    lab = rng.integers(0, 30, n)
    Y = cent[lab] + rng.normal(size=(n, dim)) * rng.uniform(0.5, 2.0, 30)[lab, None] * scale / 20
    return Y.T.copy()


def _ws(backend, P, Y0, dim, tol, mixed):
    p = backend.create_default_params()
    p.verbose = 0
    p.h2_tol = tol
    p.h2_rebuild_freq = 3
    ws = backend.init_workspace(P, dim, p, init_Y=Y0.T.copy())
    assert backend.lib.tsne_set_h2_precision(ws, 1 if mixed else 0) == 0
    return ws


def _grad(backend, ws, dim, n, exag):
    g = np.zeros((dim, n))
    Z = ctypes.c_double()
    assert backend.lib.tsne_compute_gradient(ws, exag, 0, g, ctypes.byref(Z)) == 0
    return g, Z.value


def _check(backend, P, dim, scale, tol, seed):
    n = P.shape[0]
    Y0 = _embedding(n, dim, scale, seed)
    rep64, Z64, _, _ = oracle.repulsive_exact(Y0)
    reps = {}
    for mixed in (0, 1):
        ws = _ws(backend, P, Y0, dim, tol, mixed)
        try:
            for exag in (1.0, 12.0):
                g, Z = _grad(backend, ws, dim, n, exag)
                g64 = oracle.attractive(P, Y0, exag) - rep64
                err = np.linalg.norm(g - g64) / np.linalg.norm(g64)
                assert err <= tol, (mixed, exag, err)
                assert abs(Z / Z64 - 1) <= tol
            g0, _ = _grad(backend, ws, dim, n, 0.0)
            reps[mixed] = -g0
            st = H2Stats()
            assert backend.lib.tsne_h2_stats(ws, ctypes.byref(st)) == 0
            if mixed:
                fb = st.mixed_fallback_nodes
                g2, _ = _grad(backend, ws, dim, n, 0.0)
                assert np.array_equal(g0, g2)  # repeatable bit for bit
        finally:
            backend.lib.tsne_destroy(ws)
    rep_err = {m: np.linalg.norm(reps[m] - rep64) / np.linalg.norm(rep64) for m in (0, 1)}
    diff = np.linalg.norm(reps[1] - reps[0]) / np.linalg.norm(reps[0])
    _say(f"  d={dim} scale={scale}: rep err plain {rep_err[0]:.2e} mixed {rep_err[1]:.2e} "
         f"mixed-vs-plain {diff:.2e}; fallback nodes {fb} of {st.n_nodes}")
    assert diff <= tol / 4
    return fb, st.n_nodes


@pytest.mark.parametrize("scale", [0.05, 2.0, 8.0, 40.0, 200.0])
def test_mixed_gradient_within_tol_2d(backend, problem, scale):
    _say(f"mixed 2d scale={scale}")
    _check(backend, problem, 2, scale, 1e-2, 0)


@pytest.mark.parametrize("scale", [0.05, 3.0, 30.0])
def test_mixed_gradient_within_tol_3d(backend, problem, scale):
    _say(f"mixed 3d scale={scale}")
    _check(backend, problem, 3, scale, 3e-2, 4)


def test_mixed_full_run_kl_within_one_percent(backend, problem):
    """1000 iterations with the mixed product at every rebuild: final KL within 1% of
    the plain run from the same initial embedding."""
    _say("mixed full run")
    n = problem.shape[0]
    Y0 = init_normal(n, 2, 41)
    kls = []
    for mixed in (0, 1):
        p = backend.create_default_params()
        p.verbose = 0
        p.h2_tol = 1e-2
        p.h2_rebuild_freq = 3
        p.learning_rate = max(n / 12.0, 200.0)
        ws = backend.init_workspace(problem, 2, p, init_Y=Y0.T.copy())
        try:
            assert backend.lib.tsne_set_h2_precision(ws, mixed) == 0
            assert backend.lib.tsne_run(ws) == 0
            kls.append(backend.lib.compute_kl_divergence(ws))
        finally:
            backend.lib.tsne_destroy(ws)
    _say(f"  KL plain {kls[0]:.6f} mixed {kls[1]:.6f}")
    assert abs(kls[1] - kls[0]) <= 1e-2 * abs(kls[0])

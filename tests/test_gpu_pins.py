"""Direct pins of the schedule, the KL, the FP64 check kernel, the proxy cache and the
rank cap, on the GPU through the C ABI.

* Schedule: a 1000-iteration run at N >= 2000 makes 292 rebuilds inside the loop and
  renumbers the points 6 times (reference src/tsne_h2pack.c:1706,1765-1776).
* KL at a fixed state: with the embedding set and only the last iteration run, the
  library's KL equals the oracle's FP64 KL of the same P, Y and Z (1e-10 relative).
* The FP64 all-pairs kernel that the large-N tests compare against, against the CPU
  oracle at N = 100,000: its Z and 400 sampled rows of the exact repulsive force, with
  the oracle's own Z (1e-10).
* Proxy cache (reference tests/test_optimizations.py:114-124): an H2 gradient computed
  with the proxy table generated in the process, read back from the cache file, and with
  the cache off, is bit-identical.
* Rank cap: at h2_tol = 1e-3 (the C default) the gradient error stays <= h2_tol against
  FP64, and builds whose ranks reach the cap are counted.
"""
import ctypes
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
from datasets import gaussian_mixture, init_normal, joint_P_oracle, knn_graph
from tsne_h2pack._backend import H2Stats

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def problem():
    X, _ = gaussian_mixture(6000, dim=20, n_clusters=12, seed=3)
    return joint_P_oracle(knn_graph(X, 45), 15.0)


def _params(backend, n, n_iter=1000, tol=1e-2):
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = n_iter
    p.h2_tol = tol
    p.h2_rebuild_freq = 3
    p.h2_rebuild_freq_early = 10
    p.learning_rate = max(n / 12.0, 200.0)
    return p


def test_schedule_rebuilds_and_renumberings(backend, problem):
    n = problem.shape[0]
    ws = backend.init_workspace(problem, 2, _params(backend, n), init_Y=init_normal(n, 2, 1).T.copy())
    try:
        assert backend.lib.tsne_run(ws) == 0
        nr, nn = ctypes.c_int(), ctypes.c_int()
        assert backend.lib.tsne_run_state(ws, ctypes.byref(nr), ctypes.byref(nn), None, None) == 0
    finally:
        backend.lib.tsne_destroy(ws)
    assert nr.value == sum(oracle.is_rebuild_iter(it, 3, 10, 250, True) for it in range(1000))
    assert nr.value == 292
    assert nn.value == 6


@pytest.mark.parametrize("dim", [2, 3])
def test_kl_at_fixed_state_equals_oracle(backend, problem, dim):
    n = problem.shape[0]
    rng = np.random.default_rng(4 + dim)
    Y = rng.normal(size=(dim, n)) * 6.0 + rng.integers(0, 4, size=(1, n)) * 15.0
    ws = backend.init_workspace(problem, dim, _params(backend, n, tol=1e-2 if dim == 2 else 3e-2),
                                init_Y=init_normal(n, dim, 2).T.copy())
    try:
        assert backend.lib.tsne_set_embedding(ws, np.ascontiguousarray(Y)) == 0
        assert backend.lib.tsne_run_range(ws, 999, 1000) == 0  # a rebuild iteration
        kl = backend.lib.compute_kl_divergence(ws)
        Z = ctypes.c_double()
        assert backend.lib.tsne_run_state(ws, None, None, ctypes.byref(Z), None) == 0
    finally:
        backend.lib.tsne_destroy(ws)
    kl_o = oracle.kl(problem, Y, Z.value, 1.0)
    _say(f"KL at fixed state ({dim}D): library {kl:.12f} oracle {kl_o:.12f} (Z {Z.value:.6e})")
    assert abs(kl - kl_o) <= 1e-10 * abs(kl_o)


def test_fp64_check_kernel_vs_cpu_oracle_100k(backend):
    n = 100_000
    X, _ = gaussian_mixture(n, dim=50, n_clusters=20, seed=0, spread=5.0, dirichlet=True,
                            std_range=(0.5, 2.0), dtype=np.float32)
    P, _ = backend.joint_P_from_points(X.astype(np.float64), 30, 10.0)
    del X
    rng = np.random.default_rng(6)
    Y = rng.normal(size=(2, n)) * 20.0
    ws = backend.init_workspace(P, 2, _params(backend, n), init_Y=init_normal(n, 2, 3).T.copy())
    rows = np.sort(rng.choice(n, size=400, replace=False)).astype(np.int32)
    try:
        assert backend.lib.tsne_set_embedding(ws, np.ascontiguousarray(Y)) == 0
        rep = np.zeros((rows.size, 2))
        Zd = ctypes.c_double()
        assert backend.lib.tsne_exact_repulsive_rows(ws, rows, rows.size, rep, ctypes.byref(Zd)) == 0
    finally:
        backend.lib.tsne_destroy(ws)
    rep_o, Z_o, _, _ = oracle.repulsive_exact(Y)
    ref = rep_o[:, rows].T
    rel = np.abs(rep - ref).max() / np.abs(ref).max()
    _say(f"FP64 device vs CPU oracle at 100K: Z {Zd.value:.15e} / {Z_o:.15e}, rows rel {rel:.2e}")
    assert abs(Zd.value / Z_o - 1) <= 1e-10
    assert rel <= 1e-10


_CHILD = r"""
import ctypes, json, os, sys
import numpy as np
sys.path[:0] = [os.path.join(sys.argv[1], "tsne-h2pack-python_amd")]
from scipy.sparse import csr_matrix
from tsne_h2pack._backend import get_backend
be = get_backend()
z = np.load(sys.argv[2])
P = csr_matrix((z["val"], z["idx"], z["ptr"]), shape=(z["ptr"].size - 1,) * 2)
n = P.shape[0]
p = be.create_default_params()
p.verbose = 0
p.h2_tol = 1e-2
ws = be.init_workspace(P, 2, p, init_Y=z["Y"])
g = np.zeros((2, n))
Z = ctypes.c_double()
assert be.lib.tsne_compute_gradient(ws, 1.0, 0, g, ctypes.byref(Z)) == 0
be.lib.tsne_destroy(ws)
path = ctypes.create_string_buffer(4096)
dig, disk = ctypes.c_ulonglong(), ctypes.c_int()
assert be.lib.tsne_proxy_cache_info(2, ctypes.c_double(1e-2), path, 4096, ctypes.byref(dig),
                                    ctypes.byref(disk)) == 0
np.save(sys.argv[3], g)
print(json.dumps({"from_disk": disk.value, "digest": dig.value, "path": path.value.decode()}))
"""


def test_proxy_cache_gives_bit_identical_gradient(backend, problem, tmp_path):
    root = os.path.dirname(HERE)
    n = problem.shape[0]
    inp = str(tmp_path / "in.npz")
    np.savez(inp, ptr=problem.indptr.astype(np.int32), idx=problem.indices.astype(np.int32),
             val=problem.data, Y=init_normal(n, 2, 8).T.copy() * 1e3)
    cache = str(tmp_path / "cache")
    runs = {}
    for name, env in (("generated", {"TSNE_H2PACK_CACHE_DIR": cache}),
                      ("from_disk", {"TSNE_H2PACK_CACHE_DIR": cache}),
                      ("cache_off", {"TSNE_H2PACK_NO_PROXY_CACHE": "1"})):
        e = dict(os.environ)
        e.pop("TSNE_H2PACK_NO_PROXY_CACHE", None)
        e.update(env)
        out = str(tmp_path / f"g_{name}.npy")
        _say(f"proxy cache child: {name}")
        r = subprocess.run([sys.executable, "-c", _CHILD, root, inp, out], env=e,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        info = json.loads(r.stdout.strip().splitlines()[-1])
        runs[name] = (np.load(out), info)
    assert runs["generated"][1]["from_disk"] == 0
    assert runs["from_disk"][1]["from_disk"] == 1
    assert os.path.exists(runs["from_disk"][1]["path"])
    assert runs["cache_off"][1]["path"] == ""
    digests = {v[1]["digest"] for v in runs.values()}
    assert len(digests) == 1
    g0 = runs["generated"][0]
    assert np.array_equal(g0, runs["from_disk"][0])
    assert np.array_equal(g0, runs["cache_off"][0])


def test_reference_default_tol_gradient_bound(backend, problem):
    """h2_tol = 1e-3 (tsne_create_default_params): the ranks this needs approach the cap;
    the gradient must still meet the bound on all rows."""
    n = problem.shape[0]
    for scale in (2.0, 30.0):
        rng = np.random.default_rng(int(scale))
        cent = rng.normal(scale=scale, size=(30, 2))
        lab = rng.integers(0, 30, n)
        Y0 = (cent[lab] + rng.normal(size=(n, 2)) * scale / 20).T.copy()
        ws = backend.init_workspace(problem, 2, _params(backend, n, tol=1e-3), init_Y=Y0.T.copy())
        try:
            g = np.zeros((2, n))
            Z = ctypes.c_double()
            assert backend.lib.tsne_compute_gradient(ws, 1.0, 0, g, ctypes.byref(Z)) == 0
            st = H2Stats()
            assert backend.lib.tsne_h2_stats(ws, ctypes.byref(st)) == 0
        finally:
            backend.lib.tsne_destroy(ws)
        rep64, Z64, _, _ = oracle.repulsive_exact(Y0)
        g64 = oracle.attractive(problem, Y0, 1.0) - rep64
        err = np.linalg.norm(g - g64) / np.linalg.norm(g64)
        _say(f"h2_tol 1e-3 scale {scale}: grad err {err:.2e}, rank max {st.rank_max}, "
             f"rank-capped builds {st.rank_capped_builds}")
        assert err <= 1e-3

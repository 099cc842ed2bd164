"""Failure paths and the entry points around the run, on the GPU through the C ABI.

* A non-finite initial embedding at N >= 2000 makes the initial H2 build fail:
  tsne_init_with_P returns NULL and P stays with the caller (freed once, by the
  caller), as the reference returns NULL (src/tsne_h2pack.c:1518-1523).  Below 2000
  points (direct sum, no build at init) the run itself fails with -1.
* The Z guard (reference src/tsne_h2pack.c:604-611): when every pairwise kernel value
  underflows, Z is not positive; the run reports it, leaves the repulsive force out of the
  gradient and goes on with finite values.
* Identical points: a tree whose single deepest leaf holds every point cannot be used;
  the product falls back to the direct sum, which is exact here (all differences are 0).
* tsne_fit_transform_with_P end to end equals init + run + get_embedding.
* The profiling CSV has the reference's schema and fills z_computation and gradient.
* In the job server a failing job fails alone; the others still match their results
  alone.  A second wait on one job returns -1; destroying the server wakes a waiter.
"""
import ctypes
import os
import sys
import threading

import numpy as np
import pytest

import oracle
from datasets import gaussian_mixture, init_normal, joint_P_oracle, knn_graph
from tsne_h2pack._backend import H2Stats, scipy_csr_to_c_sparse_matrix

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def P3k():
    X, _ = gaussian_mixture(3000, dim=20, n_clusters=8, seed=4)
    return joint_P_oracle(knn_graph(X, 45), 15.0)


@pytest.fixture(scope="module")
def P1k():
    X, _ = gaussian_mixture(1200, dim=20, n_clusters=6, seed=5)
    return joint_P_oracle(knn_graph(X, 30), 10.0)


def _params(backend, n, n_iter=1000, **kw):
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = n_iter
    p.h2_tol = 1e-2
    p.h2_rebuild_freq = 3
    p.h2_rebuild_freq_early = 10
    p.learning_rate = max(n / 12.0, 200.0)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def test_nonfinite_init_returns_null_and_keeps_P(backend, P3k, capfd):
    n = P3k.shape[0]
    Y0 = init_normal(n, 2, 1).T.copy()
    Y0[17, 1] = np.nan
    P_c = scipy_csr_to_c_sparse_matrix(P3k, backend.lib)
    p = _params(backend, n)
    y = np.ascontiguousarray(Y0.T)
    p.init_method = 1
    p.init_Y = y.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ws = backend.lib.tsne_init_with_P(None, n, 0, 2, ctypes.byref(p), P_c)
    assert not ws
    # P untouched: still readable and released here exactly once
    m = P_c.contents
    assert m.n_rows == n and m.nnz == P3k.nnz and m.row_ptr[n] == P3k.nnz
    backend.lib.sparse_matrix_destroy(P_c)
    err = capfd.readouterr().err
    assert "NaN or Inf" in err
    # the Python layer raises cleanly (and frees its copy of P once)
    with pytest.raises(RuntimeError, match="NULL"):
        backend.init_workspace(P3k, 2, _params(backend, n), init_Y=Y0)


def test_nonfinite_init_direct_sum_fails_the_run(backend, P1k):
    n = P1k.shape[0]
    Y0 = init_normal(n, 2, 2).T.copy()
    Y0[3, 0] = np.inf
    ws = backend.init_workspace(P1k, 2, _params(backend, n, n_iter=5), init_Y=Y0)
    try:
        assert backend.lib.tsne_run(ws) == -1
    finally:
        backend.lib.tsne_destroy(ws)


def test_z_guard_leaves_repulsion_out(backend, P1k, capfd):
    """Points 1e20 apart: every (1 + |y_i - y_j|^2)^-1 underflows in float32, Z = 0."""
    n = P1k.shape[0]
    rng = np.random.default_rng(3)
    Y0 = rng.normal(size=(n, 2)) * 1e20
    ws = backend.init_workspace(P1k, 2, _params(backend, n, n_iter=3), init_Y=Y0)
    try:
        g = np.zeros((2, n))
        Z = ctypes.c_double()
        assert backend.lib.tsne_compute_gradient(ws, 1.0, 0, g, ctypes.byref(Z)) == 0
        assert backend.lib.tsne_run(ws) == 0
        Y = np.zeros((2, n))
        backend.lib.tsne_get_embedding(ws, Y)
    finally:
        backend.lib.tsne_destroy(ws)
    assert Z.value == 1.0  # the guard's replacement value
    assert np.isfinite(g).all()
    # no repulsion and (at these distances) no attraction: the gradient is the oracle's
    # attractive force, i.e. zero, and the embedding does not move
    np.testing.assert_allclose(g, oracle.attractive(P1k, Y0.T.copy(), 1.0), atol=1e-30)
    assert np.isfinite(Y).all()
    np.testing.assert_allclose(Y, Y0.T - Y0.T.mean(axis=1, keepdims=True), rtol=0, atol=1e8)
    assert "Z <= 0" in capfd.readouterr().err


def test_identical_points_fall_back_to_direct_sum(backend, P3k):
    n = P3k.shape[0]
    Y0 = np.full((n, 2), 0.25)
    ws = backend.init_workspace(P3k, 2, _params(backend, n, n_iter=20), init_Y=Y0)
    try:
        g = np.zeros((2, n))
        Z = ctypes.c_double()
        assert backend.lib.tsne_compute_gradient(ws, 12.0, 0, g, ctypes.byref(Z)) == 0
        st = H2Stats()
        assert backend.lib.tsne_h2_stats(ws, ctypes.byref(st)) == 0
        assert backend.lib.tsne_run(ws) == 0
        Y = np.zeros((2, n))
        backend.lib.tsne_get_embedding(ws, Y)
    finally:
        backend.lib.tsne_destroy(ws)
    _say(f"identical points: direct fallbacks {st.direct_fallbacks}, Z {Z.value}")
    assert st.direct_fallbacks >= 1
    assert np.array_equal(g, np.zeros_like(g))
    assert Z.value == float(n) * (n - 1)  # every pair contributes w = 1
    assert np.isfinite(Y).all()


def test_fit_transform_with_P_equals_workspace_path(backend, P3k):
    n = P3k.shape[0]
    Y0 = init_normal(n, 2, 9)
    p = _params(backend, n, n_iter=300)
    y = np.ascontiguousarray(Y0)
    p.init_method = 1
    p.init_Y = y.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    P_c = scipy_csr_to_c_sparse_matrix(P3k, backend.lib)  # taken by the call
    out = np.zeros((2, n))
    assert backend.lib.tsne_fit_transform_with_P(None, n, 0, 2, ctypes.byref(p), P_c, out) == 0
    ws = backend.init_workspace(P3k, 2, _params(backend, n, n_iter=300), init_Y=Y0.T.copy())
    try:
        assert backend.lib.tsne_run(ws) == 0
        ref = np.zeros((2, n))
        backend.lib.tsne_get_embedding(ws, ref)
    finally:
        backend.lib.tsne_destroy(ws)
    assert np.array_equal(out, ref)


def test_profiling_csv_schema_and_columns(backend, P3k, tmp_path):
    n = P3k.shape[0]
    f = str(tmp_path / "prof.csv")
    old = os.environ.get("TSNE_OVERLAP")
    os.environ["TSNE_OVERLAP"] = "1"  # the attractive pass of rebuild iterations split off
    try:
        p = _params(backend, n, n_iter=60, enable_profiling=1)
        name = f.encode()
        p.profiling_output_file = name
        ws = backend.init_workspace(P3k, 2, p, init_Y=init_normal(n, 2, 4).T.copy())
    finally:
        if old is None:
            del os.environ["TSNE_OVERLAP"]
        else:
            os.environ["TSNE_OVERLAP"] = old
    try:
        backend.lib.tsne_set_init_profiling_times(ws, 1.5, 0.25)
        assert backend.lib.tsne_run(ws) == 0
    finally:
        backend.lib.tsne_destroy(ws)
    lines = open(f).read().splitlines()
    assert lines[0] == ("iteration,p_matvec,h2_rebuild,h2_matvec_s,h2_matvec_v,"
                        "z_computation,gradient,other,total")
    rows = np.array([[float(x) for x in l.split(",")] for l in lines[1:61]])
    assert rows.shape == (60, 9) and np.array_equal(rows[:, 0], np.arange(60))
    reb = np.array([oracle.is_rebuild_iter(it, 3, 10, 250, True) for it in range(60)])
    fresh = reb.copy()
    fresh[0] = True  # iteration 0 forms the product of the initial build
    assert (rows[reb, 2] > 0).all()                                  # rebuild
    assert (rows[fresh, 3] > 0).all() and (rows[fresh, 5] > 0).all()  # product, Z
    assert (rows[reb, 6] > 0).all()        # update kernel after the overlapped pass
    # force reused: no rebuild, product or Z; fused update: counted in p_matvec
    assert (rows[~fresh][:, [2, 3, 5, 6]] < 1e-4).all()
    assert (rows[~fresh, 1] > 0).all()
    assert (rows[:, 4] == 0).all()                                   # fused product
    np.testing.assert_allclose(rows[:, 8], rows[:, 1:8].sum(axis=1), rtol=0, atol=5e-6)
    text = "\n".join(lines)
    assert "# k-NN time: 1.500 seconds" in text and "# P construction: 0.250 seconds" in text


def test_failing_job_fails_alone_in_the_mix(backend, P3k, P1k):
    from tsne_h2pack.serving import TSNEServer
    jobs = [(P3k, init_normal(P3k.shape[0], 2, 11).T.copy()),
            (P1k, init_normal(P1k.shape[0], 2, 12).T.copy()),
            (P3k, init_normal(P3k.shape[0], 2, 13).T.copy())]
    bad = init_normal(P3k.shape[0], 2, 14).T.copy()
    bad[5, 0] = np.nan
    n_iter = 200
    alone = []
    for P, Y0 in jobs:
        ws = backend.init_workspace(P, 2, _params(backend, P.shape[0], n_iter=n_iter), init_Y=Y0)
        try:
            assert backend.lib.tsne_run(ws) == 0
            Y = np.zeros((2, P.shape[0]))
            backend.lib.tsne_get_embedding(ws, Y)
            alone.append(Y.T.copy())
        finally:
            backend.lib.tsne_destroy(ws)
    with TSNEServer(slots=4) as srv:
        ids = []
        for k, (P, Y0) in enumerate(jobs):
            ids.append(srv.submit(P, 2, srv.default_params(P.shape[0], n_iter=n_iter, h2_tol=1e-2,
                                                           h2_rebuild_freq=3, h2_rebuild_freq_early=10),
                                  init_Y=Y0))
            if k == 0:
                bad_id = srv.submit(P3k, 2, srv.default_params(P3k.shape[0], n_iter=n_iter),
                                    init_Y=bad)
        with pytest.raises(RuntimeError, match="failed"):
            srv.wait(bad_id, P3k.shape[0], 2)
        got = [srv.wait(j, P.shape[0], 2)[0] for j, (P, _) in zip(ids, jobs)]
    for a, b in zip(alone, got):
        assert np.abs(a - b).max() <= 1e-9 * np.abs(a).max()


def test_server_second_waiter_and_destroy_wakes_waiter(backend, P3k, P1k):
    from tsne_h2pack._backend import JobInfo
    from tsne_h2pack.serving import TSNEServer
    lib = backend.lib
    srv = TSNEServer(slots=1)
    try:
        a = srv.submit(P3k, 2, srv.default_params(P3k.shape[0], n_iter=400),
                       init_Y=init_normal(P3k.shape[0], 2, 21).T.copy())
        b = srv.submit(P1k, 2, srv.default_params(P1k.shape[0], n_iter=50),
                       init_Y=init_normal(P1k.shape[0], 2, 22).T.copy())
        res = {}

        def waiter(key, jid, n):
            Y = np.zeros((2, n))
            res[key] = lib.tsne_server_wait(srv._srv, jid, Y, ctypes.byref(JobInfo()))

        t1 = threading.Thread(target=waiter, args=("a1", a, P3k.shape[0]))
        t1.start()
        while srv.poll(a)["status"] == "queued":
            pass
        # the first waiter has claimed job a (or will the moment it runs): a second
        # wait on it returns -1 without blocking, once the claim is in place
        import time
        time.sleep(0.05)
        Y = np.zeros((2, P3k.shape[0]))
        assert lib.tsne_server_wait(srv._srv, a, Y, ctypes.byref(JobInfo())) == -1
        t1.join()
        assert res["a1"] == 0
        # job b is waited for by a thread while the server is destroyed: the waiter
        # returns (b ran to completion or was dropped) and destroy returns after it
        c = srv.submit(P3k, 2, srv.default_params(P3k.shape[0], n_iter=300),
                       init_Y=init_normal(P3k.shape[0], 2, 23).T.copy())
        entering = threading.Event()

        def waiter_c():
            entering.set()
            waiter("c", c, P3k.shape[0])

        t2 = threading.Thread(target=waiter_c)
        t2.start()
        entering.wait()
        time.sleep(0.2)  # the waiter is inside tsne_server_wait (ctypes drops the GIL)
    finally:
        srv.close()
    t2.join(timeout=60)
    assert not t2.is_alive()
    assert res["c"] in (0, -1)

"""Job server (continuous batching) on the GPU, through the C ABI.

* Each job's embedding is the same when it runs alone (tsne_init_with_P + tsne_run) and
  when it runs in a mix of concurrent jobs of uneven size, 2D and 3D: tolerance 1e-9
  relative to the embedding's scale (the kernels are deterministic, so the runs are
  expected to agree exactly; the tolerance only allows for that to change).
* Admission by device footprint: with a budget below one job's footprint, jobs run one
  at a time, in arrival order.
* Invalid submissions are rejected without taking P.
* A server set to the mixed H2 precision runs each job as that job alone with
  tsne_set_h2_precision(ws, 1) (same tolerance), and differently from the plain run.
"""
import ctypes
import sys

import numpy as np
import pytest

from datasets import gaussian_mixture, joint_P_oracle, knn_graph

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def jobs():
    out = []
    for k, (n, dim, seed) in enumerate([(2500, 2, 1), (6000, 2, 2), (3200, 3, 3),
                                        (1800, 2, 4), (4500, 2, 5), (2600, 3, 6)]):
        X, _ = gaussian_mixture(n, dim=20, n_clusters=5 + k, seed=seed)
        P = joint_P_oracle(knn_graph(X, 45), 15.0)
        Y0 = np.random.default_rng(seed).normal(scale=1e-2, size=(n, dim))
        out.append((P, dim, Y0))
    return out


def _params(backend, n, n_iter):
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = n_iter
    p.learning_rate = max(n / 12.0, 200.0)
    return p


def _alone(backend, P, dim, Y0, n_iter, mixed=False):
    n = P.shape[0]
    ws = backend.init_workspace(P, dim, _params(backend, n, n_iter), init_Y=Y0)
    try:
        if mixed:
            assert backend.lib.tsne_set_h2_precision(ws, 1) == 0
        assert backend.lib.tsne_run(ws) == 0
        Y = np.zeros((dim, n))
        backend.lib.tsne_get_embedding(ws, Y)
        kl = backend.lib.compute_kl_divergence(ws)
    finally:
        backend.lib.tsne_destroy(ws)
    return Y.T, kl


def test_results_same_alone_and_in_the_mix(backend, jobs):
    from tsne_h2pack.serving import TSNEServer
    n_iter = 400
    alone = []
    for P, dim, Y0 in jobs:
        _say(f"alone n={P.shape[0]} dim={dim}")
        alone.append(_alone(backend, P, dim, Y0, n_iter))
    with TSNEServer(slots=4) as srv:
        ids = [srv.submit(P, dim, srv.default_params(P.shape[0], n_iter=n_iter), init_Y=Y0)
               for P, dim, Y0 in jobs]
        _say(f"submitted {len(ids)} jobs")
        got = [srv.wait(j, P.shape[0], dim) for j, (P, dim, _) in zip(ids, jobs)]
    overlap = 0
    infos = [g[1] for g in got]
    for a in infos:
        for b in infos:
            if a is not b and a["start_s"] < b["end_s"] and b["start_s"] < a["end_s"]:
                overlap += 1
    _say(f"overlapping job pairs: {overlap // 2}")
    assert overlap > 0, "jobs never ran concurrently"
    for (Ya, kla), (Ym, info) in zip(alone, got):
        rel = np.abs(Ya - Ym).max() / np.abs(Ya).max()
        _say(f"  n={info['n_point']} d={info['low_dim']} rel diff {rel:.2e} KL {kla:.5f}/{info['kl']:.5f}")
        assert rel <= 1e-9
        assert abs(kla - info["kl"]) <= 1e-9 * abs(kla)


def test_footprint_admission_serialises(backend, jobs):
    from tsne_h2pack.serving import TSNEServer
    sel = jobs[:3]
    with TSNEServer(slots=4, hbm_budget_bytes=1) as srv:
        ids = [srv.submit(P, dim, srv.default_params(P.shape[0], n_iter=60), init_Y=Y0)
               for P, dim, Y0 in sel]
        infos = [srv.wait(j, P.shape[0], dim)[1] for j, (P, dim, _) in zip(ids, sel)]
    for a, b in zip(infos, infos[1:]):
        assert b["start_s"] >= a["end_s"] - 1e-6
    assert all(i["footprint_bytes"] > 1 for i in infos)


def test_invalid_submission_rejected(backend, jobs):
    from tsne_h2pack._backend import scipy_csr_to_c_sparse_matrix
    from tsne_h2pack.serving import TSNEServer
    P, dim, _ = jobs[0]
    with TSNEServer(slots=1) as srv:
        P_c = scipy_csr_to_c_sparse_matrix(P, backend.lib)
        p = srv.default_params(P.shape[0])
        rc = backend.lib.tsne_server_submit(srv._srv, P.shape[0], 4, ctypes.byref(p), P_c, None)
        assert rc == -1
        backend.lib.sparse_matrix_destroy(P_c)   # not taken
        assert srv.status()["queued"] == 0


def test_mixed_precision_server_equals_mixed_alone(backend, jobs):
    from tsne_h2pack.serving import TSNEServer
    n_iter = 300
    sel = [j for j in jobs if j[0].shape[0] >= 2000][:4]  # H2 jobs (N >= 2000)
    mixed = [_alone(backend, P, dim, Y0, n_iter, mixed=True) for P, dim, Y0 in sel]
    plain = [_alone(backend, P, dim, Y0, n_iter) for P, dim, Y0 in sel]
    with TSNEServer(slots=4, h2_precision="mixed") as srv:
        ids = [srv.submit(P, dim, srv.default_params(P.shape[0], n_iter=n_iter), init_Y=Y0)
               for P, dim, Y0 in sel]
        got = [srv.wait(j, P.shape[0], dim) for j, (P, dim, _) in zip(ids, sel)]
    differs = 0
    for (Ya, kla), (Yp, _), (Ym, info) in zip(mixed, plain, got):
        rel = np.abs(Ya - Ym).max() / np.abs(Ya).max()
        _say(f"  mixed n={info['n_point']} d={info['low_dim']} rel diff {rel:.2e} "
             f"KL {kla:.6f}/{info['kl']:.6f}")
        assert rel <= 1e-9
        assert abs(kla - info["kl"]) <= 1e-9 * abs(kla)
        differs += int(not np.array_equal(Ya, Yp))
    assert differs > 0, "mixed jobs ran as plain"
    with pytest.raises(ValueError):
        TSNEServer(slots=1, h2_precision="half")

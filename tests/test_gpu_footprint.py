"""The job server admits jobs by their estimated device footprint
(tsne_job_footprint): the estimate must bound what a job really holds on the device.

Each job runs its whole 1000-iteration schedule (the initial build, 292 rebuilds, the six
renumberings with their second CSR and sort scratch) while a host thread polls the
device's free memory (hipMemGetInfo) every half millisecond; the job's peak is the
largest drop below the free memory measured before its workspace was made.  A first
small job loads the library's code objects and tables, so that they are not counted.
"""
import ctypes
import threading
import time

import numpy as np
import pytest

from datasets import gaussian_mixture

pytestmark = pytest.mark.gpu


def _hip():
    # the runtime instance the library already loaded (same soname)
    for name in ("libamdhip64.so.7", "libamdhip64.so"):
        try:
            lib = ctypes.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.fail("the HIP runtime library is not loadable")
    lib.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    lib.hipMemGetInfo.restype = ctypes.c_int
    return lib


def _free(hip):
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
    return f.value


def _run(backend, P, dim, seed):
    n = P.shape[0]
    p = backend.create_default_params()
    p.verbose = 0
    p.n_iter = 1000
    p.h2_tol = 1e-2 if dim == 2 else 3e-2
    p.h2_rebuild_freq = 3
    p.learning_rate = max(n / 12.0, 200.0)
    Y0 = np.random.default_rng(seed).normal(scale=1e-4, size=(n, dim))
    ws = backend.init_workspace(P, dim, p, init_Y=Y0)
    try:
        assert backend.lib.tsne_run(ws) == 0
    finally:
        backend.lib.tsne_destroy(ws)


@pytest.mark.parametrize("n,dim,k,perp", [(60_000, 2, 90, 30.0), (40_000, 3, 150, 50.0)])
def test_job_footprint_bounds_measured_peak(backend, n, dim, k, perp):
    hip = _hip()
    X, _ = gaussian_mixture(n, dim=50, n_clusters=20, seed=11, spread=5.0, dirichlet=True,
                            std_range=(0.5, 2.0), dtype=np.float32)
    P, _ = backend.joint_P_from_points(X.astype(np.float64), k, perp)
    Pw, _ = backend.joint_P_from_points(X[:3000].astype(np.float64), k, perp)
    del X
    _run(backend, Pw, dim, 0)  # loads code objects and tables
    nnz = int(P.nnz)
    fp = backend.lib.tsne_job_footprint(n, dim, nnz)
    free0 = _free(hip)
    low = [free0]
    stop = threading.Event()

    def poll():
        while not stop.is_set():
            low[0] = min(low[0], _free(hip))
            time.sleep(0.0005)

    th = threading.Thread(target=poll, daemon=True)
    th.start()
    try:
        _run(backend, P, dim, 1)
    finally:
        stop.set()
        th.join()
    peak = free0 - low[0]
    print(f"\n  n={n} {dim}D nnz={nnz}: measured peak {peak / 2**20:.1f} MiB, "
          f"footprint {fp / 2**20:.1f} MiB ({fp / max(peak, 1):.2f}x)", flush=True)
    assert peak > 0
    assert peak <= fp

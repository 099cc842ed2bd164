"""Serving: many t-SNE jobs of uneven size on one GPU, with continuous batching.

`TSNEServer` wraps the library's job server (tsne_server_* in the C ABI).  Jobs are
joint-P matrices (as `_backend.joint_P` returns them) with their run parameters and
an optional initial embedding.  They are admitted in arrival order into a fixed number
of concurrent slots while their estimated device footprint fits the HBM budget, and
each one runs on its own HIP stream with all of its state resident on the device.  A
job's result is the same as that of a standalone `tsne_init_with_P` + `tsne_run` on the
same inputs.

The reference has no serving mode (one job per process); the job parameters are the
reference's `tsne_params_t`.
"""
import ctypes
from typing import Optional

import numpy as np

from ._backend import JobInfo, NP_DTYPE, get_backend, scipy_csr_to_c_sparse_matrix

STATUS = {0: "queued", 1: "running", 2: "done", -1: "failed"}


class TSNEServer:
    """Job server over the library's continuous-batching scheduler.

    slots: jobs that may run at once; hbm_budget_bytes: device memory that running
    jobs may hold together (0: 90% of the free memory at creation); h2_precision:
    'plain' or 'mixed' for every job, or None for the library default
    (TSNE_H2_PRECISION)."""

    def __init__(self, slots: int = 8, hbm_budget_bytes: int = 0,
                 h2_precision: Optional[str] = None):
        if h2_precision not in (None, 'plain', 'mixed'):
            raise ValueError("h2_precision must be 'plain', 'mixed' or None")
        self._be = get_backend()
        self._lib = self._be.lib
        self._srv = self._lib.tsne_server_create(int(slots), int(hbm_budget_bytes))
        if not self._srv:
            raise RuntimeError("tsne_server_create failed (no HIP device?)")
        self.slots = slots
        if h2_precision is not None:
            self._lib.tsne_server_set_h2_precision(self._srv, 1 if h2_precision == 'mixed' else 0)

    def default_params(self, n_point: int, **overrides):
        """Reference defaults (learning rate max(N/12, 200)), verbose off, with overrides
        by field name of tsne_params_t."""
        p = self._be.create_default_params()
        p.verbose = 0
        p.learning_rate = max(n_point / 12.0, 200.0)
        for k, v in overrides.items():
            if not hasattr(p, k):
                raise AttributeError(f"tsne_params_t has no field {k!r}")
            setattr(p, k, v)
        return p

    def submit(self, P, n_components: int = 2, params=None,
               init_Y: Optional[np.ndarray] = None) -> int:
        """Queue a job; P is the joint P (scipy CSR, n x n).  init_Y: (n, n_components)
        or None for the seeded random init.  Returns the job id."""
        n = P.shape[0]
        if params is None:
            params = self.default_params(n)
        y = None
        if init_Y is not None:
            init_Y = np.asarray(init_Y, dtype=NP_DTYPE)
            if init_Y.shape != (n, n_components):
                raise ValueError(f"init_Y must have shape {(n, n_components)}")
            y = np.ascontiguousarray(init_Y.T)
        P_c = scipy_csr_to_c_sparse_matrix(P, self._lib)
        params.init_Y = None
        jid = self._lib.tsne_server_submit(self._srv, n, n_components, ctypes.byref(params),
                                           P_c, y)
        if jid < 0:
            self._lib.sparse_matrix_destroy(P_c)
            raise RuntimeError("tsne_server_submit rejected the job")
        return jid

    def poll(self, job: int) -> dict:
        info = JobInfo()
        st = self._lib.tsne_server_poll(self._srv, job, ctypes.byref(info))
        if st == -2:
            raise KeyError(job)
        return _info_dict(info)

    def wait(self, job: int, n_point: int, n_components: int):
        """Block until the job ends; returns (Y of shape (n, n_components), info dict).
        Raises RuntimeError if the job failed."""
        Y = np.zeros((n_components, n_point), dtype=NP_DTYPE)
        info = JobInfo()
        rc = self._lib.tsne_server_wait(self._srv, job, Y, ctypes.byref(info))
        d = _info_dict(info)
        if rc != 0:
            raise RuntimeError(f"job {job} failed")
        return Y.T.copy(), d

    def status(self) -> dict:
        q, r, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        self._lib.tsne_server_status(self._srv, ctypes.byref(q), ctypes.byref(r),
                                     ctypes.byref(b))
        return {"queued": q.value, "running": r.value, "hbm_in_use": b.value}

    def close(self):
        if self._srv:
            self._lib.tsne_server_destroy(self._srv)
            self._srv = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _info_dict(info: JobInfo) -> dict:
    return {"job_id": info.job_id, "status": STATUS.get(info.status, info.status),
            "n_point": info.n_point, "low_dim": info.low_dim, "kl": info.kl,
            "submit_s": info.submit_s, "start_s": info.start_s, "end_s": info.end_s,
            "latency_s": info.end_s - info.submit_s,
            "footprint_bytes": info.footprint_bytes}


def autotune_sizes(sizes, dims=(2, 3), n_iter: int = 400) -> dict:
    """Tune the launch shapes of every size bucket (round(log2 N)) that the job sizes
    `sizes` fall in, for each embedding dimension in `dims`, with the library's autotuner
    (tsne_autotune on a synthetic job of 2^bucket points; none of the tuned shapes changes a
    result).  The table persists in the cache directory, and later jobs of those sizes --
    served or not -- start with their bucket's shapes.  Returns {(dim, bucket): shapes}."""
    be = get_backend()
    buckets = sorted({int(round(np.log2(max(int(n), 2)))) for n in sizes})
    out = {}
    for d in dims:
        for b in buckets:
            n = max(2000, 1 << b)
            out[(d, b)] = be.autotune(d, n, n_iter)
    return out

"""C-ABI surface of libtsne_h2pack.so: the library loads on a machine without a GPU,
exports every entry point include/tsne_h2pack.h declares, and the struct mirrors
match the reference's layout (reference tests/test_backend.py:53-90).  No compute
calls (those need a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from tsne_h2pack import _backend as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "tsne_h2pack.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(\w+)\s*\([^;{]*\)\s*;", text)))


def test_header_declares_the_thirteen_reference_symbols():
    ref13 = {"tsne_create_default_params", "tsne_init", "tsne_init_with_P", "tsne_run",
             "tsne_get_embedding", "tsne_destroy", "tsne_fit_transform",
             "tsne_fit_transform_with_P", "compute_P_from_knn_distances",
             "sparse_matrix_create", "sparse_matrix_destroy",
             "tsne_set_init_profiling_times", "compute_kl_divergence"}
    assert ref13 <= set(_declared_symbols())


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(B.library_path()))
    missing = [s for s in _declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_backend_configures_signatures():
    be = B.get_backend()
    assert be.backend_type == "hip"


def test_dtype_is_float64():
    """Reference tests/test_backend.py:53-64."""
    assert B.DTYPE == ctypes.c_double
    assert B.NP_DTYPE == np.float64


def test_struct_sizes_and_offsets():
    assert ctypes.sizeof(B.TSNEParams) == 160
    assert ctypes.sizeof(B.SparseMatrix) == 40
    off = {name: getattr(B.TSNEParams, name).offset for name, _ in B.TSNEParams._fields_}
    expect = dict(n_iter=0, perplexity=8, learning_rate=16, momentum=24, initial_momentum=32,
                  momentum_switch_iter=40, early_exag=48, early_exag_iter=56, use_vptree=60,
                  h2_tol=64, h2_rebuild_freq=72, h2pack_print_level=76, init_method=80,
                  random_seed=84, verbose=88, print_freq=92, enable_early_stopping=96,
                  convergence_tol=104, convergence_patience=112, min_iter=116,
                  sklearn_compatible=120, enable_profiling=124, profiling_output_file=128,
                  num_threads=136, h2_leaf_size=140, h2_rebuild_freq_early=144, init_Y=152)
    for k, v in expect.items():
        assert off[k] == v, k


def test_c_defaults_match_python_defaults():
    """tsne_create_default_params returns the reference's defaults by value
    (reference tests/test_backend.py:162-177); no device is touched."""
    be = B.get_backend()
    c = be.lib.tsne_create_default_params()
    p = be.create_default_params()
    for name in ("n_iter", "perplexity", "learning_rate", "momentum", "initial_momentum",
                 "momentum_switch_iter", "early_exag", "early_exag_iter", "h2_tol",
                 "h2_rebuild_freq", "verbose", "print_freq", "sklearn_compatible",
                 "h2_leaf_size", "h2_rebuild_freq_early", "num_threads"):
        assert getattr(c, name) == getattr(p, name), name


def test_sparse_matrix_roundtrip_host_only():
    """sparse_matrix_create / destroy are host allocations; marshalling of a scipy CSR
    into library memory (reference tests/test_backend.py:209-243)."""
    from scipy.sparse import csr_matrix
    be = B.get_backend()
    A = csr_matrix(np.array([[0, 1.5, 0], [2.0, 0, 3.0], [0, 0, 0]]))
    ptr = B.scipy_csr_to_c_sparse_matrix(A, be.lib)
    m = ptr.contents
    assert (m.n_rows, m.n_cols, m.nnz) == (3, 3, 3)
    assert list(np.ctypeslib.as_array(m.row_ptr, shape=(4,))) == [0, 1, 3, 3]
    assert list(np.ctypeslib.as_array(m.col_idx, shape=(3,))) == [1, 0, 2]
    assert list(np.ctypeslib.as_array(m.values, shape=(3,))) == [1.5, 2.0, 3.0]
    be.lib.sparse_matrix_destroy(ptr)


def test_init_rejects_bad_dims_without_device():
    """low_dim outside {2,3} and mismatched P are rejected before any device call
    (reference src/tsne_h2pack.c:1365-1376)."""
    be = B.get_backend()
    p = be.create_default_params()
    ptr = be.lib.sparse_matrix_create(4, 4, 0)
    assert not be.lib.tsne_init_with_P(None, 4, 0, 4, ctypes.byref(p), ptr)
    assert not be.lib.tsne_init_with_P(None, 5, 0, 2, ctypes.byref(p), ptr)
    be.lib.sparse_matrix_destroy(ptr)


def test_job_footprint_and_server_without_device():
    """The admission charge grows with N and nnz; without a device the server cannot be
    created (no host fallback)."""
    be = B.get_backend()
    f = be.lib.tsne_job_footprint
    assert f(1000, 2, 90000) < f(2000, 2, 90000) < f(2000, 2, 180000) < f(2000, 3, 180000)
    assert f(1_000_000, 2, 150_000_000) > 7_000_000_000
    import os
    if not os.path.exists("/dev/kfd"):
        assert not be.lib.tsne_server_create(4, 0)


def test_numpy_sum_replica_is_bit_exact():
    """tsne_numpy_sum (host code of the library, no device needed) reproduces numpy's
    float64 sum bit for bit: blocks of 8192 summed pairwise, in order.  The joint P's
    normalisation depends on it."""
    import ctypes

    import numpy as np

    from tsne_h2pack._backend import get_backend
    lib = get_backend().lib
    rng = np.random.default_rng(11)
    for n in [1, 2, 7, 8, 9, 127, 128, 129, 1000, 8191, 8192, 8193, 10000, 65537, 250001]:
        a = rng.random(n) * rng.choice([1e-9, 1.0, 1e7], n)
        got = lib.tsne_numpy_sum(a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
        assert got == np.sum(a), n


def test_tuning_entry_points_without_device():
    """Launch-shape tuning: invalid sizes are rejected; without a device there is no table
    (no device kind to key it), so a lookup gives the built-in shapes and the autotuner
    refuses sizes below the H2 path."""
    be = B.get_backend()
    t = B.Tuning()
    assert be.lib.tsne_tuning_lookup(4, 1000, ctypes.byref(t)) == -1
    assert be.lib.tsne_tuning_lookup(2, 0, ctypes.byref(t)) == -1
    assert be.lib.tsne_autotune(2, 1000, 400, ctypes.byref(t)) == -1
    assert be.lib.tsne_autotune(2, 5000, 10, ctypes.byref(t)) == -1
    if not os.path.exists("/dev/kfd"):
        shapes, found = be.tuning_lookup(2, 300_000)
        assert not found
        assert shapes["overlap"] == 1 and shapes["level_batch"] == 4 and shapes["graph"] == 0
        assert be.tuning_lookup(3, 5000)[0]["overlap"] == 0
        assert be.tuning_path() is None

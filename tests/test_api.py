"""TSNE constructor, validation and parameter mapping (reference tests/test_tsne.py:9-278,
702-762; tests/test_validation.py).  CPU only: nothing here reaches the device."""
import numpy as np
import pytest

from tsne_h2pack import H2_DEFAULTS, TSNE, h2_defaults
from tsne_h2pack._knn import _drop_self


def test_defaults():
    t = TSNE()
    assert t.n_components == 2 and t.perplexity == 30.0 and t.max_iter == 1000
    assert t.h2_tol == 1e-2 and t.h2_rebuild_freq == 3 and t.h2_rebuild_freq_early == 10
    assert t.h2_leaf_size == 100 and t.init == 'pca' and t.method == 'h2pack'


def test_3d_defaults():
    t = TSNE(n_components=3)
    assert t.h2_tol == 3e-2
    assert h2_defaults(3) == H2_DEFAULTS[3]


@pytest.mark.parametrize("kw,msg", [
    (dict(metric='cosine'), "not supported"),
    (dict(init='spectral'), "not supported"),
    (dict(method='barnes_hut'), "not supported"),
    (dict(perplexity=0), "perplexity"),
    (dict(learning_rate='fast'), "learning_rate"),
    (dict(learning_rate=-1.0), "learning_rate"),
    (dict(h2_rebuild_freq=0), "h2_rebuild_freq"),
    (dict(h2_leaf_size=0), "h2_leaf_size"),
    (dict(max_iter=0), "max_iter"),
    (dict(n_components=0), "n_components"),
    (dict(h2_precision='fp8'), "h2_precision"),
])
def test_validation_errors(kw, msg):
    with pytest.raises(ValueError, match=msg):
        TSNE(**kw)


def test_init_array_validation():
    with pytest.raises(ValueError):
        TSNE(init=np.zeros((10, 3)))
    with pytest.raises(ValueError, match="finite"):
        TSNE(init=np.full((10, 2), np.nan))


def test_learning_rate_auto_mapping():
    """lr auto = max(N/12, 200) (reference tests/test_tsne.py:705-710)."""
    t = TSNE()
    assert t._map_params_to_c_struct(n_samples=1200).learning_rate == 200.0
    assert t._map_params_to_c_struct(n_samples=120000).learning_rate == 10000.0
    assert TSNE(learning_rate=50.0)._map_params_to_c_struct(10).learning_rate == 50.0


def test_rebuild_early_is_at_least_rebuild():
    """reference tests/test_tsne.py:728-731."""
    p = TSNE(h2_rebuild_freq=20, h2_rebuild_freq_early=10)._map_params_to_c_struct(1000)
    assert p.h2_rebuild_freq_early == 20


def test_exact_method_mapping():
    p = TSNE(method='exact')._map_params_to_c_struct(1000)
    assert p.h2_rebuild_freq == 999999 and p.h2_tol == 1e-4


def test_schedule_constants():
    p = TSNE()._map_params_to_c_struct(1000)
    assert (p.early_exag, p.early_exag_iter, p.momentum, p.initial_momentum,
            p.momentum_switch_iter, p.sklearn_compatible) == (12.0, 250, 0.8, 0.5, 250, 1)


def test_validate_data():
    t = TSNE(perplexity=5)
    with pytest.raises(ValueError, match="2-dimensional"):
        t._validate_data(np.zeros(5))
    with pytest.raises(ValueError, match="n_samples"):
        t._validate_data(np.zeros((1, 3)))
    with pytest.raises(ValueError, match="perplexity"):
        t._validate_data(np.zeros((5, 3)))
    with pytest.raises(ValueError, match="NaN"):
        t._validate_data(np.full((10, 3), np.inf))


def test_pca_init_matches_sklearn():
    """reference tests/test_tsne.py:501-510."""
    from sklearn.decomposition import PCA
    rng = np.random.default_rng(0)
    X = rng.normal(size=(200, 10))
    Y = TSNE(random_state=3)._initial_embedding(X, 200)
    ref = PCA(n_components=2, svd_solver='randomized', random_state=3).fit_transform(X)
    ref = ref / np.std(ref[:, 0]) * 1e-4
    np.testing.assert_allclose(Y, ref, rtol=1e-10)


def test_precomputed_pca_falls_back_to_random():
    with pytest.warns(UserWarning):
        assert TSNE(metric='precomputed')._initial_embedding(None, 50) is None


def test_drop_self_vector():
    """reference tests/test_knn_methods.py:293-300: the point itself, wherever it sits,
    is removed; rows without it lose their last neighbour."""
    idx = np.array([[0, 3, 5], [2, 1, 4], [7, 8, 9]])
    dist = np.array([[0.0, 1.0, 2.0], [0.5, 0.0, 1.5], [1.0, 2.0, 3.0]])
    i2, d2 = _drop_self(idx, dist)
    assert i2.tolist() == [[3, 5], [2, 4], [7, 8]]
    assert d2.tolist() == [[1.0, 2.0], [0.5, 1.5], [1.0, 2.0]]


def test_precomputed_validation():
    from scipy.sparse import csr_matrix
    t = TSNE(metric='precomputed', perplexity=2)
    with pytest.raises(ValueError, match="sparse"):
        t.fit_transform(np.zeros((4, 4)))
    with pytest.raises(ValueError, match="square"):
        t.fit_transform(csr_matrix(np.ones((3, 4))))
    with pytest.raises(ValueError, match="more than perplexity"):
        t.fit_transform(csr_matrix(np.eye(4, k=1) + np.eye(4, k=-1)))


@pytest.mark.parametrize("kw,pattern", [
    # the fragments the reference's own tests match (tests/test_validation.py,
    # tests/test_tsne.py:104-185)
    (dict(metric='manhattan'), "metric=.*not supported"),
    (dict(init='spectral'), "init=.*not supported"),
    (dict(init=np.zeros(10)), "init"),
    (dict(init=np.zeros((50, 3))), "n_components"),
    (dict(method='invalid'), "method=.*not supported"),
    (dict(n_components=-10), "n_components must be >= 1"),
    (dict(perplexity=-30.0), "perplexity must be > 0"),
    (dict(learning_rate=0), "learning_rate must be > 0"),
    (dict(learning_rate=-200.0), "learning_rate must be > 0"),
    (dict(max_iter=-100), "max_iter must be >= 1"),
])
def test_reference_message_fragments(kw, pattern):
    with pytest.raises(ValueError, match=pattern):
        TSNE(**kw)


def test_reference_data_and_init_fragments():
    t = TSNE(perplexity=5)
    with pytest.raises(ValueError, match="n_samples=0 should be >= 2"):
        t._validate_data(np.zeros((0, 3)))
    with pytest.raises(ValueError, match="perplexity=.* must be less than"):
        t._validate_data(np.zeros((5, 3)))
    with pytest.raises(ValueError, match="contains NaN or inf"):
        t._validate_data(np.full((10, 3), np.nan))
    X = np.random.default_rng(0).normal(size=(50, 2))
    with pytest.raises(ValueError, match="n_components"):
        TSNE(init='pca', n_components=3, perplexity=5)._initial_embedding(X, 50)
    with pytest.raises(ValueError, match="shape"):
        TSNE(init=np.zeros((40, 2)))._initial_embedding(np.zeros((50, 5)), 50)
    with pytest.warns(UserWarning, match="init='random'"):
        TSNE(metric='precomputed')._initial_embedding(None, 50)
    with pytest.raises(NotImplementedError, match="fit\\(\\) is not implemented"):
        TSNE().fit(X)


def test_precomputed_short_rows_fragment():
    from scipy.sparse import csr_matrix
    with pytest.raises(ValueError, match="neighbors"):
        TSNE(metric='precomputed', perplexity=2).fit_transform(
            csr_matrix(np.eye(4, k=1) + np.eye(4, k=-1)))


def test_profiling_file_reaches_the_c_struct():
    t = TSNE(enable_profiling=True, profiling_output_file="/tmp/x.csv", random_state=7,
             n_iter_without_progress=123, verbose=2)
    p = t._map_params_to_c_struct(n_samples=5000)
    assert p.profiling_output_file == b"/tmp/x.csv" and p.enable_profiling == 1
    assert p.random_seed == 7 and p.convergence_patience == 123 and p.verbose == 1
    assert p.learning_rate == pytest.approx(5000 / 12.0) and p.print_freq == 50

"""The sklearn-style front end (TSNE.fit_transform) end to end on the device: exact k-NN
graph, joint P, optimisation.

knn_method='gpu' (tsne_knn_exact) builds the same graph as the reference's exact provider
(sklearn brute force), so on data without distance ties both runs are identical bit for
bit; the result is finite, its KL is reported, and 3D and mixed precision run too.
"""
import sys

import numpy as np
import pytest

from datasets import gaussian_mixture
from tsne_h2pack import TSNE

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def data():
    X, _ = gaussian_mixture(3000, dim=30, n_clusters=6, seed=21)
    return X


def test_gpu_knn_run_equals_exact_knn_run(backend, data):
    _say("tsne api: knn_method gpu vs exact")
    kw = dict(n_components=2, perplexity=20.0, random_state=0, max_iter=400, init='random')
    Y_gpu = TSNE(knn_method='gpu', **kw).fit_transform(data)
    Y_ref = TSNE(knn_method='exact', **kw).fit_transform(data)
    assert np.isfinite(Y_gpu).all()
    assert np.array_equal(Y_gpu, Y_ref)


@pytest.mark.parametrize("dim,precision", [(2, "mixed"), (3, "plain")])
def test_fit_transform_runs(backend, data, dim, precision):
    _say(f"tsne api: {dim}D {precision}")
    t = TSNE(n_components=dim, perplexity=20.0, random_state=1, max_iter=300, knn_method='gpu',
             h2_precision=precision)
    Y = t.fit_transform(data)
    assert Y.shape == (data.shape[0], dim) and np.isfinite(Y).all()
    assert t.kl_divergence_ is not None and 0.0 < t.kl_divergence_ < 10.0

"""tsne_h2pack on MI355X: sklearn-compatible t-SNE whose per-iteration gradient path
(H2 repulsive product and rebuild, CSR attractive force, perplexity bisection, fused
gains/momentum update) runs as hand-written gfx950 HIP kernels in libtsne_h2pack.so.

Public surface as the reference package (tsne_h2pack/__init__.py of
xinxing02/tsne-h2pack-python 3.3.1): TSNE, H2_DEFAULTS, h2_defaults.
"""
from .manifold import H2_DEFAULTS, TSNE, h2_defaults

__version__ = "3.3.1+mi355x.0"
__all__ = ["TSNE", "H2_DEFAULTS", "h2_defaults", "__version__"]

"""
mars_amd -- MI355X-native forward pass of Multivariate Adaptive Regression Splines.

    import mars_amd as mars
    model = mars.fit(X, y, max_terms=41)          # device route (cuda=True default)

`fit` / `marslib.MarsAlgo` keep the reference's API (mars.py:60, marslib.cc:152-179);
the scoring runs in hand-written HIP kernels for gfx950 behind the C ABI of
include/mars_hip.h (libmars_hip.so).
"""

from . import marslib  # noqa: F401  (raises ImportError if the native parts are missing)
from .mars import MODEL_DTYPE, compact, expand, fit, fit_batch, gram, pprint, prune  # noqa: F401

__version__ = marslib.__version__
__all__ = ["fit", "fit_batch", "expand", "gram", "prune", "compact", "pprint", "marslib", "MODEL_DTYPE",
           "__version__"]

"""
Loader of the native `marslib` module (pybind11 over libmars_hip.so).

Fails loudly when the extension has not been built: there is no fallback
implementation of the forward pass in this package.
"""

import glob
import os

_HERE = os.path.dirname(os.path.abspath(__file__))

try:
    from ._marslib import (  # noqa: F401
        Batch, MarsAlgo, __version__, default_tuning, device_bytes, expand_f32, expand_f64, gram_f32, gram_f64,
        prune_gram,
    )
except ImportError as e:  # pragma: no cover - exercised only on unbuilt trees
    raise ImportError(
        "mars_amd native extension is not built (expected %s and libmars_hip.so next to it); "
        "run `python -m mars_amd.build` or `python -c 'import __graft_entry__ as g; g.build()'`: %s"
        % (os.path.join(_HERE, "_marslib*.so"), e)
    ) from e


def library_path():
    """Absolute path of libmars_hip.so (the C ABI of include/mars_hip.h)."""
    p = os.path.join(_HERE, "libmars_hip.so")
    if not os.path.exists(p):
        raise FileNotFoundError(p)
    return p


def loaded_objects():
    return sorted(glob.glob(os.path.join(_HERE, "*.so")))

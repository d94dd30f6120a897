"""cuscan_amd: MI355X-native modern UMAP layout (HIP kernels behind a C ABI).

The package mirrors cuscan's layout operator surface and Python ``UMAP``
estimator; all numerical work runs in the HIP library ``libumaphip.so``.
"""
from . import _lib  # noqa: F401

__all__ = ["UMAP", "layout"]
__version__ = "0.1.0"


def __getattr__(name):
    if name == "UMAP":
        from .umap import UMAP

        return UMAP
    raise AttributeError(name)

"""ARBTools.ARBInterp: the same classes as ARBInterp.ARBInterp (the GPU
drop-in `tricubic` and `quadcubic` of arbinterp_amd)."""
from arbinterp_amd.ARBInterp import __version__, quadcubic, tricubic  # noqa: F401

__all__ = ["tricubic", "quadcubic", "__version__"]

"""The reference module ARBInterp.ARBInterp (src/ARBInterp/ARBInterp.py),
bound to the MI355X implementation: `tricubic` (3D, ARBInterp.py:8-621) and
`quadcubic` (4D, ARBInterp.py:627-1381) are the GPU drop-in classes of
arbinterp_amd.ARBInterp, with the reference's constructor, Query API,
attributes and error behaviour."""
from arbinterp_amd.ARBInterp import __version__, quadcubic, tricubic  # noqa: F401

__all__ = ["tricubic", "quadcubic", "__version__"]

"""MI355X-native Lekien-Marsden tricubic / quadcubic interpolation.

`from arbinterp_amd.ARBInterp import tricubic, quadcubic` is the drop-in for
the reference's `from ARBInterp.ARBInterp import tricubic, quadcubic`.
"""
from .ARBInterp import tricubic, quadcubic, __version__  # noqa: F401

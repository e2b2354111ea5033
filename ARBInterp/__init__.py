"""Import path of the reference package (src/ARBInterp/__init__.py).

`from ARBInterp import ARBInterp` (the reference's tests/test_ARBInterp.py:1)
then `ARBInterp.tricubic(...)` / `ARBInterp.quadcubic(...)` reach the MI355X
classes of arbinterp_amd.
"""

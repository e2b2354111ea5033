"""Import path the reference's examples use (`from ARBTools.ARBInterp import
tricubic`, examples/3D_ARBInterpExample.py:2; `quadcubic`,
examples/4D_ARBInterpExample.py:2): the package name of earlier ARBInterp
releases."""

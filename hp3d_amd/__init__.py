"""hp3d_amd -- MI355X (gfx950) element computation for hp3D hexahedra.

The HIP library (libhp3d_gpu.so, C ABI in include/hp3d_gpu.h) computes
element stiffness matrices and loads (H1 Poisson, real H(curl) curl-curl +
mass), static condensation of interior bubbles, back substitution, element
errors and adjoint coefficient gradients; this package is its host-side
mirror of hp3D's elem / stc_fwd / element-loop interface (elem.py), the
global numbering / assembly / solve around it (solve.py), the adjoint
objective (adjoint.py) and synthetic meshes carrying hp3D's order and
orientation codes (mesh.py).
"""
from . import mesh  # noqa: F401
from ._lib import (BF16, FIELD_ADJOINT, FIELD_ERROR, FIELD_L2, FP32, FP64, H1, HCURL,  # noqa: F401
                   LOAD_CONV, LOAD_CURL_SIN, LOAD_ONE, LOAD_POLY2, LOAD_SINSINSIN, LOAD_ZERO,
                   Hp3dError, init)
from .elem import ElementBatch, celem, elem, ndof, stc_fwd  # noqa: F401

"""numbacs.diagnostics on the GPU path (diagnostics.py:21-269 of the
reference)."""
from numbacs_amd.diagnostics import (  # noqa: F401
    C_eig_2D,
    C_eig_aux_2D,
    C_tensor_2D,
    ftle_from_eig,
    ftle_grid_2D,
)

from ._outside import outside

__getattr__ = outside(__name__)

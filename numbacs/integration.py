"""numbacs.integration on the GPU path (integration.py:7-61, 123-182,
249-464, 609-737 of the reference)."""
from numbacs_amd.integration import (  # noqa: F401
    flowmap,
    flowmap_aux_grid_2D,
    flowmap_composition,
    flowmap_composition_initial,
    flowmap_composition_step,
    flowmap_grid_2D,
)

from ._outside import outside

__getattr__ = outside(__name__)

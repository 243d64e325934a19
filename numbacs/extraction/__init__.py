"""numbacs.extraction on the GPU path: the FTLE ridge functions
(extraction/ridges.py:9-229, 720-1054 of the reference)."""
from numbacs_amd.extraction import ftle_ordered_ridges, ftle_ridge_pts, ftle_ridges  # noqa: F401

from . import ridges  # noqa: F401
from .._outside import outside

__getattr__ = outside(__name__)

"""numbacs.extraction.ridges on the GPU path."""
from numbacs_amd.extraction.ridges import (  # noqa: F401
    ftle_ordered_ridges,
    ftle_ridge_pts,
    ftle_ridges,
)

from .._outside import outside

__getattr__ = outside(__name__)

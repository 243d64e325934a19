"""Ridge extraction (mirrors numbacs.extraction for the FTLE ridge path)."""
from .ridges import ftle_ordered_ridges, ftle_ridge_pts, ftle_ridges, label8

__all__ = ["ftle_ridge_pts", "ftle_ridges", "ftle_ordered_ridges", "label8"]

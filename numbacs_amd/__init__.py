"""numbacs_amd: the FTLE hot path of NumbaCS (flow maps, FTLE, FTLE ridges)
on AMD Instinct MI355X (gfx950), behind the reference's Python entry points.

Import paths mirror the reference package:
    numbacs_amd.flows        get_predefined_flow, get_interp_arrays_2D, get_flow_2D
    numbacs_amd.integration  flowmap_grid_2D, flowmap
    numbacs_amd.diagnostics  ftle_grid_2D, C_eig_2D, ftle_from_eig
    numbacs_amd.extraction   ftle_ridge_pts, ftle_ridges
"""
__version__ = "0.1.0"

from . import diagnostics, extraction, flows, integration  # noqa: F401

"""``numbacs`` import path for the MI355X FTLE path.

Code written against the reference package (``from numbacs.integration
import flowmap_grid_2D`` and the like, src/numbacs/__init__.py and its
submodules) runs unchanged: each submodule below resolves the reference's
names on the FTLE path to the GPU implementations in :mod:`numbacs_amd`.

    numbacs.flows        get_predefined_flow, get_interp_arrays_2D, get_flow_2D,
                         get_flow_linear_2D
    numbacs.integration  flowmap_grid_2D, flowmap, flowmap_aux_grid_2D,
                         flowmap_composition, flowmap_composition_initial,
                         flowmap_composition_step
    numbacs.diagnostics  ftle_grid_2D, C_eig_2D, ftle_from_eig, C_tensor_2D,
                         C_eig_aux_2D
    numbacs.extraction   ftle_ridge_pts, ftle_ridges, ftle_ordered_ridges

Other names of the reference package raise ImportError.
"""
from numbacs_amd import __version__  # noqa: F401

from . import diagnostics, extraction, flows, integration  # noqa: F401

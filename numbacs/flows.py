"""numbacs.flows on the GPU path (flows.py:9-46, 121-258, 418-506,
1104-1295 of the reference)."""
from numbacs_amd.flows import (  # noqa: F401
    get_flow_2D,
    get_flow_linear_2D,
    get_interp_arrays_2D,
    get_predefined_flow,
)

from ._outside import outside

__getattr__ = outside(__name__)

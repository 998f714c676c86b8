"""fastearth3d_amd: MI355X-native VILMA v2 viscoelastic time step."""
from ._abi import VhParams  # noqa: F401
from .namelist import vilma_par_load, params_defaults  # noqa: F401

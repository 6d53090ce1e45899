"""MI355X-native compressible fluctuating-hydrodynamics RK3 step (FHDeX src_compressible).

Host-side mirror of the reference interface over the C ABI of ``include/fhdx.h``.
"""
from .params import Params, Fab, make_params, argon_gas, four_species, species_prefix  # noqa: F401

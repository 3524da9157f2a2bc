"""inmsom_bt: MI355X implementation of INMSOM's explicit leap-frog
barotropic step (host side; kernels in ../csrc, C ABI in include/inmsom_bt.h)."""
from .cases import Basin, make_basin, randomize_state  # noqa: F401
from .model import BarotropicModel, gridcon, rayleigh  # noqa: F401

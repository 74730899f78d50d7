"""MI355X-native HBT pair-correlation engine (hadronic_afterburner_toolkit's
HBT_correlation path) -- HIP kernels for gfx950 behind a C ABI
(include/hbt_hip.h), with a Python mirror of the reference interface."""
from . import samples  # noqa: F401
from ._lib import HBT_KERNEL_FAST, HBT_KERNEL_PLAIN, HBT_PREC_FP32C, HBT_PREC_FP64  # noqa: F401

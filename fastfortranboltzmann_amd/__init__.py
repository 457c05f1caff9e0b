"""MI355X-native (gfx950) D2Q9 pull-stream-collide BGK solver.

The compute path is libffb_hip.so (hand-written HIP, C ABI in include/ffb.h);
this package is its host-side mirror of the reference's single-domain
interface (see lattice.py).
"""
from ._lib import (COUETTE_FLOW, POISEUILLE_FLOW, PREC_BF16_DEV, PREC_E3M12_DEV, PREC_FP16_DEV, PREC_FP32,
                   PREC_FP32_DEV, PREC_FP64, SHEAR_WAVE, SLIDING_LID, FfbError)
from .lattice import Lattice, Settings, omega_for_reynolds, reynolds, version

__all__ = ["Lattice", "Settings", "FfbError", "reynolds", "omega_for_reynolds", "version",
           "SHEAR_WAVE", "COUETTE_FLOW", "POISEUILLE_FLOW", "SLIDING_LID",
           "PREC_FP32", "PREC_FP64", "PREC_FP16_DEV", "PREC_BF16_DEV", "PREC_FP32_DEV", "PREC_E3M12_DEV"]

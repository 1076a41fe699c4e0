"""profess_amd: MI355X-native PROFESS energy/potential evaluation and NTN
density optimizer (HIP for gfx950) behind the reference's call surface.

Importing the package does not touch the GPU; the first call that needs the
library loads ``libprofess_amd.so`` and raises ``OFDError`` if it is missing.
"""
from ._lib import OFDError, lib, LIB_PATH  # noqa: F401
from .calpot import (CalculatePotentialPlus, CalculatePotentialPlusDevice, ConjugateGradient, EwaldEnergy,  # noqa: F401
                     IonElectronForcesSpline,
                     OFDSystem, ProjNormRhoMinimization, request_bytes, ETABLE_NAMES)

__all__ = ["OFDError", "OFDSystem", "CalculatePotentialPlus", "CalculatePotentialPlusDevice",
           "ProjNormRhoMinimization", "IonElectronForcesSpline", "EwaldEnergy", "ConjugateGradient", "request_bytes", "ETABLE_NAMES"]

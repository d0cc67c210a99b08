"""MI355X-native DEM particle hot path of MFiX (Python plumbing).

`abi` mirrors include/mfixdem.h, `system.DemSystem` drives one DEM system
through the C ABI, `workloads` builds the seeded synthetic inputs.
"""
from . import abi  # noqa: F401

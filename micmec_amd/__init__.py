"""micmec_amd: MI355X-native micromechanical force field and integrators (FP64, HIP).

Drop-in for MicMec's hot path (molmod/micmec): ``System``, ``ForcePartMechanical``,
``MicMecForceField``, ``VerletIntegrator`` with NHC/MTK hooks and ``CGOptimizer``, with the
arithmetic in hand-written HIP kernels for gfx950 behind the C ABI of ``libmicmec_amd.so``.
"""

from . import units
from .system import System, Domain
from .topology import Grid, build_system, build_type

__all__ = ["units", "System", "Domain", "Grid", "build_system", "build_type"]
__version__ = "0.1.0"

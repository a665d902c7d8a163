"""nefi_amd — nefi's acoustic wave leapfrog propagator on AMD MI355X (gfx950).

A drop-in for ``nefi.physics.wave`` (``WaveOperator``, ``WaveInitialConditionOperator``,
``time_reversal`` and their helpers) whose time loop, fused step and hand-written adjoint run as
HIP kernels in ``nefi_amd/_lib/libnefi_wave.so`` behind the C ABI of ``include/nefi_wave.h``.
"""

from . import errors, registry
from .domain import Domain

__version__ = "0.1.0"

__all__ = ["Domain", "errors", "registry", "__version__"]

from .timestep import (
    auto_checkpoint_every,
    laplacian_eig_max,
    run_timestepping,
    stable_dt_diffusion,
    stable_dt_wave,
)
from .wave import (
    ABSORBING,
    GRAD_MODES,
    WaveGrid,
    WaveInitialConditionOperator,
    WaveOperator,
    laplacian,
    laplacian_kernel,
    ricker,
    stable_dt,
    time_reversal,
    wave_energy,
)

__all__ = [
    "ABSORBING",
    "GRAD_MODES",
    "WaveGrid",
    "WaveInitialConditionOperator",
    "WaveOperator",
    "auto_checkpoint_every",
    "laplacian",
    "laplacian_eig_max",
    "laplacian_kernel",
    "ricker",
    "run_timestepping",
    "stable_dt",
    "stable_dt_diffusion",
    "stable_dt_wave",
    "time_reversal",
    "wave_energy",
]

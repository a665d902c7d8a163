from .base import Fields, Operator
from .timestepping import TimeStepper, laplacian, stable_dt, wave_initial_state, wave_step

__all__ = ["Fields", "Operator", "TimeStepper", "laplacian", "stable_dt", "wave_initial_state",
           "wave_step"]

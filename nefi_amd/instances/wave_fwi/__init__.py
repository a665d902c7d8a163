"""wave_fwi's trace misfit (``nefi/instances/wave_fwi/__init__.py:47``)."""

from .losses import BandLimitedMSE, lowpass

__all__ = ["BandLimitedMSE", "lowpass"]

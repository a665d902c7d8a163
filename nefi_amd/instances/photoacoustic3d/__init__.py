"""photoacoustic3d's sensor model (``nefi/instances/photoacoustic3d/sensor.py``)."""

from .sensor import SensorBandwidth, gaussian_lowpass

__all__ = ["SensorBandwidth", "gaussian_lowpass"]

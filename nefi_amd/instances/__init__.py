"""The parts of nefi's two wave instances that sit on the trace path.

``wave_fwi``: the band-limited misfit and its low-pass (``nefi/instances/wave_fwi/losses.py``);
``photoacoustic3d``: the sensors' Gaussian low-pass operator
(``nefi/instances/photoacoustic3d/sensor.py``). Both run on the HIP trace filters of
``nefi_amd.physics.filters``; the instances' acquisition, scenes and solver loops are nefi's own.
"""

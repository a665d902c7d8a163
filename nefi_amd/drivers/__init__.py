"""Drivers of whole workloads built on the propagator (the neural-field FWI of config 5)."""

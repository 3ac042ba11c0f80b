"""psx -- Python binding of the MI355X backend of particle-stabilizer's per-step hot path.

The compute path is libparticle_stabilizer_hip.so (C ABI: include/psx.h); this package only
builds inputs, calls the C ABI through ctypes and converts arrays.
"""
from ._lib import PsxError, load
from .sim import Batch, Params, Scene, Simulation, mass_of, run_simulate

__all__ = ["Batch", "Params", "PsxError", "Scene", "Simulation", "load", "mass_of", "run_simulate"]

"""MI355X-native BEM mat-vec for Poisson-Boltzmann solvation (the GMRES hot
path of libBEEP).  Host scene code and HIP kernels live in ``libbeep3.so``
(include/beep_host.h, include/beep_hip.h); this package is the Python view.
"""
from .beep import BEEP
from .device import Device
from .scene import Mesh, Scene

__all__ = ["BEEP", "Device", "Mesh", "Scene"]

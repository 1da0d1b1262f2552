"""photonoxide_amd: an MI355X-native FDTD time-stepper behind photonoxide's FDTD interface.

The compute path is hand-written HIP for gfx950 in ``csrc/`` behind a C ABI
(``include/pxfdtd.h``); this package binds it with ctypes and mirrors the reference's
``fdtd`` API in :mod:`photonoxide_amd.fdtd`.
"""
from .fdtd import (Axis, Box, Boundaries, Cpml, Current, Cylinder, Design, Dipole, Field, GaussianBeam, Gpu, Grid3d,
                   Periodic, PlaneWave, Pml, PortMode, Precision, Resonance, Scene, Simulation, Source, Term, Waveform,
                   harmonic_inversion)

__all__ = ["Axis", "Box", "Boundaries", "Cpml", "Current", "Cylinder", "Design", "Dipole", "Field", "GaussianBeam", "Gpu",
           "Grid3d", "Periodic", "PlaneWave", "Pml", "PortMode", "Precision", "Resonance", "Scene", "Simulation", "Source", "Term",
           "Waveform", "harmonic_inversion"]

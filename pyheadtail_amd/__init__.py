"""pyheadtail_amd: PyHEADTAIL's per-turn macroparticle path on MI355X.

Drop-in classes with the reference's names and constructor signatures
(Particles, UniformBinSlicer, TransverseMap/TransverseSegmentMap,
Chromaticity, AmplitudeDetuning, LinearMap, RFSystems, WakeField, Resonator
and ResistiveWall families) backed by the HIP library libphx.so, plus the
fused engine FusedOneTurnMap.
"""
__version__ = "0.1.0"

from .particles.particles import Particles  # noqa: F401

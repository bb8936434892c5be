"""virtualfluids_amd -- MI355X-native D3Q27 lattice-Boltzmann stepper.

The product is ``libvfx.so`` (hand-written HIP for gfx950 behind the C ABI in
``include/vfx.h``).  This package is its Python host binding: ``Lattice``,
``BoundaryCondition`` and ``Stepper`` hold device arrays (PyTorch tensors) and
call the library; nothing here computes populations on the CPU.
"""
from ._lib import VfxError, check, lib  # noqa: F401
from .batch import Ensemble, Graph, capture, cavity_ensemble  # noqa: F401
from .refinement import ICells, NestedBox, RefinedStepper, fine_omega, scale_cf, scale_fc, scale_fc_cf  # noqa: F401
from .lattice import (BYTES, CX, CY, CZ, INV, W, BoundaryCondition, Lattice, PrecursorPlanes, Statistics,  # noqa: F401
                      Stepper, FusedGroup, FusedPlan, MeanValues, BCSet, step_levels, pack_mask)
from .precursor_files import VTKFile, VTKFileCollection, VTKReader, precursor_from_files  # noqa: F401

__all__ = ["Ensemble", "Graph", "capture", "cavity_ensemble", "ICells", "NestedBox", "RefinedStepper", "fine_omega", "scale_cf",
           "scale_fc", "scale_fc_cf", "Lattice", "BoundaryCondition", "PrecursorPlanes", "Stepper", "FusedGroup", "FusedPlan", "MeanValues", "BCSet", "step_levels", "Statistics", "VfxError", "lib", "check", "pack_mask",
           "VTKFile", "VTKFileCollection", "VTKReader", "precursor_from_files",
           "CX", "CY", "CZ", "INV", "W", "BYTES"]

"""MI355X-native NNPP time step (nonlocal Nernst-Planck-Poisson corrosion of a Mg implant).

Python host mirror over libnnpp.so's C ABI (include/nnpp.h). The compute path is the HIP
library; importing this package loads it and fails if it is missing.
"""
from . import _abi
from ._abi import (CHEM_GENERIC, CHEM_REFERENCE, FORMAT_AUTO, FORMAT_NO_BALL,
                   FORMAT_NO_ZMARCH, FORMAT_NO_DICTIONARY, FORMAT_NO_STENCIL, FORMAT_WIDE_COLUMNS, FP64, MIXED, SLOT_N, SLOT_N1,
                   NnppError, Params, StepStats)
from .problem import Problem
from . import synth, formats

lib = _abi.load()

__all__ = ["CHEM_GENERIC", "CHEM_REFERENCE", "FORMAT_AUTO", "FORMAT_NO_BALL", "FORMAT_NO_ZMARCH",
           "FORMAT_NO_DICTIONARY",
           "FORMAT_NO_STENCIL", "FORMAT_WIDE_COLUMNS", "FP64", "MIXED", "SLOT_N", "SLOT_N1",
           "NnppError", "Params", "StepStats", "Problem", "synth", "formats", "lib", "autograd"]


def __getattr__(name):  # the torch binding is imported on first use
    if name == "autograd":
        import importlib
        return importlib.import_module(__name__ + ".autograd")
    raise AttributeError(name)

"""MI355X-native trace transform: sinograms and circus functions on gfx950.

A drop-in for the reference C++ implementation's getSinograms /
getCircusFunctions path (c++/src/sinogram.hpp:57-59,
c++/src/circus.hpp:69-71).  The compute runs in hand-written HIP kernels of
``libtt_amd.so`` behind the C ABI of ``include/tt.h``; this package is the
Python host mirror of the reference's interface over torch device memory.
"""
from .functionals import (PFunctional, PFunctionalArguments, PFunctionalWrapper,
                          TFunctional, TFunctionalArguments, TFunctionalWrapper)
from . import inputs

__all__ = [
    "TFunctional", "PFunctional", "TFunctionalWrapper", "PFunctionalWrapper",
    "TFunctionalArguments", "PFunctionalArguments", "inputs",
    "getSinograms", "getCircusFunctions", "Transformer",
]


def __getattr__(name):
    # The torch-backed API is imported lazily so that the enums, parsers and
    # input generators stay usable without importing torch.
    import importlib
    if name in ("getSinograms", "getCircusFunctions", "Transformer"):
        return getattr(importlib.import_module(".transform", __name__), name)
    if name in ("core", "transform"):
        return importlib.import_module("." + name, __name__)
    raise AttributeError(name)

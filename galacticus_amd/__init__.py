"""MI355X-native batched node-ODE evolution for Galacticus.

The product is the HIP library ``galacticus_amd/lib/libgalacticus_amd.so``
(C ABI in ``include/gal_ode.h``); this package is its host side: parameter
handling in the reference's XML vocabulary (``params``), the batched
node-evolver interface (``evolver``), workload generators (``workloads``),
the cooling table (``cooling``), tree building and wave scheduling
(``trees``) and the learned surrogate (``surrogate``).
"""
from . import params  # noqa: F401

__all__ = ["params", "cooling", "workloads", "evolver"]

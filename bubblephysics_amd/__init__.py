"""MI355X-native world step of BubblePhysics 4.0.0 (HIP kernels behind a C ABI).

The product is the HIP library ``lib/libbubblephysics_hip.so`` with the C ABI declared in
``include/bubblephysics.h``; this package holds its ctypes binding (:mod:`abi`), the solver-level
mirror of the reference interface (:mod:`solver`) and synthetic scene builders (:mod:`scenes`).
"""
from . import abi  # noqa: F401

__all__ = ["abi"]

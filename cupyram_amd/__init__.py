"""MI355X-native split-step Pade parabolic-equation range march.

Public names follow cupyram/__init__.py:15-19: ``CuPyRAM``,
``compute_pade_coefficients`` and ``AsyncBatcher``.
"""
from .batching import AsyncBatcher
from .cupyram import CuPyRAM
from .pade import compute_pade_coefficients

__version__ = "0.1.0"
__all__ = ["CuPyRAM", "compute_pade_coefficients", "AsyncBatcher", "__version__"]

"""Learned TL surrogate: a 1-D Fourier-operator network over depth trained on
the PE march's own TL fields, an ensemble router with PE fallback, and the
query domains it is trained on."""
from .encode import Domain, Queries, encode, normalise_tl, denormalise_tl  # noqa: F401
from .fno import FNO, FNOConfig, dft_matrices  # noqa: F401

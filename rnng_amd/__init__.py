"""rnng_amd: MI355X (gfx950) implementation of RNNG's epsilon-neighbor-graph path.

The compute lives in librnng_amd.so (hand-written HIP kernels behind the C ABI
of include/rnng_amd.h); this package is its Python binding plus the seeded
input generators.  There is no CPU fallback.
"""
from ._lib import MODE_MFMA, MODE_PLAIN, RnngError, Stats, build, declared_symbols, lib  # noqa: F401

__all__ = ["MODE_MFMA", "MODE_PLAIN", "RnngError", "Stats", "build", "declared_symbols", "lib"]

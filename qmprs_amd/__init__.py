"""qmprs_amd -- MI355X-native Sequential MPS state-preparation encoder.

Drop-in for the reference package's encoding path: ``qmprs.primitives.MPS``
becomes ``qmprs_amd.primitives.MPS`` and
``qmprs.synthesis.mps_encoding.{MPSEncoder, Sequential}`` becomes
``qmprs_amd.synthesis.mps_encoding.{MPSEncoder, Sequential}``.  The arithmetic
runs in hand-written HIP kernels (libqmx.so) on one AMD GPU.
"""

__all__ = ["primitives", "synthesis", "circuit", "engine"]

from qmprs_amd import circuit, engine, primitives, synthesis  # noqa: E402

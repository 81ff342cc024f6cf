__all__ = ["MPSEncoder", "Sequential"]

from qmprs_amd.synthesis.mps_encoding.base import MPSEncoder
from qmprs_amd.synthesis.mps_encoding.sequential import Sequential

__all__ = ["mps_encoding"]

from qmprs_amd.synthesis import mps_encoding

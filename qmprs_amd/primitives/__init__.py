__all__ = ["MPS", "Tensor", "UnitaryBlock", "UnitaryLayer"]

from qmprs_amd.primitives.mps import MPS, Tensor, UnitaryBlock, UnitaryLayer

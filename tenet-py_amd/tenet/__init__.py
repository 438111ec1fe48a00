"""tenet (MI355X build): block-sparse symmetric tensors whose blocks live in HBM, contracted
by hand-written gfx950 kernels behind the reference TeNeT-py API (``SymmetricTensor``,
``tenet.einsum``, ``tenet.linalg``, ``tenet.network``).
"""

from tenet import linalg, symmetry
from tenet.dense import from_dense, to_dense
from tenet.engine import gemm_precision, set_gemm_precision
from tenet.fusion_tree import FusionTree, coupled_sectors, fusion_trees
from tenet.ops import (
    block_sqrt,
    add,
    adjoint,
    allclose,
    astype,
    compose,
    conj,
    divide,
    einsum,
    inner,
    multiply,
    norm,
    repartition,
    subtract,
    tensordot,
    transpose,
)
from tenet.plans import map_layout
from tenet.serialize import load, save
from tenet.space import IN, OUT, GradedSpace, Leg, Side
from tenet.structure import FusionBlockKey, TensorStructure
from tenet.symmetry import CapabilityError, StructureChangingError
from tenet.tensor import SymmetricTensor

__all__ = [
    "IN",
    "OUT",
    "CapabilityError",
    "FusionBlockKey",
    "FusionTree",
    "GradedSpace",
    "Leg",
    "Side",
    "StructureChangingError",
    "SymmetricTensor",
    "TensorStructure",
    "add",
    "adjoint",
    "allclose",
    "astype",
    "block_sqrt",
    "compose",
    "conj",
    "coupled_sectors",
    "divide",
    "einsum",
    "from_dense",
    "fusion_trees",
    "gemm_precision",
    "inner",
    "linalg",
    "load",
    "map_layout",
    "multiply",
    "norm",
    "repartition",
    "save",
    "set_gemm_precision",
    "subtract",
    "symmetry",
    "tensordot",
    "to_dense",
    "transpose",
]

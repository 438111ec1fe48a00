"""``SymmetricTensor``: a frozen structure plus ONE flat device buffer of blocks.

The reference stores a tuple of backend arrays, one per key of ``block_order``
(reference ``tensor.py:62-112``).  Here the blocks are laid out back to back in that
order in a single 1-D torch tensor (C-contiguous blocks in public axis order), which is
what lets one kernel launch move or contract every block of a tensor.  ``blocks`` still
returns one view per key, so code that reads ``t.blocks[i]`` keeps working; a tensor can
also be built from a tuple of blocks exactly as in the reference.
"""

from __future__ import annotations

from collections.abc import Sequence

import numpy as np
import torch

from tenet.space import Leg
from tenet.structure import TensorStructure

__all__ = ["SymmetricTensor", "default_device"]


def default_device() -> torch.device:
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


_TORCH_DTYPES = {
    np.dtype(np.float64): torch.float64,
    np.dtype(np.float32): torch.float32,
    torch.float64: torch.float64,
    torch.float32: torch.float32,
}


def _torch_dtype(dtype) -> torch.dtype:
    if dtype is None:
        return torch.float64
    if isinstance(dtype, torch.dtype):
        key = dtype
    else:
        key = np.dtype(dtype)
    if key not in _TORCH_DTYPES:
        raise TypeError(f"unsupported block dtype {dtype!r}; tenet blocks are float64 or float32")
    return _TORCH_DTYPES[key]


class SymmetricTensor:
    """A block-sparse symmetric tensor: ``structure`` + flat block buffer ``data``.

    Parameters
    ----------
    structure : TensorStructure
        The legs and everything derived from them.
    blocks : sequence of arrays, or a 1-D torch tensor
        Either one array per key of ``structure.block_order`` (numpy or torch, shaped as
        ``structure.block_shape(key)``), copied into a flat buffer; or the flat buffer
        itself (length ``structure.size``), adopted without copying.
    """

    __slots__ = ("structure", "data")

    def __init__(self, structure: TensorStructure, blocks, *, device=None, dtype=None) -> None:
        if isinstance(blocks, torch.Tensor) and blocks.dim() == 1:
            data = blocks
            if data.numel() != structure.size:
                raise ValueError(
                    f"flat buffer has {data.numel()} elements, structure needs {structure.size}"
                )
        else:
            blocks = tuple(blocks)
            n = structure.num_blocks
            if len(blocks) != n:
                raise ValueError(f"structure has {n} blocks, got {len(blocks)}")
            shapes = structure.shapes
            parts = []
            for i, b in enumerate(blocks):
                arr = b if isinstance(b, torch.Tensor) else torch.as_tensor(np.asarray(b))
                if tuple(arr.shape) != tuple(shapes[i]):
                    raise ValueError(
                        f"block {i} has shape {tuple(arr.shape)}, key needs {tuple(shapes[i])}"
                    )
                parts.append(arr.reshape(-1))
            if dtype is None:
                dtype = parts[0].dtype if parts else torch.float64
            dt = _torch_dtype(dtype)
            dev = torch.device(device) if device is not None else (
                parts[0].device if parts else default_device()
            )
            data = (
                torch.cat([p.to(device=dev, dtype=dt) for p in parts])
                if parts
                else torch.zeros(0, dtype=dt, device=dev)
            )
        if data.dtype not in (torch.float64, torch.float32):
            raise TypeError(f"tenet blocks are float64 or float32, got {data.dtype}")
        self.structure = structure
        self.data = data.contiguous()

    # ---------------------------------------------------------------- constructors
    @classmethod
    def from_flat(cls, structure: TensorStructure, data: torch.Tensor) -> "SymmetricTensor":
        return cls(structure, data)

    @classmethod
    def zeros(cls, legs: Sequence[Leg], *, dtype=None, device=None) -> "SymmetricTensor":
        s = TensorStructure(tuple(legs))
        dev = torch.device(device) if device is not None else default_device()
        return cls(s, torch.zeros(s.size, dtype=_torch_dtype(dtype), device=dev))

    @classmethod
    def random(cls, legs: Sequence[Leg], *, seed: int | None = None, dtype=None, device=None):
        """Standard-normal blocks drawn from ``numpy.random.default_rng(seed)``, one block at
        a time in block order (reference ``tensor.py:232-262``), so the values are the
        reference's own for the same seed."""
        s = TensorStructure(tuple(legs))
        rng = np.random.default_rng(seed)
        flat = np.empty(s.size, dtype=np.float64)
        lay = s.layout
        for i, shape in enumerate(s.shapes):
            o = int(lay.offsets[i])
            flat[o : o + int(lay.sizes[i])] = rng.standard_normal(shape).reshape(-1)
        dev = torch.device(device) if device is not None else default_device()
        return cls(s, torch.from_numpy(flat).to(device=dev, dtype=_torch_dtype(dtype)))

    @classmethod
    def from_dense(cls, dense, legs: Sequence[Leg], *, atol: float | None = 1e-10, dtype=None,
                   device=None) -> "SymmetricTensor":
        """Project a dense array onto the symmetric subspace of ``legs`` (reference
        ``tensor.py:1053``); refuses a non-symmetric array (``atol=None`` projects)."""
        from tenet.dense import from_dense

        return from_dense(dense, legs, atol=atol, dtype=_torch_dtype(dtype), device=device)

    def to_dense(self) -> np.ndarray:
        """Host float64 dense array: ``sum_tau A^(tau) (x) C^(tau)`` (reference ``tensor.py:1028``)."""
        from tenet.dense import to_dense

        return to_dense(self)

    # ---------------------------------------------------------------- views
    @property
    def legs(self) -> tuple:
        return self.structure.legs

    @property
    def ndim(self) -> int:
        return self.structure.ndim

    @property
    def provider(self):
        return self.structure.provider

    @property
    def dtype(self) -> torch.dtype:
        return self.data.dtype

    @property
    def device(self) -> torch.device:
        return self.data.device

    @property
    def codomain(self) -> tuple:
        return tuple(self.legs[i] for i in self.structure.out_axes)

    @property
    def domain(self) -> tuple:
        return tuple(self.legs[i] for i in self.structure.in_axes)

    @property
    def blocks(self) -> tuple:
        """One C-contiguous view per key of ``block_order``."""
        lay = self.structure.layout
        return tuple(
            self.data[int(o) : int(o) + int(n)].view(tuple(shape))
            for o, n, shape in zip(lay.offsets, lay.sizes, self.structure.shapes)
        )

    def items(self):
        return zip(self.structure.block_order, self.blocks)

    def block(self, key):
        return self.blocks[self.structure.index_of(key)]

    def numpy_blocks(self) -> tuple:
        """Host copies of every block, as float64 numpy arrays."""
        flat = self.data.detach().to("cpu", torch.float64).numpy()
        lay = self.structure.layout
        return tuple(
            flat[int(o) : int(o) + int(n)].reshape(shape)
            for o, n, shape in zip(lay.offsets, lay.sizes, self.structure.shapes)
        )

    @property
    def shape(self) -> tuple:
        return tuple(leg.space.dim for leg in self.legs)

    @property
    def reduced_shape(self) -> tuple:
        return tuple(leg.space.reduced_dim for leg in self.legs)

    def with_data(self, data: torch.Tensor) -> "SymmetricTensor":
        return SymmetricTensor(self.structure, data)

    def copy(self) -> "SymmetricTensor":
        return SymmetricTensor(self.structure, self.data)

    def clone(self) -> "SymmetricTensor":
        return SymmetricTensor(self.structure, self.data.clone())

    def to(self, device=None, dtype=None) -> "SymmetricTensor":
        """Move/cast the flat buffer (torch copy; casting between float32/64 on the device
        runs the library's cast kernel)."""
        from tenet import ops

        out = self
        if device is not None and torch.device(device) != self.device:
            out = SymmetricTensor(self.structure, self.data.to(device))
        if dtype is not None and _torch_dtype(dtype) != out.dtype:
            out = ops.astype(out, _torch_dtype(dtype))
        return out

    def to_backend(self, backend: str = "torch") -> "SymmetricTensor":
        if backend != "torch":
            raise ValueError("blocks live in torch-ROCm device memory only")
        return self

    # ---------------------------------------------------------------- algebra
    def __mul__(self, s):
        from tenet import ops

        return ops.multiply(self, s)

    __rmul__ = __mul__

    def __truediv__(self, s):
        from tenet import ops

        return ops.divide(self, s)

    def __add__(self, other):
        from tenet import ops

        return ops.add(self, other)

    def __sub__(self, other):
        from tenet import ops

        return ops.subtract(self, other)

    def __neg__(self):
        from tenet import ops

        return ops.multiply(self, -1.0)

    def __matmul__(self, other):
        from tenet import ops

        return ops.compose(self, other)

    def transpose(self, axes=None):
        from tenet import ops

        return ops.transpose(self, axes)

    def repartition(self, outputs, inputs):
        from tenet import ops

        return ops.repartition(self, outputs, inputs)

    def __repr__(self) -> str:
        return (
            f"SymmetricTensor(ndim={self.ndim}, blocks={self.structure.num_blocks}, "
            f"size={self.structure.size}, dtype={self.dtype}, device={self.device})"
        )

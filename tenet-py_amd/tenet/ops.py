"""The public operations on device-resident ``SymmetricTensor``s.

Every function keeps the reference's signature and refusals (reference
``ops/contraction.py``, ``ops/permutation.py``, ``ops/repartition.py``, ``ops/map.py``,
``ops/basic.py``); planning happens on the host (``tenet.plans``), execution in the gfx950
kernels (``tenet.engine``).  Nothing here touches a block on the CPU.
"""

from __future__ import annotations

import ctypes
from collections import Counter
from dataclasses import replace
from functools import lru_cache

import numpy as np
import torch

from tenet import _native
from tenet.engine import (
    RecoupleProgram,
    contraction_exec,
    inplace_map,
    recouple_program,
    require_device,
    upload,
    wait_upload,
)
from tenet.plans import (
    contraction_plan,
    permutation_plan,
    repartition_plan,
    validated_axes,
    validated_contraction_axes,
    validated_partition,
)
from tenet.space import IN, OUT
from tenet.structure import FusionBlockKey, TensorStructure
from tenet.tensor import SymmetricTensor

__all__ = [
    "transpose",
    "repartition",
    "tensordot",
    "einsum",
    "compose",
    "adjoint",
    "inner",
    "norm",
    "add",
    "subtract",
    "multiply",
    "divide",
    "conj",
    "astype",
    "allclose",
]


def _dt(t: SymmetricTensor) -> int:
    return _native.TN_F64 if t.dtype == torch.float64 else _native.TN_F32


def _apply_plan(t: SymmetricTensor, new: TensorStructure, perm, terms) -> SymmetricTensor:
    """Execute a block map: share the buffer when it is a no-op, else one recouple."""
    offs = inplace_map(t.structure, new, perm, terms)
    if offs is not None and np.array_equal(offs, new.layout.offsets):
        return SymmetricTensor(new, t.data)
    require_device(t)
    prog = recouple_program(t.structure, new, tuple(perm), tuple(terms))
    out = torch.empty(new.size, dtype=t.dtype, device=t.device)
    prog.run(t.data, out)
    return SymmetricTensor(new, out)


# ----------------------------------------------------------------------------- permutations


def transpose(t: SymmetricTensor, axes=None) -> SymmetricTensor:
    """Reorder public axes (NumPy convention); braids pay the provider's coefficients
    (reference ``ops/permutation.py:200-272``)."""
    axes = tuple(reversed(range(t.ndim))) if axes is None else validated_axes(t.ndim, axes)
    plan = permutation_plan(t.structure, axes)
    return _apply_plan(t, plan.new_structure, plan.axes, plan.terms)


def repartition(t: SymmetricTensor, outputs, inputs) -> SymmetricTensor:
    """``outputs`` become OUT, ``inputs`` IN, public order ``(*outputs, *inputs)``; legs
    that cross sides are bent (reference ``ops/repartition.py:388-476``)."""
    outputs, inputs = validated_partition(t.ndim, outputs, inputs)
    plan = repartition_plan(t.structure, outputs, inputs)
    return _apply_plan(t, plan.new_structure, plan.perm, plan.terms)


@lru_cache(maxsize=1024)
def _adjoint_plan(structure: TensorStructure):
    flipped = tuple(replace(leg, side=IN if leg.side is OUT else OUT) for leg in structure.legs)
    new = TensorStructure(flipped)
    terms = tuple(
        (structure.index_of(FusionBlockKey(k.input_tree, k.output_tree)), i, 1.0)
        for i, k in enumerate(new.block_order)
    )
    return new, terms


def adjoint(t: SymmetricTensor) -> SymmetricTensor:
    """``T^dagger``: every leg flips side, block ``(ot, it)`` becomes block ``(it, ot)``
    conjugated (real blocks here), public order unchanged (reference ``ops/map.py:446-512``)."""
    new, terms = _adjoint_plan(t.structure)
    return _apply_plan(t, new, tuple(range(t.ndim)), terms)


# ----------------------------------------------------------------------------- contraction


def tensordot(a: SymmetricTensor, b: SymmetricTensor, axes) -> SymmetricTensor:
    """Contract ``axes[0]`` of ``a`` against ``axes[1]`` of ``b``; output is ``a``'s free
    legs then ``b``'s, unchanged (reference ``ops/contraction.py:309-376``)."""
    return _tensordot(a, b, axes, None)


def _tensordot(a, b, axes, out_axes):
    ax = validated_contraction_axes(a.structure, b.structure, axes)
    contraction_plan(a.structure, b.structure, ax)  # refusals before any block moves
    require_device(a, b)
    if a.dtype != b.dtype:
        raise TypeError(f"tensordot: operand dtypes differ ({a.dtype} vs {b.dtype})")
    ex = contraction_exec(a.structure, b.structure, ax, out_axes)
    return SymmetricTensor(ex.out_structure, ex.run(a.data, b.data))


def compose(a: SymmetricTensor, b: SymmetricTensor) -> SymmetricTensor:
    """``a o b``: ``b``'s codomain consumed by ``a``'s domain; output legs are ``a``'s OUT
    legs then ``b``'s IN legs (reference ``ops/map.py:105-172``)."""
    dom = a.domain
    cod = b.codomain
    if len(dom) != len(cod):
        raise ValueError(
            f"compose: a has {len(dom)} IN legs but b has {len(cod)} OUT legs"
        )
    for i, (x, y) in enumerate(zip(dom, cod)):
        if x.space != y.space or x.dual != y.dual:
            raise ValueError(
                f"compose: a's domain and b's codomain differ at position {i} "
                f"(public axis {a.structure.in_axes[i]} of a, public axis "
                f"{b.structure.out_axes[i]} of b): composition requires the same "
                "(space, dual) in the same order"
            )
    ax = (tuple(a.structure.in_axes), tuple(b.structure.out_axes))
    # tensordot of a's domain against b's codomain yields (a OUT legs, b IN legs) in the
    # order compose promises, with no leg crossing sides.
    return _tensordot(a, b, ax, None)


def _parse(equation: str, operands) -> tuple[list, str]:
    """Refusals and implicit-output rule of reference ``ops/contraction.py:592-677``."""
    if not operands:
        raise ValueError(f"einsum: equation {equation!r} was given no operands")
    if "." in equation:
        raise ValueError(f"einsum: equation {equation!r} contains '...'; label every axis")
    lhs, arrow, rhs = equation.replace(" ", "").partition("->")
    terms = lhs.split(",")
    for label in lhs.replace(",", "") + rhs:
        if not label.isalpha() or not label.isascii():
            raise ValueError(f"einsum: label {label!r} is not an ASCII letter")
    if len(terms) != len(operands):
        raise ValueError(
            f"einsum: equation {equation!r} has {len(terms)} terms but {len(operands)} operands"
        )
    for k, (term, t) in enumerate(zip(terms, operands)):
        if len(term) != t.ndim:
            raise ValueError(
                f"einsum: term {term!r} labels {len(term)} axes but operand {k} is "
                f"{t.ndim}-dimensional"
            )
    counts = Counter(lhs.replace(",", ""))
    over = [label for label, n in counts.items() if n > 2]
    if over:
        raise ValueError(f"einsum: label {over[0]!r} occurs more than twice in {equation!r}")
    for k, term in enumerate(terms):
        rep = [label for label, n in Counter(term).items() if n > 1]
        if rep:
            raise ValueError(
                f"einsum: label {rep[0]!r} is repeated inside term {term!r}; a trace or "
                "diagonal within one operand is refused (use tenet.trace)"
            )
    out = rhs if arrow else "".join(sorted(label for label, n in counts.items() if n == 1))
    seen = set()
    for label in out:
        if label in seen:
            raise ValueError(f"einsum: label {label!r} is repeated in the output {out!r}")
        seen.add(label)
        if label not in counts:
            raise ValueError(f"einsum: output label {label!r} appears in no input term")
    for label, n in counts.items():
        if n == 1 and label not in seen:
            raise ValueError(
                f"einsum: input label {label!r} is missing from the output, which would sum "
                "that axis away"
            )
    return terms, out


def einsum(equation: str, *operands: SymmetricTensor, optimize="auto") -> SymmetricTensor:
    """``tenet.einsum("abc,cde->abde", A, B)``.  Two operands lower to one tensordot whose
    final transpose is fused into the output recouple; shared labels are contracted in
    order of first appearance in the first operand (reference
    ``ops/contraction.py:695-794``).  Three or more operands are contracted pairwise left
    to right (the reference asks opt_einsum for an order; the result is the same tensor)."""
    terms, out = _parse(equation, operands)
    if len(operands) == 1:
        return transpose(operands[0], tuple(terms[0].index(label) for label in out))
    if len(operands) > 2:
        acc, acc_term = operands[0], terms[0]
        for k in range(1, len(operands)):
            rest = "".join(terms[k + 1 :]) + out
            t = terms[k]
            keep = [lab for lab in acc_term + t if (lab in out or lab in rest) and (acc_term + t).count(lab) == 1]
            step_out = "".join(keep)
            acc = einsum(f"{acc_term},{t}->{step_out}", acc, operands[k])
            acc_term = step_out
        return transpose(acc, tuple(acc_term.index(label) for label in out))
    t0, t1 = terms
    shared = [label for label in t0 if label in t1]
    axes = (tuple(t0.index(lab) for lab in shared), tuple(t1.index(lab) for lab in shared))
    free = [lab for lab in t0 if lab not in shared] + [lab for lab in t1 if lab not in shared]
    return _tensordot(operands[0], operands[1], axes, tuple(free.index(lab) for lab in out))


# ----------------------------------------------------------------------------- vector space


def _check_same(a: SymmetricTensor, b: SymmetricTensor, op: str) -> None:
    if a.structure != b.structure:
        if a.ndim != b.ndim:
            raise ValueError(f"{op}: tensors have {a.ndim} and {b.ndim} legs")
        for i, (x, y) in enumerate(zip(a.legs, b.legs)):
            if x != y:
                raise ValueError(f"{op}: axis {i} differs: {x!r} vs {y!r}")
        raise ValueError(f"{op}: structures differ")
    if a.dtype != b.dtype:
        raise TypeError(f"{op}: dtypes differ ({a.dtype} vs {b.dtype})")


@lru_cache(maxsize=256)
def _qdim_weights_host(structure: TensorStructure):
    """Per-element quantum-dimension weights, or None when every weight is 1."""
    provider = structure.provider
    q = [provider.qdim(k.coupled) for k in structure.block_order]
    if all(x == 1.0 for x in q):
        return None
    return np.repeat(np.asarray(q, dtype=np.float64), structure.layout.sizes)


_WEIGHTS_DEV: dict = {}
_WORK: dict = {}


def _weights(structure, device):
    w = _qdim_weights_host(structure)
    if w is None:
        return None
    key = (structure, device)
    d = _WEIGHTS_DEV.get(key)
    if d is None:
        d = _WEIGHTS_DEV[key] = upload(w, device)
    else:
        wait_upload(d[1], d[2])
    return d[0]


def _work(device) -> torch.Tensor:
    """Partials buffer of the two-pass dot, one per (device, stream): reductions issued on
    different streams must not share it."""
    key = (device.type, device.index, _native.stream_ptr())
    w = _WORK.get(key)
    if w is None:
        w = _WORK[key] = torch.empty(1024, dtype=torch.float64, device=device)
    return w


ORTH_MAX = 8  # basis vectors one tn_krylov_orth call takes
_ORTH_WORK: dict = {}


def _orth_work(device) -> torch.Tensor:
    """Partials buffer of ``tn_krylov_orth``, one per (device, stream)."""
    key = (device.type, device.index, _native.stream_ptr())
    w = _ORTH_WORK.get(key)
    if w is None:
        n = int(_native.lib().tn_krylov_orth_work())
        w = _ORTH_WORK[key] = torch.empty(n, dtype=torch.float64, device=device)
    return w


def orthogonalize_(w: SymmetricTensor, basis, *, passes: int = 2, coef: torch.Tensor | None = None,
                   last_coef: torch.Tensor | None = None, norm: torch.Tensor | None = None):
    """Gram-Schmidt of ``w`` against the (orthonormal) ``basis``, in place: ``passes``
    classical passes ``w <- w - sum_i <v_i, w> v_i`` over every basis vector at once
    (qdim-weighted inner products), with FP64 accumulation (``tn_krylov_orth``).  Returns
    ``(coef, norm)``: the first pass's coefficients and ``|w|`` after the update, as device
    tensors; ``last_coef`` (a 0-d device tensor) also receives the last basis vector's
    coefficient.  At most ``ORTH_MAX`` basis vectors."""
    basis = list(basis)
    if not 1 <= len(basis) <= ORTH_MAX:
        raise ValueError(f"orthogonalize_: 1..{ORTH_MAX} basis vectors, got {len(basis)}")
    for v in basis:
        _check_same(v, w, "orthogonalize_")
    require_device(w, *basis)
    dev = w.device
    if any(v.data.data_ptr() == w.data.data_ptr() for v in basis):
        raise ValueError("orthogonalize_: w shares its storage with a basis vector")
    coef = torch.empty(ORTH_MAX, dtype=torch.float64, device=dev) if coef is None else coef
    norm = torch.empty((), dtype=torch.float64, device=dev) if norm is None else norm
    ptrs = (ctypes.c_void_p * len(basis))(*[v.data.data_ptr() for v in basis])
    wt = _weights(w.structure, dev)
    _native.check(_native.lib().tn_krylov_orth(
        _dt(w), w.data.numel(), len(basis), ptrs, w.data.data_ptr(), None if wt is None else wt.data_ptr(),
        passes, coef.data_ptr(), None if last_coef is None else last_coef.data_ptr(), norm.data_ptr(),
        _orth_work(dev).data_ptr(), _native.stream_ptr()))
    return coef, norm


def inner_device(a: SymmetricTensor, b: SymmetricTensor, out: torch.Tensor | None = None):
    """``<a|b>`` as a 0-d float64 device tensor (no host synchronization)."""
    _check_same(a, b, "inner")
    require_device(a, b)
    res = out if out is not None else torch.empty((), dtype=torch.float64, device=a.device)
    w = _weights(a.structure, a.device)
    _native.check(
        _native.lib().tn_dot(
            _dt(a), a.data.data_ptr(), b.data.data_ptr(), 0 if w is None else w.data_ptr(),
            a.data.numel(), _work(a.device).data_ptr(), res.data_ptr(), _native.stream_ptr(),
        )
    )
    return res


def inner(a: SymmetricTensor, b: SymmetricTensor):
    """``<a|b> = sum_tau qdim(c_tau) <A_tau, B_tau>`` (reference ``contraction.py:517-589``),
    returned as a 0-d float64 device tensor."""
    return inner_device(a, b)


def norm(t: SymmetricTensor):
    """``sqrt(sum_tau qdim(c_tau) |A_tau|^2)`` (reference ``ops/basic.py:309-367``) as a 0-d
    float64 device tensor."""
    return torch.sqrt(inner_device(t, t))


def _scalar(s):
    if isinstance(s, torch.Tensor):
        return s
    return float(s)


def _axpby(alpha, x: SymmetricTensor, beta, y_data: torch.Tensor) -> None:
    a_h, a_d, a_inv = _split_scalar(alpha)
    b_h, b_d, b_inv = _split_scalar(beta)
    _native.check(
        _native.lib().tn_axpby(
            _dt(x), x.data.numel(), a_h, a_d, a_inv, x.data.data_ptr(), b_h, b_d, b_inv,
            y_data.data_ptr(), _native.stream_ptr(),
        )
    )


def _split_scalar(s):
    """(host factor, device pointer or None, invert flag) for host floats, 0-d device
    tensors, and ``("inv", tensor)`` reciprocals."""
    if isinstance(s, tuple) and s[0] == "inv":
        t = s[1].to(torch.float64)
        return 1.0, t.data_ptr(), 1
    if isinstance(s, torch.Tensor):
        t = s.to(torch.float64)
        if not t.is_cuda:
            return float(t), None, 0
        return 1.0, t.data_ptr(), 0
    return float(s), None, 0


def multiply(t: SymmetricTensor, s) -> SymmetricTensor:
    """``s * t`` for a host scalar or a 0-d device tensor."""
    require_device(t)
    out = torch.empty_like(t.data)
    _axpby(_scalar(s), t, 0.0, out)
    return SymmetricTensor(t.structure, out)


def divide(t: SymmetricTensor, s) -> SymmetricTensor:
    require_device(t)
    out = torch.empty_like(t.data)
    if isinstance(s, torch.Tensor) and s.is_cuda:
        _axpby(("inv", s), t, 0.0, out)
    else:
        _axpby(1.0 / float(s), t, 0.0, out)
    return SymmetricTensor(t.structure, out)


def add(a: SymmetricTensor, b: SymmetricTensor) -> SymmetricTensor:
    _check_same(a, b, "add")
    require_device(a, b)
    out = b.data.clone()
    _axpby(1.0, a, 1.0, out)
    return SymmetricTensor(a.structure, out)


def subtract(a: SymmetricTensor, b: SymmetricTensor) -> SymmetricTensor:
    _check_same(a, b, "subtract")
    require_device(a, b)
    out = a.data.clone()
    _axpby(-1.0, b, 1.0, out)
    return SymmetricTensor(a.structure, out)


def conj(t: SymmetricTensor) -> SymmetricTensor:
    """Real blocks: the conjugate is the tensor itself (shares the buffer)."""
    return t


def cast_data(x: torch.Tensor, dtype: torch.dtype) -> torch.Tensor:
    """A flat FP64/FP32 device buffer converted by ``tn_cast`` (``x`` itself when it already
    has ``dtype``)."""
    if x.dtype == dtype:
        return x
    codes = {torch.float64: _native.TN_F64, torch.float32: _native.TN_F32}
    if x.dtype not in codes or dtype not in codes:
        raise TypeError(f"cast_data: {x.dtype} -> {dtype} is not an FP64/FP32 conversion")
    if x.device.type != "cuda" or not x.is_contiguous():
        raise RuntimeError("cast_data: needs a contiguous buffer on the HIP device")
    out = torch.empty(x.numel(), dtype=dtype, device=x.device)
    _native.check(
        _native.lib().tn_cast(codes[x.dtype], x.data_ptr(), codes[dtype], out.data_ptr(), x.numel(),
                              _native.stream_ptr())
    )
    return out


def astype(t: SymmetricTensor, dtype: torch.dtype) -> SymmetricTensor:
    if t.dtype == dtype:
        return t
    require_device(t)
    return SymmetricTensor(t.structure, cast_data(t.data.reshape(-1), dtype))


def allclose(a: SymmetricTensor, b: SymmetricTensor, rtol: float = 1e-5, atol: float = 1e-8) -> bool:
    """Norm-wise closeness: ``|a - b| <= atol + rtol |b|``."""
    _check_same(a, b, "allclose")
    d = float(norm(subtract(a, b)))
    return d <= atol + rtol * float(norm(b))


__all__ += ["inner_device", "RecoupleProgram"]


def block_sqrt(t: SymmetricTensor) -> SymmetricTensor:
    """Elementwise square root of every block (reference ``ops/blocks.py``): on a diagonal
    singular-value or density spectrum tensor, the spectrum's square root."""
    require_device(t)
    return SymmetricTensor(t.structure, torch.sqrt(t.data))

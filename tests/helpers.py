"""Shared test helpers: building U(1) legs for both the product and the oracle, and
converting product tensors into oracle tensors."""

import math

import numpy as np

from oracle import u1 as O


def gaussian_degs(qs, scale, width):
    return {q: max(1, int(round(scale * math.exp(-(q * q) / (2.0 * width * width))))) for q in qs}


def product_space(degs):
    from tenet import GradedSpace
    from tenet.symmetry import U1, U1Sector

    return GradedSpace.new(U1, {U1Sector(q): m for q, m in degs.items()})


def leg(degs, out, dual=False):
    from tenet import IN, OUT, Leg

    return Leg(product_space(degs), OUT if out else IN, dual)


def oleg(degs, out, dual=False):
    return O.OLeg(O.space(degs), out, dual)


def to_oracle(t):
    """An oracle tensor with the product tensor's legs and block values."""
    from tenet import OUT

    legs = tuple(
        O.OLeg(tuple((a.charge, m) for a, m in l.space.sectors), l.side is OUT, l.dual) for l in t.legs
    )
    return O.OTensor.from_flat(legs, t.data.detach().to("cpu").double().numpy())


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))

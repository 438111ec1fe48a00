"""``linalg.lq`` on the ``tn_lq`` Householder kernel against LAPACK on the host: per sector
matrix, ``L`` and ``Q`` equal NumPy's QR of the transpose (``np.linalg.qr(B.T)``, the
reference's ``linalg.py:549-620`` and the oracle's ``lq``; same sign convention), ``Q`` has
orthonormal rows, ``L`` is lower trapezoidal, and ``L Q`` rebuilds the tensor -- for square,
tall, wide and rank-deficient sectors, FP64 and FP32."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _tensor(dl, dr, seed, zero_block=False, dtype=None):
    from tenet.network.models import PHYS, gaussian_bond
    from tenet.space import IN, OUT, Leg
    from tenet.tensor import SymmetricTensor

    legs = (Leg(gaussian_bond(9, 32, dl, 2.0), OUT), Leg(PHYS, OUT), Leg(gaussian_bond(10, 32, dr, 2.0), IN))
    t = SymmetricTensor.random(legs, seed=seed, device="cuda")
    if zero_block:
        lay = t.structure.layout
        o, shp = int(lay.offsets[0]), lay.shapes[0]
        t.data[o: o + int(np.prod(shp))] = 0.0
    if dtype is not None:
        t = SymmetricTensor(t.structure, t.data.to(dtype))
    return t


@pytest.mark.parametrize("dl,dr,zero_block", [(40, 40, False), (90, 30, False), (20, 120, False), (60, 60, True)])
def test_lq_matches_lapack_per_sector(dl, dr, zero_block):
    from tenet import linalg, ops

    t = _tensor(dl, dr, seed=dl + dr, zero_block=zero_block)
    L, Q = linalg.lq(t, ((0,), (1, 2)))
    mats = {c: b.cpu().numpy() for c, b in linalg.to_matrices(ops.repartition(t, (0,), (1, 2))).items()}
    lm = {c: b.cpu().numpy() for c, b in linalg.to_matrices(L).items()}
    qm = {c: b.cpu().numpy() for c, b in linalg.to_matrices(Q).items()}
    for c, b in mats.items():
        q, r = np.linalg.qr(b.T)
        scale = max(np.abs(b).max(), 1e-300)
        assert lm[c].shape == r.T.shape and qm[c].shape == q.T.shape
        assert np.abs(lm[c] - r.T).max() <= 1e-12 * scale, c
        assert np.allclose(np.triu(lm[c], 1), 0.0)
        assert np.abs(lm[c] @ qm[c] - b).max() <= 1e-12 * scale
        k = qm[c].shape[0]
        assert np.abs(qm[c] @ qm[c].T - np.eye(k)).max() <= 1e-13
        # the rows of Q follow LAPACK's wherever L's column is not (numerically) null
        live = np.abs(np.diag(r)) > 1e-10 * scale
        if live.any():
            assert np.abs(qm[c][live] - q.T[live]).max() <= 1e-10, c


def test_lq_float32_tensors_factorise_in_fp64():
    from tenet import linalg, ops

    t = _tensor(50, 50, seed=3, dtype=torch.float32)
    L, Q = linalg.lq(t, ((0,), (1, 2)))
    assert L.dtype == Q.dtype == torch.float32
    mats = linalg.to_matrices(ops.repartition(t, (0,), (1, 2)))
    lm, qm = linalg.to_matrices(L), linalg.to_matrices(Q)
    for c, b in mats.items():
        b = b.double().cpu().numpy()
        err = np.abs(lm[c].double().cpu().numpy() @ qm[c].double().cpu().numpy() - b).max()
        assert err <= 1e-6 * np.abs(b).max(), c


def test_canonize_gives_right_isometries():
    """``MPS.canonize_`` on the kernel: every site but the first is a right isometry."""
    from tenet import linalg, ops
    from tenet.network.models import random_chain

    psi = random_chain(12, 24, 2.0, seed=4, device="cuda")
    psi.canonize_(0)
    for n in range(1, len(psi)):
        for c, b in linalg.to_matrices(ops.repartition(psi[n], (0,), (1, 2))).items():
            b = b.cpu().numpy()
            assert np.abs(b @ b.T - np.eye(b.shape[0])).max() <= 1e-12, (n, c)

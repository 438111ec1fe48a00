"""``tn_krylov_orth`` (``ops.orthogonalize_``): classical Gram-Schmidt passes of a Krylov
vector against the Lanczos basis, checked against NumPy FP64 on the host -- coefficients,
the orthogonalised vector, its norm -- with and without qdim weights, in FP64 and FP32,
and the Lanczos step built on it against the oracle's Lanczos."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _u1_structure():
    """A U(1) three-leg structure of ~90k elements (unit qdims: plain inner products)."""
    from tenet.network.models import PHYS, gaussian_bond
    from tenet.space import IN, OUT, Leg
    from tenet.structure import TensorStructure

    return TensorStructure((Leg(gaussian_bond(15, 32, 512, 3.0), OUT), Leg(PHYS, OUT),
                            Leg(gaussian_bond(16, 32, 512, 3.0), IN)))


@pytest.mark.parametrize("k,passes,dtype", [(1, 1, torch.float64), (3, 2, torch.float64), (8, 2, torch.float64),
                                           (5, 2, torch.float32)])
def test_gram_schmidt_passes_match_numpy(k, passes, dtype):
    from tenet import ops
    from tenet.tensor import SymmetricTensor

    st = _u1_structure()
    rng = np.random.default_rng(k)
    n = st.size
    q, _ = np.linalg.qr(rng.standard_normal((n, k)))
    w0 = rng.standard_normal(n) + 3.0 * q @ rng.standard_normal(k)
    basis = [SymmetricTensor(st, torch.from_numpy(q[:, i].copy()).to(dtype).cuda()) for i in range(k)]
    w = SymmetricTensor(st, torch.from_numpy(w0).to(dtype).cuda())
    last = torch.zeros((), dtype=torch.float64, device="cuda")
    coef, norm = ops.orthogonalize_(w, basis, passes=passes, last_coef=last)
    qd = np.stack([b.data.double().cpu().numpy() for b in basis], 1)
    wd = w0.astype(np.float32).astype(np.float64) if dtype == torch.float32 else w0
    c_ref = qd.T @ wd
    ref = wd - qd @ c_ref
    if passes == 2:
        ref = ref - qd @ (qd.T @ ref)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    got = w.data.double().cpu().numpy()
    assert np.abs(coef.cpu().numpy()[:k] - c_ref).max() <= tol * np.abs(c_ref).max()
    assert float(last) == float(coef[k - 1])
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(wd)
    assert abs(float(norm) - np.linalg.norm(got)) <= tol * np.linalg.norm(got)
    if dtype == torch.float64:
        assert np.abs(qd.T @ got).max() <= 1e-13 * np.linalg.norm(got)


def test_weighted_inner_products_follow_the_qdim_weights():
    """SU(2) vectors: inner products carry the qdim weight of each block's coupled sector."""
    from tenet import ops
    from tenet.network.models import SU2_PHYS, su2_chain_bonds
    from tenet.space import IN, OUT, Leg
    from tenet.tensor import SymmetricTensor

    bonds = su2_chain_bonds(8, 12, 2.0, max_two_j=4)
    legs = (Leg(bonds[3], OUT), Leg(SU2_PHYS, OUT), Leg(bonds[4], IN))
    vs = [SymmetricTensor.random(legs, seed=s, device="cuda") for s in range(4)]
    # orthonormalise the basis with the library's own weighted inner product
    basis = []
    for v in vs[:3]:
        for b in basis:
            v = ops.subtract(v, ops.multiply(b, float(ops.inner(b, v))))
        basis.append(ops.divide(v, float(ops.norm(v))))
    w = vs[3]
    w0 = w.data.clone()
    ops.orthogonalize_(w, basis, passes=2)
    for b in basis:
        assert abs(float(ops.inner(b, w))) <= 1e-13 * float(ops.norm(w))
    # the change lies in the basis span: w0 - w = sum_i c_i b_i with c_i = <b_i, w0>
    w_orig = SymmetricTensor(w.structure, w0)
    rebuilt = w
    for b in basis:
        rebuilt = ops.add(rebuilt, ops.multiply(b, float(ops.inner(b, w_orig))))
    assert float(ops.norm(ops.subtract(rebuilt, w_orig))) <= 1e-12 * float(ops.norm(w_orig))


def test_refuses_a_vector_aliasing_the_basis():
    from tenet import ops
    from tenet.tensor import SymmetricTensor

    st = _u1_structure()
    v = SymmetricTensor(st, torch.ones(st.size, dtype=torch.float64, device="cuda"))
    with pytest.raises(ValueError):
        ops.orthogonalize_(v, [v])
    with pytest.raises(ValueError):
        ops.orthogonalize_(v, [])

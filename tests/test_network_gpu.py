"""GPU parity of the DMRG path: H_eff (engine at every precision and the per-block arm),
Lanczos, truncated SVD, canonical form and full DMRG against the oracle and the exact
energies the reference's tests hold."""

import numpy as np
import pytest
import torch

from oracle import ed
from oracle import u1 as O
from oracle import u1_network as N
from tests.helpers import rel_err, to_oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_problem():
    from tenet.network.workloads import heff2_problem

    return heff2_problem(12, 24, 2.0, seed=3, device="cuda")


def test_heff2_fp64_matches_oracle(small_problem):
    from tenet.network import Heff2

    fl, w1, w2, fr, aa = small_problem
    got = Heff2(fl, w1, w2, fr, "fp64")(aa)
    ref = N.heff2(*(to_oracle(t) for t in (fl, w1, w2, fr, aa)))
    g = to_oracle(got)
    assert g.keys() == ref.keys()
    assert rel_err(g.flat(), ref.flat()) < 1e-12


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16x3", 5e-5)])
def test_heff2_reduced_precision(small_problem, prec, tol):
    from tenet.network import Heff2

    fl, w1, w2, fr, aa = small_problem
    ref = Heff2(fl, w1, w2, fr, "fp64")(aa)
    got = Heff2(fl, w1, w2, fr, prec)(aa)
    assert got.dtype == torch.float64
    assert rel_err(got.data.cpu().numpy(), ref.data.cpu().numpy()) < tol


def test_per_block_arm_matches_oracle(small_problem):
    from tenet.baseline_blocks import BlockTensor, heff2

    fl, w1, w2, fr, aa = small_problem
    out = heff2(*(BlockTensor.from_symmetric(t) for t in (fl, w1, w2, fr, aa)))
    got = torch.cat([b.reshape(-1) for b in out.blocks]).cpu().numpy()
    ref = N.heff2(*(to_oracle(t) for t in (fl, w1, w2, fr, aa)))
    assert rel_err(got, ref.flat()) < 1e-12


def test_env_heff2_and_lanczos_match_oracle(small_problem):
    from tenet.network import Heff2, lanczos

    fl, w1, w2, fr, aa = small_problem
    op = Heff2(fl, w1, w2, fr, "fp64")
    e, v = lanczos(op, aa, ncv=3)
    ofl, ow1, ow2, ofr, oaa = (to_oracle(t) for t in (fl, w1, w2, fr, aa))
    eo, vo = N.lanczos(lambda x: N.heff2(ofl, ow1, ow2, ofr, x), oaa, 3)
    assert abs(e - eo) < 1e-10 * abs(eo)
    assert rel_err(to_oracle(v).flat(), vo.flat()) < 1e-8


def test_svd_truncated_matches_oracle():
    import tenet
    from tests.helpers import leg, oleg

    d = {-1: 5, 0: 7, 1: 4, 2: 2}
    legs = (leg(d, 1), leg(d, 1), leg(d, 1), leg(d, 0))
    t = tenet.SymmetricTensor.random(legs, seed=11, device="cuda")
    u, s, vh = tenet.linalg.svd_truncated(t, ((0, 1), (2, 3)), max_bond=20, cutoff=1e-12)
    ou, os_, ovh, spec, dw = N.svd_truncated(to_oracle(t), (0, 1), (2, 3), 20, 1e-12)
    assert [(a.charge, m) for a, m in s.legs[0].space.sectors] == list(os_.legs[0].sectors)
    assert rel_err(sorted(np.concatenate([np.diag(b) for b in s.numpy_blocks()])),
                   sorted(np.concatenate([np.diag(b) for b in os_.blocks.values()]))) < 1e-12
    rec = tenet.compose(tenet.compose(u, s), vh)
    m = tenet.repartition(t, (0, 1), (2, 3))
    err2 = float(tenet.norm(tenet.subtract(rec, m))) ** 2
    assert abs(err2 - dw) < 1e-9 * float(tenet.norm(m)) ** 2


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("fp32", 1e-8)])
def test_dmrg_heisenberg_12_matches_ed(prec, tol):
    from tenet.network import dmrg_
    from tenet.network.models import heisenberg_mpo, random_chain

    psi = random_chain(12, 16, 3.0, seed=1, device="cuda")
    out = dmrg_(psi, heisenberg_mpo(12, device="cuda"), chi=32, max_sweeps=10, precision=prec)
    assert abs(out.energy - ed.E_OBC_12) < tol * abs(ed.E_OBC_12)


def test_dmrg_xx_chain_matches_closed_form():
    from tenet.network import dmrg_
    from tenet.network.models import heisenberg_mpo, random_chain

    L = 14
    psi = random_chain(L, 16, 3.0, seed=2, device="cuda")
    out = dmrg_(psi, heisenberg_mpo(L, delta=0.0, device="cuda"), chi=48, max_sweeps=12)
    assert abs(out.energy - ed.xx_energy(L)) < 1e-10 * abs(ed.xx_energy(L))


@pytest.mark.parametrize("layout", ["sector", "blocks"])
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-13), ("fp32", 1e-6)])
def test_batched_heff2_matches_individual_operators(prec, tol, layout):
    from tenet.network import Heff2
    from tenet.network.batched import BatchedHeff2, synthetic_operator

    # the 24-site D=600 instance makes the merged step-3 GEMM split its K into slices
    insts = [synthetic_operator(n, d, sg, dl, s, device="cuda")
             for n, d, sg, dl, s in ((16, 40, 3.0, 0.5, 1), (24, 600, 4.0, 1.3, 2), (16, 24, 3.0, 1.9, 3))]
    b = BatchedHeff2(insts, prec, layout=layout)
    if layout == "sector":
        assert b.gemm4.ws_size > b.x_size
    out = b.unpack(b.apply(b.pack([i[4] for i in insts])))
    for inst, got in zip(insts, out):
        ref = Heff2(*inst[:4], "fp64")(inst[4])
        assert got.structure == ref.structure
        assert rel_err(got.data.double().cpu().numpy(), ref.data.cpu().numpy()) < tol


def test_synthetic_operator_is_symmetric():
    from tenet.network import Heff2
    from tenet.network.batched import synthetic_operator
    import tenet

    fl, w1, w2, fr, aa = synthetic_operator(16, 40, 3.0, 0.7, 5, device="cuda")
    op = Heff2(fl, w1, w2, fr, "fp64")
    x = tenet.SymmetricTensor.random(aa.legs, seed=8, device="cuda")
    y = tenet.SymmetricTensor.random(aa.legs, seed=9, device="cuda")
    a, b = float(tenet.inner(x, op(y))), float(tenet.inner(op(x), y))
    assert abs(a - b) < 1e-12 * max(abs(a), 1e-300)

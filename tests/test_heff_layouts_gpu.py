"""GPU parity of the two abelian H_eff layouts: sector-stacked GEMMs (the default) and one
GEMM problem per block pair, against the FP64 oracle (reference ``network/env.py:416-432``,
``_heff2_full``) on the centre bond of a small chain and on an asymmetric bond pair (left
and right bonds with different sector sets, so some stacked rows and columns have no
partner in the other operand)."""

import numpy as np
import pytest
import torch

from oracle import u1_network as N
from tests.helpers import rel_err, to_oracle

pytestmark = pytest.mark.gpu


def _asymmetric_problem(seed: int):
    """H_eff on sites (2, 3) of a 12-site chain: left bond 2 (charges -2, 0, 2), right bond
    4 (charges -4..4), random Hermitian environments."""
    from tenet import ops
    from tenet.network.batched import _symmetric_env
    from tenet.network.models import PHYS, gaussian_bond, heisenberg_mpo
    from tenet.space import IN, OUT, Leg
    from tenet.tensor import SymmetricTensor

    rng = np.random.default_rng(seed)
    bl = gaussian_bond(2, 12, 64, 3.0)
    br = gaussian_bond(4, 12, 64, 3.0)
    h = heisenberg_mpo(12, 0.7, device="cuda")
    fl = _symmetric_env(bl, True, rng, "cuda")
    fr = _symmetric_env(br, False, rng, "cuda")
    aa = SymmetricTensor.random((Leg(bl, OUT), Leg(PHYS, OUT), Leg(PHYS, OUT), Leg(br, IN)),
                                seed=seed, device="cuda")
    aa = ops.divide(aa, ops.norm(aa))
    return fl, h[2], h[3], fr, aa


@pytest.fixture(scope="module", params=["centre", "asymmetric"])
def problem(request):
    if request.param == "centre":
        from tenet.network.workloads import heff2_problem

        return heff2_problem(12, 40, 1.5, seed=5, device="cuda")
    return _asymmetric_problem(7)


def test_sector_and_block_layouts_match_oracle(problem):
    from tenet.network import Heff2

    fl, w1, w2, fr, aa = problem
    ref = N.heff2(*(to_oracle(t) for t in (fl, w1, w2, fr, aa)))
    for layout in ("sector", "blocks"):
        op = Heff2(fl, w1, w2, fr, "fp64", layout=layout)
        got = to_oracle(op(aa))
        assert got.keys() == ref.keys(), layout
        assert rel_err(got.flat(), ref.flat()) < 1e-12, layout
        # a second call reuses the compiled chain and the stacked operands
        assert rel_err(to_oracle(op(aa)).flat(), ref.flat()) < 1e-12, layout


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16x3", 5e-5)])
def test_sector_layout_reduced_precision(problem, prec, tol):
    from tenet.network import Heff2

    fl, w1, w2, fr, aa = problem
    ref = Heff2(fl, w1, w2, fr, "fp64", layout="blocks")(aa).data
    got = Heff2(fl, w1, w2, fr, prec, layout="sector")(aa)
    assert got.dtype == torch.float64
    assert rel_err(got.data.cpu().numpy(), ref.cpu().numpy()) < tol
    # an FP32 vector comes back FP32
    op = Heff2(fl, w1, w2, fr, prec, layout="sector")
    y32 = op.apply_data(aa.structure, aa.data.to(torch.float32))
    assert y32.dtype == torch.float32
    assert rel_err(y32.double().cpu().numpy(), ref.cpu().numpy()) < tol


def test_unknown_layout_is_refused(problem):
    from tenet.network import Heff2

    fl, w1, w2, fr, _ = problem
    with pytest.raises(ValueError):
        Heff2(fl, w1, w2, fr, "fp64", layout="dense")


def test_sector_layout_with_split_k_matches_block_layout():
    """A bond large enough that step 3 is split into K slices: the exit pass sums them."""
    from tenet.network import Heff2
    from tenet.network.workloads import heff2_problem

    fl, w1, w2, fr, aa = heff2_problem(20, 512, 4.0, seed=2, device="cuda")
    sec = Heff2(fl, w1, w2, fr, "fp64", layout="sector")
    ch = sec.chain(aa.structure)
    assert ch.gemm4.ws_size > aa.structure.size  # at least one problem split
    ref = Heff2(fl, w1, w2, fr, "fp64", layout="blocks")(aa).data
    got = sec(aa).data
    assert rel_err(got.cpu().numpy(), ref.cpu().numpy()) < 1e-12
    got32 = Heff2(fl, w1, w2, fr, "fp32", layout="sector")(aa).data
    assert rel_err(got32.cpu().numpy(), ref.cpu().numpy()) < 2e-6


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_row_gathered_step_one_equals_the_entry_gather(prec):
    """Step 1 of the FP32 / BF16x3 sector chain reads aa's FP64 blocks in place through its
    row table and rounds them while staging (tn_gemm_grouped_rows); the same GEMM on the
    entry pass's gathered, cast copy gives bit-identical t1 (same roundings, same
    accumulation order), on a bond whose sectors span several 64-row tiles."""
    from tenet.network import Heff2
    from tenet.network.workloads import heff2_problem

    fl, w1, w2, fr, aa = heff2_problem(20, 512, 4.0, seed=3, device="cuda")
    op = Heff2(fl, w1, w2, fr, prec, layout="sector")
    ch = op.chain(aa.structure)
    frs, _ = ch.prepare(op._datas)
    t1_rows = torch.full((max(ch.t1_size, 1),), float("nan"), dtype=torch.float32, device="cuda")
    ch.gemm1_rows.run(aa.data, frs, t1_rows, prec)
    xs = torch.empty(max(ch.xs_size, 1), dtype=torch.float32, device="cuda")
    ch.entry.run(aa.data, xs)
    t1 = torch.full_like(t1_rows, float("nan"))
    ch.gemm1.run(xs, frs, t1, prec)
    torch.cuda.synchronize()
    assert not torch.isnan(t1_rows).any()
    assert torch.equal(t1_rows, t1)
    # an FP32 vector through the same row table
    t1_32 = torch.full_like(t1_rows, float("nan"))
    ch.gemm1_rows.run(aa.data.to(torch.float32), frs, t1_32, prec)
    assert torch.equal(t1_32, t1)

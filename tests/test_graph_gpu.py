"""hipGraph capture of the H_eff matvec and of the Lanczos iterations.

Each graph is replayed several times with the program's own eager work and
synchronisations between replays, and every replay is checked against the eager result
on the same input (a replay must not depend on what the previous replay left behind)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _problem(precision, L=12, D=40):
    from tenet.network import Heff2
    from tenet.network.workloads import heff2_problem

    fl, w1, w2, fr, aa = heff2_problem(L, D, 2.0, seed=4, device="cuda")
    return Heff2(fl, w1, w2, fr, precision), aa


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graphed_matvec_replays_match_eager(precision):
    from tenet.network.graph import GraphedMatvec

    op, aa = _problem(precision)
    gm = GraphedMatvec(op, aa.structure)
    for k in range(4):
        x = torch.randn(aa.structure.size, dtype=torch.float64, device="cuda")
        want = op.apply_data(aa.structure, x).clone()
        got = gm(x).clone()
        torch.cuda.synchronize()
        scratch = torch.randn(1 << 16, device="cuda").sum()  # eager work between replays
        assert float(scratch) == float(scratch)
        assert torch.equal(got, want), f"replay {k} differs from the eager matvec"


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_krylov_graph_matches_eager_lanczos(precision):
    from tenet import ops
    from tenet.network import lanczos
    from tenet.network.graph import KrylovGraph

    op, aa = _problem(precision)
    kg = KrylovGraph(op, aa.structure, ncv=4)
    v = aa
    for k in range(3):
        e_ref, v_ref = lanczos(op, v, ncv=4)
        e_g, v_g = kg.solve(v)
        torch.cuda.synchronize()
        assert abs(e_g - e_ref) <= 1e-12 * abs(e_ref), f"replay {k}: {e_g} vs {e_ref}"
        d = float(ops.norm(ops.subtract(v_g, v_ref)))
        assert d < 1e-10, f"replay {k}: Ritz vectors differ by {d:.3e}"
        v = ops.add(v_ref, ops.multiply(aa, 0.1))  # a new start vector for the next replay


def test_krylov_graph_restarted_to_convergence():
    """Restarted graph Lanczos converges to the same ground energy as the eager loop."""
    from tenet.network import lanczos
    from tenet.network.graph import KrylovGraph

    op, aa = _problem("fp64", L=10, D=24)
    kg = KrylovGraph(op, aa.structure, ncv=6)
    v, e = aa, None
    for _ in range(60):
        e, v = kg.solve(v)
    ve, ee = aa, None
    for _ in range(60):
        ee, ve = lanczos(op, ve, ncv=6)
    assert abs(e - ee) < 1e-11 * abs(ee)

"""Host logic of the generic two-site H_eff chain (``tenet.network.heff._Chain``, the path
of non-abelian operators): the composed boundary recouples (output scatter, ``_composed``
bend and operand gather of ``network/env.py:60-90,429-432`` fused into one map), the step
GEMM tables and the final scatter -- with the last GEMM's K slices folded into it when K is
split -- are executed by the NumPy table interpreter and compared with the dense SU(2)
contraction of the oracle (``oracle/su2.py``, pinned to sympy's Clebsch-Gordan values)."""

import numpy as np
import pytest
import torch

from oracle import su2 as O
from tests.interp import run_gemm, run_recouple


def _dense(t):
    from tenet import OUT

    ol = [O.SLeg([(a.two_j, m) for a, m in l.space.sectors], l.side is OUT, l.dual) for l in t.legs]
    return O.to_dense(O.SU2Rules, ol, [np.asarray(b, dtype=np.float64) for b in t.numpy_blocks()])


def _problem(n_sites, mult, seed):
    from tenet.network.models import SU2_MPO_BOND, SU2_PHYS, su2_chain_bonds, su2_heisenberg_mpo
    from tenet.space import IN, OUT, Leg
    from tenet.tensor import SymmetricTensor

    bonds = su2_chain_bonds(n_sites, mult, 3.0)
    c = n_sites // 2 - 1
    bl, br = bonds[c], bonds[c + 2]
    h = su2_heisenberg_mpo(n_sites, device="cpu")
    fl = SymmetricTensor.random((Leg(bl, IN), Leg(SU2_MPO_BOND, OUT), Leg(bl, OUT)), seed=seed, device="cpu")
    fr = SymmetricTensor.random((Leg(br, OUT), Leg(SU2_MPO_BOND, IN), Leg(br, IN)), seed=seed + 1, device="cpu")
    aa = SymmetricTensor.random((Leg(bl, OUT), Leg(SU2_PHYS, OUT), Leg(SU2_PHYS, OUT), Leg(br, IN)),
                                seed=seed + 2, device="cpu")
    return fl, h[c], h[c + 1], fr, aa


def _bent_statics(fl, w1, w2, fr):
    """The static operands bent as ``Heff2`` bends them, data moved by the interpreter."""
    from tenet.engine import inplace_map, recouple_program
    from tenet.network.heff import HEFF2_STEPS, _bent
    from tenet.plans import repartition_plan

    out = []
    for (eq, bend), t in zip(HEFF2_STEPS, (fr, w2, w1, fl)):
        tb = eq.split("->")[0].split(",")[1]
        data = t.data.double().numpy()
        s = t.structure
        if bend:
            outs, ins, tb = _bent(s, tb, bend)
            plan = repartition_plan(s, outs, ins)
            D = plan.new_structure
            offs = inplace_map(s, D, plan.perm, plan.terms)
            if offs is None or not (offs == D.layout.offsets).all():
                moved = np.zeros(D.size)
                run_recouple(recouple_program(s, D, tuple(plan.perm), tuple(plan.terms)), data, moved)
                data = moved
            s = D
        out.append(((s, tb), data))
    return out


def _emulate(chain, statics, x):
    cur = x
    n = len(chain.stages)
    for k, ((prog, a_struct, ex), (_, data)) in enumerate(zip(chain.stages, statics)):
        if prog is not None:
            nxt = np.zeros(a_struct.size)
            run_recouple(prog, cur, nxt)
            cur = nxt
        bp = data
        if ex.b.program is not None:
            bp = np.zeros(ex.b.structure.size)
            run_recouple(ex.b.program, data, bp)
        g = chain.last_gemm if k == n - 1 else ex.gemm
        C = np.zeros(max(g.c_size, 1))
        W = run_gemm(g, cur, bp, C)
        cur = W if g.gather else C
    if chain.final is None:
        return cur[: chain.s_out.size]
    out = np.zeros(chain.s_out.size)
    run_recouple(chain.final, cur, out)
    return out


@pytest.fixture
def split_everything():
    """Make every long enough GEMM split its K, so the gathered last step is exercised."""
    from tenet.engine import GemmProgram, contraction_exec

    old = GemmProgram.SPLIT_MIN_K, GemmProgram.SPLIT_TARGET_TILES
    contraction_exec.cache_clear()
    GemmProgram.SPLIT_MIN_K, GemmProgram.SPLIT_TARGET_TILES = 8, 1 << 20
    yield
    GemmProgram.SPLIT_MIN_K, GemmProgram.SPLIT_TARGET_TILES = old
    contraction_exec.cache_clear()


def _check(n_sites, mult, seed, expect_split):
    from tenet.network.heff import _Chain
    from tenet.tensor import SymmetricTensor

    fl, w1, w2, fr, aa = _problem(n_sites, mult, seed)
    statics = _bent_statics(fl, w1, w2, fr)
    chain = _Chain(aa.structure, tuple(st for st, _ in statics))
    assert (chain.last_gemm.gather) == expect_split
    # at most one recouple launch between two GEMMs
    assert len(chain.stages) == 4
    got = _emulate(chain, statics, aa.data.double().numpy())
    want = np.einsum("axB,apqr,xPpm,mQqy,rys->BPQs", _dense(fl), _dense(aa), _dense(w1),
                     _dense(w2), _dense(fr), optimize=True)
    gd = _dense(SymmetricTensor(aa.structure, torch.from_numpy(got)))
    assert np.linalg.norm(gd - want) <= 1e-12 * np.linalg.norm(want)


@pytest.mark.parametrize("n_sites,mult,seed", [(10, 6, 0), (12, 10, 4)])
def test_generic_chain_tables_reproduce_the_dense_su2_contraction(n_sites, mult, seed):
    _check(n_sites, mult, seed, expect_split=False)


def test_generic_chain_with_split_last_gemm(split_everything):
    _check(16, 20, 2, expect_split=True)

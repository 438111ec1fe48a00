"""``tn_mpo_apply`` (the fused two-site MPO pass) against NumPy on synthetic slab tables:
every input-channel count the kernel compiles a body for (0 ... 32, including the counts
where it changes how many positions a thread batches), slab pairs whose position count is
not a multiple of the tile, and strided slabs."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_mpo_apply_matches_numpy(dtype):
    from tenet import _native

    rng = np.random.default_rng(5)
    jobs, ins, outs, coef, tiles = [], [], [], [], []
    src_size = dst_size = 0
    ref_jobs = []
    for j, nin in enumerate((0, 1, 3, 5, 8, 9, 12, 16, 17, 25, 32)):
        R, S = int(rng.integers(1, 40)), int(rng.integers(1, 70))
        nout = int(rng.integers(1, 6))
        stride_in, stride_out = S + int(rng.integers(0, 3)), S + int(rng.integers(0, 3))
        i_offs = []
        for _ in range(nin):
            i_offs.append(src_size)
            src_size += R * stride_in
        o_offs = []
        for _ in range(nout):
            o_offs.append(dst_size)
            dst_size += R * stride_out
        cm = np.zeros((nout, 32))
        cm[:, :nin] = rng.standard_normal((nout, nin))
        jobs.append((R, S, len(ins), len(ins) + nin, len(outs), len(outs) + nout, len(coef), 0))
        ins += [(o, stride_in) for o in i_offs]
        outs += [(o, stride_out) for o in o_offs]
        coef += cm.reshape(-1).tolist()
        nt = -(-(R * S) // _native.MPO_TILE)
        tiles += [(j, t) for t in range(nt)]
        ref_jobs.append((R, S, i_offs, stride_in, o_offs, stride_out, cm[:, :nin]))
    src = rng.standard_normal(max(src_size, 1))
    out = np.full(dst_size, np.nan)
    for R, S, i_offs, si, o_offs, so, cm in ref_jobs:
        x = np.stack([src[o:o + R * si].reshape(R, si)[:, :S] for o in i_offs]) if i_offs else np.zeros((0, R, S))
        y = np.einsum("ok,krs->ors", cm, x)
        for k, o in enumerate(o_offs):
            out[o:o + R * so].reshape(R, so)[:, :S] = y[k]
    dev = "cuda"
    t = lambda a, d: torch.as_tensor(np.asarray(a, dtype=d).reshape(-1), device=dev)
    J, I, O, T = t(jobs, np.int64), t(ins, np.int64), t(outs, np.int64), t(tiles, np.int64)
    Cf = torch.as_tensor(np.asarray(coef), dtype=dtype, device=dev)
    X = torch.as_tensor(src, dtype=dtype, device=dev)
    Y = torch.full((dst_size,), float("nan"), dtype=dtype, device=dev)
    code = _native.TN_F64 if dtype == torch.float64 else _native.TN_F32
    _native.check(_native.lib().tn_mpo_apply(code, X.data_ptr(), Y.data_ptr(), J.data_ptr(), len(jobs),
                                             I.data_ptr(), O.data_ptr(), Cf.data_ptr(), T.data_ptr(), len(tiles),
                                             _native.stream_ptr()))
    got = Y.double().cpu().numpy()
    written = ~np.isnan(out)
    assert np.array_equal(np.isnan(got), np.isnan(out))  # exactly the slab positions written
    tol = 1e-13 if dtype == torch.float64 else 2e-6
    assert np.abs(got[written] - out[written]).max() <= tol * np.abs(out[written]).max()

"""``tn_gemm_grouped_static`` (the grouped GEMM that allocates nothing: split-K partials in
the library's static device scratch) against ``tn_gemm_grouped`` with a caller workspace and
its separate slice-sum launch.  Two calls in a row, the second on a smaller problem, must
each equal the allocating entry point bit for bit (same kernels, same summation order); a
workspace beyond the scratch is refused before anything runs."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _program(m_sizes, k, n, dtype, seed):
    from tenet.engine import GemmProgram

    rng = np.random.default_rng(seed)
    problems, a_off, b_off, c_off = [], 0, 0, 0
    for m in m_sizes:
        problems.append((m, n, c_off, n, [(a_off, b_off, k, k, n)]))
        a_off += m * k
        b_off += k * n
        c_off += m * n
    prog = GemmProgram(problems, a_off, b_off, c_off)
    A = torch.from_numpy(rng.standard_normal(a_off)).to("cuda", dtype)
    B = torch.from_numpy(rng.standard_normal(b_off)).to("cuda", dtype)
    return prog, A, B, c_off


@pytest.mark.parametrize("dtype,prec", [(torch.float64, "fp64"), (torch.float32, "fp32"),
                                        (torch.float32, "bf16x3")])
def test_static_entry_equals_the_allocating_one(dtype, prec):
    from tenet import _native

    for m_sizes, k, n, seed in (((130, 70, 200), 1500, 96, 1), ((40, 33), 700, 50, 2)):
        prog, A, B, csize = _program(m_sizes, k, n, dtype, seed)
        print(f"[no-alloc] {prec} M={m_sizes} K={k} N={n} ws={prog.ws_size}", flush=True)
        assert prog.reduce is not None and prog.ws_size > 0  # the split-K path is exercised
        assert prog.ws_size * A.element_size() <= _native.lib().tn_static_scratch_bytes()
        want = torch.full((csize,), float("nan"), dtype=dtype, device="cuda")
        prog.run(A, B, want, prec)
        got = torch.full((csize,), float("nan"), dtype=dtype, device="cuda")
        prog.run_static(A, B, got, prec)
        torch.cuda.synchronize()
        assert torch.equal(got, want)
        ref = np.concatenate([
            (np.asarray(A[o:o + m * k].double().cpu()).reshape(m, k) @
             np.asarray(B[ob:ob + k * n].double().cpu()).reshape(k, n)).ravel()
            for (m, _, _, _, [(o, ob, *_)]) in prog.problems])
        err = np.abs(got.double().cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < (1e-13 if prec == "fp64" else 1e-5)


def test_static_entry_refuses_a_workspace_beyond_its_scratch():
    import ctypes

    from tenet import _native

    lib = _native.lib()
    cap = lib.tn_static_scratch_bytes()
    rc = lib.tn_gemm_grouped_static(1, 1, 0, 0, None, None, None, None, None, None, 1, 64, 64, 0,
                                    cap // 4 + 1, None, None, None, None, 0, _native.stream_ptr())
    assert rc != 0
    assert b"static scratch" in lib.tn_last_error()
    _ = ctypes

"""NumPy interpreter of the engine's compiled device tables (recouple jobs, grouped-GEMM
problem/segment tables, MPO slab jobs), following the semantics documented in
``include/tenet_amd.h``.  Tests run compiled programs through it on the CPU and compare the
result with the oracle, so a wrong offset, stride or coefficient in a table fails without a
device."""

import numpy as np


def run_recouple(prog, src, dst):
    """``dst[region] = sum_t coeff_t * src[term_t + permuted index]`` for every job."""
    from tenet.engine import JOB_WORDS

    h = prog._tables.host
    jobs = np.asarray(h["jobs"]).reshape(-1, JOB_WORDS)
    for j in jobs:
        doff, tb, te, rank = (int(v) for v in j[:4])
        shape, dstr, sstr = j[4:4 + rank], j[10:10 + rank], j[16:16 + rank]
        idx = np.indices(shape).reshape(rank, -1)
        d = doff + (dstr[:, None] * idx).sum(0)
        s = (sstr[:, None] * idx).sum(0)
        acc = np.zeros(d.shape[0])
        for t in range(tb, te):
            acc += h["tco"][t] * src[h["tsrc"][t] + s]
        dst[d] = acc


def run_gemm(g, A, B, C):
    """``C_p = sum_s op(A_s) op(B_s)`` over the compiled problem and segment tables; flagged
    problems write the workspace (allocated here), then the program's own reduction runs."""
    W = np.zeros(max(g.ws_size, 1))
    for M, N_, c_off, ldc, sb, se, flag, rb in g._probs:
        acc = np.zeros((M, N_))
        for a_off, b_off, K, lda, ldb, _ in g._segs[sb:se]:
            m, n, k = np.arange(M)[:, None], np.arange(N_)[None, :], np.arange(K)
            if g.a_rows is not None:  # tn_gemm_grouped_rows: row m at a_rows[word 7 + m]
                Am = A[a_off + g._arows[rb + np.arange(M)][:, None] + k[None, :]]
            elif g.trans_a:
                Am = A[a_off + k[None, :] * lda + m]
            else:
                Am = A[a_off + m * lda + k[None, :]]
            Bm = B[b_off + n * ldb + k[:, None]] if g.trans_b else B[b_off + k[:, None] * ldb + n]
            acc += Am @ Bm
        out = W if flag else C
        out[c_off + np.arange(M)[:, None] * ldc + np.arange(N_)[None, :]] = acc
    if g.reduce is not None:
        run_recouple(g.reduce, W, C)
    return W


def run_mpo(jobs, ins, outs, coef, t1, t3, width=32):
    for R, S, ib, ie, ob, oe, cf, _ in jobs:
        r, s = np.arange(R)[:, None], np.arange(S)[None, :]
        v = [t1[off + r * rs + s] for off, rs in ins[ib:ie]]
        for o, (off, rs) in enumerate(outs[ob:oe]):
            acc = np.zeros((R, S))
            for k, vk in enumerate(v):
                acc += coef[cf + o * width + k] * vk
            t3[off + r * rs + s] = acc

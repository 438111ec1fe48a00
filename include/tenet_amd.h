/*
 * tenet_amd.h -- C ABI of the MI355X (gfx950) block-sparse contraction engine.
 *
 * Every entry point returns 0 on success and a nonzero code on failure; the message is
 * read with tn_last_error().  No exception crosses the ABI.  All tensor and table
 * pointers are DEVICE pointers owned by the caller (the Python host keeps them in torch
 * tensors).  Every call takes a hipStream_t (passed as void*), enqueues asynchronously
 * and never synchronizes; calls on distinct streams are independent.
 *
 * Data model: a block-sparse tensor is ONE flat buffer holding its blocks back to back in
 * the reference's canonical block order, each block C-contiguous in public axis order.
 * Plans (block offsets, permutations, recoupling coefficients, GEMM problem lists) are
 * compiled once on the host and uploaded as int64/double tables.
 *
 * Reference interfaces replaced (paths in Ryo-wtnb11/TeNeT-py):
 *   tn_recouple      <- the block transposes + coeff-scaled accumulation executed by
 *                       ops/repartition.py:458-468, ops/permutation.py:254-264 and the
 *                       gather/scatter of map_view.py:274-278 (to_matrices) and
 *                       map_view.py:331-334 (from_matrices).
 *   tn_gemm_grouped  <- the per-coupled-sector ar.do("matmul") loop of ops/map.py:160-161
 *                       (compose), i.e. the GEMM batch under tenet.einsum / tensordot.
 *   tn_dot, tn_axpby, tn_scale
 *                    <- tenet.inner / tenet.norm / add / subtract / scalar multiply
 *                       (ops/contraction.py:584-589, ops/basic.py:119,157,361-367) that
 *                       network/dmrg.py:106-227 (lanczos) runs on.
 *   tn_cast          <- ar.do("astype") on blocks (precision staging for FP32 GEMMs).
 *   tn_mpo_apply     <- the two W contractions of network/env.py:430-431 (_heff2_full),
 *                       fused into one pass with a host-merged two-site MPO table.
 */
#ifndef TENET_AMD_H
#define TENET_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* element types */
#define TN_F64 0
#define TN_F32 1

/* GEMM arithmetic: FP64 MFMA, FP32 MFMA, and FP32 emulated by three BF16 MFMA products
 * (hi*hi + hi*lo + lo*hi) with FP32 accumulation. */
#define TN_PREC_FP64 0
#define TN_PREC_FP32 1
#define TN_PREC_BF16X3 2

/* Recouple job table: TN_JOB_WORDS int64 per destination region.
 *   [0] dst_off   [1] term_begin  [2] term_end  [3] rank (<= TN_MAX_RANK)
 *   [4..9] shape  [10..15] dst_stride  [16..21] src_stride (dims in dst C order)
 *   [22] mode: -1 = linear (dst-ordered, 2048 elements per tile; -2: 512 per tile), else
 *        the index J of the src-contiguous dim used as the second axis of a 64x64 LDS
 *        transpose tile (J + 8: a 32x32 tile)
 *   [23] element count
 * Term table: per term one int64 source offset (elements) and one double coefficient.
 * Tile table: 2 int64 per tile: (job, tile index within the job). */
#define TN_JOB_WORDS 24
#define TN_MAX_RANK 6

/* GEMM problem table: TN_PROB_WORDS int64 per problem (one output block, or one K slice
 * of one when the host splits K):
 *   [0] M [1] N [2] c_off [3] ldc [4] seg_begin [5] seg_end
 *   [6] 0 = write to C, 1 = write to the workspace W (split-K partial)
 *   [7] 0 (tn_gemm_grouped_rows: the problem's first entry in its row table)
 * Segment table: TN_SEG_WORDS int64 per K segment:
 *   [0] a_off [1] b_off [2] K [3] lda [4] ldb [5] 0
 * Tile table: 4 int32 per output tile: (problem, tile row, tile col, the problem's first
 * segment [4]). */
#define TN_PROB_WORDS 8
#define TN_SEG_WORDS 6

const char* tn_last_error(void);
int tn_version(void);

/* Number of compute units of the current device (for persistent-grid sizing). */
int tn_device_cus(int* out);

/* dst[job region] = sum_t coeff_t * permute(src + term_src_off_t), over every job. */
int tn_recouple(int dtype, const void* src, void* dst, const int64_t* jobs,
                const int64_t* term_src, const double* term_coeff, const int64_t* tiles,
                int64_t ntiles, void* stream);
/* The same with a type conversion on the way (src_dtype -> dst_dtype, arithmetic in the
 * wider type): the precision staging of an FP32 H_eff (ar.do("astype") on the blocks)
 * fused with the block permutation into the layout its GEMMs read. */
int tn_recouple_cast(int src_dtype, int dst_dtype, const void* src, void* dst,
                     const int64_t* jobs, const int64_t* term_src, const double* term_coeff,
                     const int64_t* tiles, int64_t ntiles, void* stream);

/* Host-side plan compilation (no device call): the job table of one tn_recouple launch from
 * its n destination regions, replacing the per-term block loop of the reference's plan
 * application (ops/permutation.py:254-262, ops/repartition.py's apply) with one compiled
 * table.  Region j: dst_off[j], and rank dims (rows j of the n x rank tables shape,
 * dst_strides, src_strides, destination C order, unit dims allowed); its terms are
 * term_src[term_ptr[j] .. term_ptr[j+1]).  Writes jobs (n x TN_JOB_WORDS, unit dims dropped,
 * dims contiguous on both sides merged, the mode word chosen) and tile_counts (n), and
 * stats = {elements, element-terms, total tiles}; lin_tile / lin_small_tile / lds_tile /
 * small_program_tiles are the kernel's tile sizes and the small-program threshold below
 * which the small tiles are used.  Returns 1 (a write outside dst_size), 2 (a region
 * spanning 2^31 elements), 3 (a read outside src_size), 4 (rank above TN_MAX_RANK after
 * merging) or 5 (bad tile sizes), with tn_last_error set. */
int tn_recouple_plan(int64_t n, int64_t rank, const int64_t* dst_off, const int64_t* shape,
                     const int64_t* dst_strides, const int64_t* src_strides, const int64_t* term_ptr,
                     const int64_t* term_src, int64_t src_size, int64_t dst_size, int64_t lin_tile,
                     int64_t lin_small_tile, int64_t lds_tile, int64_t small_program_tiles,
                     int64_t* jobs, int64_t* tile_counts, int64_t* stats);

/* For every problem p: C_p = sum_s op(A_s) op(B_s) with op = transpose when trans_a/b,
 * written to C, or to the workspace W when the problem is flagged as a split-K partial.
 * A/B/C/W are base pointers; every offset in the tables is in elements.  bm x bn is the
 * workgroup tile (one of 32x32, 64x64, 128x64, 64x128, 128x128); grid 0 = one workgroup
 * per tile, otherwise a persistent loop over tiles. */
int tn_gemm_grouped(int precision, int dtype, int trans_a, int trans_b, const void* A,
                    const void* B, void* C, void* W, const int64_t* problems,
                    const int64_t* segs, const int32_t* tiles, int64_t ntiles, int bm, int bn,
                    int grid, void* stream);

/* tn_gemm_grouped with A's rows gathered through a row table, replacing the to_matrices
 * gather of the left operand (map_view.py:274-278) that precedes the per-sector matmul of
 * ops/map.py:160-161 when a sector matrix's rows lie in several blocks: element (m, k) of
 * problem p's segment s is A[seg.a_off + a_rows[problem word 7 + m] + k] (trans_a = 0).
 * A is FP64 or FP32 (a_dtype; FP64 is rounded to FP32 as it is staged, the precision
 * staging of an FP32 GEMM fused with the gather); B (row-major, trans_b = 0), C and W are
 * FP32; precision is TN_PREC_FP32 or TN_PREC_BF16X3; one workgroup per tile (bm x bn one of
 * 64x64, 128x64, 64x128, 128x128).  Problem word 7 must index a_rows so that every row of
 * the problem has an entry; byte offsets within a segment are 32-bit. */
int tn_gemm_grouped_rows(int precision, int a_dtype, const void* A, const int64_t* a_rows,
                         const void* B, void* C, void* W, const int64_t* problems,
                         const int64_t* segs, const int32_t* tiles, int64_t ntiles, int bm, int bn,
                         void* stream);

/* No counterpart in the reference: tn_gemm_grouped for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * split-K partial products go to a static device array of the library, at most
 * tn_static_scratch_bytes() bytes (64 MiB; larger workspaces are refused with an error),
 * instead of a caller-provided W; when red_ntiles > 0 a tn_recouple launch with the given
 * job/term/tile tables (source = the scratch, destination = C) then sums the slices into C.
 * ws_elems is the workspace the problem tables address, in elements of dtype.  Every slice
 * element is written before it is read, so nothing needs clearing, but all calls on one
 * device share that device's array: calls that overlap in time on one GPU (two streams,
 * two threads) race and must be serialized by the caller.  No hipMalloc/hipFree. */
int tn_gemm_grouped_static(int precision, int dtype, int trans_a, int trans_b, const void* A,
                           const void* B, void* C, const int64_t* problems, const int64_t* segs,
                           const int32_t* tiles, int64_t ntiles, int bm, int bn, int grid,
                           int64_t ws_elems, const int64_t* red_jobs, const int64_t* red_src,
                           const double* red_coeff, const int64_t* red_tiles, int64_t red_ntiles,
                           void* stream);
int64_t tn_static_scratch_bytes(void);

/* The same grouped product for skinny problems (the W contractions of the generic two-site
 * H_eff, network/env.py:430-431: K and one output dimension of a few MPO/physical
 * channels): rows != 0 -- every problem has N <= 8, one thread per output row; rows == 0
 * -- every problem has M <= 8, one thread per output column.  Same problem and segment
 * tables as tn_gemm_grouped; arithmetic in the operands' type.  Tile table: 2 int32 per
 * tile: (problem, block of 256 output lines). */
int tn_gemm_skinny(int dtype, int trans_a, int trans_b, int rows, const void* A, const void* B,
                   void* C, void* W, const int64_t* problems, const int64_t* segs,
                   const int32_t* tiles, int64_t ntiles, void* stream);

/* out[0] = sum_i w_i x_i y_i (w may be NULL for unit weights), FP64 accumulation,
 * deterministic two-pass reduction; `work` must hold >= 1024 doubles. */
int tn_dot(int dtype, const void* x, const void* y, const double* w, int64_t n,
           double* work, double* out, void* stream);

/* y = alpha * x + beta * y with alpha/beta read from device scalars when the pointer
 * form is used (alpha_dev/beta_dev non-NULL) -- no host synchronization. The scalar
 * used is  s_host * (s_dev ? *s_dev : 1) , optionally inverted when inv_* is set.  A
 * factor equal to 0 drops its term without reading the operand (so 0 * inf is 0). */
int tn_axpby(int dtype, int64_t n, double alpha, const double* alpha_dev, int inv_alpha,
             const void* x, double beta, const double* beta_dev, int inv_beta, void* y,
             void* stream);

/* Gram-Schmidt of a Krylov vector against the basis (the re-orthogonalising three-term
 * step of network/dmrg.py:106-227 (lanczos), which there runs as one inner and one
 * subtract per basis vector): `passes` (1 or 2) classical Gram-Schmidt passes of w against
 * the k (1..8) basis vectors basis[0..k) (a host array of device pointers, dtype like w):
 * c_i = <v_i, w>_W (weights W may be NULL), w <- w - sum_i c_i v_i.  coef[0..k) (device)
 * receives the first pass's c_i and *last_coef (device, may be NULL) its c_{k-1}; *norm
 * (device) receives |w|_W after the last pass.  FP64 accumulation, deterministic order;
 * `work` must hold tn_krylov_orth_work() doubles.  Launches only (graph-capturable). */
int tn_krylov_orth(int dtype, int64_t n, int k, const void* const* basis, void* w,
                   const double* weights, int passes, double* coef, double* last_coef,
                   double* norm, double* work, void* stream);
int64_t tn_krylov_orth_work(void);

/* x *= s (s host scalar times optional device scalar, optionally inverted). */
int tn_scale(int dtype, int64_t n, double s, const double* s_dev, int inv, void* x,
             void* stream);

/* dst = (dst_dtype) src elementwise. */
int tn_cast(int src_dtype, const void* src, int dst_dtype, void* dst, int64_t n, void* stream);

/* Fused two-site MPO application (the middle of the two-site H_eff).  For each job
 * (a pair of outer slabs) with R rows and S contiguous columns:
 *   out[o][r][s] = sum_k coef[o][k] * in[k][r][s]
 * where in/out slabs are addressed by per-job offset lists.  Job table per job (int64):
 *   [0] R [1] S [2] in_begin [3] in_end [4] out_begin [5] out_end [6] coef_off [7] 0
 * in/out lists hold (offset, row_stride) int64 pairs (at most TN_MPO_MAX_IN inputs per
 * job); coef holds, per job, nout rows of TN_MPO_MAX_IN coefficients (zero-padded) in the
 * element type.  Tile table: 2 int64 per tile: (job, block of 1024 positions). */
#define TN_MPO_JOB_WORDS 8
#define TN_MPO_MAX_IN 32
int tn_mpo_apply(int dtype, const void* in, void* out, const int64_t* jobs, int64_t njobs,
                 const int64_t* in_slabs, const int64_t* out_slabs, const void* coef,
                 const int64_t* tiles, int64_t ntiles, void* stream);

/* Batched one-sided Jacobi SVD (FP64) of the sector matrices of one blockwise SVD: replaces
 * the per-coupled-sector ar.do("linalg.svd", ...) of ops/linalg.py:229 (svd), :1436
 * (select_bond) and :1593 (svd_truncated).  Problem p factors "k vectors of length L"
 * stored row-major at x + x_off (the rows of A when rows <= cols, else the rows of A^T);
 * q + q_off holds its k x k orthogonal factor.  Problem table, TN_SVD_PROB_WORDS int64:
 *   [0] k [1] L [2] x_off [3] q_off [4] nb (even block count, blocks of w vectors)
 *   [5] cols (0: x holds A, 1: x holds A^T) [6] u_off [7] s_off [8] vh_off
 *   [9] inner sweeps per block-pair visit  [10] the problem's slot in fro2
 * tn_svd_jacobi_norms writes fro2[slot] = |A|_F^2 of every problem (run after x is filled,
 * before the first round).  With null_scale > 0 a vector whose squared norm is at or below
 * floor^2 = null_scale k^2 fro2 is numerically null: it still rotates, but its rotations
 * do not count towards convergence, and tn_svd_jacobi_finish (same null_scale) leaves its
 * row of the length-L factor zero unless asked to complete it (complete != 0: Gram-Schmidt
 * of unit vectors against the rows already set, so U and Vh are isometries on every
 * column).  null_scale = 0: plain one-sided Jacobi, every rotation counts.
 * tn_svd_jacobi_round runs round `round` of the block tournament: workgroup i takes block
 * pair slot wgs[2i+1] of problem wgs[2i]; LDS is 2 w span doubles, at least 2 w (L + k)
 * for every problem of the launch with nb > 2 and (k + k % 2) (L + k) for one with nb = 2
 * (held by one block pair, which converges inside the launch); first != 0 initialises
 * Q = I on load (round 0 of the first sweep);
 * rotations[i] receives the rotation count of its last inner sweep.  A pair rotates when
 * |<x_a,x_b>| > tol * |x_a| |x_b|.
 * tn_svd_jacobi_finish writes, per problem, S (k, descending), U (rows x k, row-major) and
 * Vh (k x cols, row-major) of the thin SVD. */
#define TN_SVD_PROB_WORDS 11
#define TN_LQ_PROB_WORDS 6
#define TN_SVD_MAX_W 64
#define TN_SVD_MAX_K 4096
int tn_svd_jacobi_norms(const int64_t* probs, int64_t nprobs, const double* x, double* fro2,
                        void* stream);
int tn_svd_jacobi_round(const int64_t* probs, const int32_t* wgs, int64_t nwgs, int64_t span,
                        double* x, double* q, int32_t* rotations, int w, int round, int first,
                        double tol, const int32_t* stop, const double* fro2, double null_scale,
                        void* stream);
/* *stop = 1 when all n rotation counts of a sweep (rotations of every round of one launch
 * group) are zero.  With `stop` passed to tn_svd_jacobi_round (NULL: always run), the host
 * queues several sweeps ahead and reads one flag instead of the counts of every sweep;
 * rounds past convergence return at once. */
int tn_svd_jacobi_converged(const int32_t* rotations, int64_t n, int32_t* stop, void* stream);
int tn_svd_jacobi_finish(const int64_t* probs, int64_t nprobs, const double* x, const double* q,
                         double* U, double* S, double* Vh, const double* fro2, double null_scale,
                         int complete, void* stream);

/* LQ of every sector matrix, replacing linalg.py:549-620 (lq: QR of each sector matrix's
 * transpose per coupled sector).  Problem p (TN_LQ_PROB_WORDS int64: m, n, a_off, l_off,
 * q_off, 0): A_p (m x n, row-major at A + a_off, overwritten by the Householder vectors)
 * = L_p Q_p with k = min(m, n), L_p (m x k, lower trapezoidal, at L + l_off) and Q_p (k x n,
 * orthonormal rows, at Q + q_off), FP64; the diagonal of L carries LAPACK dgeqrf's signs
 * (beta_j = -sign(alpha) |x|).  One workgroup per problem; max_lds_doubles >= n + 2k of
 * every problem. */
int tn_lq(const int64_t* probs, int64_t nprobs, int64_t max_lds_doubles, double* A, double* L,
          double* Q, void* stream);

/* Ground eigenpair of the Lanczos tridiagonal matrix, on the device: replaces the
 * np.linalg.eigh(tri) of network/dmrg.py:216-222 (lanczos) and the host read of the Krylov
 * scalars before it (:212).  alpha[0..n) is the diagonal, beta[0..n-1) the off-diagonal;
 * the matrix is truncated at the first |beta[j]| < tol (happy breakdown, :187-190).
 * Writes value[0] (lowest eigenvalue) and coef[0..n) (its unit eigenvector, zero past the
 * truncation, first nonzero entry positive).  One workgroup; n <= TN_TRIDIAG_MAX.
 * tn_tridiag_ground_host runs the same routine on the host (host pointers, no stream). */
#define TN_TRIDIAG_MAX 64
int tn_tridiag_ground(const double* alpha, const double* beta, int n, double tol, double* value,
                      double* coef, void* stream);
int tn_tridiag_ground_host(const double* alpha, const double* beta, int n, double tol,
                           double* value, double* coef);

/* Many independent problems in one arena (the batched Lanczos of many H_eff problems,
 * reference network/dmrg.py:106-227 run for each): problem s owns the elements
 * [off_s, off_{s+1}).  `chunks` lists (segment, begin, end) int64 triples that cover every
 * problem's range in pieces of bounded size; seg_chunks[s] .. seg_chunks[s+1] are problem
 * s's chunks, in order.
 * tn_dot_segmented: out[s * out_stride] = sum_i w_i x_i y_i over problem s (w may be NULL),
 * FP64, deterministic (one partial per chunk into `partial` (nchunks doubles), then the
 * partials of a problem summed in chunk order).
 * tn_axpby_segmented: y = (a * A_s) x + (b * B_s) y with A_s = a_dev[s * a_stride] (or
 * its inverse when a_inv; 1 when a_dev is NULL), likewise B_s; a zero factor drops its
 * operand unread, as tn_axpby.
 * tn_tridiag_ground_batched: tn_tridiag_ground for nprob problems in one launch, problem p's
 * matrix at alpha + p ld / beta + p ld, its vector at coefs + p ld, its value at values[p]. */
int tn_dot_segmented(int dtype, const void* x, const void* y, const double* w, const int64_t* chunks,
                     int64_t nchunks, double* partial, const int64_t* seg_chunks, int64_t nseg,
                     double* out, int64_t out_stride, void* stream);
int tn_axpby_segmented(int dtype, const int64_t* chunks, int64_t nchunks, double a, const double* a_dev,
                       int64_t a_stride, int a_inv, const void* x, double b, const double* b_dev,
                       int64_t b_stride, int b_inv, void* y, void* stream);
int tn_tridiag_ground_batched(const double* alpha, const double* beta, int n, int64_t ld, int64_t nprob,
                              double tol, double* values, double* coefs, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TENET_AMD_H */

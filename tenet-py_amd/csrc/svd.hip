// Batched one-sided Jacobi SVD in FP64 for the truncating blockwise SVD.
//
// Every sector matrix of one SVD call is a problem "k vectors of length L" stored row-major
// (rows of A when rows <= cols, else rows of A^T).  One-sided (Hestenes) Jacobi rotates
// pairs of vectors until all are mutually orthogonal; the same rotations applied to the
// rows of Q (k x k, starts as I) record the orthogonal factor.  At convergence
// X = Q A_rows with orthogonal rows, so A_rows = Q^T diag(sigma) W, W's rows orthonormal.
//
// Vectors are grouped in blocks of w.  One launch is one round of a circle-method
// tournament over the blocks of every problem: each workgroup takes one block pair, stages
// its 2w vectors and their Q rows in LDS, rotates the pairs it is due (round 0 of a sweep:
// every pair of the 2w vectors; later rounds: the w x w pairs across the two blocks; the
// pairs of an inner round side by side, one 16-lane group each, at most 64 groups), and
// writes both back.
// A problem whose vectors fit in one block pair converges inside a single launch.
// Each workgroup reports the rotations of its last inner sweep, so the host stops when a
// whole sweep reports none.

#include <algorithm>

#include "tn_common.hpp"

namespace {

// Singular values at or below this are treated as exact zeros (their W rows are completed
// to an orthonormal set instead of being divided by sigma).
constexpr double kNullSigma = 1e-290;

__device__ inline double wave_sum(double v) {
  // butterfly: every lane ends with the same value (IEEE addition commutes exactly)
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

typedef int i32x2 __attribute__((ext_vector_type(2)));

// A double moved across lanes by one DPP pattern (both 32-bit halves), VALU only.
template <int CTRL>
__device__ inline double dpp_move(double v) {
  const i32x2 p = __builtin_bit_cast(i32x2, v);
  const int lo = __builtin_amdgcn_update_dpp(0, p.x, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, p.y, CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, i32x2{lo, hi});
}

// all-reduce within each 16-lane row of the wave: lane ^ 1 and lane ^ 2 (quad_perm), the
// other quad of 8 lanes (row_half_mirror), the other 8 of 16 (row_mirror) -- DPP moves
// only, no LDS crossbar.  Every step adds two partial sums that both partners hold, so all
// 16 lanes end with the same bits.
__device__ inline double row_sum(double v) {
  v += dpp_move<0xB1>(v);
  v += dpp_move<0x4E>(v);
  v += dpp_move<0x141>(v);
  v += dpp_move<0x140>(v);
  return v;
}

// lanes per vector pair: one 16-lane DPP row (a wave works on four pairs at once; fewer
// lanes per pair, with longer per-lane chains, measured slower), or two rows for the long
// vectors of block tournaments (jacobi_round's PL)
constexpr int P = 16;

// the sum over a pair's PL lanes, the same bits in every lane: the 16-lane row sum, then
// for two rows the other row's sum (IEEE addition commutes, so both rows agree)
template <int PL>
__device__ inline double group_sum(double v) {
  v = row_sum(v);
  if constexpr (PL >= 32) v += __shfl_xor(v, 16, 64);
  if constexpr (PL == 64) v += __shfl_xor(v, 32, 64);
  return v;
}
// jacobi_round's lanes per pair: at most TN_SVD_LANES (16, 32 or 64), widened while a lane
// keeps at least TN_SVD_LANE_ELEMS elements of a vector
#ifndef TN_SVD_LANES
#define TN_SVD_LANES 64
#endif
#ifndef TN_SVD_LANE_ELEMS
#define TN_SVD_LANE_ELEMS 12
#endif
// global loads a thread keeps in flight while staging a block pair into LDS
constexpr int kStage = 8;

// Numerically null vectors.  With null_scale > 0 (a truncating caller), a vector whose
// squared norm is at or below floor2 = null_scale k^2 |A|_F^2 (null_scale = eps^2: the
// floor is k eps |A|_F) is numerically null: its direction is rounding noise, and the
// overlap of such a pair is re-created by rounding every sweep, so it never stops
// rotating.  Those pairs still rotate (the null vectors' norms keep converging to their
// singular values) but do not count towards convergence: the sweeps stop once every pair
// of vectors above the floor is orthogonal.  The length-L factor rows of null vectors are
// then not orthonormal, so they are left zero unless the caller asks for their
// completion.  null_scale = 0 (an exact SVD): every rotation counts, as in plain
// one-sided Jacobi.
__device__ inline double null_floor2(const double* fro2, const int64_t* pw, double null_scale) {
  const double k = static_cast<double>(pw[0]);
  return fro2[pw[10]] * (null_scale * k * k);
}

// |A|_F^2 of every problem (one workgroup each, fixed summation order).
__global__ __launch_bounds__(256) void jacobi_norms(const int64_t* __restrict__ probs,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ fro2) {
  __shared__ double part[4];
  const int64_t* pw = probs + TN_SVD_PROB_WORDS * blockIdx.x;
  const int64_t n = pw[0] * pw[1];
  const double* xp = x + pw[2];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) acc = fma(xp[i], xp[i], acc);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) fro2[pw[10]] = (part[0] + part[1]) + (part[2] + part[3]);
}

// LDS elements a lane of a pair group moves per step of its dot products and rotations
constexpr int kChunk = 8;

// x_a <- x_a + ((c - 1) x_a - s x_b), x_b <- x_b + ((c - 1) x_b + s x_a) over n elements,
// lane gl of the pair's PL-lane group taking i = gl + PL j.  Each chunk's loads are issued
// together; u0/v0 (when given) already hold the first chunk.
template <int PL>
__device__ inline void rotate_rows(double* xa, double* xb, int n, int gl, double cm1, double s,
                                   const double* u0, const double* v0) {
  for (int base = 0; base < n; base += PL * kChunk) {
    double u[kChunk], v[kChunk];
    if (base == 0 && u0 != nullptr) {
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        u[j] = u0[j];
        v[j] = v0[j];
      }
    } else {
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        const int i = base + gl + PL * j;
        const int ii = i < n ? i : 0;
        u[j] = xa[ii];
        v[j] = xb[ii];
      }
    }
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const int i = base + gl + PL * j;
      if (i < n) {
        xa[i] = u[j] + fma(cm1, u[j], -s * v[j]);
        xb[i] = v[j] + fma(cm1, v[j], s * u[j]);
      }
    }
  }
}

// circle method over n (even) players: round r in [0, n-1), pair p in [0, n/2)
__device__ inline void pair_of(int n, int r, int p, int& a, int& b) {
  // both sums lie in [0, 2 (n - 1)): one conditional subtraction replaces the modulo
  int ia = p - 1 + r, ib = n - 2 - p + r;
  ia -= ia >= n - 1 ? n - 1 : 0;
  ib -= ib >= n - 1 ? n - 1 : 0;
  a = p == 0 ? 0 : ia + 1;
  b = ib + 1;
}

template <int PL>
__global__ void jacobi_round(const int64_t* __restrict__ probs, const int32_t* __restrict__ wgs,
                             double* __restrict__ x, double* __restrict__ q,
                             int32_t* __restrict__ rot_out, int w, int round, int first,
                             double tol, const int32_t* __restrict__ stop,
                             const double* __restrict__ fro2, double null_scale) {
  extern __shared__ double lds[];
  __shared__ int rot_s;
  // the launch's problems have converged (set by jacobi_converged after an earlier sweep):
  // the host queues sweeps ahead without reading the rotation counts back, and those past
  // convergence leave every vector as it is
  if (stop != nullptr && *stop != 0) return;
  const int64_t* pw = probs + TN_SVD_PROB_WORDS * wgs[2 * blockIdx.x];
  const int slot = wgs[2 * blockIdx.x + 1];
  const int k = static_cast<int>(pw[0]);
  const int L = static_cast<int>(pw[1]);
  const int64_t xo = pw[2], qo = pw[3];
  const int nb = static_cast<int>(pw[4]);
  const int inner_sweeps = static_cast<int>(pw[9]);
  if (round >= nb - 1) {
    if (threadIdx.x == 0) rot_out[blockIdx.x] = 0;
    return;
  }
  int bi, bj;
  pair_of(nb, round, slot, bi, bj);
  const int n2 = 2 * w;
  // a problem held by one block pair (bi = 0, bj = 1, vector t is vector t) stages and
  // runs its inner tournament over its own k vectors (one zero vector when k is odd), not
  // over the 2w slots: the zero slots would only cost LDS and dot products that never
  // rotate
  const int ne = nb == 2 ? min(n2, k + (k & 1)) : n2;
  double* xs = lds;                                  // ne x L
  double* qs = lds + static_cast<size_t>(ne) * L;    // ne x k
  // staging: the ne x L vector slab and the ne x k Q slab walked as flat index spaces,
  // kStage loads of a thread issued before any of them is stored, so a workgroup has
  // many global loads in flight instead of one vector row at a time
  {
    const int nx = ne * L, nq = ne * k;
    for (int base = 0; base < nx; base += kStage * blockDim.x) {
      double v[kStage];
#pragma unroll
      for (int u = 0; u < kStage; ++u) {
        const int e = base + u * blockDim.x + threadIdx.x;
        const int t = e / L, i = e - (e / L) * L;
        const int g = t < w ? bi * w + t : bj * w + (t - w);
        // unconditional load (an element outside the problem reads its first element)
        const bool ok = e < nx && g < k;
        const double val = x[xo + (ok ? static_cast<int64_t>(g) * L + i : 0)];
        v[u] = ok ? val : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kStage; ++u) {
        const int e = base + u * blockDim.x + threadIdx.x;
        if (e < nx) xs[e] = v[u];
      }
    }
    if (first) {  // Q starts as the identity
      for (int e = threadIdx.x; e < nq; e += blockDim.x) {
        const int t = e / k, i = e - (e / k) * k;
        const int g = t < w ? bi * w + t : bj * w + (t - w);
        qs[e] = g < k && i == g ? 1.0 : 0.0;
      }
    } else {
      for (int base = 0; base < nq; base += kStage * blockDim.x) {
        double v[kStage];
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
          const int e = base + u * blockDim.x + threadIdx.x;
          const int t = e / k, i = e - (e / k) * k;
          const int g = t < w ? bi * w + t : bj * w + (t - w);
          const bool ok = e < nq && g < k;
          const double val = q[qo + (ok ? static_cast<int64_t>(g) * k + i : 0)];
          v[u] = ok ? val : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kStage; ++u) {
          const int e = base + u * blockDim.x + threadIdx.x;
          if (e < nq) qs[e] = v[u];
        }
      }
    }
  }
  // one pair per PL-lane group: a round's pairs run side by side
  const int gl = threadIdx.x % PL;
  const int grp = threadIdx.x / PL;
  const int ngrp = blockDim.x / PL;
  const double tol2 = tol * tol;
  const double floor2 = null_floor2(fro2, pw, null_scale);
  // Round 0 of a sweep meets every block once, so its inner tournament (all pairs of the
  // 2w vectors) visits the pairs inside every block of the problem.  Later rounds visit
  // only the w x w pairs across their two blocks, w inner rounds of w disjoint pairs
  // (slot p of the first block with slot w + (p + r) mod w of the second): every pair of
  // the problem is visited exactly once per sweep, and a sweep costs about half the inner
  // rounds of one that re-runs the pairs inside both blocks every round.
  const bool cross = nb > 2 && round > 0;
  const int nrounds = cross ? w : ne - 1;
  const int npairs = cross ? w : ne / 2;
  int last = 0;
  for (int it = 0; it < inner_sweeps; ++it) {
    if (threadIdx.x == 0) rot_s = 0;
    __syncthreads();
    int mine = 0;
    for (int r = 0; r < nrounds; ++r) {
      for (int p = grp; p < npairs; p += ngrp) {
        int a, b;
        if (cross) {
          a = p;
          b = w + (p + r >= w ? p + r - w : p + r);
        } else {
          pair_of(ne, r, p, a, b);
        }
        double* xa = xs + a * L;
        double* xb = xs + b * L;
        // a lane's elements i = gl + PL j, taken kChunk at a time with every LDS load of a
        // chunk issued before its arithmetic (a lane past the end reads element 0 and
        // contributes zero); the first chunk stays in registers for the rotation
        double al = 0.0, be = 0.0, ga = 0.0;
        double u0[kChunk], v0[kChunk];
        for (int base = 0; base < L; base += PL * kChunk) {
          double u[kChunk], v[kChunk];
#pragma unroll
          for (int j = 0; j < kChunk; ++j) {
            const int i = base + gl + PL * j;
            const int ii = i < L ? i : 0;
            u[j] = xa[ii];
            v[j] = xb[ii];
          }
#pragma unroll
          for (int j = 0; j < kChunk; ++j) {
            const bool ok = base + gl + PL * j < L;
            const double uu = ok ? u[j] : 0.0, vv = ok ? v[j] : 0.0;
            al = fma(uu, uu, al);
            be = fma(vv, vv, be);
            ga = fma(uu, vv, ga);
            if (base == 0) {
              u0[j] = uu;
              v0[j] = vv;
            }
          }
        }
        al = group_sum<PL>(al);
        be = group_sum<PL>(be);
        ga = group_sum<PL>(ga);
        // uniform within the group: every lane of it holds the same sums
        if (al > 0.0 && be > 0.0 && ga * ga > tol2 * al * be) {
          // t = tan(theta), |theta| <= pi/4, zeroes <x_a, x_b>; applied as
          // x + ((c - 1) x - s y) with c - 1 = -t^2 / (1 + t^2 + r), r = sqrt(1 + t^2),
          // formed directly so that small-angle rotations (c rounding to exactly 1) do
          // not grow every norm by t^2 / 2 a step
          const double d = be - al;
          const double g2 = 2.0 * ga;
          const double hyp = sqrt(fma(d, d, g2 * g2));
          const double t = g2 / (d + copysign(hyp, d));
          const double tt = t * t;
          const double rr = sqrt(1.0 + tt);
          const double den = 1.0 + tt + rr;
          const double qd = 1.0 / (rr * den);
          const double s = t * den * qd;
          const double cm1 = -tt * rr * qd;
          rotate_rows<PL>(xa, xb, L, gl, cm1, s, u0, v0);
          rotate_rows<PL>(qs + a * k, qs + b * k, k, gl, cm1, s, nullptr, nullptr);
          mine += al > floor2 && be > floor2;
        }
      }
      __syncthreads();
    }
    if (gl == 0 && mine) atomicAdd(&rot_s, mine);
    __syncthreads();
    last = rot_s;
    __syncthreads();
    if (last == 0) break;
  }
  // write-back over the same flat index spaces (stores do not stall the issue)
  for (int e = threadIdx.x; e < ne * L; e += blockDim.x) {
    const int t = e / L, i = e - (e / L) * L;
    const int g = t < w ? bi * w + t : bj * w + (t - w);
    if (g < k) x[xo + static_cast<int64_t>(g) * L + i] = xs[e];
  }
  for (int e = threadIdx.x; e < ne * k; e += blockDim.x) {
    const int t = e / k, i = e - (e / k) * k;
    const int g = t < w ? bi * w + t : bj * w + (t - w);
    if (g < k) q[qo + static_cast<int64_t>(g) * k + i] = qs[e];
  }
  if (threadIdx.x == 0) rot_out[blockIdx.x] = last;
}

// *stop = 1 once a whole sweep of the launch group rotated nothing (n counts, one per
// workgroup of every round of the sweep).
__global__ __launch_bounds__(256) void jacobi_converged(const int32_t* __restrict__ rot, int64_t n,
                                                       int32_t* __restrict__ stop) {
  __shared__ int any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  int mine = 0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) mine |= rot[i];
  if (mine) atomicOr(&any, 1);
  __syncthreads();
  if (threadIdx.x == 0 && any == 0) *stop = 1;
}

// One workgroup per problem: singular values = row norms, descending order by rank
// counting, normalized rows and Q rows scattered into U / S / Vh.
__global__ void jacobi_finish(const int64_t* __restrict__ probs, const double* __restrict__ x,
                              const double* __restrict__ q, double* __restrict__ U,
                              double* __restrict__ S, double* __restrict__ Vh,
                              const double* __restrict__ fro2, double null_scale, int complete) {
  extern __shared__ double sig[];
  int* rank = reinterpret_cast<int*>(sig + TN_SVD_MAX_K);
  const int64_t* pw = probs + TN_SVD_PROB_WORDS * blockIdx.x;
  const int k = static_cast<int>(pw[0]);
  const int L = static_cast<int>(pw[1]);
  const int64_t xo = pw[2], qo = pw[3];
  const int cols = static_cast<int>(pw[5]);
  const int64_t uo = pw[6], so = pw[7], vo = pw[8];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  // rows at or below the numerical-null floor carry no direction of their own
  const double null_sigma = fmax(kNullSigma, sqrt(null_floor2(fro2, pw, null_scale)));
  for (int i = wave; i < k; i += nw) {
    const double* xr = x + xo + static_cast<int64_t>(i) * L;
    double acc = 0.0;
    for (int j = lane; j < L; j += 64) acc = fma(xr[j], xr[j], acc);
    acc = wave_sum(acc);
    if (lane == 0) sig[i] = sqrt(acc);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    const double si = sig[i];
    int rk = 0;
    for (int j = 0; j < k; ++j) rk += (sig[j] > si) || (sig[j] == si && j < i);
    rank[i] = rk;
    S[so + rk] = si;
  }
  __syncthreads();
  for (int i = 0; i < k; ++i) {
    const int rk = rank[i];
    const double si = sig[i];
    const double inv = si > null_sigma ? 1.0 / si : 0.0;
    const double* xr = x + xo + static_cast<int64_t>(i) * L;
    const double* qr = q + qo + static_cast<int64_t>(i) * k;
    if (!cols) {
      // A (k x L) = Q^T diag(sigma) W: U[:, rk] = Q row i, Vh[rk, :] = W row i
      for (int r = threadIdx.x; r < k; r += blockDim.x) U[uo + static_cast<int64_t>(r) * k + rk] = qr[r];
      for (int j = threadIdx.x; j < L; j += blockDim.x) Vh[vo + static_cast<int64_t>(rk) * L + j] = xr[j] * inv;
    } else {
      // A^T (k x L) = Q^T diag(sigma) W: U[:, rk] = W row i, Vh[rk, :] = Q row i
      for (int j = threadIdx.x; j < L; j += blockDim.x) U[uo + static_cast<int64_t>(j) * k + rk] = xr[j] * inv;
      for (int r = threadIdx.x; r < k; r += blockDim.x) Vh[vo + static_cast<int64_t>(rk) * k + r] = qr[r];
    }
  }
  // Null space.  A zero singular value leaves its W row zero; the factor on that side must
  // still be an isometry, so those rows (ranked last) are completed by Gram-Schmidt of unit
  // vectors e_0, e_1, ... against every row already set (two passes), taking the first
  // candidate whose residual keeps more than half the mean residual (L - d) / L, which the
  // best unit vector always reaches.  Row r of W, element j, lives at
  // Vh[r, j] (rows <= cols) or U[j, r] (rows > cols).
  __syncthreads();
  int live = 0;
  for (int i = 0; i < k; ++i) live += sig[i] > null_sigma;
  if (live == k || !complete) return;
  const int64_t wrow = cols ? 1 : L, wcol = cols ? k : 1;
  double* Wb = cols ? U + uo : Vh + vo;
  double* coef = sig;  // sigma is not needed past this point
  __shared__ double part[16];
  __syncthreads();
  for (int d = live; d < k; ++d) {
    double* vd = Wb + d * wrow;
    for (int e = 0; e < L; ++e) {
      for (int j = threadIdx.x; j < L; j += blockDim.x) vd[j * wcol] = j == e ? 1.0 : 0.0;
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass) {
        for (int p = threadIdx.x; p < d; p += blockDim.x) {
          const double* wp = Wb + p * wrow;
          double acc = 0.0;
          for (int j = 0; j < L; ++j) acc = fma(wp[j * wcol], vd[j * wcol], acc);
          coef[p] = acc;
        }
        __syncthreads();
        for (int j = threadIdx.x; j < L; j += blockDim.x) {
          double v = vd[j * wcol];
          for (int p = 0; p < d; ++p) v = fma(-coef[p], Wb[p * wrow + j * wcol], v);
          vd[j * wcol] = v;
        }
        __syncthreads();
      }
      double acc = 0.0;
      for (int j = threadIdx.x; j < L; j += blockDim.x) acc = fma(vd[j * wcol], vd[j * wcol], acc);
      acc = wave_sum(acc);
      if (lane == 0) part[wave] = acc;
      __syncthreads();
      double nrm2 = 0.0;
      for (int w2 = 0; w2 < nw; ++w2) nrm2 += part[w2];
      __syncthreads();
      if (nrm2 > 0.5 * static_cast<double>(L - d) / static_cast<double>(L)) {
        const double inv = 1.0 / sqrt(nrm2);
        for (int j = threadIdx.x; j < L; j += blockDim.x) vd[j * wcol] *= inv;
        __syncthreads();
        break;
      }
    }
  }
}

// ---- LQ of sector matrices (Householder) ----------------------------------------------
// One workgroup per matrix A (m x n, row-major, overwritten), k = min(m, n).  Step j forms
// the reflector H_j = I - tau_j u_j u_j^T that maps row j's tail A[j][j:n) onto beta_j e_j
// (u_j[j] = 1, the rest of u_j stored over A[j][j+1:n); beta_j = -sign(alpha) |x|, the
// convention of LAPACK's dgeqrf on A^T), and applies it to the rows below.  Then
// A = L Q with L[i][j] = A[i][j] (i > j), L[j][j] = beta_j, and Q = [I_k 0] H_{k-1} ... H_0,
// formed by applying the reflectors to the unit rows in reverse order.  Rows are handled
// by 16-lane groups (DPP row sums); the current reflector and the taus/betas sit in LDS.
__global__ __launch_bounds__(256) void lq_kernel(const int64_t* __restrict__ probs, double* __restrict__ A,
                                                 double* __restrict__ Lo, double* __restrict__ Qo,
                                                 int64_t lds_doubles) {
  extern __shared__ double lds[];
  __shared__ double part[4];
  __shared__ double coef_s[2];
  const int64_t* pw = probs + TN_LQ_PROB_WORDS * blockIdx.x;
  const int m = static_cast<int>(pw[0]), n = static_cast<int>(pw[1]);
  double* a = A + pw[2];
  double* L = Lo + pw[3];
  double* Q = Qo + pw[4];
  const int k = m < n ? m : n;
  // a problem whose reflector and tau/beta arrays exceed the launch's LDS is left untouched
  // (the caller sizes the launch for its largest problem)
  if (static_cast<int64_t>(n) + 2 * static_cast<int64_t>(k) > lds_doubles) return;
  double* v = lds;          // n: the reflector being applied
  double* tau = lds + n;    // k
  double* beta = tau + k;   // k
  const int gl = threadIdx.x % P, grp = threadIdx.x / P, ngrp = blockDim.x / P;
  for (int j = 0; j < k; ++j) {
    double* rj = a + static_cast<int64_t>(j) * n;
    double acc = 0.0;
    for (int l = j + 1 + threadIdx.x; l < n; l += blockDim.x) acc = fma(rj[l], rj[l], acc);
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double xn2 = (part[0] + part[1]) + (part[2] + part[3]);
      const double alpha = rj[j];
      if (xn2 == 0.0) {  // nothing to annihilate: H_j = I
        tau[j] = 0.0;
        beta[j] = alpha;
        coef_s[0] = 0.0;
      } else {
        const double b = -copysign(sqrt(fma(alpha, alpha, xn2)), alpha);
        tau[j] = (b - alpha) / b;
        beta[j] = b;
        coef_s[0] = 1.0 / (alpha - b);
      }
    }
    __syncthreads();
    const double scale = coef_s[0], tj = tau[j];
    for (int l = j + 1 + threadIdx.x; l < n; l += blockDim.x) {
      const double u = rj[l] * scale;
      v[l] = u;
      rj[l] = u;
    }
    if (threadIdx.x == 0) v[j] = 1.0;
    __syncthreads();
    if (tj != 0.0) {
      for (int i = j + 1 + grp; i < m; i += ngrp) {
        double* ri = a + static_cast<int64_t>(i) * n;
        double s = 0.0;
        for (int l = j + gl; l < n; l += P) s = fma(ri[l], v[l], s);
        const double ts = tj * row_sum(s);
        for (int l = j + gl; l < n; l += P) ri[l] = fma(-ts, v[l], ri[l]);
      }
    }
    __syncthreads();
  }
  // L (m x k): the columns below the diagonal, beta on it, zeros above
  for (int e = threadIdx.x; e < m * k; e += blockDim.x) {
    const int i = e / k, c = e - (e / k) * k;
    L[e] = c < i ? a[static_cast<int64_t>(i) * n + c] : (c == i ? beta[c] : 0.0);
  }
  // Q (k x n) = [I_k 0] H_{k-1} ... H_0: unit rows, then each reflector applied on the right
  for (int e = threadIdx.x; e < k * n; e += blockDim.x) {
    const int i = e / n, c = e - (e / n) * n;
    Q[e] = c == i ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int j = k - 1; j >= 0; --j) {
    const double tj = tau[j];
    if (tj == 0.0) continue;  // uniform: tau sits in LDS
    const double* rj = a + static_cast<int64_t>(j) * n;
    for (int l = j + 1 + threadIdx.x; l < n; l += blockDim.x) v[l] = rj[l];
    if (threadIdx.x == 0) v[j] = 1.0;
    __syncthreads();
    for (int i = grp; i < k; i += ngrp) {
      double* qi = Q + static_cast<int64_t>(i) * n;
      double s = 0.0;
      for (int l = j + gl; l < n; l += P) s = fma(qi[l], v[l], s);
      const double ts = tj * row_sum(s);
      for (int l = j + gl; l < n; l += P) qi[l] = fma(-ts, v[l], qi[l]);
    }
    __syncthreads();
  }
}

// Dynamic LDS above the default needs the per-function opt-in; the request must stay
// within the device's per-workgroup limit (less what the kernel declares statically).
int allow_lds(const void* fn, size_t bytes, const char* who) {
  int dev = 0, limit = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
    return tn::fail(std::string(who) + ": cannot query the LDS limit");
  if (bytes + 256 > static_cast<size_t>(limit))
    return tn::fail(std::string(who) + ": " + std::to_string(bytes) + " bytes of LDS requested, device allows " +
                    std::to_string(limit));
  if (bytes > 65536 &&
      hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)) != hipSuccess)
    return tn::fail(std::string(who) + ": cannot raise the LDS limit to " + std::to_string(bytes));
  return 0;
}

}  // namespace

extern "C" int tn_svd_jacobi_converged(const int32_t* rotations, int64_t n, int32_t* stop, void* stream) {
  if (n <= 0) return 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  jacobi_converged<<<1, 256, 0, s>>>(rotations, n, stop);
  return tn::check_launch("jacobi_converged");
}

extern "C" int tn_svd_jacobi_norms(const int64_t* probs, int64_t nprobs, const double* x, double* fro2,
                                   void* stream) {
  if (nprobs <= 0) return 0;
  if (nprobs > 0x7fffffff) return tn::fail("tn_svd_jacobi_norms: too many problems");
  hipStream_t s = static_cast<hipStream_t>(stream);
  jacobi_norms<<<dim3(static_cast<unsigned>(nprobs)), 256, 0, s>>>(probs, x, fro2);
  return tn::check_launch("jacobi_norms");
}

extern "C" int tn_svd_jacobi_round(const int64_t* probs, const int32_t* wgs, int64_t nwgs,
                                   int64_t span, double* x, double* q, int32_t* rotations, int w,
                                   int round, int first, double tol, const int32_t* stop,
                                   const double* fro2, double null_scale, void* stream) {
  if (nwgs <= 0) return 0;
  if (w < 1 || w > TN_SVD_MAX_W) return tn::fail("tn_svd_jacobi_round: block width out of range");
  if (nwgs > 0x7fffffff) return tn::fail("tn_svd_jacobi_round: too many workgroups");
  // LDS per workgroup: 2w span doubles, at least the 2w (L + k) of a tournament problem
  // and the (k + k % 2) (L + k) of a problem held by one block pair
  const size_t bytes = static_cast<size_t>(2 * w) * static_cast<size_t>(span) * sizeof(double);
  if (span <= 0) return tn::fail("tn_svd_jacobi_round: empty launch span");
  if (fro2 == nullptr) return tn::fail("tn_svd_jacobi_round: no problem norms");
  hipStream_t s = static_cast<hipStream_t>(stream);
  // lanes per pair: more DPP rows per pair when the launch's vectors are long (span LDS
  // doubles per vector), as many as keep at least TN_SVD_LANE_ELEMS elements per lane and
  // one workgroup of at most 1024 threads
  int pl = P;
  while (pl < TN_SVD_LANES && 2 * pl * w <= 1024 && span >= 2 * pl * TN_SVD_LANE_ELEMS) pl *= 2;
  const int threads = std::max(64, std::min(1024, pl * w));  // one pl-lane group per pair
  const void* fn = pl == 64   ? reinterpret_cast<const void*>(jacobi_round<64>)
                   : pl == 32 ? reinterpret_cast<const void*>(jacobi_round<32>)
                              : reinterpret_cast<const void*>(jacobi_round<P>);
  if (int rc = allow_lds(fn, bytes, "tn_svd_jacobi_round")) return rc;
  const dim3 grid(static_cast<unsigned>(nwgs));
  if (pl == 64)
    hipLaunchKernelGGL(jacobi_round<64>, grid, dim3(threads), bytes, s, probs, wgs, x, q, rotations, w, round,
                       first, tol, stop, fro2, null_scale);
  else if (pl == 32)
    hipLaunchKernelGGL(jacobi_round<32>, grid, dim3(threads), bytes, s, probs, wgs, x, q, rotations, w, round,
                       first, tol, stop, fro2, null_scale);
  else
    hipLaunchKernelGGL(jacobi_round<P>, grid, dim3(threads), bytes, s, probs, wgs, x, q, rotations, w, round,
                       first, tol, stop, fro2, null_scale);
  return tn::check_launch("jacobi_round");
}

extern "C" int tn_svd_jacobi_finish(const int64_t* probs, int64_t nprobs, const double* x,
                                    const double* q, double* U, double* S, double* Vh,
                                    const double* fro2, double null_scale, int complete,
                                    void* stream) {
  if (nprobs <= 0) return 0;
  if (fro2 == nullptr) return tn::fail("tn_svd_jacobi_finish: no problem norms");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t bytes = TN_SVD_MAX_K * (sizeof(double) + sizeof(int));
  if (int rc = allow_lds(reinterpret_cast<const void*>(jacobi_finish), bytes, "tn_svd_jacobi_finish"))
    return rc;
  hipLaunchKernelGGL(jacobi_finish, dim3(static_cast<unsigned>(nprobs)), dim3(1024), bytes, s, probs,
                     x, q, U, S, Vh, fro2, null_scale, complete);
  return tn::check_launch("jacobi_finish");
}

extern "C" int tn_lq(const int64_t* probs, int64_t nprobs, int64_t max_lds_doubles, double* A, double* L,
                     double* Q, void* stream) {
  if (nprobs <= 0) return 0;
  if (nprobs > 0x7fffffff) return tn::fail("tn_lq: too many problems");
  if (max_lds_doubles <= 0) return tn::fail("tn_lq: empty LDS request");
  const size_t bytes = static_cast<size_t>(max_lds_doubles) * sizeof(double);
  if (int rc = allow_lds(reinterpret_cast<const void*>(lq_kernel), bytes, "tn_lq")) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(lq_kernel, dim3(static_cast<unsigned>(nprobs)), dim3(256), bytes, s, probs, A, L, Q,
                     max_lds_doubles);
  return tn::check_launch("tn_lq");
}

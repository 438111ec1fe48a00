// The small dense step of the Lanczos solver, on the device: the ground eigenpair of the
// symmetric tridiagonal matrix that the recurrence leaves in device memory.
//
// tn_tridiag_ground reads alpha[0..n) (diagonal) and beta[0..n-1) (beta[j] couples j and
// j+1), truncates at the first beta below `tol` (the happy breakdown of the recurrence:
// the Krylov space is invariant there, so the leading block holds the answer), and writes
// the lowest eigenvalue and its unit eigenvector (zeros past the truncation, first nonzero
// component positive).  The eigenvalue comes from the Sturm count (Givens' method):
// 64-way multisection across one wavefront on the device, plain bisection in the host twin
// (tn_tridiag_ground_host, which is how the CPU test suite checks the routine against
// LAPACK without a GPU), both converged to the spacing of doubles; the eigenvector from
// three steps of inverse iteration on T - lambda I with a pivoted tridiagonal LU (the
// scheme of LAPACK's dgttrf/dgtts2) on one lane, with the matrix in LDS.
#include <math.h>

#include "tn_common.hpp"

namespace {

constexpr int kMax = TN_TRIDIAG_MAX;

struct TridiagWork {
  double b2[kMax], dl[kMax], dd[kMax], du[kMax], du2[kMax];
  int piv[kMax];
};

__host__ __device__ inline int sturm_below(const double* a, const double* b2, int m, double x,
                                           double pivmin) {
  // number of eigenvalues of T[:m,:m] below x
  int cnt = 0;
  double d = 1.0;
  for (int i = 0; i < m; ++i) {
    d = a[i] - x - (i ? b2[i - 1] / d : 0.0);
    if (fabs(d) < pivmin) d = -pivmin;  // x hit an eigenvalue of a leading block: count it below
    if (d < 0.0) ++cnt;
  }
  return cnt;
}

struct TridiagSetup {
  int m;                 // leading block kept (happy breakdown at the first |beta| < tol)
  double lo, hi;         // Gershgorin interval, widened a little
  double pivmin, tnorm;  // smallest pivot allowed; scale of the matrix
};

// a[0..n), b[0..n-1): the matrix.  Fills w.b2 and returns the truncation and interval.
__host__ __device__ inline TridiagSetup tridiag_setup(const double* a, const double* b, int n, double tol,
                                                      TridiagWork& w) {
  TridiagSetup st{};
  int m = n;
  for (int j = 0; j + 1 < n; ++j)
    if (fabs(b[j]) < tol) {
      m = j + 1;
      break;
    }
  for (int i = 0; i + 1 < m; ++i) w.b2[i] = b[i] * b[i];
  double lo = INFINITY, hi = -INFINITY;
  for (int i = 0; i < m; ++i) {
    const double r = (i ? fabs(b[i - 1]) : 0.0) + (i + 1 < m ? fabs(b[i]) : 0.0);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
  }
  const double span = fmax(hi - lo, 1e-300);
  st.m = m;
  st.tnorm = fmax(fabs(lo), fabs(hi));
  // a zero pivot (lambda exactly an eigenvalue of a leading block) is replaced by pivmin
  st.pivmin = fmax(1e-290, 2.2e-16 * st.tnorm);
  st.lo = lo - 1e-15 * span;
  st.hi = hi + 1e-15 * span;
  return st;
}

// The bracket is resolved: the relative resolution of doubles, with a floor far below the
// eps * |T| accuracy that any eigenvalue of T carries.
__host__ __device__ inline bool bracket_done(double lo, double hi, double tnorm) {
  return hi - lo <= fmax(2.2e-16 * fmax(fabs(lo), fabs(hi)), 2.2e-19 * tnorm);
}

// Unit eigenvector of T[:m,:m] at eigenvalue lam: three steps of inverse iteration on
// T - lam I with the pivoted tridiagonal LU; first nonzero entry made positive.
__host__ __device__ inline void tridiag_vector(const double* a, const double* b, int m, double lam,
                                               double pivmin, double* x, TridiagWork& w) {
  double *dl = w.dl, *dd = w.dd, *du = w.du, *du2 = w.du2;
  int* piv = w.piv;
  for (int i = 0; i < m; ++i) {
    dd[i] = a[i] - lam;
    du[i] = i + 1 < m ? b[i] : 0.0;
    dl[i] = i + 1 < m ? b[i] : 0.0;
    du2[i] = 0.0;
    piv[i] = i;
  }
  for (int i = 0; i + 1 < m; ++i) {
    if (fabs(dd[i]) >= fabs(dl[i])) {
      if (fabs(dd[i]) < pivmin) dd[i] = pivmin;
      const double f = dl[i] / dd[i];
      dl[i] = f;
      dd[i + 1] -= f * du[i];
    } else {
      const double f = dd[i] / dl[i];
      dd[i] = dl[i];
      dl[i] = f;
      const double tmp = du[i];
      du[i] = dd[i + 1];
      dd[i + 1] = tmp - f * dd[i + 1];
      if (i + 2 < m) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
      piv[i] = i + 1;
    }
  }
  if (fabs(dd[m - 1]) < pivmin) dd[m - 1] = pivmin;
  for (int i = 0; i < m; ++i) x[i] = 1.0;
  for (int pass = 0; pass < 3; ++pass) {
    for (int i = 0; i + 1 < m; ++i) {  // L y = x
      if (piv[i] == i) {
        x[i + 1] -= dl[i] * x[i];
      } else {
        const double tmp = x[i];
        x[i] = x[i + 1];
        x[i + 1] = tmp - dl[i] * x[i];
      }
    }
    x[m - 1] /= dd[m - 1];  // U x = y
    if (m > 1) x[m - 2] = (x[m - 2] - du[m - 2] * x[m - 1]) / dd[m - 2];
    for (int i = m - 3; i >= 0; --i) x[i] = (x[i] - du[i] * x[i + 1] - du2[i] * x[i + 2]) / dd[i];
    double s = 0.0;
    for (int i = 0; i < m; ++i) s = fmax(s, fabs(x[i]));
    double nn = 0.0;
    for (int i = 0; i < m; ++i) {
      x[i] /= s;
      nn += x[i] * x[i];
    }
    nn = 1.0 / sqrt(nn);
    for (int i = 0; i < m; ++i) x[i] *= nn;
  }
  int first = 0;
  while (first + 1 < m && x[first] == 0.0) ++first;
  if (x[first] < 0.0)
    for (int i = 0; i < m; ++i) x[i] = -x[i];
}

// The whole routine on one thread (host twin): bisection on the Sturm count.
inline int tridiag_ground_serial(const double* a, const double* b, int n, double tol, double* value,
                                 double* x, TridiagWork& w) {
  const TridiagSetup st = tridiag_setup(a, b, n, tol, w);
  double lo = st.lo, hi = st.hi;
  for (int it = 0; it < 2200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi || bracket_done(lo, hi, st.tnorm)) break;
    if (sturm_below(a, w.b2, st.m, mid, st.pivmin) >= 1)
      hi = mid;
    else
      lo = mid;
  }
  const double lam = 0.5 * (lo + hi);
  tridiag_vector(a, b, st.m, lam, st.pivmin, x, w);
  *value = lam;
  return st.m;
}

// Device: 64-way multisection -- every lane evaluates the Sturm count at one of 64 interior
// points of the bracket, a ballot finds the first point with an eigenvalue below it, and the
// bracket shrinks 65-fold per round (about nine rounds to the resolution of doubles).  The
// vector then comes from lane 0.
__device__ void tridiag_ground_device(const double* __restrict__ alpha, const double* __restrict__ beta, int n,
                                      double tol, double* __restrict__ value, double* __restrict__ coef) {
  __shared__ double a[kMax], b[kMax], x[kMax];
  __shared__ TridiagWork w;
  __shared__ TridiagSetup st;
  const int t = threadIdx.x;
  for (int i = t; i < n; i += 64) {
    a[i] = alpha[i];
    b[i] = i + 1 < n ? beta[i] : 0.0;
  }
  __syncthreads();
  if (t == 0) st = tridiag_setup(a, b, n, tol, w);
  __syncthreads();
  const int m = st.m;
  double lo = st.lo, hi = st.hi;  // identical on every lane: every update below is uniform
  for (int round = 0; round < 64; ++round) {
    if (bracket_done(lo, hi, st.tnorm)) break;
    const double step = (hi - lo) / 65.0;
    if (lo + step <= lo) break;  // the bracket spans a few doubles
    const bool below = sturm_below(a, w.b2, m, lo + (t + 1) * step, st.pivmin) >= 1;
    const unsigned long long mask = __ballot(below);
    const int k = mask ? __ffsll(mask) - 1 : 64;  // first of the 64 points with a count >= 1
    const double nlo = k > 0 ? lo + k * step : lo;
    const double nhi = k < 64 ? lo + (k + 1) * step : hi;
    lo = nlo;
    hi = nhi;
  }
  if (t == 0) {
    const double lam = 0.5 * (lo + hi);
    tridiag_vector(a, b, m, lam, st.pivmin, x, w);
    value[0] = lam;
  }
  __syncthreads();
  for (int i = t; i < n; i += 64) coef[i] = i < m ? x[i] : 0.0;
}

__global__ __launch_bounds__(64) void tridiag_ground_kernel(const double* __restrict__ alpha,
                                                            const double* __restrict__ beta, int n,
                                                            double tol, double* __restrict__ value,
                                                            double* __restrict__ coef) {
  tridiag_ground_device(alpha, beta, n, tol, value, coef);
}

// One workgroup per problem: problem p's diagonal at alpha + p ld, off-diagonal at
// beta + p ld, its vector written to coef + p ld and its value to value[p].
__global__ __launch_bounds__(64) void tridiag_batched_kernel(const double* __restrict__ alpha,
                                                             const double* __restrict__ beta, int n,
                                                             int64_t ld, double tol,
                                                             double* __restrict__ value,
                                                             double* __restrict__ coef) {
  const int64_t p = blockIdx.x;
  tridiag_ground_device(alpha + p * ld, beta + p * ld, n, tol, value + p, coef + p * ld);
}

}  // namespace

extern "C" int tn_tridiag_ground(const double* alpha, const double* beta, int n, double tol,
                                 double* value, double* coef, void* stream) {
  if (n < 1 || n > kMax) return tn::fail("tn_tridiag_ground: n must lie in [1, TN_TRIDIAG_MAX]");
  hipStream_t s = static_cast<hipStream_t>(stream);
  tridiag_ground_kernel<<<1, 64, 0, s>>>(alpha, beta, n, tol, value, coef);
  return tn::check_launch("tn_tridiag_ground");
}

extern "C" int tn_tridiag_ground_host(const double* alpha, const double* beta, int n, double tol,
                                      double* value, double* coef) {
  if (n < 1 || n > kMax) return tn::fail("tn_tridiag_ground_host: n must lie in [1, TN_TRIDIAG_MAX]");
  double b[kMax], x[kMax];
  TridiagWork w;
  for (int i = 0; i < n; ++i) b[i] = i + 1 < n ? beta[i] : 0.0;
  const int m = tridiag_ground_serial(alpha, b, n, tol, value, x, w);
  for (int i = 0; i < n; ++i) coef[i] = i < m ? x[i] : 0.0;
  return 0;
}

extern "C" int tn_tridiag_ground_batched(const double* alpha, const double* beta, int n, int64_t ld,
                                         int64_t nprob, double tol, double* values, double* coefs,
                                         void* stream) {
  if (nprob <= 0) return 0;
  if (n < 1 || n > kMax) return tn::fail("tn_tridiag_ground_batched: n must lie in [1, TN_TRIDIAG_MAX]");
  if (ld < n) return tn::fail("tn_tridiag_ground_batched: leading dimension below n");
  if (nprob > 0x7fffffffLL) return tn::fail("tn_tridiag_ground_batched: too many problems");
  hipStream_t s = static_cast<hipStream_t>(stream);
  tridiag_batched_kernel<<<static_cast<unsigned>(nprob), 64, 0, s>>>(alpha, beta, n, ld, tol, values, coefs);
  return tn::check_launch("tn_tridiag_ground_batched");
}

// Vector-space kernels for the Krylov solver, precision staging, the fused two-site MPO
// application, and the C-ABI error state.
//
// tn_dot is a deterministic two-pass reduction in FP64; scalars that the Lanczos
// recurrence needs stay in device memory so a matvec-orthogonalise step needs no host
// round trip.
#include <string>

#include "tn_common.hpp"

namespace tn {
namespace {
thread_local std::string g_error;
}
void set_error(const std::string& msg) { g_error = msg; }
}  // namespace tn

extern "C" const char* tn_last_error(void) { return tn::g_error.c_str(); }
extern "C" int tn_version(void) { return 1; }

extern "C" int tn_device_cus(int* out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return tn::fail("tn_device_cus: no HIP device");
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return tn::fail("tn_device_cus: hipGetDeviceProperties failed");
  *out = prop.multiProcessorCount;
  return 0;
}

namespace {

constexpr int kThreads = 256;
constexpr int kMaxPartials = 1024;

template <typename T>
__global__ __launch_bounds__(kThreads) void dot_partial(const T* __restrict__ x,
                                                        const T* __restrict__ y,
                                                        const double* __restrict__ w, int64_t n,
                                                        double* __restrict__ part) {
  __shared__ double red[kThreads];
  double acc = 0.0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride) {
    double v = static_cast<double>(x[i]) * static_cast<double>(y[i]);
    if (w) v *= w[i];
    acc += v;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kThreads) void dot_final(const double* __restrict__ part, int np,
                                                      double* __restrict__ out) {
  __shared__ double red[kThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += kThreads) acc += part[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__device__ __forceinline__ double scalar(double h, const double* d, int inv) {
  if (!d) return h;
  const double v = *d;
  return h * (inv ? 1.0 / v : v);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void axpby_kernel(int64_t n, double a, const double* ad,
                                                         int ainv, const T* __restrict__ x,
                                                         double b, const double* bd, int binv,
                                                         T* __restrict__ y) {
  const double alpha = scalar(a, ad, ainv);
  const double beta = scalar(b, bd, binv);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride) {
    // a zero factor contributes nothing and its operand is not read: a Lanczos basis
    // vector past a happy breakdown (non-finite after the division by beta ~ 0) enters the
    // Ritz recombination with coefficient exactly 0
    const double yv = beta == 0.0 ? 0.0 : beta * static_cast<double>(y[i]);
    const double xv = alpha == 0.0 ? 0.0 : alpha * static_cast<double>(x[i]);
    y[i] = static_cast<T>(xv + yv);
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void scale_kernel(int64_t n, double s, const double* sd,
                                                         int inv, T* __restrict__ x) {
  const double f = scalar(s, sd, inv);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride)
    x[i] = static_cast<T>(f * static_cast<double>(x[i]));
}

template <typename S, typename D>
__global__ __launch_bounds__(kThreads) void cast_kernel(const S* __restrict__ src, D* __restrict__ dst,
                                                        int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride)
    dst[i] = static_cast<D>(src[i]);
}

unsigned grid_for(int64_t n) {
  int64_t g = (n + kThreads * 4 - 1) / (kThreads * 4);
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return static_cast<unsigned>(g);
}

// ---- fused two-site MPO application --------------------------------------------------
// One workgroup owns 1024 positions (r, s) of one slab-pair job.  The job's input count
// nin picks, once per workgroup, a body compiled for exactly that count: every index into
// the per-thread input array is a compile-time constant (the array lives in VGPRs) and a
// position costs nin x nout multiply-adds -- the coefficient rows are zero-padded to
// kMpoMax columns, but only their first nin entries are read.  Coefficients and slab
// tables are uniform across the wave and come in through scalar loads.
constexpr int kMpoMax = 32;   // max input slabs of one job (coefficient row length)
constexpr int kMpoPer = 4;    // positions per thread

template <typename T, int NIN>
__device__ __forceinline__ void mpo_body(const T* __restrict__ in, T* __restrict__ out,
                                         const int64_t* __restrict__ is, const int64_t* __restrict__ os,
                                         const T* __restrict__ cf, int nout, int64_t base, int64_t n,
                                         uint32_t S32) {
  // PER positions at a time: all PER x NIN loads are issued before the first multiply, so a
  // thread keeps that many loads in flight (a position past the job's end reads position
  // 0, in bounds, and stores nothing)
  constexpr int PER = NIN <= 8 ? kMpoPer : (NIN <= 16 ? 2 : 1);
#pragma unroll
  for (int j0 = 0; j0 < kMpoPer; j0 += PER) {
    int64_t pos[PER];
    bool live[PER];
    uint32_t rr[PER];
    int64_t ss[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int64_t e = base + (j0 + u) * kThreads + threadIdx.x;
      live[u] = e < n;
      pos[u] = live[u] ? e : 0;
      rr[u] = static_cast<uint32_t>(pos[u]) / S32;
      ss[u] = static_cast<int64_t>(static_cast<uint32_t>(pos[u]) - rr[u] * S32);
    }
    T v[PER][NIN > 0 ? NIN : 1];
#pragma unroll
    for (int k = 0; k < NIN; ++k) {
      const int64_t o = is[2 * k], st = is[2 * k + 1];
#pragma unroll
      for (int u = 0; u < PER; ++u) v[u][k] = in[o + rr[u] * st + ss[u]];
    }
    for (int o = 0; o < nout; ++o) {
      const T* row = cf + o * kMpoMax;
      const int64_t ob = os[2 * o], ost = os[2 * o + 1];
      T acc[PER];
#pragma unroll
      for (int u = 0; u < PER; ++u) acc[u] = T(0);
#pragma unroll
      for (int k = 0; k < NIN; ++k) {
        const T c = row[k];
#pragma unroll
        for (int u = 0; u < PER; ++u) acc[u] += c * v[u][k];
      }
#pragma unroll
      for (int u = 0; u < PER; ++u)
        if (live[u]) out[ob + rr[u] * ost + ss[u]] = acc[u];
    }
  }
}

template <typename T, int N>
__device__ __forceinline__ void mpo_dispatch(int nin, const T* __restrict__ in, T* __restrict__ out,
                                             const int64_t* __restrict__ is, const int64_t* __restrict__ os,
                                             const T* __restrict__ cf, int nout, int64_t base, int64_t n,
                                             uint32_t S32) {
  if constexpr (N <= kMpoMax) {
    if (nin == N)
      mpo_body<T, N>(in, out, is, os, cf, nout, base, n, S32);
    else
      mpo_dispatch<T, N + 1>(nin, in, out, is, os, cf, nout, base, n, S32);
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void mpo_kernel(const T* __restrict__ in, T* __restrict__ out,
                                                       const int64_t* __restrict__ jobs,
                                                       const int64_t* __restrict__ in_slabs,
                                                       const int64_t* __restrict__ out_slabs,
                                                       const T* __restrict__ coef,
                                                       const int64_t* __restrict__ tiles) {
  const int64_t job = tiles[2 * blockIdx.x];
  const int64_t base = tiles[2 * blockIdx.x + 1] * (kThreads * kMpoPer);
  const int64_t* jw = jobs + job * TN_MPO_JOB_WORDS;
  const int64_t R = jw[0], S = jw[1];
  const int nin = static_cast<int>(jw[3] - jw[2]);
  const int nout = static_cast<int>(jw[5] - jw[4]);
  if (nin < 0 || nin > kMpoMax) return;  // the host refuses such tables; never index past v[]
  mpo_dispatch<T, 0>(nin, in, out, in_slabs + 2 * jw[2], out_slabs + 2 * jw[4], coef + jw[6], nout,
                     base, R * S, static_cast<uint32_t>(S));
}

// ---- Gram-Schmidt pass of a Krylov vector against the basis ---------------------------
// One classical Gram-Schmidt pass over k <= kOrthMax basis vectors: c_i = <v_i, w>_W for all
// i in one sweep over the elements, then w <- w - sum_i c_i v_i in a second, which also
// accumulates |w|_W^2 of the result.  Partials per workgroup, reduced by one workgroup in a
// fixed order (deterministic, like tn_dot).
constexpr int kOrthMax = 8;

struct OrthBasis {
  const void* v[kOrthMax];
};

// deterministic tree sum of one value per thread over the workgroup
__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

template <typename T, int K>
__global__ __launch_bounds__(kThreads) void multidot_partial(OrthBasis b, const T* __restrict__ w,
                                                             const double* __restrict__ wt, int64_t n,
                                                             double* __restrict__ part) {
  __shared__ double red[kThreads];
  double acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride) {
    double x = static_cast<double>(w[i]);
    if (wt) x *= wt[i];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] += static_cast<double>(static_cast<const T*>(b.v[k])[i]) * x;
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double t = block_sum(acc[k], red);
    if (threadIdx.x == 0) part[static_cast<int64_t>(blockIdx.x) * K + k] = t;
  }
}

// coef[k] = sum over workgroups of part[g * K + k]; last (if given) = coef[K - 1]
template <int K>
__global__ __launch_bounds__(kThreads) void multidot_final(const double* __restrict__ part, int np,
                                                           double* __restrict__ coef, double* __restrict__ last) {
  __shared__ double red[kThreads];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double acc = 0.0;
    for (int g = threadIdx.x; g < np; g += kThreads) acc += part[static_cast<int64_t>(g) * K + k];
    const double t = block_sum(acc, red);
    if (threadIdx.x == 0) {
      coef[k] = t;
      if (last != nullptr && k == K - 1) *last = t;
    }
  }
}

template <typename T, int K>
__global__ __launch_bounds__(kThreads) void cgs_update(OrthBasis b, T* __restrict__ w, const double* __restrict__ coef,
                                                       const double* __restrict__ wt, int64_t n,
                                                       double* __restrict__ part) {
  __shared__ double red[kThreads];
  double c[K];
#pragma unroll
  for (int k = 0; k < K; ++k) c[k] = coef[k];
  double acc = 0.0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n; i += stride) {
    double x = static_cast<double>(w[i]);
#pragma unroll
    for (int k = 0; k < K; ++k) x -= c[k] * static_cast<double>(static_cast<const T*>(b.v[k])[i]);
    const T xs = static_cast<T>(x);
    w[i] = xs;
    double q = static_cast<double>(xs) * static_cast<double>(xs);
    if (wt) q *= wt[i];
    acc += q;
  }
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(kThreads) void norm_final(const double* __restrict__ part, int np,
                                                       double* __restrict__ out) {
  __shared__ double red[kThreads];
  double acc = 0.0;
  for (int g = threadIdx.x; g < np; g += kThreads) acc += part[g];
  const double t = block_sum(acc, red);
  if (threadIdx.x == 0) out[0] = sqrt(t);
}

template <typename T, int K>
int orth_passes(OrthBasis b, T* w, const double* wt, int64_t n, int passes, double* coef, double* last,
                double* norm, double* work, hipStream_t s) {
  int64_t g = (n + kThreads * 8 - 1) / (kThreads * 8);
  if (g < 1) g = 1;
  if (g > kMaxPartials) g = kMaxPartials;
  const unsigned gu = static_cast<unsigned>(g);
  double* part = work;                                   // kOrthMax * kMaxPartials
  double* coef2 = work + kOrthMax * kMaxPartials;        // kOrthMax
  for (int p = 0; p < passes; ++p) {
    double* c = p == 0 ? coef : coef2;
    multidot_partial<T, K><<<gu, kThreads, 0, s>>>(b, w, wt, n, part);
    multidot_final<K><<<1, kThreads, 0, s>>>(part, static_cast<int>(g), c, p == 0 ? last : nullptr);
    cgs_update<T, K><<<gu, kThreads, 0, s>>>(b, w, c, wt, n, part);
  }
  norm_final<<<1, kThreads, 0, s>>>(part, static_cast<int>(g), norm);
  return 0;
}

template <typename T, int K = 1>
int orth_dispatch(int k, OrthBasis b, T* w, const double* wt, int64_t n, int passes, double* coef,
                  double* last, double* norm, double* work, hipStream_t s) {
  if constexpr (K > kOrthMax) {
    return tn::fail("tn_krylov_orth: too many basis vectors");
  } else {
    if (k == K) return orth_passes<T, K>(b, w, wt, n, passes, coef, last, norm, work, s);
    return orth_dispatch<T, K + 1>(k, b, w, wt, n, passes, coef, last, norm, work, s);
  }
}

}  // namespace

extern "C" int tn_dot(int dtype, const void* x, const void* y, const double* w, int64_t n,
                      double* work, double* out, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  int64_t g = (n + kThreads * 8 - 1) / (kThreads * 8);
  if (g < 1) g = 1;
  if (g > kMaxPartials) g = kMaxPartials;
  if (dtype == TN_F64)
    dot_partial<double><<<static_cast<unsigned>(g), kThreads, 0, s>>>(
        static_cast<const double*>(x), static_cast<const double*>(y), w, n, work);
  else if (dtype == TN_F32)
    dot_partial<float><<<static_cast<unsigned>(g), kThreads, 0, s>>>(
        static_cast<const float*>(x), static_cast<const float*>(y), w, n, work);
  else
    return tn::fail("tn_dot: unknown dtype");
  dot_final<<<1, kThreads, 0, s>>>(work, static_cast<int>(g), out);
  return tn::check_launch("tn_dot");
}

extern "C" int tn_axpby(int dtype, int64_t n, double alpha, const double* alpha_dev, int inv_alpha,
                        const void* x, double beta, const double* beta_dev, int inv_beta, void* y,
                        void* stream) {
  if (n <= 0) return 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == TN_F64)
    axpby_kernel<double><<<grid_for(n), kThreads, 0, s>>>(n, alpha, alpha_dev, inv_alpha,
                                                          static_cast<const double*>(x), beta,
                                                          beta_dev, inv_beta, static_cast<double*>(y));
  else if (dtype == TN_F32)
    axpby_kernel<float><<<grid_for(n), kThreads, 0, s>>>(n, alpha, alpha_dev, inv_alpha,
                                                         static_cast<const float*>(x), beta,
                                                         beta_dev, inv_beta, static_cast<float*>(y));
  else
    return tn::fail("tn_axpby: unknown dtype");
  return tn::check_launch("tn_axpby");
}

extern "C" int tn_scale(int dtype, int64_t n, double sh, const double* sd, int inv, void* x,
                        void* stream) {
  if (n <= 0) return 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == TN_F64)
    scale_kernel<double><<<grid_for(n), kThreads, 0, s>>>(n, sh, sd, inv, static_cast<double*>(x));
  else if (dtype == TN_F32)
    scale_kernel<float><<<grid_for(n), kThreads, 0, s>>>(n, sh, sd, inv, static_cast<float*>(x));
  else
    return tn::fail("tn_scale: unknown dtype");
  return tn::check_launch("tn_scale");
}

extern "C" int tn_cast(int src_dtype, const void* src, int dst_dtype, void* dst, int64_t n,
                       void* stream) {
  if (n <= 0) return 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (src_dtype == TN_F64 && dst_dtype == TN_F32)
    cast_kernel<double, float><<<grid_for(n), kThreads, 0, s>>>(static_cast<const double*>(src),
                                                                static_cast<float*>(dst), n);
  else if (src_dtype == TN_F32 && dst_dtype == TN_F64)
    cast_kernel<float, double><<<grid_for(n), kThreads, 0, s>>>(static_cast<const float*>(src),
                                                                static_cast<double*>(dst), n);
  else
    return tn::fail("tn_cast: unsupported dtype pair");
  return tn::check_launch("tn_cast");
}

extern "C" int tn_mpo_apply(int dtype, const void* in, void* out, const int64_t* jobs,
                            int64_t njobs, const int64_t* in_slabs, const int64_t* out_slabs,
                            const void* coef, const int64_t* tiles, int64_t ntiles,
                            void* stream) {
  (void)njobs;
  if (ntiles <= 0) return 0;
  if (ntiles > 0x7fffffffLL) return tn::fail("tn_mpo_apply: too many tiles");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned g = static_cast<unsigned>(ntiles);
  if (dtype == TN_F64)
    mpo_kernel<double><<<g, kThreads, 0, s>>>(static_cast<const double*>(in), static_cast<double*>(out),
                                              jobs, in_slabs, out_slabs,
                                              static_cast<const double*>(coef), tiles);
  else if (dtype == TN_F32)
    mpo_kernel<float><<<g, kThreads, 0, s>>>(static_cast<const float*>(in), static_cast<float*>(out),
                                             jobs, in_slabs, out_slabs,
                                             static_cast<const float*>(coef), tiles);
  else
    return tn::fail("tn_mpo_apply: unknown dtype");
  return tn::check_launch("tn_mpo_apply");
}

extern "C" int64_t tn_krylov_orth_work(void) {
  return static_cast<int64_t>(kOrthMax) * kMaxPartials + kOrthMax;
}

extern "C" int tn_krylov_orth(int dtype, int64_t n, int k, const void* const* basis, void* w,
                              const double* weights, int passes, double* coef, double* last_coef,
                              double* norm, double* work, void* stream) {
  if (k < 1 || k > kOrthMax) return tn::fail("tn_krylov_orth: k must lie in [1, 8]");
  if (passes < 1 || passes > 2) return tn::fail("tn_krylov_orth: passes must be 1 or 2");
  if (basis == nullptr || w == nullptr || coef == nullptr || norm == nullptr || work == nullptr)
    return tn::fail("tn_krylov_orth: null pointer");
  if (n <= 0) return tn::fail("tn_krylov_orth: empty vectors");
  OrthBasis b{};
  for (int i = 0; i < k; ++i) {
    if (basis[i] == nullptr) return tn::fail("tn_krylov_orth: null basis vector");
    b.v[i] = basis[i];
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  int rc;
  if (dtype == TN_F64)
    rc = orth_dispatch<double>(k, b, static_cast<double*>(w), weights, n, passes, coef, last_coef, norm, work, s);
  else if (dtype == TN_F32)
    rc = orth_dispatch<float>(k, b, static_cast<float*>(w), weights, n, passes, coef, last_coef, norm, work, s);
  else
    return tn::fail("tn_krylov_orth: unknown dtype");
  if (rc) return rc;
  return tn::check_launch("tn_krylov_orth");
}

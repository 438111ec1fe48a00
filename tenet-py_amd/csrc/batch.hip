// Krylov vector kernels for many independent problems packed into one arena.
//
// A batch of H_eff problems keeps its Lanczos vectors back to back in one buffer; problem
// s owns the element range [off_s, off_{s+1}).  The host cuts every range into chunks of
// at most kChunk elements and lists them as (segment, begin, end) triples, so one launch
// covers the whole batch with workgroups of equal work whatever the problems' sizes:
//
//   tn_dot_segmented    out[s * stride] = sum over problem s of w x y   (FP64, two passes:
//                       one partial per chunk, then each problem's partials summed in
//                       chunk order -- deterministic)
//   tn_axpby_segmented  y = (a * alpha_s) x + (b * beta_s) y per problem, the scalars read
//                       from device arrays (optionally inverted), a zero factor dropping
//                       its operand unread (as tn_axpby)
//
// and the tridiagonal ground solve of every problem's recurrence is one launch of
// tn_tridiag_ground_batched (krylov.hip).
#include "tn_common.hpp"

namespace {

constexpr int kThreads = 256;

template <typename T>
__global__ __launch_bounds__(kThreads) void segdot_partial(const T* __restrict__ x, const T* __restrict__ y,
                                                           const double* __restrict__ w,
                                                           const int64_t* __restrict__ chunks,
                                                           double* __restrict__ part) {
  __shared__ double red[kThreads];
  const int64_t* c = chunks + 3 * static_cast<int64_t>(blockIdx.x);
  const int64_t b = c[1], e = c[2];
  double acc = 0.0;
  for (int64_t i = b + threadIdx.x; i < e; i += kThreads) {
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

__global__ __launch_bounds__(kThreads) void segdot_final(const double* __restrict__ part,
                                                         const int64_t* __restrict__ seg_chunks, int64_t nseg,
                                                         double* __restrict__ out, int64_t stride) {
  const int64_t s = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (s >= nseg) return;
  double acc = 0.0;
  for (int64_t k = seg_chunks[s]; k < seg_chunks[s + 1]; ++k) acc += part[k];
  out[s * stride] = acc;
}

__device__ __forceinline__ double seg_scalar(double h, const double* d, int64_t stride, int inv, int64_t s) {
  if (!d) return h;
  const double v = d[s * stride];
  return h * (inv ? 1.0 / v : v);
}

template <typename T>
__global__ __launch_bounds__(kThreads) void segaxpby_kernel(const int64_t* __restrict__ chunks, double a,
                                                            const double* ad, int64_t as, int ainv,
                                                            const T* __restrict__ x, double b,
                                                            const double* bd, int64_t bs, int binv,
                                                            T* __restrict__ y) {
  const int64_t* c = chunks + 3 * static_cast<int64_t>(blockIdx.x);
  const int64_t s = c[0], lo = c[1], hi = c[2];
  const double alpha = seg_scalar(a, ad, as, ainv, s);
  const double beta = seg_scalar(b, bd, bs, binv, s);
  for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads) {
    const double yv = beta == 0.0 ? 0.0 : beta * static_cast<double>(y[i]);
    const double xv = alpha == 0.0 ? 0.0 : alpha * static_cast<double>(x[i]);
    y[i] = static_cast<T>(xv + yv);
  }
}

}  // namespace

extern "C" int tn_dot_segmented(int dtype, const void* x, const void* y, const double* w,
                                const int64_t* chunks, int64_t nchunks, double* partial,
                                const int64_t* seg_chunks, int64_t nseg, double* out,
                                int64_t out_stride, void* stream) {
  if (nchunks <= 0 || nseg <= 0) return 0;
  if (nchunks > 0x7fffffffLL) return tn::fail("tn_dot_segmented: too many chunks");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned g = static_cast<unsigned>(nchunks);
  if (dtype == TN_F64)
    segdot_partial<double><<<g, kThreads, 0, s>>>(static_cast<const double*>(x), static_cast<const double*>(y),
                                                  w, chunks, partial);
  else if (dtype == TN_F32)
    segdot_partial<float><<<g, kThreads, 0, s>>>(static_cast<const float*>(x), static_cast<const float*>(y),
                                                 w, chunks, partial);
  else
    return tn::fail("tn_dot_segmented: unknown dtype");
  const unsigned gf = static_cast<unsigned>((nseg + kThreads - 1) / kThreads);
  segdot_final<<<gf, kThreads, 0, s>>>(partial, seg_chunks, nseg, out, out_stride);
  return tn::check_launch("tn_dot_segmented");
}

extern "C" int tn_axpby_segmented(int dtype, const int64_t* chunks, int64_t nchunks, double a,
                                  const double* a_dev, int64_t a_stride, int a_inv, const void* x,
                                  double b, const double* b_dev, int64_t b_stride, int b_inv, void* y,
                                  void* stream) {
  if (nchunks <= 0) return 0;
  if (nchunks > 0x7fffffffLL) return tn::fail("tn_axpby_segmented: too many chunks");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned g = static_cast<unsigned>(nchunks);
  if (dtype == TN_F64)
    segaxpby_kernel<double><<<g, kThreads, 0, s>>>(chunks, a, a_dev, a_stride, a_inv,
                                                   static_cast<const double*>(x), b, b_dev, b_stride, b_inv,
                                                   static_cast<double*>(y));
  else if (dtype == TN_F32)
    segaxpby_kernel<float><<<g, kThreads, 0, s>>>(chunks, a, a_dev, a_stride, a_inv,
                                                  static_cast<const float*>(x), b, b_dev, b_stride, b_inv,
                                                  static_cast<float*>(y));
  else
    return tn::fail("tn_axpby_segmented: unknown dtype");
  return tn::check_launch("tn_axpby_segmented");
}

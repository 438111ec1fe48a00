// Permute-and-recouple gather: dst_region = sum_t coeff_t * permute(src + off_t).
//
// One kernel moves every block of a repartition / transpose / to_matrices /
// from_matrices step in a single launch, driven by a host-compiled job table (one job per
// destination region; see tenet_amd.h for the word layout).  Each workgroup owns one tile
// of one job:
//  * linear tiles walk the destination in C order (coalesced stores); used when the
//    destination-contiguous dim is also source-contiguous, or dims are tiny;
//  * LDS tiles stage a 64x64 slab through LDS so that the loads run along the
//    source-contiguous dim J and the stores along the destination-contiguous dim -- both
//    coalesced -- with the batch over all remaining dims.
// Multi-term jobs (SU(2) recoupling) accumulate all of a region's terms in registers
// before the single store, so no atomics are needed and summation order is fixed.
#include "tn_common.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kLinPerThread = 8;
constexpr int kLinTile = kThreads * kLinPerThread;
constexpr int kT = 64;  // LDS tile edge

template <typename T, typename TS>
__device__ __forceinline__ T gather_terms(const TS* __restrict__ src, const int64_t* __restrict__ tsrc,
                                          const double* __restrict__ tco, int64_t tb, int64_t te,
                                          int64_t soff) {
  T acc = T(0);
  for (int64_t t = tb; t < te; ++t) acc += static_cast<T>(tco[t]) * static_cast<T>(src[tsrc[t] + soff]);
  return acc;
}

// Linear tile: element e of the job in destination C order, PER elements per thread.  All
// of a thread's elements are decoded first and their loads issued together (independent,
// one round of memory latency per term instead of one per element).
template <typename T, typename TS, typename TD, int PER>
__device__ __forceinline__ void linear_tile(const TS* __restrict__ src, TD* __restrict__ dst,
                                            const int64_t* __restrict__ tsrc,
                                            const double* __restrict__ tco, int64_t tb, int64_t te,
                                            int rank, const int64_t* shape, const int64_t* dstr,
                                            const int64_t* sstr, int64_t dst_off, int64_t nelem,
                                            int64_t tile) {
  const int tid = threadIdx.x;
  const int64_t base = tile * (kThreads * PER);
  int64_t so[PER], dof[PER];
  bool live[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int64_t e = base + r * kThreads + tid;
    live[r] = e < nelem;
    so[r] = 0;
    dof[r] = 0;
    uint32_t rem = live[r] ? static_cast<uint32_t>(e) : 0u;
    for (int d = rank - 1; d >= 0; --d) {
      const uint32_t n = static_cast<uint32_t>(shape[d]);
      const uint32_t q = rem / n;
      const int64_t i = rem - q * n;
      rem = q;
      so[r] += i * sstr[d];
      dof[r] += i * dstr[d];
    }
  }
  T acc[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) acc[r] = T(0);
  for (int64_t t = tb; t < te; ++t) {
    const T c = static_cast<T>(tco[t]);
    const TS* __restrict__ s0 = src + tsrc[t];
#pragma unroll
    for (int r = 0; r < PER; ++r)
      if (live[r]) acc[r] += c * static_cast<T>(s0[so[r]]);
  }
#pragma unroll
  for (int r = 0; r < PER; ++r)
    if (live[r]) dst[dst_off + dof[r]] = static_cast<TD>(acc[r]);
}

// TS: source element type, TD: destination element type (a cast rides along with the
// permutation); T: the arithmetic type, the wider of the two.
template <typename TS, typename TD>
__global__ __launch_bounds__(kThreads) void recouple_kernel(const TS* __restrict__ src,
                                                            TD* __restrict__ dst,
                                                            const int64_t* __restrict__ jobs,
                                                            const int64_t* __restrict__ tsrc,
                                                            const double* __restrict__ tco,
                                                            const int64_t* __restrict__ tiles) {
  using T = decltype(TS(0) * TD(0));
  __shared__ T lds[kT][kT + 1];
  const int64_t job_id = tiles[2 * blockIdx.x];
  const int64_t tile = tiles[2 * blockIdx.x + 1];
  const int64_t* jw = jobs + job_id * TN_JOB_WORDS;
  const int64_t dst_off = jw[0], tb = jw[1], te = jw[2];
  const int rank = static_cast<int>(jw[3]);
  const int64_t* shape = jw + 4;
  const int64_t* dstr = jw + 10;
  const int64_t* sstr = jw + 16;
  const int64_t mode = jw[22];
  const int64_t nelem = jw[23];
  const int tid = threadIdx.x;

  if (mode < 0) {
    // linear tiles: 2048 elements (mode -1) or, for programs with few elements, 512 (mode -2)
    if (mode == -2)
      linear_tile<T, TS, TD, 2>(src, dst, tsrc, tco, tb, te, rank, shape, dstr, sstr, dst_off, nelem, tile);
    else
      linear_tile<T, TS, TD, kLinPerThread>(src, dst, tsrc, tco, tb, te, rank, shape, dstr, sstr, dst_off,
                                            nelem, tile);
    return;
  }

  // LDS transpose tile over (I = rank-1, J = mode); remaining dims are batch.  Mode J + 8
  // asks for a 32x32 tile (four times the workgroups for a program with few large regions).
  const int E = mode >= 8 ? kT / 2 : kT;
  const int I = rank - 1;
  const int J = static_cast<int>(mode >= 8 ? mode - 8 : mode);
  const int64_t nI = (shape[I] + E - 1) / E;
  const int64_t nJ = (shape[J] + E - 1) / E;
  const int64_t ti = tile % nI;
  const int64_t tj = (tile / nI) % nJ;
  uint32_t b = static_cast<uint32_t>(tile / (nI * nJ));
  int64_t sbase = 0, dbase = 0;
  for (int d = rank - 1; d >= 0; --d) {
    if (d == I || d == J) continue;
    const uint32_t n = static_cast<uint32_t>(shape[d]);
    const uint32_t q = b / n;
    const int64_t i = b - q * n;
    b = q;
    sbase += i * sstr[d];
    dbase += i * dstr[d];
  }
  const int64_t i0 = ti * E, j0 = tj * E;
  const int64_t ni = min<int64_t>(E, shape[I] - i0);
  const int64_t nj = min<int64_t>(E, shape[J] - j0);
  const int64_t sI = sstr[I], sJ = sstr[J], dJ = dstr[J], dI = dstr[I];
  // load: threads run along J (source-contiguous)
  for (int r = 0; r < E * E / kThreads; ++r) {
    const int idx = r * kThreads + tid;
    const int j = idx % E, i = idx / E;
    if (i < ni && j < nj)
      lds[i][j] = gather_terms<T>(src, tsrc, tco, tb, te, sbase + (i0 + i) * sI + (j0 + j) * sJ);
  }
  __syncthreads();
  // store: threads run along I (destination-contiguous)
  for (int r = 0; r < E * E / kThreads; ++r) {
    const int idx = r * kThreads + tid;
    const int i = idx % E, j = idx / E;
    if (i < ni && j < nj) dst[dst_off + dbase + (j0 + j) * dJ + (i0 + i) * dI] = static_cast<TD>(lds[i][j]);
  }
}

}  // namespace

namespace {

template <typename TS, typename TD>
void launch_recouple(const void* src, void* dst, const int64_t* jobs, const int64_t* term_src,
                     const double* term_coeff, const int64_t* tiles, int64_t ntiles, hipStream_t s) {
  recouple_kernel<TS, TD><<<dim3(static_cast<unsigned>(ntiles)), dim3(kThreads), 0, s>>>(
      static_cast<const TS*>(src), static_cast<TD*>(dst), jobs, term_src, term_coeff, tiles);
}

}  // namespace

extern "C" int tn_recouple_cast(int src_dtype, int dst_dtype, const void* src, void* dst,
                                const int64_t* jobs, const int64_t* term_src,
                                const double* term_coeff, const int64_t* tiles, int64_t ntiles,
                                void* stream) {
  if (ntiles <= 0) return 0;
  if (ntiles > 0x7fffffffLL) return tn::fail("tn_recouple: too many tiles");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (src_dtype == TN_F64 && dst_dtype == TN_F64)
    launch_recouple<double, double>(src, dst, jobs, term_src, term_coeff, tiles, ntiles, s);
  else if (src_dtype == TN_F32 && dst_dtype == TN_F32)
    launch_recouple<float, float>(src, dst, jobs, term_src, term_coeff, tiles, ntiles, s);
  else if (src_dtype == TN_F64 && dst_dtype == TN_F32)
    launch_recouple<double, float>(src, dst, jobs, term_src, term_coeff, tiles, ntiles, s);
  else if (src_dtype == TN_F32 && dst_dtype == TN_F64)
    launch_recouple<float, double>(src, dst, jobs, term_src, term_coeff, tiles, ntiles, s);
  else
    return tn::fail("tn_recouple: unknown dtype");
  return tn::check_launch("tn_recouple");
}

extern "C" int tn_recouple(int dtype, const void* src, void* dst, const int64_t* jobs,
                           const int64_t* term_src, const double* term_coeff,
                           const int64_t* tiles, int64_t ntiles, void* stream) {
  return tn_recouple_cast(dtype, dtype, src, dst, jobs, term_src, term_coeff, tiles, ntiles, stream);
}

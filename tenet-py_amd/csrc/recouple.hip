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

// Linear tile: element e of the job in destination C order, PER elements per thread.  All
// of a thread's elements are decoded first and their loads issued unconditionally (a lane
// past the region's end reads the region's first element, in bounds by the host's checks,
// and discards it): a load under a per-lane branch would be waited for before the next one
// is issued, one memory round trip per element instead of one per term.
template <typename T, typename TS, typename TD, int PER>
__device__ __forceinline__ void linear_tile(const TS* __restrict__ src, TD* __restrict__ dst,
                                            const int64_t* __restrict__ tsrc,
                                            const double* __restrict__ tco, int64_t tb, int64_t te,
                                            int rank, const int64_t* shape, const int64_t* dstr,
                                            const int64_t* sstr, int64_t dst_off, int64_t nelem,
                                            int64_t tile) {
  const int tid = threadIdx.x;
  const int64_t base = tile * (kThreads * PER);
  // offsets within the region: below 2^31 (the host refuses larger regions)
  int32_t so[PER], dof[PER];
  bool live[PER];
#pragma unroll
  for (int r = 0; r < PER; ++r) {
    const int64_t e = base + r * kThreads + tid;
    live[r] = e < nelem;
    uint32_t rem = live[r] ? static_cast<uint32_t>(e) : 0u;
    int64_t s = 0, d = 0;
    for (int k = rank - 1; k > 0; --k) {
      const uint32_t n = static_cast<uint32_t>(shape[k]);
      const uint32_t q = rem / n;
      const int64_t i = rem - q * n;
      rem = q;
      s += i * sstr[k];
      d += i * dstr[k];
    }
    // what is left indexes dim 0 (e < nelem bounds it by shape[0])
    so[r] = static_cast<int32_t>(s + static_cast<int64_t>(rem) * sstr[0]);
    dof[r] = static_cast<int32_t>(d + static_cast<int64_t>(rem) * dstr[0]);
  }
  // acc = ((0 + c_0 v_0) + c_1 v_1) + ... in term order.  A single-term job loads straight
  // into the accumulators; a multi-term job (SU(2) recouplings, summed K slices) issues
  // the loads of up to TG terms together -- the terms of a group read through clamped
  // indices, so every load is unconditional -- and only then accumulates them in order.
  constexpr int TG = 4;
  T acc[PER];
  if (te - tb == 1) {
    const T c = static_cast<T>(tco[tb]);
    const TS* __restrict__ s0 = src + tsrc[tb];
#pragma unroll
    for (int r = 0; r < PER; ++r) acc[r] = static_cast<T>(s0[so[r]]);
#pragma unroll
    for (int r = 0; r < PER; ++r) acc[r] = T(0) + c * acc[r];
  } else {
#pragma unroll
    for (int r = 0; r < PER; ++r) acc[r] = T(0);
    for (int64_t t0 = tb; t0 < te; t0 += TG) {
      const int nt = te - t0 < TG ? static_cast<int>(te - t0) : TG;
      T v[TG][PER];
#pragma unroll
      for (int q = 0; q < TG; ++q) {
        const TS* __restrict__ sq = src + tsrc[t0 + (q < nt ? q : 0)];
#pragma unroll
        for (int r = 0; r < PER; ++r) v[q][r] = static_cast<T>(sq[so[r]]);
      }
#pragma unroll
      for (int q = 0; q < TG; ++q) {
        if (q < nt) {
          const T c = static_cast<T>(tco[t0 + q]);
#pragma unroll
          for (int r = 0; r < PER; ++r) acc[r] += c * v[q][r];
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < PER; ++r)
    if (live[r]) dst[dst_off + dof[r]] = static_cast<TD>(acc[r]);
}

// LDS transpose tile of edge E over (I = rank-1, J); remaining dims are batch.  Loads run
// along J (source-contiguous), all of a thread's loads of one term issued together (lanes
// outside the tile read its first element and drop it); stores run along I
// (destination-contiguous).
template <typename T, typename TS, typename TD, int E>
__device__ __forceinline__ void lds_tile(T (*lds)[kT + 1], const TS* __restrict__ src, TD* __restrict__ dst,
                                         const int64_t* __restrict__ tsrc, const double* __restrict__ tco,
                                         int64_t tb, int64_t te, int rank, int J, const int64_t* shape,
                                         const int64_t* dstr, const int64_t* sstr, int64_t dst_off,
                                         int64_t tile) {
  constexpr int R = E * E / kThreads;
  const int tid = threadIdx.x;
  const int I = rank - 1;
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
  const int64_t first = sbase + i0 * sI + j0 * sJ;
  // C loads in flight per thread at a time: enough to cover the latency, few registers
  constexpr int C = R < 8 ? R : 8;
#pragma unroll
  for (int r0 = 0; r0 < R; r0 += C) {
    int64_t off[C];
    bool ok[C];
    T v[C];
#pragma unroll
    for (int r = 0; r < C; ++r) {
      const int idx = (r0 + r) * kThreads + tid;
      const int j = idx % E, i = idx / E;
      ok[r] = i < ni && j < nj;
      off[r] = ok[r] ? first + i * sI + j * sJ : first;
      v[r] = T(0);
    }
    for (int64_t t = tb; t < te; ++t) {
      const T c = static_cast<T>(tco[t]);
      const TS* __restrict__ s0 = src + tsrc[t];
      T x[C];
#pragma unroll
      for (int r = 0; r < C; ++r) x[r] = static_cast<T>(s0[off[r]]);
#pragma unroll
      for (int r = 0; r < C; ++r) v[r] += c * x[r];
    }
#pragma unroll
    for (int r = 0; r < C; ++r) {
      const int idx = (r0 + r) * kThreads + tid;
      if (ok[r]) lds[idx / E][idx % E] = v[r];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int idx = r * kThreads + tid;
    const int i = idx % E, j = idx / E;
    if (i < ni && j < nj) dst[dst_off + dbase + (j0 + j) * dJ + (i0 + i) * dI] = static_cast<TD>(lds[i][j]);
  }
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

  if (mode < 0) {
    // linear tiles: 2048 elements (mode -1) or, for programs with few elements, 512 (mode -2)
    if (mode == -2)
      linear_tile<T, TS, TD, 2>(src, dst, tsrc, tco, tb, te, rank, shape, dstr, sstr, dst_off, nelem, tile);
    else
      linear_tile<T, TS, TD, kLinPerThread>(src, dst, tsrc, tco, tb, te, rank, shape, dstr, sstr, dst_off,
                                            nelem, tile);
    return;
  }

  // LDS transpose tile over (I = rank-1, J = mode); mode J + 8 asks for a 32x32 tile (four
  // times the workgroups for a program with few large regions)
  if (mode >= 8)
    lds_tile<T, TS, TD, kT / 2>(lds, src, dst, tsrc, tco, tb, te, rank, static_cast<int>(mode - 8), shape, dstr,
                                sstr, dst_off, tile);
  else
    lds_tile<T, TS, TD, kT>(lds, src, dst, tsrc, tco, tb, te, rank, static_cast<int>(mode), shape, dstr, sstr,
                            dst_off, tile);
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

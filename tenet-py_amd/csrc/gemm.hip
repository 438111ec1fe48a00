// Block-sparse grouped GEMM on gfx950 MFMA.
//
// A launch covers a list of independent problems (one per output block of a
// contraction).  Problem p writes C_p (M x N, row-major with ldc) as the sum over its K
// segments of op(A_s) op(B_s); every operand lives in place inside the flat block
// buffers of the tensors, so no sector matrix is ever gathered for operands whose block
// layout already is the matrix (or its transpose).  A host-built tile table lists every
// (problem, tile row, tile col); workgroups walk it persistently, with consecutive
// workgroups remapped onto one XCD so tiles sharing operand panels share an L2.
//
// Arithmetic (all on MFMA, 64-wide waves, 4 waves per workgroup in a 2x2 grid, each
// wave owning (BM/2)x(BN/2) as 16x16 MFMA sub-tiles):
//   TN_PREC_FP64   v_mfma_f64_16x16x4_f64, BK = 16
//   TN_PREC_FP32   v_mfma_f32_16x16x4_f32, BK = 16
//   TN_PREC_BF16X3 v_mfma_f32_16x16x32_bf16, BK = 32; each FP32 operand is split into
//                  hi = bf16(x), lo = bf16(x - hi) while staging to LDS and the product is
//                  hi*hi + hi*lo + lo*hi with FP32 accumulation.
// Accumulator layout (16x16 tile): FP32 results -- lane l, register j holds
// C[(l/16)*4 + j][l%16]; FP64 results -- C[4*j + l/16][l%16].
#include "tn_common.hpp"

namespace {

constexpr int kThreads = 256;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

template <int PREC>
struct Traits;
template <>
struct Traits<TN_PREC_FP64> {
  using T = double;
  using Acc = f64x4;
  static constexpr int BK = 16;
};
template <>
struct Traits<TN_PREC_FP32> {
  using T = float;
  using Acc = f32x4;
  static constexpr int BK = 16;
};
template <>
struct Traits<TN_PREC_BF16X3> {
  using T = float;
  using Acc = f32x4;
  static constexpr int BK = 32;
};

__device__ __forceinline__ unsigned short bf16_rne(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
  return __uint_as_float(static_cast<unsigned int>(h) << 16);
}

// LDS staging buffers for one K step.
template <int PREC, int BM, int BN>
struct Smem;
template <int BM, int BN>
struct Smem<TN_PREC_FP64, BM, BN> {
  double a[16][BM + 2];
  double b[16][BN + 2];
};
template <int BM, int BN>
struct Smem<TN_PREC_FP32, BM, BN> {
  float a[16][BM + 4];
  float b[16][BN + 4];
};
template <int BM, int BN>
struct Smem<TN_PREC_BF16X3, BM, BN> {
  // k contiguous so that one lane reads its 8 consecutive k values with one 16B load
  unsigned short ah[BM][32 + 8];
  unsigned short al[BM][32 + 8];
  unsigned short bh[BN][32 + 8];
  unsigned short bl[BN][32 + 8];
};

template <int PREC, bool TA, bool TB, int BM, int BN>
__global__ __launch_bounds__(kThreads) void gemm_kernel(
    const typename Traits<PREC>::T* __restrict__ A, const typename Traits<PREC>::T* __restrict__ B,
    typename Traits<PREC>::T* __restrict__ C, const int64_t* __restrict__ probs,
    const int64_t* __restrict__ segs, const int32_t* __restrict__ tiles, int ntiles) {
  using T = typename Traits<PREC>::T;
  using Acc = typename Traits<PREC>::Acc;
  constexpr int BK = Traits<PREC>::BK;
  constexpr int WM = BM / 2, WN = BN / 2;  // wave tile
  constexpr int SM = WM / 16, SN = WN / 16;
  __shared__ Smem<PREC, BM, BN> sm;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM;
  const int wn0 = (wave & 1) * WN;

  // XCD-aware order: the hardware deals consecutive workgroup ids round-robin over the
  // 8 XCDs; give each XCD a contiguous run of logical ids.
  const int G = gridDim.x;
  int logical = blockIdx.x;
  if ((G & 7) == 0) logical = (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);

  for (int t = logical; t < ntiles; t += G) {
    const int p = tiles[4 * t];
    const int m0 = tiles[4 * t + 1] * BM;
    const int n0 = tiles[4 * t + 2] * BN;
    const int64_t* pw = probs + static_cast<int64_t>(p) * TN_PROB_WORDS;
    const int M = static_cast<int>(pw[0]);
    const int N = static_cast<int>(pw[1]);
    const int64_t c_off = pw[2];
    const int64_t ldc = pw[3];
    const int64_t sb = pw[4], se = pw[5];

    Acc acc[SM][SN];
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SN; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    for (int64_t s = sb; s < se; ++s) {
      const int64_t* sw = segs + s * TN_SEG_WORDS;
      const T* Ap = A + sw[0];
      const T* Bp = B + sw[1];
      const int K = static_cast<int>(sw[2]);
      const int64_t lda = sw[3], ldb = sw[4];
      for (int k0 = 0; k0 < K; k0 += BK) {
        // ---- stage A (BM x BK) and B (BK x BN) into LDS ----
#pragma unroll
        for (int r = 0; r < BM * BK / kThreads; ++r) {
          const int idx = r * kThreads + tid;
          int m, k;
          if (!TA) { k = idx % BK; m = idx / BK; } else { m = idx % BM; k = idx / BM; }
          const int gm = m0 + m, gk = k0 + k;
          T v = T(0);
          if (gm < M && gk < K) v = TA ? Ap[gk * lda + gm] : Ap[gm * lda + gk];
          if constexpr (PREC == TN_PREC_BF16X3) {
            const unsigned short h = bf16_rne(v);
            sm.ah[m][k] = h;
            sm.al[m][k] = bf16_rne(v - bf16_to_f32(h));
          } else {
            sm.a[k][m] = v;
          }
        }
#pragma unroll
        for (int r = 0; r < BN * BK / kThreads; ++r) {
          const int idx = r * kThreads + tid;
          int n, k;
          if (!TB) { n = idx % BN; k = idx / BN; } else { k = idx % BK; n = idx / BK; }
          const int gn = n0 + n, gk = k0 + k;
          T v = T(0);
          if (gn < N && gk < K) v = TB ? Bp[gn * ldb + gk] : Bp[gk * ldb + gn];
          if constexpr (PREC == TN_PREC_BF16X3) {
            const unsigned short h = bf16_rne(v);
            sm.bh[n][k] = h;
            sm.bl[n][k] = bf16_rne(v - bf16_to_f32(h));
          } else {
            sm.b[k][n] = v;
          }
        }
        __syncthreads();
        // ---- MFMA over the staged K step ----
        if constexpr (PREC == TN_PREC_BF16X3) {
          const int kq = (lane >> 4) * 8;
          bf16x8 ah[SM], al[SM];
#pragma unroll
          for (int i = 0; i < SM; ++i) {
            const int row = wm0 + i * 16 + (lane & 15);
            ah[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(&sm.ah[row][kq]));
            al[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(&sm.al[row][kq]));
          }
#pragma unroll
          for (int j = 0; j < SN; ++j) {
            const int col = wn0 + j * 16 + (lane & 15);
            const bf16x8 bh = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(&sm.bh[col][kq]));
            const bf16x8 bl = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(&sm.bl[col][kq]));
#pragma unroll
            for (int i = 0; i < SM; ++i) {
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, acc[i][j], 0, 0, 0);
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
            }
          }
        } else {
#pragma unroll
          for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4);
            T a[SM], b[SN];
#pragma unroll
            for (int i = 0; i < SM; ++i) a[i] = sm.a[kr][wm0 + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < SN; ++j) b[j] = sm.b[kr][wn0 + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < SM; ++i)
#pragma unroll
              for (int j = 0; j < SN; ++j) {
                if constexpr (PREC == TN_PREC_FP64)
                  acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
                else
                  acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
              }
          }
        }
        __syncthreads();
      }
    }
    // ---- epilogue ----
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SN; ++j) {
        const int col = n0 + wn0 + j * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // v_mfma_f64_16x16x4 spreads rows over registers (row 4r + l/16); the FP32
          // accumulators keep four consecutive rows per lane group (row 4(l/16) + r)
          const int rr = PREC == TN_PREC_FP64 ? r * 4 + (lane >> 4) : (lane >> 4) * 4 + r;
          const int row = m0 + wm0 + i * 16 + rr;
          if (row < M && col < N) C[c_off + row * ldc + col] = acc[i][j][r];
        }
      }
  }
}

template <int PREC, bool TA, bool TB>
int launch_tile(int bm, int bn, const void* A, const void* B, void* C, const int64_t* probs,
                const int64_t* segs, const int32_t* tiles, int ntiles, int grid, hipStream_t s) {
  using T = typename Traits<PREC>::T;
  const T* a = static_cast<const T*>(A);
  const T* b = static_cast<const T*>(B);
  T* c = static_cast<T*>(C);
  const int g = grid > 0 ? grid : ntiles;
#define TN_GEMM_CASE(BM_, BN_)                                                              \
  if (bm == BM_ && bn == BN_) {                                                             \
    gemm_kernel<PREC, TA, TB, BM_, BN_><<<g, kThreads, 0, s>>>(a, b, c, probs, segs, tiles, \
                                                               ntiles);                    \
    return tn::check_launch("tn_gemm_grouped");                                            \
  }
  TN_GEMM_CASE(32, 32)
  TN_GEMM_CASE(64, 64)
  TN_GEMM_CASE(128, 64)
  TN_GEMM_CASE(64, 128)
  TN_GEMM_CASE(128, 128)
#undef TN_GEMM_CASE
  return tn::fail("tn_gemm_grouped: unsupported tile shape");
}

template <int PREC>
int launch_prec(int ta, int tb, int bm, int bn, const void* A, const void* B, void* C,
                const int64_t* probs, const int64_t* segs, const int32_t* tiles, int ntiles,
                int grid, hipStream_t s) {
  if (!ta && !tb) return launch_tile<PREC, false, false>(bm, bn, A, B, C, probs, segs, tiles, ntiles, grid, s);
  if (ta && !tb) return launch_tile<PREC, true, false>(bm, bn, A, B, C, probs, segs, tiles, ntiles, grid, s);
  if (!ta && tb) return launch_tile<PREC, false, true>(bm, bn, A, B, C, probs, segs, tiles, ntiles, grid, s);
  return launch_tile<PREC, true, true>(bm, bn, A, B, C, probs, segs, tiles, ntiles, grid, s);
}

}  // namespace

extern "C" int tn_gemm_grouped(int precision, int dtype, int trans_a, int trans_b,
                               const void* A, const void* B, void* C, const int64_t* problems,
                               const int64_t* segs, const int32_t* tiles, int64_t ntiles,
                               int bm, int bn, int grid, void* stream) {
  if (ntiles <= 0) return 0;
  if (ntiles > 0x7fffffffLL) return tn::fail("tn_gemm_grouped: too many tiles");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nt = static_cast<int>(ntiles);
  if (precision == TN_PREC_FP64) {
    if (dtype != TN_F64) return tn::fail("tn_gemm_grouped: FP64 arithmetic needs FP64 operands");
    return launch_prec<TN_PREC_FP64>(trans_a, trans_b, bm, bn, A, B, C, problems, segs, tiles, nt, grid, s);
  }
  if (dtype != TN_F32) return tn::fail("tn_gemm_grouped: FP32/BF16x3 arithmetic needs FP32 operands");
  if (precision == TN_PREC_FP32)
    return launch_prec<TN_PREC_FP32>(trans_a, trans_b, bm, bn, A, B, C, problems, segs, tiles, nt, grid, s);
  if (precision == TN_PREC_BF16X3)
    return launch_prec<TN_PREC_BF16X3>(trans_a, trans_b, bm, bn, A, B, C, problems, segs, tiles, nt, grid, s);
  return tn::fail("tn_gemm_grouped: unknown precision");
}

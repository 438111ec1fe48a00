// Block-sparse grouped GEMM on gfx950 MFMA.
//
// A launch covers a list of independent problems (one per output block of a
// contraction, or one K-slice of one when the host splits K).  Problem p writes C_p
// (M x N, row-major with ldc) = sum over its K segments of op(A_s) op(B_s); operands are
// read in place inside the flat block buffers (no sector matrix is gathered), and the
// output goes either to C or -- for split-K partials -- to the workspace W.  A host-built
// tile table lists every (problem, tile row, tile col); workgroups walk it persistently,
// with consecutive workgroups remapped onto one XCD so tiles sharing operand panels share
// an L2.
//
// Per workgroup: 4 waves (64 lanes) in a 2x2 grid, each owning (BM/2)x(BN/2) as 16x16
// MFMA sub-tiles.  The K loop is software pipelined: the next K step's global loads are
// issued into registers before the MFMAs on the current LDS buffer run, and LDS is double
// buffered so one barrier per K step suffices.  The k order inside a K step is permuted
// (lane group g of an MFMA step t takes k = 4g + t) so that every lane fetches its four
// k values with one 16-byte LDS read from k-contiguous LDS rows.
//
//   TN_PREC_FP64   v_mfma_f64_16x16x4_f64,  BK = 16
//   TN_PREC_FP32   v_mfma_f32_16x16x4_f32,  BK = 32
//   TN_PREC_BF16X3 v_mfma_f32_16x16x32_bf16, BK = 32; each FP32 operand is split into
//                  hi = bf16(x), lo = bf16(x - hi) while staging to LDS and the product is
//                  lo*hi + hi*lo + hi*hi with FP32 accumulation.
// Accumulator layout (16x16 tile): FP32 results -- lane l, register j holds
// C[(l/16)*4 + j][l%16]; FP64 results -- C[4*j + l/16][l%16].
#include "tn_common.hpp"

namespace {

constexpr int kThreads = 256;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
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
  static constexpr int PAD = 2;
};
template <>
struct Traits<TN_PREC_FP32> {
  using T = float;
  using Acc = f32x4;
  static constexpr int BK = 32;
  static constexpr int PAD = 4;
};
template <>
struct Traits<TN_PREC_BF16X3> {
  using T = float;
  using Acc = f32x4;
  static constexpr int BK = 32;
  static constexpr int PAD = 8;
};

__device__ __forceinline__ unsigned short bf16_rne(float x) {
  unsigned int u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(unsigned short h) {
  return __uint_as_float(static_cast<unsigned int>(h) << 16);
}

// Double-buffered LDS, rows k-contiguous: a[buf][m][k], b[buf][n][k].
template <int PREC, int BM, int BN>
struct Smem {
  using T = typename Traits<PREC>::T;
  static constexpr int LK = Traits<PREC>::BK + Traits<PREC>::PAD;
  T a[2][BM][LK];
  T b[2][BN][LK];
};
template <int BM, int BN>
struct Smem<TN_PREC_BF16X3, BM, BN> {
  static constexpr int LK = 32 + 8;
  unsigned short ah[2][BM][LK];
  unsigned short al[2][BM][LK];
  unsigned short bh[2][BN][LK];
  unsigned short bl[2][BN][LK];
};

struct Seg {
  int64_t a_off, b_off, lda, ldb;
  int K;
};

__device__ __forceinline__ Seg load_seg(const int64_t* __restrict__ segs, int64_t s) {
  const int64_t* sw = segs + s * TN_SEG_WORDS;
  return Seg{sw[0], sw[1], sw[3], sw[4], static_cast<int>(sw[2])};
}

template <int PREC, bool TA, bool TB, int BM, int BN>
__global__ __launch_bounds__(kThreads) void gemm_kernel(
    const typename Traits<PREC>::T* __restrict__ A, const typename Traits<PREC>::T* __restrict__ B,
    typename Traits<PREC>::T* __restrict__ C, typename Traits<PREC>::T* __restrict__ W,
    const int64_t* __restrict__ probs, const int64_t* __restrict__ segs,
    const int32_t* __restrict__ tiles, int ntiles) {
  using T = typename Traits<PREC>::T;
  using Acc = typename Traits<PREC>::Acc;
  constexpr int BK = Traits<PREC>::BK;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int SM = WM / 16, SN = WN / 16;
  constexpr int LA = BM * BK / kThreads;  // prefetch registers per thread
  constexpr int LB = BN * BK / kThreads;
  __shared__ Smem<PREC, BM, BN> sm;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM;
  const int wn0 = (wave & 1) * WN;
  const int g = lane >> 4;
  const int r16 = lane & 15;

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
    T* __restrict__ out = pw[6] ? W : C;

    Acc acc[SM][SN];
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SN; ++j) acc[i][j] = Acc{0, 0, 0, 0};

    T ra[LA], rb[LB];
    auto fetch = [&](const Seg& sg, int k0) {
      const T* Ap = A + sg.a_off;
      const T* Bp = B + sg.b_off;
#pragma unroll
      for (int r = 0; r < LA; ++r) {
        const int idx = r * kThreads + tid;
        int m, k;
        if (!TA) { k = idx % BK; m = idx / BK; } else { m = idx % BM; k = idx / BM; }
        const int gm = m0 + m, gk = k0 + k;
        ra[r] = (gm < M && gk < sg.K) ? (TA ? Ap[gk * sg.lda + gm] : Ap[gm * sg.lda + gk]) : T(0);
      }
#pragma unroll
      for (int r = 0; r < LB; ++r) {
        const int idx = r * kThreads + tid;
        int n, k;
        if (!TB) { n = idx % BN; k = idx / BN; } else { k = idx % BK; n = idx / BK; }
        const int gn = n0 + n, gk = k0 + k;
        rb[r] = (gn < N && gk < sg.K) ? (TB ? Bp[gn * sg.ldb + gk] : Bp[gk * sg.ldb + gn]) : T(0);
      }
    };
    auto stash = [&](int buf) {
#pragma unroll
      for (int r = 0; r < LA; ++r) {
        const int idx = r * kThreads + tid;
        int m, k;
        if (!TA) { k = idx % BK; m = idx / BK; } else { m = idx % BM; k = idx / BM; }
        if constexpr (PREC == TN_PREC_BF16X3) {
          const unsigned short h = bf16_rne(ra[r]);
          sm.ah[buf][m][k] = h;
          sm.al[buf][m][k] = bf16_rne(ra[r] - bf16_to_f32(h));
        } else {
          sm.a[buf][m][k] = ra[r];
        }
      }
#pragma unroll
      for (int r = 0; r < LB; ++r) {
        const int idx = r * kThreads + tid;
        int n, k;
        if (!TB) { n = idx % BN; k = idx / BN; } else { k = idx % BK; n = idx / BK; }
        if constexpr (PREC == TN_PREC_BF16X3) {
          const unsigned short h = bf16_rne(rb[r]);
          sm.bh[buf][n][k] = h;
          sm.bl[buf][n][k] = bf16_rne(rb[r] - bf16_to_f32(h));
        } else {
          sm.b[buf][n][k] = rb[r];
        }
      }
    };

    int64_t s = sb;
    int k0 = 0;
    Seg sg{};
    if (s < se) {
      sg = load_seg(segs, s);
      fetch(sg, 0);
    }
    int buf = 0;
    while (s < se) {
      stash(buf);
      __syncthreads();
      // advance and prefetch the next K step while this one computes
      int64_t ns = s;
      int nk = k0 + BK;
      Seg nsg = sg;
      if (nk >= sg.K) {
        ++ns;
        nk = 0;
        if (ns < se) nsg = load_seg(segs, ns);
      }
      if (ns < se) fetch(nsg, nk);

      if constexpr (PREC == TN_PREC_BF16X3) {
        const int kq = g * 8;
        bf16x8 ah[SM], al[SM];
#pragma unroll
        for (int i = 0; i < SM; ++i) {
          const int row = wm0 + i * 16 + r16;
          ah[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(&sm.ah[buf][row][kq]));
          al[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(&sm.al[buf][row][kq]));
        }
#pragma unroll
        for (int j = 0; j < SN; ++j) {
          const int col = wn0 + j * 16 + r16;
          const bf16x8 bh = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(&sm.bh[buf][col][kq]));
          const bf16x8 bl = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(&sm.bl[buf][col][kq]));
#pragma unroll
          for (int i = 0; i < SM; ++i) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh, acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl, acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh, acc[i][j], 0, 0, 0);
          }
        }
      } else if constexpr (PREC == TN_PREC_FP32) {
#pragma unroll
        for (int h = 0; h < BK / 16; ++h) {
          const int kq = h * 16 + g * 4;
          f32x4 a4[SM], b4[SN];
#pragma unroll
          for (int i = 0; i < SM; ++i) a4[i] = *reinterpret_cast<const f32x4*>(&sm.a[buf][wm0 + i * 16 + r16][kq]);
#pragma unroll
          for (int j = 0; j < SN; ++j) b4[j] = *reinterpret_cast<const f32x4*>(&sm.b[buf][wn0 + j * 16 + r16][kq]);
#pragma unroll
          for (int tt = 0; tt < 4; ++tt)
#pragma unroll
            for (int i = 0; i < SM; ++i)
#pragma unroll
              for (int j = 0; j < SN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[i][tt], b4[j][tt], acc[i][j], 0, 0, 0);
        }
      } else {
        const int kq = g * 4;
        f64x2 a0[SM], a1[SM], b0[SN], b1[SN];
#pragma unroll
        for (int i = 0; i < SM; ++i) {
          const double* pa = &sm.a[buf][wm0 + i * 16 + r16][kq];
          a0[i] = *reinterpret_cast<const f64x2*>(pa);
          a1[i] = *reinterpret_cast<const f64x2*>(pa + 2);
        }
#pragma unroll
        for (int j = 0; j < SN; ++j) {
          const double* pb = &sm.b[buf][wn0 + j * 16 + r16][kq];
          b0[j] = *reinterpret_cast<const f64x2*>(pb);
          b1[j] = *reinterpret_cast<const f64x2*>(pb + 2);
        }
#pragma unroll
        for (int i = 0; i < SM; ++i)
#pragma unroll
          for (int j = 0; j < SN; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[i][0], b0[j][0], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[i][1], b0[j][1], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[i][0], b1[j][0], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[i][1], b1[j][1], acc[i][j], 0, 0, 0);
          }
      }
      buf ^= 1;
      s = ns;
      k0 = nk;
      sg = nsg;
    }
    // ---- epilogue ----
#pragma unroll
    for (int i = 0; i < SM; ++i)
#pragma unroll
      for (int j = 0; j < SN; ++j) {
        const int col = n0 + wn0 + j * 16 + r16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // v_mfma_f64_16x16x4 spreads rows over registers (row 4r + l/16); the FP32
          // accumulators keep four consecutive rows per lane group (row 4(l/16) + r)
          const int rr = PREC == TN_PREC_FP64 ? r * 4 + g : g * 4 + r;
          const int row = m0 + wm0 + i * 16 + rr;
          if (row < M && col < N) out[c_off + row * ldc + col] = acc[i][j][r];
        }
      }
    // the next tile's first stash reuses buffer 0/1: make sure every wave is done reading
    __syncthreads();
  }
}

template <int PREC, bool TA, bool TB>
int launch_tile(int bm, int bn, const void* A, const void* B, void* C, void* W,
                const int64_t* probs, const int64_t* segs, const int32_t* tiles, int ntiles,
                int grid, hipStream_t s) {
  using T = typename Traits<PREC>::T;
  const T* a = static_cast<const T*>(A);
  const T* b = static_cast<const T*>(B);
  T* c = static_cast<T*>(C);
  T* w = static_cast<T*>(W);
  const int g = grid > 0 ? grid : ntiles;
#define TN_GEMM_CASE(BM_, BN_)                                                              \
  if (bm == BM_ && bn == BN_) {                                                             \
    gemm_kernel<PREC, TA, TB, BM_, BN_><<<g, kThreads, 0, s>>>(a, b, c, w, probs, segs,     \
                                                               tiles, ntiles);             \
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
int launch_prec(int ta, int tb, int bm, int bn, const void* A, const void* B, void* C, void* W,
                const int64_t* probs, const int64_t* segs, const int32_t* tiles, int ntiles,
                int grid, hipStream_t s) {
  if (!ta && !tb) return launch_tile<PREC, false, false>(bm, bn, A, B, C, W, probs, segs, tiles, ntiles, grid, s);
  if (ta && !tb) return launch_tile<PREC, true, false>(bm, bn, A, B, C, W, probs, segs, tiles, ntiles, grid, s);
  if (!ta && tb) return launch_tile<PREC, false, true>(bm, bn, A, B, C, W, probs, segs, tiles, ntiles, grid, s);
  return launch_tile<PREC, true, true>(bm, bn, A, B, C, W, probs, segs, tiles, ntiles, grid, s);
}

}  // namespace

extern "C" int tn_gemm_grouped(int precision, int dtype, int trans_a, int trans_b,
                               const void* A, const void* B, void* C, void* W,
                               const int64_t* problems, const int64_t* segs,
                               const int32_t* tiles, int64_t ntiles, int bm, int bn, int grid,
                               void* stream) {
  if (ntiles <= 0) return 0;
  if (ntiles > 0x7fffffffLL) return tn::fail("tn_gemm_grouped: too many tiles");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nt = static_cast<int>(ntiles);
  if (precision == TN_PREC_FP64) {
    if (dtype != TN_F64) return tn::fail("tn_gemm_grouped: FP64 arithmetic needs FP64 operands");
    return launch_prec<TN_PREC_FP64>(trans_a, trans_b, bm, bn, A, B, C, W, problems, segs, tiles, nt, grid, s);
  }
  if (dtype != TN_F32) return tn::fail("tn_gemm_grouped: FP32/BF16x3 arithmetic needs FP32 operands");
  if (precision == TN_PREC_FP32)
    return launch_prec<TN_PREC_FP32>(trans_a, trans_b, bm, bn, A, B, C, W, problems, segs, tiles, nt, grid, s);
  if (precision == TN_PREC_BF16X3)
    return launch_prec<TN_PREC_BF16X3>(trans_a, trans_b, bm, bn, A, B, C, W, problems, segs, tiles, nt, grid, s);
  return tn::fail("tn_gemm_grouped: unknown precision");
}

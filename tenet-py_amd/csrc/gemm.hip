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

#include <mutex>

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
#ifndef TN_FP32_BK
#define TN_FP32_BK 32
#endif
#ifndef TN_PIPE_WAVES_PER_EU
#define TN_PIPE_WAVES_PER_EU 4
#endif
template <>
struct Traits<TN_PREC_FP32> {
  using T = float;
  using Acc = f32x4;
  static constexpr int BK = TN_FP32_BK;
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
  __shared__ __attribute__((aligned(16))) Smem<PREC, BM, BN> sm;

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

  // Tile descriptor of the tile being fetched (may run one tile ahead of the one computed).
  struct Tile {
    int m0, n0, M, N;
    int64_t c_off, ldc, sb, se;
    int flag;
  };
  auto tile_of = [&](int t) -> Tile {
    const int p = tiles[4 * t];
    const int64_t* pw = probs + static_cast<int64_t>(p) * TN_PROB_WORDS;
    return Tile{tiles[4 * t + 1] * BM, tiles[4 * t + 2] * BN, static_cast<int>(pw[0]),
                static_cast<int>(pw[1]), pw[2], pw[3], pw[4], pw[5], static_cast<int>(pw[6])};
  };
  {
    // ---- operand staging in quads: 4 consecutive k of one row (m of A, n of B) ----
    // A quad is fetched with one 16-byte buffer load when k is contiguous in memory, or
    // with 4 scalar loads (coalesced across lanes along m/n) when it is not; either way it
    // is written to LDS with one conflict-free 16-byte store.  Offsets inside a segment
    // are computed once per segment; the K-step advance moves the descriptor base
    // (scalar); elements outside the problem read through an offset past the
    // descriptor's range, for which the hardware returns zero.
    using V4 = T __attribute__((ext_vector_type(4)));
    constexpr int QK = BK / 4;  // quads per row of a K step
    constexpr int QA = (BM * QK + kThreads - 1) / kThreads;
    constexpr int QB = (BN * QK + kThreads - 1) / kThreads;
    constexpr bool QA_EXACT = BM * QK % kThreads == 0;  // else some threads own no A quad
    constexpr bool QB_EXACT = BN * QK % kThreads == 0;
    V4 ra[QA], rb[QB];
    uint32_t ao[QA], bo[QB];
    bool arow[QA], brow[QB];
    int akq[QA], bkq[QB];
    int am[QA], bn_[QB];
    uint32_t astr = 0, bstr = 0;  // byte stride between the 4 k of a quad when not contiguous
    int64_t fseg = -1;
    int ftile = -1;
    auto offsets = [&](const Tile& tl, const Seg& sg, int64_t sid) {
      fseg = sid;
      const int m0 = tl.m0, n0 = tl.n0, M = tl.M, N = tl.N;
      astr = static_cast<uint32_t>(sg.lda * sizeof(T));
      bstr = static_cast<uint32_t>(sg.ldb * sizeof(T));
#pragma unroll
      for (int r = 0; r < QA; ++r) {
        const int q = r * kThreads + tid;
        int m, kq;
        if (!TA) { m = q / QK; kq = (q % QK) * 4; } else { m = q % BM; kq = (q / BM) * 4; }
        const bool live = QA_EXACT || q < BM * QK;
        if (!live) { m = 0; kq = 0; }
        am[r] = live ? m : -1;
        akq[r] = kq;
        arow[r] = live && m0 + m < M;
        const int64_t e = TA ? (int64_t)kq * sg.lda + (m0 + m) : (int64_t)(m0 + m) * sg.lda + kq;
        ao[r] = static_cast<uint32_t>(e * sizeof(T));
      }
#pragma unroll
      for (int r = 0; r < QB; ++r) {
        const int q = r * kThreads + tid;
        int n, kq;
        if (TB) { n = q / QK; kq = (q % QK) * 4; } else { n = q % BN; kq = (q / BN) * 4; }
        const bool live = QB_EXACT || q < BN * QK;
        if (!live) { n = 0; kq = 0; }
        bn_[r] = live ? n : -1;
        bkq[r] = kq;
        brow[r] = live && n0 + n < N;
        const int64_t e = TB ? (int64_t)(n0 + n) * sg.ldb + kq : (int64_t)kq * sg.ldb + (n0 + n);
        bo[r] = static_cast<uint32_t>(e * sizeof(T));
      }
    };
    auto load1 = [&](__amdgpu_buffer_rsrc_t d, uint32_t off) -> T {
      if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(d, off, 0, 0));
      else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(d, off, 0, 0));
    };
    // quad of 4 consecutive k starting at byte offset `off`, k-contiguous in memory
    auto load_contig = [&](__amdgpu_buffer_rsrc_t d, uint32_t off, bool row_ok, int kq, int kleft) -> V4 {
      V4 v;
      if (row_ok && kq + 3 < kleft) {
        if constexpr (sizeof(T) == 8) {
          typedef double d2 __attribute__((ext_vector_type(2)));
          const d2 lo = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(d, off, 0, 0));
          const d2 hi = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(d, off + 16, 0, 0));
          v = V4{lo[0], lo[1], hi[0], hi[1]};
        } else {
          v = __builtin_bit_cast(V4, __builtin_amdgcn_raw_buffer_load_b128(d, off, 0, 0));
        }
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          v[j] = load1(d, (row_ok && kq + j < kleft) ? off + j * (uint32_t)sizeof(T) : 0x80000000u);
      }
      return v;
    };
    // quad of 4 k values strided by `str` bytes (lanes run along m/n)
    auto load_strided = [&](__amdgpu_buffer_rsrc_t d, uint32_t off, uint32_t str, bool row_ok, int kq, int kleft) -> V4 {
      V4 v;
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = load1(d, (row_ok && kq + j < kleft) ? off + j * str : 0x80000000u);
      return v;
    };
    auto fetch = [&](int t, const Tile& tl, const Seg& sg, int64_t sid, int k0) {
      if (sid != fseg || t != ftile) {
        ftile = t;
        offsets(tl, sg, sid);
      }
      const T* Ap = A + sg.a_off + (TA ? (int64_t)k0 * sg.lda : k0);
      const T* Bp = B + sg.b_off + (TB ? k0 : (int64_t)k0 * sg.ldb);
      const __amdgpu_buffer_rsrc_t ra_d = __builtin_amdgcn_make_buffer_rsrc((void*)Ap, (short)0, 0x7fffffff, 0x00020000);
      const __amdgpu_buffer_rsrc_t rb_d = __builtin_amdgcn_make_buffer_rsrc((void*)Bp, (short)0, 0x7fffffff, 0x00020000);
      const int kleft = sg.K - k0;
#pragma unroll
      for (int r = 0; r < QA; ++r)
        ra[r] = TA ? load_strided(ra_d, ao[r], astr, arow[r], akq[r], kleft)
                   : load_contig(ra_d, ao[r], arow[r], akq[r], kleft);
#pragma unroll
      for (int r = 0; r < QB; ++r)
        rb[r] = TB ? load_contig(rb_d, bo[r], brow[r], bkq[r], kleft)
                   : load_strided(rb_d, bo[r], bstr, brow[r], bkq[r], kleft);
    };
    auto stash = [&](int buf) {
#pragma unroll
      for (int r = 0; r < QA; ++r) {
        if (!QA_EXACT && am[r] < 0) continue;
        if constexpr (PREC == TN_PREC_BF16X3) {
          typedef unsigned short u4 __attribute__((ext_vector_type(4)));
          u4 h, l;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            h[j] = bf16_rne(ra[r][j]);
            l[j] = bf16_rne(ra[r][j] - bf16_to_f32(h[j]));
          }
          *reinterpret_cast<u4*>(&sm.ah[buf][am[r]][akq[r]]) = h;
          *reinterpret_cast<u4*>(&sm.al[buf][am[r]][akq[r]]) = l;
        } else {
          *static_cast<V4*>(__builtin_assume_aligned(&sm.a[buf][am[r]][akq[r]], 16)) = ra[r];
        }
      }
#pragma unroll
      for (int r = 0; r < QB; ++r) {
        if (!QB_EXACT && bn_[r] < 0) continue;
        if constexpr (PREC == TN_PREC_BF16X3) {
          typedef unsigned short u4 __attribute__((ext_vector_type(4)));
          u4 h, l;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            h[j] = bf16_rne(rb[r][j]);
            l[j] = bf16_rne(rb[r][j] - bf16_to_f32(h[j]));
          }
          *reinterpret_cast<u4*>(&sm.bh[buf][bn_[r]][bkq[r]]) = h;
          *reinterpret_cast<u4*>(&sm.bl[buf][bn_[r]][bkq[r]]) = l;
        } else {
          *static_cast<V4*>(__builtin_assume_aligned(&sm.b[buf][bn_[r]][bkq[r]], 16)) = rb[r];
        }
      }
    };

    int buf = 0;
    int t = logical;
    Tile cur{};
    // fetch position: tile, segment, k0 of the data sitting in the prefetch registers
    int pt = -1;
    int64_t ps = 0;
    int pk = 0;
    Seg psg{};
    if (t < ntiles) {
      cur = tile_of(t);
      if (cur.sb < cur.se) {
        psg = load_seg(segs, cur.sb);
        fetch(t, cur, psg, cur.sb, 0);
        pt = t; ps = cur.sb; pk = 0;
      }
    }
    while (t < ntiles) {
      Acc acc[SM][SN];
#pragma unroll
      for (int i = 0; i < SM; ++i)
#pragma unroll
        for (int j = 0; j < SN; ++j) acc[i][j] = Acc{0, 0, 0, 0};
      if (cur.sb < cur.se && pt != t) {  // nothing prefetched for this tile yet
        psg = load_seg(segs, cur.sb);
        fetch(t, cur, psg, cur.sb, 0);
        pt = t; ps = cur.sb; pk = 0;
      }
      const int tn = t + G;
      const Tile nxt = tn < ntiles ? tile_of(tn) : Tile{};
      while (pt == t) {
        stash(buf);
        __syncthreads();
        // advance the fetch position: next K step of this tile, else the next tile's first
        int64_t ns = ps;
        int nk = pk + BK;
        Seg nsg = psg;
        if (nk >= psg.K) {
          ++ns;
          nk = 0;
          if (ns < cur.se) nsg = load_seg(segs, ns);
        }
        if (ns < cur.se) {
          fetch(t, cur, nsg, ns, nk);
          ps = ns; pk = nk; psg = nsg;
        } else if (tn < ntiles && nxt.sb < nxt.se) {
          nsg = load_seg(segs, nxt.sb);
          fetch(tn, nxt, nsg, nxt.sb, 0);
          pt = tn; ps = nxt.sb; pk = 0; psg = nsg;
        } else {
          pt = -1;
        }
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
      }
      {
        const int m0 = cur.m0, n0 = cur.n0, M = cur.M, N = cur.N;
        const int64_t c_off = cur.c_off, ldc = cur.ldc;
        T* __restrict__ out = cur.flag ? W : C;
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
      }
      t = tn;
      cur = nxt;
    }
  }
}

// ---- one tile per workgroup, two K steps of global loads in flight ----------------------
// The launch form of tn_gemm_grouped with grid = 0 (FP32 and BF16x3; FP64 and the
// persistent grid > 0 form run gemm_kernel above).  Each workgroup owns one output tile
// and walks its K steps (across the problem's segments) with TWO register sets of staged
// operands: while the MFMAs consume K step s from LDS, the loads of steps s + 1 and s + 2
// are both in flight, so a K step has two steps' worth of MFMA time to hide its memory
// latency instead of one.  The register sets alternate with the two LDS buffers (the loop
// is unrolled by two so both sets stay in registers); one barrier per K step.
//
// Operand staging.  A quad is 4 elements moved by one thread.  When k is contiguous in
// memory the quad is 4 consecutive k of one row (m of A, n of B), fetched with one 16-byte
// load and stored k-contiguous in LDS ([row][k]); each lane then reads its 4 k values of an
// MFMA step with one 16-byte LDS read.  When the row index is contiguous instead (A
// transposed, B not transposed), the FP32 path fetches 4 consecutive rows at one k with one
// 16-byte load and keeps LDS row-major in k ([k][row], row stride BM/BN + 4 so a lane group
// of 16 rows at two k lands on all 32 banks); each lane reads its 4 k values with four
// 4-byte LDS reads.  Either way one vector load moves 16 bytes, so the load issue rate
// does not bound the K loop.  (BF16x3 keeps four 4-byte loads for row-contiguous operands:
// its MFMA reads 8 consecutive k per lane.)
template <int PREC, bool TA, bool TB, int BM, int BN>
struct PipeSmem {
  using T = typename Traits<PREC>::T;
  static constexpr bool AMN = PREC == TN_PREC_FP32 && TA;
  static constexpr bool BMN = PREC == TN_PREC_FP32 && !TB;
  static constexpr int BK = Traits<PREC>::BK;
  static constexpr int LK = BK + Traits<PREC>::PAD;
  // [buf][row][k] for k-contiguous staging, [buf][k][row] for row-contiguous staging
  static constexpr int A0 = AMN ? BK : BM, A1 = AMN ? BM + 4 : LK;
  static constexpr int B0 = BMN ? BK : BN, B1 = BMN ? BN + 4 : LK;
  T a[2][A0][A1];
  T b[2][B0][B1];
};
template <bool TA, bool TB, int BM, int BN>
struct PipeSmem<TN_PREC_BF16X3, TA, TB, BM, BN> : Smem<TN_PREC_BF16X3, BM, BN> {
  static constexpr bool AMN = false;
  static constexpr bool BMN = false;
};

template <int PREC, bool TA, bool TB, int BM, int BN, typename AT = typename Traits<PREC>::T, bool AROWS = false>
struct Stage {
  using T = typename Traits<PREC>::T;
  using V4 = T __attribute__((ext_vector_type(4)));
  using V4A = AT __attribute__((ext_vector_type(4)));
  static_assert(!AROWS || !TA, "a row table gathers k-contiguous rows of A");
  static_assert(sizeof(AT) == sizeof(T) || !TA, "a wider A than the arithmetic is read k-contiguous");
  using SM_ = PipeSmem<PREC, TA, TB, BM, BN>;
  static constexpr bool AMN = SM_::AMN, BMN = SM_::BMN;
  static constexpr int BK = Traits<PREC>::BK;
  static constexpr int QK = BK / 4;  // k quads per row of a K step
  static constexpr int QA = (BM * QK + kThreads - 1) / kThreads;
  static constexpr int QB = (BN * QK + kThreads - 1) / kThreads;
  static constexpr bool QA_EXACT = BM * QK % kThreads == 0;
  static constexpr bool QB_EXACT = BN * QK % kThreads == 0;
  struct Regs {
    V4A a[QA];
    V4 b[QB];
  };
  // per quad: byte offset in the segment, row (or first row) and k (or first k) in the
  // tile, and how many of its rows lie inside the problem (0: none, 4: all)
  uint32_t ao[QA], bo[QB];
  int am[QA], bn[QB];
  int akq[QA], bkq[QB];
  int arows[QA], brows[QB];
  int64_t abase[QA];  // AROWS: the quad's row start from the row table (elements)
  uint32_t astr, bstr;

  // rt: the problem's row table (AROWS), rt[m] = element offset of row m of A from the
  // segment's a_off; read once per tile, only for rows inside the problem
  __device__ __forceinline__ void positions(int tid, int m0, int n0, int M, int N,
                                            const int64_t* __restrict__ rt = nullptr) {
#pragma unroll
    for (int r = 0; r < QA; ++r) {
      const int q = r * kThreads + tid;
      int m, kq;
      if (AMN) { m = (q % (BM / 4)) * 4; kq = q / (BM / 4); }
      else if (!TA) { m = q / QK; kq = (q % QK) * 4; }
      else { m = q % BM; kq = (q / BM) * 4; }
      const bool live = QA_EXACT || q < BM * QK;
      if (!live) { m = 0; kq = 0; }
      am[r] = live ? m : -1;
      akq[r] = kq;
      const int in = M - (m0 + m);
      arows[r] = !live || in <= 0 ? 0 : (AMN ? (in < 4 ? in : 4) : 4);
      if constexpr (AROWS) abase[r] = arows[r] > 0 ? rt[m0 + m] : 0;
    }
#pragma unroll
    for (int r = 0; r < QB; ++r) {
      const int q = r * kThreads + tid;
      int n, kq;
      if (BMN) { n = (q % (BN / 4)) * 4; kq = q / (BN / 4); }
      else if (TB) { n = q / QK; kq = (q % QK) * 4; }
      else { n = q % BN; kq = (q / BN) * 4; }
      const bool live = QB_EXACT || q < BN * QK;
      if (!live) { n = 0; kq = 0; }
      bn[r] = live ? n : -1;
      bkq[r] = kq;
      const int in = N - (n0 + n);
      brows[r] = !live || in <= 0 ? 0 : (BMN ? (in < 4 ? in : 4) : 4);
    }
  }
  // byte offsets of every quad inside a segment (32-bit: the host refuses segments that
  // span 2 GiB or more); recomputed when the fetch cursor enters a new segment
  __device__ __forceinline__ void offsets(int m0, int n0, int64_t lda, int64_t ldb) {
    astr = static_cast<uint32_t>(lda * sizeof(T));
    bstr = static_cast<uint32_t>(ldb * sizeof(T));
#pragma unroll
    for (int r = 0; r < QA; ++r) {
      const int m = am[r] < 0 ? 0 : am[r];
      int64_t e;
      if constexpr (AROWS) e = abase[r] + akq[r];
      else e = TA ? (int64_t)akq[r] * lda + (m0 + m) : (int64_t)(m0 + m) * lda + akq[r];
      ao[r] = static_cast<uint32_t>(e * sizeof(AT));
    }
#pragma unroll
    for (int r = 0; r < QB; ++r) {
      const int n = bn[r] < 0 ? 0 : bn[r];
      const int64_t e = TB ? (int64_t)(n0 + n) * ldb + bkq[r] : (int64_t)bkq[r] * ldb + (n0 + n);
      bo[r] = static_cast<uint32_t>(e * sizeof(T));
    }
  }
  template <typename E>
  static __device__ __forceinline__ E load1(__amdgpu_buffer_rsrc_t d, uint32_t off) {
    if constexpr (sizeof(E) == 8)
      return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b64(d, off, 0, 0));
    else
      return __builtin_bit_cast(E, __builtin_amdgcn_raw_buffer_load_b32(d, off, 0, 0));
  }
  // 4 consecutive elements in memory (4 k of one row, or 4 rows at one k); `valid` of them
  // lie inside the problem.  Elements outside read through an offset beyond the
  // descriptor's range, which the hardware returns as zero.
  template <typename E>
  static __device__ __forceinline__ E __attribute__((ext_vector_type(4)))
  contig(__amdgpu_buffer_rsrc_t d, uint32_t off, int valid) {
    typedef E V4E __attribute__((ext_vector_type(4)));
    V4E v;
    if (valid >= 4) {
      if constexpr (sizeof(E) == 8) {
        typedef double d2 __attribute__((ext_vector_type(2)));
        const d2 lo = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(d, off, 0, 0));
        const d2 hi = __builtin_bit_cast(d2, __builtin_amdgcn_raw_buffer_load_b128(d, off + 16, 0, 0));
        v = V4E{lo[0], lo[1], hi[0], hi[1]};
      } else {
        v = __builtin_bit_cast(V4E, __builtin_amdgcn_raw_buffer_load_b128(d, off, 0, 0));
      }
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = load1<E>(d, j < valid ? off + j * (uint32_t)sizeof(E) : 0x80000000u);
    }
    return v;
  }
  // 4 k values `str` bytes apart (lanes run along m / n, so each load is coalesced)
  static __device__ __forceinline__ V4 strided(__amdgpu_buffer_rsrc_t d, uint32_t off, uint32_t str, bool ok,
                                               int kq, int kleft) {
    V4 v;
#pragma unroll
    for (int j = 0; j < 4; ++j) v[j] = load1<T>(d, (ok && kq + j < kleft) ? off + j * str : 0x80000000u);
    return v;
  }
  __device__ __forceinline__ void fetch(Regs& rg, const AT* Ap, const T* Bp, int kleft) const {
    const __amdgpu_buffer_rsrc_t da = __builtin_amdgcn_make_buffer_rsrc((void*)Ap, (short)0, 0x7fffffff, 0x00020000);
    const __amdgpu_buffer_rsrc_t db = __builtin_amdgcn_make_buffer_rsrc((void*)Bp, (short)0, 0x7fffffff, 0x00020000);
#pragma unroll
    for (int r = 0; r < QA; ++r) {
      if constexpr (AMN) {
        rg.a[r] = contig<AT>(da, ao[r], akq[r] < kleft ? arows[r] : 0);
      } else if constexpr (TA) {
        rg.a[r] = strided(da, ao[r], astr, arows[r] > 0, akq[r], kleft);
      } else {
        const int kv = kleft - akq[r];
        rg.a[r] = contig<AT>(da, ao[r], arows[r] > 0 ? (kv < 4 ? (kv < 0 ? 0 : kv) : 4) : 0);
      }
    }
#pragma unroll
    for (int r = 0; r < QB; ++r) {
      if (BMN) {
        rg.b[r] = contig<T>(db, bo[r], bkq[r] < kleft ? brows[r] : 0);
      } else if (!TB) {
        rg.b[r] = strided(db, bo[r], bstr, brows[r] > 0, bkq[r], kleft);
      } else {
        const int kv = kleft - bkq[r];
        rg.b[r] = contig<T>(db, bo[r], brows[r] > 0 ? (kv < 4 ? (kv < 0 ? 0 : kv) : 4) : 0);
      }
    }
  }
  __device__ __forceinline__ void stash(const Regs& rg, SM_& sm, int buf) const {
#pragma unroll
    for (int r = 0; r < QA; ++r) {
      if (!QA_EXACT && am[r] < 0) continue;
      if constexpr (PREC == TN_PREC_BF16X3) {
        typedef unsigned short u4 __attribute__((ext_vector_type(4)));
        u4 h, l;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float x = static_cast<float>(rg.a[r][j]);
          h[j] = bf16_rne(x);
          l[j] = bf16_rne(x - bf16_to_f32(h[j]));
        }
        *reinterpret_cast<u4*>(&sm.ah[buf][am[r]][akq[r]]) = h;
        *reinterpret_cast<u4*>(&sm.al[buf][am[r]][akq[r]]) = l;
      } else if constexpr (AMN) {
        *static_cast<V4*>(__builtin_assume_aligned(&sm.a[buf][akq[r]][am[r]], 16)) = rg.a[r];
      } else {
        // an FP64 A is rounded to the arithmetic's FP32 here, after its loads have landed
        *static_cast<V4*>(__builtin_assume_aligned(&sm.a[buf][am[r]][akq[r]], 16)) =
            __builtin_convertvector(rg.a[r], V4);
      }
    }
#pragma unroll
    for (int r = 0; r < QB; ++r) {
      if (!QB_EXACT && bn[r] < 0) continue;
      if constexpr (PREC == TN_PREC_BF16X3) {
        typedef unsigned short u4 __attribute__((ext_vector_type(4)));
        u4 h, l;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          h[j] = bf16_rne(rg.b[r][j]);
          l[j] = bf16_rne(rg.b[r][j] - bf16_to_f32(h[j]));
        }
        *reinterpret_cast<u4*>(&sm.bh[buf][bn[r]][bkq[r]]) = h;
        *reinterpret_cast<u4*>(&sm.bl[buf][bn[r]][bkq[r]]) = l;
      } else if constexpr (BMN) {
        *static_cast<V4*>(__builtin_assume_aligned(&sm.b[buf][bkq[r]][bn[r]], 16)) = rg.b[r];
      } else {
        *static_cast<V4*>(__builtin_assume_aligned(&sm.b[buf][bn[r]][bkq[r]], 16)) = rg.b[r];
      }
    }
  }
};

// The MFMAs of one K step on LDS buffer `buf` for a wave owning rows wm0.. and columns wn0..
template <int PREC, bool TA, bool TB, int BM, int BN, int SM, int SN>
__device__ __forceinline__ void mfma_step(const PipeSmem<PREC, TA, TB, BM, BN>& sm, int buf, int wm0, int wn0,
                                          int g, int r16, f32x4 (&acc)[SM][SN]) {
  using SMT = PipeSmem<PREC, TA, TB, BM, BN>;
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
  } else {
    constexpr int BK = Traits<PREC>::BK;
#pragma unroll
    for (int h = 0; h < BK / 16; ++h) {
      const int kq = h * 16 + g * 4;
      f32x4 a4[SM], b4[SN];
#pragma unroll
      for (int i = 0; i < SM; ++i) {
        const int row = wm0 + i * 16 + r16;
        if constexpr (SMT::AMN) {
#pragma unroll
          for (int tt = 0; tt < 4; ++tt) a4[i][tt] = sm.a[buf][kq + tt][row];
        } else {
          a4[i] = *reinterpret_cast<const f32x4*>(&sm.a[buf][row][kq]);
        }
      }
#pragma unroll
      for (int j = 0; j < SN; ++j) {
        const int col = wn0 + j * 16 + r16;
        if constexpr (SMT::BMN) {
#pragma unroll
          for (int tt = 0; tt < 4; ++tt) b4[j][tt] = sm.b[buf][kq + tt][col];
        } else {
          b4[j] = *reinterpret_cast<const f32x4*>(&sm.b[buf][col][kq]);
        }
      }
#pragma unroll
      for (int tt = 0; tt < 4; ++tt)
#pragma unroll
        for (int i = 0; i < SM; ++i)
#pragma unroll
          for (int j = 0; j < SN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[i][tt], b4[j][tt], acc[i][j], 0, 0, 0);
    }
  }
}

template <int PREC, bool TA, bool TB, int BM, int BN, typename AT = float, bool AROWS = false>
__global__ __launch_bounds__(kThreads, TN_PIPE_WAVES_PER_EU) void gemm_pipe_kernel(
    const AT* __restrict__ A, const float* __restrict__ B, float* __restrict__ C, float* __restrict__ W,
    const int64_t* __restrict__ probs, const int64_t* __restrict__ segs,
    const int32_t* __restrict__ tiles, int ntiles, const int64_t* __restrict__ arow_tab = nullptr) {
  static_assert(PREC != TN_PREC_FP64, "FP64 runs gemm_kernel");
  using St = Stage<PREC, TA, TB, BM, BN, AT, AROWS>;
  constexpr int BK = Traits<PREC>::BK;
  constexpr int WM = BM / 2, WN = BN / 2;
  constexpr int SM = WM / 16, SN = WN / 16;
  __shared__ __attribute__((aligned(16))) PipeSmem<PREC, TA, TB, BM, BN> sm;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm0 = (wave >> 1) * WM;
  const int wn0 = (wave & 1) * WN;
  const int g = lane >> 4;
  const int r16 = lane & 15;

  // workgroup b runs on XCD b % 8: hand XCD x the table slice [x G/8, (x+1) G/8)
  const int G = gridDim.x;
  int t = blockIdx.x;
  if ((G & 7) == 0) t = (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3);
  if (t >= ntiles) return;
  // tile word 3 holds the problem's first segment, so the problem record and that segment
  // are fetched side by side rather than one after the other
  const int p = tiles[4 * t];
  const int64_t* pw = probs + static_cast<int64_t>(p) * TN_PROB_WORDS;
  const int m0 = tiles[4 * t + 1] * BM, n0 = tiles[4 * t + 2] * BN;
  int64_t fs = tiles[4 * t + 3];
  const int M = static_cast<int>(pw[0]), N = static_cast<int>(pw[1]);
  const int64_t c_off = pw[2], ldc = pw[3];
  const int64_t se = pw[5];
  float* __restrict__ out = pw[6] ? W : C;

  f32x4 acc[SM][SN];
#pragma unroll
  for (int i = 0; i < SM; ++i)
#pragma unroll
    for (int j = 0; j < SN; ++j) acc[i][j] = f32x4{0, 0, 0, 0};

  St st;
  // problem word 7 (AROWS): where the problem's rows start in the row table
  st.positions(tid, m0, n0, M, N, AROWS ? arow_tab + pw[7] : nullptr);
  // fetch cursor: segment fs, k offset fk inside it (runs two K steps ahead of the MFMAs)
  int fk = 0;
  Seg fsg{};
  if (fs < se) {
    fsg = load_seg(segs, fs);
    st.offsets(m0, n0, fsg.lda, fsg.ldb);
  }
  auto issue = [&](typename St::Regs& rg) -> bool {
    // skip exhausted (or empty) segments; false once the problem's K range is done
    while (fs < se && fk >= fsg.K) {
      ++fs;
      fk = 0;
      if (fs < se) {
        fsg = load_seg(segs, fs);
        st.offsets(m0, n0, fsg.lda, fsg.ldb);
      }
    }
    if (fs >= se) return false;
    const AT* Ap = A + fsg.a_off + (TA ? (int64_t)fk * fsg.lda : fk);
    const float* Bp = B + fsg.b_off + (TB ? fk : (int64_t)fk * fsg.ldb);
    st.fetch(rg, Ap, Bp, fsg.K - fk);
    fk += BK;
    return true;
  };
  typename St::Regs r0, r1;
  bool h0 = issue(r0);
  bool h1 = h0 && issue(r1);
  while (h0) {
    st.stash(r0, sm, 0);
    __syncthreads();
    h0 = issue(r0);
    mfma_step<PREC, TA, TB, BM, BN, SM, SN>(sm, 0, wm0, wn0, g, r16, acc);
    if (!h1) break;
    st.stash(r1, sm, 1);
    __syncthreads();
    h1 = issue(r1);
    mfma_step<PREC, TA, TB, BM, BN, SM, SN>(sm, 1, wm0, wn0, g, r16, acc);
  }
  // ---- epilogue: straight from the accumulators ----
#pragma unroll
  for (int i = 0; i < SM; ++i)
#pragma unroll
    for (int j = 0; j < SN; ++j) {
      const int col = n0 + wn0 + j * 16 + r16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm0 + i * 16 + g * 4 + r;
        if (row < M && col < N) out[c_off + row * ldc + col] = acc[i][j][r];
      }
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
    if constexpr (PREC == TN_PREC_FP64) {                                                   \
      gemm_kernel<PREC, TA, TB, BM_, BN_><<<g, kThreads, 0, s>>>(a, b, c, w, probs, segs,   \
                                                                 tiles, ntiles);           \
    } else {                                                                                \
      if (grid > 0)                                                                         \
        gemm_kernel<PREC, TA, TB, BM_, BN_><<<g, kThreads, 0, s>>>(a, b, c, w, probs, segs, \
                                                                   tiles, ntiles);         \
      else                                                                                  \
        gemm_pipe_kernel<PREC, TA, TB, BM_, BN_><<<g, kThreads, 0, s>>>(a, b, c, w, probs,  \
                                                                        segs, tiles,        \
                                                                        ntiles);            \
    }                                                                                       \
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

// A gathered through a row table (problem word 7 = the problem's first entry in it), read
// in FP64 or FP32, B row-major (not transposed); FP32 or BF16x3 arithmetic, one workgroup
// per tile.
template <int PREC, typename AT>
int launch_rows(int bm, int bn, const void* A, const int64_t* rows, const void* B, void* C, void* W,
                const int64_t* probs, const int64_t* segs, const int32_t* tiles, int ntiles, hipStream_t s) {
  const AT* a = static_cast<const AT*>(A);
  const float* b = static_cast<const float*>(B);
  float* c = static_cast<float*>(C);
  float* w = static_cast<float*>(W);
#define TN_ROWS_CASE(BM_, BN_)                                                                  \
  if (bm == BM_ && bn == BN_) {                                                                 \
    gemm_pipe_kernel<PREC, false, false, BM_, BN_, AT, true><<<ntiles, kThreads, 0, s>>>(      \
        a, b, c, w, probs, segs, tiles, ntiles, rows);                                          \
    return tn::check_launch("tn_gemm_grouped_rows");                                           \
  }
  TN_ROWS_CASE(64, 64)
  TN_ROWS_CASE(128, 64)
  TN_ROWS_CASE(64, 128)
  TN_ROWS_CASE(128, 128)
#undef TN_ROWS_CASE
  return tn::fail("tn_gemm_grouped_rows: unsupported tile shape");
}

// ---- skinny problems: one output line per thread ----------------------------------------
// A problem whose output has at most kSkinnyMax columns (ROWS) or rows (!ROWS) and a short
// K is pure data movement: one thread owns one output row (column), keeps its <= 8
// accumulators in registers and streams the K values of its A row (B column); the small
// operand is uniform across the workgroup and comes through the scalar cache.  Tile table:
// 2 int32 per tile: (problem, block of 256 lines).
constexpr int kSkinnyMax = 8;

// One line (output row or column) of a skinny problem with exactly NW <= 8 entries across:
// every load of a k step is issued unconditionally (a load under a per-entry branch is
// waited for before the next one is issued), four k steps unrolled.
template <typename T, bool TA, bool TB, bool ROWS, int NW>
__device__ __forceinline__ void skinny_line(const T* __restrict__ A, const T* __restrict__ B, T* __restrict__ out,
                                            const int64_t* __restrict__ pw, const int64_t* __restrict__ segs,
                                            int64_t line, int64_t c_off, int64_t ldc) {
  T acc[NW];
#pragma unroll
  for (int j = 0; j < NW; ++j) acc[j] = T(0);
  for (int64_t s = pw[4]; s < pw[5]; ++s) {
    const int64_t* sw = segs + s * TN_SEG_WORDS;
    const T* a = A + sw[0];
    const T* b = B + sw[1];
    const int64_t K = sw[2], lda = sw[3], ldb = sw[4];
#pragma unroll 4
    for (int64_t k = 0; k < K; ++k) {
      T w[NW];
      if (ROWS) {  // line = row m; j runs over the NW columns
        const T av = TA ? a[k * lda + line] : a[line * lda + k];
#pragma unroll
        for (int j = 0; j < NW; ++j) w[j] = TB ? b[j * ldb + k] : b[k * ldb + j];
#pragma unroll
        for (int j = 0; j < NW; ++j) acc[j] += av * w[j];
      } else {  // line = column n; j runs over the NW rows
        const T bv = TB ? b[line * ldb + k] : b[k * ldb + line];
#pragma unroll
        for (int j = 0; j < NW; ++j) w[j] = TA ? a[k * lda + j] : a[j * lda + k];
#pragma unroll
        for (int j = 0; j < NW; ++j) acc[j] += w[j] * bv;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NW; ++j) out[c_off + (ROWS ? line * ldc + j : j * ldc + line)] = acc[j];
}

template <typename T, bool TA, bool TB, bool ROWS>
__global__ __launch_bounds__(kThreads) void gemm_skinny_kernel(
    const T* __restrict__ A, const T* __restrict__ B, T* __restrict__ C, T* __restrict__ W,
    const int64_t* __restrict__ probs, const int64_t* __restrict__ segs, const int32_t* __restrict__ tiles) {
  const int p = tiles[2 * blockIdx.x];
  const int64_t line = static_cast<int64_t>(tiles[2 * blockIdx.x + 1]) * kThreads + threadIdx.x;
  const int64_t* pw = probs + static_cast<int64_t>(p) * TN_PROB_WORDS;
  const int64_t M = pw[0], N = pw[1], c_off = pw[2], ldc = pw[3];
  const int64_t nl = ROWS ? M : N;   // lines: output rows (ROWS) or columns
  const int nw = static_cast<int>(ROWS ? N : M);  // width of a line, 1..kSkinnyMax
  if (line >= nl) return;
  T* __restrict__ out = pw[6] ? W : C;
  switch (nw) {  // uniform over the workgroup: one problem per tile
#define TN_SKINNY_NW(n) \
  case n:               \
    skinny_line<T, TA, TB, ROWS, n>(A, B, out, pw, segs, line, c_off, ldc); \
    break;
    TN_SKINNY_NW(1) TN_SKINNY_NW(2) TN_SKINNY_NW(3) TN_SKINNY_NW(4)
    TN_SKINNY_NW(5) TN_SKINNY_NW(6) TN_SKINNY_NW(7) TN_SKINNY_NW(8)
#undef TN_SKINNY_NW
    default:
      break;  // the host never plans a skinny problem outside 1..kSkinnyMax
  }
}

template <typename T, bool ROWS>
int launch_skinny(int ta, int tb, const void* A, const void* B, void* C, void* W, const int64_t* probs,
                  const int64_t* segs, const int32_t* tiles, int ntiles, hipStream_t s) {
  const T* a = static_cast<const T*>(A);
  const T* b = static_cast<const T*>(B);
  T* c = static_cast<T*>(C);
  T* w = static_cast<T*>(W);
#define TN_SKINNY_CASE(TA_, TB_)                                                                   \
  if (ta == TA_ && tb == TB_) {                                                                    \
    gemm_skinny_kernel<T, TA_, TB_, ROWS><<<ntiles, kThreads, 0, s>>>(a, b, c, w, probs, segs, tiles); \
    return tn::check_launch("tn_gemm_skinny");                                                     \
  }
  TN_SKINNY_CASE(0, 0)
  TN_SKINNY_CASE(1, 0)
  TN_SKINNY_CASE(0, 1)
  TN_SKINNY_CASE(1, 1)
#undef TN_SKINNY_CASE
  return tn::fail("tn_gemm_skinny: bad transpose flags");
}

}  // namespace

extern "C" int tn_gemm_skinny(int dtype, int trans_a, int trans_b, int rows, const void* A,
                              const void* B, void* C, void* W, const int64_t* problems,
                              const int64_t* segs, const int32_t* tiles, int64_t ntiles,
                              void* stream) {
  if (ntiles <= 0) return 0;
  if (ntiles > 0x7fffffffLL) return tn::fail("tn_gemm_skinny: too many tiles");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nt = static_cast<int>(ntiles);
  const int ta = trans_a ? 1 : 0, tb = trans_b ? 1 : 0;
  if (dtype == TN_F64)
    return rows ? launch_skinny<double, true>(ta, tb, A, B, C, W, problems, segs, tiles, nt, s)
                : launch_skinny<double, false>(ta, tb, A, B, C, W, problems, segs, tiles, nt, s);
  if (dtype == TN_F32)
    return rows ? launch_skinny<float, true>(ta, tb, A, B, C, W, problems, segs, tiles, nt, s)
                : launch_skinny<float, false>(ta, tb, A, B, C, W, problems, segs, tiles, nt, s);
  return tn::fail("tn_gemm_skinny: unknown dtype");
}

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

extern "C" int tn_gemm_grouped_rows(int precision, int a_dtype, const void* A, const int64_t* a_rows,
                                    const void* B, void* C, void* W, const int64_t* problems,
                                    const int64_t* segs, const int32_t* tiles, int64_t ntiles, int bm,
                                    int bn, void* stream) {
  if (ntiles <= 0) return 0;
  if (ntiles > 0x7fffffffLL) return tn::fail("tn_gemm_grouped_rows: too many tiles");
  if (a_rows == nullptr) return tn::fail("tn_gemm_grouped_rows: no row table");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nt = static_cast<int>(ntiles);
  if (precision == TN_PREC_FP32) {
    if (a_dtype == TN_F64) return launch_rows<TN_PREC_FP32, double>(bm, bn, A, a_rows, B, C, W, problems, segs, tiles, nt, s);
    if (a_dtype == TN_F32) return launch_rows<TN_PREC_FP32, float>(bm, bn, A, a_rows, B, C, W, problems, segs, tiles, nt, s);
  } else if (precision == TN_PREC_BF16X3) {
    if (a_dtype == TN_F64) return launch_rows<TN_PREC_BF16X3, double>(bm, bn, A, a_rows, B, C, W, problems, segs, tiles, nt, s);
    if (a_dtype == TN_F32) return launch_rows<TN_PREC_BF16X3, float>(bm, bn, A, a_rows, B, C, W, problems, segs, tiles, nt, s);
  } else {
    return tn::fail("tn_gemm_grouped_rows: FP32 or BF16x3 arithmetic only");
  }
  return tn::fail("tn_gemm_grouped_rows: A must be FP64 or FP32");
}

// ---- the entry point that allocates nothing ------------------------------------------
// The split-K partial products of tn_gemm_grouped_static go to this file-scope array
// instead of a caller-provided workspace.  Each device the library runs on holds its own
// copy until the process ends; host code reaches it only through the per-device address
// returned by hipGetSymbolAddress (the array's name is a host-side stand-in there).
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return tn::fail("tn_gemm_grouped_static: no current device");
  if (dev < 0 || dev >= kMaxDevices) return tn::fail("tn_gemm_grouped_static: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      return tn::fail(std::string("tn_gemm_grouped_static: hipGetSymbolAddress(scratch) failed: ") +
                      hipGetErrorString(e));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" int64_t tn_static_scratch_bytes(void) { return static_cast<int64_t>(kScratchBytes); }

extern "C" int tn_gemm_grouped_static(int precision, int dtype, int trans_a, int trans_b,
                                      const void* A, const void* B, void* C,
                                      const int64_t* problems, const int64_t* segs,
                                      const int32_t* tiles, int64_t ntiles, int bm, int bn, int grid,
                                      int64_t ws_elems, const int64_t* red_jobs,
                                      const int64_t* red_src, const double* red_coeff,
                                      const int64_t* red_tiles, int64_t red_ntiles, void* stream) {
  if (ws_elems < 0) return tn::fail("tn_gemm_grouped_static: negative workspace size");
  const int64_t item = dtype == TN_F64 ? 8 : 4;
  if (dtype != TN_F64 && dtype != TN_F32) return tn::fail("tn_gemm_grouped_static: unknown dtype");
  if (ws_elems > static_cast<int64_t>(kScratchBytes) / item)
    return tn::fail("tn_gemm_grouped_static: the split-K workspace needs " + std::to_string(ws_elems * item) +
                    " bytes, more than the library's static scratch of " + std::to_string(kScratchBytes) +
                    " bytes; use tn_gemm_grouped with a caller-provided workspace");
  if (red_ntiles > 0 && ws_elems == 0)
    return tn::fail("tn_gemm_grouped_static: a reduction needs a workspace");
  void* w = nullptr;
  if (int rc = scratch_for_current_device(&w)) return rc;
  if (int rc = tn_gemm_grouped(precision, dtype, trans_a, trans_b, A, B, C, w, problems, segs, tiles,
                               ntiles, bm, bn, grid, stream))
    return rc;
  if (red_ntiles <= 0) return 0;
  return tn_recouple(dtype, w, C, red_jobs, red_src, red_coeff, red_tiles, red_ntiles, stream);
}

// Host-side plan compilation: the job table of one tn_recouple launch from its destination
// regions, in one pass over the regions (no device code here).
//
// A region is a dense box of the destination, dims in destination C order, each with a
// destination and a source stride, plus a list of terms (source offset, coefficient).  Per
// region: drop unit dims and merge each dim into the previous kept one when the two are
// contiguous in both source and destination (a region with no dim left keeps one unit
// dim); pick the kernel's mode (a 64 x 64 LDS transpose tile over the last dim and the
// first src-contiguous dim when the last dim is dst- but not src-contiguous, else the
// linear path); check every access against the buffers; count the region's tiles.  A
// program whose tiles would leave most of the device idle switches to the small tiles
// (32 x 32 LDS tiles, 512-element linear tiles).

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/tenet_amd.h"

namespace tn {
void set_error(const std::string& msg);
}

namespace {

constexpr int kMaxInRank = 64;

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace

extern "C" int tn_recouple_plan(int64_t n, int64_t rank, const int64_t* dst_off, const int64_t* shape,
                                const int64_t* dst_strides, const int64_t* src_strides,
                                const int64_t* term_ptr, const int64_t* term_src, int64_t src_size,
                                int64_t dst_size, int64_t lin_tile, int64_t lin_small_tile, int64_t lds_tile,
                                int64_t small_program_tiles, int64_t* jobs, int64_t* tile_counts,
                                int64_t* stats) {
  if (n < 0 || rank < 0 || rank > kMaxInRank) {
    tn::set_error("tn_recouple_plan: region rank out of range");
    return 4;
  }
  if (lin_tile <= 0 || lin_small_tile <= 0 || lds_tile <= 1) {
    tn::set_error("tn_recouple_plan: tile sizes must be positive");
    return 5;
  }
  const int64_t W = TN_MAX_RANK;
  int64_t elements = 0, moved = 0, total = 0;
  int64_t on[kMaxInRank], od[kMaxInRank], os[kMaxInRank];
  for (int64_t j = 0; j < n; ++j) {
    const int64_t* shp = shape + j * rank;
    const int64_t* ds = dst_strides + j * rank;
    const int64_t* ss = src_strides + j * rank;
    int r = 0;
    bool has = false;
    int64_t cn = 0, cd = 0, cs = 0;
    for (int64_t k = 0; k < rank; ++k) {
      const int64_t nk = shp[k], dk = ds[k], sk = ss[k];
      if (nk == 1) continue;
      if (has && cd == dk * nk && cs == sk * nk) {
        cn *= nk;
        cd = dk;
        cs = sk;
        continue;
      }
      if (has) {
        on[r] = cn, od[r] = cd, os[r] = cs;
        ++r;
      }
      cn = nk, cd = dk, cs = sk;
      has = true;
    }
    if (has) {
      on[r] = cn, od[r] = cd, os[r] = cs;
      ++r;
    }
    if (r == 0) {
      on[0] = od[0] = os[0] = 1;
      r = 1;
    }
    if (r > W) {
      tn::set_error("recouple region of rank " + std::to_string(r) + " exceeds " + std::to_string(W));
      return 4;
    }
    int64_t nelem = 1, dspan = 0, span = 0;
    for (int d = 0; d < r; ++d) {
      nelem *= on[d];
      dspan += (on[d] - 1) * od[d];
      span += (on[d] - 1) * os[d];
    }
    int64_t mode = -1;
    const int last = r - 1;
    if (os[last] != 1 && od[last] == 1 && on[last] >= 16) {
      for (int d = 0; d < last; ++d) {
        if (os[d] == 1 && on[d] >= 16) {
          mode = d;
          break;
        }
      }
    }
    // every access stays in bounds: checked here once, on the host
    if (dst_off[j] < 0 || dst_off[j] + dspan >= dst_size) {
      tn::set_error("recouple region writes outside its destination buffer");
      return 1;
    }
    // the kernel addresses elements within a region with 32-bit offsets
    if ((span > dspan ? span : dspan) >= (int64_t(1) << 31)) {
      tn::set_error("recouple region spans 2^31 elements or more");
      return 2;
    }
    const int64_t t0 = term_ptr[j], t1 = term_ptr[j + 1];
    for (int64_t t = t0; t < t1; ++t) {
      if (term_src[t] < 0 || term_src[t] + span >= src_size) {
        tn::set_error("recouple term reads outside its source buffer");
        return 3;
      }
    }
    int64_t* row = jobs + j * TN_JOB_WORDS;
    row[0] = dst_off[j];
    row[1] = t0;
    row[2] = t1;
    row[3] = r;
    for (int d = 0; d < W; ++d) {
      row[4 + d] = d < r ? on[d] : 0;
      row[4 + W + d] = d < r ? od[d] : 0;
      row[4 + 2 * W + d] = d < r ? os[d] : 0;
    }
    row[22] = mode;
    row[23] = nelem;
    elements += nelem;
    moved += nelem * (t1 - t0);
    int64_t c;
    if (mode < 0) {
      c = ceil_div(nelem, lin_tile);
    } else {
      c = ceil_div(on[last], lds_tile) * ceil_div(on[mode], lds_tile) * (nelem / (on[last] * on[mode]));
    }
    tile_counts[j] = c;
    total += c;
  }
  if (total < small_program_tiles) {
    const int64_t e = lds_tile / 2;
    total = 0;
    for (int64_t j = 0; j < n; ++j) {
      int64_t* row = jobs + j * TN_JOB_WORDS;
      const int64_t mode = row[22], r = row[3], nelem = row[23];
      int64_t c;
      if (mode >= 0) {
        const int64_t nl = row[4 + r - 1], nm = row[4 + mode];
        row[22] = mode + 8;
        c = ceil_div(nl, e) * ceil_div(nm, e) * (nelem / (nl * nm));
      } else {
        row[22] = -2;
        c = ceil_div(nelem, lin_small_tile);
      }
      tile_counts[j] = c;
      total += c;
    }
  }
  stats[0] = elements;
  stats[1] = moved;
  stats[2] = total;
  return 0;
}

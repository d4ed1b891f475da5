// Gradient / hessian histograms over segmented row lists (gfx950).
//
// A histogram build is described by a position list (perm: position -> row
// id) cut into SEGMENTS (start, count, output slot).  A segment is one
// (node, era) run of positions; every row of a segment lands in histogram
// slot `slot`, laid out [slot][n_feat][n_bins] for G and for H (the
// reference's [E, F, B] per node, warpgbm/cuda/histogram_kernel.cu:51-60).
//
// Work decomposition (one wave = 64 lanes):
//   * a FEATURE SLAB is 64*FPL consecutive local features; lane l owns
//     features [l*FPL, l*FPL+FPL) of the slab and loads their FPL code bytes
//     of a row with one load (1..16 bytes; byte gathers when the features
//     are not the identity columns);
//   * a CHUNK is up to C consecutive positions of one segment; a workgroup
//     of W waves takes a work item (chunk, slab), splits the chunk into W
//     contiguous sub-ranges, one per wave, and every wave adds its rows into
//     ONE shared LDS histogram of the item;
//   * a segment's chunk partials are summed by k_hist_merge.
// Sums are exact 64-bit fixed point (each row adds 2^44 + round(g * 2^sg)
// with one ds_add_u64, the row count riding in the top bits, and
// round(h * 2^sh) into a second plane unless the node's hessian is
// constant), so the result is independent of scheduling, chunking and the
// order of atomics: plain per-node builds and batched level builds of the
// same rows are bit-identical, and every bin is rounded to float32 once.
//
// LDS layout of a plane: [FPL][B + 1][64] words, word (j*(B+1) + bin)*64 +
// lane.  Lane l only touches its own column, so the 64 lanes of every add
// hit 64 different banks whatever bins they carry; bin B is a spare row
// that absorbs codes outside [0, B).
#include "common.h"
#include "wgbm_internal.h"
#include "wgbm_stats.h"
#include "wgbm_plan.h"
#include <hip/hip_ext.h>
#include <algorithm>
#include <mutex>
#include <type_traits>
#include <utility>
#include <vector>

namespace wgbm {

template <int FPL> struct CodeVec;
template <> struct CodeVec<1>  { using T = uint8_t; };
template <> struct CodeVec<2>  { using T = uint16_t; };
template <> struct CodeVec<4>  { using T = uint32_t; };
template <> struct CodeVec<8>  { using T = uint64_t; };
template <> struct CodeVec<16> { using T = uint4; };

template <int FPL>
__device__ __forceinline__ uint32_t code_u8(const typename CodeVec<FPL>::T& v, int j) {
  if constexpr (FPL == 16) {
    const uint32_t w = j < 4 ? v.x : j < 8 ? v.y : j < 12 ? v.z : v.w;
    return (w >> (8 * (j & 3))) & 0xffu;
  } else {
    return (uint32_t)((uint64_t)v >> (8 * j)) & 0xffu;
  }
}

template <int FPL>
__device__ __forceinline__ int code_byte(const typename CodeVec<FPL>::T& v, int j) {
  if constexpr (FPL == 16) {
    uint32_t w = j < 4 ? v.x : j < 8 ? v.y : j < 12 ? v.z : v.w;
    return (int)(int8_t)(uint8_t)(w >> (8 * (j & 3)));
  } else {
    return (int)(int8_t)(uint8_t)((uint64_t)v >> (8 * j));
  }
}

// ---------------------------------------------------------------------------
// Plan: chunk counts per segment (exclusive scan) for the histogram grid.
// seg_first[s] = first global chunk of segment s, seg_first[nseg] = total.
// A segment with count 0 still owns one (empty) chunk so that its slot is
// written; count < 0 marks a padding segment that owns no chunk.
// Segments with more than one chunk are listed in `multi` for the merge.
// Row count of segment s for the three kinds of plan.
__global__ void __launch_bounds__(1024)
k_hist_plan(const PlanSrc src, int32_t chunk_rows, int32_t* __restrict__ seg_first,
            int32_t* __restrict__ seg_pfirst, int32_t* __restrict__ multi, int32_t* __restrict__ n_multi) {
  plan_body(src, chunk_rows, seg_first, seg_pfirst, multi, n_multi);
}


// ---------------------------------------------------------------------------
// Shifts of a segment's node and the partial word of one bin.
struct NodeShift {
  int sg, sh, sf;   // gradient, hessian and fine-gradient shifts
  bool count_mode;  // every row of the node has the same hessian
  bool packed;      // gradient words carry the row count (34-bit gradient, fine rows)
  float h_const;    // that hessian (used by the one-plane variant)
};

// one_plane: the node is built by the one-plane variant, whose words always
// carry the count.  Otherwise the words carry it exactly when the node is
// in count mode, so both variants encode a count-mode node alike.
__device__ __forceinline__ NodeShift node_shift(const int32_t* st, int h_bits, int fine_x, bool one_plane) {
  NodeShift n;
  const float gmax = __builtin_bit_cast(float, st[0]);
  const float hmin = ord_float(st[1]);
  const float habs = __builtin_bit_cast(float, st[3]);
  n.count_mode = st[1] == st[2];
  n.packed = one_plane || n.count_mode;
  n.sg = fixed_shift(gmax, n.packed ? kGBitsCount : kGBitsWide);
  n.sf = n.sg + fine_x;
  n.sh = fixed_shift(habs, h_bits);
  n.h_const = hmin;
  return n;
}

// Gradient sum and row count of a gradient word (count 0 when the word
// carries none).
__device__ __forceinline__ void decode_word(unsigned long long w, bool packed, long long& gsum, long long& count) {
  if (packed) {
    gsum = ((long long)(w << (64 - kCountShift))) >> (64 - kCountShift);
    count = (long long)((w - (unsigned long long)gsum) >> kCountShift);
  } else {
    gsum = (long long)w;
    count = 0;
  }
}

// A bin's float32 gradient: the coarse sum (scale 2^-sg) plus, when a fine
// row landed in the bin, the fine sum (scale 2^-sf); one rounding to float32
// after at most two in float64.
__device__ __forceinline__ float bin_grad(const NodeShift& ns, double coarse, long long fine) {
  double v = ldexp(coarse, -ns.sg);
  if (fine != 0) v += ldexp((double)fine, -ns.sf);
  return (float)v;
}

// Fine slot of a histogram slot (see HistArgs::fine).
__device__ __forceinline__ int fine_slot(const HistArgs& a, int slot) {
  if (!a.sub_g) return slot;
  const int E = a.slots_per_node;
  const int c = slot / E;
  return (c >> 1) * E + (slot - c * E);
}

// Takes a fine word back to zero and returns it.  No fine row adds to the
// word while it is taken (the adds of its slot came before a barrier or an
// earlier launch), so most words -- those no fine row reached -- are only
// read; a non-zero one is exchanged for zero.
__device__ __forceinline__ long long take_fine(const HistArgs& a, int fslot, int f, int bin) {
  unsigned long long* p = reinterpret_cast<unsigned long long*>(a.fine) +
                          ((int64_t)fslot * a.n_feat + f) * a.n_bins + bin;
  const unsigned long long v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (v == 0ull) return 0;
  return (long long)__hip_atomic_exchange(p, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A bin's float32 hessian: count * h in count mode (exact in float64, one
// rounding; no fixed-point product that could wrap for a bin of many rows),
// else the fixed-point sum.
__device__ __forceinline__ float bin_hess(const NodeShift& ns, bool count_mode, long long x) {
  return count_mode ? (float)((double)x * (double)ns.h_const) : from_fixed(x, ns.sh);
}

// Sibling derivation: the parent's bin (loaded ahead of the result, so its
// latency overlaps the build's own loads) and where the sibling's bin goes.
struct SibBin {
  float pg = 0.0f, ph = 0.0f;
  int64_t ds = -1;  // -1: no sibling to write
};

__device__ __forceinline__ SibBin load_sib(const HistArgs& a, int slot, int f, int bin) {
  SibBin b;
  if (a.sub_g) {
    const int E = a.slots_per_node;
    const int c = slot / E, e = slot - c * E;
    const int64_t dp = ((int64_t)((c >> 1) * E + e) * a.n_feat + f) * a.n_bins + bin;
    b.ds = ((int64_t)((c ^ 1) * E + e) * a.n_feat + f) * a.n_bins + bin;
    b.pg = a.sub_g[dp];
    b.ph = a.sub_h[dp];
  }
  return b;
}

// Store one finished bin of slot `slot` and, when the launch derives
// siblings, the sibling's bin = parent - this (float32).
__device__ __forceinline__ void store_bin(const HistArgs& a, int slot, int f, int bin, float fg, float fh,
                                          const SibBin& sb) {
  const int64_t d = ((int64_t)slot * a.n_feat + f) * a.n_bins + bin;
  if (a.accumulate) { a.out_g[d] += fg; a.out_h[d] += fh; }
  else { a.out_g[d] = fg; a.out_h[d] = fh; }
  if (sb.ds >= 0) {
    a.out_g[sb.ds] = sb.pg - fg;
    a.out_h[sb.ds] = sb.ph - fh;
  }
}

constexpr int kDirectStage = 8;  // words per thread of a staged single-chunk read-out

// ONE_PLANE: only the T plane is allocated and H = count * h, launched when
// the caller guarantees one hessian value for every row (a.hess_const);
// otherwise the two-plane variant builds every node, or with a.dual both
// variants run over the same grid and each workgroup builds only the nodes
// whose statistics select its variant.
// WIDE: 16-wave workgroups, two per CU: every SIMD then runs 8 waves (the
// second bound is waves per execution unit), so the kernel must fit in 64
// VGPRs.
template <int FPL, int L, bool GH_POS, bool ONE_PLANE, bool WIDE>
__global__ void __launch_bounds__(WIDE ? 1024 : 512, WIDE ? 8 : 1)
k_hist_build(const HistArgs a) {
  extern __shared__ unsigned long long s64[];
  const int lane = lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int W = blockDim.x >> 6;
  const int total = *a.chunk_total;  // chunks of this launch (the plan's count)
  // Work items b = (chunk, slab), taken round-robin by a grid of at most
  // one workgroup per slot (gridDim.x a multiple of 8).  XCD-aware mapping:
  // dispatch sends block b to XCD b % 8, and b % 8 is fixed per workgroup,
  // so all slabs of a chunk run on the same XCD, back to back, and the rows
  // of a chunk are read from HBM once into that XCD's L2.  chunk grows with
  // b, so the first item past the last chunk ends the loop.
  for (int b = blockIdx.x;; b += gridDim.x) {
  const int i = b >> 3;
  const int slab = i % a.nslab;
  const int chunk = ((i / a.nslab) << 3) + (b & 7);
  if (chunk >= total) break;
  // The plan's descriptor of the chunk: positions [c0, c1), output slot and
  // partial index (-1: the segment has this one chunk; write the result).
  const int4 dsc = a.chunk_desc[chunk];
  const int c0 = dsc.x, c1 = dsc.y, slot = dsc.z, pidx = dsc.w;
  const int len = max(c1 - c0, 0);
  const int per = (len + W - 1) / W;
  const int w0 = min(c0 + wave * per, c1);
  const int w1 = min(w0 + per, c1);
  // The wave's first row ids (and position-ordered g, h) are loaded before
  // the node statistics and the LDS clear, so the three overlap.
  const bool has_rows = a.n_rows > 0 && w0 < w1;
  int first_raw = 0;
  float2 first_gh = make_float2(0.0f, 0.0f);
  if (has_rows) {
    first_raw = a.perm[min(w0 + lane, w1 - 1)];
    if constexpr (GH_POS) first_gh = a.gh[min(w0 + lane, w1 - 1)];
  }
  const NodeShift ns = node_shift(a.stats + (int64_t)(slot / a.slots_per_node) * kStatWords, a.h_bits, a.fine_x,
                                  ONE_PLANE);
  if (a.dual && ONE_PLANE != ns.count_mode) continue;  // the other variant builds this node
  const bool packed = ONE_PLANE || ns.packed;
  const int fslot = fine_slot(a, slot);
  __shared__ int s_fine;  // some wave added a fine row to this item's sums
  if (threadIdx.x == 0) s_fine = 0;

  // One shared histogram per workgroup: T plane (count << 44 | sum qg) and,
  // unless the hessian is constant, an H plane (sum qh).  A plane holds
  // [FPL][B + 1][64] words: feature lane*FPL + j, bin b, lane at
  // ((j * (B + 1) + b) * 64 + lane); bin B is a spare row that absorbs codes
  // outside [0, B), so the adds need no branch.  Each lane owns its own
  // column, so the 64 lanes of an add hit 64 different banks.
  const int B = a.n_bins;
  const int plane = FPL * (B + 1) * kWave;
  unsigned long long* T = s64;
  unsigned long long* Hq = s64 + plane;
  for (int q = threadIdx.x; q < (ONE_PLANE ? 1 : 2) * plane; q += blockDim.x) s64[q] = 0ull;
  __syncthreads();

  const int fl0 = slab * kWave * FPL + lane * FPL;  // local feature of byte 0
  const bool lane_on = fl0 < a.n_feat;
  // Wide loads need the slab's features to be the identity columns.
  // Features past n_feat (and a lane with none) add into their own lane's
  // columns, which the read-out skips.
  bool ident = a.wide_ok;
  uint32_t cols[FPL];  // byte offsets of the lane's columns in a row
  bool jon[FPL];
#pragma unroll
  for (int j = 0; j < FPL; ++j) {
    const int fl = fl0 + j;
    int c = (fl < a.n_feat) ? (a.feat_idx ? a.feat_idx[fl] : fl) : 0;
    // A column outside the matrix contributes nothing (and is never read).
    jon[j] = fl < a.n_feat && (uint64_t)c < (uint64_t)a.row_stride;
    if (!jon[j]) c = 0;
    cols[j] = (uint32_t)c;
    if (fl < a.n_feat && c != fl) ident = false;
  }
  // Byte offset of (j, bin 0) in this lane's column; a column outside the
  // matrix points at its spare row so that it adds nothing readable.
  uint32_t jbase[FPL], jcap[FPL];
#pragma unroll
  for (int j = 0; j < FPL; ++j) {
    jbase[j] = ((uint32_t)(j * (B + 1) + (jon[j] ? 0 : B)) * kWave + lane) * 8u;
    jcap[j] = jon[j] ? (uint32_t)B : 0u;
  }
  ident = __all(ident || !lane_on);

  using TV = typename CodeVec<FPL>::T;
  const int8_t* __restrict__ codes = a.codes;
  const int64_t stride = a.row_stride;
  static_assert(L <= 32 && (kWave % L) == 0, "load block: an even number of groups per block");

  auto add_row = [&](const TV& v, unsigned long long qp, unsigned long long qh) {
#pragma unroll
    for (int j = 0; j < FPL; ++j) {
      // Codes are bytes; anything outside [0, B) (negative included) lands in
      // the spare row.  A column outside the matrix has jbase at its spare
      // row and a cap of 0.
      const uint32_t bin = min(code_u8<FPL>(v, j), jcap[j]);
      const uint32_t off = jbase[j] + bin * (uint32_t)(kWave * 8);
      unsigned long long* tp = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(T) + off);
      __hip_atomic_fetch_add(tp, qp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if constexpr (!ONE_PLANE) {
        unsigned long long* hp = reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(Hq) + off);
        __hip_atomic_fetch_add(hp, qh, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  };

  // Row id at position pn of this wave's range: the load is always issued
  // (at a clamped position) and checked only where the id is used, so that
  // it stays in flight meanwhile.  valid_row gives the id, or -1 outside the
  // range or the matrix.
  auto perm_at = [&](int pn) -> int { return a.perm[min(pn, w1 - 1)]; };
  auto valid_row = [&](int pn, int t) -> int {
    return (pn < w1 && (uint64_t)t < (uint64_t)a.n_rows) ? t : -1;
  };
  auto gh_at = [&](int pn) -> float2 { return a.gh[min(pn, w1 - 1)]; };

  // Software pipeline over the wave's positions in blocks of 64: while the
  // adds of one group of L rows run, the next group's code loads and the next
  // block's row indices (and position-ordered g, h) are in flight.  Every
  // load is issued unconditionally -- a lane without features reads the
  // row's first bytes, a position without a row reads row 0 and adds a zero
  // word -- so the compiler can count loads in flight and each group waits
  // only for its own rows.
  auto run = [&](auto ident_tag) {
    constexpr bool ID = decltype(ident_tag)::value;
    // A byte gather holds FPL registers per row in flight: groups of L/FPL rows.
    constexpr int LG = ID ? L : (L / FPL > 0 ? L / FPL : 1);
    constexpr int NG = kWave / LG;  // groups per 64-position block (even)
    const uint32_t colo = lane_on ? (uint32_t)fl0 : 0u;
    auto load_codes = [&](int row) -> TV {
      // The row's address is uniform: formed in scalar registers (a 64-bit
      // product on the scalar unit), so the load takes it as its scalar base
      // and the lane's column offset as its vector offset -- no vector
      // instruction per row for the address.
      const uint64_t ra = (uint64_t)(uintptr_t)codes + (uint64_t)(uint32_t)max(row, 0) * (uint64_t)stride;
      const int8_t* rp = reinterpret_cast<const int8_t*>(
          ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(ra >> 32)) << 32) |
          (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)ra));
      TV v;
      if constexpr (ID) {
        // global (not generic) address space: a global load with a scalar
        // base, not a flat load (which would also count as an LDS access)
        const uintptr_t a16 = reinterpret_cast<uintptr_t>(rp) + colo;
        if constexpr (FPL == 16) {
          using G4 = __attribute__((address_space(1))) const uint32_t;
          G4* p4 = reinterpret_cast<G4*>(a16);
          v = make_uint4(p4[0], p4[1], p4[2], p4[3]);
        } else {
          using GTV = __attribute__((address_space(1))) const TV;
          v = *reinterpret_cast<GTV*>(a16);
        }
      } else {
        using G8 = __attribute__((address_space(1))) const uint8_t;
        G8* gp = reinterpret_cast<G8*>(reinterpret_cast<uintptr_t>(rp));
        uint8_t tmp[FPL];
#pragma unroll
        for (int j = 0; j < FPL; ++j) tmp[j] = gp[cols[j]];
        __builtin_memcpy(&v, tmp, sizeof(TV));
      }
      return v;
    };
    // Fixed-point words of a block's rows (lane = position in the block);
    // a fine row's word carries only its count, and lane_fine records that
    // the lane met one (the fine pass after the loop adds their gradients).
    unsigned long long myqp, myqh;
    bool lane_fine = false;
    auto block_words = [&](int rowv_, float2 ghv_) {
      const bool okr = rowv_ >= 0;
      float g = ghv_.x, h = ghv_.y;
      if constexpr (!GH_POS) {
        const int r0 = max(rowv_, 0);
        g = a.g[r0]; h = a.h[r0];
      }
      const long long q = to_fixed(g, ns.sg);
      if (packed) {
        const bool fine = okr && g != 0.0f && q < kFineLimit && q > -kFineLimit;
        myqp = okr ? (1ull << kCountShift) + (fine ? 0ull : (unsigned long long)q) : 0ull;
        lane_fine = lane_fine || fine;
      } else {
        myqp = okr ? (unsigned long long)q : 0ull;
      }
      myqh = okr ? (unsigned long long)to_fixed(h, ns.sh) : 0ull;
    };
    // Fine rows of the wave's positions, after its main loop (they are
    // rare: count-mode rows with |g| below 2^-16 of the node's max): each
    // lane adds a fine row's gradient at the fine scale to the global words
    // of its features' bins (codes outside [0, B) and columns outside the
    // matrix add nothing).
    auto fine_pass = [&]() {
      if (lane == 0) s_fine = 1;
      for (int bpos = w0; bpos < w1; bpos += kWave) {
        const int pos = bpos + lane;
        const int row = valid_row(pos, perm_at(pos));
        const float g = GH_POS ? gh_at(pos).x : a.g[max(row, 0)];
        const long long q = to_fixed(g, ns.sg);
        unsigned long long m = __ballot(row >= 0 && g != 0.0f && q < kFineLimit && q > -kFineLimit);
        while (m) {
          const int r = __builtin_ctzll(m);
          m &= m - 1ull;
          const int rr = __builtin_amdgcn_readlane(row, r);
          const float gr = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, g), r));
          const unsigned long long qf = (unsigned long long)to_fixed(gr, ns.sf);
          const TV v = load_codes(rr);
#pragma unroll
          for (int j = 0; j < FPL; ++j) {
            const uint32_t bin = code_u8<FPL>(v, j);
            if (bin < jcap[j]) {  // a column outside the matrix has a cap of 0
              unsigned long long* p = reinterpret_cast<unsigned long long*>(a.fine) +
                                      ((int64_t)fslot * a.n_feat + fl0 + j) * B + bin;
              __hip_atomic_fetch_add(p, qf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
          }
        }
      }
      // The adds are complete at L2 before this wave reaches the barrier
      // that precedes the read-out (which takes the words back to zero).
      __builtin_amdgcn_s_waitcnt(0);
    };
    int rowv = valid_row(w0 + lane, first_raw);
    block_words(rowv, first_gh);
    TV A[LG], Bv[LG];
#pragma unroll
    for (int q = 0; q < LG; ++q) A[q] = load_codes(__builtin_amdgcn_readlane(rowv, q));
    // Loop invariant: rowv, myqp, myqh describe this block and its first
    // group's loads are in flight; nothing loaded is used at the loop head.
    for (int base = w0; base < w1; base += kWave) {
      // The next block's row ids (and g, h) go in flight now.
      const int pn = base + kWave + lane;
      const int rown_raw = perm_at(pn);
      const float2 ghn = GH_POS ? gh_at(pn) : make_float2(0.0f, 0.0f);
      const int cnt = min(kWave, w1 - base);
      auto add_group = [&](const TV* V, int g0) {
        if (g0 >= cnt) return;  // uniform: a partial last block
        // No per-row check: a position past the wave's range carries a zero
        // word (and row 0's codes), so its adds change nothing.
#pragma unroll
        for (int q = 0; q < LG; ++q) {
          const int r = g0 + q;
          const unsigned long long qp =
              ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(myqp >> 32), r) << 32) |
              (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)myqp, r);
          unsigned long long qh = 0ull;
          if constexpr (!ONE_PLANE)
            qh = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(myqh >> 32), r) << 32) |
                 (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)myqh, r);
          add_row(V[q], qp, qh);
        }
      };
      // Pairs of groups; the loop stays rolled (a kernel that fits the
      // instruction cache), the last pair is peeled: its second load fetches
      // the next block's first group.
#pragma unroll 1
      for (int gq = 0; gq < NG - 2; gq += 2) {
#pragma unroll
        for (int q = 0; q < LG; ++q) Bv[q] = load_codes(__builtin_amdgcn_readlane(rowv, (gq + 1) * LG + q));
        add_group(A, gq * LG);
#pragma unroll
        for (int q = 0; q < LG; ++q) A[q] = load_codes(__builtin_amdgcn_readlane(rowv, (gq + 2) * LG + q));
        add_group(Bv, (gq + 1) * LG);
      }
#pragma unroll
      for (int q = 0; q < LG; ++q) Bv[q] = load_codes(__builtin_amdgcn_readlane(rowv, (NG - 1) * LG + q));
      add_group(A, (NG - 2) * LG);
      const int rown = valid_row(pn, rown_raw);
#pragma unroll
      for (int q = 0; q < LG; ++q) A[q] = load_codes(__builtin_amdgcn_readlane(rown, q));
      add_group(Bv, (NG - 1) * LG);
      rowv = rown;
      block_words(rowv, ghn);
    }
    if (__ballot(lane_fine)) fine_pass();
  };
  if (has_rows) {
    if (ident) run(std::true_type{});
    else run(std::false_type{});
  }
  __syncthreads();

  // Decode and write the chunk result.  Thread t reads word t (lane-fastest:
  // conflict free) = feature lane*FPL + j, bin, with (j, bin) = divmod(t/64, B).
  // Chunk partials are stored in this word order (consecutive threads write
  // consecutive words); k_hist_merge maps words to (feature, bin).
  const int SL = kWave * FPL * B;
  const int f_base = slab * kWave * FPL;
  const bool direct = pidx < 0;
  const bool item_fine = s_fine != 0;  // fine words to take in (the row loop ended at a barrier)
  constexpr int kFlushBatch = 4;  // words per thread whose sibling loads are in flight together
  if (!WIDE && direct && !item_fine && SL <= kDirectStage * (int)blockDim.x) {
    // A single-chunk segment's float32 bins go to [feature][bin] of its slot,
    // where word order would put consecutive lanes FPL * B floats apart.  The
    // decoded bins are staged in the freed LDS plane in output order, then
    // the result, the parent's bins and the sibling's are read and written
    // contiguously.  (Not in the 16-wave variant: its 64 VGPRs are kept for
    // the row loop, and its launches have few single-chunk segments.)
    float rg[kDirectStage], rh[kDirectStage];
#pragma unroll
    for (int k = 0; k < kDirectStage; ++k) {
      const int t = threadIdx.x + k * blockDim.x;
      rg[k] = 0.0f;
      rh[k] = 0.0f;
      if (t < SL) {
        const int j = (t >> 6) / B;
        const int word = t + j * kWave;  // skip the spare rows of features 0..j-1
        const unsigned long long w = T[word];
        long long gsum, count;
        decode_word(w, packed, gsum, count);
        rg[k] = bin_grad(ns, (double)gsum, 0);
        rh[k] = bin_hess(ns, ONE_PLANE, ONE_PLANE ? count : (long long)Hq[word]);
      }
    }
    __syncthreads();  // every word is read before the plane is reused
    float* og = reinterpret_cast<float*>(s64);  // [64 * FPL][B] of the slab, then the hessians
    float* oh = og + SL;
#pragma unroll
    for (int k = 0; k < kDirectStage; ++k) {
      const int t = threadIdx.x + k * blockDim.x;
      if (t < SL) {
        const int r = t >> 6;
        const int j = r / B;
        const int o = ((t & 63) * FPL + j) * B + (r - j * B);
        og[o] = rg[k];
        oh[o] = rh[k];
      }
    }
    __syncthreads();
    const int n_out = min(kWave * FPL, a.n_feat - f_base) * B;
    const int64_t d0 = ((int64_t)slot * a.n_feat + f_base) * B;
    int64_t p0 = -1, s0 = -1;
    if (a.sub_g) {
      const int E = a.slots_per_node;
      const int c = slot / E, e = slot - c * E;
      p0 = ((int64_t)((c >> 1) * E + e) * a.n_feat + f_base) * B;
      s0 = ((int64_t)((c ^ 1) * E + e) * a.n_feat + f_base) * B;
    }
    for (int o0 = threadIdx.x; o0 < n_out; o0 += kFlushBatch * blockDim.x) {
      float pg[kFlushBatch], ph[kFlushBatch];
      if (p0 >= 0) {
#pragma unroll
        for (int u = 0; u < kFlushBatch; ++u) {
          const int o = min(o0 + u * (int)blockDim.x, n_out - 1);
          pg[u] = a.sub_g[p0 + o];
          ph[u] = a.sub_h[p0 + o];
        }
      }
#pragma unroll
      for (int u = 0; u < kFlushBatch; ++u) {
        const int o = o0 + u * blockDim.x;
        if (o >= n_out) break;
        const float fg = og[o], fh = oh[o];
        if (a.accumulate) { a.out_g[d0 + o] += fg; a.out_h[d0 + o] += fh; }
        else { a.out_g[d0 + o] = fg; a.out_h[d0 + o] = fh; }
        if (p0 >= 0) {
          a.out_g[s0 + o] = pg[u] - fg;
          a.out_h[s0 + o] = ph[u] - fh;
        }
      }
    }
    __syncthreads();  // the read-out is done before the next item clears the LDS
    continue;
  }
  if (direct && item_fine) {
    // A single-chunk segment that took fine rows: every bin adds its fine word.
    for (int t = threadIdx.x; t < SL; t += blockDim.x) {
      const int r = t >> 6;
      const int j = r / B;
      const int bin = r - j * B;
      const int fl = (t & 63) * FPL + j;
      if (f_base + fl >= a.n_feat) continue;
      const SibBin sb = a.sub_g ? load_sib(a, slot, f_base + fl, bin) : SibBin{};
      const int word = t + j * kWave;
      long long gsum, count;
      decode_word(T[word], packed, gsum, count);
      const long long hsum = ONE_PLANE ? count : (long long)Hq[word];
      const long long fw = take_fine(a, fslot, f_base + fl, bin);
      store_bin(a, slot, f_base + fl, bin, bin_grad(ns, (double)gsum, fw), bin_hess(ns, ONE_PLANE, hsum), sb);
    }
    __syncthreads();  // the read-out is done before the next item clears the LDS
    continue;
  }
  for (int t0 = threadIdx.x; t0 < SL; t0 += kFlushBatch * blockDim.x) {
    SibBin sb[kFlushBatch];
    if (direct && a.sub_g) {
#pragma unroll
      for (int u = 0; u < kFlushBatch; ++u) {
        const int t = t0 + u * blockDim.x;
        const int r = t >> 6;
        const int j = r / B;
        const int fl = (t & 63) * FPL + j;
        if (t < SL && f_base + fl < a.n_feat) sb[u] = load_sib(a, slot, f_base + fl, r - j * B);
      }
    }
#pragma unroll
    for (int u = 0; u < kFlushBatch; ++u) {
      const int t = t0 + u * blockDim.x;
      if (t >= SL) break;
      const int r = t >> 6;
      const int ln = t & 63;
      const int j = r / B;
      const int bin = r - j * B;
      const int fl = ln * FPL + j;
      if (f_base + fl >= a.n_feat) continue;
      const int word = t + j * kWave;  // skip the spare rows of features 0..j-1
      const unsigned long long w = T[word];
      long long gsum, count;
      decode_word(w, packed, gsum, count);
      const long long hsum = ONE_PLANE ? count : (long long)Hq[word];  // count, or the fixed-point sum
      if (direct) {
        store_bin(a, slot, f_base + fl, bin, bin_grad(ns, (double)gsum, 0), bin_hess(ns, ONE_PLANE, hsum), sb[u]);
      } else {
        const int64_t d = ((int64_t)pidx * a.nslab + slab) * SL + t;
        // Partials are read once, by the merge: streaming stores keep them
        // out of this XCD's L2 (less to write back at the end of the launch).
        if constexpr (ONE_PLANE) {
          __builtin_nontemporal_store(w, reinterpret_cast<unsigned long long*>(a.part) + d);  // packed count | qg sum
        } else {
          __builtin_nontemporal_store(gsum, &a.part[d].x);
          __builtin_nontemporal_store(hsum, &a.part[d].y);
        }
      }
    }
  }
  if (!direct && threadIdx.x == 0) a.pflag[(int64_t)pidx * a.nslab + slab] = item_fine ? 1 : 0;
  __syncthreads();  // the read-out is done before the next item clears the LDS
  }
}

// Sum the chunk partials of every multi-chunk segment (exact integers) and
// write the float32 histogram.  A workgroup of 256 threads takes a work item
// (segment, slab, block of 256 / P words): P thread groups share a word, group
// p summing chunks p, p + P, ... with eight loads in flight per thread, and
// reduce through LDS.  P (1, 2, 4 or 8) comes from the plan: the fewest
// groups that keep each group at or below 16 chunks (at most 8).  Partials
// are in the build's word order: word e = (j, bin, lane) with
// (j, bin) = divmod(e / 64, B), feature lane*FPL + j.
constexpr int kMergeThreads = 256;
constexpr int kMergeMaxParts = 8;
constexpr int kMergeUnroll = 8;

struct MergeSum {
  I128 g;          // gradient sum
  long long x = 0; // count (one plane) or hessian sum
  __device__ __forceinline__ void add_packed(unsigned long long w) {
    long long gs, c;
    decode_word(w, true, gs, c);
    g.add(gs);
    x += c;
  }
  __device__ __forceinline__ void add_pair(longlong2 v) { g.add(v.x); x += v.y; }
};

template <int FPL>
__global__ void __launch_bounds__(kMergeThreads)
k_hist_merge(const HistArgs a, const int32_t* __restrict__ n_multi) {
  __shared__ unsigned long long s_glo[kMergeThreads];
  __shared__ long long s_ghi[kMergeThreads];
  __shared__ long long s_h[kMergeThreads];
  const int P = *a.merge_parts;
  const int WPI = kMergeThreads / P;  // words per item
  const int B = a.n_bins;
  const int SL = kWave * FPL * B;
  const int neb = (SL + WPI - 1) / WPI;
  const int64_t n_items = (int64_t)(*n_multi) * a.nslab * neb;
  const int el = threadIdx.x % WPI;
  const int part = threadIdx.x / WPI;
  // Consecutive items cover consecutive (j, bin) rows of one slab, so eight
  // of them together write (and, with siblings, read the parent's) whole
  // lines of the [feature][bin] output.  Virtual item vi runs on workgroup
  // vi % gridDim.x, whose XCD is vi % 8 (gridDim.x is a multiple of 8); vi maps
  // to items so that each group of eight consecutive items shares one XCD
  // and runs side by side: those lines then meet in one L2.
  const int64_t n_virtual = (n_items + 63) & ~(int64_t)63;
  for (int64_t vi = blockIdx.x; vi < n_virtual; vi += gridDim.x) {
    const int64_t q = vi >> 3;
    const int64_t it = (((q >> 3) << 6) | ((vi & 7) << 3)) | (q & 7);
    if (it >= n_items) continue;  // uniform
    const int m = (int)(it / (a.nslab * neb));
    const int r = (int)(it - (int64_t)m * a.nslab * neb);
    const int slab = r / neb;
    const int e = (r - slab * neb) * WPI + el;
    const int4 md = a.multi_desc[m];
    const int first = md.x, nch = md.y, slot = md.z;
    NodeShift ns = node_shift(a.stats + (int64_t)(slot / a.slots_per_node) * kStatWords, a.h_bits, a.fine_x, false);
    const bool one_plane = a.hess_const || (a.dual && ns.count_mode);  // the variant that built the partials
    if (one_plane && !ns.packed) ns = node_shift(a.stats + (int64_t)(slot / a.slots_per_node) * kStatWords,
                                                 a.h_bits, a.fine_x, true);
    const int rr = e >> 6;
    const int j = rr / B;
    const int bin = rr - j * B;
    const int fl = (e & 63) * FPL + j;
    const int f_base = slab * kWave * FPL;
    SibBin sb;
    if (part == 0 && e < SL && f_base + fl < a.n_feat) sb = load_sib(a, slot, f_base + fl, bin);
    MergeSum acc;
    if (e < SL) {
      const int64_t off = ((int64_t)first * a.nslab + slab) * SL + e;
      const int64_t step = (int64_t)a.nslab * SL;
      // kMergeUnroll chunks P apart at once; a chunk past the last one
      // re-reads the last chunk and is dropped.
      if (one_plane) {
        const unsigned long long* p = reinterpret_cast<const unsigned long long*>(a.part) + off;
        for (int k = part; k < nch; k += kMergeUnroll * P) {
          unsigned long long w[kMergeUnroll];
#pragma unroll
          for (int u = 0; u < kMergeUnroll; ++u) w[u] = p[(int64_t)min(k + u * P, nch - 1) * step];
#pragma unroll
          for (int u = 0; u < kMergeUnroll; ++u)
            if (k + u * P < nch) acc.add_packed(w[u]);
        }
      } else {
        const longlong2* p = a.part + off;
        for (int k = part; k < nch; k += kMergeUnroll * P) {
          longlong2 v[kMergeUnroll];
#pragma unroll
          for (int u = 0; u < kMergeUnroll; ++u) v[u] = p[(int64_t)min(k + u * P, nch - 1) * step];
#pragma unroll
          for (int u = 0; u < kMergeUnroll; ++u)
            if (k + u * P < nch) acc.add_pair(v[u]);
        }
      }
    }
    if (P > 1) {
      s_glo[threadIdx.x] = acc.g.lo;
      s_ghi[threadIdx.x] = acc.g.hi;
      s_h[threadIdx.x] = acc.x;
    }
    // Fine rows in any of the segment's chunks (count-mode nodes only); the
    // barrier also publishes the groups' sums.
    int flagged = 0;
    if (ns.packed)
      for (int k = threadIdx.x; k < nch; k += kMergeThreads) flagged |= a.pflag[(int64_t)(first + k) * a.nslab + slab];
    const bool any_fine = __syncthreads_or(flagged) != 0;
    if (part == 0 && e < SL) {
      I128 sg = acc.g;
      long long sx = acc.x;
      for (int gp = 1; gp < P; ++gp) {
        I128 o;
        o.lo = s_glo[gp * WPI + el];
        o.hi = s_ghi[gp * WPI + el];
        sg.add(o);
        sx += s_h[gp * WPI + el];
      }
      if (f_base + fl < a.n_feat) {
        const long long fw = any_fine ? take_fine(a, fine_slot(a, slot), f_base + fl, bin) : 0;
        store_bin(a, slot, f_base + fl, bin, bin_grad(ns, sg.to_double(), fw), bin_hess(ns, one_plane, sx), sb);
      }
    }
    if (P > 1) __syncthreads();
  }
}

// Statistics of one row list (reference op): stats[0..3] over g[rows], h[rows].
int stats_blocks(int64_t n) {
  const int64_t b = ceil_div(n, 1024);
  return (int)(b < 1 ? 1 : b > kStatBlocks ? kStatBlocks : b);
}

// Partial statistics of a row list (gathered g, h): one per block.
__global__ void __launch_bounds__(256)
k_list_stats(const int32_t* __restrict__ rows, int n, int64_t n_rows, const float* __restrict__ g,
             const float* __restrict__ h, int32_t* __restrict__ part) {
  StatAcc acc;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
    const int r = rows[p];
    if ((uint64_t)r < (uint64_t)n_rows) acc.add(g[r], h[r]);
  }
  acc.block_reduce();
  if (threadIdx.x == 0) acc.store(part + (int64_t)blockIdx.x * kStatWords);
}

__global__ void __launch_bounds__(256)
k_gh_stats(const float2* __restrict__ gh, int64_t n, int32_t* __restrict__ part) {
  const int k = blockIdx.y;
  const float2* v = gh + (int64_t)k * n;
  StatAcc acc;
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const float2 x = v[p];
    acc.add(x.x, x.y);
  }
  acc.block_reduce();
  if (threadIdx.x == 0) acc.store(part + ((int64_t)k * gridDim.x + blockIdx.x) * kStatWords);
}

__global__ void __launch_bounds__(256)
k_stats_reduce(const int32_t* __restrict__ part, int n_part, int32_t* __restrict__ out) {
  const int32_t* p = part + (int64_t)blockIdx.x * n_part * kStatWords;
  StatAcc acc;
  for (int i = threadIdx.x; i < n_part; i += blockDim.x) acc.load_merge(p + (int64_t)i * kStatWords);
  acc.block_reduce();
  if (threadIdx.x == 0) acc.store(out + (int64_t)blockIdx.x * kStatWords);
}

// ---------------------------------------------------------------------------
// Stable counting sort of a row list by era (reference op with E > 1).
// Positions are cut into blocks of kSortBlock; era counts per block, an
// era-major exclusive scan, then an in-order scatter per block.  Rows whose
// era lies outside [0, E) are dropped (they contribute to no histogram).
constexpr int kSortBlock = 4096;

__global__ void __launch_bounds__(256)
k_era_count(const int32_t* __restrict__ rows, int n, int64_t n_rows,
            const int32_t* __restrict__ era, int E, int32_t* __restrict__ cnt) {
  extern __shared__ int32_t sc[];
  for (int e = threadIdx.x; e < E; e += blockDim.x) sc[e] = 0;
  __syncthreads();
  const int p0 = blockIdx.x * kSortBlock;
  const int p1 = min(p0 + kSortBlock, n);
  for (int p = p0 + threadIdx.x; p < p1; p += blockDim.x) {
    const int r = rows[p];
    if ((uint64_t)r >= (uint64_t)n_rows) continue;
    const int e = era[r];
    if ((unsigned)e < (unsigned)E) atomicAdd(&sc[e], 1);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < E; e += blockDim.x) cnt[(int64_t)blockIdx.x * E + e] = sc[e];
}

__global__ void __launch_bounds__(1024)
k_era_scan(int32_t* __restrict__ cnt, int nblocks, int E, int32_t* __restrict__ seg_start,
           int32_t* __restrict__ seg_count, int32_t* __restrict__ seg_slot) {
  __shared__ int32_t s_tot[1024];
  __shared__ int32_t s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  for (int base = 0; base < E; base += 1024) {
    const int e = base + threadIdx.x;
    int tot = 0;
    if (e < E)
      for (int c = 0; c < nblocks; ++c) tot += cnt[(int64_t)c * E + e];
    s_tot[threadIdx.x] = tot;
    __syncthreads();
    if (threadIdx.x == 0) {
      int run = s_carry;
      for (int k = 0; k < 1024; ++k) { const int v = s_tot[k]; s_tot[k] = run; run += v; }
      s_carry = run;
    }
    __syncthreads();
    if (e < E) {
      int run = s_tot[threadIdx.x];
      seg_start[e] = run;
      seg_count[e] = tot;
      seg_slot[e] = e;
      for (int c = 0; c < nblocks; ++c) {
        const int v = cnt[(int64_t)c * E + e];
        cnt[(int64_t)c * E + e] = run;
        run += v;
      }
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(64)
k_era_scatter(const int32_t* __restrict__ rows, int n, int64_t n_rows, const int32_t* __restrict__ era, int E,
              const int32_t* __restrict__ off, int32_t* __restrict__ out) {
  extern __shared__ int32_t run[];
  for (int e = threadIdx.x; e < E; e += 64) run[e] = off[(int64_t)blockIdx.x * E + e];
  __syncthreads();
  const int lane = threadIdx.x;
  const int p0 = blockIdx.x * kSortBlock;
  const int p1 = min(p0 + kSortBlock, n);
  for (int base = p0; base < p1; base += 64) {
    const int p = base + lane;
    int row = 0, e = -1;
    if (p < p1) {
      row = rows[p];
      e = (uint64_t)row < (uint64_t)n_rows ? era[row] : -1;
      if ((unsigned)e >= (unsigned)E) e = -1;
    }
    bool pending = e >= 0;
    while (true) {
      const uint64_t act = __ballot(pending);
      if (act == 0) break;
      const int leader = __builtin_ctzll(act);
      const int le = __shfl(e, leader);
      const uint64_t mask = __ballot(pending && e == le);
      if (pending && e == le) {
        out[run[le] + lanes_below(mask)] = row;
        pending = false;
      }
      __syncthreads();
      if (lane == leader) run[le] += __popcll(mask);
      __syncthreads();
    }
  }
}

__global__ void k_single_segment(int32_t* seg_start, int32_t* seg_count, int32_t* seg_slot,
                                 int32_t* nseg, int n) {
  if (threadIdx.x == 0) { seg_start[0] = 0; seg_count[0] = n; seg_slot[0] = 0; *nseg = 1; }
}

__global__ void k_set_scalar(int32_t* p, int v) {
  if (threadIdx.x == 0) *p = v;
}

void plan_launch(const PlanSrc& src, int chunk_rows, int32_t* seg_first, int32_t* multi, int32_t* n_multi,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_hist_plan, dim3(1), dim3(kPlanThreads), 0, st, src, chunk_rows, seg_first, (int32_t*)nullptr, multi,
                     n_multi);
}

void plan_job_launch(const PlanJob& j, hipStream_t st) {
  hipLaunchKernelGGL(k_hist_plan, dim3(1), dim3(kPlanThreads), 0, st, j.src, j.chunk_rows, j.seg_first, j.seg_pfirst,
                     j.multi, j.n_multi);
}


// ---------------------------------------------------------------------------
// Automatic configurations keep two workgroups' sub-histograms per CU
// (64 KB each); explicit ones may use a workgroup's full 160 KB.
constexpr size_t kLdsBudget = 64 * 1024;
// A workgroup's LDS is 160 KB; the build kernel's own static variables take
// a few bytes of it, so its dynamic histogram may use 160 KB less 1 KB.
constexpr size_t kLdsMax = 160 * 1024 - 1024;

int lds_pitch(const HistConfig&, int) { return 64; }

size_t lds_bytes(const HistConfig& c, int n_bins, bool count_mode) {
  return (size_t)(count_mode ? 1 : 2) * c.fpl * (n_bins + 1) * kWave * sizeof(unsigned long long);
}

static HistConfig g_override;   // all-zero: automatic
// Configurations installed for one histogram shape (the autotuner's
// choices); they take precedence over the process-wide override.
struct ShapeConfig {
  int n_feat, n_bins;
  int64_t n_cols;
  bool one_plane;
  HistConfig c;
};
static std::vector<ShapeConfig> g_shape_cfg;
static std::mutex g_cfg_mu;
int clamp_depth(int fpl, int waves, int depth);

void set_hist_override(const HistConfig& c) {
  std::lock_guard<std::mutex> lock(g_cfg_mu);
  g_override = c;
}

void set_hist_shape_config(int n_feat, int n_bins, int64_t n_cols, bool one_plane, const HistConfig& c) {
  std::lock_guard<std::mutex> lock(g_cfg_mu);
  for (size_t i = 0; i < g_shape_cfg.size(); ++i) {
    ShapeConfig& e = g_shape_cfg[i];
    if (e.n_feat == n_feat && e.n_bins == n_bins && e.n_cols == n_cols && e.one_plane == one_plane) {
      if (c.chunk_rows == 0 && c.waves == 0 && c.fpl == 0 && c.depth == 0 && c.slots == 0)
        g_shape_cfg.erase(g_shape_cfg.begin() + i);
      else e.c = c;
      return;
    }
  }
  if (c.chunk_rows || c.waves || c.fpl || c.depth || c.slots) g_shape_cfg.push_back({n_feat, n_bins, n_cols, one_plane, c});
}

static HistConfig apply_override(HistConfig c, const HistConfig& o, int64_t row_stride) {
  if (o.chunk_rows > 0) c.chunk_rows = o.chunk_rows < kMaxChunkRows ? o.chunk_rows : kMaxChunkRows;
  if (o.fpl > 0 && row_stride % o.fpl == 0) c.fpl = o.fpl;
  // 16-wave workgroups (two per CU, 64 VGPRs) for narrow lanes; wider lanes
  // need more registers than that.  Measured at 1M x 1000 x 32 bins, FPL 4:
  // 16 waves with 8 rows in flight build the root 10 % faster than 8 waves
  // with 16.
  c.waves = o.waves > 0 ? o.waves : (c.fpl <= 4 ? 16 : 8);
  if (c.fpl > 4 && c.waves > 8) c.waves = 8;
  c.depth = clamp_depth(c.fpl, c.waves, o.depth);
  c.slots = o.slots > 0 ? o.slots : kHistSlots;
  return c;
}

HistConfig pick_hist_config(int n_feat, int n_bins, int64_t row_stride, bool one_plane) {
  const HistConfig a = auto_hist_config(n_feat, n_bins, row_stride, one_plane);
  HistConfig o;
  {
    std::lock_guard<std::mutex> lock(g_cfg_mu);
    o = g_override;
    for (const ShapeConfig& e : g_shape_cfg)
      if (e.n_feat == n_feat && e.n_bins == n_bins && e.n_cols == row_stride && e.one_plane == one_plane) o = e.c;
  }
  const HistConfig c = apply_override(a, o, row_stride);
  // An override chosen for another shape may not fit this one's LDS (e.g.
  // 16 features per lane tuned for 5 bins, then 127 bins): the automatic
  // configuration builds it instead.
  if (lds_bytes(c, n_bins, one_plane) > kLdsMax) return apply_override(a, HistConfig{}, row_stride);
  return c;
}

HistConfig auto_hist_config(int n_feat, int n_bins, int64_t row_stride, bool one_plane) {
  HistConfig c;
  (void)one_plane;
  c.chunk_rows = 8192;
  c.waves = 8;
  // Per-lane load width: the widest whose one-plane histogram lets two
  // workgroups share a CU's 160 KB of LDS (the two-plane variant then fits
  // one) and whose slabs keep at least 3/4 of the lanes on a feature (an idle
  // lane's adds still take LDS cycles); with none, the best lane use.
  // Measured at 1M x 1000 x 32 bins: FPL 4 (67 KB) beats FPL 2 by 1.4x and
  // FPL 8 (135 KB, one workgroup) by 1.1x; at 500k x 400 x 5 bins FPL 8
  // (78 % of lanes) builds 1.28x faster than FPL 16 (39 %) and byte loads
  // (FPL 1, 89 %) are slower than both.
  const int F = n_feat > 0 ? n_feat : 1;
  int best = 1;
  double best_use = -1.0;
  for (int fpl = 16; fpl >= 1; fpl >>= 1) {
    c.fpl = fpl;
    if (fpl > 1 && (lds_bytes(c, n_bins, true) > 80 * 1024 || row_stride % fpl != 0)) continue;
    const int64_t lanes = ceil_div(F, kWave * fpl) * kWave * fpl;
    const double use = (double)F / (double)lanes;
    if (use >= 0.75) { best = fpl; break; }
    if (use > best_use) { best_use = use; best = fpl; }
  }
  c.fpl = best;
  return c;
}

static int64_t hist_chunk_target(const HistConfig& c, int n_feat) {
  const int64_t nslab = ceil_div(n_feat > 0 ? n_feat : 1, kWave * c.fpl);
  const int64_t t = hist_slots(c) / nslab;
  return t > 0 ? t : 1;
}

// With C >= ceil(rows / target) (or C = chunk_rows), rows / C <= cap.
static int64_t hist_chunk_cap(const HistConfig& c, int n_feat, int64_t rows_max) {
  const int64_t a = hist_chunk_target(c, n_feat), b = ceil_div(rows_max, c.chunk_rows);
  return a > b ? a : b;
}

int64_t hist_max_chunks(const HistConfig& c, int n_feat, int64_t rows_max, int64_t nseg_max) {
  return hist_chunk_cap(c, n_feat, rows_max) + nseg_max + 1;
}

int64_t hist_max_multi(const HistConfig& c, int n_feat, int64_t rows_max) {
  return hist_chunk_cap(c, n_feat, rows_max) + 1;  // each holds more than C rows
}

// Multi-chunk segments hold more than C rows each, so their chunks number at
// most 2 * rows / C <= 2 * cap.
static int64_t partial_chunks(const HistConfig& c, int n_feat, int64_t rows_max) {
  return 2 * hist_chunk_cap(c, n_feat, rows_max) + 2;
}

// Words of the partials: per chunk the configuration's slabs of 64 * FPL
// features (a launch with fewer features per lane pads less).
size_t hist_partial_words(const HistConfig& c, int n_feat, int n_bins, int64_t rows_max) {
  const int64_t padded = ceil_div(n_feat, kWave * c.fpl) * kWave * c.fpl;
  return (size_t)(partial_chunks(c, n_feat, rows_max) * padded * n_bins);
}

// Flags of the partials, [partial chunk][slab], for any features per lane
// (at most one slab per 64 features).
size_t hist_partial_flags(const HistConfig& c, int n_feat, int64_t rows_max) {
  return (size_t)(partial_chunks(c, n_feat, rows_max) * ceil_div(n_feat > 0 ? n_feat : 1, kWave));
}

// ---------------------------------------------------------------------------
// Live kernel timing: when enabled, every histogram build and merge launch
// carries a pair of HIP events on its stream, recorded by the launch itself
// (hipExtLaunchKernel: the kernel's own start and end, not the queue's
// packets around it).
struct KernelTimer {
  bool on = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> build, merge;
};
static KernelTimer g_timer;

bool hist_timer_on() { return g_timer.on; }

// A new event pair for one timed launch (false: timing is off).
static bool timer_pair(std::vector<std::pair<hipEvent_t, hipEvent_t>>& v, hipEvent_t& a, hipEvent_t& b) {
  a = nullptr;
  b = nullptr;
  if (!g_timer.on) return false;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    if (a) (void)hipEventDestroy(a);
    a = nullptr;
    b = nullptr;
    g_timer.on = false;  // timing is best effort: stop rather than record garbage
    return false;
  }
  v.emplace_back(a, b);
  return true;
}

static void timer_clear() {
  for (auto* v : {&g_timer.build, &g_timer.merge})
    for (auto& p : *v) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
  g_timer.build.clear();
  g_timer.merge.clear();
}

template <int FPL, int L, bool CNT, bool WIDE>
static int launch_waves(const HistArgs& a, const HistConfig& c, int64_t groups, hipStream_t st) {
  const size_t lds = lds_bytes(c, a.n_bins, CNT);
  // Opt in to more than 64 KB of dynamic LDS, once per size and device (a
  // second call may come while the stream is being captured into a graph).
  static size_t opted[64] = {};
  int dev = 0;
  WGBM_HIP(hipGetDevice(&dev));
  if (lds > 64 * 1024 && (dev < 0 || dev >= 64 || opted[dev] < lds)) {
    WGBM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_hist_build<FPL, L, true, CNT, WIDE>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsMax));
    WGBM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_hist_build<FPL, L, false, CNT, WIDE>),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsMax));
    if (dev >= 0 && dev < 64) opted[dev] = kLdsMax;
  }
  hipEvent_t e0, e1;
  if (timer_pair(g_timer.build, e0, e1)) {
    if (a.gh)
      hipExtLaunchKernelGGL((k_hist_build<FPL, L, true, CNT, WIDE>), dim3((unsigned)groups), dim3(64 * c.waves), lds,
                            st, e0, e1, 0, a);
    else
      hipExtLaunchKernelGGL((k_hist_build<FPL, L, false, CNT, WIDE>), dim3((unsigned)groups), dim3(64 * c.waves), lds,
                            st, e0, e1, 0, a);
  } else if (a.gh) {
    hipLaunchKernelGGL((k_hist_build<FPL, L, true, CNT, WIDE>), dim3((unsigned)groups), dim3(64 * c.waves), lds, st,
                       a);
  } else {
    hipLaunchKernelGGL((k_hist_build<FPL, L, false, CNT, WIDE>), dim3((unsigned)groups), dim3(64 * c.waves), lds, st,
                       a);
  }
  WGBM_LAUNCHED("k_hist_build");
  return kOk;
}

template <int FPL, int L, bool CNT>
static int launch_mode(const HistArgs& a, const HistConfig& c, int64_t groups, hipStream_t st) {
  // 16-wave workgroups exist for FPL <= 4 at the shallower depth only (the
  // configuration picks them so): they must fit 64 VGPRs.
  if constexpr (FPL <= 4 && L == (FPL == 1 ? 16 : 8)) {
    if (c.waves > 8) return launch_waves<FPL, L, CNT, true>(a, c, groups, st);
  }
  return launch_waves<FPL, L, CNT, false>(a, c, groups, st);
}

template <int FPL, int L>
static int launch_fpl(const HistArgs& a, const HistConfig& c, int64_t max_chunks, int64_t max_multi,
                      int32_t* multi, int32_t* n_multi, hipStream_t st) {
  // Work items (chunk, slab); a grid of at most hist_slots(c) workgroups
  // (a multiple of 8) loops over them.
  const int64_t items = ceil_div(max_chunks, 8) * 8 * a.nslab;
  if (items > 0x7fffffff) { set_error("histogram grid too large"); return kTooBig; }
  const int64_t groups = items < hist_slots(c) ? items : hist_slots(c);
  // One plane when the caller guarantees one hessian value for every row;
  // otherwise the two-plane variant builds every node.  For a node whose rows
  // happen to share one hessian h, the hessian plane sums count copies of
  // h * 2^sh = mantissa(h) * 2^(h_bits - 24), an integer for h_bits >= 24, so
  // the bin is count * h exactly, rounded once: the one-plane variant's value
  // bit for bit, without a second launch for such nodes.
  int rc;
  if (a.hess_const) {
    rc = launch_mode<FPL, L, true>(a, c, groups, st);
  } else {
    rc = a.dual ? launch_mode<FPL, L, true>(a, c, groups, st) : kOk;
    if (rc == kOk) rc = launch_mode<FPL, L, false>(a, c, groups, st);
  }
  if (rc != kOk) return rc;
  if (max_multi > 0) {
    const int SL = kWave * FPL * a.n_bins;
    const int64_t merge_items = max_multi * a.nslab * ceil_div(SL, kMergeThreads / kMergeMaxParts);
    const int64_t mg = std::min<int64_t>((merge_items + 7) & ~(int64_t)7, 2048);  // a multiple of 8
    hipEvent_t e0, e1;
    if (timer_pair(g_timer.merge, e0, e1))
      hipExtLaunchKernelGGL((k_hist_merge<FPL>), dim3((unsigned)mg), dim3(kMergeThreads), 0, st, e0, e1, 0, a,
                            n_multi);
    else
      hipLaunchKernelGGL((k_hist_merge<FPL>), dim3((unsigned)mg), dim3(kMergeThreads), 0, st, a, n_multi);
    WGBM_LAUNCHED("k_hist_merge");
  }
  return kOk;
}

// Load depth (rows in flight per wave) allowed per features-per-lane.
int clamp_depth(int fpl, int waves, int depth) {
  int hi = fpl == 1 ? 32 : fpl <= 4 ? 16 : fpl == 8 ? 8 : 4;
  const int lo = hi / 2;
  // 16-wave workgroups keep the shallower depth within their 64 VGPRs (at
  // FPL 2 the two-plane variant spilled with 16 rows per group)
  if (waves > 8) hi = lo;
  const int def = hi;
  int d = depth > 0 ? depth : def;
  if (d > hi) d = hi;
  if (d < lo) d = lo;
  int p = 1;
  while (p * 2 <= d) p *= 2;
  return p;
}

PlanJob hist_plan_job(const HistArgs& a, const HistConfig& c, int32_t* multi, int32_t* n_multi) {
  PlanJob j{};
  j.src.mode = kPlanSegments;
  j.src.seg_count = a.seg_count;
  j.src.nseg_dev = a.nseg_dev;
  j.src.chunk_out = a.chunk_dev;
  j.src.chunk_min = kChunkMin < c.chunk_rows ? kChunkMin : c.chunk_rows;
  j.src.chunk_target = (int32_t)hist_chunk_target(c, a.n_feat);
  j.src.seg_start = a.seg_start;
  j.src.seg_slot = a.seg_slot;
  j.src.chunk_desc = a.chunk_desc;
  j.src.chunk_total = a.chunk_total;
  j.src.multi_desc = a.multi_desc;
  j.src.merge_parts = a.merge_parts;
  j.chunk_rows = c.chunk_rows;
  j.seg_first = a.seg_first;
  j.seg_pfirst = a.seg_pfirst;
  j.multi = multi;
  j.n_multi = n_multi;
  return j;
}

// Compute units of the current device (cached per device).
static int device_cus() {
  static int cus[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cus[dev] == 0) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

// The launch configuration of a build.  A launch of many small segments -- a
// deep level of many eras: thousands of (node, era) segments of a few rows,
// each its own work item -- is bound by each item's chain of dependent loads
// and its clear and read-out, not by the LDS adds: it runs one wave per work
// item over one feature per lane (the smallest LDS histogram) and as many
// workgroups as the CUs hold at once, so that many more items are in flight.
// The plan, and so every result, is the configuration's own; the partials'
// words and flags are sized for any features per lane.
constexpr int64_t kFragSegments = 1024;  // segments of a launch from which it may count as fragmented
constexpr int64_t kFragRows = 256;       // ... with at most this many rows per segment on average
static HistConfig fragmented_config(const HistArgs& a, const HistConfig& c, int64_t rows_max, int64_t nseg_max) {
  if (nseg_max < kFragSegments || rows_max > kFragRows * nseg_max) return c;
  HistConfig cl = c;
  cl.fpl = 1;
  cl.waves = 1;
  cl.depth = 16;
  const size_t lds = lds_bytes(cl, a.n_bins, a.hess_const != 0);
  const int64_t per_cu = std::min<int64_t>((int64_t)((160 * 1024) / (lds + 1024)), 16);
  cl.slots = (int32_t)std::max<int64_t>(((int64_t)device_cus() * per_cu) & ~(int64_t)7, 8);
  return cl;
}

int hist_launch(HistArgs a, const HistConfig& c, int64_t rows_max, int64_t nseg_max,
                int32_t* multi, int32_t* n_multi, hipStream_t st) {
  WGBM_REQUIRE(!a.sub_g || (a.sub_h && !a.accumulate && a.slots_per_node > 0),
               "histogram: sibling subtraction needs both parent planes and overwrite mode");
  WGBM_REQUIRE(a.chunk_desc && a.chunk_total && a.multi_desc && a.merge_parts,
               "histogram: missing chunk descriptor scratch");
  WGBM_REQUIRE(a.codes && a.perm && (a.gh || (a.g && a.h)) && a.seg_start && a.seg_count && a.seg_slot &&
                   a.seg_first && a.seg_pfirst && a.nseg_dev && a.out_g && a.out_h && a.part && a.pflag &&
                   a.fine && a.stats && multi && n_multi,
               "histogram: missing table or scratch pointer");
  const HistConfig cl = fragmented_config(a, c, rows_max, nseg_max);  // the launch's tiling
  a.chunk_rows = c.chunk_rows;
  a.nslab = (int)ceil_div(a.n_feat, kWave * cl.fpl);
  a.h_bits = hess_bits(a.n_rows);
  a.fine_x = fine_extra(a.n_rows);
  a.wide_ok = (a.row_stride % cl.fpl == 0) && ((uintptr_t)a.codes % (uintptr_t)cl.fpl == 0);
  a.lds_pitch = lds_pitch(cl, a.n_bins);
  if (lds_bytes(cl, a.n_bins, a.hess_const != 0) > kLdsMax) {
    set_error("histogram LDS budget exceeded (%d bins, %d waves, %d features per lane)", a.n_bins, cl.waves, cl.fpl);
    return kTooBig;
  }
  if (!a.planned) {
    plan_job_launch(hist_plan_job(a, c, multi, n_multi), st);
    WGBM_LAUNCHED("k_hist_plan");
  }
  const int64_t max_chunks = hist_max_chunks(c, a.n_feat, rows_max, nseg_max);
  const int64_t max_multi = hist_max_multi(c, a.n_feat, rows_max);
  if (a.n_feat == 0) return kOk;
#define WGBM_D2(FPL, D0, D1)                                                         \
  return cl.depth <= D0 ? launch_fpl<FPL, D0>(a, cl, max_chunks, max_multi, multi, n_multi, st) \
                        : launch_fpl<FPL, D1>(a, cl, max_chunks, max_multi, multi, n_multi, st);
  switch (cl.fpl) {
    case 1: WGBM_D2(1, 16, 32)
    case 2: WGBM_D2(2, 8, 16)
    case 4: WGBM_D2(4, 8, 16)
    case 8: WGBM_D2(8, 4, 8)
    case 16: WGBM_D2(16, 2, 4)
  }
#undef WGBM_D2
  set_error("unsupported features-per-lane %d", cl.fpl);
  return kBadArg;
}

}  // namespace wgbm

using namespace wgbm;

// Start timing histogram launches (clears earlier records).
extern "C" int wgbm_profile_begin(void) {
  timer_clear();
  g_timer.on = true;
  return kOk;
}

// Stop timing; out4 = {build ms, build launches, merge ms, merge launches}.
// Waits for the recorded events.
extern "C" int wgbm_profile_end(double* out4) {
  g_timer.on = false;
  double t[2] = {0.0, 0.0};
  int k = 0;
  for (auto* v : {&g_timer.build, &g_timer.merge}) {
    for (auto& p : *v) {
      WGBM_HIP(hipEventSynchronize(p.second));
      float ms = 0.0f;
      WGBM_HIP(hipEventElapsedTime(&ms, p.first, p.second));
      t[k] += ms;
    }
    ++k;
  }
  out4[0] = t[0];
  out4[1] = (double)g_timer.build.size();
  out4[2] = t[1];
  out4[3] = (double)g_timer.merge.size();
  timer_clear();
  return kOk;
}

// Workspace carve-up for the reference-style histogram op.
struct Hist3Layout {
  size_t seg_start, seg_count, seg_slot, seg_first, seg_pfirst, nseg, multi, n_multi, chunk, desc, total_chunks,
      mdesc, mparts, sorted, cnt, stats, stat_part, part, pflag, fine, total;
};

static Hist3Layout hist3_layout(int64_t n, int n_feat, int n_bins, int n_eras, const HistConfig& c) {
  Hist3Layout L{};
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  size_t o = 0;
  const int64_t E = n_eras;
  L.seg_start = o; o = al(o + 4 * E);
  L.seg_count = o; o = al(o + 4 * E);
  L.seg_slot = o;  o = al(o + 4 * E);
  L.seg_first = o; o = al(o + 4 * (E + 1));
  L.seg_pfirst = o; o = al(o + 4 * E);
  L.nseg = o;      o = al(o + 4);
  const int64_t max_multi = hist_max_multi(c, n_feat, n);
  L.multi = o;     o = al(o + 4 * max_multi);
  L.chunk = o;     o = al(o + 4);
  L.n_multi = o;   o = al(o + 4);
  L.desc = o;      o = al(o + sizeof(int4) * hist_max_chunks(c, n_feat, n, E));
  L.total_chunks = o; o = al(o + 4);
  L.mdesc = o;     o = al(o + sizeof(int4) * max_multi);
  L.mparts = o;    o = al(o + 4);
  const int64_t nb = ceil_div(n, kSortBlock);
  L.sorted = o;    o = al(o + (E > 1 ? 4 * n : 0));
  L.cnt = o;       o = al(o + (E > 1 ? 4 * nb * E : 0));
  L.stats = o;     o = al(o + 4 * kStatWords);
  L.stat_part = o; o = al(o + 4 * kStatWords * kStatBlocks);
  L.part = o;      o = al(o + sizeof(longlong2) * hist_partial_words(c, n_feat, n_bins, n));
  L.pflag = o;     o = al(o + 4 * hist_partial_flags(c, n_feat, n));
  L.fine = o;      o = al(o + 8 * (size_t)E * n_feat * n_bins);
  L.total = o;
  return L;
}

extern "C" size_t wgbm_histogram3_workspace_bytes(int64_t n_rows, int32_t n_feat, int32_t n_bins,
                                                  int32_t n_eras, int64_t n_master_feat) {
  if (n_rows < 0 || n_feat < 0 || n_bins < 1 || n_eras < 1 || n_master_feat < 0) return 0;
  // the same configuration wgbm_compute_histogram3 picks for this matrix
  const HistConfig c = pick_hist_config(n_feat, n_bins, n_master_feat, false);
  return hist3_layout(n_rows, n_feat, n_bins, n_eras, c).total;
}

extern "C" int wgbm_compute_histogram3(const int8_t* bin_indices, int64_t n_total_rows,
                                       int64_t n_master_feat, const float* grads, const float* hess,
                                       const int32_t* sample_indices, int64_t n_samples,
                                       const int32_t* feature_indices, int32_t n_feat,
                                       const int32_t* era_indices, float* grad_hist, float* hess_hist,
                                       int32_t n_eras, int32_t num_bins, void* workspace,
                                       size_t workspace_bytes, hipStream_t stream) {
  WGBM_REQUIRE(n_samples >= 0 && n_samples < 0x7fffffff, "sample count out of range: %lld",
               (long long)n_samples);
  WGBM_REQUIRE(num_bins >= 1 && num_bins <= 127, "num_bins must be in [1, 127], got %d", num_bins);
  WGBM_REQUIRE(n_eras >= 1 && n_eras <= 32768, "era count must be in [1, 32768], got %d", n_eras);
  WGBM_REQUIRE(n_feat >= 0, "feature count must be non-negative");
  WGBM_REQUIRE(n_total_rows >= 0 && n_master_feat >= 0, "bad bin matrix shape");
  const HistConfig c = pick_hist_config(n_feat, num_bins, n_master_feat, false);
  const Hist3Layout L = hist3_layout(n_samples, n_feat, num_bins, n_eras, c);
  WGBM_REQUIRE(workspace_bytes >= L.total, "workspace too small: need %zu bytes, got %zu", L.total,
               workspace_bytes);
  if (n_samples == 0 || n_feat == 0) return kOk;  // histograms accumulate nothing
  char* ws = static_cast<char*>(workspace);
  auto P = [&](size_t off) { return reinterpret_cast<int32_t*>(ws + off); };
  HistArgs a{};
  a.codes = bin_indices;
  a.row_stride = n_master_feat;
  a.n_rows = n_total_rows;
  a.feat_idx = feature_indices;
  a.n_feat = n_feat;
  a.n_bins = num_bins;
  a.g = grads;
  a.h = hess;
  a.gh = nullptr;
  a.seg_start = P(L.seg_start);
  a.seg_count = P(L.seg_count);
  a.seg_slot = P(L.seg_slot);
  a.seg_first = P(L.seg_first);
  a.seg_pfirst = P(L.seg_pfirst);
  a.nseg_dev = P(L.nseg);
  a.chunk_dev = P(L.chunk);
  a.chunk_desc = reinterpret_cast<int4*>(ws + L.desc);
  a.chunk_total = P(L.total_chunks);
  a.multi_desc = reinterpret_cast<int4*>(ws + L.mdesc);
  a.merge_parts = P(L.mparts);
  a.out_g = grad_hist;
  a.out_h = hess_hist;
  a.part = reinterpret_cast<longlong2*>(ws + L.part);
  a.pflag = P(L.pflag);
  a.fine = reinterpret_cast<long long*>(ws + L.fine);
  a.stats = P(L.stats);
  a.slots_per_node = n_eras;
  a.accumulate = 1;
  a.dual = 1;  // per-node op: constant-hessian nodes (regression) take one plane
  // The fine sums start at zero (the workspace may be fresh memory); the
  // read-outs leave them at zero.
  WGBM_HIP(hipMemsetAsync(ws + L.fine, 0, L.total - L.fine, stream));
  const int n = (int)n_samples;
  const int nsb = stats_blocks(n);
  hipLaunchKernelGGL(k_list_stats, dim3((unsigned)nsb), dim3(256), 0, stream, sample_indices, n, n_total_rows, grads,
                     hess, P(L.stat_part));
  WGBM_LAUNCHED("k_list_stats");
  hipLaunchKernelGGL(k_stats_reduce, dim3(1), dim3(256), 0, stream, P(L.stat_part), nsb, P(L.stats));
  WGBM_LAUNCHED("k_stats_reduce");
  if (n_eras == 1) {
    a.perm = sample_indices;
    hipLaunchKernelGGL(k_single_segment, dim3(1), dim3(64), 0, stream, P(L.seg_start), P(L.seg_count),
                       P(L.seg_slot), P(L.nseg), n);
    WGBM_LAUNCHED("k_single_segment");
  } else {
    const int nb = (int)ceil_div(n, kSortBlock);
    const size_t lds = (size_t)n_eras * 4;  // one counter per era (at most 128 KB)
    if (lds > 64 * 1024) {  // beyond the default dynamic LDS limit: opt in
      WGBM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_era_count),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      WGBM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_era_scatter),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    hipLaunchKernelGGL(k_era_count, dim3(nb), dim3(256), lds, stream, sample_indices, n, n_total_rows, era_indices,
                       (int)n_eras, P(L.cnt));
    WGBM_LAUNCHED("k_era_count");
    hipLaunchKernelGGL(k_era_scan, dim3(1), dim3(1024), 0, stream, P(L.cnt), nb, (int)n_eras,
                       P(L.seg_start), P(L.seg_count), P(L.seg_slot));
    WGBM_LAUNCHED("k_era_scan");
    hipLaunchKernelGGL(k_era_scatter, dim3(nb), dim3(64), lds, stream, sample_indices, n, n_total_rows, era_indices,
                       (int)n_eras, P(L.cnt), P(L.sorted));
    WGBM_LAUNCHED("k_era_scatter");
    hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, stream, P(L.nseg), (int)n_eras);
    WGBM_LAUNCHED("k_set_scalar");
    a.perm = P(L.sorted);
  }
  return hist_launch(a, c, n, n_eras, P(L.multi), P(L.n_multi), stream);
}

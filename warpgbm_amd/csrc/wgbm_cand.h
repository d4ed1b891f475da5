// Split candidates: the strict total order the selection reduces with, and
// the one-era reduction of a node's per-workgroup candidates (shared by the
// level scan's pick and the level tail that folds the decisions in).
#pragma once
#include "common.h"
#include <math.h>

namespace wgbm {

struct Cand {
  int agree;    // |sum over eras of direction| (0 when E == 1)
  float crit;   // -inf when the candidate is masked out
  int idx;      // flat index f_local*(B-1) + b, INT_MAX for "none"
};

__device__ __forceinline__ bool cand_better(const Cand& a, const Cand& b) {
  if (a.agree != b.agree) return a.agree > b.agree;
  if (a.crit != b.crit) return a.crit > b.crit;
  return a.idx < b.idx;
}

__device__ __forceinline__ Cand cand_none() { return Cand{-1, -INFINITY, 0x7fffffff}; }

__device__ __forceinline__ Cand wave_best(Cand c) {
  for (int o = 32; o >= 1; o >>= 1) {
    Cand d;
    d.agree = __shfl_xor(c.agree, o);
    d.crit = __shfl_xor(c.crit, o);
    d.idx = __shfl_xor(c.idx, o);
    if (cand_better(d, c)) c = d;
  }
  return c;
}

// Best of node `node`'s n per-workgroup candidates (slots node * n .. + n - 1),
// run by one whole wave (lane = 0..63); every lane returns the winner.
// Four candidates' loads per lane are in flight before the comparisons.
__device__ __forceinline__ Cand node_best(const int32_t* __restrict__ agree, const float* __restrict__ crit,
                                          const int32_t* __restrict__ idx, int node, int n, int lane) {
  Cand b = cand_none();
  constexpr int U = 4;
  for (int w0 = lane; w0 < n; w0 += 64 * U) {
    Cand k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t o = (int64_t)node * n + min(w0 + 64 * u, n - 1);
      k[u] = Cand{agree[o], crit[o], idx[o]};
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (w0 + 64 * u < n && cand_better(k[u], b)) b = k[u];
  }
  return wave_best(b);
}

// A winner's (local feature, bin), or (-1, -1) for "leaf" (no valid
// candidate).
__device__ __forceinline__ void cand_split(const Cand& b, int n_bins, int& fl, int& bin) {
  fl = -1;
  bin = -1;
  if (b.idx != 0x7fffffff && b.crit != -INFINITY) {
    fl = b.idx / (n_bins - 1);
    bin = b.idx % (n_bins - 1);
  }
}

}  // namespace wgbm

// Device helpers shared by the tile-sweep kernels (msv_kernels.hip) and the
// batched ansatz kernels (msv_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "msv_internal.h"

namespace msv {

template <typename V>
__device__ __forceinline__ V cmul(V a, V b) {
  V r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename V>
__device__ __forceinline__ V cadd(V a, V b) {
  V r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename V>
__device__ __forceinline__ V csub(V a, V b) {
  V r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}

template <int RB>
__device__ __forceinline__ uint64_t reg_goff(int r, const uint64_t *gstr) {
  uint64_t o = 0;
#pragma unroll
  for (int j = 0; j < RB; ++j)
    if ((r >> j) & 1) o += gstr[j];
  return o;
}

// XOR swizzle of LDS slots: when lanes step by 2^RB slots (registers on the
// lowest local bits) they would all hit one bank pair; folding bits above
// RB into the low bits spreads them while staying a bijection. The map is
// linear over GF(2), so swz(tl | roff) = swz(tl) ^ swz(roff) for disjoint
// thread and register bits: the register part is wave-uniform.
template <int RB>
__device__ __forceinline__ uint32_t lds_swz(uint32_t l) {
  constexpr uint32_t kMask = (RB >= 4) ? 15u : 0u;
  return l ^ ((l >> RB) & kMask);
}

// Args: any argument block with nruns / run_src / run_width / run_dst.
// The run parameters are read in a rolled loop (scalar loads per tile):
// hoisting all kMaxRuns triples into SGPRs would spill them.
template <typename Args>
__device__ __forceinline__ uint64_t tile_base(const Args &A, uint64_t tile) {
  uint64_t base = 0;
  const int nr = static_cast<int>(A.nruns);
#pragma unroll 1
  for (int k = 0; k < nr; ++k)
    base |= ((tile >> A.run_src[k]) & ((1ull << A.run_width[k]) - 1ull)) << A.run_dst[k];
  return base;
}

// Position of this thread in segment `sg`: tl = its local bits, gb = its
// amplitude index bits (tile base included), tm = its thread-id bits in use.
__device__ __forceinline__ void seg_thread(const TileSeg &sg, uint32_t tid, uint64_t base,
                                           uint32_t &tl, uint64_t &gb, uint32_t &tm) {
  tl = 0;
  gb = base;
  tm = tid & sg.tmask;
#pragma unroll
  for (int k = 0; k < kTileThreadBits; ++k) {
    const uint32_t b = (tm >> k) & 1u;
    tl |= b << sg.tl_bit[k];
    gb |= static_cast<uint64_t>(b) << sg.gi_bit[k];
  }
}

}  // namespace msv

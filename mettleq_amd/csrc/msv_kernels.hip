// gfx950 kernels of the statevector engine.
//
// Every structured layer is an in-place sweep over the amplitude buffer
// except the XOR permutation, which gathers into a second buffer. The
// reference engine launches one out-of-place Metal kernel per 2-4 qubits
// (shaders/single_qubit.py:165-210, shaders/qft.py:68-109); here one "tile
// sweep" applies up to a dozen qubits' worth of gates per HBM round trip:
//
//   tile      = 2^nlocal amplitudes whose index differs only in the tile's
//               local bits (the lowest "column" bits for coalescing plus the
//               gate bits); the remaining index bits number the tiles.
//   segment   = one assignment of RB local bits to registers: each thread
//               holds the 2^RB amplitudes that differ in those bits, so any
//               gate whose target is a register bit is a register-only
//               butterfly. Segments exchange through LDS.
//   program   = ordered ops (2x2 with optional controls, QFT stages,
//               one/two-bit diagonals) executed segment by segment.
#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>

#include "msv_device.h"
#include "msv_internal.h"
#include "msv_philox.h"

namespace msv {

namespace {

template <typename T> struct CisPi;
template <> struct CisPi<float> {
  // exp(i*pi*x/2^w) for an integer x < 2^w; x/2^w is formed exactly up to the
  // 24-bit rounding of x, and sincospif reduces the argument in units of pi.
  __device__ static float2 eval(uint64_t x, uint32_t w) {
    float f = ldexpf(static_cast<float>(x), -static_cast<int>(w));
    float s, c;
    sincospif(f, &s, &c);
    return make_float2(c, s);
  }
};
template <> struct CisPi<double> {
  __device__ static double2 eval(uint64_t x, uint32_t w) {
    double f = ldexp(static_cast<double>(x), -static_cast<int>(w));
    double s, c;
    sincospi(f, &s, &c);
    return make_double2(c, s);
  }
};

// ---------------------------------------------------------------------------
// tile-program building blocks (RB register bits, J = target register bit)
// ---------------------------------------------------------------------------

template <typename T, int RB, int J, bool CTRL>
__device__ __forceinline__ void u2_pairs(typename Vec2<T>::type *a, const T *m, bool tok,
                                         uint32_t creg) {
  using V = typename Vec2<T>::type;
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    const int r1 = r | (1 << J);
    const V v0 = a[r], v1 = a[r1];
    V n0, n1;
    n0.x = m[0] * v0.x - m[1] * v0.y + m[2] * v1.x - m[3] * v1.y;
    n0.y = m[0] * v0.y + m[1] * v0.x + m[2] * v1.y + m[3] * v1.x;
    n1.x = m[4] * v0.x - m[5] * v0.y + m[6] * v1.x - m[7] * v1.y;
    n1.y = m[4] * v0.y + m[5] * v0.x + m[6] * v1.y + m[7] * v1.x;
    if (CTRL) {
      const bool ok = tok && ((static_cast<uint32_t>(r) & creg) == creg);
      a[r].x = ok ? n0.x : v0.x;
      a[r].y = ok ? n0.y : v0.y;
      a[r1].x = ok ? n1.x : v1.x;
      a[r1].y = ok ? n1.y : v1.y;
    } else {
      a[r] = n0;
      a[r1] = n1;
    }
  }
}

template <typename T, int RB, bool CTRL>
__device__ __forceinline__ void u2_dispatch(typename Vec2<T>::type *a, const T *m, int j, bool tok,
                                            uint32_t creg) {
  switch (j) {
    case 0: u2_pairs<T, RB, 0, CTRL>(a, m, tok, creg); break;
    case 1: if constexpr (RB > 1) u2_pairs<T, RB, 1, CTRL>(a, m, tok, creg); break;
    case 2: if constexpr (RB > 2) u2_pairs<T, RB, 2, CTRL>(a, m, tok, creg); break;
    case 3: if constexpr (RB > 3) u2_pairs<T, RB, 3, CTRL>(a, m, tok, creg); break;
    case 4: if constexpr (RB > 4) u2_pairs<T, RB, 4, CTRL>(a, m, tok, creg); break;
    default: break;
  }
}

// X on register bit J: the pairs swap (under the controls).
template <typename T, int RB, int J, bool CTRL>
__device__ __forceinline__ void x_pairs(typename Vec2<T>::type *a, bool tok, uint32_t creg) {
  using V = typename Vec2<T>::type;
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    const int r1 = r | (1 << J);
    const V v0 = a[r], v1 = a[r1];
    if (CTRL) {
      const bool ok = tok && ((static_cast<uint32_t>(r) & creg) == creg);
      a[r].x = ok ? v1.x : v0.x;
      a[r].y = ok ? v1.y : v0.y;
      a[r1].x = ok ? v0.x : v1.x;
      a[r1].y = ok ? v0.y : v1.y;
    } else {
      a[r] = v1;
      a[r1] = v0;
    }
  }
}

template <typename T, int RB, bool CTRL>
__device__ __forceinline__ void x_dispatch(typename Vec2<T>::type *a, int j, bool tok, uint32_t creg) {
  switch (j) {
    case 0: x_pairs<T, RB, 0, CTRL>(a, tok, creg); break;
    case 1: if constexpr (RB > 1) x_pairs<T, RB, 1, CTRL>(a, tok, creg); break;
    case 2: if constexpr (RB > 2) x_pairs<T, RB, 2, CTRL>(a, tok, creg); break;
    case 3: if constexpr (RB > 3) x_pairs<T, RB, 3, CTRL>(a, tok, creg); break;
    case 4: if constexpr (RB > 4) x_pairs<T, RB, 4, CTRL>(a, tok, creg); break;
    default: break;
  }
}

template <typename T>
__device__ __forceinline__ T fma_t(T a, T b, T c) {
  if constexpr (std::is_same<T, float>::value) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

// QFT stage on register bit J. The ladder phase of the amplitude with the
// stage bit set is exp(i*pi*L/2^(p-lb)), L = index bits [lb, p). It factors
// into the thread part wt (bits outside the registers) and the register
// part from the JL register bits directly below J, which are index bits
// p-1 .. p-JL: exp(i*pi*v/2^JL) = w32[v << (4-JL)].
template <typename T, int RB, int J, int JL, bool INV>
__device__ __forceinline__ void qft_pairs(typename Vec2<T>::type *a, typename Vec2<T>::type wt,
                                          const T *w32) {
  using V = typename Vec2<T>::type;
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    const int r1 = r | (1 << J);
    V w = wt;
    if constexpr (JL > 0) {
      const int v = (r >> (J - JL)) & ((1 << JL) - 1);
      const int mi = v << (4 - JL);
      if (mi != 0) {
        V e;
        e.x = w32[2 * mi];
        e.y = w32[2 * mi + 1];
        w = cmul(wt, e);
      }
    }
    if (INV) w.y = -w.y;
    const V v0 = a[r], v1 = a[r1];
    if (!INV) {
      a[r] = cadd(v0, v1);
      a[r1] = cmul(csub(v0, v1), w);
    } else {
      const V t = cmul(v1, w);
      a[r] = cadd(v0, t);
      a[r1] = csub(v0, t);
    }
  }
}

#define MSV_QFT_CASE(JJ, LL)                                                   \
  case (JJ) * 8 + (LL):                                                        \
    if constexpr (RB > (JJ)) {                                                 \
      if (inv) qft_pairs<T, RB, JJ, LL, true>(a, wt, w32);                     \
      else qft_pairs<T, RB, JJ, LL, false>(a, wt, w32);                        \
    }                                                                          \
    break;

template <typename T, int RB>
__device__ __forceinline__ void qft_dispatch(typename Vec2<T>::type *a, typename Vec2<T>::type wt,
                                             const T *w32, int j, int jl, bool inv) {
  switch (j * 8 + jl) {
    MSV_QFT_CASE(0, 0)
    MSV_QFT_CASE(1, 0) MSV_QFT_CASE(1, 1)
    MSV_QFT_CASE(2, 0) MSV_QFT_CASE(2, 1) MSV_QFT_CASE(2, 2)
    MSV_QFT_CASE(3, 0) MSV_QFT_CASE(3, 1) MSV_QFT_CASE(3, 2) MSV_QFT_CASE(3, 3)
    MSV_QFT_CASE(4, 0) MSV_QFT_CASE(4, 1) MSV_QFT_CASE(4, 2) MSV_QFT_CASE(4, 3)
    MSV_QFT_CASE(4, 4)
    default: break;
  }
}
#undef MSV_QFT_CASE

// Registers whose amplitude has the op's bit set, as a mask over r: the
// register pattern `pat` of a register bit, or every register when a
// thread / tile bit (index mask `imask`) is set in this thread's index.
template <int RB>
__device__ __forceinline__ uint32_t regs_with_bit(uint32_t pat, uint64_t imask, uint64_t gb) {
  constexpr uint32_t all = (RB >= 5) ? ~0u : ((1u << (1 << RB)) - 1u);
  return pat | ((gb & imask) ? all : 0u);
}

// Diagonal over 1 or 2 index bits (entries d0..d3 by index 2 ba + bb, or
// d0, d1 by ba). Three forms, chosen per op (wave-uniform): both bits in
// registers (the entry of each register is uniform, selected on the scalar
// unit), one in registers (two candidate entries per thread, one uniform
// choice per register), none (one entry per thread). Each amplitude then
// costs one complex multiply.
template <typename T, int RB>
__device__ __forceinline__ void diag_apply(typename Vec2<T>::type *a, const TileOp<T> &op,
                                           uint64_t gb) {
  using V = typename Vec2<T>::type;
  const bool two = op.nbits == 2;
  const bool areg = op.ra >= 0;
  const bool breg = !two || op.rb >= 0;
  V d0, d1, d2, d3;
  d0.x = op.m[0]; d0.y = op.m[1];
  d1.x = op.m[2]; d1.y = op.m[3];
  d2.x = op.m[4]; d2.y = op.m[5];
  d3.x = op.m[6]; d3.y = op.m[7];
  if (!two) {  // one bit: entries d0 (bit 0), d1 (bit 1)
    d2 = d1;
    d3 = d1;
    d1 = d0;
  }
  const uint32_t pa = op.pa, pb = op.pb;
  if (areg && breg) {
#pragma unroll
    for (int r = 0; r < (1 << RB); ++r) {
      const bool ba = (pa >> r) & 1u, bb = (pb >> r) & 1u;
      const V lo = bb ? d1 : d0, hi = bb ? d3 : d2;
      a[r] = cmul(a[r], ba ? hi : lo);
    }
  } else if (areg || breg) {
    // the thread's value of the non-register bit fixes the other half
    const uint64_t tmask = areg ? op.bmask : op.amask;
    const bool t = (gb & tmask) != 0;
    V lo, hi;  // entries for the register bit = 0 / 1
    if (areg) {
      lo = t ? d1 : d0;
      hi = t ? d3 : d2;
    } else {
      lo = t ? d2 : d0;
      hi = t ? d3 : d1;
    }
    const uint32_t p = areg ? pa : pb;
#pragma unroll
    for (int r = 0; r < (1 << RB); ++r) a[r] = cmul(a[r], ((p >> r) & 1u) ? hi : lo);
  } else {
    const bool ta = (gb & op.amask) != 0, tb = two && (gb & op.bmask) != 0;
    const V lo = tb ? d1 : d0, hi = tb ? d3 : d2;
    const V d = ta ? hi : lo;
#pragma unroll
    for (int r = 0; r < (1 << RB); ++r) a[r] = cmul(a[r], d);
  }
}

template <typename T> struct SignBits;
template <> struct SignBits<float> {
  __device__ static float flip(float v, uint32_t sb) { return __uint_as_float(__float_as_uint(v) ^ sb); }
};
template <> struct SignBits<double> {
  __device__ static double flip(double v, uint32_t sb) {
    return __longlong_as_double(__double_as_longlong(v) ^ (static_cast<long long>(sb) << 32));
  }
};

// Diagonal with entries +1 / -1 (bit k of jl: entry k, index 2 ba + bb or
// ba is -1): the thread's registers to negate as a mask over r, then each
// amplitude's sign bits flipped.
template <typename T, int RB>
__device__ __forceinline__ void sign_apply(typename Vec2<T>::type *a, const TileOp<T> &op, uint64_t gb) {
  const uint32_t A = regs_with_bit<RB>(op.pa, op.amask, gb);
  const uint32_t s = op.jl;
  uint32_t M;
  if (op.nbits == 2) {
    const uint32_t B = regs_with_bit<RB>(op.pb, op.bmask, gb);
    M = ((s & 1u) ? (~A & ~B) : 0u) | ((s & 2u) ? (~A & B) : 0u) | ((s & 4u) ? (A & ~B) : 0u) |
        ((s & 8u) ? (A & B) : 0u);
  } else {
    M = ((s & 1u) ? ~A : 0u) | ((s & 2u) ? A : 0u);
  }
  if ((op.pa | op.pb) == 0u) {
    // no register bits: one sign for all of the thread's amplitudes
    const uint32_t sb = M << 31;
#pragma unroll
    for (int r = 0; r < (1 << RB); ++r) {
      a[r].x = SignBits<T>::flip(a[r].x, sb);
      a[r].y = SignBits<T>::flip(a[r].y, sb);
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    const uint32_t sb = (M << (31 - r)) & 0x80000000u;
    a[r].x = SignBits<T>::flip(a[r].x, sb);
    a[r].y = SignBits<T>::flip(a[r].y, sb);
  }
}

template <typename T, int RB>
__device__ __forceinline__ void run_ops(typename Vec2<T>::type *a, const TileArgs<T> &A,
                                        uint32_t ob, uint32_t oe, uint64_t gb, uint32_t tm) {
  using V = typename Vec2<T>::type;
  for (uint32_t o = ob; o < oe; ++o) {
    const TileOp<T> &op = A.ops[o];
    const uint32_t kind = op.kind;
    if (kind == OP_U2) {
      T m[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) m[k] = op.m[k];
      if (!op.has_ctrl) {
        u2_dispatch<T, RB, false>(a, m, op.rpos, true, 0u);
      } else {
        // controls only outside the registers: creg = 0 at compile time, so
        // no per-pair scalar test of the register controls
        const bool tok = (gb & op.cmask_t) == op.cmask_t;
        if (op.cmask_r == 0u) u2_dispatch<T, RB, true>(a, m, op.rpos, tok, 0u);
        else u2_dispatch<T, RB, true>(a, m, op.rpos, tok, op.cmask_r);
      }
    } else if (kind == OP_QFT_F || kind == OP_QFT_I) {
      const uint32_t w = static_cast<uint32_t>(op.pbit) - op.lb;
      const uint64_t x = (gb >> op.lb) & ((w >= 64) ? ~0ull : ((1ull << w) - 1ull));
      const V wt = CisPi<T>::eval(x, w);
      qft_dispatch<T, RB>(a, wt, A.tab, op.rpos, op.jl, kind == OP_QFT_I);
    } else if (kind == OP_DIAG) {
      diag_apply<T, RB>(a, op, gb);
    } else if (kind == OP_X) {
      if (!op.has_ctrl) {
        x_dispatch<T, RB, false>(a, op.rpos, true, 0u);
      } else {
        const bool tok = (gb & op.cmask_t) == op.cmask_t;
        if (op.cmask_r == 0u) x_dispatch<T, RB, true>(a, op.rpos, tok, 0u);
        else x_dispatch<T, RB, true>(a, op.rpos, tok, op.cmask_r);
      }
    } else if (kind == OP_SIGN) {
      sign_apply<T, RB>(a, op, gb);
    }
  }
}

// Amplitude loads and stores of the full-state sweeps. With MSV_NT (the
// default) they carry the nontemporal hint: every amplitude is read and
// written once per sweep, so keeping it in L2 only displaces other lines
// (measured on MI355X: 3-4% faster in-place sweeps).
#ifndef MSV_NT
#define MSV_NT 1
#endif
template <typename V> struct NtVec;
template <> struct NtVec<float2> { typedef float type __attribute__((ext_vector_type(2))); };
template <> struct NtVec<double2> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct NtVec<float4> { typedef float type __attribute__((ext_vector_type(4))); };

template <typename V>
__device__ __forceinline__ V ld_amp(const V *p) {
  if constexpr (MSV_NT) {
    using N = typename NtVec<V>::type;
    const N v = __builtin_nontemporal_load(reinterpret_cast<const N *>(p));
    V r;
    r.x = v.x;
    r.y = v.y;
    return r;
  } else {
    return *p;
  }
}

template <typename V>
__device__ __forceinline__ void st_amp(V *p, V a) {
  if constexpr (MSV_NT) {
    using N = typename NtVec<V>::type;
    N v;
    v.x = a.x;
    v.y = a.y;
    __builtin_nontemporal_store(v, reinterpret_cast<N *>(p));
  } else {
    *p = a;
  }
}

__device__ __forceinline__ float4 ld_f4(const float4 *p) {
  if constexpr (MSV_NT) {
    using N = NtVec<float4>::type;
    const N v = __builtin_nontemporal_load(reinterpret_cast<const N *>(p));
    return make_float4(v.x, v.y, v.z, v.w);
  } else {
    return *p;
  }
}

__device__ __forceinline__ void st_f4(float4 *p, float4 a) {
  if constexpr (MSV_NT) {
    using N = NtVec<float4>::type;
    N v = {a.x, a.y, a.z, a.w};
    __builtin_nontemporal_store(v, reinterpret_cast<N *>(p));
  } else {
    *p = a;
  }
}

// First-segment amplitudes of a tile: from HBM, or generated when the input
// is a basis state (a deferred reset: nothing is read).
template <typename T, int RB>
__device__ __forceinline__ void load_seg0(const typename Vec2<T>::type *__restrict__ st,
                                          const TileArgs<T> &A, uint64_t gb, const uint64_t *gstr,
                                          typename Vec2<T>::type *a) {
  if (A.from_basis) {
#pragma unroll
    for (int r = 0; r < (1 << RB); ++r) {
      a[r].x = (gb + reg_goff<RB>(r, gstr)) == A.basis ? T(1) : T(0);
      a[r].y = T(0);
    }
  } else {
#pragma unroll
    for (int r = 0; r < (1 << RB); ++r) a[r] = ld_amp(st + gb + reg_goff<RB>(r, gstr));
  }
}

template <typename T, int RB>
__device__ __forceinline__ void load_first(const typename Vec2<T>::type *__restrict__ st,
                                           const TileArgs<T> &A, uint32_t tid, uint64_t tile,
                                           typename Vec2<T>::type *a) {
  const TileSeg &sg = A.seg[0];
  uint64_t gstr[RB];
#pragma unroll
  for (int j = 0; j < RB; ++j) gstr[j] = sg.gstr[j];
  uint32_t tl, tm;
  uint64_t gb;
  seg_thread(sg, tid, tile_base(A, tile), tl, gb, tm);
  load_seg0<T, RB>(st, A, gb, gstr, a);
}

// Lookup-table index of phase window w for amplitude index i.
// A window's table index without its offset: linear (over XOR) in i.
__device__ __forceinline__ uint32_t win_bits(const PhaseWin &w, uint64_t i) {
  const uint32_t a = static_cast<uint32_t>(i >> w.s1) & ((1u << w.w1) - 1u);
  const uint32_t b = static_cast<uint32_t>(i >> w.s2) & ((1u << w.w2) - 1u);
  return a | (b << w.w1);
}

__device__ __forceinline__ uint32_t win_index(const PhaseWin &w, uint64_t i) { return w.off + win_bits(w, i); }

// Segment `s` of a tile: its thread's local and index bits.
struct SegPos {
  uint32_t tl, tm, tsw;
  uint64_t gb;
};

// Local index bits / amplitude index of a thread (thread-id bits in use:
// tm) in segment sg.
__device__ __forceinline__ uint32_t seg_local(const TileSeg &sg, uint32_t tm) {
  uint32_t tl = 0;
#pragma unroll
  for (int k = 0; k < kTileThreadBits; ++k) tl |= ((tm >> k) & 1u) << sg.tl_bit[k];
  return tl;
}
__device__ __forceinline__ uint64_t seg_index(const TileSeg &sg, uint32_t tm, uint64_t base) {
  uint64_t gb = base;
#pragma unroll
  for (int k = 0; k < kTileThreadBits; ++k) gb |= static_cast<uint64_t>((tm >> k) & 1u) << sg.gi_bit[k];
  return gb;
}

template <int RB>
__device__ __forceinline__ SegPos seg_pos(const TileSeg &sg, uint32_t tid, uint64_t base) {
  SegPos p;
  seg_thread(sg, tid, base, p.tl, p.gb, p.tm);
  p.tsw = lds_swz<RB>(p.tl);
  return p;
}


// One tile per block and iteration (persistent grid): the first segment
// loads from HBM, later segments exchange through LDS, the last one stores.
// PF: the first segment of the block's next tile is loaded into a second
// register set at the top of each iteration, so its HBM reads overlap this
// tile's ops, LDS exchanges and stores. Each tile belongs to one block, so
// reading it early is safe. The loads are issued before this tile's stores
// and consumed after them, and the last segment (the one that stores) is
// straight-line code after the segment loop, so that the wait for the
// prefetched amplitudes does not also wait for the stores. On the block's
// last tile the "next" load re-reads the current tile and is discarded:
// the host uses this form only with many tiles per block.
template <typename T, int RB, bool PF>
__device__ __forceinline__ void tile_sweep(typename Vec2<T>::type *__restrict__ st,
                                           const TileArgs<T> &A) {
  using V = typename Vec2<T>::type;
  constexpr int R = 1 << RB;
  extern __shared__ unsigned char smem_raw[];
  V *lds = reinterpret_cast<V *>(smem_raw);
  const uint32_t tid = threadIdx.x;
  // Blocks are dispatched round-robin over the XCDs (block b on XCD
  // b mod G). With xcd_groups = G the block's virtual rank makes the
  // blocks of one XCD take consecutive tiles (adjacent rows in HBM).
  uint64_t vrank = blockIdx.x;
  if (A.xcd_groups > 1 && gridDim.x % A.xcd_groups == 0)
    vrank = (blockIdx.x % A.xcd_groups) * (gridDim.x / A.xcd_groups) + blockIdx.x / A.xcd_groups;
  V a[R];
  V nx[PF ? R : 1];
  if constexpr (PF) {
    if (vrank < A.n_tiles) {
      load_first<T, RB>(st, A, tid, vrank, a);
      // complete these loads here, so that the loop never sees them pending
#pragma unroll
      for (int r = 0; r < R; ++r) asm volatile("" : "+v"(a[r].x), "+v"(a[r].y));
    }
  }
  const uint32_t last = A.nseg - 1;
  for (uint64_t tile = vrank; tile < A.n_tiles; tile += gridDim.x) {
    const uint64_t base = tile_base(A, tile);
    if constexpr (PF) {
      const uint64_t next = tile + gridDim.x;
      load_first<T, RB>(st, A, tid, next < A.n_tiles ? next : tile, nx);
    }
    // segments 0 .. last; all but the last end with an LDS exchange
    for (uint32_t s = 0;; ++s) {
      const TileSeg &sg = A.seg[s];
      const SegPos p = seg_pos<RB>(sg, tid, base);
      if (s == 0) {
        if constexpr (!PF) {
          uint64_t gstr[RB];
#pragma unroll
          for (int j = 0; j < RB; ++j) gstr[j] = sg.gstr[j];
          load_seg0<T, RB>(st, A, p.gb, gstr, a);
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
          a[r] = lds[p.tsw ^ lds_swz<RB>(static_cast<uint32_t>(r) << sg.reg0)];
      }
      run_ops<T, RB>(a, A, sg.op_begin, sg.op_end, p.gb, p.tm);
      if (s == last) break;
      __syncthreads();
#pragma unroll
      for (int r = 0; r < R; ++r) lds[p.tsw ^ lds_swz<RB>(static_cast<uint32_t>(r) << sg.reg0)] = a[r];
      __syncthreads();
    }
    // the last segment's stores (addresses recomputed here instead of being
    // kept live across the ops)
    {
      const TileSeg &sg = A.seg[last];
      const SegPos p = seg_pos<RB>(sg, tid, base);
      uint64_t gstr[RB];
#pragma unroll
      for (int j = 0; j < RB; ++j) gstr[j] = sg.gstr[j];
      if (A.scale != T(1)) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          a[r].x *= A.scale;
          a[r].y *= A.scale;
        }
      }
#pragma unroll
      for (int r = 0; r < R; ++r) st_amp(st + p.gb + reg_goff<RB>(r, gstr), a[r]);
    }
    if constexpr (PF) {
#pragma unroll
      for (int r = 0; r < R; ++r) a[r] = nx[r];
    }
  }
}

template <typename T, int RB>
__global__ void __launch_bounds__(512) k_tile(typename Vec2<T>::type *__restrict__ st,
                                              const TileArgs<T> A) {
  tile_sweep<T, RB, false>(st, A);
}

// The prefetching form holds two register sets; it is limited to 128 VGPRs
// so that two 512-thread blocks (the LDS limit) stay resident per CU.
template <typename T, int RB>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_tile_pf(typename Vec2<T>::type *__restrict__ st, const TileArgs<T> A) {
  tile_sweep<T, RB, true>(st, A);
}

// ---------------------------------------------------------------------------
// commuting rotation sweep (1q layers in Z-Y-Z split form)
// ---------------------------------------------------------------------------

// Rotation [[c, -s], [s, c]] on register bit J as three shears (Paeth):
//   v0 += a v1;  v1 += b v0;  v0 += a v1,   a = -tan(theta/2), b = sin(theta).
// Every step is one in-place FMA per component, so no temporaries (and no
// register moves) are needed: 3 FMAs per amplitude component per gate.
// a = b = 0 is the exact identity.
template <typename T, int RB, int J>
__device__ __forceinline__ void shear_pairs(typename Vec2<T>::type *a, T sa, T sb) {
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    const int r1 = r | (1 << J);
    a[r].x = fma_t(sa, a[r1].x, a[r].x);
    a[r].y = fma_t(sa, a[r1].y, a[r].y);
    a[r1].x = fma_t(sb, a[r].x, a[r1].x);
    a[r1].y = fma_t(sb, a[r].y, a[r1].y);
    a[r].x = fma_t(sa, a[r1].x, a[r].x);
    a[r].y = fma_t(sa, a[r1].y, a[r].y);
  }
}

// Exchanges register bit J with lane bit LB (4 or 5) inside the wavefront:
// for each register pair (r, r | 2^J) the lanes whose bit LB is 1 trade
// their r amplitude for the r | 2^J amplitude of the lane whose bit LB is
// 0. v_permlane16_swap trades the odd 16-lane rows of its first operand
// with the even rows of its second, v_permlane32_swap the upper 32 lanes of
// the first with the lower 32 of the second: afterwards register bit J
// holds what lane bit LB held and the reverse. Needs full waves.
template <int LB>
__device__ __forceinline__ void swap_word(uint32_t &lo, uint32_t &hi) {
  if constexpr (LB == 4) {
    const auto p = __builtin_amdgcn_permlane16_swap(lo, hi, false, false);
    lo = p[0];
    hi = p[1];
  } else {
    const auto p = __builtin_amdgcn_permlane32_swap(lo, hi, false, false);
    lo = p[0];
    hi = p[1];
  }
}

template <int LB>
__device__ __forceinline__ void swap_val(float &lo, float &hi) {
  uint32_t a = __float_as_uint(lo), b = __float_as_uint(hi);
  swap_word<LB>(a, b);
  lo = __uint_as_float(a);
  hi = __uint_as_float(b);
}

template <int LB>
__device__ __forceinline__ void swap_val(double &lo, double &hi) {
  const uint64_t a = static_cast<uint64_t>(__double_as_longlong(lo));
  const uint64_t b = static_cast<uint64_t>(__double_as_longlong(hi));
  uint32_t al = static_cast<uint32_t>(a), ah = static_cast<uint32_t>(a >> 32);
  uint32_t bl = static_cast<uint32_t>(b), bh = static_cast<uint32_t>(b >> 32);
  swap_word<LB>(al, bl);
  swap_word<LB>(ah, bh);
  lo = __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(ah) << 32) | al));
  hi = __longlong_as_double(static_cast<long long>((static_cast<uint64_t>(bh) << 32) | bl));
}

template <typename T, int RB, int J, int LB>
__device__ __forceinline__ void lane_swap(typename Vec2<T>::type *a) {
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    const int r1 = r | (1 << J);
    swap_val<LB>(a[r].x, a[r1].x);
    swap_val<LB>(a[r].y, a[r1].y);
  }
}

// Diagonal over all index bits, applied to a segment's amplitudes: tab[0..63]
// holds the register part d(r) for the segment's register bits (constant
// included); `f` is the thread part (one entry of a host table indexed by
// the thread-id bits in use, formed in double precision) times the tile's
// factor.
template <typename T, int RB>
__device__ __forceinline__ void dtab_apply_f(typename Vec2<T>::type *a, const T *tab, typename Vec2<T>::type f) {
  using V = typename Vec2<T>::type;
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    V d;
    d.x = tab[2 * r];
    d.y = tab[2 * r + 1];
    a[r] = cmul(a[r], cmul(d, f));
  }
}

// Per-tile factor of a diagonal: product over the 4-bit groups of the tile
// number (wave-uniform reads of a read-only table: scalar loads).
// The host pads the table to kMaxTileGroups groups (unit entries past
// ngrp), so every group is read unconditionally: the loads are independent
// and issue together.
template <typename T>
__device__ __forceinline__ typename Vec2<T>::type tile_factor(const typename Vec2<T>::type *tab,
                                                              uint32_t ngrp, uint64_t tile) {
  using V = typename Vec2<T>::type;
  (void)ngrp;
  V e[kMaxTileGroups];
#pragma unroll
  for (int g = 0; g < kMaxTileGroups; ++g) e[g] = tab[g * 16 + ((tile >> (4 * g)) & 15u)];
  V f = e[0];
#pragma unroll
  for (int g = 1; g < kMaxTileGroups; ++g) f = cmul(f, e[g]);
  return f;
}

// The deferred diagonal of a fused sweep on one thread's amplitudes: ix[w]
// is window w's table index without the register bits; windows [0, ph_rdep)
// also see the register bits (index ix[w] ^ ph_rimg[w][r]). All kFusedWins
// slots are read (the host points slots past nwin at a unit entry), so the
// thread's window reads issue together and multiply into one factor; each
// register-dependent window's reads then issue in groups of eight before
// their products.
template <typename T, int RB>
__device__ __forceinline__ void fused_diag(typename Vec2<T>::type (&a)[1 << RB], const FusedRead &F,
                                           const typename Vec2<T>::type *lut,
                                           const uint32_t (&ix)[kFusedWins]) {
  using V = typename Vec2<T>::type;
  constexpr int R = 1 << RB;
  constexpr int G = R < 8 ? R : 8;
  const uint32_t nrdep = F.ph_rdep;
  V e[kFusedWins];
#pragma unroll
  for (int w = 0; w < kFusedWins; ++w) e[w] = lut[F.ph.win[w].off + ix[w]];
  V pt;
  pt.x = T(1);
  pt.y = T(0);
#pragma unroll
  for (int w = 0; w < kFusedWins; ++w) {
    const V f = cmul(pt, e[w]);
    const bool use = static_cast<uint32_t>(w) >= nrdep;
    pt.x = use ? f.x : pt.x;
    pt.y = use ? f.y : pt.y;
  }
#pragma unroll
  for (int r = 0; r < R; ++r) a[r] = cmul(a[r], pt);
#pragma unroll
  for (int w = 0; w < kFusedRdep; ++w) {
    if (static_cast<uint32_t>(w) < nrdep) {
      const uint32_t off = F.ph.win[w].off, b = ix[w];
#pragma unroll
      for (int h = 0; h < R; h += G) {
        V g[G];
#pragma unroll
        for (int k = 0; k < G; ++k) g[k] = lut[off + (b ^ F.ph_rimg[w][h + k])];
#pragma unroll
        for (int k = 0; k < G; ++k) a[h + k] = cmul(a[h + k], g[k]);
      }
    }
  }
}

// Fused read of deferred work (FusedRead), per block: the diagonal's
// tables into LDS, and the thread parts of the source index and of each
// window's table index. Both are linear (over XOR) in the thread bits, so
// each tile adds its wave-uniform part with one XOR.
template <typename T>
__device__ __forceinline__ void fused_setup(const FusedRead &F, const TileSeg &s0, uint32_t tid,
                                            const typename Vec2<T>::type *__restrict__ ph_lut,
                                            typename Vec2<T>::type *lut, uint64_t &th,
                                            uint32_t (&thw)[kFusedWins]) {
  for (uint32_t i = tid; i < F.ph.lut_len; i += blockDim.x) lut[i] = ph_lut[i];
  const uint32_t tm = tid & s0.tmask;
  th = 0;
#pragma unroll
  for (int k = 0; k < kTileThreadBits; ++k)
    if ((tm >> k) & 1u) th ^= F.pthr[k];
#pragma unroll
  for (int w = 0; w < kFusedWins; ++w)
    thw[w] = (w < static_cast<int>(F.ph.nwin)) ? win_bits(F.ph.win[w], th) : 0u;
  __syncthreads();
}

// Fused read, per tile: segment 0's amplitudes from src through the
// deferred permutation, times the deferred diagonal. Windows [0, ph_rdep)
// see the register bits (the host orders them first; their register parts
// are in ph_rimg), the others give one product per thread.
template <typename T, int RB>
__device__ __forceinline__ void fused_load(typename Vec2<T>::type (&a)[1 << RB],
                                           const typename Vec2<T>::type *src,
                                           const FusedRead &F, const uint64_t *__restrict__ ptab,
                                           uint64_t tile, uint64_t th, const uint32_t (&thw)[kFusedWins],
                                           const typename Vec2<T>::type *lut) {
  using V = typename Vec2<T>::type;
  constexpr int R = 1 << RB;
  // image of the tile number: four byte tables (scalar loads)
  const uint64_t tsb = F.pc ^ ptab[tile & 255u] ^ ptab[256 + ((tile >> 8) & 255u)] ^
                       ptab[512 + ((tile >> 16) & 255u)] ^ ptab[768 + ((tile >> 24) & 255u)];
  const uint64_t sb = tsb ^ th;
  uint64_t ri[R];
  ri[0] = 0;
#pragma unroll
  for (int r = 1; r < R; ++r) ri[r] = ri[r & (r - 1)] ^ F.preg[__builtin_ctz(r)];
#pragma unroll
  for (int r = 0; r < R; ++r) a[r] = ld_amp(src + (sb ^ ri[r]));
  if (F.ph.nwin) {
    uint32_t ix[kFusedWins];
#pragma unroll
    for (int w = 0; w < kFusedWins; ++w) ix[w] = win_bits(F.ph.win[w], tsb) ^ thw[w];
    fused_diag<T, RB>(a, F, lut, ix);
  }
}

// PERM: the first segment reads through deferred work (fused_load), with
// the deferred diagonal's window tables in LDS behind the exchange buffer.
// Reads come from `src` and stores go to `st`: the other buffer when a
// permutation is read, so no tile reads what another tile writes; the same
// buffer for a lone diagonal, where each tile reads exactly the amplitudes
// it writes, all of them before its first store.
// Waves per SIMD a rotation or QFT sweep must keep: complex64 tiles with
// 4 register bits run 512-thread blocks, two per CU (64 KiB of LDS each),
// which needs 4 waves per SIMD, i.e. at most 128 VGPRs.
template <typename T, int RB>
struct SweepWaves {
  static constexpr int v = (RB == 4 && sizeof(T) == 4) ? 4 : 1;
};

template <typename T, int RB, bool PERM>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(SweepWaves<T, RB>::v))) k_rot(typename Vec2<T>::type *st,
                                             const typename Vec2<T>::type *__restrict__ ttab,
                                             const RotArgs<T> A,
                                             const typename Vec2<T>::type *src,
                                             const typename Vec2<T>::type *__restrict__ ph_lut,
                                             const uint64_t *__restrict__ ptab) {
  using V = typename Vec2<T>::type;
  constexpr int R = 1 << RB;
  extern __shared__ unsigned char smem_raw[];
  V *lds = reinterpret_cast<V *>(smem_raw);
  const uint32_t tid = threadIdx.x;
  uint64_t vrank = blockIdx.x;
  if (A.xcd_groups > 1 && gridDim.x % A.xcd_groups == 0)
    vrank = (blockIdx.x % A.xcd_groups) * (gridDim.x / A.xcd_groups) + blockIdx.x / A.xcd_groups;
  V *lut = lds + (A.nseg > 1 ? (1u << A.nlocal) : 0u);
  uint64_t th = 0;
  uint32_t thw[kFusedWins];
  if constexpr (PERM) fused_setup<T>(A.fr, A.seg[0], tid, ph_lut, lut, th, thw);
  const uint32_t last = A.nseg - 1;
  for (uint64_t tile = vrank; tile < A.n_tiles; tile += gridDim.x) {
    const uint64_t base = tile_base(A, tile);
    V a[R];
    for (uint32_t s = 0;; ++s) {
      const TileSeg &sg = A.seg[s];
      const SegPos p = seg_pos<RB>(sg, tid, base);
      if (s == 0) {
        if constexpr (PERM) {
          fused_load<T, RB>(a, src, A.fr, ptab, tile, th, thw, lut);
        } else {
          uint64_t gstr[RB];
#pragma unroll
          for (int j = 0; j < RB; ++j) gstr[j] = sg.gstr[j];
#pragma unroll
          for (int r = 0; r < R; ++r) a[r] = ld_amp(st + p.gb + reg_goff<RB>(r, gstr));
        }
        if (A.has_pre)
          dtab_apply_f<T, RB>(a, A.pre, cmul(ttab[kThrTabOff + p.tm], tile_factor<T>(ttab, A.ntgrp, tile)));
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
          a[r] = lds[p.tsw ^ lds_swz<RB>(static_cast<uint32_t>(r) << sg.reg0)];
      }
      const uint32_t gm = A.gmask[s];
#pragma unroll
      for (int j = 0; j < RB; ++j) {
        if ((gm >> j) & 1u) {
          const T sa = A.shear[s][j][0], sb2 = A.shear[s][j][1];
          switch (j) {
            case 0: shear_pairs<T, RB, 0>(a, sa, sb2); break;
            case 1: if constexpr (RB > 1) shear_pairs<T, RB, 1>(a, sa, sb2); break;
            case 2: if constexpr (RB > 2) shear_pairs<T, RB, 2>(a, sa, sb2); break;
            case 3: if constexpr (RB > 3) shear_pairs<T, RB, 3>(a, sa, sb2); break;
            case 4: if constexpr (RB > 4) shear_pairs<T, RB, 4>(a, sa, sb2); break;
            default: break;
          }
        }
      }
      // lane bits 4 / 5 into register bits 0 / 1, and their rotations
      if constexpr (RB >= 2) {
        const uint32_t sw = A.swp[s];
        if (sw & 1u) {
          lane_swap<T, RB, 0, 4>(a);
          shear_pairs<T, RB, 0>(a, A.shear2[s][0][0], A.shear2[s][0][1]);
        }
        if (sw & 2u) {
          lane_swap<T, RB, 1, 5>(a);
          shear_pairs<T, RB, 1>(a, A.shear2[s][1][0], A.shear2[s][1][1]);
        }
      }
      if (s == last) break;
      // the LDS write in the layout after the swaps
      const uint32_t tw = lds_swz<RB>(p.tl ^ ((tid & 16u) ? A.tfix[s][0] : 0u) ^ ((tid & 32u) ? A.tfix[s][1] : 0u));
      uint32_t rw[RB];
#pragma unroll
      for (int j = 0; j < RB; ++j) rw[j] = lds_swz<RB>(A.wimg[s][j]);
      __syncthreads();
#pragma unroll
      for (int r = 0; r < R; ++r) {
        uint32_t ro = 0;
#pragma unroll
        for (int j = 0; j < RB; ++j)
          if ((r >> j) & 1) ro ^= rw[j];
        lds[tw ^ ro] = a[r];
      }
      __syncthreads();
    }
    {
      // the last segment's layout after its swaps
      const TileSeg &sg = A.out;
      const SegPos p = seg_pos<RB>(sg, tid, base);
      if (A.has_post)
        dtab_apply_f<T, RB>(a, A.post,
                            cmul(ttab[kThrTabOff + kThrTabLen + p.tm],
                                 tile_factor<T>(ttab + kMaxTileGroups * 16, A.ntgrp, tile)));
      uint64_t gstr[RB];
#pragma unroll
      for (int j = 0; j < RB; ++j) gstr[j] = sg.gstr[j];
#pragma unroll
      for (int r = 0; r < R; ++r) st_amp(st + p.gb + reg_goff<RB>(r, gstr), a[r]);
    }
  }
}

// Rotation sweep across a deferred permutation: the pending rotations of one
// 1q layer (segments [0, so], old frame, tiles of src on coordinate bits),
// that layer's post diagonal and the deferred diagonal at segment so, then
// one exchange that writes each amplitude to its slot in the new frame
// (local index Q u ^ delta, Q linear: the slot is an XOR of per-bit images,
// swizzled the same way as every exchange), then the next layer's rotations
// on the head bits (segments (so, last], their pre diagonal first). The new
// frame's tile is the image of the old one under the permutation; its bits
// are the head bits plus "extra" vectors, and its store address is
// rep(tile) ^ the images of the last segment's bits. Tile numbers map to
// old-frame bases and to rep | delta << 48 through byte tables (ptab[1024..]
// and ptab[0..1023]): the host picks tile-number bits whose images are short
// in both frames, so that the tiles running together read and write nearby
// rows. Both diagonals are local to the tile (no per-tile factor). Reads and writes go to
// different buffers when a permutation is pending (each output tile is the
// image of exactly one input tile, so no tile reads what another writes).
template <typename T, int RB>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(SweepWaves<T, RB>::v)))
k_rot_relabel(typename Vec2<T>::type *dst, const RotArgs<T> A, const typename Vec2<T>::type *src,
              const typename Vec2<T>::type *__restrict__ ph_lut, const uint64_t *__restrict__ ptab) {
  using V = typename Vec2<T>::type;
  constexpr int R = 1 << RB;
  extern __shared__ unsigned char smem_raw[];
  V *lds = reinterpret_cast<V *>(smem_raw);
  const uint32_t tid = threadIdx.x;
  uint64_t vrank = blockIdx.x;
  if (A.xcd_groups > 1 && gridDim.x % A.xcd_groups == 0)
    vrank = (blockIdx.x % A.xcd_groups) * (gridDim.x / A.xcd_groups) + blockIdx.x / A.xcd_groups;
  const FusedRead &F = A.fr;
  V *lut = lds + (1u << A.nlocal);
  for (uint32_t i = tid; i < F.ph.lut_len; i += blockDim.x) lut[i] = ph_lut[i];
  const uint32_t so = A.rl_seg, last = A.nseg - 1;
  __syncthreads();
  for (uint64_t tile = vrank; tile < A.n_tiles; tile += gridDim.x) {
    const uint64_t base = ptab[1024 + (tile & 255u)] ^ ptab[1280 + ((tile >> 8) & 255u)] ^
                          ptab[1536 + ((tile >> 16) & 255u)] ^ ptab[1792 + ((tile >> 24) & 255u)];
    V a[R];
    for (uint32_t s = 0;; ++s) {
      const TileSeg &sg = A.seg[s];
      // the index bits only where they are used (the load, the windows)
      SegPos p;
      p.tm = tid & sg.tmask;
      p.tl = seg_local(sg, p.tm);
      p.tsw = lds_swz<RB>(p.tl);
      if (s == 0) {
        const uint64_t gb = seg_index(sg, p.tm, base);
        uint64_t gstr[RB];
#pragma unroll
        for (int j = 0; j < RB; ++j) gstr[j] = sg.gstr[j];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = ld_amp(src + gb + reg_goff<RB>(r, gstr));
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
          a[r] = lds[p.tsw ^ lds_swz<RB>(static_cast<uint32_t>(r) << sg.reg0)];
        if (s == so + 1 && A.has_pre) {
          // the thread part is read per tile (a table hit): kept across tiles it spills registers
          dtab_apply_f<T, RB>(a, A.pre, A.tile_tab[kThrTabOff + p.tm]);
        }
      }
      const uint32_t gm = A.gmask[s];
#pragma unroll
      for (int j = 0; j < RB; ++j) {
        if ((gm >> j) & 1u) {
          const T sa = A.shear[s][j][0], sb2 = A.shear[s][j][1];
          switch (j) {
            case 0: shear_pairs<T, RB, 0>(a, sa, sb2); break;
            case 1: if constexpr (RB > 1) shear_pairs<T, RB, 1>(a, sa, sb2); break;
            case 2: if constexpr (RB > 2) shear_pairs<T, RB, 2>(a, sa, sb2); break;
            case 3: if constexpr (RB > 3) shear_pairs<T, RB, 3>(a, sa, sb2); break;
            case 4: if constexpr (RB > 4) shear_pairs<T, RB, 4>(a, sa, sb2); break;
            default: break;
          }
        }
      }
      if (s == last) break;
      if (s == so) {
        if (A.has_post) {
          dtab_apply_f<T, RB>(a, A.post, A.tile_tab[kThrTabOff + kThrTabLen + p.tm]);
        }
        if (F.ph.nwin) {
          // this thread's index in segment so (old frame) without its register bits
          const uint64_t gb = seg_index(sg, p.tm, base);
          uint32_t ix[kFusedWins];
#pragma unroll
          for (int w = 0; w < kFusedWins; ++w) ix[w] = win_bits(F.ph.win[w], gb);
          fused_diag<T, RB>(a, F, lut, ix);
        }
        const uint64_t e = F.pc ^ ptab[tile & 255u] ^ ptab[256 + ((tile >> 8) & 255u)] ^
                           ptab[512 + ((tile >> 16) & 255u)] ^ ptab[768 + ((tile >> 24) & 255u)];
        uint32_t qt = static_cast<uint32_t>(e >> 48);
#pragma unroll
        for (int k = 0; k < kTileThreadBits; ++k)
          if ((p.tm >> k) & 1u) qt ^= A.rl_thr[k];
        const uint32_t wbase = lds_swz<RB>(qt);
        uint32_t ri[R];
        ri[0] = 0;
#pragma unroll
        for (int r = 1; r < R; ++r) ri[r] = ri[r & (r - 1)] ^ A.rl_reg[__builtin_ctz(r)];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) lds[wbase ^ lds_swz<RB>(ri[r])] = a[r];
        __syncthreads();
      } else {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) lds[p.tsw ^ lds_swz<RB>(static_cast<uint32_t>(r) << sg.reg0)] = a[r];
        __syncthreads();
      }
    }
    {
      const uint64_t e = F.pc ^ ptab[tile & 255u] ^ ptab[256 + ((tile >> 8) & 255u)] ^
                         ptab[512 + ((tile >> 16) & 255u)] ^ ptab[768 + ((tile >> 24) & 255u)];
      uint64_t sb = e & ((1ull << 48) - 1ull);
      const uint32_t tm = tid & A.seg[last].tmask;
#pragma unroll
      for (int k = 0; k < kTileThreadBits; ++k)
        if ((tm >> k) & 1u) sb ^= F.pthr[k];
      uint64_t gi[R];
      gi[0] = 0;
#pragma unroll
      for (int r = 1; r < R; ++r) gi[r] = gi[r & (r - 1)] ^ F.preg[__builtin_ctz(r)];
#pragma unroll
      for (int r = 0; r < R; ++r) st_amp(dst + (sb ^ gi[r]), a[r]);
    }
  }
}

// ---------------------------------------------------------------------------
// QFT sweep
// ---------------------------------------------------------------------------

// exp(i pi k / 16), k = 0 .. 15, as compile-time constants.
__host__ __device__ constexpr double cos16(int k) {
  constexpr double c[9] = {1.0,
                           0.98078528040323044912618223613424,
                           0.92387953251128675612818318939679,
                           0.83146961230254523707878837761791,
                           0.70710678118654752440084436210485,
                           0.55557023301960222474283081394853,
                           0.38268343236508977172845998403040,
                           0.19509032201612826784828486847702,
                           0.0};
  return k <= 8 ? c[k] : -c[16 - k];
}
__host__ __device__ constexpr double sin16(int k) { return k <= 8 ? cos16(8 - k) : cos16(k - 8); }

// In-place complex product a *= (c, s) with constant or wave-uniform factors.
template <typename V, typename T>
__device__ __forceinline__ void cmul_in(V &a, T c, T s) {
  const T t = a.x * s;
  a.x = fma_t(a.x, c, -(a.y * s));
  a.y = fma_t(a.y, c, t);
}

// QFT stage on register bit J with thread twiddle wt: forward: butterfly,
// then the |1> branch times exp(i pi v / 2^J) * wt (v = register bits below
// J); inverse: the conjugate phase first, then the butterfly. All steps are
// in place (a0 + a1, then a1 <- a0' - 2 a1), so the stage needs no register
// moves; the 1/sqrt(2) factors are applied at the store.
template <typename T, int RB, int J, bool INV>
__device__ __forceinline__ void qft_stage(typename Vec2<T>::type *a, typename Vec2<T>::type wt) {
  using V = typename Vec2<T>::type;
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    const int r1 = r | (1 << J);
    constexpr int kShift = 4 - J;
    const int k = (r & ((1 << J) - 1)) << kShift;  // exp(i pi v / 2^J) = cis16(k)
    V &a0 = a[r];
    V &a1 = a[r1];
    if (!INV) {
      a0.x = a0.x + a1.x;
      a0.y = a0.y + a1.y;
      a1.x = fma_t(T(-2), a1.x, a0.x);
      a1.y = fma_t(T(-2), a1.y, a0.y);
      if (k != 0) cmul_in(a1, static_cast<T>(cos16(k)), static_cast<T>(sin16(k)));
      cmul_in(a1, wt.x, wt.y);
    } else {
      cmul_in(a1, wt.x, -wt.y);
      if (k != 0) cmul_in(a1, static_cast<T>(cos16(k)), static_cast<T>(-sin16(k)));
      a0.x = a0.x + a1.x;
      a0.y = a0.y + a1.y;
      a1.x = fma_t(T(-2), a1.x, a0.x);
      a1.y = fma_t(T(-2), a1.y, a0.y);
    }
  }
}

template <typename T, int RB, bool INV>
__device__ __forceinline__ void qft_segment(typename Vec2<T>::type *a, const QftArgs<T> &A, uint32_t s,
                                            uint64_t gb) {
  using V = typename Vec2<T>::type;
  const uint32_t sm = A.smask[s];
  const uint32_t xb = A.xbits[s];
  const uint64_t x = (gb >> A.lb) & ((xb >= 64) ? ~0ull : ((1ull << xb) - 1ull));
  // thread twiddles, top stage first; a stage directly below an active one
  // squares its twiddle (exp(i pi x / 2^(w-1)) = exp(i pi x / 2^w)^2)
  V wt[RB];
#pragma unroll
  for (int j = RB - 1; j >= 0; --j) {
    if ((sm >> j) & 1u) {
      if (j + 1 < RB && ((sm >> (j + 1)) & 1u)) {
        const V h = wt[j + 1 < RB ? j + 1 : j];
        wt[j].x = fma_t(h.x, h.x, -(h.y * h.y));
        wt[j].y = T(2) * h.x * h.y;
      } else {
        wt[j] = CisPi<T>::eval(x, A.wbits[s][j]);
      }
    }
  }
  if (!INV) {
#pragma unroll
    for (int j = RB - 1; j >= 0; --j) {
      if ((sm >> j) & 1u) {
        switch (j) {
          case 0: qft_stage<T, RB, 0, false>(a, wt[0]); break;
          case 1: if constexpr (RB > 1) qft_stage<T, RB, 1, false>(a, wt[j]); break;
          case 2: if constexpr (RB > 2) qft_stage<T, RB, 2, false>(a, wt[j]); break;
          case 3: if constexpr (RB > 3) qft_stage<T, RB, 3, false>(a, wt[j]); break;
          case 4: if constexpr (RB > 4) qft_stage<T, RB, 4, false>(a, wt[j]); break;
          default: break;
        }
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      if ((sm >> j) & 1u) {
        switch (j) {
          case 0: qft_stage<T, RB, 0, true>(a, wt[0]); break;
          case 1: if constexpr (RB > 1) qft_stage<T, RB, 1, true>(a, wt[j]); break;
          case 2: if constexpr (RB > 2) qft_stage<T, RB, 2, true>(a, wt[j]); break;
          case 3: if constexpr (RB > 3) qft_stage<T, RB, 3, true>(a, wt[j]); break;
          case 4: if constexpr (RB > 4) qft_stage<T, RB, 4, true>(a, wt[j]); break;
          default: break;
        }
      }
    }
  }
}

template <typename T, int RB, bool INV, bool PERM>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(SweepWaves<T, RB>::v))) k_qft(typename Vec2<T>::type *st,
                                             const QftArgs<T> A,
                                             const typename Vec2<T>::type *src,
                                             const typename Vec2<T>::type *__restrict__ ph_lut,
                                             const uint64_t *__restrict__ ptab) {
  using V = typename Vec2<T>::type;
  constexpr int R = 1 << RB;
  extern __shared__ unsigned char smem_raw[];
  V *lds = reinterpret_cast<V *>(smem_raw);
  const uint32_t tid = threadIdx.x;
  uint64_t vrank = blockIdx.x;
  if (A.xcd_groups > 1 && gridDim.x % A.xcd_groups == 0)
    vrank = (blockIdx.x % A.xcd_groups) * (gridDim.x / A.xcd_groups) + blockIdx.x / A.xcd_groups;
  V *lut = lds + (A.nseg > 1 ? (1u << A.nlocal) : 0u);
  uint64_t th = 0;
  uint32_t thw[kFusedWins];
  if constexpr (PERM) fused_setup<T>(A.fr, A.seg[0], tid, ph_lut, lut, th, thw);
  const uint32_t last = A.nseg - 1;
  for (uint64_t tile = vrank; tile < A.n_tiles; tile += gridDim.x) {
    const uint64_t base = tile_base(A, tile);
    V a[R];
    for (uint32_t s = 0;; ++s) {
      const TileSeg &sg = A.seg[s];
      const SegPos p = seg_pos<RB>(sg, tid, base);
      if (s == 0) {
        if constexpr (PERM) {
          fused_load<T, RB>(a, src, A.fr, ptab, tile, th, thw, lut);
        } else {
          uint64_t gstr[RB];
#pragma unroll
          for (int j = 0; j < RB; ++j) gstr[j] = sg.gstr[j];
#pragma unroll
          for (int r = 0; r < R; ++r) a[r] = ld_amp(st + p.gb + reg_goff<RB>(r, gstr));
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r)
          a[r] = lds[p.tsw ^ lds_swz<RB>(static_cast<uint32_t>(r) << sg.reg0)];
      }
      if (A.smask[s]) qft_segment<T, RB, INV>(a, A, s, p.gb);
      if (s == last) break;
      __syncthreads();
#pragma unroll
      for (int r = 0; r < R; ++r) lds[p.tsw ^ lds_swz<RB>(static_cast<uint32_t>(r) << sg.reg0)] = a[r];
      __syncthreads();
    }
    {
      const TileSeg &sg = A.seg[last];
      const SegPos p = seg_pos<RB>(sg, tid, base);
#pragma unroll
      for (int r = 0; r < R; ++r) {
        a[r].x *= A.scale;
        a[r].y *= A.scale;
      }
      uint64_t gstr[RB];
#pragma unroll
      for (int j = 0; j < RB; ++j) gstr[j] = sg.gstr[j];
#pragma unroll
      for (int r = 0; r < R; ++r) st_amp(st + p.gb + reg_goff<RB>(r, gstr), a[r]);
    }
  }
}

// ---------------------------------------------------------------------------
// diagonal phase layer: product of per-window lookup tables
// ---------------------------------------------------------------------------

template <typename T> struct Pack;
template <> struct Pack<float> { using P = float4; static constexpr int kAmps = 2; };
template <> struct Pack<double> { using P = double2; static constexpr int kAmps = 1; };


template <typename T>
__device__ __forceinline__ void load_pack(const typename Vec2<T>::type *st, uint64_t pk,
                                          typename Vec2<T>::type *amp) {
  if constexpr (Pack<T>::kAmps == 2) {
    const float4 p = ld_f4(reinterpret_cast<const float4 *>(st) + pk);
    amp[0] = make_float2(p.x, p.y);
    amp[1] = make_float2(p.z, p.w);
  } else {
    amp[0] = ld_amp(st + pk);
  }
}

template <typename T>
__device__ __forceinline__ void store_pack(typename Vec2<T>::type *st, uint64_t pk,
                                           const typename Vec2<T>::type *amp) {
  if constexpr (Pack<T>::kAmps == 2) {
    st_f4(reinterpret_cast<float4 *>(st) + pk, make_float4(amp[0].x, amp[0].y, amp[1].x, amp[1].y));
  } else {
    st_amp(st + pk, amp[0]);
  }
}

// Streaming kernels walk the packs in chunks of kStreamThreads x kChunkPer
// consecutive packs (64 KiB of complex64), one chunk per block and
// iteration, in the XCD-grouped order of the tile sweeps; within a chunk a
// thread keeps kUnroll packs in flight and every load instruction of a
// wave reads consecutive addresses.
constexpr int kUnroll = 4;
constexpr int kStreamThreads = 256;
constexpr int kChunkPer = 16;
constexpr uint64_t kChunkPacks = static_cast<uint64_t>(kStreamThreads) * kChunkPer;
static_assert(kChunkPacks == kStreamChunkPacks, "streaming chunk size");
constexpr uint32_t kStreamXcdGroups = 8;  // gfx950: 8 XCDs

__device__ __forceinline__ uint64_t stream_vrank(uint32_t groups) {
  uint64_t vrank = blockIdx.x;
  if (groups > 1 && gridDim.x % groups == 0)
    vrank = (blockIdx.x % groups) * (gridDim.x / groups) + blockIdx.x / groups;
  return vrank;
}

template <typename T>
__global__ void __launch_bounds__(256) k_phase(typename Vec2<T>::type *__restrict__ st,
                                               const typename Vec2<T>::type *__restrict__ lut_g,
                                               const PhaseArgs A) {
  using V = typename Vec2<T>::type;
  extern __shared__ unsigned char smem_raw[];
  V *lut = reinterpret_cast<V *>(smem_raw);
  for (uint32_t i = threadIdx.x; i < A.lut_len; i += blockDim.x) lut[i] = lut_g[i];
  __syncthreads();
  constexpr int KA = Pack<T>::kAmps;
  const uint64_t npack = A.n_amps / KA;
  const uint64_t nchunks = (npack + kChunkPacks - 1) / kChunkPacks;
  for (uint64_t c = stream_vrank(A.xcd_groups); c < nchunks; c += gridDim.x) {
#pragma unroll 1
    for (int k0 = 0; k0 < kChunkPer; k0 += kUnroll) {
      const uint64_t base = c * kChunkPacks + static_cast<uint64_t>(k0) * kStreamThreads + threadIdx.x;
      V amp[kUnroll][KA];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const uint64_t pk = base + static_cast<uint64_t>(u) * kStreamThreads;
        if (pk < npack) load_pack<T>(st, pk, amp[u]);
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const uint64_t pk = base + static_cast<uint64_t>(u) * kStreamThreads;
        if (pk < npack) {
          const uint64_t i0 = pk * KA;
          V ph;
          ph.x = T(1);
          ph.y = T(0);
          for (uint32_t w = A.nlow; w < A.nwin; ++w) ph = cmul(ph, lut[win_index(A.win[w], i0)]);
#pragma unroll
          for (int d = 0; d < KA; ++d) {
            V p = ph;
            for (uint32_t w = 0; w < A.nlow; ++w) p = cmul(p, lut[win_index(A.win[w], i0 + d)]);
            amp[u][d] = cmul(amp[u][d], p);
          }
          store_pack<T>(st, pk, amp[u]);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// GF(2) affine gather: out[y] = in[c ^ T0[y&255] ^ T1[(y>>8)&255] ^ ...]
// ---------------------------------------------------------------------------

// PH: the gather also applies a deferred diagonal (a phase layer given as
// window lookup tables, as k_phase evaluates it) at the source index:
// out[y] = D(src(y)) in[src(y)], i.e. the phase layer followed by the
// permutation in one sweep.
template <typename T, bool PH>
__global__ void __launch_bounds__(256) k_xor_gather(const typename Vec2<T>::type *__restrict__ src,
                                                    typename Vec2<T>::type *__restrict__ dst,
                                                    uint64_t n_amps,
                                                    const uint64_t *__restrict__ tables, int ntab,
                                                    uint64_t c, uint32_t groups,
                                                    const typename Vec2<T>::type *__restrict__ lut_g,
                                                    const PhaseArgs P) {
  using V = typename Vec2<T>::type;
  __shared__ uint64_t tab[8 * 256];
  extern __shared__ unsigned char smem_raw[];
  V *lut = reinterpret_cast<V *>(smem_raw);
  for (int i = threadIdx.x; i < ntab * 256; i += blockDim.x) tab[i] = tables[i];
  if constexpr (PH)
    for (uint32_t i = threadIdx.x; i < P.lut_len; i += blockDim.x) lut[i] = lut_g[i];
  __syncthreads();
  constexpr int KA = Pack<T>::kAmps;
  const uint64_t npack = n_amps / KA;
  const uint64_t nchunks = (npack + kChunkPacks - 1) / kChunkPacks;
  for (uint64_t ch = stream_vrank(groups); ch < nchunks; ch += gridDim.x) {
#pragma unroll 1
    for (int k0 = 0; k0 < kChunkPer; k0 += kUnroll) {
      const uint64_t base = ch * kChunkPacks + static_cast<uint64_t>(k0) * kStreamThreads + threadIdx.x;
      V amp[kUnroll][KA];
      uint64_t sidx[kUnroll][KA];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const uint64_t pk = base + static_cast<uint64_t>(u) * kStreamThreads;
        if (pk < npack) {
          const uint64_t y0 = pk * KA;
          uint64_t s0 = c;
          for (int t = 0; t < ntab; ++t) s0 ^= tab[t * 256 + ((y0 >> (8 * t)) & 255u)];
          sidx[u][0] = s0;
          amp[u][0] = ld_amp(src + s0);
          // y0 is even when KA = 2: src(y0 + 1) = src(y0) ^ column of index bit 0
          if constexpr (KA == 2) {
            sidx[u][1] = s0 ^ tab[1];
            amp[u][1] = ld_amp(src + sidx[u][1]);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const uint64_t pk = base + static_cast<uint64_t>(u) * kStreamThreads;
        if (pk < npack) {
          if constexpr (PH) {
#pragma unroll
            for (int d = 0; d < KA; ++d) {
              V ph;
              ph.x = T(1);
              ph.y = T(0);
              for (uint32_t w = 0; w < P.nwin; ++w) ph = cmul(ph, lut[win_index(P.win[w], sidx[u][d])]);
              amp[u][d] = cmul(amp[u][d], ph);
            }
          }
          store_pack<T>(dst, pk, amp[u]);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// elementwise helpers
// ---------------------------------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(256) k_init_basis(typename Vec2<T>::type *__restrict__ st,
                                                    uint64_t n_amps, uint64_t index) {
  using V = typename Vec2<T>::type;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_amps;
       i += stride) {
    V v;
    v.x = (i == index) ? T(1) : T(0);
    v.y = T(0);
    st[i] = v;
  }
}

// Product state: amplitude i = lo[i & (2^l1 - 1)] * mid[(i >> l1) & (2^l2 - 1)]
// * hi[i >> (l1 + l2)], the per-bit factors of each index-bit group
// multiplied out on the host. One write pass (lo and mid from LDS, hi a
// broadcast read: it changes every 2^(l1+l2) amplitudes).
template <typename T>
__global__ void __launch_bounds__(256) k_product(typename Vec2<T>::type *__restrict__ st, uint64_t n_amps,
                                                 const typename Vec2<T>::type *__restrict__ lo,
                                                 const typename Vec2<T>::type *__restrict__ mid,
                                                 const typename Vec2<T>::type *__restrict__ hi, uint32_t l1,
                                                 uint32_t l2, uint32_t groups) {
  using V = typename Vec2<T>::type;
  __shared__ V tlo[1u << kProductGroupBits], tmid[1u << kProductGroupBits];
  for (uint32_t i = threadIdx.x; i < (1u << l1); i += blockDim.x) tlo[i] = lo[i];
  for (uint32_t i = threadIdx.x; i < (1u << l2); i += blockDim.x) tmid[i] = mid[i];
  __syncthreads();
  const uint64_t m1 = (1ull << l1) - 1ull, m2 = (1ull << l2) - 1ull;
  const uint32_t sh = l1 + l2;
  constexpr int KA = Pack<T>::kAmps;
  const uint64_t npack = n_amps / KA;
  const uint64_t nchunks = (npack + kChunkPacks - 1) / kChunkPacks;
  for (uint64_t ch = stream_vrank(groups); ch < nchunks; ch += gridDim.x) {
#pragma unroll 1
    for (int k0 = 0; k0 < kChunkPer; k0 += kUnroll) {
      const uint64_t base = ch * kChunkPacks + static_cast<uint64_t>(k0) * kStreamThreads + threadIdx.x;
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        const uint64_t pk = base + u * kStreamThreads;
        if (pk < npack) {
          V amp[KA];
#pragma unroll
          for (int d = 0; d < KA; ++d) {
            const uint64_t i = pk * KA + d;
            amp[d] = cmul(cmul(tlo[i & m1], tmid[(i >> l1) & m2]), hi[i >> sh]);
          }
          store_pack<T>(st, pk, amp);
        }
      }
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_scale(typename Vec2<T>::type *__restrict__ st,
                                               uint64_t n_amps, T re, T im, uint32_t groups) {
  using V = typename Vec2<T>::type;
  V f;
  f.x = re;
  f.y = im;
  constexpr int KA = Pack<T>::kAmps;
  const uint64_t npack = n_amps / KA;
  const uint64_t nchunks = (npack + kChunkPacks - 1) / kChunkPacks;
  for (uint64_t ch = stream_vrank(groups); ch < nchunks; ch += gridDim.x) {
#pragma unroll 1
    for (int k0 = 0; k0 < kChunkPer; k0 += kUnroll) {
      const uint64_t base = ch * kChunkPacks + static_cast<uint64_t>(k0) * kStreamThreads + threadIdx.x;
      V amp[kUnroll][KA];
#pragma unroll
      for (int u = 0; u < kUnroll; ++u)
        if (base + u * kStreamThreads < npack) load_pack<T>(st, base + u * kStreamThreads, amp[u]);
#pragma unroll
      for (int u = 0; u < kUnroll; ++u) {
        if (base + u * kStreamThreads < npack) {
#pragma unroll
          for (int d = 0; d < KA; ++d) amp[u][d] = cmul(amp[u][d], f);
          store_pack<T>(st, base + u * kStreamThreads, amp[u]);
        }
      }
    }
  }
}

template <typename Ts, typename Td>
__global__ void __launch_bounds__(256) k_convert(const typename Vec2<Ts>::type *__restrict__ src,
                                                 typename Vec2<Td>::type *__restrict__ dst,
                                                 uint64_t n_amps) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_amps;
       i += stride) {
    typename Vec2<Td>::type v;
    v.x = static_cast<Td>(src[i].x);
    v.y = static_cast<Td>(src[i].y);
    dst[i] = v;
  }
}

// Copy probe (calibration of achievable bandwidth and of the FETCH_SIZE /
// WRITE_SIZE counters): the access form of the engine's sweeps. Each block
// copies one contiguous 128 KiB chunk (512 threads x 16 vectors of 16 B,
// lane-consecutive), all loads in flight before the stores, nontemporal
// like the sweeps' loads and stores.
constexpr int kCopyVecs = 16;
__global__ void __launch_bounds__(512) k_copy_probe(const float4 *__restrict__ src,
                                                    float4 *__restrict__ dst, uint64_t n_vec) {
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * (512u * kCopyVecs) + threadIdx.x;
  float4 v[kCopyVecs];
#pragma unroll
  for (int k = 0; k < kCopyVecs; ++k)
    if (base + k * 512u < n_vec) v[k] = ld_f4(src + base + k * 512u);
#pragma unroll
  for (int k = 0; k < kCopyVecs; ++k)
    if (base + k * 512u < n_vec) st_f4(dst + base + k * 512u, v[k]);
}

// ---------------------------------------------------------------------------
// FP64 reductions. Summation orders are fixed so that a CPU restatement of
// the same order reproduces the sums bit for bit (sampling depends on it).
// ---------------------------------------------------------------------------

template <typename V>
__device__ __forceinline__ double amp_prob(V v) {
#pragma clang fp contract(off)
  const double re = static_cast<double>(v.x);
  const double im = static_cast<double>(v.y);
  const double a = re * re;
  const double b = im * im;
  return a + b;
}

// Lane l of the wave owning block b sums elements b*4096 + j*64 + l for
// j = 0..63 in order; the block sum adds the 64 lane sums in lane order.
template <typename T>
__device__ __forceinline__ double lane_sum(const typename Vec2<T>::type *st, uint64_t n_amps,
                                           uint64_t b, uint32_t lane) {
#pragma clang fp contract(off)
  double s = 0.0;
  const uint64_t base = b * kSampleBlock + lane;
  for (int j = 0; j < kSampleBlock / kSampleLanes; ++j) {
    const uint64_t idx = base + static_cast<uint64_t>(j) * kSampleLanes;
    const double p = (idx < n_amps) ? amp_prob(st[idx]) : 0.0;
    s = s + p;
  }
  return s;
}

template <typename T>
__global__ void __launch_bounds__(256) k_block_probsum(const typename Vec2<T>::type *__restrict__ st,
                                                       uint64_t n_amps, double *__restrict__ out,
                                                       uint64_t n_blocks) {
#pragma clang fp contract(off)
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (b >= n_blocks) return;  // uniform per wave
  const double s = lane_sum<T>(st, n_amps, b, lane);
  double tot = 0.0;
  for (int l = 0; l < 64; ++l) tot = tot + __shfl(s, l, 64);
  if (lane == 0) out[b] = tot;
}

template <typename T>
__global__ void __launch_bounds__(256) k_sample_resolve(const typename Vec2<T>::type *__restrict__ st,
                                                        uint64_t n_amps,
                                                        const uint64_t *__restrict__ shot_block,
                                                        const double *__restrict__ shot_resid,
                                                        uint64_t *__restrict__ shot_index,
                                                        uint64_t shots) {
#pragma clang fp contract(off)
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t shot = static_cast<uint64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (shot >= shots) return;  // uniform per wave
  const uint64_t b = shot_block[shot];
  const double r = shot_resid[shot];
  const double s = lane_sum<T>(st, n_amps, b, lane);
  double c = 0.0, r2 = 0.0;
  int L = -1, last_pos = -1;
  for (int l = 0; l < 64; ++l) {
    const double v = __shfl(s, l, 64);
    if (v > 0.0) last_pos = l;
    const double cn = c + v;
    if (L < 0 && cn > r) {
      L = l;
      r2 = r - c;
    }
    c = cn;
  }
  if (L < 0) {
    L = (last_pos < 0) ? 0 : last_pos;
    r2 = __builtin_inf();
  }
  if (static_cast<int>(lane) == L) {
    double c2 = 0.0;
    int J = -1, last_j = -1;
    const uint64_t base = b * kSampleBlock + lane;
    for (int j = 0; j < kSampleBlock / kSampleLanes; ++j) {
      const uint64_t idx = base + static_cast<uint64_t>(j) * kSampleLanes;
      const double p = (idx < n_amps) ? amp_prob(st[idx]) : 0.0;
      if (p > 0.0) last_j = j;
      const double cn = c2 + p;
      if (J < 0 && cn > r2) J = j;
      c2 = cn;
    }
    if (J < 0) J = (last_j < 0) ? 0 : last_j;
    shot_index[shot] = base + static_cast<uint64_t>(J) * kSampleLanes;
  }
}

// Device side of the sampler's host steps, for callers that hand over no
// host buffers but the output (msv_sample_noalloc). Both reproduce the host
// code of msv_sample bit for bit.
//
// Prefix of the block sums, in place and strictly left to right (the order
// of the host's sequential prefix): p[b + 1] holds block sum b on entry and
// p[0] + ... + (block sum b) on exit, p[0] = 0. One wave: each round loads 64
// block sums (the next round's load is issued before this round's adds), and
// every lane replays the 64 adds in lane order on wave-uniform values (read
// with readlane), keeping the partial sum at its own position.
__global__ void __launch_bounds__(64) k_prefix_seq(double *p, uint64_t nb) {
#pragma clang fp contract(off)
  const uint32_t lane = threadIdx.x;
  double c = 0.0;
  const uint64_t rounds = (nb + 63) / 64;
  double cur = (lane < nb) ? p[1 + lane] : 0.0;
  for (uint64_t r = 0; r < rounds; ++r) {
    const uint64_t i = r * 64 + lane;
    const uint64_t inext = i + 64;
    const double nxt = (inext < nb) ? p[1 + inext] : 0.0;
    const uint64_t bits = static_cast<uint64_t>(__double_as_longlong(cur));
    const int lo32 = static_cast<int>(static_cast<uint32_t>(bits));
    const int hi32 = static_cast<int>(static_cast<uint32_t>(bits >> 32));
    double mine = 0.0;
#pragma unroll
    for (int l = 0; l < 64; ++l) {
      const double v = __hiloint2double(__builtin_amdgcn_readlane(hi32, l),
                                        __builtin_amdgcn_readlane(lo32, l));
      c = c + v;
      if (static_cast<int>(lane) == l) mine = c;
    }
    if (i < nb) p[1 + i] = mine;
    cur = nxt;
  }
  if (lane == 0) p[0] = 0.0;
}

// Per shot: its uniform from the Philox stream, the block whose prefix
// interval holds it (first b with p[b + 1] > t, the host's upper_bound) and
// the residual t - p[b].
__global__ void __launch_bounds__(256) k_sample_search(const double *__restrict__ p, uint64_t nb,
                                                       double total, uint64_t seed, uint64_t stream_id,
                                                       uint64_t shot0, uint64_t count,
                                                       uint64_t *__restrict__ shot_block,
                                                       double *__restrict__ shot_resid) {
#pragma clang fp contract(off)
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  double t = msv::philox_u01(seed, stream_id, shot0 + k) * total;
  if (t >= total) t = nextafter(total, 0.0);
  uint64_t lo = 0, hi = nb;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (p[mid + 1] > t)
      hi = mid;
    else
      lo = mid + 1;
  }
  const uint64_t b = lo < nb ? lo : nb - 1;
  shot_block[k] = b;
  shot_resid[k] = t - p[b];
}

template <typename T>
__global__ void __launch_bounds__(256) k_probs_full(const typename Vec2<T>::type *__restrict__ st,
                                                    uint64_t n_amps, double *__restrict__ out) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_amps;
       i += stride)
    out[i] = amp_prob(st[i]);
}

// partials[chunk * 2^nw + k] = sum over the chunk's "other" indices o of
// p[scatter(k) | pdep(o)]; thread = (k, chunk), k fastest.
template <typename T>
__global__ void __launch_bounds__(256) k_marginal(const typename Vec2<T>::type *__restrict__ st,
                                                  int n, const uint64_t *__restrict__ out_bits,
                                                  int nw, double *__restrict__ partials,
                                                  uint64_t chunks, uint64_t chunk_len) {
#pragma clang fp contract(off)
  const uint64_t K = 1ull << nw;
  const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= K * chunks) return;
  const uint64_t k = t % K, chunk = t / K;
  uint64_t kb = 0, wmask = 0;
  for (int w = 0; w < nw; ++w) {
    const uint64_t bit = 1ull << out_bits[w];
    wmask |= bit;
    if ((k >> (nw - 1 - w)) & 1ull) kb |= bit;
  }
  const uint64_t all = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
  const uint64_t omask = all & ~wmask;
  // o0 = pdep(chunk * chunk_len, omask)
  uint64_t o = 0;
  {
    uint64_t v = chunk * chunk_len, m = omask;
    while (m) {
      if (v & 1ull) o |= m & (0ull - m);
      v >>= 1;
      m &= m - 1ull;
    }
  }
  double s = 0.0;
  for (uint64_t q = 0; q < chunk_len; ++q) {
    s = s + amp_prob(st[kb | o]);
    o = (o - omask) & omask;  // next subset of omask in increasing order
  }
  partials[chunk * K + k] = s;
}

template <typename T>
__global__ void __launch_bounds__(256) k_pauli(const typename Vec2<T>::type *__restrict__ st,
                                               uint64_t n_amps, const uint64_t *__restrict__ xm,
                                               const uint64_t *__restrict__ zm, int nwords,
                                               double *__restrict__ partials) {
#pragma clang fp contract(off)
  using V = typename Vec2<T>::type;
  __shared__ double red[2 * 256];
  double acc_re[kMaxWords], acc_im[kMaxWords];
  uint64_t xs[kMaxWords], zs[kMaxWords];
#pragma unroll
  for (int w = 0; w < kMaxWords; ++w) {
    acc_re[w] = 0.0;
    acc_im[w] = 0.0;
    xs[w] = (w < nwords) ? xm[w] : 0ull;
    zs[w] = (w < nwords) ? zm[w] : 0ull;
  }
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_amps;
       i += stride) {
    const V ai = st[i];
    const double ar = ai.x, aim = ai.y;
#pragma unroll
    for (int w = 0; w < kMaxWords; ++w) {
      if (w < nwords) {
        double br = ar, bi = aim;
        if (xs[w]) {
          const V aj = st[i ^ xs[w]];
          br = aj.x;
          bi = aj.y;
        }
        // conj(b) * a
        double tr = br * ar + bi * aim;
        double ti = br * aim - bi * ar;
        if (__popcll(i & zs[w]) & 1) {
          tr = -tr;
          ti = -ti;
        }
        acc_re[w] = acc_re[w] + tr;
        acc_im[w] = acc_im[w] + ti;
      }
    }
  }
  for (int w = 0; w < nwords; ++w) {
    red[threadIdx.x] = acc_re[w];
    red[256 + threadIdx.x] = acc_im[w];
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) {
        red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
        red[256 + threadIdx.x] = red[256 + threadIdx.x] + red[256 + threadIdx.x + s];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      partials[(static_cast<uint64_t>(blockIdx.x) * nwords + w) * 2] = red[0];
      partials[(static_cast<uint64_t>(blockIdx.x) * nwords + w) * 2 + 1] = red[256];
    }
    __syncthreads();
  }
}

template <typename Ta, typename Tb>
__global__ void __launch_bounds__(256) k_inner(const typename Vec2<Ta>::type *__restrict__ a,
                                               const typename Vec2<Tb>::type *__restrict__ b,
                                               uint64_t n_amps, double *__restrict__ partials) {
#pragma clang fp contract(off)
  __shared__ double red[2 * 256];
  double sr = 0.0, si = 0.0;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_amps;
       i += stride) {
    const double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
    const double tr = ar * br + ai * bi;
    const double ti = ar * bi - ai * br;
    sr = sr + tr;
    si = si + ti;
  }
  red[threadIdx.x] = sr;
  red[256 + threadIdx.x] = si;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
      red[256 + threadIdx.x] = red[256 + threadIdx.x] + red[256 + threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = red[0];
    partials[2 * blockIdx.x + 1] = red[256];
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

// Dynamic LDS beyond 64 KiB (gfx950 has 160 KiB per CU) needs the kernel's
// limit raised; done once per kernel.
template <typename K>
hipError_t allow_lds(K kernel, size_t lds) {
  if (lds <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
}

template <typename T>
hipError_t launch_tile(typename Vec2<T>::type *state, const TileArgs<T> &args, int rb,
                       uint32_t grid, hipStream_t stream) {
  using V = typename Vec2<T>::type;
  const uint32_t threads = 1u << (args.nlocal - rb);
  const size_t lds = args.nseg > 1 ? (sizeof(V) << args.nlocal) : 0;
  // The prefetching form keeps two register sets: built for 16 complex64
  // amplitudes per thread only (the 512-thread tile).
  const bool pf = args.prefetch != 0 && std::is_same<T, float>::value && rb == 4;
  switch (rb) {
    case 1: hipLaunchKernelGGL((k_tile<T, 1>), dim3(grid), dim3(threads), lds, stream, state, args); break;
    case 2: hipLaunchKernelGGL((k_tile<T, 2>), dim3(grid), dim3(threads), lds, stream, state, args); break;
    case 3: hipLaunchKernelGGL((k_tile<T, 3>), dim3(grid), dim3(threads), lds, stream, state, args); break;
    case 4:
      if constexpr (std::is_same<T, float>::value) {
        if (pf) {
          hipLaunchKernelGGL((k_tile_pf<T, 4>), dim3(grid), dim3(threads), lds, stream, state, args);
          break;
        }
      }
      hipLaunchKernelGGL((k_tile<T, 4>), dim3(grid), dim3(threads), lds, stream, state, args);
      break;
    case 5:
      if constexpr (TileCfg<T>::kRB >= 5) {
        const hipError_t e = allow_lds(k_tile<T, 5>, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_tile<T, 5>), dim3(grid), dim3(threads), lds, stream, state, args);
        break;
      }
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T, int RB>
hipError_t launch_rot_rb(typename Vec2<T>::type *state, const RotArgs<T> &args, uint32_t grid,
                         hipStream_t stream, const typename Vec2<T>::type *perm_src,
                         const typename Vec2<T>::type *ph_lut, const uint64_t *ptab) {
  using V = typename Vec2<T>::type;
  const uint32_t threads = 1u << (args.nlocal - RB);
  const size_t xlds = args.nseg > 1 ? (sizeof(V) << args.nlocal) : 0;
  if (perm_src) {
    const size_t lds = xlds + sizeof(V) * args.fr.ph.lut_len;
    const hipError_t e = allow_lds(k_rot<T, RB, true>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_rot<T, RB, true>), dim3(grid), dim3(threads), lds, stream, state,
                       args.tile_tab, args, perm_src, ph_lut, ptab);
  } else {
    const hipError_t e = allow_lds(k_rot<T, RB, false>, xlds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_rot<T, RB, false>), dim3(grid), dim3(threads), xlds, stream, state,
                       args.tile_tab, args, static_cast<const V *>(nullptr), static_cast<const V *>(nullptr),
                       static_cast<const uint64_t *>(nullptr));
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rot(typename Vec2<T>::type *state, const RotArgs<T> &args, int rb, uint32_t grid,
                      hipStream_t stream, const typename Vec2<T>::type *perm_src,
                      const typename Vec2<T>::type *ph_lut, const uint64_t *ptab) {
  switch (rb) {
    case 1: return launch_rot_rb<T, 1>(state, args, grid, stream, perm_src, ph_lut, ptab);
    case 2: return launch_rot_rb<T, 2>(state, args, grid, stream, perm_src, ph_lut, ptab);
    case 3: return launch_rot_rb<T, 3>(state, args, grid, stream, perm_src, ph_lut, ptab);
    case 4: return launch_rot_rb<T, 4>(state, args, grid, stream, perm_src, ph_lut, ptab);
    case 5:
      if constexpr (TileCfg<T>::kRB >= 5) return launch_rot_rb<T, 5>(state, args, grid, stream, perm_src, ph_lut, ptab);
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

template <typename T, int RB>
hipError_t launch_rot_relabel_rb(typename Vec2<T>::type *dst, const RotArgs<T> &args, uint32_t grid,
                                 hipStream_t stream, const typename Vec2<T>::type *src,
                                 const typename Vec2<T>::type *ph_lut, const uint64_t *ptab) {
  using V = typename Vec2<T>::type;
  const uint32_t threads = 1u << (args.nlocal - RB);
  const size_t lds = (sizeof(V) << args.nlocal) + sizeof(V) * args.fr.ph.lut_len;
  const hipError_t e = allow_lds(k_rot_relabel<T, RB>, lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_rot_relabel<T, RB>), dim3(grid), dim3(threads), lds, stream, dst, args, src, ph_lut,
                     ptab);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_rot_relabel(typename Vec2<T>::type *dst, const RotArgs<T> &args, int rb, uint32_t grid,
                              hipStream_t stream, const typename Vec2<T>::type *src,
                              const typename Vec2<T>::type *ph_lut, const uint64_t *ptab) {
  if (args.nseg < 2 || args.rl_seg + 1 >= args.nseg) return hipErrorInvalidValue;
  switch (rb) {
    case 3: return launch_rot_relabel_rb<T, 3>(dst, args, grid, stream, src, ph_lut, ptab);
    case 4: return launch_rot_relabel_rb<T, 4>(dst, args, grid, stream, src, ph_lut, ptab);
    case 5:
      if constexpr (TileCfg<T>::kRB >= 5) return launch_rot_relabel_rb<T, 5>(dst, args, grid, stream, src, ph_lut, ptab);
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
hipError_t launch_qft(typename Vec2<T>::type *state, const QftArgs<T> &args, int rb, bool inverse,
                      uint32_t grid, hipStream_t stream, const typename Vec2<T>::type *perm_src,
                      const typename Vec2<T>::type *ph_lut, const uint64_t *ptab) {
  using V = typename Vec2<T>::type;
  const uint32_t threads = 1u << (args.nlocal - rb);
  const size_t lds = args.nseg > 1 ? (sizeof(V) << args.nlocal) : 0;
  if (perm_src) {
    // the fused read exists for 4 register bits (both types' default window)
    constexpr int RBF = 4;
    if (rb != RBF) return hipErrorInvalidValue;
    const size_t flds = lds + sizeof(V) * args.fr.ph.lut_len;
    const hipError_t e = inverse ? allow_lds(k_qft<T, RBF, true, true>, flds)
                                 : allow_lds(k_qft<T, RBF, false, true>, flds);
    if (e != hipSuccess) return e;
    if (inverse)
      hipLaunchKernelGGL((k_qft<T, RBF, true, true>), dim3(grid), dim3(threads), flds, stream, state, args,
                         perm_src, ph_lut, ptab);
    else
      hipLaunchKernelGGL((k_qft<T, RBF, false, true>), dim3(grid), dim3(threads), flds, stream, state, args,
                         perm_src, ph_lut, ptab);
    return hipGetLastError();
  }
  const V *nov = nullptr;
  const uint64_t *nou = nullptr;
#define MSV_QFT_LAUNCH(RBV)                                                                          \
  {                                                                                                  \
    const hipError_t e = inverse ? allow_lds(k_qft<T, RBV, true, false>, lds)                        \
                                 : allow_lds(k_qft<T, RBV, false, false>, lds);                      \
    if (e != hipSuccess) return e;                                                                   \
  }                                                                                                  \
  if (inverse)                                                                                       \
    hipLaunchKernelGGL((k_qft<T, RBV, true, false>), dim3(grid), dim3(threads), lds, stream, state,  \
                       args, nov, nov, nou);                                                         \
  else                                                                                               \
    hipLaunchKernelGGL((k_qft<T, RBV, false, false>), dim3(grid), dim3(threads), lds, stream, state, \
                       args, nov, nov, nou);
  switch (rb) {
    case 1: MSV_QFT_LAUNCH(1) break;
    case 2: MSV_QFT_LAUNCH(2) break;
    case 3: MSV_QFT_LAUNCH(3) break;
    case 4: MSV_QFT_LAUNCH(4) break;
    case 5:
      if constexpr (TileCfg<T>::kRB >= 5) {
        MSV_QFT_LAUNCH(5)
        break;
      }
      return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
  }
#undef MSV_QFT_LAUNCH
  return hipGetLastError();
}

template <typename T>
hipError_t launch_phase(typename Vec2<T>::type *state, const typename Vec2<T>::type *lut,
                        const PhaseArgs &args, uint32_t grid, hipStream_t stream) {
  using V = typename Vec2<T>::type;
  hipLaunchKernelGGL((k_phase<T>), dim3(grid), dim3(256), sizeof(V) * args.lut_len, stream, state,
                     lut, args);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_xor_gather(const typename Vec2<T>::type *src, typename Vec2<T>::type *dst,
                             uint64_t n_amps, const uint64_t *tables, int ntab, uint64_t c,
                             uint32_t grid, hipStream_t stream) {
  PhaseArgs none;
  std::memset(&none, 0, sizeof(none));
  hipLaunchKernelGGL((k_xor_gather<T, false>), dim3(grid), dim3(256), 0, stream, src, dst, n_amps,
                     tables, ntab, c, kStreamXcdGroups, static_cast<const typename Vec2<T>::type *>(nullptr),
                     none);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_xor_gather_phase(const typename Vec2<T>::type *src, typename Vec2<T>::type *dst,
                                   uint64_t n_amps, const uint64_t *tables, int ntab, uint64_t c,
                                   const typename Vec2<T>::type *lut, const PhaseArgs &phase,
                                   uint32_t grid, hipStream_t stream) {
  using V = typename Vec2<T>::type;
  hipLaunchKernelGGL((k_xor_gather<T, true>), dim3(grid), dim3(256), sizeof(V) * phase.lut_len, stream,
                     src, dst, n_amps, tables, ntab, c, kStreamXcdGroups, lut, phase);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_init_basis(typename Vec2<T>::type *state, uint64_t n_amps, uint64_t index,
                             uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL((k_init_basis<T>), dim3(grid), dim3(256), 0, stream, state, n_amps, index);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_product(typename Vec2<T>::type *state, uint64_t n_amps, const typename Vec2<T>::type *lo,
                          const typename Vec2<T>::type *mid, const typename Vec2<T>::type *hi, uint32_t l1,
                          uint32_t l2, uint32_t grid, hipStream_t stream) {
  if (l1 > static_cast<uint32_t>(kProductGroupBits) || l2 > static_cast<uint32_t>(kProductGroupBits))
    return hipErrorInvalidValue;
  hipLaunchKernelGGL((k_product<T>), dim3(grid), dim3(256), 0, stream, state, n_amps, lo, mid, hi, l1, l2,
                     kStreamXcdGroups);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_scale(typename Vec2<T>::type *state, uint64_t n_amps, double re, double im,
                        uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL((k_scale<T>), dim3(grid), dim3(256), 0, stream, state, n_amps,
                     static_cast<T>(re), static_cast<T>(im), kStreamXcdGroups);
  return hipGetLastError();
}

template <typename Ts, typename Td>
hipError_t launch_convert(const typename Vec2<Ts>::type *src, typename Vec2<Td>::type *dst,
                          uint64_t n_amps, uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL((k_convert<Ts, Td>), dim3(grid), dim3(256), 0, stream, src, dst, n_amps);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_block_probsum(const typename Vec2<T>::type *state, uint64_t n_amps,
                                double *block_sums, uint64_t n_blocks, hipStream_t stream) {
  const uint64_t grid = (n_blocks + 3) / 4;
  hipLaunchKernelGGL((k_block_probsum<T>), dim3(static_cast<uint32_t>(grid)), dim3(256), 0, stream,
                     state, n_amps, block_sums, n_blocks);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_sample_resolve(const typename Vec2<T>::type *state, uint64_t n_amps,
                                 const uint64_t *shot_block, const double *shot_resid,
                                 uint64_t *shot_index, uint64_t shots, hipStream_t stream) {
  const uint64_t grid = (shots + 3) / 4;
  hipLaunchKernelGGL((k_sample_resolve<T>), dim3(static_cast<uint32_t>(grid)), dim3(256), 0,
                     stream, state, n_amps, shot_block, shot_resid, shot_index, shots);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_probs_full(const typename Vec2<T>::type *state, uint64_t n_amps, double *out,
                             uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL((k_probs_full<T>), dim3(grid), dim3(256), 0, stream, state, n_amps, out);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_marginal(const typename Vec2<T>::type *state, int n, const uint64_t *out_bits,
                           int nw, double *partials, uint64_t chunks, uint64_t chunk_len,
                           hipStream_t stream) {
  const uint64_t threads = (1ull << nw) * chunks;
  const uint64_t grid = (threads + 255) / 256;
  hipLaunchKernelGGL((k_marginal<T>), dim3(static_cast<uint32_t>(grid)), dim3(256), 0, stream,
                     state, n, out_bits, nw, partials, chunks, chunk_len);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_pauli(const typename Vec2<T>::type *state, uint64_t n_amps,
                        const uint64_t *xmask, const uint64_t *zmask, int nwords,
                        double *partials, uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL((k_pauli<T>), dim3(grid), dim3(256), 0, stream, state, n_amps, xmask, zmask,
                     nwords, partials);
  return hipGetLastError();
}

template <typename Ta, typename Tb>
hipError_t launch_inner(const typename Vec2<Ta>::type *a, const typename Vec2<Tb>::type *b,
                        uint64_t n_amps, double *partials, uint32_t grid, hipStream_t stream) {
  hipLaunchKernelGGL((k_inner<Ta, Tb>), dim3(grid), dim3(256), 0, stream, a, b, n_amps, partials);
  return hipGetLastError();
}

hipError_t launch_prefix_seq(double *p, uint64_t nb, hipStream_t stream) {
  hipLaunchKernelGGL(k_prefix_seq, dim3(1), dim3(64), 0, stream, p, nb);
  return hipGetLastError();
}

hipError_t launch_sample_search(const double *prefix, uint64_t nb, double total, uint64_t seed,
                                uint64_t stream_id, uint64_t shot0, uint64_t count, uint64_t *shot_block,
                                double *shot_resid, hipStream_t stream) {
  const uint64_t grid = (count + 255) / 256;
  hipLaunchKernelGGL(k_sample_search, dim3(static_cast<uint32_t>(grid)), dim3(256), 0, stream, prefix, nb,
                     total, seed, stream_id, shot0, count, shot_block, shot_resid);
  return hipGetLastError();
}

hipError_t launch_copy_probe(const float4 *src, float4 *dst, uint64_t n_vec, hipStream_t stream) {
  const uint64_t per = 512ull * kCopyVecs;
  const uint64_t grid = (n_vec + per - 1) / per;
  if (grid > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_copy_probe, dim3(static_cast<uint32_t>(grid)), dim3(512), 0, stream, src, dst, n_vec);
  return hipGetLastError();
}

#define MSV_INST_T(T)                                                                            \
  template hipError_t launch_tile<T>(Vec2<T>::type *, const TileArgs<T> &, int, uint32_t,        \
                                     hipStream_t);                                               \
  template hipError_t launch_rot<T>(Vec2<T>::type *, const RotArgs<T> &, int, uint32_t,          \
                                    hipStream_t, const Vec2<T>::type *, const Vec2<T>::type *,   \
                                    const uint64_t *);                                           \
  template hipError_t launch_rot_relabel<T>(Vec2<T>::type *, const RotArgs<T> &, int, uint32_t,  \
                                            hipStream_t, const Vec2<T>::type *,                  \
                                            const Vec2<T>::type *, const uint64_t *);            \
  template hipError_t launch_qft<T>(Vec2<T>::type *, const QftArgs<T> &, int, bool, uint32_t,    \
                                    hipStream_t, const Vec2<T>::type *, const Vec2<T>::type *,   \
                                    const uint64_t *);                                           \
  template hipError_t launch_phase<T>(Vec2<T>::type *, const Vec2<T>::type *, const PhaseArgs &, \
                                      uint32_t, hipStream_t);                                    \
  template hipError_t launch_xor_gather<T>(const Vec2<T>::type *, Vec2<T>::type *, uint64_t,     \
                                           const uint64_t *, int, uint64_t, uint32_t,            \
                                           hipStream_t);                                         \
  template hipError_t launch_xor_gather_phase<T>(const Vec2<T>::type *, Vec2<T>::type *, uint64_t, \
                                                 const uint64_t *, int, uint64_t,                \
                                                 const Vec2<T>::type *, const PhaseArgs &,       \
                                                 uint32_t, hipStream_t);                         \
  template hipError_t launch_init_basis<T>(Vec2<T>::type *, uint64_t, uint64_t, uint32_t,        \
                                           hipStream_t);                                         \
  template hipError_t launch_product<T>(Vec2<T>::type *, uint64_t, const Vec2<T>::type *,       \
                                        const Vec2<T>::type *, const Vec2<T>::type *, uint32_t,  \
                                        uint32_t, uint32_t, hipStream_t);                        \
  template hipError_t launch_scale<T>(Vec2<T>::type *, uint64_t, double, double, uint32_t,       \
                                      hipStream_t);                                              \
  template hipError_t launch_block_probsum<T>(const Vec2<T>::type *, uint64_t, double *,         \
                                              uint64_t, hipStream_t);                            \
  template hipError_t launch_sample_resolve<T>(const Vec2<T>::type *, uint64_t,                  \
                                               const uint64_t *, const double *, uint64_t *,     \
                                               uint64_t, hipStream_t);                           \
  template hipError_t launch_probs_full<T>(const Vec2<T>::type *, uint64_t, double *, uint32_t,  \
                                           hipStream_t);                                         \
  template hipError_t launch_marginal<T>(const Vec2<T>::type *, int, const uint64_t *, int,      \
                                         double *, uint64_t, uint64_t, hipStream_t);             \
  template hipError_t launch_pauli<T>(const Vec2<T>::type *, uint64_t, const uint64_t *,         \
                                      const uint64_t *, int, double *, uint32_t, hipStream_t);

MSV_INST_T(float)
MSV_INST_T(double)
#undef MSV_INST_T

template hipError_t launch_convert<float, double>(const float2 *, double2 *, uint64_t, uint32_t,
                                                  hipStream_t);
template hipError_t launch_convert<double, float>(const double2 *, float2 *, uint64_t, uint32_t,
                                                  hipStream_t);
template hipError_t launch_convert<float, float>(const float2 *, float2 *, uint64_t, uint32_t,
                                                 hipStream_t);
template hipError_t launch_convert<double, double>(const double2 *, double2 *, uint64_t, uint32_t,
                                                   hipStream_t);
template hipError_t launch_inner<float, float>(const float2 *, const float2 *, uint64_t, double *,
                                               uint32_t, hipStream_t);
template hipError_t launch_inner<float, double>(const float2 *, const double2 *, uint64_t,
                                                double *, uint32_t, hipStream_t);
template hipError_t launch_inner<double, float>(const double2 *, const float2 *, uint64_t,
                                                double *, uint32_t, hipStream_t);
template hipError_t launch_inner<double, double>(const double2 *, const double2 *, uint64_t,
                                                 double *, uint32_t, hipStream_t);

}  // namespace msv

// Batched parameter bindings of one layered rotation ansatz, with energies
// <psi_b|H|psi_b> of a diagonal observable and their gradients by the
// adjoint method (gfx950).
//
// The B bindings' complex64 states sit in one buffer, binding b at
// amplitudes [b 2^n, (b+1) 2^n): the binding number is a block of index
// bits above the circuit's n qubits that no gate touches, so one tile
// sweep over the buffer applies a layer to every binding at once. Tiles
// never straddle bindings; each tile reads its binding's gate matrices
// (built on the device from the binding's angles) with scalar loads.
//
// Backward pass (adjoint method): with phi the final state and
// lambda = H phi, gates are undone in reverse order on both; for a
// rotation R_P(theta) = exp(-i theta P / 2) the derivative of the energy is
// Im <lambda|P|phi> taken just after the gate. One fused tile sweep over
// (phi, lambda) undoes a whole layer and emits every gradient of the
// layer, as per-wave partial sums reduced afterwards in a fixed order
// (bit-reproducible).
//
// Reference: the estimator / gradient paths this replaces are
// integrations/qiskit.py:906-1090 (EstimatorV2, one circuit per binding)
// and PennyLane parameter-shift through integrations/pennylane.py:558-625.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mettleq_sv.h"
#include "msv_device.h"
#include "msv_internal.h"

namespace msv {
namespace {

#define BHIP(expr)                                                                          \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return set_error(MSV_E_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

#define BTRY(expr)                 \
  do {                             \
    int rc_ = (expr);              \
    if (rc_ != MSV_OK) return rc_; \
  } while (0)

constexpr int kBRB = 4;        // register bits per state (16 amplitudes of phi, 16 of lambda)
constexpr int kBWaves = 4;     // 256 threads
constexpr int kMaxTerms = 64;  // observable terms
constexpr int kMinBits = 10;   // index bits per binding (>= 6 thread bits + kBRB)

enum BOpKind : uint32_t {
  BOP_STAGE = 1,  // rotations of generator `gen`, at most one per register bit (see bstage)
  BOP_ROT = 2,    // rotation record of a stage (never a top-level op)
  BOP_DIAG = 4,   // diagonal on index bits a (MSB) and b of phi (and lambda)
  BOP_CZ = 5,     // product of CZ signs over a bond list (phi and lambda)
  BOP_ZDIAG = 6,  // the RZ rotations that end their qubits' rotations in a sweep (see bzdiag)
};

constexpr int kMaxCzList = 16; // non register-register CZ bonds per op

// Device-resident op record (a reverse layer of a 16-qubit ansatz has
// ~40 ops: more than the kernel-argument space).
struct BOp {
  uint32_t kind;
  int32_t rpos;    // ROT: register bit
  uint32_t gen;    // STAGE: 0 = X, 1 = Y, 2 = Z
  uint32_t gate;   // ROT: rotation number minus the sweep's first rotation
  uint32_t count;  // ROT: gradient row (the rotation number); CZ: list entries
  int32_t ra, rb;  // DIAG: register bits of index bits a / b, or -1
  uint32_t qmask;  // CZ: parity of the register-register bonds, bit r = register index r
  uint64_t amask, bmask;  // DIAG: index masks of a / b when not register bits
  float d[8];             // DIAG: d(a,b) for ab = 00, 01, 10, 11 as (re, im)
  // STAGE: list[j] = the rotation record on register bit j, or ~0; CZ: bond k = a | b << 8 | kind << 16 (kind 0: index bits a, b; kind 1:
  // register bit a, index bit b); ZDIAG: list[k] = index bit | rotation << 8 (rotation number minus the
  // sweep's first), the k < count rotations whose qubit is not a register bit
  uint32_t list[kMaxCzList];
  uint32_t zreg[4];  // ZDIAG: rotation (minus the sweep's first) on register bit j, or ~0
};

struct BArgs {
  uint64_t n_tiles;
  uint32_t n;        // index bits per binding
  uint32_t nlocal;
  uint32_t nseg;
  uint32_t nruns;
  uint32_t n_gates;  // rotations per binding
  uint32_t mat_lo;   // this sweep's rotations: [mat_lo, mat_lo + mat_cnt) of the binding's table
  uint32_t mat_cnt;
  uint32_t zop;      // the sweep's BOP_ZDIAG op, or ~0
  uint32_t run_src[kMaxRuns];
  uint32_t run_width[kMaxRuns];
  uint32_t run_dst[kMaxRuns];
  TileSeg seg[kMaxSegs];
  uint32_t op_begin[kMaxSegs];  // segment s: top-level ops [op_begin, op_end)
  uint32_t op_end[kMaxSegs];
  // pattern sweeps: CZ ops before the stages, stages, CZ ops after
  uint32_t n_pre[kMaxSegs];
  uint32_t n_stage[kMaxSegs];
  uint32_t n_post[kMaxSegs];
};

struct ObsArgs {
  uint64_t n_amps;      // all bindings
  uint32_t n;           // bits per binding
  uint32_t n_terms;
  uint32_t chunk_log;   // amplitudes per block = 2^chunk_log (within one binding)
  uint32_t pad;
  uint64_t zmask[kMaxTerms];
  double weight[kMaxTerms];
};

// ---------------------------------------------------------------------------
// device code
// ---------------------------------------------------------------------------

// Rotations act in place on the register pairs of register bit J, with
// coefficients (t, s) = (tan(phi/2), sin(phi)) from the binding's table
// (built by k_bmats), phi = theta/2 reduced to [-pi/2, pi/2]: theta/2 - phi
// is a multiple of pi, a global sign. Global signs cancel in <psi|H|psi>,
// and in Im <lambda|P|phi> of the adjoint pass because phi and lambda carry
// the same one, so neither the energies nor the gradients see them.
// Every form is three shears (Paeth) of fused multiply-adds that update
// one register from another, with no temporaries ((1 - t s) = cos phi and
// t (1 + cos phi) = s):
//   RY = [[c, -s], [s, c]]:           x0 -= t x1; x1 += s x0; x0 -= t x1;
//   RX = [[c, -i s], [-i s, c]]:      x0 += -i t x1; x1 += -i s x0; x0 += -i t x1;
//   RZ = diag(e^{-i phi}, e^{i phi}): each amplitude's (re, im) plane rotated
//                                     by -phi (x0) and +phi (x1).
// 6 FMAs per pair; the inverse negates (t, s).
template <int RB, int J>
__device__ __forceinline__ void bshear_y(float2 *a, float t, float s) {
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    float2 &x0 = a[r];
    float2 &x1 = a[r | (1 << J)];
    x0.x = fmaf(-t, x1.x, x0.x);
    x0.y = fmaf(-t, x1.y, x0.y);
    x1.x = fmaf(s, x0.x, x1.x);
    x1.y = fmaf(s, x0.y, x1.y);
    x0.x = fmaf(-t, x1.x, x0.x);
    x0.y = fmaf(-t, x1.y, x0.y);
  }
}

template <int RB, int J>
__device__ __forceinline__ void bshear_x(float2 *a, float t, float s) {
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    float2 &x0 = a[r];
    float2 &x1 = a[r | (1 << J)];
    x0.x = fmaf(t, x1.y, x0.x);
    x0.y = fmaf(-t, x1.x, x0.y);
    x1.x = fmaf(s, x0.y, x1.x);
    x1.y = fmaf(-s, x0.x, x1.y);
    x0.x = fmaf(t, x1.y, x0.x);
    x0.y = fmaf(-t, x1.x, x0.y);
  }
}

template <int RB, int J>
__device__ __forceinline__ void bshear_z(float2 *a, float t, float s) {
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    float2 &x0 = a[r];
    float2 &x1 = a[r | (1 << J)];
    x0.x = fmaf(t, x0.y, x0.x);  // x0 e^{-i phi}
    x1.x = fmaf(-t, x1.y, x1.x);  // x1 e^{i phi}
    x0.y = fmaf(-s, x0.x, x0.y);
    x1.y = fmaf(s, x1.x, x1.y);
    x0.x = fmaf(t, x0.y, x0.x);
    x1.x = fmaf(-t, x1.y, x1.x);
  }
}

template <int RB, int J, int GEN, bool DAG>
__device__ __forceinline__ void brot(float2 *a, float2 ts) {
  const float t = DAG ? -ts.x : ts.x;
  const float s = DAG ? -ts.y : ts.y;
  if (GEN == 2) bshear_z<RB, J>(a, t, s);
  else if (GEN == 1) bshear_y<RB, J>(a, t, s);
  else bshear_x<RB, J>(a, t, s);
}

// Per-thread partial of Im <lambda|P|phi> over the register pairs of bit J
// (P = X / Y / Z by GEN), as two FMA chains.
template <int RB, int J, int GEN>
__device__ __forceinline__ float bgrad(const float2 *a, const float2 *l) {
  float g0 = 0.f, g1 = 0.f;
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    const int r1 = r | (1 << J);
    const float2 a0 = a[r], a1 = a[r1], l0 = l[r], l1 = l[r1];
    if (GEN == 2) {  // Z: Im(conj(l0) a0) - Im(conj(l1) a1)
      g0 = fmaf(l0.x, a0.y, g0);
      g0 = fmaf(-l0.y, a0.x, g0);
      g1 = fmaf(-l1.x, a1.y, g1);
      g1 = fmaf(l1.y, a1.x, g1);
    } else if (GEN == 0) {  // X: Im(conj(l0) a1) + Im(conj(l1) a0)
      g0 = fmaf(l0.x, a1.y, g0);
      g0 = fmaf(-l0.y, a1.x, g0);
      g1 = fmaf(l1.x, a0.y, g1);
      g1 = fmaf(-l1.y, a0.x, g1);
    } else {  // Y: Re(conj(l1) a0) - Re(conj(l0) a1)
      g0 = fmaf(l1.x, a0.x, g0);
      g0 = fmaf(l1.y, a0.y, g0);
      g1 = fmaf(-l0.x, a1.x, g1);
      g1 = fmaf(-l0.y, a1.y, g1);
    }
  }
  return g0 + g1;
}

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}

// Sum over the 64 lanes in a fixed order, uniform result: sums within each
// 16-lane row by DPP (quad_perm [1,0,3,2], quad_perm [2,3,0,1],
// row_half_mirror, row_mirror: after each step every lane holds the sum of
// its 2, 4, 8, 16 lanes), then the four row sums in lane order.
__device__ __forceinline__ float wave_sum_dpp(float v) {
  v += dpp_f<0xB1>(v);
  v += dpp_f<0x4E>(v);
  v += dpp_f<0x141>(v);
  v += dpp_f<0x140>(v);
  const int b = __float_as_int(v);
  const float r0 = __int_as_float(__builtin_amdgcn_readlane(b, 0));
  const float r1 = __int_as_float(__builtin_amdgcn_readlane(b, 16));
  const float r2 = __int_as_float(__builtin_amdgcn_readlane(b, 32));
  const float r3 = __int_as_float(__builtin_amdgcn_readlane(b, 48));
  return (r0 + r1) + (r2 + r3);
}

template <int RB>
__device__ __forceinline__ void bdiag(float2 *a, int ra, int rb, uint64_t amask, uint64_t bmask,
                                      const float *dd, uint64_t gb) {
  const uint32_t ta = (gb & amask) ? 1u : 0u;
  const uint32_t tb = (gb & bmask) ? 1u : 0u;
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    const uint32_t ba = (ra >= 0) ? ((static_cast<uint32_t>(r) >> ra) & 1u) : ta;
    const uint32_t bb = (rb >= 0) ? ((static_cast<uint32_t>(r) >> rb) & 1u) : tb;
    const uint32_t k = 2u * ba + bb;
    float2 d;
    d.x = (k == 0) ? dd[0] : (k == 1) ? dd[2] : (k == 2) ? dd[4] : dd[6];
    d.y = (k == 0) ? dd[1] : (k == 1) ? dd[3] : (k == 2) ? dd[5] : dd[7];
    a[r] = cmul(a[r], d);
  }
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Read-only tables indexed by wave-uniform values: through the constant
// address space the loads are scalar (s_load), not per-lane vector loads.
typedef const __attribute__((address_space(4))) BOp *BOpCP;
typedef const __attribute__((address_space(4))) float *FCP;

// CZ sign products of one op on phi (and lambda).
template <int RB, bool ADJ>
__device__ __forceinline__ void bcz(BOpCP opp, float2 *a, float2 *l, uint64_t gb) {
  constexpr int R = 1 << RB;
  uint32_t c = 0, lin = 0;
  for (uint32_t k = 0; k < opp->count; ++k) {
    const uint32_t e = opp->list[k];
    const uint32_t ba = e & 255u, bb = (e >> 8) & 255u;
    const uint32_t tb = static_cast<uint32_t>(gb >> bb) & 1u;
    if (e >> 16) lin ^= tb << ba;
    else c ^= static_cast<uint32_t>(gb >> ba) & tb;
  }
  const uint32_t qm = opp->qmask;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const uint32_t par = (c ^ __popc(static_cast<uint32_t>(r) & lin) ^ (qm >> r)) & 1u;
    a[r].x = par ? -a[r].x : a[r].x;
    a[r].y = par ? -a[r].y : a[r].y;
    if (ADJ) {
      l[r].x = par ? -l[r].x : l[r].x;
      l[r].y = par ? -l[r].y : l[r].y;
    }
  }
}

// A fixed-phase op (CZ sign products or a two-qubit diagonal) on phi (and
// lambda).
template <int RB, bool ADJ>
__device__ __forceinline__ void bphase_op(BOpCP opp, float2 *a, float2 *l, uint64_t gb) {
  constexpr int R = 1 << RB;
  if (opp->kind == BOP_CZ) {
    uint32_t c = 0, lin = 0;
    for (uint32_t k = 0; k < opp->count; ++k) {
      const uint32_t e = opp->list[k];
      const uint32_t ba = e & 255u, bb = (e >> 8) & 255u;
      const uint32_t tb = static_cast<uint32_t>(gb >> bb) & 1u;
      if (e >> 16) lin ^= tb << ba;
      else c ^= static_cast<uint32_t>(gb >> ba) & tb;
    }
    const uint32_t qm = opp->qmask;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const uint32_t par = (c ^ __popc(static_cast<uint32_t>(r) & lin) ^ (qm >> r)) & 1u;
      a[r].x = par ? -a[r].x : a[r].x;
      a[r].y = par ? -a[r].y : a[r].y;
      if (ADJ) {
        l[r].x = par ? -l[r].x : l[r].x;
        l[r].y = par ? -l[r].y : l[r].y;
      }
    }
  } else {  // BOP_DIAG
    float dd[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) dd[k] = opp->d[k];
    bdiag<RB>(a, opp->ra, opp->rb, opp->amask, opp->bmask, dd, gb);
    if (ADJ) bdiag<RB>(l, opp->ra, opp->rb, opp->amask, opp->bmask, dd, gb);
  }
}

// The stage's rotation on register bit J, if it has one (list[J] = its
// record, or ~0). Forward: phi <- R phi. Adjoint: the gradient partial of
// the rotation (one wave sum, stored by lane 0 at the rotation's row),
// then phi <- R^dagger phi, lambda <- R^dagger lambda.
template <int RB, int J, int GEN, bool ADJ>
__device__ __forceinline__ void bstage_bit(BOpCP st, BOpCP cops, FCP mrow, float2 *a, float2 *l,
                                           double *__restrict__ part_t, uint64_t row_stride) {
  const uint32_t o = st->list[J];
  if (o == ~0u) return;
  const BOpCP opp = cops + o;
  const float2 ts = make_float2(mrow[2u * opp->gate], mrow[2u * opp->gate + 1u]);
  if (ADJ) {
    const float g = wave_sum_dpp(bgrad<RB, J, GEN>(a, l));
    if ((threadIdx.x & 63u) == 0u) part_t[static_cast<uint64_t>(opp->count) * row_stride] = g;
    brot<RB, J, GEN, true>(a, ts);
    brot<RB, J, GEN, true>(l, ts);
  } else {
    brot<RB, J, GEN, false>(a, ts);
  }
}

// A stage: at most one rotation per register bit, all of one generator
// (rotations on different qubits commute), as straight-line code.
template <int RB, int GEN, bool ADJ>
__device__ __forceinline__ void bstage(BOpCP st, BOpCP cops, FCP mrow, float2 *a, float2 *l,
                                       double *__restrict__ part_t, uint64_t row_stride) {
  bstage_bit<RB, 0, GEN, ADJ>(st, cops, mrow, a, l, part_t, row_stride);
  bstage_bit<RB, 1, GEN, ADJ>(st, cops, mrow, a, l, part_t, row_stride);
  bstage_bit<RB, 2, GEN, ADJ>(st, cops, mrow, a, l, part_t, row_stride);
  bstage_bit<RB, 3, GEN, ADJ>(st, cops, mrow, a, l, part_t, row_stride);
}

// The RZ rotations that end their qubits' rotations in a sweep commute with
// everything else in it (rotations on other qubits, the fixed phases), so
// they run as one diagonal D: amplitude x is multiplied by e^{-i phi_k} or
// e^{+i phi_k} for each such rotation k (its qubit's bit of x clear / set),
// with e^{i phi} = (1 - t s, s) from the rotation's (t, s). D is a thread
// factor (rotations on thread and tile bits) spread over the registers by
// doubling with the register bits' factors: 30 complex products where the
// shears of 12 rotations cost 576 FMAs.
// Forward: phi <- D phi after the last segment's ops. Adjoint, before the
// first segment's ops: the gradient partial of rotation k,
// sum over the thread's amplitudes of z_k(x) Im(conj(lambda(x)) phi(x))
// (z_k = +1 / -1 for its bit clear / set; D's other factors commute with
// Z_k, so this is the partial just after the rotation), as signed register
// sums of w = Im(conj(lambda) phi) for register bits and +-(sum of w) for
// the others, one wave sum each; then phi, lambda <- D^dagger (phi, lambda).
// a conj(b)
__device__ __forceinline__ float2 cmul_conj(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -a.x * b.y));
}

template <int RB, bool ADJ>
__device__ __forceinline__ void bzdiag(BOpCP zop, FCP mrow, float2 *a, float2 *l, uint64_t gb,
                                       double *__restrict__ part_t, uint64_t row_stride, uint32_t mat_lo) {
  static_assert(RB == 4, "the register sums and factors cover register bits 0..3");
  constexpr int R = 1 << RB;
  const uint32_t nt = zop->count;
  if (ADJ) {
    float w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) w[r] = fmaf(l[r].x, a[r].y, -l[r].y * a[r].x);
    // one level per register bit: pair sums go on, pair differences sum
    // to that bit's signed sum
    float p1[8], m1[8], p2[4], m2[4];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      p1[r] = w[2 * r] + w[2 * r + 1];
      m1[r] = w[2 * r] - w[2 * r + 1];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      p2[r] = p1[2 * r] + p1[2 * r + 1];
      m2[r] = p1[2 * r] - p1[2 * r + 1];
    }
    float sj[RB];
    sj[0] = ((m1[0] + m1[1]) + (m1[2] + m1[3])) + ((m1[4] + m1[5]) + (m1[6] + m1[7]));
    sj[1] = (m2[0] + m2[1]) + (m2[2] + m2[3]);
    sj[2] = (p2[0] - p2[1]) + (p2[2] - p2[3]);
    sj[3] = (p2[0] + p2[1]) - (p2[2] + p2[3]);
    const float wsum = (p2[0] + p2[1]) + (p2[2] + p2[3]);
#pragma unroll 1
    for (uint32_t k = 0; k < nt; ++k) {
      const uint32_t e = zop->list[k];
      const float g = wave_sum_dpp(((gb >> (e & 63u)) & 1ull) ? -wsum : wsum);
      if ((threadIdx.x & 63u) == 0u) part_t[static_cast<uint64_t>(mat_lo + (e >> 8)) * row_stride] = g;
    }
#pragma unroll
    for (int j = 0; j < RB; ++j) {
      const uint32_t o = zop->zreg[j];
      if (o == ~0u) continue;
      const float g = wave_sum_dpp(sj[j]);
      if ((threadIdx.x & 63u) == 0u) part_t[static_cast<uint64_t>(mat_lo + o) * row_stride] = g;
    }
  }
  float2 f = make_float2(1.f, 0.f);
#pragma unroll 1
  for (uint32_t k = 0; k < nt; ++k) {
    const uint32_t e = zop->list[k];
    const uint32_t o = e >> 8;
    const float t = mrow[2u * o], s = mrow[2u * o + 1u];
    f = cmul(f, make_float2(fmaf(-t, s, 1.f), ((gb >> (e & 63u)) & 1ull) ? s : -s));
  }
  // register factors e^{i phi_j}; rotation 0 of the sweep stands in for a
  // register bit without one (always a valid row), its factor replaced by 1
  float2 e[RB];
#pragma unroll
  for (int j = 0; j < RB; ++j) {
    const uint32_t o = zop->zreg[j];
    const uint32_t oo = o == ~0u ? 0u : o;
    const float t = mrow[2u * oo], sv = mrow[2u * oo + 1u];
    e[j] = make_float2(o == ~0u ? 1.f : fmaf(-t, sv, 1.f), o == ~0u ? 0.f : sv);
  }
  // D at register index r = G[r & 3] H[r >> 2]: G holds the thread factor
  // and register bits 0, 1, H register bits 2, 3 (two 4-entry tables keep
  // the registers a 16-entry one would take)
  float2 G[4], H[4];
  {
    const float2 f0 = cmul_conj(f, e[0]), f1 = cmul(f, e[0]);
    G[0] = cmul_conj(f0, e[1]);
    G[1] = cmul_conj(f1, e[1]);
    G[2] = cmul(f0, e[1]);
    G[3] = cmul(f1, e[1]);
    const float2 h0 = make_float2(e[2].x, -e[2].y);
    H[0] = cmul_conj(h0, e[3]);
    H[1] = cmul_conj(e[2], e[3]);
    H[2] = cmul(h0, e[3]);
    H[3] = cmul(e[2], e[3]);
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (ADJ) {
      a[r] = cmul_conj(cmul_conj(a[r], G[r & 3]), H[r >> 2]);
      l[r] = cmul_conj(cmul_conj(l[r], G[r & 3]), H[r >> 2]);
    } else {
      a[r] = cmul(cmul(a[r], G[r & 3]), H[r >> 2]);
    }
  }
}

// Stage-generator programs known at compile time: PAT holds the
// generators of a segment's stages in base-4 digits, first stage lowest
// (digit = generator + 1, 0 ends the program); PAT = 0 is the generic
// interpreter.
constexpr uint32_t pat_len(uint32_t p) { return p == 0 ? 0 : 1 + pat_len(p >> 2); }
constexpr int pat_gen(uint32_t p, int k) { return static_cast<int>((p >> (2 * k)) & 3u) - 1; }
constexpr uint32_t pat_of(int g0, int g1 = -1, int g2 = -1) {
  return static_cast<uint32_t>(g0 + 1) | static_cast<uint32_t>(g1 + 1) << 2 | static_cast<uint32_t>(g2 + 1) << 4;
}
constexpr int kMaxPatCz = 2;  // CZ ops before / after the stages in a pattern sweep

// One sweep over every binding's tiles. Per segment the ops
// [op_begin, op_end): fixed phases (backward: before the rotations), the
// rotation stages, fixed phases (forward: after the rotations).
// With PAT != 0 every segment holds n_pre[s] CZ ops, then no stages or the
// stages of PAT, then n_post[s] CZ ops, and the program is straight-line
// code: a loop over the ops would carry the amplitudes through a loop
// latch, where the register allocator permutes them with copies.
// The adjoint form holds phi and lambda (64 VGPRs of amplitudes): both
// states exchange through one LDS buffer in turn.
template <int RB, bool ADJ, uint32_t PAT>
__global__ void __launch_bounds__(256) k_btile(float2 *__restrict__ phi, float2 *__restrict__ lam,
                                               const BOp *__restrict__ ops,
                                               const float2 *__restrict__ mats,
                                               double *__restrict__ part, const BArgs A) {
  static_assert(RB == 4, "the rotation dispatch covers register bits 0..3");
  constexpr int R = 1 << RB;
  extern __shared__ unsigned char smem_raw[];
  float2 *lds_a = reinterpret_cast<float2 *>(smem_raw);
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const BOpCP cops = (BOpCP)(ops);
  // gradient partials: row g at part[(g * n_tiles + tile) * kBWaves + wave]
  const uint64_t row_stride = A.n_tiles * kBWaves;
  for (uint64_t tile = blockIdx.x; tile < A.n_tiles; tile += gridDim.x) {
    const uint64_t base = tile_base(A, tile);
    double *const part_t = part + tile * kBWaves + wave;
    // the binding's coefficients (tiles never straddle bindings)
    const FCP mrow = (FCP)(mats) + 2u * ((base >> A.n) * A.n_gates + A.mat_lo);
    float2 a[R], l[R];
    for (uint32_t s = 0; s < A.nseg; ++s) {
      const TileSeg &sg = A.seg[s];
      uint64_t gstr[RB];
#pragma unroll
      for (int j = 0; j < RB; ++j) gstr[j] = sg.gstr[j];
      uint32_t tl, tm;
      uint64_t gb;
      seg_thread(sg, tid, base, tl, gb, tm);
      const uint32_t reg0 = sg.reg0;
      const uint32_t tsw = lds_swz<RB>(tl);
      if (s == 0) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint64_t i = gb + reg_goff<RB>(r, gstr);
          a[r] = phi[i];
          if (ADJ) l[r] = lam[i];
        }
      } else if (!ADJ) {
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = lds_a[tsw ^ lds_swz<RB>(static_cast<uint32_t>(r) << reg0)];
      }  // ADJ: phi and lambda were read at the end of the previous segment
      if (ADJ && s == 0 && A.zop != ~0u)
        bzdiag<RB, true>(cops + A.zop, mrow, a, l, gb, part_t, row_stride, A.mat_lo);
      if constexpr (PAT == 0) {
        for (uint32_t o = A.op_begin[s]; o < A.op_end[s]; ++o) {
          const BOpCP opp = cops + o;
          if (opp->kind == BOP_STAGE) {
            const uint32_t gen = opp->gen;
            if (gen == 2u) bstage<RB, 2, ADJ>(opp, cops, mrow, a, l, part_t, row_stride);
            else if (gen == 1u) bstage<RB, 1, ADJ>(opp, cops, mrow, a, l, part_t, row_stride);
            else bstage<RB, 0, ADJ>(opp, cops, mrow, a, l, part_t, row_stride);
          } else {
            bphase_op<RB, ADJ>(opp, a, l, gb);
          }
        }
      } else {
        constexpr uint32_t L = pat_len(PAT);
        const BOpCP o0 = cops + A.op_begin[s];
        const uint32_t npre = A.n_pre[s];
        if (npre > 0u) bcz<RB, ADJ>(o0, a, l, gb);
        if (npre > 1u) bcz<RB, ADJ>(o0 + 1, a, l, gb);
        const BOpCP st = o0 + npre;
        if (A.n_stage[s] != 0u) {
          bstage<RB, pat_gen(PAT, 0), ADJ>(st, cops, mrow, a, l, part_t, row_stride);
          if constexpr (L > 1) bstage<RB, pat_gen(PAT, 1), ADJ>(st + 1, cops, mrow, a, l, part_t, row_stride);
          if constexpr (L > 2) bstage<RB, pat_gen(PAT, 2), ADJ>(st + 2, cops, mrow, a, l, part_t, row_stride);
        }
        const BOpCP post = st + A.n_stage[s];
        const uint32_t npost = A.n_post[s];
        if (npost > 0u) bcz<RB, ADJ>(post, a, l, gb);
        if (npost > 1u) bcz<RB, ADJ>(post + 1, a, l, gb);
      }
      if (!ADJ && s + 1 == A.nseg && A.zop != ~0u)
        bzdiag<RB, false>(cops + A.zop, mrow, a, l, gb, part_t, row_stride, A.mat_lo);
      if (s + 1 == A.nseg) {
        uint64_t gbs = gb;
        asm volatile("" : "+v"(gbs));
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const uint64_t i = gbs + reg_goff<RB>(r, gstr);
          phi[i] = a[r];
          if (ADJ) lam[i] = l[r];
        }
      } else {
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) lds_a[tsw ^ lds_swz<RB>(static_cast<uint32_t>(r) << reg0)] = a[r];
        __syncthreads();
        if (ADJ) {
          // phi into the next segment's registers, then lambda through the
          // same buffer
          const TileSeg &sn = A.seg[s + 1];
          uint32_t tln, tmn;
          uint64_t gbn;
          seg_thread(sn, tid, base, tln, gbn, tmn);
          const uint32_t tswn = lds_swz<RB>(tln);
#pragma unroll
          for (int r = 0; r < R; ++r) a[r] = lds_a[tswn ^ lds_swz<RB>(static_cast<uint32_t>(r) << sn.reg0)];
          __syncthreads();
#pragma unroll
          for (int r = 0; r < R; ++r) lds_a[tsw ^ lds_swz<RB>(static_cast<uint32_t>(r) << reg0)] = l[r];
          __syncthreads();
#pragma unroll
          for (int r = 0; r < R; ++r) l[r] = lds_a[tswn ^ lds_swz<RB>(static_cast<uint32_t>(r) << sn.reg0)];
        }
      }
    }
  }
}

// phi <- |0...0> in every binding
__global__ void __launch_bounds__(256) k_binit(float2 *__restrict__ phi, uint64_t n_amps, uint32_t n) {
  const uint64_t mask = (1ull << n) - 1ull;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n_amps; i += stride)
    phi[i] = make_float2((i & mask) == 0 ? 1.f : 0.f, 0.f);
}

// mats[b][g] = (tan(phi/2), sin(phi)) for rotation g of binding b, with
// phi = theta/2 reduced to [-pi/2, pi/2] and theta = params[b][pidx[g]]
// (the coefficients of brot).
__global__ void __launch_bounds__(256) k_bmats(const double *__restrict__ params, int n_params,
                                               const int32_t *__restrict__ pidx, int n_gates, int batch,
                                               float2 *__restrict__ mats) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(batch) * n_gates) return;
  const int64_t b = t / n_gates;
  const int g = static_cast<int>(t % n_gates);
  const double ph = remainder(0.5 * params[b * n_params + pidx[g]], M_PI);
  mats[t] = make_float2(static_cast<float>(tan(0.5 * ph)), static_cast<float>(sin(ph)));
}

// lambda = H phi for H = sum_t w_t Z^{zmask_t}; per-block FP64 partials of
// <phi|H|phi> (blocks lie inside one binding).
__global__ void __launch_bounds__(256) k_bobs(const float2 *__restrict__ phi, float2 *__restrict__ lam,
                                              double *__restrict__ epart, const ObsArgs A) {
  __shared__ double red[4];
  const uint64_t chunk = 1ull << A.chunk_log;
  const uint64_t start = static_cast<uint64_t>(blockIdx.x) * chunk;
  const uint64_t mask = (1ull << A.n) - 1ull;
  double e = 0.0;
  for (uint64_t i = start + threadIdx.x; i < start + chunk; i += blockDim.x) {
    const uint64_t x = i & mask;
    double h = 0.0;
    for (uint32_t t = 0; t < A.n_terms; ++t)
      h += (__popcll(x & A.zmask[t]) & 1) ? -A.weight[t] : A.weight[t];
    const float2 v = phi[i];
    lam[i] = make_float2(static_cast<float>(h * v.x), static_cast<float>(h * v.y));
    e += h * (static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y);
  }
  e = wave_sum(e);
  if ((threadIdx.x & 63u) == 0u) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) epart[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// energies[b] = sum of the binding's block partials (in order);
// grads[b][p] = sum over the gates of parameter p (CSR, in order) of the
// sum over the binding's tiles and waves (in order) of the partials.
__global__ void __launch_bounds__(256) k_breduce(const double *__restrict__ epart, int blocks_per_b,
                                                 const double *__restrict__ part, uint64_t n_tiles,
                                                 int tiles_per_b, int waves, const int32_t *__restrict__ p_off,
                                                 const int32_t *__restrict__ p_gates, int n_params,
                                                 int batch, double *__restrict__ energies,
                                                 double *__restrict__ grads) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < batch) {
    double e = 0.0;
    for (int k = 0; k < blocks_per_b; ++k) e += epart[t * blocks_per_b + k];
    energies[t] = e;
  }
  if (!grads || t >= static_cast<int64_t>(batch) * n_params) return;
  const int64_t b = t / n_params;
  const int p = static_cast<int>(t % n_params);
  double g = 0.0;
  for (int k = p_off[p]; k < p_off[p + 1]; ++k) {
    const uint64_t row = static_cast<uint64_t>(p_gates[k]) * n_tiles + static_cast<uint64_t>(b) * tiles_per_b;
    for (int i = 0; i < tiles_per_b; ++i)
      for (int w = 0; w < waves; ++w) g += part[(row + i) * kBWaves + w];
  }
  grads[t] = g;
}

}  // namespace

// ---------------------------------------------------------------------------
// host planning
// ---------------------------------------------------------------------------

namespace {

struct BGate {
  int kind;      // MSV_ROT_X/Y/Z, or 100 + MSV_PHASE_* for a fixed phase
  int q0, q1;    // qubits
  int param;     // rotations: parameter index
  double theta;  // phases: fixed angle
  int rot_no;    // rotations: gate number (matrix slot)
};

struct BSweep {
  BArgs args;
  uint32_t op_off;  // first op in the device op buffer
  bool adj;
  uint32_t pat;     // stage program of every segment (k_btile PAT), 0 = generic
  size_t lds;
};

struct BTune {
  int nlocal = 12;
  int blocks_per_cu = 4;
};

// Shortest register-window cover with coalesced first and last windows
// (same rule as the statevector tile sweeps).
std::vector<int> windows_for(uint32_t need, int nlocal, int rb) {
  const int gf = std::max(0, std::min(4, nlocal - rb));
  const int nw = nlocal - rb + 1;
  const uint32_t wmask = (1u << rb) - 1u;
  for (int m = 1; m <= kMaxSegs; ++m) {
    std::vector<int> w(m, gf);
    for (int i = 1; i + 1 < m; ++i) w[i] = 0;
    while (true) {
      uint32_t cov = 0;
      for (int x : w) cov |= wmask << x;
      if ((cov & need) == need) return w;
      int i = 0;
      for (; i < m; ++i) {
        const int lo = (i == 0 || i == m - 1) ? gf : 0;
        if (++w[i] < nw) break;
        w[i] = lo;
      }
      if (i == m) break;
    }
  }
  return {};
}

}  // namespace
}  // namespace msv

using namespace msv;

struct msv_ansatz {
  int n = 0;       // circuit qubits
  int n_eff = 0;   // index bits per binding (>= register bits)
  int device = 0;
  std::vector<BGate> gates;
  int n_rot = 0;
  int n_params = 0;
  std::vector<uint64_t> zmask;  // observable over index bits
  std::vector<double> weight;
  // device buffers for the current batch capacity
  int cap = 0;
  hipStream_t stream = nullptr;
  float2 *phi = nullptr, *lam = nullptr;
  float2 *mats = nullptr;  // [batch][rotations] (cos, sin) of half angles
  double *params_d = nullptr, *energies_d = nullptr, *grads_d = nullptr, *part = nullptr, *epart = nullptr;
  size_t part_cap = 0;
  int32_t *pidx_d = nullptr, *poff_d = nullptr, *pgates_d = nullptr;
  BOp *ops_d = nullptr;
  size_t ops_cap = 0;
  double *pin_params = nullptr, *pin_out = nullptr;  // pinned staging
  size_t pin_params_cap = 0, pin_out_cap = 0;
  bool tables_ok = false;
  // plan / graph cache per (batch, with_grad)
  struct Plan {
    BTune tune;
    std::vector<BSweep> fwd, bwd;
    std::vector<BOp> ops;
    ObsArgs obs;
    uint32_t obs_blocks = 0;
    int tiles_per_b = 0;
    int waves = 1;  // waves per tile
    uint64_t n_tiles = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    double tuned_ms = 0.0;
  };
  std::map<std::pair<int, int>, Plan> plans;
  std::mutex mu;
};

namespace {

int bit_of(const msv_ansatz *a, int q) { return a->n_eff - 1 - q; }

void free_buffers(msv_ansatz *a) {
  for (void *p : {static_cast<void *>(a->phi), static_cast<void *>(a->lam), static_cast<void *>(a->mats),
                  static_cast<void *>(a->params_d), static_cast<void *>(a->energies_d),
                  static_cast<void *>(a->grads_d), static_cast<void *>(a->part), static_cast<void *>(a->epart)})
    if (p) (void)hipFree(p);
  a->phi = a->lam = nullptr;
  a->mats = nullptr;
  a->params_d = a->energies_d = a->grads_d = a->part = a->epart = nullptr;
  a->part_cap = 0;
  a->cap = 0;
}

void drop_plans(msv_ansatz *a) {
  for (auto &kv : a->plans) {
    if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
  }
  a->plans.clear();
}

// Sweeps of one block (rotations on some qubits, then fixed phases) at a
// tile of `nl` local bits. Forward: ROT ops in time order per qubit, the
// phases at the end of the last sweep. Backward: the phases undone first,
// then per qubit in reverse time order GRAD + ROT_ADJ.
int plan_block(const msv_ansatz *a, const std::vector<int> &rots, const std::vector<int> &phases,
               int nl, bool backward, int batch, msv_ansatz::Plan *P, std::vector<BSweep> *out) {
  const int n = a->n_eff;
  const int rb = kBRB;
  const int NL = std::min(nl, n);
  // index-bit partition: low sweep [0, L), high chunks of <= NL - 4 bits;
  // every sweep's tile has NL local bits (equal tile counts)
  std::vector<std::vector<int>> sweeps;  // gate bits per sweep
  {
    int L = NL;
    int H = n - NL;
    if (H > 0 && H < rb) {
      L = n - rb;
      H = rb;
    }
    std::vector<int> low;
    for (int b = 0; b < L; ++b) low.push_back(b);
    sweeps.push_back(low);
    if (H > 0) {
      const int G = std::max(rb, NL - 4);
      const int k = (H + G - 1) / G;
      int start = L;
      for (int c = 0; c < k; ++c) {
        const int sz = H / k + (c < H % k ? 1 : 0);
        std::vector<int> g;
        for (int b = start; b < start + sz; ++b) g.push_back(b);
        sweeps.push_back(g);
        start += sz;
      }
    }
  }
  // which rotations go to which sweep (by their qubit's bit)
  std::vector<std::vector<int>> sweep_rots(sweeps.size());
  for (int gi : rots) {
    const int bit = bit_of(a, a->gates[gi].q0);
    for (size_t s = 0; s < sweeps.size(); ++s)
      if (std::find(sweeps[s].begin(), sweeps[s].end(), bit) != sweeps[s].end()) sweep_rots[s].push_back(gi);
  }
  int last_used = -1;
  for (size_t s = 0; s < sweeps.size(); ++s)
    if (!sweep_rots[s].empty()) last_used = static_cast<int>(s);
  // an RZ rotation that ends its qubit's rotations in the block commutes
  // with the rest of its sweep: it joins the sweep's diagonal (bzdiag)
  std::vector<char> in_zdiag(a->gates.size(), 0);
  {
    std::vector<int> last(static_cast<size_t>(a->n), -1);
    for (int gi : rots) last[static_cast<size_t>(a->gates[gi].q0)] = gi;
    for (int gi : last)
      if (gi >= 0 && a->gates[gi].kind == MSV_ROT_Z) in_zdiag[static_cast<size_t>(gi)] = 1;
  }
  const int phase_sweep = backward ? -1 : last_used;  // forward: phases after the last rotations
  bool phases_done = phases.empty();
  const uint64_t total_bits = static_cast<uint64_t>(n);
  int nb = 0;
  while ((1 << nb) < batch) ++nb;
  for (size_t s = 0; s < sweeps.size(); ++s) {
    const bool put_phases = !phases_done && (backward ? true : (static_cast<int>(s) == phase_sweep ||
                                                                 last_used < 0));
    if (sweep_rots[s].empty() && !put_phases) continue;
    // local bits: low sweep = its gate bits; high sweeps = columns + gates
    std::vector<int> local;
    if (s == 0) {
      for (int b = 0; b < NL; ++b) local.push_back(b);
    } else {
      const int C = std::min(NL - static_cast<int>(sweeps[s].size()), sweeps[s].front());
      for (int b = 0; b < C; ++b) local.push_back(b);
      for (int b : sweeps[s]) local.push_back(b);
    }
    const int nlocal = static_cast<int>(local.size());
    if (nlocal < rb || nlocal - rb > kTileThreadBits)
      return set_error(MSV_E_INVALID, "internal: batched tile shape");
    int loc_of[64];
    for (int i = 0; i < 64; ++i) loc_of[i] = -1;
    for (int k = 0; k < nlocal; ++k) loc_of[local[k]] = k;
    uint32_t need = 0;
    for (int gi : sweep_rots[s])
      if (!in_zdiag[static_cast<size_t>(gi)]) need |= 1u << loc_of[bit_of(a, a->gates[gi].q0)];
    std::vector<int> wins = need ? windows_for(need, nlocal, rb) : std::vector<int>{nlocal - rb};
    if (wins.empty()) return set_error(MSV_E_INVALID, "internal: no register window cover");
    BSweep sw;
    std::memset(&sw.args, 0, sizeof(sw.args));
    BArgs &A = sw.args;
    A.n = static_cast<uint32_t>(n);
    A.nlocal = static_cast<uint32_t>(nlocal);
    A.nseg = static_cast<uint32_t>(wins.size());
    A.n_gates = static_cast<uint32_t>(a->n_rot);
    // runs of non-local bits over n + nb bits (binding bits on top)
    const int all = n + nb;
    uint64_t lmask = 0;
    for (int b : local) lmask |= 1ull << b;
    int src = 0, nr = 0;
    for (int b = 0; b < all;) {
      if (b < n && (lmask >> b & 1ull)) {
        ++b;
        continue;
      }
      int e = b;
      while (e < all && !(e < n && (lmask >> e & 1ull))) ++e;
      if (nr >= kMaxRuns) return set_error(MSV_E_INVALID, "internal: too many bit runs");
      A.run_src[nr] = static_cast<uint32_t>(src);
      A.run_width[nr] = static_cast<uint32_t>(e - b);
      A.run_dst[nr] = static_cast<uint32_t>(b);
      src += e - b;
      ++nr;
      b = e;
    }
    A.nruns = static_cast<uint32_t>(nr);
    A.n_tiles = static_cast<uint64_t>(batch) << (n - nlocal);
    (void)total_bits;
    // this sweep's rotations are a contiguous gate range (a block's
    // rotations are consecutive gates)
    int mat_lo = 0, mat_hi = -1;
    for (int gi : sweep_rots[s]) {
      const int r = a->gates[gi].rot_no;
      if (mat_hi < 0 || r < mat_lo) mat_lo = r;
      mat_hi = std::max(mat_hi, r);
    }
    const int mat_cnt = mat_hi < 0 ? 0 : mat_hi - mat_lo + 1;
    A.mat_lo = static_cast<uint32_t>(mat_lo);
    A.mat_cnt = static_cast<uint32_t>(mat_cnt);
    bool all_cz = true;
    for (int gi : phases) {
      const BGate &g = a->gates[gi];
      all_cz = all_cz && g.kind - 100 == MSV_PHASE_CPHASE && std::abs(std::fabs(g.theta) - M_PI) < 1e-15;
    }
    std::vector<bool> placed(sweep_rots[s].size(), false);
    for (size_t i = 0; i < sweep_rots[s].size(); ++i) placed[i] = in_zdiag[static_cast<size_t>(sweep_rots[s][i])] != 0;
    BOp zd;
    std::memset(&zd, 0, sizeof(zd));
    zd.kind = BOP_ZDIAG;
    for (int j = 0; j < 4; ++j) zd.zreg[j] = ~0u;
    bool has_zd = false;
    uint32_t pat = ~0u;  // the sweep's stage program (~0: none seen yet, 0: generic)
    for (size_t w = 0; w < wins.size(); ++w) {
      TileSeg &sg = A.seg[w];
      const int reg0 = wins[w];
      sg.reg0 = static_cast<uint32_t>(reg0);
      for (int j = 0; j < rb; ++j) sg.gstr[j] = 1ull << local[reg0 + j];
      sg.tmask = (1u << (nlocal - rb)) - 1u;
      for (int k = 0, t = 0; k < nlocal; ++k) {
        if (k >= reg0 && k < reg0 + rb) continue;
        sg.tl_bit[t] = static_cast<uint32_t>(k);
        sg.gi_bit[t] = static_cast<uint32_t>(local[k]);
        ++t;
      }
      auto rpos = [&](int bit) -> int {
        const int k = loc_of[bit];
        return (k >= 0 && k >= reg0 && k < reg0 + rb) ? k - reg0 : -1;
      };
      auto emit_phases = [&]() {
        if (all_cz) {
          // signs: register-register bonds as a parity table over r, the
          // others as (register bit or index bit, index bit) pairs
          BOp op;
          std::memset(&op, 0, sizeof(op));
          op.kind = BOP_CZ;
          for (int gi : phases) {
            const BGate &g = a->gates[gi];
            const int ba = bit_of(a, g.q0), bb = bit_of(a, g.q1);
            const int ja = rpos(ba), jb = rpos(bb);
            uint32_t e;
            if (ja >= 0 && jb >= 0) {
              for (int r = 0; r < (1 << rb); ++r)
                if (((r >> ja) & 1) && ((r >> jb) & 1)) op.qmask ^= 1u << r;
              continue;
            } else if (ja >= 0) {
              e = static_cast<uint32_t>(ja) | static_cast<uint32_t>(bb) << 8 | 1u << 16;
            } else if (jb >= 0) {
              e = static_cast<uint32_t>(jb) | static_cast<uint32_t>(ba) << 8 | 1u << 16;
            } else {
              e = static_cast<uint32_t>(ba) | static_cast<uint32_t>(bb) << 8;
            }
            if (op.count == kMaxCzList) {
              P->ops.push_back(op);
              std::memset(&op, 0, sizeof(op));
              op.kind = BOP_CZ;
            }
            op.list[op.count++] = e;
          }
          P->ops.push_back(op);
          return;
        }
        for (int gi : phases) {
          const BGate &g = a->gates[gi];
          BOp op;
          std::memset(&op, 0, sizeof(op));
          op.kind = BOP_DIAG;
          const int ba = bit_of(a, g.q0), bb = bit_of(a, g.q1);
          op.ra = rpos(ba);
          op.rb = rpos(bb);
          op.amask = op.ra >= 0 ? 0ull : (1ull << ba);
          op.bmask = op.rb >= 0 ? 0ull : (1ull << bb);
          std::complex<double> d[4] = {1.0, 1.0, 1.0, 1.0};
          if (g.kind - 100 == MSV_PHASE_CPHASE) {
            d[3] = std::abs(std::fabs(g.theta) - M_PI) < 1e-15 ? std::complex<double>(-1.0, 0.0)
                                                             : std::polar(1.0, g.theta);
          } else {  // ZZ: exp(-i theta Z Z)
            d[0] = d[3] = std::polar(1.0, -g.theta);
            d[1] = d[2] = std::polar(1.0, g.theta);
          }
          for (int k = 0; k < 4; ++k) {
            const std::complex<double> v = backward ? std::conj(d[k]) : d[k];
            op.d[2 * k] = static_cast<float>(v.real());
            op.d[2 * k + 1] = static_cast<float>(v.imag());
          }
          P->ops.push_back(op);
        }
      };
      // the diagonal in the layout of the segment that runs it: the last
      // (forward) or the first (adjoint)
      if (w == (backward ? 0 : wins.size() - 1)) {
        for (int gi : sweep_rots[s]) {
          if (!in_zdiag[static_cast<size_t>(gi)]) continue;
          const BGate &g = a->gates[gi];
          const int bit = bit_of(a, g.q0);
          const int j = rpos(bit);
          const uint32_t o = static_cast<uint32_t>(g.rot_no - mat_lo);
          if (o >= (1u << 24)) return set_error(MSV_E_INVALID, "internal: diagonal rotation range");
          if (j >= 0) {
            if (zd.zreg[j] != ~0u) return set_error(MSV_E_INVALID, "internal: two diagonal rotations on one bit");
            zd.zreg[j] = o;
          } else {
            if (zd.count == kMaxCzList) return set_error(MSV_E_INVALID, "internal: diagonal list");
            zd.list[zd.count++] = static_cast<uint32_t>(bit) | o << 8;
          }
          has_zd = true;
        }
      }
      // rotations whose bit is a register bit of this window, not yet placed
      std::vector<int> here;
      for (size_t i = 0; i < sweep_rots[s].size(); ++i) {
        if (placed[i]) continue;
        if (rpos(bit_of(a, a->gates[sweep_rots[s][i]].q0)) >= 0) {
          placed[i] = true;
          here.push_back(sweep_rots[s][i]);
        }
      }
      if (backward) std::reverse(here.begin(), here.end());
      // per register bit, in application order
      std::vector<std::vector<int>> seq(rb);
      for (int gi : here) seq[rpos(bit_of(a, a->gates[gi].q0))].push_back(gi);
      // stages: repeatedly take, on every bit whose next rotation has the
      // generator most common among the bits' next rotations, that
      // rotation; the rotation records go first, the segment's top-level
      // ops after them
      std::vector<BOp> stages;
      std::vector<size_t> head(rb, 0);
      while (true) {
        int cnt[3] = {0, 0, 0};
        for (int j = 0; j < rb; ++j)
          if (head[j] < seq[j].size()) ++cnt[a->gates[seq[j][head[j]]].kind];
        if (cnt[0] + cnt[1] + cnt[2] == 0) break;
        const int gen = cnt[1] >= cnt[0] && cnt[1] >= cnt[2] ? 1 : (cnt[2] >= cnt[0] ? 2 : 0);
        BOp st;
        std::memset(&st, 0, sizeof(st));
        st.kind = BOP_STAGE;
        st.gen = static_cast<uint32_t>(gen);
        for (int j = 0; j < rb; ++j) {
          st.list[j] = ~0u;
          if (head[j] >= seq[j].size() || a->gates[seq[j][head[j]]].kind != gen) continue;
          const BGate &g = a->gates[seq[j][head[j]++]];
          BOp op;
          std::memset(&op, 0, sizeof(op));
          op.kind = BOP_ROT;
          op.rpos = j;
          op.gen = static_cast<uint32_t>(gen);
          op.gate = static_cast<uint32_t>(g.rot_no - mat_lo);
          op.count = static_cast<uint32_t>(g.rot_no);
          st.list[j] = static_cast<uint32_t>(P->ops.size());
          P->ops.push_back(op);
        }
        stages.push_back(st);
      }
      A.op_begin[w] = static_cast<uint32_t>(P->ops.size());
      if (w == 0 && backward && put_phases) emit_phases();
      const size_t pre_end = P->ops.size();
      for (const BOp &st : stages) P->ops.push_back(st);
      const size_t st_end = P->ops.size();
      if (w + 1 == wins.size() && !backward && put_phases) emit_phases();
      A.op_end[w] = static_cast<uint32_t>(P->ops.size());
      // pattern form: CZ ops around one stage-generator program
      A.n_pre[w] = static_cast<uint32_t>(pre_end - A.op_begin[w]);
      A.n_stage[w] = static_cast<uint32_t>(st_end - pre_end);
      A.n_post[w] = static_cast<uint32_t>(P->ops.size() - st_end);
      uint32_t prog = 0;
      for (size_t k = 0; k < stages.size() && k < 3; ++k) prog |= (stages[k].gen + 1u) << (2 * k);
      bool cz_only = true;
      for (size_t k = A.op_begin[w]; k < P->ops.size(); ++k)
        if (P->ops[k].kind != BOP_STAGE && P->ops[k].kind != BOP_CZ) cz_only = false;
      if (stages.size() > 3 || !cz_only || A.n_pre[w] > kMaxPatCz || A.n_post[w] > kMaxPatCz ||
          (!stages.empty() && pat != ~0u && pat != 0u && pat != prog))
        pat = 0;
      else if (!stages.empty() && pat == ~0u)
        pat = prog;
    }
    // segments without stages fit any program; a sweep without rotations
    // runs the generic form
    sw.pat = pat == ~0u ? 0u : pat;
    A.zop = ~0u;
    if (has_zd) {
      A.zop = static_cast<uint32_t>(P->ops.size());
      P->ops.push_back(zd);
    }
    for (bool p : placed)
      if (!p) return set_error(MSV_E_INVALID, "internal: rotation left unscheduled");
    if (put_phases) phases_done = true;
    sw.adj = backward;
    sw.lds = A.nseg > 1 ? (sizeof(float2) << nlocal) : 0;
    out->push_back(sw);
  }
  return MSV_OK;
}

int build_plan(msv_ansatz *a, int batch, const BTune &tune, bool with_grad, msv_ansatz::Plan *P) {
  P->tune = tune;
  P->fwd.clear();
  P->bwd.clear();
  P->ops.clear();
  // blocks: a run of rotations followed by a run of phases
  std::vector<std::pair<std::vector<int>, std::vector<int>>> blocks;
  for (size_t i = 0; i < a->gates.size(); ++i) {
    const bool rot = a->gates[i].kind < 100;
    if (blocks.empty() || (rot && !blocks.back().second.empty())) blocks.push_back({});
    (rot ? blocks.back().first : blocks.back().second).push_back(static_cast<int>(i));
  }
  for (auto &bl : blocks) BTRY(plan_block(a, bl.first, bl.second, tune.nlocal, false, batch, P, &P->fwd));
  if (with_grad)
    for (auto it = blocks.rbegin(); it != blocks.rend(); ++it)
      BTRY(plan_block(a, it->first, it->second, tune.nlocal, true, batch, P, &P->bwd));
  // every gradient sweep must use the same tile count (partial rows)
  const int n = a->n_eff;
  const int nl = std::min(tune.nlocal, n);
  P->n_tiles = static_cast<uint64_t>(batch) << (n - nl);
  P->tiles_per_b = 1 << (n - nl);
  P->waves = 1 << (nl - kBRB - 6);
  for (const BSweep &sw : P->bwd)
    if (sw.args.n_tiles != P->n_tiles) return set_error(MSV_E_INVALID, "internal: uneven gradient tiles");
  // observable pass
  std::memset(&P->obs, 0, sizeof(P->obs));
  P->obs.n = static_cast<uint32_t>(n);
  P->obs.n_amps = static_cast<uint64_t>(batch) << n;
  P->obs.n_terms = static_cast<uint32_t>(a->zmask.size());
  P->obs.chunk_log = static_cast<uint32_t>(std::min(n, 12));
  for (size_t t = 0; t < a->zmask.size(); ++t) {
    P->obs.zmask[t] = a->zmask[t];
    P->obs.weight[t] = a->weight[t];
  }
  P->obs_blocks = static_cast<uint32_t>(P->obs.n_amps >> P->obs.chunk_log);
  return MSV_OK;
}

// Device buffers for `batch` bindings and plan P. Cached graphs hold raw
// buffer addresses, so any reallocation first drops every cached plan.
int ensure_buffers(msv_ansatz *a, int batch, const msv_ansatz::Plan &P) {
  const size_t need_part = static_cast<size_t>(std::max(1, a->n_rot)) * P.n_tiles * kBWaves;
  const size_t need_ops = std::max<size_t>(P.ops.size(), 256);
  const size_t pin_p = static_cast<size_t>(batch) * std::max(1, a->n_params);
  const size_t pin_o = static_cast<size_t>(batch) * (1 + std::max(1, a->n_params));
  const bool grow = batch > a->cap || need_part > a->part_cap || P.ops.size() > a->ops_cap ||
                    pin_p > a->pin_params_cap || pin_o > a->pin_out_cap;
  if (grow) {
    BHIP(hipStreamSynchronize(a->stream));
    drop_plans(a);
  }
  if (batch > a->cap) {
    free_buffers(a);
    const size_t amps = static_cast<size_t>(batch) << a->n_eff;
    BHIP(hipMalloc(&a->phi, amps * sizeof(float2)));
    BHIP(hipMalloc(&a->lam, amps * sizeof(float2)));
    BHIP(hipMalloc(&a->mats, static_cast<size_t>(batch) * std::max(1, a->n_rot) * sizeof(float2)));
    BHIP(hipMalloc(&a->params_d, static_cast<size_t>(batch) * std::max(1, a->n_params) * sizeof(double)));
    BHIP(hipMalloc(&a->energies_d, static_cast<size_t>(batch) * sizeof(double)));
    BHIP(hipMalloc(&a->grads_d, static_cast<size_t>(batch) * std::max(1, a->n_params) * sizeof(double)));
    BHIP(hipMalloc(&a->epart, (amps >> std::min(a->n_eff, 12)) * sizeof(double)));
    a->cap = batch;
  }
  if (need_part > a->part_cap) {
    if (a->part) (void)hipFree(a->part);
    a->part = nullptr;
    a->part_cap = 0;
    BHIP(hipMalloc(&a->part, need_part * sizeof(double)));
    a->part_cap = need_part;
  }
  if (P.ops.size() > a->ops_cap) {
    if (a->ops_d) (void)hipFree(a->ops_d);
    a->ops_d = nullptr;
    a->ops_cap = 0;
    BHIP(hipMalloc(&a->ops_d, need_ops * sizeof(BOp)));
    a->ops_cap = need_ops;
  }
  if (pin_p > a->pin_params_cap) {
    // keep the staged parameters across the reallocation
    double *old = a->pin_params;
    const size_t old_n = a->pin_params_cap;
    a->pin_params = nullptr;
    a->pin_params_cap = 0;
    BHIP(hipHostMalloc(reinterpret_cast<void **>(&a->pin_params), pin_p * sizeof(double), hipHostMallocDefault));
    a->pin_params_cap = pin_p;
    if (old) {
      std::memcpy(a->pin_params, old, old_n * sizeof(double));
      (void)hipHostFree(old);
    }
  }
  if (pin_o > a->pin_out_cap) {
    if (a->pin_out) (void)hipHostFree(a->pin_out);
    a->pin_out = nullptr;
    a->pin_out_cap = 0;
    BHIP(hipHostMalloc(reinterpret_cast<void **>(&a->pin_out), pin_o * sizeof(double), hipHostMallocDefault));
    a->pin_out_cap = pin_o;
  }
  if (!a->tables_ok) {
    std::vector<int32_t> pidx, poff(a->n_params + 1, 0), pg;
    for (const BGate &g : a->gates)
      if (g.kind < 100) pidx.push_back(g.param);
    for (int p = 0; p < a->n_params; ++p) {
      poff[p] = static_cast<int32_t>(pg.size());
      for (const BGate &g : a->gates)
        if (g.kind < 100 && g.param == p) pg.push_back(g.rot_no);
    }
    poff[a->n_params] = static_cast<int32_t>(pg.size());
    for (int32_t **pp : {&a->pidx_d, &a->poff_d, &a->pgates_d}) {
      if (*pp) (void)hipFree(*pp);
      *pp = nullptr;
    }
    BHIP(hipMalloc(&a->pidx_d, std::max<size_t>(1, pidx.size()) * sizeof(int32_t)));
    BHIP(hipMalloc(&a->poff_d, poff.size() * sizeof(int32_t)));
    BHIP(hipMalloc(&a->pgates_d, std::max<size_t>(1, pg.size()) * sizeof(int32_t)));
    if (!pidx.empty())
      BHIP(hipMemcpy(a->pidx_d, pidx.data(), pidx.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    BHIP(hipMemcpy(a->poff_d, poff.data(), poff.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!pg.empty())
      BHIP(hipMemcpy(a->pgates_d, pg.data(), pg.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    a->tables_ok = true;
  }
  return MSV_OK;
}

// Stage programs with a compiled kernel (single rotation layers, Euler
// pairs and ZYZ, and their reverses for the adjoint sweeps).
constexpr uint32_t kPats[] = {pat_of(0), pat_of(1), pat_of(2), pat_of(1, 2), pat_of(2, 1),
                              pat_of(0, 2), pat_of(2, 0), pat_of(2, 1, 2)};

template <bool ADJ, uint32_t PAT>
void launch_k_btile(msv_ansatz *a, const BSweep &sw, uint32_t grid, uint32_t threads) {
  hipLaunchKernelGGL((k_btile<kBRB, ADJ, PAT>), dim3(grid), dim3(threads), sw.lds, a->stream, a->phi, a->lam,
                     a->ops_d, a->mats, a->part, sw.args);
}

template <bool ADJ, size_t I = 0>
void launch_pat(msv_ansatz *a, const BSweep &sw, uint32_t grid, uint32_t threads) {
  if constexpr (I == sizeof(kPats) / sizeof(kPats[0])) {
    launch_k_btile<ADJ, 0>(a, sw, grid, threads);
  } else {
    if (sw.pat == kPats[I]) launch_k_btile<ADJ, kPats[I]>(a, sw, grid, threads);
    else launch_pat<ADJ, I + 1>(a, sw, grid, threads);
  }
}

bool pat_compiled(uint32_t pat) {
  for (uint32_t p : kPats)
    if (p == pat) return true;
  return false;
}

int launch_sweep(msv_ansatz *a, const msv_ansatz::Plan &P, const BSweep &sw, int cu) {
  const uint64_t cap = static_cast<uint64_t>(cu) * P.tune.blocks_per_cu;
  const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(sw.args.n_tiles, cap)));
  const uint32_t threads = 1u << (sw.args.nlocal - kBRB);
  if (sw.adj) launch_pat<true>(a, sw, grid, threads);
  else launch_pat<false>(a, sw, grid, threads);
  BHIP(hipGetLastError());
  return MSV_OK;
}

// Enqueues one evaluation on a->stream: params (pinned) -> device, matrices,
// |0>, forward sweeps, observable, backward sweeps, reductions, results ->
// pinned. Everything is a kernel or a memcpy node: capturable as a graph.
int enqueue_eval(msv_ansatz *a, int batch, const msv_ansatz::Plan &P, bool with_grad) {
  hipStream_t st = a->stream;
  const int cu = device_cu_count(a->device);
  const size_t np = static_cast<size_t>(batch) * std::max(1, a->n_params);
  BHIP(hipMemcpyAsync(a->params_d, a->pin_params, np * sizeof(double), hipMemcpyHostToDevice, st));
  if (a->n_rot > 0) {
    const int64_t work = static_cast<int64_t>(batch) * a->n_rot;
    hipLaunchKernelGGL(k_bmats, dim3(static_cast<uint32_t>((work + 255) / 256)), dim3(256), 0, st,
                       a->params_d, a->n_params, a->pidx_d, a->n_rot, batch, a->mats);
    BHIP(hipGetLastError());
  }
  const uint64_t amps = static_cast<uint64_t>(batch) << a->n_eff;
  hipLaunchKernelGGL(k_binit, dim3(static_cast<uint32_t>(std::min<uint64_t>((amps + 255) / 256, cu * 32ull))),
                     dim3(256), 0, st, a->phi, amps, static_cast<uint32_t>(a->n_eff));
  BHIP(hipGetLastError());
  auto launch = [&](const BSweep &sw) { return launch_sweep(a, P, sw, cu); };
  for (const BSweep &sw : P.fwd) BTRY(launch(sw));
  hipLaunchKernelGGL(k_bobs, dim3(P.obs_blocks), dim3(256), 0, st, a->phi, a->lam, a->epart, P.obs);
  BHIP(hipGetLastError());
  if (with_grad)
    for (const BSweep &sw : P.bwd) BTRY(launch(sw));
  const int64_t rwork = std::max<int64_t>(batch, with_grad ? static_cast<int64_t>(batch) * a->n_params : 0);
  hipLaunchKernelGGL(k_breduce, dim3(static_cast<uint32_t>((rwork + 255) / 256)), dim3(256), 0, st, a->epart,
                     static_cast<int>(1u << (a->n_eff - std::min(a->n_eff, 12))), a->part, P.n_tiles,
                     P.tiles_per_b, P.waves, a->poff_d, a->pgates_d, a->n_params, batch, a->energies_d,
                     with_grad ? a->grads_d : nullptr);
  BHIP(hipGetLastError());
  BHIP(hipMemcpyAsync(a->pin_out, a->energies_d, batch * sizeof(double), hipMemcpyDeviceToHost, st));
  if (with_grad)
    BHIP(hipMemcpyAsync(a->pin_out + batch, a->grads_d, np * sizeof(double), hipMemcpyDeviceToHost, st));
  return MSV_OK;
}

int upload_ops(msv_ansatz *a, const msv_ansatz::Plan &P) {
  if (!P.ops.empty())
    BHIP(hipMemcpy(a->ops_d, P.ops.data(), P.ops.size() * sizeof(BOp), hipMemcpyHostToDevice));
  return MSV_OK;
}

double time_eval(msv_ansatz *a, int batch, const msv_ansatz::Plan &P, bool with_grad, int reps) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1e30;
  double best = 1e30;
  for (int r = 0; r < reps; ++r) {
    (void)hipEventRecord(e0, a->stream);
    if (enqueue_eval(a, batch, P, with_grad) != MSV_OK) {
      best = 1e30;
      break;
    }
    (void)hipEventRecord(e1, a->stream);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    best = std::min(best, static_cast<double>(ms));
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

// Autotuner: tile size x persistent blocks per CU, timed on the real batch
// with the caller's parameters; the winner's sequence is captured into a
// HIP graph and replayed by later calls of the same batch size.
int get_plan(msv_ansatz *a, int batch, bool with_grad, msv_ansatz::Plan **out) {
  const auto key = std::make_pair(batch, with_grad ? 1 : 0);
  auto it = a->plans.find(key);
  if (it != a->plans.end()) {
    *out = &it->second;
    return MSV_OK;
  }
  std::vector<BTune> cands;
  for (int nl : {12, 11, 10})
    for (int bpc : {2, 4, 8})
      if (nl <= a->n_eff && nl >= kBRB) cands.push_back({nl, bpc});
  if (cands.empty()) cands.push_back({a->n_eff, 4});
  msv_ansatz::Plan best;
  double best_ms = 1e30;
  // METTLEQ_SV_BATCH_TUNE=0 skips the search (first candidate);
  // METTLEQ_SV_BATCH_PLAN="<tile bits>,<blocks per CU>" fixes the plan.
  const char *env = std::getenv("METTLEQ_SV_BATCH_TUNE");
  bool tune = !(env && env[0] == '0') && cands.size() > 1;
  if (const char *fixed = std::getenv("METTLEQ_SV_BATCH_PLAN")) {
    int nl = 0, bpc = 0;
    if (std::sscanf(fixed, "%d,%d", &nl, &bpc) == 2 && nl >= kMinBits && nl <= 12 && bpc >= 1 && bpc <= 64) {
      cands.assign(1, BTune{std::min(nl, a->n_eff), bpc});
      tune = false;
    }
  }
  for (const BTune &t : tune ? cands : std::vector<BTune>{cands.front()}) {
    msv_ansatz::Plan P;
    BTRY(build_plan(a, batch, t, with_grad, &P));
    BTRY(ensure_buffers(a, batch, P));
    BTRY(upload_ops(a, P));
    const double ms = tune ? time_eval(a, batch, P, with_grad, 2) : 0.0;
    if (ms < best_ms || !tune) {
      best_ms = ms;
      best = P;
    }
  }
  best.tuned_ms = best_ms;
  BTRY(ensure_buffers(a, batch, best));  // may drop other cached plans, never this one
  BTRY(upload_ops(a, best));
  BHIP(hipStreamSynchronize(a->stream));
  // capture the winner
  BHIP(hipStreamBeginCapture(a->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_eval(a, batch, best, with_grad);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(a->stream, &g);
  if (rc != MSV_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ec != hipSuccess) return set_error(MSV_E_HIP, std::string("graph capture failed: ") + hipGetErrorString(ec));
  hipGraphExec_t ex = nullptr;
  const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (ei != hipSuccess) {
    (void)hipGraphDestroy(g);
    return set_error(MSV_E_HIP, std::string("graph instantiate failed: ") + hipGetErrorString(ei));
  }
  best.graph = g;
  best.exec = ex;
  // ops of different plans share the device op buffer: only one plan's ops
  // are resident, so a cached graph is re-bound by re-uploading its ops.
  auto res = a->plans.emplace(key, best);
  *out = &res.first->second;
  return MSV_OK;
}

}  // namespace

extern "C" {

int msv_ansatz_create(int n_qubits, int device, msv_ansatz **out) {
  if (!out) return set_error(MSV_E_INVALID, "null output");
  *out = nullptr;
  if (n_qubits < 1 || n_qubits > 30) return set_error(MSV_E_INVALID, "n_qubits must be in [1, 30]");
  int count = 0;
  BHIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return set_error(MSV_E_INVALID, "device out of range");
  msv_ansatz *a = new msv_ansatz();
  a->n = n_qubits;
  // at least 10 index bits per binding: every tile then has 64 x 2^k
  // threads (whole waves for the wave-level gradient sums); padding qubits
  // stay |0> and are never touched
  a->n_eff = std::max(n_qubits, kMinBits);
  a->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&a->stream) != hipSuccess) {
    delete a;
    return set_error(MSV_E_HIP, "cannot create the ansatz stream");
  }
  *out = a;
  return MSV_OK;
}

int msv_ansatz_destroy(msv_ansatz *a) {
  if (!a) return MSV_OK;
  (void)hipSetDevice(a->device);
  (void)hipStreamSynchronize(a->stream);
  drop_plans(a);
  free_buffers(a);
  for (int32_t *p : {a->pidx_d, a->poff_d, a->pgates_d})
    if (p) (void)hipFree(p);
  if (a->ops_d) (void)hipFree(a->ops_d);
  if (a->pin_params) (void)hipHostFree(a->pin_params);
  if (a->pin_out) (void)hipHostFree(a->pin_out);
  (void)hipStreamDestroy(a->stream);
  delete a;
  return MSV_OK;
}

int msv_ansatz_rotation(msv_ansatz *a, int kind, int qubit, int param_index) {
  if (!a) return set_error(MSV_E_INVALID, "null ansatz");
  if (kind < MSV_ROT_X || kind > MSV_ROT_Z) return set_error(MSV_E_INVALID, "rotation kind must be MSV_ROT_X/Y/Z");
  if (qubit < 0 || qubit >= a->n) return set_error(MSV_E_INVALID, "rotation qubit out of range");
  if (param_index < 0) return set_error(MSV_E_INVALID, "negative parameter index");
  std::lock_guard<std::mutex> lk(a->mu);
  BGate g{kind, qubit, -1, param_index, 0.0, a->n_rot++};
  a->gates.push_back(g);
  a->n_params = std::max(a->n_params, param_index + 1);
  drop_plans(a);
  a->tables_ok = false;
  return MSV_OK;
}

int msv_ansatz_phase(msv_ansatz *a, int kind, int q0, int q1, double theta) {
  if (!a) return set_error(MSV_E_INVALID, "null ansatz");
  if (kind != MSV_PHASE_ZZ && kind != MSV_PHASE_CPHASE)
    return set_error(MSV_E_INVALID, "phase kind must be MSV_PHASE_ZZ or MSV_PHASE_CPHASE");
  if (q0 < 0 || q0 >= a->n || q1 < 0 || q1 >= a->n || q0 == q1)
    return set_error(MSV_E_INVALID, "phase qubits must be two distinct valid qubits");
  if (!std::isfinite(theta)) return set_error(MSV_E_INVALID, "phase angle must be finite");
  std::lock_guard<std::mutex> lk(a->mu);
  a->gates.push_back(BGate{100 + kind, q0, q1, -1, theta, -1});
  drop_plans(a);
  return MSV_OK;
}

int msv_ansatz_observable(msv_ansatz *a, const uint64_t *zmasks, const double *weights, int n_terms) {
  if (!a) return set_error(MSV_E_INVALID, "null ansatz");
  if (n_terms < 1 || n_terms > kMaxTerms || !zmasks || !weights)
    return set_error(MSV_E_INVALID, "observable needs 1..64 terms");
  std::lock_guard<std::mutex> lk(a->mu);
  a->zmask.clear();
  a->weight.clear();
  for (int t = 0; t < n_terms; ++t) {
    if (a->n < 64 && (zmasks[t] >> a->n) != 0)
      return set_error(MSV_E_INVALID, "observable mask names a qubit outside the register");
    if (!std::isfinite(weights[t])) return set_error(MSV_E_INVALID, "observable weight must be finite");
    uint64_t m = 0;
    for (int q = 0; q < a->n; ++q)
      if ((zmasks[t] >> q) & 1ull) m |= 1ull << bit_of(a, q);
    a->zmask.push_back(m);
    a->weight.push_back(weights[t]);
  }
  drop_plans(a);
  return MSV_OK;
}

int msv_ansatz_evaluate(msv_ansatz *a, const double *params, int batch, int n_params, double *energies,
                        double *grads) {
  if (!a) return set_error(MSV_E_INVALID, "null ansatz");
  if (batch < 1) return set_error(MSV_E_INVALID, "batch must be >= 1");
  if (n_params != a->n_params)
    return set_error(MSV_E_INVALID, "n_params does not match the ansatz (" + std::to_string(a->n_params) + ")");
  if ((a->n_params > 0 && !params) || !energies) return set_error(MSV_E_INVALID, "null buffer");
  if (a->zmask.empty()) return set_error(MSV_E_INVALID, "observable not set");
  const double bytes = 2.0 * static_cast<double>(batch) * std::ldexp(8.0, a->n_eff);
  if (bytes > 200e9) return set_error(MSV_E_REFUSED, "batch states exceed the device budget");
  std::lock_guard<std::mutex> lk(a->mu);
  BHIP(hipSetDevice(a->device));
  const bool with_grad = grads != nullptr;
  const size_t np = static_cast<size_t>(batch) * a->n_params;
  for (size_t i = 0; i < np; ++i)
    if (!std::isfinite(params[i])) return set_error(MSV_E_INVALID, "parameters must be finite");
  // the autotuner times real evaluations: stage the caller's parameters first
  if (np > a->pin_params_cap) {
    BHIP(hipStreamSynchronize(a->stream));
    drop_plans(a);
    if (a->pin_params) (void)hipHostFree(a->pin_params);
    a->pin_params = nullptr;
    a->pin_params_cap = 0;
    BHIP(hipHostMalloc(reinterpret_cast<void **>(&a->pin_params), np * sizeof(double), hipHostMallocDefault));
    a->pin_params_cap = np;
  }
  if (np > 0) std::memcpy(a->pin_params, params, np * sizeof(double));
  msv_ansatz::Plan *P = nullptr;
  BTRY(get_plan(a, batch, with_grad, &P));
  BTRY(upload_ops(a, *P));
  BHIP(hipGraphLaunch(P->exec, a->stream));
  BHIP(hipStreamSynchronize(a->stream));
  std::memcpy(energies, a->pin_out, batch * sizeof(double));
  if (with_grad && a->n_params > 0)
    std::memcpy(grads, a->pin_out + batch, static_cast<size_t>(batch) * a->n_params * sizeof(double));
  return MSV_OK;
}

int msv_ansatz_info(msv_ansatz *a, int batch, int with_grad, int64_t *out, int n_out) {
  if (!a || !out || n_out < 1) return set_error(MSV_E_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(a->mu);
  auto it = a->plans.find(std::make_pair(batch, with_grad ? 1 : 0));
  if (it == a->plans.end()) return set_error(MSV_E_INVALID, "no plan for this batch size yet");
  const msv_ansatz::Plan &P = it->second;
  int64_t straight = 0;
  for (const auto *v : {&P.fwd, &P.bwd})
    for (const BSweep &sw : *v) straight += pat_compiled(sw.pat) ? 1 : 0;
  const int64_t vals[7] = {P.tune.nlocal, P.tune.blocks_per_cu, static_cast<int64_t>(P.fwd.size()),
                           static_cast<int64_t>(P.bwd.size()), static_cast<int64_t>(P.ops.size()),
                           static_cast<int64_t>(P.tuned_ms * 1000.0), straight};
  for (int i = 0; i < n_out && i < 7; ++i) out[i] = vals[i];
  return MSV_OK;
}

int msv_ansatz_time_sweeps(msv_ansatz *a, int batch, int with_grad, int reps, double *out_ms) {
  if (!a || !out_ms || reps < 1) return set_error(MSV_E_INVALID, "null argument or reps < 1");
  std::lock_guard<std::mutex> lk(a->mu);
  auto it = a->plans.find(std::make_pair(batch, with_grad ? 1 : 0));
  if (it == a->plans.end()) return set_error(MSV_E_INVALID, "no plan for this batch size yet");
  const msv_ansatz::Plan &P = it->second;
  BHIP(hipSetDevice(a->device));
  BTRY(upload_ops(a, P));
  const int cu = device_cu_count(a->device);
  hipEvent_t ev[3];
  for (hipEvent_t &e : ev) BHIP(hipEventCreate(&e));
  double tot[2] = {0.0, 0.0};
  int rc = MSV_OK;
  for (int r = 0; r < reps && rc == MSV_OK; ++r) {
    (void)hipEventRecord(ev[0], a->stream);
    for (const BSweep &sw : P.fwd)
      if (rc == MSV_OK) rc = launch_sweep(a, P, sw, cu);
    (void)hipEventRecord(ev[1], a->stream);
    for (const BSweep &sw : P.bwd)
      if (rc == MSV_OK) rc = launch_sweep(a, P, sw, cu);
    (void)hipEventRecord(ev[2], a->stream);
    if (hipEventSynchronize(ev[2]) != hipSuccess && rc == MSV_OK) rc = set_error(MSV_E_HIP, "sweep timing failed");
    float f = 0.f, b = 0.f;
    (void)hipEventElapsedTime(&f, ev[0], ev[1]);
    (void)hipEventElapsedTime(&b, ev[1], ev[2]);
    tot[0] += f;
    tot[1] += b;
  }
  for (hipEvent_t &e : ev) (void)hipEventDestroy(e);
  if (rc != MSV_OK) return rc;
  out_ms[0] = tot[0] / reps;
  out_ms[1] = tot[1] / reps;
  return MSV_OK;
}

}  // extern "C"

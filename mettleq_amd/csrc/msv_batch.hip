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
  BOP_ROT = 1,      // phi <- R phi (R = the binding's rotation `gate`)
  BOP_ROT_ADJ = 2,  // phi <- R^dagger phi, lambda <- R^dagger lambda
  BOP_GRAD = 3,     // acc[slot] += per-thread Im <lambda|P|phi>, P on register bit rpos
  BOP_DIAG = 4,     // diagonal on index bits a (MSB) and b of phi (and lambda)
  BOP_CZ = 5,       // product of CZ signs over a bond list (phi and lambda)
  BOP_FLUSH = 6,    // wave sums of acc[0..count) -> partials of gates[0..count)
};

constexpr int kMaxAcc = 8;     // gradient accumulators per flush
constexpr int kMaxCzList = 16; // non register-register CZ bonds per op
constexpr int kMaxSweepGates = 256;

// Device-resident op record (a reverse layer of a 16-qubit ansatz has
// ~60 ops: more than the kernel-argument space).
struct BOp {
  uint32_t kind;
  int32_t rpos;    // ROT/GRAD: register bit
  uint32_t gen;    // ROT/GRAD: 0 = X, 1 = Y, 2 = Z
  uint32_t gate;   // ROT/GRAD: gate number minus the sweep's first gate
  uint32_t count;  // GRAD: accumulator slot; FLUSH: slots; CZ: list entries
  int32_t ra, rb;  // DIAG: register bits of index bits a / b, or -1
  uint32_t qmask;  // CZ: parity of the register-register bonds, bit r = register index r
  uint64_t amask, bmask;  // DIAG: index masks of a / b when not register bits
  float d[8];             // DIAG: d(a,b) for ab = 00, 01, 10, 11 as (re, im)
  // FLUSH: gate of slot k; CZ: bond k = a | b << 8 | kind << 16 (kind 0:
  // index bits a, b; kind 1: register bit a, index bit b)
  uint32_t list[kMaxCzList];
};

struct BArgs {
  uint64_t n_tiles;
  uint32_t n;        // index bits per binding
  uint32_t nlocal;
  uint32_t nseg;
  uint32_t nruns;
  uint32_t n_gates;  // rotations per binding
  uint32_t mat_lo;   // this sweep's rotations: [mat_lo, mat_lo + mat_cnt)
  uint32_t mat_cnt;
  uint32_t pad;
  uint32_t run_src[kMaxRuns];
  uint32_t run_width[kMaxRuns];
  uint32_t run_dst[kMaxRuns];
  TileSeg seg[kMaxSegs];
  uint32_t op_begin[kMaxSegs];
  uint32_t op_end[kMaxSegs];
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

// Rotation exp(-i theta P / 2) with (c, s) = (cos theta/2, sin theta/2) on
// register bit J, P = X/Y/Z by GEN; DAG applies the inverse (s -> -s).
// 8 FMAs per pair; the generator is a template argument so that the pair
// loop has no branches.
template <int RB, int J, bool DAG, int GEN>
__device__ __forceinline__ void brot(float2 *a, float c, float s) {
  if (DAG) s = -s;
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    const int r1 = r | (1 << J);
    const float2 v0 = a[r], v1 = a[r1];
    if (GEN == 0) {  // RX = [[c, -i s], [-i s, c]]
      a[r].x = c * v0.x + s * v1.y;
      a[r].y = c * v0.y - s * v1.x;
      a[r1].x = c * v1.x + s * v0.y;
      a[r1].y = c * v1.y - s * v0.x;
    } else if (GEN == 1) {  // RY = [[c, -s], [s, c]]
      a[r].x = c * v0.x - s * v1.x;
      a[r].y = c * v0.y - s * v1.y;
      a[r1].x = s * v0.x + c * v1.x;
      a[r1].y = s * v0.y + c * v1.y;
    } else {  // RZ = diag(c - i s, c + i s)
      a[r].x = c * v0.x + s * v0.y;
      a[r].y = c * v0.y - s * v0.x;
      a[r1].x = c * v1.x - s * v1.y;
      a[r1].y = c * v1.y + s * v1.x;
    }
  }
}

template <int RB, bool DAG, int GEN>
__device__ __forceinline__ void brot_j(float2 *a, float c, float s, int j) {
  switch (j) {
    case 0: brot<RB, 0, DAG, GEN>(a, c, s); break;
    case 1: brot<RB, 1, DAG, GEN>(a, c, s); break;
    case 2: brot<RB, 2, DAG, GEN>(a, c, s); break;
    case 3: brot<RB, 3, DAG, GEN>(a, c, s); break;
    default: break;
  }
}

template <int RB, bool DAG>
__device__ __forceinline__ void brot_dispatch(float2 *a, uint32_t gen, float c, float s, int j) {
  if (gen == 0u) brot_j<RB, DAG, 0>(a, c, s, j);
  else if (gen == 1u) brot_j<RB, DAG, 1>(a, c, s, j);
  else brot_j<RB, DAG, 2>(a, c, s, j);
}

// Per-thread partial of Im <lambda|P|phi> over the register pairs of bit J.
template <int RB, int J, int GEN>
__device__ __forceinline__ float bgrad(const float2 *a, const float2 *l) {
  float g = 0.f;
#pragma unroll
  for (int r = 0; r < (1 << RB); ++r) {
    if (r & (1 << J)) continue;
    const int r1 = r | (1 << J);
    const float2 a0 = a[r], a1 = a[r1], l0 = l[r], l1 = l[r1];
    if (GEN == 2) {  // Z: Im(conj(l0) a0) - Im(conj(l1) a1)
      g += (l0.x * a0.y - l0.y * a0.x) - (l1.x * a1.y - l1.y * a1.x);
    } else if (GEN == 0) {  // X: Im(conj(l0) a1) + Im(conj(l1) a0)
      g += (l0.x * a1.y - l0.y * a1.x) + (l1.x * a0.y - l1.y * a0.x);
    } else {  // Y: Re(conj(l1) a0) - Re(conj(l0) a1)
      g += (l1.x * a0.x + l1.y * a0.y) - (l0.x * a1.x + l0.y * a1.y);
    }
  }
  return g;
}

template <int RB, int GEN>
__device__ __forceinline__ float bgrad_j(const float2 *a, const float2 *l, int j) {
  switch (j) {
    case 0: return bgrad<RB, 0, GEN>(a, l);
    case 1: return bgrad<RB, 1, GEN>(a, l);
    case 2: return bgrad<RB, 2, GEN>(a, l);
    case 3: return bgrad<RB, 3, GEN>(a, l);
    default: return 0.f;
  }
}

template <int RB>
__device__ __forceinline__ float bgrad_dispatch(const float2 *a, const float2 *l, uint32_t gen, int j) {
  if (gen == 0u) return bgrad_j<RB, 0>(a, l, j);
  if (gen == 1u) return bgrad_j<RB, 1>(a, l, j);
  return bgrad_j<RB, 2>(a, l, j);
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

__device__ __forceinline__ float wave_sum_f(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Read-only tables indexed by wave-uniform values: through the constant
// address space the loads are scalar (s_load), not per-lane vector loads.
typedef const __attribute__((address_space(4))) BOp *BOpCP;

// The adjoint form holds phi and lambda (64 VGPRs of amplitudes, 160 in
// all): both states exchange through one LDS buffer in turn, so its
// registers, not the LDS, set the blocks per CU (three; holding it to 128
// VGPRs for four spills and was 2.6x slower).
template <int RB, bool ADJ>
__global__ void __launch_bounds__(256) k_btile(float2 *__restrict__ phi, float2 *__restrict__ lam,
                                               const BOp *__restrict__ ops,
                                               const float2 *__restrict__ mats,
                                               double *__restrict__ part, const BArgs A) {
  constexpr int R = 1 << RB;
  extern __shared__ unsigned char smem_raw[];
  float2 *lds_a = reinterpret_cast<float2 *>(smem_raw);
  // the tile's rotations (cos, sin) staged once per tile
  float2 *lds_m = lds_a + (A.nseg > 1 ? (1u << A.nlocal) : 0u);
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid >> 6;
  const BOpCP cops = (BOpCP)(ops);
  for (uint64_t tile = blockIdx.x; tile < A.n_tiles; tile += gridDim.x) {
    const uint64_t base = tile_base(A, tile);
    {
      const float2 *src = mats + (base >> A.n) * A.n_gates + A.mat_lo;
      __syncthreads();  // previous tile's readers are done
      for (uint32_t i = tid; i < A.mat_cnt; i += blockDim.x) lds_m[i] = src[i];
      __syncthreads();
    }
    float2 a[R], l[R];
    float acc[kMaxAcc];
#pragma unroll
    for (int k = 0; k < kMaxAcc; ++k) acc[k] = 0.f;
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
      for (uint32_t o = A.op_begin[s]; o < A.op_end[s]; ++o) {
        const BOpCP opp = cops + o;
        const uint32_t kind = opp->kind;
        if (kind == BOP_ROT) {
          const float2 cs = lds_m[opp->gate];
          brot_dispatch<RB, false>(a, opp->gen, cs.x, cs.y, opp->rpos);
        } else if (kind == BOP_ROT_ADJ) {
          const float2 cs = lds_m[opp->gate];
          brot_dispatch<RB, true>(a, opp->gen, cs.x, cs.y, opp->rpos);
          if (ADJ) brot_dispatch<RB, true>(l, opp->gen, cs.x, cs.y, opp->rpos);
        } else if (kind == BOP_GRAD) {
          if (ADJ) {
            const float g = bgrad_dispatch<RB>(a, l, opp->gen, opp->rpos);
            switch (opp->count) {
              case 0: acc[0] += g; break;
              case 1: acc[1] += g; break;
              case 2: acc[2] += g; break;
              case 3: acc[3] += g; break;
              case 4: acc[4] += g; break;
              case 5: acc[5] += g; break;
              case 6: acc[6] += g; break;
              default: acc[7] += g; break;
            }
          }
        } else if (kind == BOP_FLUSH) {
          if (ADJ) {
            const uint32_t cnt = opp->count;
            float w[kMaxAcc];
#pragma unroll
            for (int k = 0; k < kMaxAcc; ++k) w[k] = (k < static_cast<int>(cnt)) ? wave_sum_f(acc[k]) : 0.f;
            if ((tid & 63u) == 0u) {
#pragma unroll
              for (int k = 0; k < kMaxAcc; ++k)
                if (k < static_cast<int>(cnt))
                  part[(static_cast<uint64_t>(opp->list[k]) * A.n_tiles + tile) * kBWaves + wave] = w[k];
            }
#pragma unroll
            for (int k = 0; k < kMaxAcc; ++k) acc[k] = 0.f;
          }
        } else if (kind == BOP_CZ) {
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
        } else if (kind == BOP_DIAG) {
          float dd[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) dd[k] = opp->d[k];
          bdiag<RB>(a, opp->ra, opp->rb, opp->amask, opp->bmask, dd, gb);
          if (ADJ) bdiag<RB>(l, opp->ra, opp->rb, opp->amask, opp->bmask, dd, gb);
        }
      }
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

// mats[b][g] = (cos, sin) of half the binding's angle params[b][pidx[g]]
__global__ void __launch_bounds__(256) k_bmats(const double *__restrict__ params, int n_params,
                                               const int32_t *__restrict__ pidx, int n_gates,
                                               int batch, float2 *__restrict__ mats) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= static_cast<int64_t>(batch) * n_gates) return;
  const int64_t b = t / n_gates;
  const int g = static_cast<int>(t % n_gates);
  double s, c;
  sincos(0.5 * params[b * n_params + pidx[g]], &s, &c);
  mats[t] = make_float2(static_cast<float>(c), static_cast<float>(s));
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
    for (int gi : sweep_rots[s]) need |= 1u << loc_of[bit_of(a, a->gates[gi].q0)];
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
    // rotations are consecutive gates); staged per tile in LDS
    int mat_lo = 0, mat_hi = -1;
    for (int gi : sweep_rots[s]) {
      const int r = a->gates[gi].rot_no;
      if (mat_hi < 0 || r < mat_lo) mat_lo = r;
      mat_hi = std::max(mat_hi, r);
    }
    const int mat_cnt = mat_hi < 0 ? 0 : mat_hi - mat_lo + 1;
    if (mat_cnt > kMaxSweepGates)
      return set_error(MSV_E_UNSUPPORTED, "more than 256 rotations in one layer of the ansatz");
    A.mat_lo = static_cast<uint32_t>(mat_lo);
    A.mat_cnt = static_cast<uint32_t>(mat_cnt);
    bool all_cz = true;
    for (int gi : phases) {
      const BGate &g = a->gates[gi];
      all_cz = all_cz && g.kind - 100 == MSV_PHASE_CPHASE && std::abs(std::fabs(g.theta) - M_PI) < 1e-15;
    }
    std::vector<bool> placed(sweep_rots[s].size(), false);
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
      A.op_begin[w] = static_cast<uint32_t>(P->ops.size());
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
      if (w == 0 && backward && put_phases) emit_phases();
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
      BOp flush;
      std::memset(&flush, 0, sizeof(flush));
      flush.kind = BOP_FLUSH;
      for (int gi : here) {
        const BGate &g = a->gates[gi];
        BOp op;
        std::memset(&op, 0, sizeof(op));
        op.rpos = rpos(bit_of(a, g.q0));
        op.gen = static_cast<uint32_t>(g.kind);  // MSV_ROT_X/Y/Z = 0/1/2
        op.gate = static_cast<uint32_t>(g.rot_no - mat_lo);
        if (backward) {
          op.kind = BOP_GRAD;
          op.count = flush.count;
          P->ops.push_back(op);
          flush.list[flush.count++] = static_cast<uint32_t>(g.rot_no);
          op.kind = BOP_ROT_ADJ;
          op.count = 0;
          P->ops.push_back(op);
          if (flush.count == kMaxAcc) {
            P->ops.push_back(flush);
            flush.count = 0;
          }
        } else {
          op.kind = BOP_ROT;
          P->ops.push_back(op);
        }
      }
      if (flush.count > 0) P->ops.push_back(flush);
      if (w + 1 == wins.size() && !backward && put_phases) emit_phases();
      A.op_end[w] = static_cast<uint32_t>(P->ops.size());
    }
    for (bool p : placed)
      if (!p) return set_error(MSV_E_INVALID, "internal: rotation left unscheduled");
    if (put_phases) phases_done = true;
    sw.adj = backward;
    sw.lds = (A.nseg > 1 ? (sizeof(float2) << nlocal) : 0) +
             sizeof(float2) * static_cast<size_t>(mat_cnt);
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
  auto launch = [&](const BSweep &sw) -> int {
    const uint64_t cap = static_cast<uint64_t>(cu) * P.tune.blocks_per_cu;
    const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(sw.args.n_tiles, cap)));
    const uint32_t threads = 1u << (sw.args.nlocal - kBRB);
    const BOp *ops = a->ops_d + sw.op_off;
    (void)ops;
    if (sw.adj)
      hipLaunchKernelGGL((k_btile<kBRB, true>), dim3(grid), dim3(threads), sw.lds, st, a->phi, a->lam,
                         a->ops_d, a->mats, a->part, sw.args);
    else
      hipLaunchKernelGGL((k_btile<kBRB, false>), dim3(grid), dim3(threads), sw.lds, st, a->phi, a->lam,
                         a->ops_d, a->mats, a->part, sw.args);
    BHIP(hipGetLastError());
    return MSV_OK;
  };
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
  const int64_t vals[6] = {P.tune.nlocal, P.tune.blocks_per_cu, static_cast<int64_t>(P.fwd.size()),
                           static_cast<int64_t>(P.bwd.size()), static_cast<int64_t>(P.ops.size()),
                           static_cast<int64_t>(P.tuned_ms * 1000.0)};
  for (int i = 0; i < n_out && i < 6; ++i) out[i] = vals[i];
  return MSV_OK;
}

}  // extern "C"

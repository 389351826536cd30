// Internal types shared by the gfx950 kernels (msv_kernels.hip) and the
// host runtime (msv_runtime.hip). Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace msv {

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

// Register bits per thread in a tile sweep: 32 complex64 or 16 complex128
// amplitudes live in registers (64 VGPRs either way).
template <typename T> struct TileCfg;
template <> struct TileCfg<float> { static constexpr int kRB = 5; static constexpr int kMaxLocal = 13; };
template <> struct TileCfg<double> { static constexpr int kRB = 4; static constexpr int kMaxLocal = 12; };

constexpr int kMaxTileOps = 24;
constexpr int kMaxSegs = 6;
constexpr int kMaxPhaseWins = 32;
constexpr int kPhaseLutMax = 4096;   // entries of all windows together
constexpr int kSampleBlock = 4096;   // amplitudes per sampling block
constexpr int kSampleLanes = 64;     // lanes per block (one wave64)
constexpr int kMaxWords = 16;        // Pauli words per reduction launch

enum TileOpKind : uint8_t {
  OP_U2 = 1,     // 2x2 on a register bit, optional controls
  OP_QFT_F = 2,  // forward QFT stage: butterfly then ladder phase on the 1 branch
  OP_QFT_I = 3,  // inverse QFT stage: conjugate ladder phase then butterfly
  OP_DIAG = 4,   // diagonal over 1 or 2 index bits (anywhere in the index)
};

// One operation of a tile program. Matrix entries are stored in the
// amplitude precision; the host rounds its double-precision values once.
template <typename T>
struct TileOp {
  uint8_t kind;
  int8_t rpos;      // target register bit (U2, QFT)
  int8_t ra;        // DIAG: register bit of index bit a, or -1
  int8_t rb;        // DIAG: register bit of index bit b, or -1
  uint8_t jl;       // QFT: ladder register bits directly below rpos
  uint8_t nbits;    // DIAG: 1 or 2
  uint8_t has_ctrl; // U2: controls present
  uint8_t pbit;     // QFT: index bit p of the stage
  uint8_t lb;       // QFT: lowest index bit of the ladder
  uint8_t pad[3];
  uint32_t cmask_r; // U2: control bits that are register bits (register positions)
  uint64_t cmask_t; // U2: control bits that are not register bits (index mask)
  uint64_t amask;   // DIAG: index mask of bit a (when ra < 0)
  uint64_t bmask;   // DIAG: index mask of bit b (when rb < 0)
  T m[8];
};

struct TileSeg {
  uint8_t reg[5];   // local bit held by register bit j
  uint8_t op_begin; // ops [op_begin, op_end) run in this segment
  uint8_t op_end;
  uint8_t pad;
};

template <typename T>
struct TileArgs {
  uint64_t n_tiles;        // tiles = 2^(n - nlocal)
  uint64_t nonlocal_mask;  // index bits enumerated by the tile number
  uint32_t nlocal;         // local bits per tile
  uint32_t nseg;
  uint8_t lpos[16];        // index bit of local bit k (increasing)
  TileSeg seg[kMaxSegs];
  T scale;                 // applied at the final store (QFT 1/sqrt2 factors)
  uint32_t pad;
  T w32[64];               // exp(2*pi*i*m/32), m = 0..31, (re, im)
  TileOp<T> ops[kMaxTileOps];
};

struct PhaseWin {
  uint8_t s1, w1, s2, w2;  // lut index = ((i>>s1)&(2^w1-1)) | (((i>>s2)&(2^w2-1))<<w1)
  uint32_t off;            // offset of this window's table
};

struct PhaseArgs {
  uint64_t n_amps;
  uint32_t nwin;
  uint32_t nlow;           // windows [0, nlow) read index bits 0..1
  uint32_t lut_len;
  uint32_t pad;
  PhaseWin win[kMaxPhaseWins];
};

// Kernel launchers (msv_kernels.hip). All enqueue on `stream`.
template <typename T>
hipError_t launch_tile(typename Vec2<T>::type *state, const TileArgs<T> &args, int rb,
                       uint32_t grid, hipStream_t stream);
template <typename T>
hipError_t launch_phase(typename Vec2<T>::type *state, const typename Vec2<T>::type *lut,
                        const PhaseArgs &args, uint32_t grid, hipStream_t stream);
template <typename T>
hipError_t launch_xor_gather(const typename Vec2<T>::type *src, typename Vec2<T>::type *dst,
                             uint64_t n_amps, const uint64_t *tables, int ntab, uint64_t c,
                             uint32_t grid, hipStream_t stream);
template <typename T>
hipError_t launch_init_basis(typename Vec2<T>::type *state, uint64_t n_amps, uint64_t index,
                             uint32_t grid, hipStream_t stream);
template <typename T>
hipError_t launch_scale(typename Vec2<T>::type *state, uint64_t n_amps, double re, double im,
                        uint32_t grid, hipStream_t stream);
template <typename Ts, typename Td>
hipError_t launch_convert(const typename Vec2<Ts>::type *src, typename Vec2<Td>::type *dst,
                          uint64_t n_amps, uint32_t grid, hipStream_t stream);
template <typename T>
hipError_t launch_block_probsum(const typename Vec2<T>::type *state, uint64_t n_amps,
                                double *block_sums, uint64_t n_blocks, hipStream_t stream);
template <typename T>
hipError_t launch_sample_resolve(const typename Vec2<T>::type *state, uint64_t n_amps,
                                 const uint64_t *shot_block, const double *shot_resid,
                                 uint64_t *shot_index, uint64_t shots, hipStream_t stream);
template <typename T>
hipError_t launch_probs_full(const typename Vec2<T>::type *state, uint64_t n_amps, double *out,
                             uint32_t grid, hipStream_t stream);
template <typename T>
hipError_t launch_marginal(const typename Vec2<T>::type *state, int n, const uint64_t *out_bits,
                           int nw, double *partials, uint64_t chunks, uint64_t chunk_len,
                           hipStream_t stream);
template <typename T>
hipError_t launch_pauli(const typename Vec2<T>::type *state, uint64_t n_amps,
                        const uint64_t *xmask, const uint64_t *zmask, int nwords,
                        double *partials, uint32_t grid, hipStream_t stream);
template <typename Ta, typename Tb>
hipError_t launch_inner(const typename Vec2<Ta>::type *a, const typename Vec2<Tb>::type *b,
                        uint64_t n_amps, double *partials, uint32_t grid, hipStream_t stream);
hipError_t launch_copy_probe(const float4 *src, float4 *dst, uint64_t n_vec, uint32_t grid,
                             hipStream_t stream);

constexpr uint32_t kReduceThreads = 256;

}  // namespace msv

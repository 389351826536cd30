// Internal types shared by the gfx950 kernels (msv_kernels.hip) and the
// host runtime (msv_runtime.hip). Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace msv {

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

// Register bits per thread in a tile sweep: 32 complex64 or 16 complex128
// amplitudes live in registers (64 VGPRs either way).
template <typename T> struct TileCfg;
// kMaxOps bounds a tile program so that TileArgs stays under 4 KiB of
// kernel arguments.
// kTabVals sizes the per-launch table pool: the QFT twiddles (32 complex).
template <> struct TileCfg<float> {
  static constexpr int kRB = 5;
  static constexpr int kMaxLocal = 13;
  static constexpr int kMaxOps = 24;
  static constexpr int kTabVals = 64;
};
template <> struct TileCfg<double> {
  static constexpr int kRB = 4;
  static constexpr int kMaxLocal = 12;
  static constexpr int kMaxOps = 16;
  static constexpr int kTabVals = 64;
};
constexpr int kDtabVals = 84;  // one layer diagonal of k_rot: 32 register + 9 thread-bit complex (+ pad)

constexpr int kMaxSegs = 6;
constexpr int kMaxPhaseWins = 32;
constexpr int kPhaseLutMax = 4096;   // entries of all windows together
constexpr int kSampleBlock = 4096;   // amplitudes per sampling block
constexpr int kSampleLanes = 64;     // lanes per block (one wave64)
constexpr int kMaxWords = 16;        // Pauli words per reduction launch
constexpr uint64_t kStreamChunkPacks = 4096;  // packs (16 B) per block iteration of the streaming kernels
constexpr int kProductGroupBits = 10;         // index bits per LDS factor table of k_product

enum TileOpKind : uint8_t {
  OP_U2 = 1,     // 2x2 on a register bit, optional controls
  OP_QFT_F = 2,  // forward QFT stage: butterfly then ladder phase on the 1 branch
  OP_QFT_I = 3,  // inverse QFT stage: conjugate ladder phase then butterfly
  OP_DIAG = 4,   // diagonal over 1 or 2 index bits (anywhere in the index)
  OP_X = 5,      // X on a register bit, optional controls (CNOT, Toffoli): a register swap
  OP_SIGN = 6,   // diagonal whose entries are +1 or -1 (CZ, Z): sign flips, mask in jl
};

// One operation of a tile program. Matrix entries are stored in the
// amplitude precision; the host rounds its double-precision values once.
// Every field is a whole dword: the kernel indexes ops with a run-time
// (wave-uniform) op number, and dword fields load with scalar loads, while
// byte fields at a run-time offset would need vector loads.
template <typename T>
struct TileOp {
  uint32_t kind;    // TileOpKind
  int32_t rpos;     // target register bit (U2, QFT)
  int32_t ra;       // DIAG: register bit of index bit a, or -1
  int32_t rb;       // DIAG: register bit of index bit b, or -1
  uint32_t jl;      // QFT: ladder register bits directly below rpos; SIGN: bit k set = entry k is -1
  uint32_t nbits;   // DIAG: 1 or 2
  uint32_t has_ctrl;// U2: controls present
  uint32_t pbit;    // QFT: index bit p of the stage
  uint32_t lb;      // QFT: lowest index bit of the ladder
  uint32_t cmask_r; // U2: control bits that are register bits (register positions)
  uint64_t cmask_t; // U2: control bits that are not register bits (index mask)
  uint64_t amask;   // DIAG: index mask of bit a (when ra < 0)
  uint64_t bmask;   // DIAG: index mask of bit b (when rb < 0)
  uint32_t pa;      // DIAG / SIGN: registers holding bit a = 1 (bit r set), 0 when ra < 0
  uint32_t pb;      // DIAG / SIGN: same for bit b
  T m[8];
};

constexpr int kTileThreadBits = 9;  // up to 512 threads per tile

// One segment: registers hold the local bits [reg0, reg0 + RB). The host
// precomputes where each thread-id bit and each register bit lands, in the
// tile (local bit) and in the amplitude index, so the kernel never looks
// up a table with a run-time index.
struct TileSeg {
  uint32_t reg0;
  uint32_t op_begin;  // ops [op_begin, op_end) run in this segment
  uint32_t op_end;
  uint32_t tmask;     // thread-id bits in use (tiles smaller than 256 x 2^RB
                      // amplitudes run duplicate threads with equal values)
  uint32_t tl_bit[kTileThreadBits];  // local bit of thread-id bit k
  uint32_t gi_bit[kTileThreadBits];  // index bit of thread-id bit k
  uint64_t gstr[5];                  // index stride of register bit j
};

// The tile number's bits are scattered onto the non-local index bits in
// at most kMaxRuns runs: base |= ((tile >> src) & (2^width - 1)) << dst.
constexpr int kMaxRuns = 16;

template <typename T>
struct TileArgs {
  uint64_t n_tiles;        // tiles = 2^(n - nlocal)
  uint32_t nlocal;         // local bits per tile
  uint32_t nseg;
  uint32_t nruns;
  uint32_t xcd_groups;     // > 1: consecutive tiles go to blocks of one XCD
  uint32_t run_src[kMaxRuns];
  uint32_t run_width[kMaxRuns];
  uint32_t run_dst[kMaxRuns];
  TileSeg seg[kMaxSegs];
  T scale;                 // applied at the final store (QFT 1/sqrt2 factors)
  uint32_t prefetch;       // 1: load the next tile's first segment while this one runs
  uint32_t from_basis;     // 1: the input is the basis state |basis>, generated instead of read
  uint32_t pad;
  uint64_t basis;
  T tab[TileCfg<T>::kTabVals];  // QFT: exp(2*pi*i*m/32), m = 0..31, (re, im) at 0
  TileOp<T> ops[TileCfg<T>::kMaxOps];
};
static_assert(sizeof(TileArgs<float>) <= 4096, "tile kernel arguments exceed 4 KiB");
static_assert(sizeof(TileArgs<double>) <= 4096, "tile kernel arguments exceed 4 KiB");

struct PhaseWin {
  uint8_t s1, w1, s2, w2;  // lut index = ((i>>s1)&(2^w1-1)) | (((i>>s2)&(2^w2-1))<<w1)
  uint32_t off;            // offset of this window's table
};

struct PhaseArgs {
  uint64_t n_amps;
  uint32_t nwin;
  uint32_t nlow;           // windows [0, nlow) read index bits 0..1
  uint32_t lut_len;
  uint32_t xcd_groups;     // chunk-to-block mapping groups (as in the tile sweeps)
  PhaseWin win[kMaxPhaseWins];
};

constexpr int kMaxTileBits = 32;   // bits of a tile number
constexpr int kMaxTileGroups = 8;  // 4-bit groups of a tile number
constexpr int kThrTabOff = 2 * kMaxTileGroups * 16;  // thread parts in RotArgs::tile_tab
constexpr int kThrTabLen = 512;          // one entry per thread-id value of a 512-thread tile
constexpr int kFusedRdep = 4;      // deferred-diagonal windows a fused sweep evaluates per register
constexpr int kFusedWins = 8;      // deferred-diagonal windows a fused sweep takes in all

// First segment of a sweep that reads through deferred work (k_rot and k_qft
// with PERM): it loads amplitude y from src(y) = M'y ^ c of the source
// buffer (identity M' and c = 0 for a lone diagonal), times a deferred
// diagonal D(src(y)), and the sweep stores into the other buffer. src(y) =
// c ^ image of the tile number (four byte tables, ptab) ^ pthr(thread bits)
// ^ preg(register bits), each the XOR of the images of the set bits.
struct FusedRead {
  uint64_t pc;                          // c
  uint64_t pthr[kTileThreadBits];       // image of segment 0's thread-id bit k
  uint64_t preg[5];                     // image of segment 0's register bit j
  uint32_t ph_rdep;                     // windows [0, ph_rdep) of ph see the register images
  uint32_t ph_pad;
  uint16_t ph_rimg[kFusedRdep][32];     // their table-index bits of register image r
  PhaseArgs ph;                         // the deferred diagonal (nwin = 0: none; at most kFusedWins)
};

// Commuting rotation sweep (1q layers in Z-Y-Z split form, k_rot): per
// segment, a rotation on each register bit set in gmask, given by its three
// shear factors (a, b); a diagonal over all n index bits before the first
// segment (pre) and after the last one (post). Each diagonal is a table
// over the segment's register bits (with the constant) x factors of the
// segment's thread bits (table layout of OP_DTAB) x a per-tile factor, the
// product over 4-bit groups of the tile number of tile_tab[g][16] entries.
template <typename T>
struct RotArgs {
  uint64_t n_tiles;
  uint32_t nlocal;
  uint32_t nseg;
  uint32_t nruns;
  uint32_t xcd_groups;
  uint32_t run_src[kMaxRuns];
  uint32_t run_width[kMaxRuns];
  uint32_t run_dst[kMaxRuns];
  uint32_t gmask[kMaxSegs];
  uint32_t has_pre;
  uint32_t has_post;
  uint32_t ntgrp;             // 4-bit groups of the tile number in tile_tab
  uint32_t pad;
  // [2][kMaxTileGroups][16]: pre, post tile-number groups (unit entries past
  // ntgrp); then at kThrTabOff the thread parts, [2][512]: pre, post, entry
  // tm = the product of the factors of the set thread-id bits of tm
  const typename Vec2<T>::type *tile_tab;
  TileSeg seg[kMaxSegs];
  T shear[kMaxSegs][5][2];
  // In-wavefront swaps (k_rot only; zero elsewhere). After segment s's
  // rotations, swp[s] bit 0 swaps register bit 0 with lane bit 4 and bit 1
  // register bit 1 with lane bit 5 (v_permlane16/32_swap), and the bit each
  // brings into its register is rotated by shear2[s][0|1]. The segment's
  // LDS write then uses its layout after the swaps: thread-id bit 4 / 5
  // adds tfix[s][0] / [1] to the thread's local index (XORed), register
  // bit j sits at local bit mask wimg[s][j]. `out` is the last segment's
  // layout after its swaps (store addresses and the post diagonal).
  uint32_t swp[kMaxSegs];
  uint32_t tfix[kMaxSegs][2];
  uint32_t wimg[kMaxSegs][5];
  T shear2[kMaxSegs][2][2];
  TileSeg out;
  T pre[kDtabVals];
  T post[kDtabVals];
  FusedRead fr;                         // k_rot<.., PERM = true> and k_rot_relabel only
  // k_rot_relabel only: after segment rl_seg the tile changes frame through
  // the deferred permutation. Local index u of that segment moves to local
  // index Q u ^ delta(tile) of the next one; rl_thr[k] / rl_reg[j] are the
  // images Q u of thread-id bit k / register bit j of segment rl_seg.
  uint32_t rl_seg;
  uint16_t rl_thr[kTileThreadBits];
  uint16_t rl_reg[5];
};
static_assert(sizeof(RotArgs<float>) <= 4096, "rotation sweep arguments exceed 4 KiB");
static_assert(sizeof(RotArgs<double>) <= 4096, "rotation sweep arguments exceed 4 KiB");

// QFT sweep (k_qft): per segment, QFT stages on the register bits set in
// smask, the window holding consecutive index bits p0 .. p0 + RB - 1 with
// p0 >= lb. The stage at register bit J (index bit p = p0 + J) puts the
// phase exp(i pi L / 2^(p - lb)) on its |1> branch, L = index bits [lb, p):
// a register part exp(i pi v / 2^J) (v = the J register bits below) and a
// thread part exp(i pi x / 2^wbits[J]), x = (gb >> lb) & (2^xbits - 1), the
// thread and tile bits below the window.
template <typename T>
struct QftArgs {
  uint64_t n_tiles;
  uint32_t nlocal;
  uint32_t nseg;
  uint32_t nruns;
  uint32_t xcd_groups;
  uint32_t run_src[kMaxRuns];
  uint32_t run_width[kMaxRuns];
  uint32_t run_dst[kMaxRuns];
  uint32_t smask[kMaxSegs];
  uint32_t xbits[kMaxSegs];
  uint32_t wbits[kMaxSegs][5];
  uint32_t lb;
  uint32_t pad;
  T scale;
  TileSeg seg[kMaxSegs];
  FusedRead fr;            // k_qft<.., PERM = true> only
};
static_assert(sizeof(QftArgs<float>) <= 4096, "QFT sweep arguments exceed 4 KiB");
static_assert(sizeof(QftArgs<double>) <= 4096, "QFT sweep arguments exceed 4 KiB");

// Kernel launchers (msv_kernels.hip). All enqueue on `stream`.
template <typename T>
hipError_t launch_tile(typename Vec2<T>::type *state, const TileArgs<T> &args, int rb,
                       uint32_t grid, hipStream_t stream);
// perm_src != nullptr: the first segment reads through deferred work
// (args.fr) from perm_src, with the diagonal's tables at ph_lut and the
// tile-number images in four 256-entry byte tables at ptab, and the sweep
// writes `state`.
template <typename T>
hipError_t launch_rot(typename Vec2<T>::type *state, const RotArgs<T> &args, int rb, uint32_t grid,
                      hipStream_t stream, const typename Vec2<T>::type *perm_src = nullptr,
                      const typename Vec2<T>::type *ph_lut = nullptr, const uint64_t *ptab = nullptr);
// Rotation sweep that changes frame mid-tile (k_rot_relabel): reads tiles of
// src in the old frame, runs segments [0, rl_seg] there, multiplies by the
// deferred diagonal (args.fr.ph, tables at ph_lut), moves every amplitude
// to its new-frame slot, runs the remaining segments and stores into dst at
// rep(tile) ^ images of the last segment's bits (args.fr.pthr / preg). The
// tile-number byte tables at ptab hold rep | delta << 48 (entries 0..1023)
// and the old-frame tile base (1024..2047).
template <typename T>
hipError_t launch_rot_relabel(typename Vec2<T>::type *dst, const RotArgs<T> &args, int rb, uint32_t grid,
                              hipStream_t stream, const typename Vec2<T>::type *src,
                              const typename Vec2<T>::type *ph_lut, const uint64_t *ptab);
template <typename T>
hipError_t launch_qft(typename Vec2<T>::type *state, const QftArgs<T> &args, int rb, bool inverse,
                      uint32_t grid, hipStream_t stream, const typename Vec2<T>::type *perm_src = nullptr,
                      const typename Vec2<T>::type *ph_lut = nullptr, const uint64_t *ptab = nullptr);
template <typename T>
hipError_t launch_phase(typename Vec2<T>::type *state, const typename Vec2<T>::type *lut,
                        const PhaseArgs &args, uint32_t grid, hipStream_t stream);
template <typename T>
hipError_t launch_xor_gather(const typename Vec2<T>::type *src, typename Vec2<T>::type *dst,
                             uint64_t n_amps, const uint64_t *tables, int ntab, uint64_t c,
                             uint32_t grid, hipStream_t stream);
template <typename T>
hipError_t launch_xor_gather_phase(const typename Vec2<T>::type *src, typename Vec2<T>::type *dst,
                                   uint64_t n_amps, const uint64_t *tables, int ntab, uint64_t c,
                                   const typename Vec2<T>::type *lut, const PhaseArgs &phase,
                                   uint32_t grid, hipStream_t stream);
// Product state from three index-bit-group factor tables (k_product):
// groups of l1 and l2 <= kProductGroupBits low bits, the rest in `hi`.
template <typename T>
hipError_t launch_product(typename Vec2<T>::type *state, uint64_t n_amps, const typename Vec2<T>::type *lo,
                          const typename Vec2<T>::type *mid, const typename Vec2<T>::type *hi, uint32_t l1,
                          uint32_t l2, uint32_t grid, hipStream_t stream);
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
hipError_t launch_copy_probe(const float4 *src, float4 *dst, uint64_t n_vec, hipStream_t stream);
// Sampler steps on the device (msv_sample_noalloc): the in-place sequential
// prefix of the block sums at p[1..nb] (p[0] = 0), and per shot k of
// [shot0, shot0 + count) its block and residual.
hipError_t launch_prefix_seq(double *p, uint64_t nb, hipStream_t stream);
hipError_t launch_sample_search(const double *prefix, uint64_t nb, double total, uint64_t seed,
                                uint64_t stream_id, uint64_t shot0, uint64_t count, uint64_t *shot_block,
                                double *shot_resid, hipStream_t stream);


constexpr uint32_t kReduceThreads = 256;

// Host helpers of the runtime shared with the batched ansatz engine.
int set_error(int code, const std::string &msg);  // records the thread's last error, returns code
int device_cu_count(int device);
uint64_t device_max_alloc(int device);  // largest single allocation (HSA pool), 0 if unknown

}  // namespace msv

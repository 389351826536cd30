// Host runtime of the statevector engine: state lifecycle, sweep planning,
// parameter staging and the C ABI declared in include/mettleq_sv.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <set>
#include <vector>

#include "../../include/mettleq_sv.h"
#include "msv_internal.h"
#include "msv_philox.h"

using namespace msv;

// Device scratch of the allocation-free sampler (msv_sample_noalloc): the
// block-sum prefix and per-shot staging. Each GPU the library runs on holds
// its own copy until the process ends. Host code never uses this name as an
// address: it takes the device address from hipGetSymbolAddress once per
// device (noalloc_scratch below).
constexpr size_t kNoAllocScratchBytes = size_t(8) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define MSV_HIP(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(MSV_E_HIP, std::string(#expr) + " failed: " + hipGetErrorString(e_)); \
  } while (0)

#define MSV_TRY(expr)          \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != MSV_OK) return rc_; \
  } while (0)

// A kernel launch on state `st`: logged (kind MSV_K_*), and enqueued unless
// the state is virtual.
#define MSV_LAUNCH(st, kind, expr)        \
  do {                                    \
    record_dispatch((st), (kind));        \
    if (!(st)->dry) MSV_HIP(expr);        \
  } while (0)

// ---------------------------------------------------------------------------
// tuning
// ---------------------------------------------------------------------------

struct Tuning {
  int local_bits_c64 = 13;  // 2^13 complex64 = 64 KiB LDS per tile
  int local_bits_c128 = 12; // 2^12 complex128 = 64 KiB
  int gate_bits = 9;        // index bits above the columns per high sweep
  int phase_blocks = 0;     // 0 = CUs * 8
  int sweep_split = 0;      // high-sweep bit sets of 1q layers: 0 contiguous, 1 page-balanced
  int tile_blocks_per_cu = 4096;  // large: one tile per block up to 2^20 tiles (measured best)
  int elem_blocks_per_cu = 1024;  // one 64 KiB chunk per block up to 2^18 chunks
  int col_bits_c64 = 4;     // contiguous low index bits per tile row in high sweeps
  int rot_split_min = 2;    // min active gates of a 1q layer for the rotation-split form (0 = never)
  int xcd_groups = 8;       // tile-to-block mapping groups (gfx950: 8 XCDs); 0/1 = plain
  int reg_bits_c64 = 4;     // register bits per thread in c64 tile sweeps (4: 512-thread tiles)
  int tile_prefetch = 0;    // c64 512-thread tiles: next tile's loads overlap the current tile
  int defer_diag = 1;       // phase layers wait for the next operation (an XOR permutation absorbs them)
  int defer_perm = 1;       // XOR permutations wait for the next operation (a 1q layer's first sweep reads through them)
  int fused_blocks_per_cu = 32;  // grid of a first sweep that reads through deferred work (tables staged per block)
  int reorder_gates = 1;    // gate lists: commuting gates are regrouped into fewer tile programs
  int defer_tail = 1;       // 1q layers leave their last sweep's rotations to the next layer's first sweep
                            // (1: without a pending permutation only up to 64 GiB states, 2: always)
  int lane_swap = 1;        // rotation sweeps reach lane bits 4 and 5 through in-wavefront swaps
};
Tuning g_tune_store;  // written by msv_set_tuning under g_tune_mu
std::mutex g_tune_mu;
thread_local Tuning t_tune;
thread_local int t_tune_depth = 0;

// One snapshot of the tuning knobs per entry point: every planner of a call
// reads the same values even when another thread sets a knob meanwhile
// (msv_set_tuning takes effect from the next call on).
struct TuneScope {
  TuneScope() {
    if (t_tune_depth++ == 0) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      t_tune = g_tune_store;
    }
  }
  ~TuneScope() { --t_tune_depth; }
  TuneScope(const TuneScope &) = delete;
  TuneScope &operator=(const TuneScope &) = delete;
};

const Tuning &tune() {
  if (t_tune_depth == 0) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    t_tune = g_tune_store;
  }
  return t_tune;
}

// One factor of a diagonal: a table over one or two index bits.
struct DiagTerm {
  int bits[2];  // index bits; bits[1] < 0 for one-bit terms; bits[0] is the MSB of the table
  double re[4], im[4];
};

struct DevProps {
  int cu = 0;
  bool ok = false;
};
DevProps g_props[64];
std::mutex g_props_mu;

int dev_cu(int device) {
  std::lock_guard<std::mutex> lk(g_props_mu);
  if (device < 0 || device >= 64) return 256;
  if (!g_props[device].ok) {
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        cu <= 0)
      cu = 256;
    g_props[device].cu = cu;
    g_props[device].ok = true;
  }
  return g_props[device].cu;
}

}  // namespace

namespace msv {
int set_error(int code, const std::string &msg) { return fail(code, msg); }
int device_cu_count(int device) { return dev_cu(device); }
}  // namespace msv

// ---------------------------------------------------------------------------
// state
// ---------------------------------------------------------------------------

struct msv_state {
  // virtual (planning) state: no device memory; every layer call plans its
  // sweeps and logs its kernels without launching them
  bool dry = false;
  std::vector<int8_t> dlog;  // kinds of the kernels launched (MSV_K_*), in order
  uint64_t dlog_total = 0;   // launches since the last reset (the log keeps the first kDispatchLogMax)
  int n = 0;
  int dtype = MSV_C64;
  int device = 0;
  int cu = 256;
  uint64_t amps = 0;
  void *buf = nullptr;  // current amplitudes
  void *alt = nullptr;  // scratch for out-of-place permutations
  hipStream_t stream = nullptr;
  char *h_stage = nullptr;  // pinned host staging ring for kernel parameters
  char *d_stage = nullptr;  // device side of the ring
  size_t stage_cap = 0;
  size_t stage_head = 0;
  void *d_red = nullptr;  // reduction scratch
  size_t red_cap = 0;
  uint64_t sweeps = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // timer on this state's stream
  // Deferred diagonal: a phase layer that fits one k_phase launch is kept
  // here (windows + host copy of its lookup tables) and applied by the next
  // operation on the amplitudes. An XOR permutation absorbs it (the gather
  // multiplies by the diagonal at the source index, one sweep for both);
  // every other entry point launches it first (flush_pending).
  bool pend_on = false;
  PhaseArgs pend_args;
  std::vector<unsigned char> pend_lut;
  std::vector<DiagTerm> pend_terms;  // its factors (a crossing sweep re-windows them)
  // Deferred XOR permutation: the amplitudes are y -> D(src(y)) buf[src(y)]
  // with src(y) = M'y ^ c (perm_col[j] = M' column j); D is the deferred
  // diagonal above when pend_on. The next 1q layer reads its first sweep
  // through it; every other entry point runs the gather first.
  bool perm_on = false;
  std::vector<uint64_t> perm_col;
  uint64_t perm_c = 0;
  // Deferred basis state: reset / set_basis keep the index here instead of
  // writing the amplitudes; a 1q layer then writes its product state in one
  // pass, and every other operation writes the basis state first. Never on
  // together with the deferred work above.
  bool basis_on = false;
  uint64_t basis_k = 0;
  // Deferred rotation tail: the rotations of a 1q layer's last gate bits
  // (shear factors per bit) and that layer's post diagonal (ratio per index
  // bit, times tail_cst), not yet applied to buf. They act before the
  // deferred diagonal and permutation above. The next 1q layer runs them in
  // its first sweep ahead of its own rotations, crossing the permutation
  // inside the tile (k_rot_relabel); every other entry point runs them
  // first. Never on together with a deferred basis state.
  bool tail_on = false;
  std::vector<int> tail_bits;                   // ascending
  std::vector<std::array<double, 2>> tail_sh;   // per index bit
  std::vector<std::complex<double>> tail_post;  // per index bit
  std::complex<double> tail_cst = 1.0;
};

namespace {

size_t esize(int dtype) { return dtype == MSV_C128 ? 16 : 8; }

constexpr size_t kDispatchLogMax = size_t(1) << 16;

void record_dispatch(msv_state *s, int kind) {
  if (s->dlog.size() < kDispatchLogMax) s->dlog.push_back(static_cast<int8_t>(kind));
  ++s->dlog_total;
}

int check_state(const msv_state *s) {
  if (!s) return fail(MSV_E_INVALID, "null state handle");
  if (!s->buf) return fail(MSV_E_INVALID, "state has no amplitude buffer");
  return MSV_OK;
}

int set_dev(const msv_state *s) {
  if (s->dry) return MSV_OK;
  MSV_HIP(hipSetDevice(s->device));
  return MSV_OK;
}

// Set while an allocation-free entry point runs on this thread: the buffer
// helpers below then refuse to allocate instead of growing a buffer.
thread_local bool t_noalloc = false;

int refuse_alloc(const char *what, size_t bytes) {
  return fail(MSV_E_REFUSED, std::string("allocation-free call: the state's deferred work needs a new ") +
                                 what + " of " + std::to_string(bytes) +
                                 " bytes; call msv_state_sync first");
}

// Copies `bytes` of host data into the staging ring and enqueues the upload
// on the state's stream; returns the device address. Wrapping the ring first
// drains the stream, so no in-flight copy or kernel still reads a slot.
int stage(msv_state *s, const void *src, size_t bytes, void **dev) {
  const size_t need = (bytes + 255) & ~size_t(255);
  if (s->dry) {  // nothing is copied: a virtual state launches nothing
    *dev = reinterpret_cast<void *>(uintptr_t(256));
    return MSV_OK;
  }
  if (need > s->stage_cap) {
    if (t_noalloc) return refuse_alloc("staging ring", std::max<size_t>(need, size_t(1) << 20));
    MSV_HIP(hipStreamSynchronize(s->stream));
    if (s->h_stage) MSV_HIP(hipHostFree(s->h_stage));
    if (s->d_stage) MSV_HIP(hipFree(s->d_stage));
    s->h_stage = nullptr;
    s->d_stage = nullptr;
    size_t cap = std::max<size_t>(need, size_t(1) << 20);
    if (hipHostMalloc(reinterpret_cast<void **>(&s->h_stage), cap, hipHostMallocDefault) !=
        hipSuccess)
      return fail(MSV_E_OOM, "pinned staging allocation failed");
    if (hipMalloc(reinterpret_cast<void **>(&s->d_stage), cap) != hipSuccess)
      return fail(MSV_E_OOM, "device staging allocation failed");
    s->stage_cap = cap;
    s->stage_head = 0;
  }
  if (s->stage_head + need > s->stage_cap) {
    MSV_HIP(hipStreamSynchronize(s->stream));
    s->stage_head = 0;
  }
  char *h = s->h_stage + s->stage_head;
  char *d = s->d_stage + s->stage_head;
  std::memcpy(h, src, bytes);
  MSV_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s->stream));
  s->stage_head += need;
  *dev = d;
  return MSV_OK;
}

int ensure_red(msv_state *s, size_t bytes) {
  if (bytes <= s->red_cap) return MSV_OK;
  if (s->dry) return fail(MSV_E_UNSUPPORTED, "a virtual state holds no amplitudes to reduce");
  if (t_noalloc) return refuse_alloc("reduction buffer", bytes);
  MSV_HIP(hipStreamSynchronize(s->stream));
  if (s->d_red) MSV_HIP(hipFree(s->d_red));
  s->d_red = nullptr;
  s->red_cap = 0;
  const size_t cap = std::max<size_t>(bytes, size_t(1) << 20);
  if (hipMalloc(&s->d_red, cap) != hipSuccess)
    return fail(MSV_E_OOM, "reduction scratch allocation failed (" + std::to_string(cap) + " B)");
  s->red_cap = cap;
  return MSV_OK;
}

int ensure_alt(msv_state *s) {
  if (s->alt) return MSV_OK;
  if (s->dry) {
    s->alt = reinterpret_cast<void *>(uintptr_t(512));  // never dereferenced: nothing launches
    return MSV_OK;
  }
  if (t_noalloc) return refuse_alloc("second state buffer", s->amps * esize(s->dtype));
  const size_t bytes = s->amps * esize(s->dtype);
  if (hipMalloc(&s->alt, bytes) != hipSuccess) {
    s->alt = nullptr;
    (void)hipGetLastError();
    return fail(MSV_E_OOM, "out-of-place permutation needs a second " + std::to_string(bytes) +
                               "-byte state buffer and the allocation failed");
  }
  return MSV_OK;
}

// Grid of the chunked streaming kernels (phase, gather, scale): one block
// per chunk of kStreamChunkPacks 16-byte packs, at most elem_blocks_per_cu
// blocks per CU (each then takes several chunks).
uint32_t stream_grid(const msv_state *s) {
  const uint64_t packs = s->amps * (s->dtype == MSV_C128 ? 16 : 8) / 16;
  const uint64_t want = (packs + kStreamChunkPacks - 1) / kStreamChunkPacks;
  const uint64_t cap = static_cast<uint64_t>(s->cu) * tune().elem_blocks_per_cu;
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min(want, cap)));
}

uint32_t elem_grid(const msv_state *s, uint64_t work) {
  const uint64_t want = (work + 255) / 256;
  const uint64_t cap = static_cast<uint64_t>(s->cu) * tune().elem_blocks_per_cu;
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min(want, cap)));
}

int bit_of(const msv_state *s, int q) { return s->n - 1 - q; }

int check_qubit(const msv_state *s, int q, const char *what) {
  if (q < 0 || q >= s->n)
    return fail(MSV_E_INVALID, std::string(what) + ": qubit " + std::to_string(q) +
                                   " out of range for " + std::to_string(s->n) + " qubits");
  return MSV_OK;
}

// Deferred work (defined with the phase and permutation planners below).
template <typename T>
int flush_pending_T(msv_state *s);
template <typename T>
int init_basis_T(msv_state *s, uint64_t index);
template <typename T>
int launch_gather(msv_state *s, const std::vector<uint64_t> &col, uint64_t c);

// ---------------------------------------------------------------------------
// tile-sweep planning
// ---------------------------------------------------------------------------

struct PlanOp {
  uint8_t kind = 0;   // OP_U2 / OP_QFT_F / OP_QFT_I / OP_DIAG
  int tbit = -1;      // target index bit (U2, QFT)
  uint64_t cmask = 0; // controls (U2), index bits
  int abit = -1, bbit = -1;  // DIAG bits (bbit < 0: one-bit diagonal)
  int lb = 0;         // QFT ladder low bit
  double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

struct SegPlan {
  int reg0 = 0;  // registers hold local bits [reg0, reg0 + rb)
  std::vector<PlanOp> ops;
};

struct SweepPlan {
  std::vector<int> local;  // index bits of the tile, ascending
  std::vector<SegPlan> segs;
  double scale = 1.0;
};

template <typename T> int rb_of() {
  return std::is_same<T, float>::value ? tune().reg_bits_c64 : TileCfg<T>::kRB;
}
template <typename T> int local_max() {
  return std::is_same<T, float>::value ? tune().local_bits_c64 : tune().local_bits_c128;
}
// Local bits of a tile: the tuned size, at most what one 512-thread block
// holds (9 thread-id bits + the register bits).
template <typename T> int max_local() {
  return std::min(std::min(local_max<T>(), TileCfg<T>::kMaxLocal), kTileThreadBits + rb_of<T>());
}
template <typename T> int col_bits() { return std::is_same<T, float>::value ? tune().col_bits_c64 : 3; }

// Partition of index bits [0, n) into a low sweep [0, L) and high chunks of
// between rb and gate_bits bits each (ascending).
template <typename T>
void partition(int n, int *low_len, std::vector<std::pair<int, int>> *chunks) {
  const int NL = max_local<T>();
  const int RB = rb_of<T>();
  const int G = std::max(RB, std::min(tune().gate_bits, NL - col_bits<T>()));
  chunks->clear();
  if (n <= NL) {
    *low_len = n;
    return;
  }
  int H = n - NL;
  int L = NL;
  if (H < RB) {
    L = n - RB;
    H = RB;
  }
  const int k = (H + G - 1) / G;
  int start = L;
  for (int c = 0; c < k; ++c) {
    const int sz = H / k + (c < H % k ? 1 : 0);
    chunks->push_back({start, sz});
    start += sz;
  }
  *low_len = L;
}

// Local bits of a high sweep whose gate bits are `gates` (ascending): the
// lowest index bits as columns (at least col_bits(), more when the gates
// leave room in the tile, for longer contiguous rows), then the gates.
template <typename T>
std::vector<int> high_sweep_local_bits(const std::vector<int> &gates) {
  const int NL = max_local<T>();
  const int lowest = gates.empty() ? 0 : gates.front();
  const int C = std::min(std::max(col_bits<T>(), NL - static_cast<int>(gates.size())), lowest);
  std::vector<int> local;
  for (int b = 0; b < C; ++b) local.push_back(b);
  for (int b : gates) local.push_back(b);
  return local;
}

template <typename T>
std::vector<int> sweep_local_bits(int n, int low_len, bool is_low, std::pair<int, int> chunk) {
  (void)n;
  if (is_low) {
    std::vector<int> local;
    for (int b = 0; b < low_len; ++b) local.push_back(b);
    return local;
  }
  std::vector<int> gates;
  for (int b = chunk.first; b < chunk.first + chunk.second; ++b) gates.push_back(b);
  return high_sweep_local_bits<T>(gates);
}

// Gate-bit sets of the high sweeps of a commuting layer. Policy 0: the
// contiguous chunks of partition(). Policy 1: same sizes, but the bits at
// or above the 2 MiB page boundary are spread evenly over the sweeps, so
// that no sweep's tile touches many more pages than another's.
template <typename T>
std::vector<std::vector<int>> high_gate_sets(const std::vector<std::pair<int, int>> &chunks) {
  std::vector<std::vector<int>> sets;
  if (tune().sweep_split != 1 || chunks.size() < 2) {
    for (auto &ch : chunks) {
      std::vector<int> g;
      for (int b = ch.first; b < ch.first + ch.second; ++b) g.push_back(b);
      sets.push_back(g);
    }
    return sets;
  }
  const int page_bit = std::is_same<T, float>::value ? 18 : 17;  // 2 MiB / amplitude size
  std::vector<int> below, above;
  for (auto &ch : chunks)
    for (int b = ch.first; b < ch.first + ch.second; ++b) (b < page_bit ? below : above).push_back(b);
  const size_t k = chunks.size();
  size_t ib = 0, ia = 0;
  for (size_t c = 0; c < k; ++c) {
    const size_t want_below = below.size() / k + (c < below.size() % k ? 1 : 0);
    std::vector<int> g;
    for (size_t t = 0; t < want_below && ib < below.size(); ++t) g.push_back(below[ib++]);
    while (static_cast<int>(g.size()) < chunks[c].second && ia < above.size()) g.push_back(above[ia++]);
    while (static_cast<int>(g.size()) < chunks[c].second && ib < below.size()) g.push_back(below[ib++]);
    std::sort(g.begin(), g.end());
    sets.push_back(g);
  }
  return sets;
}

// Windows of rb consecutive local bits covering the given local bits,
// taken from the top down; windows never start below `floor`.
std::vector<int> cover_top_down(const std::vector<int> &gate_locals, int nlocal, int rb, int floor) {
  std::vector<int> sorted = gate_locals;
  std::sort(sorted.begin(), sorted.end());
  std::vector<int> wins;
  std::vector<bool> covered(nlocal, false);
  for (int i = static_cast<int>(sorted.size()) - 1; i >= 0; --i) {
    const int b = sorted[i];
    if (covered[b]) continue;
    int lo = std::max(floor, b - rb + 1);
    lo = std::min(lo, nlocal - rb);
    lo = std::max(lo, 0);
    wins.push_back(lo);
    for (int k = lo; k < lo + rb && k < nlocal; ++k) covered[k] = true;
  }
  return wins;
}

// Global loads/stores of a segment are coalesced when the lanes run over
// the lowest local bits; that holds when the register window starts at or
// above local bit `good_from`.
int good_from(int nlocal, int rb) { return std::max(0, std::min(4, nlocal - rb)); }

template <typename T>
int launch_sweep(msv_state *s, const SweepPlan &plan) {
  using V = typename Vec2<T>::type;
  const int n = s->n;
  const int nlocal = static_cast<int>(plan.local.size());
  const int rb = std::min(rb_of<T>(), nlocal);
  if (plan.segs.empty()) return MSV_OK;
  if (static_cast<int>(plan.segs.size()) > kMaxSegs)
    return fail(MSV_E_INVALID, "internal: too many tile segments");
  TileArgs<T> A;
  std::memset(&A, 0, sizeof(A));
  A.nlocal = static_cast<uint32_t>(nlocal);
  A.nseg = static_cast<uint32_t>(plan.segs.size());
  int loc_of[64];
  for (int i = 0; i < 64; ++i) loc_of[i] = -1;
  uint64_t lmask = 0;
  for (int k = 0; k < nlocal; ++k) {
    loc_of[plan.local[k]] = k;
    lmask |= 1ull << plan.local[k];
  }
  // runs of non-local index bits, numbered by the tile index from bit 0 up
  {
    int src = 0, nr = 0;
    for (int b = 0; b < n;) {
      if (lmask >> b & 1ull) {
        ++b;
        continue;
      }
      int e = b;
      while (e < n && !(lmask >> e & 1ull)) ++e;
      if (nr >= kMaxRuns) return fail(MSV_E_INVALID, "internal: too many non-local bit runs");
      A.run_src[nr] = static_cast<uint32_t>(src);
      A.run_width[nr] = static_cast<uint32_t>(e - b);
      A.run_dst[nr] = static_cast<uint32_t>(b);
      src += e - b;
      ++nr;
      b = e;
    }
    A.nruns = static_cast<uint32_t>(nr);
  }
  A.n_tiles = 1ull << (n - nlocal);
  A.xcd_groups = static_cast<uint32_t>(std::max(0, tune().xcd_groups));
  A.scale = static_cast<T>(plan.scale);
  // table pool: QFT twiddles at 0, or the segment diagonals
  bool has_qft = false;
  for (const SegPlan &sp : plan.segs)
    for (const PlanOp &po : sp.ops) has_qft |= po.kind == OP_QFT_F || po.kind == OP_QFT_I;
  if (has_qft) {
    for (int m = 0; m < 32; ++m) {
      A.tab[2 * m] = static_cast<T>(std::cos(2.0 * M_PI * m / 32.0));
      A.tab[2 * m + 1] = static_cast<T>(std::sin(2.0 * M_PI * m / 32.0));
    }
  }
  if (nlocal - rb > kTileThreadBits)
    return fail(MSV_E_INVALID, "internal: tile larger than one block");
  int nops = 0;
  for (size_t si = 0; si < plan.segs.size(); ++si) {
    const SegPlan &sp = plan.segs[si];
    TileSeg &sg = A.seg[si];
    if (sp.reg0 < 0 || sp.reg0 + rb > nlocal)
      return fail(MSV_E_INVALID, "internal: register window outside the tile");
    sg.reg0 = static_cast<uint32_t>(sp.reg0);
    for (int j = 0; j < rb; ++j) sg.gstr[j] = 1ull << plan.local[sp.reg0 + j];
    // thread-id bits take the non-register local bits in increasing order
    sg.tmask = (1u << (nlocal - rb)) - 1u;
    for (int k = 0, t = 0; k < nlocal; ++k) {
      if (k >= sp.reg0 && k < sp.reg0 + rb) continue;
      sg.tl_bit[t] = static_cast<uint32_t>(k);
      sg.gi_bit[t] = static_cast<uint32_t>(plan.local[k]);
      ++t;
    }
    sg.op_begin = static_cast<uint32_t>(nops);
    for (const PlanOp &po : sp.ops) {
      if (nops >= TileCfg<T>::kMaxOps) return fail(MSV_E_INVALID, "internal: too many tile ops");
      TileOp<T> &op = A.ops[nops++];
      op.kind = po.kind;
      op.rpos = -1;
      op.ra = -1;
      op.rb = -1;
      for (int k = 0; k < 8; ++k) op.m[k] = static_cast<T>(po.m[k]);
      auto reg_pos = [&](int bit) -> int {
        const int k = loc_of[bit];
        if (k < 0) return -1;
        const int j = k - sp.reg0;
        return (j >= 0 && j < rb) ? j : -1;
      };
      if (po.kind == OP_U2 || po.kind == OP_QFT_F || po.kind == OP_QFT_I) {
        const int j = reg_pos(po.tbit);
        if (j < 0) return fail(MSV_E_INVALID, "internal: op target is not a register bit");
        op.rpos = static_cast<int8_t>(j);
      }
      if (po.kind == OP_U2) {
        uint64_t cm = po.cmask;
        while (cm) {
          const int bit = __builtin_ctzll(cm);
          cm &= cm - 1;
          const int j = reg_pos(bit);
          if (j >= 0) op.cmask_r |= 1u << j;
          else op.cmask_t |= 1ull << bit;
        }
        op.has_ctrl = po.cmask != 0 ? 1u : 0u;
        // an exact X (CNOT, Toffoli targets) runs as a register swap
        const double xm[8] = {0, 0, 1, 0, 1, 0, 0, 0};
        bool is_x = true;
        for (int k = 0; k < 8; ++k) is_x = is_x && po.m[k] == xm[k];
        if (is_x) op.kind = OP_X;
      } else if (po.kind == OP_QFT_F || po.kind == OP_QFT_I) {
        const int j = op.rpos;
        const int p = po.tbit;
        int jl = 0;
        for (int d = 1; d <= j; ++d) {
          const int idx = plan.local[sp.reg0 + j - d];
          if (idx == p - d && idx >= po.lb) ++jl;
          else break;
        }
        for (int d = jl + 1; d <= j; ++d) {
          if (plan.local[sp.reg0 + j - d] >= po.lb)
            return fail(MSV_E_INVALID, "internal: QFT register window not contiguous");
        }
        op.jl = static_cast<uint32_t>(jl);
        op.pbit = static_cast<uint32_t>(p);
        op.lb = static_cast<uint32_t>(po.lb);
      } else if (po.kind == OP_DIAG) {
        // registers whose index has register bit j set
        auto reg_pat = [&](int j) -> uint32_t {
          uint32_t m = 0;
          if (j >= 0)
            for (int r = 0; r < (1 << rb); ++r)
              if ((r >> j) & 1) m |= 1u << r;
          return m;
        };
        op.nbits = po.bbit >= 0 ? 2u : 1u;
        const int ja = reg_pos(po.abit);
        op.ra = static_cast<int8_t>(ja);
        op.amask = ja >= 0 ? 0ull : (1ull << po.abit);
        op.pa = reg_pat(ja);
        op.pb = 0;
        if (po.bbit >= 0) {
          const int jb = reg_pos(po.bbit);
          op.rb = static_cast<int8_t>(jb);
          op.bmask = jb >= 0 ? 0ull : (1ull << po.bbit);
          op.pb = reg_pat(jb);
        }
        // entries of exactly +1 / -1 (CZ, Z) run as sign flips
        uint32_t smask = 0;
        bool is_sign = true;
        for (uint32_t k = 0; k < (1u << op.nbits); ++k) {
          const double re = po.m[2 * k], im = po.m[2 * k + 1];
          is_sign = is_sign && im == 0.0 && (re == 1.0 || re == -1.0);
          if (re == -1.0) smask |= 1u << k;
        }
        if (is_sign) {
          op.kind = OP_SIGN;
          op.jl = smask;
        }
      }
    }
    sg.op_end = static_cast<uint32_t>(nops);
  }
  uint64_t cap = static_cast<uint64_t>(s->cu) * tune().tile_blocks_per_cu;
  // The prefetching form runs a resident grid (two 512-thread blocks per CU,
  // the LDS limit) and pays one discarded tile load per block, so it is used
  // only when every block gets at least 32 tiles.
  A.prefetch = 0;
  if (tune().tile_prefetch && std::is_same<T, float>::value && rb == 4) {
    const uint64_t resident = static_cast<uint64_t>(s->cu) * 2;
    if (A.n_tiles >= 32 * resident) {
      A.prefetch = 1;
      cap = resident;
    }
  }
  const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(A.n_tiles, cap)));
  // a deferred reset is consumed here: the sweep generates |basis> instead
  // of reading the state (callers flush every other kind of deferred work)
  if (s->basis_on) {
    if (s->pend_on || s->perm_on || s->tail_on) return fail(MSV_E_INVALID, "internal: deferred work before a tile sweep");
    A.from_basis = 1;
    A.basis = s->basis_k;
  }
  MSV_LAUNCH(s, MSV_K_TILE, launch_tile<T>(reinterpret_cast<V *>(s->buf), A, rb, grid, s->stream));
  s->basis_on = false;  // cleared once enqueued
  s->sweeps += 1;
  return MSV_OK;
}

// Shortest sequence of register windows (window = local bits [w, w + rb))
// that covers every bit of `need` and whose first and last windows start at
// or above `gf`, so that the global loads and stores stay coalesced. For a
// 13-bit tile with 5 register bits this is 3 windows, e.g. [8,13) [0,5)
// [4,9): two LDS exchanges per tile.
std::vector<int> commuting_windows(uint32_t need, int nlocal, int rb, int gf) {
  const int nw = nlocal - rb + 1;
  const uint32_t wmask = (1u << rb) - 1u;
  for (int m = 1; m <= kMaxSegs; ++m) {
    std::vector<int> w(m, gf);
    for (int i = 1; i + 1 < m; ++i) w[i] = 0;
    while (true) {
      uint32_t cov = 0;
      for (int x : w) cov |= wmask << x;
      if ((cov & need) == need) return w;
      // next sequence: first and last in [gf, nw), middle in [0, nw)
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

// Segments of a rotation sweep (k_rot): register windows and, per segment,
// the local bits that lane bits 4 and 5 bring into register bits 0 and 1
// by in-wavefront swaps after the window's rotations (-1: no swap). With
// the swaps a segment reaches 6 of the 13 bits of a 512-thread tile
// without LDS, so a sweep of 9 gate bits needs one LDS exchange instead of
// two. Swap-ins are bits no window covers, taken from the segment's
// non-window bits above its lowest 4 (thread bits 0-3 keep those, so the
// first and last segments stay coalesced).
struct RotSegs {
  std::vector<int> wins;
  std::vector<std::array<int, 2>> swaps;
};

// Thread-bit order of a segment with register window [w, w + rb): the
// non-window local bits ascending, except that thread bits 4 and 5 hold
// the swap-ins x[0] and x[1] when they are set.
std::vector<int> seg_thread_order(int nlocal, int w, int rb, const std::array<int, 2> &x) {
  std::vector<int> nw;
  for (int k = 0; k < nlocal; ++k)
    if (k < w || k >= w + rb) nw.push_back(k);
  const size_t lo = std::min<size_t>(4, nw.size());
  std::vector<int> ord(nw.begin(), nw.begin() + lo);
  std::vector<int> rest;
  for (size_t i = lo; i < nw.size(); ++i)
    if (nw[i] != x[0] && nw[i] != x[1]) rest.push_back(nw[i]);
  size_t at = 0;
  for (int i = 0; i < 2 && ord.size() < nw.size(); ++i) ord.push_back(x[i] >= 0 ? x[i] : rest[at++]);
  for (; at < rest.size(); ++at) ord.push_back(rest[at]);
  return ord;
}

RotSegs rot_segments(uint32_t need, int nlocal, int rb, int gf, bool lanes) {
  RotSegs plain;
  plain.wins = commuting_windows(need, nlocal, rb, gf);
  plain.swaps.assign(plain.wins.size(), {-1, -1});
  if (!lanes || rb < 2 || nlocal - rb < 6 || plain.wins.size() < 2) return plain;
  const int nw = nlocal - rb + 1;
  const uint32_t wmask = (1u << rb) - 1u;
  // only plans with fewer segments than the plain cover are worth swaps
  for (int m = 1; m < static_cast<int>(plain.wins.size()); ++m) {
    std::vector<int> w(m, gf);
    for (int i = 1; i + 1 < m; ++i) w[i] = 0;
    while (true) {
      uint32_t cov = 0;
      for (int x : w) cov |= wmask << x;
      std::vector<std::array<int, 2>> sw(m, {-1, -1});
      for (int i = 0; i < m; ++i) {
        int got = 0, seen = 0;
        for (int k = 0; k < nlocal && got < 2; ++k) {
          if (k >= w[i] && k < w[i] + rb) continue;
          if (seen++ < 4) continue;  // the segment's thread bits 0-3
          if (((need >> k) & 1u) && !((cov >> k) & 1u)) {
            sw[i][got++] = k;
            cov |= 1u << k;
          }
        }
      }
      if ((cov & need) == need) return {w, sw};
      int i = 0;
      for (; i < m; ++i) {
        const int lo = (i == 0 || i == m - 1) ? gf : 0;
        if (++w[i] < nw) break;
        w[i] = lo;
      }
      if (i == m) break;
    }
  }
  return plain;
}

// Commuting per-bit ops (one op per gate local bit, any order).
// `pre` runs before every op of the sweep (first segment), `post` after
// every op (last segment); both are diagonals (or absent: kind 0).
template <typename T>
int run_commuting_sweep(msv_state *s, const std::vector<int> &local,
                        const std::vector<std::pair<int, PlanOp>> &ops_by_local,
                        const PlanOp *pre = nullptr, const PlanOp *post = nullptr) {
  const int nlocal = static_cast<int>(local.size());
  const int rb = std::min(rb_of<T>(), nlocal);
  uint32_t need = 0;
  for (auto &p : ops_by_local) need |= 1u << p.first;
  const std::vector<int> order = commuting_windows(need, nlocal, rb, good_from(nlocal, rb));
  if (order.empty()) return fail(MSV_E_INVALID, "internal: no register window cover");
  SweepPlan plan;
  plan.local = local;
  std::vector<bool> used(ops_by_local.size(), false);
  for (int w : order) {
    SegPlan sp;
    sp.reg0 = w;
    if (plan.segs.empty() && pre) sp.ops.push_back(*pre);
    for (size_t i = 0; i < ops_by_local.size(); ++i) {
      const int l = ops_by_local[i].first;
      if (!used[i] && l >= w && l < w + rb) {
        sp.ops.push_back(ops_by_local[i].second);
        used[i] = true;
      }
    }
    plan.segs.push_back(sp);
  }
  if (post) plan.segs.back().ops.push_back(*post);
  return launch_sweep<T>(s, plan);
}

// U = diag(a1, b1) * [[c, -s], [s, c]] * diag(1, b2) for a unitary 2x2 U
// (the Z-Y-Z Euler form with the phases gathered into two diagonals).
struct RotSplit {
  std::complex<double> a1, b1, b2;
  double c, s;
};

bool split_rotation(const double *m, RotSplit *z) {
  using C = std::complex<double>;
  const C u00(m[0], m[1]), u01(m[2], m[3]), u10(m[4], m[5]), u11(m[6], m[7]);
  double c = 0.5 * (std::abs(u00) + std::abs(u11));
  double sn = 0.5 * (std::abs(u01) + std::abs(u10));
  const double nrm = std::hypot(c, sn);
  if (!(nrm > 0.5 && nrm < 2.0)) return false;
  c /= nrm;
  sn /= nrm;
  C a1 = 1.0, b1 = 1.0, b2 = 1.0;
  if (std::abs(u00) > 0.0) a1 = u00 / std::abs(u00);
  if (std::abs(u10) > 0.0) {
    b1 = u10 / std::abs(u10);
    if (std::abs(u01) > 0.0) b2 = -u01 / std::abs(u01) / a1;
    else if (std::abs(u11) > 0.0) b2 = u11 / std::abs(u11) / b1;
  } else if (std::abs(u11) > 0.0) {
    b2 = u11 / std::abs(u11);
  }
  // the split must reproduce U (it does for unitaries)
  const C r00 = a1 * c, r01 = -a1 * sn * b2, r10 = b1 * sn, r11 = b1 * c * b2;
  const double err = std::max(std::max(std::abs(r00 - u00), std::abs(r01 - u01)),
                              std::max(std::abs(r10 - u10), std::abs(r11 - u11)));
  if (!(err < 1e-9)) return false;
  z->a1 = a1;
  z->b1 = b1;
  z->b2 = b2;
  z->c = c;
  z->s = sn;
  return true;
}

// One commuting 1q sweep over `local` with general 2x2 ops on the active
// gate bits (the form for matrices that do not split into diagonals and a
// real rotation, e.g. non-unitary ones).
template <typename T>
int run_1q_sweep(msv_state *s, const std::vector<int> &local, const std::vector<int> &gate_locals,
                 const double *mats) {
  const int n = s->n;
  const int k = static_cast<int>(gate_locals.size());
  if (k == 0) return MSV_OK;
  std::vector<std::pair<int, PlanOp>> ops;
  for (int i = 0; i < k; ++i) {
    PlanOp po;
    po.kind = OP_U2;
    po.tbit = local[gate_locals[i]];
    const int q = n - 1 - po.tbit;
    for (int t = 0; t < 8; ++t) po.m[t] = mats[8 * q + t];
    ops.push_back({gate_locals[i], po});
  }
  return run_commuting_sweep<T>(s, local, ops);
}

// Tile geometry shared by every tile-sweep argument block: the runs that
// scatter the tile number onto the non-local index bits, and, per register
// window, where each thread-id bit and register bit lands.
template <typename Args>
int fill_tile_geometry(int n, const std::vector<int> &local, const std::vector<int> &windows, int rb,
                       Args *A) {
  const int nlocal = static_cast<int>(local.size());
  if (nlocal - rb > kTileThreadBits)
    return fail(MSV_E_INVALID, "internal: tile larger than one block");
  if (static_cast<int>(windows.size()) > kMaxSegs)
    return fail(MSV_E_INVALID, "internal: too many tile segments");
  uint64_t lmask = 0;
  for (int b : local) lmask |= 1ull << b;
  int src = 0, nr = 0;
  for (int b = 0; b < n;) {
    if (lmask >> b & 1ull) {
      ++b;
      continue;
    }
    int e = b;
    while (e < n && !(lmask >> e & 1ull)) ++e;
    if (nr >= kMaxRuns) return fail(MSV_E_INVALID, "internal: too many non-local bit runs");
    A->run_src[nr] = static_cast<uint32_t>(src);
    A->run_width[nr] = static_cast<uint32_t>(e - b);
    A->run_dst[nr] = static_cast<uint32_t>(b);
    src += e - b;
    ++nr;
    b = e;
  }
  A->nruns = static_cast<uint32_t>(nr);
  A->nlocal = static_cast<uint32_t>(nlocal);
  A->nseg = static_cast<uint32_t>(windows.size());
  A->n_tiles = 1ull << (n - nlocal);
  A->xcd_groups = static_cast<uint32_t>(std::max(0, tune().xcd_groups));
  for (size_t si = 0; si < windows.size(); ++si) {
    const int w = windows[si];
    if (w < 0 || w + rb > nlocal) return fail(MSV_E_INVALID, "internal: register window outside the tile");
    TileSeg &sg = A->seg[si];
    sg.reg0 = static_cast<uint32_t>(w);
    for (int j = 0; j < rb; ++j) sg.gstr[j] = 1ull << local[w + j];
    sg.tmask = (1u << (nlocal - rb)) - 1u;
    for (int k = 0, t = 0; k < nlocal; ++k) {
      if (k >= w && k < w + rb) continue;
      sg.tl_bit[t] = static_cast<uint32_t>(k);
      sg.gi_bit[t] = static_cast<uint32_t>(local[k]);
      ++t;
    }
  }
  return MSV_OK;
}

// Diagonal prod_b ratio[b]^bit_b(i) * cst over all n index bits, laid out for
// segment `sg` of a tile: register table (constant included) and thread-bit
// factors in `tab` (register entries at 0..63, thread-bit factors at
// 64..81, as dtab_apply_g reads them), tile-number groups in `tgrp`
// (kMaxTileGroups x 16 entries, unit past ntgrp).
// rloc: local bit of each register bit (null: the window sg.reg0 + j).
// thr: receives the thread part, kThrTabLen entries (entry tm = the product
// of the factors of the set bits of tm), the table the kernels index.
template <typename T, typename Args>
void build_layer_diag(const Args &A, const TileSeg &sg, const std::vector<int> &local, int rb,
                      const std::vector<std::complex<double>> &ratio, std::complex<double> cst,
                      T *tab, int ntgrp, std::vector<typename Vec2<T>::type> *tgrp,
                      std::vector<typename Vec2<T>::type> *thr, const int *rloc = nullptr) {
  using C = std::complex<double>;
  using V = typename Vec2<T>::type;
  const int nlocal = static_cast<int>(local.size());
  for (int r = 0; r < 32; ++r) {
    C d = 0.0;
    if (r < (1 << rb)) {
      d = cst;
      for (int j = 0; j < rb; ++j)
        if ((r >> j) & 1) d *= ratio[local[rloc ? rloc[j] : static_cast<int>(sg.reg0) + j]];
    }
    tab[2 * r] = static_cast<T>(d.real());
    tab[2 * r + 1] = static_cast<T>(d.imag());
  }
  C q[kTileThreadBits];
  for (int k = 0; k < kTileThreadBits; ++k) {
    q[k] = 1.0;
    if (k < nlocal - rb) q[k] = ratio[local[sg.tl_bit[k]]];
    tab[64 + 2 * k] = static_cast<T>(q[k].real());
    tab[65 + 2 * k] = static_cast<T>(q[k].imag());
  }
  {
    std::vector<C> f(kThrTabLen, 1.0);
    for (int tm = 1; tm < kThrTabLen; ++tm) f[tm] = f[tm & (tm - 1)] * q[__builtin_ctz(tm)];
    for (int tm = 0; tm < kThrTabLen; ++tm) {
      V e;
      e.x = static_cast<T>(f[tm].real());
      e.y = static_cast<T>(f[tm].imag());
      thr->push_back(e);
    }
  }
  // tile-number bit t lands on index bit run_dst + (t - run_src)
  std::vector<int> tbit_index;
  for (uint32_t r = 0; r < A.nruns; ++r)
    for (uint32_t w = 0; w < A.run_width[r]; ++w) tbit_index.push_back(static_cast<int>(A.run_dst[r] + w));
  for (int g = 0; g < kMaxTileGroups; ++g) {  // unit entries past ntgrp
    for (int v = 0; v < 16; ++v) {
      C f = 1.0;
      for (int i = 0; g < ntgrp && i < 4; ++i) {
        const int t = 4 * g + i;
        if (((v >> i) & 1) && t < static_cast<int>(tbit_index.size())) f *= ratio[tbit_index[t]];
      }
      V e;
      e.x = static_cast<T>(f.real());
      e.y = static_cast<T>(f.imag());
      tgrp->push_back(e);
    }
  }
}

// A 1q layer whose active gates all split as U = diag(a1, b1) R diag(1, b2)
// (Z-Y-Z form, R a real rotation): every sweep applies only the rotations,
// as three shears each (k_rot), and the layer's two diagonals, over all n
// index bits, are applied once: diag(1, b2) on every gate bit before the
// first sweep's rotations, diag(a1, b1) after the last sweep's. The
// diagonals commute with the rotations on other bits, so this equals the
// gate-by-gate layer.
// Largest deferred-diagonal table a fused first sweep keeps in LDS: two
// blocks of 64 KiB exchange buffer + 16 KiB tables fill a CU's 160 KiB.
constexpr size_t kFusedLutMaxBytes = 16 * 1024 - 16;  // plus the unit entry (pad_fused_windows)

// The fused sweeps read all kFusedWins window slots unconditionally, so that
// the table reads issue together: slots past nwin get a window of no bits
// whose table is a unit entry appended to `lut`.
template <typename V>
void pad_fused_windows(PhaseArgs &ph, std::vector<V> &lut) {
  V one;
  one.x = 1;
  one.y = 0;
  const uint32_t unit = static_cast<uint32_t>(lut.size());
  lut.push_back(one);
  for (uint32_t w = ph.nwin; w < static_cast<uint32_t>(kFusedWins); ++w) {
    ph.win[w].s1 = ph.win[w].w1 = ph.win[w].s2 = ph.win[w].w2 = 0;
    ph.win[w].off = unit;
  }
  ph.lut_len = static_cast<uint32_t>(lut.size());
}

// Whether a sweep whose segment 0 holds local bits [reg0, reg0 + rb) of
// `local` in registers can read through the state's deferred work
// (FusedRead): the diagonal's tables fit in LDS beside the exchange buffer
// and the register images touch at most kFusedRdep of its windows; the
// permutation keeps the lowest col_bits address bits among themselves, so
// every 2^col_bits-amplitude run of the sweep reads one run of the source.
// A permutation is read out of place (into the second buffer); a lone
// diagonal in place, since each tile then reads exactly what it writes.
template <typename T>
bool can_fuse_read(const msv_state *s, const std::vector<int> &local, int reg0, int rb) {
  if (!s->perm_on && !s->pend_on) return false;
  if (reg0 < 0 || reg0 + rb > static_cast<int>(local.size())) return false;
  if (s->pend_on) {
    if (s->pend_lut.size() > kFusedLutMaxBytes || s->pend_args.nwin > static_cast<uint32_t>(kFusedWins))
      return false;
    uint64_t umask = 0;
    for (int j = 0; j < rb; ++j) {
      const int b = local[reg0 + j];
      umask |= s->perm_on ? s->perm_col[b] : (1ull << b);
    }
    int ndep = 0;
    for (uint32_t w = 0; w < s->pend_args.nwin; ++w) {
      const PhaseWin &pw = s->pend_args.win[w];
      const uint64_t wm = (((1ull << pw.w1) - 1ull) << pw.s1) | (((1ull << pw.w2) - 1ull) << pw.s2);
      ndep += (wm & umask) ? 1 : 0;
    }
    if (ndep > kFusedRdep) return false;
  }
  if (s->perm_on) {
    const int cb = std::max(0, std::min(col_bits<T>(), s->n));
    const uint64_t low = (1ull << cb) - 1ull;
    for (int b = 0; b < cb; ++b)
      if (s->perm_col[b] & ~low) return false;
  }
  return true;
}

// Fill A.fr for a fused read of the state's deferred work by a sweep with
// tile geometry A (runs, segment 0) over `local`, and stage its tables.
template <typename T, typename Args>
int prepare_fused_read(msv_state *s, Args &A, const std::vector<int> &local, int rb, const uint64_t **d_ptab,
                       const typename Vec2<T>::type **d_lut) {
  using V = typename Vec2<T>::type;
  const int n = s->n;
  const int nlocal = static_cast<int>(local.size());
  std::vector<uint64_t> col(n);
  for (int b = 0; b < n; ++b) col[b] = s->perm_on ? s->perm_col[b] : (1ull << b);
  FusedRead &F = A.fr;
  F.pc = s->perm_on ? s->perm_c : 0;
  // images of the tile-number bits, as four byte tables
  std::vector<uint64_t> timg;
  for (uint32_t r = 0; r < A.nruns; ++r)
    for (uint32_t w = 0; w < A.run_width[r]; ++w) timg.push_back(col[A.run_dst[r] + w]);
  if (timg.size() > static_cast<size_t>(kMaxTileBits)) return fail(MSV_E_INVALID, "internal: tile number too wide");
  std::vector<uint64_t> ptab(4 * 256, 0);
  for (int b = 0; b < 4; ++b)
    for (int v = 0; v < 256; ++v) {
      uint64_t x = 0;
      for (int i = 0; i < 8; ++i) {
        const size_t k = 8 * static_cast<size_t>(b) + i;
        if (((v >> i) & 1) && k < timg.size()) x ^= timg[k];
      }
      ptab[256 * b + v] = x;
    }
  void *dp = nullptr;
  MSV_TRY(stage(s, ptab.data(), ptab.size() * sizeof(uint64_t), &dp));
  *d_ptab = reinterpret_cast<const uint64_t *>(dp);
  const TileSeg &s0 = A.seg[0];
  for (int k = 0; k < kTileThreadBits; ++k) F.pthr[k] = (k < nlocal - rb) ? col[s0.gi_bit[k]] : 0ull;
  uint64_t umask = 0;
  for (int j = 0; j < rb; ++j) {
    F.preg[j] = col[local[s0.reg0 + j]];
    umask |= F.preg[j];
  }
  *d_lut = nullptr;
  if (!s->pend_on) return MSV_OK;
  // windows that see the register images first (their count in ph_rdep)
  F.ph = s->pend_args;
  std::vector<PhaseWin> dep, rest;
  for (uint32_t w = 0; w < F.ph.nwin; ++w) {
    const PhaseWin &pw = s->pend_args.win[w];
    const uint64_t wm = (((1ull << pw.w1) - 1ull) << pw.s1) | (((1ull << pw.w2) - 1ull) << pw.s2);
    (wm & umask ? dep : rest).push_back(pw);
  }
  if (dep.size() > static_cast<size_t>(kFusedRdep) || F.ph.nwin > static_cast<uint32_t>(kFusedWins))
    return fail(MSV_E_INVALID, "internal: deferred diagonal too wide for the fused sweep");
  F.ph_rdep = static_cast<uint32_t>(dep.size());
  uint32_t w = 0;
  for (const PhaseWin &pw : dep) F.ph.win[w++] = pw;
  for (const PhaseWin &pw : rest) F.ph.win[w++] = pw;
  // table-index bits of the register images (XOR of the set bits' images)
  for (size_t k = 0; k < dep.size(); ++k) {
    const PhaseWin &pw = dep[k];
    for (int r = 0; r < (1 << rb); ++r) {
      uint64_t img = 0;
      for (int j = 0; j < rb; ++j)
        if ((r >> j) & 1) img ^= F.preg[j];
      const uint32_t lo = static_cast<uint32_t>(img >> pw.s1) & ((1u << pw.w1) - 1u);
      const uint32_t hi = static_cast<uint32_t>(img >> pw.s2) & ((1u << pw.w2) - 1u);
      F.ph_rimg[k][r] = static_cast<uint16_t>(lo | (hi << pw.w1));
    }
  }
  std::vector<V> lut(s->pend_lut.size() / sizeof(V));
  std::memcpy(lut.data(), s->pend_lut.data(), lut.size() * sizeof(V));
  pad_fused_windows(F.ph, lut);
  void *dl = nullptr;
  MSV_TRY(stage(s, lut.data(), lut.size() * sizeof(V), &dl));
  *d_lut = reinterpret_cast<const V *>(dl);
  return MSV_OK;
}

using Shear = std::array<double, 2>;

// Shear factors of a real rotation: [[c, -s], [s, c]] = X(a) Y(b) X(a) with
// a = -tan(theta/2) = -s/(1+c), b = s.
Shear shear_of(const RotSplit &z) { return {-z.s / (1.0 + z.c), z.s}; }

// Segments of a rotation sweep over `local` with rotations on the local
// positions `gl`: register windows, and each gate bit assigned to the first
// window that holds it (A.gmask / A.shear from sh[index bit]).
template <typename T>
void assign_rotations(RotArgs<T> &A, const std::vector<int> &local, const std::vector<int> &gl,
                      const std::vector<int> &wins, size_t seg0, int rb, const std::vector<Shear> &sh) {
  std::vector<bool> done(local.size(), false);
  for (size_t si = 0; si < wins.size(); ++si)
    for (int k : gl) {
      if (done[k] || k < wins[si] || k >= wins[si] + rb) continue;
      done[k] = true;
      const int j = k - wins[si];
      A.gmask[seg0 + si] |= 1u << j;
      A.shear[seg0 + si][j][0] = static_cast<T>(sh[local[k]][0]);
      A.shear[seg0 + si][j][1] = static_cast<T>(sh[local[k]][1]);
    }
}

// One k_rot sweep over the tile bits `local`: rotations on the index bits
// `gates` (shear factors sh[bit]), the diagonal `pre` (ratio per index bit)
// before them and `post` (ratio per index bit, times cst) after them, null
// for none. fuse: the sweep reads through the state's deferred diagonal and
// permutation (out of place when a permutation is pending).
template <typename T>
int rot_sweep_T(msv_state *s, const std::vector<int> &local, const std::vector<int> &gates,
                const std::vector<Shear> &sh, const std::vector<std::complex<double>> *pre,
                const std::vector<std::complex<double>> *post, std::complex<double> cst, bool fuse) {
  using V = typename Vec2<T>::type;
  using C = std::complex<double>;
  const int n = s->n;
  const int nlocal = static_cast<int>(local.size());
  const int rb = std::min(rb_of<T>(), nlocal);
  std::vector<int> gl;
  for (int k = 0; k < nlocal; ++k)
    if (std::find(gates.begin(), gates.end(), local[k]) != gates.end()) gl.push_back(k);
  uint32_t need = 0;
  for (int k : gl) need |= 1u << k;
  const RotSegs plan = rot_segments(need, nlocal, rb, good_from(nlocal, rb), tune().lane_swap != 0);
  const std::vector<int> &wins = plan.wins;
  if (wins.empty()) return fail(MSV_E_INVALID, "internal: no register window cover");
  RotArgs<T> A;
  std::memset(&A, 0, sizeof(A));
  MSV_TRY(fill_tile_geometry(n, local, wins, rb, &A));
  assign_rotations<T>(A, local, gl, wins, 0, rb, sh);
  int rloc[5] = {0, 0, 0, 0, 0};  // the last segment's register bits after its swaps
  for (size_t si = 0; si < wins.size(); ++si) {
    TileSeg &sg = A.seg[si];
    const std::array<int, 2> &x = plan.swaps[si];
    // thread bits 4 and 5 hold the swap-ins before the swaps
    const std::vector<int> ord = seg_thread_order(nlocal, wins[si], rb, x);
    for (size_t t = 0; t < ord.size(); ++t) {
      sg.tl_bit[t] = static_cast<uint32_t>(ord[t]);
      sg.gi_bit[t] = static_cast<uint32_t>(local[ord[t]]);
    }
    int rl[5];
    for (int j = 0; j < rb; ++j) rl[j] = wins[si] + j;
    TileSeg after = sg;
    for (int i = 0; i < 2; ++i) {
      if (x[i] < 0) continue;
      // register bit i <-> thread bit 4 + i
      A.swp[si] |= 1u << i;
      A.tfix[si][i] = (1u << x[i]) ^ (1u << (wins[si] + i));
      A.shear2[si][i][0] = static_cast<T>(sh[local[x[i]]][0]);
      A.shear2[si][i][1] = static_cast<T>(sh[local[x[i]]][1]);
      rl[i] = x[i];
      after.tl_bit[4 + i] = static_cast<uint32_t>(wins[si] + i);
      after.gi_bit[4 + i] = static_cast<uint32_t>(local[wins[si] + i]);
      after.gstr[i] = 1ull << local[x[i]];
    }
    for (int j = 0; j < rb; ++j) A.wimg[si][j] = 1u << rl[j];
    if (si + 1 == wins.size()) {
      A.out = after;
      for (int j = 0; j < rb; ++j) rloc[j] = rl[j];
    }
  }
  const int ntgrp = (n - nlocal + 3) / 4;
  if (ntgrp > kMaxTileGroups) return fail(MSV_E_INVALID, "internal: too many tile-number bits");
  A.ntgrp = static_cast<uint32_t>(ntgrp);
  std::vector<V> tgrp, thr;
  tgrp.reserve(2 * 16 * kMaxTileGroups + 2 * kThrTabLen);
  const std::vector<C> one(n, 1.0);
  A.has_pre = pre ? 1u : 0u;
  build_layer_diag<T>(A, A.seg[0], local, rb, pre ? *pre : one, 1.0, A.pre, ntgrp, &tgrp, &thr);
  A.has_post = post ? 1u : 0u;
  build_layer_diag<T>(A, A.out, local, rb, post ? *post : one, post ? cst : C(1.0), A.post, ntgrp, &tgrp,
                      &thr, rloc);
  if (pre || post) {
    tgrp.insert(tgrp.end(), thr.begin(), thr.end());  // thread parts at kThrTabOff
    void *d = nullptr;
    MSV_TRY(stage(s, tgrp.data(), tgrp.size() * sizeof(V), &d));
    A.tile_tab = reinterpret_cast<const V *>(d);
  }
  if (fuse) {
    const uint64_t *d_ptab = nullptr;
    const V *d_lut = nullptr;
    MSV_TRY(prepare_fused_read<T>(s, A, local, rb, &d_ptab, &d_lut));
    const bool oop = s->perm_on;  // a lone diagonal runs in place
    if (oop) MSV_TRY(ensure_alt(s));
    const uint64_t fcap = static_cast<uint64_t>(s->cu) * tune().fused_blocks_per_cu;
    const uint32_t fgrid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(A.n_tiles, fcap)));
    MSV_LAUNCH(s, MSV_K_ROT_FUSED, launch_rot<T>(reinterpret_cast<V *>(oop ? s->alt : s->buf), A, rb, fgrid, s->stream,
                          reinterpret_cast<const V *>(s->buf), d_lut, d_ptab));
    if (oop) std::swap(s->buf, s->alt);
    s->perm_on = false;
    s->pend_on = false;
  } else {
    const uint64_t cap = static_cast<uint64_t>(s->cu) * tune().tile_blocks_per_cu;
    const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(A.n_tiles, cap)));
    MSV_LAUNCH(s, MSV_K_ROT, launch_rot<T>(reinterpret_cast<V *>(s->buf), A, rb, grid, s->stream));
  }
  s->sweeps += 1;
  return MSV_OK;
}

// Runs a deferred rotation tail on its own (one sweep, in place): the
// deferred diagonal and permutation after it stay pending.
template <typename T>
int flush_tail_T(msv_state *s) {
  if (!s->tail_on) return MSV_OK;
  s->tail_on = false;
  const std::vector<int> local = high_sweep_local_bits<T>(s->tail_bits);
  const int rc =
      rot_sweep_T<T>(s, local, s->tail_bits, s->tail_sh, nullptr, &s->tail_post, s->tail_cst, false);
  if (rc != MSV_OK) s->tail_on = true;  // not enqueued: the tail stays pending
  return rc;
}

// Inverse of the deferred permutation's linear part: img[g] = the new-frame
// index vector whose source is old index bit g alone (col[q] = source of new
// bit q), and img_c = the new-frame index of old index 0.
bool perm_images(int n, const std::vector<uint64_t> &col, uint64_t c, std::vector<uint64_t> *img,
                 uint64_t *img_c) {
  // rows (old vector, new combination) reduced to the identity on the old side
  std::vector<std::pair<uint64_t, uint64_t>> rows(n);
  for (int q = 0; q < n; ++q) rows[q] = {col[q], 1ull << q};
  for (int g = 0; g < n; ++g) {
    int piv = -1;
    for (int r = g; r < n; ++r)
      if ((rows[r].first >> g) & 1ull) {
        piv = r;
        break;
      }
    if (piv < 0) return false;
    std::swap(rows[g], rows[piv]);
    for (int r = 0; r < n; ++r)
      if (r != g && ((rows[r].first >> g) & 1ull)) {
        rows[r].first ^= rows[g].first;
        rows[r].second ^= rows[g].second;
      }
  }
  img->assign(n, 0);
  for (int g = 0; g < n; ++g) (*img)[g] = rows[g].second;
  uint64_t y = 0;
  for (int g = 0; g < n; ++g)
    if ((c >> g) & 1ull) y ^= rows[g].second;
  *img_c = y;
  return true;
}

// Tile of a sweep that crosses the deferred permutation: old-frame tile bits
// V (the columns, the pending tail's bits and spares chosen so that the new
// frame gets the most coordinate bits), the new-frame coordinate ("head")
// bits q, those whose source col[q] lies inside V, and the extra basis
// vectors that complete the image of V.
struct CrossTile {
  std::vector<int> vloc;        // old-frame tile bits, ascending
  uint64_t vmask = 0;
  std::vector<int> head;        // new-frame coordinate bits, ascending
  uint64_t hmask = 0;
  std::vector<uint64_t> extra;  // further basis vectors (no head bits, reduced, pivots in piv)
  std::vector<int> piv;
};

// Head bits of old-frame tile bits `vmask` under the permutation with source
// columns col.
uint64_t head_of(int n, const std::vector<uint64_t> &col, uint64_t vmask) {
  uint64_t h = 0;
  for (int q = 0; q < n; ++q)
    if ((col[q] & ~vmask) == 0) h |= 1ull << q;
  return h;
}

// V = columns + `fixed` + greedy spares (each adds the bit that makes the
// most active bits head bits), up to nl bits.
uint64_t choose_tile_bits(int n, int nl, int cb, const std::vector<int> &fixed, const std::vector<uint64_t> &col,
                          const std::vector<bool> &act) {
  uint64_t v = (1ull << cb) - 1ull;
  for (int b : fixed) v |= 1ull << b;
  auto score = [&](uint64_t m) {
    const uint64_t h = head_of(n, col, m);
    int c = 0;
    for (int q = 0; q < n; ++q) c += ((h >> q) & 1ull) && act[q];
    return c;
  };
  while (__builtin_popcountll(v) < nl) {
    int best = -1, best_sc = -1;
    for (int x = 0; x < n; ++x) {
      if ((v >> x) & 1ull) continue;
      const int sc = score(v | (1ull << x));
      if (sc > best_sc) {
        best_sc = sc;
        best = x;
      }
    }
    if (best < 0) break;
    v |= 1ull << best;
  }
  return v;
}

bool cross_tile(int n, uint64_t vmask, const std::vector<uint64_t> &col, const std::vector<uint64_t> &img,
                int cb, CrossTile *t) {
  t->vmask = vmask;
  t->vloc.clear();
  for (int b = 0; b < n; ++b)
    if ((vmask >> b) & 1ull) t->vloc.push_back(b);
  t->hmask = head_of(n, col, vmask);
  t->head.clear();
  for (int q = 0; q < n; ++q)
    if ((t->hmask >> q) & 1ull) t->head.push_back(q);
  for (int q = 0; q < cb; ++q)
    if (!((t->hmask >> q) & 1ull)) return false;  // the new tile's rows must stay contiguous
  // extras: images of V without their head bits, reduced
  std::vector<uint64_t> rows;
  for (int g : t->vloc) rows.push_back(img[g] & ~t->hmask);
  t->extra.clear();
  t->piv.clear();
  for (uint64_t r : rows) {
    for (size_t k = 0; k < t->extra.size(); ++k)
      if ((r >> t->piv[k]) & 1ull) r ^= t->extra[k];
    if (!r) continue;
    const int p = __builtin_ctzll(r);
    for (size_t k = 0; k < t->extra.size(); ++k)
      if ((t->extra[k] >> p) & 1ull) t->extra[k] ^= r;
    t->extra.push_back(r);
    t->piv.push_back(p);
  }
  return t->head.size() + t->extra.size() == t->vloc.size();
}

// Coordinates of new-frame index vector v in the cross tile's basis (head
// bits in ascending order, then the extras), and what is left (zero on the
// head bits and the pivots): v = rest ^ sum of the selected basis vectors.
uint32_t cross_coords(const CrossTile &t, uint64_t v, uint64_t *rest) {
  uint32_t u = 0;
  for (size_t k = 0; k < t.head.size(); ++k)
    if ((v >> t.head[k]) & 1ull) u |= 1u << k;
  v &= ~t.hmask;
  for (size_t k = 0; k < t.extra.size(); ++k)
    if ((v >> t.piv[k]) & 1ull) {
      u |= 1u << (t.head.size() + k);
      v ^= t.extra[k];
    }
  *rest = v;
  return u;
}

// The sweep that runs the pending tail (old frame: its rotations, its
// layer's post diagonal, then the deferred diagonal), crosses the deferred
// permutation inside the tile and runs this layer's rotations on the
// active head bits (their pre diagonal first). Fails (MSV_E_INVALID,
// nothing launched) when the tile does not fit the kernel's limits.
struct Win {
  std::vector<int> bits;  // distinct index bits, ascending
  int s1 = 0, w1 = 0, s2 = 0, w2 = 0;
  std::vector<int> terms;
};

// Cover a sorted bit set with at most two runs of total width <= wmax.
bool fit_runs(const std::vector<int> &bits, int wmax, int *s1, int *w1, int *s2, int *w2) {
  if (bits.empty()) return false;
  const int lo = bits.front(), hi = bits.back();
  if (hi - lo + 1 <= wmax) {
    *s1 = lo;
    *w1 = hi - lo + 1;
    *s2 = 0;
    *w2 = 0;
    return true;
  }
  int gap_at = -1, gap = -1;
  for (size_t i = 0; i + 1 < bits.size(); ++i) {
    const int g = bits[i + 1] - bits[i];
    if (g > gap) {
      gap = g;
      gap_at = static_cast<int>(i);
    }
  }
  const int a_lo = lo, a_hi = bits[gap_at], b_lo = bits[gap_at + 1], b_hi = hi;
  const int wa = a_hi - a_lo + 1, wb = b_hi - b_lo + 1;
  if (wa + wb > wmax) return false;
  *s1 = a_lo;
  *w1 = wa;
  *s2 = b_lo;
  *w2 = wb;
  return true;
}

constexpr int kWinMaxBits = 8;  // index bits of one window's table

// Greedy windows: term ti joins the first window whose bits, with the
// term's, still fit two runs of kWinMaxBits in all; else it opens a new one.
// Windows never cover a bit of `avoid` (their tables then do not depend on
// those bits), except window `keep` (< 0: none) placed first, which takes
// the terms listed in `first`.
bool place_terms(const std::vector<DiagTerm> &terms, const std::vector<int> &first, uint64_t avoid,
                 std::vector<Win> *out) {
  std::vector<Win> &wins = *out;
  auto covers_avoid = [&](const Win &w) {
    for (int b = w.s1; b < w.s1 + w.w1; ++b)
      if ((avoid >> b) & 1ull) return true;
    for (int b = w.s2; b < w.s2 + w.w2; ++b)
      if ((avoid >> b) & 1ull) return true;
    return false;
  };
  std::vector<bool> done(terms.size(), false);
  if (!first.empty()) {
    Win w;
    for (int ti : first) {
      const DiagTerm &t = terms[ti];
      for (int k = 0; k < 2; ++k)
        if (t.bits[k] >= 0 && std::find(w.bits.begin(), w.bits.end(), t.bits[k]) == w.bits.end())
          w.bits.push_back(t.bits[k]);
      w.terms.push_back(ti);
      done[ti] = true;
    }
    std::sort(w.bits.begin(), w.bits.end());
    if (!fit_runs(w.bits, kWinMaxBits, &w.s1, &w.w1, &w.s2, &w.w2)) return false;
    wins.push_back(w);
  }
  const size_t fixed = wins.size();
  for (size_t ti = 0; ti < terms.size(); ++ti) {
    if (done[ti]) continue;
    const DiagTerm &t = terms[ti];
    std::vector<int> tb = {t.bits[0]};
    if (t.bits[1] >= 0) tb.push_back(t.bits[1]);
    bool placed = false;
    for (size_t wi = fixed; wi < wins.size() && !placed; ++wi) {
      Win &w = wins[wi];
      Win c = w;
      for (int b : tb)
        if (std::find(c.bits.begin(), c.bits.end(), b) == c.bits.end()) c.bits.push_back(b);
      std::sort(c.bits.begin(), c.bits.end());
      if (fit_runs(c.bits, kWinMaxBits, &c.s1, &c.w1, &c.s2, &c.w2) && !covers_avoid(c)) {
        c.terms.push_back(static_cast<int>(ti));
        w = c;
        placed = true;
      }
    }
    if (!placed) {
      Win w;
      w.bits = tb;
      std::sort(w.bits.begin(), w.bits.end());
      fit_runs(w.bits, kWinMaxBits, &w.s1, &w.w1, &w.s2, &w.w2);
      if (covers_avoid(w)) return false;
      w.terms.push_back(static_cast<int>(ti));
      wins.push_back(w);
    }
  }
  return true;
}

// Tables of windows [wi, ...) into one PhaseArgs (up to kMaxPhaseWins
// windows and kPhaseLutMax entries), products formed in double precision
// and rounded once; `cst` multiplies the first window's table. Returns the
// first window not taken.
template <typename T>
size_t window_tables(const msv_state *s, const std::vector<DiagTerm> &terms, const std::vector<Win> &wins,
                     size_t wi, std::complex<double> cst, PhaseArgs *out,
                     std::vector<typename Vec2<T>::type> *lut) {
  using V = typename Vec2<T>::type;
  PhaseArgs &A = *out;
  std::memset(&A, 0, sizeof(A));
  A.n_amps = s->amps;
  uint32_t nw = 0, nlow = 0;
  while (wi < wins.size() && nw < static_cast<uint32_t>(kMaxPhaseWins)) {
    const Win &w = wins[wi];
    const uint32_t len = 1u << (w.w1 + w.w2);
    if (lut->size() + len > static_cast<size_t>(kPhaseLutMax) && nw > 0) break;
    PhaseWin pw;
    pw.s1 = static_cast<uint8_t>(w.s1);
    pw.w1 = static_cast<uint8_t>(w.w1);
    pw.s2 = static_cast<uint8_t>(w.s2);
    pw.w2 = static_cast<uint8_t>(w.w2);
    pw.off = static_cast<uint32_t>(lut->size());
    for (uint32_t v = 0; v < len; ++v) {
      auto bit_val = [&](int bit) -> int {
        if (bit >= w.s1 && bit < w.s1 + w.w1) return (v >> (bit - w.s1)) & 1;
        return (v >> (w.w1 + bit - w.s2)) & 1;
      };
      double pr = nw == 0 ? cst.real() : 1.0, pi = nw == 0 ? cst.imag() : 0.0;
      for (int ti : w.terms) {
        const DiagTerm &t = terms[ti];
        int sel = bit_val(t.bits[0]);
        if (t.bits[1] >= 0) sel = sel * 2 + bit_val(t.bits[1]);
        const double nr = pr * t.re[sel] - pi * t.im[sel];
        const double ni = pr * t.im[sel] + pi * t.re[sel];
        pr = nr;
        pi = ni;
      }
      V e;
      e.x = static_cast<T>(pr);
      e.y = static_cast<T>(pi);
      lut->push_back(e);
    }
    A.win[nw++] = pw;
    if (w.bits.front() == 0) nlow = nw;
    ++wi;
  }
  A.nwin = nw;
  A.nlow = nlow;
  A.xcd_groups = static_cast<uint32_t>(std::max(0, tune().xcd_groups));
  A.lut_len = static_cast<uint32_t>(lut->size());
  return wi;
}

// The deferred diagonal of a crossing sweep, windowed for the crossing
// segment's register bits (umask): every factor that reads one of them goes
// into window 0, evaluated per register; the other windows avoid those bits
// and give one product per thread and tile. The pending tail's post diagonal
// joins as one-bit factors (its constant scales window 0's table), so the
// sweep has no post pass of its own. False when the factors do not fit the
// fused sweep's limits; the deferred diagonal's own windows are used then.
template <typename T>
bool crossing_windows(msv_state *s, const std::vector<int> &vloc, const TileSeg &sgo, int rb,
                      uint64_t umask, FusedRead *F, std::vector<typename Vec2<T>::type> *lut) {
  using V = typename Vec2<T>::type;
  using C = std::complex<double>;
  if (s->pend_terms.empty()) return false;
  std::vector<DiagTerm> terms = s->pend_terms;
  for (int b : s->tail_bits) {
    if (s->tail_post[b] == C(1.0)) continue;
    DiagTerm t;
    std::memset(&t, 0, sizeof(t));
    t.bits[0] = b;
    t.bits[1] = -1;
    t.re[0] = 1.0;
    t.re[1] = s->tail_post[b].real();
    t.im[1] = s->tail_post[b].imag();
    terms.push_back(t);
  }
  std::vector<int> first;
  for (size_t i = 0; i < terms.size(); ++i) {
    const DiagTerm &t = terms[i];
    if (((umask >> t.bits[0]) & 1ull) || (t.bits[1] >= 0 && ((umask >> t.bits[1]) & 1ull)))
      first.push_back(static_cast<int>(i));
  }
  std::vector<Win> wins;
  if (!place_terms(terms, first, umask, &wins)) return false;
  if (wins.size() > static_cast<size_t>(kFusedWins)) return false;
  PhaseArgs A;
  lut->clear();
  if (window_tables<T>(s, terms, wins, 0, s->tail_cst, &A, lut) != wins.size()) return false;
  if (lut->size() * sizeof(V) > kFusedLutMaxBytes) return false;
  F->ph = A;
  F->ph_rdep = first.empty() ? 0u : 1u;
  if (!first.empty()) {
    const PhaseWin &pw = A.win[0];
    for (int r = 0; r < (1 << rb); ++r) {
      uint64_t im = 0;
      for (int j = 0; j < rb; ++j)
        if ((r >> j) & 1) im |= 1ull << vloc[sgo.reg0 + j];
      const uint32_t lo = static_cast<uint32_t>(im >> pw.s1) & ((1u << pw.w1) - 1u);
      const uint32_t hi = static_cast<uint32_t>(im >> pw.s2) & ((1u << pw.w2) - 1u);
      F->ph_rimg[0][r] = static_cast<uint16_t>(lo | (hi << pw.w1));
    }
  }
  return true;
}

template <typename T>
int cross_sweep_T(msv_state *s, const CrossTile &t, const std::vector<bool> &act, const std::vector<Shear> &sh,
                  const std::vector<std::complex<double>> &pre, bool pre_on, bool launch,
                  uint64_t *rotated = nullptr) {
  using V = typename Vec2<T>::type;
  using C = std::complex<double>;
  const int n = s->n;
  const int nlocal = static_cast<int>(t.vloc.size());
  const int nh = static_cast<int>(t.head.size());
  const int rb = std::min(rb_of<T>(), nlocal);
  const int gf = good_from(nlocal, rb);
  const int cb = std::max(0, std::min(col_bits<T>(), n));
  std::vector<uint64_t> col(n), img;
  uint64_t img_c = 0;
  for (int q = 0; q < n; ++q) col[q] = s->perm_on ? s->perm_col[q] : (1ull << q);
  if (!perm_images(n, col, s->perm_on ? s->perm_c : 0, &img, &img_c))
    return fail(MSV_E_INVALID, "internal: deferred permutation not invertible");
  // old frame: the tail's rotations
  std::vector<int> gl_o;
  for (int k = 0; k < nlocal; ++k)
    if (std::find(s->tail_bits.begin(), s->tail_bits.end(), t.vloc[k]) != s->tail_bits.end()) gl_o.push_back(k);
  if (gl_o.size() != s->tail_bits.size()) return fail(MSV_E_INVALID, "internal: cross tile");
  uint32_t need_o = 0;
  for (int k : gl_o) need_o |= 1u << k;
  const std::vector<int> wo = commuting_windows(need_o, nlocal, rb, gf);
  if (wo.empty()) return fail(MSV_E_INVALID, "internal: cross tile needs too many segments");
  // Canonical coordinates of the new tile (head bits ascending, then the
  // extras, as cross_coords gives them) go to local positions: the columns
  // first; then, for each lane bit of the crossing segment in turn, one
  // coordinate its image brings in (so that the lanes of a wave write slots
  // that differ in the lowest bits, as in a plain exchange); then the rest.
  std::vector<uint32_t> qc(nlocal);
  for (int k = 0; k < nlocal; ++k) {
    uint64_t rest = 0;
    qc[k] = cross_coords(t, img[t.vloc[k]], &rest);
    if (rest) return fail(MSV_E_INVALID, "internal: cross tile basis");
  }
  std::vector<int> cpos(nlocal, -1);
  int next = 0;
  for (int c = 0; c < cb; ++c) cpos[c] = next++;  // head bits 0 .. cb-1 are the columns
  {
    int lane = 0;
    for (int k = 0; k < nlocal && lane < 6; ++k) {
      if (k >= wo.back() && k < wo.back() + rb) continue;  // a register bit of the crossing segment
      ++lane;
      for (int c = 0; c < nlocal; ++c)
        if (((qc[k] >> c) & 1u) && cpos[c] < 0) {
          cpos[c] = next++;
          break;
        }
    }
  }
  for (int c = 0; c < nlocal; ++c)
    if (cpos[c] < 0) cpos[c] = next++;
  auto to_local = [&](uint32_t uc) {
    uint32_t u = 0;
    for (int c = 0; c < nlocal; ++c)
      if ((uc >> c) & 1u) u |= 1u << cpos[c];
    return u;
  };
  // new-frame local positions as pseudo index bits (head bit q, extras n)
  // and as index vectors
  std::vector<int> nloc(nlocal, n);
  std::vector<uint64_t> bvec(nlocal, 0);
  for (int c = 0; c < nlocal; ++c) {
    nloc[cpos[c]] = c < nh ? t.head[c] : n;
    bvec[cpos[c]] = c < nh ? (1ull << t.head[c]) : t.extra[c - nh];
  }
  // the active head bits this sweep rotates: all of them, less the highest
  // positions while the segments would not fit (the rest rotate in the
  // plain sweeps that follow)
  std::vector<int> gl_n;
  for (int k = 0; k < nlocal; ++k)
    if (nloc[k] < n && act[nloc[k]]) gl_n.push_back(k);
  std::vector<int> wn;
  while (!gl_n.empty()) {
    uint32_t need_n = 0;
    for (int k : gl_n) need_n |= 1u << k;
    wn = commuting_windows(need_n, nlocal, rb, gf);
    if (!wn.empty() && wo.size() + wn.size() <= static_cast<size_t>(kMaxSegs)) break;
    gl_n.pop_back();
  }
  if (gl_n.empty()) return fail(MSV_E_INVALID, "internal: cross tile needs too many segments");
  if (rotated) {
    *rotated = 0;
    for (int k : gl_n) *rotated |= 1ull << nloc[k];
  }
  std::vector<int> wins = wo;
  wins.insert(wins.end(), wn.begin(), wn.end());
  RotArgs<T> A;
  std::memset(&A, 0, sizeof(A));
  MSV_TRY(fill_tile_geometry(n, t.vloc, wins, rb, &A));
  const uint32_t so = static_cast<uint32_t>(wo.size() - 1);
  A.rl_seg = so;
  assign_rotations<T>(A, t.vloc, gl_o, wo, 0, rb, s->tail_sh);
  std::vector<Shear> sh_n(n + 1, Shear{0.0, 0.0});
  for (int q : t.head) sh_n[q] = sh[q];
  assign_rotations<T>(A, nloc, gl_n, wn, wo.size(), rb, sh_n);
  // diagonals, both local to the tile: this layer's pre on the active head
  // bits after the crossing, the tail's post (on the tail bits) before it
  A.ntgrp = 0;
  std::vector<C> pre_n(n + 1, 1.0);
  for (int k : gl_n) pre_n[nloc[k]] = pre[nloc[k]];
  std::vector<V> tgrp, thr;
  tgrp.reserve(2 * 16 * kMaxTileGroups + 2 * kThrTabLen);
  A.has_pre = pre_on ? 1u : 0u;
  build_layer_diag<T>(A, A.seg[so + 1], nloc, rb, pre_n, 1.0, A.pre, 0, &tgrp, &thr);
  A.has_post = 1u;
  build_layer_diag<T>(A, A.seg[so], t.vloc, rb, s->tail_post, s->tail_cst, A.post, 0, &tgrp, &thr);
  tgrp.insert(tgrp.end(), thr.begin(), thr.end());  // thread parts at kThrTabOff
  // the crossing: images Q u of segment so's bits, per-tile rep | delta << 48
  const TileSeg &sgo = A.seg[so];
  for (int k = 0; k < kTileThreadBits; ++k)
    A.rl_thr[k] = (k < nlocal - rb) ? static_cast<uint16_t>(to_local(qc[sgo.tl_bit[k]])) : 0;
  for (int j = 0; j < rb; ++j) A.rl_reg[j] = static_cast<uint16_t>(to_local(qc[sgo.reg0 + j]));
  auto packed = [&](uint64_t v) {
    uint64_t rest = 0;
    const uint32_t u = to_local(cross_coords(t, v, &rest));
    return rest | (static_cast<uint64_t>(u) << 48);
  };
  FusedRead &F = A.fr;
  F.pc = packed(img_c);
  // Tile-number bits: one per new-frame index bit h outside the new tile
  // (neither a head bit nor an extra's pivot), ascending; its old-frame base
  // vector is the source of h with the tile's own bits cleared, so that
  // consecutive tiles differ in nearby bits in both frames.
  uint64_t pivm = 0;
  for (int pv : t.piv) pivm |= 1ull << pv;
  std::vector<uint64_t> tbase, timg;
  for (int h = 0; h < n; ++h) {
    if (((t.hmask | pivm) >> h) & 1ull) continue;
    const uint64_t o = col[h] & ~t.vmask;
    uint64_t y = 0;
    for (int g = 0; g < n; ++g)
      if ((o >> g) & 1ull) y ^= img[g];
    tbase.push_back(o);
    timg.push_back(packed(y));
  }
  if (static_cast<int>(tbase.size()) != n - nlocal || tbase.size() > static_cast<size_t>(kMaxTileBits))
    return fail(MSV_E_INVALID, "internal: cross tile numbering");
  {
    // the bases must reach every coset of the tile: full rank on the bits outside it
    std::vector<uint64_t> m(tbase);
    int rank = 0;
    for (int g = 0; g < n; ++g) {
      if ((t.vmask >> g) & 1ull) continue;
      int piv = -1;
      for (size_t r = rank; r < m.size(); ++r)
        if ((m[r] >> g) & 1ull) {
          piv = static_cast<int>(r);
          break;
        }
      if (piv < 0) continue;
      std::swap(m[rank], m[piv]);
      for (size_t r = 0; r < m.size(); ++r)
        if (static_cast<int>(r) != rank && ((m[r] >> g) & 1ull)) m[r] ^= m[rank];
      ++rank;
    }
    if (rank != n - nlocal) return fail(MSV_E_INVALID, "internal: cross tile numbering");
  }
  // store images: the new-frame index vector of each local position
  const TileSeg &sgl = A.seg[wins.size() - 1];
  for (int k = 0; k < kTileThreadBits; ++k) F.pthr[k] = (k < nlocal - rb) ? bvec[sgl.tl_bit[k]] : 0ull;
  for (int j = 0; j < rb; ++j) F.preg[j] = bvec[sgl.reg0 + j];
  // the deferred diagonal at segment so (old frame): windows that see its
  // register bits first
  uint64_t umask = 0;
  for (int j = 0; j < rb; ++j) umask |= 1ull << t.vloc[sgo.reg0 + j];
  std::vector<V> lut;
  const bool rewin = s->pend_on && crossing_windows<T>(s, t.vloc, sgo, rb, umask, &F, &lut);
  if (rewin) {
    A.has_post = 0u;  // folded into the windows
  } else if (s->pend_on) {
    if (s->pend_lut.size() > kFusedLutMaxBytes || s->pend_args.nwin > static_cast<uint32_t>(kFusedWins))
      return fail(MSV_E_INVALID, "internal: deferred diagonal too wide for a cross sweep");
    F.ph = s->pend_args;
    std::vector<PhaseWin> dep, rest;
    for (uint32_t w = 0; w < F.ph.nwin; ++w) {
      const PhaseWin &pw = s->pend_args.win[w];
      const uint64_t wm = (((1ull << pw.w1) - 1ull) << pw.s1) | (((1ull << pw.w2) - 1ull) << pw.s2);
      (wm & umask ? dep : rest).push_back(pw);
    }
    if (dep.size() > static_cast<size_t>(kFusedRdep))
      return fail(MSV_E_INVALID, "internal: deferred diagonal too wide for a cross sweep");
    F.ph_rdep = static_cast<uint32_t>(dep.size());
    uint32_t w = 0;
    for (const PhaseWin &pw : dep) F.ph.win[w++] = pw;
    for (const PhaseWin &pw : rest) F.ph.win[w++] = pw;
    for (size_t k = 0; k < dep.size(); ++k) {
      const PhaseWin &pw = dep[k];
      for (int r = 0; r < (1 << rb); ++r) {
        uint64_t im = 0;
        for (int j = 0; j < rb; ++j)
          if ((r >> j) & 1) im |= 1ull << t.vloc[sgo.reg0 + j];
        const uint32_t lo = static_cast<uint32_t>(im >> pw.s1) & ((1u << pw.w1) - 1u);
        const uint32_t hi = static_cast<uint32_t>(im >> pw.s2) & ((1u << pw.w2) - 1u);
        F.ph_rimg[k][r] = static_cast<uint16_t>(lo | (hi << pw.w1));
      }
    }
  } else {
    F.ph.nwin = 0;
    F.ph.lut_len = 0;
  }
  // The tail's post diagonal (its bits' factors and the constant) folds into
  // the deferred diagonal's tables when a window holds each of those bits:
  // the diagonals multiply at the same point, and the sweep then skips one
  // diagonal pass over its registers.
  if (s->pend_on && !rewin) {
    lut.resize(s->pend_lut.size() / sizeof(V));
    std::memcpy(lut.data(), s->pend_lut.data(), lut.size() * sizeof(V));
    std::vector<std::pair<uint32_t, int>> at;  // (window, bit) per folded factor
    bool all = true;
    for (int b : s->tail_bits) {
      if (s->tail_post[b] == C(1.0)) continue;
      int wf = -1;
      for (uint32_t w = 0; w < F.ph.nwin && wf < 0; ++w) {
        const PhaseWin &pw = F.ph.win[w];
        if ((b >= pw.s1 && b < pw.s1 + pw.w1) || (b >= pw.s2 && b < pw.s2 + pw.w2)) wf = static_cast<int>(w);
      }
      if (wf < 0) {
        all = false;
        break;
      }
      at.push_back({static_cast<uint32_t>(wf), b});
    }
    if (all && F.ph.nwin > 0) {
      auto scale = [&](const PhaseWin &pw, uint32_t idx, C f) {
        V &e = lut[pw.off + idx];
        const C v = C(e.x, e.y) * f;
        e.x = static_cast<T>(v.real());
        e.y = static_cast<T>(v.imag());
      };
      const uint32_t len0 = 1u << (F.ph.win[0].w1 + F.ph.win[0].w2);
      for (uint32_t i = 0; i < len0; ++i) scale(F.ph.win[0], i, s->tail_cst);
      for (auto &wb : at) {
        const PhaseWin &pw = F.ph.win[wb.first];
        const int b = wb.second;
        const int pos = (b >= pw.s1 && b < pw.s1 + pw.w1) ? b - pw.s1 : pw.w1 + (b - pw.s2);
        const uint32_t len = 1u << (pw.w1 + pw.w2);
        for (uint32_t i = 0; i < len; ++i)
          if ((i >> pos) & 1u) scale(pw, i, s->tail_post[b]);
      }
      A.has_post = 0u;
    }
  }
  if (!launch) return MSV_OK;
  {
    void *dt = nullptr;
    MSV_TRY(stage(s, tgrp.data(), tgrp.size() * sizeof(V), &dt));
    A.tile_tab = reinterpret_cast<const V *>(dt);
  }
  std::vector<uint64_t> ptab(8 * 256, 0);
  for (int b = 0; b < 4; ++b)
    for (int v = 0; v < 256; ++v) {
      uint64_t x = 0, o = 0;
      for (int i = 0; i < 8; ++i) {
        const size_t k = 8 * static_cast<size_t>(b) + i;
        if (((v >> i) & 1) && k < timg.size()) {
          x ^= timg[k];
          o ^= tbase[k];
        }
      }
      ptab[256 * b + v] = x;
      ptab[1024 + 256 * b + v] = o;
    }
  void *dp = nullptr;
  MSV_TRY(stage(s, ptab.data(), ptab.size() * sizeof(uint64_t), &dp));
  const V *d_lut = nullptr;
  if (s->pend_on) {
    pad_fused_windows(F.ph, lut);
    void *dl = nullptr;
    MSV_TRY(stage(s, lut.data(), lut.size() * sizeof(V), &dl));
    d_lut = reinterpret_cast<const V *>(dl);
  }
  const bool oop = s->perm_on;
  if (oop) MSV_TRY(ensure_alt(s));
  const uint64_t fcap = static_cast<uint64_t>(s->cu) * tune().fused_blocks_per_cu;
  const uint32_t fgrid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(A.n_tiles, fcap)));
  MSV_LAUNCH(s, MSV_K_ROT_RELABEL, launch_rot_relabel<T>(reinterpret_cast<V *>(oop ? s->alt : s->buf), A, rb, fgrid, s->stream,
                                reinterpret_cast<const V *>(s->buf), d_lut,
                                reinterpret_cast<const uint64_t *>(dp)));
  if (oop) std::swap(s->buf, s->alt);
  s->perm_on = false;
  s->pend_on = false;
  s->tail_on = false;
  s->sweeps += 1;
  return MSV_OK;
}

// Gate-bit sets of a 1q layer's sweeps without a crossing: the low sweep
// (tile = bits [0, L)), then the high sweeps (columns + a chunk).
template <typename T>
std::vector<std::vector<int>> layer_gate_sets(int n, const std::vector<bool> &act) {
  int L = 0;
  std::vector<std::pair<int, int>> chunks;
  partition<T>(n, &L, &chunks);
  std::vector<std::vector<int>> sets;
  std::vector<int> low;
  for (int b = 0; b < L; ++b)
    if (act[b]) low.push_back(b);
  if (!low.empty()) sets.push_back(low);
  for (const std::vector<int> &g : high_gate_sets<T>(chunks)) {
    std::vector<int> a;
    for (int b : g)
      if (act[b]) a.push_back(b);
    if (!a.empty()) sets.push_back(a);
  }
  return sets;
}

// Splits the bits `rest` (ascending) left after a crossing sweep into a tail
// (up to G bits, consecutive in `rest`, chosen so that the next crossing
// through a permutation like `pred` gets the most head bits) and sweeps of
// at most G bits.
template <typename T>
void split_rest(int n, const std::vector<int> &rest, int G, bool defer, const std::vector<uint64_t> &pred,
                std::vector<std::vector<int>> *sets, std::vector<int> *tail) {
  sets->clear();
  tail->clear();
  std::vector<int> plain = rest;
  const int NL = max_local<T>();
  const int cb = std::max(0, std::min(col_bits<T>(), n));
  if (defer && rest.size() > 1) {
    const int k = std::min<int>(G, static_cast<int>(rest.size()) - 1);
    const std::vector<bool> all(n, true);
    int best = -1, best_sc = -1;
    for (int i = 0; i + k <= static_cast<int>(rest.size()); ++i) {
      const std::vector<int> cand(rest.begin() + i, rest.begin() + i + k);
      const uint64_t v = choose_tile_bits(n, NL, cb, cand, pred, all);
      const int sc = __builtin_popcountll(head_of(n, pred, v));
      if (sc >= best_sc) {
        best_sc = sc;
        best = i;
      }
    }
    tail->assign(rest.begin() + best, rest.begin() + best + k);
    plain.clear();
    for (int b : rest)
      if (std::find(tail->begin(), tail->end(), b) == tail->end()) plain.push_back(b);
  }
  const int m = static_cast<int>(plain.size());
  const int ns = (m + G - 1) / G;
  for (int c = 0, at = 0; c < ns; ++c) {
    const int sz = m / ns + (c < m % ns ? 1 : 0);
    sets->emplace_back(plain.begin() + at, plain.begin() + at + sz);
    at += sz;
  }
}

// A layer's post diagonal split into the part on the bits `tail` (post_t)
// and the rest (post_k).
void split_post(const std::vector<std::complex<double>> &post, const std::vector<int> &tail,
                std::vector<std::complex<double>> *post_k, std::vector<std::complex<double>> *post_t) {
  *post_k = post;
  post_t->assign(post.size(), 1.0);
  for (int b : tail) {
    (*post_t)[b] = post[b];
    (*post_k)[b] = 1.0;
  }
}

// Leaves the rotations on `tail` pending, with their post diagonal post_t
// and the layer's constant.
void keep_tail(msv_state *s, const std::vector<int> &tail, const std::vector<Shear> &sh,
               const std::vector<std::complex<double>> &post_t, std::complex<double> cst) {
  s->tail_on = true;
  s->tail_bits = tail;
  s->tail_sh = sh;
  s->tail_post = post_t;
  s->tail_cst = cst;
}

template <typename T>
int run_rot_layer(msv_state *s, const std::vector<bool> &act_bit, const std::vector<RotSplit> &zb) {
  using C = std::complex<double>;
  const int n = s->n;
  std::vector<Shear> sh(n, Shear{0.0, 0.0});
  std::vector<C> pre(n, 1.0), post(n, 1.0);
  C cst = 1.0;
  bool pre_on = false;
  for (int b = 0; b < n; ++b) {
    if (!act_bit[b]) continue;
    sh[b] = shear_of(zb[b]);
    pre[b] = zb[b].b2;
    post[b] = zb[b].b1 / zb[b].a1;
    cst *= zb[b].a1;
    pre_on |= std::abs(zb[b].b2 - 1.0) > 0.0;
  }
  const int NL = max_local<T>();
  const int cb = std::max(0, std::min(col_bits<T>(), n));
  const int G = std::max(rb_of<T>(), std::min(tune().gate_bits, NL - col_bits<T>()));
  const bool defer = tune().defer_tail != 0;
  std::vector<uint64_t> pred(n);
  for (int q = 0; q < n; ++q) pred[q] = s->perm_on ? s->perm_col[q] : (1ull << q);
  const std::vector<std::vector<int>> std_sets = layer_gate_sets<T>(n, act_bit);
  if (std_sets.empty()) return flush_pending_T<T>(s);
  // A pending tail: cross the deferred permutation in one sweep when that
  // launches fewer sweeps than running the tail on its own first. Without a
  // pending permutation (a lone diagonal or nothing), crossing measured
  // faster for states up to 64 GiB (brick layers at 30 and 32 qubits: 5%)
  // and slower for the 128 GiB complex64 state (34 qubits: 9%), so by
  // default it is taken there only up to 64 GiB.
  if (s->tail_on) {
    bool crossed = false;
    const uint64_t state_bytes = s->amps * static_cast<uint64_t>(sizeof(typename Vec2<T>::type));
    const bool try_cross = s->perm_on || tune().defer_tail == 2 || state_bytes <= (1ull << 36);
    const uint64_t vmask = try_cross ? choose_tile_bits(n, NL, cb, s->tail_bits, pred, act_bit) : 0;
    CrossTile t;
    std::vector<uint64_t> img;
    uint64_t img0 = 0, rot = 0;
    if (try_cross && __builtin_popcountll(vmask) == NL && perm_images(n, pred, 0, &img, &img0) &&
        cross_tile(n, vmask, pred, img, cb, &t) &&
        cross_sweep_T<T>(s, t, act_bit, sh, pre, pre_on, false, &rot) == MSV_OK) {
      std::vector<int> rest;  // the active bits the crossing does not rotate
      for (int b = 0; b < n; ++b)
        if (act_bit[b] && !((rot >> b) & 1ull)) rest.push_back(b);
      std::vector<std::vector<int>> sets;
      std::vector<int> tail;
      split_rest<T>(n, rest, G, defer, pred, &sets, &tail);
      const size_t cross_cost = 1 + sets.size();
      const size_t plain_cost = 1 + std_sets.size() - ((defer && std_sets.size() > 1) ? 1 : 0);
      if (!rest.empty() && !sets.empty() && cross_cost < plain_cost) {
        MSV_TRY(cross_sweep_T<T>(s, t, act_bit, sh, pre, pre_on, true));
        crossed = true;
        // the rest: the first sweep also applies the pre diagonal of every
        // remaining bit (tail included); the last one the post diagonal of
        // every bit but the tail's, which the tail keeps
        std::vector<C> pre_r(n, 1.0);
        bool pre_r_on = false;
        for (int b : rest) {
          pre_r[b] = pre[b];
          pre_r_on |= std::abs(pre[b] - 1.0) > 0.0;
        }
        std::vector<C> post_k, post_t;
        split_post(post, tail, &post_k, &post_t);
        for (size_t i = 0; i < sets.size(); ++i) {
          const bool last = i + 1 == sets.size();
          MSV_TRY(rot_sweep_T<T>(s, high_sweep_local_bits<T>(sets[i]), sets[i], sh,
                                 (i == 0 && pre_r_on) ? &pre_r : nullptr, last ? &post_k : nullptr,
                                 tail.empty() ? cst : C(1.0), false));
        }
        if (!tail.empty()) keep_tail(s, tail, sh, post_t, cst);
        return MSV_OK;
      }
    }
    if (!crossed) MSV_TRY(flush_tail_T<T>(s));
  }
  // sweep tiles: the low sweep's tile is bits [0, L), a high sweep's the
  // columns + its chunk
  int L = 0;
  std::vector<std::pair<int, int>> chunks;
  partition<T>(n, &L, &chunks);
  auto tile_of = [&](const std::vector<int> &set) {
    if (set.back() < L) {
      std::vector<int> local;
      for (int b = 0; b < L; ++b) local.push_back(b);
      return local;
    }
    return high_sweep_local_bits<T>(set);
  };
  // the last sweep's rotations become the pending tail when the layer has
  // several sweeps (and its tile is a high sweep's)
  size_t nrun = std_sets.size();
  const bool defer_last = defer && nrun > 1 && std_sets.back().front() >= L;
  if (defer_last) --nrun;
  // Deferred permutation and/or diagonal: the first sweep reads through
  // them (out of place) when can_fuse_read allows; otherwise they run first.
  bool fuse = false;
  if (s->perm_on || s->pend_on) {
    const std::vector<int> local0 = tile_of(std_sets[0]);
    const int nl0 = static_cast<int>(local0.size());
    const int rb0 = std::min(rb_of<T>(), nl0);
    uint32_t need = 0;
    for (int k = 0; k < nl0; ++k)
      if (std::find(std_sets[0].begin(), std_sets[0].end(), local0[k]) != std_sets[0].end()) need |= 1u << k;
    const std::vector<int> w0 =
        rot_segments(need, nl0, rb0, good_from(nl0, rb0), tune().lane_swap != 0).wins;
    fuse = !w0.empty() && can_fuse_read<T>(s, local0, w0[0], rb0);
  }
  if (!fuse) MSV_TRY(flush_pending_T<T>(s));
  std::vector<C> post_k, post_t;
  split_post(post, defer_last ? std_sets.back() : std::vector<int>(), &post_k, &post_t);
  for (size_t i = 0; i < nrun; ++i) {
    const bool last = i + 1 == nrun;
    MSV_TRY(rot_sweep_T<T>(s, tile_of(std_sets[i]), std_sets[i], sh, (i == 0 && pre_on) ? &pre : nullptr,
                           last ? &post_k : nullptr, defer_last ? C(1.0) : cst, i == 0 && fuse));
  }
  if (defer_last) keep_tail(s, std_sets.back(), sh, post_t, cst);
  return MSV_OK;
}

// A 1q layer on a deferred basis state |k>: the result is the product
// state with amplitude i = prod over index bits b of g_b(bit b of i), g_b(x)
// = U_q[x][bit b of k] for the gate U_q on the qubit at bit b (the identity
// for inactive qubits). The factors of each of three index-bit groups are
// multiplied out in double precision on the host and k_product writes the
// state in one pass.
template <typename T>
int product_layer_T(msv_state *s, const double *mats, const uint8_t *active) {
  using V = typename Vec2<T>::type;
  using C = std::complex<double>;
  const int n = s->n;
  std::vector<std::array<C, 2>> g(n);
  for (int q = 0; q < n; ++q) {
    const int b = bit_of(s, q);
    const int kb = static_cast<int>((s->basis_k >> b) & 1ull);
    for (int x = 0; x < 2; ++x) {
      if (!active || active[q]) {
        const double *m = mats + 8 * q + 2 * (2 * x + kb);
        g[b][x] = C(m[0], m[1]);
      } else {
        g[b][x] = (x == kb) ? C(1.0) : C(0.0);
      }
    }
  }
  const int l1 = std::min(kProductGroupBits, n);
  const int l2 = std::min(kProductGroupBits, n - l1);
  const int l3 = n - l1 - l2;
  std::vector<V> tabs;
  tabs.reserve((size_t(1) << l1) + (size_t(1) << l2) + (size_t(1) << l3));
  auto group = [&](int b0, int w) {
    for (uint64_t v = 0; v < (1ull << w); ++v) {
      C f = 1.0;
      for (int k = 0; k < w; ++k) f *= g[b0 + k][(v >> k) & 1ull];
      V e;
      e.x = static_cast<T>(f.real());
      e.y = static_cast<T>(f.imag());
      tabs.push_back(e);
    }
  };
  group(0, l1);
  group(l1, l2);
  group(l1 + l2, l3);
  void *d = nullptr;
  MSV_TRY(stage(s, tabs.data(), tabs.size() * sizeof(V), &d));
  const V *lo = reinterpret_cast<const V *>(d);
  const V *mid = lo + (size_t(1) << l1);
  const V *hi = mid + (size_t(1) << l2);
  MSV_LAUNCH(s, MSV_K_PRODUCT, launch_product<T>(reinterpret_cast<V *>(s->buf), s->amps, lo, mid, hi, static_cast<uint32_t>(l1),
                            static_cast<uint32_t>(l2), stream_grid(s), s->stream));
  s->basis_on = false;  // cleared once enqueued
  return MSV_OK;
}

template <typename T>
int apply_1q_layer_T(msv_state *s, const double *mats, const uint8_t *active) {
  if (s->basis_on && !s->pend_on && !s->perm_on) return product_layer_T<T>(s, mats, active);
  const int n = s->n;
  int L = 0;
  std::vector<std::pair<int, int>> chunks;
  partition<T>(n, &L, &chunks);
  std::vector<bool> act_bit(n, false);
  int n_act = 0;
  for (int q = 0; q < n; ++q)
    if (!active || active[q]) {
      act_bit[bit_of(s, q)] = true;
      ++n_act;
    }
  // rotation-split form when every active gate is a (scaled) unitary
  if (tune().rot_split_min > 0 && n_act >= tune().rot_split_min) {
    std::vector<RotSplit> zb(n);
    bool ok = true;
    for (int b = 0; ok && b < n; ++b)
      if (act_bit[b]) ok = split_rotation(mats + 8 * (n - 1 - b), &zb[b]);
    if (ok) return run_rot_layer<T>(s, act_bit, zb);
  }
  MSV_TRY(flush_pending_T<T>(s));
  // low sweep
  {
    std::vector<int> local = sweep_local_bits<T>(n, L, true, {0, 0});
    std::vector<int> gl;
    for (int k = 0; k < static_cast<int>(local.size()); ++k)
      if (act_bit[local[k]]) gl.push_back(k);
    MSV_TRY(run_1q_sweep<T>(s, local, gl, mats));
  }
  for (const std::vector<int> &gates : high_gate_sets<T>(chunks)) {
    std::vector<int> local = high_sweep_local_bits<T>(gates);
    std::vector<int> gl;
    for (int k = 0; k < static_cast<int>(local.size()); ++k)
      if (local[k] >= L && act_bit[local[k]]) gl.push_back(k);
    MSV_TRY(run_1q_sweep<T>(s, local, gl, mats));
  }
  return MSV_OK;
}

// A QFT sweep plan runs as k_qft when every segment with stages holds
// consecutive index bits at or above lb in its registers (then the stage
// twiddles split into a register part known at compile time and one thread
// part per stage); otherwise through the generic tile program.
template <typename T>
int launch_qft_plan(msv_state *s, const SweepPlan &plan, int lb, bool inverse) {
  using V = typename Vec2<T>::type;
  const int nlocal = static_cast<int>(plan.local.size());
  const int rb = std::min(rb_of<T>(), nlocal);
  QftArgs<T> A;
  std::memset(&A, 0, sizeof(A));
  std::vector<int> wins;
  bool ok = static_cast<int>(plan.segs.size()) <= kMaxSegs;
  for (size_t si = 0; ok && si < plan.segs.size(); ++si) {
    const SegPlan &sp = plan.segs[si];
    wins.push_back(sp.reg0);
    if (sp.ops.empty()) continue;
    const int p0 = plan.local[sp.reg0];
    for (int j = 0; j < rb; ++j) ok &= plan.local[sp.reg0 + j] == p0 + j;
    ok &= p0 >= lb;
    for (const PlanOp &po : sp.ops) {
      const int j = po.tbit - p0;
      ok &= (po.kind == OP_QFT_F || po.kind == OP_QFT_I) && j >= 0 && j < rb;
      if (ok) A.smask[si] |= 1u << j;
    }
    if (!ok) break;
    A.xbits[si] = static_cast<uint32_t>(p0 - lb);
    for (int j = 0; j < rb; ++j) A.wbits[si][j] = static_cast<uint32_t>(p0 + j - lb);
  }
  if (!ok) {
    MSV_TRY(flush_pending_T<T>(s));
    return launch_sweep<T>(s, plan);
  }
  MSV_TRY(fill_tile_geometry(s->n, plan.local, wins, rb, &A));
  A.lb = static_cast<uint32_t>(lb);
  A.scale = static_cast<T>(plan.scale);
  // deferred work: read through it in the first segment (k_qft's fused
  // read exists for 4 register bits), or run it first
  if ((s->perm_on || s->pend_on) &&
      !(rb == 4 && can_fuse_read<T>(s, plan.local, static_cast<int>(A.seg[0].reg0), rb)))
    MSV_TRY(flush_pending_T<T>(s));
  if (s->perm_on || s->pend_on) {
    const uint64_t *d_ptab = nullptr;
    const V *d_lut = nullptr;
    MSV_TRY(prepare_fused_read<T>(s, A, plan.local, rb, &d_ptab, &d_lut));
    const bool oop = s->perm_on;  // a lone diagonal runs in place
    if (oop) MSV_TRY(ensure_alt(s));
    const uint64_t fcap = static_cast<uint64_t>(s->cu) * tune().fused_blocks_per_cu;
    const uint32_t fgrid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(A.n_tiles, fcap)));
    MSV_LAUNCH(s, MSV_K_QFT_FUSED, launch_qft<T>(reinterpret_cast<V *>(oop ? s->alt : s->buf), A, rb, inverse, fgrid, s->stream,
                          reinterpret_cast<const V *>(s->buf), d_lut, d_ptab));
    if (oop) std::swap(s->buf, s->alt);
    s->perm_on = false;
    s->pend_on = false;
  } else {
    const uint64_t cap = static_cast<uint64_t>(s->cu) * tune().tile_blocks_per_cu;
    const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min(A.n_tiles, cap)));
    MSV_LAUNCH(s, MSV_K_QFT, launch_qft<T>(reinterpret_cast<V *>(s->buf), A, rb, inverse, grid, s->stream));
  }
  s->sweeps += 1;
  return MSV_OK;
}

// QFT (no terminal swaps) on index bits [lb, hb]: forward stages run from
// hb down to lb; the inverse runs the adjoint stages from lb up to hb.
template <typename T>
int apply_qft_T(msv_state *s, int hb, int lb, bool inverse, int only_stage_bit) {
  const int n = s->n;
  int L = 0;
  std::vector<std::pair<int, int>> chunks;
  partition<T>(n, &L, &chunks);
  struct Sw {
    std::vector<int> local;
    int gate_lo, gate_hi;  // index range of gate bits in this sweep
  };
  std::vector<Sw> sweeps;
  sweeps.push_back({sweep_local_bits<T>(n, L, true, {0, 0}), 0, L - 1});
  for (auto &ch : chunks)
    sweeps.push_back({sweep_local_bits<T>(n, L, false, ch), ch.first, ch.first + ch.second - 1});
  // forward: high sweeps first
  std::vector<int> order;
  for (int i = static_cast<int>(sweeps.size()) - 1; i >= 0; --i) order.push_back(i);
  if (inverse) std::reverse(order.begin(), order.end());
  for (int si : order) {
    const Sw &sw = sweeps[si];
    const int nlocal = static_cast<int>(sw.local.size());
    const int rb = std::min(rb_of<T>(), nlocal);
    std::vector<int> stage_bits;  // index bits, in execution order
    for (int p = std::min(hb, sw.gate_hi); p >= std::max(lb, sw.gate_lo); --p)
      if (only_stage_bit < 0 || p == only_stage_bit) stage_bits.push_back(p);
    if (stage_bits.empty()) continue;
    if (inverse) std::reverse(stage_bits.begin(), stage_bits.end());
    int loc_of[64];
    for (int i = 0; i < 64; ++i) loc_of[i] = -1;
    for (int k = 0; k < nlocal; ++k) loc_of[sw.local[k]] = k;
    std::vector<int> gl;
    for (int p : stage_bits) gl.push_back(loc_of[p]);
    // windows stay inside the gate bits so that register bits are
    // contiguous index bits
    const int floor = (si == 0) ? 0 : nlocal - (sw.gate_hi - sw.gate_lo + 1);
    std::vector<int> wins = cover_top_down(gl, nlocal, rb, floor);
    if (inverse) std::reverse(wins.begin(), wins.end());
    SweepPlan plan;
    plan.local = sw.local;
    std::vector<bool> used(stage_bits.size(), false);
    for (int w : wins) {
      SegPlan sp;
      sp.reg0 = w;
      for (size_t i = 0; i < stage_bits.size(); ++i) {
        const int l = loc_of[stage_bits[i]];
        if (!used[i] && l >= w && l < w + rb) {
          PlanOp po;
          po.kind = inverse ? OP_QFT_I : OP_QFT_F;
          po.tbit = stage_bits[i];
          po.lb = lb;
          sp.ops.push_back(po);
          used[i] = true;
        } else if (!used[i]) {
          break;  // keep stage order: later stages wait for a later window
        }
      }
      plan.segs.push_back(sp);
    }
    for (size_t i = 0; i < used.size(); ++i)
      if (!used[i]) return fail(MSV_E_INVALID, "internal: QFT stage left unscheduled");
    const int gf = good_from(nlocal, rb);
    if (plan.segs.front().reg0 < gf) {
      SegPlan ld;
      ld.reg0 = nlocal - rb;
      plan.segs.insert(plan.segs.begin(), ld);
    }
    if (plan.segs.back().reg0 < gf) {
      SegPlan st;
      st.reg0 = nlocal - rb;
      plan.segs.push_back(st);
    }
    plan.scale = std::pow(2.0, -0.5 * static_cast<double>(stage_bits.size()));
    MSV_TRY(launch_qft_plan<T>(s, plan, lb, inverse));
  }
  return MSV_OK;
}

// Generic ordered gate list -> tile programs. A program holds at most
// `max_targets` distinct non-diagonal target bits and kMaxTileOps ops.
template <typename T>
int flush_program(msv_state *s, const std::vector<PlanOp> &ops) {
  if (ops.empty()) return MSV_OK;
  const int n = s->n;
  const int NL = std::min(max_local<T>(), n);
  std::vector<int> targets;
  for (const PlanOp &po : ops)
    if (po.tbit >= 0 && std::find(targets.begin(), targets.end(), po.tbit) == targets.end())
      targets.push_back(po.tbit);
  std::vector<int> local = targets;
  for (int b = 0; b < n && static_cast<int>(local.size()) < NL; ++b)
    if (std::find(local.begin(), local.end(), b) == local.end()) local.push_back(b);
  std::sort(local.begin(), local.end());
  const int nlocal = static_cast<int>(local.size());
  const int rb = std::min(rb_of<T>(), nlocal);
  int loc_of[64];
  for (int i = 0; i < 64; ++i) loc_of[i] = -1;
  for (int k = 0; k < nlocal; ++k) loc_of[local[k]] = k;
  SweepPlan plan;
  plan.local = local;
  std::vector<PlanOp> pending_diag;
  for (size_t i = 0; i < ops.size(); ++i) {
    const PlanOp &po = ops[i];
    if (po.tbit < 0) {
      if (plan.segs.empty()) pending_diag.push_back(po);
      else plan.segs.back().ops.push_back(po);
      continue;
    }
    const int kt = loc_of[po.tbit];
    if (!plan.segs.empty()) {
      const int w = plan.segs.back().reg0;
      if (kt >= w && kt < w + rb) {
        plan.segs.back().ops.push_back(po);
        continue;
      }
    }
    int best = -1, best_score = -1;
    const int lo_min = std::max(0, kt - rb + 1), lo_max = std::min(kt, nlocal - rb);
    for (int lo = lo_max; lo >= lo_min; --lo) {
      int score = 0;
      for (size_t k = i; k < ops.size(); ++k) {
        if (ops[k].tbit < 0) continue;
        const int l = loc_of[ops[k].tbit];
        if (l >= lo && l < lo + rb) ++score;
        else break;
      }
      if (score > best_score) {
        best_score = score;
        best = lo;
      }
    }
    SegPlan sp;
    sp.reg0 = best;
    if (plan.segs.empty()) {
      sp.ops = pending_diag;
      pending_diag.clear();
    }
    sp.ops.push_back(po);
    plan.segs.push_back(sp);
  }
  if (plan.segs.empty()) {
    SegPlan sp;
    sp.reg0 = nlocal - rb;
    sp.ops = pending_diag;
    plan.segs.push_back(sp);
  }
  const int gf = good_from(nlocal, rb);
  if (plan.segs.front().reg0 < gf) {
    SegPlan ld;
    ld.reg0 = nlocal - rb;
    plan.segs.insert(plan.segs.begin(), ld);
  }
  if (plan.segs.back().reg0 < gf) {
    SegPlan st;
    st.reg0 = nlocal - rb;
    plan.segs.push_back(st);
  }
  return launch_sweep<T>(s, plan);
}

// number of segments flush_program would produce (without launching)
template <typename T>
int count_segments(const msv_state *s, const std::vector<PlanOp> &ops) {
  const int n = s->n;
  const int NL = std::min(max_local<T>(), n);
  std::vector<int> targets;
  for (const PlanOp &po : ops)
    if (po.tbit >= 0 && std::find(targets.begin(), targets.end(), po.tbit) == targets.end())
      targets.push_back(po.tbit);
  std::vector<int> local = targets;
  for (int b = 0; b < n && static_cast<int>(local.size()) < NL; ++b)
    if (std::find(local.begin(), local.end(), b) == local.end()) local.push_back(b);
  std::sort(local.begin(), local.end());
  const int nlocal = static_cast<int>(local.size());
  const int rb = std::min(rb_of<T>(), nlocal);
  int loc_of[64];
  for (int i = 0; i < 64; ++i) loc_of[i] = -1;
  for (int k = 0; k < nlocal; ++k) loc_of[local[k]] = k;
  int segs = 0, cur = -1, first = -1, last = -1;
  for (size_t i = 0; i < ops.size(); ++i) {
    if (ops[i].tbit < 0) continue;
    const int kt = loc_of[ops[i].tbit];
    if (cur >= 0 && kt >= cur && kt < cur + rb) continue;
    int best = -1, best_score = -1;
    const int lo_min = std::max(0, kt - rb + 1), lo_max = std::min(kt, nlocal - rb);
    for (int lo = lo_max; lo >= lo_min; --lo) {
      int score = 0;
      for (size_t k = i; k < ops.size(); ++k) {
        if (ops[k].tbit < 0) continue;
        const int l = loc_of[ops[k].tbit];
        if (l >= lo && l < lo + rb) ++score;
        else break;
      }
      if (score > best_score) {
        best_score = score;
        best = lo;
      }
    }
    cur = best;
    if (first < 0) first = best;
    last = best;
    ++segs;
  }
  if (segs == 0) return 1;
  const int gf = good_from(nlocal, rb);
  if (first < gf) ++segs;
  if (last < gf) ++segs;
  return segs;
}

int max_program_targets() { return 10; }

// Gate list -> tile programs, with commuting gates regrouped. A gate's
// non-diagonal action on its target bit commutes with every action of
// another gate that does not touch that bit; diagonal actions (controls,
// diagonal gates) commute with each other. The dependency graph keeps, per
// bit, the last gate acting non-diagonally on it and the diagonal users
// since. Programs are filled by list scheduling: of the gates whose
// predecessors have run, diagonal gates first, then gates on a target the
// program already has, then a new target while the program has fewer than
// max_program_targets(), up to kMaxOps gates. Within a program the gates
// are reordered (dependencies permitting) by target bit, so the register
// windows advance in one direction, adjacent gates with the same target
// and controls are multiplied into one, and the program is cut where the
// segment count would exceed kMaxSegs.
template <typename T>
int apply_plan_ops_reordered(msv_state *s, const std::vector<PlanOp> &all) {
  const int N = static_cast<int>(all.size());
  std::vector<std::vector<int>> succ(N);
  std::vector<int> npred(N, 0);
  std::vector<int> last_nd(64, -1);
  std::vector<std::vector<int>> diag_since(64);
  auto edge = [&](int from, int to) {
    succ[from].push_back(to);
    ++npred[to];
  };
  for (int i = 0; i < N; ++i) {
    const PlanOp &op = all[i];
    if (op.kind != OP_U2 && op.kind != OP_DIAG) return fail(MSV_E_INVALID, "internal: unexpected gate-list op");
    uint64_t dmask = op.kind == OP_U2 ? op.cmask : 0;
    if (op.kind == OP_DIAG) {
      dmask |= 1ull << op.abit;
      if (op.bbit >= 0) dmask |= 1ull << op.bbit;
    }
    if (op.tbit >= 0) {
      if (last_nd[op.tbit] >= 0) edge(last_nd[op.tbit], i);
      for (int j : diag_since[op.tbit]) edge(j, i);
      diag_since[op.tbit].clear();
    }
    for (uint64_t m = dmask; m; m &= m - 1) {
      const int d = __builtin_ctzll(m);
      if (last_nd[d] >= 0) edge(last_nd[d], i);
    }
    if (op.tbit >= 0) last_nd[op.tbit] = i;
    for (uint64_t m = dmask; m; m &= m - 1) diag_since[__builtin_ctzll(m)].push_back(i);
  }
  // ready gates in list order; a pick looks at the first kScan of them
  constexpr int kScan = 256;
  std::set<int> ready;
  for (int i = 0; i < N; ++i)
    if (npred[i] == 0) ready.insert(i);
  std::vector<int> in_prog(N, -1), lp(N, 0);
  int prog_id = 0;
  while (!ready.empty()) {
    // fill one program
    std::vector<int> prog;
    uint64_t tmask = 0;
    int nt = 0;
    while (static_cast<int>(prog.size()) < TileCfg<T>::kMaxOps) {
      int pick = -1, pick_rank = 3, seen = 0;
      for (auto it = ready.begin(); it != ready.end() && seen < kScan && pick_rank > 0; ++it, ++seen) {
        const PlanOp &op = all[*it];
        int rank;
        if (op.tbit < 0) rank = 0;
        else if ((tmask >> op.tbit) & 1ull) rank = 1;
        else if (nt < max_program_targets()) rank = 2;
        else continue;
        if (rank < pick_rank) {  // the set is in list order: the first of a rank wins
          pick = *it;
          pick_rank = rank;
        }
      }
      if (pick < 0) break;
      const int i = pick;
      ready.erase(i);
      prog.push_back(i);
      in_prog[i] = prog_id;
      if (all[i].tbit >= 0 && !((tmask >> all[i].tbit) & 1ull)) {
        tmask |= 1ull << all[i].tbit;
        ++nt;
      }
      for (int j : succ[i])
        if (--npred[j] == 0) ready.insert(j);
    }
    // order inside the program: by target bit (diagonal gates first),
    // dependencies within the program permitting
    for (int i : prog)
      for (int j : succ[i])
        if (in_prog[j] == prog_id) ++lp[j];
    std::vector<int> lready;
    for (int i : prog)
      if (lp[i] == 0) lready.push_back(i);
    std::vector<PlanOp> ordered;
    ordered.reserve(prog.size());
    while (!lready.empty()) {
      size_t best = 0;
      for (size_t r = 1; r < lready.size(); ++r) {
        const int a = all[lready[r]].tbit, b = all[lready[best]].tbit;
        if (a < b || (a == b && lready[r] < lready[best])) best = r;
      }
      const int i = lready[best];
      lready.erase(lready.begin() + best);
      ordered.push_back(all[i]);
      for (int j : succ[i])
        if (in_prog[j] == prog_id && --lp[j] == 0) lready.push_back(j);
    }
    if (ordered.size() != prog.size()) return fail(MSV_E_INVALID, "internal: gate reorder lost an op");
    // adjacent gates with the same target and controls (or diagonals on the
    // same bits) become one: their product, formed in double precision
    std::vector<PlanOp> merged;
    merged.reserve(ordered.size());
    for (const PlanOp &po : ordered) {
      if (!merged.empty()) {
        PlanOp &b = merged.back();
        if (po.kind == OP_U2 && b.kind == OP_U2 && po.tbit == b.tbit && po.cmask == b.cmask) {
          // po after b: M = M_po * M_b
          std::complex<double> x[4], y[4], z[4];
          for (int k = 0; k < 4; ++k) {
            x[k] = {po.m[2 * k], po.m[2 * k + 1]};
            y[k] = {b.m[2 * k], b.m[2 * k + 1]};
          }
          z[0] = x[0] * y[0] + x[1] * y[2];
          z[1] = x[0] * y[1] + x[1] * y[3];
          z[2] = x[2] * y[0] + x[3] * y[2];
          z[3] = x[2] * y[1] + x[3] * y[3];
          for (int k = 0; k < 4; ++k) {
            b.m[2 * k] = z[k].real();
            b.m[2 * k + 1] = z[k].imag();
          }
          const double id[8] = {1, 0, 0, 0, 0, 0, 1, 0};
          bool is_id = true;
          for (int k = 0; k < 8; ++k) is_id = is_id && b.m[k] == id[k];
          if (is_id) merged.pop_back();  // e.g. X X
          continue;
        }
        if (po.kind == OP_DIAG && b.kind == OP_DIAG && po.abit == b.abit && po.bbit == b.bbit) {
          const int ne = po.bbit >= 0 ? 4 : 2;
          for (int k = 0; k < ne; ++k) {
            const std::complex<double> e = std::complex<double>(po.m[2 * k], po.m[2 * k + 1]) *
                                           std::complex<double>(b.m[2 * k], b.m[2 * k + 1]);
            b.m[2 * k] = e.real();
            b.m[2 * k + 1] = e.imag();
          }
          continue;
        }
      }
      merged.push_back(po);
    }
    // cut where the segment count would exceed kMaxSegs
    std::vector<PlanOp> cur;
    for (const PlanOp &po : merged) {
      cur.push_back(po);
      if (cur.size() > 1 && count_segments<T>(s, cur) > kMaxSegs) {
        cur.pop_back();
        MSV_TRY(flush_program<T>(s, cur));
        cur.assign(1, po);
      }
    }
    MSV_TRY(flush_program<T>(s, cur));
    ++prog_id;
  }
  for (int i = 0; i < N; ++i)
    if (in_prog[i] < 0) return fail(MSV_E_INVALID, "internal: gate left unscheduled");
  return MSV_OK;
}

template <typename T>
int apply_plan_ops(msv_state *s, const std::vector<PlanOp> &all) {
  if (tune().reorder_gates) return apply_plan_ops_reordered<T>(s, all);
  std::vector<PlanOp> cur;
  std::vector<int> targets;
  for (const PlanOp &po : all) {
    std::vector<PlanOp> trial = cur;
    trial.push_back(po);
    std::vector<int> t2 = targets;
    if (po.tbit >= 0 && std::find(t2.begin(), t2.end(), po.tbit) == t2.end()) t2.push_back(po.tbit);
    const bool fits = static_cast<int>(trial.size()) <= TileCfg<T>::kMaxOps &&
                      static_cast<int>(t2.size()) <= max_program_targets() &&
                      count_segments<T>(s, trial) <= kMaxSegs;
    if (!fits) {
      MSV_TRY(flush_program<T>(s, cur));
      cur.clear();
      targets.clear();
      cur.push_back(po);
      if (po.tbit >= 0) targets.push_back(po.tbit);
    } else {
      cur = std::move(trial);
      targets = std::move(t2);
    }
  }
  return flush_program<T>(s, cur);
}

// ---------------------------------------------------------------------------
// diagonal phase layers
// ---------------------------------------------------------------------------

template <typename T>
int launch_phase_chunk(msv_state *s, const PhaseArgs &A, const void *lut_host, size_t lut_bytes) {
  using V = typename Vec2<T>::type;
  void *d_lut = nullptr;
  MSV_TRY(stage(s, lut_host, lut_bytes, &d_lut));
  const uint32_t grid = tune().phase_blocks > 0 ? static_cast<uint32_t>(tune().phase_blocks)
                                                : stream_grid(s);
  MSV_LAUNCH(s, MSV_K_PHASE, launch_phase<T>(reinterpret_cast<V *>(s->buf), reinterpret_cast<const V *>(d_lut), A, grid,
                          s->stream));
  s->sweeps += 1;
  return MSV_OK;
}

template <typename T>
int flush_pending_T(msv_state *s) {
  // Each flag is cleared only once its work is enqueued: a failed launch or
  // staging allocation leaves the work pending, so the error is not lost.
  if (s->basis_on) {
    MSV_TRY(init_basis_T<T>(s, s->basis_k));
    s->basis_on = false;
  }
  MSV_TRY(flush_tail_T<T>(s));  // the tail acts before the diagonal and the permutation
  if (s->perm_on) {
    MSV_TRY(launch_gather<T>(s, s->perm_col, s->perm_c));  // applies a pending diagonal too
    s->perm_on = false;
    return MSV_OK;
  }
  if (!s->pend_on) return MSV_OK;
  MSV_TRY(launch_phase_chunk<T>(s, s->pend_args, s->pend_lut.data(), s->pend_lut.size()));
  s->pend_on = false;
  return MSV_OK;
}

// Largest deferred lookup table: the absorbing gather keeps it in LDS next
// to its 16 KiB of permutation tables.
constexpr size_t kPendLutMaxBytes = 48 * 1024;

template <typename T>
int apply_diag_terms_T(msv_state *s, const std::vector<DiagTerm> &terms, bool defer) {
  using V = typename Vec2<T>::type;
  using C = std::complex<double>;
  // a pending tail stays when this diagonal is deferred behind it
  if (s->pend_on || s->perm_on || s->basis_on) MSV_TRY(flush_pending_T<T>(s));
  if (terms.empty()) return MSV_OK;
  std::vector<Win> wins;
  place_terms(terms, std::vector<int>(), 0, &wins);
  // order: windows reading index bit 0 first (evaluated per amplitude)
  std::stable_sort(wins.begin(), wins.end(), [](const Win &a, const Win &b) {
    const bool la = a.bits.front() == 0, lb2 = b.bits.front() == 0;
    return la && !lb2;
  });
  std::vector<std::pair<PhaseArgs, std::vector<V>>> chunks;
  size_t wi = 0;
  while (wi < wins.size()) {
    PhaseArgs A;
    std::vector<V> lut;
    wi = window_tables<T>(s, terms, wins, wi, C(1.0), &A, &lut);
    chunks.emplace_back(A, std::move(lut));
  }
  if (defer && tune().defer_diag && chunks.size() == 1 &&
      chunks[0].second.size() * sizeof(V) <= kPendLutMaxBytes) {
    s->pend_on = true;
    s->pend_terms = terms;
    s->pend_args = chunks[0].first;
    const auto *b = reinterpret_cast<const unsigned char *>(chunks[0].second.data());
    s->pend_lut.assign(b, b + chunks[0].second.size() * sizeof(V));
    return MSV_OK;
  }
  MSV_TRY(flush_tail_T<T>(s));
  for (auto &ch : chunks)
    MSV_TRY(launch_phase_chunk<T>(s, ch.first, ch.second.data(), ch.second.size() * sizeof(V)));
  return MSV_OK;
}

// ---------------------------------------------------------------------------
// XOR permutation
// ---------------------------------------------------------------------------

template <typename T>
int apply_xor_T(msv_state *s, const uint64_t *rows, uint64_t c) {
  using V = typename Vec2<T>::type;
  const int n = s->n;
  bool ident = (c == 0);
  for (int k = 0; k < n && ident; ++k) ident = rows[k] == (1ull << (n - 1 - k));
  if (ident) return MSV_OK;
  const uint64_t all = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
  for (int k = 0; k < n; ++k)
    if (rows[k] & ~all) return fail(MSV_E_INVALID, "xor_affine: row mask exceeds the register");
  if (c & ~all) return fail(MSV_E_INVALID, "xor_affine: flip vector exceeds the register");
  // column j of the inverse map: source bits set by index bit j of y
  std::vector<uint64_t> col(n, 0);
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      if ((rows[k] >> j) & 1ull) col[j] |= 1ull << (n - 1 - k);
  // invertibility check (a non-invertible M' would not be a permutation)
  {
    std::vector<uint64_t> m(col);
    int rank = 0;
    for (int bit = 0; bit < n && rank < n; ++bit) {
      int piv = -1;
      for (int r = rank; r < n; ++r)
        if ((m[r] >> bit) & 1ull) {
          piv = r;
          break;
        }
      if (piv < 0) continue;
      std::swap(m[rank], m[piv]);
      for (int r = 0; r < n; ++r)
        if (r != rank && ((m[r] >> bit) & 1ull)) m[r] ^= m[rank];
      ++rank;
    }
    if (rank != n) return fail(MSV_E_INVALID, "xor_affine: matrix is singular over GF(2)");
  }
  if (tune().defer_perm) {
    // A pending permutation is run first (two maps are not composed); a
    // pending diagonal stays and now acts at the source index; a pending
    // basis state is written first.
    if (s->basis_on) MSV_TRY(flush_pending_T<T>(s));
    if (s->perm_on) {
      MSV_TRY(flush_tail_T<T>(s));  // a pending tail acts before the pending permutation
      MSV_TRY(launch_gather<T>(s, s->perm_col, s->perm_c));
      s->perm_on = false;
    }
    MSV_TRY(ensure_alt(s));  // reserve the second buffer now: no allocation later
    s->perm_on = true;
    s->perm_col = col;
    s->perm_c = c;
    return MSV_OK;
  }
  MSV_TRY(flush_pending_T<T>(s));
  return launch_gather<T>(s, col, c);
}

// out[y] = D(src(y)) in[src(y)], src(y) = c ^ XOR of col[j] over the set bits
// j of y (D only when a diagonal is pending), into the second buffer.
template <typename T>
int launch_gather(msv_state *s, const std::vector<uint64_t> &col, uint64_t c) {
  using V = typename Vec2<T>::type;
  const int n = s->n;
  const int ntab = (n + 7) / 8;
  std::vector<uint64_t> tab(static_cast<size_t>(ntab) * 256, 0);
  for (int t = 0; t < ntab; ++t)
    for (int v = 0; v < 256; ++v) {
      uint64_t x = 0;
      for (int b = 0; b < 8; ++b) {
        const int j = 8 * t + b;
        if (j < n && ((v >> b) & 1)) x ^= col[j];
      }
      tab[static_cast<size_t>(t) * 256 + v] = x;
    }
  MSV_TRY(ensure_alt(s));
  void *d_tab = nullptr;
  MSV_TRY(stage(s, tab.data(), tab.size() * sizeof(uint64_t), &d_tab));
  if (s->pend_on) {
    // the deferred phase layer rides along: out[y] = D(src(y)) in[src(y)]
    void *d_lut = nullptr;
    MSV_TRY(stage(s, s->pend_lut.data(), s->pend_lut.size(), &d_lut));
    MSV_LAUNCH(s, MSV_K_XOR_GATHER_PHASE, launch_xor_gather_phase<T>(reinterpret_cast<const V *>(s->buf), reinterpret_cast<V *>(s->alt),
                                       s->amps, reinterpret_cast<const uint64_t *>(d_tab), ntab, c,
                                       reinterpret_cast<const V *>(d_lut), s->pend_args,
                                       stream_grid(s), s->stream));
    s->pend_on = false;  // cleared once enqueued
  } else {
    MSV_LAUNCH(s, MSV_K_XOR_GATHER, launch_xor_gather<T>(reinterpret_cast<const V *>(s->buf), reinterpret_cast<V *>(s->alt),
                                 s->amps, reinterpret_cast<const uint64_t *>(d_tab), ntab, c,
                                 stream_grid(s), s->stream));
  }
  std::swap(s->buf, s->alt);
  s->sweeps += 1;
  return MSV_OK;
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------

template <typename T>
int block_sums_T(msv_state *s, std::vector<double> *out) {
  using V = typename Vec2<T>::type;
  const uint64_t nb = (s->amps + kSampleBlock - 1) / kSampleBlock;
  MSV_TRY(ensure_red(s, nb * sizeof(double)));
  MSV_LAUNCH(s, MSV_K_BLOCK_PROBSUM, launch_block_probsum<T>(reinterpret_cast<const V *>(s->buf), s->amps,
                                  reinterpret_cast<double *>(s->d_red), nb, s->stream));
  out->resize(nb);
  MSV_HIP(hipMemcpyAsync(out->data(), s->d_red, nb * sizeof(double), hipMemcpyDeviceToHost,
                         s->stream));
  MSV_HIP(hipStreamSynchronize(s->stream));
  return MSV_OK;
}

template <typename T>
int sample_T(msv_state *s, const int32_t *wires, int nw, uint64_t shots, uint64_t seed,
             uint64_t stream_id, uint64_t *out) {
  using V = typename Vec2<T>::type;
  if (shots == 0) return MSV_OK;
  std::vector<double> bs;
  MSV_TRY(block_sums_T<T>(s, &bs));
  std::vector<double> prefix(bs.size() + 1, 0.0);
  for (size_t b = 0; b < bs.size(); ++b) prefix[b + 1] = prefix[b] + bs[b];
  const double total = prefix.back();
  if (!(total > 0.0)) return fail(MSV_E_INVALID, "sample: state has zero norm");
  std::vector<uint64_t> shot_block(shots);
  std::vector<double> shot_resid(shots);
  for (uint64_t k = 0; k < shots; ++k) {
    double t = philox_u01(seed, stream_id, k) * total;
    if (t >= total) t = std::nextafter(total, 0.0);
    // first block b with prefix[b+1] > t
    const size_t b = static_cast<size_t>(
        std::upper_bound(prefix.begin() + 1, prefix.end(), t) - (prefix.begin() + 1));
    const size_t bb = std::min(b, bs.size() - 1);
    shot_block[k] = bb;
    shot_resid[k] = t - prefix[bb];
  }
  const size_t need = shots * (sizeof(uint64_t) * 2 + sizeof(double));
  MSV_TRY(ensure_red(s, need));
  char *base = reinterpret_cast<char *>(s->d_red);
  uint64_t *d_block = reinterpret_cast<uint64_t *>(base);
  double *d_resid = reinterpret_cast<double *>(base + shots * sizeof(uint64_t));
  uint64_t *d_index = reinterpret_cast<uint64_t *>(base + shots * (sizeof(uint64_t) + sizeof(double)));
  MSV_HIP(hipMemcpyAsync(d_block, shot_block.data(), shots * sizeof(uint64_t),
                         hipMemcpyHostToDevice, s->stream));
  MSV_HIP(hipMemcpyAsync(d_resid, shot_resid.data(), shots * sizeof(double),
                         hipMemcpyHostToDevice, s->stream));
  MSV_LAUNCH(s, MSV_K_SAMPLE_RESOLVE, launch_sample_resolve<T>(reinterpret_cast<const V *>(s->buf), s->amps, d_block, d_resid,
                                   d_index, shots, s->stream));
  std::vector<uint64_t> idx(shots);
  MSV_HIP(hipMemcpyAsync(idx.data(), d_index, shots * sizeof(uint64_t), hipMemcpyDeviceToHost,
                         s->stream));
  MSV_HIP(hipStreamSynchronize(s->stream));
  const int n = s->n;
  for (uint64_t k = 0; k < shots; ++k) {
    uint64_t v = 0;
    for (int w = 0; w < nw; ++w) {
      const int bit = n - 1 - wires[w];
      v = (v << 1) | ((idx[k] >> bit) & 1ull);
    }
    out[k] = v;
  }
  return MSV_OK;
}

template <typename T>
int probabilities_T(msv_state *s, const int32_t *wires, int nw, double *out) {
  using V = typename Vec2<T>::type;
  const int n = s->n;
  bool identity = (wires == nullptr);
  if (!identity && nw == n) {
    identity = true;
    for (int w = 0; w < n; ++w) identity = identity && wires[w] == w;
  }
  if (identity) {
    MSV_TRY(ensure_red(s, s->amps * sizeof(double)));
    MSV_LAUNCH(s, MSV_K_PROBS_FULL, launch_probs_full<T>(reinterpret_cast<const V *>(s->buf), s->amps,
                                 reinterpret_cast<double *>(s->d_red), elem_grid(s, s->amps),
                                 s->stream));
    MSV_HIP(hipMemcpyAsync(out, s->d_red, s->amps * sizeof(double), hipMemcpyDeviceToHost,
                           s->stream));
    MSV_HIP(hipStreamSynchronize(s->stream));
    return MSV_OK;
  }
  if (nw == 0) {
    double tot = 0.0;
    std::vector<double> bs;
    MSV_TRY(block_sums_T<T>(s, &bs));
    for (double v : bs) tot += v;
    out[0] = tot;
    return MSV_OK;
  }
  std::vector<uint64_t> obits(nw);
  for (int w = 0; w < nw; ++w) obits[w] = static_cast<uint64_t>(n - 1 - wires[w]);
  const uint64_t K = 1ull << nw;
  const uint64_t O = 1ull << (n - nw);
  uint64_t chunks = std::max<uint64_t>(1, (1ull << 18) / K);
  chunks = std::min(chunks, O);
  const uint64_t chunk_len = O / chunks;
  const size_t part_bytes = K * chunks * sizeof(double);
  MSV_TRY(ensure_red(s, part_bytes + 64 * sizeof(uint64_t)));
  double *d_part = reinterpret_cast<double *>(s->d_red);
  uint64_t *d_bits = reinterpret_cast<uint64_t *>(reinterpret_cast<char *>(s->d_red) + part_bytes);
  MSV_HIP(hipMemcpyAsync(d_bits, obits.data(), nw * sizeof(uint64_t), hipMemcpyHostToDevice,
                         s->stream));
  MSV_LAUNCH(s, MSV_K_MARGINAL, launch_marginal<T>(reinterpret_cast<const V *>(s->buf), n, d_bits, nw, d_part, chunks,
                             chunk_len, s->stream));
  std::vector<double> part(K * chunks);
  MSV_HIP(hipMemcpyAsync(part.data(), d_part, part_bytes, hipMemcpyDeviceToHost, s->stream));
  MSV_HIP(hipStreamSynchronize(s->stream));
  for (uint64_t k = 0; k < K; ++k) {
    double acc = 0.0;
    for (uint64_t c = 0; c < chunks; ++c) acc += part[c * K + k];
    out[k] = acc;
  }
  return MSV_OK;
}

template <typename T>
int pauli_T(msv_state *s, const uint64_t *xq, const uint64_t *zq, int nwords, double *ore,
            double *oim) {
  using V = typename Vec2<T>::type;
  const int n = s->n;
  const uint64_t all = (n >= 64) ? ~0ull : ((1ull << n) - 1ull);
  for (int w0 = 0; w0 < nwords; w0 += kMaxWords) {
    const int cnt = std::min(kMaxWords, nwords - w0);
    std::vector<uint64_t> xm(cnt), zm(cnt);
    for (int w = 0; w < cnt; ++w) {
      if ((xq[w0 + w] & ~all) || (zq[w0 + w] & ~all))
        return fail(MSV_E_INVALID, "expval_pauli: mask names a qubit outside the register");
      uint64_t xi = 0, zi = 0;
      for (int q = 0; q < n; ++q) {
        if ((xq[w0 + w] >> q) & 1ull) xi |= 1ull << (n - 1 - q);
        if ((zq[w0 + w] >> q) & 1ull) zi |= 1ull << (n - 1 - q);
      }
      xm[w] = xi;
      zm[w] = zi;
    }
    const uint32_t grid = std::min<uint32_t>(elem_grid(s, s->amps), static_cast<uint32_t>(s->cu * 4));
    const size_t part_bytes = static_cast<size_t>(grid) * cnt * 2 * sizeof(double);
    MSV_TRY(ensure_red(s, part_bytes + 2 * kMaxWords * sizeof(uint64_t)));
    double *d_part = reinterpret_cast<double *>(s->d_red);
    uint64_t *d_x = reinterpret_cast<uint64_t *>(reinterpret_cast<char *>(s->d_red) + part_bytes);
    uint64_t *d_z = d_x + kMaxWords;
    MSV_HIP(hipMemcpyAsync(d_x, xm.data(), cnt * sizeof(uint64_t), hipMemcpyHostToDevice, s->stream));
    MSV_HIP(hipMemcpyAsync(d_z, zm.data(), cnt * sizeof(uint64_t), hipMemcpyHostToDevice, s->stream));
    MSV_LAUNCH(s, MSV_K_PAULI, launch_pauli<T>(reinterpret_cast<const V *>(s->buf), s->amps, d_x, d_z, cnt, d_part,
                            grid, s->stream));
    std::vector<double> part(static_cast<size_t>(grid) * cnt * 2);
    MSV_HIP(hipMemcpyAsync(part.data(), d_part, part_bytes, hipMemcpyDeviceToHost, s->stream));
    MSV_HIP(hipStreamSynchronize(s->stream));
    for (int w = 0; w < cnt; ++w) {
      double sr = 0.0, si = 0.0;
      for (uint32_t b = 0; b < grid; ++b) {
        sr += part[(static_cast<size_t>(b) * cnt + w) * 2];
        si += part[(static_cast<size_t>(b) * cnt + w) * 2 + 1];
      }
      // multiply by i^{ny}
      const int ny = __builtin_popcountll(xq[w0 + w] & zq[w0 + w]) & 3;
      double rr = sr, ri = si;
      for (int k = 0; k < ny; ++k) {
        const double t = rr;
        rr = -ri;
        ri = t;
      }
      ore[w0 + w] = rr;
      oim[w0 + w] = ri;
    }
  }
  return MSV_OK;
}

template <typename Ta, typename Tb>
int inner_T(msv_state *a, msv_state *b, double *re, double *im) {
  using Va = typename Vec2<Ta>::type;
  using Vb = typename Vec2<Tb>::type;
  MSV_HIP(hipStreamSynchronize(b->stream));
  const uint32_t grid = std::min<uint32_t>(elem_grid(a, a->amps), static_cast<uint32_t>(a->cu * 4));
  MSV_TRY(ensure_red(a, static_cast<size_t>(grid) * 2 * sizeof(double)));
  MSV_LAUNCH(a, MSV_K_INNER, (launch_inner<Ta, Tb>(reinterpret_cast<const Va *>(a->buf),
                                reinterpret_cast<const Vb *>(b->buf), a->amps,
                                reinterpret_cast<double *>(a->d_red), grid, a->stream)));
  std::vector<double> part(static_cast<size_t>(grid) * 2);
  MSV_HIP(hipMemcpyAsync(part.data(), a->d_red, part.size() * sizeof(double),
                         hipMemcpyDeviceToHost, a->stream));
  MSV_HIP(hipStreamSynchronize(a->stream));
  double sr = 0.0, si = 0.0;
  for (uint32_t k = 0; k < grid; ++k) {
    sr += part[2 * k];
    si += part[2 * k + 1];
  }
  *re = sr;
  *im = si;
  return MSV_OK;
}

template <typename T>
int init_basis_T(msv_state *s, uint64_t index) {
  using V = typename Vec2<T>::type;
  MSV_LAUNCH(s, MSV_K_INIT_BASIS, launch_init_basis<T>(reinterpret_cast<V *>(s->buf), s->amps, index,
                               elem_grid(s, s->amps), s->stream));
  return MSV_OK;
}

// Device address of `scratch` on `device` (the current device), looked up
// once per device with hipGetSymbolAddress and kept: each GPU holds its own
// copy of the array at its own address.
std::mutex g_scratch_mu;
void *g_scratch_ptr[64];

int noalloc_scratch(int device, char **out) {
  if (device < 0 || device >= 64) return fail(MSV_E_INVALID, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      (void)hipGetLastError();
      return fail(MSV_E_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    }
    g_scratch_ptr[device] = p;
  }
  *out = static_cast<char *>(g_scratch_ptr[device]);
  return MSV_OK;
}

// msv_sample without host buffers or device allocations: the block sums and
// their prefix (device, same order as the host's), then shots in chunks that
// fit the scratch behind the prefix; each chunk's indices are copied straight
// into `out` and packed there. Same results as sample_T.
template <typename T>
int sample_noalloc_T(msv_state *s, const int32_t *wires, int nw, uint64_t shots, uint64_t seed,
                     uint64_t stream_id, uint64_t *out, char *scr) {
  using V = typename Vec2<T>::type;
  const uint64_t nb = (s->amps + kSampleBlock - 1) / kSampleBlock;
  const size_t prefix_bytes = ((nb + 1) * sizeof(double) + 255) & ~size_t(255);
  double *d_prefix = reinterpret_cast<double *>(scr);
  MSV_LAUNCH(s, MSV_K_BLOCK_PROBSUM, launch_block_probsum<T>(reinterpret_cast<const V *>(s->buf), s->amps, d_prefix + 1, nb,
                                  s->stream));
  MSV_LAUNCH(s, MSV_K_PREFIX, launch_prefix_seq(d_prefix, nb, s->stream));
  double total = 0.0;
  MSV_HIP(hipMemcpyAsync(&total, d_prefix + nb, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  MSV_HIP(hipStreamSynchronize(s->stream));
  if (!(total > 0.0)) return fail(MSV_E_INVALID, "sample: state has zero norm");
  const uint64_t per_shot = 2 * sizeof(uint64_t) + sizeof(double);
  const uint64_t chunk = (kNoAllocScratchBytes - prefix_bytes) / per_shot;
  uint64_t *d_block = reinterpret_cast<uint64_t *>(scr + prefix_bytes);
  double *d_resid = reinterpret_cast<double *>(d_block + chunk);
  uint64_t *d_index = reinterpret_cast<uint64_t *>(d_resid + chunk);
  const int n = s->n;
  for (uint64_t k0 = 0; k0 < shots; k0 += chunk) {
    const uint64_t cnt = std::min(chunk, shots - k0);
    MSV_LAUNCH(s, MSV_K_SAMPLE_SEARCH, launch_sample_search(d_prefix, nb, total, seed, stream_id, k0, cnt, d_block, d_resid,
                                 s->stream));
    MSV_LAUNCH(s, MSV_K_SAMPLE_RESOLVE, launch_sample_resolve<T>(reinterpret_cast<const V *>(s->buf), s->amps, d_block, d_resid,
                                     d_index, cnt, s->stream));
    MSV_HIP(hipMemcpyAsync(out + k0, d_index, cnt * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
    MSV_HIP(hipStreamSynchronize(s->stream));
    for (uint64_t k = k0; k < k0 + cnt; ++k) {
      const uint64_t idx = out[k];
      uint64_t v = 0;
      for (int w = 0; w < nw; ++w) {
        const int bit = n - 1 - (wires ? wires[w] : w);
        v = (v << 1) | ((idx >> bit) & 1ull);
      }
      out[k] = v;
    }
  }
  return MSV_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

#define MSV_DISPATCH(s, fn, ...) \
  ((s)->dtype == MSV_C128 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

// Launches a deferred phase layer (no-op when none is pending). Every entry
// point that reads or writes amplitudes calls it first, except the XOR
// permutation, which absorbs it, and the phase layers themselves.
int flush_pending(msv_state *s) {
  if (!s->pend_on && !s->perm_on && !s->basis_on && !s->tail_on) return MSV_OK;
  MSV_TRY(set_dev(s));
  return MSV_DISPATCH(s, flush_pending_T, s);
}

namespace {

using cd = std::complex<double>;
struct M2 {
  cd a, b, c, d;  // [[a, b], [c, d]]
};
M2 mul(const M2 &x, const M2 &y) {
  return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}
M2 dagger(const M2 &x) { return {std::conj(x.a), std::conj(x.c), std::conj(x.b), std::conj(x.d)}; }
M2 rx(double t) {
  const double c = std::cos(t / 2), s = std::sin(t / 2);
  return {c, cd(0, -s), cd(0, -s), c};
}
M2 ry(double t) {
  const double c = std::cos(t / 2), s = std::sin(t / 2);
  return {c, -s, s, c};
}
M2 rz(double t) { return {std::polar(1.0, -t / 2), 0.0, 0.0, std::polar(1.0, t / 2)}; }

msv_gate g_u(int t, const M2 &m, int c0 = -1, int c1 = -1) {
  msv_gate g;
  std::memset(&g, 0, sizeof(g));
  g.kind = MSV_G_U;
  g.q[0] = t;
  if (c0 >= 0) g.q[1 + g.ncontrols++] = c0;
  if (c1 >= 0) g.q[1 + g.ncontrols++] = c1;
  const cd e[4] = {m.a, m.b, m.c, m.d};
  for (int k = 0; k < 4; ++k) {
    g.m[2 * k] = e[k].real();
    g.m[2 * k + 1] = e[k].imag();
  }
  return g;
}
msv_gate g_d1(int q, cd d0, cd d1) {
  msv_gate g;
  std::memset(&g, 0, sizeof(g));
  g.kind = MSV_G_DIAG1;
  g.q[0] = q;
  g.m[0] = d0.real(); g.m[1] = d0.imag(); g.m[2] = d1.real(); g.m[3] = d1.imag();
  return g;
}
msv_gate g_d2(int a, int b, cd d0, cd d1, cd d2, cd d3) {
  msv_gate g;
  std::memset(&g, 0, sizeof(g));
  g.kind = MSV_G_DIAG2;
  g.q[0] = a;
  g.q[1] = b;
  const cd e[4] = {d0, d1, d2, d3};
  for (int k = 0; k < 4; ++k) {
    g.m[2 * k] = e[k].real();
    g.m[2 * k + 1] = e[k].imag();
  }
  return g;
}
msv_gate g_swap(int a, int b, int c = -1) {
  msv_gate g;
  std::memset(&g, 0, sizeof(g));
  g.kind = MSV_G_SWAP;
  g.q[0] = a;
  g.q[1] = b;
  if (c >= 0) {
    g.ncontrols = 1;
    g.q[2] = c;
  }
  return g;
}

// Device gates of one op (the routing of mettleq_amd.fusion.gates_for_op).
int circuit_op_gates(int code, const int32_t *w, const double *p, std::vector<msv_gate> *out) {
  const double r2 = std::sqrt(0.5);
  const M2 X{0.0, 1.0, 1.0, 0.0}, Y{0.0, cd(0, -1), cd(0, 1), 0.0}, H{r2, r2, r2, -r2};
  const M2 S{1.0, 0.0, 0.0, cd(0, 1)};
  switch (code) {
    case MSV_OP_I: return MSV_OK;
    case MSV_OP_X: out->push_back(g_u(w[0], X)); return MSV_OK;
    case MSV_OP_Y: out->push_back(g_u(w[0], Y)); return MSV_OK;
    case MSV_OP_H: out->push_back(g_u(w[0], H)); return MSV_OK;
    case MSV_OP_SX:
      out->push_back(g_u(w[0], {cd(0.5, 0.5), cd(0.5, -0.5), cd(0.5, -0.5), cd(0.5, 0.5)}));
      return MSV_OK;
    case MSV_OP_Z: out->push_back(g_d1(w[0], 1.0, -1.0)); return MSV_OK;
    case MSV_OP_S: out->push_back(g_d1(w[0], 1.0, cd(0, 1))); return MSV_OK;
    case MSV_OP_SDG: out->push_back(g_d1(w[0], 1.0, cd(0, -1))); return MSV_OK;
    case MSV_OP_T: out->push_back(g_d1(w[0], 1.0, std::polar(1.0, M_PI / 4))); return MSV_OK;
    case MSV_OP_TDG: out->push_back(g_d1(w[0], 1.0, std::polar(1.0, -M_PI / 4))); return MSV_OK;
    case MSV_OP_RX: out->push_back(g_u(w[0], rx(p[0]))); return MSV_OK;
    case MSV_OP_RY: out->push_back(g_u(w[0], ry(p[0]))); return MSV_OK;
    case MSV_OP_RZ: {
      const M2 m = rz(p[0]);
      out->push_back(g_d1(w[0], m.a, m.d));
      return MSV_OK;
    }
    case MSV_OP_U1: out->push_back(g_d1(w[0], 1.0, std::polar(1.0, p[0]))); return MSV_OK;
    case MSV_OP_U2: out->push_back(g_u(w[0], mul(mul(rz(p[0]), ry(M_PI / 2)), rz(p[1])))); return MSV_OK;
    case MSV_OP_U3: out->push_back(g_u(w[0], mul(mul(rz(p[1]), ry(p[0])), rz(p[2])))); return MSV_OK;
    case MSV_OP_CNOT: out->push_back(g_u(w[1], X, w[0])); return MSV_OK;
    case MSV_OP_CH: out->push_back(g_u(w[1], H, w[0])); return MSV_OK;
    case MSV_OP_CRX: out->push_back(g_u(w[1], rx(p[0]), w[0])); return MSV_OK;
    case MSV_OP_CRY: out->push_back(g_u(w[1], ry(p[0]), w[0])); return MSV_OK;
    case MSV_OP_CZ: out->push_back(g_d2(w[0], w[1], 1.0, 1.0, 1.0, -1.0)); return MSV_OK;
    case MSV_OP_CPHASE:
      out->push_back(g_d2(w[0], w[1], 1.0, 1.0, 1.0, cd(std::cos(p[0]), std::sin(p[0]))));
      return MSV_OK;
    case MSV_OP_CRZ: {
      const M2 m = rz(p[0]);
      out->push_back(g_d2(w[0], w[1], 1.0, 1.0, m.a, m.d));
      return MSV_OK;
    }
    case MSV_OP_ZZPHASE: {
      const cd em(std::cos(p[0]), -std::sin(p[0])), ep(std::cos(p[0]), std::sin(p[0]));
      out->push_back(g_d2(w[0], w[1], em, ep, ep, em));
      return MSV_OK;
    }
    case MSV_OP_XXPHASE:
    case MSV_OP_YYPHASE: {
      // exp(-i t PP) = (V x V) exp(-i t ZZ) (V^dag x V^dag), V Z V^dag = P
      const M2 v = code == MSV_OP_XXPHASE ? H : mul(S, H);
      const M2 vd = dagger(v);
      const cd em(std::cos(p[0]), -std::sin(p[0])), ep(std::cos(p[0]), std::sin(p[0]));
      out->push_back(g_u(w[0], vd));
      out->push_back(g_u(w[1], vd));
      out->push_back(g_d2(w[0], w[1], em, ep, ep, em));
      out->push_back(g_u(w[0], v));
      out->push_back(g_u(w[1], v));
      return MSV_OK;
    }
    case MSV_OP_SWAP: out->push_back(g_swap(w[0], w[1])); return MSV_OK;
    case MSV_OP_ISWAP:
      out->push_back(g_d2(w[0], w[1], 1.0, cd(0, 1), cd(0, 1), 1.0));
      out->push_back(g_swap(w[0], w[1]));
      return MSV_OK;
    case MSV_OP_TOFFOLI: out->push_back(g_u(w[2], X, w[0], w[1])); return MSV_OK;
    case MSV_OP_FREDKIN: out->push_back(g_swap(w[1], w[2], w[0])); return MSV_OK;
    default: return fail(MSV_E_INVALID, "circuit: unknown gate code " + std::to_string(code));
  }
}

int circuit_arity(int code) {
  if (code <= MSV_OP_U3) return 1;
  if (code <= MSV_OP_YYPHASE) return 2;
  return 3;
}

}  // namespace

extern "C" {

const char *msv_last_error(void) { return g_err.c_str(); }

int msv_version(void) { return 10000; }

int msv_device_count(int *out) {
  if (!out) return fail(MSV_E_INVALID, "null output");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *out = 0;
    (void)hipGetLastError();
    return fail(MSV_E_HIP, std::string("hipGetDeviceCount failed: ") + hipGetErrorString(e));
  }
  *out = c;
  return MSV_OK;
}

int msv_get_device_info(int device, msv_device_info *out) {
  if (!out) return fail(MSV_E_INVALID, "null output");
  std::memset(out, 0, sizeof(*out));
  MSV_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  MSV_HIP(hipGetDeviceProperties(&prop, device));
  size_t fr = 0, tot = 0;
  MSV_HIP(hipMemGetInfo(&fr, &tot));
  out->total_bytes = tot;
  out->free_bytes = fr;
  out->max_alloc_bytes = msv::device_max_alloc(device);
  out->compute_units = prop.multiProcessorCount;
  out->lds_bytes_per_block = static_cast<int32_t>(prop.sharedMemPerBlock);
  std::snprintf(out->name, sizeof(out->name), "%s", prop.name);
  std::snprintf(out->arch, sizeof(out->arch), "%s", prop.gcnArchName);
  return MSV_OK;
}

int msv_state_create(int n_qubits, int dtype, int device, msv_state **out) {
  TuneScope ts_;
  if (!out) return fail(MSV_E_INVALID, "null output handle");
  *out = nullptr;
  if (n_qubits < 1 || n_qubits > 40)
    return fail(MSV_E_INVALID, "n_qubits must be in [1, 40], got " + std::to_string(n_qubits));
  if (dtype != MSV_C64 && dtype != MSV_C128) return fail(MSV_E_INVALID, "unknown dtype");
  MSV_HIP(hipSetDevice(device));
  msv_state *s = new msv_state();
  s->n = n_qubits;
  s->dtype = dtype;
  s->device = device;
  s->cu = dev_cu(device);
  s->amps = 1ull << n_qubits;
  const size_t bytes = s->amps * esize(dtype);
  if (hipMalloc(&s->buf, bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete s;
    return fail(MSV_E_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes for a " +
                               std::to_string(n_qubits) + "-qubit state failed");
  }
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipFree(s->buf);
    delete s;
    return fail(MSV_E_HIP, "hipStreamCreate failed");
  }
  int rc = MSV_DISPATCH(s, init_basis_T, s, 0ull);
  if (rc == MSV_OK && hipStreamSynchronize(s->stream) != hipSuccess)
    rc = fail(MSV_E_HIP, "state initialization failed");
  if (rc != MSV_OK) {
    msv_state_destroy(s);
    return rc;
  }
  *out = s;
  return MSV_OK;
}

int msv_state_create_virtual(int n_qubits, int dtype, int device, msv_state **out) {
  TuneScope ts_;
  if (!out) return fail(MSV_E_INVALID, "null output handle");
  *out = nullptr;
  if (n_qubits < 1 || n_qubits > 40)
    return fail(MSV_E_INVALID, "n_qubits must be in [1, 40], got " + std::to_string(n_qubits));
  if (dtype != MSV_C64 && dtype != MSV_C128) return fail(MSV_E_INVALID, "unknown dtype");
  msv_state *s = new msv_state();
  s->dry = true;
  s->n = n_qubits;
  s->dtype = dtype;
  s->device = device;
  s->cu = 256;  // grid sizes only: the plan's sweeps do not depend on them
  s->amps = 1ull << n_qubits;
  s->buf = reinterpret_cast<void *>(uintptr_t(256));  // never dereferenced
  *out = s;
  return MSV_OK;
}

int msv_state_dispatch_log(msv_state *s, int32_t *kinds, int cap, uint64_t *count, int reset) {
  if (!s) return fail(MSV_E_INVALID, "null state handle");
  if (cap < 0 || (cap > 0 && !kinds)) return fail(MSV_E_INVALID, "bad output buffer");
  if (count) *count = s->dlog_total;
  const size_t m = std::min(s->dlog.size(), static_cast<size_t>(cap));
  for (size_t i = 0; i < m; ++i) kinds[i] = s->dlog[i];
  if (reset) {
    s->dlog.clear();
    s->dlog_total = 0;
  }
  return MSV_OK;
}

int msv_state_destroy(msv_state *s) {
  TuneScope ts_;
  if (!s) return MSV_OK;
  if (s->dry) {  // a virtual state owns no device resources
    delete s;
    return MSV_OK;
  }
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->buf) (void)hipFree(s->buf);
  if (s->alt) (void)hipFree(s->alt);
  if (s->d_stage) (void)hipFree(s->d_stage);
  if (s->h_stage) (void)hipHostFree(s->h_stage);
  if (s->d_red) (void)hipFree(s->d_red);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return MSV_OK;
}

int msv_state_reset(msv_state *s) { return msv_state_set_basis(s, 0); }

int msv_state_set_basis(msv_state *s, uint64_t index) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (index >= s->amps) return fail(MSV_E_INVALID, "basis index out of range");
  MSV_TRY(set_dev(s));
  s->pend_on = false;  // the new basis state replaces what deferred work would act on
  s->perm_on = false;
  s->tail_on = false;
  s->basis_on = true;  // written by the next operation (a 1q layer: as its product state)
  s->basis_k = index;
  return MSV_OK;
}

int msv_state_info(const msv_state *s, int *n_qubits, int *dtype, uint64_t *bytes) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (n_qubits) *n_qubits = s->n;
  if (dtype) *dtype = s->dtype;
  if (bytes) *bytes = s->amps * esize(s->dtype);
  return MSV_OK;
}

int msv_state_upload(msv_state *s, const void *host, uint64_t offset, uint64_t count) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (s->dry) return fail(MSV_E_UNSUPPORTED, "a virtual state holds no amplitudes");
  if (!host && count) return fail(MSV_E_INVALID, "null host buffer");
  if (offset > s->amps || count > s->amps - offset)
    return fail(MSV_E_INVALID, "upload range exceeds the state");
  MSV_TRY(set_dev(s));
  MSV_TRY(flush_pending(s));
  const size_t es = esize(s->dtype);
  MSV_HIP(hipMemcpyAsync(static_cast<char *>(s->buf) + offset * es, host, count * es,
                         hipMemcpyHostToDevice, s->stream));
  MSV_HIP(hipStreamSynchronize(s->stream));
  return MSV_OK;
}

int msv_state_download(msv_state *s, void *host, uint64_t offset, uint64_t count) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (s->dry) return fail(MSV_E_UNSUPPORTED, "a virtual state holds no amplitudes");
  if (!host && count) return fail(MSV_E_INVALID, "null host buffer");
  if (offset > s->amps || count > s->amps - offset)
    return fail(MSV_E_INVALID, "download range exceeds the state");
  MSV_TRY(set_dev(s));
  MSV_TRY(flush_pending(s));
  const size_t es = esize(s->dtype);
  MSV_HIP(hipMemcpyAsync(host, static_cast<const char *>(s->buf) + offset * es, count * es,
                         hipMemcpyDeviceToHost, s->stream));
  MSV_HIP(hipStreamSynchronize(s->stream));
  return MSV_OK;
}

int msv_state_copy(msv_state *dst, msv_state *src) {
  TuneScope ts_;
  MSV_TRY(check_state(dst));
  MSV_TRY(check_state(src));
  if (dst->dry || src->dry) return fail(MSV_E_UNSUPPORTED, "a virtual state holds no amplitudes");
  if (dst->n != src->n) return fail(MSV_E_INVALID, "state_copy: qubit counts differ");
  MSV_TRY(flush_pending(src));
  MSV_TRY(set_dev(dst));
  if (dst != src) dst->pend_on = dst->perm_on = dst->basis_on = dst->tail_on = false;  // overwritten
  MSV_HIP(hipStreamSynchronize(src->stream));
  const uint32_t grid = elem_grid(dst, dst->amps);
  if (src->dtype == MSV_C64 && dst->dtype == MSV_C64)
    MSV_LAUNCH(dst, MSV_K_CONVERT, (launch_convert<float, float>(static_cast<const float2 *>(src->buf),
                                          static_cast<float2 *>(dst->buf), dst->amps, grid,
                                          dst->stream)));
  else if (src->dtype == MSV_C64)
    MSV_LAUNCH(dst, MSV_K_CONVERT, (launch_convert<float, double>(static_cast<const float2 *>(src->buf),
                                           static_cast<double2 *>(dst->buf), dst->amps, grid,
                                           dst->stream)));
  else if (dst->dtype == MSV_C64)
    MSV_LAUNCH(dst, MSV_K_CONVERT, (launch_convert<double, float>(static_cast<const double2 *>(src->buf),
                                           static_cast<float2 *>(dst->buf), dst->amps, grid,
                                           dst->stream)));
  else
    MSV_LAUNCH(dst, MSV_K_CONVERT, (launch_convert<double, double>(static_cast<const double2 *>(src->buf),
                                            static_cast<double2 *>(dst->buf), dst->amps, grid,
                                            dst->stream)));
  MSV_HIP(hipStreamSynchronize(dst->stream));
  return MSV_OK;
}

int msv_state_scale(msv_state *s, double re, double im) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  MSV_TRY(set_dev(s));
  MSV_TRY(flush_pending(s));
  if (s->dtype == MSV_C128)
    MSV_LAUNCH(s, MSV_K_SCALE, launch_scale<double>(static_cast<double2 *>(s->buf), s->amps, re, im,
                                 stream_grid(s), s->stream));
  else
    MSV_LAUNCH(s, MSV_K_SCALE, launch_scale<float>(static_cast<float2 *>(s->buf), s->amps, re, im,
                                stream_grid(s), s->stream));
  s->sweeps += 1;
  return MSV_OK;
}

int msv_state_sync(msv_state *s) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  MSV_TRY(set_dev(s));
  MSV_TRY(flush_pending(s));
  if (s->dry) return MSV_OK;
  MSV_HIP(hipStreamSynchronize(s->stream));
  return MSV_OK;
}

int msv_state_release_scratch(msv_state *s) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  MSV_TRY(set_dev(s));
  MSV_TRY(flush_pending(s));  // a deferred permutation still needs the buffer
  if (s->dry) {
    s->alt = nullptr;
    return MSV_OK;
  }
  MSV_HIP(hipStreamSynchronize(s->stream));
  if (s->alt) MSV_HIP(hipFree(s->alt));
  s->alt = nullptr;
  return MSV_OK;
}

void *msv_state_device_ptr(msv_state *s) {
  TuneScope ts_;
  if (!s) return nullptr;
  // the caller reads the amplitudes directly: a deferred diagonal goes first
  if (flush_pending(s) != MSV_OK) return nullptr;
  return s->dry ? nullptr : s->buf;
}

int msv_apply_1q_layer(msv_state *s, const double *mats, const uint8_t *active) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (!mats) return fail(MSV_E_INVALID, "null matrix array");
  MSV_TRY(set_dev(s));
  return MSV_DISPATCH(s, apply_1q_layer_T, s, mats, active);
}

int msv_apply_phase_layer(msv_state *s, const int32_t *kinds, const int32_t *bonds,
                          const double *thetas, int n_terms) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (n_terms < 0 || (n_terms > 0 && (!kinds || !bonds || !thetas)))
    return fail(MSV_E_INVALID, "phase_layer: null term arrays");
  std::vector<DiagTerm> terms;
  for (int t = 0; t < n_terms; ++t) {
    const int a = bonds[2 * t], b = bonds[2 * t + 1];
    if (a == b || a < 0 || b < 0 || a >= s->n || b >= s->n)
      return fail(MSV_E_INVALID, "phase_layer: invalid bond (" + std::to_string(a) + ", " +
                                     std::to_string(b) + ")");
    DiagTerm d;
    d.bits[0] = bit_of(s, a);
    d.bits[1] = bit_of(s, b);
    const double th = thetas[t];
    if (kinds[t] == MSV_PHASE_ZZ) {
      const double c = std::cos(th), sn = std::sin(th);
      // exp(-i th Z_a Z_b): e^{-i th} when the bits agree, e^{+i th} otherwise
      d.re[0] = c; d.im[0] = -sn;
      d.re[1] = c; d.im[1] = sn;
      d.re[2] = c; d.im[2] = sn;
      d.re[3] = c; d.im[3] = -sn;
    } else if (kinds[t] == MSV_PHASE_CPHASE) {
      for (int k = 0; k < 3; ++k) {
        d.re[k] = 1.0;
        d.im[k] = 0.0;
      }
      if (std::fabs(th - M_PI) < 1e-12) {
        d.re[3] = -1.0;
        d.im[3] = 0.0;
      } else {
        d.re[3] = std::cos(th);
        d.im[3] = std::sin(th);
      }
    } else {
      return fail(MSV_E_INVALID, "phase_layer: unknown term kind");
    }
    terms.push_back(d);
  }
  MSV_TRY(set_dev(s));
  return MSV_DISPATCH(s, apply_diag_terms_T, s, terms, true);
}

int msv_apply_diagonal(msv_state *s, const int32_t *q1, const double *d1, int n1,
                       const int32_t *q2, const double *d2, int n2) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  std::vector<DiagTerm> terms;
  for (int t = 0; t < n1; ++t) {
    MSV_TRY(check_qubit(s, q1[t], "diagonal"));
    DiagTerm d;
    d.bits[0] = bit_of(s, q1[t]);
    d.bits[1] = -1;
    for (int k = 0; k < 2; ++k) {
      d.re[k] = d1[4 * t + 2 * k];
      d.im[k] = d1[4 * t + 2 * k + 1];
    }
    terms.push_back(d);
  }
  for (int t = 0; t < n2; ++t) {
    MSV_TRY(check_qubit(s, q2[2 * t], "diagonal"));
    MSV_TRY(check_qubit(s, q2[2 * t + 1], "diagonal"));
    if (q2[2 * t] == q2[2 * t + 1]) return fail(MSV_E_INVALID, "diagonal: duplicate qubits");
    DiagTerm d;
    d.bits[0] = bit_of(s, q2[2 * t]);
    d.bits[1] = bit_of(s, q2[2 * t + 1]);
    for (int k = 0; k < 4; ++k) {
      d.re[k] = d2[8 * t + 2 * k];
      d.im[k] = d2[8 * t + 2 * k + 1];
    }
    terms.push_back(d);
  }
  MSV_TRY(set_dev(s));
  return MSV_DISPATCH(s, apply_diag_terms_T, s, terms, true);
}

int msv_apply_xor_affine(msv_state *s, const uint64_t *rows, uint64_t c) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (!rows) return fail(MSV_E_INVALID, "null rows");
  MSV_TRY(set_dev(s));
  return MSV_DISPATCH(s, apply_xor_T, s, rows, c);
}

int msv_apply_qft(msv_state *s, int first, int m, int inverse) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (!(0 <= first && first < m && m <= s->n))
    return fail(MSV_E_INVALID, "qft: invalid qubit range [" + std::to_string(first) + ", " +
                                   std::to_string(m) + ") for " + std::to_string(s->n) +
                                   " qubits");
  MSV_TRY(set_dev(s));
  // deferred work is read by the first QFT sweep or flushed before it; a
  // deferred basis state is written first, a deferred rotation tail run first
  if (s->basis_on) MSV_TRY(flush_pending(s));
  if (s->tail_on) MSV_TRY(MSV_DISPATCH(s, flush_tail_T, s));
  const int hb = s->n - 1 - first, lb = s->n - m;
  return MSV_DISPATCH(s, apply_qft_T, s, hb, lb, inverse != 0, -1);
}

int msv_apply_qft_stage(msv_state *s, int j, int m, int inverse) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (!(0 <= j && j < m && m <= s->n))
    return fail(MSV_E_INVALID, "qft_stage: invalid j=" + std::to_string(j) + ", m=" +
                                   std::to_string(m) + ", n=" + std::to_string(s->n));
  MSV_TRY(set_dev(s));
  MSV_TRY(flush_pending(s));
  const int p = s->n - 1 - j, lb = s->n - m;
  return MSV_DISPATCH(s, apply_qft_T, s, p, lb, inverse != 0, p);
}

int msv_apply_gates(msv_state *s, const msv_gate *gates, int n_gates) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (n_gates < 0 || (n_gates > 0 && !gates)) return fail(MSV_E_INVALID, "null gate list");
  std::vector<PlanOp> ops;
  ops.reserve(n_gates);
  for (int g = 0; g < n_gates; ++g) {
    const msv_gate &G = gates[g];
    if (G.kind == MSV_G_U) {
      if (G.ncontrols < 0 || G.ncontrols > 5) return fail(MSV_E_INVALID, "gate: bad control count");
      MSV_TRY(check_qubit(s, G.q[0], "gate"));
      PlanOp po;
      po.kind = OP_U2;
      po.tbit = bit_of(s, G.q[0]);
      for (int c = 0; c < G.ncontrols; ++c) {
        MSV_TRY(check_qubit(s, G.q[1 + c], "gate control"));
        const int cb = bit_of(s, G.q[1 + c]);
        if (cb == po.tbit || ((po.cmask >> cb) & 1ull))
          return fail(MSV_E_INVALID, "gate: control and target wires must be distinct");
        po.cmask |= 1ull << cb;
      }
      for (int k = 0; k < 8; ++k) po.m[k] = G.m[k];
      ops.push_back(po);
    } else if (G.kind == MSV_G_DIAG1) {
      MSV_TRY(check_qubit(s, G.q[0], "diagonal gate"));
      PlanOp po;
      po.kind = OP_DIAG;
      po.abit = bit_of(s, G.q[0]);
      for (int k = 0; k < 4; ++k) po.m[k] = G.m[k];
      ops.push_back(po);
    } else if (G.kind == MSV_G_DIAG2) {
      MSV_TRY(check_qubit(s, G.q[0], "diagonal gate"));
      MSV_TRY(check_qubit(s, G.q[1], "diagonal gate"));
      if (G.q[0] == G.q[1]) return fail(MSV_E_INVALID, "diagonal gate: duplicate qubits");
      PlanOp po;
      po.kind = OP_DIAG;
      po.abit = bit_of(s, G.q[0]);
      po.bbit = bit_of(s, G.q[1]);
      for (int k = 0; k < 8; ++k) po.m[k] = G.m[k];
      ops.push_back(po);
    } else if (G.kind == MSV_G_SWAP) {
      if (G.ncontrols < 0 || G.ncontrols > 4) return fail(MSV_E_INVALID, "swap: bad control count");
      MSV_TRY(check_qubit(s, G.q[0], "swap"));
      MSV_TRY(check_qubit(s, G.q[1], "swap"));
      if (G.q[0] == G.q[1]) continue;
      const int a = bit_of(s, G.q[0]), b = bit_of(s, G.q[1]);
      uint64_t cm = 0;
      for (int c = 0; c < G.ncontrols; ++c) {
        MSV_TRY(check_qubit(s, G.q[2 + c], "swap control"));
        const int cb = bit_of(s, G.q[2 + c]);
        if (cb == a || cb == b) return fail(MSV_E_INVALID, "swap: control overlaps a target");
        cm |= 1ull << cb;
      }
      // SWAP = CX(a->b) CX(b->a) CX(a->b), each carrying the extra controls
      const int seq[3][2] = {{a, b}, {b, a}, {a, b}};
      for (auto &cx : seq) {
        PlanOp po;
        po.kind = OP_U2;
        po.tbit = cx[1];
        po.cmask = cm | (1ull << cx[0]);
        po.m[2] = 1.0;
        po.m[4] = 1.0;
        ops.push_back(po);
      }
    } else {
      return fail(MSV_E_INVALID, "gate: unknown kind " + std::to_string(G.kind));
    }
  }
  MSV_TRY(set_dev(s));
  // A lone deferred basis state stays deferred: the gate list runs as tile
  // sweeps only, and the first one generates the basis state instead of
  // reading a written one. Other deferred work is flushed first.
  const bool lone_basis = s->basis_on && !s->pend_on && !s->perm_on && !s->tail_on;
  if (!lone_basis || ops.empty()) MSV_TRY(flush_pending(s));
  return MSV_DISPATCH(s, apply_plan_ops, s, ops);
}

int msv_apply_circuit(msv_state *s, const int32_t *codes, const int32_t *wires, const double *params,
                      int n_ops) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (n_ops < 0 || (n_ops > 0 && (!codes || !wires || !params)))
    return fail(MSV_E_INVALID, "circuit: null arrays");
  std::vector<msv_gate> gates;
  gates.reserve(static_cast<size_t>(n_ops) + 8);
  for (int k = 0; k < n_ops; ++k) {
    const int code = codes[k];
    if (code < MSV_OP_I || code > MSV_OP_FREDKIN)
      return fail(MSV_E_INVALID, "circuit: unknown gate code " + std::to_string(code));
    const int32_t *w = wires + 3 * k;
    const int ar = circuit_arity(code);
    for (int i = 0; i < ar; ++i) {
      if (w[i] < 0 || w[i] >= s->n) return fail(MSV_E_INVALID, "circuit: qubit index out of range");
      for (int j = 0; j < i; ++j)
        if (w[i] == w[j]) return fail(MSV_E_INVALID, "circuit: repeated wire in one gate");
    }
    const double *p = params + 3 * k;
    for (int i = 0; i < 3; ++i)
      if (!std::isfinite(p[i])) return fail(MSV_E_INVALID, "circuit: non-finite angle");
    MSV_TRY(circuit_op_gates(code, w, p, &gates));
  }
  return msv_apply_gates(s, gates.data(), static_cast<int>(gates.size()));
}

int msv_probabilities(msv_state *s, const int32_t *wires, int nw, double *out) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (!out) return fail(MSV_E_INVALID, "null output");
  if (wires) {
    if (nw < 0 || nw > s->n) return fail(MSV_E_INVALID, "probabilities: bad wire count");
    uint64_t seen = 0;
    for (int w = 0; w < nw; ++w) {
      MSV_TRY(check_qubit(s, wires[w], "probabilities"));
      if ((seen >> wires[w]) & 1ull) return fail(MSV_E_INVALID, "Duplicate qubit indices");
      seen |= 1ull << wires[w];
    }
  }
  MSV_TRY(set_dev(s));
  MSV_TRY(flush_pending(s));
  return MSV_DISPATCH(s, probabilities_T, s, wires, nw, out);
}

int msv_sample(msv_state *s, const int32_t *wires, int nw, uint64_t shots, uint64_t seed,
               uint64_t stream_id, uint64_t *out) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (shots && !out) return fail(MSV_E_INVALID, "null output");
  std::vector<int32_t> all;
  if (!wires) {
    for (int q = 0; q < s->n; ++q) all.push_back(q);
    wires = all.data();
    nw = s->n;
  }
  if (nw < 0 || nw > s->n || nw > 64) return fail(MSV_E_INVALID, "sample: bad wire count");
  uint64_t seen = 0;
  for (int w = 0; w < nw; ++w) {
    MSV_TRY(check_qubit(s, wires[w], "sample"));
    if ((seen >> wires[w]) & 1ull) return fail(MSV_E_INVALID, "Duplicate qubit indices");
    seen |= 1ull << wires[w];
  }
  MSV_TRY(set_dev(s));
  MSV_TRY(flush_pending(s));
  return MSV_DISPATCH(s, sample_T, s, wires, nw, shots, seed, stream_id, out);
}

int msv_sample_noalloc(msv_state *s, const int32_t *wires, int nw, uint64_t shots, uint64_t seed,
                       uint64_t stream_id, uint64_t *out) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (s->dry) return fail(MSV_E_UNSUPPORTED, "a virtual state holds no amplitudes");
  if (shots && !out) return fail(MSV_E_INVALID, "null output");
  if (!wires) nw = s->n;
  if (nw < 0 || nw > s->n || nw > 64) return fail(MSV_E_INVALID, "sample: bad wire count");
  uint64_t seen = 0;
  for (int w = 0; wires && w < nw; ++w) {
    MSV_TRY(check_qubit(s, wires[w], "sample"));
    if ((seen >> wires[w]) & 1ull) return fail(MSV_E_INVALID, "Duplicate qubit indices");
    seen |= 1ull << wires[w];
  }
  const uint64_t nb = (s->amps + kSampleBlock - 1) / kSampleBlock;
  const size_t prefix_bytes = ((nb + 1) * sizeof(double) + 255) & ~size_t(255);
  if (prefix_bytes + 64 * (2 * sizeof(uint64_t) + sizeof(double)) > kNoAllocScratchBytes)
    return fail(MSV_E_REFUSED, "sample_noalloc: a " + std::to_string(s->n) +
                                   "-qubit state needs " + std::to_string(prefix_bytes) +
                                   " bytes of prefix scratch; the static scratch holds " +
                                   std::to_string(kNoAllocScratchBytes) + " (at most 31 qubits)");
  MSV_TRY(set_dev(s));
  char *scr = nullptr;
  MSV_TRY(noalloc_scratch(s->device, &scr));
  t_noalloc = true;
  int rc = flush_pending(s);
  t_noalloc = false;
  MSV_TRY(rc);
  if (shots == 0) return MSV_OK;
  return MSV_DISPATCH(s, sample_noalloc_T, s, wires, nw, shots, seed, stream_id, out, scr);
}

int msv_expval_pauli(msv_state *s, const uint64_t *xmask, const uint64_t *zmask, int n_words,
                     double *out_re, double *out_im) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (n_words < 0 || (n_words > 0 && (!xmask || !zmask || !out_re || !out_im)))
    return fail(MSV_E_INVALID, "expval_pauli: null arrays");
  if (n_words == 0) return MSV_OK;
  MSV_TRY(set_dev(s));
  MSV_TRY(flush_pending(s));
  return MSV_DISPATCH(s, pauli_T, s, xmask, zmask, n_words, out_re, out_im);
}

int msv_inner(msv_state *a, msv_state *b, double *re, double *im) {
  TuneScope ts_;
  MSV_TRY(check_state(a));
  MSV_TRY(check_state(b));
  if (!re || !im) return fail(MSV_E_INVALID, "null output");
  if (a->dry || b->dry) return fail(MSV_E_UNSUPPORTED, "a virtual state holds no amplitudes");
  if (a->n != b->n) return fail(MSV_E_INVALID, "inner: qubit counts differ");
  if (a->device != b->device) return fail(MSV_E_INVALID, "inner: states on different devices");
  MSV_TRY(flush_pending(a));
  MSV_TRY(flush_pending(b));
  MSV_TRY(set_dev(a));
  if (a->dtype == MSV_C64 && b->dtype == MSV_C64) return inner_T<float, float>(a, b, re, im);
  if (a->dtype == MSV_C64) return inner_T<float, double>(a, b, re, im);
  if (b->dtype == MSV_C64) return inner_T<double, float>(a, b, re, im);
  return inner_T<double, double>(a, b, re, im);
}

int msv_norm2(msv_state *s, double *out) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (!out) return fail(MSV_E_INVALID, "null output");
  MSV_TRY(set_dev(s));
  MSV_TRY(flush_pending(s));
  std::vector<double> bs;
  int rc = s->dtype == MSV_C128 ? block_sums_T<double>(s, &bs) : block_sums_T<float>(s, &bs);
  if (rc != MSV_OK) return rc;
  double t = 0.0;
  for (double v : bs) t += v;
  *out = t;
  return MSV_OK;
}

int msv_set_tuning(int key, int64_t value) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  switch (key) {
    case MSV_TUNE_TILE_LOCAL_BITS:
      if (value < 6 || value > TileCfg<float>::kMaxLocal)
        return fail(MSV_E_INVALID, "tile local bits out of range");
      g_tune_store.local_bits_c64 = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_TILE_GATE_BITS:
      if (value < 5 || value > 10) return fail(MSV_E_INVALID, "gate bits out of range");
      g_tune_store.gate_bits = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_PHASE_BLOCKS:
      if (value < 0 || value > (1 << 20)) return fail(MSV_E_INVALID, "phase blocks out of range");
      g_tune_store.phase_blocks = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_SWEEP_SPLIT:
      if (value < 0 || value > 1) return fail(MSV_E_INVALID, "sweep split policy out of range");
      g_tune_store.sweep_split = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_TILE_BLOCKS_PER_CU:
      if (value < 1 || value > 4096) return fail(MSV_E_INVALID, "tile blocks per CU out of range");
      g_tune_store.tile_blocks_per_cu = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_ELEM_BLOCKS_PER_CU:
      if (value < 1 || value > 4096) return fail(MSV_E_INVALID, "elementwise blocks per CU out of range");
      g_tune_store.elem_blocks_per_cu = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_COL_BITS:
      if (value < 0 || value > 8) return fail(MSV_E_INVALID, "column bits out of range");
      g_tune_store.col_bits_c64 = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_ROT_SPLIT_MIN:
      if (value < 0 || value > 64) return fail(MSV_E_INVALID, "rotation split threshold out of range");
      g_tune_store.rot_split_min = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_XCD_GROUPS:
      if (value < 0 || value > 64) return fail(MSV_E_INVALID, "XCD groups out of range");
      g_tune_store.xcd_groups = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_TILE_REG_BITS:
      if (value != 4 && value != 5) return fail(MSV_E_INVALID, "tile register bits must be 4 or 5");
      g_tune_store.reg_bits_c64 = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_TILE_PREFETCH:
      if (value < 0 || value > 1) return fail(MSV_E_INVALID, "tile prefetch must be 0 or 1");
      g_tune_store.tile_prefetch = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_DEFER_DIAG:
      if (value < 0 || value > 1) return fail(MSV_E_INVALID, "defer diag must be 0 or 1");
      g_tune_store.defer_diag = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_DEFER_PERM:
      if (value < 0 || value > 1) return fail(MSV_E_INVALID, "defer perm must be 0 or 1");
      g_tune_store.defer_perm = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_FUSED_BLOCKS_PER_CU:
      if (value < 1 || value > 4096) return fail(MSV_E_INVALID, "fused blocks per CU out of range");
      g_tune_store.fused_blocks_per_cu = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_REORDER_GATES:
      if (value < 0 || value > 1) return fail(MSV_E_INVALID, "reorder gates must be 0 or 1");
      g_tune_store.reorder_gates = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_DEFER_TAIL:
      if (value < 0 || value > 2) return fail(MSV_E_INVALID, "defer tail must be 0, 1 or 2");
      g_tune_store.defer_tail = static_cast<int>(value);
      return MSV_OK;
    case MSV_TUNE_LANE_SWAP:
      if (value < 0 || value > 1) return fail(MSV_E_INVALID, "lane swap must be 0 or 1");
      g_tune_store.lane_swap = static_cast<int>(value);
      return MSV_OK;
    default: return fail(MSV_E_INVALID, "unknown tuning key");
  }
}

int msv_get_tuning(int key, int64_t *value) {
  if (!value) return fail(MSV_E_INVALID, "null output");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  switch (key) {
    case MSV_TUNE_TILE_LOCAL_BITS: *value = g_tune_store.local_bits_c64; return MSV_OK;
    case MSV_TUNE_TILE_GATE_BITS: *value = g_tune_store.gate_bits; return MSV_OK;
    case MSV_TUNE_PHASE_BLOCKS: *value = g_tune_store.phase_blocks; return MSV_OK;
    case MSV_TUNE_SWEEP_SPLIT: *value = g_tune_store.sweep_split; return MSV_OK;
    case MSV_TUNE_TILE_BLOCKS_PER_CU: *value = g_tune_store.tile_blocks_per_cu; return MSV_OK;
    case MSV_TUNE_ELEM_BLOCKS_PER_CU: *value = g_tune_store.elem_blocks_per_cu; return MSV_OK;
    case MSV_TUNE_COL_BITS: *value = g_tune_store.col_bits_c64; return MSV_OK;
    case MSV_TUNE_ROT_SPLIT_MIN: *value = g_tune_store.rot_split_min; return MSV_OK;
    case MSV_TUNE_XCD_GROUPS: *value = g_tune_store.xcd_groups; return MSV_OK;
    case MSV_TUNE_TILE_REG_BITS: *value = g_tune_store.reg_bits_c64; return MSV_OK;
    case MSV_TUNE_TILE_PREFETCH: *value = g_tune_store.tile_prefetch; return MSV_OK;
    case MSV_TUNE_DEFER_DIAG: *value = g_tune_store.defer_diag; return MSV_OK;
    case MSV_TUNE_DEFER_PERM: *value = g_tune_store.defer_perm; return MSV_OK;
    case MSV_TUNE_FUSED_BLOCKS_PER_CU: *value = g_tune_store.fused_blocks_per_cu; return MSV_OK;
    case MSV_TUNE_REORDER_GATES: *value = g_tune_store.reorder_gates; return MSV_OK;
    case MSV_TUNE_DEFER_TAIL: *value = g_tune_store.defer_tail; return MSV_OK;
    case MSV_TUNE_LANE_SWAP: *value = g_tune_store.lane_swap; return MSV_OK;
    default: return fail(MSV_E_INVALID, "unknown tuning key");
  }
}

int msv_state_sweep_count(msv_state *s, uint64_t *out, int reset) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (out) *out = s->sweeps;
  if (reset) s->sweeps = 0;
  return MSV_OK;
}

int msv_state_pending(const msv_state *s, int *out) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (!out) return fail(MSV_E_INVALID, "null output");
  *out = (s->basis_on ? MSV_PENDING_BASIS : 0) | (s->pend_on ? MSV_PENDING_DIAG : 0) |
         (s->perm_on ? MSV_PENDING_PERM : 0) | (s->tail_on ? MSV_PENDING_TAIL : 0);
  return MSV_OK;
}

int msv_state_timer_start(msv_state *s) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (s->dry) return fail(MSV_E_UNSUPPORTED, "a virtual state holds no amplitudes");
  MSV_TRY(set_dev(s));
  if (!s->ev0) MSV_HIP(hipEventCreate(&s->ev0));
  if (!s->ev1) MSV_HIP(hipEventCreate(&s->ev1));
  MSV_HIP(hipEventRecord(s->ev0, s->stream));
  return MSV_OK;
}

int msv_state_timer_stop(msv_state *s, double *ms) {
  TuneScope ts_;
  MSV_TRY(check_state(s));
  if (s->dry) return fail(MSV_E_UNSUPPORTED, "a virtual state holds no amplitudes");
  if (!ms) return fail(MSV_E_INVALID, "null output");
  if (!s->ev0 || !s->ev1) return fail(MSV_E_INVALID, "timer was not started");
  MSV_TRY(set_dev(s));
  MSV_HIP(hipEventRecord(s->ev1, s->stream));
  MSV_HIP(hipEventSynchronize(s->ev1));
  float f = 0.f;
  MSV_HIP(hipEventElapsedTime(&f, s->ev0, s->ev1));
  *ms = static_cast<double>(f);
  return MSV_OK;
}

int msv_probe_copy_bandwidth(int device, uint64_t bytes, int iters, double *ms_per_copy) {
  TuneScope ts_;
  if (!ms_per_copy || iters < 1 || bytes < 16) return fail(MSV_E_INVALID, "probe: bad arguments");
  MSV_HIP(hipSetDevice(device));
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    if (a) (void)hipFree(a);
    (void)hipGetLastError();
    return fail(MSV_E_OOM, "probe: allocation failed");
  }
  hipStream_t st;
  MSV_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  MSV_HIP(hipMemsetAsync(a, 0, bytes, st));
  const uint64_t nvec = bytes / 16;
  MSV_HIP(launch_copy_probe(static_cast<const float4 *>(a), static_cast<float4 *>(b), nvec, st));
  hipEvent_t e0, e1;
  MSV_HIP(hipEventCreate(&e0));
  MSV_HIP(hipEventCreate(&e1));
  MSV_HIP(hipEventRecord(e0, st));
  for (int i = 0; i < iters; ++i)
    MSV_HIP(launch_copy_probe(static_cast<const float4 *>(a), static_cast<float4 *>(b), nvec, st));
  MSV_HIP(hipEventRecord(e1, st));
  MSV_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  MSV_HIP(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_copy = static_cast<double>(ms) / iters;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(st);
  (void)hipFree(a);
  (void)hipFree(b);
  return MSV_OK;
}

}  // extern "C"

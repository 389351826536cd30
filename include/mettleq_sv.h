/*
 * mettleq_sv.h -- C ABI of the MI355X (gfx950) statevector engine.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.
 * The Python host binds it with ctypes (mettleq_amd/_native.py); any other
 * FFI (cffi, a C++ runtime, a serving process) can bind the same symbols.
 *
 * Conventions (identical to the reference engine):
 *   - qubit q of an n-qubit register is bit (n-1-q) of the basis index,
 *     so qubit 0 is the most significant bit (reference sim.py:10-14,
 *     shaders/xor_affine.py:12);
 *   - a 2x2 matrix is passed as 8 doubles, row-major, (re, im) pairs:
 *     u00.re u00.im u01.re u01.im u10.re u10.im u11.re u11.im;
 *   - amplitudes are interleaved (re, im) complex64 (MSV_C64) or complex128
 *     (MSV_C128); indices are 64-bit (the reference stops at 31 qubits
 *     through uint32 indices, execution.py:21-29).
 *
 * Errors: every call returns MSV_OK (0) or a negative-free positive code;
 * msv_last_error() returns the calling thread's last message.
 * Ownership: the library owns device buffers, callers own host buffers.
 * Threading: calls on distinct states may run concurrently from different
 * host threads; each state owns one HIP stream and all its work is ordered
 * on it. Calls on ONE state must not overlap.
 */
#ifndef METTLEQ_SV_H
#define METTLEQ_SV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define MSV_OK 0
#define MSV_E_INVALID 1     /* bad argument (ValueError on the Python side)  */
#define MSV_E_OOM 2         /* device allocation failed                       */
#define MSV_E_HIP 3         /* HIP runtime error                              */
#define MSV_E_REFUSED 4     /* refused by a size/limit check                  */
#define MSV_E_UNSUPPORTED 5 /* valid request this build does not implement    */

/* ---- amplitude types --------------------------------------------------- */
#define MSV_C64 0
#define MSV_C128 1

/* ---- phase-term kinds for msv_apply_phase_layer ------------------------ */
#define MSV_PHASE_ZZ 0     /* exp(-i*theta*Z_a Z_b)   (ZZPHASE, zz.py:78-84)  */
#define MSV_PHASE_CPHASE 1 /* diag(1,1,1,e^{i theta}) (CPHASE/CZ, diag.py:89) */

/* ---- gate kinds for msv_apply_gates ------------------------------------ */
#define MSV_G_U 0        /* 2x2 on q0 with ncontrols controls in q[1..]      */
#define MSV_G_DIAG1 1    /* diag(d0, d1) on q0: m[0..3]                      */
#define MSV_G_DIAG2 2    /* diag(d00,d01,d10,d11) on (q0 = MSB, q1): m[0..7] */
#define MSV_G_SWAP 3     /* SWAP(q0, q1), optionally controlled by q[2..]    */
#define MSV_G_QFT_STAGE 4 /* one QFT stage, see msv_apply_qft_stage          */

typedef struct msv_state msv_state;

/* One gate of a generic gate list (msv_apply_gates). */
typedef struct msv_gate {
  int32_t kind;       /* MSV_G_*                                           */
  int32_t ncontrols;  /* MSV_G_U / MSV_G_SWAP: number of control qubits    */
  int32_t q[6];       /* target(s) first, then controls                    */
  double m[8];        /* matrix / diagonal entries (see MSV_G_*)           */
} msv_gate;

/* Device memory and identity, for the statevector preflight. */
typedef struct msv_device_info {
  uint64_t total_bytes;      /* HBM capacity                              */
  uint64_t free_bytes;       /* free HBM reported by the runtime now      */
  uint64_t max_alloc_bytes;  /* largest single allocation the runtime allows */
  int32_t compute_units;
  int32_t lds_bytes_per_block;
  char name[96];
  char arch[32];
} msv_device_info;

/* ---- library ----------------------------------------------------------- */
const char *msv_last_error(void);
int msv_version(void);
int msv_device_count(int *out);
/* max_alloc_bytes is the runtime's largest single allocation (ALLOC_MAX_SIZE
 * of the device's coarse-grained HSA memory pool), 0 when not reported. */
int msv_get_device_info(int device, msv_device_info *out);

/* ---- statevector preflight (reference execution.py:186-407) ------------- */
/* Lower bound of the check: two states (the reference's out-of-place model,
 * execution.py:195-223; circuits with XOR permutations) or one state plus
 * the library's fixed 8 MiB of scratch (in-place layer kernels). */
#define MSV_PF_OUT_OF_PLACE 0
#define MSV_PF_IN_PLACE 1
/* Decisions (the reference's decision strings, execution.py:316-331). */
#define MSV_PF_ALLOWED 0            /* "allowed_within_reported_lower_bounds"  */
#define MSV_PF_ALLOWED_UNVERIFIED 1 /* "allowed_with_unverified_device_limits" */
#define MSV_PF_ALLOWED_OVERRIDE 2   /* "allowed_with_unsafe_override"          */
#define MSV_PF_REFUSED 3            /* "refused"                               */
/* Failure bits (failure_reasons, execution.py:306-311). */
#define MSV_PF_FAIL_MAX_BUFFER 1    /* "single_state_exceeds_max_buffer_length" */
#define MSV_PF_FAIL_WORKING_SET 2   /* "two_state_lower_bound_exceeds_recommended_working_set"
                                       or "in_place_lower_bound_exceeds_recommended_working_set" */
#define MSV_PF_FAIL_QUBITS 4        /* "qubit_count_exceeds_engine_limit" (above 40)  */
typedef struct msv_preflight_limits {
  uint64_t max_buffer_bytes;   /* largest single allocation; 0 = unknown            */
  uint64_t working_set_bytes;  /* memory available for new states; 0 = unknown     */
} msv_preflight_limits;
typedef struct msv_preflight_report {
  int32_t decision;            /* MSV_PF_ALLOWED ... MSV_PF_REFUSED                 */
  int32_t failures;            /* MSV_PF_FAIL_* bits                                */
  int32_t allowed;             /* 1 when the state may be allocated                 */
  int32_t overridden;          /* 1 when allowed only by allow_unsafe               */
  int32_t buffer_check;        /* state <= max buffer: 1 yes, 0 no, -1 unknown      */
  int32_t working_check;       /* lower bound <= working set: 1, 0, -1              */
  uint64_t state_bytes;        /* 2^n x 8 (complex64) or x 16 (complex128)          */
  uint64_t lower_bound_bytes;  /* per `mode`                                        */
  uint64_t max_buffer_bytes;   /* the limit checked (0 = unknown)                   */
  uint64_t working_set_bytes;  /* working set minus reserved_bytes (0 = unknown)    */
} msv_preflight_report;
/* Checks an n-qubit state of `dtype` before allocation, like the reference's
 * statevector_preflight (execution.py:270-374). `limits` == NULL reads the
 * live limits of `device`: its largest single allocation and its free memory
 * now, which already excludes every state allocated so far, so
 * `reserved_bytes` is only memory admitted but not yet allocated (the
 * serving loop's admissions in flight). allow_unsafe != 0 is the reference's
 * override (METTLEQ_ALLOW_UNSAFE_STATEVECTOR): it admits a refusal, except a
 * qubit count above the engine limit. Returns MSV_OK when allowed and
 * MSV_E_REFUSED when refused; `out` is filled in both cases and a refusal's
 * message (msv_last_error) says "before allocation" and names the override
 * variable, as the reference's StatevectorMemoryError does
 * (execution.py:377-407). */
int msv_preflight(int n_qubits, int dtype, int mode, uint64_t reserved_bytes, int allow_unsafe,
                  int device, const msv_preflight_limits *limits, msv_preflight_report *out);

/* ---- state lifecycle (reference sim.py:17-45) -------------------------- */
/* Allocates 2^n amplitudes on `device` and sets |0...0>. */
int msv_state_create(int n_qubits, int dtype, int device, msv_state **out);
int msv_state_destroy(msv_state *s);
/* |0...0> / |index>. Nothing is written until the state's next operation:
 * a 1q layer then writes its product state in one pass, any other call
 * writes the basis state first (msv_state_sync and msv_state_device_ptr
 * included). */
int msv_state_reset(msv_state *s);
int msv_state_set_basis(msv_state *s, uint64_t index);
int msv_state_info(const msv_state *s, int *n_qubits, int *dtype, uint64_t *bytes);
/* Copies `count` amplitudes at amplitude `offset` between host and device
 * (host layout = the state's dtype). Synchronous. */
int msv_state_upload(msv_state *s, const void *host, uint64_t offset, uint64_t count);
int msv_state_download(msv_state *s, void *host, uint64_t offset, uint64_t count);
/* dst <- src (same qubit count; converts between c64 and c128). */
int msv_state_copy(msv_state *dst, msv_state *src);
/* Multiplies every amplitude by (re + i im): global phase (_common.py:231-238). */
int msv_state_scale(msv_state *s, double re, double im);
int msv_state_sync(msv_state *s);
/* Releases the out-of-place scratch buffer kept after an XOR permutation. */
int msv_state_release_scratch(msv_state *s);
/* Raw device pointer of the current amplitude buffer (for zero-copy
 * interop, e.g. DLPack); valid until the next call that permutes out of
 * place on this state. */
void *msv_state_device_ptr(msv_state *s);

/* ---- structured layers -------------------------------------------------- */
/* Per-qubit 2x2 layer: mats[q*8 .. q*8+7] acts on qubit q; active[q] == 0
 * skips qubit q (NULL = all active). Replaces u2_list_layer_all /
 * u2_layer_all / rx_layer_all (shaders/single_qubit.py:420-673). */
int msv_apply_1q_layer(msv_state *s, const double *mats, const uint8_t *active);

/* Commuting two-qubit phase terms in one pass: term b acts on qubits
 * (bonds[2b], bonds[2b+1]) with angle thetas[b]; kinds[b] is MSV_PHASE_*.
 * Replaces diag_pair_layer / diag_weighted_layer (shaders/diag.py:123-217)
 * and zz_chain_layer / zz_weighted_layer (shaders/zz.py:99-185). */
int msv_apply_phase_layer(msv_state *s, const int32_t *kinds, const int32_t *bonds,
                          const double *thetas, int n_terms);

/* Generic diagonal: n1 one-qubit diagonals (qubit q1[k], entries d1[4k..4k+3]
 * = d0.re d0.im d1.re d1.im) and n2 two-qubit diagonals (qubits q2[2k] (MSB),
 * q2[2k+1], entries d2[8k..8k+7]) applied as one pass.
 *
 * Both diagonal entry points may defer their pass (MSV_TUNE_DEFER_DIAG, on
 * by default): the state then holds the diagonal until its next operation.
 * msv_apply_xor_affine applies it inside the permutation's own sweep, and
 * msv_apply_1q_layer and msv_apply_qft may read it in their first sweep;
 * every other call launches it first, so results never differ from
 * immediate application (up to rounding of the phase products). */
int msv_apply_diagonal(msv_state *s, const int32_t *q1, const double *d1, int n1,
                       const int32_t *q2, const double *d2, int n2);

/* GF(2) affine permutation out[y] = state[M'y ^ c]: qubit k of the source
 * index is parity(rows[k] & y) ^ bit(n-1-k) of c. Replaces
 * xor_affine_gather (shaders/xor_affine.py:79-116).
 *
 * May defer its pass (MSV_TUNE_DEFER_PERM, on by default), together with a
 * pending diagonal: msv_apply_1q_layer and msv_apply_qft then read their
 * first sweep through the permutation when it keeps the lowest 4 (complex64)
 * or 3 (complex128) index bits among themselves, as a CNOT chain does;
 * every other call runs the gather first. */
int msv_apply_xor_affine(msv_state *s, const uint64_t *rows, uint64_t c);

/* QFT without terminal swaps on the contiguous qubits first..m-1 (first =
 * MSB of the subregister); inverse != 0 applies its exact inverse.
 * Replaces qft_stage_all (shaders/qft.py:207-236) and sim.qft/iqft
 * (sim.py:522-563) on contiguous ascending wires. */
int msv_apply_qft(msv_state *s, int first, int m, int inverse);

/* One QFT stage on qubit j of the subregister [j..m-1]: Hadamard butterfly
 * plus the controlled-phase ladder toward j+1..m-1 (inverse: adjoint).
 * Replaces qft_stage_sub (shaders/qft.py:144-181). */
int msv_apply_qft_stage(msv_state *s, int j, int m, int inverse);

/* Ordered list of generic gates, fused into tile sweeps by the library.
 * Replaces the per-gate StateVectorSimulator.apply_* calls dispatched by
 * Device._apply (device.py:1138-1290) and controlled_u_groups
 * (shaders/controlled_single.py:89-134).
 *
 * With MSV_TUNE_REORDER_GATES (on by default) the library may reorder gates
 * that commute (a gate's non-diagonal action on its target commutes with
 * any gate not touching that qubit; diagonal actions commute with each
 * other) and multiply adjacent gates on the same target and controls, so
 * results equal in-order application up to floating-point rounding. */
int msv_apply_gates(msv_state *s, const msv_gate *gates, int n_gates);

/* Gate codes of msv_apply_circuit: names, wire order and angle conventions
 * of the reference (gates.py:5-234, device.py:1138-1290): U3(t, p, l) =
 * RZ(p) RY(t) RZ(l); RZ(t) = diag(e^{-it/2}, e^{it/2}); U1 = P = phase;
 * ZZ/XX/YYPHASE(t) = exp(-i t P P); two-qubit gates take (control, target)
 * or (MSB, LSB) wires. */
#define MSV_OP_I 0
#define MSV_OP_X 1
#define MSV_OP_Y 2
#define MSV_OP_Z 3
#define MSV_OP_H 4
#define MSV_OP_S 5
#define MSV_OP_SDG 6
#define MSV_OP_T 7
#define MSV_OP_TDG 8
#define MSV_OP_SX 9
#define MSV_OP_RX 10
#define MSV_OP_RY 11
#define MSV_OP_RZ 12
#define MSV_OP_U1 13
#define MSV_OP_U2 14
#define MSV_OP_U3 15
#define MSV_OP_CNOT 16
#define MSV_OP_CZ 17
#define MSV_OP_CPHASE 18
#define MSV_OP_CRX 19
#define MSV_OP_CRY 20
#define MSV_OP_CRZ 21
#define MSV_OP_CH 22
#define MSV_OP_SWAP 23
#define MSV_OP_ISWAP 24
#define MSV_OP_ZZPHASE 25
#define MSV_OP_XXPHASE 26
#define MSV_OP_YYPHASE 27
#define MSV_OP_TOFFOLI 28
#define MSV_OP_FREDKIN 29
/* Ordered circuit in compact form: op k is codes[k] on wires[3k..3k+2]
 * (unused = -1) with angles params[3k..3k+2]. The gate matrices are built
 * natively and the gates fused into tile sweeps as by msv_apply_gates.
 * Replaces the per-op translation of integrations/_common.py:29-115 plus
 * the gate routing of device.py:1138-1290 for callers that submit many
 * small circuits (the serving loop). */
int msv_apply_circuit(msv_state *s, const int32_t *codes, const int32_t *wires,
                      const double *params, int n_ops);

/* ---- reductions ----------------------------------------------------------- */
/* FP64 probabilities of the requested wires, first wire = MSB of the output
 * index (sim.py:322-343); wires == NULL means all qubits in order. `out`
 * holds 2^nw doubles. Not renormalized. */
int msv_probabilities(msv_state *s, const int32_t *wires, int nw, double *out);

/* Draws `shots` basis indices from |a|^2 with a counter-based Philox4x32-10
 * stream keyed by (seed, stream_id) and writes, per shot, the requested
 * wires' bits packed MSB-first (wire 0 of the list = highest bit) into
 * out[shot]. Deterministic for a given state, seed and stream_id. */
int msv_sample(msv_state *s, const int32_t *wires, int nw, uint64_t shots,
               uint64_t seed, uint64_t stream_id, uint64_t *out);

/* msv_sample for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No
 * counterpart in the reference. Same arguments and the same output as
 * msv_sample, bit for bit, but it calls no hipMalloc/hipFree and uses no
 * host buffer besides `out`: the block-sum prefix and the per-shot staging
 * live in a static 8 MiB device array per GPU (held until the process ends),
 * shots run in chunks that fit behind the prefix. States of at most 31
 * qubits (MSV_E_REFUSED above). Deferred work of the state is flushed with
 * the state's existing buffers; if it would need a new one, the call
 * returns MSV_E_REFUSED (run msv_state_sync once beforehand). Every call on
 * one GPU shares that GPU's scratch: calls that overlap on one GPU race. */
int msv_sample_noalloc(msv_state *s, const int32_t *wires, int nw, uint64_t shots,
                       uint64_t seed, uint64_t stream_id, uint64_t *out);

/* <psi|P_w|psi> for Pauli words P_w = i^{ny_w} X^{xmask_w} Z^{zmask_w}
 * (masks over qubits: bit q = qubit q; ny_w = number of Y factors). Returns
 * real and imaginary parts. Replaces pauli_product_expectations
 * (integrations/_common.py:384-427). */
int msv_expval_pauli(msv_state *s, const uint64_t *xmask, const uint64_t *zmask,
                     int n_words, double *out_re, double *out_im);

/* FP64 <a|b> over two states of equal qubit count (any dtypes). */
int msv_inner(msv_state *a, msv_state *b, double *re, double *im);
/* FP64 sum |a_i|^2. */
int msv_norm2(msv_state *s, double *out);

/* ---- batched bindings of a layered rotation ansatz, adjoint gradients ---- */
/* An ansatz is an ordered gate list of parameterized rotations
 * R_P(theta) = exp(-i theta P / 2), P = X, Y, Z (theta = params[b][param_index]
 * for binding b) and fixed two-qubit phases (MSV_PHASE_ZZ / MSV_PHASE_CPHASE,
 * e.g. CZ = CPHASE(pi)), with a diagonal observable H = sum_t w_t Z^{zmask_t}
 * (masks over qubits: bit q = qubit q). msv_ansatz_evaluate runs every
 * binding from |0...0> in one graph launch and returns the energies
 * <psi_b|H|psi_b> and, when grads != NULL, d energy_b / d params[b][p] by
 * the adjoint method. The first call for a batch size autotunes the tile
 * shape and captures the launch sequence as a HIP graph. Replaces the
 * per-binding estimator loop (integrations/qiskit.py:906-1090) and
 * parameter-shift differentiation of the PennyLane device
 * (integrations/pennylane.py:558-625). */
#define MSV_ROT_X 0
#define MSV_ROT_Y 1
#define MSV_ROT_Z 2
typedef struct msv_ansatz msv_ansatz;
int msv_ansatz_create(int n_qubits, int device, msv_ansatz **out);
int msv_ansatz_destroy(msv_ansatz *a);
int msv_ansatz_rotation(msv_ansatz *a, int kind, int qubit, int param_index);
int msv_ansatz_phase(msv_ansatz *a, int kind, int q0, int q1, double theta);
int msv_ansatz_observable(msv_ansatz *a, const uint64_t *zmasks, const double *weights, int n_terms);
/* params: host [batch][n_params]; energies: host [batch]; grads: host
 * [batch][n_params] or NULL. Synchronous. */
int msv_ansatz_evaluate(msv_ansatz *a, const double *params, int batch, int n_params,
                        double *energies, double *grads);
/* Chosen plan of a batch size: out[0] tile local bits, out[1] blocks per CU,
 * out[2] forward sweeps, out[3] backward sweeps, out[4] tile ops,
 * out[5] autotuned time in microseconds, out[6] sweeps run by a kernel
 * compiled for their stage-generator program (the rest interpret it). */
int msv_ansatz_info(msv_ansatz *a, int batch, int with_grad, int64_t *out, int n_out);
/* Times the cached plan's sweeps alone (no graph), averaged over `reps`:
 * out_ms[0] the forward sweeps, out_ms[1] the backward sweeps (0 without
 * gradients), measured with HIP events on the ansatz's stream. The states
 * are left holding no meaningful values until the next evaluation. No
 * counterpart in the reference: it is the measurement behind the sweeps'
 * HBM fractions. */
int msv_ansatz_time_sweeps(msv_ansatz *a, int batch, int with_grad, int reps, double *out_ms);

/* ---- tuning knobs ------------------------------------------------------- */
/* Sets an integer tuning parameter (MSV_TUNE_*); returns MSV_E_INVALID for
 * unknown keys. Values apply to subsequent launches from any state. */
#define MSV_TUNE_TILE_LOCAL_BITS 1   /* local bits per tile sweep (c64)        */
#define MSV_TUNE_TILE_GATE_BITS 2    /* max non-column bits per high sweep     */
#define MSV_TUNE_PHASE_BLOCKS 3      /* persistent grid size of phase kernels   */
#define MSV_TUNE_SWEEP_SPLIT 4       /* 1q-layer high sweeps: 0 contiguous bits, 1 page-balanced */
#define MSV_TUNE_TILE_BLOCKS_PER_CU 5 /* persistent tile-sweep blocks per CU     */
#define MSV_TUNE_ELEM_BLOCKS_PER_CU 6 /* elementwise/gather blocks per CU        */
#define MSV_TUNE_COL_BITS 7           /* contiguous index bits per tile row (c64) */
#define MSV_TUNE_ROT_SPLIT_MIN 8      /* gates per sweep from which 1q gates run as
                                         diagonal x real rotation x diagonal (c64;
                                         0 = never) */
#define MSV_TUNE_XCD_GROUPS 9         /* tile sweeps: blocks of one XCD take
                                         consecutive tiles (0 = plain order) */
#define MSV_TUNE_TILE_REG_BITS 10     /* c64 tile sweeps: amplitudes per thread
                                         2^bits (5: 256-thread tiles, 4: 512) */
#define MSV_TUNE_TILE_PREFETCH 11     /* c64 512-thread tile sweeps: load the next
                                         tile while the current one runs (0/1) */
#define MSV_TUNE_DEFER_DIAG 12        /* phase layers / diagonals wait for the next
                                         operation on the state; an XOR permutation
                                         then applies them in its own sweep (0/1) */
#define MSV_TUNE_DEFER_PERM 13        /* XOR permutations wait for the next operation;
                                         a 1q layer or QFT then reads its first sweep
                                         through the permutation (and a pending
                                         diagonal) (0/1) */
#define MSV_TUNE_FUSED_BLOCKS_PER_CU 14 /* grid of such a fused first sweep (blocks per CU) */
#define MSV_TUNE_REORDER_GATES 15     /* gate lists (msv_apply_gates, msv_apply_circuit):
                                         commuting gates are regrouped into fewer tile
                                         programs (0/1) */
#define MSV_TUNE_DEFER_TAIL 16        /* 1q layers of several sweeps leave their last
                                         sweep's rotations pending; the next 1q layer
                                         runs them in its first sweep, across a pending
                                         XOR permutation (0 off; 1 on, and without a
                                         pending permutation for states up to 64 GiB;
                                         2 on, also above) */
#define MSV_TUNE_LANE_SWAP 17         /* rotation sweeps: lane bits 4 and 5 trade
                                         places with register bits inside the
                                         wavefront (v_permlane16/32_swap), so a
                                         sweep needs fewer LDS exchanges (0/1) */
int msv_set_tuning(int key, int64_t value);
int msv_get_tuning(int key, int64_t *value);

/* ---- launch accounting (execution report) ------------------------------- */
/* Kernel kinds of the dispatch log. Sweeps over the full state: */
#define MSV_K_TILE 1              /* generic tile program (gate lists, controlled groups) */
#define MSV_K_ROT 2               /* 1q-layer rotation sweep                              */
#define MSV_K_ROT_FUSED 3         /* rotation sweep reading deferred work (gather form)   */
#define MSV_K_ROT_RELABEL 4       /* pending rotations, deferred work, next layer's first */
#define MSV_K_PRODUCT 5           /* product state of a 1q layer on a basis state         */
#define MSV_K_QFT 6               /* QFT sweep                                            */
#define MSV_K_QFT_FUSED 7         /* QFT sweep reading deferred work                      */
#define MSV_K_PHASE 8             /* diagonal phase layer                                 */
#define MSV_K_XOR_GATHER 9        /* XOR permutation                                      */
#define MSV_K_XOR_GATHER_PHASE 10 /* XOR permutation with a deferred diagonal             */
#define MSV_K_INIT_BASIS 11       /* basis state                                          */
#define MSV_K_SCALE 12            /* global factor                                        */
#define MSV_K_CONVERT 13          /* state copy / precision conversion                    */
/* Reductions: */
#define MSV_K_BLOCK_PROBSUM 14
#define MSV_K_SAMPLE_RESOLVE 15
#define MSV_K_PROBS_FULL 16
#define MSV_K_MARGINAL 17
#define MSV_K_PAULI 18
#define MSV_K_INNER 19
#define MSV_K_PREFIX 20
#define MSV_K_SAMPLE_SEARCH 21
/* A virtual state: the planner of a real state of that width and dtype,
 * without device memory. Layer calls (and msv_state_sync) plan their sweeps,
 * defer work and log their kernels exactly as on a real state but launch
 * nothing; calls that read or write amplitudes return MSV_E_UNSUPPORTED.
 * The execution plan's predicted launches come from it: the counterpart of
 * the reference's launch prediction (execution.py:783-1029), measured on the
 * planner itself instead of a separate cost model. */
int msv_state_create_virtual(int n_qubits, int dtype, int device, msv_state **out);
/* Kinds (MSV_K_*) of the kernels launched on `s` since creation or the last
 * reset, in launch order: up to `cap` into `kinds` (the log keeps the first
 * 65536), their total in *count; reset != 0 clears the log. The
 * counterpart of the reference's observed_custom_dispatches
 * (execution.py:1032-1049). */
int msv_state_dispatch_log(msv_state *s, int32_t *kinds, int cap, uint64_t *count, int reset);
/* Number of full-state sweep kernels launched on this state since creation
 * or the last reset of the counter. */
int msv_state_sweep_count(msv_state *s, uint64_t *out, int reset);
/* Work the state holds back for the next operation (*out, a bit mask):
 * MSV_PENDING_BASIS a basis state not yet written, MSV_PENDING_DIAG a phase
 * layer, MSV_PENDING_PERM an XOR permutation, MSV_PENDING_TAIL a 1q layer's
 * last rotations. No counterpart in the reference (its layers run eagerly);
 * for execution reports and tests. */
#define MSV_PENDING_BASIS 1
#define MSV_PENDING_DIAG 2
#define MSV_PENDING_PERM 4
#define MSV_PENDING_TAIL 8
int msv_state_pending(const msv_state *s, int *out);

/* ---- timing ------------------------------------------------------------- */
/* HIP events on the state's own stream: start records the first event,
 * stop records the second, waits for it and returns the elapsed ms. */
int msv_state_timer_start(msv_state *s);
int msv_state_timer_stop(msv_state *s, double *ms);

/* ---- diagnostics -------------------------------------------------------- */
/* Times a plain device-to-device copy kernel of `bytes` (read + write) and
 * returns the mean milliseconds per copy; used to calibrate the achievable
 * HBM bandwidth and hardware byte counters. No reference counterpart. */
int msv_probe_copy_bandwidth(int device, uint64_t bytes, int iters, double *ms_per_copy);

/* ---- matrix product states ---------------------------------------------- */
/* The reference's MPS method (mps_state.py:250-1549 with its native core
 * _mps_native.cpp): site tensors (dl, 2, dr) in complex64, resident on the
 * device; two-site updates through an SVD on the device (one-sided Jacobi,
 * one workgroup) with the reference's truncation rule; the orthogonality center moves before each
 * update like the reference's QR moves. Site q is qubit q. */
typedef struct msv_mps msv_mps;

typedef struct {
  int64_t svd_calls;        /* SVDs run: one per two-site update, as the reference counts them */
  int64_t two_site_calls;   /* two-site updates (gates and swaps) */
  int64_t swap_calls;       /* swaps of the routing network */
  int64_t trunc_events;     /* two-site updates that kept fewer than their rank */
  int64_t truncated;        /* 1 once any update truncated */
  int64_t max_bond_ever;    /* largest bond dimension reached */
  int64_t renormalization_count;
  int64_t center;           /* site of the orthogonality center */
  double local_discarded_weight_sum;    /* sum of dropped s^2 over events */
  double local_discarded_weight_max;
  double relative_discarded_weight_sum; /* dropped s^2 / total s^2 */
  double relative_discarded_weight_max;
} msv_mps_stats;

/* MPSState.__init__ / MPSOptions(dmax, eps, renormalize_splits)
 * (mps_state.py:57-64, 256-286): |0...0> with every bond 1. */
int msv_mps_create(int n_qubits, int dmax, double eps, int renormalize, int device, msv_mps **out);
int msv_mps_destroy(msv_mps *m);
/* MPSState.reset (mps_state.py:288-332). */
int msv_mps_reset(msv_mps *m);
/* MPSState.apply_single (mps_state.py:546-566): u is the 2x2 row-major
 * complex matrix as 8 doubles (re, im). */
int msv_mps_apply_1q(msv_mps *m, int q, const double *u);
/* MPSState.apply_two (mps_state.py:744-768): u is the 4x4 row-major complex
 * matrix (32 doubles) in the basis |q0 q1>; non-adjacent operands are
 * brought together by a swap network and restored after the gate. */
int msv_mps_apply_2q(msv_mps *m, int q0, int q1, const double *u);
/* MPSState.bond_dims (mps_state.py:1119-1120): n - 1 entries. */
int msv_mps_bond_dims(msv_mps *m, int32_t *out);
/* MPSState.truncation_diagnostics counters (mps_state.py:1137-1230). */
int msv_mps_get_stats(msv_mps *m, msv_mps_stats *out);
/* Copies site q (dl x 2 x dr complex64, row-major) to host; host may be null
 * to query the shape. Test and diagnostics path. */
int msv_mps_site_tensor(msv_mps *m, int q, void *host, int32_t *dl, int32_t *dr);
/* MPSState.norm (mps_state.py:1284-1288): the norm of the center tensor. */
int msv_mps_norm(msv_mps *m, double *out);
/* MPSState.expectation_product (mps_state.py:1301-1330): <psi| x_i O_i |psi>
 * / <psi|psi> for 2x2 operators ops (8 doubles each) on distinct wires,
 * by left-to-right transfer matrices on the device. */
int msv_mps_expval_product(msv_mps *m, const int32_t *wires, const double *ops, int k, double *re,
                           double *im);
/* MPSState.to_statevector (mps_state.py:1417-1440): contracts the MPS into
 * dst (a complex64 state of the same width), normalized. */
int msv_mps_to_state(msv_mps *m, msv_state *dst);
int msv_mps_sync(msv_mps *m);

#ifdef __cplusplus
}
#endif

#endif /* METTLEQ_SV_H */

/* hsv.h -- C ABI of the MI355X hierarchical state-vector library (libhsv.so).
 *
 * Plain C: pointers, sizes and status codes only. Every function returns 0 on
 * success or a negative hsv_status; hsv_last_error() gives the message of the
 * last failure on the calling thread. A handle owns its device memory and one
 * HIP stream; calls on one handle are not thread-safe. Host buffers belong to
 * the caller. Amplitudes are interleaved (re, im) complex128 (or complex64),
 * and qubit q is bit q of the amplitude index (qubit 0 = least significant),
 * as in the reference (state_vector.hpp:60-61, 1017-1023).
 *
 * Each entry point names the reference interface it replaces (file:line in
 * the HiSVSIM sources).
 */
#ifndef HSV_H_
#define HSV_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hsv_state hsv_state;

enum hsv_status {
  HSV_OK = 0,
  HSV_ERR_ARG = -1,      /* bad argument (qubit out of range, duplicate, count) */
  HSV_ERR_GATE = -2,     /* unknown gate type or wrong number of targets */
  HSV_ERR_ALLOC = -3,    /* device or host allocation failed */
  HSV_ERR_DEVICE = -4,   /* HIP runtime error */
  HSV_ERR_LIMIT = -5,    /* part too wide for the selected kernel */
  HSV_ERR_UNITARY = -6   /* fused matrix failed its unitarity check */
};

enum hsv_dtype { HSV_C128 = 0, HSV_C64 = 1 };

/* Gate type numbering mirrors SvSim::GateType (basic_gates.hpp:12-45). */
enum hsv_gate_type {
  HSV_G_I = 3, HSV_G_X = 4, HSV_G_Y = 5, HSV_G_Z = 6, HSV_G_H = 7, HSV_G_S = 8,
  HSV_G_T = 9, HSV_G_RX = 10, HSV_G_RY = 11, HSV_G_RZ = 12, HSV_G_RI = 13,
  HSV_G_R1 = 14, HSV_G_RXFRAC = 15, HSV_G_RYFRAC = 16, HSV_G_RZFRAC = 17,
  HSV_G_RIFRAC = 18, HSV_G_R1FRAC = 19, HSV_G_SWAP = 20, HSV_G_C1 = 21,
  HSV_G_C2 = 22, HSV_G_U1 = 23, HSV_G_U2 = 24, HSV_G_U3 = 25, HSV_G_RXX = 26,
  HSV_G_RYY = 27, HSV_G_RZZ = 28, HSV_G_TDG = 29, HSV_G_SDG = 30,
  /* Not in the reference: a caller-supplied dense 2x2 or 4x4 matrix in params
   * (row-major, index = 2*bit(targets[0]) + bit(targets[1])). */
  HSV_G_MATRIX = 64
};

/* One bound gate: the reference's BoundGate (state_vector.hpp:1332-1379)
 * holding an AggregateGateFactory (basic_gates.hpp:1779-1845). params holds
 * the complex parameters interleaved (re, im) in constructor order; the
 * *Frac gates take params[0] = numerator and params[2] = power. At most 5
 * controls and 5 targets (types.hpp:8-9). */
typedef struct hsv_gate {
  int32_t type;
  int32_t num_controls;
  int32_t num_targets;
  int32_t controls[5];
  int32_t targets[5];
  double params[32];
} hsv_gate;

/* Fused-part precision modes. */
enum hsv_fused_mode {
  HSV_FUSED_F64 = 0,        /* FP64 matrix cores (v_mfma_f64)            */
  HSV_FUSED_F32SPLIT = 1,   /* FP32 matrix cores, hi/lo split products   */
  HSV_FUSED_BF16SPLIT = 2,  /* BF16 matrix cores, hi/lo split products
                               accumulated in FP32; parts of 5..14 qubits */
  HSV_FUSED_BF16X3 = 3,     /* BF16 matrix cores, three-piece split (six
                               products, 24 significant bits: FP32-class
                               compensated products) accumulated in FP32;
                               parts whose qubits and qubits 0-4 fit 13 */
  /* hsv_apply_fused_chain only, OR-ed into the mode: end with the state
   * renormalised (as hsv_renormalize, but folded into the layout-restoring
   * swap passes when there are two; a zero state stays as it is). */
  HSV_CHAIN_RENORMALIZE = 0x100
};

/* Replaces StateVector(Qubit, int alloc) (state_vector.hpp:60-61, 74-119,
 * allocate 178-248): 2^n amplitudes in HBM set to |0...0>. The reference's
 * NUMA policy argument has no device meaning and is not taken. */
int hsv_create(int num_qubits, int dtype, hsv_state** out);
/* Frees the state (the reference never frees; state_vector.hpp:173-176). */
int hsv_destroy(hsv_state* s);

int hsv_num_qubits(const hsv_state* s);
int hsv_dtype(const hsv_state* s);

/* Replaces BoundGate::apply -> apply_gate -> StateVector::apply
 * (state_vector.hpp:1349, 1202-1320, 402-510): one pass over the state. */
int hsv_apply_gate(hsv_state* s, const hsv_gate* g);

/* Replaces gather_and_execute_on (py_wrapper.cpp:156 -> execute.hpp:171-316,
 * with the live gather/execute/scatter body of execute.hpp:318-470): all
 * gates of one part applied in one pass, the part's sub-state staged in LDS.
 * opt_slots adds qubits 0..3 to the part (execute.hpp:184-199). */
int hsv_apply_part(hsv_state* s, const hsv_gate* gates, int num_gates, int opt_slots);

/* Not in the reference: the part's gates fused on the host into one
 * 2^k x 2^k FP64 unitary (checked for unitarity) and applied as a batched
 * complex GEMM on the matrix cores. k <= 7. HSV_FUSED_BF16SPLIT pads parts
 * to at least 5 qubits and needs at least 4 state qubits outside the part;
 * it also takes parts of 8..14 qubits, cut on the host into steps of at most
 * 5 qubits (each fused and checked) that one pass applies to a 12- to
 * 14-qubit tile holding qubits 0-4. */
int hsv_apply_fused(hsv_state* s, const hsv_gate* gates, int num_gates, int mode);

/* Not in the reference: a sequence of parts on the BF16-split path
 * (mode HSV_FUSED_BF16SPLIT only) with a moving qubit layout, in the role
 * of the reference's part loop (qasm_assembler_standalone.py:1783-2014, one
 * gather_and_execute_on per part). gates holds the parts one after another,
 * part p being part_sizes[p] gates; lows[5p .. 5p+4] names the 5 qubits that
 * sit in slots 0-4 while part p runs (for p = 0: qubits 0-4, in any order).
 * Each part is one pass of a 12- to 14-qubit tile holding slots 0-4 (so its
 * loads and stores cover 512-byte runs) and the part's qubits; the pass also
 * moves the next part's lows, which must be qubits of part p or of its lows,
 * into slots 0-4. So a part spends no tile slots on qubits it does not use
 * beyond those 5, and parts can be wider in useful qubits than with fixed
 * low qubits. At the end, one tile pass (when the qubits still away from
 * their slots and those in slots 0-4 fit 14) or two in-place swap passes
 * (the layout change as two involutions) restore the standard layout. A
 * tile pass rounds through FP32 like any pass of this path; the swap passes
 * are exact. Argument errors are
 * found before any pass; a part that fails later (a gate on more than 5
 * qubits) leaves the parts before it applied, in the standard layout.
 * Right after hsv_create or hsv_reset (state |0...0>), a qubit that no
 * earlier part has touched is |0> everywhere, so the first passes skip the
 * tiles with a 1 in such a qubit. */
int hsv_apply_fused_chain(hsv_state* s, const hsv_gate* gates, int num_gates, const int* part_sizes,
                          int num_parts, const int* lows, int mode);

/* No counterpart in the reference. For callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc):
 * hsv_apply_fused_chain with the same arguments and results, but calls no
 * hipMalloc, hipFree or allocating library routine. Its device scratch (the
 * renormalisation's reduction partials and one part's fused-step tables and
 * BF16 fragments) is a static 256 KiB array per GPU, held until the process
 * ends; a part that needs more than its 240 KiB of step data (more than 28
 * fused 5-qubit steps; 19 with HSV_FUSED_BF16X3, whose fragments take 12 KB
 * a step) is refused with HSV_ERR_LIMIT before any pass of that part runs. Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (handles on different streams, or threads) race. Not
 * capturable into a HIP graph (it uploads the step data from host memory and
 * its renormalisation reads partials back to the host). */
int hsv_apply_fused_chain_noalloc(hsv_state* s, const hsv_gate* gates, int num_gates, const int* part_sizes,
                                  int num_parts, const int* lows, int mode);

/* Replaces the qubit exchange of SvSimMPI::gather_qubits / gather_slots
 * (svsim-mpi.hpp:123-173, mpi_redistributer.hpp:265-344) on one device:
 * amplitude bit s moves to bit new_slot[s]. Out of place (one pass into a
 * second state-sized buffer, which replaces the first) when that buffer fits
 * in free HBM with 1 GiB to spare and n < 32; otherwise in place, as two
 * tiled swap passes (the permutation as two involutions; no allocation,
 * twice the traffic, needs n >= 10). hsv_tune "permute_in_place" forces
 * either form. Exact in both. */
int hsv_permute_qubits(hsv_state* s, const int* new_slot);

/* Amplitude access (the reference has none beyond dump_state,
 * state_vector.hpp:128-140). count amplitudes from offset, interleaved
 * (re, im) doubles on the host, whatever the state's dtype. */
int hsv_read(hsv_state* s, uint64_t offset, uint64_t count, double* host_out);
int hsv_write(hsv_state* s, uint64_t offset, uint64_t count, const double* host_in);
/* Sets |0...0> again. */
int hsv_reset(hsv_state* s);

/* FP64 reductions. */
int hsv_norm2(hsv_state* s, double* out);
/* Scales the state to unit norm (FP64 norm, one pass); the norm^2 before
 * scaling goes to *norm2_before (may be NULL). Not in the reference (its
 * gates preserve the norm in FP64); the split-precision fused paths drift
 * by ~1e-5 over a circuit and end with this step. A zero state is an error
 * (HSV_ERR_ARG). */
int hsv_renormalize(hsv_state* s, double* norm2_before);
/* <a|b> as (re, im). */
int hsv_inner(hsv_state* a, hsv_state* b, double* out2);
/* max |a_i - b_i|; replaces the amplitude loop of operator== and state_equal
 * (state_vector.hpp:1003-1010, 1053-1072; threshold 1e-10, types.hpp:10).
 * b = NULL gives max |a_i| (the scale for relative error bars). */
int hsv_max_abs_diff(hsv_state* a, hsv_state* b, double* out);
/* Replaces StateVector::probability (state_vector.hpp:897-923). */
int hsv_probability(hsv_state* s, int qubit, int value, double* out);
/* Replaces StateVector::measure (state_vector.hpp:925-981): the same walk
 * over cumulative probabilities for the given uniforms (sorted here).
 * out[i*num_qubits + j] = bit of qubits[j] in sample i. */
int hsv_measure(hsv_state* s, const int* qubits, int num_qubits,
                const double* uniforms, int num_samples, int32_t* out);
/* Replaces StateVector::project (state_vector.hpp:829-895): keeps the
 * amplitudes whose bits under qubit_mask equal `values` (bit q = qubit q),
 * zeroes the others and renormalises by 1/sqrt(p), p = the outcome's
 * probability before the projection, written to *p_out. Deviations: the
 * reference divides by p (:880-882) and removes the projected qubits from the
 * state; here they stay, fixed at their values. p = 0 is an error
 * (HSV_ERR_ARG) and leaves the state unchanged. */
int hsv_project(hsv_state* s, uint64_t qubit_mask, uint64_t values, double* p_out);
/* <psi|P|psi> for the Pauli string P with X on the qubits of x_mask only, Z
 * on z_mask only and Y on both (bit q = qubit q), FP64 accumulation. The
 * reference has no expectation read-out; its tests measure instead. out2 =
 * (re, im); im is rounding noise for a Hermitian P. */
int hsv_expect_pauli(hsv_state* s, uint64_t x_mask, uint64_t z_mask, double* out2);

/* Host-only introspection (no device needed; not in the reference).
 * hsv_gate_matrix: the 2x2 or 4x4 matrix a gate is lowered to, row-major
 * interleaved (re, im) into out (32 doubles), its size in *dim.
 * hsv_fused_unitary: the FP64 unitary hsv_apply_fused would apply for these
 * gates on an n-qubit state: qubits_out (<= 7 entries, sorted) and U row-major
 * interleaved into u_out (2 * 4^7 doubles); returns its unitarity error. */
int hsv_gate_matrix(const hsv_gate* g, double* out, int* dim);
/* Host-only, not in the reference (in the role of its external dagP
 * partitioner): the parts and low-slot sets for hsv_apply_fused_chain.
 * gate_masks[i] = the qubits of gate i as a bit mask. Each part grows by the
 * qubit-set closure from the `low` qubits in the low slots (at most k qubits
 * with them); the next part's lows are the `low` of the `lookahead` qubits of
 * the part's tile whose next gates come first that let the next part absorb
 * the most gates (beam_width = beam_branch = 1: equals
 * hisvsim_amd.partition.chain_parts gate for gate). With a wider beam, each
 * of beam_width chains in flight branches into its beam_branch best low
 * sets and the beam_width children with the fewest gates left go on; the
 * result is the shorter of that and the greedy chain. Out: part_of_gate[i]
 * (parts keep the gate order), lows_out[p*low + j] ascending (room for one
 * part per gate), *num_parts. HSV_ERR_LIMIT: a gate does not fit. */
int hsv_chain_parts(const uint64_t* gate_masks, int num_gates, int k, int low, int lookahead, int candidates,
                    int beam_width, int beam_branch, int32_t* part_of_gate, int32_t* lows_out, int* num_parts);
/* Host-only, not in the reference: the fused steps hsv_apply_fused and
 * hsv_apply_fused_chain cut a wide BF16-split part into (k = 5, width 16).
 * gate_masks[i] = the qubits of gate i (identities left out). Each step is
 * the blocking scan of one k-qubit set (the gates, in order, inside it and
 * behind no skipped gate on a shared qubit); a beam of `width` partial
 * groupings over every k-subset of the part's qubits looks for the fewest
 * steps. Out: group_of_gate[i], *num_groups. HSV_ERR_LIMIT: a gate has more
 * than k qubits, or the gates span so many qubits that there are more than
 * 65536 k-subsets (a wide part has at most 14). */
int hsv_part_steps(const uint64_t* gate_masks, int num_gates, int k, int width, int32_t* group_of_gate,
                   int* num_groups);
int hsv_fused_unitary(int num_qubits, const hsv_gate* gates, int num_gates, int* qubits_out,
                      int* num_part_qubits, double* u_out, double* unitarity_err);

/* Kernel-shape knobs for the autotuner (not in the reference):
 *   "part_kernel" 2 (the default) = FP64 tile kernel for every part whose
 *                 qubits and slots 0-4 fit a 2^12-amplitude tile (2^13 in
 *                 complex64; loads and stores in 512-byte runs, gates in
 *                 registers segment by segment), else as 1;
 *                 1 = register-blocked part kernel for parts of >= 8 qubits,
 *                 0 = LDS kernel for every part;
 *   "tile_bits"   width of the FP64 tile kernel's tile: 0 or 12 (13 in
 *                 complex64, the default) = 512 threads, two workgroups per
 *                 CU; 11 (12 in complex64) = 256 threads, 32 KB of LDS,
 *                 four workgroups per CU (faster per pass, but parts of one
 *                 qubit less); other values are refused;
 *   "reg_bits"    preferred values per thread of the register kernel, as
 *                 log2 (2..5; 0 = default for the dtype);
 *   "reg_waves"   persistent workgroups per CU and occupancy slot (1..64);
 *   "low_slots"   lowest qubits added to every LDS-staged part for
 *                 contiguous HBM runs (0..6);
 *   "io_bits"     low part bits the register kernel keeps on the lanes in
 *                 its first and last layouts, for coalesced loads and
 *                 stores (0..6; 0 = off, the default);
 *   "fused_variant" form of the BF16-split fused kernel: 2 (the default,
 *                 states of >= 12 qubits) = 4096-amplitude tiles loaded and
 *                 stored coalesced through LDS, U rows in registers;
 *                 1 = per-wave column blocks streamed into MFMA operands,
 *                 U in LDS, 4 consecutive rows per column and instruction;
 *                 0 = the same with 8 rows per lane;
 *   "fused_tiles_per_wg" 4096-amplitude tiles each workgroup of the tile
 *                 form takes in turn (1..64, default 8);
 *   "part_tiles_per_wg"  the same for the wide BF16-split part kernel
 *                 (parts of 8..14 qubits; 0..64; 0, the default, = 8
 *                 where the next tile is prefetched during the steps, 1
 *                 where the U fragments are streamed);
 *   "part13_u"    where 13-qubit tiles of the wide part kernel keep the
 *                 steps' U fragments: 0 (the default) = 512 threads, U
 *                 streamed through one LDS buffer, two workgroups per CU;
 *                 1 = 1024 threads, every step's U in LDS, the next tile
 *                 prefetched into registers during the steps (parts of up
 *                 to 11 steps); 2 = as 0 but each wave reads U from L2
 *                 straight into registers (no LDS buffer);
 *   "permute_in_place" hsv_permute_qubits: 0 (the default) = in place from
 *                 32 qubits on or when a second state does not fit, 1 =
 *                 always in place, 2 = always out of place;
 *   "tile_order"  order in which the BF16 tile kernels' workgroups take
 *                 tiles: m = 1..16 = the highest m slots outside the tile
 *                 vary fastest, so concurrent tiles spread over 2^m parts
 *                 of the state (default 4); 0 = natural order. */
int hsv_tune(hsv_state* s, const char* key, int value);

/* ---- Batched QAOA MaxCut (not in the reference) ------------------------
 * Many instances of uneven qubit counts in one batch: instance i has
 * num_qubits[i] qubits, edges[2*e], edges[2*e+1] for e in
 * [edge_offsets[i], edge_offsets[i+1]), and layers[i] QAOA layers. The state
 * is complex64 per instance; the circuit is H on every qubit, then per layer
 * ZZ(gamma) = exp(-i gamma/2 Z_u Z_v) on every edge (the reference's RZZ,
 * basic_gates.hpp:1249-1252) and RX(2 beta) on every qubit. Expectations
 * <C>, C = sum_edges (1 - Z_u Z_v)/2, are FP64 reductions. The batch runs
 * as one hipGraph (captured on first run, replayed after); angles are
 * per instance and layer, flattened in instance order. */
typedef struct hsv_qaoa_batch hsv_qaoa_batch;
int hsv_qaoa_create(int num_instances, const int* num_qubits, const int* edge_offsets, const int* edges,
                    const int* layers, hsv_qaoa_batch** out);
int hsv_qaoa_destroy(hsv_qaoa_batch* b);
int hsv_qaoa_set_angles(hsv_qaoa_batch* b, const double* gammas, const double* betas);
int hsv_qaoa_run(hsv_qaoa_batch* b, double* expectations);
/* Tile shape for the autotuner: "chunk_bits" (8..13), "threads"
 * (128/256/512), "grid_per_cu" (1..64). Invalidates the captured graph. */
int hsv_qaoa_tune(hsv_qaoa_batch* b, const char* key, int value);
/* Device time (ms) of the last hsv_qaoa_run, from HIP events. */
int hsv_qaoa_last_ms(hsv_qaoa_batch* b, double* ms);
int hsv_qaoa_read_state(hsv_qaoa_batch* b, int instance, double* amps);
int hsv_qaoa_total_amplitudes(hsv_qaoa_batch* b, uint64_t* out);

/* ---- Batched small circuits (not in the reference) -----------------------
 * The reference runs one circuit per call (qasm_assembler_standalone.py:
 * 1377-1617 per gate, 1768-2030 per part); here many circuits, or parameter
 * sets of one circuit, of uneven qubit counts run together. Instance i has
 * num_qubits[i] (1..30) qubits and the gates gates[gate_offsets[i] ..
 * gate_offsets[i+1]) in order (the reference's gate numbering, controls and
 * targets as in hsv_apply_gate; at most chunk_bits qubits per gate), and
 * starts from |0...0>. All states share one HBM buffer in `dtype`. Each
 * instance is cut into segments whose qubits fit a 2^chunk_bits-amplitude
 * chunk (an instance of at most chunk_bits qubits is a single segment);
 * stage s applies every instance's s-th segment in one persistent kernel
 * (chunks gathered into LDS, gates applied there), and the stages are
 * captured once into a hipGraph. Bad gates are reported at creation. */
typedef struct hsv_batch hsv_batch;
int hsv_batch_create(int num_instances, const int* num_qubits, const int* gate_offsets, const hsv_gate* gates,
                     int dtype, hsv_batch** out);
int hsv_batch_destroy(hsv_batch* b);
/* New parameters for the same gate structure (types, controls, targets,
 * gate for gate): the matrices are rewritten in device memory and the
 * captured graph replays unchanged. HSV_ERR_ARG if the structure differs. */
int hsv_batch_set_gates(hsv_batch* b, const hsv_gate* gates);
/* Runs every instance from |0...0> (one graph replay; captured on first
 * use and after hsv_batch_tune). */
int hsv_batch_run(hsv_batch* b);
/* <Z_mask> per instance after a run: sum_x |a_x|^2 (-1)^popcount(x & mask),
 * z_masks[i] for instance i (bit q = qubit q), FP64 accumulation. */
int hsv_batch_expect_z(hsv_batch* b, const uint64_t* z_masks, double* out);
/* The 2^num_qubits[instance] amplitudes of an instance, interleaved (re, im)
 * doubles. */
int hsv_batch_read(hsv_batch* b, int instance, double* amps);
/* Shape for the autotuner: "chunk_bits" (6..12 complex128, ..13 complex64;
 * default 12), "threads" (64/128/256/512), "grid_per_cu" (1..64).
 * Invalidates the captured graph. */
int hsv_batch_tune(hsv_batch* b, const char* key, int value);
/* Device time (ms) of the last run, the number of stages, and the
 * amplitudes the stages read and write (each segment with gates moves its
 * instance's state once), for the traffic figure. */
int hsv_batch_info(hsv_batch* b, double* last_ms, int* stages, uint64_t* pass_amplitudes);

/* ---- Cost-model MLP (not in the reference) -------------------------------
 * in -> hidden -> hidden -> 1, ReLU, FP32, trained on the GPU (MSE, Adam) by
 * this library's forward/backward kernels. Parameters are one flat array:
 * W1[in][H] | b1[H] | W2[H][H] | b2[H] | W3[H] | b3. X is row-major
 * [n][in]. Deterministic. */
typedef struct hsv_mlp hsv_mlp;
int hsv_mlp_create(int in_features, int hidden, uint64_t seed, hsv_mlp** out);
int hsv_mlp_destroy(hsv_mlp* m);
int hsv_mlp_num_params(hsv_mlp* m, uint64_t* count);
int hsv_mlp_get_params(hsv_mlp* m, float* out);
int hsv_mlp_set_params(hsv_mlp* m, const float* in);
int hsv_mlp_forward(hsv_mlp* m, const float* X, int n, float* y_out);
/* Loss (mean squared error) and, if grads_out != NULL, all gradients. */
int hsv_mlp_gradient(hsv_mlp* m, const float* X, const float* target, int n, float* loss, float* grads_out);
/* `epochs` passes of Adam over mini-batches of `batch` samples in order. */
int hsv_mlp_train(hsv_mlp* m, const float* X, const float* target, int n, int epochs, int batch, float lr,
                  float* loss);

/* Waits for all work queued on the handle. */
int hsv_synchronize(hsv_state* s);
/* Launch timing: while on, every apply/permute call records a HIP event pair
 * around its kernels on the handle's stream. hsv_kernel_times waits for them,
 * writes up to `capacity` durations (ms, in call order), sets *count to the
 * number recorded, and starts a new collection. */
int hsv_set_timing(hsv_state* s, int on);
int hsv_kernel_times(hsv_state* s, double* ms_out, int capacity, int* count);

/* Replaces the knobs and counters of py_wrapper.cpp:163-175
 * (set_opt_slots execute.hpp:35-37; obtain_* state_vector.hpp:1381-1391,
 * execute.hpp:31-33). Thread counts have no device meaning and are kept
 * only as values the caller can read back. */
void hsv_set_opt_slots(int on);
int hsv_get_opt_slots(void);
void hsv_get_counters(long long* apply_us, long long* gates, long long* parts,
                      long long* part_us);

const char* hsv_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* HSV_H_ */

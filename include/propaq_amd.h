/*
 * propaq_amd.h -- C ABI of the MI355X term-propagation engine.
 *
 * One engine holds one device-resident term sum (sorted SoA u64 key planes
 * plus FP64 coefficients) for one basis (Pauli or Majorana) and one key width
 * W in {1, 2, 4, 8} (n_units <= 32 W, the width ladder of
 * crates/pauli/src/engine.rs:289-299).  It is driven from one host thread.
 *
 * Every entry point returns PQ_OK (0) or an error code; pq_last_error()
 * returns the message of the calling thread's last failure.  Keys cross the
 * boundary as row-major arrays of W u64 words in the reference's interleaved
 * layout (crates/core/src/algebra/strings.rs:30-41): unit u owns bits 2u and
 * 2u+1; Pauli bit 2q = X, 2q+1 = Z; Majorana bit k = gamma_k.
 *
 * The engine has no CPU fallback: without a usable gfx950 device
 * pq_engine_create fails with PQ_ERR_DEVICE.
 */
#ifndef PROPAQ_AMD_H
#define PROPAQ_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PQ_ABI_VERSION 1u

#define PQ_OK 0
#define PQ_ERR_ARG 1      /* invalid argument (ValueError at the Python layer) */
#define PQ_ERR_DEVICE 2   /* HIP runtime failure or no device (RuntimeError) */
#define PQ_ERR_MEMORY 3   /* device allocation failed (MemoryError) */
#define PQ_ERR_STATE 4    /* call not valid in the engine's current state */

#define PQ_BASIS_PAULI 0u    /* BasisKind::Pauli, crates/core/src/algebra/basis.rs:303 */
#define PQ_BASIS_MAJORANA 1u /* BasisKind::Majorana */

typedef struct pq_engine pq_engine;

/* Running totals, the device counterpart of PhaseStats
 * (crates/core/src/engine/partitioned_termsum.rs:80-109) and of the
 * discard statistics (crates/core/src/engine/termsum.rs:225-231). */
typedef struct pq_counters {
  uint64_t visited;        /* anticommuting rows scanned */
  uint64_t emitted;        /* branches emitted, including pair-rule rescues */
  uint64_t declined;       /* branches held back by the coefficient test */
  uint64_t hits;           /* emitted branches that landed on a live key */
  uint64_t created;        /* rows created by branching */
  uint64_t dropped_weight; /* branches refused by the weight bound */
  double discarded_coeff_l1;  /* cumulative |coeff| of held-back, unclaimed branches */
  double discarded_coeff_max; /* largest such |coeff| of the most recent gate */
  uint64_t gates;          /* gates applied (deferred Cliffords included) */
  uint64_t deferred;       /* Clifford gates folded into the frame */
  uint64_t term_gate_updates; /* sum over gates of the live count at gate start */
  /* Which device path applied the gates (no reference counterpart; the
   * reference has one CPU path).  Diagnostics for tests and benchmarks. */
  uint64_t small_gates;    /* gates consumed by small-store runs (one workgroup, rows in LDS) */
  uint64_t chained_gates;  /* branching gates through the sorted chained kernels */
  uint64_t generic_sorts;  /* of those, gates whose children took the generic LDS sort */
  uint64_t sparse_gates;   /* branching gates through the sparse in-place path */
  uint64_t sparse_fallbacks; /* gates sent to the chained kernels because the sparse buffers did not fit */
  uint64_t sparse_pairs;   /* passes that applied two sparse gates (commuting generators) at once */
  uint64_t sparse_syncs;   /* waits for the device before a sparse pass because its buffer bound was not met */
  uint64_t sparse_grows;   /* reallocations of the store, lists and tables for sparse passes */
} pq_counters;

uint32_t pq_abi_version(void);
const char* pq_last_error(void);

/* Number of usable devices (0 when none). */
int pq_device_count(int* out);

/* Replaces PartitionedTermSum::with_inline_positions
 * (crates/core/src/engine/partitioned_termsum.rs:173) and the width dispatch
 * runner_for (crates/pauli/src/engine.rs:289, crates/majorana/src/engine.rs:289).
 * n_units = qubits (Pauli) or sites = modes / 2 (Majorana). */
int pq_engine_create(uint32_t basis_kind, uint32_t n_units, pq_engine** out);
void pq_engine_destroy(pq_engine* e);
int pq_engine_width(const pq_engine* e, uint32_t* n_words);

/* EmitCutoff from a resolved truncation pipeline
 * (crates/core/src/engine/termsum.rs:46-100, policy/truncators.rs:200-211). */
int pq_set_cutoff(pq_engine* e, int has_max_weight, uint32_t max_weight, int has_min_coeff, double min_coeff,
                  int has_min_terms, uint64_t min_terms);

/* A truncation kernel evaluated by the caller: EmitCutoff::term
 * (crates/core/src/engine/termsum.rs:46-58, 129-148), e.g. a NativeTruncator
 * plugin.  The callback has the shape of the plugin ABI's
 * propaq_truncator_keep_batch (crates/core/src/policy/native_truncator.rs:13-33):
 * words holds n_terms keys of n_words u64 (interleaved layout), weights and
 * magnitudes one entry per key, (layer_index, n_layers) the layer context set
 * by pq_set_layer; it fills out_keep (nonzero = keep) and returns 0, anything
 * else fails the engine call.  While a kernel is set (and the term budget does
 * not make the cutoff lossless):
 *  - branches: no coefficient precheck and no weight bound (admits_key is
 *    true, termsum.rs:115-121); every anticommuting row forms its branch and
 *    the kernel decides emit or hold back (admits_child, :129-148).  Gates run
 *    one at a time with one host round trip each;
 *  - seeding (pq_seed) and the reclaim of pq_scale_by_weight /
 *    pq_reclaim(NULL) admit rows by the kernel alone (admits_initial, :152;
 *    reclaim_by_kernel, :387-447).
 * The weight bound still decides whether a Clifford may be deferred (as in
 * the reference).  fn == NULL removes the kernel.  Not for batched engines or
 * while recording. */
typedef int32_t (*pq_keep_batch_fn)(void* ctx, uint32_t basis_kind, const uint64_t* words, size_t n_words,
                                    uint32_t n_units, const uint32_t* weights, const double* magnitudes,
                                    uint32_t layer_index, uint32_t n_layers, uint8_t* out_keep, size_t n_terms);
int pq_set_keep_kernel(pq_engine* e, pq_keep_batch_fn fn, void* ctx);
/* Layer context passed to the keep kernel (LayerContext,
 * crates/pauli/src/engine.rs:118). */
int pq_set_layer(pq_engine* e, uint32_t layer_index, uint32_t n_layers);

/* Small stores (at most 3/4 of the small-store row limit: 8192 rows for keys
 * of one word, 5632 for two, 3584 for four) apply runs of up to 1024 gates in one
 * workgroup, with rows and a key table in LDS, instead of a chain of kernels
 * per gate; the results are the same bits.  On by default; not used while
 * recording, with a keep kernel, a term budget or on a batched engine.  No
 * reference counterpart (an execution choice of this engine). */
int pq_set_small_runs(pq_engine* e, int on);

/* Sparse in-place gates (on by default): a gate over a store too large for a
 * small-store run rewrites its anticommuting rows in place and appends its new
 * rows, instead of merging into a sorted copy; readouts that depend on the row
 * order (expectation, export) sort the store first.  Recording, batched
 * engines and keep kernels always use the sorted chained kernels.  No
 * counterpart in the reference (its store is a hash index). */
int pq_set_sparse(pq_engine* e, int on);

/* At most `bytes` of device memory for the sparse gates' entry lists, slot
 * lists and hash tables (0, the default: whatever the device has free, less
 * a margin).  A gate whose buffers would not fit runs on the sorted chained
 * kernels instead (pq_counters.sparse_fallbacks).  No counterpart in the
 * reference. */
int pq_set_sparse_budget(pq_engine* e, uint64_t bytes);

/* Pair passes (on by default): two consecutive sparse gates whose generators
 * commute (and differ) run as one gather and one join, each row's group of
 * strings taken through both gates in order, so results are those of the
 * gates one at a time, bit for bit (pq_counters.sparse_pairs).  No
 * counterpart in the reference. */
int pq_set_pairs(pq_engine* e, int on);

/* PartitionedTermSum::set_defer_cliffords (partitioned_termsum.rs:666). */
int pq_set_defer_cliffords(pq_engine* e, int on);

/* Seed admission of run_at_width (crates/pauli/src/engine.rs:104-111):
 * terms the full cutoff does not admit are skipped, duplicates are summed in
 * the given order.  Replaces any previous contents and resets the frame. */
int pq_seed(pq_engine* e, const uint64_t* keys, const double* coeffs, uint64_t n);

/* PartitionedTermSum::apply_rotation (partitioned_termsum.rs:276-350): one
 * rotation exp(-i angle G / 2), generator as W words.  *n_terms_after (may be
 * NULL) receives the live count after the gate. */
int pq_apply_rotation(pq_engine* e, const uint64_t* generator, double angle, uint64_t* n_terms_after);

/* A sequence of rotations in application (Heisenberg) order, generators
 * row-major n_gates x W; n_terms_after (may be NULL) receives one live count
 * per gate (PropagationResult.n_terms, crates/core/src/interface/results.rs:8). */
int pq_apply_gates(pq_engine* e, const uint64_t* generators, const double* angles, uint64_t n_gates,
                   uint64_t* n_terms_after);

/* Layer-start noise of run_at_width (crates/pauli/src/engine.rs:119-160):
 * c *= table[min(weight, len - 1)] (PartitionedTermSum::scale_by_weight,
 * partitioned_termsum.rs:478), then, when reclaim != 0, drop rows the full
 * cutoff no longer admits (PartitionedTermSum::reclaim, :581). */
int pq_scale_by_weight(pq_engine* e, const double* table, uint64_t len, int reclaim);

/* Key-aware noise evaluated by the caller (scale_by_key, termsum.rs:340):
 * one factor per row, in the row order of pq_export_raw. */
int pq_scale_by_factors(pq_engine* e, const double* factors, uint64_t n);

/* Reclaim with caller-evaluated keep flags (reclaim_by_kernel,
 * termsum.rs:387), one byte per row in the row order of pq_export_raw;
 * flags == NULL applies the engine's own cutoff (reclaim, :472). */
int pq_reclaim(pq_engine* e, const uint8_t* keep_flags, uint64_t n);

int pq_len(const pq_engine* e, uint64_t* n);

/* Per-row weights in the row order of pq_export_raw (for key-aware hooks). */
int pq_weights(pq_engine* e, uint32_t* out, uint64_t capacity);

/* PartitionedTermSum::expectation (partitioned_termsum.rs:676): sum of
 * c * sign * trace(frame(key)) against a basis state given as bits per unit. */
int pq_expectation(pq_engine* e, const uint64_t* diag_state, uint64_t n_diag_words, double* out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): pq_expectation with its partial sums, the frame image and the
 * total kept in a fixed device array of PQ_NOALLOC_SCRATCH_BYTES per device
 * instead of engine buffers; it calls no hipMalloc, hipFree or other
 * allocating routine.  It sums the rows in store order, with the same
 * partition and order as pq_expectation, so on a sorted store (after
 * pq_expectation or pq_export, or gates through the sorted paths) the values
 * agree bit for bit.  PQ_ERR_ARG when the partials and the frame image need
 * more than the array holds.  Every engine on a device shares that device's
 * array: calls that overlap on one device race. */
#define PQ_NOALLOC_SCRATCH_BYTES 65536
int pq_expectation_noalloc(pq_engine* e, const uint64_t* diag_state, uint64_t n_diag_words, double* out);

/* PartitionedTermSum::terms_below (partitioned_termsum.rs:611). */
int pq_terms_below(pq_engine* e, double threshold, uint64_t* out);

/* PartitionedTermSum::iter (partitioned_termsum.rs:265) with the frame sign
 * folded into the coefficient (crates/pauli/src/engine.rs:410-417), rows in
 * store order; keys row-major n x W.  *n_written receives the row count. */
int pq_export(pq_engine* e, uint64_t* keys, double* coeffs, uint64_t capacity, uint64_t* n_written);

/* Rows as stored, without the frame (for key-aware host hooks). */
int pq_export_raw(pq_engine* e, uint64_t* keys, double* coeffs, uint64_t capacity, uint64_t* n_written);

int pq_frame_is_identity(const pq_engine* e, int* out);
int pq_counters_get(const pq_engine* e, pq_counters* out);
int pq_counters_reset(pq_engine* e);

/* Per-phase device time, measured with HIP events on the engine's own stream
 * while profiling is on.  Phases: 0 branch (count + scan + scatter), 1 child
 * sort (bucket merges or block sort), 2 store merge, 3 noise / reclaim pass,
 * 4 in-place scaling, 5 sparse gather, 6 sparse gate, 7 sparse join.
 * bytes = algorithmic bytes of the phase: for the store merge (N_in + E)
 * (8W + 8) read plus N_out (8W + 8) written; for the noise pass N_in (8W + 8)
 * read plus N_out (8W + 8) written; for the sparse kernels the bytes their
 * design must move, with A the anticommuting rows and w_t the key words the
 * anticommutation test reads (gather: N_in 8 w_t, the rest of the A rows, A
 * list entries, slot indices and table slots written; join: A list entries,
 * table slots and the previous gate's slot indices read, A coefficients and
 * the new rows written); phase 6 counts sampled sparse gates with bytes
 * (N_in + N_out)(8W + 8), one read and one write of the store (its ms stay
 * 0: a gate's device time is phases 5 + 7); 0 elsewhere.
 * Events are recorded on sampled gates only: on = k > 0 samples every k-th
 * gate (1 = every gate), so unsampled gates run without event markers (each
 * marker idles the queue for several microseconds); launches, ms and bytes
 * cover the sampled launches.  Noise passes are always sampled.  on = 0 turns
 * profiling off. */
#define PQ_N_PHASES 8
typedef struct pq_profile {
  uint64_t launches[PQ_N_PHASES];
  double ms[PQ_N_PHASES];
  double bytes[PQ_N_PHASES];
  /* Host side, every gate while profiling is on: wall time spent enqueueing
   * gates (launch calls, frame conjugation, buffer growth) and waiting for
   * the device to publish row counts. */
  uint64_t host_gates;
  double host_enqueue_ms;
  double host_wait_ms;
} pq_profile;

int pq_profile_enable(pq_engine* e, int on);
int pq_profile_get(const pq_engine* e, pq_profile* out);

/* ---- Differentiation.  The reference has no gradient API (its nearest
 * equivalent is parameter shift: two forward runs at theta +- pi/2 per angle
 * through PauliPropagator.expectation_value, crates/pauli/src/propagator.rs:270).
 * These entry points give d<O>/dtheta of every rotation exp(-i theta G / 2) by
 * replaying a recorded forward run backwards, with every truncation decision
 * of the forward run held fixed.
 *
 * pq_set_differentiable(e, 1) makes every rotation run through the branch
 * path: Cliffords are not deferred into the frame and gates with
 * |sin theta| < 1e-9 form their branches (of value ~0) instead of taking the
 * phase-only shortcut (coeff.rs:200), so each gate's derivative is present
 * at every angle.  A Clifford the frame would have absorbed branches with no
 * truncation predicate, so it loses nothing the frame would have kept.
 * Values then agree with the unrecorded run apart from rows of coefficient
 * ~cos(theta) c ~ 1e-17 c that such a Clifford leaves behind.
 *
 * pq_tape_begin (differentiable mode only) clears any previous tape and
 * records every later gate and noise / reclaim pass as a gather tape in
 * device memory (8 bytes per input row for the slot maps plus 8 for the
 * coefficients at the gate's start); while recording, gates run one at a
 * time.  pq_tape_end stops recording. */
int pq_set_differentiable(pq_engine* e, int on);
int pq_tape_begin(pq_engine* e);
int pq_tape_end(pq_engine* e);
int pq_tape_info(const pq_engine* e, uint64_t* n_gates, uint64_t* n_bytes);

/* Device copies of the current rows and frame, for recomputation of the
 * forward run segment by segment (checkpoint every k layers). */
int pq_checkpoint_save(pq_engine* e, uint32_t* id);
int pq_checkpoint_restore(pq_engine* e, uint32_t id);
int pq_checkpoint_free(pq_engine* e, uint32_t id);

/* Sets the adjoint to d<O>/dc of every current row: the readout weight
 * sign * trace(frame(key)) of PartitionedTermSum::expectation
 * (partitioned_termsum.rs:676) against the basis state. */
int pq_adjoint_seed(pq_engine* e, const uint64_t* diag_state, uint64_t n_diag_words);

/* Consumes the tape, last entry first.  The tape must end at the rows the
 * adjoint is defined on (the rows of pq_adjoint_seed, or the rows a previous
 * sweep stopped at).  grads receives one d<O>/dtheta per recorded gate in
 * application order (n_gates must equal the recorded count); the adjoint is
 * left on the rows the tape started from. */
int pq_adjoint_backward(pq_engine* e, double* grads, uint64_t n_gates);

/* The adjoint d<O>/dc per row, in the row order of pq_export_raw at the
 * point the last sweep stopped (for gradients by the seed coefficients).
 * out == NULL only reports the row count in *n_written. */
int pq_adjoint_values(pq_engine* e, double* out, uint64_t capacity, uint64_t* n_written);

/* ---- Batching.  The reference runs parameter sets one after another (its
 * only batching API, SurrogateModel::evaluate_batch,
 * crates/surrogate/src/model.rs:80-102, evaluates a symbolic tape instead).
 * A batched engine holds n_instances independent term sums of one circuit
 * structure in one sorted store: instance b's rows carry b in the top bits of
 * key word W-1 (above every unit bit; W is chosen to leave room), so every
 * gate runs once for all instances, with per-instance angles.  Each instance
 * evolves exactly as it would alone (same keys, same coefficients, same
 * expectation summation order), except that a Clifford angle is folded into
 * the frame only when every live instance has that same angle.  Term budgets
 * (min_terms) are per run and are refused here. */
int pq_engine_create_batched(uint32_t basis_kind, uint32_t n_units, uint32_t n_instances, pq_engine** out);
/* n_instances, and the shift of the instance id within key word W-1. */
int pq_batch_info(const pq_engine* e, uint32_t* n_instances, uint32_t* instance_shift);
/* Like pq_seed; row i belongs to instance instances[i]. */
int pq_seed_batched(pq_engine* e, const uint32_t* instances, const uint64_t* keys, const double* coeffs, uint64_t n);
/* Like pq_apply_gates with one generator per gate shared by all instances
 * and angles row-major n_gates x n_instances; n_terms_after gets the total
 * live count after each gate. */
int pq_apply_gates_batched(pq_engine* e, const uint64_t* generators, const double* angles, uint64_t n_gates,
                           uint64_t* n_terms_after);
/* One expectation per instance; diag_states is n_instances x n_diag_words. */
int pq_expectation_batched(pq_engine* e, const uint64_t* diag_states, uint64_t n_diag_words, double* out);
/* Live rows per instance. */
int pq_instance_counts(pq_engine* e, uint64_t* counts);
/* Removes every row of the instances whose flag is set (finished runs);
 * later gates ignore their angles.  Refused while recording (a recorded
 * batch freezes finished instances instead, see pq_freeze_instances). */
int pq_retire_instances(pq_engine* e, const uint8_t* flags);

/* pq_expectation_batched and the row counts for the flagged instances only
 * (values[b], counts[b] written where flags[b] != 0), then
 * pq_retire_instances(flags): one call for the instances a batch has
 * finished.  On a store the sparse passes left unsorted only the flagged
 * instances' rows are sorted for the readout.  No counterpart in the
 * reference (batching). */
int pq_readout_retire(pq_engine* e, const uint64_t* diag_states, uint64_t n_diag_words, const uint8_t* flags,
                      double* values, uint64_t* counts);
/* Sets the frozen instances to the flagged ones: later gates leave their
 * rows unchanged (they count as commuting rows, in a recorded tape too), so a
 * recorded batch keeps instances that have finished their own layers.
 * Seeding unfreezes all. */
int pq_freeze_instances(pq_engine* e, const uint8_t* flags);

/* Batched differentiation (no reference counterpart, as above).  A batched
 * engine records like an unbatched one (pq_set_differentiable, pq_tape_begin);
 * each recorded gate keeps its per-instance angles and the instances' row
 * ranges.  pq_adjoint_seed_batched seeds every row against its instance's
 * basis state (diag_states n_instances x n_diag_words).
 * pq_adjoint_backward_batched consumes the tape like pq_adjoint_backward and
 * writes d<O_b>/dtheta_{g,b}, gate-major: grads[g * n_instances + b].  Every
 * instance's rows are reversed in the blocks a run of that instance alone
 * would use, so its gradient equals that run's bit for bit. */
int pq_adjoint_seed_batched(pq_engine* e, const uint64_t* diag_states, uint64_t n_diag_words);
int pq_adjoint_backward_batched(pq_engine* e, double* grads, uint64_t n_gates);

/* Bytes of device memory held by the store (keys + coefficients). */
int pq_key_bytes(const pq_engine* e, uint64_t* out);

/* Blocks until all device work of the engine is done. */
int pq_synchronize(pq_engine* e);

/* ---- Symbolic surrogate (crates/surrogate) ----
 *
 * A surrogate build propagates the observable once with symbolic
 * coefficients (a shared expression DAG of cos/sin of gate parameters) and
 * compiles the surviving terms into one tape; the tape is then evaluated for
 * any number of parameter sets on the device.  The builder is host code, as
 * the reference's is; the model lives in HBM.  A builder or a model is driven
 * from one host thread at a time (evaluation reuses the model's buffers). */
typedef struct pq_surrogate pq_surrogate;
typedef struct pq_model pq_model;

/* SurrogatePropagator::build's engine (crates/surrogate/src/engine.rs:137-235)
 * over n_units qubits (Pauli) or sites (Majorana), n_units <= 256. */
int pq_surrogate_create(uint32_t basis_kind, uint32_t n_units, pq_surrogate** out);
void pq_surrogate_destroy(pq_surrogate* s);
int pq_surrogate_width(const pq_surrogate* s, uint32_t* n_words);

/* The resolved truncation config (truncators.rs:187-211 ResolvedConfig:
 * frequency, coefficient, weight, min_terms, simplify), applied after every
 * gate as apply_truncation does (engine.rs:20-57). */
int pq_surrogate_set_truncation(pq_surrogate* s, int has_frequency, uint32_t frequency, int has_coefficient,
                                double coefficient, int has_weight, uint32_t weight, int has_min_terms,
                                uint64_t min_terms, int simplify);

/* Coefficient nodes no live term reaches are dropped once the node arena has
 * doubled since the last sweep, holds at least min_nodes nodes (default
 * 2^20) and a truncation pass has dropped something.  Results do not depend
 * on it.  No reference counterpart (the reference frees nodes by reference
 * counting). */
int pq_surrogate_set_compaction(pq_surrogate* s, uint64_t min_nodes);

/* Adds the observable's terms, each a scalar leaf (engine.rs:171-178). */
int pq_surrogate_seed(pq_surrogate* s, const uint64_t* keys, const double* coeffs, uint64_t n);

/* Applies n_gates rotations in the given (Heisenberg) order: gate g is
 * symbolic in parameter param_index[g] when that is >= 0, else numeric at
 * angles[g] (GateParam, symcoeff.rs:709-777; PartitionedTermSum::
 * apply_rotation, partitioned_termsum.rs:276-350).  n_terms_after (optional,
 * n_gates entries) receives the live count after each gate, monomials_after
 * (optional, 2 n_gates entries, low then high 64 bits) the monomial count. */
int pq_surrogate_apply(pq_surrogate* s, const uint64_t* generators, const int64_t* param_index, const double* angles,
                       uint64_t n_gates, uint64_t* n_terms_after, uint64_t* monomials_after);

/* Per-gate record of a logged build (engine.rs GateRecord / MergeRecord,
 * crates/surrogate/src/engine.rs:110-135): live terms after the rotation and
 * after the truncation pass, the monomial count after it (low, high 64
 * bits), and the wall time of each step. */
typedef struct pq_surrogate_gate_stats {
  uint64_t terms_rotated;
  uint64_t terms;
  uint64_t monomials_lo, monomials_hi;
  double rotation_ms, truncation_ms;
} pq_surrogate_gate_stats;

/* pq_surrogate_apply that also fills stats[n_gates] (for the verbose log,
 * crates/surrogate/src/propagator.rs:94-160). */
int pq_surrogate_apply_logged(pq_surrogate* s, const uint64_t* generators, const int64_t* param_index,
                              const double* angles, uint64_t n_gates, pq_surrogate_gate_stats* stats);

/* Totals since creation for the closing engine_phases record: rows visited
 * by rotations (anticommuting rows), branches emitted, branches landing on a
 * live term, and the seconds spent rotating and truncating; nodes receives
 * the coefficient nodes the builder holds now.  Any output may be NULL. */
int pq_surrogate_counters(const pq_surrogate* s, uint64_t* visited, uint64_t* emitted, uint64_t* hits,
                          double* rotation_s, double* truncation_s, uint64_t* nodes);

/* Live terms of the builder. */
int pq_surrogate_len(const pq_surrogate* s, uint64_t* n);

/* compile (engine.rs:60-107): terms in store order whose overlap with the
 * basis state diag_state is nonzero, one shared tape, uploaded to the device.
 * The builder stays usable. */
int pq_surrogate_compile(pq_surrogate* s, const uint64_t* diag_state, uint64_t n_diag_words, uint32_t n_params,
                         pq_model** out);

/* A model from a tape (SurrogateModel::load, model.rs:183-262): op i has
 * kind 0..4 = Scalar, Add, Scale, Cos, Sin (symcoeff.rs:748-755), operands
 * a[i], b[i] < i, param[i] for Cos / Sin, val[i] the scalar or factor; term t
 * has overlaps[t] and root op roots[t] (UINT64_MAX: empty). */
int pq_model_create(uint32_t n_params, uint64_t n_ops, const uint8_t* kinds, const uint32_t* params, const uint64_t* a,
                    const uint64_t* b, const double* vals, uint64_t n_terms, const double* overlaps,
                    const uint64_t* roots, pq_model** out);
void pq_model_destroy(pq_model* m);

/* n_terms, n_ops, n_levels, the monomial count (low, high 64 bits,
 * model.rs:113-120) and n_params. */
int pq_model_info(const pq_model* m, uint64_t* n_terms, uint64_t* n_ops, uint64_t* n_levels, uint64_t* monomials_lo,
                  uint64_t* monomials_hi, uint32_t* n_params);

/* The tape in topological order, for saving (model.rs:124-172): arrays of
 * n_ops and n_terms entries. */
int pq_model_export(const pq_model* m, uint8_t* kinds, uint32_t* params, uint64_t* a, uint64_t* b, double* vals,
                    double* overlaps, uint64_t* roots);

/* evaluate_batch (model.rs:80-103): params is n_sets rows of stride doubles
 * (stride >= n_params), out n_sets values.  Term contributions
 * overlap * value are summed pairwise in store order. */
int pq_model_evaluate(pq_model* m, const double* params, uint64_t n_sets, uint64_t stride, double* out);

/* At most max_sets parameter sets per device pass (0, the default: as many
 * as a quarter of free device memory holds); bounds the value table, n_ops x
 * max_sets doubles.  Results do not depend on it.  No reference counterpart. */
int pq_model_set_batch_limit(pq_model* m, uint64_t max_sets);

/* Device time of the last pq_model_evaluate (no reference counterpart: the
 * measurement hook): ms in the tape kernels and in the term sum, the
 * algorithmic HBM bytes of the tape kernels (8 per value written, 8 per
 * operand value read) and the number of kernel launches.  Timing is taken
 * only after pq_model_set_profiling(m, 1). */
int pq_model_set_profiling(pq_model* m, int on);
int pq_model_timing(const pq_model* m, double* tape_ms, double* sum_ms, double* tape_bytes, uint64_t* launches);

#ifdef __cplusplus
}
#endif

#endif /* PROPAQ_AMD_H */

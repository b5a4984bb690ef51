// pq_kernels.h -- launch interface of the gfx950 kernels (host side).
//
// Device layout of a term store (SoA): W key planes of u64, plane i at
// key + i * cap, plus one FP64 coefficient column.  Rows are kept sorted by
// key (word W-1 most significant), unique, so duplicates merge by a linear
// merge instead of a hash probe.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pq_algebra.h"

namespace pq {

constexpr int kBlock = 256;       // threads per block (4 wave64s)
// Tile parameters are compile-time; propaq_amd/autotune.py measures builds
// with other values (-D PQ_BRANCH_ITEMS=... etc.) and records the fastest in
// propaq_amd/tuned.json, which the build recipe passes on.
#ifndef PQ_BRANCH_ITEMS
#define PQ_BRANCH_ITEMS 4
#endif
#ifndef PQ_MERGE_ITEMS
#define PQ_MERGE_ITEMS 4
#endif
constexpr int kItems = PQ_BRANCH_ITEMS;  // store rows per thread in the branch kernels
constexpr int kTile = kBlock * kItems;
constexpr int kMaxClassBits = 8;  // generator bits used to bucket children

// Merge tiles hold W key words + value + flag per element in LDS; wide keys
// use half tiles so a block stays within 64 KiB of static LDS.
template <int W>
struct MergeCfg {
  static constexpr int items = W <= 4 ? PQ_MERGE_ITEMS : 2;
  static constexpr int tile = kBlock * items;
};
// Store-merge tiles (k_store_splits / k_store_merge): outputs per thread.
#ifndef PQ_STORE_ITEMS
#define PQ_STORE_ITEMS 4
#endif
// narrower keys take more outputs per thread for the same LDS per tile
// (measured on the Majorana UCJ (W = 1) and HEA (W = 2) workloads: 8 and 6
// against 4 gave 4 % and 16 % on the whole run)
#ifndef PQ_STORE_ITEMS_W1
#define PQ_STORE_ITEMS_W1 8
#endif
#ifndef PQ_STORE_ITEMS_W2
#define PQ_STORE_ITEMS_W2 6
#endif
template <int W>
struct StoreCfg {
  static constexpr int items =
      W == 1 ? PQ_STORE_ITEMS_W1 : (W == 2 ? PQ_STORE_ITEMS_W2 : (W <= 4 ? PQ_STORE_ITEMS : 2));
  static constexpr int tile = kBlock * items;
};
constexpr int kMaxClasses = 1 << kMaxClassBits;
// Small-store runs: one workgroup of kSmallThreads keeps up to rows live rows
// in LDS (W words of key, the coefficient, 1.5 16-bit table slots and a bit
// each, within the 160 KiB of a CU) for a whole run of gates; 0 = no
// small-store path at this width.
constexpr int kSmallThreads = 1024;
template <int W>
struct SmallCfg {
  static constexpr int rows = W == 1 ? 8192 : (W == 2 ? 5632 : (W == 4 ? 3584 : 0));
};

enum ChildFlag : uint8_t { kEmit = 1, kPending = 2 };

// Sparse in-place gate: rows per block of the gather, and list entries per
// tile of the join.
#ifndef PQ_SPARSE_ITEMS
#define PQ_SPARSE_ITEMS 8
#endif
constexpr int kSparseItems = PQ_SPARSE_ITEMS;
constexpr int kSparseTile = kBlock * kSparseItems;
#ifndef PQ_JOIN_ITEMS
#define PQ_JOIN_ITEMS 4
#endif
constexpr int kJoinItems = PQ_JOIN_ITEMS;  // list entries per thread and tile pass in the join
constexpr int kJoinTile = kBlock * kJoinItems;
// One anticommuting row in the sparse gate's list: W key words, the row's
// coefficient c, its branch value v (phase applied), and meta = source row
// (bits 0-31) | decision (bits 32-33, kDec*).  Padded to 32, 64 or 128
// bytes so an entry never straddles more cache lines than it must.
template <int W>
struct alignas(8 * (W + 3 <= 4 ? 4 : (W + 3 <= 8 ? 8 : 16))) SparseEnt {
  static constexpr int kWords = W + 3 <= 4 ? 4 : (W + 3 <= 8 ? 8 : 16);
  uint64_t w[kWords];
};
// Hash-table slots for a gate over n rows: a power of two >= 2n, so the
// open-addressed table is at most half full even if every row anticommutes.
PQ_HD uint64_t sparse_table_slots(uint64_t n) {
  uint64_t t = 64;
  while (t < 2 * n) t <<= 1;
  return t;
}
// list entries keep the source row, and the slot lists the table slot, in 32 bits
constexpr uint64_t kSparseMaxRows = 1ULL << 30;

// A sparse pass's list is split into up to kSubLists sub-lists so that the
// gather blocks reserve their entries with atomics on different words:
// atomics on one device word are serialized, about 12 ns apiece on MI355X,
// which one counter for thousands of blocks turns into tens of microseconds.
// Over n rows the gather has nb = ceil(n / kSparseTile) blocks; block b
// appends to sub-list b mod S (S = min(nb, kSubLists)), whose region of the
// entry array is [q R, q R + R) with R = ceil(nb / S) kSparseTile, room for
// every row of its blocks.  A pass's record holds the S counters (each on its
// own 256-byte line) and, in its header, S and R for the join that clears the
// pass's slots later.
constexpr int kSubLists = 64;
constexpr int kRecStride = 32;                          // words between counters
constexpr int kRecWords = kSubLists * kRecStride + 32;  // counters, then S and R
// Per pass: rows the first gate of a pair made, summed by the join blocks into
// kMidSlots words kMidStride apart.
constexpr int kMidSlots = 16;
constexpr int kMidStride = 8;
constexpr int kMidWords = kMidSlots * kMidStride;
// Device words kept between sparse passes: three records (by pass number mod
// 3: this pass, the previous one to clear, the next one to zero) and the join
// word on a line of its own.
constexpr int kSparseSyncWords = 3 * kRecWords + 32;
struct SubLists {
  uint64_t S, R;
};
PQ_HD SubLists sparse_sublists(uint64_t n) {
  uint64_t nb = (n + kSparseTile - 1) / kSparseTile;
  if (nb == 0) nb = 1;
  const uint64_t S = nb < (uint64_t)kSubLists ? nb : (uint64_t)kSubLists;
  return SubLists{S, (nb + S - 1) / S * (uint64_t)kSparseTile};
}
// Entries a list over n rows may need (its sub-list regions).
PQ_HD uint64_t sparse_ent_need(uint64_t n) {
  const SubLists l = sparse_sublists(n);
  return l.S * l.R;
}

struct StoreDev {
  uint64_t* key;    // W planes of cap words
  double* coeff;    // cap
  uint64_t cap;
};

struct ChildDev {
  uint64_t* key;    // W planes of cap words
  double* val;      // sine branch with its phase applied
  uint8_t* flag;    // kEmit or kPending
  uint32_t* origin; // while recording: source row | kRecSign if -Im(phase) < 0; else null
  uint64_t cap;
};

// Gather tape of one recorded step (differentiation).  For a gate, per input
// row r: out[r] = the row's output slot, child[r] = kRecCommute, kRecNoChild
// (anticommuting, no branch reached the output) or the output slot its branch
// landed in, with kRecSign set when -Im(phase) = -1.  For a noise pass:
// out[r] = output slot or kRecNoChild (reclaimed), factor[r] = the damping.
constexpr uint32_t kRecSign = 0x80000000u;
constexpr uint32_t kRecNoChild = 0x7FFFFFFFu;
constexpr uint32_t kRecCommute = 0x7FFFFFFEu;
constexpr uint64_t kRecMaxRows = 0x7FFFFFFDu;

struct TapeDev {
  uint32_t* out;
  uint32_t* child;
  double* factor;  // noise passes; null: every factor is 1 (a permutation of the rows)
};

// A recorded sparse gate's tape, one record per list position (the rows that
// anticommute with the gate; every other row keeps its place and value): the
// store row (it keeps its place), where its branch landed (a row | kRecSign,
// or kRecNoChild) and its coefficient at the gate's start.
struct SparseRec {
  uint32_t row, child;
  double c;
};

// Per-gate constants, passed by value.
template <int W>
struct GateArgs {
  Key<W> gen;    // generator after the deferred frame
  Key<W> fold;   // Pauli: pair-swapped generator; Majorana: the generator
  Key<W> above;  // Majorana: bit b set iff odd count of generator bits above b
  int gen_y, gen_len, neg, n_units;
  double sin_t, cos_t, min_coeff;
  uint32_t max_weight;
  int has_min, has_maxw;
  int n_cls_bits;
  int cls_pos[kMaxClassBits];
  // Batched engines: the instance id sits in the top bits of key word W-1
  // (row >> inst_shift); unit_mask keeps the unit bits of that word for
  // weights; sc (per instance: sin, cos, phase-only flag) replaces
  // sin_t / cos_t.  Unbatched: unit_mask = ~0, sc = null.
  uint64_t unit_mask;
  const double* sc;
  int inst_shift;
  // A caller-side keep kernel decided this gate's branches (no precheck or
  // weight bound applied): own_keep[r] = 1 when store row r's own branch was
  // emitted, which the pair rule's claims read.  Null otherwise.
  const uint8_t* own_keep;
};

// Per-layer noise / reclaim epilogue.
struct NoiseArgs {
  const double* table;  // device, len entries (factor by weight), or null
  uint32_t len;
  const double* factors;     // device, one factor per row, or null
  const uint8_t* keep_flags; // device, one keep byte per row, or null
  int reclaim;          // apply the admits_initial filter afterwards
  int has_min, has_maxw;
  double min_coeff;
  uint32_t max_weight;
  uint64_t unit_mask;        // as in GateArgs
  const uint8_t* inst_drop;  // batched: drop every row of flagged instances, or null
  int inst_shift;
};

// Device counters, accumulated by the kernels into kCtrSlots slots of one
// L2 line each (block b adds into slot b % kCtrSlots), so thousands of waves
// finishing together do not serialize their atomics on one line.  The host
// sums the slots (max for disc_max_bits).
struct alignas(128) Counters {
  unsigned long long visited, emitted, declined, hits, created, dropped_weight;
  unsigned long long disc_max_bits;  // bit pattern of the largest discarded |v|
  double disc_l1;
};
constexpr int kCtrSlots = 256;
__device__ __forceinline__ Counters* ctr_slot(Counters* ctr) { return ctr + (blockIdx.x & (kCtrSlots - 1)); }

// Gate kernels read the live row count from the device (n_dev); n_max /
// e_max are host-side upper bounds used only to size the grids.
template <int KIND, int W>
struct Launch {
  static hipError_t branch_count(const GateArgs<W>& g, StoreDev s, const uint64_t* n_dev, uint64_t nb,
                                 uint32_t* counts, int m, Counters* ctr, hipStream_t st);
  // tape.child (may be null) receives kRecCommute / kRecNoChild per row; the
  // store merge fills in the slots of branches that reach the output.
  static hipError_t branch_scatter(const GateArgs<W>& g, StoreDev s, const uint64_t* n_dev, uint64_t nb,
                                   const uint64_t* offs, int m, ChildDev c, uint32_t* tape_child, hipStream_t st);
  // Sparse in-place gate (gather, join; see pq_kernels.hip): the
  // anticommuting rows are rewritten in place and new rows appended at
  // [n, n_out); the output count goes to n_out (device) and, with the list
  // length, to the host-mapped n_out_host / e_host.  Gate number seq uses list,
  // slot list and table seq mod 2 and clears the other table's slots of gate
  // seq - 1.
  // sync_words: kSparseSyncWords device words kept between gates (three
  // sub-list records used by gate number mod 3, the join word), zero, like
  // both tables, before the engine's first sparse gate.  The lists hold
  // sparse_ent_need(n_max) entries.  Events, when given, stamp the gather
  // (g0, g1) and join (j0, j1) kernels' own dispatches.  Recording: tape gets
  // one record per list position (n_max of them at most) and, on batched
  // engines, tape_inst the instance of each (else null).
  static hipError_t sparse_gate(const GateArgs<W>& g, StoreDev s, const uint64_t* n_dev, uint64_t n_max,
                                SparseEnt<W>* const ents[2], uint32_t* const slots[2], uint64_t ent_cap,
                                uint64_t* const tables[2],
                                uint64_t table_cap, uint64_t* sync_words, uint64_t seq, uint64_t* n_out,
                                uint64_t* n_out_host, uint64_t* e_host, uint32_t pub_tag, Counters* ctr,
                                SparseRec* tape, uint32_t* tape_inst, hipStream_t st, hipEvent_t g0 = nullptr,
                                hipEvent_t g1 = nullptr, hipEvent_t j0 = nullptr, hipEvent_t j1 = nullptr);
  // Two sparse gates A then B with commuting, different generators in one
  // gather and one join (see pq_kernels.hip), sharing sparse_gate's lists,
  // tables and sync words (a pass takes one gate number, seq).  The store
  // must hold 4 n_max rows.  The count after B goes to n_out and n_out_host;
  // the rows gate A made are added to mid's kMidSlots words (kMidStride
  // apart), which the gather zeroes.
  static hipError_t pair_gate(const GateArgs<W>& ga, const GateArgs<W>& gb, StoreDev s, const uint64_t* n_dev,
                              uint64_t n_max, SparseEnt<W>* const ents[2], uint32_t* const slots[2], uint64_t ent_cap,
                              uint64_t* const tables[2], uint64_t table_cap, uint64_t* sync_words, uint64_t seq,
                              uint64_t* mid, uint64_t* n_out, uint64_t* n_out_host, uint64_t* e_host,
                              uint32_t pub_tag, Counters* ctr, hipStream_t st, hipEvent_t g0 = nullptr,
                              hipEvent_t g1 = nullptr, hipEvent_t j0 = nullptr, hipEvent_t j1 = nullptr);
  static hipError_t scale_anticommuting(const GateArgs<W>& g, StoreDev s, const uint64_t* n_dev, uint64_t n_max,
                                        double f, hipStream_t st);
  static hipError_t store_merge(const GateArgs<W>& g, StoreDev s, const uint64_t* n_dev, uint64_t n_max,
                                ChildDev c, const uint64_t* e_dev, uint64_t e_max, StoreDev out, uint64_t* splits,
                                uint64_t* status, uint32_t* ticket, uint32_t* overflow, uint64_t* n_out,
                                uint64_t* n_out_host, uint64_t* e_host, uint32_t tag, Counters* ctr, TapeDev tape,
                                hipStream_t st, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
  // A run of gates over a store of at most SmallCfg<W>::rows rows in one
  // workgroup (kinds[j]: 0 = branch, 1 = phase-only).  Stops before a gate
  // whose new rows might not fit: *done = gates applied, n_after[j] = live
  // rows after gate j; the store is left sorted with *n_dev rows.
  static hipError_t small_run(const GateArgs<W>* gates, const uint8_t* kinds, int n_gates, StoreDev s,
                              uint64_t* n_dev, uint64_t* n_after, uint32_t* done, Counters* ctr, hipStream_t st);
  // Noise / reclaim over the n rows of s, kept rows compacted in order into
  // out, which may be s itself (each tile loads its rows before it publishes
  // its count, and writes only below its first row once every earlier tile
  // has published).
  // The count goes to n_out and, given n_host, as (tag << 48) | count to
  // that host-mapped word.  status holds ceil(n / kTile) + 1 words: the
  // tiles' look-back words and, last, the tile ticket (one clear for both).
  static hipError_t noise_compact(const NoiseArgs& a, StoreDev s, uint64_t n, StoreDev out, uint64_t* status,
                                  uint64_t* n_out, int n_units, TapeDev tape, hipStream_t st,
                                  uint64_t* n_host = nullptr, uint32_t tag = 0);
  // Readout cotangent of every row: d<O>/dc = sign * trace(frame(key)).
  static hipError_t adjoint_seed(StoreDev s, uint64_t n, Key<W> diag_mask, int n_units, const uint64_t* frame_img,
                                 const double* frame_sign, int n_rows, double* lam, hipStream_t st);
  // Batched engines: each row's instance (key word W-1 >> shift) picks its
  // basis-state mask from dms (W words per instance).
  static hipError_t adjoint_seed_batched(StoreDev s, uint64_t n, const uint64_t* dms, int shift, uint64_t unit_mask,
                                         int n_units, const uint64_t* frame_img, const double* frame_sign, int n_rows,
                                         double* lam, hipStream_t st);
  static hipError_t expectation(StoreDev s, uint64_t n, Key<W> diag_mask, int n_units, const uint64_t* frame_img,
                                const double* frame_sign, int n_rows, double* partials, int n_blocks,
                                hipStream_t st, uint64_t unit_mask = ~0ULL);
  // Batched expectation: work item i covers blocks of instance inst[i]
  // (rows [lo, lo + len)), block item blk[i] of nblk[i]; each writes one
  // partial, in the partition k_expectation would use for that instance
  // alone.  diag_masks holds one Key<W> per instance.
  static hipError_t expectation_segments(StoreDev s, const uint32_t* work, uint64_t n_work, const uint64_t* bounds,
                                         const uint64_t* diag_masks, int n_units, const uint64_t* frame_img,
                                         const double* frame_sign, int n_rows, double* partials,
                                         uint64_t unit_mask, hipStream_t st);
  static hipError_t weights(StoreDev s, uint64_t n, int n_units, uint32_t* out, hipStream_t st,
                            uint64_t unit_mask = ~0ULL);
  static hipError_t conjugate(StoreDev s, uint64_t n, const uint64_t* frame_img, const double* frame_sign,
                              int n_rows, uint64_t* out_keys_rowmajor, double* out_coeff, hipStream_t st);
};

// log2(m) rounds of pairwise merges of the class buckets, whose bounds are
// offs[c * nb] on the device (c = 0..m); e_max bounds the child count.  The
// merged run ends in b when *result_in_b is set, else in a.
template <int W>
hipError_t launch_merge_rounds(ChildDev a, ChildDev b, const uint64_t* offs, uint64_t nb, int m, uint64_t e_max,
                               int* result_in_b, hipStream_t st);

// Generic child sort for generators with too many bits to bucket: LDS
// block sort of every tile, then regular pairwise merge rounds.  The sorted
// run ends in b when *result_in_b is set, else in a.
template <int W>
hipError_t launch_sort_children(ChildDev a, ChildDev b, uint64_t e, int* result_in_b, hipStream_t st);

// Batched recording: the adjoint work list (instance, block, blocks) of a
// gate from its instances' row ranges, per-instance item offsets off[0..n_inst]
// and the item count (also written to *total, host-visible).
hipError_t launch_instance_work(const uint64_t* bounds, int n_inst, uint32_t* work, uint32_t* off, uint32_t* total,
                                hipStream_t st);
// grads[rows[g] * n_inst + b] = sum of instance b's partials of gate g in
// block order (gate g's partials start at part_off[g], its item offsets are
// offs[g]).
hipError_t launch_instance_grads(const uint64_t* part_off, const uint64_t* rows, const uint32_t* const* offs,
                                 uint64_t n_gates, const double* partials, int n_inst, double* grads, hipStream_t st);

// Keep decisions of a caller-side kernel for the e sorted children: flag =
// emit or pending, and own_keep[source row] = the decision (origins as
// recorded by branch_scatter).
hipError_t launch_apply_keep(uint64_t e, const uint8_t* keep, uint8_t* flag, const uint32_t* origin,
                             uint8_t* own_keep, hipStream_t st);

// Reverse step of one recorded gate: lam_in[r] = d<O>/dc_r before the gate
// from lam_out over the gate's output rows, and one partial of d<O>/dtheta
// per block (n_blocks of them) into partials.
hipError_t launch_adjoint_gate(uint64_t n_in, TapeDev tape, const double* c_in, const double* lam_out, double* lam_in,
                               double sin_t, double cos_t, double* partials, int n_blocks, hipStream_t st);
// Batched engines: work[3 i .. 3 i + 2] = (instance, block, blocks of that
// instance); each item reverses its block of the instance's rows
// [bounds[b], bounds[b + 1]) with the instance's sin / cos (sc: 3 per
// instance) and writes one partial of d<O_b>/dtheta_b.
hipError_t launch_adjoint_gate_segments(uint64_t n_work, const uint32_t* work, const uint64_t* bounds,
                                       const double* sc, TapeDev tape, const double* c_in, const double* lam_out,
                                       double* lam_in, double* partials, hipStream_t st);
// Reverse step of a recorded noise / reclaim pass (or row permutation).
hipError_t launch_adjoint_noise(uint64_t n_in, TapeDev tape, const double* lam_out, double* lam_in, hipStream_t st);
// Reverse step of a recorded sparse gate over its m records, in place on lam
// (tmp: m doubles), one partial of d<O>/dtheta per block (n_blocks).
hipError_t launch_adjoint_sparse(uint64_t m, const SparseRec* tape, double* lam, double* tmp, double sin_t,
                                 double cos_t, double* partials, int n_blocks, hipStream_t st);
// The same on a batched engine: record i belongs to instance inst[i], whose
// (sin, cos) are sc[3 inst], sc[3 inst + 1]; d<O_b>/dtheta goes to grads[b]
// for every instance b < n_inst (zero where no record), through per-block
// partials (part: sparse_inst_blocks(m, n_inst) * n_inst doubles).
constexpr int kSparseRecInstMax = 4096;  // instances a block's LDS partials hold
int sparse_inst_blocks(uint64_t m, int n_inst);
hipError_t launch_adjoint_sparse_batched(uint64_t m, const SparseRec* tape, const uint32_t* inst, const double* sc,
                                         int n_inst, double* lam, double* tmp, double* part, double* grads,
                                         hipStream_t st);
// A recorded sort: o[j] = j, then out[origin[j]] = j.
hipError_t launch_iota_u32(uint32_t* o, uint64_t n, hipStream_t st);
hipError_t launch_perm_out(const uint32_t* origin, uint32_t* out, uint64_t n, hipStream_t st);
int adjoint_blocks(uint64_t n);

// bounds[b] = first row whose instance id (key word W-1 >> shift) is >= b,
// for b = 0..n_inst; rows are sorted, so each instance is one contiguous run.
hipError_t launch_instance_bounds(const uint64_t* top_plane, const uint64_t* n_dev, int shift, int n_inst,
                                  uint64_t* bounds, hipStream_t st);

hipError_t launch_gather_bounds(const uint64_t* offs, uint64_t n_blocks, int m, uint64_t* out, hipStream_t st);
hipError_t launch_count_below(const double* coeff, uint64_t n, double thr, unsigned long long* out, hipStream_t st);
hipError_t launch_exclusive_scan(const uint32_t* in, uint64_t n, uint64_t* out, uint64_t* scratch, hipStream_t st);
uint64_t scan_scratch_words(uint64_t n);

}  // namespace pq

// pq_kernels.hip -- gfx950 kernels of the per-gate term-propagation path.
//
// One gate G (generator, angle) over a sorted, unique SoA store S runs as:
//   branch_count / branch_scatter  anticommuting rows P emit their child
//       Q = G P (key P ^ G, value (c sin)(-Im phase), flag emit|pending) into
//       per-class buckets, class = P's bits at G's support.  XOR with G keeps
//       the order of rows of one class, so every bucket is already sorted.
//   merge rounds                   pairwise merge-path of buckets -> one run C.
//   store_merge                    merge-path of S and C: every row P is
//       rewritten as c cos + (what its partner G P sent), new keys are spliced
//       in, and the output is compacted with a decoupled look-back.
// Reference semantics: crates/core/src/engine/termsum.rs:561-800 and
// crates/core/src/engine/partitioned_termsum.rs:276-350.  Under P -> G P each
// key receives at most one addition per gate, so the result does not depend
// on the order rows are visited in, and the FP64 values match the reference
// bit for bit when built with -ffp-contract=off.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "pq_kernels.h"

namespace pq {

namespace {

constexpr uint64_t kFlagAgg = 1ULL << 62;
constexpr uint64_t kFlagPre = 2ULL << 62;
constexpr uint64_t kValMask = (1ULL << 62) - 1;

__device__ __forceinline__ uint64_t ld_status(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Publish a gate's row counts to host-mapped memory: the child count, then
// (tag << 48) | n_out with release order at system scope, so a host that
// sees the tag also sees the child count.
constexpr uint64_t kPubRefused = (1ULL << 48) - 1;
__device__ __forceinline__ void publish_counts(uint64_t* n_host, uint64_t* e_host, uint32_t tag, uint64_t n_out,
                                               uint64_t e) {
  if (!n_host) return;
  if (e_host) __hip_atomic_store(e_host, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  __hip_atomic_store(n_host, ((uint64_t)tag << 48) | n_out, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor((int)lo, o, 64);
    hi = __shfl_xor((int)hi, o, 64);
    v += ((uint64_t)hi << 32) | lo;
  }
  return v;
}

// Decoupled look-back over tile aggregates: run by all 64 lanes of wave 0.
// Predecessor tiles took smaller tickets, so they are resident or finished
// and every wait ends.  Returns the exclusive prefix of this tile.  Each
// lane reads kLookG statuses per step (lane l: tile - 1 - (l kLookG + j)).
// Measured on MI355X (store merge, 8M rows): width 1 beats 4 (330 vs 382 us):
// a wider step spins on more unpublished neighbours per lane, one at a time.
#ifndef PQ_LOOKBACK_WIDTH
#define PQ_LOOKBACK_WIDTH 1
#endif
constexpr int kLookG = PQ_LOOKBACK_WIDTH;
__device__ uint64_t lookback(uint64_t* status, uint32_t tile, uint64_t agg) {
  const int lane = threadIdx.x & 63;
  if (tile == 0) {
    if (lane == 0) st_status(&status[0], kFlagPre | agg);
    return 0;
  }
  if (lane == 0) st_status(&status[tile], kFlagAgg | agg);
  uint64_t excl = 0;
  int64_t base = (int64_t)tile - 1;
  for (;;) {
    uint64_t st[kLookG];
#pragma unroll
    for (int j = 0; j < kLookG; ++j) {
      const int64_t idx = base - (lane * kLookG + j);
      st[j] = idx >= 0 ? ld_status(&status[idx]) : kFlagPre;  // virtual zero prefix before tile 0
    }
#pragma unroll
    for (int j = 0; j < kLookG; ++j) {
      const int64_t idx = base - (lane * kLookG + j);
      while ((st[j] >> 62) == 0) st[j] = ld_status(&status[idx]);
    }
    // the nearest inclusive prefix: lowest lane, then lowest j in that lane
    int jfirst = kLookG;
#pragma unroll
    for (int j = kLookG - 1; j >= 0; --j)
      if ((st[j] >> 62) == 2) jfirst = j;
    const uint64_t pmask = __ballot(jfirst < kLookG);
    const int first = pmask ? (__ffsll((unsigned long long)pmask) - 1) : 64;
    uint64_t v = 0;
#pragma unroll
    for (int j = 0; j < kLookG; ++j)
      if (lane < first || (lane == first && j <= jfirst)) v += st[j] & kValMask;
    excl += wave_sum_u64(v);
    if (pmask) break;
    base -= 64 * kLookG;
  }
  if (lane == 0) st_status(&status[tile], kFlagPre | (excl + agg));
  return excl;
}

// Block-wide exclusive scan of one u32 per thread (256 threads).  Returns
// the exclusive prefix; *total receives the block sum.
__device__ uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up((int)x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  uint32_t wbase = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kBlock / 64; ++i) {
    if (i < wave) wbase += wsum[i];
    tot += wsum[i];
  }
  *total = tot;
  __syncthreads();
  return wbase + x - v;
}

template <int W>
__device__ __forceinline__ Key<W> load_key(const uint64_t* key, uint64_t cap, uint64_t r) {
  Key<W> k;
#pragma unroll
  for (int i = 0; i < W; ++i) k.w[i] = key[(uint64_t)i * cap + r];
  return k;
}

template <int W>
__device__ __forceinline__ void store_key(uint64_t* key, uint64_t cap, uint64_t r, const Key<W>& k) {
#pragma unroll
  for (int i = 0; i < W; ++i) key[(uint64_t)i * cap + r] = k.w[i];
}

// Anticommutation of a stored row with the gate generator.  Pauli rows read
// only the key words the pair-swapped generator touches.
template <int KIND, int W>
__device__ __forceinline__ bool row_anti(const GateArgs<W>& g, const Key<W>& k) {
  if (KIND == kPauli) {
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) acc ^= k.w[i] & g.fold.w[i];
    return popc64(acc) & 1;
  }
  int ov = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) ov += popc64(k.w[i] & g.gen.w[i]);
  if (g.gen_len & 1) ov += key_count(k);
  return (ov & 1) != 0;
}

template <int KIND, int W>
__device__ __forceinline__ Key<W> load_key_for_gate(const GateArgs<W>& g, const uint64_t* key, uint64_t cap,
                                                    uint64_t r, bool& anti) {
  Key<W> k;
  if (KIND == kPauli) {
#pragma unroll
    for (int i = 0; i < W; ++i) k.w[i] = g.fold.w[i] ? key[(uint64_t)i * cap + r] : 0;
    anti = row_anti<KIND, W>(g, k);
    if (anti) {
#pragma unroll
      for (int i = 0; i < W; ++i)
        if (!g.fold.w[i]) k.w[i] = key[(uint64_t)i * cap + r];
    }
  } else {
    k = load_key<W>(key, cap, r);
    anti = row_anti<KIND, W>(g, k);
  }
  return k;
}

// Phase exponent of G * P with the frame sign folded in.
template <int KIND, int W>
__device__ __forceinline__ int child_exp(const GateArgs<W>& g, const Key<W>& p, const Key<W>& q) {
  int e;
  if (KIND == kPauli) {
    int zx = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) zx += popc64((g.gen.w[i] >> 1) & p.w[i] & kEven);
    e = g.gen_y + pauli_y(p) - pauli_y(q) + 2 * zx;
  } else {
    int par = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) par += popc64(p.w[i] & g.above.w[i]);
    e = herm_exp(g.gen_len) + herm_exp(key_count(p)) - herm_exp(key_count(q)) + 2 * (par & 1);
  }
  return (e + (g.neg ? 2 : 0)) & 3;
}

enum : int { kDecNone = 0, kDecEmit = 1, kDecPending = 2, kDecDrop = 3 };

// Key with the instance bits of a batched engine cleared (weights, traces).
template <int W>
__device__ __forceinline__ Key<W> unit_bits(const Key<W>& k, uint64_t unit_mask) {
  Key<W> m = k;
  m.w[W - 1] &= unit_mask;
  return m;
}

// sin / cos of the rotation for a row: per instance on a batched engine.
struct RowRot {
  double s, c;
  bool branch;  // false: phase-only for this instance (no branch is formed)
};
// Batched engines: an instance whose flag is negative is frozen (a recorded
// batch keeps instances that have finished their own layers): its rows
// count as commuting, so they pass every later gate unchanged.
template <int W>
__device__ __forceinline__ bool row_frozen(const GateArgs<W>& g, const Key<W>& p) {
  return g.sc && g.sc[3 * (p.w[W - 1] >> g.inst_shift) + 2] < 0.0;
}

// The table is read through a global-address-space pointer: the loads then
// differ in type from the kernel argument's sin_t / cos_t, so the compiler
// cannot merge the two cases into one load through a selected pointer (which
// puts the argument in scratch and turns the loads into flat loads).
typedef const __attribute__((address_space(1))) double global_double;

template <int W>
__device__ __forceinline__ RowRot row_rot(const GateArgs<W>& g, const Key<W>& p) {
  if (!g.sc) return RowRot{g.sin_t, g.cos_t, true};
  const uint64_t b = p.w[W - 1] >> g.inst_shift;
  global_double* t = (global_double*)g.sc + 3 * b;
  return RowRot{t[0], t[1], t[2] == 0.0};
}

// emit_from_row (termsum.rs:701-751) for the device: the precheck holds a
// branch back when |c sin| < min (before the weight test), a child above the
// weight bound is dropped, anything else is emitted.  v carries the phase in
// both the emit and the pending case (drain_claims applies it, :635).
// *ph receives -Im(phase) (+1 or -1) unless the branch is dropped.
template <int KIND, int W>
__device__ __forceinline__ int decide(const GateArgs<W>& g, const Key<W>& p, double c, Key<W>& q, double& v,
                                      double* ph = nullptr) {
#pragma unroll
  for (int i = 0; i < W; ++i) q.w[i] = p.w[i] ^ g.gen.w[i];
  const RowRot rr = row_rot<W>(g, p);
  if (!rr.branch) return kDecNone;
  const double s = c * rr.s;
  if (g.has_min && fabs(s) < g.min_coeff) {
    const double f = neg_im(child_exp<KIND, W>(g, p, q));
    if (ph) *ph = f;
    v = s * f;
    return kDecPending;
  }
  if (g.has_maxw && (uint32_t)weight<KIND, W>(unit_bits<W>(q, g.unit_mask), g.n_units) > g.max_weight)
    return kDecDrop;
  const double f = neg_im(child_exp<KIND, W>(g, p, q));
  if (ph) *ph = f;
  v = s * f;
  return kDecEmit;
}

// Class bucket of a row: its bits at the generator's support, packed.  The
// word is picked with a static select so the key stays in registers.
template <int W>
__device__ __forceinline__ int class_of(const GateArgs<W>& g, const Key<W>& p) {
  int c = 0;
  for (int j = 0; j < g.n_cls_bits; ++j) {
    const int pos = g.cls_pos[j];
    const int wd = pos >> 6;
    uint64_t word = p.w[0];
#pragma unroll
    for (int i = 1; i < W; ++i) word = (i == wd) ? p.w[i] : word;
    c |= (int)((word >> (pos & 63)) & 1ULL) << j;
  }
  return c;
}

// ------------------------------------------------------------------ branch

// Lanes of one wave whose class equals this lane's: intersect the ballots of
// every class bit (invalid lanes never match a valid one).
__device__ __forceinline__ uint64_t same_class_lanes(int cls, int n_bits) {
  const bool valid = cls >= 0;
  uint64_t same = __ballot(valid);
  for (int b = 0; b < n_bits; ++b) {
    const bool bit = valid && ((cls >> b) & 1);
    const uint64_t mb = __ballot(bit);
    same &= bit ? mb : ~mb;
  }
  return same;
}

// Branch decision of kItems rows of one tile (rows base + it * kBlock + tid):
// the key words the anticommutation test needs are loaded for every item
// first, then the rest of the key and the coefficient of the anticommuting
// rows, so each lane keeps several independent loads in flight.
template <int W>
struct BranchItem {
  Key<W> q;     // child key G P
  double v;     // sine branch with its phase
  int cls;      // class bucket, -1 when no child
  uint8_t flag; // kEmit / kPending
  bool anti, pend, drop;
  bool neg_ph;  // -Im(phase) = -1
};

template <int KIND, int W>
__device__ __forceinline__ bool word_needed(const GateArgs<W>& g, int w) {
  if (KIND == kPauli) return g.fold.w[w] != 0;
  return (g.gen_len & 1) || g.gen.w[w] != 0;
}

template <int KIND, int W>
__device__ __forceinline__ void branch_items(const GateArgs<W>& g, const StoreDev& s, uint64_t n, uint64_t base,
                                             BranchItem<W> (&d)[kItems]) {
  Key<W> p[kItems];
  double c[kItems];
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint64_t r = base + (uint64_t)it * kBlock + threadIdx.x;
#pragma unroll
    for (int w = 0; w < W; ++w) p[it].w[w] = (r < n && word_needed<KIND, W>(g, w)) ? s.key[(uint64_t)w * s.cap + r] : 0;
  }
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint64_t r = base + (uint64_t)it * kBlock + threadIdx.x;
    d[it].anti = r < n && row_anti<KIND, W>(g, p[it]);
    c[it] = 0.0;
    if (d[it].anti) {
#pragma unroll
      for (int w = 0; w < W; ++w)
        if (!word_needed<KIND, W>(g, w)) p[it].w[w] = s.key[(uint64_t)w * s.cap + r];
      c[it] = s.coeff[r];
      if (row_frozen<W>(g, p[it])) d[it].anti = false;
    }
  }
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    d[it].cls = -1;
    d[it].flag = 0;
    d[it].pend = d[it].drop = d[it].neg_ph = false;
    if (!d[it].anti) continue;
    double ph = 1.0;
    const int dec = decide<KIND, W>(g, p[it], c[it], d[it].q, d[it].v, &ph);
    d[it].neg_ph = ph < 0.0;
    d[it].pend = dec == kDecPending;
    d[it].drop = dec == kDecDrop;
    if (dec == kDecEmit || dec == kDecPending) {
      d[it].cls = class_of<W>(g, p[it]);
      d[it].flag = dec == kDecEmit ? kEmit : kPending;
    }
  }
}

// Per-block child counts by class, class-major: counts[c * nb + block].
// Blocks past the live rows write zeros, so the grid can follow a host-side
// bound of the row count while the count itself stays on the device.
template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_branch_count(GateArgs<W> g, StoreDev s, const uint64_t* n_ptr,
                                                         uint32_t* counts, int m, uint64_t n_blocks, Counters* ctr) {
  constexpr int kWaves = kBlock / 64;
  __shared__ uint32_t wcnt[kWaves][kMaxClasses];
  __shared__ uint32_t wstat[kWaves][3];
  const uint64_t n = *n_ptr;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < kWaves * kMaxClasses; i += kBlock) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  BranchItem<W> d[kItems];
  branch_items<KIND, W>(g, s, n, base, d);
  uint32_t vis = 0, dec = 0, drop = 0;
  const uint64_t lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    vis += popc64(__ballot(d[it].anti));
    dec += popc64(__ballot(d[it].pend));
    drop += popc64(__ballot(d[it].drop));
    const uint64_t same = same_class_lanes(d[it].cls, g.n_cls_bits);
    if (d[it].cls >= 0 && (same & lt) == 0) wcnt[wave][d[it].cls] += popc64(same);  // one leader per class
  }
  if (lane == 0) {
    wstat[wave][0] = vis;
    wstat[wave][1] = dec;
    wstat[wave][2] = drop;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < m; c += kBlock) {
    uint32_t t = 0;
    for (int wv = 0; wv < kWaves; ++wv) t += wcnt[wv][c];
    counts[(uint64_t)c * n_blocks + blockIdx.x] = t;
  }
  if (threadIdx.x == 0 && ctr) {
    uint32_t a = 0, b = 0, dd = 0;
    for (int wv = 0; wv < kWaves; ++wv) {
      a += wstat[wv][0];
      b += wstat[wv][1];
      dd += wstat[wv][2];
    }
    Counters* cs = ctr_slot(ctr);
    if (a) atomicAdd(&cs->visited, (unsigned long long)a);
    if (b) atomicAdd(&cs->declined, (unsigned long long)b);
    if (dd) atomicAdd(&cs->dropped_weight, (unsigned long long)dd);
  }
}

// Same walk as k_branch_count; each child lands at
// offs[class * nb + block] + (its stable rank among this block's children of
// that class), so every class bucket is in store order.
template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_branch_scatter(GateArgs<W> g, StoreDev s, const uint64_t* n_ptr,
                                                           const uint64_t* offs, int m, uint64_t n_blocks,
                                                           ChildDev c, uint32_t* tape_child) {
  constexpr int kWaves = kBlock / 64;
  __shared__ uint32_t wcnt[kItems * kWaves][kMaxClasses];  // counts, then block-local bases
  const uint64_t n = *n_ptr;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < kItems * kWaves * kMaxClasses; i += kBlock) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * kTile;
  BranchItem<W> d[kItems];
  branch_items<KIND, W>(g, s, n, base, d);
  uint32_t rank[kItems];
  const uint64_t lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint64_t same = same_class_lanes(d[it].cls, g.n_cls_bits);
    rank[it] = popc64(same & lt);
    if (d[it].cls >= 0) wcnt[it * kWaves + wave][d[it].cls] = popc64(same);
  }
  __syncthreads();
  for (int cl = threadIdx.x; cl < m; cl += kBlock) {
    uint32_t acc = 0;
    for (int i = 0; i < kItems * kWaves; ++i) {
      const uint32_t t = wcnt[i][cl];
      wcnt[i][cl] = acc;
      acc += t;
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint64_t r = base + (uint64_t)it * kBlock + threadIdx.x;
    if (tape_child && r < n) tape_child[r] = d[it].anti ? kRecNoChild : kRecCommute;
    if (d[it].cls < 0) continue;
    const uint64_t pos =
        offs[(uint64_t)d[it].cls * n_blocks + blockIdx.x] + wcnt[it * kWaves + wave][d[it].cls] + rank[it];
    store_key<W>(c.key, c.cap, pos, d[it].q);
    c.val[pos] = d[it].v;
    c.flag[pos] = d[it].flag;
    if (c.origin) c.origin[pos] = (uint32_t)r | (d[it].neg_ph ? kRecSign : 0u);
  }
}

// Phase-only gates (|sin| < 1e-9, termsum.rs:551): anticommuting rows are
// scaled in place.
template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_scale_anti(GateArgs<W> g, StoreDev s, const uint64_t* n_ptr, double f) {
  const uint64_t n = *n_ptr;
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += (uint64_t)gridDim.x * kBlock) {
    bool anti;
    (void)load_key_for_gate<KIND, W>(g, s.key, s.cap, r, anti);
    if (anti) s.coeff[r] = s.coeff[r] * f;
  }
}

// ------------------------------------------------------------ merge rounds

// Merge path, one wave cooperating: the number of A elements among the first
// d of merge(A, B), A first on equal keys.  63 probes per round shrink the
// candidate range 63-fold, so a store of 10^8 rows needs 5 dependent rounds of
// loads instead of 27.  Every lane of the wave must call it; all get the
// result.  a(i) / b(i) load the i-th key of each run.
template <int W, typename FA, typename FB>
__device__ uint64_t wave_mp_search(FA a, FB b, uint64_t na, uint64_t nb, uint64_t d) {
  const int lane = threadIdx.x & 63;
  uint64_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
  while (lo < hi) {
    const uint64_t span = hi - lo;
    const uint64_t i = lane == 63 ? hi : lo + (span * (uint64_t)lane) / 63;
    bool pred = true;  // B[d-i-1] < A[i]: enough A elements taken
    if (i < hi) pred = key_less<W>(b(d - i - 1), a(i));
    const uint64_t mask = __ballot(pred);
    const int f = __ffsll((unsigned long long)mask) - 1;
    const uint64_t i_f = __shfl(i, f, 64);
    const uint64_t i_prev = __shfl(i, f > 0 ? f - 1 : 0, 64);
    if (f == 0) return i_f;
    lo = i_prev + 1;
    hi = i_f;
  }
  return lo;
}

// Sort each tile of kT children by key in LDS (bitonic network over
// indices; slots past the tile's count compare as +infinity).
template <int W>
__global__ __launch_bounds__(kBlock) void k_block_sort(ChildDev src, ChildDev dst, uint64_t e) {
  constexpr int kT = MergeCfg<W>::tile;
  __shared__ uint64_t lkey[W][kT];
  __shared__ uint16_t idx[kT];
  const uint64_t base = (uint64_t)blockIdx.x * kT;
  const int cnt = (int)(e - base < (uint64_t)kT ? e - base : (uint64_t)kT);
  for (int t = threadIdx.x; t < kT; t += kBlock) {
    if (t < cnt) {
#pragma unroll
      for (int w = 0; w < W; ++w) lkey[w][t] = src.key[(uint64_t)w * src.cap + base + t];
    }
    idx[t] = (uint16_t)t;
  }
  __syncthreads();
  for (int size = 2; size <= kT; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < kT / 2; t += kBlock) {
        const int i = 2 * stride * (t / stride) + (t % stride);
        const int j = i + stride;
        const int a = idx[i], b = idx[j];
        bool a_gt_b;
        if (a >= cnt || b >= cnt) {
          a_gt_b = a >= cnt && (b < cnt || a > b);
        } else {
          Key<W> ka, kb;
#pragma unroll
          for (int w = 0; w < W; ++w) {
            ka.w[w] = lkey[w][a];
            kb.w[w] = lkey[w][b];
          }
          a_gt_b = key_less<W>(kb, ka);
        }
        const bool up = (i & size) == 0;
        if (a_gt_b == up) {
          idx[i] = (uint16_t)b;
          idx[j] = (uint16_t)a;
        }
      }
      __syncthreads();
    }
  }
  for (int t = threadIdx.x; t < cnt; t += kBlock) {
    const int sidx = idx[t];
#pragma unroll
    for (int w = 0; w < W; ++w) dst.key[(uint64_t)w * dst.cap + base + t] = lkey[w][sidx];
    dst.val[base + t] = src.val[base + sidx];
    dst.flag[base + t] = src.flag[base + sidx];
    if (src.origin) dst.origin[base + t] = src.origin[base + sidx];
  }
}

// LDS working set of one child-merge tile: inputs (A then B) and the slot
// map of the merged order.
template <int W>
struct MergeTileSmem {
  uint64_t key[W][MergeCfg<W>::tile];
  double val[MergeCfg<W>::tile];
  uint8_t flag[MergeCfg<W>::tile];
  uint16_t slot[MergeCfg<W>::tile];
  uint64_t split[2];
};

// Merge output positions [d0, d1) of runs A = [a0, a1) and B = [a1, b1) of
// src into dst[a0 + d0, a0 + d1), staging through LDS so reads and writes are
// coalesced.
template <int W>
__device__ void merge_tile(MergeTileSmem<W>& sm, ChildDev src, ChildDev dst, uint64_t a0, uint64_t a1,
                           uint64_t b1, uint64_t d0, uint64_t d1) {
  constexpr int kI = MergeCfg<W>::items;
  const uint64_t na = a1 - a0, nb = b1 - a1;
  const int wave = threadIdx.x >> 6;
  if (wave < 2) {
    const uint64_t d = wave == 0 ? d0 : d1;
    auto ga = [&](uint64_t i) { return load_key<W>(src.key, src.cap, a0 + i); };
    auto gb = [&](uint64_t i) { return load_key<W>(src.key, src.cap, a1 + i); };
    const uint64_t sp = wave_mp_search<W>(ga, gb, na, nb, d);
    if ((threadIdx.x & 63) == 0) sm.split[wave] = sp;
  }
  __syncthreads();
  const uint64_t ia0 = sm.split[0], ia1 = sm.split[1];
  const uint64_t ib0 = d0 - ia0, ib1 = d1 - ia1;
  const int nA = (int)(ia1 - ia0), nB = (int)(ib1 - ib0);
  {
    // every load of this thread's elements is issued before any LDS store
    uint64_t rk[kI][W];
    double rv[kI];
    uint8_t rf[kI];
#pragma unroll
    for (int k = 0; k < kI; ++k) {
      const int t = threadIdx.x + k * kBlock;
      if (t < nA + nB) {
        const uint64_t gi = t < nA ? a0 + ia0 + t : a1 + ib0 + (t - nA);
#pragma unroll
        for (int w = 0; w < W; ++w) rk[k][w] = src.key[(uint64_t)w * src.cap + gi];
        rv[k] = src.val[gi];
        rf[k] = src.flag[gi];
      }
    }
#pragma unroll
    for (int k = 0; k < kI; ++k) {
      const int t = threadIdx.x + k * kBlock;
      if (t < nA + nB) {
#pragma unroll
        for (int w = 0; w < W; ++w) sm.key[w][t] = rk[k][w];
        sm.val[t] = rv[k];
        sm.flag[t] = rf[k];
      }
    }
  }
  __syncthreads();
  const int tot = nA + nB;
  const int dl = threadIdx.x * kI;
  if (dl < tot) {
    int lo = dl > nB ? dl - nB : 0, hi = dl < nA ? dl : nA;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      Key<W> ka, kb;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        ka.w[w] = sm.key[w][mid];
        kb.w[w] = sm.key[w][nA + dl - mid - 1];
      }
      if (!key_less<W>(kb, ka))
        lo = mid + 1;
      else
        hi = mid;
    }
    int ia = lo, ib = dl - lo;
    for (int k = 0; k < kI && dl + k < tot; ++k) {
      bool takeA;
      if (ia >= nA)
        takeA = false;
      else if (ib >= nB)
        takeA = true;
      else {
        Key<W> ka, kb;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          ka.w[w] = sm.key[w][ia];
          kb.w[w] = sm.key[w][nA + ib];
        }
        takeA = !key_less<W>(kb, ka);
      }
      sm.slot[dl + k] = (uint16_t)(takeA ? ia++ : nA + ib++);
    }
  }
  __syncthreads();
  const uint64_t o0 = a0 + d0;
  for (int t = threadIdx.x; t < tot; t += kBlock) {
    const int li = sm.slot[t];
#pragma unroll
    for (int w = 0; w < W; ++w) dst.key[(uint64_t)w * dst.cap + o0 + t] = sm.key[w][li];
    dst.val[o0 + t] = sm.val[li];
    dst.flag[o0 + t] = sm.flag[li];
    if (src.origin) dst.origin[o0 + t] = src.origin[li < nA ? a0 + ia0 + li : a1 + ib0 + (li - nA)];
  }
  __syncthreads();
}

// One round of pairwise merges of the class buckets.  Bucket c spans
// [offs[c * nb], offs[(c + 1) * nb]) (the exclusive scan of the branch
// counts), so the plan is read from the device and needs no host round trip.
// After r rounds run j covers buckets [j 2^r, (j + 1) 2^r); this round merges
// runs 2p and 2p + 1.  Blocks stride over the round's tiles.
template <int W>
__global__ __launch_bounds__(kBlock) void k_merge_round(ChildDev src, ChildDev dst, const uint64_t* offs,
                                                        uint64_t nb, int m, int r) {
  constexpr int kT = MergeCfg<W>::tile;
  __shared__ MergeTileSmem<W> sm;
  __shared__ uint64_t tile0[kMaxClasses / 2 + 1];
  const int step = 1 << r;
  const int n_pairs = m / (2 * step);
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (int p = 0; p < n_pairs; ++p) {
      tile0[p] = t;
      const uint64_t a0 = offs[(uint64_t)(2 * p * step) * nb], b1 = offs[(uint64_t)((2 * p + 2) * step) * nb];
      t += (b1 - a0 + kT - 1) / kT;
    }
    tile0[n_pairs] = t;
  }
  __syncthreads();
  const uint64_t total = tile0[n_pairs];
  for (uint64_t tile = blockIdx.x; tile < total; tile += gridDim.x) {
    int p = 0;
    while (p + 1 < n_pairs && tile >= tile0[p + 1]) ++p;
    const uint64_t a0 = offs[(uint64_t)(2 * p * step) * nb];
    const uint64_t a1 = offs[(uint64_t)((2 * p + 1) * step) * nb];
    const uint64_t b1 = offs[(uint64_t)((2 * p + 2) * step) * nb];
    const uint64_t d0 = (tile - tile0[p]) * kT;
    const uint64_t d1 = d0 + kT < b1 - a0 ? d0 + kT : b1 - a0;
    merge_tile<W>(sm, src, dst, a0, a1, b1, d0, d1);
  }
}

// Regular rounds of the generic sort: runs [2p L, (2p+1) L), [(2p+1) L,
// (2p+2) L) clipped to e.
template <int W>
__global__ __launch_bounds__(kBlock) void k_merge_regular(ChildDev src, ChildDev dst, uint64_t run_len, uint64_t e) {
  constexpr int kT = MergeCfg<W>::tile;
  __shared__ MergeTileSmem<W> sm;
  const uint64_t tpp = 2 * run_len / kT;
  const uint64_t pairs = (e + 2 * run_len - 1) / (2 * run_len);
  const uint64_t last_len = e - (pairs - 1) * 2 * run_len;
  const uint64_t total = (pairs - 1) * tpp + (last_len + kT - 1) / kT;
  for (uint64_t tile = blockIdx.x; tile < total; tile += gridDim.x) {
    const uint64_t p = tile / tpp;
    const uint64_t a0 = p * 2 * run_len;
    const uint64_t a1 = a0 + run_len < e ? a0 + run_len : e;
    const uint64_t b1 = a0 + 2 * run_len < e ? a0 + 2 * run_len : e;
    const uint64_t d0 = (tile - p * tpp) * kT;
    const uint64_t d1 = d0 + kT < b1 - a0 ? d0 + kT : b1 - a0;
    merge_tile<W>(sm, src, dst, a0, a1, b1, d0, d1);
  }
}

// ------------------------------------------------------------- store merge

// Tile boundaries of merge(S, C) (S first on equal keys), one wave each; a
// child equal to the last S row of a tile joins that tile so a matched pair
// never straddles a boundary.  Also clears the look-back state of the merge
// that follows (status words, ticket).  e is read from the device.
template <int W>
__global__ __launch_bounds__(kBlock) void k_store_splits(StoreDev s, const uint64_t* n_ptr, ChildDev c,
                                                         const uint64_t* e_ptr, uint64_t out_cap, uint64_t* splits,
                                                         uint64_t* status, uint32_t* ticket, uint32_t* overflow) {
  constexpr uint64_t kT = StoreCfg<W>::tile;
  const uint64_t n = *n_ptr, e = *e_ptr;
  if (n + e > out_cap) {  // the host's capacity bound was violated: refuse, never write past the end
    if (blockIdx.x == 0 && threadIdx.x == 0) *overflow = 1;
    return;
  }
  const uint64_t n_tiles = (n + e + kT - 1) / kT;
  const uint64_t waves = (uint64_t)gridDim.x * (kBlock / 64);
  for (uint64_t t = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); t <= n_tiles; t += waves) {
    uint64_t d = t * kT;
    if (d > n + e) d = n + e;
    auto ga = [&](uint64_t i) { return load_key<W>(s.key, s.cap, i); };
    auto gb = [&](uint64_t i) { return load_key<W>(c.key, c.cap, i); };
    const uint64_t i = wave_mp_search<W>(ga, gb, n, e, d);
    uint64_t j = d - i;
    if ((threadIdx.x & 63) == 0) {
      if (i > 0 && j < e && key_eq<W>(ga(i - 1), gb(j))) ++j;
      splits[2 * t] = i;
      splits[2 * t + 1] = j;
      if (t < n_tiles) status[t] = 0;
      if (t == 0) *ticket = 0;
    }
  }
}

struct TileSplit {
  uint64_t i0, j0, i1, j1;  // store rows [i0, i1) and children [j0, j1)
};

// LDS working set of one store-merge tile.
template <int W>
struct StoreTileSmem {
  static constexpr int kCap = StoreCfg<W>::tile + 1;
  uint64_t key[W][kCap];
  // input values while merging; after the block scan (which every thread
  // enters only once done merging) the output coefficients, compacted
  double val[kCap];
  uint8_t flag[kCap];  // 0 for S rows, the child flag for C
  uint16_t oslot[kCap];
  uint32_t tile;  // ticket broadcast
  uint64_t base;
};

// Register copy of one store-merge tile's input elements t = tid + k kBlock
// (k < kI): W key words, the value, and the child flag (0 for store rows).
// A tile holds at most kT + 1 elements; the rare element kT (a child pulled
// into the tile to stay next to its equal store row) is staged by thread 0
// directly, so no lane carries registers for it.
template <int W, int kI>
struct TileRegs {
  uint64_t key[kI][W];
  double val[kI];
  uint8_t flag[kI];
};

__device__ __forceinline__ TileSplit load_split(const uint64_t* splits, uint32_t t) {
  return TileSplit{splits[2 * (uint64_t)t], splits[2 * (uint64_t)t + 1], splits[2 * (uint64_t)t + 2],
                   splits[2 * (uint64_t)t + 3]};
}

// Issue every load of a tile's elements into registers; an element reads the
// store (t < nS) or the children through one select of base and plane
// stride, so lanes do not diverge.
template <int W, int kI>
__device__ __forceinline__ void tile_fetch(TileRegs<W, kI>& r, const StoreDev& s, const ChildDev& c,
                                           const TileSplit& sp) {
  const int nS = (int)(sp.i1 - sp.i0), tot = nS + (int)(sp.j1 - sp.j0);
#pragma unroll
  for (int k = 0; k < kI; ++k) {
    const int t = threadIdx.x + k * kBlock;
    if (t < tot) {
      const bool fromS = t < nS;
      const uint64_t gi = fromS ? sp.i0 + t : sp.j0 + (t - nS);
      const uint64_t* kp = (fromS ? s.key : c.key) + gi;
      const uint64_t stride = fromS ? s.cap : c.cap;
#pragma unroll
      for (int w = 0; w < W; ++w) r.key[k][w] = kp[(uint64_t)w * stride];
      r.val[k] = (fromS ? s.coeff : c.val)[gi];
      r.flag[k] = fromS ? (uint8_t)0 : c.flag[gi];
    }
  }
}

// Whether store row a branches under this gate (anticommutes and its
// instance is not frozen), evaluated once while its key is in registers.
template <int KIND, int W>
__device__ __forceinline__ uint8_t stage_store_row(const GateArgs<W>& g, const Key<W>& a) {
  return row_anti<KIND, W>(g, a) && !row_frozen<W>(g, a) ? 1 : 0;
}

// Three-way key comparison of LDS elements i and j (sign of key_i - key_j in
// store order) that starts at word sw: every key of a tile agrees on the
// words above its most significant differing word, so those are skipped.
template <int W, int kCap>
__device__ __forceinline__ int lds_cmp(const uint64_t (&key)[W][kCap], int i, int j, int sw) {
  for (int w = sw; w >= 0; --w) {
    const uint64_t x = key[w][i], y = key[w][j];
    if (x != y) return x < y ? -1 : 1;
  }
  return 0;
}

// The most significant word in which the keys of a staged tile differ (0 if
// none): S rows [0, nS) and children [nS, nS + nC) are each sorted, so their
// end points bound every key of the tile.
template <int W, int kCap>
__device__ __forceinline__ int tile_sig_word(const uint64_t (&key)[W][kCap], int nS, int nC) {
  int sw = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    uint64_t d = 0;
    if (nS) d |= key[w][0] ^ key[w][nS - 1];
    if (nC) d |= key[w][nS] ^ key[w][nS + nC - 1];
    if (nS && nC) d |= key[w][0] ^ key[w][nS];
    if (d) sw = w;
  }
  return sw;
}

// Merge the store with the sorted children into the output store.
//  S row P (anticommuting): c' = c cos, plus the partner's branch when it
//    emitted (absorb_routed, termsum.rs:788), or its held-back branch when P
//    emitted into the partner (claim + drain_claims, termsum.rs:604-647).
//  unmatched emitted child: a new row; unmatched / unclaimed pending: dropped.
// Blocks are persistent: they take tiles in ticket order until all
// ceil((n + e) / T) tiles are done, so the grid does not depend on e.
// A tile starts as soon as its ticket is taken: taking tickets ahead of
// time (to prefetch) delays the aggregates that later tiles' look-backs wait
// for, which costs more than the prefetch saves.
// REC also writes the gather tape: every S row's output slot, and for every
// branch that reached the output (new row, hit, or claimed pending) the slot
// it landed in, under its source row.
template <int KIND, int W, bool REC>
__global__ __launch_bounds__(kBlock) void k_store_merge(GateArgs<W> g, StoreDev s, const uint64_t* n_ptr, ChildDev c,
                                                        const uint64_t* e_ptr, StoreDev out, const uint64_t* splits,
                                                        uint64_t* status, uint32_t* ticket, uint64_t* n_out,
                                                        uint64_t* n_out_host, uint64_t* e_host, uint32_t tag,
                                                        Counters* ctr, TapeDev tape) {
  constexpr int kI = StoreCfg<W>::items;
  constexpr uint64_t kT = StoreCfg<W>::tile;
  __shared__ StoreTileSmem<W> sm;
  const uint64_t n = *n_ptr, e = *e_ptr;
  if (n + e > out.cap) {  // refused by k_store_splits (overflow flag set)
    if (blockIdx.x == 0 && threadIdx.x == 0) publish_counts(n_out_host, e_host, tag, kPubRefused, e);
    return;
  }
  const uint64_t n_tiles = (n + e + kT - 1) / kT;
  if (n_tiles == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *n_out = 0;
      publish_counts(n_out_host, e_host, tag, 0, 0);
    }
    return;
  }
  unsigned long long hits = 0, emitted = 0, created = 0;
  double disc_l1 = 0.0, disc_max = 0.0;
  for (;;) {
    if (threadIdx.x == 0) sm.tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = sm.tile;
    if (tile >= n_tiles) break;
    const TileSplit sp = load_split(splits, tile);
    TileRegs<W, kI> r;
    tile_fetch<W, kI>(r, s, c, sp);
    const uint64_t i0 = sp.i0, j0 = sp.j0;
    const int nS = (int)(sp.i1 - sp.i0), nC = (int)(sp.j1 - sp.j0), tot = nS + nC;
#pragma unroll
    for (int k = 0; k < kI; ++k) {
      const int t = threadIdx.x + k * kBlock;
      if (t < tot) {
#pragma unroll
        for (int w = 0; w < W; ++w) sm.key[w][t] = r.key[k][w];
        sm.val[t] = r.val[k];
        if (t < nS) {
          Key<W> a;
#pragma unroll
          for (int w = 0; w < W; ++w) a.w[w] = r.key[k][w];
          sm.flag[t] = stage_store_row<KIND, W>(g, a);
        } else {
          sm.flag[t] = r.flag[k];
        }
      }
    }
    if (threadIdx.x == 0) {
      if (tot > (int)kT) {  // element kT (a child in practice: the pulled-in one)
        const bool fromS = (int)kT < nS;
        const uint64_t gi = fromS ? i0 + kT : j0 + (kT - nS);
        const uint64_t* kp = (fromS ? s.key : c.key) + gi;
        const uint64_t stride = fromS ? s.cap : c.cap;
        Key<W> a;
#pragma unroll
        for (int w = 0; w < W; ++w) sm.key[w][kT] = a.w[w] = kp[(uint64_t)w * stride];
        sm.val[kT] = (fromS ? s.coeff : c.val)[gi];
        sm.flag[kT] = fromS ? stage_store_row<KIND, W>(g, a) : c.flag[gi];
      }
    }
    __syncthreads();
    const int sw = tile_sig_word(sm.key, nS, nC);

    // this thread's range of the merged order; the last thread takes the rest
    const int dl = threadIdx.x * kI;
    const int dh = threadIdx.x == kBlock - 1 ? tot : (dl + kI < tot ? dl + kI : tot);
    int oslot[kI + 1];
    double ocoef[kI + 1];
    int ochild[kI + 1];  // REC: the child (local index) whose branch an S row absorbed
    uint32_t n_mine = 0;
    int is = 0, ic = 0;
    if (dl < dh) {
      int lo = dl > nC ? dl - nC : 0, hi = dl < nS ? dl : nS;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (lds_cmp(sm.key, nS + dl - mid - 1, mid, sw) >= 0)  // S first on equal keys
          lo = mid + 1;
        else
          hi = mid;
      }
      is = lo;
      ic = dl - lo;
    }
    // the S row taken last equals the current child (a child is absorbed by
    // an equal S row, which the merge order puts right before it)
    bool last_eq = dl < dh && is > 0 && ic < nC && lds_cmp(sm.key, is - 1, nS + ic, sw) == 0;
    // the head of each run in registers (word sw of the key, flag, value), so
    // a step costs one LDS round trip: the refill of the side it took
    uint64_t hs = 0, hc = 0;
    uint8_t fs = 0, fc = 0;
    double vs = 0.0, vc = 0.0;
    if (dl < dh) {
      if (is < nS) {
        hs = sm.key[sw][is];
        fs = sm.flag[is];
        vs = sm.val[is];
      }
      if (ic < nC) {
        hc = sm.key[sw][nS + ic];
        fc = sm.flag[nS + ic];
        vc = sm.val[nS + ic];
      }
    }
#pragma unroll
    for (int k = 0; k <= kI; ++k) {
      oslot[k] = -1;
      ocoef[k] = 0.0;
      ochild[k] = -1;
      if (dl + k >= dh) continue;
      // ord > 0: S row is < child ic; 0: equal keys (S first); < 0: the child first
      int ord;
      if (is >= nS)
        ord = -1;
      else if (ic >= nC)
        ord = 1;
      else if (hc != hs)
        ord = hc < hs ? -1 : 1;
      else
        ord = lds_cmp(sm.key, nS + ic, is, sw - 1);
      if (ord >= 0) {
        double cn = vs;
        const bool eq = ord == 0;
        if (fs & 1) {  // the row branches: c cos, plus what its partner sent
          if (g.sc) {  // batched: the instance's own cosine
            Key<W> a;
            a.w[W - 1] = sm.key[W - 1][is];
            cn = vs * row_rot<W>(g, a).c;
          } else {
            cn = vs * g.cos_t;
          }
          if (eq) {
            if (fc == kEmit) {
              cn = cn + vc;
              ++hits;
              if (REC) ochild[k] = ic;
            } else {
              // a held-back branch is claimed when this row's own branch was
              // emitted into the partner
              Key<W> a, q;
              double vv;
#pragma unroll
              for (int w = 0; w < W; ++w) a.w[w] = sm.key[w][is];
              const bool own = g.own_keep ? g.own_keep[i0 + is] != 0 : decide<KIND, W>(g, a, vs, q, vv) == kDecEmit;
              if (own) {
                cn = cn + vc;
                ++hits;
                ++emitted;
                if (REC) ochild[k] = ic;
              } else {
                disc_l1 += fabs(vc);
                disc_max = fmax(disc_max, fabs(vc));
              }
            }
          }
        }
        last_eq = eq;
        oslot[k] = is;
        ocoef[k] = cn;
        ++is;
        if (is < nS) {
          hs = sm.key[sw][is];
          fs = sm.flag[is];
          vs = sm.val[is];
        }
      } else {
        if (!last_eq) {
          if (fc == kEmit) {
            oslot[k] = nS + ic;
            ocoef[k] = vc;
            ++created;
          } else {
            disc_l1 += fabs(vc);
            disc_max = fmax(disc_max, fabs(vc));
          }
        }
        if (fc == kEmit) ++emitted;
        last_eq = false;
        ++ic;
        if (ic < nC) {
          hc = sm.key[sw][nS + ic];
          fc = sm.flag[nS + ic];
          vc = sm.val[nS + ic];
        }
      }
      n_mine += oslot[k] >= 0;
    }
    uint32_t total;
    uint32_t o = block_excl_scan(n_mine, &total);
    const uint32_t o_first = o;
#pragma unroll
    for (int k = 0; k <= kI; ++k) {
      if (oslot[k] < 0) continue;
      sm.oslot[o] = (uint16_t)oslot[k];
      sm.val[o] = ocoef[k];
      ++o;
    }
    if (threadIdx.x < 64) {
      const uint64_t ex = lookback(status, tile, total);
      if (threadIdx.x == 0) sm.base = ex;
    }
    __syncthreads();
    const uint64_t ob = sm.base;
    if (REC) {
      uint32_t oo = o_first;
#pragma unroll
      for (int k = 0; k <= kI; ++k) {
        if (oslot[k] < 0) continue;
        const uint32_t pos = (uint32_t)(ob + oo);
        ++oo;
        int ci = -1;
        if (oslot[k] < nS) {
          tape.out[i0 + oslot[k]] = pos;
          ci = ochild[k];
        } else {
          ci = oslot[k] - nS;
        }
        if (ci >= 0) {
          const uint32_t og = c.origin[j0 + ci];
          tape.child[og & ~kRecSign] = (og & kRecSign) | pos;
        }
      }
    }
    {
      // all LDS reads of this thread's outputs first, then the stores, so the
      // reads overlap instead of forming one round trip per output
      int li[kI + 1];
      uint64_t kv[kI + 1][W];
      double cv[kI + 1];
#pragma unroll
      for (int k = 0; k <= kI; ++k) {
        const int t = threadIdx.x + k * kBlock;
        li[k] = t < (int)total ? (int)sm.oslot[t] : 0;
        cv[k] = t < (int)total ? sm.val[t] : 0.0;
      }
#pragma unroll
      for (int k = 0; k <= kI; ++k)
#pragma unroll
        for (int w = 0; w < W; ++w) kv[k][w] = sm.key[w][li[k]];
#pragma unroll
      for (int k = 0; k <= kI; ++k) {
        const int t = threadIdx.x + k * kBlock;
        if (t < (int)total) {
#pragma unroll
          for (int w = 0; w < W; ++w) out.key[(uint64_t)w * out.cap + ob + t] = kv[k][w];
          out.coeff[ob + t] = cv[k];
        }
      }
    }
    if (tile == n_tiles - 1 && threadIdx.x == 0) {
      *n_out = ob + total;
      publish_counts(n_out_host, e_host, tag, ob + total, e);
    }
    __syncthreads();
  }
  if (ctr) {
    hits = wave_sum_u64(hits);
    emitted = wave_sum_u64(emitted);
    created = wave_sum_u64(created);
    double l1 = disc_l1, mx = disc_max;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      l1 += __shfl_xor(l1, o, 64);
      mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
      Counters* cs = ctr_slot(ctr);
      if (hits) atomicAdd(&cs->hits, hits);
      if (emitted) atomicAdd(&cs->emitted, emitted);
      if (created) atomicAdd(&cs->created, created);
      if (l1 != 0.0) atomicAdd(&cs->disc_l1, l1);
      if (mx > 0.0) atomicMax(&cs->disc_max_bits, (unsigned long long)__double_as_longlong(mx));
    }
  }
}

// ------------------------------------------------------ small-store runs

// A store of at most SmallCfg<W>::rows rows lives in one workgroup's LDS for
// a whole run of gates: rows stay where they are (new rows are appended) and
// an open-addressed table maps keys to rows, so a gate is two passes over
// the rows with no sort and no launch of its own.  Per gate the result is the
// one the sorted-merge path computes: every row gets c cos plus at most one
// branch (its partner's, or its own held-back one by the pair rule), and the
// unmatched emitted branches become new rows.  The run ends sorted, so the
// store the next gate or readout sees is the same array.
template <int KIND, int W>
struct SmallRun {
  static constexpr int R = SmallCfg<W>::rows;
  static constexpr int T = R + R / 2;            // table slots (load <= 2/3)
  static constexpr int kArgs = 8;                // gates staged per refill
  uint64_t key[W][R];
  double coeff[R];
  uint16_t table[T + (T & 1)];                   // row + 1, 0 = empty
  uint32_t own[(R + 31) / 32];                   // per row: its own branch was emitted (this gate)
  GateArgs<W> args[kArgs];
  uint8_t kinds[kArgs];
  uint32_t created;
};

template <int W>
__device__ __forceinline__ uint32_t small_slot(const Key<W>& k, uint32_t T) {
  uint64_t h = 0x9e3779b97f4a7c15ULL;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    h ^= k.w[w] + 0x632be59bd9b4e019ULL + (h << 6) + (h >> 2);
    h *= 0xbf58476d1ce4e5b9ULL;
  }
  const uint32_t x = (uint32_t)(h >> 32) ^ (uint32_t)h;
  return (uint32_t)(((uint64_t)x * T) >> 32);
}

template <int KIND, int W>
__device__ __forceinline__ Key<W> small_key(const SmallRun<KIND, W>& sm, int r) {
  Key<W> k;
#pragma unroll
  for (int w = 0; w < W; ++w) k.w[w] = sm.key[w][r];
  return k;
}

// 16-bit table slot i, read and claimed through its aligned 32-bit word.
template <int KIND, int W>
__device__ __forceinline__ uint32_t small_slot_get(const SmallRun<KIND, W>& sm, uint32_t i) {
  const uint32_t* word = reinterpret_cast<const uint32_t*>(sm.table) + (i >> 1);
  const uint32_t v = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  return (v >> ((i & 1u) * 16)) & 0xFFFFu;
}

// Insert row r (its key already written); keys are unique.
template <int KIND, int W>
__device__ __forceinline__ void small_insert(SmallRun<KIND, W>& sm, int r) {
  constexpr uint32_t T = SmallRun<KIND, W>::T;
  uint32_t i = small_slot<W>(small_key(sm, r), T);
  for (;;) {
    uint32_t* word = reinterpret_cast<uint32_t*>(sm.table) + (i >> 1);
    const uint32_t sh = (i & 1u) * 16;
    uint32_t old = __hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    bool placed = false;
    while (((old >> sh) & 0xFFFFu) == 0) {
      const uint32_t prev = atomicCAS(word, old, old | ((uint32_t)(r + 1) << sh));
      if (prev == old) {
        placed = true;
        break;
      }
      old = prev;  // the other half of the word changed, or this slot was taken
    }
    if (placed) return;
    i = i + 1 == T ? 0 : i + 1;
  }
}

// Row holding key k, or -1.  Rows inserted during the same gate are never
// looked for (branches are unique), and inserts only fill empty slots, so a
// probe sees every row present at the gate's start.
template <int KIND, int W>
__device__ __forceinline__ int small_find(const SmallRun<KIND, W>& sm, const Key<W>& k) {
  constexpr uint32_t T = SmallRun<KIND, W>::T;
  uint32_t i = small_slot<W>(k, T);
  for (;;) {
    const uint32_t e = small_slot_get(sm, i);
    if (e == 0) return -1;
    if (key_eq<W>(small_key(sm, (int)e - 1), k)) return (int)e - 1;
    i = i + 1 == T ? 0 : i + 1;
  }
}

__device__ __forceinline__ uint32_t block_sum_small(uint32_t v, uint32_t* red) {
  v = (uint32_t)wave_sum_u64(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  uint32_t t = 0;
#pragma unroll
  for (int i = 0; i < kSmallThreads / 64; ++i) t += red[i];
  __syncthreads();
  return t;
}

// Compare-exchange rows i < l into ascending order.
template <int KIND, int W>
__device__ __forceinline__ void small_order(SmallRun<KIND, W>& sm, int i, int l) {
  if (!key_less<W>(small_key(sm, l), small_key(sm, i))) return;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const uint64_t t = sm.key[w][i];
    sm.key[w][i] = sm.key[w][l];
    sm.key[w][l] = t;
  }
  const double c = sm.coeff[i];
  sm.coeff[i] = sm.coeff[l];
  sm.coeff[l] = c;
}

// One run of gates over a small store, in one workgroup.  kinds[j]: 0 =
// branch, 1 = phase-only (anticommuting rows scaled by cos).  The run stops
// before a gate whose new rows would not fit; *done = gates applied,
// n_after[j] = live rows after gate j, and the store (sorted) and *n_ptr are
// written back.
template <int KIND, int W>
__global__ __launch_bounds__(kSmallThreads) void k_small_run(const GateArgs<W>* gates, const uint8_t* kinds,
                                                             int n_gates, StoreDev s, uint64_t* n_ptr,
                                                             uint64_t* n_after, uint32_t* done, Counters* ctr) {
  using SM = SmallRun<KIND, W>;
  constexpr int R = SM::R, T = SM::T;
  constexpr int kPer = (R + kSmallThreads - 1) / kSmallThreads;
  __shared__ SM sm;
  __shared__ uint32_t red[kSmallThreads / 64];
  const int tid = threadIdx.x;
  int n = (int)*n_ptr;
  for (int r = tid; r < n; r += kSmallThreads) {
#pragma unroll
    for (int w = 0; w < W; ++w) sm.key[w][r] = s.key[(uint64_t)w * s.cap + r];
    sm.coeff[r] = s.coeff[r];
  }
  for (int i = tid; i < (T + 1) / 2; i += kSmallThreads) reinterpret_cast<uint32_t*>(sm.table)[i] = 0;
  __syncthreads();
  for (int r = tid; r < n; r += kSmallThreads) small_insert(sm, r);

  unsigned long long visited = 0, declined = 0, dropped = 0, hits = 0, emitted = 0, created = 0;
  double disc_l1 = 0.0, disc_max = 0.0;
  int j = 0;
  for (; j < n_gates; ++j) {
    if (j % SM::kArgs == 0) {  // stage the next gates' arguments
      __syncthreads();
      const int m = min(SM::kArgs, n_gates - j);
      const int cnt = m * (int)(sizeof(GateArgs<W>) / sizeof(uint32_t));
      const uint32_t* src = reinterpret_cast<const uint32_t*>(gates + j);
      uint32_t* dst = reinterpret_cast<uint32_t*>(sm.args);
      for (int i = tid; i < cnt; i += kSmallThreads) dst[i] = src[i];
      if (tid < m) sm.kinds[tid] = kinds[j + tid];
    }
    __syncthreads();
    const GateArgs<W>& g = sm.args[j % SM::kArgs];
    if (sm.kinds[j % SM::kArgs]) {  // phase-only (coeff.rs:200, termsum.rs:551)
      for (int r = tid; r < n; r += kSmallThreads)
        if (row_anti<KIND, W>(g, small_key(sm, r))) sm.coeff[r] = sm.coeff[r] * g.cos_t;
      if (tid == 0) n_after[j] = (uint64_t)n;
      continue;
    }
    // decisions first, without writing, so a gate that would not fit is not begun
    Key<W> q[kPer];
    double v[kPer];
    int dec[kPer];
    uint32_t my_emits = 0;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int r = tid + i * kSmallThreads;
      dec[i] = -1;
      v[i] = 0.0;
      if (r >= n) continue;
      const Key<W> p = small_key(sm, r);
      if (!row_anti<KIND, W>(g, p)) continue;  // commutes
      dec[i] = decide<KIND, W>(g, p, sm.coeff[r], q[i], v[i]);
      my_emits += dec[i] == kDecEmit;
    }
    for (int i = tid; i < (R + 31) / 32; i += kSmallThreads) sm.own[i] = 0;
    if (tid == 0) sm.created = 0;
    const uint32_t emits = block_sum_small(my_emits, red);
    if (n + (int)emits > R) break;
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int r = tid + i * kSmallThreads;
      if (dec[i] < 0) continue;
      ++visited;
      sm.coeff[r] = sm.coeff[r] * g.cos_t;
      if (dec[i] == kDecEmit) atomicOr(&sm.own[r >> 5], 1u << (r & 31));
      declined += dec[i] == kDecPending;
      dropped += dec[i] == kDecDrop;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      if (dec[i] != kDecEmit && dec[i] != kDecPending) continue;
      const int p = small_find(sm, q[i]);
      if (p >= 0) {
        if (dec[i] == kDecEmit || ((sm.own[p >> 5] >> (p & 31)) & 1u)) {  // absorbed, or claimed (pair rule)
          sm.coeff[p] = sm.coeff[p] + v[i];
          ++hits;
          ++emitted;
        } else {
          disc_l1 += fabs(v[i]);
          disc_max = fmax(disc_max, fabs(v[i]));
        }
      } else if (dec[i] == kDecEmit) {
        const int r = n + (int)atomicAdd(&sm.created, 1u);
#pragma unroll
        for (int w = 0; w < W; ++w) sm.key[w][r] = q[i].w[w];
        sm.coeff[r] = v[i];
        __threadfence_block();
        small_insert(sm, r);
        ++created;
        ++emitted;
      } else {
        disc_l1 += fabs(v[i]);
        disc_max = fmax(disc_max, fabs(v[i]));
      }
    }
    __syncthreads();
    n += (int)sm.created;
    if (tid == 0) n_after[j] = (uint64_t)n;
  }
  if (tid == 0) *done = (uint32_t)j;
  __syncthreads();

  // sort: bitonic network of ascending compare-exchanges only (each merge
  // starts with the mirrored step), so the virtual +inf rows past n never
  // move and any pair reaching past n is skipped
  int P = 1;
  while (P < n) P <<= 1;
  for (int k = 2; k <= P; k <<= 1) {
    for (int i = tid; i < P / 2; i += kSmallThreads) {  // mirrored step within blocks of k
      const int lo = (i / (k / 2)) * k + (i % (k / 2));
      const int hi = (lo / k) * k + (k - 1 - (lo % k));
      if (hi < n) small_order(sm, lo, hi);
    }
    __syncthreads();
    for (int jj = k / 4; jj > 0; jj >>= 1) {
      for (int i = tid; i < P / 2; i += kSmallThreads) {  // half-cleaners
        const int lo = (i / jj) * 2 * jj + (i % jj);
        const int hi = lo + jj;
        if (hi < n) small_order(sm, lo, hi);
      }
      __syncthreads();
    }
  }
  for (int r = tid; r < n; r += kSmallThreads) {
#pragma unroll
    for (int w = 0; w < W; ++w) s.key[(uint64_t)w * s.cap + r] = sm.key[w][r];
    s.coeff[r] = sm.coeff[r];
  }
  if (tid == 0) *n_ptr = (uint64_t)n;
  if (ctr) {
    Counters* cs = ctr_slot(ctr);
    const unsigned long long vals[6] = {visited, declined, dropped, hits, emitted, created};
    unsigned long long* dsts[6] = {&cs->visited, &cs->declined, &cs->dropped_weight, &cs->hits, &cs->emitted,
                                   &cs->created};
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const unsigned long long t = wave_sum_u64(vals[i]);
      if ((tid & 63) == 0 && t) atomicAdd(dsts[i], t);
    }
    double l1 = disc_l1, mx = disc_max;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      l1 += __shfl_xor(l1, o, 64);
      mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    if ((tid & 63) == 0) {
      if (l1 != 0.0) atomicAdd(&cs->disc_l1, l1);
      if (mx > 0.0) atomicMax(&cs->disc_max_bits, (unsigned long long)__double_as_longlong(mx));
    }
  }
}

// ------------------------------------------------------- noise + reclaim

// Layer-start pass (crates/pauli/src/engine.rs:119-160): c *= t[min(w,
// len-1)] (termsum.rs:321), then keep rows the full cutoff still admits
// (termsum.rs:152, :472).  Order is preserved, so the store stays sorted.
template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_noise_compact(NoiseArgs a, StoreDev s, uint64_t n, StoreDev out,
                                                          uint64_t* status, uint32_t* ticket, uint64_t* n_out,
                                                          uint64_t n_tiles, int n_units, TapeDev tape,
                                                          uint64_t* n_host, uint32_t tag) {
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_base;
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint64_t r0 = (uint64_t)tile * kTile + (uint64_t)threadIdx.x * kItems;
  Key<W> k[kItems];
  double cf[kItems];
  bool keep[kItems];
  uint32_t nk = 0;
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint64_t r = r0 + it;
    keep[it] = false;
    if (r >= n) continue;
    k[it] = load_key<W>(s.key, s.cap, r);
    double c = s.coeff[r];
    uint32_t w = 0;
    double fct = 1.0;  // linear factor of this row (tape)
    if (a.table || a.has_maxw) w = (uint32_t)weight<KIND, W>(unit_bits<W>(k[it], a.unit_mask), n_units);
    if (a.table) {
      const double t = a.table[w < a.len - 1 ? w : a.len - 1];
      c = c * t;
      fct = t;
    }
    if (a.factors) {
      c = c * a.factors[r];
      fct = fct * a.factors[r];
    }
    if (tape.factor) tape.factor[r] = fct;
    bool ok = true;
    if (a.inst_drop && a.inst_drop[k[it].w[W - 1] >> a.inst_shift]) {
      ok = false;
    } else if (a.keep_flags) {
      ok = a.keep_flags[r] != 0;
    } else if (a.reclaim) {
      if (a.has_maxw && w > a.max_weight) ok = false;
      if (a.has_min && !(fabs(c) >= a.min_coeff)) ok = false;
    }
    cf[it] = c;
    keep[it] = ok;
    nk += ok;
  }
  uint32_t total;
  const uint32_t mybase = block_excl_scan(nk, &total);
  if (threadIdx.x < 64) {
    const uint64_t ex = lookback(status, tile, total);
    if (threadIdx.x == 0) s_base = ex;
  }
  __syncthreads();
  uint64_t o = s_base + mybase;
#pragma unroll
  for (int it = 0; it < kItems; ++it) {
    const uint64_t r = r0 + it;
    if (tape.out && r < n) tape.out[r] = keep[it] ? (uint32_t)o : kRecNoChild;
    if (!keep[it]) continue;
    store_key<W>(out.key, out.cap, o, k[it]);
    out.coeff[o] = cf[it];
    ++o;
  }
  if (tile == n_tiles - 1 && threadIdx.x == 0) {
    *n_out = s_base + total;
    publish_counts(n_host, nullptr, tag, s_base + total, 0);
  }
}

// -------------------------------------------------- sparse in-place gate
//
// A gate changes only the rows that anticommute with its generator, a few per
// cent of the store on the benchmark circuits, so this path leaves every other
// row where it lies instead of rewriting the sorted store.  Two kernels:
//   k_sparse_gather  per block of rows, the anticommuting ones (found from the
//                    key words the anticommutation test needs) go to a dense
//                    list of (key, c, branch value, decision, row, table
//                    slot), at places the block reserves with one atomic, and
//                    each is entered in an open-addressed hash table keyed by
//                    its key;
//   k_sparse_join    every listed row P looks its partner G P up in the table
//                    (the partner of an anticommuting row anticommutes, so it
//                    is listed when it is live), rewrites its own coefficient
//                    in place as c cos plus what the partner sent it (the pair
//                    rule included), and appends its own branch as a new row
//                    when the partner is absent.
// Per key the arithmetic is the store merge's (termsum.rs:604-695, 788-800):
// c cos, then + the partner's emitted branch, or + its held-back branch when
// this row's own branch was emitted.  Each key's result depends only on its
// pair, so the list order and the order new rows are appended in (both set by
// atomics) change no value; the store is no longer sorted, and readouts that
// depend on row order sort it first (Engine::sort_store).
//
// Lists and tables alternate between gates: gate k fills list and table k mod
// 2, and its join also clears the slots gate k - 1 filled in the other table
// (the gather keeps every entry's slot in a u32 list beside it), so a gather
// always inserts into an empty table without reading a slot first.  The table has
// sparse_table_slots(n_in) slots: at most half full even when every row
// anticommutes, a few per cent on the benchmark circuits.

template <int W>
__device__ __forceinline__ uint64_t sparse_hash(const Key<W>& k) {
  uint64_t h = 0x243F6A8885A308D3ULL;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    h = (h ^ k.w[i]) * 0x9E3779B97F4A7C15ULL;
    h ^= h >> 31;
  }
  h ^= h >> 33;
  h *= 0xFF51AFD7ED558CCDULL;
  h ^= h >> 33;
  return h;
}

// Table slot of list entry idx with key hash h: a valid bit, 16 bits of the
// hash (so most mismatches need no key read) and the index.  0 = empty.
constexpr uint64_t kSlotValid = 1ULL << 63;
__device__ __forceinline__ uint64_t sparse_slot(uint64_t h, uint64_t idx) {
  return kSlotValid | ((h >> 48) << 32) | idx;
}
__device__ __forceinline__ bool slot_fp_match(uint64_t sl, uint64_t h) {
  return (sl & kSlotValid) && ((sl >> 32) & 0xFFFFULL) == (h >> 48);
}
// entry meta: source row (bits 0-31) | decision (32-33) | -Im(phase) < 0 (34)
constexpr uint64_t kEntNeg = 1ULL << 34;

constexpr int kSparseWaves = kBlock / 64;
// The join's tile word: finished tiles (high 24 bits) | new rows so far (low 40).
constexpr int kJoinCountBits = 40;
constexpr uint64_t kJoinCountMask = (1ULL << kJoinCountBits) - 1;
// A sub-list record's counter q and its header (S, R).
__device__ __forceinline__ unsigned long long* rec_len(uint64_t* rec, uint64_t q) {
  return reinterpret_cast<unsigned long long*>(rec + q * kRecStride);
}
__device__ __forceinline__ uint64_t* rec_hdr(uint64_t* rec) { return rec + (uint64_t)kSubLists * kRecStride; }


// Exclusive prefix, in place, over cnt[kI][kSparseWaves] in (item, wave)
// order, the order rows (or list entries) base + it kBlock + tid have;
// returns the block total.  Every thread calls it.
template <int kI>
__device__ __forceinline__ uint32_t slab_prefix(uint32_t (&cnt)[kI][kSparseWaves]) {
  __shared__ uint32_t tot;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (int i = 0; i < kI; ++i)
      for (int w = 0; w < kSparseWaves; ++w) {
        const uint32_t t = cnt[i][w];
        cnt[i][w] = acc;
        acc += t;
      }
    tot = acc;
  }
  __syncthreads();
  return tot;
}

// One list entry from a store row (its slot filled in after the insert).
template <int KIND, int W>
__device__ __forceinline__ SparseEnt<W> sparse_entry(const GateArgs<W>& g, const Key<W>& p, double c, uint64_t r,
                                                     int& dec) {
  Key<W> q;
  double v = 0.0, ph = 1.0;
  dec = decide<KIND, W>(g, p, c, q, v, &ph);
  SparseEnt<W> e;
#pragma unroll
  for (int w = 0; w < SparseEnt<W>::kWords; ++w) e.w[w] = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) e.w[w] = p.w[w];
  e.w[W] = (uint64_t)__double_as_longlong(c);
  e.w[W + 1] = (uint64_t)__double_as_longlong(v);
  e.w[W + 2] = r | ((uint64_t)dec << 32) | (ph < 0.0 ? kEntNeg : 0ULL);
  return e;
}

// Insert list entry a (key hash h) into an empty-before-this-gate table;
// keys are unique, so an insert only looks for an empty slot.  Returns the slot.
__device__ __forceinline__ uint64_t sparse_insert(uint64_t* table, uint64_t mask, uint64_t h, uint64_t a) {
  const unsigned long long val = sparse_slot(h, a);
  uint64_t i = h & mask;
  while (atomicCAS(reinterpret_cast<unsigned long long*>(table + i), 0ULL, val) != 0ULL) i = (i + 1) & mask;
  return i;
}

// Block b takes rows [b T, (b + 1) T): tests each for anticommutation with
// the words the test needs, lists the anticommuting ones in LDS, reserves
// their list entries with one atomic on its sub-list's counter in rec, and
// writes and inserts them, each thread handling about one (a few per cent of
// the rows anticommute); slots[a] keeps entry a's table slot.  Block 0 also
// clears the join word and writes the record's header (S, R).
template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_sparse_gather(GateArgs<W> g, StoreDev s, const uint64_t* n_ptr,
                                                          SparseEnt<W>* ents, uint32_t* slots, uint64_t ent_cap,
                                                          uint64_t* rec, uint64_t* table, uint64_t table_cap,
                                                          unsigned long long* join_word, Counters* ctr) {
  __shared__ uint32_t wcnt[kSparseItems][kSparseWaves];
  __shared__ uint16_t rows[kSparseTile];
  __shared__ uint64_t s_base;
  __shared__ uint32_t wstat[kSparseWaves][2];
  const uint64_t n = *n_ptr;
  const SubLists sl = sparse_sublists(n);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *join_word = 0;
    rec_hdr(rec)[0] = sl.S;
    rec_hdr(rec)[1] = sl.R;
  }
  const uint64_t T = sparse_table_slots(n);
  if (T > table_cap) return;  // the join refuses the gate
  const uint64_t base = (uint64_t)blockIdx.x * kSparseTile;
  if (base >= n) return;  // the grid is sized by the host's bound of n, which a gate in flight may overstate
  const uint64_t mask = T - 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
  // the anticommutation test: Pauli parity(P & fold), Majorana
  // |P & G| + [|G| odd] |P|, from the needed words only
  uint64_t x[kSparseItems][W];
#pragma unroll
  for (int it = 0; it < kSparseItems; ++it) {
    const uint64_t r = base + (uint64_t)it * kBlock + threadIdx.x;
#pragma unroll
    for (int w = 0; w < W; ++w)
      x[it][w] = (r < n && (word_needed<KIND, W>(g, w) || (g.sc && w == W - 1))) ? s.key[(uint64_t)w * s.cap + r]
                                                                                 : 0;
  }
  uint32_t rank[kSparseItems];
  bool anti[kSparseItems];
#pragma unroll
  for (int it = 0; it < kSparseItems; ++it) {
    const uint64_t r = base + (uint64_t)it * kBlock + threadIdx.x;
    Key<W> p;
#pragma unroll
    for (int w = 0; w < W; ++w) p.w[w] = x[it][w];
    // batched engines: the rows of a frozen instance count as commuting
    anti[it] = r < n && row_anti<KIND, W>(g, p) && !row_frozen<W>(g, p);
    const uint64_t m = __ballot(anti[it]);
    rank[it] = popc64(m & lt);
    if (lane == 0) wcnt[it][wave] = popc64(m);
  }
  const uint32_t cnt = slab_prefix<kSparseItems>(wcnt);
#pragma unroll
  for (int it = 0; it < kSparseItems; ++it)
    if (anti[it]) rows[wcnt[it][wave] + rank[it]] = (uint16_t)(it * kBlock + threadIdx.x);
  __syncthreads();
  // the reservation (in this block's sub-list) is in flight while the first
  // round of rows loads
  const uint64_t sq = blockIdx.x % sl.S;
  unsigned long long reserved = 0;
  if (threadIdx.x == 0 && cnt) reserved = atomicAdd(rec_len(rec, sq), (unsigned long long)cnt);
  uint32_t pend = 0, drop = 0;
  SparseEnt<W> e0;
  int dec0 = kDecNone;
  const bool has0 = threadIdx.x < cnt;
  uint64_t h0 = 0;
  if (has0) {
    const uint64_t r = base + rows[threadIdx.x];
    const Key<W> p = load_key<W>(s.key, s.cap, r);
    e0 = sparse_entry<KIND, W>(g, p, s.coeff[r], r, dec0);
    h0 = sparse_hash<W>(p);
  }
  if (threadIdx.x == 0) s_base = reserved;
  __syncthreads();
  // the region's room holds every row of its blocks; else the join refuses the gate
  const bool room = s_base + cnt <= sl.R && sl.S * sl.R <= ent_cap;
  const uint64_t b0 = sq * sl.R + s_base;
  if (has0) {
    pend += dec0 == kDecPending;
    drop += dec0 == kDecDrop;
    if (room) {
      const uint64_t a = b0 + threadIdx.x;
      ents[a] = e0;
      slots[a] = (uint32_t)sparse_insert(table, mask, h0, a);
    }
  }
  for (uint32_t li = threadIdx.x + kBlock; li < cnt; li += kBlock) {  // more than one row per thread: rare
    const uint64_t r = base + rows[li];
    const Key<W> p = load_key<W>(s.key, s.cap, r);
    int dec;
    SparseEnt<W> e = sparse_entry<KIND, W>(g, p, s.coeff[r], r, dec);
    pend += dec == kDecPending;
    drop += dec == kDecDrop;
    if (room) {
      const uint64_t a = b0 + li;
      ents[a] = e;
      slots[a] = (uint32_t)sparse_insert(table, mask, sparse_hash<W>(p), a);
    }
  }
  pend = (uint32_t)wave_sum_u64(pend);
  drop = (uint32_t)wave_sum_u64(drop);
  if (lane == 0) {
    wstat[wave][0] = pend;
    wstat[wave][1] = drop;
  }
  __syncthreads();
  if (threadIdx.x == 0 && ctr) {
    uint32_t b = 0, d = 0;
    for (int w = 0; w < kSparseWaves; ++w) {
      b += wstat[w][0];
      d += wstat[w][1];
    }
    Counters* cs = ctr_slot(ctr);
    if (cnt) atomicAdd(&cs->visited, (unsigned long long)cnt);
    if (b) atomicAdd(&cs->declined, (unsigned long long)b);
    if (d) atomicAdd(&cs->dropped_weight, (unsigned long long)d);
  }
}

// ---- sub-lists and the appended-row stage shared by the joins

// Entry index of list position g: pre[0..S] is the prefix of the sub-list
// lengths (pre[q] <= g < pre[q + 1] for the sub-list q holding it).
__device__ __forceinline__ uint64_t sub_entry(const uint64_t* pre, uint64_t S, uint64_t R, uint64_t g) {
  uint64_t lo = 0, hi = S;
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (pre[mid] <= g)
      lo = mid;
    else
      hi = mid;
  }
  return lo * R + (g - pre[lo]);
}

// pre[0..S] = prefix of the S sub-list lengths of rec (wave 0 of the block);
// every thread calls it.
__device__ __forceinline__ void sub_prefix(const uint64_t* rec, uint64_t S, uint64_t* pre) {
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    // every counter is loaded (those past S are zero), so the load does not
    // wait for S, which comes from the live count's own load
    uint64_t v = rec[(uint64_t)lane * kRecStride];
    v = (uint64_t)lane < S ? v : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, o, 64);
      const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), o, 64);
      if (lane >= o) v += ((uint64_t)hi << 32) | lo;
    }
    pre[lane + 1] = v;
    if (lane == 0) pre[0] = 0;
  }
  __syncthreads();
}

// Rows a join block appends are staged in LDS and written after one
// reservation on the join word per flush (a flush when a tile's rows would
// not fit, and one at the block's end, which also counts its tiles), instead
// of one reservation per tile.  R: the most rows one tile appends (a larger
// stage would only cost occupancy); kRec: the join records a tape.
template <int W_, int R, bool kRec>
struct JoinStage {
  static constexpr int W = W_;
  static constexpr int kRows = R;
  static constexpr bool kRecords = kRec;
  uint64_t key[W][kRows];
  double c[kRows];
  uint32_t rec[kRec ? kRows : 1];  // recording: the tape record of the branch that made the row | kRecSign
  uint64_t pre[kSubLists + 1], ppre[kSubLists + 1];
  uint64_t pS, pR;  // the previous pass's sub-lists
  uint64_t base;
  uint32_t len;
  bool last;
};
constexpr int kJoinTilesPerBlock = 1;

// Join blocks that take tiles: about kJoinTilesPerBlock tiles each, but at
// least one per CU while there are tiles for them.
__device__ __forceinline__ uint64_t join_blocks(uint64_t n_tiles) {
  uint64_t a = (n_tiles + kJoinTilesPerBlock - 1) / kJoinTilesPerBlock;
  const uint64_t lo = n_tiles < 256 ? n_tiles : 256;
  if (a < lo) a = lo;
  return a < gridDim.x ? a : gridDim.x;
}

// The join prologue: prefixes of this pass's and the previous pass's
// sub-lists, the next pass's counters zeroed.  Returns this pass's list
// length.
template <class Stage>
__device__ uint64_t join_prologue(const uint64_t* rec, SubLists sl, const uint64_t* prev_rec, uint64_t* next_rec,
                                  Stage& sh) {
  const uint64_t* ph = prev_rec + (uint64_t)kSubLists * kRecStride;
  sh.pS = ph[0] <= (uint64_t)kSubLists ? ph[0] : 0;
  sh.pR = ph[1];
  sub_prefix(rec, sl.S, sh.pre);
  sub_prefix(prev_rec, sh.pS, sh.ppre);
  if (blockIdx.x == 0 && threadIdx.x < kSubLists) next_rec[(uint64_t)threadIdx.x * kRecStride] = 0;
  return sh.pre[sl.S];
}

// The previous pass's table slots cleared, grid-stride over its list: every
// block calls it on its way out, after its own work, so the random clearing
// writes trail the join's loads instead of delaying them.
template <class Stage>
__device__ void clear_prev(const Stage& sh, const uint32_t* prev_slots, uint64_t* prev_table, uint64_t table_cap,
                           uint64_t ent_cap, bool have_first = false, uint32_t first = 0) {
  const uint64_t pS = sh.pS, pR = sh.pR;
  const uint64_t pl = pS ? sh.ppre[pS] : 0, stride = (uint64_t)gridDim.x * kBlock;
  const uint64_t i0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  for (uint64_t i = i0; i < pl; i += stride) {
    uint64_t slot;
    if (have_first && i == i0) {
      slot = first;
    } else {
      const uint64_t a = sub_entry(sh.ppre, pS, pR, i);
      if (a >= ent_cap) continue;
      slot = prev_slots[a];
    }
    if (slot < table_cap) prev_table[slot] = 0;
  }
}
// The slot clear_prev takes first in this thread, loaded before the join's
// work so that a small pass does not wait for it at the end (~0u: none).
template <class Stage>
__device__ uint32_t prev_slot_first(const Stage& sh, const uint32_t* prev_slots, uint64_t ent_cap) {
  const uint64_t pS = sh.pS, pl = pS ? sh.ppre[pS] : 0;
  const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= pl) return ~0u;
  const uint64_t a = sub_entry(sh.ppre, pS, sh.pR, i);
  return a < ent_cap ? prev_slots[a] : ~0u;
}

// Write the staged rows after the n old rows at a place reserved with one
// atomic on the join word; the final flush also adds the block's finished
// tiles, and the block that finishes the last tile publishes the count.
template <class Stage>
__device__ void stage_flush(Stage& sh, bool final, uint64_t tiles_done, uint64_t n_tiles, uint64_t n, StoreDev s,
                            unsigned long long* join_word, uint64_t* n_out, uint64_t* n_out_host, uint64_t* e_host,
                            uint32_t pub_tag, uint64_t total, SparseRec* tape = nullptr) {
  constexpr int W = Stage::W;
  __syncthreads();
  const uint32_t L = sh.len;
  if (threadIdx.x == 0) {
    const unsigned long long add = (final ? (unsigned long long)tiles_done << kJoinCountBits : 0ULL) | L;
    const unsigned long long old = add ? atomicAdd(join_word, add) : 0ULL;
    sh.base = old & kJoinCountMask;
    sh.last = final && (old >> kJoinCountBits) + tiles_done == n_tiles;
  }
  __syncthreads();
  const uint64_t ob = n + sh.base;
  for (uint32_t i = threadIdx.x; i < L; i += kBlock) {
#pragma unroll
    for (int w = 0; w < W; ++w) s.key[(uint64_t)w * s.cap + ob + i] = sh.key[w][i];
    s.coeff[ob + i] = sh.c[i];
    if (Stage::kRecords && tape) {  // the branch that made the row landed here
      const uint32_t rc = sh.rec[i];
      tape[rc & ~kRecSign].child = (uint32_t)(ob + i) | (rc & kRecSign);
    }
  }
  if (sh.last && threadIdx.x == 0) {  // every other block has reserved its rows: the count is final
    *n_out = ob + L;
    publish_counts(n_out_host, e_host, pub_tag, ob + L, total);
  }
  __syncthreads();
  if (threadIdx.x == 0) sh.len = 0;
  __syncthreads();
}

// Every list entry P: partner Q = G P by table lookup; P's row gets
// c cos + (Q's branch if Q emitted, or if Q held it back and P emitted), in
// place; P's own branch becomes a new row when Q is absent and P emitted, and
// goes to the discard statistics when Q is absent and P held it back.  Blocks
// stride over tiles of kJoinTile list positions (mapped to entries through
// the sub-list prefix); each thread issues the loads of its kJoinItems
// entries together, phase by phase (entries, table slots, candidate
// partners).  New rows go through the block's LDS stage (stage_flush).  The
// blocks also clear the slots the previous sparse pass filled in the other
// table and the counters of the pass after this one.
template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_sparse_join(GateArgs<W> g, const SparseEnt<W>* ents, uint64_t ent_cap,
                                                        const uint64_t* rec, const uint64_t* table,
                                                        uint64_t table_cap, const uint32_t* prev_slots,
                                                        const uint64_t* prev_rec, uint64_t* prev_table,
                                                        uint64_t* next_rec, StoreDev s, const uint64_t* n_ptr,
                                                        uint64_t* n_out, uint64_t* n_out_host, uint64_t* e_host,
                                                        uint32_t pub_tag, unsigned long long* join_word,
                                                        Counters* ctr, SparseRec* tape, uint32_t* tape_inst) {
  using Stage = JoinStage<W, kJoinTile, true>;  // a tile appends at most one row per entry
  __shared__ Stage sh;
  __shared__ uint32_t wcnt[kJoinItems][kSparseWaves];
  const uint64_t n = *n_ptr, T = sparse_table_slots(n);
  const SubLists sl = sparse_sublists(n);
  const uint64_t total =
      join_prologue(rec, sl, prev_rec, next_rec, sh);
  if (sl.S * sl.R > ent_cap || total > n || n + total > s.cap || T > table_cap) {  // a capacity bound was violated
    if (blockIdx.x == 0 && threadIdx.x == 0) publish_counts(n_out_host, e_host, pub_tag, kPubRefused, total);
    clear_prev(sh, prev_slots, prev_table, table_cap, ent_cap);
    return;
  }
  const uint64_t n_tiles = (total + kJoinTile - 1) / kJoinTile;
  if (n_tiles == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *n_out = n;
      publish_counts(n_out_host, e_host, pub_tag, n, 0);
    }
    clear_prev(sh, prev_slots, prev_table, table_cap, ent_cap);
    return;
  }
  const uint64_t active = join_blocks(n_tiles);
  if (blockIdx.x >= active) {
    clear_prev(sh, prev_slots, prev_table, table_cap, ent_cap);
    return;
  }
  // narrow keys leave the registers for it; wide ones would spill
  constexpr bool kPrefetch = W <= 2;
  const uint32_t first_slot = kPrefetch ? prev_slot_first(sh, prev_slots, ent_cap) : 0u;
  if (threadIdx.x == 0) sh.len = 0;
  const uint64_t mask = T - 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
  unsigned long long hits = 0, emitted = 0, created = 0;
  double disc_l1 = 0.0, disc_max = 0.0;
  uint64_t tiles_done = 0;
  for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += active) {
    Key<W> q[kJoinItems];
    double c[kJoinItems], v[kJoinItems];
    uint64_t meta[kJoinItems], h[kJoinItems], sl_[kJoinItems];
    bool ok[kJoinItems];
    // phase 1: the entries
#pragma unroll
    for (int k = 0; k < kJoinItems; ++k) {
      const uint64_t gpos = tile * kJoinTile + (uint64_t)k * kBlock + threadIdx.x;
      ok[k] = gpos < total;
      if (!ok[k]) continue;
      const SparseEnt<W>& e = ents[sub_entry(sh.pre, sl.S, sl.R, gpos)];
#pragma unroll
      for (int w = 0; w < W; ++w) q[k].w[w] = e.w[w];
      c[k] = __longlong_as_double((long long)e.w[W]);
      v[k] = __longlong_as_double((long long)e.w[W + 1]);
      meta[k] = e.w[W + 2];
    }
    // phase 2: the partners' first table slots
#pragma unroll
    for (int k = 0; k < kJoinItems; ++k) {
      if (!ok[k]) continue;
#pragma unroll
      for (int w = 0; w < W; ++w) q[k].w[w] ^= g.gen.w[w];
      h[k] = sparse_hash<W>(q[k]);
      sl_[k] = table[h[k] & mask];
    }
    // phase 3: the candidates those slots name (a fingerprint match), then
    // the rare probes past a first slot of another key
    int64_t j[kJoinItems];
    uint64_t pm[kJoinItems];
    double pv[kJoinItems];
#pragma unroll
    for (int k = 0; k < kJoinItems; ++k) {
      j[k] = -1;
      pm[k] = 0;
      pv[k] = 0.0;
      if (!ok[k] || !slot_fp_match(sl_[k], h[k])) continue;
      const SparseEnt<W>& e = ents[sl_[k] & 0xFFFFFFFFULL];
      bool eq = true;
#pragma unroll
      for (int w = 0; w < W; ++w) eq = eq && e.w[w] == q[k].w[w];
      if (eq) {
        j[k] = (int64_t)(sl_[k] & 0xFFFFFFFFULL);
        pm[k] = e.w[W + 2];
        pv[k] = __longlong_as_double((long long)e.w[W + 1]);
      }
    }
#pragma unroll
    for (int k = 0; k < kJoinItems; ++k) {
      if (!ok[k] || j[k] >= 0 || !(sl_[k] & kSlotValid)) continue;
      for (uint64_t i = ((h[k] & mask) + 1) & mask;; i = (i + 1) & mask) {
        const uint64_t x = table[i];
        if (!(x & kSlotValid)) break;
        if (!slot_fp_match(x, h[k])) continue;
        const SparseEnt<W>& e = ents[x & 0xFFFFFFFFULL];
        bool eq = true;
#pragma unroll
        for (int w = 0; w < W; ++w) eq = eq && e.w[w] == q[k].w[w];
        if (eq) {
          j[k] = (int64_t)(x & 0xFFFFFFFFULL);
          pm[k] = e.w[W + 2];
          pv[k] = __longlong_as_double((long long)e.w[W + 1]);
          break;
        }
      }
    }
    // phase 4: the pair arithmetic, in place
    bool fresh[kJoinItems];
#pragma unroll
    for (int k = 0; k < kJoinItems; ++k) {
      fresh[k] = false;
      if (!ok[k]) continue;
      const int dp = (int)((meta[k] >> 32) & 3);
      double cn = c[k] * row_rot<W>(g, q[k]).c;  // the instance of the row (and of its partner)
      if (j[k] >= 0) {
        const int dq = (int)((pm[k] >> 32) & 3);
        if (dq == kDecEmit) {
          cn = cn + pv[k];
          ++hits;
        } else if (dq == kDecPending) {
          if (dp == kDecEmit) {  // the pair rule: P's own branch went into Q, so Q's is claimed
            cn = cn + pv[k];
            ++hits;
            ++emitted;
          } else {
            disc_l1 += fabs(pv[k]);
            disc_max = fmax(disc_max, fabs(pv[k]));
          }
        }
      }
      s.coeff[meta[k] & 0xFFFFFFFFULL] = cn;
      if (dp == kDecEmit) {
        ++emitted;
        if (j[k] < 0) {
          fresh[k] = true;
          ++created;
        }
      } else if (dp == kDecPending && j[k] < 0) {
        disc_l1 += fabs(v[k]);
        disc_max = fmax(disc_max, fabs(v[k]));
      }
      if (tape) {  // recording: the row, its coefficient and where its branch landed
        SparseRec& tr = tape[tile * kJoinTile + (uint64_t)k * kBlock + threadIdx.x];
        tr.row = (uint32_t)(meta[k] & 0xFFFFFFFFULL);
        tr.c = c[k];
        const int dq = j[k] >= 0 ? (int)((pm[k] >> 32) & 3) : kDecNone;
        const bool lands = j[k] >= 0 && (dp == kDecEmit || (dp == kDecPending && dq == kDecEmit));
        // a branch that made a new row gets that row when the stage is placed
        if (!fresh[k])
          tr.child = lands ? (uint32_t)(pm[k] & 0xFFFFFFFFULL) | ((meta[k] & kEntNeg) ? kRecSign : 0u) : kRecNoChild;
        if (tape_inst)  // batched engines: the row's instance (its angle and gradient slot)
          tape_inst[tile * kJoinTile + (uint64_t)k * kBlock + threadIdx.x] = (uint32_t)(q[k].w[W - 1] >> g.inst_shift);
      }
    }
    uint32_t rank[kJoinItems];
#pragma unroll
    for (int k = 0; k < kJoinItems; ++k) {
      const uint64_t m = __ballot(fresh[k]);
      rank[k] = popc64(m & lt);
      if (lane == 0) wcnt[k][wave] = popc64(m);
    }
    const uint32_t tile_new = slab_prefix<kJoinItems>(wcnt);
    if (sh.len + tile_new > (uint32_t)Stage::kRows)
      stage_flush(sh, false, 0, n_tiles, n, s, join_word, n_out, n_out_host, e_host, pub_tag, total, tape);
    const uint32_t sb = sh.len;
#pragma unroll
    for (int k = 0; k < kJoinItems; ++k) {
      if (!fresh[k]) continue;
      const uint32_t at = sb + wcnt[k][wave] + rank[k];
#pragma unroll
      for (int w = 0; w < W; ++w) sh.key[w][at] = q[k].w[w];
      sh.c[at] = v[k];
      sh.rec[at] = (uint32_t)(tile * kJoinTile + (uint64_t)k * kBlock + threadIdx.x) |
                   ((meta[k] & kEntNeg) ? kRecSign : 0u);
    }
    __syncthreads();
    if (threadIdx.x == 0) sh.len = sb + tile_new;
    ++tiles_done;
  }
  stage_flush(sh, true, tiles_done, n_tiles, n, s, join_word, n_out, n_out_host, e_host, pub_tag, total, tape);
  clear_prev(sh, prev_slots, prev_table, table_cap, ent_cap, kPrefetch, first_slot);
  if (ctr) {
    hits = wave_sum_u64(hits);
    emitted = wave_sum_u64(emitted);
    created = wave_sum_u64(created);
    double l1 = disc_l1, mx = disc_max;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      l1 += __shfl_xor(l1, o, 64);
      mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
      Counters* cs = ctr_slot(ctr);
      if (hits) atomicAdd(&cs->hits, hits);
      if (emitted) atomicAdd(&cs->emitted, emitted);
      if (created) atomicAdd(&cs->created, created);
      if (l1 != 0.0) atomicAdd(&cs->disc_l1, l1);
      if (mx > 0.0) atomicMax(&cs->disc_max_bits, (unsigned long long)__double_as_longlong(mx));
    }
  }
}

// ---------------------------------------------------- sparse gate pairs
//
// Two consecutive gates A then B whose generators commute (and differ) run
// as one gather and one join.  For commuting A and B the four strings P, AP,
// BP, ABP (keys p, p^a, p^b, p^a^b) anticommute with A alike and with B
// alike, and the two rotations map their span onto itself.  So a row that
// anticommutes with A only changes with its partner under A exactly as in a
// single gate (B leaves both alone), one that anticommutes with B only with
// its partner under B (A left both alone), and one that anticommutes with
// both with the three others: gate A on the pairs (P, AP) and (BP, ABP), then
// gate B on (P, BP) and (AP, ABP).  The join replays the single gate's pair
// rule (k_sparse_join's arithmetic, decide() for every present row of the
// pair at that gate's start) on each such group, so every coefficient is the
// one the two gates give one after the other, bit for bit.
//
// The gather lists every row anticommuting with A or B (key, c, row, the two
// flags) and enters it in the table; the join probes the group's other keys.
// Each present row writes its own coefficient; new rows (absent before the
// pass, present after it) are appended by the row that made them in a pair
// of one gate, and by the group's least present key in a group of four.
// Rows made by gate A are counted into mid (the host reads the counts of
// the pairs when it next waits for the device), the others only through the
// join word, so the last tile publishes the count after B.

struct PairTally {  // one thread's counts (a few per tile: 32 bits hold them)
  uint32_t visited = 0, emitted = 0, declined = 0, hits = 0, created = 0, dropped = 0;
  double l1 = 0.0, mx = 0.0;
};

template <int W>
struct PairMember {
  Key<W> k;
  double c;
  bool live;
};

// One present row of a pair under one gate: its coefficient afterwards
// (c cos plus what its partner sends it, as k_sparse_join), with the counters
// k_sparse_gather and k_sparse_join keep for it added to t when count is set.
__device__ __forceinline__ double pair_side(double c, double cos_t, int d_own, double v_own, bool partner_live,
                                            int d_par, double v_par, PairTally& t, bool count) {
  double cn = c * cos_t;
  if (count) {
    t.visited++;
    t.declined += d_own == kDecPending;
    t.dropped += d_own == kDecDrop;
  }
  if (partner_live) {
    if (d_par == kDecEmit) {
      cn = cn + v_par;
      if (count) t.hits++;
    } else if (d_par == kDecPending) {
      if (d_own == kDecEmit) {  // the pair rule: the own branch went in, the partner's is claimed
        cn = cn + v_par;
        if (count) {
          t.hits++;
          t.emitted++;
        }
      } else if (count) {
        t.l1 += fabs(v_par);
        t.mx = fmax(t.mx, fabs(v_par));
      }
    }
  }
  if (d_own == kDecEmit) {
    if (count) {
      t.emitted++;
      if (!partner_live) t.created++;
    }
  } else if (d_own == kDecPending && !partner_live && count) {
    t.l1 += fabs(v_own);
    t.mx = fmax(t.mx, fabs(v_own));
  }
  return cn;
}

// Gate g on the pair (x, y = g x): both sides' decisions from their
// coefficients at the gate's start, then both coefficients; an absent side
// becomes present with the other's emitted branch.  The sides flagged by
// count_x / count_y add their counters to t.
template <int KIND, int W>
__device__ __forceinline__ void pair_rule(const GateArgs<W>& g, PairMember<W>& x, PairMember<W>& y, PairTally& t,
                                          bool count_x, bool count_y) {
  Key<W> q;
  double vx = 0.0, vy = 0.0;
  const int dx = x.live ? decide<KIND, W>(g, x.k, x.c, q, vx) : kDecNone;
  const int dy = y.live ? decide<KIND, W>(g, y.k, y.c, q, vy) : kDecNone;
  const double cos_t = row_rot<W>(g, x.k).c;  // batched engines: the pair's instance
  const double cx = x.live ? pair_side(x.c, cos_t, dx, vx, y.live, dy, vy, t, count_x) : 0.0;
  const double cy = y.live ? pair_side(y.c, cos_t, dy, vy, x.live, dx, vx, t, count_y) : 0.0;
  const bool lx = x.live, ly = y.live;
  if (lx) {
    x.c = cx;
  } else if (dy == kDecEmit) {
    x.c = vy;
    x.live = true;
  }
  if (ly) {
    y.c = cy;
  } else if (dx == kDecEmit) {
    y.c = vx;
    y.live = true;
  }
}

// The list entry of key k, or -1: the table's first slot for k's hash has
// been read (first); linear probing from there (the table has empty slots).
template <int W>
__device__ __forceinline__ int64_t pair_probe(const uint64_t* table, uint64_t mask, const SparseEnt<W>* ents,
                                              const Key<W>& k, uint64_t h, uint64_t first) {
  uint64_t i = h & mask, x = first;
  while (x & kSlotValid) {
    if (slot_fp_match(x, h)) {
      const SparseEnt<W>& e = ents[x & 0xFFFFFFFFULL];
      bool eq = true;
#pragma unroll
      for (int w = 0; w < W; ++w) eq = eq && e.w[w] == k.w[w];
      if (eq) return (int64_t)(x & 0xFFFFFFFFULL);
    }
    i = (i + 1) & mask;
    x = table[i];
  }
  return -1;
}

// Block b takes rows [b T, (b + 1) T) and lists those anticommuting with A or
// B, as k_sparse_gather does for one gate (entry meta: row | A flag << 32 |
// B flag << 33); block 0 also clears the join word and this pass's mid
// counts and writes the record's header.
template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_pair_gather(GateArgs<W> ga, GateArgs<W> gb, StoreDev s,
                                                        const uint64_t* n_ptr, SparseEnt<W>* ents, uint32_t* slots,
                                                        uint64_t ent_cap, uint64_t* rec, uint64_t* table,
                                                        uint64_t table_cap, unsigned long long* join_word,
                                                        unsigned long long* mid) {
  __shared__ uint32_t wcnt[kSparseItems][kSparseWaves];
  __shared__ uint16_t rows[kSparseTile];
  __shared__ uint64_t s_base;
  const uint64_t n = *n_ptr;
  const SubLists sl = sparse_sublists(n);
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      *join_word = 0;
      rec_hdr(rec)[0] = sl.S;
      rec_hdr(rec)[1] = sl.R;
    }
    if (threadIdx.x < kMidSlots) mid[threadIdx.x * kMidStride] = 0;
  }
  const uint64_t T = sparse_table_slots(n);
  if (T > table_cap) return;  // the join refuses the pass
  const uint64_t base = (uint64_t)blockIdx.x * kSparseTile;
  if (base >= n) return;  // the grid is sized by the host's bound of n, which a pass in flight may overstate
  const uint64_t mask = T - 1;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt = lane == 0 ? 0ULL : (~0ULL >> (64 - lane));
  uint64_t x[kSparseItems][W];
#pragma unroll
  for (int it = 0; it < kSparseItems; ++it) {
    const uint64_t r = base + (uint64_t)it * kBlock + threadIdx.x;
#pragma unroll
    for (int w = 0; w < W; ++w)
      x[it][w] = (r < n && (word_needed<KIND, W>(ga, w) || word_needed<KIND, W>(gb, w) || (ga.sc && w == W - 1)))
                     ? s.key[(uint64_t)w * s.cap + r] : 0;
  }
  uint32_t rank[kSparseItems];
  uint8_t flags[kSparseItems];
#pragma unroll
  for (int it = 0; it < kSparseItems; ++it) {
    const uint64_t r = base + (uint64_t)it * kBlock + threadIdx.x;
    Key<W> p;
#pragma unroll
    for (int w = 0; w < W; ++w) p.w[w] = x[it][w];
    // batched engines: the rows of a frozen instance count as commuting
    flags[it] = r < n && !row_frozen<W>(ga, p)
                    ? (uint8_t)((row_anti<KIND, W>(ga, p) ? 1 : 0) | (row_anti<KIND, W>(gb, p) ? 2 : 0)) : 0;
    const uint64_t m = __ballot(flags[it] != 0);
    rank[it] = popc64(m & lt);
    if (lane == 0) wcnt[it][wave] = popc64(m);
  }
  const uint32_t cnt = slab_prefix<kSparseItems>(wcnt);
  __shared__ uint8_t rflags[kSparseTile];
#pragma unroll
  for (int it = 0; it < kSparseItems; ++it)
    if (flags[it]) {
      const uint32_t at = wcnt[it][wave] + rank[it];
      rows[at] = (uint16_t)(it * kBlock + threadIdx.x);
      rflags[at] = flags[it];
    }
  __syncthreads();
  const uint64_t sq = blockIdx.x % sl.S;
  unsigned long long reserved = 0;
  if (threadIdx.x == 0 && cnt) reserved = atomicAdd(rec_len(rec, sq), (unsigned long long)cnt);
  // the first round of listed rows loads while the reservation is in flight
  auto entry = [&](uint32_t li, uint64_t& h) {
    const uint64_t r = base + rows[li];
    const Key<W> p = load_key<W>(s.key, s.cap, r);
    SparseEnt<W> e;
#pragma unroll
    for (int w = 0; w < SparseEnt<W>::kWords; ++w) e.w[w] = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) e.w[w] = p.w[w];
    e.w[W] = (uint64_t)__double_as_longlong(s.coeff[r]);
    e.w[W + 2] = r | ((uint64_t)rflags[li] << 32);
    h = sparse_hash<W>(p);
    return e;
  };
  const bool has0 = threadIdx.x < cnt;
  uint64_t h0 = 0;
  SparseEnt<W> e0;
  if (has0) e0 = entry(threadIdx.x, h0);
  if (threadIdx.x == 0) s_base = reserved;
  __syncthreads();
  // the region's room holds every row of its blocks; else the join refuses the pass
  if (!(s_base + cnt <= sl.R && sl.S * sl.R <= ent_cap)) return;
  const uint64_t b0 = sq * sl.R + s_base;
  if (has0) {
    const uint64_t a = b0 + threadIdx.x;
    ents[a] = e0;
    slots[a] = (uint32_t)sparse_insert(table, mask, h0, a);
  }
  for (uint32_t li = threadIdx.x + kBlock; li < cnt; li += kBlock) {  // more than one row per thread: rare
    uint64_t h;
    const SparseEnt<W> e = entry(li, h);
    const uint64_t a = b0 + li;
    ents[a] = e;
    slots[a] = (uint32_t)sparse_insert(table, mask, h, a);
  }
}

// Exclusive prefix of each thread's count over the block (thread order) and
// the block total.  Every thread calls it.
__device__ __forceinline__ uint32_t block_prefix(uint32_t v, uint32_t& total) {
  __shared__ uint32_t wsum[kSparseWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)inc, o, 64);
    if (lane >= o) inc += t;
  }
  __syncthreads();
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < kSparseWaves; ++w) {
    before += w < wave ? wsum[w] : 0;
    all += wsum[w];
  }
  total = all;
  return before + inc - v;
}

// Every list entry: its group's other rows by table lookup, the two gates on
// the group as described above, its own coefficient in place, and the rows
// it is to append (after the n_in old rows, through the block's LDS stage, as
// k_sparse_join does).  Also clears the previous sparse pass's slots and the
// next one's counters, as k_sparse_join does.
constexpr int kPairTile = kBlock;  // list entries per pair-join tile, one per thread

template <int KIND, int W>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_pair_join(GateArgs<W> ga, GateArgs<W> gb, const SparseEnt<W>* ents,
                                                      uint64_t ent_cap, const uint64_t* rec, const uint64_t* table,
                                                      uint64_t table_cap, const uint32_t* prev_slots,
                                                      const uint64_t* prev_rec, uint64_t* prev_table,
                                                      uint64_t* next_rec, StoreDev s, const uint64_t* n_ptr,
                                                      uint64_t* n_out, uint64_t* n_out_host, uint64_t* e_host,
                                                      uint32_t pub_tag, unsigned long long* join_word,
                                                      unsigned long long* mid, Counters* ctr) {
  using Stage = JoinStage<W, 3 * kPairTile, false>;  // a tile appends at most three rows per entry
  __shared__ Stage sh;
  __shared__ unsigned long long s_made[kSparseWaves];
  const uint64_t n = *n_ptr, T = sparse_table_slots(n);
  const SubLists sl = sparse_sublists(n);
  const uint64_t total =
      join_prologue(rec, sl, prev_rec, next_rec, sh);
  if (sl.S * sl.R > ent_cap || total > n || n + 3 * total > s.cap || T > table_cap) {  // a capacity bound was violated
    if (blockIdx.x == 0 && threadIdx.x == 0) publish_counts(n_out_host, e_host, pub_tag, kPubRefused, total);
    clear_prev(sh, prev_slots, prev_table, table_cap, ent_cap);
    return;
  }
  const uint64_t n_tiles = (total + kPairTile - 1) / kPairTile;
  if (n_tiles == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      *n_out = n;
      publish_counts(n_out_host, e_host, pub_tag, n, 0);
    }
    clear_prev(sh, prev_slots, prev_table, table_cap, ent_cap);
    return;
  }
  const uint64_t active = join_blocks(n_tiles);
  if (blockIdx.x >= active) {
    clear_prev(sh, prev_slots, prev_table, table_cap, ent_cap);
    return;
  }
  // narrow keys leave the registers for it; wide ones would spill
  constexpr bool kPrefetch = W <= 2;
  const uint32_t first_slot = kPrefetch ? prev_slot_first(sh, prev_slots, ent_cap) : 0u;
  if (threadIdx.x == 0) sh.len = 0;
  const uint64_t mask = T - 1;
  PairTally ta;  // both gates' counts: the pass adds them to one counter set
  unsigned long long made_a_sum = 0;  // rows gate A made, summed per block into a mid slot at the end
  uint64_t tiles_done = 0;
  for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += active) {
    PairMember<W> m[4];
    int fl = 0;
    uint64_t row = 0;
    const uint64_t gpos = tile * kPairTile + threadIdx.x;
    if (gpos < total) {
      const SparseEnt<W>& e = ents[sub_entry(sh.pre, sl.S, sl.R, gpos)];
#pragma unroll
      for (int w = 0; w < W; ++w) m[0].k.w[w] = e.w[w];
      m[0].c = __longlong_as_double((long long)e.w[W]);
      m[0].live = true;
      row = e.w[W + 2] & 0xFFFFFFFFULL;
      fl = (int)((e.w[W + 2] >> 32) & 3);
    }
    // the group's other keys: member 1 is the partner under the one gate of
    // a one-gate row, or AP; members 2 and 3 are BP and ABP
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const uint64_t p = m[0].k.w[w];
      m[1].k.w[w] = p ^ (fl == 2 ? gb.gen.w[w] : ga.gen.w[w]);
      m[2].k.w[w] = p ^ gb.gen.w[w];
      m[3].k.w[w] = p ^ ga.gen.w[w] ^ gb.gen.w[w];
    }
    const int others = fl == 3 ? 3 : (fl ? 1 : 0);
    uint64_t h[3], first[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      h[j] = sparse_hash<W>(m[j + 1].k);
      first[j] = j < others ? table[h[j] & mask] : 0;
    }
    // narrow keys: the entry each first slot names, when its fingerprint
    // matches (key words and coefficient), is loaded for all three members
    // before any is compared, so the members' round trips overlap; wide keys
    // would spill these registers and probe one member after another
    constexpr bool kCand = W <= 2;
    uint64_t cand[3][W + 1];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
#pragma unroll
      for (int w = 0; w <= W; ++w) cand[j][w] = 0;
      if (kCand && j < others && slot_fp_match(first[j], h[j])) {
        const SparseEnt<W>& e = ents[first[j] & 0xFFFFFFFFULL];
#pragma unroll
        for (int w = 0; w <= W; ++w) cand[j][w] = e.w[w];
      }
    }
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      m[j + 1].live = false;
      m[j + 1].c = 0.0;
      if (j >= others) continue;
      bool eq = kCand && slot_fp_match(first[j], h[j]);
#pragma unroll
      for (int w = 0; w < W; ++w) eq = eq && cand[j][w] == m[j + 1].k.w[w];
      if (eq) {
        m[j + 1].live = true;
        m[j + 1].c = __longlong_as_double((long long)cand[j][W]);
        continue;
      }
      // wide keys, or the first slot empty or holding another key
      const int64_t at = pair_probe<W>(table, mask, ents, m[j + 1].k, h[j], first[j]);
      if (at >= 0) {
        m[j + 1].live = true;
        m[j + 1].c = __longlong_as_double((long long)ents[at].w[W]);
      }
    }
    // the gates on the group; the rows this thread appends
    uint32_t made = 0, made_a = 0;
    bool app[3] = {false, false, false};
    if (fl == 1 || fl == 2) {  // one gate: this row and its partner, counted for this row
      const bool was = m[1].live;
      if (fl == 1)
        pair_rule<KIND, W>(ga, m[0], m[1], ta, true, false);
      else
        pair_rule<KIND, W>(gb, m[0], m[1], ta, true, false);
      if (!was && m[1].live) {
        app[0] = true;
        made = 1;
        made_a = fl == 1;
      }
    } else if (fl == 3) {  // A, then B, on the group of four; the least present key counts and appends
      bool lead = true;
#pragma unroll
      for (int j = 1; j < 4; ++j) lead = lead && !(m[j].live && key_less<W>(m[j].k, m[0].k));
      const bool was[4] = {true, m[1].live, m[2].live, m[3].live};
      pair_rule<KIND, W>(ga, m[0], m[1], ta, lead, lead);
      pair_rule<KIND, W>(ga, m[2], m[3], ta, lead, lead);
      const bool after_a[4] = {true, m[1].live, m[2].live, m[3].live};
      pair_rule<KIND, W>(gb, m[0], m[2], ta, lead, lead);
      pair_rule<KIND, W>(gb, m[1], m[3], ta, lead, lead);
      if (lead) {
#pragma unroll
        for (int j = 1; j < 4; ++j) {
          if (!was[j] && m[j].live) {
            app[j - 1] = true;
            ++made;
            made_a += after_a[j];
          }
        }
      }
    }
    if (fl) s.coeff[row] = m[0].c;
    uint32_t tile_new = 0;
    const uint32_t off = block_prefix(made, tile_new);
    made_a_sum += made_a;
    if (sh.len + tile_new > (uint32_t)Stage::kRows)
      stage_flush(sh, false, 0, n_tiles, n, s, join_word, n_out, n_out_host, e_host, pub_tag, total);
    uint32_t at = sh.len + off;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (!app[j]) continue;
      // values, not a reference into m chosen at run time (that would put m in scratch)
      const bool grp = fl == 3;
#pragma unroll
      for (int w = 0; w < W; ++w) sh.key[w][at] = grp ? m[j + 1].k.w[w] : m[1].k.w[w];
      sh.c[at] = grp ? m[j + 1].c : m[1].c;
      ++at;
    }
    __syncthreads();
    if (threadIdx.x == 0) sh.len += tile_new;
    ++tiles_done;
  }
  stage_flush(sh, true, tiles_done, n_tiles, n, s, join_word, n_out, n_out_host, e_host, pub_tag, total);
  clear_prev(sh, prev_slots, prev_table, table_cap, ent_cap, kPrefetch, first_slot);
  made_a_sum = wave_sum_u64(made_a_sum);
  if ((threadIdx.x & 63) == 0) s_made[threadIdx.x >> 6] = made_a_sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kSparseWaves; ++w) t += s_made[w];
    if (t) atomicAdd(mid + (blockIdx.x % kMidSlots) * kMidStride, t);
  }
  if (ctr) {
    unsigned long long v[6] = {ta.visited, ta.emitted, ta.declined, ta.hits, ta.created, ta.dropped};
#pragma unroll
    for (int i = 0; i < 6; ++i) v[i] = wave_sum_u64(v[i]);
    double l1 = ta.l1, mx = ta.mx;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      l1 += __shfl_xor(l1, o, 64);
      mx = fmax(mx, __shfl_xor(mx, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
      Counters* cs = ctr_slot(ctr);
      if (v[0]) atomicAdd(&cs->visited, v[0]);
      if (v[1]) atomicAdd(&cs->emitted, v[1]);
      if (v[2]) atomicAdd(&cs->declined, v[2]);
      if (v[3]) atomicAdd(&cs->hits, v[3]);
      if (v[4]) atomicAdd(&cs->created, v[4]);
      if (v[5]) atomicAdd(&cs->dropped_weight, v[5]);
      if (l1 != 0.0) atomicAdd(&cs->disc_l1, l1);
      if (mx > 0.0) atomicMax(&cs->disc_max_bits, (unsigned long long)__double_as_longlong(mx));
    }
  }
}

// ------------------------------------------------------ readout / expect

// Readout conjugation through a deferred Clifford frame (tableau.rs:109-143):
// multiply the rows of the key's set bits in ascending order, tracking the
// phase of the source product and of the image product.
template <int KIND, int W>
__device__ __forceinline__ double frame_apply(const uint64_t* img, const double* sgn, int n_rows, const Key<W>& k,
                                              Key<W>& out) {
  Key<W> src, im;
#pragma unroll
  for (int i = 0; i < W; ++i) src.w[i] = im.w[i] = 0;
  int se = 0, ie = 0;
#pragma unroll
  for (int wd = 0; wd < W; ++wd) {
    uint64_t bits = k.w[wd];
    while (bits) {
      const int b = __ffsll((unsigned long long)bits) - 1;
      bits &= bits - 1;
      const int p = wd * 64 + b;
      if (p >= n_rows) continue;
      Key<W> e1;
#pragma unroll
      for (int i = 0; i < W; ++i) e1.w[i] = 0;
      e1.w[wd] = 1ULL << b;
      se += product_exp<KIND, W>(e1, src);
      src.w[wd] ^= e1.w[wd];
      Key<W> row;
#pragma unroll
      for (int i = 0; i < W; ++i) row.w[i] = img[(uint64_t)p * W + i];
      ie += product_exp<KIND, W>(row, im) + (sgn[p] < 0.0 ? 2 : 0);
#pragma unroll
      for (int i = 0; i < W; ++i) im.w[i] ^= row.w[i];
    }
  }
  out = im;
  return (((ie - se) & 3) == 2) ? -1.0 : 1.0;
}

template <int KIND, int W>
__device__ __forceinline__ double trace_key(const Key<W>& k, const Key<W>& dm, int n_units) {
  return basis_trace<KIND, W>(k, dm, n_units);
}

// One block's share of sum c * sign * trace over rows [lo, lo + n): block
// blk of nblk, grid-stride, then a fixed tree in LDS.  The unbatched and the
// per-instance expectation use the same partition, so a batched instance sums
// in exactly the order of a run of its own.
template <int KIND, int W>
__device__ double expectation_block(StoreDev s, uint64_t lo, uint64_t n, uint64_t blk, uint64_t nblk,
                                    const Key<W>& dm, int n_units, const uint64_t* img, const double* sgn,
                                    int n_rows, uint64_t unit_mask) {
  __shared__ double red[kBlock];
  double acc = 0.0;
  for (uint64_t r = blk * kBlock + threadIdx.x; r < n; r += nblk * kBlock) {
    Key<W> k = unit_bits<W>(load_key<W>(s.key, s.cap, lo + r), unit_mask);
    double sg = 1.0;
    if (img) {
      Key<W> o;
      sg = frame_apply<KIND, W>(img, sgn, n_rows, k, o);
      k = o;
    }
    const double t = trace_key<KIND, W>(k, dm, n_units);
    if (t != 0.0) acc += s.coeff[lo + r] * sg * t;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  return red[0];
}

template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_expectation(StoreDev s, uint64_t n, Key<W> dm, int n_units,
                                                        const uint64_t* img, const double* sgn, int n_rows,
                                                        double* partials, uint64_t unit_mask) {
  const double v = expectation_block<KIND, W>(s, 0, n, blockIdx.x, gridDim.x, dm, n_units, img, sgn, n_rows,
                                              unit_mask);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// work[3 i .. 3 i + 2] = (instance, block, blocks of that instance).
template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_expectation_segments(StoreDev s, const uint32_t* work,
                                                                 const uint64_t* bounds, const uint64_t* dms,
                                                                 int n_units, const uint64_t* img, const double* sgn,
                                                                 int n_rows, double* partials, uint64_t unit_mask) {
  const uint32_t b = work[3 * blockIdx.x], blk = work[3 * blockIdx.x + 1], nblk = work[3 * blockIdx.x + 2];
  Key<W> dm;
#pragma unroll
  for (int i = 0; i < W; ++i) dm.w[i] = dms[(uint64_t)b * W + i];
  const uint64_t lo = bounds[b], hi = bounds[b + 1];
  const double v = expectation_block<KIND, W>(s, lo, hi - lo, blk, nblk, dm, n_units, img, sgn, n_rows, unit_mask);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// Binary search of each instance's first row on the most significant plane.
__global__ void k_instance_bounds(const uint64_t* top, const uint64_t* n_ptr, int shift, int n_inst,
                                  uint64_t* bounds) {
  const uint64_t n = *n_ptr;
  for (int b = blockIdx.x * blockDim.x + threadIdx.x; b <= n_inst; b += gridDim.x * blockDim.x) {
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if ((top[mid] >> shift) < (uint64_t)b)
        lo = mid + 1;
      else
        hi = mid;
    }
    bounds[b] = lo;
  }
}

template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_conjugate(StoreDev s, uint64_t n, const uint64_t* img,
                                                      const double* sgn, int n_rows, uint64_t* out_keys,
                                                      double* out_coeff) {
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += (uint64_t)gridDim.x * kBlock) {
    Key<W> k = load_key<W>(s.key, s.cap, r), o;
    double sg = 1.0;
    if (img) {
      sg = frame_apply<KIND, W>(img, sgn, n_rows, k, o);
      // bits above the frame's rows (a batched engine's instance id) pass through
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const int lo = 64 * i;
        const uint64_t above = n_rows <= lo ? ~0ULL : (n_rows - lo >= 64 ? 0ULL : ~((1ULL << (n_rows - lo)) - 1));
        o.w[i] |= k.w[i] & above;
      }
    } else {
      o = k;
    }
#pragma unroll
    for (int i = 0; i < W; ++i) out_keys[r * W + i] = o.w[i];
    out_coeff[r] = s.coeff[r] * sg;
  }
}

template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_weights(StoreDev s, uint64_t n, int n_units, uint32_t* out,
                                                    uint64_t unit_mask) {
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += (uint64_t)gridDim.x * kBlock)
    out[r] = (uint32_t)weight<KIND, W>(unit_bits<W>(load_key<W>(s.key, s.cap, r), unit_mask), n_units);
}

__global__ void k_gather_bounds(const uint64_t* offs, uint64_t n_blocks, int m, uint64_t* out) {
  for (int c = threadIdx.x; c <= m; c += blockDim.x) out[c] = offs[(uint64_t)c * n_blocks];
}

__global__ __launch_bounds__(kBlock) void k_count_below(const double* coeff, uint64_t n, double thr,
                                                        unsigned long long* out) {
  unsigned long long cnt = 0;
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += (uint64_t)gridDim.x * kBlock)
    cnt += fabs(coeff[r]) < thr;
  cnt = wave_sum_u64(cnt);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(out, cnt);
}

// ---------------------------------------------------------------- adjoint

// d<O>/dc of every row at the end of the forward run: the readout weight of
// the expectation (k_expectation without the coefficient).
template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_adjoint_seed(StoreDev s, uint64_t n, Key<W> dm, int n_units,
                                                         const uint64_t* img, const double* sgn, int n_rows,
                                                         double* lam) {
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += (uint64_t)gridDim.x * kBlock) {
    Key<W> k = load_key<W>(s.key, s.cap, r);
    double sg = 1.0;
    if (img) {
      Key<W> o;
      sg = frame_apply<KIND, W>(img, sgn, n_rows, k, o);
      k = o;
    }
    lam[r] = sg * trace_key<KIND, W>(k, dm, n_units);
  }
}

// Batched engines: row r of instance b = (key word W-1) >> shift reads with
// b's own basis-state mask dms[b W .. b W + W).
template <int KIND, int W>
__global__ __launch_bounds__(kBlock) void k_adjoint_seed_batched(StoreDev s, uint64_t n, const uint64_t* dms,
                                                                 int shift, uint64_t unit_mask, int n_units,
                                                                 const uint64_t* img, const double* sgn,
                                                                 int n_rows, double* lam) {
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += (uint64_t)gridDim.x * kBlock) {
    const Key<W> raw = load_key<W>(s.key, s.cap, r);
    const uint64_t b = raw.w[W - 1] >> shift;
    Key<W> k = unit_bits<W>(raw, unit_mask), dm;
#pragma unroll
    for (int i = 0; i < W; ++i) dm.w[i] = dms[b * W + i];
    double sg = 1.0;
    if (img) {
      Key<W> o;
      sg = frame_apply<KIND, W>(img, sgn, n_rows, k, o);
      k = o;
    }
    lam[r] = sg * trace_key<KIND, W>(k, dm, n_units);
  }
}

// Reverse of one recorded rotation.  Forward, per input row r with
// coefficient c_r: out[o_r] += cos c_r (anticommuting) or c_r (commuting),
// and out[q_r] += sin (-Im phase) c_r when its branch reached the output.
// With the truncation decisions held fixed this is linear in c, so
//   lam_in[r]   = cos lam[o_r] + sin s_r lam[q_r]   (lam[o_r] if commuting)
//   d<O>/dtheta = sum_r  -sin c_r lam[o_r] + cos s_r c_r lam[q_r].
// The merge becomes two gathers; the partial sums of each block are reduced
// in a fixed order, so the result does not depend on scheduling.
// One block's share of the reverse step of a recorded gate over rows
// [lo, lo + n): block blk of nblk, grid-stride, so a batched instance's work
// items reproduce the blocks of a run of that instance alone.
__device__ double adjoint_gate_block(uint64_t lo, uint64_t n, uint64_t blk, uint64_t nblk, TapeDev tape,
                                     const double* c_in, const double* lam_out, double* lam_in, double sin_t,
                                     double cos_t) {
  __shared__ double red[kBlock / 64];
  double acc = 0.0;
  for (uint64_t rr = blk * kBlock + threadIdx.x; rr < n; rr += nblk * kBlock) {
    const uint64_t r = lo + rr;
    const double l = lam_out[tape.out[r]];
    const uint32_t ch = tape.child[r];
    double li = l;
    if (ch != kRecCommute) {
      const double cr = c_in[r];
      li = cos_t * l;
      acc += -sin_t * cr * l;
      const uint32_t q = ch & ~kRecSign;
      if (q != kRecNoChild) {
        const double sg = (ch & kRecSign) ? -1.0 : 1.0;
        const double lq = lam_out[q];
        li += sin_t * sg * lq;
        acc += cos_t * sg * cr * lq;
      }
    }
    lam_in[r] = li;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kBlock / 64; ++i) t += red[i];
  return t;
}

__global__ __launch_bounds__(kBlock) void k_adjoint_gate(uint64_t n, TapeDev tape, const double* c_in,
                                                         const double* lam_out, double* lam_in, double sin_t,
                                                         double cos_t, double* partials) {
  const double t = adjoint_gate_block(0, n, blockIdx.x, gridDim.x, tape, c_in, lam_out, lam_in, sin_t, cos_t);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Batched engines: work item i = (instance b, block, blocks of b) over b's
// rows [bounds[b], bounds[b + 1]) with b's own sin / cos (sc[3 b], sc[3 b + 1]).
__global__ __launch_bounds__(kBlock) void k_adjoint_gate_segments(const uint32_t* work, const uint64_t* bounds,
                                                                  const double* sc, TapeDev tape, const double* c_in,
                                                                  const double* lam_out, double* lam_in,
                                                                  double* partials) {
  const uint32_t b = work[3 * blockIdx.x], blk = work[3 * blockIdx.x + 1], nblk = work[3 * blockIdx.x + 2];
  const uint64_t lo = bounds[b], hi = bounds[b + 1];
  const double t =
      adjoint_gate_block(lo, hi - lo, blk, nblk, tape, c_in, lam_out, lam_in, sc[3 * b], sc[3 * b + 1]);
  if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// Adjoint work list of a recorded batched gate, built on the device from the
// instances' row ranges: instance b gets adjoint_blocks(rows) items (b, j,
// blocks), in instance order, from offset off[b]; off[n_inst] = the total, also
// written to *total.  One workgroup; instances are taken 1024 at a time.
constexpr int kWorkThreads = 1024;
__global__ __launch_bounds__(kWorkThreads) void k_instance_work(const uint64_t* bounds, int n_inst, uint32_t* work,
                                                                uint32_t* off, uint32_t* total) {
  __shared__ uint32_t red[kWorkThreads / 64];
  uint32_t base = 0;
  for (int b0 = 0; b0 < n_inst; b0 += kWorkThreads) {
    const int b = b0 + (int)threadIdx.x;
    uint32_t nblk = 0;
    if (b < n_inst) {
      const uint64_t rows = bounds[b + 1] - bounds[b];
      if (rows) {
        const uint64_t nb = (rows + kBlock - 1) / kBlock;
        nblk = (uint32_t)(nb > 2048 ? 2048 : nb);
      }
    }
    // block-wide exclusive scan of nblk
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t x = nblk;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) red[wave] = x;
    __syncthreads();
    uint32_t ex = x - nblk, tot = 0;
#pragma unroll
    for (int i = 0; i < kWorkThreads / 64; ++i) {
      if (i < wave) ex += red[i];
      tot += red[i];
    }
    __syncthreads();
    if (b < n_inst) {
      const uint32_t o = base + ex;
      off[b] = o;
      for (uint32_t j = 0; j < nblk; ++j) {
        work[3 * (uint64_t)(o + j)] = (uint32_t)b;
        work[3 * (uint64_t)(o + j) + 1] = j;
        work[3 * (uint64_t)(o + j) + 2] = nblk;
      }
    }
    base += tot;
  }
  if (threadIdx.x == 0) {
    off[n_inst] = base;
    __hip_atomic_store(total, base, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// Per recorded batched gate g and instance b: the sum of b's work items'
// partials in block order (the order the unbatched sweep sums its blocks).
__global__ __launch_bounds__(kBlock) void k_instance_grads(const uint64_t* part_off, const uint64_t* rows,
                                                           const uint32_t* const* offs, const double* partials,
                                                           int n_inst, double* grads) {
  const uint64_t g = blockIdx.y;
  const int b = (int)(blockIdx.x * kBlock + threadIdx.x);
  if (b >= n_inst) return;
  const uint32_t* off = offs[g];
  const double* p = partials + part_off[g];
  double acc = 0.0;
  for (uint32_t k = off[b]; k < off[b + 1]; ++k) acc += p[k];
  grads[rows[g] * (uint64_t)n_inst + b] = acc;
}

// Reverse of a recorded noise / reclaim pass: reclaimed rows get zero.
__global__ __launch_bounds__(kBlock) void k_adjoint_noise(uint64_t n, TapeDev tape, const double* lam_out,
                                                          double* lam_in) {
  for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += (uint64_t)gridDim.x * kBlock) {
    const uint32_t o = tape.out[r];
    lam_in[r] = o == kRecNoChild ? 0.0 : (tape.factor ? tape.factor[r] * lam_out[o] : lam_out[o]);
  }
}

// ------------------------------------------------------------------- scan

constexpr int kScanTile = 1024;

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const uint32_t* in, uint64_t n, uint64_t* part) {
  const uint64_t b = (uint64_t)blockIdx.x * kScanTile;
  uint32_t acc = 0;
  for (int i = threadIdx.x; i < kScanTile; i += kBlock)
    if (b + i < n) acc += in[b + i];
  uint32_t total;
  (void)block_excl_scan(acc, &total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// Exclusive scan of one 1024-element chunk: the chunk's base is the sum of
// the partials before it (summed by the whole block), then a block scan of
// four contiguous elements per thread.  out[n] receives the grand total.
__global__ __launch_bounds__(kBlock) void k_scan_apply(const uint32_t* in, uint64_t n, const uint64_t* part,
                                                       uint64_t* out) {
  __shared__ uint64_t red[kBlock / 64];
  const uint64_t b = (uint64_t)blockIdx.x * kScanTile;
  uint64_t acc = 0;
  for (uint64_t j = threadIdx.x; j < blockIdx.x; j += kBlock) acc += part[j];
  acc = wave_sum_u64(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  uint64_t base = 0;
#pragma unroll
  for (int i = 0; i < kBlock / 64; ++i) base += red[i];
  uint32_t v[4], sum = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint64_t i = b + threadIdx.x * 4 + k;
    v[k] = i < n ? in[i] : 0;
    sum += v[k];
  }
  uint32_t total;
  uint64_t run = base + block_excl_scan(sum, &total);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint64_t i = b + threadIdx.x * 4 + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = base + total;
}

// Exclusive scan of up to kScanOneMax counts in one workgroup (one launch
// instead of reduce + apply): the counts are staged in LDS with coalesced
// loads, each thread scans a contiguous chunk of at most kScanOnePer, and the
// block scans the chunk sums.
constexpr int kScanOneThreads = 1024;
constexpr int kScanOnePer = 16;
constexpr uint64_t kScanOneMax = (uint64_t)kScanOneThreads * kScanOnePer;
__global__ __launch_bounds__(kScanOneThreads) void k_scan_one(const uint32_t* in, uint64_t n, uint64_t* out) {
  __shared__ uint32_t vals[kScanOneMax];
  __shared__ uint64_t red[kScanOneThreads / 64];
  const int tid = threadIdx.x;
  uint32_t v[kScanOnePer];
#pragma unroll
  for (int k = 0; k < kScanOnePer; ++k) {  // every load issued before any is used
    const uint64_t i = (uint64_t)tid + (uint64_t)k * kScanOneThreads;
    v[k] = i < n ? in[i] : 0u;
  }
#pragma unroll
  for (int k = 0; k < kScanOnePer; ++k) vals[tid + k * kScanOneThreads] = v[k];
  __syncthreads();
  const int chunk = (int)((n + kScanOneThreads - 1) / kScanOneThreads);
  const int b = tid * chunk, e = b + chunk < (int)n ? b + chunk : (int)n;
  uint64_t sum = 0;
  for (int i = b; i < e; ++i) sum += vals[i];
  const int lane = tid & 63, wave = tid >> 6;
  uint64_t x = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    const uint64_t y = ((uint64_t)(uint32_t)__shfl_up((int)hi, o, 64) << 32) | (uint32_t)__shfl_up((int)lo, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) red[wave] = x;
  __syncthreads();
  uint64_t run = x - sum, tot = 0;
#pragma unroll
  for (int i = 0; i < kScanOneThreads / 64; ++i) {
    if (i < wave) run += red[i];
    tot += red[i];
  }
  for (int i = b; i < e; ++i) {
    out[i] = run;
    run += vals[i];
  }
  if (tid == 0) out[n] = tot;
}

uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

}  // namespace

uint64_t scan_scratch_words(uint64_t n) { return div_up(n, kScanTile) + 1; }

int adjoint_blocks(uint64_t n) {
  const uint64_t nb = div_up(n, kBlock);
  return (int)(nb < 1 ? 1 : (nb > 2048 ? 2048 : nb));
}

// A recorded sparse gate, reversed in place on the adjoint: only the listed
// rows change (lam'[r] = cos lam[r] + sin s lam[q] for the record of row r
// whose branch landed in q with sign s), so the new values go to tmp first
// (a partner's old value may still be read) and k_adjoint_scatter writes them
// back.  One partial of d<O>/dtheta per block, as k_adjoint_gate.
__global__ __launch_bounds__(kBlock) void k_adjoint_sparse(uint64_t m, const SparseRec* tape, const double* lam,
                                                           double* tmp, double sin_t, double cos_t,
                                                           double* partials) {
  __shared__ double red[kBlock / 64];
  double acc = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kBlock) {
    const SparseRec t = tape[i];
    const double l = lam[t.row];
    double li = cos_t * l;
    acc += -sin_t * t.c * l;
    const uint32_t q = t.child & ~kRecSign;
    if (q != kRecNoChild) {
      const double sg = (t.child & kRecSign) ? -1.0 : 1.0;
      const double lq = lam[q];
      li += sin_t * sg * lq;
      acc += cos_t * sg * t.c * lq;
    }
    tmp[i] = li;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) t += red[i];
    partials[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_adjoint_scatter(uint64_t m, const SparseRec* tape, const double* tmp,
                                                            double* lam) {
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kBlock)
    lam[tape[i].row] = tmp[i];
}

hipError_t launch_adjoint_sparse(uint64_t m, const SparseRec* tape, double* lam, double* tmp, double sin_t,
                                 double cos_t, double* partials, int n_blocks, hipStream_t st) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adjoint_sparse, dim3(n_blocks), dim3(kBlock), 0, st, m, tape, lam, tmp, sin_t, cos_t,
                     partials);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(k_adjoint_scatter, dim3(n_blocks), dim3(kBlock), 0, st, m, tape, tmp, lam);
  return hipGetLastError();
}

// Batched engines: as k_adjoint_sparse with each record's instance angles;
// the partials of d<O_b>/dtheta gather per instance in LDS (a wave whose
// records share one instance adds once) and each block writes all n_inst of
// them, reduced per instance by k_inst_part_sum.
int sparse_inst_blocks(uint64_t m, int n_inst) {
  const int per = (1 << 18) / (n_inst > 1 ? n_inst : 1);
  const int cap = per > 64 ? per : 64, nb = adjoint_blocks(m);
  return nb < cap ? nb : cap;
}

__global__ __launch_bounds__(kBlock) void k_adjoint_sparse_batched(uint64_t m, const SparseRec* tape,
                                                                   const uint32_t* inst, const double* sc, int n_inst,
                                                                   const double* lam, double* tmp, double* part) {
  extern __shared__ double acc[];  // n_inst
  for (int b = threadIdx.x; b < n_inst; b += kBlock) acc[b] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  // the bound is uniform over the block, so every lane takes part in the
  // wave's ballots and shuffles
  for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < m; base += (uint64_t)gridDim.x * kBlock) {
    const uint64_t i = base + threadIdx.x;
    const bool ok = i < m;
    uint32_t b = 0;
    double gr = 0.0;
    if (ok) {
      const SparseRec t = tape[i];
      b = inst[i];
      const double sn = sc[3 * b], cs = sc[3 * b + 1];
      const double l = lam[t.row];
      double li = cs * l;
      gr = -sn * t.c * l;
      const uint32_t q = t.child & ~kRecSign;
      if (q != kRecNoChild) {
        const double sg = (t.child & kRecSign) ? -1.0 : 1.0;
        const double lq = lam[q];
        li += sn * sg * lq;
        gr += cs * sg * t.c * lq;
      }
      tmp[i] = li;
    }
    const uint64_t live = __ballot(ok);
    if (live == 0) continue;
    const int lead = __ffsll((unsigned long long)live) - 1;
    const uint32_t b0 = (uint32_t)__shfl((int)b, lead, 64);
    if (__ballot(ok && b != b0) == 0) {  // one instance in the wave: one add
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) gr += __shfl_xor(gr, o, 64);
      if (lane == lead) atomicAdd(&acc[b0], gr);
    } else if (ok) {
      atomicAdd(&acc[b], gr);
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < n_inst; b += kBlock) part[(uint64_t)blockIdx.x * n_inst + b] = acc[b];
}

// grads[b] = sum over the nb blocks of part[blk * n_inst + b]: 64 instances
// per block, four row groups, combined in LDS in a fixed order.
__global__ __launch_bounds__(kBlock) void k_inst_part_sum(const double* part, int nb, int n_inst, double* grads) {
  __shared__ double red[4][64];
  const int b = blockIdx.x * 64 + (threadIdx.x & 63), grp = threadIdx.x >> 6;
  double acc = 0.0;
  if (b < n_inst)
    for (int k = grp; k < nb; k += 4) acc += part[(uint64_t)k * n_inst + b];
  red[grp][threadIdx.x & 63] = acc;
  __syncthreads();
  if (grp == 0 && b < n_inst) grads[b] = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

hipError_t launch_adjoint_sparse_batched(uint64_t m, const SparseRec* tape, const uint32_t* inst, const double* sc,
                                         int n_inst, double* lam, double* tmp, double* part, double* grads,
                                         hipStream_t st) {
  if (n_inst <= 0 || n_inst > kSparseRecInstMax) return hipErrorInvalidValue;
  const int nb = sparse_inst_blocks(m, n_inst);
  if (m) {
    hipLaunchKernelGGL(k_adjoint_sparse_batched, dim3(nb), dim3(kBlock), (size_t)n_inst * sizeof(double), st, m, tape,
                       inst, sc, n_inst, lam, tmp, part);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_adjoint_scatter, dim3(adjoint_blocks(m)), dim3(kBlock), 0, st, m, tape, tmp, lam);
    err = hipGetLastError();
    if (err != hipSuccess) return err;
  }
  hipLaunchKernelGGL(k_inst_part_sum, dim3((unsigned)div_up((uint64_t)n_inst, 64)), dim3(kBlock), 0, st, part,
                     m ? nb : 0, n_inst, grads);
  return hipGetLastError();
}

// Row permutations of a recorded sort: origin[j] = j before the sort; after
// it origin[j] is the row now at j, and out[origin[j]] = j maps each row to
// its new place (a noise-style tape entry with every factor 1).
__global__ void k_iota_u32(uint32_t* o, uint64_t n) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
    o[j] = (uint32_t)j;
}
__global__ void k_perm_out(const uint32_t* origin, uint32_t* out, uint64_t n) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x)
    out[origin[j]] = (uint32_t)j;
}

hipError_t launch_iota_u32(uint32_t* o, uint64_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota_u32, dim3(adjoint_blocks(n)), dim3(kBlock), 0, st, o, n);
  return hipGetLastError();
}
hipError_t launch_perm_out(const uint32_t* origin, uint32_t* out, uint64_t n, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_perm_out, dim3(adjoint_blocks(n)), dim3(kBlock), 0, st, origin, out, n);
  return hipGetLastError();
}

hipError_t launch_adjoint_gate(uint64_t n_in, TapeDev tape, const double* c_in, const double* lam_out, double* lam_in,
                               double sin_t, double cos_t, double* partials, int n_blocks, hipStream_t st) {
  hipLaunchKernelGGL(k_adjoint_gate, dim3(n_blocks), dim3(kBlock), 0, st, n_in, tape, c_in, lam_out, lam_in, sin_t,
                     cos_t, partials);
  return hipGetLastError();
}

hipError_t launch_adjoint_gate_segments(uint64_t n_work, const uint32_t* work, const uint64_t* bounds,
                                       const double* sc, TapeDev tape, const double* c_in, const double* lam_out,
                                       double* lam_in, double* partials, hipStream_t st) {
  if (n_work == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adjoint_gate_segments, dim3((unsigned)n_work), dim3(kBlock), 0, st, work, bounds, sc, tape,
                     c_in, lam_out, lam_in, partials);
  return hipGetLastError();
}

hipError_t launch_instance_work(const uint64_t* bounds, int n_inst, uint32_t* work, uint32_t* off, uint32_t* total,
                                hipStream_t st) {
  hipLaunchKernelGGL(k_instance_work, dim3(1), dim3(kWorkThreads), 0, st, bounds, n_inst, work, off, total);
  return hipGetLastError();
}

hipError_t launch_instance_grads(const uint64_t* part_off, const uint64_t* rows, const uint32_t* const* offs,
                                 uint64_t n_gates, const double* partials, int n_inst, double* grads, hipStream_t st) {
  if (n_gates == 0) return hipSuccess;
  hipLaunchKernelGGL(k_instance_grads, dim3((unsigned)div_up((uint64_t)n_inst, kBlock), (unsigned)n_gates),
                     dim3(kBlock), 0, st, part_off, rows, offs, partials, n_inst, grads);
  return hipGetLastError();
}

__global__ void k_apply_keep(uint64_t e, const uint8_t* keep, uint8_t* flag, const uint32_t* origin,
                             uint8_t* own_keep) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < e; j += (uint64_t)gridDim.x * blockDim.x) {
    const uint8_t k = keep[j] ? 1 : 0;
    flag[j] = k ? kEmit : kPending;
    own_keep[origin[j] & ~kRecSign] = k;
  }
}

hipError_t launch_apply_keep(uint64_t e, const uint8_t* keep, uint8_t* flag, const uint32_t* origin,
                             uint8_t* own_keep, hipStream_t st) {
  if (e == 0) return hipSuccess;
  uint64_t nb = div_up(e, kBlock);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_apply_keep, dim3((unsigned)nb), dim3(kBlock), 0, st, e, keep, flag, origin, own_keep);
  return hipGetLastError();
}

hipError_t launch_adjoint_noise(uint64_t n_in, TapeDev tape, const double* lam_out, double* lam_in, hipStream_t st) {
  if (n_in == 0) return hipSuccess;
  hipLaunchKernelGGL(k_adjoint_noise, dim3(adjoint_blocks(n_in)), dim3(kBlock), 0, st, n_in, tape, lam_out, lam_in);
  return hipGetLastError();
}

hipError_t launch_gather_bounds(const uint64_t* offs, uint64_t n_blocks, int m, uint64_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_bounds, dim3(1), dim3(kBlock), 0, st, offs, n_blocks, m, out);
  return hipGetLastError();
}

hipError_t launch_count_below(const double* coeff, uint64_t n, double thr, unsigned long long* out,
                              hipStream_t st) {
  hipError_t err = hipMemsetAsync(out, 0, sizeof(unsigned long long), st);
  if (err != hipSuccess || n == 0) return err;
  uint64_t nb = div_up(n, kBlock);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL(k_count_below, dim3((unsigned)nb), dim3(kBlock), 0, st, coeff, n, thr, out);
  return hipGetLastError();
}

hipError_t launch_exclusive_scan(const uint32_t* in, uint64_t n, uint64_t* out, uint64_t* scratch,
                                 hipStream_t st) {
  const uint64_t nb = div_up(n, kScanTile);
  if (nb == 0) return hipMemsetAsync(out, 0, sizeof(uint64_t), st);
  if (n <= kScanOneMax) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(kScanOneThreads), 0, st, in, n, out);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kBlock), 0, st, in, n, scratch);
  hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)nb), dim3(kBlock), 0, st, in, n, scratch, out);
  return hipGetLastError();
}

// Persistent grids: enough blocks to fill the device, capped by the work.
static unsigned grid_for(uint64_t tiles) {
  const uint64_t cap = 2048;
  return (unsigned)(tiles < 1 ? 1 : (tiles < cap ? tiles : cap));
}

template <int W>
hipError_t launch_merge_rounds(ChildDev a, ChildDev b, const uint64_t* offs, uint64_t nb, int m, uint64_t e_max,
                               int* result_in_b, hipStream_t st) {
  constexpr uint64_t T = MergeCfg<W>::tile;
  int in_b = 0;
  for (int r = 0; (1 << r) < m; ++r) {
    const int pairs = m >> (r + 1);
    ChildDev src = in_b ? b : a, dst = in_b ? a : b;
    hipLaunchKernelGGL(k_merge_round<W>, dim3(grid_for(div_up(e_max, T) + pairs)), dim3(kBlock), 0, st, src, dst,
                       offs, nb, m, r);
    in_b ^= 1;
  }
  *result_in_b = in_b;
  return hipGetLastError();
}

template <int W>
hipError_t launch_sort_children(ChildDev a, ChildDev b, uint64_t e, int* result_in_b, hipStream_t st) {
  constexpr uint64_t T = MergeCfg<W>::tile;
  const uint64_t nt = div_up(e, T);
  if (nt == 0) {
    *result_in_b = 0;
    return hipSuccess;
  }
  hipLaunchKernelGGL(k_block_sort<W>, dim3((unsigned)nt), dim3(kBlock), 0, st, a, b, e);
  ChildDev src = b, dst = a;
  int in_b = 1;
  for (uint64_t L = T; L < e; L *= 2) {
    hipLaunchKernelGGL(k_merge_regular<W>, dim3(grid_for(nt + 1)), dim3(kBlock), 0, st, src, dst, L, e);
    ChildDev t = src;
    src = dst;
    dst = t;
    in_b ^= 1;
  }
  *result_in_b = in_b;
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::branch_count(const GateArgs<W>& g, StoreDev s, const uint64_t* n_dev, uint64_t nb,
                                         uint32_t* counts, int m, Counters* ctr, hipStream_t st) {
  hipLaunchKernelGGL((k_branch_count<KIND, W>), dim3((unsigned)nb), dim3(kBlock), 0, st, g, s, n_dev, counts, m, nb,
                     ctr);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::branch_scatter(const GateArgs<W>& g, StoreDev s, const uint64_t* n_dev, uint64_t nb,
                                           const uint64_t* offs, int m, ChildDev c, uint32_t* tape_child,
                                           hipStream_t st) {
  hipLaunchKernelGGL((k_branch_scatter<KIND, W>), dim3((unsigned)nb), dim3(kBlock), 0, st, g, s, n_dev, offs, m, nb,
                     c, tape_child);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::scale_anticommuting(const GateArgs<W>& g, StoreDev s, const uint64_t* n_dev,
                                                uint64_t n_max, double f, hipStream_t st) {
  uint64_t nb = div_up(n_max, kBlock);
  if (nb > 8192) nb = 8192;
  if (nb == 0) nb = 1;
  hipLaunchKernelGGL((k_scale_anti<KIND, W>), dim3((unsigned)nb), dim3(kBlock), 0, st, g, s, n_dev, f);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::store_merge(const GateArgs<W>& g, StoreDev s, const uint64_t* n_dev, uint64_t n_max,
                                        ChildDev c, const uint64_t* e_dev, uint64_t e_max, StoreDev out,
                                        uint64_t* splits, uint64_t* status, uint32_t* ticket, uint32_t* overflow,
                                        uint64_t* n_out, uint64_t* n_out_host, uint64_t* e_host, uint32_t tag,
                                        Counters* ctr, TapeDev tape, hipStream_t st, hipEvent_t ev0,
                                        hipEvent_t ev1) {
  const uint64_t tiles_max = div_up(n_max + e_max, StoreCfg<W>::tile);
  hipLaunchKernelGGL(k_store_splits<W>, dim3((unsigned)div_up(tiles_max + 1, kBlock / 64)), dim3(kBlock), 0, st,
                     s, n_dev, c, e_dev, out.cap, splits, status, ticket, overflow);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  // With events, hipExtLaunchKernelGGL stamps them from the kernel's own
  // dispatch start and end, so their interval is the kernel's run time
  // without the queue's launch latency.
  const dim3 grid(grid_for(tiles_max)), block(kBlock);
  if (tape.out)
    hipExtLaunchKernelGGL((k_store_merge<KIND, W, true>), grid, block, 0, st, ev0, ev1, 0, g, s, n_dev, c, e_dev, out,
                          splits, status, ticket, n_out, n_out_host, e_host, tag, ctr, tape);
  else
    hipExtLaunchKernelGGL((k_store_merge<KIND, W, false>), grid, block, 0, st, ev0, ev1, 0, g, s, n_dev, c, e_dev,
                          out, splits, status, ticket, n_out, n_out_host, e_host, tag, ctr, tape);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::sparse_gate(const GateArgs<W>& g, StoreDev s, const uint64_t* n_dev, uint64_t n_max,
                                        SparseEnt<W>* const ents[2], uint32_t* const slots[2], uint64_t ent_cap,
                                        uint64_t* const tables[2],
                                        uint64_t table_cap, uint64_t* sync_words, uint64_t seq, uint64_t* n_out,
                                        uint64_t* n_out_host, uint64_t* e_host, uint32_t pub_tag, Counters* ctr,
                                        SparseRec* tape, uint32_t* tape_inst, hipStream_t st, hipEvent_t g0,
                                        hipEvent_t g1, hipEvent_t j0, hipEvent_t j1) {
  if (ent_cap == 0 || table_cap == 0) return hipErrorInvalidValue;  // both kernels index the lists and tables
  const uint64_t nb = div_up(n_max, kSparseTile) < 1 ? 1 : div_up(n_max, kSparseTile);
  const int cur = (int)(seq & 1), prev = 1 - cur;
  uint64_t* rec = sync_words + (seq % 3) * kRecWords;  // three records, by pass number mod 3
  uint64_t* prev_rec = sync_words + ((seq + 2) % 3) * kRecWords;
  uint64_t* next_rec = sync_words + ((seq + 1) % 3) * kRecWords;
  unsigned long long* join_word = reinterpret_cast<unsigned long long*>(sync_words + 3 * kRecWords);
  hipExtLaunchKernelGGL((k_sparse_gather<KIND, W>), dim3((unsigned)nb), dim3(kBlock), 0, st, g0, g1, 0, g, s, n_dev,
                        ents[cur], slots[cur], ent_cap, rec, tables[cur], table_cap, join_word, ctr);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipExtLaunchKernelGGL((k_sparse_join<KIND, W>), dim3(grid_for(div_up(n_max, kJoinTile))), dim3(kBlock), 0, st, j0,
                        j1, 0, g, ents[cur], ent_cap, rec, tables[cur], table_cap, slots[prev], prev_rec,
                        tables[prev], next_rec, s, n_dev, n_out, n_out_host, e_host, pub_tag, join_word, ctr, tape,
                        tape_inst);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::pair_gate(const GateArgs<W>& ga, const GateArgs<W>& gb, StoreDev s, const uint64_t* n_dev,
                                      uint64_t n_max, SparseEnt<W>* const ents[2], uint32_t* const slots[2],
                                      uint64_t ent_cap, uint64_t* const tables[2], uint64_t table_cap,
                                      uint64_t* sync_words, uint64_t seq, uint64_t* mid, uint64_t* n_out,
                                      uint64_t* n_out_host, uint64_t* e_host, uint32_t pub_tag, Counters* ctr,
                                      hipStream_t st, hipEvent_t g0, hipEvent_t g1, hipEvent_t j0, hipEvent_t j1) {
  if (ent_cap == 0 || table_cap == 0) return hipErrorInvalidValue;  // both kernels index the lists and tables
  const uint64_t nb = div_up(n_max, kSparseTile) < 1 ? 1 : div_up(n_max, kSparseTile);
  const int cur = (int)(seq & 1), prev = 1 - cur;
  uint64_t* rec = sync_words + (seq % 3) * kRecWords;  // three records, by pass number mod 3
  uint64_t* prev_rec = sync_words + ((seq + 2) % 3) * kRecWords;
  uint64_t* next_rec = sync_words + ((seq + 1) % 3) * kRecWords;
  unsigned long long* join_word = reinterpret_cast<unsigned long long*>(sync_words + 3 * kRecWords);
  unsigned long long* mid_word = reinterpret_cast<unsigned long long*>(mid);
  hipExtLaunchKernelGGL((k_pair_gather<KIND, W>), dim3((unsigned)nb), dim3(kBlock), 0, st, g0, g1, 0, ga, gb, s,
                        n_dev, ents[cur], slots[cur], ent_cap, rec, tables[cur], table_cap, join_word, mid_word);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipExtLaunchKernelGGL((k_pair_join<KIND, W>), dim3(grid_for(div_up(n_max, kPairTile))), dim3(kBlock), 0, st, j0,
                        j1, 0, ga, gb, ents[cur], ent_cap, rec, tables[cur], table_cap, slots[prev], prev_rec,
                        tables[prev], next_rec, s, n_dev, n_out, n_out_host, e_host, pub_tag, join_word, mid_word,
                        ctr);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::small_run(const GateArgs<W>* gates, const uint8_t* kinds, int n_gates, StoreDev s,
                                      uint64_t* n_dev, uint64_t* n_after, uint32_t* done, Counters* ctr,
                                      hipStream_t st) {
  if constexpr (SmallCfg<W>::rows > 0) {
    hipLaunchKernelGGL((k_small_run<KIND, W>), dim3(1), dim3(kSmallThreads), 0, st, gates, kinds, n_gates, s, n_dev,
                       n_after, done, ctr);
    return hipGetLastError();
  } else {
    return hipErrorInvalidValue;
  }
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::noise_compact(const NoiseArgs& a, StoreDev s, uint64_t n, StoreDev out,
                                          uint64_t* status, uint64_t* n_out, int n_units, TapeDev tape,
                                          hipStream_t st, uint64_t* n_host, uint32_t tag) {
  const uint64_t n_tiles = div_up(n, kTile);
  if (n_tiles == 0) return hipMemsetAsync(n_out, 0, sizeof(uint64_t), st);
  // the look-back words and the ticket after them, cleared together
  uint32_t* ticket = reinterpret_cast<uint32_t*>(status + n_tiles);
  hipError_t err = hipMemsetAsync(status, 0, sizeof(uint64_t) * (n_tiles + 1), st);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL((k_noise_compact<KIND, W>), dim3((unsigned)n_tiles), dim3(kBlock), 0, st, a, s, n, out,
                     status, ticket, n_out, n_tiles, n_units, tape, n_host, tag);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::adjoint_seed_batched(StoreDev s, uint64_t n, const uint64_t* dms, int shift,
                                                 uint64_t unit_mask, int n_units, const uint64_t* img,
                                                 const double* sgn, int n_rows, double* lam, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((k_adjoint_seed_batched<KIND, W>), dim3(adjoint_blocks(n)), dim3(kBlock), 0, st, s, n, dms,
                     shift, unit_mask, n_units, img, sgn, n_rows, lam);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::adjoint_seed(StoreDev s, uint64_t n, Key<W> dm, int n_units, const uint64_t* img,
                                         const double* sgn, int n_rows, double* lam, hipStream_t st) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL((k_adjoint_seed<KIND, W>), dim3(adjoint_blocks(n)), dim3(kBlock), 0, st, s, n, dm, n_units, img,
                     sgn, n_rows, lam);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::expectation(StoreDev s, uint64_t n, Key<W> dm, int n_units, const uint64_t* img,
                                        const double* sgn, int n_rows, double* partials, int n_blocks,
                                        hipStream_t st, uint64_t unit_mask) {
  hipLaunchKernelGGL((k_expectation<KIND, W>), dim3(n_blocks), dim3(kBlock), 0, st, s, n, dm, n_units, img, sgn,
                     n_rows, partials, unit_mask);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::expectation_segments(StoreDev s, const uint32_t* work, uint64_t n_work,
                                                 const uint64_t* bounds, const uint64_t* dms, int n_units,
                                                 const uint64_t* img, const double* sgn, int n_rows,
                                                 double* partials, uint64_t unit_mask, hipStream_t st) {
  if (n_work == 0) return hipSuccess;
  hipLaunchKernelGGL((k_expectation_segments<KIND, W>), dim3((unsigned)n_work), dim3(kBlock), 0, st, s, work, bounds,
                     dms, n_units, img, sgn, n_rows, partials, unit_mask);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::weights(StoreDev s, uint64_t n, int n_units, uint32_t* out, hipStream_t st,
                                    uint64_t unit_mask) {
  uint64_t nb = div_up(n, kBlock);
  if (nb > 8192) nb = 8192;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL((k_weights<KIND, W>), dim3((unsigned)nb), dim3(kBlock), 0, st, s, n, n_units, out, unit_mask);
  return hipGetLastError();
}

hipError_t launch_instance_bounds(const uint64_t* top_plane, const uint64_t* n_dev, int shift, int n_inst,
                                  uint64_t* bounds, hipStream_t st) {
  hipLaunchKernelGGL(k_instance_bounds, dim3((unsigned)div_up((uint64_t)n_inst + 1, kBlock)), dim3(kBlock), 0, st,
                     top_plane, n_dev, shift, n_inst, bounds);
  return hipGetLastError();
}

template <int KIND, int W>
hipError_t Launch<KIND, W>::conjugate(StoreDev s, uint64_t n, const uint64_t* img, const double* sgn, int n_rows,
                                      uint64_t* out_keys, double* out_coeff, hipStream_t st) {
  uint64_t nb = div_up(n, kBlock);
  if (nb > 8192) nb = 8192;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL((k_conjugate<KIND, W>), dim3((unsigned)nb), dim3(kBlock), 0, st, s, n, img, sgn, n_rows,
                     out_keys, out_coeff);
  return hipGetLastError();
}

#define PQ_INSTANTIATE(W)                                                                                       \
  template struct Launch<kPauli, W>;                                                                            \
  template struct Launch<kMajorana, W>;                                                                         \
  template hipError_t launch_merge_rounds<W>(ChildDev, ChildDev, const uint64_t*, uint64_t, int, uint64_t, int*, \
                                             hipStream_t);                                                      \
  template hipError_t launch_sort_children<W>(ChildDev, ChildDev, uint64_t, int*, hipStream_t);
PQ_INSTANTIATE(1)
PQ_INSTANTIATE(2)
PQ_INSTANTIATE(4)
PQ_INSTANTIATE(8)

}  // namespace pq

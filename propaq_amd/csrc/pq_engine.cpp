// pq_engine.cpp -- host orchestration of the MI355X engine and its C ABI.
//
// The host keeps what is small and sequential: the truncation settings, the
// deferred Clifford frame (a stabilizer tableau of 2 n_units rows), the
// per-gate generator conjugation and the launch plan.  Everything that scales
// with the number of terms lives on the device (pq_kernels.hip).
//
// Reference call structure being replaced:
//   PartitionedTermSum::apply_rotation  crates/core/src/engine/partitioned_termsum.rs:276-350
//   CliffordTableau                     crates/core/src/algebra/tableau.rs:29-224
//   EmitCutoff::at_size                 crates/core/src/engine/termsum.rs:70
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <chrono>
#include <vector>

#include "../../include/propaq_amd.h"
#include "pq_kernels.h"
#include "pq_tableau.h"

namespace pq {
namespace {

thread_local std::string g_last_error;

struct Failure {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, const std::string& msg) { throw Failure{code, msg}; }

void hip_check(hipError_t e, const char* what) {
  if (e == hipSuccess) return;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    fail(PQ_ERR_MEMORY, std::string(what) + ": " + hipGetErrorString(e));
  fail(PQ_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPCHK(x) hip_check((x), #x)

constexpr double kCliffordEps = 1e-9;  // partitioned_termsum.rs:41
// Small device/pinned scratch: class bounds [0, 257), ticket + n_out, counts.
constexpr int kPinnedWords = 512;
constexpr int kSlotTicket = 300;
// the row count of a batched readout's extracted instances
constexpr int kSlotExtract = 304;
constexpr int kSlotCount = 310;
// Host-mapped row-count slots: (tag << 48) | count; a refused merge
// publishes kMirrorRefused.
constexpr uint64_t kMirrorMask = (1ULL << 48) - 1;
constexpr uint64_t kMirrorRefused = kMirrorMask;

// Adds the wall time of its scope to *acc (ms) when acc is set.
struct HostTimer {
  double* acc;
  std::chrono::steady_clock::time_point t0;
  explicit HostTimer(double* a) : acc(a) {
    if (acc) t0 = std::chrono::steady_clock::now();
  }
  ~HostTimer() {
    if (acc) *acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

// Device allocation that only grows.
struct DevMem {
  void* p = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  DevMem(const DevMem&) = delete;
  DevMem& operator=(const DevMem&) = delete;
  ~DevMem() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void ensure(size_t want) {
    if (want <= bytes) return;
    release();
    size_t b = std::max<size_t>(want, 256);
    HIPCHK(hipMalloc(&p, b));
    bytes = b;
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

// ------------------------------------------- allocation-free expectation
// pq_expectation_noalloc keeps its partial sums, the frame image and the
// total in this array instead of engine buffers, so the call allocates
// nothing.  Each device holds its own copy (hipGetSymbolAddress per device);
// in host code the array's name is not a device address, so only the looked
// up pointer is ever passed to kernels or copies.
constexpr size_t kScratchBytes = PQ_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up on the
// first call on that device (with it current) and kept.
void* scratch_for_current_device() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(PQ_ERR_DEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

// The partial sums added in index order by one thread: the same additions,
// in the same order, as the host loop of Engine::expectation.
__global__ void k_sum_in_order(const double* partials, int n, double* out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double acc = 0.0;
  for (int i = 0; i < n; ++i) acc += partials[i];
  *out = acc;
}

// ---------------------------------------------------------------- engine

struct EngineBase {
  virtual ~EngineBase() = default;
  int kind = 0, n_units = 0, n_words = 1;
  int has_maxw = 0, has_min = 0, has_min_terms = 0, defer = 1;
  int small_runs = 1;  // small stores: runs of gates in one workgroup (pq_set_small_runs)
  int sparse_runs = 1;  // larger stores: sparse in-place gates (pq_set_sparse)
  uint64_t sparse_budget = 0;  // bytes for their lists and tables, 0: free memory (pq_set_sparse_budget)
  int pair_runs = 1;  // consecutive commuting sparse gates share one pass (pq_set_pairs)
  uint64_t sparse_refused_at = ~0ULL;  // live count at which their buffers last did not fit
  // Sparse gates append rows out of key order; the sorted paths and the
  // order-dependent readouts sort the store first (sort_store).
  bool store_sorted = true;
  // Differentiable mode: every rotation runs through the branch path (no
  // Clifford deferral, no phase-only shortcut), so the branching of each gate
  // at any angle is in the record and the replay matches the forward run.
  int differentiable = 0;
  // Batched engine: n_inst instances share the store; instance b's rows
  // carry b in the top bits of key word W-1 (word >> inst_shift), above every
  // unit bit, so each instance is one contiguous, sorted run of rows.
  int n_inst = 0, inst_shift = 64;
  uint64_t unit_mask = ~0ULL;  // unit bits of key word W-1
  uint32_t maxw = 0;
  double min_coeff = 0.0;
  uint64_t min_terms = 0;
  uint64_t n = 0;
  // caller-side truncation kernel (pq_set_keep_kernel) and its layer context
  pq_keep_batch_fn keep_fn = nullptr;
  void* keep_ctx = nullptr;
  uint32_t layer_index = 0, n_layers_ctx = 0;
  uint64_t keep_declined = 0;  // branches the kernel held back
  pq_counters host_ctr{};
  bool profiling = false;
  uint64_t prof_every = 1;  // sample every k-th gate
  uint64_t prof_gate = 0;   // gates seen while profiling
  pq_profile prof{};
  virtual void seed(const uint64_t* keys, const double* coeffs, uint64_t cnt) = 0;
  // Applies n_gates rotations; n_terms (may be null) receives the live count
  // after each.
  virtual void apply_gates(const uint64_t* gens, const double* angles, uint64_t n_gates, uint64_t* n_terms) = 0;
  virtual void noise(const double* table, uint64_t len, const double* factors, const uint8_t* flags,
                     int reclaim) = 0;
  virtual double expectation(const uint64_t* diag, uint64_t nd) = 0;
  virtual double expectation_noalloc(const uint64_t* diag, uint64_t nd) = 0;
  virtual uint64_t terms_below(double thr) = 0;
  virtual uint64_t export_rows(uint64_t* keys, double* coeffs, uint64_t cap, bool frame) = 0;
  virtual void weights(uint32_t* out, uint64_t cap) = 0;
  virtual bool frame_identity() const = 0;
  virtual void counters(pq_counters* out) = 0;
  virtual void counters_reset() = 0;
  virtual uint64_t key_bytes() const = 0;
  virtual void sync() = 0;
  // differentiation
  virtual void tape_begin() = 0;
  virtual void tape_end() = 0;
  virtual void tape_info(uint64_t* n_gates, uint64_t* bytes) const = 0;
  virtual uint32_t checkpoint_save() = 0;
  virtual void checkpoint_restore(uint32_t id) = 0;
  virtual void checkpoint_free(uint32_t id) = 0;
  virtual void adjoint_seed(const uint64_t* diag, uint64_t nd) = 0;
  virtual void adjoint_backward(double* grads, uint64_t n_gates) = 0;
  virtual uint64_t adjoint_values(double* out, uint64_t cap) = 0;
  // batching
  virtual void seed_batched(const uint32_t* inst, const uint64_t* keys, const double* coeffs, uint64_t cnt) = 0;
  virtual void apply_gates_batched(const uint64_t* gens, const double* angles, uint64_t n_gates,
                                   uint64_t* n_terms) = 0;
  virtual void expectation_batched(const uint64_t* diag, uint64_t nd, double* out) = 0;
  virtual void instance_counts(uint64_t* out) = 0;
  virtual void retire(const uint8_t* flags) = 0;
  virtual void readout_retire(const uint64_t* diag, uint64_t nd, const uint8_t* flags, double* values,
                              uint64_t* counts) = 0;
  virtual void adjoint_seed_batched(const uint64_t* diag, uint64_t nd) = 0;
  virtual void freeze(const uint8_t* flags) = 0;
  virtual void adjoint_backward_batched(double* grads, uint64_t n_gates) = 0;
};

template <int KIND, int W>
struct Engine final : EngineBase {
  using L = Launch<KIND, W>;
  hipStream_t st = nullptr;
  struct Store {
    DevMem mem;
    uint64_t cap = 0;
    StoreDev dev() const {
      StoreDev s;
      s.key = mem.as<uint64_t>();
      s.coeff = reinterpret_cast<double*>(mem.as<uint64_t>() + (size_t)W * cap);
      s.cap = cap;
      return s;
    }
    void ensure(uint64_t want, bool slack = true) {
      if (want <= cap && mem.p) return;
      uint64_t c = std::max<uint64_t>(slack ? want + want / 2 : want, 1024);
      mem.ensure((size_t)c * (W + 1) * sizeof(uint64_t));
      cap = c;
    }
  } store[2];
  int cur = 0;
  struct Children {
    DevMem mem;
    uint64_t cap = 0;
    // key planes | values | origins (u32) | flags
    ChildDev dev(bool with_origin = false) const {
      ChildDev c;
      c.key = mem.as<uint64_t>();
      c.val = reinterpret_cast<double*>(c.key + (size_t)W * cap);
      uint32_t* org = reinterpret_cast<uint32_t*>(c.val + cap);
      c.origin = with_origin ? org : nullptr;
      c.flag = reinterpret_cast<uint8_t*>(org + cap);
      c.cap = cap;
      return c;
    }
    void ensure(uint64_t want) {
      if (want <= cap && mem.p) return;
      uint64_t c = std::max<uint64_t>(want + want / 2, 1024);
      mem.ensure((size_t)c * ((W + 1) * sizeof(uint64_t) + sizeof(uint32_t) + 1) + 64);
      cap = c;
    }
  } child[2];
  DevMem keep_dev;  // keep kernel: per-child decisions, then per-row own decisions
  std::vector<uint8_t> keep_host;
  std::vector<double> table_host;  // the noise table table_dev holds
  DevMem counts, offs, scan_scratch, splits, status, small, ctr_dev, partials, table_dev, frame_img, frame_sgn,
      row_aux, export_keys, export_coeff, sc_dev, inst_bounds, inst_work, inst_dm, inst_flags;
  std::vector<uint8_t> retired;  // batched: instances whose rows were removed
  std::vector<uint8_t> frozen;   // batched: instances whose rows later gates leave unchanged
  uint64_t* pinned = nullptr;  // host-pinned readback slots

  // Row counts live on the device in a ring of slots: a gate reads its input
  // count from nslot and its merge writes the output count to the next slot
  // (and to a host-mapped mirror), so gates are enqueued without waiting for
  // the host to learn the count.
  static constexpr int kSlots = 64;
  static constexpr int kRing = 8;  // gates in flight at most (events, mirrors)
  DevMem nslots;
  int nslot = 0;
  // host-mapped: [kRing] output counts, [kRing] child counts, overflow flag
  uint64_t* mirror_h = nullptr;
  uint64_t* mirror_d = nullptr;
  uint64_t n_ub = 0;             // upper bound of the live count at the end of the enqueued work
  int unresolved_merges = 0;     // enqueued store merges whose count the host has not read yet

  struct Pending {
    bool launched;    // kernels were enqueued (else the count is unchanged)
    bool merged;      // a store merge ran (the count comes from the mirror)
    bool sampled;     // phase events were recorded (profiling)
    int ring;         // event / mirror slot
    uint32_t tag;     // the merge publishes (tag << 48) | n_out to mirror[ring]
    uint64_t* n_terms_out;
    bool sparse = false;  // a sparse in-place gate (its mirror count is the join's)
    int nw = 0;           // sparse: key words the anticommutation test reads
    // A pair pass (two gates in one gather + join): the first gate's count is
    // read back from the device later (mid_slot), the second's is the pass's
    // published count; the pass's phase events sit in ev_ring.
    bool pair_first = false, pair_second = false;
    uint64_t mid_slot = 0;
    int ev_ring = -1;
    bool noise = false;  // a noise / reclaim pass (not a gate; its count comes from the mirror)
  };
  std::vector<Pending> pending;  // FIFO, front = oldest
  size_t pending_head = 0;
  uint64_t seq = 0;
  hipEvent_t ph_ev[kRing][2 * PQ_N_PHASES] = {};
  bool ph_open[kRing][PQ_N_PHASES] = {};
  double ph_bytes_pending[kRing] = {};

  int ring_of(uint64_t q) const { return (int)(q % kRing); }
  void* scratch_dev = nullptr;  // this device's allocation-free scratch array

  // Phase events are recorded on sampled gates only: every event marker
  // idles the queue for several microseconds between the kernels it sits
  // between, so recording them on every gate would slow the run it measures.
  bool sample_next() {
    if (!profiling) return false;
    return prof_gate++ % prof_every == 0;
  }
  void phase_begin(bool sampled, int ring, int ph) {
    if (!sampled) return;
    HIPCHK(hipEventRecord(ph_ev[ring][2 * ph], st));
  }
  void phase_end(bool sampled, int ring, int ph) {
    if (!sampled) return;
    HIPCHK(hipEventRecord(ph_ev[ring][2 * ph + 1], st));
    ph_open[ring][ph] = true;
    prof.launches[ph]++;
  }
  // Wait for the sampled gate's events and fold its phase intervals in.
  void phase_collect(int ring) {
    for (int ph = 0; ph < PQ_N_PHASES; ++ph) {
      if (!ph_open[ring][ph]) continue;
      HIPCHK(hipEventSynchronize(ph_ev[ring][2 * ph + 1]));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, ph_ev[ring][2 * ph], ph_ev[ring][2 * ph + 1]));
      prof.ms[ph] += ms;
      ph_open[ring][ph] = false;
    }
  }

  // Row count published by the store merge of a gate: the merge's last tile
  // stores (tag << 48) | n_out to the host-mapped mirror slot, so the host
  // learns it by polling, without an event marker in the queue.  After 50 ms
  // without the tag (and every 50 ms after that) the stream is queried, so a
  // failed launch surfaces as an error instead of a wait without end; the
  // query is kept off the fast path because it is not free for the queue.
  uint64_t wait_mirror(int ring, uint32_t tag) {
    HostTimer tm(profiling ? &prof.host_wait_ms : nullptr);
    auto next_check = std::chrono::steady_clock::now() + std::chrono::milliseconds(50);
    for (uint64_t spin = 1;; ++spin) {
      const uint64_t v = __atomic_load_n(&mirror_h[ring], __ATOMIC_ACQUIRE);
      if ((v >> 48) == tag) return v & kMirrorMask;
      if (spin % 256 == 0 && std::chrono::steady_clock::now() >= next_check) {
        next_check += std::chrono::milliseconds(50);
        const hipError_t q = hipStreamQuery(st);
        if (q == hipSuccess) {
          const uint64_t w = __atomic_load_n(&mirror_h[ring], __ATOMIC_ACQUIRE);
          if ((w >> 48) == tag) return w & kMirrorMask;
          fail(PQ_ERR_STATE, "the store merge did not publish its row count (engine bug)");
        }
        if (q != hipErrorNotReady) hip_check(q, "store merge");
      }
      __builtin_ia32_pause();
    }
  }
  Tableau<KIND, W> frame;
  bool frame_dirty = true;

  // ---- differentiation: gather tape, adjoint vector, checkpoints
  // kTapeGate: a gate on the chained kernels (every row's output slot and
  // branch slot); kTapeSparse: a sparse gate (records of its listed rows,
  // in place); kTapeNoise: a noise / reclaim pass or a row permutation.
  enum { kTapeGate = 0, kTapeNoise = 1, kTapeSparse = 2 };
  static bool is_gate(int type) { return type != kTapeNoise; }
  struct TapeEntry {
    int type;
    uint64_t n_in, n_out;
    double s, c;        // sin / cos of the rotation
    TapeDev t;
    const double* c_in; // coefficients at the gate's start
    SparseRec* srec = nullptr;  // kTapeSparse: n_list records
    uint32_t* rinst = nullptr;  // kTapeSparse on batched engines: each record's instance
    uint64_t n_list = 0;
    // batched engines: per-instance (sin, cos, flag), the instances' input
    // row ranges, and the adjoint work list (instance, block, blocks) with
    // its per-instance offsets, all built on the device; the item count lands
    // in host-mapped slot tot_slot
    const double* sc = nullptr;
    const uint64_t* bounds = nullptr;
    const uint32_t* work = nullptr;
    const uint32_t* work_off = nullptr;
    uint32_t tot_slot = 0;
  };
  // host-mapped work-item counts of the recorded batched gates (one per gate)
  uint32_t* work_tot_h = nullptr;
  uint32_t* work_tot_d = nullptr;
  static constexpr size_t kWorkTotSlots = (size_t)1 << 20;
  size_t work_tot_used = 0;
  DevMem grads_dev, grad_meta, zero_offs, inst_part;
  // work-item offsets of a gate with no rows: n_inst + 1 zeros
  const uint32_t* empty_offsets() {
    if (zero_offs.bytes < (size_t)(n_inst + 1) * sizeof(uint32_t)) {
      zero_offs.ensure((size_t)(n_inst + 1) * sizeof(uint32_t));
      HIPCHK(hipMemsetAsync(zero_offs.p, 0, (size_t)(n_inst + 1) * sizeof(uint32_t), st));
    }
    return zero_offs.as<uint32_t>();
  }
  bool recording = false;
  std::vector<TapeEntry> tape;
  std::vector<std::unique_ptr<DevMem>> arena;  // chunks, reused across tapes
  size_t arena_chunk = 0, arena_used = 0;
  uint64_t tape_bytes_used = 0;
  DevMem lam[2], grad_partials, lam_tmp;
  int lam_cur = 0;
  uint64_t lam_n = 0;
  bool lam_valid = false;
  struct Checkpoint {
    DevMem mem;
    uint64_t n = 0;
    bool sorted = true;
    Tableau<KIND, W> frame;
  };
  std::map<uint32_t, std::unique_ptr<Checkpoint>> ckpts;
  uint32_t next_ckpt = 1;

  void* tape_alloc(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    while (arena_chunk < arena.size() && arena_used + bytes > arena[arena_chunk]->bytes) {
      ++arena_chunk;
      arena_used = 0;
    }
    if (arena_chunk == arena.size()) {
      // room for several allocations of this size: a sparse gate reserves a
      // record for every row and returns what its list did not use
      // (tape_trim), so the next gate's reservation fits the same chunk
      std::unique_ptr<DevMem> m(new DevMem);
      m->ensure(std::max<size_t>(4 * bytes, (size_t)256 << 20));
      arena.push_back(std::move(m));
      arena_used = 0;
    }
    void* p = static_cast<char*>(arena[arena_chunk]->p) + arena_used;
    arena_used += bytes;
    tape_bytes_used += bytes;
    return p;
  }
  // Shrink the arena's last allocation p (of `bytes`) to `keep` bytes.
  void tape_trim(void* p, size_t bytes, size_t keep) {
    bytes = (bytes + 255) & ~(size_t)255;
    keep = (keep + 255) & ~(size_t)255;
    if (arena_chunk >= arena.size() || keep >= bytes) return;
    char* base = static_cast<char*>(arena[arena_chunk]->p);
    if (static_cast<char*>(p) + bytes != base + arena_used) return;  // not the last allocation
    arena_used -= bytes - keep;
    tape_bytes_used -= bytes - keep;
  }

  // Batched engines, while recording a gate (the live count is exact): keep
  // the gate's per-instance sin / cos, its input rows per instance, and the
  // adjoint work list that splits every instance's rows into the blocks a run
  // of that instance alone would use.
  void record_instances(TapeEntry& te, const double* batch_sc) {
    double* sc = static_cast<double*>(tape_alloc((size_t)n_inst * 3 * sizeof(double)));
    HIPCHK(hipMemcpyAsync(sc, batch_sc, (size_t)n_inst * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    uint64_t* bd = static_cast<uint64_t*>(tape_alloc((size_t)(n_inst + 1) * sizeof(uint64_t)));
    const StoreDev sd = store[cur].dev();
    HIPCHK(launch_instance_bounds(sd.key + (size_t)(W - 1) * sd.cap, nslot_ptr(nslot), inst_shift, n_inst, bd, st));
    // at most adjoint_blocks(rows) <= rows / kBlock + 1 items per instance
    const size_t items = (size_t)(n / kBlock) + (size_t)n_inst + 1;
    uint32_t* wd = static_cast<uint32_t*>(tape_alloc(items * 3 * sizeof(uint32_t)));
    uint32_t* off = static_cast<uint32_t*>(tape_alloc((size_t)(n_inst + 1) * sizeof(uint32_t)));
    if (!work_tot_h) {
      HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&work_tot_h), kWorkTotSlots * sizeof(uint32_t),
                           hipHostMallocMapped | hipHostMallocCoherent));
      HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&work_tot_d), work_tot_h, 0));
    }
    if (work_tot_used >= kWorkTotSlots) fail(PQ_ERR_STATE, "too many recorded batched gates in one tape");
    te.tot_slot = (uint32_t)work_tot_used++;
    HIPCHK(launch_instance_work(bd, n_inst, wd, off, work_tot_d + te.tot_slot, st));
    te.sc = sc;
    te.bounds = bd;
    te.work = wd;
    te.work_off = off;
  }

  void tape_reset() {
    tape.clear();
    arena_chunk = 0;
    arena_used = 0;
    tape_bytes_used = 0;
    work_tot_used = 0;
  }

  Engine(int n_units_) {
    kind = KIND;
    n_units = n_units_;
    n_words = W;
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&pinned), kPinnedWords * sizeof(uint64_t), hipHostMallocDefault));
    small.ensure(kPinnedWords * sizeof(uint64_t));
    HIPCHK(hipMemsetAsync(small.p, 0, kPinnedWords * sizeof(uint64_t), st));
    ctr_dev.ensure(kCtrSlots * sizeof(Counters));
    HIPCHK(hipMemsetAsync(ctr_dev.p, 0, kCtrSlots * sizeof(Counters), st));
    for (auto& r : ph_ev)
      for (auto& x : r) HIPCHK(hipEventCreate(&x));
    HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&mirror_h), (2 * kRing + 1) * sizeof(uint64_t),
                         hipHostMallocMapped | hipHostMallocCoherent));
    HIPCHK(hipHostGetDevicePointer(reinterpret_cast<void**>(&mirror_d), mirror_h, 0));
    std::memset(mirror_h, 0, (2 * kRing + 1) * sizeof(uint64_t));
    nslots.ensure(kSlots * sizeof(uint64_t));
    HIPCHK(hipMemsetAsync(nslots.p, 0, kSlots * sizeof(uint64_t), st));
    frame.init(n_units);
    store[0].ensure(1024);
    store[1].ensure(1024);
    scratch_dev = scratch_for_current_device();
    HIPCHK(hipStreamSynchronize(st));
  }

  ~Engine() override {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& r : ph_ev)
      for (auto& x : r)
        if (x) (void)hipEventDestroy(x);
    if (mirror_h) (void)hipHostFree(mirror_h);
    if (mid_pinned) (void)hipHostFree(mid_pinned);
    if (small_host) (void)hipHostFree(small_host);
    if (work_tot_h) (void)hipHostFree(work_tot_h);
    if (pinned) (void)hipHostFree(pinned);
    if (st) (void)hipStreamDestroy(st);
  }

  // Wait for every enqueued gate and fold its results in; afterwards n is
  // exact and n_ub == n.
  void sync() override {
    flush_held();
    sync_enqueued();
  }

  // sync() without launching a held gate first.
  void sync_enqueued() {
    // the pair passes' first-gate counts: copied behind the last pass, so
    // the stream wait below covers the copy too
    if (pair_mid_n) {
      if (!mid_pinned) HIPCHK(hipHostMalloc(reinterpret_cast<void**>(&mid_pinned),
                                            kPairMidSlots * kMidWords * sizeof(uint64_t), hipHostMallocDefault));
      HIPCHK(hipMemcpyAsync(mid_pinned, pair_mid.p, pair_mid_n * kMidWords * sizeof(uint64_t), hipMemcpyDeviceToHost,
                            st));
    }
    while (pending_head < pending.size()) resolve_front();
    pending.clear();
    pending_head = 0;
    HIPCHK(hipStreamSynchronize(st));
    n_ub = n;
    unresolved_merges = 0;
    read_mids();
  }

  void resolve_front() {
    Pending& p = pending[pending_head++];
    if (p.noise) {  // rows only drop: the bound of the work after it stays valid
      const uint64_t v = wait_mirror(p.ring, p.tag);
      if (p.sampled) {
        prof.bytes[3] += (double)(n + v) * (8.0 * W + 8.0);
        phase_collect(p.ring);
      }
      n = v;
      n_ub = unresolved_merges < 40 ? n << unresolved_merges : ~0ULL >> 1;
      return;
    }
    if (p.pair_first) {  // its count after the gate comes from the device later (read_mids)
      host_ctr.term_gate_updates += n;
      mid_fix.push_back(MidFix{p.mid_slot, n, 0, p.n_terms_out, p.sampled});
      return;
    }
    if (p.pair_second) {
      const uint64_t v = wait_mirror(p.ring, p.tag);
      if (v == kMirrorRefused || mirror_h[2 * kRing])
        fail(PQ_ERR_STATE, "store capacity bound violated (engine bug)");
      n = v;
      mid_fix.back().n_after = v;  // this gate's start count is added with the first gate's result
      unresolved_merges -= 2;
      n_ub = unresolved_merges < 40 ? n << unresolved_merges : ~0ULL >> 1;
      if (p.n_terms_out) *p.n_terms_out = n;
      if (p.sampled) phase_collect(p.ev_ring);
      return;
    }
    host_ctr.term_gate_updates += n;  // live count at this gate's start
    if (p.merged) {
      const uint64_t v = wait_mirror(p.ring, p.tag);
      if (v == kMirrorRefused || mirror_h[2 * kRing])
        fail(PQ_ERR_STATE, "store capacity bound violated (engine bug)");
      const uint64_t n_in = n, e = __atomic_load_n(&mirror_h[kRing + p.ring], __ATOMIC_ACQUIRE);
      n = v;
      last_list_len = e;
      if (p.sampled && p.sparse) {  // e = the gate's anticommuting rows
        const double ent = sizeof(SparseEnt<W>), a = (double)e, rows = (double)n_in;
        // gather: the test words of every row, the rest of the A rows, A
        // entries, slot indices and table slots written; join: A entries,
        // table slots and slot indices read, A coefficients and the new rows
        // written
        prof.bytes[5] += rows * 8.0 * p.nw + a * (8.0 * (W - p.nw) + 8.0) + a * (ent + 4.0 + 8.0);
        prof.bytes[7] += a * (ent + 4.0 + 8.0 + 8.0) + (double)(n - n_in) * (8.0 * W + 8.0);
        // the gate's algorithmic bytes: one read and one write of the store
        prof.bytes[6] += (double)(n_in + n) * (8.0 * W + 8.0);
        prof.launches[6]++;
      } else if (p.sampled) {
        prof.bytes[2] += (double)(n_in + e + n) * (8.0 * W + 8.0);
      }
      // each merge at most doubles the rows: tighten the bound of the work
      // still in flight to the exact count times 2^(merges in flight)
      --unresolved_merges;
      n_ub = unresolved_merges < 40 ? n << unresolved_merges : ~0ULL >> 1;
    }
    if (p.n_terms_out) *p.n_terms_out = n;
    if (p.sampled) phase_collect(p.ring);
  }

  uint64_t* nslot_ptr(int k) { return nslots.as<uint64_t>() + (k % kSlots); }
  uint64_t last_list_len = 0;  // the list length the last resolved sparse gate published

  // A recorded gate has been resolved: its output count, and for a sparse
  // gate the records its list used (the rest of its allocation is returned).
  void close_recorded_gate() {
    if (!recording || tape.empty() || !is_gate(tape.back().type)) return;
    TapeEntry& te = tape.back();
    te.n_out = n;
    if (te.type == kTapeSparse) {
      te.n_list = last_list_len;
      tape_trim(te.srec, te.n_in * sizeof(SparseRec), te.n_list * sizeof(SparseRec));
    }
  }

  Key<W> key_from(const uint64_t* w) const {
    Key<W> k;
    for (int i = 0; i < W; ++i) k.w[i] = w[i];
    return k;
  }

  Key<W> unit_key_bits(const Key<W>& k) const {
    Key<W> m = k;
    m.w[W - 1] &= unit_mask;
    return m;
  }

  bool admits_initial(const Key<W>& k, double c) const {
    if (has_maxw && (uint32_t)weight<KIND, W>(unit_key_bits(k), n_units) > maxw) return false;
    if (has_min && !(std::fabs(c) >= min_coeff)) return false;
    return true;
  }

  // run_at_width seed admission (crates/pauli/src/engine.rs:104-111)
  void seed(const uint64_t* keys, const double* coeffs, uint64_t cnt) override {
    sync();
    recording = false;
    tape_reset();
    lam_valid = false;
    std::vector<std::pair<Key<W>, double>> rows;
    std::map<std::vector<uint64_t>, size_t> index;
    std::vector<uint8_t> kk;
    if (keep_fn && cnt) {  // admits_initial = the kernel's keep (termsum.rs:152)
      std::vector<uint32_t> wt(cnt);
      std::vector<double> mag(cnt);
      for (uint64_t i = 0; i < cnt; ++i) {
        wt[i] = (uint32_t)weight<KIND, W>(unit_key_bits(key_from(keys + i * W)), n_units);
        mag[i] = std::fabs(coeffs[i]);
      }
      kk = call_keep(keys, wt.data(), mag.data(), cnt);
    }
    for (uint64_t i = 0; i < cnt; ++i) {
      Key<W> k = key_from(keys + i * W);
      if (keep_fn ? !kk[i] : !admits_initial(k, coeffs[i])) continue;
      std::vector<uint64_t> kk(k.w, k.w + W);
      auto it = index.find(kk);
      if (it != index.end()) {
        rows[it->second].second += coeffs[i];
      } else {
        index.emplace(std::move(kk), rows.size());
        rows.emplace_back(k, coeffs[i]);
      }
    }
    std::sort(rows.begin(), rows.end(),
              [](const auto& a, const auto& b) { return key_less<W>(a.first, b.first); });
    n = rows.size();
    cur = 0;
    store[0].ensure(std::max<uint64_t>(n, 1));
    // only the n seeded rows of each plane cross the bus, not the capacity
    const StoreDev sd = store[0].dev();
    std::vector<uint64_t> host((size_t)(W + 1) * std::max<uint64_t>(n, 1));
    for (uint64_t r = 0; r < n; ++r) {
      for (int i = 0; i < W; ++i) host[(size_t)i * n + r] = rows[r].first.w[i];
      std::memcpy(&host[(size_t)W * n + r], &rows[r].second, sizeof(double));
    }
    if (n) {
      for (int i = 0; i < W; ++i)
        HIPCHK(hipMemcpyAsync(sd.key + (size_t)i * sd.cap, &host[(size_t)i * n], n * sizeof(uint64_t),
                              hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(sd.coeff, &host[(size_t)W * n], n * sizeof(double), hipMemcpyHostToDevice, st));
    }
    pinned[0] = n;
    HIPCHK(hipMemcpyAsync(nslot_ptr(nslot), pinned, sizeof(uint64_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    n_ub = n;
    store_sorted = true;
    frame.init(n_units);
    frame_dirty = true;
  }

  struct Eff {
    int has_maxw, has_min;
  };
  Eff at_size() const {  // termsum.rs:70
    Eff e{has_maxw, has_min};
    if (has_min_terms && n < min_terms) e.has_maxw = e.has_min = 0;
    return e;
  }

  GateArgs<W> make_args(const Key<W>& g, double sign, double s, double c, const Eff& eff) const {
    GateArgs<W> a{};
    a.gen = g;
    for (int i = 0; i < W; ++i)
      a.fold.w[i] = KIND == kPauli ? (((g.w[i] & kEven) << 1) | ((g.w[i] >> 1) & kEven)) : g.w[i];
    a.above = above_parity_mask<W>(g);
    a.gen_y = pauli_y<W>(g);
    a.gen_len = key_count<W>(g);
    a.neg = sign < 0.0;
    a.n_units = n_units;
    a.sin_t = s;
    a.cos_t = c;
    a.has_min = eff.has_min;
    a.min_coeff = min_coeff;
    a.has_maxw = eff.has_maxw;
    a.max_weight = maxw;
    a.unit_mask = unit_mask;
    a.sc = nullptr;
    a.inst_shift = inst_shift;
    // children are bucketed by the row's bits at the generator's support;
    // a generator with more bits than buckets allow uses the generic sort
    const int k = key_count<W>(g);
    a.n_cls_bits = 0;
    if (k <= kMaxClassBits) {
      for (int wd = 0; wd < W; ++wd) {
        uint64_t bits = g.w[wd];
        while (bits) {
          const int b = __builtin_ctzll(bits);
          bits &= bits - 1;
          a.cls_pos[a.n_cls_bits++] = wd * 64 + b;
        }
      }
    }
    return a;
  }

  uint64_t read_u64(const uint64_t* dev) {
    HIPCHK(hipMemcpyAsync(pinned, dev, sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return pinned[0];
  }

  // Grow the work buffers so a gate whose input holds at most n_in rows fits:
  // e <= n_in children, at most 2 n_in output rows.  Growing waits for the
  // enqueued gates first (their buffers are in use), which also tightens the
  // bound to the exact count.
  void ensure_capacity(uint64_t& n_in, int m) {
    auto fits = [&](uint64_t b) {
      const uint64_t nb = (b + kTile - 1) / kTile;
      const uint64_t tiles = (2 * b + StoreCfg<W>::tile - 1) / StoreCfg<W>::tile;
      return store[0].cap >= 2 * b && store[1].cap >= 2 * b && child[0].cap >= b && child[1].cap >= b &&
             counts.bytes >= (size_t)m * nb * sizeof(uint32_t) &&
             offs.bytes >= ((size_t)m * nb + 1) * sizeof(uint64_t) &&
             scan_scratch.bytes >= scan_scratch_words((uint64_t)m * nb) * sizeof(uint64_t) &&
             splits.bytes >= (size_t)(tiles + 2) * 2 * sizeof(uint64_t) &&
             status.bytes >= (size_t)(tiles + 1) * sizeof(uint64_t);
    };
    if (fits(n_in)) return;
    sync();
    n_in = n;
    // Grow unless the bound of one more gate in flight (2n) fits too:
    // growing only when n itself no longer fits would leave every later gate
    // between cap/4 and cap/2 live rows to drain the pipeline again.
    if (fits(2 * n_in)) return;
    const uint64_t b = std::max<uint64_t>(n_in * 2, 4096);
    const uint64_t nb = (b + kTile - 1) / kTile;
    const uint64_t tiles = (2 * b + StoreCfg<W>::tile - 1) / StoreCfg<W>::tile;
    grow_store(cur, 2 * b, true);
    grow_store(1 - cur, 2 * b, false);
    child[0].ensure(b);
    child[1].ensure(b);
    counts.ensure((size_t)kMaxClasses * nb * sizeof(uint32_t));
    offs.ensure(((size_t)kMaxClasses * nb + 1) * sizeof(uint64_t));
    scan_scratch.ensure(scan_scratch_words((uint64_t)kMaxClasses * nb) * sizeof(uint64_t));
    splits.ensure((size_t)(tiles + 2) * 2 * sizeof(uint64_t));
    status.ensure((size_t)(tiles + 1) * sizeof(uint64_t));
  }

  // Keep decisions of the caller's kernel for cnt rows given as W planes of
  // keys (plane stride `stride`) and coefficients, all in device memory.
  std::vector<uint8_t> keep_rows(const uint64_t* key_planes, uint64_t stride, const double* vals, uint64_t cnt) {
    if (!cnt) return {};
    std::vector<uint64_t> planes((size_t)W * cnt);
    std::vector<double> v(cnt);
    for (int w = 0; w < W; ++w)
      HIPCHK(hipMemcpyAsync(planes.data() + (size_t)w * cnt, key_planes + (size_t)w * stride, cnt * sizeof(uint64_t),
                            hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(v.data(), vals, cnt * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    std::vector<uint64_t> words((size_t)W * cnt);
    std::vector<uint32_t> wt(cnt);
    std::vector<double> mag(cnt);
    for (uint64_t r = 0; r < cnt; ++r) {
      Key<W> k;
      for (int w = 0; w < W; ++w) k.w[w] = words[r * W + w] = planes[(size_t)w * cnt + r];
      wt[r] = (uint32_t)weight<KIND, W>(unit_key_bits(k), n_units);
      mag[r] = std::fabs(v[r]);
    }
    return call_keep(words.data(), wt.data(), mag.data(), cnt);
  }

  std::vector<uint8_t> call_keep(const uint64_t* words, const uint32_t* wt, const double* mag, uint64_t cnt) {
    std::vector<uint8_t> keep(cnt, 1);
    if (cnt && keep_fn(keep_ctx, (uint32_t)KIND, words, (size_t)W, (uint32_t)n_units, wt, mag, layer_index,
                       n_layers_ctx, keep.data(), (size_t)cnt) != 0)
      fail(PQ_ERR_STATE, "the keep kernel reported a failure");
    return keep;
  }

  // Reallocate store[i] to hold want rows (and half again unless exact),
  // keeping its n live rows if asked.
  void grow_store(int i, uint64_t want, bool keep, bool exact = false) {
    if (store[i].cap >= want) return;
    Store fresh;
    fresh.ensure(want, !exact);
    if (keep && n) {
      const StoreDev a = store[i].dev(), b = fresh.dev();
      for (int w = 0; w < W; ++w)
        HIPCHK(hipMemcpyAsync(b.key + (size_t)w * b.cap, a.key + (size_t)w * a.cap, n * sizeof(uint64_t),
                              hipMemcpyDeviceToDevice, st));
      HIPCHK(hipMemcpyAsync(b.coeff, a.coeff, n * sizeof(double), hipMemcpyDeviceToDevice, st));
      HIPCHK(hipStreamSynchronize(st));
    }
    std::swap(store[i].mem.p, fresh.mem.p);
    std::swap(store[i].mem.bytes, fresh.mem.bytes);
    std::swap(store[i].cap, fresh.cap);
  }

  // ---- small-store runs
  // one device blob per run: gate arguments, kinds, then *done and n_after;
  // staged through a pinned host copy of the same layout
  DevMem small_dev;
  void* small_host = nullptr;
  size_t small_host_bytes = 0;
  static constexpr uint64_t kSmallRunGates = 1024;  // gates per launch

  // A run is tried while the store holds at most 3/4 of the row limit (the
  // kernel checks every gate's new rows against it); one that could not
  // apply a single gate blocks further tries until the next apply_gates.
  bool small_blocked = false;
  bool small_path() const {
    return SmallCfg<W>::rows > 0 && small_runs && !small_blocked && !recording && !keep_fn && n_inst == 0 &&
           !has_min_terms && n_ub <= (uint64_t)SmallCfg<W>::rows * 3 / 4;
  }

  // The host side of enqueue_gate for one gate of a small run: counts it,
  // folds a deferred Clifford into the frame (false), or fills the gate's
  // arguments and kind (0 = branch, 1 = phase-only).
  bool small_gate(const uint64_t* gen_words, double angle, GateArgs<W>& a, uint8_t& kind) {
    host_ctr.gates++;
    const double s = std::sin(angle), c = std::cos(angle);
    Eff eff = at_size();
    const Key<W> gin = key_from(gen_words);
    if (defer && std::fabs(c) < kCliffordEps &&
        (!eff.has_maxw || !Tableau<KIND, W>::rotation_changes_weight(frame.n_rows, gin))) {
      if (!differentiable) {
        frame.compose_rotation(gin, angle);
        frame_dirty = true;
        host_ctr.deferred++;
        return false;
      }
      eff.has_maxw = eff.has_min = 0;
    }
    Key<W> g;
    const double sign = frame.conjugate_generator(gin, g);
    a = make_args(g, sign, s, c, eff);
    kind = !differentiable && std::fabs(s) < kCliffordEps ? 1 : 0;
    return true;
  }

  // Apply up to kSmallRunGates gates in one launch of the small-store kernel
  // (the store holds at most half its row limit).  Returns the gates consumed;
  // the kernel stops before a gate that could overflow it, and the host state
  // (frame, counters) is then replayed up to that gate.
  uint64_t small_run(const uint64_t* gens, const double* angles, uint64_t count, uint64_t* n_terms) {
    sync();
    if (n > (uint64_t)SmallCfg<W>::rows * 3 / 4) return 0;
    const uint64_t take = std::min(count, kSmallRunGates);
    if (n == 0) {  // empty store: the gates do nothing, not even fold into the frame (partitioned_termsum.rs:282)
      host_ctr.gates += take;
      if (n_terms) std::fill(n_terms, n_terms + take, 0);
      return take;
    }
    const Tableau<KIND, W> frame0 = frame;
    const bool dirty0 = frame_dirty;
    const uint64_t gates0 = host_ctr.gates, deferred0 = host_ctr.deferred;
    const size_t args_bytes = (size_t)take * sizeof(GateArgs<W>);
    const size_t kinds_off = args_bytes, res_off = args_bytes + ((take + 7) & ~(uint64_t)7);
    const size_t blob = res_off + (take + 1) * sizeof(uint64_t);
    if (small_host_bytes < blob) {
      if (small_host) HIPCHK(hipHostFree(small_host));
      small_host = nullptr;
      HIPCHK(hipHostMalloc(&small_host, blob, hipHostMallocDefault));
      small_host_bytes = blob;
    }
    char* hb = static_cast<char*>(small_host);
    GateArgs<W>* args = reinterpret_cast<GateArgs<W>*>(hb);
    uint8_t* kinds = reinterpret_cast<uint8_t*>(hb + kinds_off);
    std::vector<int64_t> dev_of(take, -1);
    int n_args = 0;
    for (uint64_t i = 0; i < take; ++i) {
      GateArgs<W> a;
      uint8_t kind;
      if (small_gate(gens + i * W, angles[i], a, kind)) {
        dev_of[i] = n_args;
        args[n_args] = a;
        kinds[n_args] = kind;
        ++n_args;
      }
    }
    const uint64_t* res_h = reinterpret_cast<const uint64_t*>(hb + res_off);
    uint32_t done = 0;
    if (n_args) {
      grow_store(cur, (uint64_t)SmallCfg<W>::rows, true);
      small_dev.ensure(blob);
      char* db = small_dev.as<char>();
      HIPCHK(hipMemcpyAsync(db, hb, res_off, hipMemcpyHostToDevice, st));
      uint64_t* res = reinterpret_cast<uint64_t*>(db + res_off);
      HIPCHK(L::small_run(reinterpret_cast<const GateArgs<W>*>(db), reinterpret_cast<const uint8_t*>(db + kinds_off),
                          n_args, store[cur].dev(), nslot_ptr(nslot), res + 1, reinterpret_cast<uint32_t*>(res),
                          ctr_dev.as<Counters>(), st));
      HIPCHK(hipMemcpyAsync(hb + res_off, res, (n_args + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      done = (uint32_t)(res_h[0] & 0xFFFFFFFFu);
      store_sorted = true;  // the run writes its rows back sorted
    }
    const uint64_t* after = res_h + 1;
    uint64_t used = take;
    for (uint64_t i = 0; i < take; ++i) {
      if (dev_of[i] >= (int64_t)done) {  // not applied: the run stopped here
        used = i;
        break;
      }
      host_ctr.term_gate_updates += n;  // live count at this gate's start
      if (dev_of[i] >= 0) n = after[dev_of[i]];
      if (n_terms) n_terms[i] = n;
    }
    if (used < take) {  // replay the host side of the applied gates only
      frame = frame0;
      frame_dirty = dirty0;
      host_ctr.gates = gates0;
      host_ctr.deferred = deferred0;
      for (uint64_t i = 0; i < used; ++i) {
        GateArgs<W> a;
        uint8_t kind;
        (void)small_gate(gens + i * W, angles[i], a, kind);
      }
    }
    n_ub = n;
    host_ctr.small_gates += used;
    if (used == 0) small_blocked = true;
    return used;
  }

  void apply_gates(const uint64_t* gens, const double* angles, uint64_t n_gates, uint64_t* n_terms) override {
    small_blocked = false;
    // a noise pass in flight leaves the bound at its input count: near the
    // small-store limit its exact count decides the path of these gates
    if (SmallCfg<W>::rows > 0 && small_runs && pending_head < pending.size() && pending.back().noise &&
        n_ub <= (uint64_t)SmallCfg<W>::rows * 16)
      sync();
    for (uint64_t g = 0; g < n_gates; ++g) {
      if (small_path()) {
        const uint64_t used = small_run(gens + g * W, angles + g, n_gates - g, n_terms ? n_terms + g : nullptr);
        if (used) {
          g += used - 1;
          continue;
        }
      }
      enqueue_gate(gens + g * W, angles[g], n_terms ? n_terms + g : nullptr);
      // keep one gate queued behind the running one; learn counts lagged.
      // A recorded gate is resolved at once: its tape slices are sized by
      // the exact count.
      const size_t lag = recording ? 0 : 1;
      while (pending.size() - pending_head > lag) resolve_front();
      close_recorded_gate();
      if (pending_head > 64) {
        pending.erase(pending.begin(), pending.begin() + pending_head);
        pending_head = 0;
      }
    }
    flush_held();
    if (!mid_fix.empty() || pair_mid_n) {  // the pairs' first counts: read once the device is done
      sync_enqueued();
      return;
    }
    while (pending_head < pending.size()) resolve_front();
    pending.clear();
    pending_head = 0;
    n_ub = n;
  }

  // PartitionedTermSum::apply_rotation (partitioned_termsum.rs:276-350),
  // enqueued without waiting for the device.
  // batch_sc: the gate's device table (sin, cos, phase-only flag per
  // instance) on a batched engine; batch_angles: the same angles on the host.
  void enqueue_gate(const uint64_t* gen_words, double angle, uint64_t* n_terms_out, const double* batch_sc = nullptr,
                    const double* batch_angles = nullptr) {
    host_ctr.gates++;
    HostTimer tm(profiling ? &prof.host_enqueue_ms : nullptr);
    if (profiling) prof.host_gates++;
    bool uniform = true;  // every live instance rotates by the same angle
    if (batch_angles) {
      int first = -1;
      for (int b = 0; b < n_inst; ++b) {
        if (retired[b] || frozen[b]) continue;
        if (first < 0) {
          first = b;
        } else if (std::memcmp(&batch_angles[b], &batch_angles[first], sizeof(double)) != 0) {
          uniform = false;
        }
      }
      angle = first >= 0 ? batch_angles[first] : 0.0;
    }
    if (has_min_terms || recording || keep_fn) sync();  // at_size / the tape / the keep kernel need the exact count
    const int ring = ring_of(seq);
    const uint32_t tag = (uint32_t)(seq % 0xFFFF) + 1;  // never 0, differs from the slot's previous user
    ++seq;
    Pending P{false, false, false, ring, tag, n_terms_out};
    const double s = std::sin(angle), c = std::cos(angle);
    TapeDev td{nullptr, nullptr, nullptr};
    if (recording) {  // the path the gate takes fills in its tape (tape_chained / enqueue_sparse)
      if (2 * n > kRecMaxRows) fail(PQ_ERR_STATE, "too many rows to record a gate (2^31 limit)");
      tape.push_back(TapeEntry{kTapeGate, n, n, s, c, td, nullptr});
      rec_sc = batch_sc;
    }
    if (n_ub == 0) {  // empty store: nothing to do (partitioned_termsum.rs:282)
      flush_held();
      pending.push_back(P);
      return;
    }
    Eff eff = at_size();
    const Key<W> gin = key_from(gen_words);
    // A Clifford shared by every live instance folds into the common frame.
    // In differentiable mode it stays on the branch path (its derivative is
    // in the tape) but branches with no truncation predicate, so no branch
    // is lost that the folded run would have kept (tableau.rs:180-209).
    if (defer && uniform && std::fabs(c) < kCliffordEps &&
        (!eff.has_maxw || !Tableau<KIND, W>::rotation_changes_weight(frame.n_rows, gin))) {
      if (!differentiable) {
        flush_held();  // its count is the one after the held gate
        frame.compose_rotation(gin, angle);
        frame_dirty = true;
        host_ctr.deferred++;
        pending.push_back(P);
        return;
      }
      eff.has_maxw = eff.has_min = 0;
    }
    Key<W> g;
    const double sign = frame.conjugate_generator(gin, g);
    GateArgs<W> a = make_args(g, sign, s, c, eff);
    a.sc = batch_sc;
    // a keep kernel replaces the precheck, the weight bound and the
    // coefficient test (lossless below the term budget: at_size drops it too)
    const bool keep_mode = keep_fn && !(has_min_terms && n < min_terms);
    if (keep_mode) a.has_min = a.has_maxw = 0;
    P.launched = true;
    P.sampled = sample_next();
    if (!batch_sc && !differentiable && std::fabs(s) < kCliffordEps) {  // phase-only (coeff.rs:200)
      flush_held();
      phase_begin(P.sampled, ring, 4);
      HIPCHK(L::scale_anticommuting(a, store[cur].dev(), nslot_ptr(nslot), n_ub, c, st));
      phase_end(P.sampled, ring, 4);
      pending.push_back(P);
      return;
    }
    // the sparse passes take stores below kSparseMaxRows / 2 rows; a bound
    // that only passes in flight made loose is tightened first (a wait for
    // the device), rather than sending a large store to the chained kernels
    // (which would sort it)
    if (sparse_runs && !keep_fn && 2 * n_ub > kSparseMaxRows && 2 * n <= kSparseMaxRows) sync();
    // two sparse gates in a row whose generators commute share one pass: the
    // first waits here until the next gate shows whether it pairs with it
    if (pair_path()) {
      if (held.on && pairs_with(held.a, a)) {
        enqueue_pair(a, P);
        return;
      }
      flush_held();
      held.a = a;
      held.P = P;
      held.on = true;
      return;
    }
    flush_held();
    enqueue_branch(a, P, keep_mode);
  }

  // The recorded gate's tape on the chained kernels: every row's output and
  // branch slot and its coefficient at the gate's start.
  TapeDev tape_chained() {
    if (!recording || tape.empty()) return TapeDev{nullptr, nullptr, nullptr};
    TapeEntry& te = tape.back();
    if (te.n_in == 0 || te.t.out) return te.t;
    te.t.out = static_cast<uint32_t*>(tape_alloc(te.n_in * sizeof(uint32_t)));
    te.t.child = static_cast<uint32_t*>(tape_alloc(te.n_in * sizeof(uint32_t)));
    double* ci = static_cast<double*>(tape_alloc(te.n_in * sizeof(double)));
    HIPCHK(hipMemcpyAsync(ci, store[cur].dev().coeff, te.n_in * sizeof(double), hipMemcpyDeviceToDevice, st));
    te.c_in = ci;
    if (rec_sc) record_instances(te, rec_sc);
    return te.t;
  }
  const double* rec_sc = nullptr;  // the recorded batched gate's device sin / cos table

  // One branching gate: in place on the sparse path when its buffers fit,
  // else through the sorted chained kernels.
  void enqueue_branch(GateArgs<W>& a, Pending& P, bool keep_mode) {
    const int ring = P.ring;
    const uint32_t tag = P.tag;
    const uint64_t* n_dev = nslot_ptr(nslot);
    if (sparse_path()) {
      // after a refusal, stores at least that large go straight to the
      // chained kernels (no wait for the count on every gate)
      uint64_t n_sp = n_ub;
      if (n_ub < sparse_refused_at && ensure_sparse_capacity(n_sp)) {
        enqueue_sparse(a, P, ring, n_sp);
        return;
      }
      if (n_ub < sparse_refused_at) {  // refused now: no room for the lists and tables
        sparse_refused_at = std::max<uint64_t>(n, 1);
        release_sparse();
      }
      host_ctr.sparse_fallbacks++;
    }
    if (!store_sorted) {  // the chained kernels merge sorted runs
      if (recording && !tape.empty() && is_gate(tape.back().type)) {
        // the sort's permutation goes on the tape before this gate's entry
        const TapeEntry te = tape.back();
        tape.pop_back();
        sort_store(true);
        tape.push_back(te);
      } else {
        sort_store();
      }
      n_dev = nslot_ptr(nslot);
    }
    const TapeDev td = tape_chained();
    const bool generic = key_count<W>(a.gen) > kMaxClassBits;
    host_ctr.chained_gates++;
    if (generic) host_ctr.generic_sorts++;
    const int m = 1 << a.n_cls_bits;
    uint64_t n_in = n_ub;
    ensure_capacity(n_in, m);
    n_dev = nslot_ptr(nslot);
    StoreDev S = store[cur].dev();
    Store& out = store[1 - cur];
    const uint64_t nb = (n_in + kTile - 1) / kTile;
    Counters* ctr = ctr_dev.as<Counters>();
    const uint64_t* e_dev = offs.as<uint64_t>() + (size_t)m * nb;

    phase_begin(P.sampled, ring, 0);
    HIPCHK(L::branch_count(a, S, n_dev, nb, counts.as<uint32_t>(), m, ctr, st));
    HIPCHK(launch_exclusive_scan(counts.as<uint32_t>(), (uint64_t)m * nb, offs.as<uint64_t>(),
                                 scan_scratch.as<uint64_t>(), st));
    const bool rec = td.out != nullptr || keep_mode;  // keep mode needs the children's origins
    HIPCHK(L::branch_scatter(a, S, n_dev, nb, offs.as<uint64_t>(), m, child[0].dev(rec), td.child, st));
    phase_end(P.sampled, ring, 0);

    phase_begin(P.sampled, ring, 1);
    ChildDev C = child[0].dev(rec);
    if (generic) {  // too many generator bits to bucket: sort the children outright
      const uint64_t e = read_u64(e_dev);
      int in_b = 0;
      HIPCHK(launch_sort_children<W>(child[0].dev(rec), child[1].dev(rec), e, &in_b, st));
      C = child[in_b].dev(rec);
    } else if (m > 1) {
      int in_b = 0;
      HIPCHK(launch_merge_rounds<W>(child[0].dev(rec), child[1].dev(rec), offs.as<uint64_t>(), nb, m, n_in, &in_b,
                                    st));
      C = child[in_b].dev(rec);
    }
    phase_end(P.sampled, ring, 1);

    if (keep_mode) {  // the kernel decides every branch on the host
      const uint64_t e = read_u64(e_dev);
      keep_host = keep_rows(C.key, C.cap, C.val, e);  // outlives the copy: the next keep_rows syncs first
      for (uint8_t k : keep_host) keep_declined += k ? 0 : 1;
      const uint64_t e1 = std::max<uint64_t>(e, 1);
      keep_dev.ensure(e1 + std::max<uint64_t>(n, 1));
      uint8_t* kd = keep_dev.as<uint8_t>();
      uint8_t* own = kd + e1;
      if (e) HIPCHK(hipMemcpyAsync(kd, keep_host.data(), e, hipMemcpyHostToDevice, st));
      HIPCHK(launch_apply_keep(e, kd, C.flag, C.origin, own, st));
      a.own_keep = own;
    }
    uint64_t* ticket = small.as<uint64_t>() + kSlotTicket;
    HIPCHK(L::store_merge(a, S, n_dev, n_in, C, e_dev, n_in, out.dev(), splits.as<uint64_t>(),
                          status.as<uint64_t>(), reinterpret_cast<uint32_t*>(ticket),
                          reinterpret_cast<uint32_t*>(mirror_d + 2 * kRing), nslot_ptr(nslot + 1), mirror_d + ring,
                          mirror_d + kRing + ring, tag, ctr, td, st, P.sampled ? ph_ev[ring][4] : nullptr,
                          P.sampled ? ph_ev[ring][5] : nullptr));
    if (P.sampled) {
      ph_open[ring][2] = true;
      prof.launches[2]++;
    }
    P.merged = true;
    ++unresolved_merges;
    pending.push_back(P);
    nslot = (nslot + 1) % kSlots;
    cur = 1 - cur;
    n_ub = 2 * n_in;
  }

  // ---- sparse gate pairs
  struct Held {
    bool on = false;
    GateArgs<W> a;
    Pending P;
  } held;
  // Rows each pair pass's first gate made, one device word per pass since
  // the host last read them (read_mids, after the device is done).
  static constexpr uint64_t kPairMidSlots = 4096;
  DevMem pair_mid;
  uint64_t* mid_pinned = nullptr;  // host copy of pair_mid (pinned)
  uint64_t pair_mid_n = 0;
  struct MidFix {
    uint64_t slot, n_before, n_after;
    uint64_t* n_terms_out;  // the first gate's
    bool sampled;
  };
  std::vector<MidFix> mid_fix;

  bool pair_path() const {
    return pair_runs && !recording && sparse_path() && !has_min_terms && n_ub < sparse_refused_at;
  }

  // Gates that can share a pass: different generators that commute.
  static bool pairs_with(const GateArgs<W>& x, const GateArgs<W>& y) {
    return !key_eq<W>(x.gen, y.gen) && !anticommute<KIND, W>(x.gen, y.gen);
  }

  // Launch the held gate alone.
  void flush_held() {
    if (!held.on) return;
    held.on = false;
    GateArgs<W> a = held.a;
    Pending P = held.P;
    enqueue_branch(a, P, false);
  }

  // The held gate A and gate B (commuting) in one pass.
  void enqueue_pair(const GateArgs<W>& b, Pending& PB) {
    GateArgs<W> a = held.a;
    Pending PA = held.P;
    held.on = false;
    if (pair_mid_n == kPairMidSlots) sync_enqueued();  // every slot is taken: read them back first
    uint64_t n_in = n_ub;
    if (!ensure_sparse_capacity(n_in, 4)) {  // no room for the pass: one gate at a time
      enqueue_branch(a, PA, false);
      GateArgs<W> bb = b;
      enqueue_branch(bb, PB, false);
      return;
    }
    pair_mid.ensure(kPairMidSlots * kMidWords * sizeof(uint64_t));
    host_ctr.sparse_gates += 2;
    host_ctr.sparse_pairs++;
    const bool sm = PA.sampled;
    PA.sparse = PB.sparse = true;
    PA.pair_first = true;
    PA.mid_slot = pair_mid_n;
    PB.pair_second = true;
    PB.merged = true;
    PB.sampled = sm;
    PB.ev_ring = PA.ring;
    const int er = PA.ring;
    SparseEnt<W>* const ents[2] = {sparse_ents[0].as<SparseEnt<W>>(), sparse_ents[1].as<SparseEnt<W>>()};
    uint32_t* const slot_lists[2] = {sparse_slots[0].as<uint32_t>(), sparse_slots[1].as<uint32_t>()};
    uint64_t* const tables[2] = {sparse_table[0].as<uint64_t>(), sparse_table[1].as<uint64_t>()};
    HIPCHK(L::pair_gate(a, b, store[cur].dev(), nslot_ptr(nslot), n_in, ents, slot_lists, ent_cap, tables, table_cap,
                        sparse_sync.as<uint64_t>(), sparse_seq++, pair_mid.as<uint64_t>() + pair_mid_n * kMidWords,
                        nslot_ptr(nslot + 1), mirror_d + PB.ring, mirror_d + kRing + PB.ring, PB.tag,
                        ctr_dev.as<Counters>(), st, sm ? ph_ev[er][10] : nullptr, sm ? ph_ev[er][11] : nullptr,
                        sm ? ph_ev[er][14] : nullptr, sm ? ph_ev[er][15] : nullptr));
    ++pair_mid_n;
    if (sm) {
      for (int ph : {5, 7}) {
        ph_open[er][ph] = true;
        prof.launches[ph]++;
      }
    }
    unresolved_merges += 2;
    pending.push_back(PA);
    pending.push_back(PB);
    nslot = (nslot + 1) % kSlots;
    n_ub = 4 * n_in;
    store_sorted = false;
  }

  // The first gates' counts of the pair passes resolved so far: the device
  // has finished and their counters are in mid_pinned (sync_enqueued).
  void read_mids() {
    if (mid_fix.empty()) {
      pair_mid_n = 0;
      return;
    }
    std::vector<uint64_t> made(pair_mid_n, 0);
    for (uint64_t p = 0; p < pair_mid_n; ++p)
      for (int q = 0; q < kMidSlots; ++q) made[p] += mid_pinned[p * kMidWords + q * kMidStride];
    const double row = 8.0 * W + 8.0;
    for (const MidFix& f : mid_fix) {
      const uint64_t n_mid = f.n_before + made[f.slot];
      if (f.n_terms_out) *f.n_terms_out = n_mid;
      host_ctr.term_gate_updates += n_mid;  // the second gate's start
      if (f.sampled) {  // both gates' algorithmic bytes: one read and one write of the store each
        prof.bytes[6] += (double)(f.n_before + n_mid) * row + (double)(n_mid + f.n_after) * row;
        prof.launches[6] += 2;
      }
    }
    mid_fix.clear();
    pair_mid_n = 0;
  }

  // ---- sparse in-place gates (k_sparse_count / gather / join)
  DevMem sparse_ents[2], sparse_slots[2], sparse_table[2];  // alternate between gates
  DevMem sparse_sync;  // the sub-list records and the join word (kSparseSyncWords)
  uint64_t ent_cap = 0, table_cap = 0;
  uint64_t sparse_seq = 0;  // sparse gates enqueued: picks the list, table and list-length words

  bool sparse_path() const {
    // a recorded batched gate keeps its gradient partials per instance in
    // LDS, for at most kSparseRecInstMax instances (more: chained kernels)
    return sparse_runs && !keep_fn && (!recording || n_inst <= kSparseRecInstMax) && 2 * n_ub <= kSparseMaxRows;
  }

  // Device bytes the sparse gates' lists, slot lists and tables take for
  // gates over at most b rows.
  static uint64_t sparse_list_bytes(uint64_t b) {
    return 2 * (sparse_ent_need(b) * (sizeof(SparseEnt<W>) + sizeof(uint32_t)) +
                sparse_table_slots(b) * sizeof(uint64_t));
  }

  // Whether growing the store to 2b rows and the lists and tables to b
  // entries fits the sparse budget (pq_set_sparse_budget) and the device's
  // free memory, less a margin.
  // share: the part of the free memory the buffers may take (the passes'
  // slack for a pass in flight is only worth half of it: the store's own
  // compaction, sorts and readouts need room too).
  bool sparse_room(uint64_t b, uint64_t grow, double share = 1.0) {
    if (sparse_budget && sparse_list_bytes(b) > sparse_budget) return false;
    uint64_t need = 0;
    const uint64_t rows = grow * b;
    if (store[cur].cap < rows) need += rows * (W + 1) * sizeof(uint64_t);
    auto add = [&](const DevMem& m, uint64_t want) { need += want > m.bytes ? want - m.bytes : 0; };
    for (int i = 0; i < 2; ++i) {
      add(sparse_ents[i], sparse_ent_need(b) * sizeof(SparseEnt<W>));
      add(sparse_slots[i], sparse_ent_need(b) * sizeof(uint32_t));
      add(sparse_table[i], sparse_table_slots(b) * sizeof(uint64_t));
    }
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    return need + kSparseMargin <= (uint64_t)(share * (double)free_b);
  }
  static constexpr uint64_t kSparseMargin = 512ull << 20;

  void release_sparse() {
    for (int i = 0; i < 2; ++i) {
      sparse_ents[i].release();
      sparse_slots[i].release();
      sparse_table[i].release();
    }
    ent_cap = table_cap = 0;
  }

  // Buffers for a sparse gate (grow = 2) or pair pass (grow = 4) over at
  // most n_in rows: the store holds grow n_in rows (a gate at most doubles
  // them, a pair quadruples them), the list at most n_in entries, the table
  // a quarter-full table of them.  While one pass is in flight the host only
  // knows n_in <= grow n, so the buffers are sized for the pass after it as
  // well (b = grow n entries, a store of grow b rows) plus half again for
  // growth: a host that enqueues one pass behind the running one then never
  // has to wait for the device here, which would leave it idle while the
  // next pass is launched.  Near the memory limit the sizes drop to the two
  // passes, then to this pass alone.  Growing waits for the enqueued gates.
  // False when even that does not fit (the gate then runs chained).
  bool ensure_sparse_capacity(uint64_t& n_in, uint64_t grow = 2) {
    auto fits = [&](uint64_t b) {
      return store[cur].cap >= grow * b && ent_cap >= sparse_ent_need(b) && table_cap >= sparse_table_slots(b);
    };
    if (fits(n_in)) return true;
    host_ctr.sparse_syncs++;
    sync();
    n_in = n;
    if (fits(grow * n_in)) return true;  // room for the next pass in flight behind this one
    uint64_t b = std::max<uint64_t>(grow * n_in + grow * n_in / 2, 4096);
    if (!sparse_room(b, grow, 0.5)) {
      b = std::max<uint64_t>(grow * n_in, 4096);
      if (!sparse_room(b, grow, 0.5)) {
        if (fits(n_in)) return true;
        b = std::max<uint64_t>(n_in, 4096);
        if (!sparse_room(b, grow)) return false;
      }
    }
    host_ctr.sparse_grows++;
    grow_store(cur, grow * b, true, true);
    if (ent_cap < sparse_ent_need(b) || table_cap < sparse_table_slots(b)) {
      // new lists and tables together: both tables empty, and no earlier
      // gate's slots left to clear (the sub-list counters are zero)
      const uint64_t e = sparse_ent_need(b);
      for (int i = 0; i < 2; ++i) {
        sparse_ents[i].ensure((size_t)e * sizeof(SparseEnt<W>));
        sparse_slots[i].ensure((size_t)e * sizeof(uint32_t));
        sparse_table[i].ensure((size_t)sparse_table_slots(b) * sizeof(uint64_t));
        HIPCHK(hipMemsetAsync(sparse_table[i].p, 0, sparse_table[i].bytes, st));
      }
      ent_cap = std::min({sparse_ents[0].bytes / sizeof(SparseEnt<W>), sparse_ents[1].bytes / sizeof(SparseEnt<W>),
                          sparse_slots[0].bytes / sizeof(uint32_t), sparse_slots[1].bytes / sizeof(uint32_t)});
      table_cap = std::min(sparse_table[0].bytes, sparse_table[1].bytes) / sizeof(uint64_t);
      sparse_sync.ensure(kSparseSyncWords * sizeof(uint64_t));
      HIPCHK(hipMemsetAsync(sparse_sync.p, 0, kSparseSyncWords * sizeof(uint64_t), st));
    }
    return true;
  }

  // Enqueue one sparse gate (its arguments made by enqueue_gate).
  void enqueue_sparse(const GateArgs<W>& a, Pending& P, int ring, uint64_t n_in) {
    host_ctr.sparse_gates++;
    int nw = 0;
    for (int w = 0; w < W; ++w)
      nw += KIND == kPauli ? (a.fold.w[w] != 0) : ((a.gen_len & 1) || a.gen.w[w] != 0);
    P.sparse = true;
    P.nw = nw;
    const bool sm = P.sampled;
    SparseRec* rec = nullptr;
    uint32_t* rinst = nullptr;
    if (recording && !tape.empty() && tape.back().n_in) {  // a record per listed row, at most one per row
      TapeEntry& te = tape.back();
      te.type = kTapeSparse;
      if (rec_sc) {  // batched: the gate's per-instance sin / cos and each record's instance
        double* sc = static_cast<double*>(tape_alloc((size_t)n_inst * 3 * sizeof(double)));
        HIPCHK(hipMemcpyAsync(sc, rec_sc, (size_t)n_inst * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
        te.sc = sc;
        te.rinst = rinst = static_cast<uint32_t*>(tape_alloc(te.n_in * sizeof(uint32_t)));
      }
      // last, so that it can be trimmed to the list's length
      te.srec = rec = static_cast<SparseRec*>(tape_alloc(te.n_in * sizeof(SparseRec)));
    }
    SparseEnt<W>* const ents[2] = {sparse_ents[0].as<SparseEnt<W>>(), sparse_ents[1].as<SparseEnt<W>>()};
    uint32_t* const slot_lists[2] = {sparse_slots[0].as<uint32_t>(), sparse_slots[1].as<uint32_t>()};
    uint64_t* const tables[2] = {sparse_table[0].as<uint64_t>(), sparse_table[1].as<uint64_t>()};
    HIPCHK(L::sparse_gate(a, store[cur].dev(), nslot_ptr(nslot), n_in, ents, slot_lists, ent_cap, tables, table_cap,
                          sparse_sync.as<uint64_t>(), sparse_seq++, nslot_ptr(nslot + 1), mirror_d + ring,
                          mirror_d + kRing + ring, P.tag, ctr_dev.as<Counters>(), rec, rinst, st,
                          sm ? ph_ev[ring][10] : nullptr,
                          sm ? ph_ev[ring][11] : nullptr, sm ? ph_ev[ring][14] : nullptr,
                          sm ? ph_ev[ring][15] : nullptr));
    if (sm) {
      for (int ph : {5, 7}) {
        ph_open[ring][ph] = true;
        prof.launches[ph]++;
      }
    }
    P.merged = true;
    ++unresolved_merges;
    pending.push_back(P);
    nslot = (nslot + 1) % kSlots;
    n_ub = 2 * n_in;
    store_sorted = false;
  }

  // Sort the live rows by key (the store's order on the sorted paths): a copy
  // into the child buffers, the generic child sort, and a copy back.
  // record: also push the permutation onto the tape (a noise-style entry
  // mapping every row to its sorted place, factor 1); every sort while
  // recording is recorded, since the tape's later row indices follow it.
  void sort_store(bool record = false) {
    sync();
    if (store_sorted) return;
    record = record || recording;
    if (record && n > kRecMaxRows) fail(PQ_ERR_STATE, "too many rows to record a sort (2^31 limit)");
    if (n > 1) {
      child[0].ensure(n);
      child[1].ensure(n);
      const ChildDev c0 = child[0].dev(record), c1 = child[1].dev(record);
      if (record) HIPCHK(launch_iota_u32(c0.origin, n, st));
      const StoreDev sd = store[cur].dev();
      for (int w = 0; w < W; ++w)
        HIPCHK(hipMemcpyAsync(c0.key + (size_t)w * c0.cap, sd.key + (size_t)w * sd.cap, n * sizeof(uint64_t),
                              hipMemcpyDeviceToDevice, st));
      HIPCHK(hipMemcpyAsync(c0.val, sd.coeff, n * sizeof(double), hipMemcpyDeviceToDevice, st));
      int in_b = 0;
      HIPCHK(launch_sort_children<W>(c0, c1, n, &in_b, st));
      const ChildDev r = in_b ? c1 : c0;
      if (record) {
        uint32_t* out = static_cast<uint32_t*>(tape_alloc(n * sizeof(uint32_t)));
        HIPCHK(launch_perm_out(r.origin, out, n, st));
        tape.push_back(TapeEntry{kTapeNoise, n, n, 0.0, 1.0, TapeDev{out, nullptr, nullptr}, nullptr});
      }
      for (int w = 0; w < W; ++w)
        HIPCHK(hipMemcpyAsync(sd.key + (size_t)w * sd.cap, r.key + (size_t)w * r.cap, n * sizeof(uint64_t),
                              hipMemcpyDeviceToDevice, st));
      HIPCHK(hipMemcpyAsync(sd.coeff, r.val, n * sizeof(double), hipMemcpyDeviceToDevice, st));
      HIPCHK(hipStreamSynchronize(st));
    }
    store_sorted = true;
  }

  void noise(const double* table, uint64_t len, const double* factors, const uint8_t* flags,
             int reclaim) override {
    sync();
    if (n == 0) return;
    NoiseArgs a{};
    if (table) {
      if (len == 0) fail(PQ_ERR_ARG, "empty noise table");
      // a run damps every layer by the same table: upload it only when it changed
      if (table_host.size() != len || std::memcmp(table_host.data(), table, len * sizeof(double)) != 0) {
        table_dev.ensure(len * sizeof(double));
        HIPCHK(hipMemcpy(table_dev.p, table, len * sizeof(double), hipMemcpyHostToDevice));
        table_host.assign(table, table + len);
      }
      a.table = table_dev.as<double>();
      a.len = (uint32_t)len;
    }
    if (factors || flags) {
      row_aux.ensure((size_t)n * sizeof(double));
      if (factors) {
        HIPCHK(hipMemcpy(row_aux.p, factors, n * sizeof(double), hipMemcpyHostToDevice));
        a.factors = row_aux.as<double>();
      } else {
        HIPCHK(hipMemcpy(row_aux.p, flags, n, hipMemcpyHostToDevice));
        a.keep_flags = row_aux.as<uint8_t>();
      }
    }
    a.unit_mask = unit_mask;
    a.inst_shift = inst_shift;
    a.reclaim = reclaim && (has_maxw || has_min) && !keep_fn;
    if (keep_fn && reclaim && !flags) {  // reclaim_by_kernel (termsum.rs:387-447): scale, then the kernel's keep
      if (a.table || a.factors) compact(a);
      kernel_reclaim();
      return;
    }
    a.has_min = has_min;
    a.min_coeff = min_coeff;
    a.has_maxw = has_maxw;
    a.max_weight = maxw;
    if (!a.table && !a.factors && !a.keep_flags && !a.reclaim) return;
    compact(a);
  }

  // Reclaim with the caller's keep kernel over every live row.
  void kernel_reclaim() {
    sync();
    if (n == 0) return;
    const StoreDev sd = store[cur].dev();
    const std::vector<uint8_t> keep = keep_rows(sd.key, sd.cap, sd.coeff, n);
    row_aux.ensure((size_t)n);
    HIPCHK(hipMemcpy(row_aux.p, keep.data(), n, hipMemcpyHostToDevice));
    NoiseArgs a{};
    a.keep_flags = row_aux.as<uint8_t>();
    a.unit_mask = unit_mask;
    a.inst_shift = inst_shift;
    compact(a);
  }

  // One order-preserving pass of k_noise_compact over the live rows.
  void compact(const NoiseArgs& a) {
    TapeDev td{nullptr, nullptr, nullptr};
    if (recording) {
      if (n > kRecMaxRows) fail(PQ_ERR_STATE, "too many rows to record a noise pass (2^31 limit)");
      td.out = static_cast<uint32_t*>(tape_alloc(n * sizeof(uint32_t)));
      td.factor = static_cast<double*>(tape_alloc(n * sizeof(double)));
      tape.push_back(TapeEntry{kTapeNoise, n, n, 0.0, 1.0, td, nullptr});
    }
    // in place: every tile has loaded its rows before it publishes its count,
    // and a tile writes only after every earlier tile has published, below
    // its own first row, so no row is overwritten before it is read
    const Store& out = store[cur];
    const uint64_t n_tiles = (n + kTile - 1) / kTile;
    status.ensure((size_t)(n_tiles + 1) * sizeof(uint64_t));  // look-back words, then the tile ticket
    // Unrecorded passes run like gates: the count is published to a mirror
    // slot and read when the host next needs it, so the gates after the pass
    // are enqueued behind it without a wait (a recorded pass needs its count
    // for the tape at once).
    const bool async = !recording && n_tiles > 0;
    const int ring = ring_of(seq);
    const uint32_t tag = (uint32_t)(seq % 0xFFFF) + 1;
    ++seq;
    phase_begin(profiling, ring, 3);
    HIPCHK(L::noise_compact(a, store[cur].dev(), n, out.dev(), status.as<uint64_t>(), nslot_ptr(nslot + 1), n_units,
                            td, st, async ? mirror_d + ring : nullptr, tag));
    phase_end(profiling, ring, 3);
    if (async) {
      Pending P{true, true, profiling, ring, tag, nullptr};
      P.noise = true;
      pending.push_back(P);
      nslot = (nslot + 1) % kSlots;
      n_ub = n;  // n stays the input count (a bound) until the pass is resolved
      return;
    }
    const uint64_t n_in = n;
    n = read_u64(nslot_ptr(nslot + 1));
    if (profiling) {
      prof.bytes[3] += (double)(n_in + n) * (8.0 * W + 8.0);
      phase_collect(ring);
    }
    nslot = (nslot + 1) % kSlots;
    n_ub = n;
    if (td.out) tape.back().n_out = n;
  }

  void upload_frame() {
    if (!frame_dirty) return;
    frame_dirty = false;
    if (frame.identity) return;
    const int R = frame.n_rows;
    std::vector<uint64_t> img((size_t)R * W);
    for (int p = 0; p < R; ++p)
      for (int i = 0; i < W; ++i) img[(size_t)p * W + i] = frame.ro[p].w[i];
    frame_img.ensure(img.size() * sizeof(uint64_t) + 8);
    frame_sgn.ensure((size_t)R * sizeof(double) + 8);
    HIPCHK(hipMemcpy(frame_img.p, img.data(), img.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(frame_sgn.p, frame.ro_s.data(), (size_t)R * sizeof(double), hipMemcpyHostToDevice));
  }

  // Basis state (one bit per unit) -> mask of the key bits its trace reads.
  Key<W> diag_mask(const uint64_t* diag, uint64_t nd) const {
    Key<W> dm{};
    for (int u = 0; u < n_units && u < 32 * W; ++u) {
      const uint64_t bit = (u / 64) < (int)nd ? (diag[u / 64] >> (u % 64)) & 1 : 0;
      if (!bit) continue;
      const int pos = KIND == kPauli ? 2 * u + 1 : 2 * u;
      dm.w[pos / 64] |= 1ULL << (pos % 64);
    }
    return dm;
  }

  double expectation(const uint64_t* diag, uint64_t nd) override {
    sort_store();  // the sum runs in store order: the same bits on every path
    if (n == 0) return 0.0;
    const Key<W> dm = diag_mask(diag, nd);
    upload_frame();
    const bool fr = !frame.identity;
    int nb = (int)std::min<uint64_t>((n + kBlock - 1) / kBlock, 2048);
    partials.ensure((size_t)nb * sizeof(double));
    HIPCHK(L::expectation(store[cur].dev(), n, dm, n_units, fr ? frame_img.as<uint64_t>() : nullptr,
                          fr ? frame_sgn.as<double>() : nullptr, frame.n_rows, partials.as<double>(), nb, st,
                          unit_mask));
    std::vector<double> host(nb);
    HIPCHK(hipMemcpyAsync(host.data(), partials.p, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double acc = 0.0;
    for (double v : host) acc += v;
    return acc;
  }

  // expectation() without allocating: partials, frame image and the total
  // live in the device's scratch array, and the sum is read back through the
  // pinned word allocated at creation.  Calls that overlap on one device
  // share the array and race.
  double expectation_noalloc(const uint64_t* diag, uint64_t nd) override {
    sync();
    if (n == 0) return 0.0;
    const Key<W> dm = diag_mask(diag, nd);
    const bool fr = !frame.identity;
    const int nb = (int)std::min<uint64_t>((n + kBlock - 1) / kBlock, 2048);
    const size_t R = fr ? (size_t)frame.n_rows : 0;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t off_part = 256, off_img = off_part + up((size_t)nb * sizeof(double));
    const size_t off_sgn = off_img + up(R * W * sizeof(uint64_t)), need = off_sgn + up(R * sizeof(double));
    if (need > kScratchBytes)
      fail(PQ_ERR_ARG, "expectation needs " + std::to_string(need) + " bytes of scratch, more than the " +
                           std::to_string(kScratchBytes) + " the allocation-free entry point holds");
    char* base = static_cast<char*>(scratch_dev);
    double* total = reinterpret_cast<double*>(base);
    double* part = reinterpret_cast<double*>(base + off_part);
    uint64_t* img = reinterpret_cast<uint64_t*>(base + off_img);
    double* sgn = reinterpret_cast<double*>(base + off_sgn);
    if (fr) {  // the tableau's rows are W contiguous words each, as the kernels read them
      HIPCHK(hipMemcpyAsync(img, frame.ro.data(), R * W * sizeof(uint64_t), hipMemcpyHostToDevice, st));
      HIPCHK(hipMemcpyAsync(sgn, frame.ro_s.data(), R * sizeof(double), hipMemcpyHostToDevice, st));
    }
    HIPCHK(L::expectation(store[cur].dev(), n, dm, n_units, fr ? img : nullptr, fr ? sgn : nullptr, frame.n_rows,
                          part, nb, st, unit_mask));
    hipLaunchKernelGGL(k_sum_in_order, dim3(1), dim3(64), 0, st, part, nb, total);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(pinned, total, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double v;
    std::memcpy(&v, pinned, sizeof(double));
    return v;
  }

  uint64_t terms_below(double thr) override {
    sync();
    unsigned long long* out = reinterpret_cast<unsigned long long*>(small.as<uint64_t>() + kSlotCount);
    HIPCHK(launch_count_below(store[cur].dev().coeff, n, thr, out, st));
    return read_u64(reinterpret_cast<uint64_t*>(out));
  }

  uint64_t export_rows(uint64_t* keys, double* coeffs, uint64_t cap, bool with_frame) override {
    sort_store();
    if (cap < n) fail(PQ_ERR_ARG, "export buffer too small");
    if (n == 0) return 0;
    upload_frame();
    const bool fr = with_frame && !frame.identity;
    export_keys.ensure((size_t)n * W * sizeof(uint64_t));
    export_coeff.ensure((size_t)n * sizeof(double));
    HIPCHK(L::conjugate(store[cur].dev(), n, fr ? frame_img.as<uint64_t>() : nullptr,
                        fr ? frame_sgn.as<double>() : nullptr, frame.n_rows, export_keys.as<uint64_t>(),
                        export_coeff.as<double>(), st));
    HIPCHK(hipMemcpyAsync(keys, export_keys.p, (size_t)n * W * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(coeffs, export_coeff.p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return n;
  }

  void weights(uint32_t* out, uint64_t cap) override {
    sync();
    if (cap < n) fail(PQ_ERR_ARG, "weights buffer too small");
    if (n == 0) return;
    row_aux.ensure((size_t)n * sizeof(uint32_t));
    HIPCHK(L::weights(store[cur].dev(), n, n_units, row_aux.as<uint32_t>(), st, unit_mask));
    HIPCHK(hipMemcpyAsync(out, row_aux.p, (size_t)n * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }

  bool frame_identity() const override { return frame.identity; }

  void counters(pq_counters* out) override {
    sync();
    std::vector<Counters> slots(kCtrSlots);
    HIPCHK(hipMemcpyAsync(slots.data(), ctr_dev.p, kCtrSlots * sizeof(Counters), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    Counters c{};
    double mxs = 0.0;
    for (const Counters& x : slots) {
      c.visited += x.visited;
      c.emitted += x.emitted;
      c.declined += x.declined;
      c.hits += x.hits;
      c.created += x.created;
      c.dropped_weight += x.dropped_weight;
      c.disc_l1 += x.disc_l1;
      double m;
      std::memcpy(&m, &x.disc_max_bits, sizeof(double));
      mxs = std::max(mxs, m);
    }
    std::memcpy(&c.disc_max_bits, &mxs, sizeof(double));
    *out = host_ctr;
    out->visited = c.visited;
    out->emitted = c.emitted;
    out->declined = c.declined + keep_declined;
    out->hits = c.hits;
    out->created = c.created;
    out->dropped_weight = c.dropped_weight;
    out->discarded_coeff_l1 = c.disc_l1;
    double mx;
    std::memcpy(&mx, &c.disc_max_bits, sizeof(double));
    out->discarded_coeff_max = mx;
  }

  void counters_reset() override {
    sync();
    HIPCHK(hipMemsetAsync(ctr_dev.p, 0, kCtrSlots * sizeof(Counters), st));
    host_ctr = pq_counters{};
    keep_declined = 0;
  }

  uint64_t key_bytes() const override { return n * (uint64_t)W * sizeof(uint64_t); }

  // ---------------------------------------------------- differentiation

  void tape_begin() override {
    sync();
    if (!differentiable) fail(PQ_ERR_STATE, "recording needs differentiable mode (pq_set_differentiable)");
    if (keep_fn) fail(PQ_ERR_STATE, "recording is not supported with a keep kernel");
    sort_store();  // a tape starts on the sorted store, the order its replays start from
    tape_reset();
    recording = true;
  }

  // The tape ends on the sorted store (a recorded permutation when sparse
  // gates left it unsorted): a replayed segment then ends in the order the
  // next segment's tape starts from (tape_begin sorts), whatever order the
  // sparse gates' appends took.
  void tape_end() override {
    sync();
    if (recording && !store_sorted) sort_store(true);
    recording = false;
  }

  void tape_info(uint64_t* n_gates, uint64_t* bytes) const override {
    uint64_t g = 0;
    for (const auto& t : tape) g += is_gate(t.type);
    if (n_gates) *n_gates = g;
    if (bytes) *bytes = tape_bytes_used;
  }

  uint32_t checkpoint_save() override {
    sync();
    std::unique_ptr<Checkpoint> cp(new Checkpoint);
    cp->n = n;
    cp->frame = frame;
    cp->sorted = store_sorted;
    cp->mem.ensure(std::max<uint64_t>(n, 1) * (W + 1) * sizeof(uint64_t));
    const StoreDev a = store[cur].dev();
    uint64_t* dst = static_cast<uint64_t*>(cp->mem.p);
    for (int w = 0; w < W; ++w)
      HIPCHK(hipMemcpyAsync(dst + (size_t)w * n, a.key + (size_t)w * a.cap, n * sizeof(uint64_t),
                            hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(dst + (size_t)W * n, a.coeff, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    const uint32_t id = next_ckpt++;
    ckpts[id] = std::move(cp);
    return id;
  }

  void checkpoint_restore(uint32_t id) override {
    auto it = ckpts.find(id);
    if (it == ckpts.end()) fail(PQ_ERR_ARG, "unknown checkpoint id");
    sync();
    if (recording) fail(PQ_ERR_STATE, "cannot restore a checkpoint while recording");
    const Checkpoint& cp = *it->second;
    grow_store(cur, std::max<uint64_t>(cp.n, 1), false);
    const StoreDev b = store[cur].dev();
    const uint64_t* src = static_cast<const uint64_t*>(cp.mem.p);
    for (int w = 0; w < W; ++w)
      HIPCHK(hipMemcpyAsync(b.key + (size_t)w * b.cap, src + (size_t)w * cp.n, cp.n * sizeof(uint64_t),
                            hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(b.coeff, src + (size_t)W * cp.n, cp.n * sizeof(double), hipMemcpyDeviceToDevice, st));
    n = n_ub = cp.n;
    pinned[0] = n;
    HIPCHK(hipMemcpyAsync(nslot_ptr(nslot), pinned, sizeof(uint64_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    frame = cp.frame;
    frame_dirty = true;
    store_sorted = cp.sorted;
  }

  void checkpoint_free(uint32_t id) override {
    if (!ckpts.erase(id)) fail(PQ_ERR_ARG, "unknown checkpoint id");
  }

  void adjoint_seed(const uint64_t* diag, uint64_t nd) override {
    sync();
    lam[0].ensure(std::max<uint64_t>(n, 1) * sizeof(double));
    lam_cur = 0;
    lam_n = n;
    lam_valid = true;
    if (n == 0) return;
    const Key<W> dm = diag_mask(diag, nd);
    upload_frame();
    const bool fr = !frame.identity;
    HIPCHK(L::adjoint_seed(store[cur].dev(), n, dm, n_units, fr ? frame_img.as<uint64_t>() : nullptr,
                           fr ? frame_sgn.as<double>() : nullptr, frame.n_rows, lam[0].as<double>(), st));
    HIPCHK(hipStreamSynchronize(st));
  }

  // Readout cotangents of a batched store: instance b reads against its own
  // basis state (diag + b nd).
  void adjoint_seed_batched(const uint64_t* diag, uint64_t nd) override {
    need_batched();
    sync();
    lam[0].ensure(std::max<uint64_t>(n, 1) * sizeof(double));
    lam_cur = 0;
    lam_n = n;
    lam_valid = true;
    if (n == 0) return;
    std::vector<uint64_t> dms((size_t)n_inst * W);
    for (int b = 0; b < n_inst; ++b) {
      const Key<W> dm = diag_mask(diag ? diag + (size_t)b * nd : nullptr, diag ? nd : 0);
      for (int i = 0; i < W; ++i) dms[(size_t)b * W + i] = dm.w[i];
    }
    inst_dm.ensure(dms.size() * sizeof(uint64_t));
    HIPCHK(hipMemcpy(inst_dm.p, dms.data(), dms.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    upload_frame();
    const bool fr = !frame.identity;
    HIPCHK(L::adjoint_seed_batched(store[cur].dev(), n, inst_dm.as<uint64_t>(), inst_shift, unit_mask, n_units,
                                   fr ? frame_img.as<uint64_t>() : nullptr, fr ? frame_sgn.as<double>() : nullptr,
                                   frame.n_rows, lam[0].as<double>(), st));
    HIPCHK(hipStreamSynchronize(st));
  }

  // Sweep the tape backwards: d<O>/dtheta of every recorded gate (in the
  // order the gates were applied) and the cotangent of the rows the tape
  // started from.  The tape is consumed.
  void adjoint_backward(double* grads, uint64_t n_gates) override { backward_impl(grads, n_gates, false); }

  void adjoint_backward_batched(double* grads, uint64_t n_gates) override {
    need_batched();
    backward_impl(grads, n_gates, true);
  }

  // grads: one value per recorded gate, or (batched) n_inst per gate, gate-major.
  void backward_impl(double* grads, uint64_t n_gates, bool batched) {
    sync();
    if (recording) fail(PQ_ERR_STATE, "end the tape before the backward sweep");
    if (!lam_valid) fail(PQ_ERR_STATE, "no adjoint: call pq_adjoint_seed first");
    uint64_t gates = 0;
    tape_info(&gates, nullptr);
    if (gates != n_gates) fail(PQ_ERR_ARG, "gradient buffer length differs from the recorded gate count");
    if (tape.empty()) return;
    if (tape.back().n_out != lam_n) fail(PQ_ERR_STATE, "the tape does not end at the adjoint's rows");
    // poff: each gate's block partials; grow: its row of the gradient
    std::vector<uint64_t> poff(tape.size() + 1, 0), grow(tape.size(), 0);
    uint64_t n_rows = 0;
    for (size_t i = 0; i < tape.size(); ++i) {
      const TapeEntry& te = tape[i];
      if (is_gate(te.type)) grow[i] = n_rows++;
      uint64_t np = 0;
      if (te.type == kTapeGate && te.n_in) {
        if (batched != (te.work != nullptr))
          fail(PQ_ERR_STATE, batched ? "the tape holds unbatched gates" : "the tape holds batched gates");
        np = batched ? (uint64_t)work_tot_h[te.tot_slot] : (uint64_t)adjoint_blocks(te.n_in);
      } else if (te.type == kTapeSparse && te.n_list) {
        if (batched != (te.rinst != nullptr))
          fail(PQ_ERR_STATE, batched ? "the tape holds unbatched gates" : "the tape holds batched gates");
        np = batched ? 0 : (uint64_t)adjoint_blocks(te.n_list);  // batched: summed per gate
      }
      poff[i + 1] = poff[i] + np;
    }
    grad_partials.ensure(std::max<uint64_t>(poff.back(), 1) * sizeof(double));
    if (batched && n_rows) {  // sparse gates sum into their rows during the sweep (none listed: zero)
      grads_dev.ensure(n_rows * (uint64_t)n_inst * sizeof(double));
      HIPCHK(hipMemsetAsync(grads_dev.p, 0, n_rows * (uint64_t)n_inst * sizeof(double), st));
    }
    for (size_t i = tape.size(); i-- > 0;) {
      const TapeEntry& te = tape[i];
      if (te.type == kTapeSparse) {  // in place: only the listed rows change
        if (te.n_list) {
          lam_tmp.ensure(te.n_list * sizeof(double));
          if (batched) {
            inst_part.ensure((uint64_t)sparse_inst_blocks(te.n_list, n_inst) * n_inst * sizeof(double));
            HIPCHK(launch_adjoint_sparse_batched(te.n_list, te.srec, te.rinst, te.sc, n_inst,
                                                 lam[lam_cur].as<double>(), lam_tmp.as<double>(),
                                                 inst_part.as<double>(),
                                                 grads_dev.as<double>() + grow[i] * (uint64_t)n_inst, st));
          } else {
            HIPCHK(launch_adjoint_sparse(te.n_list, te.srec, lam[lam_cur].as<double>(), lam_tmp.as<double>(), te.s,
                                         te.c, grad_partials.as<double>() + poff[i], adjoint_blocks(te.n_list), st));
          }
        }
        lam_n = te.n_in;
        continue;
      }
      DevMem& dst = lam[1 - lam_cur];
      dst.ensure(std::max<uint64_t>(te.n_in, 1) * sizeof(double));
      const double* lo = lam[lam_cur].as<double>();
      if (te.type == kTapeGate) {
        if (te.n_in && te.work)
          HIPCHK(launch_adjoint_gate_segments(poff[i + 1] - poff[i], te.work, te.bounds, te.sc, te.t, te.c_in, lo,
                                              dst.as<double>(), grad_partials.as<double>() + poff[i], st));
        else if (te.n_in)
          HIPCHK(launch_adjoint_gate(te.n_in, te.t, te.c_in, lo, dst.as<double>(), te.s, te.c,
                                     grad_partials.as<double>() + poff[i], adjoint_blocks(te.n_in), st));
      } else {
        HIPCHK(launch_adjoint_noise(te.n_in, te.t, lo, dst.as<double>(), st));
      }
      lam_cur = 1 - lam_cur;
      lam_n = te.n_in;
    }
    if (batched) {
      // per gate and instance, its work items' partials in block order, as
      // the unbatched sweep sums its blocks (on the device: the work lists
      // never come to the host)
      // (the chained gates; the sparse gates' rows are already summed)
      std::vector<uint64_t> part_off, rows;
      std::vector<const uint32_t*> offs;
      for (size_t i = 0; i < tape.size(); ++i) {
        if (tape[i].type != kTapeGate) continue;
        part_off.push_back(poff[i]);
        rows.push_back(grow[i]);
        offs.push_back(tape[i].work_off ? tape[i].work_off : empty_offsets());
      }
      const uint64_t ng = part_off.size();
      if (ng) {
        grad_meta.ensure(ng * (2 * sizeof(uint64_t) + sizeof(void*)));
        uint64_t* po = grad_meta.as<uint64_t>();
        uint64_t* ro = po + ng;
        const uint32_t** of = reinterpret_cast<const uint32_t**>(ro + ng);
        HIPCHK(hipMemcpyAsync(po, part_off.data(), ng * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(ro, rows.data(), ng * sizeof(uint64_t), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(of, offs.data(), ng * sizeof(void*), hipMemcpyHostToDevice, st));
        HIPCHK(launch_instance_grads(po, ro, of, ng, grad_partials.as<double>(), n_inst, grads_dev.as<double>(), st));
      }
      if (n_rows)
        HIPCHK(hipMemcpyAsync(grads, grads_dev.p, n_rows * (uint64_t)n_inst * sizeof(double), hipMemcpyDeviceToHost,
                              st));
      HIPCHK(hipStreamSynchronize(st));
      tape_reset();
      return;
    }
    std::vector<double> part(poff.back());
    if (!part.empty())
      HIPCHK(hipMemcpyAsync(part.data(), grad_partials.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    uint64_t gi = 0;
    for (size_t i = 0; i < tape.size(); ++i) {
      if (!is_gate(tape[i].type)) continue;
      double acc = 0.0;
      for (uint64_t b = poff[i]; b < poff[i + 1]; ++b) acc += part[b];
      grads[gi++] = acc;
    }
    tape_reset();
  }

  uint64_t adjoint_values(double* out, uint64_t cap) override {
    sync();
    if (!lam_valid) fail(PQ_ERR_STATE, "no adjoint: call pq_adjoint_seed first");
    if (!out) return lam_n;  // size query
    if (cap < lam_n) fail(PQ_ERR_ARG, "adjoint buffer too small");
    if (lam_n)
      HIPCHK(hipMemcpy(out, lam[lam_cur].p, lam_n * sizeof(double), hipMemcpyDeviceToHost));
    return lam_n;
  }

  // -------------------------------------------------------------- batching

  void need_batched() const {
    if (!n_inst) fail(PQ_ERR_STATE, "not a batched engine (pq_engine_create_batched)");
  }

  void seed_batched(const uint32_t* inst, const uint64_t* keys, const double* coeffs, uint64_t cnt) override {
    need_batched();
    if (has_min_terms) fail(PQ_ERR_ARG, "a term budget (min_terms) is per run; it is not supported on a batched engine");
    std::vector<uint64_t> k((size_t)cnt * W);
    for (uint64_t i = 0; i < cnt; ++i) {
      if (inst[i] >= (uint32_t)n_inst) fail(PQ_ERR_ARG, "instance id out of range");
      for (int w = 0; w < W; ++w) k[i * W + w] = keys[i * W + w];
      if (k[i * W + W - 1] & ~unit_mask) fail(PQ_ERR_ARG, "key has bits beyond n_units");
      k[i * W + W - 1] |= (uint64_t)inst[i] << inst_shift;
    }
    seed(k.data(), coeffs, cnt);
    retired.assign(n_inst, 0);
    frozen.assign(n_inst, 0);
  }

  void apply_gates_batched(const uint64_t* gens, const double* angles, uint64_t n_gates,
                           uint64_t* n_terms) override {
    need_batched();
    if (!n_gates) return;
    std::vector<double> tab((size_t)n_gates * n_inst * 3);
    for (uint64_t g = 0; g < n_gates; ++g)
      for (int b = 0; b < n_inst; ++b) {
        const double a = angles[g * n_inst + b], sn = std::sin(a);
        double* t = &tab[((size_t)g * n_inst + b) * 3];
        t[0] = sn;
        t[1] = std::cos(a);
        t[2] = (!differentiable && std::fabs(sn) < kCliffordEps) ? 1.0 : 0.0;  // phase-only
        if (frozen[b]) {  // rows pass unchanged: no branch, factor 1
          t[0] = 0.0;
          t[1] = 1.0;
          t[2] = -1.0;
        }
      }
    sync();  // the table buffer is reused by every call
    sc_dev.ensure(tab.size() * sizeof(double));
    HIPCHK(hipMemcpy(sc_dev.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    for (uint64_t g = 0; g < n_gates; ++g) {
      enqueue_gate(gens + g * W, 0.0, n_terms ? n_terms + g : nullptr, sc_dev.as<double>() + g * n_inst * 3,
                   angles + g * n_inst);
      const size_t lag = recording ? 0 : 1;
      while (pending.size() - pending_head > lag) resolve_front();
      close_recorded_gate();
      if (pending_head > 64) {
        pending.erase(pending.begin(), pending.begin() + pending_head);
        pending_head = 0;
      }
    }
    sync();
  }

  // Instance row ranges: bounds[b] .. bounds[b + 1].
  std::vector<uint64_t> instance_bounds() {
    sync();
    sort_store();  // each instance's rows are one run of the sorted store
    inst_bounds.ensure((size_t)(n_inst + 1) * sizeof(uint64_t));
    std::vector<uint64_t> h(n_inst + 1, 0);
    if (n) {
      const StoreDev sd = store[cur].dev();
      HIPCHK(launch_instance_bounds(sd.key + (size_t)(W - 1) * sd.cap, nslot_ptr(nslot), inst_shift, n_inst,
                                    inst_bounds.as<uint64_t>(), st));
      HIPCHK(hipMemcpyAsync(h.data(), inst_bounds.p, h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
    }
    return h;
  }

  void instance_counts(uint64_t* out) override {
    need_batched();
    const std::vector<uint64_t> bd = instance_bounds();
    for (int b = 0; b < n_inst; ++b) out[b] = bd[b + 1] - bd[b];
  }

  // Per instance, the partition and summation order of expectation() for a
  // run of that instance alone, so the values agree bit for bit.
  void expectation_batched(const uint64_t* diag, uint64_t nd, double* out) override {
    need_batched();
    const std::vector<uint64_t> bd = instance_bounds();
    segments_expectation(store[cur].dev(), bd, diag, nd, out);
  }

  // Per instance, the expectation value over its rows [bd[b], bd[b + 1]) of
  // the sorted rows sd (device bounds in inst_bounds).
  void segments_expectation(const StoreDev& sd, const std::vector<uint64_t>& bd, const uint64_t* diag, uint64_t nd,
                            double* out) {
    std::vector<uint32_t> work;
    std::vector<uint64_t> dms((size_t)n_inst * W);
    for (int b = 0; b < n_inst; ++b) {
      const Key<W> dm = diag_mask(diag ? diag + (size_t)b * nd : nullptr, diag ? nd : 0);
      for (int i = 0; i < W; ++i) dms[(size_t)b * W + i] = dm.w[i];
      const uint64_t nb_rows = bd[b + 1] - bd[b];
      const uint32_t nblk = (uint32_t)std::min<uint64_t>((nb_rows + kBlock - 1) / kBlock, 2048);
      for (uint32_t j = 0; j < nblk; ++j) {
        work.push_back((uint32_t)b);
        work.push_back(j);
        work.push_back(nblk);
      }
      out[b] = 0.0;
    }
    const uint64_t n_work = work.size() / 3;
    if (!n_work) return;
    upload_frame();
    const bool fr = !frame.identity;
    inst_work.ensure(work.size() * sizeof(uint32_t));
    inst_dm.ensure(dms.size() * sizeof(uint64_t));
    partials.ensure(n_work * sizeof(double));
    HIPCHK(hipMemcpy(inst_work.p, work.data(), work.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(inst_dm.p, dms.data(), dms.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
    HIPCHK(L::expectation_segments(sd, inst_work.as<uint32_t>(), n_work, inst_bounds.as<uint64_t>(),
                                   inst_dm.as<uint64_t>(), n_units, fr ? frame_img.as<uint64_t>() : nullptr,
                                   fr ? frame_sgn.as<double>() : nullptr, frame.n_rows, partials.as<double>(),
                                   unit_mask, st));
    std::vector<double> part(n_work);
    HIPCHK(hipMemcpyAsync(part.data(), partials.p, n_work * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (uint64_t i = 0; i < n_work; ++i) out[work[3 * i]] += part[i];
  }

  // Later gates leave the rows of the flagged instances unchanged (allowed
  // while recording: frozen rows are commuting rows in the tape).  Sets the
  // frozen set, so a replay from a checkpoint can restore the set it had.
  void freeze(const uint8_t* flags) override {
    need_batched();
    sync();
    for (int b = 0; b < n_inst; ++b) frozen[b] = flags[b] ? 1 : 0;
  }

  // The values and row counts of the flagged instances, then their rows
  // removed (retire).  On an unsorted store (sparse passes) only the flagged
  // instances' rows are sorted: they are copied out in order, sorted in the
  // child buffers and read there, as a sorted store would give them.
  void readout_retire(const uint64_t* diag, uint64_t nd, const uint8_t* flags, double* values,
                      uint64_t* counts) override {
    need_batched();
    sync();
    if (recording)
      fail(PQ_ERR_STATE, "instances cannot be retired while recording: freeze them instead (pq_freeze_instances)");
    std::vector<double> v(n_inst, 0.0);
    std::vector<uint64_t> bd;
    if (store_sorted || n == 0) {
      bd = instance_bounds();
      segments_expectation(store[cur].dev(), bd, diag, nd, v.data());
    } else {
      std::vector<uint8_t> drop(n_inst);
      for (int b = 0; b < n_inst; ++b) drop[b] = flags[b] ? 0 : 1;
      inst_flags.ensure((size_t)n_inst);
      HIPCHK(hipMemcpy(inst_flags.p, drop.data(), (size_t)n_inst, hipMemcpyHostToDevice));
      NoiseArgs a{};
      a.unit_mask = unit_mask;
      a.inst_shift = inst_shift;
      a.inst_drop = inst_flags.as<uint8_t>();
      child[0].ensure(n);
      child[1].ensure(n);
      const ChildDev c0 = child[0].dev(false), c1 = child[1].dev(false);
      const uint64_t n_tiles = (n + kTile - 1) / kTile;
      status.ensure((size_t)(n_tiles + 1) * sizeof(uint64_t));  // look-back words, then the tile ticket
      uint64_t* m_dev = small.as<uint64_t>() + kSlotExtract;
      HIPCHK(L::noise_compact(a, store[cur].dev(), n, StoreDev{c0.key, c0.val, c0.cap}, status.as<uint64_t>(), m_dev,
                              n_units, TapeDev{nullptr, nullptr, nullptr}, st));
      const uint64_t m = read_u64(m_dev);
      int in_b = 0;
      if (m > 1) HIPCHK(launch_sort_children<W>(c0, c1, m, &in_b, st));
      const ChildDev r = in_b ? c1 : c0;
      const StoreDev rs{r.key, r.val, r.cap};
      inst_bounds.ensure((size_t)(n_inst + 1) * sizeof(uint64_t));
      bd.assign(n_inst + 1, 0);
      if (m) {
        HIPCHK(launch_instance_bounds(rs.key + (size_t)(W - 1) * rs.cap, m_dev, inst_shift, n_inst,
                                      inst_bounds.as<uint64_t>(), st));
        HIPCHK(hipMemcpyAsync(bd.data(), inst_bounds.p, bd.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        segments_expectation(rs, bd, diag, nd, v.data());
      }
    }
    for (int b = 0; b < n_inst; ++b)
      if (flags[b]) {
        values[b] = v[b];
        counts[b] = bd[b + 1] - bd[b];
      }
    retire(flags);
  }

  // Remove every row of the flagged instances (finished parameter sets).
  void retire(const uint8_t* flags) override {
    need_batched();
    sync();
    if (recording)
      fail(PQ_ERR_STATE, "instances cannot be retired while recording: freeze them instead (pq_freeze_instances)");
    bool any = false;
    for (int b = 0; b < n_inst; ++b)
      if (flags[b] && !retired[b]) any = true;
    for (int b = 0; b < n_inst; ++b) retired[b] |= flags[b] ? 1 : 0;
    if (!any || n == 0) return;
    inst_flags.ensure((size_t)n_inst);
    HIPCHK(hipMemcpy(inst_flags.p, flags, (size_t)n_inst, hipMemcpyHostToDevice));
    NoiseArgs a{};
    a.unit_mask = unit_mask;
    a.inst_shift = inst_shift;
    a.inst_drop = inst_flags.as<uint8_t>();
    compact(a);
  }
};

// Width ladder (crates/pauli/src/engine.rs:289-299, plus W = 8); a batched
// engine also needs ceil(log2 n_inst) instance bits above the units.
template <int KIND>
EngineBase* make_engine(int n_units, int n_inst = 0) {
  int ib = 0;
  while (n_inst > 0 && (1 << ib) < n_inst) ++ib;
  if (n_inst > 0 && ib == 0) ib = 1;
  const int bits = 2 * n_units + ib;
  EngineBase* e = nullptr;
  if (bits <= 64)
    e = new Engine<KIND, 1>(n_units);
  else if (bits <= 128)
    e = new Engine<KIND, 2>(n_units);
  else if (bits <= 256)
    e = new Engine<KIND, 4>(n_units);
  else if (bits <= 512)
    e = new Engine<KIND, 8>(n_units);
  else
    fail(PQ_ERR_ARG, n_inst ? "too many units and instances for the widest key (512 bits)"
                            : "the device engine supports at most 256 units (qubits or sites)");
  if (n_inst > 0) {
    e->n_inst = n_inst;
    e->inst_shift = 64 - ib;
    e->unit_mask = (ib == 64) ? 0 : ((1ULL << e->inst_shift) - 1);
  }
  return e;
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return PQ_OK;
  } catch (const Failure& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return PQ_ERR_MEMORY;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return PQ_ERR_DEVICE;
  }
}

}  // namespace

// The error text pq_last_error returns, for the other C ABI units.
void set_last_error(const std::string& msg) { g_last_error = msg; }

}  // namespace pq

struct pq_engine {
  std::unique_ptr<pq::EngineBase> impl;
};

using pq::fail;
using pq::guarded;

#define PQ_NEED(cond, msg)                 \
  do {                                     \
    if (!(cond)) fail(PQ_ERR_ARG, (msg)); \
  } while (0)

extern "C" {

uint32_t pq_abi_version(void) { return PQ_ABI_VERSION; }

const char* pq_last_error(void) { return pq::g_last_error.c_str(); }

int pq_device_count(int* out) {
  return guarded([&] {
    PQ_NEED(out, "null output");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *out = c;
  });
}

int pq_engine_create_batched(uint32_t basis_kind, uint32_t n_units, uint32_t n_instances, pq_engine** out) {
  return guarded([&] {
    PQ_NEED(out, "null output");
    *out = nullptr;
    PQ_NEED(basis_kind <= 1, "basis_kind must be 0 (Pauli) or 1 (Majorana)");
    PQ_NEED(n_units >= 1, "n_units must be positive");
    PQ_NEED(n_instances >= 1 && n_instances <= (1u << 20), "n_instances must be in [1, 2^20]");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c == 0) fail(PQ_ERR_DEVICE, "no HIP device available");
    std::unique_ptr<pq_engine> e(new pq_engine);
    e->impl.reset(basis_kind == 0 ? pq::make_engine<pq::kPauli>((int)n_units, (int)n_instances)
                                  : pq::make_engine<pq::kMajorana>((int)n_units, (int)n_instances));
    *out = e.release();
  });
}

int pq_batch_info(const pq_engine* e, uint32_t* n_instances, uint32_t* instance_shift) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    if (n_instances) *n_instances = (uint32_t)e->impl->n_inst;
    if (instance_shift) *instance_shift = (uint32_t)e->impl->inst_shift;
  });
}

int pq_seed_batched(pq_engine* e, const uint32_t* instances, const uint64_t* keys, const double* coeffs,
                    uint64_t n) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    PQ_NEED(n == 0 || (instances && keys && coeffs), "null argument");
    e->impl->seed_batched(instances, keys, coeffs, n);
  });
}

int pq_apply_gates_batched(pq_engine* e, const uint64_t* generators, const double* angles, uint64_t n_gates,
                           uint64_t* n_terms_after) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    PQ_NEED(n_gates == 0 || (generators && angles), "null generator or angle array");
    e->impl->apply_gates_batched(generators, angles, n_gates, n_terms_after);
  });
}

int pq_expectation_batched(pq_engine* e, const uint64_t* diag_states, uint64_t n_diag_words, double* out) {
  return guarded([&] {
    PQ_NEED(e && out, "null argument");
    PQ_NEED(n_diag_words == 0 || diag_states, "null diag states");
    e->impl->expectation_batched(diag_states, n_diag_words, out);
  });
}

int pq_instance_counts(pq_engine* e, uint64_t* counts) {
  return guarded([&] {
    PQ_NEED(e && counts, "null argument");
    e->impl->instance_counts(counts);
  });
}

int pq_retire_instances(pq_engine* e, const uint8_t* flags) {
  return guarded([&] {
    PQ_NEED(e && flags, "null argument");
    e->impl->retire(flags);
  });
}

int pq_readout_retire(pq_engine* e, const uint64_t* diag_states, uint64_t n_diag_words, const uint8_t* flags,
                      double* values, uint64_t* counts) {
  return guarded([&] {
    PQ_NEED(e && flags && values && counts, "null argument");
    e->impl->readout_retire(diag_states, n_diag_words, flags, values, counts);
  });
}

int pq_engine_create(uint32_t basis_kind, uint32_t n_units, pq_engine** out) {
  return guarded([&] {
    PQ_NEED(out, "null output");
    *out = nullptr;
    PQ_NEED(basis_kind <= 1, "basis_kind must be 0 (Pauli) or 1 (Majorana)");
    PQ_NEED(n_units >= 1, "n_units must be positive");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c == 0) fail(PQ_ERR_DEVICE, "no HIP device available");
    std::unique_ptr<pq_engine> e(new pq_engine);
    e->impl.reset(basis_kind == 0 ? pq::make_engine<pq::kPauli>((int)n_units)
                                  : pq::make_engine<pq::kMajorana>((int)n_units));
    *out = e.release();
  });
}

void pq_engine_destroy(pq_engine* e) { delete e; }

int pq_engine_width(const pq_engine* e, uint32_t* n_words) {
  return guarded([&] {
    PQ_NEED(e && n_words, "null argument");
    *n_words = (uint32_t)e->impl->n_words;
  });
}

int pq_set_cutoff(pq_engine* e, int has_max_weight, uint32_t max_weight, int has_min_coeff, double min_coeff,
                  int has_min_terms, uint64_t min_terms) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    PQ_NEED(!has_min_coeff || !(min_coeff != min_coeff), "min_coeff is NaN");
    auto& x = *e->impl;
    x.has_maxw = has_max_weight != 0;
    x.maxw = max_weight;
    x.has_min = has_min_coeff != 0;
    x.min_coeff = min_coeff;
    x.has_min_terms = has_min_terms != 0;
    x.min_terms = min_terms;
  });
}

int pq_set_keep_kernel(pq_engine* e, pq_keep_batch_fn fn, void* ctx) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    if (fn && e->impl->n_inst) pq::fail(PQ_ERR_STATE, "a keep kernel is not supported on a batched engine");
    e->impl->sync();
    e->impl->keep_fn = fn;
    e->impl->keep_ctx = ctx;
  });
}

int pq_set_layer(pq_engine* e, uint32_t layer_index, uint32_t n_layers) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->sync();
    e->impl->layer_index = layer_index;
    e->impl->n_layers_ctx = n_layers;
  });
}

int pq_set_sparse(pq_engine* e, int on) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->sync();
    e->impl->sparse_runs = on ? 1 : 0;
  });
}

int pq_set_pairs(pq_engine* e, int on) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->sync();
    e->impl->pair_runs = on != 0;
  });
}

int pq_set_sparse_budget(pq_engine* e, uint64_t bytes) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->sync();
    e->impl->sparse_budget = bytes;
    e->impl->sparse_refused_at = ~0ULL;
  });
}

int pq_set_small_runs(pq_engine* e, int on) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->sync();
    e->impl->small_runs = on != 0;
  });
}

int pq_set_defer_cliffords(pq_engine* e, int on) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->defer = on != 0;
  });
}

int pq_set_differentiable(pq_engine* e, int on) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->sync();
    e->impl->differentiable = on != 0;
  });
}

int pq_seed(pq_engine* e, const uint64_t* keys, const double* coeffs, uint64_t n) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    PQ_NEED(n == 0 || (keys && coeffs), "null key or coefficient array");
    e->impl->seed(keys, coeffs, n);
  });
}

int pq_apply_rotation(pq_engine* e, const uint64_t* generator, double angle, uint64_t* n_terms_after) {
  return guarded([&] {
    PQ_NEED(e && generator, "null argument");
    e->impl->apply_gates(generator, &angle, 1, n_terms_after);
  });
}

int pq_apply_gates(pq_engine* e, const uint64_t* generators, const double* angles, uint64_t n_gates,
                   uint64_t* n_terms_after) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    PQ_NEED(n_gates == 0 || (generators && angles), "null generator or angle array");
    e->impl->apply_gates(generators, angles, n_gates, n_terms_after);
  });
}

int pq_scale_by_weight(pq_engine* e, const double* table, uint64_t len, int reclaim) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    PQ_NEED(table == nullptr || len > 0, "empty table");
    e->impl->noise(table, len, nullptr, nullptr, reclaim);
  });
}

int pq_scale_by_factors(pq_engine* e, const double* factors, uint64_t n) {
  return guarded([&] {
    PQ_NEED(e && factors, "null argument");
    e->impl->sync();
    PQ_NEED(n == e->impl->n, "one factor per row is required");
    e->impl->noise(nullptr, 0, factors, nullptr, 0);
  });
}

int pq_reclaim(pq_engine* e, const uint8_t* keep_flags, uint64_t n) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->sync();
    PQ_NEED(!keep_flags || n == e->impl->n, "one keep flag per row is required");
    e->impl->noise(nullptr, 0, nullptr, keep_flags, 1);
  });
}

int pq_len(const pq_engine* e, uint64_t* n) {
  return guarded([&] {
    PQ_NEED(e && n, "null argument");
    e->impl->sync();
    *n = e->impl->n;
  });
}

int pq_weights(pq_engine* e, uint32_t* out, uint64_t capacity) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->sync();
    PQ_NEED(out || e->impl->n == 0, "null argument");
    e->impl->weights(out, capacity);
  });
}

int pq_expectation(pq_engine* e, const uint64_t* diag_state, uint64_t n_diag_words, double* out) {
  return guarded([&] {
    PQ_NEED(e && out, "null argument");
    PQ_NEED(n_diag_words == 0 || diag_state, "null diag state");
    *out = e->impl->expectation(diag_state, n_diag_words);
  });
}

int pq_expectation_noalloc(pq_engine* e, const uint64_t* diag_state, uint64_t n_diag_words, double* out) {
  return guarded([&] {
    PQ_NEED(e && out, "null argument");
    PQ_NEED(n_diag_words == 0 || diag_state, "null basis state");
    *out = e->impl->expectation_noalloc(diag_state, n_diag_words);
  });
}

int pq_terms_below(pq_engine* e, double threshold, uint64_t* out) {
  return guarded([&] {
    PQ_NEED(e && out, "null argument");
    *out = e->impl->terms_below(threshold);
  });
}

int pq_export(pq_engine* e, uint64_t* keys, double* coeffs, uint64_t capacity, uint64_t* n_written) {
  return guarded([&] {
    PQ_NEED(e && n_written, "null argument");
    *n_written = e->impl->export_rows(keys, coeffs, capacity, true);
  });
}

int pq_export_raw(pq_engine* e, uint64_t* keys, double* coeffs, uint64_t capacity, uint64_t* n_written) {
  return guarded([&] {
    PQ_NEED(e && n_written, "null argument");
    *n_written = e->impl->export_rows(keys, coeffs, capacity, false);
  });
}

int pq_frame_is_identity(const pq_engine* e, int* out) {
  return guarded([&] {
    PQ_NEED(e && out, "null argument");
    *out = e->impl->frame_identity() ? 1 : 0;
  });
}

int pq_counters_get(const pq_engine* e, pq_counters* out) {
  return guarded([&] {
    PQ_NEED(e && out, "null argument");
    e->impl->counters(out);
  });
}

int pq_counters_reset(pq_engine* e) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->counters_reset();
  });
}

int pq_profile_enable(pq_engine* e, int on) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->sync();
    PQ_NEED(on >= 0, "negative sampling interval");
    e->impl->profiling = on != 0;
    e->impl->prof_every = on > 0 ? (uint64_t)on : 1;
    e->impl->prof_gate = 0;
    e->impl->prof = pq_profile{};
  });
}

int pq_profile_get(const pq_engine* e, pq_profile* out) {
  return guarded([&] {
    PQ_NEED(e && out, "null argument");
    *out = e->impl->prof;
  });
}

int pq_key_bytes(const pq_engine* e, uint64_t* out) {
  return guarded([&] {
    PQ_NEED(e && out, "null argument");
    e->impl->sync();
    *out = e->impl->key_bytes();
  });
}

int pq_tape_begin(pq_engine* e) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->tape_begin();
  });
}

int pq_tape_end(pq_engine* e) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->tape_end();
  });
}

int pq_tape_info(const pq_engine* e, uint64_t* n_gates, uint64_t* n_bytes) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->tape_info(n_gates, n_bytes);
  });
}

int pq_checkpoint_save(pq_engine* e, uint32_t* id) {
  return guarded([&] {
    PQ_NEED(e && id, "null argument");
    *id = e->impl->checkpoint_save();
  });
}

int pq_checkpoint_restore(pq_engine* e, uint32_t id) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->checkpoint_restore(id);
  });
}

int pq_checkpoint_free(pq_engine* e, uint32_t id) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->checkpoint_free(id);
  });
}

int pq_adjoint_seed(pq_engine* e, const uint64_t* diag_state, uint64_t n_diag_words) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    PQ_NEED(n_diag_words == 0 || diag_state, "null diag state");
    e->impl->adjoint_seed(diag_state, n_diag_words);
  });
}

int pq_adjoint_backward(pq_engine* e, double* grads, uint64_t n_gates) {
  return guarded([&] {
    PQ_NEED(e && (grads || n_gates == 0), "null argument");
    e->impl->adjoint_backward(grads, n_gates);
  });
}

int pq_freeze_instances(pq_engine* e, const uint8_t* flags) {
  return guarded([&] {
    PQ_NEED(e && flags, "null argument");
    e->impl->freeze(flags);
  });
}

int pq_adjoint_seed_batched(pq_engine* e, const uint64_t* diag_states, uint64_t n_diag_words) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->adjoint_seed_batched(diag_states, n_diag_words);
  });
}

int pq_adjoint_backward_batched(pq_engine* e, double* grads, uint64_t n_gates) {
  return guarded([&] {
    PQ_NEED(e && (grads || n_gates == 0), "null argument");
    e->impl->adjoint_backward_batched(grads, n_gates);
  });
}

int pq_adjoint_values(pq_engine* e, double* out, uint64_t capacity, uint64_t* n_written) {
  return guarded([&] {
    PQ_NEED(e && n_written && (out || capacity == 0), "null argument");
    *n_written = e->impl->adjoint_values(out, capacity);
  });
}

int pq_synchronize(pq_engine* e) {
  return guarded([&] {
    PQ_NEED(e, "null engine");
    e->impl->sync();
  });
}

}  // extern "C"

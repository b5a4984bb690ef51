// pq_surrogate.cpp -- the symbolic surrogate: host builder and device model.
//
// Reference: crates/surrogate/src (engine.rs build / apply_truncation /
// compile, symcoeff.rs SymbolicCoeff and CompiledCoeff, model.rs
// SurrogateModel).
//
// The builder propagates the observable once with symbolic coefficients.  A
// coefficient is a node of one append-only arena (Scalar, Add, Scale, Cos,
// Sin with the monomial count, the frequency range and the coefficient upper
// bound of its subtree, symcoeff.rs:22-150); nodes are shared between terms
// exactly where the reference shares Arcs.  Building is pointer-shaped host
// work, as in the reference.  Compiling flattens the surviving DAG into a
// tape, orders it by level and uploads it; evaluation runs on the device for
// a whole batch of parameter sets at once (pq_tape.hip).
//
// Where the reference leaves an order to hashing or threads, this code fixes
// one: terms in store order (key ascending, word W-1 most significant), the
// sum over terms pairwise, and simplify's monomials in first-seen order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <system_error>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/propaq_amd.h"
#include "pq_tableau.h"
#include "pq_tape.h"

namespace pq {
void set_last_error(const std::string& msg);

namespace sur {
namespace {

using u128 = unsigned __int128;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr u128 kU128Max = ~(u128)0;
constexpr uint32_t kU32Max = 0xFFFFFFFFu;
constexpr double kEps = 1e-9;          // PHASE_ONLY_EPS, partitioned_termsum.rs:41
constexpr double kOverlapEps = 1e-15;  // engine.rs:80

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
#define SHIPCHK(x) hip_check((x), #x)

u128 sat_add(u128 a, u128 b) { return a > kU128Max - b ? kU128Max : a + b; }
uint32_t sat_inc(uint32_t a) { return a == kU32Max ? a : a + 1; }
uint32_t sat_add32(uint32_t a, uint32_t b) { return a > kU32Max - b ? kU32Max : a + b; }

// ------------------------------------------------------------------ nodes

struct Node {
  double val = 0.0;    // Scalar value or Scale factor
  double upper = 0.0;  // max |scalar| times the Scale factors above it
  u128 count = 1;      // monomial instances (saturating)
  uint32_t a = kNone, b = kNone, p = 0, minf = 0, maxf = 0;
  uint8_t kind = kTapeScalar;
};

struct Arena {
  std::vector<Node> n;

  uint32_t push(const Node& x) {
    if (n.size() >= (size_t)kNone - 1) fail(PQ_ERR_MEMORY, "surrogate: more than 2^32 - 2 coefficient nodes");
    n.push_back(x);
    return (uint32_t)(n.size() - 1);
  }
  uint32_t scalar(double c) {
    Node x;
    x.kind = kTapeScalar;
    x.val = c;
    x.upper = std::fabs(c);
    return push(x);
  }
  uint32_t add(uint32_t a, uint32_t b) {
    const Node &A = n[a], &B = n[b];
    Node x;
    x.kind = kTapeAdd;
    x.a = a;
    x.b = b;
    x.count = sat_add(A.count, B.count);
    x.minf = std::min(A.minf, B.minf);
    x.maxf = std::max(A.maxf, B.maxf);
    x.upper = std::max(A.upper, B.upper);
    return push(x);
  }
  uint32_t scale(double k, uint32_t a) {
    const Node& A = n[a];
    Node x;
    x.kind = kTapeScale;
    x.a = a;
    x.val = k;
    x.count = A.count;
    x.minf = A.minf;
    x.maxf = A.maxf;
    x.upper = std::fabs(k) * A.upper;
    return push(x);
  }
  uint32_t trig(uint8_t kind, uint32_t p, uint32_t a) {
    const Node& A = n[a];
    Node x;
    x.kind = kind;
    x.a = a;
    x.p = p;
    x.count = A.count;
    x.minf = sat_inc(A.minf);
    x.maxf = sat_inc(A.maxf);
    x.upper = A.upper;
    return push(x);
  }
  uint32_t add_opt(uint32_t a, uint32_t b) {
    if (a == kNone) return b;
    if (b == kNone) return a;
    return add(a, b);
  }
  uint32_t signed_(double s, uint32_t x) { return s == 1.0 ? x : scale(s, x); }
};

// ---------------------------------------------------------------- pruning

// combine_scale_exp (symcoeff.rs:309): exp + ceil(log2 |k|), both saturating
// to i32 as Rust's `as i32` and saturating_add do.
int32_t sat_i32(double v) {
  if (std::isnan(v)) return 0;
  if (v >= 2147483647.0) return INT32_MAX;
  if (v <= -2147483648.0) return INT32_MIN;
  return (int32_t)v;
}
int32_t combine_scale_exp(int32_t e, double k) {
  const int32_t ke = sat_i32(std::ceil(std::log2(std::fabs(k))));
  const int64_t s = (int64_t)e + ke;
  return (int32_t)std::max<int64_t>(INT32_MIN, std::min<int64_t>(INT32_MAX, s));
}
bool doomed_by_coeff(int32_t se, double upper, double cutoff) {
  if (cutoff <= 0.0) return false;
  return (double)se + std::log2(upper) < std::log2(cutoff);
}

struct PKey {
  uint32_t id, depth;
  int32_t se;
  bool operator==(const PKey& o) const { return id == o.id && depth == o.depth && se == o.se; }
};
struct PKeyHash {
  size_t operator()(const PKey& k) const {
    uint64_t h = ((uint64_t)k.id << 32) ^ ((uint64_t)k.depth * 0x9E3779B97F4A7C15ULL) ^ (uint32_t)k.se;
    h ^= h >> 29;
    h *= 0xBF58476D1CE4E5B9ULL;
    return (size_t)(h ^ (h >> 32));
  }
};

// prune_node (symcoeff.rs:330-520): drops every monomial with more than cap
// trig factors or whose coefficient bound is below cut, keeping sharing.  The
// memo is keyed like the reference's (node, depth, scale exponent) and may be
// shared by all coefficients of one pass: a key's result depends on the key
// alone.  The walk keeps its own stack: a coefficient's DAG is as deep as the
// gates that touched it.
struct Pruner {
  Arena& A;
  bool has_f, has_c;
  uint32_t cap;
  double cut;
  std::unordered_map<PKey, uint32_t, PKeyHash> memo;
  struct Frame {
    uint32_t id, depth;
    int32_t se;
    bool exit;
  };
  std::vector<Frame> stack;

  PKey key_of(uint32_t id, uint32_t depth, int32_t se) const { return PKey{id, has_f ? depth : 0u, has_c ? se : 0}; }
  static uint32_t deeper(uint32_t d) { return d == kU32Max ? d : d + 1; }

  void push(uint32_t id, uint32_t depth, int32_t se) {
    if (!memo.count(key_of(id, depth, se))) stack.push_back(Frame{id, depth, se, false});
  }

  uint32_t go(uint32_t root, uint32_t depth0, int32_t se0) {
    const PKey rk = key_of(root, depth0, se0);
    auto found = memo.find(rk);
    if (found != memo.end()) return found->second;
    stack.clear();
    stack.push_back(Frame{root, depth0, se0, false});
    while (!stack.empty()) {
      const Frame fr = stack.back();
      stack.pop_back();
      const PKey key = key_of(fr.id, fr.depth, fr.se);
      const Node nd = A.n[fr.id];
      if (!fr.exit) {
        if (memo.count(key)) continue;
        if ((has_f && sat_add32(fr.depth, nd.minf) > cap) || (has_c && doomed_by_coeff(fr.se, nd.upper, cut))) {
          memo.emplace(key, kNone);
          continue;
        }
        if ((has_f && !has_c && sat_add32(fr.depth, nd.maxf) <= cap) || nd.kind == kTapeScalar) {
          memo.emplace(key, fr.id);
          continue;
        }
        stack.push_back(Frame{fr.id, fr.depth, fr.se, true});
        if (nd.kind == kTapeAdd) {
          push(nd.b, fr.depth, fr.se);
          push(nd.a, fr.depth, fr.se);
        } else if (nd.kind == kTapeScale) {
          push(nd.a, fr.depth, has_c ? combine_scale_exp(fr.se, nd.val) : fr.se);
        } else {
          push(nd.a, deeper(fr.depth), fr.se);
        }
        continue;
      }
      uint32_t r;
      if (nd.kind == kTapeAdd) {
        const uint32_t x = memo.at(key_of(nd.a, fr.depth, fr.se));
        const uint32_t y = memo.at(key_of(nd.b, fr.depth, fr.se));
        if (x == kNone || y == kNone)
          r = (y == kNone) ? x : y;
        else
          r = (x == nd.a && y == nd.b) ? fr.id : A.add(x, y);
      } else if (nd.kind == kTapeScale) {
        const uint32_t x = memo.at(key_of(nd.a, fr.depth, has_c ? combine_scale_exp(fr.se, nd.val) : fr.se));
        r = x == kNone ? kNone : (x == nd.a ? fr.id : A.scale(nd.val, x));
      } else {
        const uint32_t x = memo.at(key_of(nd.a, deeper(fr.depth), fr.se));
        r = x == kNone ? kNone : (x == nd.a ? fr.id : A.trig(nd.kind, nd.p, x));
      }
      memo.emplace(key, r);
    }
    return memo.at(rk);
  }
};

// --------------------------------------------------------------- simplify

// A monomial's trig factors: (param, cos power, sin power) triples, params
// ascending (the reference's BTreeMap FactorRun, symcoeff.rs:522-540).
using Run = std::vector<uint32_t>;
struct RunHash {
  size_t operator()(const Run& r) const {
    uint64_t h = 0xcbf29ce484222325ULL;
    for (uint32_t v : r) h = (h ^ v) * 0x100000001b3ULL;
    return (size_t)h;
  }
};
using Terms = std::vector<std::pair<Run, double>>;

// One more cos (or sin) of parameter p, in place.
void bump(Run& r, uint32_t p, bool is_sin) {
  size_t lo = 0, hi = r.size() / 3;  // first triple with param >= p
  while (lo < hi) {
    const size_t mid = (lo + hi) / 2;
    if (r[3 * mid] < p)
      lo = mid + 1;
    else
      hi = mid;
  }
  if (3 * lo < r.size() && r[3 * lo] == p) {
    r[3 * lo + (is_sin ? 2 : 1)] += 1;
    return;
  }
  const uint32_t add[3] = {p, is_sin ? 0u : 1u, is_sin ? 1u : 0u};
  r.insert(r.begin() + 3 * lo, add, add + 3);
}

// group (symcoeff.rs:547-553): scalars of equal runs summed from 0.0, in
// first-seen order.
Terms group(Terms&& in) {
  std::unordered_map<Run, size_t, RunHash> at;
  Terms out;
  out.reserve(in.size());
  for (auto& t : in) {
    auto it = at.find(t.first);
    if (it == at.end()) {
      at.emplace(t.first, out.size());
      out.emplace_back(std::move(t.first), 0.0 + t.second);
    } else {
      out[it->second].second += t.second;
    }
  }
  return out;
}

// simplify_batch (symcoeff.rs:555-666) over one pass's coefficients: every
// node's monomials once, in a post-order walk with its own stack; a child's
// list is moved into its last reader and copied for the others (the
// reference's take_or_clone); then each coefficient is rebuilt as a balanced
// sum of its nonzero monomials.
struct Simplifier {
  Arena& A;
  std::unordered_map<uint32_t, Terms> memo;
  std::unordered_map<uint32_t, uint32_t> remaining;  // readers not yet served, root uses included
  std::unordered_map<uint32_t, uint32_t> rebuilt;

  Terms take(uint32_t id) {
    uint32_t& r = remaining.at(id);
    --r;
    if (r == 0) {
      auto it = memo.find(id);
      Terms t = std::move(it->second);
      memo.erase(it);
      return t;
    }
    return memo.at(id);
  }

  void expand_all(const std::vector<uint32_t>& roots) {
    std::vector<uint32_t> work;
    for (uint32_t r : roots) {
      if (r == kNone) continue;
      if (remaining[r]++ == 0) work.push_back(r);
    }
    while (!work.empty()) {  // readers per node
      const uint32_t x = work.back();
      work.pop_back();
      const Node& nd = A.n[x];
      if (nd.kind == kTapeScalar) continue;
      if (remaining[nd.a]++ == 0) work.push_back(nd.a);
      if (nd.kind == kTapeAdd && remaining[nd.b]++ == 0) work.push_back(nd.b);
    }
    std::vector<std::pair<uint32_t, bool>> stack;
    std::unordered_map<uint32_t, uint8_t> scheduled;
    for (uint32_t r : roots) {
      if (r == kNone || scheduled.count(r)) continue;
      scheduled.emplace(r, 1);
      stack.emplace_back(r, false);
      while (!stack.empty()) {
        const auto [id, exit] = stack.back();
        stack.pop_back();
        const Node nd = A.n[id];
        if (!exit) {
          stack.emplace_back(id, true);
          if (nd.kind == kTapeAdd && scheduled.emplace(nd.b, 1).second) stack.emplace_back(nd.b, false);
          if (nd.kind != kTapeScalar && scheduled.emplace(nd.a, 1).second) stack.emplace_back(nd.a, false);
          continue;
        }
        Terms t;
        if (nd.kind == kTapeScalar) {
          t.emplace_back(Run{}, nd.val);
        } else if (nd.kind == kTapeAdd) {
          t = take(nd.a);
          Terms tb = take(nd.b);
          t.insert(t.end(), std::make_move_iterator(tb.begin()), std::make_move_iterator(tb.end()));
          t = group(std::move(t));
        } else if (nd.kind == kTapeScale) {
          t = take(nd.a);
          for (auto& m : t) m.second *= nd.val;
        } else {
          t = take(nd.a);
          for (auto& m : t) bump(m.first, nd.p, nd.kind == kTapeSin);
        }
        memo[id] = std::move(t);
      }
    }
  }

  uint32_t leaf(const Run& run, double scalar) {
    uint32_t node = A.scalar(scalar);
    for (size_t i = 0; i < run.size(); i += 3) {
      for (uint32_t c = 0; c < run[i + 1]; ++c) node = A.trig(kTapeCos, run[i], node);
      for (uint32_t s = 0; s < run[i + 2]; ++s) node = A.trig(kTapeSin, run[i], node);
    }
    return node;
  }

  // The rebuilt coefficient of a root given to expand_all.
  uint32_t simplify(uint32_t id) {
    if (id == kNone) return kNone;
    auto it = rebuilt.find(id);
    if (it != rebuilt.end()) return it->second;
    const Terms& t = memo.at(id);
    std::vector<uint32_t> leaves;
    for (const auto& m : t)
      if (m.second != 0.0) leaves.push_back(leaf(m.first, m.second));
    uint32_t root = kNone;
    if (!leaves.empty()) {
      while (leaves.size() > 1) {  // rebuild_balanced, symcoeff.rs:668-688
        std::vector<uint32_t> nxt;
        for (size_t i = 0; i + 1 < leaves.size(); i += 2) nxt.push_back(A.add(leaves[i], leaves[i + 1]));
        if (leaves.size() % 2) nxt.push_back(leaves.back());
        leaves.swap(nxt);
      }
      root = leaves[0];
    }
    rebuilt.emplace(id, root);
    return root;
  }
};

// ------------------------------------------------------------------ model

struct Model {
  uint32_t n_params = 0;
  // host tape in topological (compile) order, for export
  std::vector<TapeOp> tape;
  std::vector<double> overlaps;
  std::vector<uint32_t> roots;  // compile order index or kTapeEmpty
  u128 monomials = 0;
  // device: the tape by level
  std::vector<uint32_t> level_off;
  std::vector<uint32_t> level_of;  // compile index -> device index
  double tape_bytes_per_set = 0.0;
  size_t n_stored = 0;  // ops kept in device memory after folding
  FusedOp* d_ops = nullptr;
  uint32_t* d_level_off = nullptr;
  uint32_t* d_roots = nullptr;
  double* d_overlap = nullptr;
  // Two passes in flight, each on its own stream with its own value table,
  // cos/sin table and partial sums: independent parameter sets, so one
  // pass's levels fill the chip while the other's narrow levels drain.
  struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t copied = nullptr;  // this slot's host cos/sin table has reached the device
    double* val = nullptr;
    size_t val_bytes = 0;
    double* lut = nullptr;
    size_t lut_bytes = 0;
    double* part[2] = {nullptr, nullptr};
    size_t part_bytes = 0;
  };
  static constexpr int kSlots = 2;
  Slot slot[kSlots];
  double* h_buf = nullptr;  // pinned: the cos/sin table and the results of one pass
  size_t host_bytes = 0;
  std::vector<hipEvent_t> prof_ev;  // per pass: tape start, tape end, sum end
  bool profiling = false;
  uint64_t max_sets = 0;   // caller's bound on parameter sets per pass (0: memory decides)
  uint64_t sets_held = 0;  // sets per pass the device tables are sized for
  double t_tape = 0.0, t_sum = 0.0, bytes = 0.0;
  uint64_t launches = 0;

  ~Model() {
    if (d_ops) (void)hipFree(d_ops);
    if (d_level_off) (void)hipFree(d_level_off);
    if (d_roots) (void)hipFree(d_roots);
    if (d_overlap) (void)hipFree(d_overlap);
    for (Slot& sl : slot) {
      if (sl.val) (void)hipFree(sl.val);
      if (sl.lut) (void)hipFree(sl.lut);
      for (double* p : sl.part)
        if (p) (void)hipFree(p);
      if (sl.copied) (void)hipEventDestroy(sl.copied);
      if (sl.st) (void)hipStreamDestroy(sl.st);
    }
    for (hipEvent_t e : prof_ev)
      if (e) (void)hipEventDestroy(e);
    if (h_buf) (void)hipHostFree(h_buf);
  }

  static void grow(void** p, size_t* have, size_t want) {
    if (want <= *have) return;
    if (*p) SHIPCHK(hipFree(*p));
    *p = nullptr;
    *have = 0;
    SHIPCHK(hipMalloc(p, want));
    *have = want;
  }

  // Validates the tape, sums the monomial count, folds single-use chains into
  // their reader, orders the stored ops by level and uploads them.
  void finalize() {
    const size_t n = tape.size();
    std::vector<u128> cnt(n, 1);
    std::vector<uint32_t> use(n, 0);
    for (size_t i = 0; i < n; ++i) {
      const TapeOp& op = tape[i];
      if (op.kind > kTapeSin) fail(PQ_ERR_ARG, "unknown tape op kind");
      if (op.kind != kTapeScalar && (op.a >= i || (op.kind == kTapeAdd && op.b >= i)))
        fail(PQ_ERR_ARG, "tape operand does not precede its op");
      if ((op.kind == kTapeCos || op.kind == kTapeSin) && op.p >= n_params)
        fail(PQ_ERR_ARG, "tape parameter index beyond n_params");
      if ((op.kind == kTapeCos || op.kind == kTapeSin) && op.p >= (1u << 29))
        fail(PQ_ERR_ARG, "tape parameter index beyond 2^29");
      if (op.kind == kTapeAdd) {
        cnt[i] = sat_add(cnt[op.a], cnt[op.b]);
        use[op.a]++;
        use[op.b]++;
      } else if (op.kind != kTapeScalar) {
        cnt[i] = cnt[op.a];
        use[op.a]++;
      }
    }
    monomials = 0;
    std::vector<uint8_t> is_root(n, 0);
    for (size_t t = 0; t < roots.size(); ++t) {
      if (roots[t] == kTapeEmpty) continue;
      if (roots[t] >= n) fail(PQ_ERR_ARG, "term root beyond the tape");
      monomials = sat_add(monomials, cnt[roots[t]]);
      is_root[roots[t]] = 1;
    }
    // inl[i]: op i is folded into its only reader; d[i] the length of the
    // folded chain ending at i (at most kChain - 1, so a stored
    // single-operand op with its own step fits kChain steps)
    std::vector<uint8_t> inl(n, 0), d(n, 0);
    for (size_t i = 0; i < n; ++i) {
      const TapeOp& op = tape[i];
      const bool unary = op.kind == kTapeScale || op.kind == kTapeCos || op.kind == kTapeSin;
      if (!unary || use[i] != 1 || is_root[i]) continue;
      const int len = 1 + (inl[op.a] ? d[op.a] : 0);
      if (len > kChain - 1) continue;
      inl[i] = 1;
      d[i] = (uint8_t)len;
    }
    // stored ops, their levels and device order
    auto base_of = [&](uint32_t x) {
      while (inl[x]) x = tape[x].a;
      return x;
    };
    std::vector<uint32_t> lev(n, 0);
    uint32_t max_lev = 0;
    size_t n_store = 0;
    for (size_t i = 0; i < n; ++i) {
      if (inl[i]) continue;
      ++n_store;
      const TapeOp& op = tape[i];
      if (op.kind == kTapeScalar) continue;
      uint32_t l = 1 + lev[base_of(op.a)];
      if (op.kind == kTapeAdd) l = std::max(l, 1 + lev[base_of(op.b)]);
      lev[i] = l;
      max_lev = std::max(max_lev, l);
    }
    const uint32_t n_lev = n_store ? max_lev + 1 : 0;
    level_off.assign(n_lev + 1, 0);
    for (size_t i = 0; i < n; ++i)
      if (!inl[i]) level_off[lev[i] + 1]++;
    for (uint32_t l = 0; l < n_lev; ++l) level_off[l + 1] += level_off[l];
    std::vector<uint32_t> fill(level_off.begin(), level_off.end() - (n_lev ? 1 : 0));
    level_of.assign(n, kTapeEmpty);
    for (size_t i = 0; i < n; ++i)
      if (!inl[i]) level_of[i] = fill[lev[i]]++;
    // the folded chain from x down to its stored base, innermost step first
    auto chain = [&](uint32_t x, bool include_self, FusedOp& f, int off, uint32_t& len) {
      uint32_t st[kChain];
      double fa[kChain];
      int k = 0;
      auto push = [&](uint32_t y) {
        const TapeOp& o = tape[y];
        st[k] = (o.p << 3) | o.kind;
        fa[k] = o.val;
        ++k;
      };
      if (include_self) {
        push(x);
        x = tape[x].a;
      }
      while (inl[x]) {
        push(x);
        x = tape[x].a;
      }
      for (int j = 0; j < k; ++j) {
        f.step[off + j] = st[k - 1 - j];
        f.f[off + j] = fa[k - 1 - j];
      }
      len = (uint32_t)k;
      return level_of[x];
    };
    std::vector<FusedOp> dev(n_store);
    tape_bytes_per_set = 0.0;
    for (size_t i = 0; i < n; ++i) {
      if (inl[i]) continue;
      const TapeOp& op = tape[i];
      FusedOp f;
      std::memset(&f, 0, sizeof(f));
      if (op.kind == kTapeScalar) {
        f.kind = kTapeScalar;
        f.val = op.val;
        tape_bytes_per_set += 8.0;
      } else if (op.kind == kTapeAdd) {
        f.kind = kTapeAdd;
        f.a = chain(op.a, false, f, 0, f.na);
        f.b = chain(op.b, false, f, kChain, f.nb);
        tape_bytes_per_set += 24.0;
      } else {
        f.kind = kTapeScale;
        f.a = chain((uint32_t)i, true, f, 0, f.na);
        tape_bytes_per_set += 16.0;
      }
      dev[level_of[i]] = f;
    }
    std::vector<uint32_t> droots(roots.size());
    for (size_t t = 0; t < roots.size(); ++t) droots[t] = roots[t] == kTapeEmpty ? kTapeEmpty : level_of[roots[t]];
    n_stored = n_store;
    for (Slot& sl : slot) {
      SHIPCHK(hipStreamCreate(&sl.st));
      SHIPCHK(hipEventCreate(&sl.copied));
    }
    if (n_store) {
      SHIPCHK(hipMalloc((void**)&d_ops, n_store * sizeof(FusedOp)));
      SHIPCHK(hipMemcpy(d_ops, dev.data(), n_store * sizeof(FusedOp), hipMemcpyHostToDevice));
    }
    SHIPCHK(hipMalloc((void**)&d_level_off, level_off.size() * sizeof(uint32_t)));
    SHIPCHK(hipMemcpy(d_level_off, level_off.data(), level_off.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (!roots.empty()) {
      SHIPCHK(hipMalloc((void**)&d_roots, droots.size() * sizeof(uint32_t)));
      SHIPCHK(hipMemcpy(d_roots, droots.data(), droots.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
      SHIPCHK(hipMalloc((void**)&d_overlap, overlaps.size() * sizeof(double)));
      SHIPCHK(hipMemcpy(d_overlap, overlaps.data(), overlaps.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  }

  // make_lut (model.rs:106): cos and sin of every angle of b sets with the C
  // library's cos / sin, as the reference's f64::cos / f64::sin; a large
  // table is split over host threads by set.
  void fill_lut(double* lut, const double* params, uint32_t b, uint64_t stride) const {
    const uint64_t work = (uint64_t)n_params * b;
    unsigned nt = std::thread::hardware_concurrency();
    nt = std::max(1u, std::min(nt, 16u));
    if (work < 32768 || b < 2 * nt) nt = 1;
    auto body = [&](unsigned k) {
      const uint32_t lo = (uint32_t)((uint64_t)b * k / nt), hi = (uint32_t)((uint64_t)b * (k + 1) / nt);
      for (uint32_t s = lo; s < hi; ++s)
        for (uint32_t p = 0; p < n_params; ++p) {
          const double th = params[(uint64_t)s * stride + p];
          lut[(2ull * p) * b + s] = std::cos(th);
          lut[(2ull * p + 1) * b + s] = std::sin(th);
        }
    };
    if (nt == 1) {
      body(0);
      return;
    }
    // every started thread is joined on the way out, also when starting a
    // later one throws; slices whose thread could not start run here
    std::vector<std::thread> pool;
    pool.reserve(nt - 1);
    struct Joiner {
      std::vector<std::thread>& threads;
      ~Joiner() {
        for (auto& t : threads)
          if (t.joinable()) t.join();
      }
    } joiner{pool};
    unsigned started = 1;
    try {
      for (; started < nt; ++started) pool.emplace_back(body, started);
    } catch (const std::system_error&) {
    }
    for (unsigned k = started; k < nt; ++k) body(k);
    body(0);
  }

  // Levels whose work (ops x sets) is at most this run inside one workgroup.
  static constexpr uint64_t kBlockLevelWork = 1024;

  // A large request runs in at least kPipePasses passes of at least kPipeMin
  // sets over the two slots: the host computes and uploads the next pass's
  // cos/sin table while the device evaluates, and two passes share the chip.
  static constexpr uint64_t kPipeMin = 4096;
  static constexpr uint64_t kPipePasses = 4;

  hipEvent_t prof_event(size_t i) {
    while (prof_ev.size() <= i) {
      hipEvent_t e = nullptr;
      SHIPCHK(hipEventCreate(&e));
      prof_ev.push_back(e);
    }
    return prof_ev[i];
  }

  void evaluate(const double* params, uint64_t n_sets, uint64_t stride, double* out) {
    if (n_sets == 0) return;
    const uint64_t n_terms = roots.size();
    const uint64_t n_ops = n_stored;
    if (n_terms == 0) {
      for (uint64_t s = 0; s < n_sets; ++s) out[s] = 0.0;
      return;
    }
    // sets per pass: the value table n_ops x B doubles within a quarter of free
    // memory (asked only when the tables already held are too small)
    const uint64_t per_set = std::max<uint64_t>(n_ops, 1) * sizeof(double) +
                             2 * ((n_terms + kTermChunk - 1) / kTermChunk) * sizeof(double) +
                             4 * (uint64_t)n_params * sizeof(double);
    uint64_t fit = sets_held;
    if (n_sets > sets_held) {
      size_t free_b = 0, total_b = 0;
      SHIPCHK(hipMemGetInfo(&free_b, &total_b));
      const uint64_t budget = std::max<uint64_t>(free_b / 4, 64ull << 20);
      if (per_set > budget) fail(PQ_ERR_MEMORY, "surrogate tape does not fit device memory for one parameter set");
      fit = std::max<uint64_t>(sets_held, budget / per_set);
    }
    uint64_t B = std::min<uint64_t>(n_sets, std::max<uint64_t>(1, fit));
    B = std::min<uint64_t>(B, 1u << 20);
    if (n_sets >= 2 * kPipeMin)
      B = std::min<uint64_t>(B, std::max<uint64_t>(kPipeMin, (n_sets + kPipePasses - 1) / kPipePasses));
    if (max_sets) B = std::min<uint64_t>(B, max_sets);
    const uint64_t lut_n = std::max<uint64_t>(2ull * n_params * B, 1);
    const uint64_t n1 = (n_terms + kTermChunk - 1) / kTermChunk;
    for (Slot& sl : slot) {
      grow((void**)&sl.val, &sl.val_bytes, std::max<uint64_t>(n_ops, 1) * B * sizeof(double));
      grow((void**)&sl.lut, &sl.lut_bytes, lut_n * sizeof(double));
      if (n1 * B * sizeof(double) > sl.part_bytes) {
        for (double*& p : sl.part) {
          if (p) SHIPCHK(hipFree(p));
          p = nullptr;
        }
        sl.part_bytes = 0;
        for (double*& p : sl.part) SHIPCHK(hipMalloc((void**)&p, n1 * B * sizeof(double)));
        sl.part_bytes = n1 * B * sizeof(double);
      }
    }
    sets_held = std::max(sets_held, B);
    const size_t host_want = (kSlots * lut_n + n_sets) * sizeof(double);
    if (host_want > host_bytes) {
      if (h_buf) SHIPCHK(hipHostFree(h_buf));
      h_buf = nullptr;
      host_bytes = 0;
      SHIPCHK(hipHostMalloc((void**)&h_buf, host_want, hipHostMallocDefault));
      host_bytes = host_want;
    }
    double* res = h_buf + kSlots * lut_n;
    t_tape = t_sum = bytes = 0.0;
    launches = 0;
    const uint64_t n_pass = (n_sets + B - 1) / B;
    const uint32_t n_lev = (uint32_t)level_off.size() - 1;
    for (uint64_t i = 0; i < n_pass; ++i) {
      Slot& sl = slot[i % kSlots];
      hipStream_t st = sl.st;
      const uint64_t s0 = i * B;
      const uint32_t b = (uint32_t)std::min<uint64_t>(B, n_sets - s0);
      double* lut_h = h_buf + (i % kSlots) * lut_n;
      // this slot's previous pass has copied its table (and, stream-ordered,
      // finishes before this pass's kernels overwrite its buffers)
      if (i >= (uint64_t)kSlots) SHIPCHK(hipEventSynchronize(sl.copied));
      fill_lut(lut_h, params + s0 * stride, b, stride);
      if (n_params)
        SHIPCHK(hipMemcpyAsync(sl.lut, lut_h, 2ull * n_params * b * sizeof(double), hipMemcpyHostToDevice, st));
      SHIPCHK(hipEventRecord(sl.copied, st));
      if (profiling) SHIPCHK(hipEventRecord(prof_event(3 * i), st));
      uint32_t l = 0;
      while (l < n_lev) {
        uint32_t e = l;
        while (e < n_lev && (uint64_t)(level_off[e + 1] - level_off[e]) * b <= kBlockLevelWork) ++e;
        if (e > l) {
          SHIPCHK(launch_tape_levels_block(d_ops, d_level_off, l, e, sl.lut, sl.val, b, st));
          ++launches;
          l = e;
        } else {
          SHIPCHK(launch_tape_level(d_ops, level_off[l], level_off[l + 1] - level_off[l], sl.lut, sl.val, b, st));
          ++launches;
          ++l;
        }
      }
      if (profiling) SHIPCHK(hipEventRecord(prof_event(3 * i + 1), st));
      uint64_t n = n_terms;
      SHIPCHK(launch_term_sum(d_roots, d_overlap, sl.val, n, b, sl.part[0], st));
      int cur = 0;
      n = (n + kTermChunk - 1) / kTermChunk;
      while (n > 1) {
        SHIPCHK(launch_term_sum(nullptr, nullptr, sl.part[cur], n, b, sl.part[cur ^ 1], st));
        cur ^= 1;
        n = (n + kTermChunk - 1) / kTermChunk;
      }
      if (profiling) SHIPCHK(hipEventRecord(prof_event(3 * i + 2), st));
      SHIPCHK(hipMemcpyAsync(res + s0, sl.part[cur], b * sizeof(double), hipMemcpyDeviceToHost, st));
      if (profiling) bytes += tape_bytes_per_set * b;
    }
    for (Slot& sl : slot) SHIPCHK(hipStreamSynchronize(sl.st));
    std::memcpy(out, res, n_sets * sizeof(double));
    if (profiling) {
      // passes overlap: device time is the span from the first pass's start
      // to the last end, split between tape and term sum by their own times
      float span = 0.f, tape_sum = 0.f, sum_sum = 0.f;
      const uint64_t last = n_pass - 1;
      const uint64_t prev = n_pass >= 2 ? n_pass - 2 : last;
      float a = 0.f, c = 0.f;
      SHIPCHK(hipEventElapsedTime(&span, prof_ev[0], prof_ev[3 * last + 2]));
      if (prev != last) {
        SHIPCHK(hipEventElapsedTime(&a, prof_ev[0], prof_ev[3 * prev + 2]));
        span = std::max(span, a);
      }
      for (uint64_t i = 0; i < n_pass; ++i) {
        SHIPCHK(hipEventElapsedTime(&a, prof_ev[3 * i], prof_ev[3 * i + 1]));
        SHIPCHK(hipEventElapsedTime(&c, prof_ev[3 * i + 1], prof_ev[3 * i + 2]));
        tape_sum += a;
        sum_sum += c;
      }
      const float tot = tape_sum + sum_sum;
      t_tape = tot > 0.f ? span * (tape_sum / tot) : 0.0;
      t_sum = tot > 0.f ? span * (sum_sum / tot) : 0.0;
    }
  }
};

// ---------------------------------------------------------------- builder

struct Cfg {
  bool has_f = false, has_c = false, has_w = false, has_min = false, simplify = false;
  uint32_t freq = 0, weight = 0;
  double coeff = 0.0;
  uint64_t min_terms = 0;
};

struct BuilderBase {
  virtual ~BuilderBase() = default;
  int kind = 0, n_units = 0, n_words = 1;
  Cfg cfg;
  uint64_t compact_min = 1u << 20;  // arena size below which it is never compacted
  uint64_t visited = 0, emitted = 0, hits = 0;
  double rotation_s = 0.0, truncation_s = 0.0;
  virtual void seed(const uint64_t* keys, const double* coeffs, uint64_t n) = 0;
  virtual void apply(const uint64_t* gens, const int64_t* pidx, const double* angles, uint64_t count,
                     uint64_t* n_after, uint64_t* mono_after, pq_surrogate_gate_stats* stats) = 0;
  virtual uint64_t len() const = 0;
  virtual uint64_t arena_nodes() const = 0;
  virtual Model* compile(const uint64_t* diag, uint64_t nd, uint32_t n_params) = 0;
};

template <int W>
struct KeyHash {
  size_t operator()(const Key<W>& k) const {
    uint64_t h = 0x9E3779B97F4A7C15ULL;
    for (int i = 0; i < W; ++i) {
      h ^= k.w[i] + 0x9E3779B97F4A7C15ULL + (h << 6) + (h >> 2);
      h *= 0xBF58476D1CE4E5B9ULL;
    }
    return (size_t)(h ^ (h >> 31));
  }
};
template <int W>
struct KeyEq {
  bool operator()(const Key<W>& a, const Key<W>& b) const { return key_eq<W>(a, b); }
};

// The live terms: (key, coefficient node) pairs in a dense array, walked in
// order, with an open-addressing index from key to position (linear probing,
// at most half full).  Removal compacts the array and rebuilds the index.
template <int W>
struct TermTable {
  using Entry = std::pair<Key<W>, uint32_t>;
  std::vector<Entry> rows;
  std::vector<uint32_t> slot;  // position + 1, 0 = empty
  uint64_t mask = 0;

  size_t size() const { return rows.size(); }
  bool empty() const { return rows.empty(); }
  typename std::vector<Entry>::iterator begin() { return rows.begin(); }
  typename std::vector<Entry>::iterator end() { return rows.end(); }
  typename std::vector<Entry>::const_iterator begin() const { return rows.begin(); }
  typename std::vector<Entry>::const_iterator end() const { return rows.end(); }

  uint64_t home(const Key<W>& k) const { return (uint64_t)KeyHash<W>()(k) & mask; }

  void rebuild(size_t want) {
    size_t cap = 16;
    while (cap < 2 * want) cap <<= 1;
    slot.assign(cap, 0);
    mask = cap - 1;
    for (size_t i = 0; i < rows.size(); ++i) {
      uint64_t h = home(rows[i].first);
      while (slot[h]) h = (h + 1) & mask;
      slot[h] = (uint32_t)(i + 1);
    }
  }

  uint32_t* find(const Key<W>& k) {
    if (slot.empty()) return nullptr;
    for (uint64_t h = home(k);; h = (h + 1) & mask) {
      const uint32_t s = slot[h];
      if (!s) return nullptr;
      if (key_eq<W>(rows[s - 1].first, k)) return &rows[s - 1].second;
    }
  }

  // k must not be present
  void emplace(const Key<W>& k, uint32_t v) {
    if (rows.size() >= (size_t)0xFFFFFFFEu) fail(PQ_ERR_MEMORY, "surrogate: more than 2^32 - 2 live terms");
    rows.emplace_back(k, v);
    if (2 * rows.size() > slot.size()) {
      rebuild(rows.size());
      return;
    }
    uint64_t h = home(k);
    while (slot[h]) h = (h + 1) & mask;
    slot[h] = (uint32_t)rows.size();
  }

  template <typename Keep>
  void retain(Keep keep) {
    size_t j = 0;
    for (size_t i = 0; i < rows.size(); ++i)
      if (keep(rows[i].first, rows[i].second)) rows[j++] = rows[i];
    if (j == rows.size()) return;
    rows.resize(j);
    rebuild(j);
  }
};

template <int KIND, int W>
struct Builder : BuilderBase {
  Arena A;
  TermTable<W> terms;
  Tableau<KIND, W> frame;
  size_t compacted_at = 0;
  bool garbage = false;  // a term or a subtree was dropped since the last compaction

  explicit Builder(int nu) {
    kind = KIND;
    n_units = nu;
    n_words = W;
    frame.init(nu);
  }

  Key<W> key_from(const uint64_t* w) const {
    Key<W> k;
    for (int i = 0; i < W; ++i) k.w[i] = w[i];
    return k;
  }

  void seed(const uint64_t* keys, const double* coeffs, uint64_t n) override {
    for (uint64_t r = 0; r < n; ++r) {
      const Key<W> k = key_from(keys + r * W);
      const uint32_t leaf = A.scalar(coeffs[r]);
      uint32_t* v = terms.find(k);
      if (!v)
        terms.emplace(k, leaf);
      else
        *v = A.add_opt(*v, leaf);
    }
  }

  u128 monomials() const {
    u128 m = 0;
    for (const auto& kv : terms)
      if (kv.second != kNone) m = sat_add(m, A.n[kv.second].count);
    return m;
  }

  // One rotation (partitioned_termsum.rs:276-350 with SymbolicCoeff's rule,
  // symcoeff.rs:276-305).
  void gate(const Key<W>& gin, int64_t pidx, double angle) {
    if (terms.empty()) return;  // partitioned_termsum.rs:282
    if (pidx < 0 && std::fabs(std::cos(angle)) < kEps) {
      // a numeric Clifford angle folds into the frame; the surrogate's emit
      // gate has no weight bound, so it always may (partitioned_termsum.rs:287-296)
      frame.compose_rotation(gin, angle);
      return;
    }
    Key<W> g;
    const double sign = frame.conjugate_generator(gin, g);
    if (pidx < 0 && std::fabs(std::sin(angle)) < kEps) {  // phase_only_scale (symcoeff.rs:781)
      const double c_t = std::cos(angle);
      for (auto& kv : terms)
        if (anticommute<KIND, W>(g, kv.first)) {
          ++visited;
          if (kv.second != kNone) kv.second = A.scale(c_t, kv.second);
        }
      return;
    }
    const double c_t = pidx < 0 ? std::cos(angle) : 0.0, s_t = pidx < 0 ? std::sin(angle) : 0.0;
    std::vector<std::pair<Key<W>, uint32_t>> branches;
    for (auto& kv : terms) {
      if (!anticommute<KIND, W>(g, kv.first)) continue;
      const uint32_t old = kv.second;
      Key<W> child;
      for (int i = 0; i < W; ++i) child.w[i] = g.w[i] ^ kv.first.w[i];
      // -Im(i * phase) of the signed generator: the branch's real factor
      const double ph = sign * neg_im(product_exp<KIND, W>(g, kv.first));
      uint32_t br = kNone;
      if (old != kNone) {
        if (pidx >= 0) {
          kv.second = A.trig(kTapeCos, (uint32_t)pidx, old);
          br = A.signed_(ph, A.trig(kTapeSin, (uint32_t)pidx, old));
        } else {
          kv.second = A.scale(c_t, old);
          br = A.scale(s_t * ph, old);
        }
      }
      branches.emplace_back(child, br);
    }
    visited += branches.size();
    emitted += branches.size();
    for (const auto& b : branches) {  // absorb: existing + incoming
      uint32_t* v = terms.find(b.first);
      if (!v) {
        terms.emplace(b.first, b.second);
      } else {
        *v = A.add_opt(*v, b.second);
        ++hits;
      }
    }
  }

  // apply_truncation (engine.rs:20-57).
  void truncate() {
    const bool lossy = terms.size() >= (cfg.has_min ? cfg.min_terms : 0);
    if (cfg.simplify) {
      garbage = true;  // every coefficient is rebuilt
      Simplifier S{A, {}, {}, {}};
      std::vector<uint32_t> roots;
      roots.reserve(terms.size());
      for (const auto& kv : terms) roots.push_back(kv.second);
      S.expand_all(roots);
      for (auto& kv : terms) kv.second = S.simplify(kv.second);
    }
    if (lossy) {
      if (cfg.has_f || cfg.has_c) {
        Pruner P{A, cfg.has_f, cfg.has_c, cfg.has_f ? cfg.freq : kU32Max, cfg.has_c ? cfg.coeff : 0.0, {}, {}};
        for (auto& kv : terms)
          if (kv.second != kNone) {
            const uint32_t r = P.go(kv.second, 0, 0);
            garbage = garbage || r != kv.second;
            kv.second = r;
          }
      }
      const size_t before = terms.size();
      terms.retain([&](const Key<W>& k, uint32_t id) {
        return id != kNone && (!cfg.has_w || (uint32_t)weight<KIND, W>(k, n_units) <= cfg.weight);
      });
      garbage = garbage || terms.size() != before;
    }
    maybe_compact();
  }

  // Drops nodes no live term reaches.  Operands are always older than the
  // node using them, so one backward sweep marks and one forward sweep
  // renumbers, keeping the relative order.
  void maybe_compact() {
    const size_t N = A.n.size();
    // only truncation makes garbage: without it every node stays reachable
    if (!garbage || N < compact_min || N < 2 * compacted_at) return;
    garbage = false;
    std::vector<uint8_t> live(N, 0);
    for (const auto& kv : terms)
      if (kv.second != kNone) live[kv.second] = 1;
    for (size_t i = N; i-- > 0;) {
      if (!live[i]) continue;
      const Node& nd = A.n[i];
      if (nd.kind != kTapeScalar) live[nd.a] = 1;
      if (nd.kind == kTapeAdd) live[nd.b] = 1;
    }
    std::vector<uint32_t> nid(N, kNone);
    std::vector<Node> kept;
    for (size_t i = 0; i < N; ++i) {
      if (!live[i]) continue;
      Node nd = A.n[i];
      if (nd.kind != kTapeScalar) nd.a = nid[nd.a];
      if (nd.kind == kTapeAdd) nd.b = nid[nd.b];
      nid[i] = (uint32_t)kept.size();
      kept.push_back(nd);
    }
    A.n.swap(kept);
    for (auto& kv : terms)
      if (kv.second != kNone) kv.second = nid[kv.second];
    compacted_at = A.n.size();
  }

  void apply(const uint64_t* gens, const int64_t* pidx, const double* angles, uint64_t count, uint64_t* n_after,
             uint64_t* mono_after, pq_surrogate_gate_stats* stats) override {
    using clk = std::chrono::steady_clock;
    for (uint64_t g = 0; g < count; ++g) {
      const auto t0 = clk::now();
      gate(key_from(gens + g * W), pidx ? pidx[g] : -1, angles ? angles[g] : 0.0);
      const auto t1 = clk::now();
      const uint64_t rotated = terms.size();
      truncate();
      const auto t2 = clk::now();
      const double r_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
      const double t_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
      rotation_s += r_ms * 1e-3;
      truncation_s += t_ms * 1e-3;
      if (n_after) n_after[g] = terms.size();
      if (mono_after || stats) {
        const u128 m = monomials();
        if (mono_after) {
          mono_after[2 * g] = (uint64_t)m;
          mono_after[2 * g + 1] = (uint64_t)(m >> 64);
        }
        if (stats) {
          stats[g].terms_rotated = rotated;
          stats[g].terms = terms.size();
          stats[g].monomials_lo = (uint64_t)m;
          stats[g].monomials_hi = (uint64_t)(m >> 64);
          stats[g].rotation_ms = r_ms;
          stats[g].truncation_ms = t_ms;
        }
      }
    }
  }

  uint64_t len() const override { return terms.size(); }
  uint64_t arena_nodes() const override { return A.n.size(); }

  Key<W> diag_mask(const uint64_t* diag, uint64_t nd) const {
    Key<W> dm{};
    for (int u = 0; u < n_units && u < 32 * W; ++u) {
      const uint64_t bit = (uint64_t)(u / 64) < nd ? (diag[u / 64] >> (u % 64)) & 1 : 0;
      if (!bit) continue;
      const int pos = KIND == kPauli ? 2 * u + 1 : 2 * u;
      dm.w[pos / 64] |= 1ULL << (pos % 64);
    }
    return dm;
  }

  // compile (engine.rs:60-107) + compile_batch (symcoeff.rs:176-266).
  Model* compile(const uint64_t* diag, uint64_t nd, uint32_t n_params) override {
    std::vector<std::pair<Key<W>, uint32_t>> rows(terms.begin(), terms.end());
    std::sort(rows.begin(), rows.end(),
              [](const std::pair<Key<W>, uint32_t>& x, const std::pair<Key<W>, uint32_t>& y) {
                return key_less<W>(x.first, y.first);
              });
    const Key<W> dm = diag_mask(diag, nd);
    std::unique_ptr<Model> m(new Model);
    m->n_params = n_params;
    std::vector<uint32_t> at(A.n.size(), kNone);  // arena node -> compile index
    std::vector<std::pair<uint32_t, bool>> stack;
    for (const auto& r : rows) {
      Key<W> img;
      const double sgn = frame.conjugate(r.first, img);
      const double ov = sgn * basis_trace<KIND, W>(img, dm, n_units);
      if (!(std::fabs(ov) > kOverlapEps)) continue;
      m->overlaps.push_back(ov);
      if (r.second == kNone) {
        m->roots.push_back(kTapeEmpty);
        continue;
      }
      // post-order walk, each shared node emitted once
      if (at[r.second] == kNone) {
        stack.emplace_back(r.second, false);
        while (!stack.empty()) {
          const auto [id, exit] = stack.back();
          stack.pop_back();
          if (at[id] != kNone) continue;
          const Node& nd = A.n[id];
          if (!exit) {
            stack.emplace_back(id, true);
            if (nd.kind == kTapeAdd && at[nd.b] == kNone) stack.emplace_back(nd.b, false);
            if (nd.kind != kTapeScalar && at[nd.a] == kNone) stack.emplace_back(nd.a, false);
            continue;
          }
          TapeOp op;
          op.val = nd.val;
          op.kind = nd.kind;
          op.p = nd.p;
          op.a = nd.kind == kTapeScalar ? 0 : at[nd.a];
          op.b = nd.kind == kTapeAdd ? at[nd.b] : 0;
          if (m->tape.size() >= (size_t)kTapeEmpty) fail(PQ_ERR_MEMORY, "surrogate tape beyond 2^32 - 1 ops");
          at[id] = (uint32_t)m->tape.size();
          m->tape.push_back(op);
        }
      }
      m->roots.push_back(at[r.second]);
    }
    for (const auto& op : m->tape)
      if ((op.kind == kTapeCos || op.kind == kTapeSin) && op.p >= n_params)
        fail(PQ_ERR_ARG, "a gate's parameter index is beyond n_params");
    m->finalize();
    return m.release();
  }
};

template <int KIND>
BuilderBase* make_builder(int n_units) {
  const int bits = 2 * n_units;
  if (bits <= 64) return new Builder<KIND, 1>(n_units);
  if (bits <= 128) return new Builder<KIND, 2>(n_units);
  if (bits <= 256) return new Builder<KIND, 4>(n_units);
  return new Builder<KIND, 8>(n_units);
}

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return PQ_OK;
  } catch (const Failure& e) {
    set_last_error(e.msg);
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return PQ_ERR_MEMORY;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return PQ_ERR_DEVICE;
  }
}

}  // namespace
}  // namespace sur
}  // namespace pq

struct pq_surrogate {
  std::unique_ptr<pq::sur::BuilderBase> impl;
};
struct pq_model {
  std::unique_ptr<pq::sur::Model> impl;
};

#define SUR_NEED(cond, msg)                                \
  do {                                                     \
    if (!(cond)) pq::sur::fail(PQ_ERR_ARG, (msg));         \
  } while (0)

extern "C" {

int pq_surrogate_create(uint32_t basis_kind, uint32_t n_units, pq_surrogate** out) {
  return pq::sur::guarded([&] {
    SUR_NEED(out, "null output");
    *out = nullptr;
    SUR_NEED(basis_kind <= 1, "basis_kind must be 0 (Pauli) or 1 (Majorana)");
    SUR_NEED(n_units >= 1 && n_units <= 256, "n_units must be in [1, 256]");
    std::unique_ptr<pq_surrogate> s(new pq_surrogate);
    s->impl.reset(basis_kind == 0 ? pq::sur::make_builder<pq::kPauli>((int)n_units)
                                  : pq::sur::make_builder<pq::kMajorana>((int)n_units));
    *out = s.release();
  });
}

void pq_surrogate_destroy(pq_surrogate* s) { delete s; }

int pq_surrogate_width(const pq_surrogate* s, uint32_t* n_words) {
  return pq::sur::guarded([&] {
    SUR_NEED(s && n_words, "null argument");
    *n_words = (uint32_t)s->impl->n_words;
  });
}

int pq_surrogate_set_truncation(pq_surrogate* s, int has_frequency, uint32_t frequency, int has_coefficient,
                                double coefficient, int has_weight, uint32_t weight, int has_min_terms,
                                uint64_t min_terms, int simplify) {
  return pq::sur::guarded([&] {
    SUR_NEED(s, "null surrogate");
    SUR_NEED(!has_coefficient || !std::isnan(coefficient), "coefficient cutoff is NaN");
    pq::sur::Cfg& c = s->impl->cfg;
    c.has_f = has_frequency != 0;
    c.freq = frequency;
    c.has_c = has_coefficient != 0;
    c.coeff = coefficient;
    c.has_w = has_weight != 0;
    c.weight = weight;
    c.has_min = has_min_terms != 0;
    c.min_terms = min_terms;
    c.simplify = simplify != 0;
  });
}

int pq_surrogate_set_compaction(pq_surrogate* s, uint64_t min_nodes) {
  return pq::sur::guarded([&] {
    SUR_NEED(s, "null surrogate");
    s->impl->compact_min = min_nodes;
  });
}

int pq_surrogate_seed(pq_surrogate* s, const uint64_t* keys, const double* coeffs, uint64_t n) {
  return pq::sur::guarded([&] {
    SUR_NEED(s, "null surrogate");
    SUR_NEED(n == 0 || (keys && coeffs), "null keys or coefficients");
    s->impl->seed(keys, coeffs, n);
  });
}

int pq_surrogate_apply(pq_surrogate* s, const uint64_t* generators, const int64_t* param_index, const double* angles,
                       uint64_t n_gates, uint64_t* n_terms_after, uint64_t* monomials_after) {
  return pq::sur::guarded([&] {
    SUR_NEED(s, "null surrogate");
    SUR_NEED(n_gates == 0 || (generators && param_index && angles), "null gate arrays");
    for (uint64_t g = 0; g < n_gates; ++g)
      SUR_NEED(param_index[g] < (int64_t)0xFFFFFFFFu, "parameter index beyond 2^32 - 2");
    s->impl->apply(generators, param_index, angles, n_gates, n_terms_after, monomials_after, nullptr);
  });
}

int pq_surrogate_apply_logged(pq_surrogate* s, const uint64_t* generators, const int64_t* param_index,
                              const double* angles, uint64_t n_gates, pq_surrogate_gate_stats* stats) {
  return pq::sur::guarded([&] {
    SUR_NEED(s, "null surrogate");
    SUR_NEED(n_gates == 0 || (generators && param_index && angles && stats), "null gate arrays");
    for (uint64_t g = 0; g < n_gates; ++g)
      SUR_NEED(param_index[g] < (int64_t)0xFFFFFFFFu, "parameter index beyond 2^32 - 2");
    s->impl->apply(generators, param_index, angles, n_gates, nullptr, nullptr, stats);
  });
}

int pq_surrogate_counters(const pq_surrogate* s, uint64_t* visited, uint64_t* emitted, uint64_t* hits,
                          double* rotation_s, double* truncation_s, uint64_t* nodes) {
  return pq::sur::guarded([&] {
    SUR_NEED(s, "null surrogate");
    if (nodes) *nodes = s->impl->arena_nodes();
    if (visited) *visited = s->impl->visited;
    if (emitted) *emitted = s->impl->emitted;
    if (hits) *hits = s->impl->hits;
    if (rotation_s) *rotation_s = s->impl->rotation_s;
    if (truncation_s) *truncation_s = s->impl->truncation_s;
  });
}

int pq_surrogate_len(const pq_surrogate* s, uint64_t* n) {
  return pq::sur::guarded([&] {
    SUR_NEED(s && n, "null argument");
    *n = s->impl->len();
  });
}

int pq_surrogate_compile(pq_surrogate* s, const uint64_t* diag_state, uint64_t n_diag_words, uint32_t n_params,
                         pq_model** out) {
  return pq::sur::guarded([&] {
    SUR_NEED(s && out, "null argument");
    *out = nullptr;
    SUR_NEED(n_diag_words == 0 || diag_state, "null diag_state");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c == 0) pq::sur::fail(PQ_ERR_DEVICE, "no HIP device available");
    std::unique_ptr<pq_model> m(new pq_model);
    m->impl.reset(s->impl->compile(diag_state, n_diag_words, n_params));
    *out = m.release();
  });
}

int pq_model_create(uint32_t n_params, uint64_t n_ops, const uint8_t* kinds, const uint32_t* params, const uint64_t* a,
                    const uint64_t* b, const double* vals, uint64_t n_terms, const double* overlaps,
                    const uint64_t* roots, pq_model** out) {
  return pq::sur::guarded([&] {
    SUR_NEED(out, "null output");
    *out = nullptr;
    SUR_NEED(n_ops == 0 || (kinds && params && a && b && vals), "null tape arrays");
    SUR_NEED(n_terms == 0 || (overlaps && roots), "null term arrays");
    SUR_NEED(n_ops < (uint64_t)pq::kTapeEmpty, "tape beyond 2^32 - 1 ops");
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess || c == 0) pq::sur::fail(PQ_ERR_DEVICE, "no HIP device available");
    std::unique_ptr<pq::sur::Model> m(new pq::sur::Model);
    m->n_params = n_params;
    m->tape.resize(n_ops);
    for (uint64_t i = 0; i < n_ops; ++i) {
      SUR_NEED(kinds[i] <= pq::kTapeSin, "unknown tape op kind");
      pq::TapeOp& op = m->tape[i];
      op.kind = kinds[i];
      op.val = vals[i];
      op.p = params[i];
      const bool unary = op.kind != pq::kTapeScalar;
      SUR_NEED(!unary || a[i] < i, "tape operand does not precede its op");
      SUR_NEED(op.kind != pq::kTapeAdd || b[i] < i, "tape operand does not precede its op");
      op.a = unary ? (uint32_t)a[i] : 0;
      op.b = op.kind == pq::kTapeAdd ? (uint32_t)b[i] : 0;
    }
    m->overlaps.assign(overlaps, overlaps + n_terms);
    m->roots.resize(n_terms);
    for (uint64_t t = 0; t < n_terms; ++t) {
      SUR_NEED(roots[t] == UINT64_MAX || roots[t] < n_ops, "term root beyond the tape");
      m->roots[t] = roots[t] == UINT64_MAX ? pq::kTapeEmpty : (uint32_t)roots[t];
    }
    m->finalize();
    std::unique_ptr<pq_model> h(new pq_model);
    h->impl = std::move(m);
    *out = h.release();
  });
}

void pq_model_destroy(pq_model* m) { delete m; }

int pq_model_info(const pq_model* m, uint64_t* n_terms, uint64_t* n_ops, uint64_t* n_levels, uint64_t* monomials_lo,
                  uint64_t* monomials_hi, uint32_t* n_params) {
  return pq::sur::guarded([&] {
    SUR_NEED(m, "null model");
    const pq::sur::Model& x = *m->impl;
    if (n_terms) *n_terms = x.roots.size();
    if (n_ops) *n_ops = x.tape.size();
    if (n_levels) *n_levels = x.level_off.size() - 1;
    if (monomials_lo) *monomials_lo = (uint64_t)x.monomials;
    if (monomials_hi) *monomials_hi = (uint64_t)(x.monomials >> 64);
    if (n_params) *n_params = x.n_params;
  });
}

int pq_model_export(const pq_model* m, uint8_t* kinds, uint32_t* params, uint64_t* a, uint64_t* b, double* vals,
                    double* overlaps, uint64_t* roots) {
  return pq::sur::guarded([&] {
    SUR_NEED(m, "null model");
    const pq::sur::Model& x = *m->impl;
    SUR_NEED(x.tape.empty() || (kinds && params && a && b && vals), "null tape arrays");
    SUR_NEED(x.roots.empty() || (overlaps && roots), "null term arrays");
    for (size_t i = 0; i < x.tape.size(); ++i) {
      const pq::TapeOp& op = x.tape[i];
      kinds[i] = (uint8_t)op.kind;
      params[i] = op.p;
      a[i] = op.a;
      b[i] = op.b;
      vals[i] = op.val;
    }
    for (size_t t = 0; t < x.roots.size(); ++t) {
      overlaps[t] = x.overlaps[t];
      roots[t] = x.roots[t] == pq::kTapeEmpty ? UINT64_MAX : x.roots[t];
    }
  });
}

int pq_model_evaluate(pq_model* m, const double* params, uint64_t n_sets, uint64_t stride, double* out) {
  return pq::sur::guarded([&] {
    SUR_NEED(m, "null model");
    SUR_NEED(n_sets == 0 || out, "null output");
    SUR_NEED(stride >= m->impl->n_params, "parameter rows shorter than n_params");
    SUR_NEED(n_sets == 0 || m->impl->n_params == 0 || params, "null parameters");
    m->impl->evaluate(params, n_sets, stride, out);
  });
}

int pq_model_set_batch_limit(pq_model* m, uint64_t max_sets) {
  return pq::sur::guarded([&] {
    SUR_NEED(m, "null model");
    m->impl->max_sets = max_sets;
  });
}

int pq_model_set_profiling(pq_model* m, int on) {
  return pq::sur::guarded([&] {
    SUR_NEED(m, "null model");
    m->impl->profiling = on != 0;
  });
}

int pq_model_timing(const pq_model* m, double* tape_ms, double* sum_ms, double* tape_bytes, uint64_t* launches) {
  return pq::sur::guarded([&] {
    SUR_NEED(m, "null model");
    if (tape_ms) *tape_ms = m->impl->t_tape;
    if (sum_ms) *sum_ms = m->impl->t_sum;
    if (tape_bytes) *tape_bytes = m->impl->bytes;
    if (launches) *launches = m->impl->launches;
  });
}

}  // extern "C"

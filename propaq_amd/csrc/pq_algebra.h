// pq_algebra.h -- Pauli / Majorana algebra on packed W-word keys, shared by
// the host orchestration (tableau, seeding) and the gfx950 kernels.
//
// Key layout (crates/core/src/algebra/strings.rs:30-41): unit u owns bits
// (2u, 2u+1) of a little-endian array of W u64 words.  Pauli: bit 2q = X,
// bit 2q+1 = Z (crates/pauli/src/algebra.rs:145-160).  Majorana: bit k is
// gamma_k, site s = (gamma_2s, gamma_2s+1) (crates/majorana/src/algebra.rs:220).
//
// Phases are carried as an exponent e in Z/4 (phase = i^e).  Every product of
// two basis strings has a phase in {1, i, -1, -i}, so the integer form is
// exact and needs no complex arithmetic.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define PQ_HD __host__ __device__ __forceinline__
#else
#define PQ_HD inline
#endif

namespace pq {

constexpr uint64_t kEven = 0x5555555555555555ULL;
enum Kind : int { kPauli = 0, kMajorana = 1 };

PQ_HD int popc64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __popcll(x);
#else
  return __builtin_popcountll(x);
#endif
}

template <int W>
struct Key {
  uint64_t w[W];
};

template <int W>
PQ_HD bool key_eq(const Key<W>& a, const Key<W>& b) {
  bool eq = true;
#pragma unroll
  for (int i = 0; i < W; ++i) eq = eq && (a.w[i] == b.w[i]);
  return eq;
}

// Store order: word W-1 is the most significant.
template <int W>
PQ_HD bool key_less(const Key<W>& a, const Key<W>& b) {
#pragma unroll
  for (int i = W - 1; i >= 0; --i) {
    if (a.w[i] != b.w[i]) return a.w[i] < b.w[i];
  }
  return false;
}

template <int W>
PQ_HD int key_count(const Key<W>& k) {
  int n = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) n += popc64(k.w[i]);
  return n;
}

// Number of Y sites of a Pauli key (crates/pauli/src/algebra.rs:35).
template <int W>
PQ_HD int pauli_y(const Key<W>& k) {
  int n = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) n += popc64(k.w[i] & (k.w[i] >> 1) & kEven);
  return n;
}

// Support = Pauli weight (strings.rs:140).
template <int W>
PQ_HD int support(const Key<W>& k) {
  int n = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) n += popc64((k.w[i] | (k.w[i] >> 1)) & kEven);
  return n;
}

// Gather the even bits of x into its low 32 bits.
PQ_HD uint64_t gather_even(uint64_t x) {
  x &= kEven;
  x = (x | (x >> 1)) & 0x3333333333333333ULL;
  x = (x | (x >> 2)) & 0x0f0f0f0f0f0f0f0fULL;
  x = (x | (x >> 4)) & 0x00ff00ff00ff00ffULL;
  x = (x | (x >> 8)) & 0x0000ffff0000ffffULL;
  return (x | (x >> 16)) & 0x00000000ffffffffULL;
}

// Inclusive prefix XOR of the low 32 bits (bit s = XOR of bits 0..s).
PQ_HD uint32_t prefix_xor32(uint32_t x) {
  x ^= x << 1;
  x ^= x << 2;
  x ^= x << 4;
  x ^= x << 8;
  x ^= x << 16;
  return x;
}

// Jordan-Wigner qubit weight of a Majorana key over n_sites sites
// (crates/majorana/src/algebra.rs:85-123).  Each word holds 32 sites; the
// prefix parity of unpaired sites is carried across words.
template <int W>
PQ_HD int jw_weight(const Key<W>& k, int n_sites) {
  int total_single = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    uint32_t x = (uint32_t)gather_even(k.w[i]);
    uint32_t y = (uint32_t)gather_even(k.w[i] >> 1);
    total_single += popc64((uint64_t)(x ^ y));
  }
  const uint32_t flip = (total_single & 1) ? 0xFFFFFFFFu : 0u;
  uint32_t carry = 0;
  int w = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    uint32_t x = (uint32_t)gather_even(k.w[i]);
    uint32_t y = (uint32_t)gather_even(k.w[i] >> 1);
    uint32_t single = x ^ y, occ = x | y;
    int lo = 32 * i;
    uint32_t live = lo >= n_sites ? 0u : (n_sites - lo >= 32 ? 0xFFFFFFFFu : ((1u << (n_sites - lo)) - 1u));
    uint32_t p = prefix_xor32(single) ^ carry;
    carry = (p >> 31) ? 0xFFFFFFFFu : 0u;
    uint32_t str = (p & live) ^ (flip & live);
    w += popc64((uint64_t)(single | (occ ^ str)));
  }
  return w;
}

template <int KIND, int W>
PQ_HD int weight(const Key<W>& k, int n_units) {
  if (KIND == kPauli) return support(k);
  return jw_weight(k, n_units);
}

// Bit b of the result = parity of the bits of x strictly above b.
PQ_HD uint64_t above_parity64(uint64_t x) {
  uint64_t t = x >> 1;
  t ^= t >> 1;
  t ^= t >> 2;
  t ^= t >> 4;
  t ^= t >> 8;
  t ^= t >> 16;
  t ^= t >> 32;
  return t;
}

// Mask M with bit b set iff an odd number of bits of g lie strictly above b
// (over all W words).  parity(popcount(P & M)) is the resorting parity of
// (g, P), crates/majorana/src/monomial.rs:696.
template <int W>
PQ_HD Key<W> above_parity_mask(const Key<W>& g) {
  Key<W> m;
  uint64_t hi = 0;  // all-ones if the words above hold an odd count
#pragma unroll
  for (int i = W - 1; i >= 0; --i) {
    m.w[i] = above_parity64(g.w[i]) ^ hi;
    if (popc64(g.w[i]) & 1) hi = ~hi;
  }
  return m;
}

PQ_HD int herm_exp(int len) { return ((len & 3) <= 1) ? 0 : 1; }

// Phase exponent of a * b (a multiplied on the left), with the product key
// a ^ b.  Pauli: crates/pauli/src/algebra.rs:109; Majorana:
// crates/majorana/src/algebra.rs:180.
template <int KIND, int W>
PQ_HD int product_exp(const Key<W>& a, const Key<W>& b) {
  Key<W> o;
#pragma unroll
  for (int i = 0; i < W; ++i) o.w[i] = a.w[i] ^ b.w[i];
  int p;
  if (KIND == kPauli) {
    int zx = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) zx += popc64((a.w[i] >> 1) & b.w[i] & kEven);
    p = pauli_y(a) + pauli_y(b) - pauli_y(o) + 2 * zx;
  } else {
    Key<W> m = above_parity_mask(a);
    int par = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) par += popc64(b.w[i] & m.w[i]);
    p = herm_exp(key_count(a)) + herm_exp(key_count(b)) - herm_exp(key_count(o)) + 2 * (par & 1);
  }
  return p & 3;
}

// a anticommutes with b?  Pauli: parity(b & pair_swap(a)); Majorana:
// |a & b| + |a||b| odd (crates/majorana/src/algebra.rs:162).
template <int KIND, int W>
PQ_HD bool anticommute(const Key<W>& a, const Key<W>& b) {
  if (KIND == kPauli) {
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      uint64_t s = ((a.w[i] & kEven) << 1) | ((a.w[i] >> 1) & kEven);
      acc ^= s & b.w[i];
    }
    return popc64(acc) & 1;
  }
  int ov = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) ov += popc64(a.w[i] & b.w[i]);
  return ((ov + key_count(a) * key_count(b)) & 1) != 0;
}

// -Im(i^e): the real factor the sine branch carries (coeff.rs:165).
PQ_HD double neg_im(int e) { return e == 1 ? -1.0 : (e == 3 ? 1.0 : 0.0); }

// <s| basis(k) |s> on a computational basis state whose key-bit mask is dm
// (Pauli: the Z bit 2u+1 of every |1> qubit; Majorana: bit 2u of every
// occupied site).  Pauli: zero with any X/Y factor, else (-1)^|Z & s|.
// Majorana: zero unless every site holds both or neither gamma; a paired site
// is i Z_s under Jordan-Wigner, so the value is the normalisation's sign times
// (-1)^(empty paired sites) (crates/pauli/src/algebra.rs:123, crates/majorana/src/algebra.rs:200).
template <int KIND, int W>
PQ_HD double basis_trace(const Key<W>& k, const Key<W>& dm, int n_units) {
  if (KIND == kPauli) {
    uint64_t x = 0, par = 0;
#pragma unroll
    for (int i = 0; i < W; ++i) {
      x |= k.w[i] & kEven;
      par += popc64(k.w[i] & dm.w[i]);
    }
    if (x) return 0.0;
    return (par & 1) ? -1.0 : 1.0;
  }
  int paired = 0, occ = 0;
#pragma unroll
  for (int i = 0; i < W; ++i) {
    const int lo = 32 * i;
    const uint64_t live = lo >= n_units ? 0ULL : (n_units - lo >= 32 ? kEven : (kEven & ((1ULL << (2 * (n_units - lo))) - 1)));
    if ((k.w[i] ^ (k.w[i] >> 1)) & live) return 0.0;
    const uint64_t pr = k.w[i] & (k.w[i] >> 1) & live;
    paired += popc64(pr);
    occ += popc64(pr & dm.w[i]);
  }
  const int empties = paired - occ;
  const double prod = (empties & 1) ? -1.0 : 1.0;
  return (((paired / 2) & 1) ? -1.0 : 1.0) * prod;
}

}  // namespace pq

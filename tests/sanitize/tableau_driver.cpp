// Sanitizer driver for the engine's host-side frame code (test
// infrastructure): pq_tableau.h and pq_algebra.h compiled with g++ under
// -fsanitize=address,undefined by tests/test_sanitizers.py.  Composes random
// Clifford rotations into a Tableau and checks, for random keys, that the
// readout conjugation of a product is the product of the conjugations up to
// the phase the algebra fixes, that conjugation is injective on the keys
// tried, and that composing a quarter turn four times gives the identity.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <set>
#include <vector>

#include "pq_tableau.h"

using namespace pq;

static uint64_t rs = 0x2545F4914F6CDD1DULL;
static uint64_t rnd() {
  rs ^= rs << 13;
  rs ^= rs >> 7;
  rs ^= rs << 17;
  return rs;
}

template <int W>
static Key<W> random_key(int n_units, int max_bits) {
  Key<W> k{};
  const int nb = 1 + (int)(rnd() % (uint64_t)max_bits);
  for (int i = 0; i < nb; ++i) {
    const int b = (int)(rnd() % (uint64_t)(2 * n_units));
    k.w[b >> 6] |= 1ULL << (b & 63);
  }
  return k;
}

template <int KIND, int W>
static int check(int n_units) {
  int bad = 0;
  Tableau<KIND, W> t;
  t.init(n_units);
  for (int g = 0; g < 40; ++g) {
    const Key<W> gen = random_key<W>(n_units, KIND == kPauli ? 2 : 4);
    t.compose_rotation(gen, (rnd() & 1) ? M_PI / 2 : -M_PI / 2);
  }
  std::set<std::vector<uint64_t>> images;
  for (int i = 0; i < 200; ++i) {
    const Key<W> a = random_key<W>(n_units, 4), b = random_key<W>(n_units, 4);
    Key<W> ab, ia, ib, iab, prod;
    for (int w = 0; w < W; ++w) ab.w[w] = a.w[w] ^ b.w[w];
    const double sa = t.conjugate(a, ia), sb = t.conjugate(b, ib), sab = t.conjugate(ab, iab);
    for (int w = 0; w < W; ++w) prod.w[w] = ia.w[w] ^ ib.w[w];
    if (!key_eq<W>(prod, iab)) bad = 1;  // conjugation maps products to products
    if (std::fabs(sa) != 1.0 || std::fabs(sb) != 1.0 || std::fabs(sab) != 1.0) bad = 1;
    // commutation is preserved
    if (anticommute<KIND, W>(a, b) != anticommute<KIND, W>(ia, ib)) bad = 1;
    images.insert(std::vector<uint64_t>(ia.w, ia.w + W));
  }
  // four quarter turns about one generator are the identity
  Tableau<KIND, W> q;
  q.init(n_units);
  const Key<W> gen = random_key<W>(n_units, 2);
  for (int i = 0; i < 4; ++i) q.compose_rotation(gen, M_PI / 2);
  q.refresh_identity();
  if (!q.identity) bad = 1;
  std::printf("kind %d W %d units %d: %zu distinct images %s\n", KIND, W, n_units, images.size(), bad ? "MISMATCH" : "ok");
  return bad;
}

int main() {
  int bad = 0;
  bad |= check<kPauli, 1>(20);
  bad |= check<kPauli, 2>(50);
  bad |= check<kPauli, 4>(100);
  bad |= check<kMajorana, 1>(24);
  bad |= check<kMajorana, 2>(60);
  return bad;
}

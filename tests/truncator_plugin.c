/* A truncation plugin for the tests, written to the plugin ABI of
 * crates/core/src/policy/native_truncator.rs:13-33.  Compiled by the tests
 * with gcc; -D options pick the rule:
 *
 *   MODE=0  keep iff coeff_magnitude >= THR (no key: the same decisions as a
 *           coefficient cutoff of THR without a weight bound)
 *   MODE=1  reads keys: keep with probability magnitude / THR, decided by a
 *           hash of the key words (a pure function of its arguments)
 *   MODE=2  layer- and weight-aware: keep iff magnitude >= THR * (1 + w) *
 *           (1 + layer_index) / (1 + n_layers)
 *
 * WITH_BATCH=1 adds propaq_truncator_keep_batch.  Every call also checks its
 * arguments: a Pauli key's weight must be its number of non-identity qubits,
 * words must be NULL exactly when the plugin does not read keys, and
 * layer_index < n_layers whenever n_layers > 0; tp_violations() counts the
 * calls that broke one of these, tp_calls() all calls, tp_max_layers() the
 * largest n_layers seen. */
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#ifndef MODE
#define MODE 0
#endif
#ifndef THR
#define THR 1e-3
#endif
#ifndef WITH_BATCH
#define WITH_BATCH 0
#endif

static uint64_t n_calls, n_violations;
static uint32_t max_layers;

uint64_t tp_calls(void) { return n_calls; }
uint64_t tp_violations(void) { return n_violations; }
uint32_t tp_max_layers(void) { return max_layers; }
void tp_reset(void) {
  n_calls = n_violations = 0;
  max_layers = 0;
}

uint32_t propaq_truncator_abi_version(void) { return 1; }
uint32_t propaq_truncator_depends(void) { return MODE == 1 ? 1u : 0u; }

static uint64_t mix(uint64_t x) { /* a 64-bit finaliser: xorshift-multiply rounds */
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}

static void check(uint32_t kind, const uint64_t *words, size_t n_words, uint32_t n_units, uint32_t weight,
                  uint32_t layer_index, uint32_t n_layers) {
  n_calls++;
  if (n_layers > max_layers) max_layers = n_layers;
  int bad = 0;
  if ((MODE == 1) != (words != NULL)) bad = 1;
  if (n_layers > 0 && layer_index >= n_layers) bad = 1;
  if (words && kind == 0) {
    uint32_t w = 0;
    for (uint32_t q = 0; q < n_units; q++) {
      size_t b = 2 * (size_t)q;
      if (b / 64 < n_words && ((words[b / 64] >> (b % 64)) & 3u)) w++;
    }
    if (w != weight) bad = 1;
  }
  n_violations += bad;
}

static int decide(const uint64_t *words, size_t n_words, uint32_t weight, double mag, uint32_t layer_index,
                  uint32_t n_layers) {
  if (MODE == 0) return mag >= THR;
  if (MODE == 1) {
    double p = mag / THR;
    if (p >= 1.0) return 1;
    uint64_t h = 0x243f6a8885a308d3ULL;
    for (size_t i = 0; i < n_words; i++) h = mix(h ^ words[i]);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0) < p;
  }
  return mag >= THR * (1.0 + weight) * (1.0 + layer_index) / (1.0 + n_layers);
}

int32_t propaq_truncator_keep(void *ctx, uint32_t basis_kind, const uint64_t *words, size_t n_words, uint32_t n_units,
                              uint32_t weight, double coeff_magnitude, uint32_t layer_index, uint32_t n_layers) {
  (void)ctx;
  check(basis_kind, words, n_words, n_units, weight, layer_index, n_layers);
  return decide(words, n_words, weight, coeff_magnitude, layer_index, n_layers);
}

#if WITH_BATCH
int32_t propaq_truncator_keep_batch(void *ctx, uint32_t basis_kind, const uint64_t *words, size_t n_words_per_term,
                                    uint32_t n_units, const uint32_t *weights, const double *coeff_magnitudes,
                                    uint32_t layer_index, uint32_t n_layers, uint8_t *out_keep, size_t n_terms) {
  (void)ctx;
  for (size_t i = 0; i < n_terms; i++) {
    const uint64_t *k = words ? words + i * n_words_per_term : NULL;
    check(basis_kind, k, n_words_per_term, n_units, weights[i], layer_index, n_layers);
    out_keep[i] = (uint8_t)decide(k, n_words_per_term, weights[i], coeff_magnitudes[i], layer_index, n_layers);
  }
  return 0;
}
#endif

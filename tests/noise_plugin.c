/* A noise plugin for the tests, written to the plugin ABI of
 * crates/core/src/policy/native_noise.rs:12-28.  Compiled by the tests with
 * gcc; -D DEPENDS=0..3 sets the bits it declares (KEY = 1, LAYER = 2):
 *
 *   factor = exp(-0.03 w)                         always
 *          * 1 / (1 + 0.1 layer_index)           if it declares LAYER
 *          * exp(-0.05 |words & MASK|)            if it declares KEY
 *
 * with MASK the bits of the first two units.  WITH_BATCH=1 adds
 * propaq_noise_factor_batch.  Every call checks its arguments: words must be
 * NULL exactly when the plugin does not declare KEY, and layer_index <
 * n_layers when n_layers > 0.  np_calls() counts calls (a batch call counts
 * its terms), np_saw_words() the calls handed a key, np_violations() the
 * calls that broke a rule, np_max_layer() the largest layer index seen. */
#include <math.h>
#include <stdatomic.h>
#include <stddef.h>
#include <stdint.h>

#ifndef DEPENDS
#define DEPENDS 0
#endif
#ifndef WITH_BATCH
#define WITH_BATCH 0
#endif
#define MASK 0xFULL

static _Atomic uint64_t g_calls, g_words, g_violations;
static _Atomic uint32_t g_max_layer;

uint64_t np_calls(void) { return atomic_load(&g_calls); }
uint64_t np_saw_words(void) { return atomic_load(&g_words); }
uint64_t np_violations(void) { return atomic_load(&g_violations); }
uint32_t np_max_layer(void) { return atomic_load(&g_max_layer); }
void np_reset(void) {
  atomic_store(&g_calls, 0);
  atomic_store(&g_words, 0);
  atomic_store(&g_violations, 0);
  atomic_store(&g_max_layer, 0);
}

uint32_t propaq_noise_abi_version(void) { return 1u; }
uint32_t propaq_noise_depends(void) { return (uint32_t)DEPENDS; }

static double factor_of(const uint64_t *words, size_t n_words, uint32_t weight, uint32_t layer_index,
                        uint32_t n_layers) {
  atomic_fetch_add(&g_calls, 1);
  if ((words != NULL) != ((DEPENDS & 1) != 0)) atomic_fetch_add(&g_violations, 1);
  if (words != NULL) atomic_fetch_add(&g_words, 1);
  if (n_layers > 0 && layer_index >= n_layers) atomic_fetch_add(&g_violations, 1);
  uint32_t seen = atomic_load(&g_max_layer);
  while (layer_index > seen && !atomic_compare_exchange_weak(&g_max_layer, &seen, layer_index)) {
  }
  double f = exp(-0.03 * (double)weight);
  if (DEPENDS & 2) f = f / (1.0 + 0.1 * (double)layer_index);
  if ((DEPENDS & 1) && words != NULL && n_words > 0) f = f * exp(-0.05 * (double)__builtin_popcountll(words[0] & MASK));
  return f;
}

double propaq_noise_factor(void *ctx, uint32_t basis_kind, const uint64_t *words, size_t n_words, uint32_t n_units,
                           uint32_t weight, uint32_t layer_index, uint32_t n_layers) {
  (void)ctx;
  (void)basis_kind;
  (void)n_units;
  return factor_of(words, n_words, weight, layer_index, n_layers);
}

#if WITH_BATCH
int32_t propaq_noise_factor_batch(void *ctx, uint32_t basis_kind, const uint64_t *words, size_t n_words,
                                  uint32_t n_units, const uint32_t *weights, uint32_t layer_index, uint32_t n_layers,
                                  double *out, size_t n_terms) {
  (void)ctx;
  (void)basis_kind;
  (void)n_units;
  for (size_t i = 0; i < n_terms; i++)
    out[i] = factor_of(words ? words + i * n_words : NULL, n_words, weights[i], layer_index, n_layers);
  return 0;
}
#endif

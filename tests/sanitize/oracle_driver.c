/* Sanitizer driver for the C oracle (test infrastructure): compiled together
 * with oracle/pq_oracle.c under -fsanitize=address,undefined by
 * tests/test_sanitizers.py.  Runs the single-partition and the partitioned
 * engines over random truncated, noisy Pauli and Majorana circuits at one,
 * two and four key words, with deferred Cliffords, and checks that both give
 * the same live counts and terms; any memory or undefined-behaviour error
 * aborts the run. */
#include "../../oracle/pq_oracle.c"

#include <stdio.h>

static uint64_t rs = 0x9E3779B97F4A7C15ULL;
static uint64_t rnd(void) {
  rs ^= rs << 13;
  rs ^= rs >> 7;
  rs ^= rs << 17;
  return rs;
}

/* a key of up to max_bits set bits among the units [lo, lo + 10) (so
 * branches meet again and the store grows), lo in the top word when W > 1 */
static void random_key(uint64_t *k, int W, int n_units, int max_bits) {
  memset(k, 0, sizeof(uint64_t) * W);
  const int span = n_units < 10 ? n_units : 10, lo = n_units - span;
  const int nb = 1 + (int)(rnd() % (uint64_t)max_bits);
  for (int i = 0; i < nb; i++) {
    const int b = 2 * lo + (int)(rnd() % (uint64_t)(2 * span));
    k[b >> 6] |= 1ULL << (b & 63);
  }
}

static int run_case(int kind, int n_units, int W, int S, int defer, int n_layers, int per_layer) {
  pqo_engine *a = pqo_create(kind, n_units, W);
  pqo_part *b = pqo_part_create(kind, n_units, W, S, S < 4 ? S : 4);
  const uint32_t maxw = kind == PQO_PAULI ? 6 : (uint32_t)n_units; /* JW weights of the window's keys run long */
  pqo_set_cutoff(a, 1, maxw, 1, 1e-4, 0, 0);
  pqo_part_set_cutoff(b, 1, maxw, 1, 1e-4, 0, 0);
  pqo_set_defer(a, defer);
  pqo_part_set_defer(b, defer);
  uint64_t k[PQO_MAXW];
  for (int i = 0; i < 3; i++) {
    random_key(k, W, n_units, 3);
    const double c = 0.5 + (double)(rnd() % 1000) / 1000.0;
    if (pqo_admits_initial(a, k, c)) {
      pqo_add(a, k, c);
      pqo_part_add(b, k, c);
    }
  }
  double table[PQO_MAXW * 32 + 1];
  for (int w = 0; w <= n_units; w++) table[w] = exp(-0.02 * w);
  int bad = 0;
  for (int l = 0; l < n_layers; l++) {
    pqo_scale_by_table(a, table, (size_t)n_units + 1);
    pqo_part_scale_by_table(b, table, (size_t)n_units + 1);
    pqo_reclaim(a, NULL);
    pqo_part_reclaim(b);
    for (int g = 0; g < per_layer; g++) {
      random_key(k, W, n_units, kind == PQO_PAULI ? 3 : 4);
      const uint64_t r = rnd() % 10;
      const double ang = r == 0 ? 1.5707963267948966 : (r == 1 ? 3.141592653589793 : (double)(rnd() % 6283) / 1000.0 - 3.1415);
      pqo_apply_rotation(a, k, ang);
      pqo_part_apply_rotation(b, k, ang);
      if (pqo_len(a) != pqo_part_len(b)) bad = 1;
    }
  }
  const size_t n = pqo_len(a);
  /* every row of every partition is the single engine's row of that key, bit for bit */
  for (int q = 0; q < b->S && !bad; q++) {
    const pqo_engine *e = b->p[q];
    for (size_t i = 0; i < e->n; i++) {
      const long r = index_find(a, &e->keys[i * W]);
      if (r < 0 || a->coeffs[r] != e->coeffs[i]) {
        bad = 1;
        break;
      }
    }
  }
  /* and both export the same rows with the frame applied */
  uint64_t *kb = malloc(sizeof(uint64_t) * W * (n + 1));
  double *cb = malloc(sizeof(double) * (n + 1));
  pqo_part_export(b, kb, cb);
  free(kb);
  free(cb);
  uint64_t diag[4] = {0x5, 0, 0, 0};
  const double ea = pqo_expectation(a, diag, 1), eb = pqo_part_expectation(b, diag, 1);
  if (fabs(ea - eb) > 1e-12) bad = 1;
  pqo_destroy(a);
  pqo_part_destroy(b);
  printf("kind %d units %d W %d S %d defer %d: %zu terms %s\n", kind, n_units, W, S, defer, n, bad ? "MISMATCH" : "ok");
  return bad;
}

int main(void) {
  int bad = 0;
  bad |= run_case(PQO_PAULI, 20, 1, 3, 1, 6, 10);
  bad |= run_case(PQO_PAULI, 40, 2, 4, 1, 6, 10);
  bad |= run_case(PQO_PAULI, 100, 4, 8, 0, 6, 10);
  bad |= run_case(PQO_MAJORANA, 24, 1, 2, 1, 6, 10);
  bad |= run_case(PQO_MAJORANA, 60, 2, 5, 1, 6, 10);
  return bad;
}

/*
 * pq_oracle.c -- FP64 CPU oracle for propaq's per-gate term propagation.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker against which the
 * MI355X engine in propaq_amd/ is compared.  It is never linked into, loaded
 * by, or called from the product path; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it.
 *
 * It restates, in plain C with a runtime word count W (1..64), the single-
 * partition semantics of the reference Rust engine:
 *   keys          crates/core/src/algebra/strings.rs:34-223
 *   Pauli algebra crates/pauli/src/algebra.rs:32-142
 *   Majorana alg. crates/majorana/src/algebra.rs:32-217,
 *                 crates/majorana/src/monomial.rs:688-722
 *   f64 rotation  crates/core/src/algebra/coeff.rs:115-209
 *   tableau       crates/core/src/algebra/tableau.rs:29-224
 *   emit cutoff   crates/core/src/engine/termsum.rs:46-197
 *   scan/absorb   crates/core/src/engine/termsum.rs:561-800
 *   gate          crates/core/src/engine/partitioned_termsum.rs:276-350
 *   noise/reclaim crates/core/src/engine/termsum.rs:321-497,
 *                 crates/core/src/engine/partitioned_termsum.rs:478-608
 *   expectation   crates/core/src/engine/partitioned_termsum.rs:676-697
 * A third algebra, the "toy" algebra of the reference's own engine unit tests
 * (crates/core/tests/unit/engine/termsum.rs:6-40), is included so those
 * known-answer tests can be restated against this oracle verbatim.
 *
 * Storage mirrors the reference's row store + open-addressed key index
 * (crates/core/src/storage/operator_index.rs:129-473) in the simplest form:
 * rows are appended in arrival order and never move during a gate.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define PQO_MAXW 64
#define EVEN 0x5555555555555555ULL

enum { PQO_PAULI = 0, PQO_MAJORANA = 1, PQO_TOY = 2 };

/* ------------------------------------------------------------------ keys */

static int popc(uint64_t x) { return __builtin_popcountll(x); }

static int key_count(const uint64_t *k, int W) {
  int n = 0;
  for (int i = 0; i < W; i++) n += popc(k[i]);
  return n;
}

static int key_is_bit(const uint64_t *k, int pos) { return (int)((k[pos >> 6] >> (pos & 63)) & 1u); }

/* strings.rs:140 support(): units with at least one of their two bits set. */
static int key_support(const uint64_t *k, int W) {
  int n = 0;
  for (int i = 0; i < W; i++) n += popc((k[i] | (k[i] >> 1)) & EVEN);
  return n;
}

/* strings.rs:155 pair_swap(): exchange the two bits of every unit. */
static void key_pair_swap(const uint64_t *k, uint64_t *out, int W) {
  for (int i = 0; i < W; i++) out[i] = ((k[i] & EVEN) << 1) | ((k[i] >> 1) & EVEN);
}

static int key_eq(const uint64_t *a, const uint64_t *b, int W) {
  for (int i = 0; i < W; i++)
    if (a[i] != b[i]) return 0;
  return 1;
}

static uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static uint64_t key_hash(const uint64_t *k, int W) {
  uint64_t h = 0x243F6A8885A308D3ULL;
  for (int i = 0; i < W; i++) h = mix64(h ^ k[i]);
  return h;
}

/* ------------------------------------------------------------- algebra */

/* Per-gate generator context (pauli/src/algebra.rs:66-91,
 * majorana/src/algebra.rs:125-154). The phase exponent of G*P is kept as an
 * integer e with phase = i^e (sign folded in as +2), which is exact. */
typedef struct {
  int kind, W, n_units;
  uint64_t gen[PQO_MAXW];
  uint64_t fold[PQO_MAXW];
  int gen_y, gen_len;
  int neg; /* generator sign -1 from the Clifford frame */
} genctx;

/* pauli/src/algebra.rs:35 count_y_sites */
static int y_sites(const uint64_t *k, int W) {
  int n = 0;
  for (int i = 0; i < W; i++) n += popc(k[i] & (k[i] >> 1) & EVEN);
  return n;
}

/* majorana/src/monomial.rs:688 hermiticity_exp */
static int herm_exp(int len) { return (len % 4 == 0 || len % 4 == 1) ? 0 : 1; }

/* majorana/src/monomial.rs:696 resorting_parity(a, b): parity of the number of
 * pairs (x in a, y in b) with x > y.  Counted here one bit of b at a time
 * against the bits of a strictly above it. */
static int resort_parity(const uint64_t *a, const uint64_t *b, int W) {
  unsigned long long count = 0;
  for (int wb = 0; wb < W; wb++) {
    uint64_t bw = b[wb];
    while (bw) {
      int bit = __builtin_ctzll(bw);
      bw &= bw - 1;
      uint64_t above = bit == 63 ? 0 : (a[wb] >> (bit + 1));
      count += (unsigned long long)popc(above);
      for (int wa = wb + 1; wa < W; wa++) count += (unsigned long long)popc(a[wa]);
    }
  }
  return (int)(count & 1);
}

static void ctx_make(genctx *c, int kind, int W, int n_units, const uint64_t *gen, double sign) {
  c->kind = kind;
  c->W = W;
  c->n_units = n_units;
  memcpy(c->gen, gen, sizeof(uint64_t) * W);
  if (kind == PQO_PAULI)
    key_pair_swap(gen, c->fold, W);
  else
    memcpy(c->fold, gen, sizeof(uint64_t) * W);
  c->gen_y = y_sites(gen, W);
  c->gen_len = key_count(gen, W);
  c->neg = sign < 0.0;
}

/* Basis::anticommutes (pauli/src/algebra.rs:99, majorana/src/algebra.rs:162,
 * toy algebra termsum.rs test :20). */
static int anticommutes(const genctx *c, const uint64_t *k) {
  int W = c->W;
  if (c->kind == PQO_MAJORANA) {
    int overlap = 0;
    for (int i = 0; i < W; i++) overlap += popc(k[i] & c->gen[i]);
    int cross = (c->gen_len & 1) ? key_count(k, W) : 0;
    return (overlap + cross) & 1;
  }
  uint64_t acc = 0;
  for (int i = 0; i < W; i++) acc ^= k[i] & c->fold[i];
  return popc(acc) & 1;
}

/* Basis::product: out = k ^ gen; returns phase exponent e (phase = i^e). */
static int product(const genctx *c, const uint64_t *k, uint64_t *out) {
  int W = c->W, p;
  for (int i = 0; i < W; i++) out[i] = k[i] ^ c->gen[i];
  if (c->kind == PQO_PAULI) {
    int zx = 0;
    for (int i = 0; i < W; i++) zx += popc((c->gen[i] >> 1) & k[i] & EVEN);
    p = c->gen_y + y_sites(k, W) - y_sites(out, W) + 2 * zx;
  } else if (c->kind == PQO_MAJORANA) {
    p = herm_exp(c->gen_len) + herm_exp(key_count(k, W)) - herm_exp(key_count(out, W)) +
        2 * resort_parity(c->gen, k, W);
  } else {
    p = 1; /* toy algebra: phase +i */
  }
  p = ((p % 4) + 4) % 4;
  if (c->neg) p = (p + 2) % 4;
  return p;
}

/* -Im(i^e): the real factor of the sine branch (coeff.rs:165). */
static double neg_im(int e) { return e == 1 ? -1.0 : (e == 3 ? 1.0 : 0.0); }

/* Majorana JW weight (majorana/src/algebra.rs:85-123), restated per site:
 * single = x xor y, occupied = x or y, p = inclusive prefix XOR of single,
 * string = p xor [odd number of singles], weight counts single | (occ ^ str). */
static int jw_weight(const uint64_t *k, int W, int n_sites) {
  int n_single = 0;
  for (int s = 0; s < n_sites && 2 * s + 1 < 64 * W; s++)
    n_single += key_is_bit(k, 2 * s) ^ key_is_bit(k, 2 * s + 1);
  int flip = n_single & 1, pref = 0, w = 0;
  for (int s = 0; s < n_sites && 2 * s + 1 < 64 * W; s++) {
    int x = key_is_bit(k, 2 * s), y = key_is_bit(k, 2 * s + 1);
    int single = x ^ y, occ = x | y;
    pref ^= single;
    int str = pref ^ flip;
    w += single | (occ ^ str);
  }
  return w;
}

static uint32_t weight(int kind, const uint64_t *k, int W, int n_units) {
  if (kind == PQO_PAULI) return (uint32_t)key_support(k, W);
  if (kind == PQO_MAJORANA) return (uint32_t)jw_weight(k, W, n_units);
  return (uint32_t)key_count(k, W);
}

/* Basis::trace against a computational-basis / Fock state with one bit per
 * unit (pauli/src/algebra.rs:123, majorana/src/algebra.rs:200, toy :32). */
static double trace(int kind, const uint64_t *k, int W, int n_units, const uint64_t *diag, int n_diag) {
  if (kind == PQO_PAULI) {
    for (int i = 0; i < W; i++)
      if (k[i] & EVEN) return 0.0;
    int par = 0;
    for (int pos = 0; pos < 64 * W; pos++) {
      if (!key_is_bit(k, pos)) continue;
      int u = pos / 2;
      uint64_t dw = (u / 64) < n_diag ? diag[u / 64] : 0;
      par ^= (int)((dw >> (u % 64)) & 1);
    }
    return par ? -1.0 : 1.0;
  }
  if (kind == PQO_MAJORANA) {
    int paired = 0, prod = 1;
    for (int s = 0; s < n_units; s++) {
      int lo = key_is_bit(k, 2 * s), hi = key_is_bit(k, 2 * s + 1);
      if (lo != hi) return 0.0;
      if (lo) {
        uint64_t dw = (s / 64) < n_diag ? diag[s / 64] : 0;
        int nk = (int)((dw >> (s % 64)) & 1);
        prod *= 2 * nk - 1;
        paired++;
      }
    }
    return (double)(((paired / 2) % 2 == 0 ? 1 : -1) * prod);
  }
  uint64_t f = n_diag > 0 ? diag[0] : 0;
  return (k[0] & f) == 0 ? 1.0 : -1.0;
}

/* ------------------------------------------------------------- tableau */

/* CliffordTableau (tableau.rs:29-224): 2*n_units rows of (image, sign) for
 * the readout direction and the generator direction. */
typedef struct {
  int n_rows, W;
  uint64_t *ro_img, *gen_img; /* n_rows x W */
  double *ro_sign, *gen_sign;
  int identity;
} tableau;

static void tab_init(tableau *t, int n_units, int W) {
  t->n_rows = 2 * n_units;
  t->W = W;
  t->ro_img = calloc((size_t)t->n_rows * W + 1, sizeof(uint64_t));
  t->gen_img = calloc((size_t)t->n_rows * W + 1, sizeof(uint64_t));
  t->ro_sign = malloc(sizeof(double) * (t->n_rows + 1));
  t->gen_sign = malloc(sizeof(double) * (t->n_rows + 1));
  for (int p = 0; p < t->n_rows; p++) {
    if (p < 64 * W) {
      t->ro_img[(size_t)p * W + p / 64] = 1ULL << (p % 64);
      t->gen_img[(size_t)p * W + p / 64] = 1ULL << (p % 64);
    }
    t->ro_sign[p] = 1.0;
    t->gen_sign[p] = 1.0;
  }
  t->identity = 1;
}

static void tab_free(tableau *t) {
  free(t->ro_img);
  free(t->gen_img);
  free(t->ro_sign);
  free(t->gen_sign);
}

static void tab_refresh_identity(tableau *t) {
  int W = t->W;
  t->identity = 1;
  for (int p = 0; p < t->n_rows && t->identity; p++) {
    uint64_t want[PQO_MAXW] = {0};
    if (p < 64 * W) want[p / 64] = 1ULL << (p % 64);
    if (!key_eq(&t->ro_img[(size_t)p * W], want, W) || t->ro_sign[p] != 1.0) t->identity = 0;
  }
}

/* tableau.rs:109 apply_rows: fold the rows selected by the set bits of m,
 * tracking the phase of the source product and of the image product; the
 * sign is that of Re(image_phase / source_phase). */
static double tab_apply_rows(int kind, int n_units, int W, const uint64_t *img, const double *sign, int n_rows,
                             const uint64_t *m, uint64_t *out) {
  uint64_t src[PQO_MAXW] = {0}, im[PQO_MAXW] = {0}, tmp[PQO_MAXW], single[PQO_MAXW];
  int src_e = 0, im_e = 0;
  genctx c;
  for (int pos = 0; pos < 64 * W; pos++) {
    if (!key_is_bit(m, pos) || pos >= n_rows) continue;
    memset(single, 0, sizeof(uint64_t) * W);
    single[pos / 64] = 1ULL << (pos % 64);
    ctx_make(&c, kind, W, n_units, single, 1.0);
    src_e += product(&c, src, tmp);
    memcpy(src, tmp, sizeof(uint64_t) * W);
    ctx_make(&c, kind, W, n_units, &img[(size_t)pos * W], 1.0);
    im_e += product(&c, im, tmp);
    memcpy(im, tmp, sizeof(uint64_t) * W);
    if (sign[pos] < 0.0) im_e += 2;
  }
  memcpy(out, im, sizeof(uint64_t) * W);
  int ratio = (((im_e - src_e) % 4) + 4) % 4;
  return ratio == 2 ? -1.0 : 1.0;
}

static double tab_conjugate(const tableau *t, int kind, int n_units, const uint64_t *m, uint64_t *out) {
  if (t->identity) {
    memcpy(out, m, sizeof(uint64_t) * t->W);
    return 1.0;
  }
  return tab_apply_rows(kind, n_units, t->W, t->ro_img, t->ro_sign, t->n_rows, m, out);
}

static double tab_conjugate_generator(const tableau *t, int kind, int n_units, const uint64_t *m, uint64_t *out) {
  if (t->identity) {
    memcpy(out, m, sizeof(uint64_t) * t->W);
    return 1.0;
  }
  return tab_apply_rows(kind, n_units, t->W, t->gen_img, t->gen_sign, t->n_rows, m, out);
}

/* tableau.rs:146 compose(next): readout_new = next.readout o self.readout,
 * generator_new = self.generator o next.generator. */
static void tab_compose(tableau *self, const tableau *next, int kind, int n_units) {
  if (next->identity) return;
  int W = self->W, R = self->n_rows;
  uint64_t *ro = calloc((size_t)R * W + 1, sizeof(uint64_t)), *ge = calloc((size_t)R * W + 1, sizeof(uint64_t));
  double *ros = malloc(sizeof(double) * (R + 1)), *ges = malloc(sizeof(double) * (R + 1));
  for (int p = 0; p < R; p++) {
    double s = tab_apply_rows(kind, n_units, W, next->ro_img, next->ro_sign, R, &self->ro_img[(size_t)p * W],
                              &ro[(size_t)p * W]);
    ros[p] = s * self->ro_sign[p];
    s = tab_apply_rows(kind, n_units, W, self->gen_img, self->gen_sign, R, &next->gen_img[(size_t)p * W],
                       &ge[(size_t)p * W]);
    ges[p] = s * next->gen_sign[p];
  }
  free(self->ro_img);
  free(self->gen_img);
  free(self->ro_sign);
  free(self->gen_sign);
  self->ro_img = ro;
  self->gen_img = ge;
  self->ro_sign = ros;
  self->gen_sign = ges;
  tab_refresh_identity(self);
}

/* tableau.rs:180 for_rotation at a Clifford angle: for each single-bit
 * generator e_p anticommuting with G, readout[p] = (G e_p, s) and
 * generator[p] = (G e_p, -s) with s = sin(theta) * (-Im phase). */
static void tab_for_rotation(tableau *t, int kind, int n_units, int W, const uint64_t *gen, double angle) {
  tab_init(t, n_units, W);
  genctx c;
  ctx_make(&c, kind, W, n_units, gen, 1.0);
  double s_t = sin(angle);
  uint64_t single[PQO_MAXW], img[PQO_MAXW];
  for (int p = 0; p < t->n_rows && p < 64 * W; p++) {
    memset(single, 0, sizeof(uint64_t) * W);
    single[p / 64] = 1ULL << (p % 64);
    if (!anticommutes(&c, single)) continue;
    int e = product(&c, single, img);
    double s = s_t * neg_im(e);
    memcpy(&t->ro_img[(size_t)p * W], img, sizeof(uint64_t) * W);
    memcpy(&t->gen_img[(size_t)p * W], img, sizeof(uint64_t) * W);
    t->ro_sign[p] = s;
    t->gen_sign[p] = -s;
  }
  tab_refresh_identity(t);
}

/* tableau.rs:215 changes_weight: any readout row whose image support differs
 * from the support of its origin bit. */
static int tab_changes_weight(const tableau *t) {
  int W = t->W;
  for (int p = 0; p < t->n_rows; p++) {
    uint64_t o[PQO_MAXW] = {0};
    if (p < 64 * W) o[p / 64] = 1ULL << (p % 64);
    if (key_support(&t->ro_img[(size_t)p * W], W) != key_support(o, W)) return 1;
  }
  return 0;
}

/* -------------------------------------------------------------- engine */

/* propaq_truncator_keep of the truncation plugin ABI
 * (crates/core/src/policy/native_truncator.rs:19-23); nonzero = keep. */
typedef int32_t (*pqo_keep_fn)(void *ctx, uint32_t basis_kind, const uint64_t *words, size_t n_words,
                               uint32_t n_units, uint32_t weight, double coeff_magnitude, uint32_t layer_index,
                               uint32_t n_layers);

typedef struct pqo_engine {
  int kind, n_units, W;
  /* EmitCutoff (termsum.rs:46) */
  int has_maxw;
  uint32_t maxw;
  int has_min;
  double min;
  int has_min_terms;
  uint64_t min_terms;
  /* EmitCutoff::term (termsum.rs:50) and its LayerContext (:56) */
  pqo_keep_fn keep;
  void *keep_ctx;
  int keep_reads_key;
  uint32_t layer_index, n_layers;
  int defer;
  /* rows */
  size_t n, cap;
  uint64_t *keys;
  double *coeffs;
  /* open-addressed index of rows by key */
  uint32_t *slots;
  size_t nslots;
  /* held-back branches (termsum.rs:215-220) */
  uint32_t *pending_slot;
  size_t pending_slot_len;
  uint32_t *pending_rows;
  double *pending_vals;
  size_t n_pending, cap_pending;
  uint32_t *claimed;
  size_t n_claimed, cap_claimed;
  int pending_needs_phase;
  /* routed children */
  uint64_t *out_keys;
  double *out_vals;
  size_t n_out, cap_out;
  uint32_t *scan_rows;
  size_t cap_scan;
  tableau frame;
  /* counters (termsum.rs:221-231) */
  uint64_t emitted, hits, visited, declined;
  double disc_l1, disc_max;
} pqo_engine;

#define EMPTY_SLOT 0xFFFFFFFFu

static void index_rebuild(pqo_engine *e, size_t want_rows) {
  size_t ns = 16;
  while ((double)ns * 0.7 < (double)(want_rows + 1)) ns <<= 1;
  free(e->slots);
  e->slots = malloc(sizeof(uint32_t) * ns);
  memset(e->slots, 0xFF, sizeof(uint32_t) * ns);
  e->nslots = ns;
  for (size_t r = 0; r < e->n; r++) {
    size_t h = (size_t)key_hash(&e->keys[r * e->W], e->W) & (ns - 1);
    while (e->slots[h] != EMPTY_SLOT) h = (h + 1) & (ns - 1);
    e->slots[h] = (uint32_t)r;
  }
}

static long index_find(const pqo_engine *e, const uint64_t *k) {
  size_t h = (size_t)key_hash(k, e->W) & (e->nslots - 1);
  for (;;) {
    uint32_t r = e->slots[h];
    if (r == EMPTY_SLOT) return -1;
    if (key_eq(&e->keys[(size_t)r * e->W], k, e->W)) return (long)r;
    h = (h + 1) & (e->nslots - 1);
  }
}

static void rows_reserve(pqo_engine *e, size_t need) {
  if (need <= e->cap) return;
  size_t c = e->cap ? e->cap : 64;
  while (c < need) c *= 2;
  e->keys = realloc(e->keys, sizeof(uint64_t) * c * e->W);
  e->coeffs = realloc(e->coeffs, sizeof(double) * c);
  e->cap = c;
}

/* Append a row known to be absent, keeping the index load below 0.7. */
static size_t push_row(pqo_engine *e, const uint64_t *k, double c) {
  rows_reserve(e, e->n + 1);
  size_t r = e->n++;
  memcpy(&e->keys[r * e->W], k, sizeof(uint64_t) * e->W);
  e->coeffs[r] = c;
  if ((double)e->nslots * 0.7 < (double)(e->n + 1)) {
    index_rebuild(e, e->n * 2);
  } else {
    size_t h = (size_t)key_hash(k, e->W) & (e->nslots - 1);
    while (e->slots[h] != EMPTY_SLOT) h = (h + 1) & (e->nslots - 1);
    e->slots[h] = (uint32_t)r;
  }
  return r;
}

pqo_engine *pqo_create(int kind, int n_units, int W) {
  if (W < 1 || W > PQO_MAXW || kind < 0 || kind > 2) return NULL;
  pqo_engine *e = calloc(1, sizeof(pqo_engine));
  e->kind = kind;
  e->n_units = n_units;
  e->W = W;
  e->defer = 1;
  tab_init(&e->frame, n_units, W);
  index_rebuild(e, 0);
  return e;
}

void pqo_destroy(pqo_engine *e) {
  if (!e) return;
  free(e->keys);
  free(e->coeffs);
  free(e->slots);
  free(e->pending_slot);
  free(e->pending_rows);
  free(e->pending_vals);
  free(e->claimed);
  free(e->out_keys);
  free(e->out_vals);
  free(e->scan_rows);
  tab_free(&e->frame);
  free(e);
}

void pqo_set_cutoff(pqo_engine *e, int has_maxw, uint32_t maxw, int has_min, double min, int has_min_terms,
                    uint64_t min_terms) {
  e->has_maxw = has_maxw;
  e->maxw = maxw;
  e->has_min = has_min;
  e->min = min;
  e->has_min_terms = has_min_terms;
  e->min_terms = min_terms;
}

/* A truncation plugin's scalar keep; reads_key = the plugin declared
 * PROPAQ_DEPENDS_KEY (words are NULL otherwise, native_truncator.rs:125-131). */
void pqo_set_truncator(pqo_engine *e, pqo_keep_fn keep, void *ctx, int reads_key) {
  e->keep = keep;
  e->keep_ctx = ctx;
  e->keep_reads_key = reads_key;
}

/* cutoff.layer = LayerContext::new(layer_idx, n_layers) (pauli/src/engine.rs:118). */
void pqo_set_layer(pqo_engine *e, uint32_t layer_index, uint32_t n_layers) {
  e->layer_index = layer_index;
  e->n_layers = n_layers;
}

/* NativeTruncator::keep (native_truncator.rs:141-156) on a raw row key. */
static int kernel_keeps(const pqo_engine *e, const uint64_t *k, double mag) {
  uint32_t w = weight(e->kind, k, e->W, e->n_units);
  return e->keep(e->keep_ctx, (uint32_t)e->kind, e->keep_reads_key ? k : NULL, e->keep_reads_key ? (size_t)e->W : 0,
                 (uint32_t)e->n_units, w, mag, e->layer_index, e->n_layers) != 0;
}

void pqo_set_defer(pqo_engine *e, int on) { e->defer = on; }

size_t pqo_len(const pqo_engine *e) { return e->n; }

int pqo_frame_is_identity(const pqo_engine *e) { return e->frame.identity; }

/* TermSum::add (termsum.rs:534): fold into an existing row or append. */
void pqo_add(pqo_engine *e, const uint64_t *k, double c) {
  long r = index_find(e, k);
  if (r >= 0) {
    e->coeffs[r] += c;
    return;
  }
  push_row(e, k, c);
}

/* EmitCutoff::admits_initial (termsum.rs:152) with the full cutoff. */
int pqo_admits_initial(const pqo_engine *e, const uint64_t *k, double c) {
  if (e->keep) return kernel_keeps(e, k, fabs(c)); /* admits_key is true with a kernel (:115) */
  if (e->has_maxw && weight(e->kind, k, e->W, e->n_units) > e->maxw) return 0;
  if (e->has_min && !(fabs(c) >= e->min)) return 0;
  return 1;
}

uint32_t pqo_weight(int kind, const uint64_t *k, int W, int n_units) { return weight(kind, k, W, n_units); }

double pqo_trace(int kind, const uint64_t *k, int W, int n_units, const uint64_t *diag, int n_diag) {
  return trace(kind, k, W, n_units, diag, n_diag);
}

int pqo_product(int kind, const uint64_t *gen, const uint64_t *k, int W, int n_units, uint64_t *out) {
  genctx c;
  ctx_make(&c, kind, W, n_units, gen, 1.0);
  return product(&c, k, out);
}

int pqo_anticommutes(int kind, const uint64_t *gen, const uint64_t *k, int W, int n_units) {
  genctx c;
  ctx_make(&c, kind, W, n_units, gen, 1.0);
  return anticommutes(&c, k);
}

static void out_push(pqo_engine *e, const uint64_t *k, double v) {
  if (e->n_out == e->cap_out) {
    e->cap_out = e->cap_out ? e->cap_out * 2 : 256;
    e->out_keys = realloc(e->out_keys, sizeof(uint64_t) * e->cap_out * e->W);
    e->out_vals = realloc(e->out_vals, sizeof(double) * e->cap_out);
  }
  memcpy(&e->out_keys[e->n_out * e->W], k, sizeof(uint64_t) * e->W);
  e->out_vals[e->n_out++] = v;
}

/* termsum.rs:666 reset_pending */
static void reset_pending(pqo_engine *e) {
  for (size_t i = 0; i < e->n_pending; i++) e->pending_slot[e->pending_rows[i]] = EMPTY_SLOT;
  e->n_pending = 0;
  e->n_claimed = 0;
  if (e->pending_slot_len < e->n) {
    e->pending_slot = realloc(e->pending_slot, sizeof(uint32_t) * e->n);
    for (size_t i = e->pending_slot_len; i < e->n; i++) e->pending_slot[i] = EMPTY_SLOT;
    e->pending_slot_len = e->n;
  }
}

/* termsum.rs:680 hold_back */
static void hold_back(pqo_engine *e, size_t row, double v) {
  if (row >= e->pending_slot_len) return;
  if (e->n_pending == e->cap_pending) {
    e->cap_pending = e->cap_pending ? e->cap_pending * 2 : 256;
    e->pending_rows = realloc(e->pending_rows, sizeof(uint32_t) * e->cap_pending);
    e->pending_vals = realloc(e->pending_vals, sizeof(double) * e->cap_pending);
  }
  e->pending_slot[row] = (uint32_t)e->n_pending;
  e->pending_rows[e->n_pending] = (uint32_t)row;
  e->pending_vals[e->n_pending++] = v;
}

/* termsum.rs:691 claim */
static void claim(pqo_engine *e, size_t row) {
  if (row >= e->pending_slot_len || e->pending_slot[row] == EMPTY_SLOT) return;
  if (e->n_claimed == e->cap_claimed) {
    e->cap_claimed = e->cap_claimed ? e->cap_claimed * 2 : 256;
    e->claimed = realloc(e->claimed, sizeof(uint32_t) * e->cap_claimed);
  }
  e->claimed[e->n_claimed++] = (uint32_t)row;
}

/* termsum.rs:788 absorb_routed over the current outbox; returns rows added. */
static size_t absorb_outbox(pqo_engine *e) {
  size_t added = 0;
  for (size_t i = 0; i < e->n_out; i++) {
    const uint64_t *k = &e->out_keys[i * e->W];
    long r = index_find(e, k);
    if (r >= 0) {
      e->hits++;
      e->coeffs[r] += e->out_vals[i];
      claim(e, (size_t)r);
    } else {
      push_row(e, k, e->out_vals[i]);
      added++;
    }
  }
  e->n_out = 0;
  return added;
}

/* Effective cutoff for this gate (termsum.rs:70 at_size): below the term
 * floor every lossy predicate is off. */
typedef struct {
  int has_maxw;
  uint32_t maxw;
  int has_min;
  double min;
  int kernel;
} eff_cutoff;

static eff_cutoff at_size(const pqo_engine *e) {
  eff_cutoff c = {e->has_maxw, e->maxw, e->has_min, e->min, e->keep != NULL};
  if (e->has_min_terms && e->n < e->min_terms) {
    c.has_maxw = 0;
    c.has_min = 0;
    c.kernel = 0;
  }
  return c;
}

/* PartitionedTermSum::apply_rotation (partitioned_termsum.rs:276-350) for one
 * partition, with TermSum::scan_into / emit_from_row / absorb / drain_claims
 * (termsum.rs:561-800).  Returns the number of rows created. */
size_t pqo_apply_rotation(pqo_engine *e, const uint64_t *gen_in, double angle) {
  int W = e->W;
  if (e->n == 0) return 0;
  eff_cutoff cut = at_size(e);

  if (e->defer && fabs(cos(angle)) < 1e-9) {
    tableau step;
    tab_for_rotation(&step, e->kind, e->n_units, W, gen_in, angle);
    if (!cut.has_maxw || !tab_changes_weight(&step)) {
      tab_compose(&e->frame, &step, e->kind, e->n_units);
      tab_free(&step);
      return 0;
    }
    tab_free(&step);
  }

  uint64_t gen[PQO_MAXW];
  double gsign = tab_conjugate_generator(&e->frame, e->kind, e->n_units, gen_in, gen);
  genctx ctx;
  ctx_make(&ctx, e->kind, W, e->n_units, gen, gsign);

  double s_t = sin(angle), c_t = cos(angle);
  if (fabs(s_t) < 1e-9) { /* phase-only (coeff.rs:200, termsum.rs:551) */
    for (size_t i = 0; i < e->n; i++)
      if (anticommutes(&ctx, &e->keys[i * W])) e->coeffs[i] *= c_t;
    return 0;
  }

  /* scan (termsum.rs:561) */
  reset_pending(e);
  int precheck = cut.has_min && !cut.kernel; /* EmitPrecheck::for_gate: min set, no kernel */
  e->pending_needs_phase = precheck;
  size_t n_scan = 0;
  if (e->cap_scan < e->n) {
    e->cap_scan = e->n;
    e->scan_rows = realloc(e->scan_rows, sizeof(uint32_t) * e->cap_scan);
  }
  for (size_t i = 0; i < e->n; i++)
    if (anticommutes(&ctx, &e->keys[i * W])) e->scan_rows[n_scan++] = (uint32_t)i;

  uint64_t child[PQO_MAXW];
  e->n_out = 0;
  for (size_t j = 0; j < n_scan; j++) { /* emit_from_row (termsum.rs:701) */
    size_t i = e->scan_rows[j];
    e->visited++;
    double c = e->coeffs[i];
    if (precheck && fabs(c * s_t) < cut.min) {
      e->declined++;
      hold_back(e, i, c * s_t);
      e->coeffs[i] = c * c_t;
      continue;
    }
    int ph = product(&ctx, &e->keys[i * W], child);
    double b = c * s_t * neg_im(ph);
    e->coeffs[i] = c * c_t;
    if (cut.kernel) { /* admits_key is true; admits_child = the kernel's keep (termsum.rs:115-148) */
      if (!kernel_keeps(e, child, fabs(b))) {
        hold_back(e, i, b);
        e->declined++;
        continue;
      }
      e->emitted++;
      out_push(e, child, b);
      continue;
    }
    if (cut.has_maxw && weight(e->kind, child, W, e->n_units) > cut.maxw) continue;
    if (cut.has_min && !(fabs(b) >= cut.min)) {
      hold_back(e, i, b);
      e->declined++;
      continue;
    }
    e->emitted++;
    out_push(e, child, b);
  }

  /* absorb, then drain claims (termsum.rs:604) and absorb the rescues */
  size_t added = absorb_outbox(e);
  e->disc_max = 0.0;
  if (e->n_claimed != e->n_pending) {
    /* a pending row is unclaimed iff it is not in the claimed list */
    unsigned char *was = calloc(e->n_pending + 1, 1);
    for (size_t j = 0; j < e->n_claimed; j++) was[e->pending_slot[e->claimed[j]]] = 1;
    for (size_t j = 0; j < e->n_pending; j++) {
      if (was[j]) continue;
      double m = fabs(e->pending_vals[j]);
      e->disc_l1 += m;
      if (m > e->disc_max) e->disc_max = m;
    }
    free(was);
  }
  size_t nc = e->n_claimed;
  for (size_t j = 0; j < nc; j++) {
    size_t row = e->claimed[j];
    uint32_t slot = e->pending_slot[row];
    int ph = product(&ctx, &e->keys[row * W], child);
    double v = e->pending_vals[slot];
    if (e->pending_needs_phase) v *= neg_im(ph);
    e->emitted++;
    out_push(e, child, v);
  }
  added += absorb_outbox(e);
  return added;
}

/* A layer's rotations in order, with the live count after each: the run
 * driver's gate loop (crates/pauli/src/engine.rs:165-193) without a call per
 * gate from the host language. */
void pqo_apply_rotations(pqo_engine *e, const uint64_t *gens, const double *angles, size_t count, uint64_t *n_after) {
  for (size_t i = 0; i < count; i++) {
    pqo_apply_rotation(e, gens + i * (size_t)e->W, angles[i]);
    n_after[i] = e->n;
  }
}

/* TermSum::scale_by_weight (termsum.rs:321) with factor table[min(w, len-1)]
 * as the run driver indexes it (pauli/src/engine.rs:121-124). */
void pqo_scale_by_table(pqo_engine *e, const double *table, size_t len) {
  for (size_t i = 0; i < e->n; i++) {
    size_t w = weight(e->kind, &e->keys[i * e->W], e->W, e->n_units);
    if (w > len - 1) w = len - 1;
    e->coeffs[i] *= table[w];
  }
}

/* Per-row factors already evaluated by a key-aware model, in row order
 * (termsum.rs:340 scale_by_key). */
void pqo_scale_by_factors(pqo_engine *e, const double *f) {
  for (size_t i = 0; i < e->n; i++) e->coeffs[i] *= f[i];
}

static void clear_pending_state(pqo_engine *e) {
  e->pending_slot_len = 0;
  e->n_pending = 0;
  e->n_claimed = 0;
}

/* PartitionedTermSum::reclaim (partitioned_termsum.rs:581) + TermSum::reclaim
 * (termsum.rs:472): keep rows that the full cutoff still admits. keep_flags,
 * when given, replaces the predicate; a truncation plugin decides row by row
 * with its scalar keep (reclaim_by_kernel, termsum.rs:387-447, calls the batch
 * form when the plugin has one: the same decisions for a plugin whose keep is
 * a function of its arguments). */
size_t pqo_reclaim(pqo_engine *e, const unsigned char *keep_flags) {
  if (!keep_flags && !e->has_maxw && !e->has_min && !e->keep) return 0;
  size_t before = e->n, w = 0;
  for (size_t i = 0; i < before; i++) {
    int keep = keep_flags ? keep_flags[i] != 0 : pqo_admits_initial(e, &e->keys[i * e->W], e->coeffs[i]);
    if (!keep) continue;
    if (w != i) {
      memmove(&e->keys[w * e->W], &e->keys[i * e->W], sizeof(uint64_t) * e->W);
      e->coeffs[w] = e->coeffs[i];
    }
    w++;
  }
  e->n = w;
  index_rebuild(e, e->n * 2);
  clear_pending_state(e);
  return before - w;
}

/* partitioned_termsum.rs:676 expectation, summed in row order. */
double pqo_expectation(const pqo_engine *e, const uint64_t *diag, int n_diag) {
  double acc = 0.0;
  uint64_t img[PQO_MAXW];
  for (size_t i = 0; i < e->n; i++) {
    const uint64_t *k = &e->keys[i * e->W];
    if (e->frame.identity) {
      acc += e->coeffs[i] * trace(e->kind, k, e->W, e->n_units, diag, n_diag);
    } else {
      double s = tab_conjugate(&e->frame, e->kind, e->n_units, k, img);
      acc += e->coeffs[i] * s * trace(e->kind, img, e->W, e->n_units, diag, n_diag);
    }
  }
  return acc;
}

/* partitioned_termsum.rs:611 terms_below */
size_t pqo_terms_below(const pqo_engine *e, double thr) {
  size_t n = 0;
  for (size_t i = 0; i < e->n; i++)
    if (fabs(e->coeffs[i]) < thr) n++;
  return n;
}

/* Rows as stored (no frame), in row order. */
void pqo_export_raw(const pqo_engine *e, uint64_t *keys, double *coeffs) {
  memcpy(keys, e->keys, sizeof(uint64_t) * e->W * e->n);
  memcpy(coeffs, e->coeffs, sizeof(double) * e->n);
}

/* PartitionedTermSum::iter (partitioned_termsum.rs:265): rows with the
 * deferred frame applied, coefficient times the frame sign
 * (pauli/src/engine.rs:410-417). */
void pqo_export(const pqo_engine *e, uint64_t *keys, double *coeffs) {
  for (size_t i = 0; i < e->n; i++) {
    double s = tab_conjugate(&e->frame, e->kind, e->n_units, &e->keys[i * e->W], &keys[i * e->W]);
    coeffs[i] = e->coeffs[i] * s;
  }
}

/* Readout conjugation of an arbitrary key through the current frame. */
double pqo_frame_conjugate(const pqo_engine *e, const uint64_t *k, uint64_t *out) {
  return tab_conjugate(&e->frame, e->kind, e->n_units, k, out);
}

double pqo_frame_conjugate_generator(const pqo_engine *e, const uint64_t *k, uint64_t *out) {
  return tab_conjugate_generator(&e->frame, e->kind, e->n_units, k, out);
}

/* Compose a standalone Clifford step (test hook for the tableau pins). */
int pqo_frame_compose_rotation(pqo_engine *e, const uint64_t *gen, double angle, int *changes_weight) {
  if (!(fabs(cos(angle)) < 1e-9)) return 0;
  tableau step;
  tab_for_rotation(&step, e->kind, e->n_units, e->W, gen, angle);
  if (changes_weight) *changes_weight = tab_changes_weight(&step);
  tab_compose(&e->frame, &step, e->kind, e->n_units);
  tab_free(&step);
  return 1;
}

void pqo_counters(const pqo_engine *e, uint64_t *out4, double *disc2) {
  out4[0] = e->visited;
  out4[1] = e->emitted;
  out4[2] = e->declined;
  out4[3] = e->hits;
  disc2[0] = e->disc_l1;
  disc2[1] = e->disc_max;
}

/* ------------------------------------------------------ partitioned engine
 *
 * PartitionedTermSum (partitioned_termsum.rs:111-350, 478-697): the term set
 * split over S partitions by the owner of each key, partition_from_hash
 * (termsum.rs:872-877): (hash >> 32) mod S, with hash_value (strings.rs:185-
 * 197) = splitmix64 of the word for one-word keys.  For wider keys the
 * reference uses rustc-hash's FxHasher, which is not restated here; this file
 * uses its own mix (key_hash), which changes only which partition owns a key,
 * never a coefficient (every key receives at most one branch per gate).
 * One gate:
 *   scan    every partition runs emit_from_row over its anticommuting rows and
 *           routes each child to its owner's inbox, holding pending branches
 *           back locally;                                       (barrier)
 *   absorb  every owner absorbs its inboxes in source order (hits claim the
 *           hit row's held-back branch), drains its claims and routes the
 *           rescued branches to their owners;                   (barrier)
 *   rescue  every owner absorbs the rescued branches.
 * The partitions run on threads (OpenMP), one partition per thread.  This is
 * the CPU baseline of the benchmark; its results equal the single-partition
 * engine's key for key and bit for bit (tested for several S). */
#include <omp.h>

typedef struct {
  uint64_t *keys;
  double *vals;
  size_t n, cap;
} obox;

typedef struct pqo_part {
  int kind, n_units, W, S, threads;
  int defer;
  int has_maxw, has_min, has_min_terms;
  uint32_t maxw;
  double min;
  uint64_t min_terms;
  pqo_engine **p;
  obox *out;  /* S x S, [src * S + dst] */
  obox *res;  /* rescued branches, same layout */
  tableau frame;
} pqo_part;

static uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static int owner_of(const pqo_part *t, const uint64_t *k) {
  if (t->S <= 1) return 0;
  const uint64_t h = t->W == 1 ? splitmix64(k[0]) : key_hash(k, t->W);
  return (int)((h >> 32) % (uint64_t)t->S);
}

static void obox_push(obox *b, const uint64_t *k, double v, int W) {
  if (b->n == b->cap) {
    b->cap = b->cap ? b->cap * 2 : 256;
    b->keys = realloc(b->keys, sizeof(uint64_t) * b->cap * W);
    b->vals = realloc(b->vals, sizeof(double) * b->cap);
  }
  memcpy(&b->keys[b->n * W], k, sizeof(uint64_t) * W);
  b->vals[b->n++] = v;
}

pqo_part *pqo_part_create(int kind, int n_units, int W, int n_parts, int n_threads) {
  if (n_parts < 1 || n_parts > 1024) return NULL;
  pqo_part *t = calloc(1, sizeof(pqo_part));
  t->kind = kind;
  t->n_units = n_units;
  t->W = W;
  t->S = n_parts;
  t->threads = n_threads < 1 ? 1 : n_threads;
  t->defer = 1;
  t->p = calloc((size_t)n_parts, sizeof(pqo_engine *));
  for (int i = 0; i < n_parts; i++) {
    t->p[i] = pqo_create(kind, n_units, W);
    if (!t->p[i]) {
      for (int j = 0; j < i; j++) pqo_destroy(t->p[j]);
      free(t->p);
      free(t);
      return NULL;
    }
  }
  t->out = calloc((size_t)n_parts * n_parts, sizeof(obox));
  t->res = calloc((size_t)n_parts * n_parts, sizeof(obox));
  tab_init(&t->frame, n_units, W);
  return t;
}

void pqo_part_destroy(pqo_part *t) {
  if (!t) return;
  for (int i = 0; i < t->S; i++) pqo_destroy(t->p[i]);
  for (int i = 0; i < t->S * t->S; i++) {
    free(t->out[i].keys);
    free(t->out[i].vals);
    free(t->res[i].keys);
    free(t->res[i].vals);
  }
  free(t->out);
  free(t->res);
  free(t->p);
  tab_free(&t->frame);
  free(t);
}

void pqo_part_set_cutoff(pqo_part *t, int has_maxw, uint32_t maxw, int has_min, double min, int has_min_terms,
                         uint64_t min_terms) {
  t->has_maxw = has_maxw;
  t->maxw = maxw;
  t->has_min = has_min;
  t->min = min;
  t->has_min_terms = has_min_terms;
  t->min_terms = min_terms;
  for (int i = 0; i < t->S; i++) pqo_set_cutoff(t->p[i], has_maxw, maxw, has_min, min, has_min_terms, min_terms);
}

void pqo_part_set_defer(pqo_part *t, int on) { t->defer = on; }

size_t pqo_part_len(const pqo_part *t) {
  size_t n = 0;
  for (int i = 0; i < t->S; i++) n += t->p[i]->n;
  return n;
}

int pqo_part_admits_initial(const pqo_part *t, const uint64_t *k, double c) { return pqo_admits_initial(t->p[0], k, c); }

void pqo_part_add(pqo_part *t, const uint64_t *k, double c) { pqo_add(t->p[owner_of(t, k)], k, c); }

/* scan phase of partition q: emit_from_row (termsum.rs:701-751), routed */
static void part_scan(pqo_part *t, int q, const genctx *ctx, eff_cutoff cut, double s_t, double c_t) {
  pqo_engine *e = t->p[q];
  const int W = t->W;
  for (int d = 0; d < t->S; d++) t->out[q * t->S + d].n = 0;
  reset_pending(e);
  const int precheck = cut.has_min && !cut.kernel;
  e->pending_needs_phase = precheck;
  uint64_t child[PQO_MAXW];
  for (size_t i = 0; i < e->n; i++) {
    if (!anticommutes(ctx, &e->keys[i * W])) continue;
    e->visited++;
    const double c = e->coeffs[i];
    if (precheck && fabs(c * s_t) < cut.min) {
      e->declined++;
      hold_back(e, i, c * s_t);
      e->coeffs[i] = c * c_t;
      continue;
    }
    const int ph = product(ctx, &e->keys[i * W], child);
    const double b = c * s_t * neg_im(ph);
    e->coeffs[i] = c * c_t;
    if (cut.has_maxw && weight(e->kind, child, W, e->n_units) > cut.maxw) continue;
    if (cut.has_min && !(fabs(b) >= cut.min)) {
      hold_back(e, i, b);
      e->declined++;
      continue;
    }
    e->emitted++;
    obox_push(&t->out[q * t->S + owner_of(t, child)], child, b, W);
  }
}

/* absorb_routed (termsum.rs:788) of the boxes [src][d] for every src, in order */
static size_t part_absorb(pqo_part *t, int d, obox *boxes) {
  pqo_engine *e = t->p[d];
  size_t added = 0;
  for (int q = 0; q < t->S; q++) {
    obox *b = &boxes[q * t->S + d];
    for (size_t i = 0; i < b->n; i++) {
      const uint64_t *k = &b->keys[i * t->W];
      const long r = index_find(e, k);
      if (r >= 0) {
        e->hits++;
        e->coeffs[r] += b->vals[i];
        claim(e, (size_t)r);
      } else {
        push_row(e, k, b->vals[i]);
        added++;
      }
    }
  }
  return added;
}

/* unclaimed held-back branches to the discard statistics, then the claimed
 * ones re-emitted to their owners (drain_claims, termsum.rs:604-647) */
static void part_drain(pqo_part *t, int d, const genctx *ctx) {
  pqo_engine *e = t->p[d];
  for (int x = 0; x < t->S; x++) t->res[d * t->S + x].n = 0;
  e->disc_max = 0.0;
  if (e->n_claimed != e->n_pending) {
    unsigned char *was = calloc(e->n_pending + 1, 1);
    for (size_t j = 0; j < e->n_claimed; j++) was[e->pending_slot[e->claimed[j]]] = 1;
    for (size_t j = 0; j < e->n_pending; j++) {
      if (was[j]) continue;
      const double m = fabs(e->pending_vals[j]);
      e->disc_l1 += m;
      if (m > e->disc_max) e->disc_max = m;
    }
    free(was);
  }
  uint64_t child[PQO_MAXW];
  const size_t nc = e->n_claimed;
  for (size_t j = 0; j < nc; j++) {
    const size_t row = e->claimed[j];
    const uint32_t slot = e->pending_slot[row];
    const int ph = product(ctx, &e->keys[row * t->W], child);
    double v = e->pending_vals[slot];
    if (e->pending_needs_phase) v *= neg_im(ph);
    e->emitted++;
    obox_push(&t->res[d * t->S + owner_of(t, child)], child, v, t->W);
  }
}

/* PartitionedTermSum::apply_rotation (partitioned_termsum.rs:276-350) */
size_t pqo_part_apply_rotation(pqo_part *t, const uint64_t *gen_in, double angle) {
  const int W = t->W, S = t->S;
  const size_t n_all = pqo_part_len(t);
  if (n_all == 0) return 0;
  eff_cutoff cut = {t->has_maxw, t->maxw, t->has_min, t->min, 0};
  if (t->has_min_terms && n_all < t->min_terms) cut.has_maxw = cut.has_min = 0; /* at_size over all partitions */
  if (t->defer && fabs(cos(angle)) < 1e-9) {
    tableau step;
    tab_for_rotation(&step, t->kind, t->n_units, W, gen_in, angle);
    if (!cut.has_maxw || !tab_changes_weight(&step)) {
      tab_compose(&t->frame, &step, t->kind, t->n_units);
      tab_free(&step);
      return 0;
    }
    tab_free(&step);
  }
  uint64_t gen[PQO_MAXW];
  const double gsign = tab_conjugate_generator(&t->frame, t->kind, t->n_units, gen_in, gen);
  genctx ctx;
  ctx_make(&ctx, t->kind, W, t->n_units, gen, gsign);
  const double s_t = sin(angle), c_t = cos(angle);
  if (fabs(s_t) < 1e-9) {
#pragma omp parallel for num_threads(t->threads) schedule(static, 1)
    for (int q = 0; q < S; q++) {
      pqo_engine *e = t->p[q];
      for (size_t i = 0; i < e->n; i++)
        if (anticommutes(&ctx, &e->keys[i * W])) e->coeffs[i] *= c_t;
    }
    return 0;
  }
  size_t added = 0;
#pragma omp parallel num_threads(t->threads) reduction(+ : added)
  {
#pragma omp for schedule(static, 1)
    for (int q = 0; q < S; q++) part_scan(t, q, &ctx, cut, s_t, c_t);
#pragma omp for schedule(static, 1)
    for (int d = 0; d < S; d++) {
      added += part_absorb(t, d, t->out);
      part_drain(t, d, &ctx);
    }
#pragma omp for schedule(static, 1)
    for (int d = 0; d < S; d++) added += part_absorb(t, d, t->res);
  }
  return added;
}

void pqo_part_apply_rotations(pqo_part *t, const uint64_t *gens, const double *angles, size_t count,
                              uint64_t *n_after) {
  for (size_t i = 0; i < count; i++) {
    pqo_part_apply_rotation(t, gens + i * (size_t)t->W, angles[i]);
    n_after[i] = pqo_part_len(t);
  }
}

void pqo_part_scale_by_table(pqo_part *t, const double *table, size_t len) {
#pragma omp parallel for num_threads(t->threads) schedule(static, 1)
  for (int q = 0; q < t->S; q++) pqo_scale_by_table(t->p[q], table, len);
}

size_t pqo_part_reclaim(pqo_part *t) {
  size_t gone = 0;
#pragma omp parallel for num_threads(t->threads) schedule(static, 1) reduction(+ : gone)
  for (int q = 0; q < t->S; q++) gone += pqo_reclaim(t->p[q], NULL);
  return gone;
}

/* partitioned_termsum.rs:676-697: per-partition sums, added in partition order */
double pqo_part_expectation(const pqo_part *t, const uint64_t *diag, int n_diag) {
  double acc = 0.0;
  uint64_t img[PQO_MAXW];
  for (int q = 0; q < t->S; q++) {
    const pqo_engine *e = t->p[q];
    double part = 0.0;
    for (size_t i = 0; i < e->n; i++) {
      const double s = tab_conjugate(&t->frame, t->kind, t->n_units, &e->keys[i * t->W], img);
      part += e->coeffs[i] * s * trace(t->kind, img, t->W, t->n_units, diag, n_diag);
    }
    acc += part;
  }
  return acc;
}

size_t pqo_part_terms_below(const pqo_part *t, double thr) {
  size_t n = 0;
  for (int q = 0; q < t->S; q++) n += pqo_terms_below(t->p[q], thr);
  return n;
}

/* Rows of every partition in partition order, the shared frame applied. */
void pqo_part_export(const pqo_part *t, uint64_t *keys, double *coeffs) {
  size_t o = 0;
  for (int q = 0; q < t->S; q++) {
    const pqo_engine *e = t->p[q];
    for (size_t i = 0; i < e->n; i++, o++) {
      const double s = tab_conjugate(&t->frame, t->kind, t->n_units, &e->keys[i * t->W], &keys[o * t->W]);
      coeffs[o] = e->coeffs[i] * s;
    }
  }
}

void pqo_part_counters(const pqo_part *t, uint64_t *out4, double *disc2) {
  memset(out4, 0, 4 * sizeof(uint64_t));
  disc2[0] = disc2[1] = 0.0;
  for (int q = 0; q < t->S; q++) {
    const pqo_engine *e = t->p[q];
    out4[0] += e->visited;
    out4[1] += e->emitted;
    out4[2] += e->declined;
    out4[3] += e->hits;
    disc2[0] += e->disc_l1;
    if (e->disc_max > disc2[1]) disc2[1] = e->disc_max;
  }
}

/* Owner partition of a key, and partition q's row count (test hooks for the
 * ownership pins; export lists the partitions' rows in partition order). */
int pqo_part_owner(const pqo_part *t, const uint64_t *k) { return owner_of(t, k); }
size_t pqo_part_partition_len(const pqo_part *t, int q) { return q >= 0 && q < t->S ? t->p[q]->n : 0; }

/* ------------------------------------------------------------ surrogate tape */

/* Pairwise (adjacent pairs, level by level, zero padded) sum of v[0..n),
 * in place. */
static double pairwise_sum(double *v, size_t n) {
  if (n == 0) return 0.0;
  while (n > 1) {
    size_t m = (n + 1) / 2;
    for (size_t i = 0; i < m; ++i) v[i] = v[2 * i] + (2 * i + 1 < n ? v[2 * i + 1] : 0.0);
    n = m;
  }
  return v[0];
}

/* CompiledCoeff::evaluate_into (crates/surrogate/src/symcoeff.rs:836-850)
 * for each of n_sets parameter rows (stride doubles apart), one set after
 * another on one core as SurrogateModel::evaluate_batch's per-set closure
 * (model.rs:80-103); the term contributions overlap * value are then summed
 * pairwise in term order.  kinds 0..4 = Scalar, Add, Scale, Cos, Sin; roots
 * < 0 mark empty coefficients.  Returns 0, or -1 when out of memory. */
int pqo_tape_eval(size_t n_ops, const uint8_t *kinds, const uint32_t *params, const uint64_t *a,
                  const uint64_t *b, const double *vals, size_t n_terms, const double *overlaps,
                  const int64_t *roots, size_t n_params, size_t n_sets, size_t stride,
                  const double *sets, double *out) {
  double *res = malloc((n_ops ? n_ops : 1) * sizeof(double));
  double *lut = malloc((2 * n_params + 1) * sizeof(double));
  double *acc = malloc((n_terms ? n_terms : 1) * sizeof(double));
  if (!res || !lut || !acc) {
    free(res);
    free(lut);
    free(acc);
    return -1;
  }
  for (size_t s = 0; s < n_sets; ++s) {
    const double *th = sets + s * stride;
    /* separate cos and sin calls, as the engine's: two loops keep the
     * compiler from fusing them into sincos, whose results differ from
     * theirs in the last bit for about one angle in a thousand */
    for (size_t p = 0; p < n_params; ++p) lut[2 * p] = cos(th[p]);
    for (size_t p = 0; p < n_params; ++p) lut[2 * p + 1] = sin(th[p]);
    for (size_t i = 0; i < n_ops; ++i) {
      switch (kinds[i]) {
        case 0: res[i] = vals[i]; break;
        case 1: res[i] = res[a[i]] + res[b[i]]; break;
        case 2: res[i] = vals[i] * res[a[i]]; break;
        case 3: res[i] = lut[2 * params[i]] * res[a[i]]; break;
        default: res[i] = lut[2 * params[i] + 1] * res[a[i]]; break;
      }
    }
    for (size_t t = 0; t < n_terms; ++t) acc[t] = roots[t] < 0 ? 0.0 : overlaps[t] * res[roots[t]];
    out[s] = pairwise_sum(acc, n_terms);
  }
  free(res);
  free(lut);
  free(acc);
  return 0;
}

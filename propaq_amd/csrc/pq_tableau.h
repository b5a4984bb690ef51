// pq_tableau.h -- the deferred Clifford frame (a stabilizer tableau of
// 2 n_units rows), shared by the propagation engine and the surrogate builder.
//
// Reference: CliffordTableau, crates/core/src/algebra/tableau.rs:29-224.
#pragma once
#include <utility>
#include <vector>

#include "pq_algebra.h"

namespace pq {

// Deferred Clifford frame: readout rows P -> C^dag P C and generator rows
// P -> C P C^dag, one (image, sign) per bit position (tableau.rs:29-45).
template <int KIND, int W>
struct Tableau {
  int n_rows = 0;
  std::vector<Key<W>> ro, ge;
  std::vector<double> ro_s, ge_s;
  bool identity = true;

  static Key<W> unit_key(int p) {
    Key<W> k{};
    if (p < 64 * W) k.w[p / 64] = 1ULL << (p % 64);
    return k;
  }

  void init(int n_units) {
    n_rows = 2 * n_units;
    ro.assign(n_rows, Key<W>{});
    ge.assign(n_rows, Key<W>{});
    ro_s.assign(n_rows, 1.0);
    ge_s.assign(n_rows, 1.0);
    for (int p = 0; p < n_rows; ++p) ro[p] = ge[p] = unit_key(p);
    identity = true;
  }

  void refresh_identity() {
    identity = true;
    for (int p = 0; p < n_rows && identity; ++p)
      if (!key_eq<W>(ro[p], unit_key(p)) || ro_s[p] != 1.0) identity = false;
  }

  // tableau.rs:109 apply_rows
  static double apply_rows(const std::vector<Key<W>>& img, const std::vector<double>& sg, int n_rows,
                           const Key<W>& m, Key<W>& out) {
    Key<W> src{}, im{};
    int se = 0, ie = 0;
    for (int wd = 0; wd < W; ++wd) {
      uint64_t bits = m.w[wd];
      while (bits) {
        const int b = __builtin_ctzll(bits);
        bits &= bits - 1;
        const int p = wd * 64 + b;
        if (p >= n_rows) continue;
        const Key<W> e1 = unit_key(p);
        se += product_exp<KIND, W>(e1, src);
        src.w[wd] ^= e1.w[wd];
        ie += product_exp<KIND, W>(img[p], im) + (sg[p] < 0.0 ? 2 : 0);
        for (int i = 0; i < W; ++i) im.w[i] ^= img[p].w[i];
      }
    }
    out = im;
    return (((ie - se) % 4 + 4) % 4 == 2) ? -1.0 : 1.0;
  }

  double conjugate(const Key<W>& m, Key<W>& out) const {
    if (identity) {
      out = m;
      return 1.0;
    }
    return apply_rows(ro, ro_s, n_rows, m, out);
  }

  double conjugate_generator(const Key<W>& m, Key<W>& out) const {
    if (identity) {
      out = m;
      return 1.0;
    }
    return apply_rows(ge, ge_s, n_rows, m, out);
  }

  // for_rotation followed by compose (tableau.rs:146-209) for one rotation
  // exp(-i angle gen / 2) at a Clifford angle, touching only what the
  // rotation moves.  Readout ("this, then the rotation"): a row image that
  // anticommutes with gen becomes gen * image, its sign times
  // sin(angle) (-Im phase(gen, image)); images that commute are unchanged.
  // Generator rows (the rotation's inverse, then this): only the unit
  // positions that anticommute with gen get new images.  Signs stay +-1.
  void compose_rotation(const Key<W>& gen, double angle) {
    const double s_t = std::sin(angle);
    for (int p = 0; p < n_rows; ++p) {
      if (!anticommute<KIND, W>(gen, ro[p])) continue;
      const double s = s_t * neg_im(product_exp<KIND, W>(gen, ro[p]));
      for (int i = 0; i < W; ++i) ro[p].w[i] ^= gen.w[i];
      if (s < 0.0) ro_s[p] = -ro_s[p];
    }
    std::vector<std::pair<int, Key<W>>> upd;
    std::vector<double> upd_s;
    for (int p = 0; p < n_rows && p < 64 * W; ++p) {
      const Key<W> e1 = unit_key(p);
      if (!anticommute<KIND, W>(gen, e1)) continue;
      Key<W> m, img;
      for (int i = 0; i < W; ++i) m.w[i] = gen.w[i] ^ e1.w[i];
      const double s = -(s_t * neg_im(product_exp<KIND, W>(gen, e1)));
      upd_s.push_back(apply_rows(ge, ge_s, n_rows, m, img) * s);
      upd.emplace_back(p, img);
    }
    for (size_t k = 0; k < upd.size(); ++k) {
      ge[upd[k].first] = upd[k].second;
      ge_s[upd[k].first] = upd_s[k];
    }
    refresh_identity();
  }

  // changes_weight (tableau.rs:215) of the rotation's own tableau: a unit
  // position that anticommutes with gen maps to gen * e_p, whose support
  // must still be one unit.
  static bool rotation_changes_weight(int n_rows, const Key<W>& gen) {
    for (int p = 0; p < n_rows && p < 64 * W; ++p) {
      const Key<W> e1 = unit_key(p);
      if (!anticommute<KIND, W>(gen, e1)) continue;
      Key<W> m;
      for (int i = 0; i < W; ++i) m.w[i] = gen.w[i] ^ e1.w[i];
      if (support<W>(m) != support<W>(e1)) return true;
    }
    return false;
  }
};

}  // namespace pq

// tiled.cpp — host TiledHamiltonian: the tables the device sweep walks.
//
// Builds, from a MultipoleTerm list, the reference's flat tables
// (src/hamiltonian.jl:234-266) with the same ordering contracts, because the
// colouring, the CSR order and the program order are part of the bit-exact
// schedule:
//   * instances cell-outer (c3, c2, c1; c1 fastest), term-inner     :333-353
//   * site CSR by ascending instance id, member slot recorded        :355-375
//   * site programs per (template, member slot), CartesianIndices
//     column-major over `folded`, skipping coef*folded == 0;
//     energy programs per template skipping folded == 0              :99-133
//   * pair / triplet / general tags in site_col                      :138-181
//   * greedy colouring in ascending site order                       :456-499
//   * Zeeman body-1 templates appended after the fitted terms        :437-448
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sce_hip.h"
#include "sce_internal.h"

namespace {

struct Term {
  int body = 0;
  double coef = 0.0;
  std::vector<int> atoms, ls;
  std::vector<std::array<int, 3>> shifts;
  std::vector<double> folded;  // column-major
};

int wrap(int x, int d) {
  int r = x % d;
  return r < 0 ? r + d : r;
}

}  // namespace

struct sce_tiled {
  int n_cell_atoms = 0, lmax = 0, nlm = 0, n_sites = 0, n_active = 0, n_colors = 0;
  int n_fitted = 0;
  int dims[3] = {1, 1, 1};
  std::vector<Term> terms;
  std::vector<int32_t> inst_term, inst_ptr, inst_sites, site_ptr, site_inst, site_prog,
      sprog_ptr, sent_tgt, sfac_ptr, sfac_row, site_col, site_col2, pent_row, pent_row2,
      eprog_ptr, efac_ptr, efac_row, color_ptr, color_sites;
  std::vector<int8_t> site_slot, sfac_slot;
  std::vector<uint8_t> site_has_l1, site_active;
  std::vector<double> sent_w, eent_w, term_coef;
};

static int lm_index(int l, int m) { return l * l + l + m + 1; }

// One template's site programs and energy program (hamiltonian.jl:99-133).
static void append_programs(sce_tiled* t, const Term& tm) {
  const int D = tm.body;
  std::vector<int> ext(D);
  int total = 1;
  for (int k = 0; k < D; ++k) {
    ext[k] = 2 * tm.ls[k] + 1;
    total *= ext[k];
  }
  std::vector<int> mu(D);  // μ_k = idx_k - l_k (0-based idx)
  auto decode = [&](int lin) {
    for (int k = 0; k < D; ++k) {
      mu[k] = lin % ext[k] - tm.ls[k];
      lin /= ext[k];
    }
  };
  for (int v = 0; v < D; ++v) {
    for (int lin = 0; lin < total; ++lin) {
      const double w = tm.coef * tm.folded[lin];
      if (w == 0.0) continue;
      decode(lin);
      t->sent_w.push_back(w);
      t->sent_tgt.push_back(lm_index(tm.ls[v], mu[v]));
      int32_t rows[8];
      int nr = 0;
      for (int k = 0; k < D; ++k) {
        if (k == v) continue;
        rows[nr++] = lm_index(tm.ls[k], mu[k]);
        t->sfac_row.push_back(rows[nr - 1]);
        t->sfac_slot.push_back((int8_t)(k + 1));
      }
      t->sfac_ptr.push_back((int32_t)t->sfac_row.size() + 1);
      t->pent_row.push_back(D == 2 ? rows[0] : D == 3 ? rows[0] : 0);
      t->pent_row2.push_back(D == 3 ? rows[1] : 0);
    }
    t->sprog_ptr.push_back((int32_t)t->sent_w.size() + 1);
  }
  for (int lin = 0; lin < total; ++lin) {
    const double w = tm.folded[lin];
    if (w == 0.0) continue;
    decode(lin);
    t->eent_w.push_back(w);
    for (int k = 0; k < D; ++k) t->efac_row.push_back(lm_index(tm.ls[k], mu[k]));
    t->efac_ptr.push_back((int32_t)t->efac_row.size() + 1);
  }
  t->eprog_ptr.push_back((int32_t)t->eent_w.size() + 1);
}

extern "C" int32_t sce_tiled_build(int32_t na, int32_t n_terms, const int32_t* body,
                                   const double* coef, const int32_t* atoms,
                                   const int32_t* shifts, const int32_t* ls,
                                   const double* folded, const int32_t dims[3],
                                   const double* magmoms, const double* field,
                                   sce_tiled** out) {
  if (!out) return sce_fail(SCE_EINVAL, "sce_tiled_build: out is NULL");
  *out = nullptr;
  if (na < 1) return sce_fail(SCE_EINVAL, "n_cell_atoms must be >= 1; got %d", na);
  const int d[3] = {dims[0], dims[1], dims[2]};
  if (d[0] < 1 || d[1] < 1 || d[2] < 1)
    return sce_fail(SCE_EINVAL, "dims must be >= 1; got (%d, %d, %d)", d[0], d[1], d[2]);
  if (field && !magmoms)
    return sce_fail(SCE_EINVAL, "field given without magmoms: the Zeeman energy needs moment magnitudes");
  const double FOURPI = 4.0 * 3.141592653589793;
  auto* t = new sce_tiled();
  t->n_cell_atoms = na;
  std::memcpy(t->dims, d, sizeof d);
  int lmax = 0;
  size_t ao = 0, fo = 0;
  for (int k = 0; k < n_terms; ++k) {
    const int D = body[k];
    if (D < 1 || D > 8) {
      delete t;
      return sce_fail(SCE_EINVAL, "term %d: body %d unsupported", k + 1, D);
    }
    size_t total = 1;
    for (int i = 0; i < D; ++i) total *= (size_t)(2 * ls[ao + i] + 1);
    if (coef[k] != 0.0) {  // hamiltonian.jl:280
      Term tm;
      tm.body = D;
      for (int i = 0; i < D; ++i) {
        tm.atoms.push_back(atoms[ao + i]);
        tm.ls.push_back(ls[ao + i]);
        tm.shifts.push_back({shifts[3 * (ao + i)], shifts[3 * (ao + i) + 1], shifts[3 * (ao + i) + 2]});
        if (atoms[ao + i] < 1 || atoms[ao + i] > na) {
          delete t;
          return sce_fail(SCE_EINVAL, "term %d: atoms outside 1:%d", k + 1, na);
        }
        if (ls[ao + i] < 0) {
          delete t;
          return sce_fail(SCE_EINVAL, "term %d: negative angular momentum", k + 1);
        }
        lmax = std::max(lmax, ls[ao + i]);
      }
      if (tm.shifts[0][0] || tm.shifts[0][1] || tm.shifts[0][2]) {
        delete t;
        return sce_fail(SCE_EINVAL, "term %d: shifts[1] != 0 breaks the home-cell anchoring", k + 1);
      }
      for (int i = 0; i < D; ++i)
        for (int j = i + 1; j < D; ++j) {
          bool same = tm.atoms[i] == tm.atoms[j];
          for (int x = 0; x < 3; ++x) same = same && wrap(tm.shifts[i][x], d[x]) == wrap(tm.shifts[j][x], d[x]);
          if (same) {
            delete t;
            return sce_fail(SCE_EINVAL,
                            "term %d folds two member sites onto one supercell site under "
                            "dims = (%d, %d, %d); enlarge dims", k + 1, d[0], d[1], d[2]);
          }
        }
      tm.coef = coef[k] * std::pow(FOURPI, (double)D / 2.0);  // hamiltonian.jl:316
      tm.folded.assign(folded + fo, folded + fo + total);
      t->terms.push_back(std::move(tm));
    }
    ao += (size_t)D;
    fo += total;
  }
  t->n_fitted = (int)t->terms.size();
  const bool has_field = field && (field[0] != 0.0 || field[1] != 0.0 || field[2] != 0.0);
  if (magmoms && has_field) {  // hamiltonian.jl:437-448
    const double MU_B_EV_T = 9.2740100783e-24 / 1.602176634e-19;
    const double N1 = std::sqrt(3.0 / FOURPI);
    for (int a = 0; a < na; ++a) {
      const double m = magmoms[a];
      if (!(std::isfinite(m) && m >= 0.0)) {
        delete t;
        return sce_fail(SCE_EINVAL, "magmoms must be finite and >= 0");
      }
      if (m == 0.0) continue;
      Term tm;
      tm.body = 1;
      tm.coef = 1.0;
      tm.atoms = {a + 1};
      tm.ls = {1};
      tm.shifts = {{0, 0, 0}};
      const double s = -(MU_B_EV_T * m / N1);
      tm.folded = {s * field[1], s * field[2], s * field[0]};
      t->terms.push_back(std::move(tm));
    }
    if ((int)t->terms.size() > t->n_fitted) lmax = std::max(lmax, 1);
  }
  if (t->terms.empty()) {
    delete t;
    return sce_fail(SCE_EINVAL, "the term list is empty (no nonzero coefficients and no Zeeman term)");
  }
  if (lmax > 6) {
    delete t;
    return sce_fail(SCE_EINVAL, "lmax = %d unsupported (<= 6)", lmax);
  }
  t->lmax = lmax;
  t->nlm = (lmax + 1) * (lmax + 1);
  const int ncells = d[0] * d[1] * d[2];
  const int N = na * ncells;
  t->n_sites = N;
  const int nt = (int)t->terms.size();

  // instances
  t->inst_ptr.push_back(1);
  for (int c3 = 0; c3 < d[2]; ++c3)
    for (int c2 = 0; c2 < d[1]; ++c2)
      for (int c1 = 0; c1 < d[0]; ++c1)
        for (int k = 0; k < nt; ++k) {
          const Term& tm = t->terms[k];
          t->inst_term.push_back(k + 1);
          for (int i = 0; i < tm.body; ++i) {
            const int w1 = wrap(c1 + tm.shifts[i][0], d[0]);
            const int w2 = wrap(c2 + tm.shifts[i][1], d[1]);
            const int w3 = wrap(c3 + tm.shifts[i][2], d[2]);
            t->inst_sites.push_back(tm.atoms[i] + na * (w1 + d[0] * (w2 + d[1] * w3)));
          }
          t->inst_ptr.push_back((int32_t)t->inst_sites.size() + 1);
        }
  const int n_inst = (int)t->inst_term.size();
  const size_t tot = t->inst_sites.size();

  // site CSR
  t->site_ptr.assign((size_t)N + 1, 0);
  for (int32_t s : t->inst_sites) t->site_ptr[s]++;
  t->site_ptr[0] = 1;
  for (int s = 0; s < N; ++s) t->site_ptr[s + 1] += t->site_ptr[s];
  t->site_inst.resize(tot);
  t->site_slot.resize(tot);
  {
    std::vector<int32_t> cur(t->site_ptr.begin(), t->site_ptr.end() - 1);
    for (int i = 0; i < n_inst; ++i)
      for (int q = t->inst_ptr[i]; q < t->inst_ptr[i + 1]; ++q) {
        const int s = t->inst_sites[q - 1];
        const int32_t pos = cur[s - 1]++;
        t->site_inst[pos - 1] = i + 1;
        t->site_slot[pos - 1] = (int8_t)(q - t->inst_ptr[i] + 1);
      }
  }
  t->site_has_l1.assign(N, 0);
  t->site_active.assign(N, 0);
  t->n_active = 0;
  for (int s = 0; s < N; ++s) {
    for (int j = t->site_ptr[s]; j < t->site_ptr[s + 1]; ++j) {
      const Term& tm = t->terms[t->inst_term[t->site_inst[j - 1] - 1] - 1];
      if (tm.ls[t->site_slot[j - 1] - 1] == 1) t->site_has_l1[s] = 1;
    }
    t->site_active[s] = t->site_ptr[s + 1] > t->site_ptr[s] ? 1 : 0;
    t->n_active += t->site_active[s];
  }

  // programs
  t->sprog_ptr.push_back(1);
  t->sfac_ptr.push_back(1);
  t->eprog_ptr.push_back(1);
  t->efac_ptr.push_back(1);
  std::vector<int> pbase(nt);
  std::vector<std::array<int8_t, 2>> hoist;  // per pid: the fast paths' other slots
  int np = 0;
  for (int k = 0; k < nt; ++k) {
    const int D = t->terms[k].body;
    pbase[k] = np;
    np += D;
    for (int v = 1; v <= D; ++v) {
      if (D == 2) hoist.push_back({(int8_t)(3 - v), 0});
      else if (D == 3) hoist.push_back({(int8_t)(v == 1 ? 2 : 1), (int8_t)(v == 3 ? 2 : 3)});
      else hoist.push_back({0, 0});
    }
    append_programs(t, t->terms[k]);
    t->term_coef.push_back(t->terms[k].coef);
  }
  t->site_prog.resize(tot);
  t->site_col.resize(tot);
  t->site_col2.resize(tot);
  for (size_t j = 0; j < tot; ++j) {
    const int inst = t->site_inst[j];
    const int pid = pbase[t->inst_term[inst - 1] - 1] + t->site_slot[j];
    t->site_prog[j] = pid;
    const int off = t->inst_ptr[inst - 1] - 1;
    const int sl1 = hoist[pid - 1][0], sl2 = hoist[pid - 1][1];
    if (sl1 == 0) {
      t->site_col[j] = 0;
    } else {
      const int32_t c1 = t->inst_sites[off + sl1 - 1];
      t->site_col[j] = sl2 == 0 ? c1 : -c1;
    }
    t->site_col2[j] = sl2 == 0 ? 0 : t->inst_sites[off + sl2 - 1];
  }

  // greedy colouring in ascending site order
  std::vector<int32_t> colour(N, 0);
  std::vector<int32_t> taken;  // taken[c-1] == s: colour c held by a conflictor of s
  int ncol = 0;
  for (int s = 1; s <= N; ++s) {
    if (!t->site_active[s - 1]) continue;
    for (int j = t->site_ptr[s - 1]; j < t->site_ptr[s]; ++j) {
      const int i = t->site_inst[j - 1];
      for (int q = t->inst_ptr[i - 1]; q < t->inst_ptr[i]; ++q) {
        const int c = colour[t->inst_sites[q - 1] - 1];
        if (c > 0) taken[c - 1] = s;
      }
    }
    int c = 1;
    while (c <= ncol && taken[c - 1] == s) ++c;
    if (c > ncol) {
      ncol = c;
      taken.push_back(0);
    }
    colour[s - 1] = c;
  }
  t->n_colors = ncol;
  t->color_ptr.assign((size_t)ncol + 1, 0);
  for (int s = 0; s < N; ++s)
    if (colour[s] > 0) t->color_ptr[colour[s]]++;
  t->color_ptr[0] = 1;
  for (int c = 0; c < ncol; ++c) t->color_ptr[c + 1] += t->color_ptr[c];
  t->color_sites.resize(t->n_active);
  {
    std::vector<int32_t> cur(t->color_ptr.begin(), t->color_ptr.end() - 1);
    for (int s = 1; s <= N; ++s)
      if (colour[s - 1] > 0) t->color_sites[cur[colour[s - 1] - 1]++ - 1] = s;
  }
  *out = t;
  return SCE_OK;
}

extern "C" int32_t sce_tiled_desc(const sce_tiled* t, sce_tables_desc* o) {
  if (!t || !o) return sce_fail(SCE_EINVAL, "sce_tiled_desc: NULL argument");
  std::memset(o, 0, sizeof *o);
  o->n_cell_atoms = t->n_cell_atoms;
  std::memcpy(o->dims, t->dims, sizeof t->dims);
  o->n_sites = t->n_sites;
  o->lmax = t->lmax;
  o->nlm = t->nlm;
  o->n_terms = (int32_t)t->terms.size();
  o->n_inst = (int32_t)t->inst_term.size();
  o->n_active = t->n_active;
  o->n_colors = t->n_colors;
  o->n_prog = (int32_t)t->sprog_ptr.size() - 1;
  o->n_inst_sites = (int64_t)t->inst_sites.size();
  o->n_sent = (int64_t)t->sent_w.size();
  o->n_sfac = (int64_t)t->sfac_row.size();
  o->n_eent = (int64_t)t->eent_w.size();
  o->n_efac = (int64_t)t->efac_row.size();
  o->inst_term = t->inst_term.data();
  o->inst_ptr = t->inst_ptr.data();
  o->inst_sites = t->inst_sites.data();
  o->site_ptr = t->site_ptr.data();
  o->site_inst = t->site_inst.data();
  o->site_slot = t->site_slot.data();
  o->site_has_l1 = t->site_has_l1.data();
  o->site_active = t->site_active.data();
  o->site_prog = t->site_prog.data();
  o->sprog_ptr = t->sprog_ptr.data();
  o->sent_w = t->sent_w.data();
  o->sent_tgt = t->sent_tgt.data();
  o->sfac_ptr = t->sfac_ptr.data();
  o->sfac_row = t->sfac_row.data();
  o->sfac_slot = t->sfac_slot.data();
  o->site_col = t->site_col.data();
  o->site_col2 = t->site_col2.data();
  o->pent_row = t->pent_row.data();
  o->pent_row2 = t->pent_row2.data();
  o->eprog_ptr = t->eprog_ptr.data();
  o->eent_w = t->eent_w.data();
  o->efac_ptr = t->efac_ptr.data();
  o->efac_row = t->efac_row.data();
  o->term_coef = t->term_coef.data();
  o->color_ptr = t->color_ptr.data();
  o->color_sites = t->color_sites.data();
  return SCE_OK;
}

extern "C" int32_t sce_tiled_free(sce_tiled* t) {
  delete t;
  return SCE_OK;
}

extern "C" int32_t sce_host_zrows(int32_t lmax, int64_t n, const double* config, double* zrows) {
  if (lmax < 0 || lmax > 6) return sce_fail(SCE_EINVAL, "lmax = %d unsupported (<= 6)", lmax);
  const int nlm = (lmax + 1) * (lmax + 1);
  for (int64_t s = 0; s < n; ++s) {
    const double* e = config + 3 * s;
    sce::zlm_row_dyn(lmax, e[0], e[1], e[2], zrows + s * nlm);
  }
  return SCE_OK;
}

// _total_energy (src/energy.jl:55-75): serial over instances in instance order.
extern "C" int32_t sce_host_total_energy(const sce_tables_desc* d, const double* zrows, double* E) {
  if (!d || !zrows || !E) return sce_fail(SCE_EINVAL, "sce_host_total_energy: NULL argument");
  double acc = 0.0;
  const int nlm = d->nlm;
  for (int i = 0; i < d->n_inst; ++i) {
    const int k = d->inst_term[i];
    const int off = d->inst_ptr[i] - 1;
    double Ei = 0.0;
    for (int e = d->eprog_ptr[k - 1]; e < d->eprog_ptr[k]; ++e) {
      double p = 1.0;
      const int f0 = d->efac_ptr[e - 1] - 1;
      const int nf = d->efac_ptr[e] - 1 - f0;
      for (int m = 0; m < nf; ++m)
        p *= zrows[(size_t)(d->inst_sites[off + m] - 1) * nlm + (d->efac_row[f0 + m] - 1)];
      Ei += d->eent_w[e - 1] * p;
    }
    acc += d->term_coef[k - 1] * Ei;
  }
  *E = acc;
  return SCE_OK;
}

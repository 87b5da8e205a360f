// layout.cpp — per-sublattice walk layout of a tiled Hamiltonian (see
// sce_internal.h).  The factor products it encodes are those of site_coeffs!
// (src/energy.jl:113-136): pair p = z, triplet p = z1*z2, general
// p = ((1*z1)*z2)*z3 — the device evaluates (Z[f0]*Z[f1])*Z[f2] with unused
// factors reading 1.0, which is bit-identical for every body <= 4.
#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>

#include "sce_internal.h"

namespace {

struct Geo {
  int na, d[3];
  void cell_of(int s0, int c[3]) const {  // s0: 0-based site
    int cell = s0 / na;
    c[0] = cell % d[0];
    c[1] = (cell / d[0]) % d[1];
    c[2] = cell / (d[0] * d[1]);
  }
  int site(int atom0, const int c[3]) const {
    return atom0 + na * (c[0] + d[0] * (c[1] + d[1] * c[2]));
  }
};

// Factor sites of adjacency entry j (1-based), in the product order of the
// entry's program: pair [col]; triplet [-col, col2]; general the instance's
// members at the factor slots (ascending slot, all but the site's own).
int factor_sites(const sce_tables_desc* d, int j, int sites[8]) {
  const int col = d->site_col[j - 1];
  if (col > 0) {
    sites[0] = col;
    return 1;
  }
  if (col < 0) {
    sites[0] = -col;
    sites[1] = d->site_col2[j - 1];
    return 2;
  }
  const int inst = d->site_inst[j - 1];
  const int off = d->inst_ptr[inst - 1] - 1;
  const int body = d->inst_ptr[inst] - d->inst_ptr[inst - 1];
  const int own = d->site_slot[j - 1];
  int n = 0;
  for (int k = 1; k <= body; ++k)
    if (k != own) sites[n++] = d->inst_sites[off + k - 1];
  return n;
}

}  // namespace

int32_t sce_derive_layout(const sce_tables_desc* d, HostLayout* L, bool verify) {
  const Geo g{d->n_cell_atoms, {d->dims[0], d->dims[1], d->dims[2]}};
  const int na = g.na, nlm = d->nlm, N = d->n_sites;
  if (na * g.d[0] * g.d[1] * g.d[2] != N)
    return sce_fail(SCE_ELAYOUT, "n_sites %d != n_cell_atoms * prod(dims)", N);
  L->na = na;
  L->walk_ptr.assign(na + 1, 0);
  L->nb_ptr.assign(na + 1, 0);
  L->prog_ptr.assign(na + 1, 0);
  L->prog_off.clear();
  L->walk.clear();
  L->nb.clear();
  L->feat_ptr.assign(na + 1, 0);
  L->feat.clear();
  L->nfeat_max = 0;
  L->nb_max = 0;
  L->w_max = 0;
  L->rep_ptr.assign(na + 1, 0);
  L->rep_pid.clear();
  L->rep_cls.clear();
  L->rep_nb.clear();
  // representative site of sublattice a: atom a in cell 0 (0-based site id a)
  std::vector<std::vector<int>> rep_js(na);      // pid-sorted adjacency of the rep
  std::vector<std::vector<int>> rep_nbidx(na);   // per rep adjacency entry: nb index per factor
  auto sorted_adj = [&](int s0) {
    std::vector<int> js;
    for (int j = d->site_ptr[s0]; j < d->site_ptr[s0 + 1]; ++j) js.push_back(j);
    std::stable_sort(js.begin(), js.end(),
                     [&](int a, int b) { return d->site_prog[a - 1] < d->site_prog[b - 1]; });
    return js;
  };
  for (int a = 0; a < na; ++a) {
    L->walk_ptr[a] = (int32_t)L->walk.size();
    L->nb_ptr[a] = (int32_t)(L->nb.size() / 4);
    L->feat_ptr[a] = (int32_t)(L->feat.size() / 4);
    L->prog_ptr[a] = (int32_t)L->prog_off.size();
    L->rep_ptr[a] = (int32_t)L->rep_pid.size();
    if (!d->site_active[a]) continue;
    std::map<std::tuple<int, int, int, int>, int> nbmap;
    std::vector<std::array<int, 4>> nbl;
    auto nb_index = [&](int site1) {
      int c[3];
      g.cell_of(site1 - 1, c);
      const auto key = std::make_tuple((site1 - 1) % na, c[0], c[1], c[2]);
      auto it = nbmap.find(key);
      if (it != nbmap.end()) return it->second;
      const int k = (int)nbl.size();
      nbmap.emplace(key, k);
      nbl.push_back({std::get<0>(key), c[0], c[1], c[2]});
      return k;
    };
    rep_js[a] = sorted_adj(a);
    for (int j : rep_js[a]) {
      int fs[8];
      const int nf = factor_sites(d, j, fs);
      if (nf > 3)
        return sce_fail(SCE_EINVAL, "cluster body %d > 4 is not supported by the device walk", nf + 1);
      for (int f = 0; f < nf; ++f) rep_nbidx[a].push_back(nb_index(fs[f]));
      rep_nbidx[a].push_back(-1);  // separator
    }
    const int NB = (int)nbl.size();
    if ((NB + 1) * nlm + 1 > 65535)
      return sce_fail(SCE_EINVAL, "neighbour stencil of %d sites too large for the device walk", NB);
    const uint16_t ONE = (uint16_t)(NB * nlm);
    size_t pos = 0;
    for (int j : rep_js[a]) {
      int nbi[3];
      int nf = 0;
      while (rep_nbidx[a][pos] >= 0) nbi[nf++] = rep_nbidx[a][pos++];
      ++pos;
      const int pid = d->site_prog[j - 1];
      const int col = d->site_col[j - 1];
      L->prog_off.push_back((int32_t)(L->walk.size() - L->walk_ptr[a]));
      for (int e = d->sprog_ptr[pid - 1]; e < d->sprog_ptr[pid]; ++e) {
        WalkEntry we;
        we.w = d->sent_w[e - 1];
        we.tgt = (uint16_t)(d->sent_tgt[e - 1] - 1);
        int rows[3];
        int nr = 0;
        if (col > 0) {
          rows[nr++] = d->pent_row[e - 1];
        } else if (col < 0) {
          rows[nr++] = d->pent_row[e - 1];
          rows[nr++] = d->pent_row2[e - 1];
        } else {
          for (int f = d->sfac_ptr[e - 1]; f < d->sfac_ptr[e]; ++f) rows[nr++] = d->sfac_row[f - 1];
        }
        if (nr != nf) return sce_fail(SCE_ELAYOUT, "program %d: factor count mismatch", pid);
        uint16_t fi[3] = {ONE, ONE, ONE};
        for (int f = 0; f < nr; ++f) fi[f] = (uint16_t)(nbi[f] * nlm + rows[f] - 1);
        we.f0 = fi[0];
        we.f1 = fi[1];
        we.f2 = fi[2];
        L->walk.push_back(we);
      }
    }
    L->prog_off.push_back((int32_t)(L->walk.size() - L->walk_ptr[a]));
    pos = 0;
    for (int j : rep_js[a]) {
      int nbi[3] = {-1, -1, -1};
      int nf = 0;
      while (rep_nbidx[a][pos] >= 0) nbi[nf++] = rep_nbidx[a][pos++];
      ++pos;
      const int col = d->site_col[j - 1];
      L->rep_pid.push_back(d->site_prog[j - 1]);
      L->rep_cls.push_back((int8_t)(col > 0 ? 1 : col < 0 ? 2 : 0));
      L->rep_nb.insert(L->rep_nb.end(), {nf, nbi[0], nbi[1], nbi[2]});
    }
    for (auto& v : nbl) L->nb.insert(L->nb.end(), v.begin(), v.end());
    L->nb_max = std::max(L->nb_max, NB);
    // descriptor cluster specs, in walk order
    L->feat_ptr[a] = (int32_t)(L->feat.size() / 4);
    bool body1 = false;
    pos = 0;
    for (int j : rep_js[a]) {
      int nbi[3];
      int nf = 0;
      while (rep_nbidx[a][pos] >= 0) nbi[nf++] = rep_nbidx[a][pos++];
      ++pos;
      if (d->site_col[j - 1] < 0) {
        L->feat.insert(L->feat.end(), {1, nbi[0], nbi[1], 0});
      } else if (d->site_col[j - 1] == 0 && nf == 3) {
        L->feat.insert(L->feat.end(), {2, nbi[0], nbi[1], nbi[2]});
      } else if (d->site_col[j - 1] == 0 && nf == 0) {
        body1 = true;
      }
    }
    if (body1)
      for (int x = 0; x < 3; ++x) L->feat.insert(L->feat.end(), {3, x, 0, 0});
    const int nspec = (int)(L->feat.size() / 4) - L->feat_ptr[a];
    L->nfeat_max = std::max(L->nfeat_max, 2 * NB + nspec);
  }
  L->walk_ptr[na] = (int32_t)L->walk.size();
  L->nb_ptr[na] = (int32_t)(L->nb.size() / 4);
  L->prog_ptr[na] = (int32_t)L->prog_off.size();
  L->feat_ptr[na] = (int32_t)(L->feat.size() / 4);
  L->rep_ptr[na] = (int32_t)L->rep_pid.size();
  for (int a = 0; a < na; ++a) L->w_max = std::max(L->w_max, L->walk_ptr[a + 1] - L->walk_ptr[a]);
  if (!verify) return SCE_OK;

  // verification: every active site is its sublattice stencil, translated
  for (int s0 = 0; s0 < N; ++s0) {
    const int a = s0 % na;
    const bool act = d->site_active[s0] != 0;
    if (act != (d->site_active[a] != 0))
      return sce_fail(SCE_ELAYOUT, "site %d: activity differs from its sublattice", s0 + 1);
    if (!act) continue;
    const std::vector<int> js = sorted_adj(s0);
    if (js.size() != rep_js[a].size())
      return sce_fail(SCE_ELAYOUT, "site %d: adjacency length differs from its sublattice", s0 + 1);
    int c[3];
    g.cell_of(s0, c);
    size_t pos = 0;
    for (size_t q = 0; q < js.size(); ++q) {
      const int j = js[q], jr = rep_js[a][q];
      if (d->site_prog[j - 1] != d->site_prog[jr - 1] ||
          ((d->site_col[j - 1] > 0) != (d->site_col[jr - 1] > 0)) ||
          ((d->site_col[j - 1] < 0) != (d->site_col[jr - 1] < 0)))
        return sce_fail(SCE_ELAYOUT, "site %d: program sequence differs from its sublattice", s0 + 1);
      int fs[8];
      const int nf = factor_sites(d, j, fs);
      for (int f = 0; f < nf; ++f) {
        const int k = rep_nbidx[a][pos++];
        const int32_t* o = &L->nb[4 * (L->nb_ptr[a] + k)];
        int cc[3];
        for (int x = 0; x < 3; ++x) {
          cc[x] = c[x] + o[1 + x];
          if (cc[x] >= g.d[x]) cc[x] -= g.d[x];
        }
        if (g.site(o[0], cc) + 1 != fs[f])
          return sce_fail(SCE_ELAYOUT, "site %d: neighbour %d is not the translated stencil", s0 + 1, f + 1);
      }
      if (rep_nbidx[a][pos] != -1)
        return sce_fail(SCE_ELAYOUT, "site %d: factor count differs from its sublattice", s0 + 1);
      ++pos;
    }
  }
  return SCE_OK;
}

int32_t sce_position_tables_host(const sce_tables_desc* d, const HostLayout& L, std::vector<int32_t>& nq,
                                 std::vector<unsigned long long>& low, std::vector<uint16_t>& perm, int* adj_max) {
  const int na = L.na, nbm = L.nb_max, nact = d->n_active;
  std::vector<int32_t> cs(nact);
  for (int q = 0; q < nact; ++q) cs[q] = d->color_sites[q] - 1;
  // the neighbour sites of every colour position
  nq.assign((size_t)nact * nbm, 0);
  for (int q = 0; q < nact; ++q) {
    const int s = cs[q], a = s % na, cell = s / na;
    const int c1 = cell % d->dims[0], c2 = (cell / d->dims[0]) % d->dims[1], c3 = cell / (d->dims[0] * d->dims[1]);
    for (int j = 0; j < nbm; ++j) {
      int n = s;
      if (j < L.nb_ptr[a + 1] - L.nb_ptr[a]) {
        const int32_t* o = &L.nb[4 * (size_t)(L.nb_ptr[a] + j)];
        n = o[0] + na * ((c1 + o[1]) % d->dims[0] +
                         d->dims[0] * ((c2 + o[2]) % d->dims[1] + d->dims[1] * ((c3 + o[3]) % d->dims[2])));
      }
      nq[(size_t)q * nbm + j] = n;
    }
  }
  // which neighbour slots of a position hold lower colours
  low.clear();
  if (nbm <= 64) {
    std::vector<int> col_of((size_t)d->n_sites, -1);
    for (int c = 0; c < d->n_colors; ++c)
      for (int q = d->color_ptr[c] - 1; q < d->color_ptr[c + 1] - 1; ++q) col_of[cs[q]] = c;
    low.assign((size_t)std::max(nact, 1), 0ull);
    for (int q = 0; q < nact; ++q) {
      const int cq = col_of[cs[q]];
      for (int j = 0; j < nbm; ++j) {
        const int cn = col_of[nq[(size_t)q * nbm + j]];
        if (cn >= 0 && cn < cq) low[q] |= 1ull << j;
      }
    }
  }
  // raw adjacency index -> pid-sorted index, per colour position
  int amax = 0;
  for (int a = 0; a < na; ++a) amax = std::max(amax, L.prog_ptr[a + 1] - L.prog_ptr[a] - 1);
  if (amax > 65535) return sce_fail(SCE_EINVAL, "%d adjacency entries per site exceed the gradient's index", amax);
  const int am = std::max(amax, 1);
  perm.assign((size_t)nact * am, 0);
  std::vector<int> js;
  for (int q = 0; q < nact; ++q) {
    const int s0 = cs[q], j0 = d->site_ptr[s0], n = d->site_ptr[s0 + 1] - j0;
    js.resize(n);
    for (int k = 0; k < n; ++k) js[k] = k;
    std::stable_sort(js.begin(), js.end(),
                     [&](int x, int y) { return d->site_prog[j0 + x - 1] < d->site_prog[j0 + y - 1]; });
    for (int k = 0; k < n; ++k) perm[(size_t)q * am + js[k]] = (uint16_t)k;
  }
  *adj_max = am;
  return SCE_OK;
}

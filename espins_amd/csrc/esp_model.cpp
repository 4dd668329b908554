// esp_model.cpp -- host-side model preparation (tables, merged bonds, colouring).
#include "esp_model.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <unordered_map>

namespace esp {

// ---------------------------------------------------------------------------------
// Coupling-table expansion: the semantics of jij_get (jij.F90:123-446) and its B / D
// twins.  A row is a directed bond from the atom at fractional position f1 to the atom
// at f2, both possibly outside the home cell.  For every supercell cell (l fastest,
// jij.F90:176-186) the row lands in the next free slot of its first atom; the slot
// value is value/|m_i m_j| (jij.F90:290-291); empty slots (zero value, or a partner
// cut by the open boundary, jij.F90:264-267) are squeezed out afterwards
// (jij.F90:394-446, dij.F90:343-391 test all three D components).
// ---------------------------------------------------------------------------------
namespace {
struct Locator {
  int nb;
  const double *basis;
  int L[3];
  bool open;
  // Returns 1-based site, 0 = dropped by open boundary, -1 = unknown position.
  int operator()(const double *f, const int cell[3], bool second) const {
    double p[3];
    int c[3];
    for (int k = 0; k < 3; ++k) {
      p[k] = f[k] + cell[k];
      const double r = std::nearbyint(p[k]);
      if (std::fabs(p[k] - r) < 1e-5) p[k] = r;  // jij.F90:202-204 (eps5)
      c[k] = (int)std::floor(p[k]);
    }
    if (second && open)
      for (int k = 0; k < 3; ++k)
        if (c[k] < 0 || c[k] == L[k]) return 0;
    double in[3];
    for (int k = 0; k < 3; ++k) {
      in[k] = p[k] - c[k];
      c[k] = ((c[k] % L[k]) + L[k]) % L[k];
    }
    const int cellid = (c[2] * L[1] + c[1]) * L[0] + c[0];
    for (int a = 0; a < nb; ++a) {  // basis match within eps4 (jij.F90:226-229)
      const double *b = basis + 3 * a;
      if (std::fabs(in[0] - b[0]) <= 1e-4 && std::fabs(in[1] - b[1]) <= 1e-4 &&
          std::fabs(in[2] - b[2]) <= 1e-4)
        return cellid * nb + a + 1;
    }
    return -1;
  }
};
}  // namespace

int build_table(int nb, const double *basis_frac, const double *basis_mag, const int *L,
                int open_boundary, int nrows, const double *f1, const double *f2,
                const double *value, const double *dvec, int ncomp, int zcap, int *cnt,
                int *nbr, double *val, int *zmax_out) {
  const int ncell = L[0] * L[1] * L[2];
  const long N = (long)ncell * nb;
  Locator loc{nb, basis_frac, {L[0], L[1], L[2]}, open_boundary != 0};
  // The squeeze keeps row order, so a slot of site i is a (row, partner) pair; collect
  // per site in row order, then emit.
  struct Bond { int j; double v[3]; };
  std::vector<std::vector<Bond>> per(N);
  for (int n = 0; n < L[2]; ++n)
    for (int m = 0; m < L[1]; ++m)
      for (int l = 0; l < L[0]; ++l) {
        const int cell[3] = {l, m, n};
        for (int r = 0; r < nrows; ++r) {
          const int i = loc(f1 + 3 * r, cell, false);
          if (i < 0) return -5;
          const int j = loc(f2 + 3 * r, cell, true);
          if (j < 0) return -5;
          if (j == 0) continue;
          const double scale = std::fabs(basis_mag[(i - 1) % nb] * basis_mag[(j - 1) % nb]);
          Bond b{j, {0, 0, 0}};
          bool nz = false;
          for (int k = 0; k < ncomp; ++k) {
            b.v[k] = (ncomp == 1 ? value[r] : dvec[3 * r + k] * value[r]) / scale;
            nz |= b.v[k] != 0.0;
          }
          if (nz) per[i - 1].push_back(b);
        }
      }
  int zmax = 0;
  for (long i = 0; i < N; ++i) {
    const int c = (int)per[i].size();
    if (c > zcap) return -1;
    zmax = std::max(zmax, c);
    cnt[i] = c;
    for (int s = 0; s < c; ++s) {
      nbr[(long)s * N + i] = per[i][s].j;
      for (int k = 0; k < ncomp; ++k) val[ncomp * ((long)s * N + i) + k] = per[i][s].v[k];
    }
  }
  if (zmax_out) *zmax_out = zmax;
  return 0;
}

// ---------------------------------------------------------------------------------
// Device model.
// ---------------------------------------------------------------------------------
namespace {
struct MBond { int j; double c[kCoef]; };

// Merge the J, B and D lists of one site into one list of distinct partners, in
// first-appearance order (J slots, then B, then D), summing coefficients per partner.
// Each term's own slot order is preserved within the merged order, so per-term sums
// over the merged list visit bonds in the reference's order.
void merge_site(int i, int N, int zJ, const int *jc, const int *jn, const double *jv, int zB,
                const int *bc, const int *bn, const double *bv, int zD, const int *dc,
                const int *dn, const double *dv, std::vector<MBond> &out) {
  out.clear();
  auto find = [&](int j) -> MBond & {
    for (auto &b : out)
      if (b.j == j) return b;
    out.push_back(MBond{j, {0, 0, 0, 0, 0}});
    return out.back();
  };
  if (zJ > 0)
    for (int l = 0; l < jc[i]; ++l) find(jn[(long)l * N + i] - 1).c[0] += jv[(long)l * N + i];
  if (zB > 0)
    for (int l = 0; l < bc[i]; ++l) find(bn[(long)l * N + i] - 1).c[1] += bv[(long)l * N + i];
  if (zD > 0)
    for (int l = 0; l < dc[i]; ++l) {
      MBond &b = find(dn[(long)l * N + i] - 1);
      for (int k = 0; k < 3; ++k) b.c[2 + k] += dv[3 * ((long)l * N + i) + k];
    }
}

// DSatur colouring of a small graph; `order` breaks ties.  Returns the colour count.
int dsatur(int nv, const std::vector<std::vector<int>> &adj, const std::vector<int> &order,
           std::vector<int> &col) {
  col.assign(nv, -1);
  std::vector<std::vector<char>> sat(nv);
  std::vector<int> nsat(nv, 0);
  int ncol = 0;
  for (int it = 0; it < nv; ++it) {
    int best = -1;
    for (int v : order) {
      if (col[v] >= 0) continue;
      if (best < 0 || nsat[v] > nsat[best] ||
          (nsat[v] == nsat[best] && adj[v].size() > adj[best].size()))
        best = v;
    }
    int c = 0;
    while (c < (int)sat[best].size() && sat[best][c]) ++c;
    col[best] = c;
    ncol = std::max(ncol, c + 1);
    for (int u : adj[best]) {
      if ((int)sat[u].size() <= c) sat[u].resize(c + 1, 0);
      if (!sat[u][c]) { sat[u][c] = 1; ++nsat[u]; }
    }
  }
  return ncol;
}

// Iterated greedy (Culberson): recolour greedily visiting the vertices class by class, in a
// permuted class order; a greedy pass in such an order never needs more colours than the
// colouring it starts from.  Returns the final colour count.
int iterated_greedy(int nv, const std::vector<std::vector<int>> &adj, std::vector<int> &col, int k,
                    uint64_t &rs) {
  std::vector<int> order(nv), perm, fresh(nv), mark;
  for (int it = 0; it < 60; ++it) {
    perm.resize(k);
    for (int c = 0; c < k; ++c) perm[c] = c;
    const int kind = it % 3;
    if (kind == 0) std::reverse(perm.begin(), perm.end());
    else if (kind == 1) {
      std::vector<int> sz(k, 0);
      for (int c : col) sz[c]++;
      std::sort(perm.begin(), perm.end(), [&](int a, int b) { return sz[a] > sz[b]; });
    } else
      for (int c = k - 1; c > 0; --c) {
        rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17;
        std::swap(perm[c], perm[(int)(rs % (uint64_t)(c + 1))]);
      }
    int n = 0;
    for (int c : perm)
      for (int v = 0; v < nv; ++v)
        if (col[v] == c) order[n++] = v;
    std::fill(fresh.begin(), fresh.end(), -1);
    int k2 = 0;
    for (int v : order) {
      mark.assign(k2 + 1, 0);
      for (int u : adj[v])
        if (fresh[u] >= 0) mark[fresh[u]] = 1;
      int c = 0;
      while (c < k2 && mark[c]) ++c;
      fresh[v] = c;
      k2 = std::max(k2, c + 1);
    }
    col = fresh;
    k = k2;
  }
  return k;
}

// Periodic colouring: colour(site) = q(basis, cell mod p) for a period p_k | L_k, with q a
// colouring of the quotient torus.  Any bond of the lattice maps to an edge of the torus,
// so q lifts to a proper colouring whenever the torus has no loop.  Returns {} if no
// candidate period works.  The lifted colouring is re-verified on the full graph.
std::vector<int> periodic_colouring(int N, int nb, const int L[3],
                                    const std::vector<std::vector<int>> &adj) {
  const int nc = N / nb;
  auto cell_of = [&](int i, int x[3]) {
    const int c = i / nb;
    x[0] = c % L[0]; x[1] = (c / L[0]) % L[1]; x[2] = c / (L[0] * L[1]);
  };
  (void)nc;
  std::vector<int> best;
  int best_k = 1 << 30, best_max = 1 << 30;
  const int cands[] = {1, 2, 3, 4, 6, 8};
  for (int p : cands) {
    int P[3];
    for (int k = 0; k < 3; ++k) P[k] = (L[k] % p == 0) ? p : std::gcd(p, L[k]);
    const int nq = nb * P[0] * P[1] * P[2];
    if (nq > 4096) continue;
    auto qid = [&](int i) {
      int x[3];
      cell_of(i, x);
      return ((x[2] % P[2] * P[1] + x[1] % P[1]) * P[0] + x[0] % P[0]) * nb + i % nb;
    };
    std::vector<std::vector<int>> qadj(nq);
    bool loop = false;
    for (int i = 0; i < N && !loop; ++i) {
      const int u = qid(i);
      for (int j : adj[i]) {
        const int v = qid(j);
        if (u == v) { loop = true; break; }
        qadj[u].push_back(v);
      }
    }
    if (loop) continue;
    for (auto &a : qadj) {
      std::sort(a.begin(), a.end());
      a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    std::vector<int> order(nq), qcol, bestq;
    for (int v = 0; v < nq; ++v) order[v] = v;
    int kq = 1 << 30, kmax = 1 << 30;
    uint64_t rs = 0x9E3779B97F4A7C15ull;
    const int tries = nq <= 512 ? 24 : 4;
    for (int t = 0; t < tries; ++t) {
      if (t) {  // deterministic shuffle (xorshift)
        for (int v = nq - 1; v > 0; --v) {
          rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17;
          std::swap(order[v], order[(int)(rs % (uint64_t)(v + 1))]);
        }
      }
      int k = dsatur(nq, qadj, order, qcol);
      k = iterated_greedy(nq, qadj, qcol, k, rs);
      std::vector<int> sz(k, 0);
      for (int c : qcol) sz[c]++;
      const int mx = *std::max_element(sz.begin(), sz.end());
      if (k < kq || (k == kq && mx < kmax)) { kq = k; kmax = mx; bestq = qcol; }
    }
    // class sizes on the full lattice scale with nq; compare max class as a fraction
    const long full_max = (long)kmax * (N / nq);
    if (kq < best_k || (kq == best_k && full_max < best_max)) {
      std::vector<int> col(N);
      for (int i = 0; i < N; ++i) col[i] = bestq[qid(i)];
      bool ok = true;
      for (int i = 0; i < N && ok; ++i)
        for (int j : adj[i])
          if (col[j] == col[i]) { ok = false; break; }
      if (ok) { best = std::move(col); best_k = kq; best_max = (int)full_max; }
    }
  }
  return best;
}

// Smallest period P (per axis, from the candidates of periodic_colouring) such that a site's
// colour depends only on (basis, cell mod P); P = L when none does.
void colour_period(int N, int nb, const int L[3], const std::vector<int> &colour, int P[3]) {
  for (int k = 0; k < 3; ++k) P[k] = L[k];
  const int cands[] = {1, 2, 3, 4, 6, 8};
  for (int p : cands) {
    int Q[3];
    for (int k = 0; k < 3; ++k) Q[k] = (L[k] % p == 0) ? p : std::gcd(p, L[k]);
    const int nq = nb * Q[0] * Q[1] * Q[2];
    std::vector<int> qc(nq, -1);
    bool ok = true;
    for (int i = 0; i < N && ok; ++i) {
      const int c = i / nb, x = c % L[0], y = (c / L[0]) % L[1], z = c / (L[0] * L[1]);
      const int q = ((z % Q[2] * Q[1] + y % Q[1]) * Q[0] + x % Q[0]) * nb + i % nb;
      if (qc[q] < 0) qc[q] = colour[i];
      else ok = qc[q] == colour[i];
    }
    if (ok) {
      for (int k = 0; k < 3; ++k) P[k] = Q[k];
      return;
    }
  }
}

// Greedy colouring in reference order, then empty the smallest classes where every site
// can move to a larger class no neighbour uses.
std::vector<int> greedy_colouring(int N, const std::vector<std::vector<int>> &adj) {
  std::vector<int> col(N, -1), mark;
  int ncol = 0;
  for (int i = 0; i < N; ++i) {
    mark.assign(ncol + 1, 0);
    for (int j : adj[i])
      if (col[j] >= 0) mark[col[j]] = 1;
    int c = 0;
    while (c < ncol && mark[c]) ++c;
    col[i] = c;
    ncol = std::max(ncol, c + 1);
  }
  for (int pass = 0; pass < 4; ++pass) {
    std::vector<int> sz(ncol, 0);
    for (int c : col) sz[c]++;
    std::vector<int> byz(ncol);
    for (int c = 0; c < ncol; ++c) byz[c] = c;
    std::sort(byz.begin(), byz.end(), [&](int a, int b) { return sz[a] < sz[b]; });
    bool changed = false;
    for (int small : byz) {
      if (sz[small] == 0) continue;
      for (int i = 0; i < N; ++i) {
        if (col[i] != small) continue;
        mark.assign(ncol, 0);
        for (int j : adj[i]) mark[col[j]] = 1;
        for (auto it = byz.rbegin(); it != byz.rend(); ++it) {
          const int c = *it;
          if (c == small || mark[c] || sz[c] == 0) continue;
          col[i] = c; sz[c]++; sz[small]--; changed = true;
          break;
        }
      }
    }
    // renumber non-empty classes densely, larger first
    std::vector<int> order;
    for (int c = 0; c < ncol; ++c)
      if (sz[c]) order.push_back(c);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return sz[a] > sz[b]; });
    std::vector<int> remap(ncol, -1);
    for (int k = 0; k < (int)order.size(); ++k) remap[order[k]] = k;
    for (int &c : col) c = remap[c];
    ncol = (int)order.size();
    if (!changed) break;
  }
  return col;
}

// Implicit geometry (esp_geo.h) of a model whose device order groups every colour by basis
// atom and cell residue modulo the period P.  The per-(group, slot) records are read off the
// first site of each group; the result is accepted only if the computed partner, reference
// index, basis atom and moment equal the explicit tables for every site and slot.
void build_implicit(HostModel &hm, const int L[3], const int P[3], const double *mag) {
  hm.implicit = false;
  const int N = hm.N, nb = hm.nb, zp = hm.zmax;
  int M[3];
  for (int k = 0; k < 3; ++k) {
    if (P[k] < 1 || L[k] % P[k] != 0 || P[k] > 255) return;
    M[k] = L[k] / P[k];
    if (M[k] > 255) return;
  }
  const long mt = (long)M[0] * M[1] * M[2];
  if (mt <= 0 || N % mt != 0) return;
  const int ngroups = (int)(N / mt);
  if (ngroups != nb * P[0] * P[1] * P[2] || ngroups > 4096) return;
  espg::Geo &G = hm.geo;
  G.mt = espg::make_fastdiv((uint32_t)mt);
  G.m0 = espg::make_fastdiv((uint32_t)M[0]);
  G.m1 = espg::make_fastdiv((uint32_t)M[1]);
  G.M0 = M[0]; G.M1 = M[1]; G.M2 = M[2];
  G.P0 = P[0]; G.P1 = P[1]; G.P2 = P[2];
  G.L0 = L[0]; G.L1 = L[1];
  G.nb = nb; G.ngroups = ngroups; G.zmax = zp; G.N = N;
  auto cell = [&](int i, int x[3]) {
    const int c = i / nb;
    x[0] = c % L[0]; x[1] = (c / L[0]) % L[1]; x[2] = c / (L[0] * L[1]);
  };
  hm.grp.assign(ngroups, 0);
  hm.gslot.assign((size_t)ngroups * zp * 2, 0);
  for (int g = 0; g < ngroups; ++g) {
    const int i0 = hm.dev2ref[(size_t)g * mt];
    int x[3];
    cell(i0, x);
    hm.grp[g] = espg::grp_pack(i0 % nb, x[0] % P[0], x[1] % P[1], x[2] % P[2]);
    const int d0 = (int)((size_t)g * mt);
    for (int l = 0; l < zp; ++l) {
      const int jd = hm.nbr[(size_t)l * N + d0];
      int *e = &hm.gslot[((size_t)g * zp + l) * 2];
      if (jd == N) { e[0] = -1; e[1] = espg::q_pack(0, 0, 0); continue; }
      const int j = hm.dev2ref[jd];
      int y[3], q[3];
      cell(j, y);
      for (int k = 0; k < 3; ++k) {
        const int s_i = x[k] / P[k], s_j = y[k] / P[k];
        q[k] = ((s_j - s_i) % M[k] + M[k]) % M[k];   // in [0, M)
      }
      e[0] = (int)((jd / mt) * mt);
      e[1] = espg::q_pack(q[0], q[1], q[2]);
    }
  }
  hm.magb.assign(nb, 0.0);
  for (int a = 0; a < nb; ++a) hm.magb[a] = mag[a];
  // verify against the explicit tables
  for (int d = 0; d < N; ++d) {
    const espg::SiteGeo sg = espg::geo_site(G, hm.grp.data(), d);
    if (sg.ir != hm.dev2ref[d] || sg.a != hm.basis[d] || hm.mag[d] != hm.magb[sg.a]) return;
    for (int l = 0; l < zp; ++l) {
      const int *e = &hm.gslot[((size_t)sg.g * zp + l) * 2];
      if (espg::geo_partner(G, sg, e[0], e[1]) != hm.nbr[(size_t)l * N + d]) return;
    }
  }
  // groups touched by the sites of one 256-lane block (one lane per site at most)
  hm.rows_max = (int)std::min<long>(ngroups, (256 - 1) / mt + 2);
  hm.implicit = true;
}
}  // namespace

std::string prepare_model(HostModel &hm, int N, int nb, const double *mag, int zJ,
                          const int *j_cnt, const int *j_nbr, const double *j_val, int zB,
                          const int *b_cnt, const int *b_nbr, const double *b_val, int zD,
                          const int *d_cnt, const int *d_nbr, const double *d_vec, int have_si,
                          const double *si_par, const double *si_axis, int have_field,
                          const double *h_par, const double *h_axis, const int *Lhint,
                          const std::vector<int> *colour_hint) {
  if (N <= 0 || nb <= 0 || N % nb != 0) return "N must be a positive multiple of nb";
  if (!mag) return "mag is NULL";
  if (zJ < 0 || zB < 0 || zD < 0) return "negative table width";
  if (zJ > 0 && (!j_cnt || !j_nbr || !j_val)) return "J table pointer is NULL";
  if (zB > 0 && (!b_cnt || !b_nbr || !b_val)) return "B table pointer is NULL";
  if (zD > 0 && (!d_cnt || !d_nbr || !d_vec)) return "D table pointer is NULL";
  if (have_si && (!si_par || !si_axis)) return "single-ion pointer is NULL";
  if (have_field && (!h_par || !h_axis)) return "field pointer is NULL";
  auto check = [&](int z, const int *c, const int *n) -> bool {
    for (int i = 0; i < N; ++i) {
      if (c[i] < 0 || c[i] > z) return false;
      for (int l = 0; l < c[i]; ++l) {
        const int j = n[(long)l * N + i];
        if (j < 1 || j > N) return false;
      }
    }
    return true;
  };
  if (zJ > 0 && !check(zJ, j_cnt, j_nbr)) return "J neighbour index or count out of range";
  if (zB > 0 && !check(zB, b_cnt, b_nbr)) return "B neighbour index or count out of range";
  if (zD > 0 && !check(zD, d_cnt, d_nbr)) return "D neighbour index or count out of range";

  hm = HostModel();
  hm.N = N;
  hm.nb = nb;
  hm.has_si = have_si != 0;
  hm.has_h = have_field != 0;

  // merged bond lists (reference site order)
  std::vector<std::vector<MBond>> bonds(N);
  std::vector<MBond> tmp;
  int zmax = 0;
  for (int i = 0; i < N; ++i) {
    merge_site(i, N, zJ, j_cnt, j_nbr, j_val, zB, b_cnt, b_nbr, b_val, zD, d_cnt, d_nbr, d_vec,
               tmp);
    bonds[i] = tmp;
    zmax = std::max(zmax, (int)tmp.size());
    for (auto &b : tmp) {
      hm.has_B |= b.c[1] != 0.0;
      hm.has_D |= (b.c[2] != 0.0 || b.c[3] != 0.0 || b.c[4] != 0.0);
    }
  }
  hm.zmax = zmax;

  // Colouring of the bond graph (self bonds ignored: they only involve the site being
  // updated).  Periodic quotient colourings first, greedy + tail merge otherwise.
  std::vector<std::vector<int>> adj(N);
  for (int i = 0; i < N; ++i)
    for (auto &b : bonds[i])
      if (b.j != i) {
        adj[i].push_back(b.j);
        adj[b.j].push_back(i);
      }
  for (auto &a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  int Lc[3] = {0, 0, 0};
  if (Lhint && Lhint[0] > 0 && Lhint[1] > 0 && Lhint[2] > 0 &&
      (long)Lhint[0] * Lhint[1] * Lhint[2] * nb == N) {
    Lc[0] = Lhint[0]; Lc[1] = Lhint[1]; Lc[2] = Lhint[2];
  } else {
    const int nc = N / nb;
    int c = (int)std::lround(std::cbrt((double)nc));
    if (c > 0 && c * c * c == nc) Lc[0] = Lc[1] = Lc[2] = c;
  }
  hm.colour.clear();
  if (colour_hint && (int)colour_hint->size() == N) {  // a colouring of the same structure
    bool ok = true;
    for (int i = 0; i < N && ok; ++i)
      for (int j : adj[i])
        if ((*colour_hint)[j] == (*colour_hint)[i]) { ok = false; break; }
    if (ok) hm.colour = *colour_hint;
  }
  if (hm.colour.empty() && Lc[0] > 0) hm.colour = periodic_colouring(N, nb, Lc, adj);
  if (hm.colour.empty()) hm.colour = greedy_colouring(N, adj);
  int ncol = 0;
  for (int c : hm.colour) ncol = std::max(ncol, c + 1);
  hm.ncol = ncol;

  // Colour-major device order.  Inside a colour, sites are grouped by basis atom and by
  // residue of the cell modulo the colouring's period, cells in order inside a group: the
  // lanes of a wave then share the bond geometry of each slot, so their partners are
  // consecutive sites of one group of the partner colour and a gather touches few cache
  // lines.  Without a cell grid: by basis atom, then reference order.
  hm.col_off.assign(ncol + 1, 0);
  for (int i = 0; i < N; ++i) hm.col_off[hm.colour[i] + 1]++;
  for (int c = 0; c < ncol; ++c) hm.col_off[c + 1] += hm.col_off[c];
  hm.dev2ref.resize(N);
  hm.ref2dev.resize(N);
  int P[3] = {1, 1, 1};
  {
    std::vector<int64_t> key(N);
    if (Lc[0] > 0) colour_period(N, nb, Lc, hm.colour, P);
    for (int i = 0; i < N; ++i) {
      int64_t res = 0, sub = i / nb;
      if (Lc[0] > 0) {
        const int c = i / nb, x = c % Lc[0], y = (c / Lc[0]) % Lc[1], z = c / (Lc[0] * Lc[1]);
        res = (z % P[2] * P[1] + y % P[1]) * P[0] + x % P[0];
        sub = ((int64_t)(z / P[2]) * ((Lc[1] + P[1] - 1) / P[1]) + y / P[1]) * ((Lc[0] + P[0] - 1) / P[0]) + x / P[0];
      }
      key[i] = (((int64_t)hm.colour[i] * nb + i % nb) * ((int64_t)P[0] * P[1] * P[2]) + res) * (int64_t)N + sub;
    }
    std::vector<int> order(N);
    for (int i = 0; i < N; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return key[u] < key[v]; });
    for (int d = 0; d < N; ++d) {
      hm.dev2ref[d] = order[d];
      hm.ref2dev[order[d]] = d;
    }
  }

  // compressibility: every site of basis a has the same partner count and coefficients
  bool comp = true;
  std::vector<int> cnt_a(nb, -1);
  std::vector<std::vector<MBond>> proto(nb);
  for (int i = 0; i < N && comp; ++i) {
    const int a = i % nb;
    if (cnt_a[a] < 0) {
      cnt_a[a] = (int)bonds[i].size();
      proto[a] = bonds[i];
      continue;
    }
    if ((int)bonds[i].size() != cnt_a[a]) { comp = false; break; }
    for (size_t l = 0; l < bonds[i].size() && comp; ++l)
      comp = std::memcmp(bonds[i][l].c, proto[a][l].c, sizeof(double) * kCoef) == 0;
  }
  hm.compressed = comp;

  const int zp = std::max(zmax, 1);
  hm.zmax = zp;
  hm.has_self = false;
  hm.nbr.assign((size_t)zp * N, 0);
  if (!comp) hm.coef.assign((size_t)zp * N * kCoef, 0.0);
  else hm.ctab.assign((size_t)nb * zp * kCoef, 0.0);
  hm.mag.resize(N);
  hm.basis.resize(N);
  for (int d = 0; d < N; ++d) {
    const int i = hm.dev2ref[d];
    hm.mag[d] = mag[i];
    hm.basis[d] = (uint8_t)(i % nb);
    for (int l = 0; l < zp; ++l) {
      const bool real = l < (int)bonds[i].size();
      // padding slots point at the all-zero spin kept at device index N
      hm.nbr[(size_t)l * N + d] = real ? hm.ref2dev[bonds[i][l].j] : N;
      if (real && bonds[i][l].j == i) hm.has_self = true;
      if (!comp && real)
        for (int k = 0; k < kCoef; ++k) hm.coef[((size_t)l * N + d) * kCoef + k] = bonds[i][l].c[k];
    }
  }
  if (comp)
    for (int a = 0; a < nb; ++a)
      for (int l = 0; l < (int)proto[a].size(); ++l)
        for (int k = 0; k < kCoef; ++k)
          hm.ctab[((size_t)a * zp + l) * kCoef + k] = proto[a][l].c[k];

  hm.si4.assign((size_t)nb * 4, 0.0);
  hm.h4.assign((size_t)nb * 4, 0.0);
  for (int a = 0; a < nb; ++a) {
    if (have_si) {
      hm.si4[4 * a] = si_par[a];
      for (int k = 0; k < 3; ++k) hm.si4[4 * a + 1 + k] = si_axis[3 * a + k];
    }
    if (have_field) {
      hm.h4[4 * a] = h_par[a];
      for (int k = 0; k < 3; ++k) hm.h4[4 * a + 1 + k] = h_axis[3 * a + k];
    }
  }
  if (nb > 255) return "more than 255 basis atoms";
  if (comp && Lc[0] > 0) {
    std::vector<double> mb(nb, 0.0);
    for (int a = 0; a < nb; ++a) mb[a] = mag[a];   // site a of cell 0 (reference order)
    build_implicit(hm, Lc, P, mb.data());
  }
  return "";
}

uint64_t structure_hash(int N, int nb, int zJ, const int *j_cnt, const int *j_nbr, int zB,
                        const int *b_cnt, const int *b_nbr, int zD, const int *d_cnt,
                        const int *d_nbr) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void *p, size_t n) {
    const unsigned char *c = static_cast<const unsigned char *>(p);
    for (size_t k = 0; k < n; ++k) { h ^= c[k]; h *= 1099511628211ull; }
  };
  const int head[5] = {N, nb, zJ, zB, zD};
  mix(head, sizeof head);
  if (zJ > 0) { mix(j_cnt, sizeof(int) * N); mix(j_nbr, sizeof(int) * (size_t)zJ * N); }
  if (zB > 0) { mix(b_cnt, sizeof(int) * N); mix(b_nbr, sizeof(int) * (size_t)zB * N); }
  if (zD > 0) { mix(d_cnt, sizeof(int) * N); mix(d_nbr, sizeof(int) * (size_t)zD * N); }
  return h;
}

}  // namespace esp

// esp_setup.cpp -- host-only set-up routines (no GPU): the neighbour-shell finder of the
// reference's input generation and the spin-glass disorder of the exchange table.
//
// esp_neighbor_shells follows neighbors_get (neighbors.F90:48-168): image cells l, m, n in
// [-5, 5] ordered by the length of l a1 + m a2 + n a3 with the reference's predictable
// tie rule (neighbors_supercell_sort / internal_maxloc, neighbors.F90:1266-1360: repeatedly take
// the farthest cell, the lowest index among near-equal ones, and fill the order from the end);
// per atom, shells by increasing distance within a tolerance, their multiplicity, and the
// neighbours of each shell in (cell, kind, atom) order.
//
// esp_spin_glass follows jij_spin_glass (jij.F90:506-556): J_ij += sigma_ij x N(0,1) for every
// bond with i < j in table order, copied to the first (j, i) slot; normal numbers from MT19937
// by the polar method of mtprng_normal (mtprng.f90:381-407), the second number of each pair
// kept for the next call.  MT19937 is written from the published algorithm (Matsumoto &
// Nishimura 1998); real1 = u32 / (2^32 - 1) as mtprng_rand_real1 (mtprng.f90:328-342).
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/espins_hip.h"

namespace {

struct MT {
  uint32_t x[624];
  int i = 625;
  explicit MT(uint32_t seed) {
    x[0] = seed;
    for (int k = 1; k < 624; ++k) x[k] = 1812433253u * (x[k - 1] ^ (x[k - 1] >> 30)) + (uint32_t)k;
    i = 624;
  }
  uint32_t u32() {
    if (i >= 624) {
      for (int k = 0; k < 624; ++k) {
        const uint32_t v = (x[k] & 0x80000000u) | (x[(k + 1) % 624] & 0x7fffffffu);
        x[k] = x[(k + 397) % 624] ^ (v >> 1) ^ ((v & 1u) ? 0x9908b0dfu : 0u);
      }
      i = 0;
    }
    uint32_t y = x[i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
  }
  double real1() { return (double)u32() * (1.0 / 4294967295.0); }
};

// mtprng_normal: Marsaglia's polar method; one pair per two accepted uniforms, the second
// number of a pair returned by the next call
struct Normal {
  bool left = false;
  double other = 0.0;
  double operator()(MT &mt) {
    if (left) {
      left = false;
      return other;
    }
    double u1, u2, rad;
    do {
      u1 = -1.0 + 2.0 * mt.real1();
      u2 = -1.0 + 2.0 * mt.real1();
      rad = u1 * u1 + u2 * u2;
    } while (!(rad < 1.0 && rad > 0.0));
    const double f = std::sqrt(-2.0 * std::log(rad) / rad);
    other = u1 * f;
    left = true;
    return u2 * f;
  }
};

}  // namespace

extern "C" {

int esp_neighbor_shells(int nb, const double *lat, const double *pos, const int *nshell, double tol,
                        int cap_shell, int cap_multi, double *dnn, int *multi, int *list) {
  if (nb <= 0 || !lat || !pos || !nshell || !dnn || !multi || !list || cap_shell < 1 || cap_multi < 1 ||
      !(tol > 0.0))
    return ESP_EINVAL;
  for (int a = 0; a < nb; ++a)
    if (nshell[a] < 0 || nshell[a] > cap_shell) return ESP_EINVAL;
  constexpr int ns = 5, side = 2 * ns + 1, ncell = side * side * side;
  // image cells: (0, 0, 0) first, then l, m, n loops (n fastest), as the reference lists them
  std::vector<int> lmn(3 * ncell);
  std::vector<double> dist(ncell);
  int c = 0;
  lmn[0] = lmn[1] = lmn[2] = 0;
  dist[0] = 0.0;
  for (int l = -ns; l <= ns; ++l)
    for (int m = -ns; m <= ns; ++m)
      for (int n = -ns; n <= ns; ++n) {
        if (l == 0 && m == 0 && n == 0) continue;
        ++c;
        lmn[3 * c] = l; lmn[3 * c + 1] = m; lmn[3 * c + 2] = n;
        double v[3];
        for (int k = 0; k < 3; ++k) v[k] = l * lat[k] + m * lat[3 + k] + n * lat[6 + k];
        dist[c] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      }
  // neighbors_supercell_sort with internal_maxloc: from the last position down, the farthest
  // remaining cell, the lowest index among those within 1e-8 of it
  std::vector<int> order(ncell);
  std::vector<double> d = dist;
  for (int pos_out = ncell - 1; pos_out >= 0; --pos_out) {
    int g = 0;
    for (int k = 1; k < ncell; ++k)
      if (d[k] > d[g]) g = k;          // Fortran maxloc: the first maximum
    int lo = g;
    for (int k = 0; k < ncell; ++k)
      if (k != g && std::fabs(d[k] - d[g]) < 1e-8 && k < lo) lo = k;
    order[pos_out] = lo;
    d[lo] = -1.0;
  }
  std::vector<double> img(3 * ncell);
  for (int q = 0; q < ncell; ++q) {
    const int *t = &lmn[3 * order[q]];
    for (int k = 0; k < 3; ++k) img[3 * q + k] = t[0] * lat[k] + t[1] * lat[3 + k] + t[2] * lat[6 + k];
  }
  auto distance = [&](int a, int b, int q) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double v = pos[3 * a + k] - (pos[3 * b + k] + img[3 * q + k]);
      s += v * v;
    }
    return std::sqrt(s);
  };
  const double eta = 99999999.0;
  for (int a = 0; a < nb; ++a) {
    double d0 = 0.0, d1 = eta;
    for (int s = 0; s < nshell[a]; ++s) {
      int counter = 0;
      for (int q = 0; q < ncell; ++q)
        for (int b = 0; b < nb; ++b) {
          const double r = distance(a, b, q);
          if (r > tol && r > d0 + tol) {
            if (r < d1 - tol) {
              d1 = r;
              counter = 0;
            }
            if (r > d1 - tol && r < d1 + tol) ++counter;
          }
        }
      dnn[(size_t)a * cap_shell + s] = d1;
      multi[(size_t)a * cap_shell + s] = counter;
      d0 = d1;
      d1 = eta;
      if (counter > cap_multi) return ESP_EINVAL;
    }
    for (int s = 0; s < nshell[a]; ++s) {
      const double ds = dnn[(size_t)a * cap_shell + s];
      const int want = multi[(size_t)a * cap_shell + s];
      int nnx = 0;
      for (int q = 0; q < ncell && nnx < want; ++q)
        for (int b = 0; b < nb && nnx < want; ++b) {
          const double r = distance(a, b, q);
          if (r >= ds - tol && r <= ds + tol) {
            int *o = list + (((size_t)a * cap_shell + s) * cap_multi + nnx) * 4;
            const int *t = &lmn[3 * order[q]];
            o[0] = b; o[1] = t[0]; o[2] = t[1]; o[3] = t[2];
            ++nnx;
          }
        }
    }
  }
  return ESP_OK;
}

int esp_spin_glass(int N, int z, const int *cnt, const int *nbr, double *val, const double *sigma, int seed,
                   int *bad_pair) {
  if (N <= 0 || z < 0 || !cnt || !nbr || !val || !sigma) return ESP_EINVAL;
  MT mt((uint32_t)seed);
  Normal normal;
  for (int i = 1; i <= N; ++i)
    for (int l = 0; l < cnt[i - 1]; ++l) {
      const int j = nbr[(size_t)l * N + (i - 1)];
      if (i >= j) continue;
      const size_t e = (size_t)l * N + (i - 1);
      val[e] = normal(mt) * sigma[e] + val[e];
      bool found = false;
      for (int lj = 0; lj < cnt[j - 1]; ++lj)
        if (nbr[(size_t)lj * N + (j - 1)] == i) {
          val[(size_t)lj * N + (j - 1)] = val[e];
          found = true;
          break;
        }
      if (!found) {
        if (bad_pair) { bad_pair[0] = i; bad_pair[1] = j; }
        return ESP_ETABLE;
      }
    }
  return ESP_OK;
}

}  // extern "C"

/*
 * espins_oracle.c -- CPU restatement of the ESpinS hot path (TEST INFRASTRUCTURE).
 *
 * This file is the checker for libespins_hip.so and the CPU timing leg of bench.py.
 * It is never linked into, loaded by, or called from the product path.
 *
 * Every function follows the reference routine cited beside it (paths relative to
 * the ESpinS source tree).  Arithmetic is written in the reference's evaluation order
 * so that reference-order runs reproduce the reference's trajectory up to the libm
 * used for sin/cos/acos/exp (the reference leaves that to its compiler's runtime).
 */
#include "espins_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI    3.141592653589793238462643383279   /* constants.F90:30 */
#define ORC_TWOPI (2.0 * ORC_PI)                      /* constants.F90:31 */
#define ORC_EPS8  1.0e-8                              /* constants.F90:45 */
#define ORC_EPS4  1.0e-4
#define ORC_EPS5  1.0e-5

/* ===================== MT19937: mtprng.f90 ===================== */
/* mtprng_init (mtprng.f90:152-173).  The reference keeps the words in int64; for
 * non-negative seeds the low 32 bits, which are all that reach the output, follow
 * the standard init_genrand recurrence.  Negative seeds are sign-extended in the
 * reference and are not restated (parity unpinned for seed < 0). */
void orc_mt_init(orc_mt_state *st, int32_t seed) {
  st->mt[0] = (uint32_t)seed;
  for (int i = 1; i < 624; ++i) {
    uint32_t p = st->mt[i - 1];
    st->mt[i] = (uint32_t)(1812433253u * (p ^ (p >> 30)) + (uint32_t)i);
  }
  st->mti = 624;
}

/* mtprng_rand64 (mtprng.f90:227-282): regenerate the block, then temper. */
uint32_t orc_mt_u32(orc_mt_state *st) {
  static const uint32_t mag01[2] = {0u, 0x9908B0DFu};
  if (st->mti >= 624) {
    int kk;
    uint32_t y;
    for (kk = 0; kk < 624 - 397; ++kk) {
      y = (st->mt[kk] & 0x80000000u) | (st->mt[kk + 1] & 0x7FFFFFFFu);
      st->mt[kk] = st->mt[kk + 397] ^ (y >> 1) ^ mag01[y & 1u];
    }
    for (; kk < 623; ++kk) {
      y = (st->mt[kk] & 0x80000000u) | (st->mt[kk + 1] & 0x7FFFFFFFu);
      st->mt[kk] = st->mt[kk + (397 - 624)] ^ (y >> 1) ^ mag01[y & 1u];
    }
    y = (st->mt[623] & 0x80000000u) | (st->mt[0] & 0x7FFFFFFFu);
    st->mt[623] = st->mt[396] ^ (y >> 1) ^ mag01[y & 1u];
    st->mti = 0;
  }
  uint32_t r = st->mt[st->mti++];
  r ^= r >> 11;
  r ^= (r << 7) & 0x9D2C5680u;
  r ^= (r << 15) & 0xEFC60000u;
  r ^= r >> 18;
  return r;
}

/* mtprng_rand_real2 (mtprng.f90:347-361): r * (1/4294967296) */
double orc_mt_real2(orc_mt_state *st) {
  return (double)orc_mt_u32(st) * (1.0 / 4294967296.0);
}

void orc_mt_fill_u32(int32_t seed, int64_t n, uint32_t *out) {
  orc_mt_state st;
  orc_mt_init(&st, seed);
  for (int64_t i = 0; i < n; ++i) out[i] = orc_mt_u32(&st);
}

/* ===================== Philox4x32-10 ===================== */
/* Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11:
 * ten rounds of two 32x32->64 multiplies with a Weyl key schedule.  Not part of the
 * reference (it draws MT19937); this is the GPU's stream restated for colour-order parity. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double orc_u01(uint32_t x) { return (double)x * (1.0 / 4294967296.0); }

/* ===================== energies ===================== */
static inline int basis_of(const orc_model *m, int i0) { return i0 % m->nb; }

static inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
/* utility_cross (utility.F90:212-234) */
static inline void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* energy_get (get_quant.F90:186-261): ordered double sum over sites and slots,
 * J first, then field, single-ion, biquadratic, DM, then halved. */
double orc_energy(const orc_model *m, const double *s) {
  const int N = m->N;
  double e = 0.0;
  if (m->zJ > 0)
    for (int i = 0; i < N; ++i)
      for (int l = 0; l < m->j_cnt[i]; ++l) {
      int j = m->j_nbr[(size_t)l * N + i] - 1;
      e = e - m->j_val[(size_t)l * N + i] * dot3(s + 3 * i, s + 3 * j);
    }
  if (m->have_field)
    for (int i = 0; i < N; ++i) {
      int a = basis_of(m, i);
      e = e + 2.0 * m->h_par[a] * dot3(s + 3 * i, m->h_axis + 3 * a);
    }
  if (m->have_si)
    for (int i = 0; i < N; ++i) {
      int a = basis_of(m, i);
      double sa = dot3(s + 3 * i, m->si_axis + 3 * a);
      e = e + 2.0 * m->si_par[a] * sa * sa;
    }
  if (m->zB > 0)
    for (int i = 0; i < N; ++i)
      for (int l = 0; l < m->b_cnt[i]; ++l) {
        int j = m->b_nbr[(size_t)l * N + i] - 1;
        double d = dot3(s + 3 * i, s + 3 * j);
        e = e + m->b_val[(size_t)l * N + i] * d * d;
      }
  if (m->zD > 0)
    for (int i = 0; i < N; ++i)
      for (int l = 0; l < m->d_cnt[i]; ++l) {
        int j = m->d_nbr[(size_t)l * N + i] - 1;
        double c[3];
        cross3(s + 3 * i, s + 3 * j, c);
        e = e + dot3(m->d_vec + 3 * ((size_t)l * N + i), c);
      }
  return e / 2.0;
}

/* The terms energy_get sums, grouped by site: e_site[i] = the J, field, single-ion, B and D
 * terms of site i in the order energy_get adds them (get_quant.F90:210-257), so that
 * energy = (sum_i e_site[i]) / 2.  Summed exactly (e.g. math.fsum) this gives the correctly
 * rounded energy, free of the rounding of one long sequential sum. */
void orc_energy_sites(const orc_model *m, const double *s, double *e_site) {
  const int N = m->N;
  for (int i = 0; i < N; ++i) {
    double e = 0.0;
    const int a = basis_of(m, i);
    if (m->zJ > 0)
      for (int l = 0; l < m->j_cnt[i]; ++l) {
        int j = m->j_nbr[(size_t)l * N + i] - 1;
        e = e - m->j_val[(size_t)l * N + i] * dot3(s + 3 * i, s + 3 * j);
      }
    if (m->have_field) e = e + 2.0 * m->h_par[a] * dot3(s + 3 * i, m->h_axis + 3 * a);
    if (m->have_si) {
      double sa = dot3(s + 3 * i, m->si_axis + 3 * a);
      e = e + 2.0 * m->si_par[a] * sa * sa;
    }
    if (m->zB > 0)
      for (int l = 0; l < m->b_cnt[i]; ++l) {
        int j = m->b_nbr[(size_t)l * N + i] - 1;
        double d = dot3(s + 3 * i, s + 3 * j);
        e = e + m->b_val[(size_t)l * N + i] * d * d;
      }
    if (m->zD > 0)
      for (int l = 0; l < m->d_cnt[i]; ++l) {
        int j = m->d_nbr[(size_t)l * N + i] - 1;
        double cc[3];
        cross3(s + 3 * i, s + 3 * j, cc);
        e = e + dot3(m->d_vec + 3 * ((size_t)l * N + i), cc);
      }
    e_site[i] = e;
  }
}

/* delta_energy_get (get_quant.F90:264-337).  s holds the NEW spin at i0; the old
 * spin is recovered as s_i - ds exactly as the reference does (:306, :320). */
double orc_delta_energy(const orc_model *m, int i0, const double *s, const double ds[3]) {
  const int N = m->N;
  double de = 0.0;
  if (m->zJ > 0)
    for (int l = 0; l < m->j_cnt[i0]; ++l) {
    int j = m->j_nbr[(size_t)l * N + i0] - 1;
    de = de - m->j_val[(size_t)l * N + i0] * dot3(ds, s + 3 * j);
  }
  if (m->have_field) {
    int a = basis_of(m, i0);
    de = de + m->h_par[a] * dot3(ds, m->h_axis + 3 * a);
  }
  if (m->have_si) {
    int a = basis_of(m, i0);
    double si[3] = {s[3 * i0] - ds[0], s[3 * i0 + 1] - ds[1], s[3 * i0 + 2] - ds[2]};
    double sa = dot3(si, m->si_axis + 3 * a);
    double dsa = dot3(ds, m->si_axis + 3 * a);
    de = de + m->si_par[a] * (dsa * dsa + 2.0 * sa * dsa);
  }
  if (m->zB > 0)
    for (int l = 0; l < m->b_cnt[i0]; ++l) {
      int j = m->b_nbr[(size_t)l * N + i0] - 1;
      double si[3] = {s[3 * i0] - ds[0], s[3 * i0 + 1] - ds[1], s[3 * i0 + 2] - ds[2]};
      double sij = dot3(si, s + 3 * j);
      double dsj = dot3(ds, s + 3 * j);
      de = de + m->b_val[(size_t)l * N + i0] * (dsj * dsj + 2.0 * sij * dsj);
    }
  if (m->zD > 0)
    for (int l = 0; l < m->d_cnt[i0]; ++l) {
      int j = m->d_nbr[(size_t)l * N + i0] - 1;
      double c[3];
      cross3(ds, s + 3 * j, c);
      de = de + dot3(m->d_vec + 3 * ((size_t)l * N + i0), c);
    }
  return de;
}

/* magnetization_get (get_quant.F90:340-363): m accumulates S_j / N term by term. */
double orc_magnetization2(const orc_model *m, const double *s) {
  double mx = 0.0, my = 0.0, mz = 0.0, n = (double)m->N;
  for (int j = 0; j < m->N; ++j) {
    mx = mx + s[3 * j] / n;
    my = my + s[3 * j + 1] / n;
    mz = mz + s[3 * j + 2] / n;
  }
  return mx * mx + my * my + mz * mz;
}

/* ===================== proposals ===================== */
/* 'random': get_quant.F90:80-93.  'const': get_quant.F90:94-161 (cone of half-angle
 * tilt*acos(2u1-1) around the renormalised old spin, rotated by Rz(phi0) Ry(theta0)). */
void orc_propose(int mode, double tilt, double mag, const double old[3],
                 double u1, double u2, double out[3]) {
  if (mode == 0) {
    double costheta = 2.0 * u1 - 1.0;
    double sintheta = sqrt(1.0 - costheta * costheta);
    double phi = u2 * ORC_TWOPI;
    out[0] = mag * sintheta * cos(phi);
    out[1] = mag * sintheta * sin(phi);
    out[2] = mag * costheta;
  } else {
    double theta = tilt * acos(2.0 * u1 - 1.0);
    double phi = u2 * ORC_TWOPI;
    double sintheta = sin(theta);
    double spinx = sintheta * cos(phi);
    double spiny = sintheta * sin(phi);
    double spinz = cos(theta);
    double spinxy = sqrt(old[0] * old[0] + old[1] * old[1]);
    double cosphi, sinphi;
    if (spinxy > ORC_EPS8) { cosphi = old[0] / spinxy; sinphi = old[1] / spinxy; }
    else { cosphi = 1.0; sinphi = 0.0; }
    double inv = 1.0 / sqrt(old[0] * old[0] + old[1] * old[1] + old[2] * old[2]);
    double n0 = inv * old[0], n1 = inv * old[1], n2 = inv * old[2];
    out[0] = mag * (cosphi * n2 * spinx - sinphi * spiny + n0 * spinz);
    out[1] = mag * (sinphi * n2 * spinx + cosphi * spiny + n1 * spinz);
    out[2] = mag * (-sqrt(n0 * n0 + n1 * n1) * spinx + n2 * spinz);
  }
}

/* One metropolis() call (get_quant.F90:52-182) with its uniforms supplied.
 * u3 is consumed only if dE > 0; *need_u3 reports whether it was. */
static void metropolis_site(const orc_model *m, int i0, double beta, double *s,
                            const double prop[3], double u3, double *e_run,
                            int64_t *rejected, int *accepted) {
  double old[3] = {s[3 * i0], s[3 * i0 + 1], s[3 * i0 + 2]};
  s[3 * i0] = prop[0]; s[3 * i0 + 1] = prop[1]; s[3 * i0 + 2] = prop[2];
  double ds[3] = {s[3 * i0] - old[0], s[3 * i0 + 1] - old[1], s[3 * i0 + 2] - old[2]};
  double de = orc_delta_energy(m, i0, s, ds);
  *e_run = *e_run + de;
  *accepted = 1;
  if (de > 0.0) {
    double fac = exp(-beta * de);
    if (fac < u3) {
      s[3 * i0] = old[0]; s[3 * i0 + 1] = old[1]; s[3 * i0 + 2] = old[2];
      *e_run = *e_run - de;
      *rejected += 1;
      *accepted = 0;
    }
  }
}

/* Reference order, MT19937 stream: mcarlo.F90:511-519 calling get_quant.F90:52-182. */
void orc_sweep_ref(const orc_model *m, double *s, double beta, int mode, double tilt,
                   orc_mt_state *st, int64_t nsweeps, double *e_run, int64_t *rejected) {
  for (int64_t sw = 0; sw < nsweeps; ++sw)
    for (int i = 0; i < m->N; ++i) {
      double u1 = orc_mt_real2(st);
      double u2 = orc_mt_real2(st);
      double prop[3];
      orc_propose(mode, tilt, m->mag[i], s + 3 * i, u1, u2, prop);
      /* replicate metropolis_site but with a lazily drawn u3 (get_quant.F90:169-171) */
      double old[3] = {s[3 * i], s[3 * i + 1], s[3 * i + 2]};
      s[3 * i] = prop[0]; s[3 * i + 1] = prop[1]; s[3 * i + 2] = prop[2];
      double ds[3] = {s[3 * i] - old[0], s[3 * i + 1] - old[1], s[3 * i + 2] - old[2]};
      double de = orc_delta_energy(m, i, s, ds);
      *e_run = *e_run + de;
      if (de > 0.0) {
        double fac = exp(-beta * de);
        double u3 = orc_mt_real2(st);
        if (fac < u3) {
          s[3 * i] = old[0]; s[3 * i + 1] = old[1]; s[3 * i + 2] = old[2];
          *e_run = *e_run - de;
          *rejected += 1;
        }
      }
    }
}

void orc_sweep_colour_philox(const orc_model *m, int ncol, const int *col_off, const int *col_sites,
                             double *s, double beta, int mode, double tilt,
                             uint64_t seed, uint64_t sweep, uint32_t stream,
                             double *e_run, int64_t *rejected) {
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  for (int c = 0; c < ncol; ++c)
    for (int k = col_off[c]; k < col_off[c + 1]; ++k) {
      int i = col_sites[k];
      uint32_t ctr[4] = {(uint32_t)i, (uint32_t)sweep, (uint32_t)(sweep >> 32), stream};
      uint32_t r[4];
      orc_philox4x32_10(ctr, key, r);
      double prop[3];
      orc_propose(mode, tilt, m->mag[i], s + 3 * i, orc_u01(r[0]), orc_u01(r[1]), prop);
      int acc;
      metropolis_site(m, i, beta, s, prop, orc_u01(r[2]), e_run, rejected, &acc);
    }
}

void orc_sweep_colour_hostfed(const orc_model *m, int ncol, const int *col_off, const int *col_sites,
                              double *s, double beta, const double *prop, const double *unif,
                              uint8_t *accepted, double *e_run, int64_t *rejected) {
  for (int c = 0; c < ncol; ++c)
    for (int k = col_off[c]; k < col_off[c + 1]; ++k) {
      int i = col_sites[k];
      int acc;
      metropolis_site(m, i, beta, s, prop + 3 * i, unif[i], e_run, rejected, &acc);
      accepted[i] = (uint8_t)acc;
    }
}

/* ===================== parallel tempering ===================== */
/* pt_swap (pt.F90:1235-1288) in label form.  f = (1/T_t - 1/T_{t+1})(E_t - E_{t+1});
 * accept if f > 0, else draw u and accept iff u < exp(f); attempts counted on both
 * partners, accepts likewise. */
void orc_pt_swap(int start, int R, const double *tems, const double *energy, int *slot_rep,
                 int64_t *nswap, orc_mt_state *st, uint64_t seed, uint64_t step, uint32_t stream) {
  uint32_t key[2] = {(uint32_t)seed ^ 0x5bd1e995u, (uint32_t)(seed >> 32)};
  for (int t = start - 1; t + 1 < R; t += 2) {
    int ra = slot_rep[t], rb = slot_rep[t + 1];
    double f = (1.0 / tems[t] - 1.0 / tems[t + 1]) * (energy[ra] - energy[rb]);
    nswap[2 * t] += 1;
    nswap[2 * (t + 1)] += 1;
    int swap = 0;
    if (f > 0.0) swap = 1;
    else {
      double u;
      if (st) u = orc_mt_real2(st);
      else {
        uint32_t ctr[4] = {(uint32_t)t, (uint32_t)step, (uint32_t)(step >> 32), 0xFFFFFFFFu - stream};
        uint32_t r[4];
        orc_philox4x32_10(ctr, key, r);
        u = orc_u01(r[0]);
      }
      if (u < exp(f)) swap = 1;
    }
    if (swap) {
      slot_rep[t] = rb; slot_rep[t + 1] = ra;
      nswap[2 * t + 1] += 1;
      nswap[2 * (t + 1) + 1] += 1;
    }
  }
}

/* ===================== estimators ===================== */
/* mcarlo.F90:536-543: E, E^2, E^4 of the running energy; |m|, m^2, m^4. */
void orc_accumulate(double e, double m2, double acc[6]) {
  acc[0] = acc[0] + e;
  acc[1] = acc[1] + e * e;
  acc[2] = acc[2] + e * e * e * e;
  acc[3] = acc[3] + sqrt(m2);
  acc[4] = acc[4] + m2;
  acc[5] = acc[5] + m2 * m2;
}

/* mcarlo.F90:713-733 (pt.F90:995-1011): out = {<|m|>, <E>/N, C, chi, U_E, U_M, dE, dM} minus dM:
 * out[0]=<|m|>, out[1]=<E>/N, out[2]=C, out[3]=chi, out[4]=U_E, out[5]=U_M, out[6]=Delta_E. */
void orc_finalize(const double acc[6], int64_t nsteps, double T, int N, double out[7]) {
  double norm = 1.0 / (double)nsteps;
  double E = acc[0] * norm, E2 = acc[1] * norm, E4 = acc[2] * norm;
  double M = acc[3] * norm, M2 = acc[4] * norm, M4 = acc[5] * norm;
  double beta = 1.0 / T;
  out[0] = M;
  out[1] = E / (double)N;
  out[2] = beta * beta * (E2 - E * E) / (double)N;
  out[3] = beta * (M2 - M * M) * N;
  out[4] = 1.0 - (1.0 / 3.0) * (E4 / (E2 * E2));
  out[5] = 1.0 - (1.0 / 3.0) * (M4 / (M2 * M2));
  out[6] = sqrt((E2 - E * E) * norm) / (double)N;
}

/* ===================== coupling tables ===================== */
/* Locate the site of a fractional position (jij.F90:197-243 for the first atom,
 * :249-300 for the second): snap near-integers (eps5), split into cell + in-cell
 * position, wrap the cell periodically, match the basis atom within eps4.  Returns the
 * 1-based site, 0 if open-boundary says skip (second atom only), -1 if unknown. */
static int locate(const double *p_in, const int *cell, const int L[3], int nb,
                  const double *basis_frac, int second, int open_boundary) {
  double p[3];
  int lmn[3];
  for (int k = 0; k < 3; ++k) {
    p[k] = p_in[k] + (double)cell[k];
    if (fabs(p[k] - nearbyint(p[k])) < ORC_EPS5) p[k] = nearbyint(p[k]);
    lmn[k] = (int)floor(p[k]);
  }
  if (second && open_boundary)
    for (int k = 0; k < 3; ++k)
      if (lmn[k] < 0 || lmn[k] == L[k]) return 0;
  double pos[3];
  for (int k = 0; k < 3; ++k) {
    pos[k] = p[k] - (double)lmn[k];
    lmn[k] = (lmn[k] + 10 * L[k]) % L[k];
  }
  int sc = lmn[2] * L[1] * L[0] + lmn[1] * L[0] + lmn[0];
  for (int a = 0; a < nb; ++a)
    if (fabs(pos[0] - basis_frac[3 * a]) <= ORC_EPS4 && fabs(pos[1] - basis_frac[3 * a + 1]) <= ORC_EPS4 &&
        fabs(pos[2] - basis_frac[3 * a + 2]) <= ORC_EPS4)
      return sc * nb + a + 1;
  return -1;
}

int orc_build_table(int nb, const double *basis_frac, const double *basis_mag,
                    const int L[3], int open_boundary,
                    int nrows, const double *f1, const double *f2, const double *value,
                    const double *dvec, int ncomp, int zcap,
                    int *cnt, int *nbr, double *val) {
  const int ncell = L[0] * L[1] * L[2];
  const int N = ncell * nb;
  /* dense per-site slot lists in row order (jij_tot_matrix), then zero-drop
   * (jij_neighbors_get, jij.F90:394-446) */
  int *tot_n = (int *)calloc((size_t)N * (nrows > 0 ? nrows : 1), sizeof(int));
  double *tot_v = (double *)calloc((size_t)N * (nrows > 0 ? nrows : 1) * ncomp, sizeof(double));
  int *num = (int *)calloc((size_t)N, sizeof(int));
  int rc = 0;
  for (int n = 0; n < L[2] && rc == 0; ++n)
    for (int mm = 0; mm < L[1] && rc == 0; ++mm)
      for (int l = 0; l < L[0] && rc == 0; ++l) {
        int cell[3] = {l, mm, n};
        for (int r = 0; r < nrows; ++r) {
          int i = locate(f1 + 3 * r, cell, L, nb, basis_frac, 0, open_boundary);
          if (i < 0) { rc = -1; break; }
          int slot = num[i - 1]++;
          int j = locate(f2 + 3 * r, cell, L, nb, basis_frac, 1, open_boundary);
          if (j < 0) { rc = -1; break; }
          if (j == 0) continue; /* open boundary: slot stays zero and is dropped */
          double mi = basis_mag[(i - 1) % nb], mj = basis_mag[(j - 1) % nb];
          size_t at = (size_t)(i - 1) * nrows + slot;
          tot_n[at] = j;
          if (ncomp == 1) tot_v[at] = value[r] / fabs(mi * mj);
          else
            for (int k = 0; k < 3; ++k) tot_v[3 * at + k] = dvec[3 * r + k] * value[r] / fabs(mi * mj);
        }
      }
  int zmax = 0;
  if (rc == 0) {
    for (int i = 0; i < N && rc == 0; ++i) {
      int c = 0;
      for (int s = 0; s < nrows; ++s) {
        size_t at = (size_t)i * nrows + s;
        int nz = 0;
        for (int k = 0; k < ncomp; ++k) nz |= (tot_v[ncomp * at + k] != 0.0);
        if (!nz) continue;
        if (c >= zcap) { rc = -2; break; }
        nbr[(size_t)c * N + i] = tot_n[at];
        for (int k = 0; k < ncomp; ++k) val[ncomp * ((size_t)c * N + i) + k] = tot_v[ncomp * at + k];
        ++c;
      }
      cnt[i] = c;
      if (c > zmax) zmax = c;
    }
  }
  free(tot_n); free(tot_v); free(num);
  return rc < 0 ? rc : zmax;
}

/* A colouring is proper iff no J, B or D bond joins two distinct sites of one colour
 * (a self-bond, possible on a 1-cell supercell, only involves the updated site). */
int orc_check_colouring(const orc_model *m, const int *colour) {
  const int N = m->N;
  for (int i = 0; i < N; ++i) {
    if (m->zJ > 0)
      for (int l = 0; l < m->j_cnt[i]; ++l) {
        int j = m->j_nbr[(size_t)l * N + i] - 1;
        if (j != i && colour[j] == colour[i]) return 0;
      }
    if (m->zB > 0)
      for (int l = 0; l < m->b_cnt[i]; ++l) {
        int j = m->b_nbr[(size_t)l * N + i] - 1;
        if (j != i && colour[j] == colour[i]) return 0;
      }
    if (m->zD > 0)
      for (int l = 0; l < m->d_cnt[i]; ++l) {
        int j = m->d_nbr[(size_t)l * N + i] - 1;
        if (j != i && colour[j] == colour[i]) return 0;
      }
  }
  return 1;
}

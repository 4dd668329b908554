// esp_obs.hip -- optional observables measured on the device at every measurement of the
// global-memory path, accumulated by temperature slot:
//   order parameter        OP = sum_i S_i . axes(basis i) / N   (orderparam_get, get_quant.F90:396-422)
//   staggered moment       SM = | sum_i c(basis i) S_i / N |    (staggered_get, get_quant.F90:366-392)
//   P(m) of the cell moments (and of the cell OP / SM)          (Pm_get, get_quant.F90:425-493)
//   spin correlations per exchange shell and kind pair          (spincorr_get, get_quant.F90:496-540)
//   magnetic structure factor on a list of q-points            (mcarlo_structure_factor,
//                                                               get_quant.F90:649-717)
// The reference evaluates these on the host copy of one configuration per rank and
// measurement; here one measurement of all replicas is a few launches over the spins in HBM:
// k_obs_cells (one thread per cell: the cell sums, block partials of OP and SM, the P(m)
// counts through a histogram in LDS), k_obs_spincorr (one thread per site and several
// replicas, device order), k_sf_pack / k_sf_main / k_sf_fin (S(q) as a phase-matrix times
// spin-matrix product: one phase per q-point, site and 8 replicas; k_sf_lat for lattice
// positions: phases from per-q tables) and k_obs_final (per replica: fixed-order sums of the
// block partials, accumulation into the replica's slot).
// Histogram counts are integers, so they do not depend on the order of the atomic additions;
// every floating-point sum has a fixed order.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "esp_ctx.h"

namespace {
using namespace espk;

constexpr int kObsBlock = 256;
constexpr int kCellsPerThread = 4;    // cells per thread of k_obs_cells
constexpr int kMaxSpecies = 4;        // kinds in the spin-correlation tables
constexpr int kBondKindShift = 30;    // spin-correlation bond: partner | kind of partner << 30
constexpr unsigned kBondIdx = (1u << kBondKindShift) - 1u;
constexpr int kSfQ = 64;              // q-points per structure-factor wave (one per lane)
constexpr int kSfRep = 8;             // replicas per structure-factor wave (sincos path)
constexpr int kSfRepLat = 4;          // replicas per structure-factor wave (lattice path)
constexpr int kSfWaves = 4096;        // structure-factor waves aimed at (site chunks fill up to it)
constexpr int kSfMaxChunks = 64;
constexpr int kPmLdsBins = 12288;     // P(m) bins counted in LDS (48 KiB); more use global atomics

struct ObsArgs {
  const double *sp;
  int Ns, N, nb, R, ncell, flags;
  const int *ref2dev, *rep_slot;
  const double *mag;                  // device order
  const unsigned long long *nmeas;    // measurements so far (this one not counted yet)
  const double *axes, *coeff;
  double *part;
  int cblocks;
  const int *cdev;                    // [nb][ncell] device index of atom a of the k-th cell
  unsigned long long *pm;
  int nfreq, ndim, pm_lds;
  double m_min, m_max;
  const int *skind, *soff;
  const unsigned *sbond;
  const double *sw;
  int nshell, nsp, sblocks, srows;
  double *scpart, *sc;
  const double *q, *pos;               // pos by device site
  double polar[3];
  int have_polar, nq, every;
  double *sfspin, *sfpart;
  int sf_chunk, sf_nchunk;
  // lattice positions r = B_a + x0 A0 + x1 A1 + x2 A2 (reference order ((x2 L1 + x1) L0 + x0) nb
  // + a): phase tables exp(i q.x0 A0), exp(i q.x1 A1), exp(i q.x2 A2), exp(i q.B_a), [k][q] each
  int sf_lat, sfL[3], sf_rep;         // sf_rep: replicas per group of the packed spins
  const double2 *sfex, *sfey, *sfez, *sfeb;
  double *sf;
  unsigned long long *sfn;
  double *acc;
  double *trace;
  int trace_cap;
};

__device__ __forceinline__ V3 spin_at(const ObsArgs &o, int r, int d) {
  const double *b = o.sp + (size_t)r * 3 * o.Ns;
  return {gldi(b, d), gldi(b, (size_t)o.Ns + d), gldi(b, 2 * (size_t)o.Ns + d)};
}

// Bin of value x on the grid m_min + k (m_max - m_min) / (nfreq - 1): Fortran nint (half away
// from zero), the reference's 1-based max(., 1) as a 0-based index, clamped to the grid.
__device__ __forceinline__ int pm_bin(const ObsArgs &o, double x) {
  const double t = (x - o.m_min) / (o.m_max - o.m_min) * (double)(o.nfreq - 1);
  int k = (int)round(t);
  k = k < 0 ? 0 : k;
  return k > o.nfreq - 1 ? o.nfreq - 1 : k;
}

// kCellsPerThread cells per thread (nb consecutive reference sites each), cells in the order
// of cdev: a block covers kObsBlock x kCellsPerThread cells, so the P(m) histogram in LDS is
// cleared and flushed (one global atomic per non-empty bin) once per that many cells.
__global__ __launch_bounds__(kObsBlock) void k_obs_cells(ObsArgs o) {
  extern __shared__ unsigned s_hist[];
  __shared__ double red[4 * (kObsBlock / 64)];
  const int r = blockIdx.y;
  const bool PM = o.flags & ESP_OBS_PM, OP = o.flags & ESP_OBS_ORDER, SM = o.flags & ESP_OBS_STAGGERED;
  const int nh = o.ndim * o.nfreq;
  if (PM && o.pm_lds) {
    for (int k = threadIdx.x; k < nh; k += kObsBlock) s_hist[k] = 0u;
    __syncthreads();
  }
  const int slot = gldi(o.rep_slot, r);
  unsigned long long *g = PM ? o.pm + (size_t)slot * nh : nullptr;
  double v[4] = {0.0, 0.0, 0.0, 0.0};     // OP, SM x, y, z of this thread's cells
  for (int t = 0; t < kCellsPerThread; ++t) {
    const int c = (blockIdx.x * kCellsPerThread + t) * kObsBlock + threadIdx.x;
    if (c >= o.ncell) break;
    V3 m = {0.0, 0.0, 0.0}, st = {0.0, 0.0, 0.0};
    double op = 0.0;
    for (int a = 0; a < o.nb; ++a) {
      const int d = gldi(o.cdev, (size_t)a * o.ncell + c);
      const V3 s = spin_at(o, r, d);
      m.x += s.x; m.y += s.y; m.z += s.z;
      if (OP) op += s.x * gldi(o.axes, 3 * a) + s.y * gldi(o.axes, 3 * a + 1) + s.z * gldi(o.axes, 3 * a + 2);
      if (SM) {
        const double w = gldi(o.coeff, a);
        st.x += w * s.x; st.y += w * s.y; st.z += w * s.z;
      }
    }
    v[0] += op; v[1] += st.x; v[2] += st.y; v[3] += st.z;
    if (PM) {
      int bins[3], nbin = 0;
      bins[nbin++] = pm_bin(o, sqrt(m.x * m.x + m.y * m.y + m.z * m.z));
      if (OP) bins[nbin++] = o.nfreq + pm_bin(o, op);
      const int off = (OP ? 2 : 1) * o.nfreq;
      if (SM) bins[nbin++] = off + pm_bin(o, sqrt(st.x * st.x + st.y * st.y + st.z * st.z));
      for (int k = 0; k < nbin; ++k) {
        if (o.pm_lds) atomicAdd(&s_hist[bins[k]], 1u);
        else gadd(g + bins[k], 1ull);
      }
    }
  }
  if (PM && o.pm_lds) {
    __syncthreads();
    for (int k = threadIdx.x; k < nh; k += kObsBlock) {
      const unsigned h = s_hist[k];
      if (h) gadd(g + k, (unsigned long long)h);
    }
  }
  if (OP || SM) {
    block_sum<4>(v, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 4; ++k) gsti(o.part, ((size_t)r * o.cblocks + blockIdx.x) * 4 + k, v[k]);
  }
}

// One thread per site (device order) and REP replicas: per shell, the sums of
// S_i . S_j / (n m_i m_j) and of |S_i . S_j| / (n m_i m_j) over the shell's bonds, by kind
// pair, as block partials [shell][2][kind i][kind j].  A table entry of site i names a partner
// j (device index, its kind in the top two bits) and the weights of i -> j and, when the
// shell is symmetric, of j -> i (both directions of the bond from one gather); shell s has
// rows soff[s] .. soff[s+1]-1.  A bond costs coalesced table loads shared by the REP
// replicas and one spin gather per replica.  (Keeping a replica group on one
// XCD, as the sweep does, measured slower here: 468 against 404 us on FCC 32^3 x 64.)
template <int NS, int REP>
__global__ __launch_bounds__(kObsBlock) void k_obs_spincorr(ObsArgs o) {
  constexpr int NV = 2 * NS * NS * REP;
  __shared__ double red[NV * (kObsBlock / 64)];
  const int b = blockIdx.x;
  const int r0 = blockIdx.y * REP;
  const int d0 = b * kObsBlock + threadIdx.x;
  const bool on = d0 < o.N;
  const int d = on ? d0 : 0;
  const int ki = NS > 1 ? gldi(o.skind, d) : 0;
  V3 si[REP];
  int ru[REP];
#pragma unroll
  for (int u = 0; u < REP; ++u) {
    ru[u] = min(r0 + u, o.R - 1);              // replicas past R read a valid one, weighted 0
    const V3 s = spin_at(o, ru[u], d);
    const bool use = on && r0 + u < o.R;
    si[u] = {use ? s.x : 0.0, use ? s.y : 0.0, use ? s.z : 0.0};
  }
  for (int s = 0; s < o.nshell; ++s) {
    double v[REP][2][NS], vr[REP][2][NS];     // [sum, sum of |.|][kind j]: (i, j) and (j, i)
#pragma unroll
    for (int u = 0; u < REP; ++u)
#pragma unroll
      for (int y = 0; y < NS; ++y) { v[u][0][y] = v[u][1][y] = vr[u][0][y] = vr[u][1][y] = 0.0; }
    const int l1 = gldi(o.soff, s + 1);
    for (int l = gldi(o.soff, s); l < l1; ++l) {
      const unsigned e = gldi(o.sbond, (size_t)l * o.N + d);
      const double w = gldi(o.sw, (size_t)l * o.N + d);
      const double wr = NS > 1 ? gldi(o.sw, ((size_t)o.srows + l) * o.N + d) : 0.0;
      const int dj = (int)(e & kBondIdx);
      const int kj = NS > 1 ? (int)(e >> kBondKindShift) : 0;
#pragma unroll
      for (int u = 0; u < REP; ++u) {
        const V3 sj = spin_at(o, ru[u], dj);
        const double x = si[u].x * sj.x + si[u].y * sj.y + si[u].z * sj.z;
        const double t = x * w, ta = fabs(x) * w;
#pragma unroll
        for (int y = 0; y < NS; ++y)
          if (y == kj) {
            v[u][0][y] += t; v[u][1][y] += ta;
            if (NS > 1) { vr[u][0][y] += x * wr; vr[u][1][y] += fabs(x) * wr; }
          }
      }
    }
    double f[NV];
#pragma unroll
    for (int u = 0; u < REP; ++u)
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int x = 0; x < NS; ++x)
#pragma unroll
          for (int y = 0; y < NS; ++y)
            f[((u * 2 + t) * NS + x) * NS + y] = (x == ki ? v[u][t][y] : 0.0) + (y == ki ? vr[u][t][x] : 0.0);
    block_sum<NV>(f, red);
    if (threadIdx.x == 0)
#pragma unroll
      for (int u = 0; u < REP; ++u)
        if (r0 + u < o.R) {
          double *p = o.scpart + (((size_t)(r0 + u) * o.sblocks + b) * o.nshell + s) * 2 * NS * NS;
#pragma unroll
          for (int k = 0; k < 2 * NS * NS; ++k) gsti(p, k, f[u * 2 * NS * NS + k]);
        }
  }
}

// The structure factor F0(q) = sum_j S_j exp(i q.r_j) of every replica is a product of the
// phase matrix [q][site] with the spin matrix [site][replica, component]; a phase is computed
// once per (q, site) and kSfRep replicas.  Three launches, each a no-op unless this
// measurement is one that evaluates S(q) (uniform over the grid):
//   k_sf_pack   the spins, site-major in groups of kSfRep replicas: [group][site][kSfRep][3]
//               (device order; for lattice positions reference order and groups of kSfRepLat);
//   k_sf_main   one wave per (64 q-points, replica group, site chunk): lane = q-point, the
//               site's position and its kSfRep x 3 spins are wave-uniform loads (SGPRs),
//               one sincos per site serves kSfRep x 6 FMAs (the wave is bound by its FP64
//               instructions; 16 replicas per wave (248 VGPRs) and phases shared by four
//               waves through LDS both measured slower);
//               partial sums [chunk][replica][6][q];
//   k_sf_fin    one thread per (q, replica): the chunk partials in order, then F, the
//               component of F0 perpendicular to q, and NSF = Re(polar . F) / |polar|; adds
//               |F|^2, |F|^2 - NSF^2 and NSF^2 to the slot's sums.  q = 0 adds nothing.
__device__ __forceinline__ bool sf_due(const ObsArgs &o) {
  return (gldi(o.nmeas, 0) + 1ull) % (unsigned long long)o.every == 0ull;
}

__global__ __launch_bounds__(kObsBlock) void k_sf_pack(ObsArgs o) {
  if (!sf_due(o)) return;
  const int j = blockIdx.x * kObsBlock + threadIdx.x;    // device site, or reference site (lattice)
  if (j >= o.N) return;
  const int d = o.sf_lat ? gldi(o.ref2dev, j) : j;
  const int g = blockIdx.y;
  double *out = o.sfspin + ((size_t)g * o.N + j) * (o.sf_rep * 3);
#pragma unroll
  for (int u = 0; u < kSfRep; ++u) {
    if (u >= o.sf_rep) break;
    const int r = g * o.sf_rep + u;
    const V3 s = r < o.R ? spin_at(o, r, d) : V3{0.0, 0.0, 0.0};
    gsti(out, 3 * u, s.x); gsti(out, 3 * u + 1, s.y); gsti(out, 3 * u + 2, s.z);
  }
}

__global__ __launch_bounds__(kSfQ) void k_sf_main(ObsArgs o) {
  if (!sf_due(o)) return;
  const int iq = blockIdx.x * kSfQ + threadIdx.x;
  const int g = blockIdx.y, ch = blockIdx.z;
  const int d0 = ch * o.sf_chunk, d1 = min(o.N, d0 + o.sf_chunk);
  double qx = 0.0, qy = 0.0, qz = 0.0;
  if (iq < o.nq) { qx = gldi(o.q, 3 * iq); qy = gldi(o.q, 3 * iq + 1); qz = gldi(o.q, 3 * iq + 2); }
  double fr[kSfRep][3], fi[kSfRep][3];
#pragma unroll
  for (int u = 0; u < kSfRep; ++u)
#pragma unroll
    for (int k = 0; k < 3; ++k) { fr[u][k] = 0.0; fi[u][k] = 0.0; }
  const double *spg = o.sfspin + (size_t)g * o.N * (kSfRep * 3);
  for (int d = d0; d < d1; ++d) {
    const double *p = o.pos + 3 * (size_t)d;
    double sn, cs;
    sincos(qx * p[0] + qy * p[1] + qz * p[2], &sn, &cs);
    const double *sv = spg + (size_t)d * (kSfRep * 3);
#pragma unroll
    for (int u = 0; u < kSfRep; ++u)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        fr[u][k] += sv[3 * u + k] * cs;
        fi[u][k] += sv[3 * u + k] * sn;
      }
  }
  if (iq >= o.nq) return;
#pragma unroll
  for (int u = 0; u < kSfRep; ++u) {
    const int r = g * kSfRep + u;
    if (r >= o.R) break;
    double *out = o.sfpart + ((size_t)ch * o.R + r) * 6 * o.nq + iq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      gsti(out, (size_t)k * o.nq, fr[u][k]);
      gsti(out, (size_t)(3 + k) * o.nq, fi[u][k]);
    }
  }
}

// k_sf_main for lattice positions: the phase of site (x0, x1, x2, a) is the product of four
// table entries (exp(i q.x0 A0) exp(i q.x1 A1) exp(i q.x2 A2) exp(i q.B_a)): per site one
// complex product instead of a sincos.  A chunk is a range of (x1, x2) rows, sites in
// reference order inside a row.
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 ld2(const double2 *p, size_t i) {
  const double *d = reinterpret_cast<const double *>(p);
  return make_double2(gldi(d, 2 * i), gldi(d, 2 * i + 1));
}

__global__ __launch_bounds__(kSfQ) void k_sf_lat(ObsArgs o) {
  constexpr int RV = kSfRepLat * 3;                 // spin values of one site and group
  if (!sf_due(o)) return;
  const int iq0 = blockIdx.x * kSfQ + threadIdx.x;
  const int iq = min(iq0, o.nq - 1);                 // lanes past nq compute q nq-1, store nothing
  const int g = blockIdx.y, ch = blockIdx.z;
  const int L0 = o.sfL[0], L1 = o.sfL[1], nb = o.nb;
  const int rows = L1 * o.sfL[2];
  const int w0 = ch * o.sf_chunk, w1 = min(rows, w0 + o.sf_chunk);
  double fr[kSfRepLat][3], fi[kSfRepLat][3];
#pragma unroll
  for (int u = 0; u < kSfRepLat; ++u)
#pragma unroll
    for (int k = 0; k < 3; ++k) { fr[u][k] = 0.0; fi[u][k] = 0.0; }
  auto add = [&](const double (&v)[RV], double2 ph) {
#pragma unroll
    for (int u = 0; u < kSfRepLat; ++u)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        fr[u][k] += v[3 * u + k] * ph.x;
        fi[u][k] += v[3 * u + k] * ph.y;
      }
  };
  const double *spg = o.sfspin + (size_t)g * o.N * RV;
  for (int w = w0; w < w1; ++w) {
    const double2 byz = cmul(ld2(o.sfey, (size_t)(w % L1) * o.nq + iq), ld2(o.sfez, (size_t)(w / L1) * o.nq + iq));
    for (int x0 = 0; x0 < L0; ++x0) {
      const double2 bx = cmul(ld2(o.sfex, (size_t)x0 * o.nq + iq), byz);
      const double *sv0 = spg + ((size_t)w * L0 + x0) * nb * RV;
      int a = 0;
      for (; a + 1 < nb; a += 2) {     // two sites per wait on the (out-of-order) scalar loads
        double v0[RV], v1[RV];
#pragma unroll
        for (int k = 0; k < RV; ++k) { v0[k] = sv0[(size_t)a * RV + k]; v1[k] = sv0[(size_t)(a + 1) * RV + k]; }
        add(v0, cmul(bx, ld2(o.sfeb, (size_t)a * o.nq + iq)));
        add(v1, cmul(bx, ld2(o.sfeb, (size_t)(a + 1) * o.nq + iq)));
      }
      if (a < nb) {
        double v0[RV];
#pragma unroll
        for (int k = 0; k < RV; ++k) v0[k] = sv0[(size_t)a * RV + k];
        add(v0, cmul(bx, ld2(o.sfeb, (size_t)a * o.nq + iq)));
      }
    }
  }
  if (iq0 >= o.nq) return;
#pragma unroll
  for (int u = 0; u < kSfRepLat; ++u) {
    const int r = g * kSfRepLat + u;
    if (r >= o.R) break;
    double *out = o.sfpart + ((size_t)ch * o.R + r) * 6 * o.nq + iq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      gsti(out, (size_t)k * o.nq, fr[u][k]);
      gsti(out, (size_t)(3 + k) * o.nq, fi[u][k]);
    }
  }
}

__global__ __launch_bounds__(kObsBlock) void k_sf_fin(ObsArgs o) {
  if (!sf_due(o)) return;
  const int iq = blockIdx.x * kObsBlock + threadIdx.x;
  const int r = blockIdx.y;
  if (iq >= o.nq) return;
  const double qx = gldi(o.q, 3 * iq), qy = gldi(o.q, 3 * iq + 1), qz = gldi(o.q, 3 * iq + 2);
  const double q2 = qx * qx + qy * qy + qz * qz;
  if (q2 == 0.0) return;
  double fr[3] = {0.0, 0.0, 0.0}, fi[3] = {0.0, 0.0, 0.0};
  for (int ch = 0; ch < o.sf_nchunk; ++ch) {
    const double *p = o.sfpart + ((size_t)ch * o.R + r) * 6 * o.nq + iq;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      fr[k] += gldi(p, (size_t)k * o.nq);
      fi[k] += gldi(p, (size_t)(3 + k) * o.nq);
    }
  }
  const double pn = o.have_polar ? sqrt(o.polar[0] * o.polar[0] + o.polar[1] * o.polar[1] +
                                        o.polar[2] * o.polar[2]) : 1.0;
  // perpendicular part: F - (F . q / q^2) q (real and imaginary parts separately)
  const double ar = (fr[0] * qx + fr[1] * qy + fr[2] * qz) / q2;
  const double ai = (fi[0] * qx + fi[1] * qy + fi[2] * qz) / q2;
  const double pr[3] = {fr[0] - ar * qx, fr[1] - ar * qy, fr[2] - ar * qz};
  const double pi[3] = {fi[0] - ai * qx, fi[1] - ai * qy, fi[2] - ai * qz};
  const double f2 = pr[0] * pr[0] + pr[1] * pr[1] + pr[2] * pr[2] + pi[0] * pi[0] + pi[1] * pi[1] +
                    pi[2] * pi[2];
  double nsf = 0.0;
  if (o.have_polar) nsf = (o.polar[0] * pr[0] + o.polar[1] * pr[1] + o.polar[2] * pr[2]) / pn;
  const int slot = gldi(o.rep_slot, r);
  double *acc = o.sf + (size_t)slot * 3 * o.nq;
  gsti(acc, iq, gldi(acc, iq) + f2);
  gsti(acc, (size_t)o.nq + iq, gldi(acc, (size_t)o.nq + iq) + (f2 - nsf * nsf));
  gsti(acc, 2 * (size_t)o.nq + iq, gldi(acc, 2 * (size_t)o.nq + iq) + nsf * nsf);
}

// One block per replica: fixed-order sums of the block partials, accumulation by slot.
__global__ __launch_bounds__(kObsBlock) void k_obs_final(ObsArgs o) {
  __shared__ double red[4 * (kObsBlock / 64)];
  const int r = blockIdx.x;
  const int slot = gldi(o.rep_slot, r);
  const unsigned long long nm = gldi(o.nmeas, 0);
  double v[4] = {0.0, 0.0, 0.0, 0.0};    // OP, SM x, y, z: strided over the cell blocks
  if (o.flags & (ESP_OBS_ORDER | ESP_OBS_STAGGERED)) {   // uniform: block_sum needs all threads
    for (int b = threadIdx.x; b < o.cblocks; b += kObsBlock)
      for (int k = 0; k < 4; ++k) v[k] += gldi(o.part, ((size_t)r * o.cblocks + b) * 4 + k);
    block_sum<4>(v, red);
  }
  if ((o.flags & (ESP_OBS_ORDER | ESP_OBS_STAGGERED)) && threadIdx.x == 0) {
    const double n = (double)o.N;
    const double op = v[0] / n;
    const double sx = v[1] / n, sy = v[2] / n, sz = v[3] / n;
    const double op2 = op * op, sm2 = sx * sx + sy * sy + sz * sz;
    double *a = o.acc + (size_t)slot * 6;
    if (o.flags & ESP_OBS_ORDER) {
      gsti(a, 0, gldi(a, 0) + sqrt(op2));
      gsti(a, 1, gldi(a, 1) + op2);
      gsti(a, 2, gldi(a, 2) + op2 * op2);
    }
    if (o.flags & ESP_OBS_STAGGERED) {
      gsti(a, 3, gldi(a, 3) + sqrt(sm2));
      gsti(a, 4, gldi(a, 4) + sm2);
      gsti(a, 5, gldi(a, 5) + sm2 * sm2);
    }
    if (o.trace && nm < (unsigned long long)o.trace_cap) {
      double *t = o.trace + ((size_t)nm * o.R + slot) * 2;
      gsti(t, 0, sqrt(op2));
      gsti(t, 1, sqrt(sm2));
    }
  }
  if (o.flags & ESP_OBS_SPINCORR) {      // uniform over the block: block_sum needs all threads
    const int ns = o.nsp, kk = ns * ns;
    for (int e = 0; e < o.nshell * kk; ++e) {
      const int s = e / kk, ab = e % kk;
      double w[2] = {0.0, 0.0};               // dot, sum of |dot|: strided over the site blocks
      for (int b = threadIdx.x; b < o.sblocks; b += kObsBlock) {
        const double *p = o.scpart + (((size_t)r * o.sblocks + b) * o.nshell + s) * 2 * kk;
        w[0] += gldi(p, ab);
        w[1] += gldi(p, (size_t)kk + ab);
      }
      block_sum<2>(w, red);
      if (threadIdx.x == 0) {
        double *a = o.sc + (size_t)slot * 3 * o.nshell * kk;
        const size_t q = (size_t)s * kk + ab;
        gsti(a, q, gldi(a, q) + w[0]);
        gsti(a, (size_t)o.nshell * kk + q, gldi(a, (size_t)o.nshell * kk + q) + fabs(w[0]));
        gsti(a, 2 * (size_t)o.nshell * kk + q, gldi(a, 2 * (size_t)o.nshell * kk + q) + w[1]);
      }
    }
  }
  if ((o.flags & ESP_OBS_SFACTOR) && r == 0 && threadIdx.x == 0 && (nm + 1ull) % (unsigned long long)o.every == 0)
    gsti(o.sfn, 0, gldi(o.sfn, 0) + 1ull);
}

int ndim_pm(const esp_ctx *c) {
  return 1 + ((c->obs.flags & ESP_OBS_ORDER) ? 1 : 0) + ((c->obs.flags & ESP_OBS_STAGGERED) ? 1 : 0);
}

ObsArgs obs_args(esp_ctx *c) {
  ObsArgs o{};
  auto &b = c->obs;
  const esp::HostModel &m = c->hm;
  o.sp = c->d_sp; o.Ns = c->Ns; o.N = m.N; o.nb = m.nb; o.R = c->R; o.ncell = m.N / m.nb;
  o.flags = b.flags;
  o.ref2dev = c->d_ref2dev; o.rep_slot = c->d_rep_slot; o.mag = c->d_mag; o.nmeas = c->d_nmeas;
  o.axes = b.d_axes; o.coeff = b.d_coeff; o.part = b.d_part; o.cblocks = b.cblocks; o.cdev = b.d_cdev;
  o.pm = b.d_pm; o.nfreq = b.nfreq; o.ndim = ndim_pm(c); o.m_min = b.m_min; o.m_max = b.m_max;
  o.pm_lds = o.ndim * o.nfreq <= kPmLdsBins;
  o.skind = b.d_skind; o.soff = b.d_soff; o.sbond = b.d_sbond; o.sw = b.d_sw; o.srows = b.srows;
  o.nshell = b.nshell; o.nsp = b.nsp; o.sblocks = b.sblocks;
  o.scpart = b.d_scpart; o.sc = b.d_sc;
  o.q = b.d_q; o.pos = b.d_pos;
  for (int k = 0; k < 3; ++k) o.polar[k] = b.polar[k];
  o.have_polar = b.have_polar; o.nq = b.nq; o.every = b.sf_every < 1 ? 1 : b.sf_every;
  o.sfspin = b.d_sfspin; o.sfpart = b.d_sfpart; o.sf_chunk = b.sf_chunk; o.sf_nchunk = b.sf_nchunk;
  o.sf_lat = b.sf_lat;
  o.sf_rep = b.sf_lat ? kSfRepLat : kSfRep;
  for (int k = 0; k < 3; ++k) o.sfL[k] = b.sfL[k];
  o.sfex = b.d_sftab;
  o.sfey = o.sfex ? o.sfex + (size_t)b.sfL[0] * b.nq : nullptr;
  o.sfez = o.sfey ? o.sfey + (size_t)b.sfL[1] * b.nq : nullptr;
  o.sfeb = o.sfez ? o.sfez + (size_t)b.sfL[2] * b.nq : nullptr;
  o.sf = b.d_sf; o.sfn = b.d_sfn;
  o.acc = b.d_acc; o.trace = b.d_trace; o.trace_cap = c->trace_cap;
  return o;
}

int need_model(esp_ctx *c) {
  if (!c) return ESP_EINVAL;
  if (!c->have_model) return fail(c, ESP_ESTATE, "no model");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ESP_EHIP, "hipSetDevice");
  return ESP_OK;
}

// Model-level tables of one observable changed: drop the graph, resize the accumulators.
int changed(esp_ctx *c) {
  drop_graph(c);
  return esp::obs_resize(c);
}

}  // namespace

namespace esp {

void obs_free(esp_ctx *c) {
  auto &b = c->obs;
  dfree(b.d_axes); dfree(b.d_coeff); dfree(b.d_acc); dfree(b.d_part); dfree(b.d_pm);
  dfree(b.d_skind); dfree(b.d_soff); dfree(b.d_sbond); dfree(b.d_sw); dfree(b.d_sc);
  dfree(b.d_scpart); dfree(b.d_q); dfree(b.d_pos); dfree(b.d_sf); dfree(b.d_sfn); dfree(b.d_trace);
  dfree(b.d_sfspin); dfree(b.d_sfpart); dfree(b.d_sftab); dfree(b.d_cdev);
  b = esp_ctx::Obs{};
}

int obs_resize(esp_ctx *c) {
  auto &b = c->obs;
  dfree(b.d_acc); dfree(b.d_part); dfree(b.d_pm); dfree(b.d_sc); dfree(b.d_scpart);
  dfree(b.d_sf); dfree(b.d_sfn); dfree(b.d_trace); dfree(b.d_sfspin); dfree(b.d_sfpart); dfree(b.d_cdev);
  if (!b.flags || c->R == 0) return ESP_OK;
  const int R = c->R, N = c->hm.N, nb = c->hm.nb;
  const int ncell = N / nb;
  int rc;
  if (b.flags & (ESP_OBS_ORDER | ESP_OBS_STAGGERED | ESP_OBS_PM)) {
    // Cells in the device order of their first atom: the device order groups the sites of a
    // colour by basis atom and cell residue, cells in order inside a group, so atom a of
    // consecutive cells in this order sits at consecutive device indices (coalesced gathers).
    const std::vector<int> &r2d = c->hm.ref2dev;
    std::vector<int> ord(ncell), cd((size_t)nb * ncell);
    for (int k = 0; k < ncell; ++k) ord[k] = k;
    std::sort(ord.begin(), ord.end(), [&](int x, int y) { return r2d[(size_t)x * nb] < r2d[(size_t)y * nb]; });
    for (int a = 0; a < nb; ++a)
      for (int k = 0; k < ncell; ++k) cd[(size_t)a * ncell + k] = r2d[(size_t)ord[k] * nb + a];
    if ((rc = upload(c, b.d_cdev, cd))) return rc;
  }
  b.cblocks = (ncell + kObsBlock * kCellsPerThread - 1) / (kObsBlock * kCellsPerThread);
  b.sblocks = (N + kObsBlock - 1) / kObsBlock;
  if ((rc = dalloc(c, b.d_acc, (size_t)R * 6))) return rc;
  if ((rc = dalloc(c, b.d_part, (size_t)R * b.cblocks * 4))) return rc;
  if (b.flags & ESP_OBS_PM)
    if ((rc = dalloc(c, b.d_pm, (size_t)R * ndim_pm(c) * b.nfreq))) return rc;
  if (b.flags & ESP_OBS_SPINCORR) {
    const size_t kk = (size_t)b.nshell * b.nsp * b.nsp;
    if ((rc = dalloc(c, b.d_sc, (size_t)R * 3 * kk))) return rc;
    if ((rc = dalloc(c, b.d_scpart, (size_t)R * b.sblocks * 2 * kk))) return rc;
  }
  if (b.flags & ESP_OBS_SFACTOR) {
    if ((rc = dalloc(c, b.d_sf, (size_t)R * 3 * b.nq))) return rc;
    if ((rc = dalloc(c, b.d_sfn, 1))) return rc;
    // site chunks (lattice: chunks of (x1, x2) rows): enough waves to fill the chip, at least
    // 64 sites each
    const int srep = b.sf_lat ? kSfRepLat : kSfRep;
    const int ng = (R + srep - 1) / srep, nqb = (b.nq + kSfQ - 1) / kSfQ;
    const int units = b.sf_lat ? b.sfL[1] * b.sfL[2] : N;
    int nch = (kSfWaves + ng * nqb - 1) / (ng * nqb);
    nch = std::max(1, std::min({nch, kSfMaxChunks, units, (N + 63) / 64}));
    b.sf_chunk = (units + nch - 1) / nch;
    b.sf_nchunk = (units + b.sf_chunk - 1) / b.sf_chunk;
    if ((rc = dalloc(c, b.d_sfspin, (size_t)ng * N * srep * 3))) return rc;
    if ((rc = dalloc(c, b.d_sfpart, (size_t)b.sf_nchunk * R * 6 * b.nq))) return rc;
  }
  if (c->trace_cap > 0 && (b.flags & (ESP_OBS_ORDER | ESP_OBS_STAGGERED)))
    if ((rc = dalloc(c, b.d_trace, (size_t)c->trace_cap * R * 2))) return rc;
  return obs_clear(c);
}

int obs_clear(esp_ctx *c) {
  auto &b = c->obs;
  const int R = c->R;
  if (!b.flags || R == 0) return ESP_OK;
  HIPCHK(c, hipMemsetAsync(b.d_acc, 0, sizeof(double) * 6 * R, c->stream));
  if (b.d_pm)
    HIPCHK(c, hipMemsetAsync(b.d_pm, 0, sizeof(unsigned long long) * R * ndim_pm(c) * b.nfreq, c->stream));
  if (b.d_sc)
    HIPCHK(c, hipMemsetAsync(b.d_sc, 0, sizeof(double) * R * 3 * b.nshell * b.nsp * b.nsp, c->stream));
  if (b.d_sf) {
    HIPCHK(c, hipMemsetAsync(b.d_sf, 0, sizeof(double) * R * 3 * b.nq, c->stream));
    HIPCHK(c, hipMemsetAsync(b.d_sfn, 0, sizeof(unsigned long long), c->stream));
  }
  return ESP_OK;
}

int obs_enqueue(esp_ctx *c) {
  const auto &b = c->obs;
  if (!b.flags || c->R == 0) return ESP_OK;
  const ObsArgs o = obs_args(c);
  if (b.flags & (ESP_OBS_ORDER | ESP_OBS_STAGGERED | ESP_OBS_PM)) {
    const size_t shm = (b.flags & ESP_OBS_PM) && o.pm_lds ? sizeof(unsigned) * o.ndim * o.nfreq : 0;
    hipLaunchKernelGGL(k_obs_cells, dim3(b.cblocks, c->R), dim3(kObsBlock), shm, c->stream, o);
  }
  if (b.flags & ESP_OBS_SPINCORR) {
    const int rep = 2;     // replicas per thread: 2 beat 4 and 8 (one kind and two kinds measured)
    const dim3 grid(b.sblocks, (c->R + rep - 1) / rep);
    switch (b.nsp) {
      case 1: hipLaunchKernelGGL((k_obs_spincorr<1, 2>), grid, dim3(kObsBlock), 0, c->stream, o); break;
      case 2: hipLaunchKernelGGL((k_obs_spincorr<2, 2>), grid, dim3(kObsBlock), 0, c->stream, o); break;
      case 3: hipLaunchKernelGGL((k_obs_spincorr<3, 2>), grid, dim3(kObsBlock), 0, c->stream, o); break;
      default: hipLaunchKernelGGL((k_obs_spincorr<4, 2>), grid, dim3(kObsBlock), 0, c->stream, o); break;
    }
  }
  if (b.flags & ESP_OBS_SFACTOR) {
    const int srep = b.sf_lat ? kSfRepLat : kSfRep;
    const int ng = (c->R + srep - 1) / srep;
    hipLaunchKernelGGL(k_sf_pack, dim3((o.N + kObsBlock - 1) / kObsBlock, ng), dim3(kObsBlock), 0, c->stream, o);
    if (b.sf_lat)
      hipLaunchKernelGGL(k_sf_lat, dim3((b.nq + kSfQ - 1) / kSfQ, ng, b.sf_nchunk), dim3(kSfQ), 0, c->stream, o);
    else
      hipLaunchKernelGGL(k_sf_main, dim3((b.nq + kSfQ - 1) / kSfQ, ng, b.sf_nchunk), dim3(kSfQ), 0, c->stream, o);
    hipLaunchKernelGGL(k_sf_fin, dim3((b.nq + kObsBlock - 1) / kObsBlock, c->R), dim3(kObsBlock), 0, c->stream, o);
  }
  hipLaunchKernelGGL(k_obs_final, dim3(c->R), dim3(kObsBlock), 0, c->stream, o);
  HIPCHK(c, hipGetLastError());
  return ESP_OK;
}

}  // namespace esp

// ===================================================================================
// C ABI
// ===================================================================================
extern "C" {

int esp_set_orderparam(esp_ctx *c, const double *axes) {
  int rc = need_model(c);
  if (rc) return rc;
  auto &b = c->obs;
  if (!axes) {
    b.flags &= ~ESP_OBS_ORDER;
    dfree(b.d_axes);
    return changed(c);
  }
  std::vector<double> v(axes, axes + 3 * (size_t)c->hm.nb);
  if ((rc = upload(c, b.d_axes, v))) return rc;
  b.flags |= ESP_OBS_ORDER;
  return changed(c);
}

int esp_set_staggered(esp_ctx *c, int enable, const double *coeff) {
  int rc = need_model(c);
  if (rc) return rc;
  auto &b = c->obs;
  if (!enable) {
    b.flags &= ~ESP_OBS_STAGGERED;
    dfree(b.d_coeff);
    return changed(c);
  }
  if (!coeff) return fail(c, ESP_EINVAL, "esp_set_staggered: null coefficients");
  std::vector<double> v(coeff, coeff + c->hm.nb);
  if ((rc = upload(c, b.d_coeff, v))) return rc;
  b.flags |= ESP_OBS_STAGGERED;
  return changed(c);
}

int esp_set_pm(esp_ctx *c, int enable, int *nfreq) {
  int rc = need_model(c);
  if (rc) return rc;
  auto &b = c->obs;
  if (!enable) {
    b.flags &= ~ESP_OBS_PM;
    b.nfreq = 0;
    return changed(c);
  }
  // grid of mcarlo.F90:432-445: m_max = sum over the basis of |moment|, step 0.01
  const esp::HostModel &m = c->hm;
  double mmax = 0.0;
  for (int a = 0; a < m.nb; ++a) mmax += std::fabs(m.mag[m.ref2dev[a]]);
  const double mmin = -mmax;
  int nf = (int)std::lround((mmax - mmin) / 0.01) + 1;
  if (nf == 1) nf = 2;
  if (nf > (1 << 22)) return fail(c, ESP_EINVAL, "esp_set_pm: P(m) grid too large");
  b.nfreq = nf; b.m_min = mmin; b.m_max = mmax;
  b.flags |= ESP_OBS_PM;
  if (nfreq) *nfreq = nf;
  return changed(c);
}

int esp_set_spincorr(esp_ctx *c, int nshell, int nspecies, const int *species, int zs, const int *cnt,
                     const int *nbr) {
  int rc = need_model(c);
  if (rc) return rc;
  auto &b = c->obs;
  dfree(b.d_skind); dfree(b.d_soff); dfree(b.d_sbond); dfree(b.d_sw);
  b.flags &= ~ESP_OBS_SPINCORR;
  b.nshell = b.nsp = 0;
  if (nshell == 0) return changed(c);
  const int N = c->hm.N;
  if (nshell < 0 || nspecies < 1 || nspecies > kMaxSpecies || zs < 1 || !species || !cnt || !nbr)
    return fail(c, ESP_EINVAL, "esp_set_spincorr: bad tables (at most 4 kinds)");
  std::vector<int> sp(species, species + N), ct(cnt, cnt + (size_t)nshell * N);
  std::vector<int> nb((size_t)nshell * zs * N);
  for (int i = 0; i < N; ++i)
    if (sp[i] < 0 || sp[i] >= nspecies) return fail(c, ESP_EINVAL, "esp_set_spincorr: kind out of range");
  for (int s = 0; s < nshell; ++s)
    for (int i = 0; i < N; ++i) {
      const int n = ct[(size_t)s * N + i];
      if (n < 0 || n > zs) return fail(c, ESP_EINVAL, "esp_set_spincorr: count out of range");
      for (int l = 0; l < zs; ++l) {
        const int j = nbr[((size_t)s * zs + l) * N + i];
        if (l < n && (j < 1 || j > N)) return fail(c, ESP_EINVAL, "esp_set_spincorr: site id out of range");
        nb[((size_t)s * zs + l) * N + i] = l < n ? j - 1 : 0;
      }
    }
  // shell_num_nbors_kind(shell, kind of neighbour, kind of site) as shell_neighbors_get leaves
  // it (jij.F90:459-476): the per-kind counts of the last site of each kind; spincorr_get
  // reads it as (shell, kind i, kind j).
  std::vector<double> norm((size_t)nshell * nspecies * nspecies, 0.0);
  std::vector<int> ck(nspecies);
  for (int i = 0; i < N; ++i)
    for (int s = 0; s < nshell; ++s) {
      std::fill(ck.begin(), ck.end(), 0);
      for (int l = 0; l < ct[(size_t)s * N + i]; ++l) ++ck[sp[nb[((size_t)s * zs + l) * N + i]]];
      for (int x = 0; x < nspecies; ++x) norm[((size_t)s * nspecies + x) * nspecies + sp[i]] = ck[x];
    }
  // Bond entries of each shell: (owner, partner, owner's slot, weight owner->partner, weight
  // partner->owner) with the weight 1 / ((n m_i) m_j) of the directed bond i -> j, n the table
  // above read as (shell, kind i, kind j).  A shell whose table is symmetric as a multiset (j
  // listed c times for i exactly when i is listed c times for j) pairs opposite directed bonds
  // (the k-th listing of j by i with the k-th of i by j) into one entry owned by one of its
  // sites, which adds both terms.  Owner: the site whose slot is lower (ties: the lower id),
  // which on a translation-invariant table gives every site the same half of the slots, so
  // a table row points the same way for all sites (coalesced gathers); if that leaves some
  // site more entries than closed trails would (odd-degree sites joined to one extra vertex,
  // a site owns half its bonds rounded up), the trails decide.  Any other shell keeps one
  // entry per directed bond (reverse weight 0).
  if (N > (int)kBondIdx) return fail(c, ESP_EINVAL, "esp_set_spincorr: too many sites");
  const std::vector<int> &r2d = c->hm.ref2dev;
  const std::vector<double> &mag = c->hm.mag;
  auto wgt = [&](int s, int i, int j) {
    return 1.0 / (norm[((size_t)s * nspecies + sp[i]) * nspecies + sp[j]] * mag[r2d[i]] * mag[r2d[j]]);
  };
  struct Ent { int o, p, l; double wf, wr; };
  std::vector<std::vector<Ent>> ents(nshell);
  for (int s = 0; s < nshell; ++s) {
    using Key = std::pair<unsigned long long, int>;     // ((min id, max id), slot)
    std::vector<Key> up, down;                          // slots of the lower / higher id
    std::vector<std::pair<int, int>> self;
    for (int i = 0; i < N; ++i)
      for (int l = 0; l < ct[(size_t)s * N + i]; ++l) {
        const int j = nb[((size_t)s * zs + l) * N + i];
        if (i < j) up.push_back({(unsigned long long)i << 32 | (unsigned)j, l});
        else if (i > j) down.push_back({(unsigned long long)j << 32 | (unsigned)i, l});
        else self.push_back({i, l});
      }
    std::sort(up.begin(), up.end());
    std::sort(down.begin(), down.end());
    bool sym = up.size() == down.size();
    for (size_t e = 0; sym && e < up.size(); ++e) sym = up[e].first == down[e].first;
    std::vector<Ent> &es = ents[s];
    if (sym) {
      const int E = (int)up.size();
      std::vector<int> eu(E), ev(E), deg(N + 1, 0), owner(E), nown(N, 0);
      for (const auto &x : self) ++nown[x.first];
      std::vector<int> nself(nown);
      for (int e = 0; e < E; ++e) {
        eu[e] = (int)(up[e].first >> 32); ev[e] = (int)(up[e].first & 0xffffffffu);
        ++deg[eu[e]]; ++deg[ev[e]];
        owner[e] = down[e].second < up[e].second ? ev[e] : eu[e];
        ++nown[owner[e]];
      }
      const int rows_slot = N ? *std::max_element(nown.begin(), nown.end()) : 0;
      int rows_trail = 0;
      for (int v = 0; v < N; ++v) rows_trail = std::max(rows_trail, nself[v] + (deg[v] + 1) / 2);
      if (rows_slot > rows_trail) {
        for (int v = 0; v < N; ++v)
          if (deg[v] & 1) { eu.push_back(v); ev.push_back(N); ++deg[v]; ++deg[N]; }
        const int Et = (int)eu.size();
        std::vector<int> start(N + 2, 0), adj(2 * (size_t)Et), tr(Et, -1);
        for (int v = 0; v <= N; ++v) start[v + 1] = start[v] + deg[v];
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int e = 0; e < Et; ++e) { adj[fill[eu[e]]++] = e; adj[fill[ev[e]]++] = e; }
        std::vector<int> ptr(start.begin(), start.end() - 1);
        for (int v0 = 0; v0 <= N; ++v0) {  // every degree is even: a trail from v0 ends at v0
          int cur = v0;
          for (;;) {
            int e = -1;
            while (ptr[cur] < start[cur + 1]) {
              const int x = adj[ptr[cur]++];
              if (tr[x] < 0) { e = x; break; }
            }
            if (e < 0) break;
            tr[e] = cur;
            cur = eu[e] == cur ? ev[e] : eu[e];
          }
        }
        for (int e = 0; e < E; ++e) owner[e] = tr[e];
      }
      for (int e = 0; e < E; ++e) {
        const bool lo = owner[e] == eu[e];
        const int o = owner[e], p = lo ? ev[e] : eu[e];
        es.push_back({o, p, lo ? up[e].second : down[e].second, wgt(s, o, p), wgt(s, p, o)});
      }
    } else {
      for (int i = 0; i < N; ++i)
        for (int l = 0; l < ct[(size_t)s * N + i]; ++l) {
          const int j = nb[((size_t)s * zs + l) * N + i];
          if (i != j) es.push_back({i, j, l, wgt(s, i, j), 0.0});
        }
    }
    for (const auto &x : self) es.push_back({x.first, x.first, x.second, wgt(s, x.first, x.first), 0.0});
    std::sort(es.begin(), es.end(), [](const Ent &a, const Ent &b) { return a.o != b.o ? a.o < b.o : a.l < b.l; });
  }
  // Device-order tables: shell s has rows soff[s] .. soff[s+1]-1, as many as its largest
  // number of entries per site; a row holds the partner's device index with its kind in the
  // top bits and the weights (one kind: their sum; more kinds: forward, then reverse after
  // all rows).  Slots past a site's entries point at the site itself with weights 0.
  std::vector<int> kind(N), off(nshell + 1, 0);
  for (int i = 0; i < N; ++i) kind[r2d[i]] = sp[i];
  std::vector<std::vector<int>> per(nshell, std::vector<int>(N, 0));
  for (int s = 0; s < nshell; ++s) {
    for (const Ent &x : ents[s]) ++per[s][x.o];
    off[s + 1] = off[s] + *std::max_element(per[s].begin(), per[s].end());
  }
  const size_t rows = (size_t)std::max(off[nshell], 1);
  std::vector<unsigned> bond(rows * N);
  std::vector<double> wt((nspecies > 1 ? 2 : 1) * rows * N, 0.0);
  for (int s = 0; s < nshell; ++s) {
    for (int i = 0; i < N; ++i)
      for (int l = off[s]; l < off[s + 1]; ++l) bond[(size_t)l * N + r2d[i]] = (unsigned)r2d[i] | (unsigned)sp[i] << kBondKindShift;
    std::vector<int> used(N, 0);
    for (const Ent &x : ents[s]) {
      const size_t t = (size_t)(off[s] + used[x.o]++) * N + r2d[x.o];
      bond[t] = (unsigned)r2d[x.p] | (unsigned)sp[x.p] << kBondKindShift;
      if (nspecies > 1) { wt[t] = x.wf; wt[rows * N + t] = x.wr; }
      else wt[t] = x.wf + x.wr;
    }
  }
  if ((rc = upload(c, b.d_skind, kind))) return rc;
  if ((rc = upload(c, b.d_soff, off))) return rc;
  if ((rc = upload(c, b.d_sbond, bond))) return rc;
  if ((rc = upload(c, b.d_sw, wt))) return rc;
  b.nshell = nshell; b.nsp = nspecies; b.srows = (int)rows;
  b.flags |= ESP_OBS_SPINCORR;
  return changed(c);
}

int esp_set_sfactor(esp_ctx *c, int nq, const double *q, const double *pos, const double *polar, int every) {
  int rc = need_model(c);
  if (rc) return rc;
  auto &b = c->obs;
  dfree(b.d_q); dfree(b.d_pos); dfree(b.d_sftab);
  b.flags &= ~ESP_OBS_SFACTOR;
  b.nq = 0;
  b.sf_lat = 0;
  if (nq == 0) return changed(c);
  if (nq < 0 || !q || !pos || every < 1) return fail(c, ESP_EINVAL, "esp_set_sfactor: bad arguments");
  const int N = c->hm.N;
  std::vector<double> qv(q, q + 3 * (size_t)nq), pv(3 * (size_t)N);
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < 3; ++k) pv[3 * (size_t)c->hm.ref2dev[i] + k] = pos[3 * (size_t)i + k];
  if ((rc = upload(c, b.d_q, qv))) return rc;
  if ((rc = upload(c, b.d_pos, pv))) return rc;
  // Lattice positions: with the supercell L0 x L1 x L2 (as given to esp_set_supercell, else
  // a cube) and r_i = B_a + x0 A0 + x1 A1 + x2 A2 for every site to within a few rounding
  // errors (A_k, B_a read off the cell at the origin and its neighbours), the phases factor
  // into per-q tables, computed here once.
  const int nb = c->hm.nb, nc = N / nb;
  int L[3] = {0, 0, 0};
  if (c->Lhint[0] > 0 && (long)c->Lhint[0] * c->Lhint[1] * c->Lhint[2] == nc) {
    for (int k = 0; k < 3; ++k) L[k] = c->Lhint[k];
  } else {
    const int e = (int)std::lround(std::cbrt((double)nc));
    if (e > 0 && (long)e * e * e == nc) L[0] = L[1] = L[2] = e;
  }
  if (L[0] > 0) {
    auto at = [&](int x0, int x1, int x2, int a) { return pos + 3 * ((((size_t)x2 * L[1] + x1) * L[0] + x0) * nb + a); };
    double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, pmax = 0.0;
    for (int k = 0; k < 3; ++k)
      if (L[k] > 1) {
        const int x[3] = {k == 0, k == 1, k == 2};
        for (int m = 0; m < 3; ++m) A[k][m] = at(x[0], x[1], x[2], 0)[m] - pos[m];
      }
    for (size_t t = 0; t < 3 * (size_t)N; ++t) pmax = std::max(pmax, std::fabs(pos[t]));
    const double tol = 8.0 * 2.220446049250313e-16 * (1.0 + pmax);
    bool lat = true;
    for (int x2 = 0; x2 < L[2] && lat; ++x2)
      for (int x1 = 0; x1 < L[1] && lat; ++x1)
        for (int x0 = 0; x0 < L[0] && lat; ++x0)
          for (int a = 0; a < nb && lat; ++a)
            for (int m = 0; m < 3; ++m) {
              const double r = pos[3 * (size_t)a + m] + x0 * A[0][m] + x1 * A[1][m] + x2 * A[2][m];
              if (std::fabs(at(x0, x1, x2, a)[m] - r) > tol) { lat = false; break; }
            }
    if (lat) {
      std::vector<double2> tab((size_t)(L[0] + L[1] + L[2] + nb) * nq);
      size_t row = 0;
      for (int k = 0; k < 3; ++k)
        for (int x = 0; x < L[k]; ++x, ++row)
          for (int iq = 0; iq < nq; ++iq) {
            const double *qq = q + 3 * (size_t)iq;
            const double ph = qq[0] * (x * A[k][0]) + qq[1] * (x * A[k][1]) + qq[2] * (x * A[k][2]);
            tab[row * nq + iq] = make_double2(std::cos(ph), std::sin(ph));
          }
      for (int a = 0; a < nb; ++a, ++row)
        for (int iq = 0; iq < nq; ++iq) {
          const double *qq = q + 3 * (size_t)iq, *ba = pos + 3 * (size_t)a;
          const double ph = qq[0] * ba[0] + qq[1] * ba[1] + qq[2] * ba[2];
          tab[row * nq + iq] = make_double2(std::cos(ph), std::sin(ph));
        }
      if ((rc = upload(c, b.d_sftab, tab))) return rc;
      b.sf_lat = 1;
      for (int k = 0; k < 3; ++k) b.sfL[k] = L[k];
    }
  }
  b.nq = nq;
  b.sf_every = every;
  b.have_polar = polar != nullptr;
  for (int k = 0; k < 3; ++k) b.polar[k] = polar ? polar[k] : 0.0;
  b.flags |= ESP_OBS_SFACTOR;
  return changed(c);
}

int esp_get_obs(esp_ctx *c, int which, double *out, int64_t *count) {
  if (!c || !out) return ESP_EINVAL;
  auto &b = c->obs;
  if (!(b.flags & which) || c->R == 0) return fail(c, ESP_ESTATE, "esp_get_obs: observable not on");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ESP_EHIP, "hipSetDevice");
  const int R = c->R;
  if (count) {
    unsigned long long n = 0;
    HIPCHK(c, hipMemcpyAsync(&n, which == ESP_OBS_SFACTOR ? b.d_sfn : c->d_nmeas, sizeof n,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *count = (int64_t)n;
  }
  if (which == ESP_OBS_ORDER || which == ESP_OBS_STAGGERED) {
    std::vector<double> a((size_t)R * 6);
    HIPCHK(c, hipMemcpyAsync(a.data(), b.d_acc, sizeof(double) * a.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int o = which == ESP_OBS_ORDER ? 0 : 3;
    for (int r = 0; r < R; ++r)
      for (int k = 0; k < 3; ++k) out[3 * r + k] = a[6 * (size_t)r + o + k];
    return ESP_OK;
  }
  if (which == ESP_OBS_PM) {
    std::vector<unsigned long long> h((size_t)R * ndim_pm(c) * b.nfreq);
    HIPCHK(c, hipMemcpyAsync(h.data(), b.d_pm, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < h.size(); ++k) out[k] = (double)h[k];
    return ESP_OK;
  }
  if (which == ESP_OBS_SPINCORR) {
    HIPCHK(c, hipMemcpyAsync(out, b.d_sc, sizeof(double) * R * 3 * b.nshell * b.nsp * b.nsp,
                             hipMemcpyDeviceToHost, c->stream));
  } else if (which == ESP_OBS_SFACTOR) {
    HIPCHK(c, hipMemcpyAsync(out, b.d_sf, sizeof(double) * R * 3 * b.nq, hipMemcpyDeviceToHost, c->stream));
  } else {
    return fail(c, ESP_EINVAL, "esp_get_obs: unknown observable");
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_get_obs_trace(esp_ctx *c, double *trace, int *n) {
  if (!c || !n) return ESP_EINVAL;
  auto &b = c->obs;
  if (!b.d_trace) return fail(c, ESP_ESTATE, "esp_get_obs_trace: no trace (esp_set_trace, order/staggered on)");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ESP_EHIP, "hipSetDevice");
  unsigned long long nm = 0;
  HIPCHK(c, hipMemcpyAsync(&nm, c->d_nmeas, sizeof nm, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int k = (int)std::min<unsigned long long>(nm, (unsigned long long)c->trace_cap);
  *n = k;
  if (k > 0 && trace) {
    HIPCHK(c, hipMemcpyAsync(trace, b.d_trace, sizeof(double) * 2 * (size_t)k * c->R, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return ESP_OK;
}

}  // extern "C"

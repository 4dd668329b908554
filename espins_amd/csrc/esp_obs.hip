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
// counts through a histogram in LDS), k_obs_spincorr (one thread per site), k_obs_sfactor
// (one thread per q-point and four replicas: the phase of a site is computed once for all
// four) and k_obs_final (per replica: fixed-order sums of the block partials, accumulation
// into the replica's current slot).  Histogram counts are integers, so they do not depend on
// the order of the atomic additions; every floating-point sum has a fixed order.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <vector>

#include "esp_ctx.h"

namespace {
using namespace espk;

constexpr int kObsBlock = 256;
constexpr int kMaxSpecies = 4;        // kinds in the spin-correlation tables
constexpr int kSfBlock = 128;         // q-points per structure-factor block
constexpr int kSfRep = 4;             // replicas per structure-factor block
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
  unsigned long long *pm;
  int nfreq, ndim, pm_lds;
  double m_min, m_max;
  const int *species, *scnt, *snbr;
  const double *snorm;
  int nshell, nsp, zs, sblocks;
  double *scpart, *sc;
  const double *q, *pos;
  double polar[3];
  int have_polar, nq, every;
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

// One thread per cell (nb consecutive reference sites).
__global__ __launch_bounds__(kObsBlock) void k_obs_cells(ObsArgs o) {
  extern __shared__ unsigned s_hist[];
  __shared__ double red[4 * (kObsBlock / 64)];
  const int r = blockIdx.y;
  const int c = blockIdx.x * kObsBlock + threadIdx.x;
  const bool PM = o.flags & ESP_OBS_PM, OP = o.flags & ESP_OBS_ORDER, SM = o.flags & ESP_OBS_STAGGERED;
  const int nh = o.ndim * o.nfreq;
  if (PM && o.pm_lds) {
    for (int k = threadIdx.x; k < nh; k += kObsBlock) s_hist[k] = 0u;
    __syncthreads();
  }
  V3 m = {0.0, 0.0, 0.0}, st = {0.0, 0.0, 0.0};
  double op = 0.0;
  if (c < o.ncell) {
    for (int a = 0; a < o.nb; ++a) {
      const int d = gldi(o.ref2dev, (size_t)c * o.nb + a);
      const V3 s = spin_at(o, r, d);
      m.x += s.x; m.y += s.y; m.z += s.z;
      if (OP) op += s.x * gldi(o.axes, 3 * a) + s.y * gldi(o.axes, 3 * a + 1) + s.z * gldi(o.axes, 3 * a + 2);
      if (SM) {
        const double w = gldi(o.coeff, a);
        st.x += w * s.x; st.y += w * s.y; st.z += w * s.z;
      }
    }
  }
  const int slot = gldi(o.rep_slot, r);
  if (PM) {
    unsigned long long *g = o.pm + (size_t)slot * nh;
    int bins[3], nbin = 0;
    if (c < o.ncell) {
      bins[nbin++] = pm_bin(o, sqrt(m.x * m.x + m.y * m.y + m.z * m.z));
      if (OP) bins[nbin++] = o.nfreq + pm_bin(o, op);
      const int off = (OP ? 2 : 1) * o.nfreq;
      if (SM) bins[nbin++] = off + pm_bin(o, sqrt(st.x * st.x + st.y * st.y + st.z * st.z));
    }
    if (o.pm_lds) {
      for (int k = 0; k < nbin; ++k) atomicAdd(&s_hist[bins[k]], 1u);
      __syncthreads();
      for (int k = threadIdx.x; k < nh; k += kObsBlock) {
        const unsigned v = s_hist[k];
        if (v) gadd(g + k, (unsigned long long)v);
      }
    } else {
      for (int k = 0; k < nbin; ++k) gadd(g + bins[k], 1ull);
    }
  }
  if (OP || SM) {
    double v[4] = {op, st.x, st.y, st.z};
    block_sum<4>(v, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < 4; ++k) gsti(o.part, ((size_t)r * o.cblocks + blockIdx.x) * 4 + k, v[k]);
  }
}

// One thread per site i (reference order): per shell, the sums of S_i . S_j / (n m_i m_j) and
// of |S_i . S_j| / (n m_i m_j) over the shell's neighbours j of each kind, as block partials
// [shell][2][kind i][kind j].
__global__ __launch_bounds__(kObsBlock) void k_obs_spincorr(ObsArgs o) {
  __shared__ double red[2 * kMaxSpecies * kMaxSpecies * (kObsBlock / 64)];
  const int r = blockIdx.y;
  const int i = blockIdx.x * kObsBlock + threadIdx.x;
  const bool on = i < o.N;
  const int ns = o.nsp;
  int ki = 0;
  V3 si = {0.0, 0.0, 0.0};
  double mi = 1.0;
  if (on) {
    const int d = gldi(o.ref2dev, i);
    ki = gldi(o.species, i);
    si = spin_at(o, r, d);
    mi = gldi(o.mag, d);
  }
  for (int s = 0; s < o.nshell; ++s) {
    double v[2 * kMaxSpecies * kMaxSpecies];
#pragma unroll
    for (int k = 0; k < 2 * kMaxSpecies * kMaxSpecies; ++k) v[k] = 0.0;
    if (on) {
      const int n = gldi(o.scnt, (size_t)s * o.N + i);
      for (int l = 0; l < n; ++l) {
        const int j = gldi(o.snbr, ((size_t)s * o.zs + l) * o.N + i);
        const int dj = gldi(o.ref2dev, j);
        const int kj = gldi(o.species, j);
        const V3 sj = spin_at(o, r, dj);
        const double den = gldi(o.snorm, ((size_t)s * ns + ki) * ns + kj) * mi * gldi(o.mag, dj);
        const double dd = si.x * sj.x + si.y * sj.y + si.z * sj.z;
        const double t = dd / den, ta = fabs(dd) / den;
#pragma unroll
        for (int x = 0; x < kMaxSpecies; ++x)   // static register indices
#pragma unroll
          for (int y = 0; y < kMaxSpecies; ++y)
            if (x == ki && y == kj) {
              v[x * kMaxSpecies + y] += t;
              v[kMaxSpecies * kMaxSpecies + x * kMaxSpecies + y] += ta;
            }
      }
    }
    block_sum<2 * kMaxSpecies * kMaxSpecies>(v, red);
    if (threadIdx.x == 0) {
      double *p = o.scpart + (((size_t)r * o.sblocks + blockIdx.x) * o.nshell + s) * 2 * ns * ns;
#pragma unroll
      for (int w = 0; w < 2; ++w)
#pragma unroll
        for (int a = 0; a < kMaxSpecies; ++a)
#pragma unroll
          for (int b = 0; b < kMaxSpecies; ++b)
            if (a < ns && b < ns)
              gsti(p, (size_t)(w * ns + a) * ns + b, v[w * kMaxSpecies * kMaxSpecies + a * kMaxSpecies + b]);
    }
  }
}

// One thread per q-point, kSfRep replicas per block: F0(q) = sum_j S_j exp(i q.r_j), then the
// component of F0 perpendicular to q, F, and NSF = Re(polar . F) / |polar|; adds |F|^2,
// |F|^2 - NSF^2 and NSF^2 to the slot's sums (one writer per (slot, q)).  q = 0 adds nothing.
__global__ __launch_bounds__(kSfBlock) void k_obs_sfactor(ObsArgs o) {
  __shared__ double s_pos[kSfBlock][3];
  __shared__ double s_sp[kSfRep][kSfBlock][3];
  const unsigned long long cnt = gldi(o.nmeas, 0) + 1ull;
  if (cnt % (unsigned long long)o.every) return;       // uniform over the grid
  const int iq = blockIdx.x * kSfBlock + threadIdx.x;
  const int r0 = blockIdx.y * kSfRep;
  double qx = 0.0, qy = 0.0, qz = 0.0;
  if (iq < o.nq) { qx = gldi(o.q, 3 * iq); qy = gldi(o.q, 3 * iq + 1); qz = gldi(o.q, 3 * iq + 2); }
  double fr[kSfRep][3], fi[kSfRep][3];
#pragma unroll
  for (int u = 0; u < kSfRep; ++u)
#pragma unroll
    for (int k = 0; k < 3; ++k) { fr[u][k] = 0.0; fi[u][k] = 0.0; }
  for (int j0 = 0; j0 < o.N; j0 += kSfBlock) {
    const int j = j0 + threadIdx.x;
    __syncthreads();
    if (j < o.N) {
      for (int k = 0; k < 3; ++k) s_pos[threadIdx.x][k] = gldi(o.pos, 3 * (size_t)j + k);
      const int d = gldi(o.ref2dev, j);
      for (int u = 0; u < kSfRep; ++u) {
        const V3 s = r0 + u < o.R ? spin_at(o, r0 + u, d) : V3{0.0, 0.0, 0.0};
        s_sp[u][threadIdx.x][0] = s.x; s_sp[u][threadIdx.x][1] = s.y; s_sp[u][threadIdx.x][2] = s.z;
      }
    }
    __syncthreads();
    const int nt = min(kSfBlock, o.N - j0);
    for (int t = 0; t < nt; ++t) {
      double sn, cs;
      sincos(qx * s_pos[t][0] + qy * s_pos[t][1] + qz * s_pos[t][2], &sn, &cs);
#pragma unroll
      for (int u = 0; u < kSfRep; ++u)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          fr[u][k] += s_sp[u][t][k] * cs;
          fi[u][k] += s_sp[u][t][k] * sn;
        }
    }
  }
  const double q2 = qx * qx + qy * qy + qz * qz;
  if (iq >= o.nq || q2 == 0.0) return;
  const double pn = o.have_polar ? sqrt(o.polar[0] * o.polar[0] + o.polar[1] * o.polar[1] +
                                        o.polar[2] * o.polar[2]) : 1.0;
  for (int u = 0; u < kSfRep && r0 + u < o.R; ++u) {
    // perpendicular part: F - (F . q / q^2) q (real and imaginary parts separately)
    const double ar = (fr[u][0] * qx + fr[u][1] * qy + fr[u][2] * qz) / q2;
    const double ai = (fi[u][0] * qx + fi[u][1] * qy + fi[u][2] * qz) / q2;
    const double pr[3] = {fr[u][0] - ar * qx, fr[u][1] - ar * qy, fr[u][2] - ar * qz};
    const double pi[3] = {fi[u][0] - ai * qx, fi[u][1] - ai * qy, fi[u][2] - ai * qz};
    const double f2 = pr[0] * pr[0] + pr[1] * pr[1] + pr[2] * pr[2] + pi[0] * pi[0] + pi[1] * pi[1] +
                      pi[2] * pi[2];
    double nsf = 0.0;
    if (o.have_polar) nsf = (o.polar[0] * pr[0] + o.polar[1] * pr[1] + o.polar[2] * pr[2]) / pn;
    const int slot = gldi(o.rep_slot, r0 + u);
    double *acc = o.sf + (size_t)slot * 3 * o.nq;
    gsti(acc, iq, gldi(acc, iq) + f2);
    gsti(acc, (size_t)o.nq + iq, gldi(acc, (size_t)o.nq + iq) + (f2 - nsf * nsf));
    gsti(acc, 2 * (size_t)o.nq + iq, gldi(acc, 2 * (size_t)o.nq + iq) + nsf * nsf);
  }
}

// One block per replica: fixed-order sums of the block partials, accumulation by slot.
__global__ __launch_bounds__(kObsBlock) void k_obs_final(ObsArgs o) {
  const int r = blockIdx.x;
  const int slot = gldi(o.rep_slot, r);
  const unsigned long long nm = gldi(o.nmeas, 0);
  if ((o.flags & (ESP_OBS_ORDER | ESP_OBS_STAGGERED)) && threadIdx.x == 0) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < o.cblocks; ++b)
      for (int k = 0; k < 4; ++k) v[k] += gldi(o.part, ((size_t)r * o.cblocks + b) * 4 + k);
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
  if (o.flags & ESP_OBS_SPINCORR) {
    const int ns = o.nsp, kk = ns * ns;
    for (int e = threadIdx.x; e < o.nshell * kk; e += kObsBlock) {
      const int s = e / kk, ab = e % kk;
      double dot = 0.0, ab_sum = 0.0;
      for (int b = 0; b < o.sblocks; ++b) {
        const double *p = o.scpart + (((size_t)r * o.sblocks + b) * o.nshell + s) * 2 * kk;
        dot += gldi(p, ab);
        ab_sum += gldi(p, (size_t)kk + ab);
      }
      double *a = o.sc + (size_t)slot * 3 * o.nshell * kk;
      const size_t q = (size_t)s * kk + ab;
      gsti(a, q, gldi(a, q) + dot);
      gsti(a, (size_t)o.nshell * kk + q, gldi(a, (size_t)o.nshell * kk + q) + fabs(dot));
      gsti(a, 2 * (size_t)o.nshell * kk + q, gldi(a, 2 * (size_t)o.nshell * kk + q) + ab_sum);
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
  o.axes = b.d_axes; o.coeff = b.d_coeff; o.part = b.d_part; o.cblocks = b.cblocks;
  o.pm = b.d_pm; o.nfreq = b.nfreq; o.ndim = ndim_pm(c); o.m_min = b.m_min; o.m_max = b.m_max;
  o.pm_lds = o.ndim * o.nfreq <= kPmLdsBins;
  o.species = b.d_species; o.scnt = b.d_scnt; o.snbr = b.d_snbr; o.snorm = b.d_snorm;
  o.nshell = b.nshell; o.nsp = b.nsp; o.zs = b.zs; o.sblocks = b.sblocks;
  o.scpart = b.d_scpart; o.sc = b.d_sc;
  o.q = b.d_q; o.pos = b.d_pos;
  for (int k = 0; k < 3; ++k) o.polar[k] = b.polar[k];
  o.have_polar = b.have_polar; o.nq = b.nq; o.every = b.sf_every < 1 ? 1 : b.sf_every;
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
  dfree(b.d_species); dfree(b.d_scnt); dfree(b.d_snbr); dfree(b.d_snorm); dfree(b.d_sc);
  dfree(b.d_scpart); dfree(b.d_q); dfree(b.d_pos); dfree(b.d_sf); dfree(b.d_sfn); dfree(b.d_trace);
  b = esp_ctx::Obs{};
}

int obs_resize(esp_ctx *c) {
  auto &b = c->obs;
  dfree(b.d_acc); dfree(b.d_part); dfree(b.d_pm); dfree(b.d_sc); dfree(b.d_scpart);
  dfree(b.d_sf); dfree(b.d_sfn); dfree(b.d_trace);
  if (!b.flags || c->R == 0) return ESP_OK;
  const int R = c->R, N = c->hm.N;
  const int ncell = N / c->hm.nb;
  b.cblocks = (ncell + kObsBlock - 1) / kObsBlock;
  b.sblocks = (N + kObsBlock - 1) / kObsBlock;
  int rc;
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
  if (b.flags & ESP_OBS_SPINCORR)
    hipLaunchKernelGGL(k_obs_spincorr, dim3(b.sblocks, c->R), dim3(kObsBlock), 0, c->stream, o);
  if (b.flags & ESP_OBS_SFACTOR)
    hipLaunchKernelGGL(k_obs_sfactor, dim3((b.nq + kSfBlock - 1) / kSfBlock, (c->R + kSfRep - 1) / kSfRep),
                       dim3(kSfBlock), 0, c->stream, o);
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
  dfree(b.d_species); dfree(b.d_scnt); dfree(b.d_snbr); dfree(b.d_snorm);
  b.flags &= ~ESP_OBS_SPINCORR;
  b.nshell = b.nsp = b.zs = 0;
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
  if ((rc = upload(c, b.d_species, sp))) return rc;
  if ((rc = upload(c, b.d_scnt, ct))) return rc;
  if ((rc = upload(c, b.d_snbr, nb))) return rc;
  if ((rc = upload(c, b.d_snorm, norm))) return rc;
  b.nshell = nshell; b.nsp = nspecies; b.zs = zs;
  b.flags |= ESP_OBS_SPINCORR;
  return changed(c);
}

int esp_set_sfactor(esp_ctx *c, int nq, const double *q, const double *pos, const double *polar, int every) {
  int rc = need_model(c);
  if (rc) return rc;
  auto &b = c->obs;
  dfree(b.d_q); dfree(b.d_pos);
  b.flags &= ~ESP_OBS_SFACTOR;
  b.nq = 0;
  if (nq == 0) return changed(c);
  if (nq < 0 || !q || !pos || every < 1) return fail(c, ESP_EINVAL, "esp_set_sfactor: bad arguments");
  std::vector<double> qv(q, q + 3 * (size_t)nq), pv(pos, pos + 3 * (size_t)c->hm.N);
  if ((rc = upload(c, b.d_q, qv))) return rc;
  if ((rc = upload(c, b.d_pos, pv))) return rc;
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

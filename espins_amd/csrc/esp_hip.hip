// esp_hip.hip -- MI355X (gfx950) engine for the ESpinS Metropolis sweep and
// parallel-tempering exchange, behind the C ABI of include/espins_hip.h.
//
// Device layout (one context):
//   spins   double4 sp[R][N]   x,y,z = spin, w = sum of accepted dE at this site since the
//                              last rebase (running energy = e_base[r] + sum_i w_i, so the
//                              "energy by adding differences" of mcarlo.F90:748 is kept
//                              without atomics and is reproducible run to run)
//   sites   colour-major: colour c owns device indices [col_off[c], col_off[c+1])
//   bonds   nbr[zmax][N] (device indices, self-padded), coefficients per (basis, slot)
//           in LDS when the lattice is periodic, else coefJ/B/D[zmax][N]
//   PT      slot_rep[R] / rep_slot[R]: replica <-> temperature slot labels; a swap
//           (pt.F90:1275-1285) exchanges labels, never spins
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/espins_hip.h"
#include "esp_device.h"
#include "esp_model.h"

using namespace espd;

namespace {

constexpr int kBlock = 256;       // phase / reduce kernels: 4 waves
constexpr int kLdsBlock = 512;    // LDS-resident kernel: 8 waves per replica
constexpr int kRedBlocks = 64;    // reduce blocks per replica (upper bound)

struct Clock {                    // device-side schedule counters (graph-replayable)
  unsigned long long sweep;       // Philox sweep counter of the next sweep
  unsigned long long step;        // exchange step counter
  int parity;                     // next exchange start (1 or 2)
  int pad;
};

struct KArgs {
  double4 *sp;
  const int *nbr;
  const double *cJ, *cB, *cD;     // non-compressed coefficients
  const double *ctab;             // compressed [nb][zmax][5]
  const double *mag;
  const uint8_t *basis;
  const int *dev2ref;
  const int *col_off;
  const double *si4, *h4;
  const double *beta_slot;
  const int *rep_slot;
  unsigned long long *rej;
  const Clock *clock;
  const double *prop, *unif;      // host-fed mode, reference order
  uint8_t *accepted;
  int N, nb, zmax, R, ncol;
  int c_begin, c_count, nblk;
  int has_si, has_h, xcd_remap;
  int sweep_off;
  double tilt;
  uint32_t key0, key1, stream_base;
};

enum { M_RANDOM = 0, M_CONST = 1, M_HOSTFED = 2 };

// One attempted update of device site d of replica r (metropolis, get_quant.F90:52-182,
// with delta_energy_get, get_quant.F90:264-337).  sp points at replica r's spins (global
// or LDS).  Returns true when the move is rejected.
template <int MODE, bool CMP, bool HB, bool HD, typename SPP>
__device__ __forceinline__ bool update_site(const KArgs &a, SPP sp, const double *ctab, int d,
                                            int r, double beta, unsigned long long sweep) {
  const double4 s = sp[d];
  const V3 so = xyz(s);
  const int ir = a.dev2ref[d];
  V3 sn;
  double u3;
  if (MODE == M_HOSTFED) {
    const double *p = a.prop + 3 * ((size_t)r * a.N + ir);
    sn = {p[0], p[1], p[2]};
    u3 = a.unif[(size_t)r * a.N + ir];
  } else {
    const uint4 rn = philox10(make_uint4((uint32_t)ir, (uint32_t)sweep, (uint32_t)(sweep >> 32),
                                         a.stream_base + (uint32_t)r),
                              a.key0, a.key1);
    const double mag = a.mag[d];
    sn = MODE == M_RANDOM ? propose_random(mag, u01(rn.x), u01(rn.y))
                          : propose_const(mag, so, a.tilt, u01(rn.x), u01(rn.y));
    u3 = u01(rn.z);
  }
  const V3 ds = sub(sn, so);
  const V3 si = sub(sn, ds);  // old spin as the reference recovers it (get_quant.F90:306,320)
  double eJ = 0.0, eB = 0.0, eD = 0.0;
  const int b = a.basis[d];
  const double *crow = CMP ? ctab + (size_t)b * a.zmax * 5 : nullptr;
#pragma unroll 2
  for (int l = 0; l < a.zmax; ++l) {
    const int j = a.nbr[(size_t)l * a.N + d];
    const V3 sj = (j == d) ? sn : xyz(sp[j]);
    const double dsj = dot(ds, sj);
    double cj, cb = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    if (CMP) {
      cj = crow[5 * l];
      if (HB) cb = crow[5 * l + 1];
      if (HD) { cx = crow[5 * l + 2]; cy = crow[5 * l + 3]; cz = crow[5 * l + 4]; }
    } else {
      cj = a.cJ[(size_t)l * a.N + d];
      if (HB) cb = a.cB[(size_t)l * a.N + d];
      if (HD) {
        const double *dv = a.cD + 3 * ((size_t)l * a.N + d);
        cx = dv[0]; cy = dv[1]; cz = dv[2];
      }
    }
    eJ = eJ - cj * dsj;
    if (HB) {
      const double sij = dot(si, sj);
      eB = eB + cb * (dsj * dsj + 2.0 * sij * dsj);
    }
    if (HD) {
      // D . (dS x S_j)  (utility_cross, utility.F90:212-234)
      eD = eD + (cx * (ds.y * sj.z - ds.z * sj.y) + cy * (ds.z * sj.x - ds.x * sj.z) +
                 cz * (ds.x * sj.y - ds.y * sj.x));
    }
  }
  double de = eJ;
  if (a.has_h) {
    const double *h = a.h4 + 4 * b;
    de = de + h[0] * (ds.x * h[1] + ds.y * h[2] + ds.z * h[3]);
  }
  if (a.has_si) {
    const double *k = a.si4 + 4 * b;
    const double sa = si.x * k[1] + si.y * k[2] + si.z * k[3];
    const double dsa = ds.x * k[1] + ds.y * k[2] + ds.z * k[3];
    de = de + k[0] * (dsa * dsa + 2.0 * sa * dsa);
  }
  if (HB) de = de + eB;
  if (HD) de = de + eD;
  // reject iff dE > 0 and exp(-beta dE) < u (get_quant.F90:169-177)
  const bool rej = (de > 0.0) && (exp(-beta * de) < u3);
  if (MODE == M_HOSTFED) a.accepted[(size_t)r * a.N + ir] = rej ? 0 : 1;
  if (!rej) sp[d] = make_double4(sn.x, sn.y, sn.z, s.w + de);
  return rej;
}

__device__ __forceinline__ void decode_block(const KArgs &a, int &r, int &b) {
  const int L = blockIdx.x;
  if (a.xcd_remap) {  // blocks go round-robin over the 8 XCDs: keep a replica on one XCD
    const int x = L & 7, q = L >> 3;
    r = x + 8 * (q / a.nblk);
    b = q % a.nblk;
  } else {
    r = L / a.nblk;
    b = L % a.nblk;
  }
}

// One colour phase over every replica (global-memory path).
template <int MODE, bool CMP, bool HB, bool HD>
__global__ __launch_bounds__(kBlock) void k_phase(KArgs a) {
  extern __shared__ double s_ctab[];
  int r, b;
  decode_block(a, r, b);
  if (r >= a.R) return;
  if (CMP) {
    const int n = a.nb * a.zmax * 5;
    for (int k = threadIdx.x; k < n; k += kBlock) s_ctab[k] = a.ctab[k];
    __syncthreads();
  }
  const int k = b * kBlock + threadIdx.x;
  const int slot = a.rep_slot[r];
  const double beta = a.beta_slot[slot];
  const unsigned long long sweep = a.clock->sweep + (unsigned long long)a.sweep_off;
  bool rej = false;
  if (k < a.c_count)
    rej = update_site<MODE, CMP, HB, HD>(a, a.sp + (size_t)r * a.N, s_ctab, a.c_begin + k, r, beta,
                                         sweep);
  const unsigned long long m = __ballot(rej);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&a.rej[slot], (unsigned long long)__popcll(m));
}

struct LdsMeas {
  double *acc;                    // [R][6] by slot
  double *e_run;                  // [R] by replica
  const double *e_base;
  unsigned long long *nmeas;
  int measure_every;              // 0 = never
};

// All colours of nsweeps sweeps of one replica per block, spins resident in LDS,
// measurement (mcarlo.F90:536-543) fused every measure_every sweeps.
template <int MODE, bool CMP, bool HB, bool HD>
__global__ __launch_bounds__(kLdsBlock) void k_lds(KArgs a, int nsweeps, LdsMeas m) {
  extern __shared__ double4 s_sp[];
  __shared__ double red[4 * (kLdsBlock / 64)];
  const int r = blockIdx.x;
  double *ctab = reinterpret_cast<double *>(s_sp + a.N);
  double4 *g = a.sp + (size_t)r * a.N;
  for (int d = threadIdx.x; d < a.N; d += kLdsBlock) s_sp[d] = g[d];
  if (CMP)
    for (int k = threadIdx.x; k < a.nb * a.zmax * 5; k += kLdsBlock) ctab[k] = a.ctab[k];
  __syncthreads();
  const int slot = a.rep_slot[r];
  const double beta = a.beta_slot[slot];
  const unsigned long long sweep0 = a.clock->sweep;
  unsigned int nrej = 0;
  for (int sw = 0; sw < nsweeps; ++sw) {
    const unsigned long long sweep = sweep0 + (unsigned long long)sw;
    for (int c = 0; c < a.ncol; ++c) {
      const int beg = a.col_off[c], end = a.col_off[c + 1];
      for (int d = beg + threadIdx.x; d < end; d += kLdsBlock)
        nrej += update_site<MODE, CMP, HB, HD>(a, s_sp, ctab, d, r, beta, sweep) ? 1u : 0u;
      __syncthreads();
    }
    const bool last = sw + 1 == nsweeps;
    const bool meas = m.measure_every > 0 && (sw + 1) % m.measure_every == 0;
    if (meas || last) {
      double v[4] = {0, 0, 0, 0};
      for (int d = threadIdx.x; d < a.N; d += kLdsBlock) {
        const double4 s = s_sp[d];
        v[0] += s.w; v[1] += s.x; v[2] += s.y; v[3] += s.z;
      }
      block_sum<4>(v, red);
      if (threadIdx.x == 0) {
        const double e = m.e_base[r] + v[0];
        if (meas) {
          const double n = (double)a.N;
          const double mx = v[1] / n, my = v[2] / n, mz = v[3] / n;
          const double m2 = mx * mx + my * my + mz * mz;
          double *acc = m.acc + 6 * slot;
          const double e2 = e * e;
          acc[0] += e; acc[1] += e2; acc[2] += e2 * e2;
          acc[3] += sqrt(m2); acc[4] += m2; acc[5] += m2 * m2;
          if (r == 0) m.nmeas[0] += 1;
        }
        if (last) m.e_run[r] = e;
      }
    }
  }
  for (int d = threadIdx.x; d < a.N; d += kLdsBlock) g[d] = s_sp[d];
  // per-wave sum of rejections, one atomic per wave
  unsigned int v = nrej;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(&a.rej[slot], (unsigned long long)v);
}

// Partial sums of (w, x, y, z) per (replica, block) for the running energy and m.
__global__ __launch_bounds__(kBlock) void k_reduce(const double4 *sp, int N, int nblk,
                                                   double *partial) {
  __shared__ double red[4 * (kBlock / 64)];
  const int r = blockIdx.y, b = blockIdx.x;
  const int chunk = (N + nblk - 1) / nblk;
  const int beg = b * chunk, end = min(N, beg + chunk);
  double v[4] = {0, 0, 0, 0};
  const double4 *s = sp + (size_t)r * N;
  for (int d = beg + threadIdx.x; d < end; d += kBlock) {
    const double4 t = s[d];
    v[0] += t.w; v[1] += t.x; v[2] += t.y; v[3] += t.z;
  }
  block_sum<4>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) partial[((size_t)r * nblk + b) * 4 + k] = v[k];
}

struct FinArgs {
  const double *partial;          // [R][nblk][4] or null (e_run already fresh)
  int nblk, N, R;
  const double *e_base;
  double *e_run;
  double *acc;
  unsigned long long *nmeas;
  int *slot_rep, *rep_slot;
  const double *beta_slot;
  unsigned long long *nswap;      // [R][2]
  Clock *clock;
  int measure, exchange, advance_sweeps;
  uint32_t key0, key1, stream_base;
};

// Running energies / measurement, then the replica exchange (pt_swap, pt.F90:1235-1288),
// then the schedule clock.  One block.
__global__ __launch_bounds__(1024) void k_finalize(FinArgs f) {
  if (f.partial)
    for (int r = threadIdx.x; r < f.R; r += blockDim.x) {
      double v[4] = {0, 0, 0, 0};
      const double *p = f.partial + (size_t)r * f.nblk * 4;
      for (int b = 0; b < f.nblk; ++b)
        for (int k = 0; k < 4; ++k) v[k] += p[4 * b + k];
      const double e = f.e_base[r] + v[0];
      f.e_run[r] = e;
      if (f.measure) {
        const double n = (double)f.N;
        const double mx = v[1] / n, my = v[2] / n, mz = v[3] / n;
        const double m2 = mx * mx + my * my + mz * mz;
        double *acc = f.acc + 6 * f.rep_slot[r];
        const double e2 = e * e;
        acc[0] += e; acc[1] += e2; acc[2] += e2 * e2;
        acc[3] += sqrt(m2); acc[4] += m2; acc[5] += m2 * m2;
      }
    }
  __syncthreads();
  if (f.exchange) {
    const int start = f.clock->parity;  // 1 or 2 (pt.F90:679,898)
    const unsigned long long step = f.clock->step;
    for (int t = start - 1 + 2 * threadIdx.x; t + 1 < f.R; t += 2 * blockDim.x) {
      const int ra = f.slot_rep[t], rb = f.slot_rep[t + 1];
      const double fac = (f.beta_slot[t] - f.beta_slot[t + 1]) * (f.e_run[ra] - f.e_run[rb]);
      bool sw = fac > 0.0;
      if (!sw) {
        const uint4 rn = philox10(make_uint4((uint32_t)t, (uint32_t)step, (uint32_t)(step >> 32),
                                             0xFFFFFFFFu - f.stream_base),
                                  f.key0 ^ 0x5bd1e995u, f.key1);
        sw = u01(rn.x) < exp(fac);
      }
      f.nswap[2 * t] += 1;
      f.nswap[2 * (t + 1)] += 1;
      if (sw) {
        f.slot_rep[t] = rb; f.slot_rep[t + 1] = ra;
        f.rep_slot[rb] = t; f.rep_slot[ra] = t + 1;
        f.nswap[2 * t + 1] += 1;
        f.nswap[2 * (t + 1) + 1] += 1;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (f.measure) f.nmeas[0] += 1;
    if (f.exchange) {
      f.clock->step += 1;
      f.clock->parity = 3 - f.clock->parity;
    }
    f.clock->sweep += (unsigned long long)f.advance_sweeps;
  }
}

// Direct energy (energy_get, get_quant.F90:186-261): per-site terms, block partials.
template <bool CMP, bool HB, bool HD>
__global__ __launch_bounds__(kBlock) void k_energy(KArgs a, int nblk, double *partial) {
  extern __shared__ double s_ctab[];
  __shared__ double red[kBlock / 64];
  const int r = blockIdx.y, bk = blockIdx.x;
  if (CMP) {
    for (int k = threadIdx.x; k < a.nb * a.zmax * 5; k += kBlock) s_ctab[k] = a.ctab[k];
    __syncthreads();
  }
  const int chunk = (a.N + nblk - 1) / nblk;
  const int beg = bk * chunk, end = min(a.N, beg + chunk);
  const double4 *sp = a.sp + (size_t)r * a.N;
  double v[1] = {0.0};
  for (int d = beg + threadIdx.x; d < end; d += kBlock) {
    const V3 s = xyz(sp[d]);
    const int b = a.basis[d];
    double e = 0.0;
    for (int l = 0; l < a.zmax; ++l) {
      const int j = a.nbr[(size_t)l * a.N + d];
      const V3 t = xyz(sp[j]);
      const double dd = dot(s, t);
      double cj, cb = 0, cx = 0, cy = 0, cz = 0;
      if (CMP) {
        const double *c = s_ctab + ((size_t)b * a.zmax + l) * 5;
        cj = c[0]; cb = c[1]; cx = c[2]; cy = c[3]; cz = c[4];
      } else {
        cj = a.cJ[(size_t)l * a.N + d];
        if (HB) cb = a.cB[(size_t)l * a.N + d];
        if (HD) {
          const double *dv = a.cD + 3 * ((size_t)l * a.N + d);
          cx = dv[0]; cy = dv[1]; cz = dv[2];
        }
      }
      e = e - cj * dd;
      if (HB) e = e + cb * dd * dd;
      if (HD)
        e = e + (cx * (s.y * t.z - s.z * t.y) + cy * (s.z * t.x - s.x * t.z) +
                 cz * (s.x * t.y - s.y * t.x));
    }
    if (a.has_h) {
      const double *h = a.h4 + 4 * b;
      e = e + 2.0 * h[0] * (s.x * h[1] + s.y * h[2] + s.z * h[3]);
    }
    if (a.has_si) {
      const double *k = a.si4 + 4 * b;
      const double sa = s.x * k[1] + s.y * k[2] + s.z * k[3];
      e = e + 2.0 * k[0] * sa * sa;
    }
    v[0] += e;
  }
  block_sum<1>(v, red);
  if (threadIdx.x == 0) partial[(size_t)r * nblk + bk] = v[0];
}

__global__ void k_energy_fin(const double *partial, int nblk, int R, double *E) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += partial[(size_t)r * nblk + b];
  E[r] = s / 2.0;
}

// Spins in: reference order [R][N][3] -> device order double4 with w = 0.
__global__ void k_put_spins(const double *in, const int *ref2dev, int N, int R, double4 *sp) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * R) return;
  const int r = (int)(t / N), i = (int)(t % N);
  const double *p = in + 3 * t;
  sp[(size_t)r * N + ref2dev[i]] = make_double4(p[0], p[1], p[2], 0.0);
}

__global__ void k_get_spins(const double4 *sp, const int *ref2dev, const int *src_rep, int N,
                            int R, double *out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * R) return;
  const int q = (int)(t / N), i = (int)(t % N);
  const int r = src_rep ? src_rep[q] : q;
  const double4 s = sp[(size_t)r * N + ref2dev[i]];
  double *p = out + 3 * t;
  p[0] = s.x; p[1] = s.y; p[2] = s.z;
}

// Fold the per-site dE accumulators into e_base (after a direct-energy rebase).
__global__ void k_zero_w(double4 *sp, size_t n) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) sp[t].w = 0.0;
}

}  // namespace

// ===================================================================================
// Context
// ===================================================================================
struct esp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  size_t lds_max = 65536;
  int Lhint[3] = {0, 0, 0};       // supercell size for the periodic colouring
  int path = 0;                   // 0 auto, 1 global-memory phases, 2 LDS-resident
  bool lds_fits = false;
  // model
  esp::HostModel hm;
  bool have_model = false;
  int *d_nbr = nullptr, *d_dev2ref = nullptr, *d_ref2dev = nullptr, *d_col_off = nullptr;
  double *d_cJ = nullptr, *d_cB = nullptr, *d_cD = nullptr, *d_ctab = nullptr;
  double *d_mag = nullptr, *d_si4 = nullptr, *d_h4 = nullptr;
  uint8_t *d_basis = nullptr;
  bool use_lds = false;
  size_t lds_bytes = 0, ctab_bytes = 0;
  // replicas
  int R = 0;
  double4 *d_sp = nullptr;
  double *d_beta = nullptr, *d_ebase = nullptr, *d_erun = nullptr, *d_acc = nullptr;
  double *d_partial = nullptr;
  int red_blocks = 1;
  int *d_slot_rep = nullptr, *d_rep_slot = nullptr;
  unsigned long long *d_rej = nullptr, *d_nswap = nullptr, *d_nmeas = nullptr;
  Clock *d_clock = nullptr;
  std::vector<double> tems;
  // host-fed staging
  double *d_prop = nullptr, *d_unif = nullptr;
  uint8_t *d_accd = nullptr;
  // graph cache
  hipGraphExec_t gexec = nullptr;
  std::string gkey;
};

namespace {

int fail(esp_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                              \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(ctx, ESP_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

template <class T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

template <class T>
int dalloc(esp_ctx *c, T *&p, size_t n) {
  dfree(p);
  if (n == 0) n = 1;
  if (hipMalloc((void **)&p, n * sizeof(T)) != hipSuccess) {
    p = nullptr;
    return fail(c, ESP_ENOMEM, "hipMalloc failed");
  }
  return ESP_OK;
}

template <class T>
int upload(esp_ctx *c, T *&p, const std::vector<T> &v) {
  int rc = dalloc(c, p, v.size());
  if (rc) return rc;
  if (!v.empty())
    HIPCHK(c, hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return ESP_OK;
}

void drop_graph(esp_ctx *c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
  c->gkey.clear();
}

void free_replicas(esp_ctx *c) {
  dfree(c->d_sp); dfree(c->d_beta); dfree(c->d_ebase); dfree(c->d_erun); dfree(c->d_acc);
  dfree(c->d_partial); dfree(c->d_slot_rep); dfree(c->d_rep_slot); dfree(c->d_rej);
  dfree(c->d_nswap); dfree(c->d_nmeas); dfree(c->d_prop); dfree(c->d_unif); dfree(c->d_accd);
  c->R = 0;
  drop_graph(c);
}

void free_model(esp_ctx *c) {
  dfree(c->d_nbr); dfree(c->d_dev2ref); dfree(c->d_ref2dev); dfree(c->d_col_off);
  dfree(c->d_cJ); dfree(c->d_cB); dfree(c->d_cD); dfree(c->d_ctab); dfree(c->d_mag);
  dfree(c->d_si4); dfree(c->d_h4); dfree(c->d_basis);
  c->have_model = false;
  drop_graph(c);
}

KArgs base_args(esp_ctx *c) {
  KArgs a{};
  const esp::HostModel &m = c->hm;
  a.sp = c->d_sp;
  a.nbr = c->d_nbr;
  a.cJ = c->d_cJ; a.cB = c->d_cB; a.cD = c->d_cD; a.ctab = c->d_ctab;
  a.mag = c->d_mag; a.basis = c->d_basis; a.dev2ref = c->d_dev2ref; a.col_off = c->d_col_off;
  a.si4 = c->d_si4; a.h4 = c->d_h4;
  a.beta_slot = c->d_beta; a.rep_slot = c->d_rep_slot; a.rej = c->d_rej; a.clock = c->d_clock;
  a.N = m.N; a.nb = m.nb; a.zmax = m.zmax; a.R = c->R; a.ncol = m.ncol;
  a.has_si = m.has_si; a.has_h = m.has_h;
  return a;
}

// kernel table: [mode][cmp][hb][hd]
using PhaseFn = void (*)(KArgs);
using LdsFn = void (*)(KArgs, int, LdsMeas);
using EnergyFn = void (*)(KArgs, int, double *);

template <int MODE>
PhaseFn pick_phase(bool cmp, bool hb, bool hd) {
  if (cmp) {
    if (hb) return hd ? k_phase<MODE, true, true, true> : k_phase<MODE, true, true, false>;
    return hd ? k_phase<MODE, true, false, true> : k_phase<MODE, true, false, false>;
  }
  if (hb) return hd ? k_phase<MODE, false, true, true> : k_phase<MODE, false, true, false>;
  return hd ? k_phase<MODE, false, false, true> : k_phase<MODE, false, false, false>;
}
PhaseFn phase_fn(int mode, bool cmp, bool hb, bool hd) {
  if (mode == M_RANDOM) return pick_phase<M_RANDOM>(cmp, hb, hd);
  if (mode == M_CONST) return pick_phase<M_CONST>(cmp, hb, hd);
  return pick_phase<M_HOSTFED>(cmp, hb, hd);
}
template <int MODE>
LdsFn pick_lds(bool cmp, bool hb, bool hd) {
  if (cmp) {
    if (hb) return hd ? k_lds<MODE, true, true, true> : k_lds<MODE, true, true, false>;
    return hd ? k_lds<MODE, true, false, true> : k_lds<MODE, true, false, false>;
  }
  if (hb) return hd ? k_lds<MODE, false, true, true> : k_lds<MODE, false, true, false>;
  return hd ? k_lds<MODE, false, false, true> : k_lds<MODE, false, false, false>;
}
LdsFn lds_fn(int mode, bool cmp, bool hb, bool hd) {
  if (mode == M_RANDOM) return pick_lds<M_RANDOM>(cmp, hb, hd);
  return pick_lds<M_CONST>(cmp, hb, hd);
}
EnergyFn energy_fn(bool cmp, bool hb, bool hd) {
  if (cmp) {
    if (hb) return hd ? k_energy<true, true, true> : k_energy<true, true, false>;
    return hd ? k_energy<true, false, true> : k_energy<true, false, false>;
  }
  if (hb) return hd ? k_energy<false, true, true> : k_energy<false, true, false>;
  return hd ? k_energy<false, false, true> : k_energy<false, false, false>;
}

int check_ready(esp_ctx *c, bool need_replicas) {
  if (!c) return ESP_EINVAL;
  if (!c->have_model) return fail(c, ESP_ESTATE, "no model: call esp_set_model first");
  if (need_replicas && c->R <= 0) return fail(c, ESP_ESTATE, "no replicas: call esp_set_replicas first");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ESP_EHIP, "hipSetDevice failed");
  return ESP_OK;
}

// Enqueue `nsweeps` sweeps starting at clock->sweep + off0 (global-memory path).
int enqueue_phase_sweeps(esp_ctx *c, int mode, double tilt, uint64_t seed, int nsweeps, int off0) {
  const esp::HostModel &m = c->hm;
  KArgs a = base_args(c);
  a.tilt = tilt;
  a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32);
  a.stream_base = 0;
  PhaseFn fn = phase_fn(mode, m.compressed, m.has_B, m.has_D);
  const size_t shm = m.compressed ? c->ctab_bytes : 0;
  for (int s = 0; s < nsweeps; ++s)
    for (int col = 0; col < m.ncol; ++col) {
      a.c_begin = m.col_off[col];
      a.c_count = m.col_off[col + 1] - m.col_off[col];
      if (a.c_count == 0) continue;
      a.nblk = (a.c_count + kBlock - 1) / kBlock;
      a.xcd_remap = (c->R % 8 == 0) ? 1 : 0;
      a.sweep_off = off0 + s;
      hipLaunchKernelGGL(fn, dim3(a.nblk * c->R), dim3(kBlock), shm, c->stream, a);
    }
  HIPCHK(c, hipGetLastError());
  return ESP_OK;
}

int enqueue_lds_sweeps(esp_ctx *c, int mode, double tilt, uint64_t seed, int nsweeps,
                       int measure_every) {
  const esp::HostModel &m = c->hm;
  KArgs a = base_args(c);
  a.tilt = tilt;
  a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32);
  LdsMeas lm{c->d_acc, c->d_erun, c->d_ebase, c->d_nmeas, measure_every};
  LdsFn fn = lds_fn(mode, m.compressed, m.has_B, m.has_D);
  hipLaunchKernelGGL(fn, dim3(c->R), dim3(kLdsBlock), c->lds_bytes, c->stream, a, nsweeps, lm);
  HIPCHK(c, hipGetLastError());
  return ESP_OK;
}

int enqueue_reduce(esp_ctx *c) {
  hipLaunchKernelGGL(k_reduce, dim3(c->red_blocks, c->R), dim3(kBlock), 0, c->stream, c->d_sp,
                     c->hm.N, c->red_blocks, c->d_partial);
  HIPCHK(c, hipGetLastError());
  return ESP_OK;
}

int enqueue_finalize(esp_ctx *c, bool use_partials, bool measure, bool exchange, int advance,
                     uint64_t seed) {
  FinArgs f{};
  f.partial = use_partials ? c->d_partial : nullptr;
  f.nblk = c->red_blocks; f.N = c->hm.N; f.R = c->R;
  f.e_base = c->d_ebase; f.e_run = c->d_erun; f.acc = c->d_acc; f.nmeas = c->d_nmeas;
  f.slot_rep = c->d_slot_rep; f.rep_slot = c->d_rep_slot; f.beta_slot = c->d_beta;
  f.nswap = c->d_nswap; f.clock = c->d_clock;
  f.measure = measure; f.exchange = exchange; f.advance_sweeps = advance;
  f.key0 = (uint32_t)seed; f.key1 = (uint32_t)(seed >> 32); f.stream_base = 0;
  const int nt = std::min(1024, std::max(64, ((c->R + 63) / 64) * 64));
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(nt), 0, c->stream, f);
  HIPCHK(c, hipGetLastError());
  return ESP_OK;
}

int set_clock(esp_ctx *c, uint64_t sweep, uint64_t step, int parity) {
  Clock h{sweep, step, parity, 0};
  HIPCHK(c, hipMemcpyAsync(c->d_clock, &h, sizeof(Clock), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

// Direct energies into d_out[R].
int enqueue_energy(esp_ctx *c, double *d_out) {
  const esp::HostModel &m = c->hm;
  KArgs a = base_args(c);
  EnergyFn fn = energy_fn(m.compressed, m.has_B, m.has_D);
  const int nblk = c->red_blocks;
  hipLaunchKernelGGL(fn, dim3(nblk, c->R), dim3(kBlock), m.compressed ? c->ctab_bytes : 0,
                     c->stream, a, nblk, c->d_partial);
  hipLaunchKernelGGL(k_energy_fin, dim3((c->R + 255) / 256), dim3(256), 0, c->stream,
                     (const double *)c->d_partial, nblk, c->R, d_out);
  HIPCHK(c, hipGetLastError());
  return ESP_OK;
}

// e_base = direct energy, w = 0, e_run = e_base.
int rebase(esp_ctx *c) {
  int rc = enqueue_energy(c, c->d_ebase);
  if (rc) return rc;
  const size_t n = (size_t)c->R * c->hm.N;
  hipLaunchKernelGGL(k_zero_w, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->d_sp, n);
  HIPCHK(c, hipMemcpyAsync(c->d_erun, c->d_ebase, sizeof(double) * c->R, hipMemcpyDeviceToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

// One MC "round": steps x (M sweeps [+ measure]) [+ exchange]; advances the clock.
int enqueue_round(esp_ctx *c, int steps, int M, bool measure, bool exchange, int mode,
                  double tilt, uint64_t seed) {
  int rc;
  if (c->use_lds) {
    if (measure) {
      rc = enqueue_lds_sweeps(c, mode, tilt, seed, steps * M, M);
      if (rc) return rc;
      if (exchange) return enqueue_finalize(c, false, false, true, steps * M, seed);
      return enqueue_finalize(c, false, false, false, steps * M, seed);
    }
    rc = enqueue_lds_sweeps(c, mode, tilt, seed, steps * M, 0);
    if (rc) return rc;
    return enqueue_finalize(c, false, false, exchange, steps * M, seed);
  }
  if (!measure) {
    rc = enqueue_phase_sweeps(c, mode, tilt, seed, steps * M, 0);
    if (rc) return rc;
    if (exchange) {
      rc = enqueue_reduce(c);
      if (rc) return rc;
      return enqueue_finalize(c, true, false, true, steps * M, seed);
    }
    return enqueue_finalize(c, false, false, false, steps * M, seed);
  }
  for (int s = 0; s < steps; ++s) {
    rc = enqueue_phase_sweeps(c, mode, tilt, seed, M, s * M);
    if (rc) return rc;
    rc = enqueue_reduce(c);
    if (rc) return rc;
    const bool last = s + 1 == steps;
    rc = enqueue_finalize(c, true, true, exchange && last, last ? steps * M : 0, seed);
    if (rc) return rc;
  }
  return ESP_OK;
}

}  // namespace

// ===================================================================================
// C ABI
// ===================================================================================
extern "C" {

int esp_version(void) { return 100; }

const char *esp_last_error(const esp_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int esp_create(esp_ctx **out, int device) {
  if (!out) return ESP_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return ESP_EHIP;
  if (device < 0 || device >= n) return ESP_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return ESP_EHIP;
  esp_ctx *c = new esp_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return ESP_EHIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->lds_max = prop.sharedMemPerBlock;
  if (hipMalloc((void **)&c->d_clock, sizeof(Clock)) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return ESP_ENOMEM;
  }
  *out = c;
  return ESP_OK;
}

int esp_destroy(esp_ctx *c) {
  if (!c) return ESP_EINVAL;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_replicas(c);
  free_model(c);
  dfree(c->d_clock);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return ESP_OK;
}

int esp_build_table(int nb, const double *basis_frac, const double *basis_mag, const int *L,
                    int open_boundary, int nrows, const double *f1, const double *f2,
                    const double *value, const double *dvec, int ncomp, int zcap, int *cnt,
                    int *nbr, double *val, int *zmax) {
  if (nb <= 0 || !basis_frac || !basis_mag || !L || nrows < 0 || (ncomp != 1 && ncomp != 3) ||
      zcap < 0 || !cnt || !nbr || !val)
    return ESP_EINVAL;
  if (L[0] < 1 || L[1] < 1 || L[2] < 1) return ESP_EINVAL;
  if (nrows > 0 && (!f1 || !f2 || !value)) return ESP_EINVAL;
  if (ncomp == 3 && nrows > 0 && !dvec) return ESP_EINVAL;
  int rc = esp::build_table(nb, basis_frac, basis_mag, L, open_boundary, nrows, f1, f2, value, dvec,
                            ncomp, zcap, cnt, nbr, val, zmax);
  if (rc == -5) return ESP_ETABLE;
  if (rc < 0) return ESP_EINVAL;
  return ESP_OK;
}

int esp_set_model(esp_ctx *c, int N, int nb, const double *mag, int zJ, const int *j_cnt,
                  const int *j_nbr, const double *j_val, int zB, const int *b_cnt, const int *b_nbr,
                  const double *b_val, int zD, const int *d_cnt, const int *d_nbr,
                  const double *d_vec, int have_si, const double *si_par, const double *si_axis,
                  int have_field, const double *h_par, const double *h_axis) {
  if (!c) return ESP_EINVAL;
  HIPCHK(c, hipSetDevice(c->device));
  esp::HostModel hm;
  std::string e = esp::prepare_model(hm, N, nb, mag, zJ, j_cnt, j_nbr, j_val, zB, b_cnt, b_nbr,
                                     b_val, zD, d_cnt, d_nbr, d_vec, have_si, si_par, si_axis,
                                     have_field, h_par, h_axis, c->Lhint[0] > 0 ? c->Lhint : nullptr);
  if (!e.empty()) return fail(c, ESP_EINVAL, e);
  // Per-basis coefficient tables live in LDS; keep them small.
  const size_t ctab_bytes = (size_t)hm.nb * hm.zmax * 5 * sizeof(double);
  if (hm.compressed && ctab_bytes > 32768) {
    // too large for LDS staging: fall back to per-bond coefficients
    hm.compressed = false;
    hm.coef.assign((size_t)hm.zmax * N * esp::kCoef, 0.0);
    for (int d = 0; d < N; ++d) {
      const int a = hm.basis[d];
      for (int l = 0; l < hm.zmax; ++l)
        for (int k = 0; k < esp::kCoef; ++k)
          hm.coef[((size_t)l * N + d) * esp::kCoef + k] = hm.ctab[((size_t)a * hm.zmax + l) * esp::kCoef + k];
    }
    hm.ctab.clear();
  }
  free_replicas(c);
  free_model(c);
  c->hm = std::move(hm);
  const esp::HostModel &m = c->hm;
  int rc;
  if ((rc = upload(c, c->d_nbr, m.nbr))) return rc;
  if ((rc = upload(c, c->d_dev2ref, m.dev2ref))) return rc;
  if ((rc = upload(c, c->d_ref2dev, m.ref2dev))) return rc;
  if ((rc = upload(c, c->d_col_off, m.col_off))) return rc;
  if ((rc = upload(c, c->d_mag, m.mag))) return rc;
  if ((rc = upload(c, c->d_basis, m.basis))) return rc;
  if ((rc = upload(c, c->d_si4, m.si4))) return rc;
  if ((rc = upload(c, c->d_h4, m.h4))) return rc;
  if (m.compressed) {
    if ((rc = upload(c, c->d_ctab, m.ctab))) return rc;
  } else {
    std::vector<double> cj((size_t)m.zmax * N), cb, cd;
    if (m.has_B) cb.resize((size_t)m.zmax * N);
    if (m.has_D) cd.resize((size_t)m.zmax * N * 3);
    for (size_t t = 0; t < (size_t)m.zmax * N; ++t) {
      cj[t] = m.coef[t * esp::kCoef];
      if (m.has_B) cb[t] = m.coef[t * esp::kCoef + 1];
      if (m.has_D)
        for (int k = 0; k < 3; ++k) cd[3 * t + k] = m.coef[t * esp::kCoef + 2 + k];
    }
    if ((rc = upload(c, c->d_cJ, cj))) return rc;
    if (m.has_B && (rc = upload(c, c->d_cB, cb))) return rc;
    if (m.has_D && (rc = upload(c, c->d_cD, cd))) return rc;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ctab_bytes = m.compressed ? ctab_bytes : 0;
  c->lds_bytes = (size_t)N * sizeof(double4) + c->ctab_bytes;
  c->lds_fits = c->lds_bytes + 4096 <= c->lds_max;
  c->use_lds = c->lds_fits && c->path != 1;
  if (c->lds_fits && c->lds_bytes > 65536) {
    // opt in to the large LDS allocation for every LDS kernel instance
    for (int mode = 0; mode < 2; ++mode) {
      LdsFn fn = lds_fn(mode, m.compressed, m.has_B, m.has_D);
      if (hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)c->lds_bytes) != hipSuccess) {
        c->use_lds = false;
        c->lds_fits = false;
        (void)hipGetLastError();
      }
    }
  }
  c->have_model = true;
  return ESP_OK;
}

int esp_plan_model(int N, int nb, const double *mag, int zJ, const int *j_cnt, const int *j_nbr,
                   const double *j_val, int zB, const int *b_cnt, const int *b_nbr,
                   const double *b_val, int zD, const int *d_cnt, const int *d_nbr,
                   const double *d_vec, const int *L, int *ncol, int *colour, int *info) {
  if (!ncol || !colour) return ESP_EINVAL;
  esp::HostModel hm;
  std::string e = esp::prepare_model(hm, N, nb, mag, zJ, j_cnt, j_nbr, j_val, zB, b_cnt, b_nbr,
                                     b_val, zD, d_cnt, d_nbr, d_vec, 0, nullptr, nullptr, 0,
                                     nullptr, nullptr, L);
  if (!e.empty()) return ESP_EINVAL;
  *ncol = hm.ncol;
  std::memcpy(colour, hm.colour.data(), sizeof(int) * N);
  if (info) {
    info[0] = hm.N; info[1] = hm.zmax; info[2] = hm.ncol; info[3] = hm.compressed ? 1 : 0;
    info[4] = 0;
  }
  return ESP_OK;
}

int esp_set_path(esp_ctx *c, int path) {
  if (!c || path < 0 || path > 2) return ESP_EINVAL;
  c->path = path;
  if (c->have_model) {
    if (path == 2 && !c->lds_fits) return fail(c, ESP_EINVAL, "replica does not fit in LDS");
    c->use_lds = c->lds_fits && path != 1;
    drop_graph(c);
  }
  return ESP_OK;
}

int esp_set_supercell(esp_ctx *c, const int *L) {
  if (!c) return ESP_EINVAL;
  if (!L) { c->Lhint[0] = c->Lhint[1] = c->Lhint[2] = 0; return ESP_OK; }
  if (L[0] < 1 || L[1] < 1 || L[2] < 1) return fail(c, ESP_EINVAL, "supercell sizes must be >= 1");
  for (int k = 0; k < 3; ++k) c->Lhint[k] = L[k];
  return ESP_OK;
}

int esp_get_coloring(esp_ctx *c, int *ncol, int *colour) {
  if (!c || !ncol || !colour) return ESP_EINVAL;
  if (!c->have_model) return fail(c, ESP_ESTATE, "no model");
  *ncol = c->hm.ncol;
  std::memcpy(colour, c->hm.colour.data(), sizeof(int) * c->hm.N);
  return ESP_OK;
}

int esp_get_model_info(esp_ctx *c, int *out) {
  if (!c || !out) return ESP_EINVAL;
  if (!c->have_model) return fail(c, ESP_ESTATE, "no model");
  out[0] = c->hm.N; out[1] = c->hm.zmax; out[2] = c->hm.ncol;
  out[3] = c->hm.compressed ? 1 : 0; out[4] = c->use_lds ? 1 : 0;
  return ESP_OK;
}

int esp_set_replicas(esp_ctx *c, int R, const double *tems, const double *spins) {
  int rc = check_ready(c, false);
  if (rc) return rc;
  if (R <= 0 || R > 65536 || !tems || !spins) return fail(c, ESP_EINVAL, "bad R / tems / spins");
  for (int t = 0; t < R; ++t)
    if (!(tems[t] > 0.0)) return fail(c, ESP_EINVAL, "temperatures must be > 0");
  free_replicas(c);
  const int N = c->hm.N;
  c->R = R;
  c->tems.assign(tems, tems + R);
  c->red_blocks = std::max(1, std::min(kRedBlocks, (N + 4095) / 4096));
  if ((rc = dalloc(c, c->d_sp, (size_t)R * N))) return rc;
  if ((rc = dalloc(c, c->d_beta, R))) return rc;
  if ((rc = dalloc(c, c->d_ebase, R))) return rc;
  if ((rc = dalloc(c, c->d_erun, R))) return rc;
  if ((rc = dalloc(c, c->d_acc, (size_t)6 * R))) return rc;
  if ((rc = dalloc(c, c->d_partial, (size_t)4 * R * c->red_blocks))) return rc;
  if ((rc = dalloc(c, c->d_slot_rep, R))) return rc;
  if ((rc = dalloc(c, c->d_rep_slot, R))) return rc;
  if ((rc = dalloc(c, c->d_rej, R))) return rc;
  if ((rc = dalloc(c, c->d_nswap, (size_t)2 * R))) return rc;
  if ((rc = dalloc(c, c->d_nmeas, 1))) return rc;
  std::vector<double> beta(R);
  for (int t = 0; t < R; ++t) beta[t] = 1.0 / tems[t];  // 1.0_dp/tems(i) (pt.F90:1257)
  std::vector<int> id(R);
  for (int t = 0; t < R; ++t) id[t] = t;
  HIPCHK(c, hipMemcpyAsync(c->d_beta, beta.data(), sizeof(double) * R, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_slot_rep, id.data(), sizeof(int) * R, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_rep_slot, id.data(), sizeof(int) * R, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_acc, 0, sizeof(double) * 6 * R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_rej, 0, sizeof(unsigned long long) * R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_nswap, 0, sizeof(unsigned long long) * 2 * R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_nmeas, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return esp_put_spins(c, spins);
}

int esp_put_spins(esp_ctx *c, const double *spins) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (!spins) return fail(c, ESP_EINVAL, "spins is NULL");
  const int N = c->hm.N, R = c->R;
  double *d_in = nullptr;
  if ((rc = dalloc(c, d_in, (size_t)3 * N * R))) return rc;
  HIPCHK(c, hipMemcpyAsync(d_in, spins, sizeof(double) * 3 * N * R, hipMemcpyHostToDevice, c->stream));
  const size_t n = (size_t)N * R;
  hipLaunchKernelGGL(k_put_spins, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     (const double *)d_in, (const int *)c->d_ref2dev, N, R, c->d_sp);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(d_in);
  return rebase(c);
}

int esp_get_spins(esp_ctx *c, int by_slot, double *spins) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (!spins) return fail(c, ESP_EINVAL, "spins is NULL");
  const int N = c->hm.N, R = c->R;
  double *d_out = nullptr;
  if ((rc = dalloc(c, d_out, (size_t)3 * N * R))) return rc;
  const size_t n = (size_t)N * R;
  hipLaunchKernelGGL(k_get_spins, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                     (const double4 *)c->d_sp, (const int *)c->d_ref2dev,
                     by_slot ? (const int *)c->d_slot_rep : (const int *)nullptr, N, R, d_out);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(spins, d_out, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(d_out);
  return ESP_OK;
}

int esp_energy(esp_ctx *c, double *E) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (!E) return fail(c, ESP_EINVAL, "E is NULL");
  double *d_e = nullptr;
  if ((rc = dalloc(c, d_e, c->R))) return rc;
  if ((rc = enqueue_energy(c, d_e))) return rc;
  HIPCHK(c, hipMemcpyAsync(E, d_e, sizeof(double) * c->R, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(d_e);
  return ESP_OK;
}

int esp_sweep(esp_ctx *c, int nsweeps, int mode, double tilt, uint64_t seed, uint64_t sweep0) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (nsweeps < 0 || (mode != ESP_MODE_RANDOM && mode != ESP_MODE_CONST))
    return fail(c, ESP_EINVAL, "bad nsweeps / mode");
  if (nsweeps == 0) return ESP_OK;
  if ((rc = set_clock(c, sweep0, 0, 1))) return rc;
  if (c->use_lds) rc = enqueue_lds_sweeps(c, mode, tilt, seed, nsweeps, 0);
  else rc = enqueue_phase_sweeps(c, mode, tilt, seed, nsweeps, 0);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_sweep_hostfed(esp_ctx *c, const double *prop, const double *unif, uint8_t *accepted) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (!prop || !unif || !accepted) return fail(c, ESP_EINVAL, "NULL host-fed array");
  const esp::HostModel &m = c->hm;
  const size_t n = (size_t)m.N * c->R;
  if (!c->d_prop) {
    if ((rc = dalloc(c, c->d_prop, 3 * n))) return rc;
    if ((rc = dalloc(c, c->d_unif, n))) return rc;
    if ((rc = dalloc(c, c->d_accd, n))) return rc;
  }
  HIPCHK(c, hipMemcpyAsync(c->d_prop, prop, sizeof(double) * 3 * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_unif, unif, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  KArgs a = base_args(c);
  a.prop = c->d_prop; a.unif = c->d_unif; a.accepted = c->d_accd;
  PhaseFn fn = phase_fn(M_HOSTFED, m.compressed, m.has_B, m.has_D);
  if ((rc = set_clock(c, 0, 0, 1))) return rc;
  for (int col = 0; col < m.ncol; ++col) {
    a.c_begin = m.col_off[col];
    a.c_count = m.col_off[col + 1] - m.col_off[col];
    if (!a.c_count) continue;
    a.nblk = (a.c_count + kBlock - 1) / kBlock;
    a.xcd_remap = (c->R % 8 == 0);
    hipLaunchKernelGGL(fn, dim3(a.nblk * c->R), dim3(kBlock), m.compressed ? c->ctab_bytes : 0,
                       c->stream, a);
  }
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(accepted, c->d_accd, n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_measure(esp_ctx *c) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if ((rc = enqueue_reduce(c))) return rc;
  if ((rc = enqueue_finalize(c, true, true, false, 0, 0))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_pt_exchange(esp_ctx *c, int start, uint64_t seed, uint64_t step) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (start != 1 && start != 2) return fail(c, ESP_EINVAL, "start must be 1 or 2");
  if ((rc = set_clock(c, 0, step, start))) return rc;
  if ((rc = enqueue_reduce(c))) return rc;
  if ((rc = enqueue_finalize(c, true, false, true, 0, seed))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_run(esp_ctx *c, int n_rounds, int steps_per_round, int steps_measure, int do_measure,
            int do_exchange, int start_parity, int mode, double tilt, uint64_t seed,
            uint64_t sweep0, uint64_t step0, int use_graph) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (n_rounds < 0 || steps_per_round < 1 || steps_measure < 1 ||
      (start_parity != 1 && start_parity != 2) || (mode != ESP_MODE_RANDOM && mode != ESP_MODE_CONST))
    return fail(c, ESP_EINVAL, "bad esp_run arguments");
  if (n_rounds == 0) return ESP_OK;
  if ((rc = set_clock(c, sweep0, step0, start_parity))) return rc;
  if (!use_graph) {
    for (int k = 0; k < n_rounds; ++k) {
      rc = enqueue_round(c, steps_per_round, steps_measure, do_measure, do_exchange, mode, tilt, seed);
      if (rc) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ESP_OK;
  }
  char key[256];
  std::snprintf(key, sizeof key, "%d/%d/%d/%d/%d/%.17g/%llu", steps_per_round, steps_measure,
                do_measure, do_exchange, mode, tilt, (unsigned long long)seed);
  if (c->gkey != key) {
    drop_graph(c);
    hipGraph_t g = nullptr;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_round(c, steps_per_round, steps_measure, do_measure, do_exchange, mode, tilt, seed);
    hipError_t ce = hipStreamEndCapture(c->stream, &g);
    if (rc) return rc;
    if (ce != hipSuccess) return fail(c, ESP_EHIP, std::string("capture: ") + hipGetErrorString(ce));
    hipError_t ie = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      c->gexec = nullptr;
      return fail(c, ESP_EHIP, std::string("instantiate: ") + hipGetErrorString(ie));
    }
    c->gkey = key;
  }
  for (int k = 0; k < n_rounds; ++k) HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_time_sweeps(esp_ctx *c, int nsweeps, int mode, double tilt, uint64_t seed, uint64_t sweep0,
                    double *kernel_ms, int *nlaunch, double *wall_ms, int64_t *updates_per_launch) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (nsweeps < 1 || (mode != ESP_MODE_RANDOM && mode != ESP_MODE_CONST) || !kernel_ms || !nlaunch)
    return fail(c, ESP_EINVAL, "bad esp_time_sweeps arguments");
  if ((rc = set_clock(c, sweep0, 0, 1))) return rc;
  const esp::HostModel &m = c->hm;
  // one event pair around every sweep-kernel launch, on the context's stream
  const int per = c->use_lds ? 1 : nsweeps * m.ncol;
  std::vector<hipEvent_t> ev(2 * per, nullptr);
  for (auto &e : ev) HIPCHK(c, hipEventCreate(&e));
  int k = 0;
  if (c->use_lds) {
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    rc = enqueue_lds_sweeps(c, mode, tilt, seed, nsweeps, 0);
    HIPCHK(c, hipEventRecord(ev[1], c->stream));
    k = 1;
  } else {
    KArgs a = base_args(c);
    a.tilt = tilt;
    a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32);
    PhaseFn fn = phase_fn(mode, m.compressed, m.has_B, m.has_D);
    for (int sw = 0; sw < nsweeps; ++sw)
      for (int col = 0; col < m.ncol; ++col) {
        a.c_begin = m.col_off[col];
        a.c_count = m.col_off[col + 1] - m.col_off[col];
        if (a.c_count == 0) continue;
        a.nblk = (a.c_count + kBlock - 1) / kBlock;
        a.xcd_remap = (c->R % 8 == 0) ? 1 : 0;
        a.sweep_off = sw;
        HIPCHK(c, hipEventRecord(ev[2 * k], c->stream));
        hipLaunchKernelGGL(fn, dim3(a.nblk * c->R), dim3(kBlock),
                           m.compressed ? c->ctab_bytes : 0, c->stream, a);
        HIPCHK(c, hipEventRecord(ev[2 * k + 1], c->stream));
        ++k;
      }
    HIPCHK(c, hipGetLastError());
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double sum = 0.0;
  for (int q = 0; q < k; ++q) {
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, ev[2 * q], ev[2 * q + 1]));
    sum += ms;
  }
  float wall = 0.f;
  HIPCHK(c, hipEventElapsedTime(&wall, ev[0], ev[2 * k - 1]));
  for (auto &e : ev) (void)hipEventDestroy(e);
  *kernel_ms = sum;
  *nlaunch = k;
  if (wall_ms) *wall_ms = wall;
  if (updates_per_launch)
    *updates_per_launch = c->use_lds ? (int64_t)m.N * c->R * nsweeps : (int64_t)(m.N / m.ncol) * c->R;
  return rc;
}

int esp_get_stats(esp_ctx *c, double *acc, int64_t *rejected, double *e_run, int64_t *nmeas) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  const int R = c->R;
  if (e_run) {
    if ((rc = enqueue_reduce(c))) return rc;
    if ((rc = enqueue_finalize(c, true, false, false, 0, 0))) return rc;
    HIPCHK(c, hipMemcpyAsync(e_run, c->d_erun, sizeof(double) * R, hipMemcpyDeviceToHost, c->stream));
  }
  if (acc) HIPCHK(c, hipMemcpyAsync(acc, c->d_acc, sizeof(double) * 6 * R, hipMemcpyDeviceToHost, c->stream));
  if (rejected)
    HIPCHK(c, hipMemcpyAsync(rejected, c->d_rej, sizeof(int64_t) * R, hipMemcpyDeviceToHost, c->stream));
  if (nmeas) HIPCHK(c, hipMemcpyAsync(nmeas, c->d_nmeas, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_get_map(esp_ctx *c, int *slot_rep, int64_t *nswap) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  const int R = c->R;
  std::vector<int> sr(R);
  HIPCHK(c, hipMemcpyAsync(sr.data(), c->d_slot_rep, sizeof(int) * R, hipMemcpyDeviceToHost, c->stream));
  if (nswap)
    HIPCHK(c, hipMemcpyAsync(nswap, c->d_nswap, sizeof(int64_t) * 2 * R, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (slot_rep)
    for (int t = 0; t < R; ++t) slot_rep[t] = sr[t] + 1;
  return ESP_OK;
}

int esp_reset_stats(esp_ctx *c) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  const int R = c->R;
  HIPCHK(c, hipMemsetAsync(c->d_acc, 0, sizeof(double) * 6 * R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_rej, 0, sizeof(unsigned long long) * R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_nswap, 0, sizeof(unsigned long long) * 2 * R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_nmeas, 0, sizeof(unsigned long long), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_sync(esp_ctx *c) {
  if (!c) return ESP_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

void *esp_stream(esp_ctx *c) { return c ? (void *)c->stream : nullptr; }

}  // extern "C"

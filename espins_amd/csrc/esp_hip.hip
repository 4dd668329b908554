// esp_hip.hip -- MI355X (gfx950) engine for the ESpinS Metropolis sweep and
// parallel-tempering exchange, behind the C ABI of include/espins_hip.h.
//
// Kernels live in esp_kernels.h; this file owns the context, memory and launch logic.
//   sites   colour-major: colour c owns device indices [col_off[c], col_off[c+1])
//   PT      slot_rep[R] / rep_slot[R]: replica <-> temperature slot labels; a swap
//           (pt.F90:1275-1285) exchanges labels, never spins
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/espins_hip.h"
#include "esp_kernels.h"
#include "esp_model.h"

using namespace espd;
using namespace espk;

namespace {
constexpr int kRedBlocks = 64;    // reduce blocks per replica (upper bound)

// Scratch of the allocation-free entry point (esp_energy_noalloc): the block partials of the
// direct energy.  A file-scope device array: every GPU the library runs on holds its own
// copy until the process ends.  Host code never uses the array's name as an address; it
// uses the per-device address looked up once by hipGetSymbolAddress (scratch_address).
constexpr size_t kNoallocScratch = size_t(1) << 18;   // doubles: 2 MiB per GPU
static __device__ __align__(256) double scratch[kNoallocScratch];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on `device` (which must be current): looked up on the first
// call for that device, then cached.  Returns null if the lookup fails.
double *scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch_dev[device] = static_cast<double *>(p);
  }
  return g_scratch_dev[device];
}
}  // namespace

#include "esp_ctx.h"

namespace {

// Lanes per site: the caller's choice, else one lane per site on the global-memory path
// (measured fastest there for z = 6..42: more sites in flight beat shorter bond loops) and,
// in LDS, more lanes for wide bond lists, where a block has few sites per colour.
int auto_tps(const esp_ctx *c) {
  if (c->tps_user) return c->tps_user;
  if (!c->use_lds) return 1;
  const int z = c->hm.zmax;
  return z <= 8 ? 1 : (z <= 20 ? 2 : 4);
}

// kernel tables: [mode][layout][hb][hd][self]
using PhaseFn = void (*)(KArgs);
using LdsFn = void (*)(KArgs, int, LdsMeas);
using EnergyFn = void (*)(KArgs, int, double *);

template <int MODE, int LAY, bool HB, bool HD, int CH>
PhaseFn ph4(bool self) { return self ? k_phase<MODE, LAY, HB, HD, true, CH> : k_phase<MODE, LAY, HB, HD, false, CH>; }
template <int MODE, int LAY, bool HB, int CH>
PhaseFn ph3(bool hd, bool self) { return hd ? ph4<MODE, LAY, HB, true, CH>(self) : ph4<MODE, LAY, HB, false, CH>(self); }
template <int MODE, int LAY, int CH>
PhaseFn ph2(bool hb, bool hd, bool self) { return hb ? ph3<MODE, LAY, true, CH>(hd, self) : ph3<MODE, LAY, false, CH>(hd, self); }
template <int MODE, int CH>
PhaseFn ph1(int lay, bool hb, bool hd, bool self) {
  if (lay == LAY_IMP) return ph2<MODE, LAY_IMP, CH>(hb, hd, self);
  return lay == LAY_CMP ? ph2<MODE, LAY_CMP, CH>(hb, hd, self) : ph2<MODE, LAY_BOND, CH>(hb, hd, self);
}
template <int CH>
PhaseFn phase_fn_ch(int mode, int lay, const esp::HostModel &m) {
  if (mode == M_RANDOM) return ph1<M_RANDOM, CH>(lay, m.has_B, m.has_D, m.has_self);
  if (mode == M_CONST) return ph1<M_CONST, CH>(lay, m.has_B, m.has_D, m.has_self);
  return ph1<M_HOSTFED, CH>(lay, m.has_B, m.has_D, m.has_self);
}
// Lanes of one colour phase of every replica (mean over colours).
int64_t launch_lanes(const esp_ctx *c) {
  return (int64_t)c->hm.N * c->tps / std::max(1, c->hm.ncol) * c->R;
}

// Coefficient layout of the colour phases of this context.
int phase_layout(const esp_ctx *c) {
  const esp::HostModel &m = c->hm;
  if (m.implicit && m.compressed && c->indexing == 0) return LAY_IMP;
  return m.compressed ? LAY_CMP : LAY_BOND;
}
size_t phase_shm(const esp_ctx *c) {
  return phase_lds_bytes(phase_layout(c), c->hm.nb, c->hm.zmax, c->hm.rows_max);
}

// Wide chunks (more gathers in flight per lane, more registers) when one colour phase of all
// replicas has too few waves to hide memory latency by occupancy: < 4 waves per SIMD.
PhaseFn phase_fn(int mode, const esp_ctx *c, int64_t lanes_per_launch) {
  const bool wide = lanes_per_launch < 4 * 1024 * 64;
  const int lay = phase_layout(c);
  return wide ? phase_fn_ch<kChunkWide>(mode, lay, c->hm) : phase_fn_ch<kChunk>(mode, lay, c->hm);
}
template <int MODE, bool CMP, bool HB, bool HD>
LdsFn ld4(bool self) { return self ? k_lds<MODE, CMP, HB, HD, true> : k_lds<MODE, CMP, HB, HD, false>; }
template <int MODE, bool CMP, bool HB>
LdsFn ld3(bool hd, bool self) { return hd ? ld4<MODE, CMP, HB, true>(self) : ld4<MODE, CMP, HB, false>(self); }
template <int MODE, bool CMP>
LdsFn ld2(bool hb, bool hd, bool self) { return hb ? ld3<MODE, CMP, true>(hd, self) : ld3<MODE, CMP, false>(hd, self); }
template <int MODE>
LdsFn ld1(bool cmp, bool hb, bool hd, bool self) { return cmp ? ld2<MODE, true>(hb, hd, self) : ld2<MODE, false>(hb, hd, self); }
LdsFn lds_fn(int mode, const esp::HostModel &m) {
  if (mode == M_RANDOM) return ld1<M_RANDOM>(m.compressed, m.has_B, m.has_D, m.has_self);
  return ld1<M_CONST>(m.compressed, m.has_B, m.has_D, m.has_self);
}
EnergyFn energy_fn(bool cmp, bool hb, bool hd) {
  if (cmp) {
    if (hb) return hd ? k_energy<true, true, true> : k_energy<true, true, false>;
    return hd ? k_energy<true, false, true> : k_energy<true, false, false>;
  }
  if (hb) return hd ? k_energy<false, true, true> : k_energy<false, true, false>;
  return hd ? k_energy<false, false, true> : k_energy<false, false, false>;
}

int ensure_mpart(esp_ctx *c);

int check_ready(esp_ctx *c, bool need_replicas) {
  if (!c) return ESP_EINVAL;
  if (!c->have_model) return fail(c, ESP_ESTATE, "no model: call esp_set_model first");
  if (need_replicas && c->R <= 0) return fail(c, ESP_ESTATE, "no replicas: call esp_set_replicas first");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ESP_EHIP, "hipSetDevice failed");
  return need_replicas ? ensure_mpart(c) : ESP_OK;
}

// Per-wave sums of the colour phases -- spins (mpart) and accepted dE (d_w) -- sized for the
// current lanes per site and zeroed on (re)allocation: entries of waves a colour does not
// have stay zero.  Before the accumulators are re-laid out their sums move into e_base.
// Called before any capture, never inside one.
int ensure_mpart(esp_ctx *c) {
  const esp::HostModel &m = c->hm;
  int need = 1;
  for (int col = 0; col < m.ncol; ++col)
    need = std::max(need, ((m.col_off[col + 1] - m.col_off[col]) * c->tps + 63) / 64);
  if (c->d_mpart && c->d_w && need == c->mpart_nw) return ESP_OK;
  drop_graph(c);   // a captured round may hold the old buffers
  if (c->d_w && c->mpart_nw > 0) {
    hipLaunchKernelGGL(k_fold_w, dim3((c->R + 255) / 256), dim3(256), 0, c->stream, c->d_w,
                       m.ncol * c->mpart_nw, c->R, c->d_ebase);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  const size_t n = (size_t)c->R * m.ncol * need;
  int rc = dalloc(c, c->d_mpart, 3 * n);
  if (rc) return rc;
  if ((rc = dalloc(c, c->d_w, n))) return rc;
  HIPCHK(c, hipMemsetAsync(c->d_mpart, 0, 3 * n * sizeof(double), c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_w, 0, n * sizeof(double), c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->mpart_nw = need;
  c->mpart_valid = false;
  return ESP_OK;
}

// Kernel arguments of the colour phases, with the per-wave spin sums attached.
KArgs phase_args(esp_ctx *c, double tilt, uint64_t seed) {
  KArgs a = base_args(c);
  a.tilt = tilt;
  a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32);
  a.xcd_remap = (c->R % 8 == 0) ? 1 : 0;
  a.mpart = c->d_mpart;
  a.mpart_nw = c->mpart_nw;
  return a;
}

// Colour `col` of `a` (block count for the lanes per site).
void set_colour(KArgs &a, const esp::HostModel &m, int col) {
  a.col = col;
  a.c_begin = m.col_off[col];
  a.c_count = m.col_off[col + 1] - m.col_off[col];
  a.nblk = (a.c_count * a.tps + kBlock - 1) / kBlock;
}

// Enqueue `nsweeps` sweeps starting at clock->sweep + off0 (global-memory path).
int enqueue_phase_sweeps(esp_ctx *c, int mode, double tilt, uint64_t seed, int nsweeps, int off0) {
  if (mode == ESP_MODE_LEARNED) {
    c->mpart_valid = false;
    return esp::learn_enqueue_sweeps(c, seed, nsweeps, off0);
  }
  const esp::HostModel &m = c->hm;
  KArgs a = phase_args(c, tilt, seed);
  PhaseFn fn = phase_fn(mode, c, launch_lanes(c));
  const size_t shm = phase_shm(c);
  for (int s = 0; s < nsweeps; ++s)
    for (int col = 0; col < m.ncol; ++col) {
      set_colour(a, m, col);
      if (a.c_count == 0) continue;
      a.sweep_off = off0 + s;
      hipLaunchKernelGGL(fn, dim3(a.nblk * c->R), dim3(kBlock), shm, c->stream, a);
    }
  HIPCHK(c, hipGetLastError());
  if (nsweeps > 0) c->mpart_valid = a.mpart != nullptr;
  return ESP_OK;
}

int enqueue_lds_sweeps(esp_ctx *c, int mode, double tilt, uint64_t seed, int nsweeps,
                       int measure_every) {
  const esp::HostModel &m = c->hm;
  KArgs a = base_args(c);
  a.tilt = tilt;
  a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32);
  LdsMeas lm{c->d_acc, c->d_erun, c->d_ebase, c->d_nmeas, measure_every};
  LdsFn fn = lds_fn(mode, m);
  hipLaunchKernelGGL(fn, dim3(c->R), dim3(kLdsBlock), c->lds_bytes, c->stream, a, nsweeps, lm);
  HIPCHK(c, hipGetLastError());
  c->mpart_valid = false;
  return ESP_OK;
}

// Running-energy and magnetisation sums: the spins' sums come from the colour phases'
// per-wave sums when the last sweep left them, else from the spins.
int enqueue_reduce(esp_ctx *c) {
  const bool mp = c->mpart_valid;
  hipLaunchKernelGGL(k_reduce, dim3(c->red_blocks, c->R), dim3(kBlock), 0, c->stream,
                     (const double *)c->d_sp, (const double *)c->d_w, c->hm.N, c->Ns,
                     c->red_blocks, c->d_partial, mp ? (const double *)c->d_mpart : nullptr,
                     c->hm.ncol * c->mpart_nw);   // ncol * mpart_nw entries of both
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
  f.key0 = (uint32_t)seed; f.key1 = (uint32_t)(seed >> 32); f.stream_base = c->stream_base;
  f.trace = c->d_trace; f.trace_cap = c->trace_cap;
  const int nt = std::min(1024, std::max(64, ((c->R + 63) / 64) * 64));
  hipLaunchKernelGGL(k_finalize, dim3(1), dim3(nt), 0, c->stream, f);
  HIPCHK(c, hipGetLastError());
  return ESP_OK;
}

// enqueue_reduce + enqueue_finalize(use_partials = true) in one launch (k_reduce_fin): with the
// colour phases' per-wave sums, one block per 4096 entries of a replica, else the spin-sum blocks.
int enqueue_reduce_finalize(esp_ctx *c, bool measure, bool exchange, int advance, uint64_t seed) {
  const bool mp = c->mpart_valid;
  const int ne = c->hm.ncol * c->mpart_nw;
  const int nblk = mp ? std::max(1, std::min(c->red_blocks, (ne + 4095) / 4096)) : c->red_blocks;
  FinArgs f{};
  f.partial = c->d_partial;
  f.nblk = nblk; f.N = c->hm.N; f.R = c->R;
  f.e_base = c->d_ebase; f.e_run = c->d_erun; f.acc = c->d_acc; f.nmeas = c->d_nmeas;
  f.slot_rep = c->d_slot_rep; f.rep_slot = c->d_rep_slot; f.beta_slot = c->d_beta;
  f.nswap = c->d_nswap; f.clock = c->d_clock;
  f.measure = measure; f.exchange = exchange; f.advance_sweeps = advance;
  f.key0 = (uint32_t)seed; f.key1 = (uint32_t)(seed >> 32); f.stream_base = c->stream_base;
  f.trace = c->d_trace; f.trace_cap = c->trace_cap;
  hipLaunchKernelGGL(k_reduce_fin, dim3(nblk, c->R), dim3(kBlock), 0, c->stream, (const double *)c->d_sp,
                     (const double *)c->d_w, c->hm.N, c->Ns, mp ? (const double *)c->d_mpart : nullptr, ne,
                     c->d_partial, f, c->d_done);
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
  hipLaunchKernelGGL(fn, dim3(nblk, c->R), dim3(kBlock),
                     (m.compressed ? c->ctab_bytes : 0) + (kBlock / 64) * sizeof(double), c->stream, a,
                     nblk, c->d_partial);
  hipLaunchKernelGGL(k_energy_fin, dim3((c->R + 255) / 256), dim3(256), 0, c->stream,
                     (const double *)c->d_partial, nblk, c->R, d_out);
  HIPCHK(c, hipGetLastError());
  return ESP_OK;
}

// e_base = direct energy, accepted-dE accumulators = 0, e_run = e_base.
int rebase(esp_ctx *c) {
  int rc = ensure_mpart(c);
  if (rc) return rc;
  if ((rc = enqueue_energy(c, c->d_ebase))) return rc;
  const size_t n = (size_t)c->R * c->hm.ncol * c->mpart_nw;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->d_w, n, 0.0);
  c->mpart_valid = false;
  HIPCHK(c, hipMemcpyAsync(c->d_erun, c->d_ebase, sizeof(double) * c->R, hipMemcpyDeviceToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

// The LDS-resident kernel serves the plain proposals when the replica fits and no trace
// is recorded; learned proposals and traced runs use the global-memory path.
bool lds_sweeps(const esp_ctx *c, int mode) {
  return c->use_lds && mode != ESP_MODE_LEARNED && !c->d_trace && !c->obs.flags;
}

// One MC "round": steps x (M sweeps [+ measure]) [+ exchange]; advances the clock.
int enqueue_round(esp_ctx *c, int steps, int M, bool measure, bool exchange, int mode,
                  double tilt, uint64_t seed) {
  int rc;
  if (lds_sweeps(c, mode)) {
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
    if (exchange) return enqueue_reduce_finalize(c, false, true, steps * M, seed);
    return enqueue_finalize(c, false, false, false, steps * M, seed);
  }
  for (int s = 0; s < steps; ++s) {
    rc = enqueue_phase_sweeps(c, mode, tilt, seed, M, s * M);
    if (rc) return rc;
    if ((rc = esp::obs_enqueue(c))) return rc;   // reads spins and slot labels only
    const bool last = s + 1 == steps;
    rc = enqueue_reduce_finalize(c, true, exchange && last, last ? steps * M : 0, seed);
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
  if (hipMalloc((void **)&c->d_clock, sizeof(Clock)) != hipSuccess ||
      hipMalloc((void **)&c->d_done, sizeof(unsigned)) != hipSuccess ||
      hipMemset(c->d_done, 0, sizeof(unsigned)) != hipSuccess) {
    dfree(c->d_clock);
    dfree(c->d_done);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return ESP_ENOMEM;
  }
  // set-up of the allocation-free entry point on this device (its scratch address); a
  // failure here only disables esp_energy_noalloc, which then reports it
  (void)scratch_address(device);
  *out = c;
  return ESP_OK;
}

int esp_destroy(esp_ctx *c) {
  if (!c) return ESP_EINVAL;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  free_replicas(c);
  free_model(c);
  esp::obs_free(c);
  esp::learn_free(c);
  dfree(c->d_clock);
  dfree(c->d_done);
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
    hm.implicit = false;
  }
  if (hm.implicit && phase_lds_bytes(LAY_IMP, hm.nb, hm.zmax, hm.rows_max) > 49152) hm.implicit = false;
  if ((size_t)hm.zmax * N * sizeof(int) >= (1ull << 32) || (size_t)(N + 64) * 8 >= (1ull << 32))
    return fail(c, ESP_EINVAL, "bond table or replica beyond 4 GiB (32-bit device offsets)");
  free_replicas(c);
  free_model(c);
  esp::obs_free(c);             // observables belong to a model: all off
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
  if (m.implicit) {
    if ((rc = upload(c, c->d_grp, m.grp))) return rc;
    if ((rc = upload(c, c->d_gslot, m.gslot))) return rc;
    if ((rc = upload(c, c->d_magb, m.magb))) return rc;
  }
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
  c->lds_bytes = (size_t)(N + 1) * sizeof(double4) + c->ctab_bytes + kLdsRed * sizeof(double);
  c->Ns = ((N + 1 + 63) / 64) * 64;
  c->lds_fits = c->lds_bytes + 4096 <= c->lds_max;
  c->use_lds = c->lds_fits && c->path != 1;
  c->tps = auto_tps(c);
  if (c->lds_fits && c->lds_bytes > 65536) {
    // opt in to the large LDS allocation for every LDS kernel instance
    for (int mode = 0; mode < 2; ++mode) {
      LdsFn fn = lds_fn(mode, m);
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
    const bool imp = hm.implicit && hm.compressed &&
                     (size_t)hm.nb * hm.zmax * 5 * sizeof(double) <= 32768 &&
                     phase_lds_bytes(LAY_IMP, hm.nb, hm.zmax, hm.rows_max) <= 49152;
    info[0] = hm.N; info[1] = hm.zmax; info[2] = hm.ncol;
    info[3] = hm.compressed ? (imp ? LAY_IMP : LAY_CMP) : LAY_BOND;
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
    c->tps = auto_tps(c);
    drop_graph(c);
  }
  return ESP_OK;
}

int esp_set_tps(esp_ctx *c, int tps) {
  if (!c || !(tps == 0 || tps == 1 || tps == 2 || tps == 4 || tps == 8)) return ESP_EINVAL;
  c->tps_user = tps;
  if (c->have_model) {
    c->tps = auto_tps(c);
    drop_graph(c);
  }
  return ESP_OK;
}

int esp_set_indexing(esp_ctx *c, int mode) {
  if (!c || (mode != 0 && mode != 1)) return ESP_EINVAL;
  c->indexing = mode;
  drop_graph(c);
  return ESP_OK;
}

int esp_set_stream_base(esp_ctx *c, uint32_t base) {
  if (!c) return ESP_EINVAL;
  c->stream_base = base;
  drop_graph(c);
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
  out[3] = phase_layout(c); out[4] = c->use_lds ? 1 : 0;
  out[5] = c->tps;
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
  if ((rc = dalloc(c, c->d_sp, (size_t)3 * R * c->Ns))) return rc;
  // the padding sites (index N: the zero spin bond padding points at) stay zero
  HIPCHK(c, hipMemsetAsync(c->d_sp, 0, sizeof(double) * 3 * R * c->Ns, c->stream));
  if ((rc = dalloc(c, c->d_beta, R))) return rc;
  if ((rc = dalloc(c, c->d_ebase, R))) return rc;
  if ((rc = dalloc(c, c->d_erun, R))) return rc;
  if ((rc = dalloc(c, c->d_acc, (size_t)6 * R))) return rc;
  if ((rc = dalloc(c, c->d_partial, (size_t)4 * R * c->red_blocks))) return rc;
  if ((rc = dalloc(c, c->d_slot_rep, R))) return rc;
  if ((rc = dalloc(c, c->d_rep_slot, R))) return rc;
  if ((rc = dalloc(c, c->d_rej, (size_t)kRejSub * R))) return rc;
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
  HIPCHK(c, hipMemsetAsync(c->d_rej, 0, sizeof(unsigned long long) * kRejSub * R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_nswap, 0, sizeof(unsigned long long) * 2 * R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_nmeas, 0, sizeof(unsigned long long), c->stream));
  if ((rc = esp::obs_resize(c))) return rc;
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
                     (const double *)d_in, (const int *)c->d_ref2dev, N, c->Ns, R, c->d_sp);
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
                     (const double *)c->d_sp, (const int *)c->d_ref2dev,
                     by_slot ? (const int *)c->d_slot_rep : (const int *)nullptr, N, c->Ns, R, d_out);
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

int esp_energy_noalloc(esp_ctx *c, double *E_dev) {
  if (!c) return ESP_EINVAL;
  if (!c->have_model) return fail(c, ESP_ESTATE, "no model: call esp_set_model first");
  if (c->R <= 0) return fail(c, ESP_ESTATE, "no replicas: call esp_set_replicas first");
  if (!E_dev) return fail(c, ESP_EINVAL, "E is NULL");
  const size_t need = (size_t)c->R * c->red_blocks;
  if (need > kNoallocScratch)
    return fail(c, ESP_EINVAL, "esp_energy_noalloc: replicas x energy blocks (" + std::to_string(need) +
                                   ") exceed the scratch of " + std::to_string(kNoallocScratch) + " doubles");
  HIPCHK(c, hipSetDevice(c->device));
  double *part = scratch_address(c->device);
  if (!part) return fail(c, ESP_EHIP, "esp_energy_noalloc: hipGetSymbolAddress of the scratch failed");
  const esp::HostModel &m = c->hm;
  KArgs a = base_args(c);
  EnergyFn fn = energy_fn(m.compressed, m.has_B, m.has_D);
  const int nblk = c->red_blocks;
  // every partial the finishing kernel reads is written by the first: nothing to clear
  hipLaunchKernelGGL(fn, dim3(nblk, c->R), dim3(kBlock),
                     (m.compressed ? c->ctab_bytes : 0) + (kBlock / 64) * sizeof(double), c->stream, a,
                     nblk, part);
  hipLaunchKernelGGL(k_energy_fin, dim3((c->R + 255) / 256), dim3(256), 0, c->stream, (const double *)part,
                     nblk, c->R, E_dev);
  HIPCHK(c, hipGetLastError());
  return ESP_OK;
}

int esp_sweep(esp_ctx *c, int nsweeps, int mode, double tilt, uint64_t seed, uint64_t sweep0) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (nsweeps < 0 || (mode != ESP_MODE_RANDOM && mode != ESP_MODE_CONST && mode != ESP_MODE_LEARNED))
    return fail(c, ESP_EINVAL, "bad nsweeps / mode");
  if (nsweeps == 0) return ESP_OK;
  if ((rc = set_clock(c, sweep0, 0, 1))) return rc;
  if (lds_sweeps(c, mode)) rc = enqueue_lds_sweeps(c, mode, tilt, seed, nsweeps, 0);
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
  KArgs a = base_args(c);   // one lane per site here: no per-wave sums
  a.prop = c->d_prop; a.unif = c->d_unif; a.accepted = c->d_accd;
  c->mpart_valid = false;
  PhaseFn fn = phase_fn(M_HOSTFED, c, launch_lanes(c));
  if ((rc = set_clock(c, 0, 0, 1))) return rc;
  for (int col = 0; col < m.ncol; ++col) {
    a.col = col;
    a.c_begin = m.col_off[col];
    a.c_count = m.col_off[col + 1] - m.col_off[col];
    if (!a.c_count) continue;
    a.tps = 1;
    a.nblk = (a.c_count + kBlock - 1) / kBlock;
    a.xcd_remap = (c->R % 8 == 0);
    hipLaunchKernelGGL(fn, dim3(a.nblk * c->R), dim3(kBlock), phase_shm(c), c->stream, a);
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
      (start_parity != 1 && start_parity != 2) ||
      (mode != ESP_MODE_RANDOM && mode != ESP_MODE_CONST && mode != ESP_MODE_LEARNED))
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
  if (nsweeps < 1 || (mode != ESP_MODE_RANDOM && mode != ESP_MODE_CONST && mode != ESP_MODE_LEARNED) ||
      !kernel_ms || !nlaunch)
    return fail(c, ESP_EINVAL, "bad esp_time_sweeps arguments");
  if ((rc = set_clock(c, sweep0, 0, 1))) return rc;
  const esp::HostModel &m = c->hm;
  const bool lds = lds_sweeps(c, mode);
  // one event pair around every sweep-kernel launch, on the context's stream
  const int per = lds ? 1 : nsweeps * m.ncol;
  std::vector<hipEvent_t> ev(2 * per, nullptr);
  for (auto &e : ev) HIPCHK(c, hipEventCreate(&e));
  int k = 0;
  if (mode == ESP_MODE_LEARNED) {
    if ((rc = esp::learn_enqueue_sweeps(c, seed, nsweeps, 0, ev.data(), &k))) {
      for (auto &e : ev) (void)hipEventDestroy(e);
      return rc;
    }
  } else if (lds) {
    KArgs a = base_args(c);
    a.tilt = tilt;
    a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32);
    LdsMeas lm{c->d_acc, c->d_erun, c->d_ebase, c->d_nmeas, 0};
    LdsFn fn = lds_fn(mode, m);
    hipExtLaunchKernelGGL(fn, dim3(c->R), dim3(kLdsBlock), c->lds_bytes, c->stream, ev[0], ev[1], 0,
                          a, nsweeps, lm);
    HIPCHK(c, hipGetLastError());
    k = 1;
  } else {
    KArgs a = phase_args(c, tilt, seed);
    PhaseFn fn = phase_fn(mode, c, launch_lanes(c));
    const size_t shm = phase_shm(c);
    for (int sw = 0; sw < nsweeps; ++sw)
      for (int col = 0; col < m.ncol; ++col) {
        set_colour(a, m, col);
        if (a.c_count == 0) continue;
        a.sweep_off = sw;
        // start/stop stamped by the dispatch packet itself
        hipExtLaunchKernelGGL(fn, dim3(a.nblk * c->R), dim3(kBlock), shm, c->stream, ev[2 * k],
                              ev[2 * k + 1], 0, a);
        ++k;
      }
    HIPCHK(c, hipGetLastError());
  }
  c->mpart_valid = !lds && mode != ESP_MODE_LEARNED;
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
    *updates_per_launch = lds ? (int64_t)m.N * c->R * nsweeps : (int64_t)m.N * c->R * nsweeps / k;
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
  std::vector<int64_t> rsub(rejected ? (size_t)kRejSub * R : 0);
  if (rejected)
    HIPCHK(c, hipMemcpyAsync(rsub.data(), c->d_rej, sizeof(int64_t) * rsub.size(), hipMemcpyDeviceToHost, c->stream));
  if (nmeas) HIPCHK(c, hipMemcpyAsync(nmeas, c->d_nmeas, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (rejected) sum_rej_sub(rsub, R, rejected);
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

int esp_set_trace(esp_ctx *c, int capacity) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (capacity < 0 || (int64_t)capacity * c->R > ((int64_t)1 << 28))
    return fail(c, ESP_EINVAL, "esp_set_trace: capacity out of range");
  dfree(c->d_trace);
  c->trace_cap = 0;
  if (capacity > 0) {
    if ((rc = dalloc(c, c->d_trace, (size_t)capacity * c->R * 2))) return rc;
    c->trace_cap = capacity;
  }
  drop_graph(c);
  return esp::obs_resize(c);
}

int esp_get_trace(esp_ctx *c, double *trace, int *n) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  if (!n) return fail(c, ESP_EINVAL, "esp_get_trace: null count");
  unsigned long long nm = 0;
  HIPCHK(c, hipMemcpyAsync(&nm, c->d_nmeas, sizeof nm, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int k = (int)std::min<unsigned long long>(nm, (unsigned long long)c->trace_cap);
  *n = k;
  if (k > 0 && trace) {
    HIPCHK(c, hipMemcpyAsync(trace, c->d_trace, sizeof(double) * 2 * (size_t)k * c->R, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return ESP_OK;
}

int esp_reset_stats(esp_ctx *c) {
  int rc = check_ready(c, true);
  if (rc) return rc;
  const int R = c->R;
  HIPCHK(c, hipMemsetAsync(c->d_acc, 0, sizeof(double) * 6 * R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_rej, 0, sizeof(unsigned long long) * kRejSub * R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_nswap, 0, sizeof(unsigned long long) * 2 * R, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_nmeas, 0, sizeof(unsigned long long), c->stream));
  if ((rc = esp::obs_clear(c))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_sync(esp_ctx *c) {
  if (!c) return ESP_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

void *esp_stream(esp_ctx *c) { return c ? (void *)c->stream : nullptr; }

int esp_copy_probe(esp_ctx *c, int64_t n_bytes, int reps, double *ms) {
  if (!c) return ESP_EINVAL;
  if (n_bytes < 32 || n_bytes % 32 != 0 || n_bytes > ((int64_t)1 << 34) || reps < 1 || !ms)
    return fail(c, ESP_EINVAL, "esp_copy_probe: n_bytes must be a positive multiple of 32 (<= 16 GiB)");
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ESP_EHIP, "hipSetDevice failed");
  const size_t n = (size_t)n_bytes / sizeof(double4);
  double4 *src = nullptr, *dst = nullptr;
  int rc;
  if ((rc = dalloc(c, src, n))) return rc;
  if ((rc = dalloc(c, dst, n))) {
    dfree(src);
    return rc;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 1u << 20);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, (const double4 *)src, dst, n);  // first touch
    e = hipEventRecord(e0, c->stream);
  }
  for (int k = 0; e == hipSuccess && k < reps; ++k) {
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(256), 0, c->stream, (const double4 *)src, dst, n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  dfree(src);
  dfree(dst);
  if (e != hipSuccess) return fail(c, ESP_EHIP, std::string("esp_copy_probe: ") + hipGetErrorString(e));
  *ms = t / reps;
  return ESP_OK;
}

}  // extern "C"

// esp_batch.hip -- many independent PT chains (lattices of uneven size and kind, each with
// its own ladder) advanced together: one launch per colour index covers every chain, the
// chains whose replica fits in LDS run in a parallel graph branch, and a whole round
// (sweeps, measurements, exchanges of every chain) is captured in one hipGraph.
//
// Per chain the update rule, RNG counters and exchange are those of the single-context
// engine (esp_hip.hip); chain k draws from Philox stream base k << 16, so a chain advanced
// here reproduces a single context given the same stream base.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/espins_hip.h"
#include "esp_kernels.h"
#include "esp_model.h"

using namespace espd;
using namespace espk;

namespace {

// ---- batched kernels -------------------------------------------------------------------
// tiles: (chain, replica, block, 0); chain < 0 = padding tile (XCD interleave)
template <int MODE, int LAY, bool HB, bool HD, bool SELF, int CH>
__global__ __launch_bounds__(kBlock) ESP_PHASE_ATTR void k_bphase(const KArgs *chains, const int4 *tiles, int colour,
                                                   int sweep_off) {
  extern __shared__ double s_ctab[];
  const int4 t = tiles[blockIdx.x];
  if (t.x < 0) return;
  KArgs a = chains[t.x];
  a.col = colour;
  a.c_begin = gldi(a.col_off, colour);
  a.c_count = gldi(a.col_off, colour + 1) - a.c_begin;
  a.sweep_off = sweep_off;
  phase_body<MODE, LAY, HB, HD, SELF, CH>(a, t.y, t.z, s_ctab);
}

// LDS-resident chains: one block per (chain, replica)
template <int MODE, bool CMP, bool HB, bool HD, bool SELF>
__global__ __launch_bounds__(kLdsBlock) void k_blds(const KArgs *chains, const LdsMeas *meas,
                                                    const int2 *items, int nsweeps, int measure_every) {
  extern __shared__ double4 s_sp[];
  const int2 it = items[blockIdx.x];
  LdsMeas m = meas[it.x];
  m.measure_every = measure_every;
  lds_body<MODE, CMP, HB, HD, SELF>(chains[it.x], it.y, nsweeps, m, s_sp);
}

// partial (w, x, y, z) sums of global-path replicas: grid (max blocks, items).  Every round
// sweeps all colours first, so the spin sums come from the colour phases' per-wave sums.
__global__ __launch_bounds__(kBlock) void k_breduce(const KArgs *chains, const int *red_blocks,
                                                    double *const *partials, const int2 *items) {
  const int2 it = items[blockIdx.y];
  const int nblk = red_blocks[it.x];
  const int b = blockIdx.x;
  if (b >= nblk) return;   // uniform per block
  const KArgs &a = chains[it.x];
  reduce_body(a.sp, a.w, a.N, a.Ns, nblk, partials[it.x], a.mpart, a.ncol * a.mpart_nw, it.y, b);
}

// one block per chain; the last block to finish advances the shared clock
__global__ __launch_bounds__(1024) void k_bfinal(const FinArgs *fa, Clock *clock, unsigned *done,
                                                 int measure, int exchange, int advance,
                                                 const int *measured_in_lds) {
  const int start = clock->parity;
  const unsigned long long step = clock->step;
  FinArgs f = fa[blockIdx.x];
  const bool lds = measured_in_lds[blockIdx.x] != 0;
  f.measure = measure && !lds;          // LDS chains measured inside their kernel
  f.exchange = exchange;
  if (lds) f.partial = nullptr;          // their running energies are already fresh
  finalize_body(f, start, step);  // (LDS chains count their measurements in their kernel)
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned t = atomicAdd(done, 1u);
    if (t == gridDim.x - 1) {
      advance_clock(clock, exchange, advance);
      *done = 0u;
      __threadfence();
    }
  }
}

using BPhaseFn = void (*)(const KArgs *, const int4 *, int, int);
using BLdsFn = void (*)(const KArgs *, const LdsMeas *, const int2 *, int, int);

// variant bits: 0-1 coefficient layout (LAY_*), 2 biquadratic, 3 DM, 4 self bonds; launch
// variants add bit 5: partner gathers in flight per lane 4 instead of 2 (global path)
constexpr int kVarChunk4 = 32;
template <int MODE, int LAY, bool HB, bool HD, int CH>
BPhaseFn bp4(bool self) { return self ? k_bphase<MODE, LAY, HB, HD, true, CH> : k_bphase<MODE, LAY, HB, HD, false, CH>; }
template <int MODE, int LAY, bool HB, int CH>
BPhaseFn bp3(bool hd, bool self) { return hd ? bp4<MODE, LAY, HB, true, CH>(self) : bp4<MODE, LAY, HB, false, CH>(self); }
template <int MODE, int LAY, int CH>
BPhaseFn bp2(bool hb, bool hd, bool self) {
  return hb ? bp3<MODE, LAY, true, CH>(hd, self) : bp3<MODE, LAY, false, CH>(hd, self);
}
template <int MODE, int CH>
BPhaseFn bp1(int v) {
  const int lay = v & 3;
  if (lay == LAY_IMP) return bp2<MODE, LAY_IMP, CH>(v & 4, v & 8, v & 16);
  return lay == LAY_CMP ? bp2<MODE, LAY_CMP, CH>(v & 4, v & 8, v & 16) : bp2<MODE, LAY_BOND, CH>(v & 4, v & 8, v & 16);
}
template <int MODE>
BPhaseFn bp0(int v) { return (v & kVarChunk4) ? bp1<MODE, 4>(v) : bp1<MODE, kChunk>(v); }
BPhaseFn bphase_fn(int mode, int variant) { return mode == M_RANDOM ? bp0<M_RANDOM>(variant) : bp0<M_CONST>(variant); }

template <int MODE, bool CMP, bool HB, bool HD>
BLdsFn bl4(bool self) { return self ? k_blds<MODE, CMP, HB, HD, true> : k_blds<MODE, CMP, HB, HD, false>; }
template <int MODE, bool CMP, bool HB>
BLdsFn bl3(bool hd, bool self) { return hd ? bl4<MODE, CMP, HB, true>(self) : bl4<MODE, CMP, HB, false>(self); }
template <int MODE, bool CMP>
BLdsFn bl2(bool hb, bool hd, bool self) { return hb ? bl3<MODE, CMP, true>(hd, self) : bl3<MODE, CMP, false>(hd, self); }
template <int MODE>
BLdsFn bl1(int v) { return (v & 3) ? bl2<MODE, true>(v & 4, v & 8, v & 16) : bl2<MODE, false>(v & 4, v & 8, v & 16); }
BLdsFn blds_fn(int mode, int variant) { return mode == M_RANDOM ? bl1<M_RANDOM>(variant) : bl1<M_CONST>(variant); }

int variant_of(const esp::HostModel &m) {
  const int lay = m.compressed ? (m.implicit ? LAY_IMP : LAY_CMP) : LAY_BOND;
  return lay | (m.has_B ? 4 : 0) | (m.has_D ? 8 : 0) | (m.has_self ? 16 : 0);
}

// energy of one chain's replicas (single-context kernels reused)
template <bool CMP, bool HB, bool HD>
void energy_launch(const KArgs &a, int nblk, double *partial, size_t shm, hipStream_t s) {
  hipLaunchKernelGGL((k_energy<CMP, HB, HD>), dim3(nblk, a.R), dim3(kBlock), shm, s, a, nblk, partial);
}
void launch_energy(const KArgs &a, int variant, int nblk, double *partial, size_t shm, hipStream_t s) {
  const bool c = (variant & 3) != 0, b = variant & 4, d = variant & 8;
  if (c) {
    if (b) d ? energy_launch<true, true, true>(a, nblk, partial, shm, s) : energy_launch<true, true, false>(a, nblk, partial, shm, s);
    else d ? energy_launch<true, false, true>(a, nblk, partial, shm, s) : energy_launch<true, false, false>(a, nblk, partial, shm, s);
  } else {
    if (b) d ? energy_launch<false, true, true>(a, nblk, partial, shm, s) : energy_launch<false, true, false>(a, nblk, partial, shm, s);
    else d ? energy_launch<false, false, true>(a, nblk, partial, shm, s) : energy_launch<false, false, false>(a, nblk, partial, shm, s);
  }
}

struct Chain {
  esp::HostModel hm;
  int L[3] = {0, 0, 0};
  int R = 0, Ns = 0;
  std::vector<double> tems, spins;       // initial spins, reference order
  int variant = 0, tps = 1, lds = 0, lds_fits = 0, red_blocks = 1;
  int chunk = kChunk;                    // partner gathers in flight per lane (global path): 2 or 4
  // blocks per replica of the per-step sums (k_breduce): one per 4096 per-wave entries; the
  // direct energy's red_blocks cover 4096 spins each, 64 times the work of a per-wave sum
  int sum_blocks = 1;
  size_t ctab_bytes = 0, lds_bytes = 0;
  // device allocations (owned)
  double *sp = nullptr, *w = nullptr, *beta = nullptr, *ebase = nullptr, *erun = nullptr,
         *acc = nullptr, *partial = nullptr, *cJ = nullptr, *cB = nullptr, *cD = nullptr,
         *ctab = nullptr, *mag = nullptr, *si4 = nullptr, *h4 = nullptr;
  int *nbr = nullptr, *dev2ref = nullptr, *ref2dev = nullptr, *col_off = nullptr,
      *slot_rep = nullptr, *rep_slot = nullptr;
  uint8_t *basis = nullptr;
  unsigned long long *rej = nullptr, *nswap = nullptr, *nmeas = nullptr;
  double *mpart = nullptr;   // per-wave spin sums of the colour phases (global-path chains)
  int mpart_nw = 0;
  int *grp = nullptr, *gslot = nullptr;   // implicit geometry (esp_geo.h)
  double *magb = nullptr;
};

// Kernel grouping of a chain's colour phases: its Hamiltonian variant and, on the global path,
// the chunk width.
int launch_variant(const Chain &c) { return c.variant | (!c.lds && c.chunk == 4 ? kVarChunk4 : 0); }

}  // namespace

struct esp_batch {
  int device = 0;
  hipStream_t stream = nullptr, side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  std::string err;
  size_t lds_max = 65536;
  std::vector<Chain> chains;
  std::map<uint64_t, std::vector<int>> colour_cache;  // structure hash -> colouring
  bool committed = false;
  // device tables
  KArgs *d_chains = nullptr;
  LdsMeas *d_meas = nullptr;
  FinArgs *d_fin = nullptr;
  int *d_red_blocks = nullptr, *d_lds_flag = nullptr;
  double **d_partials = nullptr;
  int2 *d_red_items = nullptr;
  int n_red_items = 0, max_red_blocks = 1;
  Clock *d_clock = nullptr;
  unsigned *d_done = nullptr;
  int maxcol = 0;
  // per variant: tiles per colour, LDS items
  struct Var {
    int variant = 0;
    std::vector<int4 *> d_tiles;          // per colour
    std::vector<int> ntiles;
    int2 *d_lds = nullptr;
    int nlds = 0;
    size_t shm_phase = 0, shm_lds = 0;
  };
  std::vector<Var> vars;
  hipGraphExec_t gexec = nullptr;
  std::string gkey;
};

namespace {

int bfail(esp_batch *b, int code, const std::string &m) {
  if (b) b->err = m;
  return code;
}

#define BCHK(b, expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return bfail(b, ESP_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

template <class T>
int balloc(esp_batch *b, T *&p, size_t n) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
  if (hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    p = nullptr;
    return bfail(b, ESP_ENOMEM, "hipMalloc failed");
  }
  return ESP_OK;
}

template <class T>
int bupload(esp_batch *b, T *&p, const T *src, size_t n) {
  int rc = balloc(b, p, n);
  if (rc) return rc;
  if (n) BCHK(b, hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyHostToDevice, b->stream));
  return ESP_OK;
}

template <class T>
void bfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

void free_chain(Chain &c) {
  bfree(c.sp); bfree(c.w); bfree(c.beta); bfree(c.ebase); bfree(c.erun); bfree(c.acc);
  bfree(c.partial); bfree(c.cJ); bfree(c.cB); bfree(c.cD); bfree(c.ctab); bfree(c.mag);
  bfree(c.si4); bfree(c.h4); bfree(c.nbr); bfree(c.dev2ref); bfree(c.ref2dev); bfree(c.col_off);
  bfree(c.slot_rep); bfree(c.rep_slot); bfree(c.basis); bfree(c.rej); bfree(c.nswap); bfree(c.nmeas);
  bfree(c.mpart);
  c.mpart_nw = 0;
  bfree(c.grp); bfree(c.gslot); bfree(c.magb);
}

void drop_plan(esp_batch *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->gkey.clear();
  for (auto &v : b->vars) {
    for (auto *t : v.d_tiles) bfree(t);
    bfree(v.d_lds);
  }
  b->vars.clear();
  bfree(b->d_chains); bfree(b->d_meas); bfree(b->d_fin); bfree(b->d_red_blocks);
  bfree(b->d_lds_flag); bfree(b->d_partials); bfree(b->d_red_items);
}

KArgs chain_args(const Chain &c, int k) {
  KArgs a{};
  a.sp = c.sp; a.w = c.w; a.nbr = c.nbr; a.cJ = c.cJ; a.cB = c.cB; a.cD = c.cD; a.ctab = c.ctab;
  a.mag = c.mag; a.basis = c.basis; a.dev2ref = c.dev2ref; a.col_off = c.col_off;
  a.si4 = c.si4; a.h4 = c.h4; a.beta_slot = c.beta; a.rep_slot = c.rep_slot; a.rej = c.rej;
  a.N = c.hm.N; a.Ns = c.Ns; a.nb = c.hm.nb; a.zmax = c.hm.zmax; a.R = c.R; a.ncol = c.hm.ncol;
  a.has_si = c.hm.has_si; a.has_h = c.hm.has_h; a.tps = c.tps;
  a.mpart_nw = c.mpart_nw;
  a.stream_base = (uint32_t)k << 16;
  a.geo = c.hm.geo;
  a.grp = c.grp; a.gslot = c.gslot; a.magb = c.magb;
  a.rows_max = c.hm.rows_max;
  return a;
}

// Per-wave sums of a chain's colour phases -- spins (mpart, rewritten by every sweep) and
// accepted dE (w, accumulated) -- laid out for its current lanes per site.  Before the
// accumulators are re-laid out their sums move into e_base.
int size_wave_sums(esp_batch *b, Chain &c) {
  int nw = 1;
  for (int col = 0; col < c.hm.ncol; ++col)
    nw = std::max(nw, ((c.hm.col_off[col + 1] - c.hm.col_off[col]) * c.tps + 63) / 64);
  const size_t n = (size_t)c.R * c.hm.ncol * nw;
  if (nw == c.mpart_nw && c.mpart && c.w) {
    BCHK(b, hipMemsetAsync(c.mpart, 0, 3 * n * sizeof(double), b->stream));
    return ESP_OK;
  }
  if (c.w && c.mpart_nw > 0) {
    hipLaunchKernelGGL(k_fold_w, dim3((c.R + 255) / 256), dim3(256), 0, b->stream, c.w,
                       c.hm.ncol * c.mpart_nw, c.R, c.ebase);
    BCHK(b, hipGetLastError());
  }
  int rc;
  if ((rc = balloc(b, c.mpart, 3 * n)) || (rc = balloc(b, c.w, n))) return rc;
  BCHK(b, hipMemsetAsync(c.mpart, 0, 3 * n * sizeof(double), b->stream));
  BCHK(b, hipMemsetAsync(c.w, 0, n * sizeof(double), b->stream));
  c.mpart_nw = nw;
  return ESP_OK;
}

// Build device descriptor tables and the per-colour tile lists for the current plan.
int build_plan(esp_batch *b, double tilt, uint64_t seed) {
  drop_plan(b);
  const int nc = (int)b->chains.size();
  std::vector<KArgs> ka(nc);
  std::vector<LdsMeas> lm(nc);
  std::vector<FinArgs> fa(nc);
  std::vector<int> rb(nc), ldsf(nc);
  std::vector<double *> parts(nc);
  std::vector<int2> red_items;
  b->maxcol = 0;
  b->max_red_blocks = 1;
  for (int k = 0; k < nc; ++k) {
    Chain &c = b->chains[k];
    int rc0;
    if ((rc0 = size_wave_sums(b, c))) return rc0;
    c.sum_blocks = std::max(1, std::min(c.red_blocks, (c.hm.ncol * c.mpart_nw + 4095) / 4096));
    ka[k] = chain_args(c, k);
    if (!c.lds) ka[k].mpart = c.mpart;     // spin sums only where colour phases run
    ka[k].clock = b->d_clock;
    ka[k].tilt = tilt;
    ka[k].key0 = (uint32_t)seed; ka[k].key1 = (uint32_t)(seed >> 32);
    lm[k] = LdsMeas{c.acc, c.erun, c.ebase, c.nmeas, 0};
    FinArgs f{};
    f.partial = c.partial; f.nblk = c.sum_blocks; f.N = c.hm.N; f.R = c.R;
    f.e_base = c.ebase; f.e_run = c.erun; f.acc = c.acc; f.nmeas = c.nmeas;
    f.slot_rep = c.slot_rep; f.rep_slot = c.rep_slot; f.beta_slot = c.beta; f.nswap = c.nswap;
    f.clock = b->d_clock; f.key0 = (uint32_t)seed; f.key1 = (uint32_t)(seed >> 32);
    f.stream_base = (uint32_t)k << 16;
    fa[k] = f;
    rb[k] = c.sum_blocks;
    ldsf[k] = c.lds;
    parts[k] = c.partial;
    b->max_red_blocks = std::max(b->max_red_blocks, c.sum_blocks);
    if (!c.lds)
      for (int r = 0; r < c.R; ++r) red_items.push_back(make_int2(k, r));
    b->maxcol = std::max(b->maxcol, c.hm.ncol);
  }
  int rc;
  if ((rc = bupload(b, b->d_chains, ka.data(), nc))) return rc;
  if ((rc = bupload(b, b->d_meas, lm.data(), nc))) return rc;
  if ((rc = bupload(b, b->d_fin, fa.data(), nc))) return rc;
  if ((rc = bupload(b, b->d_red_blocks, rb.data(), nc))) return rc;
  if ((rc = bupload(b, b->d_lds_flag, ldsf.data(), nc))) return rc;
  if ((rc = bupload(b, b->d_partials, parts.data(), nc))) return rc;
  if ((rc = bupload(b, b->d_red_items, red_items.data(), red_items.size()))) return rc;
  b->n_red_items = (int)red_items.size();
  // launch variants: the kernel variant and, for global-path chains, the chunk width
  std::map<int, int> vid;
  for (auto &c : b->chains)
    if (!vid.count(launch_variant(c))) {
      vid[launch_variant(c)] = (int)b->vars.size();
      b->vars.emplace_back();
      b->vars.back().variant = launch_variant(c);
    }
  for (auto &v : b->vars) {
    // global-path tiles per colour, interleaved so that tile t runs on XCD t % 8 and all
    // tiles of one replica share an XCD (and its L2)
    v.d_tiles.assign(b->maxcol, nullptr);
    v.ntiles.assign(b->maxcol, 0);
    for (int col = 0; col < b->maxcol; ++col) {
      std::vector<std::vector<int4>> q(8);
      int qi = 0;
      for (int k = 0; k < nc; ++k) {
        const Chain &c = b->chains[k];
        if (launch_variant(c) != v.variant || c.lds || col >= c.hm.ncol) continue;
        const int cnt = c.hm.col_off[col + 1] - c.hm.col_off[col];
        const int nblk = (cnt * c.tps + kBlock - 1) / kBlock;
        // Implicit geometry: every group of the colour has the same sub-cell layout, so the
        // blocks go in sub-cell order across the groups (block (g, t) of every group g before
        // any block t + 1).  Blocks that run together then gather from the same region of the
        // lattice and share its lines in L2 (about 30 % fewer L2 misses than group order).
        std::vector<int> order(nblk);
        for (int bb = 0; bb < nblk; ++bb) order[bb] = bb;
        if (c.hm.implicit) {
          const int per = kBlock / c.tps, mt = (int)c.hm.geo.mt.d;
          std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
            const int ox = x * per, oy = y * per;
            return std::make_pair(ox % mt, ox / mt) < std::make_pair(oy % mt, oy / mt);
          });
        }
        for (int r = 0; r < c.R; ++r, ++qi)
          for (int bb = 0; bb < nblk; ++bb) q[qi % 8].push_back(make_int4(k, r, order[bb], 0));
      }
      size_t longest = 0;
      for (auto &x : q) longest = std::max(longest, x.size());
      std::vector<int4> tiles;
      for (size_t i = 0; i < longest; ++i)
        for (int x = 0; x < 8; ++x)
          tiles.push_back(i < q[x].size() ? q[x][i] : make_int4(-1, 0, 0, 0));
      // trim trailing padding
      while (!tiles.empty() && tiles.back().x < 0) tiles.pop_back();
      v.ntiles[col] = (int)tiles.size();
      if (!tiles.empty() && (rc = bupload(b, v.d_tiles[col], tiles.data(), tiles.size()))) return rc;
    }
    std::vector<int2> li;
    size_t shm_l = 0, shm_p = 0;
    for (int k = 0; k < nc; ++k) {
      const Chain &c = b->chains[k];
      if (launch_variant(c) != v.variant) continue;
      shm_p = std::max(shm_p, phase_lds_bytes(c.variant & 3, c.hm.nb, c.hm.zmax, c.hm.rows_max));
      if (!c.lds) continue;
      shm_l = std::max(shm_l, c.lds_bytes);
      for (int r = 0; r < c.R; ++r) li.push_back(make_int2(k, r));
    }
    v.shm_phase = shm_p;
    v.shm_lds = shm_l;
    v.nlds = (int)li.size();
    if (!li.empty() && (rc = bupload(b, v.d_lds, li.data(), li.size()))) return rc;
    if (shm_l > 65536)
      for (int mode = 0; mode < 2; ++mode)
        BCHK(b, hipFuncSetAttribute((const void *)blds_fn(mode, v.variant & (kVarChunk4 - 1)),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm_l));
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  return ESP_OK;
}

// One round for every chain: steps x (M sweeps [+ measure]) [+ exchange].
int enqueue_round(esp_batch *b, int steps, int M, bool measure, bool exchange, int mode) {
  bool any_lds = false;
  for (auto &v : b->vars) any_lds |= v.nlds > 0;
  if (any_lds) {  // fork: LDS chains on the side stream
    BCHK(b, hipEventRecord(b->fork, b->stream));
    BCHK(b, hipStreamWaitEvent(b->side, b->fork, 0));
    for (auto &v : b->vars)
      if (v.nlds > 0)
        hipLaunchKernelGGL(blds_fn(mode, v.variant & (kVarChunk4 - 1)), dim3(v.nlds), dim3(kLdsBlock), v.shm_lds, b->side,
                           (const KArgs *)b->d_chains, (const LdsMeas *)b->d_meas,
                           (const int2 *)v.d_lds, steps * M, measure ? M : 0);
    BCHK(b, hipGetLastError());
    BCHK(b, hipEventRecord(b->join, b->side));
  }
  for (int s = 0; s < steps; ++s) {
    for (int m = 0; m < M; ++m)
      for (int col = 0; col < b->maxcol; ++col)
        for (auto &v : b->vars)
          if (v.ntiles[col] > 0)
            hipLaunchKernelGGL(bphase_fn(mode, v.variant), dim3(v.ntiles[col]), dim3(kBlock), v.shm_phase,
                               b->stream, (const KArgs *)b->d_chains, (const int4 *)v.d_tiles[col], col,
                               s * M + m);
    BCHK(b, hipGetLastError());
    const bool last = s + 1 == steps;
    if (measure && b->n_red_items > 0)
      hipLaunchKernelGGL(k_breduce, dim3(b->max_red_blocks, b->n_red_items), dim3(kBlock), 0, b->stream,
                         (const KArgs *)b->d_chains, (const int *)b->d_red_blocks,
                         (double *const *)b->d_partials, (const int2 *)b->d_red_items);
    if (last) break;
    if (measure)
      hipLaunchKernelGGL(k_bfinal, dim3((unsigned)b->chains.size()), dim3(256), 0, b->stream,
                         (const FinArgs *)b->d_fin, b->d_clock, b->d_done, 1, 0, 0,
                         (const int *)b->d_lds_flag);
    BCHK(b, hipGetLastError());
  }
  if (any_lds) BCHK(b, hipStreamWaitEvent(b->stream, b->join, 0));
  // last step: energies of global chains (reduce) + measurement + exchange + clock
  if (!measure && exchange && b->n_red_items > 0)
    hipLaunchKernelGGL(k_breduce, dim3(b->max_red_blocks, b->n_red_items), dim3(kBlock), 0, b->stream,
                       (const KArgs *)b->d_chains, (const int *)b->d_red_blocks,
                       (double *const *)b->d_partials, (const int2 *)b->d_red_items);
  hipLaunchKernelGGL(k_bfinal, dim3((unsigned)b->chains.size()), dim3(256), 0, b->stream,
                     (const FinArgs *)b->d_fin, b->d_clock, b->d_done, measure ? 1 : 0, exchange ? 1 : 0,
                     steps * M, (const int *)b->d_lds_flag);
  BCHK(b, hipGetLastError());
  return ESP_OK;
}

int chain_energy(esp_batch *b, int k, double *d_out) {
  Chain &c = b->chains[k];
  KArgs a = chain_args(c, k);
  launch_energy(a, c.variant, c.red_blocks, c.partial,
                (c.hm.compressed ? c.ctab_bytes : 0) + (kBlock / 64) * sizeof(double), b->stream);
  hipLaunchKernelGGL(k_energy_fin, dim3((c.R + 255) / 256), dim3(256), 0, b->stream,
                     (const double *)c.partial, c.red_blocks, c.R, d_out);
  BCHK(b, hipGetLastError());
  return ESP_OK;
}

int check_chain(esp_batch *b, int k) {
  if (!b) return ESP_EINVAL;
  if (!b->committed) return bfail(b, ESP_ESTATE, "batch not committed");
  if (k < 0 || k >= (int)b->chains.size()) return bfail(b, ESP_EINVAL, "chain index out of range");
  if (hipSetDevice(b->device) != hipSuccess) return bfail(b, ESP_EHIP, "hipSetDevice failed");
  return ESP_OK;
}

}  // namespace

extern "C" {

int esp_batch_create(esp_batch **out, int device) {
  if (!out) return ESP_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return ESP_EHIP;
  if (device < 0 || device >= n) return ESP_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return ESP_EHIP;
  esp_batch *b = new esp_batch();
  b->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) b->lds_max = prop.sharedMemPerBlock;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess ||
      hipMalloc((void **)&b->d_clock, sizeof(Clock)) != hipSuccess ||
      hipMalloc((void **)&b->d_done, sizeof(unsigned)) != hipSuccess) {
    delete b;
    return ESP_EHIP;
  }
  (void)hipMemset(b->d_done, 0, sizeof(unsigned));
  *out = b;
  return ESP_OK;
}

int esp_batch_destroy(esp_batch *b) {
  if (!b) return ESP_EINVAL;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  (void)hipStreamSynchronize(b->side);
  drop_plan(b);
  for (auto &c : b->chains) free_chain(c);
  bfree(b->d_clock);
  bfree(b->d_done);
  (void)hipEventDestroy(b->fork);
  (void)hipEventDestroy(b->join);
  (void)hipStreamDestroy(b->side);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return ESP_OK;
}

const char *esp_batch_last_error(const esp_batch *b) { return b ? b->err.c_str() : "null batch"; }

int esp_batch_add_chain(esp_batch *b, int N, int nb, const double *mag, int zJ, const int *j_cnt,
                        const int *j_nbr, const double *j_val, int zB, const int *b_cnt,
                        const int *b_nbr, const double *b_val, int zD, const int *d_cnt,
                        const int *d_nbr, const double *d_vec, int have_si, const double *si_par,
                        const double *si_axis, int have_field, const double *h_par,
                        const double *h_axis, const int *L3, int R, const double *tems,
                        const double *spins, int *chain_id) {
  if (!b) return ESP_EINVAL;
  if (b->committed) return bfail(b, ESP_ESTATE, "batch already committed");
  if (R <= 0 || R > 65535 || !tems || !spins) return bfail(b, ESP_EINVAL, "bad R / tems / spins");
  for (int t = 0; t < R; ++t)
    if (!(tems[t] > 0.0)) return bfail(b, ESP_EINVAL, "temperatures must be > 0");
  if (b->chains.size() >= 65535) return bfail(b, ESP_EINVAL, "too many chains");
  Chain c;
  const uint64_t h = (zJ >= 0 && j_cnt && j_nbr)
                         ? esp::structure_hash(N, nb, zJ, j_cnt, j_nbr, zB, b_cnt, b_nbr, zD, d_cnt, d_nbr)
                         : 0;
  auto it = b->colour_cache.find(h);
  std::string e = esp::prepare_model(c.hm, N, nb, mag, zJ, j_cnt, j_nbr, j_val, zB, b_cnt, b_nbr,
                                     b_val, zD, d_cnt, d_nbr, d_vec, have_si, si_par, si_axis,
                                     have_field, h_par, h_axis, L3,
                                     it != b->colour_cache.end() ? &it->second : nullptr);
  if (!e.empty()) return bfail(b, ESP_EINVAL, e);
  if (h && it == b->colour_cache.end()) b->colour_cache[h] = c.hm.colour;
  c.ctab_bytes = (size_t)c.hm.nb * c.hm.zmax * 5 * sizeof(double);
  if (c.hm.compressed && c.ctab_bytes > 32768) {
    c.hm.compressed = false;
    c.hm.coef.assign((size_t)c.hm.zmax * N * esp::kCoef, 0.0);
    for (int d = 0; d < N; ++d)
      for (int l = 0; l < c.hm.zmax; ++l)
        for (int q = 0; q < esp::kCoef; ++q)
          c.hm.coef[((size_t)l * N + d) * esp::kCoef + q] =
              c.hm.ctab[((size_t)c.hm.basis[d] * c.hm.zmax + l) * esp::kCoef + q];
    c.hm.ctab.clear();
    c.hm.implicit = false;
  }
  if (!c.hm.compressed) c.ctab_bytes = 0;
  if (c.hm.implicit && phase_lds_bytes(LAY_IMP, c.hm.nb, c.hm.zmax, c.hm.rows_max) > 49152)
    c.hm.implicit = false;
  if ((size_t)c.hm.zmax * N * sizeof(int) >= (1ull << 32)) return bfail(b, ESP_EINVAL, "bond table beyond 4 GiB");
  if (L3) std::memcpy(c.L, L3, sizeof c.L);
  c.R = R;
  c.tems.assign(tems, tems + R);
  c.spins.assign(spins, spins + (size_t)3 * N * R);
  c.variant = variant_of(c.hm);
  c.Ns = ((N + 1 + 63) / 64) * 64;
  c.lds_bytes = (size_t)(N + 1) * sizeof(double4) + c.ctab_bytes + kLdsRed * sizeof(double);
  c.lds_fits = c.lds_bytes + 4096 <= b->lds_max;
  c.lds = c.lds_fits;
  c.tps = 1;  // measured fastest on gfx950 for every lattice tried; the autotuner may change it
  c.red_blocks = std::max(1, std::min(64, (N + 4095) / 4096));
  b->chains.push_back(std::move(c));
  if (chain_id) *chain_id = (int)b->chains.size() - 1;
  return ESP_OK;
}

int esp_batch_set_chain_config(esp_batch *b, int k, int path, int tps) {
  if (!b || k < 0 || k >= (int)b->chains.size()) return ESP_EINVAL;
  if (!(tps == 1 || tps == 2 || tps == 4 || tps == 8) || path < 0 || path > 2)
    return bfail(b, ESP_EINVAL, "bad path / tps");
  Chain &c = b->chains[k];
  if (path == 2 && !c.lds_fits) return bfail(b, ESP_EINVAL, "replica does not fit in LDS");
  c.lds = path == 0 ? c.lds_fits : (path == 2);
  c.tps = tps;
  b->gkey.clear();
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  return ESP_OK;
}

int esp_batch_set_chain_chunk(esp_batch *b, int k, int chunk) {
  if (!b || k < 0 || k >= (int)b->chains.size()) return ESP_EINVAL;
  if (chunk != 2 && chunk != 4) return bfail(b, ESP_EINVAL, "chunk must be 2 or 4");
  b->chains[k].chunk = chunk;
  b->gkey.clear();
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  return ESP_OK;
}

int esp_batch_get_chain_chunk(esp_batch *b, int k, int *chunk) {
  if (!b || !chunk || k < 0 || k >= (int)b->chains.size()) return ESP_EINVAL;
  *chunk = b->chains[k].chunk;
  return ESP_OK;
}

int esp_batch_commit(esp_batch *b) {
  if (!b) return ESP_EINVAL;
  if (b->committed) return bfail(b, ESP_ESTATE, "batch already committed");
  if (b->chains.empty()) return bfail(b, ESP_ESTATE, "no chains");
  BCHK(b, hipSetDevice(b->device));
  int rc;
  for (int k = 0; k < (int)b->chains.size(); ++k) {
    Chain &c = b->chains[k];
    const esp::HostModel &m = c.hm;
    const int N = m.N, R = c.R;
    if ((rc = bupload(b, c.nbr, m.nbr.data(), m.nbr.size()))) return rc;
    if ((rc = bupload(b, c.dev2ref, m.dev2ref.data(), m.dev2ref.size()))) return rc;
    if ((rc = bupload(b, c.ref2dev, m.ref2dev.data(), m.ref2dev.size()))) return rc;
    if ((rc = bupload(b, c.col_off, m.col_off.data(), m.col_off.size()))) return rc;
    if ((rc = bupload(b, c.mag, m.mag.data(), m.mag.size()))) return rc;
    if ((rc = bupload(b, c.basis, m.basis.data(), m.basis.size()))) return rc;
    if ((rc = bupload(b, c.si4, m.si4.data(), m.si4.size()))) return rc;
    if ((rc = bupload(b, c.h4, m.h4.data(), m.h4.size()))) return rc;
    if (m.implicit) {
      if ((rc = bupload(b, c.grp, m.grp.data(), m.grp.size()))) return rc;
      if ((rc = bupload(b, c.gslot, m.gslot.data(), m.gslot.size()))) return rc;
      if ((rc = bupload(b, c.magb, m.magb.data(), m.magb.size()))) return rc;
    }
    if (m.compressed) {
      if ((rc = bupload(b, c.ctab, m.ctab.data(), m.ctab.size()))) return rc;
    } else {
      std::vector<double> cj((size_t)m.zmax * N), cb, cd;
      if (m.has_B) cb.resize((size_t)m.zmax * N);
      if (m.has_D) cd.resize((size_t)m.zmax * N * 3);
      for (size_t t = 0; t < (size_t)m.zmax * N; ++t) {
        cj[t] = m.coef[t * esp::kCoef];
        if (m.has_B) cb[t] = m.coef[t * esp::kCoef + 1];
        if (m.has_D)
          for (int q = 0; q < 3; ++q) cd[3 * t + q] = m.coef[t * esp::kCoef + 2 + q];
      }
      if ((rc = bupload(b, c.cJ, cj.data(), cj.size()))) return rc;
      if (m.has_B && (rc = bupload(b, c.cB, cb.data(), cb.size()))) return rc;
      if (m.has_D && (rc = bupload(b, c.cD, cd.data(), cd.size()))) return rc;
    }
    // replicas
    if ((rc = balloc(b, c.sp, (size_t)3 * R * c.Ns))) return rc;
    BCHK(b, hipMemsetAsync(c.sp, 0, sizeof(double) * 3 * R * c.Ns, b->stream));
    std::vector<double> beta(R);
    for (int t = 0; t < R; ++t) beta[t] = 1.0 / c.tems[t];
    std::vector<int> id(R);
    for (int t = 0; t < R; ++t) id[t] = t;
    if ((rc = bupload(b, c.beta, beta.data(), R))) return rc;
    if ((rc = bupload(b, c.slot_rep, id.data(), R))) return rc;
    if ((rc = bupload(b, c.rep_slot, id.data(), R))) return rc;
    if ((rc = balloc(b, c.ebase, R))) return rc;
    if ((rc = balloc(b, c.erun, R))) return rc;
    if ((rc = balloc(b, c.acc, (size_t)6 * R))) return rc;
    if ((rc = balloc(b, c.partial, (size_t)4 * R * c.red_blocks))) return rc;
    if ((rc = balloc(b, c.rej, (size_t)kRejSub * R))) return rc;
    if ((rc = balloc(b, c.nswap, (size_t)2 * R))) return rc;
    if ((rc = balloc(b, c.nmeas, 1))) return rc;
    BCHK(b, hipMemsetAsync(c.acc, 0, sizeof(double) * 6 * R, b->stream));
    BCHK(b, hipMemsetAsync(c.rej, 0, sizeof(unsigned long long) * kRejSub * R, b->stream));
    BCHK(b, hipMemsetAsync(c.nswap, 0, sizeof(unsigned long long) * 2 * R, b->stream));
    BCHK(b, hipMemsetAsync(c.nmeas, 0, sizeof(unsigned long long), b->stream));
    double *d_in = nullptr;
    if ((rc = bupload(b, d_in, c.spins.data(), c.spins.size()))) return rc;
    const size_t n = (size_t)N * R;
    hipLaunchKernelGGL(k_put_spins, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream,
                       (const double *)d_in, (const int *)c.ref2dev, N, c.Ns, R, c.sp);
    BCHK(b, hipGetLastError());
    if ((rc = size_wave_sums(b, c))) return rc;
    if ((rc = chain_energy(b, k, c.ebase))) return rc;
    BCHK(b, hipMemcpyAsync(c.erun, c.ebase, sizeof(double) * R, hipMemcpyDeviceToDevice, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    bfree(d_in);
    std::vector<double>().swap(c.spins);
  }
  b->committed = true;
  return ESP_OK;
}

int esp_batch_run(esp_batch *b, int n_rounds, int steps_per_round, int steps_measure, int do_measure,
                  int do_exchange, int start_parity, int mode, double tilt, uint64_t seed,
                  uint64_t sweep0, uint64_t step0, int use_graph) {
  if (!b) return ESP_EINVAL;
  if (!b->committed) return bfail(b, ESP_ESTATE, "batch not committed");
  if (n_rounds < 0 || steps_per_round < 1 || steps_measure < 1 ||
      (start_parity != 1 && start_parity != 2) || (mode != ESP_MODE_RANDOM && mode != ESP_MODE_CONST))
    return bfail(b, ESP_EINVAL, "bad esp_batch_run arguments");
  BCHK(b, hipSetDevice(b->device));
  if (n_rounds == 0) return ESP_OK;
  char key[256];
  std::snprintf(key, sizeof key, "%d/%d/%d/%d/%d/%.17g/%llu", steps_per_round, steps_measure,
                do_measure, do_exchange, mode, tilt, (unsigned long long)seed);
  int rc;
  if (b->gkey != key) {
    if ((rc = build_plan(b, tilt, seed))) return rc;
  }
  Clock h{sweep0, step0, start_parity, 0};
  BCHK(b, hipMemcpyAsync(b->d_clock, &h, sizeof(Clock), hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  if (!use_graph) {
    for (int k = 0; k < n_rounds; ++k)
      if ((rc = enqueue_round(b, steps_per_round, steps_measure, do_measure, do_exchange, mode))) return rc;
    b->gkey = key;
    BCHK(b, hipStreamSynchronize(b->stream));
    return ESP_OK;
  }
  if (b->gkey != key || !b->gexec) {
    hipGraph_t g = nullptr;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_round(b, steps_per_round, steps_measure, do_measure, do_exchange, mode);
    const hipError_t ce = hipStreamEndCapture(b->stream, &g);
    if (rc) return rc;
    if (ce != hipSuccess) return bfail(b, ESP_EHIP, std::string("capture: ") + hipGetErrorString(ce));
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    const hipError_t ie = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
      b->gexec = nullptr;
      return bfail(b, ESP_EHIP, std::string("instantiate: ") + hipGetErrorString(ie));
    }
  }
  b->gkey = key;
  for (int k = 0; k < n_rounds; ++k) BCHK(b, hipGraphLaunch(b->gexec, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return ESP_OK;
}

}  // extern "C"

// Autotuner -------------------------------------------------------------------------------
namespace {
std::string shape_key(const Chain &c) {
  char k[128];
  std::snprintf(k, sizeof k, "N%d_nb%d_z%d_col%d_v%d", c.hm.N, c.hm.nb, c.hm.zmax, c.hm.ncol, c.variant);
  return k;
}

// Mean time of one round (one sweep of every chain + exchange) of the current plan, after one
// untimed round.  Advances the chains; the tuning clock keeps its own Philox counters.
int time_rounds(esp_batch *b, int nrounds, double *ms) {
  int rc = build_plan(b, 0.125, 0x74756e65ull);
  if (rc) return rc;
  Clock h{(1ull << 61), 0, 1, 0};
  BCHK(b, hipMemcpyAsync(b->d_clock, &h, sizeof(Clock), hipMemcpyHostToDevice, b->stream));
  if ((rc = enqueue_round(b, 1, 1, false, true, ESP_MODE_RANDOM))) return rc;
  hipEvent_t e0, e1;
  BCHK(b, hipEventCreate(&e0));
  BCHK(b, hipEventCreate(&e1));
  BCHK(b, hipEventRecord(e0, b->stream));
  for (int k = 0; k < nrounds && rc == ESP_OK; ++k) rc = enqueue_round(b, 1, 1, false, true, ESP_MODE_RANDOM);
  BCHK(b, hipEventRecord(e1, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  float t = 0.f;
  BCHK(b, hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = t / nrounds;
  return rc;
}
}  // namespace

// The autotuner times real sweeps and exchange rounds on the chains' own buffers; the chain
// state it changes (spins, energy accumulators, running energies, replica <-> slot labels) is
// copied aside first and put back afterwards, so tuning never changes a trajectory.
namespace {
struct ChainSnap {
  double *sp = nullptr, *ebase = nullptr, *erun = nullptr;
  int *slot_rep = nullptr, *rep_slot = nullptr;
};

// Save (save = true) or restore the state of every chain.  Saving first folds the per-wave
// accepted-dE accumulators into e_base (tuning may re-lay them out); restoring zeroes them in
// whatever layout the tuned plan left.
int snap_copy(esp_batch *b, std::vector<ChainSnap> &s, bool save) {
  for (size_t k = 0; k < b->chains.size(); ++k) {
    Chain &c = b->chains[k];
    ChainSnap &q = s[k];
    const size_t nsp = (size_t)3 * c.R * c.Ns, R = (size_t)c.R;
    const int ne = c.hm.ncol * c.mpart_nw;
    if (save) {
      int rc;
      if ((rc = balloc(b, q.sp, nsp)) || (rc = balloc(b, q.ebase, R)) || (rc = balloc(b, q.erun, R)) ||
          (rc = balloc(b, q.slot_rep, R)) || (rc = balloc(b, q.rep_slot, R)))
        return rc;
      if (c.w && ne > 0)
        hipLaunchKernelGGL(k_fold_w, dim3((c.R + 255) / 256), dim3(256), 0, b->stream, c.w, ne, c.R, c.ebase);
    } else if (c.w && ne > 0) {
      BCHK(b, hipMemsetAsync(c.w, 0, sizeof(double) * R * ne, b->stream));
    }
    auto cp = [&](void *dst, const void *src, size_t bytes) {
      return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, b->stream);
    };
    const hipError_t e[5] = {
        save ? cp(q.sp, c.sp, nsp * 8) : cp(c.sp, q.sp, nsp * 8),
        save ? cp(q.ebase, c.ebase, R * 8) : cp(c.ebase, q.ebase, R * 8),
        save ? cp(q.erun, c.erun, R * 8) : cp(c.erun, q.erun, R * 8),
        save ? cp(q.slot_rep, c.slot_rep, R * 4) : cp(c.slot_rep, q.slot_rep, R * 4),
        save ? cp(q.rep_slot, c.rep_slot, R * 4) : cp(c.rep_slot, q.rep_slot, R * 4)};
    for (hipError_t x : e)
      if (x != hipSuccess) return bfail(b, ESP_EHIP, std::string("autotune snapshot: ") + hipGetErrorString(x));
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  return ESP_OK;
}

void snap_free(std::vector<ChainSnap> &s) {
  for (auto &q : s) {
    bfree(q.sp); bfree(q.ebase); bfree(q.erun); bfree(q.slot_rep); bfree(q.rep_slot);
  }
}

int autotune_body(esp_batch *b, int nrounds, const char *cache_path);
}  // namespace

extern "C" int esp_batch_autotune(esp_batch *b, int nsweeps, const char *cache_path) {
  if (!b) return ESP_EINVAL;
  if (!b->committed) return bfail(b, ESP_ESTATE, "commit the batch before tuning");
  if (nsweeps < 1) return bfail(b, ESP_EINVAL, "nsweeps must be >= 1");
  BCHK(b, hipSetDevice(b->device));
  std::vector<ChainSnap> snap(b->chains.size());
  int rc = snap_copy(b, snap, true);
  if (rc == ESP_OK) rc = autotune_body(b, std::max(2, nsweeps), cache_path);
  if (rc == ESP_OK) rc = snap_copy(b, snap, false);
  (void)hipStreamSynchronize(b->stream);
  snap_free(snap);
  if (rc) return rc;
  return esp_batch_reset_stats(b);
}

namespace {
// Whole-batch search.  Chains of one shape share a kernel configuration (global path or
// LDS-resident, lanes per site, partner gathers in flight per lane).  Timing a shape alone
// favours what hides latency on an idle GPU (more lanes per site, the LDS branch) although
// the batch as a whole is throughput-bound,
// so every candidate is judged by the measured time of a whole round of the batch: start from
// the throughput default (global path, one lane per site) or the cached choices, then, shape
// by shape in order of work, keep a configuration only when it makes the round at least 0.5 %
// faster.  Each measurement is nrounds rounds after an untimed one.  A shape is the lattice
// and Hamiltonian (N, basis, bond width, colours, kernel variant); the replica count only
// scales the work.
int autotune_body(esp_batch *b, int nrounds, const char *cache_path) {
  struct Shape {
    int lds = 0, tps = 1, chunk = kChunk;
  };
  std::map<std::string, Shape> decided;
  if (cache_path) {
    if (FILE *f = std::fopen(cache_path, "r")) {
      char line[256], key[128];
      int l, t, ch;
      double ms;
      while (std::fgets(line, sizeof line, f)) {
        const int n = std::sscanf(line, "%127s %d %d %lf %d", key, &l, &t, &ms, &ch);
        if (n >= 4 && (t == 1 || t == 2 || t == 4))
          decided[key] = Shape{l, t, n == 5 && (ch == 2 || ch == 4) ? ch : kChunk};
      }
      std::fclose(f);
    }
  }
  // shape classes, largest share of the work first
  std::map<std::string, std::vector<int>> cls;
  std::map<std::string, double> work;
  for (int k = 0; k < (int)b->chains.size(); ++k) {
    const Chain &c = b->chains[k];
    const std::string key = shape_key(c);
    cls[key].push_back(k);
    work[key] += (double)c.hm.N * c.R * c.hm.zmax;
  }
  std::vector<std::string> order;
  for (auto &kv : cls) order.push_back(kv.first);
  std::sort(order.begin(), order.end(), [&](const std::string &x, const std::string &y) {
    return work[x] > work[y];
  });
  auto apply = [&](const std::string &key, const Shape &sh) {
    for (int k : cls[key]) {
      Chain &c = b->chains[k];
      c.lds = sh.lds && c.lds_fits;
      c.tps = sh.tps;
      c.chunk = sh.chunk;
    }
  };
  std::vector<std::string> todo;
  for (auto &key : order) {
    auto it = decided.find(key);
    const Chain &c0 = b->chains[cls[key][0]];
    if (it != decided.end() && (!it->second.lds || c0.lds_fits) && it->second.tps <= std::max(1, c0.hm.zmax)) {
      apply(key, it->second);
    } else {
      decided[key] = Shape{};
      apply(key, decided[key]);
      todo.push_back(key);
    }
  }
  std::map<std::string, double> round_ms;
  if (!todo.empty()) {
    double base = 0;
    int rc = time_rounds(b, nrounds, &base);
    if (rc) return rc;
    for (auto &key : todo) {
      const Chain &c0 = b->chains[cls[key][0]];
      std::vector<Shape> cand;
      for (int tps = 1; tps <= 4; tps *= 2) {
        if (tps > 1 && tps > c0.hm.zmax) break;
        for (int ch = 2; ch <= 4; ch += 2)
          if (tps > 1 || ch != kChunk) cand.push_back(Shape{0, tps, ch});
        if (c0.lds_fits) cand.push_back(Shape{1, tps, kChunk});
      }
      for (const Shape &sh : cand) {
        const Shape prev = decided[key];
        apply(key, sh);
        double t = 0;
        if ((rc = time_rounds(b, nrounds, &t))) return rc;
        if (t < 0.995 * base) {
          base = t;
          decided[key] = sh;
        } else {
          apply(key, prev);
        }
      }
      round_ms[key] = base;
    }
  }
  if (cache_path) {
    if (FILE *f = std::fopen(cache_path, "w")) {
      for (auto &kv : decided)
        std::fprintf(f, "%s %d %d %.6f %d\n", kv.first.c_str(), kv.second.lds, kv.second.tps,
                     round_ms.count(kv.first) ? round_ms[kv.first] : 0.0, kv.second.chunk);
      std::fclose(f);
    }
  }
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->gkey.clear();
  BCHK(b, hipStreamSynchronize(b->stream));
  return ESP_OK;
}
}  // namespace

extern "C" {

// HIP-event timing of the batched colour phases (the dominant kernel): nsweeps sweeps of the
// global-memory chains, one event pair recorded by the dispatch packets around every launch.
int esp_batch_time_sweeps(esp_batch *b, int nsweeps, int mode, uint64_t seed, uint64_t sweep0,
                          double *kernel_ms, int *nlaunch, double *wall_ms, int64_t *updates) {
  if (!b || !kernel_ms || !nlaunch) return ESP_EINVAL;
  if (!b->committed) return bfail(b, ESP_ESTATE, "batch not committed");
  if (nsweeps < 1 || (mode != ESP_MODE_RANDOM && mode != ESP_MODE_CONST))
    return bfail(b, ESP_EINVAL, "bad esp_batch_time_sweeps arguments");
  BCHK(b, hipSetDevice(b->device));
  int rc = build_plan(b, 0.125, seed);
  if (rc) return rc;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->gkey.clear();
  Clock h{sweep0, 0, 1, 0};
  BCHK(b, hipMemcpyAsync(b->d_clock, &h, sizeof(Clock), hipMemcpyHostToDevice, b->stream));
  int per = 0;
  for (int col = 0; col < b->maxcol; ++col)
    for (auto &v : b->vars) per += v.ntiles[col] > 0;
  if (per == 0) return bfail(b, ESP_ESTATE, "no global-memory chains to time");
  std::vector<hipEvent_t> ev(2 * (size_t)per * nsweeps, nullptr);
  for (auto &e : ev) BCHK(b, hipEventCreate(&e));
  int k = 0;
  for (int s = 0; s < nsweeps; ++s)
    for (int col = 0; col < b->maxcol; ++col)
      for (auto &v : b->vars)
        if (v.ntiles[col] > 0) {
          hipExtLaunchKernelGGL(bphase_fn(mode, v.variant), dim3(v.ntiles[col]), dim3(kBlock), v.shm_phase,
                                b->stream, ev[2 * k], ev[2 * k + 1], 0, (const KArgs *)b->d_chains,
                                (const int4 *)v.d_tiles[col], col, s);
          ++k;
        }
  BCHK(b, hipGetLastError());
  BCHK(b, hipStreamSynchronize(b->stream));
  double sum = 0.0;
  for (int q = 0; q < k; ++q) {
    float ms = 0.f;
    BCHK(b, hipEventElapsedTime(&ms, ev[2 * q], ev[2 * q + 1]));
    sum += ms;
  }
  float wall = 0.f;
  BCHK(b, hipEventElapsedTime(&wall, ev[0], ev[2 * k - 1]));
  for (auto &e : ev) (void)hipEventDestroy(e);
  int64_t upd = 0;
  for (auto &c : b->chains)
    if (!c.lds) upd += (int64_t)c.hm.N * c.R;
  *kernel_ms = sum;
  *nlaunch = k;
  if (wall_ms) *wall_ms = wall;
  if (updates) *updates = upd * nsweeps;
  return ESP_OK;
}

int esp_batch_num_chains(esp_batch *b) { return b ? (int)b->chains.size() : ESP_EINVAL; }

int esp_batch_info(esp_batch *b, int k, int *out) {
  if (!b || !out || k < 0 || k >= (int)b->chains.size()) return ESP_EINVAL;
  const Chain &c = b->chains[k];
  out[0] = c.hm.N; out[1] = c.R; out[2] = c.hm.ncol; out[3] = c.tps; out[4] = c.lds;
  out[5] = c.hm.zmax; out[6] = c.variant; out[7] = c.lds_fits;
  return ESP_OK;
}

int esp_batch_get_stats(esp_batch *b, int k, double *acc, int64_t *rejected, double *e_run,
                        int64_t *nmeas) {
  int rc = check_chain(b, k);
  if (rc) return rc;
  Chain &c = b->chains[k];
  if (e_run) {
    // fresh running energies: partial sums over the replica, e_base + sum w
    KArgs a = chain_args(c, k);
    hipLaunchKernelGGL(k_reduce, dim3(c.red_blocks, c.R), dim3(kBlock), 0, b->stream,
                       (const double *)a.sp, (const double *)a.w, a.N, a.Ns, c.red_blocks, c.partial,
                       (const double *)nullptr, c.hm.ncol * c.mpart_nw);
    FinArgs f{};
    f.partial = c.partial; f.nblk = c.red_blocks; f.N = c.hm.N; f.R = c.R;
    f.e_base = c.ebase; f.e_run = c.erun; f.acc = c.acc; f.nmeas = c.nmeas;
    f.slot_rep = c.slot_rep; f.rep_slot = c.rep_slot; f.beta_slot = c.beta; f.nswap = c.nswap;
    f.clock = b->d_clock;
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, b->stream, f);
    BCHK(b, hipGetLastError());
    BCHK(b, hipMemcpyAsync(e_run, c.erun, sizeof(double) * c.R, hipMemcpyDeviceToHost, b->stream));
  }
  if (acc) BCHK(b, hipMemcpyAsync(acc, c.acc, sizeof(double) * 6 * c.R, hipMemcpyDeviceToHost, b->stream));
  std::vector<int64_t> rsub(rejected ? (size_t)kRejSub * c.R : 0);
  if (rejected)
    BCHK(b, hipMemcpyAsync(rsub.data(), c.rej, sizeof(int64_t) * rsub.size(), hipMemcpyDeviceToHost, b->stream));
  if (nmeas) BCHK(b, hipMemcpyAsync(nmeas, c.nmeas, sizeof(int64_t), hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  if (rejected) sum_rej_sub(rsub, c.R, rejected);
  return ESP_OK;
}

int esp_batch_get_map(esp_batch *b, int k, int *slot_rep, int64_t *nswap) {
  int rc = check_chain(b, k);
  if (rc) return rc;
  Chain &c = b->chains[k];
  std::vector<int> sr(c.R);
  BCHK(b, hipMemcpyAsync(sr.data(), c.slot_rep, sizeof(int) * c.R, hipMemcpyDeviceToHost, b->stream));
  if (nswap) BCHK(b, hipMemcpyAsync(nswap, c.nswap, sizeof(int64_t) * 2 * c.R, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  if (slot_rep)
    for (int t = 0; t < c.R; ++t) slot_rep[t] = sr[t] + 1;
  return ESP_OK;
}

int esp_batch_get_spins(esp_batch *b, int k, int by_slot, double *spins) {
  int rc = check_chain(b, k);
  if (rc) return rc;
  if (!spins) return bfail(b, ESP_EINVAL, "spins is NULL");
  Chain &c = b->chains[k];
  const int N = c.hm.N, R = c.R;
  double *d_out = nullptr;
  if ((rc = balloc(b, d_out, (size_t)3 * N * R))) return rc;
  const size_t n = (size_t)N * R;
  hipLaunchKernelGGL(k_get_spins, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, b->stream,
                     (const double *)c.sp, (const int *)c.ref2dev,
                     by_slot ? (const int *)c.slot_rep : (const int *)nullptr, N, c.Ns, R, d_out);
  BCHK(b, hipGetLastError());
  BCHK(b, hipMemcpyAsync(spins, d_out, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  bfree(d_out);
  return ESP_OK;
}

int esp_batch_energy(esp_batch *b, int k, double *E) {
  int rc = check_chain(b, k);
  if (rc) return rc;
  if (!E) return bfail(b, ESP_EINVAL, "E is NULL");
  Chain &c = b->chains[k];
  double *d_e = nullptr;
  if ((rc = balloc(b, d_e, c.R))) return rc;
  if ((rc = chain_energy(b, k, d_e))) return rc;
  BCHK(b, hipMemcpyAsync(E, d_e, sizeof(double) * c.R, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  bfree(d_e);
  return ESP_OK;
}

int esp_batch_reset_stats(esp_batch *b) {
  if (!b) return ESP_EINVAL;
  if (!b->committed) return bfail(b, ESP_ESTATE, "batch not committed");
  BCHK(b, hipSetDevice(b->device));
  for (auto &c : b->chains) {
    BCHK(b, hipMemsetAsync(c.acc, 0, sizeof(double) * 6 * c.R, b->stream));
    BCHK(b, hipMemsetAsync(c.rej, 0, sizeof(unsigned long long) * kRejSub * c.R, b->stream));
    BCHK(b, hipMemsetAsync(c.nswap, 0, sizeof(unsigned long long) * 2 * c.R, b->stream));
    BCHK(b, hipMemsetAsync(c.nmeas, 0, sizeof(unsigned long long), b->stream));
  }
  BCHK(b, hipStreamSynchronize(b->stream));
  return ESP_OK;
}

int esp_batch_sync(esp_batch *b) {
  if (!b) return ESP_EINVAL;
  BCHK(b, hipStreamSynchronize(b->stream));
  BCHK(b, hipStreamSynchronize(b->side));
  return ESP_OK;
}

}  // extern "C"

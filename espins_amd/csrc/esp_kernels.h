// esp_kernels.h -- sweep, measurement, exchange and energy kernels (gfx950, wave64).
//
// Device layout of one replica r (N sites, colour-major device order, stride Ns > N):
//   x[Ns], y[Ns], z[Ns]  spin components (structure of arrays: a gather of a partner spin
//                     is three 8-byte loads whose addresses are contiguous across a wave);
//                     index N holds a zero spin that padded bond slots point at
//   wpart[ncol][nw]   sum of the accepted dE of the sites of one wave of one colour phase
//                     since the last rebase, so the running energy ("by adding differences",
//                     mcarlo.F90:748) is e_base[r] + sum wpart, reproducibly (one writer per
//                     entry per phase, phases ordered by launches)
// Bonds: nbr[zmax][N] device partner indices (self-padded); coefficients per (basis, slot)
// in LDS when the lattice is periodic, else cJ/cB[zmax][N], cD[zmax][N][3].
#pragma once
#include <cstdint>
#include <vector>

#include "esp_device.h"
#include "esp_geo.h"

namespace espk {
struct Clock {                   // device-side schedule counters (graph-replayable)
  unsigned long long sweep;      // Philox sweep counter of the next sweep
  unsigned long long step;       // exchange step counter
  int parity;                    // next exchange start (1 or 2)
  int pad;
};

namespace {  // internal linkage: each translation unit that launches these owns its copies
using namespace espd;

constexpr int kBlock = 256;      // phase / reduce kernels (4 waves)
constexpr int kLdsBlock = 512;   // LDS-resident kernel (8 waves per replica)
// Tuning constants (compile-time; the defaults are the measured best on MI355X).  Fewer
// gathers in flight per lane means fewer registers and more waves per SIMD, which hides the
// gather latency better than a wide chunk does.
#ifndef ESP_CHUNK
#define ESP_CHUNK 2
#endif
#ifndef ESP_CHUNK_WIDE
#define ESP_CHUNK_WIDE 4
#endif
#ifndef ESP_PHASE_WAVES
#define ESP_PHASE_WAVES 0        // > 0: ask the compiler for this many waves per SIMD
#endif
#if ESP_PHASE_WAVES > 0
#define ESP_PHASE_ATTR __attribute__((amdgpu_waves_per_eu(ESP_PHASE_WAVES)))
#else
#define ESP_PHASE_ATTR
#endif
constexpr int kChunk = ESP_CHUNK;  // partner gathers issued together per lane
constexpr int kChunkWide = ESP_CHUNK_WIDE;  // ... for launches too small to fill the GPU with waves
constexpr int kChunkLds = 8;     // ... in the LDS-resident sweep
#ifndef ESP_CHUNK_LEARN
#define ESP_CHUNK_LEARN 8
#endif
constexpr int kChunkLearn = ESP_CHUNK_LEARN;  // ... in the learned sweep
// Rejections are counted per slot in kRejSub sub-counters, rej[slot][kRejSub]: a wave adds
// to the sub-counter its global wave index selects, so the thousands of waves of one replica
// do not queue at one address (a single counter per slot cost 12-14 % of a colour phase).
// Readers sum the sub-counters.
constexpr int kRejSub = 16;
// Dynamic LDS of the LDS-resident kernel: [N+1 spins as double4][coefficient table]
// [kLdsRed doubles of reduction scratch]; every byte of LDS it uses is in that one region.
constexpr int kLdsRed = 4 * (512 / 64);

struct KArgs {
  double *sp;                    // [R][3][Ns]
  double *w;                     // [R][ncol][mpart_nw] per-wave accepted-dE sums
  const int *nbr;
  const double *cJ, *cB, *cD;
  const double *ctab;            // [nb][zmax][5]
  const double *mag;
  const uint8_t *basis;
  const int *dev2ref;
  const int *col_off;
  const double *si4, *h4;
  const double *beta_slot;
  const int *rep_slot;
  unsigned long long *rej;
  const Clock *clock;
  const double *prop, *unif;     // host-fed mode, reference order
  uint8_t *accepted;
  int N, Ns, nb, zmax, R, ncol;
  int c_begin, c_count, nblk;
  int has_si, has_h, xcd_remap;
  int sweep_off, tps;            // tps: lanes cooperating on one site (1, 2, 4, 8)
  double tilt;
  uint32_t key0, key1, stream_base;
  // per-wave sums of the spins after each colour phase, or null: mpart[R][ncol][mpart_nw][3];
  // phase `col` overwrites its waves' entries
  double *mpart;
  int col, mpart_nw;
  // implicit geometry (esp_geo.h), used by the global-memory colour phases when set
  espg::Geo geo;
  const int *grp;                // [ngroups]
  const int *gslot;              // [ngroups][zmax][2]
  const double *magb;            // [nb]
  int rows_max;                  // group rows one block stages in LDS
};

enum { M_RANDOM = 0, M_CONST = 1, M_HOSTFED = 2 };

// Global-memory access through address space 1.  A pointer that a kernel reads out of a
// struct in memory (the batched kernels' per-chain arguments) is a generic pointer to the
// compiler, which then emits flat instructions; a flat load also counts against the LDS wait
// counter, so every LDS read after it waits for all gathers in flight.  Accesses through these
// helpers are global loads, stores and atomics.
#define ESP_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T gldi(const T *p, size_t i) {
  return ((const ESP_GLOBAL T *)p)[i];
}
template <class T>
__device__ __forceinline__ void gsti(T *p, size_t i, T v) {
  ((ESP_GLOBAL T *)p)[i] = v;
}
template <class T>
__device__ __forceinline__ void gadd(T *p, T v) {     // no-return atomic add, device scope
  __hip_atomic_fetch_add((ESP_GLOBAL T *)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 32-bit byte offset from a wave-uniform base: lets the compiler use the SGPR-base
// addressing form (one VGPR offset per lane, no 64-bit address arithmetic).  The host keeps
// every table and replica below 4 GiB.
template <class T>
__device__ __forceinline__ T ld_off(const T *base, unsigned byte_off) {
  return *(const ESP_GLOBAL T *)((const ESP_GLOBAL char *)base + byte_off);
}

struct GSpins {                  // one replica in global memory
  double *x, *y, *z;
  __device__ __forceinline__ V3 ld(unsigned j) const {
    const unsigned o = j * 8u;
    return {ld_off(x, o), ld_off(y, o), ld_off(z, o)};
  }
  __device__ __forceinline__ void st(unsigned j, const V3 &v) const {
    gsti(x, j, v.x); gsti(y, j, v.y); gsti(z, j, v.z);
  }
};
struct LSpins {                  // one replica in LDS (x, y, z, w)
  double4 *p;
  __device__ __forceinline__ V3 ld(unsigned j) const { const double4 t = p[j]; return {t.x, t.y, t.z}; }
};

struct Proposal {
  V3 so, sn;
  double u3;
};

// What the update of one site needs besides the spins: device index, reference index (the
// Philox counter and the host-fed arrays use it), basis atom and moment.  With the implicit
// geometry they come from the device index and the group table in LDS (s_grp holds the rows
// of groups g0..), else from the per-site tables.
struct Site {
  int d, ir, a;
  double mag;
  espg::SiteGeo sg;
};

template <bool IMP>
__device__ __forceinline__ Site site_info(const KArgs &a, int d, const int *s_grp, int g0) {
  Site s;
  s.d = d;
  if (IMP) {
    s.sg = espg::geo_site(a.geo, s_grp - g0, d);
    s.ir = s.sg.ir;
    s.a = s.sg.a;
    s.mag = gldi(a.magb, s.a);
  } else {
    s.ir = gldi(a.dev2ref, d);
    s.a = gldi(a.basis, d);
    s.mag = gldi(a.mag, d);
  }
  return s;
}

// Proposal of one attempted update (identical on every lane of the site's group).
template <int MODE, class SP>
__device__ __forceinline__ Proposal propose(const KArgs &a, const SP &sp, const Site &st, int r,
                                            unsigned long long sweep) {
  Proposal p;
  p.so = sp.ld(st.d);
  const int ir = st.ir;
  if (MODE == M_HOSTFED) {
    const size_t q = 3 * ((size_t)r * a.N + ir);
    p.sn = {gldi(a.prop, q), gldi(a.prop, q + 1), gldi(a.prop, q + 2)};
    p.u3 = gldi(a.unif, (size_t)r * a.N + ir);
  } else {
    const uint4 rn = philox10(make_uint4((uint32_t)ir, (uint32_t)sweep, (uint32_t)(sweep >> 32),
                                         a.stream_base + (uint32_t)r),
                              a.key0, a.key1);
    p.sn = MODE == M_RANDOM ? propose_random(st.mag, u01(rn.x), u01(rn.y))
                            : propose_const(st.mag, p.so, a.tilt, u01(rn.x), u01(rn.y));
    p.u3 = u01(rn.z);
  }
  return p;
}

// One bond's contribution to the J, B, D sums of delta_energy_get (get_quant.F90:288-336).
template <bool CMP, bool HB, bool HD>
__device__ __forceinline__ void bond_term(const KArgs &a, const double *crow, int l, int d,
                                          const V3 &ds, const V3 &si, const V3 &sj, double &eJ,
                                          double &eB, double &eD) {
  double cj, cb = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
  if (CMP) {
    cj = crow[5 * l];
    if (HB) cb = crow[5 * l + 1];
    if (HD) { cx = crow[5 * l + 2]; cy = crow[5 * l + 3]; cz = crow[5 * l + 4]; }
  } else {
    cj = gldi(a.cJ, (size_t)l * a.N + d);
    if (HB) cb = gldi(a.cB, (size_t)l * a.N + d);
    if (HD) {
      const size_t q = 3 * ((size_t)l * a.N + d);
      cx = gldi(a.cD, q); cy = gldi(a.cD, q + 1); cz = gldi(a.cD, q + 2);
    }
  }
  const double dsj = dot(ds, sj);
  eJ = eJ - cj * dsj;
  if (HB) {
    const double sij = dot(si, sj);
    eB = eB + cb * (dsj * dsj + 2.0 * sij * dsj);
  }
  if (HD)  // D . (dS x S_j)  (utility_cross, utility.F90:212-234)
    eD = eD + (cx * (ds.y * sj.z - ds.z * sj.y) + cy * (ds.z * sj.x - ds.x * sj.z) +
               cz * (ds.x * sj.y - ds.y * sj.x));
}

// Partial bond sums over the slots l = g, g+tps, ... of site st.  Inactive lanes return
// zeros.  Slots go in chunks of CH: the CH partner indices of the next chunk are loaded (or,
// with the implicit geometry, computed from the (group, slot) rows staged in LDS at s_gs,
// rows of groups g0..) while the current chunk's CH partner gathers are in flight, so a
// wave waits for one gather round trip per chunk.  Full chunks need no masks; only the last
// partial chunk is masked (its spare lanes re-read slot g).  Padded slots point at the zero
// spin (index N), so they need no test; a self bond (SELF) reads the new spin.
template <int CH, bool CMP, bool HB, bool HD, bool SELF, bool IMP, class SP>
__device__ __forceinline__ void bond_sums(const KArgs &a, const SP &sp, const double *ctab, const Site &st,
                                          const int *s_gs, int g0, bool active, int g,
                                          const Proposal &p, double &eJ, double &eB, double &eD) {
  eJ = 0.0; eB = 0.0; eD = 0.0;
  if (!active) return;
  const int d = st.d;
  const V3 ds = sub(p.sn, p.so);
  const V3 si = sub(p.sn, ds);  // old spin as the reference recovers it (get_quant.F90:306,320)
  const int tps = a.tps;
  const int nk = (a.zmax - g + tps - 1) / tps;
  const int nfull = nk - nk % CH;
  const double *crow = CMP ? ctab + (size_t)st.a * a.zmax * 5 : nullptr;
  const unsigned ud = (unsigned)d;
  const int *grow = IMP ? s_gs + 2 * (st.sg.g - g0) * a.zmax : nullptr;
  unsigned jn[CH];   // partner indices of the next chunk
  auto load_idx = [&](int k0) {
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int l = k0 + u < nk ? g + tps * (k0 + u) : g;
      if (IMP) {
        const int2 e = *reinterpret_cast<const int2 *>(grow + 2 * l);
        jn[u] = (unsigned)espg::geo_partner(a.geo, st.sg, e.x, e.y);
      } else {
        jn[u] = (unsigned)ld_off(a.nbr, ((unsigned)l * (unsigned)a.N + ud) * 4u);
      }
    }
  };
  if (nk > 0) load_idx(0);
  for (int k0 = 0; k0 < nfull; k0 += CH) {
    V3 sv[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const V3 t = sp.ld(jn[u]);
      sv[u] = (SELF && jn[u] == ud) ? p.sn : t;
    }
    if (k0 + CH < nk) load_idx(k0 + CH);
#pragma unroll
    for (int u = 0; u < CH; ++u) bond_term<CMP, HB, HD>(a, crow, g + tps * (k0 + u), d, ds, si, sv[u], eJ, eB, eD);
  }
  const int rem = nk - nfull;
  if (rem > 0) {
    V3 sv[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const V3 t = sp.ld(jn[u]);
      sv[u] = (SELF && jn[u] == ud) ? p.sn : t;
    }
#pragma unroll
    for (int u = 0; u < CH; ++u)
      if (u < rem) bond_term<CMP, HB, HD>(a, crow, g + tps * (nfull + u), d, ds, si, sv[u], eJ, eB, eD);
  }
}

// Sum over the tps lanes of a site's group: xor butterfly, bitwise identical on every lane.
__device__ __forceinline__ double group_sum(double v, int tps) {
  for (int o = 1; o < tps; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ void rej_add(unsigned long long *rej, int slot, unsigned long long n) {
  const unsigned wv = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (kRejSub - 1);
  gadd(&rej[(size_t)slot * kRejSub + wv], n);
}

// Local terms and the accept test of metropolis (get_quant.F90:166-178): reject iff
// dE > 0 and exp(-beta dE) < u3.  Returns dE through *de.
__device__ __forceinline__ bool decide(const KArgs &a, int b, const Proposal &p, double eJ,
                                       double eB, double eD, double beta, double *de_out) {
  const V3 ds = sub(p.sn, p.so);
  const V3 si = sub(p.sn, ds);
  double de = eJ;
  if (a.has_h) {
    const size_t q = 4 * (size_t)b;
    const double h0 = gldi(a.h4, q), h1 = gldi(a.h4, q + 1), h2 = gldi(a.h4, q + 2),
                 h3 = gldi(a.h4, q + 3);
    de = de + h0 * (ds.x * h1 + ds.y * h2 + ds.z * h3);
  }
  if (a.has_si) {
    const size_t q = 4 * (size_t)b;
    const double k0 = gldi(a.si4, q), k1 = gldi(a.si4, q + 1), k2 = gldi(a.si4, q + 2),
                 k3 = gldi(a.si4, q + 3);
    const double sa = si.x * k1 + si.y * k2 + si.z * k3;
    const double dsa = ds.x * k1 + ds.y * k2 + ds.z * k3;
    de = de + k0 * (dsa * dsa + 2.0 * sa * dsa);
  }
  de = de + eB;
  de = de + eD;
  *de_out = de;
  return (de > 0.0) && (exp(-beta * de) < p.u3);
}

__device__ __forceinline__ void decode_block(const KArgs &a, int &r, int &b) {
  const int L = blockIdx.x;
  if (a.xcd_remap) {  // blocks go round-robin over the 8 XCDs: keep a replica on one XCD/L2
    const int x = L & 7, q = L >> 3;
    r = x + 8 * (q / a.nblk);
    b = q % a.nblk;
  } else {
    r = L / a.nblk;
    b = L % a.nblk;
  }
}

// Sum over the 64 lanes of a wave with DPP lane moves (no LDS traffic): pairs, quads, half
// rows, rows, then rows 0+1 into row 1 and rows 0..2 into row 3; lane 63 ends with the total,
// which readlane makes wave-uniform.  Every lane must be active.  The order of the additions
// is fixed, so the result is reproducible.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const int2 h = __builtin_bit_cast(int2, v);
  int2 r;
  r.x = __builtin_amdgcn_update_dpp(0, h.x, CTRL, ROW_MASK, 0xf, false);
  r.y = __builtin_amdgcn_update_dpp(0, h.y, CTRL, ROW_MASK, 0xf, false);
  return __builtin_bit_cast(double, r);
}
__device__ __forceinline__ double wave_total(double v) {
  v += dpp_f64<0xb1, 0xf>(v);    // quad_perm [1,0,3,2]
  v += dpp_f64<0x4e, 0xf>(v);    // quad_perm [2,3,0,1]
  v += dpp_f64<0x141, 0xf>(v);   // row_half_mirror
  v += dpp_f64<0x140, 0xf>(v);   // row_mirror
  v += dpp_f64<0x142, 0xa>(v);   // row_bcast:15 into rows 1, 3
  v += dpp_f64<0x143, 0xc>(v);   // row_bcast:31 into rows 2, 3
  const int2 h = __builtin_bit_cast(int2, v);
  return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_readlane(h.x, 63),
                                              __builtin_amdgcn_readlane(h.y, 63)));
}

// Per-wave sums of one colour phase: the accepted dE (added to the wave's accumulator, which
// the running energy sums) and the spins after the phase (stored; a site changes only in its
// own colour phase, so after a sweep these add up to the replica's magnetisation and the
// measurement reads them instead of every spin).  Fixed butterflies, one writer per entry.
// The dE sum is added with a no-return atomic (one per wave and phase, so the order of the
// additions to an entry is that of the launches): the wave does not wait for a read.
__device__ __forceinline__ void wave_sums(const KArgs &a, int r, int b, bool owner, bool rej, double de,
                                          const Proposal &p, int tps) {
  double w = owner && !rej ? de : 0.0;
  V3 sf = {0.0, 0.0, 0.0};
  if (owner) sf = rej ? p.so : p.sn;
  w = wave_total(w);
  if (a.mpart) {
    sf.x = wave_total(sf.x);
    sf.y = wave_total(sf.y);
    sf.z = wave_total(sf.z);
  }
  // waves past the colour's last lane own no entry (mpart_nw counts the colour's waves)
  const int wave = (b * kBlock + threadIdx.x) >> 6;
  if ((threadIdx.x & 63) == 0 && wave * 64 < a.c_count * tps) {
    const size_t e = ((size_t)r * a.ncol + a.col) * a.mpart_nw + wave;
    if (w != 0.0) gadd(a.w + e, w);
    if (a.mpart) {
      gsti(a.mpart, e * 3, sf.x); gsti(a.mpart, e * 3 + 1, sf.y); gsti(a.mpart, e * 3 + 2, sf.z);
    }
  }
}

// Coefficient layouts of the colour phases: per-bond tables, per-(basis, slot) table in LDS,
// or the per-(basis, slot) table plus the implicit geometry (no per-site tables read).
enum { LAY_BOND = 0, LAY_CMP = 1, LAY_IMP = 2 };

// Dynamic LDS of a colour-phase block: [coefficient table][group rows][group records].
__host__ __device__ inline size_t phase_lds_bytes(int lay, int nb, int zmax, int rows_max) {
  size_t n = lay >= LAY_CMP ? (size_t)nb * zmax * 5 * sizeof(double) : 0;
  if (lay == LAY_IMP) n += (size_t)rows_max * zmax * 2 * sizeof(int) + (size_t)rows_max * sizeof(int);
  return n;
}

// Block b of colour phase (a.c_begin, a.c_count) of replica r (global-memory path);
// a.tps lanes per site.  Shared by the single-context and the batched kernels.
template <int MODE, int LAY, bool HB, bool HD, bool SELF, int CH = kChunk>
__device__ __forceinline__ void phase_body(const KArgs &a, int r, int b, double *s_ctab) {
  constexpr bool CMP = LAY >= LAY_CMP, IMP = LAY == LAY_IMP;
  const int tps = a.tps;
  const int per = kBlock / tps;                        // sites of this block
  int g0 = 0;
  int *s_gs = nullptr, *s_grp = nullptr;
  if (CMP) {
    const int n = a.nb * a.zmax * 5;
    for (int k = threadIdx.x; k < n; k += kBlock) s_ctab[k] = gldi(a.ctab, k);
  }
  if (IMP) {   // (group, slot) rows and group records of the groups this block's sites lie in
    const int first = a.c_begin + b * per;
    const int last = min(a.c_begin + a.c_count, first + per) - 1;
    const int mt = (int)a.geo.mt.d;
    g0 = first / mt;
    const int nrow = last / mt - g0 + 1;
    s_gs = reinterpret_cast<int *>(s_ctab + a.nb * a.zmax * 5);
    s_grp = s_gs + 2 * a.rows_max * a.zmax;
    const int n = 2 * nrow * a.zmax;
    const int *src = a.gslot + (size_t)2 * g0 * a.zmax;
    for (int k = threadIdx.x; k < n; k += kBlock) s_gs[k] = gldi(src, k);
    for (int k = threadIdx.x; k < nrow; k += kBlock) s_grp[k] = gldi(a.grp, g0 + k);
  }
  if (CMP) __syncthreads();
  const int g = threadIdx.x & (tps - 1);
  const int k = (b * kBlock + threadIdx.x) / tps;
  const bool active = k < a.c_count;
  const int d = a.c_begin + (active ? k : 0);
  const int slot = gldi(a.rep_slot, r);
  const double beta = gldi(a.beta_slot, slot);
  const unsigned long long sweep = gldi(&a.clock->sweep, 0) + (unsigned long long)a.sweep_off;
  double *base = a.sp + (size_t)r * 3 * a.Ns;
  const GSpins sp{base, base + a.Ns, base + 2 * a.Ns};
  // inactive lanes of a block sit on its first site, which lies in row g0
  const Site st = site_info<IMP>(a, active ? d : a.c_begin + b * per, s_grp, g0);
  Proposal p{};
  if (active) p = propose<MODE>(a, sp, st, r, sweep);
  double eJ, eB, eD;
  bond_sums<CH, CMP, HB, HD, SELF, IMP>(a, sp, s_ctab, st, s_gs, g0, active, g, p, eJ, eB, eD);
  eJ = group_sum(eJ, tps);
  if (HB) eB = group_sum(eB, tps);
  if (HD) eD = group_sum(eD, tps);
  bool rej = false;
  double de = 0.0;
  if (active) {
    rej = decide(a, st.a, p, eJ, eB, eD, beta, &de);
    if (g == 0) {
      if (MODE == M_HOSTFED) gsti(a.accepted, (size_t)r * a.N + st.ir, (uint8_t)(rej ? 0 : 1));
      if (!rej) sp.st(d, p.sn);
    }
  }
  const unsigned long long m = __ballot(rej && g == 0);
  if ((threadIdx.x & 63) == 0 && m) rej_add(a.rej, slot, (unsigned long long)__popcll(m));
  wave_sums(a, r, b, active && g == 0, rej, de, p, tps);
}

// One colour phase over every replica of one context.
template <int MODE, int LAY, bool HB, bool HD, bool SELF, int CH>
__global__ __launch_bounds__(kBlock) ESP_PHASE_ATTR void k_phase(KArgs a) {
  extern __shared__ double s_ctab[];
  int r, b;
  decode_block(a, r, b);
  if (r >= a.R) return;  // uniform per block
  phase_body<MODE, LAY, HB, HD, SELF, CH>(a, r, b, s_ctab);
}

struct LdsMeas {
  double *acc;                   // [R][6] by slot
  double *e_run;                 // [R] by replica
  double *e_base;                // [R]: the launch's accepted dE is added at its end
  unsigned long long *nmeas;
  int measure_every;             // 0 = never
};

// All colours of nsweeps sweeps of one replica per block, spins resident in LDS,
// measurement (mcarlo.F90:536-543) fused every measure_every sweeps.
template <int MODE, bool CMP, bool HB, bool HD, bool SELF>
__device__ __forceinline__ void lds_body(const KArgs &a, int r, int nsweeps, const LdsMeas &m,
                                         double4 *s_sp) {
  double *ctab = reinterpret_cast<double *>(s_sp + a.N + 1);
  double *red = ctab + (CMP ? a.nb * a.zmax * 5 : 0);
  const double *gx = a.sp + (size_t)r * 3 * a.Ns;
  // w of a site in LDS: its accepted dE during this launch
  for (int d = threadIdx.x; d <= a.N; d += kLdsBlock)   // index N: the zero spin
    s_sp[d] = d < a.N ? make_double4(gldi(gx, d), gldi(gx, a.Ns + d), gldi(gx, 2 * (size_t)a.Ns + d), 0.0)
                      : make_double4(0.0, 0.0, 0.0, 0.0);
  if (CMP)
    for (int k = threadIdx.x; k < a.nb * a.zmax * 5; k += kLdsBlock) ctab[k] = gldi(a.ctab, k);
  // accepted dE the global-memory phases accumulated since the last rebase
  double wg[1] = {0.0};
  {
    const int ne = a.ncol * a.mpart_nw;
    const double *wr = a.w + (size_t)r * ne;
    for (int k = threadIdx.x; k < ne; k += kLdsBlock) wg[0] += gldi(wr, k);
  }
  block_sum<1>(wg, red);     // (its barriers also order the LDS fills above)
  const LSpins sp{s_sp};
  const int slot = gldi(a.rep_slot, r);
  const double beta = gldi(a.beta_slot, slot);
  const unsigned long long sweep0 = gldi(&a.clock->sweep, 0);
  const int tps = a.tps;
  const int g = threadIdx.x & (tps - 1);
  const int per = kLdsBlock / tps;
  unsigned long long nrej = 0;   // 64-bit: one launch may run many sweeps at high rejection
  // The estimator sums of this replica's slot live in thread 0's registers during the
  // launch (same additions, same order as adding to memory each time) and are stored once
  // at the end: the sweep loop itself writes nothing but LDS.
  double A[6] = {0, 0, 0, 0, 0, 0};
  double e_last = 0.0;
  unsigned nmeas_local = 0;
  if (threadIdx.x == 0 && m.measure_every > 0)
    for (int k = 0; k < 6; ++k) A[k] = gldi(m.acc, 6 * slot + k);
  const double e_base = gldi(m.e_base, r);
  double w_launch = 0.0;      // thread 0: the launch's accepted dE at the last block sum
  for (int sw = 0; sw < nsweeps; ++sw) {
    const unsigned long long sweep = sweep0 + (unsigned long long)sw;
    for (int c = 0; c < a.ncol; ++c) {
      const int beg = gldi(a.col_off, c), end = gldi(a.col_off, c + 1);
      for (int base = beg; base < end; base += per) {  // uniform trip count in the block
        const int d0 = base + threadIdx.x / tps;
        const bool active = d0 < end;
        const int d = active ? d0 : beg;
        const Site st = site_info<false>(a, d, nullptr, 0);
        Proposal p{};
        if (active) p = propose<MODE>(a, sp, st, r, sweep);
        double eJ, eB, eD;
        bond_sums<kChunkLds, CMP, HB, HD, SELF, false>(a, sp, ctab, st, nullptr, 0, active, g, p, eJ, eB, eD);
        eJ = group_sum(eJ, tps);
        if (HB) eB = group_sum(eB, tps);
        if (HD) eD = group_sum(eD, tps);
        if (active) {
          double de;
          const bool rej = decide(a, st.a, p, eJ, eB, eD, beta, &de);
          if (g == 0) {
            if (rej) ++nrej;
            else {
              const double4 old = s_sp[d];
              s_sp[d] = make_double4(p.sn.x, p.sn.y, p.sn.z, old.w + de);
            }
          }
        }
      }
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
        const double e = e_base + wg[0] + v[0];
        w_launch = v[0];
        if (meas) {
          const double n = (double)a.N;
          const double mx = v[1] / n, my = v[2] / n, mz = v[3] / n;
          const double m2 = mx * mx + my * my + mz * mz;
          const double e2 = e * e;
          A[0] += e; A[1] += e2; A[2] += e2 * e2;
          A[3] += sqrt(m2); A[4] += m2; A[5] += m2 * m2;
          ++nmeas_local;
        }
        e_last = e;
      }
    }
  }
  if (threadIdx.x == 0) {
    if (nmeas_local) {
      for (int k = 0; k < 6; ++k) gsti(m.acc, 6 * slot + k, A[k]);
      if (r == 0) gsti(m.nmeas, 0, gldi(m.nmeas, 0) + nmeas_local);
    }
    if (nsweeps > 0) {
      gsti(m.e_run, r, e_last);
      gsti(m.e_base, r, e_base + w_launch);   // one block per replica: the only writer
    }
  }
  double *ox = a.sp + (size_t)r * 3 * a.Ns;
  for (int d = threadIdx.x; d < a.N; d += kLdsBlock) {
    const double4 s = s_sp[d];
    gsti(ox, d, s.x); gsti(ox, a.Ns + d, s.y); gsti(ox, 2 * (size_t)a.Ns + d, s.z);
  }
  unsigned long long v = nrej;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((threadIdx.x & 63) == 0 && v) rej_add(a.rej, slot, v);
}

template <int MODE, bool CMP, bool HB, bool HD, bool SELF>
__global__ __launch_bounds__(kLdsBlock) void k_lds(KArgs a, int nsweeps, LdsMeas m) {
  extern __shared__ double4 s_sp[];
  lds_body<MODE, CMP, HB, HD, SELF>(a, blockIdx.x, nsweeps, m, s_sp);
}

// Partial sums per (replica, block) for the running energy and m: the per-wave accepted-dE
// accumulators (ne entries per replica) and either the colour phases' per-wave spin sums
// (mpart, ne entries of 3) or, without them, the spins.
__device__ __forceinline__ void reduce_body(const double *sp, const double *wpart, int N, int Ns, int nblk,
                                            double *partial, const double *mpart, int ne, int r, int b) {
  __shared__ double red[4 * (kBlock / 64)];
  double v[4] = {0, 0, 0, 0};
  const int mc = (ne + nblk - 1) / nblk;
  const double *wr = wpart + (size_t)r * ne;
  for (int q = b * mc + threadIdx.x; q < min(ne, (b + 1) * mc); q += kBlock) v[0] += wr[q];
  if (mpart) {
    const double *mr = mpart + (size_t)r * ne * 3;
    for (int q = b * mc + threadIdx.x; q < min(ne, (b + 1) * mc); q += kBlock) {
      v[1] += mr[3 * q]; v[2] += mr[3 * q + 1]; v[3] += mr[3 * q + 2];
    }
  } else {
    const int chunk = (N + nblk - 1) / nblk;
    const int beg = b * chunk, end = min(N, beg + chunk);
    const double *x = sp + (size_t)r * 3 * Ns;
    for (int d = beg + threadIdx.x; d < end; d += kBlock) {
      v[1] += x[d]; v[2] += x[Ns + d]; v[3] += x[2 * Ns + d];
    }
  }
  block_sum<4>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) partial[((size_t)r * nblk + b) * 4 + k] = v[k];
}

__global__ __launch_bounds__(kBlock) void k_reduce(const double *sp, const double *wpart, int N, int Ns,
                                                   int nblk, double *partial, const double *mpart, int ne) {
  reduce_body(sp, wpart, N, Ns, nblk, partial, mpart, ne, blockIdx.y, blockIdx.x);
}

// e_base[r] += the sum of replica r's per-wave accumulators (fixed order); they are then zero.
// Used before the accumulators are re-laid out.
__global__ void k_fold_w(double *wpart, int ne, int R, double *e_base) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  double *w = wpart + (size_t)r * ne;
  double s = 0.0;
  for (int q = 0; q < ne; ++q) { s += w[q]; w[q] = 0.0; }
  e_base[r] += s;
}

struct FinArgs {
  const double *partial;         // [R][nblk][4] or null (e_run already fresh)
  int nblk, N, R;
  const double *e_base;
  double *e_run;
  double *acc;
  unsigned long long *nmeas;
  int *slot_rep, *rep_slot;
  const double *beta_slot;
  unsigned long long *nswap;     // [R][2]
  Clock *clock;
  int measure, exchange, advance_sweeps;
  uint32_t key0, key1, stream_base;
  double *trace;                 // [trace_cap][R][2] by slot: E, |m| per measurement, or null
  int trace_cap;
};

// Running energies / measurement, then the replica exchange (pt_swap, pt.F90:1235-1288)
// in label form.  One block per context; start parity and step come from the clock values
// the caller read.
__device__ __forceinline__ void finalize_body(const FinArgs &f, int start, unsigned long long step) {
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
        const unsigned long long k = f.nmeas[0];   // incremented after the barrier below
        if (f.trace && k < (unsigned long long)f.trace_cap) {
          double *t = f.trace + 2 * ((size_t)k * f.R + f.rep_slot[r]);
          t[0] = e;
          t[1] = sqrt(m2);
        }
      }
    }
  __syncthreads();
  if (f.exchange) {
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
  if (threadIdx.x == 0 && f.measure) f.nmeas[0] += 1;
}

__device__ __forceinline__ void advance_clock(Clock *c, bool exchange, int advance_sweeps) {
  if (exchange) {
    c->step += 1;
    c->parity = 3 - c->parity;
  }
  c->sweep += (unsigned long long)advance_sweeps;
}

__global__ __launch_bounds__(1024) void k_finalize(FinArgs f) {
  const int start = f.clock->parity;  // 1 or 2 (pt.F90:679,898)
  const unsigned long long step = f.clock->step;
  finalize_body(f, start, step);
  if (threadIdx.x == 0) advance_clock(f.clock, f.exchange, f.advance_sweeps);
}

// k_reduce followed by k_finalize in one launch: every block adds its partial sums
// (reduce_body); the last block to finish -- counted on a device-scope counter that it resets
// -- then runs finalize_body and advances the clock.  Each block's thread 0 releases its
// partials with a device fence before counting; the last block's threads fence (acquire, which
// also drops stale L1 lines) before reading them.
__global__ __launch_bounds__(kBlock) void k_reduce_fin(const double *sp, const double *wpart, int N, int Ns,
                                                       const double *mpart, int ne, double *partial, FinArgs f,
                                                       unsigned *done) {
  reduce_body(sp, wpart, N, Ns, f.nblk, partial, mpart, ne, blockIdx.y, blockIdx.x);
  __shared__ int s_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    s_last = atomicAdd(done, 1u) == gridDim.x * gridDim.y - 1;
  }
  __syncthreads();
  if (!s_last) return;   // uniform per block
  __threadfence();
  const int start = f.clock->parity;  // 1 or 2 (pt.F90:679,898)
  const unsigned long long step = f.clock->step;
  finalize_body(f, start, step);
  if (threadIdx.x == 0) {
    advance_clock(f.clock, f.exchange, f.advance_sweeps);
    *done = 0u;
  }
}

// Direct energy (energy_get, get_quant.F90:186-261): per-site terms, block partials.
template <bool CMP, bool HB, bool HD>
__global__ __launch_bounds__(kBlock) void k_energy(KArgs a, int nblk, double *partial) {
  extern __shared__ double s_ctab[];   // [coefficient table][kBlock/64 scratch]
  double *red = s_ctab + (CMP ? a.nb * a.zmax * 5 : 0);
  const int r = blockIdx.y, bk = blockIdx.x;
  if (CMP) {
    for (int k = threadIdx.x; k < a.nb * a.zmax * 5; k += kBlock) s_ctab[k] = a.ctab[k];
    __syncthreads();
  }
  const int chunk = (a.N + nblk - 1) / nblk;
  const int beg = bk * chunk, end = min(a.N, beg + chunk);
  double *base = a.sp + (size_t)r * 3 * a.Ns;
  const GSpins sp{base, base + a.Ns, base + 2 * a.Ns};
  double v[1] = {0.0};
  for (int d = beg + threadIdx.x; d < end; d += kBlock) {
    const V3 s = sp.ld(d);
    const int b = a.basis[d];
    double e = 0.0;
    for (int l = 0; l < a.zmax; ++l) {
      const unsigned j = (unsigned)a.nbr[(size_t)l * a.N + d];
      const V3 t = sp.ld(j);
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

// Spins in: reference order [R][N][3] -> device order SoA.
__global__ void k_put_spins(const double *in, const int *ref2dev, int N, int Ns, int R, double *sp) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * R) return;
  const int r = (int)(t / N), i = (int)(t % N);
  const double *p = in + 3 * t;
  const int d = ref2dev[i];
  double *x = sp + (size_t)r * 3 * Ns;
  x[d] = p[0]; x[Ns + d] = p[1]; x[2 * Ns + d] = p[2];
}

__global__ void k_get_spins(const double *sp, const int *ref2dev, const int *src_rep, int N, int Ns,
                            int R, double *out) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)N * R) return;
  const int q = (int)(t / N), i = (int)(t % N);
  const int r = src_rep ? src_rep[q] : q;
  const double *x = sp + (size_t)r * 3 * Ns;
  const int d = ref2dev[i];
  double *p = out + 3 * t;
  p[0] = x[d]; p[1] = x[Ns + d]; p[2] = x[2 * Ns + d];
}

// Plain streaming copy (grid-stride, 32 B per lane per iteration): calibrates the HBM
// traffic counters against a kernel whose bytes are known.
__global__ void k_copy(const double4 *src, double4 *dst, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    dst[t] = src[t];
}

__global__ void k_fill(double *p, size_t n, double v) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

// Per-slot totals of the rejection sub-counters rej[slot][kRejSub] (host side).
inline void sum_rej_sub(const std::vector<int64_t> &sub, int R, int64_t *out) {
  for (int s = 0; s < R; ++s) {
    int64_t t = 0;
    for (int k = 0; k < kRejSub; ++k) t += sub[(size_t)s * kRejSub + k];
    out[s] = t;
  }
}

}  // namespace
}  // namespace espk

// esp_geo.h -- implicit lattice geometry: partner indices, reference site indices and basis
// atoms computed from the device index instead of loaded from per-site tables.
//
// It applies to a periodic supercell L0 x L1 x L2 with nb basis atoms whose colouring has a
// period P (a site's colour depends only on its basis atom and its cell modulo P, P_k | L_k).
// The device order (esp_model.cpp) then consists of "groups": for every (basis atom a,
// residue r = cell mod P), in colour order, the M0 x M1 x M2 cells x = s * P + r (M = L / P),
// s in lexicographic order (s0 fastest).  Every group has Mt = M0 M1 M2 sites, so
//     device index d = g * Mt + (s2 * M1 + s1) * M0 + s0.
// The bond in slot l of a site of group g goes to the cell x + delta_l of the basis atom
// b_l, i.e. to group g'(g, l) at sub-cell (s + q(g, l)) mod M with q = floor((r + delta) / P).
// One table row per (group, slot) -- {first device index of g', q packed in bytes} -- gives
// every partner index with integer arithmetic: the per-site neighbour list (4 B per bond of
// HBM traffic) is not read at all.
//
// The same functions run on the host (to verify them against the explicit tables for every
// site and slot before a context uses them) and in the kernels.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ESP_HD __host__ __device__ __forceinline__
#else
#define ESP_HD inline
#endif

namespace espg {

// Unsigned division by a run-time constant D >= 1 for 32-bit numerators: multiply-high,
// add, shift (Granlund & Montgomery 1994, the "round-up" method with a 33-bit multiplier
// split as 2^32 + m).
struct FastDiv {
  uint32_t m, s1, s2, d;
};

ESP_HD uint32_t umulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

inline FastDiv make_fastdiv(uint32_t D) {
  uint32_t l = 0;
  while (l < 32 && (1ull << l) < D) ++l;            // l = ceil(log2 D)
  FastDiv f;
  f.m = (uint32_t)(((1ull << 32) * ((1ull << l) - D)) / D + 1);
  f.s1 = l > 0 ? 1u : 0u;
  f.s2 = l > 0 ? l - 1 : 0u;
  f.d = D;
  return f;
}

ESP_HD uint32_t fdiv(uint32_t n, const FastDiv &f) {
  const uint32_t t = umulhi32(f.m, n);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

// Geometry of one model (host-built, passed by value to the kernels).
struct Geo {
  FastDiv mt, m0, m1;            // Mt, M0, M1
  int M0, M1, M2, P0, P1, P2, L0, L1, nb;
  int ngroups, zmax, N;
};

// Per-group record: basis atom and residue, packed a | r0 << 8 | r1 << 16 | r2 << 24.
ESP_HD int grp_pack(int a, int r0, int r1, int r2) { return a | (r0 << 8) | (r1 << 16) | (r2 << 24); }

// Per-(group, slot) record: x = first device index of the partner group (-1: padding slot,
// which points at the zero spin N); y = the sub-cell shift reduced mod M, q_k in [0, M_k),
// packed as q0 | q1 << 8 | q2 << 16 (M_k <= 255).
ESP_HD int q_pack(int q0, int q1, int q2) { return q0 | (q1 << 8) | (q2 << 16); }

// 24-bit multiply-add (full-rate v_mad_u32_u24 on the device; operands < 2^24).
ESP_HD uint32_t umad24(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul24(a, b) + c;
#else
  return a * b + c;
#endif
}

ESP_HD uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }

struct SiteGeo {
  int g;              // group
  int s0, s1, s2;     // sub-cell
  int a;              // basis atom
  int ir;             // reference site index (0-based): ((x2 L1 + x1) L0 + x0) nb + a
};

ESP_HD SiteGeo geo_site(const Geo &G, const int *grp, int d) {
  SiteGeo s;
  const uint32_t ud = (uint32_t)d;
  const uint32_t g = fdiv(ud, G.mt);
  const uint32_t k = ud - g * (uint32_t)G.mt.d;
  const uint32_t t = fdiv(k, G.m0);
  s.g = (int)g;
  s.s0 = (int)(k - t * (uint32_t)G.M0);
  const uint32_t u = fdiv(t, G.m1);
  s.s1 = (int)(t - u * (uint32_t)G.M1);
  s.s2 = (int)u;
  const int e = grp[g];
  s.a = e & 0xff;
  const int x0 = s.s0 * G.P0 + ((e >> 8) & 0xff);
  const int x1 = s.s1 * G.P1 + ((e >> 16) & 0xff);
  const int x2 = s.s2 * G.P2 + ((e >> 24) & 0xff);
  s.ir = ((x2 * G.L1 + x1) * G.L0 + x0) * G.nb + s.a;
  return s;
}

// (s + q) mod M for s, q in [0, M): t = s + q < 2M, and t - M wraps to a huge unsigned value
// when t < M, so the unsigned minimum is the reduced sum (three VALU instructions with the
// byte extraction folded into the add).
ESP_HD uint32_t wrap_add(uint32_t s, uint32_t q, uint32_t M) {
  const uint32_t t = s + q;
  return umin32(t, t - M);
}

// Device index of the partner of site s through the (group, slot) record (pg, qp), computed
// without branches: padding slots (pg < 0) select the zero spin N.
ESP_HD int geo_partner(const Geo &G, const SiteGeo &s, int pg, int qp) {
  const uint32_t uq = (uint32_t)qp;
  const uint32_t v0 = wrap_add((uint32_t)s.s0, uq & 0xffu, (uint32_t)G.M0);
  const uint32_t v1 = wrap_add((uint32_t)s.s1, (uq >> 8) & 0xffu, (uint32_t)G.M1);
  const uint32_t v2 = wrap_add((uint32_t)s.s2, (uq >> 16) & 0xffu, (uint32_t)G.M2);
  const uint32_t j = (uint32_t)pg + umad24(umad24(v2, (uint32_t)G.M1, v1), (uint32_t)G.M0, v0);
  return pg < 0 ? G.N : (int)j;
}

}  // namespace espg

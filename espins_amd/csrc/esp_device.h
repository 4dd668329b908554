// esp_device.h -- device building blocks shared by the sweep, energy and batch kernels
// (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace espd {

// Philox4x32-10 (Salmon et al., SC'11): counter-based, so no RNG state is stored; every
// attempted update draws from the counter (site, sweep, stream) under the run's key.
__device__ __forceinline__ uint4 philox10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
  }
  return c;
}

__device__ __forceinline__ double u01(uint32_t x) { return (double)x * 2.3283064365386962890625e-10; }

struct V3 { double x, y, z; };
__device__ __forceinline__ double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 sub(const V3 &a, const V3 &b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 xyz(const double4 &s) { return {s.x, s.y, s.z}; }

// 'random' proposal (get_quant.F90:80-93): cos(theta) = 2u1-1, phi = 2 pi u2.
__device__ __forceinline__ V3 propose_random(double mag, double u1, double u2) {
  const double ct = 2.0 * u1 - 1.0;
  const double st = sqrt(1.0 - ct * ct);
  double sp, cp;
  sincospi(2.0 * u2, &sp, &cp);
  return {mag * st * cp, mag * st * sp, mag * ct};
}

// 'const' proposal (get_quant.F90:94-161): a cone of half-angle tilt*acos(2u1-1) around
// the renormalised old spin.
__device__ __forceinline__ V3 propose_const(double mag, const V3 &o, double tilt, double u1, double u2) {
  const double theta = tilt * acos(2.0 * u1 - 1.0);
  double sp, cp, sth, cth;
  sincospi(2.0 * u2, &sp, &cp);
  sincos(theta, &sth, &cth);
  const double px = sth * cp, py = sth * sp, pz = cth;
  const double sxy = sqrt(o.x * o.x + o.y * o.y);
  double cphi = 1.0, sphi = 0.0;
  if (sxy > 1.0e-8) { cphi = o.x / sxy; sphi = o.y / sxy; }
  const double inv = 1.0 / sqrt(o.x * o.x + o.y * o.y + o.z * o.z);
  const double nx = inv * o.x, ny = inv * o.y, nz = inv * o.z;
  return {mag * (cphi * nz * px - sphi * py + nx * pz),
          mag * (sphi * nz * px + cphi * py + ny * pz),
          mag * (-sqrt(nx * nx + ny * ny) * px + nz * pz)};
}

// Deterministic block sum of NV doubles (fixed shuffle tree inside each wave, then waves
// in order).  All threads must call it; the result is valid on thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *scratch /* NV*nwaves */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k)
    for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
  if (lane == 0)
    for (int k = 0; k < NV; ++k) scratch[wid * NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; ++k) {
      double s = scratch[k];
      for (int w = 1; w < nw; ++w) s += scratch[w * NV + k];
      v[k] = s;
    }
  __syncthreads();
}

}  // namespace espd

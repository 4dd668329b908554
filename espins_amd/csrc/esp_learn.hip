// esp_learn.hip -- learned proposal for the Metropolis-Hastings sweep (gfx950).
//
// The energy of site i as a function of its own spin S (every other spin fixed) is
//     E_i(S) = -G.S + S^T Q S
// with G = sum_j J_ij S_j - sum_j S_j x D_ij - h a_h   (exchange, DM, Zeeman) and
//      Q = sum_j B_ij S_j S_j^T + Delta a a^T           (biquadratic, single ion),
// the same terms energy_get / delta_energy_get evaluate (get_quant.F90:186-337).  A small MLP
// maps five features of (G, Q, T) to a two-component von Mises-Fisher mixture q(S | env);
// moves are accepted with min(1, exp(-beta dE) q(S_old) / q(S_new)).  q depends only on the
// neighbours, which a colour phase never changes, so detailed balance holds for any weights:
// the network only decides how often moves are accepted.
//
// Network: 5 -> 32 -> 32 -> 32 -> 6, GELU (tanh form), BF16 weights for sampling (FP32 accumulate), FP32
// master weights for training with Adam.  Outputs per component k: log-scale s_k of the
// concentration kappa_k = beta|G| exp(s_k), weight logit, tilt t_k = (pi/2) tanh(.) of the
// mean direction from G-hat toward the nearer pole of the anisotropy axis.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "esp_ctx.h"

using namespace espd;
using namespace espk;

namespace {

constexpr int kIn = 5, kH = 32, kOut = 6;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kOut * kH, kNP = oB4 + kOut;
static_assert(kNP == ESP_LEARN_NPARAMS, "parameter count");
constexpr int kRec = 16;   // record: x[5], kappa0, ghat[3], axis[3], s[3], pad
constexpr int kTrainBlock = 128;
constexpr double kPi = 3.14159265358979323846;
// Weight of a fixed uniform-sphere component beside the two network components: the density
// of any direction stays >= kUniform / (4 pi), so a poorly fitted network slows the chain by
// at most a factor kUniform against uniform proposals instead of freezing it.
constexpr double kUniform = 1.0 / 32.0;

// Activation: GELU in its tanh form (Hendrycks & Gimpel 2016),
//   g(x) = x sigma(u),  u = c (x + 0.044715 x^3),  c = 2 sqrt(2 / pi),
// which equals 0.5 x (1 + tanh(u / 2)).  Evaluated as x / (1 + 2^(-u log2 e)) with the
// hardware exp2 and reciprocal: six VALU instructions, two of them transcendental, the same
// FP32 expression in training and in the sweep.  g'(x) = sigma + x sigma (1 - sigma) u'(x).
constexpr float kGeluC = 1.5957691216057308f;             // 2 sqrt(2 / pi)
constexpr float kGeluK = 2.3022081983067571f;             // 2 sqrt(2 / pi) log2(e)
__device__ __forceinline__ float gelu_sigma(float x) {
  const float a = x * fmaf(x * x, 0.044715f * kGeluK, kGeluK);
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-a));
}
__device__ __forceinline__ float act(float x) { return x * gelu_sigma(x); }
__device__ __forceinline__ float act_d(float x) {
  const float sg = gelu_sigma(x);
  return sg + x * sg * (1.0f - sg) * kGeluC * fmaf(x * x, 3.0f * 0.044715f, 1.0f);
}

// MLP forward (weights w in any address space).
__device__ __forceinline__ void mlp_forward(const float *w, const float x[kIn], float out[kOut]) {
  float h1[kH], h2[kH];
#pragma unroll 4
  for (int i = 0; i < kH; ++i) {
    float s = w[oB1 + i];
#pragma unroll
    for (int k = 0; k < kIn; ++k) s += w[oW1 + i * kIn + k] * x[k];
    h1[i] = act(s);
  }
#pragma unroll 2
  for (int i = 0; i < kH; ++i) {
    float s = w[oB2 + i];
#pragma unroll 8
    for (int k = 0; k < kH; ++k) s += w[oW2 + i * kH + k] * h1[k];
    h2[i] = act(s);
  }
#pragma unroll 2
  for (int i = 0; i < kH; ++i) {
    float s = w[oB3 + i];
#pragma unroll 8
    for (int k = 0; k < kH; ++k) s += w[oW3 + i * kH + k] * h2[k];
    h1[i] = act(s);  // h3
  }
#pragma unroll
  for (int i = 0; i < kOut; ++i) {
    float s = w[oB4 + i];
#pragma unroll 8
    for (int k = 0; k < kH; ++k) s += w[oW4 + i * kH + k] * h1[k];
    out[i] = s;
  }
}

// log(1 - kUniform), log(kUniform / (4 pi)), log(2 pi)
constexpr double kLog1mU = -0.0317486983145803;
constexpr double kLogUni = -5.996760149769017;
constexpr double kLog2Pi = 1.8378770664093453;

// The proposal density of one site:
//   q(x) = (1 - kUniform) sum_k w_k vMF(x; kappa_k, mu_k) + kUniform / (4 pi).
// The parameters (w, kappa, mu) are computed from the network outputs in FP32 -- any
// values work, provided sampling and density use the same ones -- and promoted to FP64;
// the density, its normaliser and the sampler are FP64.
// ---- the sweep's network: hidden layers on the matrix cores --------------------------------
// One wave evaluates the network for its 64 lanes' inputs.  Layer 1 (5 -> 32) and the output
// layer (32 -> 6) run per lane in FP32; the two 32 x 32 layers run as
//   Z^T = W H^T   (v_mfma_f32_32x32x16_bf16: A = W, B = H^T, D = Z^T)
// with BF16 weights and activations and FP32 accumulation, two 32-proposal column blocks per
// wave.  Operand layouts of the instruction (lane l, half h = l / 32):
//   A[32][16]: A[l % 32][8 h + e], e < 8        B[16][32]: B[8 h + e][l % 32]
//   D[32][32]: register v holds D[8 (v / 4) + 4 h + v % 4][l % 32]
// A lane's D registers cover hidden units {8 q + 4 h + r} of one proposal; the B fragment of
// the next layer needs units {16 ks + 8 h + e}: half of them are the partner lane's (l ^ 32),
// exchanged with one cross-half shuffle per register pair.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));
// Two floats to BF16, round to nearest even, lo in bits 0-15: one v_cvt_pk_bf16_f32 on gfx950.
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  const floatx2 v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

__device__ __forceinline__ bf16x8 frag(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3) {
  uint4 u = make_uint4(d0, d1, d2, d3);
  return __builtin_bit_cast(bf16x8, u);
}

// W2 / W3 A-fragments (BF16, from global memory): row l % 32, columns 16 ks + 8 h + 0..7.
struct NetA {
  bf16x8 a[2][2];   // [layer 2 / 3][ks]
};
__device__ __forceinline__ NetA load_net_a(const uint16_t *wbf16) {
  NetA n;
  const int l = threadIdx.x & 63, h = l >> 5, i = l & 31;
#pragma unroll
  for (int layer = 0; layer < 2; ++layer)
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const uint4 u = *reinterpret_cast<const uint4 *>(wbf16 + (layer ? oW3 : oW2) + i * kH + 16 * ks + 8 * h);
      n.a[layer][ks] = __builtin_bit_cast(bf16x8, u);
    }
  return n;
}

// B fragments of one 32-proposal column block from D registers (after the activation):
// keep q in {h, 2 + h}, exchange q in {1 - h, 3 - h} with the partner lane.
__device__ __forceinline__ void d_to_b(const floatx16 &z, int h, bf16x8 b[2]) {
  uint32_t p[8];   // p[2 q + r / 2]: units 8 q + 4 h + r, r = 0..3, as BF16 pairs
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    p[2 * q] = pack_bf16(z[4 * q], z[4 * q + 1]);
    p[2 * q + 1] = pack_bf16(z[4 * q + 2], z[4 * q + 3]);
  }
  // send q = 1 - h and q = 3 - h; receive the partner's q = h' for the same slots
  const uint32_t s0 = h ? p[0] : p[2], s1 = h ? p[1] : p[3];
  const uint32_t s2 = h ? p[4] : p[6], s3 = h ? p[5] : p[7];
  const uint32_t r0 = __shfl_xor(s0, 32, 64), r1 = __shfl_xor(s1, 32, 64);
  const uint32_t r2 = __shfl_xor(s2, 32, 64), r3 = __shfl_xor(s3, 32, 64);
  // ks = 0: units 8 h + 0..7; ks = 1: units 16 + 8 h + 0..7
  //   h = 0: [own q = 2 ks | partner q = 2 ks];   h = 1: [partner q = 2 ks + 1 | own q = 2 ks + 1]
  b[0] = h ? frag(r0, r1, p[2], p[3]) : frag(p[0], p[1], r0, r1);
  b[1] = h ? frag(r2, r3, p[6], p[7]) : frag(p[4], p[5], r2, r3);
}

__device__ __forceinline__ floatx16 bias_d(const float *b, int h) {
  floatx16 z;
#pragma unroll
  for (int v = 0; v < 16; ++v) z[v] = b[8 * (v >> 2) + 4 * h + (v & 3)];
  return z;
}

// Network of the sweep for the calling lane's input x (every lane of the wave must call).
// wl: FP32 weights in LDS (layer 1, biases, layer 4 from the BF16 copy); na: W2/W3 fragments.
__device__ __forceinline__ void net_forward_mfma(const float *wl, const NetA &na, const float x[kIn],
                                                 float out[kOut]) {
  const int l = threadIdx.x & 63, h = l >> 5;
  // layer 1 per lane, then B fragments: own units 16 ks + 8 h + e for column block h, the
  // partner's same units for column block 1 - h
  uint32_t own[8], snd[8];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int half = 0; half < 2; ++half)
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        float z[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          const int u = 16 * ks + 8 * (half ^ h) + e + t;   // half 0: own units, 1: the partner's
          float acc = wl[oB1 + u];
#pragma unroll
          for (int k = 0; k < kIn; ++k) acc += wl[oW1 + u * kIn + k] * x[k];
          z[t] = act(acc);
        }
        (half ? snd : own)[4 * ks + e / 2] = pack_bf16(z[0], z[1]);
      }
  uint32_t rcv[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) rcv[k] = __shfl_xor(snd[k], 32, 64);
  bf16x8 b[2][2];   // [column block][ks]
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    const bf16x8 mine = frag(own[4 * ks], own[4 * ks + 1], own[4 * ks + 2], own[4 * ks + 3]);
    const bf16x8 theirs = frag(rcv[4 * ks], rcv[4 * ks + 1], rcv[4 * ks + 2], rcv[4 * ks + 3]);
    b[0][ks] = h ? theirs : mine;
    b[1][ks] = h ? mine : theirs;
  }
  // layers 2 and 3 on the matrix cores
  floatx16 z[2];
#pragma unroll
  for (int layer = 0; layer < 2; ++layer) {
    const float *bias = wl + (layer ? oB3 : oB2);
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      z[cb] = bias_d(bias, h);
      z[cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(na.a[layer][0], b[cb][0], z[cb], 0, 0, 0);
      z[cb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(na.a[layer][1], b[cb][1], z[cb], 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 16; ++v) z[cb][v] = act(z[cb][v]);
    }
    if (layer == 0)
#pragma unroll
      for (int cb = 0; cb < 2; ++cb) d_to_b(z[cb], h, b[cb]);
  }
  // output layer: partial sums over this lane's 16 units of each column block, completed
  // with the partner's; the lane's own proposal is column block h
#pragma unroll
  for (int q = 0; q < kOut; ++q) {
    float part[2];
#pragma unroll
    for (int cb = 0; cb < 2; ++cb) {
      float acc = 0.0f;
#pragma unroll
      for (int v = 0; v < 16; ++v) acc += wl[oW4 + q * kH + 8 * (v >> 2) + 4 * h + (v & 3)] * z[cb][v];
      part[cb] = acc;
    }
    const float give = h ? part[0] : part[1];     // the partner's column block
    const float got = __shfl_xor(give, 32, 64);
    out[q] = wl[oB4 + q] + (h ? part[1] : part[0]) + got;
  }
}

struct Mix {
  double logw[2];   // log of (1 - kUniform) w_k
  double kappa[2];
  double lnorm[2];  // log of kappa / (2 pi (1 - e^{-2 kappa})): vMF density = exp(lnorm + kappa (mu.x - 1))
  double e2k[2];    // e^{-2 kappa}
  double t[2];      // tilt angle of mu_k from ghat
  V3 mu[2];
};

__device__ __forceinline__ V3 vnorm(V3 v) {
  const double r = rsqrt(dot(v, v));
  return {v.x * r, v.y * r, v.z * r};
}

// Unit vector in the (ghat, axis) plane, normal to ghat, toward the pole of the axis nearer
// to ghat (any normal when ghat is along the axis).
__device__ __forceinline__ V3 tilt_frame(const V3 &ghat, const V3 &ax) {
  const double c = dot(ghat, ax);
  const double sgn = c < 0.0 ? -1.0 : 1.0;
  V3 e = {sgn * ax.x - sgn * c * ghat.x, sgn * ax.y - sgn * c * ghat.y, sgn * ax.z - sgn * c * ghat.z};
  if (dot(e, e) < 1e-24) e = fabs(ghat.x) < 0.9 ? V3{0.0, -ghat.z, ghat.y} : V3{-ghat.z, 0.0, ghat.x};
  return vnorm(e);
}

__device__ __forceinline__ float softplusf(float z) { return fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z))); }

// Mixture from the network outputs and the local frame (ghat, pole axis).
__device__ __forceinline__ Mix make_mix(const float o[kOut], float kappa0, const V3 &ghat, const V3 &ax) {
  Mix m;
  const V3 e = tilt_frame(ghat, ax);
  const float d = o[4] - o[1];                 // log softmax of the two weight logits
  m.logw[0] = (double)(-softplusf(d)) + kLog1mU;
  m.logw[1] = (double)(-softplusf(-d)) + kLog1mU;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const float sc = fminf(fmaxf(o[3 * k], -8.0f), 8.0f);
    const double kap = (double)(kappa0 * expf(sc)) + 1e-6;
    m.kappa[k] = kap;
    const double em1 = expm1(-2.0 * kap);
    m.e2k[k] = 1.0 + em1;
    m.lnorm[k] = log(kap / -em1) - kLog2Pi;
    const float th = tanhf(o[3 * k + 2]);
    float sn, cs;
    sincospif(0.5f * th, &sn, &cs);
    m.t[k] = 0.5 * kPi * (double)th;
    m.mu[k] = vnorm(V3{cs * ghat.x + sn * e.x, cs * ghat.y + sn * e.y, cs * ghat.z + sn * e.z});
  }
  return m;
}

// The mixture as the sweep uses it, without logarithms: selection weights ws_k = (1 - kUniform)
// w_k, concentrations, e^{-2 kappa_k} and density prefactors c_k = ws_k kappa_k / (2 pi (1 -
// e^{-2 kappa_k})), so that q(x) = sum_k c_k exp(kappa_k (mu_k . x - 1)) + kUniform / (4 pi).
// Sampling and density use the same ws_k, kappa_k and mu_k (the log-domain Mix of training
// differs from it only by rounding, which no sweep quantity depends on).
struct MixS {
  double ws0;       // selection weight of component 0 (component 1: 1 - kUniform - ws0)
  double kappa[2], e2k[2], c[2];
  V3 mu[2];
};
__device__ __forceinline__ MixS make_mix_sweep(const float o[kOut], float kappa0, const V3 &ghat, const V3 &ax) {
  MixS m;
  const V3 e = tilt_frame(ghat, ax);
  const double w0 = (double)__builtin_amdgcn_rcpf(1.0f + __expf(o[4] - o[1]));   // softmax of the logits
  const double ws[2] = {(1.0 - kUniform) * w0, (1.0 - kUniform) * (1.0 - w0)};
  m.ws0 = ws[0];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const float sc = fminf(fmaxf(o[3 * k], -8.0f), 8.0f);
    const double kap = (double)(kappa0 * expf(sc)) + 1e-6;
    m.kappa[k] = kap;
    // e^{-2 kappa} < 2^-54 for kappa > 19: expm1 would round to exactly -1 there
    const double em1 = kap > 19.0 ? -1.0 : expm1(-2.0 * kap);
    m.e2k[k] = 1.0 + em1;
    m.c[k] = ws[k] * kap / (2.0 * kPi * -em1);
    const float th = tanhf(o[3 * k + 2]);
    float sn, cs;
    sincospif(0.5f * th, &sn, &cs);
    m.mu[k] = vnorm(V3{cs * ghat.x + sn * e.x, cs * ghat.y + sn * e.y, cs * ghat.z + sn * e.z});
  }
  return m;
}

// log vMF_k(x) = lnorm_k + kappa_k (mu_k . x - 1)
__device__ __forceinline__ double log_vmf(const Mix &m, int k, const V3 &x) {
  return m.lnorm[k] + m.kappa[k] * (dot(m.mu[k], x) - 1.0);
}

__device__ __forceinline__ double log_q(const Mix &m, const V3 &x) {
  const double l0 = m.logw[0] + log_vmf(m, 0, x);
  const double l1 = m.logw[1] + log_vmf(m, 1, x);
  const double mx = fmax(fmax(l0, l1), kLogUni);
  return mx + log(exp(l0 - mx) + exp(l1 - mx) + exp(kLogUni - mx));
}

// Wood (1994) for S^2: w = 1 + log(v + (1 - v) e^{-2 kappa}) / kappa, azimuth 2 pi u; kappa = 0
// is the uniform density (w = 2 v - 1).
__device__ __forceinline__ V3 sample_vmf(double kappa, double e2k, const V3 &mu, double v, double u) {
  double w = kappa > 0.0 ? 1.0 + log(v + (1.0 - v) * e2k) / kappa : 2.0 * v - 1.0;
  w = fmin(1.0, fmax(-1.0, w));
  const double st = sqrt(fmax(0.0, 1.0 - w * w));
  const V3 e1 = vnorm(fabs(mu.x) < 0.9 ? V3{0.0, -mu.z, mu.y} : V3{-mu.z, 0.0, mu.x});
  const V3 e2 = {mu.y * e1.z - mu.z * e1.y, mu.z * e1.x - mu.x * e1.z, mu.x * e1.y - mu.y * e1.x};
  double sp, cp;
  sincospi(2.0 * u, &sp, &cp);
  return {w * mu.x + st * (cp * e1.x + sp * e2.x), w * mu.y + st * (cp * e1.y + sp * e2.y),
          w * mu.z + st * (cp * e1.z + sp * e2.z)};
}

// Local environment of device site d of one replica: G, Q_B (xx,yy,zz,xy,xz,yz).
template <bool CMP, bool HB, bool HD, class SP>
__device__ __forceinline__ void local_env(const KArgs &a, const SP &sp, const double *ctab, int d,
                                          V3 &G, double QB[6]) {
  G = {0, 0, 0};
  for (int k = 0; k < 6; ++k) QB[k] = 0.0;
  const double *crow = CMP ? ctab + (size_t)a.basis[d] * a.zmax * 5 : nullptr;
  const unsigned ud = (unsigned)d;
  for (int l0 = 0; l0 < a.zmax; l0 += kChunkLearn) {
    unsigned jj[kChunkLearn];
    V3 sv[kChunkLearn];
#pragma unroll
    for (int u = 0; u < kChunkLearn; ++u) {
      const int l = l0 + u < a.zmax ? l0 + u : 0;
      jj[u] = (unsigned)ld_off(a.nbr, ((unsigned)l * (unsigned)a.N + ud) * 4u);
    }
#pragma unroll
    for (int u = 0; u < kChunkLearn; ++u) sv[u] = sp.ld(jj[u]);
#pragma unroll
    for (int u = 0; u < kChunkLearn; ++u) {
      const int l = l0 + u;
      if (l >= a.zmax) continue;
      const V3 sj = sv[u];
      double cj, cb = 0, cx = 0, cy = 0, cz = 0;
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
      G.x += cj * sj.x; G.y += cj * sj.y; G.z += cj * sj.z;
      if (HD) {  // -(S_j x D)
        G.x -= sj.y * cz - sj.z * cy;
        G.y -= sj.z * cx - sj.x * cz;
        G.z -= sj.x * cy - sj.y * cx;
      }
      if (HB) {
        QB[0] += cb * sj.x * sj.x; QB[1] += cb * sj.y * sj.y; QB[2] += cb * sj.z * sj.z;
        QB[3] += cb * sj.x * sj.y; QB[4] += cb * sj.x * sj.z; QB[5] += cb * sj.y * sj.z;
      }
    }
  }
  if (a.has_h) {
    const double *h = a.h4 + 4 * a.basis[d];
    G.x -= h[0] * h[1]; G.y -= h[0] * h[2]; G.z -= h[0] * h[3];
  }
}

__device__ __forceinline__ double quad(const double Q[6], const V3 &s) {
  return Q[0] * s.x * s.x + Q[1] * s.y * s.y + Q[2] * s.z * s.z +
         2.0 * (Q[3] * s.x * s.y + Q[4] * s.x * s.z + Q[5] * s.y * s.z);
}

// Features of the network from the environment of a spin of length mag: the site energy of
// direction x is beta E = -kappa0 (ghat.x) + beta mag^2 x^T Q x with kappa0 = beta mag |G|.
//   x0 = log(1 + kappa0)                    strength of the local field
//   x1 = |ghat . a|                         field direction against the anisotropy axis
//   x2 = log T
//   x3 = beta mag^2 (ghat^T Q_B ghat - tr Q_B / 3)   biquadratic projection on the field
//   x4 = beta mag^2 Delta                   single-ion anisotropy
__device__ __forceinline__ void features(const V3 &G, const double QB[6], const V3 &ax, double Delta,
                                         double beta, double mag, float x[kIn], float &kappa0, V3 &ghat) {
  const double g2 = dot(G, G);
  double g = 0.0;
  if (g2 > 1e-300) {
    g = sqrt(g2);
    const double r = 1.0 / g;
    ghat = {G.x * r, G.y * r, G.z * r};
  } else {
    ghat = ax;
  }
  kappa0 = (float)(beta * mag * g);
  const float bm2 = (float)(beta * mag * mag);
  const float trq = (float)(QB[0] + QB[1] + QB[2]);
  x[0] = log1pf(kappa0);
  x[1] = fabsf((float)dot(ghat, ax));
  x[2] = logf((float)(1.0 / beta));
  x[3] = bm2 * ((float)quad(QB, ghat) - trq * (1.0f / 3.0f));
  x[4] = bm2 * (float)Delta;
}

__device__ __forceinline__ void axis_of(const KArgs &a, int d, V3 &ax, double &Delta) {
  if (a.has_si) {
    const double *k = a.si4 + 4 * a.basis[d];
    Delta = k[0];
    ax = {k[1], k[2], k[3]};
  } else {
    Delta = 0.0;
    ax = {0.0, 0.0, 1.0};
  }
}

// One colour phase with learned proposals (global-memory path, one lane per site).
#ifndef ESP_LEARN_WAVES
#define ESP_LEARN_WAVES 4        // > 0: ask the compiler for this many waves per SIMD (4: no spills, no AGPRs)
#endif
#if ESP_LEARN_WAVES > 0
#define ESP_LEARN_ATTR __attribute__((amdgpu_waves_per_eu(ESP_LEARN_WAVES)))
#else
#define ESP_LEARN_ATTR
#endif

template <bool CMP, bool HB, bool HD>
__global__ __launch_bounds__(kBlock) ESP_LEARN_ATTR void k_phase_learned(KArgs a, const uint16_t *wbf16) {
  extern __shared__ double s_dyn[];  // [ctab][weights as float]
  int r, b;
  decode_block(a, r, b);
  if (r >= a.R) return;
  const int nct = CMP ? a.nb * a.zmax * 5 : 0;
  double *ctab = s_dyn;
  float *w = reinterpret_cast<float *>(s_dyn + nct);
  for (int k = threadIdx.x; k < nct; k += kBlock) ctab[k] = a.ctab[k];
  for (int k = threadIdx.x; k < kNP; k += kBlock) w[k] = __uint_as_float((uint32_t)wbf16[k] << 16);
  __syncthreads();
  // every lane runs the network (matrix instructions need the whole wave); lanes past the
  // colour class evaluate a stand-in site and write nothing
  const int k = b * kBlock + threadIdx.x;
  const bool active = k < a.c_count;
  const int d = a.c_begin + (active ? k : 0);
  const int slot = a.rep_slot[r];
  const double beta = a.beta_slot[slot];
  const unsigned long long sweep = a.clock->sweep + (unsigned long long)a.sweep_off;
  double *base = a.sp + (size_t)r * 3 * a.Ns;
  const GSpins sp{base, base + a.Ns, base + 2 * a.Ns};
  V3 G;
  double QB[6];
  local_env<CMP, HB, HD>(a, sp, ctab, d, G, QB);
  V3 ax;
  double Delta;
  axis_of(a, d, ax, Delta);
  double Q[6] = {QB[0] + Delta * ax.x * ax.x, QB[1] + Delta * ax.y * ax.y, QB[2] + Delta * ax.z * ax.z,
                 QB[3] + Delta * ax.x * ax.y, QB[4] + Delta * ax.x * ax.z, QB[5] + Delta * ax.y * ax.z};
  const double mag = a.mag[d];
  float x[kIn], o[kOut];
  float kappa0;
  V3 ghat;
  features(G, QB, ax, Delta, beta, mag, x, kappa0, ghat);
  net_forward_mfma(w, load_net_a(wbf16), x, o);
  const MixS mix = make_mix_sweep(o, kappa0, ghat, ax);
  const int ir = a.dev2ref[d];
  const uint4 rn = philox10(make_uint4((uint32_t)ir, (uint32_t)sweep, (uint32_t)(sweep >> 32),
                                       a.stream_base + (uint32_t)r),
                            a.key0, a.key1);
  // component: uniform with probability kUniform, else network component 0 or 1
  const double uc = u01(rn.x);
  const int comp = uc < kUniform ? 2 : (uc - kUniform < mix.ws0 ? 0 : 1);
  // select by value (a dynamic index into the register-resident mixture would spill it)
  const bool one = comp == 1;
  const V3 xn = sample_vmf(comp == 2 ? 0.0 : (one ? mix.kappa[1] : mix.kappa[0]),
                           comp == 2 ? 1.0 : (one ? mix.e2k[1] : mix.e2k[0]),
                           comp == 2 ? ghat : (one ? mix.mu[1] : mix.mu[0]), u01(rn.y), u01(rn.z));
  const V3 so = sp.ld((unsigned)d);
  const V3 sn = {mag * xn.x, mag * xn.y, mag * xn.z};
  const V3 xo = vnorm(so);
  const double dE = -(G.x * (sn.x - so.x) + G.y * (sn.y - so.y) + G.z * (sn.z - so.z)) + quad(Q, sn) - quad(Q, so);
  // Metropolis-Hastings ratio exp(-beta dE) q(xo) / q(xn) with q in the linear domain: the
  // uniform component keeps both densities >= kUniform / (4 pi) > 0
  const double c0 = mix.c[0], c1 = mix.c[1];
  const double cu = kUniform / (4.0 * kPi);
  const double qo = c0 * exp(mix.kappa[0] * (dot(mix.mu[0], xo) - 1.0)) +
                    c1 * exp(mix.kappa[1] * (dot(mix.mu[1], xo) - 1.0)) + cu;
  const double qn = c0 * exp(mix.kappa[0] * (dot(mix.mu[0], xn) - 1.0)) +
                    c1 * exp(mix.kappa[1] * (dot(mix.mu[1], xn) - 1.0)) + cu;
  const double ratio = exp(-beta * dE) * qo / qn;
  const bool acc = ratio >= 1.0 || u01(rn.w) < ratio;
  if (active && acc) {
    sp.x[d] = sn.x; sp.y[d] = sn.y; sp.z[d] = sn.z;
  }
  const unsigned long long m = __ballot(active && !acc);
  if ((threadIdx.x & 63) == 0 && m) rej_add(a.rej, slot, (unsigned long long)__popcll(m));
  Proposal pr;
  pr.so = so; pr.sn = sn; pr.u3 = 0.0;
  wave_sums(a, r, b, active, !acc, dE, pr, 1);   // accepted dE into the wave's accumulator
}

// Records of local conditionals: n random (replica, site) pairs of the current state.
template <bool CMP, bool HB, bool HD>
__global__ __launch_bounds__(kBlock) void k_collect(KArgs a, int n, uint32_t s0, uint32_t s1, float *out) {
  extern __shared__ double s_dyn[];
  const int nct = CMP ? a.nb * a.zmax * 5 : 0;
  for (int k = threadIdx.x; k < nct; k += kBlock) s_dyn[k] = a.ctab[k];
  __syncthreads();
  const int t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= n) return;
  const uint4 rn = philox10(make_uint4((uint32_t)t, 0x636f6cu, 0u, 0u), s0, s1);
  const int r = (int)(rn.x % (uint32_t)a.R);
  const int d = (int)(rn.y % (uint32_t)a.N);
  double *base = a.sp + (size_t)r * 3 * a.Ns;
  const GSpins sp{base, base + a.Ns, base + 2 * a.Ns};
  V3 G;
  double QB[6];
  local_env<CMP, HB, HD>(a, sp, s_dyn, d, G, QB);
  V3 ax;
  double Delta;
  axis_of(a, d, ax, Delta);
  const double beta = a.beta_slot[a.rep_slot[r]];
  float x[kIn];
  float kappa0;
  V3 ghat;
  features(G, QB, ax, Delta, beta, a.mag[d], x, kappa0, ghat);
  const V3 s = vnorm(sp.ld((unsigned)d));
  float *rec = out + (size_t)t * kRec;
  for (int k = 0; k < kIn; ++k) rec[k] = x[k];
  rec[5] = kappa0;
  rec[6] = (float)ghat.x; rec[7] = (float)ghat.y; rec[8] = (float)ghat.z;
  rec[9] = (float)ax.x; rec[10] = (float)ax.y; rec[11] = (float)ax.z;
  rec[12] = (float)s.x; rec[13] = (float)s.y; rec[14] = (float)s.z;
  rec[15] = (float)((size_t)r * a.N + a.dev2ref[d]);  // replica * N + site (exact below 2^24)
}

// Per-block gradient of the mean negative log-likelihood over records [off, off + n):
// each thread runs one record forward and backward; the block then sums the parameter
// gradients over its records in a fixed order (deterministic).
// Unit vector stored at rec[o..o+2] (records hold FP32; renormalised in FP64).
__device__ __forceinline__ V3 rec_vec(const float *rec, int o) {
  return vnorm(V3{(double)rec[o], (double)rec[o + 1], (double)rec[o + 2]});
}

__global__ __launch_bounds__(kTrainBlock) void k_train_block(const float *w, const float *recs, int n,
                                                             const int *perm, int off, float *part_grad,
                                                             double *part_loss) {
  extern __shared__ float s_tr[];
  // per record (one thread): a0[5] a1[32] a2[32] a3[32] | g4[6] d3[32] d2[32] d1[32]; the
  // delta slots hold the pre-activations z_l until the backward pass overwrites them
  constexpr int kA = kIn + 3 * kH, kD = kOut + 3 * kH, kS = kA + kD;
  float *my = s_tr + threadIdx.x * kS;  // odd stride: lanes hit distinct banks
  float *a0 = my, *a1 = a0 + kIn, *a2 = a1 + kH, *a3 = a2 + kH;
  float *g4 = my + kA, *d3 = g4 + kOut, *d2 = d3 + kH, *d1 = d2 + kH;
  __shared__ double s_loss[kTrainBlock];
  const int i = blockIdx.x * kTrainBlock + threadIdx.x;
  double loss = 0.0;
  if (i < n) {
    const float *rec = recs + (size_t)(perm ? perm[off + i] : off + i) * kRec;
    for (int k = 0; k < kIn; ++k) a0[k] = rec[k];
    for (int q = 0; q < kH; ++q) {
      float z = w[oB1 + q];
      for (int k = 0; k < kIn; ++k) z += w[oW1 + q * kIn + k] * a0[k];
      d1[q] = z;
      a1[q] = act(z);
    }
    for (int q = 0; q < kH; ++q) {
      float z = w[oB2 + q];
      for (int k = 0; k < kH; ++k) z += w[oW2 + q * kH + k] * a1[k];
      d2[q] = z;
      a2[q] = act(z);
    }
    for (int q = 0; q < kH; ++q) {
      float z = w[oB3 + q];
      for (int k = 0; k < kH; ++k) z += w[oW3 + q * kH + k] * a2[k];
      d3[q] = z;
      a3[q] = act(z);
    }
    float o[kOut];
    for (int q = 0; q < kOut; ++q) {
      float z = w[oB4 + q];
      for (int k = 0; k < kH; ++k) z += w[oW4 + q * kH + k] * a3[k];
      o[q] = z;
    }
    const V3 ghat = rec_vec(rec, 6), ax = rec_vec(rec, 9), s = rec_vec(rec, 12);
    const Mix m = make_mix(o, rec[5], ghat, ax);
    double lv[2];
    for (int k = 0; k < 2; ++k) lv[k] = m.logw[k] + log_vmf(m, k, s);
    const double mx = fmax(fmax(lv[0], lv[1]), kLogUni);
    const double lse = mx + log(exp(lv[0] - mx) + exp(lv[1] - mx) + exp(kLogUni - mx));
    loss = -lse;
    const double rnet = exp(lv[0] - lse) + exp(lv[1] - lse);  // responsibility of the network part
    // d loss / d outputs
    const V3 e = tilt_frame(ghat, ax);
    for (int k = 0; k < 2; ++k) {
      const double rk = exp(lv[k] - lse);      // responsibility of component k
      const double kap = m.kappa[k];
      const double dlk = 1.0 / kap - 2.0 / expm1(2.0 * kap) + (dot(m.mu[k], s) - 1.0);  // d log f / d kappa
      const double sraw = (double)o[3 * k];
      const double dks = (sraw > -8.0 && sraw < 8.0) ? (kap - 1e-6) : 0.0;              // d kappa / d o
      g4[3 * k] = (float)(-rk * dlk * dks);
      g4[3 * k + 1] = (float)(exp(m.logw[k] - kLog1mU) * rnet - rk);
      const double t = m.t[k];
      const double th = t / (0.5 * kPi);
      const V3 dmu = {-sin(t) * ghat.x + cos(t) * e.x, -sin(t) * ghat.y + cos(t) * e.y,
                      -sin(t) * ghat.z + cos(t) * e.z};
      g4[3 * k + 2] = (float)(-rk * kap * dot(dmu, s) * 0.5 * kPi * (1.0 - th * th));
    }
    for (int k = 0; k < kH; ++k) {
      float acc = 0.f;
      for (int q = 0; q < kOut; ++q) acc += w[oW4 + q * kH + k] * g4[q];
      d3[k] = acc * act_d(d3[k]);
    }
    for (int k = 0; k < kH; ++k) {
      float acc = 0.f;
      for (int q = 0; q < kH; ++q) acc += w[oW3 + q * kH + k] * d3[q];
      d2[k] = acc * act_d(d2[k]);
    }
    for (int k = 0; k < kH; ++k) {
      float acc = 0.f;
      for (int q = 0; q < kH; ++q) acc += w[oW2 + q * kH + k] * d2[q];
      d1[k] = acc * act_d(d1[k]);
    }
  } else {
    for (int k = 0; k < kS; ++k) my[k] = 0.f;
  }
  s_loss[threadIdx.x] = loss;
  __syncthreads();
  // parameter gradients: sum over the block's records in record order
  for (int p = threadIdx.x; p < kNP; p += kTrainBlock) {
    // layer l: weight gradient = delta_out[row] * act_in[col], bias gradient = delta_out[row]
    int dout_off, ain_off, row, col = 0;
    bool bias = false;
    if (p < oB1) { row = p / kIn; col = p % kIn; dout_off = kA + kOut + 2 * kH; ain_off = 0; }
    else if (p < oW2) { row = p - oB1; dout_off = kA + kOut + 2 * kH; ain_off = 0; bias = true; }
    else if (p < oB2) { row = (p - oW2) / kH; col = (p - oW2) % kH; dout_off = kA + kOut + kH; ain_off = kIn; }
    else if (p < oW3) { row = p - oB2; dout_off = kA + kOut + kH; ain_off = 0; bias = true; }
    else if (p < oB3) { row = (p - oW3) / kH; col = (p - oW3) % kH; dout_off = kA + kOut; ain_off = kIn + kH; }
    else if (p < oW4) { row = p - oB3; dout_off = kA + kOut; ain_off = 0; bias = true; }
    else if (p < oB4) { row = (p - oW4) / kH; col = (p - oW4) % kH; dout_off = kA; ain_off = kIn + 2 * kH; }
    else { row = p - oB4; dout_off = kA; ain_off = 0; bias = true; }
    float g = 0.f;
    for (int q = 0; q < kTrainBlock; ++q) {
      const float *rq = s_tr + q * kS;
      const float dq = rq[dout_off + row];
      g += bias ? dq : dq * rq[ain_off + col];
    }
    part_grad[(size_t)blockIdx.x * kNP + p] = g;
  }
  if (threadIdx.x == 0) {
    double sl = 0.0;
    for (int q = 0; q < kTrainBlock; ++q) sl += s_loss[q];
    part_loss[blockIdx.x] = sl;
  }
}

__global__ void k_train_reduce(const float *part_grad, const double *part_loss, int nblk, int n,
                               float *grad, double *loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kNP) {
    float g = 0.f;
    for (int b = 0; b < nblk; ++b) g += part_grad[(size_t)b * kNP + p];
    grad[p] = g / (float)n;
  }
  if (p == 0) {
    double l = 0.0;
    for (int b = 0; b < nblk; ++b) l += part_loss[b];
    *loss = l / (double)n;
  }
}

// Adam on the FP32 master weights; writes the BF16 copy used by the sweep.
__global__ void k_adam(float *w, float *mom, float *vel, const float *g, float lr, float b1, float b2,
                       float eps, int t, uint16_t *wbf16) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNP) return;
  const float gp = g[p];
  const float m = b1 * mom[p] + (1.f - b1) * gp;
  const float v = b2 * vel[p] + (1.f - b2) * gp * gp;
  mom[p] = m;
  vel[p] = v;
  const float mh = m / (1.f - powf(b1, (float)t));
  const float vh = v / (1.f - powf(b2, (float)t));
  const float nw = w[p] - lr * mh / (sqrtf(vh) + eps);
  w[p] = nw;
  // round to nearest even BF16
  const uint32_t u = __float_as_uint(nw);
  const uint32_t rb = ((u >> 16) & 1u) + 0x7FFFu;
  wbf16[p] = (uint16_t)((u + rb) >> 16);
}

__global__ void k_to_bf16(const float *w, uint16_t *wbf16) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kNP) return;
  const uint32_t u = __float_as_uint(w[p]);
  const uint32_t rb = ((u >> 16) & 1u) + 0x7FFFu;
  wbf16[p] = (uint16_t)((u + rb) >> 16);
}

// Evaluation: per record the six outputs, kappa_0, kappa_1, log q(s | env) and w_0, from the
// FP32 master network (sweep_net = 0) or from the network the sweep runs (BF16 weights and
// hidden activations on the matrix cores, sweep_net = 1).
constexpr int kEvalBlock = 128;
__global__ __launch_bounds__(kEvalBlock) void k_eval(const float *w, const uint16_t *wbf16, int sweep_net,
                                                     const float *recs, int n, double *out) {
  __shared__ float wl[kNP];
  const int i = blockIdx.x * kEvalBlock + threadIdx.x;
  const int ic = i < n ? i : n - 1;            // whole waves run the matrix-core network
  const float *rec = recs + (size_t)ic * kRec;
  float x[kIn], o[kOut];
  for (int k = 0; k < kIn; ++k) x[k] = rec[k];
  if (sweep_net) {
    for (int k = threadIdx.x; k < kNP; k += kEvalBlock) wl[k] = __uint_as_float((uint32_t)wbf16[k] << 16);
    const NetA na = load_net_a(wbf16);
    __syncthreads();
    net_forward_mfma(wl, na, x, o);
  } else {
    mlp_forward(w, x, o);
  }
  if (i >= n) return;
  const V3 ghat = rec_vec(rec, 6), ax = rec_vec(rec, 9), s = rec_vec(rec, 12);
  const Mix m = make_mix(o, rec[5], ghat, ax);
  double *q = out + (size_t)i * 10;
  for (int k = 0; k < kOut; ++k) q[k] = o[k];
  q[6] = m.kappa[0];
  q[7] = m.kappa[1];
  q[8] = log_q(m, s);
  q[9] = exp(m.logw[0] - kLog1mU);
}

using LearnFn = void (*)(KArgs, const uint16_t *);
using CollectFn = void (*)(KArgs, int, uint32_t, uint32_t, float *);
template <bool CMP, bool HB>
LearnFn lf2(bool hd) { return hd ? k_phase_learned<CMP, HB, true> : k_phase_learned<CMP, HB, false>; }
LearnFn learn_fn(const esp::HostModel &m) {
  if (m.compressed) return m.has_B ? lf2<true, true>(m.has_D) : lf2<true, false>(m.has_D);
  return m.has_B ? lf2<false, true>(m.has_D) : lf2<false, false>(m.has_D);
}
template <bool CMP, bool HB>
CollectFn cf2(bool hd) { return hd ? k_collect<CMP, HB, true> : k_collect<CMP, HB, false>; }
CollectFn collect_fn(const esp::HostModel &m) {
  if (m.compressed) return m.has_B ? cf2<true, true>(m.has_D) : cf2<true, false>(m.has_D);
  return m.has_B ? cf2<false, true>(m.has_D) : cf2<false, false>(m.has_D);
}

int learn_ready(esp_ctx *c, bool need_state) {
  if (!c) return ESP_EINVAL;
  if (hipSetDevice(c->device) != hipSuccess) return fail(c, ESP_EHIP, "hipSetDevice failed");
  if (need_state) {
    if (!c->have_model) return fail(c, ESP_ESTATE, "no model: call esp_set_model first");
    if (c->R <= 0) return fail(c, ESP_ESTATE, "no replicas: call esp_set_replicas first");
    if (c->hm.has_self)
      return fail(c, ESP_EINVAL, "learned proposals need a model without self bonds");
  }
  return ESP_OK;
}

int ensure_weights(esp_ctx *c) {
  if (c->d_lw) return ESP_OK;
  int rc;
  if ((rc = dalloc(c, c->d_lw, kNP)) || (rc = dalloc(c, c->d_lm, kNP)) || (rc = dalloc(c, c->d_lv, kNP)) ||
      (rc = dalloc(c, c->d_lbf, kNP)) || (rc = dalloc(c, c->d_grad, kNP)))
    return rc;
  return ESP_OK;
}

int ensure_records(esp_ctx *c, int need) {
  if (need <= c->rec_cap) return ESP_OK;
  int cap = std::max(need, 2 * c->rec_cap);
  float *nr = nullptr;
  int rc = dalloc(c, nr, (size_t)cap * kRec);
  if (rc) return rc;
  if (c->nrec > 0)
    HIPCHK(c, hipMemcpyAsync(nr, c->d_recs, sizeof(float) * kRec * c->nrec, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(c->d_recs);
  c->d_recs = nr;
  c->rec_cap = cap;
  return ESP_OK;
}

// Dynamic LDS of k_train_block: activations and deltas of every record of the block.
constexpr size_t train_shmem() { return (size_t)kTrainBlock * (kIn + 3 * kH + kOut + 3 * kH) * sizeof(float); }

// Gradient scratch for minibatches of up to n records and up to nsteps losses.
int ensure_grad_scratch(esp_ctx *c, int n, int nsteps) {
  const int nblk = (n + kTrainBlock - 1) / kTrainBlock;
  int rc;
  if (nblk > c->pgrad_cap) {
    if ((rc = dalloc(c, c->d_pgrad, (size_t)nblk * kNP)) || (rc = dalloc(c, c->d_ploss, nblk))) return rc;
    c->pgrad_cap = nblk;
  }
  if (nsteps > c->loss_cap) {
    if ((rc = dalloc(c, c->d_loss, nsteps))) return rc;
    c->loss_cap = nsteps;
  }
  HIPCHK(c, hipFuncSetAttribute((const void *)k_train_block, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)train_shmem()));
  return ESP_OK;
}

// Gradient of the mean loss over records (perm ? perm[off + i] : off + i), i < n; loss to d_loss[k].
int enqueue_grad(esp_ctx *c, const int *perm, int off, int n, int k) {
  const int nblk = (n + kTrainBlock - 1) / kTrainBlock;
  hipLaunchKernelGGL(k_train_block, dim3(nblk), dim3(kTrainBlock), train_shmem(), c->stream,
                     (const float *)c->d_lw, (const float *)c->d_recs, n, perm, off, c->d_pgrad, c->d_ploss);
  hipLaunchKernelGGL(k_train_reduce, dim3((kNP + 255) / 256), dim3(256), 0, c->stream, (const float *)c->d_pgrad,
                     (const double *)c->d_ploss, nblk, n, c->d_grad, c->d_loss + k);
  HIPCHK(c, hipGetLastError());
  return ESP_OK;
}

}  // namespace

namespace esp {

// nsweeps sweeps of learned proposals from clock->sweep + off0 (global-memory path).
int learn_enqueue_sweeps(esp_ctx *c, uint64_t seed, int nsweeps, int off0, hipEvent_t *ev, int *nlaunch) {
  int rc = learn_ready(c, true);
  if (rc) return rc;
  if (!c->have_learn) return fail(c, ESP_ESTATE, "no proposal weights: call esp_learn_set_weights first");
  const esp::HostModel &m = c->hm;
  KArgs a = base_args(c);
  a.tps = 1;
  a.mpart = nullptr;                  // the learned phases keep no spin sums
  a.key0 = (uint32_t)seed; a.key1 = (uint32_t)(seed >> 32);
  LearnFn fn = learn_fn(m);
  const size_t shm = (m.compressed ? c->ctab_bytes : 0) + kNP * sizeof(float);
  int k = 0;
  for (int s = 0; s < nsweeps; ++s)
    for (int col = 0; col < m.ncol; ++col) {
      a.col = col;
      a.c_begin = m.col_off[col];
      a.c_count = m.col_off[col + 1] - m.col_off[col];
      if (a.c_count == 0) continue;
      a.nblk = (a.c_count + kBlock - 1) / kBlock;
      a.xcd_remap = (c->R % 8 == 0) ? 1 : 0;
      a.sweep_off = off0 + s;
      if (ev)
        hipExtLaunchKernelGGL(fn, dim3(a.nblk * c->R), dim3(kBlock), shm, c->stream, ev[2 * k], ev[2 * k + 1], 0,
                              a, (const uint16_t *)c->d_lbf);
      else
        hipLaunchKernelGGL(fn, dim3(a.nblk * c->R), dim3(kBlock), shm, c->stream, a, (const uint16_t *)c->d_lbf);
      ++k;
    }
  HIPCHK(c, hipGetLastError());
  if (nlaunch) *nlaunch = k;
  return ESP_OK;
}

void learn_free(esp_ctx *c) {
  dfree(c->d_lw); dfree(c->d_lm); dfree(c->d_lv); dfree(c->d_lbf); dfree(c->d_recs); dfree(c->d_pgrad);
  dfree(c->d_grad); dfree(c->d_ploss); dfree(c->d_loss); dfree(c->d_perm);
  c->have_learn = false;
  c->nrec = c->rec_cap = c->pgrad_cap = c->loss_cap = c->perm_cap = 0;
}

}  // namespace esp

extern "C" {

int esp_learn_set_weights(esp_ctx *c, const float *w, int nparams) {
  int rc = learn_ready(c, false);
  if (rc) return rc;
  if (!w || nparams != kNP) return fail(c, ESP_EINVAL, "esp_learn_set_weights: need ESP_LEARN_NPARAMS weights");
  for (int p = 0; p < kNP; ++p)
    if (!std::isfinite(w[p])) return fail(c, ESP_EINVAL, "esp_learn_set_weights: non-finite weight");
  if ((rc = ensure_weights(c))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_lw, w, sizeof(float) * kNP, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_lm, 0, sizeof(float) * kNP, c->stream));
  HIPCHK(c, hipMemsetAsync(c->d_lv, 0, sizeof(float) * kNP, c->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3((kNP + 255) / 256), dim3(256), 0, c->stream, (const float *)c->d_lw, c->d_lbf);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->adam_t = 0;
  c->have_learn = true;
  drop_graph(c);
  return ESP_OK;
}

int esp_learn_get_weights(esp_ctx *c, float *w, int nparams) {
  int rc = learn_ready(c, false);
  if (rc) return rc;
  if (!w || nparams != kNP) return fail(c, ESP_EINVAL, "esp_learn_get_weights: need ESP_LEARN_NPARAMS weights");
  if (!c->have_learn) return fail(c, ESP_ESTATE, "no proposal weights");
  HIPCHK(c, hipMemcpyAsync(w, c->d_lw, sizeof(float) * kNP, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_learn_collect(esp_ctx *c, int n, uint64_t seed) {
  int rc = learn_ready(c, true);
  if (rc) return rc;
  if (n < 0 || (int64_t)c->nrec + n > (int64_t)1 << 27)
    return fail(c, ESP_EINVAL, "esp_learn_collect: record count out of range (max 2^27)");
  if (n == 0) return ESP_OK;
  if ((rc = ensure_records(c, c->nrec + n))) return rc;
  KArgs a = base_args(c);
  const size_t shm = c->hm.compressed ? c->ctab_bytes : 0;
  hipLaunchKernelGGL(collect_fn(c->hm), dim3((n + kBlock - 1) / kBlock), dim3(kBlock), shm, c->stream, a, n,
                     (uint32_t)seed, (uint32_t)(seed >> 32), c->d_recs + (size_t)kRec * c->nrec);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->nrec += n;
  return ESP_OK;
}

int esp_learn_clear(esp_ctx *c) {
  int rc = learn_ready(c, false);
  if (rc) return rc;
  c->nrec = 0;
  return ESP_OK;
}

int esp_learn_num_records(esp_ctx *c, int *n) {
  if (!c || !n) return ESP_EINVAL;
  *n = c->nrec;
  return ESP_OK;
}

int esp_learn_get_records(esp_ctx *c, int first, int n, float *out) {
  int rc = learn_ready(c, false);
  if (rc) return rc;
  if (!out || first < 0 || n < 0 || first + n > c->nrec) return fail(c, ESP_EINVAL, "record range");
  if (n == 0) return ESP_OK;
  HIPCHK(c, hipMemcpyAsync(out, c->d_recs + (size_t)kRec * first, sizeof(float) * kRec * n,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_learn_put_records(esp_ctx *c, const float *recs, int n) {
  int rc = learn_ready(c, false);
  if (rc) return rc;
  if ((!recs && n > 0) || n < 0 || n > (1 << 27)) return fail(c, ESP_EINVAL, "esp_learn_put_records: bad size");
  c->nrec = 0;
  if (n == 0) return ESP_OK;
  if ((rc = ensure_records(c, n))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_recs, recs, sizeof(float) * kRec * n, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->nrec = n;
  return ESP_OK;
}

int esp_learn_grad(esp_ctx *c, int first, int n, float *grad, double *loss) {
  int rc = learn_ready(c, false);
  if (rc) return rc;
  if (!c->have_learn) return fail(c, ESP_ESTATE, "no proposal weights");
  if (!grad || !loss || first < 0 || n <= 0 || first + n > c->nrec) return fail(c, ESP_EINVAL, "record range");
  if ((rc = ensure_grad_scratch(c, n, 1))) return rc;
  if ((rc = enqueue_grad(c, nullptr, first, n, 0))) return rc;
  HIPCHK(c, hipMemcpyAsync(grad, c->d_grad, sizeof(float) * kNP, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(loss, c->d_loss, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return ESP_OK;
}

int esp_learn_train(esp_ctx *c, int epochs, int batch, float lr, uint64_t seed, double *epoch_loss) {
  int rc = learn_ready(c, false);
  if (rc) return rc;
  if (!c->have_learn) return fail(c, ESP_ESTATE, "no proposal weights");
  if (epochs < 0 || batch < 1 || batch > (1 << 20) || !(lr > 0.0f) || c->nrec < 1)
    return fail(c, ESP_EINVAL, "esp_learn_train: bad epochs / batch / lr, or no records");
  const int n = c->nrec;
  const int nb = (n + batch - 1) / batch;
  if ((rc = ensure_grad_scratch(c, std::min(batch, n), nb))) return rc;
  if (n > c->perm_cap) {
    if ((rc = dalloc(c, c->d_perm, n))) return rc;
    c->perm_cap = n;
  }
  std::vector<int> perm(n);
  std::vector<double> losses(nb);
  uint64_t st = seed ^ 0x9E3779B97F4A7C15ull;
  auto next = [&st]() {  // splitmix64
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  for (int e = 0; e < epochs; ++e) {
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = n - 1; i > 0; --i) std::swap(perm[i], perm[(int)(next() % (uint64_t)(i + 1))]);
    HIPCHK(c, hipMemcpyAsync(c->d_perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    for (int b = 0; b < nb; ++b) {
      const int off = b * batch, cnt = std::min(batch, n - off);
      if ((rc = enqueue_grad(c, c->d_perm, off, cnt, b))) return rc;
      c->adam_t += 1;
      hipLaunchKernelGGL(k_adam, dim3((kNP + 255) / 256), dim3(256), 0, c->stream, c->d_lw, c->d_lm, c->d_lv,
                         (const float *)c->d_grad, lr, 0.9f, 0.999f, 1e-8f, c->adam_t, c->d_lbf);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(losses.data(), c->d_loss, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));  // the host permutation buffer is reused
    if (epoch_loss) {
      double s = 0.0, wsum = 0.0;
      for (int b = 0; b < nb; ++b) {
        const double cnt = std::min(batch, n - b * batch);
        s += losses[b] * cnt;
        wsum += cnt;
      }
      epoch_loss[e] = s / wsum;
    }
  }
  drop_graph(c);
  return ESP_OK;
}

int esp_learn_eval(esp_ctx *c, int first, int n, int sweep_net, double *out) {
  int rc = learn_ready(c, false);
  if (rc) return rc;
  if (!c->have_learn) return fail(c, ESP_ESTATE, "no proposal weights");
  if (!out || first < 0 || n < 0 || first + n > c->nrec) return fail(c, ESP_EINVAL, "record range");
  if (n == 0) return ESP_OK;
  double *d_out = nullptr;
  if ((rc = dalloc(c, d_out, (size_t)n * 10))) return rc;
  hipLaunchKernelGGL(k_eval, dim3((n + kEvalBlock - 1) / kEvalBlock), dim3(kEvalBlock), 0, c->stream,
                     (const float *)c->d_lw, (const uint16_t *)c->d_lbf, sweep_net ? 1 : 0,
                     (const float *)c->d_recs + (size_t)kRec * first, n, d_out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(double) * 10 * n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dfree(d_out);
  if (e != hipSuccess) return fail(c, ESP_EHIP, std::string("esp_learn_eval: ") + hipGetErrorString(e));
  return ESP_OK;
}

}  // extern "C"

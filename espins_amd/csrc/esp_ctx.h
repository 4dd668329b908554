// esp_ctx.h -- the engine context behind the opaque esp_ctx of include/espins_hip.h, and
// the host helpers every translation unit of the context uses (internal linkage).
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/espins_hip.h"
#include "esp_kernels.h"
#include "esp_model.h"

using espk::Clock;
using espk::KArgs;

struct esp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  size_t lds_max = 65536;
  int Lhint[3] = {0, 0, 0};       // supercell size for the periodic colouring
  int path = 0;                   // 0 auto, 1 global-memory phases, 2 LDS-resident
  uint32_t stream_base = 0;       // Philox stream of replica r: stream_base + r
  int indexing = 0;               // 0 implicit geometry when the lattice allows, 1 explicit tables
  bool lds_fits = false;
  // model
  esp::HostModel hm;
  bool have_model = false;
  int *d_nbr = nullptr, *d_dev2ref = nullptr, *d_ref2dev = nullptr, *d_col_off = nullptr;
  double *d_cJ = nullptr, *d_cB = nullptr, *d_cD = nullptr, *d_ctab = nullptr;
  double *d_mag = nullptr, *d_si4 = nullptr, *d_h4 = nullptr;
  uint8_t *d_basis = nullptr;
  int *d_grp = nullptr, *d_gslot = nullptr;   // implicit geometry (esp_geo.h)
  double *d_magb = nullptr;
  bool use_lds = false;
  size_t lds_bytes = 0, ctab_bytes = 0;
  // replicas
  int R = 0;
  double *d_sp = nullptr;         // [R][3][N]
  double *d_w = nullptr;          // [R][N]
  int tps = 1;                    // lanes per site in the sweep kernels
  int Ns = 0;                     // replica stride (> N; index N is the zero spin)
  int tps_user = 0;               // 0 = automatic
  double *d_beta = nullptr, *d_ebase = nullptr, *d_erun = nullptr, *d_acc = nullptr;
  double *d_partial = nullptr;
  int red_blocks = 1;
  int *d_slot_rep = nullptr, *d_rep_slot = nullptr;
  unsigned long long *d_rej = nullptr, *d_nswap = nullptr, *d_nmeas = nullptr;
  Clock *d_clock = nullptr;
  unsigned *d_done = nullptr;    // blocks of k_reduce_fin done this step (the last one resets it)
  std::vector<double> tems;
  // host-fed staging
  double *d_prop = nullptr, *d_unif = nullptr;
  uint8_t *d_accd = nullptr;
  // per-wave spin sums of the colour phases: [R][ncol][mpart_nw][3]; valid while the last
  // change of the spins was a full global-path sweep (d_w has the same wave layout)
  double *d_mpart = nullptr;
  int mpart_nw = 0;
  bool mpart_valid = false;
  // per-measurement trace of E and |m| by slot (global-memory path only)
  double *d_trace = nullptr;      // [trace_cap][R][2]
  int trace_cap = 0;
  // learned proposal (esp_learn.hip)
  float *d_lw = nullptr;          // FP32 master weights [ESP_LEARN_NPARAMS]
  float *d_lm = nullptr, *d_lv = nullptr;  // Adam moments
  uint16_t *d_lbf = nullptr;      // BF16 copy the sweep reads
  bool have_learn = false;
  int adam_t = 0;
  float *d_recs = nullptr;        // local-conditional records [rec_cap][ESP_LEARN_RECORD]
  int nrec = 0, rec_cap = 0;
  float *d_pgrad = nullptr, *d_grad = nullptr;
  double *d_ploss = nullptr, *d_loss = nullptr;
  int pgrad_cap = 0, loss_cap = 0;
  int *d_perm = nullptr;
  int perm_cap = 0;
  // optional observables (esp_obs.hip), measured with the energy and m at every measurement
  // of the global-memory path; accumulators by slot, cleared with the statistics
  struct Obs {
    int flags = 0;                 // ESP_OBS_* bits that are on
    // order parameter / staggered magnetisation: [nb][3] axes, [nb] coefficients
    double *d_axes = nullptr, *d_coeff = nullptr;
    double *d_acc = nullptr;       // [R][6]: |OP|, OP^2, OP^4, |SM|, SM^2, SM^4
    double *d_part = nullptr;      // [R][cell blocks][4]
    int *d_cdev = nullptr;         // [nb][ncell] device sites of the cells, cells by first atom's device index
    int cblocks = 0;
    // P(m): grid m_min + k (m_max - m_min) / (nfreq - 1); counts [R][3][nfreq]
    int nfreq = 0;
    double m_min = 0.0, m_max = 0.0;
    unsigned long long *d_pm = nullptr;
    // spin correlations per shell: bond tables by device site
    int nshell = 0, nsp = 0, srows = 0;
    int *d_skind = nullptr;        // [N] kind of each site
    int *d_soff = nullptr;         // [nshell+1] first table row of each shell
    unsigned *d_sbond = nullptr;   // [srows][N] partner | kind of partner << 30
    double *d_sw = nullptr;        // [1 or 2][srows][N] weights (sum; or forward, reverse)
    double *d_sc = nullptr;        // [R][3][nshell][nsp][nsp]: sum, |sum|, sum |.|
    double *d_scpart = nullptr;    // [R][site blocks][nshell][2][nsp][nsp]
    int sblocks = 0;
    // structure factor
    int nq = 0, sf_every = 1;
    double polar[3] = {0, 0, 0};
    int have_polar = 0;
    double *d_q = nullptr, *d_pos = nullptr;   // [nq][3], [N][3] by device site
    double *d_sfspin = nullptr;    // [replica groups][N][8 or 4][3] spins of an S(q) evaluation
    double *d_sfpart = nullptr;    // [site chunks][R][6][nq] partial F0
    int sf_chunk = 0, sf_nchunk = 0;
    int sf_lat = 0, sfL[3] = {0, 0, 0};   // lattice positions: phases from per-q tables
    double2 *d_sftab = nullptr;    // [L0 + L1 + L2 + nb][nq] exp(i q.x A_k), exp(i q.B_a)
    double *d_sf = nullptr;        // [R][3][nq]
    unsigned long long *d_sfn = nullptr;      // structure-factor measurements
    // per-measurement |OP| and |SM| by slot while the trace is on: [trace_cap][R][2]
    double *d_trace = nullptr;
  } obs;
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
  dfree(c->d_sp); dfree(c->d_w); dfree(c->d_beta); dfree(c->d_ebase); dfree(c->d_erun); dfree(c->d_acc);
  dfree(c->d_partial); dfree(c->d_slot_rep); dfree(c->d_rep_slot); dfree(c->d_rej);
  dfree(c->d_nswap); dfree(c->d_nmeas); dfree(c->d_prop); dfree(c->d_unif); dfree(c->d_accd);
  dfree(c->d_trace);
  c->trace_cap = 0;
  dfree(c->d_mpart);
  c->mpart_nw = 0;
  c->mpart_valid = false;
  c->R = 0;
  drop_graph(c);
}

void free_model(esp_ctx *c) {
  dfree(c->d_nbr); dfree(c->d_dev2ref); dfree(c->d_ref2dev); dfree(c->d_col_off);
  dfree(c->d_cJ); dfree(c->d_cB); dfree(c->d_cD); dfree(c->d_ctab); dfree(c->d_mag);
  dfree(c->d_si4); dfree(c->d_h4); dfree(c->d_basis);
  dfree(c->d_grp); dfree(c->d_gslot); dfree(c->d_magb);
  c->have_model = false;
  drop_graph(c);
}

KArgs base_args(esp_ctx *c) {
  KArgs a{};
  const esp::HostModel &m = c->hm;
  a.sp = c->d_sp;
  a.w = c->d_w;
  a.mpart_nw = c->mpart_nw;
  a.tps = c->tps;
  a.nbr = c->d_nbr;
  a.cJ = c->d_cJ; a.cB = c->d_cB; a.cD = c->d_cD; a.ctab = c->d_ctab;
  a.mag = c->d_mag; a.basis = c->d_basis; a.dev2ref = c->d_dev2ref; a.col_off = c->d_col_off;
  a.si4 = c->d_si4; a.h4 = c->d_h4;
  a.beta_slot = c->d_beta; a.rep_slot = c->d_rep_slot; a.rej = c->d_rej; a.clock = c->d_clock;
  a.N = m.N; a.Ns = c->Ns; a.nb = m.nb; a.zmax = m.zmax; a.R = c->R; a.ncol = m.ncol;
  a.has_si = m.has_si; a.has_h = m.has_h;
  a.stream_base = c->stream_base;
  a.geo = m.geo;
  a.grp = c->d_grp; a.gslot = c->d_gslot; a.magb = c->d_magb;
  a.rows_max = m.rows_max;
  return a;
}

}  // namespace

namespace esp {
// Optional observables (esp_obs.hip): the measurement kernels of one measurement, enqueued
// after the partial sums of the energy and m and before the finalisation (which counts the
// measurement); state freed with the replicas / the model; accumulators cleared.
int obs_enqueue(esp_ctx *c);
void obs_free(esp_ctx *c);
int obs_resize(esp_ctx *c);      // (re)allocate the per-replica accumulators for c->R
int obs_clear(esp_ctx *c);
// Learned-proposal state (esp_learn.hip).
// ev (optional): 2 events per launch, recorded by the dispatch packets (timing); *nlaunch
// receives the number of launches.
int learn_enqueue_sweeps(esp_ctx *c, uint64_t seed, int nsweeps, int off0, hipEvent_t *ev = nullptr,
                         int *nlaunch = nullptr);
void learn_free(esp_ctx *c);
}  // namespace esp

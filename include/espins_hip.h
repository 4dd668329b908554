/*
 * espins_hip.h -- C ABI of libespins_hip.so, the MI355X (gfx950) engine for the ESpinS
 * Metropolis sweep and parallel-tempering replica exchange.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an int
 * status (ESP_OK = 0, negative on error; esp_last_error() gives the text) and never
 * aborts the process.  Arrays use the reference's Fortran layouts, so a Fortran caller
 * passes its own arrays through ISO_C_BINDING without copies:
 *   spin_matrix(3,N[,R])     double[R][N][3]
 *   jij_matrix(N,z)          double[z][N]     (slot-major, site fastest)
 *   jij_nbors_matrix(N,z)    int[z][N]        1-based site ids
 *   jij_num_nbors_matrix(N)  int[N]
 *   dij_matrix(3,N,z)        double[z][N][3]
 * Energies and temperatures are in the reference's internal units (Kelvin: couplings
 * already divided by k_B and by |m_i m_j|, as jij.F90:163,290 leave them).
 *
 * Threading: one host thread per context; calls on a context are ordered on its own
 * HIP stream and are synchronous at return unless their name ends in _async.
 */
#ifndef ESPINS_HIP_H
#define ESPINS_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ESP_OK        0
#define ESP_EINVAL   -1   /* bad argument (null pointer, size out of range)      */
#define ESP_EHIP     -2   /* HIP runtime error (no device, launch failure, ...)  */
#define ESP_ESTATE   -3   /* call out of order (no model / no replicas yet)      */
#define ESP_ENOMEM   -4   /* device or host allocation failed                     */
#define ESP_ETABLE   -5   /* coupling rows name an unknown atomic position        */

#define ESP_MODE_RANDOM 0 /* mcarlo_mode = random (get_quant.F90:80-93)          */
#define ESP_MODE_CONST  1 /* mcarlo_mode = const  (get_quant.F90:94-161)         */
#define ESP_MODE_LEARNED 2 /* learned vMF-mixture proposal, Metropolis-Hastings  */

#define ESP_LEARN_NPARAMS 2502 /* MLP 5 -> 32 -> 32 -> 32 -> 6 (weights + biases)   */
#define ESP_LEARN_RECORD  16   /* floats per local-conditional record               */

typedef struct esp_ctx esp_ctx;

/* ---- lifetime ------------------------------------------------------------------ */
/* Replaces the module-global state of mc_parameters/mc_jij (parameters.F90:93,
 * jij.F90:32-35): one opaque context per engine instance on HIP device `device`. */
int esp_create(esp_ctx **ctx, int device);
int esp_destroy(esp_ctx *ctx);
const char *esp_last_error(const esp_ctx *ctx);
/* Library version as 10000*major + 100*minor + patch. */
int esp_version(void);

/* ---- coupling tables (host only; no GPU needed) ---------------------------------- */
/* jij_get / bij_get / dij_get (jij.F90:123-446, bij.F90:88-359, dij.F90:87-391):
 * expand per-cell coupling rows over an L[0] x L[1] x L[2] supercell into ELL tables.
 * basis_frac[nb][3]: basis positions (species-sorted order the site index uses);
 * basis_mag[nb]: moments; rows: f1[nrows][3], f2[nrows][3], value[nrows] (internal units),
 * dvec[nrows][3] (unit DM vectors, ncomp==3 only, else NULL).  Zero bonds are dropped.
 * Outputs need capacity N*zcap (val: N*zcap*ncomp).  *zmax receives the realised width. */
int esp_build_table(int nb, const double *basis_frac, const double *basis_mag,
                    const int *L, int open_boundary,
                    int nrows, const double *f1, const double *f2, const double *value,
                    const double *dvec, int ncomp, int zcap,
                    int *cnt, int *nbr, double *val, int *zmax);

/* neighbors_get (neighbors.F90:48-168) for the input generation: for atom a of the cell
 * (Cartesian pos[nb][3], lattice vectors lat[3][3] as rows), its first nshell[a] neighbour
 * shells among the images l a1 + m a2 + n a3, |l|,|m|,|n| <= 5, distances equal within tol:
 * dnn[nb][cap_shell] shell radii, multi[nb][cap_shell] multiplicities, list[nb][cap_shell]
 * [cap_multi][4] = (atom b, l, m, n) in the reference's order.  Host only. */
int esp_neighbor_shells(int nb, const double *lat, const double *pos, const int *nshell, double tol,
                        int cap_shell, int cap_multi, double *dnn, int *multi, int *list);
/* jij_spin_glass (jij.F90:506-556): J += sigma x N(0,1) (MT19937 stream of seed, polar
 * method) for every bond i < j of the ELL table in table order, copied to the first (j, i)
 * slot.  sigma[z][N] as val.  ESP_ETABLE if a bond has no reverse (bad_pair = i, j).
 * Host only. */
int esp_spin_glass(int N, int z, const int *cnt, const int *nbr, double *val, const double *sigma, int seed,
                   int *bad_pair);

/* ---- model ------------------------------------------------------------------------ */
/* Upload the Hamiltonian (the tables energy_get / delta_energy_get read,
 * get_quant.F90:186-337).  zB = 0 / zD = 0 / have_si = 0 / have_field = 0 switch a term
 * off (their pointers may then be NULL).  si_par/h_par[nb], si_axis/h_axis[nb][3].
 * Computes a proper colouring of the union of the J, B and D graphs. */
int esp_set_model(esp_ctx *ctx, int N, int nb, const double *mag,
                  int zJ, const int *j_cnt, const int *j_nbr, const double *j_val,
                  int zB, const int *b_cnt, const int *b_nbr, const double *b_val,
                  int zD, const int *d_cnt, const int *d_nbr, const double *d_vec,
                  int have_si, const double *si_par, const double *si_axis,
                  int have_field, const double *h_par, const double *h_axis);
/* Optional, before esp_set_model: the supercell_size (parameters.F90:352-357) the tables
 * were expanded over.  Lets the colouring use a period of the lattice (colour classes of
 * equal size, few colours); any result is verified and greedy colouring is the fallback. */
int esp_set_supercell(esp_ctx *ctx, const int *L3);
/* Host-only planning (no device needed): the colouring and layout facts esp_set_model
 * would compute for these tables (L may be NULL).  colour[N] as in esp_get_coloring,
 * info[5] as in esp_get_model_info (info[4] = 0). */
int esp_plan_model(int N, int nb, const double *mag,
                   int zJ, const int *j_cnt, const int *j_nbr, const double *j_val,
                   int zB, const int *b_cnt, const int *b_nbr, const double *b_val,
                   int zD, const int *d_cnt, const int *d_nbr, const double *d_vec,
                   const int *L3, int *ncol, int *colour, int *info);
/* Sweep kernel choice: 0 = automatic (LDS-resident when a replica fits in LDS), 1 = global-
 * memory colour phases, 2 = LDS-resident (EINVAL if the replica does not fit). */
int esp_set_path(esp_ctx *ctx, int path);
/* Philox stream of replica r is base + r (default 0); the exchange stream is
 * 0xFFFFFFFF - base.  A batch gives chain k the base k << 16. */
int esp_set_stream_base(esp_ctx *ctx, uint32_t base);
/* colour[N]: 0-based colour of 1-based site i at colour[i-1]; *ncol colours. */
int esp_get_coloring(esp_ctx *ctx, int *ncol, int *colour);
/* Model facts: out[0]=N, out[1]=merged bond width, out[2]=colours, out[3]=coefficient
 * layout of the colour phases (0 per-bond tables, 1 per-basis tables, 2 per-basis tables and
 * the implicit lattice geometry), out[4]=1 if the LDS-resident sweep is used, out[5]=lanes per
 * site in the sweep kernels. */
int esp_get_model_info(esp_ctx *ctx, int *out6);
/* Neighbour indexing of the global-memory colour phases: 0 = compute partner and reference
 * site indices from the device index when the supercell is periodic and translation
 * invariant (the geometry is verified against the tables for every site and slot first),
 * 1 = always read the per-site neighbour tables.  The replaced loads are those of
 * jij_nbors_matrix in delta_energy_get (get_quant.F90:288-336). */
int esp_set_indexing(esp_ctx *ctx, int mode);
/* Lanes cooperating on one site's bond sum in the sweep kernels: 1, 2, 4 or 8; 0 = automatic
 * (from the bond width). */
int esp_set_tps(esp_ctx *ctx, int tps);

/* ---- replicas ---------------------------------------------------------------------- */
/* R replicas (configurations), temperature slot t holds tems[t] (ascending for PT,
 * parameters.F90:601).  spins[R][N][3]; replica r starts in slot r.  Resets the running
 * energies to the direct energies, the swap map, and all counters. */
int esp_set_replicas(esp_ctx *ctx, int R, const double *tems, const double *spins);
/* spins[R][N][3] of each replica (configuration), or of each slot when by_slot != 0. */
int esp_get_spins(esp_ctx *ctx, int by_slot, double *spins);
/* Overwrite the spins (e.g. after a checkpoint read); resets the running energies. */
int esp_put_spins(esp_ctx *ctx, const double *spins);

/* ---- hot path -------------------------------------------------------------------- */
/* energy_get (get_quant.F90:186-261) of every replica: E[R] (by replica). */
int esp_energy(esp_ctx *ctx, double *E);
/* Allocation-free variant of esp_energy, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  E is DEVICE-accessible memory of R doubles; the direct
 * energies of the R replicas are enqueued on the context's stream (esp_stream) and the call
 * returns without waiting.  The block partials live in a fixed device scratch array of
 * 262144 doubles (2 MiB) per GPU, looked up at esp_create: R x energy blocks (one block per
 * 4096 sites, at most 64) must not exceed it, else ESP_EINVAL.  The call makes no
 * hipMalloc/hipFree.  All contexts on one GPU share that array: calls that overlap on one
 * GPU (different streams, or threads) race. */
int esp_energy_noalloc(esp_ctx *ctx, double *E);
/* nsweeps Metropolis sweeps of every replica at its slot's temperature, in colour order:
 * metropolis + delta_energy_get (get_quant.F90:52-182, 264-337) with Philox4x32-10
 * uniforms, counter (site, sweep, replica), key seed.  Sweep k uses counter sweep0+k. */
int esp_sweep(esp_ctx *ctx, int nsweeps, int mode, double tilt, uint64_t seed, uint64_t sweep0);
/* Deterministic mode: one sweep in colour order with proposals prop[R][N][3] (new spin
 * vectors) and accept uniforms unif[R][N] given by the host; accepted[R][N] out (1/0). */
int esp_sweep_hostfed(esp_ctx *ctx, const double *prop, const double *unif, uint8_t *accepted);
/* One measurement (mcarlo.F90:536-543): accumulate E, E^2, E^4 of the running energy and
 * |m|, m^2, m^4 (magnetization_get, get_quant.F90:340-363) into each replica's slot. */
int esp_measure(esp_ctx *ctx);
/* pt_swap (pt.F90:1235-1288) on the device in label form: pairs (t,t+1) for
 * t = start-1, start+1, ... (start = 1 or 2, the reference's 1-based parity); uniform of
 * pair t from Philox counter (t, step, 0xFFFFFFFF-stream), key seed^0x5bd1e995. */
int esp_pt_exchange(esp_ctx *ctx, int start, uint64_t seed, uint64_t step);

/* Fused block of the PT / MC schedule (pt.F90:654-688 and 773-908 without I/O):
 * n_rounds x { steps_per_round x { steps_measure sweeps; [measure] }; [exchange] }.
 * Exchange parity starts at start_parity (1 or 2) and alternates per round.
 * Kernels are enqueued back-to-back (captured in a hipGraph when use_graph != 0). */
int esp_run(esp_ctx *ctx, int n_rounds, int steps_per_round, int steps_measure,
            int do_measure, int do_exchange, int start_parity,
            int mode, double tilt, uint64_t seed, uint64_t sweep0, uint64_t step0,
            int use_graph);

/* Timing of the sweep kernels alone: nsweeps sweeps with a HIP event pair recorded on the
 * context's stream around every sweep-kernel launch.  kernel_ms = sum of the per-launch
 * elapsed times, nlaunch = launches, wall_ms = first start to last stop,
 * updates_per_launch = attempted updates one launch performs (mean over colours). */
int esp_time_sweeps(esp_ctx *ctx, int nsweeps, int mode, double tilt, uint64_t seed,
                    uint64_t sweep0, double *kernel_ms, int *nlaunch, double *wall_ms,
                    int64_t *updates_per_launch);

/* ---- statistics ------------------------------------------------------------------ */
/* acc[R][6] by slot: sums of E, E^2, E^4, |m|, m^2, m^4; rejected[R] by slot;
 * e_run[R] by replica (running energy "by adding differences"); nmeas = measurements. */
int esp_get_stats(esp_ctx *ctx, double *acc, int64_t *rejected, double *e_run, int64_t *nmeas);
/* slot_rep[R]: 1-based replica id at slot t (the reference's map, pt.F90:1282-1284);
 * nswap[R][2]: attempts, accepts per slot (num_swap(2,R)). */
int esp_get_map(esp_ctx *ctx, int *slot_rep, int64_t *nswap);
/* Zero acc, rejected, nswap and the measurement count (phase boundaries, pt.F90:650). */
int esp_reset_stats(esp_ctx *ctx);
/* Record E and |m| of every slot at each measurement (capacity measurements, counted by
 * nmeas from the last esp_reset_stats; 0 switches the trace off).  While the trace is on,
 * sweeps use the global-memory path.  trace[n][R][2] by slot; *n = measurements stored. */
int esp_set_trace(esp_ctx *ctx, int capacity);
int esp_get_trace(esp_ctx *ctx, double *trace, int *n);
/* ---- optional observables (measured on the device at every measurement) ------------
 * Each setter switches one observable on (or off with a null table / zero size) for the
 * current model and replicas (set them again after esp_set_model / esp_set_replicas).
 * They are measured with E and m at every measurement of esp_run (the measured steps of
 * mcarlo.F90:526-590 and pt.F90:715-908), accumulated by slot (temperature), cleared by
 * esp_reset_stats.  While any is on, sweeps use the global-memory path.
 * Sites are in the reference order; a cell is nb consecutive sites. */
#define ESP_OBS_ORDER     1   /* order parameter, orderparam_get (get_quant.F90:396-422)  */
#define ESP_OBS_STAGGERED 2   /* staggered magnetisation, staggered_get (get_quant.F90:366-392) */
#define ESP_OBS_PM        4   /* P(m), Pm_get (get_quant.F90:425-493)                      */
#define ESP_OBS_SPINCORR  8   /* spin correlations per shell, spincorr_get (get_quant.F90:496-540) */
#define ESP_OBS_SFACTOR  16   /* structure factor, mcarlo_structure_factor (get_quant.F90:649-717) */
/* axes[nb][3] (orderparam_axes_cart, parameters.F90); OP = sum_i S_i . axes[basis i] / N. */
int esp_set_orderparam(esp_ctx *ctx, const double *axes);
/* coeff[nb] (staggered_m_coeff); SM = | sum_i coeff[basis i] S_i / N |. */
int esp_set_staggered(esp_ctx *ctx, int enable, const double *coeff);
/* P(m) of the cell moments |sum of a cell's spins| (and of the cell order parameter and
 * cell staggered moment when those are on) on the grid of mcarlo.F90:432-456:
 * m_max = sum_a |moment_a|, m_min = -m_max, nfreq = nint((m_max - m_min) / 0.01) + 1
 * (at least 2).  *nfreq receives the grid size. */
int esp_set_pm(esp_ctx *ctx, int enable, int *nfreq);
/* Spin correlations per shell (shell_neighbors_get, jij.F90:449-503): species[N] (0-based
 * kind of every site, at most 4 kinds), cnt[nshell][N] and nbr[nshell][zs][N] (1-based site
 * ids): the neighbours of each site in each shell.  The normalisation is the reference's
 * shell_num_nbors_kind (its per-kind count of the last site of each kind).  The call builds
 * device-order bond tables on the host (a shell whose table is symmetric stores each bond
 * once, with both directions' weights); the tables belong to the current model. */
int esp_set_spincorr(esp_ctx *ctx, int nshell, int nspecies, const int *species, int zs,
                     const int *cnt, const int *nbr);
/* Structure factor on nq q-points q[nq][3] (Cartesian, the phase is q . r) for site
 * positions pos[N][3] (Cartesian), measured when the measurement count (from 1) is a
 * multiple of every (sfactor_steps_measure); polar[3] (sfactor_polar) or NULL.  Positions
 * that form the supercell's lattice (esp_set_supercell, else a cube) get per-q phase tables
 * at this call; others are evaluated with one sincos per (q, site) and 8 replicas. */
int esp_set_sfactor(esp_ctx *ctx, int nq, const double *q, const double *pos, const double *polar,
                    int every);
/* Accumulators by slot (sums over the measurements; divide by the counts):
 *   ESP_OBS_ORDER, ESP_OBS_STAGGERED: out[R][3] = sum |X|, sum X^2, sum X^4
 *   ESP_OBS_PM:       out[R][ndim][nfreq] counts (as doubles) of the cell values per bin,
 *                     ndim = 1 + (order on) + (staggered on): m, then OP, then SM
 *   ESP_OBS_SPINCORR: out[R][3][nshell][nsp][nsp] = sum of spincorr_tmp, of |spincorr_tmp|,
 *                     of spincorr_abs_tmp (get_quant.F90:517-536)
 *   ESP_OBS_SFACTOR:  out[R][3][nq] = sum |F|^2, |F|^2 - NSF^2, NSF^2; *count = measurements
 * count receives the number of measurements that contributed (may be NULL). */
int esp_get_obs(esp_ctx *ctx, int which, double *out, int64_t *count);
/* Per-measurement |OP| and |SM| by slot while the trace is on: trace[n][R][2]. */
int esp_get_obs_trace(esp_ctx *ctx, double *trace, int *n);

/* Wait for the context's stream. */
int esp_sync(esp_ctx *ctx);
/* The context's HIP stream (hipStream_t) for callers that time or order work on it. */
void *esp_stream(esp_ctx *ctx);
/* Diagnostics, no counterpart in the reference: a plain device-to-device copy kernel
 * (k_copy) of n_bytes, launched reps times after one untimed launch; *ms = mean time per
 * launch.  Calibrates HBM traffic counters (FETCH_SIZE / WRITE_SIZE) and the copy
 * bandwidth the sweep kernels are compared with. */
int esp_copy_probe(esp_ctx *ctx, int64_t n_bytes, int reps, double *ms);

/* ---- learned proposal (no counterpart in the reference) --------------------------------
 * mode ESP_MODE_LEARNED in esp_sweep / esp_run draws each new spin from a two-component
 * von Mises-Fisher mixture whose parameters an MLP computes from the site's environment
 * (local field, temperature, biquadratic and single-ion projections), plus a fixed
 * uniform-sphere component of weight 1/32, and accepts with
 * min(1, exp(-beta dE) q(S_old) / q(S_new)).  The proposal depends on the neighbours only,
 * so detailed balance holds for any weights.  Global-memory path, one lane per site;
 * models with self bonds are refused.
 * Record layout (ESP_LEARN_RECORD floats): features x[5], kappa0 = beta |m| |G|,
 * ghat[3] (field direction), axis[3] (single-ion axis, else z), s[3] (spin direction),
 * replica * N + site (1 float; exact below 2^24).
 * Parameter layout: W1[32][5] b1[32] W2[32][32] b2[32] W3[32][32] b3[32] W4[6][32] b4[6];
 * outputs per component k: log kappa scale, weight logit, tilt (pi/2 tanh) toward the axis. */
/* FP32 master weights (copied to BF16 for the sweep); resets the Adam state. */
int esp_learn_set_weights(esp_ctx *ctx, const float *w, int nparams);
int esp_learn_get_weights(esp_ctx *ctx, float *w, int nparams);
/* Append n records of the current configurations: (replica, site) drawn from Philox key seed. */
int esp_learn_collect(esp_ctx *ctx, int n, uint64_t seed);
int esp_learn_clear(esp_ctx *ctx);
int esp_learn_num_records(esp_ctx *ctx, int *n);
int esp_learn_get_records(esp_ctx *ctx, int first, int n, float *recs);
/* Replace the record set with n host records. */
int esp_learn_put_records(esp_ctx *ctx, const float *recs, int n);
/* Gradient of the mean negative log-likelihood -log q(s | env) over records
 * [first, first+n) at the current master weights: grad[ESP_LEARN_NPARAMS], *loss. */
int esp_learn_grad(esp_ctx *ctx, int first, int n, float *grad, double *loss);
/* Adam (lr, beta 0.9 / 0.999, eps 1e-8) over the record set: epochs x shuffled minibatches
 * of `batch`; epoch_loss[epochs] (may be NULL) = mean loss seen in each epoch. */
int esp_learn_train(esp_ctx *ctx, int epochs, int batch, float lr, uint64_t seed,
                    double *epoch_loss);
/* Per record [n][10]: the six outputs, kappa_0, kappa_1, log q(s | env), w_0; from the FP32
 * master weights (sweep_net = 0) or from the network the sweep runs (sweep_net = 1: BF16
 * weights and hidden activations on the matrix cores, FP32 accumulation). */
int esp_learn_eval(esp_ctx *ctx, int first, int n, int sweep_net, double *out);

/* ---- batches of independent chains -------------------------------------------------- */
/* Many PT chains (lattices of different kind and size, each with its own ladder) advanced
 * together: one launch per colour index covers every chain, chains whose replica fits in
 * LDS run in a parallel graph branch, and a round of every chain is one hipGraph launch.
 * Chain k uses Philox stream base k << 16 (see esp_set_stream_base). */
typedef struct esp_batch esp_batch;
int esp_batch_create(esp_batch **b, int device);
int esp_batch_destroy(esp_batch *b);
const char *esp_batch_last_error(const esp_batch *b);
/* Arguments as esp_set_model + esp_set_supercell + esp_set_replicas; *chain_id receives
 * the chain's index.  Chains with identical bond structure share one colouring. */
int esp_batch_add_chain(esp_batch *b, int N, int nb, const double *mag,
                        int zJ, const int *j_cnt, const int *j_nbr, const double *j_val,
                        int zB, const int *b_cnt, const int *b_nbr, const double *b_val,
                        int zD, const int *d_cnt, const int *d_nbr, const double *d_vec,
                        int have_si, const double *si_par, const double *si_axis,
                        int have_field, const double *h_par, const double *h_axis,
                        const int *L3, int R, const double *tems, const double *spins,
                        int *chain_id);
/* Per-chain kernel shape: path 0 auto / 1 global / 2 LDS, tps lanes per site. */
int esp_batch_set_chain_config(esp_batch *b, int chain, int path, int tps);
/* Partner gathers in flight per lane on chain k's global-memory colour phases: 2 (default)
 * or 4 (more memory-level parallelism per wave, more registers).  A kernel shape with no
 * counterpart in the reference; the autotuner searches it. */
int esp_batch_set_chain_chunk(esp_batch *b, int chain, int chunk);
int esp_batch_get_chain_chunk(esp_batch *b, int chain, int *chunk);
/* Upload every chain (after the last esp_batch_add_chain). */
int esp_batch_commit(esp_batch *b);
/* Autotuner: for every distinct chain shape (largest share of the work first), try the
 * candidate kernel shapes (global or LDS-resident path x lanes per site x partner gathers in
 * flight per lane) and keep one only
 * when the measured time of a whole round of the batch (nsweeps rounds, at least 2) drops by
 * 0.5 % or more; the start is the global path with one lane per site.  The chains' state is
 * restored afterwards.  cache_path (may be NULL) is a text file of earlier decisions, read
 * first (shapes found there are not timed again) and rewritten. */
int esp_batch_autotune(esp_batch *b, int nsweeps, const char *cache_path);
/* esp_run for every chain at once (same schedule for all chains). */
int esp_batch_run(esp_batch *b, int n_rounds, int steps_per_round, int steps_measure,
                  int do_measure, int do_exchange, int start_parity, int mode, double tilt,
                  uint64_t seed, uint64_t sweep0, uint64_t step0, int use_graph);
int esp_batch_num_chains(esp_batch *b);
/* out[8]: N, R, colours, lanes per site, LDS path, bond width, kernel variant, LDS fits. */
int esp_batch_info(esp_batch *b, int chain, int *out8);
/* HIP-event timing of the batched colour-phase kernel over nsweeps sweeps of every chain on
 * the global-memory path (diagnostics for the roofline; no reference counterpart): summed
 * kernel time, number of launches, wall time first start -> last end, and the attempted
 * updates of those launches.  The chains advance (a Philox range from sweep0). */
int esp_batch_time_sweeps(esp_batch *b, int nsweeps, int mode, uint64_t seed, uint64_t sweep0,
                          double *kernel_ms, int *nlaunch, double *wall_ms, int64_t *updates);
int esp_batch_get_stats(esp_batch *b, int chain, double *acc, int64_t *rejected, double *e_run,
                        int64_t *nmeas);
int esp_batch_get_map(esp_batch *b, int chain, int *slot_rep, int64_t *nswap);
int esp_batch_get_spins(esp_batch *b, int chain, int by_slot, double *spins);
int esp_batch_energy(esp_batch *b, int chain, double *E);
int esp_batch_reset_stats(esp_batch *b);
int esp_batch_sync(esp_batch *b);

#ifdef __cplusplus
}
#endif
#endif

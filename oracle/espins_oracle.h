/*
 * espins_oracle.h -- CPU restatement of the ESpinS Metropolis / parallel-tempering
 * hot path.  TEST INFRASTRUCTURE ONLY: the product (libespins_hip.so) never links,
 * loads or calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the CPU timing leg.
 *
 * Array conventions follow the reference's Fortran arrays (column-major), seen
 * from C:
 *   spins        spin_matrix(3,N)        -> double s[N][3]
 *   J table      jij_matrix(N,zJ)        -> double j_val[zJ][N]   (slot-major)
 *   J neighbours jij_nbors_matrix(N,zJ)  -> int    j_nbr[zJ][N]   (1-based site ids)
 *   J counts     jij_num_nbors_matrix(N) -> int    j_cnt[N]
 *   D table      dij_matrix(3,N,zD)      -> double d_vec[zD][N][3]
 *   single ion   singleion_parameters(nb), singleion_axes_cart(3,nb)
 *   field        field_parameters(nb),     field_axes_cart(3,nb)
 * The basis index of 1-based site i is mod(i,nb) with 0 -> nb
 * (get_quant.F90:300-314), i.e. (i-1) % nb for a 0-based site.
 *
 * Parity status of each function is stated in its comment; "pinned" names the
 * check in tests/ that would fail on a plausible mistake.
 */
#ifndef ESPINS_ORACLE_H
#define ESPINS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- MT19937 (mtprng.f90:152-173, 227-282, 347-361) ---------------- */
typedef struct { int32_t mti; uint32_t mt[624]; } orc_mt_state;
void     orc_mt_init(orc_mt_state *st, int32_t seed);
uint32_t orc_mt_u32(orc_mt_state *st);
double   orc_mt_real2(orc_mt_state *st);            /* u32 * 2^-32, in [0,1) */
void     orc_mt_fill_u32(int32_t seed, int64_t n, uint32_t *out);

/* ---------------- Philox4x32-10 (Salmon et al., SC'11) ---------------- */
void   orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u01(uint32_t x);                          /* x * 2^-32 */

/* ---------------- Model ---------------- */
typedef struct {
  int N, nb;
  const double *mag;                                        /* N    */
  int zJ; const int *j_cnt; const int *j_nbr; const double *j_val;
  int zB; const int *b_cnt; const int *b_nbr; const double *b_val;
  int zD; const int *d_cnt; const int *d_nbr; const double *d_vec;
  int have_si;    const double *si_par, *si_axis;           /* nb, 3*nb */
  int have_field; const double *h_par,  *h_axis;            /* nb, 3*nb */
} orc_model;

/* energy_get (get_quant.F90:186-261) */
double orc_energy(const orc_model *m, const double *s);
/* The same terms grouped by site, e_site[N]: energy = (sum_i e_site[i]) / 2 (exact summation
 * gives the correctly rounded energy; the long sequential sum above rounds as the reference's). */
void orc_energy_sites(const orc_model *m, const double *s, double *e_site);
/* delta_energy_get (get_quant.F90:264-337); s already holds the proposed spin at site i0 */
double orc_delta_energy(const orc_model *m, int i0, const double *s, const double ds[3]);
/* magnetization_get (get_quant.F90:340-363): |m|^2 with m = sum_j S_j / N */
double orc_magnetization2(const orc_model *m, const double *s);

/* Proposals.  mode 0 = 'random' (get_quant.F90:80-93), 1 = 'const' (get_quant.F90:94-161). */
void orc_propose(int mode, double tilt, double mag, const double old[3],
                 double u1, double u2, double out[3]);

/* metropolis (get_quant.F90:52-182) in reference site order 1..N with MT19937
 * (mcarlo.F90:511-519).  nsweeps full sweeps. */
void orc_sweep_ref(const orc_model *m, double *s, double beta, int mode, double tilt,
                   orc_mt_state *st, int64_t nsweeps, double *e_run, int64_t *rejected);

/* Same update rule in colour order (colour c holds sites col_sites[col_off[c]..col_off[c+1]) ,
 * 0-based) with a Philox stream: counter (site0, sweep_lo, sweep_hi, stream), key (seed_lo, seed_hi);
 * u1 = out[0], u2 = out[1], u3 = out[2] of one Philox call per attempted update. */
void orc_sweep_colour_philox(const orc_model *m, int ncol, const int *col_off, const int *col_sites,
                             double *s, double beta, int mode, double tilt,
                             uint64_t seed, uint64_t sweep, uint32_t stream,
                             double *e_run, int64_t *rejected);

/* Host-fed deterministic mode: prop[N][3] proposed spins, unif[N] accept uniforms,
 * accepted[N] out (1/0).  Colour order as above. */
void orc_sweep_colour_hostfed(const orc_model *m, int ncol, const int *col_off, const int *col_sites,
                              double *s, double beta, const double *prop, const double *unif,
                              uint8_t *accepted, double *e_run, int64_t *rejected);

/* pt_swap (pt.F90:1235-1288), label form: configurations stay in place, slot_rep[t]
 * names the replica at temperature slot t (0-based, temperatures ascending).  energy is
 * per replica.  start is the reference's 1-based start (1 or 2).  The uniform of pair t
 * comes from the MT stream (u drawn only when f <= 0, like the reference) when st != NULL,
 * otherwise from Philox counter (t, step_lo, step_hi, 0xFFFFFFFF - stream), key (seed_lo ^ 0x5bd1e995, seed_hi). */
void orc_pt_swap(int start, int R, const double *tems, const double *energy, int *slot_rep,
                 int64_t *nswap /* [R][2]: attempts, accepts */, orc_mt_state *st,
                 uint64_t seed, uint64_t step, uint32_t stream);

/* Estimator accumulation (mcarlo.F90:536-543) and final formulas (mcarlo.F90:713-733). */
void orc_accumulate(double e, double m2, double acc[6]);
void orc_finalize(const double acc[6], int64_t nsteps, double T, int N, double out[7]);

/* Coupling-table builder (jij.F90:123-446, bij.F90:88-359, dij.F90:87-391).
 * Rows (f1,f2 fractional, value[, dvec unit vector]) are directed bonds from the atom
 * at f1 to the atom at f2 (cell offsets allowed).  basis_frac[nb][3] are the basis
 * positions in the species-sorted order the site index uses.  Site index (1-based):
 * ((n*L2 + m)*L1 + l)*nb + a + 1.  Values are divided by |m_i m_j| (moments per basis
 * atom) and zero bonds are dropped; ncomp = 1 (J, B) or 3 (D).  Output arrays have
 * capacity N*zcap; the realised z_max is returned (or -1 on an unknown position,
 * -2 on overflow of zcap). */
int orc_build_table(int nb, const double *basis_frac, const double *basis_mag,
                    const int L[3], int open_boundary,
                    int nrows, const double *f1, const double *f2, const double *value,
                    const double *dvec, int ncomp, int zcap,
                    int *cnt, int *nbr, double *val);

/* Greedy proper colouring of the union of the J, B and D graphs in site order
 * (a restatement used only to check the product's colouring). */
int orc_check_colouring(const orc_model *m, const int *colour);

#ifdef __cplusplus
}
#endif
#endif

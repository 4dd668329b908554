// esp_model.h -- host-side model preparation for the MI355X sweep kernels.
//
// The reference keeps three separate padded neighbour tables (J, B, D) in reference
// site order (jij.F90:394-446, bij.F90:314-359, dij.F90:343-391).  The device wants:
//  * one merged bond list per site (a neighbour that carries J, B and D is gathered once),
//  * sites renumbered colour-major so a colour phase reads and writes a contiguous range,
//  * bond coefficients collapsed to per-basis tables when the lattice is periodic,
//  * slot-major (site-fastest) layout so neighbour-index loads coalesce.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "esp_geo.h"

namespace esp {

// Coefficients of one merged bond: J, B, D (3).  Summed over duplicates of a pair.
constexpr int kCoef = 5;

struct HostModel {
  int N = 0, nb = 0;
  int zmax = 0;                 // merged bond width (padded)
  int ncol = 0;
  bool has_B = false, has_D = false, has_si = false, has_h = false;
  bool compressed = false;      // coefficients depend only on (basis, slot)
  bool has_self = false;        // some bond joins a site to itself (1-cell supercell axis)
  std::vector<int> colour;      // [N] by reference site (0-based)
  std::vector<int> col_off;     // [ncol+1] device ranges
  std::vector<int> dev2ref, ref2dev;
  std::vector<int> nbr;         // [zmax][N] device partner index (N = the zero spin)
  std::vector<double> coef;     // [zmax][N][kCoef] when !compressed
  std::vector<double> ctab;     // [nb][zmax][kCoef] when compressed
  std::vector<double> mag;      // [N] device order
  std::vector<uint8_t> basis;   // [N] device order
  std::vector<double> si4, h4;  // [nb][4]: parameter, axis x, y, z
  // Implicit geometry (esp_geo.h): partner and reference indices computed from the device
  // index.  Set only when the lattice is periodic, the coefficients are per (basis, slot) and
  // the computed indices equal the explicit tables for every site and slot.
  bool implicit = false;
  espg::Geo geo{};
  std::vector<int> grp;         // [ngroups] packed basis atom and residue
  std::vector<int> gslot;       // [ngroups][zmax][2]: partner group start, packed shift
  std::vector<double> magb;     // [nb] moment of each basis atom
  int rows_max = 0;             // groups one 256-lane block of a colour phase can touch
};

// Build the device-ready model from reference-layout tables.  Returns "" on success,
// else an error message.
std::string prepare_model(HostModel &hm, int N, int nb, const double *mag,
                          int zJ, const int *j_cnt, const int *j_nbr, const double *j_val,
                          int zB, const int *b_cnt, const int *b_nbr, const double *b_val,
                          int zD, const int *d_cnt, const int *d_nbr, const double *d_vec,
                          int have_si, const double *si_par, const double *si_axis,
                          int have_field, const double *h_par, const double *h_axis,
                          const int *Lhint /* supercell size or NULL */,
                          const std::vector<int> *colour_hint = nullptr);

// 64-bit FNV-1a over the bond structure (counts and partner lists) of a model.
uint64_t structure_hash(int N, int nb, int zJ, const int *j_cnt, const int *j_nbr, int zB,
                        const int *b_cnt, const int *b_nbr, int zD, const int *d_cnt,
                        const int *d_nbr);

// jij_get-style ELL expansion (see espins_hip.h, esp_build_table).
int build_table(int nb, const double *basis_frac, const double *basis_mag, const int *L,
                int open_boundary, int nrows, const double *f1, const double *f2,
                const double *value, const double *dvec, int ncomp, int zcap, int *cnt,
                int *nbr, double *val, int *zmax);

}  // namespace esp

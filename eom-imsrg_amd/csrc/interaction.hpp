// Host-side set-up of a physical Hamiltonian on the basis: the harmonic
// oscillator one-body part and the isospin-coupled two-body matrix elements
// of an me2j file (with the centre-of-mass and Lawson two-body terms).
//   h0_harm_osc   basic_IMSRG.f90:2093-2191 (calculate_h0_harm_osc)
//   read_me2j     interaction_IO.f90:282-557 (read_me2j_interaction)
//   p1_p2, r1_r2  basic_IMSRG.f90:4445-4565, reduced_p / reduced_r 4567-4657
#pragma once
#include <string>
#include <vector>

#include "basis.hpp"

namespace imsrg {

// The Hamiltonian's scalars that the reference keeps on the sq_op
// (hospace, lawson_beta, com_hw, Aprot + Aneut) and the Hamiltonian type
// (1: T - Tcm + V, 2: T + Uho + V, 3: T + V, 4: T / A, 5: Hcm one-body
// part without (com_hw/hw)^2).
struct HamParams {
  double hw = 20.0;
  int htype = 1;
  double lawson_beta = 0.0;
  double com_hw = 20.0;
  int mass = 16;
};

// fhh, fph, fpp of the flat operator `op` (host, oplen doubles) set to the
// harmonic-oscillator one-body part; other entries untouched.
std::string h0_harm_osc(const Basis& b, const HamParams& p, double* op);

// The file's orbit table: nfile entries in .sps order, proton (2tz = -1)
// and neutron entries of one (n, l, j) adjacent, proton first.
struct FileOrbits {
  std::vector<int> n, l, jj, tz;
  std::vector<int> con;   // basis orbit of each entry, -1 when the basis lacks it
};

// Checks the table and maps it onto the basis orbits.
std::string map_file_orbits(const Basis& b, int nfile, const int* fn, const int* fl, const int* fjj,
                            const int* ftz, FileOrbits& f);

// Number of values an me2j file over these orbits holds (4 per isospin
// quadruple and 2J).
long long me2j_count(const FileOrbits& f);

// The values of an me2j(.gz) file: one header line, then lines of up to
// ten fixed-width fields (f12.7, then f13.7 each).
std::string me2j_read_values(const char* path, long long count, std::vector<double>& out);

// Every two-body element the values give, written into `op` (host) in the
// reference's loop order (a later write of one element replaces an earlier).
std::string me2j_to_op(const Basis& b, const FileOrbits& f, const std::vector<double>& me, const HamParams& p,
                       double* op);

// b^2 <ab J| r1.r2 |cd J> and b^2/hbar^2 <ab J| p1.p2 |cd J>, antisymmetrised
// and normalised, HO length b.
double r1_r2(const Basis& b, int a, int bb, int c, int d, int J);
double p1_p2(const Basis& b, int a, int bb, int c, int d, int J);

}  // namespace imsrg

// Hartree-Fock basis of a vacuum Hamiltonian (calc_HF, HF_mod.f90:14-105,
// without three-body forces).  The self-consistent loop runs on the host over
// the single-particle (tz, l, j) blocks (a few orbits each); the two-body
// transform to the HF basis is two GEMMs per channel on the device (capi.hip).
#pragma once
#include <string>
#include <vector>

#include "basis.hpp"

namespace imsrg {

struct HFResult {
  std::vector<i64> boff;     // block q's matrices at [boff[q], boff[q+1]), n x n column-major
  std::vector<double> D;     // HF orbits: column i of block q = HF orbit i in the block's HO orbits
  std::vector<double> eig;   // single-particle energies, per block ascending (N values)
  double ehf = 0.0;          // e_HF
  int iterations = 0;
  // per channel: M = C - (-1)^J Cr over the (hh, ph, pp) pairs, nt x nt
  std::vector<i64> moff;     // channel q at [moff[q], moff[q+1])
};

// The SCF loop (HF_mod.f90:46-104) on the host copy `op` of the vacuum
// Hamiltonian (herm 1, all oplen doubles); on return E0 and the one-body part
// of `op` hold e_HF and the Fock matrix in the HF basis (transform_1b_to_HF,
// e_HF), and `M` the per-channel two-body transform matrices.
std::string hartree_fock_host(const Basis& b, double* op, int maxit, HFResult& r, std::vector<double>& M);

// The per-channel matrices M of the two-body transform by the HF orbits D
// (blocks at boff), channel q at [moff[q], moff[q+1]).
void transform_matrices(const Basis& b, const std::vector<i64>& boff, const std::vector<double>& D,
                        std::vector<i64>& moff, std::vector<double>& M);

// sqop_to_HF (HF_mod.f90:181-209) for a hermitian scalar operator, host
// part: the one-body part of `op` (host copy, all oplen doubles) rotated by
// the HF orbits D (IMSRG_SIZE_SPMAT doubles), and the two-body transform
// matrices; the caller transforms the two-body part and normal orders.
std::string sqop_to_hf_host(const Basis& b, const double* D, double* op, std::vector<i64>& moff,
                            std::vector<double>& M);

}  // namespace imsrg

// Host-side basis for the IMSRG(2) scalar-commutator path: orbit table,
// (2J, parity, Tz) two-body channels with their pp / ph / hh pair lists,
// the cross-coupled (Pandya) block layout and a dense 6j table, plus the
// device copies of all of it.  The channel, pair and cross-coupled orders
// are the reference's (they define the operator storage the Fortran caller
// hands over):
//   channels and pairs   basic_IMSRG.f90:521-685 (allocate_blocks)
//   channel index        basic_IMSRG.f90:1108-1114 (block_index)
//   sp blocks            basic_IMSRG.f90:340-380 (read_sp_basis)
//   cross-coupled blocks cross_coupled.f90:61-177 (allocate_CCMAT)
//   6j symbols           basic_IMSRG.f90:2227-2332, anglib.f:263-379
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace imsrg {

typedef long long i64;

// pair classes and the six gam blocks (gam1..gam6 = pppp, ppph, pphh, phph,
// hhhh, phhh; rows/cols by class, basic_IMSRG.f90:131-133 sea1/sea2/sqs)
constexpr int kRowCls[6] = {0, 0, 0, 1, 2, 1};
constexpr int kColCls[6] = {0, 1, 2, 1, 2, 2};
constexpr bool kSquare[6] = {true, false, false, true, true, false};

// Everything a kernel needs to address an operator, passed by value.
struct DevBasis {
  int N, A, P, JTM, nJ;            // nJ = JTM+1 = number of 2J/2 values of a pair
  int nblocks, ncc;
  i64 off_fhh, off_fph, off_fpp, oplen, cclen, wclen;
  const int* jj;                   // [N] 2j
  const int* ll;                   // [N]
  const int* tz;                   // [N] 2tz
  const int* con;                  // [N] 1 hole, 0 particle
  const int* cidx;                 // [N] index among holes / particles
  const int* holes;                // [A] hole orbits in orbit order
  const int* parts;                // [P] particle orbits in orbit order
  const int* spstart;              // [nsp+1] CSR of single-particle blocks
  const int* spmem;                // [N]
  const int* spblk;                // [N] orbit -> sp block
  const int* xmap;                 // [N*N*nJ] pair (i,j) at 2J -> index in class list (symmetric), -1
  const int* dims;                 // [nblocks*3] npp, nph, nhh
  const int* chJ;                  // [nblocks] 2J
  const i64* goff;                 // [nblocks*6] offset of gam g of channel q
  const int* pairs;                // packed (i | j<<16), by (q, class)
  const i64* pairoff;              // [nblocks*3]
  const int* rlen;                 // [ncc]
  const int* ncol;                 // [ncc]
  const int* ccJ;                  // [ncc]
  const i64* ccoff;                // [ncc] offset of CCX(q) (rlen x ncol, col-major)
  const i64* wcoff;                // [ncc] offset of WCC(q) (rlen x rlen, col-major)
  const int* rmap;                 // [N*N*nJ] ordered pair (i,j) at 2J -> CC row, -1
  const int* nbmap;                // [N*N*nJ] ordered (p,h) at 2J -> CC column, -1
  const int* rmapJ;                // [(J/2*N + a)*N + b] -> CC row of (a,b)
  const int* rmapJT;               // [(J/2*N + b)*N + a] -> CC row of (a,b)
  const int* ccrows;               // packed row pairs per CC block
  const i64* ccrowoff;             // [ncc]
  const int* cccols;               // packed column (p,h) pairs per CC block
  const i64* cccoloff;             // [ncc]
  const double* six;               // 6j table [j1/2][j2/2][j4/2][j5/2][J6/2][J3/2]
  int nh6, nJ6;
  // many-instance batches: blockIdx.y is the instance of a launch, and each
  // array of an instance follows the previous instance's at a stride of one
  // instance's length (oplen, cclen, wclen, N*N, and the lengths below)
  const int* php;                  // [2*nphp] (particle index, hole index) pairs in one (l, j, tz) block
  int nphp;
  int jsmax;                       // 1 + max 2j: bound on the J count of any orbit pair
  int ninst;                       // grid.y of batched launches (1: one instance)
  const int* act;                  // act[inst] == 0 skips the instance; null: all run
  i64 spart, spscr, sk2scr;        // partial sums, 222_pp_hh split-K scratch (P, K2)
};

struct Basis {
  // orbits
  int N = 0, A = 0, P = 0, JTM = 0, lmax = 0, nJ = 0;
  std::vector<int> nn, ll, jj, tz, con, cidx, holes, parts;
  // sp blocks
  std::vector<int> spstart, spmem, spblk;
  // two-body channels
  int nblocks = 0;
  std::vector<int> chJ, chT, chP, dims;
  std::vector<std::vector<int>> pl;          // [q*3+c] flattened (i,j)
  std::vector<int> pairs;                    // packed
  std::vector<i64> pairoff;
  std::vector<i64> goff;
  i64 off_fhh = 0, off_fph = 0, off_fpp = 0, oplen = 0;
  std::vector<int> xmap;
  // cross-coupled
  int ncc = 0;
  std::vector<int> ccJ, ccT, ccP, rlen, ncol;
  std::vector<i64> ccoff, wcoff, ccrowoff, cccoloff;
  std::vector<int> ccrows, cccols, rmap, nbmap, rmapJ, rmapJT;
  i64 cclen = 0, wclen = 0;
  // 6j
  int nh6 = 0, nJ6 = 0;
  std::vector<double> six;

  int rows(int q, int g) const { return dims[q * 3 + kRowCls[g]]; }
  int cols(int q, int g) const { return dims[q * 3 + kColCls[g]]; }
  i64 blocksize(int q, int g) const { return (i64)rows(q, g) * cols(q, g); }
};

// Builds every host table; returns an empty string or an error message.
std::string build_basis(Basis& b, int N, const int* n, const int* l, const int* jj,
                        const int* tz, const int* hole);

// 6j symbol {j1 j2 j3; j4 j5 j6} (all arguments twice the angular momentum)
double sixj_value(int j1, int j2, int j3, int j4, int j5, int j6);

}  // namespace imsrg

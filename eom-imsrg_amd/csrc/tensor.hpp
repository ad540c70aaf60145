// Tensor operators (rank > 0) on the device: layout tables and the kernels of
// the scalar-tensor commutator [L, R] (L scalar, R a tensor of rank
// lambda = rank/2 and parity change dpar/2), TS_commutators.f90:10-1849.
//
// Flat tensor layout (allocate_tensor, basic_IMSRG.f90:688-951):
//   [E0 | fhh (A x A) | fph (P x A) | fpp (P x P) | per block q:
//    tgam1..tgam9, tgam g = (pairs of class sea1(g) in channel (J1,Tz,P1)) x
//    (pairs of class sea2(g) in channel (J2,Tz,P2)), column-major]
// blocks in the reference's loop order (2J1 outer, 2J2 = max(|2J1 - rank|,
// 2J1) .. 2J1 + rank, Tz = -1..1, P1 = 0..1); blocks with 2J2 > 2 jmax are
// empty.  The one-body offsets are the scalar operator's.
#pragma once

#include "basis.hpp"
#include "kernels.hpp"

namespace imsrg {

// pair class (0 pp, 1 ph, 2 hh) of the rows / columns of tgam g (sea1 / sea2,
// basic_IMSRG.f90:127)
constexpr int kTSea1[9] = {0, 0, 0, 1, 2, 1, 2, 1, 2};
constexpr int kTSea2[9] = {0, 1, 2, 1, 2, 2, 0, 0, 1};

struct DevTensor {
  int rank, dpar, nblocks, ncc;
  i64 oplen, off2b;                // doubles of an operator; first two-body element
  const int* blk;                  // [nblocks*8] 2J1 2J2 Tz P1 P2 q1 q2 -
  const int* tdim;                 // [nblocks*6] np1 nb1 nh1 np2 nb2 nh2
  const i64* tgoff;                // [nblocks*9]
  const int* look;                 // [((2J1/2)*(rank/2+1) + (2J2-2J1)/2)*6 + 2(Tz+1) + P1] -> q
  const int* cblk;                 // [ncc*8] 2J3 2J4 Tz PAR r1 r2 nb1 nb2
  const int* cscal;                // [ncc*2] scalar cross-coupled blocks of the rows (J3) and columns (J4)
  const i64* coff;                 // [ncc*3] CCX, CCR, W offsets
  const int* cclook;               // [(2J3/2)*(rank/2+1) + (2J4-2J3)/2] -> qx of (J3, J4, Tz 0, PAR 0)
  const double* x6;                // {J1 J2 rank; j4 j5 j6}: [J1/2][J2/2][j4/2][j5/2][j6/2]
  const double* s6;                // {c d J2; b x J4}:       [c/2][d/2][b/2][x/2][J2/2][J4/2]
  const i64* wtab;                 // [(((2J3/2)*(rank/2+1) + (2J4-2J3)/2)*4 + 2 Tz + PAR)*2]: offset of the
                                   // product W of that cross-coupled block, its leading dimension
  const double* n9;                // 9j {u1 u2 J3; u4 u5 J4; J1 J2 rank} * sqrt((J3+1)(J4+1)) of the
                                   // 222_ph back-transform, J3 <= J4 <= J3 + rank, J1 <= J2 <= J1 + rank:
                                   // [u1/2][u2/2][u4/2][u5/2][J3/2][(J4-J3)/2][J1/2][(J2-J1)/2]
                                   // (the outputs of one orbit quadruple read one contiguous run)
  int nJ, nhx, nho;                // 2J/2 values; half-integers up to jmax + rank; up to jmax
};

// one sub-block g of block q and its pair lists
struct TSub { int q, g, n1, n2; i64 off, roff, coff; };
// 256 consecutive elements of sub-block sb
struct TElem { int sb; int start; };
// one output element of the 222_ph back-transform: its offset, its orbits
// (a | b << 8 | c << 16 | d << 24) and 2J1 | 2J2 << 16; the list is ordered
// by orbit quadruple, so the lanes of a wave share the quadruple's loops
struct TPh { int off, abcd, jj; };
// one-body output: cls 0 fhh (p <= q, mirrored), 1 fpp (p <= q, mirrored), 2 fph
struct TOne { int cls, p, q, pad; };
// 256 consecutive elements of CCX (term 1) or CCR (term 2) of block qx
struct TCC { int qx, term, start, pad; };

void launch_ts111(const DevBasis& B, const DevTensor& T, const double* L, int Lh, const double* R, int Rh,
                  double* RES, hipStream_t s);
// mode 0: TS_commutator_121, 1: TS_commutator_211 (X = the scalar
// cross-coupled set of L), 2: TS_commutator_221 (X = w1, Y = w2, pm = Lh)
void launch_ts_one(int mode, const DevBasis& B, const DevTensor& T, const TOne* t, int nt, const double* L, int Lh,
                   const double* X, const double* Y, const double* R, int Rh, double* RES, int RESh, hipStream_t s);
// TS_commutator_122 (assigns) over element tiles
void launch_ts122(const DevBasis& B, const DevTensor& T, const TSub* sb, const TElem* t, int nt, const double* L,
                  int Lh, const double* R, int Rh, double* RES, hipStream_t s);
// TS_commutator_212 (adds) over the quadruple-ordered output list
void launch_ts212(const DevBasis& B, const DevTensor& T, const TPh* t, i64 nt, const double* L, int Lh,
                  const double* R, int Rh, double* RES, hipStream_t s);
// the back-transform of TS_commutator_222_ph from the products Wx / Wy, one
// thread per listed output element
void launch_ts_ph(const DevBasis& B, const DevTensor& T, const TPh* t, i64 nt, int Lh, int Rh, const double* Wx,
                  const double* Wy, double* RES, hipStream_t s);
// calculate_single_pandya of every listed (block, term) into CC
void launch_ts_pandya(const DevBasis& B, const DevTensor& T, const TCC* t, int nt, const double* R, int Rh,
                      double* CC, hipStream_t s);
// RES[x] += w1[x] - w2[x] over [off, off + n)
void launch_ts_epilogue(i64 off, i64 n, const double* w1, const double* w2, double* RES, hipStream_t s);
// weighted sum of squares over segments (off, len, weight): part[k] per
// chunk k of 4096 elements (chunk table ck: segment offset, length, weight),
// then dst = sqrt(sum part) in chunk order
void launch_wnorm(const i64* ck_off, const int* ck_len, const double* ck_w, int nck, const double* op, double* part,
                  double* dst, hipStream_t s);

}  // namespace imsrg

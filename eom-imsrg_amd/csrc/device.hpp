// Device-side accessors shared by the kernels: one-body and two-body element
// fetches with the reference's phase / normalisation conventions, the 6j
// table lookup and small index helpers.
#pragma once
#include <hip/hip_runtime.h>

#include "basis.hpp"

namespace imsrg {

typedef double v4d __attribute__((ext_vector_type(4)));

// 1/sqrt(2) exactly as the reference types it (commutators.f90:834,847)
#define IMSRG_INV_SQRT2_REF 0.70710678118

__device__ __forceinline__ int m1pow(int e) { return (e & 1) ? -1 : 1; }
__device__ __forceinline__ bool dtri(int a, int b, int c) { return c <= a + b && c >= abs(a - b); }
__device__ __forceinline__ int row_cls(int g) { return g < 3 ? 0 : (g == 4 ? 2 : 1); }
__device__ __forceinline__ int col_cls(int g) { return g == 0 ? 0 : (g == 1 || g == 3 ? 1 : 2); }
__device__ __forceinline__ bool is_square(int g) { return g == 0 || g == 3 || g == 4; }
__device__ __forceinline__ int brows(const DevBasis& B, int q, int g) { return B.dims[q * 3 + row_cls(g)]; }
__device__ __forceinline__ int bcols(const DevBasis& B, int q, int g) { return B.dims[q * 3 + col_cls(g)]; }
__device__ __forceinline__ void unpack(int p, int& a, int& b) { a = p & 0xffff; b = p >> 16; }

// f_elem, basic_IMSRG.f90:1191-1221
__device__ __forceinline__ double f_elem(const DevBasis& B, const double* op, int herm, int a, int c) {
  const int c1 = B.con[a], c2 = B.con[c];
  if (c1 + c2 == 0) return op[B.off_fpp + B.cidx[a] + (i64)B.cidx[c] * B.P];
  if (c1 + c2 == 2) return op[B.off_fhh + B.cidx[a] + (i64)B.cidx[c] * B.A];
  if (c1 > c2) return op[B.off_fph + B.cidx[c] + (i64)B.cidx[a] * B.P] * herm;
  return op[B.off_fph + B.cidx[a] + (i64)B.cidx[c] * B.P];
}

// Position of the two-body element <ab|V|cd>_J: the lookup half of v_elem
// (basic_IMSRG.f90:1314-1429).  Reading it from two operators gives the
// reference's v_elem / v_same pair (basic_IMSRG.f90:2051-2072).
struct VPos {
  i64 off;      // < 0: element is zero by symmetry
  double pre;   // phase (-1)^(1+(ja+jb-J)/2) for swapped pairs, sqrt(2) for a=b
  int swp;      // stored as the transposed element times herm
};

__device__ __forceinline__ VPos vpos(const DevBasis& B, int a, int b, int c, int d, int J) {
  VPos p;
  p.off = -1; p.pre = 0.0; p.swp = 0;
  const int ja = B.jj[a], jb = B.jj[b], jc = B.jj[c], jd = B.jj[d];
  if (!(dtri(ja, jb, J) && dtri(jc, jd, J))) return p;
  const int P = (B.ll[a] + B.ll[b]) & 1;
  if (((B.ll[c] + B.ll[d]) & 1) != P) return p;
  const int T2 = B.tz[a] + B.tz[b];
  if (B.tz[c] + B.tz[d] != T2) return p;
  const int q = 3 * J + 2 * (T2 / 2 + 1) + P;
  const int C1 = B.con[a] + B.con[b], C2 = B.con[c] + B.con[d];
  const int lo = C1 < C2 ? C1 : C2, hi = C1 < C2 ? C2 : C1;
  // (lo,hi): (0,0) gam1, (0,1) gam2, (0,2) gam3, (1,1) gam4, (2,2) gam5, (1,2) gam6
  int g;
  if (lo == 0) g = hi;
  else if (lo == 1) g = (hi == 1) ? 3 : 5;
  else g = 4;
  double pre = 1.0;
  const int i1 = B.xmap[((i64)a * B.N + b) * B.nJ + (J >> 1)];
  if (a > b) pre = (double)m1pow(1 + (ja + jb - J) / 2);
  else if (a == b) pre = pre * 1.4142135623730951;
  const int i2 = B.xmap[((i64)c * B.N + d) * B.nJ + (J >> 1)];
  if (c > d) pre = pre * m1pow(1 + (jc + jd - J) / 2);
  else if (c == d) pre = pre * 1.4142135623730951;
  const int ld = B.dims[q * 3 + row_cls(g)];
  if (C1 > C2) { p.off = B.goff[q * 6 + g] + i2 + (i64)i1 * ld; p.swp = 1; }
  else { p.off = B.goff[q * 6 + g] + i1 + (i64)i2 * ld; p.swp = 0; }
  p.pre = pre;
  return p;
}

__device__ __forceinline__ double vat(const VPos& p, const double* op, int herm) {
  if (p.off < 0) return 0.0;
  double x = op[p.off];
  if (p.swp) x = x * herm;
  return x * p.pre;
}

__device__ __forceinline__ double v_elem(const DevBasis& B, const double* op, int herm, int a, int b, int c,
                                         int d, int J) {
  return vat(vpos(B, a, b, c, d, J), op, herm);
}

// sixj, basic_IMSRG.f90:2553-2599: j1,j2,j4,j5 half-integer; j3,j6 integer
__device__ __forceinline__ double sixj(const DevBasis& B, int j1, int j2, int j3, int j4, int j5, int j6) {
  if (!(dtri(j1, j2, j3) && dtri(j4, j5, j3) && dtri(j1, j5, j6) && dtri(j4, j2, j6))) return 0.0;
  const int nh = B.nh6;
  const i64 idx = (((((i64)(j1 >> 1) * nh + (j2 >> 1)) * nh + (j4 >> 1)) * nh + (j5 >> 1)) * B.nJ6 + (j6 >> 1)) *
                      B.nJ6 + (j3 >> 1);
  return B.six[idx];
}

// sum over the 64 lanes of a wave (every lane receives the total)
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// instance of a batched launch; false when it is masked off (the whole
// workgroup returns together)
__device__ __forceinline__ bool batch_inst(const DevBasis& B, int& inst) {
  inst = blockIdx.y;
  return !(B.act && !B.act[inst]);
}
__device__ __forceinline__ bool batch_on(const int* act, int& inst) {
  inst = blockIdx.y;
  return !(act && !act[inst]);
}

// commutator_122 partner terms (commutators.f90:283-396): the stored block
// (g index) of classes lo <= hi of one channel
__device__ __forceinline__ int blk_of(int lo, int hi) { return lo == 0 ? hi : (lo == 1 ? (hi == 1 ? 3 : 5) : 4); }

// Addressing of the partner terms of one output element (IX, JX) of block
// (rc, cc): a row-partner term at class cl reads the full channel matrix at
// (partner, JX), a column-partner term at (IX, partner).  The stored block
// holding that element and the element's offset are linear in the partner
// index: off = base[cl] + idx * stride[cl], with the herm factor sg[cl]
// (1 in a stored block, Lh / Rh in the transposed one).  Base and stride
// depend on the thread's IX / JX only, so they are formed once per thread
// and each term costs its partner record and the two operand loads.
// Offsets are 32-bit (an operator holds < 2^31 doubles, checked when the
// plan is built).
struct PartnerAddr {
  int b0, b1, b2;   // base per class
  int s0, s1, s2;   // stride per class
  int t0, t1, t2;   // 1: element in the transposed stored block
};
__device__ __forceinline__ int pick3(int cl, int v0, int v1, int v2) { return cl == 0 ? v0 : (cl == 1 ? v1 : v2); }
__device__ __forceinline__ i64 pa_off(const PartnerAddr A, int cl, int idx) {
  return (i64)(unsigned)(pick3(cl, A.b0, A.b1, A.b2) + idx * pick3(cl, A.s0, A.s1, A.s2));
}
__device__ __forceinline__ int pa_trans(const PartnerAddr A, int cl) { return pick3(cl, A.t0, A.t1, A.t2); }

// the row- and column-partner addressing of output (IX, JX) of the block of
// classes (rc, cc), from the channel's block offsets go[6] and class
// dimensions dm[3]
__device__ __forceinline__ void partner_addr(const i64* go, const int* dm, int rc, int cc, int IX, int JX,
                                             PartnerAddr& AR, PartnerAddr& AC) {
  // row partner at class cl: column JX of class cc; column partner at class
  // cl: row IX of class rc
#define IMSRG_PA(cl, bR, sR, tR, bC, sC, tC)                                                  \
  if (cl <= cc) { bR = (int)(go[blk_of(cl, cc)] + (i64)JX * dm[cl]); sR = 1; tR = 0; }       \
  else { bR = (int)(go[blk_of(cc, cl)] + JX); sR = dm[cc]; tR = 1; }                          \
  if (rc <= cl) { bC = (int)(go[blk_of(rc, cl)] + IX); sC = dm[rc]; tC = 0; }                 \
  else { bC = (int)(go[blk_of(cl, rc)] + (i64)IX * dm[cl]); sC = 1; tC = 1; }
  IMSRG_PA(0, AR.b0, AR.s0, AR.t0, AC.b0, AC.s0, AC.t0)
  IMSRG_PA(1, AR.b1, AR.s1, AR.t1, AC.b1, AC.s1, AC.t1)
  IMSRG_PA(2, AR.b2, AR.s2, AR.t2, AC.b2, AC.s2, AC.t2)
#undef IMSRG_PA
}

// (row, col) of element t of the upper triangle (row <= col) of an n x n
// block, enumerated column by column.
__device__ __forceinline__ void upper_decode(i64 t, int& row, int& col) {
  int c = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((i64)c * (c + 1) / 2 > t) --c;
  while ((i64)(c + 1) * (c + 2) / 2 <= t) ++c;
  col = c;
  row = (int)(t - (i64)c * (c + 1) / 2);
}

}  // namespace imsrg

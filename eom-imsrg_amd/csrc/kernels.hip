// HIP kernels of the IMSRG(2) scalar-commutator path for gfx950 (CDNA4).
//
//  * k_comm122p, k_comm222ph_scatter, k_pandya: per-element gathers, one
//    thread per output element, 256-element tiles from a task list that
//    spans every channel (grouped) or one channel (plain path);
//    k_comm222ph_chunk: the grouped inverse Pandya by quadruple chunks.
//  * the FP64 MFMA GEMMs live in gemm.hip; k_pppp_sym applies the
//    reference's  W - pm W^T  epilogue of 222_pp_hh to the pppp product.
//  * one-body terms, norms, axpy, generator, mbpt2 and normal ordering.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"

namespace imsrg {

// ============================================================ commutator_122
// commutators.f90:283-396.  RES two-body = 1b x 2b part (assigned): for
// every stored element (ab,cd) at J, the sums over the (l,j,tz) block
// partners i of a, b, c, d of f(L)*v(R) - f(R)*v(L), then divided by
// sqrt(1+delta_ab) sqrt(1+delta_cd).
// The partner lists are precomputed per stored pair (a-, b-, c-, d-replaced,
// in that order), so no term performs an index search.

// one partner term: f (from LDS) times the operand elements r = R[o],
// l = L[o], accumulated as sm + fl*vr - fr*vl (row partners) or
// sm - fl*vr + fr*vl (column partners), vr = r * pre * (Rh if the element
// is read from the transposed block), vl = l * pre * (Lh likewise), fl / fr
// times Lh / Rh when the record's one-body element is read transposed.
// Those herm factors and the record's sign are +-1: they combine into one
// sign bit per product, applied to pre = pre_other * (1 or sqrt 2)
// (pa1 / pa2).  lneg / rneg: Lh / Rh < 0.
template <bool ROW>
__device__ __forceinline__ double term122(double sm, const Partner p, double r, double l, const PartnerAddr A,
                                          double pa1, double pa2, const double* s_f, int nf1, int lneg, int rneg) {
  const int code = p.code;
  const int tr = pa_trans(A, (code >> 22) & 3);
  const int fs = (code >> 24) & 1, neg = (code >> 25) & 1;
  const double a = ((code >> 26) & 1) ? pa2 : pa1;
  const int sr = neg ^ (fs & lneg) ^ (tr & rneg);   // sign of the R-operand product
  const int sl = neg ^ (fs & rneg) ^ (tr & lneg);   // sign of the L-operand product
  const double cr = (sr ^ (ROW ? 0 : 1)) ? -a : a, cl = (sl ^ (ROW ? 1 : 0)) ? -a : a;
  sm = fma(s_f[p.fk] * cr, r, sm);
  return fma(s_f[nf1 + p.fk] * cl, l, sm);
}

// the n partner terms of one list, record k at part[k * ps], IMSRG_122_UNROLL
// records and their operand loads in flight at a time; the terms are added
// in list order
#ifndef IMSRG_122_UNROLL
#define IMSRG_122_UNROLL 2
#endif
template <bool ROW>
__device__ __forceinline__ double sum122(double sm, const Partner* part, i64 ps, int n, const PartnerAddr A,
                                         double pa1, double pa2, const double* s_f, int nf1, const double* L,
                                         int lneg, const double* R, int rneg) {
  constexpr int U = IMSRG_122_UNROLL;
  int x = 0;
  for (; x + U - 1 < n; x += U) {
    Partner p[U];
    double r[U], l[U];
#pragma unroll
    for (int u = 0; u < U; ++u) p[u] = part[(x + u) * ps];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const i64 o = pa_off(A, (p[u].code >> 22) & 3, p[u].code & 0x3fffff);
      r[u] = R[o];
      l[u] = L[o];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) sm = term122<ROW>(sm, p[u], r[u], l[u], A, pa1, pa2, s_f, nf1, lneg, rneg);
  }
  for (; x < n; ++x) {
    const Partner p = part[x * ps];
    const i64 o = pa_off(A, (p.code >> 22) & 3, p.code & 0x3fffff);
    sm = term122<ROW>(sm, p, R[o], L[o], A, pa1, pa2, s_f, nf1, lneg, rneg);
  }
  return sm;
}

__global__ __launch_bounds__(256) void k_comm122p(DevBasis B, const SegTask* tasks, const Partner* part,
                                                  const Partner* partT, int pstride, const int* rowp,
                                                  const int* colp, const int* f1off, int nf1,
                                                  const double* L, int Lh, const double* R, int Rh, double* RES,
                                                  int RESh, int mirror) {
  int inst;
  if (!batch_inst(B, inst)) return;
  L += (i64)inst * B.oplen;
  R += (i64)inst * B.oplen;
  RES += (i64)inst * B.oplen;
  // the one-body elements the partner terms read, L then R, staged once
  extern __shared__ double s_f[];
  for (int k = threadIdx.x; k < nf1; k += blockDim.x) {
    const int o = f1off[k];
    s_f[k] = L[o];
    s_f[nf1 + k] = R[o];
  }
  __syncthreads();
  const SegTask t = tasks[xcd_contig(blockIdx.x, gridDim.x)];   // neighbouring tiles share partner rows
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= t.cnt) return;
  const int q = t.q, g = t.g;
  const int rc = row_cls(g), cc = col_cls(g);
  const int n1 = B.dims[q * 3 + rc];
  const bool sq = is_square(g);
  int IX, JX;
  if (sq) upper_decode(e, IX, JX);
  else { IX = (int)(e % n1); JX = (int)(e / n1); }
  const i64 gr = B.pairoff[q * 3 + rc] + IX, gc = B.pairoff[q * 3 + cc] + JX;
  int a, b, c, d;
  unpack(B.pairs[gr], a, b);
  unpack(B.pairs[gc], c, d);
  const double pre_ab = (a == b) ? 1.4142135623730951 : 1.0;
  const double pre_cd = (c == d) ? 1.4142135623730951 : 1.0;
  PartnerAddr AR, AC;
  partner_addr(B.goff + q * 6, B.dims + q * 3, rc, cc, IX, JX, AR, AC);   // uniform (scalar) loads
  double sm = 0.0;
  const int lneg = Lh < 0, rneg = Rh < 0;
  sm = sum122<true>(sm, partT + gr, pstride, rowp[gr + 1] - rowp[gr], AR, pre_cd, pre_cd * 1.4142135623730951, s_f,
                    nf1, L, lneg, R, rneg);
  // column partners: when every active lane of the wave holds an element of
  // one column pair (the common case), the list is read through the scalar
  // cache instead of once per lane
  const int gc0 = __builtin_amdgcn_readfirstlane((int)gc);
  if (__all((int)gc == gc0)) {
    const int x0 = colp[gc0], nx = colp[gc0 + 1] - x0;
    sm = sum122<false>(sm, part + x0, 1, nx, AC, pre_ab, pre_ab * 1.4142135623730951, s_f, nf1, L, lneg, R, rneg);
  } else {
    sm = sum122<false>(sm, part + colp[gc], 1, colp[gc + 1] - colp[gc], AC, pre_ab, pre_ab * 1.4142135623730951, s_f,
                       nf1, L, lneg, R, rneg);
  }
  sm = sm / sqrt(1.0 + (a == b ? 1.0 : 0.0)) / sqrt(1.0 + (c == d ? 1.0 : 0.0));
  const i64 base = B.goff[q * 6 + g];
  RES[base + IX + (i64)JX * n1] = sm;
  if (sq && mirror) RES[base + JX + (i64)IX * n1] = sm * RESh;
}

void launch_comm122p(const DevBasis& B, const SegTask* t, int nt, const Partner* part, const Partner* partT,
                     int pstride, const int* rowp, const int* colp, const int* f1off, int nf1, const double* L, int Lh,
                     const double* R, int Rh, double* RES, int RESh, hipStream_t s, int mirror) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_comm122p, dim3(nt, B.ninst), dim3(kTile), 2 * nf1 * sizeof(double), s, B, t, part, partT,
                     pstride, rowp, colp, f1off, nf1, L, Lh, R, Rh, RES, RESh, mirror);
}

// ================================================== commutator_222_ph scatter
// commutators.f90:804-923.  Every J' term of the reference reads WCC and
// its transpose only in the combination
//     W'(x,y) = Lh W(x,y) - Rh W(y,x),    W = LCC RCC^T,
// inside:  (W(gik,rjl) + ph W(gjl,rik)) Lh - (ph W(rik,gjl) + W(rjl,gik)) Rh
//        =  W'(gik,rjl) + ph W'(gjl,rik)
// outside: W'(gil,rjk) + ph W'(gjk,ril).
// W' satisfies W'(y,x) = -Lh Rh W'(x,y), so the GEMM stores its upper
// triangle only and the scatter reads the element (min, max).  (Storing
// both triangles, for reads along columns only, costs the GEMM more in
// transposed stores than it saves here.)

// sixj for arguments whose four triangle conditions hold by construction
// (the J' ranges below), so the table read is unconditional
__device__ __forceinline__ double sixj_tri(const DevBasis& B, int j1, int j2, int j3, int j4, int j5, int j6) {
  const int nh = B.nh6;
  const i64 idx = (((((i64)(j1 >> 1) * nh + (j2 >> 1)) * nh + (j4 >> 1)) * nh + (j5 >> 1)) * B.nJ6 + (j6 >> 1)) *
                      B.nJ6 + (j3 >> 1);
  return B.six[idx];
}

// One recoupling's sum over J' (jmin..jmax step 2), added onto sm:
//   inside:  - (-1)^((jl+jk+J)/2) (W'(gik,rjl) + phase W'(gjl,rik)) {ji jj J; jl jk J'}
//   outside: + (W'(gil,rjk) + phase W'(gjk,ril)) {ji jj J; jk jl J'}
// ph_pair gives the W' combination of one J' term.
// JFAST: the row tables read with J' fastest (rmap; the grouped form runs
// the consecutive J' terms of one quadruple in neighbouring lanes, which
// then read neighbouring entries), else with J' slowest (rmapJ / rmapJT;
// the per-channel form runs neighbouring i in neighbouring lanes)
template <bool INSIDE, bool JFAST>
__device__ __forceinline__ double ph_pair(const DevBasis& B, const i64* wcoff, const int* rlen, const double* WCC,
                                          double mpm, int phase, int i, int j, int k, int l, int JP, int Tz, int PAR) {
  // inside reads rows (k,i) / cols (j,l); outside rows (l,i) / cols (j,k).
  // 32-bit element offsets from uniform bases (WCC holds < 2^32 doubles,
  // checked at plan time)
  const int N = B.N, JTM = B.JTM, nJ = B.nJ, h = JP >> 1;
  const int a = INSIDE ? k : l, c = INSIDE ? l : k;
  const int qx = h + Tz * (JTM + 1) + 2 * PAR * (JTM + 1);
  const unsigned wo = (unsigned)wcoff[qx], rr = (unsigned)rlen[qx];
  unsigned x1, y1, x2, y2;
  if (JFAST) {
    x1 = B.rmap[(a * N + i) * nJ + h];   // gik | gil
    y1 = B.rmap[(j * N + c) * nJ + h];   // rjl | rjk
    x2 = B.rmap[(i * N + a) * nJ + h];   // rik | ril
    y2 = B.rmap[(c * N + j) * nJ + h];   // gjl | gjk
  } else {
    const unsigned t = (unsigned)h * (unsigned)(N * N);
    x1 = B.rmapJ[t + a * N + i];
    y1 = B.rmapJ[t + j * N + c];
    x2 = B.rmapJT[t + a * N + i];
    y2 = B.rmapJ[t + c * N + j];
  }
  // W'(x1, y1) + phase W'(y2, x2), each read from the upper triangle
  const double A1 = x1 <= y1 ? WCC[wo + x1 + y1 * rr] : mpm * WCC[wo + y1 + x1 * rr];
  const double A2 = y2 <= x2 ? WCC[wo + y2 + x2 * rr] : mpm * WCC[wo + x2 + y2 * rr];
  return A1 + phase * A2;
}

template <bool INSIDE>
__device__ __forceinline__ double ph_jsum(const DevBasis& B, const double* WCC, double mpm, int i, int j, int k, int l,
                                          int Jtot, int jmin, int jmax, int phase, int Tz, int PAR, double sm) {
  const int ji = B.jj[i], jj = B.jj[j], jk = B.jj[k], jl = B.jj[l];
  const double sg = INSIDE ? -(double)m1pow((jl + jk + Jtot) / 2) : 1.0;
  for (int JP = jmin; JP <= jmax; JP += 2) {
    const double v = ph_pair<INSIDE, false>(B, B.wcoff, B.rlen, WCC, mpm, phase, i, j, k, l, JP, Tz, PAR);
    sm = sm + sg * v * (INSIDE ? sixj_tri(B, ji, jj, Jtot, jl, jk, JP) : sixj_tri(B, ji, jj, Jtot, jk, jl, JP));
  }
  return sm;
}

__device__ __forceinline__ void ph_element(const DevBasis& B, const double* WCC, int Lh, int Rh, double* RES,
                                           int RESh, int q, int g, int IX, int JX, int i, int j, int k, int l,
                                           bool mirror = true) {
  const int n1 = brows(B, q, g);
  const bool sq = is_square(g);
  const double pre = (i == j) ? IMSRG_INV_SQRT2_REF : 1.0;
  const double pre2 = (k == l) ? IMSRG_INV_SQRT2_REF : 1.0;
  const double mpm = -(double)(Lh * Rh);
  const int Jtot = B.chJ[q];
  const int ji = B.jj[i], jj = B.jj[j], jk = B.jj[k], jl = B.jj[l];
  const int li = B.ll[i], lj = B.ll[j], lk = B.ll[k], ll = B.ll[l];
  const int ti = B.tz[i], tj = B.tz[j], tk = B.tz[k], tl = B.tz[l];
  const int phase = m1pow((ji + jj + jk + jl) / 2);
  double sm = 0.0;
  // "inside" recoupling
  int Tz = abs(ti - tk) / 2;
  if (abs(tl - tj) != Tz * 2) return;
  int PAR = (li + lk) % 2;
  if ((ll + lj) % 2 != PAR) return;
  sm = ph_jsum<true>(B, WCC, mpm, i, j, k, l, Jtot, max(abs(jj - jl), abs(ji - jk)),
                                min(jj + jl, ji + jk), phase, Tz, PAR, sm);
  // "outside" recoupling
  Tz = abs(ti - tl) / 2;
  if (abs(tk - tj) != Tz * 2) return;
  PAR = (li + ll) % 2;
  if ((lk + lj) % 2 != PAR) return;
  sm = ph_jsum<false>(B, WCC, mpm, i, j, k, l, Jtot, max(abs(ji - jl), abs(jj - jk)),
                                 min(ji + jl, jj + jk), phase, Tz, PAR, sm);
  double* X = RES + B.goff[q * 6 + g];
  const i64 o = IX + (i64)JX * n1;
  X[o] = X[o] + sm * pre * pre2;
  if (sq && mirror) X[JX + (i64)IX * n1] = X[o] * RESh;
}

// per-channel form: segment tiles of one channel (plain path)
__global__ __launch_bounds__(256) void k_comm222ph_scatter(DevBasis B, const SegTask* tasks, const double* WCC,
                                                           int Lh, int Rh, double* RES, int RESh) {
  int inst;
  if (!batch_inst(B, inst)) return;
  WCC += (i64)inst * B.wclen;
  RES += (i64)inst * B.oplen;
  const SegTask t = tasks[blockIdx.x];
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= t.cnt) return;
  const int q = t.q, g = t.g;
  const int n1 = brows(B, q, g);
  int IX, JX;
  if (is_square(g)) upper_decode(e, IX, JX);
  else { IX = (int)(e % n1); JX = (int)(e / n1); }
  int i, j, k, l;
  unpack(B.pairs[B.pairoff[q * 3 + row_cls(g)] + IX], i, j);
  unpack(B.pairs[B.pairoff[q * 3 + col_cls(g)] + JX], k, l);
  ph_element(B, WCC, Lh, Rh, RES, RESh, q, g, IX, JX, i, j, k, l);
}

void launch_comm222ph_scatter(const DevBasis& B, const SegTask* t, int nt, const double* W, int Lh, int Rh,
                              double* RES, int RESh, hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_comm222ph_scatter, dim3(nt, B.ninst), dim3(kTile), 0, s, B, t, W, Lh, Rh, RES, RESh);
}

// J' ranges of the two recouplings of quadruple (i,j,k,l) (empty when
// jmin > jmax); ok = false when an isospin / parity check fails, in which
// case the element receives nothing (as in ph_element)
struct PhRanges { int jin, nin, jout, nout; bool ok; };
__device__ __forceinline__ PhRanges ph_ranges(const DevBasis& B, int i, int j, int k, int l) {
  PhRanges r{0, 0, 0, 0, false};
  const int ji = B.jj[i], jj = B.jj[j], jk = B.jj[k], jl = B.jj[l];
  const int li = B.ll[i], lj = B.ll[j], lk = B.ll[k], ll = B.ll[l];
  const int ti = B.tz[i], tj = B.tz[j], tk = B.tz[k], tl = B.tz[l];
  if (abs(tl - tj) != abs(ti - tk) || (ll + lj) % 2 != (li + lk) % 2) return r;
  if (abs(tk - tj) != abs(ti - tl) || (lk + lj) % 2 != (li + ll) % 2) return r;
  r.ok = true;
  r.jin = max(abs(jj - jl), abs(ji - jk));
  r.nin = max(0, (min(jj + jl, ji + jk) - r.jin) / 2 + 1);
  r.jout = max(abs(ji - jl), abs(jj - jk));
  r.nout = max(0, (min(ji + jl, jj + jk) - r.jout) / 2 + 1);
  return r;
}

// largest x in [0, n) with a[x] <= t (a ascending, a[0] <= t)
__device__ __forceinline__ int lds_floor(const int* a, int n, int t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// grouped form by quadruple chunks.  Phase 1: one thread per J' term of
// the chunk's quadruples loads its two W' values (index tables, then W')
// into LDS; phase 2: one thread per output (quadruple, J) sums its
// quadruple's terms from LDS with the 6j weights, in the order ph_element
// adds them (same result bits).  Each W' term is fetched once per
// quadruple instead of once per (quadruple, J).
__global__ __launch_bounds__(256) void k_comm222ph_chunk(DevBasis B, const PhChunk* chunks, const int* fq,
                                                         const int* fql, const int* fqi, const QuadItem* items,
                                                         const double* WCC, int Lh, int Rh, double* RES, int RESh,
                                                         int mirror) {
  int inst;
  if (!batch_inst(B, inst)) return;
  WCC += (i64)inst * B.wclen;
  RES += (i64)inst * B.oplen;
  __shared__ int s_ijkl[kPhChunkItems], s_loff[kPhChunkItems], s_ioff[kPhChunkItems];
  __shared__ int s_jin[kPhChunkItems], s_nin[kPhChunkItems], s_jout[kPhChunkItems];
  // per quadruple, from the orbit tables once: 2j of i, j, k, l (8 bits
  // each), and the recoupling / output data (phase, Tz and parity of the
  // inside and outside recouplings, the output channel's Tz / parity part,
  // block g and the two (a,a) factors)
  __shared__ int s_jq[kPhChunkItems], s_meta[kPhChunkItems];
  __shared__ double s_val[kPhChunkLanes];
  __shared__ i64 s_wcoff[kPhMaxCC];
  __shared__ int s_rlen[kPhMaxCC];
  const PhChunk c = chunks[xcd_contig(blockIdx.x, gridDim.x)];   // neighbouring chunks share W' lines
  const int tid = threadIdx.x;
  const double mpm = -(double)(Lh * Rh);
  for (int x = tid; x < B.ncc; x += blockDim.x) {   // cross-coupled block offsets and row counts
    s_wcoff[x] = B.wcoff[x];
    s_rlen[x] = B.rlen[x];
  }
  if (tid < c.nq) {
    const int u = fq[c.q0 + tid];
    const int i = u & 255, j = (u >> 8) & 255, k = (u >> 16) & 255, l = (u >> 24) & 255;
    const PhRanges r = ph_ranges(B, i, j, k, l);
    s_ijkl[tid] = u;
    s_loff[tid] = fql[c.q0 + tid];
    s_ioff[tid] = fqi[c.q0 + tid] - c.item0;
    s_jin[tid] = r.jin;
    s_nin[tid] = r.ok ? r.nin : 0;
    s_jout[tid] = r.ok ? r.jout : -1;   // -1: no contribution at all
    const int ji = B.jj[i], jj = B.jj[j], jk = B.jj[k], jl = B.jj[l];
    const int ti = B.tz[i], tj = B.tz[j], tk = B.tz[k], tl = B.tz[l];
    const int li = B.ll[i], lj = B.ll[j], lk = B.ll[k], ll = B.ll[l];
    const int rc = B.con[i] + B.con[j], cc = B.con[k] + B.con[l];
    const int g = (rc == 0) ? cc : (rc == 1 ? (cc == 1 ? 3 : 5) : 4);
    s_jq[tid] = ji | (jj << 8) | (jk << 16) | (jl << 24);
    // bit 0: phase (-1)^((ji+jj+jk+jl)/2) < 0; 1: Tz in; 2: parity in;
    // 3: Tz out; 4: parity out; 5-7: output block g; 8: i == j; 9: k == l;
    // 10+: the output channel's 2 (Tz + 1) + P
    s_meta[tid] = (((ji + jj + jk + jl) / 2) & 1) | ((abs(ti - tk) / 2) << 1) | (((li + lk) & 1) << 2) |
                  ((abs(ti - tl) / 2) << 3) | (((li + ll) & 1) << 4) | (g << 5) | ((i == j) << 8) |
                  ((k == l) << 9) | ((2 * ((ti + tj) / 2 + 1) + ((li + lj) & 1)) << 10);
  }
  __syncthreads();
  for (int t = tid; t < c.nlanes; t += blockDim.x) {
    const int q = lds_floor(s_loff, c.nq, t);
    const int u = s_ijkl[q], m = s_meta[q];
    const int i = u & 255, j = (u >> 8) & 255, k = (u >> 16) & 255, l = (u >> 24) & 255;
    const int r = t - s_loff[q];
    const int phase = (m & 1) ? -1 : 1;
    double v;
    if (r < s_nin[q])
      v = ph_pair<true, true>(B, s_wcoff, s_rlen, WCC, mpm, phase, i, j, k, l, s_jin[q] + 2 * r, (m >> 1) & 1,
                        (m >> 2) & 1);
    else
      v = ph_pair<false, true>(B, s_wcoff, s_rlen, WCC, mpm, phase, i, j, k, l, s_jout[q] + 2 * (r - s_nin[q]),
                         (m >> 3) & 1, (m >> 4) & 1);
    s_val[t] = v;
  }
  __syncthreads();
  if (tid >= c.nitems) return;
  const int q = lds_floor(s_ioff, c.nq, tid);
  if (s_jout[q] < 0) return;
  const QuadItem it = items[c.item0 + tid];
  const int u = it.ijkl;
  const int i = u & 255, j = (u >> 8) & 255, k = (u >> 16) & 255, l = (u >> 24) & 255;
  const int J = it.J;
  const int jq = s_jq[q], m = s_meta[q];
  const int ji = jq & 255, jj = (jq >> 8) & 255, jk = (jq >> 16) & 255, jl = (jq >> 24) & 255;
  const double* v = s_val + s_loff[q];
  const int nin = s_nin[q], jin = s_jin[q], jout = s_jout[q];
  const int nout = (q + 1 < c.nq ? s_loff[q + 1] : c.nlanes) - s_loff[q] - nin;
  double sm = 0.0;
  const double sg = -(double)m1pow((jl + jk + J) / 2);
  for (int r = 0; r < nin; ++r) sm = sm + sg * v[r] * sixj_tri(B, ji, jj, J, jl, jk, jin + 2 * r);
  for (int r = 0; r < nout; ++r) sm = sm + 1.0 * v[nin + r] * sixj_tri(B, ji, jj, J, jk, jl, jout + 2 * r);
  // output position: channel from (J, Tz, parity), block from the pair classes
  const int qc = 3 * J + (m >> 10);
  const int g = (m >> 5) & 7;
  const int IX = B.xmap[((i64)i * B.N + j) * B.nJ + (J >> 1)];
  const int JX = B.xmap[((i64)k * B.N + l) * B.nJ + (J >> 1)];
  const int n1 = brows(B, qc, g);
  const double pre = ((m >> 8) & 1) ? IMSRG_INV_SQRT2_REF : 1.0;
  const double pre2 = ((m >> 9) & 1) ? IMSRG_INV_SQRT2_REF : 1.0;
  double* X = RES + B.goff[qc * 6 + g];
  const i64 o = IX + (i64)JX * n1;
  X[o] = X[o] + sm * pre * pre2;
  if (is_square(g) && mirror) X[JX + (i64)IX * n1] = X[o] * RESh;
}

void launch_comm222ph_chunk(const DevBasis& B, const PhChunk* ch, int nch, const int* fq, const int* fql,
                            const int* fqi, const QuadItem* items, const double* W, int Lh, int Rh, double* RES,
                            int RESh, hipStream_t s, int mirror) {
  if (nch <= 0) return;
  hipLaunchKernelGGL(k_comm222ph_chunk, dim3(nch, B.ninst), dim3(256), 0, s, B, ch, fq, fql, fqi, items, W, Lh, Rh,
                     RES, RESh, mirror);
}

// lower triangle of every square block (pppp, phph, hhhh) from its upper
// triangle: X(y, x) = herm X(x, y) for x < y, by 64 x 64 tile pairs (tile
// rows ti <= tile columns tj; task.pad = block g) through LDS, coalesced
// both ways.  Lets the fused commutator's 122 and 222_ph write only the
// upper triangles.
__global__ __launch_bounds__(256) void k_sym_lower(DevBasis B, const PairTask* tasks, double* RES, int herm) {
  __shared__ double T[64][65];
  int inst;
  if (!batch_inst(B, inst)) return;
  RES += (i64)inst * B.oplen;
  const PairTask t = tasks[blockIdx.x];
  const int q = t.q, g = t.pad;
  const int n = brows(B, q, g);
  double* X = RES + B.goff[q * 6 + g];
  const int x0 = t.ti * 64, y0 = t.tj * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int c = ty; c < 64; c += 4) {   // upper tile, column y0 + c
    const int x = x0 + tx, y = y0 + c;
    T[c][tx] = (x < n && y < n && x < y) ? X[x + (i64)y * n] : 0.0;
  }
  __syncthreads();
  for (int c = ty; c < 64; c += 4) {   // lower tile, column x0 + c
    const int yy = y0 + tx, xx = x0 + c;
    if (yy < n && xx < n && xx < yy) X[yy + (i64)xx * n] = herm * T[tx][c];
  }
}

void launch_sym_lower(const DevBasis& B, const PairTask* t, int nt, double* RES, int herm, hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_sym_lower, dim3(nt, B.ninst), dim3(256), 0, s, B, t, RES, herm);
}

// =========================================================== Pandya forward
// calculate_cross_coupled, cross_coupled.f90:1190-1300.  One thread per
// CCX(q1)(r, nb) element; r = ordered pair (a,b), nb = (p,h).
__global__ __launch_bounds__(256) void k_pandya(DevBasis B, const CCTask* tasks, const double* op, int herm,
                                                double* CC) {
  int inst;
  if (!batch_inst(B, inst)) return;
  op += (i64)inst * B.oplen;
  CC += (i64)inst * B.cclen;
  const CCTask t = tasks[xcd_contig(blockIdx.x, gridDim.x)];
  const int q1 = t.q1;
  const int r = B.rlen[q1], nc = B.ncol[q1];
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= (i64)r * nc) return;
  const int R = (int)(e % r), NB = (int)(e / r);
  int a, b, p, h;
  unpack(B.ccrows[B.ccrowoff[q1] + R], a, b);
  unpack(B.cccols[B.cccoloff[q1] + NB], p, h);
  const int JC = B.ccJ[q1];
  const int ja = B.jj[a], jb = B.jj[b], jp = B.jj[p], jh = B.jj[h];
  const int la = B.ll[a], lb = B.ll[b], lp = B.ll[p], lh = B.ll[h];
  const int ta = B.tz[a], tb = B.tz[b], tp = B.tz[p], th = B.tz[h];
  double val = 0.0;
  if ((((la + lh) & 1) == ((lb + lp) & 1)) && (ta + th == tb + tp)) {
    const int Jmin = max(abs(jp - jb), abs(ja - jh));
    const int Jmax = min(jp + jb, ja + jh);
    double sm = 0.0;
    for (int JT = Jmin; JT <= Jmax; JT += 2)
      sm = sm + (double)(m1pow(JT / 2) * (JT + 1)) * sixj(B, jp, jh, JC, ja, jb, JT) *
                    v_elem(B, op, herm, h, a, p, b, JT);
    val = sm * m1pow((jh + jb + JC) / 2) * 1.0 * sqrt(JC + 1.0);
  }
  CC[B.ccoff[q1] + R + (i64)NB * r] = val;
}

void launch_pandya(const DevBasis& B, const CCTask* t, int nt, const double* op, int herm, double* CC,
                   hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_pandya, dim3(nt, B.ninst), dim3(kTile), 0, s, B, t, op, herm, CC);
}

// ============================================================ commutator_111
// commutators.f90:36-80.  One wave per one-body output (assigns); the lanes
// split the intermediate orbit k and the partial dot products are reduced
// across the wave.
__global__ __launch_bounds__(256) void k_comm111(DevBasis B, const double* L, int Lh, const double* R, int Rh,
                                                 double* RES) {
  int inst;
  if (!batch_inst(B, inst)) return;
  L += (i64)inst * B.oplen;
  R += (i64)inst * B.oplen;
  RES += (i64)inst * B.oplen;
  const int A = B.A, P = B.P;
  const i64 idx = (i64)blockIdx.x * 4 + (threadIdx.x >> 6);   // wave-uniform
  const int lane = threadIdx.x & 63;
  const double *Lhh = L + B.off_fhh, *Lph = L + B.off_fph, *Lpp = L + B.off_fpp;
  const double *Rhh = R + B.off_fhh, *Rph = R + B.off_fph, *Rpp = R + B.off_fpp;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  if (idx < (i64)A * A) {
    const int i = (int)(idx % A), j = (int)(idx / A);
    for (int k = lane; k < A; k += 64) {
      s0 += Lhh[i + k * A] * Rhh[k + j * A];
      s1 += Lhh[j + k * A] * Rhh[k + i * A];
    }
    for (int k = lane; k < P; k += 64) {
      s2 += Lph[k + (i64)i * P] * Rph[k + (i64)j * P];
      s3 += Lph[k + (i64)j * P] * Rph[k + (i64)i * P];
    }
    s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2); s3 = wave_sum(s3);
    if (lane == 0) RES[B.off_fhh + idx] = s0 + s2 * Lh - (s1 * Lh + s3) * Rh;
    return;
  }
  if (idx < (i64)A * A + (i64)P * A) {
    const i64 x = idx - (i64)A * A;
    const int a = (int)(x % P), i = (int)(x / P);
    for (int k = lane; k < A; k += 64) {
      s0 += Lph[a + (i64)k * P] * Rhh[k + i * A];
      s1 += Rph[a + (i64)k * P] * Lhh[k + i * A];
    }
    for (int k = lane; k < P; k += 64) {
      s2 += Lpp[a + (i64)k * P] * Rph[k + (i64)i * P];
      s3 += Rpp[a + (i64)k * P] * Lph[k + (i64)i * P];
    }
    s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2); s3 = wave_sum(s3);
    if (lane == 0) RES[B.off_fph + x] = (s0 + s2) - s1 - s3;
    return;
  }
  if (idx < (i64)A * A + (i64)P * A + (i64)P * P) {
    const i64 x = idx - (i64)A * A - (i64)P * A;
    const int a = (int)(x % P), b = (int)(x / P);
    for (int k = lane; k < A; k += 64) {
      s0 += Lph[a + (i64)k * P] * Rph[b + (i64)k * P];
      s1 += Lph[b + (i64)k * P] * Rph[a + (i64)k * P];
    }
    for (int k = lane; k < P; k += 64) {
      s2 += Lpp[a + (i64)k * P] * Rpp[k + (i64)b * P];
      s3 += Lpp[b + (i64)k * P] * Rpp[k + (i64)a * P];
    }
    s0 = wave_sum(s0); s1 = wave_sum(s1); s2 = wave_sum(s2); s3 = wave_sum(s3);
    if (lane == 0) RES[B.off_fpp + x] = s0 * Rh + s2 - (s1 + s3 * Rh) * Lh;
  }
}

void launch_comm111(const DevBasis& B, const double* L, int Lh, const double* R, int Rh, double* RES,
                    hipStream_t s) {
  const i64 n = (i64)B.A * B.A + (i64)B.P * B.A + (i64)B.P * B.P;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_comm111, dim3((unsigned)((n + 3) / 4), B.ninst), dim3(256), 0, s, B, L, Lh, R, Rh, RES);
}

__device__ double block_sum(double v, double* red);

// one-body result of 121 / 221: RES(p,q) += add and the herm-symmetric fill
__device__ __forceinline__ void onebody_update(const DevBasis& B, const OneTask& t, double add, double* RES,
                                               int RESh) {
  if (t.kind == 0) {
    double* F = RES + B.off_fhh;
    const i64 o = t.pi + (i64)t.qi * B.A;
    F[o] = F[o] + add;
    F[t.qi + (i64)t.pi * B.A] = F[o] * RESh;
  } else if (t.kind == 1) {
    double* F = RES + B.off_fpp;
    const i64 o = t.pi + (i64)t.qi * B.P;
    F[o] = F[o] + add;
    F[t.qi + (i64)t.pi * B.P] = F[o] * RESh;
  } else {
    double* F = RES + B.off_fph;
    const i64 o = t.pi + (i64)t.qi * B.P;
    F[o] = F[o] + add;
  }
}

// ============================================================ commutator_121
// commutators.f90:83-280 (accumulates).  One workgroup per one-body output
// (p, q); its threads take the (particle a, hole i, J) terms, with (a, i)
// from the list of particle-hole pairs inside one (l, j, tz) block (the
// only pairs the reference's loop keeps), tree-reduced in LDS.
__global__ __launch_bounds__(256) void k_comm121(DevBasis B, const OneTask* tasks, const double* L, int Lh,
                                                 const double* R, int Rh, double* RES, int RESh) {
  __shared__ double red[256];
  int inst;
  if (!batch_inst(B, inst)) return;
  L += (i64)inst * B.oplen;
  R += (i64)inst * B.oplen;
  RES += (i64)inst * B.oplen;
  const OneTask t = tasks[blockIdx.x];
  const int pk = t.pk, qk = t.qk, jq = B.jj[qk];
  const double* Lph = L + B.off_fph;
  const double* Rph = R + B.off_fph;
  const int JS = B.jsmax;
  double v = 0.0;
  for (int x = threadIdx.x; x < B.nphp * JS; x += blockDim.x) {
    const int pr = x / JS, js = x % JS;
    const int a = B.php[2 * pr], i = B.php[2 * pr + 1];
    const int ak = B.parts[a], ik = B.holes[i];
    const int ji = B.jj[ik];
    const int JT = abs(ji - jq) + 2 * js;
    if (JT > ji + jq) continue;
    const VPos p1 = vpos(B, ak, pk, ik, qk, JT);
    const VPos p2 = vpos(B, ik, pk, ak, qk, JT);
    const double la = Lph[a + (i64)i * B.P], ra = Rph[a + (i64)i * B.P];
    v = v + (JT + 1) * (la * (Lh * vat(p1, R, Rh) - vat(p2, R, Rh)) - ra * (Rh * vat(p1, L, Lh) - vat(p2, L, Lh)));
  }
  const double sm = block_sum(v, red);
  if (threadIdx.x == 0) onebody_update(B, t, sm / (B.jj[pk] + 1.0), RES, RESh);
}

void launch_comm121(const DevBasis& B, const OneTask* t, int nt, const double* L, int Lh, const double* R,
                    int Rh, double* RES, int RESh, hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_comm121, dim3(nt, B.ninst), dim3(256), 0, s, B, t, L, Lh, R, Rh, RES, RESh);
}

// ============================================================ commutator_221
// commutators.f90:435-578 (accumulates; reads w1/w2 of 222_pp_hh, herm 1).
// One workgroup per one-body output; its threads take the (intermediate
// orbit c, 2J) terms (holes c with w1, particles c with w2), tree-reduced.
// Elements of w2's pppp blocks (W2 = Rh L3 R3^T, rank nh) are formed from L
// and R here, so 222_pp_hh need not store that block.
__device__ __forceinline__ double w2_elem(const DevBasis& B, const double* w2, const double* L, const double* R,
                                          int Rh, int a, int b, int c, int d, int J) {
  const VPos p = vpos(B, a, b, c, d, J);
  if (p.off < 0) return 0.0;
  const int C1 = B.con[a] + B.con[b], C2 = B.con[c] + B.con[d];
  if (C1 != 0 || C2 != 0) return vat(p, w2, 1);
  const int q = 3 * J + 2 * ((B.tz[a] + B.tz[b]) / 2 + 1) + ((B.ll[a] + B.ll[b]) & 1);
  const int np = B.dims[q * 3], nh = B.dims[q * 3 + 2];
  const i64 o3 = B.goff[q * 6 + 2];
  const i64 rel = p.off - B.goff[q * 6 + 0];
  const int x = (int)(rel % np), y = (int)(rel / np);
  double s = 0.0;
  for (int h = 0; h < nh; ++h) s += L[o3 + x + (i64)h * np] * R[o3 + y + (i64)h * np];
  return (s * Rh) * p.pre;
}

__global__ __launch_bounds__(256) void k_comm221(DevBasis B, const OneTask* tasks, const double* L, int Lh,
                                                 const double* R, int Rh, const double* w1, const double* w2,
                                                 double* RES, int RESh) {
  __shared__ double red[256];
  int inst;
  if (!batch_inst(B, inst)) return;
  L += (i64)inst * B.oplen;
  R += (i64)inst * B.oplen;
  w1 += (i64)inst * B.oplen;
  w2 += (i64)inst * B.oplen;
  RES += (i64)inst * B.oplen;
  const OneTask t = tasks[blockIdx.x];
  const int ik = t.pk, jk = t.qk, ji = B.jj[ik];
  const int pm = Lh * Rh;
  const bool both = (t.kind != 2);
  const int JS = B.jsmax;
  double v = 0.0;
  // thread item (c, J): intermediate orbit c (holes first, with w1; then
  // particles, with w2) and one 2J of its coupling to i
  for (int y = threadIdx.x; y < B.N * JS; y += blockDim.x) {
    const int x = y / JS;
    const bool hole = x < B.A;
    const int ck = hole ? B.holes[x] : B.parts[x - B.A];
    const int jc = B.jj[ck];
    const int JT = abs(jc - ji) + 2 * (y % JS);
    if (JT > jc + ji) continue;
    if (hole) {
      if (both)
        v = v + (v_elem(B, w1, 1, ck, ik, ck, jk, JT) - pm * v_elem(B, w1, 1, ck, jk, ck, ik, JT)) * (JT + 1);
      else
        v = v + v_elem(B, w1, 1, ck, ik, ck, jk, JT) * (JT + 1);
    } else {
      if (both)
        v = v + (w2_elem(B, w2, L, R, Rh, ck, ik, ck, jk, JT) - pm * w2_elem(B, w2, L, R, Rh, ck, jk, ck, ik, JT)) *
                    (JT + 1);
      else
        v = v + w2_elem(B, w2, L, R, Rh, ck, ik, ck, jk, JT) * (JT + 1);
    }
  }
  const double sm = block_sum(v, red);
  if (threadIdx.x == 0) onebody_update(B, t, sm / (ji + 1.0), RES, RESh);
}

void launch_comm221(const DevBasis& B, const OneTask* t, int nt, const double* L, int Lh, const double* R, int Rh,
                    const double* w1, const double* w2, double* RES, int RESh, hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_comm221, dim3(nt, B.ninst), dim3(256), 0, s, B, t, L, Lh, R, Rh, w1, w2, RES, RESh);
}

// ================================================= commutator_110 + 220 (E0)
// commutators.f90:10-33 and 399-432.  Workgroup q < nblocks: channel q's
// 220 sum times (2J+1); workgroup nblocks: the 110 sum.
__device__ double block_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) red[tid] += red[tid + s];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void k_e0_partials(DevBasis B, const double* L, int Lh, const double* R, int Rh,
                                                     double* part) {
  int inst;
  if (!batch_inst(B, inst)) return;
  L += (i64)inst * B.oplen;
  R += (i64)inst * B.oplen;
  part += (i64)inst * B.spart;
  __shared__ double red[256];
  const int q = blockIdx.x;
  double v = 0.0;
  if (q < B.nblocks) {
    const int np = B.dims[q * 3], nh = B.dims[q * 3 + 2];
    const double* L3 = L + B.goff[q * 6 + 2];
    const double* R3 = R + B.goff[q * 6 + 2];
    for (i64 x = threadIdx.x; x < (i64)np * nh; x += blockDim.x) v = v - L3[x] * R3[x] * Rh;
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) part[q] = s * (B.chJ[q] + 1);
  } else {
    const double* Lph = L + B.off_fph;
    const double* Rph = R + B.off_fph;
    for (i64 x = threadIdx.x; x < (i64)B.P * B.A; x += blockDim.x) {
      const int hk = B.holes[(int)(x / B.P)];
      v = v + (Lph[x] * Rph[x] * Lh - Rph[x] * Lph[x] * Rh) * (B.jj[hk] + 1);
    }
    const double s = block_sum(v, red);
    if (threadIdx.x == 0) part[B.nblocks] = s;
  }
}

// dst = [use110] part[n] + [use220] 2 * sum(part[0..n-1])   (one workgroup)
__global__ __launch_bounds__(256) void k_e0_final(const double* part, int n, int use220, int use110, double* dst,
                                                  int accumulate, i64 spart, i64 sdst, const int* act) {
  __shared__ double red[256];
  int inst;
  if (!batch_on(act, inst)) return;
  part += (i64)inst * spart;
  dst += (i64)inst * sdst;
  double v = 0.0;
  for (int x = threadIdx.x; x < n; x += blockDim.x) v += part[x];
  const double s = block_sum(v, red);
  if (threadIdx.x == 0) {
    const double e110 = use110 ? part[n] : 0.0;
    const double e220 = use220 ? s * 2.0 : 0.0;
    const double e = e110 + e220;
    dst[0] = accumulate ? dst[0] + e : e;
  }
}

void launch_e0_partials(const DevBasis& B, const double* L, int Lh, const double* R, int Rh, double* part,
                        hipStream_t s) {
  hipLaunchKernelGGL(k_e0_partials, dim3(B.nblocks + 1, B.ninst), dim3(256), 0, s, B, L, Lh, R, Rh, part);
}

void launch_e0_final(const double* part, int nblocks, int use220, int use110, double* dst, int accumulate,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_e0_final, dim3(1), dim3(256), 0, s, part, nblocks, use220, use110, dst, accumulate, (i64)0,
                     (i64)0, (const int*)nullptr);
}
void launch_e0_final_b(const DevBasis& B, const double* part, int use220, int use110, double* dst, int accumulate,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_e0_final, dim3(1, B.ninst), dim3(256), 0, s, part, B.nblocks, use220, use110, dst, accumulate,
                     B.spart, B.oplen, B.act);
}

// ======================================== 222_pp_hh pppp block epilogue
// commutators.f90:611-668.  The grouped GEMM leaves C = W1 - W2 =
// L1 R1 - Rh L3 R3^T in w1's pppp block; RES1 += C - pm C^T.  Tile pair
// (I,J), I <= J, of 64 x 64: both C tiles are staged through LDS so every
// global access is a contiguous column run.  With ref_outputs the
// reference's W1 and W2 are left in w1 / w2 (W2 formed here, rank nh).
__device__ __forceinline__ double w2pp(const double* L3, const double* R3, int np, int nh, int Rh, int x, int y) {
  double s = 0.0;
  for (int h = 0; h < nh; ++h) s += L3[x + (i64)h * np] * R3[y + (i64)h * np];
  return s * Rh;
}

__global__ __launch_bounds__(256) void k_pppp_sym(DevBasis B, const PairTask* tasks, const double* L, int Lh,
                                                  const double* R, int Rh, double* RES, double* w1, double* w2,
                                                  int ref_outputs) {
  int inst;
  if (!batch_inst(B, inst)) return;
  L += (i64)inst * B.oplen;
  R += (i64)inst * B.oplen;
  RES += (i64)inst * B.oplen;
  w1 += (i64)inst * B.oplen;
  w2 += (i64)inst * B.oplen;
  __shared__ double T1[64][65];   // C(I0 + x, J0 + y)
  __shared__ double T2[64][65];   // C(J0 + x, I0 + y)
  const PairTask t = tasks[blockIdx.x];
  const int q = t.q;
  const int np = B.dims[q * 3], nh = B.dims[q * 3 + 2];
  const i64 o1 = B.goff[q * 6 + 0], o3 = B.goff[q * 6 + 2];
  const double *L3 = L + o3, *R3 = R + o3;
  double* Cm = w1 + o1;
  double* X1 = RES + o1;
  double* W21 = w2 + o1;
  const int pm = Lh * Rh;
  const int I0 = t.ti * 64, J0 = t.tj * 64;
  const bool diag = (t.ti == t.tj);
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int x = e & 63, y = e >> 6;
    T1[x][y] = (I0 + x < np && J0 + y < np) ? Cm[(I0 + x) + (i64)(J0 + y) * np] : 0.0;
    if (!diag) T2[x][y] = (J0 + x < np && I0 + y < np) ? Cm[(J0 + x) + (i64)(I0 + y) * np] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int x = e & 63, y = e >> 6;
    const int gx = I0 + x, gy = J0 + y;
    if (gx < np && gy < np) {
      const double cxy = T1[x][y];
      const double cyx = diag ? T1[y][x] : T2[y][x];
      const i64 o = gx + (i64)gy * np;
      X1[o] = X1[o] + cxy - cyx * pm;
      if (ref_outputs) {
        const double w2v = w2pp(L3, R3, np, nh, Rh, gx, gy);
        Cm[o] = cxy + w2v;
        W21[o] = w2v;
      }
    }
    if (!diag) {
      const int hx = J0 + x, hy = I0 + y;   // element (J0 + x, I0 + y) of tile (J, I)
      if (hx < np && hy < np) {
        const double cxy = T2[x][y];
        const double cyx = T1[y][x];
        const i64 o = hx + (i64)hy * np;
        X1[o] = X1[o] + cxy - cyx * pm;
        if (ref_outputs) {
          const double w2v = w2pp(L3, R3, np, nh, Rh, hx, hy);
          Cm[o] = cxy + w2v;
          W21[o] = w2v;
        }
      }
    }
  }
}

// The same epilogue when w1 is not rewritten (the flow's commutator): one
// workgroup per output tile (two per off-diagonal pair, blockIdx.x = 2 task
// + side), its own C tile read straight from memory and only the tile whose
// transpose it needs staged in LDS (33 KB instead of 66 KB: twice the
// workgroups per CU).  Same arithmetic as k_pppp_sym.
__global__ __launch_bounds__(256) void k_pppp_sym1(DevBasis B, const PairTask* tasks, double* RES, const double* w1,
                                                   int pm, int upper_only) {
  int inst;
  if (!batch_inst(B, inst)) return;
  RES += (i64)inst * B.oplen;
  w1 += (i64)inst * B.oplen;
  __shared__ double S[64][65];   // S[x][y] = C(A0 + x, B0 + y), the transposed source tile
  // upper_only: one workgroup per tile pair for its upper tile (the caller
  // fills the lower triangles afterwards), else two, the second for the
  // lower tile
  const PairTask t = tasks[upper_only ? blockIdx.x : blockIdx.x >> 1];
  const int side = upper_only ? 0 : (blockIdx.x & 1);
  if (side == 1 && t.ti == t.tj) return;
  const int q = t.q;
  const int np = B.dims[q * 3];
  const i64 o1 = B.goff[q * 6 + 0];
  const double* Cm = w1 + o1;
  double* X1 = RES + o1;
  // output tile rows R0.., cols C0..; source tile rows C0.., cols R0..
  const int R0 = (side == 0 ? t.ti : t.tj) * 64, C0 = (side == 0 ? t.tj : t.ti) * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int x = e & 63, y = e >> 6;
    S[x][y] = (C0 + x < np && R0 + y < np) ? Cm[(C0 + x) + (i64)(R0 + y) * np] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int x = e & 63, y = e >> 6;
    const int gx = R0 + x, gy = C0 + y;
    if (gx < np && gy < np) {
      const i64 o = gx + (i64)gy * np;
      const double cxy = Cm[o];
      const double cyx = S[y][x];
      X1[o] = X1[o] + cxy - cyx * pm;
    }
  }
}

void launch_pppp_sym(const DevBasis& B, const PairTask* t, int nt, const double* L, int Lh, const double* R, int Rh,
                     double* RES, double* w1, double* w2, int ref_outputs, hipStream_t s) {
  if (nt <= 0) return;
  if (ref_outputs != 1) {
    const int upper_only = ref_outputs == 2 ? 1 : 0;
    hipLaunchKernelGGL(k_pppp_sym1, dim3(upper_only ? nt : 2 * nt, B.ninst), dim3(256), 0, s, B, t, RES,
                       (const double*)w1, Lh * Rh, upper_only);
    return;
  }
  hipLaunchKernelGGL(k_pppp_sym, dim3(nt, B.ninst), dim3(256), 0, s, B, t, L, Lh, R, Rh, RES, w1, w2, ref_outputs);
}

// ================================ 222_pp_hh epilogue of the small blocks
// commutators.f90:636-764 after the grouped GEMMs:
//   g == 1: region gam2|gam3 (np x (nb+nh)); w1 already holds
//           L1 [R2|R3] - R1 [L2|L3]; adds the rank-nh parts W2 and RES.
//   g == 3 / 4 / 5: gam4 / gam5 / gam6 from P = [L2|L3]^T [R2|R3].
// P = [L2|L3]^T [R2|R3] (nbh x nbh) element (x, y): sum of the split-K slabs
__device__ __forceinline__ double pval(const double* Pq, int nbh, int ks, int x, int y) {
  double v = 0.0;
  for (int s = 0; s < ks; ++s) v += Pq[(i64)s * nbh * nbh + x + (i64)y * nbh];
  return v;
}

__global__ __launch_bounds__(256) void k_pphh_epilogue(DevBasis B, const SegTask* tasks, const double* L, int Lh,
                                                       const double* R, int Rh, double* RES, double* w1,
                                                       double* w2, const double* Pscr, const i64* poff,
                                                       const int* k3ks, const double* K2scr, const i64* k2off,
                                                       const int* k2ks) {
  int inst;
  if (!batch_inst(B, inst)) return;
  L += (i64)inst * B.oplen;
  R += (i64)inst * B.oplen;
  RES += (i64)inst * B.oplen;
  w1 += (i64)inst * B.oplen;
  w2 += (i64)inst * B.oplen;
  Pscr += (i64)inst * B.spscr;
  K2scr += (i64)inst * B.sk2scr;
  const SegTask t = tasks[blockIdx.x];
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= t.cnt) return;
  const int q = t.q;
  const int np = B.dims[q * 3], nb = B.dims[q * 3 + 1], nh = B.dims[q * 3 + 2];
  const int pm = Lh * Rh;
  const i64* go = B.goff + q * 6;
  const double *L3 = L + go[2], *R3 = R + go[2], *L5 = L + go[4], *R5 = R + go[4], *L6 = L + go[5],
               *R6 = R + go[5];
  if (t.g == 1) {
    const int x = (int)(e % np), y = (int)(e / np);
    double w2v;
    if (y < nb) {
      double s1 = 0.0, s2 = 0.0;
      for (int h = 0; h < nh; ++h) {
        s1 += L3[x + (i64)h * np] * R6[y + (i64)h * nb];
        s2 += R3[x + (i64)h * np] * L6[y + (i64)h * nb];
      }
      w2v = Rh * s1 + (-1.0 * Lh) * s2;
    } else {
      const int hp = y - nb;
      double s1 = 0.0, s2 = 0.0;
      for (int h = 0; h < nh; ++h) {
        s1 += L3[x + (i64)h * np] * R5[h + (i64)hp * nh];
        s2 += R3[x + (i64)h * np] * L5[h + (i64)hp * nh];
      }
      w2v = s1 + (-1.0) * s2;
    }
    // W1 = L1 [R2|R3] - R1 [L2|L3]: sum of the split-K slabs
    double w1v = 0.0;
    const double* K2q = K2scr + k2off[q];
    const i64 slab = (i64)np * (nb + nh);
    for (int s = 0; s < k2ks[q]; ++s) w1v += K2q[s * slab + e];
    const i64 o = go[1] + e;
    RES[o] = RES[o] + w1v - w2v;
    w1[o] = w1v;
    w2[o] = w2v;
    return;
  }
  const int nbh = nb + nh;
  const double* Pq = Pscr + poff[q];
  const int ks = k3ks[q];
  if (t.g == 3) {   // gam4, nb x nb
    const int x = (int)(e % nb), y = (int)(e / nb);
    double sxy = 0.0, syx = 0.0;
    for (int h = 0; h < nh; ++h) {
      sxy += L6[x + (i64)h * nb] * R6[y + (i64)h * nb];
      syx += L6[y + (i64)h * nb] * R6[x + (i64)h * nb];
    }
    const double w1xy = pval(Pq, nbh, ks, x, y) * Lh, w1yx = pval(Pq, nbh, ks, y, x) * Lh;
    const double w2xy = sxy * Rh, w2yx = syx * Rh;
    const i64 o = go[3] + e;
    RES[o] = RES[o] + w1xy - w2xy - (w1yx - w2yx) * pm;
    w1[o] = w1xy;
    w2[o] = w2xy;
    return;
  }
  if (t.g == 4) {   // gam5, nh x nh
    const int a = (int)(e % nh), b = (int)(e / nh);
    double sab = 0.0, sba = 0.0;
    for (int h = 0; h < nh; ++h) {
      sab += L5[a + (i64)h * nh] * R5[h + (i64)b * nh];
      sba += L5[b + (i64)h * nh] * R5[h + (i64)a * nh];
    }
    const double w1ab = pval(Pq, nbh, ks, nb + a, nb + b) * Lh, w1ba = pval(Pq, nbh, ks, nb + b, nb + a) * Lh;
    const i64 o = go[4] + e;
    RES[o] = RES[o] + w1ab - sab - (w1ba - sba) * pm;
    w1[o] = w1ab;
    w2[o] = sab;
    return;
  }
  {   // gam6, nb x nh
    const int x = (int)(e % nb), h = (int)(e / nb);
    double s1 = 0.0, s2 = 0.0;
    for (int hp = 0; hp < nh; ++hp) {
      s1 += L6[x + (i64)hp * nb] * R5[hp + (i64)h * nh];
      s2 += R6[x + (i64)hp * nb] * L5[hp + (i64)h * nh];
    }
    const double w1v = Lh * pval(Pq, nbh, ks, x, nb + h) + (-1.0 * Rh) * pval(Pq, nbh, ks, nb + h, x);
    const double w2v = s1 + (-1.0) * s2;
    const i64 o = go[5] + e;
    RES[o] = RES[o] + w1v - w2v;
    w1[o] = w1v;
    w2[o] = w2v;
  }
}

void launch_pphh_epilogue(const DevBasis& B, const SegTask* t, int nt, const double* L, int Lh, const double* R,
                          int Rh, double* RES, double* w1, double* w2, const double* Pscr, const i64* poff,
                          const int* k3ks, const double* K2scr, const i64* k2off, const int* k2ks, hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_pphh_epilogue, dim3(nt, B.ninst), dim3(256), 0, s, B, t, L, Lh, R, Rh, RES, w1, w2, Pscr, poff, k3ks,
                     K2scr, k2off, k2ks);
}

// ================================================ streaming algebra / norm
// add_sq_op, basic_IMSRG.f90:3551-3583:  C = A*ax + B*bx
__global__ __launch_bounds__(256) void k_axpy(i64 n, const double* A, double ax, const double* Bv, double bx,
                                              double* C, const int* act) {
  int inst;
  if (!batch_on(act, inst)) return;
  A += (i64)inst * n; Bv += (i64)inst * n; C += (i64)inst * n;
  for (i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (i64)gridDim.x * blockDim.x)
    C[x] = A[x] * ax + Bv[x] * bx;
}
static unsigned stream_grid(i64 n) {
  i64 g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  if (g < 1) g = 1;
  return (unsigned)g;
}

void launch_axpy(i64 n, const double* A, double ax, const double* Bv, double bx, double* C, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy, dim3(stream_grid(n)), dim3(256), 0, s, n, A, ax, Bv, bx, C, (const int*)nullptr);
}
// the same on ninst instances back to back (stride n), act-masked
void launch_axpy_b(i64 n, const double* A, double ax, const double* Bv, double bx, double* C, int ninst,
                   const int* act, hipStream_t s) {
  unsigned g = stream_grid(n);
  if (ninst > 1 && g > 1024) g = 1024;
  hipLaunchKernelGGL(k_axpy, dim3(g, ninst), dim3(256), 0, s, n, A, ax, Bv, bx, C, act);
}

// mat_frob_norm, basic_IMSRG.f90:4056-4087: weighted sum of squares per chunk
__global__ __launch_bounds__(256) void k_norm_partials(const double* op, const i64* coff, const int* clen,
                                                       const double* cw, double* part, i64 sop, i64 spart,
                                                       const int* act) {
  __shared__ double red[256];
  int inst;
  if (!batch_on(act, inst)) return;
  op += (i64)inst * sop;
  part += (i64)inst * spart;
  const int c = blockIdx.x;
  const double* x = op + coff[c];
  const int n = clen[c];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += x[i] * x[i];
  const double s = block_sum(v, red);
  if (threadIdx.x == 0) part[c] = s * cw[c];
}
// X += c * T (every element, E0 by chunk 0's first thread) and the norm
// partials of T in the same pass: the axpy and norm of one nested
// commutator (append_operator + mat_frob_norm, IMSRG_MAGNUS.f90:227-229),
// each element's update and each chunk's partial sum exactly as k_axpy and
// k_norm_partials form them
__global__ __launch_bounds__(256) void k_axpy_norm(double* X, const double* T, double c, const i64* coff,
                                                   const int* clen, const double* cw, double* part, i64 sop, i64 spart,
                                                   const int* act) {
  __shared__ double red[256];
  int inst;
  if (!batch_on(act, inst)) return;
  X += (i64)inst * sop;
  T += (i64)inst * sop;
  part += (i64)inst * spart;
  const int ch = blockIdx.x;
  if (ch == 0 && threadIdx.x == 0) X[0] = X[0] * 1.0 + T[0] * c;
  const i64 o = coff[ch];
  const int n = clen[ch];
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double y = T[o + i];
    v += y * y;
    X[o + i] = X[o + i] * 1.0 + y * c;
  }
  const double s = block_sum(v, red);
  if (threadIdx.x == 0) part[ch] = s * cw[ch];
}
void launch_axpy_norm_b(const DevBasis& B, double* X, const double* T, double c, const i64* chunk_off,
                        const int* chunk_len, const double* chunk_w, int nchunk, double* part, hipStream_t s) {
  if (nchunk <= 0) return;
  hipLaunchKernelGGL(k_axpy_norm, dim3(nchunk, B.ninst), dim3(256), 0, s, X, T, c, chunk_off, chunk_len, chunk_w, part,
                     B.oplen, B.spart, B.act);
}

__global__ __launch_bounds__(1024) void k_sum_sqrt(const double* part, int n, double* dst, int do_sqrt, i64 spart,
                                                   i64 sdst, const int* act) {
  __shared__ double red[1024];
  int inst;
  if (!batch_on(act, inst)) return;
  part += (i64)inst * spart;
  dst += (i64)inst * sdst;
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
  const double s = block_sum(v, red);
  if (threadIdx.x == 0) dst[0] = do_sqrt ? sqrt(s) : s;
}

void launch_norm_partials_b(const DevBasis& B, const double* op, const i64* chunk_off, const int* chunk_len,
                            const double* chunk_w, int nchunk, double* part, hipStream_t s) {
  if (nchunk <= 0) return;
  hipLaunchKernelGGL(k_norm_partials, dim3(nchunk, B.ninst), dim3(256), 0, s, op, chunk_off, chunk_len, chunk_w, part,
                     B.oplen, B.spart, B.act);
}
void launch_sum_sqrt_b(const DevBasis& B, const double* part, int n, double* dst, i64 sdst, int do_sqrt,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_sum_sqrt, dim3(1, B.ninst), dim3(1024), 0, s, part, n, dst, do_sqrt, B.spart, sdst, B.act);
}

// ======================================================== white generator
// twobody_monopole, basic_IMSRG.f90:2075-2090, for every orbit pair
__global__ __launch_bounds__(256) void k_monopole(DevBasis B, const double* H, double* M) {
  int inst;
  if (!batch_inst(B, inst)) return;
  H += (i64)inst * B.oplen;
  M += (i64)inst * B.N * B.N;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B.N * B.N) return;
  const int a = idx % B.N, b = idx / B.N;
  const int ja = B.jj[a], jb = B.jj[b];
  double sm = 0.0;
  for (int JT = abs(ja - jb); JT <= ja + jb; JT += 2) sm = sm + (JT + 1) * v_elem(B, H, 1, a, b, a, b, JT);
  M[idx] = sm / (ja + 1.0) / (jb + 1.0);
}
void launch_monopole(const DevBasis& B, const double* H, double* M, hipStream_t s) {
  hipLaunchKernelGGL(k_monopole, dim3((B.N * B.N + 255) / 256, B.ninst), dim3(256), 0, s, B, H, M);
}

// Ground-state generators: the energy denominators of build_gs_white
// (generators.f90:4-106) from Hd, the numerators from Hn:
//   mode 0 (white, generators.f90:54,100):  eta = Hn / Eden
//   mode 1 (atan / atanxx, generators.f90:213,263 and 318,366):
//                                            eta = 0.5 atan(2 Hn / Eden)
// (build_gs_atan is mode 1 with Hd = Hn; build_gs_atanxx passes Hold as Hd)
__device__ __forceinline__ double gen_value(double x, double Eden, int mode) {
  return mode == 0 ? x / Eden : atan(2 * x / Eden) * 0.5;
}

// one-body part (also zeroes fhh, fpp)
__global__ __launch_bounds__(256) void k_gen_1b(DevBasis B, const double* Hd, const double* Hn, double* ETA,
                                                int mode) {
  int inst;
  if (!batch_inst(B, inst)) return;
  Hd += (i64)inst * B.oplen;
  Hn += (i64)inst * B.oplen;
  ETA += (i64)inst * B.oplen;
  const i64 idx = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  const i64 nA = (i64)B.A * B.A, nPA = (i64)B.P * B.A, nP = (i64)B.P * B.P;
  if (idx < nA) { ETA[B.off_fhh + idx] = 0.0; return; }
  if (idx < nA + nP) { ETA[B.off_fpp + (idx - nA)] = 0.0; return; }
  if (idx >= nA + nP + nPA) return;
  const i64 x = idx - nA - nP;
  const int a = (int)(x % B.P), i = (int)(x / B.P);
  const int ak = B.parts[a], ik = B.holes[i];
  if (B.jj[ik] != B.jj[ak] || B.ll[ik] != B.ll[ak] || B.tz[ik] != B.tz[ak]) return;
  const int ji = B.jj[ik];
  double Eden = 0.0;
  for (int JT = 0; JT <= 2 * ji; JT += 2) Eden = Eden - (JT + 1) * v_elem(B, Hd, 1, ak, ik, ak, ik, JT);
  Eden = Eden / (ji + 1.0) / (ji + 1.0);
  Eden = Eden + Hd[B.off_fpp + a + (i64)a * B.P] - Hd[B.off_fhh + i + (i64)i * B.A];
  ETA[B.off_fph + x] = gen_value(Hn[B.off_fph + x], Eden, mode);
}
void launch_gen_1b(const DevBasis& B, const double* Hd, const double* Hn, double* ETA, int mode, hipStream_t s) {
  const i64 n = (i64)B.A * B.A + (i64)B.P * B.P + (i64)B.P * B.A;
  hipLaunchKernelGGL(k_gen_1b, dim3((unsigned)((n + 255) / 256), B.ninst), dim3(256), 0, s, B, Hd, Hn, ETA, mode);
}

// two-body part, full-segment tasks (gam3 from the formula, the rest zero);
// M = the monopoles of Hd
__global__ __launch_bounds__(256) void k_gen_2b(DevBasis B, const SegTask* tasks, const double* Hd, const double* M,
                                                const double* Hn, double* ETA, int mode) {
  int inst;
  if (!batch_inst(B, inst)) return;
  Hd += (i64)inst * B.oplen;
  Hn += (i64)inst * B.oplen;
  ETA += (i64)inst * B.oplen;
  M += (i64)inst * B.N * B.N;
  const SegTask t = tasks[blockIdx.x];
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= t.cnt) return;
  const int q = t.q, g = t.g;
  const i64 o = B.goff[q * 6 + g] + e;
  if (g != 2) { ETA[o] = 0.0; return; }
  const int np = B.dims[q * 3];
  const int IX = (int)(e % np), JX = (int)(e / np);
  int a, b, i, j;
  unpack(B.pairs[B.pairoff[q * 3 + 0] + IX], a, b);
  unpack(B.pairs[B.pairoff[q * 3 + 2] + JX], i, j);
  const int N = B.N;
  double Eden = 0.0;
  Eden = Eden + M[a + b * N];
  Eden = Eden + M[i + j * N];
  Eden = Eden - M[a + i * N];
  Eden = Eden - M[a + j * N];
  Eden = Eden - M[i + b * N];
  Eden = Eden - M[j + b * N];
  Eden = Eden + f_elem(B, Hd, 1, a, a) + f_elem(B, Hd, 1, b, b) - f_elem(B, Hd, 1, i, i) - f_elem(B, Hd, 1, j, j);
  ETA[o] = gen_value(Hn[o], Eden, mode);
}
void launch_gen_2b(const DevBasis& B, const SegTask* t, int nt, const double* Hd, const double* M, const double* Hn,
                   double* ETA, int mode, hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_gen_2b, dim3(nt, B.ninst), dim3(256), 0, s, B, t, Hd, M, Hn, ETA, mode);
}

// mbpt2, basic_IMSRG.f90:4165-4199: one partial per gam3 tile
__global__ __launch_bounds__(256) void k_mbpt2_partials(DevBasis B, const SegTask* tasks, const double* H,
                                                        double* part) {
  int inst;
  if (!batch_inst(B, inst)) return;
  H += (i64)inst * B.oplen;
  part += (i64)inst * B.spart;
  __shared__ double red[256];
  const SegTask t = tasks[blockIdx.x];
  const i64 e = (i64)t.start + threadIdx.x;
  double v = 0.0;
  const int q = t.q;
  if (e < t.cnt) {
    const int np = B.dims[q * 3];
    const int II = (int)(e % np), JJ = (int)(e / np);
    int i, j, k, l;
    unpack(B.pairs[B.pairoff[q * 3 + 0] + II], i, j);
    unpack(B.pairs[B.pairoff[q * 3 + 2] + JJ], k, l);
    const double fi = f_elem(B, H, 1, i, i), fj = f_elem(B, H, 1, j, j);
    const double fk = f_elem(B, H, 1, k, k), fl = f_elem(B, H, 1, l, l);
    const double eden = fk + fl - fi - fj;
    const double x = H[B.goff[q * 6 + 2] + e];
    v = x * x / eden;
  }
  const double s = block_sum(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s * (B.chJ[q] + 1.0);
}
void launch_mbpt2_partials(const DevBasis& B, const SegTask* t, int nt, const double* H, double* part,
                           hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_mbpt2_partials, dim3(nt, B.ninst), dim3(256), 0, s, B, t, H, part);
}

// ============================================================ normal_order
// basic_IMSRG.f90:2419-2550.  scratch[0] receives E0; then the one-body
// corrections are applied in place (every thread reads only the element it
// writes, so the in-place update equals the reference's TEMP copy).
__global__ __launch_bounds__(256) void k_no_e0(DevBasis B, const double* H, double* E0) {
  __shared__ double red[256];
  double v = 0.0;
  // 2 * sum_h f_hh (2j+1)
  for (int i = threadIdx.x; i < B.N; i += blockDim.x)
    if (B.con[i] == 1) v += f_elem(B, H, 1, i, i) * (B.jj[i] + 1);
  double s1 = block_sum(v, red) * 2.0;
  v = 0.0;
  for (int x = threadIdx.x; x < B.N * B.N; x += blockDim.x) {
    const int i = x / B.N, j = x % B.N;
    if (B.con[i] * B.con[j] == 0) continue;
    for (int JT = abs(B.jj[i] - B.jj[j]); JT <= B.jj[i] + B.jj[j]; JT += 2)
      v += v_elem(B, H, 1, i, j, i, j, JT) * (JT + 1);
  }
  const double s2 = block_sum(v, red);
  if (threadIdx.x == 0) E0[0] = (s1 + s2) * 0.5;
}

__global__ __launch_bounds__(256) void k_no_1b(DevBasis B, double* H) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B.N * B.N) return;
  const int ax = idx % B.N, bx = idx / B.N;
  if (B.jj[ax] != B.jj[bx] || B.ll[ax] != B.ll[bx] || B.tz[ax] != B.tz[bx]) return;
  const int ca = B.con[ax], cb = B.con[bx];
  // the reference visits hh (a<=b), ph (a particle, b hole), pp (a<=b)
  if (ca == 1 && cb == 1 && B.cidx[ax] > B.cidx[bx]) return;
  if (ca == 0 && cb == 0 && B.cidx[ax] > B.cidx[bx]) return;
  if (ca == 1 && cb == 0) return;
  double sm = 0.0;
  for (int ix = 0; ix < B.N; ++ix) {
    if (B.con[ix] != 1) continue;
    for (int JT = abs(B.jj[ix] - B.jj[ax]); JT <= B.jj[ix] + B.jj[ax]; JT += 2)
      sm = sm + (JT + 1.0) / (B.jj[ax] + 1) * v_elem(B, H, 1, ax, ix, bx, ix, JT);
  }
  const double v = f_elem(B, H, 1, ax, bx) + sm;
  const int a = B.cidx[ax], b = B.cidx[bx];
  if (ca == 1) {
    H[B.off_fhh + a + (i64)b * B.A] = v;
    H[B.off_fhh + b + (i64)a * B.A] = v;
  } else if (cb == 1) {
    H[B.off_fph + a + (i64)b * B.P] = v;
  } else {
    H[B.off_fpp + a + (i64)b * B.P] = v;
    H[B.off_fpp + b + (i64)a * B.P] = v;
  }
}

__global__ void k_set_e0(double* H, const double* E0) { H[0] = E0[0]; }

// vectorize / repackage, basic_IMSRG.f90:3929-4053: one workgroup per run.
// In repackage the lower element X(j,i) is stored before its partner
// X(i,j) = herm * v by the same thread, so a diagonal element ends as
// herm * v, as in the reference's loop.
__global__ __launch_bounds__(256) void k_vectorize(const VecTask* tasks, double* op, double* vec, int dir,
                                                   double herm) {
  const VecTask t = tasks[blockIdx.x];
  for (int x = threadIdx.x; x < t.len; x += blockDim.x) {
    if (dir == 0) {
      vec[t.vec_off + x] = op[t.op_off + x];
    } else {
      const double v = vec[t.vec_off + x];
      op[t.op_off + x] = v;
      if (t.mirror) op[t.op_off + (i64)x * t.mirror] = v * herm;
    }
  }
}
void launch_vectorize(const VecTask* t, int nt, double* op, double* vec, int dir, int herm, hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_vectorize, dim3(nt), dim3(256), 0, s, t, op, vec, dir, (double)herm);
}

// E0 = v for every instance of an operator array
__global__ void k_fill_e0(DevBasis B, double* op, double v) {
  int inst;
  if (!batch_inst(B, inst)) return;
  if (threadIdx.x == 0) op[(i64)inst * B.oplen] = v;
}
void launch_fill_e0(const DevBasis& B, double* op, double v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_e0, dim3(1, B.ninst), dim3(64), 0, s, B, op, v);
}

void launch_normal_order(const DevBasis& B, double* H, double* scratch, hipStream_t s) {
  hipLaunchKernelGGL(k_no_e0, dim3(1), dim3(256), 0, s, B, H, scratch);
  hipLaunchKernelGGL(k_no_1b, dim3((B.N * B.N + 255) / 256), dim3(256), 0, s, B, H);
  hipLaunchKernelGGL(k_set_e0, dim3(1), dim3(1), 0, s, H, scratch);
}

// ================================================================ helpers
__global__ void k_fill(double* p, i64 n, double v, const int* act) {
  int inst;
  if (!batch_on(act, inst)) return;
  p += (i64)inst * n;
  for (i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (i64)gridDim.x * blockDim.x) p[x] = v;
}
__global__ void k_copy(double* d, const double* s, i64 n, const int* act) {
  int inst;
  if (!batch_on(act, inst)) return;
  d += (i64)inst * n;
  s += (i64)inst * n;
  for (i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (i64)gridDim.x * blockDim.x) d[x] = s[x];
}
void launch_fill(double* p, i64 n, double v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill, dim3(stream_grid(n)), dim3(256), 0, s, p, n, v, (const int*)nullptr);
}
void launch_copy(double* dst, const double* src, i64 n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy, dim3(stream_grid(n)), dim3(256), 0, s, dst, src, n, (const int*)nullptr);
}
void launch_fill_b(double* p, i64 n, double v, int ninst, const int* act, hipStream_t s) {
  if (n <= 0) return;
  unsigned g = stream_grid(n);
  if (ninst > 1 && g > 1024) g = 1024;
  hipLaunchKernelGGL(k_fill, dim3(g, ninst), dim3(256), 0, s, p, n, v, act);
}
void launch_copy_b(double* dst, const double* src, i64 n, int ninst, const int* act, hipStream_t s) {
  if (n <= 0) return;
  unsigned g = stream_grid(n);
  if (ninst > 1 && g > 1024) g = 1024;
  hipLaunchKernelGGL(k_copy, dim3(g, ninst), dim3(256), 0, s, dst, src, n, act);
}

// ====================================================== transform_2b_to_HF
// HF_mod.f90:567-692: each channel's two-body blocks as one square matrix
// over its pairs in the order (hh, ph, pp), nt x nt column-major at
// V + voff[q] (square blocks as stored, rectangular blocks in both
// orientations); unpack writes the six blocks back from it.  One thread per
// stored element (the full-block segment tasks).
__device__ __forceinline__ void hf_pos(const DevBasis& B, int q, int g, i64 e, int& n1, i64& row, i64& col, i64& nt) {
  const int np = B.dims[q * 3], nb = B.dims[q * 3 + 1], nh = B.dims[q * 3 + 2];
  const int r0[3] = {nh + nb, nh, 0};
  n1 = brows(B, q, g);
  row = r0[row_cls(g)] + e % n1;
  col = r0[col_cls(g)] + e / n1;
  nt = (i64)np + nb + nh;
}

__global__ __launch_bounds__(256) void k_hf_pack(DevBasis B, const SegTask* tasks, const double* op, double* V,
                                                 const i64* voff) {
  const SegTask t = tasks[blockIdx.x];
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= t.cnt) return;
  int n1;
  i64 row, col, nt;
  hf_pos(B, t.q, t.g, e, n1, row, col, nt);
  const double x = op[B.goff[t.q * 6 + t.g] + e];
  double* Vq = V + voff[t.q];
  Vq[row + col * nt] = x;
  if (!is_square(t.g)) Vq[col + row * nt] = x;
}

__global__ __launch_bounds__(256) void k_hf_unpack(DevBasis B, const SegTask* tasks, const double* V,
                                                   const i64* voff, double* op) {
  const SegTask t = tasks[blockIdx.x];
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= t.cnt) return;
  int n1;
  i64 row, col, nt;
  hf_pos(B, t.q, t.g, e, n1, row, col, nt);
  op[B.goff[t.q * 6 + t.g] + e] = V[voff[t.q] + row + col * nt];
}

void launch_hf_pack(const DevBasis& B, const SegTask* t, int nt, const double* op, double* V, const i64* voff,
                    hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_hf_pack, dim3(nt), dim3(kTile), 0, s, B, t, op, V, voff);
}

void launch_hf_unpack(const DevBasis& B, const SegTask* t, int nt, const double* V, const i64* voff, double* op,
                      hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_hf_unpack, dim3(nt), dim3(kTile), 0, s, B, t, V, voff, op);
}

}  // namespace imsrg

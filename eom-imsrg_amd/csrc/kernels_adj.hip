// Reverse-mode (vector-Jacobian) kernels of the scalar-commutator path, the
// transposes of the forward kernels in kernels.hip.  The reference has no
// adjoint; these give d<RESbar, [L,R]>/dL and /dR on the (anti)hermitian,
// Pauli-respecting operators the flow produces.  Every term's transpose reads
// the same commutator cotangent RESb: each forward term maps such operators
// to such operators, so its cotangent is RESb restricted to them.
//
//  * element gathers of the forward pass (121, 221, Pandya, 222_ph scatter)
//    become scatters with FP64 atomics into their sparse sources;
//  * 222_pp_hh's products are transposed by the grouped MFMA GEMM;
//  * 122's two-body part reuses the forward 122 kernel on the symmetrised,
//    Pauli-masked cotangent (a 1b x 2b commutator is its own transpose up
//    to herm factors); its one-body part sums per workgroup in LDS.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"

namespace imsrg {

namespace {

__device__ __forceinline__ void aadd(double* p, double v) { unsafeAtomicAdd(p, v); }

// add v into the storage position vat(p, op, herm) reads
__device__ __forceinline__ void vadj(const VPos& p, double* bar, int herm, double v) {
  if (p.off < 0) return;
  aadd(bar + p.off, (p.swp ? v * herm : v) * p.pre);
}

// add v into the storage position f_elem(B, op, herm, a, c) reads
__device__ __forceinline__ void fadj(const DevBasis& B, double* bar, int herm, int a, int c, double v) {
  const int c1 = B.con[a], c2 = B.con[c];
  i64 o;
  double s = 1.0;
  if (c1 + c2 == 0) o = B.off_fpp + B.cidx[a] + (i64)B.cidx[c] * B.P;
  else if (c1 + c2 == 2) o = B.off_fhh + B.cidx[a] + (i64)B.cidx[c] * B.A;
  else if (c1 > c2) { o = B.off_fph + B.cidx[c] + (i64)B.cidx[a] * B.P; s = herm; }
  else o = B.off_fph + B.cidx[a] + (i64)B.cidx[c] * B.P;
  aadd(bar + o, v * s);
}

// cotangent of the value u written at (IX,JX) of block g, the lower partner
// of a square block being filled with u*herm afterwards
__device__ __forceinline__ double fold2(const DevBasis& B, const double* Xb, int q, int g, int IX, int JX, int herm) {
  const int n1 = brows(B, q, g);
  const double* X = Xb + B.goff[q * 6 + g];
  if (!is_square(g)) return X[IX + (i64)JX * n1];
  if (IX == JX) return X[IX + (i64)IX * n1] * herm;
  return X[IX + (i64)JX * n1] + herm * X[JX + (i64)IX * n1];
}

// cotangent of the one-body value added by onebody_update
__device__ __forceinline__ double fold1(const DevBasis& B, const OneTask& t, const double* Xb, int herm) {
  if (t.kind == 2) return Xb[B.off_fph + t.pi + (i64)t.qi * B.P];
  const double* F = Xb + (t.kind == 0 ? B.off_fhh : B.off_fpp);
  const int ld = (t.kind == 0) ? B.A : B.P;
  if (t.pi == t.qi) return F[t.pi + (i64)t.pi * ld] * herm;
  return F[t.pi + (i64)t.qi * ld] + herm * F[t.qi + (i64)t.pi * ld];
}

__device__ double bsum(double v, double* red) {
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

}  // namespace

// ------------------------------------------------ 110 / 220 (E0) transpose
__global__ __launch_bounds__(256) void k_e0_vjp_1b(DevBasis B, const double* L, int Lh, const double* R, int Rh,
                                                   const double* RESb, double* Lb, double* Rb) {
  const i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (i64)B.P * B.A) return;
  const double f = RESb[0] * (Lh - Rh) * (B.jj[B.holes[(int)(x / B.P)]] + 1);
  const i64 o = B.off_fph + x;
  Lb[o] += f * R[o];
  Rb[o] += f * L[o];
}

__global__ __launch_bounds__(256) void k_e0_vjp_2b(DevBasis B, const SegTask* tasks, const double* L, int Lh,
                                                   const double* R, int Rh, const double* RESb, double* Lb,
                                                   double* Rb) {
  const SegTask t = tasks[blockIdx.x];
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= t.cnt) return;
  const i64 o = B.goff[t.q * 6 + 2] + e;
  const double c = -2.0 * RESb[0] * (B.chJ[t.q] + 1) * Rh;
  Lb[o] += c * R[o];
  Rb[o] += c * L[o];
}

void launch_e0_vjp(const DevBasis& B, const SegTask* g3, int ng3, const double* L, int Lh, const double* R, int Rh,
                   const double* RESb, double* Lb, double* Rb, hipStream_t s) {
  const i64 n = (i64)B.P * B.A;
  if (n > 0)
    hipLaunchKernelGGL(k_e0_vjp_1b, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, B, L, Lh, R, Rh, RESb, Lb,
                       Rb);
  if (Lh * Rh == -1 && ng3 > 0)
    hipLaunchKernelGGL(k_e0_vjp_2b, dim3(ng3), dim3(256), 0, s, B, g3, L, Lh, R, Rh, RESb, Lb, Rb);
}

// ------------------------------------------------------- 111 transpose
// One thread per one-body element of Lb (first E1 threads) or Rb.
// One wave per one-body cotangent element (Lb block elements first, then
// Rb); the lanes split the intermediate orbit and reduce across the wave.
__global__ __launch_bounds__(256) void k_111_vjp(DevBasis B, const double* L, int Lh, const double* R, int Rh,
                                                 const double* RESb, double* Lb, double* Rb) {
  const int A = B.A, P = B.P;
  const i64 E1 = (i64)A * A + (i64)P * A + (i64)P * P;
  const i64 idx = (i64)blockIdx.x * 4 + (threadIdx.x >> 6);   // wave-uniform
  const int lane = threadIdx.x & 63;
  if (idx >= 2 * E1) return;
  const bool side = idx >= E1;   // false: Lb, true: Rb
  const i64 x = side ? idx - E1 : idx;
  const double *Lhh = L + B.off_fhh, *Lph = L + B.off_fph, *Lpp = L + B.off_fpp;
  const double *Rhh = R + B.off_fhh, *Rph = R + B.off_fph, *Rpp = R + B.off_fpp;
  const double *Bhh = RESb + B.off_fhh, *Bph = RESb + B.off_fph, *Bpp = RESb + B.off_fpp;
  auto th1b = [&](int i, int j) { return Bhh[i + j * A] - Lh * Rh * Bhh[j + i * A]; };
  auto th2b = [&](int i, int j) { return Lh * Bhh[i + j * A] - Rh * Bhh[j + i * A]; };
  auto tp1b = [&](int p, int r) { return Rh * Bpp[p + (i64)r * P] - Lh * Bpp[r + (i64)p * P]; };
  auto tp2b = [&](int p, int r) { return Bpp[p + (i64)r * P] - Rh * Lh * Bpp[r + (i64)p * P]; };
  double s = 0.0;
  double* dst;
  if (x < (i64)A * A) {
    const int i = (int)(x % A), j = (int)(x / A);
    if (!side) {
      for (int k = lane; k < A; k += 64) s += th1b(i, k) * Rhh[j + k * A];
      for (int p = lane; p < P; p += 64) s -= Rph[p + (i64)i * P] * Bph[p + (i64)j * P];
      dst = Lb + B.off_fhh + x;
    } else {
      for (int k = lane; k < A; k += 64) s += Lhh[k + i * A] * th1b(k, j);
      for (int p = lane; p < P; p += 64) s += Lph[p + (i64)i * P] * Bph[p + (i64)j * P];
      dst = Rb + B.off_fhh + x;
    }
  } else if (x < (i64)A * A + (i64)P * A) {
    const i64 y = x - (i64)A * A;
    const int p = (int)(y % P), i = (int)(y / P);
    if (!side) {
      for (int j = lane; j < A; j += 64) s += Rph[p + (i64)j * P] * th2b(i, j) + Bph[p + (i64)j * P] * Rhh[i + j * A];
      for (int r = lane; r < P; r += 64)
        s += -Rpp[r + (i64)p * P] * Bph[r + (i64)i * P] + tp1b(p, r) * Rph[r + (i64)i * P];
      dst = Lb + B.off_fph + y;
    } else {
      for (int k = lane; k < A; k += 64) s += Lph[p + (i64)k * P] * th2b(k, i) - Bph[p + (i64)k * P] * Lhh[i + k * A];
      for (int r = lane; r < P; r += 64)
        s += Lpp[r + (i64)p * P] * Bph[r + (i64)i * P] + tp1b(r, p) * Lph[r + (i64)i * P];
      dst = Rb + B.off_fph + y;
    }
  } else {
    const i64 y = x - (i64)A * A - (i64)P * A;
    const int p = (int)(y % P), r = (int)(y / P);
    if (!side) {
      for (int i = lane; i < A; i += 64) s += Bph[p + (i64)i * P] * Rph[r + (i64)i * P];
      for (int k = lane; k < P; k += 64) s += tp2b(p, k) * Rpp[r + (i64)k * P];
      dst = Lb + B.off_fpp + y;
    } else {
      for (int i = lane; i < A; i += 64) s -= Bph[p + (i64)i * P] * Lph[r + (i64)i * P];
      for (int k = lane; k < P; k += 64) s += Lpp[k + (i64)p * P] * tp2b(k, r);
      dst = Rb + B.off_fpp + y;
    }
  }
  s = wave_sum(s);
  if (lane == 0) *dst += s;
}

void launch_111_vjp(const DevBasis& B, const double* L, int Lh, const double* R, int Rh, const double* RESb,
                    double* Lb, double* Rb, hipStream_t s) {
  const i64 n = 2 * ((i64)B.A * B.A + (i64)B.P * B.A + (i64)B.P * B.P);
  if (n <= 0) return;
  hipLaunchKernelGGL(k_111_vjp, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, B, L, Lh, R, Rh, RESb, Lb, Rb);
}

// ------------------------------------------------------- 121 transpose
// one workgroup per one-body output, threads over (same-block
// particle-hole pair, 2J) as in the forward kernel
__global__ __launch_bounds__(256) void k_121_vjp(DevBasis B, const OneTask* tasks, const double* L, int Lh,
                                                 const double* R, int Rh, const double* RESb, int RESh, double* Lb,
                                                 double* Rb) {
  const OneTask t = tasks[blockIdx.x];
  const int pk = t.pk, qk = t.qk, jq = B.jj[qk];
  const double sb = fold1(B, t, RESb, RESh) / (B.jj[pk] + 1.0);
  if (sb == 0.0) return;
  const double* Lph = L + B.off_fph;
  const double* Rph = R + B.off_fph;
  const int JS = B.jsmax;
  for (int x = threadIdx.x; x < B.nphp * JS; x += blockDim.x) {
    const int pr = x / JS, js = x % JS;
    const int a = B.php[2 * pr], i = B.php[2 * pr + 1];
    const int ak = B.parts[a], ik = B.holes[i];
    const int ji = B.jj[ik];
    const int JT = abs(ji - jq) + 2 * js;
    if (JT > ji + jq) continue;
    const i64 o = a + (i64)i * B.P;
    const double lf = Lph[o], rf = Rph[o];
    const VPos p1 = vpos(B, ak, pk, ik, qk, JT);
    const VPos p2 = vpos(B, ik, pk, ak, qk, JT);
    const double w = sb * (JT + 1);
    vadj(p1, Rb, Rh, w * lf * Lh);
    vadj(p2, Rb, Rh, -w * lf);
    vadj(p1, Lb, Lh, -w * rf * Rh);
    vadj(p2, Lb, Lh, w * rf);
    aadd(Lb + B.off_fph + o, w * (Lh * vat(p1, R, Rh) - vat(p2, R, Rh)));
    aadd(Rb + B.off_fph + o, -w * (Rh * vat(p1, L, Lh) - vat(p2, L, Lh)));
  }
}

void launch_121_vjp(const DevBasis& B, const OneTask* t, int nt, const double* L, int Lh, const double* R, int Rh,
                    const double* RESb, int RESh, double* Lb, double* Rb, hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_121_vjp, dim3(nt), dim3(256), 0, s, B, t, L, Lh, R, Rh, RESb, RESh, Lb, Rb);
}

// ------------------------------------------------------- 221 transpose
// RESb (one-body) -> cotangents of the w1 / w2 elements 221 reads, added
// into Wb1 / Wb2 (which already hold the 222_pp_hh part); threads over
// (intermediate orbit c, 2J)
__global__ __launch_bounds__(256) void k_221_vjp(DevBasis B, const OneTask* tasks, int Lh, int Rh,
                                                 const double* RESb, int RESh, double* Wb1, double* Wb2) {
  const OneTask t = tasks[blockIdx.x];
  const int ik = t.pk, jk = t.qk, ji = B.jj[ik];
  const int pm = Lh * Rh;
  const bool both = (t.kind != 2);
  const double sb = fold1(B, t, RESb, RESh) / (ji + 1.0);
  if (sb == 0.0) return;
  const int JS = B.jsmax;
  for (int y = threadIdx.x; y < B.N * JS; y += blockDim.x) {
    const int x = y / JS;
    const bool hole = x < B.A;
    const int ck = hole ? B.holes[x] : B.parts[x - B.A];
    double* W = hole ? Wb1 : Wb2;
    const int jc = B.jj[ck];
    const int JT = abs(jc - ji) + 2 * (y % JS);
    if (JT > jc + ji) continue;
    vadj(vpos(B, ck, ik, ck, jk, JT), W, 1, sb * (JT + 1));
    if (both) vadj(vpos(B, ck, jk, ck, ik, JT), W, 1, -pm * sb * (JT + 1));
  }
}

void launch_221_vjp(const DevBasis& B, const OneTask* t, int nt, int Lh, int Rh, const double* RESb, int RESh,
                    double* Wb1, double* Wb2, hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_221_vjp, dim3(nt), dim3(256), 0, s, B, t, Lh, Rh, RESb, RESh, Wb1, Wb2);
}

// ------------------------------------------- 222_pp_hh cotangent blocks
// Wb1 / Wb2 block g = cotangent of the w1 / w2 product of block g:
// square g: Y = X - pm X^T (X = RESb block), Wb1 = Y, Wb2 = -Y;
// rectangular g: Wb1 = X, Wb2 = -X.
__global__ __launch_bounds__(256) void k_pphh_wbar(DevBasis B, const SegTask* tasks, const double* RESb, int pm,
                                                   double* Wb1, double* Wb2) {
  const SegTask t = tasks[xcd_contig(blockIdx.x, gridDim.x)];
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= t.cnt) return;
  const int q = t.q, g = t.g;
  const i64 base = B.goff[q * 6 + g];
  double y;
  if (is_square(g)) {
    const int n1 = brows(B, q, g);
    const int IX = (int)(e % n1), JX = (int)(e / n1);
    y = RESb[base + e] - pm * RESb[base + JX + (i64)IX * n1];
  } else {
    y = RESb[base + e];
  }
  Wb1[base + e] = y;
  Wb2[base + e] = -y;
}

void launch_pphh_wbar(const DevBasis& B, const SegTask* t, int nt, const double* RESb, int pm, double* Wb1,
                      double* Wb2, hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_pphh_wbar, dim3(nt), dim3(256), 0, s, B, t, RESb, pm, Wb1, Wb2);
}

// ----------------------------------------------- 122 transpose helpers
// cbar = symmetrised cotangent: square blocks (X + h X^T)/2, rectangular
// X/2, rows and columns of Pauli-forbidden pairs (a,a) at odd J zeroed
__global__ __launch_bounds__(256) void k_to_dense(DevBasis B, const SegTask* tasks, const double* X, int herm,
                                                  double* cbar) {
  const SegTask t = tasks[xcd_contig(blockIdx.x, gridDim.x)];
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= t.cnt) return;
  const int q = t.q, g = t.g;
  const int n1 = brows(B, q, g);
  const int IX = (int)(e % n1), JX = (int)(e / n1);
  const i64 base = B.goff[q * 6 + g];
  double v;
  if (is_square(g)) v = 0.5 * (X[base + e] + herm * X[base + JX + (i64)IX * n1]);
  else v = 0.5 * X[base + e];
  if ((B.chJ[q] >> 1) & 1) {
    int a, b, c, d;
    unpack(B.pairs[B.pairoff[q * 3 + row_cls(g)] + IX], a, b);
    unpack(B.pairs[B.pairoff[q * 3 + col_cls(g)] + JX], c, d);
    if (a == b || c == d) v = 0.0;
  }
  cbar[base + e] = v;
}

void launch_to_dense(const DevBasis& B, const SegTask* t, int nt, const double* X, int herm, double* cbar,
                     hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_to_dense, dim3(nt), dim3(256), 0, s, B, t, X, herm, cbar);
}




// Both 122 transposes in one pass over the upper elements (the partner
// lists, element decoding and partner positions are shared):
// * two-body part: with cbar the symmetrised cotangent (one-body part
//   zero), y1 = [R_1b, cbar] and y2 = [L_1b, cbar] element by element (the
//   partner sums of k_comm122p), then Lb += -Rh y1 and Rb += Lh y2 on
//   stored elements, the lower partners of square blocks through their herm
//   (Lh for y1, Rh for y2), rectangular blocks twice (their transposed
//   partners are not stored);
// * one-body part: the partner sums of k_comm122p weighted by the cotangent
//   of each upper element, accumulated per workgroup in LDS (compact index
//   fk) and added to Lb / Rb once at the end.
// One workgroup per task, in XCD-contiguous order (neighbouring tiles share
// partner rows in one L2); LDS holds the one-body values of L and R and the
// workgroup's one-body accumulators.
// the transposed partner terms [x, xe) of one list (row partners: signs +,
// column partners: -), IMSRG_122_UNROLL records and their loads in flight
#ifndef IMSRG_122_UNROLL
#define IMSRG_122_UNROLL 2
#endif
#ifndef IMSRG_122_COPIES
#define IMSRG_122_COPIES 4
#endif
// The +-1 factors of a term (the record's sign, the herm factor of a
// transposed element: ch for cbar, Rh / Lh for R / L, and Lh / Rh of a
// transposed one-body element) combine into one sign bit per product,
// applied to pre = pre_other * (1 or sqrt 2) (pa1 / pa2); chneg / lneg /
// rneg: RESh / Lh / Rh < 0.
// RED: the list is the same for every lane of the wave (all active), so each
// term's one-body sums are reduced across the wave and added by one lane
template <bool ROW, bool RED = false>
__device__ __forceinline__ void vjp122(double& y1, double& y2, const Partner* part, i64 ps, int n,
                                       const PartnerAddr A, double pa1, double pa2, double sb, int chneg,
                                       const double* fL, const double* fR, double* lbar, double* rbar,
                                       const double* cbar, const double* L, int lneg, const double* R, int rneg) {
  constexpr int U = IMSRG_122_UNROLL;
  constexpr int col = ROW ? 0 : 1;
  const bool ob = sb != 0.0;   // the one-body part of this element
  int x = 0;
  while (x < n) {
    const int nu = min(U, n - x);
    Partner p[U];
    i64 o[U];
    double c[U], r[U], l[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (u < nu) p[u] = part[(x + u) * ps];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (u < nu) {
        o[u] = pa_off(A, (p[u].code >> 22) & 3, p[u].code & 0x3fffff);
        c[u] = cbar[o[u]];
        if (ob) { r[u] = R[o[u]]; l[u] = L[o[u]]; }
        else if (RED) { r[u] = 0.0; l[u] = 0.0; }
      }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (u < nu) {
        const int code = p[u].code;
        const int tr = pa_trans(A, (code >> 22) & 3);
        const int fs = (code >> 24) & 1, neg = ((code >> 25) & 1) ^ col;
        const double a = ((code >> 26) & 1) ? pa2 : pa1;
        const int fk = p[u].fk;
        const int tc = neg ^ (tr & chneg);
        y1 = fma(fR[fk] * ((tc ^ (fs & rneg)) ? -a : a), c[u], y1);
        y2 = fma(fL[fk] * ((tc ^ (fs & lneg)) ? -a : a), c[u], y2);
        if (RED) {
          const double al = (neg ^ (tr & rneg) ^ (fs & lneg)) ? -a : a;
          const double ar = (neg ^ 1 ^ (tr & lneg) ^ (fs & rneg)) ? -a : a;
          const double vl = wave_sum((sb * al) * r[u]), vr = wave_sum((sb * ar) * l[u]);
          if ((threadIdx.x & 63) == 0) {
            unsafeAtomicAdd(lbar + fk, vl);
            unsafeAtomicAdd(rbar + fk, vr);
          }
        } else if (ob) {
          const double al = (neg ^ (tr & rneg) ^ (fs & lneg)) ? -a : a;
          const double ar = (neg ^ 1 ^ (tr & lneg) ^ (fs & rneg)) ? -a : a;
          unsafeAtomicAdd(lbar + fk, (sb * al) * r[u]);
          unsafeAtomicAdd(rbar + fk, (sb * ar) * l[u]);
        }
      }
    x += nu;
  }
}

__global__ __launch_bounds__(256) void k_122_vjp(DevBasis B, const SegTask* tasks, int ntasks, const Partner* part,
                                                 const Partner* partT, int pstride, const int* rowp,
                                                 const int* colp, const int* f1off, int nf1,
                                                 const double* L, int Lh, const double* R, int Rh,
                                                 const double* cbar, const double* RESb, int RESh, double* Lb,
                                                 double* Rb, int ncopy) {
  // one-body sums of the row lists in ncopy copies (lane k adds to copy
  // k % ncopy): neighbouring lanes, rows that share their first orbit, often
  // add to the same element at once, and the LDS serialises such adds per
  // copy
  extern __shared__ double sh[];
  double* fL = sh;
  double* fR = sh + nf1;
  const int cs = 2 * nf1 + (ncopy > 1);   // stride of the copies
  double* lbar = sh + 2 * nf1 + (threadIdx.x & (ncopy - 1)) * cs;
  double* rbar = lbar + nf1;
  for (int k = threadIdx.x; k < nf1; k += blockDim.x) {
    const int o = f1off[k];
    fL[k] = L[o];
    fR[k] = R[o];
  }
  for (int k = threadIdx.x; k < ncopy * cs; k += blockDim.x) sh[2 * nf1 + k] = 0.0;
  __syncthreads();
  const int ch = RESh;
  for (int ti = xcd_contig(blockIdx.x, gridDim.x); ti < ntasks; ti += gridDim.x) {
    const SegTask t = tasks[ti];
    // lanes past the task's end stay active (the wave-wide sums below) on
    // its last element, with a zero cotangent and no stores
    const bool valid = (i64)t.start + threadIdx.x < t.cnt;
    const i64 e = valid ? (i64)t.start + threadIdx.x : t.cnt - 1;
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
    const double nrm = 1.0 / (sqrt(1.0 + (a == b ? 1.0 : 0.0)) * sqrt(1.0 + (c == d ? 1.0 : 0.0)));
    double sb = valid ? fold2(B, RESb, q, g, IX, JX, RESh) : 0.0;
    sb = sb / sqrt(1.0 + (a == b ? 1.0 : 0.0)) / sqrt(1.0 + (c == d ? 1.0 : 0.0));
    PartnerAddr AR, AC;
    partner_addr(B.goff + q * 6, B.dims + q * 3, rc, cc, IX, JX, AR, AC);
    double y1 = 0.0, y2 = 0.0;
    const int lneg = Lh < 0, rneg = Rh < 0, chneg = ch < 0;
    vjp122<true>(y1, y2, partT + gr, pstride, rowp[gr + 1] - rowp[gr], AR, pre_cd, pre_cd * 1.4142135623730951, sb,
                 chneg, fL, fR, lbar, rbar, cbar, L, lneg, R, rneg);
    // column partners through the scalar cache when the wave's active lanes
    // share one column pair
    const int gc0 = __builtin_amdgcn_readfirstlane((int)gc);
    if (__all((int)gc == gc0)) {
      const int x0 = colp[gc0], nx = colp[gc0 + 1] - x0;
      vjp122<false, true>(y1, y2, part + x0, 1, nx, AC, pre_ab, pre_ab * 1.4142135623730951, sb, chneg, fL, fR, lbar,
                          rbar, cbar, L, lneg, R, rneg);
    } else {
      vjp122<false>(y1, y2, part + colp[gc], 1, colp[gc + 1] - colp[gc], AC, pre_ab, pre_ab * 1.4142135623730951, sb,
                    chneg, fL, fR, lbar, rbar, cbar, L, lneg, R, rneg);
    }
    if (!valid) continue;
    y1 *= nrm;
    y2 *= nrm;
    const i64 o = B.goff[q * 6 + g] + IX + (i64)JX * n1;
    if (!sq) {
      Lb[o] += -2.0 * Rh * y1;
      Rb[o] += 2.0 * Lh * y2;
    } else if (IX == JX) {
      Lb[o] += -Rh * Lh * y1;
      Rb[o] += Lh * Rh * y2;
    } else {
      const i64 ol = B.goff[q * 6 + g] + JX + (i64)IX * n1;
      Lb[o] += -Rh * y1;
      Lb[ol] += -Rh * Lh * y1;
      Rb[o] += Lh * y2;
      Rb[ol] += Lh * Rh * y2;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nf1; k += blockDim.x) {
    double vl = 0.0, vr = 0.0;
    for (int c = 0; c < ncopy; ++c) {
      vl += sh[2 * nf1 + c * cs + k];
      vr += sh[2 * nf1 + c * cs + nf1 + k];
    }
    if (vl != 0.0) aadd(Lb + f1off[k], vl);
    if (vr != 0.0) aadd(Rb + f1off[k], vr);
  }
}

void launch_122_vjp(const DevBasis& B, const SegTask* t, int nt, const Partner* part, const Partner* partT,
                    int pstride, const int* rowp, const int* colp, const int* f1off, int nf1, const double* L, int Lh,
                    const double* R, int Rh, const double* cbar, const double* RESb, int RESh, double* Lb, double* Rb,
                    hipStream_t s) {
  if (nt <= 0) return;
  const int grid = nt;
  // as many copies of the one-body sums (4, 2 or 1) as 64 KB of LDS hold
  int ncopy = IMSRG_122_COPIES;
  auto bytes = [&](int nc) { return (2 * nf1 + nc * (2 * nf1 + (nc > 1))) * sizeof(double); };
  while (ncopy > 1 && bytes(ncopy) > 65536) ncopy >>= 1;
  const size_t lds = bytes(ncopy);
  hipLaunchKernelGGL(k_122_vjp, dim3(grid), dim3(256), lds, s, B, t, nt, part, partT, pstride, rowp, colp, f1off,
                     nf1, L, Lh, R, Rh, cbar, RESb, RESh, Lb, Rb, ncopy);
}

// -------------------------------------------------- 222_ph transpose
// The scatter of 222_ph (one output per quadruple (ij,kl) and J) reads W'
// of the cross-coupled block at 2J' through its upper triangle U:
// wprime(x, y) = U(x, y) for x <= y and -Lh Rh U(y, x) below.  Its
// transpose runs in three steps:
//   k_ph_sb:   the folded cotangent of every (quadruple, J) output, one lane
//              per output as in the forward scatter;
//   k_ph_coef: one lane per (quadruple, recoupling, J'): its J column of 6j
//              symbols against those cotangents;
//   k_ph_tile: one workgroup per 64 x 64 tile of U: the (lane, sign) terms
//              that land in the tile, summed in LDS and stored once.
// Where each term lands depends on the basis only, so the host sorts the
// terms by tile once; no global atomics, and every upper tile of U is
// written (no clearing pass).

__global__ __launch_bounds__(256) void k_ph_sb(DevBasis B, const QuadItem* items, int n, const double* RESb, int RESh,
                                               double* SB) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const QuadItem it = items[t];
  const int i = it.ijkl & 255, j = (it.ijkl >> 8) & 255, k = (it.ijkl >> 16) & 255, l = (it.ijkl >> 24) & 255;
  const int J = it.J;
  const int P = (B.ll[i] + B.ll[j]) & 1;
  const int q = 3 * J + 2 * ((B.tz[i] + B.tz[j]) / 2 + 1) + P;
  const int rc = B.con[i] + B.con[j], cc = B.con[k] + B.con[l];
  const int g = (rc == 0) ? cc : (rc == 1 ? (cc == 1 ? 3 : 5) : 4);
  const int IX = B.xmap[((i64)i * B.N + j) * B.nJ + (J >> 1)];
  const int JX = B.xmap[((i64)k * B.N + l) * B.nJ + (J >> 1)];
  const double pre = ((i == j) ? IMSRG_INV_SQRT2_REF : 1.0) * ((k == l) ? IMSRG_INV_SQRT2_REF : 1.0);
  SB[t] = fold2(B, RESb, q, g, IX, JX, RESh) * pre;
}

// lanes[t] = quadruple index * 64 + slot; slots [0, n1) are the J' of the
// first recoupling (coefficient -(-1)^((jl+jk+J)/2) {ji jj J; jl jk J'}),
// [n1, n1 + n2) those of the second ({ji jj J; jk jl J'})
__global__ __launch_bounds__(256) void k_ph_coef(DevBasis B, const int* lanes, int nl, const int* quads,
                                                 const int* qstart, const double* SB, double* CV) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nl) return;
  const int v = lanes[t];
  const int qd = v >> 6, slot = v & 63;
  const int ijkl = quads[qd];
  const double* sb = SB + qstart[qd];
  const int i = ijkl & 255, j = (ijkl >> 8) & 255, k = (ijkl >> 16) & 255, l = (ijkl >> 24) & 255;
  const int ji = B.jj[i], jj = B.jj[j], jk = B.jj[k], jl = B.jj[l];
  const int lo = max(abs(ji - jj), abs(jk - jl));
  const int nj = (min(ji + jj, jk + jl) - lo) / 2 + 1;
  const int a1 = max(abs(jj - jl), abs(ji - jk)), b1 = min(jj + jl, ji + jk);
  const int n1 = b1 >= a1 ? (b1 - a1) / 2 + 1 : 0;
  double c = 0.0;
  if (slot < n1) {
    const int JP = a1 + 2 * slot;
    for (int x = 0; x < nj; ++x) {
      const int J = lo + 2 * x;
      c -= m1pow((jl + jk + J) / 2) * sixj(B, ji, jj, J, jl, jk, JP) * sb[x];
    }
  } else {
    const int JP = max(abs(ji - jl), abs(jj - jk)) + 2 * (slot - n1);
    for (int x = 0; x < nj; ++x) c += sixj(B, ji, jj, lo + 2 * x, jk, jl, JP) * sb[x];
  }
  CV[t] = c;
}

// terms[tstart[w] .. tstart[w+1]) land in tile w: (lane, bits) with bits
// 0-11 the position (x - x0) + 64 (y - y0), bit 12 a factor -1 (phase of the
// second read of a recoupling), bit 13 a factor -Lh Rh (read below the
// diagonal)
__global__ __launch_bounds__(256) void k_ph_tile(DevBasis B, const PairTask* tiles, const int* tstart,
                                                 const int2* terms, const double* CV, double mpm, double* U) {
  __shared__ double acc[64 * 64];
  for (int e = threadIdx.x; e < 64 * 64; e += 256) acc[e] = 0.0;
  __syncthreads();
  const int wb = xcd_contig(blockIdx.x, gridDim.x);
  const int e0 = tstart[wb], e1 = tstart[wb + 1];
  for (int e = e0 + threadIdx.x; e < e1; e += 256) {
    const int2 tm = terms[e];
    double v = CV[tm.x];
    if (tm.y & 4096) v = -v;
    if (tm.y & 8192) v *= mpm;
    atomicAdd(&acc[tm.y & 4095], v);
  }
  __syncthreads();
  const PairTask tk = tiles[wb];
  const int r = B.rlen[tk.q];
  double* W = U + B.wcoff[tk.q];
  const int x0 = tk.ti * 64, y0 = tk.tj * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int x = x0 + (e & 63), y = y0 + (e >> 6);
    if (x < r && y < r) W[x + (i64)y * r] = acc[e];
  }
}

void launch_ph_vjp(const DevBasis& B, const QuadItem* items, int nitems, const int* lanes, int nl, const int* quads,
                   const int* qstart, const PairTask* tiles, int ntiles, const int* tstart, const int2* terms,
                   const double* RESb, int RESh, int Lh, int Rh, double* SB, double* CV, double* U, hipStream_t s) {
  if (nitems <= 0 || nl <= 0 || ntiles <= 0) return;
  hipLaunchKernelGGL(k_ph_sb, dim3((nitems + 255) / 256), dim3(256), 0, s, B, items, nitems, RESb, RESh, SB);
  hipLaunchKernelGGL(k_ph_coef, dim3((nl + 255) / 256), dim3(256), 0, s, B, lanes, nl, quads, qstart, SB, CV);
  hipLaunchKernelGGL(k_ph_tile, dim3(ntiles), dim3(256), 0, s, B, tiles, tstart, terms, CV, -(double)(Lh * Rh), U);
}


// calculate_cross_coupled transpose by gathering (no atomics, fixed order):
// each stored two-body element that some Pandya term reads sums the
// cotangents of the cross-coupled elements reading it.  A term of CC element
// (JC; row (a,b), column (p,h)) at 2J' reads v_elem(h,a,p,b,J'), stored where
// vpos puts it, so the readers of position o of channel J' are the orbit
// assignments (h,a) / (p,b) from its row pair / column pair (either order,
// either orientation) that vpos maps to o, with h a hole and p a particle;
// blocks pppp and hhhh are never read.  Every reader adds, over JC,
//   CCb * (-1)^((jh+jb+JC)/2) sqrt(JC+1) (-1)^(J'/2) (J'+1) {jp jh JC; ja jb J'}
// times vpos's phase (and herm when stored transposed), as the forward adds
// the element into the CC element.
__device__ __forceinline__ double pandya_reader(const DevBasis& B, const double* CCb, int h, int a, int p, int b,
                                                int Jp) {
  const int N = B.N, nJ = B.nJ, JTM = B.JTM;
  const int ja = B.jj[a], jb = B.jj[b], jp = B.jj[p], jh = B.jj[h];
  const int TZ = abs(B.tz[a] - B.tz[b]) / 2, PAR = (B.ll[a] + B.ll[b]) & 1;
  if (abs(B.tz[p] - B.tz[h]) / 2 != TZ || ((B.ll[p] + B.ll[h]) & 1) != PAR) return 0.0;
  const int JCmin = max(abs(ja - jb), abs(jp - jh)), JCmax = min(ja + jb, jp + jh);
  double sm = 0.0;
  for (int JC = JCmin; JC <= JCmax; JC += 2) {
    const int r = B.rmap[((i64)a * N + b) * nJ + (JC >> 1)];
    const int c = B.nbmap[((i64)p * N + h) * nJ + (JC >> 1)];
    if (r < 0 || c < 0) continue;
    const int q1 = (JC >> 1) + TZ * (JTM + 1) + 2 * PAR * (JTM + 1);
    const double cb = CCb[B.ccoff[q1] + r + (i64)c * B.rlen[q1]];
    sm = sm + cb * m1pow((jh + jb + JC) / 2) * sqrt(JC + 1.0) * sixj(B, jp, jh, JC, ja, jb, Jp);
  }
  return sm * (double)(m1pow(Jp / 2) * (Jp + 1));
}

__global__ __launch_bounds__(256) void k_pandya_vjp_gather(DevBasis B, const SegTask* tasks, const double* CCb,
                                                           int herm, double* opb) {
  const SegTask t = tasks[xcd_contig(blockIdx.x, gridDim.x)];
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= t.cnt) return;
  const int q = t.q, g = t.g;
  const int n1 = brows(B, q, g);
  const int IX = (int)(e % n1), JX = (int)(e / n1);
  int x1, x2, y1, y2;
  unpack(B.pairs[B.pairoff[q * 3 + row_cls(g)] + IX], x1, x2);
  unpack(B.pairs[B.pairoff[q * 3 + col_cls(g)] + JX], y1, y2);
  const int Jp = B.chJ[q];
  const i64 o = B.goff[q * 6 + g] + e;
  // the eight assignments (h,a,p,b); repeats (x1 = x2, y1 = y2, or the row
  // pair equal to the column pair) are counted once
  const int H[8] = {x1, x2, x1, x2, y1, y2, y1, y2};
  const int A[8] = {x2, x1, x2, x1, y2, y1, y2, y1};
  const int Pp[8] = {y1, y1, y2, y2, x1, x1, x2, x2};
  const int Bb[8] = {y2, y2, y1, y1, x2, x2, x1, x1};
  double tot = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int h = H[k], a = A[k], p = Pp[k], b = Bb[k];
    if (B.con[h] != 1 || B.con[p] != 0) continue;
    bool dup = false;
#pragma unroll
    for (int m = 0; m < k; ++m) dup = dup || (H[m] == h && A[m] == a && Pp[m] == p && Bb[m] == b);
    if (dup) continue;
    const VPos vp = vpos(B, h, a, p, b, Jp);
    if (vp.off != o) continue;
    const double v = pandya_reader(B, CCb, h, a, p, b, Jp);
    tot = tot + (vp.swp ? v * herm : v) * vp.pre;
  }
  if (tot != 0.0) opb[o] += tot;
}

void launch_pandya_vjp_gather(const DevBasis& B, const SegTask* t, int nt, const double* CCb, int herm, double* opb,
                              hipStream_t s) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_pandya_vjp_gather, dim3(nt), dim3(256), 0, s, B, t, CCb, herm, opb);
}

// ------------------------------------------------- white / normal order
// gen_core transpose (build_gs_white, mode 0: eta = Hn / Eden;
// build_gs_atan / atanxx, mode 1: eta = 0.5 atan(2 Hn / Eden), both with the
// denominators of Hd): Hnb += deta/dHn ETAb, the cotangent of Eden goes to
// Hdb (one-body) and Mb (monopoles of Hd).  Hdb and Hnb may be the same
// operator.  Mode 1 scales both partial derivatives by 1 / (1 + u^2),
// u = 2 Hn / Eden.
__device__ __forceinline__ double gen_damp(double x, double Eden, int mode) {
  if (mode == 0) return 1.0;
  const double u = 2 * x / Eden;
  return 1.0 / (1.0 + u * u);
}

__global__ __launch_bounds__(256) void k_gen_vjp_1b(DevBasis B, const double* Hd, const double* Hn, const double* ETAb,
                                                    double* Hdb, double* Hnb, int mode) {
  const i64 x = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= (i64)B.P * B.A) return;
  const int a = (int)(x % B.P), i = (int)(x / B.P);
  const int ak = B.parts[a], ik = B.holes[i];
  if (B.jj[ik] != B.jj[ak] || B.ll[ik] != B.ll[ak] || B.tz[ik] != B.tz[ak]) return;
  const double g = ETAb[B.off_fph + x];
  if (g == 0.0) return;
  const int ji = B.jj[ik];
  double Eden = 0.0;
  for (int JT = 0; JT <= 2 * ji; JT += 2) Eden = Eden - (JT + 1) * v_elem(B, Hd, 1, ak, ik, ak, ik, JT);
  Eden = Eden / (ji + 1.0) / (ji + 1.0);
  Eden = Eden + Hd[B.off_fpp + a + (i64)a * B.P] - Hd[B.off_fhh + i + (i64)i * B.A];
  const double xv = Hn[B.off_fph + x], d = gen_damp(xv, Eden, mode);
  aadd(Hnb + B.off_fph + x, g * d / Eden);
  const double Eb = -g * xv * d / (Eden * Eden);
  aadd(Hdb + B.off_fpp + a + (i64)a * B.P, Eb);
  aadd(Hdb + B.off_fhh + i + (i64)i * B.A, -Eb);
  for (int JT = 0; JT <= 2 * ji; JT += 2)
    vadj(vpos(B, ak, ik, ak, ik, JT), Hdb, 1, -Eb * (JT + 1) / (ji + 1.0) / (ji + 1.0));
}

// two-body part; the cotangents of the monopoles go to Mb (N x N)
__global__ __launch_bounds__(256) void k_gen_vjp_2b(DevBasis B, const SegTask* tasks, const double* Hd,
                                                    const double* Hn, const double* M, const double* ETAb,
                                                    double* Hdb, double* Hnb, double* Mb, int mode) {
  const SegTask t = tasks[blockIdx.x];
  const i64 e = (i64)t.start + threadIdx.x;
  if (e >= t.cnt) return;
  const int q = t.q;
  const i64 o = B.goff[q * 6 + 2] + e;
  const double g = ETAb[o];
  if (g == 0.0) return;
  const int np = B.dims[q * 3];
  const int IX = (int)(e % np), JX = (int)(e / np);
  int a, b, i, j;
  unpack(B.pairs[B.pairoff[q * 3 + 0] + IX], a, b);
  unpack(B.pairs[B.pairoff[q * 3 + 2] + JX], i, j);
  const int N = B.N;
  double Eden = M[a + b * N] + M[i + j * N] - M[a + i * N] - M[a + j * N] - M[i + b * N] - M[j + b * N];
  Eden = Eden + f_elem(B, Hd, 1, a, a) + f_elem(B, Hd, 1, b, b) - f_elem(B, Hd, 1, i, i) - f_elem(B, Hd, 1, j, j);
  const double xv = Hn[o], d = gen_damp(xv, Eden, mode);
  aadd(Hnb + o, g * d / Eden);
  const double Eb = -g * xv * d / (Eden * Eden);
  aadd(Mb + a + b * N, Eb);
  aadd(Mb + i + j * N, Eb);
  aadd(Mb + a + i * N, -Eb);
  aadd(Mb + a + j * N, -Eb);
  aadd(Mb + i + b * N, -Eb);
  aadd(Mb + j + b * N, -Eb);
  fadj(B, Hdb, 1, a, a, Eb);
  fadj(B, Hdb, 1, b, b, Eb);
  fadj(B, Hdb, 1, i, i, -Eb);
  fadj(B, Hdb, 1, j, j, -Eb);
}

__global__ __launch_bounds__(256) void k_monopole_vjp(DevBasis B, const double* Mb, double* Hb) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B.N * B.N) return;
  const double m = Mb[idx];
  if (m == 0.0) return;
  const int a = idx % B.N, b = idx / B.N;
  const int ja = B.jj[a], jb = B.jj[b];
  for (int JT = abs(ja - jb); JT <= ja + jb; JT += 2)
    vadj(vpos(B, a, b, a, b, JT), Hb, 1, m * (JT + 1) / (ja + 1.0) / (jb + 1.0));
}

void launch_gen_vjp(const DevBasis& B, const SegTask* g3, int ng3, const double* Hd, const double* Hn,
                    const double* M, const double* ETAb, double* Hdb, double* Hnb, double* Mb, int mode,
                    hipStream_t s) {
  const i64 n = (i64)B.P * B.A;
  if (n > 0)
    hipLaunchKernelGGL(k_gen_vjp_1b, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, B, Hd, Hn, ETAb, Hdb, Hnb,
                       mode);
  if (ng3 > 0) hipLaunchKernelGGL(k_gen_vjp_2b, dim3(ng3), dim3(256), 0, s, B, g3, Hd, Hn, M, ETAb, Hdb, Hnb, Mb, mode);
  hipLaunchKernelGGL(k_monopole_vjp, dim3((B.N * B.N + 255) / 256), dim3(256), 0, s, B, Mb, Hdb);
}

// normal_order transpose.  Tb must hold a copy of Hb on entry.  Pass 1: the
// one-body elements normal_order replaced (their T values do not reach H);
// pass 2: E0.
__global__ __launch_bounds__(256) void k_no_vjp_1b(DevBasis B, const double* Hb, double* Tb) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B.N * B.N) return;
  const int ax = idx % B.N, bx = idx / B.N;
  if (B.jj[ax] != B.jj[bx] || B.ll[ax] != B.ll[bx] || B.tz[ax] != B.tz[bx]) return;
  const int ca = B.con[ax], cb = B.con[bx];
  if (ca == 1 && cb == 1 && B.cidx[ax] > B.cidx[bx]) return;
  if (ca == 0 && cb == 0 && B.cidx[ax] > B.cidx[bx]) return;
  if (ca == 1 && cb == 0) return;
  const int a = B.cidx[ax], b = B.cidx[bx];
  double vb;
  if (ca == 1) {
    const i64 o1 = B.off_fhh + a + (i64)b * B.A, o2 = B.off_fhh + b + (i64)a * B.A;
    vb = Hb[o1] + (a != b ? Hb[o2] : 0.0);
    Tb[o2] = 0.0;
    Tb[o1] = vb;          // f_elem(ax, bx) of T reads fhh(a, b)
  } else if (cb == 1) {
    const i64 o1 = B.off_fph + a + (i64)b * B.P;
    vb = Hb[o1];
    Tb[o1] = vb;
  } else {
    const i64 o1 = B.off_fpp + a + (i64)b * B.P, o2 = B.off_fpp + b + (i64)a * B.P;
    vb = Hb[o1] + (a != b ? Hb[o2] : 0.0);
    Tb[o2] = 0.0;
    Tb[o1] = vb;
  }
  if (vb == 0.0) return;
  for (int ix = 0; ix < B.N; ++ix) {
    if (B.con[ix] != 1) continue;
    for (int JT = abs(B.jj[ix] - B.jj[ax]); JT <= B.jj[ix] + B.jj[ax]; JT += 2)
      vadj(vpos(B, ax, ix, bx, ix, JT), Tb, 1, vb * (JT + 1.0) / (B.jj[ax] + 1));
  }
}

__global__ __launch_bounds__(256) void k_no_vjp_e0(DevBasis B, const double* Hb, double* Tb) {
  const double e0 = Hb[0];
  if (threadIdx.x == 0 && blockIdx.x == 0) Tb[0] = 0.0;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B.N * B.N || e0 == 0.0) return;
  const int i = idx % B.N, j = idx / B.N;
  if (i == j && B.con[i] == 1) fadj(B, Tb, 1, i, i, e0 * (B.jj[i] + 1));
  if (B.con[i] * B.con[j] == 0) return;
  for (int JT = abs(B.jj[i] - B.jj[j]); JT <= B.jj[i] + B.jj[j]; JT += 2)
    vadj(vpos(B, i, j, i, j, JT), Tb, 1, 0.5 * e0 * (JT + 1));
}

void launch_normal_order_vjp(const DevBasis& B, const double* Hb, double* Tb, hipStream_t s) {
  const int n = B.N * B.N;
  hipLaunchKernelGGL(k_no_vjp_1b, dim3((n + 255) / 256), dim3(256), 0, s, B, Hb, Tb);
  hipLaunchKernelGGL(k_no_vjp_e0, dim3((n + 255) / 256), dim3(256), 0, s, B, Hb, Tb);
}

// dst[0] += factor * src[0]
__global__ void k_add_e0(double* dst, const double* src, double factor) { dst[0] += factor * src[0]; }
void launch_add_e0(double* dst, const double* src, double factor, hipStream_t s) {
  hipLaunchKernelGGL(k_add_e0, dim3(1), dim3(1), 0, s, dst, src, factor);
}

}  // namespace imsrg

// Kernels of the scalar-tensor commutator (TS_commutators.f90:10-1849) and
// the tensor cross-coupled transform (calculate_single_pandya,
// cross_coupled.f90:1416-1634).  One thread per two-body output element
// (gathers, no atomics: bit-reproducible), one wave per one-body output
// element; the angular-momentum recouplings read host-built 6j tables.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "tensor.hpp"

namespace imsrg {

namespace {

// xxxsixj, basic_IMSRG.f90:2602-2640: {J1 J2 rank; j4 j5 j6} (zero where a
// triangle fails, as stored)
__device__ __forceinline__ double x6(const DevTensor& T, int J1, int J2, int j4, int j5, int j6) {
  return T.x6[(((((i64)(J1 >> 1) * T.nJ + (J2 >> 1)) * T.nhx + (j4 >> 1)) * T.nhx + (j5 >> 1)) * T.nhx) +
              (j6 >> 1)];
}
// sixj {c d J2; b x J4} (basic_IMSRG.f90:2553-2599) with x up to jmax + rank
__device__ __forceinline__ double s6(const DevTensor& T, int c, int d, int J2, int b, int x, int J4) {
  return T.s6[(((((i64)(c >> 1) * T.nho + (d >> 1)) * T.nho + (b >> 1)) * T.nhx + (x >> 1)) * T.nJ + (J2 >> 1)) *
                  T.nJ +
              (J4 >> 1)];
}
// ninej, basic_IMSRG.f90:4425-4440: the 9j {a b J1; c d J2; J3 J4 rank}
__device__ __forceinline__ double ninej(const DevTensor& T, int a, int b, int J1, int c, int d, int J2, int J3,
                                        int J4) {
  const int R = T.rank;
  const int xmin = max(max(abs(J2 - b), abs(a - R)), abs(J4 - c));
  const int xmax = min(min(J2 + b, a + R), J4 + c);
  double sm = 0.0;
  for (int x = xmin; x <= xmax; x += 2)
    sm = sm - (x + 1.0) * x6(T, J1, J2, x, a, b) * s6(T, c, d, J2, b, x, J4) * x6(T, J3, J4, x, a, c);
  return sm;
}

// the 9j {u1 u2 J3; u4 u5 J4; J1 J2 rank} sqrt((J3+1)(J4+1)) for J3 <= J4,
// J1 <= J2 (host table)
__device__ __forceinline__ double n9(const DevTensor& T, int u1, int u2, int J3, int u4, int u5, int J4, int J1,
                                     int J2) {
  const int nR = T.rank / 2 + 1;
  return T.n9[(((((((i64)(u1 >> 1) * T.nho + (u2 >> 1)) * T.nho + (u4 >> 1)) * T.nho + (u5 >> 1)) * T.nJ +
                   (J3 >> 1)) * nR + ((J4 - J3) >> 1)) * T.nJ + (J1 >> 1)) * nR + ((J2 - J1) >> 1)];
}

__device__ __forceinline__ int tsub(int C1, int C2) {   // C = con + con + 1; tensor_elem's qx arithmetic
  int qx = C1 * C2;
  if (qx == 9) qx = 5;
  if (C1 > C2) qx += (qx == 2 ? 6 : (qx == 3 ? 4 : (qx == 6 ? 3 : 0)));
  return qx - 1;
}

// tensor_elem, basic_IMSRG.f90:1548-1685
__device__ double tensor_elem(const DevBasis& B, const DevTensor& T, const double* op, int herm, int ax, int bx,
                              int cx, int dx, int J1x, int J2x) {
  int a, b, c, d, J1, J2;
  double phase = 1.0;
  bool sw;
  if (J1x <= J2x) { sw = false; J1 = J1x; J2 = J2x; a = ax; b = bx; c = cx; d = dx; }
  else { sw = true; J1 = J2x; J2 = J1x; a = cx; b = dx; c = ax; d = bx; }
  const int R = T.rank;
  if (!(dtri(J1, J2, R) && J2 - J1 <= R)) return 0.0;
  const int ja = B.jj[a], jb = B.jj[b], jc = B.jj[c], jd = B.jj[d];
  if (!(dtri(ja, jb, J1) && dtri(jc, jd, J2))) return 0.0;
  const int P = (B.ll[a] + B.ll[b]) & 1;
  if (((B.ll[c] + B.ll[d]) & 1) != (P ^ ((T.dpar >> 1) & 1))) return 0.0;
  const int T2 = B.tz[a] + B.tz[b];
  if (B.tz[c] + B.tz[d] != T2) return 0.0;
  if (J1 > 2 * B.JTM || J2 > 2 * B.JTM) return 0.0;
  const int q = T.look[((J1 >> 1) * (R / 2 + 1) + ((J2 - J1) >> 1)) * 6 + 2 * (T2 / 2 + 1) + P];
  if (sw) phase = (double)(m1pow((J1 - J2) / 2) * herm);
  const int g = tsub(B.con[a] + B.con[b] + 1, B.con[c] + B.con[d] + 1);
  double pre = 1.0;
  int i1, i2;
  if (a > b) {
    i1 = B.xmap[((i64)b * B.N + a) * B.nJ + (J1 >> 1)];
    pre = m1pow(1 + (ja + jb - J1) / 2);
  } else {
    if (a == b) pre = 1.4142135623730951;
    i1 = B.xmap[((i64)a * B.N + b) * B.nJ + (J1 >> 1)];
  }
  if (c > d) {
    i2 = B.xmap[((i64)d * B.N + c) * B.nJ + (J2 >> 1)];
    pre = pre * m1pow(1 + (jc + jd - J2) / 2);
  } else {
    if (c == d) pre = pre * 1.4142135623730951;
    i2 = B.xmap[((i64)c * B.N + d) * B.nJ + (J2 >> 1)];
  }
  const int ld = T.tdim[q * 6 + kTSea1[g]];
  return op[T.tgoff[q * 9 + g] + i1 + (i64)i2 * ld] * pre * phase;
}

// f_tensor_elem, basic_IMSRG.f90:1261-1289
__device__ __forceinline__ double f_tensor_elem(const DevBasis& B, const double* op, int herm, int a, int c) {
  const int c1 = B.con[a], c2 = B.con[c];
  if (c1 + c2 == 0) return op[B.off_fpp + B.cidx[a] + (i64)B.cidx[c] * B.P];
  if (c1 + c2 == 2) return op[B.off_fhh + B.cidx[a] + (i64)B.cidx[c] * B.A];
  if (c1 > c2) return op[B.off_fph + B.cidx[c] + (i64)B.cidx[a] * B.P] * herm * m1pow((B.jj[a] - B.jj[c]) / 2);
  return op[B.off_fph + B.cidx[a] + (i64)B.cidx[c] * B.P];
}

// TS_commutator_111, TS_commutators.f90:10-55: one thread per one-body
// element (fhh, fph, fpp), the reference's products summed directly
__global__ __launch_bounds__(256) void k_ts111(DevBasis B, DevTensor T, const double* L, int Lh, const double* R,
                                               int Rh, double* RES) {
  const i64 A = B.A, P = B.P;
  i64 e = (i64)blockIdx.x * 256 + threadIdx.x;
  const double *Lhh = L + B.off_fhh, *Lph = L + B.off_fph, *Lpp = L + B.off_fpp;
  const double *Rhh = R + B.off_fhh, *Rph = R + B.off_fph, *Rpp = R + B.off_fpp;
  if (e < A * A) {
    const int i = (int)(e % A), j = (int)(e / A);
    double t1ij = 0.0, t2ij = 0.0, t1ji = 0.0, t2ji = 0.0;
    for (int k = 0; k < A; ++k) {
      t1ij += Lhh[i + k * A] * Rhh[k + j * A];
      t1ji += Lhh[j + k * A] * Rhh[k + i * A];
    }
    for (int a = 0; a < P; ++a) {
      t2ij += Lph[a + i * P] * Rph[a + j * P];
      t2ji += Lph[a + j * P] * Rph[a + i * P];
    }
    const double ph = m1pow((B.jj[B.holes[i]] - B.jj[B.holes[j]]) / 2);
    RES[B.off_fhh + e] = t1ij + t2ij * Lh - (t1ji * Lh + t2ji) * Rh * ph;
    return;
  }
  e -= A * A;
  if (e < P * A) {
    const int a = (int)(e % P), i = (int)(e / P);
    double s = 0.0;
    for (int k = 0; k < A; ++k) s += Lph[a + k * P] * Rhh[k + i * A];
    for (int b = 0; b < P; ++b) s += Lpp[a + b * P] * Rph[b + i * P];
    double u = 0.0;
    for (int k = 0; k < A; ++k) u += Rph[a + k * P] * Lhh[k + i * A];
    for (int b = 0; b < P; ++b) u += Rpp[a + b * P] * Lph[b + i * P];
    RES[B.off_fph + e] = s - u;
    return;
  }
  e -= P * A;
  if (e < P * P) {
    const int i = (int)(e % P), j = (int)(e / P);
    double p1ij = 0.0, p2ij = 0.0, p1ji = 0.0, p2ji = 0.0;
    for (int k = 0; k < A; ++k) {
      p1ij += Lph[i + k * P] * Rph[j + k * P];
      p1ji += Lph[j + k * P] * Rph[i + k * P];
    }
    for (int b = 0; b < P; ++b) {
      p2ij += Lpp[i + b * P] * Rpp[b + j * P];
      p2ji += Lpp[j + b * P] * Rpp[b + i * P];
    }
    const double ph = m1pow((B.jj[B.parts[i]] - B.jj[B.parts[j]]) / 2);
    RES[B.off_fpp + e] = p1ij * Rh * ph + p2ij - (p1ji + p2ji * Rh * ph) * Lh;
  }
}

// one wave per one-body output element: mode 0 TS_commutator_121
// (TS_commutators.f90:58-303), 1 TS_commutator_211 (307-552), 2
// TS_commutator_221 (809-1067); lane 0 adds the wave's sum and writes the
// mirrored element as the reference does
template <int MODE>
__global__ __launch_bounds__(256) void k_ts_one(DevBasis B, DevTensor T, const TOne* tasks, int nt, const double* L,
                                                int Lh, const double* X, const double* Y, const double* R, int Rh,
                                                double* RES, int RESh) {
  const int lane = threadIdx.x & 63;
  const int ti = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (ti >= nt) return;
  const TOne t = tasks[ti];
  const int pk = t.cls == 0 ? B.holes[t.p] : B.parts[t.p];
  const int qk = t.cls == 1 ? B.parts[t.q] : B.holes[t.q];
  const int jp = B.jj[pk], jq = B.jj[qk];
  const int rank = T.rank;
  double sm = 0.0;
  if constexpr (MODE == 0) {
    for (int k = lane; k < B.nphp; k += 64) {
      const int a = B.php[2 * k], i = B.php[2 * k + 1];
      const int ak = B.parts[a], ik = B.holes[i], ja = B.jj[ak], ji = B.jj[ik];
      double smx = 0.0, smy = 0.0;
      for (int J2 = abs(ji - jq); J2 <= ji + jq; J2 += 2)
        for (int J1 = abs(ja - jp); J1 <= ja + jp; J1 += 2)
          smx = smx + tensor_elem(B, T, R, Rh, ak, pk, ik, qk, J1, J2) * sqrt((J1 + 1.0) * (J2 + 1.0)) *
                          m1pow(J1 / 2) * x6(T, J1, J2, jq, jp, ji);
      for (int J1 = abs(ji - jp); J1 <= ji + jp; J1 += 2)
        for (int J2 = abs(ja - jq); J2 <= ja + jq; J2 += 2)
          smy = smy + tensor_elem(B, T, R, Rh, ik, pk, ak, qk, J1, J2) * sqrt((J1 + 1.0) * (J2 + 1.0)) *
                          m1pow(J1 / 2) * x6(T, J1, J2, jq, jp, ji);
      sm = sm + L[B.off_fph + a + (i64)i * B.P] * (Lh * smx - smy) * m1pow((rank + jq + ji) / 2);
    }
  } else if constexpr (MODE == 1) {
    const int tp = B.tz[pk], tq = B.tz[qk];
    for (int k = lane; k < B.P * B.A; k += 64) {
      const int a = k % B.P, i = k / B.P;
      const int ak = B.parts[a], ik = B.holes[i];
      const int ja = B.jj[ak], ji = B.jj[ik], la = B.ll[ak], li = B.ll[ik], ta = B.tz[ak], ti_ = B.tz[ik];
      if ((li & 1) != ((la + (T.dpar >> 1)) & 1)) continue;
      if (ti_ != ta) continue;
      if (!dtri(ja, ji, rank)) continue;
      if (tp != tq) continue;
      const int PAR = (la + li) & 1;
      if (((B.ll[pk] + B.ll[qk]) & 1) != PAR) continue;
      const int qx = (rank >> 1) + 2 * PAR * (B.JTM + 1);
      const int rai = B.nbmap[((i64)ak * B.N + ik) * B.nJ + (rank >> 1)];
      const int rpq = B.rmap[((i64)pk * B.N + qk) * B.nJ + (rank >> 1)];
      const int rqp = B.rmap[((i64)qk * B.N + pk) * B.nJ + (rank >> 1)];
      const double* CX = X + B.ccoff[qx];
      const i64 rl = B.rlen[qx];
      sm = sm + R[B.off_fph + a + (i64)i * B.P] *
                    (m1pow((jp - jq + rank) / 2) * CX[rpq + rai * rl] -
                     Rh * Lh * m1pow(rank / 2) * CX[rqp + rai * rl]) /
                    sqrt(rank + 1.0);
    }
  } else {
    // X = w1 (c over holes), Y = w2 (c over particles); pm = Lh here
    for (int ck = lane; ck < B.N; ck += 64) {
      const bool hole = ck < B.A;
      const int c = hole ? B.holes[ck] : B.parts[ck - B.A];
      const double* w = hole ? X : Y;
      const int jc = B.jj[c];
      double s = 0.0;
      for (int J1 = abs(jc - jp); J1 <= jc + jp; J1 += 2) {
        for (int J2 = abs(jc - jq); J2 <= min(J1 - 2, jc + jq); J2 += 2)
          s = s - sqrt((J1 + 1.0) * (J2 + 1.0)) * x6(T, J1, J2, jq, jp, jc) *
                      tensor_elem(B, T, w, 1, c, qk, c, pk, J2, J1) * m1pow(J2 / 2) * Lh;
        for (int J2 = max(J1, abs(jc - jq)); J2 <= jc + jq; J2 += 2)
          s = s + sqrt((J1 + 1.0) * (J2 + 1.0)) * x6(T, J1, J2, jq, jp, jc) *
                      tensor_elem(B, T, w, 1, c, pk, c, qk, J1, J2) * m1pow(J1 / 2);
      }
      sm = sm + s * m1pow((jc + jq) / 2);
    }
  }
  sm = wave_sum(sm);
  if (MODE == 2) sm = sm * m1pow(rank / 2);
  if (lane != 0) return;
  if (t.cls == 2) {
    RES[B.off_fph + t.p + (i64)t.q * B.P] += sm;
    return;
  }
  const i64 n = t.cls == 0 ? B.A : B.P;
  double* F = RES + (t.cls == 0 ? B.off_fhh : B.off_fpp);
  const double x = F[t.p + t.q * n] + sm;
  F[t.p + t.q * n] = x;
  F[t.q + t.p * n] = x * RESh * m1pow((jp - jq) / 2);
}

// the back-transform of TS_commutator_222_ph (TS_commutators.f90:1532-1843)
// as a gather: the reference adds each product element Wx / Wy(IX = (d,a),
// JX = (c,b)) of every cross-coupled block to two stored elements; here the
// output element (P Q | Rr S; J1 J2) collects the sixteen (case, target)
// forms it can be, with the same coefficients and the same cut |W| > 1e-6
__device__ double ph_gather(const DevBasis& B, const DevTensor& T, const double* Wx, const double* Wy, int Lh, int Rh,
                            int P, int Q, int Rr, int S, int J1, int J2) {
  if ((P == Q && ((J1 >> 1) & 1)) || (Rr == S && ((J2 >> 1) & 1))) return 0.0;
  const int R = T.rank;
  const double thr = (double)1e-6f;
  const int nR = R / 2 + 1;
  double sm = 0.0;
  for (int pat = 0; pat < 16; ++pat) {
    int a, b, c, d;
    bool ok;
    switch (pat) {
      case 0: a = P; b = Q; c = Rr; d = S; ok = true; break;                  // b>=a d>=c: (a b c d)
      case 1: c = P; d = Q; a = Rr; b = S; ok = true; break;                  //            (c d a b)
      case 2: a = P; b = Q; d = Rr; c = S; ok = Rr < S; break;                // b>=a d<c:  (a b d c)
      case 3: d = P; c = Q; a = Rr; b = S; ok = P < Q; break;                 //            (d c a b)
      case 4: b = P; a = Q; c = Rr; d = S; ok = P < Q; break;                 // b<a d>=c:  (b a c d)
      case 5: c = P; d = Q; b = Rr; a = S; ok = Rr < S; break;                //            (c d b a)
      case 6: b = P; a = Q; d = Rr; c = S; ok = P < Q && Rr < S; break;       // b<a d<c:   (b a d c)
      case 7: d = P; c = Q; b = Rr; a = S; ok = P < Q && Rr < S; break;       //            (d c b a)
      case 8: d = P; b = Q; c = Rr; a = S; ok = true; break;                  // b>=d a>=c: (d b c a)
      case 9: c = P; a = Q; d = Rr; b = S; ok = true; break;                  //            (c a d b)
      case 10: d = P; b = Q; a = Rr; c = S; ok = Rr < S; break;               // b>=d a<c:  (d b a c)
      case 11: a = P; c = Q; d = Rr; b = S; ok = P < Q; break;                //            (a c d b)
      case 12: b = P; d = Q; c = Rr; a = S; ok = P < Q; break;                // b<d a>=c:  (b d c a)
      case 13: c = P; a = Q; b = Rr; d = S; ok = Rr < S; break;               //            (c a b d)
      case 14: b = P; d = Q; a = Rr; c = S; ok = P < Q && Rr < S; break;      // b<d a<c:   (b d a c)
      default: a = P; c = Q; b = Rr; d = S; ok = P < Q && Rr < S; break;      //            (a c b d)
    }
    if (!ok) continue;
    const int ja = B.jj[a], jb = B.jj[b], jc = B.jj[c], jd = B.jj[d];
    const int Tz = abs(B.tz[d] - B.tz[a]) >> 1;
    if ((abs(B.tz[c] - B.tz[b]) >> 1) != Tz) continue;
    const int PAR = (B.ll[d] + B.ll[a]) & 1;
    if (((B.ll[c] + B.ll[b]) & 1) != ((PAR + (T.dpar >> 1)) & 1)) continue;
    const double* W = pat < 8 ? Wx : Wy;
    const double phabcd = m1pow((ja + jb + jc + jd) / 2);
    // the J3 / J4-independent part of the pattern's coefficient; the 9j is
    // {u1 u2 J3; u4 u5 J4; J1 J2 rank} with (u1 u2) = (ja jd) or (jd ja)
    // and (u4 u5) = (jb jc) or (jc jb)
    double c0;
    bool use34;
    switch (pat) {
      case 0: c0 = m1pow((jb + jd + J2) / 2) * Lh; use34 = true; break;
      case 1: c0 = -m1pow((jc + ja + J1 + R) / 2) * Rh; use34 = false; break;
      case 2: c0 = m1pow((jb + jc) / 2) * Lh; use34 = true; break;
      case 3: c0 = m1pow((ja - jd + R) / 2) * Rh; use34 = false; break;
      case 4: c0 = m1pow((ja + jd + J1 + J2) / 2) * Lh; use34 = true; break;
      case 5: c0 = m1pow((jc - jb + J1 + J2 + R) / 2) * Rh; use34 = false; break;
      case 6: c0 = m1pow((ja + jc + J1) / 2) * Lh; use34 = true; break;
      case 7: c0 = m1pow((jb - jd + J2 + R) / 2) * Rh; use34 = false; break;
      case 8: c0 = m1pow((jc + ja + J2) / 2) * Lh; use34 = false; break;
      case 9: c0 = m1pow((jd - jb + J1 + R) / 2) * Rh; use34 = true; break;
      case 10: c0 = -Lh; use34 = false; break;
      case 11: c0 = phabcd * m1pow(R / 2) * Rh; use34 = true; break;
      case 12: c0 = -phabcd * m1pow((J1 + J2) / 2) * Lh; use34 = false; break;
      case 13: c0 = m1pow((J1 + J2 + R) / 2) * Rh; use34 = true; break;
      case 14: c0 = m1pow((jb + jd + J1) / 2) * Lh; use34 = false; break;
      default: c0 = m1pow((ja - jc + J2 + R) / 2) * Rh; use34 = true; break;
    }
    const bool t2 = pat & 1;                       // the (c d a b)-type target of a case
    const bool sw12 = (pat < 8) ? t2 : !t2;
    const int u1 = sw12 ? jd : ja, u2 = sw12 ? ja : jd;
    const int u4 = t2 ? jc : jb, u5 = t2 ? jb : jc;
    double sp = 0.0;
    for (int J3 = abs(jd - ja); J3 <= jd + ja; J3 += 2) {
      const int IX = B.rmap[((i64)d * B.N + a) * B.nJ + (J3 >> 1)];
      // Wy of a block with J3 = J4 is never formed (zero)
      const int J4lo = max(max(max(abs(J3 - R), J3), abs(jc - jb)), pat < 8 ? J3 : J3 + 2);
      const int J4hi = min(min(J3 + R, jc + jb), 2 * B.JTM);
      for (int J4 = J4lo; J4 <= J4hi; J4 += 2) {
        const i64* wt = T.wtab + ((((J3 >> 1) * nR + ((J4 - J3) >> 1)) << 2) + 2 * Tz + PAR) * 2;
        const int JX = B.rmap[((i64)c * B.N + b) * B.nJ + (J4 >> 1)];
        const double w = W[wt[0] + IX + JX * wt[1]];
        if (!(fabs(w) > thr)) continue;
        const double v = n9(T, u1, u2, J3, u4, u5, J4, J1, J2) * w;
        sp += (use34 && (((J3 + J4) >> 1) & 1)) ? -v : v;
      }
    }
    sm += c0 * sp;
  }
  sm *= sqrt((J1 + 1.0) * (J2 + 1.0));
  if (P == Q) sm *= 1.4142135623730951;
  if (Rr == S) sm *= 1.4142135623730951;
  return sm;
}

// TS_commutator_122 of one output element (TS_commutators.f90:555-663)
__device__ double ts122(const DevBasis& B, const DevTensor& T, const double* L, int Lh, const double* R, int Rh, int a,
                        int b, int c, int d, int J1, int J2) {
  double sm = 0.0;
  const int sa = B.spblk[a], sbk = B.spblk[b], sc = B.spblk[c], sd = B.spblk[d];
  for (int k = B.spstart[sa]; k < B.spstart[sa + 1]; ++k) {
    const int i = B.spmem[k];
    sm = sm + f_elem(B, L, Lh, a, i) * tensor_elem(B, T, R, Rh, i, b, c, d, J1, J2);
  }
  for (int k = B.spstart[sbk]; k < B.spstart[sbk + 1]; ++k) {
    const int i = B.spmem[k];
    sm = sm + f_elem(B, L, Lh, b, i) * tensor_elem(B, T, R, Rh, a, i, c, d, J1, J2);
  }
  for (int k = B.spstart[sc]; k < B.spstart[sc + 1]; ++k) {
    const int i = B.spmem[k];
    sm = sm - f_elem(B, L, Lh, i, c) * tensor_elem(B, T, R, Rh, a, b, i, d, J1, J2);
  }
  for (int k = B.spstart[sd]; k < B.spstart[sd + 1]; ++k) {
    const int i = B.spmem[k];
    sm = sm - f_elem(B, L, Lh, i, d) * tensor_elem(B, T, R, Rh, a, b, c, i, J1, J2);
  }
  sm = sm / sqrt(1.0 + (a == b)) / sqrt(1.0 + (c == d));
  return sm;
}

// TS_commutator_212 of one output element (TS_commutators.f90:666-806)
__device__ double ts212(const DevBasis& B, const DevTensor& T, const double* L, int Lh, const double* R, int Rh, int a,
                        int b, int c, int d, int J1, int J2) {
  const int rank = T.rank;
  const int ja = B.jj[a], jb = B.jj[b], jc = B.jj[c], jd = B.jj[d];
  const int ta = B.tz[a], tb = B.tz[b], tc = B.tz[c], td = B.tz[d];
  const int mla = B.ll[a] & 1, mlb = B.ll[b] & 1, mlc = B.ll[c] & 1, mld = B.ll[d] & 1;
  double sm = 0.0;
  for (int i = 0; i < B.N; ++i) {
    const int ji = B.jj[i], ti = B.tz[i], mli = (B.ll[i] + (T.dpar >> 1)) & 1;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0, s4 = 0.0;
    if (ti == ta && mli == mla && dtri(ji, ja, rank))
      s1 = -x6(T, J1, J2, ji, ja, jb) * f_tensor_elem(B, R, Rh, a, i) * v_elem(B, L, Lh, i, b, c, d, J2);
    s1 = s1 * m1pow((ja + jb - J2) / 2);
    if (ti == tb && mli == mlb && dtri(ji, jb, rank))
      s2 = x6(T, J1, J2, ji, jb, ja) * f_tensor_elem(B, R, Rh, b, i) * v_elem(B, L, Lh, i, a, c, d, J2);
    s2 = s2 * m1pow((J1 + J2) / 2);
    if (ti == td && mli == mld && dtri(ji, jd, rank))
      s3 = x6(T, J1, J2, jd, ji, jc) * f_tensor_elem(B, R, Rh, i, d) * v_elem(B, L, Lh, a, b, c, i, J1);
    s3 = s3 * m1pow((jc + jd - J1) / 2);
    if (ti == tc && mli == mlc && dtri(ji, jc, rank))
      s4 = -x6(T, J1, J2, jc, ji, jd) * f_tensor_elem(B, R, Rh, i, c) * v_elem(B, L, Lh, a, b, d, i, J1);
    s4 = s4 * m1pow((J2 + J1) / 2);
    sm = sm + (s1 + s2 + s3 + s4);
  }
  sm = sm * sqrt((J1 + 1.0) * (J2 + 1.0) / (1.0 + (a == b)) / (1.0 + (c == d))) * m1pow(rank / 2);
  return sm;
}

// TS_commutator_122 (assigns), one thread per stored two-body element in
// tiles of 256 consecutive elements of a sub-block
__global__ __launch_bounds__(256) void k_ts122(DevBasis B, DevTensor T, const TSub* subs, const TElem* tasks,
                                               const double* L, int Lh, const double* R, int Rh, double* RES) {
  const TElem t = tasks[blockIdx.x];
  const TSub sb = subs[t.sb];
  const int e = t.start + threadIdx.x;
  if (e >= sb.n1 * sb.n2) return;
  const int IX = e % sb.n1, JX = e / sb.n1;
  int a, b, c, d;
  unpack(B.pairs[sb.roff + IX], a, b);
  unpack(B.pairs[sb.coff + JX], c, d);
  const int J1 = T.blk[sb.q * 8], J2 = T.blk[sb.q * 8 + 1];
  RES[sb.off + IX + (i64)JX * sb.n1] = ts122(B, T, L, Lh, R, Rh, a, b, c, d, J1, J2);
}

// TS_commutator_212 (adds), one thread per element of the quadruple-ordered
// output list (neighbouring lanes share the orbit quadruple and so the orbit
// loop and its conditions)
__global__ __launch_bounds__(256) void k_ts212(DevBasis B, DevTensor T, const TPh* tasks, i64 nt, const double* L,
                                               int Lh, const double* R, int Rh, double* RES) {
  const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  const TPh t = tasks[e];
  const int u = t.abcd;
  RES[t.off] += ts212(B, T, L, Lh, R, Rh, u & 255, (u >> 8) & 255, (u >> 16) & 255, (u >> 24) & 255, t.jj & 0xffff,
                      t.jj >> 16);
}

// the 222_ph back-transform, one thread per output element of the
// quadruple-ordered list: neighbouring lanes share the orbit quadruple, so
// they run the same pattern and (J3, J4) loops and read the same W elements
// and one run of 9j symbols
__global__ __launch_bounds__(256) void k_ts_ph(DevBasis B, DevTensor T, const TPh* tasks, i64 nt, int Lh, int Rh,
                                               const double* Wx, const double* Wy, double* RES) {
  const i64 e = (i64)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt) return;
  const TPh t = tasks[e];
  const int u = t.abcd;
  RES[t.off] += ph_gather(B, T, Wx, Wy, Lh, Rh, u & 255, (u >> 8) & 255, (u >> 16) & 255, (u >> 24) & 255,
                          t.jj & 0xffff, t.jj >> 16);
}

// calculate_single_pandya, cross_coupled.f90:1416-1634, one thread per
// element of CCX(qx) (term 1: r1 x nb2) or CCR(qx) (term 2: nb1 x r2)
__global__ __launch_bounds__(256) void k_ts_pandya(DevBasis B, DevTensor T, const TCC* tasks, const double* R, int Rh,
                                                   double* CC) {
  const TCC t = tasks[blockIdx.x];
  const int qx = t.qx;
  const int J3c = T.cblk[qx * 8], J4c = T.cblk[qx * 8 + 1];
  const int r1 = T.cblk[qx * 8 + 4], r2 = T.cblk[qx * 8 + 5], nb1 = T.cblk[qx * 8 + 6], nb2 = T.cblk[qx * 8 + 7];
  const int s1 = T.cscal[qx * 2], s2 = T.cscal[qx * 2 + 1];
  const int e = t.start + threadIdx.x;
  const int dp = T.dpar >> 1;
  int a, b, p, h;
  if (t.term == 1) {
    if (e >= r1 * nb2) return;
    const int RI = e % r1, NB2 = e / r1;
    unpack(B.ccrows[B.ccrowoff[s1] + RI], a, b);
    unpack(B.cccols[B.cccoloff[s2] + NB2], p, h);
    double* out = CC + T.coff[qx * 3] + RI + (i64)NB2 * r1;
    const bool cons = (((B.ll[a] + B.ll[h]) & 1) == ((B.ll[b] + B.ll[p] + dp) & 1)) &&
                      (B.tz[a] + B.tz[h] == B.tz[b] + B.tz[p]);
    double sm = 0.0;
    if (cons) {
      const int ja = B.jj[a], jb = B.jj[b], jp = B.jj[p], jh = B.jj[h];
      for (int J3 = abs(ja - jh); J3 <= ja + jh; J3 += 2)
        for (int J4 = abs(jp - jb); J4 <= jp + jb; J4 += 2)
          sm = sm - m1pow(J4 / 2) * sqrt((J3 + 1.0) * (J4 + 1.0)) * ninej(T, ja, jb, J3c, jh, jp, J4c, J3, J4) *
                        tensor_elem(B, T, R, Rh, a, h, p, b, J3, J4);
      sm = sm * m1pow((jh + jb + J4c) / 2) * sqrt((J3c + 1.0) * (J4c + 1.0));
    }
    *out = sm;
  } else {
    if (e >= nb1 * r2) return;
    const int NB1 = e % nb1, GI = e / nb1;
    unpack(B.cccols[B.cccoloff[s1] + NB1], p, h);
    unpack(B.ccrows[B.ccrowoff[s2] + GI], b, a);
    double* out = CC + T.coff[qx * 3 + 1] + NB1 + (i64)GI * nb1;
    const bool cons = (((B.ll[a] + B.ll[h]) & 1) == ((B.ll[b] + B.ll[p] + dp) & 1)) &&
                      (B.tz[a] + B.tz[h] == B.tz[b] + B.tz[p]);
    double sm = 0.0;
    if (cons) {
      const int ja = B.jj[a], jb = B.jj[b], jp = B.jj[p], jh = B.jj[h];
      for (int J3 = abs(ja - jh); J3 <= ja + jh; J3 += 2)
        for (int J4 = abs(jp - jb); J4 <= jp + jb; J4 += 2)
          sm = sm - m1pow(J4 / 2) * sqrt((J3 + 1.0) * (J4 + 1.0)) * ninej(T, jh, jp, J3c, ja, jb, J4c, J3, J4) *
                        tensor_elem(B, T, R, Rh, h, a, b, p, J3, J4);
      sm = sm * m1pow((jp + ja + J4c) / 2) * sqrt((J3c + 1.0) * (J4c + 1.0)) * m1pow((jp + jh) / 2);
    }
    *out = sm;
  }
}

__global__ __launch_bounds__(256) void k_ts_epilogue(i64 off, i64 n, const double* w1, const double* w2,
                                                     double* RES) {
  for (i64 x = (i64)blockIdx.x * 256 + threadIdx.x; x < n; x += (i64)gridDim.x * 256)
    RES[off + x] = RES[off + x] + w1[off + x] - w2[off + x];
}

constexpr int kNormChunk = 4096;

__global__ __launch_bounds__(256) void k_wnorm_part(const i64* ck_off, const int* ck_len, const double* ck_w,
                                                    const double* op, double* part) {
  const int k = blockIdx.x;
  const i64 o = ck_off[k];
  const int n = ck_len[k];
  double s = 0.0;
  for (int x = threadIdx.x; x < n; x += 256) s += op[o + x] * op[o + x];
  __shared__ double red[4];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[k] = (red[0] + red[1] + red[2] + red[3]) * ck_w[k];
}

__global__ __launch_bounds__(256) void k_wnorm_final(const double* part, int n, double* dst) {
  __shared__ double red[256];
  double s = 0.0;
  for (int x = threadIdx.x; x < n; x += 256) s += part[x];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *dst = sqrt(red[0]);
}

int grid_of(i64 n) { return (int)std::min<i64>((n + 255) / 256, 65536); }

}  // namespace

void launch_ts111(const DevBasis& B, const DevTensor& T, const double* L, int Lh, const double* R, int Rh,
                  double* RES, hipStream_t s) {
  const i64 n = (i64)B.A * B.A + (i64)B.P * B.A + (i64)B.P * B.P;
  if (n) hipLaunchKernelGGL(k_ts111, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, B, T, L, Lh, R, Rh, RES);
}

void launch_ts_one(int mode, const DevBasis& B, const DevTensor& T, const TOne* t, int nt, const double* L, int Lh,
                   const double* X, const double* Y, const double* R, int Rh, double* RES, int RESh, hipStream_t s) {
  if (nt == 0) return;
  const dim3 g((nt + 3) / 4);
  if (mode == 0) hipLaunchKernelGGL(k_ts_one<0>, g, dim3(256), 0, s, B, T, t, nt, L, Lh, X, Y, R, Rh, RES, RESh);
  else if (mode == 1) hipLaunchKernelGGL(k_ts_one<1>, g, dim3(256), 0, s, B, T, t, nt, L, Lh, X, Y, R, Rh, RES, RESh);
  else hipLaunchKernelGGL(k_ts_one<2>, g, dim3(256), 0, s, B, T, t, nt, L, Lh, X, Y, R, Rh, RES, RESh);
}

void launch_ts122(const DevBasis& B, const DevTensor& T, const TSub* sb, const TElem* t, int nt, const double* L,
                  int Lh, const double* R, int Rh, double* RES, hipStream_t s) {
  if (nt > 0) hipLaunchKernelGGL(k_ts122, dim3(nt), dim3(256), 0, s, B, T, sb, t, L, Lh, R, Rh, RES);
}

void launch_ts212(const DevBasis& B, const DevTensor& T, const TPh* t, i64 nt, const double* L, int Lh,
                  const double* R, int Rh, double* RES, hipStream_t s) {
  if (nt > 0)
    hipLaunchKernelGGL(k_ts212, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, B, T, t, nt, L, Lh, R, Rh, RES);
}

void launch_ts_ph(const DevBasis& B, const DevTensor& T, const TPh* t, i64 nt, int Lh, int Rh, const double* Wx,
                  const double* Wy, double* RES, hipStream_t s) {
  if (nt > 0)
    hipLaunchKernelGGL(k_ts_ph, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, s, B, T, t, nt, Lh, Rh, Wx, Wy,
                       RES);
}

void launch_ts_pandya(const DevBasis& B, const DevTensor& T, const TCC* t, int nt, const double* R, int Rh,
                      double* CC, hipStream_t s) {
  if (nt) hipLaunchKernelGGL(k_ts_pandya, dim3(nt), dim3(256), 0, s, B, T, t, R, Rh, CC);
}

void launch_ts_epilogue(i64 off, i64 n, const double* w1, const double* w2, double* RES, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_ts_epilogue, dim3(grid_of(n)), dim3(256), 0, s, off, n, w1, w2, RES);
}

void launch_wnorm(const i64* ck_off, const int* ck_len, const double* ck_w, int nck, const double* op, double* part,
                  double* dst, hipStream_t s) {
  if (nck) hipLaunchKernelGGL(k_wnorm_part, dim3(nck), dim3(256), 0, s, ck_off, ck_len, ck_w, op, part);
  hipLaunchKernelGGL(k_wnorm_final, dim3(1), dim3(256), 0, s, part, nck, dst);
}

}  // namespace imsrg

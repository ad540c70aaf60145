// Host basis construction for the scalar-commutator path.  See basis.hpp
// for the reference routines each table follows.
#include "basis.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace imsrg {

namespace {

bool triangle(int a, int b, int c) { return c <= a + b && c >= std::abs(a - b); }

// (n)! in long double, n <= 170 is plenty for 2j <= 60
long double lfact(int n) {
  static std::vector<long double> t;
  if (t.empty()) {
    t.resize(200);
    t[0] = 1.0L;
    for (int i = 1; i < 200; ++i) t[i] = t[i - 1] * (long double)i;
  }
  return t[n];
}

// Racah's single-sum formula for the 6j symbol, evaluated with exact
// factorial ratios in extended precision (Edmonds, eq. 6.3.7).
long double tri_coef(int a, int b, int c) {   // Delta(abc), twice-j arguments
  return std::sqrt(lfact((a + b - c) / 2) * lfact((a - b + c) / 2) * lfact((-a + b + c) / 2) /
                   lfact((a + b + c) / 2 + 1));
}

}  // namespace

double sixj_value(int j1, int j2, int j3, int j4, int j5, int j6) {
  if (j1 < 0 || j2 < 0 || j3 < 0 || j4 < 0 || j5 < 0 || j6 < 0) return 0.0;
  if (!triangle(j1, j2, j3) || !triangle(j1, j5, j6) || !triangle(j4, j2, j6) ||
      !triangle(j4, j5, j3))
    return 0.0;
  if (((j1 + j2 + j3) | (j1 + j5 + j6) | (j4 + j2 + j6) | (j4 + j5 + j3)) & 1) return 0.0;
  const int a1 = (j1 + j2 + j3) / 2, a2 = (j1 + j5 + j6) / 2, a3 = (j4 + j2 + j6) / 2,
            a4 = (j4 + j5 + j3) / 2;
  const int b1 = (j1 + j2 + j4 + j5) / 2, b2 = (j2 + j3 + j5 + j6) / 2,
            b3 = (j3 + j1 + j6 + j4) / 2;
  const int tmin = std::max(std::max(a1, a2), std::max(a3, a4));
  const int tmax = std::min(b1, std::min(b2, b3));
  long double s = 0.0L;
  for (int t = tmin; t <= tmax; ++t) {
    long double term = lfact(t + 1) / (lfact(t - a1) * lfact(t - a2) * lfact(t - a3) *
                                       lfact(t - a4) * lfact(b1 - t) * lfact(b2 - t) *
                                       lfact(b3 - t));
    s += (t & 1) ? -term : term;
  }
  s *= tri_coef(j1, j2, j3) * tri_coef(j1, j5, j6) * tri_coef(j4, j2, j6) * tri_coef(j4, j5, j3);
  return (double)s;
}

std::string build_basis(Basis& b, int N, const int* n, const int* l, const int* jj,
                        const int* tz, const int* hole) {
  if (N <= 0 || N > 4096) return "orbit count out of range";
  b = Basis();
  b.N = N;
  b.nn.assign(n, n + N);
  b.ll.assign(l, l + N);
  b.jj.assign(jj, jj + N);
  b.tz.assign(tz, tz + N);
  b.con.assign(hole, hole + N);
  for (int i = 0; i < N; ++i) {
    if (jj[i] <= 0 || (jj[i] & 1) == 0) return "2j must be a positive odd integer";
    if (tz[i] != 1 && tz[i] != -1) return "2tz must be +-1";
    if (hole[i] != 0 && hole[i] != 1) return "hole flag must be 0 or 1";
    if (l[i] < 0) return "negative l";
  }
  // hole / particle indices (find_holes, basic_IMSRG.f90:420-437)
  b.cidx.resize(N);
  for (int i = 0; i < N; ++i) {
    if (hole[i]) { b.cidx[i] = (int)b.holes.size(); b.holes.push_back(i); }
    else { b.cidx[i] = (int)b.parts.size(); b.parts.push_back(i); }
  }
  b.A = (int)b.holes.size();
  b.P = (int)b.parts.size();
  b.JTM = *std::max_element(jj, jj + N);
  b.lmax = *std::max_element(l, l + N);
  b.nJ = b.JTM + 1;
  const int nJ = b.nJ;

  // single-particle blocks in the order of read_sp_basis (tz, l, j)
  b.spblk.assign(N, -1);
  b.spstart.push_back(0);
  for (int t = -1; t <= 1; t += 2)
    for (int li = 0; li <= b.lmax; ++li)
      for (int ji = 1; ji <= b.JTM; ji += 2) {
        for (int i = 0; i < N; ++i)
          if (jj[i] == ji && l[i] == li && tz[i] == t) {
            b.spblk[i] = (int)b.spstart.size() - 1;
            b.spmem.push_back(i);
          }
        b.spstart.push_back((int)b.spmem.size());
      }

  // two-body channels, q = 3*(2J) + 2*(Tz+1) + P
  b.nblocks = (b.JTM + 1) * 6;
  b.chJ.resize(b.nblocks); b.chT.resize(b.nblocks); b.chP.resize(b.nblocks);
  b.dims.assign(b.nblocks * 3, 0);
  b.pl.assign(b.nblocks * 3, {});
  b.xmap.assign((size_t)N * N * nJ, -1);
  for (int q = 0; q < b.nblocks; ++q) {
    const int J = 2 * (q / 6), T = (q % 6) / 2 - 1, Pr = q % 2;
    b.chJ[q] = J; b.chT[q] = T; b.chP[q] = Pr;
    for (int i = 0; i < N; ++i)
      for (int j = i; j < N; ++j) {
        if ((tz[i] + tz[j]) / 2 != T) continue;
        if ((l[i] + l[j]) % 2 != Pr) continue;
        if (!triangle(jj[i], jj[j], J)) continue;
        const int c = hole[i] + hole[j];
        const int idx = b.dims[q * 3 + c]++;
        b.xmap[((size_t)i * N + j) * nJ + J / 2] = idx;
        b.xmap[((size_t)j * N + i) * nJ + J / 2] = idx;
        b.pl[q * 3 + c].push_back(i);
        b.pl[q * 3 + c].push_back(j);
      }
  }
  b.pairoff.resize(b.nblocks * 3);
  for (int x = 0; x < b.nblocks * 3; ++x) {
    b.pairoff[x] = (i64)b.pairs.size();
    for (size_t k = 0; k < b.pl[x].size(); k += 2)
      b.pairs.push_back(b.pl[x][k] | (b.pl[x][k + 1] << 16));
  }
  // flat operator layout: E0, fhh, fph, fpp, then gam1..6 per channel
  b.off_fhh = 1;
  b.off_fph = b.off_fhh + (i64)b.A * b.A;
  b.off_fpp = b.off_fph + (i64)b.P * b.A;
  i64 off = b.off_fpp + (i64)b.P * b.P;
  b.goff.resize(b.nblocks * 6);
  for (int q = 0; q < b.nblocks; ++q)
    for (int g = 0; g < 6; ++g) {
      b.goff[q * 6 + g] = off;
      off += b.blocksize(q, g);
    }
  b.oplen = off;

  // cross-coupled blocks, q1 = J + Tz*(JTM+1) + 2*P*(JTM+1)
  b.ncc = (b.JTM + 1) * 4;
  b.ccJ.resize(b.ncc); b.ccT.resize(b.ncc); b.ccP.resize(b.ncc);
  b.rlen.assign(b.ncc, 0); b.ncol.assign(b.ncc, 0);
  b.ccoff.resize(b.ncc); b.wcoff.resize(b.ncc);
  b.ccrowoff.resize(b.ncc); b.cccoloff.resize(b.ncc);
  b.rmap.assign((size_t)N * N * nJ, -1);
  b.nbmap.assign((size_t)N * N * nJ, -1);
  i64 co = 0, wo = 0;
  for (int q1 = 0; q1 < b.ncc; ++q1) {
    const int JC = 2 * (q1 % (b.JTM + 1));
    const int TZ = (q1 % (2 * b.JTM + 2)) / (b.JTM + 1);
    const int PAR = q1 / (2 * b.JTM + 2);
    b.ccJ[q1] = JC; b.ccT[q1] = TZ; b.ccP[q1] = PAR;
    b.ccrowoff[q1] = (i64)b.ccrows.size();
    b.cccoloff[q1] = (i64)b.cccols.size();
    int r = 0, nc = 0;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        if (!triangle(jj[i], jj[j], JC)) continue;
        if ((l[i] + l[j]) % 2 != PAR) continue;
        if (std::abs(tz[i] - tz[j]) / 2 != TZ) continue;
        const size_t x = ((size_t)i * N + j) * nJ + JC / 2;
        if (hole[i] == 0 && hole[j] == 1) {
          b.nbmap[x] = nc++;
          b.cccols.push_back(i | (j << 16));
        }
        b.rmap[x] = r++;
        b.ccrows.push_back(i | (j << 16));
      }
    b.rlen[q1] = r; b.ncol[q1] = nc;
    b.ccoff[q1] = co; co += (i64)r * nc;
    b.wcoff[q1] = wo; wo += (i64)r * r;
  }
  b.cclen = co;
  b.wclen = wo;

  // 6j table indexed [j1/2][j2/2][j4/2][j5/2][J6/2][J3/2]: the symbols that
  // share their four half-integer arguments form one contiguous tile, and
  // for fixed J6 the channel 2J (= J3) is the fastest index, which is what
  // the lanes of one orbit quadruple read together.
  b.nh6 = (b.JTM + 1) / 2;
  b.nJ6 = b.JTM + 1;
  const int nh = b.nh6, nJ6 = b.nJ6;
  b.six.assign((size_t)nJ6 * nh * nh * nh * nh * nJ6, 0.0);
  for (int a1 = 0; a1 < nh; ++a1)
    for (int a2 = 0; a2 < nh; ++a2)
      for (int a4 = 0; a4 < nh; ++a4)
        for (int a5 = 0; a5 < nh; ++a5)
          for (int J6 = 0; J6 < nJ6; ++J6)
            for (int J3 = 0; J3 < nJ6; ++J3) {
              const size_t idx =
                  ((((((size_t)a1 * nh + a2) * nh + a4) * nh + a5) * nJ6 + J6) * nJ6) + J3;
              b.six[idx] = sixj_value(2 * a1 + 1, 2 * a2 + 1, 2 * J3, 2 * a4 + 1, 2 * a5 + 1, 2 * J6);
            }
  // cross-coupled row maps with 2J/2 slowest: rmapJ[(J*N + a)*N + b] and the
  // transposed rmapJT[(J*N + b)*N + a] (the inverse Pandya reads rows (k,i)
  // and (i,k) for neighbouring i)
  b.rmapJ.assign((size_t)N * N * nJ, -1);
  b.rmapJT.assign((size_t)N * N * nJ, -1);
  for (int x = 0; x < N; ++x)
    for (int y = 0; y < N; ++y)
      for (int jo = 0; jo < nJ; ++jo) {
        const int v = b.rmap[((size_t)x * N + y) * nJ + jo];
        b.rmapJ[((size_t)jo * N + x) * N + y] = v;
        b.rmapJT[((size_t)jo * N + y) * N + x] = v;
      }
  return "";
}

}  // namespace imsrg

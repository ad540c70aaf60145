// Harmonic-oscillator one-body part and me2j two-body matrix elements of a
// physical Hamiltonian, built on the host in the flat operator layout (the
// C-ABI uploads the result).  Arithmetic follows the reference statement by
// statement, including the single-precision square roots of the off-diagonal
// kinetic elements (default-real Fortran expressions) and the order of the
// two-body writes.
#include "interaction.hpp"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace imsrg {

namespace {

inline int kron(int i, int j) { return i == j ? 1 : 0; }
inline int m1pow(int e) { return (e & 1) ? -1 : 1; }   // (-1)**e, any sign

// one-body entry (i, j) of the flat layout, both orbits in the basis:
// fpp(P x P), fph (P x A: particle row, hole column), fhh (A x A)
void set_f(const Basis& b, double* op, int i, int j, double v) {
  const int ci = b.con[i], cj = b.con[j];
  if (ci + cj == 0) {
    op[b.off_fpp + b.cidx[i] + (i64)b.cidx[j] * b.P] = v;
    op[b.off_fpp + b.cidx[j] + (i64)b.cidx[i] * b.P] = v;
  } else if (ci + cj == 1) {
    if (cj > ci) op[b.off_fph + b.cidx[i] + (i64)b.cidx[j] * b.P] = v;
    else op[b.off_fph + b.cidx[j] + (i64)b.cidx[i] * b.P] = v;
  } else {
    op[b.off_fhh + b.cidx[i] + (i64)b.cidx[j] * b.A] = v;
    op[b.off_fhh + b.cidx[j] + (i64)b.cidx[i] * b.A] = v;
  }
}

// <a || r || c> in HO length units, basic_IMSRG.f90:4613-4657
double reduced_r(const Basis& b, int a, int c) {
  const int la = b.ll[a], lb = b.ll[c], na = b.nn[a], nb = b.nn[c];
  double c1 = 0.0, p1 = 0.0;
  if (la == lb + 1) {
    c1 = std::sqrt(lb + 1.0);
    if (na == nb) p1 = std::sqrt(nb + lb + 1.5);
    else if (na == nb - 1) p1 = -1 * std::sqrt(nb * 1.0);
  } else if (la == lb - 1) {
    c1 = -1 * std::sqrt(1.0 * lb);
    if (na == nb) p1 = std::sqrt(nb + lb + 0.5);
    else if (na == nb + 1) p1 = -1 * std::sqrt(nb + 1.0);
  }
  return p1 * c1;
}

// <a || p || c> in units of hbar / b, basic_IMSRG.f90:4567-4611
double reduced_p(const Basis& b, int a, int c) {
  const int la = b.ll[a], lb = b.ll[c], na = b.nn[a], nb = b.nn[c];
  double c1 = 0.0, p1 = 0.0;
  if (la == lb + 1) {
    c1 = std::sqrt(lb + 1.0);
    if (na == nb) p1 = -1 * std::sqrt(nb + lb + 1.5);
    else if (na == nb - 1) p1 = -1 * std::sqrt(nb * 1.0);
  } else if (la == lb - 1) {
    c1 = -1 * std::sqrt(1.0 * lb);
    if (na == nb) p1 = std::sqrt(nb + lb + 0.5);
    else if (na == nb + 1) p1 = std::sqrt(nb + 1.0);
  }
  return p1 * c1;
}

// the j-coupled one-body factor of a vector operator between orbits x, y
// (zero across proton / neutron)
template <class RED>
double jfac(const Basis& b, int x, int y, RED red) {
  if (b.tz[x] != b.tz[y]) return 0.0;
  const int jx = b.jj[x], jy = b.jj[y], lx = b.ll[x], ly = b.ll[y];
  return std::sqrt((jx + 1.0) * (jy + 1.0)) * (double)m1pow((3 + 2 * lx + jy) / 2) *
         sixj_value(jx, 2, jy, 2 * ly, 1, 2 * lx) * red(b, x, y);
}

// The two vector-product elements share everything but the reduced element
// and the signs of the (jb -+ jc - J) phases.
template <class RED>
double vdot(const Basis& b, int a, int bb, int c, int d, int J, RED red, int sgn) {
  const int ja = b.jj[a], jb = b.jj[bb], jc = b.jj[c], jd = b.jj[d];
  const double j_ac = jfac(b, a, c, red), j_ad = jfac(b, a, d, red);
  const double j_bc = jfac(b, bb, c, red), j_bd = jfac(b, bb, d, red);
  // (-1)**((jb + sgn*jc - J)/2) and (-1)**((jb + sgn*jd - J)/2); the
  // numerators are even, so the parity of the half is exact
  const int e1 = (jb + sgn * jc - J) / 2, e2 = (jc + jd - J) / 2, e3 = (jb + sgn * jd - J) / 2;
  const double t1 = (double)m1pow(e1) * sixj_value(ja, jb, J, jd, jc, 2) * j_ac * j_bd;
  const double t2 = (double)(m1pow(e2) * m1pow(e3)) * sixj_value(ja, jb, J, jc, jd, 2) * j_ad * j_bc;
  return (t1 - t2) / std::sqrt((1.0 + kron(a, bb)) * (1.0 + kron(c, d)));
}

}  // namespace

double r1_r2(const Basis& b, int a, int bb, int c, int d, int J) { return vdot(b, a, bb, c, d, J, reduced_r, 1); }
double p1_p2(const Basis& b, int a, int bb, int c, int d, int J) { return vdot(b, a, bb, c, d, J, reduced_p, -1); }

std::string h0_harm_osc(const Basis& b, const HamParams& p, double* op) {
  if (p.htype < 1 || p.htype > 5) return "htype must be 1..5";
  if (p.mass < 1 && (p.htype == 1 || p.htype == 4 || p.htype == 5)) return "mass must be positive";
  if (!(p.hw > 0.0)) return "hw must be positive";
  std::memset(op + b.off_fhh, 0, sizeof(double) * (size_t)(b.off_fpp + (i64)b.P * b.P - b.off_fhh));
  const double hw = p.hw, beta = p.lawson_beta, cmhw = p.com_hw;
  const int mass = p.mass;
  for (int i = 0; i < b.N; ++i) {
    const int ni = b.nn[i], li = b.ll[i], ji = b.jj[i], tzi = b.tz[i];
    for (int j = i; j < b.N; ++j) {
      if (b.ll[j] != li || b.jj[j] != ji || b.tz[j] != tzi) continue;
      const int nj = b.nn[j], lj = b.ll[j];
      if (std::abs(ni - nj) > 1) continue;
      // kij = 0.5*hw*(...): the off-diagonal square roots are taken of
      // default-real (single precision) arguments, as the reference does
      double kij;
      if (ni == nj) kij = 0.5 * hw * (double)((float)(2 * nj + lj) + 1.5f);
      else if (ni == nj - 1) kij = 0.5 * hw * (double)std::sqrt((float)nj * ((float)(nj + lj) + 0.5f));
      else kij = 0.5 * hw * (double)std::sqrt((float)(nj + 1) * ((float)(nj + lj) + 1.5f));
      double T = 0.0;
      switch (p.htype) {
        case 1:
          if (std::fabs(beta) < (double)1e-5f) {
            T = kij * (1.0 - 1.0 / mass);
          } else {
            T = kij * (1.0 + (beta - 1.0) / mass) +
                beta * cmhw * cmhw / (hw * hw) * kij * (double)(kron(ni, nj) - kron(ni, nj + 1) - kron(ni, nj - 1)) /
                    mass;
          }
          break;
        case 2: T = 2 * kij * (double)kron(ni, nj); break;
        case 3: T = kij; break;
        case 4: T = kij / mass; break;
        default: T = kij * (double)(kron(ni, nj) - kron(ni, nj + 1) - kron(ni, nj - 1)) / mass; break;
      }
      set_f(b, op, i, j, T);
    }
  }
  return "";
}

std::string map_file_orbits(const Basis& b, int nfile, const int* fn, const int* fl, const int* fjj,
                            const int* ftz, FileOrbits& f) {
  if (nfile <= 0 || (nfile & 1)) return "the file orbit count must be positive and even";
  f.n.assign(fn, fn + nfile); f.l.assign(fl, fl + nfile); f.jj.assign(fjj, fjj + nfile);
  f.tz.assign(ftz, ftz + nfile);
  f.con.assign(nfile, -1);
  for (int x = 0; x < nfile; x += 2) {
    if (f.tz[x] != -1 || f.tz[x + 1] != 1 || f.n[x] != f.n[x + 1] || f.l[x] != f.l[x + 1] ||
        f.jj[x] != f.jj[x + 1])
      return "file orbits must come in (proton, neutron) pairs of one (n, l, j)";
    if (f.n[x] < 0 || f.l[x] < 0 || (f.jj[x] != 2 * f.l[x] - 1 && f.jj[x] != 2 * f.l[x] + 1))
      return "file orbit " + std::to_string(x) + " is not an (n, l, j) orbit";
  }
  std::vector<int> seen(b.N, 0);
  for (int x = 0; x < nfile; ++x)
    for (int i = 0; i < b.N; ++i)
      if (b.nn[i] == f.n[x] && b.ll[i] == f.l[x] && b.jj[i] == f.jj[x] && b.tz[i] == f.tz[x]) {
        if (f.con[x] >= 0) return "two basis orbits match file orbit " + std::to_string(x);
        f.con[x] = i;
        ++seen[i];
      }
  for (int x = 0; x < nfile; x += 2)
    if ((f.con[x] < 0) != (f.con[x + 1] < 0))
      return "the basis holds only one of the proton / neutron orbits of file orbit " + std::to_string(x);
  for (int i = 0; i < b.N; ++i)
    if (seen[i] > 1) return "file orbits repeat basis orbit " + std::to_string(i);
  return "";
}

// The loop nest over isospin-coupled quadruples of read_me2j_interaction
// (interaction_IO.f90:318-350 counting, 396-430 conversion): orbits k1 >=
// k2, kk1 <= k1, kk2 <= kk1 (kk2 <= k2 when kk1 == k1), equal parity, then
// every common 2J.  `body(k1, k2, kk1, kk2, J)` per (quadruple, 2J).
template <class F>
void me2j_loop(const FileOrbits& f, F&& body) {
  const int nt = (int)f.n.size() / 2;
  for (int k1 = 0; k1 < nt; ++k1) {
    const int l1 = f.l[2 * k1], j1 = f.jj[2 * k1];
    for (int k2 = 0; k2 <= k1; ++k2) {
      const int l2 = f.l[2 * k2], j2 = f.jj[2 * k2];
      for (int kk1 = 0; kk1 <= k1; ++kk1) {
        const int ll1 = f.l[2 * kk1], jj1 = f.jj[2 * kk1];
        const int endpoint = (k1 == kk1) ? k2 : kk1;
        for (int kk2 = 0; kk2 <= endpoint; ++kk2) {
          const int ll2 = f.l[2 * kk2], jj2 = f.jj[2 * kk2];
          if ((l1 + l2) % 2 != (ll1 + ll2) % 2) continue;
          const int jmin = std::max(std::abs(j1 - j2), std::abs(jj1 - jj2));
          const int jmax = std::min(j1 + j2, jj1 + jj2);
          for (int JT = jmin; JT <= jmax; JT += 2) body(k1, k2, kk1, kk2, JT);
        }
      }
    }
  }
}

long long me2j_count(const FileOrbits& f) {
  long long n = 0;
  me2j_loop(f, [&](int, int, int, int, int) { n += 4; });
  return n;
}

namespace {

// one Fortran fixed-width real field (Fw.7): blanks read as zero, a field
// without a decimal point carries 7 implied decimals
bool parse_field(const char* s, int w, double& v) {
  char buf[64];
  int k = 0;
  for (int x = 0; x < w && s[x] && s[x] != '\n' && s[x] != '\r'; ++x)
    if (s[x] != ' ' && s[x] != '\t') {
      if (k >= 63) return false;
      buf[k++] = s[x];
    }
  buf[k] = 0;
  if (k == 0) { v = 0.0; return true; }
  char* end = nullptr;
  v = std::strtod(buf, &end);
  if (end != buf + k) return false;
  if (!std::strpbrk(buf, ".eEdD")) v /= 1e7;
  return true;
}

}  // namespace

std::string me2j_read_values(const char* path, long long count, std::vector<double>& out) {
  gzFile g = gzopen(path, "rb");
  if (!g) return std::string("cannot open ") + path;
  std::vector<char> line(1 << 16);
  out.assign((size_t)count, 0.0);
  std::string err;
  if (!gzgets(g, line.data(), (int)line.size())) err = std::string(path) + ": no header line";
  for (long long i = 0; err.empty() && i < count; i += 10) {
    if (!gzgets(g, line.data(), (int)line.size())) {
      err = std::string(path) + ": ends after " + std::to_string(i) + " of " + std::to_string(count) + " values";
      break;
    }
    const size_t len = std::strlen(line.data());
    const int nv = (int)std::min<long long>(10, count - i);
    for (int x = 0; x < nv; ++x) {
      const size_t at = x == 0 ? 0 : 12 + 13 * (size_t)(x - 1);
      const int w = x == 0 ? 12 : 13;
      double v = 0.0;
      if (at < len && !parse_field(line.data() + at, w, v)) {
        err = std::string(path) + ": bad number on data line " + std::to_string(i / 10 + 1);
        break;
      }
      out[(size_t)(i + x)] = v;
    }
  }
  gzclose(g);
  return err;
}

std::string me2j_to_op(const Basis& b, const FileOrbits& f, const std::vector<double>& me, const HamParams& p,
                       double* op) {
  if ((long long)me.size() != me2j_count(f)) return "value count does not match the file orbits";
  if (p.mass < 1) return "mass must be positive";
  if (!(p.hw > 0.0)) return "hw must be positive";
  const int COM = p.htype == 1 ? 1 : 0;
  const double hw = p.hw, beta = p.lawson_beta, cmhw = p.com_hw;
  const int N = b.N, nJ = b.nJ;
  // r1.r2 and p1.p2 from per-orbit-pair factors and the basis's 6j table
  // (the values r1_r2 / p1_p2 compute, in the same operation order); each
  // is evaluated only where its coefficient is nonzero
  const bool need_r = beta != 0.0, need_p = beta != 0.0 || COM != 0;
  std::vector<double> jr((size_t)N * N, 0.0), jp((size_t)N * N, 0.0);
  for (int x = 0; x < N; ++x)
    for (int y = 0; y < N; ++y) {
      if (need_r) jr[(size_t)x * N + y] = jfac(b, x, y, reduced_r);
      if (need_p) jp[(size_t)x * N + y] = jfac(b, x, y, reduced_p);
    }
  const int nh = b.nh6, nJ6 = b.nJ6;
  auto six2 = [&](int j1, int j2, int J, int j4, int j5) {   // {j1 j2 J; j4 j5 1}
    return b.six[(((((size_t)(j1 >> 1) * nh + (j2 >> 1)) * nh + (j4 >> 1)) * nh + (j5 >> 1)) * nJ6 + 1) * nJ6 +
                 J / 2];
  };
  auto vdot_tab = [&](const std::vector<double>& t, int a, int bb, int c, int d, int J, int sgn) {
    const int ja = b.jj[a], jb = b.jj[bb], jc = b.jj[c], jd = b.jj[d];
    const double j_ac = t[(size_t)a * N + c], j_ad = t[(size_t)a * N + d];
    const double j_bc = t[(size_t)bb * N + c], j_bd = t[(size_t)bb * N + d];
    const int e1 = (jb + sgn * jc - J) / 2, e2 = (jc + jd - J) / 2, e3 = (jb + sgn * jd - J) / 2;
    const double t1 = (double)m1pow(e1) * six2(ja, jb, J, jd, jc) * j_ac * j_bd;
    const double t2 = (double)(m1pow(e2) * m1pow(e3)) * six2(ja, jb, J, jc, jd) * j_ad * j_bc;
    return (t1 - t2) / std::sqrt((1.0 + kron(a, bb)) * (1.0 + kron(c, d)));
  };
  long long i = 0;
  me2j_loop(f, [&](int k1, int k2, int kk1, int kk2, int JT) {
    const double* m = me.data() + i;
    i += 4;
    if (f.con[2 * k1] < 0 || f.con[2 * k2] < 0 || f.con[2 * kk1] < 0 || f.con[2 * kk2] < 0) return;
    const int PAR = (f.l[2 * k1] + f.l[2 * k2]) % 2;
    for (int ax = 2 * k1; ax <= 2 * k1 + 1; ++ax) {
      const int a = f.con[ax];
      for (int bx = 2 * k2; bx <= 2 * k2 + 1; ++bx) {
        const int bb = f.con[bx];
        for (int cx = 2 * kk1; cx <= 2 * kk1 + 1; ++cx) {
          const int c = f.con[cx];
          for (int dx = 2 * kk2; dx <= 2 * kk2 + 1; ++dx) {
            const int d = f.con[dx];
            double pre2 = 1.0;
            if (a == bb) pre2 = pre2 * std::sqrt(0.5);
            if (c == d) pre2 = pre2 * std::sqrt(0.5);
            // the file's isospin projection is the opposite sign
            const int ta = -b.tz[a], tb = -b.tz[bb], tc = -b.tz[c], td = -b.tz[d];
            if (tc + td != ta + tb) continue;
            const int T = -(ta + tb) / 2;
            const int q = 3 * JT + 2 * (T + 1) + PAR;
            double V = 0.125 * (ta - tb) * (tc - td) * m[0] +
                       (double)(kron(ta + tb, -2) * kron(tc + td, -2)) * m[1] +
                       (double)(kron(ta + tb, 2) * kron(tc + td, 2)) * m[3] +
                       0.125 * std::abs((ta - tb) * (tc - td)) * m[2];
            const double g2 = need_r ? vdot_tab(jr, a, bb, c, d, JT, 1) : 0.0;
            const double g3 = need_p ? vdot_tab(jp, a, bb, c, d, JT, -1) : 0.0;
            const double g1 = (g3 + cmhw * cmhw / (hw * hw) * g2) * beta;
            V = V * pre2 + (g1 - g3 * COM) * hw / p.mass;
            const int C1 = b.con[a] + b.con[bb] + 1, C2 = b.con[c] + b.con[d] + 1;
            int qx = C1 * C2;
            if (qx == 9) qx = 5;
            const int g = qx - 1;
            double pre = 1;
            int i1, i2;
            if (a > bb) {
              i1 = b.xmap[((size_t)bb * N + a) * nJ + JT / 2];
              pre = m1pow(1 + (b.jj[a] + b.jj[bb] - JT) / 2);
            } else {
              i1 = b.xmap[((size_t)a * N + bb) * nJ + JT / 2];
            }
            if (c > d) {
              i2 = b.xmap[((size_t)d * N + c) * nJ + JT / 2];
              pre = pre * m1pow(1 + (b.jj[c] + b.jj[d] - JT) / 2);
            } else {
              i2 = b.xmap[((size_t)c * N + d) * nJ + JT / 2];
            }
            double* X = op + b.goff[q * 6 + g];
            const i64 ld = b.rows(q, g);
            if (qx == 1 || qx == 5 || qx == 4) {
              X[i2 + (i64)i1 * ld] = V * pre;
              X[i1 + (i64)i2 * ld] = V * pre;
            } else if (C1 > C2) {
              X[i2 + (i64)i1 * ld] = V * pre;
            } else {
              X[i1 + (i64)i2 * ld] = V * pre;
            }
          }
        }
      }
    }
  });
  return "";
}

}  // namespace imsrg

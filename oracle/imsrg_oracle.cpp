// ============================================================================
// TEST INFRASTRUCTURE ONLY - NOT PART OF THE PRODUCT.
//
// CPU oracle for the IMSRG(2) scalar-commutator / BCH / Magnus hot path of
// n-parzuchowski/EOM-IMSRG.  This file restates, in plain C++ (OpenMP + an
// optional host BLAS dgemm), the reference algorithm so that the HIP product
// can be checked element by element on the same inputs.  Only tests/,
// __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
//
// Every routine cites the reference routine (file:line) it follows.  Layout
// conventions (shared with the product, fixed by the reference's storage):
//   * orbits 0-based; jj = 2j, tz = 2tz (+-1), con = 1 hole / 0 particle
//   * operator = one flat FP64 array:
//       [E0 | fhh (A x A) | fph (P x A) | fpp (P x P) | per channel q:
//        gam1 np x np, gam2 np x nb, gam3 np x nh, gam4 nb x nb, gam5 nh x nh,
//        gam6 nb x nh]      (all column-major, as Fortran stores them)
//   * channels in the reference order q = 3*(2J) + 2*(Tz+1) + P
//     (basic_IMSRG.f90:572-681, block_index basic_IMSRG.f90:1108-1114)
//   * cross-coupled blocks q = J + Tz*(JTM+1) + 2*P*(JTM+1)
//     (cross_coupled.f90:61-177), CCX(q) r x nb and WCC(q) r x r col-major.
//
// Pinning (what checks this restatement independently of itself):
//   * brute-force J-scheme commutator formulas restated from
//     brute_force_testing.f90:1765-2076 (orc_bf_*), which share no code path
//     with the production-order routines below except f_elem / v_elem;
//   * 6j / 9j closed forms, symmetries and orthogonality (tests);
//   * BCH closed forms (Omega = 0; one-body Omega = orbital rotation), the
//     Magnus identity dOmega = eta for Omega = 0, antisymmetry of [A,B];
//   * frob_norm against the Frobenius norm of the full (anti)hermitian
//     matrices assembled independently, the all-ones closed form and one
//     element per block class (weights 1 and 2);
//   * the magnus_decouple loop: each step recomposed from the pinned
//     primitives (Omega_1 = ds eta, Euler of the Magnus series, BCH from the
//     bare H, mbpt2 stop), planted mistakes rejected, and the exact energy
//     of a Hamiltonian without two-body part (sum of lowest eigenvalues);
//   * the generators (white, atan, atanxx) against closed-form Epstein-Nesbet
//     denominators, w2 as the composition of pinned parts and its
//     Moller-Plesset small-coupling limit, the three flows' common E0;
//   * vectorize / repackage against the reference's own neq count, the
//     visiting order read off an index-valued operator, and round trips;
//   * every transpose (imsrg_adjoint.inc) against central differences of
//     its forward routine, including the whole flow gradient;
//   * the Hamiltonian set-up (imsrg_interaction.inc): HO one-body elements
//     against quadrature of Laguerre radial functions, r1.r2 + p1.p2 through
//     the exact two-nucleon centre-of-mass spectrum, the me2j isospin to
//     proton-neutron conversion against the Clebsch-Gordan decomposition;
//   * Hartree-Fock (imsrg_hf.inc): the two-body basis transform against the
//     four-index sum over v_elem, the HF Hamiltonian against normal_order of
//     the rotated vacuum Hamiltonian, the Brillouin condition.
// The reference itself is unbuildable here (it needs an external BLAS/LAPACK
// library that the image lacks), so there is no oracle/_ref.
// ============================================================================
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <dlfcn.h>
#include <omp.h>

typedef long long i64;

namespace {

// ---------------------------------------------------------------- host BLAS
typedef void (*dgemm64_fn)(const char*, const char*, const i64*, const i64*,
                           const i64*, const double*, const double*,
                           const i64*, const double*, const i64*,
                           const double*, double*, const i64*, size_t, size_t);
dgemm64_fn g_dgemm = nullptr;

// C := alpha*op(A)*op(B) + beta*C   (column-major, BLAS dgemm semantics)
void gemm(char ta, char tb, i64 m, i64 n, i64 k, double al, const double* A,
          i64 lda, const double* B, i64 ldb, double be, double* C, i64 ldc) {
  if (m <= 0 || n <= 0) return;
  if (g_dgemm) {
    g_dgemm(&ta, &tb, &m, &n, &k, &al, A, &lda, B, &ldb, &be, C, &ldc, 1, 1);
    return;
  }
  const bool tA = (ta == 'T'), tB = (tb == 'T');
#pragma omp parallel for schedule(static) if (m * n * k > 100000)
  for (i64 j = 0; j < n; ++j)
    for (i64 i = 0; i < m; ++i) {
      double s = 0.0;
      for (i64 p = 0; p < k; ++p) {
        double a = tA ? A[p + i * lda] : A[i + p * lda];
        double b = tB ? B[j + p * ldb] : B[p + j * ldb];
        s += a * b;
      }
      C[i + j * ldc] = (be == 0.0) ? al * s : al * s + be * C[i + j * ldc];
    }
}

// ------------------------------------------------------------ angular momentum
// Log-factorial table and Racah's closed form for the 6j symbol (Edmonds
// eq. 6.3.7), the method of anglib.f:263-379 (dfact0, d6ji, d, wwwwww).
std::vector<double> g_lf;
void init_logfact() {
  if (!g_lf.empty()) return;
  g_lf.resize(1001);
  g_lf[0] = 0.0;
  for (int i = 1; i <= 1000; ++i) g_lf[i] = g_lf[i - 1] + std::log((double)i);
}
// arguments are twice the angular momenta
double tri_delta(int a, int b, int c) {
  return 0.5 * (g_lf[(a + b - c) / 2] + g_lf[(a - b + c) / 2] +
                g_lf[(-a + b + c) / 2] - g_lf[(a + b + c) / 2 + 1]);
}
bool tri(int a, int b, int c) { return c <= a + b && c >= std::abs(a - b); }
double racah6j(int j1, int j2, int j3, int l1, int l2, int l3) {
  if (j1 < 0 || j2 < 0 || j3 < 0 || l1 < 0 || l2 < 0 || l3 < 0) return 0.0;
  if (!tri(j2, j3, j1) || !tri(l1, l2, j3) || !tri(l1, l3, j2) ||
      !tri(l2, l3, j1))
    return 0.0;
  if (((j1 + j2 + j3) & 1) || ((j3 + l1 + l2) & 1) || ((j2 + l1 + l3) & 1) ||
      ((j1 + l2 + l3) & 1))
    return 0.0;
  double df = tri_delta(j1, j2, j3) + tri_delta(j1, l2, l3) +
              tri_delta(l1, j2, l3) + tri_delta(l1, l2, j3);
  int m1 = (j1 + j2 + j3) / 2, m2 = (j1 + l2 + l3) / 2, m3 = (l1 + j2 + l3) / 2,
      m4 = (l1 + l2 + j3) / 2;
  int m5 = (j1 + j2 + l1 + l2) / 2, m6 = (j2 + j3 + l2 + l3) / 2,
      m7 = (j3 + j1 + l3 + l1) / 2;
  int imin = std::max(std::max(m1, m2), std::max(m3, m4));
  int imax = std::min(m5, std::min(m6, m7));
  double w = 0.0;
  for (int i = imin; i <= imax; ++i) {
    double sg = (i & 1) ? -1.0 : 1.0;
    w += sg * std::exp(df + g_lf[i + 1] - g_lf[i - m1] - g_lf[i - m2] -
                       g_lf[i - m3] - g_lf[i - m4] - g_lf[m5 - i] -
                       g_lf[m6 - i] - g_lf[m7 - i]);
  }
  return w;
}
// Racah W and the 9j symbol as a sum over k (anglib.f:531-566, wcoeff/coef9)
double wcoef(int a, int b, int c, int d, int e, int f) {
  double w = racah6j(a, b, e, d, c, f);
  if (((a + b + c + d) / 2) & 1) w = -w;
  return w;
}
double ninej(int j1, int j2, int j3, int j4, int j5, int j6, int j7, int j8,
             int j9) {
  double s = 0.0;
  int kmin = std::max(std::abs(j1 - j9), std::max(std::abs(j4 - j8), std::abs(j2 - j6)));
  int kmax = std::min(j1 + j9, std::min(j4 + j8, j2 + j6));
  for (int k = kmin; k <= kmax; k += 2)
    s += (k + 1) * wcoef(j1, j9, j4, j8, k, j7) * wcoef(j2, j6, j8, j4, k, j5) *
         wcoef(j1, j9, j2, j6, k, j3);
  return s;
}

inline int ipow_m1(int e) { return (e & 1) ? -1 : 1; }   // (-1)**e, any sign

// ------------------------------------------------------------------- basis
const int SEA1[6] = {0, 0, 0, 1, 2, 1};   // pair class of the rows (0 pp,1 ph,2 hh)
const int SEA2[6] = {0, 1, 2, 1, 2, 2};   // pair class of the columns
const bool SQS[6] = {true, false, false, true, true, false};

struct Basis {
  int N = 0, A = 0, P = 0, JTM = 0, lmax = 0;
  std::vector<int> nn, ll, jj, tz, con, holes, parts, cidx;
  // two-body channels
  int nblocks = 0;
  std::vector<int> chJ, chT, chP;            // 2J, Tz, parity
  std::vector<int> dim;                      // [q*3 + c]: npp, nph, nhh
  std::vector<std::vector<int>> qn;          // [q*3 + c]: pairs (i,j) flattened
  std::vector<i64> goff;                     // [q*6 + g]
  i64 off_fhh = 0, off_fph = 0, off_fpp = 0, oplen = 0;
  std::vector<int> xmap;                     // [(i*N+j)*(JTM+1) + J/2], i<=j
  // single-particle blocks (same j,l,tz), basic_IMSRG.f90:340-380
  std::vector<std::vector<int>> states;
  std::vector<int> spblk;                    // orbit -> its block
  // cross-coupled layout, cross_coupled.f90:61-177
  int ncc = 0;
  std::vector<int> ccJ, ccT, ccP, rlen, ncol;
  std::vector<i64> ccoff, wcoff;
  i64 cclen = 0, wclen = 0;
  std::vector<int> rmap, nbmap;              // [(i*N+j)*(JTM+1) + JC/2]
  std::vector<int> ccpair;                   // ordered pairs (i, j) of the rows, block by block
  std::vector<i64> ccpoff;                   // [q] first row of block q in ccpair / 2
  // 6j cache over (j1,j2,j4,j5 half-integer; j3,j6 integer), 2j units
  int nh6 = 0, nJ6 = 0;
  std::vector<double> six;

  int rows(int q, int g) const { return dim[q * 3 + SEA1[g]]; }
  int cols(int q, int g) const { return dim[q * 3 + SEA2[g]]; }
  int block_index(int J, int T, int P_) const { return 3 * J + 2 * (T + 1) + P_; }
  double sixj(int j1, int j2, int j3, int j4, int j5, int j6) const {
    // sixj of basic_IMSRG.f90:2553-2599: j1,j2,j4,j5 half-integer, j3,j6 integer
    if (!(tri(j1, j2, j3) && tri(j4, j5, j3) && tri(j1, j5, j6) && tri(j4, j2, j6)))
      return 0.0;
    size_t idx = ((((size_t)(j1 / 2) * nh6 + j2 / 2) * nh6 + j4 / 2) * nh6 + j5 / 2);
    idx = (idx * nJ6 + j3 / 2) * nJ6 + j6 / 2;
    return six[idx];
  }
};

void build_basis(Basis& b, int N, const int* n, const int* l, const int* jj,
                 const int* tz, const int* con) {
  init_logfact();
  b.N = N;
  b.nn.assign(n, n + N); b.ll.assign(l, l + N); b.jj.assign(jj, jj + N);
  b.tz.assign(tz, tz + N); b.con.assign(con, con + N);
  b.cidx.resize(N);
  int h = 0, p = 0;
  for (int i = 0; i < N; ++i) {       // find_holes, basic_IMSRG.f90:420-437
    if (con[i] == 1) { b.cidx[i] = h++; b.holes.push_back(i); }
    else { b.cidx[i] = p++; b.parts.push_back(i); }
  }
  b.A = h; b.P = p;
  b.JTM = *std::max_element(jj, jj + N);
  b.lmax = *std::max_element(l, l + N);
  const int nJ = b.JTM + 1, JTM = b.JTM;

  // sp blocks, loop order of read_sp_basis (basic_IMSRG.f90:345-379)
  b.spblk.assign(N, -1);
  for (int t = -1; t <= 1; t += 2)
    for (int li = 0; li <= b.lmax; ++li)
      for (int ji = 1; ji <= JTM; ji += 2) {
        std::vector<int> z;
        for (int i = 0; i < N; ++i)
          if (jj[i] == ji && l[i] == li && tz[i] == t) { z.push_back(i); b.spblk[i] = (int)b.states.size(); }
        b.states.push_back(z);
      }

  // two-body channels, allocate_blocks (basic_IMSRG.f90:521-685)
  b.nblocks = (JTM + 1) * 6;
  b.chJ.resize(b.nblocks); b.chT.resize(b.nblocks); b.chP.resize(b.nblocks);
  b.dim.assign(b.nblocks * 3, 0);
  b.qn.assign(b.nblocks * 3, {});
  b.xmap.assign((size_t)N * N * nJ, -1);
  b.off_fhh = 1;
  b.off_fph = b.off_fhh + (i64)b.A * b.A;
  b.off_fpp = b.off_fph + (i64)b.P * b.A;
  i64 off = b.off_fpp + (i64)b.P * b.P;
  int q = 0;
  for (int J = 0; J <= 2 * JTM; J += 2)
    for (int T = -1; T <= 1; ++T)
      for (int Pr = 0; Pr <= 1; ++Pr) {
        b.chJ[q] = J; b.chT[q] = T; b.chP[q] = Pr;
        for (int i = 0; i < N; ++i)
          for (int j = i; j < N; ++j) {
            if (T != (tz[i] + tz[j]) / 2) continue;
            if (Pr != (l[i] + l[j]) % 2) continue;
            if (!tri(jj[i], jj[j], J)) continue;
            int c = con[i] + con[j];           // 0 pp, 1 ph, 2 hh
            b.xmap[((size_t)i * N + j) * nJ + J / 2] = b.dim[q * 3 + c]++;
            b.qn[q * 3 + c].push_back(i);
            b.qn[q * 3 + c].push_back(j);
          }
        ++q;
      }
  b.goff.resize(b.nblocks * 6);
  for (q = 0; q < b.nblocks; ++q)
    for (int g = 0; g < 6; ++g) {
      b.goff[q * 6 + g] = off;
      off += (i64)b.rows(q, g) * b.cols(q, g);
    }
  b.oplen = off;

  // cross-coupled blocks, allocate_CCMAT (cross_coupled.f90:61-177)
  b.ncc = (JTM + 1) * 4;
  b.ccJ.resize(b.ncc); b.ccT.resize(b.ncc); b.ccP.resize(b.ncc);
  b.rlen.assign(b.ncc, 0); b.ncol.assign(b.ncc, 0);
  b.rmap.assign((size_t)N * N * nJ, -1);
  b.nbmap.assign((size_t)N * N * nJ, -1);
  b.ccoff.resize(b.ncc); b.wcoff.resize(b.ncc); b.ccpoff.resize(b.ncc);
  i64 co = 0, wo = 0;
  for (int q1 = 0; q1 < b.ncc; ++q1) {
    int JC = (q1 % (JTM + 1)) * 2;
    int PAR = q1 / (2 * JTM + 2);
    int TZ = (q1 % (2 * JTM + 2)) / (JTM + 1);
    b.ccJ[q1] = JC; b.ccT[q1] = TZ; b.ccP[q1] = PAR;
    b.ccpoff[q1] = (i64)b.ccpair.size() / 2;
    int r = 0, nbc = 0;
    for (int i = 0; i < N; ++i)
      for (int j = 0; j < N; ++j) {
        if (!tri(jj[i], jj[j], JC)) continue;
        if ((l[i] + l[j]) % 2 != PAR) continue;
        if (std::abs(tz[i] - tz[j]) / 2 != TZ) continue;
        size_t x = ((size_t)i * N + j) * nJ + JC / 2;
        if (con[i] == 0 && con[j] == 1) b.nbmap[x] = nbc++;
        b.rmap[x] = r++;
        b.ccpair.push_back(i);
        b.ccpair.push_back(j);
      }
    b.rlen[q1] = r; b.ncol[q1] = nbc;
    b.ccoff[q1] = co; co += (i64)r * nbc;
    b.wcoff[q1] = wo; wo += (i64)r * r;
  }
  b.cclen = co; b.wclen = wo;

  // 6j cache (store_6j, basic_IMSRG.f90:2227-2332): full table on the
  // half-integer / integer grid used by the hot path
  b.nh6 = (JTM + 1) / 2 + 1;
  b.nJ6 = JTM + 2;
  b.six.assign((size_t)b.nh6 * b.nh6 * b.nh6 * b.nh6 * b.nJ6 * b.nJ6, 0.0);
  for (int a1 = 1; a1 <= JTM; a1 += 2)
    for (int a2 = 1; a2 <= JTM; a2 += 2)
      for (int a4 = 1; a4 <= JTM; a4 += 2)
        for (int a5 = 1; a5 <= JTM; a5 += 2)
          for (int a3 = 0; a3 <= 2 * JTM; a3 += 2)
            for (int a6 = 0; a6 <= 2 * JTM; a6 += 2) {
              if (!(tri(a1, a2, a3) && tri(a4, a5, a3) && tri(a1, a5, a6) && tri(a4, a2, a6))) continue;
              size_t idx = ((((size_t)(a1 / 2) * b.nh6 + a2 / 2) * b.nh6 + a4 / 2) * b.nh6 + a5 / 2);
              idx = (idx * b.nJ6 + a3 / 2) * b.nJ6 + a6 / 2;
              b.six[idx] = racah6j(a1, a2, a3, a4, a5, a6);
            }
}

// ------------------------------------------------------------- accessors
struct Op {
  const Basis* b;
  double* d;
  int herm;
  double& E0() { return d[0]; }
  double* fhh() const { return d + b->off_fhh; }
  double* fph() const { return d + b->off_fph; }
  double* fpp() const { return d + b->off_fpp; }
  double* gam(int q, int g) const { return d + b->goff[q * 6 + g]; }
};

// f_elem, basic_IMSRG.f90:1191-1221
double f_elem(int a, int c, const Op& op) {
  const Basis& b = *op.b;
  int c1 = b.con[a], c2 = b.con[c];
  switch (c1 + c2) {
    case 0: return op.fpp()[b.cidx[a] + (i64)b.cidx[c] * b.P];
    case 1:
      if (c1 > c2) return op.fph()[b.cidx[c] + (i64)b.cidx[a] * b.P] * op.herm;
      return op.fph()[b.cidx[a] + (i64)b.cidx[c] * b.P];
    default: return op.fhh()[b.cidx[a] + (i64)b.cidx[c] * b.A];
  }
}

// Position of a two-body element, the lookup part of v_elem
// (basic_IMSRG.f90:1314-1429).  v_same (2051-2072) re-reads the same
// position on another operator; here both are expressed by returning it.
struct VPos {
  bool ok;
  int q, g;
  int i1, i2;     // row / column in gam g when !swap
  bool swap;      // element is X(i2,i1)*herm
  double pre;
};
VPos v_pos(const Basis& b, int a, int bb, int c, int d, int J) {
  VPos p; p.ok = false;
  if (!(tri(b.jj[a], b.jj[bb], J) && tri(b.jj[c], b.jj[d], J))) return p;
  int P_ = (b.ll[a] + b.ll[bb]) % 2;
  if ((b.ll[c] + b.ll[d]) % 2 != P_) return p;
  int T = (b.tz[a] + b.tz[bb]) / 2;
  if (b.tz[c] + b.tz[d] != 2 * T) return p;
  p.q = b.block_index(J, T, P_);
  int C1 = b.con[a] + b.con[bb] + 1, C2 = b.con[c] + b.con[d] + 1;
  int qx = C1 * C2; if (qx == 9) qx = 5;
  p.g = qx - 1;
  const int nJ = b.JTM + 1, N = b.N;
  double pre = 1.0;
  if (a > bb) {
    p.i1 = b.xmap[((size_t)bb * N + a) * nJ + J / 2];
    pre = ipow_m1(1 + (b.jj[a] + b.jj[bb] - J) / 2);
  } else {
    if (a == bb) pre *= std::sqrt(2.0);
    p.i1 = b.xmap[((size_t)a * N + bb) * nJ + J / 2];
  }
  if (c > d) {
    p.i2 = b.xmap[((size_t)d * N + c) * nJ + J / 2];
    pre *= ipow_m1(1 + (b.jj[c] + b.jj[d] - J) / 2);
  } else {
    if (c == d) pre *= std::sqrt(2.0);
    p.i2 = b.xmap[((size_t)c * N + d) * nJ + J / 2];
  }
  p.swap = (C1 > C2);
  p.pre = pre;
  p.ok = true;
  return p;
}
inline double v_at(const VPos& p, const Op& op) {
  if (!p.ok) return 0.0;
  const Basis& b = *op.b;
  const double* X = op.gam(p.q, p.g);
  if (p.swap) {
    int ld = b.rows(p.q, p.g);
    return X[p.i2 + (i64)p.i1 * ld] * op.herm * p.pre;
  }
  int ld = b.rows(p.q, p.g);
  return X[p.i1 + (i64)p.i2 * ld] * p.pre;
}
double v_elem(int a, int bb, int c, int d, int J, const Op& op) {
  return v_at(v_pos(*op.b, a, bb, c, d, J), op);
}

// ------------------------------------------------------------- commutators
// commutator_110, commutators.f90:10-33
double comm110(const Op& L, const Op& R) {
  const Basis& b = *L.b;
  double sm = 0.0;
  for (int i = 0; i < b.A; ++i) {
    int ji = b.jj[b.holes[i]];
    for (int a = 0; a < b.P; ++a) {
      double lf = L.fph()[a + (i64)i * b.P], rf = R.fph()[a + (i64)i * b.P];
      sm = sm + (lf * rf * L.herm - rf * lf * R.herm) * (ji + 1);
    }
  }
  return sm;
}

// commutator_220, commutators.f90:399-432
double comm220(const Op& L, const Op& R) {
  const Basis& b = *L.b;
  if (R.herm * L.herm != -1) return 0.0;
  double sm = 0.0;
  for (int q = 0; q < b.nblocks; ++q) {
    int np = b.dim[q * 3], nh = b.dim[q * 3 + 2];
    const double* L3 = L.gam(q, 2); const double* R3 = R.gam(q, 2);
    double smx = 0.0;
    for (int IX = 0; IX < np; ++IX)
      for (int JX = 0; JX < nh; ++JX)
        smx = smx - L3[IX + (i64)JX * np] * R3[IX + (i64)JX * np] * R.herm;
    sm = sm + smx * (b.chJ[q] + 1);
  }
  return sm * 2.0;
}

// commutator_111, commutators.f90:36-80 (assigns the one-body part of RES)
void comm111(const Op& L, const Op& R, Op& RES) {
  const Basis& b = *L.b;
  const i64 hol = b.A, par = b.P;
  std::vector<double> th1(hol * hol), th2(hol * hol), tb1(par * hol), tb2(par * hol),
      tp1(par * par), tp2(par * par);
  gemm('N', 'N', hol, hol, hol, 1.0, L.fhh(), hol, R.fhh(), hol, 0.0, th1.data(), hol);
  gemm('T', 'N', hol, hol, par, 1.0, L.fph(), par, R.fph(), par, 0.0, th2.data(), hol);
  if (par == 0) std::fill(th2.begin(), th2.end(), 0.0);
  for (i64 j = 0; j < hol; ++j)
    for (i64 i = 0; i < hol; ++i)
      RES.fhh()[i + j * hol] = th1[i + j * hol] + th2[i + j * hol] * L.herm -
                               (th1[j + i * hol] * L.herm + th2[j + i * hol]) * R.herm;
  gemm('N', 'N', par, hol, hol, 1.0, L.fph(), par, R.fhh(), hol, 0.0, tb1.data(), par);
  gemm('N', 'N', par, hol, par, 1.0, L.fpp(), par, R.fph(), par, 0.0, tb2.data(), par);
  for (i64 x = 0; x < par * hol; ++x) RES.fph()[x] = tb1[x] + tb2[x];
  gemm('N', 'N', par, hol, hol, 1.0, R.fph(), par, L.fhh(), hol, 0.0, tb1.data(), par);
  gemm('N', 'N', par, hol, par, 1.0, R.fpp(), par, L.fph(), par, 0.0, tb2.data(), par);
  for (i64 x = 0; x < par * hol; ++x) RES.fph()[x] = RES.fph()[x] - tb1[x] - tb2[x];
  gemm('N', 'T', par, par, hol, 1.0, L.fph(), par, R.fph(), par, 0.0, tp1.data(), par);
  gemm('N', 'N', par, par, par, 1.0, L.fpp(), par, R.fpp(), par, 0.0, tp2.data(), par);
  for (i64 j = 0; j < par; ++j)
    for (i64 i = 0; i < par; ++i)
      RES.fpp()[i + j * par] = tp1[i + j * par] * R.herm + tp2[i + j * par] -
                               (tp1[j + i * par] + tp2[j + i * par] * R.herm) * L.herm;
}

// inner sum of commutator_121 for one (p,q) (commutators.f90:127-154)
double c121_sum(const Op& L, const Op& R, int pk, int qk) {
  const Basis& b = *L.b;
  int jq = b.jj[qk];
  double sm = 0.0;
  for (int a = 0; a < b.P; ++a) {
    int ak = b.parts[a];
    for (int i = 0; i < b.A; ++i) {
      int ik = b.holes[i];
      if (b.jj[ik] != b.jj[ak] || b.ll[ik] != b.ll[ak] || b.tz[ik] != b.tz[ak]) continue;
      int ji = b.jj[ik];
      double smx = 0, smy = 0, smx2 = 0, smy2 = 0;
      for (int JT = std::abs(ji - jq); JT <= ji + jq; JT += 2) {
        VPos p1 = v_pos(b, ak, pk, ik, qk, JT);
        smx = smx + v_at(p1, R) * (JT + 1);
        smx2 = smx2 + v_at(p1, L) * (JT + 1);
        VPos p2 = v_pos(b, ik, pk, ak, qk, JT);
        smy = smy + v_at(p2, R) * (JT + 1);
        smy2 = smy2 + v_at(p2, L) * (JT + 1);
      }
      sm = sm + L.fph()[a + (i64)i * b.P] * (L.herm * smx - smy) -
           R.fph()[a + (i64)i * b.P] * (R.herm * smx2 - smy2);
    }
  }
  return sm;
}

// commutator_121, commutators.f90:83-280 (accumulates)
void comm121(const Op& L, const Op& R, Op& RES) {
  const Basis& b = *L.b;
  auto same = [&](int x, int y) {
    return b.jj[x] == b.jj[y] && b.ll[x] == b.ll[y] && b.tz[x] == b.tz[y];
  };
  for (int p = 0; p < b.A; ++p)
    for (int q = p; q < b.A; ++q) {
      int pk = b.holes[p], qk = b.holes[q];
      if (!same(pk, qk)) continue;
      double sm = c121_sum(L, R, pk, qk);
      double* F = RES.fhh();
      F[p + (i64)q * b.A] = F[p + (i64)q * b.A] + sm / (b.jj[pk] + 1.0);
      F[q + (i64)p * b.A] = F[p + (i64)q * b.A] * RES.herm;
    }
  for (int p = 0; p < b.P; ++p)
    for (int q = p; q < b.P; ++q) {
      int pk = b.parts[p], qk = b.parts[q];
      if (!same(pk, qk)) continue;
      double sm = c121_sum(L, R, pk, qk);
      double* F = RES.fpp();
      F[p + (i64)q * b.P] = F[p + (i64)q * b.P] + sm / (b.jj[pk] + 1.0);
      F[q + (i64)p * b.P] = F[p + (i64)q * b.P] * RES.herm;
    }
  for (int p = 0; p < b.P; ++p)
    for (int q = 0; q < b.A; ++q) {
      int pk = b.parts[p], qk = b.holes[q];
      if (!same(pk, qk)) continue;
      double sm = c121_sum(L, R, pk, qk);
      double* F = RES.fph();
      F[p + (i64)q * b.P] = F[p + (i64)q * b.P] + sm / (b.jj[pk] + 1.0);
    }
}

// commutator_122, commutators.f90:283-396 (assigns the two-body part)
void comm122(const Op& L, const Op& R, Op& RES) {
  const Basis& b = *L.b;
  for (int q = 0; q < b.nblocks; ++q) {
    int JT = b.chJ[q];
    for (int g = 0; g < 6; ++g) {
      int n1 = b.rows(q, g), n2 = b.cols(q, g);
      if (n1 * n2 == 0) continue;
      const std::vector<int>& P1 = b.qn[q * 3 + SEA1[g]];
      const std::vector<int>& P2 = b.qn[q * 3 + SEA2[g]];
      double* X = RES.gam(q, g);
      for (int IX = 0; IX < n1; ++IX) {
        int a = P1[2 * IX], bb = P1[2 * IX + 1];
        for (int JX = (SQS[g] ? IX : 0); JX < n2; ++JX) {
          int c = P2[2 * JX], d = P2[2 * JX + 1];
          double sm = 0.0;
          for (int i_sp : b.states[b.spblk[a]]) {
            VPos p = v_pos(b, i_sp, bb, c, d, JT);
            sm = sm + f_elem(a, i_sp, L) * v_at(p, R) - f_elem(a, i_sp, R) * v_at(p, L);
          }
          for (int i_sp : b.states[b.spblk[bb]]) {
            VPos p = v_pos(b, a, i_sp, c, d, JT);
            sm = sm + f_elem(bb, i_sp, L) * v_at(p, R) - f_elem(bb, i_sp, R) * v_at(p, L);
          }
          for (int i_sp : b.states[b.spblk[c]]) {
            VPos p = v_pos(b, a, bb, i_sp, d, JT);
            sm = sm - f_elem(i_sp, c, L) * v_at(p, R) + f_elem(i_sp, c, R) * v_at(p, L);
          }
          for (int i_sp : b.states[b.spblk[d]]) {
            VPos p = v_pos(b, a, bb, c, i_sp, JT);
            sm = sm - f_elem(i_sp, d, L) * v_at(p, R) + f_elem(i_sp, d, R) * v_at(p, L);
          }
          sm = sm / std::sqrt(1.0 + (a == bb)) / std::sqrt(1.0 + (c == d));
          X[IX + (i64)JX * n1] = sm;
          if (SQS[g]) X[JX + (i64)IX * n1] = sm * RES.herm;
        }
      }
    }
  }
}

// commutator_222_pp_hh, commutators.f90:581-768 (accumulates, fills w1, w2)
void comm222pphh(const Op& L, const Op& R, Op& RES, Op& w1, Op& w2) {
  const Basis& b = *L.b;
  const int pm = R.herm * L.herm;
  for (int q = 0; q < b.nblocks; ++q) {
    const i64 np = b.dim[q * 3], nb = b.dim[q * 3 + 1], nh = b.dim[q * 3 + 2];
    for (int g = 0; g < 6; ++g) {
      std::fill(w1.gam(q, g), w1.gam(q, g) + (i64)b.rows(q, g) * b.cols(q, g), 0.0);
      std::fill(w2.gam(q, g), w2.gam(q, g) + (i64)b.rows(q, g) * b.cols(q, g), 0.0);
    }
    double *L1 = L.gam(q, 0), *L2 = L.gam(q, 1), *L3 = L.gam(q, 2), *L5 = L.gam(q, 4), *L6 = L.gam(q, 5);
    double *R1 = R.gam(q, 0), *R2 = R.gam(q, 1), *R3 = R.gam(q, 2), *R5 = R.gam(q, 4), *R6 = R.gam(q, 5);
    double *W11 = w1.gam(q, 0), *W12 = w1.gam(q, 1), *W13 = w1.gam(q, 2), *W14 = w1.gam(q, 3),
           *W15 = w1.gam(q, 4), *W16 = w1.gam(q, 5);
    double *W21 = w2.gam(q, 0), *W22 = w2.gam(q, 1), *W23 = w2.gam(q, 2), *W24 = w2.gam(q, 3),
           *W25 = w2.gam(q, 4), *W26 = w2.gam(q, 5);
    if (np != 0) gemm('N', 'N', np, np, np, 1.0, L1, np, R1, np, 0.0, W11, np);
    if (nh != 0) gemm('N', 'N', nh, nh, nh, 1.0, L5, nh, R5, nh, 0.0, W25, nh);
    if (np * nh != 0) {
      gemm('T', 'N', nh, nh, np, 1.0, L3, np, R3, np, 0.0, W15, nh);
      for (i64 x = 0; x < nh * nh; ++x) W15[x] *= L.herm;
      gemm('N', 'T', np, np, nh, 1.0, L3, np, R3, np, 0.0, W21, np);
      for (i64 x = 0; x < np * np; ++x) W21[x] *= R.herm;
      gemm('N', 'N', np, nh, np, -1.0, R1, np, L3, np, 0.0, W13, np);
      gemm('N', 'N', np, nh, np, 1.0, L1, np, R3, np, 1.0, W13, np);
      gemm('N', 'N', np, nh, nh, 1.0, R3, np, L5, nh, 0.0, W23, np);
      gemm('N', 'N', np, nh, nh, 1.0, L3, np, R5, nh, -1.0, W23, np);
    }
    double* X1 = RES.gam(q, 0);
    for (i64 j = 0; j < np; ++j)
      for (i64 i = 0; i < np; ++i)
        X1[i + j * np] = X1[i + j * np] + W11[i + j * np] - W21[i + j * np] -
                         (W11[j + i * np] - W21[j + i * np]) * pm;
    double* X5 = RES.gam(q, 4);
    for (i64 j = 0; j < nh; ++j)
      for (i64 i = 0; i < nh; ++i)
        X5[i + j * nh] = X5[i + j * nh] + W15[i + j * nh] - W25[i + j * nh] -
                         (W15[j + i * nh] - W25[j + i * nh]) * pm;
    double* X3 = RES.gam(q, 2);
    for (i64 x = 0; x < np * nh; ++x) X3[x] = X3[x] + W13[x] - W23[x];
    if (nb * nh != 0) {
      gemm('N', 'T', nb, nb, nh, 1.0, L6, nb, R6, nb, 0.0, W24, nb);
      for (i64 x = 0; x < nb * nb; ++x) W24[x] *= R.herm;
    }
    if (nb * np != 0) {
      gemm('T', 'N', nb, nb, np, 1.0, L2, np, R2, np, 0.0, W14, nb);
      for (i64 x = 0; x < nb * nb; ++x) W14[x] *= L.herm;
    }
    double* X4 = RES.gam(q, 3);
    for (i64 j = 0; j < nb; ++j)
      for (i64 i = 0; i < nb; ++i)
        X4[i + j * nb] = X4[i + j * nb] + W14[i + j * nb] - W24[i + j * nb] -
                         (W14[j + i * nb] - W24[j + i * nb]) * pm;
    if (nb * nh != 0) {
      gemm('N', 'N', nb, nh, nh, 1.0, R6, nb, L5, nh, 0.0, W26, nb);
      gemm('N', 'N', nb, nh, nh, 1.0, L6, nb, R5, nh, -1.0, W26, nb);
    }
    if (nb * np != 0) {
      gemm('N', 'N', np, nb, np, 1.0, R1, np, L2, np, 0.0, W12, np);
      gemm('N', 'N', np, nb, np, 1.0, L1, np, R2, np, -1.0, W12, np);
    }
    if (nb * np * nh != 0) {
      gemm('T', 'N', nb, nh, np, -1.0 * R.herm, R2, np, L3, np, 0.0, W16, nb);
      gemm('T', 'N', nb, nh, np, (double)L.herm, L2, np, R3, np, 1.0, W16, nb);
      gemm('N', 'T', np, nb, nh, 1.0, R3, np, L6, nb, 0.0, W22, np);
      gemm('N', 'T', np, nb, nh, (double)R.herm, L3, np, R6, nb, -1.0 * L.herm, W22, np);
    }
    double* X6 = RES.gam(q, 5);
    for (i64 x = 0; x < nb * nh; ++x) X6[x] = X6[x] + W16[x] - W26[x];
    double* X2 = RES.gam(q, 1);
    for (i64 x = 0; x < np * nb; ++x) X2[x] = X2[x] + W12[x] - W22[x];
  }
}

// commutator_221, commutators.f90:435-578 (accumulates; after 222_pp_hh)
void comm221(const Op& L, const Op& R, Op& RES, const Op& w1, const Op& w2) {
  const Basis& b = *L.b;
  const int pm = R.herm * L.herm;
  auto sum_c = [&](int ik, int jk, bool both) {
    int ji = b.jj[ik];
    double sm = 0.0;
    for (int c = 0; c < b.A; ++c) {
      int ck = b.holes[c], jc = b.jj[ck];
      for (int JT = std::abs(jc - ji); JT <= jc + ji; JT += 2) {
        if (both)
          sm = sm + (v_elem(ck, ik, ck, jk, JT, w1) - pm * v_elem(ck, jk, ck, ik, JT, w1)) * (JT + 1);
        else
          sm = sm + v_elem(ck, ik, ck, jk, JT, w1) * (JT + 1);
      }
    }
    for (int c = 0; c < b.P; ++c) {
      int ck = b.parts[c], jc = b.jj[ck];
      for (int JT = std::abs(jc - ji); JT <= jc + ji; JT += 2) {
        if (both)
          sm = sm + (v_elem(ck, ik, ck, jk, JT, w2) - pm * v_elem(ck, jk, ck, ik, JT, w2)) * (JT + 1);
        else
          sm = sm + v_elem(ck, ik, ck, jk, JT, w2) * (JT + 1);
      }
    }
    return sm;
  };
  auto same = [&](int x, int y) {
    return b.jj[x] == b.jj[y] && b.ll[x] == b.ll[y] && b.tz[x] == b.tz[y];
  };
  for (int i = 0; i < b.A; ++i)
    for (int j = i; j < b.A; ++j) {
      int ik = b.holes[i], jk = b.holes[j];
      if (!same(ik, jk)) continue;
      double sm = sum_c(ik, jk, true);
      double* F = RES.fhh();
      F[i + (i64)j * b.A] = F[i + (i64)j * b.A] + sm / (b.jj[ik] + 1.0);
      F[j + (i64)i * b.A] = F[i + (i64)j * b.A] * RES.herm;
    }
  for (int i = 0; i < b.P; ++i)
    for (int j = i; j < b.P; ++j) {
      int ik = b.parts[i], jk = b.parts[j];
      if (!same(ik, jk)) continue;
      double sm = sum_c(ik, jk, true);
      double* F = RES.fpp();
      F[i + (i64)j * b.P] = F[i + (i64)j * b.P] + sm / (b.jj[ik] + 1.0);
      F[j + (i64)i * b.P] = F[i + (i64)j * b.P] * RES.herm;
    }
  for (int i = 0; i < b.P; ++i)
    for (int j = 0; j < b.A; ++j) {
      int ik = b.parts[i], jk = b.holes[j];
      if (!same(ik, jk)) continue;
      double sm = sum_c(ik, jk, false);
      double* F = RES.fph();
      F[i + (i64)j * b.P] = F[i + (i64)j * b.P] + sm / (b.jj[ik] + 1.0);
    }
}

// calculate_cross_coupled, cross_coupled.f90:1190-1300
void cross_coupled(const Op& H, double* CC) {
  const Basis& b = *H.b;
  const int N = b.N, nJ = b.JTM + 1;
#pragma omp parallel for schedule(dynamic)
  for (int q1 = 0; q1 < b.ncc; ++q1) {
    int JC = b.ccJ[q1], PAR = b.ccP[q1], TZ = b.ccT[q1];
    int r = b.rlen[q1];
    double* X = CC + b.ccoff[q1];
    std::fill(X, X + (i64)r * b.ncol[q1], 0.0);
    for (int hg = 0; hg < b.A; ++hg)
      for (int pg = 0; pg < b.P; ++pg) {
        int h = b.holes[hg], p = b.parts[pg];
        int jp = b.jj[p], jh = b.jj[h], lp = b.ll[p], lh = b.ll[h], tp = b.tz[p], th = b.tz[h];
        if (!tri(jp, jh, JC)) continue;
        if ((lp + lh) % 2 != PAR) continue;
        if (std::abs(tp - th) / 2 != TZ) continue;
        int NB = b.nbmap[((size_t)p * N + h) * nJ + JC / 2];
        for (int a = 0; a < N; ++a)
          for (int bb = 0; bb < N; ++bb) {
            int ja = b.jj[a], jb = b.jj[bb], la = b.ll[a], lb = b.ll[bb], ta = b.tz[a], tb = b.tz[bb];
            if (!tri(ja, jb, JC)) continue;
            if ((la + lb) % 2 != PAR) continue;
            if (std::abs(ta - tb) / 2 != TZ) continue;
            int RI = b.rmap[((size_t)a * N + bb) * nJ + JC / 2];
            if (((la + lh) % 2 == (lb + lp) % 2) && (ta + th == tb + tp)) {
              int Jmin = std::max(std::abs(jp - jb), std::abs(ja - jh));
              int Jmax = std::min(jp + jb, ja + jh);
              double sm = 0.0;
              for (int JT = Jmin; JT <= Jmax; JT += 2)
                sm = sm + ipow_m1(JT / 2) * (JT + 1) * b.sixj(jp, jh, JC, ja, jb, JT) *
                              v_elem(h, a, p, bb, JT, H);
              X[RI + (i64)NB * r] = sm * ipow_m1((jh + jb + JC) / 2) * 1.0 * std::sqrt(JC + 1.0);
            }
          }
      }
  }
}

// commutator_222_ph, commutators.f90:771-925 (accumulates)
void comm222ph(const double* LCC, int Lh, const double* RCC, int Rh, Op& RES, double* WCC) {
  const Basis& b = *RES.b;
  const int N = b.N, nJ = b.JTM + 1, JTM = b.JTM;
  for (int q = 0; q < b.ncc; ++q) {
    int nb = b.ncol[q], r = b.rlen[q];
    if (nb * r == 0) continue;
    gemm('N', 'T', r, r, nb, 1.0, LCC + b.ccoff[q], r, RCC + b.ccoff[q], r, 0.0, WCC + b.wcoff[q], r);
  }
  auto row = [&](int x, int y, int qx) { return b.rmap[((size_t)x * N + y) * nJ + b.ccJ[qx] / 2]; };
#pragma omp parallel for schedule(dynamic)
  for (int q = 0; q < b.nblocks; ++q) {
    int Jtot = b.chJ[q];
    for (int g = 0; g < 6; ++g) {
      int n1 = b.rows(q, g), n2 = b.cols(q, g);
      if (n1 * n2 == 0) continue;
      const std::vector<int>& P1 = b.qn[q * 3 + SEA1[g]];
      const std::vector<int>& P2 = b.qn[q * 3 + SEA2[g]];
      double* X = RES.gam(q, g);
      for (int IX = 0; IX < n1; ++IX) {
        double pre = 1.0;
        int i = P1[2 * IX], j = P1[2 * IX + 1];
        if (i == j) pre = .70710678118;
        int ji = b.jj[i], jj = b.jj[j], li = b.ll[i], lj = b.ll[j], ti = b.tz[i], tj = b.tz[j];
        for (int JX = (SQS[g] ? IX : 0); JX < n2; ++JX) {
          double pre2 = 1.0;
          int k = P2[2 * JX], l = P2[2 * JX + 1];
          if (k == l) pre2 = .70710678118;
          int jk = b.jj[k], jl = b.jj[l], lk = b.ll[k], ll = b.ll[l], tk = b.tz[k], tl = b.tz[l];
          double sm = 0.0;
          int jmin = std::max(std::abs(jj - jl), std::abs(ji - jk));
          int jmax = std::min(jj + jl, ji + jk);
          int Tz = std::abs(ti - tk) / 2;
          if (std::abs(tl - tj) != Tz * 2) continue;
          int PAR = (li + lk) % 2;
          if ((ll + lj) % 2 != PAR) continue;
          int phase = ipow_m1((ji + jj + jk + jl) / 2);
          for (int JP = jmin; JP <= jmax; JP += 2) {
            int qx = JP / 2 + Tz * (JTM + 1) + 2 * PAR * (JTM + 1);
            const double* W = WCC + b.wcoff[qx];
            i64 rr = b.rlen[qx];
            int rjl = row(j, l, qx), rik = row(i, k, qx), gjl = row(l, j, qx), gik = row(k, i, qx);
            sm = sm - ipow_m1((jl + jk + Jtot) / 2) *
                          ((W[gik + rjl * rr] + phase * W[gjl + rik * rr]) * Lh -
                           (phase * W[rik + gjl * rr] + W[rjl + gik * rr]) * Rh) *
                          b.sixj(ji, jj, Jtot, jl, jk, JP);
          }
          Tz = std::abs(ti - tl) / 2;
          if (std::abs(tk - tj) != Tz * 2) continue;
          PAR = (li + ll) % 2;
          if ((lk + lj) % 2 != PAR) continue;
          jmin = std::max(std::abs(ji - jl), std::abs(jj - jk));
          jmax = std::min(ji + jl, jj + jk);
          phase = ipow_m1((ji + jj + jk + jl) / 2);
          for (int JP = jmin; JP <= jmax; JP += 2) {
            int qx = JP / 2 + Tz * (JTM + 1) + 2 * PAR * (JTM + 1);
            const double* W = WCC + b.wcoff[qx];
            i64 rr = b.rlen[qx];
            int ril = row(i, l, qx), rjk = row(j, k, qx), gil = row(l, i, qx), gjk = row(k, j, qx);
            sm = sm + ((W[gil + rjk * rr] + phase * W[gjk + ril * rr]) * Lh -
                       (phase * W[ril + gjk * rr] + W[rjk + gil * rr]) * Rh) *
                          b.sixj(ji, jj, Jtot, jk, jl, JP);
          }
          X[IX + (i64)JX * n1] = X[IX + (i64)JX * n1] + sm * pre * pre2;
          if (SQS[g]) X[JX + (i64)IX * n1] = X[IX + (i64)JX * n1] * RES.herm;
        }
      }
    }
  }
}

// mat_frob_norm, basic_IMSRG.f90:4056-4087
double frob_norm(const Op& op) {
  const Basis& b = *op.b;
  double s1 = 0, s2 = 0, s3 = 0;
  for (i64 x = 0; x < (i64)b.A * b.A; ++x) s1 += op.fhh()[x] * op.fhh()[x];
  for (i64 x = 0; x < (i64)b.P * b.A; ++x) s2 += op.fph()[x] * op.fph()[x];
  for (i64 x = 0; x < (i64)b.P * b.P; ++x) s3 += op.fpp()[x] * op.fpp()[x];
  double sm = s1 + 2 * s2 + s3;
  for (int q = 0; q < b.nblocks; ++q)
    for (int g = 0; g < 6; ++g) {
      const double* X = op.gam(q, g);
      i64 n = (i64)b.rows(q, g) * b.cols(q, g);
      double s = 0.0;
      for (i64 x = 0; x < n; ++x) s += X[x] * X[x];
      sm = sm + s;
      if (!SQS[g]) sm = sm + s;
    }
  return std::sqrt(sm);
}

// add_sq_op, basic_IMSRG.f90:3551-3583: C = ax*A + bx*B over every element
void add_op(const Basis& b, const double* A, double ax, const double* B, double bx, double* C) {
  for (i64 x = 0; x < b.oplen; ++x) C[x] = A[x] * ax + B[x] * bx;
}

// twobody_monopole, basic_IMSRG.f90:2075-2090
double monopole(int a, int bb, const Op& H) {
  const Basis& b = *H.b;
  int ja = b.jj[a], jb = b.jj[bb];
  double sm = 0.0;
  for (int JT = std::abs(ja - jb); JT <= ja + jb; JT += 2) sm = sm + (JT + 1) * v_elem(a, bb, a, bb, JT, H);
  return sm / (ja + 1.0) / (jb + 1.0);
}

// Ground-state generators with the denominators of Hd and the numerators
// of Hn: build_gs_white (generators.f90:4-106; mode 0, eta = Hn / Eden),
// build_gs_atan (163-269; mode 1, eta = 0.5 atan(2 Hn / Eden), Hd = Hn),
// build_gs_atanxx (271-374; mode 1, Hd = Hold).  ETA herm = -1; E0 kept.
void gen_core(const Op& Hd, const Op& Hn, Op& ETA, int mode) {
  const Basis& b = *Hd.b;
  auto val = [mode](double x, double Eden) { return mode == 0 ? x / Eden : std::atan(2 * x / Eden) * 0.5; };
  std::fill(ETA.fpp(), ETA.fpp() + (i64)b.P * b.P, 0.0);
  std::fill(ETA.fhh(), ETA.fhh() + (i64)b.A * b.A, 0.0);
  for (int a = 0; a < b.P; ++a) {
    int ak = b.parts[a];
    for (int i = 0; i < b.A; ++i) {
      int ik = b.holes[i];
      if (b.jj[ik] != b.jj[ak] || b.ll[ik] != b.ll[ak] || b.tz[ik] != b.tz[ak]) continue;
      int ji = b.jj[ik];
      double Eden = 0.0;
      for (int JT = 0; JT <= 2 * ji; JT += 2) Eden = Eden - (JT + 1) * v_elem(ak, ik, ak, ik, JT, Hd);
      Eden = Eden / (ji + 1.0) / (ji + 1.0);
      Eden = Eden + Hd.fpp()[a + (i64)a * b.P] - Hd.fhh()[i + (i64)i * b.A];
      ETA.fph()[a + (i64)i * b.P] = val(Hn.fph()[a + (i64)i * b.P], Eden);
    }
  }
  for (int q = 0; q < b.nblocks; ++q) {
    for (int g = 0; g < 6; ++g)
      std::fill(ETA.gam(q, g), ETA.gam(q, g) + (i64)b.rows(q, g) * b.cols(q, g), 0.0);
    int np = b.dim[q * 3], nh = b.dim[q * 3 + 2];
    const std::vector<int>& Pp = b.qn[q * 3 + 0];
    const std::vector<int>& Ph = b.qn[q * 3 + 2];
    for (int IX = 0; IX < np; ++IX) {
      int a = Pp[2 * IX], bb = Pp[2 * IX + 1];
      for (int JX = 0; JX < nh; ++JX) {
        int i = Ph[2 * JX], j = Ph[2 * JX + 1];
        double Eden = 0.0;
        Eden = Eden + monopole(a, bb, Hd);
        Eden = Eden + monopole(i, j, Hd);
        Eden = Eden - monopole(a, i, Hd);
        Eden = Eden - monopole(a, j, Hd);
        Eden = Eden - monopole(i, bb, Hd);
        Eden = Eden - monopole(j, bb, Hd);
        Eden = Eden + f_elem(a, a, Hd) + f_elem(bb, bb, Hd) - f_elem(i, i, Hd) - f_elem(j, j, Hd);
        ETA.gam(q, 2)[IX + (i64)JX * np] = val(Hn.gam(q, 2)[IX + (i64)JX * np], Eden);
      }
    }
  }
}

void white(const Op& H, Op& ETA) { gen_core(H, H, ETA, 0); }

// mbpt2, basic_IMSRG.f90:4165-4199
double mbpt2(const Op& H) {
  const Basis& b = *H.b;
  double sm = 0.0;
  for (int q = 0; q < b.nblocks; ++q) {
    int np = b.dim[q * 3], nh = b.dim[q * 3 + 2];
    const std::vector<int>& Pp = b.qn[q * 3 + 0];
    const std::vector<int>& Ph = b.qn[q * 3 + 2];
    double s1 = 0.0;
    for (int II = 0; II < np; ++II) {
      double fi = f_elem(Pp[2 * II], Pp[2 * II], H), fj = f_elem(Pp[2 * II + 1], Pp[2 * II + 1], H);
      for (int JJ = 0; JJ < nh; ++JJ) {
        double fk = f_elem(Ph[2 * JJ], Ph[2 * JJ], H), fl = f_elem(Ph[2 * JJ + 1], Ph[2 * JJ + 1], H);
        double eden = fk + fl - fi - fj;
        double x = H.gam(q, 2)[II + (i64)JJ * np];
        s1 = s1 + x * x / eden;
      }
    }
    sm = sm + (b.chJ[q] + 1.0) * s1;
  }
  return sm;
}

// normal_order, basic_IMSRG.f90:2419-2550 (vacuum operator -> normal ordered)
void normal_order(Op& H) {
  const Basis& b = *H.b;
  std::vector<double> tmp(H.d, H.d + b.oplen);
  Op T{&b, tmp.data(), H.herm};
  double sm = 0.0;
  for (int i = 0; i < b.N; ++i) sm = sm + f_elem(i, i, T) * b.con[i] * (b.jj[i] + 1);
  sm = sm * 2.0;
  for (int i = 0; i < b.N; ++i)
    for (int j = 0; j < b.N; ++j) {
      if (b.con[i] * b.con[j] == 0) continue;
      for (int JT = std::abs(b.jj[i] - b.jj[j]); JT <= b.jj[i] + b.jj[j]; JT += 2)
        sm = sm + v_elem(i, j, i, j, JT, T) * (JT + 1);
    }
  H.d[0] = sm * 0.5;
  auto corr = [&](int ax, int bx) {
    double s = 0.0;
    for (int i = 0; i < b.A; ++i) {
      int ix = b.holes[i];
      for (int JT = std::abs(b.jj[ix] - b.jj[ax]); JT <= b.jj[ix] + b.jj[ax]; JT += 2)
        s = s + (JT + 1.0) / (b.jj[ax] + 1) * v_elem(ax, ix, bx, ix, JT, T);
    }
    return s;
  };
  auto same = [&](int x, int y) {
    return b.jj[x] == b.jj[y] && b.ll[x] == b.ll[y] && b.tz[x] == b.tz[y];
  };
  for (int a = 0; a < b.A; ++a)
    for (int c = a; c < b.A; ++c) {
      int ax = b.holes[a], bx = b.holes[c];
      if (!same(ax, bx)) continue;
      double v = f_elem(ax, bx, T) + corr(ax, bx);
      H.fhh()[a + (i64)c * b.A] = v;
      H.fhh()[c + (i64)a * b.A] = v;
    }
  for (int a = 0; a < b.P; ++a)
    for (int c = 0; c < b.A; ++c) {
      int ax = b.parts[a], bx = b.holes[c];
      if (!same(ax, bx)) continue;
      H.fph()[a + (i64)c * b.P] = f_elem(ax, bx, T) + corr(ax, bx);
    }
  for (int a = 0; a < b.P; ++a)
    for (int c = a; c < b.P; ++c) {
      int ax = b.parts[a], bx = b.parts[c];
      if (!same(ax, bx)) continue;
      double v = f_elem(ax, bx, T) + corr(ax, bx);
      H.fpp()[a + (i64)c * b.P] = v;
      H.fpp()[c + (i64)a * b.P] = v;
    }
}

// Scratch for the nested-commutator drivers.
struct Work {
  std::vector<double> w1, w2, LCC, RCC, WCC;
  void fit(const Basis& b) {
    if ((i64)w1.size() != b.oplen) { w1.assign(b.oplen, 0.0); w2.assign(b.oplen, 0.0); }
    if ((i64)LCC.size() != b.cclen) { LCC.assign(b.cclen, 0.0); RCC.assign(b.cclen, 0.0); }
    if ((i64)WCC.size() != b.wclen) WCC.assign(b.wclen, 0.0);
  }
};

// The sequence of term calls used by BCH_EXPAND / MAGNUS_EXPAND
// (IMSRG_MAGNUS.f90:207-219, 677-683): INT2 = [G, AD] without E0.
void full_commutator(const Op& G, const Op& AD, Op& INT2, Work& w) {
  const Basis& b = *G.b;
  w.fit(b);
  Op w1{&b, w.w1.data(), 1}, w2{&b, w.w2.data(), 1};
  cross_coupled(AD, w.RCC.data());
  cross_coupled(G, w.LCC.data());
  comm111(G, AD, INT2);
  comm121(G, AD, INT2);
  comm122(G, AD, INT2);
  comm222pphh(G, AD, INT2, w1, w2);
  comm221(G, AD, INT2, w1, w2);
  comm222ph(w.LCC.data(), G.herm, w.RCC.data(), AD.herm, INT2, w.WCC.data());
}

// build_gs_w2, generators.f90:109-160: eta1 = build_gs_atan(H); H2 =
// [eta1, H] with E0 = 110 + 220; eta = 1.5 eta1 + 0.5 build_gs_atanxx(H, H2)
// (add_sq_op, basic_IMSRG.f90:3551-3579; eta1 and eta2 start from zero)
void gs_w2(const Op& H, Op& ETA, Work& w) {
  const Basis& b = *H.b;
  std::vector<double> e1(b.oplen, 0.0), h2(b.oplen, 0.0), e2(b.oplen, 0.0);
  Op eta1{&b, e1.data(), -1}, H2{&b, h2.data(), 1}, eta2{&b, e2.data(), -1};
  gen_core(H, H, eta1, 1);
  h2[0] = comm110(eta1, H) + comm220(eta1, H);
  full_commutator(eta1, H, H2, w);
  gen_core(H, H2, eta2, 1);
  for (i64 x = 0; x < b.oplen; ++x) ETA.d[x] = e1[x] * 1.5 + e2[x] * 0.5;
}

// the generator of a flow: 0 white, 1 atan, 2 w2
void build_generator(int gen, const Op& H, Op& ETA, Work& w) {
  if (gen == 1) gen_core(H, H, ETA, 1);
  else if (gen == 2) gs_w2(H, ETA, w);
  else white(H, ETA);
}

// BCH_EXPAND, IMSRG_MAGNUS.f90:149-234.  HS = exp(G) H exp(-G); returns
// the number of loop iterations executed (iw of the exit) or -1 on clipping.
int bch(Op& HS, const Op& G, const Op& H, double conv, int maxit, Work& w,
        double* advals_out) {
  const Basis& b = *G.b;
  std::vector<double> int1(b.oplen), int2(b.oplen), ad(b.oplen);
  Op INT1{&b, int1.data(), 1}, INT2{&b, int2.data(), 1}, AD{&b, ad.data(), 1};
  std::copy(H.d, H.d + b.oplen, HS.d); HS.herm = H.herm;
  std::copy(HS.d, HS.d + b.oplen, int2.data()); INT2.herm = HS.herm;
  double coef = 1.0;
  int iw;
  for (iw = 2; iw <= maxit; ++iw) {
    coef = coef / (iw - 1.0);
    std::copy(HS.d, HS.d + b.oplen, int1.data()); INT1.herm = HS.herm;
    std::copy(int2.begin(), int2.end(), ad.begin()); AD.herm = INT2.herm;
    std::fill(int2.begin(), int2.end(), 0.0);
    int2[0] = comm110(G, AD) + comm220(G, AD);
    full_commutator(G, AD, INT2, w);
    add_op(b, int1.data(), 1.0, int2.data(), coef, HS.d);
    double adv = frob_norm(INT2) * coef;
    if (advals_out) advals_out[iw - 2] = adv;
    if (adv < conv) break;
  }
  if (iw > maxit) return -1;
  return iw;
}

// MAGNUS_EXPAND, IMSRG_MAGNUS.f90:620-691.  Returns the number of nested
// commutators evaluated.
int magnus(Op& DG, const Op& G, double conv, Work& w) {
  const Basis& b = *G.b;
  const double cof[7] = {1.0, -0.5, .0833333333, 0.0, -0.00138888888, 0.0, 3.306878e-5};
  std::vector<double> int1(b.oplen), int2(b.oplen), ad(b.oplen);
  Op INT1{&b, int1.data(), -1}, INT2{&b, int2.data(), -1}, AD{&b, ad.data(), -1};
  double advals[7] = {0, 0, 0, 0, 0, 0, 0};
  std::copy(DG.d, DG.d + b.oplen, int2.data()); INT2.herm = DG.herm;
  advals[0] = frob_norm(INT2);
  double fullnorm = frob_norm(G);
  if (fullnorm < (double)1e-9f) return 0;
  int q = 1, ncomm = 0;
  for (int i = 2; i <= 7; ++i) {
    std::copy(DG.d, DG.d + b.oplen, int1.data()); INT1.herm = DG.herm;
    std::copy(int2.begin(), int2.end(), ad.begin()); AD.herm = INT2.herm;
    double adnorm = advals[i - q - 1];
    q = (std::fabs(cof[i - 1]) > (double)1e-6f) ? 1 : 2;
    if (std::fabs(adnorm / fullnorm) < conv) break;
    full_commutator(G, AD, INT2, w);
    ++ncomm;
    add_op(b, int1.data(), 1.0, int2.data(), cof[i - 1], DG.d);
    advals[i - 1] = frob_norm(INT2) * std::fabs(cof[i - 1]);
  }
  return ncomm;
}

#include "imsrg_adjoint.inc"
#include "imsrg_interaction.inc"
#include "imsrg_hf.inc"
#include "imsrg_tensor.inc"

}  // namespace

// =============================================================== C interface
extern "C" {

int orc_set_blas(const char* path) {
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return 1;
  void* f = dlsym(h, "scipy_dgemm_64_");
  if (!f) return 2;
  g_dgemm = (dgemm64_fn)f;
  return 0;
}
void orc_unset_blas() { g_dgemm = nullptr; }
int orc_set_threads(int n) { omp_set_num_threads(n); return omp_get_max_threads(); }

void* orc_basis_new(int N, const int* n, const int* l, const int* jj, const int* tz, const int* con) {
  Basis* b = new Basis();
  build_basis(*b, N, n, l, jj, tz, con);
  return b;
}
void orc_basis_free(void* h) { delete (Basis*)h; }

// sizes: [N, A, P, JTM, nblocks, oplen, ncc, cclen, wclen]
void orc_basis_sizes(void* h, i64* out) {
  Basis& b = *(Basis*)h;
  out[0] = b.N; out[1] = b.A; out[2] = b.P; out[3] = b.JTM; out[4] = b.nblocks;
  out[5] = b.oplen; out[6] = b.ncc; out[7] = b.cclen; out[8] = b.wclen;
}
// per channel: dims [nblocks*3], goff [nblocks*6], J/T/P [nblocks*3]
void orc_basis_channels(void* h, int* dims, i64* goff, int* jtp) {
  Basis& b = *(Basis*)h;
  for (int q = 0; q < b.nblocks; ++q) {
    for (int c = 0; c < 3; ++c) dims[q * 3 + c] = b.dim[q * 3 + c];
    for (int g = 0; g < 6; ++g) goff[q * 6 + g] = b.goff[q * 6 + g];
    jtp[q * 3] = b.chJ[q]; jtp[q * 3 + 1] = b.chT[q]; jtp[q * 3 + 2] = b.chP[q];
  }
}
// pairs of channel q, class c (0 pp, 1 ph, 2 hh): out[2*n]
int orc_basis_pairs(void* h, int q, int c, int* out) {
  Basis& b = *(Basis*)h;
  const std::vector<int>& v = b.qn[q * 3 + c];
  std::copy(v.begin(), v.end(), out);
  return (int)v.size() / 2;
}
// CC layout: rlen, ncol, ccoff, wcoff, J/T/P
void orc_basis_cc(void* h, int* rlen, int* ncol, i64* ccoff, i64* wcoff, int* jtp) {
  Basis& b = *(Basis*)h;
  for (int q = 0; q < b.ncc; ++q) {
    rlen[q] = b.rlen[q]; ncol[q] = b.ncol[q]; ccoff[q] = b.ccoff[q]; wcoff[q] = b.wcoff[q];
    jtp[q * 3] = b.ccJ[q]; jtp[q * 3 + 1] = b.ccT[q]; jtp[q * 3 + 2] = b.ccP[q];
  }
}
int orc_cc_row(void* h, int i, int j, int JC) {
  Basis& b = *(Basis*)h;
  return b.rmap[((size_t)i * b.N + j) * (b.JTM + 1) + JC / 2];
}
int orc_cc_col(void* h, int i, int j, int JC) {
  Basis& b = *(Basis*)h;
  return b.nbmap[((size_t)i * b.N + j) * (b.JTM + 1) + JC / 2];
}

double orc_sixj(int j1, int j2, int j3, int j4, int j5, int j6) { init_logfact(); return racah6j(j1, j2, j3, j4, j5, j6); }
double orc_ninej(int j1, int j2, int j3, int j4, int j5, int j6, int j7, int j8, int j9) {
  init_logfact(); return ninej(j1, j2, j3, j4, j5, j6, j7, j8, j9);
}
double orc_sixj_table(void* h, int j1, int j2, int j3, int j4, int j5, int j6) {
  return ((Basis*)h)->sixj(j1, j2, j3, j4, j5, j6);
}

double orc_v_elem(void* h, const double* op, int herm, int a, int bb, int c, int d, int J) {
  Basis& b = *(Basis*)h; Op o{&b, (double*)op, herm};
  return v_elem(a, bb, c, d, J, o);
}
double orc_f_elem(void* h, const double* op, int herm, int a, int c) {
  Basis& b = *(Basis*)h; Op o{&b, (double*)op, herm};
  return f_elem(a, c, o);
}

#define OPS(h) Basis& b = *(Basis*)h;
double orc_comm110(void* h, const double* L, int Lh, const double* R, int Rh) {
  OPS(h) return comm110(Op{&b, (double*)L, Lh}, Op{&b, (double*)R, Rh});
}
double orc_comm220(void* h, const double* L, int Lh, const double* R, int Rh) {
  OPS(h) return comm220(Op{&b, (double*)L, Lh}, Op{&b, (double*)R, Rh});
}
void orc_comm111(void* h, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh) {
  OPS(h) Op res{&b, RES, RESh}; comm111(Op{&b, (double*)L, Lh}, Op{&b, (double*)R, Rh}, res);
}
void orc_comm121(void* h, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh) {
  OPS(h) Op res{&b, RES, RESh}; comm121(Op{&b, (double*)L, Lh}, Op{&b, (double*)R, Rh}, res);
}
void orc_comm122(void* h, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh) {
  OPS(h) Op res{&b, RES, RESh}; comm122(Op{&b, (double*)L, Lh}, Op{&b, (double*)R, Rh}, res);
}
void orc_comm222pphh(void* h, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh,
                     double* w1, double* w2) {
  OPS(h) Op res{&b, RES, RESh}, W1{&b, w1, 1}, W2{&b, w2, 1};
  comm222pphh(Op{&b, (double*)L, Lh}, Op{&b, (double*)R, Rh}, res, W1, W2);
}
void orc_comm221(void* h, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh,
                 const double* w1, const double* w2) {
  OPS(h) Op res{&b, RES, RESh};
  comm221(Op{&b, (double*)L, Lh}, Op{&b, (double*)R, Rh}, res, Op{&b, (double*)w1, 1}, Op{&b, (double*)w2, 1});
}
void orc_cross_coupled(void* h, const double* op, int herm, double* CC) {
  OPS(h) cross_coupled(Op{&b, (double*)op, herm}, CC);
}
void orc_comm222ph(void* h, const double* LCC, int Lh, const double* RCC, int Rh, double* RES, int RESh,
                   double* WCC) {
  OPS(h) Op res{&b, RES, RESh}; comm222ph(LCC, Lh, RCC, Rh, res, WCC);
}
void orc_h0_harm_osc(void* h, double hw, int htype, double beta, double cmhw, int mass, double* op) {
  OPS(h) Op o{&b, op, 1};
  h0_harm_osc(o, hw, htype, beta, cmhw, mass);
}
i64 orc_me2j_count(int nfile, const int* l, const int* jj) { return me2j_values(nfile / 2, l, jj); }
int orc_me2j_convert(void* h, const double* me, i64 nme, int nfile, const int* l, const int* jj, const int* con,
                     double hw, int htype, double beta, double cmhw, int mass, double* op) {
  OPS(h) Op o{&b, op, 1};
  return me2j_convert(o, me, nme, nfile, l, jj, con, hw, htype, beta, cmhw, mass);
}
double orc_r1_r2(void* h, int a, int bb, int c, int d, int J) { OPS(h) return vec_dot(b, a, bb, c, d, J, false); }
double orc_p1_p2(void* h, int a, int bb, int c, int d, int J) { OPS(h) return vec_dot(b, a, bb, c, d, J, true); }
int orc_sp_layout(void* h, int* sizes) {
  OPS(h) for (size_t q = 0; q < b.states.size(); ++q) if (sizes) sizes[q] = (int)b.states[q].size();
  return (int)b.states.size();
}
void orc_hf_kin(void* h, const double* H, double* T) { OPS(h) hf_kin(Op{&b, (double*)H, 1}, T); }
void orc_hf_density(void* h, const double* D, double* rho) { OPS(h) hf_density(b, D, rho); }
void orc_hf_gamma(void* h, const double* H, const double* rho, double* gam) {
  OPS(h) hf_gamma(Op{&b, (double*)H, 1}, rho, gam);
}
void orc_hf_finish(void* h, double* H, const double* D, double* F, double* T, double* V) {
  OPS(h) Op o{&b, H, 1};
  hf_transform_1b(o, D, F, T, V);
  o.E0() = hf_energy(b, T, V);
  hf_transform_2b(o, D);
}
void orc_hf_transform_2b(void* h, double* H, const double* D) { OPS(h) Op o{&b, H, 1}; hf_transform_2b(o, D); }
void orc_sqop_to_hf(void* h, double* O, const double* D) { OPS(h) Op o{&b, O, 1}; hf_sqop_to_hf(o, D); }
double orc_frob_norm(void* h, const double* op) { OPS(h) return frob_norm(Op{&b, (double*)op, 1}); }
void orc_add(void* h, const double* A, double ax, const double* B, double bx, double* C) {
  OPS(h) add_op(b, A, ax, B, bx, C);
}
void orc_white(void* h, const double* H, double* ETA) {
  OPS(h) Op eta{&b, ETA, -1}; white(Op{&b, (double*)H, 1}, eta);
}
double orc_mbpt2(void* h, const double* H) { OPS(h) return mbpt2(Op{&b, (double*)H, 1}); }
void orc_normal_order(void* h, double* H) { OPS(h) Op o{&b, H, 1}; normal_order(o); }
double orc_monopole(void* h, const double* H, int a, int c) { OPS(h) return monopole(a, c, Op{&b, (double*)H, 1}); }

// ---- reverse-mode derivatives (imsrg_adjoint.inc); gradients accumulate
void orc_commutator_vjp(void* h, const double* L, int Lh, const double* R, int Rh, const double* RESb, int RESh,
                        double* Lb, double* Rb) {
  OPS(h) Op lb{&b, Lb, Lh}, rb{&b, Rb, Rh};
  adj_commutator(Op{&b, (double*)L, Lh}, Op{&b, (double*)R, Rh}, RESb, RESh, lb, rb);
}
int orc_bch_vjp(void* h, const double* G, const double* H, double conv, int maxit, const double* HSb, double* Gb,
                double* Hb) {
  OPS(h) Op gb{&b, Gb, -1}, hb{&b, Hb, 1};
  return adj_bch(Op{&b, (double*)G, -1}, Op{&b, (double*)H, 1}, conv, maxit, HSb, gb, hb);
}
int orc_magnus_vjp(void* h, const double* ETA, const double* G, double conv, const double* DGb, double* ETAb,
                   double* Gb) {
  OPS(h) Op eb{&b, ETAb, -1}, gb{&b, Gb, -1};
  return adj_magnus(Op{&b, (double*)ETA, -1}, Op{&b, (double*)G, -1}, conv, DGb, eb, gb);
}
void orc_white_vjp(void* h, const double* H, const double* ETAb, double* Hb) {
  OPS(h) Op hb{&b, Hb, 1};
  adj_white(Op{&b, (double*)H, 1}, Op{&b, (double*)ETAb, -1}, hb);
}
void orc_normal_order_vjp(void* h, const double* Hb, double* Tb) {
  OPS(h) Op tb{&b, Tb, 1};
  adj_normal_order(Op{&b, (double*)Hb, 1}, tb);
}
int orc_flow_grad(void* h, const double* T, double ds, double bch_conv, double mag_conv, double crit_tol,
                  int maxsteps, int ckpt, double* Tbar, double* E0) {
  OPS(h) return flow_grad(b, T, ds, bch_conv, mag_conv, crit_tol, maxsteps, ckpt, Tbar, E0, 0);
}
int orc_flow_grad_gen(void* h, const double* T, double ds, double bch_conv, double mag_conv, double crit_tol,
                      int maxsteps, int ckpt, double* Tbar, double* E0, int gen) {
  OPS(h) return flow_grad(b, T, ds, bch_conv, mag_conv, crit_tol, maxsteps, ckpt, Tbar, E0, gen);
}
void orc_gen_vjp(void* h, const double* Hd, const double* Hn, const double* ETAb, double* Hdb, double* Hnb,
                 int mode) {
  OPS(h) Op hdb{&b, Hdb, 1}, hnb{&b, Hnb, 1};
  adj_gen(Op{&b, (double*)Hd, 1}, Op{&b, (double*)Hn, 1}, Op{&b, (double*)ETAb, -1}, hdb, hnb, mode);
}
void orc_gs_w2_vjp(void* h, const double* H, const double* ETAb, double* Hb) {
  OPS(h) Op hb{&b, Hb, 1}; Work w;
  adj_gs_w2(Op{&b, (double*)H, 1}, Op{&b, (double*)ETAb, -1}, hb, w);
}

// One full scalar commutator RES = [L, R] in the BCH term order
// (E0 = 110 + 220, then 111, 121, 122, 222_pp_hh, 221, 222_ph).
void orc_commutator(void* h, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh) {
  OPS(h) Op Lo{&b, (double*)L, Lh}, Ro{&b, (double*)R, Rh}, res{&b, RES, RESh};
  Work w;
  std::fill(RES, RES + b.oplen, 0.0);
  RES[0] = comm110(Lo, Ro) + comm220(Lo, Ro);
  full_commutator(Lo, Ro, res, w);
}

int orc_bch(void* h, double* HS, const double* G, const double* H, double conv, int maxit, double* advals) {
  OPS(h) Op hs{&b, HS, 1}; Work w;
  return bch(hs, Op{&b, (double*)G, -1}, Op{&b, (double*)H, 1}, conv, maxit, w, advals);
}
int orc_magnus(void* h, double* DG, const double* G, double conv) {
  OPS(h) Op dg{&b, DG, -1}; Work w;
  return magnus(dg, Op{&b, (double*)G, -1}, conv, w);
}

// magnus_decouple, IMSRG_MAGNUS.f90:15-128, with the white generator.
// HS (in: bare normal-ordered H; out: decoupled H), G (out: Omega).
// crit_mode 0: loop while |mbpt2| > crit_tol (reference);
//           1: loop until |E0(s) - E0(s-ds)| < crit_tol.
// log[step*4 + {s, E0, E0+mbpt2, n_comm}] for up to maxsteps steps.
int orc_magnus_decouple_gen(void* h, double* HS, double* G, double ds, double bch_conv, double mag_conv,
                            int crit_mode, double crit_tol, int maxsteps, double* log, int gen) {
  OPS(h)
  Work w;
  std::vector<double> Hb(HS, HS + b.oplen), dg(b.oplen, 0.0);
  Op hs{&b, HS, 1}, g{&b, G, -1}, DG{&b, dg.data(), -1}, H{&b, Hb.data(), 1};
  build_generator(gen, hs, DG, w);
  double s = 0.0;
  double E_mbpt2 = mbpt2(hs);
  double crit = std::fabs(E_mbpt2);
  double Eold = HS[0];
  int steps = 0;
  while (true) {
    if (crit_mode == 0 && !(crit > crit_tol)) break;
    if (steps >= maxsteps) return -steps;
    ++steps;
    int nm = magnus(DG, g, mag_conv, w);
    add_op(b, G, 1.0, dg.data(), ds, G);      // euler_step, IMSRG_MAGNUS.f90:754-764
    s += ds;
    int nb = bch(hs, g, H, bch_conv, 40, w, nullptr);
    if (nb < 0) return -100000;
    build_generator(gen, hs, DG, w);
    E_mbpt2 = mbpt2(hs);
    crit = std::fabs(E_mbpt2);
    if (log) {
      log[(steps - 1) * 4 + 0] = s;
      log[(steps - 1) * 4 + 1] = HS[0];
      log[(steps - 1) * 4 + 2] = HS[0] + E_mbpt2;
      log[(steps - 1) * 4 + 3] = nm + (nb - 1);
    }
    if (crit_mode == 1) {
      if (std::fabs(HS[0] - Eold) < crit_tol) break;
      Eold = HS[0];
    }
  }
  return steps;
}

int orc_magnus_decouple(void* h, double* HS, double* G, double ds, double bch_conv, double mag_conv,
                        int crit_mode, double crit_tol, int maxsteps, double* log) {
  return orc_magnus_decouple_gen(h, HS, G, ds, bch_conv, mag_conv, crit_mode, crit_tol, maxsteps, log, 0);
}
void orc_gs_atan(void* h, const double* H, double* ETA) {
  OPS(h) Op eta{&b, ETA, -1}; gen_core(Op{&b, (double*)H, 1}, Op{&b, (double*)H, 1}, eta, 1);
}
void orc_gs_atanxx(void* h, const double* Hold, const double* H, double* ETA) {
  OPS(h) Op eta{&b, ETA, -1}; gen_core(Op{&b, (double*)Hold, 1}, Op{&b, (double*)H, 1}, eta, 1);
}
void orc_gs_w2(void* h, const double* H, double* ETA) {
  OPS(h) Op eta{&b, ETA, -1}; Work w; gs_w2(Op{&b, (double*)H, 1}, eta, w);
}

// --------------------------------------------- brute-force formulas (pins)
// scalar_scalar_0body_comm, brute_force_testing.f90:1887-1928
double orc_bf_0body(void* h, const double* A_, int Ah, const double* B_, int Bh) {
  OPS(h) Op AA{&b, (double*)A_, Ah}, BB{&b, (double*)B_, Bh};
  const int JTM2 = b.JTM * 2, T = b.N;
  double sm = 0.0;
  for (int i = 0; i < T; ++i)
    for (int j = 0; j < T; ++j)
      sm += (b.con[i] - b.con[j]) * (b.jj[i] + 1.0) * f_elem(i, j, AA) * f_elem(j, i, BB);
  for (int i = 0; i < T; ++i)
    for (int j = 0; j < T; ++j)
      for (int k = 0; k < T; ++k)
        for (int l = 0; l < T; ++l) {
          int occ = b.con[i] * b.con[j] * (1 - b.con[k]) * (1 - b.con[l]) -
                    b.con[k] * b.con[l] * (1 - b.con[i]) * (1 - b.con[j]);
          if (occ == 0) continue;
          for (int J = 0; J <= JTM2; J += 2)
            sm += occ * (J + 1.0) * v_elem(i, j, k, l, J, AA) * v_elem(k, l, i, j, J, BB) / 4.0;
        }
  return sm;
}
// scalar_scalar_1body_comm, brute_force_testing.f90:1765-1823
double orc_bf_1body(void* h, const double* A_, int Ah, const double* B_, int Bh, int a, int bb) {
  OPS(h) Op AA{&b, (double*)A_, Ah}, BB{&b, (double*)B_, Bh};
  const int JTM2 = b.JTM * 2, T = b.N;
  int ja = b.jj[a], jb = b.jj[bb];
  double sm = 0.0;
  for (int i = 0; i < T; ++i) sm += f_elem(a, i, AA) * f_elem(i, bb, BB) - f_elem(a, i, BB) * f_elem(i, bb, AA);
  if (ja == jb) {
    for (int i = 0; i < T; ++i)
      for (int j = 0; j < T; ++j) {
        int occ = b.con[i] - b.con[j];
        if (occ == 0) continue;
        for (int J = 0; J <= JTM2; J += 2)
          sm += occ * (J + 1.0) / (ja + 1.0) *
                (f_elem(i, j, AA) * v_elem(j, a, i, bb, J, BB) - f_elem(i, j, BB) * v_elem(j, a, i, bb, J, AA));
      }
    for (int i = 0; i < T; ++i)
      for (int j = 0; j < T; ++j)
        for (int k = 0; k < T; ++k) {
          int occ = b.con[i] * b.con[j] * (1 - b.con[k]) + (1 - b.con[i]) * (1 - b.con[j]) * b.con[k];
          if (occ == 0) continue;
          for (int J = 0; J <= JTM2; J += 2)
            sm += occ * (J + 1.0) / (ja + 1.0) *
                  (v_elem(a, k, i, j, J, AA) * v_elem(i, j, bb, k, J, BB) -
                   v_elem(a, k, i, j, J, BB) * v_elem(i, j, bb, k, J, AA)) / 2.0;
        }
  }
  return sm;
}
// scalar_scalar_2body_comm, brute_force_testing.f90:1975-2076.  The three
// groups (1b x 2b, ladder, ph via 9j) are returned separately in parts[3].
double orc_bf_2body(void* h, const double* A_, int Ah, const double* B_, int Bh, int a, int bb, int c,
                    int d, int Jtot, double* parts) {
  OPS(h) Op AA{&b, (double*)A_, Ah}, BB{&b, (double*)B_, Bh};
  const int JTM2 = b.JTM * 2, T = b.N;
  int ja = b.jj[a], jb = b.jj[bb], jc = b.jj[c], jd = b.jj[d];
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int i = 0; i < T; ++i) {
    s1 += f_elem(a, i, AA) * v_elem(i, bb, c, d, Jtot, BB) + f_elem(bb, i, AA) * v_elem(a, i, c, d, Jtot, BB) -
          f_elem(i, c, AA) * v_elem(a, bb, i, d, Jtot, BB) - f_elem(i, d, AA) * v_elem(a, bb, c, i, Jtot, BB);
    s1 -= f_elem(a, i, BB) * v_elem(i, bb, c, d, Jtot, AA) + f_elem(bb, i, BB) * v_elem(a, i, c, d, Jtot, AA) -
          f_elem(i, c, BB) * v_elem(a, bb, i, d, Jtot, AA) - f_elem(i, d, BB) * v_elem(a, bb, c, i, Jtot, AA);
  }
  for (int i = 0; i < T; ++i)
    for (int j = 0; j < T; ++j) {
      int occ = 1 - b.con[i] - b.con[j];
      if (occ == 0) continue;
      s2 += 0.5 * occ *
            (v_elem(a, bb, i, j, Jtot, AA) * v_elem(i, j, c, d, Jtot, BB) -
             v_elem(a, bb, i, j, Jtot, BB) * v_elem(i, j, c, d, Jtot, AA));
    }
  for (int i = 0; i < T; ++i) {
    int ji = b.jj[i];
    for (int j = 0; j < T; ++j) {
      int jjj = b.jj[j];
      int occ = b.con[i] - b.con[j];
      if (occ == 0) continue;
      for (int J1 = 0; J1 <= JTM2; J1 += 2)
        for (int J2 = 0; J2 <= JTM2; J2 += 2) {
          double t = 0.0;
          t += ipow_m1((J1 + J2 + ja - jc) / 2) * (J1 + 1.0) * (J2 + 1.0) *
               ninej(jjj, J1, ja, J2, ji, jb, jc, jd, Jtot) * v_elem(j, a, i, d, J1, AA) * v_elem(i, bb, j, c, J2, BB);
          t -= ipow_m1((J1 + J2 + jb - jc) / 2) * (J1 + 1.0) * (J2 + 1.0) *
               ninej(jjj, J1, jb, J2, ji, ja, jc, jd, Jtot) * v_elem(j, bb, i, d, J1, AA) *
               v_elem(i, a, j, c, J2, BB) * ipow_m1((ja + jb - Jtot) / 2);
          t -= ipow_m1((J1 + J2 + ja - jd) / 2) * (J1 + 1.0) * (J2 + 1.0) *
               ninej(jjj, J1, ja, J2, ji, jb, jd, jc, Jtot) * v_elem(j, a, i, c, J1, AA) *
               v_elem(i, bb, j, d, J2, BB) * ipow_m1((jc + jd - Jtot) / 2);
          t += ipow_m1((J1 + J2 + jb - jd) / 2) * (J1 + 1.0) * (J2 + 1.0) *
               ninej(jjj, J1, jb, J2, ji, ja, jd, jc, Jtot) * v_elem(j, bb, i, c, J1, AA) *
               v_elem(i, a, j, d, J2, BB) * ipow_m1((ja + jb + jc + jd) / 2);
          s3 += occ * t;
        }
    }
  }
  if (parts) { parts[0] = s1; parts[1] = s2; parts[2] = s3; }
  return s1 + s2 + s3;
}

// ------------------------------------------------------------- tensors
void* orc_tensor_new(void* h, int rank, int dpar) {
  TLayout* t = new TLayout();
  build_tensor(*t, *(Basis*)h, rank, dpar);
  return t;
}
void orc_tensor_free(void* th) { delete (TLayout*)th; }
// out: oplen, nblocks, ncc, cclen, wlen
void orc_tensor_sizes(void* th, i64* out) {
  const TLayout& t = *(TLayout*)th;
  out[0] = t.oplen; out[1] = t.nblocks; out[2] = t.ncc; out[3] = t.cclen; out[4] = t.wlen;
}
// per block: info[7q..] = 2J1, 2J2, Tz, P1, P2, q1, q2; dims[6q..]; goff[9q..]
void orc_tensor_blocks(void* th, int* info, int* dims, i64* goff) {
  const TLayout& t = *(TLayout*)th;
  for (int q = 0; q < t.nblocks; ++q) {
    const int v[7] = {t.J1[q], t.J2[q], t.T[q], t.P1[q], t.P2[q], t.q1[q], t.q2[q]};
    for (int k = 0; k < 7; ++k) info[7 * q + k] = v[k];
    for (int k = 0; k < 6; ++k) dims[6 * q + k] = t.dim[6 * q + k];
    for (int k = 0; k < 9; ++k) goff[9 * q + k] = t.goff[9 * q + k];
  }
}
int orc_tensor_block_index_ref(int J1, int J2, int rank, int T, int P) {
  return tensor_block_index_ref(J1, J2, rank, T, P);
}
double orc_xxxsixj(int J1, int J2, int R, int j4, int j5, int j6) { init_logfact(); return xxxsixj(J1, J2, R, j4, j5, j6); }
double orc_ninej_ref(int a, int bb, int J1, int c, int d, int J2, int J3, int J4, int R) {
  init_logfact();
  return ninej_ref(a, bb, J1, c, d, J2, J3, J4, R);
}
#define TOPS(th) const TLayout& t = *(TLayout*)th; const Basis& b = *t.b; (void)b;
double orc_tensor_elem(void* th, const double* op, int herm, int a, int bb, int c, int d, int J1, int J2) {
  TOPS(th) return tensor_elem(a, bb, c, d, J1, J2, TOp{&t, (double*)op, herm});
}
double orc_f_tensor_elem(void* th, const double* op, int herm, int a, int c) {
  TOPS(th) return f_tensor_elem(a, c, TOp{&t, (double*)op, herm});
}
void orc_ts111(void* th, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh) {
  TOPS(th) TOp res{&t, RES, RESh}; ts111(Op{&b, (double*)L, Lh}, TOp{&t, (double*)R, Rh}, res);
}
void orc_ts121(void* th, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh) {
  TOPS(th) TOp res{&t, RES, RESh}; ts121(Op{&b, (double*)L, Lh}, TOp{&t, (double*)R, Rh}, res);
}
void orc_ts211(void* th, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh) {
  TOPS(th)
  std::vector<double> LCC(b.cclen, 0.0);
  cross_coupled(Op{&b, (double*)L, Lh}, LCC.data());
  TOp res{&t, RES, RESh};
  ts211(LCC.data(), Lh, TOp{&t, (double*)R, Rh}, res);
}
void orc_ts122(void* th, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh) {
  TOPS(th) TOp res{&t, RES, RESh}; ts122(Op{&b, (double*)L, Lh}, TOp{&t, (double*)R, Rh}, res);
}
void orc_ts212(void* th, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh) {
  TOPS(th) TOp res{&t, RES, RESh}; ts212(Op{&b, (double*)L, Lh}, TOp{&t, (double*)R, Rh}, res);
}
void orc_ts222pphh(void* th, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh, double* w1,
                   double* w2) {
  TOPS(th) TOp res{&t, RES, RESh}, W1{&t, w1, Rh}, W2{&t, w2, Rh};
  ts222pphh(Op{&b, (double*)L, Lh}, TOp{&t, (double*)R, Rh}, res, W1, W2);
}
void orc_ts221(void* th, const double* w1, const double* w2, int pm, double* RES, int RESh) {
  TOPS(th) TOp res{&t, RES, RESh};
  ts221(TOp{&t, (double*)w1, 1}, TOp{&t, (double*)w2, 1}, pm, res);
}
void orc_ts222ph(void* th, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh) {
  TOPS(th)
  std::vector<double> LCC(b.cclen, 0.0), CC(t.cclen, 0.0), Wx(t.wlen, 0.0), Wy(t.wlen, 0.0);
  cross_coupled(Op{&b, (double*)L, Lh}, LCC.data());
  TOp res{&t, RES, RESh};
  ts222ph(LCC.data(), Lh, TOp{&t, (double*)R, Rh}, res, CC.data(), Wx.data(), Wy.data());
}
// calculate_single_pandya of every block (both terms) into CC (cclen doubles)
void orc_tensor_pandya(void* th, const double* R, int Rh, double* CC) {
  TOPS(th)
  TOp r{&t, (double*)R, Rh};
  for (int qx = 0; qx < t.ncc; ++qx) {
    if (t.cJ4[qx] > 2 * b.JTM) continue;
    if ((i64)t.r1[qx] * t.nb2[qx] > 0) single_pandya(r, qx, 1, CC);
    if ((i64)t.nb1[qx] * t.r2[qx] > 0) single_pandya(r, qx, 2, CC);
  }
}
void orc_ts_commutator(void* th, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh) {
  TOPS(th)
  TOp res{&t, RES, RESh};
  TWork w;
  ts_commutator(Op{&b, (double*)L, Lh}, TOp{&t, (double*)R, Rh}, res, w);
}
int orc_bch_tensor(void* th, const double* G, int Gh, double* HS, int HSh, double conv, double* advals) {
  TOPS(th)
  TOp hs{&t, HS, HSh};
  return bch_tensor(Op{&b, (double*)G, Gh}, hs, conv, advals);
}
double orc_tensor_frob_norm(void* th, const double* op) { TOPS(th) return tensor_frob_norm(TOp{&t, (double*)op, 1}); }
void orc_copy_rank0_to_tensor(void* th, const double* A, int Ah, double* B) {
  TOPS(th) TOp tb{&t, B, Ah}; copy_rank0_to_tensor(Op{&b, (double*)A, Ah}, tb);
}

}  // extern "C"

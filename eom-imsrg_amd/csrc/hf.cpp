// Host side of the Hartree-Fock basis (calc_HF, HF_mod.f90:14-105, no
// three-body forces): the self-consistent loop over the single-particle
// (tz, l, j) blocks, the one-body transform and e_HF, and the per-channel
// matrices of the two-body transform.  The blocks hold a few orbits each, so
// the eigenproblems use a cyclic Jacobi solver (eigenvalues ascending,
// eigenvector signs fixed afterwards as the reference fixes them).
#include "hf.hpp"

#include <algorithm>
#include <cmath>

namespace imsrg {

namespace {

inline int m1pow(int e) { return (e & 1) ? -1 : 1; }

// v_elem (basic_IMSRG.f90:1314-1429) on a hermitian host operator
double v_host(const Basis& b, const double* op, int a, int bb, int c, int d, int J) {
  auto tri = [](int x, int y, int z) { return z <= x + y && z >= std::abs(x - y); };
  if (!tri(b.jj[a], b.jj[bb], J) || !tri(b.jj[c], b.jj[d], J)) return 0.0;
  const int P = (b.ll[a] + b.ll[bb]) % 2;
  if ((b.ll[c] + b.ll[d]) % 2 != P) return 0.0;
  const int T = (b.tz[a] + b.tz[bb]) / 2;
  if (b.tz[c] + b.tz[d] != 2 * T) return 0.0;
  const int q = 3 * J + 2 * (T + 1) + P;
  const int C1 = b.con[a] + b.con[bb] + 1, C2 = b.con[c] + b.con[d] + 1;
  int qx = C1 * C2;
  if (qx == 9) qx = 5;
  const int g = qx - 1, N = b.N, nJ = b.nJ;
  double pre = 1.0;
  int i1, i2;
  if (a > bb) {
    i1 = b.xmap[((size_t)bb * N + a) * nJ + J / 2];
    pre = m1pow(1 + (b.jj[a] + b.jj[bb] - J) / 2);
  } else {
    if (a == bb) pre *= std::sqrt(2.0);
    i1 = b.xmap[((size_t)a * N + bb) * nJ + J / 2];
  }
  if (c > d) {
    i2 = b.xmap[((size_t)d * N + c) * nJ + J / 2];
    pre *= m1pow(1 + (b.jj[c] + b.jj[d] - J) / 2);
  } else {
    if (c == d) pre *= std::sqrt(2.0);
    i2 = b.xmap[((size_t)c * N + d) * nJ + J / 2];
  }
  const double* X = op + b.goff[q * 6 + g];
  const i64 ld = b.rows(q, g);
  if (C1 > C2) return X[i2 + (i64)i1 * ld] * pre;
  return X[i1 + (i64)i2 * ld] * pre;
}

// f_elem (basic_IMSRG.f90:1191-1221), hermitian operator
double f_host(const Basis& b, const double* op, int a, int c) {
  const int c1 = b.con[a], c2 = b.con[c];
  if (c1 + c2 == 0) return op[b.off_fpp + b.cidx[a] + (i64)b.cidx[c] * b.P];
  if (c1 + c2 == 2) return op[b.off_fhh + b.cidx[a] + (i64)b.cidx[c] * b.A];
  if (c1 > c2) return op[b.off_fph + b.cidx[c] + (i64)b.cidx[a] * b.P];
  return op[b.off_fph + b.cidx[a] + (i64)b.cidx[c] * b.P];
}

// eigenvalues (ascending) and eigenvectors (columns of V) of the symmetric
// n x n column-major A, by cyclic Jacobi rotations
void jacobi_eigen(int n, const double* A, double* w, double* V) {
  std::vector<double> a(A, A + (size_t)n * n);
  for (int x = 0; x < n * n; ++x) V[x] = 0.0;
  for (int i = 0; i < n; ++i) V[i + i * n] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        tot += a[i + j * n] * a[i + j * n];
        if (i != j) off += a[i + j * n] * a[i + j * n];
      }
    if (off <= 1e-34 * tot || off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[p + q * n];
        if (apq == 0.0) continue;
        const double theta = (a[q + q * n] - a[p + p * n]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {   // columns p, q
          const double akp = a[k + p * n], akq = a[k + q * n];
          a[k + p * n] = c * akp - s * akq;
          a[k + q * n] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {   // rows p, q
          const double apk = a[p + k * n], aqk = a[q + k * n];
          a[p + k * n] = c * apk - s * aqk;
          a[q + k * n] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[k + p * n], vkq = V[k + q * n];
          V[k + p * n] = c * vkp - s * vkq;
          V[k + q * n] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return a[x + x * n] < a[y + y * n]; });
  std::vector<double> Vs((size_t)n * n);
  for (int k = 0; k < n; ++k) {
    w[k] = a[ord[k] + ord[k] * n];
    for (int r = 0; r < n; ++r) Vs[r + k * n] = V[r + ord[k] * n];
  }
  std::copy(Vs.begin(), Vs.end(), V);
}

// X <- D^T X D, one n x n block
void rotate(int n, const double* D, double* X) {
  std::vector<double> t((size_t)n * n, 0.0), y((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += X[i + k * n] * D[k + j * n];
      t[i + j * n] = s;
    }
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int k = 0; k < n; ++k) s += D[k + i * n] * t[k + j * n];
      y[i + j * n] = s;
    }
  std::copy(y.begin(), y.end(), X);
}


// the Fock-matrix remap of transform_1b_to_HF (HF_mod.f90:531-560): block q's
// rotated matrix F (n x n) into fpp / fph / fhh of op
void write_one_body(const Basis& b, int q, const double* Fq, double* op) {
  const int n = b.spstart[q + 1] - b.spstart[q];
  auto st = [&](int i) { return b.spmem[b.spstart[q] + i]; };
  for (int a = 0; a < n; ++a)
    for (int c = a; c < n; ++c) {
      const int i = st(a), j = st(c);
      const int c1 = b.con[i], c2 = b.con[j];
      if (c1 + c2 == 0) {
        op[b.off_fpp + b.cidx[i] + (i64)b.cidx[j] * b.P] = Fq[a + c * n];
        op[b.off_fpp + b.cidx[j] + (i64)b.cidx[i] * b.P] = Fq[a + c * n];
      } else if (c1 + c2 == 1) {
        if (c2 > c1) op[b.off_fph + b.cidx[i] + (i64)b.cidx[j] * b.P] = Fq[a + c * n];
        else op[b.off_fph + b.cidx[j] + (i64)b.cidx[i] * b.P] = Fq[c + a * n];
      } else {
        op[b.off_fhh + b.cidx[i] + (i64)b.cidx[j] * b.A] = Fq[a + c * n];
        op[b.off_fhh + b.cidx[j] + (i64)b.cidx[i] * b.A] = Fq[a + c * n];
      }
    }
}

}  // namespace

// transform_2b_to_HF (HF_mod.f90:567-692): the four products V1 - s (V2 + V3)
// + V4 are one two-sided product with M = C - s Cr, s = (-1)^J, per channel
void transform_matrices(const Basis& b, const std::vector<i64>& boff, const std::vector<double>& D,
                        std::vector<i64>& moff, std::vector<double>& M) {
  const int nbk = (int)b.spstart.size() - 1;
  auto bn = [&](int q) { return b.spstart[q + 1] - b.spstart[q]; };
  auto st = [&](int q, int i) { return b.spmem[b.spstart[q] + i]; };
  std::vector<double> Ds((size_t)b.N * b.N, 0.0);
  for (int q = 0; q < nbk; ++q)
    for (int i = 0; i < bn(q); ++i)
      for (int j = 0; j < bn(q); ++j) Ds[st(q, i) + (size_t)st(q, j) * b.N] = D[boff[q] + i + (i64)j * bn(q)];
  auto dsm = [&](int x, int y) { return Ds[x + (size_t)y * b.N]; };
  moff.assign(b.nblocks + 1, 0);
  for (int q = 0; q < b.nblocks; ++q) {
    const i64 nt = (i64)b.dims[q * 3] + b.dims[q * 3 + 1] + b.dims[q * 3 + 2];
    moff[q + 1] = moff[q] + nt * nt;
  }
  M.assign((size_t)moff[b.nblocks], 0.0);
  for (int q = 0; q < b.nblocks; ++q) {
    const int np = b.dims[q * 3], nb = b.dims[q * 3 + 1], nh = b.dims[q * 3 + 2], nt = np + nb + nh;
    if (nt == 0) continue;
    std::vector<int> qa(nt), qb(nt);
    const int r0[3] = {nh + nb, nh, 0};   // rows of the pp, ph, hh pairs
    for (int c = 0; c < 3; ++c)
      for (int x = 0; x < b.dims[q * 3 + c]; ++x) {
        qa[r0[c] + x] = b.pl[q * 3 + c][2 * x];
        qb[r0[c] + x] = b.pl[q * 3 + c][2 * x + 1];
      }
    const double s = m1pow(b.chJ[q] / 2);
    double* Mq = M.data() + moff[q];
    for (int II = 0; II < nt; ++II) {
      const int i = qa[II], j = qb[II];
      for (int JJ = 0; JJ < nt; ++JJ) {
        const int a = qa[JJ], bb = qb[JJ];
        const double nrm = std::sqrt(1.0 + (a == bb)) / std::sqrt(1.0 + (i == j));
        const double c = dsm(a, i) * dsm(bb, j) * nrm;
        const double cr = a == bb ? 0.0 : dsm(bb, i) * dsm(a, j) * m1pow((b.jj[a] + b.jj[bb]) / 2) * nrm;
        Mq[JJ + (i64)II * nt] = c - s * cr;
      }
    }
  }
}

std::string hartree_fock_host(const Basis& b, double* op, int maxit, HFResult& r, std::vector<double>& M) {
  const int nbk = (int)b.spstart.size() - 1;
  const int nj = (b.JTM + 1) / 2;
  auto bn = [&](int q) { return b.spstart[q + 1] - b.spstart[q]; };
  auto st = [&](int q, int i) { return b.spmem[b.spstart[q] + i]; };
  auto bj = [&](int q) { return (q % nj) * 2 + 1; };   // (tz, l, j) order: j fastest
  r.boff.assign(nbk + 1, 0);
  for (int q = 0; q < nbk; ++q) r.boff[q + 1] = r.boff[q] + (i64)bn(q) * bn(q);
  const size_t tot = (size_t)r.boff[nbk];
  std::vector<double> T(tot, 0.0), F(tot, 0.0), Vg(tot, 0.0), rho(tot, 0.0), ev(b.N, 0.0), w(b.N, 0.0);
  r.D.assign(tot, 0.0);
  // write_kin_matrix (HF_mod.f90:211-240) and the identity start (54-61)
  for (int q = 0; q < nbk; ++q) {
    const int n = bn(q);
    double* Tq = T.data() + r.boff[q];
    for (int i = 0; i < n; ++i) {
      r.D[r.boff[q] + i + (i64)i * n] = 1.0;
      for (int j = i; j < n; ++j) {
        Tq[i + j * n] = f_host(b, op, st(q, i), st(q, j));
        Tq[j + i * n] = Tq[i + j * n];
      }
    }
  }
  double crit = 10.0;
  int it = 0;
  std::vector<double> Vec;
  while (crit > (double)1e-6f) {
    if (it >= maxit) return "Hartree-Fock did not converge in " + std::to_string(maxit) + " iterations";
    // density_matrix (HF_mod.f90:419-446)
    for (int q = 0; q < nbk; ++q) {
      const int n = bn(q);
      const double* Dq = r.D.data() + r.boff[q];
      double* R = rho.data() + r.boff[q];
      for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
          double s = 0.0;
          for (int k = 0; k < n; ++k) s += Dq[i + k * n] * b.con[st(q, k)] * (Dq[j + k * n] * b.con[st(q, k)]);
          R[i + j * n] = s * (bj(q) + 1);
        }
    }
    // gamma_matrix (HF_mod.f90:242-305)
    for (int q = 0; q < nbk; ++q) {
      const int n = bn(q), jfoc = bj(q);
      double* G = Vg.data() + r.boff[q];
      for (int i = 0; i < n; ++i)
        for (int j = i; j < n; ++j) {
          const int n1 = st(q, i), n3 = st(q, j);
          double sm = 0.0;
          for (int qr = 0; qr < nbk; ++qr) {
            const int m = bn(qr), jrho = bj(qr);
            const double* R = rho.data() + r.boff[qr];
            for (int g = 0; g < m; ++g)
              for (int h = 0; h < m; ++h) {
                const int n2 = st(qr, g), n4 = st(qr, h);
                for (int JJ = std::abs(jrho - jfoc); JJ <= jrho + jfoc; JJ += 2)
                  sm = sm + R[h + g * m] * v_host(b, op, n1, n2, n3, n4, JJ) * (JJ + 1.0) / (jrho + 1.0);
              }
          }
          G[i + j * n] = sm / (jfoc + 1.0);
          G[j + i * n] = sm / (jfoc + 1.0);
        }
    }
    // Fock matrix, diagonalize_blocks, the eigenvalue criterion (63-85)
    crit = 0.0;
    for (int q = 0; q < nbk; ++q) {
      const int n = bn(q);
      if (n == 0) continue;
      for (i64 x = r.boff[q]; x < r.boff[q + 1]; ++x) F[x] = T[x] + Vg[x] + 0.0;
      Vec.assign((size_t)n * n, 0.0);
      jacobi_eigen(n, F.data() + r.boff[q], w.data(), Vec.data());
      std::copy(Vec.begin(), Vec.end(), r.D.begin() + r.boff[q]);
      double s2 = 0.0;
      for (int i = 0; i < n; ++i) s2 += (ev[b.spstart[q] + i] - w[i]) * (ev[b.spstart[q] + i] - w[i]);
      crit = crit + std::sqrt(s2);
      for (int i = 0; i < n; ++i) ev[b.spstart[q] + i] = w[i];
    }
    ++it;
  }
  r.iterations = it;
  r.eig = ev;
  // signs (88-93), the Fock matrix of the last density, transform_1b_to_HF
  // (481-564) and e_HF (449-478)
  double ehf = 0.0;
  for (int q = 0; q < nbk; ++q) {
    const int n = bn(q);
    if (n == 0) continue;
    double* Dq = r.D.data() + r.boff[q];
    for (int i = 0; i < n; ++i)
      if (Dq[i + i * n] < 0.0)
        for (int k = 0; k < n; ++k) Dq[k + i * n] = Dq[k + i * n] * (-1);
    double *Fq = F.data() + r.boff[q], *Tq = T.data() + r.boff[q], *Vq = Vg.data() + r.boff[q];
    for (int x = 0; x < n * n; ++x) Fq[x] = Tq[x] + Vq[x] + 0.0;
    rotate(n, Dq, Fq);
    rotate(n, Dq, Tq);
    rotate(n, Dq, Vq);
    write_one_body(b, q, Fq, op);
  }
  for (int q = 0; q < nbk; ++q) {
    const int n = bn(q);
    const double *Tq = T.data() + r.boff[q], *Vq = Vg.data() + r.boff[q];
    for (int i = 0; i < n; ++i)
      ehf = ehf + (Tq[i + i * n] + 0.5 * Vq[i + i * n] + (1.0 / 3.0) * 0.0) * b.con[st(q, i)] * (bj(q) + 1.0);
  }
  r.ehf = ehf;
  op[0] = ehf;
  transform_matrices(b, r.boff, r.D, r.moff, M);
  return "";
}

std::string sqop_to_hf_host(const Basis& b, const double* D, double* op, std::vector<i64>& moff,
                            std::vector<double>& M) {
  const int nbk = (int)b.spstart.size() - 1;
  std::vector<i64> boff(nbk + 1, 0);
  for (int q = 0; q < nbk; ++q) {
    const i64 n = b.spstart[q + 1] - b.spstart[q];
    boff[q + 1] = boff[q] + n * n;
  }
  // write_kin_matrix (HF_mod.f90:211-240), then the transform_1b_to_HF pair
  for (int q = 0; q < nbk; ++q) {
    const int n = b.spstart[q + 1] - b.spstart[q];
    if (n == 0) continue;
    std::vector<double> T((size_t)n * n);
    for (int i = 0; i < n; ++i)
      for (int j = i; j < n; ++j) {
        T[i + (size_t)j * n] = f_host(b, op, b.spmem[b.spstart[q] + i], b.spmem[b.spstart[q] + j]);
        T[j + (size_t)i * n] = T[i + (size_t)j * n];
      }
    rotate(n, D + boff[q], T.data());
    write_one_body(b, q, T.data(), op);
  }
  transform_matrices(b, boff, std::vector<double>(D, D + boff[nbk]), moff, M);
  return "";
}

}  // namespace imsrg

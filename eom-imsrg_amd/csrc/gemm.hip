// Grouped FP64 MFMA GEMM for gfx950 (v_mfma_f64_16x16x4f64).
//
// One launch runs a list of tile tasks over a list of problems
//   C = alpha * (op(A1) op(B1) + s2 * op(A2) op(B2))  (+ beta C),
// operands addressed as (slot, offset) so a plan serves every call.  Tile
// shapes 64x64, 128x64, 128x128, 64x32, 32x64, 32x32 (256 threads = 4 waves
// in 2 x 2, each wave a (TBM/2) x (TBN/2) block of 16x16 MFMA tiles; the
// small shapes cut the edge padding of thin problems), K in steps of 16 through
// two LDS buffers with the next K-tile prefetched into registers; or (shapes
// 128x64 and 64x64 register-direct) each wave loading its own operands
// into registers with no LDS.  Split K:
// task.ks selects K-chunk [ks*kchunk, (ks+1)*kchunk) and slab C + ks*cstride.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"

namespace imsrg {

namespace {

#ifndef IMSRG_GEMM_BK
#define IMSRG_GEMM_BK 16
#endif
constexpr int GBK = IMSRG_GEMM_BK;   // K step through LDS

// per thread and K-tile: TBM*16/256 elements of op(A), TBN*16/256 of op(B)
template <int TBM, int TBN>
struct Regs {
  double a[TBM * GBK / 256];
  double b[TBN * GBK / 256];
};

// The transposes TA / TB are template parameters (four specialised bodies
// chosen once per workgroup), and out-of-range elements load a valid
// address and select zero, so the fetch has no per-element branches.
template <int TBM, int TBN, bool TA0, bool TB0, bool RT>
__device__ __forceinline__ void fetch(Regs<TBM, TBN>& t, const double* A, int lda, const double* Bm, int ldb, int M,
                                      int N, int K, int m0, int n0, int k0, double scale, bool ta, bool tb) {
  // RT: the transposes are the (workgroup-uniform) run-time ta / tb, one
  // code path for all four; else the template constants
  const bool TA = RT ? ta : TA0, TB = RT ? tb : TB0;
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < TBM * GBK / 256; ++r) {
    const int e = tid + 256 * r;
    const int i = TA ? e / GBK : e % TBM, kk = TA ? e % GBK : e / TBM;
    const int gi = m0 + i, gk = k0 + kk;
    const bool in = gi < M && gk < K;
    const int ci = in ? gi : 0, ck = in ? gk : 0;
    const double v = TA ? A[ck + (i64)ci * lda] : A[ci + (i64)ck * lda];
    t.a[r] = in ? v * scale : 0.0;
  }
#pragma unroll
  for (int r = 0; r < TBN * GBK / 256; ++r) {
    const int e = tid + 256 * r;
    const int j = TB ? e % TBN : e / GBK, kk = TB ? e / TBN : e % GBK;
    const int gj = n0 + j, gk = k0 + kk;
    const bool in = gj < N && gk < K;
    const int cj = in ? gj : 0, ck = in ? gk : 0;
    const double w = TB ? Bm[cj + (i64)ck * ldb] : Bm[ck + (i64)cj * ldb];
    t.b[r] = in ? w : 0.0;
  }
}

template <int TBM, int TBN, bool TA0, bool TB0, bool RT>
__device__ __forceinline__ void stash(const Regs<TBM, TBN>& t, double (*As)[TBM + 1], double (*Bs)[TBN + 1], bool ta,
                                      bool tb) {
  const bool TA = RT ? ta : TA0, TB = RT ? tb : TB0;
  const int tid = threadIdx.x;
#pragma unroll
  for (int r = 0; r < TBM * GBK / 256; ++r) {
    const int e = tid + 256 * r;
    if (!TA) As[e / TBM][e % TBM] = t.a[r];
    else As[e % GBK][e / GBK] = t.a[r];
  }
#pragma unroll
  for (int r = 0; r < TBN * GBK / 256; ++r) {
    const int e = tid + 256 * r;
    if (!TB) Bs[e % GBK][e / GBK] = t.b[r];
    else Bs[e / TBN][e % TBN] = t.b[r];
  }
}

// acc += op(A)[m0:m0+TBM, k0:k1] op(B)[k0:k1, n0:n0+TBN].  The MFMA takes
// op(B)^T as its A operand and op(A)^T as its B operand: register r of lane
// l of tile (mi, ni) holds C row wm*TBM/2 + mi*16 + l%16, column
// wn*TBN/2 + ni*16 + l/16 + 4r (16 consecutive rows per row group, so the
// column-major stores are contiguous).
template <int TBM, int TBN, bool TA, bool TB, bool RT = false>
__device__ __forceinline__ void mma_t(v4d* acc, double (*As)[GBK][TBM + 1], double (*Bs)[GBK][TBN + 1],
                                      const double* A, int lda, const double* Bm, int ldb, int M, int N, int k0s,
                                      int k1, int m0, int n0, double scale, bool ta = false, bool tb = false) {
  constexpr int MI = TBM / 32, NI = TBN / 32;   // acc[mi * NI + ni]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int lr = lane & 15, lk = lane >> 4;
  if (k0s >= k1) return;
  Regs<TBM, TBN> t;
  fetch<TBM, TBN, TA, TB, RT>(t, A, lda, Bm, ldb, M, N, k1, m0, n0, k0s, scale, ta, tb);
  stash<TBM, TBN, TA, TB, RT>(t, As[0], Bs[0], ta, tb);
  __syncthreads();
  int cur = 0;
  for (int k0 = k0s; k0 < k1; k0 += GBK) {
    const bool more = (k0 + GBK < k1);
    if (more) fetch<TBM, TBN, TA, TB, RT>(t, A, lda, Bm, ldb, M, N, k1, m0, n0, k0 + GBK, scale, ta, tb);
#pragma unroll
    for (int kk = 0; kk < GBK; kk += 4) {
      double a[MI], b[NI];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) a[mi] = As[cur][kk + lk][wm * (TBM / 2) + mi * 16 + lr];
#pragma unroll
      for (int ni = 0; ni < NI; ++ni) b[ni] = Bs[cur][kk + lk][wn * (TBN / 2) + ni * 16 + lr];
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
          acc[mi * NI + ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[ni], a[mi], acc[mi * NI + ni], 0, 0, 0);
    }
    if (more) stash<TBM, TBN, TA, TB, RT>(t, As[cur ^ 1], Bs[cur ^ 1], ta, tb);
    __syncthreads();
    cur ^= 1;
  }
}

// operand slot k (uniform): a select chain, so the slot table stays in
// scalar registers instead of an indexed copy in private memory
__device__ __forceinline__ const double* slot(const Slots& sl, int k, int inst) {
  const double* r = sl.p[0];
  i64 st = sl.st[0];
#pragma unroll
  for (int x = 1; x < 6; ++x)
    if (k == x) { r = sl.p[x]; st = sl.st[x]; }
  return r + (i64)inst * st;
}


// product P (1 or 2) of a kernel specialised on the transposes of its
// family (SIG bit0 TA1, bit1 TB1, bit2 TA2, bit3 TB2), or with the
// problem's own transposes read at run time (SIG < 0: one main loop whose
// index maps select on the uniform ta / tb).
template <int TBM, int TBN, int SIG, int P>
__device__ __forceinline__ void mma_p(v4d* acc, double (*As)[GBK][TBM + 1], double (*Bs)[GBK][TBN + 1],
                                      const double* A, int lda, int tA, const double* Bm, int ldb, int tB, int M, int N,
                                      int k0s, int k1, int m0, int n0, double scale) {
  if constexpr (SIG < 0) {
    mma_t<TBM, TBN, false, false, true>(acc, As, Bs, A, lda, Bm, ldb, M, N, k0s, k1, m0, n0, scale, tA != 0,
                                        tB != 0);
  } else {
    constexpr bool TA = (SIG >> (P == 1 ? 0 : 2)) & 1, TB = (SIG >> (P == 1 ? 1 : 3)) & 1;
    mma_t<TBM, TBN, TA, TB>(acc, As, Bs, A, lda, Bm, ldb, M, N, k0s, k1, m0, n0, scale);
  }
}

template <int TBM, int TBN, int SIG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    TBM * TBN <= 32 * 32 ? 8 : (TBM * TBN <= 64 * 64 ? 4 : (TBM * TBN <= 128 * 64 ? 2 : 1))))) void k_gemm_t(Slots sl, const GemmProb* probs, const GemmTask* tasks, XcdMap xm,
                                                double ascale, double s2scale) {
  constexpr int MI = TBM / 32, NI = TBN / 32;
  __shared__ double As[2][GBK][TBM + 1];
  __shared__ double Bs[2][GBK][TBN + 1];
  const int ti = xcd_task(xm, blockIdx.x);
  if (ti < 0) return;
  int inst;
  if (!batch_on(sl.act, inst)) return;
  const GemmTask t = tasks[ti];
  const GemmProb pr = probs[t.prob];
  const int m0 = t.tm * TBM, n0 = t.tn * TBN;
  const int kb = pr.kchunk > 0 ? t.ks * pr.kchunk : 0;
  const int ke = pr.kchunk > 0 ? min(pr.k, kb + pr.kchunk) : pr.k;
  v4d acc[MI * NI];
#pragma unroll
  for (int a = 0; a < MI * NI; ++a) acc[a] = v4d{0.0, 0.0, 0.0, 0.0};
  // mode bit3: A is upper triangular (as stored); rows m0.. of op(A) reach
  // only k >= m0 (A as stored) or k < m0 + TBM (A transposed)
  const bool tri = pr.mode & 8;
  mma_p<TBM, TBN, SIG, 1>(acc, As, Bs, slot(sl, pr.sa1, inst) + pr.a1, pr.lda, pr.ta1, slot(sl, pr.sb1, inst) + pr.b1,
                          pr.ldb, pr.tb1, pr.m, pr.n, (tri && !pr.ta1) ? max(kb, m0) : kb,
                          (tri && pr.ta1) ? min(ke, m0 + TBM) : ke, m0, n0, 1.0);
  if (pr.nprod == 2) {
    // second product: same K range, or its own K (k2 > 0) with its own
    // leading dimensions packed in a2/b2's problem (lda2 = ldb2 = m here)
    if (pr.k2 > 0)
      mma_p<TBM, TBN, SIG, 2>(acc, As, Bs, slot(sl, pr.sa2, inst) + pr.a2, pr.m, pr.ta2, slot(sl, pr.sb2, inst) + pr.b2,
                              pr.n, pr.tb2, pr.m, pr.n, 0, pr.k2, m0, n0, pr.s2 * ((pr.mode & 64) ? ascale : s2scale));
    else
      mma_p<TBM, TBN, SIG, 2>(acc, As, Bs, slot(sl, pr.sa2, inst) + pr.a2, pr.lda, pr.ta2, slot(sl, pr.sb2, inst) + pr.b2,
                              pr.ldb, pr.tb2, pr.m, pr.n, (tri && !pr.ta2) ? max(kb, m0) : kb,
                              (tri && pr.ta2) ? min(ke, m0 + TBM) : ke, m0, n0,
                              pr.s2 * ((pr.mode == 0 || (pr.mode & 16)) ? s2scale : 1.0));
  }
  double* C = const_cast<double*>(slot(sl, pr.sc, inst)) + pr.c + (i64)t.ks * pr.cstride;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const double al = pr.mode == 0 ? pr.alpha * ascale
                                  : pr.alpha * ((pr.mode & 1) ? ascale : 1.0) * ((pr.mode & 2) ? s2scale : 1.0);
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * (TBM / 2) + mi * 16 + (lane & 15);
        const int col = n0 + wn * (TBN / 2) + ni * 16 + (lane >> 4) + 4 * r;
        if (row < pr.m && col < pr.n) {
          const i64 o = row + (i64)col * pr.ldc;
          const double v = al * acc[mi * NI + ni][r];
          C[o] = (pr.beta == 0.0) ? v : v + pr.beta * C[o];
        }
      }
}

// ------------------------------------------------------------------------
// Register-direct variant (tile shapes 6 and 7): the four waves of a
// workgroup compute independent WM x WN blocks of a (2 WM) x (2 WN) tile,
// each loading its MFMA operands from global memory straight into
// registers, DS_STAGES chunks of 8 k in flight; no LDS and no barriers.
// In a chunk, MFMA g (g = 0, 1) takes k = 2 kq + g in lane group kq
// (kq = lane / 16), so an operand that is contiguous in k (op(A) with TA,
// op(B) without TB) is one 16-byte run per lane; an operand contiguous in
// its row / column index is 16 consecutive doubles per lane group.  Rows
// and columns beyond the problem read a valid row / column: their products
// reach only outputs that are not stored.
typedef double v2d __attribute__((ext_vector_type(2)));

// one product of a wave: acc += scale op(A)[m0:, k0:k1] op(B)[k0:k1, n0:];
// the transposes are uniform run-time values (strides of k and of the row /
// column), so both products of a problem run through one body
struct DProd {
  const double* A;
  const double* B;
  int lda, ldb, ta, tb, k0, k1;
  double scale;
};

// TT: 0..3 = (TA, TB) as compile-time constants, -1 = the problem's own
template <int WM, int WN, int TT>
__device__ __forceinline__ void mma_d(v4d* acc, const DProd& q, int M, int N, int m0, int n0) {
  constexpr int MI = WM / 16, NI = WN / 16, KC = 8, G = 2;
  constexpr int S = WM * WN >= 64 * 64 ? 2 : 3;   // chunks in flight
  const int k0 = q.k0, k1 = q.k1;
  if (k0 >= k1) return;
  const bool ta = TT < 0 ? q.ta != 0 : (TT >> 1) != 0, tb = TT < 0 ? q.tb != 0 : (TT & 1) != 0;
  const int lane = threadIdx.x & 63, lr = lane & 15, kq = lane >> 4;
  const double* A = q.A;
  const double* Bm = q.B;
  const double scale = q.scale;
  const i64 sa = ta ? 1 : q.lda, sb = tb ? q.ldb : 1;   // stride of k
  const i64 ra = ta ? q.lda : 1, rb = tb ? 1 : q.ldb;   // stride of the row / column
  const double* pa[MI];
  const double* pb[NI];
#pragma unroll
  for (int mi = 0; mi < MI; ++mi) pa[mi] = A + (i64)min(m0 + mi * 16 + lr, M - 1) * ra + (i64)(k0 + G * kq) * sa;
#pragma unroll
  for (int ni = 0; ni < NI; ++ni) pb[ni] = Bm + (i64)min(n0 + ni * 16 + lr, N - 1) * rb + (i64)(k0 + G * kq) * sb;
  // 16-byte loads where an operand is contiguous in k and every lane's run
  // is 16-byte aligned
  const bool va = ta && ((reinterpret_cast<uintptr_t>(A + k0) & 15) == 0) && ((q.lda & 1) == 0);
  const bool vb = !tb && ((reinterpret_cast<uintptr_t>(Bm + k0) & 15) == 0) && ((q.ldb & 1) == 0);
  double af[S][MI][G], bf[S][NI][G];
  auto load = [&](int st) {
#pragma unroll
    for (int mi = 0; mi < MI; ++mi) {
      if (va) {
        const v2d x = *reinterpret_cast<const v2d*>(pa[mi]);
        af[st][mi][0] = x[0] * scale;
        af[st][mi][1] = x[1] * scale;
      } else {
        af[st][mi][0] = pa[mi][0] * scale;
        af[st][mi][1] = pa[mi][sa] * scale;
      }
      pa[mi] += KC * sa;
    }
#pragma unroll
    for (int ni = 0; ni < NI; ++ni) {
      if (vb) {
        const v2d x = *reinterpret_cast<const v2d*>(pb[ni]);
        bf[st][ni][0] = x[0];
        bf[st][ni][1] = x[1];
      } else {
        bf[st][ni][0] = pb[ni][0];
        bf[st][ni][1] = pb[ni][sb];
      }
      pb[ni] += KC * sb;
    }
  };
  auto mma = [&](int st) {
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int mi = 0; mi < MI; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
          acc[mi * NI + ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[st][ni][g], af[st][mi][g], acc[mi * NI + ni],
                                                                   0, 0, 0);
  };
  const int nfull = (k1 - k0) / KC;
#pragma unroll
  for (int st = 0; st < S - 1; ++st)
    if (st < nfull) load(st);
  for (int c = 0; c < nfull; c += S) {
#pragma unroll
    for (int st = 0; st < S; ++st) {
      if (c + st + S - 1 < nfull) load((st + S - 1) % S);
      if (c + st < nfull) mma(st);
    }
  }
  if ((k1 - k0) % KC) {   // the last partial chunk: k >= k1 read at a valid k and zeroed
    const int kb = k0 + nfull * KC + G * kq;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const bool in = kb + g < k1;
      const i64 kc = min(kb + g, k1 - 1);
#pragma unroll
      for (int mi = 0; mi < MI; ++mi) {
        const double v = A[(i64)min(m0 + mi * 16 + lr, M - 1) * ra + kc * sa];
        af[0][mi][g] = in ? v * scale : 0.0;
      }
#pragma unroll
      for (int ni = 0; ni < NI; ++ni) {
        const double v = Bm[(i64)min(n0 + ni * 16 + lr, N - 1) * rb + kc * sb];
        bf[0][ni][g] = in ? v : 0.0;
      }
    }
    mma(0);
  }
}

// SIG as for k_gemm_t: the transposes of the family's two products as
// constants (bit0 TA1, bit1 TB1, bit2 TA2, bit3 TB2), or SIG < 0: each
// problem's own, read at run time through one body for both products
template <int WM, int WN, int SIG>
__global__ __launch_bounds__(256) void k_gemm_d(Slots sl, const GemmProb* probs, const GemmTask* tasks, XcdMap xm,
                                                double ascale, double s2scale) {
  constexpr int MI = WM / 16, NI = WN / 16;
  const int ti = xcd_task(xm, blockIdx.x);
  if (ti < 0) return;
  int inst;
  if (!batch_on(sl.act, inst)) return;
  const GemmTask t = tasks[ti];
  const GemmProb pr = probs[t.prob];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = t.tm * (2 * WM) + (wave >> 1) * WM, n0 = t.tn * (2 * WN) + (wave & 1) * WN;
  if (m0 >= pr.m || n0 >= pr.n) return;   // the waves are independent (no barrier below)
  const int kb = pr.kchunk > 0 ? t.ks * pr.kchunk : 0;
  const int ke = pr.kchunk > 0 ? min(pr.k, kb + pr.kchunk) : pr.k;
  v4d acc[MI * NI];
#pragma unroll
  for (int a = 0; a < MI * NI; ++a) acc[a] = v4d{0.0, 0.0, 0.0, 0.0};
  // mode bit3: A is upper triangular (as stored, zero below its diagonal):
  // rows m0.. of op(A) reach only k >= m0 (A as stored) or k < m0 + WM (A
  // transposed)
  const bool tri = pr.mode & 8;
  DProd p1, p2;
  p1.A = slot(sl, pr.sa1, inst) + pr.a1; p1.B = slot(sl, pr.sb1, inst) + pr.b1;
  p1.lda = pr.lda; p1.ldb = pr.ldb; p1.ta = pr.ta1; p1.tb = pr.tb1;
  p1.k0 = (tri && !pr.ta1) ? max(kb, m0) : kb;
  p1.k1 = (tri && pr.ta1) ? min(ke, m0 + WM) : ke;
  p1.scale = 1.0;
  p2.A = slot(sl, pr.sa2, inst) + pr.a2; p2.B = slot(sl, pr.sb2, inst) + pr.b2;
  p2.ta = pr.ta2; p2.tb = pr.tb2;
  if (pr.k2 > 0) {
    // second product with its own K and leading dimensions m, n
    p2.lda = pr.m; p2.ldb = pr.n; p2.k0 = 0; p2.k1 = pr.k2; p2.scale = pr.s2 * ((pr.mode & 64) ? ascale : s2scale);
  } else {
    p2.lda = pr.lda; p2.ldb = pr.ldb;
    p2.k0 = (tri && !pr.ta2) ? max(kb, m0) : kb;
    p2.k1 = (tri && pr.ta2) ? min(ke, m0 + WM) : ke;
    p2.scale = pr.s2 * ((pr.mode == 0 || (pr.mode & 16)) ? s2scale : 1.0);
  }
  if constexpr (SIG < 0) {
    for (int p = 0; p < pr.nprod; ++p) mma_d<WM, WN, -1>(acc, p == 0 ? p1 : p2, pr.m, pr.n, m0, n0);
  } else {
    mma_d<WM, WN, ((SIG & 1) << 1) | ((SIG >> 1) & 1)>(acc, p1, pr.m, pr.n, m0, n0);
    if (pr.nprod == 2) mma_d<WM, WN, (((SIG >> 2) & 1) << 1) | ((SIG >> 3) & 1)>(acc, p2, pr.m, pr.n, m0, n0);
  }
  double* C = const_cast<double*>(slot(sl, pr.sc, inst)) + pr.c + (i64)t.ks * pr.cstride;
  const double al = pr.mode == 0 ? pr.alpha * ascale
                                  : pr.alpha * ((pr.mode & 1) ? ascale : 1.0) * ((pr.mode & 2) ? s2scale : 1.0);
#pragma unroll
  for (int mi = 0; mi < MI; ++mi)
#pragma unroll
    for (int ni = 0; ni < NI; ++ni)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + mi * 16 + (lane & 15);
        const int col = n0 + ni * 16 + (lane >> 4) + 4 * r;
        if (row < pr.m && col < pr.n) {
          const i64 o = row + (i64)col * pr.ldc;
          const double v = al * acc[mi * NI + ni][r];
          C[o] = (pr.beta == 0.0) ? v : v + pr.beta * C[o];
        }
      }
}

}  // namespace

// tile shapes: 0 64x64, 1 128x64, 2 128x128, 3 64x32, 4 32x64, 5 32x32
// (LDS-staged); 6 128x64, 7 64x64 (register-direct, per wave 64x32 /
// 32x32)
int tile_rows(int tile) { return (tile == 1 || tile == 2 || tile == 6) ? 128 : (tile == 4 || tile == 5 ? 32 : 64); }
int tile_cols(int tile) { return tile == 2 ? 128 : ((tile == 3 || tile == 5) ? 32 : 64); }

template <int TBM, int TBN>
static void launch_sig(int sig, int grid, int ninst, const Slots& sl, const GemmProb* probs, const GemmTask* tasks,
                const XcdMap& xm, hipStream_t s, double ascale, double s2scale) {
  switch (sig) {
    case 0:
      hipLaunchKernelGGL((k_gemm_t<TBM, TBN, 0>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
    case 1:
      hipLaunchKernelGGL((k_gemm_t<TBM, TBN, 1>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
    case 8:
      hipLaunchKernelGGL((k_gemm_t<TBM, TBN, 8>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
    case 10:
      hipLaunchKernelGGL((k_gemm_t<TBM, TBN, 10>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
    default:
      hipLaunchKernelGGL((k_gemm_t<TBM, TBN, -1>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
  }
}

template <int WM, int WN>
static void launch_d(int sig, int grid, int ninst, const Slots& sl, const GemmProb* probs, const GemmTask* tasks,
                     const XcdMap& xm, hipStream_t s, double ascale, double s2scale) {
  switch (sig) {
    case 0:
      hipLaunchKernelGGL((k_gemm_d<WM, WN, 0>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
    case 1:
      hipLaunchKernelGGL((k_gemm_d<WM, WN, 1>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
    case 2:
      hipLaunchKernelGGL((k_gemm_d<WM, WN, 2>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
    case 3:
      hipLaunchKernelGGL((k_gemm_d<WM, WN, 3>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
    case 4:
      hipLaunchKernelGGL((k_gemm_d<WM, WN, 4>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
    case 8:
      hipLaunchKernelGGL((k_gemm_d<WM, WN, 8>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
    case 10:
      hipLaunchKernelGGL((k_gemm_d<WM, WN, 10>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
    default:
      hipLaunchKernelGGL((k_gemm_d<WM, WN, -1>), dim3(grid, ninst), dim3(256), 0, s, sl, probs, tasks, xm, ascale,
                         s2scale);
      break;
  }
}

// the signatures with specialised kernels (the LDS-staged shapes run 2, 3
// and 4 on their run-time path)
int gemm_sig_supported(int sig) { return (sig >= 0 && sig <= 4) || sig == 8 || sig == 10; }

void launch_gemm_tiled(int tile, const Slots& sl, const GemmProb* probs, const GemmTask* tasks, int grid,
                       const XcdMap& xm, hipStream_t s, double ascale, double s2scale, int ninst, int sig) {
  if (grid <= 0) return;
  if (!gemm_sig_supported(sig)) sig = -1;
  switch (tile) {
    case 0: launch_sig<64, 64>(sig, grid, ninst, sl, probs, tasks, xm, s, ascale, s2scale); break;
    case 1: launch_sig<128, 64>(sig, grid, ninst, sl, probs, tasks, xm, s, ascale, s2scale); break;
    case 3: launch_sig<64, 32>(sig, grid, ninst, sl, probs, tasks, xm, s, ascale, s2scale); break;
    case 4: launch_sig<32, 64>(sig, grid, ninst, sl, probs, tasks, xm, s, ascale, s2scale); break;
    case 5: launch_sig<32, 32>(sig, grid, ninst, sl, probs, tasks, xm, s, ascale, s2scale); break;
    case 6: launch_d<64, 32>(sig, grid, ninst, sl, probs, tasks, xm, s, ascale, s2scale); break;
    case 7: launch_d<32, 32>(sig, grid, ninst, sl, probs, tasks, xm, s, ascale, s2scale); break;
    default: launch_sig<128, 128>(sig, grid, ninst, sl, probs, tasks, xm, s, ascale, s2scale); break;
  }
}

}  // namespace imsrg

// Grouped FP64 MFMA GEMM for gfx950 (v_mfma_f64_16x16x4f64).
//
// One launch runs a list of tile tasks over a list of problems
//   C = alpha * (op(A1) op(B1) + s2 * op(A2) op(B2))  (+ beta C),
// operands addressed as (slot, offset) so a plan serves every call.  Tile
// shapes 64x64, 128x64, 128x128, 64x32, 32x64, 32x32 (256 threads = 4 waves
// in 2 x 2, each wave a (TBM/2) x (TBN/2) block of 16x16 MFMA tiles; the
// small shapes cut the edge padding of thin problems), K in steps of 16 through
// two LDS buffers with the next K-tile prefetched into registers.  Split K:
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
                              pr.n, pr.tb2, pr.m, pr.n, 0, pr.k2, m0, n0, pr.s2 * s2scale);
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

}  // namespace

// tile shapes: 0 64x64, 1 128x64, 2 128x128, 3 64x32, 4 32x64, 5 32x32
int tile_rows(int tile) { return (tile == 1 || tile == 2) ? 128 : (tile >= 4 ? 32 : 64); }
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

int gemm_sig_supported(int sig) { return sig == 0 || sig == 1 || sig == 8 || sig == 10; }

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
    default: launch_sig<128, 128>(sig, grid, ninst, sl, probs, tasks, xm, s, ascale, s2scale); break;
  }
}

}  // namespace imsrg

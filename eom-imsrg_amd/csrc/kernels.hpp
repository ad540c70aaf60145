// Work descriptors and kernel launchers of the scalar-commutator path.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "basis.hpp"

namespace imsrg {

constexpr int kTile = 256;   // work items per workgroup of the element kernels

// A tile of up to kTile elements of block g of channel q; `cnt` is the
// number of elements of the (triangular or full) segment.
struct SegTask { int q; int g; int start; int cnt; };

// One-body output (kind 0 hh, 1 pp, 2 ph): orbits p, q (p <= q for hh, pp)
// plus their indices in fhh / fpp / fph.
struct OneTask { int kind; int pk; int qk; int pi; int qi; };

// One two-body output element named by its orbits (i | j<<8 | k<<16 | l<<24,
// pairs i<=j, k<=l) and its 2J.
struct QuadItem { int ijkl; int J; };

// A pair reached from a stored pair by replacing one orbit inside its
// (l, j, tz) block (the 1b x 2b sums of commutator_122).  code: index in
// its class list (bits 0-21) | class << 22 | f stored transposed << 24 |
// negative phase << 25 | sqrt(2) factor << 26.  foff: offset of the
// one-body element in the operator.
// fk: compact index of the one-body element f(x, i) the term reads
// (operator offset f1off[fk]); the kernels stage those few elements in LDS
struct Partner { int code; int fk; };
constexpr int kF1Max = 2048;   // compact one-body elements of the 122 partner sums

// Tile of up to kTile CCX elements of cross-coupled block q1
struct CCTask { int q1; int start; };

// Operand table of a grouped GEMM launch: problems address operands as
// (slot, offset) so one problem list serves every call.
// operand slots of a GEMM launch; st[x] is slot x's stride between the
// instances of a batched launch (grid.y), act the instance mask (null: all)
struct Slots {
  const double* p[6];
  i64 st[6];
  const int* act;
};

// C(slotC) = alpha * (op(A1) op(B1) + s2 * op(A2) op(B2)) + beta * C
struct GemmProb {
  i64 a1, b1, a2, b2, c;
  int m, n, k;
  int lda, ldb, ldc;
  int nprod;
  unsigned char ta1, tb1, ta2, tb2;
  unsigned char sa1, sb1, sa2, sb2;
  unsigned char sc, mode, pad1, pad2;   // mode 0: alpha*ascale, s2*s2scale; else alpha is
                                        // scaled by ascale (bit0) and s2scale (bit1), s2 by
                                        // s2scale (bit4); bit3: A upper triangular;
                                        // bit6: a second product with its own K
                                        // (k2 > 0) scaled by ascale, not s2scale
  double alpha, s2, beta;
  int kchunk;        // > 0: split K in chunks of kchunk, chunk ks written to C + ks*cstride
  int k2;            // > 0: K of the second product when it differs (no split K then)
  i64 cstride;
};
struct GemmTask { int prob; int tm; int tn; int ks; };

// Tile pair (I,J), I <= J, of the pppp block of channel q
struct PairTask { int q; int ti; int tj; int pad; };

// 222_ph on the grouped path: a chunk of consecutive orbit quadruples of
// the output items (items item0 .. item0 + nitems, at most kPhChunkItems;
// their J' terms, at most kPhChunkLanes, are gathered once into LDS)
constexpr int kPhChunkItems = 256;
constexpr int kPhChunkLanes = 1024;
constexpr int kPhMaxCC = 256;   // cross-coupled blocks whose offsets the chunk kernel keeps in LDS
struct PhChunk { int q0, nq, item0, nitems, nlanes, pad0, pad1, pad2; };

// One contiguous run of the vectorize stream (basic_IMSRG.f90:3929-3981):
// len operator elements from op_off map to vector positions vec_off...; for
// a column of a square block (mirror = its dimension n) repackage also
// writes the transposed partners op_off + x * n, times herm.
struct VecTask { i64 op_off; i64 vec_off; int len; int mirror; };

// Workgroups are dispatched to the 8 XCDs round-robin (blockIdx mod 8), each
// XCD with its own L2.  An XcdMap gives every XCD its own contiguous task
// range so the tiles of one channel share an L2: block b runs task
// start[b % 8] + b / 8 (if below count).  count[0] < 0: identity mapping.
constexpr int kXcds = 8;
struct XcdMap { int start[kXcds]; int count[kXcds]; };

// Workgroups are dispatched to the XCDs round robin (block b on XCD b % 8);
// this maps them so that XCD x runs one contiguous range of the n tasks, and
// neighbouring tasks (which read neighbouring data) share that XCD's L2.
// A bijection of [0, n).
__device__ __forceinline__ int xcd_contig(int b, int n) {
  const int x = b & (kXcds - 1), k = b / kXcds, fl = n / kXcds, rem = n % kXcds;
  return x * fl + min(x, rem) + k;
}

__device__ __forceinline__ int xcd_task(const XcdMap& m, int block) {
  if (m.count[0] < 0) return block;
  const int x = block & (kXcds - 1), loc = block / kXcds;
  return loc < m.count[x] ? m.start[x] + loc : -1;
}

// Task list on the device with per-group ranges (the plain path launches
// one range at a time; the grouped path launches everything at once).
template <class T>
struct TaskList {
  T* dev = nullptr;
  int n = 0;
  std::vector<T> host;
  std::vector<int> chan_begin;   // [groups+1] (groups in list order)
  XcdMap xmap{{0}, {-1}};        // grouped launch mapping
  int grid = 0;                  // grouped launch size
};

// launchers (all asynchronous on `s`)
void launch_comm222ph_scatter(const DevBasis& B, const SegTask* t, int nt, const double* W, int Lh, int Rh,
                              double* RES, int RESh, hipStream_t s);
// square blocks' lower triangles from their upper triangles (herm)
void launch_sym_lower(const DevBasis& B, const PairTask* t, int nt, double* RES, int herm, hipStream_t s);
// 222_ph scatter by quadruple chunks: fq packed ijkl per quadruple, fql its
// first J' term within its chunk, fqi its first item
void launch_comm222ph_chunk(const DevBasis& B, const PhChunk* ch, int nch, const int* fq, const int* fql,
                            const int* fqi, const QuadItem* items, const double* W, int Lh, int Rh, double* RES,
                            int RESh, hipStream_t s, int mirror);
// dir 0: vec <- op (vectorize); dir 1: op <- vec (repackage, herm fill)
void launch_vectorize(const VecTask* t, int nt, double* op, double* vec, int dir, int herm, hipStream_t s);
// transform_2b_to_HF: gam blocks <-> per-channel square matrices (hh, ph, pp)
void launch_hf_pack(const DevBasis& B, const SegTask* t, int nt, const double* op, double* V, const i64* voff,
                    hipStream_t s);
void launch_hf_unpack(const DevBasis& B, const SegTask* t, int nt, const double* V, const i64* voff, double* op,
                      hipStream_t s);
// part: the partner lists (CSR rowp / colp); partT: the row lists with
// entry k of pair gid at k * pstride + gid
void launch_comm122p(const DevBasis& B, const SegTask* t, int nt, const Partner* part, const Partner* partT,
                     int pstride, const int* rowp, const int* colp, const int* f1off, int nf1, const double* L, int Lh,
                     const double* R, int Rh, double* RES, int RESh, hipStream_t s, int mirror = 1);
void launch_pandya(const DevBasis& B, const CCTask* t, int nt, const double* op, int herm, double* CC,
                   hipStream_t s);
void launch_comm111(const DevBasis& B, const double* L, int Lh, const double* R, int Rh, double* RES,
                    hipStream_t s);
void launch_comm121(const DevBasis& B, const OneTask* t, int nt, const double* L, int Lh, const double* R,
                    int Rh, double* RES, int RESh, hipStream_t s);
void launch_comm221(const DevBasis& B, const OneTask* t, int nt, const double* L, int Lh, const double* R, int Rh,
                    const double* w1, const double* w2, double* RES, int RESh, hipStream_t s);
void launch_e0_partials(const DevBasis& B, const double* L, int Lh, const double* R, int Rh, double* part,
                        hipStream_t s);
void launch_e0_final(const double* part, int nblocks, int use220, int use110, double* dst, int accumulate,
                     hipStream_t s);
// grouped MFMA GEMM (gemm.hip); tile 0: 64x64, 1: 128x64, 2: 128x128.
// alpha and s2 of every problem are multiplied by ascale / s2scale.
int tile_rows(int tile);
int tile_cols(int tile);
// sig: the family's transposes (bit0 TA1, bit1 TB1, bit2 TA2, bit3 TB2) when
// every problem shares them and a kernel is specialised for them
// (gemm_sig_supported), else -1 (each problem's own, chosen at run time)
void launch_gemm_tiled(int tile, const Slots& sl, const GemmProb* probs, const GemmTask* tasks, int grid,
                       const XcdMap& xm, hipStream_t s, double ascale = 1.0, double s2scale = 1.0, int ninst = 1,
                       int sig = -1);
int gemm_sig_supported(int sig);
// RES1 += C - pm C^T with C = L1 R1 - Rh L3 R3^T held in w1's pppp blocks;
// ref_outputs 1: also leave the reference's W1 = L1 R1 and W2 = Rh L3 R3^T
// in w1 / w2 (otherwise w1's pppp blocks keep C and w2's are untouched);
// 2: only the upper tiles of RES1 (the caller fills the lower triangles)
void launch_pppp_sym(const DevBasis& B, const PairTask* t, int nt, const double* L, int Lh, const double* R, int Rh,
                     double* RES, double* w1, double* w2, int ref_outputs, hipStream_t s);
void launch_pphh_epilogue(const DevBasis& B, const SegTask* t, int nt, const double* L, int Lh, const double* R,
                          int Rh, double* RES, double* w1, double* w2, const double* Pscr, const i64* poff,
                          const int* k3ks, const double* K2scr, const i64* k2off, const int* k2ks, hipStream_t s);
void launch_axpy(i64 n, const double* A, double ax, const double* B, double bx, double* C, hipStream_t s);
// many-instance forms: B.ninst instances back to back, masked by B.act
// (axpy / fill / copy: stride n; norms: operand stride oplen, partials
// stride spart, results stride sdst; e0_final: partials spart, dst oplen)
void launch_axpy_b(i64 n, const double* A, double ax, const double* Bv, double bx, double* C, int ninst,
                   const int* act, hipStream_t s);
void launch_fill_b(double* p, i64 n, double v, int ninst, const int* act, hipStream_t s);
void launch_copy_b(double* dst, const double* src, i64 n, int ninst, const int* act, hipStream_t s);
void launch_axpy_norm_b(const DevBasis& B, double* X, const double* T, double c, const i64* chunk_off,
                        const int* chunk_len, const double* chunk_w, int nchunk, double* part, hipStream_t s);
void launch_norm_partials_b(const DevBasis& B, const double* op, const i64* chunk_off, const int* chunk_len,
                            const double* chunk_w, int nchunk, double* part, hipStream_t s);
void launch_sum_sqrt_b(const DevBasis& B, const double* part, int n, double* dst, i64 sdst, int do_sqrt,
                       hipStream_t s);
void launch_e0_final_b(const DevBasis& B, const double* part, int use220, int use110, double* dst, int accumulate,
                       hipStream_t s);
void launch_monopole(const DevBasis& B, const double* H, double* M, hipStream_t s);
// generators: denominators of build_gs_white from Hd, numerators from Hn;
// mode 0 Hn / Eden (white), 1 0.5 atan(2 Hn / Eden) (atan, atanxx)
void launch_gen_1b(const DevBasis& B, const double* Hd, const double* Hn, double* ETA, int mode, hipStream_t s);
void launch_gen_2b(const DevBasis& B, const SegTask* t, int nt, const double* Hd, const double* M, const double* Hn,
                   double* ETA, int mode, hipStream_t s);
void launch_fill_e0(const DevBasis& B, double* op, double v, hipStream_t s);
void launch_mbpt2_partials(const DevBasis& B, const SegTask* t, int nt, const double* H, double* part,
                           hipStream_t s);
void launch_normal_order(const DevBasis& B, double* H, double* scratch, hipStream_t s);
void launch_fill(double* p, i64 n, double v, hipStream_t s);
void launch_copy(double* dst, const double* src, i64 n, hipStream_t s);

// reverse-mode kernels (kernels_adj.hip)
void launch_e0_vjp(const DevBasis& B, const SegTask* g3, int ng3, const double* L, int Lh, const double* R, int Rh,
                   const double* RESb, double* Lb, double* Rb, hipStream_t s);
void launch_111_vjp(const DevBasis& B, const double* L, int Lh, const double* R, int Rh, const double* RESb,
                    double* Lb, double* Rb, hipStream_t s);
void launch_121_vjp(const DevBasis& B, const OneTask* t, int nt, const double* L, int Lh, const double* R, int Rh,
                    const double* RESb, int RESh, double* Lb, double* Rb, hipStream_t s);
void launch_221_vjp(const DevBasis& B, const OneTask* t, int nt, int Lh, int Rh, const double* RESb, int RESh,
                    double* Wb1, double* Wb2, hipStream_t s);
void launch_pphh_wbar(const DevBasis& B, const SegTask* t, int nt, const double* RESb, int pm, double* Wb1,
                      double* Wb2, hipStream_t s);
void launch_to_dense(const DevBasis& B, const SegTask* t, int nt, const double* X, int herm, double* cbar,
                     hipStream_t s);
void launch_ph_vjp(const DevBasis& B, const QuadItem* items, int nitems, const int* lanes, int nl, const int* quads,
                   const int* qstart, const PairTask* tiles, int ntiles, const int* tstart, const int2* terms,
                   const double* RESb, int RESh, int Lh, int Rh, double* SB, double* CV, double* U, hipStream_t s);
// both 122 transposes (two-body and one-body) in one pass
void launch_122_vjp(const DevBasis& B, const SegTask* t, int nt, const Partner* part, const Partner* partT,
                    int pstride, const int* rowp, const int* colp, const int* f1off, int nf1, const double* L, int Lh,
                    const double* R, int Rh, const double* cbar, const double* RESb, int RESh, double* Lb, double* Rb,
                    hipStream_t s);
// the same transpose gathered per stored element over tasks of the blocks
// a Pandya term reads (no atomics)
void launch_pandya_vjp_gather(const DevBasis& B, const SegTask* t, int nt, const double* CCb, int herm, double* opb,
                              hipStream_t s);
void launch_gen_vjp(const DevBasis& B, const SegTask* g3, int ng3, const double* Hd, const double* Hn,
                    const double* M, const double* ETAb, double* Hdb, double* Hnb, double* Mb, int mode,
                    hipStream_t s);
void launch_normal_order_vjp(const DevBasis& B, const double* Hb, double* Tb, hipStream_t s);
void launch_add_e0(double* dst, const double* src, double factor, hipStream_t s);

}  // namespace imsrg

// C-ABI, device plan and nested-commutator drivers of the scalar-commutator
// path.  The drivers restate the control flow of IMSRG_MAGNUS.f90
// (BCH_EXPAND 149-234, MAGNUS_EXPAND 620-691, euler_step 754-764,
// magnus_decouple 15-128) on device-resident operators.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "basis.hpp"
#include "imsrg_hip.h"
#include "hf.hpp"
#include "interaction.hpp"
#include "kernels.hpp"

using namespace imsrg;

namespace {
thread_local std::string g_err;
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
// Argument check of the operator entry points: every pointer set, every
// hermiticity flag +1 (hermitian) or -1 (anti-hermitian).
int check_args(std::initializer_list<const void*> ptrs, std::initializer_list<int> herms) {
  for (const void* p : ptrs)
    if (!p) return fail(IMSRG_EINVAL, "null argument");
  for (int h : herms)
    if (h != 1 && h != -1) return fail(IMSRG_EINVAL, "herm must be +1 or -1");
  return 0;
}
}  // namespace


#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail(IMSRG_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// One family of grouped GEMM problems (one launch on the grouped path, one
// per channel / cross-coupled block on the plain path) and its tile tasks.
struct GemmFamily {
  int tile = 0;                   // 0: 64x64, 1: 128x64, 2: 128x128
  bool upper = false;             // only tiles with tm <= tn (symmetric results)
  std::vector<GemmProb> probs_h;
  std::vector<int> nks;           // split-K chunks per problem
  std::vector<int> group_end;     // problems of group g: [group_end[g-1], group_end[g])
  GemmProb* probs = nullptr;
  TaskList<GemmTask> tasks;
  GemmTask* tasks_alloc = nullptr;
  bool xcd_aware = false;         // pack groups onto XCDs (else one largest-first list)
  int sig = -1;                   // shared transposes of the problems (launch_gemm_tiled), -1 mixed
};

// the transposes every problem of a family shares, as launch_gemm_tiled's
// sig, or -1 when they differ or no kernel is specialised for them
int family_sig(const std::vector<GemmProb>& probs) {
  int sig = -2;
  for (const GemmProb& p : probs) {
    const int s = (p.ta1 ? 1 : 0) | (p.tb1 ? 2 : 0) | (p.nprod == 2 ? ((p.ta2 ? 4 : 0) | (p.tb2 ? 8 : 0)) : -1);
    if (s < 0) {   // one product: any second-product bits are compatible
      const int s1 = (p.ta1 ? 1 : 0) | (p.tb1 ? 2 : 0);
      if (sig == -2) sig = s1 | 256;   // 256: second product not yet fixed
      else if ((sig & 3) != s1) return -1;
      continue;
    }
    if (sig == -2) sig = s;
    else if (sig & 256) {
      if ((sig & 3) != (s & 3)) return -1;
      sig = s;
    } else if (sig != s) return -1;
  }
  if (sig == -2) return -1;
  sig &= 255;
  return gemm_sig_supported(sig) ? sig : -1;
}

struct imsrg_basis_s {
  Basis b;
  int path = IMSRG_PATH_GROUPED;
  bool dev_ready = false;
  // GEMM tile shapes for pp1, k2, k3, ph (64x64 until tuned) and the reverse-mode
  // 222_pp_hh rounds (32x32 through LDS), W' transpose (register-direct 128x64)
  // and fused pp products (64x32 through LDS, XCD packing): the fastest
  // commutator transpose at eMax 8 (1.54 ms against 1.62 ms with 64x32, 64x32
  // and register-direct 128x64)
  int tile_req[7] = {0, 0, 0, 0, 5, 6, 11};
  int plan_ver = 0;                 // bumped whenever a task list is rebuilt (captured graphs go stale)
  DevBasis D{};
  std::vector<void*> allocs;
  // task lists (grouped by channel / cross-coupled block)
  TaskList<SegTask> seg_tri, seg_full, seg_g3, seg_epi;
  TaskList<SegTask> seg_pv;         // reverse mode: the blocks a Pandya term reads (gam 2, 3, 4, 6)
  TaskList<CCTask> cc;
  TaskList<PairTask> pairs;
  TaskList<PairTask> symt;          // 64 x 64 tile pairs of every square block (pad = g)
  GemmFamily pp1, k2, k3, ph;       // pppp L1 R1, pp x (ph|hh), (ph|hh)^2, cross-coupled W'
  // Hartree-Fock two-body transform: T = V M, then V = M^T T per channel
  // (built on first use), with the per-channel square matrices V, T, M
  GemmFamily hf1, hf2;
  double* hf_buf = nullptr;        // [3 * hf_len]: V, T, M
  i64* hf_off = nullptr;           // [nblocks + 1] channel offsets
  i64 hf_len = 0;
  std::vector<Partner> part_h;      // 122 partner lists (CSR by global pair id)
  std::vector<int> rowp_h, colp_h;
  Partner* partners = nullptr;
  // the row lists again, entry k of pair gid at k * npairs + gid: the
  // neighbouring rows that one wave's lanes hold read neighbouring records
  std::vector<Partner> partT_h;
  Partner* partT = nullptr;
  int pstride = 0;
  int *rowp = nullptr, *colp = nullptr;
  std::vector<QuadItem> quad_h;   // 222_ph outputs, quadruple-major (grouped path)
  QuadItem* quad = nullptr;
  // the same outputs in chunks of quadruples (k_comm222ph_chunk)
  std::vector<PhChunk> phc_h;
  std::vector<int> fq_h, fql_h, fqi_h;
  PhChunk* phc = nullptr;
  int *fq = nullptr, *fql = nullptr, *fqi = nullptr;
  std::vector<OneTask> one_h;
  OneTask* one = nullptr;
  // norm chunks
  std::vector<i64> ch_off_h;
  std::vector<int> ch_len_h;
  std::vector<double> ch_w_h;
  i64* ch_off = nullptr;
  int* ch_len = nullptr;
  double* ch_w = nullptr;
  // scratch
  std::vector<i64> poff_h, k2off_h;   // split-K partial slabs of K3 (P) and K2, per channel
  std::vector<int> k3ks_h, k2ks_h;
  i64* poff = nullptr;
  i64* k2off = nullptr;
  int *k3ks = nullptr, *k2ks = nullptr;
  double* Pscr = nullptr;
  double* K2scr = nullptr;
  double* part = nullptr;
  int npart = 0;
  double* M = nullptr;       // monopole table N x N
  double* dval = nullptr;    // device scalars
  double* hval = nullptr;    // pinned host scalars
  // reverse mode, built on first use: 222_pp_hh transposed products in
  // rounds of distinct outputs, the W' transpose products, quadruple
  // starts, the compact one-body map of 122
  bool adj_ready = false;
  GemmFamily pb[6], ppb, phb;   // reverse-mode 222_pp_hh rounds, the fused pp products, the W' transpose
  std::vector<int> f1map_h, f1off_h;   // 122: one-body offset -> compact index, and back
  int* f1off = nullptr;
  int nf1 = 0;
  // 222_ph transpose: quadruples (packed ijkl) with the start of their J
  // items, one lane per (quadruple, recoupling, J'), the upper 64 x 64 tiles
  // of the W' blocks and the (lane, position) terms of each tile
  std::vector<int> phq_h, phqs_h, phl_h, pht0_h;
  std::vector<PairTask> phtile_h;
  std::vector<int2> phterm_h;
  int *phq = nullptr, *phqs = nullptr, *phl = nullptr, *pht0 = nullptr;
  PairTask* phtile = nullptr;
  int2* phterm = nullptr;
  double *SB = nullptr, *CV = nullptr;
  double *Wb1 = nullptr, *Wb2 = nullptr, *cbar = nullptr, *Y = nullptr, *U = nullptr;
  double *aLCC = nullptr, *aRCC = nullptr, *aLCCb = nullptr, *aRCCb = nullptr, *Mb = nullptr;
  // vectorize stream: runs, length, device staging vector (built on first use)
  std::vector<VecTask> vec_h;
  VecTask* vec = nullptr;
  double* vecbuf = nullptr;
  i64 neq = 0;

  // concurrent term schedule of the grouped commutator: three auxiliary
  // streams (default priority) and the events that fork and join them
  bool term_streams = true;
  bool aux_ready = false;
  hipStream_t aux[3] = {nullptr, nullptr, nullptr};
  hipEvent_t tev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};

  ~imsrg_basis_s() {
    for (hipEvent_t e : tev)
      if (e) (void)hipEventDestroy(e);
    for (hipStream_t x : aux)
      if (x) (void)hipStreamDestroy(x);
    for (void* p : allocs) (void)hipFree(p);
    if (hval) (void)hipHostFree(hval);
  }
};

namespace {

template <class T>
int upload(imsrg_basis_s* B, const std::vector<T>& v, T** out) {
  *out = nullptr;
  if (v.empty()) return 0;
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, v.size() * sizeof(T)));
  B->allocs.push_back(p);
  HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = (T*)p;
  return 0;
}

int dalloc(imsrg_basis_s* B, size_t bytes, void** out) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<size_t>(bytes, 8)));
  B->allocs.push_back(p);
  *out = p;
  return 0;
}

// channels in decreasing pppp cost, so a grouped launch issues its largest
// workgroups first (longest-processing-time order)
std::vector<int> channel_order(const Basis& b) {
  std::vector<int> o(b.nblocks);
  for (int q = 0; q < b.nblocks; ++q) o[q] = q;
  std::stable_sort(o.begin(), o.end(), [&](int x, int y) {
    const i64 cx = (i64)b.dims[x * 3] + b.dims[x * 3 + 1], cy = (i64)b.dims[y * 3] + b.dims[y * 3 + 1];
    return cx > cy;
  });
  return o;
}

enum SegMode { SEG_TRI, SEG_FULL, SEG_G3, SEG_EPI };

void build_seg(const Basis& b, TaskList<SegTask>& tl, SegMode mode, const std::vector<int>& order) {
  tl.host.clear();
  tl.chan_begin.assign(1, 0);
  for (int q : order) {
    const int np = b.dims[q * 3], nb = b.dims[q * 3 + 1], nh = b.dims[q * 3 + 2];
    for (int g = 0; g < 6; ++g) {
      i64 cnt = 0;
      const int n1 = b.rows(q, g), n2 = b.cols(q, g);
      switch (mode) {
        case SEG_TRI: cnt = kSquare[g] ? (i64)n1 * (n1 + 1) / 2 : (i64)n1 * n2; break;
        case SEG_FULL: cnt = (i64)n1 * n2; break;
        case SEG_G3: cnt = (g == 2) ? (i64)n1 * n2 : 0; break;
        case SEG_EPI:
          if (g == 1) cnt = (i64)np * (nb + nh);
          else if (g == 3) cnt = (i64)nb * nb;
          else if (g == 4) cnt = (i64)nh * nh;
          else if (g == 5) cnt = (i64)nb * nh;
          break;
      }
      for (i64 s = 0; s < cnt; s += kTile) tl.host.push_back(SegTask{q, g, (int)s, (int)cnt});
    }
    tl.chan_begin.push_back((int)tl.host.size());
  }
  tl.n = (int)tl.host.size();
}

int tiles(int n) { return (n + 63) / 64; }

// (re)build the tile tasks of a GEMM family for tile shape `tile`.  Groups
// (channels / cross-coupled blocks) are packed onto the XCDs by cost
// (largest first onto the least loaded XCD); each XCD's range lists its
// groups largest first, so per-group ranges stay contiguous for the plain
// path and the grouped launch keeps each group inside one L2.
int family_tasks(imsrg_basis_s* B, GemmFamily& f, int tile) {
  f.tile = tile;
  f.sig = family_sig(f.probs_h);
  ++B->plan_ver;
  const int bm = tile_rows(tile), bn = tile_cols(tile);
  const int ng = (int)f.group_end.size();
  std::vector<std::vector<GemmTask>> gt(ng);
  std::vector<double> gcost(ng, 0.0);
  int p0 = 0;
  for (int g = 0; g < ng; ++g) {
    const int pe = f.group_end[g];
    for (int pid = p0; pid < pe; ++pid) {
      const GemmProb& p = f.probs_h[pid];
      const int tm_n = (p.m + bm - 1) / bm, tn_n = (p.n + bn - 1) / bn;
      for (int ks = 0; ks < f.nks[pid]; ++ks) {
        const int kb = p.kchunk > 0 ? ks * p.kchunk : 0;
        const int ke = p.kchunk > 0 ? std::min(p.k, kb + p.kchunk) : p.k;
        for (int tm = 0; tm < tm_n; ++tm)
          for (int tn = 0; tn < tn_n; ++tn) {
            if (f.upper && (tn + 1) * bn <= tm * bm) continue;   // tile entirely below the diagonal
            gt[g].push_back(GemmTask{pid, tm, tn, ks});
            gcost[g] += 8.0 + (ke - kb) * (p.nprod == 2 ? 2.0 : 1.0);
          }
      }
    }
    p0 = pe;
  }
  std::vector<int> gord(ng);
  for (int g = 0; g < ng; ++g) gord[g] = g;
  std::stable_sort(gord.begin(), gord.end(), [&](int x, int y) { return gcost[x] > gcost[y]; });
  f.tasks.host.clear();
  f.tasks.chan_begin.assign(1, 0);
  if (f.xcd_aware) {
    std::vector<std::vector<int>> xg(kXcds);
    double load[kXcds] = {0};
    for (int g : gord) {
      if (gt[g].empty()) continue;
      int best = 0;
      for (int x = 1; x < kXcds; ++x)
        if (load[x] < load[best]) best = x;
      xg[best].push_back(g);
      load[best] += gcost[g];
    }
    int maxc = 0;
    for (int x = 0; x < kXcds; ++x) {
      f.tasks.xmap.start[x] = (int)f.tasks.host.size();
      for (int g : xg[x]) {
        f.tasks.host.insert(f.tasks.host.end(), gt[g].begin(), gt[g].end());
        f.tasks.chan_begin.push_back((int)f.tasks.host.size());
      }
      f.tasks.xmap.count[x] = (int)f.tasks.host.size() - f.tasks.xmap.start[x];
      maxc = std::max(maxc, f.tasks.xmap.count[x]);
    }
    f.tasks.grid = kXcds * maxc;
  } else {
    for (int g : gord) {
      f.tasks.host.insert(f.tasks.host.end(), gt[g].begin(), gt[g].end());
      f.tasks.chan_begin.push_back((int)f.tasks.host.size());
    }
    for (int x = 0; x < kXcds; ++x) { f.tasks.xmap.start[x] = 0; f.tasks.xmap.count[x] = -1; }
    f.tasks.grid = (int)f.tasks.host.size();
  }
  f.tasks.n = (int)f.tasks.host.size();
  if (f.tasks_alloc) {
    for (auto it = B->allocs.begin(); it != B->allocs.end(); ++it)
      if (*it == (void*)f.tasks_alloc) { B->allocs.erase(it); break; }
    (void)hipFree(f.tasks_alloc);
    f.tasks_alloc = nullptr;
  }
  int rc = upload(B, f.tasks.host, &f.tasks.dev);
  f.tasks_alloc = f.tasks.dev;
  return rc;
}

int build_plan(imsrg_basis_s* B) {
  const Basis& b = B->b;
  const std::vector<int> order = channel_order(b);
  build_seg(b, B->seg_tri, SEG_TRI, order);
  build_seg(b, B->seg_full, SEG_FULL, order);
  build_seg(b, B->seg_g3, SEG_G3, order);
  build_seg(b, B->seg_epi, SEG_EPI, order);
  // cross-coupled element tiles, largest blocks first
  {
    std::vector<int> co(b.ncc);
    for (int q = 0; q < b.ncc; ++q) co[q] = q;
    std::stable_sort(co.begin(), co.end(),
                     [&](int x, int y) { return (i64)b.rlen[x] * b.ncol[x] > (i64)b.rlen[y] * b.ncol[y]; });
    B->cc.host.clear();
    B->cc.chan_begin.assign(1, 0);
    B->ph.probs_h.clear(); B->ph.nks.clear(); B->ph.group_end.clear();
    B->ph.upper = true;
    std::vector<int> co2 = co;   // W' GEMM: largest r^2 * nb first
    for (int q : co) {
      const i64 cnt = (i64)b.rlen[q] * b.ncol[q];
      for (i64 s = 0; s < cnt; s += kTile) B->cc.host.push_back(CCTask{q, (int)s});
      B->cc.chan_begin.push_back((int)B->cc.host.size());
    }
    std::stable_sort(co2.begin(), co2.end(), [&](int x, int y) {
      return (i64)b.rlen[x] * b.rlen[x] * (b.ncol[x] + 1) > (i64)b.rlen[y] * b.rlen[y] * (b.ncol[y] + 1);
    });
    for (int q : co2) {
      const int r = b.rlen[q];
      if (r > 0) {
        // W' = Lh LCC RCC^T - Rh RCC LCC^T, upper tiles only (alpha = Lh and
        // s2 = -Rh/Lh are applied at launch)
        GemmProb p{};
        p.a1 = b.ccoff[q]; p.b1 = b.ccoff[q]; p.a2 = b.ccoff[q]; p.b2 = b.ccoff[q]; p.c = b.wcoff[q];
        p.m = r; p.n = r; p.k = b.ncol[q];
        p.lda = r; p.ldb = r; p.ldc = r;
        p.nprod = 2; p.ta1 = 0; p.tb1 = 1; p.ta2 = 0; p.tb2 = 1;
        p.sa1 = 0; p.sb1 = 1; p.sa2 = 1; p.sb2 = 0; p.sc = 2;
        p.alpha = 1.0; p.s2 = 1.0; p.beta = 0.0;
        B->ph.probs_h.push_back(p);
        B->ph.nks.push_back(1);
      }
      B->ph.group_end.push_back((int)B->ph.probs_h.size());
    }
    B->cc.n = (int)B->cc.host.size();
  }
  // 222_pp_hh: pppp product L1 R1 (-> w1) + tile-pair epilogue, K2 = pp x
  // (ph|hh) dual product, K3 = (ph|hh)^T (ph|hh)
  B->pairs.host.clear(); B->pairs.chan_begin.assign(1, 0);
  for (GemmFamily* f : {&B->pp1, &B->k2, &B->k3}) { f->probs_h.clear(); f->nks.clear(); f->group_end.clear(); }
  // split K (deterministic: partial slabs summed in order by the epilogue)
  const int KC2 = 128, KC3 = 64;
  B->poff_h.assign(b.nblocks, 0); B->k3ks_h.assign(b.nblocks, 1);
  B->k2off_h.assign(b.nblocks, 0); B->k2ks_h.assign(b.nblocks, 1);
  i64 pofs = 0, k2ofs = 0;
  for (int q = 0; q < b.nblocks; ++q) {
    const int np = b.dims[q * 3], nb = b.dims[q * 3 + 1], nh = b.dims[q * 3 + 2];
    B->k3ks_h[q] = std::max(1, (np + KC3 - 1) / KC3);
    B->poff_h[q] = pofs;
    pofs += (i64)B->k3ks_h[q] * (nb + nh) * (nb + nh);
    B->k2ks_h[q] = std::max(1, (np + KC2 - 1) / KC2);
    B->k2off_h[q] = k2ofs;
    k2ofs += (i64)B->k2ks_h[q] * np * (nb + nh);
  }
  for (int q : order) {
    const int np = b.dims[q * 3], nb = b.dims[q * 3 + 1], nh = b.dims[q * 3 + 2];
    const int T = tiles(np);
    for (int ti = 0; ti < T; ++ti)
      for (int tj = ti; tj < T; ++tj) B->pairs.host.push_back(PairTask{q, ti, tj, 0});
    B->pairs.chan_begin.push_back((int)B->pairs.host.size());
    if (np > 0) {
      // C = L1 R1 + s2 L3 R3^T with s2 = -Rh applied at launch (K2 = nh)
      GemmProb p{};
      p.a1 = b.goff[q * 6 + 0]; p.b1 = b.goff[q * 6 + 0]; p.c = b.goff[q * 6 + 0];
      p.a2 = b.goff[q * 6 + 2]; p.b2 = b.goff[q * 6 + 2];
      p.m = np; p.n = np; p.k = np;
      p.lda = np; p.ldb = np; p.ldc = np;
      p.nprod = nh > 0 ? 2 : 1; p.k2 = nh;
      p.ta2 = 0; p.tb2 = 1;
      p.sa1 = 0; p.sb1 = 1; p.sa2 = 0; p.sb2 = 1; p.sc = 2;
      p.alpha = 1.0; p.s2 = 1.0; p.beta = 0.0;
      B->pp1.probs_h.push_back(p);
      B->pp1.nks.push_back(1);
    }
    B->pp1.group_end.push_back((int)B->pp1.probs_h.size());
    if (np > 0 && nb + nh > 0) {
      GemmProb p{};
      p.a1 = b.goff[q * 6 + 0]; p.b1 = b.goff[q * 6 + 1];
      p.a2 = b.goff[q * 6 + 0]; p.b2 = b.goff[q * 6 + 1];
      p.c = B->k2off_h[q];
      p.m = np; p.n = nb + nh; p.k = np;
      p.lda = np; p.ldb = np; p.ldc = np;
      p.nprod = 2;
      p.sa1 = 0; p.sb1 = 1; p.sa2 = 1; p.sb2 = 0; p.sc = 2;
      p.alpha = 1.0; p.s2 = -1.0; p.beta = 0.0;
      p.kchunk = KC2; p.cstride = (i64)np * (nb + nh);
      B->k2.probs_h.push_back(p);
      B->k2.nks.push_back(B->k2ks_h[q]);
    }
    B->k2.group_end.push_back((int)B->k2.probs_h.size());
    if (nb + nh > 0) {
      GemmProb p{};
      p.a1 = b.goff[q * 6 + 1]; p.b1 = b.goff[q * 6 + 1];
      p.c = B->poff_h[q];
      p.m = nb + nh; p.n = nb + nh; p.k = np;
      p.lda = std::max(np, 1); p.ldb = std::max(np, 1); p.ldc = nb + nh;
      p.nprod = 1; p.ta1 = 1; p.tb1 = 0;
      p.sa1 = 0; p.sb1 = 1; p.sc = 2;
      p.alpha = 1.0; p.s2 = 0.0; p.beta = 0.0;
      p.kchunk = KC3; p.cstride = (i64)(nb + nh) * (nb + nh);
      B->k3.probs_h.push_back(p);
      B->k3.nks.push_back(B->k3ks_h[q]);
    }
    B->k3.group_end.push_back((int)B->k3.probs_h.size());
  }
  B->pairs.n = (int)B->pairs.host.size();
  B->pp1.tile = B->tile_req[0] & 7; B->k2.tile = B->tile_req[1] & 7; B->k3.tile = B->tile_req[2] & 7;
  B->ph.tile = B->tile_req[3] & 7;
  // one-body outputs of 121 / 221 (hh p<=q, pp p<=q, ph), same (l, j, tz)
  B->one_h.clear();
  auto same = [&](int x, int y) { return b.jj[x] == b.jj[y] && b.ll[x] == b.ll[y] && b.tz[x] == b.tz[y]; };
  for (int p = 0; p < b.A; ++p)
    for (int q = p; q < b.A; ++q)
      if (same(b.holes[p], b.holes[q])) B->one_h.push_back(OneTask{0, b.holes[p], b.holes[q], p, q});
  for (int p = 0; p < b.P; ++p)
    for (int q = p; q < b.P; ++q)
      if (same(b.parts[p], b.parts[q])) B->one_h.push_back(OneTask{1, b.parts[p], b.parts[q], p, q});
  for (int p = 0; p < b.P; ++p)
    for (int q = 0; q < b.A; ++q)
      if (same(b.parts[p], b.holes[q])) B->one_h.push_back(OneTask{2, b.parts[p], b.holes[q], p, q});
  // 122 partner lists.  For the stored pair (a,b) of channel q: row role =
  // (a -> i, f(a,i)) then (b -> i, f(b,i)); column role = (a -> i, f(i,a))
  // then (b -> i, f(i,b)); i runs over the (l,j,tz) block of the replaced
  // orbit.  Phase / sqrt(2) of the resulting pair as v_elem applies them.
  {
    // the partner terms address operator elements with 32-bit offsets
    if (b.oplen >= ((i64)1 << 31)) return fail(IMSRG_EINVAL, "operator too large for 32-bit element offsets");
    B->part_h.clear();
    const size_t npairs = b.pairs.size();
    B->rowp_h.assign(npairs + 1, 0);
    B->colp_h.assign(npairs + 1, 0);
    auto fref = [&](int x, int y, int& foff, int& fsw) {   // f_elem(x, y) location
      const int c1 = b.con[x], c2 = b.con[y];
      fsw = 0;
      if (c1 + c2 == 0) foff = (int)(b.off_fpp + b.cidx[x] + (i64)b.cidx[y] * b.P);
      else if (c1 + c2 == 2) foff = (int)(b.off_fhh + b.cidx[x] + (i64)b.cidx[y] * b.A);
      else if (c1 > c2) { foff = (int)(b.off_fph + b.cidx[y] + (i64)b.cidx[x] * b.P); fsw = 1; }
      else foff = (int)(b.off_fph + b.cidx[x] + (i64)b.cidx[y] * b.P);
    };
    auto pair_code = [&](int x, int y, int J) {        // pair (x,y) as v_elem orders it
      int neg = 0, s2 = 0, lo = x, hi = y;
      if (x > y) { lo = y; hi = x; neg = ((1 + (b.jj[x] + b.jj[y] - J) / 2) & 1); }
      else if (x == y) s2 = 1;
      const int cl = b.con[lo] + b.con[hi];
      const int idx = b.xmap[((size_t)lo * b.N + hi) * b.nJ + J / 2];
      return idx | (cl << 22) | (neg << 25) | (s2 << 26);
    };
    std::vector<std::vector<Partner>> rowl(npairs), coll(npairs);
    for (int q = 0; q < b.nblocks; ++q) {
      const int J = b.chJ[q];
      for (int c = 0; c < 3; ++c)
        for (int ix = 0; ix < b.dims[q * 3 + c]; ++ix) {
          const size_t gid = (size_t)b.pairoff[q * 3 + c] + ix;
          const int x = b.pl[q * 3 + c][2 * ix], y = b.pl[q * 3 + c][2 * ix + 1];
          int foff, fsw;
          for (int k = b.spstart[b.spblk[x]]; k < b.spstart[b.spblk[x] + 1]; ++k) {
            const int i = b.spmem[k];
            fref(x, i, foff, fsw);
            rowl[gid].push_back(Partner{pair_code(i, y, J) | (fsw << 24), foff});
            fref(i, x, foff, fsw);
            coll[gid].push_back(Partner{pair_code(i, y, J) | (fsw << 24), foff});
          }
          for (int k = b.spstart[b.spblk[y]]; k < b.spstart[b.spblk[y] + 1]; ++k) {
            const int i = b.spmem[k];
            fref(y, i, foff, fsw);
            rowl[gid].push_back(Partner{pair_code(x, i, J) | (fsw << 24), foff});
            fref(i, y, foff, fsw);
            coll[gid].push_back(Partner{pair_code(x, i, J) | (fsw << 24), foff});
          }
        }
    }
    for (size_t gi = 0; gi < npairs; ++gi) {
      B->rowp_h[gi] = (int)B->part_h.size();
      B->part_h.insert(B->part_h.end(), rowl[gi].begin(), rowl[gi].end());
    }
    B->rowp_h[npairs] = (int)B->part_h.size();
    for (size_t gi = 0; gi < npairs; ++gi) {
      B->colp_h[gi] = (int)B->part_h.size();
      B->part_h.insert(B->part_h.end(), coll[gi].begin(), coll[gi].end());
    }
    B->colp_h[npairs] = (int)B->part_h.size();
    // compact index of the one-body elements the partner terms read (the
    // kernels stage them in LDS); the lists above carried operator offsets
    B->f1map_h.assign((size_t)(b.off_fpp + (i64)b.P * b.P), -1);
    B->f1off_h.clear();
    for (Partner& p : B->part_h) {
      if (B->f1map_h[p.fk] < 0) { B->f1map_h[p.fk] = (int)B->f1off_h.size(); B->f1off_h.push_back(p.fk); }
      p.fk = B->f1map_h[p.fk];
    }
    B->nf1 = (int)B->f1off_h.size();
    if (B->nf1 > kF1Max) return fail(IMSRG_EINVAL, "one-body partner set exceeds the LDS table of 122");
    int lmax = 0;
    for (size_t gi = 0; gi < npairs; ++gi) lmax = std::max(lmax, B->rowp_h[gi + 1] - B->rowp_h[gi]);
    B->pstride = (int)npairs;
    B->partT_h.assign((size_t)std::max(lmax, 1) * npairs, Partner{0, 0});
    for (size_t gi = 0; gi < npairs; ++gi)
      for (int k = 0; k < B->rowp_h[gi + 1] - B->rowp_h[gi]; ++k)
        B->partT_h[(size_t)k * npairs + gi] = B->part_h[B->rowp_h[gi] + k];
  }
  // 222_ph outputs as (orbit quadruple, 2J) items, J fastest: every upper
  // element of every channel exactly once (checked against seg_tri)
  {
    B->quad_h.clear();
    if (b.N > 256) return fail(IMSRG_EINVAL, "more than 256 orbits");
    static const int rcls[6] = {0, 0, 0, 1, 2, 1}, ccls[6] = {0, 1, 2, 1, 2, 2};
    for (int Pr = 0; Pr <= 1; ++Pr)
      for (int T = -1; T <= 1; ++T)
        for (int g = 0; g < 6; ++g) {
          std::vector<int> rows, cols;
          for (int i = 0; i < b.N; ++i)
            for (int j = i; j < b.N; ++j) {
              if ((b.ll[i] + b.ll[j]) % 2 != Pr || (b.tz[i] + b.tz[j]) / 2 != T) continue;
              const int c = b.con[i] + b.con[j];
              // key (i << 8) | j orders pairs as the channel lists do (i, then j)
              if (c == rcls[g]) rows.push_back((i << 8) | j);
              if (c == ccls[g]) cols.push_back((i << 8) | j);
            }
          // order: column pair (k,l), then j, then i fastest, then 2J: the
          // WCC rows (k,i) and (l,i) of neighbouring lanes are neighbours
          std::vector<int> rows_j;   // row pairs sorted by (j, i)
          rows_j = rows;
          std::stable_sort(rows_j.begin(), rows_j.end(),
                           [](int x, int y) { return ((x & 255) << 8 | (x >> 8)) < ((y & 255) << 8 | (y >> 8)); });
          for (int cp : cols) {
            const int k = cp >> 8, l = cp & 255;
            for (int rp : rows_j) {
              if (kSquare[g] && cp < rp) continue;   // (k,l) >= (i,j): the upper triangle
              const int i = rp >> 8, j = rp & 255;
              const int lo = std::max(std::abs(b.jj[i] - b.jj[j]), std::abs(b.jj[k] - b.jj[l]));
              const int hi = std::min(b.jj[i] + b.jj[j], b.jj[k] + b.jj[l]);
              for (int J = lo; J <= hi; J += 2)
                B->quad_h.push_back(QuadItem{i | (j << 8) | (k << 16) | (l << 24), J});
            }
          }
        }
    i64 tri = 0;
    for (int q = 0; q < b.nblocks; ++q)
      for (int g = 0; g < 6; ++g)
        tri += kSquare[g] ? (i64)b.rows(q, g) * (b.rows(q, g) + 1) / 2 : b.blocksize(q, g);
    if ((i64)B->quad_h.size() != tri) return fail(IMSRG_EINVAL, "quadruple enumeration does not cover the channels");
    if (b.wclen >= ((i64)1 << 32)) return fail(IMSRG_EINVAL, "WCC exceeds 2^32 doubles");
    if (b.ncc > kPhMaxCC) return fail(IMSRG_EINVAL, "more cross-coupled blocks than the 222_ph chunk kernel holds");
    // chunks: consecutive quadruples with at most kPhChunkItems items and
    // kPhChunkLanes J' terms (the two recouplings' ranges, none when an
    // isospin / parity check fails: the rules of ph_ranges)
    B->phc_h.clear(); B->fq_h.clear(); B->fql_h.clear(); B->fqi_h.clear();
    auto lanes_of = [&](int u) {
      const int i = u & 255, j = (u >> 8) & 255, k = (u >> 16) & 255, l = (u >> 24) & 255;
      const int ji = b.jj[i], jj = b.jj[j], jk = b.jj[k], jl = b.jj[l];
      if (std::abs(b.tz[l] - b.tz[j]) != std::abs(b.tz[i] - b.tz[k]) ||
          (b.ll[l] + b.ll[j]) % 2 != (b.ll[i] + b.ll[k]) % 2)
        return 0;
      if (std::abs(b.tz[k] - b.tz[j]) != std::abs(b.tz[i] - b.tz[l]) ||
          (b.ll[k] + b.ll[j]) % 2 != (b.ll[i] + b.ll[l]) % 2)
        return 0;
      const int jin = std::max(std::abs(jj - jl), std::abs(ji - jk));
      const int nin = std::max(0, (std::min(jj + jl, ji + jk) - jin) / 2 + 1);
      const int jout = std::max(std::abs(ji - jl), std::abs(jj - jk));
      const int nout = std::max(0, (std::min(ji + jl, jj + jk) - jout) / 2 + 1);
      return nin + nout;
    };
    PhChunk cur{0, 0, 0, 0, 0, 0, 0, 0};
    const size_t nit = B->quad_h.size();
    for (size_t x = 0; x < nit;) {
      size_t e = x;
      while (e < nit && B->quad_h[e].ijkl == B->quad_h[x].ijkl) ++e;
      const int items = (int)(e - x), lanes = lanes_of(B->quad_h[x].ijkl);
      if (items > kPhChunkItems || lanes > kPhChunkLanes) return fail(IMSRG_EINVAL, "quadruple exceeds a 222_ph chunk");
      if (cur.nq > 0 && (cur.nitems + items > kPhChunkItems || cur.nlanes + lanes > kPhChunkLanes)) {
        B->phc_h.push_back(cur);
        cur = PhChunk{(int)B->fq_h.size(), 0, (int)x, 0, 0, 0, 0, 0};
      }
      B->fq_h.push_back(B->quad_h[x].ijkl);
      B->fql_h.push_back(cur.nlanes);
      B->fqi_h.push_back((int)x);
      cur.nq += 1;
      cur.nitems += items;
      cur.nlanes += lanes;
      x = e;
    }
    if (cur.nq > 0) B->phc_h.push_back(cur);
  }
  // norm chunks: fhh (w 1), fph (w 2), fpp (w 1), gam blocks (w 1 square, 2 rectangular)
  B->ch_off_h.clear(); B->ch_len_h.clear(); B->ch_w_h.clear();
  auto add_range = [&](i64 off, i64 len, double w) {
    for (i64 s = 0; s < len; s += 8192) {
      B->ch_off_h.push_back(off + s);
      B->ch_len_h.push_back((int)std::min<i64>(8192, len - s));
      B->ch_w_h.push_back(w);
    }
  };
  add_range(b.off_fhh, (i64)b.A * b.A, 1.0);
  add_range(b.off_fph, (i64)b.P * b.A, 2.0);
  add_range(b.off_fpp, (i64)b.P * b.P, 1.0);
  for (int q = 0; q < b.nblocks; ++q)
    for (int g = 0; g < 6; ++g) add_range(b.goff[q * 6 + g], b.blocksize(q, g), kSquare[g] ? 1.0 : 2.0);

  // upload
  int rc;
  if ((rc = upload(B, B->seg_tri.host, &B->seg_tri.dev))) return rc;
  if ((rc = upload(B, B->seg_full.host, &B->seg_full.dev))) return rc;
  if ((rc = upload(B, B->seg_g3.host, &B->seg_g3.dev))) return rc;
  if ((rc = upload(B, B->seg_epi.host, &B->seg_epi.dev))) return rc;
  if ((rc = upload(B, B->cc.host, &B->cc.dev))) return rc;
  if ((rc = upload(B, B->pairs.host, &B->pairs.dev))) return rc;
  for (GemmFamily* f : {&B->pp1, &B->k2, &B->k3, &B->ph}) {
    if ((rc = upload(B, f->probs_h, &f->probs))) return rc;
    if ((rc = family_tasks(B, *f, f->tile))) return rc;
  }
  if ((rc = upload(B, B->one_h, &B->one))) return rc;
  B->symt.host.clear();
  for (int q = 0; q < b.nblocks; ++q)
    for (int g : {0, 3, 4}) {
      const int n = b.dims[q * 3 + (g == 0 ? 0 : (g == 3 ? 1 : 2))];
      const int nt = (n + 63) / 64;
      for (int tj = 0; tj < nt; ++tj)
        for (int ti = 0; ti <= tj; ++ti) B->symt.host.push_back(PairTask{q, ti, tj, g});
    }
  B->symt.n = (int)B->symt.host.size();
  if ((rc = upload(B, B->symt.host, &B->symt.dev))) return rc;
  // particle-hole pairs inside one (l, j, tz) block (121's intermediate sums)
  {
    std::vector<int> php;
    int jmx = 0;
    for (int x = 0; x < b.N; ++x) jmx = std::max(jmx, b.jj[x]);
    for (int i = 0; i < b.A; ++i)
      for (int a = 0; a < b.P; ++a) {
        const int ak = b.parts[a], ik = b.holes[i];
        if (b.jj[ak] == b.jj[ik] && b.ll[ak] == b.ll[ik] && b.tz[ak] == b.tz[ik]) {
          php.push_back(a);
          php.push_back(i);
        }
      }
    B->D.nphp = (int)php.size() / 2;
    B->D.jsmax = jmx + 1;
    if ((rc = upload(B, php, (int**)&B->D.php))) return rc;
  }
  if ((rc = upload(B, B->quad_h, &B->quad))) return rc;
  if ((rc = upload(B, B->phc_h, &B->phc)) || (rc = upload(B, B->fq_h, &B->fq)) ||
      (rc = upload(B, B->fql_h, &B->fql)) || (rc = upload(B, B->fqi_h, &B->fqi)))
    return rc;
  if ((rc = upload(B, B->part_h, &B->partners))) return rc;
  if ((rc = upload(B, B->partT_h, &B->partT))) return rc;
  if ((rc = upload(B, B->f1off_h, &B->f1off))) return rc;
  if ((rc = upload(B, B->rowp_h, &B->rowp))) return rc;
  if ((rc = upload(B, B->colp_h, &B->colp))) return rc;
  if ((rc = upload(B, B->ch_off_h, &B->ch_off))) return rc;
  if ((rc = upload(B, B->ch_len_h, &B->ch_len))) return rc;
  if ((rc = upload(B, B->ch_w_h, &B->ch_w))) return rc;
  if ((rc = upload(B, B->poff_h, &B->poff))) return rc;
  if ((rc = upload(B, B->k2off_h, &B->k2off))) return rc;
  if ((rc = upload(B, B->k3ks_h, &B->k3ks))) return rc;
  if ((rc = upload(B, B->k2ks_h, &B->k2ks))) return rc;
  if ((rc = dalloc(B, std::max<i64>(pofs, 1) * sizeof(double), (void**)&B->Pscr))) return rc;
  if ((rc = dalloc(B, std::max<i64>(k2ofs, 1) * sizeof(double), (void**)&B->K2scr))) return rc;
  B->D.spscr = std::max<i64>(pofs, 1);
  B->D.sk2scr = std::max<i64>(k2ofs, 1);
  B->npart = std::max<int>(std::max<int>((int)B->ch_off_h.size(), B->seg_g3.n), b.nblocks + 1);
  if ((rc = dalloc(B, (size_t)B->npart * sizeof(double), (void**)&B->part))) return rc;
  B->D.spart = B->npart;
  B->D.ninst = 1;
  B->D.act = nullptr;
  if ((rc = dalloc(B, (size_t)b.N * b.N * sizeof(double), (void**)&B->M))) return rc;
  if ((rc = dalloc(B, 64 * sizeof(double), (void**)&B->dval))) return rc;
  HIPCHK(hipHostMalloc((void**)&B->hval, 64 * sizeof(double), hipHostMallocDefault));
  return 0;
}

int ensure_dev(imsrg_basis_s* B) {
  if (B->dev_ready) return 0;
  const Basis& b = B->b;
  DevBasis& D = B->D;
  int rc;
  D.N = b.N; D.A = b.A; D.P = b.P; D.JTM = b.JTM; D.nJ = b.nJ;
  D.nblocks = b.nblocks; D.ncc = b.ncc;
  D.off_fhh = b.off_fhh; D.off_fph = b.off_fph; D.off_fpp = b.off_fpp;
  D.oplen = b.oplen; D.cclen = b.cclen; D.wclen = b.wclen;
  D.nh6 = b.nh6; D.nJ6 = b.nJ6;
#define UP(field, vec)                                                   \
  if ((rc = upload(B, b.vec, (std::remove_const_t<std::remove_pointer_t<decltype(D.field)>>**)&D.field))) return rc;
  UP(jj, jj) UP(ll, ll) UP(tz, tz) UP(con, con) UP(cidx, cidx) UP(holes, holes) UP(parts, parts)
  UP(spstart, spstart) UP(spmem, spmem) UP(spblk, spblk) UP(xmap, xmap) UP(dims, dims) UP(chJ, chJ)
  UP(goff, goff) UP(pairs, pairs) UP(pairoff, pairoff) UP(rlen, rlen) UP(ncol, ncol) UP(ccJ, ccJ)
  UP(ccoff, ccoff) UP(wcoff, wcoff) UP(rmap, rmap) UP(nbmap, nbmap) UP(rmapJ, rmapJ) UP(rmapJT, rmapJT) UP(ccrows, ccrows) UP(ccrowoff, ccrowoff)
  UP(cccols, cccols) UP(cccoloff, cccoloff) UP(six, six)
#undef UP
  if ((rc = build_plan(B))) return rc;
  for (hipStream_t& x : B->aux) HIPCHK(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
  for (hipEvent_t& e : B->tev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  B->aux_ready = true;
  B->dev_ready = true;
  return 0;
}

// launch a task list either in one grouped launch or one launch per group
template <class T, class F>
void run(imsrg_basis_s* B, TaskList<T>& tl, F&& launch) {
  if (tl.n == 0) return;
  if (B->path == IMSRG_PATH_GROUPED) {
    launch(tl.dev, tl.n);
    return;
  }
  for (size_t g = 0; g + 1 < tl.chan_begin.size(); ++g) {
    const int b0 = tl.chan_begin[g], e0 = tl.chan_begin[g + 1];
    if (e0 > b0) launch(tl.dev + b0, e0 - b0);
  }
}

hipStream_t S(imsrg_stream s) { return (hipStream_t)s; }

// one GEMM family: one XCD-mapped launch (grouped) or one launch per group
void run_gemm(imsrg_basis_s* B, GemmFamily& f, const Slots& sl, hipStream_t s, double ascale = 1.0,
              double s2scale = 1.0) {
  if (f.tasks.n == 0) return;
  if (B->path == IMSRG_PATH_GROUPED) {
    launch_gemm_tiled(f.tile, sl, f.probs, f.tasks.dev, f.tasks.grid, f.tasks.xmap, s, ascale, s2scale, 1, f.sig);
    return;
  }
  const XcdMap ident{{0}, {-1}};
  for (size_t g = 0; g + 1 < f.tasks.chan_begin.size(); ++g) {
    const int b0 = f.tasks.chan_begin[g], e0 = f.tasks.chan_begin[g + 1];
    if (e0 > b0) launch_gemm_tiled(f.tile, sl, f.probs, f.tasks.dev + b0, e0 - b0, ident, s, ascale, s2scale, 1, f.sig);
  }
}

// ------------------------------------------------------------ term bodies
// The term sequence runs in an execution context: the device basis (whose
// ninst / act select the instances of batched launches) and the scratch the
// terms write (partial sums, 222_pp_hh split-K slabs, monopoles, scalars;
// one copy per instance).  Scalars of instance k sit at dval + k * kDStride.
constexpr i64 kDStride = 64;

struct Exec {
  imsrg_basis_s* B = nullptr;
  DevBasis D{};
  double *part = nullptr, *Pscr = nullptr, *K2scr = nullptr, *M = nullptr, *dval = nullptr;
};

Exec basis_exec(imsrg_basis_s* B) {
  Exec X;
  X.B = B;
  X.D = B->D;
  X.part = B->part; X.Pscr = B->Pscr; X.K2scr = B->K2scr; X.M = B->M; X.dval = B->dval;
  return X;
}

// one GEMM family in the context's instances
void run_gemm_x(const Exec& X, GemmFamily& f, Slots sl, hipStream_t s, double ascale = 1.0, double s2scale = 1.0) {
  sl.act = X.D.act;
  if (X.D.ninst > 1 && X.B->path == IMSRG_PATH_GROUPED) {
    if (f.tasks.n)
      launch_gemm_tiled(f.tile, sl, f.probs, f.tasks.dev, f.tasks.grid, f.tasks.xmap, s, ascale, s2scale, X.D.ninst,
                        f.sig);
    return;
  }
  run_gemm(X.B, f, sl, s, ascale, s2scale);
}

void t_e0(const Exec& X, const double* L, int Lh, const double* R, int Rh, double* dst, hipStream_t s) {
  launch_e0_partials(X.D, L, Lh, R, Rh, X.part, s);
  launch_e0_final_b(X.D, X.part, (Lh * Rh == -1) ? 1 : 0, 1, dst, 0, s);
}
void t_111(const Exec& X, const double* L, int Lh, const double* R, int Rh, double* RES, hipStream_t s) {
  launch_comm111(X.D, L, Lh, R, Rh, RES, s);
}
void t_121(const Exec& X, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh,
           hipStream_t s) {
  launch_comm121(X.D, X.B->one, (int)X.B->one_h.size(), L, Lh, R, Rh, RES, RESh, s);
}
void t_122(const Exec& X, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh,
           hipStream_t s, bool mirror = true) {
  imsrg_basis_s* B = X.B;
  run(B, B->seg_tri, [&](const SegTask* t, int n) {
    launch_comm122p(X.D, t, n, B->partners, B->partT, B->pstride, B->rowp, B->colp, B->f1off, B->nf1, L, Lh, R, Rh,
                    RES, RESh, s,
                    mirror ? 1 : 0);
  });
}
void t_pphh(const Exec& X, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh,
            double* w1, double* w2, hipStream_t s, int ref_outputs = 1) {
  (void)RESh;
  imsrg_basis_s* B = X.B;
  const i64 op = B->b.oplen;
  Slots s1{};
  s1.p[0] = L; s1.p[1] = R; s1.p[2] = w1;
  s1.st[0] = op; s1.st[1] = op; s1.st[2] = op;
  run_gemm_x(X, B->pp1, s1, s, 1.0, -(double)Rh);
  run(B, B->pairs,
      [&](const PairTask* t, int n) { launch_pppp_sym(X.D, t, n, L, Lh, R, Rh, RES, w1, w2, ref_outputs, s); });
  Slots s2{};
  s2.p[0] = L; s2.p[1] = R; s2.p[2] = X.K2scr;
  s2.st[0] = op; s2.st[1] = op; s2.st[2] = X.D.sk2scr;
  run_gemm_x(X, B->k2, s2, s);
  Slots s3{};
  s3.p[0] = L; s3.p[1] = R; s3.p[2] = X.Pscr;
  s3.st[0] = op; s3.st[1] = op; s3.st[2] = X.D.spscr;
  run_gemm_x(X, B->k3, s3, s);
  run(B, B->seg_epi, [&](const SegTask* t, int n) {
    launch_pphh_epilogue(X.D, t, n, L, Lh, R, Rh, RES, w1, w2, X.Pscr, B->poff, B->k3ks, X.K2scr, B->k2off,
                         B->k2ks, s);
  });
}
void t_221(const Exec& X, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh,
           const double* w1, const double* w2, hipStream_t s) {
  launch_comm221(X.D, X.B->one, (int)X.B->one_h.size(), L, Lh, R, Rh, w1, w2, RES, RESh, s);
}
void t_pandya(const Exec& X, const double* op, int herm, double* CC, hipStream_t s) {
  run(X.B, X.B->cc, [&](const CCTask* t, int n) { launch_pandya(X.D, t, n, op, herm, CC, s); });
}
void t_ph(const Exec& X, const double* LCC, int Lh, const double* RCC, int Rh, double* RES, int RESh,
          double* WCC, hipStream_t s, bool mirror = true) {
  imsrg_basis_s* B = X.B;
  Slots sl{};
  sl.p[0] = LCC; sl.p[1] = RCC; sl.p[2] = WCC;
  sl.st[0] = B->b.cclen; sl.st[1] = B->b.cclen; sl.st[2] = B->b.wclen;
  run_gemm_x(X, B->ph, sl, s, (double)Lh, -(double)(Rh * Lh));
  if (B->path == IMSRG_PATH_GROUPED)
    launch_comm222ph_chunk(X.D, B->phc, (int)B->phc_h.size(), B->fq, B->fql, B->fqi, B->quad, WCC, Lh, Rh, RES,
                           RESh, s, mirror ? 1 : 0);
  else
    run(B, B->seg_tri,
        [&](const SegTask* t, int n) { launch_comm222ph_scatter(X.D, t, n, WCC, Lh, Rh, RES, RESh, s); });
}
// op += c * T and the norm of T into dst, one pass over T (instances as t_norm)
void t_axpy_norm(const Exec& X, double* op, const double* T, double c, double* dst, hipStream_t s) {
  imsrg_basis_s* B = X.B;
  const int nch = (int)B->ch_off_h.size();
  launch_axpy_norm_b(X.D, op, T, c, B->ch_off, B->ch_len, B->ch_w, nch, X.part, s);
  launch_sum_sqrt_b(X.D, X.part, nch, dst, kDStride, 1, s);
}
// device norm of op into dst (device; instance k at dst + k * kDStride)
void t_norm(const Exec& X, const double* op, double* dst, hipStream_t s) {
  imsrg_basis_s* B = X.B;
  const int nch = (int)B->ch_off_h.size();
  launch_norm_partials_b(X.D, op, B->ch_off, B->ch_len, B->ch_w, nch, X.part, s);
  launch_sum_sqrt_b(X.D, X.part, nch, dst, kDStride, 1, s);
}

// RES = [L, R]: the term sequence of BCH_EXPAND / MAGNUS_EXPAND
// (IMSRG_MAGNUS.f90:207-219, 677-683).  LCC must already hold the
// cross-coupled L; RCC is computed here.
void t_full(const Exec& X, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh, bool e0,
            double* w1, double* w2, const double* LCC, double* RCC, double* WCC, hipStream_t s) {
  // grouped: 122 and 222_ph write the upper triangles of the square blocks
  // and one symmetrising pass fills the lower ones (every term's
  // contribution is herm-symmetric, so the lower triangles need not be
  // accumulated on the way)
  const bool mirror = X.B->path != IMSRG_PATH_GROUPED;
  imsrg_basis_s* B = X.B;
  if (!mirror && B->term_streams && B->aux_ready) {
    // Concurrent schedule, same results: RES's two-body part is still
    // written in the reference order (122, pppp, the other 222_pp_hh
    // blocks, 222_ph) on s, its one-body part in the order 111, 121, 221;
    // the GEMM families, the Pandya transform of R and the zero- and
    // one-body terms run beside them on the auxiliary streams.
    hipStream_t sg = B->aux[0], sc = B->aux[1], so = B->aux[2];
    hipEvent_t* ev = B->tev;
    (void)hipEventRecord(ev[0], s);
    for (hipStream_t x : B->aux) (void)hipStreamWaitEvent(x, ev[0], 0);
    // cross-coupled R and W' = Lh LCC RCC^T - Rh RCC LCC^T
    t_pandya(X, R, Rh, RCC, sc);
    Slots sw{};
    sw.p[0] = LCC; sw.p[1] = RCC; sw.p[2] = WCC;
    sw.st[0] = B->b.cclen; sw.st[1] = B->b.cclen; sw.st[2] = B->b.wclen;
    run_gemm_x(X, B->ph, sw, sc, (double)Lh, -(double)(Rh * Lh));
    (void)hipEventRecord(ev[1], sc);
    // the 222_pp_hh GEMM families
    const i64 op = B->b.oplen;
    Slots s1{};
    s1.p[0] = L; s1.p[1] = R; s1.p[2] = w1;
    s1.st[0] = op; s1.st[1] = op; s1.st[2] = op;
    run_gemm_x(X, B->pp1, s1, sg, 1.0, -(double)Rh);
    (void)hipEventRecord(ev[2], sg);
    Slots s2{};
    s2.p[0] = L; s2.p[1] = R; s2.p[2] = X.K2scr;
    s2.st[0] = op; s2.st[1] = op; s2.st[2] = X.D.sk2scr;
    run_gemm_x(X, B->k2, s2, sg);
    Slots s3{};
    s3.p[0] = L; s3.p[1] = R; s3.p[2] = X.Pscr;
    s3.st[0] = op; s3.st[1] = op; s3.st[2] = X.D.spscr;
    run_gemm_x(X, B->k3, s3, sg);
    (void)hipEventRecord(ev[3], sg);
    // E0 and the one-body part
    if (e0) t_e0(X, L, Lh, R, Rh, RES, so);
    t_111(X, L, Lh, R, Rh, RES, so);
    t_121(X, L, Lh, R, Rh, RES, RESh, so);
    (void)hipEventRecord(ev[4], so);
    // RES two-body in the reference order
    t_122(X, L, Lh, R, Rh, RES, RESh, s, false);
    (void)hipStreamWaitEvent(s, ev[2], 0);
    run(B, B->pairs,
        [&](const PairTask* t, int n) { launch_pppp_sym(X.D, t, n, L, Lh, R, Rh, RES, w1, w2, 2, s); });
    (void)hipStreamWaitEvent(s, ev[3], 0);
    run(B, B->seg_epi, [&](const SegTask* t, int n) {
      launch_pphh_epilogue(X.D, t, n, L, Lh, R, Rh, RES, w1, w2, X.Pscr, B->poff, B->k3ks, X.K2scr, B->k2off,
                           B->k2ks, s);
    });
    (void)hipStreamWaitEvent(s, ev[4], 0);
    t_221(X, L, Lh, R, Rh, RES, RESh, w1, w2, s);
    (void)hipStreamWaitEvent(s, ev[1], 0);
    launch_comm222ph_chunk(X.D, B->phc, (int)B->phc_h.size(), B->fq, B->fql, B->fqi, B->quad, WCC, Lh, Rh, RES,
                           RESh, s, 0);
    launch_sym_lower(X.D, B->symt.dev, B->symt.n, RES, RESh, s);
    return;
  }
  t_pandya(X, R, Rh, RCC, s);
  if (e0) t_e0(X, L, Lh, R, Rh, RES, s);
  t_111(X, L, Lh, R, Rh, RES, s);
  t_121(X, L, Lh, R, Rh, RES, RESh, s);
  t_122(X, L, Lh, R, Rh, RES, RESh, s, mirror);
  t_pphh(X, L, Lh, R, Rh, RES, RESh, w1, w2, s, mirror ? 0 : 2);
  t_221(X, L, Lh, R, Rh, RES, RESh, w1, w2, s);
  t_ph(X, LCC, Lh, RCC, Rh, RES, RESh, WCC, s, mirror);
  if (!mirror) launch_sym_lower(X.D, X.B->symt.dev, X.B->symt.n, RES, RESh, s);
}


// ============================================================ reverse mode
// One transposed 222_pp_hh product  C(out) += alpha op(A) op(B), slots
// 0 L, 1 R, 2 Lb, 3 Rb, 4 Wb1, 5 Wb2 (cotangents of the w1 / w2 blocks),
// dims 'p' np, 'b' nb, 'h' nh; mode bit0 * Lh, bit1 * Rh.
struct BwdProd {
  int out_slot, out_g, a_slot, a_g, ta, b_slot, b_g, tb;
  char m, n, k;
  double alpha;
  int mode;
};
// transposes of W11 = L1 R1, W25 = L5 R5, W15 = Lh L3^T R3, W21 = Rh L3 R3^T,
// W13 = L1 R3 - R1 L3, W23 = L3 R5 - R3 L5, W24 = Rh L6 R6^T, W26 = L6 R5 -
// R6 L5, W14 = Lh L2^T R2, W12 = L1 R2 - R1 L2, W16 = Lh L2^T R3 - Rh R2^T L3,
// W22 = Rh L3 R6^T - Lh R3 L6^T  (commutators.f90:611-755)
const BwdProd kBwd[36] = {
    {2, 0, 4, 0, 0, 1, 0, 1, 'p', 'p', 'p', 1, 0},  {3, 0, 0, 0, 1, 4, 0, 0, 'p', 'p', 'p', 1, 0},
    {2, 4, 5, 4, 0, 1, 4, 1, 'h', 'h', 'h', 1, 0},  {3, 4, 0, 4, 1, 5, 4, 0, 'h', 'h', 'h', 1, 0},
    {2, 2, 1, 2, 0, 4, 4, 1, 'p', 'h', 'h', 1, 1},  {3, 2, 0, 2, 0, 4, 4, 0, 'p', 'h', 'h', 1, 1},
    {2, 2, 5, 0, 0, 1, 2, 0, 'p', 'h', 'p', 1, 2},  {3, 2, 5, 0, 1, 0, 2, 0, 'p', 'h', 'p', 1, 2},
    {2, 0, 4, 2, 0, 1, 2, 1, 'p', 'p', 'h', 1, 0},  {3, 2, 0, 0, 1, 4, 2, 0, 'p', 'h', 'p', 1, 0},
    {3, 0, 4, 2, 0, 0, 2, 1, 'p', 'p', 'h', -1, 0}, {2, 2, 1, 0, 1, 4, 2, 0, 'p', 'h', 'p', -1, 0},
    {2, 2, 5, 2, 0, 1, 4, 1, 'p', 'h', 'h', 1, 0},  {3, 4, 0, 2, 1, 5, 2, 0, 'h', 'h', 'p', 1, 0},
    {3, 2, 5, 2, 0, 0, 4, 1, 'p', 'h', 'h', -1, 0}, {2, 4, 1, 2, 1, 5, 2, 0, 'h', 'h', 'p', -1, 0},
    {2, 5, 5, 3, 0, 1, 5, 0, 'b', 'h', 'b', 1, 2},  {3, 5, 5, 3, 1, 0, 5, 0, 'b', 'h', 'b', 1, 2},
    {2, 5, 5, 5, 0, 1, 4, 1, 'b', 'h', 'h', 1, 0},  {3, 4, 0, 5, 1, 5, 5, 0, 'h', 'h', 'b', 1, 0},
    {3, 5, 5, 5, 0, 0, 4, 1, 'b', 'h', 'h', -1, 0}, {2, 4, 1, 5, 1, 5, 5, 0, 'h', 'h', 'b', -1, 0},
    {2, 1, 1, 1, 0, 4, 3, 1, 'p', 'b', 'b', 1, 1},  {3, 1, 0, 1, 0, 4, 3, 0, 'p', 'b', 'b', 1, 1},
    {2, 0, 4, 1, 0, 1, 1, 1, 'p', 'p', 'b', 1, 0},  {3, 1, 0, 0, 1, 4, 1, 0, 'p', 'b', 'p', 1, 0},
    {3, 0, 4, 1, 0, 0, 1, 1, 'p', 'p', 'b', -1, 0}, {2, 1, 1, 0, 1, 4, 1, 0, 'p', 'b', 'p', -1, 0},
    {2, 1, 1, 2, 0, 4, 5, 1, 'p', 'b', 'h', 1, 1},  {3, 2, 0, 1, 0, 4, 5, 0, 'p', 'h', 'b', 1, 1},
    {3, 1, 0, 2, 0, 4, 5, 1, 'p', 'b', 'h', -1, 2}, {2, 2, 1, 1, 0, 4, 5, 0, 'p', 'h', 'b', -1, 2},
    {2, 2, 5, 1, 0, 1, 5, 0, 'p', 'h', 'b', 1, 2},  {3, 5, 5, 1, 1, 0, 2, 0, 'b', 'h', 'p', 1, 2},
    {3, 2, 5, 1, 0, 0, 5, 0, 'p', 'h', 'b', -1, 1}, {2, 5, 5, 1, 1, 1, 2, 0, 'b', 'h', 'p', -1, 1},
};

// kBwd products computed by the fused family ppb instead of the rounds: those
// into the pp blocks, and the L_pp^T / R_pp^T products into the pb and ph blocks
bool fused_pp(const BwdProd& x) {
  return x.out_g == 0 || (x.a_slot <= 1 && x.a_g == 0 && x.ta == 1 && x.b_slot == 4 && (x.out_g == 1 || x.out_g == 2));
}

int ensure_adj(imsrg_basis_s* B) {
  if (B->adj_ready) return 0;
  int rc = ensure_dev(B);
  if (rc) return rc;
  const Basis& b = B->b;
  const std::vector<int> order = channel_order(b);
  static const int rcls[6] = {0, 0, 0, 1, 2, 1}, ccls[6] = {0, 1, 2, 1, 2, 2};
  for (GemmFamily& f : B->pb) { f.probs_h.clear(); f.nks.clear(); f.group_end.clear(); f.tile = 0; }
  GemmFamily& pp = B->ppb;
  pp.probs_h.clear(); pp.nks.clear(); pp.group_end.clear(); pp.tile = 0;
  for (int q : order) {
    const int dd[3] = {b.dims[q * 3], b.dims[q * 3 + 1], b.dims[q * 3 + 2]};
    auto dim = [&](char c) { return c == 'p' ? dd[0] : (c == 'b' ? dd[1] : dd[2]); };
    int used[4][6] = {{0}};
    for (const BwdProd& x : kBwd) {
      const int m = dim(x.m), n = dim(x.n), k = dim(x.k);
      if (m == 0 || n == 0 || k == 0) continue;
      if (fused_pp(x)) continue;    // in ppb below
      const int r = used[x.out_slot][x.out_g]++;
      GemmProb p{};
      p.a1 = b.goff[q * 6 + x.a_g]; p.b1 = b.goff[q * 6 + x.b_g]; p.c = b.goff[q * 6 + x.out_g];
      p.m = m; p.n = n; p.k = k;
      p.lda = std::max(dd[rcls[x.a_g]], 1); p.ldb = std::max(dd[rcls[x.b_g]], 1); p.ldc = std::max(dd[rcls[x.out_g]], 1);
      p.nprod = 1; p.ta1 = (unsigned char)x.ta; p.tb1 = (unsigned char)x.tb;
      p.sa1 = (unsigned char)x.a_slot; p.sb1 = (unsigned char)x.b_slot; p.sc = (unsigned char)x.out_slot;
      p.alpha = x.alpha; p.beta = 1.0; p.mode = (unsigned char)(4 | x.mode);
      B->pb[r].probs_h.push_back(p);
      B->pb[r].nks.push_back(1);
    }
    (void)ccls;
    for (GemmFamily& f : B->pb) f.group_end.push_back((int)f.probs_h.size());
    // The six kBwd products into the pp blocks of Lb and Rb as one problem
    // per output.  The row-p blocks pp | pb | ph of an operator are one contiguous
    // np x (np + nb + nh) column-major matrix X_p (ld np), and with the
    // herm symmetry R_pp^T = Rh R_pp, L_pp^T = Lh L_pp of the operators:
    //   Lb_pp += Wb1_p R_p^T            = Rh (Wb1_pp R_pp + Rh Wb1_x R_x^T)
    //   Rb_pp += L_pp^T Wb1_pp - Wb1_x L_x^T = Lh (L_pp Wb1_pp - Lh Wb1_x L_x^T)
    // (x: the pb | ph columns, K = nb + nh): the first product untransposed,
    // the second with op(B) transposed, the pp output read and written once.
    const int np = dd[0], nx = dd[1] + dd[2];
    if (np > 0) {
      for (int side = 0; side < 2; ++side) {
        GemmProb p{};
        const i64 o0 = b.goff[q * 6 + 0], o1 = b.goff[q * 6 + 1];
        p.a1 = o0; p.b1 = o0; p.c = o0;
        p.sa1 = side ? 0 : 4; p.sb1 = side ? 4 : 1; p.sc = side ? 3 : 2;
        p.m = np; p.n = np; p.k = np; p.lda = np; p.ldb = np; p.ldc = np;
        p.ta1 = 0; p.tb1 = 0;
        p.nprod = 1;
        if (nx > 0) {
          p.nprod = 2; p.k2 = nx; p.a2 = o1; p.b2 = o1; p.ta2 = 0; p.tb2 = 1;
          p.sa2 = 4; p.sb2 = side ? 0 : 1;
        }
        // launched with ascale = Lh, s2scale = Rh: Lb al = Rh, second scale Rh;
        // Rb al = Lh, second scale -Lh (bit6)
        p.alpha = 1.0; p.beta = 1.0;
        p.s2 = side ? -1.0 : 1.0;
        p.mode = (unsigned char)(side ? (4 | 1 | 64) : (4 | 2));
        pp.probs_h.push_back(p);
        pp.nks.push_back(1);
      }
    }
    // The transposed products that stream a whole pp block for the pb and
    // ph outputs (kBwd's R_pp^T Wb1_pb, R_pp^T Wb1_ph and the L_pp ones):
    //   Lb_x += -R_pp^T Wb1_x = -Rh R_pp Wb1_x,  Rb_x += L_pp^T Wb1_x = Lh L_pp Wb1_x
    // with x = pb | ph (one np x (nb + nh) block, ld np): one product each.
    if (np > 0 && nx > 0) {
      for (int side = 0; side < 2; ++side) {
        GemmProb p{};
        const i64 o0 = b.goff[q * 6 + 0], o1 = b.goff[q * 6 + 1];
        p.a1 = o0; p.b1 = o1; p.c = o1;
        p.sa1 = side ? 0 : 1; p.sb1 = 4; p.sc = side ? 3 : 2;
        p.m = np; p.n = nx; p.k = np; p.lda = np; p.ldb = np; p.ldc = np;
        p.nprod = 1;
        p.alpha = side ? 1.0 : -1.0; p.beta = 1.0;
        p.mode = (unsigned char)(side ? (4 | 1) : (4 | 2));
        pp.probs_h.push_back(p);
        pp.nks.push_back(1);
      }
    }
    pp.group_end.push_back((int)pp.probs_h.size());
  }
  // W' transpose per cross-coupled block, U the upper-triangular cotangent
  // of W' = Lh LCC RCC^T - Rh RCC LCC^T:
  //   LCCb += Lh U RCC - Rh U^T RCC = Lh (U RCC - Rh Lh U^T RCC),
  //   RCCb += -Rh U LCC + Lh U^T LCC = -Rh (U LCC - Rh Lh U^T LCC)
  // (ascale = Lh and s2scale = -Rh Lh at launch: al = Lh, resp. Lh s2scale;
  // both sides take U then U^T, so one kernel specialised on the transposes
  // runs the family; triangular A limits each tile's K)
  B->phb.probs_h.clear(); B->phb.nks.clear(); B->phb.group_end.clear(); B->phb.tile = 0;
  for (int q = 0; q < b.ncc; ++q) {
    const int r = b.rlen[q], nb = b.ncol[q];
    if (r > 0 && nb > 0) {
      for (int side = 0; side < 2; ++side) {
        GemmProb p{};
        p.a1 = b.wcoff[q]; p.b1 = b.ccoff[q]; p.a2 = b.wcoff[q]; p.b2 = b.ccoff[q]; p.c = b.ccoff[q];
        p.m = r; p.n = nb; p.k = r;
        p.lda = r; p.ldb = r; p.ldc = r;
        p.nprod = 2; p.ta1 = 0; p.tb1 = 0; p.ta2 = 1; p.tb2 = 0;
        p.sa1 = 4; p.sb1 = side ? 0 : 1; p.sa2 = 4; p.sb2 = side ? 0 : 1; p.sc = side ? 3 : 2;
        p.alpha = 1.0; p.s2 = 1.0; p.beta = 1.0; p.mode = (unsigned char)(1 | 4 | 8 | 16 | (side ? 2 : 0));
        B->phb.probs_h.push_back(p);
        B->phb.nks.push_back(1);
      }
    }
    B->phb.group_end.push_back((int)B->phb.probs_h.size());
  }
  B->seg_pv.host.clear();
  for (const SegTask& t : B->seg_full.host)
    if (t.g != 0 && t.g != 4) B->seg_pv.host.push_back(t);
  B->seg_pv.n = (int)B->seg_pv.host.size();
  if ((rc = upload(B, B->seg_pv.host, &B->seg_pv.dev))) return rc;
  // 222_ph transpose: lanes, tiles and the terms of each tile
  {
    B->phq_h.clear(); B->phqs_h.clear(); B->phl_h.clear(); B->phtile_h.clear();
    std::vector<int> tbase(b.ncc + 1, 0);
    for (int q = 0; q < b.ncc; ++q) {
      const int nt = (b.rlen[q] + 63) / 64;
      tbase[q] = (int)B->phtile_h.size();
      for (int tj = 0; tj < nt; ++tj)
        for (int ti = 0; ti <= tj; ++ti) B->phtile_h.push_back(PairTask{q, ti, tj, 0});
    }
    const int N = b.N, nJ = b.nJ, JTM = b.JTM;
    std::vector<int> tid;   // tile of each term, in term order
    std::vector<int2> term;
    auto rmap = [&](int x, int y, int JP) { return b.rmap[((size_t)x * N + y) * nJ + JP / 2]; };
    auto add = [&](int lane, int qx, int x, int y, bool neg) {
      if (x < 0 || y < 0) return false;
      int bits = neg ? 4096 : 0;
      if (x > y) { std::swap(x, y); bits |= 8192; }
      const int ti = x / 64, tj = y / 64;
      tid.push_back(tbase[qx] + tj * (tj + 1) / 2 + ti);
      term.push_back(int2{lane, bits | ((x - 64 * ti) + 64 * (y - 64 * tj))});
      return true;
    };
    bool ok = true;
    for (size_t x = 0; x < B->quad_h.size(); ++x) {
      if (x > 0 && B->quad_h[x].ijkl == B->quad_h[x - 1].ijkl) continue;
      const int u = B->quad_h[x].ijkl;
      const int i = u & 255, j = (u >> 8) & 255, k = (u >> 16) & 255, l = (u >> 24) & 255;
      const int ji = b.jj[i], jj = b.jj[j], jk = b.jj[k], jl = b.jj[l];
      const int a1 = std::max(std::abs(jj - jl), std::abs(ji - jk)), b1 = std::min(jj + jl, ji + jk);
      const int a2 = std::max(std::abs(ji - jl), std::abs(jj - jk)), b2 = std::min(ji + jl, jj + jk);
      const int n1 = b1 >= a1 ? (b1 - a1) / 2 + 1 : 0, n2 = b2 >= a2 ? (b2 - a2) / 2 + 1 : 0;
      const int qd = (int)B->phq_h.size();
      if (n1 + n2 > 64 || qd >= (1 << 25)) return fail(IMSRG_EINVAL, "222_ph transpose lane index overflow");
      const bool neg = ((ji + jj + jk + jl) / 2) % 2 != 0;
      const int T1 = std::abs(b.tz[i] - b.tz[k]) / 2, P1 = (b.ll[i] + b.ll[k]) % 2;
      const int T2 = std::abs(b.tz[i] - b.tz[l]) / 2, P2 = (b.ll[i] + b.ll[l]) % 2;
      B->phq_h.push_back(u);
      B->phqs_h.push_back((int)x);
      for (int sl = 0; sl < n1 + n2; ++sl) {
        const int lane = (int)B->phl_h.size();
        B->phl_h.push_back(qd * 64 + sl);
        if (sl < n1) {
          const int JP = a1 + 2 * sl, qx = JP / 2 + T1 * (JTM + 1) + 2 * P1 * (JTM + 1);
          ok = ok && add(lane, qx, rmap(k, i, JP), rmap(j, l, JP), false);
          ok = ok && add(lane, qx, rmap(l, j, JP), rmap(i, k, JP), neg);
        } else {
          const int JP = a2 + 2 * (sl - n1), qx = JP / 2 + T2 * (JTM + 1) + 2 * P2 * (JTM + 1);
          ok = ok && add(lane, qx, rmap(l, i, JP), rmap(j, k, JP), false);
          ok = ok && add(lane, qx, rmap(k, j, JP), rmap(i, l, JP), neg);
        }
      }
    }
    if (!ok) return fail(IMSRG_EINVAL, "222_ph transpose: a read outside the cross-coupled rows");
    const size_t nt = B->phtile_h.size();
    B->pht0_h.assign(nt + 1, 0);
    for (int t : tid) ++B->pht0_h[t + 1];
    for (size_t t = 0; t < nt; ++t) B->pht0_h[t + 1] += B->pht0_h[t];
    B->phterm_h.resize(term.size());
    std::vector<int> fillp(B->pht0_h.begin(), B->pht0_h.end() - 1);
    for (size_t e = 0; e < term.size(); ++e) B->phterm_h[fillp[tid[e]]++] = term[e];
  }
  // tile shapes of the transposed products (imsrg_set_gemm_tile families
  // IMSRG_GEMM_ADJ / IMSRG_GEMM_ADJ_PH)
  for (GemmFamily& f : B->pb) {
    if ((rc = upload(B, f.probs_h, &f.probs))) return rc;
    f.xcd_aware = (B->tile_req[4] & IMSRG_TILE_XCD) != 0;
    if ((rc = family_tasks(B, f, B->tile_req[4] & 7))) return rc;
  }
  if ((rc = upload(B, B->ppb.probs_h, &B->ppb.probs))) return rc;
  B->ppb.xcd_aware = (B->tile_req[6] & IMSRG_TILE_XCD) != 0;
  if ((rc = family_tasks(B, B->ppb, B->tile_req[6] & 7))) return rc;
  if ((rc = upload(B, B->phb.probs_h, &B->phb.probs))) return rc;
  B->phb.xcd_aware = (B->tile_req[5] & IMSRG_TILE_XCD) != 0;
  if ((rc = family_tasks(B, B->phb, B->tile_req[5] & 7))) return rc;
  if ((rc = upload(B, B->phq_h, &B->phq)) || (rc = upload(B, B->phqs_h, &B->phqs)) ||
      (rc = upload(B, B->phl_h, &B->phl)) || (rc = upload(B, B->phtile_h, &B->phtile)) ||
      (rc = upload(B, B->pht0_h, &B->pht0)) || (rc = upload(B, B->phterm_h, &B->phterm)))
    return rc;
  const size_t op = (size_t)std::max<i64>(b.oplen, 1) * sizeof(double);
  const size_t cc = (size_t)std::max<i64>(b.cclen, 1) * sizeof(double);
  if ((rc = dalloc(B, op, (void**)&B->Wb1)) || (rc = dalloc(B, op, (void**)&B->Wb2)) ||
      (rc = dalloc(B, op, (void**)&B->cbar)) || (rc = dalloc(B, op, (void**)&B->Y)) ||
      (rc = dalloc(B, (size_t)std::max<i64>(b.wclen, 1) * sizeof(double), (void**)&B->U)) ||
      (rc = dalloc(B, cc, (void**)&B->aLCC)) || (rc = dalloc(B, cc, (void**)&B->aRCC)) ||
      (rc = dalloc(B, cc, (void**)&B->aLCCb)) || (rc = dalloc(B, cc, (void**)&B->aRCCb)) ||
      (rc = dalloc(B, (size_t)b.N * b.N * sizeof(double), (void**)&B->Mb)) ||
      (rc = dalloc(B, std::max<size_t>(B->quad_h.size(), 1) * sizeof(double), (void**)&B->SB)) ||
      (rc = dalloc(B, std::max<size_t>(B->phl_h.size(), 1) * sizeof(double), (void**)&B->CV)))
    return rc;
  // the transpose stores every upper tile of U; the tiles below stay zero,
  // as the triangular GEMMs read the lower half of their diagonal tiles
  HIPCHK(hipMemset(B->U, 0, (size_t)std::max<i64>(b.wclen, 1) * sizeof(double)));
  B->adj_ready = true;
  return 0;
}

// Lb += d<RESb, [L,R]>/dL, Rb += d<RESb, [L,R]>/dR (the term sequence of
// t_full; with e0 also 110 + 220).  LCCin: the cross-coupled L if already
// formed; LCCb_acc: accumulate L's cross-coupled cotangent there instead of
// transposing the Pandya transform here (the caller does it once).
void comm_vjp(imsrg_basis_s* B, const double* L, int Lh, const double* R, int Rh, const double* RESb, int RESh,
              double* Lb, double* Rb, bool e0, const double* LCCin, double* LCCb_acc, hipStream_t s) {
  const Basis& b = B->b;
  const DevBasis& D = B->D;
  const int pm = Lh * Rh;
  // The 222_ph chain (Pandya forms, scatter transpose, W' transpose
  // products) and the dense 122 cotangent write only their own buffers until
  // the final gathers, so they run on auxiliary streams; every accumulation
  // into Lb / Rb stays on s in the order below (identical results).
  const bool fork = B->term_streams && B->aux_ready;
  hipStream_t sph = fork ? B->aux[1] : s, s122 = fork ? B->aux[2] : s;
  if (fork) {
    (void)hipEventRecord(B->tev[0], s);
    (void)hipStreamWaitEvent(sph, B->tev[0], 0);
    (void)hipStreamWaitEvent(s122, B->tev[0], 0);
  }
  launch_pandya(D, B->cc.dev, B->cc.n, R, Rh, B->aRCC, sph);
  const double* LCC = LCCin;
  if (!LCC) {
    launch_pandya(D, B->cc.dev, B->cc.n, L, Lh, B->aLCC, sph);
    LCC = B->aLCC;
  }
  launch_ph_vjp(D, B->quad, (int)B->quad_h.size(), B->phl, (int)B->phl_h.size(), B->phq, B->phqs, B->phtile,
                (int)B->phtile_h.size(), B->pht0, B->phterm, RESb, RESh, Lh, Rh, B->SB, B->CV, B->U, sph);
  double* LCCb = LCCb_acc ? LCCb_acc : B->aLCCb;
  if (!LCCb_acc) launch_fill(B->aLCCb, b.cclen, 0.0, sph);
  launch_fill(B->aRCCb, b.cclen, 0.0, sph);
  Slots sc{};
  sc.p[0] = LCC; sc.p[1] = B->aRCC; sc.p[2] = LCCb; sc.p[3] = B->aRCCb; sc.p[4] = B->U;
  if (B->phb.tasks.n)
    launch_gemm_tiled(B->phb.tile, sc, B->phb.probs, B->phb.tasks.dev, B->phb.tasks.grid, B->phb.tasks.xmap, sph,
                      (double)Lh, -(double)(Rh * Lh), 1, B->phb.sig);
  launch_fill(B->cbar, b.off_fpp + (i64)b.P * b.P, 0.0, s122);
  launch_to_dense(D, B->seg_full.dev, B->seg_full.n, RESb, RESh, B->cbar, s122);
  if (fork) {
    (void)hipEventRecord(B->tev[1], sph);
    (void)hipEventRecord(B->tev[2], s122);
  }
  if (e0) launch_e0_vjp(D, B->seg_g3.dev, B->seg_g3.n, L, Lh, R, Rh, RESb, Lb, Rb, s);
  launch_111_vjp(D, L, Lh, R, Rh, RESb, Lb, Rb, s);
  launch_121_vjp(D, B->one, (int)B->one_h.size(), L, Lh, R, Rh, RESb, RESh, Lb, Rb, s);
  // 222_pp_hh and 221: cotangents of the w1 / w2 products, then the
  // transposed products in rounds of distinct outputs
  launch_pphh_wbar(D, B->seg_full.dev, B->seg_full.n, RESb, pm, B->Wb1, B->Wb2, s);
  launch_221_vjp(D, B->one, (int)B->one_h.size(), Lh, Rh, RESb, RESh, B->Wb1, B->Wb2, s);
  Slots sp{};
  sp.p[0] = L; sp.p[1] = R; sp.p[2] = Lb; sp.p[3] = Rb; sp.p[4] = B->Wb1; sp.p[5] = B->Wb2;
  if (B->ppb.tasks.n)
    launch_gemm_tiled(B->ppb.tile, sp, B->ppb.probs, B->ppb.tasks.dev, B->ppb.tasks.grid, B->ppb.tasks.xmap, s,
                      (double)Lh, (double)Rh, 1, B->ppb.sig);
  for (GemmFamily& f : B->pb)
    if (f.tasks.n) launch_gemm_tiled(f.tile, sp, f.probs, f.tasks.dev, f.tasks.grid, f.tasks.xmap, s, (double)Lh,
                                     (double)Rh, 1, f.sig);
  // 122: two-body parts from the symmetrised cotangent (densified on
  // s122), one-body part by the partner sums
  if (fork) (void)hipStreamWaitEvent(s, B->tev[2], 0);
  launch_122_vjp(D, B->seg_tri.dev, B->seg_tri.n, B->partners, B->partT, B->pstride, B->rowp, B->colp, B->f1off,
                 B->nf1, L, Lh, R, Rh,
                 B->cbar, RESb, RESh, Lb, Rb, s);
  // 222_ph: the Pandya transposes of the cross-coupled cotangents (the
  // scatter transpose into the W' cotangent U and the triangular-U GEMMs ran
  // on sph)
  if (fork) (void)hipStreamWaitEvent(s, B->tev[1], 0);
  launch_pandya_vjp_gather(D, B->seg_pv.dev, B->seg_pv.n, B->aRCCb, Rh, Rb, s);
  if (!LCCb_acc) launch_pandya_vjp_gather(D, B->seg_pv.dev, B->seg_pv.n, B->aLCCb, Lh, Lb, s);
}

}  // namespace

// Nested-commutator terms of one BCH / Magnus expansion kept for its
// transpose: A[0] is the expanded operator, A[k] = [Omega, A[k-1]], c[k] the
// coefficient of A[k] in the sum.
struct TermRec {
  std::vector<const double*> A;
  std::vector<double> c;
};

// =================================================================== flow
struct imsrg_flow_s {
  imsrg_basis_s* B = nullptr;
  Exec X;                                // execution context of the flow's terms
  double *buf[2] = {nullptr, nullptr};   // INT2 / AD ping-pong
  double *w1 = nullptr, *w2 = nullptr, *GCC = nullptr, *ADCC = nullptr, *WCC = nullptr;
  double *int1 = nullptr, *Hbare = nullptr;
  double* coefs = nullptr;   // device coefficient table for graph terms
  bool graphs = false;
  int generator = 0;   // IMSRG_GEN_WHITE / _ATAN / _W2 for magnus_decouple and magnus_step
  struct GraphKey {
    int kind, iw;
    const void *a, *b, *c;
    int h1, h2;
    bool operator<(const GraphKey& o) const {
      return std::tie(kind, iw, a, b, c, h1, h2) < std::tie(o.kind, o.iw, o.a, o.b, o.c, o.h1, o.h2);
    }
  };
  std::map<GraphKey, hipGraphExec_t> execs;
  int plan_ver = -1;   // the basis plan the cached graphs were captured from
  std::vector<void*> allocs;
  // reverse mode: nested-commutator terms of one BCH / Magnus expansion,
  // cotangent ping-pong, Omega's cross-coupled cotangent, and the flow
  // gradient's checkpoints and segment states
  std::vector<double*> pool;   // recorded nested-commutator terms
  double *Ab = nullptr, *Prev = nullptr, *GCCb = nullptr, *mDG = nullptr;
  std::vector<double*> ckpts, hckpts, segG, segH, segE;   // Omega and H(s) every ckpt steps; one segment's states
  std::vector<TermRec> lmrec, lbrec;   // terms of the last segment's expansions, kept by the forward sweep
  int last_n0 = -1;                    // first step of that segment (-1: none kept)
  double* decDG = nullptr;   // generator buffer of imsrg_magnus_decouple
  double* tOp = nullptr;     // result buffer of imsrg_transform_sq_op_bch, generator of imsrg_dhds
  double *vE1 = nullptr, *vH2 = nullptr, *vE1b = nullptr, *vE2b = nullptr, *vH2b = nullptr;   // w2 transpose
  bool fw_ok = false;        // a checkpointed forward sweep is held
  int fw_steps = 0, fw_ckpt = 1;
  double fw_ds = 0.5, fw_bch = 1e-6, fw_mag = 1e-6;
  double *gH = nullptr, *gHS = nullptr, *gG = nullptr, *gDG = nullptr, *gHSb = nullptr, *gGb = nullptr,
         *gHb = nullptr, *gdOb = nullptr, *getab = nullptr;
  // many-instance batch (imsrg_batch_create): ninst instances back to back;
  // X masks by the instances still flowing (actf), Xe by those still adding
  // terms to the current BCH / Magnus expansion (acte)
  bool batch = false;
  int ninst = 1;
  Exec Xe;
  int *actf = nullptr, *acte = nullptr;      // device masks
  int *hmf = nullptr, *hme = nullptr;        // pinned staging of the masks
  double* hv = nullptr;                      // pinned per-instance scalars (kDStride each)
  std::vector<int> hf;                       // host copy of the flow mask
  ~imsrg_flow_s() {
    for (auto& kv : execs) (void)hipGraphExecDestroy(kv.second);
    for (void* p : allocs) (void)hipFree(p);
    if (hmf) (void)hipHostFree(hmf);
    if (hme) (void)hipHostFree(hme);
    if (hv) (void)hipHostFree(hv);
  }
};

namespace {

int falloc(imsrg_flow_s* F, i64 n, double** out) {
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<i64>(n, 1) * sizeof(double)));
  F->allocs.push_back(p);
  *out = (double*)p;
  return 0;
}

// Run `body` on stream s, through a cached hipGraph when graphs are on.
template <class Body>
int run_term(imsrg_flow_s* F, const imsrg_flow_s::GraphKey& key, hipStream_t s, Body&& body) {
  if (!F->graphs || F->B->path != IMSRG_PATH_GROUPED) {
    body(s);
    HIPCHK(hipGetLastError());
    return 0;
  }
  if (F->plan_ver != F->B->plan_ver) {   // task lists were rebuilt: drop graphs that point at the old ones
    for (auto& kv : F->execs) (void)hipGraphExecDestroy(kv.second);
    F->execs.clear();
    F->plan_ver = F->B->plan_ver;
  }
  auto it = F->execs.find(key);
  if (it == F->execs.end()) {
    hipStream_t cs;
    HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    body(cs);
    HIPCHK(hipStreamEndCapture(cs, &g));
    hipGraphExec_t ex;
    HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    HIPCHK(hipStreamDestroy(cs));
    it = F->execs.emplace(key, ex).first;
  }
  HIPCHK(hipGraphLaunch(it->second, s));
  return 0;
}

int pool_get(imsrg_flow_s* F, size_t& idx, double** out) {
  while (F->pool.size() <= idx) {
    double* p = nullptr;
    if (int rc = falloc(F, F->B->b.oplen, &p)) return rc;
    F->pool.push_back(p);
  }
  *out = F->pool[idx++];
  return 0;
}

int read_scalar(imsrg_basis_s* B, const double* dev, double* out, hipStream_t s) {
  HIPCHK(hipMemcpyAsync(B->hval, dev, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *out = B->hval[0];
  return 0;
}

// BCH_EXPAND, IMSRG_MAGNUS.f90:149-234
int bch(imsrg_flow_s* F, double* HS, const double* G, int Gh, const double* H, int Hh, double conv, int maxit,
        int* nterms, double* advals, hipStream_t s, TermRec* rec = nullptr, size_t* pidx = nullptr) {
  imsrg_basis_s* B = F->B;
  const i64 n = B->b.oplen;
  if (rec) {   // keep every term (for a transpose of this expansion)
    rec->A.assign(1, H);
    rec->c.assign(1, 1.0);
  }
  HIPCHK(hipMemcpyAsync(HS, H, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(F->buf[0], H, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  // G is constant over the expansion: its cross-coupled form once
  t_pandya(F->X, G, Gh, F->GCC, s);
  double coef = 1.0;
  int iw;
  for (iw = 2; iw <= maxit; ++iw) {
    coef = coef / (iw - 1.0);
    double* AD = F->buf[iw % 2];
    double* INT2 = F->buf[(iw + 1) % 2];
    const double c = coef;
    imsrg_flow_s::GraphKey key{0, iw, HS, G, H, Gh, Hh};
    int rc = run_term(F, key, s, [&](hipStream_t cs) {
      t_full(F->X, G, Gh, AD, Hh, INT2, Hh, true, F->w1, F->w2, F->GCC, F->ADCC, F->WCC, cs);
      t_axpy_norm(F->X, HS, INT2, c, B->dval, cs);
    });
    if (rc) return rc;
    if (rec) {
      double* keep = nullptr;
      if ((rc = pool_get(F, *pidx, &keep))) return rc;
      HIPCHK(hipMemcpyAsync(keep, INT2, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      rec->A.push_back(keep);
      rec->c.push_back(coef);
    }
    double nrm = 0.0;
    if ((rc = read_scalar(B, B->dval, &nrm, s))) return rc;
    const double adv = nrm * coef;
    if (advals) advals[iw - 2] = adv;
    if (adv < conv) break;
  }
  if (nterms) *nterms = iw;
  if (iw > maxit) return fail(IMSRG_ECLIP, "clipping BCH");
  return 0;
}

// MAGNUS_EXPAND, IMSRG_MAGNUS.f90:620-691
int magnus(imsrg_flow_s* F, double* DG, int DGh, const double* G, int Gh, double conv, int* ncomm,
           hipStream_t s, TermRec* rec = nullptr, size_t* pidx = nullptr, const double* eta = nullptr) {
  imsrg_basis_s* B = F->B;
  const i64 n = B->b.oplen;
  const double cof[7] = {1.0, -0.5, .0833333333, 0.0, -0.00138888888, 0.0, 3.306878e-5};
  double advals[7] = {0, 0, 0, 0, 0, 0, 0};
  int rc;
  // INT2 = DG (both ping-pong buffers carry DG's E0, which no term writes)
  HIPCHK(hipMemcpyAsync(F->buf[0], DG, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(F->buf[1], DG, sizeof(double), hipMemcpyDeviceToDevice, s));
  t_norm(F->X, DG, B->dval, s);
  t_norm(F->X, G, B->dval + 1, s);
  HIPCHK(hipMemcpyAsync(B->hval, B->dval, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  advals[0] = B->hval[0];
  const double fullnorm = B->hval[1];
  if (ncomm) *ncomm = 0;
  if (rec) {   // keep every term; eta = a copy of DG on entry
    rec->A.assign(1, eta);
    rec->c.assign(1, 1.0);
  }
  if (fullnorm < (double)1e-9f) return 0;
  t_pandya(F->X, G, Gh, F->GCC, s);
  int q = 1, cnt = 0;
  for (int i = 2; i <= 7; ++i) {
    double* AD = F->buf[i % 2];
    double* INT2 = F->buf[(i + 1) % 2];
    const double adnorm = advals[i - q - 1];
    q = (std::fabs(cof[i - 1]) > (double)1e-6f) ? 1 : 2;
    if (std::fabs(adnorm / fullnorm) < conv) break;
    const double c = cof[i - 1];
    imsrg_flow_s::GraphKey key{1, i, DG, G, nullptr, Gh, DGh};
    rc = run_term(F, key, s, [&](hipStream_t cs) {
      t_full(F->X, G, Gh, AD, DGh, INT2, DGh, false, F->w1, F->w2, F->GCC, F->ADCC, F->WCC, cs);
      t_axpy_norm(F->X, DG, INT2, c, B->dval, cs);
    });
    if (rc) return rc;
    ++cnt;
    if (rec) {
      double* keep = nullptr;
      if ((rc = pool_get(F, *pidx, &keep))) return rc;
      HIPCHK(hipMemcpyAsync(keep, INT2, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      rec->A.push_back(keep);
      rec->c.push_back(c);
    }
    double nrm = 0.0;
    if ((rc = read_scalar(B, B->dval, &nrm, s))) return rc;
    advals[i - 1] = nrm * std::fabs(cof[i - 1]);
  }
  if (ncomm) *ncomm = cnt;
  return 0;
}

// generator with the denominators of Hd and the numerators of Hn (mode 0
// white, 1 atan / atanxx)
int gen(const Exec& X, const double* Hd, const double* Hn, double* ETA, int mode, hipStream_t s) {
  launch_monopole(X.D, Hd, X.M, s);
  launch_gen_1b(X.D, Hd, Hn, ETA, mode, s);
  launch_gen_2b(X.D, X.B->seg_full.dev, X.B->seg_full.n, Hd, X.M, Hn, ETA, mode, s);
  HIPCHK(hipGetLastError());
  return 0;
}

// build_gs_white, generators.f90:4-106
int white(const Exec& X, const double* H, double* ETA, hipStream_t s) { return gen(X, H, H, ETA, 0, s); }

int mbpt2_dev(const Exec& X, const double* H, double* dst, hipStream_t s) {
  launch_mbpt2_partials(X.D, X.B->seg_g3.dev, X.B->seg_g3.n, H, X.part, s);
  launch_sum_sqrt_b(X.D, X.part, X.B->seg_g3.n, dst, kDStride, 0, s);
  HIPCHK(hipGetLastError());
  return 0;
}

// build_gs_w2, generators.f90:109-160: eta1 = atan(H), H2 = [eta1, H]
// (E0 included), eta = 1.5 eta1 + 0.5 atanxx(H, H2).  Uses the flow's
// ping-pong operators and commutator workspace (free between expansions).
int gen_w2(imsrg_flow_s* F, const Exec& X, const double* H, double* ETA, hipStream_t s) {
  const i64 n = F->B->b.oplen;
  double* eta1 = F->buf[0];
  double* H2 = F->buf[1];
  int rc;
  // eta1 starts from zero as duplicate_sq_op leaves it: the generator
  // writes fph only where a and i share (l, j, tz)
  launch_fill_b(eta1, n, 0.0, X.D.ninst, X.D.act, s);
  if ((rc = gen(X, H, H, eta1, 1, s))) return rc;
  launch_fill_e0(X.D, eta1, 0.0, s);
  t_pandya(X, eta1, -1, F->GCC, s);
  t_full(X, eta1, -1, H, 1, H2, 1, true, F->w1, F->w2, F->GCC, F->ADCC, F->WCC, s);
  if ((rc = gen(X, H, H2, ETA, 1, s))) return rc;
  launch_fill_e0(X.D, ETA, 0.0, s);
  launch_axpy_b(n, eta1, 1.5, ETA, 0.5, ETA, X.D.ninst, X.D.act, s);   // add_sq_op(eta1, 1.5, eta2, 0.5, eta)
  HIPCHK(hipGetLastError());
  return 0;
}

// the flow's generator of H into ETA
int generator(imsrg_flow_s* F, const Exec& X, const double* H, double* ETA, hipStream_t s) {
  switch (F->generator) {
    case IMSRG_GEN_ATAN: return gen(X, H, H, ETA, 1, s);
    case IMSRG_GEN_W2: return gen_w2(F, X, H, ETA, s);
    default: return white(X, H, ETA, s);
  }
}

int magnus_step(imsrg_flow_s* F, double* HS, double* G, double* DG, const double* H, double ds, double bch_conv,
                double mag_conv, double* out, hipStream_t s) {
  imsrg_basis_s* B = F->B;
  int rc, nm = 0, nb = 0;
  if ((rc = magnus(F, DG, -1, G, -1, mag_conv, &nm, s))) return rc;
  launch_axpy(B->b.oplen, G, 1.0, DG, ds, G, s);   // euler_step
  if ((rc = bch(F, HS, G, -1, H, 1, bch_conv, 40, &nb, nullptr, s))) return rc;
  if ((rc = generator(F, F->X, HS, DG, s))) return rc;
  if ((rc = mbpt2_dev(F->X, HS, B->dval + 2, s))) return rc;
  HIPCHK(hipMemcpyAsync(B->dval + 3, HS, sizeof(double), hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(B->hval, B->dval + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (out) {
    out[0] = B->hval[1];
    out[1] = B->hval[0];
    out[2] = nm;
    out[3] = nb;
  }
  return 0;
}


// ------------------------------------------------- many-instance batches
// The step-level calls over F->ninst instances at once: every launch covers
// all instances (grid.y) and skips masked ones, so each instance runs
// exactly the kernels, and takes exactly the exits, of its own flow.

// upload a mask through its pinned staging (synchronises first, so the
// staging of an earlier upload is free)
int set_mask(imsrg_flow_s* F, int* dev, int* stage, const std::vector<int>& m, hipStream_t s) {
  HIPCHK(hipStreamSynchronize(s));
  for (int k = 0; k < F->ninst; ++k) stage[k] = m[k];
  HIPCHK(hipMemcpyAsync(dev, stage, F->ninst * sizeof(int), hipMemcpyHostToDevice, s));
  return 0;
}

// per-instance scalars dval[k * kDStride + 0 .. nslot) to F->hv; with E0 of
// op (optional) in slot 3
int read_inst(imsrg_flow_s* F, int nslot, const double* op, hipStream_t s) {
  HIPCHK(hipMemcpyAsync(F->hv, F->X.dval, (size_t)F->ninst * kDStride * sizeof(double), hipMemcpyDeviceToHost, s));
  (void)nslot;
  if (op)
    HIPCHK(hipMemcpy2DAsync(F->hv + 3, kDStride * sizeof(double), op, (size_t)F->B->b.oplen * sizeof(double),
                            sizeof(double), F->ninst, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

// BCH_EXPAND (IMSRG_MAGNUS.f90:149-234) per instance of the flow mask;
// nterms[k] = the reference's iw at instance k's exit
int bch_b(imsrg_flow_s* F, double* HS, const double* G, const double* H, double conv, int maxit, int* nterms,
          hipStream_t s) {
  imsrg_basis_s* B = F->B;
  const int ni = F->ninst;
  const i64 n = B->b.oplen;
  int rc;
  std::vector<int> he = F->hf;
  if ((rc = set_mask(F, F->acte, F->hme, he, s))) return rc;
  launch_copy_b(HS, H, n, ni, F->acte, s);
  launch_copy_b(F->buf[0], H, n, ni, F->acte, s);
  t_pandya(F->Xe, G, -1, F->GCC, s);
  double coef = 1.0;
  int iw;
  for (iw = 2; iw <= maxit; ++iw) {
    coef = coef / (iw - 1.0);
    double* AD = F->buf[iw % 2];
    double* INT2 = F->buf[(iw + 1) % 2];
    const double c = coef;
    imsrg_flow_s::GraphKey key{2, iw, HS, G, H, -1, 1};
    rc = run_term(F, key, s, [&](hipStream_t cs) {
      t_full(F->Xe, G, -1, AD, 1, INT2, 1, true, F->w1, F->w2, F->GCC, F->ADCC, F->WCC, cs);
      t_axpy_norm(F->Xe, HS, INT2, c, F->Xe.dval, cs);
    });
    if (rc) return rc;
    if ((rc = read_inst(F, 1, nullptr, s))) return rc;
    bool any = false, changed = false;
    for (int k = 0; k < ni; ++k) {
      if (!he[k]) continue;
      if (F->hv[k * kDStride] * coef < conv) {
        he[k] = 0;
        if (nterms) nterms[k] = iw;
        changed = true;
      } else {
        any = true;
      }
    }
    if (!any) break;
    if (changed && (rc = set_mask(F, F->acte, F->hme, he, s))) return rc;
  }
  if (iw > maxit) return fail(IMSRG_ECLIP, "clipping BCH");
  return 0;
}

// MAGNUS_EXPAND (IMSRG_MAGNUS.f90:620-691) per instance of the flow mask:
// DG holds eta on entry and dOmega on exit; ncomm[k] nested commutators
int magnus_b(imsrg_flow_s* F, double* DG, const double* G, double conv, int* ncomm, hipStream_t s) {
  imsrg_basis_s* B = F->B;
  const int ni = F->ninst;
  const i64 n = B->b.oplen;
  const double cof[7] = {1.0, -0.5, .0833333333, 0.0, -0.00138888888, 0.0, 3.306878e-5};
  int rc;
  std::vector<int> he = F->hf, q(ni, 1);
  std::vector<double> adv((size_t)ni * 7, 0.0), full(ni, 0.0);
  // both ping-pong buffers carry DG (every term assigns all but E0)
  launch_copy_b(F->buf[0], DG, n, ni, F->actf, s);
  launch_copy_b(F->buf[1], DG, n, ni, F->actf, s);
  t_norm(F->X, DG, F->X.dval, s);
  t_norm(F->X, G, F->X.dval + 1, s);
  if ((rc = read_inst(F, 2, nullptr, s))) return rc;
  bool any = false;
  for (int k = 0; k < ni; ++k) {
    if (ncomm) ncomm[k] = 0;
    if (!he[k]) continue;
    adv[(size_t)k * 7] = F->hv[k * kDStride];
    full[k] = F->hv[k * kDStride + 1];
    if (full[k] < (double)1e-9f) he[k] = 0;
    any = any || he[k];
  }
  if (!any) return 0;
  if ((rc = set_mask(F, F->acte, F->hme, he, s))) return rc;
  t_pandya(F->Xe, G, -1, F->GCC, s);
  for (int i = 2; i <= 7; ++i) {
    bool changed = false;
    any = false;
    for (int k = 0; k < ni; ++k) {
      if (!he[k]) continue;
      const double adnorm = adv[(size_t)k * 7 + (i - q[k] - 1)];
      q[k] = (std::fabs(cof[i - 1]) > (double)1e-6f) ? 1 : 2;
      if (std::fabs(adnorm / full[k]) < conv) {
        he[k] = 0;
        changed = true;
      } else {
        any = true;
      }
    }
    if (!any) break;
    if (changed && (rc = set_mask(F, F->acte, F->hme, he, s))) return rc;
    double* AD = F->buf[i % 2];
    double* INT2 = F->buf[(i + 1) % 2];
    const double c = cof[i - 1];
    imsrg_flow_s::GraphKey key{3, i, DG, G, nullptr, -1, -1};
    rc = run_term(F, key, s, [&](hipStream_t cs) {
      t_full(F->Xe, G, -1, AD, -1, INT2, -1, false, F->w1, F->w2, F->GCC, F->ADCC, F->WCC, cs);
      t_axpy_norm(F->Xe, DG, INT2, c, F->Xe.dval, cs);
    });
    if (rc) return rc;
    if ((rc = read_inst(F, 1, nullptr, s))) return rc;
    for (int k = 0; k < ni; ++k)
      if (he[k]) {
        if (ncomm) ++ncomm[k];
        adv[(size_t)k * 7 + (i - 1)] = F->hv[k * kDStride] * std::fabs(cof[i - 1]);
      }
  }
  return 0;
}

// one Magnus step per flowing instance; E0 and mbpt2 of instance k land in
// F->hv[k * kDStride + 3] and [+ 2]
int magnus_step_b(imsrg_flow_s* F, double* HS, double* G, double* DG, const double* H, double ds, double bch_conv,
                  double mag_conv, int* nm, int* nb, hipStream_t s) {
  imsrg_basis_s* B = F->B;
  int rc;
  if ((rc = magnus_b(F, DG, G, mag_conv, nm, s))) return rc;
  launch_axpy_b(B->b.oplen, G, 1.0, DG, ds, G, F->ninst, F->actf, s);   // euler_step
  if ((rc = bch_b(F, HS, G, H, bch_conv, 40, nb, s))) return rc;
  if ((rc = generator(F, F->X, HS, DG, s))) return rc;
  if ((rc = mbpt2_dev(F->X, HS, F->X.dval + 2, s))) return rc;
  return read_inst(F, 3, HS, s);
}

// ------------------------------------------------- reverse mode (flow)
int fbuf(imsrg_flow_s* F, double** slot, i64 n) {
  if (*slot) return 0;
  return falloc(F, n, slot);
}
int grow(imsrg_flow_s* F, std::vector<double*>& v, size_t k) {
  while (v.size() <= k) {
    double* p = nullptr;
    if (int rc = falloc(F, F->B->b.oplen, &p)) return rc;
    v.push_back(p);
  }
  return 0;
}

// BCH_EXPAND (IMSRG_MAGNUS.f90:149-234) keeping its terms; HS (optional)
// receives the sum.  Same kernels and order as bch(), so HS is identical.
int bch_record(imsrg_flow_s* F, const double* G, const double* H, int Hh, double conv, int maxit, TermRec& rec,
               size_t& pidx, double* HS, int* nterms, hipStream_t s) {
  imsrg_basis_s* B = F->B;
  const i64 n = B->b.oplen;
  int rc;
  rec.A.assign(1, H);
  rec.c.assign(1, 1.0);
  if (HS) HIPCHK(hipMemcpyAsync(HS, H, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  t_pandya(F->X, G, -1, F->GCC, s);
  double coef = 1.0;
  int iw;
  for (iw = 2; iw <= maxit; ++iw) {
    coef = coef / (iw - 1.0);
    double* nx = nullptr;
    if ((rc = pool_get(F, pidx, &nx))) return rc;
    t_full(F->X, G, -1, rec.A.back(), Hh, nx, Hh, true, F->w1, F->w2, F->GCC, F->ADCC, F->WCC, s);
    if (HS) launch_axpy(n, HS, 1.0, nx, coef, HS, s);
    t_norm(F->X, nx, B->dval, s);
    double nrm = 0.0;
    if ((rc = read_scalar(B, B->dval, &nrm, s))) return rc;
    rec.A.push_back(nx);
    rec.c.push_back(coef);
    if (nrm * coef < conv) break;
  }
  if (nterms) *nterms = iw;
  if (iw > maxit) return fail(IMSRG_ECLIP, "clipping BCH");
  return 0;
}

// transpose of a recorded BCH_EXPAND: Gb += d<HSb,HS>/dG, Hb += d<HSb,HS>/dH
int bch_reverse(imsrg_flow_s* F, const double* G, int Hh, const TermRec& rec, const double* HSb, double* Gb,
                double* Hb, hipStream_t s) {
  imsrg_basis_s* B = F->B;
  const i64 n = B->b.oplen;
  int rc;
  if ((rc = fbuf(F, &F->Ab, n)) || (rc = fbuf(F, &F->Prev, n)) || (rc = fbuf(F, &F->GCCb, B->b.cclen))) return rc;
  const int K = (int)rec.A.size() - 1;
  launch_axpy(n, HSb, rec.c[K], HSb, 0.0, F->Ab, s);
  if (K > 0) {
    t_pandya(F->X, G, -1, F->GCC, s);
    launch_fill(F->GCCb, B->b.cclen, 0.0, s);
  }
  for (int k = K; k >= 1; --k) {
    launch_fill(F->Prev, n, 0.0, s);
    comm_vjp(B, G, -1, rec.A[k - 1], Hh, F->Ab, Hh, Gb, F->Prev, true, F->GCC, F->GCCb, s);
    launch_axpy(n, F->Prev, 1.0, HSb, rec.c[k - 1], F->Ab, s);
  }
  launch_axpy(n, Hb, 1.0, F->Ab, 1.0, Hb, s);
  if (K > 0) launch_pandya_vjp_gather(B->D, B->seg_pv.dev, B->seg_pv.n, F->GCCb, -1, Gb, s);
  HIPCHK(hipGetLastError());
  return 0;
}

// MAGNUS_EXPAND (IMSRG_MAGNUS.f90:620-691) keeping its terms: DG holds eta
// on entry (kept unchanged in rec.A[0] = ETA) and receives dOmega.  Same
// kernels, exits and order as magnus().
int magnus_record(imsrg_flow_s* F, const double* ETA, double* DG, const double* G, double conv, TermRec& rec,
                  size_t& pidx, int* ncomm, hipStream_t s) {
  imsrg_basis_s* B = F->B;
  const i64 n = B->b.oplen;
  const double cof[7] = {1.0, -0.5, .0833333333, 0.0, -0.00138888888, 0.0, 3.306878e-5};
  double advals[7] = {0, 0, 0, 0, 0, 0, 0};
  int rc;
  rec.A.assign(1, ETA);
  rec.c.assign(1, 1.0);
  if (DG != ETA) HIPCHK(hipMemcpyAsync(DG, ETA, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  t_norm(F->X, ETA, B->dval, s);
  t_norm(F->X, G, B->dval + 1, s);
  HIPCHK(hipMemcpyAsync(B->hval, B->dval, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  advals[0] = B->hval[0];
  const double fullnorm = B->hval[1];
  if (ncomm) *ncomm = 0;
  if (fullnorm < (double)1e-9f) return 0;
  t_pandya(F->X, G, -1, F->GCC, s);
  int q = 1;
  for (int i = 2; i <= 7; ++i) {
    const double adnorm = advals[i - q - 1];
    q = (std::fabs(cof[i - 1]) > (double)1e-6f) ? 1 : 2;
    if (std::fabs(adnorm / fullnorm) < conv) break;
    double* nx = nullptr;
    if ((rc = pool_get(F, pidx, &nx))) return rc;
    HIPCHK(hipMemcpyAsync(nx, rec.A.back(), sizeof(double), hipMemcpyDeviceToDevice, s));
    t_full(F->X, G, -1, rec.A.back(), -1, nx, -1, false, F->w1, F->w2, F->GCC, F->ADCC, F->WCC, s);
    launch_axpy(n, DG, 1.0, nx, cof[i - 1], DG, s);
    t_norm(F->X, nx, B->dval, s);
    double nrm = 0.0;
    if ((rc = read_scalar(B, B->dval, &nrm, s))) return rc;
    advals[i - 1] = nrm * std::fabs(cof[i - 1]);
    rec.A.push_back(nx);
    rec.c.push_back(cof[i - 1]);
  }
  if (ncomm) *ncomm = (int)rec.A.size() - 1;
  return 0;
}

// transpose of a recorded MAGNUS_EXPAND: ETAb += d<DGb,dOmega>/d eta,
// Gb += d<DGb,dOmega>/dOmega (E0 of every term is eta's E0)
int magnus_reverse(imsrg_flow_s* F, const double* G, const TermRec& rec, const double* DGb, double* ETAb, double* Gb,
                   hipStream_t s) {
  imsrg_basis_s* B = F->B;
  const i64 n = B->b.oplen;
  int rc;
  if ((rc = fbuf(F, &F->Ab, n)) || (rc = fbuf(F, &F->Prev, n)) || (rc = fbuf(F, &F->GCCb, B->b.cclen))) return rc;
  const int K = (int)rec.A.size() - 1;
  double e0sum = 0.0;
  for (double c : rec.c) e0sum += c;
  launch_axpy(n, DGb, rec.c[K], DGb, 0.0, F->Ab, s);
  launch_fill(F->Ab, 1, 0.0, s);
  if (K > 0) {
    t_pandya(F->X, G, -1, F->GCC, s);
    launch_fill(F->GCCb, B->b.cclen, 0.0, s);
  }
  for (int k = K; k >= 1; --k) {
    launch_fill(F->Prev, n, 0.0, s);
    comm_vjp(B, G, -1, rec.A[k - 1], -1, F->Ab, -1, Gb, F->Prev, false, F->GCC, F->GCCb, s);
    launch_axpy(n, F->Prev, 1.0, DGb, rec.c[k - 1], F->Ab, s);
    launch_fill(F->Ab, 1, 0.0, s);
  }
  launch_axpy(n, ETAb, 1.0, F->Ab, 1.0, ETAb, s);
  launch_add_e0(ETAb, DGb, e0sum, s);
  if (K > 0) launch_pandya_vjp_gather(B->D, B->seg_pv.dev, B->seg_pv.n, F->GCCb, -1, Gb, s);
  HIPCHK(hipGetLastError());
  return 0;
}

// BCH_EXPAND transpose (IMSRG_MAGNUS.f90:149-234): Gb += d<HSb, HS>/dG,
// Hb += d<HSb, HS>/dH for HS = BCH(G, H), G antihermitian
int bch_vjp(imsrg_flow_s* F, const double* G, const double* H, int Hh, double conv, int maxit, const double* HSb,
            double* Gb, double* Hb, int* nterms, hipStream_t s) {
  int rc;
  if ((rc = ensure_adj(F->B))) return rc;
  TermRec rec;
  size_t pidx = 0;
  if ((rc = bch_record(F, G, H, Hh, conv, maxit, rec, pidx, nullptr, nterms, s))) return rc;
  return bch_reverse(F, G, Hh, rec, HSb, Gb, Hb, s);
}

// MAGNUS_EXPAND transpose (IMSRG_MAGNUS.f90:620-691), eta and Omega
// antihermitian: ETAb += d<DGb, dOmega>/d eta, Gb += d<DGb, dOmega>/dOmega
int magnus_vjp(imsrg_flow_s* F, const double* ETA, const double* G, double conv, const double* DGb, double* ETAb,
               double* Gb, int* ncomm, hipStream_t s) {
  int rc;
  if ((rc = ensure_adj(F->B))) return rc;
  double* dg = nullptr;
  if ((rc = fbuf(F, &F->mDG, F->B->b.oplen))) return rc;
  dg = F->mDG;
  TermRec rec;
  size_t pidx = 0;
  if ((rc = magnus_record(F, ETA, dg, G, conv, rec, pidx, ncomm, s))) return rc;
  return magnus_reverse(F, G, rec, DGb, ETAb, Gb, s);
}

// generator transpose with the denominators of Hd and the numerators of Hn
// (mode 0 white, 1 atan / atanxx): Hdb, Hnb += d<ETAb, gen(Hd, Hn)>/d(Hd, Hn)
int gen_vjp(imsrg_basis_s* B, const double* Hd, const double* Hn, const double* ETAb, double* Hdb, double* Hnb,
            int mode, hipStream_t s) {
  int rc;
  if ((rc = ensure_adj(B))) return rc;
  launch_monopole(B->D, Hd, B->M, s);
  launch_fill(B->Mb, (i64)B->b.N * B->b.N, 0.0, s);
  launch_gen_vjp(B->D, B->seg_g3.dev, B->seg_g3.n, Hd, Hn, B->M, ETAb, Hdb, Hnb, B->Mb, mode, s);
  HIPCHK(hipGetLastError());
  return 0;
}

// build_gs_white transpose (generators.f90:4-106): Hb += d<ETAb, white(H)>/dH
int white_vjp(imsrg_basis_s* B, const double* H, const double* ETAb, double* Hb, hipStream_t s) {
  return gen_vjp(B, H, H, ETAb, Hb, Hb, 0, s);
}

// build_gs_w2 transpose (generators.f90:109-160): ETA = 1.5 eta1 + 0.5 eta2
// with eta1 = atan(H), H2 = [eta1, H], eta2 = atanxx(H, H2); eta1 and H2 are
// recomputed (on the flow's buffers), then eta2 -> (H, H2), H2 -> (eta1, H)
// through the commutator transpose, eta1 -> H.  Hb accumulates.
int w2_vjp(imsrg_flow_s* F, const double* H, const double* ETAb, double* Hb, hipStream_t s) {
  imsrg_basis_s* B = F->B;
  const i64 n = B->b.oplen;
  int rc;
  for (double** p : {&F->vE1, &F->vH2, &F->vE1b, &F->vE2b, &F->vH2b})
    if ((rc = fbuf(F, p, n))) return rc;
  if ((rc = ensure_adj(B))) return rc;
  launch_fill(F->vE1, n, 0.0, s);
  if ((rc = gen(F->X, H, H, F->vE1, 1, s))) return rc;
  t_pandya(F->X, F->vE1, -1, F->GCC, s);
  t_full(F->X, F->vE1, -1, H, 1, F->vH2, 1, true, F->w1, F->w2, F->GCC, F->ADCC, F->WCC, s);
  launch_axpy(n, ETAb, 1.5, ETAb, 0.0, F->vE1b, s);
  launch_axpy(n, ETAb, 0.5, ETAb, 0.0, F->vE2b, s);
  launch_fill(F->vH2b, n, 0.0, s);
  if ((rc = gen_vjp(B, H, F->vH2, F->vE2b, Hb, F->vH2b, 1, s))) return rc;
  comm_vjp(B, F->vE1, -1, H, 1, F->vH2b, 1, F->vE1b, Hb, true, nullptr, nullptr, s);
  if ((rc = gen_vjp(B, H, H, F->vE1b, Hb, Hb, 1, s))) return rc;
  HIPCHK(hipGetLastError());
  return 0;
}

// transpose of the flow's generator
int generator_vjp(imsrg_flow_s* F, const double* H, const double* ETAb, double* Hb, hipStream_t s) {
  switch (F->generator) {
    case IMSRG_GEN_ATAN: return gen_vjp(F->B, H, H, ETAb, Hb, Hb, 1, s);
    case IMSRG_GEN_W2: return w2_vjp(F, H, ETAb, Hb, s);
    default: return white_vjp(F->B, H, ETAb, Hb, s);
  }
}

// normal_order transpose (basic_IMSRG.f90:2419-2550): Tb = d<Hb, NO(T)>/dT
int normal_order_vjp(imsrg_basis_s* B, const double* Hb, double* Tb, hipStream_t s) {
  HIPCHK(hipMemcpyAsync(Tb, Hb, B->b.oplen * sizeof(double), hipMemcpyDeviceToDevice, s));
  launch_normal_order_vjp(B->D, Hb, Tb, s);
  HIPCHK(hipGetLastError());
  return 0;
}

// Forward sweep of the flow gradient: magnus_decouple (the flow's
// generator, |mbpt2| > crit_tol loop) from H = normal_order(T), keeping
// Omega every `ckpt` steps in the flow object for flow_backward.
int flow_forward(imsrg_flow_s* F, const double* T, double ds, double bch_conv, double mag_conv, double crit_tol,
                 int maxsteps, int ckpt, double* E0, int* nsteps, hipStream_t s) {
  imsrg_basis_s* B = F->B;
  const i64 n = B->b.oplen;
  const size_t nb = (size_t)n * sizeof(double);
  int rc, nt = 0, nc = 0;
  if (ckpt < 0 || maxsteps < 1) return fail(IMSRG_EINVAL, "checkpoint interval must be >= 0");
  // ckpt 0: one segment holding every step's states and terms (the reverse
  // sweep recomputes nothing; the buffers grow with the steps)
  const bool keep_all = ckpt == 0;
  if (keep_all) ckpt = maxsteps;
  if ((rc = ensure_adj(B))) return rc;
  for (double** p : {&F->gH, &F->gHS, &F->gG, &F->gDG, &F->gHSb, &F->gGb, &F->gHb, &F->gdOb, &F->getab})
    if ((rc = fbuf(F, p, n))) return rc;
  const size_t pre = keep_all ? 1 : (size_t)ckpt;
  if ((rc = grow(F, F->segG, pre)) || (rc = grow(F, F->segH, pre)) || (rc = grow(F, F->segE, pre))) return rc;
  F->fw_ok = false;
  F->last_n0 = -1;
  F->lmrec.assign(ckpt, TermRec());
  F->lbrec.assign(ckpt, TermRec());
  HIPCHK(hipMemcpyAsync(F->gH, T, nb, hipMemcpyDeviceToDevice, s));
  launch_normal_order(B->D, F->gH, B->dval + 8, s);
  HIPCHK(hipMemcpyAsync(F->gHS, F->gH, nb, hipMemcpyDeviceToDevice, s));
  launch_fill(F->gG, n, 0.0, s);
  launch_fill(F->gDG, n, 0.0, s);
  if ((rc = generator(F, F->X, F->gHS, F->gDG, s)) || (rc = mbpt2_dev(F->X, F->gHS, B->dval + 2, s))) return rc;
  double crit = 0.0;
  if ((rc = read_scalar(B, B->dval + 2, &crit, s))) return rc;
  crit = std::fabs(crit);
  int steps = 0;
  size_t pidx = 0;
  // Omega and H(s) are checkpointed every ckpt steps; the states and the
  // nested-commutator terms of the current segment are kept as the reverse
  // sweep would recompute them, so the last segment is not recomputed
  while (crit > crit_tol) {
    if (steps >= maxsteps) return fail(IMSRG_ESTEPS, "flow gradient: step limit reached");
    const int k = steps % ckpt;
    if (k == 0) {
      const size_t c = (size_t)(steps / ckpt);
      if ((rc = grow(F, F->ckpts, c)) || (rc = grow(F, F->hckpts, c))) return rc;
      HIPCHK(hipMemcpyAsync(F->ckpts[c], F->gG, nb, hipMemcpyDeviceToDevice, s));
      HIPCHK(hipMemcpyAsync(F->hckpts[c], F->gHS, nb, hipMemcpyDeviceToDevice, s));
      pidx = 0;
    }
    if (keep_all && ((rc = grow(F, F->segG, (size_t)k + 1)) || (rc = grow(F, F->segH, (size_t)k)) ||
                     (rc = grow(F, F->segE, (size_t)k))))
      return rc;
    HIPCHK(hipMemcpyAsync(F->segG[k], F->gG, nb, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(F->segH[k], F->gHS, nb, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(F->segE[k], F->gDG, nb, hipMemcpyDeviceToDevice, s));
    if ((rc = magnus(F, F->gDG, -1, F->gG, -1, mag_conv, &nc, s, &F->lmrec[k], &pidx, F->segE[k]))) return rc;
    launch_axpy(n, F->gG, 1.0, F->gDG, ds, F->gG, s);
    if ((rc = bch(F, F->gHS, F->gG, -1, F->gH, 1, bch_conv, 40, &nt, nullptr, s, &F->lbrec[k], &pidx))) return rc;
    if ((rc = generator(F, F->X, F->gHS, F->gDG, s)) || (rc = mbpt2_dev(F->X, F->gHS, B->dval + 2, s))) return rc;
    if ((rc = read_scalar(B, B->dval + 2, &crit, s))) return rc;
    crit = std::fabs(crit);
    ++steps;
  }
  if (steps > 0) {
    const int n0 = ((steps - 1) / ckpt) * ckpt;
    HIPCHK(hipMemcpyAsync(F->segG[steps - n0], F->gG, nb, hipMemcpyDeviceToDevice, s));
    F->last_n0 = n0;
  }
  double e0 = 0.0;
  if ((rc = read_scalar(B, F->gHS, &e0, s))) return rc;
  F->fw_ok = true;
  F->fw_steps = steps; F->fw_ckpt = ckpt; F->fw_ds = ds; F->fw_bch = bch_conv; F->fw_mag = mag_conv;
  if (E0) *E0 = e0;
  if (nsteps) *nsteps = steps;
  return 0;
}

// Reverse sweep: each checkpoint segment (but the last one, which the
// forward sweep kept) is recomputed from its Omega and H(s) checkpoints,
// keeping every nested-commutator term, and swept backwards through
// BCH_EXPAND, euler_step, MAGNUS_EXPAND and the generator; Tb = dE0/dT.
int flow_backward(imsrg_flow_s* F, double* Tb, hipStream_t s) {
  imsrg_basis_s* B = F->B;
  const i64 n = B->b.oplen;
  int rc, nt = 0, nc = 0;
  if (!F->fw_ok) return fail(IMSRG_EINVAL, "flow gradient: no forward sweep to reverse");
  const int steps = F->fw_steps, ckpt = F->fw_ckpt;
  const double ds = F->fw_ds, bch_conv = F->fw_bch, mag_conv = F->fw_mag;
  launch_fill(F->gHSb, n, 0.0, s);
  launch_fill(F->gHSb, 1, 1.0, s);
  launch_fill(F->gGb, n, 0.0, s);
  launch_fill(F->gHb, n, 0.0, s);
  const int nck = (steps + ckpt - 1) / ckpt;
  std::vector<TermRec> mrec, brec;
  for (int seg = nck - 1; seg >= 0; --seg) {
    const int n0 = seg * ckpt, n1 = std::min(steps, n0 + ckpt);
    if (seg == nck - 1 && F->last_n0 == n0) {   // kept by the forward sweep
      mrec.assign(F->lmrec.begin(), F->lmrec.begin() + (n1 - n0));
      brec.assign(F->lbrec.begin(), F->lbrec.begin() + (n1 - n0));
    } else {
      mrec.assign(n1 - n0, TermRec());
      brec.assign(n1 - n0, TermRec());
      HIPCHK(hipMemcpyAsync(F->segG[0], F->ckpts[seg], n * sizeof(double), hipMemcpyDeviceToDevice, s));
      HIPCHK(hipMemcpyAsync(F->segH[0], F->hckpts[seg], n * sizeof(double), hipMemcpyDeviceToDevice, s));
      launch_fill(F->segE[0], n, 0.0, s);
      if ((rc = generator(F, F->X, F->segH[0], F->segE[0], s))) return rc;
      size_t pidx = 0;
      for (int m = n0; m < n1; ++m) {
        const int k = m - n0;
        if ((rc = magnus_record(F, F->segE[k], F->gDG, F->segG[k], mag_conv, mrec[k], pidx, &nc, s))) return rc;
        launch_axpy(n, F->segG[k], 1.0, F->gDG, ds, F->segG[k + 1], s);
        if ((rc = bch_record(F, F->segG[k + 1], F->gH, 1, bch_conv, 40, brec[k], pidx, F->segH[k + 1], &nt, s)))
          return rc;
        launch_fill(F->segE[k + 1], n, 0.0, s);
        if ((rc = generator(F, F->X, F->segH[k + 1], F->segE[k + 1], s))) return rc;
      }
    }
    for (int m = n1 - 1; m >= n0; --m) {
      const int k = m - n0;
      if ((rc = bch_reverse(F, F->segG[k + 1], 1, brec[k], F->gHSb, F->gGb, F->gHb, s))) return rc;
      launch_axpy(n, F->gGb, ds, F->gGb, 0.0, F->gdOb, s);
      launch_fill(F->getab, n, 0.0, s);
      if ((rc = magnus_reverse(F, F->segG[k], mrec[k], F->gdOb, F->getab, F->gGb, s))) return rc;
      launch_fill(F->gHSb, n, 0.0, s);
      if ((rc = generator_vjp(F, F->segH[k], F->getab, F->gHSb, s))) return rc;
      if (m == 0) {
        launch_axpy(n, F->gHb, 1.0, F->gHSb, 1.0, F->gHb, s);
        launch_fill(F->gHSb, n, 0.0, s);
      }
    }
  }
  F->last_n0 = -1;   // the kept segment's buffers were reused by the recomputation
  if ((rc = normal_order_vjp(B, F->gHb, Tb, s))) return rc;
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

// The vectorize order (basic_IMSRG.f90:3929-3981): E0, fhh, fpp, fph (each
// column-major, whole), then per channel gam(1..6): rectangular blocks whole
// (column-major), square blocks column by column from the diagonal down.
// neq as read_omega_checkpoint counts it (basic_IMSRG.f90:3338-3348).
void vec_plan(const Basis& b, std::vector<VecTask>* out, i64* neq) {
  std::vector<VecTask> v;
  i64 k = 0;
  auto run = [&](i64 off, i64 len, int mirror) {
    // split long whole-block runs so no workgroup streams more than 64 Ki
    // elements (a square-block column, len < 2^16, stays one run: its
    // mirrored partners are addressed from the run's start)
    if (mirror) {
      v.push_back(VecTask{off, k, (int)len, mirror});
      k += len;
      return;
    }
    for (i64 x = 0; x < len; x += 65536) {
      const int l = (int)std::min<i64>(65536, len - x);
      v.push_back(VecTask{off + x, k, l, mirror});
      k += l;
    }
  };
  run(0, 1, 0);
  run(b.off_fhh, (i64)b.A * b.A, 0);
  run(b.off_fpp, (i64)b.P * b.P, 0);
  run(b.off_fph, (i64)b.P * b.A, 0);
  for (int q = 0; q < b.nblocks; ++q) {
    const int np = b.dims[q * 3], nb = b.dims[q * 3 + 1], nh = b.dims[q * 3 + 2];
    const int rows[6] = {np, np, np, nb, nh, nb}, cols[6] = {np, nb, nh, nb, nh, nh};
    for (int g = 0; g < 6; ++g) {
      const i64 off = b.goff[q * 6 + g];
      const int nr = rows[g], nc = cols[g];
      if (g == 0 || g == 3 || g == 4) {
        for (int i = 0; i < nc; ++i) run(off + i + (i64)i * nr, nr - i, nr);
      } else {
        run(off, (i64)nr * nc, 0);
      }
    }
  }
  if (out) *out = std::move(v);
  if (neq) *neq = k;
}

// Name of the Omega checkpoint at flow parameter s, as write_omega_checkpoint
// and read_omega_checkpoint form it (basic_IMSRG.f90:3239-3249,3340-3347):
// s written as Fortran F6.3 (asterisks when it does not fit), left
// adjusted, with a '0' appended when s < 10, after "<prefix>_s".
std::string ckpt_name(const std::string& prefix, double s) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%6.3f", s);
  std::string f(buf);
  if (f.size() > 6) f = "******";
  if (s < 10.0) {
    const size_t a = f.find_first_not_of(' ');
    return prefix + "_s" + (a == std::string::npos ? std::string() : f.substr(a)) + "0.bin";
  }
  return prefix + "_s" + f + ".bin";
}

bool file_exists(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  std::fclose(f);
  return true;
}

// magnus_decouple, IMSRG_MAGNUS.f90:15-128, with the flow's generator.
// prefix != null switches the reference's checkpointing on: on entry the
// newest Omega file s = 95.5, 95.0, ... (> 1.0) is looked for
// (read_omega_checkpoint, basic_IMSRG.f90:3304-3407); when one exists the
// flow resumes from it (HS = BCH(Omega, H), the generator rebuilt, s = sx,
// steps = nint(sx/ds), no mbpt2 test before the first step), and every chk
// steps Omega is written before the step (not at the first such step after
// a restart) and the file 12 (eMax 14: 4) units of s older removed
// (write_omega_checkpoint, 3215-3301).
int decouple(imsrg_flow_s* F, double* HS, double* G, double ds, double bch_conv, double mag_conv, int crit_mode,
             double crit_tol, int maxsteps, const char* prefix, int chk, double* log, int* steps, double* restart_s,
             hipStream_t s) {
  if (!F) return fail(IMSRG_EINVAL, "null flow");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  if (!HS || !G) return fail(IMSRG_EINVAL, "null argument");
  imsrg_basis_s* B = F->B;
  const i64 n = B->b.oplen;
  int emax = 0;
  for (int a = 0; a < B->b.N; ++a) emax = std::max(emax, 2 * B->b.nn[a] + B->b.ll[a]);
  if (chk <= 0) chk = emax == 14 ? 12 : (emax > 14 ? 4 : 24);
  const double drop = emax == 14 ? 4.0 : 12.0;
  int rc;
  if ((rc = fbuf(F, &F->decDG, n))) return rc;
  double* DG = F->decDG;
  launch_fill(DG, n, 0.0, s);
  HIPCHK(hipMemcpyAsync(F->Hbare, HS, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  double sx = -1.0;
  if (prefix) {
    for (double t = 95.5; t > 1.0; t -= 0.5)
      if (file_exists(ckpt_name(prefix, t))) { sx = t; break; }
  }
  if (restart_s) *restart_s = sx;
  double crit, Eold, s_flow = 0.0;
  int st = 0;
  bool first = false;
  if (sx < 0.0) {
    if ((rc = generator(F, F->X, HS, DG, s))) return rc;
    if ((rc = mbpt2_dev(F->X, HS, B->dval + 2, s))) return rc;
    HIPCHK(hipMemcpyAsync(B->dval + 3, HS, sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(B->hval, B->dval + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    crit = std::fabs(B->hval[0]);
    Eold = B->hval[1];
  } else {
    if ((rc = imsrg_read_op_bin(B, ckpt_name(prefix, sx).c_str(), -1, G, s))) return rc;
    if ((rc = bch(F, HS, G, -1, F->Hbare, 1, bch_conv, 40, nullptr, nullptr, s))) return rc;
    if ((rc = generator(F, F->X, HS, DG, s))) return rc;
    HIPCHK(hipMemcpyAsync(B->hval, HS, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    crit = 10.0;
    Eold = B->hval[0];
    s_flow = sx;
    st = (int)std::lround(sx / ds);
    first = true;
  }
  const int st0 = st;
  while (true) {
    if (crit_mode == 0 && !(crit > crit_tol)) break;
    if (st - st0 >= maxsteps) {
      if (steps) *steps = st;
      return fail(IMSRG_ESTEPS, "magnus_decouple: step limit reached");
    }
    ++st;
    if (prefix && st % chk == 0) {
      if (first) {
        first = false;
      } else {
        if ((rc = imsrg_write_op_bin(B, G, ckpt_name(prefix, s_flow).c_str(), s))) return rc;
        std::remove(ckpt_name(prefix, s_flow - drop).c_str());
      }
    }
    double out[4];
    if ((rc = magnus_step(F, HS, G, DG, F->Hbare, ds, bch_conv, mag_conv, out, s))) return rc;
    s_flow += ds;
    crit = std::fabs(out[1]);
    if (log) {
      const int r = st - st0 - 1;
      log[r * 4 + 0] = s_flow;
      log[r * 4 + 1] = out[0];
      log[r * 4 + 2] = out[0] + out[1];
      log[r * 4 + 3] = out[2] + (out[3] - 1);
    }
    if (crit_mode == 1) {
      if (std::fabs(out[0] - Eold) < crit_tol) break;
      Eold = out[0];
    }
  }
  if (steps) *steps = st;
  return 0;
}

}  // namespace

// Fixed device workspace of imsrg_commutator_222_ph_noalloc: 64 MiB per GPU,
// held by every GPU the library runs on until the process ends.
constexpr i64 kNoAllocWccDoubles = (i64)(64 << 20) / (i64)sizeof(double);
static __device__ __align__(256) double scratch[kNoAllocWccDoubles];

namespace {

// device address of `scratch` per device, resolved once per device with
// that device current
int noalloc_scratch(double** out) {
  static std::mutex mu;
  static std::map<int, double*> addr;
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(mu);
  auto it = addr.find(dev);
  if (it == addr.end()) {
    double* p = nullptr;
    HIPCHK(hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(IMSRG_EHIP, "hipGetSymbolAddress returned no address for the no-allocation workspace");
    it = addr.emplace(dev, p).first;
  }
  *out = it->second;
  return 0;
}

int ensure_vec(imsrg_basis_s* B) {
  if (B->vec) return 0;
  int rc;
  vec_plan(B->b, &B->vec_h, &B->neq);
  if ((rc = upload(B, B->vec_h, &B->vec))) return rc;
  void* p = nullptr;
  if ((rc = dalloc(B, (size_t)B->neq * sizeof(double), &p))) return rc;
  B->vecbuf = (double*)p;
  return 0;
}

}  // namespace

// ================================================================== C-ABI
extern "C" {

const char* imsrg_last_error(void) { return g_err.c_str(); }

int imsrg_basis_create(int norb, const int* n, const int* l, const int* jj2, const int* tz2, const int* hole,
                       imsrg_basis* out) {
  if (!out || !n || !l || !jj2 || !tz2 || !hole) return fail(IMSRG_EINVAL, "null argument");
  imsrg_basis_s* B = new imsrg_basis_s();
  std::string e = build_basis(B->b, norb, n, l, jj2, tz2, hole);
  if (!e.empty()) {
    delete B;
    return fail(IMSRG_EINVAL, e);
  }
  *out = B;
  return 0;
}

int imsrg_basis_destroy(imsrg_basis B) {
  delete B;
  return 0;
}

int64_t imsrg_basis_size(imsrg_basis B, int what) {
  if (!B) return -1;
  const Basis& b = B->b;
  switch (what) {
    case IMSRG_SIZE_NORB: return b.N;
    case IMSRG_SIZE_NHOLE: return b.A;
    case IMSRG_SIZE_NPART: return b.P;
    case IMSRG_SIZE_JTOTAL_MAX: return b.JTM;
    case IMSRG_SIZE_NBLOCKS: return b.nblocks;
    case IMSRG_SIZE_OP: return b.oplen;
    case IMSRG_SIZE_NCC: return b.ncc;
    case IMSRG_SIZE_CC: return b.cclen;
    case IMSRG_SIZE_WCC: return b.wclen;
    case IMSRG_SIZE_NEQ: {
      i64 n = 0;
      vec_plan(b, nullptr, &n);
      return n;
    }
    case IMSRG_SIZE_SPMAT: {
      i64 n = 0;
      for (size_t q = 0; q + 1 < b.spstart.size(); ++q)
        n += (i64)(b.spstart[q + 1] - b.spstart[q]) * (b.spstart[q + 1] - b.spstart[q]);
      return n;
    }
  }
  return -1;
}

int imsrg_basis_channels(imsrg_basis B, int* dims, int64_t* goff, int* jtp) {
  if (!B) return fail(IMSRG_EINVAL, "null basis");
  const Basis& b = B->b;
  for (int q = 0; q < b.nblocks; ++q) {
    for (int c = 0; c < 3; ++c) if (dims) dims[q * 3 + c] = b.dims[q * 3 + c];
    for (int g = 0; g < 6; ++g) if (goff) goff[q * 6 + g] = b.goff[q * 6 + g];
    if (jtp) { jtp[q * 3] = b.chJ[q]; jtp[q * 3 + 1] = b.chT[q]; jtp[q * 3 + 2] = b.chP[q]; }
  }
  return 0;
}

int imsrg_basis_pairs(imsrg_basis B, int q, int cls, int* out) {
  if (!B || q < 0 || q >= B->b.nblocks || cls < 0 || cls > 2) return fail(IMSRG_EINVAL, "bad channel/class");
  const std::vector<int>& v = B->b.pl[q * 3 + cls];
  if (out) std::copy(v.begin(), v.end(), out);
  return (int)v.size() / 2;
}

int imsrg_basis_cc(imsrg_basis B, int* rlen, int* ncol, int64_t* ccoff, int64_t* wcoff, int* jtp) {
  if (!B) return fail(IMSRG_EINVAL, "null basis");
  const Basis& b = B->b;
  for (int q = 0; q < b.ncc; ++q) {
    if (rlen) rlen[q] = b.rlen[q];
    if (ncol) ncol[q] = b.ncol[q];
    if (ccoff) ccoff[q] = b.ccoff[q];
    if (wcoff) wcoff[q] = b.wcoff[q];
    if (jtp) { jtp[q * 3] = b.ccJ[q]; jtp[q * 3 + 1] = b.ccT[q]; jtp[q * 3 + 2] = b.ccP[q]; }
  }
  return 0;
}

double imsrg_basis_sixj(imsrg_basis B, int j1, int j2, int j3, int j4, int j5, int j6) {
  (void)B;
  return sixj_value(j1, j2, j3, j4, j5, j6);
}

int imsrg_alloc(int64_t count, double** out) {
  if (!out || count < 0) return fail(IMSRG_EINVAL, "bad allocation request");
  void* p = nullptr;
  HIPCHK(hipMalloc(&p, std::max<int64_t>(count, 1) * sizeof(double)));
  HIPCHK(hipMemset(p, 0, std::max<int64_t>(count, 1) * sizeof(double)));
  *out = (double*)p;
  return 0;
}

int imsrg_free(double* p) {
  if (p) HIPCHK(hipFree(p));
  return 0;
}

namespace {
// offset and length (doubles) of one sq_op component in the flat layout
int component(const Basis& b, int q, int g, i64* off, i64* len) {
  const i64 A = b.A, P = b.P;
  if (q == -1) {
    switch (g) {
      case 0: *off = 0; *len = 1; return 0;
      case 1: *off = 1; *len = A * A; return 0;
      case 2: *off = 1 + A * A; *len = P * A; return 0;
      case 3: *off = 1 + A * A + P * A; *len = P * P; return 0;
      default: return fail(IMSRG_EINVAL, "one-body component g must be 0..3");
    }
  }
  if (q < 0 || q >= b.nblocks || g < 0 || g > 5) return fail(IMSRG_EINVAL, "bad channel / block");
  *off = b.goff[q * 6 + g];
  *len = (i64)b.rows(q, g) * b.cols(q, g);
  return 0;
}
}  // namespace

int imsrg_op_put(imsrg_basis B, double* op, int q, int g, const double* host) {
  if (!B || !op || !host) return fail(IMSRG_EINVAL, "null argument");
  i64 off, len;
  if (int rc = component(B->b, q, g, &off, &len)) return rc;
  if (len > 0) HIPCHK(hipMemcpy(op + off, host, len * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int imsrg_op_get(imsrg_basis B, const double* op, int q, int g, double* host) {
  if (!B || !op || !host) return fail(IMSRG_EINVAL, "null argument");
  i64 off, len;
  if (int rc = component(B->b, q, g, &off, &len)) return rc;
  if (len > 0) HIPCHK(hipMemcpy(host, op + off, len * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int imsrg_set_path(imsrg_basis B, int path) {
  if (!B || (path != IMSRG_PATH_GROUPED && path != IMSRG_PATH_PLAIN)) return fail(IMSRG_EINVAL, "bad path");
  B->path = path;
  return 0;
}

int imsrg_set_term_streams(imsrg_basis B, int on) {
  if (!B) return fail(IMSRG_EINVAL, "null basis");
  B->term_streams = (on != 0);
  ++B->plan_ver;   // captured commutator graphs hold the other schedule
  return 0;
}

#define PREP()                          \
  if (!B) return fail(IMSRG_EINVAL, "null basis"); \
  {                                     \
    int rc_ = ensure_dev(B);            \
    if (rc_) return rc_;                \
  }                                     \
  hipStream_t s = S(stream);

int imsrg_commutator_110(imsrg_basis B, const double* L, int Lh, const double* R, int Rh, double* value,
                         imsrg_stream stream) {
  if (int a_ = check_args({L, R, value}, {Lh, Rh})) return a_;
  PREP();
  launch_e0_partials(B->D, L, Lh, R, Rh, B->part, s);
  launch_e0_final(B->part, B->b.nblocks, 0, 1, B->dval, 0, s);
  HIPCHK(hipGetLastError());
  return read_scalar(B, B->dval, value, s);
}

int imsrg_commutator_220(imsrg_basis B, const double* L, int Lh, const double* R, int Rh, double* value,
                         imsrg_stream stream) {
  if (int a_ = check_args({L, R, value}, {Lh, Rh})) return a_;
  PREP();
  launch_e0_partials(B->D, L, Lh, R, Rh, B->part, s);
  launch_e0_final(B->part, B->b.nblocks, (Lh * Rh == -1) ? 1 : 0, 0, B->dval, 0, s);
  HIPCHK(hipGetLastError());
  return read_scalar(B, B->dval, value, s);
}

int imsrg_commutator_e0(imsrg_basis B, const double* L, int Lh, const double* R, int Rh, double* RES,
                        imsrg_stream stream) {
  if (int a_ = check_args({L, R, RES}, {Lh, Rh})) return a_;
  PREP();
  t_e0(basis_exec(B), L, Lh, R, Rh, RES, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_commutator_111(imsrg_basis B, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh,
                         imsrg_stream stream) {
  if (int a_ = check_args({L, R, RES}, {Lh, Rh})) return a_;
  PREP();
  (void)RESh;
  t_111(basis_exec(B), L, Lh, R, Rh, RES, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_commutator_121(imsrg_basis B, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh,
                         imsrg_stream stream) {
  if (int a_ = check_args({L, R, RES}, {Lh, Rh, RESh})) return a_;
  PREP();
  t_121(basis_exec(B), L, Lh, R, Rh, RES, RESh, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_commutator_122(imsrg_basis B, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh,
                         imsrg_stream stream) {
  if (int a_ = check_args({L, R, RES}, {Lh, Rh, RESh})) return a_;
  PREP();
  t_122(basis_exec(B), L, Lh, R, Rh, RES, RESh, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_commutator_222_pp_hh(imsrg_basis B, const double* L, int Lh, const double* R, int Rh, double* RES,
                               int RESh, double* w1, double* w2, imsrg_stream stream) {
  if (int a_ = check_args({L, R, RES, w1, w2}, {Lh, Rh, RESh})) return a_;
  PREP();
  t_pphh(basis_exec(B), L, Lh, R, Rh, RES, RESh, w1, w2, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_commutator_221(imsrg_basis B, const double* L, int Lh, const double* R, int Rh, double* RES, int RESh,
                         const double* w1, const double* w2, imsrg_stream stream) {
  if (int a_ = check_args({L, R, RES, w1, w2}, {Lh, Rh, RESh})) return a_;
  PREP();
  t_221(basis_exec(B), L, Lh, R, Rh, RES, RESh, w1, w2, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_calculate_cross_coupled(imsrg_basis B, const double* op, int herm, double* CC, imsrg_stream stream) {
  if (int a_ = check_args({op, CC}, {herm})) return a_;
  PREP();
  t_pandya(basis_exec(B), op, herm, CC, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_commutator_222_ph(imsrg_basis B, const double* LCC, int Lh, const double* RCC, int Rh, double* RES,
                            int RESh, double* WCC, imsrg_stream stream) {
  if (int a_ = check_args({LCC, RCC, RES, WCC}, {Lh, Rh, RESh})) return a_;
  PREP();
  t_ph(basis_exec(B), LCC, Lh, RCC, Rh, RES, RESh, WCC, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_add_sq_op(imsrg_basis B, const double* A, double ax, const double* Bv, double bx, double* C,
                    imsrg_stream stream) {
  if (int a_ = check_args({A, Bv, C}, {})) return a_;
  PREP();
  launch_axpy(B->b.oplen, A, ax, Bv, bx, C, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_mat_frob_norm(imsrg_basis B, const double* op, double* value, imsrg_stream stream) {
  if (int a_ = check_args({op, value}, {})) return a_;
  PREP();
  t_norm(basis_exec(B), op, B->dval, s);
  HIPCHK(hipGetLastError());
  return read_scalar(B, B->dval, value, s);
}

int imsrg_build_gs_white(imsrg_basis B, const double* H, double* ETA, imsrg_stream stream) {
  if (int a_ = check_args({H, ETA}, {})) return a_;
  PREP();
  return white(basis_exec(B), H, ETA, s);
}

int imsrg_build_gs_atan(imsrg_basis B, const double* H, double* ETA, imsrg_stream stream) {
  if (!H || !ETA) return fail(IMSRG_EINVAL, "null argument");
  PREP();
  return gen(basis_exec(B), H, H, ETA, 1, s);
}

int imsrg_build_gs_atanxx(imsrg_basis B, const double* Hold, const double* H, double* ETA, imsrg_stream stream) {
  if (!Hold || !H || !ETA) return fail(IMSRG_EINVAL, "null argument");
  PREP();
  return gen(basis_exec(B), Hold, H, ETA, 1, s);
}

int imsrg_vectorize(imsrg_basis B, const double* op, double* vec, imsrg_stream stream) {
  if (!op || !vec) return fail(IMSRG_EINVAL, "null argument");
  PREP();
  int rc;
  if ((rc = ensure_vec(B))) return rc;
  launch_vectorize(B->vec, (int)B->vec_h.size(), (double*)op, vec, 0, 1, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_repackage(imsrg_basis B, const double* vec, int herm, double* op, imsrg_stream stream) {
  if (!op || !vec) return fail(IMSRG_EINVAL, "null argument");
  if (herm != 1 && herm != -1) return fail(IMSRG_EINVAL, "herm must be +1 or -1");
  PREP();
  int rc;
  if ((rc = ensure_vec(B))) return rc;
  launch_vectorize(B->vec, (int)B->vec_h.size(), op, (double*)vec, 1, herm, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_write_op_bin(imsrg_basis B, const double* op, const char* path, imsrg_stream stream) {
  if (!op || !path) return fail(IMSRG_EINVAL, "null argument");
  PREP();
  int rc;
  if ((rc = ensure_vec(B))) return rc;
  std::vector<double> h((size_t)B->neq);
  launch_vectorize(B->vec, (int)B->vec_h.size(), (double*)op, B->vecbuf, 0, 1, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h.data(), B->vecbuf, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(IMSRG_EINVAL, std::string("cannot open ") + path + " for writing");
  const size_t w = std::fwrite(h.data(), sizeof(double), h.size(), f);
  const int cl = std::fclose(f);
  if (w != h.size() || cl != 0) return fail(IMSRG_EINVAL, std::string("short write to ") + path);
  return 0;
}

int imsrg_read_op_bin(imsrg_basis B, const char* path, int herm, double* op, imsrg_stream stream) {
  if (!op || !path) return fail(IMSRG_EINVAL, "null argument");
  if (herm != 1 && herm != -1) return fail(IMSRG_EINVAL, "herm must be +1 or -1");
  PREP();
  int rc;
  if ((rc = ensure_vec(B))) return rc;
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(IMSRG_EINVAL, std::string("cannot open ") + path);
  std::fseek(f, 0, SEEK_END);
  const long bytes = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (bytes != (long)(B->neq * (i64)sizeof(double))) {
    std::fclose(f);
    return fail(IMSRG_EINVAL, std::string(path) + ": " + std::to_string(bytes) + " bytes, the basis needs " +
                                  std::to_string(B->neq * 8));
  }
  std::vector<double> h((size_t)B->neq);
  const size_t r = std::fread(h.data(), sizeof(double), h.size(), f);
  std::fclose(f);
  if (r != h.size()) return fail(IMSRG_EINVAL, std::string("short read from ") + path);
  HIPCHK(hipMemcpyAsync(B->vecbuf, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s));
  launch_vectorize(B->vec, (int)B->vec_h.size(), op, B->vecbuf, 1, herm, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int imsrg_h0_harm_osc(imsrg_basis B, double hw, int htype, double lawson_beta, double com_hw, int mass, double* H,
                      imsrg_stream stream) {
  if (!H) return fail(IMSRG_EINVAL, "null argument");
  if (!B) return fail(IMSRG_EINVAL, "null basis");
  const Basis& b = B->b;
  HamParams p;
  p.hw = hw; p.htype = htype; p.lawson_beta = lawson_beta; p.com_hw = com_hw; p.mass = mass;
  // host copy of the operator up to the end of fpp
  std::vector<double> h((size_t)(b.off_fpp + (i64)b.P * b.P), 0.0);
  std::string err = h0_harm_osc(b, p, h.data());
  if (!err.empty()) return fail(IMSRG_EINVAL, err);
  PREP();
  const size_t n1 = h.size() - (size_t)b.off_fhh;
  HIPCHK(hipMemcpyAsync(H + b.off_fhh, h.data() + b.off_fhh, n1 * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int imsrg_read_me2j(imsrg_basis B, const char* path, int nfile, const int* n, const int* l, const int* jj,
                    const int* tz, double hw, int htype, double lawson_beta, double com_hw, int mass, double* H,
                    imsrg_stream stream) {
  if (!path || !n || !l || !jj || !tz || !H) return fail(IMSRG_EINVAL, "null argument");
  if (!B) return fail(IMSRG_EINVAL, "null basis");
  if (htype < 1 || htype > 5) return fail(IMSRG_EINVAL, "htype must be 1..5");
  if (mass < 1) return fail(IMSRG_EINVAL, "mass must be positive");
  if (!(hw > 0.0)) return fail(IMSRG_EINVAL, "hw must be positive");
  const Basis& b = B->b;
  HamParams p;
  p.hw = hw; p.htype = htype; p.lawson_beta = lawson_beta; p.com_hw = com_hw; p.mass = mass;
  FileOrbits f;
  std::string err = map_file_orbits(b, nfile, n, l, jj, tz, f);
  if (!err.empty()) return fail(IMSRG_EINVAL, err);
  std::vector<double> me;
  err = me2j_read_values(path, me2j_count(f), me);
  if (!err.empty()) return fail(IMSRG_EINVAL, err);
  PREP();
  // the two-body part round trips through the host: elements the file
  // does not hold keep their values
  std::vector<double> h((size_t)b.oplen, 0.0);
  const i64 t0 = b.goff[0];
  const size_t n2 = (size_t)(b.oplen - t0);
  HIPCHK(hipMemcpyAsync(h.data() + t0, H + t0, n2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  err = me2j_to_op(b, f, me, p, h.data());
  if (!err.empty()) return fail(IMSRG_EINVAL, err);
  HIPCHK(hipMemcpyAsync(H + t0, h.data() + t0, n2 * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

// the two GEMM families and buffers of the Hartree-Fock two-body transform
int ensure_hf(imsrg_basis_s* B) {
  if (B->hf_buf) return 0;
  const Basis& b = B->b;
  std::vector<i64> off(b.nblocks + 1, 0);
  for (int q = 0; q < b.nblocks; ++q) {
    const i64 nt = (i64)b.dims[q * 3] + b.dims[q * 3 + 1] + b.dims[q * 3 + 2];
    off[q + 1] = off[q] + nt * nt;
  }
  B->hf_len = std::max<i64>(off[b.nblocks], 1);
  int rc;
  if ((rc = dalloc(B, (size_t)(3 * B->hf_len) * sizeof(double), (void**)&B->hf_buf))) return rc;
  if ((rc = upload(B, off, &B->hf_off))) return rc;
  for (int pass = 0; pass < 2; ++pass) {
    GemmFamily& f = pass == 0 ? B->hf1 : B->hf2;
    f.probs_h.clear(); f.nks.clear(); f.group_end.clear();
    for (int q = 0; q < b.nblocks; ++q) {
      const int nt = b.dims[q * 3] + b.dims[q * 3 + 1] + b.dims[q * 3 + 2];
      if (nt == 0) continue;
      GemmProb p{};
      p.a1 = p.b1 = p.c = off[q];
      p.m = p.n = p.k = nt;
      p.lda = p.ldb = p.ldc = nt;
      p.nprod = 1;
      // slots: 0 V, 1 T, 2 M.  pass 0: T = V M; pass 1: V = M^T T
      p.ta1 = pass == 0 ? 0 : 1;
      p.sa1 = pass == 0 ? 0 : 2;
      p.sb1 = pass == 0 ? 2 : 1;
      p.sc = pass == 0 ? 1 : 0;
      p.mode = 0;
      p.alpha = 1.0;
      p.beta = 0.0;
      f.probs_h.push_back(p);
      f.nks.push_back(1);
      f.group_end.push_back((int)f.probs_h.size());
    }
    if ((rc = upload(B, f.probs_h, &f.probs))) return rc;
    if ((rc = family_tasks(B, f, 0))) return rc;
  }
  return 0;
}

// E0 and the one-body part of the host copy h uploaded to H, then H's
// two-body part transformed by the per-channel matrices M on the device
int hf_transform_dev(imsrg_basis_s* B, const double* h, const std::vector<double>& M, double* H, hipStream_t s) {
  const Basis& b = B->b;
  int rc;
  if ((rc = ensure_hf(B))) return rc;
  double *V = B->hf_buf, *T = B->hf_buf + B->hf_len, *Md = B->hf_buf + 2 * B->hf_len;
  const size_t n1 = (size_t)(b.off_fpp + (i64)b.P * b.P);
  HIPCHK(hipMemcpyAsync(H, h, n1 * sizeof(double), hipMemcpyHostToDevice, s));
  if (!M.empty()) HIPCHK(hipMemcpyAsync(Md, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice, s));
  launch_hf_pack(B->D, B->seg_full.dev, B->seg_full.n, H, V, B->hf_off, s);
  Slots sl{};
  sl.p[0] = V; sl.p[1] = T; sl.p[2] = Md;
  run_gemm(B, B->hf1, sl, s);
  run_gemm(B, B->hf2, sl, s);
  launch_hf_unpack(B->D, B->seg_full.dev, B->seg_full.n, V, B->hf_off, H, s);
  HIPCHK(hipGetLastError());
  // the host copy must outlive the upload
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int imsrg_sqop_to_hf(imsrg_basis B, const double* coefs, double* Op, int herm, imsrg_stream stream) {
  if (!coefs || !Op) return fail(IMSRG_EINVAL, "null argument");
  if (herm != 1) return fail(IMSRG_EINVAL, "sqop_to_HF transforms hermitian scalar operators (herm = 1)");
  PREP();
  const Basis& b = B->b;
  std::vector<double> h((size_t)b.oplen);
  HIPCHK(hipMemcpyAsync(h.data(), Op, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  std::vector<i64> moff;
  std::vector<double> M;
  std::string err = sqop_to_hf_host(b, coefs, h.data(), moff, M);
  if (!err.empty()) return fail(IMSRG_EINVAL, err);
  int rc;
  if ((rc = hf_transform_dev(B, h.data(), M, Op, s))) return rc;
  launch_normal_order(B->D, Op, B->dval, s);   // normal_order (HF_mod.f90:204)
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int imsrg_hartree_fock(imsrg_basis B, double* H, int maxit, double* ehf, int* iterations, double* coefs,
                       imsrg_stream stream) {
  if (!H) return fail(IMSRG_EINVAL, "null argument");
  if (maxit < 1) return fail(IMSRG_EINVAL, "maxit must be positive");
  PREP();
  const Basis& b = B->b;
  std::vector<double> h((size_t)b.oplen);
  HIPCHK(hipMemcpyAsync(h.data(), H, h.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  HFResult r;
  std::vector<double> M;
  std::string err = hartree_fock_host(b, h.data(), maxit, r, M);
  if (!err.empty()) return fail(IMSRG_ESTEPS, err);
  int rc;
  if ((rc = hf_transform_dev(B, h.data(), M, H, s))) return rc;
  HIPCHK(hipStreamSynchronize(s));
  if (ehf) *ehf = r.ehf;
  if (iterations) *iterations = r.iterations;
  if (coefs) std::copy(r.D.begin(), r.D.end(), coefs);
  return 0;
}

int imsrg_mbpt2(imsrg_basis B, const double* H, double* value, imsrg_stream stream) {
  if (int a_ = check_args({H, value}, {})) return a_;
  PREP();
  int rc = mbpt2_dev(basis_exec(B), H, B->dval, s);
  if (rc) return rc;
  return read_scalar(B, B->dval, value, s);
}

int imsrg_normal_order(imsrg_basis B, double* H, imsrg_stream stream) {
  if (int a_ = check_args({H}, {})) return a_;
  PREP();
  launch_normal_order(B->D, H, B->dval, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_flow_create(imsrg_basis B, imsrg_flow* out) {
  if (!B || !out) return fail(IMSRG_EINVAL, "null argument");
  int rc = ensure_dev(B);
  if (rc) return rc;
  imsrg_flow_s* F = new imsrg_flow_s();
  F->B = B;
  F->X = basis_exec(B);
  const Basis& b = B->b;
  if ((rc = falloc(F, b.oplen, &F->buf[0])) || (rc = falloc(F, b.oplen, &F->buf[1])) ||
      (rc = falloc(F, b.oplen, &F->w1)) || (rc = falloc(F, b.oplen, &F->w2)) ||
      (rc = falloc(F, b.cclen, &F->GCC)) || (rc = falloc(F, b.cclen, &F->ADCC)) ||
      (rc = falloc(F, b.wclen, &F->WCC)) || (rc = falloc(F, b.oplen, &F->Hbare))) {
    delete F;
    return rc;
  }
  *out = F;
  return 0;
}

int imsrg_flow_destroy(imsrg_flow F) {
  delete F;
  return 0;
}

int imsrg_flow_set_graphs(imsrg_flow F, int on) {
  if (!F) return fail(IMSRG_EINVAL, "null flow");
  F->graphs = (on != 0);
  return 0;
}

int imsrg_flow_set_generator(imsrg_flow F, int generator) {
  if (!F) return fail(IMSRG_EINVAL, "null flow");
  if (generator != IMSRG_GEN_WHITE && generator != IMSRG_GEN_ATAN && generator != IMSRG_GEN_W2)
    return fail(IMSRG_EINVAL, "unknown generator");
  F->generator = generator;
  return 0;
}

int imsrg_build_gs_w2(imsrg_flow F, const double* H, double* ETA, imsrg_stream stream) {
  if (!F || !H || !ETA) return fail(IMSRG_EINVAL, "null argument");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  return gen_w2(F, F->X, H, ETA, S(stream));
}

int imsrg_bch_expand(imsrg_flow F, double* HS, const double* G, int Gh, const double* H, int Hh, double conv,
                     int maxit, int* nterms, double* advals, imsrg_stream stream) {
  if (!F) return fail(IMSRG_EINVAL, "null flow");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  return bch(F, HS, G, Gh, H, Hh, conv, maxit, nterms, advals, S(stream));
}

int imsrg_transform_sq_op_bch(imsrg_flow F, double* Op, int Oph, const double* G, double conv, int maxit,
                              int* nterms, imsrg_stream stream) {
  if (!F || !Op || !G) return fail(IMSRG_EINVAL, "null argument");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  hipStream_t s = S(stream);
  const i64 n = F->B->b.oplen;
  int rc;
  if ((rc = fbuf(F, &F->tOp, n))) return rc;
  // Oevolved = BCH(G, Op); copy_sq_op(Oevolved, Op)
  if ((rc = bch(F, F->tOp, G, -1, Op, Oph, conv, maxit, nterms, nullptr, s))) return rc;
  HIPCHK(hipMemcpyAsync(Op, F->tOp, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  return 0;
}

int imsrg_dhds(imsrg_flow F, const double* HS, double* DH, imsrg_stream stream) {
  if (!F || !HS || !DH) return fail(IMSRG_EINVAL, "null argument");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  if (HS == DH) return fail(IMSRG_EINVAL, "DH must not alias HS");
  hipStream_t s = S(stream);
  int rc;
  if ((rc = fbuf(F, &F->tOp, F->B->b.oplen))) return rc;
  double* eta = F->tOp;
  launch_fill(eta, F->B->b.oplen, 0.0, s);   // duplicate_sq_op(HS, ETA): a zero operator
  if ((rc = generator(F, F->X, HS, eta, s))) return rc;
  t_pandya(F->X, eta, -1, F->GCC, s);
  t_full(F->X, eta, -1, HS, 1, DH, 1, true, F->w1, F->w2, F->GCC, F->ADCC, F->WCC, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_magnus_expand(imsrg_flow F, double* DG, int DGh, const double* G, int Gh, double conv, int* ncomm,
                        imsrg_stream stream) {
  if (!F) return fail(IMSRG_EINVAL, "null flow");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  return magnus(F, DG, DGh, G, Gh, conv, ncomm, S(stream));
}

int imsrg_euler_step(imsrg_basis B, double* G, const double* DG, double ds, imsrg_stream stream) {
  if (int a_ = check_args({G, DG}, {})) return a_;
  PREP();
  launch_axpy(B->b.oplen, G, 1.0, DG, ds, G, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_magnus_step(imsrg_flow F, double* HS, double* G, double* DG, const double* H, double ds,
                      double bch_conv, double mag_conv, double* out, imsrg_stream stream) {
  if (!F) return fail(IMSRG_EINVAL, "null flow");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  return magnus_step(F, HS, G, DG, H, ds, bch_conv, mag_conv, out, S(stream));
}

int imsrg_magnus_decouple(imsrg_flow F, double* HS, double* G, double ds, double bch_conv, double mag_conv,
                          int crit_mode, double crit_tol, int maxsteps, double* log, int* steps,
                          imsrg_stream stream) {
  return decouple(F, HS, G, ds, bch_conv, mag_conv, crit_mode, crit_tol, maxsteps, nullptr, 0, log, steps, nullptr,
                  S(stream));
}

int imsrg_magnus_decouple_chk(imsrg_flow F, double* HS, double* G, double ds, double bch_conv, double mag_conv,
                              int crit_mode, double crit_tol, int maxsteps, const char* prefix, int chk,
                              double* log, int* steps, double* restart_s, imsrg_stream stream) {
  if (!prefix) return fail(IMSRG_EINVAL, "null checkpoint prefix");
  return decouple(F, HS, G, ds, bch_conv, mag_conv, crit_mode, crit_tol, maxsteps, prefix, chk, log, steps,
                  restart_s, S(stream));
}

int imsrg_omega_checkpoint_name(const char* prefix, double s, char* out, int outlen) {
  if (!prefix || !out || outlen < 1) return fail(IMSRG_EINVAL, "null argument");
  const std::string n = ckpt_name(prefix, s);
  if ((int)n.size() + 1 > outlen) return fail(IMSRG_EINVAL, "name buffer too short");
  std::memcpy(out, n.c_str(), n.size() + 1);
  return 0;
}

// ------------------------------------------------ many-instance batches
int imsrg_batch_create(imsrg_basis B, int ninst, imsrg_flow* out) {
  if (!B || !out) return fail(IMSRG_EINVAL, "null argument");
  if (ninst < 1 || ninst > 65535) return fail(IMSRG_EINVAL, "batch size out of range");
  if (B->path != IMSRG_PATH_GROUPED) return fail(IMSRG_EINVAL, "batched flows run the grouped path");
  int rc = ensure_dev(B);
  if (rc) return rc;
  imsrg_flow_s* F = new imsrg_flow_s();
  F->B = B;
  F->batch = true;
  F->ninst = ninst;
  const Basis& b = B->b;
  const i64 ni = ninst;
  Exec X = basis_exec(B);
  X.D.ninst = ninst;
  void* pm = nullptr;
  if ((rc = falloc(F, ni * b.oplen, &F->buf[0])) || (rc = falloc(F, ni * b.oplen, &F->buf[1])) ||
      (rc = falloc(F, ni * b.oplen, &F->w1)) || (rc = falloc(F, ni * b.oplen, &F->w2)) ||
      (rc = falloc(F, ni * b.cclen, &F->GCC)) || (rc = falloc(F, ni * b.cclen, &F->ADCC)) ||
      (rc = falloc(F, ni * b.wclen, &F->WCC)) || (rc = falloc(F, ni * b.oplen, &F->Hbare)) ||
      (rc = falloc(F, ni * b.oplen, &F->decDG)) || (rc = falloc(F, ni * X.D.spart, &X.part)) ||
      (rc = falloc(F, ni * X.D.spscr, &X.Pscr)) || (rc = falloc(F, ni * X.D.sk2scr, &X.K2scr)) ||
      (rc = falloc(F, ni * b.N * b.N, &X.M)) || (rc = falloc(F, ni * kDStride, &X.dval))) {
    delete F;
    return rc;
  }
  if (hipMalloc(&pm, 2 * ni * sizeof(int)) != hipSuccess) {
    delete F;
    return fail(IMSRG_EHIP, "hipMalloc of the batch masks failed");
  }
  F->allocs.push_back(pm);
  F->actf = (int*)pm;
  F->acte = F->actf + ni;
  if (hipHostMalloc((void**)&F->hmf, ni * sizeof(int), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&F->hme, ni * sizeof(int), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&F->hv, ni * kDStride * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    delete F;
    return fail(IMSRG_EHIP, "hipHostMalloc of the batch staging failed");
  }
  F->X = X;
  F->X.D.act = F->actf;
  F->Xe = X;
  F->Xe.D.act = F->acte;
  F->hf.assign(ninst, 1);
  if ((rc = set_mask(F, F->actf, F->hmf, F->hf, nullptr))) {
    delete F;
    return rc;
  }
  if (hipStreamSynchronize(nullptr) != hipSuccess) {
    delete F;
    return fail(IMSRG_EHIP, "batch mask upload failed");
  }
  *out = F;
  return 0;
}

int imsrg_batch_size(imsrg_flow F) { return F ? F->ninst : -1; }

#define PREP_BATCH()                                                     \
  if (!F || !F->batch) return fail(IMSRG_EINVAL, "not a batched flow"); \
  hipStream_t s = S(stream);                                             \
  imsrg_basis_s* B = F->B;                                               \
  (void)B;

int imsrg_batch_commutator(imsrg_flow F, const double* L, int Lh, const double* R, int Rh, double* RES,
                           imsrg_stream stream) {
  PREP_BATCH();
  if (!L || !R || !RES) return fail(IMSRG_EINVAL, "null argument");
  int rc;
  F->hf.assign(F->ninst, 1);
  if ((rc = set_mask(F, F->actf, F->hmf, F->hf, s))) return rc;
  t_pandya(F->X, L, Lh, F->GCC, s);
  t_full(F->X, L, Lh, R, Rh, RES, -Lh * Rh, true, F->w1, F->w2, F->GCC, F->ADCC, F->WCC, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_batch_bch_expand(imsrg_flow F, double* HS, const double* G, const double* H, double conv, int maxit,
                           int* nterms, imsrg_stream stream) {
  PREP_BATCH();
  if (!HS || !G || !H) return fail(IMSRG_EINVAL, "null argument");
  int rc;
  F->hf.assign(F->ninst, 1);
  if ((rc = set_mask(F, F->actf, F->hmf, F->hf, s))) return rc;
  return bch_b(F, HS, G, H, conv, maxit, nterms, s);
}

int imsrg_batch_magnus_expand(imsrg_flow F, double* DG, const double* G, double conv, int* ncomm,
                              imsrg_stream stream) {
  PREP_BATCH();
  if (!DG || !G) return fail(IMSRG_EINVAL, "null argument");
  int rc;
  F->hf.assign(F->ninst, 1);
  if ((rc = set_mask(F, F->actf, F->hmf, F->hf, s))) return rc;
  return magnus_b(F, DG, G, conv, ncomm, s);
}

int imsrg_batch_magnus_decouple(imsrg_flow F, double* HS, double* G, double ds, double bch_conv, double mag_conv,
                                double crit_tol, int maxsteps, double* E0, int* steps, int* nsteps_max,
                                imsrg_stream stream) {
  PREP_BATCH();
  if (!HS || !G) return fail(IMSRG_EINVAL, "null argument");
  const int ni = F->ninst;
  const i64 n = B->b.oplen;
  int rc;
  F->hf.assign(ni, 1);
  if ((rc = set_mask(F, F->actf, F->hmf, F->hf, s))) return rc;
  double* DG = F->decDG;
  launch_fill_b(DG, n, 0.0, ni, F->actf, s);
  launch_copy_b(F->Hbare, HS, n, ni, F->actf, s);
  if ((rc = generator(F, F->X, HS, DG, s))) return rc;
  if ((rc = mbpt2_dev(F->X, HS, F->X.dval + 2, s))) return rc;
  if ((rc = read_inst(F, 3, HS, s))) return rc;
  bool any = false;
  for (int k = 0; k < ni; ++k) {
    if (steps) steps[k] = 0;
    if (E0) E0[k] = F->hv[k * kDStride + 3];
    if (!(std::fabs(F->hv[k * kDStride + 2]) > crit_tol)) F->hf[k] = 0;
    any = any || F->hf[k];
  }
  if ((rc = set_mask(F, F->actf, F->hmf, F->hf, s))) return rc;
  std::vector<int> nm(ni), nb(ni);
  int st = 0;
  while (any) {
    if (st >= maxsteps) {
      if (nsteps_max) *nsteps_max = st;
      return fail(IMSRG_ESTEPS, "batched magnus_decouple: step limit reached");
    }
    ++st;
    if ((rc = magnus_step_b(F, HS, G, DG, F->Hbare, ds, bch_conv, mag_conv, nm.data(), nb.data(), s))) return rc;
    bool changed = false;
    any = false;
    for (int k = 0; k < ni; ++k) {
      if (!F->hf[k]) continue;
      if (E0) E0[k] = F->hv[k * kDStride + 3];
      if (steps) steps[k] = st;
      if (!(std::fabs(F->hv[k * kDStride + 2]) > crit_tol)) {
        F->hf[k] = 0;
        changed = true;
      } else {
        any = true;
      }
    }
    if (changed && (rc = set_mask(F, F->actf, F->hmf, F->hf, s))) return rc;
  }
  if (nsteps_max) *nsteps_max = st;
  return 0;
}

// ------------------------------------------------------- reverse mode
int imsrg_commutator_vjp(imsrg_basis B, const double* L, int Lh, const double* R, int Rh, const double* RESb,
                         int RESh, double* Lb, double* Rb, imsrg_stream stream) {
  if (!B || !L || !R || !RESb || !Lb || !Rb) return fail(IMSRG_EINVAL, "null argument");
  int rc = ensure_adj(B);
  if (rc) return rc;
  comm_vjp(B, L, Lh, R, Rh, RESb, RESh, Lb, Rb, true, nullptr, nullptr, S(stream));
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_bch_vjp(imsrg_flow F, const double* G, const double* H, int Hherm, double conv, int maxit,
                  const double* HSb, double* Gb, double* Hb, int* nterms, imsrg_stream stream) {
  if (!F || !G || !H || !HSb || !Gb || !Hb) return fail(IMSRG_EINVAL, "null argument");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  return bch_vjp(F, G, H, Hherm, conv, maxit, HSb, Gb, Hb, nterms, S(stream));
}

int imsrg_magnus_vjp(imsrg_flow F, const double* ETA, const double* G, double conv, const double* DGb, double* ETAb,
                     double* Gb, int* ncomm, imsrg_stream stream) {
  if (!F || !ETA || !G || !DGb || !ETAb || !Gb) return fail(IMSRG_EINVAL, "null argument");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  return magnus_vjp(F, ETA, G, conv, DGb, ETAb, Gb, ncomm, S(stream));
}

int imsrg_generator_vjp(imsrg_flow F, const double* H, const double* ETAb, double* Hb, imsrg_stream stream) {
  if (!F || !H || !ETAb || !Hb) return fail(IMSRG_EINVAL, "null argument");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  return generator_vjp(F, H, ETAb, Hb, S(stream));
}

int imsrg_white_vjp(imsrg_basis B, const double* H, const double* ETAb, double* Hb, imsrg_stream stream) {
  if (!B || !H || !ETAb || !Hb) return fail(IMSRG_EINVAL, "null argument");
  return white_vjp(B, H, ETAb, Hb, S(stream));
}

int imsrg_normal_order_vjp(imsrg_basis B, const double* Hb, double* Tb, imsrg_stream stream) {
  if (!B || !Hb || !Tb) return fail(IMSRG_EINVAL, "null argument");
  int rc = ensure_dev(B);
  if (rc) return rc;
  return normal_order_vjp(B, Hb, Tb, S(stream));
}

int imsrg_flow_forward(imsrg_flow F, const double* T, double ds, double bch_conv, double mag_conv, double crit_tol,
                       int maxsteps, int ckpt, double* E0, int* steps, imsrg_stream stream) {
  if (!F || !T) return fail(IMSRG_EINVAL, "null argument");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  return flow_forward(F, T, ds, bch_conv, mag_conv, crit_tol, maxsteps, ckpt, E0, steps, S(stream));
}

int imsrg_flow_backward(imsrg_flow F, double* Tb, imsrg_stream stream) {
  if (!F || !Tb) return fail(IMSRG_EINVAL, "null argument");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  return flow_backward(F, Tb, S(stream));
}

int imsrg_flow_grad(imsrg_flow F, const double* T, double ds, double bch_conv, double mag_conv, double crit_tol,
                    int maxsteps, int ckpt, double* Tb, double* E0, int* steps, imsrg_stream stream) {
  if (!F || !T || !Tb) return fail(IMSRG_EINVAL, "null argument");
  if (F->batch) return fail(IMSRG_EINVAL, "batched flow: use the imsrg_batch_* calls");
  int rc = flow_forward(F, T, ds, bch_conv, mag_conv, crit_tol, maxsteps, ckpt, E0, steps, S(stream));
  if (rc) return rc;
  return flow_backward(F, Tb, S(stream));
}

int imsrg_set_gemm_tile(imsrg_basis B, int family, int tile) {
  const int shape = tile & 7;
  if (!B || family < 0 || family > 6 || shape > 7 || tile < 0 || tile > (IMSRG_TILE_XCD | 7))
    return fail(IMSRG_EINVAL, "bad gemm family / tile");
  if (family >= 4) {   // the reverse-mode families: every round / side
    B->tile_req[family] = tile;
    if (!B->adj_ready) return 0;   // applied when the reverse-mode plan is built
    int rc = 0;
    auto set = [&](GemmFamily& g) {
      g.xcd_aware = (tile & IMSRG_TILE_XCD) != 0;
      if (!rc) rc = family_tasks(B, g, shape);
    };
    if (family == 4)
      for (GemmFamily& g : B->pb) set(g);
    else if (family == 5)
      set(B->phb);
    else
      set(B->ppb);
    return rc;
  }
  GemmFamily* f[4] = {&B->pp1, &B->k2, &B->k3, &B->ph};
  f[family]->xcd_aware = (tile & IMSRG_TILE_XCD) != 0;
  B->tile_req[family] = tile;
  if (!B->dev_ready) return 0;   // applied when the plan is built
  return family_tasks(B, *f[family], shape);
}

// ------------------------------------------------ no-allocation variant
// imsrg_commutator_222_ph with the WCC workspace in the library's fixed
// device array instead of a caller buffer.  The array's device address is
// looked up once per device (hipGetSymbolAddress, no allocation) on the
// first call made with that device current; host code uses only that
// address, never the array's name.
int imsrg_commutator_222_ph_noalloc(imsrg_basis B, const double* LCC, int Lh, const double* RCC, int Rh,
                                    double* RES, int RESh, imsrg_stream stream) {
  if (int a_ = check_args({LCC, RCC, RES}, {Lh, Rh, RESh})) return a_;
  if (!B) return fail(IMSRG_EINVAL, "null basis");
  if (B->b.wclen > kNoAllocWccDoubles)
    return fail(IMSRG_EINVAL, "imsrg_commutator_222_ph_noalloc: the basis needs " + std::to_string(B->b.wclen) +
                                  " doubles of WCC workspace, more than the fixed " +
                                  std::to_string(kNoAllocWccDoubles) + " (64 MiB)");
  if (!B->dev_ready)
    return fail(IMSRG_EINVAL,
                "imsrg_commutator_222_ph_noalloc: the basis has no device tables yet (any allocating call on the "
                "basis, e.g. imsrg_flow_create or imsrg_commutator_222_ph, uploads them)");
  double* wcc = nullptr;
  if (int rc = noalloc_scratch(&wcc)) return rc;
  hipStream_t s = S(stream);
  t_ph(basis_exec(B), LCC, Lh, RCC, Rh, RES, RESh, wcc, s);
  HIPCHK(hipGetLastError());
  return 0;
}

int imsrg_get_gemm_tile(imsrg_basis B, int family) {
  if (!B || family < 0 || family > 6) return -1;
  return B->tile_req[family];
}

int imsrg_profile_kernel(imsrg_basis B, int which, const double* L, int Lh, const double* R, int Rh, double* RES,
                         int RESh, double* w1, double* w2, double* LCC, double* RCC, double* WCC, int iters,
                         double* ms, imsrg_stream stream) {
  PREP();
  if (iters < 1 || which < 0 || which > IMSRG_K_COPY) return fail(IMSRG_EINVAL, "bad profile request");
  auto body = [&](int k) {
    switch (k) {
      case IMSRG_K_PPPP: {
        Slots s1{};
        s1.p[0] = L; s1.p[1] = R; s1.p[2] = w1;
        run_gemm(B, B->pp1, s1, s, 1.0, -(double)Rh);
        run(B, B->pairs,
            [&](const PairTask* t, int n) {   // as the commutator runs it (grouped: upper tiles)
              launch_pppp_sym(B->D, t, n, L, Lh, R, Rh, RES, w1, w2, B->path == IMSRG_PATH_GROUPED ? 2 : 0, s);
            });
        break;
      }
      case IMSRG_K_PPPP_GEMM: {
        Slots s1{};
        s1.p[0] = L; s1.p[1] = R; s1.p[2] = w1;
        run_gemm(B, B->pp1, s1, s, 1.0, -(double)Rh);
        break;
      }
      case IMSRG_K_PPHH_SMALL: {
        Slots s2{};
        s2.p[0] = L; s2.p[1] = R; s2.p[2] = B->K2scr;
        run_gemm(B, B->k2, s2, s);
        Slots s3{};
        s3.p[0] = L; s3.p[1] = R; s3.p[2] = B->Pscr;
        run_gemm(B, B->k3, s3, s);
        run(B, B->seg_epi, [&](const SegTask* t, int n) {
          launch_pphh_epilogue(B->D, t, n, L, Lh, R, Rh, RES, w1, w2, B->Pscr, B->poff, B->k3ks, B->K2scr, B->k2off,
                               B->k2ks, s);
        });
        break;
      }
      case IMSRG_K_PH_GEMM: {
        Slots sl{};
        sl.p[0] = LCC; sl.p[1] = RCC; sl.p[2] = WCC;
        run_gemm(B, B->ph, sl, s, (double)Lh, -(double)(Rh * Lh));
        break;
      }
      case IMSRG_K_PH_SCATTER:   // as the grouped commutator runs it: upper triangles only
        if (B->path == IMSRG_PATH_GROUPED)
          launch_comm222ph_chunk(B->D, B->phc, (int)B->phc_h.size(), B->fq, B->fql, B->fqi, B->quad, WCC, Lh, Rh,
                                 RES, RESh, s, 0);
        else
          run(B, B->seg_tri,
              [&](const SegTask* t, int n) { launch_comm222ph_scatter(B->D, t, n, WCC, Lh, Rh, RES, RESh, s); });
        break;
      case IMSRG_K_122: t_122(basis_exec(B), L, Lh, R, Rh, RES, RESh, s, B->path != IMSRG_PATH_GROUPED); break;
      case IMSRG_K_PANDYA: t_pandya(basis_exec(B), R, Rh, RCC, s); break;
      case IMSRG_K_COPY:   // first half of WCC into its second half: wclen/2 doubles each way
        launch_copy(WCC + B->b.wclen / 2, WCC, B->b.wclen / 2, s);
        break;
      case IMSRG_K_ONEBODY:
        t_111(basis_exec(B), L, Lh, R, Rh, RES, s);
        t_121(basis_exec(B), L, Lh, R, Rh, RES, RESh, s);
        t_221(basis_exec(B), L, Lh, R, Rh, RES, RESh, w1, w2, s);
        break;
      case IMSRG_K_COMMUTATOR:
        t_pandya(basis_exec(B), L, Lh, LCC, s);
        t_full(basis_exec(B), L, Lh, R, Rh, RES, RESh, true, w1, w2, LCC, RCC, WCC, s);
        break;
    }
  };
  // inputs of every class once (w1/w2, LCC/RCC, WCC), then the timed loop
  t_pandya(basis_exec(B), L, Lh, LCC, s);
  t_full(basis_exec(B), L, Lh, R, Rh, RES, RESh, true, w1, w2, LCC, RCC, WCC, s);
  body(which);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, s));
  for (int it = 0; it < iters; ++it) body(which);
  HIPCHK(hipEventRecord(e1, s));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  *ms = (double)t / iters;
  return 0;
}

}  // extern "C"

#include "capi_tensor.inc"

// spk_api.hip -- C-ABI of the SIE solve loop: handle, precalculation
// (lib_sie_ml_fmm_precalculation_tri), matvec
// (lib_sie_ml_fmm_calculate_vector_b_tri), right-hand side
// (get_vector_b_tri) and the CERFACS GMRES driver
// (lib_math_solver_gmres_run) on HBM-resident vectors.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <future>
#include <thread>
#include <type_traits>
#include <vector>

#include "spk_device.h"
#include "spk_dbuf.h"
#include "spk_quad.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include "spk_tree.h"

using namespace spk;

namespace spk {
thread_local std::string g_err;
int set_error(int code, const std::string &msg) {
  g_err = msg;
  return code;
}
}  // namespace spk

namespace {

int fail(int code, const std::string &msg) { return spk::set_error(code, msg); }

#define HIPCHK(expr)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(SPK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));     \
  } while (0)

using zc = std::complex<double>;
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// buffers of the batched matvec (R right-hand sides), one set per precision
template <class T>
struct BatchBufs {
  int R = 0;
  std::vector<DBuf<T>> B, D;  // per level, [box][it][s][r]
  DBuf<T> W, YF, YP, XI, XP;  // folded leaf coefficients; far part [r][2N] and its partial
                              // slices; interleaved x
  DBuf<uint32_t> XB;          // XI in bf16 (the fp8 preconditioner product)
  DBuf<const T *> xs;
  DBuf<T *> ys;
};

struct LevelDev {
  DBuf<zd> B, D, tlt, up, down;
  DBuf<zf> fB, fD, ftlt, fup, fdown;  // complex64 mirror (mixed precision)
  DBuf<int> child, parent, octant, i4_off, i4_box, i4_slot;
  DBuf<int> i4_slotm;  // mirrored offset slots (342 - slot): the transposed M2L
  DBuf<double> rhat;  // TL inputs: slot directions, per-slot Hankel factors, live slots
  DBuf<zd> coef;
  DBuf<int> live;
  int nb = 0;
  // the level's truncation number, samples and grid (directions, weights)
  int K = 0, S = 0;
  DBuf<double> khat, trf;
  // where the parent's K differs: the interpolation from this level's grid
  // to the parent's on one (theta, phi) pair (SS), and the map of the
  // downward pass to this level's grid (RR: its transpose, or the
  // interpolation from the parent's grid); CSR
  DBuf<int> up_row, up_col, dn_row, dn_col;
  DBuf<double> up_val, dn_val;
};

}  // namespace

struct spk_handle {
  spk_params p{};
  Material mat{};
  Tables tb;
  int ngp = 3;
  double qa[20], qb[20], qw[20], qa12[20], qb12[20], qw12[20];
  hipStream_t st = nullptr;
  bool ready = false;
  // host mesh/topology
  int N = 0;
  std::vector<int> corner, element;
  std::vector<FnRec> fn;
  std::vector<ElemRec> er;
  std::vector<double> rref;
  Tree tree;
  // device
  DBuf<FnRec> d_fn;
  DBuf<ElemRec> d_er;
  DBuf<double> d_rref;
  DBuf<int> d_perm, d_box_first, d_box_count, d_e2_off, d_e2_list, d_leaf_box, d_leafbox_of_pos;
  DBuf<int64_t> d_near_off;
  DBuf<zd> d_blocks;
  DBuf<int> d_row_box, d_row_r0, d_disagg_task;
  DBuf<int> d_mx_box, d_mx_r0;  // groups of row tiles of a box (batched fp8 product)
  int mx_groups = 0;
  int mx_tpw = 1;  // row tiles per group of that map
  // tuned launch parameters (spk_autotune, spk_set_tuning; 0: the defaults),
  // kept across spk_setup_tri calls (an ensemble's meshes are alike)
  int tune_tpw = 0, tune_inner = 0;
  // per-level truncation numbers (spk_set_level_truncation): requested
  // values by level, the downward mode, the values of the current tree
  std::vector<int> levK_req;
  int rr_mode = SPK_RR_ANTERPOLATE;
  std::vector<int> levK;
  bool multik = false;
  std::vector<Tables> tbl;  // per level (all equal to tb without multik)
  DBuf<zd> lvl_tmp;         // coefficients of one level on its parent's grid
  TreeWork tw;  // scratch of the device octree build
  int disagg_tasks = 0;
  // batched matvec: disaggregation tiles (box, first position), buffers
  DBuf<int> d_tile_box, d_tile_p0;
  int ntiles = 0;
  // bf16 copy of the complex64 near blocks (near-field preconditioner), built on first use
  DBuf<uint32_t> d_blk_bf;
  // fp8 (E4M3FN) copy with per-row, per-component scales (single-vector preconditioner)
  DBuf<uint32_t> d_blk_f8;
  DBuf<float> d_f8_scale;
  bool near_bf_ok = false, near_f8_ok = false;  // the copies match the current blocks
  // matrix-core disaggregation tiles: (box, first position) of up to 128 edges
  DBuf<int> d_dx_box, d_dx_p0;
  int dx_tiles = 0;
  BatchBufs<zd> bd;
  BatchBufs<zf> bf;
  int spmv_wgs = 0;
  DBuf<zd> d_rt, d_W;
  // 8-node quadrilaterals with 10 edge functions (spk_setup_quad): the mesh
  // and neighbour arrays, kernel constants, the patterns' quadrature points,
  // element-local solution coefficients (scattered field), fold constants.
  // Coefficient buffers of every level then hold two halves of 8 channels
  // ((TJx, TJy) and (TJz, TD), each [comp*4 + cur*2 + k]).
  bool quad = false;
  QuadMesh qmesh;
  QuadDev qdev{};
  zd q_eta_a{}, q_eta12{};
  QuadFoldC qfold{};
  DBuf<double> q_el, q_pts;
  DBuf<int> q_edge_p, q_eadj, q_eside, q_cadj;
  DBuf<zd> q_seh;
  // block-Jacobi preconditioner (SPK_PRECOND_BLOCK_LU): per group of leaf
  // boxes (the children of one box a level up) the inverse of its dense
  // near-field matrix, padded to nmax, applied as one batched GEMV; built on
  // first use for the current operator
  struct BlockLU {
    bool ok = false;
    int ngroups = 0, nmax = 0;
    DBuf<zd> A, X, Y;
    DBuf<int> p0, cnt, ipiv, info;
    rocblas_handle rb = nullptr;
    ~BlockLU() {
      if (rb) (void)rocblas_destroy_handle(rb);
    }
  } blu;
  QuadTopo qtopo() const {
    return QuadTopo{q_el.p, q_edge_p.p, q_eadj.p, q_eside.p, q_cadj.p, qmesh.nel, qmesh.medge};
  }
  DBuf<double> d_khat, d_tth, d_tph, d_q, d_leaf_centre, d_trf;
  std::vector<LevelDev> lev;
  DBuf<zd> d_x, d_y, d_yfar, d_tmp;
  // the transposed operator (spk_set_transposed): flag, edge -> position map,
  // scaled input of the transposed far field
  bool transposed = false;
  DBuf<int> d_pos;
  std::vector<double> pp_host;  // the mesh nodes of the last setup (the adjoint moves them)
  DBuf<zd> d_xt;
  DBuf<zf> f_xt;
  zd sc1{1, 0}, sc2{1, 0};
  // complex64 mirror of the operator (precision == SPK_MIXED): near blocks,
  // patterns, folded W and the far-part scratch
  DBuf<zf> f_blocks, f_rt, f_W, f_yfar, f_x, f_y;
  bool mixed_ready = false;
  // mixed-precision inner solver workspace
  DBuf<zf> i_V, i_w;
  int i_m = 0;
  int refinements = 0;  // outer steps of the last mixed solve
  DBuf<zd> d_diag;      // Jacobi diagonal, built on first use
  bool diag_ready = false;
  // gmres workspace
  DBuf<zd> g_V, g_w, g_r0, g_xc, g_b, g_h, g_partial, g_scal, g_y;
  // one-reduce MGS: T = (I + L)^-1, [c; g], T^H g, partial sums of V h and of T
  DBuf<zd> g_T, g_c, g_tg, g_pn, g_pT;
  int g_T_m = -1;
  // pinned host copy of one pass's projection coefficients and |w|^2
  struct Pinned {
    zd *p = nullptr;
    size_t n = 0;
    ~Pinned() {
      if (p) (void)hipHostFree(p);
    }
    hipError_t ensure(size_t count) {
      if (count <= n) return hipSuccess;
      if (p) (void)hipHostFree(p);
      p = nullptr;
      n = 0;
      hipError_t e = hipHostMalloc((void **)&p, count * sizeof(zd), hipHostMallocDefault);
      if (e == hipSuccess) n = count;
      return e;
    }
  } pin;
  // near-field preconditioned flexible GMRES workspace, kept between solves
  // (DBuf reuses an allocation that fits): outer bases V, Z, inner bases U
  // (complex64 on mixed handles, complex128 otherwise), pointer tables
  struct FgWork {
    DBuf<zd> V, Z, W, Hd, Rv, Tv, Dv, Yo, Yi, h1, Hi, dots, apart, part;
    DBuf<zf> Uf, Wif;
    DBuf<zd> Ud, Wid;
    DBuf<uint64_t> tin, tprod, torth;
    DBuf<int> kout;
  } fg;
  // the inner near-field solve (inner_iterations steps of near-field product
  // and Arnoldi update, no host synchronisation inside) captured once as a
  // HIP graph and replayed every outer step; rebuilt when its launch
  // configuration or any buffer address changes
  struct InnerGraph {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    void reset() {
      if (exec) (void)hipGraphExecDestroy(exec);
      exec = nullptr;
      key.clear();
    }
    ~InnerGraph() { reset(); }
  } ig;
  uint64_t setup_serial = 0;  // spk_setup_tri calls completed or started
  bool graph_ok = true;       // spk_set_inner_graph
  DBuf<unsigned long long> g_ctl;    // MGS sweep abort flag
  DBuf<unsigned char> g_slots;       // MGS sweep exchange slots
  unsigned long long sweep_seq = 0;  // sequence number of the last sweep launch
  int g_m = 0;
  std::vector<double> hist;
  void *scan_tmp = nullptr;
  size_t scan_bytes = 0;
  // near fill scratch: block classes, task counts and offsets, task lists,
  // edge-adjacent K terms and singular-pair partial blocks
  struct NearWork {
    DBuf<uint8_t> cls;
    DBuf<int> ca, cb, ba, bb;
    DBuf<NearTask> ta, tb;
    DBuf<zd> kt, zz;
  } nw;
  spk_stats stats{};
  ~spk_handle() {
    if (scan_tmp) (void)hipFree(scan_tmp);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

NearArgs near_args(const spk_handle *h) {
  NearArgs a{};
  a.fn = h->d_fn.p;
  a.er = h->d_er.p;
  a.rref = h->d_rref.p;
  a.perm = h->d_perm.p;
  a.box_first = h->d_box_first.p;
  a.box_count = h->d_box_count.p;
  a.e2_off = h->d_e2_off.p;
  a.e2_list = h->d_e2_list.p;
  a.near_off = h->d_near_off.p;
  a.nbl = (int)h->tree.box_first.size();
  a.form = h->p.formulation;
  a.ngp = h->ngp;
  a.edge_swap = (a.form == SPK_PMCHWT || a.form == SPK_MCTF || a.form == SPK_PMCHWT_S ||
                 a.form == SPK_JMCFIE);
  a.mat = h->mat;
  for (int i = 0; i < 12; ++i) {
    a.qa[i] = h->qa[i];
    a.qb[i] = h->qb[i];
    a.qw[i] = h->qw[i];
    a.qa12[i] = h->qa12[i];
    a.qb12[i] = h->qb12[i];
    a.qw12[i] = h->qw12[i];
  }
  return a;
}

int near_fill(spk_handle *h) {
  const int64_t nnz = h->tree.nnz;
  HIPCHK(h->d_blocks.alloc(4 * (size_t)nnz));
  if (nnz == 0) return SPK_OK;
  NearArgs a = near_args(h);
  auto &NW = h->nw;
  DBuf<uint8_t> &cls = NW.cls;
  DBuf<int> &ca = NW.ca, &cb = NW.cb, &ba = NW.ba, &bb = NW.bb;
  HIPCHK(cls.alloc(nnz));
  HIPCHK(ca.alloc(nnz + 1));
  HIPCHK(cb.alloc(nnz + 1));
  HIPCHK(ba.alloc(nnz + 1));
  HIPCHK(bb.alloc(nnz + 1));
  HIPCHK(hipMemsetAsync(ca.p + nnz, 0, sizeof(int), h->st));
  HIPCHK(hipMemsetAsync(cb.p + nnz, 0, sizeof(int), h->st));
  launch_near_classify(a, nnz, cls.p, ca.p, cb.p, h->st);
  HIPCHK(exclusive_scan_int(ca.p, ba.p, nnz + 1, &h->scan_tmp, &h->scan_bytes, h->st));
  HIPCHK(exclusive_scan_int(cb.p, bb.p, nnz + 1, &h->scan_tmp, &h->scan_bytes, h->st));
  int na = 0, nb = 0;
  HIPCHK(hipMemcpyAsync(&na, ba.p + nnz, sizeof(int), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(&nb, bb.p + nnz, sizeof(int), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipGetLastError());
  DBuf<NearTask> &ta = NW.ta, &tb = NW.tb;
  DBuf<zd> &kt = NW.kt, &zz = NW.zz;
  HIPCHK(ta.alloc(std::max(na, 1)));
  HIPCHK(tb.alloc(std::max(nb, 1)));
  HIPCHK(kt.alloc(2 * (size_t)std::max(nb, 1)));
  HIPCHK(zz.alloc(4 * (size_t)std::max(na, 1)));
  launch_near_emit(a, nnz, cls.p, ba.p, bb.p, ta.p, tb.p, h->st);
  launch_near_edge(a, nb, tb.p, kt.p, h->st);
  launch_near_singular(a, na, ta.p, cls.p, bb.p, kt.p, zz.p, h->st);
  launch_near_regular(a, nnz, cls.p, ba.p, zz.p, h->d_blocks.p, h->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->st));
  h->stats.singular_pairs = na;
  return SPK_OK;
}

// Near field of a quadrilateral handle.  For leaf box b, T(b) holds the
// elements of its E2 columns (each column's one or two elements); the
// pairs (row edge function, element of T(b)) are computed once, then every
// block combines the pairs of its column's elements as the dense assembly
// does (lib_sie_quad_solver_normal.f90:165-210).  Boxes go in chunks whose
// pair sums fit kQuadPairChunk pairs.
constexpr int64_t kQuadPairChunk = 1 << 22;

int near_fill_quad(spk_handle *h) {
  const Tree &t = h->tree;
  const QuadMesh &m = h->qmesh;
  const int nbl = (int)t.box_first.size();
  const int64_t nnz = t.nnz;
  HIPCHK(h->d_blocks.alloc(4 * (size_t)nnz));
  if (nnz == 0) return SPK_OK;
  std::vector<int> tb_off(nbl + 1, 0), tb_list, colA(t.e2_list.size()), colB(t.e2_list.size(), -1);
  std::vector<int64_t> pbase(nbl + 1, 0), toff(nbl + 1, 0);
  std::vector<int> els;
  for (int b = 0; b < nbl; ++b) {
    els.clear();
    for (int ci = t.e2_off[b]; ci < t.e2_off[b + 1]; ++ci) {
      const int *e = &m.edge_p[5 * (size_t)t.e2_list[ci]];
      els.push_back(e[0] - 1);
      if (e[2] > 0) els.push_back(e[2] - 1);
    }
    std::sort(els.begin(), els.end());
    els.erase(std::unique(els.begin(), els.end()), els.end());
    auto idx = [&](int el) {
      return (int)(std::lower_bound(els.begin(), els.end(), el) - els.begin());
    };
    for (int ci = t.e2_off[b]; ci < t.e2_off[b + 1]; ++ci) {
      const int *e = &m.edge_p[5 * (size_t)t.e2_list[ci]];
      colA[ci] = idx(e[0] - 1);
      if (e[2] > 0) colB[ci] = idx(e[2] - 1);
    }
    tb_off[b + 1] = tb_off[b] + (int)els.size();
    tb_list.insert(tb_list.end(), els.begin(), els.end());
    pbase[b + 1] = pbase[b] + (int64_t)t.box_count[b] * (int64_t)els.size();
    toff[b + 1] = toff[b] + (int64_t)((t.box_count[b] + 63) / 64) * (int64_t)els.size();
  }
  DBuf<int> d_tb_off, d_tb_list, d_colA, d_colB, l_es, l_et, l_side, l_cs, l_ct;
  DBuf<int64_t> d_pbase, d_toff, l_sdst, l_edst, l_cdst;
  DBuf<zd> sums, so, eo, co;
  HIPCHK(d_tb_off.upload(tb_off));
  HIPCHK(d_tb_list.upload(tb_list));
  HIPCHK(d_colA.upload(colA));
  HIPCHK(d_colB.upload(colB));
  HIPCHK(d_pbase.upload(pbase));
  HIPCHK(d_toff.upload(toff));
  const QuadTopo T = h->qtopo();
  std::vector<int> ls, es, et, eside, cs, ct;
  std::vector<int64_t> sdst, edst, cdst;
  int64_t max_pairs = 0;
  for (int b = 0; b < nbl; ++b) max_pairs = std::max(max_pairs, pbase[b + 1] - pbase[b]);
  HIPCHK(sums.alloc((size_t)std::max(max_pairs, std::min(kQuadPairChunk, pbase[nbl])) * 30));
  int singular = 0;
  for (int b0 = 0; b0 < nbl;) {
    int b1 = b0 + 1;
    while (b1 < nbl && pbase[b1 + 1] - pbase[b0] <= kQuadPairChunk) ++b1;
    ls.clear();
    sdst.clear();
    es.clear();
    et.clear();
    eside.clear();
    edst.clear();
    cs.clear();
    ct.clear();
    cdst.clear();
    for (int b = b0; b < b1; ++b) {
      const int *tl = &tb_list[tb_off[b]];
      const int nT = tb_off[b + 1] - tb_off[b];
      for (int r = 0; r < t.box_count[b]; ++r) {
        const int s = t.perm[t.box_first[b] + r];
        const int ii = m.edge_p[5 * (size_t)s] - 1;
        const int64_t q0 = pbase[b] - pbase[b0] + (int64_t)r * nT;
        const int *pi = std::lower_bound(tl, tl + nT, ii);
        ls.push_back(s);
        sdst.push_back(q0 + (pi - tl));
        for (const QuadEdgeAdj &a : m.eadj[ii]) {
          const int *pt = std::lower_bound(tl, tl + nT, a.other - 1);
          if (pt == tl + nT || *pt != a.other - 1) continue;
          es.push_back(s);
          et.push_back(a.other - 1);
          eside.push_back(a.side[0]);
          eside.push_back(a.side[1]);
          edst.push_back(q0 + (pt - tl));
        }
        // corner-adjacent pairs: the 6 x 6 rule (pair_class order: an edge
        // neighbour is never also taken as a corner neighbour)
        for (int c : m.cadj[ii]) {
          bool edge = false;
          for (const QuadEdgeAdj &a : m.eadj[ii]) edge = edge || a.other == c;
          if (edge || c - 1 == ii) continue;
          const int *pt = std::lower_bound(tl, tl + nT, c - 1);
          if (pt == tl + nT || *pt != c - 1) continue;
          cs.push_back(s);
          ct.push_back(c - 1);
          cdst.push_back(q0 + (pt - tl));
        }
      }
    }
    launch_qn_regular(h->qdev, T, b0, b1, toff[b1] - toff[b0], d_toff.p, d_pbase.p, d_tb_off.p,
                      d_tb_list.p, h->d_box_first.p, h->d_box_count.p, h->d_perm.p, sums.p, h->st);
    HIPCHK(l_sdst.upload(sdst));
    HIPCHK(so.alloc(ls.size() * 30));
    HIPCHK(quad_self_rows(h->qdev, T, m, ls, so.p, h->st));
    HIPCHK(l_es.upload(es));
    HIPCHK(l_et.upload(et));
    HIPCHK(l_side.upload(eside));
    HIPCHK(l_edst.upload(edst));
    HIPCHK(eo.alloc(std::max<size_t>(1, es.size()) * 30));
    launch_quad_edge_adj(h->qdev, T, (int)es.size(), l_es.p, l_et.p, l_side.p, eo.p, h->st);
    HIPCHK(l_cs.upload(cs));
    HIPCHK(l_ct.upload(ct));
    HIPCHK(l_cdst.upload(cdst));
    HIPCHK(co.alloc(std::max<size_t>(1, cs.size()) * 30));
    launch_quad_corner(h->qdev, T, (int)cs.size(), l_cs.p, l_ct.p, co.p, h->st);
    launch_qn_scatter((int)ls.size(), l_sdst.p, so.p, sums.p, h->st);
    launch_qn_scatter((int)es.size(), l_edst.p, eo.p, sums.p, h->st);
    launch_qn_scatter((int)cs.size(), l_cdst.p, co.p, sums.p, h->st);
    launch_qn_assemble(T, h->p.formulation, h->q_eta_a, h->q_eta12, b0, b1,
                       t.near_off[b1] - t.near_off[b0], h->d_near_off.p, h->d_e2_off.p,
                       h->d_e2_list.p, d_colA.p, d_colB.p, d_pbase.p, d_tb_off.p, sums.p,
                       h->d_blocks.p, h->st);
    HIPCHK(hipGetLastError());
    // the host lists are reused by the next chunk
    HIPCHK(hipStreamSynchronize(h->st));
    singular += (int)(ls.size() + es.size());
    b0 = b1;
  }
  h->stats.singular_pairs = singular;
  return SPK_OK;
}

// the host's per-level table inputs: TL slot factors and SS/RR phases
struct HostLevelTables {
  std::vector<double> rhat;
  std::vector<cplx> coef, up, down;
  std::vector<int> live;
};
// tbl: the sample grid of every level; SS/RR phases of child level l on
// its parent's grid
static std::vector<HostLevelTables> host_level_tables(const std::vector<Tables> &tbl,
                                                      const Material &mat, const Tree &t) {
  std::vector<HostLevelTables> v(t.l_max + 1);
  std::vector<std::thread> th;
  for (int l = t.l_min; l <= t.l_max; ++l)
    th.emplace_back([&, l]() {
      tl_coefficients(tbl[l], mat, t, l, &v[l].rhat, &v[l].coef, &v[l].live);
      make_phase_level(tbl[l > t.l_min ? l - 1 : l], mat, t, l, &v[l].up, &v[l].down);
    });
  for (auto &x : th) x.join();
  return v;
}

// pre: the host tables (computed while the near fill runs on the device)
int setup_mlfmm(spk_handle *h, std::future<std::vector<HostLevelTables>> &pre) {
  const Tree &t = h->tree;
  const int S = h->tb.S, N = h->N;
  double t0 = now_ms();
  // leaf centres
  const Level &LL = t.lev[t.l_max];
  std::vector<double> lc(3 * LL.n.size());
  for (size_t b = 0; b < LL.n.size(); ++b) box_centre(t, LL.n[b], t.l_max, &lc[3 * b]);
  HIPCHK(h->d_leaf_centre.upload(lc));
  HIPCHK(h->d_khat.upload(h->tb.khat));
  HIPCHK(h->d_trf.upload(h->tb.trf));
  HIPCHK(h->d_tth.upload(h->tb.ttheta));
  HIPCHK(h->d_tph.upload(h->tb.tphi));
  std::vector<double> q(60);
  for (int i = 0; i < 20; ++i) {
    q[i] = h->qa[i];
    q[20 + i] = h->qb[i];
    q[40 + i] = h->qw[i];
  }
  HIPCHK(h->d_q.upload(q));
  if (h->quad) {
    // quadrature points of every position's patterns, then the patterns
    HIPCHK(h->q_pts.alloc((size_t)N * kQuadPtStride));
    launch_qpoints(h->qdev, h->qtopo(), N, h->d_perm.p, h->d_leaf_box.p, h->d_leaf_centre.p,
                   h->q_pts.p, h->st);
    HIPCHK(h->d_rt.alloc((size_t)N * kQuadRtRows * S));
    launch_qpatterns(N, S, h->d_khat.p, h->q_pts.p, h->qdev.k1, h->qdev.k2, h->d_rt.p, h->st);
  } else {
    HIPCHK(h->d_rt.alloc((size_t)N * 4 * S));
    launch_patterns(N, S, h->ngp, h->d_fn.p, h->d_perm.p, h->d_leaf_box.p, h->d_leaf_centre.p,
                    h->d_khat.p, h->d_tth.p, h->d_tph.p, h->d_q.p, h->d_q.p + 20, h->d_q.p + 40,
                    zcx(h->mat.k1), zcx(h->mat.k2), h->d_rt.p, h->st);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->st));
  h->stats.patterns_ms = now_ms() - t0;
  t0 = now_ms();
  const std::vector<HostLevelTables> ht = pre.get();
  // levels
  if (h->lev.size() < (size_t)t.l_max + 1) h->lev.resize(t.l_max + 1);
  const int nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  size_t tmp_need = 0;
  for (int l = t.l_min; l <= t.l_max; ++l) {
    const Level &L = t.lev[l];
    LevelDev &D = h->lev[l];
    D.nb = (int)L.n.size();
    const Tables &tl = h->tbl[l];
    D.K = tl.K;
    D.S = tl.S;
    HIPCHK(D.khat.upload(tl.khat));
    HIPCHK(D.trf.upload(tl.trf));
    // two halves of 8 channels on quadrilateral handles
    const size_t halves = h->quad ? 2 : 1;
    HIPCHK(D.B.alloc(halves * D.nb * 8 * D.S));
    HIPCHK(D.D.alloc(halves * D.nb * 8 * D.S));
    if (l > t.l_min && h->levK[l - 1] != h->levK[l]) {
      // interpolation to the parent's grid, and the downward map
      std::vector<int> r, c, tr, tc;
      std::vector<double> v, tv;
      interp_pair_csr(3, h->levK[l], h->levK[l - 1], &r, &c, &v);
      HIPCHK(D.up_row.upload(r));
      HIPCHK(D.up_col.upload(c));
      HIPCHK(D.up_val.upload(v));
      if (h->rr_mode == SPK_RR_ANTERPOLATE) {
        csr_transpose(2 * D.S, r, c, v, &tr, &tc, &tv);
      } else {
        interp_pair_csr(3, h->levK[l - 1], h->levK[l], &tr, &tc, &tv);
      }
      HIPCHK(D.dn_row.upload(tr));
      HIPCHK(D.dn_col.upload(tc));
      HIPCHK(D.dn_val.upload(tv));
      tmp_need = std::max(tmp_need, (size_t)D.nb * 8 * h->tbl[l - 1].S);
    }
    if (l < t.l_max) HIPCHK(D.child.upload(L.child));
    HIPCHK(D.parent.upload(L.parent));
    HIPCHK(D.octant.upload(L.octant));
    HIPCHK(D.i4_off.upload(L.i4_off));
    HIPCHK(D.i4_box.upload(L.i4_box));
    HIPCHK(D.i4_slot.upload(L.i4_slot));
    {
      std::vector<int> mir(L.i4_slot.size());
      for (size_t q = 0; q < mir.size(); ++q) mir[q] = 342 - L.i4_slot[q];
      HIPCHK(D.i4_slotm.upload(mir));
    }
    {
      // TL on the device from the host's per-slot Hankel factors
      const std::vector<double> &rhat = ht[l].rhat;
      const std::vector<cplx> &coef = ht[l].coef;
      const std::vector<int> &live = ht[l].live;
      std::vector<zd> cz(coef.size());
      for (size_t i = 0; i < coef.size(); ++i) cz[i] = zd{coef[i].re, coef[i].im};
      DBuf<double> &d_rhat = D.rhat;
      DBuf<zd> &d_coef = D.coef;
      DBuf<int> &d_live = D.live;
      HIPCHK(d_rhat.upload(rhat));
      HIPCHK(d_coef.upload(cz));
      HIPCHK(d_live.upload(live));
      HIPCHK(D.tlt.alloc(343 * 2 * (size_t)D.S));
      launch_tl_level(D.S, D.K, D.khat.p, D.trf.p, d_rhat.p, d_coef.p, d_live.p, D.tlt.p, h->st);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(h->st));
    }
    const std::vector<cplx> &up = ht[l].up, &down = ht[l].down;
    std::vector<zd> uz(up.size()), dz(down.size());
    for (size_t i = 0; i < up.size(); ++i) {
      uz[i] = zd{up[i].re, up[i].im};
      dz[i] = zd{down[i].re, down[i].im};
    }
    HIPCHK(D.up.upload(uz));
    HIPCHK(D.down.upload(dz));
    h->stats.boxes_per_level[l] = D.nb;
    h->stats.i4_pairs_per_level[l] = (int64_t)L.i4_box.size();
  }
  HIPCHK(h->d_W.alloc((size_t)LL.n.size() * (h->quad ? 12 : 8) * S));
  if (tmp_need) HIPCHK(h->lvl_tmp.alloc(tmp_need));
  h->stats.tables_ms = now_ms() - t0;
  // formulation scaling of the far field (MLFMM_tri_summation_New :478-495)
  const zc e1(h->mat.eta1.re, h->mat.eta1.im), e2(h->mat.eta2.re, h->mat.eta2.im);
  const zc ea(h->mat.eta_a.re, h->mat.eta_a.im);
  zc s1 = 1.0, s2 = 1.0;
  switch (h->p.formulation) {
    case SPK_PMCHWT_S: s2 = e1; break;
    case SPK_MCTF: s2 = e1 * e2; break;
    case SPK_ICTF: s1 = 1.0 / ea; s2 = ea; break;
    case SPK_MCTF2:
      s1 = zc(h->mat.a_lc[0].re, h->mat.a_lc[0].im) / e1;
      s2 = zc(h->mat.d_lc[0].re, h->mat.d_lc[0].im) * e1;
      break;
    default: break;
  }
  h->sc1 = zd{s1.real(), s1.imag()};
  h->sc2 = zd{s2.real(), s2.imag()};
  if (h->quad) {
    // the quadrilateral rows' fold constants: the dense fill's (A, DE, DH)
    // combination (combine of ff_ABC_II_Edge's sums) in plane-wave form, and
    // the assembly's formulation scalings of the four blocks
    const QuadDev &Q = h->qdev;
    const zc iwe(Q.iwe.x, Q.iwe.y), iwm(Q.iwm.x, Q.iwm.y);
    const zc ea(h->q_eta_a.x, h->q_eta_a.y), e12(h->q_eta12.x, h->q_eta12.y);
    zc sa = 1.0, sb = 1.0, sc = 1.0, sd = 1.0;
    if (h->p.formulation == SPK_ICTF) {
      sa = sb = 1.0 / ea;
      sc = sd = ea;
    } else if (h->p.formulation == SPK_MCTF) {
      sc = sd = e12;
    }
    auto Z = [](zc c) { return zd{c.real(), c.imag()}; };
    for (int k = 0; k < 2; ++k) {
      const zc km = k ? zc(Q.k2.x, Q.k2.y) : zc(Q.k1.x, Q.k1.y);
      const zc em = k ? zc(Q.e2.x, Q.e2.y) : zc(Q.e1.x, Q.e1.y);
      const zc mik = zc(0.0, -1.0) * km;
      h->qfold.cAE[k] = Z(sa * mik);
      h->qfold.cJE[k] = Z(sb * iwm);
      h->qfold.cDE[k] = Z(sb * mik / (em * iwe));
      h->qfold.cJH[k] = Z(-sc * iwe * em);
      h->qfold.cDH[k] = Z(-sc * mik / iwm);
      h->qfold.cAH[k] = Z(sd * mik);
    }
  }
  return SPK_OK;
}

// Buffers of the operator in complex128 (zd) or its complex64 mirror (zf).
static zd *pB(LevelDev &L, zd *) { return L.B.p; }
static zf *pB(LevelDev &L, zf *) { return L.fB.p; }
static zd *pD(LevelDev &L, zd *) { return L.D.p; }
static zf *pD(LevelDev &L, zf *) { return L.fD.p; }
static const zd *pT(LevelDev &L, zd *) { return L.tlt.p; }
static const zf *pT(LevelDev &L, zf *) { return L.ftlt.p; }
static const zd *pUp(LevelDev &L, zd *) { return L.up.p; }
static const zf *pUp(LevelDev &L, zf *) { return L.fup.p; }
static const zd *pDn(LevelDev &L, zd *) { return L.down.p; }
static const zf *pDn(LevelDev &L, zf *) { return L.fdown.p; }
static const zd *pRt(spk_handle *h, zd *) { return h->d_rt.p; }
static const zf *pRt(spk_handle *h, zf *) { return h->f_rt.p; }
static zd *pW(spk_handle *h, zd *) { return h->d_W.p; }
static zf *pW(spk_handle *h, zf *) { return h->f_W.p; }
static const zd *pBlk(spk_handle *h, zd *) { return h->d_blocks.p; }
static const zf *pBlk(spk_handle *h, zf *) { return h->f_blocks.p; }
static zd *pYf(spk_handle *h, zd *) { return h->d_yfar.p; }
static zf *pYf(spk_handle *h, zf *) { return h->f_yfar.p; }
static DBuf<zd> &pXt(spk_handle *h, zd *) { return h->d_xt; }
static DBuf<zf> &pXt(spk_handle *h, zf *) { return h->f_xt; }

// y = Z x with per-level truncation numbers (complex128).  Upward: where a
// parent's K exceeds its children's, the children's coefficients are
// interpolated to the parent's grid (up CSR) before SS.  Downward, with
// SPK_RR_REFERENCE, RR shifts on the parent's grid and interpolates to the
// child's (dn CSR); with SPK_RR_ANTERPOLATE every level's M2L sums are
// folded first and RR works on folded coefficients, applying the transpose
// of the upward interpolation (dn CSR) where K changes; the leaf's folded
// coefficients then go straight to the disaggregation.
static const int kPairsC[8] = {0 * 4 + 0 * 2 + 0, 1 * 4 + 0 * 2 + 0, 0 * 4 + 0 * 2 + 1,
                               1 * 4 + 0 * 2 + 1, 0 * 4 + 1 * 2 + 0, 1 * 4 + 1 * 2 + 0,
                               0 * 4 + 1 * 2 + 1, 1 * 4 + 1 * 2 + 1};
static const int kPairsW[8] = {0, 1, 2, 3, 4, 5, 6, 7};

static int matvec_multik(spk_handle *h, const zd *x, zd *y, bool near_on, bool far_on) {
  const Tree &t = h->tree;
  const int N = h->N, lmax = t.l_max, lmin = t.l_min;
  hipStream_t st = h->st;
  zd *yfar = h->d_yfar.p;
  if (far_on) {
    LevelDev &LL = h->lev[lmax];
    const int SL = LL.S;
    launch_agg(LL.nb, SL, N, h->d_box_first.p, h->d_box_count.p, h->d_perm.p, h->d_rt.p, x, LL.B.p,
               st, true);
    for (int l = lmax - 1; l >= lmin; --l) {
      LevelDev &P = h->lev[l], &Cl = h->lev[l + 1];
      const zd *src = Cl.B.p;
      if (Cl.K != P.K) {
        launch_resample(Cl.nb, Cl.S, P.S, Cl.up_row.p, Cl.up_col.p, Cl.up_val.p, kPairsC, Cl.B.p,
                        h->lvl_tmp.p, false, st);
        src = h->lvl_tmp.p;
      }
      launch_ss(P.nb, P.S, P.child.p, src, (const zd *)Cl.up.p, P.B.p, st);
    }
    for (int l = lmin; l <= lmax; ++l)
      launch_m2l(h->lev[l].nb, h->lev[l].S, h->lev[l].i4_off.p, h->lev[l].i4_box.p,
                 h->lev[l].i4_slot.p, (const zd *)h->lev[l].B.p, (const zd *)h->lev[l].tlt.p,
                 h->lev[l].D.p, st);
    const zd I = {0, 1};
    const zd kd1 = I * zcx(h->mat.k1), kd2 = I * zcx(h->mat.k2);
    const zd my1 = zd{h->mat.my1, 0} * I * h->mat.omega, my2 = zd{h->mat.my2, 0} * I * h->mat.omega;
    const zd ep1 = zcx(h->mat.eps1) * I * h->mat.omega, ep2 = zcx(h->mat.eps2) * I * h->mat.omega;
    const zd *W = nullptr;
    if (h->rr_mode == SPK_RR_REFERENCE) {
      for (int l = lmin + 1; l <= lmax; ++l) {
        LevelDev &Cl = h->lev[l], &P = h->lev[l - 1];
        if (Cl.K != P.K) {
          launch_shift_write(Cl.nb, P.S, Cl.parent.p, Cl.octant.p, P.D.p, (const zd *)Cl.down.p,
                             h->lvl_tmp.p, false, st);
          launch_resample(Cl.nb, P.S, Cl.S, Cl.dn_row.p, Cl.dn_col.p, Cl.dn_val.p, kPairsC,
                          h->lvl_tmp.p, Cl.D.p, true, st);
        } else {
          launch_rr(Cl.nb, Cl.S, Cl.parent.p, Cl.octant.p, (const zd *)P.D.p,
                    (const zd *)Cl.down.p, Cl.D.p, st);
        }
      }
      launch_fold(LL.nb, SL, (const zd *)LL.D.p, kd1, kd2, my1, my2, ep1, ep2, h->d_W.p, st);
      W = h->d_W.p;
    } else {
      // folded coefficients of every level into its (consumed) B buffer
      for (int l = lmin; l <= lmax; ++l)
        launch_fold(h->lev[l].nb, h->lev[l].S, (const zd *)h->lev[l].D.p, kd1, kd2, my1, my2, ep1,
                    ep2, h->lev[l].B.p, st);
      for (int l = lmin + 1; l <= lmax; ++l) {
        LevelDev &Cl = h->lev[l], &P = h->lev[l - 1];
        if (Cl.K != P.K) {
          launch_shift_write(Cl.nb, P.S, Cl.parent.p, Cl.octant.p, P.B.p, (const zd *)Cl.down.p,
                             h->lvl_tmp.p, true, st);
          launch_resample(Cl.nb, P.S, Cl.S, Cl.dn_row.p, Cl.dn_col.p, Cl.dn_val.p, kPairsW,
                          h->lvl_tmp.p, Cl.B.p, true, st);
        } else {
          launch_rr_folded(Cl.nb, Cl.S, Cl.parent.p, Cl.octant.p, (const zd *)P.B.p,
                           (const zd *)Cl.down.p, Cl.B.p, st);
        }
      }
      W = LL.B.p;
    }
    launch_disagg(h->disagg_tasks, N, SL, h->d_disagg_task.p, h->d_perm.p, h->d_leafbox_of_pos.p,
                  h->d_rt.p, W, h->sc1, h->sc2, near_on ? yfar : y, st, false);
  }
  if (near_on)
    launch_near_spmv_rows(h->spmv_wgs, N, h->d_row_box.p, h->d_row_r0.p, h->d_box_first.p,
                          h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p,
                          h->d_perm.p, (const zd *)h->d_blocks.p, x, far_on ? (const zd *)yfar : nullptr, y,
                          st);
  HIPCHK(hipGetLastError());
  return SPK_OK;
}

// y = Z x of a quadrilateral handle: the radiating patterns aggregated into
// 16 channels per sample (two halves of the levels' buffers), SS, M2L and
// RR on each half, the fold into the rows' two vectors per medium, the
// receiving patterns' disaggregation; the near field as for RWG edges.
template <class T>
static int matvec_quad_t(spk_handle *h, const T *x, T *y, bool near_on, bool far_on,
                         hipEvent_t *ev) {
  T *tag = nullptr;
  const Tree &t = h->tree;
  const int S = h->tb.S, N = h->N, lmax = t.l_max, lmin = t.l_min;
  hipStream_t st = h->st;
  T *yfar = pYf(h, tag);
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
  auto half = [&](int l) { return (size_t)h->lev[l].nb * 8 * S; };
  mark(0);
  if (far_on) {
    LevelDev &LL = h->lev[lmax];
    launch_qagg(LL.nb, S, N, h->d_box_first.p, h->d_box_count.p, h->d_perm.p, pRt(h, tag), x,
                pB(LL, tag), half(lmax), st);
    mark(1);
    for (int l = lmax - 1; l >= lmin; --l)
      for (int c = 0; c < 2; ++c)
        launch_ss(h->lev[l].nb, S, h->lev[l].child.p, (const T *)pB(h->lev[l + 1], tag) + c * half(l + 1),
                  pUp(h->lev[l + 1], tag), pB(h->lev[l], tag) + c * half(l), st);
    mark(2);
    for (int l = lmin; l <= lmax; ++l)
      for (int c = 0; c < 2; ++c)
        launch_m2l(h->lev[l].nb, S, h->lev[l].i4_off.p, h->lev[l].i4_box.p, h->lev[l].i4_slot.p,
                   (const T *)pB(h->lev[l], tag) + c * half(l), pT(h->lev[l], tag),
                   pD(h->lev[l], tag) + c * half(l), st);
    mark(3);
    for (int l = lmin + 1; l <= lmax; ++l)
      for (int c = 0; c < 2; ++c)
        launch_rr(h->lev[l].nb, S, h->lev[l].parent.p, h->lev[l].octant.p,
                  (const T *)pD(h->lev[l - 1], tag) + c * half(l - 1), pDn(h->lev[l], tag),
                  pD(h->lev[l], tag) + c * half(l), st);
    mark(4);
    launch_qfold(LL.nb, S, (const T *)pD(LL, tag), half(lmax), h->d_khat.p, h->qfold, pW(h, tag), st);
    mark(5);
    launch_qdisagg(h->disagg_tasks, N, S, h->d_disagg_task.p, h->d_perm.p, h->d_leafbox_of_pos.p,
                   pRt(h, tag), (const T *)pW(h, tag), near_on ? yfar : y, st);
    mark(6);
  }
  if (near_on)
    launch_near_spmv_rows(h->spmv_wgs, N, h->d_row_box.p, h->d_row_r0.p, h->d_box_first.p,
                          h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p,
                          h->d_perm.p, pBlk(h, tag), x, far_on ? (const T *)yfar : nullptr, y,
                          st);
  mark(7);
  HIPCHK(hipGetLastError());
  return SPK_OK;
}

// y = Z x on device vectors (2N complex each), far part into yfar first.
// T = zd: the FP64 operator; T = zf: its complex64 mirror.
// ev (optional, 8 events): recorded before stage 0 and after stages 0..6.
template <class T>
int matvec_t(spk_handle *h, const T *x, T *y, bool near_on = true, bool far_on = true,
             hipEvent_t *ev = nullptr) {
  T *tag = nullptr;
  if (h->quad) {
    if (h->transposed) return fail(SPK_ERR_STATE, "quadrilateral handles: no transposed operator");
    return matvec_quad_t<T>(h, x, y, near_on, far_on, ev);
  }
  if (h->multik) {
    if constexpr (std::is_same<T, zd>::value) {
      if (h->transposed) return fail(SPK_ERR_STATE, "per-level truncation numbers: no transposed operator");
      return matvec_multik(h, x, y, near_on, far_on);
    } else {
      return fail(SPK_ERR_STATE, "per-level truncation numbers run on FP64 handles");
    }
  }
  const Tree &t = h->tree;
  const int S = h->tb.S, N = h->N, lmax = t.l_max, lmin = t.l_min;
  hipStream_t st = h->st;
  T *yfar = pYf(h, tag);
  auto mark = [&](int i) {
    if (ev) (void)hipEventRecord(ev[i], st);
  };
  mark(0);
  if (far_on && h->transposed) {
    // Z_far^T x: the stages in reverse, each transposed: the disaggregation
    // as an aggregation of [sc1 x_e; sc2 x_h] with the patterns themselves;
    // the fold's transpose; RR's transpose upwards (children's sums times
    // the down phases); M2L with each I4 pair's mirrored offset; SS's
    // transpose downwards (parents times the up phases, added to M2L's
    // output); the aggregation's transpose as a disaggregation with the
    // conjugated patterns (k_fold with kd = 0, my = ep = 1 reorders the
    // sums into its layout)
    LevelDev &LL = h->lev[lmax];
    DBuf<T> &xt = pXt(h, tag);
    HIPCHK(xt.alloc(2 * (size_t)N));
    launch_scale_halves(N, h->sc1, h->sc2, x, xt.p, st);
    launch_agg(LL.nb, S, N, h->d_box_first.p, h->d_box_count.p, h->d_perm.p, pRt(h, tag),
               (const T *)xt.p, pB(LL, tag), st, false);
    mark(1);
    const zd I = {0, 1};
    const zd kd1 = I * zcx(h->mat.k1), kd2 = I * zcx(h->mat.k2);
    const zd my1 = zd{h->mat.my1, 0} * I * h->mat.omega, my2 = zd{h->mat.my2, 0} * I * h->mat.omega;
    const zd ep1 = zcx(h->mat.eps1) * I * h->mat.omega, ep2 = zcx(h->mat.eps2) * I * h->mat.omega;
    launch_fold_t(LL.nb, S, (const T *)pB(LL, tag), kd1, kd2, my1, my2, ep1, ep2, pD(LL, tag), st);
    for (int l = lmax - 1; l >= lmin; --l)
      launch_ss(h->lev[l].nb, S, h->lev[l].child.p, (const T *)pD(h->lev[l + 1], tag),
                pDn(h->lev[l + 1], tag), pD(h->lev[l], tag), st);
    mark(2);
    for (int l = lmin; l <= lmax; ++l)
      launch_m2l(h->lev[l].nb, S, h->lev[l].i4_off.p, h->lev[l].i4_box.p, h->lev[l].i4_slotm.p,
                 (const T *)pD(h->lev[l], tag), pT(h->lev[l], tag), pB(h->lev[l], tag), st);
    mark(3);
    for (int l = lmin + 1; l <= lmax; ++l)
      launch_rr(h->lev[l].nb, S, h->lev[l].parent.p, h->lev[l].octant.p,
                (const T *)pB(h->lev[l - 1], tag), pUp(h->lev[l], tag), pB(h->lev[l], tag), st);
    mark(4);
    launch_fold(LL.nb, S, (const T *)pB(LL, tag), zd{0, 0}, zd{0, 0}, zd{1, 0}, zd{1, 0},
                zd{1, 0}, zd{1, 0}, pW(h, tag), st);
    mark(5);
    launch_disagg(h->disagg_tasks, N, S, h->d_disagg_task.p, h->d_perm.p, h->d_leafbox_of_pos.p,
                  pRt(h, tag), (const T *)pW(h, tag), zd{1, 0}, zd{1, 0}, near_on ? yfar : y, st,
                  true);
    mark(6);
  } else if (far_on) {
    LevelDev &LL = h->lev[lmax];
    launch_agg(LL.nb, S, N, h->d_box_first.p, h->d_box_count.p, h->d_perm.p, pRt(h, tag), x,
               pB(LL, tag), st);
    mark(1);
    for (int l = lmax - 1; l >= lmin; --l)
      launch_ss(h->lev[l].nb, S, h->lev[l].child.p, (const T *)pB(h->lev[l + 1], tag),
                pUp(h->lev[l + 1], tag), pB(h->lev[l], tag), st);
    mark(2);
    for (int l = lmin; l <= lmax; ++l)
      launch_m2l(h->lev[l].nb, S, h->lev[l].i4_off.p, h->lev[l].i4_box.p, h->lev[l].i4_slot.p,
                 (const T *)pB(h->lev[l], tag), pT(h->lev[l], tag), pD(h->lev[l], tag), st);
    mark(3);
    for (int l = lmin + 1; l <= lmax; ++l)
      launch_rr(h->lev[l].nb, S, h->lev[l].parent.p, h->lev[l].octant.p,
                (const T *)pD(h->lev[l - 1], tag), pDn(h->lev[l], tag), pD(h->lev[l], tag), st);
    mark(4);
    const zd I = {0, 1};
    const zd kd1 = I * zcx(h->mat.k1), kd2 = I * zcx(h->mat.k2);
    const zd my1 = zd{h->mat.my1, 0} * I * h->mat.omega, my2 = zd{h->mat.my2, 0} * I * h->mat.omega;
    const zd ep1 = zcx(h->mat.eps1) * I * h->mat.omega, ep2 = zcx(h->mat.eps2) * I * h->mat.omega;
    launch_fold(LL.nb, S, (const T *)pD(LL, tag), kd1, kd2, my1, my2, ep1, ep2, pW(h, tag), st);
    mark(5);
    launch_disagg(h->disagg_tasks, N, S, h->d_disagg_task.p, h->d_perm.p, h->d_leafbox_of_pos.p,
                  pRt(h, tag), (const T *)pW(h, tag), h->sc1, h->sc2, near_on ? yfar : y, st);
    mark(6);
  }
  if (near_on)
    launch_near_spmv_rows(h->spmv_wgs, N, h->d_row_box.p, h->d_row_r0.p, h->d_box_first.p,
                          h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p,
                          h->d_perm.p, pBlk(h, tag), x, far_on ? (const T *)yfar : nullptr, y,
                          st);
  mark(7);
  HIPCHK(hipGetLastError());
  return SPK_OK;
}

static BatchBufs<zd> &pBB(spk_handle *h, zd *) { return h->bd; }
static BatchBufs<zf> &pBB(spk_handle *h, zf *) { return h->bf; }

// batch buffers for R right-hand sides (kept for later calls with R or fewer)
template <class T>
static int batch_ensure(spk_handle *h, int R) {
  T *tag = nullptr;
  BatchBufs<T> &BB = pBB(h, tag);
  if (BB.R >= R) return SPK_OK;
  const Tree &t = h->tree;
  const size_t S = h->tb.S;
  // allocations that fit are reused (a handle set up again for a mesh of
  // the same size allocates nothing)
  if (BB.B.size() < (size_t)t.l_max + 1) {
    BB.B.resize(t.l_max + 1);
    BB.D.resize(t.l_max + 1);
  }
  for (int l = t.l_min; l <= t.l_max; ++l) {
    HIPCHK(BB.B[l].alloc((size_t)h->lev[l].nb * 8 * S * R));
    HIPCHK(BB.D[l].alloc((size_t)h->lev[l].nb * 8 * S * R));
  }
  HIPCHK(BB.W.alloc((size_t)h->lev[t.l_max].nb * 8 * S * R));
  HIPCHK(BB.YF.alloc((size_t)R * 2 * h->N));
  HIPCHK(BB.YP.alloc((size_t)kBatchSplit * R * 2 * h->N));
  HIPCHK(BB.XI.alloc((size_t)kBatchMax * 2 * h->N));
  HIPCHK(BB.XB.alloc((size_t)kBatchMax * 2 * h->N));
  HIPCHK(BB.XP.alloc((size_t)kBatchMax * 2 * h->N));
  HIPCHK(BB.xs.alloc(kBatchMax));
  HIPCHK(BB.ys.alloc(kBatchMax));
  BB.R = R;
  return SPK_OK;
}

// ys[r] = Z xs[r] for r < R <= kBatchMax; xs, ys: device arrays of R pointers.
// far_on = false: the near-field part only.
template <class T>
int matvec_batch_t(spk_handle *h, int R, const T *const *xs, T *const *ys, bool far_on = true,
                   const uint32_t *blocks_bf = nullptr, const float *f8_scale = nullptr) {
  T *tag = nullptr;
  if (R < 1 || R > kBatchMax) return fail(SPK_ERR_ARG, "batch size must be 1..16");
  if (far_on && h->transposed)
    return fail(SPK_ERR_STATE, "the batched products do not take the transposed operator");
  if (far_on && h->multik)
    return fail(SPK_ERR_STATE, "the batched products do not take per-level truncation numbers");
  if (far_on && h->quad)
    return fail(SPK_ERR_STATE, "quadrilateral handles: batched far-field products are not built");
  int rc = batch_ensure<T>(h, R);
  if (rc) return rc;
  BatchBufs<T> &BB = pBB(h, tag);
  const Tree &t = h->tree;
  const int S = h->tb.S, N = h->N, lmax = t.l_max, lmin = t.l_min;
  hipStream_t st = h->st;
  LevelDev &LL = h->lev[lmax];
  launch_batch_interleave(N, R, h->d_perm.p, xs, BB.XI.p, BB.XP.p, st);
  if (!far_on) {
    if constexpr (std::is_same<T, zf>::value) {
      if (f8_scale && R > 8) {
        launch_batch_near_f8(h->spmv_wgs, N, R, h->d_row_box.p, h->d_row_r0.p, h->mx_groups,
                             h->d_mx_box.p, h->d_mx_r0.p, h->d_box_first.p,
                             h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p,
                             h->d_perm.p, blocks_bf, f8_scale, (const zf *)BB.XI.p, BB.XB.p, ys,
                             h->mx_tpw, st);
        HIPCHK(hipGetLastError());
        return SPK_OK;
      }
    }
    launch_batch_near(h->spmv_wgs, N, R, h->d_row_box.p, h->d_row_r0.p, h->d_box_first.p,
                      h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p, h->d_perm.p,
                      pBlk(h, tag), (const T *)BB.XI.p, (const T *)nullptr, ys, st,
                      f8_scale ? nullptr : blocks_bf);
    HIPCHK(hipGetLastError());
    return SPK_OK;
  }
  launch_batch_agg(LL.nb, S, N, R, h->d_box_first.p, h->d_box_count.p, h->d_perm.p, pRt(h, tag),
                   (const T *)BB.XI.p, (const T *)BB.XP.p, BB.B[lmax].p, st);
  for (int l = lmax - 1; l >= lmin; --l)
    launch_batch_ss(h->lev[l].nb, S, R, h->lev[l].child.p, (const T *)BB.B[l + 1].p,
                    pUp(h->lev[l + 1], tag), BB.B[l].p, st);
  for (int l = lmin; l <= lmax; ++l)
    launch_batch_m2l(h->lev[l].nb, S, R, h->lev[l].i4_off.p, h->lev[l].i4_box.p,
                     h->lev[l].i4_slot.p, (const T *)BB.B[l].p, pT(h->lev[l], tag), BB.D[l].p, st);
  for (int l = lmin + 1; l <= lmax; ++l)
    launch_batch_rr(h->lev[l].nb, S, R, h->lev[l].parent.p, h->lev[l].octant.p,
                    (const T *)BB.D[l - 1].p, pDn(h->lev[l], tag), BB.D[l].p, st);
  const zd I = {0, 1};
  const zd kd1 = I * zcx(h->mat.k1), kd2 = I * zcx(h->mat.k2);
  const zd my1 = zd{h->mat.my1, 0} * I * h->mat.omega, my2 = zd{h->mat.my2, 0} * I * h->mat.omega;
  const zd ep1 = zcx(h->mat.eps1) * I * h->mat.omega, ep2 = zcx(h->mat.eps2) * I * h->mat.omega;
  launch_batch_fold(LL.nb, S, R, (const T *)BB.D[lmax].p, kd1, kd2, my1, my2, ep1, ep2, BB.W.p, st);
  launch_batch_disagg(h->ntiles, N, S, R, h->d_tile_box.p, h->d_tile_p0.p, h->d_box_first.p,
                      h->d_box_count.p, h->d_perm.p, pRt(h, tag), (const T *)BB.W.p, h->sc1,
                      h->sc2, BB.YP.p, BB.YF.p, h->dx_tiles, h->d_dx_box.p, h->d_dx_p0.p, st);
  launch_batch_near(h->spmv_wgs, N, R, h->d_row_box.p, h->d_row_r0.p, h->d_box_first.p,
                    h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p, h->d_perm.p,
                    pBlk(h, tag), (const T *)BB.XI.p, (const T *)BB.YF.p, ys, st);
  HIPCHK(hipGetLastError());
  return SPK_OK;
}

int matvec_dev(spk_handle *h, const zd *x, zd *y, bool near_on = true, bool far_on = true,
               hipEvent_t *ev = nullptr) {
  return matvec_t<zd>(h, x, y, near_on, far_on, ev);
}

// complex64 mirror of the operator: the near blocks, patterns, translation
// and phase tables rounded once from their complex128 values.
int build_mixed_mirror(spk_handle *h) {
  const size_t S = h->tb.S;
  const Tree &t = h->tree;
  HIPCHK(h->f_blocks.alloc(h->d_blocks.n));
  launch_to_zf((int64_t)h->d_blocks.n, h->d_blocks.p, h->f_blocks.p, h->st);
  HIPCHK(h->f_rt.alloc(h->d_rt.n));
  launch_to_zf((int64_t)h->d_rt.n, h->d_rt.p, h->f_rt.p, h->st);
  HIPCHK(h->f_W.alloc(h->d_W.n));
  HIPCHK(h->f_yfar.alloc(2 * (size_t)h->N));
  HIPCHK(h->f_x.alloc(2 * (size_t)h->N));
  HIPCHK(h->f_y.alloc(2 * (size_t)h->N));
  for (int l = t.l_min; l <= t.l_max; ++l) {
    LevelDev &D = h->lev[l];
    HIPCHK(D.fB.alloc(D.B.n));
    HIPCHK(D.fD.alloc(D.D.n));
    HIPCHK(D.ftlt.alloc(D.tlt.n));
    launch_to_zf((int64_t)D.tlt.n, D.tlt.p, D.ftlt.p, h->st);
    HIPCHK(D.fup.alloc(D.up.n));
    launch_to_zf((int64_t)D.up.n, D.up.p, D.fup.p, h->st);
    HIPCHK(D.fdown.alloc(D.down.n));
    launch_to_zf((int64_t)D.down.n, D.down.p, D.fdown.p, h->st);
  }
  (void)S;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->st));
  h->mixed_ready = true;
  return SPK_OK;
}

// algorithmic HBM bytes of each matvec stage (FP64, 16-byte complex)
void stage_bytes(const spk_handle *h, double *b, double c = 16.0) {
  const Tree &t = h->tree;
  const double S = h->tb.S, N = h->N;
  const double nbl = (double)t.lev[t.l_max].n.size();
  b[0] = N * 4 * S * c + 2 * N * c + nbl * 8 * S * c;  // patterns + x + C
  b[1] = 0;
  b[2] = 0;
  b[3] = 0;
  for (int l = t.l_min; l <= t.l_max; ++l) {
    const double nb = (double)t.lev[l].n.size();
    if (l < t.l_max) b[1] += (double)t.lev[l + 1].n.size() * 8 * S * c + nb * 8 * S * c;
    b[2] += 2.0 * nb * 8 * S * c + 343.0 * 2 * S * c;  // read B, write D~, TL table
    if (l > t.l_min) b[3] += (double)t.lev[l - 1].n.size() * 8 * S * c + 2.0 * nb * 8 * S * c;
  }
  b[4] = 2.0 * nbl * 8 * S * c;
  b[5] = N * 4 * S * c + nbl * 8 * S * c + 2 * N * c;  // patterns + W + y
  b[6] = (double)t.nnz * 4 * c + 2 * N * c + 2 * N * c + 2 * N * c;  // blocks + x + yfar + y
  b[7] = b[0] + b[1] + b[2] + b[3] + b[4] + b[5] + b[6];
}

}  // namespace

extern "C" {

const char *spk_last_error(void) { return spk::g_err.c_str(); }

static int scratch_init(int device);

int spk_create(const spk_params *params, spk_handle **out) {
  if (!params || !out) return fail(SPK_ERR_ARG, "null argument");
  if (params->lambda <= 0) return fail(SPK_ERR_ARG, "lambda must be positive");
  if (params->formulation < 1 || params->formulation > 8)
    return fail(SPK_ERR_ARG, "unknown formulation");
  const int ng = params->quad_points ? params->quad_points : 3;
  if (ng != 1 && ng != 3 && ng != 4 && ng != 6 && ng != 7 && ng != 12)
    return fail(SPK_ERR_ARG, "quad_points must be 1, 3, 4, 6, 7 or 12");
  if (params->precision != SPK_FP64 && params->precision != SPK_MIXED)
    return fail(SPK_ERR_ARG, "unknown precision");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (params->device < 0 || params->device >= ndev)
    return fail(SPK_ERR_ARG, "device ordinal out of range");
  HIPCHK(hipSetDevice(params->device));
  if (int rc = scratch_init(params->device)) return rc;
  std::unique_ptr<spk_handle> h(new spk_handle());
  h->p = *params;
  if (h->p.tree_s_opt <= 0) h->p.tree_s_opt = 250;
  if (h->p.truncation_number <= 0) h->p.truncation_number = 19;
  if (h->p.normal_convention != 2) h->p.normal_convention = 1;
  h->ngp = ng;
  make_material(h->p, &h->mat);
  quadrature_tri(ng, h->qa, h->qb, h->qw);
  quadrature_tri(12, h->qa12, h->qb12, h->qw12);
  HIPCHK(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
  *out = h.release();
  return SPK_OK;
}

void spk_destroy(spk_handle *h) { delete h; }

// the map of groups of tpw 16-row tiles of one leaf box (the batched fp8
// near product's workgroups)
static int build_mx_groups(spk_handle *h, int tpw) {
  const Tree &t = h->tree;
  const int grp_rows = 16 * tpw;
  std::vector<int> gb, g0;
  for (size_t b = 0; b < t.box_count.size(); ++b)
    for (int r = 0; r < t.box_count[b]; r += grp_rows) {
      gb.push_back((int)b);
      g0.push_back(r);
    }
  h->mx_groups = (int)gb.size();
  h->mx_tpw = tpw;
  HIPCHK(h->d_mx_box.upload(gb));
  HIPCHK(h->d_mx_r0.upload(g0));
  return SPK_OK;
}

// The operator of the current topology (h->fn: each unknown's node pair,
// whose midpoint places it in the octree): the octree and its lists, the
// truncation numbers, launch lists, the near field and the MLFMM tables.
static int setup_operator(spk_handle *h, double t0) {
  h->blu.ok = false;
  // the octree and its lists: on the device (the host builder with
  // SPK_HOST_TREE=1; both give the same lists)
  const char *ht = std::getenv("SPK_HOST_TREE");
  const bool host_tree = ht && ht[0] == '1';
  HIPCHK(h->d_fn.upload(h->fn));
  if (host_tree) {
    build_tree(h->fn, h->N, h->p.tree_s_opt, &h->tree);
  } else {
    const TreeDevOut out{&h->d_perm,      &h->d_pos,       &h->d_leaf_box,
                         &h->d_leafbox_of_pos, &h->d_box_first, &h->d_box_count,
                         &h->d_e2_off,    &h->d_e2_list,   &h->d_near_off};
    const char *what = "";
    const hipError_t e =
        build_tree_device(h->d_fn.p, h->N, h->p.tree_s_opt, &h->tree, h->tw, out, h->st, &what);
    if (e != hipSuccess)
      return fail(SPK_ERR_HIP, std::string("device octree (") + what + "): " + hipGetErrorString(e));
  }
  {
    // truncation number of every level (the reference sets all of them to
    // truncation_number_default, lib_sie_mlfmm.f90:696-698)
    const Tree &tr = h->tree;
    h->levK.assign(tr.l_max + 1, h->p.truncation_number);
    for (int l = 0; l <= tr.l_max; ++l)
      if (l < (int)h->levK_req.size() && h->levK_req[l] > 0) h->levK[l] = h->levK_req[l];
    h->multik = false;
    for (int l = tr.l_min; l < tr.l_max; ++l) {
      if (h->levK[l] < h->levK[l + 1])
        return fail(SPK_ERR_ARG, "a level's truncation number exceeds its parent's (level " +
                                     std::to_string(l + 1) + ")");
      if (h->levK[l] != h->levK[l + 1]) h->multik = true;
    }
    if (h->multik && h->p.precision == SPK_MIXED)
      return fail(SPK_ERR_ARG, "per-level truncation numbers run on FP64 handles");
    if (h->multik && h->quad)
      return fail(SPK_ERR_ARG, "per-level truncation numbers: triangle meshes only");
    make_tables(h->levK[tr.l_max], &h->tb);
    h->tbl.assign(tr.l_max + 1, Tables{});
    for (int l = tr.l_min; l <= tr.l_max; ++l) {
      if (h->levK[l] == h->levK[tr.l_max])
        h->tbl[l] = h->tb;
      else
        make_tables(h->levK[l], &h->tbl[l]);
    }
  }
  const Tree &t = h->tree;
  h->stats.tree_ms = now_ms() - t0;
  if (host_tree) {
    HIPCHK(h->d_perm.upload(t.perm));
    HIPCHK(h->d_box_first.upload(t.box_first));
    HIPCHK(h->d_box_count.upload(t.box_count));
    HIPCHK(h->d_e2_off.upload(t.e2_off));
    HIPCHK(h->d_e2_list.upload(t.e2_list));
    HIPCHK(h->d_near_off.upload(t.near_off));
    HIPCHK(h->d_leaf_box.upload(t.leaf_box));
    HIPCHK(h->d_pos.upload(t.pos));
    std::vector<int> lbp(h->N);
    for (int p = 0; p < h->N; ++p) lbp[p] = t.leaf_box[t.perm[p]];
    HIPCHK(h->d_leafbox_of_pos.upload(lbp));
  }
  h->transposed = false;
  std::vector<int> rb, r0;
  for (size_t b = 0; b < t.box_count.size(); ++b)
    for (int r = 0; r < t.box_count[b]; r += kSpmvRowsPerWG) {
      rb.push_back((int)b);
      r0.push_back(r);
    }
  h->spmv_wgs = (int)rb.size();
  HIPCHK(h->d_row_box.upload(rb));
  HIPCHK(h->d_row_r0.upload(r0));
  if (int e = build_mx_groups(h, h->tune_tpw > 0 ? h->tune_tpw : near_mx_tiles_per_group()))
    return e;
  // disaggregation tasks: up to kDisaggEdges consecutive positions of one box
  if (h->N >= (1 << 28)) return fail(SPK_ERR_ARG, "more than 2^28 unknowns per current");
  std::vector<int> dt;
  for (size_t b = 0; b < t.box_count.size(); ++b)
    for (int r = 0; r < t.box_count[b]; r += kDisaggEdges)
      dt.push_back((t.box_first[b] + r) * 8 + std::min(kDisaggEdges, t.box_count[b] - r));
  h->disagg_tasks = (int)dt.size();
  HIPCHK(h->d_disagg_task.upload(dt));
  {
    std::vector<int> tb, tp;
    for (size_t b = 0; b < t.box_count.size(); ++b)
      for (int r = 0; r < t.box_count[b]; r += kBatchTile) {
        tb.push_back((int)b);
        tp.push_back(t.box_first[b] + r);
      }
    h->ntiles = (int)tb.size();
    HIPCHK(h->d_tile_box.upload(tb));
    HIPCHK(h->d_tile_p0.upload(tp));
    std::vector<int> xb, xp;
    for (size_t b = 0; b < t.box_count.size(); ++b)
      for (int r = 0; r < t.box_count[b]; r += 128) {
        xb.push_back((int)b);
        xp.push_back(t.box_first[b] + r);
      }
    h->dx_tiles = (int)xb.size();
    HIPCHK(h->d_dx_box.upload(xb));
    HIPCHK(h->d_dx_p0.upload(xp));
    // the batch buffers are resized for this mesh on their next use
    h->bd.R = 0;
    h->bf.R = 0;
  }
  HIPCHK(h->d_x.alloc(2 * (size_t)h->N));
  HIPCHK(h->d_y.alloc(2 * (size_t)h->N));
  HIPCHK(h->d_yfar.alloc(2 * (size_t)h->N));
  HIPCHK(h->d_tmp.alloc(2 * (size_t)h->N));
  // the host part of the level tables runs beside the device's near fill
  std::future<std::vector<HostLevelTables>> pre =
      std::async(std::launch::async, host_level_tables, std::cref(h->tbl), std::cref(h->mat),
                 std::cref(h->tree));
  double t1 = now_ms();
  int rc = h->quad ? near_fill_quad(h) : near_fill(h);
  if (rc != SPK_OK) return rc;
  h->stats.near_fill_ms = now_ms() - t1;
  rc = setup_mlfmm(h, pre);
  if (rc != SPK_OK) return rc;
  h->stats.n_edges = h->N;
  h->stats.near_blocks = t.nnz;
  h->stats.l_min = t.l_min;
  h->stats.l_max = t.l_max;
  h->stats.samples = h->tb.S;
  h->stats.bytes_near = (double)t.nnz * 64.0;
  h->stats.bytes_patterns = (double)h->N * (h->quad ? kQuadRtRows : 4) * h->tb.S * 16.0;
  if (h->p.precision == SPK_MIXED) {
    rc = build_mixed_mirror(h);
    if (rc != SPK_OK) return rc;
  }
  h->stats.setup_ms = now_ms() - t0;
  h->ready = true;
  return SPK_OK;
}


int spk_setup_tri(spk_handle *h, int np, const double *pp, int nt, const int *tt) {
  if (!h || !pp || !tt || np < 3 || nt < 1) return fail(SPK_ERR_ARG, "bad mesh arguments");
  HIPCHK(hipSetDevice(h->p.device));
  h->ready = false;
  h->mixed_ready = false;
  h->diag_ready = false;
  ++h->setup_serial;
  h->ig.reset();
  // the low-precision near-field copies belong to the previous operator
  h->near_bf_ok = h->near_f8_ok = false;
  h->stats = spk_stats{};
  h->quad = false;
  const double t0 = now_ms();
  std::string err;
  int rc = build_rwg(np, pp, nt, tt, h->p.normal_convention, &h->corner, &h->element, &h->fn,
                     &h->er, &h->rref, &err);
  if (rc != SPK_OK) return fail(rc, err);
  h->N = (int)h->rref.size();
  if (h->N < 1) return fail(SPK_ERR_MESH, "mesh has no interior edges (no RWG pairs)");
  HIPCHK(h->d_er.upload(h->er));
  HIPCHK(h->d_rref.upload(h->rref));
  h->pp_host.assign(pp, pp + 3 * (size_t)np);
  return setup_operator(h, t0);
}

int spk_setup_quad(spk_handle *h, double Lx, double Ly, int nx, int ny, const double *z) {
  if (!h || !z || nx < 1 || ny < 1 || !(Lx > 0) || !(Ly > 0))
    return fail(SPK_ERR_ARG, "bad quadrilateral mesh arguments");
  const int f = h->p.formulation;
  if (f != SPK_PMCHWT && f != SPK_MCTF && f != SPK_ICTF)
    return fail(SPK_ERR_ARG, "quadrilateral meshes support PMCHWT, MCTF and ICTF");
  if (h->p.eps_r1_re == 0 && h->p.eps_r1_im == 0)
    return fail(SPK_ERR_ARG, "eps_r1 must be given (1 for vacuum)");
  HIPCHK(hipSetDevice(h->p.device));
  h->ready = false;
  h->mixed_ready = false;
  h->diag_ready = false;
  ++h->setup_serial;
  h->ig.reset();
  h->near_bf_ok = h->near_f8_ok = false;
  h->stats = spk_stats{};
  h->quad = true;
  const double t0 = now_ms();
  // sorting_quad, Neighbours_EdgeCorner_Full, Sorting_edge
  QuadMesh &m = h->qmesh;
  m = QuadMesh();
  quad_mesh(Lx, Ly, nx, ny, z, &m);
  quad_topology(&m);
  if (m.medge < 1) return fail(SPK_ERR_MESH, "mesh has no edge functions");
  std::vector<double> el;
  std::vector<int> eadj, eside, cadj;
  if (!quad_topology_arrays(m, &el, &eadj, &eside, &cadj))
    return fail(SPK_ERR_MESH, "element with more than 8 neighbours");
  HIPCHK(h->q_el.upload(el));
  HIPCHK(h->q_edge_p.upload(m.edge_p));
  HIPCHK(h->q_eadj.upload(eadj));
  HIPCHK(h->q_eside.upload(eside));
  HIPCHK(h->q_cadj.upload(cadj));
  quad_constants(h->p, &h->qdev, &h->q_eta_a, &h->q_eta12);
  HIPCHK(h->q_seh.alloc(2 * (size_t)m.nel * 10));
  // each edge function's node pair in the records the octree reads
  h->N = m.medge;
  h->fn.assign(2 * (size_t)h->N, FnRec{});
  for (int s = 0; s < h->N; ++s) {
    int n1, n2;
    quad_edge_nodes(m, s, &n1, &n2);
    FnRec &r = h->fn[2 * (size_t)s];
    for (int d = 0; d < 3; ++d) {
      r.c[d] = m.pp[3 * (size_t)(n1 - 1) + d];
      r.c[3 + d] = m.pp[3 * (size_t)(n2 - 1) + d];
    }
    h->fn[2 * (size_t)s + 1] = r;
  }
  h->corner.clear();
  h->element.clear();
  h->er.clear();
  h->rref.clear();
  h->pp_host.clear();
  int rc = setup_operator(h, t0);
  if (rc != SPK_OK) return rc;
  // The far part stands for the regular element-pair rule only where the
  // supports of functions in non-adjacent leaf boxes share no node: every
  // support point lies within one element diameter of its function's
  // midpoint, and such midpoints are more than a leaf edge apart.
  double diam = 0;
  for (int e = 0; e < m.nel; ++e)
    for (int a = 0; a < 8; ++a)
      for (int b = a + 1; b < 8; ++b) {
        const double *pa = &el[(size_t)e * 24 + 3 * a], *pb = &el[(size_t)e * 24 + 3 * b];
        const double d2 = (pa[0] - pb[0]) * (pa[0] - pb[0]) + (pa[1] - pb[1]) * (pa[1] - pb[1]) +
                          (pa[2] - pb[2]) * (pa[2] - pb[2]);
        diam = std::max(diam, std::sqrt(d2));
      }
  const double leaf = std::ldexp(h->tree.D, -h->tree.l_max);
  if (h->tree.l_max >= 2 && leaf <= 2 * diam) {
    h->ready = false;
    return fail(SPK_ERR_ARG, "leaf boxes (" + std::to_string(leaf) +
                                 " m) must exceed twice the element diameter (" +
                                 std::to_string(diam) + " m): raise tree_s_opt");
  }
  return SPK_OK;
}

int spk_device_alloc(int device, size_t bytes, void **d_ptr) {
  if (!d_ptr) return fail(SPK_ERR_ARG, "null argument");
  *d_ptr = nullptr;
  HIPCHK(hipSetDevice(device));
  if (bytes == 0) return SPK_OK;
  HIPCHK(hipMalloc(d_ptr, bytes));
  return SPK_OK;
}

int spk_device_free(void *d_ptr) {
  if (d_ptr) HIPCHK(hipFree(d_ptr));
  return SPK_OK;
}

int spk_device_copy(void *dst, const void *src, size_t bytes, int direction) {
  if ((!dst || !src) && bytes) return fail(SPK_ERR_ARG, "null argument");
  if (bytes == 0) return SPK_OK;
  hipMemcpyKind k = direction == SPK_COPY_H2D   ? hipMemcpyHostToDevice
                    : direction == SPK_COPY_D2H ? hipMemcpyDeviceToHost
                    : direction == SPK_COPY_D2D ? hipMemcpyDeviceToDevice
                                                : hipMemcpyDefault;
  if (direction < SPK_COPY_H2D || direction > SPK_COPY_D2D)
    return fail(SPK_ERR_ARG, "unknown copy direction");
  // the handles' streams do not synchronise with the null stream: finish
  // their work first, so that a copy after an asynchronous product (e.g.
  // spk_matvec_device) sees its result
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(dst, src, bytes, k));
  return SPK_OK;
}

int spk_device_synchronize(int device) {
  HIPCHK(hipSetDevice(device));
  HIPCHK(hipDeviceSynchronize());
  return SPK_OK;
}

int spk_get_stats(const spk_handle *h, spk_stats *out) {
  if (!h || !out) return fail(SPK_ERR_ARG, "null argument");
  *out = h->stats;
  return SPK_OK;
}

int spk_num_edges(const spk_handle *h, int64_t *n) {
  if (!h || !n) return fail(SPK_ERR_ARG, "null argument");
  *n = h->N;
  return SPK_OK;
}

// Switch the handle between Z and its transpose: the near blocks are
// transposed in HBM (through a temporary copy), the far field runs its
// transposed stages.  The low-precision near-field copies of the
// preconditioner are rebuilt from the transposed blocks on their next use.
int spk_set_transposed(spk_handle *h, int transposed) {
  if (!h) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  const bool want = transposed != 0;
  if (want == h->transposed) return SPK_OK;
  if (want && h->quad) return fail(SPK_ERR_STATE, "quadrilateral handles: no transposed operator");
  // the block-Jacobi inverses belong to the untransposed blocks
  h->blu.ok = false;
  if (want && h->multik)
    return fail(SPK_ERR_STATE, "per-level truncation numbers: no transposed operator");
  HIPCHK(hipSetDevice(h->p.device));
  const Tree &t = h->tree;
  const int nbl = (int)t.box_count.size();
  int max_rows = 0;
  for (int c : t.box_count) max_rows = std::max(max_rows, c);
  {
    DBuf<zd> tmp;
    HIPCHK(tmp.alloc(h->d_blocks.n));
    launch_near_transpose(nbl, max_rows, h->d_box_first.p, h->d_box_count.p, h->d_e2_off.p,
                          h->d_e2_list.p, h->d_near_off.p, h->d_perm.p, h->d_pos.p,
                          h->d_leaf_box.p, (const zd *)h->d_blocks.p, tmp.p, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    std::swap(h->d_blocks, tmp);
  }
  if (h->mixed_ready && h->f_blocks.p) {
    DBuf<zf> tmp;
    HIPCHK(tmp.alloc(h->f_blocks.n));
    launch_near_transpose(nbl, max_rows, h->d_box_first.p, h->d_box_count.p, h->d_e2_off.p,
                          h->d_e2_list.p, h->d_near_off.p, h->d_perm.p, h->d_pos.p,
                          h->d_leaf_box.p, (const zf *)h->f_blocks.p, tmp.p, h->st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(h->st));
    std::swap(h->f_blocks, tmp);
  }
  h->near_bf_ok = h->near_f8_ok = false;
  h->ig.reset();
  h->transposed = want;
  return SPK_OK;
}

int spk_set_inner_graph(spk_handle *h, int enable) {
  if (!h) return fail(SPK_ERR_ARG, "null argument");
  h->graph_ok = enable != 0;
  if (!h->graph_ok) h->ig.reset();
  return SPK_OK;
}

int spk_get_transposed(const spk_handle *h, int *transposed) {
  if (!h || !transposed) return fail(SPK_ERR_ARG, "null argument");
  *transposed = h->transposed ? 1 : 0;
  return SPK_OK;
}

int spk_matvec_device(spk_handle *h, const void *d_x, void *d_y) {
  if (!h || !d_x || !d_y) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  return matvec_dev(h, (const zd *)d_x, (zd *)d_y);
}

int spk_matvec_c64_device(spk_handle *h, const void *d_x, void *d_y) {
  if (!h || !d_x || !d_y) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (!h->mixed_ready) return fail(SPK_ERR_STATE, "handle was not created with SPK_MIXED");
  int rc = matvec_t<zf>(h, (const zf *)d_x, (zf *)d_y);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  return SPK_OK;
}

}  // extern "C"

// the batched C-ABI entry points: host arrays of R device pointers
template <class T>
static int matvec_batch_api(spk_handle *h, int R, const void *const *d_x, void *const *d_y) {
  if (!h || !d_x || !d_y) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (R < 1 || R > kBatchMax) return fail(SPK_ERR_ARG, "batch size must be 1..16");
  for (int r = 0; r < R; ++r)
    if (!d_x[r] || !d_y[r]) return fail(SPK_ERR_ARG, "null vector in the batch");
  if (h->quad) {
    // quadrilateral handles: one product per vector
    for (int r = 0; r < R; ++r)
      if (int e = matvec_t<T>(h, (const T *)d_x[r], (T *)d_y[r])) return e;
    HIPCHK(hipStreamSynchronize(h->st));
    return SPK_OK;
  }
  int rc = batch_ensure<T>(h, R);
  if (rc) return rc;
  T *tag = nullptr;
  BatchBufs<T> &BB = pBB(h, tag);
  HIPCHK(hipMemcpy(BB.xs.p, d_x, R * sizeof(void *), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(BB.ys.p, d_y, R * sizeof(void *), hipMemcpyHostToDevice));
  rc = matvec_batch_t<T>(h, R, BB.xs.p, BB.ys.p);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  return SPK_OK;
}

extern "C" {

int spk_matvec_batch_device(spk_handle *h, int R, const void *const *d_x, void *const *d_y) {
  return matvec_batch_api<zd>(h, R, d_x, d_y);
}

int spk_matvec_batch_c64_device(spk_handle *h, int R, const void *const *d_x, void *const *d_y) {
  if (h && !h->mixed_ready) return fail(SPK_ERR_STATE, "handle was not created with SPK_MIXED");
  return matvec_batch_api<zf>(h, R, d_x, d_y);
}

int spk_matvec_c64(spk_handle *h, const float *x, float *y) {
  if (!h || !x || !y) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (!h->mixed_ready) return fail(SPK_ERR_STATE, "handle was not created with SPK_MIXED");
  const size_t bytes = 2 * (size_t)h->N * sizeof(zf);
  HIPCHK(hipMemcpyAsync(h->f_x.p, x, bytes, hipMemcpyHostToDevice, h->st));
  int rc = matvec_t<zf>(h, h->f_x.p, h->f_y.p);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(y, h->f_y.p, bytes, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return SPK_OK;
}

int spk_matvec(spk_handle *h, const double *x, double *y) {
  if (!h || !x || !y) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  const size_t bytes = 2 * (size_t)h->N * sizeof(zd);
  HIPCHK(hipMemcpyAsync(h->d_x.p, x, bytes, hipMemcpyHostToDevice, h->st));
  int rc = matvec_dev(h, h->d_x.p, h->d_y.p);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(y, h->d_y.p, bytes, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return SPK_OK;
}

int spk_matvec_parts(spk_handle *h, const double *x, double *y_near, double *y_far) {
  if (!h || !x) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  const size_t bytes = 2 * (size_t)h->N * sizeof(zd);
  HIPCHK(hipMemcpyAsync(h->d_x.p, x, bytes, hipMemcpyHostToDevice, h->st));
  if (y_near) {
    int rc = matvec_dev(h, h->d_x.p, h->d_y.p, true, false);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(y_near, h->d_y.p, bytes, hipMemcpyDeviceToHost, h->st));
  }
  if (y_far) {
    int rc = matvec_dev(h, h->d_x.p, h->d_y.p, false, true);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(y_far, h->d_y.p, bytes, hipMemcpyDeviceToHost, h->st));
  }
  HIPCHK(hipStreamSynchronize(h->st));
  return SPK_OK;
}

// get_vector_b_tri / incident_field_tri_general, plane wave
static RhsArgs rhs_args(const spk_handle *h, const double k_in[3], const double e_in[3]) {
  RhsArgs a{};
  a.N = h->N;
  a.ng = h->ngp;
  for (int i = 0; i < 20; ++i) {
    a.qa[i] = h->qa[i];
    a.qb[i] = h->qb[i];
    a.qw[i] = h->qw[i];
  }
  // the phase is dot_product(k1*k_in, R_a) (:977), and Fortran's
  // dot_product conjugates a complex first argument: exp(-i conj(k1) k_in.R)
  a.k1 = zd{h->mat.k1.re, -h->mat.k1.im};
  for (int d = 0; d < 3; ++d) {
    a.kin[d] = k_in[d];
    a.ein[d] = e_in[d];
  }
  a.c0my = h->mat.my1 * kC0;
  const zc e1(h->mat.eta1.re, h->mat.eta1.im), e2(h->mat.eta2.re, h->mat.eta2.im);
  const zc ea(h->mat.eta_a.re, h->mat.eta_a.im);
  const zc a0(h->mat.a_lc[0].re, h->mat.a_lc[0].im), d0(h->mat.d_lc[0].re, h->mat.d_lc[0].im);
  const zc ep1(h->mat.eps1.re, h->mat.eps1.im), ep2(h->mat.eps2.re, h->mat.eps2.im);
  // the formulation's combination of (<f,E>, <f,H>, <f,nxE>, <f,nxH>)
  // (the select case of get_vector_b_tri, src/lib_sie_tri_calculation_mod.f90:391-460)
  zc ca[4] = {0, 0, 0, 0}, cb[4] = {0, 0, 0, 0};
  switch (h->p.formulation) {
    case SPK_PMCHWT: ca[0] = 1.0; cb[1] = 1.0; break;
    case SPK_PMCHWT_S: ca[0] = 1.0; cb[1] = e1; break;
    case SPK_MCTF: ca[0] = 1.0; cb[1] = e1 * e2; break;
    case SPK_JMCFIE: ca[0] = 1.0 / e1; ca[3] = 1.0; cb[2] = -1.0; cb[1] = e1; break;
    case SPK_MCTF2: ca[0] = a0 / e1; cb[1] = d0 * e1; break;
    case SPK_ICTF: ca[0] = 1.0 / ea; cb[1] = ea; break;
    case SPK_CNF: ca[3] = 1.0; cb[2] = -1.0; break;
    default:
      ca[3] = h->mat.my1 / (h->mat.my1 + h->mat.my2);
      cb[2] = -ep1 / (ep1 + ep2);
      break;
  }
  for (int q = 0; q < 4; ++q) {
    a.ca[q] = zd{ca[q].real(), ca[q].imag()};
    a.cb[q] = zd{cb[q].real(), cb[q].imag()};
  }
  return a;
}

static int rhs_plane_dev(spk_handle *h, const double k_in[3], const double e_in[3], zd *d_b) {
  if (h->quad) {
    // get_vector_b_quad (src/lib_sie_quad_calculation_mod.f90:91-138)
    const QuadRhsArgs a = quad_rhs_args(h->p.formulation, h->qdev, h->q_eta_a, h->q_eta12, k_in, e_in);
    launch_quad_rhs_plane(h->qdev, h->qtopo(), a, d_b, h->st);
    HIPCHK(hipGetLastError());
    return SPK_OK;
  }
  const RhsArgs a = rhs_args(h, k_in, e_in);
  launch_rhs_plane(a, h->d_fn.p, d_b, h->st);
  HIPCHK(hipGetLastError());
  return SPK_OK;
}

int spk_rhs_plane_wave_device(spk_handle *h, const double k_in[3], const double e_in[3],
                              void *d_b) {
  if (!h || !k_in || !e_in || !d_b) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready || h->N <= 0 || !h->d_fn.p) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  int rc = rhs_plane_dev(h, k_in, e_in, (zd *)d_b);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(h->st));
  return SPK_OK;
}

int spk_rhs_plane_wave(spk_handle *h, const double k_in[3], const double e_in[3], double *bout) {
  if (!h || !k_in || !e_in || !bout) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready || h->N <= 0 || !h->d_fn.p) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  DBuf<zd> db;
  HIPCHK(db.alloc(2 * (size_t)h->N));
  int rc = rhs_plane_dev(h, k_in, e_in, db.p);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(bout, db.p, 2 * (size_t)h->N * sizeof(zd), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return SPK_OK;
}

static ScatArgs scat_args(const spk_handle *h, int npts) {
  ScatArgs a{};
  a.N = h->N;
  a.npts = npts;
  double qa[20], qb[20], qw[20];
  quadrature_tri(3, qa, qb, qw);
  for (int i = 0; i < 3; ++i) {
    a.qa[i] = qa[i];
    a.qb[i] = qb[i];
    a.qw[i] = qw[i];
  }
  const zc iw(0.0, h->mat.omega);
  const zc eps1(h->mat.eps1.re, h->mat.eps1.im);
  const zc iwmy = iw * h->mat.my1, iweps = iw * eps1, inv_iwmy = 1.0 / (iw * h->mat.my1);
  a.k1 = zd{h->mat.k1.re, h->mat.k1.im};
  a.iwmy = zd{iwmy.real(), iwmy.imag()};
  a.iweps = zd{iweps.real(), iweps.imag()};
  a.inv_iwmy = zd{inv_iwmy.real(), inv_iwmy.imag()};
  return a;
}

// One field evaluation through device staging: the points dr[3*npts] are
// copied in, one workgroup per point writes every component of dE, dH
// [3*npts], which are copied out.
static int scattered_staged(spk_handle *h, int npts, const double *r, const zd *d_x, double *E,
                            double *H, double *dr, zd *dE, zd *dH) {
  const ScatArgs a = scat_args(h, npts);
  HIPCHK(hipMemcpyAsync(dr, r, 3 * (size_t)npts * sizeof(double), hipMemcpyHostToDevice, h->st));
  if (h->quad) {
    // element-local coefficients (every slot is some function's), then the field
    HIPCHK(hipMemsetAsync(h->q_seh.p, 0, h->q_seh.n * sizeof(zd), h->st));
    launch_quad_local(h->qtopo(), d_x, h->q_seh.p, h->st);
    launch_quad_field(h->qdev, h->qtopo(), npts, dr, h->q_seh.p, (double *)dE, (double *)dH, h->st);
  } else {
    launch_scattered(a, h->d_fn.p, d_x, dr, dE, dH, h->st);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(E, dE, 3 * (size_t)npts * sizeof(zd), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipMemcpyAsync(H, dH, 3 * (size_t)npts * sizeof(zd), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return SPK_OK;
}

static int scattered_dev(spk_handle *h, int npts, const double *r, const zd *d_x, double *E,
                         double *H) {
  DBuf<double> dr;
  DBuf<zd> dE, dH;
  HIPCHK(dr.alloc(3 * (size_t)npts));
  HIPCHK(dE.alloc(3 * (size_t)npts));
  HIPCHK(dH.alloc(3 * (size_t)npts));
  return scattered_staged(h, npts, r, d_x, E, H, dr.p, dE.p, dH.p);
}

// ---------------------------------------------------------------------------
// Staging for the scattered field without allocation: a file-scope device
// array, one copy per GPU, held until the process ends.  Host code never
// passes the array's name to a kernel or a copy: spk_create takes its
// device address once per device (hipGetSymbolAddress, with the device
// current) and the entry point uses that address.  Layout: points
// [3 * kScatScratchPoints] double, then E and H [3 * kScatScratchPoints]
// complex128 each (every offset a multiple of 256 bytes).
// ---------------------------------------------------------------------------
constexpr int kScatScratchPoints = 16384;
constexpr size_t kScatOffE = 3 * (size_t)kScatScratchPoints * sizeof(double);
constexpr size_t kScatOffH = kScatOffE + 3 * (size_t)kScatScratchPoints * sizeof(zd);
constexpr size_t kScatScratchBytes = kScatOffH + 3 * (size_t)kScatScratchPoints * sizeof(zd);
static_assert(kScatOffE % 256 == 0 && kScatOffH % 256 == 0, "scratch offsets");
static __device__ __align__(256) unsigned char scratch[kScatScratchBytes];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kScratchDevices];

// The device address of `scratch` on `device`, which must be current.
static int scratch_init(int device) {
  if (device < 0 || device >= kScratchDevices)
    return fail(SPK_ERR_ARG, "device ordinal beyond the staging table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return SPK_OK;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(SPK_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
  return SPK_OK;
}

int spk_scattered_field_device(spk_handle *h, int npts, const double *r, const void *d_x,
                               double *E, double *H) {
  if (!h || !r || !d_x || !E || !H || npts < 0) return fail(SPK_ERR_ARG, "bad argument");
  if (!h->ready || h->N <= 0 || !h->d_fn.p) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (npts == 0) return SPK_OK;
  return scattered_dev(h, npts, r, (const zd *)d_x, E, H);
}

int spk_scattered_field_device_noalloc(spk_handle *h, int npts, const double *r,
                                       const void *d_x, double *E, double *H) {
  if (!h || !r || !d_x || !E || !H || npts < 0) return fail(SPK_ERR_ARG, "bad argument");
  if (!h->ready || h->N <= 0 || !h->d_fn.p) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (npts > kScatScratchPoints)
    return fail(SPK_ERR_ARG, "npts exceeds the no-allocation maximum of " +
                                 std::to_string(kScatScratchPoints) + " points");
  if (npts == 0) return SPK_OK;
  HIPCHK(hipSetDevice(h->p.device));
  unsigned char *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = (unsigned char *)g_scratch_ptr[h->p.device];
  }
  if (!base) return fail(SPK_ERR_STATE, "no staging address for this device");
  return scattered_staged(h, npts, r, (const zd *)d_x, E, H, (double *)base,
                          (zd *)(base + kScatOffE), (zd *)(base + kScatOffH));
}

int spk_scattered_field(spk_handle *h, int npts, const double *r, const double *x, double *E,
                        double *H) {
  if (!h || !r || !x || !E || !H || npts < 0) return fail(SPK_ERR_ARG, "bad argument");
  if (!h->ready || h->N <= 0 || !h->d_fn.p) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (npts == 0) return SPK_OK;
  DBuf<zd> dx;
  HIPCHK(dx.alloc(2 * (size_t)h->N));
  HIPCHK(hipMemcpyAsync(dx.p, x, 2 * (size_t)h->N * sizeof(zd), hipMemcpyHostToDevice, h->st));
  return scattered_dev(h, npts, r, dx.p, E, H);
}

// ---------------------------------------------------------------------------
// CERFACS GMRES (zPackgmres.f90 :537-1565) on device vectors.
// ---------------------------------------------------------------------------
static void zrotg_(zc *ca, zc cb, double *cc, zc *s) {
  if (std::abs(*ca) == 0.0) {
    *cc = 0.0;
    *s = 1.0;
    *ca = cb;
    return;
  }
  const double scale = std::abs(*ca) + std::abs(cb);
  const double a1 = std::abs(*ca / zc(scale, 0.0)), b1 = std::abs(cb / zc(scale, 0.0));
  const double norm = scale * std::sqrt(a1 * a1 + b1 * b1);
  const zc alpha = *ca / std::abs(*ca);
  *cc = std::abs(*ca) / norm;
  *s = alpha * std::conj(cb) / norm;
  *ca = alpha * norm;
}
static void zrot1(zc *x, zc *y, double cc, zc s) {
  const zc t = cc * *x + s * *y;
  *y = cc * *y - std::conj(s) * *x;
  *x = t;
}

static int dev_dot(spk_handle *h, const zd *x, const zd *y, int64_t n, zc *out) {
  launch_zdotc(n, x, y, h->g_partial.p, h->g_scal.p, h->st);
  zd r;
  HIPCHK(hipMemcpyAsync(&r, h->g_scal.p, sizeof(zd), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  *out = zc(r.x, r.y);
  return SPK_OK;
}

// ---------------------------------------------------------------------------
// Mixed precision: iterative refinement around a complex64 inner GMRES.
//
// Outer loop (complex128): r = b - Z x with the FP64 operator; stop when
// |r| / |b| <= tolerance.  Inner solve: GMRES on the complex64 operator for
// Z d = r / |r|, Krylov basis stored in complex64, Gram-Schmidt by CGS with
// the CERFACS reorthogonalisation test (ICGS), dot products accumulated in
// complex128 and the Hessenberg matrix and Givens rotations kept in
// complex128 on the host; it stops at the relative residual
// max(refinement_tolerance (default 3e-4), tolerance / (2 |r|/|b|)): the
// last solves need only the reduction still missing.  x += |r| d.  No
// counterpart in the reference, which computes in FP64 only.
// ---------------------------------------------------------------------------
static int gmres_mixed(spk_handle *h, const zd *b, zd *x, const spk_gmres_params *gp,
                       spk_gmres_info *info) {
  if (!h->mixed_ready) return fail(SPK_ERR_STATE, "handle was not created with SPK_MIXED");
  if (gp->preconditioning != SPK_PRECOND_NONE)
    return fail(SPK_ERR_ARG, "preconditioning is not available with SPK_MIXED");
  const double t_start = now_ms();
  const int64_t n = 2 * (int64_t)h->N;
  const int max_it = std::max(1, gp->max_iterations);
  const int m = gp->restart > 0 ? std::min(gp->restart, max_it) : max_it;
  const double tol = gp->tolerance;
  const double inner_tol = gp->refinement_tolerance > 0 ? gp->refinement_tolerance : 3e-4;
  const int max_outer = gp->refinement_steps > 0 ? gp->refinement_steps : 30;
  if (h->i_m != m || h->i_V.n != (size_t)n * (m + 1)) {
    HIPCHK(h->i_V.alloc((size_t)n * (m + 1)));
    h->i_m = m;
  }
  if (h->i_w.n != (size_t)n) HIPCHK(h->i_w.alloc(n));
  if (h->g_w.n != (size_t)n) {
    HIPCHK(h->g_w.alloc(n));
    HIPCHK(h->g_r0.alloc(n));
    HIPCHK(h->g_xc.alloc(n));
  }
  HIPCHK(h->g_h.alloc(m + 1));
  HIPCHK(h->g_y.alloc(m + 1));
  HIPCHK(h->pin.ensure(m + 2));
  const size_t gc_elems = gemv_c_partial_elems(n, m + 1, 1);
  HIPCHK(h->g_partial.alloc(gc_elems + (size_t)reduce_blocks(n)));
  HIPCHK(h->g_pn.alloc(gemv_n_partial_elems(n, m + 1)));
  HIPCHK(h->g_scal.alloc(4));
  zf *V = h->i_V.p, *w = h->i_w.p;
  zd *r = h->g_r0.p, *tmp = h->g_w.p, *d = h->g_xc.p;
  zd *partial_nrm = h->g_partial.p + gc_elems;
  h->hist.clear();
  int rc = SPK_OK;
  auto dot_host = [&](const auto *u, const auto *v, double *out) -> int {
    launch_zdotc(n, u, v, h->g_partial.p, h->g_scal.p, h->st);
    zd t;
    HIPCHK(hipMemcpyAsync(&t, h->g_scal.p, sizeof(zd), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    *out = t.x;
    return SPK_OK;
  };
  if (!gp->use_initial_guess) HIPCHK(hipMemsetAsync(x, 0, n * sizeof(zd), h->st));
  double bb;
  if ((rc = dot_host(b, b, &bb))) return rc;
  const double bn = std::sqrt(bb);
  int matvecs = 0, total = 0, outer = 0;
  double be = 0.0, bea = 0.0;
  if (bn == 0.0) {
    HIPCHK(hipMemsetAsync(x, 0, n * sizeof(zd), h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (info) *info = spk_gmres_info{0, 0, 0.0, 0.0, 0, (now_ms() - t_start) / 1e3, 0.0};
    return SPK_OK;
  }
  auto residual = [&](double *rn) -> int {
    int e = matvec_dev(h, x, tmp);
    if (e) return e;
    ++matvecs;
    launch_zsub(n, b, tmp, r, h->st);
    double rr;
    if ((e = dot_host(r, r, &rr))) return e;
    *rn = std::sqrt(rr);
    return SPK_OK;
  };
  double rn;
  if ((rc = residual(&rn))) return rc;
  be = rn / bn;
  bea = be;
  std::vector<zc> H((size_t)(m + 1) * (m + 1)), g(m + 1), rs(m + 1), yc(m + 1);
  std::vector<double> rcs(m + 1);
  auto HH = [&](int i, int j) -> zc & { return H[(size_t)j * (m + 1) + i]; };
  while (be > tol && total < max_it && outer < max_outer) {
    // V_0 = (r / |r|) in complex64, renormalised after rounding
    launch_scale_down(n, 1.0 / rn, r, V, h->st);
    double vv;
    if ((rc = dot_host(V, V, &vv))) return rc;
    const double beta = std::sqrt(vv);
    launch_scal_copy_f(n, 1.0 / beta, V, V, h->st);
    std::fill(H.begin(), H.end(), zc(0.0));
    std::fill(g.begin(), g.end(), zc(0.0));
    g[0] = beta;
    const int jmax = std::min(m, max_it - total);
    int k = 0;
    double est = 1.0;
    // this inner solve only has to supply the reduction still missing, with
    // a factor 2 margin: stop at max(refinement tolerance, tol / (2 be))
    const double stop_at = std::max(inner_tol, std::min(0.5, 0.5 * tol / be));
    for (int j = 0; j < jmax; ++j) {
      if ((rc = matvec_t<zf>(h, V + (size_t)j * n, w))) return rc;
      ++matvecs;
      int nOrtho = 0;
      double dnormw = 0.0;
      for (;;) {
        ++nOrtho;
        launch_gemv_c(n, j + 1, (const zf *)V, n, (const zf *)w, (const zf *)nullptr,
                      h->g_partial.p, h->g_h.p, h->st);
        launch_gemv_n_sub(n, j + 1, (const zf *)V, n, h->g_h.p, w, h->g_pn.p, h->st);
        // |w|^2 lands after the coefficients: one copy to pinned memory per pass
        launch_zdotc(n, (const zf *)w, (const zf *)w, partial_nrm, h->g_h.p + (j + 1), h->st);
        const zd *hv = h->pin.p;
        HIPCHK(hipMemcpyAsync(h->pin.p, h->g_h.p, (j + 2) * sizeof(zd), hipMemcpyDeviceToHost,
                              h->st));
        HIPCHK(hipStreamSynchronize(h->st));
        double dloo = 0.0;
        for (int i = 0; i <= j; ++i) {
          const zc hi(hv[i].x, hv[i].y);
          HH(i, j) += hi;
          dloo += std::norm(hi);
        }
        dloo = std::sqrt(dloo);
        dnormw = std::sqrt(std::max(hv[j + 1].x, 0.0));
        if (dnormw <= dloo && nOrtho < 3) continue;
        break;
      }
      HH(j + 1, j) = dnormw;
      if (dnormw > 0.0) launch_scal_copy_f(n, 1.0 / dnormw, w, V + (size_t)(j + 1) * n, h->st);
      for (int i = 0; i < j; ++i) zrot1(&HH(i, j), &HH(i + 1, j), rcs[i], rs[i]);
      zc a = HH(j, j);
      double cc;
      zc ss;
      zrotg_(&a, HH(j + 1, j), &cc, &ss);
      rcs[j] = cc;
      rs[j] = ss;
      zrot1(&g[j], &g[j + 1], cc, ss);
      zrot1(&HH(j, j), &HH(j + 1, j), cc, ss);
      HH(j + 1, j) = 0.0;
      est = std::abs(g[j + 1]) / beta;
      ++total;
      k = j + 1;
      h->hist.push_back(est * be);
      if (est <= stop_at || dnormw == 0.0) break;
    }
    bea = est * be;
    // y = H^-1 g, d = V y, x += |r| d
    for (int j = k - 1; j >= 0; --j) {
      zc sum = g[j];
      for (int i = j + 1; i < k; ++i) sum -= HH(j, i) * yc[i];
      yc[j] = sum / HH(j, j);
    }
    std::vector<zd> yd(k);
    for (int j = 0; j < k; ++j) yd[j] = zd{yc[j].real(), yc[j].imag()};
    HIPCHK(hipMemcpyAsync(h->g_y.p, yd.data(), k * sizeof(zd), hipMemcpyHostToDevice, h->st));
    launch_gemv_n(n, k, (const zf *)V, n, h->g_y.p, d, h->g_pn.p, h->st);
    launch_zaxpy(n, zd{rn, 0.0}, d, x, h->st);
    if ((rc = residual(&rn))) return rc;
    be = rn / bn;
    ++outer;
  }
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipGetLastError());
  if (info) {
    info->info = be <= tol ? 0 : -4;
    info->iterations = total;
    info->arnoldi_backward_error = bea;
    info->backward_error = be;
    info->matvecs = matvecs;
    info->seconds = (now_ms() - t_start) / 1e3;
    info->true_backward_error = be;
  }
  h->refinements = outer;
  return SPK_OK;
}

// ---------------------------------------------------------------------------
// Mixed precision for R <= 16 right-hand sides of one mesh (incidence angles).
// Each right-hand side runs the refinement of gmres_mixed with its own
// complex64 Krylov basis; the inner Arnoldi steps of all right-hand sides
// that are inside an inner solve advance together, with one batched
// complex64 matvec per step, and the FP64 residuals of those that finish an
// inner solve in the same step are one batched FP64 matvec.  No counterpart
// in the reference.
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {
struct RhsState {
  // host copies of the inner solve
  std::vector<zc> H, g, rs, yc;
  std::vector<double> rcs;
  double bn = 0, rn = 0, be = 1, bea = 1, beta = 0, est = 1, stop_at = 0;
  int j = 0, total = 0, outer = 0, matvecs = 0, phase = 0;  // phase 0 inner, 1 residual, 2 done
  int relaxed = 0;  // outer products taken on the complex64 operator
};
}  // namespace

static int gmres_mixed_batch(spk_handle *h, int R, const zd *const *b, zd *const *x,
                             const spk_gmres_params *gp, spk_gmres_info *info) {
  if (!h->mixed_ready) return fail(SPK_ERR_STATE, "handle was not created with SPK_MIXED");
  if (R < 1 || R > kBatchMax) return fail(SPK_ERR_ARG, "batch size must be 1..16");
  if (gp->preconditioning != SPK_PRECOND_NONE)
    return fail(SPK_ERR_ARG, "preconditioning is not available with SPK_MIXED");
  const double t_start = now_ms();
  const int64_t n = 2 * (int64_t)h->N;
  const int max_it = std::max(1, gp->max_iterations);
  const int m = gp->restart > 0 ? std::min(gp->restart, max_it) : max_it;
  const double tol = gp->tolerance;
  const double inner_tol = gp->refinement_tolerance > 0 ? gp->refinement_tolerance : 3e-4;
  const int max_outer = gp->refinement_steps > 0 ? gp->refinement_steps : 30;
  const size_t ldh = (size_t)m + 2;
  // device state: bases, work vectors, coefficients
  DBuf<zf> V, W;
  DBuf<zd> Rv, Dv, Tv, Hd, Y, dots;
  DBuf<const zf *> xs32;
  DBuf<zf *> ys32;
  DBuf<const zd *> xs64;
  DBuf<zd *> ys64;
  HIPCHK(V.alloc((size_t)R * (m + 1) * n));
  HIPCHK(W.alloc((size_t)R * n));
  HIPCHK(Rv.alloc((size_t)R * n));
  HIPCHK(Dv.alloc(n));
  HIPCHK(Tv.alloc((size_t)R * n));
  HIPCHK(Hd.alloc((size_t)R * ldh));
  HIPCHK(Y.alloc(ldh));
  HIPCHK(dots.alloc(R));
  HIPCHK(xs32.alloc(kBatchMax));
  HIPCHK(ys32.alloc(kBatchMax));
  HIPCHK(xs64.alloc(kBatchMax));
  HIPCHK(ys64.alloc(kBatchMax));
  const size_t gc_elems = gemv_c_partial_elems(n, m + 1, 1);
  HIPCHK(h->g_partial.alloc(gc_elems + (size_t)reduce_blocks(n)));
  zd *partial_nrm = h->g_partial.p + gc_elems;
  HIPCHK(h->g_pn.alloc(gemv_n_partial_elems(n, m + 1)));
  HIPCHK(h->pin.ensure((size_t)R * ldh));
  // arguments of the batched Gram-Schmidt passes: basis, vector and
  // coefficient pointers, column counts, partial sums
  DBuf<uint64_t> orth_args;
  DBuf<int> orth_ks;
  DBuf<zd> orth_partial;
  HIPCHK(orth_args.alloc(4 * (size_t)kBatchMax));
  HIPCHK(orth_ks.alloc(kBatchMax));
  HIPCHK(orth_partial.alloc(batch_orth_partial_elems(n, R, m + 1)));
  auto Vr = [&](int r) { return V.p + (size_t)r * (m + 1) * n; };
  auto Wr = [&](int r) { return W.p + (size_t)r * n; };
  auto Rr = [&](int r) { return Rv.p + (size_t)r * n; };
  auto Tr = [&](int r) { return Tv.p + (size_t)r * n; };
  std::vector<RhsState> S(R);
  for (auto &st : S) {
    st.H.assign((size_t)(m + 1) * (m + 1), 0.0);
    st.g.assign(m + 1, 0.0);
    st.rs.assign(m + 1, 0.0);
    st.yc.assign(m + 1, 0.0);
    st.rcs.assign(m + 1, 0.0);
  }
  // |u_r|^2 for a list of right-hand sides, one transfer
  auto norms2 = [&](const std::vector<int> &rr, auto vec_of, std::vector<double> *out) -> int {
    for (size_t i = 0; i < rr.size(); ++i) {
      const auto *u = vec_of(rr[i]);
      launch_zdotc(n, u, u, partial_nrm, dots.p + i, h->st);
    }
    HIPCHK(hipMemcpyAsync(h->pin.p, dots.p, rr.size() * sizeof(zd), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    out->resize(rr.size());
    for (size_t i = 0; i < rr.size(); ++i) (*out)[i] = std::max(h->pin.p[i].x, 0.0);
    return SPK_OK;
  };
  int rc = SPK_OK;
  std::vector<int> all(R);
  for (int r = 0; r < R; ++r) all[r] = r;
  std::vector<double> nv;
  if ((rc = norms2(all, [&](int r) { return b[r]; }, &nv))) return rc;
  for (int r = 0; r < R; ++r) {
    S[r].bn = std::sqrt(nv[r]);
    if (!gp->use_initial_guess) HIPCHK(hipMemsetAsync(x[r], 0, n * sizeof(zd), h->st));
  }
  // residuals r = b - Z x (batched FP64 matvec) for a set of right-hand sides;
  // then each either finishes or starts its next inner solve
  auto residuals = [&](const std::vector<int> &set) -> int {
    if (set.empty()) return SPK_OK;
    std::vector<const zd *> xp;
    std::vector<zd *> yp;
    for (int r : set) {
      xp.push_back(x[r]);
      yp.push_back(Tr(r));
    }
    int e = SPK_OK;
    if (set.size() < 4 || h->quad) {
      // a few stragglers: the single-vector operator is faster than a narrow batch
      for (size_t i = 0; i < set.size() && !e; ++i) e = matvec_t<zd>(h, xp[i], yp[i]);
    } else {
      HIPCHK(hipMemcpy(xs64.p, xp.data(), set.size() * sizeof(void *), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(ys64.p, yp.data(), set.size() * sizeof(void *), hipMemcpyHostToDevice));
      e = matvec_batch_t<zd>(h, (int)set.size(), xs64.p, ys64.p);
    }
    if (e) return e;
    for (int r : set) {
      launch_zsub(n, b[r], Tr(r), Rr(r), h->st);
      ++S[r].matvecs;
    }
    std::vector<double> rr2;
    if ((e = norms2(set, Rr, &rr2))) return e;
    std::vector<int> start;
    for (size_t i = 0; i < set.size(); ++i) {
      RhsState &st = S[set[i]];
      st.rn = std::sqrt(rr2[i]);
      st.be = st.bn > 0 ? st.rn / st.bn : 0.0;
      if (st.outer == 0) st.bea = st.be;
      if (st.bn == 0.0 || st.be <= tol || st.total >= max_it || st.outer >= max_outer) {
        st.phase = 2;
      } else {
        st.phase = 0;
        start.push_back(set[i]);
      }
    }
    // V_0 = r / |r| in complex64, renormalised after rounding
    for (int r : start) launch_scale_down(n, 1.0 / S[r].rn, Rr(r), Vr(r), h->st);
    std::vector<double> vv;
    if ((e = norms2(start, [&](int r) { return (const zf *)Vr(r); }, &vv))) return e;
    for (size_t i = 0; i < start.size(); ++i) {
      RhsState &st = S[start[i]];
      st.beta = std::sqrt(vv[i]);
      launch_scal_copy_f(n, 1.0 / st.beta, Vr(start[i]), Vr(start[i]), h->st);
      std::fill(st.H.begin(), st.H.end(), zc(0.0));
      std::fill(st.g.begin(), st.g.end(), zc(0.0));
      st.g[0] = st.beta;
      st.j = 0;
      st.est = 1.0;
      st.stop_at = std::max(inner_tol, std::min(0.5, 0.5 * tol / st.be));
    }
    return SPK_OK;
  };
  if ((rc = residuals(all))) return rc;
  std::vector<int> act;
  for (;;) {
    act.clear();
    for (int r = 0; r < R; ++r)
      if (S[r].phase == 0) act.push_back(r);
    if (act.empty()) break;
    // one batched complex64 matvec: w_r = Z32 v_{r, j_r}
    std::vector<const zf *> xp;
    std::vector<zf *> yp;
    for (int r : act) {
      xp.push_back(Vr(r) + (size_t)S[r].j * n);
      yp.push_back(Wr(r));
    }
    if (act.size() < 4 || h->quad) {
      for (size_t i = 0; i < act.size() && !rc; ++i) rc = matvec_t<zf>(h, xp[i], yp[i]);
    } else {
      HIPCHK(hipMemcpy(xs32.p, xp.data(), act.size() * sizeof(void *), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(ys32.p, yp.data(), act.size() * sizeof(void *), hipMemcpyHostToDevice));
      rc = matvec_batch_t<zf>(h, (int)act.size(), xs32.p, ys32.p);
    }
    if (rc) return rc;
    // ICGS passes for all active right-hand sides; a second pass only for
    // those whose |w| fell below |h| (the CERFACS test)
    std::vector<int> pass = act;
    std::vector<double> dnormw(R, 0.0);
    for (int np_ = 1; !pass.empty() && np_ <= 3; ++np_) {
      // one batched ICGS pass for all right-hand sides in `pass`
      const int P = (int)pass.size();
      std::vector<uint64_t> arg(4 * (size_t)kBatchMax, 0);
      int maxk = 0;
      for (int i = 0; i < P; ++i) {
        const int r = pass[i], k = S[r].j + 1;
        maxk = std::max(maxk, k);
        arg[i] = (uint64_t)(uintptr_t)Vr(r);
        arg[kBatchMax + i] = (uint64_t)(uintptr_t)Wr(r);
        arg[2 * kBatchMax + i] = (uint64_t)(uintptr_t)(Hd.p + (size_t)r * ldh);
        arg[3 * kBatchMax + i] = (uint64_t)k;
      }
      std::vector<int> kk(P);
      for (int i = 0; i < P; ++i) kk[i] = (int)arg[3 * kBatchMax + i];
      HIPCHK(hipMemcpy(orth_args.p, arg.data(), 3 * kBatchMax * sizeof(uint64_t),
                       hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(orth_ks.p, kk.data(), P * sizeof(int), hipMemcpyHostToDevice));
      const zf *const *Vs = (const zf *const *)orth_args.p;
      zf *const *Ws = (zf *const *)(orth_args.p + kBatchMax);
      zd *const *Hs = (zd *const *)(orth_args.p + 2 * kBatchMax);
      launch_batch_orth(n, P, maxk, orth_ks.p, Vs, Ws, Hs, orth_partial.p, h->st);
      HIPCHK(hipMemcpyAsync(h->pin.p, Hd.p, (size_t)R * ldh * sizeof(zd), hipMemcpyDeviceToHost,
                            h->st));
      HIPCHK(hipStreamSynchronize(h->st));
      std::vector<int> again;
      for (int r : pass) {
        RhsState &st = S[r];
        const int j = st.j;
        const zd *hv = h->pin.p + (size_t)r * ldh;
        double dloo = 0.0;
        for (int i = 0; i <= j; ++i) {
          const zc hi(hv[i].x, hv[i].y);
          st.H[(size_t)j * (m + 1) + i] += hi;
          dloo += std::norm(hi);
        }
        dnormw[r] = std::sqrt(std::max(hv[j + 1].x, 0.0));
        if (dnormw[r] <= std::sqrt(dloo) && np_ < 3) again.push_back(r);
      }
      pass.swap(again);
    }
    // Hessenberg update, Givens rotations and the stopping test per right-hand side
    std::vector<int> fin;
    for (int r : act) {
      RhsState &st = S[r];
      const int j = st.j;
      auto HH = [&](int i, int c) -> zc & { return st.H[(size_t)c * (m + 1) + i]; };
      HH(j + 1, j) = dnormw[r];
      if (dnormw[r] > 0.0)
        launch_scal_copy_f(n, 1.0 / dnormw[r], Wr(r), Vr(r) + (size_t)(j + 1) * n, h->st);
      for (int i = 0; i < j; ++i) zrot1(&HH(i, j), &HH(i + 1, j), st.rcs[i], st.rs[i]);
      zc a = HH(j, j);
      double cc;
      zc ss;
      zrotg_(&a, HH(j + 1, j), &cc, &ss);
      st.rcs[j] = cc;
      st.rs[j] = ss;
      zrot1(&st.g[j], &st.g[j + 1], cc, ss);
      zrot1(&HH(j, j), &HH(j + 1, j), cc, ss);
      HH(j + 1, j) = 0.0;
      st.est = std::abs(st.g[j + 1]) / st.beta;
      ++st.total;
      ++st.matvecs;
      st.bea = st.est * st.be;
      st.j = j + 1;
      if (st.est <= st.stop_at || dnormw[r] == 0.0 || st.j >= m || st.total >= max_it) fin.push_back(r);
    }
    // finished inner solves: x += |r| V y, then their FP64 residuals together
    for (int r : fin) {
      RhsState &st = S[r];
      const int k = st.j;
      auto HH = [&](int i, int c) -> zc & { return st.H[(size_t)c * (m + 1) + i]; };
      for (int j = k - 1; j >= 0; --j) {
        zc sum = st.g[j];
        for (int i = j + 1; i < k; ++i) sum -= HH(j, i) * st.yc[i];
        st.yc[j] = sum / HH(j, j);
      }
      std::vector<zd> yd(k);
      for (int j = 0; j < k; ++j) yd[j] = zd{st.yc[j].real(), st.yc[j].imag()};
      HIPCHK(hipMemcpyAsync(Y.p, yd.data(), k * sizeof(zd), hipMemcpyHostToDevice, h->st));
      HIPCHK(hipStreamSynchronize(h->st));
      launch_gemv_n(n, k, (const zf *)Vr(r), n, Y.p, Dv.p, h->g_pn.p, h->st);
      launch_zaxpy(n, zd{st.rn, 0.0}, Dv.p, x[r], h->st);
      ++st.outer;
      st.phase = 1;
    }
    if ((rc = residuals(fin))) return rc;
  }
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipGetLastError());
  const double secs = (now_ms() - t_start) / 1e3;
  for (int r = 0; r < R; ++r) {
    const RhsState &st = S[r];
    info[r].info = st.be <= tol ? 0 : -4;
    info[r].iterations = st.total;
    info[r].arnoldi_backward_error = st.bea;
    info[r].backward_error = st.be;
    info[r].matvecs = st.matvecs;
    info[r].seconds = secs;
    info[r].true_backward_error = st.be;
  }
  return SPK_OK;
}

extern "C" {

extern "C++" {
// the bf16 copy of the complex64 near blocks (built once per setup)
static int ensure_near_bf(spk_handle *h) {
  if (!h->near_bf_ok) {
    HIPCHK(h->d_blk_bf.alloc(h->f_blocks.n));
    launch_to_bf((int64_t)h->f_blocks.n, h->f_blocks.p, h->d_blk_bf.p, h->st);
    HIPCHK(hipGetLastError());
    h->near_bf_ok = true;
  }
  return SPK_OK;
}

// the fp8 copy of the complex64 near blocks (built once per setup)
static int ensure_near_f8(spk_handle *h) {
  if (!h->near_f8_ok) {
    HIPCHK(h->d_blk_f8.alloc(h->f_blocks.n / 2));
    HIPCHK(h->d_f8_scale.alloc(4 * (size_t)h->N));
    launch_near_to_f8(h->N, h->d_leafbox_of_pos.p, h->d_box_first.p, h->d_e2_off.p,
                      h->d_near_off.p, h->f_blocks.p, h->d_blk_f8.p, h->d_f8_scale.p, h->st);
    HIPCHK(hipGetLastError());
    h->near_f8_ok = true;
  }
  return SPK_OK;
}

static DBuf<zf> &fgU(spk_handle::FgWork &F, zf *) { return F.Uf; }
static DBuf<zd> &fgU(spk_handle::FgWork &F, zd *) { return F.Ud; }
static DBuf<zf> &fgWi(spk_handle::FgWork &F, zf *) { return F.Wif; }
static DBuf<zd> &fgWi(spk_handle::FgWork &F, zd *) { return F.Wid; }

// Flexible GMRES on the FP64 operator for R <= kBatchMax right-hand sides,
// right-preconditioned by kin GMRES steps (from zero) on the near-field
// operator: complex64 on an SPK_MIXED handle (T = zf), FP64 otherwise.  The
// near field holds most of the operator's conditioning, so a short inner
// solve with it replaces most outer steps, each of which costs a full FP64
// product (near field plus the MLFMM far field).
//  - Inner: the active right-hand sides' inner steps share one (batched)
//    near-field product; their Arnoldi runs on the device (two classical
//    Gram-Schmidt passes, coefficients and norms stay in HBM) and the host
//    solves the small least-squares problems once per outer step.
//  - Outer: z_r = U_r y_r is kept, w = Z z_r in FP64, ICGS with the CERFACS
//    reorthogonalisation test, Givens rotations on the host; the Arnoldi
//    backward error is the stopping test, confirmed by the true residual
//    b - Z x (a new cycle starts from it otherwise, as after m steps).
// The block-Jacobi preconditioner of the current operator: groups are the
// occupied boxes one level above the leaves (leaf boxes are in Morton
// order, so a group's positions are contiguous), each group's dense
// near-field matrix is factorised and inverted (rocSOLVER getrf / getri,
// batched over the groups, padded to the largest group).
static int ensure_block_lu(spk_handle *h) {
  if (h->blu.ok) return SPK_OK;
  auto &B = h->blu;
  const Tree &t = h->tree;
  const Level &LL = t.lev[t.l_max];
  const int nbl = (int)LL.n.size();
  std::vector<int> p0, cnt;
  for (int b = 0; b < nbl;) {
    int e = b + 1;
    while (e < nbl && (LL.n[e] >> 3) == (LL.n[b] >> 3)) ++e;
    const int first = t.box_first[b], last = t.box_first[e - 1] + t.box_count[e - 1];
    p0.push_back(first);
    cnt.push_back(last - first);
    b = e;
  }
  int cmax = 0;
  for (int c : cnt) cmax = std::max(cmax, c);
  B.ngroups = (int)p0.size();
  B.nmax = 2 * cmax;
  const size_t nn = (size_t)B.nmax * B.nmax;
  const double gb = (double)B.ngroups * nn * sizeof(zd) / 1e9;
  if (gb > 64.0)
    return fail(SPK_ERR_ARG, "block-Jacobi groups too large (" + std::to_string(gb) +
                                 " GB of dense blocks): use smaller leaf boxes");
  HIPCHK(B.p0.upload(p0));
  HIPCHK(B.cnt.upload(cnt));
  HIPCHK(B.A.alloc((size_t)B.ngroups * nn));
  HIPCHK(B.X.alloc((size_t)B.ngroups * B.nmax));
  HIPCHK(B.Y.alloc((size_t)B.ngroups * B.nmax));
  HIPCHK(B.ipiv.alloc((size_t)B.ngroups * B.nmax));
  HIPCHK(B.info.alloc(B.ngroups));
  HIPCHK(hipMemsetAsync(B.A.p, 0, B.A.n * sizeof(zd), h->st));
  launch_blu_extract(B.ngroups, B.nmax, B.p0.p, B.cnt.p, h->d_perm.p, h->d_pos.p,
                     h->d_leafbox_of_pos.p, h->d_box_first.p, h->d_e2_off.p, h->d_e2_list.p,
                     h->d_near_off.p, h->d_blocks.p, B.A.p, h->st);
  HIPCHK(hipGetLastError());
  if (!B.rb && rocblas_create_handle(&B.rb) != rocblas_status_success)
    return fail(SPK_ERR_HIP, "rocblas_create_handle failed");
  if (rocblas_set_stream(B.rb, h->st) != rocblas_status_success)
    return fail(SPK_ERR_HIP, "rocblas_set_stream failed");
  auto *Ar = (rocblas_double_complex *)B.A.p;
  if (rocsolver_zgetrf_strided_batched(B.rb, B.nmax, B.nmax, Ar, B.nmax, (rocblas_stride)nn,
                                       B.ipiv.p, B.nmax, B.info.p, B.ngroups) !=
      rocblas_status_success)
    return fail(SPK_ERR_HIP, "rocsolver_zgetrf_strided_batched failed");
  std::vector<int> inf(B.ngroups);
  HIPCHK(hipMemcpyAsync(inf.data(), B.info.p, B.ngroups * sizeof(int), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  for (int g = 0; g < B.ngroups; ++g)
    if (inf[g] != 0)
      return fail(SPK_ERR_STATE, "block-Jacobi: group " + std::to_string(g) +
                                       " has a singular near-field matrix");
  if (rocsolver_zgetri_strided_batched(B.rb, B.nmax, Ar, B.nmax, (rocblas_stride)nn, B.ipiv.p,
                                       B.nmax, B.info.p, B.ngroups) != rocblas_status_success)
    return fail(SPK_ERR_HIP, "rocsolver_zgetri_strided_batched failed");
  HIPCHK(hipStreamSynchronize(h->st));
  B.ok = true;
  return SPK_OK;
}

// z = M^-1 v for the block-Jacobi preconditioner (stream-ordered)
static int apply_block_lu(spk_handle *h, const zd *v, zd *z) {
  auto &B = h->blu;
  launch_blu_gather(B.ngroups, B.nmax, h->N, B.p0.p, B.cnt.p, h->d_perm.p, v, B.X.p, h->st);
  const rocblas_double_complex one = {1.0, 0.0}, zero = {0.0, 0.0};
  const rocblas_stride nn = (rocblas_stride)B.nmax * B.nmax;
  if (rocblas_zgemv_strided_batched(B.rb, rocblas_operation_none, B.nmax, B.nmax, &one,
                                    (const rocblas_double_complex *)B.A.p, B.nmax, nn,
                                    (const rocblas_double_complex *)B.X.p, 1, B.nmax, &zero,
                                    (rocblas_double_complex *)B.Y.p, 1, B.nmax, B.ngroups) !=
      rocblas_status_success)
    return fail(SPK_ERR_HIP, "rocblas_zgemv_strided_batched failed");
  launch_blu_scatter(B.ngroups, B.nmax, h->N, B.p0.p, B.cnt.p, h->d_perm.p, B.Y.p, z, h->st);
  HIPCHK(hipGetLastError());
  return SPK_OK;
}

template <class T>
static int fgmres_near(spk_handle *h, int R, const zd *const *b, zd *const *x,
                       const spk_gmres_params *gp, spk_gmres_info *info) {
  if (R < 1 || R > kBatchMax) return fail(SPK_ERR_ARG, "batch size must be 1..16");
  if (std::is_same<T, zf>::value && !h->mixed_ready)
    return fail(SPK_ERR_STATE, "handle was not created with SPK_MIXED");
  // complex64 inner solves run on a low-precision copy of the near field:
  // the preconditioner only has to approximate the near-field solve.  The
  // fp8 copy by default (one SpMV per vector below 4 active right-hand
  // sides, the matrix-core product from 9, complex64 blocks between), the
  // bf16 copy with SPK_NEAR_F8=0.
  const uint32_t *bf = nullptr, *f8 = nullptr;
  // SPK_PRECOND_BLOCK_LU: the inner solve is the block-Jacobi inverse
  const bool blu = gp->preconditioning == SPK_PRECOND_BLOCK_LU;
  if (blu) {
    if (int e = ensure_block_lu(h)) return e;
  } else if constexpr (std::is_same<T, zf>::value) {
    const char *ev = getenv("SPK_NEAR_F8");
    if (!(ev && ev[0] == '0')) {
      const int e8 = ensure_near_f8(h);
      if (e8) return e8;
      f8 = h->d_blk_f8.p;
    } else {
      const int e = ensure_near_bf(h);
      if (e) return e;
      bf = h->d_blk_bf.p;
    }
  }
  const double t_start = now_ms();
  const int64_t n = 2 * (int64_t)h->N;
  const int max_it = std::max(1, gp->max_iterations);
  const int m = std::min(max_it, gp->restart > 0 ? gp->restart : 100);
  const int kin = gp->inner_iterations > 0 ? std::min(gp->inner_iterations, 64)
                  : h->tune_inner > 0 ? h->tune_inner
                                      : 30;
  const double tol = gp->tolerance;
  // relaxed outer products on a mixed handle (SPK_RELAX=0 keeps every one in FP64)
  const char *erx = getenv("SPK_RELAX");
  // (not with the block-Jacobi preconditioner: the systems it is for are
  // ill-conditioned enough that the complex64 operator's error stalls them)
  const bool relax_on = !(erx && erx[0] == '0') && !blu;
  // kRelaxDelta above the complex64 operator's error leaves room for the
  // residual gap (the relaxed products' errors add up to about kRelaxDelta'
  // times the sum of the later estimates): 1e-3 forced a second cycle on a
  // 5 x 5 lambda patch, 2e-3 and 5e-3 did not (and 5e-4 left a 16-angle batch
  // at 9.95e-7 of the 1e-6 tolerance)
  constexpr double kRelaxDelta = 5e-3;
  // SPK_INNER_GRAPH=0: launch the inner steps one by one instead of
  // replaying their captured graph (read at each solve)
  const char *gv = std::getenv("SPK_INNER_GRAPH");
  const bool use_graph = h->graph_ok && !(gv && gv[0] == '0');
  const size_t ldh = (size_t)m + 2, ldhi = (size_t)kin + 2, hik = (size_t)(kin + 1) * kin;
  const size_t ldu = (size_t)(kin + 1) * n;
  auto &F = h->fg;
  DBuf<zd> &V = F.V, &Z = F.Z, &W = F.W, &Hd = F.Hd, &Rv = F.Rv, &Tv = F.Tv, &Dv = F.Dv,
           &Yo = F.Yo, &Yi = F.Yi, &h1 = F.h1, &Hi = F.Hi, &dots = F.dots;
  DBuf<T> &U = fgU(F, (T *)nullptr), &Wi = fgWi(F, (T *)nullptr);
  HIPCHK(V.alloc((size_t)R * (m + 1) * n));
  HIPCHK(Z.alloc((size_t)R * m * n));
  HIPCHK(W.alloc((size_t)R * n));
  HIPCHK(Rv.alloc((size_t)R * n));
  HIPCHK(Tv.alloc((size_t)R * n));
  HIPCHK(Dv.alloc(n));
  HIPCHK(Hd.alloc((size_t)R * ldh));
  HIPCHK(Yo.alloc(ldh));
  HIPCHK(Yi.alloc((size_t)R * kin));
  HIPCHK(U.alloc((size_t)R * ldu));
  HIPCHK(Wi.alloc((size_t)R * n));
  HIPCHK(h1.alloc((size_t)R * ldhi));
  DBuf<zd> &apart = F.apart;
  HIPCHK(apart.alloc(arnoldi_partial_elems(n, R, kin)));
  HIPCHK(Hi.alloc((size_t)R * hik));
  HIPCHK(dots.alloc(R));
  const int kmax = std::max(m, kin) + 1;
  DBuf<zd> &part = F.part;
  HIPCHK(part.alloc(std::max(batch_orth_partial_elems(n, R, m + 1), (size_t)reduce_blocks(n))));
  HIPCHK(h->g_pn.alloc(gemv_n_partial_elems(n, kmax)));
  HIPCHK(h->pin.ensure(std::max((size_t)R * ldh, (size_t)R * hik)));
  // pointer tables: the inner products' inputs U_s[:, j] and outputs W_s
  // (fixed for the solve); the outer products' and Gram-Schmidt passes'
  // (per step; separate arrays: a table is rewritten only after the kernels
  // that read its last contents have completed)
  constexpr int B = kBatchMax;
  DBuf<uint64_t> &tin = F.tin, &tprod = F.tprod, &torth = F.torth;
  DBuf<int> &kout = F.kout;
  HIPCHK(tin.alloc((size_t)(1 + kin) * B));
  HIPCHK(tprod.alloc(2 * (size_t)B));
  HIPCHK(torth.alloc(3 * (size_t)B));
  HIPCHK(kout.alloc(B));
  {
    std::vector<uint64_t> t((size_t)(1 + kin) * B, 0);
    for (int s = 0; s < B; ++s) {
      const int q = s < R ? s : 0;
      t[s] = (uint64_t)(uintptr_t)(Wi.p + (size_t)q * n);
      for (int j = 0; j < kin; ++j)
        t[(size_t)(1 + j) * B + s] = (uint64_t)(uintptr_t)(U.p + q * ldu + (size_t)j * n);
    }
    HIPCHK(hipMemcpy(tin.p, t.data(), t.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  }
  T *const *tW = (T *const *)tin.p;
  auto Vr = [&](int r) { return V.p + (size_t)r * (m + 1) * n; };
  auto Zr = [&](int r) { return Z.p + (size_t)r * m * n; };
  auto Wr = [&](int r) { return W.p + (size_t)r * n; };
  auto Rr = [&](int r) { return Rv.p + (size_t)r * n; };
  auto Tr = [&](int r) { return Tv.p + (size_t)r * n; };
  std::vector<RhsState> S(R);
  for (auto &st : S) {
    st.H.assign((size_t)(m + 1) * (m + 1), 0.0);
    st.g.assign(m + 1, 0.0);
    st.rs.assign(m + 1, 0.0);
    st.yc.assign(m + 1, 0.0);
    st.rcs.assign(m + 1, 0.0);
  }
  if (R == 1) h->hist.clear();
  // |u|^2 for a list of right-hand sides, one transfer
  auto norms2 = [&](const std::vector<int> &rr, auto vec_of, std::vector<double> *out) -> int {
    for (size_t i = 0; i < rr.size(); ++i) {
      const zd *u = vec_of(rr[i]);
      launch_zdotc(n, u, u, part.p, dots.p + i, h->st);
    }
    HIPCHK(hipMemcpyAsync(h->pin.p, dots.p, rr.size() * sizeof(zd), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    out->resize(rr.size());
    for (size_t i = 0; i < rr.size(); ++i) (*out)[i] = std::max(h->pin.p[i].x, 0.0);
    return SPK_OK;
  };
  // FP64 products ys[i] = Z xs[i]: batched from 4 vectors up.  relax: on the
  // complex64 operator instead (T = zf only; slot i's first two inner basis
  // columns hold the single-precision copies, their inner solve is done)
  auto products = [&](const std::vector<const zd *> &xs, const std::vector<zd *> &ys,
                      bool relax) -> int {
    if constexpr (std::is_same<T, zf>::value) {
      if (relax) {
        for (size_t i = 0; i < xs.size(); ++i)
          launch_scale_down(n, 1.0, xs[i], (zf *)(U.p + i * ldu), h->st);
        if (xs.size() < 4 || h->quad) {
          for (size_t i = 0; i < xs.size(); ++i) {
            const int e = matvec_t<zf>(h, (const zf *)(U.p + i * ldu), (zf *)(U.p + i * ldu + n));
            if (e) return e;
          }
        } else {
          std::vector<const void *> pp(2 * B, nullptr);
          for (size_t i = 0; i < xs.size(); ++i) {
            pp[i] = U.p + i * ldu;
            pp[B + i] = U.p + i * ldu + n;
          }
          HIPCHK(hipMemcpy(tprod.p, pp.data(), 2 * B * sizeof(void *), hipMemcpyHostToDevice));
          const int e = matvec_batch_t<zf>(h, (int)xs.size(), (const zf *const *)tprod.p,
                                           (zf *const *)(tprod.p + B));
          if (e) return e;
        }
        for (size_t i = 0; i < xs.size(); ++i) {
          HIPCHK(hipMemsetAsync(ys[i], 0, n * sizeof(zd), h->st));
          launch_axpy_up(n, zd{1.0, 0.0}, (const zf *)(U.p + i * ldu + n), ys[i], h->st);
        }
        return SPK_OK;
      }
    }
    if (xs.size() < 4 || h->quad) {
      for (size_t i = 0; i < xs.size(); ++i) {
        const int e = matvec_t<zd>(h, xs[i], ys[i]);
        if (e) return e;
      }
      return SPK_OK;
    }
    // the previous product on these tables has completed (a synchronisation follows each use)
    HIPCHK(hipMemcpy(tprod.p, xs.data(), xs.size() * sizeof(void *), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(tprod.p + B, ys.data(), ys.size() * sizeof(void *), hipMemcpyHostToDevice));
    return matvec_batch_t<zd>(h, (int)xs.size(), (const zd *const *)tprod.p, (zd *const *)(tprod.p + B));
  };
  int rc = SPK_OK;
  std::vector<int> all(R);
  for (int r = 0; r < R; ++r) all[r] = r;
  std::vector<double> nv;
  if ((rc = norms2(all, [&](int r) { return b[r]; }, &nv))) return rc;
  for (int r = 0; r < R; ++r) {
    S[r].bn = std::sqrt(nv[r]);
    if (!gp->use_initial_guess) HIPCHK(hipMemsetAsync(x[r], 0, n * sizeof(zd), h->st));
  }
  bool first = true;
  // true residuals r = b - Z x of a set; each finishes or starts a cycle
  auto residuals = [&](const std::vector<int> &set) -> int {
    if (set.empty()) return SPK_OK;
    int e = SPK_OK;
    if (first && !gp->use_initial_guess) {
      for (int r : set) HIPCHK(hipMemcpyAsync(Rr(r), b[r], n * sizeof(zd), hipMemcpyDeviceToDevice, h->st));
    } else {
      std::vector<const zd *> xs;
      std::vector<zd *> ys;
      for (int r : set) {
        xs.push_back(x[r]);
        ys.push_back(Tr(r));
      }
      if ((e = products(xs, ys, false))) return e;
      for (int r : set) {
        launch_zsub(n, b[r], Tr(r), Rr(r), h->st);
        ++S[r].matvecs;
      }
    }
    first = false;
    std::vector<double> rr2;
    if ((e = norms2(set, Rr, &rr2))) return e;
    for (size_t i = 0; i < set.size(); ++i) {
      const int r = set[i];
      RhsState &st = S[r];
      st.rn = std::sqrt(rr2[i]);
      st.be = st.bn > 0 ? st.rn / st.bn : 0.0;
      if (st.outer == 0) st.bea = st.be;
      if (st.bn == 0.0 || st.be <= tol || st.total >= max_it) {
        st.phase = 2;
        continue;
      }
      st.phase = 0;
      launch_zscal_copy(n, zd{1.0 / st.rn, 0.0}, Rr(r), Vr(r), h->st);
      std::fill(st.H.begin(), st.H.end(), zc(0.0));
      std::fill(st.g.begin(), st.g.end(), zc(0.0));
      st.g[0] = st.rn;
      st.beta = st.rn;
      st.j = 0;
    }
    return SPK_OK;
  };
  if ((rc = residuals(all))) return rc;
  std::vector<int> act;
  for (;;) {
    act.clear();
    for (int r = 0; r < R; ++r)
      if (S[r].phase == 0) act.push_back(r);
    if (act.empty()) break;
    const int P = (int)act.size();
    // ---- inner near-field solves: slot s works on right-hand side act[s]
    std::vector<const zd *> zin;
    std::vector<zd *> wout;
    if (blu) {
      for (int s = 0; s < P; ++s) {
        const int r = act[s];
        zd *z = Zr(r) + (size_t)S[r].j * n;
        if (int e = apply_block_lu(h, Vr(r) + (size_t)S[r].j * n, z)) return e;
        zin.push_back(z);
        wout.push_back(Wr(r));
      }
    } else {
      for (int s = 0; s < P; ++s) {
        const zd *v = Vr(act[s]) + (size_t)S[act[s]].j * n;
        if constexpr (std::is_same<T, zf>::value)
          launch_scale_down(n, 1.0, v, (zf *)(U.p + s * ldu), h->st);
        else
          HIPCHK(hipMemcpyAsync(U.p + s * ldu, v, n * sizeof(zd), hipMemcpyDeviceToDevice, h->st));
      }
      // the inner steps' launches (near-field products, Arnoldi updates)
      auto inner_steps = [&]() -> int {
      int rc = SPK_OK;
      for (int j = 0; j < kin; ++j) {
        if (P < 4) {
          for (int s = 0; s < P && !rc; ++s) {
            T *ui = U.p + s * ldu + (size_t)j * n, *wo = Wi.p + (size_t)s * n;
            if constexpr (std::is_same<T, zf>::value) {
              if (f8) {
                launch_near_spmv_f8(h->spmv_wgs, h->N, h->d_row_box.p, h->d_row_r0.p, h->d_box_first.p,
                                    h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p,
                                    h->d_perm.p, f8, h->d_f8_scale.p, ui, wo, h->st);
                continue;
              }
              if (bf) {
                launch_near_spmv_bf(h->spmv_wgs, h->N, h->d_row_box.p, h->d_row_r0.p, h->d_box_first.p,
                                    h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p,
                                    h->d_perm.p, bf, ui, wo, h->st);
                continue;
              }
            }
            rc = matvec_t<T>(h, ui, wo, true, false);
          }
        } else {
          rc = matvec_batch_t<T>(h, P, (const T *const *)(tin.p + (size_t)(1 + j) * B), tW, false,
                                 f8 ? f8 : bf, f8 ? h->d_f8_scale.p : nullptr);
        }
        if (rc) return rc;
        launch_arnoldi_cgs<T>(n, P, j, kin, U.p, ldu, Wi.p, h1.p, (int)ldhi, Hi.p, apart.p, h->st);
      }
      return rc;
      };
      if (use_graph && (f8 || bf)) {
        // batch buffers sized before the capture (no allocation inside it)
        if (P >= 4 && (rc = batch_ensure<T>(h, P))) return rc;
        const std::vector<uint64_t> key = {
            (uint64_t)P, (uint64_t)kin, (uint64_t)n, (uint64_t)(uintptr_t)f8,
            (uint64_t)(uintptr_t)bf, (uint64_t)(uintptr_t)U.p, (uint64_t)(uintptr_t)Wi.p,
            (uint64_t)(uintptr_t)h1.p, (uint64_t)(uintptr_t)Hi.p, (uint64_t)(uintptr_t)apart.p,
            (uint64_t)(uintptr_t)tin.p, (uint64_t)(uintptr_t)pBB(h, (T *)nullptr).XI.p,
            (uint64_t)(uintptr_t)pBB(h, (T *)nullptr).XB.p, (uint64_t)h->setup_serial};
        if (h->ig.key != key || !h->ig.exec) {
          h->ig.reset();
          hipGraph_t graph = nullptr;
          HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
          const int rc_cap = inner_steps();
          const hipError_t e_end = hipStreamEndCapture(h->st, &graph);
          if (rc_cap) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc_cap;
          }
          HIPCHK(e_end);
          const hipError_t e_inst = hipGraphInstantiate(&h->ig.exec, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          HIPCHK(e_inst);
          h->ig.key = key;
        }
        HIPCHK(hipGraphLaunch(h->ig.exec, h->st));
      } else if ((rc = inner_steps())) {
        return rc;
      }
      HIPCHK(hipMemcpyAsync(h->pin.p, Hi.p, (size_t)P * hik * sizeof(zd), hipMemcpyDeviceToHost, h->st));
      HIPCHK(hipStreamSynchronize(h->st));
      // least squares min |e_0 - H y| per slot (Givens), up to a breakdown
      std::vector<zd> yall((size_t)R * kin, zd{0, 0});
      std::vector<int> klen(P);
      for (int s = 0; s < P; ++s) {
        const zd *Hs = h->pin.p + (size_t)s * hik;
        std::vector<zc> Hc((size_t)(kin + 1) * kin), gg(kin + 1, 0.0), yy(kin, 0.0);
        for (size_t i = 0; i < Hc.size(); ++i) Hc[i] = zc(Hs[i].x, Hs[i].y);
        auto HC = [&](int i, int c) -> zc & { return Hc[(size_t)c * (kin + 1) + i]; };
        int kk = kin;
        for (int c = 0; c < kin; ++c)
          if (std::abs(HC(c + 1, c)) == 0.0) {
            kk = c + 1;
            break;
          }
        gg[0] = 1.0;
        std::vector<double> cs(kk);
        std::vector<zc> sn(kk);
        for (int c = 0; c < kk; ++c) {
          for (int i = 0; i < c; ++i) zrot1(&HC(i, c), &HC(i + 1, c), cs[i], sn[i]);
          zc a = HC(c, c);
          zrotg_(&a, HC(c + 1, c), &cs[c], &sn[c]);
          zrot1(&gg[c], &gg[c + 1], cs[c], sn[c]);
          zrot1(&HC(c, c), &HC(c + 1, c), cs[c], sn[c]);
          HC(c + 1, c) = 0.0;
        }
        while (kk > 0 && std::abs(HC(kk - 1, kk - 1)) == 0.0) --kk;
        for (int c = kk - 1; c >= 0; --c) {
          zc sum = gg[c];
          for (int i = c + 1; i < kk; ++i) sum -= HC(c, i) * yy[i];
          yy[c] = sum / HC(c, c);
        }
        for (int c = 0; c < kk; ++c) yall[(size_t)s * kin + c] = zd{yy[c].real(), yy[c].imag()};
        klen[s] = kk;
      }
      // from pinned memory, which is next written by a later copy on this stream
      std::copy(yall.begin(), yall.begin() + (size_t)P * kin, h->pin.p);
      HIPCHK(hipMemcpyAsync(Yi.p, h->pin.p, (size_t)P * kin * sizeof(zd), hipMemcpyHostToDevice, h->st));
      for (int s = 0; s < P; ++s) {
        const int r = act[s];
        zd *z = Zr(r) + (size_t)S[r].j * n;
        if (klen[s] > 0)
          launch_gemv_n(n, klen[s], (const T *)(U.p + s * ldu), n, Yi.p + (size_t)s * kin, z, h->g_pn.p,
                        h->st);
        else  // no usable inner step: the identity preconditioner
          HIPCHK(hipMemcpyAsync(z, Vr(r) + (size_t)S[r].j * n, n * sizeof(zd), hipMemcpyDeviceToDevice,
                                h->st));
        zin.push_back(z);
        wout.push_back(Wr(r));
      }
    }
    // ---- outer step: w_r = Z z_r, ICGS against V_r.  Relaxed products:
    // once every active estimate is at or below tol / kRelaxDelta, the
    // complex64 operator (about 5e-4 from the FP64 one) stays within the
    // operator error the outer iteration tolerates (inexact Krylov: Bouras
    // and Fraysse 2005, Simoncini and Szyld 2003); the true residual that
    // ends a cycle is computed in FP64.
    bool relax = std::is_same<T, zf>::value && relax_on;
    for (int r : act) relax = relax && S[r].est * kRelaxDelta <= tol;
    if ((rc = products(zin, wout, relax))) return rc;
    for (int r : act) S[r].relaxed += relax ? 1 : 0;
    std::vector<int> pass = act;
    std::vector<double> dnormw(R, 0.0);
    for (int np_ = 1; !pass.empty() && np_ <= 3; ++np_) {
      const int Pp = (int)pass.size();
      std::vector<uint64_t> arg(3 * (size_t)B, 0);
      std::vector<int> kk(Pp);
      int maxk = 0;
      for (int i = 0; i < Pp; ++i) {
        const int r = pass[i];
        kk[i] = S[r].j + 1;
        maxk = std::max(maxk, kk[i]);
        arg[i] = (uint64_t)(uintptr_t)Vr(r);
        arg[B + i] = (uint64_t)(uintptr_t)Wr(r);
        arg[2 * B + i] = (uint64_t)(uintptr_t)(Hd.p + (size_t)r * ldh);
      }
      HIPCHK(hipMemcpy(torth.p, arg.data(), 3 * B * sizeof(uint64_t), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(kout.p, kk.data(), Pp * sizeof(int), hipMemcpyHostToDevice));
      launch_batch_orth<zd>(n, Pp, maxk, kout.p, (const zd *const *)torth.p, (zd *const *)(torth.p + B),
                            (zd *const *)(torth.p + 2 * B), part.p, h->st);
      HIPCHK(hipMemcpyAsync(h->pin.p, Hd.p, (size_t)R * ldh * sizeof(zd), hipMemcpyDeviceToHost, h->st));
      HIPCHK(hipStreamSynchronize(h->st));
      std::vector<int> again;
      for (int r : pass) {
        RhsState &st = S[r];
        const int j = st.j;
        const zd *hv = h->pin.p + (size_t)r * ldh;
        double dloo = 0.0;
        for (int i = 0; i <= j; ++i) {
          const zc hi(hv[i].x, hv[i].y);
          st.H[(size_t)j * (m + 1) + i] += hi;
          dloo += std::norm(hi);
        }
        dnormw[r] = std::sqrt(std::max(hv[j + 1].x, 0.0));
        if (dnormw[r] <= std::sqrt(dloo) && np_ < 3) again.push_back(r);
      }
      pass.swap(again);
    }
    std::vector<int> fin;
    for (int r : act) {
      RhsState &st = S[r];
      const int j = st.j;
      auto HH = [&](int i, int c) -> zc & { return st.H[(size_t)c * (m + 1) + i]; };
      HH(j + 1, j) = dnormw[r];
      if (dnormw[r] > 0.0) launch_zscal_copy(n, zd{1.0 / dnormw[r], 0.0}, Wr(r), Vr(r) + (size_t)(j + 1) * n, h->st);
      for (int i = 0; i < j; ++i) zrot1(&HH(i, j), &HH(i + 1, j), st.rcs[i], st.rs[i]);
      zc a = HH(j, j);
      double cc;
      zc ss;
      zrotg_(&a, HH(j + 1, j), &cc, &ss);
      st.rcs[j] = cc;
      st.rs[j] = ss;
      zrot1(&st.g[j], &st.g[j + 1], cc, ss);
      zrot1(&HH(j, j), &HH(j + 1, j), cc, ss);
      HH(j + 1, j) = 0.0;
      st.est = std::abs(st.g[j + 1]) / st.bn;
      st.bea = st.est;
      ++st.total;
      ++st.matvecs;
      st.j = j + 1;
      if (R == 1) h->hist.push_back(st.est);
      if (st.est <= tol || dnormw[r] == 0.0 || st.j >= m || st.total >= max_it) fin.push_back(r);
    }
    // finished cycles: x += Z y, then their true residuals together
    for (int r : fin) {
      RhsState &st = S[r];
      const int k = st.j;
      auto HH = [&](int i, int c) -> zc & { return st.H[(size_t)c * (m + 1) + i]; };
      for (int j = k - 1; j >= 0; --j) {
        zc sum = st.g[j];
        for (int i = j + 1; i < k; ++i) sum -= HH(j, i) * st.yc[i];
        st.yc[j] = sum / HH(j, j);
      }
      std::vector<zd> yd(k);
      for (int j = 0; j < k; ++j) yd[j] = zd{st.yc[j].real(), st.yc[j].imag()};
      HIPCHK(hipMemcpyAsync(Yo.p, yd.data(), k * sizeof(zd), hipMemcpyHostToDevice, h->st));
      HIPCHK(hipStreamSynchronize(h->st));
      launch_gemv_n(n, k, (const zd *)Zr(r), n, Yo.p, Dv.p, h->g_pn.p, h->st);
      launch_zaxpy(n, zd{1.0, 0.0}, Dv.p, x[r], h->st);
      ++st.outer;
      st.phase = 1;
    }
    if ((rc = residuals(fin))) return rc;
  }
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipGetLastError());
  const double secs = (now_ms() - t_start) / 1e3;
  for (int r = 0; r < R; ++r) {
    const RhsState &st = S[r];
    info[r].info = st.be <= tol ? 0 : -4;
    info[r].iterations = st.total;
    info[r].arnoldi_backward_error = st.bea;
    info[r].backward_error = st.be;
    info[r].matvecs = st.matvecs;
    info[r].seconds = secs;
    info[r].true_backward_error = st.be;
  }
  return SPK_OK;
}
}  // extern "C++"

// The near-field part of the operator on an SPK_MIXED handle's complex64
// blocks or on the preconditioner's bf16 / fp8 copies, for R complex64
// vectors: one launch per vector (the single-vector kernels) or one batched
// launch (the kernels the batched inner solves use).
int spk_near_product_lowp(spk_handle *h, int format, int batched, int R, const void *const *d_x,
                          void *const *d_y) {
  if (!h || !d_x || !d_y) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (!h->mixed_ready) return fail(SPK_ERR_STATE, "handle was not created with SPK_MIXED");
  if (R < 1 || R > kBatchMax) return fail(SPK_ERR_ARG, "batch size must be 1..16");
  if (format < SPK_NEAR_C64 || format > SPK_NEAR_FP8) return fail(SPK_ERR_ARG, "unknown format");
  for (int r = 0; r < R; ++r)
    if (!d_x[r] || !d_y[r]) return fail(SPK_ERR_ARG, "null vector in the batch");
  HIPCHK(hipSetDevice(h->p.device));
  int rc = SPK_OK;
  if (format == SPK_NEAR_BF16 && (rc = ensure_near_bf(h))) return rc;
  if (format == SPK_NEAR_FP8 && (rc = ensure_near_f8(h))) return rc;
  const uint32_t *bf = format == SPK_NEAR_BF16 ? h->d_blk_bf.p : nullptr;
  const uint32_t *f8 = format == SPK_NEAR_FP8 ? h->d_blk_f8.p : nullptr;
  if (!batched) {
    for (int r = 0; r < R; ++r) {
      const zf *x = (const zf *)d_x[r];
      zf *y = (zf *)d_y[r];
      if (f8)
        launch_near_spmv_f8(h->spmv_wgs, h->N, h->d_row_box.p, h->d_row_r0.p, h->d_box_first.p,
                            h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p,
                            h->d_perm.p, f8, h->d_f8_scale.p, x, y, h->st);
      else if (bf)
        launch_near_spmv_bf(h->spmv_wgs, h->N, h->d_row_box.p, h->d_row_r0.p, h->d_box_first.p,
                            h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p,
                            h->d_perm.p, bf, x, y, h->st);
      else if ((rc = matvec_t<zf>(h, x, y, true, false)))
        return rc;
    }
  } else {
    if ((rc = batch_ensure<zf>(h, R))) return rc;
    BatchBufs<zf> &BB = h->bf;
    HIPCHK(hipMemcpy(BB.xs.p, d_x, R * sizeof(void *), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(BB.ys.p, d_y, R * sizeof(void *), hipMemcpyHostToDevice));
    if ((rc = matvec_batch_t<zf>(h, R, BB.xs.p, BB.ys.p, false, f8 ? f8 : bf,
                                 f8 ? h->d_f8_scale.p : nullptr)))
      return rc;
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->st));
  return SPK_OK;
}

int spk_gmres_batch_device(spk_handle *h, int R, const void *const *d_b, void *const *d_x,
                           const spk_gmres_params *gp, spk_gmres_info *info) {
  if (!h || !d_b || !d_x || !gp || !info) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (R < 1 || R > kBatchMax) return fail(SPK_ERR_ARG, "batch size must be 1..16");
  for (int r = 0; r < R; ++r)
    if (!d_b[r] || !d_x[r]) return fail(SPK_ERR_ARG, "null vector in the batch");
  if (gp->preconditioning == SPK_PRECOND_NEAR_FIELD ||
      gp->preconditioning == SPK_PRECOND_BLOCK_LU)
    return h->p.precision == SPK_MIXED
               ? fgmres_near<zf>(h, R, (const zd *const *)d_b, (zd *const *)d_x, gp, info)
               : fgmres_near<zd>(h, R, (const zd *const *)d_b, (zd *const *)d_x, gp, info);
  if (h->p.precision != SPK_MIXED)
    return fail(SPK_ERR_STATE, "the batched refinement solve runs on an SPK_MIXED handle");
  return gmres_mixed_batch(h, R, (const zd *const *)d_b, (zd *const *)d_x, gp, info);
}

// Run-time tuning of the near-field preconditioned solver on the handle's
// mesh: the batched fp8 near product's row tiles per workgroup (1, 2, 4;
// each timed over 20 launches with HIP events at 16 vectors) and the inner
// solve length (20, 25, 30, 36, 40; each timed as one batched solve of the
// R incidences with gp).  The fastest of each is kept in the handle.
int spk_autotune(spk_handle *h, int R, const double *k_in, const double *e_in,
                 const spk_gmres_params *gp, spk_tuning *out) {
  if (!h || !k_in || !e_in || !gp || !out) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (R < 1 || R > kBatchMax) return fail(SPK_ERR_ARG, "batch size must be 1..16");
  if (h->transposed) return fail(SPK_ERR_STATE, "the handle holds the transposed operator");
  HIPCHK(hipSetDevice(h->p.device));
  *out = spk_tuning{};
  const int64_t n = 2 * (int64_t)h->N;
  int rc = SPK_OK;
  // row tiles of the batched fp8 product (mixed handles: the only ones
  // whose preconditioner runs it)
  const int tiles[3] = {1, 2, 4};
  int best_tpw = h->tune_tpw > 0 ? h->tune_tpw : near_mx_tiles_per_group();
  if (h->mixed_ready) {
    if ((rc = ensure_near_f8(h))) return rc;
    constexpr int kR = 16;
    if ((rc = batch_ensure<zf>(h, kR))) return rc;
    DBuf<zf> X, Y;
    {
      std::vector<zf> hx((size_t)kR * n);
      for (size_t i = 0; i < hx.size(); ++i)
        hx[i] = zf{(float)std::cos(0.37 * (double)i), (float)std::sin(0.59 * (double)i)};
      HIPCHK(X.upload(hx));
    }
    HIPCHK(Y.alloc((size_t)kR * n));
    std::vector<const zf *> xp(kR);
    std::vector<zf *> yp(kR);
    for (int r = 0; r < kR; ++r) {
      xp[r] = X.p + r * n;
      yp[r] = Y.p + r * n;
    }
    BatchBufs<zf> &BB = h->bf;
    HIPCHK(hipMemcpy(BB.xs.p, xp.data(), kR * sizeof(void *), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(BB.ys.p, yp.data(), kR * sizeof(void *), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    double best = 1e300;
    for (int i = 0; i < 3 && !rc; ++i) {
      if ((rc = build_mx_groups(h, tiles[i]))) break;
      constexpr int kWarm = 3, kReps = 20;
      for (int k = 0; k < kWarm + kReps && !rc; ++k) {
        if (k == kWarm) (void)hipEventRecord(e0, h->st);
        rc = matvec_batch_t<zf>(h, kR, BB.xs.p, BB.ys.p, false, h->d_blk_f8.p, h->d_f8_scale.p);
      }
      if (rc) break;
      (void)hipEventRecord(e1, h->st);
      if (hipEventSynchronize(e1) != hipSuccess) {
        rc = fail(SPK_ERR_HIP, "event synchronisation failed");
        break;
      }
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      out->tile_us[i] = 1e3 * ms / kReps;
      if (out->tile_us[i] < best) {
        best = out->tile_us[i];
        best_tpw = tiles[i];
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    h->tune_tpw = best_tpw;
    if ((rc = build_mx_groups(h, best_tpw))) return rc;
  }
  out->tiles_per_group = best_tpw;
  // inner solve length: one batched preconditioned solve per candidate
  std::vector<DBuf<zd>> B(R), X(R);
  std::vector<const void *> bp(R);
  std::vector<void *> xp(R);
  for (int r = 0; r < R; ++r) {
    HIPCHK(B[r].alloc(n));
    HIPCHK(X[r].alloc(n));
    if ((rc = rhs_plane_dev(h, k_in + 3 * r, e_in + 3 * r, B[r].p))) return rc;
    bp[r] = B[r].p;
    xp[r] = X[r].p;
  }
  spk_gmres_params g = *gp;
  g.preconditioning = SPK_PRECOND_NEAR_FIELD;
  g.use_initial_guess = 0;
  std::vector<spk_gmres_info> inf(R);
  const int cand[5] = {20, 25, 30, 36, 40};
  g.inner_iterations = 30;
  if ((rc = spk_gmres_batch_device(h, R, bp.data(), xp.data(), &g, inf.data()))) return rc;
  double best = 1e300;
  int best_in = 30;
  for (int i = 0; i < 5; ++i) {
    g.inner_iterations = cand[i];
    const double t0 = now_ms();
    if ((rc = spk_gmres_batch_device(h, R, bp.data(), xp.data(), &g, inf.data()))) return rc;
    const double ms = now_ms() - t0;
    bool conv = true;
    for (int r = 0; r < R; ++r) conv = conv && inf[r].info == 0;
    out->inner_tried[i] = cand[i];
    out->inner_ms[i] = conv ? ms : -ms;
    if (conv && ms < best) {
      best = ms;
      best_in = cand[i];
    }
  }
  out->n_inner = 5;
  h->tune_inner = best_in;
  out->inner_iterations = best_in;
  return SPK_OK;
}

// Per-level truncation numbers for the next spk_setup_tri.
int spk_set_level_truncation(spk_handle *h, int n, const int *K, int rr_mode) {
  if (!h || (n > 0 && !K)) return fail(SPK_ERR_ARG, "null argument");
  if (n < 0 || n > 32) return fail(SPK_ERR_ARG, "n must be 0..32");
  if (rr_mode != SPK_RR_REFERENCE && rr_mode != SPK_RR_ANTERPOLATE)
    return fail(SPK_ERR_ARG, "rr_mode must be SPK_RR_REFERENCE or SPK_RR_ANTERPOLATE");
  for (int l = 0; l < n; ++l)
    if (K[l] < 0 || K[l] > 120) return fail(SPK_ERR_ARG, "truncation numbers must be 0..120");
  h->levK_req.assign(K, K + n);
  h->rr_mode = rr_mode;
  return SPK_OK;
}

// Applies tuned values (spk_autotune on another handle of a similar mesh):
// row tiles per workgroup of the batched fp8 near product (1, 2 or 4) and
// the inner solve length used when spk_gmres_params.inner_iterations is 0;
// 0 restores a default.
int spk_set_tuning(spk_handle *h, int tiles_per_group, int inner_iterations) {
  if (!h) return fail(SPK_ERR_ARG, "null argument");
  if (tiles_per_group != 0 && tiles_per_group != 1 && tiles_per_group != 2 && tiles_per_group != 4)
    return fail(SPK_ERR_ARG, "tiles_per_group must be 0, 1, 2 or 4");
  if (inner_iterations < 0 || inner_iterations > 64)
    return fail(SPK_ERR_ARG, "inner_iterations must be 0..64");
  h->tune_tpw = tiles_per_group;
  h->tune_inner = inner_iterations;
  if (h->ready) {
    HIPCHK(hipSetDevice(h->p.device));
    return build_mx_groups(h, tiles_per_group > 0 ? tiles_per_group : near_mx_tiles_per_group());
  }
  return SPK_OK;
}

int spk_gmres_device(spk_handle *h, const void *d_b, void *d_x, const spk_gmres_params *gp,
                     spk_gmres_info *info) {
  if (!h || !d_b || !d_x || !gp) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (gp->preconditioning == SPK_PRECOND_NEAR_FIELD ||
      gp->preconditioning == SPK_PRECOND_BLOCK_LU) {
    const zd *bb[1] = {(const zd *)d_b};
    zd *xx[1] = {(zd *)d_x};
    spk_gmres_info tmp;
    const int e = h->p.precision == SPK_MIXED ? fgmres_near<zf>(h, 1, bb, xx, gp, &tmp)
                                              : fgmres_near<zd>(h, 1, bb, xx, gp, &tmp);
    if (!e && info) *info = tmp;
    return e;
  }
  if (h->p.precision == SPK_MIXED) return gmres_mixed(h, (const zd *)d_b, (zd *)d_x, gp, info);
  const double t_start = now_ms();
  const int64_t n = 2 * (int64_t)h->N;
  const int iterMax = std::max(1, gp->max_iterations);
  int m = gp->restart > 0 ? std::min(gp->restart, iterMax) : iterMax;
  const double tol = gp->tolerance;
  const int orth = gp->orthogonalization;
  const zd *b = (const zd *)d_b;
  zd *x = (zd *)d_x;
  if (h->g_m != m || h->g_V.n != (size_t)n * (m + 1)) {
    HIPCHK(h->g_V.alloc((size_t)n * (m + 1)));
    h->g_m = m;
  }
  if (h->g_w.n != (size_t)n) {
    HIPCHK(h->g_w.alloc(n));
    HIPCHK(h->g_r0.alloc(n));
    HIPCHK(h->g_xc.alloc(n));
  }
  HIPCHK(h->g_h.alloc(m + 1));
  HIPCHK(h->pin.ensure(m + 2));
  HIPCHK(h->g_y.alloc(m + 1));
  const size_t gc_elems = gemv_c_partial_elems(n, m + 1, 2);
  HIPCHK(h->g_partial.alloc(gc_elems + (size_t)reduce_blocks(n)));
  zd *partial_nrm = h->g_partial.p + gc_elems;
  HIPCHK(h->g_pn.alloc(gemv_n_partial_elems(n, m + 1)));
  HIPCHK(h->g_scal.alloc(4));
  // MGS and IMGS passes: by default the one-reduce form (h = (I + L)^-1 V^H w,
  // two reads of V per pass); SPK_MGS=sweep selects the cooperative
  // column-by-column sweep, SPK_MGS=steps one launch pair per column.
  const char *mgs_env = std::getenv("SPK_MGS");
  const std::string mgs_kind = mgs_env ? mgs_env : "";
  const bool is_mgs = orth == SPK_ORTH_MGS || orth == SPK_ORTH_IMGS;
  const int sweepG = (is_mgs && mgs_kind == "sweep") ? mgs_sweep_blocks(n) : 0;
  const bool mgs_steps = is_mgs && mgs_kind == "steps";
  const bool one_reduce = is_mgs && !sweepG && !mgs_steps;
  const int ldt = m + 1;
  if (one_reduce && (h->g_T_m != m || !h->g_T.p)) {
    HIPCHK(h->g_T.alloc((size_t)ldt * ldt));
    HIPCHK(hipMemsetAsync(h->g_T.p, 0, (size_t)ldt * ldt * sizeof(zd), h->st));
    HIPCHK(h->g_c.alloc(2 * (size_t)ldt));
    HIPCHK(h->g_tg.alloc(ldt));
    // T products: k <= ldt columns, at most ceil(k / 32) column chunks
    HIPCHK(h->g_pT.alloc(std::max(gemv_c_partial_elems(ldt, ldt, 1), (size_t)((ldt + 31) / 32) * ldt)));
    h->g_T_m = m;
  }
  if (sweepG) {
    const size_t need = (size_t)(m + 1) * 256 * mgs_sweep_slot_bytes();
    if (h->g_slots.n < need) {
      HIPCHK(h->g_slots.alloc(need));
      HIPCHK(hipMemsetAsync(h->g_slots.p, 0, need, h->st));
      h->sweep_seq = 0;
    }
    if (!h->g_ctl.p) {
      HIPCHK(h->g_ctl.alloc(1));
      HIPCHK(hipMemsetAsync(h->g_ctl.p, 0, sizeof(unsigned long long), h->st));
    }
  }
  zd *V = h->g_V.p, *w = h->g_w.p, *r0 = h->g_r0.p, *xc = h->g_xc.p;
  std::vector<zc> H((size_t)(m + 1) * (m + 1), 0.0), yc(m + 1), rs(m + 1), hh(m + 1);
  std::vector<double> rc(m + 1);
  auto HH = [&](int i, int j) -> zc & { return H[(size_t)j * (m + 1) + i]; };
  h->hist.clear();
  int matvecs = 0, info1 = 0, it_total = 0, compRsd = 0;
  double be = 0.0, bea = 0.0;
  zc dtmp;
  int rc_ = SPK_OK;
  // left Jacobi preconditioning: M^-1 applied where zPackgmres requests
  // precondLeft (:890-923, :955, :1054, :1310-1340, :1510)
  const bool prec = gp->preconditioning == SPK_PRECOND_LEFT_JACOBI;
  if (gp->preconditioning != SPK_PRECOND_NONE && !prec)
    return fail(SPK_ERR_ARG, "unknown preconditioning");
  if (prec && !h->diag_ready) {
    HIPCHK(h->d_diag.alloc(n));
    launch_jacobi_diag(h->N, h->d_perm.p, h->d_leafbox_of_pos.p, h->d_box_first.p, h->d_e2_off.p,
                       h->d_e2_list.p, h->d_near_off.p, h->d_blocks.p, h->d_diag.p, h->st);
    HIPCHK(hipGetLastError());
    h->diag_ready = true;
  }
  auto precond = [&](const zd *in, zd *out) {
    if (prec) launch_div_diag(n, h->d_diag.p, in, out, h->st);
    else if (in != out)
      (void)hipMemcpyAsync(out, in, n * sizeof(zd), hipMemcpyDeviceToDevice, h->st);
  };
  double tbe = 0.0;
  if (!gp->use_initial_guess) HIPCHK(hipMemsetAsync(x, 0, n * sizeof(zd), h->st));
  if ((rc_ = dev_dot(h, b, b, n, &dtmp))) return rc_;
  const double bn = std::sqrt(dtmp.real());
  if (bn == 0.0) {
    HIPCHK(hipMemsetAsync(x, 0, n * sizeof(zd), h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    if (info) *info = spk_gmres_info{0, 0, 0.0, 0.0, 0, (now_ms() - t_start) / 1e3, 0.0};
    return SPK_OK;
  }
  double sPb = bn;
  if (prec) {
    precond(b, r0);
    if ((rc_ = dev_dot(h, r0, r0, n, &dtmp))) return rc_;
    sPb = std::sqrt(dtmp.real());
  }
  if ((rc_ = matvec_dev(h, x, r0))) return rc_;
  ++matvecs;
  launch_zsub(n, b, r0, r0, h->st);
  precond(r0, w);
  if ((rc_ = dev_dot(h, w, w, n, &dtmp))) return rc_;
  double beta = std::sqrt(dtmp.real());
  if (beta == 0.0) {
    if (info) *info = spk_gmres_info{0, 0, 0.0, 0.0, matvecs, (now_ms() - t_start) / 1e3, 0.0};
    return SPK_OK;
  }
  int iterOut = 0;
  bool done = false;
  while (!done) {
    launch_zscal_copy(n, zd{1.0 / beta, 0.0}, w, V, h->st);
    HH(0, m) = beta;
    for (int j = 1; j <= m; ++j) HH(j, m) = 0.0;
    int jH = 1;
    for (;;) {
      if (prec) {
        if ((rc_ = matvec_dev(h, V + (size_t)(jH - 1) * n, r0))) return rc_;
        precond(r0, w);
      } else if ((rc_ = matvec_dev(h, V + (size_t)(jH - 1) * n, w))) {
        return rc_;
      }
      ++matvecs;
      for (int j = 0; j < jH; ++j) HH(j, jH - 1) = 0.0;
      int nOrtho = 0;
      double dnormw = 0.0;
      for (;;) {
        ++nOrtho;
        double dloo = 0.0;
        if (sweepG) {
          // the whole MGS pass and |w|^2 in one cooperative launch
          HIPCHK(launch_mgs_sweep(sweepG, n, jH, V, n, w, h->g_slots.p, h->g_h.p, h->g_h.p + jH,
                                  h->g_ctl.p, ++h->sweep_seq, h->st));
        } else if (one_reduce) {
          zd *Tm = h->g_T.p, *c = h->g_c.p;
          if (nOrtho == 1 && jH > 1) {
            // c = V^H w and g = V^H v_{jH-1} in one read of V; T grows by row jH-1
            launch_gemv_c(n, jH, (const zd *)V, n, (const zd *)w,
                          (const zd *)(V + (size_t)(jH - 1) * n), h->g_partial.p, c, h->st);
            launch_gemv_c((int64_t)(jH - 1), jH - 1, (const zd *)Tm, (int64_t)ldt,
                          (const zd *)(c + jH), (const zd *)nullptr, h->g_pT.p, h->g_tg.p, h->st);
            launch_icwy_row(jH - 1, Tm, ldt, h->g_tg.p, h->st);
          } else {
            launch_gemv_c(n, jH, (const zd *)V, n, (const zd *)w, (const zd *)nullptr,
                          h->g_partial.p, c, h->st);
            if (nOrtho == 1) launch_icwy_row(0, Tm, ldt, h->g_tg.p, h->st);
          }
          launch_gemv_n((int64_t)jH, jH, (const zd *)Tm, (int64_t)ldt, c, h->g_h.p, h->g_pT.p, h->st);
          launch_gemv_n_sub(n, jH, (const zd *)V, n, h->g_h.p, w, h->g_pn.p, h->st);
        } else if (mgs_steps) {
          for (int j = 0; j < jH; ++j)
            launch_mgs_step(n, j ? V + (size_t)(j - 1) * n : nullptr, j ? h->g_h.p + (j - 1) : nullptr,
                            V + (size_t)j * n, w, h->g_partial.p, h->g_h.p + j, h->st);
          launch_zaxpy_dev(n, h->g_h.p + (jH - 1), -1.0, V + (size_t)(jH - 1) * n, w, h->st);
        } else {
          launch_gemv_c(n, jH, (const zd *)V, n, (const zd *)w, (const zd *)nullptr,
                        h->g_partial.p, h->g_h.p, h->st);
          launch_gemv_n_sub(n, jH, (const zd *)V, n, h->g_h.p, w, h->g_pn.p, h->st);
        }
        // |w|^2 lands after the coefficients: one copy to pinned memory per pass
        if (!sweepG) launch_zdotc(n, w, w, partial_nrm, h->g_h.p + jH, h->st);
        const zd *hv = h->pin.p;
        HIPCHK(hipMemcpyAsync(h->pin.p, h->g_h.p, (jH + 1) * sizeof(zd), hipMemcpyDeviceToHost,
                              h->st));
        HIPCHK(hipStreamSynchronize(h->st));
        if (sweepG && hv[jH].y != 0.0) {
          // an exchange wait ran out of polls: the pass is void
          HIPCHK(hipMemsetAsync(h->g_ctl.p, 0, sizeof(unsigned long long), h->st));
          HIPCHK(hipStreamSynchronize(h->st));
          return fail(SPK_ERR_HIP, "Gram-Schmidt sweep: partial-sum exchange timed out");
        }
        for (int j = 0; j < jH; ++j) {
          const zc dVi(hv[j].x, hv[j].y);
          HH(j, jH - 1) += dVi;
          dloo += std::abs(dVi) * std::abs(dVi);
        }
        dnormw = std::sqrt(hv[jH].x);
        if (orth == SPK_ORTH_IMGS || orth == SPK_ORTH_ICGS) {
          dloo = std::sqrt(dloo);
          if (dnormw <= dloo && nOrtho < 3) continue;
        }
        break;
      }
      HH(jH, jH - 1) = dnormw;
      launch_zscal_copy(n, zd{1.0 / dnormw, 0.0}, w, V + (size_t)jH * n, h->st);
      for (int j = 0; j < jH - 1; ++j) zrot1(&HH(j, jH - 1), &HH(j + 1, jH - 1), rc[j], rs[j]);
      zc ahj = HH(jH - 1, jH - 1);
      const zc ahp = HH(jH, jH - 1);
      double cc;
      zc ss;
      zrotg_(&ahj, ahp, &cc, &ss);
      rc[jH - 1] = cc;
      rs[jH - 1] = ss;
      zrot1(&HH(jH - 1, m), &HH(jH, m), cc, ss);
      zrot1(&HH(jH - 1, jH - 1), &HH(jH, jH - 1), cc, ss);
      HH(jH, jH - 1) = 0.0;
      const double dnormres = std::abs(HH(jH, m));
      bea = dnormres / (0.0 * 1.0 + sPb);
      it_total = iterOut * m + jH;
      h->hist.push_back(bea);
      if (bea <= tol || it_total >= iterMax) {
        compRsd = 1;
        for (int j = 0; j < jH; ++j) yc[j] = HH(j, m);
        for (int j = jH - 1; j >= 0; --j)
          if (yc[j] != 0.0) {
            yc[j] = yc[j] / HH(j, j);
            const zc tq = yc[j];
            for (int i = j - 1; i >= 0; --i) yc[i] = yc[i] - tq * HH(i, j);
          }
        std::vector<zd> yd(jH);
        for (int j = 0; j < jH; ++j) yd[j] = zd{yc[j].real(), yc[j].imag()};
        HIPCHK(hipMemcpyAsync(h->g_y.p, yd.data(), jH * sizeof(zd), hipMemcpyHostToDevice, h->st));
        launch_gemv_n(n, jH, (const zd *)V, n, h->g_y.p, r0, h->g_pn.p, h->st);
        HIPCHK(hipMemcpyAsync(xc, x, n * sizeof(zd), hipMemcpyDeviceToDevice, h->st));
        launch_zadd(n, r0, xc, h->st);
        if ((rc_ = matvec_dev(h, xc, w))) return rc_;
        ++matvecs;
        launch_zsub(n, b, w, r0, h->st);
        if ((rc_ = dev_dot(h, r0, r0, n, &dtmp))) return rc_;
        const double true_res = std::sqrt(dtmp.real());
        if (prec) {
          precond(r0, w);
          if ((rc_ = dev_dot(h, w, w, n, &dtmp))) return rc_;
          be = std::sqrt(dtmp.real()) / sPb;
        } else {
          be = true_res / sPb;
        }
        tbe = true_res / bn;
        if (be <= tol || it_total >= iterMax) {
          HIPCHK(hipMemcpyAsync(x, xc, n * sizeof(zd), hipMemcpyDeviceToDevice, h->st));
          info1 = (be <= tol) ? 0 : -4;
          done = true;
          break;
        }
      }
      ++jH;
      if (jH <= m) continue;
      break;
    }
    if (done) break;
    ++iterOut;
    --jH;
    if (bea > tol) {
      for (int j = 0; j < jH; ++j) yc[j] = HH(j, m);
      for (int j = jH - 1; j >= 0; --j)
        if (yc[j] != 0.0) {
          yc[j] = yc[j] / HH(j, j);
          const zc tq = yc[j];
          for (int i = j - 1; i >= 0; --i) yc[i] = yc[i] - tq * HH(i, j);
        }
      std::vector<zd> yd(jH);
      for (int j = 0; j < jH; ++j) yd[j] = zd{yc[j].real(), yc[j].imag()};
      HIPCHK(hipMemcpyAsync(h->g_y.p, yd.data(), jH * sizeof(zd), hipMemcpyHostToDevice, h->st));
      launch_gemv_n(n, jH, (const zd *)V, n, h->g_y.p, r0, h->g_pn.p, h->st);
      HIPCHK(hipMemcpyAsync(xc, x, n * sizeof(zd), hipMemcpyDeviceToDevice, h->st));
      launch_zadd(n, r0, xc, h->st);
    }
    HIPCHK(hipMemcpyAsync(x, xc, n * sizeof(zd), hipMemcpyDeviceToDevice, h->st));
    if (compRsd) {
      if ((rc_ = matvec_dev(h, x, r0))) return rc_;
      ++matvecs;
      launch_zsub(n, b, r0, r0, h->st);
      precond(r0, w);
      if ((rc_ = dev_dot(h, w, w, n, &dtmp))) return rc_;
      beta = std::sqrt(dtmp.real());
    } else {
      beta = std::abs(HH(m, m));
      for (int j = m - 1; j >= 0; --j) {
        HH(j, m) = 0.0;
        zrot1(&HH(j, m), &HH(j + 1, m), rc[j], -rs[j]);
      }
      std::vector<zd> hd(m + 1);
      for (int j = 0; j <= m; ++j) hd[j] = zd{HH(j, m).real(), HH(j, m).imag()};
      HIPCHK(hipMemcpyAsync(h->g_y.p, hd.data(), (m + 1) * sizeof(zd), hipMemcpyHostToDevice,
                            h->st));
      launch_gemv_n(n, m + 1, (const zd *)V, n, h->g_y.p, w, h->g_pn.p, h->st);
    }
  }
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipGetLastError());
  if (info) {
    info->info = info1;
    info->iterations = it_total;
    info->arnoldi_backward_error = bea;
    info->backward_error = be;
    info->matvecs = matvecs;
    info->seconds = (now_ms() - t_start) / 1e3;
    info->true_backward_error = tbe;
  }
  return SPK_OK;
}

int spk_gmres(spk_handle *h, const double *b, double *x, const spk_gmres_params *p,
              spk_gmres_info *info) {
  if (!h || !b || !x || !p) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  const size_t bytes = 2 * (size_t)h->N * sizeof(zd);
  DBuf<zd> db, dx;
  HIPCHK(db.alloc(2 * (size_t)h->N));
  HIPCHK(dx.alloc(2 * (size_t)h->N));
  HIPCHK(hipMemcpyAsync(db.p, b, bytes, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(dx.p, x, bytes, hipMemcpyHostToDevice, h->st));
  int rc = spk_gmres_device(h, db.p, dx.p, p, info);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(x, dx.p, bytes, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return SPK_OK;
}

// ---------------------------------------------------------------------------
// Adjoint gradient of J = sum_p w_p |E(r_p)|^2 (the scattered field of the
// plane-wave solution I of Z I = b) with respect to the heights (z) of mesh
// nodes.  dJ = 2 Re[E^H W dF I + nu^T (db - dZ I)] with Z^T nu = F^T W conj(E):
// one forward solve, one transposed solve, then for every requested node the
// central differences of the local quantities it moves (spk_kernels.hip
// k_adj_*), contracted at once.  No counterpart in the reference.
// ---------------------------------------------------------------------------
int spk_adjoint_height_gradient(spk_handle *h, const double k_in[3], const double e_in[3],
                                int npts, const double *r, const double *w,
                                const spk_gmres_params *gp, int n_nodes, const int *nodes,
                                double dz, double *J, double *grad, double *terms,
                                spk_gmres_info *info) {
  if (!h || !k_in || !e_in || !r || !w || !gp || !J || npts < 1 || n_nodes < 0 ||
      (n_nodes > 0 && (!nodes || !grad)) || !(dz > 0))
    return fail(SPK_ERR_ARG, "bad argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (h->quad) return fail(SPK_ERR_STATE, "the adjoint gradient runs on triangle meshes");
  if (h->transposed) return fail(SPK_ERR_STATE, "the handle holds the transposed operator");
  const int np = (int)(h->pp_host.size() / 3);
  for (int i = 0; i < n_nodes; ++i)
    if (nodes[i] < 0 || nodes[i] >= np) return fail(SPK_ERR_ARG, "node index out of range");
  HIPCHK(hipSetDevice(h->p.device));
  const int N = h->N;
  const int64_t n = 2 * (int64_t)N;
  hipStream_t st = h->st;
  int rc = SPK_OK;
  DBuf<zd> b, I, nu, c, E, H, ytmp;
  DBuf<double> dr, dw;
  HIPCHK(b.alloc(n));
  HIPCHK(I.alloc(n));
  HIPCHK(nu.alloc(n));
  HIPCHK(c.alloc(n));
  HIPCHK(ytmp.alloc(n));
  HIPCHK(E.alloc(3 * (size_t)npts));
  HIPCHK(H.alloc(3 * (size_t)npts));
  HIPCHK(dr.upload(std::vector<double>(r, r + 3 * (size_t)npts)));
  HIPCHK(dw.upload(std::vector<double>(w, w + (size_t)npts)));
  spk_gmres_info inf[2] = {};
  // forward solve and its field
  if ((rc = rhs_plane_dev(h, k_in, e_in, b.p))) return rc;
  if ((rc = spk_gmres_device(h, b.p, I.p, gp, &inf[0]))) return rc;
  const ScatArgs sa = scat_args(h, npts);
  launch_scattered(sa, h->d_fn.p, I.p, dr.p, E.p, H.p, st);
  HIPCHK(hipGetLastError());
  std::vector<zd> Eh(3 * (size_t)npts);
  HIPCHK(hipMemcpyAsync(Eh.data(), E.p, Eh.size() * sizeof(zd), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double Jv = 0.0;
  for (int p = 0; p < npts; ++p)
    for (int k = 0; k < 3; ++k)
      Jv += w[p] * (Eh[3 * p + k].x * Eh[3 * p + k].x + Eh[3 * p + k].y * Eh[3 * p + k].y);
  *J = Jv;
  // the adjoint's right-hand side and the leaf field W of I
  launch_scattered_t(sa, h->d_fn.p, dr.p, dw.p, E.p, c.p, st);
  HIPCHK(hipGetLastError());
  if ((rc = matvec_t<zd>(h, I.p, ytmp.p))) return rc;
  const Tree &t = h->tree;
  const size_t S = h->tb.S, nleaf = h->lev[t.l_max].nb;
  DBuf<zd> WI, Bt;
  HIPCHK(WI.alloc(8 * S * nleaf));
  HIPCHK(Bt.alloc(8 * S * nleaf));
  HIPCHK(hipMemcpyAsync(WI.p, h->d_W.p, 8 * S * nleaf * sizeof(zd), hipMemcpyDeviceToDevice, st));
  // the transposed solve and its leaf sums
  if ((rc = spk_set_transposed(h, 1))) return rc;
  rc = spk_gmres_device(h, c.p, nu.p, gp, &inf[1]);
  if (!rc) rc = matvec_t<zd>(h, nu.p, ytmp.p);
  if (!rc) {
    const hipError_t e = hipMemcpyAsync(Bt.p, h->lev[t.l_max].B.p, 8 * S * nleaf * sizeof(zd),
                                        hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) rc = fail(SPK_ERR_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
  }
  const int rc_back = spk_set_transposed(h, 0);
  if (rc) return rc;
  if (rc_back) return rc_back;
  if (info) {
    info[0] = inf[0];
    info[1] = inf[1];
  }
  if (n_nodes == 0) return SPK_OK;
  // the edges each requested node moves (either side's triangle holds it)
  std::vector<std::vector<int>> node_edges(np);
  for (int m = 0; m < N; ++m)
    for (int q = 0; q < 2; ++q)
      for (int v = 0; v < 3; ++v) {
        auto &lst = node_edges[h->er[2 * m + q].node[v] - 1];
        if (lst.empty() || lst.back() != m) lst.push_back(m);
      }
  std::vector<int> aoff(n_nodes + 1, 0), aedge, anode;
  std::vector<int64_t> toff;
  for (int k = 0; k < n_nodes; ++k) {
    std::vector<int> lst = node_edges[nodes[k]];
    std::sort(lst.begin(), lst.end());
    lst.erase(std::unique(lst.begin(), lst.end()), lst.end());
    int64_t acc = 0;
    for (int m : lst) {
      aedge.push_back(m);
      anode.push_back(k);
      toff.push_back(acc);
      const int bx = t.leaf_box[m];
      acc += 2 * (int64_t)(t.e2_off[bx + 1] - t.e2_off[bx]);
    }
    toff.push_back(acc);  // the node's total (its last slot's end)
    aoff[k + 1] = (int)aedge.size();
  }
  // task_off is indexed per slot with one extra entry per node: rebase so
  // that node k's slots start at aoff[k] + k
  std::vector<int> aoff_t(n_nodes + 1);
  for (int k = 0; k <= n_nodes; ++k) aoff_t[k] = aoff[k] + k;
  // per-edge side nodes (p1, p2, free vertex, element nodes), 0-based
  std::vector<int> sn(12 * (size_t)N);
  for (int m = 0; m < N; ++m)
    for (int q = 0; q < 2; ++q) {
      const int *cr = h->er[2 * m + q].node;
      int nv = 0, ipc = 0;
      for (int j = 0; j < 3; ++j)
        if (cr[j] != h->corner[2 * m] && cr[j] != h->corner[2 * m + 1]) {
          nv = j + 1;
          ipc = cr[j];
        }
      int i1, i2;
      if (nv == 1) {
        i1 = cr[1];
        i2 = cr[2];
      } else if (nv == 2) {
        i1 = cr[2];
        i2 = cr[0];
      } else {
        i1 = cr[0];
        i2 = cr[1];
      }
      int *o = &sn[6 * (2 * (size_t)m + q)];
      o[0] = i1 - 1;
      o[1] = i2 - 1;
      o[2] = ipc - 1;
      o[3] = cr[0] - 1;
      o[4] = cr[1] - 1;
      o[5] = cr[2] - 1;
    }
  DBuf<double> d_pp;
  DBuf<int> d_sn, d_nodes, d_aoff, d_aedge, d_anode;
  DBuf<int64_t> d_toff;
  HIPCHK(d_pp.upload(h->pp_host));
  HIPCHK(d_sn.upload(sn));
  HIPCHK(d_nodes.upload(std::vector<int>(nodes, nodes + n_nodes)));
  HIPCHK(d_aoff.upload(aoff_t));
  HIPCHK(d_aedge.upload(aedge));
  HIPCHK(d_anode.upload(anode));
  HIPCHK(d_toff.upload(toff));
  const int nslot = (int)aedge.size();
  AdjArgs G{};
  G.N = N;
  G.nnode = n_nodes;
  G.nslot = nslot;
  G.normal_convention = h->p.normal_convention;
  G.dz = dz;
  G.pp = d_pp.p;
  G.side_nodes = d_sn.p;
  G.leaf_box = h->d_leaf_box.p;
  G.node = d_nodes.p;
  // the near kernel reads slots through aff_off (rebased, with the per-node
  // end entries of task_off); the per-slot kernels read aff_edge / aff_node
  G.aff_off = d_aoff.p;
  G.aff_node = d_anode.p;
  G.task_off = d_toff.p;
  // aff_edge as the near kernel sees it: one placeholder entry per node end
  std::vector<int> aedge_t;
  for (int k = 0; k < n_nodes; ++k) {
    for (int q = aoff[k]; q < aoff[k + 1]; ++q) aedge_t.push_back(aedge[q]);
    aedge_t.push_back(-1);
  }
  DBuf<int> d_aedge_t;
  HIPCHK(d_aedge_t.upload(aedge_t));
  const NearArgs A = near_args(h);
  constexpr int kGroups = 64;
  DBuf<zd> near_part, far_out, rhs_out, fld_out;
  HIPCHK(near_part.alloc((size_t)n_nodes * kGroups));
  HIPCHK(far_out.alloc(std::max(nslot, 1)));
  HIPCHK(rhs_out.alloc(std::max(nslot, 1)));
  HIPCHK(fld_out.alloc(std::max(nslot, 1)));
  {
    AdjArgs Gn = G;
    Gn.aff_edge = d_aedge_t.p;
    launch_adj_near(A, Gn, kGroups, nu.p, I.p, near_part.p, st);
  }
  G.aff_edge = d_aedge.p;
  AdjFar F{};
  F.S = (int)S;
  F.ngp = h->ngp;
  F.khat = h->d_khat.p;
  F.tth = h->d_tth.p;
  F.tph = h->d_tph.p;
  F.leaf_centre = h->d_leaf_centre.p;
  for (int i = 0; i < 12; ++i) {
    F.qa[i] = h->qa[i];
    F.qb[i] = h->qb[i];
    F.qw[i] = h->qw[i];
  }
  F.k1 = zcx(h->mat.k1);
  F.k2 = zcx(h->mat.k2);
  F.sc1 = h->sc1;
  F.sc2 = h->sc2;
  launch_adj_far(G, F, nu.p, I.p, WI.p, Bt.p, far_out.p, st);
  launch_adj_rhs(G, rhs_args(h, k_in, e_in), nu.p, rhs_out.p, st);
  launch_adj_field(G, sa, dr.p, dw.p, E.p, I.p, fld_out.p, st);
  HIPCHK(hipGetLastError());
  std::vector<zd> hn((size_t)n_nodes * kGroups), hf(std::max(nslot, 1)), hr(std::max(nslot, 1)),
      hq(std::max(nslot, 1));
  HIPCHK(hipMemcpyAsync(hn.data(), near_part.p, hn.size() * sizeof(zd), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(hf.data(), far_out.p, hf.size() * sizeof(zd), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(hr.data(), rhs_out.p, hr.size() * sizeof(zd), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(hq.data(), fld_out.p, hq.size() * sizeof(zd), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int k = 0; k < n_nodes; ++k) {
    zc tz(0.0, 0.0), tf(0.0, 0.0);
    for (int g = 0; g < kGroups; ++g) tz += zc(hn[(size_t)k * kGroups + g].x, hn[(size_t)k * kGroups + g].y);
    for (int q = aoff[k]; q < aoff[k + 1]; ++q) {
      tz += zc(hf[q].x, hf[q].y);
      tf += zc(hq[q].x, hq[q].y) + zc(hr[q].x, hr[q].y);
    }
    grad[k] = 2.0 * (tf - tz).real() / (2.0 * dz);
    if (terms) {
      zc tn(0.0, 0.0), tfar(0.0, 0.0), tfld(0.0, 0.0), trhs(0.0, 0.0);
      for (int g = 0; g < kGroups; ++g) tn += zc(hn[(size_t)k * kGroups + g].x, hn[(size_t)k * kGroups + g].y);
      for (int q = aoff[k]; q < aoff[k + 1]; ++q) {
        tfar += zc(hf[q].x, hf[q].y);
        tfld += zc(hq[q].x, hq[q].y);
        trhs += zc(hr[q].x, hr[q].y);
      }
      terms[4 * k + 0] = tfld.real() / dz;
      terms[4 * k + 1] = trhs.real() / dz;
      terms[4 * k + 2] = -tn.real() / dz;
      terms[4 * k + 3] = -tfar.real() / dz;
    }
  }
  return SPK_OK;
}

int spk_gmres_history(const spk_handle *h, double *hist, int cap, int *count) {
  if (!h || !count) return fail(SPK_ERR_ARG, "null argument");
  *count = (int)h->hist.size();
  if (hist)
    for (int i = 0; i < std::min(cap, *count); ++i) hist[i] = h->hist[i];
  return SPK_OK;
}

int spk_get_rwg(const spk_handle *h, int *corner, int *element) {
  if (!h) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (h->quad) return fail(SPK_ERR_STATE, "quadrilateral handles have no RWG pairs");
  if (corner) std::memcpy(corner, h->corner.data(), h->corner.size() * sizeof(int));
  if (element) std::memcpy(element, h->element.data(), h->element.size() * sizeof(int));
  return SPK_OK;
}

int spk_get_tree(const spk_handle *h, int64_t *leaf, int *l_min, int *l_max, double *D,
                 double *xmin3) {
  if (!h) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  const Tree &t = h->tree;
  if (leaf) std::memcpy(leaf, t.leaf.data(), t.leaf.size() * sizeof(int64_t));
  if (l_min) *l_min = t.l_min;
  if (l_max) *l_max = t.l_max;
  if (D) *D = t.D;
  if (xmin3) std::memcpy(xmin3, t.xmin, 3 * sizeof(double));
  return SPK_OK;
}

int spk_get_i4(const spk_handle *h, int level, int64_t n, int64_t *out, int cap, int *count) {
  if (!h || !count) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  const Tree &t = h->tree;
  if (level < t.l_min || level > t.l_max) return fail(SPK_ERR_ARG, "level outside [l_min, l_max]");
  const Level &L = t.lev[level];
  auto it = std::lower_bound(L.n.begin(), L.n.end(), n);
  if (it == L.n.end() || *it != n) return fail(SPK_ERR_ARG, "box is empty");
  const int a = (int)(it - L.n.begin());
  std::vector<int64_t> v;
  for (int q = L.i4_off[a]; q < L.i4_off[a + 1]; ++q) v.push_back(L.n[L.i4_box[q]]);
  std::sort(v.begin(), v.end());
  *count = (int)v.size();
  if (out)
    for (int i = 0; i < std::min(cap, *count); ++i) out[i] = v[i];
  return SPK_OK;
}

int spk_get_near_csr(const spk_handle *h, int64_t *ia, int32_t *ja, double *abcd) {
  if (!h) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  const Tree &t = h->tree;
  std::vector<zd> blocks;
  if (abcd) {
    blocks.resize(4 * (size_t)t.nnz);
    HIPCHK(hipMemcpy(blocks.data(), h->d_blocks.p, blocks.size() * sizeof(zd), hipMemcpyDeviceToHost));
  }
  int64_t off = 0;
  for (int m = 0; m < h->N; ++m) {
    const int b = t.leaf_box[m];
    const int cols = t.e2_off[b + 1] - t.e2_off[b];
    const int r = t.pos[m] - t.box_first[b];
    if (ia) ia[m] = off;
    if (ja) std::memcpy(ja + off, &t.e2_list[t.e2_off[b]], cols * sizeof(int));
    if (abcd) {
      const zd *src = &blocks[4 * (size_t)(t.near_off[b] + (int64_t)r * cols)];
      std::memcpy(abcd + 8 * off, src, 4 * (size_t)cols * sizeof(zd));
    }
    off += cols;
  }
  if (ia) ia[h->N] = off;
  return SPK_OK;
}

int spk_near_blocks(spk_handle *h, int npairs, const int *m, const int *n, double *out) {
  if (!h || !m || !n || !out || npairs < 0) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready || h->N <= 0 || !h->d_fn.p) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (h->quad) return fail(SPK_ERR_STATE, "quadrilateral handles: use spk_get_near_csr");
  for (int i = 0; i < npairs; ++i)
    if (m[i] < 0 || m[i] >= h->N || n[i] < 0 || n[i] >= h->N)
      return fail(SPK_ERR_ARG, "edge index out of range");
  DBuf<int> dm, dn;
  DBuf<zd> dout;
  HIPCHK(dm.upload(std::vector<int>(m, m + npairs)));
  HIPCHK(dn.upload(std::vector<int>(n, n + npairs)));
  HIPCHK(dout.alloc(4 * (size_t)std::max(npairs, 1)));
  NearArgs a = near_args(h);
  launch_near_pairs(a, npairs, dm.p, dn.p, dout.p, h->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->st));
  if (npairs) HIPCHK(hipMemcpy(out, dout.p, 4 * (size_t)npairs * sizeof(zd), hipMemcpyDeviceToHost));
  return SPK_OK;
}

int spk_get_patterns(const spk_handle *h, int first, int count, double *out) {
  if (!h || !out) return fail(SPK_ERR_ARG, "null argument");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (first < 0 || count < 0 || first + count > h->N) return fail(SPK_ERR_ARG, "edge range");
  if (h->quad) return fail(SPK_ERR_STATE, "quadrilateral handles: patterns have their own layout");
  const int S = h->tb.S;
  for (int e = 0; e < count; ++e) {
    const int p = h->tree.pos[first + e];
    HIPCHK(hipMemcpy(out + (size_t)e * 8 * S, h->d_rt.p + (size_t)p * 4 * S, 4 * (size_t)S * sizeof(zd),
                     hipMemcpyDeviceToHost));
  }
  return SPK_OK;
}

extern "C++" {
template <class T>
static int profile_matvec_t(spk_handle *h, int reps, double *ms, double *bytes) {
  hipEvent_t ev[8];
  for (int i = 0; i < 8; ++i) HIPCHK(hipEventCreate(&ev[i]));
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int rc = SPK_OK;
  T *x = nullptr, *y = nullptr;
  if constexpr (std::is_same<T, zd>::value) {
    x = h->d_x.p;
    y = h->d_y.p;
  } else {
    x = h->f_x.p;
    y = h->f_y.p;
  }
  for (int r = 0; r < reps && rc == SPK_OK; ++r) {
    rc = matvec_t<T>(h, x, y, true, true, ev);
    if (rc) break;
    HIPCHK(hipEventSynchronize(ev[7]));
    float t;
    for (int i = 0; i < 7; ++i) {
      HIPCHK(hipEventElapsedTime(&t, ev[i], ev[i + 1]));
      acc[i] += t;
    }
    HIPCHK(hipEventElapsedTime(&t, ev[0], ev[7]));
    acc[7] += t;
  }
  for (int i = 0; i < 8; ++i) (void)hipEventDestroy(ev[i]);
  if (rc) return rc;
  if (ms)
    for (int i = 0; i < 8; ++i) ms[i] = acc[i] / reps;
  if (bytes) stage_bytes(h, bytes, (double)sizeof(T));
  return SPK_OK;
}
}  // extern "C++"

int spk_probe_read_bandwidth(int device, size_t bytes, int reps, double *gbps) {
  if (!gbps || bytes < (1u << 20) || reps < 1) return fail(SPK_ERR_ARG, "bad probe arguments");
  HIPCHK(hipSetDevice(device));
  DBuf<double2> buf;
  DBuf<double> out;
  HIPCHK(buf.alloc(bytes / 16));
  HIPCHK(out.alloc(1));
  HIPCHK(hipMemset(buf.p, 0, (bytes / 16) * 16));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  float best = 0.f;
  for (int r = 0; r <= reps; ++r) {
    (void)hipEventRecord(e0, nullptr);
    launch_stream_read(buf.p, (bytes / 16) * 16, out.p, nullptr);
    (void)hipEventRecord(e1, nullptr);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    if (r > 0 && (best == 0.f || ms < best)) best = ms;  // the first launch warms up
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  *gbps = (double)((bytes / 16) * 16) / (best * 1e-3) / 1e9;
  return SPK_OK;
}

int spk_profile_matvec(spk_handle *h, int reps, double *ms, double *bytes) {
  if (!h || reps < 1) return fail(SPK_ERR_ARG, "bad arguments");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (h->multik) return fail(SPK_ERR_STATE, "stage profile: uniform truncation numbers only");
  return profile_matvec_t<zd>(h, reps, ms, bytes);
}

int spk_profile_matvec_c64(spk_handle *h, int reps, double *ms, double *bytes) {
  if (!h || reps < 1) return fail(SPK_ERR_ARG, "bad arguments");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (!h->mixed_ready) return fail(SPK_ERR_STATE, "handle was not created with SPK_MIXED");
  return profile_matvec_t<zf>(h, reps, ms, bytes);
}

} // extern "C"
// times the preconditioner's near SpMV on the bf16 (f8 false) or fp8 copy
static int profile_near_lp(spk_handle *h, int reps, double *ms, double *bytes, bool f8) {
  if (!h || reps < 1 || !ms || !bytes) return fail(SPK_ERR_ARG, "bad arguments");
  if (!h->ready) return fail(SPK_ERR_STATE, "spk_setup_tri has not completed");
  if (!h->mixed_ready) return fail(SPK_ERR_STATE, "handle was not created with SPK_MIXED");
  const int rc = f8 ? ensure_near_f8(h) : ensure_near_bf(h);
  if (rc) return rc;
  const int64_t n = 2 * (int64_t)h->N;
  DBuf<zf> x, y;
  HIPCHK(x.alloc(n));
  HIPCHK(y.alloc(n));
  HIPCHK(hipMemsetAsync(x.p, 0, n * sizeof(zf), h->st));
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  auto run = [&]() {
    if (f8)
      launch_near_spmv_f8(h->spmv_wgs, h->N, h->d_row_box.p, h->d_row_r0.p, h->d_box_first.p,
                          h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p,
                          h->d_perm.p, h->d_blk_f8.p, h->d_f8_scale.p, x.p, y.p, h->st);
    else
      launch_near_spmv_bf(h->spmv_wgs, h->N, h->d_row_box.p, h->d_row_r0.p, h->d_box_first.p,
                          h->d_box_count.p, h->d_e2_off.p, h->d_e2_list.p, h->d_near_off.p,
                          h->d_perm.p, h->d_blk_bf.p, x.p, y.p, h->st);
  };
  run();
  HIPCHK(hipEventRecord(e0, h->st));
  for (int r = 0; r < reps; ++r) run();
  HIPCHK(hipEventRecord(e1, h->st));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)t / reps;
  const double nblocks = (double)(h->f_blocks.n / 4);
  *bytes = f8 ? nblocks * 8.0 + 16.0 * (double)h->N + 3.0 * (double)n * 8.0
              : nblocks * 16.0 + 3.0 * (double)n * 8.0;
  return SPK_OK;
}
extern "C" {

int spk_profile_near_bf(spk_handle *h, int reps, double *ms, double *bytes) {
  return profile_near_lp(h, reps, ms, bytes, false);
}

int spk_profile_near_f8(spk_handle *h, int reps, double *ms, double *bytes) {
  return profile_near_lp(h, reps, ms, bytes, true);
}

// The interpolation between two sample grids as this library builds it
// (interp_pair_csr), applied on the host to one current's four channels
// (theta k1, theta k2, phi k1, phi k2; 2K^2 + 2 complex samples each).
int spk_interp_probe(int Km, int Kn, const double *in, double *out) {
  if (!in || !out || Km < 4 || Kn < 4 || Km > 120 || Kn > 120)
    return fail(SPK_ERR_ARG, "bad arguments");
  std::vector<int> r, c;
  std::vector<double> v;
  interp_pair_csr(3, Km, Kn, &r, &c, &v);
  const int Sm = 2 * Km * Km + 2, Sn = 2 * Kn * Kn + 2;
  for (int k = 0; k < 2; ++k) {  // medium: channels (k, 2 + k)
    const int ch[2] = {k, 2 + k};
    for (int row = 0; row < 2 * Sn; ++row) {
      double re = 0, im = 0;
      for (int e = r[row]; e < r[row + 1]; ++e) {
        const int ci = c[e] < Sm ? 0 : 1, si = c[e] - ci * Sm;
        re += in[2 * ((size_t)ch[ci] * Sm + si)] * v[e];
        im += in[2 * ((size_t)ch[ci] * Sm + si) + 1] * v[e];
      }
      const int co = row < Sn ? 0 : 1, so = row - co * Sn;
      out[2 * ((size_t)ch[co] * Sn + so)] = re;
      out[2 * ((size_t)ch[co] * Sn + so) + 1] = im;
    }
  }
  return SPK_OK;
}

int spk_topology_probe(int np, const double *pp, int nt, const int *tt, int normal_convention,
                       int tree_s_opt, int64_t *n_edges, int *corner, int *element, int64_t *leaf,
                       int *l_min, int *l_max, double *D, double *xmin3) {
  if (!pp || !tt || !n_edges || np < 3 || nt < 1) return fail(SPK_ERR_ARG, "bad mesh arguments");
  std::vector<int> c, e;
  std::vector<FnRec> fn;
  std::vector<ElemRec> er;
  std::vector<double> rr;
  std::string err;
  int rc = build_rwg(np, pp, nt, tt, normal_convention == 2 ? 2 : 1, &c, &e, &fn, &er, &rr, &err);
  if (rc != SPK_OK) return fail(rc, err);
  *n_edges = (int64_t)rr.size();
  if (corner) std::memcpy(corner, c.data(), c.size() * sizeof(int));
  if (element) std::memcpy(element, e.data(), e.size() * sizeof(int));
  if (leaf || l_min || l_max || D || xmin3) {
    if (rr.empty()) return fail(SPK_ERR_MESH, "mesh has no interior edges");
    Tree t;
    build_tree(fn, (int)rr.size(), tree_s_opt > 0 ? tree_s_opt : 250, &t);
    if (leaf) std::memcpy(leaf, t.leaf.data(), t.leaf.size() * sizeof(int64_t));
    if (l_min) *l_min = t.l_min;
    if (l_max) *l_max = t.l_max;
    if (D) *D = t.D;
    if (xmin3) std::memcpy(xmin3, t.xmin, 3 * sizeof(double));
  }
  return SPK_OK;
}

}  // extern "C"

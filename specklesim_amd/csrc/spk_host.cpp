// spk_host.cpp -- host-side setup of the SIE solve loop: RWG topology,
// octree and interaction lists, k-space sampling tables, M2L transfer
// tables and SS/RR phase tables.  Everything here runs once per geometry;
// the per-matvec work is in spk_kernels.hip.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <unordered_map>

#include "spk_internal.h"

namespace spk {

using zc = std::complex<double>;

static inline zc Z(const cplx &a) { return zc(a.re, a.im); }
static inline cplx C(const zc &a) { return cplx{a.real(), a.imag()}; }

// set_local_material_parameters_tri (src/lib_sie_tri_data_container.f90:212-224)
// and formulation_coefficient (src/lib_sie_tri_calculation_mod.f90:368-387).
void make_material(const spk_params &p, Material *m) {
  m->k0 = 2 * kPi / p.lambda;
  m->omega = 2 * kPi * kC0 / p.lambda;
  m->my1 = kMy0;
  m->my2 = kMy0;
  zc er1(p.eps_r1_re, p.eps_r1_im), er2(p.eps_r2_re, p.eps_r2_im);
  zc k1 = m->k0 * std::sqrt(er1), k2 = m->k0 * std::sqrt(er2);
  zc e1 = kEps0 * er1, e2 = kEps0 * er2;
  zc eta1 = std::sqrt(m->my1 / e1), eta2 = std::sqrt(m->my2 / e2);
  m->k1 = C(k1);
  m->k2 = C(k2);
  m->eps1 = C(e1);
  m->eps2 = C(e2);
  m->eta1 = C(eta1);
  m->eta2 = C(eta2);
  m->eta_a = C((eta1 + eta2) / 2.0);
  double x = 1.0, y = 0.5;
  m->a_lc[0] = C(2 * x * eta1 / (eta1 + y * eta2));
  m->a_lc[1] = C(2 * x * eta2 / (eta1 + y * eta2));
  m->d_lc[0] = C((eta1 + y * eta2) / (2 * x * eta1));
  m->d_lc[1] = C((eta1 + y * eta2) / (2 * x * eta2));
}

// Triangle rules of src/lib_sie_integration_points.f90:7-143.  The
// reference writes them as default-real literals, so each value is the
// float nearest to the printed decimal.
void quadrature_tri(int ngp, double *a, double *b, double *w) {
  for (int i = 0; i < 20; ++i) a[i] = b[i] = w[i] = 0.0;
  auto set = [&](int n, const float *A, const float *B, const float *W) {
    for (int i = 0; i < n; ++i) {
      a[i] = A[i];
      b[i] = B[i];
      w[i] = W[i];
    }
  };
  switch (ngp) {
    case 1: {
      static const float A[] = {0.33333333333333f}, W[] = {1.0f};
      set(1, A, A, W);
    } break;
    case 3: {
      static const float A[] = {0.16666666666667f, 0.16666666666667f, 0.66666666666667f};
      static const float B[] = {0.16666666666667f, 0.66666666666667f, 0.16666666666667f};
      static const float W[] = {0.33333333333333f, 0.33333333333333f, 0.33333333333333f};
      set(3, A, B, W);
    } break;
    case 4: {
      static const float A[] = {0.33333333333333f, 0.2f, 0.2f, 0.6f};
      static const float B[] = {0.33333333333333f, 0.2f, 0.6f, 0.2f};
      static const float W[] = {-0.5625f, 0.52083333333333f, 0.52083333333333f,
                                0.52083333333333f};
      set(4, A, B, W);
    } break;
    case 6: {
      static const float A[] = {0.44594849091597f, 0.44594849091597f, 0.10810301816807f,
                                0.09157621350977f, 0.09157621350977f, 0.81684757298046f};
      static const float B[] = {0.44594849091597f, 0.10810301816807f, 0.44594849091597f,
                                0.09157621350977f, 0.81684757298046f, 0.09157621350977f};
      static const float W[] = {0.22338158967801f, 0.22338158967801f, 0.22338158967801f,
                                0.10995174365532f, 0.10995174365532f, 0.10995174365532f};
      set(6, A, B, W);
    } break;
    case 7: {
      static const float A[] = {0.33333333333333f, 0.47014206410511f, 0.47014206410511f,
                                0.05971587178977f, 0.10128650732346f, 0.10128650732346f,
                                0.79742698535309f};
      static const float B[] = {0.33333333333333f, 0.47014206410511f, 0.05971587178977f,
                                0.47014206410511f, 0.10128650732346f, 0.79742698535309f,
                                0.10128650732346f};
      static const float W[] = {0.225f, 0.13239415278851f, 0.13239415278851f,
                                0.13239415278851f, 0.12593918054483f, 0.12593918054483f,
                                0.12593918054483f};
      set(7, A, B, W);
    } break;
    case 12: {
      static const float A[] = {0.24928674517091f, 0.24928674517091f, 0.50142650965818f,
                                0.0630890144915f,  0.0630890144915f,  0.873821971017f,
                                0.31035245103378f, 0.6365024991214f,  0.05314504984482f,
                                0.6365024991214f,  0.31035245103378f, 0.05314504984482f};
      static const float B[] = {0.24928674517091f, 0.50142650965818f, 0.24928674517091f,
                                0.0630890144915f,  0.873821971017f,   0.0630890144915f,
                                0.6365024991214f,  0.05314504984482f, 0.31035245103378f,
                                0.31035245103378f, 0.05314504984482f, 0.6365024991214f};
      static const float W[] = {0.11678627572638f, 0.11678627572638f, 0.11678627572638f,
                                0.05084490637021f, 0.05084490637021f, 0.05084490637021f,
                                0.08285107561837f, 0.08285107561837f, 0.08285107561837f,
                                0.08285107561837f, 0.08285107561837f, 0.08285107561837f};
      set(12, A, B, W);
    } break;
    default:
      break;
  }
}

// ---------------------------------------------------------------------------
// RWG topology: find_neighbours/common_2 (src/lib_sie_tri_calculation_mod.f90:865-928),
// fn_parameter_type (:1539-1586), norm_vec_r_new (:1006-1024).
// An edge hash replaces the O(nt^2) pair scan; the pair order (i<j by i,
// then j) and the corner order (element i's) are those of the scan.
// ---------------------------------------------------------------------------
int build_rwg(int np, const double *pp, int nt, const int *tt, int normal_convention,
              std::vector<int> *corner, std::vector<int> *element, std::vector<FnRec> *fn,
              std::vector<ElemRec> *er, std::vector<double> *rref, std::string *err) {
  for (int i = 0; i < 3 * nt; ++i)
    if (tt[i] < 1 || tt[i] > np) {
      *err = "triangle node index out of range (tt must be 1-based)";
      return SPK_ERR_MESH;
    }
  // triangles incident to each node
  std::vector<int> deg(np + 1, 0);
  for (int i = 0; i < 3 * nt; ++i) deg[tt[i]]++;
  for (int i = 0; i < np; ++i) deg[i + 1] += deg[i];
  std::vector<int> adj(3 * nt), fill(np, 0);
  for (int i = 0; i < nt; ++i)
    for (int k = 0; k < 3; ++k) {
      int v = tt[3 * i + k] - 1;
      adj[deg[v] + fill[v]++] = i;
    }
  corner->clear();
  element->clear();
  // pairs (i < j) in order of i, then j: triangle ranges scanned on threads,
  // their lists concatenated in range order
  const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int nch = std::max(1, std::min(nth, nt / 2048));
  std::vector<std::vector<int>> pcor(nch), pelm(nch);
  std::vector<char> dup(nch, 0);
  auto scan = [&](int c) {
    const int i0 = (int)((int64_t)nt * c / nch), i1 = (int)((int64_t)nt * (c + 1) / nch);
    std::vector<int> cand;
    for (int i = i0; i < i1 && !dup[c]; ++i) {
      cand.clear();
      for (int k = 0; k < 3; ++k) {
        int v = tt[3 * i + k] - 1;
        for (int a = deg[v]; a < deg[v + 1]; ++a)
          if (adj[a] > i) cand.push_back(adj[a]);
      }
      std::sort(cand.begin(), cand.end());
      cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
      for (int j : cand) {
        int com[3], ne = 0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b)
            if (tt[3 * i + a] == tt[3 * j + b] && ne < 3) com[ne++] = tt[3 * i + a];
        if (ne == 2) {
          pcor[c].push_back(com[0]);
          pcor[c].push_back(com[1]);
          pelm[c].push_back(i + 1);
          pelm[c].push_back(j + 1);
        } else if (ne == 3) {
          dup[c] = 1;
        }
      }
    }
  };
  {
    std::vector<std::thread> th;
    for (int c = 1; c < nch; ++c) th.emplace_back(scan, c);
    scan(0);
    for (auto &t : th) t.join();
  }
  for (int c = 0; c < nch; ++c) {
    if (dup[c]) {
      *err = "duplicate triangle in mesh";
      return SPK_ERR_MESH;
    }
    corner->insert(corner->end(), pcor[c].begin(), pcor[c].end());
    element->insert(element->end(), pelm[c].begin(), pelm[c].end());
  }
  const int N = (int)element->size() / 2;
  std::vector<double> cent(3 * (size_t)nt);
  for (int i = 0; i < nt; ++i)
    for (int d = 0; d < 3; ++d)
      cent[3 * i + d] = (pp[3 * (tt[3 * i] - 1) + d] + pp[3 * (tt[3 * i + 1] - 1) + d] +
                         pp[3 * (tt[3 * i + 2] - 1) + d]) /
                        3.0;
  fn->assign(2 * (size_t)N, FnRec{});
  er->assign(2 * (size_t)N, ElemRec{});
  rref->assign(N, 0.0);
  auto records = [&](int m0, int m1) {
  for (int m = m0; m < m1; ++m) {
    for (int ot = 1; ot <= 2; ++ot) {
      int e = (*element)[2 * m + ot - 1] - 1;
      const int *cr = &tt[3 * e];
      int nv = 0, ipc = 0;
      for (int j = 0; j < 3; ++j)
        if (cr[j] != (*corner)[2 * m] && cr[j] != (*corner)[2 * m + 1]) {
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
      FnRec &f = (*fn)[2 * m + ot - 1];
      for (int d = 0; d < 3; ++d) {
        f.c[d] = pp[3 * (i1 - 1) + d];
        f.c[3 + d] = pp[3 * (i2 - 1) + d];
        f.c[6 + d] = pp[3 * (ipc - 1) + d];
      }
      double dv[3] = {f.c[0] - f.c[3], f.c[1] - f.c[4], f.c[2] - f.c[5]};
      f.length = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
      double a1[3] = {f.c[0] - f.c[6], f.c[1] - f.c[7], f.c[2] - f.c[8]};
      double a2[3] = {f.c[3] - f.c[6], f.c[4] - f.c[7], f.c[5] - f.c[8]};
      double x[3] = {a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2],
                     a1[0] * a2[1] - a1[1] * a2[0]};
      f.area = 0.5 * std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
      int sg = -1 * (((ot - 1) * ot) - 1);
      f.sign = sg;
      f.df = sg * f.length / f.area;
      for (int d = 0; d < 3; ++d) f.mid[d] = cent[3 * e + d];
      double v1[3], v2[3];
      for (int d = 0; d < 3; ++d) {
        if (normal_convention == 1) {
          v1[d] = f.c[d] - f.c[3 + d];
          v2[d] = f.c[6 + d] - f.c[3 + d];
        } else {
          v1[d] = f.c[6 + d] - f.c[d];
          v2[d] = f.c[3 + d] - f.c[d];
        }
      }
      double nx[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2],
                      v1[0] * v2[1] - v1[1] * v2[0]};
      double nl = std::sqrt(nx[0] * nx[0] + nx[1] * nx[1] + nx[2] * nx[2]);
      for (int d = 0; d < 3; ++d) {
        f.nq[d] = nx[d] / nl;
        f.nqf[d] = (double)(float)f.nq[d];
      }
      ElemRec &r = (*er)[2 * m + ot - 1];
      r.elem = e;
      r.node[0] = cr[0];
      r.node[1] = cr[1];
      r.node[2] = cr[2];
    }
    int p = (*element)[2 * m] - 1, q = (*element)[2 * m + 1] - 1;
    double d[3] = {cent[3 * p] - cent[3 * q], cent[3 * p + 1] - cent[3 * q + 1],
                   cent[3 * p + 2] - cent[3 * q + 2]};
    (*rref)[m] = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  }
  };
  {
    const int nr = std::max(1, std::min(nth, N / 2048));
    std::vector<std::thread> th;
    for (int c = 1; c < nr; ++c)
      th.emplace_back(records, (int)((int64_t)N * c / nr), (int)((int64_t)N * (c + 1) / nr));
    records(0, (int)((int64_t)N / nr));
    for (auto &t : th) t.join();
  }
  return SPK_OK;
}

// ---------------------------------------------------------------------------
// Octree: lib_tree_constructor (src/lib/libtree/src/lib/lib_tree.f90:176-320),
// scaling (:370-434), Morton index from the coordinate mantissa
// (lib_tree_helper_functions.f90:217-379, 547-629; bit i of x -> 3i+2,
// y -> 3i+1, z -> 3i), l_max(s) (:1290-1344), E2 (:675-858), I4
// (:1096-1245).  Data points are the common-edge midpoints
// (src/lib_sie_tri_mlfmm_interface.f90:75-81).
// ---------------------------------------------------------------------------
static uint64_t spread3(uint64_t v) {
  uint64_t r = 0;
  for (int i = 0; i < 21; ++i) r |= ((v >> i) & 1ull) << (3 * i);
  return r;
}
static uint64_t morton(uint64_t x, uint64_t y, uint64_t z) {
  return (spread3(x) << 2) | (spread3(y) << 1) | spread3(z);
}
static void demorton(uint64_t n, int l, int64_t *x) {
  x[0] = x[1] = x[2] = 0;
  for (int i = 0; i < l; ++i) {
    x[0] |= (int64_t)((n >> (3 * i + 2)) & 1) << i;
    x[1] |= (int64_t)((n >> (3 * i + 1)) & 1) << i;
    x[2] |= (int64_t)((n >> (3 * i)) & 1) << i;
  }
}

// own box first, then neighbours in nested x, y, z order (lib_tree_get_domain_i2)
static int e2_boxes(int64_t n, int l, int64_t *out) {
  int64_t x[3];
  const int64_t ub = ((int64_t)1 << l) - 1;
  demorton((uint64_t)n, l, x);
  int cnt = 0;
  out[cnt++] = n;
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j)
      for (int k = -1; k <= 1; ++k) {
        int64_t a = x[0] + i, b = x[1] + j, c = x[2] + k;
        if (a < 0 || a > ub || b < 0 || b > ub || c < 0 || c > ub) continue;
        int64_t nb = (int64_t)morton(a, b, c);
        if (nb != n) out[cnt++] = nb;
      }
  return cnt;
}

static int find_box(const std::vector<int64_t> &v, int64_t n) {
  auto it = std::lower_bound(v.begin(), v.end(), n);
  if (it == v.end() || *it != n) return -1;
  return (int)(it - v.begin());
}

void box_centre(const Tree &t, int64_t n, int l, double *x) {
  int64_t ix[3];
  demorton((uint64_t)n, l, ix);
  for (int d = 0; d < 3; ++d) x[d] = std::ldexp(1.0, -l) * ((double)ix[d] + 0.5) * t.D + t.xmin[d];
}

void build_tree(const std::vector<FnRec> &fn, int N, int s_opt, Tree *t) {
  std::vector<double> pts(3 * (size_t)N);
  for (int m = 0; m < N; ++m)
    for (int d = 0; d < 3; ++d) pts[3 * m + d] = (fn[2 * m].c[d] + fn[2 * m].c[3 + d]) / 2;
  double xmin[3], xmax[3], Dd[3];
  for (int d = 0; d < 3; ++d) {
    xmin[d] = xmax[d] = pts[d];
    for (int m = 1; m < N; ++m) {
      xmax[d] = std::max(xmax[d], pts[3 * m + d]);
      xmin[d] = std::min(xmin[d], pts[3 * m + d]);
    }
    Dd[d] = (xmax[d] - xmin[d]) * (1 + std::ldexp(1.0, -52));
  }
  t->D = std::max(Dd[0], std::max(Dd[1], Dd[2]));
  std::memcpy(t->xmin, xmin, sizeof xmin);
  std::vector<uint64_t> u(N);
  for (int m = 0; m < N; ++m) {
    uint64_t ix[3];
    for (int d = 0; d < 3; ++d) {
      double xs = (pts[3 * m + d] - xmin[d]) / t->D;
      ix[d] = (uint64_t)(xs * 2097152.0);
    }
    u[m] = morton(ix[0], ix[1], ix[2]);
  }
  std::vector<uint64_t> us(u);
  std::sort(us.begin(), us.end());
  int l_max = 1;
  for (int i = 0; i + s_opt < N; ++i) {
    uint64_t a = us[i], b = us[i + s_opt];
    int j = kThresholdLevel + 1;
    for (;;) {
      --j;
      a >>= 3;
      b >>= 3;
      if (a == b) {
        l_max = std::max(l_max, j);
        break;
      }
    }
  }
  t->l_max = l_max;
  t->l_min = std::min(2, l_max);
  t->leaf.resize(N);
  for (int m = 0; m < N; ++m) t->leaf[m] = (int64_t)(u[m] >> (3 * (kThresholdLevel - l_max)));
  t->lev.assign(l_max + 1, Level{});
  for (int l = 0; l <= l_max; ++l) {
    std::vector<int64_t> &v = t->lev[l].n;
    v.resize(N);
    for (int m = 0; m < N; ++m) v[m] = t->leaf[m] >> (3 * (l_max - l));
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  // box-major renumbering
  const Level &LL = t->lev[l_max];
  const int nbl = (int)LL.n.size();
  t->leaf_box.resize(N);
  t->box_count.assign(nbl, 0);
  for (int m = 0; m < N; ++m) {
    t->leaf_box[m] = find_box(LL.n, t->leaf[m]);
    t->box_count[t->leaf_box[m]]++;
  }
  t->box_first.assign(nbl, 0);
  for (int b = 1; b < nbl; ++b) t->box_first[b] = t->box_first[b - 1] + t->box_count[b - 1];
  t->perm.resize(N);
  t->pos.resize(N);
  {
    std::vector<int> f(nbl, 0);
    for (int m = 0; m < N; ++m) {
      int b = t->leaf_box[m];
      int p = t->box_first[b] + f[b]++;
      t->perm[p] = m;
      t->pos[m] = p;
    }
  }
  // E2 lists per leaf box (sorted edge numbers, sort_edge_index): the
  // neighbour boxes and list sizes first, then the boxes' lists filled and
  // sorted in parallel (each box's segment is its own)
  t->e2_off.assign(nbl + 1, 0);
  t->near_off.assign(nbl + 1, 0);
  std::vector<int> nbr(27 * (size_t)nbl, -1);
  for (int b = 0; b < nbl; ++b) {
    int64_t bx[27];
    const int nb = e2_boxes(LL.n[b], l_max, bx);
    int cnt = 0, w = 0;
    for (int q = 0; q < nb; ++q) {
      const int p = find_box(LL.n, bx[q]);
      if (p < 0) continue;
      nbr[27 * (size_t)b + w++] = p;
      cnt += t->box_count[p];
    }
    t->e2_off[b + 1] = t->e2_off[b] + cnt;
    t->near_off[b + 1] = t->near_off[b] + (int64_t)t->box_count[b] * cnt;
  }
  t->e2_list.assign(t->e2_off[nbl], 0);
  {
    auto fill = [&](int b0, int b1) {
      for (int b = b0; b < b1; ++b) {
        int *dst = t->e2_list.data() + t->e2_off[b];
        for (int q = 0; q < 27 && nbr[27 * (size_t)b + q] >= 0; ++q) {
          const int p = nbr[27 * (size_t)b + q];
          for (int k = 0; k < t->box_count[p]; ++k) *dst++ = t->perm[t->box_first[p] + k];
        }
        std::sort(t->e2_list.data() + t->e2_off[b], dst);
      }
    };
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int w = 0; w < nt; ++w) th.emplace_back(fill, (int)((int64_t)nbl * w / nt), (int)((int64_t)nbl * (w + 1) / nt));
    for (auto &x : th) x.join();
  }
  t->nnz = t->near_off[nbl];
  // per-level parent/child/I4
  for (int l = t->l_min; l <= l_max; ++l) {
    Level &L = t->lev[l];
    const int nb = (int)L.n.size();
    L.parent.assign(nb, -1);
    L.octant.assign(nb, 0);
    if (l > t->l_min || l >= 1) {
      for (int b = 0; b < nb; ++b) {
        L.octant[b] = (int)(L.n[b] & 7);
        if (l > t->l_min) L.parent[b] = find_box(t->lev[l - 1].n, L.n[b] >> 3);
      }
    }
    if (l < l_max) {
      const Level &Lc = t->lev[l + 1];
      L.child.assign(8 * (size_t)nb, -1);
      for (int b = 0; b < nb; ++b)
        for (int o = 0; o < 8; ++o) L.child[8 * b + o] = find_box(Lc.n, (L.n[b] << 3) | o);
    }
    L.i4_off.assign(nb + 1, 0);
    L.i4_box.clear();
    L.i4_slot.clear();
    for (int a = 0; a < nb; ++a) {
      int64_t par[27], own[27], cand[27 * 8];
      int np_ = (l >= 1) ? e2_boxes(L.n[a] >> 3, l - 1, par) : 0;
      int no = e2_boxes(L.n[a], l, own);
      int nc = 0;
      for (int i = 0; i < np_; ++i)
        for (int o = 0; o < 8; ++o) cand[nc++] = (par[i] << 3) | o;
      std::sort(cand, cand + nc);
      int64_t xa[3];
      demorton((uint64_t)L.n[a], l, xa);
      for (int i = 0; i < nc; ++i) {
        bool skip = false;
        for (int k = 0; k < no; ++k)
          if (own[k] == cand[i]) skip = true;
        if (skip) continue;
        int bp = find_box(L.n, cand[i]);
        if (bp < 0) continue;
        int64_t xb[3];
        demorton((uint64_t)cand[i], l, xb);
        int slot = (int)((xa[0] - xb[0] + 3) * 49 + (xa[1] - xb[1] + 3) * 7 + (xa[2] - xb[2] + 3));
        L.i4_box.push_back(bp);
        L.i4_slot.push_back(slot);
      }
      L.i4_off[a + 1] = (int)L.i4_box.size();
    }
  }
}

// ---------------------------------------------------------------------------
// Special functions for the tables.
// ---------------------------------------------------------------------------
// Gauss-Legendre on [-1,1], ascending: Newton iteration on the Legendre
// three-term recurrence, weights renormalised to sum 2 (the normalisation
// of Gauss_Quadrature_Fast.f90 legendre_compute_glr).
void gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < n; ++i) {
    double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) p0 = 1.0;
      double pd = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / pd;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    double pd = n * (z * p1 - p0) / (z * z - 1.0);
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pd * pd);
  }
  double s = 0;
  for (int i = 0; i < n; ++i) s += w[i];
  for (int i = 0; i < n; ++i) w[i] = 2.0 * w[i] / s;
}

// h_l^(2)(z) = j_l(z) - i y_l(z), l = 0..n-1: j_l by normalised backward
// (Miller) recurrence, y_l by forward recurrence (the CSPHJY scheme behind
// src/lib/libmath/src/math/lib_math_bessel.f90:936-991).
void hankel2_sph(int n, double zre, double zim, cplx *out) {
  const zc z(zre, zim);
  const int nmax = n - 1;
  std::vector<zc> sj(n + 2), sy(n + 2);
  const zc sj0 = std::sin(z) / z;
  sj[0] = sj0;
  if (nmax >= 1) {
    const zc sj1 = (sj0 - std::cos(z)) / z;
    const int st = (int)std::abs(z) + nmax + 60;
    std::vector<zc> tmp(st + 2);
    zc f2 = 0.0, f1 = 1e-100, f = 0.0, f0 = 0.0, fj1 = 0.0;
    for (int k = st; k >= 0; --k) {
      f = (2.0 * k + 3.0) * f1 / z - f2;
      if (k <= nmax + 1) tmp[k] = f;
      if (k == 1) fj1 = f;
      if (k == 0) f0 = f;
      f2 = f1;
      f1 = f;
      if (std::abs(f) > 1e150) {
        f *= 1e-150;
        f1 *= 1e-150;
        f2 *= 1e-150;
        fj1 *= 1e-150;
        for (int q = k; q <= nmax + 1 && q <= st; ++q) tmp[q] *= 1e-150;
      }
    }
    const zc cs = (std::abs(sj0) > std::abs(sj1)) ? sj0 / f0 : sj1 / fj1;
    for (int k = 0; k <= nmax; ++k) sj[k] = cs * tmp[k];
  }
  sy[0] = -std::cos(z) / z;
  if (nmax >= 1) sy[1] = (sy[0] - std::sin(z)) / z;
  for (int k = 2; k <= nmax; ++k) sy[k] = (2.0 * k - 1.0) * sy[k - 1] / z - sy[k - 2];
  for (int k = 0; k <= nmax; ++k) out[k] = C(sj[k] - zc(0, 1) * sy[k]);
}

// Khat_and_TRF_ExtendedK (src/lib_sie_mlfmm.f90:190-249): K Gauss-Legendre
// nodes in cos(theta) x 2K uniform phi, plus the two poles (zero weight,
// Cartesian x/y components stored there).
void make_tables(int K, Tables *tb) {
  tb->K = K;
  const int Ns = 2 * K * K, S = Ns + 2;
  tb->S = S;
  tb->khat.assign(3 * S, 0.0);
  tb->ttheta.assign(3 * S, 0.0);
  tb->tphi.assign(3 * S, 0.0);
  tb->trf.assign(S, 0.0);
  std::vector<double> xx(K), ww(K);
  gauss_legendre(K, xx.data(), ww.data());
  const double dphi = 2 * kPi / (2 * K);
  for (int j = 0; j < K; ++j)
    for (int k = 0; k < 2 * K; ++k) {
      const double phi = k * dphi;
      const int i = k + j * 2 * K;
      const double st = std::sin(std::acos(xx[j]));
      tb->khat[3 * i] = st * std::cos(phi);
      tb->khat[3 * i + 1] = st * std::sin(phi);
      tb->khat[3 * i + 2] = xx[j];
      tb->ttheta[3 * i] = xx[j] * std::cos(phi);
      tb->ttheta[3 * i + 1] = xx[j] * std::sin(phi);
      tb->ttheta[3 * i + 2] = -st;
      tb->tphi[3 * i] = -std::sin(phi);
      tb->tphi[3 * i + 1] = std::cos(phi);
      tb->tphi[3 * i + 2] = 0.0;
      tb->trf[i] = ww[j] * dphi;
    }
  for (int p = 0; p < 2; ++p) {
    const int i = Ns + p;
    tb->khat[3 * i + 2] = p == 0 ? 1.0 : -1.0;
    tb->ttheta[3 * i] = 1.0;
    tb->tphi[3 * i + 1] = 1.0;
    tb->trf[i] = 0.0;
  }
}

// TRF * TL for one level: TL_km_ml_fmm (src/lib_sie_mlfmm.f90:129-188) on
// the 7^3 offset grid of precalculation_LT3_TL (:27-82) with the E2 cube
// left empty (lib_sie_lookup_table.f90:631-656).  For a lossy interior the
// k2 table is zero above the leaf level (:166-175).
void make_tl_level(const Tables &tb, const Material &mat, const Tree &t, int l,
                   std::vector<cplx> *out, int nthreads) {
  const int S = tb.S, K = tb.K;
  out->assign(343 * 2 * (size_t)S, cplx{0, 0});
  const double d_c = (1.0 / (double)(1LL << l)) * t.D;
  const double x_min = -(3 * d_c), x_max = -x_min, dx = (x_max - x_min) / 6;
  const bool lossy_skip = (std::fabs(mat.k2.im) > 0.0) && (l < t.l_max);
  auto work = [&](int t0, int t1) {
    std::vector<cplx> hl(K);
    std::vector<double> pm(K);
    for (int ts = t0; ts < t1; ++ts) {
      const int i = ts / 49, j = (ts / 7) % 7, k = ts % 7;
      if (i >= 2 && i <= 4 && j >= 2 && j <= 4 && k >= 2 && k <= 4) continue;
      const double r[3] = {x_min + i * dx, x_min + j * dx, x_min + k * dx};
      const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
      const double rh[3] = {r[0] / rl, r[1] / rl, r[2] / rl};
      for (int med = 0; med < 2; ++med) {
        if (med == 1 && lossy_skip) continue;
        const zc kd = med == 0 ? Z(mat.k1) : Z(mat.k2);
        const zc zr = kd * rl;
        hankel2_sph(K, zr.real(), zr.imag(), hl.data());
        cplx *o = &(*out)[((size_t)ts * 2 + med) * S];
        for (int s = 0; s < S; ++s) {
          const double x = rh[0] * tb.khat[3 * s] + rh[1] * tb.khat[3 * s + 1] + rh[2] * tb.khat[3 * s + 2];
          pm[0] = 1.0;
          if (K > 1) pm[1] = x;
          for (int q = 2; q < K; ++q) pm[q] = (2.0 * q - 1.0) / q * x * pm[q - 1] - (q - 1.0) / q * pm[q - 2];
          zc acc = 0.0, mi = 1.0;
          for (int q = 1; q <= K; ++q) {
            mi = mi * zc(0.0, -1.0);
            acc += Z(hl[q - 1]) * mi * (double)(2 * (q - 1) + 1) * pm[q - 1] * kd /
                   ((4 * kPi) * (4 * kPi));
          }
          o[s] = C(tb.trf[s] * acc);
        }
      }
    }
  };
  int nt = std::max(1, nthreads);
  std::vector<std::thread> th;
  for (int w = 0; w < nt; ++w) th.emplace_back(work, 343 * w / nt, 343 * (w + 1) / nt);
  for (auto &x : th) x.join();
}

// The per-slot factors of the TL series for the device table kernel: for
// slot ts of the 7^3 offset grid and medium med, rhat (3 doubles) and
// c_q = h_q^(2)(k r) (-i)^(q+1) (2q+1) k / (4 pi)^2, q = 0..K-1, so that
// TL(khat) = sum_q c_q P_q(rhat.khat); live = 0 for the zeroed slots (the
// 3^3 near cube, and the lossy interior medium above the leaf level).
void tl_coefficients(const Tables &tb, const Material &mat, const Tree &t, int l,
                     std::vector<double> *rhat, std::vector<cplx> *coef,
                     std::vector<int> *live) {
  const int K = tb.K;
  rhat->assign(343 * 3, 0.0);
  coef->assign(343 * 2 * (size_t)K, cplx{0, 0});
  live->assign(343 * 2, 0);
  const double d_c = (1.0 / (double)(1LL << l)) * t.D;
  const double x_min = -(3 * d_c), x_max = -x_min, dx = (x_max - x_min) / 6;
  const bool lossy_skip = (std::fabs(mat.k2.im) > 0.0) && (l < t.l_max);
  std::vector<cplx> hl(K);
  for (int ts = 0; ts < 343; ++ts) {
    const int i = ts / 49, j = (ts / 7) % 7, k = ts % 7;
    if (i >= 2 && i <= 4 && j >= 2 && j <= 4 && k >= 2 && k <= 4) continue;
    const double r[3] = {x_min + i * dx, x_min + j * dx, x_min + k * dx};
    const double rl = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    for (int d = 0; d < 3; ++d) (*rhat)[3 * ts + d] = r[d] / rl;
    for (int med = 0; med < 2; ++med) {
      if (med == 1 && lossy_skip) continue;
      const zc kd = med == 0 ? Z(mat.k1) : Z(mat.k2);
      const zc zr = kd * rl;
      hankel2_sph(K, zr.real(), zr.imag(), hl.data());
      zc mi = 1.0;
      for (int q = 1; q <= K; ++q) {
        mi = mi * zc(0.0, -1.0);
        (*coef)[((size_t)ts * 2 + med) * K + (q - 1)] =
            C(Z(hl[q - 1]) * mi * (double)(2 * (q - 1) + 1) * kd / ((4 * kPi) * (4 * kPi)));
      }
      (*live)[ts * 2 + med] = 1;
    }
  }
}

// SS/RR phases of translation_SS/RR (src/lib_sie_tri_mlfmm_interface.f90:466-655):
// child level l, octant o: delta = centre(child) - centre(parent)
// = +-2^-(l+1) D per axis; up = exp(+i k khat.delta), down = exp(-i k khat.delta).
void make_phase_level(const Tables &tb, const Material &mat, const Tree &t, int l,
                      std::vector<cplx> *up, std::vector<cplx> *down) {
  const int S = tb.S;
  up->assign(8 * 2 * (size_t)S, cplx{0, 0});
  down->assign(8 * 2 * (size_t)S, cplx{0, 0});
  const double h = std::ldexp(1.0, -(l + 1)) * t.D;
  for (int o = 0; o < 8; ++o) {
    const double dlt[3] = {((o >> 2) & 1) ? h : -h, ((o >> 1) & 1) ? h : -h, (o & 1) ? h : -h};
    for (int med = 0; med < 2; ++med) {
      const zc kd = med == 0 ? Z(mat.k1) : Z(mat.k2);
      for (int s = 0; s < S; ++s) {
        const double tmp = tb.khat[3 * s] * dlt[0] + tb.khat[3 * s + 1] * dlt[1] + tb.khat[3 * s + 2] * dlt[2];
        (*up)[((size_t)o * 2 + med) * S + s] = C(std::exp(zc(0, 1) * kd * tmp));
        (*down)[((size_t)o * 2 + med) * S + s] = C(std::exp(-zc(0, 1) * kd * tmp));
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Interpolation between the sample grids of two truncation numbers
// (Lagrange_Interpolation_New, src/lib_sie_mlfmm.f90:268-493, with
// theta_phi_extended :573-610 and Sorting_theta_phi_LenReduced_Arr_New
// :414-493) as a sparse matrix on one (theta, phi) component pair of one
// medium.  The extended grid is built symbolically: every cell records
// where its value comes from (an interior sample, a pole row's Cartesian
// formula at its column's angle, or zero) and its sign, by the same
// sequence of fills, reflections across the poles and periodic copies,
// in the same order, as the reference builds the values.  Stencil
// indices use the float arithmetic of the reference's default reals.
// ---------------------------------------------------------------------------
namespace {
struct ExtCell {
  int kind = 0;  // 0 zero, 1 interior sample, 2 pole theta = 0, 3 pole theta = pi
  int src = 0;   // interior sample index
  double c = 0, s = 0, sign = 1;
};
}  // namespace

void interp_pair_csr(int p, int Km, int Kn, std::vector<int> *rowptr, std::vector<int> *col,
                     std::vector<double> *val) {
  const int Kx = 2 * Km;  // phi samples of the source grid
  const int Sm = 2 * Km * Km + 2, Sn = 2 * Kn * Kn + 2, Nsm = 2 * Km * Km, Nsn = 2 * Kn * Kn;
  const int nth = Km + 2 * p + 2, nph = Kx + 2 * p + 1;
  // theta_phi_extended: a_theta(1..3Km+2), a_phi(1..6Km), 1-based
  std::vector<double> xx(Km), ww(Km), at(3 * Km + 3), ap(6 * Km + 1);
  gauss_legendre(Km, xx.data(), ww.data());
  const double d_phi = 6 * kPi / (3 * (2 * Km));
  for (int j = 1; j <= Km; ++j) {
    at[j] = 2 * kPi - std::acos(xx[Km - (j - 1) - 1]);
    at[j + Km + 1] = std::acos(xx[j - 1]);
    at[2 * Km + 2 + j] = -std::acos(xx[Km - (j - 1) - 1]);
  }
  at[Km + 1] = kPi;
  at[2 * Km + 2] = 0.0;
  for (int k = 1; k <= 6 * Km; ++k) ap[k] = -2 * kPi + d_phi * (k - 1);
  // the extended grid (1-based rows m, columns n)
  std::vector<ExtCell> E((size_t)(nth + 1) * (nph + 1));
  auto cell = [&](int m, int n) -> ExtCell & { return E[(size_t)m * (nph + 1) + n]; };
  for (int m = 1; m <= nph; ++m) {
    const int n = m + Kx - p;
    ExtCell pi_row, zero_row;
    pi_row.kind = 3;
    zero_row.kind = 2;
    pi_row.c = zero_row.c = std::cos(ap[n]);
    pi_row.s = zero_row.s = std::sin(ap[n]);
    cell(p + 1, m) = pi_row;
    cell(Km + p + 2, m) = zero_row;
  }
  for (int i = 1; i <= Km; ++i)
    for (int j = 1; j <= Kx; ++j) {
      ExtCell e;
      e.kind = 1;
      e.src = (i - 1) * Kx + (j - 1);
      cell(p + 1 + i, p + j) = e;
    }
  for (int m = 1; m <= nth; ++m) {
    const double xt = at[m - p + Km + 1];
    for (int n = 1; n <= nph; ++n) {
      const double xp = ap[n - p + Kx];
      int src = 0, dn = 0;
      if (xt < 0.0)
        src = 2 * (Km + p + 2) - m;
      else if (xt >= kPi)
        src = 2 * (p + 1) - m;
      if (!src) continue;
      if (xp < kPi && xp >= 0.0)
        dn = Kx / 2;
      else if (xp < 2 * kPi && xp >= kPi)
        dn = -Kx / 2;
      if (!dn) continue;
      ExtCell e = cell(src, n + dn);
      e.sign = -e.sign;
      cell(m, n) = e;
    }
  }
  for (int m = 1; m <= nth; ++m)
    for (int n = 1; n <= nph; ++n) {
      const double xp = ap[n - p + Kx];
      if (xp >= 2 * kPi)
        cell(m, n) = cell(m, n - Kx);
      else if (xp < 0)
        cell(m, n) = cell(m, n + Kx);
    }
  // Lagrange_Interpolation_Weights on the 2p nodes ext(K_arr+t+1-p ..)
  auto weights = [&](double x, int t, int K_arr, const std::vector<double> &ext, double *wk) {
    double nodes[16];
    const int kt = K_arr + t;
    for (int j = 1; j <= 2 * p; ++j) nodes[j - 1] = ext[kt + 1 - p + (j - 1)];
    for (int j = 1; j <= 2 * p; ++j) {
      wk[j - 1] = 1.0;
      for (int m = 1; m <= 2 * p; ++m)
        if (m != j) wk[j - 1] = (x - nodes[m - 1]) / (nodes[j - 1] - nodes[m - 1]) * wk[j - 1];
    }
  };
  std::vector<double> xt(Kn), wt(Kn);
  gauss_legendre(Kn, xt.data(), wt.data());
  rowptr->assign(2 * (size_t)Sn + 1, 0);
  col->clear();
  val->clear();
  std::vector<std::vector<std::pair<int, double>>> rows(2 * (size_t)Sn);
  double wkt[16], wkp[16];
  for (int it = 1; it <= Kn; ++it) {
    const double theta = std::acos(xt[it - 1]);
    const int t = (int)std::floor((float)((float)(it - 1.0f) / (float)Kn) * (float)Km) + 1;
    weights(theta, t, Km, at, wkt);
    for (int jp = 1; jp <= 2 * Kn; ++jp) {
      const int tt = (it - 1) * 2 * Kn + (jp - 1);
      const double phi = (jp - 1) * 2 * kPi / (2 * Kn);
      const int sidx = (int)std::floor((float)((float)(jp - 1.0f) / (float)Kn) * (float)Km) + 1;
      weights(phi, sidx, Kx, ap, wkp);
      for (int i = 1; i <= 2 * p; ++i)
        for (int j = 1; j <= 2 * p; ++j) {
          const double w = wkp[j - 1] * wkt[i - 1];
          const ExtCell &e = cell(p + t + i - p, p + sidx + j - p);
          for (int co = 0; co < 2; ++co) {
            auto &r = rows[(size_t)co * Sn + tt];
            if (e.kind == 1) {
              r.push_back({co * Sm + e.src, w * e.sign});
            } else if (e.kind == 2 || e.kind == 3) {
              const int ps = e.kind == 2 ? Nsm : Nsm + 1;
              const double eps = e.kind == 2 ? 1.0 : -1.0;  // theta row sign at the pole
              const double ws = w * e.sign;
              if (co == 0) {  // theta: eps (c fx + s fy)
                r.push_back({0 * Sm + ps, ws * eps * e.c});
                r.push_back({1 * Sm + ps, ws * eps * e.s});
              } else {  // phi: -s fx + c fy
                r.push_back({0 * Sm + ps, -ws * e.s});
                r.push_back({1 * Sm + ps, ws * e.c});
              }
            }
          }
        }
    }
  }
  for (int co = 0; co < 2; ++co)
    for (int q = 0; q < 2; ++q) rows[(size_t)co * Sn + Nsn + q].push_back({co * Sm + Nsm + q, 1.0});
  for (size_t r = 0; r < rows.size(); ++r) {
    for (auto &e : rows[r]) {
      col->push_back(e.first);
      val->push_back(e.second);
    }
    (*rowptr)[r + 1] = (int)col->size();
  }
}

// the transpose of a CSR matrix with ncols columns
void csr_transpose(int ncols, const std::vector<int> &rowptr, const std::vector<int> &col,
                   const std::vector<double> &val, std::vector<int> *trow, std::vector<int> *tcol,
                   std::vector<double> *tval) {
  const int nrows = (int)rowptr.size() - 1;
  trow->assign(ncols + 1, 0);
  for (int c : col) (*trow)[c + 1]++;
  for (int c = 0; c < ncols; ++c) (*trow)[c + 1] += (*trow)[c];
  tcol->assign(col.size(), 0);
  tval->assign(col.size(), 0.0);
  std::vector<int> fill(trow->begin(), trow->end() - 1);
  for (int r = 0; r < nrows; ++r)
    for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
      const int q = fill[col[e]]++;
      (*tcol)[q] = r;
      (*tval)[q] = val[e];
    }
}

}  // namespace spk

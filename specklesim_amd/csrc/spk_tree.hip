// spk_tree.hip -- the octree of the common-edge midpoints and its lists on
// the GPU: lib_tree_constructor (src/lib/libtree/src/lib/lib_tree.f90:176-320)
// with its scaling (:370-434), the Morton index from the coordinate mantissa
// (lib_tree_helper_functions.f90:217-379, 547-629), l_max(s_opt)
// (:1290-1344), E2 (:675-858) and I4 (:1096-1245), and the box-major
// renumbering and near-block offsets of this library.  Every list comes out
// identical to the host builder's (spk_host.cpp build_tree): the sorts are
// stable radix sorts, the E2 merge places each edge by its rank, and the
// I4 candidates are sorted as the host sorts them.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>

#include "spk_tree.h"

namespace spk {
namespace {

constexpr int kT = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

__host__ __device__ inline uint64_t spread3(uint64_t v) {
  uint64_t r = 0;
  for (int i = 0; i < 21; ++i) r |= ((v >> i) & 1ull) << (3 * i);
  return r;
}
__host__ __device__ inline int64_t morton3(int64_t x, int64_t y, int64_t z) {
  return (int64_t)((spread3((uint64_t)x) << 2) | (spread3((uint64_t)y) << 1) | spread3((uint64_t)z));
}
__host__ __device__ inline void demorton3(uint64_t n, int l, int64_t *x) {
  x[0] = x[1] = x[2] = 0;
  for (int i = 0; i < l; ++i) {
    x[0] |= (int64_t)((n >> (3 * i + 2)) & 1) << i;
    x[1] |= (int64_t)((n >> (3 * i + 1)) & 1) << i;
    x[2] |= (int64_t)((n >> (3 * i)) & 1) << i;
  }
}
// own box first, then the neighbours in nested x, y, z order
__device__ int e2_boxes_d(int64_t n, int l, int64_t *out) {
  int64_t x[3];
  const int64_t ub = ((int64_t)1 << l) - 1;
  demorton3((uint64_t)n, l, x);
  int cnt = 0;
  out[cnt++] = n;
  for (int i = -1; i <= 1; ++i)
    for (int j = -1; j <= 1; ++j)
      for (int k = -1; k <= 1; ++k) {
        const int64_t a = x[0] + i, b = x[1] + j, c = x[2] + k;
        if (a < 0 || a > ub || b < 0 || b > ub || c < 0 || c > ub) continue;
        const int64_t nb = morton3(a, b, c);
        if (nb != n) out[cnt++] = nb;
      }
  return cnt;
}
// position of key in the sorted v[0..nv), -1 if absent
__device__ inline int find_sorted(const int64_t *v, int nv, int64_t key) {
  int lo = 0, hi = nv;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (v[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return (lo < nv && v[lo] == key) ? lo : -1;
}
// number of entries of the ascending v[0..nv) below key
__device__ inline int lower_count(const int *v, int nv, int key) {
  int lo = 0, hi = nv;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (v[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// the data point of edge m: its common edge's midpoint (SoA, pts[d*N + m])
__global__ void k_tree_pts(int N, const FnRec *fn, double *pts) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N) return;
  for (int d = 0; d < 3; ++d) pts[(size_t)d * N + m] = (fn[2 * m].c[d] + fn[2 * m].c[3 + d]) / 2;
}

// per-block minima and maxima of the three coordinates, then one block
// reduces the partials (min and max are exact in any order)
__global__ void k_bbox_part(int N, const double *pts, double *part) {
  __shared__ double s[6][kT];
  double v[6] = {1e308, 1e308, 1e308, -1e308, -1e308, -1e308};
  for (int m = blockIdx.x * blockDim.x + threadIdx.x; m < N; m += gridDim.x * blockDim.x)
    for (int d = 0; d < 3; ++d) {
      const double x = pts[(size_t)d * N + m];
      v[d] = fmin(v[d], x);
      v[3 + d] = fmax(v[3 + d], x);
    }
  for (int q = 0; q < 6; ++q) s[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int w = kT / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int q = 0; q < 6; ++q)
        s[q][threadIdx.x] = q < 3 ? fmin(s[q][threadIdx.x], s[q][threadIdx.x + w])
                                  : fmax(s[q][threadIdx.x], s[q][threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; ++q) part[(size_t)blockIdx.x * 6 + q] = s[q][0];
}
__global__ void k_bbox_final(int nparts, const double *part, double *box) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double v[6] = {1e308, 1e308, 1e308, -1e308, -1e308, -1e308};
  for (int b = 0; b < nparts; ++b)
    for (int q = 0; q < 6; ++q)
      v[q] = q < 3 ? fmin(v[q], part[(size_t)b * 6 + q]) : fmax(v[q], part[(size_t)b * 6 + q]);
  for (int q = 0; q < 6; ++q) box[q] = v[q];
}

// Morton index of each point at the threshold level (21 bits per axis)
__global__ void k_tree_morton(int N, const double *pts, double x0, double y0, double z0, double D,
                              uint64_t *u, int *iota) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= N) return;
  const double xm[3] = {x0, y0, z0};
  int64_t ix[3];
  for (int d = 0; d < 3; ++d) {
    const double xs = (pts[(size_t)d * N + m] - xm[d]) / D;
    ix[d] = (int64_t)(uint64_t)(xs * 2097152.0);
  }
  u[m] = (uint64_t)morton3(ix[0], ix[1], ix[2]);
  iota[m] = m;
}

// l_max(s_opt): the deepest level at which the s_opt-th next point in
// Morton order shares the box (k shifts of 3 bits make them equal at level
// 22 - k)
__global__ void k_tree_lmax(int N, int s_opt, const uint64_t *us, int *lmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i + s_opt >= N) return;
  const uint64_t x = us[i] ^ us[i + s_opt];
  const int k = x ? (63 - __clzll((long long)x)) / 3 + 1 : 1;
  atomicMax(lmax, kThresholdLevel + 1 - k);
}

__global__ void k_tree_leafkey(int N, const uint64_t *u, int shift, uint64_t *lk) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < N) lk[m] = u[m] >> shift;
}

// first position of every distinct value of the sorted keys >> sh
__global__ void k_tree_flags(int N, const uint64_t *sk, int sh, int *flag) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) flag[p] = (p == 0 || (sk[p] >> sh) != (sk[p - 1] >> sh)) ? 1 : 0;
}
__global__ void k_tree_compact(int N, const uint64_t *sk, int sh, const int *flag, const int *scan,
                               int64_t *out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N && flag[p]) out[scan[p] - 1] = (int64_t)(sk[p] >> sh);
}

// leaf boxes: first positions, box of every edge and position, positions
__global__ void k_tree_leaf(int N, const int *flag, const int *scan, const int *perm,
                            int *box_first, int *leaf_box, int *pos, int *lbp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int b = scan[p] - 1;
  if (flag[p]) box_first[b] = p;
  const int e = perm[p];
  leaf_box[e] = b;
  pos[e] = p;
  lbp[p] = b;
}
__global__ void k_tree_count(int nbl, int N, const int *box_first, int *box_count) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nbl) box_count[b] = (b + 1 < nbl ? box_first[b + 1] : N) - box_first[b];
}

// E2 of every leaf box: its present neighbour boxes (27 slots, -1 after the
// last), the number of their edges and of the box's near blocks
__global__ void k_e2_count(int nbl, int l, const int64_t *Ln, const int *box_count, int *nbr,
                           int *cnt, int64_t *rows) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nbl) return;
  if (b == nbl) {  // the scans' last entry
    cnt[b] = 0;
    rows[b] = 0;
    return;
  }
  int64_t bx[27];
  const int nb = e2_boxes_d(Ln[b], l, bx);
  int w = 0, c = 0;
  for (int q = 0; q < nb; ++q) {
    const int p = find_sorted(Ln, nbl, bx[q]);
    if (p < 0) continue;
    nbr[27 * (size_t)b + w++] = p;
    c += box_count[p];
  }
  for (int q = w; q < 27; ++q) nbr[27 * (size_t)b + q] = -1;
  cnt[b] = c;
  rows[b] = (int64_t)box_count[b] * c;
}

// the sorted E2 list of box b (one workgroup per box): the neighbour boxes'
// edge lists are each ascending (positions inside a box run in edge order),
// so an edge's place in the merged list is its index in its own box plus
// the number of smaller edges in each other neighbour box
__global__ void __launch_bounds__(kT) k_e2_fill(int nbl, const int *nbr, const int *box_first,
                                                const int *box_count, const int *perm,
                                                const int *e2_off, int *e2_list) {
  const int b = blockIdx.x;
  if (b >= nbl) return;
  __shared__ int f[27], c[27], pre[28];
  __shared__ int nq;
  if (threadIdx.x == 0) {
    int q = 0, acc = 0;
    for (; q < 27 && nbr[27 * (size_t)b + q] >= 0; ++q) {
      const int p = nbr[27 * (size_t)b + q];
      f[q] = box_first[p];
      c[q] = box_count[p];
      pre[q] = acc;
      acc += c[q];
    }
    pre[q] = acc;
    nq = q;
  }
  __syncthreads();
  const int total = pre[nq];
  int *dst = e2_list + e2_off[b];
  for (int g = threadIdx.x; g < total; g += blockDim.x) {
    int q = 0;
    while (pre[q + 1] <= g) ++q;
    const int i = g - pre[q];
    const int e = perm[f[q] + i];
    int rank = i;
    for (int r = 0; r < nq; ++r)
      if (r != q) rank += lower_count(perm + f[r], c[r], e);
    dst[rank] = e;
  }
}

// I4 of box a at level l (one workgroup per box): the children of the
// parent's E2 boxes, sorted, without the box's own E2 boxes and without the
// empty ones; mode 0 counts, mode 1 writes partner positions and TL slots
__global__ void __launch_bounds__(kT) k_i4(int nb, int l, const int64_t *Ln, int mode, int *cnt,
                                           const int *off, int *box, int *slot) {
  const int a = blockIdx.x;
  if (a >= nb) return;
  __shared__ int64_t cand[kT];
  __shared__ int64_t own[27], par[27];
  __shared__ int npar, nown;
  __shared__ int keep_scan[kT];
  const int64_t na = Ln[a];
  if (threadIdx.x == 0) {
    npar = (l >= 1) ? e2_boxes_d(na >> 3, l - 1, par) : 0;
    nown = e2_boxes_d(na, l, own);
  }
  __syncthreads();
  const int t = threadIdx.x;
  cand[t] = t < 8 * npar ? ((par[t >> 3] << 3) | (t & 7)) : INT64_MAX;
  __syncthreads();
  // bitonic sort of the 256 candidates, ascending
  for (int k = 2; k <= kT; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int ixj = t ^ j;
      if (ixj > t) {
        const int64_t x = cand[t], y = cand[ixj];
        const bool up = (t & k) == 0;
        if ((x > y) == up) {
          cand[t] = y;
          cand[ixj] = x;
        }
      }
      __syncthreads();
    }
  const int64_t cnd = cand[t];
  int bp = -1;
  bool keep = cnd != INT64_MAX;
  if (keep)
    for (int q = 0; q < nown; ++q)
      if (own[q] == cnd) keep = false;
  if (keep) {
    bp = find_sorted(Ln, nb, cnd);
    keep = bp >= 0;
  }
  // exclusive prefix of keep over the workgroup, in candidate order
  keep_scan[t] = keep ? 1 : 0;
  __syncthreads();
  for (int w = 1; w < kT; w <<= 1) {
    const int v = t >= w ? keep_scan[t - w] : 0;
    __syncthreads();
    keep_scan[t] += v;
    __syncthreads();
  }
  if (mode == 0) {
    if (t == kT - 1) cnt[a] = keep_scan[t];
    return;
  }
  if (!keep) return;
  const int idx = off[a] + keep_scan[t] - 1;
  int64_t xa[3], xb[3];
  demorton3((uint64_t)na, l, xa);
  demorton3((uint64_t)cnd, l, xb);
  box[idx] = bp;
  slot[idx] = (int)((xa[0] - xb[0] + 3) * 49 + (xa[1] - xb[1] + 3) * 7 + (xa[2] - xb[2] + 3));
}

__global__ void k_tree_links(int nb, const int64_t *Ln, const int64_t *Lp, int nbp,
                             const int64_t *Lc, int nbc, int *parent, int *octant, int *child) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  octant[b] = (int)(Ln[b] & 7);
  parent[b] = Lp ? find_sorted(Lp, nbp, Ln[b] >> 3) : -1;
  if (Lc)
    for (int o = 0; o < 8; ++o) child[8 * (size_t)b + o] = find_sorted(Lc, nbc, (Ln[b] << 3) | o);
}

__global__ void k_set_int(int *p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

template <class T>
hipError_t to_host(std::vector<T> &v, const T *d, size_t n, hipStream_t st) {
  v.resize(n);
  if (n == 0) return hipSuccess;
  return hipMemcpyAsync(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, st);
}

}  // namespace

#define TRY(step, expr)                \
  do {                                 \
    const hipError_t e_ = (expr);      \
    if (e_ != hipSuccess) {            \
      *what = step;                    \
      return e_;                       \
    }                                  \
  } while (0)

hipError_t build_tree_device(const FnRec *d_fn, int N, int s_opt, Tree *t, TreeWork &w,
                             const TreeDevOut &out, hipStream_t st, const char **what) {
  *what = "";
  const int npart = 256;
  TRY("alloc", w.pts.alloc(3 * (size_t)N));
  TRY("alloc", w.part.alloc(6 * (size_t)npart));
  TRY("alloc", w.box.alloc(6));
  TRY("alloc", w.u.alloc(N));
  TRY("alloc", w.us.alloc(N));
  TRY("alloc", w.lk.alloc(N));
  TRY("alloc", w.sk.alloc(N));
  TRY("alloc", w.iota.alloc(N));
  TRY("alloc", w.flag.alloc(N));
  TRY("alloc", w.scan.alloc(N));
  TRY("alloc", w.lmax.alloc(1));
  TRY("alloc", out.perm->alloc(N));
  TRY("alloc", out.pos->alloc(N));
  TRY("alloc", out.leaf_box->alloc(N));
  TRY("alloc", out.leafbox_of_pos->alloc(N));
  // hipCUB scratch: the largest of the sorts and scans below
  {
    size_t a = 0, b = 0, c = 0, d = 0, e = 0;
    TRY("cub size", hipcub::DeviceRadixSort::SortKeys(nullptr, a, w.u.p, w.us.p, N, 0, 63, st));
    TRY("cub size", hipcub::DeviceRadixSort::SortPairs(nullptr, b, w.lk.p, w.sk.p, w.iota.p,
                                                        out.perm->p, N, 0, 63, st));
    TRY("cub size", hipcub::DeviceScan::InclusiveSum(nullptr, c, w.flag.p, w.scan.p, N, st));
    TRY("cub size", hipcub::DeviceScan::ExclusiveSum(nullptr, d, (const int64_t *)nullptr,
                                                      (int64_t *)nullptr, N + 1, st));
    TRY("cub size", hipcub::DeviceScan::ExclusiveSum(nullptr, e, (const int *)nullptr,
                                                      (int *)nullptr, N + 1, st));
    const size_t need = std::max(std::max(std::max(a, b), std::max(c, d)), e);
    TRY("alloc", w.cub.alloc(need));
  }
  size_t cub_bytes = w.cub.cap;
  // bounding box, D, xmin (host arithmetic of build_tree on the 6 extremes)
  hipLaunchKernelGGL(k_tree_pts, dim3(nblk(N)), dim3(kT), 0, st, N, d_fn, w.pts.p);
  hipLaunchKernelGGL(k_bbox_part, dim3(npart), dim3(kT), 0, st, N, w.pts.p, w.part.p);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, st, npart, w.part.p, w.box.p);
  TRY("bounding box", hipGetLastError());
  double bx[6];
  TRY("bounding box", hipMemcpyAsync(bx, w.box.p, sizeof bx, hipMemcpyDeviceToHost, st));
  TRY("bounding box", hipStreamSynchronize(st));
  double Dd[3];
  for (int d = 0; d < 3; ++d) {
    t->xmin[d] = bx[d];
    Dd[d] = (bx[3 + d] - bx[d]) * (1 + std::ldexp(1.0, -52));
  }
  t->D = std::max(Dd[0], std::max(Dd[1], Dd[2]));
  // Morton keys, sorted; l_max
  hipLaunchKernelGGL(k_tree_morton, dim3(nblk(N)), dim3(kT), 0, st, N, w.pts.p, t->xmin[0],
                     t->xmin[1], t->xmin[2], t->D, w.u.p, w.iota.p);
  TRY("sort", hipcub::DeviceRadixSort::SortKeys(w.cub.p, cub_bytes, w.u.p, w.us.p, N, 0, 63, st));
  int lmax = 1;
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, w.lmax.p, 1);
  hipLaunchKernelGGL(k_tree_lmax, dim3(nblk(N)), dim3(kT), 0, st, N, s_opt, w.us.p, w.lmax.p);
  TRY("l_max", hipGetLastError());
  TRY("l_max", hipMemcpyAsync(&lmax, w.lmax.p, sizeof(int), hipMemcpyDeviceToHost, st));
  TRY("l_max", hipStreamSynchronize(st));
  t->l_max = lmax;
  t->l_min = std::min(2, lmax);
  // leaf keys; edges sorted by leaf box, ascending inside a box (stable)
  const int shift = 3 * (kThresholdLevel - lmax);
  hipLaunchKernelGGL(k_tree_leafkey, dim3(nblk(N)), dim3(kT), 0, st, N, w.u.p, shift, w.lk.p);
  cub_bytes = w.cub.cap;
  TRY("sort", hipcub::DeviceRadixSort::SortPairs(w.cub.p, cub_bytes, w.lk.p, w.sk.p, w.iota.p,
                                                  out.perm->p, N, 0, 3 * lmax, st));
  // level lists, leaf level last
  std::vector<int> nlev(lmax + 1), lev_off(lmax + 2, 0);
  TRY("alloc", w.levn.alloc((size_t)(lmax + 1) * N));
  for (int l = 0; l <= lmax; ++l) {
    const int sh = 3 * (lmax - l);
    hipLaunchKernelGGL(k_tree_flags, dim3(nblk(N)), dim3(kT), 0, st, N, w.sk.p, sh, w.flag.p);
    cub_bytes = w.cub.cap;
    TRY("scan", hipcub::DeviceScan::InclusiveSum(w.cub.p, cub_bytes, w.flag.p, w.scan.p, N, st));
    hipLaunchKernelGGL(k_tree_compact, dim3(nblk(N)), dim3(kT), 0, st, N, w.sk.p, sh, w.flag.p,
                       w.scan.p, w.levn.p + (size_t)l * N);
    TRY("levels", hipMemcpyAsync(&nlev[l], w.scan.p + N - 1, sizeof(int), hipMemcpyDeviceToHost, st));
  }
  TRY("levels", hipStreamSynchronize(st));
  const int nbl = nlev[lmax];
  const int64_t *LL = w.levn.p + (size_t)lmax * N;
  // leaf boxes (flag and scan of the leaf level are still in place)
  TRY("alloc", out.box_first->alloc(nbl));
  TRY("alloc", out.box_count->alloc(nbl));
  hipLaunchKernelGGL(k_tree_leaf, dim3(nblk(N)), dim3(kT), 0, st, N, w.flag.p, w.scan.p,
                     out.perm->p, out.box_first->p, out.leaf_box->p, out.pos->p,
                     out.leafbox_of_pos->p);
  hipLaunchKernelGGL(k_tree_count, dim3(nblk(nbl)), dim3(kT), 0, st, nbl, N, out.box_first->p,
                     out.box_count->p);
  // E2 lists and near-block offsets
  TRY("alloc", w.nbr.alloc(27 * (size_t)nbl));
  TRY("alloc", w.cnt.alloc(nbl + 1));
  TRY("alloc", w.rows.alloc(nbl + 1));
  TRY("alloc", out.e2_off->alloc(nbl + 1));
  TRY("alloc", out.near_off->alloc(nbl + 1));
  hipLaunchKernelGGL(k_e2_count, dim3(nblk(nbl + 1)), dim3(kT), 0, st, nbl, lmax, LL,
                     out.box_count->p, w.nbr.p, w.cnt.p, w.rows.p);
  cub_bytes = w.cub.cap;
  TRY("scan", hipcub::DeviceScan::ExclusiveSum(w.cub.p, cub_bytes, w.cnt.p, out.e2_off->p, nbl + 1, st));
  cub_bytes = w.cub.cap;
  TRY("scan", hipcub::DeviceScan::ExclusiveSum(w.cub.p, cub_bytes, w.rows.p, out.near_off->p,
                                                nbl + 1, st));
  int e2_total = 0;
  TRY("e2", hipMemcpyAsync(&e2_total, out.e2_off->p + nbl, sizeof(int), hipMemcpyDeviceToHost, st));
  TRY("e2", hipStreamSynchronize(st));
  TRY("alloc", out.e2_list->alloc(std::max(e2_total, 1)));
  hipLaunchKernelGGL(k_e2_fill, dim3(nbl), dim3(kT), 0, st, nbl, w.nbr.p, out.box_first->p,
                     out.box_count->p, out.perm->p, out.e2_off->p, out.e2_list->p);
  TRY("e2", hipGetLastError());
  // per level: parent, octant, child, I4 (levels l_min..l_max)
  for (int l = 0; l <= lmax; ++l) lev_off[l + 1] = lev_off[l] + nlev[l];
  const int nall = lev_off[lmax + 1];
  TRY("alloc", w.lpar.alloc(nall));
  TRY("alloc", w.loct.alloc(nall));
  TRY("alloc", w.lchild.alloc(8 * (size_t)nall));
  TRY("alloc", w.li4off.alloc(nall + lmax + 1));
  TRY("alloc", w.i4cnt.alloc(nall + lmax + 1));
  std::vector<int> i4tot(lmax + 1, 0);
  for (int l = t->l_min; l <= lmax; ++l) {
    const int nb = nlev[l];
    const int64_t *Ln = w.levn.p + (size_t)l * N;
    const int64_t *Lp = l > t->l_min ? w.levn.p + (size_t)(l - 1) * N : nullptr;
    const int64_t *Lc = l < lmax ? w.levn.p + (size_t)(l + 1) * N : nullptr;
    hipLaunchKernelGGL(k_tree_links, dim3(nblk(nb)), dim3(kT), 0, st, nb, Ln, Lp,
                       l > 0 ? nlev[l - 1] : 0, Lc, l < lmax ? nlev[l + 1] : 0,
                       w.lpar.p + lev_off[l], w.loct.p + lev_off[l],
                       w.lchild.p + 8 * (size_t)lev_off[l]);
    int *cnt = w.i4cnt.p + lev_off[l] + l, *off = w.li4off.p + lev_off[l] + l;
    hipLaunchKernelGGL(k_i4, dim3(nb), dim3(kT), 0, st, nb, l, Ln, 0, cnt, nullptr, nullptr,
                       nullptr);
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, st, cnt + nb, 0);
    cub_bytes = w.cub.cap;
    TRY("scan", hipcub::DeviceScan::ExclusiveSum(w.cub.p, cub_bytes, cnt, off, nb + 1, st));
    TRY("i4", hipMemcpyAsync(&i4tot[l], off + nb, sizeof(int), hipMemcpyDeviceToHost, st));
  }
  TRY("i4", hipStreamSynchronize(st));
  std::vector<int> i4_at(lmax + 2, 0);
  for (int l = 0; l <= lmax; ++l) i4_at[l + 1] = i4_at[l] + i4tot[l];
  TRY("alloc", w.li4box.alloc(std::max(i4_at[lmax + 1], 1)));
  TRY("alloc", w.li4slot.alloc(std::max(i4_at[lmax + 1], 1)));
  for (int l = t->l_min; l <= lmax; ++l) {
    const int nb = nlev[l];
    const int64_t *Ln = w.levn.p + (size_t)l * N;
    hipLaunchKernelGGL(k_i4, dim3(nb), dim3(kT), 0, st, nb, l, Ln, 1, nullptr,
                       w.li4off.p + lev_off[l] + l, w.li4box.p + i4_at[l],
                       w.li4slot.p + i4_at[l]);
  }
  TRY("i4", hipGetLastError());
  // host copies
  t->lev.assign(lmax + 1, Level{});
  for (int l = 0; l <= lmax; ++l) {
    Level &L = t->lev[l];
    const int nb = nlev[l];
    TRY("copy", to_host(L.n, w.levn.p + (size_t)l * N, nb, st));
    if (l < t->l_min) continue;
    TRY("copy", to_host(L.parent, w.lpar.p + lev_off[l], nb, st));
    TRY("copy", to_host(L.octant, w.loct.p + lev_off[l], nb, st));
    if (l < lmax) TRY("copy", to_host(L.child, w.lchild.p + 8 * (size_t)lev_off[l], 8 * (size_t)nb, st));
    TRY("copy", to_host(L.i4_off, w.li4off.p + lev_off[l] + l, nb + 1, st));
    TRY("copy", to_host(L.i4_box, w.li4box.p + i4_at[l], i4tot[l], st));
    TRY("copy", to_host(L.i4_slot, w.li4slot.p + i4_at[l], i4tot[l], st));
  }
  TRY("copy", to_host(t->perm, out.perm->p, N, st));
  TRY("copy", to_host(t->pos, out.pos->p, N, st));
  TRY("copy", to_host(t->leaf_box, out.leaf_box->p, N, st));
  TRY("copy", to_host(t->box_first, out.box_first->p, nbl, st));
  TRY("copy", to_host(t->box_count, out.box_count->p, nbl, st));
  TRY("copy", to_host(t->e2_off, out.e2_off->p, nbl + 1, st));
  TRY("copy", to_host(t->e2_list, out.e2_list->p, e2_total, st));
  TRY("copy", to_host(t->near_off, out.near_off->p, nbl + 1, st));
  std::vector<uint64_t> lk;
  TRY("copy", to_host(lk, w.lk.p, N, st));
  TRY("copy", hipStreamSynchronize(st));
  t->leaf.assign(lk.begin(), lk.end());
  t->nnz = t->near_off[nbl];
  return hipSuccess;
}

}  // namespace spk

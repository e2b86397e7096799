// spk_quad.hip -- gfx950 kernels and C-ABI of the 8-node quadrilateral
// (10 Huber edge functions) impedance fill: impedance_edge_versus_element
// (src/lib_sie_quad_singularity.f90:22-98) for every (edge function,
// element) pair and the assembly of the 2M x 2M system
// (src/lib_sie_quad_solver_normal.f90:165-210).
//
// Pair classes and their rules, as in the reference:
//   regular (no shared node)  4 x 4 Gauss-Legendre on both elements
//   corner-adjacent           6 x 6 on both elements
//   self                      15 x 15 outer points, polar Taylor singularity
//                             subtraction inside (4 sectors x 15 x 5)
//   edge-adjacent             16 sub-domains x 5^4 (theta, Psi, mu, Lambda)
// Launch shapes: one thread per regular pair; one workgroup per self pair
// (225 outer points); one workgroup per edge-adjacent pair (16 cases x 5
// theta nodes).  Partial sums are reduced in a fixed order, so results are
// reproducible.
#include <algorithm>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "spk_quad.h"

namespace spk {

namespace {

constexpr double QPI = 3.14159265358979323846;

// device array owned by one call or one handle (freed on destruction)
template <class T>
struct QBuf {
  T *p = nullptr;
  size_t n = 0;
  ~QBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    reset();
    n = count;
    if (!count) return hipSuccess;
    return hipMalloc(&p, count * sizeof(T));
  }
  hipError_t upload(const std::vector<T> &v) {
    hipError_t e = alloc(v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
};


__device__ inline void qgreen(zd k, double R, zd &G, zd &DG) {
  // Diver_Green (src/lib_sie_quad_calculation_mod.f90:553-564)
  const double ph = -k.x * R, mag = exp(k.y * R);
  double sn, cs;
  sincos(ph, &sn, &cs);
  const zd E = {mag * cs, mag * sn};  // exp(-i k R)
  const zd ikr = {-k.y * R, k.x * R};
  G = E * (1.0 / (4 * QPI * R));
  DG = -((zd{1.0 + ikr.x, ikr.y} * E) * (1.0 / (4 * QPI * (R * R))));
}

// shape functions and derivatives (R_shape, N_differential, f_edge_new)
__device__ inline void shape_diff(double k, double e, double dk[8], double de[8]) {
  dk[0] = 0.25 * (1 - e) * (2 * k + e);
  dk[1] = -k * (1 - e);
  dk[2] = 0.25 * (1 - e) * (2 * k - e);
  dk[3] = 0.50 * (1 - e * e);
  dk[4] = 0.25 * (1 + e) * (2 * k + e);
  dk[5] = -k * (1 + e);
  dk[6] = 0.25 * (1 + e) * (2 * k - e);
  dk[7] = -0.50 * (1 - e * e);
  de[0] = 0.25 * (1 - k) * (2 * e + k);
  de[1] = -0.5 * (1 - k * k);
  de[2] = 0.25 * (1 + k) * (2 * e - k);
  de[3] = -e * (1 + k);
  de[4] = 0.25 * (1 + k) * (2 * e + k);
  de[5] = 0.5 * (1 - k * k);
  de[6] = 0.25 * (1 - k) * (2 * e - k);
  de[7] = -e * (1 - k);
}
__device__ inline void shape(double k, double e, double N[8]) {
  N[0] = 0.25 * (1 - k) * (1 - e) * (-1 - e - k);
  N[1] = 0.5 * (1 - k * k) * (1 - e);
  N[2] = 0.25 * (1 + k) * (1 - e) * (-1 - e + k);
  N[3] = 0.5 * (1 - e * e) * (1 + k);
  N[4] = 0.25 * (1 + k) * (1 + e) * (-1 + e + k);
  N[5] = 0.5 * (1 - k * k) * (1 + e);
  N[6] = 0.25 * (1 - k) * (1 + e) * (-1 + e - k);
  N[7] = 0.5 * (1 - e * e) * (1 - k);
}
__device__ inline double fedge(int n, double k, double e) {
  switch (n) {
    case 0: return -0.25 * (1 - e) * (2 * k + e);
    case 1: return 0.25 * (1 - e) * (2 * k - e);
    case 2: return -0.25 * (1 + e) * (2 * k - e);
    case 3: return 0.25 * (1 + e) * (2 * k + e);
    case 4: return 0.5 * (1 - e * e);
    case 5: return -0.25 * (1 - k) * (2 * e + k);
    case 6: return 0.25 * (1 - k) * (2 * e - k);
    case 7: return -0.25 * (1 + k) * (2 * e - k);
    case 8: return 0.25 * (1 + k) * (2 * e + k);
    default: return 0.5 * (1 - k * k);
  }
}
// n x curl of the edge functions' scalar parts (df/deta for 1-5, df/dksi for 6-10)
__device__ inline double fedge_rot(int n, double k, double e) {
  switch (n) {
    case 0: return -(0.25 * (2 * k + 2 * e - 1));
    case 1: return -(0.25 * (2 * e - 2 * k - 1));
    case 2: return -(0.25 * (2 * e - 2 * k + 1));
    case 3: return -(0.25 * (2 * k + 2 * e + 1));
    case 4: return -(-e);
    case 5: return 0.25 * (2 * e + 2 * k - 1);
    case 6: return 0.25 * (2 * k - 2 * e - 1);
    case 7: return 0.25 * (2 * k - 2 * e + 1);
    case 8: return 0.25 * (2 * k + 2 * e + 1);
    default: return -k;
  }
}

__device__ inline void cross3(const double a[3], const double b[3], double o[3]) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double dot3(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

struct Geo {
  double JJ, nq[3], au[3], av[3];  // unit normal, (Ev x n)/JJ and (n x Eu)/JJ
};

__device__ inline void geo(double k, double e, const double *El, Geo &g) {
  double dk[8], de[8], Eu[3] = {0, 0, 0}, Ev[3] = {0, 0, 0};
  shape_diff(k, e, dk, de);
  for (int d = 0; d < 3; ++d) {
    double su = 0, sv = 0;
    for (int q = 0; q < 8; ++q) {
      su += dk[q] * El[q * 3 + d];
      sv += de[q] * El[q * 3 + d];
    }
    Eu[d] = su;
    Ev[d] = sv;
  }
  double r[3];
  cross3(Eu, Ev, r);
  g.JJ = sqrt(dot3(r, r));
  for (int d = 0; d < 3; ++d) g.nq[d] = r[d] / g.JJ;
  double t[3];
  cross3(Ev, g.nq, t);
  for (int d = 0; d < 3; ++d) g.au[d] = t[d] / g.JJ;
  cross3(g.nq, Eu, t);
  for (int d = 0; d < 3; ++d) g.av[d] = t[d] / g.JJ;
}
__device__ inline void rloc(double k, double e, const double *El, double r[3]) {
  double N[8];
  shape(k, e, N);
  for (int d = 0; d < 3; ++d) {
    double s = 0;
    for (int q = 0; q < 8; ++q) s += N[q] * El[q * 3 + d];
    r[d] = s;
  }
}
// observation edge function pI (1-based) at (k, e)
__device__ inline void obs_fn(const Geo &g, int pI, double k, double e, double Na[3]) {
  const double f = fedge(pI - 1, k, e);
  const double *v = pI <= 5 ? g.au : g.av;
  for (int d = 0; d < 3; ++d) Na[d] = f * v[d];
}

// ff_ABC_II_Edge accumulated with weight w into S[5][10]
// (src/lib_sie_quad_singularity.f90:581-618)
__device__ inline void acc_ff(const QuadDev &D, const double Na[3], const double R[3],
                              const Geo &q, double kq, double eq, double w, zd S[5][10]) {
  const double nR = sqrt(dot3(R, R));
  const double rn[3] = {R[0] / nR, R[1] / nR, R[2] / nR};
  zd Ga, GSa, Gm, GSm;
  qgreen(D.k1, nR, Ga, GSa);
  qgreen(D.k2, nR, Gm, GSm);
  const zd gA = GSa + GSm, gBE = D.ie1 * GSa + D.ie2 * GSm, gC = Ga + Gm;
  const zd gCH = D.e1 * Ga + D.e2 * Gm;
  const double Nrn = dot3(Na, rn);
#pragma unroll
  for (int n = 0; n < 10; ++n) {
    const double f = fedge(n, kq, eq);
    const double *v = n < 5 ? q.au : q.av;
    const double Nq[3] = {f * v[0], f * v[1], f * v[2]};
    double aa[3], t[3];
    cross3(q.nq, Nq, aa);
    cross3(aa, rn, t);
    const double nrot = fedge_rot(n, kq, eq) / q.JJ;
    const double f1 = dot3(Na, t) * q.JJ;
    const double f2 = Nrn * nrot * q.JJ;
    const double f3 = dot3(Na, aa) * q.JJ;
    S[0][n] = S[0][n] + gA * (f1 * w);
    S[1][n] = S[1][n] + gBE * (f2 * w);
    S[2][n] = S[2][n] + gA * (f2 * w);
    S[3][n] = S[3][n] + gC * (f3 * w);
    S[4][n] = S[4][n] + gCH * (f3 * w);
  }
}

// (KA, KBE, KBH, KCE, KCH) -> (A, DE, DH)
__device__ inline void combine(const QuadDev &D, zd S[5][10], zd out[3][10]) {
#pragma unroll
  for (int n = 0; n < 10; ++n) {
    out[0][n] = S[0][n];
    out[1][n] = S[1][n] / D.iwe - (-(D.iwm * S[3][n]));
    out[2][n] = S[2][n] / (-D.iwm) - D.iwe * S[4][n];
  }
}

// regular pair: NG x NG Gauss on both elements (Element_integration 'Normal'
// and Integration_KABCD_Edge_new, :479-522)
template <int NG>
__device__ void regular_pair(const QuadDev &D, const double *x, const double *w, const double *Ea,
                             const double *Eq, int pI, zd out[3][10]) {
  zd S[5][10];
  for (int k = 0; k < 5; ++k)
    for (int n = 0; n < 10; ++n) S[k][n] = zd{0, 0};
  for (int j = 0; j < NG; ++j)
    for (int i = 0; i < NG; ++i) {
      Geo ga;
      geo(x[i], x[j], Ea, ga);
      double Na[3], Ra[3];
      obs_fn(ga, pI, x[i], x[j], Na);
      rloc(x[i], x[j], Ea, Ra);
      const double TRF = w[i] * w[j] * ga.JJ;
      for (int a = 0; a < NG; ++a)
        for (int b = 0; b < NG; ++b) {
          Geo gq;
          geo(x[a], x[b], Eq, gq);
          double Rq[3], R[3];
          rloc(x[a], x[b], Eq, Rq);
          for (int d = 0; d < 3; ++d) R[d] = Ra[d] - Rq[d];
          acc_ff(D, Na, R, gq, x[a], x[b], (w[a] * w[b]) * TRF, S);
        }
    }
  combine(D, S, out);
}

// Source-element data of the regular NG x NG rule at point (a, b), shared by
// every pair of a workgroup that integrates over the same element: position,
// rule weight and, per edge function n, JJ (n x N_n) and the n x curl factor
// fedge_rot (JJ times nrot of acc_ff).
struct QSrc {
  double R[3], w;
  double aJ[10][3];
  double fr[10];
};
template <int NG>
__device__ inline void qsrc_point(const QuadDev &D, const double *Eq, int a, int b, QSrc &o) {
  const double *x = NG == 6 ? D.x6 : D.x4, *wt = NG == 6 ? D.w6 : D.w4;
  const double k = x[a], e = x[b];
  Geo g;
  geo(k, e, Eq, g);
  rloc(k, e, Eq, o.R);
  o.w = wt[a] * wt[b];
#pragma unroll
  for (int n = 0; n < 10; ++n) {
    const double f = fedge(n, k, e);
    const double *v = n < 5 ? g.au : g.av;
    const double Nq[3] = {f * v[0], f * v[1], f * v[2]};
    double aa[3];
    cross3(g.nq, Nq, aa);
    for (int d = 0; d < 3; ++d) o.aJ[n][d] = aa[d] * g.JJ;
    o.fr[n] = fedge_rot(n, k, e);
  }
}

// one outer point (i, j) of the NG x NG rule on the test element against the
// source element's NG^2 points src (point a, b at src[NG a + b], the inner
// loop order of regular_pair), accumulated into S
template <int NG>
__device__ void outer_point_tab(const QuadDev &D, const double *Ea, const QSrc *src, int pI,
                                int i, int j, zd S[5][10]) {
  const double *x = NG == 6 ? D.x6 : D.x4, *wt = NG == 6 ? D.w6 : D.w4;
  Geo ga;
  geo(x[i], x[j], Ea, ga);
  double Na[3], Ra[3];
  obs_fn(ga, pI, x[i], x[j], Na);
  rloc(x[i], x[j], Ea, Ra);
  const double TRF = wt[i] * wt[j] * ga.JJ;
  for (int p = 0; p < NG * NG; ++p) {
    const QSrc &P = src[p];
    const double R[3] = {Ra[0] - P.R[0], Ra[1] - P.R[1], Ra[2] - P.R[2]};
    const double nR = sqrt(dot3(R, R));
    const double rn[3] = {R[0] / nR, R[1] / nR, R[2] / nR};
    zd Ga, GSa, Gm, GSm;
    qgreen(D.k1, nR, Ga, GSa);
    qgreen(D.k2, nR, Gm, GSm);
    const zd gA = GSa + GSm, gBE = D.ie1 * GSa + D.ie2 * GSm, gC = Ga + Gm;
    const zd gCH = D.e1 * Ga + D.e2 * Gm;
    const double Nrn = dot3(Na, rn), w = P.w * TRF;
#pragma unroll
    for (int n = 0; n < 10; ++n) {
      double t[3];
      cross3(P.aJ[n], rn, t);
      const double f1 = dot3(Na, t), f2 = Nrn * P.fr[n], f3 = dot3(Na, P.aJ[n]);
      S[0][n] = S[0][n] + gA * (f1 * w);
      S[1][n] = S[1][n] + gBE * (f2 * w);
      S[2][n] = S[2][n] + gA * (f2 * w);
      S[3][n] = S[3][n] + gC * (f3 * w);
      S[4][n] = S[4][n] + gCH * (f3 * w);
    }
  }
}

// regular_pair<4> with the source element's 16 points read from src
__device__ void regular_pair_tab(const QuadDev &D, const double *Ea, const QSrc *src, int pI,
                                 zd out[3][10]) {
  zd S[5][10];
  for (int k = 0; k < 5; ++k)
    for (int n = 0; n < 10; ++n) S[k][n] = zd{0, 0};
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i) outer_point_tab<4>(D, Ea, src, pI, i, j, S);
  combine(D, S, out);
}

__device__ inline int pair_class(const QuadTopo &T, int s, int t, int *side) {
  const int ii = T.edge_p[5 * s] - 1;
  if (t == ii) return 0;
  for (int a = 0; a < kQuadMaxAdj; ++a)
    if (T.eadj[ii * kQuadMaxAdj + a] == t + 1) {
      side[0] = T.eside[(ii * kQuadMaxAdj + a) * 2];
      side[1] = T.eside[(ii * kQuadMaxAdj + a) * 2 + 1];
      return 1;
    }
  for (int a = 0; a < kQuadMaxAdj; ++a)
    if (T.cadj[ii * kQuadMaxAdj + a] == t + 1) return 2;
  return 3;
}

__device__ inline void store30(zd out[3][10], zd *dst) {
  for (int k = 0; k < 3; ++k)
    for (int n = 0; n < 10; ++n) dst[k * 10 + n] = out[k][n];
}

// dense fill, rows [s0, s0 + rows): workgroup = (element t, 64 rows), the
// regular pairs (no shared node) on the element's source table; self,
// edge- and corner-adjacent pairs come from the list kernels
__global__ void __launch_bounds__(64) k_quad_regular_grid(QuadDev D, QuadTopo T, int s0, int rows,
                                                          int nrb, zd *sums) {
  __shared__ QSrc src[16];
  const int t = blockIdx.x / nrb, rb = blockIdx.x - t * nrb;
  if (threadIdx.x < 16)
    qsrc_point<4>(D, T.el + (size_t)t * 24, threadIdx.x >> 2, threadIdx.x & 3, src[threadIdx.x]);
  __syncthreads();
  const int r = rb * 64 + threadIdx.x;
  if (r >= rows) return;
  const int s = s0 + r;
  int side[2];
  if (pair_class(T, s, t, side) != 3) return;
  const int ii = T.edge_p[5 * s] - 1, pI = T.edge_p[5 * s + 1];
  zd out[3][10];
  regular_pair_tab(D, T.el + (size_t)ii * 24, src, pI, out);
  store30(out, sums + ((size_t)r * T.nel + t) * 30);
}

__global__ void __launch_bounds__(128) k_quad_regular_list(QuadDev D, QuadTopo T, int n,
                                                           const int *ls, const int *lt,
                                                           const int *lg, zd *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = ls[i], t = lt[i];
  const int ii = T.edge_p[5 * s] - 1, pI = T.edge_p[5 * s + 1];
  zd o[3][10];
  if (lg[i] == 6)
    regular_pair<6>(D, D.x6, D.w6, T.el + (size_t)ii * 24, T.el + (size_t)t * 24, pI, o);
  else
    regular_pair<4>(D, D.x4, D.w4, T.el + (size_t)ii * 24, T.el + (size_t)t * 24, pI, o);
  store30(o, out + (size_t)i * 30);
}

// fixed-order reduction of 30 complex values over a workgroup
__device__ inline zd wsum(zd v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  return v;
}
__device__ void block_reduce30(zd v[3][10], zd *dst, double scale = 1.0) {
  __shared__ zd red[4][30];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int k = 0; k < 3; ++k)
    for (int n = 0; n < 10; ++n) {
      const zd s = wsum(scale == 1.0 ? v[k][n] : v[k][n] * scale);
      if (lane == 0) red[wv][k * 10 + n] = s;
    }
  __syncthreads();
  if (threadIdx.x < 30) {
    zd s = red[0][threadIdx.x];
    for (int q = 1; q < nw; ++q) s = s + red[q][threadIdx.x];
    dst[threadIdx.x] = s;
  }
  __syncthreads();
}

// corner-adjacent pairs (the 6 x 6 rule of regular_pair<6>): one workgroup
// per listed pair (s[i], t[i]), the source points in LDS, one thread per
// outer point, summed in a fixed order; out[i*30]
__global__ void __launch_bounds__(64) k_quad_corner(QuadDev D, QuadTopo T, const int *ls,
                                                    const int *lt, zd *out) {
  __shared__ QSrc src[36];
  const int s = ls[blockIdx.x], t = lt[blockIdx.x];
  if (threadIdx.x < 36)
    qsrc_point<6>(D, T.el + (size_t)t * 24, threadIdx.x / 6, threadIdx.x % 6, src[threadIdx.x]);
  __syncthreads();
  const int ii = T.edge_p[5 * s] - 1, pI = T.edge_p[5 * s + 1];
  zd S[5][10];
  for (int k = 0; k < 5; ++k)
    for (int n = 0; n < 10; ++n) S[k][n] = zd{0, 0};
  if (threadIdx.x < 36)
    outer_point_tab<6>(D, T.el + (size_t)ii * 24, src, pI, threadIdx.x % 6, threadIdx.x / 6, S);
  zd r[3][10];
  combine(D, S, r);
  block_reduce30(r, out + (size_t)blockIdx.x * 30);
}

// self term at one outer point: Integration_KABCD_Taylor_ST_4FEdge (:299-380),
// for the test-function direction dir (0: au, the functions 1-5; 1: av,
// 6-10) at unit value; the term is linear in the test function's value, so
// function pI's is fedge(pI - 1, mu, nu) times the result of its direction
__device__ void taylor_point(const QuadDev &D, int dir, double mu, double nu, const double *El,
                             zd out[3][10]) {
  double dk[8], de[8], Eu[3] = {0, 0, 0}, Ev[3] = {0, 0, 0}, Ra[3];
  shape_diff(mu, nu, dk, de);
  for (int d = 0; d < 3; ++d) {
    for (int q = 0; q < 8; ++q) {
      Eu[d] += dk[q] * El[q * 3 + d];
      Ev[d] += de[q] * El[q * 3 + d];
    }
  }
  Geo gp;
  geo(mu, nu, El, gp);
  rloc(mu, nu, El, Ra);
  const double *Na = dir ? gp.av : gp.au;
  const double pb[5] = {atan((1 - nu) / (1 - mu)), QPI - atan((1 - nu) / (1 + mu)),
                        QPI + atan((1 + nu) / (1 + mu)), 2 * QPI - atan((1 + nu) / (1 - mu)),
                        2 * QPI + atan((1 - nu) / (1 - mu))};
  zd S[5][10];
  for (int k = 0; k < 5; ++k)
    for (int n = 0; n < 10; ++n) S[k][n] = zd{0, 0};
  for (int m = 0; m < 4; ++m) {
    const double a = (pb[m + 1] - pb[m]) / 2, cc = (pb[m + 1] + pb[m]) / 2;
    for (int i = 0; i < 15; ++i) {
      const double phi = cc + a * D.x15[i];
      double sp, cp;
      sincos(phi, &sp, &cp);
      double A[3];
      for (int d = 0; d < 3; ++d) A[d] = -Eu[d] * cp - Ev[d] * sp;
      const double Aamp = sqrt(dot3(A, A));
      double rho2;
      if (m == 0) rho2 = (1 - nu) / sin(QPI - phi);
      else if (m == 1) rho2 = fabs((1 + mu) / cos(QPI - phi));
      else if (m == 2) rho2 = fabs((1 + nu) / cos(1.5 * QPI - phi));
      else rho2 = fabs((1 - mu) / cos(phi));
      // ff_phi_AB_Edge (:621-637): the 1/rho parts of KA and KB
      double bb[3];
      for (int d = 0; d < 3; ++d) bb[d] = -A[d] / (Aamp * Aamp * Aamp) / (4 * QPI);
      double fA[10], fB[10];
      for (int n = 0; n < 10; ++n) {
        const double f = fedge(n, mu, nu);
        const double *v = n < 5 ? gp.au : gp.av;
        const double Nq[3] = {f * v[0], f * v[1], f * v[2]};
        double aa[3], t[3];
        cross3(gp.nq, Nq, aa);
        cross3(aa, bb, t);
        fA[n] = dot3(Na, t) * gp.JJ;
        fB[n] = dot3(Na, bb) * gp.JJ * (fedge_rot(n, mu, nu) / gp.JJ);
      }
      for (int j = 0; j < 5; ++j) {
        const double rho = rho2 / 2 * (1 + D.x5[j]);
        const double mq = mu + rho * cp, nq = nu + rho * sp;
        Geo gq;
        geo(mq, nq, El, gq);
        double Rq[3], R[3];
        rloc(mq, nq, El, Rq);
        for (int d = 0; d < 3; ++d) R[d] = Ra[d] - Rq[d];
        const double TRF = 0.5 * rho2 * D.w15[i] * D.w5[j] * a;
        acc_ff(D, Na, R, gq, mq, nq, rho * TRF, S);
        for (int n = 0; n < 10; ++n) {
          S[0][n].x -= 2 * fA[n] / rho * TRF;
          S[1][n] = S[1][n] - D.ceps * (fB[n] / rho * TRF);
          S[2][n].x -= 2 * fB[n] / rho * TRF;
        }
      }
      const double T0 = log(fabs(rho2) * Aamp) * D.w15[i] * a;
      for (int n = 0; n < 10; ++n) {
        S[0][n].x += 2 * fA[n] * T0;
        S[1][n] = S[1][n] + D.ceps * (fB[n] * T0);
        S[2][n].x += 2 * fB[n] * T0;
      }
    }
  }
  combine(D, S, out);
}

// self terms of all ten test functions of element els[blockIdx.x] against
// the element itself, out[blockIdx.x][pI - 1][30]: one thread per outer
// point of the 15 x 15 rule, each direction integrated once
__global__ void __launch_bounds__(256) k_quad_self_el(QuadDev D, QuadTopo T, const int *els,
                                                      zd *out) {
  const double *El = T.el + (size_t)els[blockIdx.x] * 24;
  const int p = threadIdx.x;
  const bool on = p < 225;
  const double mu = D.x15[on ? p % 15 : 0], nu = D.x15[on ? p / 15 : 0];
  double TRF = 0;
  if (on) {
    Geo ga;
    geo(mu, nu, El, ga);
    TRF = D.w15[p % 15] * D.w15[p / 15] * ga.JJ;
  }
#pragma unroll 1
  for (int dir = 0; dir < 2; ++dir) {
    zd r[3][10];
    for (int k = 0; k < 3; ++k)
      for (int n = 0; n < 10; ++n) r[k][n] = zd{0, 0};
    if (on) taylor_point(D, dir, mu, nu, El, r);
#pragma unroll 1
    for (int pI = 1 + 5 * dir; pI <= 5 + 5 * dir; ++pI) {
      const double f = on ? fedge(pI - 1, mu, nu) * TRF : 0.0;
      block_reduce30(r, out + ((size_t)blockIdx.x * 10 + pI - 1) * 30, f);
    }
  }
}

// x[k] = z[perm[k]]
__global__ void k_quad_permute(int64_t n, const int64_t *perm, const zd *z, zd *x) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) x[k] = z[perm[k]];
}

// out[i][:] = src[slot[i]][:], 30 complex each
__global__ void k_quad_pick(int n, const int *slot, const zd *src, zd *out) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)n * 30) return;
  const int i = (int)(id / 30), k = (int)(id % 30);
  out[(size_t)i * 30 + k] = src[(size_t)slot[i] * 30 + k];
}

// edge-adjacent sub-domain limits (Lm_theta/Psi/mu/Lambd_EA, :656-923)
__device__ inline double ea_theta(int ll, int mm, int nn) {
  if (ll == 0) return (mm == 0) ? ((nn <= 2) ? 0 : QPI / 4) : ((nn == 1) ? 0 : QPI / 4);
  if (ll == 1) return (nn == 1) ? QPI / 4 : QPI / 2;
  if (mm == 0) return (nn <= 2) ? QPI / 2 : 3 * QPI / 4;
  return (nn == 1) ? QPI / 2 : 3 * QPI / 4;
}
__device__ inline void ea_psi(int ll, int mm, int nn, double th, double ab[2]) {
  const double p1 = atan(cos(th)), p2 = atan(sin(th));
  if (ll == 0 && mm == 0) {
    if (nn == 1) { ab[0] = 0; ab[1] = p1; }
    else if (nn == 2) { ab[0] = p1; ab[1] = QPI / 2; }
    else if (nn == 3) { ab[0] = 0; ab[1] = p2; }
    else { ab[0] = p2; ab[1] = QPI / 2; }
  } else if (ll == 0) {
    ab[0] = (nn == 1) ? p1 : p2; ab[1] = QPI / 2;
  } else if (ll == 1 && mm == 0) {
    ab[0] = 0; ab[1] = p2;
  } else if (ll == 1) {
    ab[0] = p2; ab[1] = QPI / 2;
  } else if (mm == 0) {
    if (nn == 1) { ab[0] = 0; ab[1] = p2; }
    else if (nn == 2) { ab[0] = p2; ab[1] = QPI / 2; }
    else if (nn == 3) { ab[0] = 0; ab[1] = -p1; }
    else { ab[0] = -p1; ab[1] = QPI / 2; }
  } else {
    ab[0] = (nn == 1) ? p2 : -p1; ab[1] = QPI / 2;
  }
}
__device__ inline void ea_mu(int ll, int mm, int nn, double th, double ps, double ab[2]) {
  const double tp = tan(QPI / 2 - ps), tt = tan(QPI / 2 - th);
  const double u1p = 2 * cos(th) * tp - 1, u2p = 2 * cos(th) * tp + 1;
  const double u1t = 2 * tt - 1, u2t = 2 * tt + 1;
  if (ll == 0 && mm == 0) {
    ab[0] = -1;
    ab[1] = (nn == 1) ? 1 : (nn == 3) ? u1t : u1p;
  } else if (ll == 0) {
    ab[0] = u1p; ab[1] = (nn == 1) ? 1 : u1t;
  } else if (ll == 1) {
    if (nn == 1) { ab[0] = u1t; ab[1] = 1; }
    else { ab[0] = -1; ab[1] = u2t; }
  } else if (mm == 0) {
    ab[1] = 1;
    ab[0] = (nn == 1) ? u2t : (nn == 3) ? -1 : u2p;
  } else {
    ab[0] = (nn == 1) ? u2t : -1; ab[1] = u2p;
  }
}
__device__ inline double ea_lambda(int ll, int mm, double th, double ps, double u) {
  if (mm == 1) return 2 / sin(ps);
  if (ll == 0) return (1 + u) / (cos(th) * cos(ps));
  if (ll == 1) return 2 / (sin(th) * cos(ps));
  return (u - 1) / (cos(th) * cos(ps));
}

// Element_rotation_edge (:387-439)
__device__ inline int rot_node(int qI, int j) {  // output position (1-based) of input node j
  if (qI == 3 || qI == 4) return j <= 4 ? 4 + j : j - 4;
  if (qI == 6 || qI == 7) return j < 7 ? j + 2 : j - 6;
  if (qI == 8 || qI == 9) return j > 2 ? j - 2 : j + 6;
  return j;
}
__device__ inline void rot_edge(int qI, int n, int *idx, int *sgn) {  // n 1-based
  const int r34[10] = {4, 3, 2, 1, 5, 9, 8, 7, 6, 10};
  const int r67[10] = {8, 9, 6, 7, 10, 2, 1, 4, 3, 5};
  const int s67[10] = {1, 1, 1, 1, 1, -1, -1, -1, -1, -1};
  const int r89[10] = {7, 6, 9, 8, 10, 3, 4, 1, 2, 5};
  const int s89[10] = {-1, -1, -1, -1, -1, 1, 1, 1, 1, 1};
  if (qI == 3 || qI == 4) { *idx = r34[n - 1]; *sgn = -1; }
  else if (qI == 6 || qI == 7) { *idx = r67[n - 1]; *sgn = s67[n - 1]; }
  else if (qI == 8 || qI == 9) { *idx = r89[n - 1]; *sgn = s89[n - 1]; }
  else { *idx = n; *sgn = 1; }
}

__global__ void __launch_bounds__(128) k_quad_edge_adj(QuadDev D, QuadTopo T, const int *ls,
                                                       const int *lt, const int *lside,
                                                       zd *out) {
  const int task = blockIdx.x;
  const int s = ls[task], t = lt[task];
  const int sa = lside[2 * task], sq = lside[2 * task + 1];
  const int ii = T.edge_p[5 * s] - 1, pI = T.edge_p[5 * s + 1];
  __shared__ double Ea[24], Eq[24];
  if (threadIdx.x < 8) {
    const int j = threadIdx.x + 1;
    const int pa = rot_node(sa, j) - 1, pq = rot_node(sq, j) - 1;
    for (int d = 0; d < 3; ++d) {
      Ea[pa * 3 + d] = T.el[(size_t)ii * 24 + (j - 1) * 3 + d];
      Eq[pq * 3 + d] = T.el[(size_t)t * 24 + (j - 1) * 3 + d];
    }
  }
  __syncthreads();
  int pIa, sgn_a;
  rot_edge(sa, pI, &pIa, &sgn_a);
  zd v[3][10];
  for (int k = 0; k < 3; ++k)
    for (int n = 0; n < 10; ++n) v[k][n] = zd{0, 0};
  if (threadIdx.x < 80) {
    // case c: (mm, ll, nn) in the reference's loop order; theta node i4
    const int c = threadIdx.x / 5, i4 = threadIdx.x % 5;
    int mm, ll, nn;
    if (c < 10) {
      mm = 0;
      if (c < 4) { ll = 0; nn = c + 1; }
      else if (c < 6) { ll = 1; nn = c - 3; }
      else { ll = 2; nn = c - 5; }
    } else {
      mm = 1;
      ll = (c - 10) / 2;
      nn = (c - 10) % 2 + 1;
    }
    const double *x = D.x5, *w = D.w5;
    const double ta = ea_theta(ll, mm, nn), tb = ta + QPI / 4;
    const double tam = (tb - ta) / 2, tbm = (tb + ta) / 2;
    const double th = tam * x[i4] + tbm;
    double pab[2];
    ea_psi(ll, mm, nn, th, pab);
    const double pam = (pab[1] - pab[0]) / 2, pbm = (pab[1] + pab[0]) / 2;
    zd S[5][10];
    for (int k = 0; k < 5; ++k)
      for (int n = 0; n < 10; ++n) S[k][n] = zd{0, 0};
    for (int i3 = 0; i3 < 5; ++i3) {
      const double ps = pam * x[i3] + pbm;
      double mab[2];
      ea_mu(ll, mm, nn, th, ps, mab);
      const double mam = (mab[1] - mab[0]) / 2, mbm = (mab[1] + mab[0]) / 2;
      double sps, cps, sth, cth;
      sincos(ps, &sps, &cps);
      sincos(th, &sth, &cth);
      for (int i2 = 0; i2 < 5; ++i2) {
        const double u = mam * x[i2] + mbm;
        const double lb = ea_lambda(ll, mm, th, ps, u);
        const double lam = lb / 2, lbm = lb / 2;
        for (int i1 = 0; i1 < 5; ++i1) {
          const double X = lam * x[i1] + lbm;
          const double mua = u, nua = -1 + X * sps;
          const double muq = X * cps * cth - mua, nuq = X * cps * sth - 1;
          Geo ga, gq;
          geo(mua, nua, Ea, ga);
          geo(muq, nuq, Eq, gq);
          double Na[3], Ra[3], Rq[3], R[3];
          obs_fn(ga, pIa, mua, nua, Na);
          rloc(mua, nua, Ea, Ra);
          rloc(muq, nuq, Eq, Rq);
          for (int d = 0; d < 3; ++d) R[d] = Ra[d] - Rq[d];
          const double TRF = lam * w[i1] * ga.JJ * (X * X) * cps;
          acc_ff(D, Na, R, gq, muq, nuq, (tam * w[i4]) * (pam * w[i3]) * (mam * w[i2]) * TRF, S);
        }
      }
    }
    zd r[3][10];
    combine(D, S, r);
    // back to the unrotated source functions
    for (int n = 1; n <= 10; ++n) {
      int iq, sq_;
      rot_edge(sq, n, &iq, &sq_);
      for (int k = 0; k < 3; ++k) v[k][n - 1] = r[k][iq - 1] * (double)(sq_ * sgn_a);
    }
  }
  block_reduce30(v, out + (size_t)task * 30);
}

__global__ void k_quad_assemble(QuadTopo T, int form, zd eta_a, zd eta12, int s0, int rows,
                                const zd *sums, zd *MM) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int M = T.medge;
  if (id >= (int64_t)rows * M) return;
  const int r = (int)(id / M), tc = (int)(id % M);
  const int *e = T.edge_p + 5 * tc;
  const int ii = e[0] - 1, pa = e[1] - 1, jj = e[2] - 1, pq = e[3] - 1, f = e[4];
  zd m3[3];
  for (int k = 0; k < 3; ++k) {
    zd v = sums[((size_t)r * T.nel + ii) * 30 + k * 10 + pa];
    if (f != 0) v = v + sums[((size_t)r * T.nel + jj) * 30 + k * 10 + pq] * (double)f;
    m3[k] = v;
  }
  zd a = m3[0], b = m3[1], c = m3[2], d = m3[0];
  if (form == SPK_ICTF) {
    a = m3[0] / eta_a;
    b = m3[1] / eta_a;
    c = m3[2] * eta_a;
    d = m3[0] * eta_a;
  } else if (form == SPK_MCTF) {
    c = m3[2] * eta12;
    d = m3[0] * eta12;
  }
  const size_t s = (size_t)(s0 + r), n2 = 2 * (size_t)M;
  MM[s * n2 + tc] = a;
  MM[s * n2 + M + tc] = b;
  MM[(M + s) * n2 + tc] = c;
  MM[(M + s) * n2 + M + tc] = d;
}

// Plane-wave right-hand side (get_vector_b_quad with incident_int,
// src/lib_sie_quad_calculation_mod.f90:91-138, 1207-1240): edge function s
// is tested on its own element edge_p(1) only, 6 x 6 Gauss points; the
// phase is exp(-i conj(k1) k.R) (Fortran's dot_product conjugates the
// complex first argument of dot_product(k1*k_in, R_a)); H = (k x E)/(my0 c0).
// One thread per edge function.
__global__ void k_quad_rhs_plane(QuadDev D, QuadTopo T, QuadRhsArgs a, zd *b) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= T.medge) return;
  const int ii = T.edge_p[5 * s] - 1, pI = T.edge_p[5 * s + 1];
  double El[24];
  for (int q = 0; q < 24; ++q) El[q] = T.el[(size_t)ii * 24 + q];
  zd SE = {0, 0}, SH = {0, 0};
  for (int j = 0; j < 6; ++j)
    for (int i = 0; i < 6; ++i) {
      const double kk = D.x6[i], ee = D.x6[j];
      Geo g;
      geo(kk, ee, El, g);
      double R[3], Na[3];
      rloc(kk, ee, El, R);
      obs_fn(g, pI, kk, ee, Na);
      const double TRF = D.w6[i] * D.w6[j] * g.JJ;
      const double kr = a.kin[0] * R[0] + a.kin[1] * R[1] + a.kin[2] * R[2];
      // exp(-i k1c kr) = exp(Im(k1c) kr) (cos(Re k1c kr) - i sin(Re k1c kr))
      double sn, cs;
      sincos(a.k1c.x * kr, &sn, &cs);
      const double mag = exp(a.k1c.y * kr);
      const zd e = {mag * cs, -mag * sn};
      const zd E[3] = {e * a.ein[0], e * a.ein[1], e * a.ein[2]};
      const zd H[3] = {(E[2] * a.kin[1] - E[1] * a.kin[2]) * (1.0 / a.c0my),
                       (E[0] * a.kin[2] - E[2] * a.kin[0]) * (1.0 / a.c0my),
                       (E[1] * a.kin[0] - E[0] * a.kin[1]) * (1.0 / a.c0my)};
      SE = SE + (E[0] * Na[0] + E[1] * Na[1] + E[2] * Na[2]) * TRF;
      SH = SH + (H[0] * Na[0] + H[1] * Na[1] + H[2] * Na[2]) * TRF;
    }
  b[s] = SE * a.sE;
  b[T.medge + s] = SH * a.sH;
}

inline unsigned nb(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// sums[r[i]][t[i]][0:30] = src[i][0:30]
__global__ void k_quad_scatter(int n, int nel, const int *r, const int *t, const zd *src, zd *sums) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)n * 30) return;
  const int i = (int)(id / 30), k = (int)(id % 30);
  sums[((size_t)r[i] * nel + t[i]) * 30 + k] = src[(size_t)i * 30 + k];
}

// ---------------------------------------------------------------------------
// MLFMM over the quadrilateral edge functions.  The reference's quadrilateral
// path is dense only (lib_sie_quad_solver_normal.f90:62-270); its MLFMM
// data container declares a pattern slot for it (preCalculated_RT_quad,
// src/lib_sie_data_container.f90:84) that nothing fills.  Here the edge
// functions sit in the same octree as RWG edges (at the midpoint of their
// node pair, find_midpoint_edge_quad, lib_sie_quad_calculation_mod.f90:210-255)
// and the far part of the quadrilateral system is the plane-wave form of the
// regular 4 x 4 element-pair rule the dense fill uses for distant pairs
// (Integration_KABCD_Edge_new with ff_ABC_II_Edge):
//   G(ra - rq) = sum_s TL_s exp(-ik khat.(ra - ca)) exp(+ik khat.(rq - cq))
//   grad_a G   = sum_s TL_s (-ik khat) ...
// so with the receiving pattern R (the test function on its own element),
// the radiating patterns TJ (n x N over the function's one or two elements)
// and TD (its surface divergence), per medium:
//   KA = sum TL (-ik) R.(TJ x khat),  KB = sum TL (-ik)(khat.R) TD,
//   KC = sum TL R.TJ
// and the rows combine as the dense fill's (A, DE, DH) do.
// ---------------------------------------------------------------------------

// Quadrature points of the patterns, per position p (edge function perm[p]),
// relative to the leaf box centre: 16 receiving points (d, w JJ Na) and 32
// radiating points (d, w JJ (n x N), w JJ nrot); the second element's 16
// are zero-weighted when the function has one element.
constexpr int kQRecv = 16, kQRad = 32, kQPtStride = kQuadPtStride;
static_assert(kQPtStride == kQRecv * 6 + kQRad * 7, "point layout");

__global__ void k_qpoints(QuadDev D, QuadTopo T, int N, const int *perm, const int *leaf_box,
                          const double *leaf_centre, double *pts) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int m = perm[p];
  const double *c = leaf_centre + 3 * (size_t)leaf_box[m];
  const int *e = T.edge_p + 5 * (size_t)m;
  const int ii = e[0] - 1, pI = e[1], jj = e[2] - 1, pq = e[3];
  const double f = (double)e[4];
  double *o = pts + (size_t)p * kQPtStride;
  const double *Ea = T.el + (size_t)ii * 24;
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i) {
      Geo g;
      geo(D.x4[i], D.x4[j], Ea, g);
      double r[3], Na[3];
      rloc(D.x4[i], D.x4[j], Ea, r);
      obs_fn(g, pI, D.x4[i], D.x4[j], Na);
      const double w = D.w4[i] * D.w4[j] * g.JJ;
      double *q = o + (j * 4 + i) * 6;
      for (int d = 0; d < 3; ++d) {
        q[d] = r[d] - c[d];
        q[3 + d] = w * Na[d];
      }
    }
  double *o2 = o + kQRecv * 6;
  for (int h = 0; h < 2; ++h) {
    const bool on = h == 0 || (jj >= 0 && e[4] != 0);
    const double *Eq = T.el + (size_t)(h == 0 ? ii : (on ? jj : ii)) * 24;
    const int n = (h == 0 ? pI : pq) - 1;
    const double sg = h == 0 ? 1.0 : f;
    for (int j = 0; j < 4; ++j)
      for (int i = 0; i < 4; ++i) {
        double *q = o2 + (h * 16 + j * 4 + i) * 7;
        if (!on) {
          for (int d = 0; d < 7; ++d) q[d] = 0;
          continue;
        }
        const double kq = D.x4[i], eq = D.x4[j];
        Geo g;
        geo(kq, eq, Eq, g);
        double r[3];
        rloc(kq, eq, Eq, r);
        const double fv = fedge(n, kq, eq);
        const double *v = n < 5 ? g.au : g.av;
        const double Nq[3] = {fv * v[0], fv * v[1], fv * v[2]};
        double aa[3];
        cross3(g.nq, Nq, aa);
        const double w = D.w4[i] * D.w4[j] * sg;
        for (int d = 0; d < 3; ++d) {
          q[d] = r[d] - c[d];
          q[3 + d] = w * g.JJ * aa[d];
        }
        // nrot * JJ = the scalar part's derivative (fedge_rot)
        q[6] = w * fedge_rot(n, kq, eq);
      }
  }
}

// exp(sg * i k x) for complex k
__device__ inline zd expik(zd k, double x, double sg) {
  const double mag = exp(-sg * k.y * x);
  double sn, cs;
  sincos(sg * k.x * x, &sn, &cs);
  return zd{mag * cs, mag * sn};
}

// rt[p][q][s]: q = k*4 + {TJx, TJy, TJz, TD} (radiating, exp(+ik khat.d)),
// 8 + k*3 + {Rx, Ry, Rz} (receiving, exp(-ik khat.d)); one thread per
// (sample, position)
__global__ void k_qpatterns(int N, int S, const double *khat, const double *pts, zd k1, zd k2,
                            zd *rt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (s >= S || p >= N) return;
  const double kh[3] = {khat[3 * s], khat[3 * s + 1], khat[3 * s + 2]};
  const double *o = pts + (size_t)p * kQPtStride;
  zd *out = rt + (size_t)p * 14 * S + s;
  for (int k = 0; k < 2; ++k) {
    const zd kk = k ? k2 : k1;
    zd R[3] = {{0, 0}, {0, 0}, {0, 0}}, TJ[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int i = 0; i < kQRecv; ++i) {
      const double *q = o + i * 6;
      const zd ph = expik(kk, kh[0] * q[0] + kh[1] * q[1] + kh[2] * q[2], -1.0);
      for (int d = 0; d < 3; ++d) R[d] = R[d] + ph * q[3 + d];
    }
    const double *o2 = o + kQRecv * 6;
    for (int i = 0; i < kQRad; ++i) {
      const double *q = o2 + i * 7;
      const zd ph = expik(kk, kh[0] * q[0] + kh[1] * q[1] + kh[2] * q[2], 1.0);
      for (int d = 0; d < 4; ++d) TJ[d] = TJ[d] + ph * q[3 + d];
    }
    for (int d = 0; d < 4; ++d) out[(size_t)(k * 4 + d) * S] = TJ[d];
    for (int d = 0; d < 3; ++d) out[(size_t)(8 + k * 3 + d) * S] = R[d];
  }
}

// Aggregation: B[half c/2][b][(c%2)*4 + cur*2 + k][s] = sum over the box's
// positions of TJ/TD[k][c](s) * x_cur; grid (samples, leaf boxes, k*4 + c)
template <class T>
__global__ void k_qagg(int S, int N, const int *box_first, const int *box_count, const int *perm,
                       const T *rt, const T *x, T *B, size_t half) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y, q = blockIdx.z;
  if (s >= S) return;
  const int k = q >> 2, c = q & 3;
  T ae = {0, 0}, ah = {0, 0};
  const int p0 = box_first[b], cnt = box_count[b];
  for (int e0 = 0; e0 < cnt; e0 += kSumBlock<T>) {
    T ge = {0, 0}, gh = {0, 0};
    const int e1 = min(cnt, e0 + kSumBlock<T>);
    for (int e = e0; e < e1; ++e) {
      const int p = p0 + e, m = perm[p];
      const T f = rt[((size_t)p * 14 + q) * S + s];
      ge = ge + f * x[m];
      gh = gh + f * x[m + N];
    }
    ae = ae + ge;
    ah = ah + gh;
  }
  T *Bb = B + (size_t)(c >> 1) * half + (size_t)b * 8 * S;
  Bb[(size_t)((c & 1) * 4 + 0 + k) * S + s] = ae;
  Bb[(size_t)((c & 1) * 4 + 2 + k) * S + s] = ah;
}

template <class T> struct qreal;
template <> struct qreal<zd> { using type = double; };
template <> struct qreal<zf> { using type = float; };
template <class T> using qreal_t = typename qreal<T>::type;

template <class T>
__device__ inline void cross_k(const T J[3], const double kh[3], T o[3]) {
  using R = qreal_t<T>;
  o[0] = J[1] * (R)kh[2] - J[2] * (R)kh[1];
  o[1] = J[2] * (R)kh[0] - J[0] * (R)kh[2];
  o[2] = J[0] * (R)kh[1] - J[1] * (R)kh[0];
}

// The receiving sums of a leaf box combined per medium into the two
// vectors the rows read (W[b][k*6 + {E x,y,z, H x,y,z}][s]):
//   WE = cAE (DJe x khat) + cJE DJh + cDE DDh khat
//   WH = cJH DJe + cDH DDe khat + cAH (DJh x khat)
template <class T>
__global__ void k_qfold(int S, const T *D, size_t half, const double *khat, QuadFoldC C, T *W) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (s >= S) return;
  const double kh[3] = {khat[3 * s], khat[3 * s + 1], khat[3 * s + 2]};
  const T *D0 = D + (size_t)b * 8 * S, *D1 = D0 + half;
  T *Wb = W + (size_t)b * 12 * S;
  for (int k = 0; k < 2; ++k) {
    const T Je[3] = {D0[(size_t)(0 + k) * S + s], D0[(size_t)(4 + k) * S + s],
                     D1[(size_t)(0 + k) * S + s]};
    const T Jh[3] = {D0[(size_t)(2 + k) * S + s], D0[(size_t)(6 + k) * S + s],
                     D1[(size_t)(2 + k) * S + s]};
    const T De = D1[(size_t)(4 + k) * S + s], Dh = D1[(size_t)(6 + k) * S + s];
    T xe[3], xh[3];
    cross_k(Je, kh, xe);
    cross_k(Jh, kh, xh);
    const T cAE = cvt<T>(C.cAE[k]), cJE = cvt<T>(C.cJE[k]), cDE = cvt<T>(C.cDE[k]);
    const T cJH = cvt<T>(C.cJH[k]), cDH = cvt<T>(C.cDH[k]), cAH = cvt<T>(C.cAH[k]);
    const T dE = cDE * Dh, dH = cDH * De;
    for (int d = 0; d < 3; ++d) {
      Wb[(size_t)(k * 6 + d) * S + s] = cAE * xe[d] + cJE * Jh[d] + dE * (qreal_t<T>)kh[d];
      Wb[(size_t)(k * 6 + 3 + d) * S + s] = cJH * Je[d] + dH * (qreal_t<T>)kh[d] + cAH * xh[d];
    }
  }
}

__device__ inline zd qwave_sum(zd v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  return v;
}
__device__ inline zf qwave_sum(zf v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  return v;
}

// Disaggregation: one wavefront per task of up to kDisaggEdges consecutive
// positions of one leaf box (task = p0 * 8 + count), lanes over samples:
// y_e = sum_{k,s} R.WE, y_h = sum R.WH
template <class T>
__global__ void __launch_bounds__(256) k_qdisagg(int ntask, int N, int S, const int *task,
                                                 const int *perm, const int *leaf_box_of_pos,
                                                 const T *rt, const T *W, T *y) {
  constexpr int E = kDisaggEdges;
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= ntask) return;
  const int code = task[t], p0 = code >> 3, cnt = code & 7;
  const T *Wb = W + (size_t)leaf_box_of_pos[p0] * 12 * S;
  T se[E], sh[E];
  for (int e = 0; e < E; ++e) se[e] = sh[e] = T{0, 0};
  for (int k = 0; k < 2; ++k) {
    const T *f[E];
    for (int e = 0; e < E; ++e) f[e] = rt + ((size_t)(p0 + min(e, cnt - 1)) * 14 + 8 + k * 3) * S;
    const T *wk = Wb + (size_t)k * 6 * S;
    for (int s = lane; s < S; s += 64) {
      T w[6];
      for (int d = 0; d < 6; ++d) w[d] = wk[(size_t)d * S + s];
      for (int e = 0; e < E; ++e) {
        if (e >= cnt) continue;
        const T r0 = f[e][s], r1 = f[e][S + s], r2 = f[e][2 * S + s];
        se[e] = se[e] + r0 * w[0] + r1 * w[1] + r2 * w[2];
        sh[e] = sh[e] + r0 * w[3] + r1 * w[4] + r2 * w[5];
      }
    }
  }
  for (int e = 0; e < E; ++e) {
    const T a = qwave_sum(se[e]), b = qwave_sum(sh[e]);
    if (lane == 0 && e < cnt) {
      const int m = perm[p0 + e];
      y[m] = a;
      y[m + N] = b;
    }
  }
}

// Near field, per chunk of leaf boxes [b0, b1): the (row edge function,
// element) pairs of box b are rows x T(b), T(b) = the elements of its E2
// columns; pair index q = pbase[b] + r * |T(b)| + li.  Regular pairs here
// (one thread each, a workgroup sharing one element's source table); self,
// edge- and corner-adjacent pairs come from the list kernels and are
// scattered in.
__device__ inline int upper_box(const int64_t *off, int b0, int b1, int64_t g) {
  int lo = b0, hi = b1 - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(64) k_qn_regular(QuadDev D, QuadTopo T, int b0, int b1,
                                                   const int64_t *toff, const int64_t *pbase,
                                                   const int *tb_off, const int *tb_list,
                                                   const int *box_first, const int *box_count,
                                                   const int *perm, zd *sums) {
  // task = (box b, 64-row block rb, element li of T(b)), task index
  // toff[b] + rb * |T(b)| + li
  __shared__ QSrc src[16];
  const int64_t task = toff[b0] + blockIdx.x;
  const int b = upper_box(toff, b0, b1, task);
  const int nT = tb_off[b + 1] - tb_off[b];
  const int64_t rel = task - toff[b];
  const int rb = (int)(rel / nT), li = (int)(rel - (int64_t)rb * nT);
  const int t = tb_list[tb_off[b] + li];
  if (threadIdx.x < 16)
    qsrc_point<4>(D, T.el + (size_t)t * 24, threadIdx.x >> 2, threadIdx.x & 3, src[threadIdx.x]);
  __syncthreads();
  const int r = rb * 64 + threadIdx.x;
  if (r >= box_count[b]) return;
  const int s = perm[box_first[b] + r];
  int side[2];
  if (pair_class(T, s, t, side) != 3) return;
  const int ii = T.edge_p[5 * s] - 1, pI = T.edge_p[5 * s + 1];
  zd out[3][10];
  regular_pair_tab(D, T.el + (size_t)ii * 24, src, pI, out);
  store30(out, sums + (size_t)(pbase[b] - pbase[b0] + (int64_t)r * nT + li) * 30);
}

__global__ void k_qn_scatter(int n, const int64_t *dst, const zd *src, zd *sums) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)n * 30) return;
  const int i = (int)(id / 30), k = (int)(id % 30);
  sums[(size_t)dst[i] * 30 + k] = src[(size_t)i * 30 + k];
}

// blocks g of the chunk's boxes: the dense assembly's rule
// (lib_sie_quad_solver_normal.f90:165-210) on the pair sums
__global__ void k_qn_assemble(QuadTopo T, int form, zd eta_a, zd eta12, int b0, int b1,
                              const int64_t *near_off, const int *e2_off, const int *e2_list,
                              const int *colA, const int *colB, const int64_t *pbase,
                              const int *tb_off, const zd *sums, zd *blocks) {
  const int64_t g = near_off[b0] + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= near_off[b1]) return;
  const int b = upper_box(near_off, b0, b1, g);
  const int cols = e2_off[b + 1] - e2_off[b];
  const int64_t rel = g - near_off[b];
  const int r = (int)(rel / cols), c = (int)(rel - (int64_t)r * cols);
  const int ci = e2_off[b] + c;
  const int tc = e2_list[ci];
  const int *e = T.edge_p + 5 * (size_t)tc;
  const int pa = e[1] - 1, pq = e[3] - 1, f = e[4];
  const int nT = tb_off[b + 1] - tb_off[b];
  const int64_t base = pbase[b] - pbase[b0] + (int64_t)r * nT;
  const zd *sa = sums + (size_t)(base + colA[ci]) * 30;
  const zd *sb = f != 0 ? sums + (size_t)(base + colB[ci]) * 30 : nullptr;
  zd m3[3];
  for (int k = 0; k < 3; ++k) {
    zd v = sa[k * 10 + pa];
    if (f != 0) v = v + sb[k * 10 + pq] * (double)f;
    m3[k] = v;
  }
  zd a = m3[0], bb = m3[1], cc = m3[2], d = m3[0];
  if (form == SPK_ICTF) {
    a = m3[0] / eta_a;
    bb = m3[1] / eta_a;
    cc = m3[2] * eta_a;
    d = m3[0] * eta_a;
  } else if (form == SPK_MCTF) {
    cc = m3[2] * eta12;
    d = m3[0] * eta12;
  }
  zd *o = blocks + 4 * (size_t)g;
  o[0] = a;
  o[1] = bb;
  o[2] = cc;
  o[3] = d;
}

// element-local coefficients of a solution (the S_EH array of
// lib_sie_quad_solver_normal.f90:312-323): seh[cur][element][10]
__global__ void k_quad_local(QuadTopo T, const zd *x, zd *seh) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= T.medge) return;
  const int *e = T.edge_p + 5 * (size_t)m;
  const int ii = e[0] - 1, pa = e[1] - 1, jj = e[2] - 1, pq = e[3] - 1;
  const size_t nel10 = (size_t)T.nel * 10;
  const zd xe = x[m], xh = x[T.medge + m];
  seh[(size_t)ii * 10 + pa] = xe;
  seh[nel10 + (size_t)ii * 10 + pa] = xh;
  if (jj >= 0) {
    seh[(size_t)jj * 10 + pq] = xe * (double)e[4];
    seh[nel10 + (size_t)jj * 10 + pq] = xh * (double)e[4];
  }
}

// Integration_scattered_field_quad (src/lib_sie_quad_calculation_mod.f90:1281-1323)
// summed over all elements (lib_sie_quad_solver_normal.f90:388-396, ngp = 5):
// one workgroup per observation point, threads over elements, the partial
// fields added in a fixed order.  E, H: 3 complex per point.
__global__ void __launch_bounds__(256) k_quad_field(QuadDev D, QuadTopo T, const double *pts,
                                                    const zd *seh, double *Eo, double *Ho) {
  const int pt = blockIdx.x;
  const double Ra[3] = {pts[3 * pt], pts[3 * pt + 1], pts[3 * pt + 2]};
  zd acc[6] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}};
  const size_t nel10 = (size_t)T.nel * 10;
  for (int el = threadIdx.x; el < T.nel; el += blockDim.x) {
    const double *Eq = T.el + (size_t)el * 24;
    const zd *SE = seh + (size_t)el * 10, *SH = seh + nel10 + (size_t)el * 10;
    for (int i = 0; i < 5; ++i)
      for (int j = 0; j < 5; ++j) {
        const double kq = D.x5[i], eq = D.x5[j];
        Geo g;
        geo(kq, eq, Eq, g);
        double Rq[3];
        rloc(kq, eq, Eq, Rq);
        const double R[3] = {Ra[0] - Rq[0], Ra[1] - Rq[1], Ra[2] - Rq[2]};
        const double nR = sqrt(dot3(R, R));
        zd G, DG;
        qgreen(D.k1, nR, G, DG);
        const zd gg[3] = {DG * (R[0] / nR), DG * (R[1] / nR), DG * (R[2] / nR)};
        const double ff = g.JJ * D.w5[i] * D.w5[j];
        for (int n = 0; n < 10; ++n) {
          const double fv = fedge(n, kq, eq);
          const double *v = n < 5 ? g.au : g.av;
          const double Nq[3] = {fv * v[0], fv * v[1], fv * v[2]};
          double aa[3];
          cross3(g.nq, Nq, aa);
          const double nrot = fedge_rot(n, kq, eq) / g.JJ;
          // KA = (aa x gradG) ff, KB = nrot gradG ff, KC = aa G ff
          const zd KA[3] = {(gg[2] * aa[1] - gg[1] * aa[2]) * ff,
                            (gg[0] * aa[2] - gg[2] * aa[0]) * ff,
                            (gg[1] * aa[0] - gg[0] * aa[1]) * ff};
          const zd se = SE[n], sh = SH[n];
          const zd cE = sh * (1.0 / D.iwe), cH = se * (1.0 / (-D.iwm));
          const zd dE = sh * D.iwm, dH = se * D.iwe;
          for (int d = 0; d < 3; ++d) {
            const zd KB = gg[d] * (nrot * ff), KC = G * (aa[d] * ff);
            acc[d] = acc[d] + KA[d] * se + KB * cE + KC * dE;
            acc[3 + d] = acc[3 + d] + KA[d] * sh + KB * cH - KC * dH;
          }
        }
      }
  }
  __shared__ zd red[4][6];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int q = 0; q < 6; ++q) {
    const zd s = wsum(acc[q]);
    if (lane == 0) red[wv][q] = s;
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    zd s = red[0][threadIdx.x];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) s = s + red[w][threadIdx.x];
    double *o = threadIdx.x < 3 ? Eo : Ho;
    const int d = threadIdx.x % 3;
    o[(size_t)pt * 6 + 2 * d] = s.x;
    o[(size_t)pt * 6 + 2 * d + 1] = s.y;
  }
}

}  // namespace

void launch_qpoints(const QuadDev &d, const QuadTopo &T, int N, const int *perm,
                    const int *leaf_box, const double *leaf_centre, double *pts, hipStream_t st) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_qpoints, dim3(nb(N, 64)), dim3(64), 0, st, d, T, N, perm, leaf_box,
                     leaf_centre, pts);
}
void launch_qpatterns(int N, int S, const double *khat, const double *pts, zd k1, zd k2, zd *rt,
                      hipStream_t st) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_qpatterns, dim3(nb(S, 256), N), dim3(256), 0, st, N, S, khat, pts, k1, k2,
                     rt);
}
template <class T>
void launch_qagg(int nbl, int S, int N, const int *box_first, const int *box_count,
                 const int *perm, const T *rt, const T *x, T *B, size_t half, hipStream_t st) {
  if (nbl <= 0) return;
  hipLaunchKernelGGL(k_qagg<T>, dim3(nb(S, 256), nbl, 8), dim3(256), 0, st, S, N, box_first,
                     box_count, perm, rt, x, B, half);
}
template <class T>
void launch_qfold(int nbl, int S, const T *D, size_t half, const double *khat,
                  const QuadFoldC &C, T *W, hipStream_t st) {
  if (nbl <= 0) return;
  hipLaunchKernelGGL(k_qfold<T>, dim3(nb(S, 256), nbl), dim3(256), 0, st, S, D, half, khat, C, W);
}
template <class T>
void launch_qdisagg(int ntask, int N, int S, const int *task, const int *perm,
                    const int *leaf_box_of_pos, const T *rt, const T *W, T *y, hipStream_t st) {
  if (ntask <= 0) return;
  hipLaunchKernelGGL(k_qdisagg<T>, dim3(nb(ntask, 4)), dim3(256), 0, st, ntask, N, S, task, perm,
                     leaf_box_of_pos, rt, W, y);
}
template void launch_qagg<zd>(int, int, int, const int *, const int *, const int *, const zd *,
                              const zd *, zd *, size_t, hipStream_t);
template void launch_qagg<zf>(int, int, int, const int *, const int *, const int *, const zf *,
                              const zf *, zf *, size_t, hipStream_t);
template void launch_qfold<zd>(int, int, const zd *, size_t, const double *, const QuadFoldC &,
                               zd *, hipStream_t);
template void launch_qfold<zf>(int, int, const zf *, size_t, const double *, const QuadFoldC &,
                               zf *, hipStream_t);
template void launch_qdisagg<zd>(int, int, int, const int *, const int *, const int *, const zd *,
                                 const zd *, zd *, hipStream_t);
template void launch_qdisagg<zf>(int, int, int, const int *, const int *, const int *, const zf *,
                                 const zf *, zf *, hipStream_t);
void launch_qn_regular(const QuadDev &d, const QuadTopo &T, int b0, int b1, int64_t ntask,
                       const int64_t *toff, const int64_t *pbase, const int *tb_off,
                       const int *tb_list, const int *box_first, const int *box_count,
                       const int *perm, zd *sums, hipStream_t st) {
  if (ntask <= 0) return;
  hipLaunchKernelGGL(k_qn_regular, dim3((unsigned)ntask), dim3(64), 0, st, d, T, b0, b1, toff,
                     pbase, tb_off, tb_list, box_first, box_count, perm, sums);
}
void launch_qn_scatter(int n, const int64_t *dst, const zd *src, zd *sums, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_qn_scatter, dim3(nb((int64_t)n * 30, 256)), dim3(256), 0, st, n, dst, src,
                     sums);
}
void launch_qn_assemble(const QuadTopo &T, int form, zd eta_a, zd eta12, int b0, int b1,
                        int64_t nblocks, const int64_t *near_off, const int *e2_off,
                        const int *e2_list, const int *colA, const int *colB,
                        const int64_t *pbase, const int *tb_off, const zd *sums, zd *blocks,
                        hipStream_t st) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(k_qn_assemble, dim3(nb(nblocks, 256)), dim3(256), 0, st, T, form, eta_a,
                     eta12, b0, b1, near_off, e2_off, e2_list, colA, colB, pbase, tb_off, sums,
                     blocks);
}
void launch_quad_local(const QuadTopo &T, const zd *x, zd *seh, hipStream_t st) {
  if (T.medge <= 0) return;
  hipLaunchKernelGGL(k_quad_local, dim3(nb(T.medge, 256)), dim3(256), 0, st, T, x, seh);
}
void launch_quad_field(const QuadDev &d, const QuadTopo &T, int npts, const double *pts,
                       const zd *seh, double *E, double *H, hipStream_t st) {
  if (npts <= 0) return;
  hipLaunchKernelGGL(k_quad_field, dim3(npts), dim3(256), 0, st, d, T, pts, seh, E, H);
}

void launch_quad_scatter(int n, int nel, const int *r, const int *t, const zd *src, zd *sums,
                         hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_quad_scatter, dim3(nb((int64_t)n * 30, 256)), dim3(256), 0, st, n, nel, r, t,
                     src, sums);
}

void launch_quad_regular_grid(const QuadDev &d, const QuadTopo &T, int s0, int rows, zd *sums,
                              hipStream_t st) {
  if (rows <= 0 || T.nel <= 0) return;
  const int nrb = (rows + 63) / 64;
  hipLaunchKernelGGL(k_quad_regular_grid, dim3((unsigned)((int64_t)T.nel * nrb)), dim3(64), 0, st,
                     d, T, s0, rows, nrb, sums);
}
void launch_quad_regular_list(const QuadDev &d, const QuadTopo &T, int n, const int *s,
                              const int *t, const int *ngp, zd *out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_quad_regular_list, dim3(nb(n, 128)), dim3(128), 0, st, d, T, n, s, t,
                     ngp, out);
}
void launch_quad_corner(const QuadDev &d, const QuadTopo &T, int n, const int *s, const int *t,
                        zd *out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_quad_corner, dim3(n), dim3(64), 0, st, d, T, s, t, out);
}
void launch_quad_permute(int64_t n, const int64_t *perm, const zd *z, zd *x, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_quad_permute, dim3(nb(n, 256)), dim3(256), 0, st, n, perm, z, x);
}

hipError_t quad_self_rows(const QuadDev &d, const QuadTopo &T, const QuadMesh &m,
                          const std::vector<int> &ls, zd *out, hipStream_t st) {
  if (ls.empty()) return hipSuccess;
  std::vector<int> els(ls.size()), slot(ls.size());
  for (size_t i = 0; i < ls.size(); ++i) els[i] = m.edge_p[5 * (size_t)ls[i]] - 1;
  std::sort(els.begin(), els.end());
  els.erase(std::unique(els.begin(), els.end()), els.end());
  for (size_t i = 0; i < ls.size(); ++i) {
    const int *e = &m.edge_p[5 * (size_t)ls[i]];
    const int k = (int)(std::lower_bound(els.begin(), els.end(), e[0] - 1) - els.begin());
    slot[i] = k * 10 + e[1] - 1;
  }
  QBuf<int> d_els, d_slot;
  QBuf<zd> per_el;
  hipError_t e = d_els.upload(els);
  if (e == hipSuccess) e = d_slot.upload(slot);
  if (e == hipSuccess) e = per_el.alloc(els.size() * 10 * 30);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_quad_self_el, dim3((unsigned)els.size()), dim3(256), 0, st, d, T, d_els.p,
                     per_el.p);
  const int n = (int)ls.size();
  hipLaunchKernelGGL(k_quad_pick, dim3(nb((int64_t)n * 30, 256)), dim3(256), 0, st, n, d_slot.p,
                     per_el.p, out);
  e = hipGetLastError();
  // the temporaries are freed on return
  const hipError_t e2 = hipStreamSynchronize(st);
  return e != hipSuccess ? e : e2;
}
void launch_quad_edge_adj(const QuadDev &d, const QuadTopo &T, int n, const int *s, const int *t,
                          const int *side, zd *out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_quad_edge_adj, dim3(n), dim3(128), 0, st, d, T, s, t, side, out);
}
void launch_quad_rhs_plane(const QuadDev &d, const QuadTopo &T, const QuadRhsArgs &a, zd *b,
                           hipStream_t st) {
  if (T.medge <= 0) return;
  hipLaunchKernelGGL(k_quad_rhs_plane, dim3(nb(T.medge, 128)), dim3(128), 0, st, d, T, a, b);
}
void launch_quad_assemble(const QuadTopo &T, int formulation, zd eta_a, zd eta12, int s0,
                          int rows, const zd *sums, zd *MM, hipStream_t st) {
  const int64_t n = (int64_t)rows * T.medge;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_quad_assemble, dim3(nb(n, 256)), dim3(256), 0, st, T, formulation, eta_a,
                     eta12, s0, rows, sums, MM);
}

// set_local_material_parameters_quad (src/lib_sie_quad_data_container.f90:118-129)
void quad_constants(const spk_params &p, QuadDev *D, zd *eta_a, zd *eta12) {
  typedef std::complex<double> C;
  const double k0 = 2 * kPi / p.lambda, omega = 2 * kPi * kC0 / p.lambda;
  const C er1(p.eps_r1_re, p.eps_r1_im), er2(p.eps_r2_re, p.eps_r2_im);
  const C er1v = (er1 == C(0.0)) ? C(1.0) : er1;
  const C k1 = k0 * std::sqrt(er1v), k2 = k0 * std::sqrt(er2);
  const C eta1 = std::sqrt(kMy0 / (kEps0 * er1v)), eta2 = std::sqrt(kMy0 / (kEps0 * er2));
  auto Z = [](C c) { return zd{c.real(), c.imag()}; };
  D->k1 = Z(k1);
  D->k2 = Z(k2);
  D->ie1 = Z(1.0 / er1v);
  D->ie2 = Z(1.0 / er2);
  D->e1 = Z(er1v);
  D->e2 = Z(er2);
  D->ceps = Z(1.0 + 1.0 / er2);
  D->iwe = Z(C(0.0, 1.0) * omega * kEps0);
  D->iwm = Z(C(0.0, 1.0) * omega * kMy0);
  gauss_legendre(4, D->x4, D->w4);
  gauss_legendre(5, D->x5, D->w5);
  gauss_legendre(6, D->x6, D->w6);
  gauss_legendre(15, D->x15, D->w15);
  *eta_a = Z((eta1 + eta2) / 2.0);
  *eta12 = Z(eta1 * eta2);
}

QuadRhsArgs quad_rhs_args(int formulation, const QuadDev &D, zd eta_a, zd eta12,
                          const double k_in[3], const double e_in[3]) {
  QuadRhsArgs a{};
  for (int d = 0; d < 3; ++d) {
    a.kin[d] = k_in[d];
    a.ein[d] = e_in[d];
  }
  a.k1c = zd{D.k1.x, -D.k1.y};
  a.c0my = (double)kMy0 * (double)kC0;
  a.sE = zd{1.0, 0.0};
  a.sH = zd{1.0, 0.0};
  if (formulation == SPK_MCTF) {
    a.sH = eta12;
  } else if (formulation == SPK_ICTF) {
    a.sE = zd{1.0, 0.0} / eta_a;
    a.sH = eta_a;
  }
  return a;
}

bool quad_topology_arrays(const QuadMesh &m, std::vector<double> *el, std::vector<int> *eadj,
                          std::vector<int> *eside, std::vector<int> *cadj) {
  el->assign((size_t)m.nel * 24, 0.0);
  for (int e = 0; e < m.nel; ++e)
    for (int n = 0; n < 8; ++n)
      for (int d = 0; d < 3; ++d)
        (*el)[(size_t)e * 24 + n * 3 + d] = m.pp[3 * (size_t)(m.vert[8 * (size_t)e + n] - 1) + d];
  eadj->assign((size_t)m.nel * kQuadMaxAdj, 0);
  eside->assign((size_t)m.nel * kQuadMaxAdj * 2, 0);
  cadj->assign((size_t)m.nel * kQuadMaxAdj, 0);
  for (int e = 0; e < m.nel; ++e) {
    if ((int)m.eadj[e].size() > kQuadMaxAdj || (int)m.cadj[e].size() > kQuadMaxAdj) return false;
    for (size_t a = 0; a < m.eadj[e].size(); ++a) {
      (*eadj)[(size_t)e * kQuadMaxAdj + a] = m.eadj[e][a].other;
      (*eside)[((size_t)e * kQuadMaxAdj + a) * 2] = m.eadj[e][a].side[0];
      (*eside)[((size_t)e * kQuadMaxAdj + a) * 2 + 1] = m.eadj[e][a].side[1];
    }
    for (size_t a = 0; a < m.cadj[e].size(); ++a) (*cadj)[(size_t)e * kQuadMaxAdj + a] = m.cadj[e][a];
  }
  return true;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
namespace {

#define QCHK(expr)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return spk::set_error(SPK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace
}  // namespace spk

using namespace spk;

struct spk_quad {
  spk_params p{};
  hipStream_t st = nullptr;
  QuadMesh mesh;
  QuadDev dev{};
  zd eta_a{}, eta12{};
  bool ready = false;
  QBuf<double> d_el;
  QBuf<int> d_edge_p, d_eadj, d_eside, d_cadj;
  // dense solve: rocSOLVER handle on st, pivots and status of the last factorisation
  rocblas_handle rb = nullptr;
  QBuf<int64_t> ipiv, info, perm;
  QBuf<zd> lu_tmp;
  QuadTopo topo() const {
    return QuadTopo{d_el.p, d_edge_p.p, d_eadj.p, d_eside.p, d_cadj.p, mesh.nel, mesh.medge};
  }
};

extern "C" {

int spk_quad_create(const spk_params *params, spk_quad **out) {
  if (!params || !out) return set_error(SPK_ERR_ARG, "null argument");
  if (params->lambda <= 0) return set_error(SPK_ERR_ARG, "lambda must be positive");
  const int f = params->formulation;
  if (f != SPK_PMCHWT && f != SPK_MCTF && f != SPK_ICTF)
    return set_error(SPK_ERR_ARG, "quadrilateral fill supports PMCHWT, MCTF and ICTF");
  int ndev = 0;
  QCHK(hipGetDeviceCount(&ndev));
  if (params->device < 0 || params->device >= ndev)
    return set_error(SPK_ERR_ARG, "device ordinal out of range");
  QCHK(hipSetDevice(params->device));
  spk_quad *q = new spk_quad();
  q->p = *params;
  if (hipStreamCreateWithFlags(&q->st, hipStreamNonBlocking) != hipSuccess) {
    delete q;
    return set_error(SPK_ERR_HIP, "hipStreamCreate failed");
  }
  quad_constants(*params, &q->dev, &q->eta_a, &q->eta12);
  *out = q;
  return SPK_OK;
}

void spk_quad_destroy(spk_quad *q) {
  if (!q) return;
  if (q->rb) (void)rocblas_destroy_handle(q->rb);
  if (q->st) (void)hipStreamDestroy(q->st);
  delete q;
}

int spk_quad_setup(spk_quad *q, double Lx, double Ly, int nx, int ny, const double *z) {
  if (!q || !z || nx < 1 || ny < 1 || !(Lx > 0) || !(Ly > 0))
    return set_error(SPK_ERR_ARG, "bad quadrilateral mesh arguments");
  QCHK(hipSetDevice(q->p.device));
  q->ready = false;
  QuadMesh &m = q->mesh;
  m = QuadMesh();
  quad_mesh(Lx, Ly, nx, ny, z, &m);
  quad_topology(&m);
  std::vector<double> el;
  std::vector<int> eadj, eside, cadj;
  if (!quad_topology_arrays(m, &el, &eadj, &eside, &cadj))
    return set_error(SPK_ERR_MESH, "element with more than 8 neighbours");
  QCHK(q->d_el.upload(el));
  QCHK(q->d_edge_p.upload(m.edge_p));
  QCHK(q->d_eadj.upload(eadj));
  QCHK(q->d_eside.upload(eside));
  QCHK(q->d_cadj.upload(cadj));
  q->ready = true;
  return SPK_OK;
}

int spk_quad_topology_probe(double Lx, double Ly, int nx, int ny, const double *z, int *n_edges,
                            int *n_elements, int *n_nodes, double *pp, int *vert8, int *edge_p5) {
  if (!z || nx < 1 || ny < 1 || !(Lx > 0) || !(Ly > 0))
    return set_error(SPK_ERR_ARG, "bad quadrilateral mesh arguments");
  QuadMesh m;
  quad_mesh(Lx, Ly, nx, ny, z, &m);
  quad_topology(&m);
  if (n_edges) *n_edges = m.medge;
  if (n_elements) *n_elements = m.nel;
  if (n_nodes) *n_nodes = m.np;
  if (pp) std::memcpy(pp, m.pp.data(), m.pp.size() * sizeof(double));
  if (vert8) std::memcpy(vert8, m.vert.data(), m.vert.size() * sizeof(int));
  if (edge_p5) std::memcpy(edge_p5, m.edge_p.data(), m.edge_p.size() * sizeof(int));
  return SPK_OK;
}

int spk_quad_sizes(const spk_quad *q, int *n_edges, int *n_elements, int *n_nodes) {
  if (!q) return set_error(SPK_ERR_ARG, "null argument");
  if (n_edges) *n_edges = q->mesh.medge;
  if (n_elements) *n_elements = q->mesh.nel;
  if (n_nodes) *n_nodes = q->mesh.np;
  return SPK_OK;
}

int spk_quad_get_mesh(const spk_quad *q, double *pp, int *vert8) {
  if (!q) return set_error(SPK_ERR_ARG, "null argument");
  if (pp) std::memcpy(pp, q->mesh.pp.data(), q->mesh.pp.size() * sizeof(double));
  if (vert8) std::memcpy(vert8, q->mesh.vert.data(), q->mesh.vert.size() * sizeof(int));
  return SPK_OK;
}

int spk_quad_get_edges(const spk_quad *q, int *edge_p5) {
  if (!q || !edge_p5) return set_error(SPK_ERR_ARG, "null argument");
  std::memcpy(edge_p5, q->mesh.edge_p.data(), q->mesh.edge_p.size() * sizeof(int));
  return SPK_OK;
}

int spk_quad_edge_element(spk_quad *q, int n, const int *s, const int *t, double *out) {
  if (!q || n < 0 || (n && (!s || !t || !out))) return set_error(SPK_ERR_ARG, "bad arguments");
  if (!q->ready) return set_error(SPK_ERR_STATE, "spk_quad_setup has not completed");
  const QuadMesh &m = q->mesh;
  // classify on the host as impedance_edge_versus_element does
  std::vector<int> rs, rt, rg, ri, ss, si, es, et, eside, ei;
  for (int i = 0; i < n; ++i) {
    if (s[i] < 0 || s[i] >= m.medge || t[i] < 0 || t[i] >= m.nel)
      return set_error(SPK_ERR_ARG, "pair index out of range");
    const int ii = m.edge_p[5 * (size_t)s[i]] - 1;
    if (t[i] == ii) {
      ss.push_back(s[i]);
      si.push_back(i);
      continue;
    }
    bool done = false;
    for (const QuadEdgeAdj &a : m.eadj[ii])
      if (a.other == t[i] + 1) {
        es.push_back(s[i]);
        et.push_back(t[i]);
        eside.push_back(a.side[0]);
        eside.push_back(a.side[1]);
        ei.push_back(i);
        done = true;
        break;
      }
    if (done) continue;
    int g = 4;
    for (int c : m.cadj[ii])
      if (c == t[i] + 1) g = 6;
    rs.push_back(s[i]);
    rt.push_back(t[i]);
    rg.push_back(g);
    ri.push_back(i);
  }
  QBuf<int> a0, a1, a2, c0, c1, c2;
  QBuf<zd> o1, o2, o3;
  QCHK(a0.upload(rs));
  QCHK(a1.upload(rt));
  QCHK(a2.upload(rg));
  QCHK(c0.upload(es));
  QCHK(c1.upload(et));
  QCHK(c2.upload(eside));
  QCHK(o1.alloc(rs.size() * 30));
  QCHK(o2.alloc(ss.size() * 30));
  QCHK(o3.alloc(es.size() * 30));
  const QuadTopo T = q->topo();
  launch_quad_regular_list(q->dev, T, (int)rs.size(), a0.p, a1.p, a2.p, o1.p, q->st);
  QCHK(quad_self_rows(q->dev, T, m, ss, o2.p, q->st));
  launch_quad_edge_adj(q->dev, T, (int)es.size(), c0.p, c1.p, c2.p, o3.p, q->st);
  QCHK(hipGetLastError());
  std::vector<zd> h1(rs.size() * 30), h2(ss.size() * 30), h3(es.size() * 30);
  if (!h1.empty()) QCHK(hipMemcpyAsync(h1.data(), o1.p, h1.size() * sizeof(zd), hipMemcpyDeviceToHost, q->st));
  if (!h2.empty()) QCHK(hipMemcpyAsync(h2.data(), o2.p, h2.size() * sizeof(zd), hipMemcpyDeviceToHost, q->st));
  if (!h3.empty()) QCHK(hipMemcpyAsync(h3.data(), o3.p, h3.size() * sizeof(zd), hipMemcpyDeviceToHost, q->st));
  QCHK(hipStreamSynchronize(q->st));
  zd *o = (zd *)out;
  for (size_t k = 0; k < ri.size(); ++k) std::memcpy(o + (size_t)ri[k] * 30, &h1[k * 30], 30 * sizeof(zd));
  for (size_t k = 0; k < si.size(); ++k) std::memcpy(o + (size_t)si[k] * 30, &h2[k * 30], 30 * sizeof(zd));
  for (size_t k = 0; k < ei.size(); ++k) std::memcpy(o + (size_t)ei[k] * 30, &h3[k * 30], 30 * sizeof(zd));
  return SPK_OK;
}

int spk_quad_fill_device(spk_quad *q, void *d_MM) {
  if (!q || !d_MM) return set_error(SPK_ERR_ARG, "null argument");
  if (!q->ready) return set_error(SPK_ERR_STATE, "spk_quad_setup has not completed");
  const QuadMesh &m = q->mesh;
  const QuadTopo T = q->topo();
  // rows of edge functions in chunks whose [rows][elements][30] sums fit
  // 256 MiB, whole 64-row blocks of the regular-pair kernel where they can
  const size_t per_row = (size_t)m.nel * 30 * sizeof(zd);
  size_t rows_fit = std::max<size_t>(1, (256u << 20) / per_row);
  if (rows_fit >= 64) rows_fit -= rows_fit % 64;
  const int chunk = (int)std::min<size_t>((size_t)m.medge, rows_fit);
  QBuf<zd> sums, so, eo, co;
  QBuf<int> el_s, el_t, el_side, sc_r, sc_t, sc_r2, cl_s, cl_t, sc_r3;
  QCHK(sums.alloc((size_t)chunk * m.nel * 30));
  // the self terms of every row at once (one workgroup per element)
  {
    std::vector<int> all(m.medge);
    for (int i = 0; i < m.medge; ++i) all[i] = i;
    QCHK(so.alloc((size_t)m.medge * 30));
    QCHK(quad_self_rows(q->dev, T, m, all, so.p, q->st));
  }
  for (int s0 = 0; s0 < m.medge; s0 += chunk) {
    const int rows = std::min(chunk, m.medge - s0);
    launch_quad_regular_grid(q->dev, T, s0, rows, sums.p, q->st);
    // edge- and corner-adjacent pairs of these rows
    std::vector<int> es, et, eside, cs, ct, cr;
    for (int r = 0; r < rows; ++r) {
      const int ii = m.edge_p[5 * (size_t)(s0 + r)] - 1;
      for (const QuadEdgeAdj &a : m.eadj[ii]) {
        es.push_back(s0 + r);
        et.push_back(a.other - 1);
        eside.push_back(a.side[0]);
        eside.push_back(a.side[1]);
      }
      for (int c : m.cadj[ii]) {
        bool edge = false;
        for (const QuadEdgeAdj &a : m.eadj[ii]) edge = edge || a.other == c;
        if (edge || c - 1 == ii) continue;
        cs.push_back(s0 + r);
        ct.push_back(c - 1);
        cr.push_back(r);
      }
    }
    QCHK(el_s.upload(es));
    QCHK(el_t.upload(et));
    QCHK(el_side.upload(eside));
    QCHK(eo.alloc(es.size() * 30));
    launch_quad_edge_adj(q->dev, T, (int)es.size(), el_s.p, el_t.p, el_side.p, eo.p, q->st);
    QCHK(cl_s.upload(cs));
    QCHK(cl_t.upload(ct));
    QCHK(co.alloc(cs.size() * 30));
    launch_quad_corner(q->dev, T, (int)cs.size(), cl_s.p, cl_t.p, co.p, q->st);
    QCHK(hipGetLastError());
    // scatter the self and edge-adjacent pairs into the sums
    std::vector<int> sr(rows), st_(rows), er(es.size());
    for (int r = 0; r < rows; ++r) {
      sr[r] = r;
      st_[r] = m.edge_p[5 * (size_t)(s0 + r)] - 1;
    }
    for (size_t k = 0; k < es.size(); ++k) er[k] = es[k] - s0;
    QCHK(sc_r.upload(sr));
    QCHK(sc_t.upload(st_));
    launch_quad_scatter(rows, m.nel, sc_r.p, sc_t.p, so.p + (size_t)s0 * 30, sums.p, q->st);
    QCHK(sc_r2.upload(er));
    launch_quad_scatter((int)es.size(), m.nel, sc_r2.p, el_t.p, eo.p, sums.p, q->st);
    QCHK(sc_r3.upload(cr));
    launch_quad_scatter((int)cs.size(), m.nel, sc_r3.p, cl_t.p, co.p, sums.p, q->st);
    launch_quad_assemble(T, q->p.formulation, q->eta_a, q->eta12, s0, rows, sums.p, (zd *)d_MM,
                         q->st);
    QCHK(hipGetLastError());
    QCHK(hipStreamSynchronize(q->st));
  }
  return SPK_OK;
}

int spk_quad_rhs_plane_wave_device(spk_quad *q, const double k_in[3], const double e_in[3],
                                   void *d_b) {
  if (!q || !k_in || !e_in || !d_b) return set_error(SPK_ERR_ARG, "null argument");
  if (!q->ready) return set_error(SPK_ERR_STATE, "spk_quad_setup has not completed");
  QCHK(hipSetDevice(q->p.device));
  const QuadRhsArgs a = quad_rhs_args(q->p.formulation, q->dev, q->eta_a, q->eta12, k_in, e_in);
  launch_quad_rhs_plane(q->dev, q->topo(), a, (zd *)d_b, q->st);
  QCHK(hipGetLastError());
  QCHK(hipStreamSynchronize(q->st));
  return SPK_OK;
}

int spk_quad_rhs_plane_wave(spk_quad *q, const double k_in[3], const double e_in[3], double *b) {
  if (!q || !b) return set_error(SPK_ERR_ARG, "null argument");
  if (!q->ready) return set_error(SPK_ERR_STATE, "spk_quad_setup has not completed");
  QCHK(hipSetDevice(q->p.device));
  QBuf<zd> d;
  QCHK(d.alloc(2 * (size_t)q->mesh.medge));
  const int rc = spk_quad_rhs_plane_wave_device(q, k_in, e_in, d.p);
  if (rc) return rc;
  QCHK(hipMemcpy(b, d.p, 2 * (size_t)q->mesh.medge * sizeof(zd), hipMemcpyDeviceToHost));
  return SPK_OK;
}

int spk_quad_fill(spk_quad *q, double *MM) {
  if (!q || !MM) return set_error(SPK_ERR_ARG, "null argument");
  if (!q->ready) return set_error(SPK_ERR_STATE, "spk_quad_setup has not completed");
  const size_t n = 2 * (size_t)q->mesh.medge;
  QBuf<zd> d;
  QCHK(d.alloc(n * n));
  int rc = spk_quad_fill_device(q, d.p);
  if (rc) return rc;
  QCHK(hipMemcpy(MM, d.p, n * n * sizeof(zd), hipMemcpyDeviceToHost));
  return SPK_OK;
}

int spk_quad_solve_lu_device(spk_quad *q, void *d_MM, int nrhs, void *d_b) {
  if (!q || !d_MM || !d_b || nrhs < 1) return set_error(SPK_ERR_ARG, "bad arguments");
  if (!q->ready) return set_error(SPK_ERR_STATE, "spk_quad_setup has not completed");
  QCHK(hipSetDevice(q->p.device));
  const int64_t n = 2 * (int64_t)q->mesh.medge;
  if (!q->rb && rocblas_create_handle(&q->rb) != rocblas_status_success)
    return set_error(SPK_ERR_HIP, "rocblas_create_handle failed");
  if (rocblas_set_stream(q->rb, q->st) != rocblas_status_success)
    return set_error(SPK_ERR_HIP, "rocblas_set_stream failed");
  if (q->ipiv.n != (size_t)n) QCHK(q->ipiv.alloc(n));
  if (!q->info.p) QCHK(q->info.alloc(1));
  // MM is row-major, so the column-major matrix rocSOLVER sees is MM^T:
  // factor MM^T = P L U, then solve with the transposed factors, MM x = b.
  auto *A = (rocblas_double_complex *)d_MM;
  if (rocsolver_zgetrf_64(q->rb, n, n, A, n, q->ipiv.p, q->info.p) != rocblas_status_success)
    return set_error(SPK_ERR_HIP, "rocsolver_zgetrf_64 failed");
  int64_t info = 0;
  QCHK(hipMemcpyAsync(&info, q->info.p, sizeof(int64_t), hipMemcpyDeviceToHost, q->st));
  QCHK(hipStreamSynchronize(q->st));
  if (info > 0)
    return set_error(SPK_ERR_STATE, "U(" + std::to_string(info) + "," + std::to_string(info) +
                                        ") is zero: the system matrix is singular");
  // MM = U^T L^T P^T: x = P (L^T)^-1 (U^T)^-1 b.  The triangular solves go
  // by blocks of kLuBlock rows: a transposed GEMV of the solved part (its
  // reads run down contiguous columns of the factors) and a small TRSV of
  // the diagonal block; the interchanges become one gather by the
  // permutation they compose to (the interchanges of ipiv in reverse order).
  {
    std::vector<int64_t> piv(n), idx(n);
    QCHK(hipMemcpyAsync(piv.data(), q->ipiv.p, n * sizeof(int64_t), hipMemcpyDeviceToHost, q->st));
    QCHK(hipStreamSynchronize(q->st));
    for (int64_t i = 0; i < n; ++i) idx[i] = i;
    for (int64_t i = n - 1; i >= 0; --i) std::swap(idx[i], idx[piv[i] - 1]);
    QCHK(q->perm.upload(idx));
    if (q->lu_tmp.n != (size_t)n) QCHK(q->lu_tmp.alloc(n));
  }
  constexpr int64_t kLuBlock = 512;
  const rocblas_double_complex one = {1.0, 0.0}, mone = {-1.0, 0.0};
  for (int j = 0; j < nrhs; ++j) {
    auto *b = (rocblas_double_complex *)d_b + (size_t)j * n;
    bool ok = true;
    for (int64_t k0 = 0; k0 < n && ok; k0 += kLuBlock) {  // U^T y = b, forward
      const int64_t m = std::min(kLuBlock, n - k0);
      if (k0 > 0)
        ok = rocblas_zgemv_64(q->rb, rocblas_operation_transpose, k0, m, &mone, A + k0 * n, n, b,
                              1, &one, b + k0, 1) == rocblas_status_success;
      ok = ok && rocblas_ztrsv_64(q->rb, rocblas_fill_upper, rocblas_operation_transpose,
                                  rocblas_diagonal_non_unit, m, A + k0 + k0 * n, n, b + k0,
                                  1) == rocblas_status_success;
    }
    const int64_t last = ((n - 1) / kLuBlock) * kLuBlock;
    for (int64_t k0 = last; k0 >= 0 && ok; k0 -= kLuBlock) {  // L^T w = y, backward
      const int64_t m = std::min(kLuBlock, n - k0), k1 = k0 + m;
      if (k1 < n)
        ok = rocblas_zgemv_64(q->rb, rocblas_operation_transpose, n - k1, m, &mone,
                              A + k1 + k0 * n, n, b + k1, 1, &one, b + k0,
                              1) == rocblas_status_success;
      ok = ok && rocblas_ztrsv_64(q->rb, rocblas_fill_lower, rocblas_operation_transpose,
                                  rocblas_diagonal_unit, m, A + k0 + k0 * n, n, b + k0,
                                  1) == rocblas_status_success;
    }
    if (!ok) return set_error(SPK_ERR_HIP, "rocBLAS triangular solve failed");
    launch_quad_permute(n, q->perm.p, (const zd *)b, q->lu_tmp.p, q->st);
    QCHK(hipMemcpyAsync(b, q->lu_tmp.p, n * sizeof(zd), hipMemcpyDeviceToDevice, q->st));
  }
  QCHK(hipGetLastError());
  QCHK(hipStreamSynchronize(q->st));
  return SPK_OK;
}

int spk_quad_solve_plane_wave(spk_quad *q, const double k_in[3], const double e_in[3], double *x) {
  if (!q || !k_in || !e_in || !x) return set_error(SPK_ERR_ARG, "null argument");
  if (!q->ready) return set_error(SPK_ERR_STATE, "spk_quad_setup has not completed");
  QCHK(hipSetDevice(q->p.device));
  const size_t n = 2 * (size_t)q->mesh.medge;
  QBuf<zd> MM, b;
  QCHK(MM.alloc(n * n));
  QCHK(b.alloc(n));
  int rc = spk_quad_fill_device(q, MM.p);
  if (!rc) rc = spk_quad_rhs_plane_wave_device(q, k_in, e_in, b.p);
  if (!rc) rc = spk_quad_solve_lu_device(q, MM.p, 1, b.p);
  if (rc) return rc;
  QCHK(hipMemcpy(x, b.p, n * sizeof(zd), hipMemcpyDeviceToHost));
  return SPK_OK;
}

int spk_quad_scattered_field(spk_quad *q, int npts, const double *r, const double *x, double *E,
                             double *H) {
  if (!q || !x || npts < 0 || (npts && (!r || !E || !H)))
    return set_error(SPK_ERR_ARG, "bad arguments");
  if (!q->ready) return set_error(SPK_ERR_STATE, "spk_quad_setup has not completed");
  if (!npts) return SPK_OK;
  QCHK(hipSetDevice(q->p.device));
  const QuadMesh &m = q->mesh;
  QBuf<zd> dx, seh, dE, dH;
  QBuf<double> dr;
  QCHK(dx.alloc(2 * (size_t)m.medge));
  QCHK(seh.alloc(2 * (size_t)m.nel * 10));
  QCHK(dr.alloc(3 * (size_t)npts));
  QCHK(dE.alloc(3 * (size_t)npts));
  QCHK(dH.alloc(3 * (size_t)npts));
  QCHK(hipMemcpyAsync(dx.p, x, dx.n * sizeof(zd), hipMemcpyHostToDevice, q->st));
  QCHK(hipMemcpyAsync(dr.p, r, dr.n * sizeof(double), hipMemcpyHostToDevice, q->st));
  // element-local coefficients S_EH, then the field of every element at every point
  QCHK(hipMemsetAsync(seh.p, 0, seh.n * sizeof(zd), q->st));
  const QuadTopo T = q->topo();
  launch_quad_local(T, dx.p, seh.p, q->st);
  launch_quad_field(q->dev, T, npts, dr.p, seh.p, (double *)dE.p, (double *)dH.p, q->st);
  QCHK(hipGetLastError());
  QCHK(hipMemcpyAsync(E, dE.p, dE.n * sizeof(zd), hipMemcpyDeviceToHost, q->st));
  QCHK(hipMemcpyAsync(H, dH.p, dH.n * sizeof(zd), hipMemcpyDeviceToHost, q->st));
  QCHK(hipStreamSynchronize(q->st));
  return SPK_OK;
}

}  // extern "C"

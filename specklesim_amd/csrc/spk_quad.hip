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

#include "spk_quad.h"

namespace spk {

namespace {

constexpr double QPI = 3.14159265358979323846;

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

__global__ void __launch_bounds__(128) k_quad_regular_grid(QuadDev D, QuadTopo T, int s0, int rows,
                                                           zd *sums) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)rows * T.nel) return;
  const int r = (int)(id / T.nel), t = (int)(id % T.nel);
  const int s = s0 + r;
  int side[2];
  const int cls = pair_class(T, s, t, side);
  if (cls < 2) return;
  const int ii = T.edge_p[5 * s] - 1, pI = T.edge_p[5 * s + 1];
  zd out[3][10];
  if (cls == 2)
    regular_pair<6>(D, D.x6, D.w6, T.el + (size_t)ii * 24, T.el + (size_t)t * 24, pI, out);
  else
    regular_pair<4>(D, D.x4, D.w4, T.el + (size_t)ii * 24, T.el + (size_t)t * 24, pI, out);
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
__device__ void block_reduce30(zd v[3][10], zd *dst) {
  __shared__ zd red[4][30];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int k = 0; k < 3; ++k)
    for (int n = 0; n < 10; ++n) {
      const zd s = wsum(v[k][n]);
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

// self term at one outer point: Integration_KABCD_Taylor_ST_4FEdge (:299-380)
__device__ void taylor_point(const QuadDev &D, int pI, double mu, double nu, const double *El,
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
  double Na[3];
  obs_fn(gp, pI, mu, nu, Na);
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

__global__ void __launch_bounds__(256) k_quad_self(QuadDev D, QuadTopo T, const int *ls,
                                                   zd *out) {
  const int s = ls[blockIdx.x];
  const int ii = T.edge_p[5 * s] - 1, pI = T.edge_p[5 * s + 1];
  const double *El = T.el + (size_t)ii * 24;
  zd v[3][10];
  for (int k = 0; k < 3; ++k)
    for (int n = 0; n < 10; ++n) v[k][n] = zd{0, 0};
  const int p = threadIdx.x;
  if (p < 225) {
    const int i = p % 15, j = p / 15;
    Geo ga;
    geo(D.x15[i], D.x15[j], El, ga);
    const double TRF = D.w15[i] * D.w15[j] * ga.JJ;
    zd r[3][10];
    taylor_point(D, pI, D.x15[i], D.x15[j], El, r);
    for (int k = 0; k < 3; ++k)
      for (int n = 0; n < 10; ++n) v[k][n] = r[k][n] * TRF;
  }
  block_reduce30(v, out + (size_t)blockIdx.x * 30);
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

}  // namespace

void launch_quad_scatter(int n, int nel, const int *r, const int *t, const zd *src, zd *sums,
                         hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_quad_scatter, dim3(nb((int64_t)n * 30, 256)), dim3(256), 0, st, n, nel, r, t,
                     src, sums);
}

void launch_quad_regular_grid(const QuadDev &d, const QuadTopo &T, int s0, int rows, zd *sums,
                              hipStream_t st) {
  const int64_t n = (int64_t)rows * T.nel;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_quad_regular_grid, dim3(nb(n, 128)), dim3(128), 0, st, d, T, s0, rows,
                     sums);
}
void launch_quad_regular_list(const QuadDev &d, const QuadTopo &T, int n, const int *s,
                              const int *t, const int *ngp, zd *out, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_quad_regular_list, dim3(nb(n, 128)), dim3(128), 0, st, d, T, n, s, t,
                     ngp, out);
}
void launch_quad_self(const QuadDev &d, const QuadTopo &T, int n, const int *s, zd *out,
                      hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_quad_self, dim3(n), dim3(256), 0, st, d, T, s, out);
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

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
namespace {

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
  // set_local_material_parameters_quad (src/lib_sie_quad_data_container.f90:118-129)
  typedef std::complex<double> C;
  const double k0 = 2 * kPi / params->lambda, omega = 2 * kPi * kC0 / params->lambda;
  const C er1(params->eps_r1_re, params->eps_r1_im), er2(params->eps_r2_re, params->eps_r2_im);
  const C er1v = (er1 == C(0.0)) ? C(1.0) : er1;
  const C k1 = k0 * std::sqrt(er1v), k2 = k0 * std::sqrt(er2);
  const C eta1 = std::sqrt(kMy0 / (kEps0 * er1v)), eta2 = std::sqrt(kMy0 / (kEps0 * er2));
  auto Z = [](C c) { return zd{c.real(), c.imag()}; };
  QuadDev &D = q->dev;
  D.k1 = Z(k1);
  D.k2 = Z(k2);
  D.ie1 = Z(1.0 / er1v);
  D.ie2 = Z(1.0 / er2);
  D.e1 = Z(er1v);
  D.e2 = Z(er2);
  D.ceps = Z(1.0 + 1.0 / er2);
  D.iwe = Z(C(0.0, 1.0) * omega * kEps0);
  D.iwm = Z(C(0.0, 1.0) * omega * kMy0);
  gauss_legendre(4, D.x4, D.w4);
  gauss_legendre(5, D.x5, D.w5);
  gauss_legendre(6, D.x6, D.w6);
  gauss_legendre(15, D.x15, D.w15);
  q->eta_a = Z((eta1 + eta2) / 2.0);
  q->eta12 = Z(eta1 * eta2);
  *out = q;
  return SPK_OK;
}

void spk_quad_destroy(spk_quad *q) {
  if (!q) return;
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
  std::vector<double> el((size_t)m.nel * 24);
  for (int e = 0; e < m.nel; ++e)
    for (int n = 0; n < 8; ++n)
      for (int d = 0; d < 3; ++d)
        el[(size_t)e * 24 + n * 3 + d] = m.pp[3 * (size_t)(m.vert[8 * (size_t)e + n] - 1) + d];
  std::vector<int> eadj((size_t)m.nel * kQuadMaxAdj, 0), eside((size_t)m.nel * kQuadMaxAdj * 2, 0),
      cadj((size_t)m.nel * kQuadMaxAdj, 0);
  for (int e = 0; e < m.nel; ++e) {
    if ((int)m.eadj[e].size() > kQuadMaxAdj || (int)m.cadj[e].size() > kQuadMaxAdj)
      return set_error(SPK_ERR_MESH, "element with more than 8 neighbours");
    for (size_t a = 0; a < m.eadj[e].size(); ++a) {
      eadj[(size_t)e * kQuadMaxAdj + a] = m.eadj[e][a].other;
      eside[((size_t)e * kQuadMaxAdj + a) * 2] = m.eadj[e][a].side[0];
      eside[((size_t)e * kQuadMaxAdj + a) * 2 + 1] = m.eadj[e][a].side[1];
    }
    for (size_t a = 0; a < m.cadj[e].size(); ++a) cadj[(size_t)e * kQuadMaxAdj + a] = m.cadj[e][a];
  }
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
  QBuf<int> a0, a1, a2, b0, c0, c1, c2;
  QBuf<zd> o1, o2, o3;
  QCHK(a0.upload(rs));
  QCHK(a1.upload(rt));
  QCHK(a2.upload(rg));
  QCHK(b0.upload(ss));
  QCHK(c0.upload(es));
  QCHK(c1.upload(et));
  QCHK(c2.upload(eside));
  QCHK(o1.alloc(rs.size() * 30));
  QCHK(o2.alloc(ss.size() * 30));
  QCHK(o3.alloc(es.size() * 30));
  const QuadTopo T = q->topo();
  launch_quad_regular_list(q->dev, T, (int)rs.size(), a0.p, a1.p, a2.p, o1.p, q->st);
  launch_quad_self(q->dev, T, (int)ss.size(), b0.p, o2.p, q->st);
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
  // rows of edge functions in chunks whose [rows][elements][30] sums fit 256 MiB
  const size_t per_row = (size_t)m.nel * 30 * sizeof(zd);
  const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)m.medge, (256u << 20) / per_row));
  QBuf<zd> sums, so, eo;
  QBuf<int> sl, el_s, el_t, el_side, sc_r, sc_t, sc_r2;
  QCHK(sums.alloc((size_t)chunk * m.nel * 30));
  for (int s0 = 0; s0 < m.medge; s0 += chunk) {
    const int rows = std::min(chunk, m.medge - s0);
    launch_quad_regular_grid(q->dev, T, s0, rows, sums.p, q->st);
    // self and edge-adjacent pairs of these rows
    std::vector<int> ss(rows), es, et, eside;
    for (int r = 0; r < rows; ++r) {
      ss[r] = s0 + r;
      const int ii = m.edge_p[5 * (size_t)(s0 + r)] - 1;
      for (const QuadEdgeAdj &a : m.eadj[ii]) {
        es.push_back(s0 + r);
        et.push_back(a.other - 1);
        eside.push_back(a.side[0]);
        eside.push_back(a.side[1]);
      }
    }
    QCHK(sl.upload(ss));
    QCHK(so.alloc((size_t)rows * 30));
    launch_quad_self(q->dev, T, rows, sl.p, so.p, q->st);
    QCHK(el_s.upload(es));
    QCHK(el_t.upload(et));
    QCHK(el_side.upload(eside));
    QCHK(eo.alloc(es.size() * 30));
    launch_quad_edge_adj(q->dev, T, (int)es.size(), el_s.p, el_t.p, el_side.p, eo.p, q->st);
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
    launch_quad_scatter(rows, m.nel, sc_r.p, sc_t.p, so.p, sums.p, q->st);
    QCHK(sc_r2.upload(er));
    launch_quad_scatter((int)es.size(), m.nel, sc_r2.p, el_t.p, eo.p, sums.p, q->st);
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
  QuadRhsArgs a{};
  for (int d = 0; d < 3; ++d) {
    a.kin[d] = k_in[d];
    a.ein[d] = e_in[d];
  }
  a.k1c = zd{q->dev.k1.x, -q->dev.k1.y};
  a.c0my = (double)kMy0 * (double)kC0;
  a.sE = zd{1.0, 0.0};
  a.sH = zd{1.0, 0.0};
  if (q->p.formulation == SPK_MCTF) {
    a.sH = q->eta12;
  } else if (q->p.formulation == SPK_ICTF) {
    a.sE = zd{1.0, 0.0} / q->eta_a;
    a.sH = q->eta_a;
  }
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

}  // extern "C"

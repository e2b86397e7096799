// spk_kernels.hip -- gfx950 kernels of the SIE solve loop.
//
// Near-field fill: one thread per 2x2 impedance block for the regular
// (3x3-rule) triangle pairs, separate task kernels for the singular
// (singularity-subtracted, Hanninen 2006) pairs and for the 12x12
// edge-adjacent K term, so that each launch runs one branch type.
// MLFMM: aggregation, SS, M2L, RR, fold + disaggregation, with the sample
// index s mapped to consecutive lanes everywhere (coalesced [..][S] rows).
// Near SpMV: one workgroup per (leaf box, row chunk) over the box's dense
// block matrix, x gathered once into LDS.
#include <hipcub/hipcub.hpp>

#include "spk_device.h"
#include "spk_fp8.h"

namespace spk {

constexpr double PI = kPi;

// ---------------------------------------------------------------------------
// geometry helpers (src/lib_sie_math.f90:550-597)
// ---------------------------------------------------------------------------
__device__ inline void simplex(double a, double b, const double *c, double *o) {
  const double w0 = 1 - a - b;
  o[0] = w0 * c[0] + a * c[3] + b * c[6];
  o[1] = w0 * c[1] + a * c[4] + b * c[7];
  o[2] = w0 * c[2] + a * c[5] + b * c[8];
}
__device__ inline void cross3(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ inline double len3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }
// cross_rc with the real operand already rounded to float (nqf)
__device__ inline void cross_rc(const double *a, const zd *b, zd *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ inline zd dot_rc(const double *a, const zd *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// exp(-i k R)
__device__ inline zd expmikr(zd k, double R) {
  double s, c;
  sincos(k.x * R, &s, &c);
  const double e = (k.y != 0.0) ? exp(k.y * R) : 1.0;
  return zd{e * c, -e * s};
}

// Diver_Green (src/lib_sie_tri_calculation_mod.f90:1421-1432); rinv = 1 / R
// (the fill kernels divide once per point pair)
__device__ inline void green(zd k, double R, double rinv, zd &G, zd &DG) {
  const zd E = expmikr(k, R);
  const zd ikr = zd{-k.y * R, k.x * R};
  const double inv = rinv * (1.0 / (4 * PI));
  G = E * inv;
  DG = -((zd{1.0 + ikr.x, ikr.y} * E) * (inv * rinv));
}
// Diver_Green_s (:1435-1451)
__device__ inline void green_s(zd k, double R, zd &Gs, zd &DGs) {
  const zd E = expmikr(k, R);
  const zd k2 = k * k;
  if (R == 0.0) {
    Gs = zd{k.y, -k.x} / (4 * PI);
    DGs = k2 / (8 * PI);
  } else {
    const zd ikr = zd{-k.y * R, k.x * R};
    Gs = (E - zd{1.0, 0.0}) / (4 * PI * R) + k2 * R / (8 * PI);
    DGs = (zd{1.0, 0.0} - zd{1.0 + ikr.x, ikr.y} * E) / (4 * PI * (R * R)) + k2 / (8 * PI);
  }
}

// ---------------------------------------------------------------------------
// Hanninen analytic integrals (PIER 63, 2006): Iq_L (:1074-1119),
// Iq_s (:1256-1331), K2q (:1333-1358), K3q (:1395-1419), K4q (:1360-1393).
// ---------------------------------------------------------------------------
__device__ void iq_l_all(const double *Ra, const double *p1, const double *p2, double *o) {
  double d[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const double l = len3(d);
  const double s[3] = {d[0] / l, d[1] / l, d[2] / l};
  const double t2[3] = {p2[0] - Ra[0], p2[1] - Ra[1], p2[2] - Ra[2]};
  const double t1[3] = {p1[0] - Ra[0], p1[1] - Ra[1], p1[2] - Ra[2]};
  const double Sp = dot3(t2, s), Sn = dot3(t1, s);
  const double a1[3] = {Ra[0] - p1[0], Ra[1] - p1[1], Ra[2] - p1[2]};
  const double a2[3] = {Ra[0] - p2[0], Ra[1] - p2[1], Ra[2] - p2[2]};
  const double Rn = len3(a1), Rp = len3(a2);
  const double Sum_p = Rp + Sp, Sum_n = Rn + Sn, Dif_p = Rp - Sp, Dif_n = Rn - Sn;
  const double R0 = sqrt(Rp * Rp - Sp * Sp);
  double n1;
  if (R0 == 0.0)
    n1 = log(Sp / Sn);
  else if (fabs(Dif_p) > fabs(Sum_n))
    n1 = log(Dif_n / Dif_p);
  else
    n1 = log(Sum_p / Sum_n);
  const double p1v = 0.5 * (R0 * R0) * n1 + 0.5 * (Sp * Rp - Sn * Rn);
  const double p3v = 0.75 * (R0 * R0) * p1v + 0.25 * (Sp * (Rp * Rp * Rp) - Sn * (Rn * Rn * Rn));
  o[0] = n1;
  o[1] = p1v;
  o[2] = p3v;
}

struct Anal {
  double Is_n3, Is_n1, Is_p1, h;
  double m[3][3];
  double ILst[3][3];
};

__device__ inline void edge_ps_pt(const double *c, int i, const double *&ps, const double *&pt) {
  if (i == 0) {
    ps = c + 6;
    pt = c + 3;
  } else if (i == 1) {
    ps = c;
    pt = c + 6;
  } else {
    ps = c + 3;
    pt = c;
  }
}

__device__ void analytic_eval(const double *Ra, const FnRec &f, bool iqs, Anal &A) {
  double arr[3][3];
  for (int i = 0; i < 3; ++i) {
    double d[3] = {f.c[3 * i] - Ra[0], f.c[3 * i + 1] - Ra[1], f.c[3 * i + 2] - Ra[2]};
    const double l = len3(d);
    arr[i][0] = d[0] / l;
    arr[i][1] = d[1] / l;
    arr[i][2] = d[2] / l;
  }
  const double x = 1 + dot3(arr[0], arr[1]) + dot3(arr[0], arr[2]) + dot3(arr[1], arr[2]);
  double r[3];
  cross3(arr[1], arr[2], r);
  const double y = fabs(dot3(arr[0], r));
  const double Omeg = 2 * atan2(y, x);
  const double dm[3] = {Ra[0] - f.mid[0], Ra[1] - f.mid[1], Ra[2] - f.mid[2]};
  const double h = dot3(f.nq, dm);
  double Is_n3 = 0.0;
  if (h != 0.0 && !iqs && Omeg <= PI && Omeg > -PI) Is_n3 = (-1 / h) * Omeg;
  // m_vector (:1040-1050)
  {
    const double *c = f.c;
    double d[3], xx[3];
    d[0] = c[6] - c[3]; d[1] = c[7] - c[4]; d[2] = c[8] - c[5];
    cross3(f.nq, d, xx);
    double l = len3(xx);
    A.m[0][0] = xx[0] / l; A.m[0][1] = xx[1] / l; A.m[0][2] = xx[2] / l;
    d[0] = c[0] - c[6]; d[1] = c[1] - c[7]; d[2] = c[2] - c[8];
    cross3(f.nq, d, xx);
    l = len3(xx);
    A.m[1][0] = xx[0] / l; A.m[1][1] = xx[1] / l; A.m[1][2] = xx[2] / l;
    d[0] = c[3] - c[0]; d[1] = c[4] - c[1]; d[2] = c[5] - c[2];
    cross3(f.nq, d, xx);
    l = len3(xx);
    A.m[2][0] = xx[0] / l; A.m[2][1] = xx[1] / l; A.m[2][2] = xx[2] / l;
  }
  double tv[3], ILts[3][3], sum_il = 0.0;
  for (int i = 0; i < 3; ++i) {
    const double *ps, *pt;
    edge_ps_pt(f.c, i, ps, pt);
    const double mp[3] = {Ra[0] - (pt[0] + ps[0]) / 2, Ra[1] - (pt[1] + ps[1]) / 2,
                          Ra[2] - (pt[2] + ps[2]) / 2};
    tv[i] = dot3(A.m[i], mp);
    iq_l_all(Ra, pt, ps, ILts[i]);
    iq_l_all(Ra, ps, pt, A.ILst[i]);
    sum_il = sum_il + tv[i] * ILts[i][0];
  }
  const double Is_n1 = 1.0 * ((-1 * (h * h)) * Is_n3 - sum_il);
  sum_il = 0.0;
  for (int i = 0; i < 3; ++i) sum_il = sum_il + tv[i] * ILts[i][1];
  const double Is_p1 = (double)(1.0f / 3.0f) * ((h * h) * Is_n1 - sum_il);
  A.Is_n3 = Is_n3;
  A.Is_n1 = Is_n1;
  A.Is_p1 = Is_p1;
  A.h = h;
}

__device__ inline double iq_s_val(const Anal &A, int q, bool iqs) {
  if (iqs && q > -1) return 0.0;
  if (q == -3) return A.Is_n3;
  if (q == -1) return A.Is_n1;
  return A.Is_p1;
}

__device__ void k2q_eval(int q, const double *Ra, const FnRec &f, bool iqs, const Anal &A, double *o) {
  const int col = (q == -1) ? 1 : 2;
  double sum[3] = {0, 0, 0};
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) sum[k] = sum[k] + A.ILst[i][col] * A.m[i][k];
  const double Iqs = iq_s_val(A, q, iqs);
  const double coef = (q == -1) ? 1.0 : (double)(1.0f / 3.0f);
  for (int k = 0; k < 3; ++k) {
    const double rho = Ra[k] - A.h * f.nq[k];
    o[k] = (coef * sum[k] + (rho - f.c[6 + k]) * Iqs) * f.df * 0.5;
  }
}
__device__ void k3q_eval(int q, const FnRec &f, bool iqs, const Anal &A, double *o) {
  const int col = (q == -1) ? 0 : 1;
  double sum[3] = {0, 0, 0};
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) sum[k] = sum[k] + A.m[i][k] * A.ILst[i][col];
  const double Is = iq_s_val(A, q - 2, iqs);
  for (int k = 0; k < 3; ++k) o[k] = sum[k] - ((A.h * (double)q) * f.nq[k]) * Is;
}
__device__ void k4q_eval(int q, const double *Ra, const FnRec &f, bool iqs, const Anal &A, double *o) {
  if (iqs) {
    o[0] = o[1] = o[2] = 0.0;
    return;
  }
  double k3[3], x[3];
  k3q_eval(q, f, iqs, A, k3);
  const double d[3] = {Ra[0] - f.c[6], Ra[1] - f.c[7], Ra[2] - f.c[8]};
  cross3(d, k3, x);
  for (int k = 0; k < 3; ++k) o[k] = (-x[k] * f.df) * 0.5;
}

// result_element_JMCFIE per medium
struct Rel {
  zd Kn, Kt, Tn, Tt;
};

// tri_outer_integration (:2255-2302), both media at once.
template <bool NEED_N>
__device__ void normal_pair(const FnRec &fm, const FnRec &fn, const NearArgs &A, Rel *rv) {
  const zd kk[2] = {zcx(A.mat.k1), zcx(A.mat.k2)};
  const zd ik[2] = {1.0 / kk[0], 1.0 / kk[1]};
  for (int md = 0; md < 2; ++md) rv[md] = Rel{{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  double x_nt = 0.0, x_nn = 0.0;
  const int ngp = A.ngp;
  for (int j = 0; j < ngp; ++j) {
    double ra[3], rfm[3], rfn[3];
    simplex(A.qa[j], A.qb[j], fm.c, ra);
    for (int k = 0; k < 3; ++k) {
      rfm[k] = 0.5 * (ra[k] - fm.c[6 + k]) * fm.df;
      rfn[k] = 0.5 * (ra[k] - fn.c[6 + k]) * fn.df;
    }
    const double TRF = A.qw[j] * fm.area;
    zd IK[2][3], ITa[2][3], ITbn[2][3], ITbt[2];
    for (int md = 0; md < 2; ++md) {
      ITbt[md] = zd{0, 0};
      for (int k = 0; k < 3; ++k) IK[md][k] = ITa[md][k] = ITbn[md][k] = zd{0, 0};
    }
    for (int i = 0; i < ngp; ++i) {
      double rq[3], rfq[3], R[3], Rn[3], x[3];
      simplex(A.qa[i], A.qb[i], fn.c, rq);
      for (int k = 0; k < 3; ++k) {
        rfq[k] = 0.5 * (rq[k] - fn.c[6 + k]) * fn.df;
        R[k] = ra[k] - rq[k];
      }
      const double Rl = len3(R), rinv = 1.0 / Rl;
      for (int k = 0; k < 3; ++k) Rn[k] = R[k] * rinv;
      cross3(rfq, Rn, x);
      const double TRFi = A.qw[i] * fn.area;
#pragma unroll
      for (int md = 0; md < 2; ++md) {
        zd G, DG;
        green(kk[md], Rl, rinv, G, DG);
        for (int k = 0; k < 3; ++k) {
          IK[md][k] = IK[md][k] + x[k] * DG * TRFi;
          ITa[md][k] = ITa[md][k] + rfq[k] * G * TRFi;
          if (NEED_N) ITbn[md][k] = ITbn[md][k] + (fn.df * Rn[k]) * DG * TRFi;
        }
        ITbt[md] = ITbt[md] + fn.df * G * TRFi;
      }
    }
#pragma unroll
    for (int md = 0; md < 2; ++md) {
      rv[md].Kt = rv[md].Kt + dot_rc(rfm, IK[md]) * TRF;
      const zd tmp_s = (ik[md] * fm.df) * ITbt[md];
      rv[md].Tt = rv[md].Tt + zd{0.0, TRF} * (kk[md] * dot_rc(rfm, ITa[md]) - tmp_s);
      if (NEED_N) {
        zd nk[3], av[3], tv[3];
        cross_rc(fm.nqf, IK[md], nk);
        rv[md].Kn = rv[md].Kn + dot_rc(rfm, nk) * TRF;
        for (int k = 0; k < 3; ++k) av[k] = kk[md] * ITa[md][k] + ik[md] * ITbn[md][k];
        cross_rc(fm.nqf, av, tv);
        rv[md].Tn = rv[md].Tn + zd{0.0, TRF} * dot_rc(rfm, tv);
      }
    }
    double nx[3];
    cross3(fm.nq, rfm, nx);
    x_nt = x_nt - dot3(nx, rfn) * TRF;
    x_nn = x_nn - dot3(rfm, rfn) * TRF;
  }
  const double fnt = (double)(float)x_nt, fnn = (double)(float)x_nn;
  rv[0].Kt = rv[0].Kt + zd{-0.5 * fnt, 0};
  rv[1].Kt = rv[1].Kt + zd{0.5 * fnt, 0};
  rv[0].Kn = rv[0].Kn + zd{-0.5 * fnn, 0};
  rv[1].Kn = rv[1].Kn + zd{0.5 * fnn, 0};
}

// tri_outer_integration_singular (:2199-2253) with the analytic part
// (:2115-2145) shared by both media and the smoothed kernel (:2080-2113).
__device__ void singular_pair(const FnRec &fm, const FnRec &fn, bool iqs, const NearArgs &A,
                              Rel *rv) {
  const zd kk[2] = {zcx(A.mat.k1), zcx(A.mat.k2)};
  const zd ik[2] = {1.0 / kk[0], 1.0 / kk[1]};
  for (int md = 0; md < 2; ++md) rv[md] = Rel{{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  double x_nt = 0.0, x_nn = 0.0;
  const int ngp = A.ngp;
  for (int j = 0; j < ngp; ++j) {
    double ra[3], rfm[3], rfn[3];
    simplex(A.qa[j], A.qb[j], fm.c, ra);
    for (int k = 0; k < 3; ++k) {
      rfm[k] = 0.5 * (ra[k] - fm.c[6 + k]) * fm.df;
      rfn[k] = 0.5 * (ra[k] - fn.c[6 + k]) * fn.df;
    }
    const double TRF = A.qw[j] * fm.area;
    Anal an;
    analytic_eval(ra, fn, iqs, an);
    double k2n[3], k2p[3], k3n[3], k3p[3], k4n[3], k4p[3];
    k2q_eval(-1, ra, fn, iqs, an, k2n);
    k3q_eval(-1, fn, iqs, an, k3n);
    k4q_eval(-1, ra, fn, iqs, an, k4n);
    k2q_eval(1, ra, fn, iqs, an, k2p);
    k3q_eval(1, fn, iqs, an, k3p);
    k4q_eval(1, ra, fn, iqs, an, k4p);
    const double Iqs_n1 = iq_s_val(an, -1, iqs), Iqs_p1 = iq_s_val(an, 1, iqs);
    for (int md = 0; md < 2; ++md) {
      const zd kd = kk[md], kd2 = kd * kd;
      zd IK[3], ITa[3], ITbn[3], ITbt;
      // analytic part
      for (int k = 0; k < 3; ++k) {
        IK[k] = iqs ? zd{0, 0} : zd{(1 / (4 * PI)) * k4n[k], 0} - kd2 / (8 * PI) * k4p[k];
        ITa[k] = zd{(1 / (4 * PI)) * k2n[k], 0} - kd2 / (8 * PI) * k2p[k];
        ITbn[k] = -(zd{(fn.df / (4 * PI)) * k3n[k], 0} - kd2 * fn.df / (8 * PI)) * k3p[k];
      }
      ITbt = zd{(fn.df / (4 * PI)) * Iqs_n1, 0} - kd2 * fn.df / (8 * PI) * Iqs_p1;
      // smooth part
      zd sIK[3] = {{0, 0}, {0, 0}, {0, 0}}, sITa[3] = {{0, 0}, {0, 0}, {0, 0}},
         sITbn[3] = {{0, 0}, {0, 0}, {0, 0}}, sITbt = {0, 0};
      for (int i = 0; i < ngp; ++i) {
        double rq[3], rfq[3], R[3];
        simplex(A.qa[i], A.qb[i], fn.c, rq);
        for (int k = 0; k < 3; ++k) {
          rfq[k] = 0.5 * (rq[k] - fn.c[6 + k]) * fn.df;
          R[k] = ra[k] - rq[k];
        }
        const double Rl = len3(R);
        zd Gs, DGs;
        green_s(kd, Rl, Gs, DGs);
        const double TRFi = A.qw[i] * fn.area;
        if (!iqs) {
          double Rn[3] = {R[0] / Rl, R[1] / Rl, R[2] / Rl}, x[3];
          cross3(rfq, Rn, x);
          for (int k = 0; k < 3; ++k) {
            sIK[k] = sIK[k] + x[k] * DGs * TRFi;
            sITbn[k] = sITbn[k] + (fn.df * Rn[k]) * DGs * TRFi;
          }
        }
        for (int k = 0; k < 3; ++k) sITa[k] = sITa[k] + rfq[k] * Gs * TRFi;
        sITbt = sITbt + fn.df * Gs * TRFi;
      }
      for (int k = 0; k < 3; ++k) {
        ITbn[k] = ITbn[k] + sITbn[k];
        IK[k] = IK[k] + sIK[k];
        ITa[k] = ITa[k] + sITa[k];
      }
      ITbt = ITbt + sITbt;
      zd nk[3], av[3], t1[3], t2[3], tv[3];
      cross_rc(fm.nqf, IK, nk);
      rv[md].Kn = rv[md].Kn + dot_rc(rfm, nk) * TRF;
      rv[md].Kt = rv[md].Kt + dot_rc(rfm, IK) * TRF;
      for (int k = 0; k < 3; ++k) av[k] = kd * ITa[k];
      cross_rc(fm.nqf, av, t1);
      for (int k = 0; k < 3; ++k) av[k] = ik[md] * ITbn[k];
      cross_rc(fm.nqf, av, t2);
      for (int k = 0; k < 3; ++k) tv[k] = t1[k] + t2[k];
      rv[md].Tn = rv[md].Tn + zd{0.0, TRF} * dot_rc(rfm, tv);
      const zd tmp_s = (ik[md] * fm.df) * ITbt;
      rv[md].Tt = rv[md].Tt + zd{0.0, TRF} * (kd * dot_rc(rfm, ITa) - tmp_s);
    }
    double nx[3];
    cross3(fm.nq, rfm, nx);
    x_nt = x_nt - dot3(nx, rfn) * TRF;
    x_nn = x_nn - dot3(rfm, rfn) * TRF;
  }
  const double fnt = (double)(float)x_nt, fnn = (double)(float)x_nn;
  rv[0].Kt = rv[0].Kt + zd{-0.5 * fnt, 0};
  rv[1].Kt = rv[1].Kt + zd{0.5 * fnt, 0};
  rv[0].Kn = rv[0].Kn + zd{-0.5 * fnn, 0};
  rv[1].Kn = rv[1].Kn + zd{0.5 * fnn, 0};
}

// tri_outer_integration_edge_adjacent (:2165-2197), 12x12 rule, both media.
__device__ void edge_adjacent_kt(const FnRec &fm, const FnRec &fn, const NearArgs &A, zd *kt) {
  const zd kk[2] = {zcx(A.mat.k1), zcx(A.mat.k2)};
  zd K[2] = {{0, 0}, {0, 0}};
  double x_nt = 0.0;
  for (int j = 0; j < 12; ++j) {
    double ra[3], rfm[3], rfn[3];
    simplex(A.qa12[j], A.qb12[j], fm.c, ra);
    for (int k = 0; k < 3; ++k) {
      rfm[k] = 0.5 * (ra[k] - fm.c[6 + k]) * fm.df;
      rfn[k] = 0.5 * (ra[k] - fn.c[6 + k]) * fn.df;
    }
    const double TRF = A.qw12[j] * fm.area;
    zd Ik[2][3] = {{{0, 0}, {0, 0}, {0, 0}}, {{0, 0}, {0, 0}, {0, 0}}};
    for (int i = 0; i < 12; ++i) {
      double rq[3], rfq[3], R[3], Rn[3], x[3];
      simplex(A.qa12[i], A.qb12[i], fn.c, rq);
      for (int k = 0; k < 3; ++k) {
        rfq[k] = 0.5 * (rq[k] - fn.c[6 + k]) * fn.df;
        R[k] = ra[k] - rq[k];
      }
      const double Rl = len3(R), rinv = 1.0 / Rl;
      for (int k = 0; k < 3; ++k) Rn[k] = R[k] * rinv;
      cross3(rfq, Rn, x);
      const double TRFi = A.qw12[i] * fn.area;
#pragma unroll
      for (int md = 0; md < 2; ++md) {
        zd G, DG;
        green(kk[md], Rl, rinv, G, DG);
        for (int k = 0; k < 3; ++k) Ik[md][k] = Ik[md][k] + x[k] * DG * TRFi;
      }
    }
    for (int md = 0; md < 2; ++md) K[md] = K[md] + dot_rc(rfm, Ik[md]) * TRF;
    double nx[3];
    cross3(fm.nq, rfm, nx);
    x_nt = x_nt - dot3(rfn, nx) * TRF;
  }
  const double fnt = (double)(float)x_nt;
  kt[0] = K[0] + zd{-0.5 * fnt, 0};
  kt[1] = K[1] + zd{0.5 * fnt, 0};
}

// tri_impedance_building_* (:1713-2009): per-pair zz
__device__ void assemble_zz(const NearArgs &A, const Rel *r, zd *zz) {
  const zd e1 = zcx(A.mat.eta1), e2 = zcx(A.mat.eta2);
  const zd one = {1.0, 0.0};
  switch (A.form) {
    case SPK_PMCHWT:
    case SPK_MCTF:
    case SPK_PMCHWT_S:
      zz[0] = e1 * r[0].Tt + e2 * r[1].Tt;
      zz[1] = -(r[0].Kt + r[1].Kt);
      zz[2] = (one / e1) * r[0].Tt + (one / e2) * r[1].Tt;
      zz[3] = zd{0, 0};
      break;
    case SPK_JMCFIE:
      zz[0] = r[0].Kn - r[1].Kn + r[0].Tt + r[1].Tt;
      zz[1] = -((one / e1) * (r[0].Kt - r[0].Tn) + (one / e2) * (r[1].Kt + r[1].Tn));
      zz[2] = -(e1 * (r[0].Tn - r[0].Kt)) + e2 * (r[1].Kt + r[1].Tn);
      zz[3] = zd{0, 0};
      break;
    case SPK_MCTF2: {
      const zd a0 = zcx(A.mat.a_lc[0]), a1 = zcx(A.mat.a_lc[1]);
      const zd d0 = zcx(A.mat.d_lc[0]), d1 = zcx(A.mat.d_lc[1]);
      zz[0] = r[0].Tt * a0 + r[1].Tt * a1;
      zz[1] = -(r[0].Kt * a0 / e1 + r[1].Kt * a1 / e2);
      zz[2] = e1 * d0 * r[0].Kt + r[1].Kt * e2 * d1;
      zz[3] = d0 * r[0].Tt + d1 * r[1].Tt;
    } break;
    case SPK_ICTF: {
      const zd ea = zcx(A.mat.eta_a);
      zz[0] = (e1 * r[0].Tt + e2 * r[1].Tt) / ea;
      zz[1] = -(r[0].Kt + r[1].Kt) / ea;
      zz[2] = (r[0].Kt + r[1].Kt) * ea;
      zz[3] = (r[0].Tt / e1 + r[1].Tt / e2) * ea;
    } break;
    case SPK_CNF:
      zz[0] = r[0].Kn - r[1].Kn;
      zz[1] = r[0].Tn / e1 - r[1].Tn / e2;
      zz[2] = e2 * r[1].Tn - e1 * r[0].Tn;
      zz[3] = zd{0, 0};
      break;
    default: {  // MNMF
      const zd ep1 = zcx(A.mat.eps1), ep2 = zcx(A.mat.eps2);
      const zd my1 = {A.mat.my1, 0}, my2 = {A.mat.my2, 0};
      const zd b1 = ep1 / (ep1 + ep2), b2 = ep2 / (ep1 + ep2);
      const zd c1 = my1 / (my1 + my2), c2 = my2 / (my1 + my2);
      zz[0] = b1 * r[0].Kn - b2 * r[1].Kn;
      zz[1] = b1 / e1 * r[0].Tn - b2 / e2 * r[1].Tn;
      zz[2] = e2 * c2 * r[1].Tn - e1 * c1 * r[0].Tn;
      zz[3] = c1 * r[0].Kn - c2 * r[1].Kn;
    } break;
  }
}

// lib_sie_tri_get_impedance_edges block assembly (:463-644)
__device__ inline void accumulate_D(const NearArgs &A, zd *D, const zd *zz) {
  switch (A.form) {
    case SPK_PMCHWT:
      D[0] = D[0] + zz[0];
      D[1] = D[1] + zz[1];
      D[2] = -D[1];
      D[3] = D[3] + zz[2];
      break;
    case SPK_MCTF: {
      const zd e12 = zcx(A.mat.eta1) * zcx(A.mat.eta2);
      D[0] = D[0] + zz[0];
      D[1] = D[1] + zz[1];
      D[2] = -D[1] * e12;
      D[3] = D[3] + zz[2] * e12;
    } break;
    case SPK_PMCHWT_S:
      D[0] = D[0] + zz[0];
      D[1] = D[1] + zz[1];
      D[2] = -D[2] + zz[1];
      D[3] = D[3] + zz[2];
      break;
    case SPK_JMCFIE:
      D[0] = D[0] + zz[0];
      D[1] = D[1] + zz[1];
      D[2] = D[2] + zz[2];
      D[3] = D[0];
      break;
    default:
      D[0] = D[0] + zz[0];
      D[1] = D[1] + zz[1];
      D[2] = D[2] + zz[2];
      D[3] = D[3] + zz[3];
      break;
  }
}

__device__ inline void finish_D(const NearArgs &A, zd *D) {
  if (A.form == SPK_PMCHWT_S) {
    const zd e1 = zcx(A.mat.eta1);
    D[1] = D[1] * e1;
    D[2] = D[2] * e1;
    D[3] = D[3] * (e1 * e1);
  }
}

// block index -> (row edge m, column edge n)
__device__ inline void decode_block(const NearArgs &A, int64_t g, int &m, int &n) {
  int lo = 0, hi = A.nbl - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (A.near_off[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int b = lo;
  const int cols = A.e2_off[b + 1] - A.e2_off[b];
  const int64_t rel = g - A.near_off[b];
  const int r = (int)(rel / cols), c = (int)(rel - (int64_t)r * cols);
  m = A.perm[A.box_first[b] + r];
  n = A.e2_list[A.e2_off[b] + c];
}

__device__ inline bool edge_adj(const ElemRec &a, const ElemRec &b) {
  int cnt = 0;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) cnt += (a.node[j] == b.node[i]);
  return cnt == 2;
}

__global__ void k_near_classify(NearArgs A, int64_t nnz, uint8_t *cls, int *cnt_a, int *cnt_b) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nnz) return;
  int m, n;
  decode_block(A, g, m, n);
  const double rr = 1.5 * A.rref[m];
  unsigned bits = 0;
  for (int p = 0; p < 4; ++p) {
    const int t = p >> 1, s = p & 1;
    const FnRec &fm = A.fn[2 * m + t];
    const FnRec &fq = A.fn[2 * n + s];
    const double d[3] = {fm.mid[0] - fq.mid[0], fm.mid[1] - fq.mid[1], fm.mid[2] - fq.mid[2]};
    const double r_ob = len3(d);
    if (!(r_ob > rr)) {
      bits |= 1u << p;
      const ElemRec &ea = A.er[2 * m + t], &eb = A.er[2 * n + s];
      if (A.edge_swap && ea.elem != eb.elem && edge_adj(ea, eb)) bits |= 16u << p;
    }
  }
  cls[g] = (uint8_t)bits;
  cnt_a[g] = __popc(bits & 15u);
  cnt_b[g] = __popc(bits >> 4);
}

__global__ void k_near_emit(NearArgs A, int64_t nnz, const uint8_t *cls, const int *base_a,
                            const int *base_b, NearTask *ta, NearTask *tb) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nnz) return;
  const unsigned bits = cls[g];
  if (!bits) return;
  int ra = base_a[g], rb = base_b[g];
  for (int p = 0; p < 4; ++p) {
    if (bits & (1u << p)) ta[ra++] = NearTask{g, p, 0};
    if (bits & (16u << p)) tb[rb++] = NearTask{g, p, 0};
  }
}

__global__ void __launch_bounds__(128) k_near_edge(NearArgs A, int ntask, const NearTask *tb, zd *kt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntask) return;
  int m, n;
  decode_block(A, tb[t].g, m, n);
  const int p = tb[t].p;
  const FnRec fm = A.fn[2 * m + (p >> 1)];
  const FnRec fq = A.fn[2 * n + (p & 1)];
  zd k2[2];
  edge_adjacent_kt(fm, fq, A, k2);
  kt[2 * t] = k2[0];
  kt[2 * t + 1] = k2[1];
}

__global__ void __launch_bounds__(128) k_near_singular(NearArgs A, int ntask, const NearTask *ta,
                                                       const uint8_t *cls, const int *base_b,
                                                       const zd *kt, zd *zz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntask) return;
  const int64_t g = ta[t].g;
  const int p = ta[t].p;
  int m, n;
  decode_block(A, g, m, n);
  const FnRec fm = A.fn[2 * m + (p >> 1)];
  const FnRec fq = A.fn[2 * n + (p & 1)];
  const bool iqs = A.er[2 * m + (p >> 1)].elem == A.er[2 * n + (p & 1)].elem;
  Rel rv[2];
  singular_pair(fm, fq, iqs, A, rv);
  const unsigned bits = cls[g];
  if (bits & (16u << p)) {
    const int rank = __popc((bits >> 4) & ((1u << p) - 1u));
    const int idx = base_b[g] + rank;
    rv[0].Kt = kt[2 * idx];
    rv[1].Kt = kt[2 * idx + 1];
  }
  zd z[4];
  assemble_zz(A, rv, z);
  for (int k = 0; k < 4; ++k) zz[4 * (int64_t)t + k] = z[k];
}

template <bool NEED_N>
__global__ void __launch_bounds__(256) k_near_regular(NearArgs A, int64_t nnz, const uint8_t *cls,
                                                      const int *base_a, const zd *zz, zd *blocks) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nnz) return;
  int m, n;
  decode_block(A, g, m, n);
  const unsigned bits = cls[g];
  zd D[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (int p = 0; p < 4; ++p) {
    zd z[4];
    if (bits & (1u << p)) {
      const int rank = __popc(bits & ((1u << p) - 1u));
      const int64_t idx = base_a[g] + rank;
      for (int k = 0; k < 4; ++k) z[k] = zz[4 * idx + k];
    } else {
      const FnRec &fm = A.fn[2 * m + (p >> 1)];
      const FnRec &fq = A.fn[2 * n + (p & 1)];
      Rel rv[2];
      normal_pair<NEED_N>(fm, fq, A, rv);
      assemble_zz(A, rv, z);
    }
    accumulate_D(A, D, z);
  }
  finish_D(A, D);
  for (int k = 0; k < 4; ++k) blocks[4 * g + k] = D[k];
}

static inline unsigned nblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }
static inline unsigned gcap(int64_t n) { unsigned b = nblk(n, 256); return b < 2048u ? (b ? b : 1u) : 2048u; }

// One thread per requested (m, n): all four triangle pairs with the
// reference's branch rules inline (parity probe for arbitrary pairs).
__global__ void __launch_bounds__(64) k_near_pairs(NearArgs A, int np_, const int *pm,
                                                   const int *pn, zd *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= np_) return;
  const int m = pm[t], n = pn[t];
  const double rr = 1.5 * A.rref[m];
  zd D[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (int p = 0; p < 4; ++p) {
    const FnRec fm = A.fn[2 * m + (p >> 1)];
    const FnRec fq = A.fn[2 * n + (p & 1)];
    const ElemRec ea = A.er[2 * m + (p >> 1)], eb = A.er[2 * n + (p & 1)];
    const double d[3] = {fm.mid[0] - fq.mid[0], fm.mid[1] - fq.mid[1], fm.mid[2] - fq.mid[2]};
    Rel rv[2];
    if (!(len3(d) > rr)) {
      const bool iqs = ea.elem == eb.elem;
      singular_pair(fm, fq, iqs, A, rv);
      if (A.edge_swap && !iqs && edge_adj(ea, eb)) {
        zd kt[2];
        edge_adjacent_kt(fm, fq, A, kt);
        rv[0].Kt = kt[0];
        rv[1].Kt = kt[1];
      }
    } else {
      normal_pair<true>(fm, fq, A, rv);
    }
    zd z[4];
    assemble_zz(A, rv, z);
    accumulate_D(A, D, z);
  }
  finish_D(A, D);
  for (int k = 0; k < 4; ++k) out[4 * t + k] = D[k];
}

void launch_near_pairs(const NearArgs &a, int npairs, const int *m, const int *n, zd *out,
                       hipStream_t st) {
  if (npairs <= 0) return;
  hipLaunchKernelGGL(k_near_pairs, dim3(nblk(npairs, 64)), dim3(64), 0, st, a, npairs, m, n, out);
}

void launch_near_classify(const NearArgs &a, int64_t nnz, uint8_t *cls, int *cnt_a, int *cnt_b,
                          hipStream_t st) {
  if (nnz <= 0) return;
  hipLaunchKernelGGL(k_near_classify, dim3(nblk(nnz, 256)), dim3(256), 0, st, a, nnz, cls, cnt_a,
                     cnt_b);
}
void launch_near_emit(const NearArgs &a, int64_t nnz, const uint8_t *cls, const int *base_a,
                      const int *base_b, NearTask *ta, NearTask *tb, hipStream_t st) {
  if (nnz <= 0) return;
  hipLaunchKernelGGL(k_near_emit, dim3(nblk(nnz, 256)), dim3(256), 0, st, a, nnz, cls, base_a,
                     base_b, ta, tb);
}
void launch_near_edge(const NearArgs &a, int ntask, const NearTask *tb, zd *kt, hipStream_t st) {
  if (ntask <= 0) return;
  hipLaunchKernelGGL(k_near_edge, dim3(nblk(ntask, 128)), dim3(128), 0, st, a, ntask, tb, kt);
}
void launch_near_singular(const NearArgs &a, int ntask, const NearTask *ta, const uint8_t *cls,
                          const int *base_b, const zd *kt, zd *zz, hipStream_t st) {
  if (ntask <= 0) return;
  hipLaunchKernelGGL(k_near_singular, dim3(nblk(ntask, 128)), dim3(128), 0, st, a, ntask, ta, cls,
                     base_b, kt, zz);
}
void launch_near_regular(const NearArgs &a, int64_t nnz, const uint8_t *cls, const int *base_a,
                         const zd *zz, zd *blocks, hipStream_t st) {
  if (nnz <= 0) return;
  const bool need_n = (a.form == SPK_JMCFIE || a.form == SPK_CNF || a.form == SPK_MNMF);
  if (need_n)
    hipLaunchKernelGGL(k_near_regular<true>, dim3(nblk(nnz, 256)), dim3(256), 0, st, a, nnz, cls,
                       base_a, zz, blocks);
  else
    hipLaunchKernelGGL(k_near_regular<false>, dim3(nblk(nnz, 256)), dim3(256), 0, st, a, nnz, cls,
                       base_a, zz, blocks);
}

hipError_t exclusive_scan_int(const int *in, int *out, int64_t n, void **tmp, size_t *tmp_bytes,
                              hipStream_t st) {
  size_t need = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, need, in, out, (int)n, st);
  if (e != hipSuccess) return e;
  if (need > *tmp_bytes || !*tmp) {
    if (*tmp) (void)hipFree(*tmp);
    *tmp = nullptr;
    *tmp_bytes = 0;
    e = hipMalloc(tmp, need ? need : 1);
    if (e != hipSuccess) {
      *tmp = nullptr;
      return e;
    }
    *tmp_bytes = need;
  }
  return hipcub::DeviceScan::ExclusiveSum(*tmp, need, in, out, (int)n, st);
}

// ---------------------------------------------------------------------------
// Block-Jacobi preconditioner: dense group matrices from the near blocks
// ---------------------------------------------------------------------------
// One workgroup per (group, row): the row's near columns that fall inside
// the group are written into A_g; the padding diagonal is set by the row
// workgroups beyond 2 cnt.
__global__ void __launch_bounds__(256) k_blu_extract(int nmax, const int *p0, const int *cnt,
                                                     const int *perm, const int *pos,
                                                     const int *leafbox_of_pos,
                                                     const int *box_first, const int *e2_off,
                                                     const int *e2_list, const int64_t *near_off,
                                                     const zd *blocks, zd *A) {
  const int g = blockIdx.y, r = blockIdx.x;
  const int c0 = cnt[g], q0 = p0[g];
  zd *Ag = A + (size_t)g * nmax * nmax;
  if (r >= c0) {
    // padding rows/columns 2 c0 .. nmax-1: identity (two per workgroup)
    const int i0 = 2 * c0 + 2 * (r - c0);
    if (threadIdx.x < 2 && i0 + (int)threadIdx.x < nmax) {
      const int i = i0 + threadIdx.x;
      Ag[(size_t)i * nmax + i] = zd{1.0, 0.0};
    }
    return;
  }
  const int p = q0 + r;
  const int b = leafbox_of_pos[p];
  const int cols = e2_off[b + 1] - e2_off[b];
  const int rr = p - box_first[b];
  const zd *row = blocks + 4 * (size_t)(near_off[b] + (int64_t)rr * cols);
  for (int c = threadIdx.x; c < cols; c += blockDim.x) {
    const int q = pos[e2_list[e2_off[b] + c]] - q0;
    if (q < 0 || q >= c0) continue;
    const zd *bl = row + 4 * (size_t)c;
    Ag[(size_t)q * nmax + r] = bl[0];
    Ag[(size_t)(c0 + q) * nmax + r] = bl[1];
    Ag[(size_t)q * nmax + c0 + r] = bl[2];
    Ag[(size_t)(c0 + q) * nmax + c0 + r] = bl[3];
  }
}

__global__ void k_blu_gather(int nmax, int N, const int *p0, const int *cnt, const int *perm,
                             const zd *v, zd *X) {
  const int g = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmax) return;
  const int c0 = cnt[g];
  zd val = {0, 0};
  if (i < c0)
    val = v[perm[p0[g] + i]];
  else if (i < 2 * c0)
    val = v[N + perm[p0[g] + i - c0]];
  X[(size_t)g * nmax + i] = val;
}

__global__ void k_blu_scatter(int nmax, int N, const int *p0, const int *cnt, const int *perm,
                              const zd *Y, zd *z) {
  const int g = blockIdx.y, i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c0 = cnt[g];
  if (i >= 2 * c0) return;
  const zd val = Y[(size_t)g * nmax + i];
  if (i < c0)
    z[perm[p0[g] + i]] = val;
  else
    z[N + perm[p0[g] + i - c0]] = val;
}

void launch_blu_extract(int ngroups, int nmax, const int *p0, const int *cnt, const int *perm,
                        const int *pos, const int *leafbox_of_pos, const int *box_first,
                        const int *e2_off, const int *e2_list, const int64_t *near_off,
                        const zd *blocks, zd *A, hipStream_t st) {
  if (ngroups <= 0) return;
  // rows 0 .. nmax/2: real rows below cnt, padding pairs above
  hipLaunchKernelGGL(k_blu_extract, dim3((nmax + 1) / 2, ngroups), dim3(256), 0, st, nmax, p0, cnt,
                     perm, pos, leafbox_of_pos, box_first, e2_off, e2_list, near_off, blocks, A);
}
void launch_blu_gather(int ngroups, int nmax, int N, const int *p0, const int *cnt,
                       const int *perm, const zd *v, zd *X, hipStream_t st) {
  if (ngroups <= 0) return;
  hipLaunchKernelGGL(k_blu_gather, dim3(nblk(nmax, 256), ngroups), dim3(256), 0, st, nmax, N, p0,
                     cnt, perm, v, X);
}
void launch_blu_scatter(int ngroups, int nmax, int N, const int *p0, const int *cnt,
                        const int *perm, const zd *Y, zd *z, hipStream_t st) {
  if (ngroups <= 0) return;
  hipLaunchKernelGGL(k_blu_scatter, dim3(nblk(nmax, 256), ngroups), dim3(256), 0, st, nmax, N, p0,
                     cnt, perm, Y, z);
}

// ---------------------------------------------------------------------------
// MLFMM kernels
// ---------------------------------------------------------------------------
// PreCalculate_RT / RT_function_SCSArr_New (src/lib_sie_tri_mlfmm.f90:503-583):
// one thread per (position, sample).
__global__ void k_patterns(int N, int S, int ngp, const FnRec *fn, const int *perm,
                           const int *leaf_box, const double *leaf_centre, const double *khat,
                           const double *tth, const double *tph, const double *qa,
                           const double *qb, const double *qw, zd k1, zd k2, zd *rt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (i >= S || p >= N) return;
  const int m = perm[p];
  const double *Rc = &leaf_centre[3 * leaf_box[m]];
  const double kh[3] = {khat[3 * i], khat[3 * i + 1], khat[3 * i + 2]};
  zd o[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (int ll = 0; ll < 2; ++ll) {
    const FnRec &f = fn[2 * m + ll];
    const double dfm = f.sign * f.length;
    zd sa[3] = {{0, 0}, {0, 0}, {0, 0}}, sb[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int j = 0; j < ngp; ++j) {
      double Ro[3], fm[3], Roc[3];
      simplex(qa[j], qb[j], f.c, Ro);
      for (int k = 0; k < 3; ++k) {
        fm[k] = 0.5 * (Ro[k] - f.c[6 + k]) * dfm;
        Roc[k] = Ro[k] - Rc[k];
      }
      const double tmp = dot3(kh, Roc);
      const zd ff1 = expmikr(k1, tmp) * qw[j];
      const zd ff2 = expmikr(k2, tmp) * qw[j];
      for (int k = 0; k < 3; ++k) {
        sa[k] = sa[k] + fm[k] * ff1;
        sb[k] = sb[k] + fm[k] * ff2;
      }
    }
    const double tt3[3] = {tth[3 * i], tth[3 * i + 1], tth[3 * i + 2]};
    const double tp3[3] = {tph[3 * i], tph[3 * i + 1], tph[3 * i + 2]};
    o[0] = o[0] + dot_rc(tt3, sa);
    o[1] = o[1] + dot_rc(tp3, sa);
    o[2] = o[2] + dot_rc(tt3, sb);
    o[3] = o[3] + dot_rc(tp3, sb);
  }
  zd *out = rt + (size_t)p * 4 * S;
  for (int q = 0; q < 4; ++q) out[(size_t)q * S + i] = o[q];
}

void launch_patterns(int N, int S, int ngp, const FnRec *fn, const int *perm, const int *leaf_box,
                     const double *leaf_centre, const double *khat, const double *tth,
                     const double *tph, const double *qa, const double *qb, const double *qw,
                     zd k1, zd k2, zd *rt, hipStream_t st) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_patterns, dim3(nblk(S, 256), N), dim3(256), 0, st, N, S, ngp, fn, perm,
                     leaf_box, leaf_centre, khat, tth, tph, qa, qb, qw, k1, k2, rt);
}

// The matvec kernels below are templates on the complex type T: zd
// (complex128) for the FP64 operator, zf (complex64) for the mixed-precision
// inner operator.  Sums accumulate in T.

// lib_sie_ml_fmm_get_c (src/lib_sie_tri_mlfmm_interface.f90:395-448):
// C[b][comp*4+cur*2+k][s] = sum_{edges of b} conj(RT[k][comp][s]) * x_cur
// complex64 values are moved in pairs (one 16-byte load covers two values)
template <class T>
constexpr int kPack = sizeof(T) == 8 ? 2 : 1;
struct zf2 {
  zf a, b;
};
__device__ inline zf2 ld2(const zf *p) {
  const float4 v = *reinterpret_cast<const float4 *>(p);
  return zf2{zf{v.x, v.y}, zf{v.z, v.w}};
}

// CONJ = false: the patterns themselves (the transposed operator's
// disaggregation, aggregated)
template <class T, bool CONJ>
__global__ void k_agg(int S, int N, const int *box_first, const int *box_count, const int *perm,
                      const T *rt, const T *x, T *C) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y, q = blockIdx.z;  // q = k*2 + comp
  if (s >= S) return;
  const int k = q >> 1, comp = q & 1;
  T ae = {0, 0}, ah = {0, 0};
  const int p0 = box_first[b], cnt = box_count[b];
  for (int e0 = 0; e0 < cnt; e0 += kSumBlock<T>) {
    T ge = {0, 0}, gh = {0, 0};
    const int e1 = min(cnt, e0 + kSumBlock<T>);
#pragma unroll 4
    for (int e = e0; e < e1; ++e) {
      const int p = p0 + e;
      const int m = perm[p];
      const T r = rt[((size_t)p * 4 + q) * S + s];
      const T f = CONJ ? conj(r) : r;
      const T xe = x[m], xh = x[m + N];
      ge = ge + f * xe;
      gh = gh + f * xh;
    }
    ae = ae + ge;
    ah = ah + gh;
  }
  T *Cb = C + (size_t)b * 8 * S;
  Cb[(size_t)(comp * 4 + 0 * 2 + k) * S + s] = ae;
  Cb[(size_t)(comp * 4 + 1 * 2 + k) * S + s] = ah;
}

template <class T>
void launch_agg(int nbl, int S, int N, const int *box_first, const int *box_count,
                const int *perm, const T *rt, const T *x, T *C, hipStream_t st, bool conj_rt) {
  if (nbl <= 0) return;
  if (conj_rt)
    hipLaunchKernelGGL((k_agg<T, true>), dim3(nblk(S, 256), nbl, 4), dim3(256), 0, st, S, N,
                       box_first, box_count, perm, rt, x, C);
  else
    hipLaunchKernelGGL((k_agg<T, false>), dim3(nblk(S, 256), nbl, 4), dim3(256), 0, st, S, N,
                       box_first, box_count, perm, rt, x, C);
}

// translation_SS (:466-528): P[it][s] = sum_o C_child[it][s] * up[o][k][s]
template <class T>
__global__ void k_ss(int S, const int *child, const T *Cc, const T *up, T *P) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (s >= S) return;
  T acc[8];
  for (int it = 0; it < 8; ++it) acc[it] = T{0, 0};
  for (int o = 0; o < 8; ++o) {
    const int c = child[8 * b + o];
    if (c < 0) continue;
    const T u0 = up[((size_t)o * 2 + 0) * S + s], u1 = up[((size_t)o * 2 + 1) * S + s];
    const T *src = Cc + (size_t)c * 8 * S;
#pragma unroll
    for (int it = 0; it < 8; ++it) acc[it] = acc[it] + src[(size_t)it * S + s] * ((it & 1) ? u1 : u0);
  }
  T *dst = P + (size_t)b * 8 * S;
  for (int it = 0; it < 8; ++it) dst[(size_t)it * S + s] = acc[it];
}

template <class T>
void launch_ss(int nb, int S, const int *child, const T *Cc, const T *up, T *P, hipStream_t st) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_ss<T>, dim3(nblk(S, 256), nb), dim3(256), 0, st, S, child, Cc, up, P);
}

// translation_SR (:547-586): Dt[a][it][s] = sum_{b in I4(a)} B[b][it][s]*TLT[slot][k][s]
template <class T>
__global__ void k_m2l(int S, const int *i4_off, const int *i4_box, const int *i4_slot,
                      const T *B, const T *tlt, T *Dt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (s >= S) return;
  T acc[8];
  for (int it = 0; it < 8; ++it) acc[it] = T{0, 0};
  const int q0 = i4_off[a], q1 = i4_off[a + 1];
  for (int g0 = q0; g0 < q1; g0 += kSumBlock<T>) {
    T gs[8];
#pragma unroll
    for (int it = 0; it < 8; ++it) gs[it] = T{0, 0};
    const int g1 = min(q1, g0 + kSumBlock<T>);
    for (int q = g0; q < g1; ++q) {
      const int b = i4_box[q], slot = i4_slot[q];
      const T t0 = tlt[((size_t)slot * 2 + 0) * S + s], t1 = tlt[((size_t)slot * 2 + 1) * S + s];
      const T *src = B + (size_t)b * 8 * S;
#pragma unroll
      for (int it = 0; it < 8; ++it) gs[it] = gs[it] + src[(size_t)it * S + s] * ((it & 1) ? t1 : t0);
    }
#pragma unroll
    for (int it = 0; it < 8; ++it) acc[it] = acc[it] + gs[it];
  }
  T *dst = Dt + (size_t)a * 8 * S;
  for (int it = 0; it < 8; ++it) dst[(size_t)it * S + s] = acc[it];
}

template <class T>
void launch_m2l(int nb, int S, const int *i4_off, const int *i4_box, const int *i4_slot,
                const T *B, const T *tlt, T *Dt, hipStream_t st) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_m2l<T>, dim3(nblk(S, 128), nb), dim3(128), 0, st, S, i4_off, i4_box,
                     i4_slot, B, tlt, Dt);
}

// translation_RR (:600-655): Dc[it][s] += Dp[it][s]*down[o][k][s]
template <class T>
__global__ void k_rr(int S, const int *parent, const int *octant, const T *Dp, const T *down,
                     T *Dc) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (s >= S) return;
  const int p = parent[b], o = octant[b];
  const T d0 = down[((size_t)o * 2 + 0) * S + s], d1 = down[((size_t)o * 2 + 1) * S + s];
  const T *src = Dp + (size_t)p * 8 * S;
  T *dst = Dc + (size_t)b * 8 * S;
#pragma unroll
  for (int it = 0; it < 8; ++it)
    dst[(size_t)it * S + s] = dst[(size_t)it * S + s] + src[(size_t)it * S + s] * ((it & 1) ? d1 : d0);
}

template <class T>
void launch_rr(int nb, int S, const int *parent, const int *octant, const T *Dp, const T *down,
               T *Dc, hipStream_t st) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_rr<T>, dim3(nblk(S, 256), nb), dim3(256), 0, st, S, parent, octant, Dp,
                     down, Dc);
}

// MLFMM_tri_summation_New (src/lib_sie_tri_mlfmm.f90:416-497) folded per box:
//   y_e = sum_{k,s} fa*W1a + fb*W1b,  y_h = sum fa*W2a + fb*W2b
// W1a = Dbh*kd+Dae*my, W1b = -Dah*kd+Dbe*my, W2a = -Dbe*kd+Dah*eps,
// W2b = Dae*kd+Dbh*eps.  W layout [box][k*4 + {1a,1b,2a,2b}][S].
template <class T>
__global__ void k_fold(int S, const T *D, T kd1, T kd2, T my1, T my2, T ep1, T ep2, T *W) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (s >= S) return;
  const T *Db = D + (size_t)b * 8 * S;
  T *Wb = W + (size_t)b * 8 * S;
  for (int k = 0; k < 2; ++k) {
    const T kd = k ? kd2 : kd1, my = k ? my2 : my1, ep = k ? ep2 : ep1;
    const T Dae = Db[(size_t)(0 + k) * S + s], Dah = Db[(size_t)(2 + k) * S + s];
    const T Dbe = Db[(size_t)(4 + k) * S + s], Dbh = Db[(size_t)(6 + k) * S + s];
    Wb[(size_t)(k * 4 + 0) * S + s] = Dbh * kd + Dae * my;
    Wb[(size_t)(k * 4 + 1) * S + s] = -(Dah * kd) + Dbe * my;
    Wb[(size_t)(k * 4 + 2) * S + s] = -(Dbe * kd) + Dah * ep;
    Wb[(size_t)(k * 4 + 3) * S + s] = Dae * kd + Dbh * ep;
  }
}

template <class T>
void launch_fold(int nbl, int S, const T *D, zd kd1, zd kd2, zd my1, zd my2, zd ep1, zd ep2,
                 T *W, hipStream_t st) {
  if (nbl <= 0) return;
  hipLaunchKernelGGL(k_fold<T>, dim3(nblk(S, 256), nbl), dim3(256), 0, st, S, D, cvt<T>(kd1),
                     cvt<T>(kd2), cvt<T>(my1), cvt<T>(my2), cvt<T>(ep1), cvt<T>(ep2), W);
}

// The transpose of k_fold per box and sample: from the aggregated adjoint
// sums C (layout [comp*4 + cur*2 + k], comp = theta/phi, cur = e/h) to the
// adjoint of D (the same layout):
//   D_ae = my C_ae + kd C_bh,  D_ah = -kd C_be + ep C_ah,
//   D_be = my C_be - kd C_ah,  D_bh = kd C_ae + ep C_bh.
template <class T>
__global__ void k_fold_t(int S, const T *C, T kd1, T kd2, T my1, T my2, T ep1, T ep2, T *D) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (s >= S) return;
  const T *Cb = C + (size_t)b * 8 * S;
  T *Db = D + (size_t)b * 8 * S;
  for (int k = 0; k < 2; ++k) {
    const T kd = k ? kd2 : kd1, my = k ? my2 : my1, ep = k ? ep2 : ep1;
    const T Cae = Cb[(size_t)(0 + k) * S + s], Cah = Cb[(size_t)(2 + k) * S + s];
    const T Cbe = Cb[(size_t)(4 + k) * S + s], Cbh = Cb[(size_t)(6 + k) * S + s];
    Db[(size_t)(0 + k) * S + s] = Cae * my + Cbh * kd;
    Db[(size_t)(2 + k) * S + s] = -(Cbe * kd) + Cah * ep;
    Db[(size_t)(4 + k) * S + s] = Cbe * my - Cah * kd;
    Db[(size_t)(6 + k) * S + s] = Cae * kd + Cbh * ep;
  }
}

template <class T>
void launch_fold_t(int nbl, int S, const T *C, zd kd1, zd kd2, zd my1, zd my2, zd ep1, zd ep2,
                   T *D, hipStream_t st) {
  if (nbl <= 0) return;
  hipLaunchKernelGGL(k_fold_t<T>, dim3(nblk(S, 256), nbl), dim3(256), 0, st, S, C, cvt<T>(kd1),
                     cvt<T>(kd2), cvt<T>(my1), cvt<T>(my2), cvt<T>(ep1), cvt<T>(ep2), D);
}

// out = [sc1 x_e; sc2 x_h]
template <class T>
__global__ void k_scale_halves(int N, T sc1, T sc2, const T *x, T *out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * (int64_t)N) return;
  out[i] = x[i] * (i < N ? sc1 : sc2);
}

template <class T>
void launch_scale_halves(int N, zd sc1, zd sc2, const T *x, T *out, hipStream_t st) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_scale_halves<T>, dim3(nblk(2 * (int64_t)N, 256)), dim3(256), 0, st, N,
                     cvt<T>(sc1), cvt<T>(sc2), x, out);
}

// The transpose of the near blocks: block (m, n) of the output is block
// (n, m) of the input with its 2 x 2 components transposed.  One workgroup
// per (leaf box, row); the column's own row lives in its leaf box, found
// through pos, and m among that box's E2 columns (sorted by edge number)
// by bisection.
template <class T>
__global__ void k_near_transpose(int nbl, const int *box_first, const int *box_count,
                                 const int *e2_off, const int *e2_list, const int64_t *near_off,
                                 const int *perm, const int *pos, const int *leaf_box,
                                 const T *in, T *out) {
  const int b = blockIdx.y;
  const int r = blockIdx.x;
  if (b >= nbl || r >= box_count[b]) return;
  const int m = perm[box_first[b] + r];
  const int c0 = e2_off[b], cols = e2_off[b + 1] - c0;
  for (int c = threadIdx.x; c < cols; c += blockDim.x) {
    const int n = e2_list[c0 + c];
    const int bq = leaf_box[n];
    const int rq = pos[n] - box_first[bq];
    const int q0 = e2_off[bq], qc = e2_off[bq + 1] - q0;
    int lo = 0, hi = qc - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (e2_list[q0 + mid] < m)
        lo = mid + 1;
      else
        hi = mid;
    }
    const T *src = in + 4 * (near_off[bq] + (int64_t)rq * qc + lo);
    T *dst = out + 4 * (near_off[b] + (int64_t)r * cols + c);
    dst[0] = src[0];
    dst[1] = src[2];
    dst[2] = src[1];
    dst[3] = src[3];
  }
}

template <class T>
void launch_near_transpose(int nbl, int max_rows, const int *box_first, const int *box_count,
                           const int *e2_off, const int *e2_list, const int64_t *near_off,
                           const int *perm, const int *pos, const int *leaf_box, const T *in,
                           T *out, hipStream_t st) {
  if (nbl <= 0 || max_rows <= 0) return;
  hipLaunchKernelGGL(k_near_transpose<T>, dim3(max_rows, nbl), dim3(256), 0, st, nbl, box_first,
                     box_count, e2_off, e2_list, near_off, perm, pos, leaf_box, in, out);
}

__device__ inline zd wave_sum(zd v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  return v;
}
__device__ inline zf wave_sum(zf v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  return v;
}

// one wavefront per position: y[m] = sc1*y_e, y[m+N] = sc2*y_h (far part)
// One wavefront per task of up to kDisaggEdges consecutive edges of one leaf
// box: each W sample is loaded once for the group (W is the box's folded
// coefficient array), the patterns of the group's edges stream from HBM.
// task[t] = p0 * 8 + count.  Per edge the sums run in the same order as one
// wavefront per edge would (media, then lane-strided samples).
// CONJ: the conjugated patterns (the transposed operator's aggregation,
// disaggregated)
template <class T, bool CONJ>
__global__ void __launch_bounds__(256) k_disagg(int ntask, int N, int S, const int *task,
                                                const int *perm, const int *leaf_box_of_pos,
                                                const T *rt, const T *W, T sc1, T sc2, T *y) {
  constexpr int E = kDisaggEdges;
  const int lane = threadIdx.x & 63;
  const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (t >= ntask) return;
  const int code = task[t], p0 = code >> 3, cnt = code & 7;
  const int b = leaf_box_of_pos[p0];
  const T *Wb = W + (size_t)b * 8 * S;
  T s1[E], s2[E];
#pragma unroll
  for (int e = 0; e < E; ++e) s1[e] = s2[e] = T{0, 0};
  for (int k = 0; k < 2; ++k) {
    const T *w1a = Wb + (size_t)(k * 4 + 0) * S, *w1b = Wb + (size_t)(k * 4 + 1) * S;
    const T *w2a = Wb + (size_t)(k * 4 + 2) * S, *w2b = Wb + (size_t)(k * 4 + 3) * S;
    // the group's pattern rows (a short group repeats its last edge, whose
    // products are then dropped), so that all loads of a step issue together
    const T *f[E];
#pragma unroll
    for (int e = 0; e < E; ++e) f[e] = rt + ((size_t)(p0 + min(e, cnt - 1)) * 4 + k * 2) * S;
    if constexpr (kPack<T> == 2) {
      // complex64: each lane takes two consecutive samples per 16-byte load
      for (int s = 2 * lane; s < S; s += 128) {
        const zf2 wa = ld2(w1a + s), wb = ld2(w1b + s), wc = ld2(w2a + s), wd = ld2(w2b + s);
        zf2 a[E], bb[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          a[e] = ld2(f[e] + s);
          bb[e] = ld2(f[e] + S + s);
          if (CONJ) {
            a[e] = zf2{conj(a[e].a), conj(a[e].b)};
            bb[e] = zf2{conj(bb[e].a), conj(bb[e].b)};
          }
        }
#pragma unroll
        for (int e = 0; e < E; ++e)
          if (e < cnt) {
            s1[e] = s1[e] + a[e].a * wa.a + bb[e].a * wb.a + a[e].b * wa.b + bb[e].b * wb.b;
            s2[e] = s2[e] + a[e].a * wc.a + bb[e].a * wd.a + a[e].b * wc.b + bb[e].b * wd.b;
          }
      }
    } else {
      for (int s = lane; s < S; s += 64) {
        const T wa = w1a[s], wb = w1b[s], wc = w2a[s], wd = w2b[s];
        T a[E], bb[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
          a[e] = CONJ ? conj(f[e][s]) : f[e][s];
          bb[e] = CONJ ? conj(f[e][S + s]) : f[e][S + s];
        }
#pragma unroll
        for (int e = 0; e < E; ++e)
          if (e < cnt) {
            s1[e] = s1[e] + a[e] * wa + bb[e] * wb;
            s2[e] = s2[e] + a[e] * wc + bb[e] * wd;
          }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const T r1 = wave_sum(s1[e]), r2 = wave_sum(s2[e]);
    if (lane == 0 && e < cnt) {
      const int m = perm[p0 + e];
      y[m] = r1 * sc1;
      y[m + N] = r2 * sc2;
    }
  }
}

template <class T>
void launch_disagg(int ntask, int N, int S, const int *task, const int *perm,
                   const int *leaf_box_of_pos, const T *rt, const T *W, zd sc1, zd sc2, T *y,
                   hipStream_t st, bool conj_rt) {
  if (ntask <= 0) return;
  if (conj_rt)
    hipLaunchKernelGGL((k_disagg<T, true>), dim3(nblk(ntask, 4)), dim3(256), 0, st, ntask, N, S,
                       task, perm, leaf_box_of_pos, rt, W, cvt<T>(sc1), cvt<T>(sc2), y);
  else
    hipLaunchKernelGGL((k_disagg<T, false>), dim3(nblk(ntask, 4)), dim3(256), 0, st, ntask, N, S,
                       task, perm, leaf_box_of_pos, rt, W, cvt<T>(sc1), cvt<T>(sc2), y);
}

// Near SpMV over the box-dense block matrices.  Workgroup = (leaf box,
// chunk of kRowsPerWG rows); x of the box's E2 columns staged in LDS in
// chunks of kColChunk; one wavefront per row, lanes over columns.
constexpr int kRowsPerWG = kSpmvRowsPerWG;
constexpr int kColChunk = 1024;

template <class T>
__global__ void __launch_bounds__(256) k_near_spmv(int N, const int *row_box, const int *row_r0,
                                                   const int *box_first, const int *box_count,
                                                   const int *e2_off, const int *e2_list,
                                                   const int64_t *near_off, const int *perm,
                                                   const T *blocks, const T *x, const T *yadd,
                                                   T *y) {
  __shared__ T xs[2 * kColChunk];
  const int b = row_box[blockIdx.x], r0 = row_r0[blockIdx.x];
  const int rows = min(kRowsPerWG, box_count[b] - r0);
  const int c0 = e2_off[b], cols = e2_off[b + 1] - c0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  T ae[kRowsPerWG / 4], ah[kRowsPerWG / 4];
  for (int q = 0; q < kRowsPerWG / 4; ++q) ae[q] = ah[q] = T{0, 0};
  for (int cc = 0; cc < cols; cc += kColChunk) {
    const int nc = min(kColChunk, cols - cc);
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
      const int n = e2_list[c0 + cc + c];
      xs[2 * c] = x[n];
      xs[2 * c + 1] = x[n + N];
    }
    __syncthreads();
    for (int q = 0; q < kRowsPerWG / 4; ++q) {
      const int r = r0 + wv + 4 * q;
      if (r >= r0 + rows) break;
      const T *Z = blocks + 4 * (near_off[b] + (int64_t)(r - 0) * cols + cc);
      for (int c = lane; c < nc; c += 64) {
        T z0, z1, z2, z3;
        if constexpr (kPack<T> == 2) {
          const zf2 u = ld2(Z + 4 * c), v = ld2(Z + 4 * c + 2);
          z0 = u.a;
          z1 = u.b;
          z2 = v.a;
          z3 = v.b;
        } else {
          z0 = Z[4 * c];
          z1 = Z[4 * c + 1];
          z2 = Z[4 * c + 2];
          z3 = Z[4 * c + 3];
        }
        const T xe = xs[2 * c], xh = xs[2 * c + 1];
        ae[q] = ae[q] + z0 * xe + z1 * xh;
        ah[q] = ah[q] + z2 * xe + z3 * xh;
      }
    }
  }
  for (int q = 0; q < kRowsPerWG / 4; ++q) {
    const int r = r0 + wv + 4 * q;
    const T se = wave_sum(ae[q]), sh = wave_sum(ah[q]);
    if (r < r0 + rows && lane == 0) {
      const int m = perm[box_first[b] + r];
      T oe = se, oh = sh;
      if (yadd) {
        oe = oe + yadd[m];
        oh = oh + yadd[m + N];
      }
      y[m] = oe;
      y[m + N] = oh;
    }
  }
}

template <class T>
void launch_near_spmv_rows(int nwg, int N, const int *row_box, const int *row_r0,
                           const int *box_first, const int *box_count, const int *e2_off,
                           const int *e2_list, const int64_t *near_off, const int *perm,
                           const T *blocks, const T *x, const T *yadd, T *y, hipStream_t st) {
  if (nwg <= 0) return;
  hipLaunchKernelGGL(k_near_spmv<T>, dim3(nwg), dim3(256), 0, st, N, row_box, row_r0, box_first,
                     box_count, e2_off, e2_list, near_off, perm, blocks, x, yadd, y);
}

// ---------------------------------------------------------------------------
// Levels of different truncation numbers.
// ---------------------------------------------------------------------------
// Y[box][pair item][s_out] (+)= sum over the CSR row of val * X[box][item
// of the column's component][s_in]: the interpolation (or its transpose)
// between two sample grids on the four (theta, phi) component pairs of a
// box's 8 coefficient items; pairs[2q + c] = item of component c of pair q.
// One thread per (output row, box, pair).
__global__ void k_resample(int Sin, int Sout, const int *rowptr, const int *col, const double *val,
                           int4 pairs_lo, int4 pairs_hi, const zd *X, zd *Y, int accumulate) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y, q = blockIdx.z;
  if (r >= 2 * Sout) return;
  const int pr[8] = {pairs_lo.x, pairs_lo.y, pairs_lo.z, pairs_lo.w,
                     pairs_hi.x, pairs_hi.y, pairs_hi.z, pairs_hi.w};
  const int it0 = pr[2 * q], it1 = pr[2 * q + 1];
  const zd *x0 = X + ((size_t)b * 8 + it0) * Sin, *x1 = X + ((size_t)b * 8 + it1) * Sin;
  zd acc = {0, 0};
  for (int e = rowptr[r]; e < rowptr[r + 1]; ++e) {
    const int c = col[e];
    const zd v = c < Sin ? x0[c] : x1[c - Sin];
    acc = acc + v * val[e];
  }
  const int co = r < Sout ? 0 : 1, so = r - co * Sout;
  zd *y = Y + ((size_t)b * 8 + (co ? it1 : it0)) * Sout + so;
  *y = accumulate ? *y + acc : acc;
}

void launch_resample(int nbox, int Sin, int Sout, const int *rowptr, const int *col,
                     const double *val, const int *pairs, const zd *X, zd *Y, bool accumulate,
                     hipStream_t st) {
  if (nbox <= 0) return;
  hipLaunchKernelGGL(k_resample, dim3(nblk(2 * (int64_t)Sout, 256), nbox, 4), dim3(256), 0, st, Sin,
                     Sout, rowptr, col, val, make_int4(pairs[0], pairs[1], pairs[2], pairs[3]),
                     make_int4(pairs[4], pairs[5], pairs[6], pairs[7]), X, Y, accumulate ? 1 : 0);
}

// T[b][it][s] = P[parent[b]][it][s] * down[octant[b]][medium][s] on the
// parent's grid; medium = it & 1 (coefficient items) or it >> 2 (folded)
__global__ void k_shift_write(int S, const int *parent, const int *octant, const zd *P,
                              const zd *down, zd *T, int folded) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (s >= S) return;
  const int p = parent[b], o = octant[b];
  const zd d0 = down[((size_t)o * 2 + 0) * S + s], d1 = down[((size_t)o * 2 + 1) * S + s];
  for (int it = 0; it < 8; ++it) {
    const int k = folded ? it >> 2 : it & 1;
    T[((size_t)b * 8 + it) * S + s] = P[((size_t)p * 8 + it) * S + s] * (k ? d1 : d0);
  }
}
void launch_shift_write(int nb, int S, const int *parent, const int *octant, const zd *P,
                        const zd *down, zd *T, bool folded, hipStream_t st) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_shift_write, dim3(nblk(S, 256), nb), dim3(256), 0, st, S, parent, octant, P,
                     down, T, folded ? 1 : 0);
}

// RR on folded coefficients (items k*4 + {1a, 1b, 2a, 2b}, medium it >> 2)
__global__ void k_rr_folded(int S, const int *parent, const int *octant, const zd *Wp,
                            const zd *down, zd *Wc) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (s >= S) return;
  const int p = parent[b], o = octant[b];
  const zd d0 = down[((size_t)o * 2 + 0) * S + s], d1 = down[((size_t)o * 2 + 1) * S + s];
  for (int it = 0; it < 8; ++it) {
    zd &w = Wc[((size_t)b * 8 + it) * S + s];
    w = w + Wp[((size_t)p * 8 + it) * S + s] * ((it >> 2) ? d1 : d0);
  }
}
void launch_rr_folded(int nb, int S, const int *parent, const int *octant, const zd *Wp,
                      const zd *down, zd *Wc, hipStream_t st) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_rr_folded, dim3(nblk(S, 256), nb), dim3(256), 0, st, S, parent, octant, Wp,
                     down, Wc);
}

#define SPK_MATVEC_INSTANTIATE(T)                                                               \
  template void launch_agg<T>(int, int, int, const int *, const int *, const int *, const T *,   \
                              const T *, T *, hipStream_t, bool);                                \
  template void launch_fold_t<T>(int, int, const T *, zd, zd, zd, zd, zd, zd, T *, hipStream_t); \
  template void launch_scale_halves<T>(int, zd, zd, const T *, T *, hipStream_t);                \
  template void launch_near_transpose<T>(int, int, const int *, const int *, const int *,        \
                                         const int *, const int64_t *, const int *, const int *, \
                                         const int *, const T *, T *, hipStream_t);              \
  template void launch_ss<T>(int, int, const int *, const T *, const T *, T *, hipStream_t);     \
  template void launch_m2l<T>(int, int, const int *, const int *, const int *, const T *,        \
                              const T *, T *, hipStream_t);                                      \
  template void launch_rr<T>(int, int, const int *, const int *, const T *, const T *, T *,      \
                             hipStream_t);                                                       \
  template void launch_fold<T>(int, int, const T *, zd, zd, zd, zd, zd, zd, T *, hipStream_t);   \
  template void launch_disagg<T>(int, int, int, const int *, const int *, const int *, const T *, \
                                 const T *, zd, zd, T *, hipStream_t, bool);                     \
  template void launch_near_spmv_rows<T>(int, int, const int *, const int *, const int *,        \
                                         const int *, const int *, const int *, const int64_t *, \
                                         const int *, const T *, const T *, const T *, T *,      \
                                         hipStream_t);
SPK_MATVEC_INSTANTIATE(zd)
SPK_MATVEC_INSTANTIATE(zf)

// Near field in bf16 (the preconditioner's copy): a complex value is one
// 32-bit word, real part in the low half; a 2x2 block is 16 bytes.
__device__ inline uint32_t f32_to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite values)
  return u >> 16;
}
__global__ void k_to_bf(int64_t n, const zf *a, uint32_t *b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = f32_to_bf16(a[i].x) | (f32_to_bf16(a[i].y) << 16);
}
void launch_to_bf(int64_t n, const zf *a, uint32_t *b, hipStream_t st) {
  hipLaunchKernelGGL(k_to_bf, dim3(2048), dim3(256), 0, st, n, a, b);
}
// the near SpMV of k_near_spmv (complex64 x and y) on bf16 blocks
__global__ void __launch_bounds__(256) k_near_spmv_bf(int N, const int *row_box, const int *row_r0,
                                                      const int *box_first, const int *box_count,
                                                      const int *e2_off, const int *e2_list,
                                                      const int64_t *near_off, const int *perm,
                                                      const uint4 *blocks, const zf *x, zf *y) {
  __shared__ zf xs[2 * kColChunk];
  const int b = row_box[blockIdx.x], r0 = row_r0[blockIdx.x];
  const int rows = min(kRowsPerWG, box_count[b] - r0);
  const int c0 = e2_off[b], cols = e2_off[b + 1] - c0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  zf ae[kRowsPerWG / 4], ah[kRowsPerWG / 4];
  for (int q = 0; q < kRowsPerWG / 4; ++q) ae[q] = ah[q] = zf{0, 0};
  auto lo = [](uint32_t u) { return __uint_as_float(u << 16); };
  auto hi = [](uint32_t u) { return __uint_as_float(u & 0xFFFF0000u); };
  for (int cc = 0; cc < cols; cc += kColChunk) {
    const int nc = min(kColChunk, cols - cc);
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
      const int n = e2_list[c0 + cc + c];
      xs[2 * c] = x[n];
      xs[2 * c + 1] = x[n + N];
    }
    __syncthreads();
    for (int q = 0; q < kRowsPerWG / 4; ++q) {
      const int r = r0 + wv + 4 * q;
      if (r >= r0 + rows) break;
      const uint4 *Z = blocks + (near_off[b] + (int64_t)r * cols + cc);
      for (int c = lane; c < nc; c += 64) {
        const uint4 v = Z[c];
        const zf z0 = {lo(v.x), hi(v.x)}, z1 = {lo(v.y), hi(v.y)};
        const zf z2 = {lo(v.z), hi(v.z)}, z3 = {lo(v.w), hi(v.w)};
        const zf xe = xs[2 * c], xh = xs[2 * c + 1];
        ae[q] = ae[q] + z0 * xe + z1 * xh;
        ah[q] = ah[q] + z2 * xe + z3 * xh;
      }
    }
  }
  for (int q = 0; q < kRowsPerWG / 4; ++q) {
    const int r = r0 + wv + 4 * q;
    const zf se = wave_sum(ae[q]), sh = wave_sum(ah[q]);
    if (r < r0 + rows && lane == 0) {
      const int m = perm[box_first[b] + r];
      y[m] = se;
      y[m + N] = sh;
    }
  }
}
void launch_near_spmv_bf(int nwg, int N, const int *row_box, const int *row_r0,
                         const int *box_first, const int *box_count, const int *e2_off,
                         const int *e2_list, const int64_t *near_off, const int *perm,
                         const uint32_t *blocks, const zf *x, zf *y, hipStream_t st) {
  if (nwg <= 0) return;
  hipLaunchKernelGGL(k_near_spmv_bf, dim3(nwg), dim3(256), 0, st, N, row_box, row_r0, box_first,
                     box_count, e2_off, e2_list, near_off, perm, (const uint4 *)blocks, x, y);
}

// Near field in fp8 (OCP E4M3FN, the single-vector preconditioner's copy):
// a 2x2 block is 8 bytes (Z0r, Z0i, Z1r, Z1i, Z2r, Z2i, Z3r, Z3i); every
// scalar row and component has its own scale (the row's largest |re|, |im|
// of that component maps to 448), scale[p] = (s0, s1, s2, s3) for position p;
// f32_to_e4m3 (spk_fp8.h) encodes.
// one workgroup per row position: the row's component maxima, then its blocks encoded
__global__ void __launch_bounds__(256) k_near_to_f8(int N, const int *leaf_box_of_pos,
                                                    const int *box_first, const int *e2_off,
                                                    const int64_t *near_off, const zf *blocks,
                                                    uint2 *out, float4 *scale) {
  __shared__ float red[4][4];
  __shared__ float inv[4];
  const int p = blockIdx.x;
  const int b = leaf_box_of_pos[p], r = p - box_first[b];
  const int cols = e2_off[b + 1] - e2_off[b];
  const int64_t g0 = near_off[b] + (int64_t)r * cols;
  float mx[4] = {0.f, 0.f, 0.f, 0.f};
  for (int c = threadIdx.x; c < cols; c += 256)
    for (int k = 0; k < 4; ++k) {
      const zf z = blocks[4 * (g0 + c) + k];
      mx[k] = fmaxf(mx[k], fmaxf(fabsf(z.x), fabsf(z.y)));
    }
  for (int k = 0; k < 4; ++k) {
    float v = mx[k];
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_down(v, off, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int k = threadIdx.x;
    const float m = fmaxf(fmaxf(red[0][k], red[1][k]), fmaxf(red[2][k], red[3][k]));
    const float sc = m > 0.f ? m / 448.f : 1.f;
    inv[k] = 1.f / sc;
    reinterpret_cast<float *>(scale + p)[k] = sc;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < cols; c += 256) {
    uint32_t w[2] = {0u, 0u};
    for (int k = 0; k < 4; ++k) {
      const zf z = blocks[4 * (g0 + c) + k];
      const uint32_t lo = f32_to_e4m3(z.x * inv[k]), hi = f32_to_e4m3(z.y * inv[k]);
      w[k >> 1] |= (lo | (hi << 8)) << (16 * (k & 1));
    }
    out[g0 + c] = uint2{w[0], w[1]};
  }
}
void launch_near_to_f8(int N, const int *leaf_box_of_pos, const int *box_first, const int *e2_off,
                       const int64_t *near_off, const zf *blocks, uint32_t *out, float *scale,
                       hipStream_t st) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_near_to_f8, dim3(N), dim3(256), 0, st, N, leaf_box_of_pos, box_first, e2_off,
                     near_off, blocks, (uint2 *)out, (float4 *)scale);
}
// the near SpMV of k_near_spmv (complex64 x and y) on the fp8 blocks: the
// four components are summed apart and scaled once per row.  A wave owns rows
// r0 + wv + 4q (q < 4) and walks their columns together, two 64-column groups
// per step, so eight 8-byte loads are in flight per lane and each x value read
// from LDS serves four rows (rows past the box end read row r0 and are not
// written; a column past the chunk end reads the last column with x = 0).
__global__ void __launch_bounds__(256) k_near_spmv_f8(int N, const int *row_box, const int *row_r0,
                                                      const int *box_first, const int *box_count,
                                                      const int *e2_off, const int *e2_list,
                                                      const int64_t *near_off, const int *perm,
                                                      const uint2 *blocks, const float4 *scale,
                                                      const zf *x, zf *y) {
  __shared__ zf xs[2 * kColChunk];
  const int b = row_box[blockIdx.x], r0 = row_r0[blockIdx.x];
  const int rows = min(kRowsPerWG, box_count[b] - r0);
  const int c0 = e2_off[b], cols = e2_off[b + 1] - c0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  constexpr int NQ = kRowsPerWG / 4;
  const uint2 *Zq[NQ];
  for (int q = 0; q < NQ; ++q) {
    const int r = r0 + wv + 4 * q;
    Zq[q] = blocks + (near_off[b] + (int64_t)(r < r0 + rows ? r : r0) * cols);
  }
  zf a0[NQ], a1[NQ], a2[NQ], a3[NQ];
  for (int q = 0; q < NQ; ++q) a0[q] = a1[q] = a2[q] = a3[q] = zf{0, 0};
  auto acc = [&](int q, uint2 v, zf xe, zf xh) {
    const auto q0 = __builtin_amdgcn_cvt_pk_f32_fp8((int)v.x, false);
    const auto q1 = __builtin_amdgcn_cvt_pk_f32_fp8((int)v.x, true);
    const auto q2 = __builtin_amdgcn_cvt_pk_f32_fp8((int)v.y, false);
    const auto q3 = __builtin_amdgcn_cvt_pk_f32_fp8((int)v.y, true);
    a0[q] = a0[q] + zf{q0[0], q0[1]} * xe;
    a1[q] = a1[q] + zf{q1[0], q1[1]} * xh;
    a2[q] = a2[q] + zf{q2[0], q2[1]} * xe;
    a3[q] = a3[q] + zf{q3[0], q3[1]} * xh;
  };
  for (int cc = 0; cc < cols; cc += kColChunk) {
    const int nc = min(kColChunk, cols - cc);
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
      const int n = e2_list[c0 + cc + c];
      xs[2 * c] = x[n];
      xs[2 * c + 1] = x[n + N];
    }
    __syncthreads();
    for (int c = lane; c < nc; c += 128) {
      const int c1 = min(c + 64, nc - 1);
      uint2 v0[NQ], v1[NQ];
      for (int q = 0; q < NQ; ++q) {
        v0[q] = Zq[q][cc + c];
        v1[q] = Zq[q][cc + c1];
      }
      const zf xe0 = xs[2 * c], xh0 = xs[2 * c + 1];
      zf xe1 = xs[2 * c1], xh1 = xs[2 * c1 + 1];
      if (c + 64 >= nc) xe1 = xh1 = zf{0, 0};
      for (int q = 0; q < NQ; ++q) {
        acc(q, v0[q], xe0, xh0);
        acc(q, v1[q], xe1, xh1);
      }
    }
  }
  for (int q = 0; q < NQ; ++q) {
    const int r = r0 + wv + 4 * q;
    const zf s0 = wave_sum(a0[q]), s1 = wave_sum(a1[q]), s2 = wave_sum(a2[q]), s3 = wave_sum(a3[q]);
    if (r < r0 + rows && lane == 0) {
      const int p = box_first[b] + r;
      const float4 sc = scale[p];
      const int m = perm[p];
      y[m] = s0 * sc.x + s1 * sc.y;
      y[m + N] = s2 * sc.z + s3 * sc.w;
    }
  }
}
void launch_near_spmv_f8(int nwg, int N, const int *row_box, const int *row_r0,
                         const int *box_first, const int *box_count, const int *e2_off,
                         const int *e2_list, const int64_t *near_off, const int *perm,
                         const uint32_t *blocks, const float *scale, const zf *x, zf *y,
                         hipStream_t st) {
  if (nwg <= 0) return;
  hipLaunchKernelGGL(k_near_spmv_f8, dim3(nwg), dim3(256), 0, st, N, row_box, row_r0, box_first,
                     box_count, e2_off, e2_list, near_off, perm, (const uint2 *)blocks,
                     (const float4 *)scale, x, y);
}

// complex128 -> complex64 copy (the mixed-precision operator's tables)
__global__ void k_to_zf(int64_t n, const zd *a, zf *b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = zf{(float)a[i].x, (float)a[i].y};
}
void launch_to_zf(int64_t n, const zd *a, zf *b, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_to_zf, dim3(gcap(n)), dim3(256), 0, st, n, a, b);
}

// ---------------------------------------------------------------------------
// Jacobi preconditioner (precalculation_diagonal_element and
// lib_sie_ml_fmm_preconditioner_left, src/lib_sie_tri_mlfmm_interface.f90:94-109,
// 321-339): diag = [Z_mm(1,1); Z_mm(2,2)], y = x / diag.  One thread per
// box-major position finds the self column in its box's sorted E2 list.
// ---------------------------------------------------------------------------
__global__ void k_jacobi_diag(int N, const int *perm, const int *leaf_box_of_pos,
                              const int *box_first, const int *e2_off, const int *e2_list,
                              const int64_t *near_off, const zd *blocks, zd *diag) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const int m = perm[p], b = leaf_box_of_pos[p];
  const int r = p - box_first[b];
  const int c0 = e2_off[b], cols = e2_off[b + 1] - c0;
  int lo = 0, hi = cols - 1, c = -1;
  while (lo <= hi) {
    const int md = (lo + hi) >> 1, v = e2_list[c0 + md];
    if (v == m) {
      c = md;
      break;
    }
    if (v < m) lo = md + 1;
    else hi = md - 1;
  }
  zd d11 = {0, 0}, d22 = {0, 0};
  if (c >= 0) {
    const zd *Z = blocks + 4 * (near_off[b] + (int64_t)r * cols + c);
    d11 = Z[0];
    d22 = Z[3];
  }
  diag[m] = d11;
  diag[m + N] = d22;
}
void launch_jacobi_diag(int N, const int *perm, const int *leaf_box_of_pos, const int *box_first,
                        const int *e2_off, const int *e2_list, const int64_t *near_off,
                        const zd *blocks, zd *diag, hipStream_t st) {
  if (N <= 0) return;
  hipLaunchKernelGGL(k_jacobi_diag, dim3(nblk(N, 256)), dim3(256), 0, st, N, perm,
                     leaf_box_of_pos, box_first, e2_off, e2_list, near_off, blocks, diag);
}
__global__ void k_div_diag(int64_t n, const zd *diag, const zd *x, zd *y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i] / diag[i];
}
void launch_div_diag(int64_t n, const zd *diag, const zd *x, zd *y, hipStream_t st) {
  hipLaunchKernelGGL(k_div_diag, dim3(gcap(n)), dim3(256), 0, st, n, diag, x, y);
}

// ---------------------------------------------------------------------------
// TL tables of one level (TL_km_ml_fmm, src/lib_sie_mlfmm.f90:129-188, on the
// 7^3 offset grid of precalculation_LT3_TL, :27-82), pre-multiplied by
// TRF: tlt[slot][med][s] = TRF_s sum_q c_q P_q(rhat.khat_s) with the slot
// factors c_q from the host.  One thread per (slot, medium, sample).
// ---------------------------------------------------------------------------
__global__ void k_tl_level(int S, int K, const double *khat, const double *trf,
                           const double *rhat, const zd *coef, const int *live, zd *tlt) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int sm = blockIdx.y;  // slot * 2 + medium
  if (s >= S) return;
  zd *o = tlt + (size_t)sm * S + s;
  if (!live[sm]) {
    *o = zd{0, 0};
    return;
  }
  const int ts = sm >> 1;
  const double x = rhat[3 * ts] * khat[3 * s] + rhat[3 * ts + 1] * khat[3 * s + 1] +
                   rhat[3 * ts + 2] * khat[3 * s + 2];
  const zd *c = coef + (size_t)sm * K;
  double p0 = 1.0, p1 = x;
  zd acc = c[0] * p0;
  if (K > 1) acc = acc + c[1] * p1;
  for (int q = 2; q < K; ++q) {
    const double p2 = (2.0 * q - 1.0) / q * x * p1 - (q - 1.0) / q * p0;
    acc = acc + c[q] * p2;
    p0 = p1;
    p1 = p2;
  }
  *o = acc * trf[s];
}
void launch_tl_level(int S, int K, const double *khat, const double *trf, const double *rhat,
                     const zd *coef, const int *live, zd *tlt, hipStream_t st) {
  hipLaunchKernelGGL(k_tl_level, dim3(nblk(S, 256), 343 * 2), dim3(256), 0, st, S, K, khat, trf,
                     rhat, coef, live, tlt);
}

// ---------------------------------------------------------------------------
// Plane-wave right-hand side (get_vector_b_tri / incident_field_tri_general,
// src/lib_sie_tri_calculation_mod.f90:391-460, 939-1002): at the ng rule
// points of both triangles of edge m, E = e_in exp(-i k1 k_in.r),
// H = k_in x E / (my1 c0), tested with f_m(r) = df/2 (r - p_c).
// ---------------------------------------------------------------------------
__global__ void k_rhs_plane(RhsArgs a, const FnRec *fn, zd *b) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= a.N) return;
  zd S4[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (int ot = 0; ot < 2; ++ot) {
    const FnRec &f = fn[2 * m + ot];
    for (int i = 0; i < a.ng; ++i) {
      double Ra[3], rfm[3];
      const double w0 = 1 - a.qa[i] - a.qb[i];
      for (int d = 0; d < 3; ++d) {
        Ra[d] = w0 * f.c[d] + a.qa[i] * f.c[3 + d] + a.qb[i] * f.c[6 + d];
        rfm[d] = (Ra[d] - f.c[6 + d]) * f.df * 0.5;
      }
      const zd ph = (a.k1 * a.kin[0]) * Ra[0] + (a.k1 * a.kin[1]) * Ra[1] +
                    (a.k1 * a.kin[2]) * Ra[2];
      // exp(-i ph) = exp(Im ph) (cos Re ph - i sin Re ph)
      double sn, cs;
      sincos(ph.x, &sn, &cs);
      const double mag = exp(ph.y);
      const zd e = {mag * cs, -mag * sn};
      const zd E[3] = {e * a.ein[0], e * a.ein[1], e * a.ein[2]};
      const zd H[3] = {(E[2] * a.kin[1] - E[1] * a.kin[2]) * (1.0 / a.c0my),
                       (E[0] * a.kin[2] - E[2] * a.kin[0]) * (1.0 / a.c0my),
                       (E[1] * a.kin[0] - E[0] * a.kin[1]) * (1.0 / a.c0my)};
      const double *nf = f.nqf;
      const zd nE[3] = {E[2] * nf[1] - E[1] * nf[2], E[0] * nf[2] - E[2] * nf[0],
                        E[1] * nf[0] - E[0] * nf[1]};
      const zd nH[3] = {H[2] * nf[1] - H[1] * nf[2], H[0] * nf[2] - H[2] * nf[0],
                        H[1] * nf[0] - H[0] * nf[1]};
      const double wa = a.qw[i] * f.area;
      S4[0] = S4[0] + (E[0] * rfm[0] + E[1] * rfm[1] + E[2] * rfm[2]) * wa;
      S4[1] = S4[1] + (H[0] * rfm[0] + H[1] * rfm[1] + H[2] * rfm[2]) * wa;
      S4[2] = S4[2] + (nE[0] * rfm[0] + nE[1] * rfm[1] + nE[2] * rfm[2]) * wa;
      S4[3] = S4[3] + (nH[0] * rfm[0] + nH[1] * rfm[1] + nH[2] * rfm[2]) * wa;
    }
  }
  zd be = {0, 0}, bh = {0, 0};
  for (int q = 0; q < 4; ++q) {
    be = be + a.ca[q] * S4[q];
    bh = bh + a.cb[q] * S4[q];
  }
  b[m] = be;
  b[m + a.N] = bh;
}
void launch_rhs_plane(const RhsArgs &a, const FnRec *fn, zd *b, hipStream_t st) {
  if (a.N <= 0) return;
  hipLaunchKernelGGL(k_rhs_plane, dim3(nblk(a.N, 128)), dim3(128), 0, st, a, fn, b);
}

// ---------------------------------------------------------------------------
// Scattered field at observation points: Integration_scattered_field_tri
// (src/lib_sie_tri_calculation_mod.f90:2782-2825) with its 3-point rule.
// One workgroup per point, threads over RWG edges, fixed-order block
// reduction of the six field components.  The reference multiplies the
// charge term of E by 0 (:2811); it is left out.
// ---------------------------------------------------------------------------
__device__ inline zd block_sum(zd v);

__global__ void __launch_bounds__(256) k_scattered(ScatArgs a, const FnRec *fn, const zd *x,
                                                   const double *r, zd *Eo, zd *Ho) {
  const int p = blockIdx.x;
  const double ra[3] = {r[3 * p], r[3 * p + 1], r[3 * p + 2]};
  zd SE[3] = {{0, 0}, {0, 0}, {0, 0}}, SH[3] = {{0, 0}, {0, 0}, {0, 0}};
  for (int m = threadIdx.x; m < a.N; m += blockDim.x) {
    const zd J = x[m], M = x[m + a.N];
    for (int t = 0; t < 2; ++t) {
      const FnRec &f = fn[2 * m + t];
      const double dfm = f.sign * f.length;
      for (int i = 0; i < 3; ++i) {
        const double w0 = 1 - a.qa[i] - a.qb[i];
        double fm[3], d[3];
        for (int k = 0; k < 3; ++k) {
          const double rq = w0 * f.c[k] + a.qa[i] * f.c[3 + k] + a.qb[i] * f.c[6 + k];
          fm[k] = 0.5 * dfm * (rq - f.c[6 + k]);
          d[k] = ra[k] - rq;
        }
        const double R = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), rinv = 1.0 / R;
        zd G, DG;
        green(a.k1, R, rinv, G, DG);
        const zd gG[3] = {DG * (d[0] * rinv), DG * (d[1] * rinv), DG * (d[2] * rinv)};
        const zd Mf[3] = {M * fm[0], M * fm[1], M * fm[2]};
        const zd cE[3] = {Mf[1] * gG[2] - Mf[2] * gG[1], Mf[2] * gG[0] - Mf[0] * gG[2],
                          Mf[0] * gG[1] - Mf[1] * gG[0]};
        const zd fxg[3] = {gG[2] * fm[1] - gG[1] * fm[2], gG[0] * fm[2] - gG[2] * fm[0],
                           gG[1] * fm[0] - gG[0] * fm[1]};
        const zd cJ = a.iwmy * J, cM = a.iweps * M, cQ = a.inv_iwmy * M * dfm;
        for (int k = 0; k < 3; ++k) {
          SE[k] = SE[k] - (cJ * fm[k] * G - cE[k]) * a.qw[i];
          SH[k] = SH[k] - (cM * fm[k] * G - cQ * gG[k] + J * fxg[k]) * a.qw[i];
        }
      }
    }
  }
  for (int k = 0; k < 3; ++k) {
    const zd e = block_sum(SE[k]);
    const zd h = block_sum(SH[k]);
    if (threadIdx.x == 0) {
      Eo[3 * p + k] = e;
      Ho[3 * p + k] = h;
    }
  }
}

// ---------------------------------------------------------------------------
// BLAS-1 with deterministic two-stage reductions (fixed grid of kRedBlocks)
// ---------------------------------------------------------------------------
constexpr int kRedBlocks = 256;
int reduce_blocks(int64_t) { return kRedBlocks; }

__device__ inline zd block_sum(zd v) {
  __shared__ zd red[4];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  zd t = {0, 0};
  if (threadIdx.x == 0) t = red[0] + red[1] + red[2] + red[3];
  __syncthreads();
  return t;
}

template <class T>
__global__ void k_dot_partial(int64_t n, const T *x, const T *y, zd *partial) {
  zd acc = {0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    acc = acc + conj(widen(x[i])) * widen(y[i]);
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}
__global__ void k_reduce_final(int nparts, const zd *partial, zd *out) {
  zd acc = {0, 0};
  for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc = acc + partial[i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) *out = acc;
}
template <class T>
void launch_zdotc(int64_t n, const T *x, const T *y, zd *partial, zd *out, hipStream_t st) {
  hipLaunchKernelGGL(k_dot_partial<T>, dim3(kRedBlocks), dim3(256), 0, st, n, x, y, partial);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, kRedBlocks, partial, out);
}
template void launch_zdotc<zd>(int64_t, const zd *, const zd *, zd *, zd *, hipStream_t);
template void launch_zdotc<zf>(int64_t, const zf *, const zf *, zd *, zd *, hipStream_t);

// x += a * d (d complex64, x complex128)
__global__ void k_axpy_up(int64_t n, zd a, const zf *d, zd *x) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    x[i] = x[i] + a * widen(d[i]);
}
void launch_axpy_up(int64_t n, zd a, const zf *d, zd *x, hipStream_t st) {
  hipLaunchKernelGGL(k_axpy_up, dim3(gcap(n)), dim3(256), 0, st, n, a, d, x);
}
// y = s * x rounded to complex64
__global__ void k_scale_down(int64_t n, double s, const zd *x, zf *y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = zf{(float)(x[i].x * s), (float)(x[i].y * s)};
}
void launch_scale_down(int64_t n, double s, const zd *x, zf *y, hipStream_t st) {
  hipLaunchKernelGGL(k_scale_down, dim3(gcap(n)), dim3(256), 0, st, n, s, x, y);
}
// y = s * x in complex64 (s computed in double)
__global__ void k_scal_copy_f(int64_t n, double s, const zf *x, zf *y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = zf{(float)(x[i].x * s), (float)(x[i].y * s)};
}
void launch_scal_copy_f(int64_t n, double s, const zf *x, zf *y, hipStream_t st) {
  hipLaunchKernelGGL(k_scal_copy_f, dim3(gcap(n)), dim3(256), 0, st, n, s, x, y);
}

__global__ void k_zaxpy(int64_t n, zd a, const zd *x, zd *y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = y[i] + a * x[i];
}
void launch_zaxpy(int64_t n, zd a, const zd *x, zd *y, hipStream_t st) {
  hipLaunchKernelGGL(k_zaxpy, dim3(gcap(n)), dim3(256), 0, st, n, a, x, y);
}
__global__ void k_zaxpy_dev(int64_t n, const zd *a, double sgn, const zd *x, zd *y) {
  const zd c = (*a) * sgn;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = y[i] + c * x[i];
}
void launch_zaxpy_dev(int64_t n, const zd *a, double sgn, const zd *x, zd *y, hipStream_t st) {
  hipLaunchKernelGGL(k_zaxpy_dev, dim3(gcap(n)), dim3(256), 0, st, n, a,
                     sgn, x, y);
}
__global__ void k_zscal_copy(int64_t n, zd a, const zd *x, zd *y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i];
}
void launch_zscal_copy(int64_t n, zd a, const zd *x, zd *y, hipStream_t st) {
  hipLaunchKernelGGL(k_zscal_copy, dim3(gcap(n)), dim3(256), 0, st, n, a,
                     x, y);
}
__global__ void k_zsub(int64_t n, const zd *b, const zd *x, zd *y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = b[i] - x[i];
}
void launch_zsub(int64_t n, const zd *b, const zd *x, zd *y, hipStream_t st) {
  hipLaunchKernelGGL(k_zsub, dim3(gcap(n)), dim3(256), 0, st, n, b, x, y);
}
__global__ void k_zadd(int64_t n, const zd *a, zd *y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    y[i] = y[i] + a[i];
}
void launch_zadd(int64_t n, const zd *a, zd *y, hipStream_t st) {
  hipLaunchKernelGGL(k_zadd, dim3(gcap(n)), dim3(256), 0, st, n, a, y);
}

// MGS step j: w -= h_{j-1} v_{j-1} (if vprev), then partial <v_j, w>.
__global__ void k_mgs_partial(int64_t n, const zd *vprev, const zd *hprev, const zd *vj, zd *w,
                              zd *partial) {
  const zd c = vprev ? -(*hprev) : zd{0, 0};
  zd acc = {0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    zd wi = w[i];
    if (vprev) {
      wi = wi + c * vprev[i];
      w[i] = wi;
    }
    acc = acc + conj(vj[i]) * wi;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}
void launch_mgs_step(int64_t n, const zd *vprev, const zd *hprev, const zd *vj, zd *w,
                     zd *partial, zd *hout, hipStream_t st) {
  hipLaunchKernelGGL(k_mgs_partial, dim3(kRedBlocks), dim3(256), 0, st, n, vprev, hprev, vj, w,
                     partial);
  hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, kRedBlocks, partial, hout);
}

// ---------------------------------------------------------------------------
// One modified Gram-Schmidt pass in a single cooperative launch.
//
// For j = 0..k-1: h_j = <v_j, w>; w -= h_j v_j (sequential, as MGS), then
// |w|^2.  Each block owns a contiguous slice of w, held in registers for the
// whole pass; the next basis vector's slice is loaded before each exchange.
//
// Exchange without fences: block b publishes its step-j partial in its own
// slot (j, b) with agent-scope atomic stores of (re, im), waits for them to
// complete (s_waitcnt vmcnt(0)) and then stores the launch's sequence number
// into the slot.  Lane q of every block's first wave polls slots q, q+64, ...
// with agent-scope atomic loads until they carry this launch's sequence
// number, then reads the values; every block sums them in the same order, so
// all blocks hold the same h_j.  Slots from earlier launches carry smaller
// sequence numbers.  A poll gives up after kPollLimit tries and raises
// ctl[0]; every later poll then stops at once, all waves exit, nrm->y reports
// the flag and the host discards the pass.
// ---------------------------------------------------------------------------
constexpr long kPollLimit = 1L << 22;

struct alignas(32) SweepSlot {
  double re, im;
  unsigned long long seq, pad;
};

__device__ inline void slot_publish(SweepSlot *s, zd v, unsigned long long seq) {
  __hip_atomic_store(&s->re, v.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(&s->im, v.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __hip_atomic_store(&s->seq, seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// lane-parallel sum of the G slots of one step (first wave only)
__device__ inline zd slots_sum(SweepSlot *slots, int G, unsigned long long seq,
                               unsigned long long *abort_flag) {
  const int lane = threadIdx.x & 63;
  zd sp = {0, 0};
  for (int q = lane; q < G; q += 64) {
    SweepSlot *s = slots + q;
    long polls = 0;
    while (__hip_atomic_load(&s->seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != seq) {
      if ((++polls & 255) == 0) {
        if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
        if (polls > kPollLimit) {
          atomicExch(abort_flag, 1ull);
          break;
        }
      }
      __builtin_amdgcn_s_sleep(1);
    }
    asm volatile("" ::: "memory");
    sp.x += __hip_atomic_load(&s->re, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    sp.y += __hip_atomic_load(&s->im, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return wave_sum(sp);
}

template <int E>
__global__ void __launch_bounds__(256) k_mgs_sweep(int64_t n, int k, const zd *V, int64_t ldv,
                                                   zd *w, SweepSlot *slots, zd *hout, zd *nrm,
                                                   unsigned long long *ctl,
                                                   unsigned long long seq) {
  const int G = gridDim.x, tid = threadIdx.x, lane = tid & 63;
  const int64_t chunk = (n + G - 1) / G;
  const int64_t i0 = (int64_t)blockIdx.x * chunk;
  const int64_t i1 = min(n, i0 + chunk);
  __shared__ zd hsh;
  zd wr[E], vr[E], vn[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int64_t i = i0 + tid + e * 256;
    wr[e] = i < i1 ? w[i] : zd{0, 0};
    vr[e] = (k > 0 && i < i1) ? V[i] : zd{0, 0};
    vn[e] = zd{0, 0};
  }
  for (int j = 0; j < k; ++j) {
    zd acc = {0, 0};
#pragma unroll
    for (int e = 0; e < E; ++e) acc = acc + conj(vr[e]) * wr[e];
    acc = block_sum(acc);
    if (tid == 0) slot_publish(slots + (size_t)j * G + blockIdx.x, acc, seq);
    if (j + 1 < k) {
      const zd *vnext = V + (size_t)(j + 1) * ldv;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int64_t i = i0 + tid + e * 256;
        vn[e] = i < i1 ? vnext[i] : zd{0, 0};
      }
    }
    if (tid < 64) {
      const zd sp = slots_sum(slots + (size_t)j * G, G, seq, ctl);
      if (lane == 0) hsh = sp;
    }
    __syncthreads();
    const zd h = hsh;
    if (blockIdx.x == 0 && tid == 0) hout[j] = h;
    const zd c = -h;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      wr[e] = wr[e] + c * vr[e];
      vr[e] = vn[e];
    }
    __syncthreads();
  }
  zd acc = {0, 0};
#pragma unroll
  for (int e = 0; e < E; ++e) {
    acc = acc + conj(wr[e]) * wr[e];
    const int64_t i = i0 + tid + e * 256;
    if (i < i1) w[i] = wr[e];
  }
  acc = block_sum(acc);
  if (tid == 0) slot_publish(slots + (size_t)k * G + blockIdx.x, acc, seq);
  if (blockIdx.x == 0 && tid < 64) {
    const zd sp = slots_sum(slots + (size_t)k * G, G, seq, ctl);
    if (lane == 0) {
      const unsigned long long ab =
          __hip_atomic_load(ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      *nrm = zd{sp.x, ab ? 1.0 : 0.0};
    }
  }
}

size_t mgs_sweep_slot_bytes() { return sizeof(SweepSlot); }

int mgs_sweep_blocks(int64_t n) {
  // 64 blocks while a slice fits 16 elements per thread, then more blocks
  int64_t g = 64;
  while (g < 256 && (n + g - 1) / g > 256 * 16) g *= 2;
  if ((n + g - 1) / g > 256 * 16) return 0;  // too large for register slices
  return (int)g;
}

template <int E>
static hipError_t launch_sweep_e(int G, int64_t n, int k, const zd *V, int64_t ldv, zd *w,
                                 void *slots, zd *hout, zd *nrm, unsigned long long *ctl,
                                 unsigned long long seq, hipStream_t st) {
  SweepSlot *sl = (SweepSlot *)slots;
  void *args[] = {&n, &k, (void *)&V, &ldv, &w, &sl, &hout, &nrm, &ctl, &seq};
  return hipLaunchCooperativeKernel((const void *)k_mgs_sweep<E>, dim3(G), dim3(256), args, 0, st);
}

hipError_t launch_mgs_sweep(int G, int64_t n, int k, const zd *V, int64_t ldv, zd *w, void *slots,
                            zd *hout, zd *nrm, unsigned long long *ctl, unsigned long long seq,
                            hipStream_t st) {
  const int64_t chunk = (n + G - 1) / G;
  const int64_t e = (chunk + 255) / 256;
  if (e <= 1) return launch_sweep_e<1>(G, n, k, V, ldv, w, slots, hout, nrm, ctl, seq, st);
  if (e <= 2) return launch_sweep_e<2>(G, n, k, V, ldv, w, slots, hout, nrm, ctl, seq, st);
  if (e <= 4) return launch_sweep_e<4>(G, n, k, V, ldv, w, slots, hout, nrm, ctl, seq, st);
  if (e <= 8) return launch_sweep_e<8>(G, n, k, V, ldv, w, slots, hout, nrm, ctl, seq, st);
  if (e <= 16) return launch_sweep_e<16>(G, n, k, V, ldv, w, slots, hout, nrm, ctl, seq, st);
  return hipErrorInvalidValue;
}

// Tall-skinny products with the Krylov basis V (n rows, column j at
// V + j*ldv), dot products accumulated in complex128 whatever the storage
// type T.  Both kernels split the work over a 2-D grid so that a launch has
// about 2048 workgroups whatever k is, and write partial sums that a second
// kernel adds in a fixed order (results do not depend on scheduling).
//
// out[v*k + j] = <V_j, w_v> (v < NV): grid (row chunks R) x (column tiles of
// CT); every thread keeps NV*CT accumulators over its rows of the chunk;
// partial[(v*k + j) * R + chunk].
constexpr int kGemvWorkgroups = 2048;
constexpr int kGemvMaxRowChunks = 256;
static inline int gemv_c_rows(int64_t n, int ctiles) {
  // at least 4 rows per thread, so that the per-tile reduction is amortised
  const int64_t by_rows = (n + 1023) / 1024;
  const int64_t by_fill = (kGemvWorkgroups + ctiles - 1) / ctiles;
  return (int)std::max<int64_t>(1, std::min<int64_t>(std::min(by_rows, by_fill), kGemvMaxRowChunks));
}
size_t gemv_c_partial_elems(int64_t n, int kmax, int nv) {
  (void)n;
  return (size_t)nv * (size_t)kmax * kGemvMaxRowChunks;
}
template <class T, int NV, int CT>
__global__ void __launch_bounds__(256) k_gemv_c_tile(int64_t n, int k, const T *V, int64_t ldv,
                                                     const T *w0, const T *w1, int64_t rows,
                                                     zd *partial) {
  __shared__ zd red[4][NV * CT];
  const int R = gridDim.x, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int j0 = blockIdx.y * CT, nt = min(CT, k - j0);
  const int64_t i0 = (int64_t)blockIdx.x * rows, i1 = min<int64_t>(n, i0 + rows);
  zd acc[NV][CT];
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int t = 0; t < CT; ++t) acc[v][t] = zd{0, 0};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    zd x[NV];
    x[0] = widen(w0[i]);
    if (NV > 1) x[NV - 1] = widen(w1[i]);
#pragma unroll
    for (int t = 0; t < CT; ++t)
      if (t < nt) {
        const zd c = conj(widen(V[(size_t)(j0 + t) * ldv + i]));
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v][t] = acc[v][t] + c * x[v];
      }
  }
#pragma unroll
  for (int v = 0; v < NV; ++v)
#pragma unroll
    for (int t = 0; t < CT; ++t) {
      const zd r = wave_sum(acc[v][t]);
      if (lane == 0) red[wv][v * CT + t] = r;
    }
  __syncthreads();
  if (threadIdx.x < NV * CT) {
    const int v = threadIdx.x / CT, t = threadIdx.x % CT;
    if (t < nt)
      partial[((size_t)v * k + j0 + t) * R + blockIdx.x] =
          red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
  }
}
// out[o] = sum_r partial[o * R + r], one wave per output, fixed order
__global__ void k_gemv_c_final(int nout, int R, const zd *partial, zd *out) {
  const int o = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (o >= nout) return;
  zd acc = {0, 0};
  for (int r = lane; r < R; r += 64) acc = acc + partial[(size_t)o * R + r];
  acc = wave_sum(acc);
  if (lane == 0) out[o] = acc;
}
template <class T>
void launch_gemv_c(int64_t n, int k, const T *V, int64_t ldv, const T *w0, const T *w1,
                   zd *partial, zd *out, hipStream_t st) {
  if (k <= 0) return;
  const int nv = w1 ? 2 : 1;
  const int CT = w1 ? 8 : 16;
  const int ctiles = (k + CT - 1) / CT;
  const int R = gemv_c_rows(n, ctiles);
  const int64_t rows = (n + R - 1) / R;
  if (w1)
    hipLaunchKernelGGL((k_gemv_c_tile<T, 2, 8>), dim3(R, ctiles), dim3(256), 0, st, n, k, V, ldv,
                       w0, w1, rows, partial);
  else
    hipLaunchKernelGGL((k_gemv_c_tile<T, 1, 16>), dim3(R, ctiles), dim3(256), 0, st, n, k, V, ldv,
                       w0, w0, rows, partial);
  hipLaunchKernelGGL(k_gemv_c_final, dim3((nv * k + 3) / 4), dim3(256), 0, st, nv * k, R, partial,
                     out);
}

// V h: grid (row tiles of 256) x (column chunks C); each thread sums its
// row over the chunk's columns into partial[chunk * n + i]; the final
// kernel adds the C partials in order and writes out = V h or w -= V h.
// A single chunk skips the partials.
static inline int gemv_n_chunks(int64_t n, int k) {
  const int64_t tiles = (n + 255) / 256;
  const int64_t by_fill = std::max<int64_t>(1, kGemvWorkgroups / tiles);
  const int64_t by_cols = std::max(1, (k + 31) / 32);
  return (int)std::min(by_fill, by_cols);
}
size_t gemv_n_partial_elems(int64_t n, int kmax) {
  return (size_t)gemv_n_chunks(n, kmax) * (size_t)n;
}
template <class T, int MODE>  // MODE 0: partial, 1: out = V h, 2: w -= V h
__global__ void __launch_bounds__(256) k_gemv_n_tile(int64_t n, int k, int cols, const T *V,
                                                     int64_t ldv, const zd *h, zd *partial,
                                                     zd *out, T *w) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int j0 = blockIdx.y * cols, j1 = min(k, j0 + cols);
  zd acc = {0, 0};
  int j = j0;
  for (; j + 8 <= j1; j += 8) {
    zd v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = widen(V[(size_t)(j + t) * ldv + i]);
#pragma unroll
    for (int t = 0; t < 8; ++t) acc = acc + v[t] * h[j + t];
  }
  for (; j < j1; ++j) acc = acc + widen(V[(size_t)j * ldv + i]) * h[j];
  if (MODE == 0) partial[(size_t)blockIdx.y * n + i] = acc;
  else if (MODE == 1) out[i] = acc;
  else w[i] = cvt<T>(widen(w[i]) - acc);
}
template <class T, int MODE>
__global__ void k_gemv_n_final(int64_t n, int C, const zd *partial, zd *out, T *w) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    zd acc = partial[i];
    for (int c = 1; c < C; ++c) acc = acc + partial[(size_t)c * n + i];
    if (MODE == 1) out[i] = acc;
    else w[i] = cvt<T>(widen(w[i]) - acc);
  }
}
template <class T, int MODE>
static void launch_gemv_n_mode(int64_t n, int k, const T *V, int64_t ldv, const zd *h, zd *out,
                               T *w, zd *partial, hipStream_t st) {
  const int C = gemv_n_chunks(n, k);
  const int cols = (k + C - 1) / C;
  const unsigned tiles = (unsigned)((n + 255) / 256);
  if (C == 1) {
    hipLaunchKernelGGL((k_gemv_n_tile<T, MODE>), dim3(tiles, 1), dim3(256), 0, st, n, k, cols, V,
                       ldv, h, partial, out, w);
    return;
  }
  hipLaunchKernelGGL((k_gemv_n_tile<T, 0>), dim3(tiles, C), dim3(256), 0, st, n, k, cols, V, ldv,
                     h, partial, out, w);
  hipLaunchKernelGGL((k_gemv_n_final<T, MODE>), dim3(gcap(n)), dim3(256), 0, st, n, C, partial,
                     out, w);
}
template <class T>
void launch_gemv_n_sub(int64_t n, int k, const T *V, int64_t ldv, const zd *h, T *w, zd *partial,
                       hipStream_t st) {
  if (k <= 0) return;
  launch_gemv_n_mode<T, 2>(n, k, V, ldv, h, nullptr, w, partial, st);
}
template <class T>
void launch_gemv_n(int64_t n, int k, const T *V, int64_t ldv, const zd *h, zd *out, zd *partial,
                   hipStream_t st) {
  // k = 0 writes zeros (the tile kernel sums an empty range)
  launch_gemv_n_mode<T, 1>(n, k, V, ldv, h, out, (T *)nullptr, partial, st);
}
#define SPK_GEMV_INSTANTIATE(T)                                                                  \
  template void launch_gemv_c<T>(int64_t, int, const T *, int64_t, const T *, const T *, zd *,    \
                                 zd *, hipStream_t);                                              \
  template void launch_gemv_n_sub<T>(int64_t, int, const T *, int64_t, const zd *, T *, zd *,     \
                                     hipStream_t);                                                \
  template void launch_gemv_n<T>(int64_t, int, const T *, int64_t, const zd *, zd *, zd *,        \
                                 hipStream_t);
SPK_GEMV_INSTANTIATE(zd)
SPK_GEMV_INSTANTIATE(zf)

// Streaming read of n 16-byte words (the HBM read ceiling that the matvec
// kernels are compared against); the sum is written only if it equals a
// value no input produces, so the loads cannot be dropped.
__global__ void __launch_bounds__(256) k_stream_read(const double2 *a, size_t n, double *out) {
  double s = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double2 v = a[i];
    s += v.x + v.y;
  }
  if (s == -1.2345e300) out[0] = s;
}
void launch_stream_read(const void *a, size_t bytes, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_stream_read, dim3(8192), dim3(256), 0, st, (const double2 *)a, bytes / 16,
                     out);
}

// One-reduce MGS (the inverse compact WY form of modified Gram-Schmidt):
// with L the strictly lower part of V^H V, the MGS coefficients of one pass
// over k basis vectors are h = (I + L)^-1 c, c = V^H w.  Tm holds
// T = (I + L)^-1 column-major (column j at Tm + j*ldt, rows j..), grown by
// one row per Arnoldi step: row k1 = -conj(T^H g) with g_i = <v_i, v_k1>.
__global__ void k_icwy_row(int k1, zd *Tm, int ldt, const zd *tg) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < k1) Tm[(size_t)j * ldt + k1] = -conj(tg[j]);
  if (j == k1) Tm[(size_t)k1 * ldt + k1] = zd{1.0, 0.0};
}
void launch_icwy_row(int k1, zd *Tm, int ldt, const zd *tg, hipStream_t st) {
  hipLaunchKernelGGL(k_icwy_row, dim3(nblk(k1 + 1, 256)), dim3(256), 0, st, k1, Tm, ldt, tg);
}

void launch_scattered(const ScatArgs &a, const FnRec *fn, const zd *x, const double *r, zd *E,
                      zd *H, hipStream_t st) {
  if (a.npts <= 0) return;
  hipLaunchKernelGGL(k_scattered, dim3(a.npts), dim3(256), 0, st, a, fn, x, r, E, H);
}


// ---------------------------------------------------------------------------
// Adjoint gradient of a far-field intensity with respect to node heights.
// Every geometry-dependent local quantity (an RWG side's record, a near
// block, a radiation pattern, a right-hand-side entry, a side's field at a
// detector point) is evaluated with one node moved by +dz and by -dz along
// z, and the central difference is contracted at once with the adjoint
// vector, the currents or the leaf fields.  Near blocks keep the singular /
// regular branch choice of the unmoved geometry.
// ---------------------------------------------------------------------------

// The record of RWG side `side` (0, 1) from node coordinates, node pn moved
// by dz along z; sn = (p1, p2, free vertex, element nodes in tt order),
// 0-based; the arithmetic of build_rwg (fn_parameter_type, norm_vec_r_new).
__device__ FnRec fn_rebuild(const double *pp, const int *sn, int side, int normal_convention,
                            int pn, double dz) {
  FnRec f;
  auto P = [&](int v, int d) -> double { return pp[3 * v + d] + ((v == pn && d == 2) ? dz : 0.0); };
  for (int d = 0; d < 3; ++d) {
    f.c[d] = P(sn[0], d);
    f.c[3 + d] = P(sn[1], d);
    f.c[6 + d] = P(sn[2], d);
  }
  const double dv[3] = {f.c[0] - f.c[3], f.c[1] - f.c[4], f.c[2] - f.c[5]};
  f.length = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
  const double a1[3] = {f.c[0] - f.c[6], f.c[1] - f.c[7], f.c[2] - f.c[8]};
  const double a2[3] = {f.c[3] - f.c[6], f.c[4] - f.c[7], f.c[5] - f.c[8]};
  double x[3];
  cross3(a1, a2, x);
  f.area = 0.5 * sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  const double sg = side == 0 ? 1.0 : -1.0;
  f.sign = sg;
  f.df = sg * f.length / f.area;
  for (int d = 0; d < 3; ++d) f.mid[d] = (P(sn[3], d) + P(sn[4], d) + P(sn[5], d)) / 3.0;
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
  double nx[3];
  cross3(v1, v2, nx);
  const double nl = len3(nx);
  for (int d = 0; d < 3; ++d) {
    f.nq[d] = nx[d] / nl;
    f.nqf[d] = (double)(float)f.nq[d];
  }
  f.pad[0] = f.pad[1] = 0.0;
  return f;
}

__device__ inline bool in_list(const int *v, int n, int x) {
  for (int i = 0; i < n; ++i)
    if (v[i] == x) return true;
  return false;
}

// The pair (m side t, n side s) of a near block from the given side
// records, with the branch choice (singular test, self element, edge
// adjacency) of the unmoved geometry; zz as assemble_zz gives it.
__device__ void pair_frozen(const NearArgs &A, int m, int n, int t, int s, const FnRec &fa,
                            const FnRec &fb, zd *zz) {
  const double rr = 1.5 * A.rref[m];
  const FnRec &bm = A.fn[2 * m + t], &bq = A.fn[2 * n + s];
  const double d[3] = {bm.mid[0] - bq.mid[0], bm.mid[1] - bq.mid[1], bm.mid[2] - bq.mid[2]};
  const ElemRec ea = A.er[2 * m + t], eb = A.er[2 * n + s];
  Rel rv[2];
  if (!(len3(d) > rr)) {
    const bool iqs = ea.elem == eb.elem;
    singular_pair(fa, fb, iqs, A, rv);
    if (A.edge_swap && !iqs && edge_adj(ea, eb)) {
      zd kt[2];
      edge_adjacent_kt(fa, fb, A, kt);
      rv[0].Kt = kt[0];
      rv[1].Kt = kt[1];
    }
  } else {
    normal_pair<true>(fa, fb, A, rv);
  }
  assemble_zz(A, rv, zz);
}

// Z_mn(+dz) - Z_mn(-dz).  The block is linear in its four triangle pairs'
// zz (accumulate_D, finish_D), so the difference is accumulated from the
// pairs' differences; a pair with neither triangle holding the moved node
// (moved_m[t], moved_n[s] false) contributes an exact zero and is skipped.
__device__ void block_frozen_diff(const NearArgs &A, int m, int n, const FnRec *fmp,
                                  const FnRec *fmm, const FnRec *fnp, const FnRec *fnm,
                                  const bool *moved_m, const bool *moved_n, zd *D) {
  D[0] = D[1] = D[2] = D[3] = zd{0, 0};
  for (int p = 0; p < 4; ++p) {
    const int t = p >> 1, s = p & 1;
    zd dz[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    if (moved_m[t] || moved_n[s]) {
      zd zp[4], zm[4];
      pair_frozen(A, m, n, t, s, fmp[t], fnp[s], zp);
      pair_frozen(A, m, n, t, s, fmm[t], fnm[s], zm);
      for (int k = 0; k < 4; ++k) dz[k] = zp[k] - zm[k];
    }
    accumulate_D(A, D, dz);
  }
  finish_D(A, D);
}

__device__ inline bool side_holds(const int *sn, int pn) {
  return sn[3] == pn || sn[4] == pn || sn[5] == pn;
}

// sum over the requested nodes' near tasks of nu_m^T (Z_mn(+dz) - Z_mn(-dz)) I_n:
// for slot a (affected edge e_a) of node k, tasks j < cols: (e_a, E2 column j);
// cols <= j < 2 cols: (E2 column j - cols, e_a) unless that row is affected
// too (then it is counted among its own slot's rows)
__global__ void __launch_bounds__(256) k_adj_near(NearArgs A, AdjArgs G, const zd *nu, const zd *I,
                                                  zd *partial) {
  const int k = blockIdx.y;
  const int pn = G.node[k];
  // node k's slots: aff_off[k] .. aff_off[k + 1] - 2, then its end entry
  const int s0 = G.aff_off[k], ns = G.aff_off[k + 1] - s0 - 1;
  const int *aff = G.aff_edge + s0;
  const int64_t *toff = G.task_off + s0;
  const int64_t ntask = toff[ns];
  zd acc = {0, 0};
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < ntask;
       t += (int64_t)gridDim.x * blockDim.x) {
    int a = 0;
    while (a + 1 < ns && toff[a + 1] <= t) ++a;
    const int ea = aff[a];
    const int b = G.leaf_box[ea];
    const int c0 = A.e2_off[b], cols = A.e2_off[b + 1] - c0;
    const int64_t loc = t - toff[a];
    const int other = A.e2_list[c0 + (int)(loc % cols)];
    int m, n;
    if (loc < cols) {
      m = ea;
      n = other;
    } else {
      if (in_list(aff, ns, other)) continue;
      m = other;
      n = ea;
    }
    const bool am = in_list(aff, ns, m), an = in_list(aff, ns, n);
    FnRec fmp[2], fmm[2], fnp[2], fnm[2];
    bool mvm[2], mvn[2];
    for (int q = 0; q < 2; ++q) {
      const int *snm = G.side_nodes + 6 * (2 * m + q), *snn = G.side_nodes + 6 * (2 * n + q);
      mvm[q] = am && side_holds(snm, pn);
      mvn[q] = an && side_holds(snn, pn);
      if (mvm[q]) {
        fmp[q] = fn_rebuild(G.pp, snm, q, G.normal_convention, pn, G.dz);
        fmm[q] = fn_rebuild(G.pp, snm, q, G.normal_convention, pn, -G.dz);
      } else {
        fmp[q] = fmm[q] = A.fn[2 * m + q];
      }
      if (mvn[q]) {
        fnp[q] = fn_rebuild(G.pp, snn, q, G.normal_convention, pn, G.dz);
        fnm[q] = fn_rebuild(G.pp, snn, q, G.normal_convention, pn, -G.dz);
      } else {
        fnp[q] = fnm[q] = A.fn[2 * n + q];
      }
    }
    zd d[4];
    block_frozen_diff(A, m, n, fmp, fmm, fnp, fnm, mvm, mvn, d);
    const zd ie = I[n], ih = I[n + G.N];
    acc = acc + nu[m] * (d[0] * ie + d[1] * ih) + nu[m + G.N] * (d[2] * ie + d[3] * ih);
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partial[(size_t)k * gridDim.x + blockIdx.x] = acc;
}

// the four pattern components of edge m (sides fs[2]) at sample s: theta
// and phi projections for media 1 and 2 (k_patterns' integrand)
__device__ void pattern_sample(const FnRec *fs, const double *Rc, const double *kh,
                               const double *tt3, const double *tp3, int ngp, const double *qa,
                               const double *qb, const double *qw, zd k1, zd k2, zd *o) {
  o[0] = o[1] = o[2] = o[3] = zd{0, 0};
  for (int ll = 0; ll < 2; ++ll) {
    const FnRec &f = fs[ll];
    const double dfm = f.sign * f.length;
    zd sa[3] = {{0, 0}, {0, 0}, {0, 0}}, sb[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int j = 0; j < ngp; ++j) {
      double Ro[3], fm[3], Roc[3];
      simplex(qa[j], qb[j], f.c, Ro);
      for (int kk = 0; kk < 3; ++kk) {
        fm[kk] = 0.5 * (Ro[kk] - f.c[6 + kk]) * dfm;
        Roc[kk] = Ro[kk] - Rc[kk];
      }
      const double tmp = dot3(kh, Roc);
      const zd ff1 = expmikr(k1, tmp) * qw[j];
      const zd ff2 = expmikr(k2, tmp) * qw[j];
      for (int kk = 0; kk < 3; ++kk) {
        sa[kk] = sa[kk] + fm[kk] * ff1;
        sb[kk] = sb[kk] + fm[kk] * ff2;
      }
    }
    o[0] = o[0] + dot_rc(tt3, sa);
    o[1] = o[1] + dot_rc(tp3, sa);
    o[2] = o[2] + dot_rc(tt3, sb);
    o[3] = o[3] + dot_rc(tp3, sb);
  }
}

// per (node, affected edge m): the far field's share, through the change of
// m's pattern: as a receiver nu_m against the leaf field W of I (the folded
// coefficients), as a source I_m against the transposed pipeline's leaf sums
// Bt of nu
__global__ void __launch_bounds__(256) k_adj_far(AdjArgs G, AdjFar F, const zd *nu, const zd *I,
                                                 const zd *W, const zd *Bt, zd *out) {
  const int g = blockIdx.x;
  const int pn = G.node[G.aff_node[g]];
  const int m = G.aff_edge[g];
  const int b = G.leaf_box[m];
  const int S = F.S;
  FnRec fp[2], fq[2];
  for (int q = 0; q < 2; ++q) {
    fp[q] = fn_rebuild(G.pp, G.side_nodes + 6 * (2 * m + q), q, G.normal_convention, pn, G.dz);
    fq[q] = fn_rebuild(G.pp, G.side_nodes + 6 * (2 * m + q), q, G.normal_convention, pn, -G.dz);
  }
  const double *Rc = F.leaf_centre + 3 * b;
  const zd *Wb = W + (size_t)b * 8 * S, *Bb = Bt + (size_t)b * 8 * S;
  const zd ne = nu[m] * F.sc1, nh = nu[m + G.N] * F.sc2, ie = I[m], ih = I[m + G.N];
  zd acc = {0, 0};
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    zd op[4], om[4];
    pattern_sample(fp, Rc, F.khat + 3 * s, F.tth + 3 * s, F.tph + 3 * s, F.ngp, F.qa, F.qb, F.qw,
                   F.k1, F.k2, op);
    pattern_sample(fq, Rc, F.khat + 3 * s, F.tth + 3 * s, F.tph + 3 * s, F.ngp, F.qa, F.qb, F.qw,
                   F.k1, F.k2, om);
    for (int kk = 0; kk < 2; ++kk) {
      const zd da = op[kk * 2 + 0] - om[kk * 2 + 0], db = op[kk * 2 + 1] - om[kk * 2 + 1];
      acc = acc + ne * (da * Wb[(size_t)(kk * 4 + 0) * S + s] + db * Wb[(size_t)(kk * 4 + 1) * S + s]);
      acc = acc + nh * (da * Wb[(size_t)(kk * 4 + 2) * S + s] + db * Wb[(size_t)(kk * 4 + 3) * S + s]);
      // source side: C[comp*4 + cur*2 + k] = conj(rt[k*2 + comp]) x_cur
      acc = acc + ie * (conj(da) * Bb[(size_t)(0 * 4 + 0 * 2 + kk) * S + s] +
                        conj(db) * Bb[(size_t)(1 * 4 + 0 * 2 + kk) * S + s]);
      acc = acc + ih * (conj(da) * Bb[(size_t)(0 * 4 + 1 * 2 + kk) * S + s] +
                        conj(db) * Bb[(size_t)(1 * 4 + 1 * 2 + kk) * S + s]);
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) out[g] = acc;
}

// b_e, b_h of edge m from its side records (k_rhs_plane's integrand)
__device__ void rhs_edge(const RhsArgs &a, const FnRec *fs, zd &be, zd &bh) {
  zd S4[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
  for (int ot = 0; ot < 2; ++ot) {
    const FnRec &f = fs[ot];
    for (int i = 0; i < a.ng; ++i) {
      double Ra[3], rfm[3];
      const double w0 = 1 - a.qa[i] - a.qb[i];
      for (int d = 0; d < 3; ++d) {
        Ra[d] = w0 * f.c[d] + a.qa[i] * f.c[3 + d] + a.qb[i] * f.c[6 + d];
        rfm[d] = (Ra[d] - f.c[6 + d]) * f.df * 0.5;
      }
      const zd ph = (a.k1 * a.kin[0]) * Ra[0] + (a.k1 * a.kin[1]) * Ra[1] + (a.k1 * a.kin[2]) * Ra[2];
      double sn, cs;
      sincos(ph.x, &sn, &cs);
      const double mag = exp(ph.y);
      const zd e = {mag * cs, -mag * sn};
      const zd E[3] = {e * a.ein[0], e * a.ein[1], e * a.ein[2]};
      const zd H[3] = {(E[2] * a.kin[1] - E[1] * a.kin[2]) * (1.0 / a.c0my),
                       (E[0] * a.kin[2] - E[2] * a.kin[0]) * (1.0 / a.c0my),
                       (E[1] * a.kin[0] - E[0] * a.kin[1]) * (1.0 / a.c0my)};
      const double *nf = f.nqf;
      const zd nE[3] = {E[2] * nf[1] - E[1] * nf[2], E[0] * nf[2] - E[2] * nf[0],
                        E[1] * nf[0] - E[0] * nf[1]};
      const zd nH[3] = {H[2] * nf[1] - H[1] * nf[2], H[0] * nf[2] - H[2] * nf[0],
                        H[1] * nf[0] - H[0] * nf[1]};
      const double wa = a.qw[i] * f.area;
      S4[0] = S4[0] + (E[0] * rfm[0] + E[1] * rfm[1] + E[2] * rfm[2]) * wa;
      S4[1] = S4[1] + (H[0] * rfm[0] + H[1] * rfm[1] + H[2] * rfm[2]) * wa;
      S4[2] = S4[2] + (nE[0] * rfm[0] + nE[1] * rfm[1] + nE[2] * rfm[2]) * wa;
      S4[3] = S4[3] + (nH[0] * rfm[0] + nH[1] * rfm[1] + nH[2] * rfm[2]) * wa;
    }
  }
  be = bh = zd{0, 0};
  for (int q = 0; q < 4; ++q) {
    be = be + a.ca[q] * S4[q];
    bh = bh + a.cb[q] * S4[q];
  }
}

// per (node, affected edge): nu_m^T (b_m(+dz) - b_m(-dz))
__global__ void k_adj_rhs(AdjArgs G, RhsArgs a, const zd *nu, zd *out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= G.nslot) return;
  const int pn = G.node[G.aff_node[g]];
  const int m = G.aff_edge[g];
  FnRec fp[2], fq[2];
  for (int q = 0; q < 2; ++q) {
    fp[q] = fn_rebuild(G.pp, G.side_nodes + 6 * (2 * m + q), q, G.normal_convention, pn, G.dz);
    fq[q] = fn_rebuild(G.pp, G.side_nodes + 6 * (2 * m + q), q, G.normal_convention, pn, -G.dz);
  }
  zd bpe, bph, bme, bmh;
  rhs_edge(a, fp, bpe, bph);
  rhs_edge(a, fq, bme, bmh);
  out[g] = nu[m] * (bpe - bme) + nu[m + G.N] * (bph - bmh);
}

// E at ra of unit J (Aj) and unit M (Am) on one side (k_scattered's
// integrand; the charge term of E is dropped as there)
__device__ void scat_side(const ScatArgs &a, const FnRec &f, const double *ra, zd *Aj, zd *Am) {
  const double dfm = f.sign * f.length;
  for (int k = 0; k < 3; ++k) Aj[k] = Am[k] = zd{0, 0};
  for (int i = 0; i < 3; ++i) {
    const double w0 = 1 - a.qa[i] - a.qb[i];
    double fm[3], d[3];
    for (int k = 0; k < 3; ++k) {
      const double rq = w0 * f.c[k] + a.qa[i] * f.c[3 + k] + a.qb[i] * f.c[6 + k];
      fm[k] = 0.5 * dfm * (rq - f.c[6 + k]);
      d[k] = ra[k] - rq;
    }
    const double R = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]), rinv = 1.0 / R;
    zd Gf, DG;
    green(a.k1, R, rinv, Gf, DG);
    const zd gG[3] = {DG * (d[0] * rinv), DG * (d[1] * rinv), DG * (d[2] * rinv)};
    const zd cx[3] = {fm[1] * gG[2] - fm[2] * gG[1], fm[2] * gG[0] - fm[0] * gG[2],
                      fm[0] * gG[1] - fm[1] * gG[0]};
    for (int k = 0; k < 3; ++k) {
      Aj[k] = Aj[k] - (a.iwmy * fm[k] * Gf) * a.qw[i];
      Am[k] = Am[k] + cx[k] * a.qw[i];
    }
  }
}

// c = F^T (w conj(E)): c_m = sum_p w_p A_pm . conj(E_p), one workgroup per edge
__global__ void __launch_bounds__(256) k_scattered_t(ScatArgs a, const FnRec *fn, const double *r,
                                                     const double *w, const zd *E, zd *c) {
  const int m = blockIdx.x;
  zd ce = {0, 0}, ch = {0, 0};
  for (int p = threadIdx.x; p < a.npts; p += blockDim.x) {
    const double ra[3] = {r[3 * p], r[3 * p + 1], r[3 * p + 2]};
    const zd Ep[3] = {conj(E[3 * p]) * w[p], conj(E[3 * p + 1]) * w[p], conj(E[3 * p + 2]) * w[p]};
    for (int t = 0; t < 2; ++t) {
      zd Aj[3], Am[3];
      scat_side(a, fn[2 * m + t], ra, Aj, Am);
      ce = ce + Aj[0] * Ep[0] + Aj[1] * Ep[1] + Aj[2] * Ep[2];
      ch = ch + Am[0] * Ep[0] + Am[1] * Ep[1] + Am[2] * Ep[2];
    }
  }
  ce = block_sum(ce);
  ch = block_sum(ch);
  if (threadIdx.x == 0) {
    c[m] = ce;
    c[m + a.N] = ch;
  }
}

// per (node, affected edge m): sum_p w_p conj(E_p) . (E_p,m(+dz) - E_p,m(-dz)),
// E_p,m the field of m's currents I_m at point p
__global__ void __launch_bounds__(256) k_adj_field(AdjArgs G, ScatArgs a, const double *r,
                                                   const double *w, const zd *E, const zd *I,
                                                   zd *out) {
  const int g = blockIdx.x;
  const int pn = G.node[G.aff_node[g]];
  const int m = G.aff_edge[g];
  FnRec fp[2], fq[2];
  for (int q = 0; q < 2; ++q) {
    fp[q] = fn_rebuild(G.pp, G.side_nodes + 6 * (2 * m + q), q, G.normal_convention, pn, G.dz);
    fq[q] = fn_rebuild(G.pp, G.side_nodes + 6 * (2 * m + q), q, G.normal_convention, pn, -G.dz);
  }
  const zd ie = I[m], ih = I[m + G.N];
  zd acc = {0, 0};
  for (int p = threadIdx.x; p < a.npts; p += blockDim.x) {
    const double ra[3] = {r[3 * p], r[3 * p + 1], r[3 * p + 2]};
    for (int t = 0; t < 2; ++t) {
      zd Ajp[3], Amp[3], Ajm[3], Amm[3];
      scat_side(a, fp[t], ra, Ajp, Amp);
      scat_side(a, fq[t], ra, Ajm, Amm);
      for (int k = 0; k < 3; ++k)
        acc = acc + conj(E[3 * p + k]) * w[p] * ((Ajp[k] - Ajm[k]) * ie + (Amp[k] - Amm[k]) * ih);
    }
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) out[g] = acc;
}

void launch_adj_near(const NearArgs &A, const AdjArgs &G, int groups, const zd *nu, const zd *I,
                     zd *partial, hipStream_t st) {
  if (G.nnode <= 0) return;
  hipLaunchKernelGGL(k_adj_near, dim3(groups, G.nnode), dim3(256), 0, st, A, G, nu, I, partial);
}
void launch_adj_far(const AdjArgs &G, const AdjFar &F, const zd *nu, const zd *I, const zd *W,
                    const zd *Bt, zd *out, hipStream_t st) {
  if (G.nslot <= 0) return;
  hipLaunchKernelGGL(k_adj_far, dim3(G.nslot), dim3(256), 0, st, G, F, nu, I, W, Bt, out);
}
void launch_adj_rhs(const AdjArgs &G, const RhsArgs &a, const zd *nu, zd *out, hipStream_t st) {
  if (G.nslot <= 0) return;
  hipLaunchKernelGGL(k_adj_rhs, dim3(nblk(G.nslot, 64)), dim3(64), 0, st, G, a, nu, out);
}
void launch_scattered_t(const ScatArgs &a, const FnRec *fn, const double *r, const double *w,
                        const zd *E, zd *c, hipStream_t st) {
  if (a.N <= 0) return;
  hipLaunchKernelGGL(k_scattered_t, dim3(a.N), dim3(256), 0, st, a, fn, r, w, E, c);
}
void launch_adj_field(const AdjArgs &G, const ScatArgs &a, const double *r, const double *w,
                      const zd *E, const zd *I, zd *out, hipStream_t st) {
  if (G.nslot <= 0 || a.npts <= 0) return;
  hipLaunchKernelGGL(k_adj_field, dim3(G.nslot), dim3(256), 0, st, G, a, r, w, E, I, out);
}

}  // namespace spk

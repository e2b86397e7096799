/*
 * spk_oracle.c -- CPU restatement of the SpeckleSim SIE solve loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the thing measured or
 * shipped).  See spk_oracle.h.  Each routine names the reference file:line
 * it restates; paths are relative to the reference repository root.
 *
 * Parity notes reproduced on purpose:
 *  - default-real literals (quadrature tables, c0, eps_0, my_0, 1.0/3.0 in
 *    Iq_s/K2q, 5.0/6.0 in Iq_L) are rounded to float before widening;
 *  - CMPLX(x) without KIND yields default (single) complex: cross_rc rounds
 *    its real operand to float (src/lib_sie_math.f90:561-571) and the
 *    0.5*(-1)**md*cmplx(x_nt,0.0) identity terms round x_nt/x_nn to float
 *    (src/lib_sie_tri_calculation_mod.f90:2195,2251-2252,2300-2301).
 */
#include "spk_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex zc;

#define ZCMPLX(a, b) CMPLX((a), (b))

/* src/lib/libmath/src/math/lib_math_constants.f90:24  PI = 4*atan(1d0) */
#define SPO_PI 3.14159265358979323846
/* src/lib_sie_constants.f90:9-11 -- single-precision literals */
#define C0_F ((double)2.99792458e+8f)
#define EPS0_F ((double)8.854187817e-12f)
#define MY0_F ((double)1.2566370614e-6f)

static inline double fr(double x) { return (double)(float)x; }

/* ------------------------------------------------------------------------ */
/* small vector helpers: src/lib_sie_math.f90:550-597                         */
/* ------------------------------------------------------------------------ */
static inline void vsub(const double *a, const double *b, double *o) {
  o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2];
}
static inline double vdot(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double vlen(const double *v) {
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}
static inline void vcross(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
/* cross_rc: cmplx(v1(i)) is default complex -> v1 rounded to float. */
static inline void vcross_rc(const double *a, const zc *b, zc *o) {
  double f0 = fr(a[0]), f1 = fr(a[1]), f2 = fr(a[2]);
  o[0] = f1 * b[2] - f2 * b[1];
  o[1] = f2 * b[0] - f0 * b[2];
  o[2] = f0 * b[1] - f1 * b[0];
}
static inline zc zdot_rc(const double *a, const zc *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* ------------------------------------------------------------------------ */
/* Triangle rules: src/lib_sie_integration_points.f90:7-143                   */
/* literals are default real -> rounded to float                             */
/* ------------------------------------------------------------------------ */
void spo_quadrature_tri(int ngp, double *al, double *be, double *w) {
  int m;
  for (m = 0; m < 20; ++m) { al[m] = 0.0; be[m] = 0.0; w[m] = 0.0; }
  if (ngp == 1) {
    al[0] = fr(0.33333333333333); be[0] = fr(0.33333333333333); w[0] = 1.0;
  } else if (ngp == 3) {
    const float a[3] = {0.16666666666667f, 0.16666666666667f, 0.66666666666667f};
    const float b[3] = {0.16666666666667f, 0.66666666666667f, 0.16666666666667f};
    for (m = 0; m < 3; ++m) { al[m] = a[m]; be[m] = b[m]; w[m] = (double)0.33333333333333f; }
  } else if (ngp == 4) {
    const float a[4] = {0.33333333333333f, 0.20000000000000f, 0.20000000000000f, 0.60000000000000f};
    const float b[4] = {0.33333333333333f, 0.20000000000000f, 0.60000000000000f, 0.20000000000000f};
    const float ww[4] = {-0.56250000000000f, 0.52083333333333f, 0.52083333333333f, 0.52083333333333f};
    for (m = 0; m < 4; ++m) { al[m] = a[m]; be[m] = b[m]; w[m] = ww[m]; }
  } else if (ngp == 6) {
    const float a[6] = {0.44594849091597f, 0.44594849091597f, 0.10810301816807f,
                        0.09157621350977f, 0.09157621350977f, 0.81684757298046f};
    const float b[6] = {0.44594849091597f, 0.10810301816807f, 0.44594849091597f,
                        0.09157621350977f, 0.81684757298046f, 0.09157621350977f};
    const float ww[6] = {0.22338158967801f, 0.22338158967801f, 0.22338158967801f,
                         0.10995174365532f, 0.10995174365532f, 0.10995174365532f};
    for (m = 0; m < 6; ++m) { al[m] = a[m]; be[m] = b[m]; w[m] = ww[m]; }
  } else if (ngp == 7) {
    const float a[7] = {0.33333333333333f, 0.47014206410511f, 0.47014206410511f,
                        0.05971587178977f, 0.10128650732346f, 0.10128650732346f,
                        0.79742698535309f};
    const float b[7] = {0.33333333333333f, 0.47014206410511f, 0.05971587178977f,
                        0.47014206410511f, 0.10128650732346f, 0.79742698535309f,
                        0.10128650732346f};
    const float ww[7] = {0.22500000000000f, 0.13239415278851f, 0.13239415278851f,
                         0.13239415278851f, 0.12593918054483f, 0.12593918054483f,
                         0.12593918054483f};
    for (m = 0; m < 7; ++m) { al[m] = a[m]; be[m] = b[m]; w[m] = ww[m]; }
  } else if (ngp == 12) {
    const float a[12] = {0.249286745170910f, 0.249286745170910f, 0.501426509658180f,
                         0.063089014491500f, 0.063089014491500f, 0.873821971017000f,
                         0.310352451033780f, 0.636502499121400f, 0.053145049844820f,
                         0.636502499121400f, 0.310352451033780f, 0.053145049844820f};
    const float b[12] = {0.249286745170910f, 0.501426509658180f, 0.249286745170910f,
                         0.063089014491500f, 0.873821971017000f, 0.063089014491500f,
                         0.636502499121400f, 0.053145049844820f, 0.310352451033780f,
                         0.310352451033780f, 0.053145049844820f, 0.636502499121400f};
    const float ww[12] = {0.116786275726380f, 0.116786275726380f, 0.116786275726380f,
                          0.0508449063702100f, 0.0508449063702100f, 0.0508449063702100f,
                          0.0828510756183700f, 0.0828510756183700f, 0.0828510756183700f,
                          0.0828510756183700f, 0.0828510756183700f, 0.0828510756183700f};
    for (m = 0; m < 12; ++m) { al[m] = a[m]; be[m] = b[m]; w[m] = ww[m]; }
  }
}

/* ------------------------------------------------------------------------ */
/* Context                                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
  double c[3][3];  /* p1, p2, pc */
  double area, length, df, mid[3], nq[3];
  int sign;
} fnrwg;

typedef struct {
  int nbox;
  int64_t *n;      /* sorted box indices */
  int *first, *count; /* leaf level only: element ranges into box_elems */
} level_t;

struct spo_ctx {
  spo_params p;
  /* materials, set_local_material_parameters_tri
   * (src/lib_sie_tri_data_container.f90:212-224) */
  double k0, omega, my1, my2;
  zc k1, k2, eps1, eps2, eta1, eta2, eta_a;
  /* mesh */
  int np, nt;
  double *pp;   /* 3*np */
  int *tt;      /* 3*nt, 1-based */
  double *cent; /* 3*nt element centroids */
  /* RWG */
  int N;
  int *pcorner, *pelem; /* 2N each, 1-based */
  fnrwg *fn;            /* 2N: fn[2m+ot-1] */
  double *rref;         /* N */
  /* tree */
  int tree_built, l_min, l_max;
  double D, xmin[3];
  int64_t *leaf;        /* N */
  level_t *lev;         /* [0..l_max] */
  int *box_elems;       /* N, grouped per leaf box */
  int *leaf_pos;        /* N: position of edge's leaf box in lev[l_max] */
  /* near CSR */
  int64_t nnz;
  int32_t *ia, *ja;
  zc *za;               /* 4*nnz: D11 D12 D21 D22 */
  /* MLFMM */
  int S;                /* 2K^2+2 */
  double *khat, *ttheta, *tphi, *trf; /* S*3, S*3, S*3, S */
  zc *rt;               /* N * 4 * S : [m][k*2+comp][i] (comp 0 = a/theta) */
  zc **tl;              /* per level: 343 * 2 * S_l */
  int mlfmm_ready;
  /* per-level truncation numbers (the reference's truncation_number(l),
   * lib_sie_mlfmm.f90:631-701; all equal to p.K unless spo_set_level_K):
   * requested values, and per level l_min..l_max: K, samples, grid */
  int Kreq[32], nKreq;
  int rr_adjoint;   /* RR to a smaller K: 0 the reference's interpolation of
                       the local coefficients, 1 the transpose of SS's
                       interpolation (anterpolation) */
  int Kl[32], Sl[32];
  double *khl[32], *trfl[32];
};

static void *xmalloc(size_t n) {
  void *p = calloc(1, n ? n : 1);
  if (!p) { fprintf(stderr, "spk_oracle: out of memory\n"); abort(); }
  return p;
}

/* find_neighbours + common_2: src/lib_sie_tri_calculation_mod.f90:865-928.
 * Pairs i<j of elements with two common corners, in (i, j) order; the two
 * corners are listed in element i's corner order. */
static void build_pairs(spo_ctx *c) {
  int nt = c->nt, np = c->np, i, k;
  int *deg = xmalloc(sizeof(int) * (np + 1));
  for (i = 0; i < nt; ++i)
    for (k = 0; k < 3; ++k) deg[c->tt[3 * i + k] - 1 + 1]++;
  for (i = 0; i < np; ++i) deg[i + 1] += deg[i];
  int *adj = xmalloc(sizeof(int) * 3 * nt);
  int *fill = xmalloc(sizeof(int) * np);
  for (i = 0; i < nt; ++i)
    for (k = 0; k < 3; ++k) {
      int v = c->tt[3 * i + k] - 1;
      adj[deg[v] + fill[v]++] = i;
    }
  int cap = 3 * nt / 2 + 16, n = 0;
  int *pc = xmalloc(sizeof(int) * 2 * cap), *pe = xmalloc(sizeof(int) * 2 * cap);
  int *cand = xmalloc(sizeof(int) * 64);
  for (i = 0; i < nt; ++i) {
    int nc = 0, a, b2;
    for (k = 0; k < 3; ++k) {
      int v = c->tt[3 * i + k] - 1;
      for (a = deg[v]; a < deg[v + 1]; ++a) {
        int j = adj[a], dup = 0;
        if (j <= i) continue;
        for (b2 = 0; b2 < nc; ++b2) if (cand[b2] == j) { dup = 1; break; }
        if (!dup && nc < 64) cand[nc++] = j;
      }
    }
    /* ascending j */
    for (a = 1; a < nc; ++a) {
      int t = cand[a], q = a - 1;
      while (q >= 0 && cand[q] > t) { cand[q + 1] = cand[q]; --q; }
      cand[q + 1] = t;
    }
    for (a = 0; a < nc; ++a) {
      int j = cand[a], ne = 0, com[3] = {-1, -1, -1}, ii, jj;
      for (ii = 0; ii < 3; ++ii)
        for (jj = 0; jj < 3; ++jj)
          if (c->tt[3 * i + ii] == c->tt[3 * j + jj] && ne < 3) com[ne++] = c->tt[3 * i + ii];
      if (ne >= 2) {
        if (n >= cap) {
          cap *= 2;
          pc = realloc(pc, sizeof(int) * 2 * cap);
          pe = realloc(pe, sizeof(int) * 2 * cap);
        }
        pc[2 * n] = com[0]; pc[2 * n + 1] = com[1];
        pe[2 * n] = i + 1; pe[2 * n + 1] = j + 1;
        ++n;
      }
    }
  }
  c->N = n; c->pcorner = pc; c->pelem = pe;
  free(deg); free(adj); free(fill); free(cand);
}

/* norm_vec_r_new: src/lib_sie_tri_calculation_mod.f90:1006-1024 */
static void normal_vec(const spo_ctx *c, const double cr[3][3], double *n) {
  double v1[3], v2[3], x[3];
  if (c->p.normal_convention == 1) { vsub(cr[0], cr[1], v1); vsub(cr[2], cr[1], v2); }
  else { vsub(cr[2], cr[0], v1); vsub(cr[1], cr[0], v2); }
  vcross(v1, v2, x);
  double l = vlen(x);
  n[0] = x[0] / l; n[1] = x[1] / l; n[2] = x[2] / l;
}

/* fn_parameter_type: src/lib_sie_tri_calculation_mod.f90:1539-1586 */
static void build_fn(spo_ctx *c) {
  int m, ot, j;
  c->fn = xmalloc(sizeof(fnrwg) * 2 * c->N);
  c->rref = xmalloc(sizeof(double) * c->N);
  for (m = 0; m < c->N; ++m) {
    for (ot = 1; ot <= 2; ++ot) {
      int e = c->pelem[2 * m + ot - 1] - 1, nv = 0;
      const int *cr = &c->tt[3 * e];
      double pcv[3] = {0, 0, 0};
      for (j = 0; j < 3; ++j)
        if (cr[j] != c->pcorner[2 * m] && cr[j] != c->pcorner[2 * m + 1]) {
          memcpy(pcv, &c->pp[3 * (cr[j] - 1)], 3 * sizeof(double));
          nv = j + 1;
        }
      int i1, i2;
      if (nv == 1) { i1 = cr[1]; i2 = cr[2]; }
      else if (nv == 2) { i1 = cr[2]; i2 = cr[0]; }
      else { i1 = cr[0]; i2 = cr[1]; }
      fnrwg *f = &c->fn[2 * m + ot - 1];
      memcpy(f->c[0], &c->pp[3 * (i1 - 1)], 3 * sizeof(double));
      memcpy(f->c[1], &c->pp[3 * (i2 - 1)], 3 * sizeof(double));
      memcpy(f->c[2], pcv, 3 * sizeof(double));
      double d[3], a1[3], a2[3], x[3];
      vsub(f->c[0], f->c[1], d);
      f->length = vlen(d);
      vsub(f->c[0], f->c[2], a1); vsub(f->c[1], f->c[2], a2);
      vcross(a1, a2, x);
      f->area = 0.5 * vlen(x);
      f->sign = -1 * (((ot - 1) * ot) - 1);
      f->df = f->sign * f->length / f->area;
      memcpy(f->mid, &c->cent[3 * e], 3 * sizeof(double));
      normal_vec(c, f->c, f->nq);
    }
    /* r_ref: src/lib_sie_tri_mlfmm_interface.f90:840-842 */
    double d[3];
    vsub(&c->cent[3 * (c->pelem[2 * m] - 1)], &c->cent[3 * (c->pelem[2 * m + 1] - 1)], d);
    c->rref[m] = vlen(d);
  }
}

spo_ctx *spo_create(const spo_params *prm, int np, const double *pp, int nt,
                    const int *tt) {
  spo_ctx *c = xmalloc(sizeof(spo_ctx));
  c->p = *prm;
  if (c->p.ngp <= 0) c->p.ngp = 3;
  if (c->p.s_opt <= 0) c->p.s_opt = 250;
  if (c->p.K <= 0) c->p.K = 19;
  if (c->p.formulation <= 0) c->p.formulation = SPO_PMCHWT;
  if (c->p.normal_convention <= 0) c->p.normal_convention = 1;
  /* set_local_material_parameters_tri */
  double lam = c->p.lambda;
  c->k0 = 2 * SPO_PI / lam;
  c->omega = 2 * SPO_PI * C0_F / lam;
  c->my1 = MY0_F; c->my2 = MY0_F;
  zc er1 = ZCMPLX(c->p.eps_r1_re, c->p.eps_r1_im);
  zc er2 = ZCMPLX(c->p.eps_r2_re, c->p.eps_r2_im);
  c->k1 = c->k0 * csqrt(er1);
  c->k2 = c->k0 * csqrt(er2);
  c->eps1 = EPS0_F * er1;
  c->eps2 = EPS0_F * er2;
  c->eta1 = csqrt(c->my1 / c->eps1);
  c->eta2 = csqrt(c->my2 / c->eps2);
  c->eta_a = (c->eta1 + c->eta2) / 2;
  c->np = np; c->nt = nt;
  c->pp = xmalloc(sizeof(double) * 3 * np);
  memcpy(c->pp, pp, sizeof(double) * 3 * np);
  c->tt = xmalloc(sizeof(int) * 3 * nt);
  memcpy(c->tt, tt, sizeof(int) * 3 * nt);
  /* find_midpoint_element: (v1+v2+v3)/3.0 */
  c->cent = xmalloc(sizeof(double) * 3 * nt);
  for (int i = 0; i < nt; ++i)
    for (int d = 0; d < 3; ++d)
      c->cent[3 * i + d] = (pp[3 * (tt[3 * i] - 1) + d] + pp[3 * (tt[3 * i + 1] - 1) + d] +
                            pp[3 * (tt[3 * i + 2] - 1) + d]) / 3.0;
  build_pairs(c);
  build_fn(c);
  return c;
}

void spo_destroy(spo_ctx *c) {
  if (!c) return;
  free(c->pp); free(c->tt); free(c->cent); free(c->pcorner); free(c->pelem);
  free(c->fn); free(c->rref); free(c->leaf); free(c->box_elems); free(c->leaf_pos);
  if (c->lev) {
    for (int l = 0; l <= c->l_max; ++l) {
      free(c->lev[l].n); free(c->lev[l].first); free(c->lev[l].count);
    }
    free(c->lev);
  }
  free(c->ia); free(c->ja); free(c->za);
  free(c->khat); free(c->ttheta); free(c->tphi); free(c->trf); free(c->rt);
  for (int l = 0; l < 32; ++l) { free(c->khl[l]); free(c->trfl[l]); }
  if (c->tl) {
    for (int l = 0; l <= c->l_max; ++l) free(c->tl[l]);
    free(c->tl);
  }
  free(c);
}

int spo_n_edges(const spo_ctx *c) { return c->N; }

void spo_get_pairs(const spo_ctx *c, int *corner, int *element) {
  memcpy(corner, c->pcorner, sizeof(int) * 2 * c->N);
  memcpy(element, c->pelem, sizeof(int) * 2 * c->N);
}

void spo_get_fn(const spo_ctx *c, int m, int ot, double *o) {
  const fnrwg *f = &c->fn[2 * m + ot - 1];
  for (int a = 0; a < 3; ++a)
    for (int d = 0; d < 3; ++d) o[3 * a + d] = f->c[a][d];
  o[9] = f->area; o[10] = f->length; o[11] = f->df;
  memcpy(o + 12, f->mid, 3 * sizeof(double));
  memcpy(o + 15, f->nq, 3 * sizeof(double));
  o[18] = f->sign;
}

/* ------------------------------------------------------------------------ */
/* Green's functions: src/lib_sie_tri_calculation_mod.f90:1421-1451          */
/* ------------------------------------------------------------------------ */
static inline zc zexp_mikr(zc km, double R) {
  /* exp(-im*km*R): -im*km = (Im km, -Re km) */
  return cexp(ZCMPLX(cimag(km) * R, -creal(km) * R));
}
static void diver_green(zc km, double R, zc *G, zc *DG) {
  zc e = zexp_mikr(km, R);
  zc ikr = ZCMPLX(-cimag(km) * R, creal(km) * R); /* im*km*R */
  *DG = -((1.0 + ikr) * e / (4 * SPO_PI * (R * R)));
  *G = e / (4 * SPO_PI * R);
}
static void diver_green_s(zc km, double R, zc *Gs, zc *DGs) {
  zc e = zexp_mikr(km, R);
  if (R == 0.0) {
    *Gs = ZCMPLX(cimag(km), -creal(km)) / (4 * SPO_PI); /* -im*km */
    *DGs = km * km / (8 * SPO_PI);
  } else {
    zc ikr = ZCMPLX(-cimag(km) * R, creal(km) * R);
    *Gs = (e - 1.0) / (4 * SPO_PI * R) + km * km * R / (8 * SPO_PI);
    *DGs = (1.0 - (1.0 + ikr) * e) / (4 * SPO_PI * (R * R)) + km * km / (8 * SPO_PI);
  }
}

/* r_simplex_new: src/lib_sie_tri_calculation_mod.f90:2770-2779 */
static inline void simplex(double a, double b, const double c[3][3], double *o) {
  for (int d = 0; d < 3; ++d) o[d] = (1 - a - b) * c[0][d] + a * c[1][d] + b * c[2][d];
}

/* ------------------------------------------------------------------------ */
/* Hanninen analytic integrals (PIER 63, 2006)                               */
/* ------------------------------------------------------------------------ */
/* Iq_L: src/lib_sie_tri_calculation_mod.f90:1074-1119.  out: n1, p1, p3 */
static void iq_l_all(const double *Ra, const double *p1, const double *p2, double *o) {
  double d[3], s[3], t1[3], t2[3];
  vsub(p2, p1, d);
  double l = vlen(d);
  s[0] = d[0] / l; s[1] = d[1] / l; s[2] = d[2] / l;
  vsub(p2, Ra, t2); vsub(p1, Ra, t1);
  double Sp = vdot(t2, s), Sn = vdot(t1, s);
  double a1[3], a2[3];
  vsub(Ra, p1, a1); vsub(Ra, p2, a2);
  double Rn = vlen(a1), Rp = vlen(a2);
  double Sum_p = Rp + Sp, Sum_n = Rn + Sn, Dif_p = Rp - Sp, Dif_n = Rn - Sn;
  double R0 = sqrt(Rp * Rp - Sp * Sp);
  double n1;
  if (R0 == 0.0) n1 = log(Sp / Sn);
  else if (fabs(Dif_p) > fabs(Sum_n)) n1 = log(Dif_n / Dif_p);
  else n1 = log(Sum_p / Sum_n);
  double p1v = 0.5 * (R0 * R0) * n1 + 0.5 * (Sp * Rp - Sn * Rn);
  double p3v = 0.75 * (R0 * R0) * p1v + 0.25 * (Sp * (Rp * Rp * Rp) - Sn * (Rn * Rn * Rn));
  o[0] = n1; o[1] = p1v; o[2] = p3v;
}

double spo_iq_l(int q, const double *ra, const double *p1, const double *p2) {
  double o[3];
  iq_l_all(ra, p1, p2, o);
  if (q == -1) return o[0];
  if (q == 1) return o[1];
  if (q == 3) return o[2];
  return NAN;
}

/* m_vector: src/lib_sie_tri_calculation_mod.f90:1040-1050 */
static void m_vector(const double *nq, const double c[3][3], double m[3][3]) {
  double d[3], x[3];
  vsub(c[2], c[1], d); vcross(nq, d, x);
  { double l = vlen(x); m[0][0] = x[0] / l; m[0][1] = x[1] / l; m[0][2] = x[2] / l; }
  vsub(c[0], c[2], d); vcross(nq, d, x);
  { double l = vlen(x); m[1][0] = x[0] / l; m[1][1] = x[1] / l; m[1][2] = x[2] / l; }
  vsub(c[1], c[0], d); vcross(nq, d, x);
  { double l = vlen(x); m[2][0] = x[0] / l; m[2][1] = x[1] / l; m[2][2] = x[2] / l; }
}

/* edge_vectors_assigned: src/lib_sie_tri_calculation_mod.f90:2706-2717
 * edge i: p_s, p_t */
static void edge_ps_pt(const double c[3][3], int i, const double **ps, const double **pt) {
  if (i == 0) { *ps = c[2]; *pt = c[1]; }
  else if (i == 1) { *ps = c[0]; *pt = c[2]; }
  else { *ps = c[1]; *pt = c[0]; }
}

/* All analytic quantities for one observation point and one source fn.
 * This restates Iq_s (:1256-1331), K2q (:1333-1358), K3q (:1395-1419) and
 * K4q (:1360-1393) with each Iq_L evaluated once per edge and ordering. */
typedef struct {
  double Is_n3, Is_n1, Is_p1;      /* Iq_s internal values (before masking) */
  double h;
  double m[3][3];
  double ILst[3][3];               /* Iq_L(.., p_s, p_t): n1, p1, p3 per edge */
} anal_t;

static void analytic_eval(const double *Ra, const fnrwg *f, int iqs_l, anal_t *A) {
  double arr[3][3], d[3];
  int i;
  for (i = 0; i < 3; ++i) {
    vsub(f->c[i], Ra, d);
    double l = vlen(d);
    arr[i][0] = d[0] / l; arr[i][1] = d[1] / l; arr[i][2] = d[2] / l;
  }
  double x = 1 + vdot(arr[0], arr[1]) + vdot(arr[0], arr[2]) + vdot(arr[1], arr[2]);
  double r[3];
  vcross(arr[1], arr[2], r);
  double y = fabs(vdot(arr[0], r));
  double Omeg = 2 * atan2(y, x);
  vsub(Ra, f->mid, d);
  double h = vdot(f->nq, d);
  double Is_n3 = 0.0;
  if (h == 0.0) Is_n3 = 0.0;
  else if (iqs_l) Is_n3 = 0.0;
  else if (Omeg <= SPO_PI && Omeg > -SPO_PI) Is_n3 = (-1 / h) * Omeg;
  m_vector(f->nq, f->c, A->m);
  double tv[3], sum_il = 0.0, ILts[3][3];
  for (i = 0; i < 3; ++i) {
    const double *ps, *pt;
    edge_ps_pt(f->c, i, &ps, &pt);
    double mp[3];
    for (int k = 0; k < 3; ++k) mp[k] = Ra[k] - (pt[k] + ps[k]) / 2;
    tv[i] = vdot(A->m[i], mp);
    iq_l_all(Ra, pt, ps, ILts[i]);   /* Iq_s uses (p_t, p_s) */
    iq_l_all(Ra, ps, pt, A->ILst[i]); /* K2q/K3q/K4q use (p_s, p_t) */
    sum_il = sum_il + tv[i] * ILts[i][0];
  }
  /* n = -1: 1/(real(n)+2) == 1.0 */
  double Is_n1 = 1.0 * ((-1 * (h * h)) * Is_n3 - sum_il);
  sum_il = 0.0;
  for (i = 0; i < 3; ++i) sum_il = sum_il + tv[i] * ILts[i][1];
  /* n = 1: 1.0/(real(n)+2.0) evaluated in single precision */
  double Is_p1 = (double)(1.0f / 3.0f) * ((1 * (h * h)) * Is_n1 - sum_il);
  A->Is_n3 = Is_n3; A->Is_n1 = Is_n1; A->Is_p1 = Is_p1; A->h = h;
}

/* Iq_s return value with the Iqs_L mask. */
static inline double iq_s_val(const anal_t *A, int q, int iqs_l) {
  if (iqs_l && q > -1) return 0.0;
  if (q == -3) return A->Is_n3;
  if (q == -1) return A->Is_n1;
  return A->Is_p1;
}

static void k2q_eval(int q, const double *Ra, const fnrwg *f, int iqs_l, const anal_t *A, double *o) {
  int col = (q == -1) ? 1 : 2; /* Iq_L(q+2): p1 for q=-1, p3 for q=1 */
  double sum[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) sum[k] = sum[k] + A->ILst[i][col] * A->m[i][k];
  double Iqs = iq_s_val(A, q, iqs_l);
  double rho[3];
  for (int k = 0; k < 3; ++k) rho[k] = Ra[k] - A->h * f->nq[k];
  double coef = (q == -1) ? 1.0 : (double)(1.0f / 3.0f); /* 1/(real(q)+2) single */
  for (int k = 0; k < 3; ++k)
    o[k] = (coef * sum[k] + (rho[k] - f->c[2][k]) * Iqs) * f->df * 0.5;
}

static void k3q_eval(int q, const fnrwg *f, int iqs_l, const anal_t *A, double *o) {
  int col = (q == -1) ? 0 : 1;
  double sum[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) sum[k] = sum[k] + A->m[i][k] * A->ILst[i][col];
  double Is = iq_s_val(A, q - 2, iqs_l);
  for (int k = 0; k < 3; ++k) o[k] = sum[k] - ((A->h * (double)q) * f->nq[k]) * Is;
}

static void k4q_eval(int q, const double *Ra, const fnrwg *f, int iqs_l, const anal_t *A, double *o) {
  if (iqs_l) { o[0] = o[1] = o[2] = 0.0; return; }
  double k3[3], d[3], x[3];
  k3q_eval(q, f, iqs_l, A, k3);
  vsub(Ra, f->c[2], d);
  vcross(d, k3, x);
  for (int k = 0; k < 3; ++k) o[k] = (-x[k] * f->df) * 0.5;
}

static void fn_from19(const double *fn19, fnrwg *f) {
  for (int a = 0; a < 3; ++a)
    for (int d = 0; d < 3; ++d) f->c[a][d] = fn19[3 * a + d];
  f->area = fn19[9]; f->length = fn19[10]; f->df = fn19[11];
  memcpy(f->mid, fn19 + 12, 3 * sizeof(double));
  memcpy(f->nq, fn19 + 15, 3 * sizeof(double));
  f->sign = (int)fn19[18];
}

void spo_iq_s_all(const double *ra, const double *corners9, const double *nq,
                  const double *mid, int iqs_l, double *out3) {
  fnrwg f;
  memset(&f, 0, sizeof f);
  for (int a = 0; a < 3; ++a)
    for (int d = 0; d < 3; ++d) f.c[a][d] = corners9[3 * a + d];
  memcpy(f.nq, nq, 3 * sizeof(double));
  memcpy(f.mid, mid, 3 * sizeof(double));
  anal_t A;
  analytic_eval(ra, &f, iqs_l, &A);
  out3[0] = A.Is_n3; out3[1] = A.Is_n1; out3[2] = A.Is_p1;
}
void spo_k2q(int q, const double *ra, const double *fn19, int iqs_l, double *o) {
  fnrwg f; anal_t A; fn_from19(fn19, &f); analytic_eval(ra, &f, iqs_l, &A);
  k2q_eval(q, ra, &f, iqs_l, &A, o);
}
void spo_k3q(int q, const double *ra, const double *fn19, int iqs_l, double *o) {
  fnrwg f; anal_t A; fn_from19(fn19, &f); analytic_eval(ra, &f, iqs_l, &A);
  k3q_eval(q, &f, iqs_l, &A, o);
}
void spo_k4q(int q, const double *ra, const double *fn19, int iqs_l, double *o) {
  fnrwg f; anal_t A; fn_from19(fn19, &f); analytic_eval(ra, &f, iqs_l, &A);
  k4q_eval(q, ra, &f, iqs_l, &A, o);
}

/* ------------------------------------------------------------------------ */
/* Element-pair integrals                                                    */
/* ------------------------------------------------------------------------ */
typedef struct { zc Kn, Kt, Tn, Tt; } relem;
typedef struct { zc IK[3], IT_a[3], IT_bn[3], IT_bt; } rinner;

/* tri_inner_integration: :2048-2078 */
static void inner_normal(const double *ra, const fnrwg *fn, int ngp, zc kd, rinner *rv,
                         const double *qa, const double *qb, const double *qw) {
  memset(rv, 0, sizeof *rv);
  for (int j = 0; j < ngp; ++j) {
    double rq[3], rfn[3], R[3], Rn[3], x[3];
    simplex(qa[j], qb[j], fn->c, rq);
    for (int k = 0; k < 3; ++k) rfn[k] = 0.5 * (rq[k] - fn->c[2][k]) * fn->df;
    vsub(ra, rq, R);
    double Rl = vlen(R);
    for (int k = 0; k < 3; ++k) Rn[k] = R[k] / Rl;
    zc G, DG;
    diver_green(kd, vlen(R), &G, &DG);
    double TRF = qw[j] * fn->area;
    vcross(rfn, Rn, x);
    for (int k = 0; k < 3; ++k) {
      rv->IK[k] = rv->IK[k] + x[k] * DG * TRF;
      rv->IT_a[k] = rv->IT_a[k] + rfn[k] * G * TRF;
      rv->IT_bn[k] = rv->IT_bn[k] + (fn->df * Rn[k]) * DG * TRF;
    }
    rv->IT_bt = rv->IT_bt + fn->df * G * TRF;
  }
}

/* tri_inner_singular_integration_smooth: :2080-2113 */
static void inner_smooth(const double *ra, const fnrwg *fn, int ngp, zc kd, int iqs_l,
                         rinner *rv, const double *qa, const double *qb, const double *qw) {
  memset(rv, 0, sizeof *rv);
  for (int j = 0; j < ngp; ++j) {
    double rq[3], rfn[3], R[3], Rn[3], x[3];
    simplex(qa[j], qb[j], fn->c, rq);
    for (int k = 0; k < 3; ++k) rfn[k] = 0.5 * (rq[k] - fn->c[2][k]) * fn->df;
    vsub(ra, rq, R);
    double Rl = vlen(R);
    zc Gs, DGs;
    diver_green_s(kd, Rl, &Gs, &DGs);
    double TRF = qw[j] * fn->area;
    if (!iqs_l) {
      for (int k = 0; k < 3; ++k) Rn[k] = R[k] / Rl;
      vcross(rfn, Rn, x);
      for (int k = 0; k < 3; ++k) {
        rv->IK[k] = rv->IK[k] + x[k] * DGs * TRF;
        rv->IT_bn[k] = rv->IT_bn[k] + (fn->df * Rn[k]) * DGs * TRF;
      }
    }
    for (int k = 0; k < 3; ++k) rv->IT_a[k] = rv->IT_a[k] + rfn[k] * Gs * TRF;
    rv->IT_bt = rv->IT_bt + fn->df * Gs * TRF;
  }
}

/* tri_inner_singular_integration_singular: :2115-2145 (analytic part) */
static void inner_singular(const double *ra, const fnrwg *fn, zc kd, int iqs_l,
                           const anal_t *A, rinner *rv) {
  double k2n[3], k2p[3], k3n[3], k3p[3], k4n[3], k4p[3];
  double Iqs_n1 = iq_s_val(A, -1, iqs_l), Iqs_p1 = iq_s_val(A, 1, iqs_l);
  k2q_eval(-1, ra, fn, iqs_l, A, k2n);
  k3q_eval(-1, fn, iqs_l, A, k3n);
  k4q_eval(-1, ra, fn, iqs_l, A, k4n);
  k2q_eval(1, ra, fn, iqs_l, A, k2p);
  k3q_eval(1, fn, iqs_l, A, k3p);
  k4q_eval(1, ra, fn, iqs_l, A, k4p);
  zc kd2 = kd * kd;
  for (int k = 0; k < 3; ++k) {
    rv->IK[k] = iqs_l ? 0.0 : (1 / (4 * SPO_PI)) * k4n[k] - kd2 / (8 * SPO_PI) * k4p[k];
    rv->IT_a[k] = (1 / (4 * SPO_PI)) * k2n[k] - kd2 / (8 * SPO_PI) * k2p[k];
    /* -(fn%df/(4*PI)*K3_n1 - kd**2*fn%df/(8*PI))*K3_p1  (elementwise) */
    rv->IT_bn[k] = -((fn->df / (4 * SPO_PI)) * k3n[k] - kd2 * fn->df / (8 * SPO_PI)) * k3p[k];
  }
  rv->IT_bt = (fn->df / (4 * SPO_PI)) * Iqs_n1 - kd2 * fn->df / (8 * SPO_PI) * Iqs_p1;
}

/* tri_outer_integration (:2255-2302) and tri_outer_integration_singular
 * (:2199-2253).  singular: 0 normal, 1 singular. */
static relem outer(const spo_ctx *c, const fnrwg *fm, const fnrwg *fn, int ngp, int md,
                   int singular, int iqs_l) {
  double qa[20], qb[20], qw[20];
  spo_quadrature_tri(ngp, qa, qb, qw);
  relem rv = {0, 0, 0, 0};
  zc kd = (md == 1) ? c->k1 : c->k2;
  double x_nn = 0.0, x_nt = 0.0;
  for (int j = 0; j < ngp; ++j) {
    double ra[3], rfm[3], rfn[3];
    simplex(qa[j], qb[j], fm->c, ra);
    for (int k = 0; k < 3; ++k) {
      rfm[k] = 0.5 * (ra[k] - fm->c[2][k]) * fm->df;
      rfn[k] = 0.5 * (ra[k] - fn->c[2][k]) * fn->df;
    }
    double TRF = qw[j] * fm->area;
    rinner In;
    zc tmpv[3], t1[3], t2[3], av[3];
    if (!singular) {
      inner_normal(ra, fn, ngp, kd, &In, qa, qb, qw);
      for (int k = 0; k < 3; ++k) av[k] = kd * In.IT_a[k] + 1.0 / kd * In.IT_bn[k];
      vcross_rc(fm->nq, av, tmpv);
    } else {
      anal_t A;
      rinner Is, Ism;
      analytic_eval(ra, fn, iqs_l, &A);
      inner_singular(ra, fn, kd, iqs_l, &A, &Is);
      inner_smooth(ra, fn, ngp, kd, iqs_l, &Ism, qa, qb, qw);
      for (int k = 0; k < 3; ++k) {
        In.IT_bn[k] = Is.IT_bn[k] + Ism.IT_bn[k];
        In.IK[k] = Is.IK[k] + Ism.IK[k];
        In.IT_a[k] = Is.IT_a[k] + Ism.IT_a[k];
      }
      In.IT_bt = Is.IT_bt + Ism.IT_bt;
      for (int k = 0; k < 3; ++k) { av[k] = kd * In.IT_a[k]; }
      vcross_rc(fm->nq, av, t1);
      for (int k = 0; k < 3; ++k) { av[k] = 1.0 / kd * In.IT_bn[k]; }
      vcross_rc(fm->nq, av, t2);
      for (int k = 0; k < 3; ++k) tmpv[k] = t1[k] + t2[k];
    }
    zc nk[3];
    vcross_rc(fm->nq, In.IK, nk);
    rv.Kn = rv.Kn + zdot_rc(rfm, nk) * TRF;
    rv.Kt = rv.Kt + zdot_rc(rfm, In.IK) * TRF;
    rv.Tn = rv.Tn + ZCMPLX(0.0, TRF) * zdot_rc(rfm, tmpv);
    zc tmp_s = 1.0 / kd * fm->df * In.IT_bt;
    rv.Tt = rv.Tt + ZCMPLX(0.0, TRF) * (kd * zdot_rc(rfm, In.IT_a) - tmp_s);
    double nx[3];
    vcross(fm->nq, rfm, nx);
    x_nt = x_nt - vdot(nx, rfn) * TRF;
    x_nn = x_nn - vdot(rfm, rfn) * TRF;
  }
  double sg = (md == 1) ? -0.5 : 0.5;
  rv.Kn = rv.Kn + sg * (double)(float)x_nn;
  rv.Kt = rv.Kt + sg * (double)(float)x_nt;
  return rv;
}

/* tri_outer_integration_edge_adjacent (:2165-2197) with
 * tri_inner_integration_edge_adjacent (:2011-2044), 12x12 rule. */
static zc outer_edge_adjacent_parts(const spo_ctx *c, const fnrwg *fm, const fnrwg *fn, int md,
                                    zc *kint, double *xnt) {
  double qa[20], qb[20], qw[20];
  const int ngp = 12;
  spo_quadrature_tri(ngp, qa, qb, qw);
  zc kd = (md == 1) ? c->k1 : c->k2;
  zc Kt = 0.0;
  double x_nt = 0.0;
  for (int j = 0; j < ngp; ++j) {
    double ra[3], rfm[3], rfn[3];
    simplex(qa[j], qb[j], fm->c, ra);
    for (int k = 0; k < 3; ++k) {
      rfm[k] = 0.5 * (ra[k] - fm->c[2][k]) * fm->df;
      rfn[k] = 0.5 * (ra[k] - fn->c[2][k]) * fn->df;
    }
    double TRF = qw[j] * fm->area;
    zc Ik[3] = {0, 0, 0};
    for (int i = 0; i < ngp; ++i) {
      double rq[3], rfq[3], R[3], Rn[3], x[3];
      simplex(qa[i], qb[i], fn->c, rq);
      for (int k = 0; k < 3; ++k) rfq[k] = 0.5 * (rq[k] - fn->c[2][k]) * fn->df;
      vsub(ra, rq, R);
      double Rl = vlen(R);
      for (int k = 0; k < 3; ++k) Rn[k] = R[k] / Rl;
      zc G, DG;
      diver_green(kd, vlen(R), &G, &DG);
      double TRFi = qw[i] * fn->area;
      vcross(rfq, Rn, x);
      for (int k = 0; k < 3; ++k) Ik[k] = Ik[k] + x[k] * DG * TRFi;
    }
    Kt = Kt + zdot_rc(rfm, Ik) * TRF;
    double nx[3];
    vcross(fm->nq, rfm, nx);
    x_nt = x_nt - vdot(rfn, nx) * TRF;
  }
  double sg = (md == 1) ? -0.5 : 0.5;
  *kint = Kt;
  *xnt = x_nt;
  return Kt + sg * (double)(float)x_nt;
}

static zc outer_edge_adjacent(const spo_ctx *c, const fnrwg *fm, const fnrwg *fn, int md) {
  zc kint;
  double xnt;
  return outer_edge_adjacent_parts(c, fm, fn, md, &kint, &xnt);
}

/* edge_adjacent_tri: :2544-2569 (2 common corners) */
static int edge_adjacent(const spo_ctx *c, int e1, int e2) {
  int cnt = 0;
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      if (c->tt[3 * e1 + j] == c->tt[3 * e2 + i]) ++cnt;
  return cnt == 2;
}

/* formulation_coefficient: :368-387 */
static void form_coef(const spo_ctx *c, zc *a, zc *d) {
  double x = 1.0, y = 0.50;
  a[0] = 2 * x * c->eta1 / (c->eta1 + y * c->eta2);
  a[1] = 2 * x * c->eta2 / (c->eta1 + y * c->eta2);
  d[0] = (c->eta1 + y * c->eta2) / (2 * x * c->eta1);
  d[1] = (c->eta1 + y * c->eta2) / (2 * x * c->eta2);
}

/* tri_impedance_building_*: :1713-2009.  Returns zz (3 or 4 entries). */
static void building(const spo_ctx *c, int m, int n, int tt, int ll, int singular, zc *zz) {
  const fnrwg *fm = &c->fn[2 * m + tt - 1];
  const fnrwg *fn = &c->fn[2 * n + ll - 1];
  int ngp = c->p.ngp, form = c->p.formulation;
  relem r1, r2;
  int edge_swap = (form == SPO_PMCHWT || form == SPO_MCTF || form == SPO_PMCHWT_S ||
                   form == SPO_JMCFIE);
  if (singular) {
    int pp = c->pelem[2 * m + tt - 1] - 1, qq = c->pelem[2 * n + ll - 1] - 1;
    int iqs_l = (pp == qq);
    r1 = outer(c, fm, fn, ngp, 1, 1, iqs_l);
    r2 = outer(c, fm, fn, ngp, 2, 1, iqs_l);
    if (edge_swap && !iqs_l && edge_adjacent(c, pp, qq)) {
      zc k1t = outer_edge_adjacent(c, fm, fn, 1);
      zc k2t = outer_edge_adjacent(c, fm, fn, 2);
      r1.Kt = k1t; r2.Kt = k2t;
    }
  } else {
    r1 = outer(c, fm, fn, ngp, 1, 0, 0);
    r2 = outer(c, fm, fn, ngp, 2, 0, 0);
  }
  zc e1 = c->eta1, e2 = c->eta2;
  switch (form) {
    case SPO_PMCHWT: case SPO_MCTF: case SPO_PMCHWT_S:
      zz[0] = e1 * r1.Tt + e2 * r2.Tt;
      zz[1] = -(r1.Kt + r2.Kt);
      zz[2] = 1.0 / e1 * r1.Tt + 1.0 / e2 * r2.Tt;
      break;
    case SPO_JMCFIE:
      zz[0] = r1.Kn - r2.Kn + r1.Tt + r2.Tt;
      zz[1] = -(1.0 / e1 * (r1.Kt - r1.Tn) + 1.0 / e2 * (r2.Kt + r2.Tn));
      zz[2] = -e1 * (r1.Tn - r1.Kt) + e2 * (r2.Kt + r2.Tn);
      break;
    case SPO_MCTF2: {
      zc a[2], d[2];
      form_coef(c, a, d);
      zz[0] = r1.Tt * a[0] + r2.Tt * a[1];
      zz[1] = -(r1.Kt * a[0] / e1 + r2.Kt * a[1] / e2);
      zz[2] = e1 * d[0] * r1.Kt + r2.Kt * e2 * d[1];
      zz[3] = d[0] * r1.Tt + d[1] * r2.Tt;
    } break;
    case SPO_ICTF:
      zz[0] = (e1 * r1.Tt + e2 * r2.Tt) / c->eta_a;
      zz[1] = -(r1.Kt + r2.Kt) / c->eta_a;
      zz[2] = (r1.Kt + r2.Kt) * c->eta_a;
      zz[3] = (r1.Tt / e1 + r2.Tt / e2) * c->eta_a;
      break;
    case SPO_CNF:
      zz[0] = r1.Kn - r2.Kn;
      zz[1] = r1.Tn / e1 - r2.Tn / e2;
      zz[2] = e2 * r2.Tn - e1 * r1.Tn;
      break;
    case SPO_MNMF: {
      zc b1 = c->eps1 / (c->eps1 + c->eps2), b2 = c->eps2 / (c->eps1 + c->eps2);
      zc c1 = c->my1 / (c->my1 + c->my2), c2 = c->my2 / (c->my1 + c->my2);
      zz[0] = b1 * r1.Kn - b2 * r2.Kn;
      zz[1] = b1 / e1 * r1.Tn - b2 / e2 * r2.Tn;
      zz[2] = e2 * c2 * r2.Tn - e1 * c1 * r1.Tn;
      zz[3] = c1 * r1.Kn - c2 * r2.Kn;
    } break;
  }
}

/* lib_sie_tri_get_impedance_edges: :463-644 */
static void impedance_edges(const spo_ctx *c, int m, int n, double r_ref, zc *D) {
  D[0] = D[1] = D[2] = D[3] = 0.0;
  int form = c->p.formulation;
  for (int t = 1; t <= 2; ++t)
    for (int s = 1; s <= 2; ++s) {
      int p = c->pelem[2 * m + t - 1] - 1, q = c->pelem[2 * n + s - 1] - 1;
      double d[3];
      vsub(&c->cent[3 * p], &c->cent[3 * q], d);
      double r_ob = vlen(d);
      int singular = !(r_ob > 1.5 * r_ref);
      zc zz[4] = {0, 0, 0, 0};
      building(c, m, n, t, s, singular, zz);
      switch (form) {
        case SPO_PMCHWT:
          D[0] = D[0] + zz[0]; D[1] = D[1] + zz[1]; D[2] = -D[1]; D[3] = D[3] + zz[2];
          break;
        case SPO_MCTF:
          D[0] = D[0] + zz[0]; D[1] = D[1] + zz[1];
          D[2] = -D[1] * c->eta1 * c->eta2; D[3] = D[3] + zz[2] * c->eta1 * c->eta2;
          break;
        case SPO_PMCHWT_S:
          D[0] = D[0] + zz[0]; D[1] = D[1] + zz[1]; D[2] = -D[2] + zz[1]; D[3] = D[3] + zz[2];
          break;
        case SPO_JMCFIE:
          D[0] = D[0] + zz[0]; D[1] = D[1] + zz[1]; D[2] = D[2] + zz[2]; D[3] = D[0];
          break;
        case SPO_CNF:
          D[0] = D[0] + zz[0]; D[1] = D[1] + zz[1]; D[2] = D[2] + zz[2]; D[3] = D[3] + zz[3];
          break;
        default: /* ICTF, MCTF2, MNMF */
          D[0] = D[0] + zz[0]; D[1] = D[1] + zz[1]; D[2] = D[2] + zz[2]; D[3] = D[3] + zz[3];
          break;
      }
    }
  if (form == SPO_PMCHWT_S) {
    D[1] = D[1] * c->eta1; D[2] = D[2] * c->eta1; D[3] = D[3] * (c->eta1 * c->eta1);
  }
}

/* Probe: the 12x12 edge-adjacent K (outer_edge_adjacent) of side tt of edge m
 * against side ll of edge n in medium md (1 outside, 2 inside):
 * out = (K re, K im, integral part re, im, x_nt before its float rounding). */
void spo_edge_adjacent_kt(const spo_ctx *c, int m, int tt, int n, int ll, int md, double *out) {
  zc kint;
  double xnt;
  zc k = outer_edge_adjacent_parts(c, &c->fn[2 * m + tt - 1], &c->fn[2 * n + ll - 1], md, &kint,
                                   &xnt);
  out[0] = creal(k);
  out[1] = cimag(k);
  out[2] = creal(kint);
  out[3] = cimag(kint);
  out[4] = xnt;
}

void spo_near_block(const spo_ctx *c, int m, int n, double *out) {
  zc D[4];
  impedance_edges(c, m, n, c->rref[m], D);
  for (int k = 0; k < 4; ++k) { out[2 * k] = creal(D[k]); out[2 * k + 1] = cimag(D[k]); }
}

/* ------------------------------------------------------------------------ */
/* Octree: src/lib/libtree/src/lib/lib_tree.f90, lib_tree_helper_functions  */
/* ------------------------------------------------------------------------ */
#define TH_LEVEL 21 /* int(8*8*1.0/3): lib_tree.f90:235 */

static uint64_t morton3(uint64_t x, uint64_t y, uint64_t z, int l) {
  /* lib_tree_hf_interleave_bits: bit i of x -> 3i+2, y -> 3i+1, z -> 3i */
  uint64_t n = 0;
  for (int i = 0; i < l; ++i)
    n |= (((x >> i) & 1) << (3 * i + 2)) | (((y >> i) & 1) << (3 * i + 1)) | (((z >> i) & 1) << (3 * i));
  return n;
}
static void demorton3(uint64_t n, int l, int64_t *x) {
  x[0] = x[1] = x[2] = 0;
  for (int i = 0; i < l; ++i) {
    x[0] |= (int64_t)((n >> (3 * i + 2)) & 1) << i;
    x[1] |= (int64_t)((n >> (3 * i + 1)) & 1) << i;
    x[2] |= (int64_t)((n >> (3 * i)) & 1) << i;
  }
}

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

static int level_find(const level_t *L, int64_t n) {
  int lo = 0, hi = L->nbox - 1;
  while (lo <= hi) {
    int mid = (lo + hi) / 2;
    if (L->n[mid] == n) return mid;
    if (L->n[mid] < n) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}

/* lib_tree_constructor with s (:176-320), scaling (:370-434), universal
 * index (lib_tree_helper_functions.f90:217-379, mantissa bits :547-629),
 * lib_tree_get_level_max (:1290-1344). */
int spo_tree_build(spo_ctx *c) {
  int N = c->N, m, d;
  double *pts = xmalloc(sizeof(double) * 3 * N);
  for (m = 0; m < N; ++m) {
    const fnrwg *f = &c->fn[2 * m];  /* fn_edge_center, ot = 1 */
    for (d = 0; d < 3; ++d) pts[3 * m + d] = (f->c[0][d] + f->c[1][d]) / 2;
  }
  double xmin[3], xmax[3], Dd[3];
  for (d = 0; d < 3; ++d) {
    xmin[d] = pts[d]; xmax[d] = pts[d];
    for (m = 1; m < N; ++m) {
      if (pts[3 * m + d] > xmax[d]) xmax[d] = pts[3 * m + d];
      if (pts[3 * m + d] < xmin[d]) xmin[d] = pts[3 * m + d];
    }
    Dd[d] = (xmax[d] - xmin[d]) * (1 + ldexp(1.0, -52));
  }
  double D = Dd[0];
  if (Dd[1] > D) D = Dd[1];
  if (Dd[2] > D) D = Dd[2];
  c->D = D;
  memcpy(c->xmin, xmin, sizeof xmin);
  uint64_t *u = xmalloc(sizeof(uint64_t) * N);
  for (m = 0; m < N; ++m) {
    uint64_t ix[3];
    for (d = 0; d < 3; ++d) {
      double xs = (pts[3 * m + d] - xmin[d]) / D;
      ix[d] = (uint64_t)(xs * 2097152.0); /* floor(xs*2^21), xs in [0,1) */
    }
    u[m] = morton3(ix[0], ix[1], ix[2], TH_LEVEL);
  }
  uint64_t *us = xmalloc(sizeof(uint64_t) * N);
  memcpy(us, u, sizeof(uint64_t) * N);
  qsort(us, N, sizeof(uint64_t), cmp_u64);
  int l_max = 1, s = c->p.s_opt;
  for (int i = 0; i + s < N; ++i) {
    uint64_t a = us[i], b = us[i + s];
    int j = TH_LEVEL + 1;
    for (;;) {
      j = j - 1;
      a >>= 3; b >>= 3;
      if (a == b) { if (j > l_max) l_max = j; break; }
    }
  }
  c->l_max = l_max;
  c->l_min = 2; /* lib_tree_get_level_min(k=1) */
  if (c->l_max < c->l_min) c->l_min = c->l_max;
  c->leaf = xmalloc(sizeof(int64_t) * N);
  for (m = 0; m < N; ++m) c->leaf[m] = (int64_t)(u[m] >> (3 * (TH_LEVEL - l_max)));
  /* boxes per level */
  c->lev = xmalloc(sizeof(level_t) * (l_max + 1));
  for (int l = 0; l <= l_max; ++l) {
    uint64_t *b = xmalloc(sizeof(uint64_t) * N);
    for (m = 0; m < N; ++m) b[m] = (uint64_t)c->leaf[m] >> (3 * (l_max - l));
    qsort(b, N, sizeof(uint64_t), cmp_u64);
    int nb = 0;
    for (m = 0; m < N; ++m) if (m == 0 || b[m] != b[m - 1]) b[nb++] = b[m];
    c->lev[l].nbox = nb;
    c->lev[l].n = xmalloc(sizeof(int64_t) * nb);
    for (m = 0; m < nb; ++m) c->lev[l].n[m] = (int64_t)b[m];
    free(b);
  }
  /* elements per leaf box, ascending element number */
  level_t *L = &c->lev[l_max];
  L->first = xmalloc(sizeof(int) * L->nbox);
  L->count = xmalloc(sizeof(int) * L->nbox);
  c->leaf_pos = xmalloc(sizeof(int) * N);
  for (m = 0; m < N; ++m) { c->leaf_pos[m] = level_find(L, c->leaf[m]); L->count[c->leaf_pos[m]]++; }
  for (int b = 1; b < L->nbox; ++b) L->first[b] = L->first[b - 1] + L->count[b - 1];
  int *fill = xmalloc(sizeof(int) * L->nbox);
  c->box_elems = xmalloc(sizeof(int) * N);
  for (m = 0; m < N; ++m) { int b = c->leaf_pos[m]; c->box_elems[L->first[b] + fill[b]++] = m; }
  free(fill); free(u); free(us); free(pts);
  c->tree_built = 1;
  return l_max;
}

void spo_tree_info(const spo_ctx *c, int *l_min, int *l_max, double *D, double *xmin3) {
  *l_min = c->l_min; *l_max = c->l_max; *D = c->D;
  memcpy(xmin3, c->xmin, 3 * sizeof(double));
}
void spo_edge_leaf(const spo_ctx *c, int64_t *u) { memcpy(u, c->leaf, sizeof(int64_t) * c->N); }

/* lib_tree_hf_get_neighbour_all_xD (:884-1027): own box first, then the
 * neighbours in nested x, y, z order (lib_tree_get_domain_e2 :710-753). */
static int e2_boxes(int64_t n, int l, int64_t *out) {
  int64_t x[3], ub = ((int64_t)1 << l) - 1;
  int cnt = 0;
  demorton3((uint64_t)n, l, x);
  out[cnt++] = n;
  for (int i = -1; i <= 1; ++i) {
    int64_t a = x[0] + i;
    if (a < 0 || a > ub) continue;
    for (int j = -1; j <= 1; ++j) {
      int64_t b = x[1] + j;
      if (b < 0 || b > ub) continue;
      for (int k = -1; k <= 1; ++k) {
        int64_t cc = x[2] + k;
        if (cc < 0 || cc > ub) continue;
        int64_t nb = (int64_t)morton3(a, b, cc, l);
        if (nb != n) out[cnt++] = nb;
      }
    }
  }
  return cnt;
}

int spo_e2_of_edge(const spo_ctx *c, int m, int *out, int cap) {
  int64_t boxes[27];
  int nb = e2_boxes(c->leaf[m], c->l_max, boxes), cnt = 0;
  const level_t *L = &c->lev[c->l_max];
  for (int i = 0; i < nb; ++i) {
    int p = level_find(L, boxes[i]);
    if (p < 0) continue;
    for (int k = 0; k < L->count[p]; ++k) {
      if (cnt < cap) out[cnt] = c->box_elems[L->first[p] + k];
      ++cnt;
    }
  }
  return cnt;
}

/* lib_tree_get_domain_i4 (:1096-1245), ascending box order. */
int spo_i4(const spo_ctx *c, int l, int64_t n, int64_t *out, int cap) {
  (void)c;
  int64_t par[27], own[27];
  int np_ = e2_boxes(n >> 3, l - 1, par), no = e2_boxes(n, l, own), cnt = 0;
  int64_t cand[27 * 8];
  for (int i = 0; i < np_; ++i)
    for (int o = 0; o < 8; ++o) cand[cnt++] = (par[i] << 3) | o;
  qsort(cand, cnt, sizeof(int64_t), cmp_u64);
  int res = 0;
  for (int i = 0; i < cnt; ++i) {
    int skip = 0;
    for (int k = 0; k < no; ++k) if (own[k] == cand[i]) { skip = 1; break; }
    if (skip) continue;
    if (res < cap) out[res] = cand[i];
    ++res;
  }
  return res;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* precalculation_near_field_csr_tri: src/lib_sie_tri_mlfmm_interface.f90:779-863 */
int64_t spo_near_csr_build(spo_ctx *c) {
  if (!c->tree_built) spo_tree_build(c);
  int N = c->N;
  free(c->ia); free(c->ja); free(c->za);
  c->ia = xmalloc(sizeof(int32_t) * (N + 1));
  int64_t nnz = 0;
  for (int m = 0; m < N; ++m) nnz += spo_e2_of_edge(c, m, NULL, 0);
  c->nnz = nnz;
  c->ja = xmalloc(sizeof(int32_t) * (nnz ? nnz : 1));
  c->za = xmalloc(sizeof(zc) * 4 * (nnz ? nnz : 1));
  int *buf = xmalloc(sizeof(int) * N);
  int64_t pos = 0;
  c->ia[0] = 0;
  for (int m = 0; m < N; ++m) {
    int k = spo_e2_of_edge(c, m, buf, N);
    qsort(buf, k, sizeof(int), cmp_int); /* sort_edge_index */
    for (int j = 0; j < k; ++j) c->ja[pos + j] = buf[j];
    pos += k;
    c->ia[m + 1] = (int32_t)pos;
  }
  free(buf);
  return nnz;
}

void spo_near_csr_get(const spo_ctx *c, int32_t *ia, int32_t *ja) {
  memcpy(ia, c->ia, sizeof(int32_t) * (c->N + 1));
  memcpy(ja, c->ja, sizeof(int32_t) * c->nnz);
}

void spo_near_fill(spo_ctx *c, int lo, int hi, int nthreads) {
  if (hi > c->N) hi = c->N;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (int m = lo; m < hi; ++m) {
    for (int64_t j = c->ia[m]; j < c->ia[m + 1]; ++j)
      impedance_edges(c, m, c->ja[j], c->rref[m], &c->za[4 * j]);
  }
  (void)nthreads;
}

void spo_near_values(const spo_ctx *c, double *abcd) {
  for (int64_t j = 0; j < 4 * c->nnz; ++j) { abcd[2 * j] = creal(c->za[j]); abcd[2 * j + 1] = cimag(c->za[j]); }
}

/* ------------------------------------------------------------------------ */
/* Special functions (precompute)                                            */
/* ------------------------------------------------------------------------ */
/* Gauss-Legendre nodes on [-1,1], ascending; Newton on the three-term
 * recurrence (the published algorithm the reference's Glaser-Liu-Rokhlin
 * legendre_compute_glr also converges to); weights renormalised to sum 2
 * as Gauss_Quadrature_Fast.f90:424-499 does. */
void spo_gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < n; ++i) {
    double z = cos(SPO_PI * (i + 0.75) / (n + 0.5)), pd = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) { double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
      if (n == 1) { p1 = z; p0 = 1.0; }
      pd = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / pd;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) { double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k; p0 = p1; p1 = p2; }
    pd = n * (z * p1 - p0) / (z * z - 1.0);
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * pd * pd);
  }
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += w[i];
  for (int i = 0; i < n; ++i) w[i] = 2.0 * w[i] / s;
}

/* Spherical Hankel h_l^(2)(z) = j_l(z) - i y_l(z), l = 0..n-1, complex z.
 * j_l by Miller backward recurrence normalised to j_0 or j_1, y_l by forward
 * recurrence -- the scheme of Zhang & Jin's CSPHJY that
 * src/lib/libmath/src/math/lib_math_bessel.f90:936-991 calls. */
void spo_hankel2_sph(int n, double zre, double zim, double *out) {
  zc z = ZCMPLX(zre, zim);
  int nmax = n - 1;
  zc *sj = xmalloc(sizeof(zc) * (n + 2)), *sy = xmalloc(sizeof(zc) * (n + 2));
  zc sj0 = csin(z) / z;
  sj[0] = sj0;
  if (nmax >= 1) {
    zc sj1 = (sj0 - ccos(z)) / z;
    sj[1] = sj1;
    /* backward recurrence from a start order well above max(nmax, |z|) */
    int st = (int)(cabs(z)) + nmax + 60;
    zc f2 = 0.0, f1 = 1e-100, f = 0.0, f0 = 0.0, fj1 = 0.0;
    zc *tmp = xmalloc(sizeof(zc) * (st + 2));
    for (int k = st; k >= 0; --k) {
      f = (2.0 * k + 3.0) * f1 / z - f2;
      if (k <= nmax + 1) tmp[k] = f;
      if (k == 1) fj1 = f;
      if (k == 0) f0 = f;
      f2 = f1; f1 = f;
      if (cabs(f) > 1e150) { /* rescale to avoid overflow */
        f *= 1e-150; f1 *= 1e-150; f2 *= 1e-150; fj1 *= 1e-150;
        for (int q = k; q <= nmax + 1 && q <= st; ++q) tmp[q] *= 1e-150;
      }
    }
    zc cs = (cabs(sj0) > cabs(sj1)) ? sj0 / f0 : sj1 / fj1;
    for (int k = 0; k <= nmax; ++k) sj[k] = cs * tmp[k];
    free(tmp);
  }
  sy[0] = -ccos(z) / z;
  if (nmax >= 1) sy[1] = (sy[0] - csin(z)) / z;
  for (int k = 2; k <= nmax; ++k) sy[k] = (2.0 * k - 1.0) * sy[k - 1] / z - sy[k - 2];
  for (int k = 0; k <= nmax; ++k) {
    zc h = sj[k] - I * sy[k];
    out[2 * k] = creal(h); out[2 * k + 1] = cimag(h);
  }
  free(sj); free(sy);
}

/* Legendre P_l(x), l = 0..n-1: recurrence of FSU lpn
 * (src/lib/libmath_external/src/fsu/fsu_special_functions.f90:19424). */
void spo_legendre_p(int n, double x, double *out) {
  if (n <= 0) return;
  out[0] = 1.0;
  if (n > 1) out[1] = x;
  double p0 = 1.0, p1 = x;
  for (int k = 2; k < n; ++k) {
    double pf = (2.0 * k - 1.0) / k * x * p1 - (k - 1.0) / k * p0;
    out[k] = pf; p0 = p1; p1 = pf;
  }
}

/* Khat_and_TRF_ExtendedK: src/lib_sie_mlfmm.f90:190-249 */
static void khat_trf(int K, double *kh, double *tth, double *tph, double *trf) {
  double *xx = xmalloc(sizeof(double) * K), *ww = xmalloc(sizeof(double) * K);
  spo_gauss_legendre(K, xx, ww);
  int Ns = 2 * K * K;
  double dphi = 2 * SPO_PI / (2 * K);
  for (int j = 0; j < K; ++j)
    for (int k = 0; k < 2 * K; ++k) {
      double phi = k * dphi;
      int i = k + j * 2 * K;
      double st = sin(acos(xx[j]));
      kh[3 * i] = st * cos(phi); kh[3 * i + 1] = st * sin(phi); kh[3 * i + 2] = xx[j];
      tth[3 * i] = xx[j] * cos(phi); tth[3 * i + 1] = xx[j] * sin(phi); tth[3 * i + 2] = -st;
      tph[3 * i] = -sin(phi); tph[3 * i + 1] = cos(phi); tph[3 * i + 2] = 0.0;
      trf[i] = ww[j] * dphi;
    }
  double np_[2][3] = {{0, 0, 1}, {0, 0, -1}};
  for (int p = 0; p < 2; ++p) {
    int i = Ns + p;
    memcpy(&kh[3 * i], np_[p], 3 * sizeof(double));
    tth[3 * i] = 1; tth[3 * i + 1] = 0; tth[3 * i + 2] = 0;
    tph[3 * i] = 0; tph[3 * i + 1] = 1; tph[3 * i + 2] = 0;
    trf[i] = 0.0;
  }
  free(xx); free(ww);
}

/* ------------------------------------------------------------------------ */
/* Lagrange interpolation of a sampled far-field pattern between the sample  */
/* grids of two truncation numbers: Lagrange_Interpolation_New and its       */
/* helpers (src/lib_sie_mlfmm.f90:251-493, 573-610), restated loop by loop   */
/* with the reference's 1-based arrays (macros below).  Default reals of the */
/* reference are single precision: the stencil indices (it-1.0)/K_n*K_m are */
/* evaluated in float.                                                       */
/* ------------------------------------------------------------------------ */

/* theta_phi_extended (:573-610): a_theta(1..3K+2) from 2pi down to -pi
 * across both poles, a_phi(1..6K) from -2pi in steps of pi/K */
static void theta_phi_ext(int Km, double *a_phi, double *a_theta) {
  int m_ph = 3 * (2 * Km);
  double d_phi = 6 * SPO_PI / m_ph;
  double *xx = xmalloc(sizeof(double) * Km), *ww = xmalloc(sizeof(double) * Km);
  spo_gauss_legendre(Km, xx, ww);
#define AT(i) a_theta[(i) - 1]
#define XX(i) xx[(i) - 1]
  for (int j = 1; j <= Km; ++j) {
    AT(j) = 2 * SPO_PI - acos(XX(Km - (j - 1)));
    AT(j + Km + 1) = acos(XX(j));
    AT(2 * Km + 2 + j) = -acos(XX(Km - (j - 1)));
  }
  AT(Km + 1) = SPO_PI;
  AT(2 * Km + 2) = 0.0;
#undef AT
#undef XX
  for (int k = 1; k <= m_ph; ++k) a_phi[k - 1] = -2 * SPO_PI + d_phi * (k - 1);
  free(xx); free(ww);
}

/* Lagrange_Interpolation_Weights (:392-412): the 2p weights of x on the
 * nodes ext(K_arr+t+1-p .. K_arr+t+p) */
static void lagrange_w(double x, int t, int p, int K_arr, const double *ext, double *wk) {
  double xx[16];
  int kt = K_arr + t;
  for (int j = 1; j <= 2 * p; ++j) xx[j - 1] = ext[kt + 1 - p + (j - 1) - 1];
  for (int j = 1; j <= 2 * p; ++j) {
    wk[j - 1] = 1.0;
    for (int m = 1; m <= 2 * p; ++m)
      if (m != j) wk[j - 1] = (x - xx[m - 1]) / (xx[j - 1] - xx[m - 1]) * wk[j - 1];
  }
}

/* Sorting_theta_phi_LenReduced_Arr_New (:414-493): the K_m grid extended by
 * p rows beyond each pole (reflection: theta -> -theta, phi -> phi + pi,
 * both components negated), the two pole rows from the Cartesian pole
 * samples, and p columns beyond 0 and 2pi (periodic).  ff_in(i, j, q):
 * interior node i (ascending cos theta), phi index j, channel q; fxy[q2][4]:
 * (fx, fy) at theta = 0 then at theta = pi for medium q2. */
static void sort_ext(int p, int Km, const zc *ff_in, const zc fxy[2][4], zc *ff_out,
                     const double *a_phi, const double *a_theta) {
  int Kn = 2 * Km, n_phi = Kn + 2 * p + 1, n_theta = Km + 2 * p + 2;
#define FO(m, n, q) ff_out[((size_t)((q) - 1) * n_theta + ((m) - 1)) * n_phi + ((n) - 1)]
#define FI(i, j, q) ff_in[((size_t)((q) - 1) * Km + ((i) - 1)) * Kn + ((j) - 1)]
#define AP(i) a_phi[(i) - 1]
#define AT(i) a_theta[(i) - 1]
  for (size_t i = 0; i < (size_t)4 * n_theta * n_phi; ++i) ff_out[i] = 0.0;
  for (int m = 1; m <= n_phi; ++m) {
    int n = m + Kn - p;
    zc Mt[2] = {cos(AP(n)), sin(AP(n))};
    zc Mp[2] = {-sin(AP(n)), cos(AP(n))};
    for (int j = 1; j <= 2; ++j) {
      const zc *f0 = &fxy[j - 1][0], *f1 = &fxy[j - 1][2];
      FO(p + 1, m, j) = -(conj(Mt[0]) * f1[0] + conj(Mt[1]) * f1[1]);
      FO(Km + p + 2, m, j) = conj(Mt[0]) * f0[0] + conj(Mt[1]) * f0[1];
      FO(p + 1, m, j + 2) = conj(Mp[0]) * f1[0] + conj(Mp[1]) * f1[1];
      FO(Km + p + 2, m, j + 2) = conj(Mp[0]) * f0[0] + conj(Mp[1]) * f0[1];
    }
  }
  for (int q = 1; q <= 4; ++q)
    for (int i = 1; i <= Km; ++i)
      for (int j = 1; j <= Kn; ++j) FO(p + 1 + i, p + j, q) = FI(i, j, q);
  for (int m = 1; m <= n_theta; ++m) {
    double xt = AT(m - p + Km + 1);
    for (int n = 1; n <= n_phi; ++n) {
      double xp = AP(n - p + Kn);
      int src = 0, dn = 0;
      if (xt < 0.0) {
        src = 2 * (Km + p + 2) - m;
      } else if (xt >= SPO_PI) {
        src = 2 * (p + 1) - m;
      }
      if (src) {
        if (xp < SPO_PI && xp >= 0.0) dn = Kn / 2;
        else if (xp < 2 * SPO_PI && xp >= SPO_PI) dn = -Kn / 2;
        if (dn)
          for (int q = 1; q <= 4; ++q) FO(m, n, q) = -FO(src, n + dn, q);
      }
    }
  }
  for (int m = 1; m <= n_theta; ++m)
    for (int n = 1; n <= n_phi; ++n) {
      double xp = AP(n - p + Kn);
      if (xp >= 2 * SPO_PI) {
        for (int q = 1; q <= 4; ++q) FO(m, n, q) = FO(m, n - Kn, q);
      } else if (xp < 0) {
        for (int q = 1; q <= 4; ++q) FO(m, n, q) = FO(m, n + Kn, q);
      }
    }
#undef FO
#undef FI
#undef AP
#undef AT
}

/* Lagrange_Interpolation_New (:268-390) for one current: in / out hold the
 * four channels (theta k1, theta k2, phi k1, phi k2), each of 2K^2 + 2
 * samples (the last two: the Cartesian x / y components at theta = 0 and
 * pi, copied) as interleaved complex doubles. */
void spo_lagrange_interp(int p, int Km, int Kn, const double *in, double *out) {
  const zc *fin = (const zc *)in;
  zc *fout = (zc *)out;
  int Sm = 2 * Km * Km + 2, Sn = 2 * Kn * Kn + 2, Ns_m = 2 * Km * Km, Ns_n = 2 * Kn * Kn;
  int n_phi = 2 * Kn, m_phi = 2 * Km;
  zc *ff_in = xmalloc(sizeof(zc) * 4 * (size_t)Km * 2 * Km);
  for (int q = 0; q < 4; ++q)
    for (int i = 0; i < Km; ++i)
      for (int j = 0; j < 2 * Km; ++j) ff_in[((size_t)q * Km + i) * 2 * Km + j] = fin[(size_t)q * Sm + i * 2 * Km + j];
  /* fx_fy (:251-266): channel 1/2 hold x, 3/4 hold y at the poles */
  zc fxy[2][4];
  for (int j = 0; j < 2; ++j) {
    fxy[j][0] = fin[(size_t)j * Sm + Ns_m];         /* fx, theta = 0 */
    fxy[j][1] = fin[(size_t)(j + 2) * Sm + Ns_m];   /* fy, theta = 0 */
    fxy[j][2] = fin[(size_t)j * Sm + Ns_m + 1];     /* fx, theta = pi */
    fxy[j][3] = fin[(size_t)(j + 2) * Sm + Ns_m + 1];
  }
  double *a_phi = xmalloc(sizeof(double) * 6 * Km), *a_theta = xmalloc(sizeof(double) * (3 * Km + 2));
  theta_phi_ext(Km, a_phi, a_theta);
  int n_theta_e = Km + 2 * p + 2, n_phi_e = 2 * Km + 2 * p + 1;
  zc *ff_out = xmalloc(sizeof(zc) * 4 * (size_t)n_theta_e * n_phi_e);
  sort_ext(p, Km, ff_in, (const zc(*)[4])fxy, ff_out, a_phi, a_theta);
  double *xt = xmalloc(sizeof(double) * Kn), *xw = xmalloc(sizeof(double) * Kn);
  spo_gauss_legendre(Kn, xt, xw);
  double wkt[16], wkp[16];
  for (int it = 1; it <= Kn; ++it) {
    double theta = acos(xt[it - 1]);
    int t = (int)floorf((float)(it - 1.0f) / (float)Kn * (float)Km) + 1;
    int kt = p + t;
    lagrange_w(theta, t, p, Km, a_theta, wkt);
    for (int jp = 1; jp <= n_phi; ++jp) {
      int tt = (it - 1) * n_phi + jp;
      zc acc[4] = {0, 0, 0, 0};
      double phi = (jp - 1) * 2 * SPO_PI / n_phi;
      int s = (int)floorf((float)(jp - 1.0f) / (float)Kn * (float)Km) + 1;
      int kp = p + s;
      lagrange_w(phi, s, p, m_phi, a_phi, wkp);
      for (int i = 1; i <= 2 * p; ++i)
        for (int j = 1; j <= 2 * p; ++j) {
          double w = wkp[j - 1] * wkt[i - 1];
          int m = kt + i - p, n = kp + j - p;
          for (int q = 0; q < 4; ++q)
            acc[q] = acc[q] + w * ff_out[((size_t)q * n_theta_e + (m - 1)) * n_phi_e + (n - 1)];
        }
      for (int q = 0; q < 4; ++q) fout[(size_t)q * Sn + tt - 1] = acc[q];
    }
  }
  for (int q = 0; q < 4; ++q) {
    fout[(size_t)q * Sn + Ns_n] = fin[(size_t)q * Sm + Ns_m];
    fout[(size_t)q * Sn + Ns_n + 1] = fin[(size_t)q * Sm + Ns_m + 1];
  }
  free(ff_in); free(ff_out); free(a_phi); free(a_theta); free(xt); free(xw);
}

/* TL_km_ml_fmm: src/lib_sie_mlfmm.f90:129-188 (one medium). */
static void tl_eval(int K, zc kd, const double *r, const double *kh, int S, zc *out) {
  double rl = vlen(r);
  double rh[3] = {r[0] / rl, r[1] / rl, r[2] / rl};
  zc rhk = kd * rl;
  double *hl = xmalloc(sizeof(double) * 2 * K), *pm = xmalloc(sizeof(double) * K);
  spo_hankel2_sph(K, creal(rhk), cimag(rhk), hl);
  for (int i = 0; i < S; ++i) out[i] = 0.0;
  for (int i = 0; i < S; ++i) {
    double x = vdot(rh, &kh[3 * i]);
    spo_legendre_p(K, x, pm);
    zc acc = 0.0;
    zc mi = 1.0;
    for (int j = 1; j <= K; ++j) {
      mi = mi * ZCMPLX(-0.0, -1.0); /* (-im)**j */
      zc h = ZCMPLX(hl[2 * (j - 1)], hl[2 * (j - 1) + 1]);
      acc = acc + h * mi * (double)(2 * (j - 1) + 1) * pm[j - 1] * kd / ((4 * SPO_PI) * (4 * SPO_PI));
    }
    out[i] = acc;
  }
  free(hl); free(pm);
}

void spo_tl(int K, double kre, double kim, const double *r, double *out) {
  int S = 2 * K * K + 2;
  double *kh = xmalloc(sizeof(double) * 3 * S), *a = xmalloc(sizeof(double) * 3 * S),
         *b = xmalloc(sizeof(double) * 3 * S), *w = xmalloc(sizeof(double) * S);
  khat_trf(K, kh, a, b, w);
  zc *o = xmalloc(sizeof(zc) * S);
  tl_eval(K, ZCMPLX(kre, kim), r, kh, S, o);
  for (int i = 0; i < S; ++i) { out[2 * i] = creal(o[i]); out[2 * i + 1] = cimag(o[i]); }
  free(kh); free(a); free(b); free(w); free(o);
}

/* box centre (lib_tree_hf_get_centre_of_box :789-814) unscaled
 * (lib_tree_get_unscaled_point, lib_tree.f90:474-487) */
static void box_centre(const spo_ctx *c, int64_t n, int l, double *x) {
  int64_t ix[3];
  demorton3((uint64_t)n, l, ix);
  for (int d = 0; d < 3; ++d) {
    double s = ldexp(1.0, -l) * ((double)ix[d] + 0.5);
    x[d] = s * c->D + c->xmin[d];
  }
}

/* PreCalculate_RT + RT_function_SCSArr_New:
 * src/lib_sie_tri_mlfmm.f90:503-583.  Rc = leaf box centre of edge m. */
static void patterns(spo_ctx *c, int nthreads) {
  int K = c->p.K, S = c->S, N = c->N, ng = c->p.ngp;
  double qa[20], qb[20], qw[20];
  spo_quadrature_tri(ng, qa, qb, qw);
  free(c->rt);
  c->rt = xmalloc(sizeof(zc) * (size_t)N * 4 * S);
  (void)K;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
  for (int m = 0; m < N; ++m) {
    double Rc[3];
    box_centre(c, c->leaf[m], c->l_max, Rc);
    zc *o = &c->rt[(size_t)m * 4 * S];
    for (int ll = 1; ll <= 2; ++ll) {
      const fnrwg *f = &c->fn[2 * m + ll - 1];
      double dfm = f->sign * f->length;
      for (int i = 0; i < S; ++i) {
        zc sa[3] = {0, 0, 0}, sb[3] = {0, 0, 0};
        for (int j = 0; j < ng; ++j) {
          double Ro[3], fm[3], Roc[3];
          simplex(qa[j], qb[j], f->c, Ro);
          for (int k = 0; k < 3; ++k) { fm[k] = 0.5 * (Ro[k] - f->c[2][k]) * dfm; Roc[k] = Ro[k] - Rc[k]; }
          double tmp = vdot(&c->khat[3 * i], Roc);
          zc ff1 = zexp_mikr(c->k1, tmp) * qw[j];
          zc ff2 = zexp_mikr(c->k2, tmp) * qw[j];
          for (int k = 0; k < 3; ++k) { sa[k] = sa[k] + fm[k] * ff1; sb[k] = sb[k] + fm[k] * ff2; }
        }
        /* [k][comp]: 0 = (k1, theta), 1 = (k1, phi), 2 = (k2, theta), 3 = (k2, phi) */
        o[0 * S + i] += zdot_rc(&c->ttheta[3 * i], sa);
        o[1 * S + i] += zdot_rc(&c->tphi[3 * i], sa);
        o[2 * S + i] += zdot_rc(&c->ttheta[3 * i], sb);
        o[3 * S + i] += zdot_rc(&c->tphi[3 * i], sb);
      }
    }
  }
}

/* precalculation_LT3_TL + PreCalc_transfer_function:
 * src/lib_sie_mlfmm.f90:27-125, grid of lib_sie_lookup_table.f90:204-213,
 * 631-656.  Table slot = (ix*7 + iy)*7 + iz, offsets -3..3 per axis. */
static void tl_tables(spo_ctx *c, int nthreads) {
  c->tl = xmalloc(sizeof(zc *) * (c->l_max + 1));
  for (int l = 0; l <= c->l_max; ++l) c->tl[l] = NULL;
  int np_ = 7;
  for (int l = c->l_min; l <= c->l_max; ++l) {
    if (l < 1) continue;
    const int K = c->Kl[l], S = c->Sl[l];
    const double *khat = c->khl[l];
    c->tl[l] = xmalloc(sizeof(zc) * 343 * 2 * S);
    double d_c = (1.0 / (double)(1LL << l)) * c->D;
    double x_min = -((np_ - 1) / 2 * d_c);
    double x_max = -x_min;
    double dx = (x_max - x_min) / (np_ - 1);
    int lossy_skip = (fabs(cimag(c->k2)) > 0.0) && (l < c->l_max);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int t = 0; t < 343; ++t) {
      int i = t / 49, j = (t / 7) % 7, k = t % 7;
      zc *o = &c->tl[l][(size_t)t * 2 * S];
      if (i >= 2 && i <= 4 && j >= 2 && j <= 4 && k >= 2 && k <= 4) continue; /* E2 cube */
      double r[3] = {x_min + i * dx, x_min + j * dx, x_min + k * dx};
      tl_eval(K, c->k1, r, khat, S, o);
      if (!lossy_skip) tl_eval(K, c->k2, r, khat, S, o + S);
    }
  }
  (void)nthreads;
}

/* Per-level truncation numbers for the next spo_mlfmm_setup: K[l] for
 * levels l < n (0: the default p.K). */
void spo_set_level_K(spo_ctx *c, int n, const int *K, int rr_adjoint) {
  c->nKreq = n < 32 ? n : 32;
  for (int l = 0; l < c->nKreq; ++l) c->Kreq[l] = K[l];
  c->rr_adjoint = rr_adjoint;
}

/* The interpolation from the K_m to the K_n grid as a sparse matrix per
 * medium: its columns are spo_lagrange_interp of unit samples (channel
 * theta or phi of one medium, one sample), so rows / cols index
 * (component * S + sample).  The transpose is the anterpolation. */
typedef struct { int nnz; int *row, *col; double *val; } spmat;
static void interp_matrix(int Km, int Kn, spmat *M) {
  int Sm = 2 * Km * Km + 2, Sn = 2 * Kn * Kn + 2;
  double *in = xmalloc(sizeof(double) * 2 * 4 * Sm), *out = xmalloc(sizeof(double) * 2 * 4 * Sn);
  int cap = 1 << 16;
  M->nnz = 0;
  M->row = xmalloc(sizeof(int) * cap); M->col = xmalloc(sizeof(int) * cap);
  M->val = xmalloc(sizeof(double) * cap);
  memset(in, 0, sizeof(double) * 2 * 4 * Sm);
  for (int cin = 0; cin < 2; ++cin)
    for (int si = 0; si < Sm; ++si) {
      int q = cin * 2; /* channel (theta or phi) of medium 1 */
      in[2 * ((size_t)q * Sm + si)] = 1.0;
      spo_lagrange_interp(3, Km, Kn, in, out);
      in[2 * ((size_t)q * Sm + si)] = 0.0;
      for (int cout = 0; cout < 2; ++cout)
        for (int so = 0; so < Sn; ++so) {
          double v = out[2 * ((size_t)(cout * 2) * Sn + so)];
          if (v == 0.0) continue;
          if (M->nnz == cap) {
            cap *= 2;
            M->row = realloc(M->row, sizeof(int) * cap); M->col = realloc(M->col, sizeof(int) * cap);
            M->val = realloc(M->val, sizeof(double) * cap);
          }
          M->row[M->nnz] = cout * Sn + so; M->col[M->nnz] = cin * Sm + si; M->val[M->nnz] = v;
          ++M->nnz;
        }
    }
  free(in); free(out);
}

/* the same on folded coefficients W (items k*4 + {1a, 1b, 2a, 2b}: the
 * (theta, phi) pairs (k*4, k*4+1) and (k*4+2, k*4+3) of medium k) */
static void anterp_w(const spmat *M, int Km, int Kn, const zc *x, zc *y) {
  int Sm = 2 * Km * Km + 2, Sn = 2 * Kn * Kn + 2;
  for (int k = 0; k < 2; ++k)
    for (int g = 0; g < 2; ++g) {
      const int it_c[2] = {k * 4 + 2 * g, k * 4 + 2 * g + 1};
      for (int e = 0; e < M->nnz; ++e) {
        int co = M->row[e] / Sn, so = M->row[e] % Sn, ci = M->col[e] / Sm, si = M->col[e] % Sm;
        y[(size_t)it_c[ci] * Sm + si] += M->val[e] * x[(size_t)it_c[co] * Sn + so];
      }
    }
}

/* y[items] (+)= M^T x[items] on the 8 coefficient items of a box, for the
 * two currents and two media; M maps the K_m grid to the K_n grid, so the
 * transpose maps K_n samples (x) to K_m samples (y) */
static void anterp_items(const spmat *M, int Km, int Kn, const zc *x, zc *y) {
  int Sm = 2 * Km * Km + 2, Sn = 2 * Kn * Kn + 2;
  for (int cur = 0; cur < 2; ++cur)
    for (int k = 0; k < 2; ++k) {
      const int it_c[2] = {0 * 4 + cur * 2 + k, 1 * 4 + cur * 2 + k};
      for (int e = 0; e < M->nnz; ++e) {
        int co = M->row[e] / Sn, so = M->row[e] % Sn, ci = M->col[e] / Sm, si = M->col[e] % Sm;
        y[(size_t)it_c[ci] * Sm + si] += M->val[e] * x[(size_t)it_c[co] * Sn + so];
      }
    }
}

int spo_mlfmm_setup_levels(spo_ctx *c, int nthreads);

void spo_mlfmm_setup(spo_ctx *c, int nthreads) { (void)spo_mlfmm_setup_levels(c, nthreads); }

/* returns 0, or -1 if a level's K exceeds its parent's (the reference's SS
 * and RR interpolate only from a smaller to a larger K going up) */
int spo_mlfmm_setup_levels(spo_ctx *c, int nthreads) {
  if (!c->tree_built) spo_tree_build(c);
  for (int l = 0; l <= c->l_max && l < 32; ++l) {
    c->Kl[l] = (l < c->nKreq && c->Kreq[l] > 0) ? c->Kreq[l] : c->p.K;
    c->Sl[l] = 2 * c->Kl[l] * c->Kl[l] + 2;
  }
  for (int l = c->l_min; l < c->l_max; ++l)
    if (c->Kl[l] < c->Kl[l + 1]) return -1;
  for (int l = c->l_min; l <= c->l_max; ++l) {
    free(c->khl[l]); free(c->trfl[l]);
    const int S = c->Sl[l];
    c->khl[l] = xmalloc(sizeof(double) * 3 * S);
    c->trfl[l] = xmalloc(sizeof(double) * S);
    double *tth = xmalloc(sizeof(double) * 3 * S), *tph = xmalloc(sizeof(double) * 3 * S);
    khat_trf(c->Kl[l], c->khl[l], tth, tph, c->trfl[l]);
    free(tth); free(tph);
  }
  int K = c->Kl[c->l_max];
  c->S = 2 * K * K + 2;
  int S = c->S;
  free(c->khat); free(c->ttheta); free(c->tphi); free(c->trf);
  c->khat = xmalloc(sizeof(double) * 3 * S);
  c->ttheta = xmalloc(sizeof(double) * 3 * S);
  c->tphi = xmalloc(sizeof(double) * 3 * S);
  c->trf = xmalloc(sizeof(double) * S);
  khat_trf(K, c->khat, c->ttheta, c->tphi, c->trf);
  patterns(c, nthreads);
  tl_tables(c, nthreads);
  c->mlfmm_ready = 1;
  return 0;
}

/* Lagrange_Interpolation_New on the two currents of a box's 8 coefficient
 * items (it = comp*4 + cur*2 + k) from the K_m to the K_n grid */
static void interp_items(int Km, int Kn, const zc *in, zc *out) {
  int Sm = 2 * Km * Km + 2, Sn = 2 * Kn * Kn + 2;
  zc *a = xmalloc(sizeof(zc) * 4 * Sm), *b = xmalloc(sizeof(zc) * 4 * Sn);
  for (int cur = 0; cur < 2; ++cur) {
    const int items[4] = {0 * 4 + cur * 2 + 0, 0 * 4 + cur * 2 + 1, 1 * 4 + cur * 2 + 0,
                          1 * 4 + cur * 2 + 1};
    for (int q = 0; q < 4; ++q) memcpy(&a[(size_t)q * Sm], &in[(size_t)items[q] * Sm], sizeof(zc) * Sm);
    spo_lagrange_interp(3, Km, Kn, (const double *)a, (double *)b);
    for (int q = 0; q < 4; ++q) memcpy(&out[(size_t)items[q] * Sn], &b[(size_t)q * Sn], sizeof(zc) * Sn);
  }
  free(a); free(b);
}

/* ------------------------------------------------------------------------ */
/* Matvec: lib_ml_fmm_run (src/lib/libmlfmm/src/lib/lib_ml_fmm.f90:269-1176) */
/* with the SpeckleSim plug-ins (src/lib_sie_tri_mlfmm_interface.f90:395-777) */
/* coefficient item it = comp*4 + cur*2 + k  (comp a/b, cur xe/xh, medium k) */
/* ------------------------------------------------------------------------ */
void spo_matvec(spo_ctx *c, const double *xin, double *yout, int nthreads) {
  int N = c->N, S = c->S, lmax = c->l_max, lmin = c->l_min;
  const zc *x = (const zc *)xin;
  zc *y = (zc *)yout;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  zc **B = xmalloc(sizeof(zc *) * (lmax + 1));
  zc **Dt = xmalloc(sizeof(zc *) * (lmax + 1));
  for (int l = lmin; l <= lmax; ++l) {
    B[l] = xmalloc(sizeof(zc) * (size_t)c->lev[l].nbox * 8 * c->Sl[l]);
    Dt[l] = xmalloc(sizeof(zc) * (size_t)c->lev[l].nbox * 8 * c->Sl[l]);
  }
  const level_t *LL = &c->lev[lmax];
  /* upward pass step 1: lib_sie_ml_fmm_get_c */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int b = 0; b < LL->nbox; ++b) {
    zc *C = &B[lmax][(size_t)b * 8 * S];
    for (int e = 0; e < LL->count[b]; ++e) {
      int m = c->box_elems[LL->first[b] + e];
      zc xe = x[m], xh = x[m + N];
      const zc *rt = &c->rt[(size_t)m * 4 * S];
      for (int k = 0; k < 2; ++k)
        for (int i = 0; i < S; ++i) {
          zc fa = conj(rt[(k * 2 + 0) * S + i]), fb = conj(rt[(k * 2 + 1) * S + i]);
          C[(0 * 4 + 0 * 2 + k) * S + i] = C[(0 * 4 + 0 * 2 + k) * S + i] + fa * xe;
          C[(1 * 4 + 0 * 2 + k) * S + i] = C[(1 * 4 + 0 * 2 + k) * S + i] + fb * xe;
          C[(0 * 4 + 1 * 2 + k) * S + i] = C[(0 * 4 + 1 * 2 + k) * S + i] + fa * xh;
          C[(1 * 4 + 1 * 2 + k) * S + i] = C[(1 * 4 + 1 * 2 + k) * S + i] + fb * xh;
        }
    }
  }
  /* upward pass step 2: translation SS (:466-528): a child of a smaller K
   * is interpolated to the parent's grid first (Lagrange_Interpolation_New,
   * stencil 2p = 6), then shifted with the parent grid's directions */
  for (int l = lmax - 1; l >= lmin; --l) {
    const level_t *L = &c->lev[l], *Lc = &c->lev[l + 1];
    const int Sp = c->Sl[l], Sc = c->Sl[l + 1];
    const int interp = c->Kl[l] > c->Kl[l + 1];
    const double *khp = c->khl[l];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int b = 0; b < L->nbox; ++b) {
      double xp[3];
      box_centre(c, L->n[b], l, xp);
      zc *P = &B[l][(size_t)b * 8 * Sp];
      zc *Ci = interp ? xmalloc(sizeof(zc) * 8 * Sp) : NULL;
      for (int o = 0; o < 8; ++o) {
        int64_t cn = (L->n[b] << 3) | o;
        int cp = level_find(Lc, cn);
        if (cp < 0) continue;
        double xc[3], r[3];
        box_centre(c, cn, l + 1, xc);
        vsub(xc, xp, r);
        const zc *Cc = &B[l + 1][(size_t)cp * 8 * Sc];
        if (interp) {
          interp_items(c->Kl[l + 1], c->Kl[l], Cc, Ci);
          Cc = Ci;
        }
        const int S = Sp;
        for (int i = 0; i < S; ++i) {
          double tmp = vdot(&khp[3 * i], r);
          zc t[2] = {cexp(ZCMPLX(-cimag(c->k1) * tmp, creal(c->k1) * tmp)),
                     cexp(ZCMPLX(-cimag(c->k2) * tmp, creal(c->k2) * tmp))};
          for (int it = 0; it < 8; ++it) P[it * S + i] = P[it * S + i] + Cc[it * S + i] * t[it & 1];
        }
      }
      free(Ci);
    }
  }
  /* downward pass step 1: translation SR over I4 (:547-586) */
  for (int l = lmin; l <= lmax; ++l) {
    const level_t *L = &c->lev[l];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int a = 0; a < L->nbox; ++a) {
      int64_t lst[216];
      int n4 = spo_i4(c, l, L->n[a], lst, 216);
      int64_t xa[3];
      demorton3((uint64_t)L->n[a], l, xa);
      const int S = c->Sl[l];
      const double *trf = c->trfl[l];
      zc *Da = &Dt[l][(size_t)a * 8 * S];
      for (int q = 0; q < n4; ++q) {
        int bp = level_find(L, lst[q]);
        if (bp < 0) continue;
        int64_t xb[3];
        demorton3((uint64_t)lst[q], l, xb);
        int slot = (int)((xa[0] - xb[0] + 3) * 49 + (xa[1] - xb[1] + 3) * 7 + (xa[2] - xb[2] + 3));
        const zc *T = &c->tl[l][(size_t)slot * 2 * S];
        const zc *Bb = &B[l][(size_t)bp * 8 * S];
        for (int it = 0; it < 8; ++it) {
          const zc *Tk = T + (it & 1) * S;
          for (int i = 0; i < S; ++i) Da[it * S + i] = Da[it * S + i] + Bb[it * S + i] * trf[i] * Tk[i];
        }
      }
    }
  }
  /* downward pass step 2: translation RR (:600-655): to a child of a
   * smaller K the parent's coefficients are shifted on the parent's grid,
   * then interpolated to the child's grid (Lagrange_Interpolation_New), or
   * (rr_adjoint) mapped by the transpose of SS's child-to-parent
   * interpolation */
  spmat Mup[32];
  memset(Mup, 0, sizeof Mup);
  int multik = 0;
  for (int l = lmin + 1; l <= lmax; ++l) multik |= c->Kl[l] != c->Kl[l - 1];
  const int wspace = multik && c->rr_adjoint;
  zc eps_d[2] = {c->eps1 * I * c->omega, c->eps2 * I * c->omega};
  zc my_d[2] = {c->my1 * I * c->omega, c->my2 * I * c->omega};
  zc kdv[2] = {c->k1 * I, c->k2 * I};
  if (wspace) {
    /* anterpolation: every level's M2L sums are folded first (the per-sample
     * map of MLFMM_tri_summation_New, src/lib_sie_tri_mlfmm.f90:416-497),
     * then shifted down and mapped by the transpose of SS's interpolation,
     * so that the final sum over the leaf grid equals the sum over each
     * coarser grid with the receiving patterns interpolated up to it */
    for (int l = lmin; l <= lmax; ++l) {
      const int Sl = c->Sl[l];
      for (int b = 0; b < c->lev[l].nbox; ++b) {
        zc *Db = &Dt[l][(size_t)b * 8 * Sl];
        for (int i = 0; i < Sl; ++i) {
          zc w[8];
          for (int k = 0; k < 2; ++k) {
            zc Dae = Db[(0 * 4 + 0 * 2 + k) * Sl + i], Dah = Db[(0 * 4 + 1 * 2 + k) * Sl + i];
            zc Dbe = Db[(1 * 4 + 0 * 2 + k) * Sl + i], Dbh = Db[(1 * 4 + 1 * 2 + k) * Sl + i];
            w[k * 4 + 0] = Dbh * kdv[k] + Dae * my_d[k];
            w[k * 4 + 1] = -(Dah * kdv[k]) + Dbe * my_d[k];
            w[k * 4 + 2] = -(Dbe * kdv[k]) + Dah * eps_d[k];
            w[k * 4 + 3] = Dae * kdv[k] + Dbh * eps_d[k];
          }
          for (int it = 0; it < 8; ++it) Db[it * Sl + i] = w[it];
        }
      }
    }
    for (int l = lmin + 1; l <= lmax; ++l)
      if (c->Kl[l] < c->Kl[l - 1]) interp_matrix(c->Kl[l], c->Kl[l - 1], &Mup[l]);
  }
  for (int l = lmin + 1; l <= lmax; ++l) {
    const level_t *L = &c->lev[l], *Lp = &c->lev[l - 1];
    const int Sc = c->Sl[l], Sp = c->Sl[l - 1];
    const int interp = c->Kl[l] < c->Kl[l - 1];
    const double *khp = c->khl[l - 1];
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int b = 0; b < L->nbox; ++b) {
      int pp = level_find(Lp, L->n[b] >> 3);
      double xc[3], xp[3], r[3];
      box_centre(c, L->n[b], l, xc);
      box_centre(c, L->n[b] >> 3, l - 1, xp);
      vsub(xc, xp, r);
      zc *Dc = &Dt[l][(size_t)b * 8 * Sc];
      const zc *Dp = &Dt[l - 1][(size_t)pp * 8 * Sp];
      if (interp) {
        zc *A = xmalloc(sizeof(zc) * 8 * Sp), *Ai = xmalloc(sizeof(zc) * 8 * Sc);
        for (int i = 0; i < Sp; ++i) {
          double tmp = vdot(&khp[3 * i], r);
          zc t[2] = {zexp_mikr(c->k1, tmp), zexp_mikr(c->k2, tmp)};
          for (int it = 0; it < 8; ++it) A[it * Sp + i] = Dp[it * Sp + i] * t[wspace ? it >> 2 : it & 1];
        }
        if (wspace) {
          for (size_t i = 0; i < (size_t)8 * Sc; ++i) Ai[i] = 0.0;
          anterp_w(&Mup[l], c->Kl[l], c->Kl[l - 1], A, Ai);
        } else {
          interp_items(c->Kl[l - 1], c->Kl[l], A, Ai);
        }
        for (size_t i = 0; i < (size_t)8 * Sc; ++i) Dc[i] = Dc[i] + Ai[i];
        free(A); free(Ai);
        continue;
      }
      const int S = Sc;
      for (int i = 0; i < S; ++i) {
        double tmp = vdot(&khp[3 * i], r);
        zc t[2] = {zexp_mikr(c->k1, tmp), zexp_mikr(c->k2, tmp)};
        for (int it = 0; it < 8; ++it)
          Dc[it * S + i] = Dc[it * S + i] + Dp[it * S + i] * t[wspace ? it >> 2 : it & 1];
      }
    }
  }
  for (int l = 0; l < 32; ++l) { free(Mup[l].row); free(Mup[l].col); free(Mup[l].val); }
  /* final summation: lib_sie_ml_fmm_get_v_y_j (:690-751) with
   * MLFMM_tri_summation_New (src/lib_sie_tri_mlfmm.f90:416-497) and the
   * near-field CSR products (:755-777). */
  zc a_lc[2], d_lc[2];
  form_coef(c, a_lc, d_lc);
  int form = c->p.formulation;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 16)
#endif
  for (int j = 0; j < N; ++j) {
    int b = c->leaf_pos[j];
    const zc *Db = &Dt[lmax][(size_t)b * 8 * S];
    const zc *rt = &c->rt[(size_t)j * 4 * S];
    zc s1 = 0.0, s2 = 0.0;
    for (int k = 0; k < 2 && wspace; ++k) {
      const zc *fa = rt + (k * 2 + 0) * S, *fb = rt + (k * 2 + 1) * S;
      const zc *W = Db + (size_t)k * 4 * S;
      for (int i = 0; i < S; ++i) {
        s1 = s1 + (fa[i] * W[0 * S + i] + fb[i] * W[1 * S + i]);
        s2 = s2 + (fa[i] * W[2 * S + i] + fb[i] * W[3 * S + i]);
      }
    }
    for (int k = 0; k < 2 && !wspace; ++k) {
      const zc *fa = rt + (k * 2 + 0) * S, *fb = rt + (k * 2 + 1) * S;
      const zc *Dae = Db + (0 * 4 + 0 * 2 + k) * S, *Dbe = Db + (1 * 4 + 0 * 2 + k) * S;
      const zc *Dah = Db + (0 * 4 + 1 * 2 + k) * S, *Dbh = Db + (1 * 4 + 1 * 2 + k) * S;
      zc sKH = 0.0, sLE = 0.0, sKE = 0.0, sLH = 0.0;
      for (int i = 0; i < S; ++i) {
        zc rka = -fb[i], rkb = fa[i];
        sKH = sKH + (rka * Dah[i] * kdv[k] + rkb * Dbh[i] * kdv[k]);
        sLE = sLE + (fa[i] * Dae[i] * my_d[k] + fb[i] * Dbe[i] * my_d[k]);
        sKE = sKE + (-rka * Dae[i] * kdv[k] + -rkb * Dbe[i] * kdv[k]);
        sLH = sLH + (fa[i] * Dah[i] * eps_d[k] + fb[i] * Dbh[i] * eps_d[k]);
      }
      s1 = s1 + sKH + sLE;
      s2 = s2 + sKE + sLH;
    }
    switch (form) {
      case SPO_PMCHWT_S: s2 = s2 * c->eta1; break;
      case SPO_MCTF: s2 = s2 * c->eta1 * c->eta2; break;
      case SPO_ICTF: s1 = s1 / c->eta_a; s2 = s2 * c->eta_a; break;
      case SPO_MCTF2: s1 = s1 * a_lc[0] / c->eta1; s2 = s2 * d_lc[0] * c->eta1; break;
      default: break;
    }
    /* near: E2 order (box order) */
    zc n1 = 0.0, n2 = 0.0;
    int64_t boxes[27];
    int nb = e2_boxes(c->leaf[j], lmax, boxes);
    for (int q = 0; q < nb; ++q) {
      int p = level_find(LL, boxes[q]);
      if (p < 0) continue;
      for (int e = 0; e < LL->count[p]; ++e) {
        int i = c->box_elems[LL->first[p] + e];
        /* find_elements_in_field_zcsr: column i of row j (binary search in
         * the sorted row gives the same entry as the linear search). */
        int lo = c->ia[j], hi = c->ia[j + 1] - 1, pos = -1;
        while (lo <= hi) {
          int md = (lo + hi) / 2;
          if (c->ja[md] == i) { pos = md; break; }
          if (c->ja[md] < i) lo = md + 1; else hi = md - 1;
        }
        const zc *Z = &c->za[4 * (int64_t)pos];
        n1 = n1 + (Z[0] * x[i] + Z[1] * x[i + N]);
        n2 = n2 + (Z[2] * x[i] + Z[3] * x[i + N]);
      }
    }
    y[j] = n1 + s1;
    y[j + N] = n2 + s2;
  }
  for (int l = lmin; l <= lmax; ++l) { free(B[l]); free(Dt[l]); }
  free(B); free(Dt);
}

/* ------------------------------------------------------------------------ */
/* Plane-wave RHS: get_vector_b_tri + incident_field_tri_general            */
/* (src/lib_sie_tri_calculation_mod.f90:391-460, 939-1002)                  */
/* ------------------------------------------------------------------------ */
void spo_rhs_plane(const spo_ctx *c, const double *kin, const double *ein, double *bout) {
  int N = c->N, ng = c->p.ngp, form = c->p.formulation;
  double qa[20], qb[20], qw[20];
  spo_quadrature_tri(ng, qa, qb, qw);
  zc *b = (zc *)bout;
  zc a_lc[2], d_lc[2];
  form_coef(c, a_lc, d_lc);
  for (int m = 0; m < N; ++m) {
    zc S4[4] = {0, 0, 0, 0};
    for (int ot = 1; ot <= 2; ++ot) {
      const fnrwg *f = &c->fn[2 * m + ot - 1];
      for (int i = 0; i < ng; ++i) {
        double Ra[3], rfm[3];
        simplex(qa[i], qb[i], f->c, Ra);
        for (int k = 0; k < 3; ++k) rfm[k] = (Ra[k] - f->c[2][k]) * f->df * 0.5;
        /* dot_product(k1*k_in, R_a) (:977): Fortran conjugates the complex
         * first argument, so the phase is conj(k1) k_in.R */
        const zc k1c = conj(c->k1);
        zc kk[3] = {k1c * kin[0], k1c * kin[1], k1c * kin[2]};
        zc ph = kk[0] * Ra[0] + kk[1] * Ra[1] + kk[2] * Ra[2];
        zc e = cexp(-I * ph);
        zc E[3] = {ein[0] * e, ein[1] * e, ein[2] * e};
        zc kh[3] = {kin[0], kin[1], kin[2]};
        zc H[3] = {(kh[1] * E[2] - kh[2] * E[1]) / (c->my1 * C0_F),
                   (kh[2] * E[0] - kh[0] * E[2]) / (c->my1 * C0_F),
                   (kh[0] * E[1] - kh[1] * E[0]) / (c->my1 * C0_F)};
        zc nE[3], nH[3];
        vcross_rc(f->nq, E, nE);
        vcross_rc(f->nq, H, nH);
        S4[0] = S4[0] + zdot_rc(rfm, E) * qw[i] * f->area;
        S4[1] = S4[1] + zdot_rc(rfm, H) * qw[i] * f->area;
        S4[2] = S4[2] + zdot_rc(rfm, nE) * qw[i] * f->area;
        S4[3] = S4[3] + zdot_rc(rfm, nH) * qw[i] * f->area;
      }
    }
    switch (form) {
      case SPO_PMCHWT: b[m] = S4[0]; b[m + N] = S4[1]; break;
      case SPO_PMCHWT_S: b[m] = S4[0]; b[m + N] = S4[1] * c->eta1; break;
      case SPO_MCTF: b[m] = S4[0]; b[m + N] = S4[1] * c->eta1 * c->eta2; break;
      case SPO_JMCFIE: b[m] = 1.0 / c->eta1 * S4[0] + S4[3]; b[m + N] = -S4[2] + c->eta1 * S4[1]; break;
      case SPO_MCTF2: b[m] = a_lc[0] / c->eta1 * S4[0]; b[m + N] = d_lc[0] * c->eta1 * S4[1]; break;
      case SPO_ICTF: b[m] = S4[0] / c->eta_a; b[m + N] = S4[1] * c->eta_a; break;
      case SPO_CNF: b[m] = S4[3]; b[m + N] = -S4[2]; break;
      case SPO_MNMF:
        b[m] = S4[3] * c->my1 / (c->my1 + c->my2);
        b[m + N] = -S4[2] * c->eps1 / (c->eps1 + c->eps2);
        break;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Scattered field at observation points: Integration_scattered_field_tri    */
/* (src/lib_sie_tri_calculation_mod.f90:2782-2825).  The rule is the 3-point */
/* one (ngp = 3 at every call site, lib_sie_tri_solver_normal.f90:331-350);  */
/* the charge term of E is multiplied by 0 in the reference (:2811) and is   */
/* left out here.                                                            */
/* ------------------------------------------------------------------------ */
void spo_scattered_field(const spo_ctx *c, int npts, const double *r, const double *xin,
                         double *Eout, double *Hout) {
  const int N = c->N, ng = 3;
  double qa[20], qb[20], qw[20];
  spo_quadrature_tri(ng, qa, qb, qw);
  const zc *x = (const zc *)xin;
  zc *E = (zc *)Eout, *H = (zc *)Hout;
  const zc iwmy = I * c->omega * c->my1, iweps = I * c->omega * c->eps1;
  const zc inv_iwmy = 1.0 / (I * c->omega * c->my1);
  for (int p = 0; p < npts; ++p) {
    const double *ra = r + 3 * p;
    zc SE[3] = {0, 0, 0}, SH[3] = {0, 0, 0};
    for (int m = 0; m < N; ++m) {
      const zc J = x[m], M = x[m + N];
      for (int t = 0; t < 2; ++t) {
        const fnrwg *f = &c->fn[2 * m + t];
        const double dfm = f->sign * f->length;
        for (int i = 0; i < ng; ++i) {
          double rq[3], fm[3], d[3];
          simplex(qa[i], qb[i], f->c, rq);
          for (int k = 0; k < 3; ++k) fm[k] = 0.5 * dfm * (rq[k] - f->c[2][k]);
          for (int k = 0; k < 3; ++k) d[k] = ra[k] - rq[k];
          const double R = vlen(d);
          zc G, DG;
          diver_green(c->k1, R, &G, &DG);
          zc gG[3], Mf[3], cE[3], fxg[3];
          for (int k = 0; k < 3; ++k) gG[k] = d[k] / R * DG;
          for (int k = 0; k < 3; ++k) Mf[k] = M * fm[k];
          cE[0] = Mf[1] * gG[2] - Mf[2] * gG[1];
          cE[1] = Mf[2] * gG[0] - Mf[0] * gG[2];
          cE[2] = Mf[0] * gG[1] - Mf[1] * gG[0];
          fxg[0] = fm[1] * gG[2] - fm[2] * gG[1];
          fxg[1] = fm[2] * gG[0] - fm[0] * gG[2];
          fxg[2] = fm[0] * gG[1] - fm[1] * gG[0];
          for (int k = 0; k < 3; ++k) {
            const zc ke = iwmy * J * fm[k] * G - cE[k];
            SE[k] = SE[k] - ke * qw[i];
            const zc kh = iweps * M * fm[k] * G - inv_iwmy * M * dfm * gG[k] + J * fxg[k];
            SH[k] = SH[k] - kh * qw[i];
          }
        }
      }
    }
    for (int k = 0; k < 3; ++k) {
      E[3 * p + k] = SE[k];
      H[3 * p + k] = SH[k];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* CERFACS GMRES (src/lib/libmath_external/src/algorithm_842/zPackgmres.f90 */
/* :537-1565), straight-line form of the reverse-communication loop, with   */
/* the netlib zrotg/zrot of algorithm_842/blas.f:266-380.                   */
/* ------------------------------------------------------------------------ */
typedef void (*spo_mv_fn)(void *ctx, const zc *x, zc *y);

static zc zdotc(int n, const zc *x, const zc *y) {
  zc s = 0.0;
  for (int i = 0; i < n; ++i) s = s + conj(x[i]) * y[i];
  return s;
}
static void zrotg_(zc *ca, zc cb, double *cc, zc *s) {
  if (cabs(*ca) == 0.0) { *cc = 0.0; *s = 1.0; *ca = cb; return; }
  double scale = cabs(*ca) + cabs(cb);
  double a1 = cabs(*ca / ZCMPLX(scale, 0.0)), b1 = cabs(cb / ZCMPLX(scale, 0.0));
  double norm = scale * sqrt(a1 * a1 + b1 * b1);
  zc alpha = *ca / cabs(*ca);
  *cc = cabs(*ca) / norm;
  *s = alpha * conj(cb) / norm;
  *ca = alpha * norm;
}
static void zrot1(zc *x, zc *y, double cc, zc s) {
  zc t = cc * *x + s * *y;
  *y = cc * *y - conj(s) * *x;
  *x = t;
}

/* Left preconditioning by a diagonal (CERFACS icntl(4) = 1 with the
 * reference's lib_sie_ml_fmm_preconditioner_left, src/lib_sie_tri_mlfmm_interface.f90:321-339):
 * y = x / diag.  diag == NULL: no preconditioning. */
static void precond_left(int n, const zc *diag, const zc *x, zc *y) {
  if (!diag) {
    memcpy(y, x, sizeof(zc) * n);
    return;
  }
  for (int i = 0; i < n; ++i) y[i] = x[i] / diag[i];
}

static int gmres_core(int n, spo_mv_fn mv, void *ctx, const zc *diag, const zc *b, zc *x,
                      int iterMax, int m, double tol, int orth, double *hist, int *iters,
                      double *be_out, double *tbe_out) {
  zc *V = xmalloc(sizeof(zc) * (size_t)n * (m + 1));
  zc *H = xmalloc(sizeof(zc) * (size_t)(m + 1) * (m + 1)); /* H(i,j) = H[(j)*(m+1)+i] */
  zc *w = xmalloc(sizeof(zc) * n), *r0 = xmalloc(sizeof(zc) * n);
  zc *xc = xmalloc(sizeof(zc) * n), *yc = xmalloc(sizeof(zc) * (m + 1));
  zc *rs = xmalloc(sizeof(zc) * (m + 1)), *cg = xmalloc(sizeof(zc) * (m + 1));
  double *rc = xmalloc(sizeof(double) * (m + 1));
#define HH(i, j) H[(size_t)(j) * (m + 1) + (i)]
  int info = 0, it_total = 0, compRsd = 0;
  double be = 0.0;
  double bn = sqrt(creal(zdotc(n, b, b)));
  if (tbe_out) *tbe_out = 0.0;
  if (bn == 0.0) {
    for (int j = 0; j < n; ++j) x[j] = 0.0;
    *iters = 0; *be_out = 0.0;
    goto done;
  }
  /* sPb: norm of the preconditioned right-hand side (zPackgmres.f90:890-923) */
  double sPb = bn;
  if (diag) {
    precond_left(n, diag, b, r0);
    sPb = sqrt(creal(zdotc(n, r0, r0)));
  }
  mv(ctx, x, r0);
  for (int j = 0; j < n; ++j) r0[j] = b[j] - r0[j];
  precond_left(n, diag, r0, w);
  double beta = sqrt(creal(zdotc(n, w, w)));
  if (beta == 0.0) { *iters = 0; *be_out = 0.0; goto done; }
  int iterOut = 0;
  for (;;) {
    zc aux = 1.0 / beta;
    for (int j = 0; j < n; ++j) V[j] = aux * w[j];
    HH(0, m) = beta;
    for (int j = 1; j <= m; ++j) HH(j, m) = 0.0;
    int jH = 1;
    double bea = 0.0;
    for (;;) {
      memcpy(w, &V[(size_t)(jH - 1) * n], sizeof(zc) * n);
      mv(ctx, w, r0);
      precond_left(n, diag, r0, w);
      for (int j = 0; j < jH; ++j) HH(j, jH - 1) = 0.0;
      int nOrtho = 0;
      double dnormw;
      for (;;) {
        nOrtho++;
        double dloo = 0.0;
        if (orth == 0 || orth == 1) {
          for (int j = 0; j < jH; ++j) {
            const zc *vj = &V[(size_t)j * n];
            zc dVi = zdotc(n, vj, w);
            HH(j, jH - 1) = HH(j, jH - 1) + dVi;
            dloo = dloo + cabs(dVi) * cabs(dVi);
            zc a = -1.0 * dVi;
            for (int i = 0; i < n; ++i) w[i] = w[i] + a * vj[i];
          }
        } else {
          for (int j = 0; j < jH; ++j) cg[j] = zdotc(n, &V[(size_t)j * n], w);
          for (int j = 0; j < jH; ++j) HH(j, jH - 1) = HH(j, jH - 1) + cg[j];
          for (int i = 0; i < n; ++i) {
            zc acc = 0.0;
            for (int j = 0; j < jH; ++j) acc = acc + V[(size_t)j * n + i] * cg[j];
            w[i] = w[i] - acc;
          }
          double s2 = 0.0;
          for (int j = 0; j < jH; ++j) s2 += cabs(cg[j]) * cabs(cg[j]);
          dloo = s2;
        }
        dnormw = sqrt(creal(zdotc(n, w, w)));
        if (orth == 1 || orth == 3) {
          dloo = sqrt(dloo);
          if (dnormw <= dloo && nOrtho < 3) continue;
        }
        break;
      }
      HH(jH, jH - 1) = dnormw;
      {
        zc a = 1.0 / dnormw;
        zc *vn = &V[(size_t)jH * n];
        for (int i = 0; i < n; ++i) vn[i] = a * w[i];
      }
      for (int j = 0; j < jH - 1; ++j) zrot1(&HH(j, jH - 1), &HH(j + 1, jH - 1), rc[j], rs[j]);
      zc ahj = HH(jH - 1, jH - 1), ahp = HH(jH, jH - 1);
      double cc; zc ss;
      zrotg_(&ahj, ahp, &cc, &ss);
      rc[jH - 1] = cc; rs[jH - 1] = ss;
      zrot1(&HH(jH - 1, m), &HH(jH, m), cc, ss);
      zrot1(&HH(jH - 1, jH - 1), &HH(jH, jH - 1), cc, ss);
      HH(jH, jH - 1) = 0.0;
      double dnormres = cabs(HH(jH, m));
      bea = dnormres / (0.0 * 1.0 + sPb);
      it_total = iterOut * m + jH;
      if (hist && it_total - 1 < iterMax) hist[it_total - 1] = bea;
      if (bea <= tol || it_total >= iterMax) {
        compRsd = 1;
        for (int j = 0; j < jH; ++j) yc[j] = HH(j, m);
        for (int j = jH - 1; j >= 0; --j) { /* ztrsv U,N,N */
          if (yc[j] != 0.0) {
            yc[j] = yc[j] / HH(j, j);
            zc t = yc[j];
            for (int i = j - 1; i >= 0; --i) yc[i] = yc[i] - t * HH(i, j);
          }
        }
        for (int i = 0; i < n; ++i) {
          zc acc = 0.0;
          for (int j = 0; j < jH; ++j) acc = acc + V[(size_t)j * n + i] * yc[j];
          r0[i] = acc;
        }
        for (int i = 0; i < n; ++i) xc[i] = x[i] + r0[i];
        memcpy(r0, xc, sizeof(zc) * n);
        mv(ctx, r0, w);
        for (int i = 0; i < n; ++i) w[i] = b[i] - w[i];
        const double true_res = sqrt(creal(zdotc(n, w, w)));
        precond_left(n, diag, w, r0);
        double dres = sqrt(creal(zdotc(n, r0, r0)));
        be = dres / sPb;
        if (be <= tol || it_total >= iterMax) {
          memcpy(x, xc, sizeof(zc) * n);
          info = (be <= tol) ? 0 : -4;
          *iters = it_total;
          *be_out = be;
          /* rinfo(2): backward error of the unpreconditioned system (:1390) */
          if (tbe_out) *tbe_out = true_res / bn;
          goto done;
        }
      }
      jH = jH + 1;
      if (jH <= m) continue;
      break;
    }
    iterOut = iterOut + 1;
    /* restart */
    jH = jH - 1;
    if (bea > tol) {
      for (int j = 0; j < jH; ++j) yc[j] = HH(j, m);
      for (int j = jH - 1; j >= 0; --j) {
        if (yc[j] != 0.0) {
          yc[j] = yc[j] / HH(j, j);
          zc t = yc[j];
          for (int i = j - 1; i >= 0; --i) yc[i] = yc[i] - t * HH(i, j);
        }
      }
      for (int i = 0; i < n; ++i) {
        zc acc = 0.0;
        for (int j = 0; j < jH; ++j) acc = acc + V[(size_t)j * n + i] * yc[j];
        xc[i] = x[i] + acc;
      }
    }
    memcpy(x, xc, sizeof(zc) * n);
    if (compRsd) {
      mv(ctx, x, r0);
      for (int i = 0; i < n; ++i) r0[i] = b[i] - r0[i];
      precond_left(n, diag, r0, w);
      beta = sqrt(creal(zdotc(n, w, w)));
    } else {
      beta = cabs(HH(m, m));
      for (int j = m - 1; j >= 0; --j) {
        HH(j, m) = 0.0;
        zrot1(&HH(j, m), &HH(j + 1, m), rc[j], -rs[j]);
      }
      for (int i = 0; i < n; ++i) {
        zc acc = 0.0;
        for (int j = 0; j <= m; ++j) acc = acc + V[(size_t)j * n + i] * HH(j, m);
        w[i] = acc;
      }
    }
  }
done:
#undef HH
  free(V); free(H); free(w); free(r0); free(xc); free(yc); free(rs); free(cg); free(rc);
  return info;
}

typedef struct { spo_ctx *c; int nt; } mvctx;
static void mv_oracle(void *p, const zc *x, zc *y) {
  mvctx *q = (mvctx *)p;
  spo_matvec(q->c, (const double *)x, (double *)y, q->nt);
}

/* precalculation_diagonal_element (src/lib_sie_tri_mlfmm_interface.f90:94-109):
 * diag = [Z_mm(1,1); Z_mm(2,2)] from the near-field CSR. */
static void jacobi_diag(const spo_ctx *c, zc *diag) {
  const int N = c->N;
  for (int m = 0; m < N; ++m) {
    zc z11 = 0.0, z22 = 0.0;
    for (int64_t j = c->ia[m]; j < c->ia[m + 1]; ++j)
      if (c->ja[j] == m) {
        z11 = c->za[4 * j + 0];
        z22 = c->za[4 * j + 3];
        break;
      }
    diag[m] = z11;
    diag[m + N] = z22;
  }
}

int spo_gmres(spo_ctx *c, const double *b, double *x, int maxit, int restart, double tol,
              int orth, int nthreads, double *hist, int *iters, double *be_out) {
  return spo_gmres_prec(c, b, x, maxit, restart, tol, orth, 0, nthreads, hist, iters, be_out,
                        NULL);
}

int spo_gmres_prec(spo_ctx *c, const double *b, double *x, int maxit, int restart, double tol,
                   int orth, int left_jacobi, int nthreads, double *hist, int *iters,
                   double *be_out, double *tbe_out) {
  mvctx q = {c, nthreads};
  if (restart <= 0 || restart > maxit) restart = maxit;
  zc *diag = NULL;
  if (left_jacobi) {
    diag = xmalloc(sizeof(zc) * 2 * (size_t)c->N);
    jacobi_diag(c, diag);
  }
  int info = gmres_core(2 * c->N, mv_oracle, &q, diag, (const zc *)b, (zc *)x, maxit, restart,
                        tol, orth, hist, iters, be_out, tbe_out);
  free(diag);
  return info;
}

typedef struct { int n; const zc *A; } densectx;
static void mv_dense(void *p, const zc *x, zc *y) {
  densectx *d = (densectx *)p;
  for (int i = 0; i < d->n; ++i) y[i] = 0.0;
  for (int j = 0; j < d->n; ++j)
    for (int i = 0; i < d->n; ++i) y[i] = y[i] + d->A[(size_t)j * d->n + i] * x[j];
}

int spo_gmres_dense(int n, const double *A, const double *b, double *x, int maxit, int restart,
                    double tol, int orth, double *hist, int *iters, double *be_out) {
  densectx d = {n, (const zc *)A};
  if (restart <= 0 || restart > maxit) restart = maxit;
  return gmres_core(n, mv_dense, &d, NULL, (const zc *)b, (zc *)x, maxit, restart, tol, orth,
                    hist, iters, be_out, NULL);
}

int spo_gmres_dense_prec(int n, const double *A, const double *diag, const double *b, double *x,
                         int maxit, int restart, double tol, int orth, double *hist, int *iters,
                         double *be_out, double *tbe_out) {
  densectx d = {n, (const zc *)A};
  if (restart <= 0 || restart > maxit) restart = maxit;
  return gmres_core(n, mv_dense, &d, (const zc *)diag, (const zc *)b, (zc *)x, maxit, restart, tol,
                    orth, hist, iters, be_out, tbe_out);
}

/* Dense impedance matrix over all pairs, the fill loop of
 * lib_sie_tri_solver_normal_I (src/lib_sie_tri_solver_normal.f90:97-125):
 * Z is 2N x 2N complex, row-major, [[D11 D12],[D21 D22]] blocks. */
void spo_dense_fill(const spo_ctx *c, double *Zout, int nthreads) {
  int N = c->N;
  zc *Z = (zc *)Zout;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int m = 0; m < N; ++m)
    for (int n = 0; n < N; ++n) {
      zc D[4];
      impedance_edges(c, m, n, c->rref[m], D);
      Z[(size_t)m * 2 * N + n] = D[0];
      Z[(size_t)m * 2 * N + n + N] = D[1];
      Z[(size_t)(m + N) * 2 * N + n] = D[2];
      Z[(size_t)(m + N) * 2 * N + n + N] = D[3];
    }
}

/* Radiation pattern of edge m (after spo_mlfmm_setup): out = 4*S complex. */
void spo_get_pattern(const spo_ctx *c, int m, double *out) {
  const zc *rt = &c->rt[(size_t)m * 4 * c->S];
  for (int i = 0; i < 4 * c->S; ++i) { out[2 * i] = creal(rt[i]); out[2 * i + 1] = cimag(rt[i]); }
}

/* ------------------------------------------------------------------------ */
/* Tree helpers for the reference's own known-answer tests                  */
/* (lib_tree_helper_functions.f90 test_lib_tree_hf_*, lib_tree.f90          */
/* test_lib_tree_get_level_max).  xs are already-scaled coordinates.        */
/* ------------------------------------------------------------------------ */
uint64_t spo_tree_uindex(const double *xs, int l) {
  uint64_t ix[3];
  for (int d = 0; d < 3; ++d) ix[d] = (uint64_t)ldexp(xs[d], l);
  return morton3(ix[0], ix[1], ix[2], l);
}
int spo_tree_neighbours(int64_t n, int l, int64_t *out) {
  int64_t b[27];
  int k = e2_boxes(n, l, b);
  for (int i = 1; i < k; ++i) out[i - 1] = b[i];
  return k - 1;
}
void spo_tree_centre(int64_t n, int l, double *out) {
  int64_t ix[3];
  demorton3((uint64_t)n, l, ix);
  for (int d = 0; d < 3; ++d) out[d] = ldexp(1.0, -l) * ((double)ix[d] + 0.5);
}
int spo_tree_level_max(int n, const double *xs, int s, int l_th) {
  uint64_t *u = xmalloc(sizeof(uint64_t) * (n ? n : 1));
  for (int m = 0; m < n; ++m) u[m] = spo_tree_uindex(&xs[3 * m], l_th);
  qsort(u, n, sizeof(uint64_t), cmp_u64);
  int l_max = 1;
  for (int i = 0; i + s < n; ++i) {
    uint64_t a = u[i], b = u[i + s];
    int j = l_th + 1;
    for (;;) {
      j = j - 1;
      a >>= 3; b >>= 3;
      if (a == b) { if (j > l_max) l_max = j; break; }
    }
  }
  free(u);
  return l_max;
}

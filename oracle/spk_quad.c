/* spk_quad.c -- CPU restatement of SpeckleSim's 8-node quadrilateral
 * (10 Huber edge functions) near/dense impedance fill.
 *
 * TEST INFRASTRUCTURE ONLY: this file is the oracle that the tests, the
 * smoke check and the benchmark's CPU baseline compare the GPU library
 * against.  Nothing in the product links or calls it.
 *
 * Follows (file:line in the reference):
 *   sorting_quad                        src/lib_sie_quad_data_container.f90:321-403
 *   set_local_material_parameters_quad  src/lib_sie_quad_data_container.f90:118-129
 *   Neighbours_EdgeCorner_Full          src/lib_sie_quad_calculation_mod.f90:952-995
 *   Common_EdgeCorner                   :997-1024
 *   Sorting_edge, find_edges_in_element :624-759, sorting_edge_sub :889-909
 *   find_edge_index                     :568-622
 *   R_shape, N_differential, f_edge_new :260-300, 352-396
 *   paras_st_a / paras_st_b / R_local   :484-551
 *   Diver_Green (quad module)           :553-564
 *   rho_2_4Feld, phi_range_4Feld        :1042-1052, 1080-1089
 *   impedance_edge_versus_element       src/lib_sie_quad_singularity.f90:22-98
 *   Element_integration                 :100-188
 *   Integration_KABCD_Taylor_ST_4FEdge  :299-380
 *   Element_rotation_edge               :387-439
 *   Integration_KABCD_Edge_new          :479-522
 *   ff_ABC_II_Edge, ff_phi_AB_Edge      :581-637
 *   Lm_*_EA, GL_Int_*_EA_8Node          :656-1045
 *   matrix assembly                     src/lib_sie_quad_solver_normal.f90:165-210
 *   get_vector_b_quad, incident_int     src/lib_sie_quad_calculation_mod.f90:91-138, 1207-1240
 *
 * The 1-D Gauss-Legendre rules are computed (spo_gauss_legendre); the
 * reference tabulates the same nodes as 16-digit double literals
 * (src/lib_sie_integration_points.f90:145-2898).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "spk_oracle.h"

typedef double complex zc;
#define QPI 3.14159265358979323846
#define QC0 ((double)2.99792458e+8f)
#define QEPS0 ((double)8.854187817e-12f)
#define QMY0 ((double)1.2566370614e-6f)

typedef struct { int other, ga[3], gq[3]; } edge_adj;   /* CommonEdge_Elements(2:8) */
typedef struct { int other, ga[2], gq[2]; } corner_adj; /* CommonCorner_Elements(2:6) */

struct spq_ctx {
  double lam, k0, omega;
  zc k1, k2, eps_r1, eps_r2, eta1, eta2;
  int np, nel, medge;
  double *pp;  /* 3*np */
  int *vert;   /* 8*nel, 1-based node ids */
  int *ne, *nc;
  edge_adj **ea;
  corner_adj **ca;
  int *edge_p; /* 5*medge, Edge_I(s)%edge_p(1:5), 1-based element/edge ids */
};

static void *qmalloc(size_t n) {
  void *p = calloc(1, n ? n : 1);
  if (!p) {
    fprintf(stderr, "spk_quad: out of memory\n");
    abort();
  }
  return p;
}

/* local edge node pairs (the edge(10,2) table of Sorting_edge/find_edge_index) */
static const int kEdge[10][2] = {{1, 2}, {2, 3}, {7, 6}, {6, 5}, {8, 4},
                                 {1, 8}, {8, 7}, {3, 4}, {4, 5}, {2, 6}};

/* ------------------------------------------------------------------------ */
/* mesh: sorting_quad                                                        */
/* ------------------------------------------------------------------------ */
static void build_mesh(spq_ctx *c, double Lx, double Ly, int nx, int ny, const double *z) {
  const int nl1 = 2 * nx + 1, nl2 = 2 * ny + 1;
  /* real :: dL(2): the spacing is single precision, and so is dL*(i-1) */
  const float dl1 = (float)(Lx / (nl1 - 1)), dl2 = (float)(Ly / (nl2 - 1));
  c->np = nl1 * nl2;
  c->pp = qmalloc(sizeof(double) * 3 * c->np);
  int *ss = qmalloc(sizeof(int) * nl1 * nl2); /* ss(i, j) -> ss[(j)*nl1 + i] */
  int np = 0;
  for (int i = 0; i < nl1; ++i)
    for (int j = 0; j < nl2; ++j) {
      ++np;
      ss[j * nl1 + i] = np;
      const float px = dl1 * (float)i, py = dl2 * (float)j;
      c->pp[3 * (np - 1) + 0] = -Lx / 2 + (double)px;
      c->pp[3 * (np - 1) + 1] = -Ly / 2 + (double)py;
      c->pp[3 * (np - 1) + 2] = z[j * nl1 + i]; /* zz(i, j), file order i fastest */
    }
  c->nel = nx * ny;
  c->vert = qmalloc(sizeof(int) * 8 * c->nel);
  int nt = 0;
  for (int j = 0; j + 2 < nl1; j += 2)
    for (int i = 0; i + 2 < nl2; i += 2) {
      int *v = c->vert + 8 * nt++;
#define SS(a, b) ss[(b) * nl1 + (a)]
      v[0] = SS(j, i);
      v[1] = SS(j + 1, i);
      v[2] = SS(j + 2, i);
      v[3] = SS(j + 2, i + 1);
      v[4] = SS(j + 2, i + 2);
      v[5] = SS(j + 1, i + 2);
      v[6] = SS(j, i + 2);
      v[7] = SS(j, i + 1);
#undef SS
    }
  free(ss);
}

/* ------------------------------------------------------------------------ */
/* topology: Neighbours_EdgeCorner_Full, Common_EdgeCorner, Sorting_edge     */
/* ------------------------------------------------------------------------ */
static int common_edge_corner(const int *a, const int *q, int *na, int *nq, int *common) {
  int n_equal = 0;
  common[0] = common[1] = common[2] = -1;
  na[0] = na[1] = na[2] = 0;
  nq[0] = nq[1] = nq[2] = 0;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j)
      if (a[i] == q[j]) {
        ++n_equal;
        if (n_equal <= 3) {
          common[n_equal - 1] = q[j];
          na[n_equal - 1] = i + 1;
          nq[n_equal - 1] = j + 1;
        }
      }
  return n_equal;
}

static void neighbours(spq_ctx *c) {
  const int n = c->nel;
  c->ne = qmalloc(sizeof(int) * n);
  c->nc = qmalloc(sizeof(int) * n);
  c->ea = qmalloc(sizeof(edge_adj *) * n);
  c->ca = qmalloc(sizeof(corner_adj *) * n);
  edge_adj *te = qmalloc(sizeof(edge_adj) * (n + 1));
  corner_adj *tc = qmalloc(sizeof(corner_adj) * (4 * n + 1));
  for (int i = 0; i < n; ++i) {
    int nne = 0, nnc = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      int na[3], nq[3], cm[3];
      common_edge_corner(c->vert + 8 * i, c->vert + 8 * j, na, nq, cm);
      if (cm[2] != -1) {
        edge_adj *e = &te[nne++];
        e->other = j + 1;
        for (int k = 0; k < 3; ++k) {
          e->ga[k] = na[k];
          e->gq[k] = nq[k];
        }
      } else if (cm[0] != -1) {
        corner_adj *e = &tc[nnc++];
        e->other = j + 1;
        e->ga[0] = na[0];
        e->ga[1] = na[1];
        e->gq[0] = nq[0];
        e->gq[1] = nq[1];
      }
    }
    c->ne[i] = nne;
    c->nc[i] = nnc;
    c->ea[i] = qmalloc(sizeof(edge_adj) * (nne + 1));
    c->ca[i] = qmalloc(sizeof(corner_adj) * (nnc + 1));
    memcpy(c->ea[i], te, sizeof(edge_adj) * nne);
    memcpy(c->ca[i], tc, sizeof(corner_adj) * nnc);
  }
  free(te);
  free(tc);
}

/* sorting_edge_sub: local edge index of a node pair and its direction */
static void edge_of_pair(int n1, int n2, int *kn, int *dir) {
  for (int k = 0; k < 10; ++k) {
    if (n1 == kEdge[k][0] && n2 == kEdge[k][1]) {
      *kn = k + 1;
      *dir = 1;
      return;
    }
    if (n2 == kEdge[k][0] && n1 == kEdge[k][1]) {
      *kn = k + 1;
      *dir = -1;
      return;
    }
  }
  /* no match: the reference leaves the pair as it was */
  *kn = n1;
  *dir = n2;
}

/* find_edges_in_element: a = (other, ga(1:3), gq(1:3)) after the node swap */
static int find_edges_in_element(const int *a, int j, int *ep) {
  int jj = 0, pI = 0, pII = 0, fI = 0, fII = 0, n1 = 0, n2 = 0;
  if (a[1] == kEdge[j - 1][0] && a[2] == kEdge[j - 1][1]) {
    pI = j; fI = 1; n1 = a[4]; n2 = a[5]; jj = a[0];
  } else if (a[2] == kEdge[j - 1][0] && a[1] == kEdge[j - 1][1]) {
    pI = j; fI = -1; n1 = a[4]; n2 = a[5]; jj = a[0];
  } else if (a[2] == kEdge[j - 1][0] && a[3] == kEdge[j - 1][1]) {
    pI = j; fI = 1; n1 = a[5]; n2 = a[6]; jj = a[0];
  } else if (a[3] == kEdge[j - 1][0] && a[2] == kEdge[j - 1][1]) {
    pI = j; fI = -1; n1 = a[5]; n2 = a[6]; jj = a[0];
  }
  if (jj == 0) return 0;
  edge_of_pair(n1, n2, &pII, &fII);
  ep[0] = pI;
  ep[1] = jj;
  ep[2] = pII;
  ep[3] = fI * fII;
  return 1;
}

static void sorting_edge(spq_ctx *c) {
  const int n = c->nel;
  int *order = qmalloc(sizeof(int) * n * 10);
  for (int i = 0; i < n * 10; ++i) order[i] = 1;
  int *tmp = qmalloc(sizeof(int) * 5 * n * 10);
  int p = 0;
  for (int s = 1; s <= n; ++s) {
    for (int j = 1; j <= 10; ++j) {
      if (order[(s - 1) * 10 + (j - 1)] == 0) continue;
      ++p;
      int *e = tmp + 5 * (p - 1);
      e[0] = s; e[1] = j; e[2] = 0; e[3] = 0; e[4] = 0;
      for (int i = 0; i < c->ne[s - 1]; ++i) {
        const edge_adj *ad = &c->ea[s - 1][i];
        int a[7] = {ad->other, ad->ga[0], ad->ga[1], ad->ga[2], ad->gq[0], ad->gq[1], ad->gq[2]};
        if (a[2] - a[1] != 1) {
          int m = a[2]; a[2] = a[3]; a[3] = m;
          m = a[5]; a[5] = a[6]; a[6] = m;
        }
        int ep[4];
        if (find_edges_in_element(a, j, ep)) {
          order[(ep[1] - 1) * 10 + (ep[2] - 1)] = 0;
          e[0] = s; e[1] = ep[0]; e[2] = ep[1]; e[3] = ep[2]; e[4] = ep[3];
        }
      }
    }
  }
  c->medge = p;
  c->edge_p = qmalloc(sizeof(int) * 5 * (p ? p : 1));
  memcpy(c->edge_p, tmp, sizeof(int) * 5 * p);
  free(tmp);
  free(order);
}

/* find_edge_index: local edge indices of the common side in both elements */
static void find_edge_index(const int *S, int *idx) {
  int A[6];
  memcpy(A, S, sizeof(A));
  if (A[1] - A[0] != 1) {
    int m = A[1]; A[1] = A[2]; A[2] = m;
    m = A[4]; A[4] = A[5]; A[5] = m;
  }
  int pI = 0, pII = 0;
  for (int j = 0; j < 10; ++j) {
    const int fwd = A[0] == kEdge[j][0] && A[1] == kEdge[j][1];
    const int rev = A[1] == kEdge[j][0] && A[0] == kEdge[j][1];
    if (fwd || rev) {
      pI = j + 1;
      for (int kn = 0; kn < 10; ++kn)
        if ((A[3] == kEdge[kn][0] && A[4] == kEdge[kn][1]) ||
            (A[4] == kEdge[kn][0] && A[3] == kEdge[kn][1])) {
          pII = kn + 1;
          break;
        }
    }
  }
  idx[0] = pI;
  idx[1] = pII;
}

/* ------------------------------------------------------------------------ */
/* geometry of the 8-node element (Huber's dissertation, Eqs. 3.7-3.65)       */
/* ------------------------------------------------------------------------ */
static void r_shape(double k, double e, double *N) {
  N[0] = 0.25 * (1 - k) * (1 - e) * (-1 - e - k);
  N[1] = 0.5 * (1 - k * k) * (1 - e);
  N[2] = 0.25 * (1 + k) * (1 - e) * (-1 - e + k);
  N[3] = 0.5 * (1 - e * e) * (1 + k);
  N[4] = 0.25 * (1 + k) * (1 + e) * (-1 + e + k);
  N[5] = 0.5 * (1 - k * k) * (1 + e);
  N[6] = 0.25 * (1 - k) * (1 + e) * (-1 + e - k);
  N[7] = 0.5 * (1 - e * e) * (1 - k);
}
static void n_diff(double k, double e, double *dk, double *de) {
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
/* f_edge_new: f(10) and the derivatives used by n_rot_N */
static void f_edge(double k, double e, double *f, double *dfk, double *dfe) {
  f[0] = -0.25 * (1 - e) * (2 * k + e);
  f[1] = 0.25 * (1 - e) * (2 * k - e);
  f[2] = -0.25 * (1 + e) * (2 * k - e);
  f[3] = 0.25 * (1 + e) * (2 * k + e);
  f[4] = 0.5 * (1 - e * e);
  f[5] = -0.25 * (1 - k) * (2 * e + k);
  f[6] = 0.25 * (1 - k) * (2 * e - k);
  f[7] = -0.25 * (1 + k) * (2 * e - k);
  f[8] = 0.25 * (1 + k) * (2 * e + k);
  f[9] = 0.5 * (1 - k * k);
  dfk[0] = -0.5 * (1 - e);
  dfk[1] = 0.5 * (1 - e);
  dfk[2] = -0.5 * (1 + e);
  dfk[3] = 0.5 * (1 + e);
  dfk[4] = 0;
  dfk[5] = 0.25 * (2 * e + 2 * k - 1);
  dfk[6] = 0.25 * (2 * k - 2 * e - 1);
  dfk[7] = 0.25 * (2 * k - 2 * e + 1);
  dfk[8] = 0.25 * (2 * k + 2 * e + 1);
  dfk[9] = -k;
  dfe[0] = 0.25 * (2 * k + 2 * e - 1);
  dfe[1] = 0.25 * (2 * e - 2 * k - 1);
  dfe[2] = 0.25 * (2 * e - 2 * k + 1);
  dfe[3] = 0.25 * (2 * k + 2 * e + 1);
  dfe[4] = -e;
  dfe[5] = -0.5 * (1 - k);
  dfe[6] = 0.5 * (1 - k);
  dfe[7] = -0.25 * (1 + k);
  dfe[8] = 0.25 * (1 + k);
  dfe[9] = 0;
}

static inline void cross3(const double *a, const double *b, double *o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline double len3(const double *a) { return sqrt(dot3(a, a)); }

/* derived_element_parameters (type B carries nq and n_rot_N as well) */
typedef struct {
  double JJ, nq[3], N[10][3], nrot[10];
} delem;

static void tangents(double k, double e, const double El[8][3], double *Eu, double *Ev) {
  double dk[8], de[8];
  n_diff(k, e, dk, de);
  for (int n = 0; n < 3; ++n) {
    double su = 0, sv = 0;
    for (int q = 0; q < 8; ++q) {
      su += dk[q] * El[q][n];
      sv += de[q] * El[q][n];
    }
    Eu[n] = su;
    Ev[n] = sv;
  }
}

/* paras_st_b (paras_st_a computes the same JJ and N) */
static void paras_b(double k, double e, const double El[8][3], delem *d) {
  double Eu[3], Ev[3], r[3], f[10], dfk[10], dfe[10], v[3];
  tangents(k, e, El, Eu, Ev);
  cross3(Eu, Ev, r);
  d->JJ = len3(r);
  for (int n = 0; n < 3; ++n) d->nq[n] = r[n] / d->JJ;
  f_edge(k, e, f, dfk, dfe);
  for (int n = 0; n < 10; ++n) {
    if (n < 5) {
      cross3(Ev, d->nq, v);
      d->nrot[n] = -dfe[n] / d->JJ;
    } else {
      cross3(d->nq, Eu, v);
      d->nrot[n] = dfk[n] / d->JJ;
    }
    for (int m = 0; m < 3; ++m) d->N[n][m] = f[n] * (v[m] / d->JJ);
  }
}

static void r_local(double k, double e, const double El[8][3], double *r) {
  double N[8];
  r_shape(k, e, N);
  for (int n = 0; n < 3; ++n) {
    double s = 0;
    for (int q = 0; q < 8; ++q) s += N[q] * El[q][n];
    r[n] = s;
  }
}

static void qgreen(zc km, double R, zc *G, zc *DG) {
  const zc ex = cexp(-I * km * R);
  *DG = -(1.0 + I * km * R) * ex / (4 * QPI * (R * R));
  *G = ex / (4 * QPI * R);
}

/* ff_ABC_II_Edge: the five kernels for the 10 source functions */
static void ff_abc(const spq_ctx *c, const double *Na, const double *R, const delem *d,
                   zc ff[5][10]) {
  const double nR = len3(R);
  double rn[3] = {R[0] / nR, R[1] / nR, R[2] / nR};
  zc Ga, GSa, Gm, GSm;
  qgreen(c->k1, nR, &Ga, &GSa);
  qgreen(c->k2, nR, &Gm, &GSm);
  for (int n = 0; n < 10; ++n) {
    double aa[3], t[3];
    cross3(d->nq, d->N[n], aa);
    cross3(aa, rn, t);
    const double f1 = dot3(Na, t) * d->JJ;
    const double f2 = dot3(Na, rn) * d->nrot[n] * d->JJ;
    const double f3 = dot3(Na, aa) * d->JJ;
    ff[0][n] = f1 * (GSa + GSm);
    ff[1][n] = f2 * (1.0 / c->eps_r1 * GSa + 1.0 / c->eps_r2 * GSm);
    ff[2][n] = f2 * (GSa + GSm);
    ff[3][n] = f3 * (Ga + Gm);
    ff[4][n] = f3 * (c->eps_r1 * Ga + c->eps_r2 * Gm);
  }
}

/* the (KA, KBE, KBH, KCE, KCH) -> (A, DE, DH) combination shared by the rules */
static void combine(const spq_ctx *c, zc S[5][10], zc out[3][10]) {
  const zc iwe = I * c->omega * QEPS0, iwm = I * c->omega * QMY0;
  for (int n = 0; n < 10; ++n) {
    out[0][n] = S[0][n];
    out[1][n] = S[1][n] / iwe - (-iwm * S[3][n]);
    out[2][n] = S[2][n] / (-iwm) - iwe * S[4][n];
  }
}

/* ------------------------------------------------------------------------ */
/* regular rule: Element_integration 'Normal' + Integration_KABCD_Edge_new   */
/* ------------------------------------------------------------------------ */
static void int_normal(const spq_ctx *c, const double Ea[8][3], const double Eq[8][3], int pI,
                       int ngp, zc sum[3][10]) {
  double x[64], w[64];
  spo_gauss_legendre(ngp, x, w);
  memset(sum, 0, sizeof(zc) * 30);
  for (int j = 0; j < ngp; ++j)
    for (int i = 0; i < ngp; ++i) {
      delem da;
      paras_b(x[i], x[j], Ea, &da);
      const double TRF = w[i] * w[j] * da.JJ;
      double Ra[3];
      r_local(x[i], x[j], Ea, Ra);
      zc S[5][10];
      memset(S, 0, sizeof(S));
      for (int a = 0; a < ngp; ++a)
        for (int b = 0; b < ngp; ++b) {
          delem dq;
          paras_b(x[a], x[b], Eq, &dq);
          double Rq[3], R[3];
          r_local(x[a], x[b], Eq, Rq);
          for (int m = 0; m < 3; ++m) R[m] = Ra[m] - Rq[m];
          zc ff[5][10];
          ff_abc(c, da.N[pI - 1], R, &dq, ff);
          const double T2 = w[a] * w[b];
          for (int k = 0; k < 5; ++k)
            for (int n = 0; n < 10; ++n) S[k][n] += ff[k][n] * T2;
        }
      zc r[3][10];
      combine(c, S, r);
      for (int k = 0; k < 3; ++k)
        for (int n = 0; n < 10; ++n) sum[k][n] += r[k][n] * TRF;
    }
}

/* ------------------------------------------------------------------------ */
/* self term: Integration_KABCD_Taylor_ST_4FEdge                              */
/* ------------------------------------------------------------------------ */
static void taylor_st_raw(const spq_ctx *c, int pI, double mu, double nu, const double El[8][3],
                          int ngp, zc rv[3][10], zc raw[5][10]) {
  double x1[64], w1[64], x2[64], w2[64];
  spo_gauss_legendre(ngp, x1, w1);
  spo_gauss_legendre(5, x2, w2);
  double Eu[3], Ev[3], Ra[3];
  tangents(mu, nu, El, Eu, Ev);
  delem dpa;
  paras_b(mu, nu, El, &dpa);
  r_local(mu, nu, El, Ra);
  const double *Na = dpa.N[pI - 1];
  const double pb[5] = {atan((1 - nu) / (1 - mu)), QPI - atan((1 - nu) / (1 + mu)),
                        QPI + atan((1 + nu) / (1 + mu)), 2 * QPI - atan((1 + nu) / (1 - mu)),
                        2 * QPI + atan((1 - nu) / (1 - mu))};
  zc KAa[10] = {0}, KAb[10] = {0}, KBEa[10] = {0}, KBEb[10] = {0}, KBHa[10] = {0},
     KBHb[10] = {0}, KCE[10] = {0}, KCH[10] = {0};
  const zc ceps = 1.0 + 1.0 / c->eps_r2;
  for (int m = 0; m < 4; ++m) {
    const double a = (pb[m + 1] - pb[m]) / 2, cc = (pb[m + 1] + pb[m]) / 2;
    for (int i = 0; i < ngp; ++i) {
      const double phi = cc + a * x1[i];
      double A[3];
      for (int n = 0; n < 3; ++n) A[n] = -Eu[n] * cos(phi) - Ev[n] * sin(phi);
      const double Aamp = len3(A);
      double rho2;
      switch (m) {
        case 0: rho2 = (1 - nu) / sin(QPI - phi); break;
        case 1: rho2 = fabs((1 + mu) / cos(QPI - phi)); break;
        case 2: rho2 = fabs((1 + nu) / cos(1.5 * QPI - phi)); break;
        default: rho2 = fabs((1 - mu) / cos(phi)); break;
      }
      /* ff_phi_AB_Edge */
      double fAB[2][10];
      {
        const double Am = len3(A);
        double bb[3];
        for (int n = 0; n < 3; ++n) bb[n] = -A[n] / (Am * Am * Am) / (4 * QPI);
        for (int n = 0; n < 10; ++n) {
          double aa[3], t[3];
          cross3(dpa.nq, dpa.N[n], aa);
          cross3(aa, bb, t);
          fAB[0][n] = dot3(Na, t) * dpa.JJ;
          fAB[1][n] = dot3(Na, bb) * dpa.JJ * dpa.nrot[n];
        }
      }
      for (int j = 0; j < 5; ++j) {
        const double rho = rho2 / 2 * (1 + x2[j]);
        const double mq = mu + rho * cos(phi), nq = nu + rho * sin(phi);
        delem dq;
        paras_b(mq, nq, El, &dq);
        double Rq[3], R[3];
        r_local(mq, nq, El, Rq);
        for (int n = 0; n < 3; ++n) R[n] = Ra[n] - Rq[n];
        const double TRF = 0.5 * rho2 * w1[i] * w2[j] * a;
        zc ff[5][10];
        ff_abc(c, Na, R, &dq, ff);
        for (int n = 0; n < 10; ++n) {
          KAa[n] += (ff[0][n] * rho - 2 * fAB[0][n] / rho) * TRF;
          KBEa[n] += (ff[1][n] * rho - ceps * fAB[1][n] / rho) * TRF;
          KBHa[n] += (ff[2][n] * rho - 2 * fAB[1][n] / rho) * TRF;
          KCE[n] += ff[3][n] * rho * TRF;
          KCH[n] += ff[4][n] * rho * TRF;
        }
      }
      const double T0 = log(fabs(rho2) * Aamp) * w1[i] * a;
      for (int n = 0; n < 10; ++n) {
        KAb[n] += 2 * fAB[0][n] * T0;
        KBEb[n] += ceps * fAB[1][n] * T0;
        KBHb[n] += 2 * fAB[1][n] * T0;
      }
    }
  }
  const zc iwe = I * c->omega * QEPS0, iwm = I * c->omega * QMY0;
  for (int n = 0; n < 10; ++n) {
    rv[0][n] = KAa[n] + KAb[n];
    rv[1][n] = (KBEa[n] + KBEb[n]) / iwe - (-iwm * KCE[n]);
    rv[2][n] = (KBHa[n] + KBHb[n]) / (-iwm) - iwe * KCH[n];
  }
  if (raw)
    for (int n = 0; n < 10; ++n) {
      raw[0][n] = KAa[n] + KAb[n];
      raw[1][n] = KBEa[n] + KBEb[n];
      raw[2][n] = KBHa[n] + KBHb[n];
      raw[3][n] = KCE[n];
      raw[4][n] = KCH[n];
    }
}

static void taylor_st(const spq_ctx *c, int pI, double mu, double nu, const double El[8][3],
                      int ngp, zc rv[3][10]) {
  taylor_st_raw(c, pI, mu, nu, El, ngp, rv, NULL);
}

static void int_self(const spq_ctx *c, const double Ea[8][3], int pI, int ngp, zc sum[3][10]) {
  double x[64], w[64];
  spo_gauss_legendre(ngp, x, w);
  memset(sum, 0, sizeof(zc) * 30);
  for (int j = 0; j < ngp; ++j)
    for (int i = 0; i < ngp; ++i) {
      delem da;
      paras_b(x[i], x[j], Ea, &da);
      const double TRF = w[i] * w[j] * da.JJ;
      zc r[3][10];
      taylor_st(c, pI, x[i], x[j], Ea, ngp, r);
      for (int k = 0; k < 3; ++k)
        for (int n = 0; n < 10; ++n) sum[k][n] += r[k][n] * TRF;
    }
}

/* ------------------------------------------------------------------------ */
/* edge-adjacent elements: GL_Int_theta/Psi/mu/Lambd_EA_8Node                 */
/* ------------------------------------------------------------------------ */
static void lm_psi(int ll, int mm, int nn, double th, double *ab) {
  const double p1 = atan(cos(th)), p2 = atan(sin(th)), p12 = atan(sin(th));
  ab[0] = ab[1] = 0;
  if (ll == 0 && mm == 0) {
    if (nn == 1) { ab[0] = 0; ab[1] = p1; }
    else if (nn == 2) { ab[0] = p1; ab[1] = QPI / 2; }
    else if (nn == 3) { ab[0] = 0; ab[1] = p2; }
    else { ab[0] = p2; ab[1] = QPI / 2; }
  } else if (ll == 0 && mm == 1) {
    if (nn == 1) { ab[0] = p1; ab[1] = QPI / 2; }
    else { ab[0] = p2; ab[1] = QPI / 2; }
  } else if (ll == 1 && mm == 0) {
    ab[0] = 0.0; ab[1] = p12;
  } else if (ll == 1 && mm == 1) {
    ab[0] = p12; ab[1] = QPI / 2;
  } else if (ll == 2 && mm == 0) {
    if (nn == 1) { ab[0] = 0; ab[1] = p2; }
    else if (nn == 2) { ab[0] = p2; ab[1] = QPI / 2; }
    else if (nn == 3) { ab[0] = 0; ab[1] = -p1; }
    else { ab[0] = -p1; ab[1] = QPI / 2; }
  } else { /* ll == 2, mm == 1 */
    if (nn == 1) { ab[0] = p2; ab[1] = QPI / 2; }
    else { ab[0] = -p1; ab[1] = QPI / 2; }
  }
}
static void lm_mu(int ll, int mm, int nn, double th, double ps, double *ab) {
  const double u1p = 2 * cos(th) * tan(QPI / 2 - ps) - 1, u2p = 2 * cos(th) * tan(QPI / 2 - ps) + 1;
  const double u1t = 2 * tan(QPI / 2 - th) - 1, u2t = 2 * tan(QPI / 2 - th) + 1;
  if (ll == 0 && mm == 0) {
    if (nn == 1) { ab[0] = -1; ab[1] = 1; }
    else if (nn == 2) { ab[0] = -1; ab[1] = u1p; }
    else if (nn == 3) { ab[0] = -1; ab[1] = u1t; }
    else { ab[0] = -1; ab[1] = u1p; }
  } else if (ll == 0 && mm == 1) {
    if (nn == 1) { ab[0] = u1p; ab[1] = 1; }
    else { ab[0] = u1p; ab[1] = u1t; }
  } else if (ll == 1) { /* mm 0 and 1 share the limits */
    if (nn == 1) { ab[0] = u1t; ab[1] = 1; }
    else { ab[0] = -1; ab[1] = u2t; }
  } else if (ll == 2 && mm == 0) {
    if (nn == 1) { ab[0] = u2t; ab[1] = 1; }
    else if (nn == 2) { ab[0] = u2p; ab[1] = 1; }
    else if (nn == 3) { ab[0] = -1; ab[1] = 1; }
    else { ab[0] = u2p; ab[1] = 1; }
  } else { /* ll == 2, mm == 1 */
    if (nn == 1) { ab[0] = u2t; ab[1] = u2p; }
    else { ab[0] = -1; ab[1] = u2p; }
  }
}
static double lm_lambda(int ll, int mm, double th, double ps, double u) {
  if (ll == 0 && mm == 0) return (1 + u) / (cos(th) * cos(ps));
  if (ll == 1 && mm == 0) return 2 / (sin(th) * cos(ps));
  if (ll == 2 && mm == 0) return (u - 1) / (cos(th) * cos(ps));
  return 2 / sin(ps); /* mm == 1 */
}
static double lm_theta(int ll, int mm, int nn) {
  if (ll == 0) return (mm == 0) ? ((nn <= 2) ? 0 : QPI / 4) : ((nn == 1) ? 0 : QPI / 4);
  if (ll == 1) return (nn == 1) ? QPI / 4 : QPI / 2;
  if (mm == 0) return (nn <= 2) ? QPI / 2 : 3 * QPI / 4;
  return (nn == 1) ? QPI / 2 : 3 * QPI / 4;
}

static void int_edge_adjacent(const spq_ctx *c, const double Ea[8][3], const double Eq[8][3],
                              int pI, int ngp, zc sum[3][10]) {
  double x[64], w[64];
  spo_gauss_legendre(ngp, x, w);
  memset(sum, 0, sizeof(zc) * 30);
  for (int s = 1; s <= 2; ++s) {
    const int mm = s - 1;
    for (int l = 1; l <= 3; ++l) {
      const int ll = l - 1;
      const int nmax = ((ll == 0 && mm == 0) || (ll == 2 && mm == 0)) ? 4 : 2;
      for (int nn = 1; nn <= nmax; ++nn) {
        /* theta layer */
        const double ta = lm_theta(ll, mm, nn), tb = ta + QPI / 4;
        const double tam = (tb - ta) / 2, tbm = (tb + ta) / 2;
        zc S4[3][10];
        memset(S4, 0, sizeof(S4));
        for (int i4 = 0; i4 < ngp; ++i4) {
          const double th = tam * x[i4] + tbm;
          double pab[2];
          lm_psi(ll, mm, nn, th, pab);
          const double pam = (pab[1] - pab[0]) / 2, pbm = (pab[1] + pab[0]) / 2;
          zc S3[3][10];
          memset(S3, 0, sizeof(S3));
          for (int i3 = 0; i3 < ngp; ++i3) {
            const double ps = pam * x[i3] + pbm;
            double mab[2];
            lm_mu(ll, mm, nn, th, ps, mab);
            const double mam = (mab[1] - mab[0]) / 2, mbm = (mab[1] + mab[0]) / 2;
            zc S2[3][10];
            memset(S2, 0, sizeof(S2));
            for (int i2 = 0; i2 < ngp; ++i2) {
              const double u = mam * x[i2] + mbm;
              /* GL_Int_Lambd_EA_8Node */
              const double lb = lm_lambda(ll, mm, th, ps, u);
              const double lam_ = (lb - 0.0) / 2, lbm = (lb + 0.0) / 2;
              zc S1[5][10];
              memset(S1, 0, sizeof(S1));
              for (int i1 = 0; i1 < ngp; ++i1) {
                const double X = lam_ * x[i1] + lbm;
                const double mu_a = u, nu_a = -1 + X * sin(ps);
                const double mu_q = X * cos(ps) * cos(th) - mu_a, nu_q = X * cos(ps) * sin(th) - 1;
                delem da, dq;
                paras_b(mu_a, nu_a, Ea, &da);
                paras_b(mu_q, nu_q, Eq, &dq);
                double Ra[3], Rq[3], R[3];
                r_local(mu_a, nu_a, Ea, Ra);
                r_local(mu_q, nu_q, Eq, Rq);
                for (int m = 0; m < 3; ++m) R[m] = Ra[m] - Rq[m];
                zc ff[5][10];
                ff_abc(c, da.N[pI - 1], R, &dq, ff);
                const double TRF = lam_ * w[i1] * da.JJ * (X * X) * cos(ps);
                for (int k = 0; k < 5; ++k)
                  for (int n = 0; n < 10; ++n) S1[k][n] += ff[k][n] * TRF;
              }
              zc r1[3][10];
              combine(c, S1, r1);
              for (int k = 0; k < 3; ++k)
                for (int n = 0; n < 10; ++n) S2[k][n] += mam * w[i2] * r1[k][n];
            }
            for (int k = 0; k < 3; ++k)
              for (int n = 0; n < 10; ++n) S3[k][n] += pam * w[i3] * S2[k][n];
          }
          for (int k = 0; k < 3; ++k)
            for (int n = 0; n < 10; ++n) S4[k][n] += tam * w[i4] * S3[k][n];
        }
        for (int k = 0; k < 3; ++k)
          for (int n = 0; n < 10; ++n) sum[k][n] += S4[k][n];
      }
    }
  }
}

/* Element_rotation_edge: node permutation that brings side qI to the bottom */
static void rotate_edge(int qI, const double in[8][3], double out[8][3], int rot[2][10]) {
  static const int r34[2][10] = {{4, 3, 2, 1, 5, 9, 8, 7, 6, 10}, {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1}};
  static const int r67[2][10] = {{8, 9, 6, 7, 10, 2, 1, 4, 3, 5}, {1, 1, 1, 1, 1, -1, -1, -1, -1, -1}};
  static const int r89[2][10] = {{7, 6, 9, 8, 10, 3, 4, 1, 2, 5}, {-1, -1, -1, -1, -1, 1, 1, 1, 1, 1}};
  static const int rid[2][10] = {{1, 2, 3, 4, 5, 6, 7, 8, 9, 10}, {1, 1, 1, 1, 1, 1, 1, 1, 1, 1}};
  const int (*r)[10] = rid;
  for (int j = 1; j <= 8; ++j) {
    int jj = j;
    if (qI == 3 || qI == 4) jj = j <= 4 ? 4 + j : j - 4;
    else if (qI == 6 || qI == 7) jj = j < 7 ? j + 2 : j - 6;
    else if (qI == 8 || qI == 9) jj = j > 2 ? j - 2 : j + 6;
    memcpy(out[jj - 1], in[j - 1], sizeof(double) * 3);
  }
  if (qI == 3 || qI == 4) r = r34;
  else if (qI == 6 || qI == 7) r = r67;
  else if (qI == 8 || qI == 9) r = r89;
  memcpy(rot, r, sizeof(int) * 20);
}

static void element_nodes(const spq_ctx *c, int e, double El[8][3]) {
  for (int n = 0; n < 8; ++n) memcpy(El[n], c->pp + 3 * (c->vert[8 * (e - 1) + n] - 1), 24);
}

/* impedance_edge_versus_element: s (RWG-like edge function), t (element), 1-based */
void spq_edge_element(const spq_ctx *c, int s, int t, double *out) {
  zc sum[3][10];
  const int ii = c->edge_p[5 * (s - 1) + 0], pI = c->edge_p[5 * (s - 1) + 1];
  double Ea[8][3], Eq[8][3];
  element_nodes(c, ii, Ea);
  element_nodes(c, t, Eq);
  int is_edge = 0, is_corner = 0, tidx[2] = {0, 0};
  for (int n = 0; n < c->ne[ii - 1]; ++n)
    if (c->ea[ii - 1][n].other == t) {
      const edge_adj *ad = &c->ea[ii - 1][n];
      const int S6[6] = {ad->ga[0], ad->ga[1], ad->ga[2], ad->gq[0], ad->gq[1], ad->gq[2]};
      find_edge_index(S6, tidx);
      is_edge = 1;
      break;
    }
  for (int n = 0; n < c->nc[ii - 1]; ++n)
    if (c->ca[ii - 1][n].other == t) is_corner = 1;
  if (t == ii) {
    int_self(c, Ea, pI, 15, sum);
  } else if (is_edge) {
    double Eaa[8][3], Eqq[8][3];
    int ra[2][10], rq[2][10];
    rotate_edge(tidx[0], Ea, Eaa, ra);
    rotate_edge(tidx[1], Eq, Eqq, rq);
    zc sa[3][10];
    int_edge_adjacent(c, Eaa, Eqq, ra[0][pI - 1], 5, sa);
    for (int n = 0; n < 10; ++n)
      for (int k = 0; k < 3; ++k)
        sum[k][n] = sa[k][rq[0][n] - 1] * (double)rq[1][n] * (double)ra[1][pI - 1];
  } else if (is_corner) {
    int_normal(c, Ea, Eq, pI, 6, sum);
  } else {
    int_normal(c, Ea, Eq, pI, 4, sum);
  }
  zc *o = (zc *)out;
  for (int k = 0; k < 3; ++k)
    for (int n = 0; n < 10; ++n) o[k * 10 + n] = sum[k][n];
}

/* ------------------------------------------------------------------------ */
/* context                                                                   */
/* ------------------------------------------------------------------------ */
spq_ctx *spq_create(double lambda, double er1_re, double er1_im, double er2_re, double er2_im,
                    double Lx, double Ly, int nx, int ny, const double *z) {
  if (nx < 1 || ny < 1 || !z) return NULL;
  spq_ctx *c = qmalloc(sizeof(spq_ctx));
  c->lam = lambda;
  c->k0 = 2 * QPI / lambda;
  c->omega = 2 * QPI * QC0 / lambda;
  c->eps_r1 = CMPLX(er1_re, er1_im);
  c->eps_r2 = CMPLX(er2_re, er2_im);
  c->k1 = c->k0 * csqrt(c->eps_r1);
  c->k2 = c->k0 * csqrt(c->eps_r2);
  c->eta1 = csqrt(QMY0 / (QEPS0 * c->eps_r1));
  c->eta2 = csqrt(QMY0 / (QEPS0 * c->eps_r2));
  build_mesh(c, Lx, Ly, nx, ny, z);
  neighbours(c);
  sorting_edge(c);
  return c;
}

void spq_destroy(spq_ctx *c) {
  if (!c) return;
  for (int i = 0; i < c->nel; ++i) {
    free(c->ea[i]);
    free(c->ca[i]);
  }
  free(c->ea); free(c->ca); free(c->ne); free(c->nc);
  free(c->pp); free(c->vert); free(c->edge_p);
  free(c);
}

int spq_n_edges(const spq_ctx *c) { return c->medge; }
int spq_n_elements(const spq_ctx *c) { return c->nel; }
int spq_n_nodes(const spq_ctx *c) { return c->np; }
void spq_nodes(const spq_ctx *c, double *pp) { memcpy(pp, c->pp, sizeof(double) * 3 * c->np); }
void spq_elements(const spq_ctx *c, int *v) { memcpy(v, c->vert, sizeof(int) * 8 * c->nel); }
void spq_edges(const spq_ctx *c, int *ep) { memcpy(ep, c->edge_p, sizeof(int) * 5 * c->medge); }

/* Sum_KA/KDE/KDH for all (s, t) and the assembly of MM (row-major, 2M x 2M)
 * for PMCHWT (1), MCTF (2) or ICTF (3), lib_sie_quad_solver_normal.f90:165-210 */
void spq_assemble(const spq_ctx *c, int formulation, double *MMout, int nthreads) {
  const int M = c->medge, ne = c->nel;
  zc *sums = qmalloc(sizeof(zc) * (size_t)M * ne * 30); /* [s][t][k][n] */
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads > 0 ? nthreads : 1)
  for (int s = 1; s <= M; ++s)
    for (int t = 1; t <= ne; ++t)
      spq_edge_element(c, s, t, (double *)(sums + ((size_t)(s - 1) * ne + (t - 1)) * 30));
  zc *MM = (zc *)MMout;
  const zc ea = (c->eta1 + c->eta2) / 2;
  for (int s = 0; s < M; ++s)
    for (int t = 0; t < M; ++t) {
      const int *e = c->edge_p + 5 * t;
      const int ii = e[0], pa = e[1], jj = e[2], pq = e[3], f = e[4];
      zc m3[3];
      for (int k = 0; k < 3; ++k) {
        zc v = sums[((size_t)s * ne + (ii - 1)) * 30 + k * 10 + (pa - 1)];
        if (f != 0) v = v + sums[((size_t)s * ne + (jj - 1)) * 30 + k * 10 + (pq - 1)] * (double)f;
        m3[k] = v;
      }
      const zc Mmm = m3[0], Mem = m3[1], Mme = m3[2];
      zc a = Mmm, b = Mem, cc = Mme, d = Mmm;
      if (formulation == 3) {
        a = Mmm / ea; b = Mem / ea; cc = Mme * ea; d = Mmm * ea;
      } else if (formulation == 2) {
        cc = Mme * c->eta1 * c->eta2; d = Mmm * c->eta1 * c->eta2;
      }
      MM[(size_t)s * 2 * M + t] = a;
      MM[(size_t)s * 2 * M + M + t] = b;
      MM[(size_t)(M + s) * 2 * M + t] = cc;
      MM[(size_t)(M + s) * 2 * M + M + t] = d;
    }
  free(sums);
}

/* ------------------------------------------------------------------------ */
/* probes for the pinning tests                                              */
/* ------------------------------------------------------------------------ */
/* geometry at (k, e) of an element El[24]: N[8], dN/dk[8], dN/de[8], JJ,
 * nq[3], the 10 edge functions Nq[10][3] and n.curl Nq[10] */
void spq_geometry_probe(double k, double e, const double *El24, double *out) {
  double El[8][3];
  memcpy(El, El24, sizeof(El));
  r_shape(k, e, out);
  n_diff(k, e, out + 8, out + 16);
  delem d;
  paras_b(k, e, El, &d);
  out[24] = d.JJ;
  memcpy(out + 25, d.nq, 3 * sizeof(double));
  memcpy(out + 28, d.N, 30 * sizeof(double));
  memcpy(out + 58, d.nrot, 10 * sizeof(double));
}

/* The self-term inner integral at one observation point (mu, nu) of element
 * El[24], test function pI (1-based), ngp angular points per sector: the five
 * singularity-subtracted sums (KA, KB-E, KB-H with their log terms, KC-E,
 * KC-H), then the combined (A, DE, DH) rows; out = 80 complex. */
void spq_self_point(const spq_ctx *c, const double *El24, int pI, double mu, double nu, int ngp,
                    double *out) {
  double El[8][3];
  memcpy(El, El24, sizeof(El));
  zc rv[3][10], raw[5][10];
  taylor_st_raw(c, pI, mu, nu, El, ngp, rv, raw);
  zc *o = (zc *)out;
  for (int k = 0; k < 5; ++k)
    for (int n = 0; n < 10; ++n) o[k * 10 + n] = raw[k][n];
  for (int k = 0; k < 3; ++k)
    for (int n = 0; n < 10; ++n) o[50 + k * 10 + n] = rv[k][n];
}

/* Plane-wave right-hand side (get_vector_b_quad with incident_int,
 * src/lib_sie_quad_calculation_mod.f90:91-138, 1207-1240): for edge function
 * s, the 6 x 6 Gauss rule over its own element ii only (edge_p(1)), of
 * N_pI . E_inc and N_pI . H_inc with E_inc = E0 exp(-i conj(k1) k.R) -- the
 * reference's dot_product(k1*k_in, R_a) conjugates its complex first
 * argument -- and H_inc = (k x E_inc)/(my_1 c0), my_1 = my_0; then the
 * formulation's scaling: PMCHWT (E, H), MCTF (E, H eta1 eta2), ICTF
 * (E / eta_a, H eta_a), eta_a = (eta1 + eta2)/2.  The edge vectors are those
 * of paras_st_b: (Ev x n)/JJ and (n x Eu)/JJ; Edge_vector_new divides by
 * Eu.(Ev x n), which equals JJ = |Eu x Ev|. */
void spq_rhs_plane(const spq_ctx *c, int formulation, const double *kin, const double *ein,
                   double *bout) {
  const int M = c->medge, ngp = 6;
  double x[64], w[64];
  spo_gauss_legendre(ngp, x, w);
  zc *b = (zc *)bout;
  const zc k1c = conj(c->k1);
  const double c0my = QMY0 * QC0;
  const zc ea = (c->eta1 + c->eta2) / 2, e12 = c->eta1 * c->eta2;
  for (int s = 0; s < M; ++s) {
    const int ii = c->edge_p[5 * s + 0], pI = c->edge_p[5 * s + 1];
    double El[8][3];
    element_nodes(c, ii, El);
    zc SE = 0, SH = 0;
    for (int j = 0; j < ngp; ++j)
      for (int i = 0; i < ngp; ++i) {
        delem d;
        paras_b(x[i], x[j], El, &d);
        double Ra[3];
        r_local(x[i], x[j], El, Ra);
        const double TRF = w[i] * w[j] * d.JJ;
        const zc e = cexp(-I * k1c * (kin[0] * Ra[0] + kin[1] * Ra[1] + kin[2] * Ra[2]));
        const zc E[3] = {ein[0] * e, ein[1] * e, ein[2] * e};
        const zc H[3] = {(kin[1] * E[2] - kin[2] * E[1]) / c0my,
                         (kin[2] * E[0] - kin[0] * E[2]) / c0my,
                         (kin[0] * E[1] - kin[1] * E[0]) / c0my};
        const double *Nv = d.N[pI - 1];
        SE += (Nv[0] * E[0] + Nv[1] * E[1] + Nv[2] * E[2]) * TRF;
        SH += (Nv[0] * H[0] + Nv[1] * H[1] + Nv[2] * H[2]) * TRF;
      }
    zc be = SE, bh = SH;
    if (formulation == 2) bh = SH * e12;
    else if (formulation == 3) {
      be = SE / ea;
      bh = SH * ea;
    }
    b[s] = be;
    b[M + s] = bh;
  }
}

/* the regular rule of impedance_edge_versus_element at any order ngp for
 * edge function s and element t (1-based) */
void spq_regular_pair(const spq_ctx *c, int s, int t, int ngp, double *out) {
  zc sum[3][10];
  const int ii = c->edge_p[5 * (s - 1) + 0], pI = c->edge_p[5 * (s - 1) + 1];
  double Ea[8][3], Eq[8][3];
  element_nodes(c, ii, Ea);
  element_nodes(c, t, Eq);
  int_normal(c, Ea, Eq, pI, ngp, sum);
  zc *o = (zc *)out;
  for (int k = 0; k < 3; ++k)
    for (int n = 0; n < 10; ++n) o[k * 10 + n] = sum[k][n];
}

/* ------------------------------------------------------------------------ */
/* scattered field of a solution                                             */
/* ------------------------------------------------------------------------ */
/* The element-local coefficients S_EH of a solution x[2M]
 * (src/lib_sie_quad_solver_normal.f90:312-323), then at each point the sum
 * over all elements of Integration_scattered_field_quad
 * (src/lib_sie_quad_calculation_mod.f90:1281-1323), ngp x ngp Gauss points
 * on the source element, exterior medium k1:
 *   E += KA SE + KB SH / (i w eps0) + i w my0 KC SH
 *   H += KA SH + KB SE / (-i w my0) - i w eps0 KC SE
 * with KA = (n x N) x gradG ff, KB = nrot gradG ff, KC = (n x N) G ff,
 * ff = JJ w_i w_j.  E, H: 3 complex per point. */
void spq_scattered_field(const spq_ctx *c, int npts, const double *r, const double *xin, int ngp,
                         double *Eout, double *Hout) {
  const int M = c->medge, nel = c->nel;
  const zc *x = (const zc *)xin;
  zc *seh = calloc((size_t)2 * nel * 10, sizeof(zc));
  for (int m = 0; m < M; ++m) {
    const int *e = c->edge_p + 5 * m;
    const int i = e[0] - 1, pa = e[1] - 1, j = e[2] - 1, pb = e[3] - 1;
    seh[(size_t)i * 10 + pa] = x[m];
    seh[(size_t)(i + nel) * 10 + pa] = x[m + M];
    if (j >= 0) {
      seh[(size_t)j * 10 + pb] = seh[(size_t)i * 10 + pa] * e[4];
      seh[(size_t)(j + nel) * 10 + pb] = seh[(size_t)(i + nel) * 10 + pa] * e[4];
    }
  }
  double xg[64], wg[64];
  spo_gauss_legendre(ngp, xg, wg);
  const zc iwe = I * c->omega * QEPS0, iwm = I * c->omega * QMY0;
  zc *E = (zc *)Eout, *H = (zc *)Hout;
  for (int p = 0; p < npts; ++p) {
    const double *Ra = r + 3 * p;
    zc sE[3] = {0, 0, 0}, sH[3] = {0, 0, 0};
    for (int s = 0; s < nel; ++s) {
      double Eq[8][3];
      element_nodes(c, s + 1, Eq);
      const zc *SE = seh + (size_t)s * 10, *SH = seh + (size_t)(s + nel) * 10;
      for (int i = 0; i < ngp; ++i)
        for (int j = 0; j < ngp; ++j) {
          delem d;
          paras_b(xg[i], xg[j], Eq, &d);
          double Rq[3], R[3];
          r_local(xg[i], xg[j], Eq, Rq);
          for (int k = 0; k < 3; ++k) R[k] = Ra[k] - Rq[k];
          const double Rn = len3(R);
          zc G, DG;
          qgreen(c->k1, Rn, &G, &DG);
          zc gg[3];
          for (int k = 0; k < 3; ++k) gg[k] = R[k] / Rn * DG;
          const double ff = d.JJ * wg[i] * wg[j];
          for (int n = 0; n < 10; ++n) {
            double aa[3];
            cross3(d.nq, d.N[n], aa);
            const zc KA[3] = {(aa[1] * gg[2] - aa[2] * gg[1]) * ff,
                              (aa[2] * gg[0] - aa[0] * gg[2]) * ff,
                              (aa[0] * gg[1] - aa[1] * gg[0]) * ff};
            for (int k = 0; k < 3; ++k) {
              const zc KB = d.nrot[n] * gg[k] * ff, KC = aa[k] * G * ff;
              sE[k] += KA[k] * SE[n] + 1.0 / iwe * KB * SH[n] - (-iwm) * KC * SH[n];
              sH[k] += KA[k] * SH[n] + 1.0 / (-iwm) * KB * SE[n] - iwe * KC * SE[n];
            }
          }
        }
    }
    for (int k = 0; k < 3; ++k) {
      E[3 * p + k] = sE[k];
      H[3 * p + k] = sH[k];
    }
  }
  free(seh);
}

// spk_internal.h -- shared definitions of the MI355X SIE solve loop.
//
// Layout in HBM (one handle = one mesh):
//   * RWG sides: FnRec[2N] (AoS, 24 doubles, one 192-byte record per side)
//     and ElemRec[2N] (element id + its three node ids) for adjacency tests.
//   * Edges are renumbered "box-major": position p in [0,N) runs over the
//     leaf boxes in Morton order, edges ascending inside a box.  perm[p] is
//     the reference edge number.
//   * Near field: every leaf box b owns a dense block matrix
//     Z_b[rows = edges of b][cols = E2(b) sorted ascending] of 2x2 complex
//     blocks (D11,D12,D21,D22), row-major, 64 B per block in FP64.  The
//     reference CSR row of edge m is row (pos(m)-first_b) of Z_b with
//     ja = E2(b), so the CSR is recovered exactly (ia, ja, values).
//   * Radiation patterns: RT[p][q][S], q = k*2+comp (k medium, comp
//     theta/phi), S = 2K^2+2 samples, sample index fastest.
//   * MLFMM box coefficients per level: [box][8][S] with item
//     it = comp*4 + cur*2 + k (cur = xe/xh), identical to the reference's
//     a_nm/b_nm item(1..4) lists.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/speckle_gpu.h"

namespace spk {

// ----------------------------------------------------------------------------
// constants (src/lib_sie_constants.f90:9-11 are single-precision literals)
// ----------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846;
constexpr double kC0 = (double)2.99792458e+8f;
constexpr double kEps0 = (double)8.854187817e-12f;
constexpr double kMy0 = (double)1.2566370614e-6f;
constexpr int kThresholdLevel = 21;  // int(8*8*1.0/3), lib_tree.f90:235

struct cplx {
  double re, im;
};

// Per RWG side record, fn_rwg of src/lib_sie_type_function.f90:12-20.
struct FnRec {
  double c[9];    // p1, p2, pc
  double area;
  double length;
  double df;      // sign*length/area
  double mid[3];  // element centroid
  double nq[3];   // unit normal (norm_vec_r_new)
  double nqf[3];  // nq rounded to float (cross_rc uses cmplx(nq) = single)
  double sign;
  double pad[2];
};
static_assert(sizeof(FnRec) == 24 * sizeof(double), "FnRec layout");

struct ElemRec {
  int elem;       // element index (0-based)
  int node[3];    // node indices of that element (1-based, as in tt)
};

struct Material {
  double k0, omega, my1, my2;
  cplx k1, k2, eps1, eps2, eta1, eta2, eta_a;
  cplx a_lc[2], d_lc[2];  // formulation_coefficient (MCTF2)
};

struct Tables {
  int K = 0, S = 0;
  std::vector<double> khat, ttheta, tphi, trf;  // S*3, S*3, S*3, S
};

// One octree level of non-empty boxes.
struct Level {
  std::vector<int64_t> n;          // sorted Morton indices
  std::vector<int> parent;         // position in level l-1 (-1 at l_min)
  std::vector<int> octant;         // child index inside the parent
  std::vector<int> child;          // 8 per box, -1 if empty (l < l_max)
  std::vector<int> i4_off;         // I4 CSR over boxes
  std::vector<int> i4_box;         // partner position in this level
  std::vector<int> i4_slot;        // TL slot (ix*7+iy)*7+iz of offset a-b
};

struct Tree {
  int l_min = 0, l_max = 0;
  double D = 0, xmin[3] = {0, 0, 0};
  std::vector<int64_t> leaf;       // per edge, Morton index at l_max
  std::vector<Level> lev;          // [0..l_max]
  std::vector<int> perm;           // position -> edge
  std::vector<int> pos;            // edge -> position
  std::vector<int> leaf_box;       // edge -> leaf box position
  std::vector<int> box_first, box_count;  // leaf level, in positions
  std::vector<int> e2_off, e2_list;       // per leaf box, sorted edge numbers
  std::vector<int64_t> near_off;          // per leaf box, first block index
  int64_t nnz = 0;
};

// host setup steps (spk_host.cpp)
void make_material(const spk_params &p, Material *m);
int build_rwg(int np, const double *pp, int nt, const int *tt, int normal_convention,
              std::vector<int> *corner, std::vector<int> *element, std::vector<FnRec> *fn,
              std::vector<ElemRec> *er, std::vector<double> *rref, std::string *err);
void build_tree(const std::vector<FnRec> &fn, int n_edges, int s_opt, Tree *t);
void box_centre(const Tree &t, int64_t n, int l, double *x);
void make_tables(int K, Tables *tb);
// TL*TRF per level: [343][2][S] complex, zero outside the live slots.
void make_tl_level(const Tables &tb, const Material &mat, const Tree &t, int l,
                   std::vector<cplx> *out, int nthreads);
// SS / RR phase tables for child level l: [8][2][S] complex each.
void tl_coefficients(const Tables &tb, const Material &mat, const Tree &t, int l,
                     std::vector<double> *rhat, std::vector<cplx> *coef, std::vector<int> *live);
void make_phase_level(const Tables &tb, const Material &mat, const Tree &t, int l,
                      std::vector<cplx> *up, std::vector<cplx> *down);
void quadrature_tri(int ngp, double *a, double *b, double *w);
// the library's error message (spk_last_error); returns code
int set_error(int code, const std::string &msg);
void gauss_legendre(int n, double *x, double *w);
void hankel2_sph(int n, double zre, double zim, cplx *out);
// Lagrange interpolation from the K_m to the K_n sample grid (stencil 2p)
// on one (theta, phi) component pair of one medium, as CSR: rows
// comp * S_n + sample, columns comp * S_m + sample
void interp_pair_csr(int p, int Km, int Kn, std::vector<int> *rowptr, std::vector<int> *col,
                     std::vector<double> *val);
void csr_transpose(int ncols, const std::vector<int> &rowptr, const std::vector<int> &col,
                   const std::vector<double> &val, std::vector<int> *trow, std::vector<int> *tcol,
                   std::vector<double> *tval);

}  // namespace spk

// spk_quad.h -- 8-node quadrilateral (10 Huber edge functions) fill:
// host topology and the device kernels' argument structs.
#pragma once

#include <vector>

#include "spk_device.h"

namespace spk {

constexpr int kQuadMaxAdj = 8;  // edge or corner neighbours per element

struct QuadEdgeAdj {
  int other;  // 1-based element
  int ga[3], gq[3];
  int side[2];  // find_edge_index: local side of the common edge in both elements
};

struct QuadMesh {
  int np = 0, nel = 0, medge = 0;
  std::vector<double> pp;  // 3*np
  std::vector<int> vert;   // 8*nel, 1-based nodes
  std::vector<std::vector<QuadEdgeAdj>> eadj;
  std::vector<std::vector<int>> cadj;  // 1-based elements
  std::vector<int> edge_p;             // 5*medge, Edge_I(s)%edge_p(1:5)
};

void quad_mesh(double Lx, double Ly, int nx, int ny, const double *z, QuadMesh *m);
void quad_topology(QuadMesh *m);
void quad_side_index(const int ga[3], const int gq[3], int side[2]);
void quad_edge_nodes(const QuadMesh &m, int s, int *n1, int *n2);

// constants of the kernels (materials and 1-D Gauss-Legendre rules)
struct QuadDev {
  zd k1, k2, ie1, ie2, e1, e2, ceps, iwe, iwm;
  double x4[4], w4[4], x5[5], w5[5], x6[6], w6[6], x15[15], w15[15];
};

// element coordinates and topology in HBM
struct QuadTopo {
  const double *el;    // nel * 24: node n of element e at el[e*24 + n*3 + d]
  const int *edge_p;   // medge * 5
  const int *eadj;     // nel * kQuadMaxAdj: other element (1-based) or 0
  const int *eside;    // nel * kQuadMaxAdj * 2
  const int *cadj;     // nel * kQuadMaxAdj
  int nel, medge;
};

// regular pairs (no shared node) of rows [s0, s0+rows) x all elements into
// sums[r][t][30]; other pairs' slots are left as they are
void launch_quad_regular_grid(const QuadDev &d, const QuadTopo &T, int s0, int rows, zd *sums,
                              hipStream_t st);
// regular pairs from a list (s, t, ngp), 0-based s and t, out[i*30]
void launch_quad_regular_list(const QuadDev &d, const QuadTopo &T, int n, const int *s,
                              const int *t, const int *ngp, zd *out, hipStream_t st);
// self pairs: rows ls[i] (0-based edge functions, host list) against their
// own element, out[i*30] (device); each element's self terms are integrated
// once for all its rows.  Synchronises the stream before returning.
hipError_t quad_self_rows(const QuadDev &d, const QuadTopo &T, const QuadMesh &m,
                          const std::vector<int> &ls, zd *out, hipStream_t st);
// edge-adjacent pairs (s, t, side_a, side_q), out[i*30]
void launch_quad_edge_adj(const QuadDev &d, const QuadTopo &T, int n, const int *s, const int *t,
                          const int *side, zd *out, hipStream_t st);
// corner-adjacent pairs (s[i], t[i]) (0-based) by the 6 x 6 rule, out[i*30]
void launch_quad_corner(const QuadDev &d, const QuadTopo &T, int n, const int *s, const int *t,
                        zd *out, hipStream_t st);
// x[k] = z[perm[k]] (the row interchanges of an LU factorisation)
void launch_quad_permute(int64_t n, const int64_t *perm, const zd *z, zd *x, hipStream_t st);
// sums[r[i]][t[i]][:] = src[i][:] (30 complex each)
void launch_quad_scatter(int n, int nel, const int *r, const int *t, const zd *src, zd *sums,
                         hipStream_t st);
// plane-wave right-hand side b[2M]: phase factor k1c = conj(k1), H = (k x E)/c0my,
// formulation scaling b = [sE * <N, E>; sH * <N, H>]
struct QuadRhsArgs {
  double kin[3], ein[3];
  zd k1c;
  double c0my;
  zd sE, sH;
};
void launch_quad_rhs_plane(const QuadDev &d, const QuadTopo &T, const QuadRhsArgs &a, zd *b,
                           hipStream_t st);
// material constants and 1-D rules of the kernels (eta_a = (eta1 + eta2)/2,
// eta12 = eta1 eta2 for the formulation scalings)
void quad_constants(const spk_params &p, QuadDev *D, zd *eta_a, zd *eta12);
QuadRhsArgs quad_rhs_args(int formulation, const QuadDev &D, zd eta_a, zd eta12,
                          const double k_in[3], const double e_in[3]);
// element nodes (nel * 24) and the neighbour arrays of QuadTopo; false when
// an element has more than kQuadMaxAdj edge or corner neighbours
bool quad_topology_arrays(const QuadMesh &m, std::vector<double> *el, std::vector<int> *eadj,
                          std::vector<int> *eside, std::vector<int> *cadj);
// MM (row-major 2M x 2M) rows [s0, s0+rows) from sums[r][t][30]
void launch_quad_assemble(const QuadTopo &T, int formulation, zd eta_a, zd eta12, int s0,
                          int rows, const zd *sums, zd *MM, hipStream_t st);


// MLFMM over the quadrilateral edge functions.  Constants of the fold per
// medium (k = 0 exterior, 1 interior), with mik = -i k_m and the row
// scalings sa..sd of the formulation:
//   cAE = sa mik, cJE = sb iwm, cDE = sb mik / (eps_m iwe),
//   cJH = -sc iwe eps_m, cDH = -sc mik / iwm, cAH = sd mik
struct QuadFoldC {
  zd cAE[2], cJE[2], cDE[2], cJH[2], cDH[2], cAH[2];
};
// doubles per position of the pattern quadrature points
constexpr int kQuadPtStride = 16 * 6 + 32 * 7;
// rows of the pattern array per position (8 radiating, 6 receiving)
constexpr int kQuadRtRows = 14;
void launch_qpoints(const QuadDev &d, const QuadTopo &T, int N, const int *perm,
                    const int *leaf_box, const double *leaf_centre, double *pts, hipStream_t st);
void launch_qpatterns(int N, int S, const double *khat, const double *pts, zd k1, zd k2, zd *rt,
                      hipStream_t st);
template <class T>
void launch_qagg(int nbl, int S, int N, const int *box_first, const int *box_count,
                 const int *perm, const T *rt, const T *x, T *B, size_t half, hipStream_t st);
template <class T>
void launch_qfold(int nbl, int S, const T *D, size_t half, const double *khat,
                  const QuadFoldC &C, T *W, hipStream_t st);
template <class T>
void launch_qdisagg(int ntask, int N, int S, const int *task, const int *perm,
                    const int *leaf_box_of_pos, const T *rt, const T *W, T *y, hipStream_t st);
// near field of leaf boxes [b0, b1): pair sums (regular pairs), the self /
// edge- / corner-adjacent list results scattered to their pair index, then
// the 2 x 2 blocks.  Regular pairs by tasks (box, 64-row block, element of
// T(b)): toff[b] = the first task of box b, ntask = toff[b1] - toff[b0].
void launch_qn_regular(const QuadDev &d, const QuadTopo &T, int b0, int b1, int64_t ntask,
                       const int64_t *toff, const int64_t *pbase, const int *tb_off,
                       const int *tb_list, const int *box_first, const int *box_count,
                       const int *perm, zd *sums, hipStream_t st);
void launch_qn_scatter(int n, const int64_t *dst, const zd *src, zd *sums, hipStream_t st);
void launch_qn_assemble(const QuadTopo &T, int form, zd eta_a, zd eta12, int b0, int b1,
                        int64_t nblocks, const int64_t *near_off, const int *e2_off,
                        const int *e2_list, const int *colA, const int *colB,
                        const int64_t *pbase, const int *tb_off, const zd *sums, zd *blocks,
                        hipStream_t st);
// scattered field of a solution at npts points: element-local coefficients
// seh[2][nel][10], then E, H (3 complex per point)
void launch_quad_local(const QuadTopo &T, const zd *x, zd *seh, hipStream_t st);
void launch_quad_field(const QuadDev &d, const QuadTopo &T, int npts, const double *pts,
                       const zd *seh, double *E, double *H, hipStream_t st);

}  // namespace spk

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

// regular pairs of rows [s0, s0+rows) x all elements into sums[r][t][30]
void launch_quad_regular_grid(const QuadDev &d, const QuadTopo &T, int s0, int rows, zd *sums,
                              hipStream_t st);
// regular pairs from a list (s, t, ngp), 0-based s and t, out[i*30]
void launch_quad_regular_list(const QuadDev &d, const QuadTopo &T, int n, const int *s,
                              const int *t, const int *ngp, zd *out, hipStream_t st);
// self pairs: rows s[i] (0-based) against their own element, out[i*30]
void launch_quad_self(const QuadDev &d, const QuadTopo &T, int n, const int *s, zd *out,
                      hipStream_t st);
// edge-adjacent pairs (s, t, side_a, side_q), out[i*30]
void launch_quad_edge_adj(const QuadDev &d, const QuadTopo &T, int n, const int *s, const int *t,
                          const int *side, zd *out, hipStream_t st);
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
// MM (row-major 2M x 2M) rows [s0, s0+rows) from sums[r][t][30]
void launch_quad_assemble(const QuadTopo &T, int formulation, zd eta_a, zd eta12, int s0,
                          int rows, const zd *sums, zd *MM, hipStream_t st);

}  // namespace spk

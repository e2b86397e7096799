// spk_quad_host.cpp -- host-side mesh and topology of the 8-node
// quadrilateral discretisation: the height-field grid (sorting_quad,
// src/lib_sie_quad_data_container.f90:321-403), edge and corner neighbours
// (Neighbours_EdgeCorner_Full / Common_EdgeCorner,
// src/lib_sie_quad_calculation_mod.f90:952-1024), the global numbering of
// the 10 edge functions per element (Sorting_edge / find_edges_in_element,
// :624-759) and the local side indices of a shared side (find_edge_index,
// :568-622).  Integer output is identical to the reference's lists; the
// neighbour search visits only elements that share a node, in ascending
// element order, which gives the same lists as the reference's all-pairs
// loop.
#include <algorithm>
#include <cstring>
#include <vector>

#include "spk_quad.h"

namespace spk {

namespace {

// node pairs of the 10 local edge functions (edge(10,2) in the reference)
constexpr int kSide[10][2] = {{1, 2}, {2, 3}, {7, 6}, {6, 5}, {8, 4},
                              {1, 8}, {8, 7}, {3, 4}, {4, 5}, {2, 6}};

int count_common(const int *a, const int *q, int ga[3], int gq[3], int cm[3]) {
  int n = 0;
  cm[0] = cm[1] = cm[2] = -1;
  ga[0] = ga[1] = ga[2] = gq[0] = gq[1] = gq[2] = 0;
  for (int i = 0; i < 8; ++i)
    for (int j = 0; j < 8; ++j)
      if (a[i] == q[j]) {
        if (n < 3) {
          cm[n] = q[j];
          ga[n] = i + 1;
          gq[n] = j + 1;
        }
        ++n;
      }
  return n;
}

void side_of_pair(int n1, int n2, int *kn, int *dir) {
  for (int k = 0; k < 10; ++k) {
    if (n1 == kSide[k][0] && n2 == kSide[k][1]) {
      *kn = k + 1;
      *dir = 1;
      return;
    }
    if (n2 == kSide[k][0] && n1 == kSide[k][1]) {
      *kn = k + 1;
      *dir = -1;
      return;
    }
  }
  *kn = n1;
  *dir = n2;
}

}  // namespace

void quad_mesh(double Lx, double Ly, int nx, int ny, const double *z, QuadMesh *m) {
  const int nl1 = 2 * nx + 1, nl2 = 2 * ny + 1;
  // the reference keeps the spacing, and its products with the index, in
  // single precision (real :: dL(2))
  const float dl1 = (float)(Lx / (nl1 - 1)), dl2 = (float)(Ly / (nl2 - 1));
  m->np = nl1 * nl2;
  m->pp.assign(3 * (size_t)m->np, 0.0);
  std::vector<int> ss((size_t)nl1 * nl2);
  int np = 0;
  for (int i = 0; i < nl1; ++i)
    for (int j = 0; j < nl2; ++j) {
      ss[(size_t)j * nl1 + i] = ++np;
      const float px = dl1 * (float)i, py = dl2 * (float)j;
      m->pp[3 * (np - 1) + 0] = -Lx / 2 + (double)px;
      m->pp[3 * (np - 1) + 1] = -Ly / 2 + (double)py;
      m->pp[3 * (np - 1) + 2] = z[(size_t)j * nl1 + i];
    }
  m->nel = nx * ny;
  m->vert.assign(8 * (size_t)m->nel, 0);
  int e = 0;
  auto S = [&](int a, int b) { return ss[(size_t)b * nl1 + a]; };
  for (int x = 0; x + 2 < nl1; x += 2)
    for (int y = 0; y + 2 < nl2; y += 2) {
      int *v = &m->vert[8 * (size_t)e++];
      v[0] = S(x, y);
      v[1] = S(x + 1, y);
      v[2] = S(x + 2, y);
      v[3] = S(x + 2, y + 1);
      v[4] = S(x + 2, y + 2);
      v[5] = S(x + 1, y + 2);
      v[6] = S(x, y + 2);
      v[7] = S(x, y + 1);
    }
}

void quad_topology(QuadMesh *m) {
  const int n = m->nel;
  std::vector<std::vector<int>> node_el((size_t)m->np + 1);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < 8; ++k) node_el[m->vert[8 * (size_t)i + k]].push_back(i);
  m->eadj.assign(n, {});
  m->cadj.assign(n, {});
  std::vector<int> cand;
  for (int i = 0; i < n; ++i) {
    cand.clear();
    for (int k = 0; k < 8; ++k)
      for (int j : node_el[m->vert[8 * (size_t)i + k]])
        if (j != i) cand.push_back(j);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    for (int j : cand) {
      int ga[3], gq[3], cm[3];
      count_common(&m->vert[8 * (size_t)i], &m->vert[8 * (size_t)j], ga, gq, cm);
      if (cm[2] != -1) {
        QuadEdgeAdj a;
        a.other = j + 1;
        for (int k = 0; k < 3; ++k) {
          a.ga[k] = ga[k];
          a.gq[k] = gq[k];
        }
        quad_side_index(a.ga, a.gq, a.side);
        m->eadj[i].push_back(a);
      } else if (cm[0] != -1) {
        m->cadj[i].push_back(j + 1);
      }
    }
  }
  // global numbering of the edge functions
  std::vector<char> used((size_t)n * 10, 1);
  m->edge_p.clear();
  for (int s = 1; s <= n; ++s)
    for (int j = 1; j <= 10; ++j) {
      if (!used[(size_t)(s - 1) * 10 + (j - 1)]) continue;
      int e[5] = {s, j, 0, 0, 0};
      for (const QuadEdgeAdj &ad : m->eadj[s - 1]) {
        int a[7] = {ad.other, ad.ga[0], ad.ga[1], ad.ga[2], ad.gq[0], ad.gq[1], ad.gq[2]};
        if (a[2] - a[1] != 1) {
          std::swap(a[2], a[3]);
          std::swap(a[5], a[6]);
        }
        int fI = 0, n1 = 0, n2 = 0;
        const int s0 = kSide[j - 1][0], s1 = kSide[j - 1][1];
        if (a[1] == s0 && a[2] == s1) {
          fI = 1; n1 = a[4]; n2 = a[5];
        } else if (a[2] == s0 && a[1] == s1) {
          fI = -1; n1 = a[4]; n2 = a[5];
        } else if (a[2] == s0 && a[3] == s1) {
          fI = 1; n1 = a[5]; n2 = a[6];
        } else if (a[3] == s0 && a[2] == s1) {
          fI = -1; n1 = a[5]; n2 = a[6];
        } else {
          continue;
        }
        int kn, dir;
        side_of_pair(n1, n2, &kn, &dir);
        used[(size_t)(a[0] - 1) * 10 + (kn - 1)] = 0;
        e[0] = s; e[1] = j; e[2] = a[0]; e[3] = kn; e[4] = fI * dir;
      }
      m->edge_p.insert(m->edge_p.end(), e, e + 5);
    }
  m->medge = (int)(m->edge_p.size() / 5);
}

// the node pair of edge function s on its first element (1-based node
// numbers), whose midpoint places the function (find_midpoint_edge_quad,
// src/lib_sie_quad_calculation_mod.f90:210-255)
void quad_edge_nodes(const QuadMesh &m, int s, int *n1, int *n2) {
  const int ii = m.edge_p[5 * (size_t)s] - 1, pI = m.edge_p[5 * (size_t)s + 1];
  *n1 = m.vert[8 * (size_t)ii + kSide[pI - 1][0] - 1];
  *n2 = m.vert[8 * (size_t)ii + kSide[pI - 1][1] - 1];
}

void quad_side_index(const int ga[3], const int gq[3], int side[2]) {
  int A[6] = {ga[0], ga[1], ga[2], gq[0], gq[1], gq[2]};
  if (A[1] - A[0] != 1) {
    std::swap(A[1], A[2]);
    std::swap(A[4], A[5]);
  }
  int pI = 0, pII = 0;
  for (int j = 0; j < 10; ++j)
    if ((A[0] == kSide[j][0] && A[1] == kSide[j][1]) ||
        (A[1] == kSide[j][0] && A[0] == kSide[j][1])) {
      pI = j + 1;
      for (int k = 0; k < 10; ++k)
        if ((A[3] == kSide[k][0] && A[4] == kSide[k][1]) ||
            (A[4] == kSide[k][0] && A[3] == kSide[k][1])) {
          pII = k + 1;
          break;
        }
    }
  side[0] = pI;
  side[1] = pII;
}

}  // namespace spk

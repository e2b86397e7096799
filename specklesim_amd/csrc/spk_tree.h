// spk_tree.h -- the octree, its box lists and the near-field (E2) and
// interaction (I4) lists built on the GPU (spk_tree.hip).  The result is the
// same Tree the host builder (spk_host.cpp build_tree) gives, bit for bit;
// the device arrays the kernels read are left in the caller's buffers.
#pragma once
#include <cstdint>
#include <string>

#include "spk_dbuf.h"
#include "spk_internal.h"

namespace spk {

// scratch of one build, kept between builds (sized by the largest mesh)
struct TreeWork {
  DBuf<double> pts, part, box;      // 3N points (SoA), per-block min/max, the 6 extremes
  DBuf<uint64_t> u, us, lk, sk;     // Morton keys, sorted; leaf keys, sorted
  DBuf<int> iota, flag, scan, cnt, nbr, lmax, i4cnt;
  DBuf<int64_t> levn, rows;         // level Morton lists (concatenated); near rows per box
  DBuf<int> lpar, loct, lchild, li4off, li4box, li4slot;
  DBuf<unsigned char> cub;          // hipCUB temporary storage
};

// the device arrays of the tree that the product kernels read
struct TreeDevOut {
  DBuf<int> *perm, *pos, *leaf_box, *leafbox_of_pos, *box_first, *box_count, *e2_off, *e2_list;
  DBuf<int64_t> *near_off;
};

// d_fn: the 2N side records on the device.  Fills *t (host copies of every
// array) and the buffers of out.  Returns a hipError_t; *what names the step.
hipError_t build_tree_device(const FnRec *d_fn, int N, int s_opt, Tree *t, TreeWork &w,
                             const TreeDevOut &out, hipStream_t st, const char **what);

}  // namespace spk

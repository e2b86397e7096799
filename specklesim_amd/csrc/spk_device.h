// spk_device.h -- device-side complex arithmetic and shared kernel argument
// structs for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "spk_internal.h"

namespace spk {

struct zd {
  double x, y;
};
struct zf {
  float x, y;
};

__host__ __device__ inline zd mk(double a, double b) { return zd{a, b}; }
__host__ __device__ inline zd operator+(zd a, zd b) { return zd{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline zd operator-(zd a, zd b) { return zd{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline zd operator-(zd a) { return zd{-a.x, -a.y}; }
__host__ __device__ inline zd operator*(zd a, zd b) {
  return zd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__host__ __device__ inline zd operator*(zd a, double s) { return zd{a.x * s, a.y * s}; }
__host__ __device__ inline zd operator*(double s, zd a) { return zd{a.x * s, a.y * s}; }
__host__ __device__ inline zd conj(zd a) { return zd{a.x, -a.y}; }
__host__ __device__ inline zd operator/(zd a, zd b) {
  double d = b.x * b.x + b.y * b.y;
  return zd{(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}
__host__ __device__ inline zd operator/(zd a, double s) { return zd{a.x / s, a.y / s}; }
__host__ __device__ inline zd operator/(double s, zd b) { return zd{s, 0.0} / b; }
__host__ __device__ inline zd zcx(const cplx &c) { return zd{c.re, c.im}; }
// i*a
__host__ __device__ inline zd mul_i(zd a) { return zd{-a.y, a.x}; }

__host__ __device__ inline zf operator+(zf a, zf b) { return zf{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline zf operator-(zf a, zf b) { return zf{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline zf operator-(zf a) { return zf{-a.x, -a.y}; }
__host__ __device__ inline zf operator*(zf a, zf b) {
  return zf{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__host__ __device__ inline zf operator*(zf a, float s) { return zf{a.x * s, a.y * s}; }
__host__ __device__ inline zf conj(zf a) { return zf{a.x, -a.y}; }
__host__ __device__ inline zd widen(zf a) { return zd{(double)a.x, (double)a.y}; }
__host__ __device__ inline zd widen(zd a) { return a; }

// value conversion to the kernel's complex type
template <class T> __host__ __device__ inline T cvt(zd a);
template <> __host__ __device__ inline zd cvt<zd>(zd a) { return a; }
template <> __host__ __device__ inline zf cvt<zf>(zd a) { return zf{(float)a.x, (float)a.y}; }

// Near-field fill arguments (one struct, passed by value).
struct NearArgs {
  const FnRec *fn;
  const ElemRec *er;
  const double *rref;
  const int *perm;
  const int *box_first;
  const int *box_count;
  const int *e2_off;
  const int *e2_list;
  const int64_t *near_off;
  int nbl;
  int form;
  int ngp;
  int edge_swap;  // PMCHWT family + JMCFIE use the 12x12 edge-adjacent Kt
  Material mat;
  double qa[12], qb[12], qw[12];
  double qa12[12], qb12[12], qw12[12];
};

struct NearTask {
  int64_t g;  // block index
  int p;      // triangle pair (t-1)*2+(s-1)
  int pad;
};

// launch wrappers (spk_kernels.hip)
void launch_near_classify(const NearArgs &a, int64_t nnz, uint8_t *cls, int *cnt_a, int *cnt_b,
                          hipStream_t st);
void launch_near_emit(const NearArgs &a, int64_t nnz, const uint8_t *cls, const int *base_a,
                      const int *base_b, NearTask *ta, NearTask *tb, hipStream_t st);
void launch_near_edge(const NearArgs &a, int ntask, const NearTask *tb, zd *kt, hipStream_t st);
void launch_near_singular(const NearArgs &a, int ntask, const NearTask *ta, const uint8_t *cls,
                          const int *base_b, const zd *kt, zd *zz, hipStream_t st);
void launch_near_regular(const NearArgs &a, int64_t nnz, const uint8_t *cls, const int *base_a,
                         const zd *zz, zd *blocks, hipStream_t st);
hipError_t exclusive_scan_int(const int *in, int *out, int64_t n, void **tmp, size_t *tmp_bytes,
                        hipStream_t st);

// MLFMM (FP64)
struct MlfmmArgs {
  int N, S, nbl;
  const int *perm;        // position -> edge
  const int *box_first, *box_count;
};
void launch_patterns(int N, int S, int ngp, const FnRec *fn, const int *perm, const int *leaf_box,
                     const double *leaf_centre, const double *khat, const double *tth,
                     const double *tph, const double *qa, const double *qb, const double *qw,
                     zd k1, zd k2, zd *rt, hipStream_t st);
// conj_rt = false: the transposed operator's aggregation (patterns not conjugated)
template <class T>
void launch_agg(int nbl, int S, int N, const int *box_first, const int *box_count,
                const int *perm, const T *rt, const T *x, T *C, hipStream_t st,
                bool conj_rt = true);
template <class T>
void launch_fold_t(int nbl, int S, const T *C, zd kd1, zd kd2, zd my1, zd my2, zd ep1, zd ep2,
                   T *D, hipStream_t st);
template <class T>
void launch_scale_halves(int N, zd sc1, zd sc2, const T *x, T *out, hipStream_t st);
template <class T>
void launch_near_transpose(int nbl, int max_rows, const int *box_first, const int *box_count,
                           const int *e2_off, const int *e2_list, const int64_t *near_off,
                           const int *perm, const int *pos, const int *leaf_box, const T *in,
                           T *out, hipStream_t st);
template <class T>
void launch_ss(int nb, int S, const int *child, const T *Cc, const T *up, T *P, hipStream_t st);
template <class T>
void launch_m2l(int nb, int S, const int *i4_off, const int *i4_box, const int *i4_slot,
                const T *B, const T *tlt, T *Dt, hipStream_t st);
template <class T>
void launch_rr(int nb, int S, const int *parent, const int *octant, const T *Dp, const T *down,
               T *Dc, hipStream_t st);
template <class T>
void launch_fold(int nbl, int S, const T *D, zd kd1, zd kd2, zd my1, zd my2, zd ep1, zd ep2,
                 T *W, hipStream_t st);
// conj_rt: the transposed operator's disaggregation (conjugated patterns)
template <class T>
void launch_disagg(int ntask, int N, int S, const int *task, const int *perm,
                   const int *leaf_box_of_pos, const T *rt, const T *W, zd sc1, zd sc2, T *y,
                   hipStream_t st, bool conj_rt = false);
template <class T>
void launch_near_spmv_rows(int nwg, int N, const int *row_box, const int *row_r0,
                           const int *box_first, const int *box_count, const int *e2_off,
                           const int *e2_list, const int64_t *near_off, const int *perm,
                           const T *blocks, const T *x, const T *yadd, T *y, hipStream_t st);
void launch_to_zf(int64_t n, const zd *a, zf *b, hipStream_t st);
// TL table of one level from the host's per-slot series factors
void launch_tl_level(int S, int K, const double *khat, const double *trf, const double *rhat,
                     const zd *coef, const int *live, zd *tlt, hipStream_t st);
constexpr int kSpmvRowsPerWG = 16;
// Complex64 sums of many terms (aggregation over a box's edges, M2L over
// the I4 list, disaggregation over samples) add each group of kSumBlock
// terms from zero before adding it to the running sum: the far field
// amplifies the rounding of these sums, and the running sum then takes one
// rounding per group instead of one per term.  Complex128 keeps the plain
// sequential order (one group).
template <class T>
inline constexpr int kSumBlock = 1 << 30;
template <>
inline constexpr int kSumBlock<zf> = 16;
// edges per disaggregation task (one wavefront, one leaf box)
constexpr int kDisaggEdges = 4;
// right-hand sides per batched matvec launch; edges per batched
// disaggregation tile
constexpr int kBatchMax = 16;
constexpr int kBatchTile = 64;
// workgroups per disaggregation tile, each summing a share of the samples
constexpr int kBatchSplit = 4;
// arbitrary (m, n) pairs, one thread per pair (parity probe)
void launch_near_pairs(const NearArgs &a, int npairs, const int *m, const int *n, zd *out,
                       hipStream_t st);

// plane-wave right-hand side, one thread per RWG edge (reference order):
// b[m] = sum_q ca[q] S_q(m), b[m+N] = sum_q cb[q] S_q(m) with the four
// tested incident-field integrals S = (<f,E>, <f,H>, <f,n x E>, <f,n x H>)
struct RhsArgs {
  int N, ng;
  double qa[20], qb[20], qw[20];
  zd k1;
  double kin[3], ein[3];
  double c0my;  // my1 * c0
  zd ca[4], cb[4];
};
void launch_rhs_plane(const RhsArgs &a, const FnRec *fn, zd *b, hipStream_t st);

// scattered E and H at observation points (Integration_scattered_field_tri)
struct ScatArgs {
  int N, npts;
  double qa[3], qb[3], qw[3];
  zd k1, iwmy, iweps, inv_iwmy;
};
void launch_scattered(const ScatArgs &a, const FnRec *fn, const zd *x, const double *r, zd *E,
                      zd *H, hipStream_t st);

// adjoint gradient with respect to node heights (central differences of
// the local quantities, node moved by +-dz along z)
struct AdjArgs {
  int N, nnode, nslot, normal_convention;
  double dz;
  const double *pp;        // node coordinates [3 * np]
  const int *side_nodes;   // per RWG side: p1, p2, free vertex, element nodes (0-based) [6 * 2N]
  const int *leaf_box;     // edge -> leaf box
  const int *node;         // requested node ids (0-based) [nnode]
  // near kernel: node k's slots are aff_off[k] .. aff_off[k + 1] - 2 of
  // aff_edge / task_off, followed by one end entry (task_off = the node's
  // task count); the per-slot kernels index aff_edge / aff_node by slot
  const int *aff_off;
  const int *aff_edge;
  const int *aff_node;
  const int64_t *task_off;
};
struct AdjFar {
  int S, ngp;
  const double *khat, *tth, *tph, *leaf_centre;
  double qa[12], qb[12], qw[12];
  zd k1, k2, sc1, sc2;
};
void launch_adj_near(const NearArgs &A, const AdjArgs &G, int groups, const zd *nu, const zd *I,
                     zd *partial, hipStream_t st);
void launch_adj_far(const AdjArgs &G, const AdjFar &F, const zd *nu, const zd *I, const zd *W,
                    const zd *Bt, zd *out, hipStream_t st);
void launch_adj_rhs(const AdjArgs &G, const RhsArgs &a, const zd *nu, zd *out, hipStream_t st);
void launch_scattered_t(const ScatArgs &a, const FnRec *fn, const double *r, const double *w,
                        const zd *E, zd *c, hipStream_t st);
void launch_adj_field(const AdjArgs &G, const ScatArgs &a, const double *r, const double *w,
                      const zd *E, const zd *I, zd *out, hipStream_t st);

// Jacobi preconditioner: diag = [Z_mm(1,1); Z_mm(2,2)] (reference order), y = x / diag
void launch_jacobi_diag(int N, const int *perm, const int *leaf_box_of_pos, const int *box_first,
                        const int *e2_off, const int *e2_list, const int64_t *near_off,
                        const zd *blocks, zd *diag, hipStream_t st);
void launch_div_diag(int64_t n, const zd *diag, const zd *x, zd *y, hipStream_t st);

// BLAS-1 (deterministic reductions)
int reduce_blocks(int64_t n);
template <class T>  // sum conj(x) y accumulated in complex128
void launch_zdotc(int64_t n, const T *x, const T *y, zd *partial, zd *out, hipStream_t st);
void launch_axpy_up(int64_t n, zd a, const zf *d, zd *x, hipStream_t st);        // x += a d
void launch_scale_down(int64_t n, double s, const zd *x, zf *y, hipStream_t st);  // y = s x
void launch_scal_copy_f(int64_t n, double s, const zf *x, zf *y, hipStream_t st);
void launch_zaxpy(int64_t n, zd a, const zd *x, zd *y, hipStream_t st);
void launch_zaxpy_dev(int64_t n, const zd *a, double sgn, const zd *x, zd *y, hipStream_t st);
void launch_zscal_copy(int64_t n, zd a, const zd *x, zd *y, hipStream_t st);
void launch_zsub(int64_t n, const zd *b, const zd *x, zd *y, hipStream_t st);  // y = b - x
void launch_mgs_step(int64_t n, const zd *vprev, const zd *hprev, const zd *vj, zd *w,
                     zd *partial, zd *hout, hipStream_t st);
// one MGS pass over V[:, 0:k] in one cooperative launch (see the kernel).
// slots: (k+1)*G records of mgs_sweep_slot_bytes(), zeroed once at
// allocation; seq: a number above every earlier launch's on these slots;
// ctl[0]: abort flag (zero); nrm = (|w|^2, abort flag)
int mgs_sweep_blocks(int64_t n);  // 0: n too large for the register slices
size_t mgs_sweep_slot_bytes();
hipError_t launch_mgs_sweep(int G, int64_t n, int k, const zd *V, int64_t ldv, zd *w, void *slots,
                            zd *hout, zd *nrm, unsigned long long *ctl, unsigned long long seq,
                            hipStream_t st);
// out[v*k + j] = <V_j, w_v> for j < k, v = 0 (w0) and, when w1 is not null,
// v = 1 (w1); partial: gemv_c_partial_elems(n, k, 1 or 2) complex128
template <class T>
void launch_gemv_c(int64_t n, int k, const T *V, int64_t ldv, const T *w0, const T *w1,
                   zd *partial, zd *out, hipStream_t st);
size_t gemv_c_partial_elems(int64_t n, int kmax, int nv);
// w -= V h and out = V h over k columns; partial: gemv_n_partial_elems(n, k)
template <class T>
void launch_gemv_n_sub(int64_t n, int k, const T *V, int64_t ldv, const zd *h, T *w, zd *partial,
                       hipStream_t st);
template <class T>
void launch_gemv_n(int64_t n, int k, const T *V, int64_t ldv, const zd *h, zd *out, zd *partial,
                   hipStream_t st);
size_t gemv_n_partial_elems(int64_t n, int kmax);
// the matvec for R <= kBatchMax right-hand sides (spk_batch.hip); xs, ys:
// device arrays of R vector pointers; k-space arrays [box][it][s][r]
// XI[(m * 2 + cur) * RP + r] = xs[r][m + cur N], RP = R rounded up to a power of two
// and XP[(p * 2 + cur) * RP + r] = xs[r][perm[p] + cur N] (edge-position order)
template <class T>
void launch_batch_interleave(int N, int R, const int *perm, const T *const *xs, T *XI, T *XP,
                             hipStream_t st);
template <class T>
void launch_batch_agg(int nbl, int S, int N, int R, const int *box_first, const int *box_count,
                      const int *perm, const T *rt, const T *XI, const T *XP, T *C,
                      hipStream_t st);
template <class T>
void launch_batch_ss(int nb, int S, int R, const int *child, const T *Cc, const T *up, T *P,
                     hipStream_t st);
template <class T>
void launch_batch_m2l(int nb, int S, int R, const int *i4_off, const int *i4_box,
                      const int *i4_slot, const T *B, const T *tlt, T *Dt, hipStream_t st);
template <class T>
void launch_batch_rr(int nb, int S, int R, const int *parent, const int *octant, const T *Dp,
                     const T *down, T *Dc, hipStream_t st);
template <class T>
void launch_batch_fold(int nbl, int S, int R, const T *D, zd kd1, zd kd2, zd my1, zd my2, zd ep1,
                       zd ep2, T *W, hipStream_t st);
// tiles of up to kBatchTile consecutive edges of one leaf box (mx tiles: up
// to 128, for complex64 with 9..16 right-hand sides on the matrix cores)
// yp: kBatchSplit partial slices of R x 2N, summed into yf
template <class T>
void launch_batch_disagg(int ntiles, int N, int S, int R, const int *tile_box, const int *tile_p0,
                         const int *box_first, const int *box_count, const int *perm, const T *rt,
                         const T *W, zd sc1, zd sc2, T *yp, T *yf, int mx_tiles,
                         const int *mx_box, const int *mx_p0, hipStream_t st);
// tiles (box, first row) of kSpmvRowsPerWG rows; complex64 with 9..16
// right-hand sides runs on the matrix cores, on the bf16 copy of the
// blocks when blocks_bf is given
template <class T>
void launch_batch_near(int nwg, int N, int R, const int *row_box, const int *row_r0,
                       const int *box_first, const int *box_count, const int *e2_off,
                       const int *e2_list, const int64_t *near_off, const int *perm,
                       const T *blocks, const T *XI, const T *yf, T *const *ys,
                       hipStream_t st, const uint32_t *blocks_bf = nullptr);
// the near-field preconditioner's product for 9..16 right-hand sides (XI as
// launch_batch_interleave writes it; XB: scratch of N * 32 words for x in
// bf16) on the fp8 copy of the blocks, on the matrix cores; workgroups = the
// row tiles (box, first row) of kSpmvRowsPerWG rows
int near_mx_tiles_per_group();
// levels of different truncation numbers (spk_kernels.hip)
void launch_resample(int nbox, int Sin, int Sout, const int *rowptr, const int *col,
                     const double *val, const int *pairs, const zd *X, zd *Y, bool accumulate,
                     hipStream_t st);
void launch_shift_write(int nb, int S, const int *parent, const int *octant, const zd *P,
                        const zd *down, zd *T, bool folded, hipStream_t st);
void launch_rr_folded(int nb, int S, const int *parent, const int *octant, const zd *Wp,
                      const zd *down, zd *Wc, hipStream_t st);

void launch_batch_near_f8(int nwg, int N, int R, const int *row_box, const int *row_r0,
                          int ngrp, const int *grp_box, const int *grp_r0,
                          const int *box_first, const int *box_count, const int *e2_off,
                          const int *e2_list, const int64_t *near_off, const int *perm,
                          const uint32_t *blocks, const float *scale, const zf *XI,
                          uint32_t *XB, zf *const *ys, int tpw, hipStream_t st);

// one ICGS pass over P bases (coefficients into hs[r][0..ks[r]), |w_r|^2
// into hs[r][ks[r]]); partial: batch_orth_partial_elems(n, P, maxk)
template <class T>
void launch_batch_orth(int64_t n, int P, int maxk, const int *ks, const T *const *Vs, T *const *ws,
                       zd *const *hs, zd *partial, hipStream_t st);
size_t batch_orth_partial_elems(int64_t n, int P, int maxk);
// the near field in bf16 for the near-field preconditioner: one 32-bit word
// per complex value (real part in the low half), blocks as in the complex64
// array; the SpMV takes and returns complex64 vectors (no far-field add)
void launch_to_bf(int64_t n, const zf *a, uint32_t *b, hipStream_t st);
void launch_near_spmv_bf(int nwg, int N, const int *row_box, const int *row_r0,
                         const int *box_first, const int *box_count, const int *e2_off,
                         const int *e2_list, const int64_t *near_off, const int *perm,
                         const uint32_t *blocks, const zf *x, zf *y, hipStream_t st);
// the near field in fp8 (OCP E4M3FN) for the single-vector preconditioner:
// 8 bytes per 2x2 block (Z0r, Z0i, Z1r, Z1i, Z2r, Z2i, Z3r, Z3i), scale[4 p + k]
// = the scale of component k of the row at position p (its largest |re|,
// |im| maps to 448); one workgroup per row position
void launch_near_to_f8(int N, const int *leaf_box_of_pos, const int *box_first, const int *e2_off,
                       const int64_t *near_off, const zf *blocks, uint32_t *out, float *scale,
                       hipStream_t st);
void launch_near_spmv_f8(int nwg, int N, const int *row_box, const int *row_r0,
                         const int *box_first, const int *box_count, const int *e2_off,
                         const int *e2_list, const int64_t *near_off, const int *perm,
                         const uint32_t *blocks, const float *scale, const zf *x, zf *y,
                         hipStream_t st);
// one Arnoldi step (one classical Gram-Schmidt pass, three launches) of P
// near-field inner solves after w_s = A U_s[:, j]: h_s = U_s[:, 0..j]^H w_s
// (also in hb + s * ldh), w_s -= U_s h_s, H_s[:, j] = (h_s, |w_s|),
// U_s[:, j+1] = w_s / |w_s|.  Slot s: U + s * ldu (column c at c * n),
// W + s * n, H + s * (kin + 1) * kin (column-major, leading dimension
// kin + 1); partial: arnoldi_partial_elems(n, P, kin); kin <= 64
template <class T>
void launch_arnoldi_cgs(int64_t n, int P, int j, int kin, T *U, size_t ldu, T *W, zd *hb, int ldh,
                        zd *H, zd *partial, hipStream_t st);
size_t arnoldi_partial_elems(int64_t n, int P, int kin);

// streaming read of a buffer (bandwidth probe)
void launch_stream_read(const void *a, size_t bytes, double *out, hipStream_t st);
// one-reduce MGS: row k1 of T = (I + L)^-1 from tg = T^H g
void launch_icwy_row(int k1, zd *Tm, int ldt, const zd *tg, hipStream_t st);
void launch_zadd(int64_t n, const zd *a, zd *y, hipStream_t st);  // y += a

// Block-Jacobi preconditioner on groups of leaf boxes (the children of one
// box a level up): group g holds positions [p0[g], p0[g] + cnt[g]); its
// dense matrix A_g (column-major, leading dimension nmax, unknowns ordered
// [x_e of the group; x_h of the group], identity beyond 2 cnt[g]) comes
// from the near blocks.  A must be zeroed before the extraction.
void launch_blu_extract(int ngroups, int nmax, const int *p0, const int *cnt, const int *perm,
                        const int *pos, const int *leafbox_of_pos, const int *box_first,
                        const int *e2_off, const int *e2_list, const int64_t *near_off,
                        const zd *blocks, zd *A, hipStream_t st);
// x_g = v gathered into the group layout (zero padded), z = y_g scattered back
void launch_blu_gather(int ngroups, int nmax, int N, const int *p0, const int *cnt,
                       const int *perm, const zd *v, zd *X, hipStream_t st);
void launch_blu_scatter(int ngroups, int nmax, int N, const int *p0, const int *cnt,
                        const int *perm, const zd *Y, zd *z, hipStream_t st);

}  // namespace spk

// spk_batch.hip -- the MLFMM matvec applied to R <= 16 right-hand sides at
// once (incidence angles of one mesh): every operator array (near blocks,
// patterns, TL and phase tables) is read once per launch for all R vectors.
//
// Batched k-space arrays keep the right-hand side fastest: a box's
// coefficients are [it][s][r], index (it * S + s) * R + r, so the
// elementwise translation kernels run over u = s * R + r with the table
// index s = u / R.  Vectors cross the launchers as device arrays of R
// pointers (each a 2N complex vector in the reference order).
// The arithmetic of each kernel is that of its single-vector counterpart in
// spk_kernels.hip; only the order of the near-field and disaggregation sums
// differs.  Complex64 batches of 9..16 run the near field, aggregation and
// disaggregation on the matrix cores (v_mfma_f32_16x16x4f32, which rounds as
// a chain of four fused multiply-adds): the aggregation and disaggregation
// sum each group of 16 terms from zero before adding it to the running sum.
#include <algorithm>
#include <type_traits>

#include <cstdlib>

#include "spk_device.h"

namespace spk {

namespace {

inline unsigned bnblk(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

template <class T>
__device__ inline T bshfl_down(T v, int off) {
  v.x = __shfl_down(v.x, off, 64);
  v.y = __shfl_down(v.y, off, 64);
  return v;
}

}  // namespace

// The R input vectors interleaved: XI[(m * 2 + cur) * RP + r] = x_r[m + cur N]
// (zero for r >= R), so that a column's 2 RP values are one contiguous run.
// XP: the same in edge-position order, XP[(p * 2 + cur) * RP + r] = x_r[perm[p] + cur N].
template <class T, int RP>
__global__ void kb_interleave(int N, int R, const int *perm, const T *const *xs, T *XI, T *XP) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)2 * N * RP) return;
  const int r = (int)(t % RP);
  const int64_t mc = t / RP;
  const int m = (int)(mc >> 1), cur = (int)(mc & 1);
  XI[t] = r < R ? xs[r][m + (int64_t)cur * N] : T{0, 0};
  XP[t] = r < R ? xs[r][perm[m] + (int64_t)cur * N] : T{0, 0};
}

// Aggregation: C[b][it][s][r] = sum_{edges of b} conj(RT[p][q][s]) x_r(cur),
// it = comp*4 + cur*2 + k.  Workgroup = (256 samples, leaf box, q); the
// box's x values are staged in LDS 64 edges at a time and read as
// broadcasts; each thread keeps 2R accumulators.  RP = R rounded up to a
// power of two (lanes beyond R are zero and not written).
template <class T, int RP>
__global__ void __launch_bounds__(256) kb_agg(int S, int N, int R, const int *box_first,
                                              const int *box_count, const T *rt, const T *XP,
                                              T *C) {
  __shared__ T xl[64 * 2 * RP];
  const int s = blockIdx.x * 256 + threadIdx.x;
  const int b = blockIdx.y, q = blockIdx.z, k = q >> 1, comp = q & 1;
  const int p0 = box_first[b], cnt = box_count[b];
  T ae[RP], ah[RP];
#pragma unroll
  for (int r = 0; r < RP; ++r) ae[r] = ah[r] = T{0, 0};
  for (int e0 = 0; e0 < cnt; e0 += 64) {
    const int ne = min(64, cnt - e0);
    __syncthreads();
    // the box's edges are consecutive positions: one contiguous run of XP
    for (int i = threadIdx.x; i < ne * 2 * RP; i += 256) xl[i] = XP[(size_t)(p0 + e0) * 2 * RP + i];
    __syncthreads();
    if (s < S) {
      for (int g0 = 0; g0 < ne; g0 += kSumBlock<T>) {
        T ge[RP], gh[RP];
#pragma unroll
        for (int r = 0; r < RP; ++r) ge[r] = gh[r] = T{0, 0};
        const int g1 = min(ne, g0 + kSumBlock<T>);
        for (int e = g0; e < g1; ++e) {
          const T f = conj(rt[((size_t)(p0 + e0 + e) * 4 + q) * S + s]);
          const T *xe = xl + e * 2 * RP;
#pragma unroll
          for (int r = 0; r < RP; ++r) {
            ge[r] = ge[r] + f * xe[r];
            gh[r] = gh[r] + f * xe[RP + r];
          }
        }
#pragma unroll
        for (int r = 0; r < RP; ++r) {
          ae[r] = ae[r] + ge[r];
          ah[r] = ah[r] + gh[r];
        }
      }
    }
  }
  if (s < S) {
    T *Cb = C + (size_t)b * 8 * S * R;
#pragma unroll
    for (int r = 0; r < RP; ++r)
      if (r < R) {
        Cb[((size_t)(comp * 4 + 0 + k) * S + s) * R + r] = ae[r];
        Cb[((size_t)(comp * 4 + 2 + k) * S + s) * R + r] = ah[r];
      }
  }
}

typedef float mxf4 __attribute__((ext_vector_type(4)));
#define SPK_MFMA(a, b, c) __builtin_amdgcn_mfma_f32_16x16x4f32((a), (b), (c), 0, 0, 0)

// Aggregation for 9..16 right-hand sides in complex64 on the matrix cores:
// per (leaf box, q) C[s][(cur, r)] = sum_p conj(RT[p][q][s]) x_r(p, cur) is a
// product of the box's pattern rows (samples x edges) and its x values
// (edges x 32).  Workgroup = (256 samples, box, q), wave = 64 samples as 4
// tiles of 16: tile 2h + par holds the samples sw + 32 h + 2 i + par, so a
// lane's one 16-byte load gives the A values of two tiles.  A k-step covers
// 4 edges (lane group g = edge), B[g][n] = x_n of that edge from LDS.  With
// a = conj(f): re += fr xr + fi xi, im += fr xi - fi xr, for cur = 0 and 1.
// Samples and edges past the end are clamped on load (their x is zero in
// LDS, their results are not written).
__global__ void __launch_bounds__(256) kb_agg_mx(int S, int R, const int *box_first,
                                                 const int *box_count, const zf *rt,
                                                 const zf *XP, zf *C) {
  __shared__ float4 xl[64 * 16];
  const int b = blockIdx.y, q = blockIdx.z, k = q >> 1, comp = q & 1;
  const int p0 = box_first[b], cnt = box_count[b];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int sw = blockIdx.x * 256 + 64 * wv;  // first sample of this wave
  const bool active = sw < S;                 // uniform over the wave
  // the two sample pairs of this lane (16-byte aligned: S is even)
  const int sa = min(sw + 2 * m, S - 2), sb = min(sw + 32 + 2 * m, S - 2);
  mxf4 acc[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[t][u] = mxf4{0.f, 0.f, 0.f, 0.f};
  float4 fa[4], fb[4];
  auto load = [&](int e0, float4 *a, float4 *bb) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = min(e0 + 4 * j + g, cnt - 1);
      const zf *f = rt + ((size_t)(p0 + e) * 4 + q) * S;
      a[j] = *reinterpret_cast<const float4 *>(f + sa);
      bb[j] = *reinterpret_cast<const float4 *>(f + sb);
    }
  };
  if (active && cnt > 0) load(0, fa, fb);
  for (int e0 = 0; e0 < cnt; e0 += 64) {
    const int ne = min(64, cnt - e0);
    __syncthreads();
    {
      const int n = threadIdx.x & 15;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = (int)(threadIdx.x + 256 * u) >> 4;
        const size_t pe = (size_t)(p0 + e0 + min(e, ne - 1)) * 2;
        const zf x0 = XP[pe * 16 + n], x1 = XP[(pe + 1) * 16 + n];
        xl[threadIdx.x + 256 * u] = e < ne ? float4{x0.x, x0.y, x1.x, x1.y} : float4{0.f, 0.f, 0.f, 0.f};
      }
    }
    __syncthreads();
    if (!active) continue;
    for (int i = 0; i < ne; i += 16) {
      float4 na[4], nb[4];
      load(e0 + i + 16, na, nb);
      __builtin_amdgcn_sched_barrier(0);
      // the group's 16 edges are summed from zero and then added to the
      // running sums, so the running sums take one rounding per group
      mxf4 gs[4][4];
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u) gs[t][u] = mxf4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float4 x = xl[(i + 4 * j + g) * 16 + m];
        const float n0 = -x.x, n1 = -x.z;
        const float ar[4] = {fa[j].x, fa[j].z, fb[j].x, fb[j].z};
        const float ai[4] = {fa[j].y, fa[j].w, fb[j].y, fb[j].w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          gs[t][0] = SPK_MFMA(ar[t], x.x, gs[t][0]);
          gs[t][1] = SPK_MFMA(ar[t], x.y, gs[t][1]);
          gs[t][2] = SPK_MFMA(ar[t], x.z, gs[t][2]);
          gs[t][3] = SPK_MFMA(ar[t], x.w, gs[t][3]);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          gs[t][0] = SPK_MFMA(ai[t], x.y, gs[t][0]);
          gs[t][1] = SPK_MFMA(ai[t], n0, gs[t][1]);
          gs[t][2] = SPK_MFMA(ai[t], x.w, gs[t][2]);
          gs[t][3] = SPK_MFMA(ai[t], n1, gs[t][3]);
        }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[t][u] += gs[t][u];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        fa[j] = na[j];
        fb[j] = nb[j];
      }
    }
  }
  // accumulator element v of lane (g, n) in tile 2h + par: sample sw + 32h + 2(4g + v) + par
  const int n = m;
  if (!active || n >= R) return;
  zf *Cb = C + (size_t)b * 8 * S * R;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int s = sw + 32 * (t >> 1) + 2 * (4 * g + v) + (t & 1);
      if (s < S) {
        Cb[((size_t)(comp * 4 + 0 + k) * S + s) * R + n] = zf{acc[t][0][v], acc[t][1][v]};
        Cb[((size_t)(comp * 4 + 2 + k) * S + s) * R + n] = zf{acc[t][2][v], acc[t][3][v]};
      }
    }
}

// translation_SS over u = s*R + r
template <class T>
__global__ void kb_ss(int S, int R, const int *child, const T *Cc, const T *up, T *P) {
  const int SR = S * R;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (u >= SR) return;
  const int s = u / R;
  T acc[8];
  for (int it = 0; it < 8; ++it) acc[it] = T{0, 0};
  for (int o = 0; o < 8; ++o) {
    const int c = child[8 * b + o];
    if (c < 0) continue;
    const T u0 = up[((size_t)o * 2 + 0) * S + s], u1 = up[((size_t)o * 2 + 1) * S + s];
    const T *src = Cc + (size_t)c * 8 * SR;
#pragma unroll
    for (int it = 0; it < 8; ++it) acc[it] = acc[it] + src[(size_t)it * SR + u] * ((it & 1) ? u1 : u0);
  }
  T *dst = P + (size_t)b * 8 * SR;
  for (int it = 0; it < 8; ++it) dst[(size_t)it * SR + u] = acc[it];
}

// translation_SR over u.  XCD-aware order: the 8 XCDs take workgroups
// round robin, so workgroup L is given the logical tile (L % 8) * per + L / 8;
// each XCD then sweeps a contiguous run of target boxes for one sample
// chunk, whose I4 sources overlap and stay in that XCD's L2.
template <class T>
__global__ void kb_m2l(int S, int R, int nb, int gx, const int *i4_off, const int *i4_box,
                       const int *i4_slot, const T *B, const T *tlt, T *Dt) {
  const int SR = S * R;
  const int per = gridDim.x / 8;
  const int lin = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (lin >= gx * nb) return;
  const int u = (lin / nb) * blockDim.x + threadIdx.x;
  const int a = lin % nb;
  if (u >= SR) return;
  const int s = u / R;
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
      const T *src = B + (size_t)b * 8 * SR;
#pragma unroll
      for (int it = 0; it < 8; ++it) gs[it] = gs[it] + src[(size_t)it * SR + u] * ((it & 1) ? t1 : t0);
    }
#pragma unroll
    for (int it = 0; it < 8; ++it) acc[it] = acc[it] + gs[it];
  }
  T *dst = Dt + (size_t)a * 8 * SR;
  for (int it = 0; it < 8; ++it) dst[(size_t)it * SR + u] = acc[it];
}

// translation_RR over u (in place on the child level's D)
template <class T>
__global__ void kb_rr(int S, int R, const int *parent, const int *octant, const T *Dp,
                      const T *down, T *Dc) {
  const int SR = S * R;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (u >= SR) return;
  const int s = u / R;
  const int p = parent[b], o = octant[b];
  const T d0 = down[((size_t)o * 2 + 0) * S + s], d1 = down[((size_t)o * 2 + 1) * S + s];
  const T *src = Dp + (size_t)p * 8 * SR;
  T *dst = Dc + (size_t)b * 8 * SR;
#pragma unroll
  for (int it = 0; it < 8; ++it)
    dst[(size_t)it * SR + u] = dst[(size_t)it * SR + u] + src[(size_t)it * SR + u] * ((it & 1) ? d1 : d0);
}

// the folded coefficients of the disaggregation over u
template <class T>
__global__ void kb_fold(int S, int R, const T *D, T kd1, T kd2, T my1, T my2, T ep1, T ep2, T *W) {
  const int SR = S * R;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (u >= SR) return;
  const T *Db = D + (size_t)b * 8 * SR;
  T *Wb = W + (size_t)b * 8 * SR;
  for (int k = 0; k < 2; ++k) {
    const T kd = k ? kd2 : kd1, my = k ? my2 : my1, ep = k ? ep2 : ep1;
    const T Dae = Db[(size_t)(0 + k) * SR + u], Dah = Db[(size_t)(2 + k) * SR + u];
    const T Dbe = Db[(size_t)(4 + k) * SR + u], Dbh = Db[(size_t)(6 + k) * SR + u];
    Wb[(size_t)(k * 4 + 0) * SR + u] = Dbh * kd + Dae * my;
    Wb[(size_t)(k * 4 + 1) * SR + u] = -(Dah * kd) + Dbe * my;
    Wb[(size_t)(k * 4 + 2) * SR + u] = -(Dbe * kd) + Dah * ep;
    Wb[(size_t)(k * 4 + 3) * SR + u] = Dae * kd + Dbh * ep;
  }
}

// Disaggregation as a small GEMM per box: workgroup = (leaf box, tile of up
// to 64 consecutive edges), 16 RP threads; thread = (group of 4 edges of the
// tile, right-hand side).  Samples are staged in LDS in chunks of kBatchSC:
// the tile's pattern rows and the box's folded W for all R; a W value read
// from LDS serves the thread's 4 edges.
// yf[r][m] = sc1 * sum_{k,s} (fa W1a + fb W1b), yf[r][m + N] = sc2 * sum (fa W2a + fb W2b).
constexpr int kBatchSC = 16;
template <class T, int RP>
__global__ void __launch_bounds__(256) kb_disagg(int N, int S, int R, const int *tile_box,
                                                 const int *tile_p0, const int *box_first,
                                                 const int *box_count, const int *perm,
                                                 const T *rt, const T *W, T sc1, T sc2, int sblk,
                                                 T *yp) {
  constexpr int NT = 16 * RP, EPT = 4;
  __shared__ T wl[8 * kBatchSC * RP];
  __shared__ T fl[kBatchTile * 4 * kBatchSC];
  const int b = tile_box[blockIdx.x], p0 = tile_p0[blockIdx.x];
  const int cnt = min(kBatchTile, box_first[b] + box_count[b] - p0);
  const int eg = threadIdx.x / RP, r = threadIdx.x % RP;  // edges eg*4 .. eg*4+3
  const int SR = S * R;
  const T *Wb = W + (size_t)b * 8 * SR;
  T se[EPT], sh[EPT];
#pragma unroll
  for (int e = 0; e < EPT; ++e) se[e] = sh[e] = T{0, 0};
  // this workgroup's sample range; its partial sums go to slice blockIdx.y
  const int s_lo = blockIdx.y * sblk, s_hi = min(S, s_lo + sblk);
  T *yf = yp + (size_t)blockIdx.y * R * 2 * N;
  for (int s0 = s_lo; s0 < s_hi; s0 += kBatchSC) {
    const int ns = min(kBatchSC, s_hi - s0);
    __syncthreads();
    // W[j][s0 + i][r] -> wl[(j * kBatchSC + i) * RP + r]
    for (int t = threadIdx.x; t < 8 * kBatchSC * RP; t += NT) {
      const int j = t / (kBatchSC * RP), i = (t / RP) % kBatchSC, rr = t % RP;
      wl[t] = (i < ns && rr < R) ? Wb[(size_t)j * SR + (size_t)(s0 + i) * R + rr] : T{0, 0};
    }
    // RT[p0 + e][row][s0 + i] -> fl[(e * 4 + row) * kBatchSC + i]
    for (int t = threadIdx.x; t < kBatchTile * 4 * kBatchSC; t += NT) {
      const int e = t / (4 * kBatchSC), row = (t / kBatchSC) & 3, i = t % kBatchSC;
      fl[t] = (e < cnt && i < ns) ? rt[((size_t)(p0 + e) * 4 + row) * S + s0 + i] : T{0, 0};
    }
    __syncthreads();
    if (eg * EPT < cnt) {
      // complex64: the chunk's samples are summed from zero, then added
      constexpr bool blk = kSumBlock<T> <= kBatchSC;
      T ge[EPT], gh[EPT];
#pragma unroll
      for (int e = 0; e < EPT; ++e) ge[e] = gh[e] = T{0, 0};
      T *te = blk ? ge : se, *th = blk ? gh : sh;
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        const T *w1a = wl + (k * 4 + 0) * kBatchSC * RP + r, *w1b = wl + (k * 4 + 1) * kBatchSC * RP + r;
        const T *w2a = wl + (k * 4 + 2) * kBatchSC * RP + r, *w2b = wl + (k * 4 + 3) * kBatchSC * RP + r;
#pragma unroll 4
        for (int i = 0; i < kBatchSC; ++i) {
          const T wa = w1a[i * RP], wb = w1b[i * RP], wc = w2a[i * RP], wd = w2b[i * RP];
#pragma unroll
          for (int e = 0; e < EPT; ++e) {
            const T *f = fl + ((eg * EPT + e) * 4 + k * 2) * kBatchSC;
            const T a = f[i], bb = f[kBatchSC + i];
            te[e] = te[e] + a * wa + bb * wb;
            th[e] = th[e] + a * wc + bb * wd;
          }
        }
      }
      if constexpr (blk) {
#pragma unroll
        for (int e = 0; e < EPT; ++e) {
          se[e] = se[e] + ge[e];
          sh[e] = sh[e] + gh[e];
        }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < EPT; ++e) {
    const int el = eg * EPT + e;
    if (el < cnt && r < R) {
      const int m = perm[p0 + el];
      yf[(size_t)r * 2 * N + m] = se[e] * sc1;
      yf[(size_t)r * 2 * N + m + N] = sh[e] * sc2;
    }
  }
}

// Disaggregation for 9..16 right-hand sides in complex64 on the matrix
// cores: per leaf box y_r(p, e) = sum_{k,s} fa W(k,e,a) + fb W(k,e,b) with
// fa = RT[p][2k][s], fb = RT[p][2k+1][s] is a product of the box's pattern
// rows (edges x (k, s)) and its folded W ((k, s) x r).  Workgroup = (tile of
// up to 128 consecutive edges of one box, sample part blockIdx.y), wave = 16
// edges.  A k-step covers 4 samples: lane group g takes sample s + 4g + j in
// step j, so a lane's two 16-byte loads per pattern row serve 4 steps.  W of
// kMxSC samples is staged in LDS as [k][sample][half][r][2]: half 0 = the e
// pair (a, b), half 1 = the h pair.  Edges and samples past the end are
// clamped on load (their W is zero in LDS, their results are not written).
// yf slice blockIdx.y: [r][m] = sc1 * sum (e part), [r][m + N] = sc2 * sum (h part).
constexpr int kMxSC = 32;
__global__ void __launch_bounds__(512) kb_disagg_mx(int N, int S, int R, const int *tile_box,
                                                    const int *tile_p0, const int *box_first,
                                                    const int *box_count, const int *perm,
                                                    const zf *rt, const zf *W, zf sc1, zf sc2,
                                                    int sblk, zf *yp) {
  __shared__ float4 wl[2 * kMxSC * 2 * 16];
  const int b = tile_box[blockIdx.x], p0 = tile_p0[blockIdx.x];
  const int cnt = min(128, box_first[b] + box_count[b] - p0);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int ew = 16 * wv;
  const bool active = ew < cnt;  // uniform over the wave
  const zf *F = rt + (size_t)(p0 + min(ew + m, cnt - 1)) * 4 * S;
  const size_t SR = (size_t)S * R;
  const zf *Wb = W + (size_t)b * 8 * SR;
  const int s_lo = blockIdx.y * sblk, s_hi = min(S, s_lo + sblk);
  mxf4 er = {0.f, 0.f, 0.f, 0.f}, ei = er, hr = er, hi = er;
  // pattern rows 2k and 2k+1 at samples s + 4g .. s + 4g + 3 (pairs clamped, S is even)
  auto load = [&](int s, int k, float4 *a) {
    const int s1 = min(s + 4 * g, S - 2), s2 = min(s + 4 * g + 2, S - 2);
    const zf *fa = F + (size_t)(2 * k) * S, *fb = fa + S;
    a[0] = *reinterpret_cast<const float4 *>(fa + s1);
    a[1] = *reinterpret_cast<const float4 *>(fa + s2);
    a[2] = *reinterpret_cast<const float4 *>(fb + s1);
    a[3] = *reinterpret_cast<const float4 *>(fb + s2);
  };
  float4 cur[4];
  if (active && s_lo < s_hi) load(s_lo, 0, cur);
  for (int s0 = s_lo; s0 < s_hi; s0 += kMxSC) {
    __syncthreads();
    {
      // W[it][s0 + i][r] -> wl as complex [((k * kMxSC + i) * 2 + half) * 16 + r][jj], it = 4k + 2half + jj
      zf *wc = reinterpret_cast<zf *>(wl);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int t = threadIdx.x + 512 * u;
        const int r = t & 15, i = (t >> 4) & (kMxSC - 1), it = t >> 9;
        const int s = s0 + i;
        const bool ok = r < R && s < s_hi;
        const zf v = Wb[((size_t)it * S + min(s, S - 1)) * R + min(r, R - 1)];
        const int k = it >> 2, half = (it >> 1) & 1, jj = it & 1;
        wc[(((k * kMxSC + i) * 2 + half) * 16 + r) * 2 + jj] = ok ? v : zf{0.f, 0.f};
      }
    }
    __syncthreads();
    if (!active) continue;
    for (int sg = 0; sg < kMxSC && s0 + sg < s_hi; sg += 16) {
#pragma unroll
      for (int k = 0; k < 2; ++k) {
        float4 nx[4];
        if (k == 0)
          load(s0 + sg, 1, nx);
        else
          load(s0 + sg + 16, 0, nx);
        __builtin_amdgcn_sched_barrier(0);
        const float fa[8] = {cur[0].x, cur[0].y, cur[0].z, cur[0].w, cur[1].x, cur[1].y, cur[1].z, cur[1].w};
        const float fb[8] = {cur[2].x, cur[2].y, cur[2].z, cur[2].w, cur[3].x, cur[3].y, cur[3].z, cur[3].w};
        // the group's 16 samples are summed from zero and then added to the
        // running sums, so the running sums take one rounding per group
        mxf4 gr = {0.f, 0.f, 0.f, 0.f}, gi = gr, gh = gr, gj = gr;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int i = sg + 4 * g + j;
          const float4 we = wl[((k * kMxSC + i) * 2 + 0) * 16 + m];  // (e,a), (e,b)
          const float4 wh = wl[((k * kMxSC + i) * 2 + 1) * 16 + m];  // (h,a), (h,b)
          const float ar = fa[2 * j], ai = fa[2 * j + 1], br = fb[2 * j], bi = fb[2 * j + 1];
          const float nai = -ai, nbi = -bi;
          gr = SPK_MFMA(ar, we.x, gr);
          gi = SPK_MFMA(ar, we.y, gi);
          gh = SPK_MFMA(ar, wh.x, gh);
          gj = SPK_MFMA(ar, wh.y, gj);
          gr = SPK_MFMA(nai, we.y, gr);
          gi = SPK_MFMA(ai, we.x, gi);
          gh = SPK_MFMA(nai, wh.y, gh);
          gj = SPK_MFMA(ai, wh.x, gj);
          gr = SPK_MFMA(br, we.z, gr);
          gi = SPK_MFMA(br, we.w, gi);
          gh = SPK_MFMA(br, wh.z, gh);
          gj = SPK_MFMA(br, wh.w, gj);
          gr = SPK_MFMA(nbi, we.w, gr);
          gi = SPK_MFMA(bi, we.z, gi);
          gh = SPK_MFMA(nbi, wh.w, gh);
          gj = SPK_MFMA(bi, wh.z, gj);
        }
        er += gr;
        ei += gi;
        hr += gh;
        hi += gj;
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 4; ++u) cur[u] = nx[u];
      }
    }
  }
  // accumulator element v of lane (g, n): edge ew + 4g + v, right-hand side n
  const int n = m;
  if (!active || n >= R) return;
  zf *yf = yp + (size_t)blockIdx.y * R * 2 * N;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int e = ew + 4 * g + v;
    if (e < cnt) {
      const int mm = perm[p0 + e];
      yf[(size_t)n * 2 * N + mm] = zf{er[v], ei[v]} * sc1;
      yf[(size_t)n * 2 * N + mm + N] = zf{hr[v], hi[v]} * sc2;
    }
  }
}

// Near field for R vectors: workgroup = (leaf box, 16 rows), 4 rows per
// wave processed together.  A wave's lanes = CL column lanes x RG groups of
// right-hand sides, each lane holding RPL = RP / RG of them: a lane loads the
// 2x2 blocks of its columns (lanes of the same column share one address) and
// applies them to its RPL vectors, read from the LDS copy of the box's E2
// columns.  Rows are reduced over the column lanes at the end.
template <class T, int RP>
__global__ void __launch_bounds__(256) kb_near(int N, int R, const int *row_box, const int *row_r0,
                                               const int *box_first, const int *box_count,
                                               const int *e2_off, const int *e2_list,
                                               const int64_t *near_off, const int *perm,
                                               const T *blocks, const T *XI, const T *yf,
                                               T *const *ys) {
  constexpr int RG = RP >= 8 ? RP / 4 : 1, RPL = RP / RG, CL = 64 / RG;
  constexpr int KC = 32768 / (2 * RP * (int)sizeof(T));  // columns per LDS chunk (32 KB)
  constexpr int NQ = kSpmvRowsPerWG / 4;
  __shared__ T xsl[KC * 2 * RP];
  const int b = row_box[blockIdx.x], r0 = row_r0[blockIdx.x];
  const int rows = min(kSpmvRowsPerWG, box_count[b] - r0);
  const int c0 = e2_off[b], cols = e2_off[b + 1] - c0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int rg = lane / CL, cl = lane % CL;
  T ae[NQ][RPL], ah[NQ][RPL];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int v = 0; v < RPL; ++v) ae[q][v] = ah[q][v] = T{0, 0};
  for (int cc = 0; cc < cols; cc += KC) {
    const int nc = min(KC, cols - cc);
    __syncthreads();
    for (int t = threadIdx.x; t < nc * 2 * RP; t += 256) {
      const int c = t / (2 * RP), w = t % (2 * RP);
      xsl[t] = XI[(size_t)e2_list[c0 + cc + c] * 2 * RP + w];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int row = r0 + wv + 4 * q;
      if (row < r0 + rows) {
        const T *Z = blocks + 4 * (near_off[b] + (int64_t)row * cols + cc);
        for (int c = cl; c < nc; c += CL) {
          const T z0 = Z[4 * c], z1 = Z[4 * c + 1], z2 = Z[4 * c + 2], z3 = Z[4 * c + 3];
          const T *xe = xsl + (c * 2 + 0) * RP + rg * RPL, *xh = xsl + (c * 2 + 1) * RP + rg * RPL;
#pragma unroll
          for (int v = 0; v < RPL; ++v) {
            ae[q][v] = ae[q][v] + z0 * xe[v] + z1 * xh[v];
            ah[q][v] = ah[q][v] + z2 * xe[v] + z3 * xh[v];
          }
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int row = r0 + wv + 4 * q;
#pragma unroll
    for (int v = 0; v < RPL; ++v) {
      T se = ae[q][v], sh = ah[q][v];
#pragma unroll
      for (int off = CL / 2; off > 0; off >>= 1) {
        se = se + bshfl_down(se, off);
        sh = sh + bshfl_down(sh, off);
      }
      const int r = rg * RPL + v;
      if (cl == 0 && r < R && row < r0 + rows) {
        const int m = perm[box_first[b] + row];
        if (yf) {
          se = se + yf[(size_t)r * 2 * N + m];
          sh = sh + yf[(size_t)r * 2 * N + m + N];
        }
        ys[r][m] = se;
        ys[r][m + N] = sh;
      }
    }
  }
}

// Near field for 9..16 right-hand sides in complex64 on the matrix cores.
// Per leaf box the near operator is dense: rows = the box's edges, columns
// = its E2 edges, each entry a 2x2 complex block.  Workgroup = (box, up to
// 16 consecutive rows), its 4 waves take every fourth group of 16 columns.
// One v_mfma_f32_16x16x4f32 multiplies A[m][g] (a real part of the block of
// row m and column c + g) by B[g][n] (a real part of x_n at column c + g),
// so each complex product is four real ones into four accumulators:
//   e_re += Z0r xer - Z0i xei + Z1r xhr - Z1i xhi,  e_im += Z0r xei + Z0i xer + ...
//   h_re, h_im the same with Z2, Z3.
// A lane loads the whole 2x2 block of its (row, column), 32 bytes; the
// columns' (xe, xh) of the 16 right-hand sides are staged in LDS, 128
// columns at a time, and a wave loads its next 16 columns' blocks before it
// multiplies the current ones.  Rows and columns past the end are clamped on
// load (their x is zero in LDS, their results are not written).  The four
// waves' sums are added in wave order through LDS.
// BF: the blocks are the bf16 copy (one 16-byte load per block, widened to
// fp32 in registers).
constexpr int kMxCols = 128;
template <bool BF>
__global__ void __launch_bounds__(256) kb_near_mx(int N, int R, const int *row_box,
                                                  const int *row_r0, const int *box_first,
                                                  const int *box_count, const int *e2_off,
                                                  const int *e2_list, const int64_t *near_off,
                                                  const int *perm, const void *blocks, const zf *XI,
                                                  const zf *yf, zf *const *ys) {
  constexpr int PER = kMxCols * 16 / 256;
  __shared__ float4 xsl[kMxCols * 16];
  const int b = row_box[blockIdx.x];
  const int pf = box_first[b], nrow = box_count[b];
  const int c0 = e2_off[b], cols = e2_off[b + 1] - c0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int rw = row_r0[blockIdx.x];  // first row of the tile (rw < nrow)
  const int row = min(rw + m, nrow - 1);
  const int64_t roff = near_off[b] + (int64_t)row * cols;  // first block of the row
  const float4 *Zr = reinterpret_cast<const float4 *>(blocks) + 2 * roff;
  const uint4 *Zb = reinterpret_cast<const uint4 *>(blocks) + roff;
  mxf4 er = {0.f, 0.f, 0.f, 0.f}, ei = er, hr = er, hi = er;
  float4 za[4], zb[4];
  auto load = [&](int cg, float4 *a, float4 *bb) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = min(cg + 4 * j + g, cols - 1);
      if constexpr (BF) {
        const uint4 v = Zb[c];
        a[j] = float4{__uint_as_float(v.x << 16), __uint_as_float(v.x & 0xFFFF0000u),
                      __uint_as_float(v.y << 16), __uint_as_float(v.y & 0xFFFF0000u)};
        bb[j] = float4{__uint_as_float(v.z << 16), __uint_as_float(v.z & 0xFFFF0000u),
                       __uint_as_float(v.w << 16), __uint_as_float(v.w & 0xFFFF0000u)};
      } else {
        a[j] = Zr[2 * c];
        bb[j] = Zr[2 * c + 1];
      }
    }
  };
  if (cols > 0) load(16 * wv, za, zb);
  for (int cc = 0; cc < cols; cc += kMxCols) {
    const int nc = min(kMxCols, cols - cc);
    __syncthreads();
    {
      // unconditional (clamped) loads so that all of them are in flight
      // together; XI is zero for right-hand sides R..15
      const int n = threadIdx.x & 15;
      int e[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k)
        e[k] = e2_list[c0 + cc + min((int)(threadIdx.x + 256 * k) >> 4, nc - 1)];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const zf xe = XI[((size_t)e[k] * 2 + 0) * 16 + n], xh = XI[((size_t)e[k] * 2 + 1) * 16 + n];
        const bool ok = ((int)(threadIdx.x + 256 * k) >> 4) < nc;
        xsl[threadIdx.x + 256 * k] = ok ? float4{xe.x, xe.y, xh.x, xh.y} : float4{0.f, 0.f, 0.f, 0.f};
      }
    }
    __syncthreads();
    for (int off = 16 * wv; off < nc; off += 64) {
      float4 na[4], nb[4];
      load(cc + off + 64, na, nb);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float4 x = xsl[(off + 4 * j + g) * 16 + m];
        const float4 z = za[j], w = zb[j];
        const float nxi = -x.y, nhi = -x.w;
        er = SPK_MFMA(z.x, x.x, er);
        ei = SPK_MFMA(z.x, x.y, ei);
        hr = SPK_MFMA(w.x, x.x, hr);
        hi = SPK_MFMA(w.x, x.y, hi);
        er = SPK_MFMA(z.y, nxi, er);
        ei = SPK_MFMA(z.y, x.x, ei);
        hr = SPK_MFMA(w.y, nxi, hr);
        hi = SPK_MFMA(w.y, x.x, hi);
        er = SPK_MFMA(z.z, x.z, er);
        ei = SPK_MFMA(z.z, x.w, ei);
        hr = SPK_MFMA(w.z, x.z, hr);
        hi = SPK_MFMA(w.z, x.w, hi);
        er = SPK_MFMA(z.w, nhi, er);
        ei = SPK_MFMA(z.w, x.z, ei);
        hr = SPK_MFMA(w.w, nhi, hr);
        hi = SPK_MFMA(w.w, x.z, hi);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        za[j] = na[j];
        zb[j] = nb[j];
      }
    }
  }
  // waves 1..3 hand their sums to wave 0 through LDS
  __syncthreads();
  mxf4 *red = reinterpret_cast<mxf4 *>(xsl);
  if (wv > 0) {
    mxf4 *o = red + ((size_t)(wv - 1) * 64 + lane) * 4;
    o[0] = er;
    o[1] = ei;
    o[2] = hr;
    o[3] = hi;
  }
  __syncthreads();
  if (wv > 0) return;
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const mxf4 *o = red + ((size_t)p * 64 + lane) * 4;
    er += o[0];
    ei += o[1];
    hr += o[2];
    hi += o[3];
  }
  // accumulator element i of lane (g, n) is row 4g + i of the tile, right-hand side n
  const int n = m;
  if (n >= R) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = rw + 4 * g + i;
    if (r < nrow) {
      const int mm = perm[pf + r];
      zf se = {er[i], ei[i]}, sh = {hr[i], hi[i]};
      if (yf) {
        se = se + yf[(size_t)n * 2 * N + mm];
        sh = sh + yf[(size_t)n * 2 * N + mm + N];
      }
      ys[n][mm] = se;
      ys[n][mm + N] = sh;
    }
  }
}

// ---------------------------------------------------------------------------
// The same three products for complex128 batches of 9..16 on the FP64
// matrix cores (v_mfma_f64_16x16x4f64: A[m][g] in lane g*16 + m and B[g][n]
// in lane g*16 + n as in the fp32 form, but D[4i + g][n] in element i of lane
// g*16 + n, as probed on the device).  Each complex value is one 16-byte
// load; sums run as one running sum (complex128 needs no grouping).
// ---------------------------------------------------------------------------
typedef double mxd4 __attribute__((ext_vector_type(4)));
#define SPK_MFMA64(a, b, c) __builtin_amdgcn_mfma_f64_16x16x4f64((a), (b), (c), 0, 0, 0)

// near field: workgroup = (box, 16 rows), 4 waves take every fourth group of
// 16 columns, x staged in LDS 64 columns at a time as (xe, xh) per (column, r)
constexpr int kMx64Cols = 64;
__global__ void __launch_bounds__(256) kb_near_mx64(int N, int R, const int *row_box,
                                                    const int *row_r0, const int *box_first,
                                                    const int *box_count, const int *e2_off,
                                                    const int *e2_list, const int64_t *near_off,
                                                    const int *perm, const zd *blocks, const zd *XI,
                                                    const zd *yf, zd *const *ys) {
  constexpr int PER = kMx64Cols * 16 / 256;
  __shared__ double4 xsl[kMx64Cols * 16];
  const int b = row_box[blockIdx.x];
  const int pf = box_first[b], nrow = box_count[b];
  const int c0 = e2_off[b], cols = e2_off[b + 1] - c0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int rw = row_r0[blockIdx.x];
  const int row = min(rw + m, nrow - 1);
  const double2 *Zr = reinterpret_cast<const double2 *>(blocks) + 4 * (near_off[b] + (int64_t)row * cols);
  mxd4 er = {0.0, 0.0, 0.0, 0.0}, ei = er, hr = er, hi = er;
  double2 cur[4][4];
  auto load = [&](int cg, double2 (*z)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = min(cg + 4 * j + g, cols - 1);
#pragma unroll
      for (int e = 0; e < 4; ++e) z[j][e] = Zr[4 * c + e];
    }
  };
  if (cols > 0) load(16 * wv, cur);
  for (int cc = 0; cc < cols; cc += kMx64Cols) {
    const int nc = min(kMx64Cols, cols - cc);
    __syncthreads();
    {
      const int n = threadIdx.x & 15;
      int e[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k)
        e[k] = e2_list[c0 + cc + min((int)(threadIdx.x + 256 * k) >> 4, nc - 1)];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const zd xe = XI[((size_t)e[k] * 2 + 0) * 16 + n], xh = XI[((size_t)e[k] * 2 + 1) * 16 + n];
        const bool ok = ((int)(threadIdx.x + 256 * k) >> 4) < nc;
        xsl[threadIdx.x + 256 * k] = ok ? double4{xe.x, xe.y, xh.x, xh.y} : double4{0.0, 0.0, 0.0, 0.0};
      }
    }
    __syncthreads();
    for (int off = 16 * wv; off < nc; off += 64) {
      double2 nx[4][4];
      load(cc + off + 64, nx);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double4 x = xsl[(off + 4 * j + g) * 16 + m];
        const double2 z0 = cur[j][0], z1 = cur[j][1], z2 = cur[j][2], z3 = cur[j][3];
        const double nxi = -x.y, nhi = -x.w;
        er = SPK_MFMA64(z0.x, x.x, er);
        ei = SPK_MFMA64(z0.x, x.y, ei);
        hr = SPK_MFMA64(z2.x, x.x, hr);
        hi = SPK_MFMA64(z2.x, x.y, hi);
        er = SPK_MFMA64(z0.y, nxi, er);
        ei = SPK_MFMA64(z0.y, x.x, ei);
        hr = SPK_MFMA64(z2.y, nxi, hr);
        hi = SPK_MFMA64(z2.y, x.x, hi);
        er = SPK_MFMA64(z1.x, x.z, er);
        ei = SPK_MFMA64(z1.x, x.w, ei);
        hr = SPK_MFMA64(z3.x, x.z, hr);
        hi = SPK_MFMA64(z3.x, x.w, hi);
        er = SPK_MFMA64(z1.y, nhi, er);
        ei = SPK_MFMA64(z1.y, x.z, ei);
        hr = SPK_MFMA64(z3.y, nhi, hr);
        hi = SPK_MFMA64(z3.y, x.z, hi);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) cur[j][e] = nx[j][e];
    }
  }
  __syncthreads();
  mxd4 *red = reinterpret_cast<mxd4 *>(xsl);
  if (wv > 0) {
    mxd4 *o = red + ((size_t)(wv - 1) * 64 + lane) * 4;
    o[0] = er;
    o[1] = ei;
    o[2] = hr;
    o[3] = hi;
  }
  __syncthreads();
  if (wv > 0) return;
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const mxd4 *o = red + ((size_t)p * 64 + lane) * 4;
    er += o[0];
    ei += o[1];
    hr += o[2];
    hi += o[3];
  }
  const int n = m;
  if (n >= R) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = rw + 4 * i + g;
    if (r < nrow) {
      const int mm = perm[pf + r];
      zd se = {er[i], ei[i]}, sh = {hr[i], hi[i]};
      if (yf) {
        se = se + yf[(size_t)n * 2 * N + mm];
        sh = sh + yf[(size_t)n * 2 * N + mm + N];
      }
      ys[n][mm] = se;
      ys[n][mm + N] = sh;
    }
  }
}

// aggregation: workgroup = (128 samples, box, q), wave = 32 samples as 2
// tiles of 16 (tile t: samples sw + 16 t + i); k-step = 4 edges
__global__ void __launch_bounds__(256) kb_agg_mx64(int S, int R, const int *box_first,
                                                   const int *box_count, const zd *rt,
                                                   const zd *XP, zd *C) {
  __shared__ double4 xl[32 * 16];
  const int b = blockIdx.y, q = blockIdx.z, k = q >> 1, comp = q & 1;
  const int p0 = box_first[b], cnt = box_count[b];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int sw = blockIdx.x * 128 + 32 * wv;
  const bool active = sw < S;
  const int sa = min(sw + m, S - 1), sb = min(sw + 16 + m, S - 1);
  mxd4 acc[2][4];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[t][u] = mxd4{0.0, 0.0, 0.0, 0.0};
  double2 fa[4], fb[4];
  auto load = [&](int e0, double2 *a, double2 *bb) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = min(e0 + 4 * j + g, cnt - 1);
      const double2 *f = reinterpret_cast<const double2 *>(rt + ((size_t)(p0 + e) * 4 + q) * S);
      a[j] = f[sa];
      bb[j] = f[sb];
    }
  };
  if (active && cnt > 0) load(0, fa, fb);
  for (int e0 = 0; e0 < cnt; e0 += 32) {
    const int ne = min(32, cnt - e0);
    __syncthreads();
    {
      const int n = threadIdx.x & 15;
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int e = (int)(threadIdx.x + 256 * u) >> 4;
        const size_t pe = (size_t)(p0 + e0 + min(e, ne - 1)) * 2;
        const zd x0 = XP[pe * 16 + n], x1 = XP[(pe + 1) * 16 + n];
        xl[threadIdx.x + 256 * u] = e < ne ? double4{x0.x, x0.y, x1.x, x1.y} : double4{0.0, 0.0, 0.0, 0.0};
      }
    }
    __syncthreads();
    if (!active) continue;
    for (int i = 0; i < ne; i += 16) {
      double2 na[4], nb[4];
      load(e0 + i + 16, na, nb);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double4 x = xl[(i + 4 * j + g) * 16 + m];
        const double n0 = -x.x, n1 = -x.z;
        const double ar[2] = {fa[j].x, fb[j].x}, ai[2] = {fa[j].y, fb[j].y};
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          acc[t][0] = SPK_MFMA64(ar[t], x.x, acc[t][0]);
          acc[t][1] = SPK_MFMA64(ar[t], x.y, acc[t][1]);
          acc[t][2] = SPK_MFMA64(ar[t], x.z, acc[t][2]);
          acc[t][3] = SPK_MFMA64(ar[t], x.w, acc[t][3]);
        }
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          acc[t][0] = SPK_MFMA64(ai[t], x.y, acc[t][0]);
          acc[t][1] = SPK_MFMA64(ai[t], n0, acc[t][1]);
          acc[t][2] = SPK_MFMA64(ai[t], x.w, acc[t][2]);
          acc[t][3] = SPK_MFMA64(ai[t], n1, acc[t][3]);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        fa[j] = na[j];
        fb[j] = nb[j];
      }
    }
  }
  const int n = m;
  if (!active || n >= R) return;
  zd *Cb = C + (size_t)b * 8 * S * R;
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int s = sw + 16 * t + 4 * v + g;
      if (s < S) {
        Cb[((size_t)(comp * 4 + 0 + k) * S + s) * R + n] = zd{acc[t][0][v], acc[t][1][v]};
        Cb[((size_t)(comp * 4 + 2 + k) * S + s) * R + n] = zd{acc[t][2][v], acc[t][3][v]};
      }
    }
}

// disaggregation: workgroup = (up to 128 edges of a box, sample part),
// wave = 16 edges; k-step = 4 samples (lane group g takes sample s + 4g + j
// in step j); W staged in LDS 16 samples at a time as [k][sample][e/h][r]
constexpr int kMx64SC = 16;
__global__ void __launch_bounds__(512) kb_disagg_mx64(int N, int S, int R, const int *tile_box,
                                                      const int *tile_p0, const int *box_first,
                                                      const int *box_count, const int *perm,
                                                      const zd *rt, const zd *W, zd sc1, zd sc2,
                                                      int sblk, zd *yp) {
  __shared__ double4 wl[2 * kMx64SC * 2 * 16];
  const int b = tile_box[blockIdx.x], p0 = tile_p0[blockIdx.x];
  const int cnt = min(128, box_first[b] + box_count[b] - p0);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int ew = 16 * wv;
  const bool active = ew < cnt;
  const double2 *F = reinterpret_cast<const double2 *>(rt + (size_t)(p0 + min(ew + m, cnt - 1)) * 4 * S);
  const size_t SR = (size_t)S * R;
  const zd *Wb = W + (size_t)b * 8 * SR;
  const int s_lo = blockIdx.y * sblk, s_hi = min(S, s_lo + sblk);
  mxd4 er = {0.0, 0.0, 0.0, 0.0}, ei = er, hr = er, hi = er;
  // pattern rows 2k and 2k+1 at samples s + 4g .. s + 4g + 3 (clamped)
  auto load = [&](int s, int k, double2 *a) {
    const double2 *fa = F + (size_t)(2 * k) * S, *fb = fa + S;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int sj = min(s + 4 * g + j, S - 1);
      a[j] = fa[sj];
      a[4 + j] = fb[sj];
    }
  };
  double2 cur[8];
  if (active && s_lo < s_hi) load(s_lo, 0, cur);
  for (int s0 = s_lo; s0 < s_hi; s0 += kMx64SC) {
    __syncthreads();
    {
      // W[it][s0 + i][r] -> wl as complex [((k * SC + i) * 2 + half) * 16 + r][jj], it = 4k + 2half + jj
      zd *wc = reinterpret_cast<zd *>(wl);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = threadIdx.x + 512 * u;
        const int r = t & 15, i = (t >> 4) & (kMx64SC - 1), it = t >> 8;
        const int s = s0 + i;
        const bool ok = r < R && s < s_hi;
        const zd v = Wb[((size_t)it * S + min(s, S - 1)) * R + min(r, R - 1)];
        const int k = it >> 2, half = (it >> 1) & 1, jj = it & 1;
        wc[(((k * kMx64SC + i) * 2 + half) * 16 + r) * 2 + jj] = ok ? v : zd{0.0, 0.0};
      }
    }
    __syncthreads();
    if (!active) continue;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      double2 nx[8];
      if (k == 0)
        load(s0, 1, nx);
      else
        load(s0 + kMx64SC, 0, nx);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int i = 4 * g + j;
        const double4 we = wl[((k * kMx64SC + i) * 2 + 0) * 16 + m];  // (e,a), (e,b)
        const double4 wh = wl[((k * kMx64SC + i) * 2 + 1) * 16 + m];  // (h,a), (h,b)
        const double ar = cur[j].x, ai = cur[j].y, br = cur[4 + j].x, bi = cur[4 + j].y;
        const double nai = -ai, nbi = -bi;
        er = SPK_MFMA64(ar, we.x, er);
        ei = SPK_MFMA64(ar, we.y, ei);
        hr = SPK_MFMA64(ar, wh.x, hr);
        hi = SPK_MFMA64(ar, wh.y, hi);
        er = SPK_MFMA64(nai, we.y, er);
        ei = SPK_MFMA64(ai, we.x, ei);
        hr = SPK_MFMA64(nai, wh.y, hr);
        hi = SPK_MFMA64(ai, wh.x, hi);
        er = SPK_MFMA64(br, we.z, er);
        ei = SPK_MFMA64(br, we.w, ei);
        hr = SPK_MFMA64(br, wh.z, hr);
        hi = SPK_MFMA64(br, wh.w, hi);
        er = SPK_MFMA64(nbi, we.w, er);
        ei = SPK_MFMA64(bi, we.z, ei);
        hr = SPK_MFMA64(nbi, wh.w, hr);
        hi = SPK_MFMA64(bi, wh.z, hi);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < 8; ++u) cur[u] = nx[u];
    }
  }
  const int n = m;
  if (!active || n >= R) return;
  zd *yf = yp + (size_t)blockIdx.y * R * 2 * N;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int e = ew + 4 * v + g;
    if (e < cnt) {
      const int mm = perm[p0 + e];
      yf[(size_t)n * 2 * N + mm] = zd{er[v], ei[v]} * sc1;
      yf[(size_t)n * 2 * N + mm + N] = zd{hr[v], hi[v]} * sc2;
    }
  }
}

// ---------------------------------------------------------------------------
// The near-field preconditioner's product for 9..16 right-hand sides on the
// bf16 matrix cores (v_mfma_f32_16x16x16bf16_1k: lane g*16 + m holds
// A[m][4g .. 4g+3] and B[4g .. 4g+3][n = m], D[4g + i][n] in element i, as
// probed on the device).  One product covers 4 columns: k = 4 * column +
// component, A = the bf16 block of (row, column) as stored: (Z0r, Z0i, Z1r,
// Z1i) for the e rows and (Z2r, Z2i, Z3r, Z3i) for the h rows; B = x in bf16
// arranged per (column, right-hand side) as (xer, -xei, xhr, -xhi) for the
// real parts and (xei, xer, xhi, xhr) for the imaginary parts.  Products
// accumulate in fp32.  Workgroup = (box, 16 rows), the 4 waves take every
// fourth group of 16 columns, x staged in LDS 128 columns at a time.
// ---------------------------------------------------------------------------
typedef short mxs4 __attribute__((ext_vector_type(4)));
__device__ inline uint32_t bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite values)
  return u >> 16;
}
__device__ inline uint32_t bf16_pair(float lo, float hi) { return bf16_bits(lo) | (bf16_bits(hi) << 16); }
__device__ inline mxs4 as_s4(uint32_t a, uint32_t b) {
  mxs4 v;
  v[0] = (short)(a & 0xFFFFu);
  v[1] = (short)(a >> 16);
  v[2] = (short)(b & 0xFFFFu);
  v[3] = (short)(b >> 16);
  return v;
}
__global__ void __launch_bounds__(256) kb_near_mxb(int N, int R, const int *row_box,
                                                   const int *row_r0, const int *box_first,
                                                   const int *box_count, const int *e2_off,
                                                   const int *e2_list, const int64_t *near_off,
                                                   const int *perm, const uint4 *blocks, const zf *XI,
                                                   zf *const *ys) {
  constexpr int PER = kMxCols * 16 / 256;
  __shared__ uint4 xsl[kMxCols * 16];
  const int b = row_box[blockIdx.x];
  const int pf = box_first[b], nrow = box_count[b];
  const int c0 = e2_off[b], cols = e2_off[b + 1] - c0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int rw = row_r0[blockIdx.x];
  const int row = min(rw + m, nrow - 1);
  const uint4 *Zr = blocks + near_off[b] + (int64_t)row * cols;
  mxf4 er = {0.f, 0.f, 0.f, 0.f}, ei = er, hr = er, hi = er;
  uint4 cur[4];
  auto load = [&](int cg, uint4 *z) {
#pragma unroll
    for (int j = 0; j < 4; ++j) z[j] = Zr[min(cg + 4 * j + g, cols - 1)];
  };
  if (cols > 0) load(16 * wv, cur);
  for (int cc = 0; cc < cols; cc += kMxCols) {
    const int nc = min(kMxCols, cols - cc);
    __syncthreads();
    {
      const int n = threadIdx.x & 15;
      int e[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k)
        e[k] = e2_list[c0 + cc + min((int)(threadIdx.x + 256 * k) >> 4, nc - 1)];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const zf xe = XI[((size_t)e[k] * 2 + 0) * 16 + n], xh = XI[((size_t)e[k] * 2 + 1) * 16 + n];
        const bool ok = ((int)(threadIdx.x + 256 * k) >> 4) < nc;
        const uint4 v = {bf16_pair(xe.x, -xe.y), bf16_pair(xh.x, -xh.y), bf16_pair(xe.y, xe.x),
                         bf16_pair(xh.y, xh.x)};
        xsl[threadIdx.x + 256 * k] = ok ? v : uint4{0u, 0u, 0u, 0u};
      }
    }
    __syncthreads();
    for (int off = 16 * wv; off < nc; off += 64) {
      uint4 nx[4];
      load(cc + off + 64, nx);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint4 x = xsl[(off + 4 * j + g) * 16 + m];
        const mxs4 be = as_s4(x.x, x.y), bi = as_s4(x.z, x.w);
        const mxs4 ae = as_s4(cur[j].x, cur[j].y), ah = as_s4(cur[j].z, cur[j].w);
        er = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ae, be, er, 0, 0, 0);
        ei = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ae, bi, ei, 0, 0, 0);
        hr = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, be, hr, 0, 0, 0);
        hi = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bi, hi, 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < 4; ++j) cur[j] = nx[j];
    }
  }
  // waves 1..3 hand their sums to wave 0 through LDS
  __syncthreads();
  mxf4 *red = reinterpret_cast<mxf4 *>(xsl);
  if (wv > 0) {
    mxf4 *o = red + ((size_t)(wv - 1) * 64 + lane) * 4;
    o[0] = er;
    o[1] = ei;
    o[2] = hr;
    o[3] = hi;
  }
  __syncthreads();
  if (wv > 0) return;
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    const mxf4 *o = red + ((size_t)p * 64 + lane) * 4;
    er += o[0];
    ei += o[1];
    hr += o[2];
    hi += o[3];
  }
  const int n = m;
  if (n >= R) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = rw + 4 * g + i;
    if (r < nrow) {
      const int mm = perm[pf + r];
      ys[n][mm] = zf{er[i], ei[i]};
      ys[n][mm + N] = zf{hr[i], hi[i]};
    }
  }
}

// The same product on the fp8 copy of the blocks (E4M3FN, 8 bytes per 2x2
// block, per-row component scales).  The scales differ per component, so each
// component gets its own accumulators: one product covers 8 columns of one
// component, k = 2 * column + (re, im); lane g * 16 + m holds A = (Zc_r, Zc_i)
// of row m at columns 2g and 2g + 1 of the 8 (decoded by v_cvt_pk_f32_fp8;
// an fp8 value has 3 mantissa bits, so its fp32 form is exactly bf16 and the
// high halves are the operand) and B = the staged x words of those columns.
// The row's scales multiply the component sums before the waves' sums are
// added.  A workgroup covers RT row tiles of 16 rows (one x staging for all).
template <bool HI>
__device__ inline mxs4 fp8_comp_pair(uint32_t w0, uint32_t w1) {
  const auto a = __builtin_amdgcn_cvt_pk_f32_fp8((int)w0, HI);
  const auto b = __builtin_amdgcn_cvt_pk_f32_fp8((int)w1, HI);
  return as_s4((__float_as_uint(a[0]) >> 16) | (__float_as_uint(a[1]) & 0xFFFF0000u),
               (__float_as_uint(b[0]) >> 16) | (__float_as_uint(b[1]) & 0xFFFF0000u));
}
template <int RT>
__global__ void __launch_bounds__(256) kb_near_mxf8(int N, int R, const int *row_box,
                                                    const int *row_r0, const int *box_first,
                                                    const int *box_count, const int *e2_off,
                                                    const int *e2_list, const int64_t *near_off,
                                                    const int *perm, const uint2 *blocks,
                                                    const float4 *scale, const uint32_t *XB,
                                                    zf *const *ys) {
  constexpr int PER = kMxCols * 16 / 256;
  __shared__ uint4 xsl[kMxCols * 16];
  const int b = row_box[blockIdx.x];
  const int pf = box_first[b], nrow = box_count[b];
  const int c0 = e2_off[b], cols = e2_off[b + 1] - c0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int m = lane & 15, g = lane >> 4;
  const int rw = row_r0[blockIdx.x];
  const uint2 *Zr[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) Zr[t] = blocks + near_off[b] + (int64_t)min(rw + 16 * t + m, nrow - 1) * cols;
  // acc[t][c][0] real, acc[t][c][1] imaginary part of component c's sum, row tile t
  mxf4 acc[RT][4][2];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[t][c][0] = acc[t][c][1] = mxf4{0.f, 0.f, 0.f, 0.f};
  // a lane's blocks of one 16-column step: columns cg + 8 jj + 2g + u
  uint2 cur[RT][4];
  auto load = [&](int cg, uint2 (*z)[4]) {
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) z[t][j] = Zr[t][min(cg + 8 * (j >> 1) + 2 * g + (j & 1), cols - 1)];
  };
  if (cols > 0) load(16 * wv, cur);
  for (int cc = 0; cc < cols; cc += kMxCols) {
    const int nc = min(kMxCols, cols - cc);
    __syncthreads();
    {
      const int n = threadIdx.x & 15;
      int e[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k)
        e[k] = e2_list[c0 + cc + min((int)(threadIdx.x + 256 * k) >> 4, nc - 1)];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        // x in bf16 (re low, im high), rounded once by kb_x_to_bf: negating a
        // half is its sign bit, (im, re) is the halves swapped
        const uint32_t xe = XB[((size_t)e[k] * 2 + 0) * 16 + n], xh = XB[((size_t)e[k] * 2 + 1) * 16 + n];
        const bool ok = ((int)(threadIdx.x + 256 * k) >> 4) < nc;
        const uint4 v = {xe ^ 0x80000000u, xh ^ 0x80000000u, (xe >> 16) | (xe << 16),
                         (xh >> 16) | (xh << 16)};
        xsl[threadIdx.x + 256 * k] = ok ? v : uint4{0u, 0u, 0u, 0u};
      }
    }
    __syncthreads();
    for (int off = 16 * wv; off < nc; off += 64) {
      uint2 nx[RT][4];
      load(cc + off + 64, nx);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        const int col = off + 8 * jj + 2 * g;
        const uint4 x0 = xsl[col * 16 + m], x1 = xsl[(col + 1) * 16 + m];
        const mxs4 ber = as_s4(x0.x, x1.x), bhr = as_s4(x0.y, x1.y);
        const mxs4 bei = as_s4(x0.z, x1.z), bhi = as_s4(x0.w, x1.w);
#pragma unroll
        for (int t = 0; t < RT; ++t) {
          const uint2 z0 = cur[t][2 * jj], z1 = cur[t][2 * jj + 1];
          const mxs4 a0 = fp8_comp_pair<false>(z0.x, z1.x), a1 = fp8_comp_pair<true>(z0.x, z1.x);
          const mxs4 a2 = fp8_comp_pair<false>(z0.y, z1.y), a3 = fp8_comp_pair<true>(z0.y, z1.y);
          mxf4(&A)[4][2] = acc[t];
          A[0][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, ber, A[0][0], 0, 0, 0);
          A[0][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, bei, A[0][1], 0, 0, 0);
          A[1][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, bhr, A[1][0], 0, 0, 0);
          A[1][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, bhi, A[1][1], 0, 0, 0);
          A[2][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, ber, A[2][0], 0, 0, 0);
          A[2][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, bei, A[2][1], 0, 0, 0);
          A[3][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, bhr, A[3][0], 0, 0, 0);
          A[3][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, bhi, A[3][1], 0, 0, 0);
        }
      }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int t = 0; t < RT; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) cur[t][j] = nx[t][j];
    }
  }
  // element i of this lane in tile t is output row rw + 16t + 4g + i: scale,
  // then sum the waves (tile by tile through LDS)
  mxf4 *red = reinterpret_cast<mxf4 *>(xsl);
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    mxf4 er, ei, hr, hi;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = min(rw + 16 * t + 4 * g + i, nrow - 1);
      const float4 sc = scale[pf + r];
      er[i] = sc.x * acc[t][0][0][i] + sc.y * acc[t][1][0][i];
      ei[i] = sc.x * acc[t][0][1][i] + sc.y * acc[t][1][1][i];
      hr[i] = sc.z * acc[t][2][0][i] + sc.w * acc[t][3][0][i];
      hi[i] = sc.z * acc[t][2][1][i] + sc.w * acc[t][3][1][i];
    }
    __syncthreads();
    if (wv > 0) {
      mxf4 *o = red + ((size_t)(wv - 1) * 64 + lane) * 4;
      o[0] = er;
      o[1] = ei;
      o[2] = hr;
      o[3] = hi;
    }
    __syncthreads();
    if (wv == 0) {
#pragma unroll
      for (int p = 0; p < 3; ++p) {
        const mxf4 *o = red + ((size_t)p * 64 + lane) * 4;
        er += o[0];
        ei += o[1];
        hr += o[2];
        hi += o[3];
      }
      const int n = m;
      if (n < R) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = rw + 16 * t + 4 * g + i;
          if (r < nrow) {
            const int mm = perm[pf + r];
            ys[n][mm] = zf{er[i], ei[i]};
            ys[n][mm + N] = zf{hr[i], hi[i]};
          }
        }
      }
    }
  }
}

// The same product with TPW row tiles of a box per workgroup sharing one x
// staging: wave w works on tile w / CS over the columns c = 16 (w % CS) +
// 16 CS k of each chunk (CS = 4 / TPW column splits), and the CS waves of a
// tile add their sums through LDS at the end.  TPW = 1 is one 16-row tile
// per workgroup; TPW = 4 stages x once for 64 rows but leaves a quarter of
// the workgroups.  Waves without rows (a box's last group) take part in the
// staging only.
template <int TPW>
__global__ void __launch_bounds__(256) kb_near_mxf8g(int N, int R, const int *grp_box,
                                                     const int *grp_r0, const int *box_first,
                                                     const int *box_count, const int *e2_off,
                                                     const int *e2_list, const int64_t *near_off,
                                                     const int *perm, const uint2 *blocks,
                                                     const float4 *scale, const uint32_t *XB,
                                                     zf *const *ys) {
  constexpr int CS = 4 / TPW;
  constexpr int PER = kMxCols * 16 / 256;
  __shared__ uint4 xsl[kMxCols * 16];
  const int b = grp_box[blockIdx.x];
  const int pf = box_first[b], nrow = box_count[b];
  const int c0 = e2_off[b], cols = e2_off[b + 1] - c0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tile = wv / CS, split = wv % CS;
  const int m = lane & 15, g = lane >> 4;
  const int rw = grp_r0[blockIdx.x] + 16 * tile;
  const bool active = rw < nrow;
  const uint2 *Zr = blocks + near_off[b] + (int64_t)min(rw + m, nrow - 1) * cols;
  // acc[c][0] real, acc[c][1] imaginary part of component c's sum
  mxf4 acc[4][2];
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c][0] = acc[c][1] = mxf4{0.f, 0.f, 0.f, 0.f};
  // a lane's blocks of one 16-column step: columns cg + 8 jj + 2g + u
  uint2 cur[4];
  auto load = [&](int cg, uint2 *z) {
#pragma unroll
    for (int j = 0; j < 4; ++j) z[j] = Zr[min(cg + 8 * (j >> 1) + 2 * g + (j & 1), cols - 1)];
  };
  if (active && cols > 0) load(16 * split, cur);
  for (int cc = 0; cc < cols; cc += kMxCols) {
    const int nc = min(kMxCols, cols - cc);
    __syncthreads();
    {
      const int n = threadIdx.x & 15;
      int e[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k)
        e[k] = e2_list[c0 + cc + min((int)(threadIdx.x + 256 * k) >> 4, nc - 1)];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const uint32_t xe = XB[((size_t)e[k] * 2 + 0) * 16 + n], xh = XB[((size_t)e[k] * 2 + 1) * 16 + n];
        const bool ok = ((int)(threadIdx.x + 256 * k) >> 4) < nc;
        const uint4 v = {xe ^ 0x80000000u, xh ^ 0x80000000u, (xe >> 16) | (xe << 16),
                         (xh >> 16) | (xh << 16)};
        xsl[threadIdx.x + 256 * k] = ok ? v : uint4{0u, 0u, 0u, 0u};
      }
    }
    __syncthreads();
    if (active) {
      for (int off = 16 * split; off < nc; off += 16 * CS) {
        uint2 nx[4];
        load(cc + off + 16 * CS, nx);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int jj = 0; jj < 2; ++jj) {
          const int col = off + 8 * jj + 2 * g;
          const uint4 x0 = xsl[col * 16 + m], x1 = xsl[(col + 1) * 16 + m];
          const mxs4 ber = as_s4(x0.x, x1.x), bhr = as_s4(x0.y, x1.y);
          const mxs4 bei = as_s4(x0.z, x1.z), bhi = as_s4(x0.w, x1.w);
          const uint2 z0 = cur[2 * jj], z1 = cur[2 * jj + 1];
          const mxs4 a0 = fp8_comp_pair<false>(z0.x, z1.x), a1 = fp8_comp_pair<true>(z0.x, z1.x);
          const mxs4 a2 = fp8_comp_pair<false>(z0.y, z1.y), a3 = fp8_comp_pair<true>(z0.y, z1.y);
          acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, ber, acc[0][0], 0, 0, 0);
          acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, bei, acc[0][1], 0, 0, 0);
          acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, bhr, acc[1][0], 0, 0, 0);
          acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, bhi, acc[1][1], 0, 0, 0);
          acc[2][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, ber, acc[2][0], 0, 0, 0);
          acc[2][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, bei, acc[2][1], 0, 0, 0);
          acc[3][0] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, bhr, acc[3][0], 0, 0, 0);
          acc[3][1] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a3, bhi, acc[3][1], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < 4; ++j) cur[j] = nx[j];
      }
    }
  }
  // element i of this lane is output row rw + 4g + i of vector m: the row's
  // scales, then the splits' sums in a fixed order (split 0 adds 1..CS-1)
  mxf4 er, ei, hr, hi;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = min(rw + 4 * g + i, nrow - 1);
    const float4 sc = scale[pf + r];
    er[i] = sc.x * acc[0][0][i] + sc.y * acc[1][0][i];
    ei[i] = sc.x * acc[0][1][i] + sc.y * acc[1][1][i];
    hr[i] = sc.z * acc[2][0][i] + sc.w * acc[3][0][i];
    hi[i] = sc.z * acc[2][1][i] + sc.w * acc[3][1][i];
  }
  if constexpr (CS > 1) {
    mxf4 *red = reinterpret_cast<mxf4 *>(xsl);
    __syncthreads();
    if (split > 0) {
      mxf4 *o = red + (((size_t)tile * (CS - 1) + split - 1) * 64 + lane) * 4;
      o[0] = er;
      o[1] = ei;
      o[2] = hr;
      o[3] = hi;
    }
    __syncthreads();
    if (split > 0) return;
#pragma unroll
    for (int p = 0; p < CS - 1; ++p) {
      const mxf4 *o = red + (((size_t)tile * (CS - 1) + p) * 64 + lane) * 4;
      er += o[0];
      ei += o[1];
      hr += o[2];
      hi += o[3];
    }
  }
  if (!active) return;
  const int n = m;
  if (n >= R) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = rw + 4 * g + i;
    if (r < nrow) {
      const int mm = perm[pf + r];
      ys[n][mm] = zf{er[i], ei[i]};
      ys[n][mm + N] = zf{hr[i], hi[i]};
    }
  }
}

// the preconditioner's product for 9..16 right-hand sides on the fp8 copy
// (blocks, scale as k_near_spmv_f8 reads them); workgroups = row tiles of
// 16 rows.  Two tiles per workgroup (x staged once for 32 rows, 156 VGPRs)
// measured the same, 105 us at 16 vectors on the 10 x 10 lambda patch.
// XB[i] = XI[i] in bf16 (re in the low half, im in the high half), rounded
// to nearest even as bf16_pair rounds
__global__ void __launch_bounds__(256) kb_x_to_bf(int64_t len, const zf *XI, uint32_t *XB) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) XB[i] = bf16_pair(XI[i].x, XI[i].y);
}
void launch_batch_near_f8(int nwg, int N, int R, const int *row_box, const int *row_r0,
                          int ngrp, const int *grp_box, const int *grp_r0,
                          const int *box_first, const int *box_count, const int *e2_off,
                          const int *e2_list, const int64_t *near_off, const int *perm,
                          const uint32_t *blocks, const float *scale, const zf *XI,
                          uint32_t *XB, zf *const *ys, int tpw, hipStream_t st) {
  if (nwg <= 0) return;
  const int64_t len = (int64_t)N * 2 * 16;
  hipLaunchKernelGGL(kb_x_to_bf, dim3(bnblk(len, 256)), dim3(256), 0, st, len, XI, XB);
  // tpw row tiles per workgroup (grp_box / grp_r0 are the map of groups of
  // tpw 16-row tiles); 1 runs the original kernel on the 16-row tile map
  if (tpw == 1 || ngrp <= 0)
    hipLaunchKernelGGL(kb_near_mxf8<1>, dim3(nwg), dim3(256), 0, st, N, R, row_box, row_r0,
                       box_first, box_count, e2_off, e2_list, near_off, perm, (const uint2 *)blocks,
                       (const float4 *)scale, (const uint32_t *)XB, ys);
  else if (tpw == 2)
    hipLaunchKernelGGL(kb_near_mxf8g<2>, dim3(ngrp), dim3(256), 0, st, N, R, grp_box, grp_r0,
                       box_first, box_count, e2_off, e2_list, near_off, perm, (const uint2 *)blocks,
                       (const float4 *)scale, (const uint32_t *)XB, ys);
  else
    hipLaunchKernelGGL(kb_near_mxf8g<4>, dim3(ngrp), dim3(256), 0, st, N, R, grp_box, grp_r0,
                       box_first, box_count, e2_off, e2_list, near_off, perm, (const uint2 *)blocks,
                       (const float4 *)scale, (const uint32_t *)XB, ys);
}

// Row tiles per workgroup of the batched fp8 near product: 1 by default,
// The process default of the row tiles per workgroup (a handle's tuned
// value, spk_autotune, takes precedence); SPK_NEAR_MXF8_TPW = 1, 2 or 4 to
// choose (read once per process).  On the
// 10 x 10 lambda patch at 16 vectors: 82.9, 84.2 and 94.1 us, with HBM
// traffic 1.75x, 1.28x and 1.2x the algorithmic bytes.
int near_mx_tiles_per_group() {
  static const int tpw = [] {
    const char *v = std::getenv("SPK_NEAR_MXF8_TPW");
    const int t = v ? std::atoi(v) : 1;
    return (t == 2 || t == 4) ? t : 1;
  }();
  return tpw;
}

// padded batch width: R rounded up to a power of two
inline int batch_pad(int R) { return R <= 1 ? 1 : R <= 2 ? 2 : R <= 4 ? 4 : R <= 8 ? 8 : 16; }

// ---------------------------------------------------------------------------
// Gram-Schmidt products of P independent Krylov bases in one launch each
// (the batched solves' orthogonalisation): basis r has k_r columns of n
// values at Vs[r] + j*n, its new vector is ws[r], its coefficients go to
// hs[r][0..k_r) and |w_r|^2 to hs[r][k_r].  Products accumulate in
// complex128, partial sums are added in a fixed order.
// ---------------------------------------------------------------------------
namespace {
template <class T>
__device__ inline zd bwiden(T v) {
  return zd{(double)v.x, (double)v.y};
}
__device__ inline zd bwave_sum(zd v) {
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_down(v.x, off, 64);
    v.y += __shfl_down(v.y, off, 64);
  }
  return v;
}
}  // namespace

constexpr int kOrthCT = 16;
// partial[(r * ldp + j) * Rc + chunk] = <V_r[:, j], w_r> over the chunk's rows
template <class T>
__global__ void __launch_bounds__(256) kb_gemv_c(int64_t n, int64_t rows, int ldp, const int *ks,
                                                 const T *const *Vs, const T *const *ws,
                                                 zd *partial) {
  __shared__ zd red[4][kOrthCT];
  const int r = blockIdx.z, k = ks[r];
  const int j0 = blockIdx.y * kOrthCT;
  if (j0 >= k) return;  // uniform for the workgroup
  const int nt = min(kOrthCT, k - j0), Rc = gridDim.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const T *V = Vs[r], *w = ws[r];
  const int64_t i0 = (int64_t)blockIdx.x * rows, i1 = min<int64_t>(n, i0 + rows);
  zd acc[kOrthCT];
#pragma unroll
  for (int t = 0; t < kOrthCT; ++t) acc[t] = zd{0, 0};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += 256) {
    const zd x = bwiden(w[i]);
#pragma unroll
    for (int t = 0; t < kOrthCT; ++t)
      if (t < nt) {
        const zd v = bwiden(V[(size_t)(j0 + t) * n + i]);
        acc[t] = acc[t] + conj(v) * x;
      }
  }
#pragma unroll
  for (int t = 0; t < kOrthCT; ++t) {
    const zd v = bwave_sum(acc[t]);
    if (lane == 0) red[wv][t] = v;
  }
  __syncthreads();
  if (threadIdx.x < nt) {
    const int t = threadIdx.x;
    partial[((size_t)r * ldp + j0 + t) * Rc + blockIdx.x] = red[0][t] + red[1][t] + red[2][t] + red[3][t];
  }
}
// hs[r][j] = sum of the Rc partials, one wave per (r, j)
__global__ void kb_gemv_c_final(int P, int ldp, int Rc, const int *ks, const zd *partial,
                                zd *const *hs) {
  const int o = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (o >= P * ldp) return;
  const int r = o / ldp, j = o % ldp;
  if (j >= ks[r]) return;
  zd acc = {0, 0};
  for (int c = lane; c < Rc; c += 64) acc = acc + partial[(size_t)o * Rc + c];
  acc = bwave_sum(acc);
  if (lane == 0) hs[r][j] = acc;
}
// w_r -= V_r h_r (thread per row, all k_r columns)
template <class T>
__global__ void __launch_bounds__(256) kb_gemv_n_sub(int64_t n, const int *ks, const T *const *Vs,
                                                     const zd *const *hs, T *const *ws) {
  const int r = blockIdx.y;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int k = ks[r];
  const T *V = Vs[r];
  const zd *h = hs[r];
  zd acc = {0, 0};
  int j = 0;
  for (; j + 8 <= k; j += 8) {
    zd v[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) v[t] = bwiden(V[(size_t)(j + t) * n + i]);
#pragma unroll
    for (int t = 0; t < 8; ++t) acc = acc + v[t] * h[j + t];
  }
  for (; j < k; ++j) acc = acc + bwiden(V[(size_t)j * n + i]) * h[j];
  const zd wv = bwiden(ws[r][i]) - acc;
  ws[r][i] = cvt<T>(wv);
}
// |w_r|^2: partial[r * G + block], then hs[r][k_r] = sum (imaginary part 0)
template <class T>
__global__ void __launch_bounds__(256) kb_norm2(int64_t n, T *const *ws, zd *partial) {
  __shared__ double red[4];
  const int r = blockIdx.y, G = gridDim.x;
  const T *w = ws[r];
  double acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)G * 256) {
    const zd v = bwiden(w[i]);
    acc += v.x * v.x + v.y * v.y;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[(size_t)r * G + blockIdx.x] = zd{red[0] + red[1] + red[2] + red[3], 0.0};
}
__global__ void kb_norm2_final(int G, const int *ks, const zd *partial, zd *const *hs) {
  const int r = blockIdx.x, lane = threadIdx.x;
  zd acc = {0, 0};
  for (int c = lane; c < G; c += 64) acc = acc + partial[(size_t)r * G + c];
  acc = bwave_sum(acc);
  if (lane == 0) hs[r][ks[r]] = acc;
}

constexpr int kOrthNormBlocks = 64;
size_t batch_orth_partial_elems(int64_t n, int P, int maxk) {
  const size_t c = (size_t)P * (size_t)maxk * 256;
  const size_t d = (size_t)P * kOrthNormBlocks;
  return c > d ? c : d;
}

template <class T>
void launch_batch_orth(int64_t n, int P, int maxk, const int *ks, const T *const *Vs, T *const *ws,
                       zd *const *hs, zd *partial, hipStream_t st) {
  if (P <= 0 || maxk <= 0) return;
  const int ctiles = (maxk + kOrthCT - 1) / kOrthCT;
  // row chunks: at least 4 rows per thread, about 2048 workgroups in all
  int64_t rc = std::max<int64_t>(1, std::min<int64_t>((n + 1023) / 1024,
                                                     (2048 + (int64_t)ctiles * P - 1) / ((int64_t)ctiles * P)));
  rc = std::min<int64_t>(rc, 256);
  const int64_t rows = (n + rc - 1) / rc;
  hipLaunchKernelGGL(kb_gemv_c<T>, dim3((unsigned)rc, ctiles, P), dim3(256), 0, st, n, rows, maxk,
                     ks, Vs, (const T *const *)ws, partial);
  hipLaunchKernelGGL(kb_gemv_c_final, dim3((unsigned)(((int64_t)P * maxk + 3) / 4)), dim3(256), 0,
                     st, P, maxk, (int)rc, ks, (const zd *)partial, hs);
  hipLaunchKernelGGL(kb_gemv_n_sub<T>, dim3(bnblk(n, 256), P), dim3(256), 0, st, n, ks, Vs,
                     (const zd *const *)hs, ws);
  hipLaunchKernelGGL(kb_norm2<T>, dim3(kOrthNormBlocks, P), dim3(256), 0, st, n, ws, partial);
  hipLaunchKernelGGL(kb_norm2_final, dim3(P), dim3(64), 0, st, kOrthNormBlocks, ks,
                     (const zd *)partial, hs);
}

// ---------------------------------------------------------------------------
// Device-side Arnoldi step of P near-field inner solves (the FGMRES
// preconditioner), one classical Gram-Schmidt pass in three launches:
//   dot:  partial[s][c][g] = sum over the rows of chunk g of conj(U_s[:, c]) w_s
//         (one workgroup per 1024-row chunk, column and slot)
//   upd:  h_s = sum_g partial (fixed order, in every workgroup), w_s -= U_s h_s,
//         pnrm[s][g] = |w_s|^2 over chunk g
//   fin:  |w_s| = sqrt(sum_g pnrm); H_s[:, j] = (h_s, |w_s|); U_s[:, j + 1] = w_s / |w_s|
// Slot s: U + s * ldu (column c at c * n), w: W + s * n, h: hb + s * ldh,
// H + s * (kin + 1) * kin (column-major, leading dimension kin + 1).
// Products accumulate in complex128.
// ---------------------------------------------------------------------------
constexpr int kArnRows = 1024;  // rows per workgroup of the dot kernel
// partial[s][c][g] = sum over rows [g * kArnRows, (g + 1) * kArnRows) of
// conj(U_s[:, c]) w_s: one workgroup per (row chunk, column, slot)
template <class T>
__global__ void __launch_bounds__(256) k_arn_dot(int64_t n, int k, int G, const T *U, size_t ldu,
                                                 const T *W, zd *partial) {
  __shared__ zd red[4];
  const int g = blockIdx.x, c = blockIdx.y, s = blockIdx.z;
  const T *u = U + (size_t)s * ldu + (size_t)c * n, *w = W + (size_t)s * n;
  const int64_t r0 = (int64_t)g * kArnRows, r1 = min(n, r0 + kArnRows);
  zd acc = {0, 0};
#pragma unroll 4
  for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {
    const zd a = bwiden(u[i]), b = bwiden(w[i]);
    acc.x += a.x * b.x + a.y * b.y;
    acc.y += a.x * b.y - a.y * b.x;
  }
  acc = bwave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[((size_t)s * k + c) * G + g] = red[0] + red[1] + red[2] + red[3];
}
template <class T>
__global__ void __launch_bounds__(256) k_arn_upd(int64_t n, int64_t rows, int k, int G, int Gd,
                                                 const T *U, size_t ldu, T *W, const zd *partial,
                                                 zd *hb, int ldh, double *pnrm) {
  __shared__ zd hs[64];
  __shared__ double red[4];
  const int s = blockIdx.y, g = blockIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // h_c = sum over the G partials: wave wv takes columns wv, wv + 4, ...;
  // lanes stride the partials, then a shuffle tree (a fixed order)
  for (int c = wv; c < k; c += 4) {
    const zd *p = partial + ((size_t)s * k + c) * Gd;
    zd acc = {0, 0};
    for (int q = lane; q < Gd; q += 64) acc = acc + p[q];
    acc = bwave_sum(acc);
    if (lane == 0) {
      hs[c] = acc;
      if (g == 0) hb[(size_t)s * ldh + c] = acc;
    }
  }
  __syncthreads();
  const T *Us = U + (size_t)s * ldu;
  T *w = W + (size_t)s * n;
  const int64_t r0 = (int64_t)g * rows, r1 = min(n, r0 + rows);
  double nrm = 0.0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {
    zd wi = bwiden(w[i]);
    for (int c = 0; c < k; ++c) {
      const zd u = bwiden(Us[(size_t)c * n + i]), hc = hs[c];
      wi.x -= u.x * hc.x - u.y * hc.y;
      wi.y -= u.x * hc.y + u.y * hc.x;
    }
    using F = decltype(w[i].x);
    const T wt = T{(F)wi.x, (F)wi.y};
    w[i] = wt;
    const zd ws = bwiden(wt);
    nrm += ws.x * ws.x + ws.y * ws.y;
  }
  for (int off = 32; off > 0; off >>= 1) nrm += __shfl_down(nrm, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nrm;
  __syncthreads();
  if (threadIdx.x == 0) pnrm[(size_t)s * G + g] = red[0] + red[1] + red[2] + red[3];
}
template <class T>
__global__ void __launch_bounds__(256) k_arn_fin(int64_t n, int64_t rows, int j, int kin, int G,
                                                 const double *pnrm, const zd *hb, int ldh,
                                                 const T *W, size_t ldu, T *U, zd *H) {
  __shared__ double nrm;
  const int s = blockIdx.y, g = blockIdx.x;
  if (threadIdx.x < 64) {
    double a = 0.0;
    for (int q = threadIdx.x; q < G; q += 64) a += pnrm[(size_t)s * G + q];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    if (threadIdx.x == 0) nrm = sqrt(fmax(a, 0.0));
  }
  __syncthreads();
  const double nv = nrm, inv = nv > 0.0 ? 1.0 / nv : 0.0;
  if (g == 0 && threadIdx.x <= j + 1) {
    const int i = threadIdx.x;
    H[(size_t)s * (kin + 1) * kin + (size_t)j * (kin + 1) + i] =
        i <= j ? hb[(size_t)s * ldh + i] : zd{nv, 0.0};
  }
  const T *w = W + (size_t)s * n;
  T *u = U + (size_t)s * ldu + (size_t)(j + 1) * n;
  const int64_t r0 = (int64_t)g * rows, r1 = min(n, r0 + rows);
  using F = decltype(w[0].x);
  for (int64_t i = r0 + threadIdx.x; i < r1; i += 256) {
    const T v = w[i];
    u[i] = T{(F)(v.x * inv), (F)(v.y * inv)};
  }
}
// row chunks of the update and normalisation (at most 256): 256 rows for
// fewer than 4 slots (more workgroups for a short grid), 512 otherwise
static int arnoldi_chunks(int64_t n, int P) {
  const int64_t rows = P < 4 ? 256 : 512;
  return (int)std::max<int64_t>(1, std::min<int64_t>(256, (n + rows - 1) / rows));
}
static int arnoldi_dot_chunks(int64_t n) { return (int)((n + kArnRows - 1) / kArnRows); }
size_t arnoldi_partial_elems(int64_t n, int P, int kin) {
  return (size_t)P * (kin + 1) * (size_t)std::max(256, arnoldi_dot_chunks(n)) + (size_t)P * 256;
}
template <class T>
void launch_arnoldi_cgs(int64_t n, int P, int j, int kin, T *U, size_t ldu, T *W, zd *hb, int ldh,
                        zd *H, zd *partial, hipStream_t st) {
  if (P <= 0) return;
  const int G = arnoldi_chunks(n, P), Gd = arnoldi_dot_chunks(n);
  const int64_t rows = (n + G - 1) / G;
  const int k = j + 1;
  double *pn = reinterpret_cast<double *>(partial + (size_t)P * (kin + 1) * std::max(256, Gd));
  hipLaunchKernelGGL(k_arn_dot<T>, dim3(Gd, k, P), dim3(256), 0, st, n, k, Gd, (const T *)U, ldu,
                     (const T *)W, partial);
  hipLaunchKernelGGL(k_arn_upd<T>, dim3(G, P), dim3(256), 0, st, n, rows, k, G, Gd, (const T *)U,
                     ldu, W, (const zd *)partial, hb, ldh, pn);
  hipLaunchKernelGGL(k_arn_fin<T>, dim3(G, P), dim3(256), 0, st, n, rows, j, kin, G,
                     (const double *)pn, (const zd *)hb, ldh, (const T *)W, ldu, U, H);
}

template <class T>
void launch_batch_interleave(int N, int R, const int *perm, const T *const *xs, T *XI, T *XP,
                             hipStream_t st) {
  const int RP = batch_pad(R);
  const unsigned g = bnblk((int64_t)2 * N * RP, 256);
  switch (RP) {
#define SPK_IL_CASE(P)                                                                          \
  case P:                                                                                       \
    hipLaunchKernelGGL((kb_interleave<T, P>), dim3(g), dim3(256), 0, st, N, R, perm, xs, XI, XP); \
    break;
    SPK_IL_CASE(1) SPK_IL_CASE(2) SPK_IL_CASE(4) SPK_IL_CASE(8) SPK_IL_CASE(16)
#undef SPK_IL_CASE
  }
}
template <class T>
void launch_batch_agg(int nbl, int S, int N, int R, const int *box_first, const int *box_count,
                      const int *perm, const T *rt, const T *XI, const T *XP, T *C,
                      hipStream_t st) {
  if (nbl <= 0) return;
  (void)XI;
  (void)perm;
  const dim3 g(bnblk(S, 256), nbl, 4);
  if constexpr (std::is_same<T, zf>::value) {
    if (batch_pad(R) == 16) {
      hipLaunchKernelGGL(kb_agg_mx, g, dim3(256), 0, st, S, R, box_first, box_count, rt, XP, C);
      return;
    }
  } else {
    if (batch_pad(R) == 16) {
      hipLaunchKernelGGL(kb_agg_mx64, dim3(bnblk(S, 128), nbl, 4), dim3(256), 0, st, S, R, box_first,
                         box_count, rt, XP, C);
      return;
    }
  }
  switch (batch_pad(R)) {
#define SPK_AGG_CASE(RP)                                                                       \
  case RP:                                                                                     \
    hipLaunchKernelGGL((kb_agg<T, RP>), g, dim3(256), 0, st, S, N, R, box_first, box_count,     \
                       rt, XP, C);                                                             \
    break;
    SPK_AGG_CASE(1) SPK_AGG_CASE(2) SPK_AGG_CASE(4) SPK_AGG_CASE(8) SPK_AGG_CASE(16)
#undef SPK_AGG_CASE
  }
}
template <class T>
void launch_batch_ss(int nb, int S, int R, const int *child, const T *Cc, const T *up, T *P,
                     hipStream_t st) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(kb_ss<T>, dim3(bnblk((int64_t)S * R, 256), nb), dim3(256), 0, st, S, R, child,
                     Cc, up, P);
}
template <class T>
void launch_batch_m2l(int nb, int S, int R, const int *i4_off, const int *i4_box,
                      const int *i4_slot, const T *B, const T *tlt, T *Dt, hipStream_t st) {
  if (nb <= 0) return;
  const int gx = (int)bnblk((int64_t)S * R, 128);
  const int64_t g = ((int64_t)gx * nb + 7) / 8 * 8;  // a multiple of the 8 XCDs
  hipLaunchKernelGGL(kb_m2l<T>, dim3((unsigned)g), dim3(128), 0, st, S, R, nb, gx, i4_off, i4_box,
                     i4_slot, B, tlt, Dt);
}
template <class T>
void launch_batch_rr(int nb, int S, int R, const int *parent, const int *octant, const T *Dp,
                     const T *down, T *Dc, hipStream_t st) {
  if (nb <= 0) return;
  hipLaunchKernelGGL(kb_rr<T>, dim3(bnblk((int64_t)S * R, 256), nb), dim3(256), 0, st, S, R,
                     parent, octant, Dp, down, Dc);
}
template <class T>
void launch_batch_fold(int nbl, int S, int R, const T *D, zd kd1, zd kd2, zd my1, zd my2, zd ep1,
                       zd ep2, T *W, hipStream_t st) {
  if (nbl <= 0) return;
  hipLaunchKernelGGL(kb_fold<T>, dim3(bnblk((int64_t)S * R, 256), nbl), dim3(256), 0, st, S, R, D,
                     cvt<T>(kd1), cvt<T>(kd2), cvt<T>(my1), cvt<T>(my2), cvt<T>(ep1), cvt<T>(ep2),
                     W);
}
// yf = sum over the kBatchSplit slices of yp, in slice order
template <class T>
__global__ void kb_sum_split(int64_t len, const T *yp, T *yf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  T acc = yp[i];
  for (int sp = 1; sp < kBatchSplit; ++sp) acc = acc + yp[(size_t)sp * len + i];
  yf[i] = acc;
}

template <class T>
void launch_batch_disagg(int ntiles, int N, int S, int R, const int *tile_box, const int *tile_p0,
                         const int *box_first, const int *box_count, const int *perm, const T *rt,
                         const T *W, zd sc1, zd sc2, T *yp, T *yf, int mx_tiles,
                         const int *mx_box, const int *mx_p0, hipStream_t st) {
  if (ntiles <= 0) return;
  // samples split over kBatchSplit workgroups per tile (multiples of 16 samples)
  const int sblk = (int)bnblk(bnblk(S, 16), kBatchSplit) * 16;
  const dim3 g(ntiles, kBatchSplit);
  bool done = false;
  if (batch_pad(R) == 16 && mx_tiles > 0) {
    if constexpr (std::is_same<T, zf>::value)
      hipLaunchKernelGGL(kb_disagg_mx, dim3(mx_tiles, kBatchSplit), dim3(512), 0, st, N, S, R,
                         mx_box, mx_p0, box_first, box_count, perm, rt, W, cvt<T>(sc1),
                         cvt<T>(sc2), sblk, yp);
    else
      hipLaunchKernelGGL(kb_disagg_mx64, dim3(mx_tiles, kBatchSplit), dim3(512), 0, st, N, S, R,
                         mx_box, mx_p0, box_first, box_count, perm, rt, W, sc1, sc2, sblk, yp);
    done = true;
  }
  if (!done) switch (batch_pad(R)) {
#define SPK_DIS_CASE(RP)                                                                        \
  case RP:                                                                                      \
    hipLaunchKernelGGL((kb_disagg<T, RP>), g, dim3(16 * RP), 0, st, N, S, R, tile_box, tile_p0, \
                       box_first, box_count, perm, rt, W, cvt<T>(sc1), cvt<T>(sc2), sblk, yp);  \
    break;
    SPK_DIS_CASE(1) SPK_DIS_CASE(2) SPK_DIS_CASE(4) SPK_DIS_CASE(8) SPK_DIS_CASE(16)
#undef SPK_DIS_CASE
  }
  const int64_t len = (int64_t)R * 2 * N;
  hipLaunchKernelGGL(kb_sum_split<T>, dim3(bnblk(len, 256)), dim3(256), 0, st, len, (const T *)yp,
                     yf);
}
template <class T>
void launch_batch_near(int nwg, int N, int R, const int *row_box, const int *row_r0,
                       const int *box_first, const int *box_count, const int *e2_off,
                       const int *e2_list, const int64_t *near_off, const int *perm,
                       const T *blocks, const T *XI, const T *yf, T *const *ys,
                       hipStream_t st, const uint32_t *blocks_bf) {
  if (nwg <= 0) return;
  if constexpr (std::is_same<T, zd>::value) {
    if (batch_pad(R) == 16) {
      hipLaunchKernelGGL(kb_near_mx64, dim3(nwg), dim3(256), 0, st, N, R, row_box, row_r0,
                         box_first, box_count, e2_off, e2_list, near_off, perm, blocks, XI, yf, ys);
      return;
    }
  }
  if constexpr (std::is_same<T, zf>::value) {
    if (batch_pad(R) == 16) {
      if (blocks_bf && !yf)  // the preconditioner's product: bf16 operands throughout
        hipLaunchKernelGGL(kb_near_mxb, dim3(nwg), dim3(256), 0, st, N, R, row_box, row_r0,
                           box_first, box_count, e2_off, e2_list, near_off, perm,
                           (const uint4 *)blocks_bf, XI, ys);
      else if (blocks_bf)
        hipLaunchKernelGGL(kb_near_mx<true>, dim3(nwg), dim3(256), 0, st, N, R, row_box, row_r0,
                           box_first, box_count, e2_off, e2_list, near_off, perm,
                           (const void *)blocks_bf, XI, yf, ys);
      else
        hipLaunchKernelGGL(kb_near_mx<false>, dim3(nwg), dim3(256), 0, st, N, R, row_box, row_r0,
                           box_first, box_count, e2_off, e2_list, near_off, perm,
                           (const void *)blocks, XI, yf, ys);
      return;
    }
  }
  switch (batch_pad(R)) {
#define SPK_NEAR_CASE(RP)                                                                       \
  case RP:                                                                                      \
    hipLaunchKernelGGL((kb_near<T, RP>), dim3(nwg), dim3(256), 0, st, N, R, row_box, row_r0,    \
                       box_first, box_count, e2_off, e2_list, near_off, perm, blocks, XI, yf,    \
                       ys);                                                                     \
    break;
    SPK_NEAR_CASE(1) SPK_NEAR_CASE(2) SPK_NEAR_CASE(4) SPK_NEAR_CASE(8) SPK_NEAR_CASE(16)
#undef SPK_NEAR_CASE
  }
}

#define SPK_BATCH_INSTANTIATE(T)                                                                  \
  template void launch_batch_interleave<T>(int, int, const int *, const T *const *, T *, T *,     \
                                           hipStream_t);                                          \
  template void launch_batch_agg<T>(int, int, int, int, const int *, const int *, const int *,     \
                                    const T *, const T *, const T *, T *, hipStream_t);            \
  template void launch_batch_ss<T>(int, int, int, const int *, const T *, const T *, T *,          \
                                   hipStream_t);                                                   \
  template void launch_batch_m2l<T>(int, int, int, const int *, const int *, const int *,          \
                                    const T *, const T *, T *, hipStream_t);                       \
  template void launch_batch_rr<T>(int, int, int, const int *, const int *, const T *, const T *,  \
                                   T *, hipStream_t);                                              \
  template void launch_batch_fold<T>(int, int, int, const T *, zd, zd, zd, zd, zd, zd, T *,        \
                                     hipStream_t);                                                 \
  template void launch_batch_disagg<T>(int, int, int, int, const int *, const int *, const int *,  \
                                       const int *, const int *, const T *, const T *, zd, zd,     \
                                       T *, T *, int, const int *, const int *, hipStream_t);      \
  template void launch_batch_near<T>(int, int, int, const int *, const int *, const int *,         \
                                     const int *, const int *, const int *, const int64_t *,       \
                                     const int *, const T *, const T *, const T *, T *const *,     \
                                     hipStream_t, const uint32_t *);
SPK_BATCH_INSTANTIATE(zd)
SPK_BATCH_INSTANTIATE(zf)
template void launch_batch_orth<zf>(int64_t, int, int, const int *, const zf *const *, zf *const *,
                                    zd *const *, zd *, hipStream_t);
template void launch_batch_orth<zd>(int64_t, int, int, const int *, const zd *const *, zd *const *,
                                    zd *const *, zd *, hipStream_t);
template void launch_arnoldi_cgs<zf>(int64_t, int, int, int, zf *, size_t, zf *, zd *, int, zd *,
                                     zd *, hipStream_t);
template void launch_arnoldi_cgs<zd>(int64_t, int, int, int, zd *, size_t, zd *, zd *, int, zd *,
                                     zd *, hipStream_t);

}  // namespace spk

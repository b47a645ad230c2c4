// HIP kernels of the KraNS-MI engine (gfx950 / CDNA4, wave64).
//
// Design rules (see engine.hpp for the launchers' contracts):
//  * No atomics anywhere: every reduction is a fixed tree (per-thread
//    sequential partial -> wave64 shuffle tree -> LDS across the block's 4
//    waves -> per-block partial -> one finishing block summing the partials
//    in index order). The block count depends only on the length, so results
//    are bitwise reproducible run to run.
//  * Sparse products are gathers: y = A x over the CSR rows, y = A^T x over
//    the CSC twin's columns. With one lane per row (short rows) each output
//    is summed sequentially in ascending inner index from 0.0 with no FMA
//    contraction, which is the rounding of the reference's Eigen products
//    (EigenSparseMatrixAdapter.hpp:34-57).
//  * Scalars (alpha, beta, norms, CG state) stay on the device; kernels read
//    them through pointers so the host can queue work without syncing.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <type_traits>

#include "engine.hpp"

namespace kh {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    v += __shfl_down(v, off, 64);
  return v;
}

// Block-wide sum; the result is valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
    lds[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < kWaves; ++w)
      t += lds[w];
  }
  __syncthreads();
  return t;
}

// Sum of pc[0 .. nb) in k_finish's order (per thread a stride-kBlock run
// with eight loads in flight, then block_sum); valid in thread 0. Every
// thread of the block must call it.
__device__ __forceinline__ double sum_partials(const double *pc, int nb, double *lds) {
  double acc = 0.0;
  int i = threadIdx.x;
  for (; i + 7 * kBlock < nb; i += 8 * kBlock) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      v[u] = pc[i + u * kBlock];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      acc += v[u];
  }
  for (; i < nb; i += kBlock)
    acc += pc[i];
  return block_sum(acc, lds);
}

template <class T> struct Acc;
template <> struct Acc<double> { using type = double; };
template <> struct Acc<float> { using type = float; };

// one thread per row, no cap (kernels that do not stride over the grid)
inline unsigned grid_rows(int64_t rows) {
  const int64_t g = (rows + kBlock - 1) / kBlock;
  if (g > ((int64_t)1 << 31) - 1)
    throw std::length_error("more rows than one launch covers");
  return (unsigned)(g < 1 ? 1 : g);
}

inline int grid_for(int64_t len, int per_thread = 1) {
  int64_t g = (len + (int64_t)kBlock * per_thread - 1) / ((int64_t)kBlock * per_thread);
  if (g < 1)
    g = 1;
  if (g > 16384)
    g = 16384;
  return (int)g;
}

// ----------------------------------------------------------------- sparse --
// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8), each with its own L2. The gather kernels renumber blocks so that
// XCD x takes the contiguous block range [x q, (x + 1) q), q = nb / 8 (the
// last nb % 8 blocks keep their numbers): an XCD's rows then touch the x
// entries of one contiguous eighth of the operator, which its L2 keeps,
// instead of every XCD fetching all of x. A bijection on [0, nb); it changes
// only which workgroup computes a row, never the row's arithmetic.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t b = blockIdx.x, nb = gridDim.x, q = nb >> 3;
  return b < (q << 3) ? (b & 7) * q + (b >> 3) : b;
}

// sum of val[p] * x(idx[p]) over p in [p0, p1), ascending from 0 with no
// contraction (the rounding of a sequential loop). Entries go in groups of
// four, then pairs, then a last single one; the index, value and x loads of
// a group are issued before its sums, so rows of two or four entries cost
// three dependent memory latencies (row pointers, entries, x) instead of two
// per entry, and no gather is issued for an entry that does not exist (a
// gather costs up to one cache line per lane).
template <class T, class XAt>
__device__ __forceinline__ typename Acc<T>::type row_dot(int32_t p0, int32_t p1,
                                                         const int32_t *__restrict__ idx,
                                                         const T *__restrict__ val, XAt xat) {
#pragma clang fp contract(off)
  using A = typename Acc<T>::type;
  A acc = (A)0;
  int32_t p = p0;
  for (; p + 4 <= p1; p += 4) {
    const int32_t j0 = idx[p], j1 = idx[p + 1], j2 = idx[p + 2], j3 = idx[p + 3];
    const T v0 = val[p], v1 = val[p + 1], v2 = val[p + 2], v3 = val[p + 3];
    const T x0 = xat(j0), x1 = xat(j1), x2 = xat(j2), x3 = xat(j3);
    acc += (A)v0 * (A)x0;
    acc += (A)v1 * (A)x1;
    acc += (A)v2 * (A)x2;
    acc += (A)v3 * (A)x3;
  }
  for (; p + 2 <= p1; p += 2) {
    const int32_t j0 = idx[p], j1 = idx[p + 1];
    const T v0 = val[p], v1 = val[p + 1];
    const T x0 = xat(j0), x1 = xat(j1);
    acc += (A)v0 * (A)x0;
    acc += (A)v1 * (A)x1;
  }
  if (p < p1)
    acc += (A)val[p] * (A)xat(idx[p]);
  return acc;
}

// One lane per row, one row per thread (the grid covers all rows, so every
// row's chain of loads is in flight at once up to the occupancy limit).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_spmv_lane(int64_t rows, const int32_t *__restrict__ ptr,
                const int32_t *__restrict__ idx, const T *__restrict__ val,
                const T *__restrict__ x, T *__restrict__ y,
                const T *__restrict__ z, const double *__restrict__ beta) {
#pragma clang fp contract(off)
  const int64_t i = xcd_block() * kBlock + threadIdx.x;
  if (i >= rows)
    return;
  const T b = z ? (T)(*beta) : (T)0;
  const T zi = z ? z[i] : (T)0;
  T out = (T)row_dot<T>(ptr[i], ptr[i + 1], idx, val, [=](int32_t j) { return x[j]; });
  if (z)
    out = out + (-b) * zi;
  y[i] = out;
}

// Rows of exactly R entries each (row pointers R i): row i's entries sit at
// R i .. R i + R - 1, so the lane loads them without the row pointers (one
// dependent memory latency fewer) and with vector loads where R allows. The
// sums are those of row_dot (ascending from 0, no contraction): results are
// bitwise those of k_spmv_lane / k_spmm_lane.
template <class T, int R>
__device__ __forceinline__ void ell_entries(int64_t i, const int32_t *__restrict__ idx,
                                            const T *__restrict__ val, int32_t (&j)[R],
                                            T (&v)[R]) {
  const int64_t p = i * R;
  if constexpr (R % 4 == 0) {
#pragma unroll
    for (int r = 0; r < R; r += 4) {
      const int4 a = *reinterpret_cast<const int4 *>(idx + p + r);
      j[r] = a.x, j[r + 1] = a.y, j[r + 2] = a.z, j[r + 3] = a.w;
    }
  } else if constexpr (R % 2 == 0) {
#pragma unroll
    for (int r = 0; r < R; r += 2) {
      const int2 a = *reinterpret_cast<const int2 *>(idx + p + r);
      j[r] = a.x, j[r + 1] = a.y;
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r)
      j[r] = idx[p + r];
  }
  if constexpr (std::is_same<T, float>::value && R % 4 == 0) {
#pragma unroll
    for (int r = 0; r < R; r += 4) {
      const float4 a = *reinterpret_cast<const float4 *>(val + p + r);
      v[r] = a.x, v[r + 1] = a.y, v[r + 2] = a.z, v[r + 3] = a.w;
    }
  } else if constexpr (R % 2 == 0) {
    using V2 = typename std::conditional<std::is_same<T, float>::value, float2, double2>::type;
#pragma unroll
    for (int r = 0; r < R; r += 2) {
      const V2 a = *reinterpret_cast<const V2 *>(val + p + r);
      v[r] = a.x, v[r + 1] = a.y;
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r)
      v[r] = val[p + r];
  }
}

// act (optional): one column's flag; the launch writes nothing when it is 0.
template <class T, int R>
__global__ void __launch_bounds__(kBlock)
    k_spmv_ell(int64_t rows, const int32_t *__restrict__ idx, const T *__restrict__ val,
               const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ z,
               const double *__restrict__ beta, const int *__restrict__ act) {
#pragma clang fp contract(off)
  const int64_t i = xcd_block() * kBlock + threadIdx.x;
  if (i >= rows || (act && !*act))
    return;
  const T b = z ? (T)(*beta) : (T)0;
  const T zi = z ? z[i] : (T)0;
  int32_t j[R];
  T v[R], xv[R];
  ell_entries<T, R>(i, idx, val, j, v);
#pragma unroll
  for (int r = 0; r < R; ++r)
    xv[r] = x[j[r]];
  T out = (T)0;
#pragma unroll
  for (int r = 0; r < R; ++r)
    out += v[r] * xv[r];
  if (z)
    out = out + (-b) * zi;
  y[i] = out;
}

// Multi-RHS form: the row's entries are loaded once for all columns.
template <class T, int R>
__global__ void __launch_bounds__(kBlock)
    k_spmm_ell(int64_t rows, const int32_t *__restrict__ idx, const T *__restrict__ val, int s,
               const T *__restrict__ x, int64_t ldx, T *__restrict__ y, int64_t ldy,
               const int *__restrict__ active) {
#pragma clang fp contract(off)
  const int64_t i = xcd_block() * kBlock + threadIdx.x;
  if (i >= rows)
    return;
  int32_t j[R];
  T v[R];
  ell_entries<T, R>(i, idx, val, j, v);
  for (int c = 0; c < s; ++c) {
    if (active && !active[c])
      continue;
    const T *xc = x + (int64_t)c * ldx;
    T xv[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
      xv[r] = xc[j[r]];
    T out = (T)0;
#pragma unroll
    for (int r = 0; r < R; ++r)
      out += v[r] * xv[r];
    y[i + (int64_t)c * ldy] = out;
  }
}

// LDS-staged gathers. Rows go in blocks of ROWS * kSpanThreads rows; the
// entries of block b reference x only within the window [s_b, s_b + len)
// (s_b a multiple of 4, from the operator's span table, len the largest
// window of the table), so the block copies its window of x into LDS and
// gathers from LDS: one L2 request per 128 bytes of window instead of one
// per entry whose cache line the gather misses in L1. A thread first issues
// the loads of its rows' entries, then its FW 16-byte loads of the window
// (all in flight together), stores them to LDS, and after the barrier
// gathers from LDS. The row arithmetic is k_spmv_ell's: y is bitwise the
// same.
constexpr int kSpanThreads = 512;

template <class T, int R, int ROWS, int FW>
__global__ void __launch_bounds__(kSpanThreads)
    k_spmv_ell_span(int64_t rows, const int32_t *__restrict__ span, int len, int64_t lenx,
                    const int32_t *__restrict__ idx, const T *__restrict__ val,
                    const T *__restrict__ x, T *__restrict__ y, const T *__restrict__ z,
                    const double *__restrict__ beta, const int *__restrict__ act) {
#pragma clang fp contract(off)
  if (act && !*act)  // one column's flag: the launch writes nothing when it is 0
    return;
  extern __shared__ __align__(16) unsigned char span_lds[];
  T *xs = reinterpret_cast<T *>(span_lds);
  const int64_t b = xcd_block();
  const int64_t r0 = b * (int64_t)(ROWS * kSpanThreads) + threadIdx.x;
  int32_t j[ROWS][R];
  T v[ROWS][R], zi[ROWS];
#pragma unroll
  for (int k = 0; k < ROWS; ++k) {
    const int64_t r = r0 + (int64_t)k * kSpanThreads;
    // rows past the end take the block's last row (inside its window); their
    // results are not stored
    const int64_t i = r < rows ? r : rows - 1;
    ell_entries<T, R>(i, idx, val, j[k], v[k]);
    zi[k] = z ? z[i] : (T)0;
  }
  const T bt = z ? (T)(*beta) : (T)0;
  const int64_t s0 = span[b];
  const int64_t L = min((int64_t)len, lenx - s0);
  using V16 = typename std::conditional<std::is_same<T, float>::value, float4, double2>::type;
  constexpr int E = 16 / (int)sizeof(T);
  const int64_t nv = L / E;  // 0 only when x ends less than 16 bytes past s_b
  const V16 *xw = reinterpret_cast<const V16 *>(x + s0);
  V16 *xs16 = reinterpret_cast<V16 *>(xs);
  if (nv > 0) {
    V16 t[FW];
#pragma unroll
    for (int u = 0; u < FW; ++u) {  // indices past the window repeat its last load
      const int64_t e = threadIdx.x + (int64_t)u * kSpanThreads;
      t[u] = xw[e < nv ? e : nv - 1];
    }
    // unconditional stores (past the window a thread rewrites the last
    // vector with the value already there): a store under a condition would
    // draw its load into the condition, one load in flight at a time
#pragma unroll
    for (int u = 0; u < FW; ++u) {
      const int64_t e = threadIdx.x + (int64_t)u * kSpanThreads;
      xs16[e < nv ? e : nv - 1] = t[u];
    }
  }
  for (int64_t e = nv * E + threadIdx.x; e < L; e += kSpanThreads)
    xs[e] = x[s0 + e];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < ROWS; ++k) {
    const int64_t i = r0 + (int64_t)k * kSpanThreads;
    T xv[R];
#pragma unroll
    for (int r = 0; r < R; ++r)
      xv[r] = xs[j[k][r] - s0];
    T out = (T)0;
#pragma unroll
    for (int r = 0; r < R; ++r)
      out += v[k][r] * xv[r];
    if (z)
      out = out + (-bt) * zi[k];
    if (i < rows)
      y[i] = out;
  }
}

// G lanes per row (G in 4,16,64): strided partial sums, shuffle tree.
template <class T, int G>
__global__ void __launch_bounds__(kBlock)
    k_spmv_group(int64_t rows, const int32_t *__restrict__ ptr,
                 const int32_t *__restrict__ idx, const T *__restrict__ val,
                 const T *__restrict__ x, T *__restrict__ y,
                 const T *__restrict__ z, const double *__restrict__ beta) {
#pragma clang fp contract(off)
  using A = typename Acc<T>::type;
  const int sub = threadIdx.x % G;
  const T b = z ? (T)(*beta) : (T)0;
  const int64_t groups_per_grid = (int64_t)gridDim.x * (kBlock / G);
  for (int64_t i = (int64_t)blockIdx.x * (kBlock / G) + threadIdx.x / G; i < rows;
       i += groups_per_grid) {
    const int32_t p0 = ptr[i], p1 = ptr[i + 1];
    A acc = (A)0;
    for (int32_t p = p0 + sub; p < p1; p += G)
      acc += (A)val[p] * (A)x[idx[p]];
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
      acc += __shfl_down(acc, off, G);
    if (sub == 0) {
      T out = (T)acc;
      if (z)
        out = out + (-b) * z[i];
      y[i] = out;
    }
  }
}

// Multi-RHS gather, one lane per row (one row per thread), columns c < s
// with active[c] != 0.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_spmm_lane(int64_t rows, const int32_t *__restrict__ ptr,
                const int32_t *__restrict__ idx, const T *__restrict__ val,
                int s, const T *__restrict__ x, int64_t ldx, T *__restrict__ y,
                int64_t ldy, const int *__restrict__ active) {
  const int64_t i = xcd_block() * kBlock + threadIdx.x;
  if (i >= rows)
    return;
  const int32_t p0 = ptr[i], p1 = ptr[i + 1];
  for (int c = 0; c < s; ++c) {
    if (active && !active[c])
      continue;
    const T *xc = x + (int64_t)c * ldx;
    y[i + (int64_t)c * ldy] = (T)row_dot<T>(p0, p1, idx, val, [=](int32_t j) { return xc[j]; });
  }
}

template <class T, int G>
__global__ void __launch_bounds__(kBlock)
    k_spmm_group(int64_t rows, const int32_t *__restrict__ ptr,
                 const int32_t *__restrict__ idx, const T *__restrict__ val,
                 int s, const T *__restrict__ x, int64_t ldx,
                 T *__restrict__ y, int64_t ldy, const int *__restrict__ active) {
#pragma clang fp contract(off)
  using A = typename Acc<T>::type;
  const int sub = threadIdx.x % G;
  const int64_t groups_per_grid = (int64_t)gridDim.x * (kBlock / G);
  for (int64_t i = (int64_t)blockIdx.x * (kBlock / G) + threadIdx.x / G; i < rows;
       i += groups_per_grid) {
    const int32_t p0 = ptr[i], p1 = ptr[i + 1];
    for (int c = 0; c < s; ++c) {
      if (active && !active[c])
        continue;
      const T *xc = x + (int64_t)c * ldx;
      A acc = (A)0;
      for (int32_t p = p0 + sub; p < p1; p += G)
        acc += (A)val[p] * (A)xc[idx[p]];
#pragma unroll
      for (int off = G / 2; off > 0; off >>= 1)
        acc += __shfl_down(acc, off, G);
      if (sub == 0)
        y[i + (int64_t)c * ldy] = (T)acc;
    }
  }
}

// ------------------------------------------------------------- reductions --
template <class TA, class TB>
__global__ void __launch_bounds__(kBlock)
    k_dot_partial(int64_t len, const TA *__restrict__ a, const TB *__restrict__ b,
                  double *__restrict__ part) {
  __shared__ double lds[kWaves];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock)
    acc += (double)a[i] * (double)b[i];
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0)
    part[blockIdx.x] = acc;
}

// out[col] = sum of part[col*nb .. col*nb+nb) in index order (one block per
// column); optional sqrt(max(0, .)).
__global__ void __launch_bounds__(kBlock)
    k_finish(const double *__restrict__ part, int nb, double *__restrict__ out,
             int do_sqrt) {
  __shared__ double lds[kWaves];
  const double acc = sum_partials(part + (int64_t)blockIdx.x * nb, nb, lds);
  if (threadIdx.x == 0)
    out[blockIdx.x] = do_sqrt ? sqrt(fmax(0.0, acc)) : acc;
}

// Whether a classical Gram-Schmidt pass needs its second pass, by the test of
// Daniel, Gragg, Kaufman and Stewart (Math. Comp. 30 (1976) 772-795) with
// eta = 1/sqrt(2), the test ARPACK's Lanczos and Arnoldi steps use: the
// second pass is skipped when the first kept ||w'|| >= eta ||w||, i.e. when
// it cancelled less than half of ||w||^2; then w' is orthogonal to V to
// working precision already and a second pass changes it only by rounding.
// g[0 .. s) are the first pass's coefficients h1 = V^T w and g[s] = ||w||^2
// (both FP64 sums); ||w'||^2 = ||w||^2 - ||h1||^2 for orthonormal V, so the
// pass is kept when ||h1||^2 < cancel ||w||^2 with cancel = 1 - eta^2. Each
// wave sums ||h1||^2 the same way (lane-strided partial sums, the shuffle
// tree, lane 0's total broadcast), so every wave of every block of every
// launch reaches the same decision.
__device__ __forceinline__ bool dgks_skip(const double *__restrict__ g, int s,
                                          double cancel = 0.5) {
  const int lane = threadIdx.x & 63;
  double a = 0.0;
  for (int i = lane; i < s; i += 64) {
    const double v = g[i];
    a += v * v;
  }
  a = __shfl(wave_sum(a), 0, 64);
  const double w2 = g[s];
  return isfinite(a) && isfinite(w2) && w2 > 0.0 && a < cancel * w2;
}

// k_finish for the second-pass coefficients of a selective CGS pair: with
// the second pass skipped (dgks_skip over g and s = gridDim.x columns) the
// coefficients are zero and the partials are not read.
__global__ void __launch_bounds__(kBlock)
    k_finish_pass2(const double *__restrict__ part, int nb, double *__restrict__ out,
                   const double *__restrict__ g) {
  __shared__ double lds[kWaves];
  if (dgks_skip(g, (int)gridDim.x)) {
    if (threadIdx.x == 0)
      out[blockIdx.x] = 0.0;
    return;
  }
  const double acc = sum_partials(part + (int64_t)blockIdx.x * nb, nb, lds);
  if (threadIdx.x == 0)
    out[blockIdx.x] = acc;
}

// Reorthogonalisation products with memory-level parallelism: a block owns
// kRB = 256*kR consecutive rows; thread t owns rows r0+t+k*256 (k < kR), so
// for every column the block issues kR coalesced loads per thread before any
// arithmetic. All s columns of the block's chunk are handled by the block
// (w is read once); per column a fixed wave64 tree and a 4-wave LDS sum give
// the block partial. part layout [c * nb + block].
constexpr int kR = 8;
constexpr int kRB = kBlock * kR;

// Columns per step: enough independent loads in flight per thread (kR rows x
// CU columns) to cover HBM latency with 2-4 blocks per CU.
template <class TV> struct ColUnroll { static constexpr int value = sizeof(TV) == 4 ? 4 : 2; };

template <class TV, class TW>
__global__ void __launch_bounds__(kBlock)
    k_vtw_partial(int64_t len, const TV *__restrict__ V, int64_t ld, int s,
                  const TW *__restrict__ w, double *__restrict__ part, int nb, int wsq) {
  constexpr int CU = ColUnroll<TV>::value;
  extern __shared__ double red[];  // [s][kWaves]
  const int64_t r0 = (int64_t)blockIdx.x * kRB + threadIdx.x;
  // Row indices are clamped into range (rows past the end get weight 0), so
  // every load below is unconditional and the loads of a step issue
  // back to back.
  int64_t rr[kR];
  double wv[kR];
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    const int64_t r = r0 + (int64_t)k * kBlock;
    rr[k] = r < len ? r : len - 1;
    const double wr = (double)w[rr[k]];
    wv[k] = r < len ? wr : 0.0;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int s_main = s - s % CU;
  for (int c0 = 0; c0 < s_main; c0 += CU) {
    TV x[CU][kR];
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const TV *vc = V + (int64_t)(c0 + u) * ld;
#pragma unroll
      for (int k = 0; k < kR; ++k)
        x[u][k] = vc[rr[k]];
    }
    // keep the CU*kR loads of this step together ahead of their uses
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < kR; ++k)
        acc += (double)x[u][k] * wv[k];
      acc = wave_sum(acc);
      if (lane == 0)
        red[(c0 + u) * kWaves + wid] = acc;
    }
  }
  for (int c = s_main; c < s; ++c) {
    const TV *vc = V + (int64_t)c * ld;
    TV x[kR];
#pragma unroll
    for (int k = 0; k < kR; ++k)
      x[k] = vc[rr[k]];
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kR; ++k)
      acc += (double)x[k] * wv[k];
    acc = wave_sum(acc);
    if (lane == 0)
      red[c * kWaves + wid] = acc;
  }
  if (wsq) {  // ||w||^2 of the block's rows as column s
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < kR; ++k)
      acc += wv[k] * wv[k];
    acc = wave_sum(acc);
    if (lane == 0)
      red[s * kWaves + wid] = acc;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < s + wsq; c += kBlock) {
    double t = 0.0;
    for (int w2 = 0; w2 < kWaves; ++w2)
      t += red[c * kWaves + w2];
    part[(int64_t)c * nb + blockIdx.x] = t;
  }
}

// FP32 V^T w with 16-byte loads: a block owns 2048 rows as 512 float4
// groups; thread t owns groups t and t + 256 (rows 4t..4t+3 and
// 1024+4t..1024+4t+3 of the block), two columns per step. Same block count
// and partial layout as k_vtw_partial; len and ld multiples of 4, V and w
// 16-byte aligned.
__global__ void __launch_bounds__(kBlock)
    k_vtw4_partial(int64_t len, const float *__restrict__ V, int64_t ld, int s,
                   const float *__restrict__ w, double *__restrict__ part, int nb, int wsq) {
  constexpr int G = 2, CU = 2;
  extern __shared__ double red[];  // [s][kWaves]
  const int64_t nq = len >> 2;
  const int64_t q0 = (int64_t)blockIdx.x * kBlock * G + threadIdx.x;
  int64_t qq[G];
  double wv[G][4];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int64_t q = q0 + (int64_t)g * kBlock;
    qq[g] = q < nq ? q : nq - 1;
    const float4 t = reinterpret_cast<const float4 *>(w)[qq[g]];
    const bool ok = q < nq;
    wv[g][0] = ok ? (double)t.x : 0.0;
    wv[g][1] = ok ? (double)t.y : 0.0;
    wv[g][2] = ok ? (double)t.z : 0.0;
    wv[g][3] = ok ? (double)t.w : 0.0;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int s_main = s - s % CU;
  for (int c0 = 0; c0 < s_main; c0 += CU) {
    float4 x[CU][G];
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const float4 *vc = reinterpret_cast<const float4 *>(V + (int64_t)(c0 + u) * ld);
#pragma unroll
      for (int g = 0; g < G; ++g)
        x[u][g] = vc[qq[g]];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      double acc = 0.0;
#pragma unroll
      for (int g = 0; g < G; ++g) {
        acc += (double)x[u][g].x * wv[g][0];
        acc += (double)x[u][g].y * wv[g][1];
        acc += (double)x[u][g].z * wv[g][2];
        acc += (double)x[u][g].w * wv[g][3];
      }
      acc = wave_sum(acc);
      if (lane == 0)
        red[(c0 + u) * kWaves + wid] = acc;
    }
  }
  for (int c = s_main; c < s; ++c) {
    const float4 *vc = reinterpret_cast<const float4 *>(V + (int64_t)c * ld);
    float4 x[G];
#pragma unroll
    for (int g = 0; g < G; ++g)
      x[g] = vc[qq[g]];
    double acc = 0.0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      acc += (double)x[g].x * wv[g][0];
      acc += (double)x[g].y * wv[g][1];
      acc += (double)x[g].z * wv[g][2];
      acc += (double)x[g].w * wv[g][3];
    }
    acc = wave_sum(acc);
    if (lane == 0)
      red[c * kWaves + wid] = acc;
  }
  if (wsq) {  // ||w||^2 of the block's rows as column s
    double acc = 0.0;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        acc += wv[g][e] * wv[g][e];
    acc = wave_sum(acc);
    if (lane == 0)
      red[s * kWaves + wid] = acc;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < s + wsq; c += kBlock) {
    double t = 0.0;
    for (int w2 = 0; w2 < kWaves; ++w2)
      t += red[c * kWaves + w2];
    part[(int64_t)c * nb + blockIdx.x] = t;
  }
}

// w -= V h over the same row ownership; optional partial of ||w_new||^2.
template <class TV, class TW>
__global__ void __launch_bounds__(kBlock)
    k_vh_sub(int64_t len, const TV *__restrict__ V, int64_t ld, int s,
             const double *__restrict__ h, TW *__restrict__ w, double *__restrict__ part,
             const double *__restrict__ g, double *__restrict__ ran, double cancel) {
  constexpr int CU = ColUnroll<TV>::value;
  extern __shared__ double hs[];  // [s]
  __shared__ double lds[kWaves];
  // With g (the last sweep of a selective CGS pair; g = first-pass
  // coefficients and ||w||^2): when dgks_skip holds, the middle sweep did
  // nothing and this launch applies the first pass's update w -= V h1
  // instead of the second pass's (the arithmetic of the middle sweep's
  // update). ran[0] records whether the second pass ran (1) or not (0).
  const bool skip = g != nullptr && dgks_skip(g, s, cancel);
  const double *hsrc = skip ? g : h;
  for (int j = threadIdx.x; j < s; j += kBlock)
    hs[j] = hsrc[j];
  if (ran && blockIdx.x == 0 && threadIdx.x == 0)
    ran[0] = skip ? 0.0 : 1.0;
  __syncthreads();
  const int64_t r0 = (int64_t)blockIdx.x * kRB + threadIdx.x;
  int64_t rr[kR];
  double acc[kR];
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    const int64_t r = r0 + (int64_t)k * kBlock;
    rr[k] = r < len ? r : len - 1;
    acc[k] = 0.0;
  }
  const int s_main = s - s % CU;
  for (int c0 = 0; c0 < s_main; c0 += CU) {
    TV x[CU][kR];
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const TV *vc = V + (int64_t)(c0 + u) * ld;
#pragma unroll
      for (int k = 0; k < kR; ++k)
        x[u][k] = vc[rr[k]];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const double hc = hs[c0 + u];
#pragma unroll
      for (int k = 0; k < kR; ++k)
        acc[k] += (double)x[u][k] * hc;
    }
  }
  for (int c = s_main; c < s; ++c) {
    const TV *vc = V + (int64_t)c * ld;
    const double hc = hs[c];
#pragma unroll
    for (int k = 0; k < kR; ++k)
      acc[k] += (double)vc[rr[k]] * hc;
  }
  TW wold[kR];
#pragma unroll
  for (int k = 0; k < kR; ++k)
    wold[k] = w[rr[k]];
  double nacc = 0.0;
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    const int64_t r = r0 + (int64_t)k * kBlock;
    const TW nw = (TW)((double)wold[k] - acc[k]);
    if (r < len) {
      w[r] = nw;
      nacc += (double)nw * (double)nw;
    }
  }
  if (part) {
    nacc = block_sum(nacc, lds);
    if (threadIdx.x == 0)
      part[blockIdx.x] = nacc;
  }
}

// The middle of a classical Gram-Schmidt pair with the block's tile of V in
// LDS: a block owns 128 rows and all s columns; thread t loads row t % 128 of
// column half t / 128 (every load issued before any use), the tile is stored
// to LDS (column-major, stride 129), threads < 128 then update their row's w
// exactly as k_vh_sub does (sequential sum over all columns in order), and
// all threads sum the tile's columns against the new w: with cw the smallest
// of 16, 32, 64, 128 not below s, thread t takes column t % cw over the
// cw / 2 rows of row group t / cw (row order), so no chain is longer than
// cw / 2 products. Each thread accumulates its (column, group) over the
// block's tiles; the groups of a column are then summed in group order. HBM
// sees V once. With MULTI a block walks tiles blockIdx.x, blockIdx.x +
// gridDim.x, ... (the next tile's loads in flight while the current one is
// processed), so the partials of V^T w are one per block: part[c * gridDim.x
// + block]. Without MULTI the grid must cover every tile (one each; the
// prefetch and its registers are compiled out).
constexpr int kLdsRows = 128, kLdsStride = kLdsRows + 1;

// With VEC (FP32 V, len and ld multiples of 4, V 16-byte aligned) the tile
// is loaded with 16-byte loads instead: thread t takes rows 4 (t % 32) ..
// 4 (t % 32) + 3 of columns t / 32, t / 32 + 8, ... (HMAX / 4 loads).
template <class TV, class TW, int HMAX, bool MULTI, bool VEC>
__global__ void __launch_bounds__(kBlock)
    k_vh_sub_vtw_lds(int64_t len, const TV *__restrict__ V, int64_t ld, int s,
                     const double *__restrict__ h, TW *__restrict__ w, double *__restrict__ part,
                     int nb, double *__restrict__ npart, const double *__restrict__ g) {
  static_assert(!VEC || (std::is_same<TV, float>::value && HMAX % 4 == 0),
                "16-byte tile loads are for FP32 blocks");
  // selective CGS pair: no second pass, nothing to do here (the last sweep
  // applies the first pass's update)
  if (g != nullptr && dgks_skip(g, s))
    return;
  constexpr int NV = VEC ? HMAX / 4 : 1;  // 16-byte loads per thread
  extern __shared__ double sh[];
  double *wn = sh;  // [kLdsRows]
  TV *tile = reinterpret_cast<TV *>(wn + kLdsRows);  // [s][kLdsStride]
  const int t = threadIdx.x, rl = t & (kLdsRows - 1), half = t / kLdsRows;
  const int sh2 = (s + 1) >> 1;
  const int c_lo = half ? sh2 : 0, c_hi = half ? s : sh2;
  const int64_t ntiles = (len + kLdsRows - 1) / kLdsRows;
  const int cw = s <= 16 ? 16 : s <= 32 ? 32 : s <= 64 ? 64 : 128;
  const int col = t & (cw - 1), grp = t / cw, rpg = cw / 2;
  double colsum = 0.0;  // (column col, row group grp) over this block's tiles
  double nsum = 0.0;    // with npart: ||updated w||^2 over this block's rows (wave 2)
  // the next tile's loads are issued as soon as the current tile is in LDS,
  // so they are in flight while its two passes run
  TV x[VEC ? 1 : HMAX];
  float4 x4[NV];
  TW wnext = (TW)0;
  const int q4 = t & 31, cg = t >> 5;  // VEC: row quad, first column
  auto load = [&](int64_t ti) {
    // an opaque copy of ld (empty asm, scalar register) keeps the compiler
    // from hoisting the column offsets out of the tile loop into vector
    // registers
    int64_t ldo = ld;
    __asm__ volatile("" : "+s"(ldo));
    if constexpr (VEC) {
      int64_t r4 = ti * kLdsRows + 4 * q4;
      if (r4 >= len)
        r4 = len - 4;  // len % 4 == 0: rows r4 .. r4 + 3 exist
      const TV *base = V + r4;
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        const int c = u * 8 + cg;
        const int cc = c < s ? c : s - 1;
        x4[u] = *reinterpret_cast<const float4 *>(base + (int64_t)cc * ldo);
      }
    } else {
      const int64_t r = ti * kLdsRows + rl;
      const TV *base = V + (r < len ? r : len - 1);
#pragma unroll
      for (int u = 0; u < HMAX; ++u) {
        const int c = c_lo + u;
        const int cc = c < c_hi ? c : (c_hi > c_lo ? c_hi - 1 : 0);
        x[u] = base[(int64_t)cc * ldo];
      }
    }
    if (half == 0) {
      const int64_t r = ti * kLdsRows + rl;
      wnext = w[r < len ? r : len - 1];
    }
  };
  if ((int64_t)blockIdx.x < ntiles)
    load(blockIdx.x);
  for (int64_t tile_i = blockIdx.x; tile_i < ntiles; tile_i += gridDim.x) {
    const int64_t r = tile_i * kLdsRows + rl;
    __syncthreads();  // the previous tile's LDS reads are done
    if constexpr (VEC) {
#pragma unroll
      for (int u = 0; u < NV; ++u) {
        const int c = u * 8 + cg;
        if (c < s) {
          TV *d = tile + c * kLdsStride + 4 * q4;
          d[0] = x4[u].x, d[1] = x4[u].y, d[2] = x4[u].z, d[3] = x4[u].w;
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < HMAX; ++u) {
        const int c = c_lo + u;
        if (c < c_hi)
          tile[c * kLdsStride + rl] = x[u];
      }
    }
    const TW wold = wnext;
    __syncthreads();
    if (MULTI && tile_i + (int64_t)gridDim.x < ntiles)
      load(tile_i + gridDim.x);
    if (half == 0) {
      // h[c] is the same for every lane: scalar loads, no LDS traffic
      double acc = 0.0;
#pragma unroll 2
      for (int c = 0; c < s; ++c)
        acc += (double)tile[c * kLdsStride + rl] * h[c];
      const TW nw = (TW)((double)wold - acc);
      if (r < len)
        w[r] = nw;
      wn[rl] = r < len ? (double)nw : 0.0;
    }
    __syncthreads();
    if (col < s) {  // s <= 128
      const TV *tc = tile + col * kLdsStride + grp * rpg;
      const double *wg = wn + grp * rpg;
      double a = 0.0;
#pragma unroll 4
      for (int i = 0; i < rpg; ++i)
        a += (double)tc[i] * wg[i];
      colsum += a;
    }
    if (npart && t >= 128 && t < 192) {  // rows l and l + 64, then a fixed wave tree
      const int l = t - 128;
      double v = wn[l] * wn[l] + wn[l + 64] * wn[l + 64];
      v = wave_sum(v);
      nsum += v;
    }
  }
  // the row groups of each column, in group order (the tile's LDS is free)
  __syncthreads();
  double *red = reinterpret_cast<double *>(tile);  // [kBlock / cw][cw]
  red[t] = colsum;
  __syncthreads();
  if (t < s) {
    double a = 0.0;
    for (int g = 0; g < kBlock / cw; ++g)
      a += red[g * cw + t];
    part[(int64_t)t * nb + blockIdx.x] = a;
  }
  if (npart && t == 128)
    npart[blockIdx.x] = nsum;
}

// The middle of a classical Gram-Schmidt pair in one read of V: w -= V h
// (the update of pass 1, exactly as k_vh_sub computes it) and then the block
// partials of V^T w for the updated w (the projection of pass 2, exactly as
// k_vtw_partial computes it from the stored w). The second column sweep
// re-reads the rows this thread just read, which are still in cache, so HBM
// sees V once instead of twice. part layout [c * nb + block].
template <class TV, class TW>
__global__ void __launch_bounds__(kBlock)
    k_vh_sub_vtw(int64_t len, const TV *__restrict__ V, int64_t ld, int s,
                 const double *__restrict__ h, TW *__restrict__ w, double *__restrict__ part,
                 int nb, const double *__restrict__ g) {
  if (g != nullptr && dgks_skip(g, s))  // selective CGS pair, no second pass
    return;
  constexpr int CU = ColUnroll<TV>::value;
  extern __shared__ double sh[];  // hs[s], then red[s][kWaves]
  double *hs = sh;
  double *red = sh + s;
  for (int j = threadIdx.x; j < s; j += kBlock)
    hs[j] = h[j];
  __syncthreads();
  const int64_t r0 = (int64_t)blockIdx.x * kRB + threadIdx.x;
  int64_t rr[kR];
  double acc[kR];
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    const int64_t r = r0 + (int64_t)k * kBlock;
    rr[k] = r < len ? r : len - 1;
    acc[k] = 0.0;
  }
  const int s_main = s - s % CU;
  // pass-1 update
  for (int c0 = 0; c0 < s_main; c0 += CU) {
    TV x[CU][kR];
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const TV *vc = V + (int64_t)(c0 + u) * ld;
#pragma unroll
      for (int k = 0; k < kR; ++k)
        x[u][k] = vc[rr[k]];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const double hc = hs[c0 + u];
#pragma unroll
      for (int k = 0; k < kR; ++k)
        acc[k] += (double)x[u][k] * hc;
    }
  }
  for (int c = s_main; c < s; ++c) {
    const TV *vc = V + (int64_t)c * ld;
    const double hc = hs[c];
#pragma unroll
    for (int k = 0; k < kR; ++k)
      acc[k] += (double)vc[rr[k]] * hc;
  }
  double wv[kR];
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    const int64_t r = r0 + (int64_t)k * kBlock;
    const TW nw = (TW)((double)w[rr[k]] - acc[k]);
    if (r < len)
      w[r] = nw;
    wv[k] = r < len ? (double)nw : 0.0;
  }
  // pass-2 projection of the updated w
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int c0 = 0; c0 < s_main; c0 += CU) {
    TV x[CU][kR];
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const TV *vc = V + (int64_t)(c0 + u) * ld;
#pragma unroll
      for (int k = 0; k < kR; ++k)
        x[u][k] = vc[rr[k]];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      double a = 0.0;
#pragma unroll
      for (int k = 0; k < kR; ++k)
        a += (double)x[u][k] * wv[k];
      a = wave_sum(a);
      if (lane == 0)
        red[(c0 + u) * kWaves + wid] = a;
    }
  }
  for (int c = s_main; c < s; ++c) {
    const TV *vc = V + (int64_t)c * ld;
    double a = 0.0;
#pragma unroll
    for (int k = 0; k < kR; ++k)
      a += (double)vc[rr[k]] * wv[k];
    a = wave_sum(a);
    if (lane == 0)
      red[c * kWaves + wid] = a;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < s; c += kBlock) {
    double t = 0.0;
    for (int w2 = 0; w2 < kWaves; ++w2)
      t += red[c * kWaves + w2];
    part[(int64_t)c * nb + blockIdx.x] = t;
  }
}

// OUT(:,c) = V H(:,c)  (overwrite), c < q.
template <class TV, class TO>
__global__ void __launch_bounds__(kBlock)
    k_gemm_nn(int64_t len, const TV *__restrict__ V, int64_t ldv, int s,
              const double *__restrict__ H, int q, TO *__restrict__ out,
              int64_t ldo) {
  const int c = blockIdx.y;
  const double *hc = H + (int64_t)c * s;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock) {
    double acc = 0.0;
    for (int j = 0; j < s; ++j)
      acc += (double)V[i + (int64_t)j * ldv] * hc[j];
    out[i + (int64_t)c * ldo] = (TO)acc;
  }
}

// Batched V^T W: blockIdx.z = column c of W; tiles of CT columns of V.
// part layout: [(c * s + j) * nb + bx].
template <class TV, class TW, int CT>
__global__ void __launch_bounds__(kBlock)
    k_gemm_tn_partial(int64_t len, const TV *__restrict__ V, int64_t ldv, int s,
                      const TW *__restrict__ W, int64_t ldw,
                      double *__restrict__ part, int nb) {
  __shared__ double lds[CT][kWaves];
  const int c0 = blockIdx.y * CT;
  const int cw = blockIdx.z;
  const TW *w = W + (int64_t)cw * ldw;
  double acc[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c)
    acc[c] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock) {
    const double wi = (double)w[i];
#pragma unroll
    for (int c = 0; c < CT; ++c)
      if (c0 + c < s)
        acc[c] += (double)V[i + (int64_t)(c0 + c) * ldv] * wi;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const double v = wave_sum(acc[c]);
    if (lane == 0)
      lds[c][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < CT && c0 + (int)threadIdx.x < s) {
    double t = 0.0;
    for (int w2 = 0; w2 < kWaves; ++w2)
      t += lds[threadIdx.x][w2];
    part[((int64_t)cw * s + c0 + threadIdx.x) * nb + blockIdx.x] = t;
  }
}

// W(:,c) -= V H(:,c) for c < q (blockIdx.y = c); optional partial norms of
// the updated columns, part layout [c * nb + bx].
template <class TV, class TW>
__global__ void __launch_bounds__(kBlock)
    k_gemm_nn_sub(int64_t len, const TV *__restrict__ V, int64_t ldv, int s,
                  const double *__restrict__ H, TW *__restrict__ W, int64_t ldw,
                  double *__restrict__ part, int nb) {
  __shared__ double hs[512];
  __shared__ double lds[kWaves];
  const int c = blockIdx.y;
  const double *hc = H + (int64_t)c * s;
  for (int j = threadIdx.x; j < s && j < 512; j += kBlock)
    hs[j] = hc[j];
  __syncthreads();
  TW *w = W + (int64_t)c * ldw;
  double nacc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock) {
    double acc = 0.0;
    for (int j = 0; j < s; ++j)
      acc += (double)V[i + (int64_t)j * ldv] * (j < 512 ? hs[j] : hc[j]);
    const TW nw = (TW)((double)w[i] - acc);
    w[i] = nw;
    nacc += (double)nw * (double)nw;
  }
  if (part) {
    nacc = block_sum(nacc, lds);
    if (threadIdx.x == 0)
      part[(int64_t)c * nb + blockIdx.x] = nacc;
  }
}

// Column sums of squares of W (q columns), part layout [c * nb + bx].
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_col_sq(int64_t len, const T *__restrict__ W, int64_t ldw,
             double *__restrict__ part, int nb) {
  __shared__ double lds[kWaves];
  const int c = blockIdx.y;
  const T *w = W + (int64_t)c * ldw;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock)
    acc += (double)w[i] * (double)w[i];
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0)
    part[(int64_t)c * nb + blockIdx.x] = acc;
}

// W(:,c) *= scale[c] (device scalars); columns with scale 0 are left as is
// when skip_zero is set.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_scale_cols(int64_t len, T *__restrict__ W, int64_t ldw,
                 const double *__restrict__ scale, int skip_zero) {
  const int c = blockIdx.y;
  const double sc = scale[c];
  if (skip_zero && sc == 0.0)
    return;
  const T st = (T)sc;
  T *w = W + (int64_t)c * ldw;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock)
    w[i] = w[i] * st;
}

// OUT(:,c) = X(:,c) - Y(:,c) for c < q with mask[c] != 0.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_sub_cols(int64_t len, const T *__restrict__ X, int64_t ldx,
               const T *__restrict__ Y, int64_t ldy, T *__restrict__ OUT,
               int64_t ldo, const int *__restrict__ mask) {
#pragma clang fp contract(off)
  const int c = blockIdx.y;
  if (mask && !mask[c])
    return;
  const T *x = X + (int64_t)c * ldx;
  const T *y = Y + (int64_t)c * ldy;
  T *o = OUT + (int64_t)c * ldo;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock)
    o[i] = x[i] + (T)(-1) * y[i];
}

// ------------------------------------------------- tall-skinny GEMM tiles --
// H(s x q) = V^T W over a 64 x 64 output tile (blockIdx.y: rows of H in
// steps of 64, blockIdx.z: columns in steps of 64). Each block reduces a
// contiguous chunk of rows into a private 64x64 partial (LDS-staged 32-row
// slabs, 4x4 register micro-tiles per thread); k_ts_finish sums the partials
// in block order. V and W are read once per tile pair.
constexpr int kTS = 64, kSlab = 32;

template <class TV, class TW>
__global__ void __launch_bounds__(kBlock)
    k_ts_tn_partial(int64_t len, int64_t rows_per_block, const TV *__restrict__ V,
                    int64_t ldv, int s, const TW *__restrict__ W, int64_t ldw, int q,
                    double *__restrict__ part) {
  __shared__ double Vs[kSlab][kTS + 1];
  __shared__ double Ws[kSlab][kTS + 1];
  const int j0 = blockIdx.y * kTS, c0 = blockIdx.z * kTS;
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < len ? r0 + rows_per_block : len;
  double acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
      acc[a][b] = 0.0;
  for (int64_t rb = r0; rb < r1; rb += kSlab) {
    // load the slab: 32 rows x 64 columns of V and W (zero padded)
    // clamped (always in-bounds) loads, zeroed afterwards: no per-load branch
    TV va[kSlab * kTS / kBlock];
    TW wa[kSlab * kTS / kBlock];
#pragma unroll
    for (int t = 0; t < kSlab * kTS / kBlock; ++t) {
      const int e = threadIdx.x + t * kBlock;
      const int r = e % kSlab, j = e / kSlab;
      const int64_t row = (rb + r) < r1 ? (rb + r) : (r1 - 1);
      const int jv = (j0 + j) < s ? (j0 + j) : (s - 1);
      const int jw = (c0 + j) < q ? (c0 + j) : (q - 1);
      va[t] = V[row + (int64_t)jv * ldv];
      wa[t] = W[row + (int64_t)jw * ldw];
    }
#pragma unroll
    for (int t = 0; t < kSlab * kTS / kBlock; ++t) {
      const int e = threadIdx.x + t * kBlock;
      const int r = e % kSlab, j = e / kSlab;
      const bool inr = (rb + r) < r1;
      Vs[r][j] = (inr && j0 + j < s) ? (double)va[t] : 0.0;
      Ws[r][j] = (inr && c0 + j < q) ? (double)wa[t] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int r = 0; r < kSlab; ++r) {
      double a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = Vs[r][tx * 4 + u];
        b[u] = Ws[r][ty * 4 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          acc[u][v] += a[u] * b[v];
    }
    __syncthreads();
  }
  // partial tile layout: [(tile_y, tile_z, block)][64 x 64]
  const int64_t tile = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  double *out = part + tile * (kTS * kTS);
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v)
      out[(ty * 4 + v) * kTS + tx * 4 + u] = acc[u][v];
}

// H(j, c) = sum over blocks of the partial tiles, in block order.
__global__ void __launch_bounds__(kBlock)
    k_ts_finish(const double *__restrict__ part, int nblk, int ty_tiles, int s, int q,
                double *__restrict__ H) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t per_tile = kTS * kTS;
  const int64_t tiles = (int64_t)ty_tiles * ((q + kTS - 1) / kTS);
  if (e >= tiles * per_tile)
    return;
  const int64_t t = e / per_tile, w = e % per_tile;
  const int tyi = (int)(t % ty_tiles), tzi = (int)(t / ty_tiles);
  const int jj = (int)(w % kTS), cc = (int)(w / kTS);
  const int j = tyi * kTS + jj, c = tzi * kTS + cc;
  if (j >= s || c >= q)
    return;
  const double *p = part + ((int64_t)tzi * ty_tiles + tyi) * nblk * per_tile + w;
  double acc = 0.0;
  for (int b = 0; b < nblk; ++b)
    acc += p[(int64_t)b * per_tile];
  H[j + (int64_t)c * s] = acc;
}

// OUT(:, c) (= or -=) V H(:, c) for a tile of kQT output columns
// (blockIdx.y); H staged in LDS, one row per thread, FP64 accumulation.
constexpr int kQT = 32;
template <class TV, class TO, bool SUB>
__global__ void __launch_bounds__(kBlock)
    k_ts_nn(int64_t len, const TV *__restrict__ V, int64_t ldv, int s,
            const double *__restrict__ H, int q, TO *__restrict__ OUT, int64_t ldo) {
  constexpr int JU = 8;  // V columns loaded per step (unconditional, clamped row)
  __shared__ double Hs[64][kQT];
  const int c0 = blockIdx.y * kQT;
  const int qc = (q - c0) < kQT ? (q - c0) : kQT;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i - threadIdx.x < len;
       i += (int64_t)gridDim.x * kBlock) {
    const int64_t ic = i < len ? i : len - 1;
    double acc[kQT];
#pragma unroll
    for (int c = 0; c < kQT; ++c)
      acc[c] = 0.0;
    for (int jb = 0; jb < s; jb += 64) {
      const int js = (s - jb) < 64 ? (s - jb) : 64;
      __syncthreads();
      for (int e = threadIdx.x; e < 64 * kQT; e += kBlock) {
        const int j = e / kQT, c = e % kQT;
        Hs[j][c] = (j < js && c < qc) ? H[(jb + j) + (int64_t)(c0 + c) * s] : 0.0;
      }
      __syncthreads();
      for (int j0 = 0; j0 < js; j0 += JU) {
        TV v[JU];
#pragma unroll
        for (int u = 0; u < JU; ++u) {
          const int j = (j0 + u) < js ? (j0 + u) : (js - 1);  // Hs is 0 past js
          v[u] = V[ic + (int64_t)(jb + j) * ldv];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < JU; ++u) {
          const double vu = (double)v[u];
#pragma unroll
          for (int c = 0; c < kQT; ++c)
            acc[c] += vu * Hs[(j0 + u) & 63][c];
        }
      }
    }
    if (i < len) {
      for (int c = 0; c < qc; ++c) {
        TO *o = OUT + (int64_t)(c0 + c) * ldo;
        o[i] = SUB ? (TO)((double)o[i] - acc[c]) : (TO)acc[c];
      }
    }
  }
}

// ------------------------------------------ shifted-difference stencil --
// A is (n-N) x n with (A x)_i = x_{i+N} - c x_i (examples/quickstart.cpp's
// StdMatrixAdapter lambda), never materialised. Expressions and operand
// order follow the reference's lambdas term by term, no contraction; the
// optional epilogue y = y + (-beta) z is the one of the CSR kernels.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_sd_mv(int64_t m, int64_t N, double c, const T *__restrict__ x, T *__restrict__ y,
            const T *__restrict__ z, const double *__restrict__ beta) {
#pragma clang fp contract(off)
  const T cc = (T)c;
  const T b = z ? (T)(*beta) : (T)0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * kBlock) {
    T out = x[i + N] - cc * x[i];
    if (z)
      out = out + (-b) * z[i];
    y[i] = out;
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_sd_mtv(int64_t n, int64_t N, double c, const T *__restrict__ x, T *__restrict__ y,
             const T *__restrict__ z, const double *__restrict__ beta) {
#pragma clang fp contract(off)
  const T cc = (T)c;
  const T b = z ? (T)(*beta) : (T)0;
  const int64_t m = n - N;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * kBlock) {
    T out = (j >= N ? x[j - N] : (T)0) - (j < m ? cc * x[j] : (T)0);
    if (z)
      out = out + (-b) * z[j];
    y[j] = out;
  }
}

// Blocks: column-major (ldx, ldy, optional active mask) or row-major (stride
// s, rm != 0); thread t covers (row t % rows, column t / rows) column-major,
// (row t / s, column t % s) row-major.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_sd_block(int64_t n, int64_t N, double c, int trans, int s, const T *__restrict__ x,
               int64_t ldx, T *__restrict__ y, int64_t ldy, const int *__restrict__ act, int rm) {
#pragma clang fp contract(off)
  const T cc = (T)c;
  const int64_t m = n - N;
  const int64_t rows = trans ? n : m;
  const int64_t total = rows * (int64_t)s;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * kBlock) {
    int64_t i;
    int col;
    if (rm) {
      i = t / s;
      col = (int)(t - i * s);
    } else {
      col = (int)(t / rows);
      i = t - (int64_t)col * rows;
    }
    if (act && !act[col])
      continue;
    // element (r, col) of x / y
    const int64_t xs = rm ? s : 1, xc = rm ? 1 : ldx, ys = rm ? s : 1, yc = rm ? 1 : ldy;
    const T *xx = x + (int64_t)col * xc;
    T out;
    if (!trans)
      out = xx[(i + N) * xs] - cc * xx[i * xs];
    else
      out = (i >= N ? xx[(i - N) * xs] : (T)0) - (i < m ? cc * xx[i * xs] : (T)0);
    y[i * ys + (int64_t)col * yc] = out;
  }
}

// y = y + (-beta) z with beta on the device (the fused epilogue of the CSR
// kernels, for operators that cannot fuse it).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_sub_beta(int64_t len, const double *__restrict__ beta, const T *__restrict__ z,
               T *__restrict__ y) {
#pragma clang fp contract(off)
  const T b = (T)(*beta);
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = y[i] + (-b) * z[i];
}

// Y(:, c) = X(:, c) for the columns with act[c] != 0.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_copy_active(int64_t len, int s, const T *__restrict__ X, int64_t ldx, T *__restrict__ Y,
                  int64_t ldy, const int *__restrict__ act) {
  const int64_t total = len * (int64_t)s;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * kBlock) {
    const int c = (int)(t / len);
    const int64_t i = t - (int64_t)c * len;
    if (act[c])
      Y[i + (int64_t)c * ldy] = X[i + (int64_t)c * ldx];
  }
}

// --------------------------------------------- row-major block products --
// Y = op(A) X for columns c0 .. c0 + cw - 1 of a block of s columns stored
// row-major (element (r, c) at r * s + c). Thread t handles (row t / cw,
// column c0 + t % cw): the threads of one row read the row's indices and
// values (cache-served broadcasts) and, per nonzero, one contiguous
// cw-element run of a row of X, so every gather is coalesced however large
// the block. Sequential ascending sum, no contraction: the same rounding as
// the one-lane-per-row kernels, whatever the chunking.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_spmm_rows(int64_t rows, const int32_t *__restrict__ ptr, const int32_t *__restrict__ idx,
                const T *__restrict__ val, int s, int c0, int cw, const T *__restrict__ X,
                T *__restrict__ Y) {
  const int64_t total = rows * (int64_t)cw;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * kBlock) {
    const int64_t i = t / cw;
    const int c = c0 + (int)(t - i * cw);
    const T *xc = X + c;
    Y[i * s + c] = (T)row_dot<T>(ptr[i], ptr[i + 1], idx, val,
                                 [=](int32_t j) { return xc[(int64_t)j * s]; });
  }
}

// Column-major (ld) <-> row-major (stride s) block transposes through LDS.
// A block owns 64 rows and a group of up to 64 columns (blockIdx.y): on the
// row-major side its rows form one contiguous run when s <= 64 (runs of 64
// elements otherwise), read or written flat; on the column-major side each
// column's 64 rows are one run. Tile stride 65 keeps the column-wise LDS
// accesses off a single bank.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_cols_to_rows(int64_t len, int s, const T *__restrict__ X, int64_t ldx, T *__restrict__ R) {
  __shared__ T tile[64 * 65];
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int c0 = blockIdx.y * 64;
  const int cw = s - c0 < 64 ? s - c0 : 64;
  const int rows = len - r0 < 64 ? (int)(len - r0) : 64;
  for (int e = threadIdx.x; e < cw * 64; e += kBlock) {
    const int c = e >> 6, r = e & 63;
    if (r < rows)
      tile[r * 65 + c] = X[r0 + r + (int64_t)(c0 + c) * ldx];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < rows * cw; e += kBlock) {
    const int r = e / cw, c = e - r * cw;
    R[(r0 + r) * s + c0 + c] = tile[r * 65 + c];
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_rows_to_cols(int64_t len, int s, const T *__restrict__ R, T *__restrict__ X, int64_t ldx) {
  __shared__ T tile[64 * 65];
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int c0 = blockIdx.y * 64;
  const int cw = s - c0 < 64 ? s - c0 : 64;
  const int rows = len - r0 < 64 ? (int)(len - r0) : 64;
  for (int e = threadIdx.x; e < rows * cw; e += kBlock) {
    const int r = e / cw, c = e - r * cw;
    tile[r * 65 + c] = R[(r0 + r) * s + c0 + c];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < cw * 64; e += kBlock) {
    const int c = e >> 6, r = e & 63;
    if (r < rows)
      X[r0 + r + (int64_t)(c0 + c) * ldx] = tile[r * 65 + c];
  }
}

// ----------------------------------------------------- FP32 MFMA Gram tiles --
// H(s x q) = V^T W for FP32 V, W on the matrix cores: v_mfma_f32_16x16x4f32.
// A block owns a contiguous chunk of rows and one 64 x 64 tile of H
// (blockIdx.y, blockIdx.z); 64-row slabs of V and W are staged in LDS with
// coalesced column reads; wave w accumulates H rows [16w, 16w+16) against the
// four 16-column tiles. MFMA lane layout (16x16x4): A[i][k] from lane
// i + 16k, B[k][j] from lane j + 16k, D[4*(l/16)+v][l%16] in element v of
// lane l. Block partials (FP32) are summed in FP64 in block order.
typedef float floatx4 __attribute__((ext_vector_type(4)));
constexpr int kMT = 64, kMSlab = 64;

__global__ void __launch_bounds__(kBlock)
    k_gram_mfma_partial(int64_t len, int64_t rows_per_block, const float *__restrict__ V,
                        int64_t ldv, int s, const float *__restrict__ W, int64_t ldw, int q,
                        float *__restrict__ part) {
  __shared__ float Vs[kMSlab][kMT + 4];
  __shared__ float Ws[kMSlab][kMT + 4];
  const int j0 = blockIdx.y * kMT, c0 = blockIdx.z * kMT;
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < len ? r0 + rows_per_block : len;
  floatx4 acc[4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
    acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  constexpr int PER = kMSlab * kMT / kBlock;  // 16 elements per thread per slab
  for (int64_t rb = r0; rb < r1; rb += kMSlab) {
    float va[PER], wa[PER];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * kBlock;
      const int r = e % kMSlab, j = e / kMSlab;
      const int64_t row = (rb + r) < r1 ? (rb + r) : (r1 - 1);
      const int jv = (j0 + j) < s ? (j0 + j) : (s - 1);
      const int jw = (c0 + j) < q ? (c0 + j) : (q - 1);
      va[u] = V[row + (int64_t)jv * ldv];
      wa[u] = W[row + (int64_t)jw * ldw];
    }
    __syncthreads();  // previous slab fully consumed
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int e = threadIdx.x + u * kBlock;
      const int r = e % kMSlab, j = e / kMSlab;
      const bool inr = (rb + r) < r1;
      Vs[r][j] = (inr && j0 + j < s) ? va[u] : 0.f;
      Ws[r][j] = (inr && c0 + j < q) ? wa[u] : 0.f;
    }
    __syncthreads();
#pragma unroll 4
    for (int k4 = 0; k4 < kMSlab / 4; ++k4) {
      const int kr = k4 * 4 + lane / 16;
      const float a = Vs[kr][16 * wid + (lane & 15)];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const float b = Ws[kr][16 * t + (lane & 15)];
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[t], 0, 0, 0);
      }
    }
  }
  const int64_t tile = ((int64_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
  float *out = part + tile * (kMT * kMT);
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 16 * wid + 4 * (lane / 16) + v;  // H row (V column)
      const int col = 16 * t + (lane & 15);            // H column (W column)
      out[col * kMT + row] = acc[t][v];
    }
}

__global__ void __launch_bounds__(kBlock)
    k_gram_finish(const float *__restrict__ part, int nblk, int ty_tiles, int s, int q,
                  double *__restrict__ H) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int64_t per_tile = kMT * kMT;
  const int64_t tiles = (int64_t)ty_tiles * ((q + kMT - 1) / kMT);
  if (e >= tiles * per_tile)
    return;
  const int64_t t = e / per_tile, w = e % per_tile;
  const int tyi = (int)(t % ty_tiles), tzi = (int)(t / ty_tiles);
  const int jj = (int)(w % kMT), cc = (int)(w / kMT);
  const int j = tyi * kMT + jj, c = tzi * kMT + cc;
  if (j >= s || c >= q)
    return;
  const float *p = part + ((int64_t)tzi * ty_tiles + tyi) * nblk * per_tile + w;
  double acc = 0.0;
  for (int b = 0; b < nblk; ++b)
    acc += (double)p[(int64_t)b * per_tile];
  H[j + (int64_t)c * s] = acc;
}

// ------------------------------------------- fused FP32 MFMA Gram pair --
// S = Z^T Z (upper-triangle tiles only) and G = Z^T AZ in one pass over the
// rows, q <= kG2MaxQ. A block of 8 waves owns a contiguous chunk of rows and
// every output tile of both products, so Z and AZ are read from HBM exactly
// once. 32-row slabs of Z and AZ are staged column-major in LDS with a row
// stride of 34 floats: the 32 lanes of a column write 32 consecutive banks,
// and the MFMA operand reads (16 columns x 2 rows per half-wave) hit 32
// distinct banks. The next slab is loaded into registers while the current
// one is consumed. Tile list: S tiles (ti <= tj) then G tiles, dealt to the
// waves round-robin. Block partials (FP32, at most 4096 rows each) are summed
// in FP64 in block order by k_gram2_finish.
constexpr int kG2Threads = 512, kG2Slab = 32, kG2MaxQ = 128;
constexpr int kG2Waves = kG2Threads / 64;

__device__ inline void g2_tile(int t, int T, int &ti, int &tj, bool &is_s) {
  const int ns = T * (T + 1) / 2;
  if (t < ns) {
    int i = 0, rem = t;
    while (rem >= T - i) {
      rem -= T - i;
      ++i;
    }
    ti = i;
    tj = i + rem;
    is_s = true;
  } else {
    ti = (t - ns) / T;
    tj = (t - ns) % T;
    is_s = false;
  }
}

// 16x16x4 MFMA in the element type: FP32 (v_mfma_f32_16x16x4f32) or FP64
// (v_mfma_f64_16x16x4f64). Both take A[i][k] from lane i + 16k and B[k][j]
// from lane j + 16k. Element v of lane l holds D[row(l, v)][l % 16] with
// row = 4*(l/16) + v for FP32 and row = l/16 + 4*v for FP64.
typedef double doublex4 __attribute__((ext_vector_type(4)));
template <class T> struct Mfma16;
template <> struct Mfma16<float> {
  using V = floatx4;
  static __device__ inline V step(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct Mfma16<double> {
  using V = doublex4;
  static __device__ inline V step(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

// mode 0: the Gram pair (s == q). mode 1: H = Z^T AZ for Z with s columns
// and AZ with q columns (s, q <= 128), all (s/16) x (q/16) tiles, A operand
// from Z and B operand from AZ (the tall-skinny V^T W of the solvers).
__device__ inline int g2_ntiles(int mode, int s, int q) {
  const int ns = (s + 15) / 16, nq = (q + 15) / 16;
  return mode == 0 ? nq * (nq + 1) / 2 + nq * nq : ns * nq;
}

// SLAB rows per LDS slab (a multiple of 4; SLAB * MAXQ a multiple of 512),
// at most MAXQ columns per operand (a multiple of 16). The Gram pair uses 32 x
// 128; the tall-skinny products of at most 64 columns use 96-row (FP32) or
// 48-row (FP64) slabs, so every column is read 384 bytes at a time instead of
// 128 (the short runs held those products near 1 TB/s).
//
// Z and AZ may be of a narrower type than the compute type T (FP32 input on
// FP64 MFMA: every product exact, the sums in FP64). Each operand's slab is
// staged in LDS in its own type (TLZ, TLA: FP32 operands as FP32, which
// halves their LDS and so allows longer column runs) and converted to T as
// the MFMA steps read it.
template <class T, int SLAB = kG2Slab, int MAXQ = kG2MaxQ, class TZ = T, class TA = T,
          class TLZ = T, class TLA = T>
__global__ void __launch_bounds__(kG2Threads)
    k_gram2_partial(int64_t len, int64_t rows_per_block, const TZ *__restrict__ Z,
                    int64_t ldz, int s, const TA *__restrict__ AZ, int64_t ldaz, int q,
                    int mode, T *__restrict__ part) {
  using V4 = typename Mfma16<T>::V;
  constexpr int LD = SLAB + 2;
  constexpr int MAXT = MAXQ / 16;
  // tiles per wave: the full-width kernel also serves the Gram pair (mode 0,
  // triangle of S plus all of G); the narrower ones only V^T W (mode 1)
  constexpr int MAX_TILES =
      MAXQ == kG2MaxQ ? MAXT * (MAXT + 1) / 2 + MAXT * MAXT : MAXT * MAXT;
  constexpr int PER_WAVE = (MAX_TILES + kG2Waves - 1) / kG2Waves;
  static_assert(SLAB % 4 == 0 && (SLAB * MAXQ) % kG2Threads == 0 && MAXQ % 16 == 0,
                "slab shape");
  static_assert(sizeof(TZ) <= sizeof(TLZ) && sizeof(TA) <= sizeof(TLA), "exact staging");
  __shared__ TLZ Zs[MAXQ * LD];
  __shared__ TLA As[MAXQ * LD];
  const int nt = (q + 15) / 16;
  const int SP = ((s + 15) / 16) * 16, QP = nt * 16;
  const int ntiles = g2_ntiles(mode, s, q);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < len ? r0 + rows_per_block : len;
  // this wave's tiles: operand offsets in LDS (floats)
  int aoff[PER_WAVE], boff[PER_WAVE];
  bool bs[PER_WAVE];
#pragma unroll
  for (int u = 0; u < PER_WAVE; ++u) {
    const int t = wid + u * kG2Waves;
    int ti = 0, tj = 0;
    bool is_s = true;
    if (t < ntiles) {
      if (mode == 0) {
        g2_tile(t, nt, ti, tj, is_s);
      } else {
        ti = t / nt;
        tj = t % nt;
        is_s = false;
      }
    }
    aoff[u] = (16 * ti + (lane & 15)) * LD + (lane >> 4);
    boff[u] = (16 * tj + (lane & 15)) * LD + (lane >> 4);
    bs[u] = is_s;
  }
  V4 acc[PER_WAVE];
#pragma unroll
  for (int u = 0; u < PER_WAVE; ++u)
    acc[u] = V4{(T)0, (T)0, (T)0, (T)0};
  // staging: element e = tid + u*512 of the (SLAB rows x MAXQ cols) slab of Z
  // and of AZ: row e % SLAB, column e / SLAB
  constexpr int kLd = SLAB * MAXQ / kG2Threads;  // loads per operand and thread
  TLZ za[kLd];
  TLA aa[kLd];
  auto load = [&](int64_t rb) {
#pragma unroll
    for (int u = 0; u < kLd; ++u) {
      const int e = threadIdx.x + u * kG2Threads;
      const int r = e % SLAB, c = e / SLAB;
      const int64_t row = (rb + r) < r1 ? (rb + r) : (r1 - 1);
      za[u] = c < s ? (TLZ)Z[row + (int64_t)c * ldz] : (TLZ)0;
      aa[u] = c < q ? (TLA)AZ[row + (int64_t)c * ldaz] : (TLA)0;
    }
  };
  if (r0 < r1)
    load(r0);
  for (int64_t rb = r0; rb < r1; rb += SLAB) {
    __syncthreads();  // previous slab consumed
#pragma unroll
    for (int u = 0; u < kLd; ++u) {
      const int e = threadIdx.x + u * kG2Threads;
      const int r = e % SLAB, c = e / SLAB;
      const bool inr = (rb + r) < r1;
      if (c < SP)
        Zs[c * LD + r] = (inr && c < s) ? za[u] : (TLZ)0;
      if (c < QP)
        As[c * LD + r] = (inr && c < q) ? aa[u] : (TLA)0;
    }
    __syncthreads();
    if (rb + SLAB < r1)
      load(rb + SLAB);
#pragma unroll 8
    for (int k = 0; k < SLAB; k += 4) {
#pragma unroll
      for (int u = 0; u < PER_WAVE; ++u) {
        if (wid + u * kG2Waves < ntiles) {
          const T a = (T)Zs[aoff[u] + k];
          const T b = bs[u] ? (T)Zs[boff[u] + k] : (T)As[boff[u] + k];
          acc[u] = Mfma16<T>::step(a, b, acc[u]);
        }
      }
    }
  }
  T *out = part + (int64_t)blockIdx.x * ntiles * 256;
#pragma unroll
  for (int u = 0; u < PER_WAVE; ++u) {
    const int t = wid + u * kG2Waves;
    if (t < ntiles)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        out[t * 256 + Mfma16<T>::row(lane, v) * 16 + (lane & 15)] = acc[u][v];
  }
}

// H = V^T W for narrow blocks (s, q <= 16: one 16 x 16 MFMA tile) of FP64 or
// FP32 (converted exactly) input, the window-against-basis products of the
// cyclic solver at small k, on FP64 MFMA. A block of
// 8 waves stages 256-row slabs of both operands in LDS (16 coalesced loads
// in flight per thread, the next slab loading while this one is consumed);
// wave w runs the MFMA steps of slab rows [32w, 32w + 32), and the 8 wave
// tiles are summed in wave order at the end. Output layout as
// k_gram2_partial with one tile (256 values per block), for k_tn2_finish.
constexpr int kTnSlab = 256, kTnLd = kTnSlab + 1;

template <class T, class TW = T>
__global__ void __launch_bounds__(kG2Threads)
    k_tn_narrow(int64_t len, int64_t rows_per_block, const T *__restrict__ V, int64_t ldv,
                int s, const TW *__restrict__ W, int64_t ldw, int q,
                double *__restrict__ part) {
  __shared__ double Vs[16 * kTnLd];
  __shared__ double Ws[16 * kTnLd];
  __shared__ double red[kG2Waves * 256];
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = r0 + rows_per_block < len ? r0 + rows_per_block : len;
  // element u of this thread: column 2u + (tid >= 256), row tid % 256
  const int rl = threadIdx.x & (kTnSlab - 1), chi = threadIdx.x >> 8;
  // raw values stay in registers (in flight during the MFMA steps of the
  // current slab) and are converted on the LDS store
  T va[8];
  TW wa[8];
  auto load = [&](int64_t rb) {
    const int64_t row = (rb + rl) < r1 ? (rb + rl) : (r1 - 1);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int c = 2 * u + chi;
      va[u] = c < s ? V[row + (int64_t)c * ldv] : (T)0;
      wa[u] = c < q ? W[row + (int64_t)c * ldw] : (TW)0;
    }
  };
  doublex4 acc = doublex4{0.0, 0.0, 0.0, 0.0};
  const int ai = (lane & 15) * kTnLd + (lane >> 4) + 32 * wid;
  if (r0 < r1)
    load(r0);
  for (int64_t rb = r0; rb < r1; rb += kTnSlab) {
    __syncthreads();  // previous slab consumed
    const bool inr = (rb + rl) < r1;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int c = 2 * u + chi;
      Vs[c * kTnLd + rl] = inr ? (double)va[u] : 0.0;
      Ws[c * kTnLd + rl] = inr ? (double)wa[u] : 0.0;
    }
    __syncthreads();
    if (rb + kTnSlab < r1)
      load(rb + kTnSlab);
#pragma unroll
    for (int k = 0; k < 32; k += 4)
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(Vs[ai + k], Ws[ai + k], acc, 0, 0, 0);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v)
    red[wid * 256 + Mfma16<double>::row(lane, v) * 16 + (lane & 15)] = acc[v];
  __syncthreads();
  if (threadIdx.x < 256) {
    double t = 0.0;
    for (int w = 0; w < kG2Waves; ++w)
      t += red[w * 256 + threadIdx.x];
    part[(int64_t)blockIdx.x * 256 + threadIdx.x] = t;
  }
}

// SG[0 : q*q] = S, SG[q*q : 2*q*q] = G, both column-major q x q; S is
// mirrored from its upper triangle. One wave per output, as k_tn2_finish.
// mode 1: H (s x q, column-major) from the VᵀW tiles. One wave per output:
// lane l sums the partials of blocks l, l+64, ... in order, then a fixed
// shuffle tree (deterministic for a given block count).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_tn2_finish(const T *__restrict__ part, int nblk, int s, int q, double *__restrict__ H) {
  const int64_t e = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (e >= (int64_t)s * q)
    return;
  const int i = (int)(e % s), j = (int)(e / s);
  const int nq = (q + 15) / 16;
  const int ntiles = ((s + 15) / 16) * nq;
  const int t = (i / 16) * nq + j / 16;
  const T *p = part + (int64_t)t * 256 + (i % 16) * 16 + (j % 16);
  double acc = 0.0;
  for (int b = lane; b < nblk; b += 64)
    acc += (double)p[(int64_t)b * ntiles * 256];
  acc = wave_sum(acc);
  if (lane == 0)
    H[e] = acc;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_gram2_finish(const T *__restrict__ part, int nblk, int q, double *__restrict__ SG) {
  const int64_t e = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int64_t qq = (int64_t)q * q;
  if (e >= 2 * qq)
    return;
  const bool is_g = e >= qq;
  const int64_t w = is_g ? e - qq : e;
  int i = (int)(w % q), j = (int)(w / q);  // H(i, j)
  if (!is_g && i > j) {
    const int t = i;
    i = j;
    j = t;
  }
  const int nt = (q + 15) / 16;
  const int ntiles = nt * (nt + 1) / 2 + nt * nt;
  const int ti = i / 16, tj = j / 16;
  int t;
  if (!is_g)
    t = ti * nt - ti * (ti - 1) / 2 + (tj - ti);
  else
    t = nt * (nt + 1) / 2 + ti * nt + tj;
  const T *p = part + (int64_t)t * 256 + (i % 16) * 16 + (j % 16);
  double acc = 0.0;
  for (int b = lane; b < nblk; b += 64)
    acc += (double)p[(int64_t)b * ntiles * 256];
  acc = wave_sum(acc);
  if (lane == 0)
    SG[e] = acc;
}

// OUT(:, c) (= or -=) V H(:, c), c < q <= 16 * NQ, on FP64 MFMA: computed as
// OUT^T = H^T V^T so that element v of lane l lands in row l % 16 of a
// 16-row group, i.e. 16 consecutive rows per column on the store. A wave owns
// G groups of 16 rows per chunk and all ceil(q/16) column tiles; H^T sits in
// LDS (column j of the output at j * sld), V is read once (FP32 V converted
// exactly), FP64 accumulation. With HOIST the old OUT values of the
// subtracting form are loaded with the chunk's V, so a chunk waits for memory
// once (narrow outputs; wide ones would need too many registers). Blocks
// stride over 64G-row chunks.
template <class TV, class TO, bool SUB, int G, int NQ, bool HOIST>
__global__ void __launch_bounds__(kBlock)
    k_nn_mfma(int64_t len, const TV *__restrict__ V, int64_t ldv, int s,
              const double *__restrict__ H, int q, TO *__restrict__ OUT, int64_t ldo, int sld) {
  extern __shared__ double Hs[];  // [qp][sld]
  const int nq = (q + 15) / 16, qp = nq * 16;
  for (int e = threadIdx.x; e < qp * sld; e += kBlock) {
    const int j = e / sld, k = e - j * sld;
    Hs[e] = (j < q && k < s) ? H[k + (int64_t)j * s] : 0.0;
  }
  __syncthreads();
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  const int s4 = (s + 3) & ~3;
  for (int64_t c0 = (int64_t)blockIdx.x * 64 * G; c0 < len;
       c0 += (int64_t)gridDim.x * 64 * G) {
    int64_t ig[G], rc[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      ig[g] = c0 + 64 * g + 16 * wid + li;
      rc[g] = ig[g] < len ? ig[g] : (len - 1);
    }
    double old[G][NQ][4];
    if (SUB && HOIST) {
#pragma unroll
      for (int g = 0; g < G; ++g)
#pragma unroll
        for (int t = 0; t < NQ; ++t)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int j = 16 * t + Mfma16<double>::row(lane, v);
            old[g][t][v] = (t < nq && j < q && ig[g] < len)
                               ? (double)OUT[ig[g] + (int64_t)j * ldo] : 0.0;
          }
    }
    doublex4 acc[G][NQ];
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
      for (int t = 0; t < NQ; ++t)
        acc[g][t] = doublex4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < s4; k0 += 16) {
      double b[G][4];
#pragma unroll
      for (int g = 0; g < G; ++g)
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int k = k0 + 4 * u + lk;
          const int kc = k < s ? k : s - 1;
          const double x = (double)V[rc[g] + (int64_t)kc * ldv];
          b[g][u] = k < s ? x : 0.0;
        }
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kb = k0 + 4 * u;
        if (kb < s4) {
#pragma unroll
          for (int t = 0; t < NQ; ++t)
            if (t < nq) {
              const double a = Hs[(16 * t + li) * sld + kb + lk];
#pragma unroll
              for (int g = 0; g < G; ++g)
                acc[g][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[g][u], acc[g][t], 0, 0, 0);
            }
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int64_t i = ig[g];
      if (i < len) {
#pragma unroll
        for (int t = 0; t < NQ; ++t)
          if (t < nq) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int j = 16 * t + Mfma16<double>::row(lane, v);
              if (j < q) {
                TO *o = OUT + i + (int64_t)j * ldo;
                const double ov = HOIST ? old[g][t][v] : (SUB ? (double)*o : 0.0);
                *o = SUB ? (TO)(ov - acc[g][t][v]) : (TO)acc[g][t][v];
              }
            }
          }
      }
    }
  }
}

// ---------------------------------------------------------------- level 1 --
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_axpby(int64_t len, double alpha, const T *__restrict__ x, double beta,
            T *__restrict__ y) {
#pragma clang fp contract(off)
  const T a = (T)alpha, b = (T)beta;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock) {
    T v = (b == (T)0) ? (T)0 : y[i];
    if (b != (T)1 && b != (T)0)
      v = v * b;
    if (x)
      v = v + a * x[i];
    y[i] = v;
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_scale_dev(int64_t len, const double *__restrict__ num, int recip,
                T *__restrict__ y) {
  const double s = recip ? 1.0 / (*num) : (*num);
  const T st = (T)s;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = y[i] * st;
}

// y = x * (1 / *den): the vnew = w / beta store of a Golub-Kahan step with
// beta on the device.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_scale_copy_dev(int64_t len, const double *__restrict__ den, const T *__restrict__ x,
                     T *__restrict__ y) {
  const T s = (T)(1.0 / (*den));
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = x[i] * s;
}

// ||w|| from the block partials of k_vh_sub (every block forms the same sum
// as k_finish does, in the same order) and y = x * (1 / ||w||) as
// k_scale_copy_dev computes it; block 0 stores the norm. One launch for the
// norm and the normalised copy of a Golub-Kahan step.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_norm_scale_copy(const double *__restrict__ part, int nb, double *__restrict__ nrm,
                      int64_t len, const T *__restrict__ x, T *__restrict__ y) {
  __shared__ double lds[kWaves];
  __shared__ double beta;
  double acc = 0.0;
  int i = threadIdx.x;
  for (; i + 7 * kBlock < nb; i += 8 * kBlock) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      v[u] = part[i + u * kBlock];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      acc += v[u];
  }
  for (; i < nb; i += kBlock)
    acc += part[i];
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) {
    const double b = sqrt(fmax(0.0, acc));
    beta = b;
    if (blockIdx.x == 0)
      *nrm = b;
  }
  __syncthreads();
  const T sc = (T)(1.0 / beta);
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < len;
       j += (int64_t)gridDim.x * kBlock)
    y[j] = x[j] * sc;
}

// R(:, j) = AX(:, j) + (T)(neg_theta[j]) X(:, j) for j < q: the LOBPCG
// residual block in one launch, per element the rounding of a copy of AX
// followed by axpby(-theta_j, X, 1) (product rounded, then the sum).
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_resid_block(int64_t len, const T *__restrict__ AX, const T *__restrict__ X,
                  const double *__restrict__ neg_theta, T *__restrict__ R, int64_t ld) {
#pragma clang fp contract(off)
  const int64_t off = (int64_t)blockIdx.y * ld;
  const T a = (T)neg_theta[blockIdx.y];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock) {
    const T prod = a * X[off + i];
    R[off + i] = AX[off + i] + prod;
  }
}

template <class TI, class TO>
__global__ void __launch_bounds__(kBlock)
    k_convert(int64_t len, const TI *__restrict__ x, TO *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock)
    y[i] = (TO)x[i];
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_diag_solve(int64_t len, const double *__restrict__ d, const T *__restrict__ r,
                 T *__restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len;
       i += (int64_t)gridDim.x * kBlock) {
    const double di = d[i];
    y[i] = di > 0.0 ? (T)((double)r[i] / di) : r[i];
  }
}

// ---------------------------------------------------------------------- CG --
// Per-column recurrence of cg_normal_eq_prec
// (CyclicLanczosSolver.cpp:48-113), one column per blockIdx.y.
enum : int32_t { kRun = 0, kDone = 1 };

// r = rhs, w = 0, z = M^{-1} r (or alias), p = z; partials of r.r and r.z.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_cg_init(int64_t m, const T *__restrict__ rhs, T *__restrict__ w,
              T *__restrict__ r, T *__restrict__ z, T *__restrict__ p,
              const double *__restrict__ d, double *__restrict__ part_rr,
              double *__restrict__ part_rz, int nb) {
  __shared__ double lds[kWaves];
  const int c = blockIdx.y;
  const int64_t o = (int64_t)c * m;
  double arr = 0.0, arz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * kBlock) {
    const T ri = rhs[o + i];
    r[o + i] = ri;
    w[o + i] = (T)0;
    T zi = ri;
    if (d) {
      const double di = d[i];
      zi = di > 0.0 ? (T)((double)ri / di) : ri;
      z[o + i] = zi;
    }
    p[o + i] = zi;
    arr += (double)ri * (double)ri;
    arz += (double)ri * (double)zi;
  }
  arr = block_sum(arr, lds);
  arz = block_sum(arz, lds);
  if (threadIdx.x == 0) {
    part_rr[(int64_t)c * nb + blockIdx.x] = arr;
    part_rz[(int64_t)c * nb + blockIdx.x] = arz;
  }
}

__global__ void k_cg_init_finish(int s, const double *__restrict__ part_rr,
                                 const double *__restrict__ part_rz, int nb,
                                 CgState *__restrict__ st, int *__restrict__ act,
                                 double rel_tol, int64_t maxit) {
  __shared__ double lds[kWaves];
  const int c = blockIdx.x;
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock) {
    a += part_rr[(int64_t)c * nb + i];
    b += part_rz[(int64_t)c * nb + i];
  }
  a = block_sum(a, lds);
  b = block_sum(b, lds);
  if (threadIdx.x == 0) {
    CgState S{};
    S.k = 0;
    S.done = kRun;
    S.reason = 0;
    const double rhsn = sqrt(fmax(0.0, a));
    S.rhsn = rhsn;
    S.rnorm = rhsn;
    if (!(rhsn > 0.0) || !isfinite(rhsn)) {
      S.done = kDone;
      S.reason = 1;
    } else {
      const double rz = b;
      if (!(rz > 0.0) || !isfinite(rz)) {
        S.done = kDone;
        S.reason = 1;
      } else {
        S.rz = rz;
        S.rnorm = sqrt(rz);
        S.best = S.rnorm;
        S.kbest = 0;
        S.tol = fmax(rel_tol * S.rnorm, 1e-30);
        if (S.rnorm <= S.tol) {
          S.done = kDone;
          S.reason = 0;
        } else if (maxit <= 0) {
          S.done = kDone;
          S.reason = 4;
        }
      }
    }
    st[c] = S;
    act[c] = S.done == kRun;
  }
}

// pAp partials per column: p . Ap
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_cg_pap(int64_t m, const T *__restrict__ p, const T *__restrict__ Ap,
             const int *__restrict__ act, double *__restrict__ part, int nb) {
  __shared__ double lds[kWaves];
  const int c = blockIdx.y;
  if (!act[c])
    return;
  const int64_t o = (int64_t)c * m;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * kBlock)
    acc += (double)p[o + i] * (double)Ap[o + i];
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0)
    part[(int64_t)c * nb + blockIdx.x] = acc;
}

// pAp = the sum of the column's p.Ap partials (every block forms it, in
// the order of a per-column sum over blocks); a breakdown (non-finite or
// |pAp| <= breakdown_tol) stops the column, otherwise w += alpha p;
// r -= alpha Ap; z = M^{-1} r and the partials of r.z (alpha = rz / pAp).
// Block 0 records pAp and the stop.
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_cg_update(int64_t m, T *__restrict__ w, T *__restrict__ r,
                T *__restrict__ z, const T *__restrict__ p,
                const T *__restrict__ Ap, const double *__restrict__ d,
                CgState *__restrict__ st, int *__restrict__ act,
                const double *__restrict__ pap_part, double *__restrict__ part, int nb,
                double breakdown_tol) {
#pragma clang fp contract(off)
  __shared__ double lds[kWaves];
  __shared__ int on;
  __shared__ double pap;
  const int c = blockIdx.y;
  if (threadIdx.x == 0)
    on = act[c];  // one read per block: block 0 may clear it meanwhile
  __syncthreads();
  if (!on)
    return;
  const double sum = sum_partials(pap_part + (int64_t)c * nb, nb, lds);
  if (threadIdx.x == 0)
    pap = sum;
  __syncthreads();
  const bool brk = !isfinite(pap) || fabs(pap) <= breakdown_tol;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st[c].pAp = pap;
    if (brk) {
      st[c].done = kDone;
      st[c].reason = 2;
      act[c] = 0;
    }
  }
  if (brk)
    return;
  const double alpha = st[c].rz / pap;
  const T a = (T)alpha, na = (T)(-alpha);
  const int64_t o = (int64_t)c * m;
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * kBlock) {
    w[o + i] = w[o + i] + a * p[o + i];
    const T ri = r[o + i] + na * Ap[o + i];
    r[o + i] = ri;
    T zi = ri;
    if (d) {
      const double di = d[i];
      zi = di > 0.0 ? (T)((double)ri / di) : ri;
      z[o + i] = zi;
    }
    acc += (double)ri * (double)zi;
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0)
    part[(int64_t)c * nb + blockIdx.x] = acc;
}

__global__ void k_cg_beta(const double *__restrict__ part, int nb,
                          CgState *__restrict__ st, int *__restrict__ act,
                          int64_t maxit, int64_t stall) {
  __shared__ double lds[kWaves];
  const int c = blockIdx.x;
  if (!act[c])
    return;
  double a = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlock)
    a += part[(int64_t)c * nb + i];
  a = block_sum(a, lds);
  if (threadIdx.x == 0) {
    CgState S = st[c];
    S.rz_new = a;
    if (!isfinite(a)) {
      S.done = kDone;
      S.reason = 3;
    } else {
      S.rnorm = sqrt(fmax(0.0, a));
      if (S.rnorm <= S.tol) {
        S.k += 1;
        S.done = kDone;
        S.reason = 0;
      } else {
        S.pAp = a / fmax(1e-300, S.rz);  // beta, consumed by k_cg_dir
        S.rz = a;
        S.k += 1;
        if (S.rnorm < 0.5 * S.best) {
          S.best = S.rnorm;
          S.kbest = S.k;
        }
        if (S.k >= maxit) {
          S.done = kDone;
          S.reason = 4;
        } else if (stall > 0 && S.k - S.kbest >= stall) {
          S.done = kDone;
          S.reason = 5;
        }
      }
    }
    // p must still be updated for a column that stopped at maxit? No: the
    // reference leaves the loop, so p is dead; act gates k_cg_dir.
    st[c] = S;
    act[c] = (S.done == kRun) ? 1 : ((S.reason == 4 || S.reason == 5) ? 2 : 0);
  }
}

// p = beta p + z (columns still running; act==2 marks a maxit or stall stop
// after the update of w and r, where p is no longer used).
// (act == 2 is cleared to 0 here: the column's p is dead.)
template <class T>
__global__ void __launch_bounds__(kBlock)
    k_cg_dir(int64_t m, T *__restrict__ p, const T *__restrict__ z,
             const CgState *__restrict__ st, int *__restrict__ act) {
#pragma clang fp contract(off)
  const int c = blockIdx.y;
  const int a = act[c];
  if (a != 1) {
    if (a == 2 && blockIdx.x == 0 && threadIdx.x == 0)
      act[c] = 0;
    return;
  }
  const T b = (T)st[c].pAp;
  const int64_t o = (int64_t)c * m;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < m;
       i += (int64_t)gridDim.x * kBlock)
    p[o + i] = p[o + i] * b + z[o + i];
}


}  // namespace

// =============================================================== launchers ==
// Records an event pair around the launches of a traced class.
// Per-launch timing of a traced class. By default an event is recorded on
// the stream before and after the launch (the pair then also spans the
// dispatch of the launch); with `stamped` the scope only provides the events
// and the launch passes them to hipExtLaunchKernelGGL, which stamps the
// kernel's own start and end (the duration a kernel-trace profiler reports).
struct TraceScope {
  Context &c;
  bool on, stamped;
  hipEvent_t a = nullptr, b = nullptr;
  double bytes;
  int cls;
  int64_t len = 0;
  int s = 0;
  TraceScope(Context &ctx, int cls_, double by, int64_t len_ = 0, int s_ = 0,
             bool stamped_ = false)
      : c(ctx), on(cls_ > 0 && ((ctx.trace_mask >> cls_) & 1u)), stamped(stamped_), bytes(by),
        cls(cls_), len(len_), s(s_) {
    if (on) {
      a = c.take_event();
      b = c.take_event();
      if (!stamped)
        KH_CHECK(hipEventRecord(a, c.stream));
    }
  }
  hipEvent_t start() const { return on ? a : nullptr; }
  hipEvent_t stop() const { return on ? b : nullptr; }
  ~TraceScope() {
    if (on) {
      if (!stamped)
        (void)hipEventRecord(b, c.stream);
      c.trace.push_back({bytes, a, b, cls, len, s});
    }
  }
};

int reduce_blocks(int64_t len) {
  int64_t g = (len + 2047) / 2048;
  if (g < 1)
    g = 1;
  if (g > 1024)
    g = 1024;
  return (int)g;
}

// Entries per row when every row has the same count R and R is one the
// fixed-length kernels take (1..4 or 8), else 0.
static int ell_width(int uniform, const int32_t *idx, const void *val) {
  if (!(uniform >= 1 && uniform <= 4) && uniform != 8)
    return 0;
  if (reinterpret_cast<uintptr_t>(idx) % 16 != 0 || reinterpret_cast<uintptr_t>(val) % 16 != 0)
    return 0;
  return uniform;
}

#define KH_ELL_LAUNCH(KERNEL, R, STREAM, ...)                                                \
  hipLaunchKernelGGL((KERNEL<T, R>), dim3(grid_rows(rows)), dim3(kBlock), 0, STREAM, __VA_ARGS__)

// act (fixed-length rows only): one column's device flag, nothing written when 0
template <class T>
static void launch_spmv(Context &ctx, int64_t rows, const int32_t *ptr,
                        const int32_t *idx, const T *val, int lpr, int ell, const T *x,
                        T *y, const T *z, const double *beta, const int *act = nullptr) {
  hipStream_t s = ctx.stream;
  if (rows <= 0)
    return;
  if (ell > 0) {
    switch (ell) {
    case 1: KH_ELL_LAUNCH(k_spmv_ell, 1, s, rows, idx, val, x, y, z, beta, act); break;
    case 2: KH_ELL_LAUNCH(k_spmv_ell, 2, s, rows, idx, val, x, y, z, beta, act); break;
    case 3: KH_ELL_LAUNCH(k_spmv_ell, 3, s, rows, idx, val, x, y, z, beta, act); break;
    case 4: KH_ELL_LAUNCH(k_spmv_ell, 4, s, rows, idx, val, x, y, z, beta, act); break;
    default: KH_ELL_LAUNCH(k_spmv_ell, 8, s, rows, idx, val, x, y, z, beta, act); break;
    }
    KH_CHECK(hipGetLastError());
    return;
  }
  switch (lpr) {
  case 1:
    hipLaunchKernelGGL(k_spmv_lane<T>, dim3(grid_rows(rows)), dim3(kBlock), 0, s,
                       rows, ptr, idx, val, x, y, z, beta);
    break;
  case 4:
    hipLaunchKernelGGL((k_spmv_group<T, 4>), dim3(grid_for(rows * 4)), dim3(kBlock), 0,
                       s, rows, ptr, idx, val, x, y, z, beta);
    break;
  case 16:
    hipLaunchKernelGGL((k_spmv_group<T, 16>), dim3(grid_for(rows * 16)), dim3(kBlock),
                       0, s, rows, ptr, idx, val, x, y, z, beta);
    break;
  default:
    hipLaunchKernelGGL((k_spmv_group<T, 64>), dim3(grid_for(rows * 64)), dim3(kBlock),
                       0, s, rows, ptr, idx, val, x, y, z, beta);
    break;
  }
  KH_CHECK(hipGetLastError());
}

#define KH_SPAN_LAUNCH(R, ROWS, FW)                                                          \
  hipLaunchKernelGGL((k_spmv_ell_span<T, R, ROWS, FW>), dim3((unsigned)nb),                   \
                     dim3(kSpanThreads), lds, ctx.stream, rows, span, len, lenx, idx, val, x, y, \
                     z, beta, act)
#define KH_SPAN_FW(R, ROWS)                                                                  \
  switch (fw) {                                                                              \
  case 8: KH_SPAN_LAUNCH(R, ROWS, 8); break;                                                 \
  case 16: KH_SPAN_LAUNCH(R, ROWS, 16); break;                                               \
  default: KH_SPAN_LAUNCH(R, ROWS, 20); break;                                               \
  }
#define KH_SPAN_ROWS(R)                                                                      \
  switch (rows_pt) {                                                                         \
  case 1: KH_SPAN_FW(R, 1) break;                                                            \
  case 2: KH_SPAN_FW(R, 2) break;                                                            \
  case 4: KH_SPAN_FW(R, 4) break;                                                            \
  default: KH_SPAN_FW(R, 8) break;                                                           \
  }

// k_spmv_ell_span over `rows` rows of R entries in blocks of rb rows
// (rb / kSpanThreads rows per thread, at most 32 / R; the window copied in
// 8, 16 or 20 16-byte loads per thread).
template <class T>
static void launch_spmv_span(Context &ctx, int64_t rows, int R, int rb, const int32_t *span,
                             int len, int64_t lenx, const int32_t *idx, const T *val,
                             const T *x, T *y, const T *z, const double *beta,
                             const int *act) {
  const int rows_pt = rb / kSpanThreads;
  const int64_t nb = (rows + rb - 1) / rb;
  const std::size_t lds = (std::size_t)len * sizeof(T);
  const int64_t nvec = (int64_t)lds / 16;
  const int fw = nvec <= 8 * kSpanThreads ? 8 : nvec <= 16 * kSpanThreads ? 16 : 20;
  if (rows_pt * kSpanThreads != rb || rows_pt > 8 || rows_pt * R > 32 || nb > INT32_MAX ||
      len < 4 || len % 4 != 0 || nvec > 20 * kSpanThreads)
    throw std::logic_error("spmv: span block size or window out of range");
  switch (R) {
  case 1: KH_SPAN_ROWS(1) break;
  case 2: KH_SPAN_ROWS(2) break;
  case 3: KH_SPAN_ROWS(3) break;
  case 4: KH_SPAN_ROWS(4) break;
  default:
    switch (rows_pt) {
    case 1: KH_SPAN_FW(8, 1) break;
    case 2: KH_SPAN_FW(8, 2) break;
    default: KH_SPAN_FW(8, 4) break;
    }
  }
  KH_CHECK(hipGetLastError());
}
#undef KH_SPAN_ROWS
#undef KH_SPAN_FW
#undef KH_SPAN_LAUNCH

template <class T>
static void spmv_dispatch(Context &ctx, const DeviceCsr &A, bool trans, int64_t rows,
                          const int32_t *ptr, const int32_t *idx, const T *val, int lpr,
                          int uni, const T *x, T *y, const T *z, const double *beta,
                          const int *act = nullptr) {
  const int ell = ell_width(uni, idx, val);
  const int di = trans ? 1 : 0, pi = std::is_same<T, double>::value ? 0 : 1;
  const int64_t lenx = trans ? A.m : A.n;
  // window k usable for this call: a table, its window within the LDS a
  // workgroup may hold, x aligned for the 16-byte window copy
  auto usable = [&](int k) {
    return ell > 0 && A.span_rb[di][k] > 0 &&
           (int64_t)A.span_len[di][k] * (int64_t)sizeof(T) <= ctx.lds_max &&
           reinterpret_cast<uintptr_t>(x) % 16 == 0;
  };
  auto run = [&](int path) {
    if (path >= 2)
      launch_spmv_span<T>(ctx, rows, ell, A.span_rb[di][path - 2], A.span[di][path - 2].as<int32_t>(),
                          A.span_len[di][path - 2], lenx, idx, val, x, y, z, beta, act);
    else
      launch_spmv<T>(ctx, rows, ptr, idx, val, lpr, ell, x, y, z, beta, act);
  };
  int path = A.spmv_path[pi][di].load(std::memory_order_relaxed);
  if (path == 0) {
    bool table = false;  // a window this operator and precision can ever use
    for (int k = 0; k < 2; ++k)
      table = table || (ell > 0 && A.span_rb[di][k] > 0 &&
                        (int64_t)A.span_len[di][k] * (int64_t)sizeof(T) <= ctx.lds_max);
    if (!table) {
      A.spmv_path[pi][di].store(1, std::memory_order_relaxed);
      run(1);
      return;
    }
    // first use: time the plain gather and each usable window (the faster
    // of two runs of five launches each, after one warm-up: single short
    // runs picked differently from one process to the next) when y is
    // written from inputs it does
    // not alias, x is aligned, no column flag may turn the launches into
    // no-ops and the stream is not being captured; every variant writes the
    // same y. Otherwise this call gathers from L2 and the choice waits for a
    // later call.
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    KH_CHECK(hipStreamIsCapturing(ctx.stream, &cs));
    const bool alias = (const void *)y == (const void *)x || (const void *)y == (const void *)z;
    if (cs != hipStreamCaptureStatusNone || alias || act ||
        reinterpret_cast<uintptr_t>(x) % 16 != 0) {
      run(1);
      return;
    }
    hipEvent_t e0, e1;
    KH_CHECK(hipEventCreate(&e0));
    KH_CHECK(hipEventCreate(&e1));
    float best = 0.0f;
    int pick = 1;
    for (int cand : {1, 2, 3}) {
      if (cand >= 2 && !usable(cand - 2))
        continue;
      run(cand);
      float ms = 0.0f;
      for (int round = 0; round < 2; ++round) {
        KH_CHECK(hipEventRecord(e0, ctx.stream));
        for (int r = 0; r < 5; ++r)
          run(cand);
        KH_CHECK(hipEventRecord(e1, ctx.stream));
        KH_CHECK(hipEventSynchronize(e1));
        float t = 0.0f;
        KH_CHECK(hipEventElapsedTime(&t, e0, e1));
        ms = round == 0 ? t : std::min(ms, t);
      }
      if (cand == 1 || ms < best) {
        best = ms;
        pick = cand;
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    A.spmv_path[pi][di].store((signed char)pick, std::memory_order_relaxed);
    return;  // y holds the (identical) product of the last run
  }
  // a window path picked for aligned x falls back to the gathers otherwise
  run(path >= 2 && usable(path - 2) ? path : 1);
}

int spmv_path(const DeviceCsr &A, bool trans, Prec p) {
  return A.spmv_path[p == Prec::F64 ? 0 : 1][trans ? 1 : 0].load(std::memory_order_relaxed);
}

void set_spmv_path(const DeviceCsr &A, bool trans, Prec p, int path) {
  A.spmv_path[p == Prec::F64 ? 0 : 1][trans ? 1 : 0].store((signed char)path,
                                                          std::memory_order_relaxed);
}

void spmv(Context &ctx, const DeviceCsr &A, bool trans, Prec p, const void *x,
          void *y, const void *z, const double *beta) {
  const int64_t rows = trans ? A.n : A.m;
  const double e = p == Prec::F64 ? 8.0 : 4.0;
  const int32_t *ptr = trans ? A.cp.as<int32_t>() : A.rp.as<int32_t>();
  const int32_t *idx = trans ? A.ri.as<int32_t>() : A.ci.as<int32_t>();
  const int lpr = trans ? A.lpr_cols : A.lpr_rows;
  const int uni = trans ? A.uniform_cols : A.uniform_rows;
  // the fixed-length kernels read no row pointers
  const void *vals = p == Prec::F64 ? (trans ? A.vt.p : A.va.p) : (trans ? A.vt32.p : A.va32.p);
  const bool fixed = ell_width(uni, idx, vals) > 0;
  TraceScope tr(ctx, kTraceSpmv,
                (e + 4.0) * A.nnz + (fixed ? 0.0 : 4.0 * (rows + 1)) + e * (trans ? A.m : A.n) +
                    e * rows * (z ? 2.0 : 1.0));
  if (rows <= 0)
    return;
  if (p == Prec::F64) {
    const double *val = trans ? A.vt.as<double>() : A.va.as<double>();
    spmv_dispatch<double>(ctx, A, trans, rows, ptr, idx, val, lpr, uni, (const double *)x,
                          (double *)y, (const double *)z, beta);
  } else {
    const float *val = trans ? A.vt32.as<float>() : A.va32.as<float>();
    spmv_dispatch<float>(ctx, A, trans, rows, ptr, idx, val, lpr, uni, (const float *)x,
                         (float *)y, (const float *)z, beta);
  }
}

template <class T>
static void launch_spmm(Context &ctx, int64_t rows, const int32_t *ptr,
                        const int32_t *idx, const T *val, int lpr, int ell, int s,
                        const T *x, int64_t ldx, T *y, int64_t ldy,
                        const int *act) {
  hipStream_t st = ctx.stream;
  if (rows <= 0 || s <= 0)
    return;
  if (ell > 0) {
    switch (ell) {
    case 1: KH_ELL_LAUNCH(k_spmm_ell, 1, st, rows, idx, val, s, x, ldx, y, ldy, act); break;
    case 2: KH_ELL_LAUNCH(k_spmm_ell, 2, st, rows, idx, val, s, x, ldx, y, ldy, act); break;
    case 3: KH_ELL_LAUNCH(k_spmm_ell, 3, st, rows, idx, val, s, x, ldx, y, ldy, act); break;
    case 4: KH_ELL_LAUNCH(k_spmm_ell, 4, st, rows, idx, val, s, x, ldx, y, ldy, act); break;
    default: KH_ELL_LAUNCH(k_spmm_ell, 8, st, rows, idx, val, s, x, ldx, y, ldy, act); break;
    }
    KH_CHECK(hipGetLastError());
    return;
  }
  if (lpr == 1)
    hipLaunchKernelGGL(k_spmm_lane<T>, dim3(grid_rows(rows)), dim3(kBlock), 0, st, rows,
                       ptr, idx, val, s, x, ldx, y, ldy, act);
  else if (lpr == 4)
    hipLaunchKernelGGL((k_spmm_group<T, 4>), dim3(grid_for(rows * 4)), dim3(kBlock), 0,
                       st, rows, ptr, idx, val, s, x, ldx, y, ldy, act);
  else if (lpr == 16)
    hipLaunchKernelGGL((k_spmm_group<T, 16>), dim3(grid_for(rows * 16)), dim3(kBlock),
                       0, st, rows, ptr, idx, val, s, x, ldx, y, ldy, act);
  else
    hipLaunchKernelGGL((k_spmm_group<T, 64>), dim3(grid_for(rows * 64)), dim3(kBlock),
                       0, st, rows, ptr, idx, val, s, x, ldx, y, ldy, act);
  KH_CHECK(hipGetLastError());
}

void spmm(Context &ctx, const DeviceCsr &A, bool trans, Prec p, int s,
          const void *x, int64_t ldx, void *y, int64_t ldy, const int *active) {
  const int64_t rows = trans ? A.n : A.m;
  const int32_t *ptr = trans ? A.cp.as<int32_t>() : A.rp.as<int32_t>();
  const int32_t *idx = trans ? A.ri.as<int32_t>() : A.ci.as<int32_t>();
  const int lpr = trans ? A.lpr_cols : A.lpr_rows;
  const int uni = trans ? A.uniform_cols : A.uniform_rows;
  // one column of an operator with fixed-length rows: the single-vector
  // kernels (bitwise the same sums; they include the LDS-window gathers),
  // which leave y untouched when the column's `active` flag is 0
  const bool one = s == 1 && rows > 0 &&
                   ell_width(uni, idx, p == Prec::F64 ? (trans ? A.vt.p : A.va.p)
                                                      : (trans ? A.vt32.p : A.va32.p)) > 0;
  if (one && p == Prec::F64) {
    const double *val = trans ? A.vt.as<double>() : A.va.as<double>();
    spmv_dispatch<double>(ctx, A, trans, rows, ptr, idx, val, lpr, uni, (const double *)x,
                          (double *)y, nullptr, nullptr, active);
  } else if (one) {
    const float *val = trans ? A.vt32.as<float>() : A.va32.as<float>();
    spmv_dispatch<float>(ctx, A, trans, rows, ptr, idx, val, lpr, uni, (const float *)x,
                         (float *)y, nullptr, nullptr, active);
  } else if (p == Prec::F64) {
    const double *val = trans ? A.vt.as<double>() : A.va.as<double>();
    launch_spmm<double>(ctx, rows, ptr, idx, val, lpr, ell_width(uni, idx, val), s,
                        (const double *)x, ldx, (double *)y, ldy, active);
  } else {
    const float *val = trans ? A.vt32.as<float>() : A.va32.as<float>();
    launch_spmm<float>(ctx, rows, ptr, idx, val, lpr, ell_width(uni, idx, val), s,
                       (const float *)x, ldx, (float *)y, ldy, active);
  }
}

void dot(Context &ctx, Prec pa, const void *a, Prec pb, const void *b,
         int64_t len, double *out, bool do_sqrt) {
  const int nb = reduce_blocks(len);
  double *part = ctx.partials(nb);
  hipStream_t s = ctx.stream;
  if (pa == Prec::F64 && pb == Prec::F64)
    hipLaunchKernelGGL((k_dot_partial<double, double>), dim3(nb), dim3(kBlock), 0, s, len,
                       (const double *)a, (const double *)b, part);
  else if (pa == Prec::F32 && pb == Prec::F32)
    hipLaunchKernelGGL((k_dot_partial<float, float>), dim3(nb), dim3(kBlock), 0, s, len,
                       (const float *)a, (const float *)b, part);
  else if (pa == Prec::F64)
    hipLaunchKernelGGL((k_dot_partial<double, float>), dim3(nb), dim3(kBlock), 0, s, len,
                       (const double *)a, (const float *)b, part);
  else
    hipLaunchKernelGGL((k_dot_partial<float, double>), dim3(nb), dim3(kBlock), 0, s, len,
                       (const float *)a, (const double *)b, part);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kBlock), 0, s, part, nb, out, do_sqrt ? 1 : 0);
  KH_CHECK(hipGetLastError());
}

inline int vtw_blocks(int64_t len) {
  int64_t g = (len + kRB - 1) / kRB;
  return (int)(g < 1 ? 1 : g);
}

// The partial kernel of V^T w: 16-byte loads for aligned FP32 blocks.
// With events ea / eb (tracing) the launch stamps the kernel's start and end.
template <class TV, class TW>
static void launch_vtw_partial(Context &ctx, const TV *V, int64_t ld, int s, const TW *w,
                               int64_t len, double *part, int nb, hipEvent_t ea = nullptr,
                               hipEvent_t eb = nullptr, int wsq = 0) {
  const std::size_t lds = (std::size_t)(s + wsq) * kWaves * sizeof(double);
  if constexpr (std::is_same<TV, float>::value && std::is_same<TW, float>::value) {
    if (len % 4 == 0 && ld % 4 == 0 && reinterpret_cast<uintptr_t>(V) % 16 == 0 &&
        reinterpret_cast<uintptr_t>(w) % 16 == 0) {
      hipExtLaunchKernelGGL(k_vtw4_partial, dim3(nb), dim3(kBlock), lds, ctx.stream, ea, eb, 0,
                            len, V, ld, s, w, part, nb, wsq);
      return;
    }
  }
  hipExtLaunchKernelGGL((k_vtw_partial<TV, TW>), dim3(nb), dim3(kBlock), lds, ctx.stream, ea, eb,
                        0, len, V, ld, s, w, part, nb, wsq);
}

template <class TV, class TW>
static void launch_gemv_t(Context &ctx, const TV *V, int64_t ld, int s, const TW *w,
                          int64_t len, double *out, bool wsq = false) {
  const int nb = vtw_blocks(len);
  if (wsq && s >= 512)
    throw std::invalid_argument("gemv_t: ||w||^2 column with at most 511 columns");
  for (int c0 = 0; c0 < s; c0 += 512) {  // LDS holds 512 x 4 partials
    const int sc = std::min(512, s - c0);
    const int sq = wsq ? 1 : 0;  // single chunk when wsq
    double *part = ctx.partials((std::size_t)nb * (sc + sq));
    {
      // class 1 is the FP32 reorthogonalisation product (k_vtw4_partial for
      // aligned blocks), the kernel the roofline reports
      constexpr bool kF32 = std::is_same<TV, float>::value && std::is_same<TW, float>::value;
      TraceScope tr(ctx, kF32 ? kTraceVtW : kTraceOther,
                    (double)len * sc * sizeof(TV) + (double)len * sizeof(TW) +
                        8.0 * nb * (sc + sq),
                    len, sc, /*stamped=*/true);
      launch_vtw_partial(ctx, V + (int64_t)c0 * ld, ld, sc, w, len, part, nb, tr.start(),
                         tr.stop(), sq);
    }
    hipLaunchKernelGGL(k_finish, dim3(sc + sq), dim3(kBlock), 0, ctx.stream, part, nb, out + c0,
                       0);
  }
  KH_CHECK(hipGetLastError());
}

template <class T>
static double time_vtw_t(Context &ctx, int64_t len, int s, int reps) {
  if (len <= 0 || s <= 0 || s > 512 || reps < 1)
    throw std::invalid_argument("time_vtw: 0 < s <= 512, len > 0, reps >= 1");
  const int nb = vtw_blocks(len);
  double *part = ctx.partials((std::size_t)nb * s);
  // rotate over copies of V and w totalling at least 1 GiB, so that no
  // launch finds its block in the 256 MB Infinity Cache (as in a solve,
  // where the block is read once per pass between other sweeps)
  const std::size_t vb = (std::size_t)len * s * sizeof(T), wb = (std::size_t)len * sizeof(T);
  const int copies = (int)std::min<std::size_t>(16, std::max<std::size_t>(
                                                        2, ((std::size_t)1 << 30) / (vb + wb) + 1));
  std::vector<DBuf> V, w;
  for (int i = 0; i < copies; ++i) {
    V.emplace_back(vb);
    w.emplace_back(wb);
    KH_CHECK(hipMemsetAsync(V.back().p, 0, vb, ctx.stream));
    KH_CHECK(hipMemsetAsync(w.back().p, 0, wb, ctx.stream));
  }
  int turn = 0;
  auto launch = [&] {
    const int i = turn++ % copies;
    launch_vtw_partial(ctx, (const T *)V[(std::size_t)i].p, len, s, (const T *)w[(std::size_t)i].p,
                       len, part, nb);
  };
  for (int i = 0; i < copies; ++i)
    launch();
  hipEvent_t e0, e1;
  KH_CHECK(hipEventCreate(&e0));
  KH_CHECK(hipEventCreate(&e1));
  KH_CHECK(hipEventRecord(e0, ctx.stream));
  for (int r = 0; r < reps; ++r)
    launch();
  KH_CHECK(hipEventRecord(e1, ctx.stream));
  KH_CHECK(hipGetLastError());
  KH_CHECK(hipEventSynchronize(e1));
  float ms = 0.0f;
  KH_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

double time_vtw(Context &ctx, Prec p, int64_t len, int s, int reps) {
  return p == Prec::F64 ? time_vtw_t<double>(ctx, len, s, reps)
                        : time_vtw_t<float>(ctx, len, s, reps);
}

std::size_t cgs2_partials_bound(int64_t len, int s) {
  const std::size_t nb = (std::size_t)vtw_blocks(len), sc = (std::size_t)std::max(s, 0);
  std::size_t need = nb * std::min<std::size_t>(sc + 1, 512);  // V^T w (+ ||w||^2)
  need = std::max(need, nb * sc);                           // two-sweep middle (k_vh_sub_vtw)
  need = std::max(need, nb);                                // norm of the last update
  const std::size_t ntl = (std::size_t)((std::max<int64_t>(len, 0) + kLdsRows - 1) / kLdsRows);
  need = std::max(need, std::min<std::size_t>(ntl, 16384) * (sc + 1));  // LDS-tiled middle
  return need;
}

void gemv_t(Context &ctx, Prec pv, const void *V, int64_t ld, int s, Prec pw,
            const void *w, int64_t len, double *out, bool wsq) {
  if (s <= 0)
    return;
  if (pv == Prec::F64 && pw == Prec::F64)
    launch_gemv_t(ctx, (const double *)V, ld, s, (const double *)w, len, out, wsq);
  else if (pv == Prec::F32 && pw == Prec::F32)
    launch_gemv_t(ctx, (const float *)V, ld, s, (const float *)w, len, out, wsq);
  else if (pv == Prec::F64)
    launch_gemv_t(ctx, (const double *)V, ld, s, (const float *)w, len, out, wsq);
  else
    launch_gemv_t(ctx, (const float *)V, ld, s, (const double *)w, len, out, wsq);
}

template <class TV, class TW>
static void launch_gemv_n_sub(Context &ctx, const TV *V, int64_t ld, int s,
                              const double *h, TW *w, int64_t len, double *nrm,
                              TW *dst = nullptr, const double *g = nullptr,
                              double *ran = nullptr, double cancel = 0.5) {
  const int nb = vtw_blocks(len);
  double *part = nrm ? ctx.partials(nb) : nullptr;
  {
  TraceScope tr(ctx, kTraceVhSub,
                (double)len * s * sizeof(TV) + 2.0 * len * sizeof(TW) + 8.0 * s);
  hipLaunchKernelGGL((k_vh_sub<TV, TW>), dim3(nb), dim3(kBlock),
                     (std::size_t)std::max(s, 1) * sizeof(double), ctx.stream, len, V, ld, s, h,
                     w, part, g, ran, cancel);
  }
  // every block of the fused norm-and-copy re-reads the nb partials: used up
  // to 1024 partials (2^21 rows) on at most 2048 blocks (<= 16 MB of L2
  // reads); longer vectors take the norm and the copy as two launches
  if (nrm && dst && nb <= 1024) {
    TraceScope tr(ctx, kTraceOther, 2.0 * len * sizeof(TW));  // w read, dst written
    hipLaunchKernelGGL(k_norm_scale_copy<TW>, dim3(std::min(grid_for(len), 2048)), dim3(kBlock),
                       0, ctx.stream, part, nb, nrm, len, (const TW *)w, dst);
  } else if (nrm) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(kBlock), 0, ctx.stream, part, nb, nrm, 1);
    if (dst)
      hipLaunchKernelGGL(k_scale_copy_dev<TW>, dim3(grid_for(len)), dim3(kBlock), 0, ctx.stream,
                         len, (const double *)nrm, (const TW *)w, dst);
  }
  KH_CHECK(hipGetLastError());
}

void gemv_n_sub_normalise(Context &ctx, Prec pv, const void *V, int64_t ld, int s,
                          const double *h, Prec pw, void *w, int64_t len, double *nrm,
                          void *dst, const double *g, double *ran, double cancel) {
  if (pv == Prec::F64 && pw == Prec::F64)
    launch_gemv_n_sub(ctx, (const double *)V, ld, s, h, (double *)w, len, nrm, (double *)dst, g,
                      ran, cancel);
  else if (pv == Prec::F32 && pw == Prec::F32)
    launch_gemv_n_sub(ctx, (const float *)V, ld, s, h, (float *)w, len, nrm, (float *)dst, g, ran,
                      cancel);
  else if (pv == Prec::F64)
    launch_gemv_n_sub(ctx, (const double *)V, ld, s, h, (float *)w, len, nrm, (float *)dst, g,
                      ran, cancel);
  else
    launch_gemv_n_sub(ctx, (const float *)V, ld, s, h, (double *)w, len, nrm, (double *)dst, g,
                      ran, cancel);
}

void gemv_n_sub(Context &ctx, Prec pv, const void *V, int64_t ld, int s,
                const double *h, Prec pw, void *w, int64_t len, double *nrm2_out,
                const double *g, double *ran, double cancel) {
  if (s <= 0 && !nrm2_out)
    return;
  if (pv == Prec::F64 && pw == Prec::F64)
    launch_gemv_n_sub(ctx, (const double *)V, ld, s, h, (double *)w, len, nrm2_out,
                      (double *)nullptr, g, ran, cancel);
  else if (pv == Prec::F32 && pw == Prec::F32)
    launch_gemv_n_sub(ctx, (const float *)V, ld, s, h, (float *)w, len, nrm2_out,
                      (float *)nullptr, g, ran, cancel);
  else if (pv == Prec::F64)
    launch_gemv_n_sub(ctx, (const double *)V, ld, s, h, (float *)w, len, nrm2_out,
                      (float *)nullptr, g, ran, cancel);
  else
    launch_gemv_n_sub(ctx, (const float *)V, ld, s, h, (double *)w, len, nrm2_out,
                      (double *)nullptr, g, ran, cancel);
}

template <class TV, class TW>
static void launch_gemv_n_sub_t(Context &ctx, const TV *V, int64_t ld, int s, const double *h,
                                TW *w, int64_t len, double *out, double *nrm1,
                                const double *g = nullptr) {
  if (g && nrm1)
    throw std::invalid_argument("gemv_n_sub_t: a selective pair takes no norm here");
  // LDS-tiled form when its tile fits the workgroup's LDS (160 KB on gfx950:
  // FP64 tiles up to 128 columns) and the columns fill it (below about 12
  // columns the two-sweep kernel is faster)
  const std::size_t lds = kLdsRows * 8 + (std::size_t)s * kLdsStride * sizeof(TV);
  const int64_t ntl = (len + kLdsRows - 1) / kLdsRows;
  constexpr bool kF32 = std::is_same<TV, float>::value;
  const bool vec = kF32 && len % 4 == 0 && ld % 4 == 0 &&
                   reinterpret_cast<uintptr_t>(V) % 16 == 0;
  // FP32 blocks with 16-byte loads: at most 4096 blocks, each walking its
  // tiles with the next tile's loads in flight (a few tiles per block at 2^21
  // rows); otherwise one 128-row tile per block up to 16384 tiles (2^21
  // rows), and beyond that 16384 blocks each take every 16384-th tile (the
  // dword-load pipelined form needs many more registers). Both bound the
  // partials per column that k_finish sums.
  const int64_t nbl = std::min<int64_t>(ntl, vec ? 4096 : 16384);
  if (s >= 12 && s <= 128 && lds <= (std::size_t)ctx.lds_max &&
      (std::size_t)nbl * (s + 1) <= ctx.part_capacity()) {
    double *part = ctx.partials((std::size_t)nbl * (s + 1));
    double *npart = nrm1 ? part + (std::size_t)nbl * s : nullptr;
    {
      TraceScope tr(ctx, kTraceVhSub,
                    (double)len * s * sizeof(TV) + 2.0 * len * sizeof(TW) + 8.0 * nbl * s);
      const int hm = (s + 1) / 2;
      const bool multi = ntl > nbl;
#define KH_LDS_LAUNCH(H, M, VC)                                                                \
  hipLaunchKernelGGL((k_vh_sub_vtw_lds<TV, TW, H, M, VC>), dim3((unsigned)nbl), dim3(kBlock),  \
                     lds, ctx.stream, len, V, ld, s, h, w, part, (int)nbl, npart, g)
#define KH_LDS_PICK(H)                                                                         \
  {                                                                                            \
    if (vec) {                                                                                 \
      if (multi) KH_LDS_LAUNCH(H, true, kF32); else KH_LDS_LAUNCH(H, false, kF32);             \
    } else {                                                                                   \
      if (multi) KH_LDS_LAUNCH(H, true, false); else KH_LDS_LAUNCH(H, false, false);           \
    }                                                                                          \
  }
      if (hm <= 16)
        KH_LDS_PICK(16)
      else if (hm <= 32)
        KH_LDS_PICK(32)
      else if (hm <= 48)
        KH_LDS_PICK(48)
      else
        KH_LDS_PICK(64)
#undef KH_LDS_PICK
#undef KH_LDS_LAUNCH
    }
    if (g)
      hipLaunchKernelGGL(k_finish_pass2, dim3(s), dim3(kBlock), 0, ctx.stream, part, (int)nbl,
                         out, g);
    else
      hipLaunchKernelGGL(k_finish, dim3(s), dim3(kBlock), 0, ctx.stream, part, (int)nbl, out, 0);
    if (nrm1)
      hipLaunchKernelGGL(k_finish, dim3(1), dim3(kBlock), 0, ctx.stream, npart, (int)nbl, nrm1, 1);
    KH_CHECK(hipGetLastError());
    return;
  }
  const int nb = vtw_blocks(len);
  double *part = ctx.partials((std::size_t)nb * s);
  {
    // V read once from HBM, w read and written, partials written
    TraceScope tr(ctx, kTraceVhSub,
                  (double)len * s * sizeof(TV) + 2.0 * len * sizeof(TW) + 8.0 * nb * s);
    hipLaunchKernelGGL((k_vh_sub_vtw<TV, TW>), dim3(nb), dim3(kBlock),
                       (std::size_t)s * (1 + kWaves) * sizeof(double), ctx.stream, len, V, ld, s,
                       h, w, part, nb, g);
  }
  if (g)
    hipLaunchKernelGGL(k_finish_pass2, dim3(s), dim3(kBlock), 0, ctx.stream, part, nb, out, g);
  else
    hipLaunchKernelGGL(k_finish, dim3(s), dim3(kBlock), 0, ctx.stream, part, nb, out, 0);
  KH_CHECK(hipGetLastError());
  if (nrm1)  // after the finish: the dot reuses the partial buffer
    dot(ctx, std::is_same<TW, double>::value ? Prec::F64 : Prec::F32, w,
        std::is_same<TW, double>::value ? Prec::F64 : Prec::F32, w, len, nrm1, true);
}

void gemv_n_sub_t(Context &ctx, Prec pv, const void *V, int64_t ld, int s, const double *h,
                  Prec pw, void *w, int64_t len, double *out, double *nrm1, const double *g) {
  if (s <= 0)
    return;
  if (s > 1024)
    throw std::invalid_argument("gemv_n_sub_t: more than 1024 columns");
  if (pv == Prec::F64 && pw == Prec::F64)
    launch_gemv_n_sub_t(ctx, (const double *)V, ld, s, h, (double *)w, len, out, nrm1, g);
  else if (pv == Prec::F32 && pw == Prec::F32)
    launch_gemv_n_sub_t(ctx, (const float *)V, ld, s, h, (float *)w, len, out, nrm1, g);
  else if (pv == Prec::F64)
    launch_gemv_n_sub_t(ctx, (const double *)V, ld, s, h, (float *)w, len, out, nrm1, g);
  else
    launch_gemv_n_sub_t(ctx, (const float *)V, ld, s, h, (double *)w, len, out, nrm1, g);
}

// k_nn_mfma when the block fits it (q <= 128 output columns, H^T in 64 KB
// of LDS); returns false otherwise.
template <class TV, class TO, bool SUB>
static bool try_nn_mfma(Context &ctx, const TV *V, int64_t ldv, int s, const double *H, int q,
                        TO *out, int64_t ldo, int64_t len) {
  if (q < 2 || q > 128 || s < 1 || len <= 0)
    return false;
  const int sld = ((s + 3) & ~3) + 1;
  const int qp = ((q + 15) / 16) * 16;
  const std::size_t lds = (std::size_t)qp * sld * sizeof(double);
  if (lds > 65536)
    return false;
  if (q <= 32) {  // two 16-row groups per wave
    const int64_t chunks = (len + 127) / 128;
    const unsigned grid = (unsigned)std::min<int64_t>(chunks, 4096);
    hipLaunchKernelGGL((k_nn_mfma<TV, TO, SUB, 2, 2, true>), dim3(grid), dim3(kBlock), lds, ctx.stream,
                       len, V, ldv, s, H, q, out, ldo, sld);
  } else {
    const int64_t chunks = (len + 63) / 64;
    const unsigned grid = (unsigned)std::min<int64_t>(chunks, 4096);
    hipLaunchKernelGGL((k_nn_mfma<TV, TO, SUB, 1, 8, false>), dim3(grid), dim3(kBlock), lds, ctx.stream,
                       len, V, ldv, s, H, q, out, ldo, sld);
  }
  KH_CHECK(hipGetLastError());
  return true;
}

template <class TV, class TO>
static void launch_gemm_nn(Context &ctx, const TV *V, int64_t ldv, int s, const double *H,
                           int q, TO *out, int64_t ldo, int64_t len) {
  if (try_nn_mfma<TV, TO, false>(ctx, V, ldv, s, H, q, out, ldo, len))
    return;
  if (q == 1) {
    dim3 grid(grid_for(len, 4), 1);
    hipLaunchKernelGGL((k_gemm_nn<TV, TO>), grid, dim3(kBlock), 0, ctx.stream, len, V, ldv,
                       s, H, q, out, ldo);
  } else {
    dim3 grid(grid_for(len, 2), (q + kQT - 1) / kQT);
    hipLaunchKernelGGL((k_ts_nn<TV, TO, false>), grid, dim3(kBlock), 0, ctx.stream, len, V,
                       ldv, s, H, q, out, ldo);
  }
  KH_CHECK(hipGetLastError());
}

void gemm_nn(Context &ctx, Prec pv, const void *V, int64_t ldv, int s,
             const double *H, int q, Prec po, void *OUT, int64_t ldo, int64_t len) {
  if (q <= 0)
    return;
  TraceScope tr(ctx, kTraceOther,
                (double)len * s * (pv == Prec::F64 ? 8 : 4) +
                    (double)len * q * (po == Prec::F64 ? 8 : 4));  // V read, OUT written
  if (pv == Prec::F64 && po == Prec::F64)
    launch_gemm_nn(ctx, (const double *)V, ldv, s, H, q, (double *)OUT, ldo, len);
  else if (pv == Prec::F32 && po == Prec::F32)
    launch_gemm_nn(ctx, (const float *)V, ldv, s, H, q, (float *)OUT, ldo, len);
  else if (pv == Prec::F32)
    launch_gemm_nn(ctx, (const float *)V, ldv, s, H, q, (double *)OUT, ldo, len);
  else
    launch_gemm_nn(ctx, (const double *)V, ldv, s, H, q, (float *)OUT, ldo, len);
}

void gemv_n(Context &ctx, Prec pv, const void *V, int64_t ld, int s, const double *h,
            Prec po, void *out, int64_t len) {
  gemm_nn(ctx, pv, V, ld, s, h, 1, po, out, len, len);
}

template <class TV, class TW>
static void launch_gemm_tn(Context &ctx, const TV *V, int64_t ldv, int s, const TW *W,
                           int64_t ldw, int q, int64_t len, double *H) {
  const int ty = (s + kTS - 1) / kTS, tz = (q + kTS - 1) / kTS;
  int nblk = (int)std::min<int64_t>(512, std::max<int64_t>(1, (len + 2047) / 2048));
  int64_t rpb = (len + nblk - 1) / nblk;
  rpb = (rpb + kSlab - 1) / kSlab * kSlab;
  nblk = (int)std::max<int64_t>(1, (len + rpb - 1) / rpb);
  double *part = ctx.partials((std::size_t)nblk * ty * tz * kTS * kTS);
  hipLaunchKernelGGL((k_ts_tn_partial<TV, TW>), dim3(nblk, ty, tz), dim3(kBlock), 0,
                     ctx.stream, len, rpb, V, ldv, s, W, ldw, q, part);
  const int64_t outs = (int64_t)ty * tz * kTS * kTS;
  hipLaunchKernelGGL(k_ts_finish, dim3((unsigned)((outs + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, ctx.stream, part, nblk, ty, s, q, H);
  KH_CHECK(hipGetLastError());
}

void gram_mfma(Context &ctx, const float *V, int64_t ldv, int s, const float *W, int64_t ldw,
               int q, int64_t len, double *H) {
  if (s <= 0 || q <= 0 || len <= 0)
    return;
  const int ty = (s + kMT - 1) / kMT, tz = (q + kMT - 1) / kMT;
  int nblk = (int)std::min<int64_t>(512, std::max<int64_t>(1, (len + 4095) / 4096));
  int64_t rpb = (len + nblk - 1) / nblk;
  rpb = (rpb + kMSlab - 1) / kMSlab * kMSlab;
  nblk = (int)std::max<int64_t>(1, (len + rpb - 1) / rpb);
  const std::size_t need = (std::size_t)nblk * ty * tz * kMT * kMT;  // floats
  float *part = reinterpret_cast<float *>(ctx.partials((need + 1) / 2));
  TraceScope tr(ctx, kTraceGram, 2.0 * (double)len * s * q);
  hipLaunchKernelGGL(k_gram_mfma_partial, dim3(nblk, ty, tz), dim3(kBlock), 0, ctx.stream, len,
                     rpb, V, ldv, s, W, ldw, q, part);
  const int64_t outs = (int64_t)ty * tz * kMT * kMT;
  hipLaunchKernelGGL(k_gram_finish, dim3((unsigned)((outs + kBlock - 1) / kBlock)), dim3(kBlock),
                     0, ctx.stream, part, nblk, ty, s, q, H);
  KH_CHECK(hipGetLastError());
}

void sd_apply(Context &ctx, int64_t n, int64_t N, double c, bool trans, Prec p, const void *x,
              void *y, const void *z, const double *beta) {
  const int64_t rows = trans ? n : n - N;
  if (rows <= 0)
    return;
  const double e = p == Prec::F64 ? 8.0 : 4.0;
  TraceScope tr(ctx, kTraceSpmv, e * (double)(n + (n - N)) + (z ? e * rows : 0.0));
  const dim3 g(grid_for(rows));
  if (p == Prec::F64) {
    if (trans)
      hipLaunchKernelGGL(k_sd_mtv<double>, g, dim3(kBlock), 0, ctx.stream, n, N, c,
                         (const double *)x, (double *)y, (const double *)z, beta);
    else
      hipLaunchKernelGGL(k_sd_mv<double>, g, dim3(kBlock), 0, ctx.stream, n - N, N, c,
                         (const double *)x, (double *)y, (const double *)z, beta);
  } else {
    if (trans)
      hipLaunchKernelGGL(k_sd_mtv<float>, g, dim3(kBlock), 0, ctx.stream, n, N, c,
                         (const float *)x, (float *)y, (const float *)z, beta);
    else
      hipLaunchKernelGGL(k_sd_mv<float>, g, dim3(kBlock), 0, ctx.stream, n - N, N, c,
                         (const float *)x, (float *)y, (const float *)z, beta);
  }
  KH_CHECK(hipGetLastError());
}

void sd_block(Context &ctx, int64_t n, int64_t N, double c, bool trans, Prec p, int s,
              const void *x, int64_t ldx, void *y, int64_t ldy, const int *act, bool row_major) {
  const int64_t rows = trans ? n : n - N;
  if (rows <= 0 || s <= 0)
    return;
  int64_t g = (rows * (int64_t)s + kBlock - 1) / kBlock;
  g = std::max<int64_t>(1, std::min<int64_t>(g, 65536));
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_sd_block<double>, dim3((unsigned)g), dim3(kBlock), 0, ctx.stream, n, N,
                       c, trans ? 1 : 0, s, (const double *)x, ldx, (double *)y, ldy, act,
                       row_major ? 1 : 0);
  else
    hipLaunchKernelGGL(k_sd_block<float>, dim3((unsigned)g), dim3(kBlock), 0, ctx.stream, n, N,
                       c, trans ? 1 : 0, s, (const float *)x, ldx, (float *)y, ldy, act,
                       row_major ? 1 : 0);
  KH_CHECK(hipGetLastError());
}

void sub_beta(Context &ctx, Prec p, int64_t len, const double *beta, const void *z, void *y) {
  if (len <= 0)
    return;
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_sub_beta<double>, dim3(grid_for(len)), dim3(kBlock), 0, ctx.stream, len,
                       beta, (const double *)z, (double *)y);
  else
    hipLaunchKernelGGL(k_sub_beta<float>, dim3(grid_for(len)), dim3(kBlock), 0, ctx.stream, len,
                       beta, (const float *)z, (float *)y);
  KH_CHECK(hipGetLastError());
}

void copy_active(Context &ctx, Prec p, int64_t len, int s, const void *X, int64_t ldx, void *Y,
                 int64_t ldy, const int *act) {
  if (len <= 0 || s <= 0)
    return;
  int64_t g = (len * (int64_t)s + kBlock - 1) / kBlock;
  g = std::max<int64_t>(1, std::min<int64_t>(g, 65536));
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_copy_active<double>, dim3((unsigned)g), dim3(kBlock), 0, ctx.stream,
                       len, s, (const double *)X, ldx, (double *)Y, ldy, act);
  else
    hipLaunchKernelGGL(k_copy_active<float>, dim3((unsigned)g), dim3(kBlock), 0, ctx.stream, len,
                       s, (const float *)X, ldx, (float *)Y, ldy, act);
  KH_CHECK(hipGetLastError());
}

void spmm_rows(Context &ctx, const DeviceCsr &A, bool trans, Prec p, int s, const void *X,
               void *Y) {
  const int64_t rows = trans ? A.n : A.m;
  if (rows <= 0 || s <= 0)
    return;
  const double e = p == Prec::F64 ? 8.0 : 4.0;
  TraceScope tr(ctx, kTraceSpmv,
                (e + 4.0) * A.nnz + 4.0 * (rows + 1) + e * s * ((trans ? A.m : A.n) + rows));
  const int32_t *ptr = trans ? A.cp.as<int32_t>() : A.rp.as<int32_t>();
  const int32_t *idx = trans ? A.ri.as<int32_t>() : A.ci.as<int32_t>();
  auto run = [&](int cw) {  // all columns, in chunks of cw
    for (int c0 = 0; c0 < s; c0 += cw) {
      const int w = std::min(cw, s - c0);
      int64_t g = (rows * (int64_t)w + kBlock - 1) / kBlock;
      g = std::max<int64_t>(1, std::min<int64_t>(g, (int64_t)1 << 30));
      if (p == Prec::F64)
        hipLaunchKernelGGL(k_spmm_rows<double>, dim3((unsigned)g), dim3(kBlock), 0, ctx.stream,
                           rows, ptr, idx, trans ? A.vt.as<double>() : A.va.as<double>(), s, c0,
                           w, (const double *)X, (double *)Y);
      else
        hipLaunchKernelGGL(k_spmm_rows<float>, dim3((unsigned)g), dim3(kBlock), 0, ctx.stream,
                           rows, ptr, idx, trans ? A.vt32.as<float>() : A.va32.as<float>(), s,
                           c0, w, (const float *)X, (float *)Y);
      KH_CHECK(hipGetLastError());
    }
  };
  // The gathered block (input length x s) beyond 160 MB does not stay in the
  // 256 MB MALL: chunks of 16 or 32 columns may then be faster. The choice
  // is timed once per precision and direction (on a block of 32 or more
  // columns, outside stream capture) and kept on the matrix.
  const double xbytes = e * s * (double)(trans ? A.m : A.n);
  int cw = s;
  if (s >= 32 && xbytes > 160e6) {
    std::atomic<int> &slot = A.rows_chunk[p == Prec::F64 ? 0 : 1][trans ? 1 : 0];
    int pick = slot.load(std::memory_order_relaxed);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    KH_CHECK(hipStreamIsCapturing(ctx.stream, &cs));
    if (pick == 0 && cs == hipStreamCaptureStatusNone) {
      hipEvent_t e0, e1;
      KH_CHECK(hipEventCreate(&e0));
      KH_CHECK(hipEventCreate(&e1));
      float best = 0.0f;
      for (int cand : {1 << 30, 32, 16}) {
        run(std::min(cand, s));  // warm
        KH_CHECK(hipEventRecord(e0, ctx.stream));
        run(std::min(cand, s));
        KH_CHECK(hipEventRecord(e1, ctx.stream));
        KH_CHECK(hipEventSynchronize(e1));
        float ms = 0.0f;
        KH_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (pick == 0 || ms < best)
          best = ms, pick = cand;
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      slot.store(pick, std::memory_order_relaxed);
      return;  // Y holds the product (every candidate computes it fully)
    }
    if (pick > 0)
      cw = std::min(pick, s);
  }
  run(cw);
}

void cols_to_rows(Context &ctx, Prec p, int64_t len, int s, const void *X, int64_t ldx, void *R) {
  if (len <= 0 || s <= 0)
    return;
  const dim3 grid((unsigned)((len + 63) / 64), (unsigned)((s + 63) / 64));
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_cols_to_rows<double>, grid, dim3(kBlock), 0, ctx.stream, len, s,
                       (const double *)X, ldx, (double *)R);
  else
    hipLaunchKernelGGL(k_cols_to_rows<float>, grid, dim3(kBlock), 0, ctx.stream, len, s,
                       (const float *)X, ldx, (float *)R);
  KH_CHECK(hipGetLastError());
}

void rows_to_cols(Context &ctx, Prec p, int64_t len, int s, const void *R, void *X, int64_t ldx) {
  if (len <= 0 || s <= 0)
    return;
  const dim3 grid((unsigned)((len + 63) / 64), (unsigned)((s + 63) / 64));
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_rows_to_cols<double>, grid, dim3(kBlock), 0, ctx.stream, len, s,
                       (const double *)R, (double *)X, ldx);
  else
    hipLaunchKernelGGL(k_rows_to_cols<float>, grid, dim3(kBlock), 0, ctx.stream, len, s,
                       (const float *)R, (float *)X, ldx);
  KH_CHECK(hipGetLastError());
}

// mode 0: SG = [Z^T Z | Z^T AZ] (s == q); mode 1: SG = Z^T AZ (s x q).
// T is the MFMA type; Z and AZ (TZ, TA) are converted to it when staged.
template <class T, class TZ = T, class TA = T>
static void launch_gram2(Context &ctx, const TZ *Z, int64_t ldz, int s, const TA *AZ,
                         int64_t ldaz, int q, int64_t len, double *SG, int mode) {
  static_assert(sizeof(TZ) <= sizeof(T) && sizeof(TA) <= sizeof(T), "exact conversion");
  if (q <= 0 || s <= 0 || len <= 0)
    return;
  if (q > kG2MaxQ || s > kG2MaxQ)
    throw std::invalid_argument("gram2_mfma: more than 128 columns");
  const int ns = (s + 15) / 16, nq = (q + 15) / 16;
  const int ntiles = mode == 0 ? nq * (nq + 1) / 2 + nq * nq : ns * nq;
  if (mode == 1 && s <= 16 && q <= 16) {
    // one tile: 256-row slabs, about 2048 blocks
    int64_t rpb = (len + 2047) / 2048;
    rpb = std::max<int64_t>((rpb + kTnSlab - 1) / kTnSlab * kTnSlab, kTnSlab);
    const int64_t nblk = (len + rpb - 1) / rpb;
    double *part = ctx.partials((std::size_t)nblk * 256);
    hipLaunchKernelGGL((k_tn_narrow<TZ, TA>), dim3((unsigned)nblk), dim3(kG2Threads), 0,
                       ctx.stream, len, rpb, Z, ldz, s, AZ, ldaz, q, part);
    const int64_t outs = (int64_t)s * q * 64;  // one wave per output
    hipLaunchKernelGGL(k_tn2_finish<double>, dim3((unsigned)((outs + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, ctx.stream, part, (int)nblk, s, q, SG);
    KH_CHECK(hipGetLastError());
    return;
  }
  // Column runs set the rate of these products (short runs at an 8 MB column
  // stride held them near 1 TB/s). FP32 operands are staged as FP32 (also on
  // FP64 MFMA), FP64 ones as FP64, in slabs as long as about 75 KB of LDS
  // allows (two blocks per CU): tall-skinny products of at most 64 columns
  // per operand take 96-row slabs unless both operands are FP64 (48 rows);
  // the Gram pair and wider products (up to 128 columns) take 32-row slabs
  // (64-row FP32 slabs measured slower for the Gram pair).
  using TLZ = typename std::conditional<sizeof(TZ) == 4, float, T>::type;
  using TLA = typename std::conditional<sizeof(TA) == 4, float, T>::type;
  constexpr int kStage = (int)(sizeof(TLZ) + sizeof(TLA));  // bytes per row and column pair
  constexpr int kWideSlab = kStage <= 12 ? 96 : 48;
  constexpr int kFullSlab = kG2Slab;
  const bool wide = mode == 1 && s <= 64 && q <= 64;
  const int slab = wide ? kWideSlab : kFullSlab;
  // about 512 blocks (1024 measured no faster for the FP64 tall-skinny
  // products), at most about 4096 rows (FP32 accumulation length) each,
  // unless the partials would not fit the context's reduction buffer
  const int64_t maxblk = std::max<int64_t>(
      1, (int64_t)(ctx.part_capacity() * 8 / sizeof(T)) / ((int64_t)ntiles * 256));
  int64_t rpb = std::max<int64_t>((len + 511) / 512, 1);
  rpb = std::min<int64_t>(rpb, 4096);
  rpb = std::max<int64_t>(rpb, (len + maxblk - 1) / maxblk);
  rpb = (rpb + slab - 1) / slab * slab;
  const int64_t nblk64 = (len + rpb - 1) / rpb;
  const std::size_t need = (std::size_t)nblk64 * ntiles * 256;  // elements
  T *part = reinterpret_cast<T *>(ctx.partials((need * sizeof(T) + 7) / 8));
  {
    TraceScope tr(ctx, (mode == 0 && sizeof(T) == 4) ? kTraceGram : kTraceNone,
                  2.0 * (double)len * s * q * (mode == 0 ? 2.0 : 1.0));
    if (wide)
      hipLaunchKernelGGL((k_gram2_partial<T, kWideSlab, 64, TZ, TA, TLZ, TLA>),
                         dim3((unsigned)nblk64), dim3(kG2Threads), 0, ctx.stream, len, rpb, Z,
                         ldz, s, AZ, ldaz, q, mode, part);
    else
      hipLaunchKernelGGL((k_gram2_partial<T, kFullSlab, kG2MaxQ, TZ, TA, TLZ, TLA>),
                         dim3((unsigned)nblk64), dim3(kG2Threads), 0, ctx.stream, len, rpb, Z,
                         ldz, s, AZ, ldaz, q, mode, part);
  }
  if (mode == 0) {
    const int64_t outs = 2 * (int64_t)q * q * 64;  // one wave per output
    hipLaunchKernelGGL(k_gram2_finish<T>, dim3((unsigned)((outs + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, ctx.stream, part, (int)nblk64, q, SG);
  } else {
    const int64_t outs = (int64_t)s * q * 64;  // one wave per output
    hipLaunchKernelGGL(k_tn2_finish<T>, dim3((unsigned)((outs + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, ctx.stream, part, (int)nblk64, s, q, SG);
  }
  KH_CHECK(hipGetLastError());
}

void gram2_mfma(Context &ctx, const float *Z, int64_t ldz, const float *AZ, int64_t ldaz, int q,
                int64_t len, double *SG) {
  launch_gram2<float>(ctx, Z, ldz, q, AZ, ldaz, q, len, SG, 0);
}

void gram2_mfma(Context &ctx, const double *Z, int64_t ldz, const double *AZ, int64_t ldaz,
                int q, int64_t len, double *SG) {
  launch_gram2<double>(ctx, Z, ldz, q, AZ, ldaz, q, len, SG, 0);
}

void gemm_tn(Context &ctx, Prec pv, const void *V, int64_t ldv, int s, Prec pw,
             const void *W, int64_t ldw, int q, int64_t len, double *H) {
  if (s <= 0 || q <= 0)
    return;
  if (q <= 4) {  // narrow: column by column (no 64-wide tile padding)
    const std::size_t ew = pw == Prec::F64 ? 8 : 4;
    for (int c = 0; c < q; ++c)
      gemv_t(ctx, pv, V, ldv, s, pw, static_cast<const char *>(W) + (std::size_t)c * ldw * ew,
             len, H + (std::size_t)c * s);
    return;
  }
  TraceScope tr(ctx, kTraceOther,
                (double)len * (s * (pv == Prec::F64 ? 8 : 4) + q * (pw == Prec::F64 ? 8 : 4)));
  // up to 128 columns a side on FP64 MFMA, FP32 operands converted exactly
  // (every product exact, FP64 sums as on the VALU path below)
  if (s <= kG2MaxQ && q <= kG2MaxQ) {
    if (pv == Prec::F64 && pw == Prec::F64)
      launch_gram2<double>(ctx, (const double *)V, ldv, s, (const double *)W, ldw, q, len, H, 1);
    else if (pv == Prec::F32 && pw == Prec::F32)
      launch_gram2<double>(ctx, (const float *)V, ldv, s, (const float *)W, ldw, q, len, H, 1);
    else if (pv == Prec::F64)
      launch_gram2<double>(ctx, (const double *)V, ldv, s, (const float *)W, ldw, q, len, H, 1);
    else
      launch_gram2<double>(ctx, (const float *)V, ldv, s, (const double *)W, ldw, q, len, H, 1);
  } else if (pv == Prec::F64 && pw == Prec::F64)
    launch_gemm_tn(ctx, (const double *)V, ldv, s, (const double *)W, ldw, q, len, H);
  else if (pv == Prec::F32 && pw == Prec::F32)
    launch_gemm_tn(ctx, (const float *)V, ldv, s, (const float *)W, ldw, q, len, H);
  else if (pv == Prec::F64)
    launch_gemm_tn(ctx, (const double *)V, ldv, s, (const float *)W, ldw, q, len, H);
  else
    launch_gemm_tn(ctx, (const float *)V, ldv, s, (const double *)W, ldw, q, len, H);
}

template <class TV, class TW>
static void launch_gemm_nn_sub(Context &ctx, const TV *V, int64_t ldv, int s,
                               const double *H, int q, TW *W, int64_t ldw, int64_t len,
                               double *norms) {
  if (q <= 4) {  // narrow: column by column
    for (int c = 0; c < q; ++c)
      launch_gemv_n_sub(ctx, V, ldv, s, H + (std::size_t)c * s, W + (int64_t)c * ldw, len,
                        norms ? norms + c : nullptr);
    return;
  }
  if (!try_nn_mfma<TV, TW, true>(ctx, V, ldv, s, H, q, W, ldw, len)) {
    dim3 grid(grid_for(len, 2), (q + kQT - 1) / kQT);
    hipLaunchKernelGGL((k_ts_nn<TV, TW, true>), grid, dim3(kBlock), 0, ctx.stream, len, V, ldv,
                       s, H, q, W, ldw);
    KH_CHECK(hipGetLastError());
  }
  if (norms) {
    const int nb = reduce_blocks(len);
    double *part = ctx.partials((std::size_t)nb * q);
    hipLaunchKernelGGL(k_col_sq<TW>, dim3(nb, q), dim3(kBlock), 0, ctx.stream, len,
                       (const TW *)W, ldw, part, nb);
    hipLaunchKernelGGL(k_finish, dim3(q), dim3(kBlock), 0, ctx.stream, part, nb, norms, 1);
    KH_CHECK(hipGetLastError());
  }
}

void gemm_nn_sub(Context &ctx, Prec pv, const void *V, int64_t ldv, int s,
                 const double *H, int q, Prec pw, void *W, int64_t ldw, int64_t len,
                 double *norms) {
  if (q <= 0)
    return;
  if (pv == Prec::F64 && pw == Prec::F64)
    launch_gemm_nn_sub(ctx, (const double *)V, ldv, s, H, q, (double *)W, ldw, len, norms);
  else if (pv == Prec::F32 && pw == Prec::F32)
    launch_gemm_nn_sub(ctx, (const float *)V, ldv, s, H, q, (float *)W, ldw, len, norms);
  else if (pv == Prec::F64)
    launch_gemm_nn_sub(ctx, (const double *)V, ldv, s, H, q, (float *)W, ldw, len, norms);
  else
    launch_gemm_nn_sub(ctx, (const float *)V, ldv, s, H, q, (double *)W, ldw, len, norms);
}

void col_norms(Context &ctx, Prec p, const void *W, int64_t ldw, int q, int64_t len,
               double *out) {
  if (q <= 0)
    return;
  const int nb = reduce_blocks(len);
  double *part = ctx.partials((std::size_t)nb * q);
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_col_sq<double>, dim3(nb, q), dim3(kBlock), 0, ctx.stream, len,
                       (const double *)W, ldw, part, nb);
  else
    hipLaunchKernelGGL(k_col_sq<float>, dim3(nb, q), dim3(kBlock), 0, ctx.stream, len,
                       (const float *)W, ldw, part, nb);
  hipLaunchKernelGGL(k_finish, dim3(q), dim3(kBlock), 0, ctx.stream, part, nb, out, 1);
  KH_CHECK(hipGetLastError());
}

void scale_cols(Context &ctx, Prec p, int64_t len, int q, void *W, int64_t ldw,
                const double *scale, bool skip_zero) {
  if (q <= 0 || len <= 0)
    return;
  dim3 grid(grid_for(len, 4), q);
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_scale_cols<double>, grid, dim3(kBlock), 0, ctx.stream, len,
                       (double *)W, ldw, scale, skip_zero ? 1 : 0);
  else
    hipLaunchKernelGGL(k_scale_cols<float>, grid, dim3(kBlock), 0, ctx.stream, len,
                       (float *)W, ldw, scale, skip_zero ? 1 : 0);
  KH_CHECK(hipGetLastError());
}

void sub_cols(Context &ctx, Prec p, int64_t len, int q, const void *X, int64_t ldx,
              const void *Y, int64_t ldy, void *OUT, int64_t ldo, const int *mask) {
  if (q <= 0 || len <= 0)
    return;
  dim3 grid(grid_for(len, 4), q);
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_sub_cols<double>, grid, dim3(kBlock), 0, ctx.stream, len,
                       (const double *)X, ldx, (const double *)Y, ldy, (double *)OUT, ldo,
                       mask);
  else
    hipLaunchKernelGGL(k_sub_cols<float>, grid, dim3(kBlock), 0, ctx.stream, len,
                       (const float *)X, ldx, (const float *)Y, ldy, (float *)OUT, ldo,
                       mask);
  KH_CHECK(hipGetLastError());
}

void axpby(Context &ctx, Prec p, int64_t len, double alpha, const void *x, double beta,
           void *y) {
  if (len <= 0)
    return;
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_axpby<double>, dim3(grid_for(len)), dim3(kBlock), 0, ctx.stream,
                       len, alpha, (const double *)x, beta, (double *)y);
  else
    hipLaunchKernelGGL(k_axpby<float>, dim3(grid_for(len)), dim3(kBlock), 0, ctx.stream,
                       len, alpha, (const float *)x, beta, (float *)y);
  KH_CHECK(hipGetLastError());
}

void scale_dev(Context &ctx, Prec p, int64_t len, const double *num, bool reciprocal,
               void *y) {
  if (len <= 0)
    return;
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_scale_dev<double>, dim3(grid_for(len)), dim3(kBlock), 0,
                       ctx.stream, len, num, reciprocal ? 1 : 0, (double *)y);
  else
    hipLaunchKernelGGL(k_scale_dev<float>, dim3(grid_for(len)), dim3(kBlock), 0,
                       ctx.stream, len, num, reciprocal ? 1 : 0, (float *)y);
  KH_CHECK(hipGetLastError());
}

void scale_copy_dev(Context &ctx, Prec p, int64_t len, const double *den, const void *x,
                    void *y) {
  if (len <= 0)
    return;
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_scale_copy_dev<double>, dim3(grid_for(len)), dim3(kBlock), 0,
                       ctx.stream, len, den, (const double *)x, (double *)y);
  else
    hipLaunchKernelGGL(k_scale_copy_dev<float>, dim3(grid_for(len)), dim3(kBlock), 0,
                       ctx.stream, len, den, (const float *)x, (float *)y);
  KH_CHECK(hipGetLastError());
}

void resid_block(Context &ctx, Prec p, int64_t len, int q, const void *AX, const void *X,
                 int64_t ld, const double *neg_theta, void *R) {
  if (len <= 0 || q <= 0)
    return;
  const dim3 grid(grid_for(len, 4), q);
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_resid_block<double>, grid, dim3(kBlock), 0, ctx.stream, len,
                       (const double *)AX, (const double *)X, neg_theta, (double *)R, ld);
  else
    hipLaunchKernelGGL(k_resid_block<float>, grid, dim3(kBlock), 0, ctx.stream, len,
                       (const float *)AX, (const float *)X, neg_theta, (float *)R, ld);
  KH_CHECK(hipGetLastError());
}

void copy(Context &ctx, Prec p, int64_t len, const void *x, void *y) {
  if (len <= 0 || x == y)
    return;
  KH_CHECK(hipMemcpyAsync(y, x, (std::size_t)len * (p == Prec::F64 ? 8 : 4),
                          hipMemcpyDeviceToDevice, ctx.stream));
}

void convert(Context &ctx, Prec from, const void *x, Prec to, void *y, int64_t len) {
  if (len <= 0)
    return;
  if (from == to) {
    copy(ctx, from, len, x, y);
    return;
  }
  if (from == Prec::F64)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(len)), dim3(kBlock), 0,
                       ctx.stream, len, (const double *)x, (float *)y);
  else
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(len)), dim3(kBlock), 0,
                       ctx.stream, len, (const float *)x, (double *)y);
  KH_CHECK(hipGetLastError());
}

void fill_zero(Context &ctx, Prec p, int64_t len, void *y) {
  if (len <= 0)
    return;
  axpby(ctx, p, len, 0.0, nullptr, 0.0, y);
}

void diag_solve(Context &ctx, Prec p, int64_t len, const double *d, const void *r,
                void *y) {
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_diag_solve<double>, dim3(grid_for(len)), dim3(kBlock), 0,
                       ctx.stream, len, d, (const double *)r, (double *)y);
  else
    hipLaunchKernelGGL(k_diag_solve<float>, dim3(grid_for(len)), dim3(kBlock), 0,
                       ctx.stream, len, d, (const float *)r, (float *)y);
  KH_CHECK(hipGetLastError());
}

// --------------------------------------------------------------------- CG --
void CgBatch::ensure(int s_, int64_t m_, int64_t n_, Prec p_) {
  const std::size_t e = p_ == Prec::F64 ? 8 : 4;
  const std::size_t need_m = (std::size_t)s_ * (std::size_t)m_ * e;
  const std::size_t need_n = (std::size_t)s_ * (std::size_t)n_ * e;
  if (r.bytes < need_m)
    r = DBuf(need_m);
  if (z.bytes < need_m)
    z = DBuf(need_m);
  if (pv.bytes < need_m)
    pv = DBuf(need_m);
  if (Ap.bytes < need_m)
    Ap = DBuf(need_m);
  if (tmp_n.bytes < need_n)
    tmp_n = DBuf(need_n);
  if (st.bytes < (std::size_t)s_ * sizeof(CgState))
    st = DBuf((std::size_t)s_ * sizeof(CgState));
  if (act.bytes < (std::size_t)s_ * sizeof(int))
    act = DBuf((std::size_t)s_ * sizeof(int));
  s = s_;
  m = m_;
  n = n_;
  p = p_;
}

template <class T>
static void cg_init_t(Context &ctx, CgBatch &B, const T *rhs, T *w, const double *d,
                      double rel_tol, int64_t maxit) {
  const int nb = reduce_blocks(B.m);
  double *part = ctx.partials((std::size_t)nb * B.s * 2);
  dim3 grid(nb, B.s);
  T *z = d ? B.z.as<T>() : B.r.as<T>();
  hipLaunchKernelGGL(k_cg_init<T>, grid, dim3(kBlock), 0, ctx.stream, B.m, rhs, w,
                     B.r.as<T>(), z, B.pv.as<T>(), d, part, part + (std::size_t)nb * B.s,
                     nb);
  hipLaunchKernelGGL(k_cg_init_finish, dim3(B.s), dim3(kBlock), 0, ctx.stream, B.s, part,
                     part + (std::size_t)nb * B.s, nb, B.st.as<CgState>(), B.act.as<int>(),
                     rel_tol, maxit);
  KH_CHECK(hipGetLastError());
}

void cg_init(Context &ctx, CgBatch &B, const void *rhs, void *w, const double *diag,
             double rel_tol, int64_t maxit) {
  B.have_prec = diag != nullptr;
  if (B.p == Prec::F64)
    cg_init_t(ctx, B, (const double *)rhs, (double *)w, diag, rel_tol, maxit);
  else
    cg_init_t(ctx, B, (const float *)rhs, (float *)w, diag, rel_tol, maxit);
}

template <class T>
static void cg_post_t(Context &ctx, CgBatch &B, T *w, const double *d, int64_t maxit,
                      double bt, int64_t stall) {
  const int nb = reduce_blocks(B.m);
  // p.Ap partials, then r.z partials (the update reads the former while it
  // writes the latter)
  double *part = ctx.partials((std::size_t)nb * B.s * 2);
  double *part_rz = part + (std::size_t)nb * B.s;
  dim3 grid(nb, B.s);
  T *r = B.r.as<T>(), *p = B.pv.as<T>(), *Ap = B.Ap.as<T>();
  T *z = d ? B.z.as<T>() : r;
  int *act = B.act.as<int>();
  CgState *st = B.st.as<CgState>();
  hipLaunchKernelGGL(k_cg_pap<T>, grid, dim3(kBlock), 0, ctx.stream, B.m, p, Ap, act, part,
                     nb);
  hipLaunchKernelGGL(k_cg_update<T>, grid, dim3(kBlock), 0, ctx.stream, B.m, w, r, z, p, Ap,
                     d, st, act, part, part_rz, nb, bt);
  hipLaunchKernelGGL(k_cg_beta, dim3(B.s), dim3(kBlock), 0, ctx.stream, part_rz, nb, st, act,
                     maxit, stall);
  hipLaunchKernelGGL(k_cg_dir<T>, grid, dim3(kBlock), 0, ctx.stream, B.m, p, z, st, act);
  KH_CHECK(hipGetLastError());
}

void cg_after_products(Context &ctx, CgBatch &B, void *w, const double *diag,
                       int64_t maxit, double breakdown_tol, int64_t stall) {
  if (B.p == Prec::F64)
    cg_post_t(ctx, B, (double *)w, diag, maxit, breakdown_tol, stall);
  else
    cg_post_t(ctx, B, (float *)w, diag, maxit, breakdown_tol, stall);
}

// ------------------------------------------------- packed Golub-Kahan steps --
namespace {

static_assert(kPackRows == kRB, "a pack segment block is a V^T w block");

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_pk_spmv(int64_t rows, const int32_t *__restrict__ ptr, const int32_t *__restrict__ idx,
              const T *__restrict__ val, const T *__restrict__ x, T *__restrict__ y,
              const T *__restrict__ z, const PackSeg *__restrict__ segs,
              const int32_t *__restrict__ seg_of_blk, int jb) {
#pragma clang fp contract(off)
  const int64_t i = xcd_block() * kBlock + threadIdx.x;
  if (i >= rows)
    return;
  const PackSeg &S = segs[seg_of_blk[i / kPackRows]];
  if (!S.active)
    return;
  const T b = z ? (T)S.bnorm[jb] : (T)0;
  const T zi = z ? z[i] : (T)0;
  T out = (T)row_dot<T>(ptr[i], ptr[i + 1], idx, val, [=](int32_t j) { return x[j]; });
  if (z)
    out = out + (-b) * zi;
  y[i] = out;
}

// k_vh_sub with the coefficients of the block's segment
template <class TV, class TW>
__global__ void __launch_bounds__(kBlock)
    k_pk_vh_sub(int64_t len, const TV *__restrict__ V, int64_t ld, int s,
                const double *__restrict__ H, int64_t hstride, TW *__restrict__ w,
                double *__restrict__ part, const PackSeg *__restrict__ segs,
                const int32_t *__restrict__ seg_of_blk) {
  constexpr int CU = ColUnroll<TV>::value;
  extern __shared__ double hs[];  // [s]
  __shared__ double lds[kWaves];
  const int sg = seg_of_blk[blockIdx.x];
  if (!segs[sg].active)
    return;  // the whole block leaves: no barrier is reached
  const double *h = H + (int64_t)sg * hstride;
  for (int j = threadIdx.x; j < s; j += kBlock)
    hs[j] = h[j];
  __syncthreads();
  const int64_t r0 = (int64_t)blockIdx.x * kRB + threadIdx.x;
  int64_t rr[kR];
  double acc[kR];
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    const int64_t r = r0 + (int64_t)k * kBlock;
    rr[k] = r < len ? r : len - 1;
    acc[k] = 0.0;
  }
  const int s_main = s - s % CU;
  for (int c0 = 0; c0 < s_main; c0 += CU) {
    TV x[CU][kR];
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const TV *vc = V + (int64_t)(c0 + u) * ld;
#pragma unroll
      for (int k = 0; k < kR; ++k)
        x[u][k] = vc[rr[k]];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < CU; ++u) {
      const double hc = hs[c0 + u];
#pragma unroll
      for (int k = 0; k < kR; ++k)
        acc[k] += (double)x[u][k] * hc;
    }
  }
  for (int c = s_main; c < s; ++c) {
    const TV *vc = V + (int64_t)c * ld;
    const double hc = hs[c];
#pragma unroll
    for (int k = 0; k < kR; ++k)
      acc[k] += (double)vc[rr[k]] * hc;
  }
  TW wold[kR];
#pragma unroll
  for (int k = 0; k < kR; ++k)
    wold[k] = w[rr[k]];
  double nacc = 0.0;
#pragma unroll
  for (int k = 0; k < kR; ++k) {
    const int64_t r = r0 + (int64_t)k * kBlock;
    const TW nw = (TW)((double)wold[k] - acc[k]);
    if (r < len) {
      w[r] = nw;
      nacc += (double)nw * (double)nw;
    }
  }
  if (part) {
    nacc = block_sum(nacc, lds);
    if (threadIdx.x == 0)
      part[blockIdx.x] = nacc;
  }
}

// k_finish over one segment's blocks: block (c, seg) sums part[c * nb +
// blk0 .. blk0 + nblk) in k_finish's order. mode 0: H[seg * hs + c];
// mode 1 (c = 0): beta = sqrt(max(0, sum)) into the segment's bnorm[jb].
__global__ void __launch_bounds__(kBlock)
    k_pk_finish(const double *__restrict__ part, int64_t nb, const PackSeg *__restrict__ segs,
                int side, double *__restrict__ H, int64_t hs, int mode, int jb) {
  __shared__ double lds[kWaves];
  const int c = blockIdx.x, sg = blockIdx.y;
  const PackSeg &S = segs[sg];
  if (!S.active)
    return;
  const double *pc = part + (int64_t)c * nb + S.blk0[side];
  const int64_t n = S.nblk[side];
  double acc = 0.0;
  int64_t i = threadIdx.x;
  for (; i + 7 * kBlock < n; i += 8 * kBlock) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      v[u] = pc[i + u * kBlock];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      acc += v[u];
  }
  for (; i < n; i += kBlock)
    acc += pc[i];
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) {
    if (mode == 0)
      H[(int64_t)sg * hs + c] = acc;
    else
      S.bnorm[jb] = sqrt(fmax(0.0, acc));
  }
}

template <class T>
__global__ void __launch_bounds__(kBlock)
    k_pk_scale_copy(int64_t len, const PackSeg *__restrict__ segs,
                    const int32_t *__restrict__ seg_of_blk, int j, const T *__restrict__ x,
                    T *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= len)
    return;
  const PackSeg &S = segs[seg_of_blk[i / kPackRows]];
  if (!S.active)
    return;
  const T sc = (T)(1.0 / S.bnorm[j]);
  y[i] = x[i] * sc;
}

}  // namespace

void pack_spmv(Context &ctx, const DeviceCsr &A, bool trans, Prec p, const PackLayout &L,
               const void *x, void *y, const void *z, int jb) {
  const int side = trans ? 1 : 0;
  const int64_t rows = trans ? A.n : A.m;
  if (rows != L.nblk[side] * kPackRows)
    throw std::invalid_argument("pack_spmv: matrix rows do not match the pack layout");
  const int32_t *ptr = trans ? A.cp.as<int32_t>() : A.rp.as<int32_t>();
  const int32_t *idx = trans ? A.ri.as<int32_t>() : A.ci.as<int32_t>();
  const double e = p == Prec::F64 ? 8.0 : 4.0;
  TraceScope tr(ctx, kTraceSpmv,
                (e + 4.0) * A.nnz + 4.0 * (rows + 1) + e * (trans ? A.m : A.n) +
                    e * rows * (z ? 2.0 : 1.0));
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_pk_spmv<double>, dim3(grid_rows(rows)), dim3(kBlock), 0, ctx.stream,
                       rows, ptr, idx, (trans ? A.vt : A.va).as<double>(), (const double *)x,
                       (double *)y, (const double *)z, L.segs, L.seg_of_blk[side], jb);
  else
    hipLaunchKernelGGL(k_pk_spmv<float>, dim3(grid_rows(rows)), dim3(kBlock), 0, ctx.stream,
                       rows, ptr, idx, (trans ? A.vt32 : A.va32).as<float>(), (const float *)x,
                       (float *)y, (const float *)z, L.segs, L.seg_of_blk[side], jb);
  KH_CHECK(hipGetLastError());
}

void pack_vtw(Context &ctx, Prec p, const PackLayout &L, int side, const void *V, int64_t ld,
              int s, const void *w, double *part) {
  const int64_t len = L.nblk[side] * kPackRows;
  const int nb = (int)L.nblk[side];
  if (s <= 0)
    return;
  if (s > 512)
    throw std::invalid_argument("pack_vtw: at most 512 columns");
  if ((std::size_t)nb * s > ctx.part_capacity())
    throw std::length_error("pack_vtw: reduction partial buffer too small for this pack");
  TraceScope tr(ctx, p == Prec::F32 ? kTraceVtW : kTraceNone,
                (double)len * s * (p == Prec::F64 ? 8 : 4) + (double)len * (p == Prec::F64 ? 8 : 4) +
                    8.0 * nb * s,
                len, s, /*stamped=*/true);
  if (p == Prec::F64)
    launch_vtw_partial(ctx, (const double *)V, ld, s, (const double *)w, len, part, nb, tr.start(),
                       tr.stop());
  else
    launch_vtw_partial(ctx, (const float *)V, ld, s, (const float *)w, len, part, nb, tr.start(),
                       tr.stop());
}

void pack_finish(Context &ctx, const PackLayout &L, int side, const double *part, int s,
                 double *H, int64_t hs) {
  if (s <= 0 || L.nseg <= 0)
    return;
  hipLaunchKernelGGL(k_pk_finish, dim3((unsigned)s, (unsigned)L.nseg), dim3(kBlock), 0,
                     ctx.stream, part, L.nblk[side], L.segs, side, H, hs, 0, 0);
  KH_CHECK(hipGetLastError());
}

void pack_vh_sub(Context &ctx, Prec p, const PackLayout &L, int side, const void *V, int64_t ld,
                 int s, const double *H, int64_t hs, void *w, double *norm_part) {
  const int64_t len = L.nblk[side] * kPackRows;
  const std::size_t lds = (std::size_t)std::max(s, 1) * sizeof(double);
  TraceScope tr(ctx, kTraceVhSub,
                (double)len * s * (p == Prec::F64 ? 8 : 4) + 2.0 * len * (p == Prec::F64 ? 8 : 4) +
                    8.0 * s);
  if (p == Prec::F64)
    hipLaunchKernelGGL((k_pk_vh_sub<double, double>), dim3((unsigned)L.nblk[side]), dim3(kBlock),
                       lds, ctx.stream, len, (const double *)V, ld, s, H, hs, (double *)w,
                       norm_part, L.segs, L.seg_of_blk[side]);
  else
    hipLaunchKernelGGL((k_pk_vh_sub<float, float>), dim3((unsigned)L.nblk[side]), dim3(kBlock),
                       lds, ctx.stream, len, (const float *)V, ld, s, H, hs, (float *)w,
                       norm_part, L.segs, L.seg_of_blk[side]);
  KH_CHECK(hipGetLastError());
}

void pack_norm(Context &ctx, const PackLayout &L, int side, const double *norm_part, int j) {
  if (L.nseg <= 0)
    return;
  hipLaunchKernelGGL(k_pk_finish, dim3(1, (unsigned)L.nseg), dim3(kBlock), 0, ctx.stream,
                     norm_part, L.nblk[side], L.segs, side, (double *)nullptr, (int64_t)0, 1, j);
  KH_CHECK(hipGetLastError());
}

void pack_scale_copy(Context &ctx, Prec p, const PackLayout &L, int side, int j, const void *w,
                     void *dst) {
  const int64_t len = L.nblk[side] * kPackRows;
  if (p == Prec::F64)
    hipLaunchKernelGGL(k_pk_scale_copy<double>, dim3(grid_rows(len)), dim3(kBlock), 0,
                       ctx.stream, len, L.segs, L.seg_of_blk[side], j, (const double *)w,
                       (double *)dst);
  else
    hipLaunchKernelGGL(k_pk_scale_copy<float>, dim3(grid_rows(len)), dim3(kBlock), 0,
                       ctx.stream, len, L.segs, L.seg_of_blk[side], j, (const float *)w,
                       (float *)dst);
  KH_CHECK(hipGetLastError());
}

}  // namespace kh

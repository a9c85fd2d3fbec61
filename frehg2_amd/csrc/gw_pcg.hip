// gw_pcg.hip — structured-grid preconditioned CG on gfx950.
//
// Replaces PETSc KSPCG (src/core/LinearSystem.cpp:149, :320) for the Richards
// head system.  Convergence test: ||b - A x||_2 <= max(rtol ||b||_2, atol) on
// the recursive residual, confirmed on the true residual (a restart if the
// two disagree).  PETSc's default tests the preconditioned norm; the true
// residual is the preconditioner-independent form of the same tolerance.
//
// Work mapping ("column waves"): one 64-lane wavefront owns one z-column, lane
// l holding rows k = l + 64 m (m < R = ceil(nz / 64)), so every load and store
// along k is coalesced, k +- 1 neighbours come from lane shuffles and the
// z-line (tridiagonal) preconditioner is two wave-level affine scans in
// registers (DPP lane moves).  Kernels without lateral neighbours split the columns into 8
// contiguous ranges, one per XCD.  Stencil kernels use the j-march mapping: a
// wave owns one i and a slab of S consecutive j, walks it in j and keeps the
// j - 1 / j / j + 1 columns in registers, so each column is fetched once; the
// i +- 1 columns belong to the waves next to it on the same XCD and come from
// that XCD's L2.  (Fetching the j +- 1 columns of a row-ordered sweep from
// cache instead costs ~1.7x in time at 512 x 512 x 100: three rows of the
// operand and stencil arrays overflow the 4 MB L2 of an XCD.)
//
// One iteration is two kernels:
//   dir:    p = z + beta p (beta from the previous update's partial sums,
//           recomputed at the 4 lateral neighbours), q = A p, partial p.q
//   update: alpha = rz / p.q, r -= alpha q, z = M^-1 r, partial r.z and
//           r.r; x takes alpha p one iteration late, two terms at a time
//           (k_flush_x adds a last deferred term)
// Each kernel reduces the other's per-block partials redundantly in every
// block, in a fixed order, so the scalars are deterministic and no host
// round trip sits inside the iteration.  Block 0 mirrors the scalars to a
// state struct the host polls every few iterations.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "../../include/frehg_gw.h"
#include "gw_pcg.h"

namespace frehg {
namespace {

struct PcgState {
  double rz[2];
  double rr;
  double alpha;
  double alphaH[2];  // alpha of iteration n in slot n & 1 (deferred x updates)
  int it;
  int done;
  // read by the direction kernels: the recursion stops once rr <= tol2, or
  // at iteration `limit` (the iteration budget of a run of launches)
  double tol2;
  int limit;
  // device-side control (Pcg::enqueueDevice): the solve's scalars and
  // decisions live here, so no host round trip sits inside the solve
  int skip;         // the current round / cycle is over before it starts (no-ops)
  int stop;         // the solve is over: converged, diverged or out of iterations
  int converged;
  int fallback;     // mixed: a cycle failed to halve the FP64 residual
  int keepNew;      // ... and its output iterate was still the better one
  int xInB;         // mixed: the current iterate is in the second x buffer
  int total;        // iterations so far
  int refinements;
  int stalls;       // FP64 rounds that made no iteration
  int maxInner;     // most iterations of one round / cycle
  int rounds;       // rounds / cycles that ran
  int firstInner;   // iterations of the first one
  int laterInner;   // most iterations of a later one
  int gate;         // 1: the solve did not converge inside the budget
  double tolF2;     // final tolerance, squared
  double rrTrue, bbTrue, rrPrev, s;
};

constexpr int kBlock = 256;
constexpr int kWpb = kBlock / kWave;
constexpr int kChunk = 32;   // rows per LDS tile in the factorization (16: +20 % time)
constexpr int kPoll = 8;      // iterations between host polls (until a rate is known)
constexpr int kPollMax = 32;  // the most iterations launched between two polls

#define HIP_OK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e_) + " at " + \
                               __FILE__ + ":" + std::to_string(__LINE__));            \
  } while (0)

__device__ __forceinline__ void column_loop(int ncol, int64_t& first, int64_t& end,
                                            int64_t& stride) {
  const int xcd = blockIdx.x % kXcds;
  const int bpx = gridDim.x / kXcds;
  const int local = blockIdx.x / kXcds;
  const int64_t c0 = static_cast<int64_t>(xcd) * ncol / kXcds;
  end = static_cast<int64_t>(xcd + 1) * ncol / kXcds;
  first = c0 + static_cast<int64_t>(local) * kWpb + threadIdx.x / kWave;
  stride = static_cast<int64_t>(bpx) * kWpb;
}

__device__ __forceinline__ int64_t column_base(const Grid& g, int64_t col) {
  const int j = static_cast<int>(col / g.nx) + 1;
  const int i = static_cast<int>(col % g.nx) + 1;
  return g.c3(j, i, 0);
}

// j-march mapping: wave -> (i, rows j0..j1).  Blocks are remapped so each
// XCD receives a contiguous range of waves (whole slabs when nx * kWpb
// divides evenly).  Returns false for waves past the last slab.
__device__ __forceinline__ bool march_span(const Grid& g, const March& mj, int& i, int& j0,
                                           int& j1) {
  const int64_t wg = xcd_block() * kWpb + threadIdx.x / kWave;
  i = static_cast<int>(wg % g.nx) + 1;
  const int64_t slab = wg / g.nx;
  j0 = static_cast<int>(slab) * mj.S + 1;
  j1 = min(g.ny, j0 + mj.S - 1);
  return slab < mj.nslab && j0 <= g.ny;
}

// j-march over pairs of columns (i, i + 1), i odd (nx even): wave -> (i, rows
// j0..j1); mj.nslab counts slabs of nx / 2 waves.
__device__ __forceinline__ bool march_span2(const Grid& g, const March& mj, int& i, int& j0,
                                            int& j1) {
  const int64_t wg = xcd_block() * kWpb + threadIdx.x / kWave;
  const int half = g.nx / 2;
  i = 2 * static_cast<int>(wg % half) + 1;
  const int64_t slab = wg / half;
  j0 = static_cast<int>(slab) * mj.S + 1;
  j1 = min(g.ny, j0 + mj.S - 1);
  return slab < mj.nslab && j0 <= g.ny;
}

// Row k of a column for a lane past the column's end: such a lane reads the
// last row (in bounds) and its caller zeroes the value.  All lanes' loads then
// issue in one block; a load under `if (k < nz)` compiles to a branch that
// waits for it before the next load issues.
__device__ __forceinline__ int row_in(int k, int nz) { return k < nz ? k : nz - 1; }

template <class T, int R>
__device__ __forceinline__ void load_col(const T* __restrict__ a, int64_t base, int lane, int nz,
                                         T (&v)[R]) {
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int k = lane + kWave * m;
    const T t = a[base + row_in(k, nz)];
    v[m] = k < nz ? t : T(0);
  }
}

// Partial sums live in parts[slot * pstride + block]; pstride is the largest
// grid of any kernel, so kernels with different grids never overlap.
// Redundant in-block reduction of the previous kernel's partial sums.  Every
// block repeats it, which measured faster in the solver than both
// alternatives tried: a one-block reduction kernel between the two (one more
// launch and kernel boundary per reduction) and a last-block-done reduction
// (its device-scope release per block writes back the XCD's L2).
template <int NV>
__device__ __forceinline__ void reduce_parts(const double* parts, int pstride, int nparts,
                                             const int (&slot)[NV], double (&out)[NV]) {
  int op[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    out[q] = 0.0;
    op[q] = 0;
  }
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] += parts[slot[q] * pstride + b];
  }
  block_reduce<NV>(out, op);
}

template <int NV>
__device__ __forceinline__ void write_parts(double* parts, int pstride, const int (&slot)[NV],
                                            double (&v)[NV]) {
  int op[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) op[q] = 0;
  block_reduce<NV>(v, op);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) parts[slot[q] * pstride + blockIdx.x] = v[q];
  }
}

// Value of row k-1 / k+1 of a column held as v[m] at row lane + 64 m.
template <class T, int R>
__device__ __forceinline__ void shift_rows(const T (&v)[R], int lane, T (&below)[R],
                                           T (&above)[R]) {
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const T up = __shfl_up(v[m], 1, kWave);
    const T upSeg = (m > 0) ? __shfl(v[m > 0 ? m - 1 : 0], kWave - 1, kWave) : T(0);
    above[m] = lane > 0 ? up : upSeg;  // row k - 1
    const T dn = __shfl_down(v[m], 1, kWave);
    const T dnSeg = (m + 1 < R) ? __shfl(v[m + 1 < R ? m + 1 : m], 0, kWave) : T(0);
    below[m] = lane < kWave - 1 ? dn : dnSeg;  // row k + 1
  }
}

// The two recurrences of the column solve with LDL^T factors held in
// registers (iv = 1 / delta, wv = Azp / delta; zero past row nz - 1).
template <class T, int R>
__device__ __forceinline__ void line_solve(int lane, int nz, const T (&iv)[R], const T (&wv)[R],
                                           const T (&r)[R], T (&z)[R]) {
  T wPrev[R], dummy[R];
  shift_rows<T, R>(wv, lane, dummy, wPrev);
  T y[R];
  T carry = T(0);
#pragma unroll
  for (int m = 0; m < R; ++m) y[m] = scan_affine_up(r[m], -wPrev[m], carry, lane);
  carry = T(0);
#pragma unroll
  for (int m = R - 1; m >= 0; --m) {
    const int k = lane + kWave * m;
    z[m] = scan_affine_down(k < nz ? y[m] * iv[m] : T(0), k < nz ? -wv[m] : T(0), carry, lane);
  }
}

// z = M^-1 r for one column: point Jacobi (LINE = false) or the exact
// tridiagonal solve of the column block (LINE = true) with the LDL^T factors
// invd_k = 1 / delta_k and w_k = Azp_k / delta_k:
//   forward  y_k = r_k - w_{k-1} y_{k-1}
//   backward z_k = y_k invd_k - w_k z_{k+1}
// Both recurrences are affine, y_k = a_k + b_k y_{k-1}, and run as DPP
// scans over the 64 lanes of each row segment.
template <class T, int R, bool LINE>
__device__ __forceinline__ void apply_precond(int lane, int nz, int64_t base, const T* __restrict__ invd,
                                              const T* __restrict__ w, const T (&r)[R], T (&z)[R]) {
  T iv[R];
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int k = lane + kWave * m;
    const T t = invd[base + row_in(k, nz)];
    iv[m] = k < nz ? t : T(0);
  }
  if constexpr (!LINE) {
#pragma unroll
    for (int m = 0; m < R; ++m) z[m] = r[m] * iv[m];
    return;
  } else {
    T wv[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int k = lane + kWave * m;
      const T t = w[base + row_in(k, nz)];
      wv[m] = k < nz ? t : T(0);
    }
    line_solve<T, R>(lane, nz, iv, wv, r, z);
  }
}

// ---------------------------------------------------------------------------
// Factorization of the column blocks (once per solve).  One wave owns 64
// consecutive columns; rows are staged through LDS in 32-row tiles so the
// global loads and stores stay coalesced while each lane runs its column's
// sequential LDL^T recurrence  delta_k = Ad_k - Azp_{k-1} w_{k-1}.
// ---------------------------------------------------------------------------
template <class T, bool LINE>
__global__ __launch_bounds__(64) void k_factor(Grid g, const T* __restrict__ Ad,
                                               const T* __restrict__ Azp, T* __restrict__ invd,
                                               T* __restrict__ w) {
  __shared__ int64_t sbase[kWave];
  __shared__ T sA[kWave][kChunk + 1];
  __shared__ T sE[kWave][kChunk + 1];
  const int lane = threadIdx.x;
  const int ncol = g.columns();
  const int64_t col0 = static_cast<int64_t>(blockIdx.x) * kWave;
  const int64_t mycol = col0 + lane;
  sbase[lane] = mycol < ncol ? column_base(g, mycol) : -1;
  __syncthreads();
  // The next tile is fetched into registers while the current one is
  // factored, so its loads are in flight during the sequential recurrence.
  T ra[kChunk], re[kChunk];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int m = 0; m < kChunk; ++m) {
      const int e = m * kWave + lane;
      const int cc = e / kChunk, kk = e % kChunk, k = k0 + kk;
      ra[m] = T(0);
      re[m] = T(0);
      if (sbase[cc] >= 0 && k < g.nz) {
        ra[m] = Ad[sbase[cc] + k];
        if (LINE) re[m] = Azp[sbase[cc] + k];
      }
    }
  };
  fetch(0);
  T ePrev = T(0), wPrev = T(0);
  for (int k0 = 0; k0 < g.nz; k0 += kChunk) {
#pragma unroll
    for (int m = 0; m < kChunk; ++m) {
      const int e = m * kWave + lane;
      const int cc = e / kChunk, kk = e % kChunk, k = k0 + kk;
      if (sbase[cc] >= 0 && k < g.nz) {
        sA[cc][kk] = ra[m];
        if (LINE) sE[cc][kk] = re[m];
      }
    }
    __syncthreads();
    if (k0 + kChunk < g.nz) fetch(k0 + kChunk);
    if (mycol < ncol) {
      for (int kk = 0; kk < kChunk && k0 + kk < g.nz; ++kk) {
        if (LINE) {
          const T e = sE[lane][kk];
          const T d = sA[lane][kk] - ePrev * wPrev;
          const T inv = T(1) / d;
          const T wv = e * inv;
          sA[lane][kk] = inv;
          sE[lane][kk] = wv;
          ePrev = e;
          wPrev = wv;
        } else {
          sA[lane][kk] = T(1) / sA[lane][kk];
        }
      }
    }
    __syncthreads();
#pragma unroll 4
    for (int m = 0; m < kChunk; ++m) {
      const int e = m * kWave + lane;
      const int cc = e / kChunk, kk = e % kChunk, k = k0 + kk;
      if (sbase[cc] >= 0 && k < g.nz) {
        invd[sbase[cc] + k] = sA[cc][kk];
        if (LINE) w[sbase[cc] + k] = sE[cc][kk];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Direct solve of laterally decoupled systems (no x or y legs: a batch of
// independent tridiagonal columns).  One wave owns one column's active rows
// k0 .. nz - 1 (row k0 + lane + 64 m), so every load is coalesced and the
// column never leaves registers.
//
// Factor: the pivots delta_k = a_k - c_{k-1}^2 / delta_{k-1} (c_k = Azp_k,
// the leg to row k + 1) are the ratios p_k / p_{k-1} of the continuants
// p_k = a_k p_{k-1} - c_{k-1}^2 p_{k-2}, i.e. v_k = M_k v_{k-1} with
// M_k = [[a_k, -c_{k-1}^2], [1, 0]] and v_{-1} = (1, 0).  The prefix products
// M_k ... M_0 are an inclusive scan over the lanes of each row segment (four
// DPP steps inside rows of 16 lanes, then the vectors entering rows 1-3 from
// the composed matrices of lanes 15, 31 and 47, as the affine scans do),
// each product rescaled by a power of two (exact) so it stays in range; the
// segment's last vector carries into the next segment.  Then the LDL^T solve
// (line_solve), one FP64 refinement x += M^-1 (b - A x) (the scan's pivots
// can lose digits to cancellation on near-singular columns) and the
// residual partials.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void rescale2(double& a, double& b, double& c, double& d) {
  const double mx = fmax(fmax(fabs(a), fabs(b)), fmax(fabs(c), fabs(d)));
  if (!(mx > 0.0) || !isfinite(mx)) return;
  int e = 0;
  (void)frexp(mx, &e);
  a = ldexp(a, -e);
  b = ldexp(b, -e);
  c = ldexp(c, -e);
  d = ldexp(d, -e);
}

// One step of the 2 x 2 product scan inside rows of 16 lanes: lane l's
// matrix times the one S lanes before it (DPP row shift; lanes without such
// a neighbour in their row take the identity).
template <int S>
__device__ __forceinline__ void mat_step(double& m00, double& m01, double& m10, double& m11) {
  constexpr int ctrl = kDppRowShr + S;
  const double p00 = dpp_move<ctrl>(1.0, m00), p01 = dpp_move<ctrl>(0.0, m01);
  const double p10 = dpp_move<ctrl>(0.0, m10), p11 = dpp_move<ctrl>(1.0, m11);
  double n00 = m00 * p00 + m01 * p10, n01 = m00 * p01 + m01 * p11;
  double n10 = m10 * p00 + m11 * p10, n11 = m10 * p01 + m11 * p11;
  rescale2(n00, n01, n10, n11);
  m00 = n00;
  m01 = n01;
  m10 = n10;
  m11 = n11;
}

// x <- (lane l's matrix) x, the matrix read with readlane (wave-uniform).
__device__ __forceinline__ void mat_apply_lane(double m00, double m01, double m10, double m11,
                                               int l, double& x0, double& x1) {
  const double a = lane_value(m00, l), b = lane_value(m01, l);
  const double c = lane_value(m10, l), d = lane_value(m11, l);
  double y0 = a * x0 + b * x1, y1 = c * x0 + d * x1, z0 = 0.0, z1 = 0.0;
  rescale2(y0, y1, z0, z1);
  x0 = y0;
  x1 = y1;
}

template <int R>
__device__ __forceinline__ void continuant_factor(int lane, int L, const double (&a)[R],
                                                  const double (&c)[R], double (&iv)[R],
                                                  double (&wv)[R]) {
  double cUp[R], dummy[R];
  shift_rows<double, R>(c, lane, dummy, cUp);  // cUp[m]: the leg of row k - 1
  double v0 = 1.0, v1 = 0.0;                   // v_{k0 - 1}
  const int row = lane >> 4;
#pragma unroll
  for (int m = 0; m < R; ++m) {
    // rows past the column take the identity map (their pivots repeat the
    // last one and are never used)
    const bool in = lane + kWave * m < L;
    double m00 = in ? a[m] : 1.0, m01 = in ? -cUp[m] * cUp[m] : 0.0;
    double m10 = in ? 1.0 : 0.0, m11 = in ? 0.0 : 1.0;
    // products from the start of each 16-lane row, then the vectors entering
    // rows 1-3 from the composed matrices of lanes 15, 31 and 47
    mat_step<1>(m00, m01, m10, m11);
    mat_step<2>(m00, m01, m10, m11);
    mat_step<4>(m00, m01, m10, m11);
    mat_step<8>(m00, m01, m10, m11);
    double e0[4], e1[4];
    e0[0] = v0;
    e1[0] = v1;
#pragma unroll
    for (int r = 1; r < 4; ++r) {
      e0[r] = e0[r - 1];
      e1[r] = e1[r - 1];
      mat_apply_lane(m00, m01, m10, m11, 16 * r - 1, e0[r], e1[r]);
    }
    const double x0 = row == 0 ? e0[0] : row == 1 ? e0[1] : row == 2 ? e0[2] : e0[3];
    const double x1 = row == 0 ? e1[0] : row == 1 ? e1[1] : row == 2 ? e1[2] : e1[3];
    const double u0 = m00 * x0 + m01 * x1, u1 = m10 * x0 + m11 * x1;
    const double delta = u0 / u1;
    iv[m] = 1.0 / delta;
    wv[m] = c[m] * iv[m];
    v0 = e0[3];
    v1 = e1[3];
    mat_apply_lane(m00, m01, m10, m11, kWave - 1, v0, v1);
  }
}

// b - A x on a column in registers (tridiagonal: a, and c to the next row).
template <int R>
__device__ __forceinline__ void column_residual(int lane, int L, const double (&a)[R],
                                                const double (&c)[R], const double (&bv)[R],
                                                const double (&x)[R], double (&r)[R]) {
  double xDn[R], xUp[R], cDn[R], cUp[R];
  shift_rows<double, R>(x, lane, xDn, xUp);
  shift_rows<double, R>(c, lane, cDn, cUp);
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int kk = lane + kWave * m;
    r[m] = kk < L ? bv[m] - (a[m] * x[m] + c[m] * xDn[m] + cUp[m] * xUp[m]) : 0.0;
  }
}

template <int R>
__device__ __forceinline__ void column_direct(const Grid& g, const Stencil<double>& A,
                                              const double* __restrict__ b,
                                              double* __restrict__ x, int64_t base, int k0,
                                              int lane, KSum& rr, KSum& bb) {
  const int L = g.nz - k0;
  double a[R], c[R], bv[R];
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int kk = lane + kWave * m;
    const int64_t cc = base + k0 + (kk < L ? kk : L - 1);
    const double ta = A.Ad[cc], tc = A.Azp[cc], tb = b[cc];
    a[m] = kk < L ? ta : 1.0;  // rows past the column: identity, decoupled
    c[m] = kk < L - 1 ? tc : 0.0;
    bv[m] = kk < L ? tb : 0.0;
  }
  double iv[R], wv[R], xv[R], r[R], d[R];
  continuant_factor<R>(lane, L, a, c, iv, wv);
  line_solve<double, R>(lane, L, iv, wv, bv, xv);
  column_residual<R>(lane, L, a, c, bv, xv, r);
  line_solve<double, R>(lane, L, iv, wv, r, d);
#pragma unroll
  for (int m = 0; m < R; ++m) xv[m] += d[m];
  column_residual<R>(lane, L, a, c, bv, xv, r);
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int kk = lane + kWave * m;
    if (kk < L) {
      x[base + k0 + kk] = xv[m];
      rr.add(r[m] * r[m]);
      bb.add(bv[m] * bv[m]);
    }
  }
}

// column_direct<R> for a run-time R <= RMAX (R in 1 .. 8, or 16 for 513 ..
// 1024 rows): the branches stop at the deepest form the launch holds, so a
// table without deep columns does not take the R = 16 form's registers.
template <int RMAX>
__device__ __forceinline__ void column_direct_upto(int R, const Grid& g,
                                                   const Stencil<double>& A,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ x, int64_t base, int k0,
                                                   int lane, KSum& rr, KSum& bb) {
  if constexpr (RMAX == 16) {
    if (R > 8) column_direct<16>(g, A, b, x, base, k0, lane, rr, bb);
    else column_direct_upto<8>(R, g, A, b, x, base, k0, lane, rr, bb);
  } else if constexpr (RMAX > 1) {
    if (R == RMAX) column_direct<RMAX>(g, A, b, x, base, k0, lane, rr, bb);
    else column_direct_upto<RMAX - 1>(R, g, A, b, x, base, k0, lane, rr, bb);
  } else {
    column_direct<1>(g, A, b, x, base, k0, lane, rr, bb);
  }
}

// One launch over every column with active rows, deepest bucket first (the
// longest waves start first and the shallow ones fill in behind them): wave
// w of bucket b solves columns order[colStart + (w - waveStart) * cpw + q],
// q < cpw.  RMAX is the deepest bucket's R.  Partials rr (slot 1), bb (slot 3).
template <int RMAX>
__global__ __launch_bounds__(kBlock) void k_column_solve(Grid g, Stencil<double> A,
                                                         const double* __restrict__ b,
                                                         double* __restrict__ x,
                                                         const int* __restrict__ kt,
                                                         const int* __restrict__ order,
                                                         ColumnTable tab,
                                                         double* __restrict__ parts, int pstride) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wg = blockIdx.x * kWpb + threadIdx.x / kWave;
  KSum rr, bb;
  int bk = 0;
  while (bk < tab.buckets && wg >= tab.waveStart[bk + 1]) ++bk;
  if (bk < tab.buckets) {
    const int first = tab.colStart[bk] + (wg - tab.waveStart[bk]) * tab.cpw[bk];
    const int last = min(first + tab.cpw[bk], tab.colStart[bk + 1]);
    for (int q = first; q < last; ++q) {
      const int col = order[q];
      const int j = col / g.nx + 1, i = col % g.nx + 1;
      column_direct_upto<RMAX>(tab.R[bk], g, A, b, x, g.c3(j, i, 0), kt[g.c2(j, i)], lane, rr,
                               bb);
    }
  }
  double v[2] = {rr.value(), bb.value()};
  const int slots[2] = {1, 3};
  write_parts<2>(parts, pstride, slots, v);
}

// r = b - A x, z = M^-1 r; partials rz (slot 0), rr (1), bb (3).  j-march.
template <int R, bool LINE>
__global__ __launch_bounds__(kBlock) void k_residual64(Grid g, March mj, Stencil<double> A,
                                                       const double* __restrict__ b,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ r,
                                                       double* __restrict__ z,
                                                       const double* __restrict__ invd,
                                                       const double* __restrict__ w,
                                                       double* __restrict__ parts, int pstride,
                                                       const PcgState* __restrict__ ctl) {
  if (ctl && ctl->skip) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t sx = g.nz;
  KSum rz, rr, bb;
  int i, j0, j1;
  if (march_span(g, mj, i, j0, j1)) {
    double xPrev[R], xCur[R], xNext[R], ayPrev[R];
    load_col<double, R>(x, g.c3(j0 - 1, i, 0), lane, g.nz, xPrev);
    load_col<double, R>(x, g.c3(j0, i, 0), lane, g.nz, xCur);
    load_col<double, R>(A.Ayp, g.c3(j0 - 1, i, 0), lane, g.nz, ayPrev);
    for (int j = j0; j <= j1; ++j) {
      const int64_t base = g.c3(j, i, 0);
      load_col<double, R>(x, base + static_cast<int64_t>(g.nx2) * g.nz, lane, g.nz, xNext);
      double xdn[R], xup[R], az[R], azUp[R], ay[R], dummy[R], rv[R], zv[R];
      load_col<double, R>(A.Azp, base, lane, g.nz, az);
      load_col<double, R>(A.Ayp, base, lane, g.nz, ay);
      shift_rows<double, R>(xCur, lane, xdn, xup);
      shift_rows<double, R>(az, lane, dummy, azUp);
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int k = lane + kWave * m;
        rv[m] = 0.0;
        if (k < g.nz) {
          const int64_t c = base + k;
          const double q = A.Ad[c] * xCur[m] + A.Axp[c] * x[c + sx] + A.Axp[c - sx] * x[c - sx] +
                           ay[m] * xNext[m] + ayPrev[m] * xPrev[m] + az[m] * xdn[m] +
                           azUp[m] * xup[m];
          const double bv = b[c];
          rv[m] = bv - q;
          bb.add(bv * bv);
        }
      }
      apply_precond<double, R, LINE>(lane, g.nz, base, invd, w, rv, zv);
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int k = lane + kWave * m;
        if (k < g.nz) {
          r[base + k] = rv[m];
          z[base + k] = zv[m];
          rz.add(rv[m] * zv[m]);
          rr.add(rv[m] * rv[m]);
        }
        xPrev[m] = xCur[m];
        xCur[m] = xNext[m];
        ayPrev[m] = ay[m];
      }
    }
  }
  double v[3] = {rz.value(), rr.value(), bb.value()};
  const int slots[3] = {0, 1, 3};
  write_parts<3>(parts, pstride, slots, v);
}

// Search direction and operator product: p = z + beta p_old, q = A p, p.q.
// j-march: p at j - 1, j, j + 1 is formed once per column and kept in
// registers; the i +- 1 values are formed from z and p_old at the neighbour.
template <class T, int R, bool RESTART>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(Grid g, March mj, Stencil<T> A,
                                                    const T* __restrict__ z,
                                                    const T* __restrict__ pOld,
                                                    T* __restrict__ pNew, T* __restrict__ q,
                                                    double* __restrict__ parts, int pstride,
                                                    int nprev, PcgState* __restrict__ st, int n) {
  // Launches queued behind the one that ended the recursion (convergence or
  // the budget) are no-ops and leave the count alone.
  if (st->done) return;
  if (n >= st->limit) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->it = n;
    }
    return;
  }
  double red[2];
  const int rs[2] = {0, 1};
  reduce_parts<2>(parts, pstride, nprev, rs, red);
  const double rzv = red[0], rrv = red[1];
  if (rrv <= st->tol2) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->it = n;
      st->rr = rrv;
    }
    return;
  }
  const double betaD = RESTART ? 0.0 : rzv / st->rz[(n + 1) & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz[n & 1] = rzv;
    st->it = n;
    st->rr = rrv;
  }
  const T beta = static_cast<T>(betaD);
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t sx = g.nz, sy = static_cast<int64_t>(g.nx2) * g.nz;
  KSum pq;
  int i, j0, j1;
  if (march_span(g, mj, i, j0, j1)) {
    auto formp = [&](int64_t base, T (&pv)[R]) {
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int k = lane + kWave * m;
        const int64_t cc = base + row_in(k, g.nz);
        const T t = RESTART ? z[cc] : z[cc] + beta * pOld[cc];
        pv[m] = k < g.nz ? t : T(0);
      }
    };
    T pPrev[R], pCur[R], pNext[R], ayPrev[R];
    formp(g.c3(j0 - 1, i, 0), pPrev);
    formp(g.c3(j0, i, 0), pCur);
    load_col<T, R>(A.Ayp, g.c3(j0 - 1, i, 0), lane, g.nz, ayPrev);
    for (int j = j0; j <= j1; ++j) {
      const int64_t base = g.c3(j, i, 0);
      formp(base + sy, pNext);
      T pdn[R], pup[R], az[R], azUp[R], ay[R], dummy[R];
      load_col<T, R>(A.Azp, base, lane, g.nz, az);
      load_col<T, R>(A.Ayp, base, lane, g.nz, ay);
      shift_rows<T, R>(pCur, lane, pdn, pup);
      shift_rows<T, R>(az, lane, dummy, azUp);
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int k = lane + kWave * m;
        const int64_t c = base + row_in(k, g.nz);
        const T pxp = RESTART ? z[c + sx] : z[c + sx] + beta * pOld[c + sx];
        const T pxm = RESTART ? z[c - sx] : z[c - sx] + beta * pOld[c - sx];
        const T qv = A.Ad[c] * pCur[m] + A.Axp[c] * pxp + A.Axp[c - sx] * pxm + ay[m] * pNext[m] +
                     ayPrev[m] * pPrev[m] + az[m] * pdn[m] + azUp[m] * pup[m];
        if (k < g.nz) {
          pNew[c] = pCur[m];
          q[c] = qv;
          pq.add(static_cast<double>(pCur[m]) * static_cast<double>(qv));
        }
        pPrev[m] = pCur[m];
        pCur[m] = pNext[m];
        ayPrev[m] = ay[m];
      }
    }
  }
  double v[1] = {pq.value()};
  const int slots[1] = {2};
  write_parts<1>(parts, pstride, slots, v);
}

// alpha = rz / p.q; x += alpha p; r -= alpha q; z = M^-1 r; partials rz, rr.
// The x update is deferred by one iteration (XM): an even iteration of a
// sweep leaves x alone (XM = 0) and the odd one after it adds both terms,
// x = (x + alpha_prev p_prev) + alpha p (XM = 2), the same two roundings as
// two single updates.  p_prev is still intact then (the next direction
// kernel overwrites it), so x is read and written every other iteration.
template <class T, int R, bool LINE, int XM>
__global__ __launch_bounds__(kBlock) void k_pcg_update(Grid g, T* __restrict__ x,
                                                       const T* __restrict__ pPrev,
                                                       const T* __restrict__ p, T* __restrict__ r,
                                                       const T* __restrict__ q, T* __restrict__ z,
                                                       const T* __restrict__ invd,
                                                       const T* __restrict__ w,
                                                       double* __restrict__ parts, int pstride,
                                                       int nprev, PcgState* __restrict__ st,
                                                       int n) {
  if (st->done) return;
  double red[1];
  const int rs[1] = {2};
  reduce_parts<1>(parts, pstride, nprev, rs, red);
  const double alphaD = st->rz[n & 1] / red[0];
  const T alphaPrev = XM == 2 ? static_cast<T>(st->alphaH[(n - 1) & 1]) : T(0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = alphaD;
    st->alphaH[n & 1] = alphaD;
  }
  const T alpha = static_cast<T>(alphaD);
  const int lane = threadIdx.x & (kWave - 1);
  KSum rz, rr;
  int64_t col, end, stride;
  column_loop(g.columns(), col, end, stride);
  for (; col < end; col += stride) {
    const int64_t base = column_base(g, col);
    T rv[R], zv[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int k = lane + kWave * m;
      const int64_t c = base + row_in(k, g.nz);
      T xNew = T(0);
      if (XM == 2) xNew = (x[c] + alphaPrev * pPrev[c]) + alpha * p[c];
      const T rNew = r[c] - alpha * q[c];
      rv[m] = k < g.nz ? rNew : T(0);
      if (XM == 2 && k < g.nz) x[c] = xNew;
    }
    apply_precond<T, R, LINE>(lane, g.nz, base, invd, w, rv, zv);
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int k = lane + kWave * m;
      if (k < g.nz) {
        r[base + k] = rv[m];
        z[base + k] = zv[m];
        rz.add(static_cast<double>(rv[m]) * static_cast<double>(zv[m]));
        rr.add(static_cast<double>(rv[m]) * static_cast<double>(rv[m]));
      }
    }
  }
  double v[2] = {rz.value(), rr.value()};
  const int slots[2] = {0, 1};
  write_parts<2>(parts, pstride, slots, v);
}

// The x term an even last iteration deferred: x += alpha_n p_n (interior).
template <class T>
__global__ __launch_bounds__(kBlock) void k_flush_x(Grid g, T* __restrict__ x,
                                                    const T* __restrict__ p,
                                                    const PcgState* __restrict__ st, int slot) {
  const T alpha = static_cast<T>(st->alphaH[slot]);
  const int64_t n = static_cast<int64_t>(g.columns()) * g.nz;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t c = column_base(g, t / g.nz) + t % g.nz;
    x[c] = x[c] + alpha * p[c];
  }
}

// Mixed form, outer sweep: xo = xi + sc d (all 7 stencil points), r = b - A xo
// in FP64; partials rr (slot 1), bb (slot 3).  Without HASD, xo is xi.  j-march.
// Under device control (ctl set) the sweep is skipped once the solve is over,
// and xi / xo swap when the current iterate is in the second buffer; sc is
// then 1 / s of the control state.
template <int R, bool HASD>
__global__ __launch_bounds__(kBlock) void k_refine64(Grid g, March mj, Stencil<double> A,
                                                     const double* __restrict__ b,
                                                     double* xa, double* xb,
                                                     const float* __restrict__ d, double sc,
                                                     double* __restrict__ r,
                                                     double* __restrict__ parts, int pstride,
                                                     const PcgState* __restrict__ ctl) {
  const double* __restrict__ xi = xa;
  double* __restrict__ xo = xb;
  if (ctl) {
    if (ctl->skip) return;
    if (ctl->xInB) {
      xi = xb;
      xo = xa;
    }
    sc = 1.0 / ctl->s;
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t sx = g.nz, sy = static_cast<int64_t>(g.nx2) * g.nz;
  KSum rr, bb;
  auto xn = [&](int64_t c) { return HASD ? xi[c] + sc * static_cast<double>(d[c]) : xi[c]; };
  auto formx = [&](int64_t base, double (&xv)[R]) {
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int k = lane + kWave * m;
      xv[m] = k < g.nz ? xn(base + k) : 0.0;
    }
  };
  int i, j0, j1;
  if (march_span(g, mj, i, j0, j1)) {
    double xPrev[R], xCur[R], xNext[R], ayPrev[R];
    formx(g.c3(j0 - 1, i, 0), xPrev);
    formx(g.c3(j0, i, 0), xCur);
    load_col<double, R>(A.Ayp, g.c3(j0 - 1, i, 0), lane, g.nz, ayPrev);
    for (int j = j0; j <= j1; ++j) {
      const int64_t base = g.c3(j, i, 0);
      formx(base + sy, xNext);
      double xdn[R], xup[R], az[R], azUp[R], ay[R], dummy[R];
      load_col<double, R>(A.Azp, base, lane, g.nz, az);
      load_col<double, R>(A.Ayp, base, lane, g.nz, ay);
      shift_rows<double, R>(xCur, lane, xdn, xup);
      shift_rows<double, R>(az, lane, dummy, azUp);
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int k = lane + kWave * m;
        if (k < g.nz) {
          const int64_t c = base + k;
          const double qv = A.Ad[c] * xCur[m] + A.Axp[c] * xn(c + sx) + A.Axp[c - sx] * xn(c - sx) +
                            ay[m] * xNext[m] + ayPrev[m] * xPrev[m] + az[m] * xdn[m] +
                            azUp[m] * xup[m];
          const double bv = b[c];
          const double rv = bv - qv;
          if (HASD) xo[c] = xCur[m];
          r[c] = rv;
          rr.add(rv * rv);
          bb.add(bv * bv);
        }
        xPrev[m] = xCur[m];
        xCur[m] = xNext[m];
        ayPrev[m] = ay[m];
      }
    }
  }
  double v[2] = {rr.value(), bb.value()};
  const int slots[2] = {1, 3};
  write_parts<2>(parts, pstride, slots, v);
}

// Mixed form, inner start: r32 = s r64, d = 0, z32 = M32^-1 r32; partials rz, rr.
template <int R, bool LINE>
__global__ __launch_bounds__(kBlock) void k_inner_init(Grid g, const double* __restrict__ r64,
                                                       double s, float* __restrict__ r32,
                                                       float* __restrict__ z32,
                                                       float* __restrict__ d32,
                                                       const float* __restrict__ invd,
                                                       const float* __restrict__ w,
                                                       double* __restrict__ parts, int pstride,
                                                       const PcgState* __restrict__ ctl) {
  if (ctl) {
    if (ctl->skip) return;
    s = ctl->s;
  }
  const int lane = threadIdx.x & (kWave - 1);
  KSum rz, rr;
  int64_t col, end, stride;
  column_loop(g.columns(), col, end, stride);
  for (; col < end; col += stride) {
    const int64_t base = column_base(g, col);
    float rv[R], zv[R];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int k = lane + kWave * m;
      const double t = r64[base + row_in(k, g.nz)];
      rv[m] = k < g.nz ? static_cast<float>(t * s) : 0.0f;
    }
    apply_precond<float, R, LINE>(lane, g.nz, base, invd, w, rv, zv);
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int k = lane + kWave * m;
      if (k < g.nz) {
        r32[base + k] = rv[m];
        z32[base + k] = zv[m];
        d32[base + k] = 0.0f;
        rz.add(static_cast<double>(rv[m]) * static_cast<double>(zv[m]));
        rr.add(static_cast<double>(rv[m]) * static_cast<double>(rv[m]));
      }
    }
  }
  double v[2] = {rz.value(), rr.value()};
  const int slots[2] = {0, 1};
  write_parts<2>(parts, pstride, slots, v);
}

// ---------------------------------------------------------------------------
// Operands in the row-pair layout: lane l of row segment m holds rows
// k0 = 128 m + 2 l and k1 = k0 + 1 (R2 = ceil(nz / 128) segments), so every
// load and store moves two elements a lane (8 bytes in FP32, 16 in FP64),
// and the z-line recurrences first compose each lane's two steps and then
// scan over half as many lanes-rows.  Used when nz is even (the pairs are
// then aligned to their size).
// ---------------------------------------------------------------------------
typedef float float4v __attribute__((ext_vector_type(4)));
template <bool NT>
__device__ __forceinline__ void store4(float* p, float4v v) {
  if constexpr (NT) __builtin_nontemporal_store(v, reinterpret_cast<float4v*>(p));
  else *reinterpret_cast<float4v*>(p) = v;
}
// The update kernels store r, z (and x) non-temporally: the direction kernel
// after them then runs 2-3 % faster (FP64 0.336 -> 0.330 ms, FP32 0.181 ->
// 0.176 ms at 512 x 512 x 100) and they lose nothing themselves.
constexpr bool kUpdateNts = true;
template <class T> struct PairOf;
template <> struct PairOf<float> { typedef float v __attribute__((ext_vector_type(2))); };
template <> struct PairOf<double> { typedef double v __attribute__((ext_vector_type(2))); };

template <int R2, class T>
__device__ __forceinline__ void load_pair(const T* __restrict__ a, int64_t base, int lane, int nz,
                                          T (&v0)[R2], T (&v1)[R2]) {
  using V = typename PairOf<T>::v;
#pragma unroll
  for (int m = 0; m < R2; ++m) {
    const int k0 = 128 * m + 2 * lane;
    // nz is even in this layout: the last pair (nz - 2, nz - 1) stands in
    const V t = *reinterpret_cast<const V*>(a + base + (k0 < nz ? k0 : nz - 2));
    v0[m] = k0 < nz ? t.x : T(0);
    v1[m] = k0 < nz ? t.y : T(0);
  }
}

// NT: non-temporal stores (streamed past the caches' retention)
template <int R2, bool NT = false, class T>
__device__ __forceinline__ void store_pair(T* __restrict__ a, int64_t base, int lane, int nz,
                                           const T (&v0)[R2], const T (&v1)[R2]) {
  using V = typename PairOf<T>::v;
#pragma unroll
  for (int m = 0; m < R2; ++m) {
    const int k0 = 128 * m + 2 * lane;
    V t;
    t.x = v0[m];
    t.y = v1[m];
    if (k0 < nz) {
      if constexpr (NT) __builtin_nontemporal_store(t, reinterpret_cast<V*>(a + base + k0));
      else *reinterpret_cast<V*>(a + base + k0) = t;
    }
  }
}

// Row k1 + 1 (dn) and row k0 - 1 (up) of a column in the pair layout.
template <int R2, class T>
__device__ __forceinline__ void pair_shift(const T (&v0)[R2], const T (&v1)[R2], int lane,
                                           T (&dn)[R2], T (&up)[R2]) {
#pragma unroll
  for (int m = 0; m < R2; ++m) {
    const T nx0 = __shfl_down(v0[m], 1, kWave);
    const T seg = (m + 1 < R2) ? __shfl(v0[m + 1 < R2 ? m + 1 : m], 0, kWave) : T(0);
    dn[m] = lane < kWave - 1 ? nx0 : seg;
    const T pv1 = __shfl_up(v1[m], 1, kWave);
    const T segp = (m > 0) ? __shfl(v1[m > 0 ? m - 1 : 0], kWave - 1, kWave) : T(0);
    up[m] = lane > 0 ? pv1 : segp;
  }
}

// z = M^-1 r in the pair layout (the recurrences of apply_precond), from the
// factors iv = invd, w of the rows (w unused without LINE).
template <int R2, bool LINE, class T>
__device__ __forceinline__ void precond_pair_f(int lane, const T (&iv0)[R2], const T (&iv1)[R2],
                                               const T (&w0)[R2], const T (&w1)[R2],
                                               const T (&r0)[R2], const T (&r1)[R2],
                                               T (&z0)[R2], T (&z1)[R2]) {
  if constexpr (!LINE) {
#pragma unroll
    for (int m = 0; m < R2; ++m) {
      z0[m] = r0[m] * iv0[m];
      z1[m] = r1[m] * iv1[m];
    }
    return;
  } else {
    T wDn[R2], wUp[R2];
    pair_shift<R2>(w0, w1, lane, wDn, wUp);
    T y0[R2], y1[R2];
    T carry = T(0);
#pragma unroll
    for (int m = 0; m < R2; ++m) {
      // y_k = r_k - w_{k-1} y_{k-1}: rows k0 then k1, composed per lane
      const T a0 = r0[m], b0 = -wUp[m];
      const T a1 = r1[m], b1 = -w0[m];
      const T cin = carry;
      const T Y1 = scan_affine_up(a1 + b1 * a0, b1 * b0, carry, lane);
      const T P = dpp_move<kDppWaveShr1>(cin, Y1);  // row k0 - 1
      y0[m] = a0 + b0 * P;
      y1[m] = Y1;
    }
    carry = T(0);
#pragma unroll
    for (int m = R2 - 1; m >= 0; --m) {
      // z_k = y_k invd_k - w_k z_{k+1}: rows k1 then k0, composed per lane
      const T a0 = y0[m] * iv0[m], b0 = -w0[m];
      const T a1 = y1[m] * iv1[m], b1 = -w1[m];
      const T cin = carry;
      const T Z0 = scan_affine_down(a0 + b0 * a1, b0 * b1, carry, lane);
      const T Nx = dpp_move<kDppWaveShl1>(cin, Z0);  // row k1 + 1
      z0[m] = Z0;
      z1[m] = a1 + b1 * Nx;
    }
  }
}

// precond_pair_f with the factors loaded from invd, w.
template <int R2, bool LINE, class T>
__device__ __forceinline__ void precond_pair(int lane, int nz, int64_t base,
                                             const T* __restrict__ invd, const T* __restrict__ w,
                                             const T (&r0)[R2], const T (&r1)[R2], T (&z0)[R2],
                                             T (&z1)[R2]) {
  T iv0[R2], iv1[R2], w0[R2], w1[R2];
  load_pair<R2>(invd, base, lane, nz, iv0, iv1);
  if constexpr (LINE) load_pair<R2>(w, base, lane, nz, w0, w1);
  precond_pair_f<R2, LINE>(lane, iv0, iv1, w0, w1, r0, r1, z0, z1);
}

// k_pcg_dir in the pair layout (one column a wave).
template <class T, int R2, bool RESTART>
__global__ __launch_bounds__(kBlock) void k_pcg_dir_p(Grid g, March mj, Stencil<T> A,
                                                      const T* __restrict__ z,
                                                      const T* __restrict__ pOld,
                                                      T* __restrict__ pNew,
                                                      T* __restrict__ q,
                                                      double* __restrict__ parts, int pstride,
                                                      int nprev, PcgState* __restrict__ st,
                                                      int n) {
  if (st->done) return;
  if (n >= st->limit) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->it = n;
    }
    return;
  }
  double red[2];
  const int rs[2] = {0, 1};
  reduce_parts<2>(parts, pstride, nprev, rs, red);
  const double rzv = red[0], rrv = red[1];
  if (rrv <= st->tol2) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->it = n;
      st->rr = rrv;
    }
    return;
  }
  const double betaD = RESTART ? 0.0 : rzv / st->rz[(n + 1) & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz[n & 1] = rzv;
    st->it = n;
    st->rr = rrv;
  }
  const T beta = static_cast<T>(betaD);
  const int lane = threadIdx.x & (kWave - 1);
  const int nz = g.nz;
  const int64_t sx = nz, sy = static_cast<int64_t>(g.nx2) * nz;
  KSum pq;
  int i, j0, j1;
  if (march_span(g, mj, i, j0, j1)) {
    auto formp = [&](int64_t base, T (&p0)[R2], T (&p1)[R2]) {
      T z0[R2], z1[R2];
      load_pair<R2>(z, base, lane, nz, z0, z1);
      if (RESTART) {
#pragma unroll
        for (int m = 0; m < R2; ++m) {
          p0[m] = z0[m];
          p1[m] = z1[m];
        }
      } else {
        T o0[R2], o1[R2];
        load_pair<R2>(pOld, base, lane, nz, o0, o1);
#pragma unroll
        for (int m = 0; m < R2; ++m) {
          p0[m] = z0[m] + beta * o0[m];
          p1[m] = z1[m] + beta * o1[m];
        }
      }
    };
    T pP0[R2], pP1[R2], pC0[R2], pC1[R2], ayP0[R2], ayP1[R2];
    formp(g.c3(j0 - 1, i, 0), pP0, pP1);
    formp(g.c3(j0, i, 0), pC0, pC1);
    load_pair<R2>(A.Ayp, g.c3(j0 - 1, i, 0), lane, nz, ayP0, ayP1);
    for (int j = j0; j <= j1; ++j) {
      const int64_t base = g.c3(j, i, 0);
      T pN0[R2], pN1[R2], px0[R2], px1[R2], pm0[R2], pm1[R2];
      formp(base + sy, pN0, pN1);
      formp(base + sx, px0, px1);
      formp(base - sx, pm0, pm1);
      T ad0[R2], ad1[R2], ax0[R2], ax1[R2], am0[R2], am1[R2], ay0[R2], ay1[R2], az0[R2],
          az1[R2];
      load_pair<R2>(A.Ad, base, lane, nz, ad0, ad1);
      load_pair<R2>(A.Axp, base, lane, nz, ax0, ax1);
      load_pair<R2>(A.Axp, base - sx, lane, nz, am0, am1);
      load_pair<R2>(A.Ayp, base, lane, nz, ay0, ay1);
      load_pair<R2>(A.Azp, base, lane, nz, az0, az1);
      T pDn[R2], pUp[R2], azDn[R2], azUp[R2];
      pair_shift<R2>(pC0, pC1, lane, pDn, pUp);
      pair_shift<R2>(az0, az1, lane, azDn, azUp);
      T q0[R2], q1[R2];
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        q0[m] = ad0[m] * pC0[m] + ax0[m] * px0[m] + am0[m] * pm0[m] + ay0[m] * pN0[m] +
                ayP0[m] * pP0[m] + az0[m] * pC1[m] + azUp[m] * pUp[m];
        q1[m] = ad1[m] * pC1[m] + ax1[m] * px1[m] + am1[m] * pm1[m] + ay1[m] * pN1[m] +
                ayP1[m] * pP1[m] + az1[m] * pDn[m] + az0[m] * pC0[m];
        if (128 * m + 2 * lane < nz) {
          pq.add(static_cast<double>(pC0[m]) * static_cast<double>(q0[m]));
          pq.add(static_cast<double>(pC1[m]) * static_cast<double>(q1[m]));
        }
      }
      store_pair<R2>(pNew, base, lane, nz, pC0, pC1);
      store_pair<R2>(q, base, lane, nz, q0, q1);
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        pP0[m] = pC0[m];
        pP1[m] = pC1[m];
        pC0[m] = pN0[m];
        pC1[m] = pN1[m];
        ayP0[m] = ay0[m];
        ayP1[m] = ay1[m];
      }
    }
  }
  double v[1] = {pq.value()};
  const int slots[1] = {2};
  write_parts<1>(parts, pstride, slots, v);
}

// ---------------------------------------------------------------------------
// Jacobi-Chebyshev preconditioner (FP32, pair layout): z = p_m(D^-1 A) D^-1 r,
// p_m the degree-m polynomial of the Chebyshev semi-iteration for D^-1 A on
// [a, b] (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed.,
// Algorithm 12.1), theta = (b + a) / 2, delta = (b - a) / 2,
// sigma = theta / delta, rho_0 = 1 / sigma:
//   z_1 = d_0 = D^-1 r / theta
//   rho_k = 1 / (2 sigma - rho_{k-1})
//   d_k = rho_k rho_{k-1} d_{k-1} + (2 rho_k / delta) D^-1 (r - A z_k)
//   z_{k+1} = z_k + d_k,          k = 1 .. m - 1.
// D^-1 A of a diagonally dominant matrix has its spectrum in (0, 2], so
// b = 2.  The residual polynomial q_m(l) = 1 - l p_m(l) = T_m((b + a - 2 l) /
// (b - a)) / T_m(sigma) is below 1 on (0, b] (it falls from q_m(0) = 1 and
// stays within 1 / T_m(sigma) on [a, b]) and, for odd m, negative beyond b;
// so p_m > 0 on the whole positive axis and the preconditioner is SPD for
// any spectrum.  The update kernel (point Jacobi)
// leaves D^-1 r in z; step k reads z_k (FIRST: D^-1 r scaled by 1 / theta,
// d_{k-1} likewise) and writes d_k and z_{k+1}; the last step leaves the
// partial sums r.z and r.r for the next direction kernel.  j-march.
// ---------------------------------------------------------------------------
template <int R2, bool FIRST, bool LAST>
__global__ __launch_bounds__(kBlock) void k_cheb_step_p(Grid g, March mj, Stencil<float> A,
                                                        const float* __restrict__ zin,
                                                        float* __restrict__ zout,
                                                        float* __restrict__ d,
                                                        const float* __restrict__ r,
                                                        const float* __restrict__ invd,
                                                        float invTheta, float c1, float c2,
                                                        double* __restrict__ parts, int pstride,
                                                        const int* __restrict__ skip) {
  if (skip && *skip) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int nz = g.nz;
  const int64_t sx = nz, sy = static_cast<int64_t>(g.nx2) * nz;
  KSum rz, rr;
  int i, j0, j1;
  if (march_span(g, mj, i, j0, j1)) {
    auto formz = [&](int64_t base, float (&z0)[R2], float (&z1)[R2]) {
      load_pair<R2>(zin, base, lane, nz, z0, z1);
      if (FIRST) {
#pragma unroll
        for (int m = 0; m < R2; ++m) {
          z0[m] *= invTheta;
          z1[m] *= invTheta;
        }
      }
    };
    float zP0[R2], zP1[R2], zC0[R2], zC1[R2], ayP0[R2], ayP1[R2];
    formz(g.c3(j0 - 1, i, 0), zP0, zP1);
    formz(g.c3(j0, i, 0), zC0, zC1);
    load_pair<R2>(A.Ayp, g.c3(j0 - 1, i, 0), lane, nz, ayP0, ayP1);
    for (int j = j0; j <= j1; ++j) {
      const int64_t base = g.c3(j, i, 0);
      float zN0[R2], zN1[R2], zx0[R2], zx1[R2], zm0[R2], zm1[R2];
      formz(base + sy, zN0, zN1);
      formz(base + sx, zx0, zx1);
      formz(base - sx, zm0, zm1);
      float ad0[R2], ad1[R2], ax0[R2], ax1[R2], am0[R2], am1[R2], ay0[R2], ay1[R2], az0[R2],
          az1[R2];
      load_pair<R2>(A.Ad, base, lane, nz, ad0, ad1);
      load_pair<R2>(A.Axp, base, lane, nz, ax0, ax1);
      load_pair<R2>(A.Axp, base - sx, lane, nz, am0, am1);
      load_pair<R2>(A.Ayp, base, lane, nz, ay0, ay1);
      load_pair<R2>(A.Azp, base, lane, nz, az0, az1);
      float zDn[R2], zUp[R2], azDn[R2], azUp[R2];
      pair_shift<R2>(zC0, zC1, lane, zDn, zUp);
      pair_shift<R2>(az0, az1, lane, azDn, azUp);
      float r0[R2], r1[R2], iv0[R2], iv1[R2], d0[R2], d1[R2];
      load_pair<R2>(r, base, lane, nz, r0, r1);
      load_pair<R2>(invd, base, lane, nz, iv0, iv1);
      if (FIRST) {
#pragma unroll
        for (int m = 0; m < R2; ++m) {
          d0[m] = zC0[m];
          d1[m] = zC1[m];
        }
      } else {
        load_pair<R2>(d, base, lane, nz, d0, d1);
      }
      float n0[R2], n1[R2];
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        const float q0 = ad0[m] * zC0[m] + ax0[m] * zx0[m] + am0[m] * zm0[m] + ay0[m] * zN0[m] +
                         ayP0[m] * zP0[m] + az0[m] * zC1[m] + azUp[m] * zUp[m];
        const float q1 = ad1[m] * zC1[m] + ax1[m] * zx1[m] + am1[m] * zm1[m] + ay1[m] * zN1[m] +
                         ayP1[m] * zP1[m] + az1[m] * zDn[m] + az0[m] * zC0[m];
        d0[m] = c1 * d0[m] + c2 * (iv0[m] * (r0[m] - q0));
        d1[m] = c1 * d1[m] + c2 * (iv1[m] * (r1[m] - q1));
        n0[m] = zC0[m] + d0[m];
        n1[m] = zC1[m] + d1[m];
        if (LAST && 128 * m + 2 * lane < nz) {
          rz.add(static_cast<double>(r0[m]) * static_cast<double>(n0[m]));
          rz.add(static_cast<double>(r1[m]) * static_cast<double>(n1[m]));
          rr.add(static_cast<double>(r0[m]) * static_cast<double>(r0[m]));
          rr.add(static_cast<double>(r1[m]) * static_cast<double>(r1[m]));
        }
      }
      if (!LAST) store_pair<R2>(d, base, lane, nz, d0, d1);
      store_pair<R2>(zout, base, lane, nz, n0, n1);
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        zP0[m] = zC0[m];
        zP1[m] = zC1[m];
        zC0[m] = zN0[m];
        zC1[m] = zN1[m];
        ayP0[m] = ay0[m];
        ayP1[m] = ay1[m];
      }
    }
  }
  if (LAST) {
    double v[2] = {rz.value(), rr.value()};
    const int slots[2] = {0, 1};
    write_parts<2>(parts, pstride, slots, v);
  }
}

// k_pcg_dir_p with two neighbouring columns (i, i + 1) per wave: their p is
// formed once for both stencils, the shared x leg Axp(i) is read once, and
// only the outer neighbours i - 1 and i + 2 come from the XCD's L2 (half the
// cache loads of the one-column form).  nx even.  NTS: p and q are stored
// non-temporally.
template <class T, int R2, bool RESTART, bool NTS>
__global__ __launch_bounds__(kBlock) void k_pcg_dir_p2(Grid g, March mj, Stencil<T> A,
                                                       const T* __restrict__ z,
                                                       const T* __restrict__ pOld,
                                                       T* __restrict__ pNew,
                                                       T* __restrict__ q,
                                                       double* __restrict__ parts, int pstride,
                                                       int nprev, PcgState* __restrict__ st,
                                                       int n) {
  if (st->done) return;
  if (n >= st->limit) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->it = n;
    }
    return;
  }
  double red[2];
  const int rs[2] = {0, 1};
  reduce_parts<2>(parts, pstride, nprev, rs, red);
  const double rzv = red[0], rrv = red[1];
  if (rrv <= st->tol2) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->it = n;
      st->rr = rrv;
    }
    return;
  }
  const double betaD = RESTART ? 0.0 : rzv / st->rz[(n + 1) & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz[n & 1] = rzv;
    st->it = n;
    st->rr = rrv;
  }
  const T beta = static_cast<T>(betaD);
  const int lane = threadIdx.x & (kWave - 1);
  const int nz = g.nz;
  const int64_t sx = nz, sy = static_cast<int64_t>(g.nx2) * nz;
  KSum pq;
  int i, j0, j1;
  if (march_span2(g, mj, i, j0, j1)) {
    // p = z + beta p_old of a column: loaded (loadp), then formed (mkp)
    struct Zo {
      T z0[R2], z1[R2], o0[R2], o1[R2];
    };
    auto loadp = [&](int64_t base, Zo& v) {
      load_pair<R2>(z, base, lane, nz, v.z0, v.z1);
      if (!RESTART) load_pair<R2>(pOld, base, lane, nz, v.o0, v.o1);
    };
    auto mkp = [&](const Zo& v, T (&p0)[R2], T (&p1)[R2]) {
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        if (RESTART) {
          p0[m] = v.z0[m];
          p1[m] = v.z1[m];
        } else {
          p0[m] = v.z0[m] + beta * v.o0[m];
          p1[m] = v.z1[m] + beta * v.o1[m];
        }
      }
    };
    auto formp = [&](int64_t base, T (&p0)[R2], T (&p1)[R2]) {
      Zo v;
      loadp(base, v);
      mkp(v, p0, p1);
    };
    // column a = i, column b = i + 1
    T aP0[R2], aP1[R2], aC0[R2], aC1[R2], aY0[R2], aY1[R2];
    T bP0[R2], bP1[R2], bC0[R2], bC1[R2], bY0[R2], bY1[R2];
    formp(g.c3(j0 - 1, i, 0), aP0, aP1);
    formp(g.c3(j0, i, 0), aC0, aC1);
    formp(g.c3(j0 - 1, i + 1, 0), bP0, bP1);
    formp(g.c3(j0, i + 1, 0), bC0, bC1);
    load_pair<R2>(A.Ayp, g.c3(j0 - 1, i, 0), lane, nz, aY0, aY1);
    load_pair<R2>(A.Ayp, g.c3(j0 - 1, i + 1, 0), lane, nz, bY0, bY1);
    for (int j = j0; j <= j1; ++j) {
      const int64_t ba = g.c3(j, i, 0), bb = ba + sx;
      Zo vaN, vbN, vW, vE;
      loadp(ba + sy, vaN);
      loadp(bb + sy, vbN);
      loadp(ba - sx, vW);  // column i - 1
      loadp(bb + sx, vE);  // column i + 2
      T adA0[R2], adA1[R2], axA0[R2], axA1[R2], axW0[R2], axW1[R2], ayA0[R2], ayA1[R2],
          azA0[R2], azA1[R2];
      T adB0[R2], adB1[R2], axB0[R2], axB1[R2], ayB0[R2], ayB1[R2], azB0[R2], azB1[R2];
      load_pair<R2>(A.Ad, ba, lane, nz, adA0, adA1);
      load_pair<R2>(A.Axp, ba, lane, nz, axA0, axA1);
      load_pair<R2>(A.Axp, ba - sx, lane, nz, axW0, axW1);
      load_pair<R2>(A.Ayp, ba, lane, nz, ayA0, ayA1);
      load_pair<R2>(A.Azp, ba, lane, nz, azA0, azA1);
      load_pair<R2>(A.Ad, bb, lane, nz, adB0, adB1);
      load_pair<R2>(A.Axp, bb, lane, nz, axB0, axB1);
      load_pair<R2>(A.Ayp, bb, lane, nz, ayB0, ayB1);
      load_pair<R2>(A.Azp, bb, lane, nz, azB0, azB1);
      // Every load of the step ahead of the arithmetic: left to itself the
      // scheduler spreads them between waits in some forms (the restart
      // form at R2 = 1: 0.43 against 0.29 ms a launch).
      __builtin_amdgcn_sched_barrier(0);
      T aN0[R2], aN1[R2], bN0[R2], bN1[R2], w0[R2], w1[R2], e0[R2], e1[R2];
      mkp(vaN, aN0, aN1);
      mkp(vbN, bN0, bN1);
      mkp(vW, w0, w1);
      mkp(vE, e0, e1);
      T aDn[R2], aUp[R2], azAUp[R2], dummyA[R2], bDn[R2], bUp[R2], azBUp[R2], dummyB[R2];
      pair_shift<R2>(aC0, aC1, lane, aDn, aUp);
      pair_shift<R2>(azA0, azA1, lane, dummyA, azAUp);
      pair_shift<R2>(bC0, bC1, lane, bDn, bUp);
      pair_shift<R2>(azB0, azB1, lane, dummyB, azBUp);
      T qa0[R2], qa1[R2], qb0[R2], qb1[R2];
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        qa0[m] = adA0[m] * aC0[m] + axA0[m] * bC0[m] + axW0[m] * w0[m] + ayA0[m] * aN0[m] +
                 aY0[m] * aP0[m] + azA0[m] * aC1[m] + azAUp[m] * aUp[m];
        qa1[m] = adA1[m] * aC1[m] + axA1[m] * bC1[m] + axW1[m] * w1[m] + ayA1[m] * aN1[m] +
                 aY1[m] * aP1[m] + azA1[m] * aDn[m] + azA0[m] * aC0[m];
        qb0[m] = adB0[m] * bC0[m] + axB0[m] * e0[m] + axA0[m] * aC0[m] + ayB0[m] * bN0[m] +
                 bY0[m] * bP0[m] + azB0[m] * bC1[m] + azBUp[m] * bUp[m];
        qb1[m] = adB1[m] * bC1[m] + axB1[m] * e1[m] + axA1[m] * aC1[m] + ayB1[m] * bN1[m] +
                 bY1[m] * bP1[m] + azB1[m] * bDn[m] + azB0[m] * bC0[m];
        if (128 * m + 2 * lane < nz) {
          pq.add(static_cast<double>(aC0[m]) * static_cast<double>(qa0[m]));
          pq.add(static_cast<double>(aC1[m]) * static_cast<double>(qa1[m]));
          pq.add(static_cast<double>(bC0[m]) * static_cast<double>(qb0[m]));
          pq.add(static_cast<double>(bC1[m]) * static_cast<double>(qb1[m]));
        }
      }
      store_pair<R2, NTS>(pNew, ba, lane, nz, aC0, aC1);
      store_pair<R2, NTS>(q, ba, lane, nz, qa0, qa1);
      store_pair<R2, NTS>(pNew, bb, lane, nz, bC0, bC1);
      store_pair<R2, NTS>(q, bb, lane, nz, qb0, qb1);
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        aP0[m] = aC0[m];
        aP1[m] = aC1[m];
        aC0[m] = aN0[m];
        aC1[m] = aN1[m];
        aY0[m] = ayA0[m];
        aY1[m] = ayA1[m];
        bP0[m] = bC0[m];
        bP1[m] = bC1[m];
        bC0[m] = bN0[m];
        bC1[m] = bN1[m];
        bY0[m] = ayB0[m];
        bY1[m] = ayB1[m];
      }
    }
  }
  double v[1] = {pq.value()};
  const int slots[1] = {2};
  write_parts<1>(parts, pstride, slots, v);
}

// FP32 direction kernel in the quad layout: lane l, half h = l / 32 of the
// wave, owns column i + h (i odd, nx even) and its rows 4 (l % 32) .. + 3
// (nz a multiple of 4, at most 128), so every load and store moves 16 bytes
// a lane and one instruction covers both columns.  Column i's east
// neighbour and column i + 1's west neighbour are the other half's own
// values (lane l ^ 32); only the outer neighbours i - 1 and i + 2 come from
// the XCD's L2.  0.166 ms against 0.175 ms for k_pcg_dir_p2 at 512 x 512 x
// 100, both at 4-8 waves a CU.
__device__ __forceinline__ float4v swap_halves(float4v v) {
  float4v r;
  r.x = __shfl_xor(v.x, 32, kWave);
  r.y = __shfl_xor(v.y, 32, kWave);
  r.z = __shfl_xor(v.z, 32, kWave);
  r.w = __shfl_xor(v.w, 32, kWave);
  return r;
}

template <bool RESTART>
__global__ __launch_bounds__(kBlock) void k_pcg_dir_q(Grid g, March mj, Stencil<float> A,
                                                      const float* __restrict__ z,
                                                      const float* __restrict__ pOld,
                                                      float* __restrict__ pNew,
                                                      float* __restrict__ q,
                                                      double* __restrict__ parts, int pstride,
                                                      int nprev, PcgState* __restrict__ st,
                                                      int n) {
  if (st->done) return;
  if (n >= st->limit) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->it = n;
    }
    return;
  }
  double red[2];
  const int rs[2] = {0, 1};
  reduce_parts<2>(parts, pstride, nprev, rs, red);
  const double rzv = red[0], rrv = red[1];
  if (rrv <= st->tol2) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->it = n;
      st->rr = rrv;
    }
    return;
  }
  const double betaD = RESTART ? 0.0 : rzv / st->rz[(n + 1) & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz[n & 1] = rzv;
    st->it = n;
    st->rr = rrv;
  }
  const float beta = static_cast<float>(betaD);
  const int lane = threadIdx.x & (kWave - 1);
  const int h = lane >> 5, hl = lane & 31;
  const int nz = g.nz;
  const int64_t sy = static_cast<int64_t>(g.nx2) * nz;
  const int k0 = 4 * hl;
  const bool in = k0 < nz;
  const int kc = in ? k0 : nz - 4;  // lanes past the column read the last quad
  KSum pq;
  int i, j0, j1;
  if (march_span2(g, mj, i, j0, j1)) {
    const int64_t own = g.c3(0, i + h, 0) + kc;                // + j sy
    const int64_t outer = g.c3(0, h ? i + 2 : i - 1, 0) + kc;  // + j sy
    auto ld = [](const float* a) { return *reinterpret_cast<const float4v*>(a); };
    auto formp = [&](int64_t c) {
      const float4v p = RESTART ? ld(z + c) : ld(z + c) + beta * ld(pOld + c);
      return in ? p : float4v(0.0f);
    };
    float4v pP = formp(own + (j0 - 1) * sy), pC = formp(own + j0 * sy);
    float4v ayP = ld(A.Ayp + own + (j0 - 1) * sy);
    for (int j = j0; j <= j1; ++j) {
      const int64_t co = own + j * sy, cx = outer + j * sy;
      const float4v pN = formp(co + sy), pO = formp(cx);
      const float4v ad = ld(A.Ad + co), ax = ld(A.Axp + co), ay = ld(A.Ayp + co);
      const float4v az = ld(A.Azp + co);
      float4v axW = float4v(0.0f);
      if (h == 0) axW = ld(A.Axp + cx);  // the leg from column i - 1
      const float4v pX = swap_halves(pC), axX = swap_halves(ax);
      if (h) axW = axX;  // column i + 1's west leg is column i's east leg
      const float4v pE = h ? pO : pX, pW = h ? pX : pO;
      // rows k0 - 1 and k0 + 4 of the own column (none across the halves)
      const float up = __shfl_up(pC.w, 1, kWave), dn = __shfl_down(pC.x, 1, kWave);
      const float azu = __shfl_up(az.w, 1, kWave);
      const float pUp = hl > 0 ? up : 0.0f, pDn = hl < 31 ? dn : 0.0f;
      const float azUp = hl > 0 ? azu : 0.0f;
      float4v qv;
      qv.x = ad.x * pC.x + ax.x * pE.x + axW.x * pW.x + ay.x * pN.x + ayP.x * pP.x +
             az.x * pC.y + azUp * pUp;
      qv.y = ad.y * pC.y + ax.y * pE.y + axW.y * pW.y + ay.y * pN.y + ayP.y * pP.y +
             az.y * pC.z + az.x * pC.x;
      qv.z = ad.z * pC.z + ax.z * pE.z + axW.z * pW.z + ay.z * pN.z + ayP.z * pP.z +
             az.z * pC.w + az.y * pC.y;
      qv.w = ad.w * pC.w + ax.w * pE.w + axW.w * pW.w + ay.w * pN.w + ayP.w * pP.w +
             az.w * pDn + az.z * pC.z;
      if (in) {
        pq.add(static_cast<double>(pC.x) * static_cast<double>(qv.x));
        pq.add(static_cast<double>(pC.y) * static_cast<double>(qv.y));
        pq.add(static_cast<double>(pC.z) * static_cast<double>(qv.z));
        pq.add(static_cast<double>(pC.w) * static_cast<double>(qv.w));
        *reinterpret_cast<float4v*>(pNew + co) = pC;
        *reinterpret_cast<float4v*>(q + co) = qv;
      }
      pP = pC;
      pC = pN;
      ayP = ay;
    }
  }
  double v[1] = {pq.value()};
  const int slots[1] = {2};
  write_parts<1>(parts, pstride, slots, v);
}

// k_pcg_update in the quad layout of k_pcg_dir_q: a wave takes a pair of
// columns (i, i + 1), one per half, four rows a lane, so the loads and
// stores move 16 bytes a lane and one pair of z-line scans serves both
// columns (scan_affine_up_halves / _down_halves): each lane first composes
// its four steps of a recurrence, the half-wave scans the composed maps, and
// the lane then walks its first three rows from the value entering it.
template <bool LINE, int XM, bool NTS = kUpdateNts>
__global__ __launch_bounds__(kBlock) void k_pcg_update_q(Grid g, float* __restrict__ x,
                                                         const float* __restrict__ pPrev,
                                                         const float* __restrict__ p,
                                                         float* __restrict__ r,
                                                         const float* __restrict__ q,
                                                         float* __restrict__ z,
                                                         const float* __restrict__ invd,
                                                         const float* __restrict__ w,
                                                         double* __restrict__ parts, int pstride,
                                                         int nprev, PcgState* __restrict__ st,
                                                         int n) {
  if (st->done) return;
  double red[1];
  const int rs[1] = {2};
  reduce_parts<1>(parts, pstride, nprev, rs, red);
  const double alphaD = st->rz[n & 1] / red[0];
  const float alphaPrev = XM == 2 ? static_cast<float>(st->alphaH[(n - 1) & 1]) : 0.0f;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = alphaD;
    st->alphaH[n & 1] = alphaD;
  }
  const float alpha = static_cast<float>(alphaD);
  const int lane = threadIdx.x & (kWave - 1);
  const int h = lane >> 5, hl = lane & 31;
  const int nz = g.nz;
  const int k0 = 4 * hl;
  const bool in = k0 < nz;
  const int kc = in ? k0 : nz - 4;
  auto ld = [](const float* a) { return *reinterpret_cast<const float4v*>(a); };
  KSum rz, rr;
  const int half = g.nx / 2;
  int64_t pr, end, stride;
  column_loop(half * g.ny, pr, end, stride);
  for (; pr < end; pr += stride) {
    const int j = static_cast<int>(pr / half) + 1;
    const int i = 2 * static_cast<int>(pr % half) + 1 + h;
    const int64_t c = g.c3(j, i, 0) + kc;
    if (XM == 2) {
      const float4v xn = (ld(x + c) + alphaPrev * ld(pPrev + c)) + alpha * ld(p + c);
      if (in) store4<NTS>(x + c, xn);
    }
    float4v rv = ld(r + c) - alpha * ld(q + c);
    if (!in) rv = float4v(0.0f);
    const float4v iv = in ? ld(invd + c) : float4v(0.0f);
    float4v zv;
    if constexpr (!LINE) {
      zv = rv * iv;
    } else {
      const float4v wv = in ? ld(w + c) : float4v(0.0f);
      const float wU0 = __shfl_up(wv.w, 1, kWave);
      const float wU = hl > 0 ? wU0 : 0.0f;  // w of row k0 - 1
      // forward y_k = r_k - w_{k-1} y_{k-1}: the lane's four steps as one map
      float a = rv.x, b = -wU;
      a = rv.y - wv.x * a;
      b = -wv.x * b;
      a = rv.z - wv.y * a;
      b = -wv.y * b;
      a = rv.w - wv.z * a;
      b = -wv.z * b;
      const float Y3 = scan_affine_up_halves(a, b, lane);
      const float P0 = dpp_move<kDppWaveShr1>(0.0f, Y3);
      const float yIn = hl > 0 ? P0 : 0.0f;  // y of row k0 - 1
      const float y0 = rv.x - wU * yIn, y1 = rv.y - wv.x * y0, y2 = rv.z - wv.y * y1;
      const float y3 = Y3;
      // backward z_k = y_k invd_k - w_k z_{k+1}
      float c2 = y3 * iv.w, d2 = -wv.w;
      c2 = y2 * iv.z - wv.z * c2;
      d2 = -wv.z * d2;
      c2 = y1 * iv.y - wv.y * c2;
      d2 = -wv.y * d2;
      c2 = y0 * iv.x - wv.x * c2;
      d2 = -wv.x * d2;
      const float Z0 = scan_affine_down_halves(c2, d2, lane);
      const float N0 = dpp_move<kDppWaveShl1>(0.0f, Z0);
      const float zIn = hl < 31 ? N0 : 0.0f;  // z of row k0 + 4
      zv.w = y3 * iv.w - wv.w * zIn;
      zv.z = y2 * iv.z - wv.z * zv.w;
      zv.y = y1 * iv.y - wv.y * zv.z;
      zv.x = Z0;
    }
    if (in) {
      store4<NTS>(r + c, rv);
      store4<NTS>(z + c, zv);
      rz.add(static_cast<double>(rv.x) * static_cast<double>(zv.x));
      rz.add(static_cast<double>(rv.y) * static_cast<double>(zv.y));
      rz.add(static_cast<double>(rv.z) * static_cast<double>(zv.z));
      rz.add(static_cast<double>(rv.w) * static_cast<double>(zv.w));
      rr.add(static_cast<double>(rv.x) * static_cast<double>(rv.x));
      rr.add(static_cast<double>(rv.y) * static_cast<double>(rv.y));
      rr.add(static_cast<double>(rv.z) * static_cast<double>(rv.z));
      rr.add(static_cast<double>(rv.w) * static_cast<double>(rv.w));
    }
  }
  double v[2] = {rz.value(), rr.value()};
  const int slots[2] = {0, 1};
  write_parts<2>(parts, pstride, slots, v);
}

// k_pcg_update in the pair layout (x deferred the same way).
template <class T, int R2, bool LINE, int XM, bool NTS = kUpdateNts>
__global__ __launch_bounds__(kBlock) void k_pcg_update_p(Grid g, T* __restrict__ x,
                                                         const T* __restrict__ pPrev,
                                                         const T* __restrict__ p,
                                                         T* __restrict__ r,
                                                         const T* __restrict__ q,
                                                         T* __restrict__ z,
                                                         const T* __restrict__ invd,
                                                         const T* __restrict__ w,
                                                         double* __restrict__ parts, int pstride,
                                                         int nprev, PcgState* __restrict__ st,
                                                         int n) {
  if (st->done) return;
  double red[1];
  const int rs[1] = {2};
  reduce_parts<1>(parts, pstride, nprev, rs, red);
  const double alphaD = st->rz[n & 1] / red[0];
  const T alphaPrev = XM == 2 ? static_cast<T>(st->alphaH[(n - 1) & 1]) : T(0);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha = alphaD;
    st->alphaH[n & 1] = alphaD;
  }
  const T alpha = static_cast<T>(alphaD);
  const int lane = threadIdx.x & (kWave - 1);
  const int nz = g.nz;
  KSum rz, rr;
  int64_t col, end, stride;
  column_loop(g.columns(), col, end, stride);
  for (; col < end; col += stride) {
    const int64_t base = column_base(g, col);
    T r0[R2], r1[R2], q0[R2], q1[R2];
    if (XM == 2) {
      T x0[R2], x1[R2], a0[R2], a1[R2], p0[R2], p1[R2];
      load_pair<R2>(x, base, lane, nz, x0, x1);
      load_pair<R2>(pPrev, base, lane, nz, a0, a1);
      load_pair<R2>(p, base, lane, nz, p0, p1);
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        x0[m] = (x0[m] + alphaPrev * a0[m]) + alpha * p0[m];
        x1[m] = (x1[m] + alphaPrev * a1[m]) + alpha * p1[m];
      }
      store_pair<R2, NTS>(x, base, lane, nz, x0, x1);
    }
    load_pair<R2>(r, base, lane, nz, r0, r1);
    load_pair<R2>(q, base, lane, nz, q0, q1);
#pragma unroll
    for (int m = 0; m < R2; ++m) {
      r0[m] = r0[m] - alpha * q0[m];
      r1[m] = r1[m] - alpha * q1[m];
    }
    T z0[R2], z1[R2];
    precond_pair<R2, LINE>(lane, nz, base, invd, w, r0, r1, z0, z1);
    store_pair<R2, NTS>(r, base, lane, nz, r0, r1);
    store_pair<R2, NTS>(z, base, lane, nz, z0, z1);
#pragma unroll
    for (int m = 0; m < R2; ++m) {
      if (128 * m + 2 * lane < nz) {
        rz.add(static_cast<double>(r0[m]) * static_cast<double>(z0[m]));
        rz.add(static_cast<double>(r1[m]) * static_cast<double>(z1[m]));
        rr.add(static_cast<double>(r0[m]) * static_cast<double>(r0[m]));
        rr.add(static_cast<double>(r1[m]) * static_cast<double>(r1[m]));
      }
    }
  }
  double v[2] = {rz.value(), rr.value()};
  const int slots[2] = {0, 1};
  write_parts<2>(parts, pstride, slots, v);
}

// k_inner_init in the pair layout: r32 = s r64, d = 0, z32 = M32^-1 r32.
template <int R2, bool LINE>
__global__ __launch_bounds__(kBlock) void k_inner_init_p(Grid g, const double* __restrict__ r64,
                                                         double s, float* __restrict__ r32,
                                                         float* __restrict__ z32,
                                                         float* __restrict__ d32,
                                                         const float* __restrict__ invd,
                                                         const float* __restrict__ w,
                                                         double* __restrict__ parts,
                                                         int pstride,
                                                         const PcgState* __restrict__ ctl) {
  if (ctl) {
    if (ctl->skip) return;
    s = ctl->s;
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int nz = g.nz;
  KSum rz, rr;
  int64_t col, end, stride;
  column_loop(g.columns(), col, end, stride);
  for (; col < end; col += stride) {
    const int64_t base = column_base(g, col);
    float r0[R2], r1[R2], zero[R2];
#pragma unroll
    for (int m = 0; m < R2; ++m) {
      const int k0 = 128 * m + 2 * lane;
      zero[m] = 0.0f;
      const double2 t = *reinterpret_cast<const double2*>(r64 + base + (k0 < nz ? k0 : nz - 2));
      r0[m] = k0 < nz ? static_cast<float>(t.x * s) : 0.0f;
      r1[m] = k0 < nz ? static_cast<float>(t.y * s) : 0.0f;
    }
    float z0[R2], z1[R2];
    precond_pair<R2, LINE>(lane, nz, base, invd, w, r0, r1, z0, z1);
    store_pair<R2>(r32, base, lane, nz, r0, r1);
    store_pair<R2>(z32, base, lane, nz, z0, z1);
    store_pair<R2>(d32, base, lane, nz, zero, zero);
#pragma unroll
    for (int m = 0; m < R2; ++m) {
      if (128 * m + 2 * lane < nz) {
        rz.add(static_cast<double>(r0[m]) * static_cast<double>(z0[m]));
        rz.add(static_cast<double>(r1[m]) * static_cast<double>(z1[m]));
        rr.add(static_cast<double>(r0[m]) * static_cast<double>(r0[m]));
        rr.add(static_cast<double>(r1[m]) * static_cast<double>(r1[m]));
      }
    }
  }
  double v[2] = {rz.value(), rr.value()};
  const int slots[2] = {0, 1};
  write_parts<2>(parts, pstride, slots, v);
}

// ---------------------------------------------------------------------------
// The FP64 sweeps over column pairs in the row-pair layout (nx and nz even,
// nz <= 256), the shape of k_pcg_dir_p2: k_residual_p2 is k_residual64 and
// k_refine_p2 is k_refine64.  XV: x at a column of the pair layout (the
// refinement adds sc d of the FP32 correction first).
// ---------------------------------------------------------------------------
// The stencil's coefficients at column i (A) and i + 1 (B), and the x leg at
// i - 1 (W), in the pair layout.
template <int R2>
struct CoefPair {
  double adA0[R2], adA1[R2], axA0[R2], axA1[R2], axW0[R2], axW1[R2], ayA0[R2], ayA1[R2],
      azA0[R2], azA1[R2];
  double adB0[R2], adB1[R2], axB0[R2], axB1[R2], ayB0[R2], ayB1[R2], azB0[R2], azB1[R2];
};
template <int R2>
__device__ __forceinline__ void load_coef_pair(const Stencil<double>& A, int64_t ba, int64_t sx,
                                               int lane, int nz, CoefPair<R2>& c) {
  const int64_t bb = ba + sx;
  load_pair<R2>(A.Ad, ba, lane, nz, c.adA0, c.adA1);
  load_pair<R2>(A.Axp, ba, lane, nz, c.axA0, c.axA1);
  load_pair<R2>(A.Axp, ba - sx, lane, nz, c.axW0, c.axW1);
  load_pair<R2>(A.Ayp, ba, lane, nz, c.ayA0, c.ayA1);
  load_pair<R2>(A.Azp, ba, lane, nz, c.azA0, c.azA1);
  load_pair<R2>(A.Ad, bb, lane, nz, c.adB0, c.adB1);
  load_pair<R2>(A.Axp, bb, lane, nz, c.axB0, c.axB1);
  load_pair<R2>(A.Ayp, bb, lane, nz, c.ayB0, c.ayB1);
  load_pair<R2>(A.Azp, bb, lane, nz, c.azB0, c.azB1);
}

// q = A x for columns i (qa) and i + 1 (qb) of the pair layout, from x at
// j - 1 (P), j (C), j + 1 (N) and the outer columns (w: i - 1, e: i + 2), the
// y leg toward j - 1 (Y) and the step's coefficients (c).
template <int R2>
__device__ __forceinline__ void stencil_pair_q(
    int lane, const double (&aP0)[R2], const double (&aP1)[R2], const double (&aC0)[R2],
    const double (&aC1)[R2], const double (&aN0)[R2], const double (&aN1)[R2],
    const double (&bP0)[R2], const double (&bP1)[R2], const double (&bC0)[R2],
    const double (&bC1)[R2], const double (&bN0)[R2], const double (&bN1)[R2],
    const double (&w0)[R2], const double (&w1)[R2], const double (&e0)[R2],
    const double (&e1)[R2], const double (&aY0)[R2], const double (&aY1)[R2],
    const double (&bY0)[R2], const double (&bY1)[R2], const CoefPair<R2>& c,
    double (&qa0)[R2], double (&qa1)[R2], double (&qb0)[R2], double (&qb1)[R2]) {
  double aDn[R2], aUp[R2], azAUp[R2], dA[R2], bDn[R2], bUp[R2], azBUp[R2], dB[R2];
  pair_shift<R2>(aC0, aC1, lane, aDn, aUp);
  pair_shift<R2>(c.azA0, c.azA1, lane, dA, azAUp);
  pair_shift<R2>(bC0, bC1, lane, bDn, bUp);
  pair_shift<R2>(c.azB0, c.azB1, lane, dB, azBUp);
#pragma unroll
  for (int m = 0; m < R2; ++m) {
    qa0[m] = c.adA0[m] * aC0[m] + c.axA0[m] * bC0[m] + c.axW0[m] * w0[m] + c.ayA0[m] * aN0[m] +
             aY0[m] * aP0[m] + c.azA0[m] * aC1[m] + azAUp[m] * aUp[m];
    qa1[m] = c.adA1[m] * aC1[m] + c.axA1[m] * bC1[m] + c.axW1[m] * w1[m] + c.ayA1[m] * aN1[m] +
             aY1[m] * aP1[m] + c.azA1[m] * aDn[m] + c.azA0[m] * aC0[m];
    qb0[m] = c.adB0[m] * bC0[m] + c.axB0[m] * e0[m] + c.axA0[m] * aC0[m] + c.ayB0[m] * bN0[m] +
             bY0[m] * bP0[m] + c.azB0[m] * bC1[m] + azBUp[m] * bUp[m];
    qb1[m] = c.adB1[m] * bC1[m] + c.axB1[m] * e1[m] + c.axA1[m] * aC1[m] + c.ayB1[m] * bN1[m] +
             bY1[m] * bP1[m] + c.azB1[m] * bDn[m] + c.azB0[m] * bC0[m];
  }
}

// KIND 0: r = b - A x, z = M^-1 r (LINE: z-line), partials rz (0), rr (1),
// bb (3).  KIND 1: xo = xi + sc d (HASD) or xi, r = b - A xo, partials rr
// (1), bb (3).
template <int R2, int KIND, bool LINE, bool HASD>
__global__ __launch_bounds__(kBlock) void k_sweep_p2(Grid g, March mj, Stencil<double> A,
                                                     const double* __restrict__ b,
                                                     double* xa, double* xb,
                                                     const float* __restrict__ d, double sc,
                                                     double* __restrict__ r,
                                                     double* __restrict__ z,
                                                     const double* __restrict__ invd,
                                                     const double* __restrict__ w,
                                                     double* __restrict__ parts, int pstride,
                                                     const PcgState* __restrict__ ctl) {
  const double* __restrict__ xi = xa;
  double* __restrict__ xo = xb;
  if (ctl) {
    if (ctl->skip) return;
    if (KIND == 1) {
      if (ctl->xInB) {
        xi = xb;
        xo = xa;
      }
      sc = 1.0 / ctl->s;
    }
  }
  const int lane = threadIdx.x & (kWave - 1);
  const int nz = g.nz;
  const int64_t sx = nz, sy = static_cast<int64_t>(g.nx2) * nz;
  KSum rz, rr, bbs;
  int i, j0, j1;
  if (march_span2(g, mj, i, j0, j1)) {
    // x (and the correction d) of a column: loaded, then formed (xi + sc d)
    struct Xd {
      double v0[R2], v1[R2];
      float d0[R2], d1[R2];
    };
    auto loadx = [&](int64_t base, Xd& o) {
      load_pair<R2>(xi, base, lane, nz, o.v0, o.v1);
      if (KIND == 1 && HASD) load_pair<R2>(d, base, lane, nz, o.d0, o.d1);
    };
    auto formxd = [&](const Xd& o, double (&v0)[R2], double (&v1)[R2]) {
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        v0[m] = o.v0[m];
        v1[m] = o.v1[m];
        if (KIND == 1 && HASD) {
          v0[m] = v0[m] + sc * static_cast<double>(o.d0[m]);
          v1[m] = v1[m] + sc * static_cast<double>(o.d1[m]);
        }
      }
    };
    auto formx = [&](int64_t base, double (&v0)[R2], double (&v1)[R2]) {
      Xd o;
      loadx(base, o);
      formxd(o, v0, v1);
    };
    double aP0[R2], aP1[R2], aC0[R2], aC1[R2], aY0[R2], aY1[R2];
    double bP0[R2], bP1[R2], bC0[R2], bC1[R2], bY0[R2], bY1[R2];
    formx(g.c3(j0 - 1, i, 0), aP0, aP1);
    formx(g.c3(j0, i, 0), aC0, aC1);
    formx(g.c3(j0 - 1, i + 1, 0), bP0, bP1);
    formx(g.c3(j0, i + 1, 0), bC0, bC1);
    load_pair<R2>(A.Ayp, g.c3(j0 - 1, i, 0), lane, nz, aY0, aY1);
    load_pair<R2>(A.Ayp, g.c3(j0 - 1, i + 1, 0), lane, nz, bY0, bY1);
    for (int j = j0; j <= j1; ++j) {
      const int64_t ba = g.c3(j, i, 0), bb = ba + sx;
      // every load of the step ahead of the arithmetic (left to itself the
      // scheduler spreads them between waits)
      Xd xaN, xbN, xW, xE;
      loadx(ba + sy, xaN);
      loadx(bb + sy, xbN);
      loadx(ba - sx, xW);  // column i - 1
      loadx(bb + sx, xE);  // column i + 2
      double bA0[R2], bA1[R2], bB0[R2], bB1[R2];
      load_pair<R2>(b, ba, lane, nz, bA0, bA1);
      load_pair<R2>(b, bb, lane, nz, bB0, bB1);
      double ivA0[R2], ivA1[R2], wA0[R2], wA1[R2], ivB0[R2], ivB1[R2], wB0[R2], wB1[R2];
      if constexpr (KIND == 0) {
        load_pair<R2>(invd, ba, lane, nz, ivA0, ivA1);
        load_pair<R2>(invd, bb, lane, nz, ivB0, ivB1);
        if constexpr (LINE) {
          load_pair<R2>(w, ba, lane, nz, wA0, wA1);
          load_pair<R2>(w, bb, lane, nz, wB0, wB1);
        }
      }
      CoefPair<R2> cp;
      load_coef_pair<R2>(A, ba, sx, lane, nz, cp);
      __builtin_amdgcn_sched_barrier(0);
      double aN0[R2], aN1[R2], bN0[R2], bN1[R2], w0[R2], w1[R2], e0[R2], e1[R2];
      formxd(xaN, aN0, aN1);
      formxd(xbN, bN0, bN1);
      formxd(xW, w0, w1);
      formxd(xE, e0, e1);
      double qa0[R2], qa1[R2], qb0[R2], qb1[R2];
      stencil_pair_q<R2>(lane, aP0, aP1, aC0, aC1, aN0, aN1, bP0, bP1, bC0, bC1, bN0, bN1, w0,
                         w1, e0, e1, aY0, aY1, bY0, bY1, cp, qa0, qa1, qb0, qb1);
      double rA0[R2], rA1[R2], rB0[R2], rB1[R2];
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        const bool in = 128 * m + 2 * lane < nz;
        rA0[m] = in ? bA0[m] - qa0[m] : 0.0;
        rA1[m] = in ? bA1[m] - qa1[m] : 0.0;
        rB0[m] = in ? bB0[m] - qb0[m] : 0.0;
        rB1[m] = in ? bB1[m] - qb1[m] : 0.0;
        if (in) {
          rr.add(rA0[m] * rA0[m]);
          rr.add(rA1[m] * rA1[m]);
          rr.add(rB0[m] * rB0[m]);
          rr.add(rB1[m] * rB1[m]);
          bbs.add(bA0[m] * bA0[m]);
          bbs.add(bA1[m] * bA1[m]);
          bbs.add(bB0[m] * bB0[m]);
          bbs.add(bB1[m] * bB1[m]);
        }
      }
      store_pair<R2>(r, ba, lane, nz, rA0, rA1);
      store_pair<R2>(r, bb, lane, nz, rB0, rB1);
      if constexpr (KIND == 0) {
        double zA0[R2], zA1[R2], zB0[R2], zB1[R2];
        precond_pair_f<R2, LINE>(lane, ivA0, ivA1, wA0, wA1, rA0, rA1, zA0, zA1);
        precond_pair_f<R2, LINE>(lane, ivB0, ivB1, wB0, wB1, rB0, rB1, zB0, zB1);
        store_pair<R2>(z, ba, lane, nz, zA0, zA1);
        store_pair<R2>(z, bb, lane, nz, zB0, zB1);
#pragma unroll
        for (int m = 0; m < R2; ++m) {
          if (128 * m + 2 * lane < nz) {
            rz.add(rA0[m] * zA0[m]);
            rz.add(rA1[m] * zA1[m]);
            rz.add(rB0[m] * zB0[m]);
            rz.add(rB1[m] * zB1[m]);
          }
        }
      } else if constexpr (HASD) {
        store_pair<R2>(xo, ba, lane, nz, aC0, aC1);
        store_pair<R2>(xo, bb, lane, nz, bC0, bC1);
      }
#pragma unroll
      for (int m = 0; m < R2; ++m) {
        aP0[m] = aC0[m];
        aP1[m] = aC1[m];
        aC0[m] = aN0[m];
        aC1[m] = aN1[m];
        aY0[m] = cp.ayA0[m];
        aY1[m] = cp.ayA1[m];
        bP0[m] = bC0[m];
        bP1[m] = bC1[m];
        bC0[m] = bN0[m];
        bC1[m] = bN1[m];
        bY0[m] = cp.ayB0[m];
        bY1[m] = cp.ayB1[m];
      }
    }
  }
  if constexpr (KIND == 0) {
    double v[3] = {rz.value(), rr.value(), bbs.value()};
    const int slots[3] = {0, 1, 3};
    write_parts<3>(parts, pstride, slots, v);
  } else {
    double v[2] = {rr.value(), bbs.value()};
    const int slots[2] = {1, 3};
    write_parts<2>(parts, pstride, slots, v);
  }
}

// FP64 stencil -> FP32 stencil over interior cells.
__global__ __launch_bounds__(kBlock) void k_cast_stencil(Grid g, Stencil<double> A,
                                                         Stencil<float> B) {
  const int64_t n = static_cast<int64_t>(g.columns()) * g.nz;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int k = static_cast<int>(t % g.nz);
    const int64_t c = column_base(g, t / g.nz) + k;
    B.Ad[c] = static_cast<float>(A.Ad[c]);
    B.Axp[c] = static_cast<float>(A.Axp[c]);
    B.Ayp[c] = static_cast<float>(A.Ayp[c]);
    B.Azp[c] = static_cast<float>(A.Azp[c]);
  }
}

// Fixed-order reduction of nvals partial slots into out[0..nvals).
__global__ __launch_bounds__(kBlock) void k_reduce_parts(const double* __restrict__ parts,
                                                         int pstride, int nparts,
                                                         double* __restrict__ out) {
  double v[4];
  const int slots[4] = {0, 1, 2, 3};
  reduce_parts<4>(parts, pstride, nparts, slots, v);
  if (threadIdx.x == 0) {
    for (int q = 0; q < 4; ++q) out[q] = v[q];
  }
}

__global__ void k_state_reset(PcgState* st, double tol2, int limit) {
  st->done = 0;
  st->it = 0;
  st->tol2 = tol2;
  st->limit = limit;
}

// ---------------------------------------------------------------------------
// Device-side control.  The host-polled solve decides between launches on
// the host; these one-block kernels take the same decisions, in the same
// order and with the same arithmetic, on the device, so a solve is a fixed
// sequence of launches (capturable in a graph): a round (FP64) or cycle
// (mixed) of B direction/update pairs, whose launches past the end of the
// recursion return at once, then the control kernel that reads the result.
// Iterations inside a round are numbered from 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int round_its(const PcgState* st, int budget) {
  return st->done ? st->it : budget;
}

// FP64 start, after the first residual: tolerance, first decisions.
__global__ __launch_bounds__(kBlock) void k_dc_begin64(const double* __restrict__ parts,
                                                       int pstride, int nparts, PcgState* st,
                                                       double rtol, double atol, int maxIt) {
  double v[2];
  const int slots[2] = {1, 3};
  reduce_parts<2>(parts, pstride, nparts, slots, v);
  if (threadIdx.x != 0) return;
  const double rr = v[0], bb = v[1];
  const double tol = fmax(rtol * sqrt(bb), atol);
  st->tolF2 = tol * tol;
  st->tol2 = st->tolF2;
  st->rrTrue = rr;
  st->bbTrue = bb;
  st->total = 0;
  st->refinements = 0;
  st->stalls = 0;
  st->maxInner = 0;
  st->rounds = 0;
  st->firstInner = 0;
  st->laterInner = 0;
  st->fallback = 0;
  st->keepNew = 0;
  st->xInB = 0;
  st->converged = rr <= st->tolF2 ? 1 : 0;
  st->stop = (st->converged || !isfinite(rr) || maxIt <= 0) ? 1 : 0;
  st->gate = 1;
}

// Before an FP64 round: its launches are no-ops if the solve is over.
__global__ void k_dc_round64(PcgState* st, int maxIt) {
  st->skip = st->stop;
  st->done = st->stop;
  st->it = 0;
  st->limit = maxIt - st->total;
}

// After a round's launches: the x term an odd count left deferred (the p of
// launch its - 1: even launches write pEven, odd ones pOdd).  With COPY the
// last p is also left in pOdd, which the next cycle's first launch reads as
// p_old (mixed cycles continue the recursion).
template <class T, bool COPY>
__global__ __launch_bounds__(kBlock) void k_dc_flush(Grid g, T* __restrict__ x, T* pEven,
                                                     T* pOdd, const PcgState* __restrict__ st,
                                                     int budget) {
  if (st->skip) return;
  const int its = round_its(st, budget);
  if (!(its & 1)) return;
  const T alpha = static_cast<T>(st->alphaH[(its - 1) & 1]);
  const int64_t n = static_cast<int64_t>(g.columns()) * g.nz;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t c = column_base(g, t / g.nz) + t % g.nz;
    const T p = pEven[c];
    x[c] = x[c] + alpha * p;
    if (COPY) pOdd[c] = p;
  }
}

// After an FP64 round's true residual: the host loop's decisions.
__global__ __launch_bounds__(kBlock) void k_dc_end64(const double* __restrict__ parts,
                                                     int pstride, int nparts, PcgState* st,
                                                     int budget, int maxIt) {
  double v[2];
  const int slots[2] = {1, 3};
  reduce_parts<2>(parts, pstride, nparts, slots, v);
  if (threadIdx.x != 0 || st->skip) return;
  const int its = round_its(st, budget);
  st->total += its;
  st->maxInner = max(st->maxInner, its);
  if (st->rounds == 0) st->firstInner = its;
  else st->laterInner = max(st->laterInner, its);
  st->rounds += 1;
  const double rr = v[0];
  st->rrTrue = rr;
  if (rr <= st->tolF2) {
    st->converged = 1;
    st->stop = 1;
  } else if (!isfinite(rr) || st->total >= maxIt) {
    st->stop = 1;
  } else if (its == 0 && ++st->stalls > 2) {
    st->stop = 1;
  }
}

// Mixed start, after the first FP64 residual: tolerance, scale, decisions.
__global__ __launch_bounds__(kBlock) void k_dc_begin_mixed(const double* __restrict__ parts,
                                                           int pstride, int nparts, PcgState* st,
                                                           double rtol, double atol, int maxIt,
                                                           int maxRef, double ncell) {
  double v[2];
  const int slots[2] = {1, 3};
  reduce_parts<2>(parts, pstride, nparts, slots, v);
  if (threadIdx.x != 0) return;
  const double rr = v[0], bb = v[1];
  const double tol = fmax(rtol * sqrt(bb), atol);
  st->tolF2 = tol * tol;
  st->rrTrue = rr;
  st->bbTrue = bb;
  st->s = rr > 0.0 ? 1.0 / sqrt(rr / ncell) : 1.0;
  st->total = 0;
  st->refinements = 0;
  st->stalls = 0;
  st->maxInner = 0;
  st->rounds = 0;
  st->firstInner = 0;
  st->laterInner = 0;
  st->fallback = 0;
  st->keepNew = 0;
  st->xInB = 0;
  st->converged = rr <= st->tolF2 ? 1 : 0;
  st->stop = (st->converged || !isfinite(rr) || maxRef <= 0 || maxIt <= 0) ? 1 : 0;
  st->gate = 1;
}

// Before a mixed cycle: the FP32 recursion's tolerance and iteration cap.
__global__ void k_dc_cycle(PcgState* st, int maxIt, double innerRtol) {
  st->skip = st->stop;
  st->done = st->stop;
  st->it = 0;
  if (st->stop) return;
  const double s = st->s;
  const double rrS = st->rrTrue * s * s;
  st->tol2 = fmax(innerRtol * innerRtol * rrS, 0.25 * st->tolF2 * s * s);
  st->limit = min(maxIt - st->total, 400);
  st->rrPrev = st->rrTrue;
}

// After a cycle's launches: count, and leave the last rz where the next
// cycle's first launch (iteration 0 of its numbering) reads the previous one.
__global__ void k_dc_mid(PcgState* st, int budget) {
  if (st->skip) return;
  const int its = round_its(st, budget);
  st->total += its;
  st->maxInner = max(st->maxInner, its);
  if (st->rounds == 0) st->firstInner = its;
  else st->laterInner = max(st->laterInner, its);
  st->rounds += 1;
  if (its > 0) st->rz[1] = st->rz[(its - 1) & 1];
}

// After a cycle's FP64 residual: the host loop's decisions.
__global__ __launch_bounds__(kBlock) void k_dc_end_mixed(const double* __restrict__ parts,
                                                         int pstride, int nparts, PcgState* st,
                                                         int maxIt, int maxRef) {
  double v[2];
  const int slots[2] = {1, 3};
  reduce_parts<2>(parts, pstride, nparts, slots, v);
  if (threadIdx.x != 0 || st->skip) return;
  const double rr = v[0];
  st->refinements += 1;
  if (!isfinite(rr) || rr > 0.25 * st->rrPrev) {
    st->fallback = 1;
    st->keepNew = (isfinite(rr) && rr < st->rrPrev) ? 1 : 0;
    st->stop = 1;
    return;
  }
  st->xInB ^= 1;
  st->rrTrue = rr;
  if (rr <= st->tolF2) {
    st->converged = 1;
    st->stop = 1;
  } else if (st->refinements >= maxRef || st->total >= maxIt) {
    st->stop = 1;
  }
}

__global__ void k_dc_finish(PcgState* st) { st->gate = st->converged ? 0 : 1; }

// After the direct column solve: one solve with one refinement.
__global__ void k_dc_direct(PcgState* st) {
  st->total = 1;
  st->refinements = 1;
  st->maxInner = 1;
}

// Converged mixed solve whose iterate is in the second buffer: x = xb.
__global__ __launch_bounds__(kBlock) void k_dc_take_x(Grid g, double* __restrict__ x,
                                                      const double* __restrict__ xb,
                                                      const PcgState* __restrict__ st) {
  if (!(st->converged && st->xInB)) return;
  const int64_t n = g.cellsH();
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[c] = xb[c];
}

template <class F>
void dispatch_R(int nz, F&& f) {
  const int R = (nz + kWave - 1) / kWave;
  switch (R) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    // 513 .. 1024 rows: one deep form (rows past the column are masked)
    case 9: case 10: case 11: case 12: case 13: case 14: case 15: case 16:
      f(std::integral_constant<int, 16>{});
      break;
    default: throw std::runtime_error("nz > 1024 is not supported by the column kernels");
  }
}

template <class F>
void dispatch_R2(int nz, F&& f) {
  switch ((nz + 127) / 128) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: case 6: case 7: case 8: f(std::integral_constant<int, 8>{}); break;
    default: throw std::runtime_error("nz > 1024 is not supported by the column kernels");
  }
}

template <class T>
T* dalloc(int64_t n) {
  void* p = nullptr;
  HIP_OK(hipMalloc(&p, static_cast<size_t>(n) * sizeof(T)));
  HIP_OK(hipMemset(p, 0, static_cast<size_t>(n) * sizeof(T)));
  return static_cast<T*>(p);
}

}  // namespace

__global__ void k_fill_vectors(int64_t n, double* __restrict__ a, double* __restrict__ b,
                               double* __restrict__ c, double* __restrict__ d) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    a[i] = 0.0;
    b[i] = 0.0;
    c[i] = 0.0;
    d[i] = 0.0;
  }
}

int device_cus() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
      cus > 0)
    return cus;
  return 256;
}

void Pcg::initCommon() {
  if (g_.nz > 1024) throw std::runtime_error("nz > 1024 is not supported by the column kernels");
  cus_ = device_cus();
  pairs_ = g_.nz % 2 == 0;
  cw2_ = pairs_ && g_.nx % 2 == 0;
  quads_ = cw2_ && g_.nz % 4 == 0 && g_.nz > 64 && g_.nz <= 128;
  defaultShapes(g_, &wpc64_, &wpc32_);
  applyShapes();
}

// A vector slot of the external workspace: whole 256-byte lines.
constexpr int64_t padded(int64_t n) { return (n + 31) / 32 * 32; }

int64_t Pcg::workspaceDoubles(const Grid& g) {
  const int cus = device_cus();
  int wpc64 = 16, wpc32 = 32;
  defaultShapes(g, &wpc64, &wpc32);
  const Pcg::Shapes sh = shapesFor(g, cus, 16, wpc64, wpc32);
  return 7 * padded(g.cellsH()) + padded(4 * static_cast<int64_t>(sh.pstride)) + padded(4) +
         padded((sizeof(PcgState) + sizeof(double) - 1) / sizeof(double));
}

size_t Pcg::hostStateBytes() { return sizeof(PcgState); }

Pcg::Pcg(const Grid& g, hipStream_t stream, const PcgWorkspace& ws)
    : g_(g), stream_(stream), external_(true) {
  initCommon();
  const int64_t nH = g.cellsH();
  int64_t off = 0;
  auto take = [&](int64_t n) {
    double* p = ws.device + off;
    off += padded(n);
    return p;
  };
  r_ = take(nH);
  z_ = take(nH);
  p0_ = take(nH);
  p1_ = take(nH);
  q_ = take(nH);
  invd_ = take(nH);
  w_ = take(nH);
  parts_ = take(4 * static_cast<int64_t>(pstride_));
  partsCap_ = pstride_;
  red_ = take(4);
  state_ = take((sizeof(PcgState) + sizeof(double) - 1) / sizeof(double));
  if (!ws.device || !ws.hostRed || !ws.hostState || off > ws.doubles)
    throw std::runtime_error("PCG workspace too small for this grid");
  redHost_ = ws.hostRed;
  stateHost_ = ws.hostState;
}

void Pcg::clearVectors() {
  const int64_t nH = g_.cellsH();
  const int blocks = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(
      int64_t(cus_) * 8, (nH + kBlock - 1) / kBlock)));
  k_fill_vectors<<<blocks, kBlock, 0, stream_>>>(nH, r_, z_, p0_, p1_);
  HIP_OK(hipGetLastError());
}

Pcg::Pcg(const Grid& g, hipStream_t stream) : g_(g), stream_(stream) {
  initCommon();
  const int64_t nH = g.cellsH();
  r_ = dalloc<double>(nH);
  z_ = dalloc<double>(nH);
  p0_ = dalloc<double>(nH);
  p1_ = dalloc<double>(nH);
  q_ = dalloc<double>(nH);
  invd_ = dalloc<double>(nH);
  w_ = dalloc<double>(nH);
  parts_ = dalloc<double>(4 * static_cast<int64_t>(pstride_));
  partsCap_ = pstride_;
  red_ = dalloc<double>(4);
  state_ = dalloc<PcgState>(1);
  HIP_OK(hipHostMalloc(&redHost_, 4 * sizeof(double), hipHostMallocDefault));
  HIP_OK(hipHostMalloc(&stateHost_, sizeof(PcgState), hipHostMallocDefault));
}

Pcg::~Pcg() {
  for (auto& e : evs_) (void)hipEventDestroy(e);
  if (external_) return;
  for (void* p : {(void*)r_, (void*)z_, (void*)p0_, (void*)p1_, (void*)q_, (void*)invd_, (void*)w_,
                  (void*)Ad32_, (void*)Axp32_, (void*)Ayp32_, (void*)Azp32_, (void*)r32_,
                  (void*)z32_, (void*)p032_, (void*)p132_, (void*)q32_, (void*)d32_,
                  (void*)invd32_, (void*)w32_, (void*)parts_, (void*)red_, state_, (void*)order_, (void*)chebD32_})
    if (p) (void)hipFree(p);
  if (redHost_) (void)hipHostFree(redHost_);
  if (stateHost_) (void)hipHostFree(stateHost_);
}

double Pcg::spmvBytesPerLaunch(bool mixed) const {
  // dir kernel per interior cell: z, p_old, Ad, Axp, Ayp, Azp in; p, q out.
  const double per = mixed ? 8.0 * 4.0 : 8.0 * 8.0;
  return per * static_cast<double>(g_.columns()) * g_.nz;
}

// The direction kernels over column pairs (nx and nz even) stream best at one
// wave per SIMD, where each wave has the memory pipeline to itself (sweep at
// 512 x 512 x 100: FP64 0.317 ms at 4 waves per CU against 0.337-0.342 at
// 8-16; FP32 0.174 against 0.181-0.186); the one-column forms at 16 (FP64)
// and 32 (FP32).
void Pcg::defaultShapes(const Grid& g, int* wpc64, int* wpc32) {
  *wpc64 = colPairs64(g) ? 4 : 16;
  *wpc32 = g.nx % 2 == 0 && g.nz % 2 == 0 && g.nz > 64 ? 4 : 32;
}

Pcg::Shapes Pcg::shapesFor(const Grid& g, int cus, int wpcCol, int wpc64, int wpc32) {
  Shapes sh;
  const int64_t ncol = g.columns();
  const int64_t wavesWanted = std::max<int64_t>(8, std::min<int64_t>(ncol, int64_t(cus) * wpcCol));
  sh.blocks = static_cast<int>(std::max<int64_t>(8, ((wavesWanted + kWpb - 1) / kWpb + 7) / 8 * 8));
  const bool cw2 = g.nz % 2 == 0 && g.nx % 2 == 0;
  // the one-column FP64 kernels (and, without column pairs, the direction
  // kernels) run one column a wave
  sh.march64 = makeMarch(g, cus, colPairs64(g) ? 16 : wpc64);
  sh.march32 = makeMarch(g, cus, cw2 ? 32 : wpc32);
  sh.march32w = makeMarch(g, cus, wpc32, 2);
  sh.march64w = makeMarch(g, cus, wpc64, 2);
  // the FP64 sweeps at 8 waves a CU (512 x 512 x 100, in the step's
  // sequence: residual 0.43-0.46 ms and refinements 0.36 / 0.38 ms, where 4
  // waves take 0.45-0.46 / 0.53 / 0.36 and 12 waves 0.45-0.47 / 0.33 / 0.39)
  sh.sweep64w = makeMarch(g, cus, 8, 2);
  sh.pstride = std::max({sh.blocks, sh.march64.blocks, sh.march32.blocks, sh.march32w.blocks,
                         sh.march64w.blocks, sh.sweep64w.blocks});
  return sh;
}

void Pcg::applyShapes() {
  const Shapes sh = shapesFor(g_, cus_, wpcCol_, wpc64_, wpc32_);
  blocks_ = sh.blocks;
  march64_ = sh.march64;
  march32_ = sh.march32;
  march32w_ = sh.march32w;
  march64w_ = sh.march64w;
  sweep64w_ = sh.sweep64w;
  pstride_ = std::max(sh.pstride, colBlocks_);
  if (parts_ && pstride_ > partsCap_ && external_)
    throw std::runtime_error("launch shapes beyond the external workspace's partial sums");
  if (parts_ && pstride_ > partsCap_) {
    HIP_OK(hipStreamSynchronize(stream_));
    HIP_OK(hipFree(parts_));
    parts_ = nullptr;
    parts_ = dalloc<double>(4 * static_cast<int64_t>(pstride_));
    partsCap_ = pstride_;
  }
}

void Pcg::setShapes(int march64WavesPerCu, int march32WavesPerCu, int columnWavesPerCu) {
  if (march64WavesPerCu > 0) wpc64_ = march64WavesPerCu;
  if (march32WavesPerCu > 0) wpc32_ = march32WavesPerCu;
  if (columnWavesPerCu > 0) wpcCol_ = columnWavesPerCu;
  applyShapes();
}

March Pcg::makeMarch(const Grid& g, int cus, int wavesPerCu, int columnsPerWave) {
  March mj;
  const int64_t perSlab = std::max(1, g.nx / columnsPerWave);
  const int64_t target = static_cast<int64_t>(cus) * wavesPerCu;
  mj.nslab = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(g.ny, target / perSlab)));
  mj.S = (g.ny + mj.nslab - 1) / mj.nslab;
  mj.nslab = (g.ny + mj.S - 1) / mj.S;  // every slab starts inside the grid
  const int64_t waves = perSlab * mj.nslab;
  mj.blocks = static_cast<int>(std::max<int64_t>(8, ((waves + kWpb - 1) / kWpb + 7) / 8 * 8));
  return mj;
}

void Pcg::reduceHost(int nparts, int nvals, double* out) {
  k_reduce_parts<<<1, kBlock, 0, stream_>>>(parts_, pstride_, nparts, red_);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(redHost_, red_, 4 * sizeof(double), hipMemcpyDeviceToHost, stream_));
  HIP_OK(hipStreamSynchronize(stream_));
  std::memcpy(out, redHost_, static_cast<size_t>(nvals) * sizeof(double));
}

int Pcg::solve(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
               PcgStats* st) {
  *st = PcgStats{};
  choosePairs64(A, b, x);
  if (direct_ && opt.zline) return solveDirect(A, b, x, opt, st);
  return opt.mixed ? solveMixed(A, b, x, opt, st) : solveFp64(A, b, x, opt, st);
}

// ---------------------------------------------------------------------------
// Direct column solve (laterally decoupled operators).
// ---------------------------------------------------------------------------
void Pcg::setColumnSolve(const int* kt, const std::vector<int>& ktHost) {
  if (external_) throw std::runtime_error("the column solve needs its own workspace");
  kt_ = kt;
  // R = ceil(active rows / 64), 1 .. 8, and one deep bucket (R = 16) for
  // 513 .. 1024 rows
  std::vector<std::vector<int>> byR(17);
  for (int j = 1; j <= g_.ny; ++j)
    for (int i = 1; i <= g_.nx; ++i) {
      const int L = g_.nz - ktHost[static_cast<size_t>(g_.c2(j, i))];
      if (L <= 0) continue;
      const int R = (L + kWave - 1) / kWave;
      byR[R <= 8 ? R : 16].push_back((j - 1) * g_.nx + (i - 1));
    }
  bucketCols_.clear();
  bucketCount_.clear();
  bucketR_.clear();
  for (int R = 16; R >= 1; --R) {  // deepest first: the longest waves start first
    if (byR[R].empty()) continue;
    bucketR_.push_back(R);
    bucketCount_.push_back(static_cast<int>(byR[R].size()));
    bucketCols_.insert(bucketCols_.end(), byR[R].begin(), byR[R].end());
  }
  bucketCpw_.assign(bucketR_.size(), 1);
  if (order_) HIP_OK(hipFree(order_));
  order_ = dalloc<int>(std::max<int64_t>(1, static_cast<int64_t>(bucketCols_.size())));
  if (!bucketCols_.empty())
    HIP_OK(hipMemcpy(order_, bucketCols_.data(), bucketCols_.size() * sizeof(int),
                     hipMemcpyHostToDevice));
  direct_ = true;
  buildColumnTable();
}

void Pcg::buildColumnTable() {
  tab_ = ColumnTable{};
  tab_.buckets = static_cast<int>(bucketR_.size());
  int col = 0, wave = 0;
  for (int b = 0; b < tab_.buckets; ++b) {
    tab_.R[b] = bucketR_[b];
    tab_.cpw[b] = bucketCpw_[b];
    tab_.colStart[b] = col;
    tab_.waveStart[b] = wave;
    col += bucketCount_[b];
    wave += (bucketCount_[b] + bucketCpw_[b] - 1) / bucketCpw_[b];
  }
  tab_.colStart[tab_.buckets] = col;
  tab_.waveStart[tab_.buckets] = wave;
  colBlocks_ = std::max(1, (wave + kWpb - 1) / kWpb);
  applyShapes();
}

void Pcg::setColumnTiles(const int* cpw, int n) {
  if (!kt_) throw std::runtime_error("no column solve on this operator");
  if (n != static_cast<int>(bucketR_.size()))
    throw std::runtime_error("one tile per depth bucket is expected");
  for (int b = 0; b < n; ++b)
    if (cpw[b] < 1 || cpw[b] > 64) throw std::runtime_error("columns per wave must be in [1, 64]");
  for (int b = 0; b < n; ++b) bucketCpw_[b] = cpw[b];
  buildColumnTable();
}

int Pcg::columnTiles(int* R, int* counts, int* cpw, int cap) const {
  const int n = static_cast<int>(bucketR_.size());
  for (int b = 0; b < n && b < cap; ++b) {
    if (R) R[b] = bucketR_[b];
    if (counts) counts[b] = bucketCount_[b];
    if (cpw) cpw[b] = bucketCpw_[b];
  }
  return n;
}

void Pcg::launchColumnSolve(const Stencil<double>& A, const double* b, double* x) {
  // buckets are deepest first; no bucket (no active cell) runs one block of
  // idle waves that writes zero partials
  auto go = [&](auto RC) {
    k_column_solve<decltype(RC)::value><<<colBlocks_, kBlock, 0, stream_>>>(
        g_, A, b, x, kt_, order_, tab_, parts_, pstride_);
  };
  switch (tab_.buckets ? tab_.R[0] : 1) {
    case 1: go(std::integral_constant<int, 1>{}); break;
    case 2: go(std::integral_constant<int, 2>{}); break;
    case 3: go(std::integral_constant<int, 3>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 5: go(std::integral_constant<int, 5>{}); break;
    case 6: go(std::integral_constant<int, 6>{}); break;
    case 7: go(std::integral_constant<int, 7>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    default: go(std::integral_constant<int, 16>{}); break;
  }
  HIP_OK(hipGetLastError());
}

int Pcg::solveDirect(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                     PcgStats* st) {
  launchColumnSolve(A, b, x);
  double red[4];
  reduceHost(colBlocks_, 4, red);
  const double rr = red[1], bb = red[3];
  const double tol = std::max(opt.rtol * std::sqrt(bb), opt.atol);
  st->rhs_norm = std::sqrt(bb);
  st->refinements = 1;
  if (rr <= tol * tol) {
    st->iterations = 1;
    st->converged = 1;
    st->residual_norm = std::sqrt(rr);
    return FREHG_OK;
  }
  // the PCG finishes from the direct solution
  launchFactor(A, opt.zline, false);
  const double rr2 = residualHost(A, b, x, opt.zline).first;
  return fp64Loop(A, b, x, opt, st, rr2, tol * tol, 1, 0);
}

// ---------------------------------------------------------------------------
// Launch helpers shared by the host-polled and the device-controlled solves.
// ---------------------------------------------------------------------------
// The FP64 iteration kernels take the pair layout when nz is even, above 64
// (a column of 64 rows or fewer fills half a wave of row pairs, and one row a
// lane fills all of it), at most 512 (deeper columns spill the pair kernels'
// registers to scratch) and every FP64 operand of the solve is 16-byte
// aligned (the pairs are loaded and stored as 16-byte vectors); the PCG's own
// vectors always are.
void Pcg::choosePairs64(const Stencil<double>& A, const double* b, const double* x) {
  auto al = [](const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  pairs64_ = pairForm64(g_) && al(A.Ad) && al(A.Axp) && al(A.Ayp) && al(A.Azp) && al(b) &&
             al(x);
}

template <class T>
void Pcg::launchDir(const Stencil<T>& A, const T* z, const T* pOld, T* pNew, T* q, int np, int n,
                    bool restart, const March& mj) {
  PcgState* dst = static_cast<PcgState*>(state_);
  constexpr bool fp32 = std::is_same<T, float>::value;
  if (fp32 ? pairs_ : pairs64_) {
    // FP64 p and q are stored non-temporally (0.319 -> 0.316 ms a launch at
    // 512 x 512 x 100); FP32 gains nothing from it
    dispatch_R2(g_.nz, [&](auto RC) {
      constexpr int R2 = decltype(RC)::value;
      constexpr bool NTS = !fp32;
      if constexpr (fp32 || R2 <= 4) {
        const March& mw = fp32 ? march32w_ : march64w_;
        if constexpr (fp32) {
          if (quads_) {
            if (restart)
              k_pcg_dir_q<true><<<mw.blocks, kBlock, 0, stream_>>>(
                  g_, mw, A, z, pOld, pNew, q, parts_, pstride_, np, dst, n);
            else
              k_pcg_dir_q<false><<<mw.blocks, kBlock, 0, stream_>>>(
                  g_, mw, A, z, pOld, pNew, q, parts_, pstride_, np, dst, n);
            return;
          }
        }
        if (cw2(fp32) && restart)
          k_pcg_dir_p2<T, R2, true, NTS><<<mw.blocks, kBlock, 0, stream_>>>(
              g_, mw, A, z, pOld, pNew, q, parts_, pstride_, np, dst, n);
        else if (cw2(fp32))
          k_pcg_dir_p2<T, R2, false, NTS><<<mw.blocks, kBlock, 0, stream_>>>(
              g_, mw, A, z, pOld, pNew, q, parts_, pstride_, np, dst, n);
        else if (restart)
          k_pcg_dir_p<T, R2, true><<<mj.blocks, kBlock, 0, stream_>>>(
              g_, mj, A, z, pOld, pNew, q, parts_, pstride_, np, dst, n);
        else
          k_pcg_dir_p<T, R2, false><<<mj.blocks, kBlock, 0, stream_>>>(
              g_, mj, A, z, pOld, pNew, q, parts_, pstride_, np, dst, n);
      } else {
        throw std::runtime_error("FP64 pair kernels take nz <= 512");
      }
    });
    HIP_OK(hipGetLastError());
    return;
  }
  dispatch_R(g_.nz, [&](auto RC) {
    constexpr int R = decltype(RC)::value;
    if (restart)
      k_pcg_dir<T, R, true><<<mj.blocks, kBlock, 0, stream_>>>(g_, mj, A, z, pOld, pNew, q,
                                                               parts_, pstride_, np, dst, n);
    else
      k_pcg_dir<T, R, false><<<mj.blocks, kBlock, 0, stream_>>>(g_, mj, A, z, pOld, pNew, q,
                                                                parts_, pstride_, np, dst, n);
  });
  HIP_OK(hipGetLastError());
}

template <class T>
void Pcg::launchUpdate(T* x, const T* pPrev, const T* p, T* r, const T* q, T* z, const T* invd,
                       const T* w, bool zline, bool two, int n, const March& mj) {
  PcgState* dst = static_cast<PcgState*>(state_);
  constexpr bool fp32 = std::is_same<T, float>::value;
  const int nd = dirBlocks(fp32, mj);
  if constexpr (fp32) {
    if (quads_) {
      if (zline && two)
        k_pcg_update_q<true, 2><<<blocks_, kBlock, 0, stream_>>>(
            g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
      else if (zline)
        k_pcg_update_q<true, 0><<<blocks_, kBlock, 0, stream_>>>(
            g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
      else if (two)
        k_pcg_update_q<false, 2><<<blocks_, kBlock, 0, stream_>>>(
            g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
      else
        k_pcg_update_q<false, 0><<<blocks_, kBlock, 0, stream_>>>(
            g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
      HIP_OK(hipGetLastError());
      return;
    }
  }
  if (fp32 ? pairs_ : pairs64_) {
    dispatch_R2(g_.nz, [&](auto RC) {
      constexpr int R2 = decltype(RC)::value;
      if constexpr (!fp32 && R2 > 4) throw std::runtime_error("FP64 pair kernels take nz <= 512");
      else if (zline && two)
        k_pcg_update_p<T, R2, true, 2><<<blocks_, kBlock, 0, stream_>>>(
            g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
      else if (zline)
        k_pcg_update_p<T, R2, true, 0><<<blocks_, kBlock, 0, stream_>>>(
            g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
      else if (two)
        k_pcg_update_p<T, R2, false, 2><<<blocks_, kBlock, 0, stream_>>>(
            g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
      else
        k_pcg_update_p<T, R2, false, 0><<<blocks_, kBlock, 0, stream_>>>(
            g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
    });
    HIP_OK(hipGetLastError());
    return;
  }
  dispatch_R(g_.nz, [&](auto RC) {
    constexpr int R = decltype(RC)::value;
    if (zline && two)
      k_pcg_update<T, R, true, 2><<<blocks_, kBlock, 0, stream_>>>(
          g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
    else if (zline)
      k_pcg_update<T, R, true, 0><<<blocks_, kBlock, 0, stream_>>>(
          g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
    else if (two)
      k_pcg_update<T, R, false, 2><<<blocks_, kBlock, 0, stream_>>>(
          g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
    else
      k_pcg_update<T, R, false, 0><<<blocks_, kBlock, 0, stream_>>>(
          g_, x, pPrev, p, r, q, z, invd, w, parts_, pstride_, nd, dst, n);
  });
  HIP_OK(hipGetLastError());
}

void Pcg::launchFactor(const Stencil<double>& A, bool zline, bool fp32) {
  const int nFactorBlocks = (g_.columns() + kWave - 1) / kWave;
  if (fp32 && zline)
    k_factor<float, true><<<nFactorBlocks, kWave, 0, stream_>>>(g_, Ad32_, Azp32_, invd32_, w32_);
  else if (fp32)
    k_factor<float, false><<<nFactorBlocks, kWave, 0, stream_>>>(g_, Ad32_, Azp32_, invd32_, w32_);
  else if (zline)
    k_factor<double, true><<<nFactorBlocks, kWave, 0, stream_>>>(g_, A.Ad, A.Azp, invd_, w_);
  else
    k_factor<double, false><<<nFactorBlocks, kWave, 0, stream_>>>(g_, A.Ad, A.Azp, invd_, w_);
  HIP_OK(hipGetLastError());
}

void Pcg::launchResidual(const Stencil<double>& A, const double* b, const double* x, bool zline,
                         const void* ctl) {
  const PcgState* c = static_cast<const PcgState*>(ctl);
  if (sweepPairs()) {
    const March& mw = sweep64w_;
    double* xm = const_cast<double*>(x);  // read only in this form
    auto go = [&](auto RC) {
      constexpr int R2 = decltype(RC)::value;
      if (zline)
        k_sweep_p2<R2, 0, true, false><<<mw.blocks, kBlock, 0, stream_>>>(
            g_, mw, A, b, xm, xm, nullptr, 0.0, r_, z_, invd_, w_, parts_, pstride_, c);
      else
        k_sweep_p2<R2, 0, false, false><<<mw.blocks, kBlock, 0, stream_>>>(
            g_, mw, A, b, xm, xm, nullptr, 0.0, r_, z_, invd_, w_, parts_, pstride_, c);
    };
    if (g_.nz <= 128) go(std::integral_constant<int, 1>{});
    else go(std::integral_constant<int, 2>{});
    HIP_OK(hipGetLastError());
    return;
  }
  dispatch_R(g_.nz, [&](auto RC) {
    constexpr int R = decltype(RC)::value;
    const March& mj = march64_;
    if (zline)
      k_residual64<R, true><<<mj.blocks, kBlock, 0, stream_>>>(g_, mj, A, b, x, r_, z_, invd_, w_,
                                                               parts_, pstride_, c);
    else
      k_residual64<R, false><<<mj.blocks, kBlock, 0, stream_>>>(g_, mj, A, b, x, r_, z_, invd_,
                                                                w_, parts_, pstride_, c);
  });
  HIP_OK(hipGetLastError());
}

void Pcg::launchRefine(const Stencil<double>& A, const double* b, double* xa, double* xb,
                       double sc, bool hasD, const void* ctl) {
  const PcgState* c = static_cast<const PcgState*>(ctl);
  if (sweepPairs()) {
    const March& mw = sweep64w_;
    auto go = [&](auto RC) {
      constexpr int R2 = decltype(RC)::value;
      if (hasD)
        k_sweep_p2<R2, 1, false, true><<<mw.blocks, kBlock, 0, stream_>>>(
            g_, mw, A, b, xa, xb, d32_, sc, r_, nullptr, nullptr, nullptr, parts_, pstride_, c);
      else
        k_sweep_p2<R2, 1, false, false><<<mw.blocks, kBlock, 0, stream_>>>(
            g_, mw, A, b, xa, xb, d32_, sc, r_, nullptr, nullptr, nullptr, parts_, pstride_, c);
    };
    if (g_.nz <= 128) go(std::integral_constant<int, 1>{});
    else go(std::integral_constant<int, 2>{});
    HIP_OK(hipGetLastError());
    return;
  }
  dispatch_R(g_.nz, [&](auto RC) {
    constexpr int R = decltype(RC)::value;
    if (hasD)
      k_refine64<R, true><<<march64_.blocks, kBlock, 0, stream_>>>(
          g_, march64_, A, b, xa, xb, d32_, sc, r_, parts_, pstride_, c);
    else
      k_refine64<R, false><<<march64_.blocks, kBlock, 0, stream_>>>(
          g_, march64_, A, b, xa, xb, d32_, sc, r_, parts_, pstride_, c);
  });
  HIP_OK(hipGetLastError());
}

void Pcg::launchInnerInit(bool zline, double s, const void* ctl) {
  const PcgState* c = static_cast<const PcgState*>(ctl);
  if (pairs_) {
    dispatch_R2(g_.nz, [&](auto RC) {
      constexpr int R2 = decltype(RC)::value;
      if (zline)
        k_inner_init_p<R2, true><<<blocks_, kBlock, 0, stream_>>>(
            g_, r_, s, r32_, z32_, d32_, invd32_, w32_, parts_, pstride_, c);
      else
        k_inner_init_p<R2, false><<<blocks_, kBlock, 0, stream_>>>(
            g_, r_, s, r32_, z32_, d32_, invd32_, w32_, parts_, pstride_, c);
    });
  } else {
    dispatch_R(g_.nz, [&](auto RC) {
      constexpr int R = decltype(RC)::value;
      if (zline)
        k_inner_init<R, true><<<blocks_, kBlock, 0, stream_>>>(g_, r_, s, r32_, z32_, d32_,
                                                               invd32_, w32_, parts_, pstride_, c);
      else
        k_inner_init<R, false><<<blocks_, kBlock, 0, stream_>>>(g_, r_, s, r32_, z32_, d32_,
                                                                invd32_, w32_, parts_, pstride_, c);
    });
  }
  HIP_OK(hipGetLastError());
}

void Pcg::launchChebSteps(const PcgOptions& opt, const int* skip) {
  const double b = 2.0, a = opt.cheb_lower * b;
  const double theta = 0.5 * (b + a), delta = 0.5 * (b - a), sigma = theta / delta;
  const Stencil<float> A32{Ad32_, Axp32_, Ayp32_, Azp32_};
  const March& mj = march32_;
  double rhoPrev = 1.0 / sigma;
  float* zin = z32_;
  float* zout = q32_;
  for (int k = 1; k < opt.cheb_degree; ++k) {
    const double rho = 1.0 / (2.0 * sigma - rhoPrev);
    const float c1 = static_cast<float>(rho * rhoPrev), c2 = static_cast<float>(2.0 * rho / delta);
    const float invTheta = static_cast<float>(1.0 / theta);
    const bool first = k == 1, last = k + 1 == opt.cheb_degree;
    dispatch_R2(g_.nz, [&](auto RC) {
      constexpr int R2 = decltype(RC)::value;
      auto go = [&](auto kern) {
        kern<<<mj.blocks, kBlock, 0, stream_>>>(g_, mj, A32, zin, zout, chebD32_, r32_, invd32_,
                                                invTheta, c1, c2, parts_, pstride_, skip);
      };
      if (first && last) go(k_cheb_step_p<R2, true, true>);
      else if (first) go(k_cheb_step_p<R2, true, false>);
      else if (last) go(k_cheb_step_p<R2, false, true>);
      else go(k_cheb_step_p<R2, false, false>);
    });
    HIP_OK(hipGetLastError());
    std::swap(zin, zout);
    rhoPrev = rho;
  }
}

template <class T>
int Pcg::innerIterations(const Stencil<T>& A, T* x, T*& pOld, T*& pNew, T* r, T* z, T* q,
                         const T* invd, const T* w, bool zline, double tol2, int maxIt,
                         const March& mj, int nprev0, int n0, bool restart,
                         const PcgOptions& opt, PcgStats* st, int* itDone, double rr0) {
  // Iterations are numbered on from n0 so that a resumed run finds the
  // previous rz in the state slot the numbering expects.
  PcgState* dst = static_cast<PcgState*>(state_);
  PcgState* hst = static_cast<PcgState*>(stateHost_);
  int n = n0;
  maxIt += n0;
  k_state_reset<<<1, 1, 0, stream_>>>(dst, tol2, maxIt);
  HIP_OK(hipGetLastError());
  T* const bufOld = pOld;  // dir launch n writes p into (n - n0 even ? bufNew : bufOld)
  T* const bufNew = pNew;
  bool sampled[kPollMax] = {};
  const int every = std::max(1, opt.timing_every);
  if (opt.timing && evs_.empty()) {
    evs_.resize(2 * kPollMax);
    for (auto& e : evs_) HIP_OK(hipEventCreate(&e));
  }
  hipEvent_t* evs = evs_.data();
  int done = 0;
  // Launches between polls: kPoll, then the iterations the recursion's rate
  // since the last poll predicts it still needs, plus the launch that sees it
  // converge.  Fewer polls (each leaves the device idle while the host
  // reacts) and fewer launches queued past convergence; the iterates are
  // the same.
  int chunkLen = kPoll;
  double rrPrev = rr0;  // r.r entering iteration itPrev (< 0: not known)
  int itPrev = n0;
  while (true) {
    const int chunk = std::min(chunkLen, maxIt - n);
    if (chunk <= 0) break;
    for (int c = 0; c < chunk; ++c, ++n) {
      // an event pair around a launch delays the next dispatch by a few
      // microseconds, so only one launch in `every` is bracketed, mid-way
      // through each group (the first launch of a solve is the restart form)
      sampled[c] = opt.timing && (n - n0) % every == every / 2;
      if (sampled[c]) HIP_OK(hipEventRecord(evs[2 * c], stream_));
      const bool cheb = std::is_same<T, float>::value && opt.cheb;
      const int np = n == n0 ? nprev0 : (cheb ? mj.blocks : blocks_);
      launchDir<T>(A, z, pOld, pNew, q, np, n, n == n0 && restart, mj);
      if (sampled[c]) HIP_OK(hipEventRecord(evs[2 * c + 1], stream_));
      const bool two = ((n - n0) & 1) != 0;  // x takes this and the previous term
      launchUpdate<T>(x, pOld, pNew, r, q, z, invd, w, zline, two, n, mj);
      if (cheb) launchChebSteps(opt, &dst->done);
      std::swap(pOld, pNew);
    }
    HIP_OK(hipMemcpyAsync(hst, dst, sizeof(PcgState), hipMemcpyDeviceToHost, stream_));
    HIP_OK(hipStreamSynchronize(stream_));
    if (opt.timing) {
      int active = hst->done ? hst->it - (n - chunk) : chunk;
      active = std::max(0, std::min(active, chunk));
      for (int c = 0; c < active; ++c) {
        if (!sampled[c]) continue;
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, evs[2 * c], evs[2 * c + 1]));
        st->spmv_ms += ms;
        st->spmv_launches += 1;
      }
    }
    if (hst->done) {
      done = 1;
      break;
    }
    const double rrNow = hst->rr;
    const int itNow = hst->it;
    chunkLen = kPoll;
    if (rrPrev > 0.0 && rrNow > 0.0 && rrNow < rrPrev && itNow > itPrev && tol2 > 0.0) {
      const double perIt = std::log(rrNow / rrPrev) / (itNow - itPrev);
      const double need = std::log(tol2 / rrNow) / perIt;
      if (std::isfinite(need))
        chunkLen = static_cast<int>(std::min<double>(kPollMax, std::max(2.0, std::ceil(need) + 1.0)));
    }
    rrPrev = rrNow;
    itPrev = itNow;
  }
  *itDone = (done ? hst->it : std::min(n, maxIt)) - n0;
  // Launches queued after convergence are no-ops, yet the loop swapped the
  // pointers for them too: point pOld at the last p actually formed.
  if (*itDone > 0) {
    const bool lastEven = ((*itDone - 1) & 1) == 0;
    pOld = lastEven ? bufNew : bufOld;
    pNew = lastEven ? bufOld : bufNew;
  } else {
    pOld = bufOld;
    pNew = bufNew;
  }
  // an odd count ends on an even iteration, whose x term was deferred
  if (*itDone & 1) {
    k_flush_x<T><<<blocks_ * 4, kBlock, 0, stream_>>>(g_, x, pOld, dst, (n0 + *itDone - 1) & 1);
    HIP_OK(hipGetLastError());
  }
  return done;
}

std::pair<double, double> Pcg::residualHost(const Stencil<double>& A, const double* b,
                                            const double* x, bool zline) {
  launchResidual(A, b, x, zline, nullptr);
  double red[4];
  reduceHost(sweepBlocks(), 4, red);
  return std::make_pair(red[1], red[3]);
}

int Pcg::solveFp64(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                   PcgStats* st) {
  launchFactor(A, opt.zline, false);
  const auto [rr, bb] = residualHost(A, b, x, opt.zline);
  const double tol = std::max(opt.rtol * std::sqrt(bb), opt.atol);
  st->rhs_norm = std::sqrt(bb);
  return fp64Loop(A, b, x, opt, st, rr, tol * tol, 0, 0);
}

// The outer loop of the FP64 solve from a true residual rr: rounds of
// recursion, each confirmed (or restarted) on the true residual.
int Pcg::fp64Loop(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                  PcgStats* st, double rr, double tol2, int total, int stalls) {
  double* pOld = p0_;
  double* pNew = p1_;
  while (true) {
    if (rr <= tol2) {
      st->converged = 1;
      st->residual_norm = std::sqrt(rr);
      break;
    }
    if (!std::isfinite(rr) || total >= opt.max_iterations) {
      st->residual_norm = std::sqrt(rr);
      break;
    }
    int its = 0;
    innerIterations<double>(A, x, pOld, pNew, r_, z_, q_, invd_, w_, opt.zline, tol2,
                            opt.max_iterations - total, march64_, sweepBlocks(), 0, true, opt,
                            st, &its, rr);
    total += its;
    if (st->rounds++ == 0) st->first_inner = its;
    else st->later_inner = std::max(st->later_inner, its);
    rr = residualHost(A, b, x, opt.zline).first;  // true residual (restart point if not converged)
    if (its == 0 && ++stalls > 2 && !(rr <= tol2)) {
      st->residual_norm = std::sqrt(rr);
      break;
    }
  }
  st->iterations = total;
  return st->converged ? FREHG_OK : FREHG_EDIVERGED;
}

Stencil<float> Pcg::stencil32() {
  if (external_) throw std::runtime_error("the mixed form allocates its FP32 workspace");
  const int64_t nH = g_.cellsH();
  if (!Ad32_) {
    Ad32_ = dalloc<float>(nH);
    Axp32_ = dalloc<float>(nH);
    Ayp32_ = dalloc<float>(nH);
    Azp32_ = dalloc<float>(nH);
    r32_ = dalloc<float>(nH);
    z32_ = dalloc<float>(nH);
    p032_ = dalloc<float>(nH);
    p132_ = dalloc<float>(nH);
    q32_ = dalloc<float>(nH);
    d32_ = dalloc<float>(nH);
    invd32_ = dalloc<float>(nH);
    w32_ = dalloc<float>(nH);
  }
  return Stencil<float>{Ad32_, Axp32_, Ayp32_, Azp32_};
}

std::pair<double, double> Pcg::refineHost(const Stencil<double>& A, const double* b, double* xi,
                                          double* xo, double sc, bool hasD) {
  launchRefine(A, b, xi, xo, sc, hasD, nullptr);
  double red[4];
  reduceHost(sweepBlocks(), 4, red);
  return std::make_pair(red[1], red[3]);
}

void Pcg::mixedSetup(const Stencil<double>& A, const PcgOptions& opt) {
  const Stencil<float> A32 = stencil32();
  if (opt.cheb) {
    if (!pairs_) throw std::runtime_error("the Chebyshev preconditioner needs an even nz");
    if (opt.zline || opt.cheb_degree < 3 || opt.cheb_degree > 15 || !(opt.cheb_degree & 1) ||
        !(opt.cheb_lower > 0.0 && opt.cheb_lower < 1.0))
      throw std::runtime_error("Chebyshev: odd degree in [3, 15], lower ratio in (0, 1)");
    if (!chebD32_) chebD32_ = dalloc<float>(g_.cellsH());
  }
  if (!opt.stencil32_ready) {
    k_cast_stencil<<<blocks_ * 4, kBlock, 0, stream_>>>(g_, A, A32);
    HIP_OK(hipGetLastError());
  }
  launchFactor(A, opt.zline, true);
}

int Pcg::solveMixed(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                    PcgStats* st) {
  mixedSetup(A, opt);
  // x ping-pongs between the caller's array and the FP64 q_ buffer (idle in this form).
  MixedCarry c;
  c.xa = x;
  c.xb = q_;
  const auto [rr, bb] = refineHost(A, b, c.xa, c.xb, 0.0, false);
  const double tol = std::max(opt.rtol * std::sqrt(bb), opt.atol);
  c.rr = rr;
  c.tol2 = tol * tol;
  st->rhs_norm = std::sqrt(bb);
  const double ncell = static_cast<double>(g_.columns()) * g_.nz;
  // Reliable updates: the FP32 CG runs on s r (one fixed scale, so the FP32
  // iterates are O(1) at the start); whenever its recursive residual has
  // fallen by inner_rtol, or to the final tolerance, the correction is added
  // to x in FP64, r = b - A x is recomputed in FP64 and handed back to the
  // FP32 recursion, which carries on with its search direction instead of
  // restarting, so the Krylov space built so far is kept.
  c.s = rr > 0.0 ? 1.0 / std::sqrt(rr / ncell) : 1.0;
  c.pOld = p032_;
  c.pNew = p132_;
  return mixedLoop(A, b, x, opt, st, c);
}

int Pcg::mixedLoop(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                   PcgStats* st, MixedCarry c) {
  const int64_t nH = g_.cellsH();
  const Stencil<float> A32{Ad32_, Axp32_, Ayp32_, Azp32_};
  while (true) {
    if (c.rr <= c.tol2) {
      st->converged = 1;
      break;
    }
    if (!std::isfinite(c.rr) || st->refinements >= opt.max_refinements ||
        c.total >= opt.max_iterations)
      break;
    launchInnerInit(opt.zline, c.s, nullptr);
    if (opt.cheb) launchChebSteps(opt, nullptr);
    const double rrS = c.rr * c.s * c.s;
    const double tolIn2 =
        std::max(opt.inner_rtol * opt.inner_rtol * rrS, 0.25 * c.tol2 * c.s * c.s);
    int its = 0;
    innerIterations<float>(A32, d32_, c.pOld, c.pNew, r32_, z32_, q32_, invd32_, w32_, opt.zline,
                           tolIn2, std::min(opt.max_iterations - c.total, 400), march32_,
                           opt.cheb ? march32_.blocks : blocks_, c.n0, c.restart, opt, st, &its,
                           rrS);
    c.total += its;
    c.n0 += its;
    c.restart = false;
    if (st->rounds++ == 0) st->first_inner = its;
    else st->later_inner = std::max(st->later_inner, its);
    const double rrPrev = c.rr;
    c.rr = refineHost(A, b, c.xa, c.xb, 1.0 / c.s, true).first;
    st->refinements += 1;
    if (!std::isfinite(c.rr) || c.rr > 0.25 * rrPrev) {
      // The FP32 cycle did not halve the FP64 residual (an operator too
      // ill-conditioned for FP32, or a breakdown): finish in FP64 from the
      // better of the two iterates.
      const bool keepNew = std::isfinite(c.rr) && c.rr < rrPrev;
      return fp64Finish(A, b, x, opt, st, keepNew ? c.xb : c.xa, c.total);
    }
    std::swap(c.xa, c.xb);
  }
  if (c.xa != x)
    HIP_OK(hipMemcpyAsync(x, c.xa, static_cast<size_t>(nH) * sizeof(double),
                          hipMemcpyDeviceToDevice, stream_));
  st->iterations = c.total;
  st->residual_norm = std::sqrt(c.rr);
  return st->converged ? FREHG_OK : FREHG_EDIVERGED;
}

int Pcg::fp64Finish(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                    PcgStats* st, const double* best, int total) {
  const int64_t nH = g_.cellsH();
  if (best != x)
    HIP_OK(hipMemcpyAsync(x, best, static_cast<size_t>(nH) * sizeof(double),
                          hipMemcpyDeviceToDevice, stream_));
  PcgOptions o64 = opt;
  o64.mixed = false;
  o64.cheb = false;  // FP64 finish: point Jacobi in place of the FP32 polynomial
  o64.max_iterations = std::max(1, opt.max_iterations - total);
  PcgStats s64;
  const int rc = solveFp64(A, b, x, o64, &s64);
  st->fp64_iterations = s64.iterations;
  st->iterations = total + s64.iterations;
  st->residual_norm = s64.residual_norm;
  st->converged = s64.converged;
  st->spmv_ms += s64.spmv_ms;
  st->spmv_launches += s64.spmv_launches;
  return rc;
}

// ---------------------------------------------------------------------------
// Device-controlled solve: the host-polled solve's launches and decisions as
// a fixed sequence (rounds / cycles of `iterations` launch pairs each), no
// host synchronization, ending with a copy of the control state to the
// host mirror.  Capturable in a graph.
// ---------------------------------------------------------------------------
void Pcg::prepareDevice(const PcgOptions& opt) {
  if (opt.mixed && !(direct_ && opt.zline)) {  // allocated before any capture
    (void)stencil32();
    if (opt.cheb && !chebD32_) chebD32_ = dalloc<float>(g_.cellsH());
  }
}

void Pcg::enqueueDevice(const Stencil<double>& A, const double* b, double* x,
                        const PcgOptions& opt, const PcgBudget& bud) {
  if (bud.rounds < 1 || bud.iterations < 1 || bud.iterations > 4096)
    throw std::runtime_error("device PCG budget out of range");
  choosePairs64(A, b, x);
  PcgState* dst = static_cast<PcgState*>(state_);
  if (bud.later < 0 || bud.later > 4096) throw std::runtime_error("device PCG budget out of range");
  auto budgetOf = [&](int round) {
    return round == 0 || bud.later <= 0 ? bud.iterations : bud.later;
  };
  if (direct_ && opt.zline) {
    launchColumnSolve(A, b, x);
    k_dc_begin64<<<1, kBlock, 0, stream_>>>(parts_, pstride_, colBlocks_, dst, opt.rtol, opt.atol,
                                            opt.max_iterations);
    k_dc_direct<<<1, 1, 0, stream_>>>(dst);
    HIP_OK(hipGetLastError());
  } else if (!opt.mixed) {
    launchFactor(A, opt.zline, false);
    launchResidual(A, b, x, opt.zline, nullptr);
    k_dc_begin64<<<1, kBlock, 0, stream_>>>(parts_, pstride_, sweepBlocks(), dst, opt.rtol,
                                            opt.atol, opt.max_iterations);
    HIP_OK(hipGetLastError());
    for (int round = 0; round < bud.rounds; ++round) {
      const int B = budgetOf(round);
      k_dc_round64<<<1, 1, 0, stream_>>>(dst, opt.max_iterations);
      HIP_OK(hipGetLastError());
      double* pOld = p0_;
      double* pNew = p1_;
      for (int n = 0; n < B; ++n) {
        launchDir<double>(A, z_, pOld, pNew, q_, n == 0 ? sweepBlocks() : blocks_, n, n == 0,
                          march64_);
        launchUpdate<double>(x, pOld, pNew, r_, q_, z_, invd_, w_, opt.zline, (n & 1) != 0, n,
                             march64_);
        std::swap(pOld, pNew);
      }
      k_dc_flush<double, false><<<blocks_ * 4, kBlock, 0, stream_>>>(g_, x, p1_, p0_, dst, B);
      HIP_OK(hipGetLastError());
      launchResidual(A, b, x, opt.zline, dst);
      k_dc_end64<<<1, kBlock, 0, stream_>>>(parts_, pstride_, sweepBlocks(), dst, B,
                                            opt.max_iterations);
      HIP_OK(hipGetLastError());
    }
  } else {
    mixedSetup(A, opt);
    const Stencil<float> A32{Ad32_, Axp32_, Ayp32_, Azp32_};
    launchRefine(A, b, x, q_, 0.0, false, nullptr);
    const double ncell = static_cast<double>(g_.columns()) * g_.nz;
    k_dc_begin_mixed<<<1, kBlock, 0, stream_>>>(parts_, pstride_, sweepBlocks(), dst, opt.rtol,
                                                opt.atol, opt.max_iterations,
                                                opt.max_refinements, ncell);
    HIP_OK(hipGetLastError());
    for (int cycle = 0; cycle < bud.rounds; ++cycle) {
      const int B = budgetOf(cycle);
      k_dc_cycle<<<1, 1, 0, stream_>>>(dst, opt.max_iterations, opt.inner_rtol);
      HIP_OK(hipGetLastError());
      launchInnerInit(opt.zline, 0.0, dst);
      if (opt.cheb) launchChebSteps(opt, &dst->skip);
      float* pOld = p032_;
      float* pNew = p132_;
      const int np = opt.cheb ? march32_.blocks : blocks_;
      for (int n = 0; n < B; ++n) {
        launchDir<float>(A32, z32_, pOld, pNew, q32_, np, n, cycle == 0 && n == 0, march32_);
        launchUpdate<float>(d32_, pOld, pNew, r32_, q32_, z32_, invd32_, w32_, opt.zline,
                            (n & 1) != 0, n, march32_);
        if (opt.cheb) launchChebSteps(opt, &dst->done);
        std::swap(pOld, pNew);
      }
      k_dc_flush<float, true><<<blocks_ * 4, kBlock, 0, stream_>>>(g_, d32_, p132_, p032_, dst,
                                                                   B);
      k_dc_mid<<<1, 1, 0, stream_>>>(dst, B);
      HIP_OK(hipGetLastError());
      launchRefine(A, b, x, q_, 0.0, true, dst);
      k_dc_end_mixed<<<1, kBlock, 0, stream_>>>(parts_, pstride_, sweepBlocks(), dst,
                                                opt.max_iterations, opt.max_refinements);
      HIP_OK(hipGetLastError());
    }
    k_dc_take_x<<<blocks_ * 4, kBlock, 0, stream_>>>(g_, x, q_, dst);
    HIP_OK(hipGetLastError());
  }
  k_dc_finish<<<1, 1, 0, stream_>>>(dst);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpyAsync(stateHost_, dst, sizeof(PcgState), hipMemcpyDeviceToHost, stream_));
}

const int* Pcg::gateFlag() const { return &static_cast<const PcgState*>(state_)->gate; }

int Pcg::deviceResult(PcgStats* st, DeviceSummary* sum) const {
  const PcgState* h = static_cast<const PcgState*>(stateHost_);
  *st = PcgStats{};
  st->iterations = h->total;
  st->refinements = h->refinements;
  st->converged = h->converged;
  st->residual_norm = std::sqrt(h->rrTrue);
  st->rhs_norm = std::sqrt(h->bbTrue);
  if (sum) {
    sum->max_inner = h->maxInner;
    sum->first_inner = h->firstInner;
    sum->later_inner = h->laterInner;
    sum->rounds = h->rounds;
    sum->stopped = h->stop;
    sum->fallback = h->fallback;
  }
  return h->gate;
}

int Pcg::resumeDevice(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                      PcgStats* st) {
  const PcgState h = *static_cast<const PcgState*>(stateHost_);
  *st = PcgStats{};
  choosePairs64(A, b, x);
  st->rhs_norm = std::sqrt(h.bbTrue);
  st->refinements = h.refinements;
  if (direct_ && opt.zline) {
    if (h.stop && !h.converged) {
      st->iterations = h.total;
      st->residual_norm = std::sqrt(h.rrTrue);
      return FREHG_EDIVERGED;
    }
    launchFactor(A, opt.zline, false);
    const double rr = residualHost(A, b, x, opt.zline).first;
    return fp64Loop(A, b, x, opt, st, rr, h.tolF2, h.total, 0);
  }
  if (!opt.mixed) {
    if (h.stop && !h.converged) {
      st->iterations = h.total;
      st->residual_norm = std::sqrt(h.rrTrue);
      return FREHG_EDIVERGED;
    }
    return fp64Loop(A, b, x, opt, st, h.rrTrue, h.tolF2, h.total, h.stalls);
  }
  MixedCarry c;
  c.xa = h.xInB ? q_ : x;
  c.xb = h.xInB ? x : q_;
  if (h.fallback) return fp64Finish(A, b, x, opt, st, h.keepNew ? c.xb : c.xa, h.total);
  if (h.stop) {
    if (c.xa != x)
      HIP_OK(hipMemcpyAsync(x, c.xa, static_cast<size_t>(g_.cellsH()) * sizeof(double),
                            hipMemcpyDeviceToDevice, stream_));
    st->iterations = h.total;
    st->residual_norm = std::sqrt(h.rrTrue);
    return st->converged ? FREHG_OK : FREHG_EDIVERGED;
  }
  // out of cycles: the host loop carries on from the device's state (the
  // last p is in p032_, the last rz in the slot an even numbering reads)
  c.rr = h.rrTrue;
  c.tol2 = h.tolF2;
  c.s = h.s;
  c.total = h.total;
  c.n0 = 0;
  c.restart = h.refinements == 0;
  c.pOld = p032_;
  c.pNew = p132_;
  return mixedLoop(A, b, x, opt, st, c);
}

template int Pcg::innerIterations<double>(const Stencil<double>&, double*, double*&, double*&,
                                          double*, double*, double*, const double*, const double*,
                                          bool, double, int, const March&, int, int, bool,
                                          const PcgOptions&, PcgStats*, int*, double);
template int Pcg::innerIterations<float>(const Stencil<float>&, float*, float*&, float*&, float*,
                                         float*, float*, const float*, const float*, bool, double,
                                         int, const March&, int, int, bool, const PcgOptions&,
                                         PcgStats*, int*, double);

}  // namespace frehg

// gw_ls.hip — the solve seam: a COO-assembled sparse system solved by PCG on
// the GPU, replacing frehg::LinearSystem (src/core/LinearSystem.hpp:109-151;
// PETSc MatSetPreallocationCOO / MatSetValuesCOO / KSPSolve at
// src/core/LinearSystem.cpp:294, 308, 320).
//
// setPattern (LinearSystem.cpp:273-296): entries with a -1 row or column are
// ignored, duplicates are merged.  Two paths:
//
//  * Structured.  The groundwater step hands the seam the fixed 7-point
//    pattern of RichardsSolver::buildCooPattern (src/gw/RichardsSolver.cpp:
//    680-727): row r's entries 7r..7r+6 are [diag, xm, xp, ym, yp, zm, zp]
//    (Predictor.cpp:689-697) with column -1 for a missing neighbour, rows in
//    the grid's gid order (j, i, k), k innermost over each column's active
//    cells (Grid.cpp:160-245).  setPattern recognises that pattern without
//    any hint: the z legs chain each column's rows, every column reaches the
//    grid's bottom layer (active cells are k >= ktop), so the bottom cells'
//    lateral legs give the column adjacency, and a breadth-first placement
//    recovers each connected group of columns on a grid.  Groups that do not
//    touch are placed side by side in y with an empty row between them.  The
//    placement is then checked leg by leg against the pattern; any mismatch
//    keeps the generic path.  Values are scattered into the symmetric stencil
//    storage the step's own PCG runs on (gw_pcg.h): the diagonal and the
//    plus legs, after checking that every minus leg equals its mirror plus
//    leg (a nonsymmetric operator, beyond 1e-12 of the diagonal, keeps the
//    generic path for those values).  The solve is that PCG: point Jacobi or
//    z-line, FP64 or FP32-inner mixed, the scalars on the device.
//  * Generic CSR for any other pattern.  Each CSR slot keeps the ordered list
//    of COO entries that feed it, so setValues sums duplicates in a fixed
//    order (deterministic, no atomics).  Point-Jacobi PCG in two kernels per
//    iteration (the SpMV reads eight lanes a row, forming p = z + beta p_old
//    of every column it gathers, so the direction update needs no kernel of
//    its own), the scalars reduced on the device from fixed-order block
//    partials; the host reads the state every few iterations.
//
// Both start from x = 0 (PETSc's default guess) and stop on the true
// residual ||b - A x||_2 <= max(rtol ||b||_2, atol); non-convergence is an
// error (LinearSystem.cpp:426-431).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/frehg_gw.h"
#include "gw_device.h"
#include "gw_pcg.h"

namespace frehg {

void set_error_message(const char* m);  // gw_step.hip

namespace {

struct LsError : std::runtime_error {
  int code;
  LsError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void ls_fail(int c, const std::string& m) { throw LsError(c, m); }

#define LS_HIP(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) ls_fail(FREHG_EDEVICE, std::string("HIP: ") + hipGetErrorString(e_)); \
  } while (0)

constexpr int kT = 256;
constexpr int kRowLanes = 8;      // lanes that share one CSR row in the SpMV
constexpr int kPollEvery = 8;     // iterations between host reads of the state

// ---------------------------------------------------------------------------
// Structured embedding of the 7-point pattern (host, once per setPattern).
// ---------------------------------------------------------------------------
struct Embedding {
  int nx = 0, ny = 0, nz = 0;
  std::vector<int64_t> cell;      // halo index of each row
  std::vector<uint8_t> legs;      // bit 0: xp, bit 1: yp, bit 2: zp leg present
  std::vector<int64_t> minus;     // 3 per row: xm, ym, zm columns (-1 if absent)
};

// Returns false (with a reason) unless the pattern is the 7-point pattern of
// a structured grid in buildCooPattern's order.
bool embed_7point(int64_t n, int64_t nCoo, const int64_t* rows, const int64_t* cols,
                  Embedding& em, std::string& why) {
  if (n <= 0 || nCoo != 7 * n) {
    why = "not 7 entries per row";
    return false;
  }
  const auto col = [&](int64_t r, int q) { return cols[7 * r + q]; };
  for (int64_t r = 0; r < n; ++r) {
    for (int q = 0; q < 7; ++q)
      if (rows[7 * r + q] != r) {
        why = "rows are not grouped in sevens";
        return false;
      }
    if (col(r, 0) != r) {
      why = "the first entry of a row is not its diagonal";
      return false;
    }
    for (int q = 1; q < 7; ++q)
      if (col(r, q) < -1 || col(r, q) >= n || col(r, q) == r) {
        why = "a leg column is out of range";
        return false;
      }
    const int64_t zm = col(r, 5), zp = col(r, 6);
    if ((zm != -1 && zm != r - 1) || (zp != -1 && zp != r + 1)) {
      why = "z legs do not chain consecutive rows";
      return false;
    }
    if (zp != -1 && col(zp, 5) != r) {
      why = "z legs are not mirrored";
      return false;
    }
  }
  // columns: maximal chains of z legs
  std::vector<int64_t> colOf(n), start, len;
  for (int64_t r = 0; r < n; ++r) {
    if (col(r, 5) == -1) {
      start.push_back(r);
      len.push_back(0);
    } else if (start.empty()) {
      why = "a column does not start with a zm-free row";
      return false;
    }
    colOf[r] = static_cast<int64_t>(start.size()) - 1;
    len.back() += 1;
  }
  const int64_t ncol = static_cast<int64_t>(start.size());
  int64_t nzMax = 0;
  for (int64_t c = 0; c < ncol; ++c) {
    if (col(start[c] + len[c] - 1, 6) != -1) {
      why = "a column runs on past its bottom";
      return false;
    }
    nzMax = std::max(nzMax, len[c]);
  }
  const auto depth = [&](int64_t r) { return start[colOf[r]] + len[colOf[r]] - 1 - r; };
  // column adjacency from the bottom cells (q: 1 xm, 2 xp, 3 ym, 4 yp)
  std::vector<int64_t> adj(4 * ncol, -1);
  for (int64_t c = 0; c < ncol; ++c) {
    const int64_t rb = start[c] + len[c] - 1;
    for (int q = 1; q <= 4; ++q) {
      const int64_t o = col(rb, q);
      if (o == -1) continue;
      if (depth(o) != 0) {
        why = "a bottom cell's lateral leg ends above the neighbour's bottom";
        return false;
      }
      adj[4 * c + q - 1] = colOf[o];
    }
  }
  // every lateral leg: same depth in the adjacent column, present exactly
  // where the adjacent column is that deep; legs mirrored
  static const int mirror[5] = {0, 2, 1, 4, 3};
  for (int64_t r = 0; r < n; ++r) {
    const int64_t c = colOf[r], d = depth(r);
    for (int q = 1; q <= 4; ++q) {
      const int64_t o = col(r, q), oc = adj[4 * c + q - 1];
      const bool expect = oc != -1 && len[oc] > d;
      if (!expect) {
        if (o != -1) {
          why = "a lateral leg where no neighbour exists";
          return false;
        }
        continue;
      }
      if (o == -1 || colOf[o] != oc || depth(o) != d || col(o, mirror[q]) != r) {
        why = "lateral legs are not those of a grid";
        return false;
      }
    }
  }
  // place connected groups of columns on a grid (breadth-first)
  std::vector<int64_t> ci(ncol, INT64_MIN), cj(ncol, INT64_MIN), group(ncol, -1);
  struct Box {
    int64_t i0, i1, j0, j1;
  };
  std::vector<Box> boxes;
  static const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
  for (int64_t s = 0; s < ncol; ++s) {
    if (group[s] != -1) continue;
    const int64_t g = static_cast<int64_t>(boxes.size());
    boxes.push_back({0, 0, 0, 0});
    std::deque<int64_t> queue{s};
    group[s] = g;
    ci[s] = 0;
    cj[s] = 0;
    while (!queue.empty()) {
      const int64_t c = queue.front();
      queue.pop_front();
      Box& b = boxes[g];
      b.i0 = std::min(b.i0, ci[c]);
      b.i1 = std::max(b.i1, ci[c]);
      b.j0 = std::min(b.j0, cj[c]);
      b.j1 = std::max(b.j1, cj[c]);
      for (int q = 0; q < 4; ++q) {
        const int64_t o = adj[4 * c + q];
        if (o == -1) continue;
        if (group[o] == -1) {
          group[o] = g;
          ci[o] = ci[c] + di[q];
          cj[o] = cj[c] + dj[q];
          queue.push_back(o);
        } else if (ci[o] != ci[c] + di[q] || cj[o] != cj[c] + dj[q]) {
          why = "the column adjacency is not planar";
          return false;
        }
      }
    }
  }
  // groups stacked in y, one empty row between neighbours
  int64_t width = 0, height = 0;
  std::vector<int64_t> yOff(boxes.size());
  for (size_t g = 0; g < boxes.size(); ++g) {
    width = std::max(width, boxes[g].i1 - boxes[g].i0 + 1);
    yOff[g] = height + (g ? 1 : 0);
    height = yOff[g] + (boxes[g].j1 - boxes[g].j0 + 1);
  }
  if (width * height > (int64_t(1) << 30) || nzMax > (int64_t(1) << 20)) {
    why = "the embedding is too large";
    return false;
  }
  em.nx = static_cast<int>(width);
  em.ny = static_cast<int>(height);
  em.nz = static_cast<int>(nzMax);
  std::vector<int64_t> owner(static_cast<size_t>(width * height), -1);
  std::vector<int64_t> gi(ncol), gj(ncol);
  for (int64_t c = 0; c < ncol; ++c) {
    const Box& b = boxes[group[c]];
    gi[c] = ci[c] - b.i0;
    gj[c] = cj[c] - b.j0 + yOff[group[c]];
    int64_t& slot = owner[gj[c] * width + gi[c]];
    if (slot != -1) {
      why = "two columns land on one place";
      return false;
    }
    slot = c;
  }
  // the placement's neighbours are exactly the pattern's
  for (int64_t c = 0; c < ncol; ++c)
    for (int q = 0; q < 4; ++q) {
      const int64_t ii = gi[c] + di[q], jj = gj[c] + dj[q];
      const int64_t at = (ii < 0 || jj < 0 || ii >= width || jj >= height)
                             ? -1
                             : owner[jj * width + ii];
      if (at != adj[4 * c + q]) {
        why = "the placement adds a neighbour the pattern does not have";
        return false;
      }
    }
  const Grid g{em.nx, em.ny, em.nz, em.nx + 2, em.ny + 2, 1.0, 1.0, 1.0};
  em.cell.resize(n);
  em.legs.resize(n);
  em.minus.resize(3 * n);
  for (int64_t r = 0; r < n; ++r) {
    const int64_t c = colOf[r];
    em.cell[r] = g.c3(static_cast<int>(gj[c]) + 1, static_cast<int>(gi[c]) + 1,
                      static_cast<int>(nzMax - 1 - depth(r)));
    em.legs[r] = static_cast<uint8_t>((col(r, 2) != -1 ? 1 : 0) | (col(r, 4) != -1 ? 2 : 0) |
                                      (col(r, 6) != -1 ? 4 : 0));
    em.minus[3 * r] = col(r, 1);
    em.minus[3 * r + 1] = col(r, 3);
    em.minus[3 * r + 2] = col(r, 5);
  }
  return true;
}

// ---------------------------------------------------------------------------
// Structured-path kernels.
// ---------------------------------------------------------------------------
// Inactive places of the embedding are identity rows (Ad = 1, legs 0).
__global__ void k_ls_identity(int64_t nH, double* __restrict__ Ad) {
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < nH;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x)
    Ad[c] = 1.0;
}

__global__ void k_ls_scatter_stencil(int64_t n, const int64_t* __restrict__ cell,
                                     const uint8_t* __restrict__ legs,
                                     const double* __restrict__ v, Stencil<double> A) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n;
       r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t c = cell[r];
    const int m = legs[r];
    A.Ad[c] = v[7 * r];
    A.Axp[c] = (m & 1) ? v[7 * r + 2] : 0.0;
    A.Ayp[c] = (m & 2) ? v[7 * r + 4] : 0.0;
    A.Azp[c] = (m & 4) ? v[7 * r + 6] : 0.0;
  }
}

// Largest |minus leg - mirror plus leg| / |diagonal| over the rows; block
// maxima in parts.
__global__ void k_ls_asymmetry(int64_t n, const int64_t* __restrict__ minus,
                               const double* __restrict__ v, double* __restrict__ parts) {
  double w[1] = {0.0};
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n;
       r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double d = fabs(v[7 * r]) > 0.0 ? fabs(v[7 * r]) : 1.0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const int64_t o = minus[3 * r + a];
      if (o < 0) continue;
      const double e = fabs(v[7 * r + 1 + 2 * a] - v[7 * o + 2 + 2 * a]) / d;
      w[0] = fmax(w[0], e != e ? INFINITY : e);
    }
  }
  const int op[1] = {1};
  block_reduce<1>(w, op);
  if (threadIdx.x == 0) parts[blockIdx.x] = w[0];
}

__global__ void k_ls_scatter_vec(int64_t n, const int64_t* __restrict__ cell,
                                 const double* __restrict__ src, double* __restrict__ dst) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n;
       r += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[cell[r]] = src[r];
}

__global__ void k_ls_gather_vec(int64_t n, const int64_t* __restrict__ cell,
                                const double* __restrict__ src, double* __restrict__ dst) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n;
       r += static_cast<int64_t>(gridDim.x) * blockDim.x)
    dst[r] = src[cell[r]];
}

__global__ void k_ls_fill(int64_t n, double* __restrict__ x, double v) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[i] = v;
}

// ---------------------------------------------------------------------------
// Generic CSR path.
// ---------------------------------------------------------------------------
__global__ void k_gather_values(int64_t nnz, const int64_t* __restrict__ slotPtr,
                                const int64_t* __restrict__ slotEntries,
                                const double* __restrict__ coo, double* __restrict__ val) {
  for (int64_t s = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; s < nnz;
       s += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double v = 0.0;
    for (int64_t e = slotPtr[s]; e < slotPtr[s + 1]; ++e) v += coo[slotEntries[e]];
    val[s] = v;
  }
}

__global__ void k_diag_inv(int64_t n, const int64_t* __restrict__ rp, const int64_t* __restrict__ ci,
                           const double* __restrict__ val, double* __restrict__ dinv) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n;
       r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double d = 1.0;
    for (int64_t e = rp[r]; e < rp[r + 1]; ++e)
      if (ci[e] == r && val[e] != 0.0) d = 1.0 / val[e];
    dinv[r] = d;
  }
}

// Device state of the generic PCG (read by the host every few iterations).
struct CsrState {
  double rz[2];
  double rr;
  int it;
  int done;
};

// Fixed-order reduction of a slot's block partials inside every block.
template <int NV>
__device__ __forceinline__ void csr_reduce(const double* parts, int pstride, int nparts,
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
__device__ __forceinline__ void csr_write(double* parts, int pstride, const int (&slot)[NV],
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

// r = b - A x (or r = b when x is null); z = D^-1 r; partials rz (0), rr (1),
// bb (3).  Eight lanes a row.
__global__ __launch_bounds__(kT) void k_csr_residual(int64_t n, const int64_t* __restrict__ rp,
                                                      const int64_t* __restrict__ ci,
                                                      const double* __restrict__ val,
                                                      const double* __restrict__ dinv,
                                                      const double* __restrict__ b,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ r,
                                                      double* __restrict__ z,
                                                      double* __restrict__ parts, int pstride) {
  const int sub = threadIdx.x % kRowLanes;
  KSum rz, rr, bb;
  const int64_t rowsPerPass = static_cast<int64_t>(gridDim.x) * (blockDim.x / kRowLanes);
  for (int64_t row = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kRowLanes) +
                     threadIdx.x / kRowLanes;
       row - (threadIdx.x / kRowLanes) < n; row += rowsPerPass) {
    double s = 0.0;
    if (row < n && x)
      for (int64_t e = rp[row] + sub; e < rp[row + 1]; e += kRowLanes) s += val[e] * x[ci[e]];
#pragma unroll
    for (int o = kRowLanes / 2; o > 0; o /= 2) s += __shfl_xor(s, o, kRowLanes);
    if (row < n && sub == 0) {
      const double rv = b[row] - s;
      const double zv = dinv[row] * rv;
      r[row] = rv;
      z[row] = zv;
      rz.add(rv * zv);
      rr.add(rv * rv);
      bb.add(b[row] * b[row]);
    }
  }
  double v[3] = {rz.value(), rr.value(), bb.value()};
  const int slots[3] = {0, 1, 3};
  csr_write<3>(parts, pstride, slots, v);
}

// p = z + beta p_old (formed wherever it is read), q = A p, partial p.q
// (slot 2); beta from the previous update's partial r.z.
__global__ __launch_bounds__(kT) void k_csr_dir(int64_t n, const int64_t* __restrict__ rp,
                                                 const int64_t* __restrict__ ci,
                                                 const double* __restrict__ val,
                                                 const double* __restrict__ z,
                                                 const double* __restrict__ pOld,
                                                 double* __restrict__ pNew,
                                                 double* __restrict__ q,
                                                 double* __restrict__ parts, int pstride,
                                                 int nprev, CsrState* __restrict__ st, int it,
                                                 int restart, double tol2) {
  double red[2];
  const int rs[2] = {0, 1};
  csr_reduce<2>(parts, pstride, nprev, rs, red);
  if (red[1] <= tol2 || !(red[1] == red[1])) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && !st->done) {
      st->done = 1;
      st->it = it;
      st->rr = red[1];
    }
    return;
  }
  const double beta = restart ? 0.0 : red[0] / st->rz[(it + 1) & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->rz[it & 1] = red[0];
    st->it = it;
    st->rr = red[1];
  }
  const int sub = threadIdx.x % kRowLanes;
  KSum pq;
  const int64_t rowsPerPass = static_cast<int64_t>(gridDim.x) * (blockDim.x / kRowLanes);
  for (int64_t row = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kRowLanes) +
                     threadIdx.x / kRowLanes;
       row - (threadIdx.x / kRowLanes) < n; row += rowsPerPass) {
    double s = 0.0;
    if (row < n)
      for (int64_t e = rp[row] + sub; e < rp[row + 1]; e += kRowLanes) {
        const int64_t c = ci[e];
        s += val[e] * (restart ? z[c] : z[c] + beta * pOld[c]);
      }
#pragma unroll
    for (int o = kRowLanes / 2; o > 0; o /= 2) s += __shfl_xor(s, o, kRowLanes);
    if (row < n && sub == 0) {
      const double pv = restart ? z[row] : z[row] + beta * pOld[row];
      pNew[row] = pv;
      q[row] = s;
      pq.add(pv * s);
    }
  }
  double v[1] = {pq.value()};
  const int slots[1] = {2};
  csr_write<1>(parts, pstride, slots, v);
}

// alpha = rz / p.q; x += alpha p; r -= alpha q; z = D^-1 r; partials rz, rr.
__global__ __launch_bounds__(kT) void k_csr_update(int64_t n, double* __restrict__ x,
                                                    const double* __restrict__ p,
                                                    double* __restrict__ r,
                                                    const double* __restrict__ q,
                                                    double* __restrict__ z,
                                                    const double* __restrict__ dinv,
                                                    double* __restrict__ parts, int pstride,
                                                    int nprev, CsrState* __restrict__ st, int it) {
  if (st->done) return;
  double red[1];
  const int rs[1] = {2};
  csr_reduce<1>(parts, pstride, nprev, rs, red);
  const double alpha = st->rz[it & 1] / red[0];
  KSum rz, rr;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    x[i] += alpha * p[i];
    const double rv = r[i] - alpha * q[i];
    const double zv = dinv[i] * rv;
    r[i] = rv;
    z[i] = zv;
    rz.add(rv * zv);
    rr.add(rv * rv);
  }
  double v[2] = {rz.value(), rr.value()};
  const int slots[2] = {0, 1};
  csr_write<2>(parts, pstride, slots, v);
}

__global__ void k_csr_reduce_slots(const double* __restrict__ parts, int pstride, int nparts,
                                   double* __restrict__ out) {
  double v[4];
  const int slots[4] = {0, 1, 2, 3};
  csr_reduce<4>(parts, pstride, nparts, slots, v);
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; ++q) out[q] = v[q];
}

__global__ void k_csr_state_reset(CsrState* st) {
  st->done = 0;
  st->it = 0;
}

__global__ void k_max_parts(int np, const double* __restrict__ parts, double* __restrict__ out) {
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < np; b += blockDim.x) v[0] = fmax(v[0], parts[b]);
  const int op[1] = {1};
  block_reduce<1>(v, op);
  if (threadIdx.x == 0) *out = v[0];
}

struct Ls {
  int64_t n = 0, nCoo = 0, nnz = 0;
  double rtol = 1e-8, atol = 1e-14;
  int maxIt = 500;
  int pc = FREHG_PC_JACOBI;
  int precision = FREHG_PREC_FP64;
  double innerRtol = 0.0;
  bool pattern = false, values = false;
  bool structuredPattern = false;  // the pattern embeds as a 7-point grid
  bool structuredValues = false;   // ... and the current values are symmetric
  std::string structuredWhy;       // why the pattern did not embed
  double asymmetry = 0.0;          // largest relative minus/plus leg difference
  std::vector<int64_t> hRp, hCi;
  // generic CSR
  int64_t *rp = nullptr, *ci = nullptr, *slotPtr = nullptr, *slotEntries = nullptr;
  double *val = nullptr, *dinv = nullptr, *r = nullptr, *z = nullptr, *p0 = nullptr,
         *p1 = nullptr, *q = nullptr;
  double *parts = nullptr, *scal = nullptr, *scalHost = nullptr;
  CsrState *state = nullptr, *stateHost = nullptr;
  int grid = 512, gridRows = 512, pstride = 512;
  // structured
  Grid g{};
  std::unique_ptr<Pcg> pcg;
  Stencil<double> A{};
  double *bH = nullptr, *xH = nullptr;
  int64_t* cell = nullptr;
  uint8_t* legs = nullptr;
  int64_t* minus = nullptr;
  bool csrWork = false;              // generic-path vectors allocated
  hipStream_t stream = nullptr;
  std::vector<void*> owned;

  ~Ls() {
    if (stream) (void)hipStreamSynchronize(stream);
    pcg.reset();
    for (void* a : owned)
      if (a) (void)hipFree(a);
    if (scalHost) (void)hipHostFree(scalHost);
    if (stateHost) (void)hipHostFree(stateHost);
    if (stream) (void)hipStreamDestroy(stream);
  }
  template <class T>
  T* alloc(int64_t count) {
    void* a = nullptr;
    LS_HIP(hipMalloc(&a, static_cast<size_t>(std::max<int64_t>(count, 1)) * sizeof(T)));
    owned.push_back(a);
    LS_HIP(hipMemset(a, 0, static_cast<size_t>(std::max<int64_t>(count, 1)) * sizeof(T)));
    return static_cast<T*>(a);
  }
  int blocksFor(int64_t count) const {
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(2048, (count + kT - 1) / kT)));
  }

  void buildCsr(const int64_t* rows, const int64_t* cols) {
    std::vector<std::vector<std::pair<int64_t, int64_t>>> perRow(n);
    for (int64_t e = 0; e < nCoo; ++e) {
      if (rows[e] < 0 || cols[e] < 0) continue;
      if (rows[e] >= n || cols[e] >= n) ls_fail(FREHG_EINVAL, "COO index out of range");
      perRow[rows[e]].push_back({cols[e], e});
    }
    std::vector<int64_t> sp(1, 0), se;
    hRp.assign(n + 1, 0);
    hCi.clear();
    for (int64_t rr = 0; rr < n; ++rr) {
      auto& v = perRow[rr];
      std::sort(v.begin(), v.end());
      for (size_t a = 0; a < v.size(); ++a) {
        if (a == 0 || v[a].first != v[a - 1].first) {
          hCi.push_back(v[a].first);
          sp.push_back(sp.back());
        }
        se.push_back(v[a].second);
        sp.back() += 1;
      }
      hRp[rr + 1] = static_cast<int64_t>(hCi.size());
    }
    nnz = static_cast<int64_t>(hCi.size());
    rp = alloc<int64_t>(n + 1);
    ci = alloc<int64_t>(nnz);
    slotPtr = alloc<int64_t>(nnz + 1);
    slotEntries = alloc<int64_t>(static_cast<int64_t>(se.size()));
    val = alloc<double>(nnz);
    LS_HIP(hipMemcpy(rp, hRp.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    LS_HIP(hipMemcpy(ci, hCi.data(), nnz * sizeof(int64_t), hipMemcpyHostToDevice));
    LS_HIP(hipMemcpy(slotPtr, sp.data(), sp.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    LS_HIP(hipMemcpy(slotEntries, se.data(), se.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  }

  void buildStructured(const Embedding& em) {
    g = Grid{em.nx, em.ny, em.nz, em.nx + 2, em.ny + 2, 1.0, 1.0, 1.0};
    const int64_t nH = g.cellsH();
    A.Ad = alloc<double>(nH);
    A.Axp = alloc<double>(nH);
    A.Ayp = alloc<double>(nH);
    A.Azp = alloc<double>(nH);
    bH = alloc<double>(nH);
    xH = alloc<double>(nH);
    cell = alloc<int64_t>(n);
    legs = alloc<uint8_t>(n);
    minus = alloc<int64_t>(3 * n);
    LS_HIP(hipMemcpy(cell, em.cell.data(), n * sizeof(int64_t), hipMemcpyHostToDevice));
    LS_HIP(hipMemcpy(legs, em.legs.data(), n * sizeof(uint8_t), hipMemcpyHostToDevice));
    LS_HIP(hipMemcpy(minus, em.minus.data(), 3 * n * sizeof(int64_t), hipMemcpyHostToDevice));
    k_ls_identity<<<blocksFor(nH), kT, 0, stream>>>(nH, A.Ad);
    LS_HIP(hipGetLastError());
    pcg = std::make_unique<Pcg>(g, stream);
  }

  // The seam's own copy of the values (the caller may reuse its buffer after
  // setValues, as PETSc copies them), merged into CSR slots, and D^-1.
  void gatherCsr(const double* coo) {
    k_gather_values<<<blocksFor(nnz), kT, 0, stream>>>(nnz, slotPtr, slotEntries, coo, val);
    k_diag_inv<<<grid, kT, 0, stream>>>(n, rp, ci, val, dinv);
    LS_HIP(hipGetLastError());
  }
  void csrVectors() {
    if (csrWork) return;
    r = alloc<double>(n);
    z = alloc<double>(n);
    p0 = alloc<double>(n);
    p1 = alloc<double>(n);
    q = alloc<double>(n);
    csrWork = true;
  }

  double readScalar() {
    LS_HIP(hipMemcpyAsync(scalHost, scal, sizeof(double), hipMemcpyDeviceToHost, stream));
    LS_HIP(hipStreamSynchronize(stream));
    return *scalHost;
  }

  void solveStructured(const double* b, double* x, frehg_gw_solve_stats* st) {
    const int64_t nH = g.cellsH();
    k_ls_fill<<<blocksFor(nH), kT, 0, stream>>>(nH, xH, 0.0);
    k_ls_scatter_vec<<<blocksFor(n), kT, 0, stream>>>(n, cell, b, bH);
    LS_HIP(hipGetLastError());
    PcgOptions o;
    o.rtol = rtol;
    o.atol = atol;
    o.max_iterations = maxIt;
    o.zline = pc == FREHG_PC_ZLINE;
    o.mixed = precision == FREHG_PREC_MIXED;
    if (innerRtol > 0.0) o.inner_rtol = innerRtol;
    PcgStats ps;
    const int rc = pcg->solve(A, bH, xH, o, &ps);
    k_ls_gather_vec<<<blocksFor(n), kT, 0, stream>>>(n, cell, xH, x);
    LS_HIP(hipGetLastError());
    LS_HIP(hipStreamSynchronize(stream));
    st->iterations = ps.iterations;
    st->refinements = ps.refinements;
    st->converged = ps.converged;
    st->residual_norm = ps.residual_norm;
    st->rhs_norm = ps.rhs_norm;
    st->fp64_iterations = ps.fp64_iterations;
    (void)rc;
  }

  void solveGeneric(const double* b, double* x, frehg_gw_solve_stats* st) {
    csrVectors();
    k_ls_fill<<<grid, kT, 0, stream>>>(n, x, 0.0);
    auto residual = [&](bool fromZero) {
      k_csr_residual<<<gridRows, kT, 0, stream>>>(n, rp, ci, val, dinv, b, fromZero ? nullptr : x,
                                                   r, z, parts, pstride);
      k_csr_reduce_slots<<<1, kT, 0, stream>>>(parts, pstride, gridRows, scal);
      LS_HIP(hipGetLastError());
      LS_HIP(hipMemcpyAsync(scalHost, scal, 4 * sizeof(double), hipMemcpyDeviceToHost, stream));
      LS_HIP(hipStreamSynchronize(stream));
      return std::make_pair(scalHost[1], scalHost[3]);
    };
    auto [rr, bb] = residual(true);
    const double tol = std::max(rtol * std::sqrt(bb), atol);
    const double tol2 = tol * tol;
    st->rhs_norm = std::sqrt(bb);
    int total = 0, stalls = 0;
    double* pOld = p0;
    double* pNew = p1;
    while (true) {
      if (rr <= tol2) {
        st->converged = 1;
        break;
      }
      if (!std::isfinite(rr) || total >= maxIt) break;
      k_csr_state_reset<<<1, 1, 0, stream>>>(state);
      int it = 0, done = 0;
      const int cap = maxIt - total;
      while (!done && it < cap) {
        const int chunk = std::min(kPollEvery, cap - it);
        for (int c = 0; c < chunk; ++c, ++it) {
          const int np = it == 0 ? gridRows : grid;
          k_csr_dir<<<gridRows, kT, 0, stream>>>(n, rp, ci, val, z, pOld, pNew, q, parts, pstride,
                                                  np, state, it, it == 0 ? 1 : 0, tol2);
          k_csr_update<<<grid, kT, 0, stream>>>(n, x, pNew, r, q, z, dinv, parts, pstride,
                                                gridRows, state, it);
          std::swap(pOld, pNew);
        }
        LS_HIP(hipGetLastError());
        LS_HIP(hipMemcpyAsync(stateHost, state, sizeof(CsrState), hipMemcpyDeviceToHost, stream));
        LS_HIP(hipStreamSynchronize(stream));
        done = stateHost->done;
      }
      const int its = done ? stateHost->it : it;
      total += its;
      std::tie(rr, bb) = residual(false);  // true residual; a restart if not converged
      if (its == 0 && ++stalls > 2 && !(rr <= tol2)) break;
    }
    st->iterations = total;
    st->residual_norm = std::sqrt(rr);
  }
};

// ---------------------------------------------------------------------------
// The allocation-free variant of the structured solve
// (frehg_ls_solve_stencil_noalloc): every work vector in one file-scope
// device array.  Each GPU holds its own copy of it from the library's load
// until the process ends.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocCells = FREHG_NOALLOC_MAX_CELLS;
constexpr int64_t kNoAllocDoubles = 7 * (kNoAllocCells + 32) + 4 * 4096 + 256;
constexpr int kMaxDevices = 64;

__device__ __align__(256) double scratch[kNoAllocDoubles];

std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
double g_scratch_red[kMaxDevices][4];
alignas(64) unsigned char g_scratch_state[kMaxDevices][256];

// The device address of the current device's scratch array, looked up once
// per device with that device current (the array's name in host code is a
// host-side stand-in, not a device address).
double* scratch_of_current_device(int* devOut) {
  int dev = 0;
  LS_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) ls_fail(FREHG_EDEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      ls_fail(FREHG_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  *devOut = dev;
  return g_scratch_ptr[dev];
}

template <class Fn>
int ls_guard(Fn fn) {
  try {
    fn();
    return FREHG_OK;
  } catch (const LsError& e) {
    set_error_message(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_error_message(e.what());
    return FREHG_EDEVICE;
  }
}

}  // namespace
}  // namespace frehg

using frehg::Ls;
using frehg::ls_fail;

extern "C" {

int frehg_ls_create(int64_t n_rows, double rtol, double atol, int max_iterations, int preconditioner,
                    frehg_linear_system** out) {
  return frehg::ls_guard([&] {
    if (!out || n_rows < 0) frehg::ls_fail(FREHG_EINVAL, "invalid arguments");
    if (preconditioner != FREHG_PC_JACOBI && preconditioner != FREHG_PC_ZLINE)
      frehg::ls_fail(FREHG_EUNSUPPORTED,
                     "the solve seam runs point Jacobi or z-line (ICC(0) is the CPU oracle's)");
    auto ls = std::make_unique<Ls>();
    ls->n = n_rows;
    ls->rtol = rtol;
    ls->atol = atol;
    ls->maxIt = max_iterations > 0 ? max_iterations : 500;
    ls->pc = preconditioner;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      frehg::ls_fail(FREHG_EDEVICE, "no HIP device is visible");
    LS_HIP(hipStreamCreateWithFlags(&ls->stream, hipStreamNonBlocking));
    ls->grid = ls->blocksFor(n_rows);
    // the SpMV kernels take kT / kRowLanes rows a block and pass
    constexpr int rowsPerBlock = frehg::kT / frehg::kRowLanes;
    ls->gridRows = static_cast<int>(std::max<int64_t>(
        1, std::min<int64_t>(2048, (n_rows + rowsPerBlock - 1) / rowsPerBlock)));
    ls->pstride = std::max(ls->grid, ls->gridRows);
    *out = reinterpret_cast<frehg_linear_system*>(ls.release());
  });
}

int frehg_ls_destroy(frehg_linear_system* h) {
  delete reinterpret_cast<Ls*>(h);
  return FREHG_OK;
}

int frehg_ls_set_pattern(frehg_linear_system* h, int64_t n_coo, const int64_t* rows,
                         const int64_t* cols) {
  return frehg::ls_guard([&] {
    Ls* ls = reinterpret_cast<Ls*>(h);
    if (!ls || n_coo < 0 || (n_coo > 0 && (!rows || !cols)))
      frehg::ls_fail(FREHG_EINVAL, "invalid arguments");
    if (ls->pattern) frehg::ls_fail(FREHG_EINVAL, "setPattern may be called once");
    ls->nCoo = n_coo;
    frehg::Embedding em;
    ls->structuredPattern = frehg::embed_7point(ls->n, n_coo, rows, cols, em, ls->structuredWhy);
    if (!ls->structuredPattern && ls->pc == FREHG_PC_ZLINE)
      frehg::ls_fail(FREHG_EUNSUPPORTED, "the z-line preconditioner needs the 7-point pattern of "
                                         "a structured grid (" + ls->structuredWhy + ")");
    ls->buildCsr(rows, cols);
    ls->dinv = ls->alloc<double>(ls->n);
    ls->parts = ls->alloc<double>(4 * static_cast<int64_t>(ls->pstride));
    ls->scal = ls->alloc<double>(4);
    ls->state = ls->alloc<frehg::CsrState>(1);
    LS_HIP(hipHostMalloc(reinterpret_cast<void**>(&ls->scalHost), 4 * sizeof(double),
                         hipHostMallocDefault));
    LS_HIP(hipHostMalloc(reinterpret_cast<void**>(&ls->stateHost), sizeof(frehg::CsrState),
                         hipHostMallocDefault));
    if (ls->structuredPattern) ls->buildStructured(em);
    LS_HIP(hipStreamSynchronize(ls->stream));
    ls->pattern = true;
  });
}

int frehg_ls_set_values(frehg_linear_system* h, const double* values_device) {
  return frehg::ls_guard([&] {
    Ls* ls = reinterpret_cast<Ls*>(h);
    if (!ls || !values_device) frehg::ls_fail(FREHG_EINVAL, "invalid arguments");
    if (!ls->pattern) frehg::ls_fail(FREHG_EINVAL, "setValues before setPattern");
    ls->structuredValues = false;
    ls->gatherCsr(values_device);
    if (ls->structuredPattern) {
      const int nb = ls->blocksFor(ls->n);
      frehg::k_ls_asymmetry<<<nb, frehg::kT, 0, ls->stream>>>(ls->n, ls->minus, values_device,
                                                              ls->parts);
      frehg::k_max_parts<<<1, frehg::kT, 0, ls->stream>>>(nb, ls->parts, ls->scal);
      LS_HIP(hipGetLastError());
      ls->asymmetry = ls->readScalar();
      ls->structuredValues = ls->asymmetry <= 1e-12;
      if (ls->structuredValues) {
        frehg::k_ls_scatter_stencil<<<nb, frehg::kT, 0, ls->stream>>>(ls->n, ls->cell, ls->legs,
                                                                      values_device, ls->A);
        LS_HIP(hipGetLastError());
      } else if (ls->pc == FREHG_PC_ZLINE) {
        frehg::ls_fail(FREHG_EUNSUPPORTED,
                       "the values are not symmetric (relative difference " +
                           std::to_string(ls->asymmetry) + "): z-line CG needs a symmetric operator");
      }
    }
    LS_HIP(hipStreamSynchronize(ls->stream));
    ls->values = true;
  });
}

int frehg_ls_solve(frehg_linear_system* h, const double* b, double* x, frehg_gw_solve_stats* stats) {
  return frehg::ls_guard([&] {
    Ls* ls = reinterpret_cast<Ls*>(h);
    if (!ls || !b || !x) frehg::ls_fail(FREHG_EINVAL, "invalid arguments");
    if (!ls->values) frehg::ls_fail(FREHG_EINVAL, "solve before setValues");
    if (ls->precision == FREHG_PREC_MIXED && !ls->structuredValues)
      frehg::ls_fail(FREHG_EUNSUPPORTED, "the mixed form runs on the 7-point structured path only");
    frehg_gw_solve_stats st{};
    if (ls->n > 0) {
      if (ls->structuredValues) ls->solveStructured(b, x, &st);
      else ls->solveGeneric(b, x, &st);
    } else {
      st.converged = 1;
    }
    if (stats) *stats = st;
    if (!st.converged) frehg::ls_fail(FREHG_EDIVERGED, "linear solve did not converge");
  });
}

int frehg_ls_set_precision(frehg_linear_system* h, int precision, double inner_rtol) {
  return frehg::ls_guard([&] {
    Ls* ls = reinterpret_cast<Ls*>(h);
    if (!ls || (precision != FREHG_PREC_FP64 && precision != FREHG_PREC_MIXED) ||
        inner_rtol < 0.0 || inner_rtol >= 1.0)
      frehg::ls_fail(FREHG_EINVAL, "invalid arguments");
    ls->precision = precision;
    ls->innerRtol = inner_rtol;
  });
}

int frehg_ls_structured(frehg_linear_system* h, int* structured, double* asymmetry) {
  return frehg::ls_guard([&] {
    Ls* ls = reinterpret_cast<Ls*>(h);
    if (!ls || !structured) frehg::ls_fail(FREHG_EINVAL, "invalid arguments");
    *structured = ls->structuredPattern ? (ls->values ? (ls->structuredValues ? 2 : 1) : 1) : 0;
    if (asymmetry) *asymmetry = ls->asymmetry;
  });
}

int frehg_ls_solve_stencil_noalloc(int nx, int ny, int nz, const double* Ad, const double* Axp,
                                   const double* Ayp, const double* Azp, const double* rhs,
                                   double* sol, double rtol, double atol, int max_iterations,
                                   int preconditioner, frehg_gw_solve_stats* stats) {
  return frehg::ls_guard([&] {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !Ad || !Axp || !Ayp || !Azp || !rhs || !sol)
      frehg::ls_fail(FREHG_EINVAL, "invalid arguments");
    if (preconditioner != FREHG_PC_JACOBI && preconditioner != FREHG_PC_ZLINE)
      frehg::ls_fail(FREHG_EUNSUPPORTED, "point Jacobi or z-line");
    if (nz > 1024)
      frehg::ls_fail(FREHG_EUNSUPPORTED, "nz > 1024 is not supported by the column kernels");
    const int64_t nH = static_cast<int64_t>(nx + 2) * (ny + 2) * nz;
    if (nH > frehg::kNoAllocCells)
      frehg::ls_fail(FREHG_EINVAL, "(nx+2)(ny+2)nz = " + std::to_string(nH) +
                                       " exceeds the allocation-free scratch maximum of " +
                                       std::to_string(frehg::kNoAllocCells) + " cells");
    const frehg::Grid g{nx, ny, nz, nx + 2, ny + 2, 1.0, 1.0, 1.0};
    const int64_t need = frehg::Pcg::workspaceDoubles(g);
    if (need > frehg::kNoAllocDoubles)
      frehg::ls_fail(FREHG_EINVAL, "the grid needs more scratch than the array holds");
    int dev = 0;
    double* base = frehg::scratch_of_current_device(&dev);
    frehg::PcgWorkspace ws;
    ws.device = base;
    ws.doubles = frehg::kNoAllocDoubles;
    ws.hostRed = frehg::g_scratch_red[dev];
    ws.hostState = frehg::g_scratch_state[dev];
    const hipStream_t s = nullptr;  // the null stream: the call returns when the solve is done
    frehg::Pcg pcg(g, s, ws);
    pcg.clearVectors();
    frehg::k_ls_fill<<<std::max(1, static_cast<int>(std::min<int64_t>(2048, (nH + 255) / 256))),
                       frehg::kT, 0, s>>>(nH, sol, 0.0);
    LS_HIP(hipGetLastError());
    frehg::PcgOptions o;
    o.rtol = rtol;
    o.atol = atol;
    o.max_iterations = max_iterations > 0 ? max_iterations : 500;
    o.zline = preconditioner == FREHG_PC_ZLINE;
    frehg::PcgStats ps;
    const frehg::Stencil<double> A{const_cast<double*>(Ad), const_cast<double*>(Axp),
                                   const_cast<double*>(Ayp), const_cast<double*>(Azp)};
    pcg.solve(A, rhs, sol, o, &ps);
    LS_HIP(hipStreamSynchronize(s));
    frehg_gw_solve_stats st{};
    st.iterations = ps.iterations;
    st.converged = ps.converged;
    st.residual_norm = ps.residual_norm;
    st.rhs_norm = ps.rhs_norm;
    if (stats) *stats = st;
    if (!st.converged) frehg::ls_fail(FREHG_EDIVERGED, "linear solve did not converge");
  });
}

int frehg_ls_is_symmetric(frehg_linear_system* h, double tol, int* symmetric) {
  return frehg::ls_guard([&] {
    Ls* ls = reinterpret_cast<Ls*>(h);
    if (!ls || !symmetric) frehg::ls_fail(FREHG_EINVAL, "invalid arguments");
    if (!ls->values) frehg::ls_fail(FREHG_EINVAL, "isSymmetric before setValues");
    std::vector<double> v(ls->nnz);
    LS_HIP(hipMemcpyAsync(v.data(), ls->val, ls->nnz * sizeof(double), hipMemcpyDeviceToHost,
                          ls->stream));
    LS_HIP(hipStreamSynchronize(ls->stream));
    auto at = [&](int64_t i, int64_t j) {
      const auto b = ls->hCi.begin() + ls->hRp[i], e = ls->hCi.begin() + ls->hRp[i + 1];
      const auto it = std::lower_bound(b, e, j);
      return (it != e && *it == j) ? v[it - ls->hCi.begin()] : 0.0;
    };
    bool ok = true;
    for (int64_t i = 0; i < ls->n && ok; ++i)
      for (int64_t e = ls->hRp[i]; e < ls->hRp[i + 1]; ++e)
        if (std::fabs(v[e] - at(ls->hCi[e], i)) > tol) {
          ok = false;
          break;
        }
    *symmetric = ok ? 1 : 0;
  });
}

}  // extern "C"

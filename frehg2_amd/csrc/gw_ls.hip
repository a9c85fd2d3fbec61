// gw_ls.hip — the solve seam: a COO-assembled sparse system solved by Jacobi
// PCG on the GPU, replacing frehg::LinearSystem (src/core/LinearSystem.hpp:
// 109-151; PETSc MatSetPreallocationCOO / MatSetValuesCOO / KSPSolve at
// src/core/LinearSystem.cpp:294, 308, 320).
//
// setPattern builds a CSR structure on the host: entries with a -1 row or
// column are ignored, duplicates are merged (LinearSystem.cpp:273-296).  Each
// CSR slot keeps the ordered list of COO entries that feed it, so setValues
// sums duplicates in a fixed order (deterministic, no atomics).  solve starts
// from x = 0 (PETSc's default guess) and stops on the true residual
// ||b - A x||_2 <= max(rtol ||b||_2, atol); non-convergence is an error
// (LinearSystem.cpp:426-431).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/frehg_gw.h"
#include "gw_device.h"

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

// y = A x (one thread per row); optionally y = b - A x.
__global__ void k_csr_mv(int64_t n, const int64_t* __restrict__ rp, const int64_t* __restrict__ ci,
                         const double* __restrict__ val, const double* __restrict__ x,
                         const double* __restrict__ b, double* __restrict__ y) {
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; r < n;
       r += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double s = 0.0;
    for (int64_t e = rp[r]; e < rp[r + 1]; ++e) s += val[e] * x[ci[e]];
    y[r] = b ? b[r] - s : s;
  }
}

// partial dot products a.b over a fixed grid; parts[blockIdx]
__global__ void k_dot(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ parts) {
  KSum s;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    s.add(a[i] * b[i]);
  double v[1] = {s.value()};
  const int op[1] = {0};
  block_reduce<1>(v, op);
  if (threadIdx.x == 0) parts[blockIdx.x] = v[0];
}

__global__ void k_sum_parts(int np, const double* __restrict__ parts, double* __restrict__ out) {
  double v[1] = {0.0};
  for (int b = threadIdx.x; b < np; b += blockDim.x) v[0] += parts[b];
  const int op[1] = {0};
  block_reduce<1>(v, op);
  if (threadIdx.x == 0) *out = v[0];
}

// z = dinv r; p = z + beta p (beta < 0 means restart: p = z)
__global__ void k_zp(int64_t n, const double* __restrict__ dinv, const double* __restrict__ r,
                     double* __restrict__ z, double* __restrict__ p, double beta) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double zi = dinv[i] * r[i];
    z[i] = zi;
    p[i] = beta < 0.0 ? zi : zi + beta * p[i];
  }
}

__global__ void k_axpy2(int64_t n, double alpha, const double* __restrict__ p,
                        const double* __restrict__ q, double* __restrict__ x, double* __restrict__ r) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    x[i] += alpha * p[i];
    r[i] -= alpha * q[i];
  }
}

__global__ void k_zero(int64_t n, double* __restrict__ x) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    x[i] = 0.0;
}

struct Ls {
  int64_t n = 0, nCoo = 0, nnz = 0;
  double rtol = 1e-8, atol = 1e-14;
  int maxIt = 500;
  bool pattern = false, values = false;
  std::vector<int64_t> hRp, hCi;
  int64_t *rp = nullptr, *ci = nullptr, *slotPtr = nullptr, *slotEntries = nullptr;
  double *val = nullptr, *dinv = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr;
  double *parts = nullptr, *scal = nullptr, *scalHost = nullptr;
  hipStream_t stream = nullptr;
  int grid = 512;

  ~Ls() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void* a : {(void*)rp, (void*)ci, (void*)slotPtr, (void*)slotEntries, (void*)val,
                    (void*)dinv, (void*)r, (void*)z, (void*)p, (void*)q, (void*)parts, (void*)scal})
      if (a) (void)hipFree(a);
    if (scalHost) (void)hipHostFree(scalHost);
    if (stream) (void)hipStreamDestroy(stream);
  }
  template <class T>
  T* alloc(int64_t count) {
    void* a = nullptr;
    LS_HIP(hipMalloc(&a, static_cast<size_t>(std::max<int64_t>(count, 1)) * sizeof(T)));
    LS_HIP(hipMemset(a, 0, static_cast<size_t>(std::max<int64_t>(count, 1)) * sizeof(T)));
    return static_cast<T*>(a);
  }
  double dot(const double* a, const double* b) {
    k_dot<<<grid, kT, 0, stream>>>(n, a, b, parts);
    k_sum_parts<<<1, kT, 0, stream>>>(grid, parts, scal);
    LS_HIP(hipGetLastError());
    LS_HIP(hipMemcpyAsync(scalHost, scal, sizeof(double), hipMemcpyDeviceToHost, stream));
    LS_HIP(hipStreamSynchronize(stream));
    return *scalHost;
  }
};

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
    if (preconditioner != FREHG_PC_JACOBI)
      frehg::ls_fail(FREHG_EUNSUPPORTED, "the COO seam uses point Jacobi");
    auto ls = std::make_unique<Ls>();
    ls->n = n_rows;
    ls->rtol = rtol;
    ls->atol = atol;
    ls->maxIt = max_iterations > 0 ? max_iterations : 500;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      frehg::ls_fail(FREHG_EDEVICE, "no HIP device is visible");
    LS_HIP(hipStreamCreateWithFlags(&ls->stream, hipStreamNonBlocking));
    ls->grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(512, (n_rows + 255) / 256)));
    ls->r = ls->alloc<double>(n_rows);
    ls->z = ls->alloc<double>(n_rows);
    ls->p = ls->alloc<double>(n_rows);
    ls->q = ls->alloc<double>(n_rows);
    ls->dinv = ls->alloc<double>(n_rows);
    ls->parts = ls->alloc<double>(ls->grid);
    ls->scal = ls->alloc<double>(1);
    LS_HIP(hipHostMalloc(reinterpret_cast<void**>(&ls->scalHost), sizeof(double),
                         hipHostMallocDefault));
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
    if (!ls || (n_coo > 0 && (!rows || !cols))) frehg::ls_fail(FREHG_EINVAL, "invalid arguments");
    if (ls->pattern) frehg::ls_fail(FREHG_EINVAL, "setPattern may be called once");
    std::vector<std::vector<std::pair<int64_t, int64_t>>> perRow(ls->n);
    for (int64_t e = 0; e < n_coo; ++e) {
      if (rows[e] < 0 || cols[e] < 0) continue;
      if (rows[e] >= ls->n || cols[e] >= ls->n)
        frehg::ls_fail(FREHG_EINVAL, "COO index out of range");
      perRow[rows[e]].push_back({cols[e], e});
    }
    std::vector<int64_t> slotPtr(1, 0), slotEntries;
    ls->hRp.assign(ls->n + 1, 0);
    ls->hCi.clear();
    for (int64_t r = 0; r < ls->n; ++r) {
      auto& v = perRow[r];
      std::sort(v.begin(), v.end());
      for (size_t a = 0; a < v.size(); ++a) {
        if (a == 0 || v[a].first != v[a - 1].first) {
          ls->hCi.push_back(v[a].first);
          slotPtr.push_back(slotPtr.back());
        }
        slotEntries.push_back(v[a].second);
        slotPtr.back() += 1;
      }
      ls->hRp[r + 1] = static_cast<int64_t>(ls->hCi.size());
    }
    ls->nCoo = n_coo;
    ls->nnz = static_cast<int64_t>(ls->hCi.size());
    ls->rp = ls->alloc<int64_t>(ls->n + 1);
    ls->ci = ls->alloc<int64_t>(ls->nnz);
    ls->slotPtr = ls->alloc<int64_t>(ls->nnz + 1);
    ls->slotEntries = ls->alloc<int64_t>(static_cast<int64_t>(slotEntries.size()));
    ls->val = ls->alloc<double>(ls->nnz);
    LS_HIP(hipMemcpy(ls->rp, ls->hRp.data(), (ls->n + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
    LS_HIP(hipMemcpy(ls->ci, ls->hCi.data(), ls->nnz * sizeof(int64_t), hipMemcpyHostToDevice));
    LS_HIP(hipMemcpy(ls->slotPtr, slotPtr.data(), slotPtr.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice));
    LS_HIP(hipMemcpy(ls->slotEntries, slotEntries.data(), slotEntries.size() * sizeof(int64_t),
                     hipMemcpyHostToDevice));
    ls->pattern = true;
  });
}

int frehg_ls_set_values(frehg_linear_system* h, const double* values_device) {
  return frehg::ls_guard([&] {
    Ls* ls = reinterpret_cast<Ls*>(h);
    if (!ls || !values_device) frehg::ls_fail(FREHG_EINVAL, "invalid arguments");
    if (!ls->pattern) frehg::ls_fail(FREHG_EINVAL, "setValues before setPattern");
    const int g = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(2048, (ls->nnz + 255) / 256)));
    frehg::k_gather_values<<<g, frehg::kT, 0, ls->stream>>>(ls->nnz, ls->slotPtr, ls->slotEntries,
                                                             values_device, ls->val);
    frehg::k_diag_inv<<<ls->grid, frehg::kT, 0, ls->stream>>>(ls->n, ls->rp, ls->ci, ls->val,
                                                              ls->dinv);
    LS_HIP(hipGetLastError());
    LS_HIP(hipStreamSynchronize(ls->stream));
    ls->values = true;
  });
}

int frehg_ls_solve(frehg_linear_system* h, const double* b, double* x, frehg_gw_solve_stats* stats) {
  return frehg::ls_guard([&] {
    Ls* ls = reinterpret_cast<Ls*>(h);
    if (!ls || !b || !x) frehg::ls_fail(FREHG_EINVAL, "invalid arguments");
    if (!ls->values) frehg::ls_fail(FREHG_EINVAL, "solve before setValues");
    const int64_t n = ls->n;
    hipStream_t s = ls->stream;
    frehg_gw_solve_stats st{};
    frehg::k_zero<<<ls->grid, frehg::kT, 0, s>>>(n, x);
    LS_HIP(hipMemcpyAsync(ls->r, b, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    const double bn = std::sqrt(ls->dot(b, b));
    const double tol = std::max(ls->rtol * bn, ls->atol);
    st.rhs_norm = bn;
    double rn = bn, rz = 0.0;
    bool restart = true;
    int it = 0;
    while (true) {
      if (rn <= tol) {
        frehg::k_csr_mv<<<ls->grid, frehg::kT, 0, s>>>(n, ls->rp, ls->ci, ls->val, x, b, ls->q);
        const double tn = std::sqrt(ls->dot(ls->q, ls->q));
        if (tn <= tol) {
          st.converged = 1;
          st.residual_norm = tn;
          break;
        }
        LS_HIP(hipMemcpyAsync(ls->r, ls->q, n * sizeof(double), hipMemcpyDeviceToDevice, s));
        rn = tn;
        restart = true;
      }
      if (it >= ls->maxIt) {
        st.residual_norm = rn;
        break;
      }
      frehg::k_zp<<<ls->grid, frehg::kT, 0, s>>>(n, ls->dinv, ls->r, ls->z, ls->p, -1.0);
      const double rzn = ls->dot(ls->r, ls->z);
      if (!restart)
        frehg::k_zp<<<ls->grid, frehg::kT, 0, s>>>(n, ls->dinv, ls->r, ls->z, ls->p, rzn / rz);
      restart = false;
      rz = rzn;
      frehg::k_csr_mv<<<ls->grid, frehg::kT, 0, s>>>(n, ls->rp, ls->ci, ls->val, ls->p, nullptr,
                                                    ls->q);
      const double alpha = rz / ls->dot(ls->p, ls->q);
      frehg::k_axpy2<<<ls->grid, frehg::kT, 0, s>>>(n, alpha, ls->p, ls->q, x, ls->r);
      LS_HIP(hipGetLastError());
      rn = std::sqrt(ls->dot(ls->r, ls->r));
      ++it;
    }
    st.iterations = it;
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
    LS_HIP(hipMemcpy(v.data(), ls->val, ls->nnz * sizeof(double), hipMemcpyDeviceToHost));
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

// Achievable-bandwidth probe: plain streaming kernels timed with HIP events
// on the caller's device, at the footprint of the PCG direction kernel.
//
// No counterpart in the reference.  It gives the measured ceiling that the
// stencil kernels are compared with beside the HBM peak: a copy (1 read,
// 1 write) and a stream with the direction kernel's mix of compulsory
// arrays (6 reads, 2 writes), all FP64, each array far larger than the
// on-package caches once the 8 arrays are counted together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/frehg_probe.h"

namespace {

__global__ __launch_bounds__(256) void k_copy(const double* __restrict__ a, double* __restrict__ b,
                                              int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_stream_6r2w(
    const double* __restrict__ a0, const double* __restrict__ a1, const double* __restrict__ a2,
    const double* __restrict__ a3, const double* __restrict__ a4, const double* __restrict__ a5,
    double* __restrict__ o0, double* __restrict__ o1, int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const double v0 = a0[i], v1 = a1[i], v2 = a2[i], v3 = a3[i], v4 = a4[i], v5 = a5[i];
    o0[i] = v0 + 0.5 * v1;
    o1[i] = v2 * v3 + v4 * v5;
  }
}

__global__ void k_fill(double* __restrict__ a, int64_t n, double v) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    a[i] = v + 1e-9 * static_cast<double>(i & 1023);
}

thread_local std::string g_probe_error;

struct Fail {
  int code;
};

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    g_probe_error = std::string(what) + ": " + hipGetErrorString(e);
    throw Fail{FREHG_PROBE_EDEVICE};
  }
}

}  // namespace

extern "C" int frehg_hbm_probe(int64_t n, int iterations, double* copy_gbs, double* stream_gbs) {
  if (n <= 0 || iterations <= 0 || !copy_gbs || !stream_gbs) {
    g_probe_error = "frehg_hbm_probe: n and iterations must be positive, outputs non-null";
    return FREHG_PROBE_EINVAL;
  }
  std::vector<double*> a(8, nullptr);
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FREHG_PROBE_OK;
  try {
    int dev = 0, cus = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
    check(hipStreamCreate(&s), "hipStreamCreate");
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    for (auto& p : a) check(hipMalloc(&p, static_cast<size_t>(n) * sizeof(double)), "hipMalloc");
    const int blocks = std::max(1, cus) * 16;
    for (int q = 0; q < 8; ++q) k_fill<<<blocks, 256, 0, s>>>(a[q], n, 1.0 + q);
    check(hipGetLastError(), "fill");
    auto time = [&](auto launch) {
      launch();  // warm-up
      check(hipEventRecord(e0, s), "hipEventRecord");
      for (int it = 0; it < iterations; ++it) launch();
      check(hipEventRecord(e1, s), "hipEventRecord");
      check(hipEventSynchronize(e1), "hipEventSynchronize");
      float ms = 0.f;
      check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      return static_cast<double>(ms) / iterations;
    };
    const double tc = time([&] { k_copy<<<blocks, 256, 0, s>>>(a[0], a[1], n); });
    const double ts = time([&] {
      k_stream_6r2w<<<blocks, 256, 0, s>>>(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], n);
    });
    check(hipGetLastError(), "probe kernels");
    const double bytes = static_cast<double>(n) * sizeof(double);
    *copy_gbs = 2.0 * bytes / (tc * 1e-3) / 1e9;
    *stream_gbs = 8.0 * bytes / (ts * 1e-3) / 1e9;
  } catch (const Fail& f) {
    rc = f.code;
  }
  for (auto p : a)
    if (p) (void)hipFree(p);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  return rc;
}

extern "C" const char* frehg_probe_last_error(void) { return g_probe_error.c_str(); }

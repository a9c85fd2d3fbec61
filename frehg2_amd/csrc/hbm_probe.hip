// Achievable-bandwidth probe: plain streaming kernels timed with HIP events
// on the caller's device, at the footprint of the PCG direction kernel.
//
// No counterpart in the reference.  It gives the measured ceiling that the
// stencil kernels are compared with beside the HBM peak.  frehg_hbm_probe
// times a copy (1 read, 1 write) and a stream with the direction kernel's mix
// of compulsory arrays (6 reads, 2 writes); frehg_hbm_probe_sweep times every
// read/write mix of the step's kernels (copy, read-only, write-only, 6r2w,
// 5r1w = an SpMV row) over the access width (8 or 16 bytes a lane), the
// cache policy (ordinary or non-temporal loads and stores) and the occupancy
// (1 to 32 blocks of 256 threads per CU), so the best of them is the part's
// own streaming rate at that mix.  All arrays are FP64, each far larger than
// the on-package caches once the arrays are counted together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/frehg_probe.h"

namespace {

constexpr int kMaxArrays = 8;
// 16 bytes a lane as a native vector (the non-temporal builtins take it)
typedef double dvec2 __attribute__((ext_vector_type(2)));

struct Arrays {
  const double* in[kMaxArrays];
  double* out[kMaxArrays];
};

template <class V, bool NT>
__device__ __forceinline__ V ld(const V* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <class V, bool NT>
__device__ __forceinline__ void st(V* p, V v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// NR reads and NW writes of n doubles; VEC = 1 (double) or 2 (double2).
// Each output is a different combination of the inputs so no load is dead;
// the read-only form folds everything into one value per thread, stored only
// if it is impossible (keeps the loads live without a write stream).
template <int NR, int NW, int VEC, bool NT>
__global__ __launch_bounds__(256) void k_mix(Arrays a, int64_t n, double* sink) {
  using V = typename std::conditional<VEC == 2, dvec2, double>::type;
  const int64_t nv = n / VEC;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  double acc = 0.0;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < nv;
       t += stride) {
    V v[NR > 0 ? NR : 1];
#pragma unroll
    for (int r = 0; r < NR; ++r) v[r] = ld<V, NT>(reinterpret_cast<const V*>(a.in[r]) + t);
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      V o;
      if constexpr (VEC == 2) {
        o.x = 0.0;
        o.y = 0.0;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const double c = 1.0 + 0.125 * ((r + w) & 3);
          o.x += c * v[r].x;
          o.y += c * v[r].y;
        }
        if (NR == 0) {
          o.x = static_cast<double>(t);
          o.y = 1.0;
        }
      } else {
        o = 0.0;
#pragma unroll
        for (int r = 0; r < NR; ++r) o += (1.0 + 0.125 * ((r + w) & 3)) * v[r];
        if (NR == 0) o = static_cast<double>(t);
      }
      st<V, NT>(reinterpret_cast<V*>(a.out[w]) + t, o);
    }
    if constexpr (NW == 0) {
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if constexpr (VEC == 2) acc += v[r].x + v[r].y;
        else acc += v[r];
      }
    }
  }
  if (NW == 0 && acc == -1.0) sink[0] = acc;  // never true for the filled data
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

using Launch = void (*)(const Arrays&, int64_t, double*, int, hipStream_t);

template <int NR, int NW, int VEC, bool NT>
void launch_mix(const Arrays& a, int64_t n, double* sink, int blocks, hipStream_t s) {
  k_mix<NR, NW, VEC, NT><<<blocks, 256, 0, s>>>(a, n, sink);
}

struct Pattern {
  int reads, writes;
  Launch fn[2][2];  // [vec 8/16 B][non-temporal]
};

template <int NR, int NW>
Pattern pattern() {
  return Pattern{NR, NW,
                 {{launch_mix<NR, NW, 1, false>, launch_mix<NR, NW, 1, true>},
                  {launch_mix<NR, NW, 2, false>, launch_mix<NR, NW, 2, true>}}};
}

// The probe's device state: eight arrays of n doubles, a stream and events.
struct Probe {
  std::vector<double*> a;
  double* sink = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int cus = 0;
  int64_t n = 0;

  Probe() : a(kMaxArrays, nullptr) {}
  // Allocation and fill; on a failure the destructor frees what was made.
  void init(int64_t n_) {
    n = n_;
    int dev = 0;
    check(hipGetDevice(&dev), "hipGetDevice");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
    check(hipStreamCreate(&s), "hipStreamCreate");
    check(hipEventCreate(&e0), "hipEventCreate");
    check(hipEventCreate(&e1), "hipEventCreate");
    for (auto& p : a) check(hipMalloc(&p, static_cast<size_t>(n) * sizeof(double)), "hipMalloc");
    check(hipMalloc(&sink, sizeof(double)), "hipMalloc");
    const int blocks = std::max(1, cus) * 16;
    for (int q = 0; q < kMaxArrays; ++q) k_fill<<<blocks, 256, 0, s>>>(a[q], n, 1.0 + q);
    check(hipGetLastError(), "fill");
  }
  ~Probe() {
    for (auto p : a)
      if (p) (void)hipFree(p);
    if (sink) (void)hipFree(sink);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
  }
  Arrays arrays(int reads, int writes) const {
    Arrays r{};
    for (int q = 0; q < reads; ++q) r.in[q] = a[q];
    for (int w = 0; w < writes; ++w) r.out[w] = a[reads + w];
    return r;
  }
  // Mean launch time in ms over `iterations` launches after one warm-up.
  template <class F>
  double time(int iterations, F&& launch) {
    launch();
    check(hipEventRecord(e0, s), "hipEventRecord");
    for (int it = 0; it < iterations; ++it) launch();
    check(hipEventRecord(e1, s), "hipEventRecord");
    check(hipEventSynchronize(e1), "hipEventSynchronize");
    check(hipGetLastError(), "probe kernels");
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    return static_cast<double>(ms) / iterations;
  }
};

}  // namespace

extern "C" int frehg_hbm_probe(int64_t n, int iterations, double* copy_gbs, double* stream_gbs) {
  if (n <= 0 || iterations <= 0 || !copy_gbs || !stream_gbs) {
    g_probe_error = "frehg_hbm_probe: n and iterations must be positive, outputs non-null";
    return FREHG_PROBE_EINVAL;
  }
  try {
    Probe p;
    p.init(n);
    const int blocks = std::max(1, p.cus) * 16;
    const Arrays c = p.arrays(1, 1), m = p.arrays(6, 2);
    const double tc = p.time(iterations, [&] { launch_mix<1, 1, 1, false>(c, n, p.sink, blocks, p.s); });
    const double ts = p.time(iterations, [&] { launch_mix<6, 2, 1, false>(m, n, p.sink, blocks, p.s); });
    const double bytes = static_cast<double>(n) * sizeof(double);
    *copy_gbs = 2.0 * bytes / (tc * 1e-3) / 1e9;
    *stream_gbs = 8.0 * bytes / (ts * 1e-3) / 1e9;
  } catch (const Fail& f) {
    return f.code;
  }
  return FREHG_PROBE_OK;
}

extern "C" int frehg_hbm_probe_sweep(int64_t n, int iterations, frehg_hbm_probe_result* out,
                                     int capacity, int* count) {
  if (n <= 1 || iterations <= 0 || !out || !count || capacity <= 0) {
    g_probe_error = "frehg_hbm_probe_sweep: n > 1, iterations > 0 and an output table are required";
    return FREHG_PROBE_EINVAL;
  }
  n &= ~int64_t(1);  // whole double2 vectors
  *count = 0;
  try {
    Probe p;
    p.init(n);
    const Pattern pats[] = {pattern<1, 1>(), pattern<1, 0>(), pattern<0, 1>(), pattern<6, 2>(),
                            pattern<5, 1>()};
    const int bpc[] = {1, 2, 4, 8, 16, 32};
    const double bytes = static_cast<double>(n) * sizeof(double);
    for (const Pattern& pt : pats) {
      const Arrays arr = p.arrays(pt.reads, pt.writes);
      for (int vec = 0; vec < 2; ++vec)
        for (int nt = 0; nt < 2; ++nt)
          for (int b : bpc) {
            if (*count >= capacity) return FREHG_PROBE_OK;
            const int blocks = std::max(1, p.cus) * b;
            const double ms = p.time(iterations, [&] { pt.fn[vec][nt](arr, n, p.sink, blocks, p.s); });
            frehg_hbm_probe_result& r = out[(*count)++];
            r.reads = pt.reads;
            r.writes = pt.writes;
            r.bytes_per_lane = vec ? 16 : 8;
            r.nontemporal = nt;
            r.blocks_per_cu = b;
            r.mean_ms = ms;
            r.gbs = (pt.reads + pt.writes) * bytes / (ms * 1e-3) / 1e9;
          }
    }
  } catch (const Fail& f) {
    return f.code;
  }
  return FREHG_PROBE_OK;
}

extern "C" const char* frehg_probe_last_error(void) { return g_probe_error.c_str(); }

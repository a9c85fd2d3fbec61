// gw_device.h — device-side building blocks of the MI355X Richards step:
// grid indexing in the halo layout, the van Genuchten-Mualem closures,
// wave64 / block reductions with compensated (Neumaier) accumulation and the
// XCD-aware work mapping.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gw_math.h"

namespace frehg {

constexpr int kWave = 64;
constexpr int kXcds = 8;  // MI355X: workgroups are dispatched round-robin over 8 XCDs

// Halo layout (src/core/Types.hpp:35-43): (j, i, k), one ghost in j and i,
// k unit stride; z-face fields carry nz + 1 planes (plane k = face above k).
struct Grid {
  int nx, ny, nz;
  int nx2, ny2;
  double dx, dy, az;
  __host__ __device__ int64_t c3(int j, int i, int k) const {
    return (static_cast<int64_t>(j) * nx2 + i) * nz + k;
  }
  __host__ __device__ int64_t cF(int j, int i, int k) const {
    return (static_cast<int64_t>(j) * nx2 + i) * (nz + 1) + k;
  }
  __host__ __device__ int c2(int j, int i) const { return j * nx2 + i; }
  __host__ __device__ int64_t cellsH() const { return static_cast<int64_t>(ny2) * nx2 * nz; }
  __host__ __device__ int64_t cellsF() const { return static_cast<int64_t>(ny2) * nx2 * (nz + 1); }
  __host__ __device__ int columns() const { return nx * ny; }
};

// ---------------------------------------------------------------------------
// van Genuchten-Mualem closures (src/gw/VanGenuchten.hpp:45-128).
//
// Branches, clamps and limiters are the reference's.  The powers are formed
// from shared logarithms instead of one pow() each: with t = (a|h|)^n,
//   Se            = (1 + t)^-m              = exp(-m log1p t)
//   Se^(1/m)      = 1 / (1 + t)             (exact identity)
//   1 - (1 - Se^(1/m))^m = 1 - (t/(1+t))^m  = -expm1(m log(t/(1+t)))
// with log(t/(1+t)) = n log(a|h|) - log1p t for t < 1 and -log1p(1/t)
// otherwise (1/t from t itself), so no term cancels.  Mualem's K then costs
// two log-class and two exp-class evaluations and a square root instead of
// five pow(), and C(h) none more:
//   C = a n m (ts - tr) (a|h|)^(n-1) / (1 + t)^(m+1)
//     = a n m (ts - tr) (t / (a|h|)) Se / (1 + t).
// h(theta) = -(1/a) (R^(1/m) - 1)^(1/n), R = (ts - tr)/(theta - tr), is
// formed as -(1/a) exp(log(expm1(log(R)/m))/n) (no cancellation in
// R^(1/m) - 1 near saturation).  Logarithms are gw_log / gw_log1p
// (gw_math.h, a few ulp).  The results agree with the pow() forms to ~1e-14
// relative (and are better conditioned for dry cells, where 1 - (1 - u)^m
// cancels).  The K/C pass is bound by its vector instructions on the GPU
// (~580 a cell and 85 % VALU busy with pow-class library calls): dropping
// C's two exponentials and one of K's, and the cheaper logarithms, took it
// from 0.52 to 0.38 ms at 512 x 512 x 100 (K alone 0.43 -> 0.32 ms).
// ---------------------------------------------------------------------------
struct VgSoil {
  double a, n, ts, tr, aev;
};

__device__ __forceinline__ double vg_theta(const VgSoil& s, double h) {
  const double m = 1.0 - 1.0 / s.n;
  double wc;
  if (h > s.aev) {
    wc = s.ts;
  } else {
    const double t = exp(s.n * gw_log(fabs(s.a * h)));
    const double se = exp(-m * gw_log1p(t));
    wc = s.tr + (s.ts - s.tr) * se;
  }
  if (wc > s.ts) wc = s.ts;
  else if (wc < s.tr) wc = s.tr;
  return wc;
}

__device__ __forceinline__ double vg_head(const VgSoil& s, double wc) {
  const double eps = 1.0e-7;
  const double m = 1.0 - 1.0 / s.n;
  if (wc - s.tr < eps) wc = s.tr + eps;
  if (wc < s.ts) {
    const double ratio = (s.ts - s.tr) / (wc - s.tr);
    const double x = expm1(gw_log(ratio) * (1.0 / m));  // ratio^(1/m) - 1
    return -(1.0 / s.a) * exp(gw_log(x) * (1.0 / s.n));
  }
  return 0.0;
}

// The Mualem factors shared by every Ks the cell carries: K(h, ks) =
// ((ks * sqrt_se) * br) * br, capped at ks and = ks for h > 0.
struct MualemFactors {
  double sq, br;
  bool sat;
};
__device__ __forceinline__ MualemFactors vg_mualem(const VgSoil& s, double h) {
  const double m = 1.0 - 1.0 / s.n;
  const double nl = s.n * gw_log(fabs(s.a * h));  // log t
  const double t = exp(nl);
  const double l1 = gw_log1p(t);
  const double se = exp(-m * l1);
  const double lr = t < 1.0 ? nl - l1 : -gw_log1p(1.0 / t);  // log(t / (1 + t))
  const double br = -expm1(m * lr);
  return MualemFactors{sqrt(se), br, h > 0.0};
}
// Mualem factors and C(h) together: C reuses t and Se (no further
// transcendental).  At h = 0, (a|h|)^(n-1) = 0 (n > 1) and so is C.
__device__ __forceinline__ MualemFactors vg_mualem_capacity(const VgSoil& s, double h,
                                                            double& cap) {
  const double m = 1.0 - 1.0 / s.n;
  const double ah = fabs(s.a * h);
  const double nl = s.n * gw_log(ah);
  const double t = exp(nl);
  const double l1 = gw_log1p(t);
  const double se = exp(-m * l1);
  const double lr = t < 1.0 ? nl - l1 : -gw_log1p(1.0 / t);
  const double br = -expm1(m * lr);
  if (h > 0.0 || !(ah > 0.0)) {
    cap = 0.0;
  } else {
    cap = s.a * s.n * m * (s.ts - s.tr) * (t / ah) * (se / (1.0 + t));
  }
  return MualemFactors{sqrt(se), br, h > 0.0};
}
__device__ __forceinline__ double vg_k_from(const MualemFactors& f, double ks) {
  double k = ks * f.sq * f.br * f.br;
  if (k > ks) k = ks;
  if (f.sat) k = ks;
  return k;
}
__device__ __forceinline__ double vg_conductivity(const VgSoil& s, double h, double ks) {
  return vg_k_from(vg_mualem(s, h), ks);
}

__device__ __forceinline__ double vg_dkdtheta(const VgSoil& s, double wc, double ks) {
  if (wc > 0.9999 * s.ts && wc < s.ts) wc = 0.9999 * s.ts;
  const double m = 1.0 - 1.0 / s.n;
  const double se = (wc - s.tr) / (s.ts - s.tr);
  const double t0 = pow(1.0 - pow(se, 1.0 / m), m);
  const double t1 = 0.5 * ks * pow(se, -0.5) * (1.0 - t0) * (1.0 - t0);
  const double t2 = 2.0 * ks * pow(se, (2.0 - m) / (2.0 * m)) * (1.0 - t0) *
                    pow(1.0 - pow(se, 1.0 / m), m - 1.0);
  return (t1 + t2) / (s.ts - s.tr);
}

// FaceConductivity.hpp:37-58 from the flanks' closures ka = K(h_a, ks_a).
__device__ __forceinline__ double face_k_harmonic(double ka, double ksa, double la, double kb,
                                                  double ksb, double lb) {
  if (ksa <= 0.0 || ksb <= 0.0) return 0.0;
  const double ksh = (la + lb) / (la / ksa + lb / ksb);
  return ksh * 0.5 * (ka / ksa + kb / ksb);
}

// ---------------------------------------------------------------------------
// Reductions.  Neumaier-compensated accumulation per thread, a fixed-shape
// wave64 butterfly, then a fixed-order combine of the block's waves: the
// result depends only on the launch shape, never on timing.
// ---------------------------------------------------------------------------
struct KSum {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double v) {
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
  }
  __device__ __forceinline__ double value() const { return s + c; }
};

// ---------------------------------------------------------------------------
// Lane movement with DPP (data-parallel primitives): a VALU move that reads
// another lane's register directly, with no LDS round trip (a ds_bpermute
// shuffle costs an LDS access and its latency on every step of a scan).
// Rows are groups of 16 lanes.  dpp_move<CTRL>(old, v): each lane takes v
// from the lane CTRL names; a lane whose source lies outside its row (row
// shifts) or outside the wave (wave shifts) keeps `old`.
// ---------------------------------------------------------------------------
constexpr int kDppRowShr = 0x110;    // row_shr:n = kDppRowShr + n: lane l <- l - n
constexpr int kDppRowShl = 0x100;    // row_shl:n = kDppRowShl + n: lane l <- l + n
constexpr int kDppWaveShr1 = 0x138;  // wave_shr:1: lane l <- l - 1
constexpr int kDppWaveShl1 = 0x130;  // wave_shl:1: lane l <- l + 1

// The builtins are typed int -> int: 4- and 8-byte values travel as 32-bit
// words (a value passed as double would be converted, not moved).
template <int CTRL, class T>
__device__ __forceinline__ T dpp_move(T old, T v) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "dpp_move moves 4- or 8-byte values");
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                             __builtin_bit_cast(int, v), CTRL,
                                                             0xF, 0xF, false));
  } else {
    const uint64_t o = __builtin_bit_cast(uint64_t, old), x = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
        static_cast<int>(static_cast<uint32_t>(o)), static_cast<int>(static_cast<uint32_t>(x)),
        CTRL, 0xF, 0xF, false));
    const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_update_dpp(
        static_cast<int>(static_cast<uint32_t>(o >> 32)),
        static_cast<int>(static_cast<uint32_t>(x >> 32)), CTRL, 0xF, 0xF, false));
    return __builtin_bit_cast(T, (static_cast<uint64_t>(hi) << 32) | lo);
  }
}

// v of lane l, wave-uniform.
template <class T>
__device__ __forceinline__ T lane_value(T v, int l) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "lane_value reads 4- or 8-byte values");
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
  } else {
    const uint64_t x = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = static_cast<uint32_t>(
        __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x)), l));
    const uint32_t hi = static_cast<uint32_t>(
        __builtin_amdgcn_readlane(static_cast<int>(static_cast<uint32_t>(x >> 32)), l));
    return __builtin_bit_cast(T, (static_cast<uint64_t>(hi) << 32) | lo);
  }
}

// One step of an affine scan inside rows: compose lane l's map y -> a + b y
// with the map S lanes before it (UP) or after it (!UP).  Lanes without such
// a neighbour in their row compose with the identity map (0, 1).
template <int S, bool UP, class T>
__device__ __forceinline__ void affine_step(T& a, T& b) {
  constexpr int ctrl = UP ? kDppRowShr + S : kDppRowShl + S;
  const T an = dpp_move<ctrl>(T(0), a);
  const T bn = dpp_move<ctrl>(T(1), b);
  a = a + b * an;
  b = b * bn;
}

// y_l = a_l + b_l y_{l-1} over the 64 lanes of a wave with y_{-1} = y0.
// Steps of 1, 2, 4 and 8 lanes compose every lane's map with all earlier maps
// of its row of 16 lanes; the values entering rows 1-3 then follow from the
// composed maps of lanes 15, 31 and 47.  Returns y_l; y0 becomes y_63.
template <class T>
__device__ __forceinline__ T scan_affine_up(T a, T b, T& y0, int lane) {
  affine_step<1, true>(a, b);
  affine_step<2, true>(a, b);
  affine_step<4, true>(a, b);
  affine_step<8, true>(a, b);
  const T e0 = y0;
  const T e1 = lane_value(a, 15) + lane_value(b, 15) * e0;
  const T e2 = lane_value(a, 31) + lane_value(b, 31) * e1;
  const T e3 = lane_value(a, 47) + lane_value(b, 47) * e2;
  y0 = lane_value(a, 63) + lane_value(b, 63) * e3;
  const T e = lane < 32 ? (lane < 16 ? e0 : e1) : (lane < 48 ? e2 : e3);
  return a + b * e;
}

// z_l = a_l + b_l z_{l+1} over the 64 lanes with z_64 = z0 (the mirror of
// scan_affine_up).  Returns z_l; z0 becomes z_0.
template <class T>
__device__ __forceinline__ T scan_affine_down(T a, T b, T& z0, int lane) {
  affine_step<1, false>(a, b);
  affine_step<2, false>(a, b);
  affine_step<4, false>(a, b);
  affine_step<8, false>(a, b);
  const T e3 = z0;
  const T e2 = lane_value(a, 48) + lane_value(b, 48) * e3;
  const T e1 = lane_value(a, 32) + lane_value(b, 32) * e2;
  const T e0 = lane_value(a, 16) + lane_value(b, 16) * e1;
  z0 = lane_value(a, 0) + lane_value(b, 0) * e0;
  const T e = lane < 32 ? (lane < 16 ? e0 : e1) : (lane < 48 ? e2 : e3);
  return a + b * e;
}

// The two scans above within each half of the wave: lanes 0-31 and 32-63
// hold two independent sequences (two columns), each entered with 0.  After
// the in-row steps only rows 1 and 3 take a value from the row before them.
template <class T>
__device__ __forceinline__ T scan_affine_up_halves(T a, T b, int lane) {
  affine_step<1, true>(a, b);
  affine_step<2, true>(a, b);
  affine_step<4, true>(a, b);
  affine_step<8, true>(a, b);
  const T e1 = lane_value(a, 15), e3 = lane_value(a, 47);
  const T e = (lane & 16) ? (lane < 32 ? e1 : e3) : T(0);
  return a + b * e;
}
template <class T>
__device__ __forceinline__ T scan_affine_down_halves(T a, T b, int lane) {
  affine_step<1, false>(a, b);
  affine_step<2, false>(a, b);
  affine_step<4, false>(a, b);
  affine_step<8, false>(a, b);
  const T e0 = lane_value(a, 16), e2 = lane_value(a, 48);
  const T e = (lane & 16) ? T(0) : (lane < 32 ? e0 : e2);
  return a + b * e;
}

// The wave reductions below and the stencil kernels' row shifts stay on LDS
// shuffles: with DPP moves in either place the FP64 PCG direction kernel
// (512 x 512 x 100) ran 5 % slower (0.398 -> 0.418 ms a launch), while the
// z-line scans above gained 21 % in the update kernel.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, kWave));
  return v;
}

// Block-wide reduction of NV values; every thread receives the results.
// op: 0 sum, 1 max, 2 min.  Requires blockDim.x a multiple of 64, <= 1024.
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], const int (&op)[NV]) {
  __shared__ double sh[16][NV];
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  const int nw = blockDim.x / kWave;
#pragma unroll
  for (int q = 0; q < NV; ++q)
    v[q] = op[q] == 0 ? wave_sum(v[q]) : (op[q] == 1 ? wave_max(v[q]) : wave_min(v[q]));
  __syncthreads();
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) sh[wid][q] = v[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double r = sh[0][q];
    for (int w = 1; w < nw; ++w)
      r = op[q] == 0 ? r + sh[w][q] : (op[q] == 1 ? fmax(r, sh[w][q]) : fmin(r, sh[w][q]));
    v[q] = r;
  }
  __syncthreads();
}

// XCD-aware remap of a 1D grid: the blocks an XCD receives (b % 8 == x)
// cover one contiguous range of logical blocks, so stencil neighbours a few
// rows apart are served from the same L2.
__device__ __forceinline__ int64_t xcd_block() {
  const int64_t b = blockIdx.x;
  const int64_t nb = gridDim.x;
  const int64_t x = b % kXcds;
  const int64_t per = nb / kXcds;
  const int64_t rem = nb % kXcds;
  return x * per + (x < rem ? x : rem) + b / kXcds;
}

}  // namespace frehg

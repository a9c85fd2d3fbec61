// gw_step.hip — the MI355X Richards groundwater step: device kernels for every
// phase of frehg::gw::RichardsSolver::step (src/gw/RichardsSolver.cpp:729-794),
// the host orchestration that stages a domain on device, and the C-ABI of
// include/frehg_gw.h.
//
// Kernel map (reference file:line each kernel restates):
//   k_ghosts            Predictor.cpp:757-817, Corrector.cpp:709-756
//   k_classify_coupled  Corrector.cpp:55-118 <PC>
//   k_cell_k            VanGenuchten.hpp:99-111, once per cell and phase
//   k_assemble          Predictor.cpp:105-698 <PC>: the predictor's face K
//                       (the rules of face_lat_at / face_z_at, specialised
//                       to an active cell) and the assembly into stencil
//                       storage (FP64, and FP32 for the mixed form)
//   (PCG)               gw_pcg.hip, replacing LinearSystem::solve
//   k_scatter           Predictor.cpp:745-754
//   k_faces             Predictor.cpp:127-386 and the Darcy fluxes of
//                       Corrector.cpp:135-533 (corrector; face_lat_at<DIR>
//                       and face_z_at for one face each)
//   k_coupled_top       Corrector.cpp:539-627
//   k_update_wc         Corrector.cpp:629-705 (with the root-uptake sink)
//   k_realloc_classify  Reallocate.cpp:181-391 (saturated_contact: head)
//   k_realloc_walk      Reallocate.cpp:410-540 and Corrector.cpp:758-792, one
//                       wave per column, the walks done in parallel where the
//                       result is the serial walk's (walk_parallel)
//   k_realloc_walk_serial  the serial walk for the columns left over
//   k_boundary_in       RichardsSolver.cpp:961-1006
//   k_adapt_inputs      AdaptiveStep.cpp:59-98
//   k_cpl_*             the coupling window around the step (include/
//                       frehg_coupler.h): src/coupling/Exchange.cpp:21-128
//                       and the surface depth rule src/swe/WetDry.cpp:24-43
//   k_density_faces     Baroclinic.cpp:46-187 (face means of the ratios)
// Density / viscosity ratios: the kernels that carry them are instantiated
// twice; without ratios (freshwater) each factor is the literal 1.0, which
// folds away, so that path reads no ratio array.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/frehg_coupler.h"
#include "../../include/frehg_gw.h"
#include "../../include/frehg_gw_testing.h"
#include "gw_device.h"
#include "gw_pcg.h"

namespace frehg {

// The last error message of this thread, shared by every entry point.
thread_local std::string g_error;
void set_error_message(const char* m) { g_error = m; }

namespace {

struct ApiError : std::runtime_error {
  int code;
  ApiError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(int code, const std::string& m) { throw ApiError(code, m); }

#define HIP_CHECK(x)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (x);                                                                \
    if (e_ != hipSuccess)                                                               \
      fail(FREHG_EDEVICE, std::string("HIP: ") + hipGetErrorString(e_) + " at " +      \
                              __FILE__ + ":" + std::to_string(__LINE__));               \
  } while (0)

constexpr int kThreads = 256;

// Reduction slots of one step (each filled by kRb per-block partials).
enum Slot {
  S_SS = 0, S_ADJ, S_DISC, S_LOST, S_VENT, S_THIN, S_VLOSS, S_BIN, S_DQMAX, S_COMIN,
  S_VOL, S_CEX, S_CBO, S_CEV, S_CLAMP, S_TR,
  S_XAPP, S_XDISC, S_XGAIN, S_XHELD, S_XSURF,  // coupling window (Exchange.cpp)
  S_COUNT
};
__constant__ int kSlotOp[S_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 2, 0, 0, 0, 0, 0, 0,
                                     0, 0, 0, 0, 0};

struct Dev {
  Grid g;
  double *h, *hn, *wc, *wcn, *ch;
  double *ksx, *ksy, *ksz, *vga, *vgn, *wcs, *wcr, *aev;
  // a van Genuchten field that is the same in every cell is kept as one value
  // (its pointer is then null): no kernel streams it
  double vgaC, vgnC, wcsC, wcrC, aevC;
  double *Kx, *Ky, *Kz;
  double *kx, *ky, *kzF, *qx, *qy, *qzF;
  double *vloss;
  double *hOut, *wcOut;  // second buffer pair written by reallocation
  double *sendUp, *sendDown;
  int* colFlag;
  unsigned char* sat;     // wc >= theta_s || h > aev entering reallocation
  double *dz3d, *bot3d, *ax, *ay, *sinx, *cosx, *siny, *cosy;
  double* bath;
  int* kt;
  int *topCode, *botCode, *sxm, *sxp, *sym, *syp, *cplMode;
  double *topValue, *botValue, *vxm, *vxp, *vym, *vyp;
  double *eta, *dept, *seep, *avail, *gain, *bounce;
  double minDepth;
  int coupled, ventAll;
  double ss, courantMax;
  int full3d, harmonic, reallocOff, redistribute, terrain, guessPrev;
  // predictor_corrector: legacy, the reference's <PC = false, HARM = false>
  // pair: a wet coupled top at 0.5 (Ksz + K), no pressurized seepage face
  // (mode 3), and corrector faces that give both flanks the minus (lateral)
  // or lower (vertical) cell's Ks
  int pcLegacy;
  // the root-uptake sink S [1/s] per cell (H), null when none is set: the
  // predictor takes dtg S V from the rhs (Predictor.cpp:700-726) and the
  // corrector from the theta update (Corrector.cpp:674-705)
  double* sink;
  // density / viscosity ratios (Baroclinic.cpp), null when none are set (all
  // 1): cells rRho, rVisc (H); face means rXp, vXp, rYp, vYp (H, the slot of
  // the face east / north of the cell) and rZp, vZp (F, plane k = the face
  // above cell k).  densGrad: the reallocation's vertical head gradient
  // subtracts rZp instead of 1 (reallocation_gradient: density)
  double *rRho, *rVisc, *rXp, *rYp, *vXp, *vYp, *rZp, *vZp;
  int densGrad;
  // the evapotranspiration soil channel (RichardsSolver.hpp:620-636), null
  // when off: columns with etCol != 0 take the applied flux etQ [m/s] at the
  // top face, guard-free (2D halo layout)
  int* etCol;
  double* etQ;
  double *Ad, *Axp, *Ayp, *Azp, *rhs, *x;
  float *Ad32, *Axp32, *Ayp32, *Azp32;  // the stencil rounded to FP32 (mixed form), or null
  double* parts;
  int rb;
  // the transport window (RichardsSolver.cpp:796-878), when tw: face volumes
  // since the last reset (winX, winY, winSink: H; winZ: F; winVent, winVap:
  // 2D) and the walks' path corrections of the step (pathCorr: F)
  int tw;
  double *winX, *winY, *winZ, *winSink, *pathCorr, *winVent, *winVap;
  // the step's dt [s], in device memory (a replayed graph reads the value of
  // the step it runs), and the gate of the kernels after the solve: when set
  // and nonzero they return at once (the solve did not converge inside a
  // device-controlled budget and the host finishes the step)
  const double* dtgp;
  const int* gate;
};

// The kernels after the solve return at once when their gate is closed.
__device__ __forceinline__ bool gated(const Dev& d) { return d.gate && *d.gate; }

__device__ __forceinline__ bool act(const Dev& d, int j, int i, int k) {
  return k >= d.kt[d.g.c2(j, i)];
}
__device__ __forceinline__ double wcs_at(const Dev& d, int64_t c) {
  return d.wcs ? d.wcs[c] : d.wcsC;
}
__device__ __forceinline__ double wcr_at(const Dev& d, int64_t c) {
  return d.wcr ? d.wcr[c] : d.wcrC;
}
__device__ __forceinline__ double aev_at(const Dev& d, int64_t c) {
  return d.aev ? d.aev[c] : d.aevC;
}
__device__ __forceinline__ VgSoil soil_at(const Dev& d, int64_t c) {
  return VgSoil{d.vga ? d.vga[c] : d.vgaC, d.vgn ? d.vgn[c] : d.vgnC, wcs_at(d, c), wcr_at(d, c),
                aev_at(d, c)};
}
__device__ __forceinline__ bool is_head(int code) {
  return code == FREHG_BC_HEAD || code == FREHG_BC_HEAD_HYDROSTATIC;
}

// XCD-aware grid-stride range over [0, n): each XCD walks one contiguous
// eighth of the index space.
struct Range {
  int64_t t, end, stride;
};
__device__ __forceinline__ Range xcd_range(int64_t n) {
  const int xcd = blockIdx.x % kXcds;
  const int bpx = gridDim.x / kXcds;
  const int local = blockIdx.x / kXcds;
  Range r;
  r.t = static_cast<int64_t>(xcd) * n / kXcds + static_cast<int64_t>(local) * blockDim.x +
        threadIdx.x;
  r.end = static_cast<int64_t>(xcd + 1) * n / kXcds;
  r.stride = static_cast<int64_t>(bpx) * blockDim.x;
  return r;
}

// Column order of the assembly and the theta update, which read their j-1
// neighbours: i-tiles of kTileW columns, each walked j-major.  An XCD's consecutive
// threads then meet a cell's j-1 neighbour kTileW * nz cells back (still in
// that XCD's L2) instead of a whole (nx+2) * nz row back (assembly 1.31 ->
// 1.25 ms at 512 x 512 x 100; the reallocation's classification, bound by
// vector instructions, and the face pass gain nothing).  A bijection of [0, nx * ny) onto the
// interior columns.
constexpr int kTileW = 64;
__device__ __forceinline__ void tiled_column(const Grid& g, int col, int& j, int& i) {
  const int tileCols = g.ny * kTileW;
  const int t = col / tileCols;
  const int base = t * kTileW;
  const int w = min(kTileW, g.nx - base);
  const int rem = col - t * tileCols;
  j = rem / w + 1;
  i = base + rem % w + 1;
}

template <int NV>
__device__ __forceinline__ void put_parts(const Dev& d, const int (&slot)[NV], double (&v)[NV]) {
  int op[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) op[q] = kSlotOp[slot[q]];
  block_reduce<NV>(v, op);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) d.parts[static_cast<int64_t>(slot[q]) * d.rb + blockIdx.x] = v[q];
  }
}

// Adds to the partials put_parts wrote in an earlier kernel of the same grid
// (sum slots only).
template <int NV>
__device__ __forceinline__ void add_parts(const Dev& d, const int (&slot)[NV], double (&v)[NV]) {
  int op[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) op[q] = kSlotOp[slot[q]];
  block_reduce<NV>(v, op);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) d.parts[static_cast<int64_t>(slot[q]) * d.rb + blockIdx.x] += v[q];
  }
}

// ---- ghosts (Predictor.cpp:66-87, :757-817; Corrector.cpp:709-756) --------
// rFace: the boundary face's density ratio; the plus edges scale the depth
// term by it, the minus edges do not (Predictor.cpp:56-62, 66-87)
__device__ __forceinline__ double ghost_head(int code, double value, double interior, double zc,
                                             bool coupled, double bed, double depth, double rFace,
                                             bool plus) {
  if (code == FREHG_BC_HEAD) return value;
  if (code == FREHG_BC_HEAD_HYDROSTATIC) {
    if (coupled && value > bed) {
      const double hydro = (bed - zc) * rFace;
      return plus ? hydro + depth * rFace : hydro + depth;
    }
    return (value - zc) * rFace;
  }
  return interior;
}

__global__ void k_ghosts(Dev d, int doH, int doWc) {
  if (gated(d)) return;
  const Grid& g = d.g;
  const int64_t nxe = static_cast<int64_t>(g.ny) * g.nz;
  const int64_t n = nxe + static_cast<int64_t>(g.nx) * g.nz;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const bool xEdge = t < nxe;
    const int64_t u = xEdge ? t : t - nxe;
    const int k = static_cast<int>(u % g.nz);
    const int a = static_cast<int>(u / g.nz) + 1;
    for (int side = 0; side < 2; ++side) {
      int ji, ii, jg, ig, code;
      double value;
      if (xEdge) {
        ji = jg = a;
        ii = side == 0 ? 1 : g.nx;
        ig = side == 0 ? 0 : g.nx + 1;
        code = side == 0 ? d.sxm[g.c2(ji, ii)] : d.sxp[g.c2(ji, ii)];
        value = side == 0 ? d.vxm[g.c2(ji, ii)] : d.vxp[g.c2(ji, ii)];
      } else {
        ii = ig = a;
        ji = side == 0 ? 1 : g.ny;
        jg = side == 0 ? 0 : g.ny + 1;
        code = side == 0 ? d.sym[g.c2(ji, ii)] : d.syp[g.c2(ji, ii)];
        value = side == 0 ? d.vym[g.c2(ji, ii)] : d.vyp[g.c2(ji, ii)];
      }
      const int64_t cin = g.c3(ji, ii, k), cg = g.c3(jg, ig, k);
      if (doH) {
        const double zc = d.bot3d[cin] + 0.5 * d.dz3d[cin];
        const double dep = d.coupled ? d.dept[g.c2(ji, ii)] : 0.0;
        const double* rf = xEdge ? d.rXp : d.rYp;  // face slot: the ghost's (minus), the cell's (plus)
        const double rFace = rf ? rf[side == 0 ? cg : cin] : 1.0;
        d.h[cg] = ghost_head(code, value, d.h[cin], zc, d.coupled != 0, d.bath[g.c2(ji, ii)], dep,
                             rFace, side == 1);
      }
      if (doWc) d.wc[cg] = is_head(code) ? wcs_at(d, cin) : d.wc[cin];
    }
  }
}

// ---- coupled-top classification (Corrector.cpp:55-118) -------------------
__global__ void k_classify_coupled(Dev d) {
  const double dtg = *d.dtgp;
  const Grid& g = d.g;
  const int n = g.columns();
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int j = t / g.nx + 1, i = t % g.nx + 1;
    const int c2 = g.c2(j, i);
    const int kt = d.kt[c2] < g.nz ? d.kt[c2] : -1;
    const double dep = d.dept[c2];
    if (kt < 0 || dep <= 0.0) {
      // a pressurized dry top is a seepage face (mode 3), not under legacy
      bool pressurized = false;
      if (!d.pcLegacy && kt >= 0) {
        const int64_t c = g.c3(j, i, kt);
        const double rz = d.rZp ? d.rZp[g.cF(j, i, kt)] : 1.0;
        pressurized = d.h[c] > rz * 0.5 * d.dz3d[c];
      }
      d.cplMode[c2] = pressurized ? 3 : 0;
      continue;
    }
    const int64_t c = g.c3(j, i, kt);
    const double ks = d.ksz[c];
    const double qEst = 2.0 * ks * (d.h[c] - dep) / d.dz3d[c] - ks;
    d.cplMode[c2] = (qEst < 0.0 && -qEst * dtg > d.avail[c2]) ? 2 : 1;
  }
}

// ---- cells gathered into full waves ----------------------------------------
// A closure's logarithms and exponentials are needed in some cells only (the
// unsaturated ones), and the cell kernels' waves run over consecutive rows,
// so nearly every wave holds a few such cells and the whole wave pays for
// them.  A kernel can instead do the cheap cells at once and append the
// others to its wave's ring of up to kRing cell indices in LDS (head and cnt
// wave-uniform): whenever kWave entries wait, the wave evaluates them
// together.  Each cell's arithmetic is unchanged.
constexpr int kRing = 2 * kWave;
struct WaveRing {
  int64_t* q;
  int head, cnt;
};
__device__ __forceinline__ void ring_push(WaveRing& r, bool want, int64_t val) {
  const unsigned long long m = __ballot(want);
  const unsigned lo = static_cast<unsigned>(m), hi = static_cast<unsigned>(m >> 32);
  const int below = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0));
  if (want) r.q[(r.head + r.cnt + below) & (kRing - 1)] = val;
  r.cnt += __popcll(m);
}
// The entry of this lane in the next full wave's worth (or in the remainder
// when last), and whether the lane has one; advances the ring.
__device__ __forceinline__ bool ring_take(WaveRing& r, int lane, int64_t& val) {
  __builtin_amdgcn_wave_barrier();
  const bool has = lane < r.cnt;
  if (has) val = r.q[(r.head + lane) & (kRing - 1)];
  __builtin_amdgcn_wave_barrier();
  const int took = r.cnt < kWave ? r.cnt : kWave;
  r.head = (r.head + took) & (kRing - 1);
  r.cnt -= took;
  return has;
}

// ---- per-cell conductivities, halo cells included -------------------------
// With CAP (the predictor, where h is h^n) it also stores C(h^n) for the
// assembly and copies the step's start state into (hn, wcn).  A cell with
// h > 0 takes K = Ks and C = 0 at once (what vg_k_from and
// vg_mualem_capacity return there); the others go through the wave's ring
// and are evaluated 64 at a time (303 / 256 us against 379 /
// 328 us for the closure in every lane, with and without C, at 512 x 512 x
// 100, where 60 % of the cells are saturated).
template <bool ISO, bool CAP>
__global__ __launch_bounds__(kThreads) void k_cell_k(Dev d) {
  if (gated(d)) return;
  __shared__ int64_t ring[kThreads / kWave][kRing];
  const int lane = threadIdx.x & (kWave - 1);
  WaveRing wr{ring[threadIdx.x / kWave], 0, 0};
  const int64_t n = d.g.cellsH();
  auto eval = [&](int64_t c) {
    MualemFactors f;
    if (CAP) {
      double cap;
      f = vg_mualem_capacity(soil_at(d, c), d.h[c], cap);
      d.ch[c] = cap;
    } else {
      f = vg_mualem(soil_at(d, c), d.h[c]);
    }
    d.Kx[c] = vg_k_from(f, d.ksx[c]);
    if (!ISO) {
      d.Ky[c] = vg_k_from(f, d.ksy[c]);
      d.Kz[c] = vg_k_from(f, d.ksz[c]);
    }
  };
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t c0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x - lane; c0 < n;
       c0 += stride) {
    const int64_t c = c0 + lane;
    bool slow = false;
    if (c < n) {
      const double h = d.h[c];
      if (CAP) {
        // the step's start state (h^n, theta^n), halo included
        d.hn[c] = h;
        d.wcn[c] = d.wc[c];
      }
      if (h > 0.0) {
        d.Kx[c] = d.ksx[c];
        if (!ISO) {
          d.Ky[c] = d.ksy[c];
          d.Kz[c] = d.ksz[c];
        }
        if (CAP) d.ch[c] = 0.0;
      } else {
        slow = true;
      }
    }
    ring_push(wr, slow, c);
    if (wr.cnt >= kWave) {
      int64_t cq;
      if (ring_take(wr, lane, cq)) eval(cq);
    }
  }
  if (wr.cnt > 0) {
    int64_t cq;
    if (ring_take(wr, lane, cq)) eval(cq);
  }
}

__device__ __forceinline__ bool unsat(const Dev& d, int j, int i, int k) {
  const int64_t c = d.g.c3(j, i, k);
  return act(d, j, i, k) && d.wc[c] < wcs_at(d, c);
}

// ---- lateral faces: K (Predictor.cpp:127-180, 236-268, 318-374) and, with
// corr, the Darcy flux (Corrector.cpp:176-366).  dir 0 = x, 1 = y.
// With DENS the Darcy flux carries the face's viscosity ratio, and its
// along-layer gravity term r v (consistent pair) (Corrector.cpp:194-247).
template <int DIR, bool DENS>
__device__ __forceinline__ void face_lat_at(const Dev& d, int corr, int j, int i,
                                            int k) {
  const Grid& g = d.g;
  const double* K = DIR == 0 ? d.Kx : d.Ky;
  const double* ks = DIR == 0 ? d.ksx : d.ksy;
  const double* area = DIR == 0 ? d.ax : d.ay;
  const double* cosv = DIR == 0 ? d.cosx : d.cosy;
  const double* sinv = DIR == 0 ? d.sinx : d.siny;
  double* kf = DIR == 0 ? d.kx : d.ky;
  double* qf = DIR == 0 ? d.qx : d.qy;
  const int nlast = DIR == 0 ? g.nx : g.ny;
  const double spacing = DIR == 0 ? g.dx : g.dy;
  const int in = DIR == 0 ? i + 1 : i, jn = DIR == 0 ? j : j + 1, s = DIR == 0 ? i : j;
  {
    const int64_t c = g.c3(j, i, k), cn = g.c3(jn, in, k);
    // every operand of the face, loaded once up front (c and cn lie in the halo)
    const double Kc = K[c], Kn = K[cn];
    const double A = area[c], cv = cosv[c], sv = sinv[c];
    const double hc = d.h[c], hn = d.h[cn], b3c = d.bot3d[c], b3n = d.bot3d[cn];
    const double base = d.harmonic ? face_k_harmonic(Kc, ks[c], 1.0, Kn, ks[cn], 1.0)
                                   : 0.5 * (Kc + Kn);
    double v = base;
    // domain-edge override (Predictor.cpp:236-268)
    if (s == 0) {
      const int jj = DIR == 0 ? j : 1, ii = DIR == 0 ? 1 : i;
      const int code = DIR == 0 ? d.sxm[g.c2(jj, ii)] : d.sym[g.c2(jj, ii)];
      if (code == FREHG_BC_NOFLUX) v = 0.0;
      else if (is_head(code)) v = Kn;  // cell (jj, ii) is the face's inner cell cn
    }
    if (s == nlast) {
      const int code = DIR == 0 ? d.sxp[g.c2(j, i)] : d.syp[g.c2(j, i)];
      if (code == FREHG_BC_NOFLUX) v = 0.0;
      else if (is_head(code)) v = Kc;
    }
    const bool uns = !d.full3d && (unsat(d, j, i, k) || unsat(d, jn, in, k));
    if (uns) v = 0.0;
    const bool inA = s >= 1 && s <= nlast, inB = s + 1 >= 1 && s + 1 <= nlast;
    const bool actA = act(d, j, i, k), actB = act(d, jn, in, k);
    if ((inA && !actA) || (inB && !actB)) v = 0.0;
    kf[c] = v;
    if (!corr) return;
    // Darcy flux (Corrector.cpp:176-366), positive toward -axis
    double q;
    const double vf = DENS ? (DIR == 0 ? d.vXp : d.vYp)[c] : 1.0;
    const double rf = DENS ? (DIR == 0 ? d.rXp : d.rYp)[c] : 1.0;
    if (s == 0) {
      // the inner cell (j, 1) / (1, i) is cn and the ghost cell is c
      const double kface = Kn;
      q = A * kface * vf * cv * (hn - hc) / (0.5 * spacing);
      const int code = DIR == 0 ? d.sxm[g.c2(j, 1)] : d.sym[g.c2(1, i)];
      const double val = DIR == 0 ? d.vxm[g.c2(j, 1)] : d.vym[g.c2(1, i)];
      if (code == FREHG_BC_NOFLUX) q = 0.0;
      else if (code == FREHG_BC_FLUX) q = -val * A;
    } else {
      double kface, delta = spacing;
      const int sgn = (b3n > b3c) ? 1 : -1;
      if (s == nlast) {
        kface = Kc;
        delta = 0.5 * spacing;
      } else {
        kface = base;
        if (d.pcLegacy && ks[cn] != ks[c])  // legacy: the minus cell's Ks on both flanks
          kface = 0.5 * (Kc + vg_conductivity(soil_at(d, cn), hn, ks[c]));
        if (!actA || !actB) kface = 0.0;
      }
      const double rv = d.pcLegacy ? 1.0 : rf * vf;  // legacy: no ratios on the gravity term
      q = A * (kface * vf * cv * (hn - hc) / delta + static_cast<double>(sgn) * kface * rv * sv);
      if (uns) q = 0.0;
      if (s == nlast) {
        const int code = DIR == 0 ? d.sxp[g.c2(j, i)] : d.syp[g.c2(j, i)];
        const double val = DIR == 0 ? d.vxp[g.c2(j, i)] : d.vyp[g.c2(j, i)];
        if (code == FREHG_BC_NOFLUX) q = 0.0;
        else if (code == FREHG_BC_FLUX) q = val * A;
      }
    }
    qf[c] = q;
  }
}

// ---- vertical faces: K (Predictor.cpp:181-232, 280-313, 375-385) and, with
// corr, the Darcy flux (Corrector.cpp:373-527) with the coupled top modes.
template <bool DENS>
__device__ __forceinline__ void face_z_at(const Dev& d, int corr, double dtg, int j, int i,
                                          int k, double& clamp) {
  const Grid& g = d.g;
  {
    const int c2 = g.c2(j, i);
    const int kt = d.kt[c2];
    const bool actK = k < g.nz && k >= kt;
    const bool actM = k >= 1 && k - 1 >= kt;
    // the cells above (ca) and below (cb) the face, loaded once up front; at
    // k = 0 and k = nz they are the neighbouring columns' end rows, in bounds
    const int64_t cb = g.c3(j, i, 0) + k, ca = cb - 1;
    const double Kza = d.Kz[ca], Kzb = d.Kz[cb], ha = d.h[ca], hb = d.h[cb];
    const double dza = d.dz3d[ca], dzb = d.dz3d[cb];
    // the column's top and bottom inputs (2D), loaded with the rest rather
    // than behind kt and the face-K store in the top and bottom faces' rules
    const int topc = d.topCode[c2], botc = d.botCode[c2];
    const double topv = d.topValue[c2], botv = d.botValue[c2];
    const bool etc = d.etCol && d.etCol[c2];
    const double etq = d.etCol ? d.etQ[c2] : 0.0;
    const double base = d.harmonic ? face_k_harmonic(Kza, d.ksz[ca], dza, Kzb, d.ksz[cb], dzb)
                                   : 0.5 * (Kza + Kzb);
    double v;
    if (k == 0) {
      v = 0.0;
    } else if (k == g.nz) {
      v = botc == FREHG_BC_NOFLUX ? 0.0 : Kza;
    } else if (actK && !actM) {
      v = Kzb;
    } else if (!actM) {
      v = 0.0;
    } else {
      v = base;
    }
    if (actK && (k == 0 || !actM)) {  // top-face rule
      const int64_t c = cb;
      const double own = Kzb;
      if (d.coupled) {
        const double wet = d.pcLegacy ? 0.5 * (d.ksz[c] + own) : d.ksz[c];
        v = d.dept[c2] > 0.0 ? wet : own;
      } else {
        const int code = topc;
        if (code == FREHG_BC_HEAD) v = d.ksz[c];
        else if (code == FREHG_BC_NOFLUX) v = 0.0;
        else v = own;
      }
    }
    if ((k >= 1 && !actM) || (k < g.nz && !actK)) v = 0.0;  // seal, last
    d.kzF[g.cF(j, i, k)] = v;
    if (!corr) return;
    double q = 0.0;
    const int64_t fz = g.cF(j, i, k);
    const double rz = DENS ? d.rZp[fz] : 1.0, vz = DENS ? d.vZp[fz] : 1.0;
    if (actK && !actM) {
      const int64_t c = cb;
      if (d.coupled) {
        const double ks = d.ksz[c];
        const double depth = d.dept[c2];
        const int mode = d.cplMode[c2];
        const double delta = 0.5 * dzb;
        if (mode == 2) {
          const double roomTop = fmax(0.0, (0.9999 * wcs_at(d, c) - d.wc[c]) * dzb);
          const double take = fmin(d.avail[c2], roomTop);
          q = -take * g.az / dtg;
          d.avail[c2] -= take;
        } else if (mode == 1) {
          q = g.az * ks * vz * ((hb - depth) / delta - rz);
          if (q < 0.0) {
            const double need = -q * dtg / g.az;
            if (need > d.avail[c2]) {
              q = -d.avail[c2] * g.az / dtg;
              d.avail[c2] = 0.0;
            } else {
              d.avail[c2] -= need;
            }
          }
        } else if (mode == 3) {
          q = g.az * ks * vz * ((hb - depth) / delta - rz);
          if (q <= 0.0) {
            clamp += -q * dtg;
            q = 0.0;
          }
        } else {
          const double kG = vg_conductivity(soil_at(d, c), depth, ks);
          const double kface = 0.5 * (kG + Kzb);
          q = g.az * kface * vz * ((hb - depth) / delta - rz);
          if (q <= 0.0) q = 0.0;
          if (etc) {
            q += etq * g.az;
          } else if (topc == FREHG_BC_FLUX) {
            const double tv = topv;
            if ((tv > 0.0 && d.wc[c] > wcr_at(d, c)) || tv < 0.0) q += tv * g.az;
          }
        }
      } else {
        const int code = topc;
        if (etc) {
          q = etq * g.az;
        } else if (code == FREHG_BC_FLUX) {
          q = topv * g.az;
        } else if (code == FREHG_BC_HEAD) {
          q = g.az * d.ksz[c] * vz * ((hb - topv) / (0.5 * dzb) - rz);
        }
      }
    } else if (k == g.nz && actM) {
      const double kface = Kza;
      const int code = botc;
      if (code == FREHG_BC_FLUX) q = botv * g.az;
      else if (code == FREHG_BC_GRAVITY) q = -kface * vz * rz * g.az;
      else if (code == FREHG_BC_HEAD)
        q = g.az * kface * vz * ((botv - ha) / (0.5 * dza) - rz);
    } else if (actM && actK) {
      double kface = base;
      if (d.pcLegacy && d.ksz[ca] != d.ksz[cb])  // legacy: the lower cell's Ks on both flanks
        kface = 0.5 * (vg_conductivity(soil_at(d, ca), ha, d.ksz[cb]) + Kzb);
      q = g.az * kface * vz * ((hb - ha) / (0.5 * (dza + dzb)) - rz);
    }
    d.qzF[fz] = q;
  }
}

// The corrector's three face passes in one: thread (j, i, k) of
// (ny+1) x (nx+1) x (nz+1) takes the x face east of (j, i) (j >= 1, k < nz),
// the y face north of it (i >= 1, k < nz) and the z face above (j, i, k)
// (j, i >= 1), so the operands the three share are fetched once (1.02 ms
// against 0.36 + 0.48 + 0.27 ms for three kernels at 512 x 512 x 100).
template <bool DENS>
__global__ __launch_bounds__(kThreads) void k_faces(Dev d, int corr) {
  if (gated(d)) return;
  const double dtg = *d.dtgp;
  const Grid& g = d.g;
  const int nzf = g.nz + 1, nxa = g.nx + 1;
  const int64_t n = static_cast<int64_t>(g.ny + 1) * nxa * nzf;
  double clamp = 0.0;
  for (Range r = xcd_range(n); r.t < r.end; r.t += r.stride) {
    const int k = static_cast<int>(r.t % nzf);
    const int64_t col = r.t / nzf;
    const int i = static_cast<int>(col % nxa), j = static_cast<int>(col / nxa);
    if (k < g.nz) {
      if (j >= 1) face_lat_at<0, DENS>(d, corr, j, i, k);
      if (i >= 1) face_lat_at<1, DENS>(d, corr, j, i, k);
    }
    if (j >= 1 && i >= 1) face_z_at<DENS>(d, corr, dtg, j, i, k, clamp);
  }
  if (corr) {
    double v[1] = {clamp};
    const int s[1] = {S_CLAMP};
    put_parts<1>(d, s, v);
  }
}

// ---- density / viscosity face means (Baroclinic.cpp:46-187) ---------------
__global__ void k_fill(double* p, int64_t n, double v, const int* gate) {
  if (gate && *gate) return;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x)
    p[t] = v;
}

// Thread (j, i, k) of (ny+1) x (nx+1) x (nz+1) sets the x face east of (j, i),
// the y face north of it and the z face above (j, i, k), as k_faces indexes
// them.  Arithmetic means between active flanks; a minus-side edge face
// averages the interior cell with its ghost (the ghost's own ratio under a
// head condition), a plus-side edge face is 1 unless the edge holds a head
// condition (then the ghost's ratio); the face above a column's top cell
// averages the top cell with the ghost above it (top[] when given, else the
// top cell's own ratio); the bottom face stays 1.  Slots no rule reaches
// keep the 1 they were filled with.
__global__ void k_density_faces(Dev d, const double* rrTop, const double* rvTop) {
  const Grid& g = d.g;
  const int nzf = g.nz + 1, nxa = g.nx + 1;
  const int64_t n = static_cast<int64_t>(g.ny + 1) * nxa * nzf;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int k = static_cast<int>(t % nzf);
    const int64_t col = t / nzf;
    const int i = static_cast<int>(col % nxa), j = static_cast<int>(col / nxa);
    if (k < g.nz && j >= 1) {  // x face east of (j, i), i in [0, nx]
      const int64_t c = g.c3(j, i, k), ce = g.c3(j, i + 1, k);
      double r = 1.0, v = 1.0;
      if (i == 0) {
        if (act(d, j, 1, k)) {
          if (is_head(d.sxm[g.c2(j, 1)])) {
            r = d.rRho[c];
            v = d.rVisc[c];
          } else {
            r = 0.5 * (d.rRho[ce] + d.rRho[c]);
            v = 0.5 * (d.rVisc[ce] + d.rVisc[c]);
          }
        }
      } else if (i == g.nx) {
        if (act(d, j, i, k) && is_head(d.sxp[g.c2(j, i)])) {
          r = d.rRho[ce];
          v = d.rVisc[ce];
        }
      } else if (act(d, j, i, k) && act(d, j, i + 1, k)) {
        r = 0.5 * (d.rRho[c] + d.rRho[ce]);
        v = 0.5 * (d.rVisc[c] + d.rVisc[ce]);
      }
      d.rXp[c] = r;
      d.vXp[c] = v;
    }
    if (k < g.nz && i >= 1) {  // y face north of (j, i), j in [0, ny]
      const int64_t c = g.c3(j, i, k), cn = g.c3(j + 1, i, k);
      double r = 1.0, v = 1.0;
      if (j == 0) {
        if (act(d, 1, i, k)) {
          if (is_head(d.sym[g.c2(1, i)])) {
            r = d.rRho[c];
            v = d.rVisc[c];
          } else {
            r = 0.5 * (d.rRho[cn] + d.rRho[c]);
            v = 0.5 * (d.rVisc[cn] + d.rVisc[c]);
          }
        }
      } else if (j == g.ny) {
        if (act(d, j, i, k) && is_head(d.syp[g.c2(j, i)])) {
          r = d.rRho[cn];
          v = d.rVisc[cn];
        }
      } else if (act(d, j, i, k) && act(d, j + 1, i, k)) {
        r = 0.5 * (d.rRho[c] + d.rRho[cn]);
        v = 0.5 * (d.rVisc[c] + d.rVisc[cn]);
      }
      d.rYp[c] = r;
      d.vYp[c] = v;
    }
    if (j >= 1 && i >= 1) {  // z face above (j, i, k), k in [0, nz]
      const bool below = k < g.nz && act(d, j, i, k);
      const bool above = k >= 1 && act(d, j, i, k - 1);
      double r = 1.0, v = 1.0;
      if (below && above) {
        const int64_t c = g.c3(j, i, k);
        r = 0.5 * (d.rRho[c - 1] + d.rRho[c]);
        v = 0.5 * (d.rVisc[c - 1] + d.rVisc[c]);
      } else if (below) {
        const int64_t c = g.c3(j, i, k);
        const int c2 = g.c2(j, i);
        r = 0.5 * (d.rRho[c] + (rrTop ? rrTop[c2] : d.rRho[c]));
        v = 0.5 * (d.rVisc[c] + (rvTop ? rvTop[c2] : d.rVisc[c]));
      }
      d.rZp[g.cF(j, i, k)] = r;
      d.vZp[g.cF(j, i, k)] = v;
    }
  }
}

// ---- assembly (Predictor.cpp:400-698) into stencil storage ---------------
// With DENS the legs carry r v of their face, the storage r of the cell and
// the gravity terms r^2 v (Predictor.cpp:474-687); without, every ratio is
// the literal 1.0 and folds away.
template <bool DENS>
__global__ __launch_bounds__(kThreads) void k_assemble(Dev d) {
  const double dtg = *d.dtgp;
  const Grid& g = d.g;
  const int64_t n = static_cast<int64_t>(g.columns()) * g.nz;
  const double dx = g.dx, dy = g.dy, az = g.az;
  for (Range r = xcd_range(n); r.t < r.end; r.t += r.stride) {
    const int k = static_cast<int>(r.t % g.nz);
    const int col = static_cast<int>(r.t / g.nz);
    int j, i;
    tiled_column(g, col, j, i);
    const int c2 = g.c2(j, i);
    const int64_t c = g.c3(j, i, k);
    const int kt = d.kt[c2];
    if (k < kt) {
      d.Ad[c] = 1.0;
      d.Axp[c] = 0.0;
      d.Ayp[c] = 0.0;
      d.Azp[c] = 0.0;
      d.rhs[c] = 0.0;
      d.x[c] = 0.0;
      if (d.Ad32) {
        d.Ad32[c] = 1.0f;
        d.Axp32[c] = 0.0f;
        d.Ayp32[c] = 0.0f;
        d.Azp32[c] = 0.0f;
      }
      continue;
    }
    // Operands, loaded together ahead of the rules that use them (the halo
    // keeps every neighbour address in bounds).  The six face conductivities
    // follow face_lat_at / face_z_at (corr = 0) specialised to an active cell.
    const int64_t sx = g.nz, sy = static_cast<int64_t>(g.nx2) * g.nz;
    const int64_t cxm = c - sx, cxp = c + sx, cym = c - sy, cyp = c + sy;
    const int ktE = d.kt[c2 + 1], ktW = d.kt[c2 - 1];
    const int ktN = d.kt[c2 + g.nx2], ktS = d.kt[c2 - g.nx2];
    const double chv = d.ch[c];  // C(h^n), stored by the predictor's k_cell_k
    const double dzc = d.dz3d[c], dzU = d.dz3d[c + 1], dzD = d.dz3d[c - 1];
    const double Kxc = d.Kx[c], KxE = d.Kx[cxp], KxW = d.Kx[cxm];
    const double Kyc = d.Ky[c], KyN = d.Ky[cyp], KyS = d.Ky[cym];
    const double Kzc = d.Kz[c], KzU = d.Kz[c + 1], KzD = d.Kz[c - 1];
    const double cosxc = d.cosx[c], cosxW = d.cosx[cxm], axc = d.ax[c], axW = d.ax[cxm];
    const double cosyc = d.cosy[c], cosyS = d.cosy[cym], ayc = d.ay[c], ayS = d.ay[cym];
    const int cE = d.sxp[c2], cW = d.sxm[c2], cN = d.syp[c2], cS = d.sym[c2];
    const int bot = d.botCode[c2];
    const double wcnc = d.wcn[c], wcsc = wcs_at(d, c), hnc = d.hn[c];
    double rc = 1.0, rxc = 1.0, rxW = 1.0, vxc = 1.0, vxW = 1.0, ryc = 1.0, ryS = 1.0, vyc = 1.0,
           vyS = 1.0, rzT = 1.0, rzB = 1.0, vzT = 1.0, vzB = 1.0;
    if (DENS) {
      const int64_t f = g.cF(j, i, k);
      rc = d.rRho[c];
      rxc = d.rXp[c], rxW = d.rXp[cxm], vxc = d.vXp[c], vxW = d.vXp[cxm];
      ryc = d.rYp[c], ryS = d.rYp[cym], vyc = d.vYp[c], vyS = d.vYp[cym];
      rzT = d.rZp[f], rzB = d.rZp[f + 1], vzT = d.vZp[f], vzB = d.vZp[f + 1];
    }
    double b3c = 0.0, b3E = 0.0, b3W = 0.0, b3N = 0.0, b3S = 0.0;
    double sinxc = 0.0, sinxW = 0.0, sinyc = 0.0, sinyS = 0.0;
    if (d.terrain) {
      b3c = d.bot3d[c];
      b3E = d.bot3d[cxp];
      b3W = d.bot3d[cxm];
      b3N = d.bot3d[cyp];
      b3S = d.bot3d[cym];
      sinxc = d.sinx[c];
      sinxW = d.sinx[cxm];
      sinyc = d.siny[c];
      sinyS = d.siny[cym];
    }
    const bool east = i == g.nx, west = i == 1, north = j == g.ny, south = j == 1;
    const bool isTop = (k == 0) || (k - 1 < kt);
    const double volume = az * dzc;
    const bool unsatC = !d.full3d && unsat(d, j, i, k);
    double kxp = d.harmonic ? face_k_harmonic(Kxc, d.ksx[c], 1.0, KxE, d.ksx[cxp], 1.0)
                            : 0.5 * (Kxc + KxE);
    if (east) {
      if (cE == FREHG_BC_NOFLUX) kxp = 0.0;
      else if (is_head(cE)) kxp = Kxc;
    }
    if (!d.full3d && (unsatC || unsat(d, j, i + 1, k))) kxp = 0.0;
    if (!east && k < ktE) kxp = 0.0;
    double kxm = d.harmonic ? face_k_harmonic(KxW, d.ksx[cxm], 1.0, Kxc, d.ksx[c], 1.0)
                            : 0.5 * (KxW + Kxc);
    if (west) {
      if (cW == FREHG_BC_NOFLUX) kxm = 0.0;
      else if (is_head(cW)) kxm = Kxc;
    }
    if (!d.full3d && (unsat(d, j, i - 1, k) || unsatC)) kxm = 0.0;
    if (!west && k < ktW) kxm = 0.0;
    double kyp = d.harmonic ? face_k_harmonic(Kyc, d.ksy[c], 1.0, KyN, d.ksy[cyp], 1.0)
                            : 0.5 * (Kyc + KyN);
    if (north) {
      if (cN == FREHG_BC_NOFLUX) kyp = 0.0;
      else if (is_head(cN)) kyp = Kyc;
    }
    if (!d.full3d && (unsatC || unsat(d, j + 1, i, k))) kyp = 0.0;
    if (!north && k < ktN) kyp = 0.0;
    double kym = d.harmonic ? face_k_harmonic(KyS, d.ksy[cym], 1.0, Kyc, d.ksy[c], 1.0)
                            : 0.5 * (KyS + Kyc);
    if (south) {
      if (cS == FREHG_BC_NOFLUX) kym = 0.0;
      else if (is_head(cS)) kym = Kyc;
    }
    if (!d.full3d && (unsat(d, j - 1, i, k) || unsatC)) kym = 0.0;
    if (!south && k < ktS) kym = 0.0;
    double gxp = -kxp * dtg * rxc * vxc * cosxc / (dx * dx) * (axc * dx);
    double gxm = -kxm * dtg * rxW * vxW * cosxW / (dx * dx) * (axW * dx);
    double gyp = -kyp * dtg * ryc * vyc * cosyc / (dy * dy) * (ayc * dy);
    double gym = -kym * dtg * ryS * vyS * cosyS / (dy * dy) * (ayS * dy);
    if (east && is_head(cE)) gxp *= 2.0;
    if (west && is_head(cW)) gxm *= 2.0;
    if (north && is_head(cN)) gyp *= 2.0;
    if (south && is_head(cS)) gym *= 2.0;
    double dzf = (k == g.nz - 1) ? 0.5 * dzc : 0.5 * (dzc + dzU);
    // face above the cell (k) and below it (k + 1)
    double kzTop;
    if (k == 0) kzTop = 0.0;
    else if (isTop) kzTop = Kzc;
    else kzTop = d.harmonic ? face_k_harmonic(KzD, d.ksz[c - 1], dzD, Kzc, d.ksz[c], dzc)
                            : 0.5 * (KzD + Kzc);
    if (isTop) {
      if (d.coupled) {
        const double wet = d.pcLegacy ? 0.5 * (d.ksz[c] + Kzc) : d.ksz[c];
        kzTop = d.dept[c2] > 0.0 ? wet : Kzc;
      } else {
        const int code = d.topCode[c2];
        if (code == FREHG_BC_HEAD) kzTop = d.ksz[c];
        else if (code == FREHG_BC_NOFLUX) kzTop = 0.0;
        else kzTop = Kzc;
      }
      if (k >= 1) kzTop = 0.0;
    }
    double kzBot;
    if (k == g.nz - 1) kzBot = bot == FREHG_BC_NOFLUX ? 0.0 : Kzc;
    else kzBot = d.harmonic ? face_k_harmonic(Kzc, d.ksz[c], dzc, KzU, d.ksz[c + 1], dzU)
                            : 0.5 * (Kzc + KzU);
    const double gzp = -kzBot * dtg * rzB * vzB / (dzc * dzf) * volume;
    dzf = isTop ? 0.5 * dzc : 0.5 * (dzc + dzD);
    const double gzm = -kzTop * dtg * rzT * vzT / (dzc * dzf) * volume;

    const double storage = (chv + d.ss * wcnc / wcsc) * rc;
    double gct = storage * volume - (gxp + gxm + gyp + gym);
    if (east && cE == FREHG_BC_FLUX) gct += gxp;
    if (west && cW == FREHG_BC_FLUX) gct += gxm;
    if (north && cN == FREHG_BC_FLUX) gct += gyp;
    if (south && cS == FREHG_BC_FLUX) gct += gym;
    if (!(k == g.nz - 1 && bot != FREHG_BC_HEAD)) gct -= gzp;
    const int mode = d.coupled ? d.cplMode[c2] : 0;
    if (isTop) {
      bool dir = d.coupled ? (mode == 1) : (d.topCode[c2] == FREHG_BC_HEAD);
      dir = dir || (d.coupled && mode == 3);  // never 3 under legacy
      if (dir) gct -= gzm;
    } else {
      gct -= gzm;
    }

    double b = storage * hnc * volume;
    b -= dtg * volume * kzBot * rzB * rzB * vzB / dzc;
    b += dtg * volume * kzTop * rzT * rzT * vzT / dzc;
    if (d.terrain) {
      // the along-layer gravity term, K r^2 v sin (legacy: K sin)
      const bool plain = d.pcLegacy;
      double sg = (b3E > b3c) ? 1.0 : -1.0;
      b += sg * dtg * axc * kxp * (plain ? 1.0 : rxc * rxc * vxc) * sinxc;
      sg = (b3c > b3W) ? -1.0 : 1.0;
      b += sg * dtg * axW * kxm * (plain ? 1.0 : rxW * rxW * vxW) * sinxW;
      sg = (b3N > b3c) ? 1.0 : -1.0;
      b += sg * dtg * ayc * kyp * (plain ? 1.0 : ryc * ryc * vyc) * sinyc;
      sg = (b3c > b3S) ? -1.0 : 1.0;
      b += sg * dtg * ayS * kym * (plain ? 1.0 : ryS * ryS * vyS) * sinyS;
    }
    if (east) {
      if (cE == FREHG_BC_FLUX) b += rxc * axc * d.vxp[c2] * dtg;
      else b -= gxp * d.hn[cxp];
    }
    if (west) {
      if (cW == FREHG_BC_FLUX) b += rxW * axW * d.vxm[c2] * dtg;
      else b -= gxm * d.hn[cxm];
    }
    if (north) {
      if (cN == FREHG_BC_FLUX) b += ryc * ayc * d.vyp[c2] * dtg;
      else b -= gyp * d.hn[cyp];
    }
    if (south) {
      if (cS == FREHG_BC_FLUX) b += ryS * ayS * d.vym[c2] * dtg;
      else b -= gym * d.hn[cym];
    }
    if (k == g.nz - 1) {
      if (bot == FREHG_BC_FLUX)
        b += dtg * rzB * volume * (d.botValue[c2] + kzBot * rzB * vzB) / dzc;
      else if (bot == FREHG_BC_HEAD) b -= gzp * d.botValue[c2];
    }
    if (isTop) {
      if (d.coupled) {
        if (mode == 1) {
          b -= gzm * d.dept[c2];
        } else if (mode == 3) {
          // Dirichlet ghost head 0: the leg sits in the diagonal only.
        } else if (mode == 2) {
          const double roomTop = fmax(0.0, (0.9999 * wcsc - wcnc) * dzc);
          const double take = fmin(d.avail[c2], roomTop);
          const double qSupply = -take / dtg;
          b += -dtg * rzT * volume * (qSupply + kzTop * rzT * vzT) / dzc;
        } else {
          double qtop = 0.0;
          if (d.etCol && d.etCol[c2]) {
            qtop = d.etQ[c2];
          } else if (d.topCode[c2] == FREHG_BC_FLUX) {
            const double tv = d.topValue[c2];
            if ((tv > 0.0 && wcnc < wcsc) || (tv < 0.0 && wcnc > wcr_at(d, c))) qtop = tv;
          }
          b += -dtg * rzT * volume * (qtop + kzTop * rzT * vzT) / dzc;
        }
      } else if (d.etCol && d.etCol[c2]) {
        b += -dtg * rzT * volume * (d.etQ[c2] + kzTop * rzT * vzT) / dzc;
      } else if (d.topCode[c2] == FREHG_BC_HEAD) {
        b -= gzm * d.topValue[c2];
      } else if (d.topCode[c2] == FREHG_BC_FLUX) {
        b += -dtg * rzT * volume * (d.topValue[c2] + kzTop * rzT * vzT) / dzc;
      }
    }
    const double axp = (!east && k >= ktE) ? gxp : 0.0;
    const double ayp = (!north && k >= ktN) ? gyp : 0.0;
    const double azp = (k + 1 < g.nz) ? gzp : 0.0;
    if (d.sink) {
      const double sk = d.sink[c];
      if (sk != 0.0) b -= dtg * rc * sk * volume;
    }
    d.Ad[c] = gct;
    d.Axp[c] = axp;
    d.Ayp[c] = ayp;
    d.Azp[c] = azp;
    d.rhs[c] = b;
    if (d.Ad32) {
      d.Ad32[c] = static_cast<float>(gct);
      d.Axp32[c] = static_cast<float>(axp);
      d.Ayp32[c] = static_cast<float>(ayp);
      d.Azp32[c] = static_cast<float>(azp);
    }
    d.x[c] = d.guessPrev ? hnc : 0.0;
  }
}

__global__ void k_scatter(Dev d) {
  if (gated(d)) return;
  const Grid& g = d.g;
  const int64_t n = static_cast<int64_t>(g.columns()) * g.nz;
  for (Range r = xcd_range(n); r.t < r.end; r.t += r.stride) {
    const int k = static_cast<int>(r.t % g.nz);
    const int col = static_cast<int>(r.t / g.nz);
    const int j = col / g.nx + 1, i = col % g.nx + 1;
    const int64_t c = g.c3(j, i, k);
    const double xv = d.x[c];  // loaded with kt, not after it
    if (act(d, j, i, k)) d.h[c] = xv;
  }
}

// ---- coupled top bookkeeping (Corrector.cpp:539-627) ----------------------
__global__ void k_coupled_top(Dev d) {
  if (gated(d)) return;
  const double dtg = *d.dtgp;
  const Grid& g = d.g;
  KSum ex, bo, ev;
  for (Range r = xcd_range(g.columns()); r.t < r.end; r.t += r.stride) {
    const int j = static_cast<int>(r.t / g.nx) + 1, i = static_cast<int>(r.t % g.nx) + 1;
    const int c2 = g.c2(j, i);
    const int kt = d.kt[c2];
    if (kt >= g.nz) continue;
    const int64_t c = g.c3(j, i, kt);
    const double q = d.qzF[g.cF(j, i, kt)];
    ex.add(q * dtg);
    if (q < 0.0 && d.dept[c2] <= d.minDepth) {
      const double volume = g.az * d.dz3d[c];
      if ((wcs_at(d, c) - d.wc[c]) < -q * dtg / volume) {
        const double dd = -q * dtg / g.az;
        d.eta[c2] += dd;
        d.dept[c2] += dd;
        d.avail[c2] += dd;
        d.gain[c2] += dd;
        if (d.bounce) d.bounce[c2] += dd;
        bo.add(dd * g.az);
      }
    }
    d.seep[c2] += q * dtg / g.az;
    if (d.etCol && d.etCol[c2]) {  // the channel's exact volume (Corrector.cpp:609-615)
      d.seep[c2] -= d.etQ[c2] * dtg;
      ev.add(d.etQ[c2] * dtg * g.az);
    } else if (d.topCode[c2] == FREHG_BC_FLUX && d.topValue[c2] > 0.0 && d.wc[c] > wcr_at(d, c) &&
               d.dept[c2] <= d.minDepth) {
      d.seep[c2] -= d.topValue[c2] * dtg;
      ev.add(d.topValue[c2] * dtg * g.az);
    }
  }
  double v[3] = {ex.value(), bo.value(), ev.value()};
  const int s[3] = {S_CEX, S_CBO, S_CEV};
  put_parts<3>(d, s, v);
}

// ---- coupling window (src/coupling/Exchange.cpp) ---------------------------
// The gw side of Coupler::step around the groundwater window, on the attached
// 2D coupling arrays.  (The held-seepage salt snapshot of Exchange.cpp:27-39
// and :75-79 belongs to transport and is not built.)

// Coupler::beginWindow (Exchange.cpp:21-52): avail = dept and gain = 0 on the
// interior, the bounce share cleared over the whole plane.
__global__ void k_cpl_begin_window(Dev d) {
  const Grid& g = d.g;
  const int64_t n = static_cast<int64_t>(g.ny2) * g.nx2;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int j = static_cast<int>(t / g.nx2), i = static_cast<int>(t % g.nx2);
    if (j >= 1 && j <= g.ny && i >= 1 && i <= g.nx) {
      d.avail[t] = d.dept[t];
      d.gain[t] = 0.0;
    }
    if (d.bounce) d.bounce[t] = 0.0;
  }
}

// SurfaceSolver::updateDepth's cell-centre depth (src/swe/WetDry.cpp:24-43)
// over the whole plane: dept = eta - bottom, 0 where that is <= min_depth.
__global__ void k_cpl_refresh_depth(Dev d, const double* __restrict__ bottom) {
  const int64_t n = static_cast<int64_t>(d.g.ny2) * d.g.nx2;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double v = d.eta[t] - bottom[t];
    if (v <= d.minDepth) v = 0.0;
    d.dept[t] = v;
  }
}

// Coupler::applySeepage (Exchange.cpp:54-110): the accumulated exchange depth
// a = seep_accum goes onto eta.  Infiltration (a < 0) applies in full down to
// the bottom (the clamp's remainder is discarded and booked); seepage applies
// onto a wet surface or once it exceeds min_depth, else it is held.  qss
// (may be null) = a / dt.  Partials: applied and discarded depths.
__global__ void k_cpl_apply_seepage(Dev d, const double* __restrict__ bottom, double dt,
                                    double* __restrict__ qss) {
  const Grid& g = d.g;
  KSum ap, dc;
  for (Range r = xcd_range(g.columns()); r.t < r.end; r.t += r.stride) {
    const int j = static_cast<int>(r.t / g.nx) + 1, i = static_cast<int>(r.t % g.nx) + 1;
    const int c2 = g.c2(j, i);
    const double a = d.seep[c2];
    if (qss) qss[c2] = a / dt;
    if (a < 0.0) {
      const double before = d.eta[c2];
      double e = before + a;
      if (e < bottom[c2]) e = bottom[c2];
      d.eta[c2] = e;
      const double change = e - before;
      d.gain[c2] += change;
      ap.add(change);
      dc.add(change - a);
      d.seep[c2] = 0.0;
    } else if (d.dept[c2] > d.minDepth || a > d.minDepth) {
      d.eta[c2] += a;
      d.gain[c2] += a;
      ap.add(a);
      d.seep[c2] = 0.0;
    }
  }
  double v[2] = {ap.value(), dc.value()};
  const int s[2] = {S_XAPP, S_XDISC};
  put_parts<2>(d, s, v);
}

// Coupler::reduceAudit (Exchange.cpp:112-128): sums of gain and of the held
// accumulator over the interior; with bottom, also the depth of a surface
// at rest, sum(eta - bottom) (SurfaceSolver::ownedVolume,
// src/swe/SurfaceSolver.cpp:831-844, counted from eta so that films below
// min_depth are included).
__global__ void k_cpl_reduce_audit(Dev d, const double* __restrict__ bottom) {
  const Grid& g = d.g;
  KSum ga, he, sv;
  for (Range r = xcd_range(g.columns()); r.t < r.end; r.t += r.stride) {
    const int j = static_cast<int>(r.t / g.nx) + 1, i = static_cast<int>(r.t % g.nx) + 1;
    const int c2 = g.c2(j, i);
    ga.add(d.gain[c2]);
    he.add(d.seep[c2]);
    if (bottom) sv.add(d.eta[c2] - bottom[c2]);
  }
  double v[3] = {ga.value(), he.value(), sv.value()};
  const int s[3] = {S_XGAIN, S_XHELD, S_XSURF};
  put_parts<3>(d, s, v);
}

// ---- theta update (Corrector.cpp:629-672) ----------------------------------
// With SINK the root-uptake sink leaves after the flux update, as a second
// rounding step (Corrector.cpp:674-705), and its volume S V dtg is summed.
// With DENS the fluxes are weighted by their face's density ratio and the
// storage by the cell's (Corrector.cpp:653-669).
template <bool SINK, bool DENS>
__global__ void k_update_wc(Dev d) {
  if (gated(d)) return;
  const double dtg = *d.dtgp;
  const Grid& g = d.g;
  const int64_t n = static_cast<int64_t>(g.columns()) * g.nz;
  const int64_t sy = static_cast<int64_t>(g.nx2) * g.nz;
  KSum ss, tr;
  for (Range r = xcd_range(n); r.t < r.end; r.t += r.stride) {
    const int k = static_cast<int>(r.t % g.nz);
    const int col = static_cast<int>(r.t / g.nz);
    int j, i;
    tiled_column(g, col, j, i);
    const int64_t c = g.c3(j, i, k);
    const int64_t f = g.cF(j, i, k);
    // the operands, loaded with kt rather than after the activity test (in
    // bounds: halo and face planes), so their latency is not added to kt's
    const double dz = d.dz3d[c], hc = d.h[c], hnc = d.hn[c], wcnc = d.wcn[c];
    const double qxc = d.qx[c], qxW = d.qx[c - g.nz], qyc = d.qy[c], qyS = d.qy[c - sy];
    const double qzB = d.qzF[f + 1], qzT = d.qzF[f];
    if (!act(d, j, i, k)) continue;
    const double volume = g.az * dz;
    const double dh = hc - hnc;
    const double rr = DENS ? d.rRho[c] : 1.0;
    const double coeff = volume * (rr + rr * d.ss * dh / wcs_at(d, c));
    double dqx, dqy, dqz;
    if (DENS) {
      dqx = dtg * (qxc * d.rXp[c] - qxW * d.rXp[c - g.nz]);
      dqy = dtg * (qyc * d.rYp[c] - qyS * d.rYp[c - sy]);
      dqz = dtg * (qzB * d.rZp[f + 1] - qzT * d.rZp[f]);
    } else {
      dqx = dtg * (qxc - qxW);
      dqy = dtg * (qyc - qyS);
      dqz = dtg * (qzB - qzT);
    }
    double upd = (wcnc * rr * volume + dqx + dqy + dqz) / coeff;
    ss.add(upd * volume * rr * d.ss * dh / wcs_at(d, c));
    if (SINK) {
      const double sk = d.sink[c];
      if (sk != 0.0) {
        const double dtheta = dtg * rr * sk * volume / coeff;
        upd -= dtheta;
        ss.add(-(dtheta * volume * rr * d.ss * dh / wcs_at(d, c)));
        tr.add(sk * volume * dtg);
      }
    }
    d.wc[c] = upd;
    d.sat[c] = (upd >= wcs_at(d, c) || hc > aev_at(d, c)) ? 1 : 0;
  }
  if (SINK) {
    double v[2] = {ss.value(), tr.value()};
    const int s[2] = {S_SS, S_TR};
    put_parts<2>(d, s, v);
  } else {
    double v[1] = {ss.value()};
    const int s[1] = {S_SS};
    put_parts<1>(d, s, v);
  }
}

// ---- post-allocation, phase 1 (Reallocate.cpp:181-391), head contact -------
struct Split {
  double up, down;
};
// The operands of a cell's gradient split (the halo and the face planes keep
// every address in bounds); h is the state entering reallocation.  The
// caller loads them with its own operands, so they are in flight together
// instead of one behind each test.
struct SplitOps {
  double h0, hE, hW, hN, hS, hU, hD, kxE, kxW, kyN, kyS, kzUp, kzDn, dz0, dzU, dzD;
  int ktE, ktW, ktN, ktS, kt0;
};
__device__ __forceinline__ SplitOps split_ops(const Dev& d, int j, int i, int k) {
  const Grid& g = d.g;
  const double* hR = d.h;
  const int64_t c = g.c3(j, i, k);
  const int64_t sx = g.nz, sy = static_cast<int64_t>(g.nx2) * g.nz;
  const int64_t f = g.cF(j, i, k);
  SplitOps o;
  o.h0 = hR[c], o.hE = hR[c + sx], o.hW = hR[c - sx], o.hN = hR[c + sy], o.hS = hR[c - sy];
  o.hU = hR[c + 1], o.hD = hR[c - 1];
  o.kxE = d.kx[c], o.kxW = d.kx[c - sx], o.kyN = d.ky[c], o.kyS = d.ky[c - sy];
  o.kzUp = d.kzF[f + 1], o.kzDn = d.kzF[f];
  o.dz0 = d.dz3d[c], o.dzU = d.dz3d[c + 1], o.dzD = d.dz3d[c - 1];
  o.ktE = d.kt[g.c2(j, i + 1)], o.ktW = d.kt[g.c2(j, i - 1)];
  o.ktN = d.kt[g.c2(j + 1, i)], o.ktS = d.kt[g.c2(j - 1, i)];
  o.kt0 = d.kt[g.c2(j, i)];
  return o;
}
__device__ Split gradient_split(const Dev& d, int j, int i, int k, const SplitOps& o) {
  const Grid& g = d.g;
  const int64_t f = g.cF(j, i, k);
  const double h0 = o.h0, hE = o.hE, hW = o.hW, hN = o.hN, hS = o.hS, hU = o.hU, hD = o.hD;
  const double kxE = o.kxE, kxW = o.kxW, kyN = o.kyN, kyS = o.kyS;
  const double kzUp = o.kzUp, kzDn = o.kzDn;
  const double dz0 = o.dz0, dzU = o.dzU, dzD = o.dzD;
  const int ktE = o.ktE, ktW = o.ktW, ktN = o.ktN, ktS = o.ktS, kt0 = o.kt0;
  double d6[6];
  d6[0] = (k >= ktE && kxE > 0.0) ? (hE - h0) / g.dx : 0.0;
  d6[1] = (k >= ktW && kxW > 0.0) ? (h0 - hW) / g.dx : 0.0;
  d6[2] = (k >= ktN && kyN > 0.0) ? (hN - h0) / g.dy : 0.0;
  d6[3] = (k >= ktS && kyS > 0.0) ? (h0 - hS) / g.dy : 0.0;
  if (k + 1 < g.nz && k + 1 >= kt0 && kzUp > 0.0) {
    const double dzp = 0.5 * (dz0 + dzU);
    d6[4] = (hU - h0) / dzp - (d.densGrad ? d.rZp[f + 1] : 1.0);
  } else {
    d6[4] = 0.0;
  }
  if (k - 1 >= 0 && k - 1 >= kt0 && kzDn > 0.0) {
    const double dzm = 0.5 * (dz0 + dzD);
    d6[5] = (h0 - hD) / dzm - (d.densGrad ? d.rZp[f] : 1.0);
  } else {
    d6[5] = 0.0;
  }
  if (!d.full3d) d6[0] = d6[1] = d6[2] = d6[3] = 0.0;
  double tot = 0.0, gz;
  if (d6[0] * d6[1] >= 0.0) tot += fabs(0.5 * (d6[0] + d6[1]));
  else tot += fabs(d6[0]) + fabs(d6[1]);
  if (d6[2] * d6[3] >= 0.0) tot += fabs(0.5 * (d6[2] + d6[3]));
  else tot += fabs(d6[2]) + fabs(d6[3]);
  if (d6[4] * d6[5] >= 0.0) {
    gz = 0.5 * (d6[4] + d6[5]);
    tot += fabs(gz);
  } else {
    gz = 0.0;
    tot += fabs(d6[4]) + fabs(d6[5]);
  }
  Split sp{0.0, 0.0};
  if (tot > 0.0) {
    if (gz > 0.0) sp.up = gz / tot;
    else if (gz < 0.0) sp.down = -gz / tot;
    else {
      sp.down = fabs(d6[4]) / tot;
      sp.up = fabs(d6[5]) / tot;
    }
  }
  return sp;
}

// ---- reallocation, final clamp and step-end inputs: one wave per column ----
// Reallocate.cpp:122-545 (saturated_contact: head), Corrector.cpp:758-792,
// RichardsSolver.cpp:961-1006, AdaptiveStep.cpp:59-98.
//
// Lane l holds rows k = l + 64 m of the wave's column.  Phase 1 (the
// per-cell classification) reads the state as it stood when the phase began
// (h, wc and the neighbours' saturation flags) and writes its result into the
// other buffer pair (hOut, wcOut), so no snapshot copy is taken.  Phase 2
// (the send walks) is sequential within a column: the wave runs it as
// wave-uniform control flow over the senders in k order, keeping the
// column's room, theta and face fluxes in registers; the cells that still
// have room are tracked as ballot masks, so a walk jumps straight to the next
// receiving cell (the reference's loop visits the full cells in between and
// leaves them untouched).  Every arithmetic operation and its order are the
// reference's.

// Value of lane l (wave-uniform) of a double.
__device__ __forceinline__ double lane_read(double x, int l) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b & 0xffffffffll), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}
// Row k (wave-uniform) of a column held as v[m] at row lane + 64 m.
template <int N>
__device__ __forceinline__ double col_get(const double (&v)[N], int k) {
  const int m = k >> 6;
  double s = v[0];
#pragma unroll
  for (int mm = 1; mm < N; ++mm)
    if (m == mm) s = v[mm];
  return lane_read(s, k & 63);
}
template <int N>
__device__ __forceinline__ void col_set(double (&v)[N], int k, double x, int lane) {
  const int m = k >> 6;
#pragma unroll
  for (int mm = 0; mm < N; ++mm)
    if (m == mm && lane == (k & 63)) v[mm] = x;
}
template <int N>
__device__ __forceinline__ void col_add(double (&v)[N], int k, double x, int lane) {
  const int m = k >> 6;
#pragma unroll
  for (int mm = 0; mm < N; ++mm)
    if (m == mm && lane == (k & 63)) v[mm] += x;
}
// Highest set row below k, or -1.
template <int R>
__device__ __forceinline__ int mask_prev(const unsigned long long (&mk)[R], int k) {
  if (k <= 0) return -1;
  const int m = (k - 1) >> 6, l = (k - 1) & 63;
  int res = -1;
#pragma unroll
  for (int mm = 0; mm < R; ++mm) {
    unsigned long long bits = mk[mm];
    if (mm > m) bits = 0ull;
    else if (mm == m && l < 63) bits &= (1ull << (l + 1)) - 1ull;
    if (bits) res = mm * 64 + 63 - __clzll(static_cast<long long>(bits));
  }
  return res;
}
// Lowest set row above k, or -1.
template <int R>
__device__ __forceinline__ int mask_next(const unsigned long long (&mk)[R], int k) {
  const int m = (k + 1) >> 6, l = (k + 1) & 63;
  int res = -1;
#pragma unroll
  for (int mm = R - 1; mm >= 0; --mm) {
    unsigned long long bits = mk[mm];
    if (mm < m) bits = 0ull;
    else if (mm == m) bits &= ~((1ull << l) - 1ull);
    if (bits) res = mm * 64 + __ffsll(static_cast<long long>(bits)) - 1;
  }
  return res;
}
template <int R>
__device__ __forceinline__ void mask_clear(unsigned long long (&mk)[R], int k) {
#pragma unroll
  for (int mm = 0; mm < R; ++mm)
    if ((k >> 6) == mm) mk[mm] &= ~(1ull << (k & 63));
}

// Rows [a, b) of mask word w as bits.
__device__ __forceinline__ unsigned long long row_bits(int w, int a, int b) {
  const int lo = max(a - 64 * w, 0), hi = min(b - 64 * w, 64);
  if (hi <= lo) return 0ull;
  const unsigned long long top = hi == 64 ? ~0ull : ((1ull << hi) - 1ull);
  return top & ~((1ull << lo) - 1ull);
}

// The send walks of one column done by all lanes at once, for the common
// case; returns false, having changed nothing, when the case does not hold.
//
// The serial walk (k_realloc_walk_serial) visits the senders in k order; an
// up-send fills the nearest cells with room above the sender, a down-send
// those below.  When (1) no sender holds room itself, (2) every up-send has a
// cell with room above it, and (3) no receiving cell runs out of room, every
// send lands whole in the nearest cell with room as the column stood before
// the walk.  A receiving cell r then gets, in the serial order, the
// down-sends of the senders between the previous cell with room and r, then
// the up-sends of the senders between r and the next cell with room.  Face r
// takes the down-sends and face r+1 the up-sends, and by (1) no face is
// written from two receiving cells.  Each sender lane forms its sends'
// quotients (volume / receiving cell volume, volume / dt) with the serial
// walk's operands, and each receiving lane adds them in the serial walk's
// order: the result is bit-identical.  Down-sends with no cell with room
// below are dropped (the serial walk's "lost" volume).
// sh holds 7 * 64 R doubles: dz, then (volume, theta increment, flux
// increment) of the up-send and of the down-send of every row.
template <int R>
__device__ __forceinline__ bool walk_parallel(const Grid& g, int lane, int kt, double dtg,
                                              double (&wcv)[R], const double (&dzv)[R],
                                              double (&qz)[R + 1], const double (&room)[R],
                                              const double (&sUp)[R], const double (&sDn)[R],
                                              const unsigned long long (&sendMask)[R],
                                              const unsigned long long (&roomMask)[R],
                                              double* sh, KSum& lost) {
  const int nz = g.nz;
  unsigned long long both = 0ull;
#pragma unroll
  for (int m = 0; m < R; ++m) both |= sendMask[m] & roomMask[m];
  if (both) return false;                                   // (1)
  const int firstRoom = mask_next<R>(roomMask, kt - 1);
  const int topEnd = firstRoom < 0 ? nz : firstRoom;
  unsigned long long topUp = 0ull;
#pragma unroll
  for (int m = 0; m < R; ++m) topUp |= __ballot(sUp[m] > 0.0) & row_bits(m, 0, topEnd);
  if (topUp) return false;                                  // (2)
  double* shDz = sh;
  double* shU = sh + kWave * R;
  double* shD = sh + 4 * kWave * R;
#pragma unroll
  for (int m = 0; m < R; ++m) shDz[lane + kWave * m] = dzv[m];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  bool dropped[R];
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int k = lane + kWave * m;
    double uW = 0.0, uQ = 0.0, dW = 0.0, dQ = 0.0;
    dropped[m] = false;
    if (sUp[m] > 0.0) {  // by (2) a cell with room lies above
      const int t = mask_prev<R>(roomMask, k);
      uW = sUp[m] / (g.az * shDz[t]);
      uQ = sUp[m] / dtg;
    }
    if (sDn[m] > 0.0) {
      const int t = mask_next<R>(roomMask, k);
      if (t >= 0) {
        dW = sDn[m] / (g.az * shDz[t]);
        dQ = sDn[m] / dtg;
      } else {
        dropped[m] = true;
      }
    }
    shU[3 * k] = sUp[m];
    shU[3 * k + 1] = uW;
    shU[3 * k + 2] = uQ;
    shD[3 * k] = sDn[m];
    shD[3 * k + 1] = dW;
    shD[3 * k + 2] = dQ;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // face r+1 of row r = lane + 64 m
  double qNext[R];
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const double down = __shfl_down(qz[m], 1, kWave);
    const double wrap = __shfl(qz[m + 1], 0, kWave);
    qNext[m] = lane == kWave - 1 ? wrap : down;
  }
  bool over = false;
  // The sends of the senders in rows [a, b) into one receiving cell, four
  // table loads in flight at a time, applied in row order.
  auto apply = [&](const double* e, int a, int b, double& w, double& rm, double& q, int& has,
                   bool add) {
#pragma unroll
    for (int wd = 0; wd < R; ++wd) {
      unsigned long long bits = sendMask[wd] & row_bits(wd, a, b);
      while (bits) {
        int ks[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          ks[u] = bits ? wd * 64 + __ffsll(static_cast<long long>(bits)) - 1 : -1;
          bits &= bits - 1ull;
        }
        double v[4][3];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int kk = ks[u] < 0 ? 0 : ks[u];
          v[u][0] = e[3 * kk];
          v[u][1] = e[3 * kk + 1];
          v[u][2] = e[3 * kk + 2];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          if (ks[u] >= 0 && v[u][0] > 0.0) {
            if (!(rm > v[u][0])) over = true;               // (3)
            w = w + v[u][1];
            rm = rm - v[u][0];
            q = add ? q + v[u][2] : q - v[u][2];
            has = 1;
          }
        }
      }
    }
  };
  double nW[R], qDn[R], qUp[R];
  int hasDn[R], hasUp[R];
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int r = lane + kWave * m;
    nW[m] = wcv[m];
    qDn[m] = qz[m];
    qUp[m] = qNext[m];
    hasDn[m] = 0;
    hasUp[m] = 0;
    if (r >= nz || !((roomMask[m] >> lane) & 1ull)) continue;
    double cRoom = room[m];
    const int prev = mask_prev<R>(roomMask, r);
    apply(shD, prev + 1, r, nW[m], cRoom, qDn[m], hasDn[m], false);
    const int next = mask_next<R>(roomMask, r);
    apply(shU, r + 1, next < 0 ? nz : next, nW[m], cRoom, qUp[m], hasUp[m], true);
  }
  if (__ballot(over)) return false;
  // commit
#pragma unroll
  for (int m = 0; m < R; ++m) {
    if (dropped[m]) lost.add(-sDn[m]);
    wcv[m] = nW[m];
    if (hasDn[m]) qz[m] = qDn[m];
  }
#pragma unroll
  for (int m = 0; m <= R; ++m) {
    // face lane + 64 m takes the up-sends of row lane + 64 m - 1
    double v = 0.0;
    int h = 0;
    if (m < R) {
      v = __shfl_up(qUp[m], 1, kWave);
      h = __shfl_up(hasUp[m], 1, kWave);
    }
    if (m > 0) {
      const double vw = __shfl(qUp[m - 1], kWave - 1, kWave);
      const int hw = __shfl(hasUp[m - 1], kWave - 1, kWave);
      if (lane == 0) {
        v = vw;
        h = hw;
      }
    } else if (lane == 0) {
      h = 0;
    }
    if (h) qz[m] = v;
  }
  return true;
}

// Saturation flag of the state entering reallocation (Reallocate.cpp:303-337
// "sat": wc >= theta_s or h > aev), for the halo columns; interior cells get
// theirs from k_update_wc.
__global__ void k_sat_halo(Dev d) {
  if (gated(d)) return;
  const Grid& g = d.g;
  const int nh = 2 * (g.nx2 + g.ny);
  const int64_t n = static_cast<int64_t>(nh) * g.nz;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int k = static_cast<int>(t % g.nz);
    const int q = static_cast<int>(t / g.nz);
    int j, i;
    if (q < g.nx2) { j = 0; i = q; }
    else if (q < 2 * g.nx2) { j = g.ny + 1; i = q - g.nx2; }
    else if (q < 2 * g.nx2 + g.ny) { j = q - 2 * g.nx2 + 1; i = 0; }
    else { j = q - 2 * g.nx2 - g.ny + 1; i = g.nx + 1; }
    const int64_t c = g.c3(j, i, k);
    d.sat[c] = (d.wc[c] >= wcs_at(d, c) || d.h[c] > aev_at(d, c)) ? 1 : 0;
  }
}

// Phase 1, per cell: the classification (Reallocate.cpp:181-391).  Writes
// the cell's outcome into the second buffer pair (wcOut, hOut), its send
// volumes, and flags the columns that have senders.
__global__ __launch_bounds__(kThreads) void k_realloc_classify(Dev d) {
  if (gated(d)) return;
  // The cells whose head is rebuilt from theta (h(theta): two logarithms and
  // two exponentials) go through the wave's ring, their closure operands
  // beside them, and are evaluated 64 at a time.
  __shared__ int64_t ringIdx[kThreads / kWave][kRing];
  __shared__ double ringVal[kThreads / kWave][5][kRing];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  WaveRing wr{ringIdx[wv], 0, 0};
  const Grid& g = d.g;
  const int nz = g.nz;
  const int64_t n = static_cast<int64_t>(g.columns()) * nz;
  const int64_t sy = static_cast<int64_t>(g.nx2) * nz;
  KSum adj, disc;
  auto headOf = [&]() {
    __builtin_amdgcn_wave_barrier();
    const int e = (wr.head + lane) & (kRing - 1);
    const bool has = lane < wr.cnt;
    int64_t c = 0;
    double a = 0.0, nn = 0.0, ts = 0.0, tr = 0.0, w = 0.0;
    if (has) {
      c = ringIdx[wv][e];
      a = ringVal[wv][0][e];
      nn = ringVal[wv][1][e];
      ts = ringVal[wv][2][e];
      tr = ringVal[wv][3][e];
      w = ringVal[wv][4][e];
    }
    __builtin_amdgcn_wave_barrier();
    const int took = wr.cnt < kWave ? wr.cnt : kWave;
    wr.head = (wr.head + took) & (kRing - 1);
    wr.cnt -= took;
    if (has) d.hOut[c] = vg_head(VgSoil{a, nn, ts, tr, 0.0}, w);
  };
  // the walk of xcd_range, wave-uniform: lane l takes t0 + l
  const Range rg = xcd_range(n);
  for (int64_t t0 = rg.t - lane; t0 < rg.end; t0 += rg.stride) {
    const int64_t t = t0 + lane;
    bool defer = false;
    int64_t cDef = 0;
    VgSoil sDef{0.0, 0.0, 0.0, 0.0, 0.0};
    double wDef = 0.0;
    if (t < rg.end) {
      const int k = static_cast<int>(t % nz);
      const int col = static_cast<int>(t / nz);
      const int j = col / g.nx + 1, i = col % g.nx + 1;
      const int c2 = g.c2(j, i);
      const int kt = d.kt[c2];
      const int64_t c = g.c3(j, i, k);
      const double before = d.wc[c];
      const double hc = d.h[c];
      // the operands of an active cell, loaded with kt rather than after it (in
      // bounds: halo), so their latency is not added to kt's
      const VgSoil s = soil_at(d, c);
      const double volume = g.az * d.dz3d[c];
      const bool satU = d.sat[c + 1], satE = d.sat[c + nz], satW = d.sat[c - nz];
      const bool satN = d.sat[c + sy], satS = d.sat[c - sy], satD = d.sat[c - 1];
      const double kxE = d.kx[c], kxW = d.kx[c - nz], kyN = d.ky[c], kyS = d.ky[c - sy];
      // and the gradient split's (a third of the cells split an excess here)
      const SplitOps so = split_ops(d, j, i, k);
      // ghost heads next to edge columns are carried into the new buffer (the
      // ghost theta is rebuilt from the interior afterwards)
      if (i == 1) d.hOut[c - nz] = d.h[c - nz];
      if (i == g.nx) d.hOut[c + nz] = d.h[c + nz];
      if (j == 1) d.hOut[c - sy] = d.h[c - sy];
      if (j == g.ny) d.hOut[c + sy] = d.h[c + sy];
      double wOut = before, hOutV = hc, su = 0.0, sd = 0.0;
      if (k >= kt) {
        if (before >= s.ts) {
          if (!d.reallocOff) {
            const double excess = (before - s.ts) * volume;
            const Split sp = gradient_split(d, j, i, k, so);
            su = excess * sp.up;
            sd = excess * sp.down;
            adj.add(-(excess - su - sd));
            disc.add(excess - su - sd);
            wOut = s.ts;
          }
        } else {
          // the reference's test chain (first true wins) is an OR of its tests
          bool adjSat = (k + 1 < nz && satU) || (kxE != 0.0 && satE) || (kxW != 0.0 && satW) ||
                        (kyN != 0.0 && satN) || (kyS != 0.0 && satS);
          if (!adjSat) {
            if (k == 0 || k - 1 < kt) {
              if (d.coupled) {
                if (d.dept[c2] > 0.0) adjSat = true;
              } else if (d.topCode[c2] == FREHG_BC_HEAD && d.topValue[c2] >= 0.0) {
                adjSat = true;
              }
            } else if (satD) {
              adjSat = true;
            }
          }
          if (!adjSat) {
            if (before < 0.9999 * s.ts) {  // hOut = vg_head(s, before), through the ring
              defer = true;
              cDef = c;
              sDef = s;
              wDef = before;
            }
          } else if (!d.reallocOff) {
            const double after = vg_theta(s, hc);
            wOut = after;
            if (d.redistribute && after < before) {
              const double surplus = (before - after) * volume;
              const Split sp = gradient_split(d, j, i, k, so);
              su = surplus * sp.up;
              sd = surplus * sp.down;
              adj.add(-(surplus - su - sd));
              disc.add(surplus - su - sd);
            } else {
              adj.add((after - before) * volume);
              if (after < before) disc.add((before - after) * volume);
            }
          }
        }
      }
      d.wcOut[c] = wOut;
      if (!defer) d.hOut[c] = hOutV;
      d.sendUp[c] = su;
      d.sendDown[c] = sd;
      if (su > 0.0 || sd > 0.0) d.colFlag[col] = 1;
    }
    {
      const unsigned long long m = __ballot(defer);
      const unsigned lo = static_cast<unsigned>(m), hi = static_cast<unsigned>(m >> 32);
      const int below = __builtin_amdgcn_mbcnt_hi(hi, __builtin_amdgcn_mbcnt_lo(lo, 0));
      if (defer) {
        const int e = (wr.head + wr.cnt + below) & (kRing - 1);
        ringIdx[wv][e] = cDef;
        ringVal[wv][0][e] = sDef.a;
        ringVal[wv][1][e] = sDef.n;
        ringVal[wv][2][e] = sDef.ts;
        ringVal[wv][3][e] = sDef.tr;
        ringVal[wv][4][e] = wDef;
      }
      wr.cnt += __popcll(m);
    }
    if (wr.cnt >= kWave) headOf();
  }
  if (wr.cnt > 0) headOf();
  double v[2] = {adj.value(), disc.value()};
  const int sl[2] = {S_ADJ, S_DISC};
  put_parts<2>(d, sl, v);
}

// Phase 2 (the send walks, Reallocate.cpp:410-540) and the final clamp
// (Corrector.cpp:758-792): one wave per column.  The room of every cell is
// what phase 1 left, (theta_s - theta) * volume, and is rebuilt here from
// wcOut with the same operands.  k_realloc_walk finishes the columns without
// senders and those walk_parallel can do (nz <= 256); the columns it leaves
// flagged are finished by k_realloc_walk_serial.
constexpr int kParallelWalkRows = 4;  // R limit of walk_parallel (LDS)

// The final clamp of a column and its theta out (Corrector.cpp:758-792).
template <int R>
__device__ __forceinline__ void clamp_column(const Dev& d, int lane, int kt, int64_t b0,
                                             double (&wcv)[R], const double (&dzv)[R], KSum& vl) {
  const Grid& g = d.g;
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int k = lane + kWave * m;
    if (k >= g.nz || k < kt) continue;
    const int64_t c = b0 + k;
    double w = wcv[m];
    const double volume = g.az * dzv[m];
    const double ts = wcs_at(d, c), tr = wcr_at(d, c);
    if (w > ts) {
      const double ex = (w - ts) * volume;
      d.vloss[c] += ex;
      vl.add(ex);
      w = ts;
    } else if (w < tr) {
      const double de = (tr - w) * volume;
      d.vloss[c] -= de;
      vl.add(-de);
      w = tr;
    }
    wcv[m] = w;
    d.wcOut[c] = w;
  }
}

// Room, sends and their ballot masks of a walking column.  upIn / dnIn,
// when given, are the column's send volumes already loaded (rows past nz
// hold 0).
template <int R>
__device__ __forceinline__ void walk_inputs(const Dev& d, int lane, int kt, int64_t b0,
                                            const double (&wcv)[R], const double (&dzv)[R],
                                            double (&room)[R], double (&sUp)[R], double (&sDn)[R],
                                            unsigned long long (&sendMask)[R],
                                            unsigned long long (&roomMask)[R],
                                            const double* upIn = nullptr,
                                            const double* dnIn = nullptr) {
  const Grid& g = d.g;
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int k = lane + kWave * m;
    room[m] = 0.0;
    sUp[m] = 0.0;
    sDn[m] = 0.0;
    if (k < g.nz && k >= kt) {
      room[m] = (wcs_at(d, b0 + k) - wcv[m]) * (g.az * dzv[m]);
      sUp[m] = upIn ? upIn[m] : d.sendUp[b0 + k];
      sDn[m] = dnIn ? dnIn[m] : d.sendDown[b0 + k];
    }
    sendMask[m] = __ballot(sUp[m] > 0.0 || sDn[m] > 0.0);
    roomMask[m] = __ballot(room[m] > 0.0);
  }
}

template <int R>
__global__ __launch_bounds__(kThreads) void k_realloc_walk(Dev d) {
  if (gated(d)) return;
  const double dtg = *d.dtgp;
  constexpr int RF = R + 1;  // face rows 0..nz
  constexpr int RS = R <= kParallelWalkRows ? R : 1;
  __shared__ double shWalk[kThreads / kWave][7 * kWave * RS];
  const Grid& g = d.g;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = static_cast<int>(threadIdx.x / kWave);
  const int nz = g.nz;
  KSum lost, vl;
  const int ncol = g.columns();
  const int xcd = blockIdx.x % kXcds, bpx = gridDim.x / kXcds, local = blockIdx.x / kXcds;
  constexpr int wpb = kThreads / kWave;
  const int cBeg = static_cast<int>(static_cast<int64_t>(xcd) * ncol / kXcds);
  const int cEnd = static_cast<int>(static_cast<int64_t>(xcd + 1) * ncol / kXcds);
  // a column's flag is loaded one column ahead, so the walk's operands are
  // loaded together with the column's state instead of after the flag
  const int step = bpx * wpb;
  int col = cBeg + local * wpb + wv;
  int flag = col < cEnd ? d.colFlag[col] : 0;
  for (; col < cEnd; col += step) {
    const int flagNext = col + step < cEnd ? d.colFlag[col + step] : 0;
    const int j = col / g.nx + 1, i = col % g.nx + 1;
    const int kt = d.kt[g.c2(j, i)];
    const int64_t b0 = g.c3(j, i, 0);
    const bool walk = flag != 0 && !d.reallocOff;
    flag = flagNext;
    const int64_t f0 = g.cF(j, i, 0);
    double wcv[R], dzv[R], upv[R], dnv[R], qz[RF];
#pragma unroll
    for (int m = 0; m < R; ++m) {
      const int k = lane + kWave * m;
      wcv[m] = k < nz ? d.wcOut[b0 + k] : 0.0;
      dzv[m] = k < nz ? d.dz3d[b0 + k] : 1.0;
      if constexpr (R <= kParallelWalkRows) {
        upv[m] = walk && k < nz ? d.sendUp[b0 + k] : 0.0;
        dnv[m] = walk && k < nz ? d.sendDown[b0 + k] : 0.0;
      }
    }
    if constexpr (R <= kParallelWalkRows) {
#pragma unroll
      for (int m = 0; m < RF; ++m) {
        const int k = lane + kWave * m;
        qz[m] = walk && k <= nz ? d.qzF[f0 + k] : 0.0;
      }
    }
    if (walk) {
      if constexpr (R > kParallelWalkRows) {
        continue;  // k_realloc_walk_serial
      } else {
        if (d.tw) continue;  // path recording: k_realloc_walk_serial
        double room[R], sUp[R], sDn[R];
        unsigned long long sendMask[R], roomMask[R];
        walk_inputs<R>(d, lane, kt, b0, wcv, dzv, room, sUp, sDn, sendMask, roomMask, upv, dnv);
        if (!walk_parallel<R>(g, lane, kt, dtg, wcv, dzv, qz, room, sUp, sDn, sendMask, roomMask,
                              shWalk[wv], lost))
          continue;  // k_realloc_walk_serial
        if (lane == 0) d.colFlag[col] = 0;
#pragma unroll
        for (int m = 0; m < RF; ++m) {
          const int k = lane + kWave * m;
          if (k <= nz) d.qzF[f0 + k] = qz[m];
        }
      }
    }
    clamp_column<R>(d, lane, kt, b0, wcv, dzv, vl);
  }
  double v[4] = {lost.value(), 0.0, 0.0, vl.value()};
  const int sl[4] = {S_LOST, S_VENT, S_THIN, S_VLOSS};
  put_parts<4>(d, sl, v);
}

// The columns k_realloc_walk left flagged: the serial walk, sender by sender
// in k order, as wave-uniform control flow.  The column's room, theta and
// face fluxes sit in registers; the cells that still have room are ballot
// masks, so a walk jumps straight to the next receiving cell (the
// reference's loop visits the full cells in between and leaves them
// untouched).  Every arithmetic operation and its order are the reference's.
// Same grid as k_realloc_walk: its audit partials are added to that kernel's.
template <int R, bool TW>
__global__ __launch_bounds__(kThreads) void k_realloc_walk_serial(Dev d) {
  if (gated(d)) return;
  const double dtg = *d.dtgp;
  constexpr int RF = R + 1;
  const Grid& g = d.g;
  const int lane = threadIdx.x & (kWave - 1);
  const int nz = g.nz;
  KSum lost, vent, thin, vl;
  const int ncol = g.columns();
  const bool lead = lane == 0;
  const int gw = static_cast<int>(blockIdx.x * (kThreads / kWave) + threadIdx.x / kWave);
  const int nw = static_cast<int>(gridDim.x * (kThreads / kWave));
  for (int base = gw * kWave; base < ncol; base += nw * kWave) {
    const bool flagged = base + lane < ncol && d.colFlag[base + lane] != 0;
    unsigned long long pend = __ballot(flagged);
    while (pend) {
      const int col = base + __ffsll(static_cast<long long>(pend)) - 1;
      pend &= pend - 1ull;
      if (lead) d.colFlag[col] = 0;
      const int j = col / g.nx + 1, i = col % g.nx + 1;
      const int c2 = g.c2(j, i);
      const int kt = d.kt[c2];
      const int64_t b0 = g.c3(j, i, 0);
      const int64_t f0 = g.cF(j, i, 0);
      double wcv[R], dzv[R], qz[RF], room[R], sUp[R], sDn[R];
      unsigned long long sendMask[R], roomMask[R];
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int k = lane + kWave * m;
        wcv[m] = k < nz ? d.wcOut[b0 + k] : 0.0;
        dzv[m] = k < nz ? d.dz3d[b0 + k] : 1.0;
      }
#pragma unroll
      for (int m = 0; m < RF; ++m) {
        const int k = lane + kWave * m;
        qz[m] = k <= nz ? d.qzF[f0 + k] : 0.0;
      }
      walk_inputs<R>(d, lane, kt, b0, wcv, dzv, room, sUp, sDn, sendMask, roomMask);
      const bool topIsHead = d.topCode[c2] == FREHG_BC_HEAD;
      double dept = d.coupled ? d.dept[c2] : 0.0;
      // The receiving cell and the face a walk is working on are cached in
      // wave-uniform registers and written back to the lane that owns them
      // only when the walk moves on: most senders of a column fill the same
      // first cell with room above the saturated zone.
      int curRow = -1, curFace = -1;
      double cRoom = 0.0, cW = 0.0, cDz = 1.0, cQ = 0.0;
      auto flush_cell = [&]() {
        if (curRow >= 0) {
          col_set<R>(room, curRow, cRoom, lane);
          col_set<R>(wcv, curRow, cW, lane);
        }
      };
      auto use_cell = [&](int rr) {
        if (rr != curRow) {
          flush_cell();
          curRow = rr;
          cRoom = col_get<R>(room, rr);
          cW = col_get<R>(wcv, rr);
          cDz = col_get<R>(dzv, rr);
        }
      };
      auto flush_face = [&]() {
        if (curFace >= 0) col_set<RF>(qz, curFace, cQ, lane);
      };
      auto use_face = [&](int f) {
        if (f != curFace) {
          flush_face();
          curFace = f;
          cQ = col_get<RF>(qz, f);
        }
      };
      // Path recording (the transport window, Reallocate.cpp:425-445): a
      // transfer from cell `from` to cell `to` (the cell above the column
      // top for a vent) takes its legacy mirror back off `legacyFace` (pt)
      // and puts its rate on every face it crosses, as a difference array
      // (df) summed along the column at the end.
      double pt[TW ? RF : 1], df[TW ? RF : 1];
      if constexpr (TW) {
#pragma unroll
        for (int m = 0; m < RF; ++m) {
          pt[m] = 0.0;
          df[m] = 0.0;
        }
      }
      auto record = [&](int from, int to, double vol, int legacyFace, double legacySign) {
        if constexpr (TW) {
          const double rate = vol / dtg;
          col_add<RF>(pt, legacyFace, -legacySign * rate, lane);
          if (to < from) {  // faces to+1 .. from, upward
            col_add<RF>(df, to + 1, rate, lane);
            col_add<RF>(df, from + 1, -rate, lane);
          } else {          // faces from+1 .. to, downward
            col_add<RF>(df, from + 1, -rate, lane);
            col_add<RF>(df, to + 1, rate, lane);
          }
        }
      };
      // One receiving cell of a walk: room, theta and the mirror face flux.
      auto take_at = [&](int from, int ll, double& dV, int faceRow, bool faceAdd) {
        use_cell(ll);
        const double take = (cRoom > dV) ? dV : cRoom;
        cW = cW + take / (g.az * cDz);
        const double left = cRoom - take;
        cRoom = left;
        if (!(left > 0.0)) mask_clear<R>(roomMask, ll);
        use_face(faceRow);
        cQ = faceAdd ? cQ + take / dtg : cQ - take / dtg;
        record(from, ll, take, faceRow, faceAdd ? 1.0 : -1.0);
        dV -= take;
      };
#pragma unroll 1
      for (int ms = 0; ms < R; ++ms) {
        unsigned long long pend = 0ull;
#pragma unroll
        for (int mm = 0; mm < R; ++mm)
          if (mm == ms) pend = sendMask[mm];
#pragma unroll 1
        while (pend) {
          const int k = ms * 64 + __ffsll(static_cast<long long>(pend)) - 1;
          pend &= pend - 1ull;
          double dV = col_get<R>(sUp, k);
          if (dV > 0.0) {
            int ll = k;
            while (dV > 0.0) {
              const int nxt = mask_prev<R>(roomMask, ll);
              if (nxt < kt) {
                ll = kt;
                break;
              }
              ll = nxt;
              take_at(k, ll, dV, ll + 1, true);
            }
            if (dV > 0.0) {
              if (d.coupled) {
                const double dd = dV / g.az;
                if (d.ventAll || dept > 0.0 || dd > d.minDepth) {
                  if (!(dept > 0.0 || dd > d.minDepth) && lead) thin.add(dV);
                  if (lead) {
                    d.eta[c2] += dd;
                    d.dept[c2] += dd;
                    d.avail[c2] += dd;
                    d.gain[c2] += dd;
                  }
                  dept += dd;
                  use_face(ll);
                  cQ = cQ + dV / dtg;
                  record(k, ll - 1, dV, ll, 1.0);
                  if (TW && lead) d.winVent[c2] += dV;
                  if (lead) vent.add(dV);
                } else if (lead) {
                  lost.add(-dV);
                }
                dV = 0.0;
              } else if (topIsHead) {
                use_face(ll);
                cQ = cQ + dV / dtg;
                record(k, ll - 1, dV, ll, 1.0);
                dV = 0.0;
              } else {
                while (dV > 0.0) {
                  const int nxt = mask_next<R>(roomMask, ll);
                  if (nxt < 0) break;
                  ll = nxt;
                  take_at(k, ll, dV, ll, false);
                }
                if (lead) lost.add(-dV);
                dV = 0.0;
              }
            }
          }
          dV = col_get<R>(sDn, k);
          if (dV > 0.0) {
            int ll = k;
            while (dV > 0.0) {
              const int nxt = mask_next<R>(roomMask, ll);
              if (nxt < 0) break;
              ll = nxt;
              take_at(k, ll, dV, ll, false);
            }
            if (lead) lost.add(-dV);
          }
        }
      }
      flush_cell();
      flush_face();
#pragma unroll
      for (int m = 0; m < RF; ++m) {
        const int k = lane + kWave * m;
        if (k <= nz) d.qzF[f0 + k] = qz[m];
      }
      if constexpr (TW) {
        double carry = 0.0;
#pragma unroll
        for (int m = 0; m < RF; ++m) {
          const int k = lane + kWave * m;
          const double run = scan_affine_up(df[m], 1.0, carry, lane);
          if (k <= nz) d.pathCorr[f0 + k] = pt[m] + run;
        }
      }
      clamp_column<R>(d, lane, kt, b0, wcv, dzv, vl);
    }
  }
  double v[4] = {lost.value(), vent.value(), thin.value(), vl.value()};
  const int sl[4] = {S_LOST, S_VENT, S_THIN, S_VLOSS};
  add_parts<4>(d, sl, v);
}

// Boundary inflow of the step (RichardsSolver.cpp:961-1006): the net volume
// through the domain's faces, one thread per column for the top and bottom
// faces and one per face for the sides.
__global__ __launch_bounds__(kThreads) void k_boundary_in(Dev d) {
  if (gated(d)) return;
  const double dtg = *d.dtgp;
  const Grid& g = d.g;
  const int nz = g.nz;
  const int64_t sy = static_cast<int64_t>(g.nx2) * nz;
  KSum bin;
  for (Range r = xcd_range(g.columns()); r.t < r.end; r.t += r.stride) {
    const int col = static_cast<int>(r.t);
    const int j = col / g.nx + 1, i = col % g.nx + 1;
    const int kt = d.kt[g.c2(j, i)];
    if (kt >= nz) continue;
    const int64_t f0 = g.cF(j, i, 0);
    bin.add(-d.qzF[f0 + kt] * dtg);
    bin.add(d.qzF[f0 + nz] * dtg);
  }
  // the domain's side faces, one thread per face: west, east, south, north
  // (a corner column takes both of its sides)
  const int64_t nSide = (2 * static_cast<int64_t>(g.ny) + 2 * static_cast<int64_t>(g.nx)) * nz;
  for (Range r = xcd_range(nSide); r.t < r.end; r.t += r.stride) {
    const int k = static_cast<int>(r.t % nz);
    const int q = static_cast<int>(r.t / nz);
    int j, i, side;
    if (q < g.ny) {
      side = 0, j = q + 1, i = 1;
    } else if (q < 2 * g.ny) {
      side = 1, j = q - g.ny + 1, i = g.nx;
    } else if (q < 2 * g.ny + g.nx) {
      side = 2, j = 1, i = q - 2 * g.ny + 1;
    } else {
      side = 3, j = g.ny, i = q - 2 * g.ny - g.nx + 1;
    }
    if (k < d.kt[g.c2(j, i)]) continue;
    const int64_t c = g.c3(j, i, k);
    if (side == 0) bin.add(-d.qx[c - nz] * dtg);
    else if (side == 1) bin.add(d.qx[c] * dtg);
    else if (side == 2) bin.add(-d.qy[c - sy] * dtg);
    else bin.add(d.qy[c] * dtg);
  }
  double v[1] = {bin.value()};
  const int sl[1] = {S_BIN};
  put_parts<1>(d, sl, v);
}

// Adaptive-step inputs (AdaptiveStep.cpp:59-98): the largest flux imbalance
// and the Courant limit of the unsaturated cells, top cells excluded.  Only
// launched when dt can change (dt_min < dt_max).
__global__ __launch_bounds__(kThreads) void k_adapt_inputs(Dev d) {
  if (gated(d)) return;
  const double dtg = *d.dtgp;
  const Grid& g = d.g;
  const int nz = g.nz;
  const int64_t n = static_cast<int64_t>(g.columns()) * nz;
  const int64_t sy = static_cast<int64_t>(g.nx2) * nz;
  double dqMax = 0.0, coMin = 1.0e8;
  for (Range r = xcd_range(n); r.t < r.end; r.t += r.stride) {
    const int k = static_cast<int>(r.t % nz);
    const int col = static_cast<int>(r.t / nz);
    const int j = col / g.nx + 1, i = col % g.nx + 1;
    const int kt = d.kt[g.c2(j, i)];
    if (k <= kt) continue;  // inactive, or the top cell
    const int64_t c = g.c3(j, i, k);
    const int64_t fk = g.cF(j, i, k);
    const double qin = d.qx[c] / d.ax[c] + d.qy[c] / d.ay[c] + d.qzF[fk + 1] / g.az;
    const double qou = d.qx[c - nz] / d.ax[c] + d.qy[c - sy] / d.ay[c] + d.qzF[fk] / g.az;
    const double dq = fabs(qin - qou) * dtg / d.dz3d[c];
    if (dq > dqMax) dqMax = dq;
    if (d.wc[c] < wcs_at(d, c)) {
      const double dk = vg_dkdtheta(soil_at(d, c), d.wc[c], d.ksz[c]);
      const double co = d.courantMax * d.dz3d[c] / dk;
      if (co < coMin) coMin = co;
    }
  }
  double v[2] = {dqMax, coMin};
  const int sl[2] = {S_DQMAX, S_COMIN};
  put_parts<2>(d, sl, v);
}

__global__ void k_volume(Dev d) {
  const Grid& g = d.g;
  const int64_t n = static_cast<int64_t>(g.columns()) * g.nz;
  KSum vol;
  for (Range r = xcd_range(n); r.t < r.end; r.t += r.stride) {
    const int k = static_cast<int>(r.t % g.nz);
    const int col = static_cast<int>(r.t / g.nz);
    const int j = col / g.nx + 1, i = col % g.nx + 1;
    if (!act(d, j, i, k)) continue;
    const int64_t c = g.c3(j, i, k);
    vol.add(d.wc[c] * g.az * d.dz3d[c]);
  }
  double v[1] = {vol.value()};
  const int s[1] = {S_VOL};
  put_parts<1>(d, s, v);
}

// Fixed-order reduction of every slot: one block per slot.
__global__ void k_reduce_slots(const double* __restrict__ parts, int rb, double* __restrict__ out,
                               const int* gate) {
  if (gate && *gate) return;
  const int slot = blockIdx.x;
  const int op = kSlotOp[slot];
  double v[1] = {op == 0 ? 0.0 : (op == 1 ? -INFINITY : INFINITY)};
  for (int b = threadIdx.x; b < rb; b += blockDim.x) {
    const double p = parts[static_cast<int64_t>(slot) * rb + b];
    v[0] = op == 0 ? v[0] + p : (op == 1 ? fmax(v[0], p) : fmin(v[0], p));
  }
  const int ops[1] = {op};
  block_reduce<1>(v, ops);
  if (threadIdx.x == 0) out[slot] = v[0];
}

__global__ void k_reset_parts(double* parts, int rb) {
  const int64_t n = static_cast<int64_t>(S_COUNT) * rb;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int slot = static_cast<int>(t / rb);
    parts[t] = kSlotOp[slot] == 0 ? 0.0 : (kSlotOp[slot] == 1 ? 0.0 : 1.0e8);
  }
}

// ---- initial conditions (RichardsSolver.cpp:381-418) ------------------------
__global__ void k_initial(Dev d, int form, const double* vals, double vconst, double offset) {
  const Grid& g = d.g;
  const int64_t n = static_cast<int64_t>(g.columns()) * g.nz;
  for (Range r = xcd_range(n); r.t < r.end; r.t += r.stride) {
    const int k = static_cast<int>(r.t % g.nz);
    const int col = static_cast<int>(r.t / g.nz);
    const int j = col / g.nx + 1, i = col % g.nx + 1;
    const int64_t c = g.c3(j, i, k);
    const VgSoil s = soil_at(d, c);
    const double zc = d.bot3d[c] + 0.5 * d.dz3d[c];
    double value = vconst;
    if (vals) value = (form == FREHG_IC_WATER_TABLE) ? vals[col] : vals[r.t];
    double hv, w;
    if (form == FREHG_IC_MOISTURE) {
      w = value;
      hv = (w >= s.ts) ? (d.bath[g.c2(j, i)] - zc) : vg_head(s, w);
    } else if (form == FREHG_IC_HEAD) {
      hv = value;
      w = vg_theta(s, hv);
    } else {
      hv = (value + offset) - zc;
      w = vg_theta(s, hv);
    }
    d.h[c] = hv;
    d.wc[c] = w;
  }
}

// ---- the transport window (RichardsSolver.cpp:826-878) ----------------------
// Every face slot, halos included: x, y, the root-uptake volumes (H), the z
// faces with the walks' path corrections (F) and the ET vapour (2D).
__global__ void k_transport_window(Dev d) {
  if (gated(d)) return;
  const double dtg = *d.dtgp;
  const Grid& g = d.g;
  const int64_t nH = g.cellsH(), nF = g.cellsF();
  const int64_t n2 = static_cast<int64_t>(g.ny2) * g.nx2;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < nF;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    if (t < nH) {
      d.winX[t] += d.qx[t] * dtg;
      d.winY[t] += d.qy[t] * dtg;
      if (d.sink) d.winSink[t] += d.sink[t] * g.az * d.dz3d[t] * dtg;
    }
    d.winZ[t] += (d.qzF[t] + d.pathCorr[t]) * dtg;
    if (t < n2 && d.etCol) d.winVap[t] += d.etQ[t] * g.az * dtg;
  }
}

// ---- test exports (include/frehg_gw_testing.h) -----------------------------
// One closure of the device path per point (soil as 5 doubles a point).
__global__ void k_test_vg(int closure, int64_t n, const double* __restrict__ soil,
                          const double* __restrict__ x, const double* __restrict__ ks,
                          double* __restrict__ out) {
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const VgSoil s{soil[5 * t], soil[5 * t + 1], soil[5 * t + 2], soil[5 * t + 3],
                   soil[5 * t + 4]};
    double v = 0.0, cap = 0.0;
    switch (closure) {
      case 0: v = vg_theta(s, x[t]); break;
      case 1: v = vg_head(s, x[t]); break;
      case 2: (void)vg_mualem_capacity(s, x[t], cap); v = cap; break;
      case 3: v = vg_conductivity(s, x[t], ks[t]); break;
      case 4: v = vg_dkdtheta(s, x[t], ks[t]); break;
      default: v = vg_k_from(vg_mualem_capacity(s, x[t], cap), ks[t]); break;
    }
    out[t] = v;
  }
}

// Saturation flags of every halo-layout cell from the installed theta and h
// (what k_update_wc and k_sat_halo leave for the classification); inactive
// interior cells are never saturated neighbours.
__global__ void k_sat_all(Dev d) {
  const Grid& g = d.g;
  const int64_t n = g.cellsH();
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < n;
       c += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int k = static_cast<int>(c % g.nz);
    const int64_t q = c / g.nz;
    const int i = static_cast<int>(q % g.nx2), j = static_cast<int>(q / g.nx2);
    const bool halo = i == 0 || j == 0 || i == g.nx + 1 || j == g.ny + 1;
    d.sat[c] = (halo || act(d, j, i, k)) &&
                       (d.wc[c] >= wcs_at(d, c) || d.h[c] > aev_at(d, c))
                   ? 1
                   : 0;
  }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
// Row segments per column: R = ceil(nz / 64) lanes-rows per wave lane.
template <class F>
void dispatch_col(int nz, F&& f) {
  switch ((nz + kWave - 1) / kWave) {
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
    default: fail(FREHG_EUNSUPPORTED, "nz > 1024 is not supported by the column kernels");
  }
}

template <class T>
T* dev_alloc(int64_t n) {
  void* p = nullptr;
  HIP_CHECK(hipMalloc(&p, static_cast<size_t>(std::max<int64_t>(n, 1)) * sizeof(T)));
  HIP_CHECK(hipMemset(p, 0, static_cast<size_t>(std::max<int64_t>(n, 1)) * sizeof(T)));
  return static_cast<T*>(p);
}

}  // namespace

struct GwSolver {
  Grid g{};
  Dev d{};
  hipStream_t stream = nullptr;
  std::unique_ptr<Pcg> pcg;
  PcgOptions popt;
  std::vector<void*> allocations;
  // geometry / options
  double offset = 0, boxTop = 0;
  double dtMin = 0, dtMax = 0, dqGrow = 0.01, dqShrink = 0.02;
  double dtgNext = 0, dtgCurrent = 0;
  int rb = 0, nb2d = 0;
  int rbCol = 0;  // blocks of the column kernels (<= rb)
  // Grids of the cell-K and face passes: the blocks that can be resident at
  // once (<= rb), so the grid-stride walk sweeps the cells in one round
  // (cell K 256 -> 231 us, faces 1012 -> 999 us at 512 x 512 x 100; the
  // assembly, the classification and the theta update gain nothing).
  int gFaces[2] = {0, 0}, gCellK[4] = {0, 0, 0, 0};
  bool iso = false;  // ksx = ksy = ksz: one Ks and one per-cell K array
  int64_t nActive = 0;
  std::vector<int> codesHost[6];
  // results
  frehg_gw_audit audit{};
  frehg_gw_solve_stats last{};
  double* slots = nullptr;      // device
  double* slotsHost = nullptr;  // pinned
  // timing
  bool timing = false;
  int timingEvery = 1;  // SpMV launches bracketed: one in timingEvery
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evs0 = nullptr, evs1 = nullptr;
  frehg_gw_timing tm{};
  // the step's dt: device word read by the kernels, pinned host word it is
  // copied from
  double* dtgDev = nullptr;
  double* dtgPinned = nullptr;
  // step mode (FREHG_STEP_*): host-polled, device-controlled, or a
  // device-controlled step captured in a graph (one per buffer parity) and
  // replayed; the budget of the device-controlled solve
  int stepMode = FREHG_STEP_HOST;
  PcgBudget budget;
  bool budgetSet = false;    // given by the caller (else sized from the last solve)
  bool budgetReady = false;  // a host step has sized it, or the caller set it
  // most iterations of the first and of later rounds, and most rounds, over
  // the device steps since the budget was last sized (it shrinks to them)
  int recentFirst = 0, recentLater = 0, recentRounds = 0, recentSteps = 0;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  Dev capturedDev[2]{};      // the device view each graph was captured with
  double* hParity0 = nullptr;
  frehg_gw_graph_stats gstats{};

  template <class T>
  T* alloc(int64_t n) {
    T* p = dev_alloc<T>(n);
    allocations.push_back(p);
    return p;
  }

  ~GwSolver() {
    if (stream) (void)hipStreamSynchronize(stream);
    invalidateGraphs();
    if (dtgPinned) (void)hipHostFree(dtgPinned);
    pcg.reset();
    for (void* p : allocations) (void)hipFree(p);
    if (slotsHost) (void)hipHostFree(slotsHost);
    for (hipEvent_t e : {ev0, ev1, evs0, evs1})
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int gridCells() const { return rb; }

  void create(const frehg_gw_desc& ds);
  void buildGeometry(const frehg_gw_desc& ds, std::vector<double>& bath, std::vector<double>& dz,
                     std::vector<double>& bot, std::vector<int>& ktop);
  void ghosts(bool doH, bool doWc) {
    const int64_t n = (static_cast<int64_t>(g.ny) + g.nx) * g.nz;
    const int blocks = static_cast<int>(std::min<int64_t>((n + kThreads - 1) / kThreads, 4096));
    k_ghosts<<<std::max(blocks, 1), kThreads, 0, stream>>>(d, doH ? 1 : 0, doWc ? 1 : 0);
    HIP_CHECK(hipGetLastError());
  }
  void faces(int corr) {
    if (iso && !corr) k_cell_k<true, true><<<gCellK[3], kThreads, 0, stream>>>(d);
    else if (iso) k_cell_k<true, false><<<gCellK[2], kThreads, 0, stream>>>(d);
    else if (!corr) k_cell_k<false, true><<<gCellK[1], kThreads, 0, stream>>>(d);
    else k_cell_k<false, false><<<gCellK[0], kThreads, 0, stream>>>(d);
    HIP_CHECK(hipGetLastError());
    // the predictor's face K are formed inside the assembly
    if (!corr) return;
    if (d.rRho) k_faces<true><<<gFaces[1], kThreads, 0, stream>>>(d, corr);
    else k_faces<false><<<gFaces[0], kThreads, 0, stream>>>(d, corr);
    HIP_CHECK(hipGetLastError());
  }
  void copy(double* dst, const double* src, int64_t n) {
    HIP_CHECK(hipMemcpyAsync(dst, src, static_cast<size_t>(n) * sizeof(double),
                             hipMemcpyDeviceToDevice, stream));
  }
  // The step's sums to the host mirror; reduceSlots() waits for them.
  void enqueueSlots() {
    k_reduce_slots<<<S_COUNT, kThreads, 0, stream>>>(d.parts, rb, slots, d.gate);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(slotsHost, slots, S_COUNT * sizeof(double), hipMemcpyDeviceToHost,
                             stream));
  }
  void reduceSlots() {
    enqueueSlots();
    HIP_CHECK(hipStreamSynchronize(stream));
  }
  // The step's dt into device memory (every kernel reads it there), from a
  // pinned host word: a replayed graph copies the value its step was given.
  void enqueueDtg() {
    HIP_CHECK(hipMemcpyAsync(dtgDev, dtgPinned, sizeof(double), hipMemcpyHostToDevice, stream));
  }
  void setDtg(double dtg) {
    *dtgPinned = dtg;
    enqueueDtg();
  }
  void step(double t, double dtg);
  // Step phases: before the solve (refresh, predictor assembly) and after it
  // (scatter, corrector, post-allocation, boundary inflow, dt inputs).
  void enqueuePre();
  void enqueuePost();
  void hostStep();
  void deviceStep();
  void finishStep(const PcgStats& ps);
  void captureStep(int parity);
  PcgOptions stepOptions() const;
  void flipParity() {
    std::swap(d.h, d.hOut);
    std::swap(d.wc, d.wcOut);
  }
  int parity() const { return d.h == hParity0 ? 0 : 1; }
  void invalidateGraphs() {
    for (int p = 0; p < 2; ++p)
      if (gexec[p]) {
        (void)hipGraphExecDestroy(gexec[p]);
        gexec[p] = nullptr;
      }
  }
  void setStepMode(int mode, int rounds, int iterations);
  void shrinkBudget(const DeviceSummary& sum);
  // Post-allocation and final clamp (Reallocate.cpp:122-545,
  // Corrector.cpp:758-792): the classification writes the second buffer pair,
  // which then becomes the state; the saturation flags must be current.
  void postAllocation() {
    k_realloc_classify<<<rb, kThreads, 0, stream>>>(d);
    HIP_CHECK(hipGetLastError());
    if (d.tw) fill(d.pathCorr, g.cellsF(), 0.0);
    dispatch_col(g.nz, [&](auto RC) {
      constexpr int R = decltype(RC)::value;
      k_realloc_walk<R><<<rbCol, kThreads, 0, stream>>>(d);
      if (d.tw) k_realloc_walk_serial<R, true><<<rbCol, kThreads, 0, stream>>>(d);
      else k_realloc_walk_serial<R, false><<<rbCol, kThreads, 0, stream>>>(d);
    });
    HIP_CHECK(hipGetLastError());
    flipParity();
    ghosts(false, true);
  }
  void fill(double* p, int64_t n, double v) {
    k_fill<<<static_cast<int>(std::min<int64_t>(2048, (n + kThreads - 1) / kThreads)), kThreads, 0,
             stream>>>(p, n, v, d.gate);
    HIP_CHECK(hipGetLastError());
  }
  // RichardsSolver::enableTransportWindow / resetTransportWindow
  // (RichardsSolver.cpp:796-824).
  void enableTransportWindow() {
    if (d.tw) return;
    const int64_t nH = g.cellsH(), nF = g.cellsF();
    const int64_t n2 = static_cast<int64_t>(g.ny2) * g.nx2;
    d.winX = alloc<double>(nH);
    d.winY = alloc<double>(nH);
    d.winSink = alloc<double>(nH);
    d.winZ = alloc<double>(nF);
    d.pathCorr = alloc<double>(nF);
    d.winVent = alloc<double>(n2);
    d.winVap = alloc<double>(n2);
    d.tw = 1;
  }
  void resetTransportWindow() {
    if (!d.tw) return;
    const int64_t nH = g.cellsH(), nF = g.cellsF();
    const int64_t n2 = static_cast<int64_t>(g.ny2) * g.nx2;
    for (double* p : {d.winX, d.winY, d.winSink}) fill(p, nH, 0.0);
    fill(d.winZ, nF, 0.0);
    fill(d.winVent, n2, 0.0);
    fill(d.winVap, n2, 0.0);
    HIP_CHECK(hipStreamSynchronize(stream));
  }
  // The sweep alone on the installed state (frehg_gw_test_post_allocate).
  void testPostAllocate(double dtg) {
    if (!(dtg > 0.0)) fail(FREHG_EINVAL, "dtg must be positive");
    dtgCurrent = dtg;
    setDtg(dtg);
    k_reset_parts<<<64, kThreads, 0, stream>>>(d.parts, rb);
    k_sat_all<<<rb, kThreads, 0, stream>>>(d);
    HIP_CHECK(hipGetLastError());
    postAllocation();
    reduceSlots();
    const double* s = slotsHost;
    audit = frehg_gw_audit{};
    if (!d.reallocOff) {
      audit.realloc_adjust = s[S_ADJ] + s[S_LOST];
      audit.realloc_dropped = s[S_DISC] - s[S_LOST];
      audit.cpl_vent = s[S_VENT];
      audit.cpl_vent_thin = s[S_THIN];
    }
    audit.vloss = s[S_VLOSS];
  }
  double volume() {
    k_reset_parts<<<64, kThreads, 0, stream>>>(d.parts, rb);
    k_volume<<<rb, kThreads, 0, stream>>>(d);
    HIP_CHECK(hipGetLastError());
    reduceSlots();
    return slotsHost[S_VOL];
  }
  // Coupling window (Exchange.cpp), on the attached arrays.
  void needCoupling(const char* who) const {
    if (!d.coupled) fail(FREHG_EINVAL, std::string(who) + ": no coupling attached");
  }
  int planeBlocks() const {
    const int64_t n = static_cast<int64_t>(d.g.ny2) * d.g.nx2;
    return static_cast<int>(std::min<int64_t>(1024, (n + kThreads - 1) / kThreads));
  }
  void beginWindow() {
    needCoupling("frehg_cpl_begin_window");
    k_cpl_begin_window<<<planeBlocks(), kThreads, 0, stream>>>(d);
    HIP_CHECK(hipGetLastError());
  }
  void refreshDepth(const double* bottom) {
    needCoupling("frehg_cpl_refresh_depth");
    if (!bottom) fail(FREHG_EINVAL, "frehg_cpl_refresh_depth: bottom is required");
    k_cpl_refresh_depth<<<planeBlocks(), kThreads, 0, stream>>>(d, bottom);
    HIP_CHECK(hipGetLastError());
  }
  void applySeepage(const double* bottom, double dt, double* qss, double* applied,
                    double* discarded) {
    needCoupling("frehg_cpl_apply_seepage");
    if (!bottom || !(dt > 0.0) || !applied || !discarded)
      fail(FREHG_EINVAL, "frehg_cpl_apply_seepage: bottom, dt > 0 and outputs are required");
    k_reset_parts<<<64, kThreads, 0, stream>>>(d.parts, rb);
    k_cpl_apply_seepage<<<rb, kThreads, 0, stream>>>(d, bottom, dt, qss);
    HIP_CHECK(hipGetLastError());
    reduceSlots();
    const double cell = d.g.dx * d.g.dy;
    *applied = slotsHost[S_XAPP] * cell;
    *discarded = slotsHost[S_XDISC] * cell;
  }
  void reduceAudit(const double* bottom, double* gain, double* held, double* surface) {
    needCoupling("frehg_cpl_reduce_audit");
    k_reset_parts<<<64, kThreads, 0, stream>>>(d.parts, rb);
    k_cpl_reduce_audit<<<rb, kThreads, 0, stream>>>(d, bottom);
    HIP_CHECK(hipGetLastError());
    reduceSlots();
    const double cell = d.g.dx * d.g.dy;
    if (gain) *gain = slotsHost[S_XGAIN] * cell;
    if (held) *held = slotsHost[S_XHELD] * cell;
    if (surface) *surface = slotsHost[S_XSURF] * cell;
  }
  double* fieldPtr(int f, int64_t* n);
  // the root-uptake sink: host_in = S [1/s] per cell (H) applies it from the
  // next step on; null detaches it
  void setSink(const double* host_in) {
    if (!host_in) {
      d.sink = nullptr;
      return;
    }
    if (!sinkBuf) sinkBuf = alloc<double>(g.cellsH());
    HIP_CHECK(hipMemcpyAsync(sinkBuf, host_in, static_cast<size_t>(g.cellsH()) * sizeof(double),
                             hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    d.sink = sinkBuf;
  }
  double* sinkBuf = nullptr;
  // density / viscosity ratios: rr, rv = cell ratios (H, ghosts included);
  // rrTop, rvTop = the ghost above each column's top cell (2D halo layout) or
  // null for the top cell's own; rr = null resets every ratio to 1
  void setDensity(const double* rr, const double* rv, const double* rrTop, const double* rvTop,
                  int gradient) {
    if (!rr) {
      d.rRho = d.rVisc = d.rXp = d.rYp = d.vXp = d.vYp = d.rZp = d.vZp = nullptr;
      d.densGrad = 0;
    } else {
      const int64_t nH = g.cellsH(), nF = g.cellsF(), n2 = static_cast<int64_t>(g.ny2) * g.nx2;
      if (!densBuf[0]) {
        for (int q = 0; q < 6; ++q) densBuf[q] = alloc<double>(nH);
        densBuf[6] = alloc<double>(nF);
        densBuf[7] = alloc<double>(nF);
        densTop[0] = alloc<double>(n2);
        densTop[1] = alloc<double>(n2);
      }
      d.rRho = densBuf[0], d.rVisc = densBuf[1], d.rXp = densBuf[2], d.rYp = densBuf[3];
      d.vXp = densBuf[4], d.vYp = densBuf[5], d.rZp = densBuf[6], d.vZp = densBuf[7];
      const int fb = static_cast<int>(std::min<int64_t>((nF + kThreads - 1) / kThreads, 4096));
      HIP_CHECK(hipMemcpyAsync(d.rRho, rr, nH * sizeof(double), hipMemcpyHostToDevice, stream));
      if (rv)
        HIP_CHECK(hipMemcpyAsync(d.rVisc, rv, nH * sizeof(double), hipMemcpyHostToDevice, stream));
      else
        k_fill<<<fb, kThreads, 0, stream>>>(d.rVisc, nH, 1.0, nullptr);
      for (int q = 2; q < 6; ++q)
        k_fill<<<fb, kThreads, 0, stream>>>(densBuf[q], nH, 1.0, nullptr);
      for (int q = 6; q < 8; ++q)
        k_fill<<<fb, kThreads, 0, stream>>>(densBuf[q], nF, 1.0, nullptr);
      HIP_CHECK(hipGetLastError());
      const double* top[2] = {rrTop, rvTop};
      const double* topDev[2] = {nullptr, nullptr};
      for (int q = 0; q < 2; ++q)
        if (top[q]) {
          HIP_CHECK(hipMemcpyAsync(densTop[q], top[q], n2 * sizeof(double), hipMemcpyHostToDevice,
                                   stream));
          topDev[q] = densTop[q];
        }
      k_density_faces<<<fb, kThreads, 0, stream>>>(d, topDev[0], topDev[1]);
      HIP_CHECK(hipGetLastError());
      d.densGrad = gradient ? 1 : 0;
    }
    ghosts(true, false);  // hydrostatic ghosts carry the boundary-face ratio
    HIP_CHECK(hipStreamSynchronize(stream));
  }
  double* densBuf[8] = {};
  // the evapotranspiration soil channel: columns (2D int mask) and the
  // applied flux qvap [m/s] (2D), both halo layout; null detaches it
  void setEt(const int* columns, const double* qvap) {
    if (!columns != !qvap)
      fail(FREHG_EINVAL, "frehg_gw_set_et: columns and qvap are given together (both NULL "
                         "detaches the channel)");
    if (!columns) {
      d.etCol = nullptr;
      d.etQ = nullptr;
      return;
    }
    const int64_t n2 = static_cast<int64_t>(g.ny2) * g.nx2;
    if (!etColBuf) {
      etColBuf = alloc<int>(n2);
      etQBuf = alloc<double>(n2);
    }
    HIP_CHECK(hipMemcpyAsync(etColBuf, columns, n2 * sizeof(int), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(etQBuf, qvap, n2 * sizeof(double), hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    d.etCol = etColBuf;
    d.etQ = etQBuf;
  }
  int* etColBuf = nullptr;
  double* etQBuf = nullptr;
  double* densTop[2] = {};
};

void GwSolver::buildGeometry(const frehg_gw_desc& ds, std::vector<double>& bath,
                             std::vector<double>& dz, std::vector<double>& bot,
                             std::vector<int>& ktop) {
  // TerrainMetric (src/gw/TerrainMetric.cpp:101-430), single rank.
  const int nx = g.nx, ny = g.ny, nz = g.nz, nx2 = g.nx2, ny2 = g.ny2;
  auto I2 = [&](int j, int i) { return static_cast<size_t>(j) * nx2 + i; };
  auto I3 = [&](int j, int i, int k) { return (static_cast<size_t>(j) * nx2 + i) * nz + k; };
  std::vector<double> dzk(nz);
  const double stretch = ds.dz_stretch > 0 ? ds.dz_stretch : 1.0;
  double totalDepth = 0;
  for (int k = 0; k < nz; ++k) {
    dzk[k] = ds.dz * std::pow(stretch, k);
    totalDepth += dzk[k];
  }
  bath.assign(static_cast<size_t>(ny2) * nx2, 0.0);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      bath[I2(j, i)] = ds.bottom_elevation ? ds.bottom_elevation[(j - 1) * nx + (i - 1)]
                                           : ds.bottom_elevation_const;
  double lo = 0.0, hi = 0.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      lo = std::min(lo, bath[I2(j, i)]);
      hi = std::max(hi, bath[I2(j, i)]);
    }
  offset = lo < 0.0 ? -lo : 0.0;
  boxTop = hi + offset;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) bath[I2(j, i)] += offset;
  for (int j = 1; j <= ny; ++j) {
    bath[I2(j, 0)] = bath[I2(j, 1)];
    bath[I2(j, nx + 1)] = bath[I2(j, nx)];
  }
  for (int i = 0; i <= nx + 1; ++i) {
    bath[I2(0, i)] = bath[I2(1, i)];
    bath[I2(ny + 1, i)] = bath[I2(ny, i)];
  }
  const size_t nH = static_cast<size_t>(ny2) * nx2 * nz;
  dz.assign(nH, 0.0);
  bot.assign(nH, 0.0);
  ktop.assign(static_cast<size_t>(ny2) * nx2, nz);
  if (!ds.follow_terrain) {
    std::vector<double> bot1d(nz);
    double level = boxTop;
    for (int k = 0; k < nz; ++k) {
      level -= dzk[k];
      bot1d[k] = level;
    }
    std::vector<int> activeK(nz);
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double b = bath[I2(j, i)];
        for (int k = 0; k < nz; ++k) {
          bot[I3(j, i, k)] = bot1d[k];
          dz[I3(j, i, k)] = dzk[k];
        }
        for (int k = 0; k < nz; ++k) {
          double& bk = bot[I3(j, i, k)];
          if (bk >= b - 1.0e-5) {
            activeK[k] = 0;
            if (bk - b < dzk[k]) bk = b;
          } else {
            activeK[k] = 1;
            if (b - bk <= dzk[k]) {
              if (b - bk >= 0.25 * dzk[k]) {
                dz[I3(j, i, k)] = b - bk;
              } else if (k + 1 < nz) {
                activeK[k] = 0;
                dz[I3(j, i, k + 1)] += b - bk;
              }
            }
          }
        }
        int top = nz;
        for (int k = 0; k < nz; ++k)
          if (activeK[k]) {
            top = k;
            break;
          }
        ktop[I2(j, i)] = top;
      }
  } else {
    const bool uniform = ds.terrain_layers == 1;
    const double boxBottom = boxTop - totalDepth;
    double increSum = 0.0, wgt = 1.0;
    for (int k = 0; k < nz; ++k) {
      increSum += wgt;
      wgt *= stretch;
    }
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double b = bath[I2(j, i)];
        if (b <= boxBottom && !uniform)
          fail(FREHG_EINVAL, "follow_terrain: a bed sits at or below the subsurface box bottom");
        double dzv = uniform ? dzk[0] : (b - boxBottom) / increSum;
        double level = b;
        for (int k = 0; k < nz; ++k) {
          if (uniform) dzv = dzk[k];
          level -= dzv;
          dz[I3(j, i, k)] = dzv;
          bot[I3(j, i, k)] = level;
          if (!uniform) dzv *= stretch;
        }
        ktop[I2(j, i)] = 0;
      }
  }
  for (std::vector<double>* f : {&dz, &bot})
    for (int k = 0; k < nz; ++k) {
      for (int j = 1; j <= ny; ++j) {
        (*f)[I3(j, 0, k)] = (*f)[I3(j, 1, k)];
        (*f)[I3(j, nx + 1, k)] = (*f)[I3(j, nx, k)];
      }
      for (int i = 0; i <= nx + 1; ++i) {
        (*f)[I3(0, i, k)] = (*f)[I3(1, i, k)];
        (*f)[I3(ny + 1, i, k)] = (*f)[I3(ny, i, k)];
      }
    }
}

void GwSolver::create(const frehg_gw_desc& ds) {
  if (ds.nx < 1 || ds.ny < 1 || ds.nz < 1) fail(FREHG_EINVAL, "extents must be positive");
  if (!(ds.dx > 0) || !(ds.dy > 0) || !(ds.dz > 0)) fail(FREHG_EINVAL, "spacings must be positive");
  if (ds.nz > 1024) fail(FREHG_EUNSUPPORTED, "nz > 1024 is not supported");
  if (ds.predictor_corrector != 0 && ds.predictor_corrector != 1)
    fail(FREHG_EINVAL, "predictor_corrector: 1 consistent or 0 legacy");
  if (ds.saturated_contact != 1)
    fail(FREHG_EUNSUPPORTED, "saturated_contact: legacy is order-dependent; head only");
  if (ds.preconditioner == FREHG_PC_ICC0)
    fail(FREHG_EUNSUPPORTED, "icc0 is a CPU-only preconditioner; use zline or jacobi");
  if (ds.preconditioner < FREHG_PC_JACOBI || ds.preconditioner > FREHG_PC_CHEBYSHEV)
    fail(FREHG_EINVAL, "unknown preconditioner");
  if (ds.preconditioner == FREHG_PC_CHEBYSHEV &&
      (ds.precision != FREHG_PREC_MIXED || ds.nz % 2 != 0))
    fail(FREHG_EINVAL, "the Chebyshev preconditioner is the mixed form's (FP32) and needs an even nz");
  if (!ds.ksx || !ds.ksy || !ds.ksz || !ds.vg_alpha || !ds.vg_n || !ds.theta_s || !ds.theta_r ||
      !ds.aev)
    fail(FREHG_EINVAL, "all eight per-cell soil arrays are required");
  if (ds.ic_form < 0 || ds.ic_form > 2) fail(FREHG_EINVAL, "unknown initial-condition form");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    fail(FREHG_EDEVICE, "no HIP device is visible");
  HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  g.nx = ds.nx;
  g.ny = ds.ny;
  g.nz = ds.nz;
  g.nx2 = ds.nx + 2;
  g.ny2 = ds.ny + 2;
  g.dx = ds.dx;
  g.dy = ds.dy;
  g.az = ds.dx * ds.dy;
  d.g = g;
  int cus = 256;
  hipDeviceProp_t prop;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    cus = prop.multiProcessorCount;
  const int64_t ncell = static_cast<int64_t>(g.columns()) * g.nz;
  rb = static_cast<int>(std::min<int64_t>(int64_t(cus) * 8, std::max<int64_t>(
                                                                 8, (ncell + kThreads - 1) / kThreads)));
  rb = (rb + 7) / 8 * 8;
  d.rb = rb;
  rbCol = static_cast<int>(std::min<int64_t>(
      rb, std::max<int64_t>(8, ((g.columns() + kThreads / kWave - 1) / (kThreads / kWave) + 7) /
                                   8 * 8)));
  {
    auto resident = [&](const void* f) {
      int per = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, kThreads, 0) != hipSuccess ||
          per <= 0)
        return rb;
      const int64_t b = (int64_t(cus) * per + 7) / 8 * 8;
      return static_cast<int>(std::min<int64_t>(rb, std::max<int64_t>(8, b)));
    };
    gFaces[0] = resident(reinterpret_cast<const void*>(&k_faces<false>));
    gFaces[1] = resident(reinterpret_cast<const void*>(&k_faces<true>));
    gCellK[0] = resident(reinterpret_cast<const void*>(&k_cell_k<false, false>));
    gCellK[1] = resident(reinterpret_cast<const void*>(&k_cell_k<false, true>));
    gCellK[2] = resident(reinterpret_cast<const void*>(&k_cell_k<true, false>));
    gCellK[3] = resident(reinterpret_cast<const void*>(&k_cell_k<true, true>));
  }

  std::vector<double> bath, dz, bot;
  std::vector<int> ktop;
  buildGeometry(ds, bath, dz, bot, ktop);
  const int nx = g.nx, ny = g.ny, nz = g.nz, nx2 = g.nx2, ny2 = g.ny2;
  auto I3 = [&](int j, int i, int k) { return (static_cast<size_t>(j) * nx2 + i) * nz + k; };
  const int64_t nH = g.cellsH(), nF = g.cellsF(), n2 = static_cast<int64_t>(ny2) * nx2;
  nActive = 0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) nActive += nz - ktop[static_cast<size_t>(j) * nx2 + i];

  // metrics (TerrainMetric.cpp:318-408)
  std::vector<double> axv(nH), ayv(nH), sxv(nH, 0.0), cxv(nH, 1.0), syv(nH, 0.0), cyv(nH, 1.0);
  if (ds.follow_terrain) {
    for (int j = 0; j <= ny + 1; ++j)
      for (int i = 1; i < nx; ++i)
        for (int k = 0; k < nz; ++k) {
          const double zc = bot[I3(j, i, k)] + 0.5 * dz[I3(j, i, k)];
          const double zp = bot[I3(j, i + 1, k)] + 0.5 * dz[I3(j, i + 1, k)];
          const double hd = std::fabs(zp - zc);
          const double dist = std::sqrt(hd * hd + g.dx * g.dx);
          sxv[I3(j, i, k)] = hd / dist;
          cxv[I3(j, i, k)] = g.dx / dist;
        }
    for (int j = 1; j < ny; ++j)
      for (int i = 0; i <= nx + 1; ++i)
        for (int k = 0; k < nz; ++k) {
          const double zc = bot[I3(j, i, k)] + 0.5 * dz[I3(j, i, k)];
          const double zp = bot[I3(j + 1, i, k)] + 0.5 * dz[I3(j + 1, i, k)];
          const double hd = std::fabs(zp - zc);
          const double dist = std::sqrt(hd * hd + g.dy * g.dy);
          syv[I3(j, i, k)] = hd / dist;
          cyv[I3(j, i, k)] = g.dy / dist;
        }
  }
  for (int j = 0; j <= ny + 1; ++j)
    for (int i = 0; i <= nx + 1; ++i)
      for (int k = 0; k < nz; ++k) {
        const size_t c = I3(j, i, k);
        if (ds.follow_terrain) {
          const double dzE = (i <= nx) ? dz[I3(j, i + 1, k)] : dz[c];
          const double dzN = (j <= ny) ? dz[I3(j + 1, i, k)] : dz[c];
          axv[c] = 0.5 * (dz[c] + dzE) * cxv[c] * g.dy;
          ayv[c] = 0.5 * (dz[c] + dzN) * cyv[c] * g.dx;
        } else {
          axv[c] = g.dy * dz[c];
          ayv[c] = g.dx * dz[c];
        }
      }

  auto up = [&](double* dst, const void* src, int64_t n, size_t sz) {
    HIP_CHECK(hipMemcpy(dst, src, static_cast<size_t>(n) * sz, hipMemcpyHostToDevice));
  };
  auto field = [&]() { return alloc<double>(nH); };
  d.dz3d = field();
  d.bot3d = field();
  d.ax = field();
  d.ay = field();
  d.sinx = field();
  d.cosx = field();
  d.siny = field();
  d.cosy = field();
  up(d.dz3d, dz.data(), nH, 8);
  up(d.bot3d, bot.data(), nH, 8);
  up(d.ax, axv.data(), nH, 8);
  up(d.ay, ayv.data(), nH, 8);
  up(d.sinx, sxv.data(), nH, 8);
  up(d.cosx, cxv.data(), nH, 8);
  up(d.siny, syv.data(), nH, 8);
  up(d.cosy, cyv.data(), nH, 8);
  d.bath = alloc<double>(n2);
  up(d.bath, bath.data(), n2, 8);
  d.kt = alloc<int>(n2);
  HIP_CHECK(hipMemcpy(d.kt, ktop.data(), n2 * sizeof(int), hipMemcpyHostToDevice));
  axv.clear();
  ayv.clear();

  // soil staging with mirrored edge ghosts (RichardsSolver.cpp:207-329, :24-57)
  {
    std::vector<double> tmp(nH);
    const double* src[8] = {ds.ksx, ds.ksy, ds.ksz, ds.vg_alpha, ds.vg_n, ds.theta_s, ds.theta_r,
                            ds.aev};
    double** dst[8] = {&d.ksx, &d.ksy, &d.ksz, &d.vga, &d.vgn, &d.wcs, &d.wcr, &d.aev};
    // Isotropic soil (ksx = ksy = ksz everywhere) keeps one Ks array and one
    // per-cell K: the three closures would be bit-identical.
    const size_t nIn = static_cast<size_t>(nx) * ny * nz;
    iso = (ds.ksx == ds.ksy || std::memcmp(ds.ksx, ds.ksy, nIn * sizeof(double)) == 0) &&
          (ds.ksx == ds.ksz || std::memcmp(ds.ksx, ds.ksz, nIn * sizeof(double)) == 0);
    for (int a = 0; a < 8; ++a) {
      if (iso && (a == 1 || a == 2)) {
        *dst[a] = d.ksx;
        continue;
      }
      std::fill(tmp.begin(), tmp.end(), 0.0);
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
          std::memcpy(&tmp[I3(j, i, 0)], src[a] + (static_cast<size_t>(j - 1) * nx + (i - 1)) * nz,
                      nz * sizeof(double));
      for (int j = 0; j <= ny + 1; ++j) {
        std::memcpy(&tmp[I3(j, 0, 0)], &tmp[I3(j, 1, 0)], nz * sizeof(double));
        std::memcpy(&tmp[I3(j, nx + 1, 0)], &tmp[I3(j, nx, 0)], nz * sizeof(double));
      }
      for (int i = 0; i <= nx + 1; ++i) {
        std::memcpy(&tmp[I3(0, i, 0)], &tmp[I3(1, i, 0)], nz * sizeof(double));
        std::memcpy(&tmp[I3(ny + 1, i, 0)], &tmp[I3(ny, i, 0)], nz * sizeof(double));
      }
      if (a >= 3) {  // van Genuchten fields: one value when uniform
        bool uniform = true;
        for (int64_t q = 1; q < nH && uniform; ++q) uniform = tmp[q] == tmp[0];
        if (uniform) {
          double* cst[5] = {&d.vgaC, &d.vgnC, &d.wcsC, &d.wcrC, &d.aevC};
          *cst[a - 3] = tmp[0];
          *dst[a] = nullptr;
          continue;
        }
      }
      *dst[a] = field();
      up(*dst[a], tmp.data(), nH, 8);
    }
  }
  // state and work fields
  d.h = field();
  d.hn = field();
  d.wc = field();
  d.wcn = field();
  d.ch = field();
  d.Kx = field();
  d.Ky = iso ? d.Kx : field();
  d.Kz = iso ? d.Kx : field();
  d.kx = field();
  d.ky = field();
  d.qx = field();
  d.qy = field();
  d.kzF = alloc<double>(nF);
  d.qzF = alloc<double>(nF);
  d.vloss = field();
  d.hOut = field();
  d.wcOut = field();
  d.sendUp = field();
  d.sendDown = field();
  d.colFlag = alloc<int>(g.columns());
  d.sat = alloc<unsigned char>(nH);
  d.Ad = field();
  d.Axp = field();
  d.Ayp = field();
  d.Azp = field();
  d.rhs = field();
  d.x = field();
  d.parts = alloc<double>(static_cast<int64_t>(S_COUNT) * rb);
  slots = alloc<double>(S_COUNT);
  HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&slotsHost), S_COUNT * sizeof(double),
                          hipHostMallocDefault));
  dtgDev = alloc<double>(1);
  HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&dtgPinned), sizeof(double),
                          hipHostMallocDefault));
  *dtgPinned = 0.0;
  d.dtgp = dtgDev;
  d.gate = nullptr;
  hParity0 = d.h;

  // boundary markers (RichardsSolver.cpp:429-566)
  {
    const int* cs[6] = {ds.top_code, ds.bot_code, ds.xm_code, ds.xp_code, ds.ym_code, ds.yp_code};
    const double* vs[6] = {ds.top_value, ds.bot_value, ds.xm_value, ds.xp_value, ds.ym_value,
                           ds.yp_value};
    int** cd[6] = {&d.topCode, &d.botCode, &d.sxm, &d.sxp, &d.sym, &d.syp};
    double** vd[6] = {&d.topValue, &d.botValue, &d.vxm, &d.vxp, &d.vym, &d.vyp};
    std::vector<double> vv(n2);
    for (int a = 0; a < 6; ++a) {
      std::vector<int>& cc = codesHost[a];
      cc.assign(n2, FREHG_BC_NOFLUX);
      std::fill(vv.begin(), vv.end(), 0.0);
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          const bool member = a < 2 || (a == 2 && i == 1) || (a == 3 && i == nx) ||
                              (a == 4 && j == 1) || (a == 5 && j == ny);
          if (!member) continue;
          const size_t c = static_cast<size_t>(j - 1) * nx + (i - 1);
          const int code = cs[a] ? cs[a][c] : FREHG_BC_NOFLUX;
          if (code < 0 || code > 4) fail(FREHG_EINVAL, "boundary code out of range");
          if (a < 2 && code == FREHG_BC_HEAD_HYDROSTATIC)
            fail(FREHG_EINVAL, "hydrostatic head is a side condition only");
          if (a != 1 && code == FREHG_BC_GRAVITY)
            fail(FREHG_EINVAL, "gravity drainage is a bottom condition only");
          double v = vs[a] ? vs[a][c] : 0.0;
          if (code == FREHG_BC_HEAD_HYDROSTATIC) v += offset;
          cc[static_cast<size_t>(j) * nx2 + i] = code;
          vv[static_cast<size_t>(j) * nx2 + i] = v;
        }
      *cd[a] = alloc<int>(n2);
      *vd[a] = alloc<double>(n2);
      HIP_CHECK(hipMemcpy(*cd[a], cc.data(), n2 * sizeof(int), hipMemcpyHostToDevice));
      up(*vd[a], vv.data(), n2, 8);
    }
    d.cplMode = alloc<int>(n2);
  }

  // options
  d.ss = ds.specific_storage;
  d.full3d = ds.use_full3d ? 1 : 0;
  d.pcLegacy = ds.predictor_corrector == 0 ? 1 : 0;
  // the legacy pair has no harmonic form (Predictor.cpp:94-102, Corrector.cpp:124-132)
  d.harmonic = ds.face_conductivity == 1 && !d.pcLegacy ? 1 : 0;
  d.reallocOff = ds.reallocation == 1 ? 1 : 0;
  d.redistribute = ds.reallocation_surplus == 1 ? 1 : 0;
  d.terrain = ds.follow_terrain ? 1 : 0;
  d.guessPrev = ds.initial_guess_previous ? 1 : 0;
  d.courantMax = ds.courant_max;
  dtMin = ds.dt_min;
  dtMax = ds.dt_max;
  dqGrow = ds.dq_grow;
  dqShrink = ds.dq_shrink;
  dtgNext = ds.dt_init;
  popt.rtol = ds.rtol;
  popt.atol = ds.atol;
  popt.max_iterations = ds.max_iterations > 0 ? ds.max_iterations : 1000;
  popt.zline = ds.preconditioner == FREHG_PC_ZLINE;
  popt.mixed = ds.precision == FREHG_PREC_MIXED;
  popt.cheb = ds.preconditioner == FREHG_PC_CHEBYSHEV;
  popt.inner_rtol = ds.inner_rtol > 0 ? ds.inner_rtol : 1e-4;

  // initial conditions (RichardsSolver.cpp:331-427)
  double* icDev = nullptr;
  if (ds.ic_values) {
    const int64_t nv = ds.ic_form == FREHG_IC_WATER_TABLE ? g.columns()
                                                         : static_cast<int64_t>(g.columns()) * nz;
    icDev = dev_alloc<double>(nv);
    up(icDev, ds.ic_values, nv, 8);
  }
  k_initial<<<rb, kThreads, 0, stream>>>(d, ds.ic_form, icDev, ds.ic_const, offset);
  HIP_CHECK(hipGetLastError());
  ghosts(true, true);
  copy(d.hn, d.h, nH);
  copy(d.wcn, d.wc, nH);
  HIP_CHECK(hipStreamSynchronize(stream));
  if (icDev) (void)hipFree(icDev);
  pcg = std::make_unique<Pcg>(g, stream);
  // No lateral conductivity anywhere: every x and y leg of the operator is
  // zero and the columns are independent tridiagonal systems, which the
  // z-line form solves directly (one launch over all columns).
  {
    const size_t nIn = static_cast<size_t>(nx) * ny * nz;
    bool lateral = false;
    for (size_t q = 0; q < nIn && !lateral; ++q) lateral = ds.ksx[q] != 0.0 || ds.ksy[q] != 0.0;
    if (!lateral && popt.zline) pcg->setColumnSolve(d.kt, ktop);
  }
  if (popt.mixed && !pcg->columnSolve()) {  // the assembly writes the FP32 stencil beside the FP64 one
    const Stencil<float> a32 = pcg->stencil32();
    d.Ad32 = a32.Ad;
    d.Axp32 = a32.Axp;
    d.Ayp32 = a32.Ayp;
    d.Azp32 = a32.Azp;
  }
  HIP_CHECK(hipEventCreate(&ev0));
  HIP_CHECK(hipEventCreate(&ev1));
  HIP_CHECK(hipEventCreate(&evs0));
  HIP_CHECK(hipEventCreate(&evs1));
  tm.spmv_bytes = pcg->spmvBytesPerLaunch(popt.mixed);
}

PcgOptions GwSolver::stepOptions() const {
  PcgOptions o = popt;
  o.timing = timing;
  o.timing_every = timingEvery;
  o.stencil32_ready = d.Ad32 != nullptr;
  return o;
}

void GwSolver::enqueuePre() {
  k_reset_parts<<<64, kThreads, 0, stream>>>(d.parts, rb);
  HIP_CHECK(hipGetLastError());
  // step-start refresh (RichardsSolver.cpp:743-748)
  ghosts(true, true);
  // predictor (its cell-K pass also copies h, theta into hn, wcn)
  if (d.coupled) {
    k_classify_coupled<<<std::max(1, std::min(4096, (g.columns() + kThreads - 1) / kThreads)),
                         kThreads, 0, stream>>>(d);
    HIP_CHECK(hipGetLastError());
  }
  faces(0);
  if (d.rRho) k_assemble<true><<<rb, kThreads, 0, stream>>>(d);
  else k_assemble<false><<<rb, kThreads, 0, stream>>>(d);
  HIP_CHECK(hipGetLastError());
}

void GwSolver::enqueuePost() {
  k_scatter<<<rb, kThreads, 0, stream>>>(d);
  HIP_CHECK(hipGetLastError());
  ghosts(true, false);
  // corrector
  faces(1);
  if (d.coupled) {
    k_coupled_top<<<rb, kThreads, 0, stream>>>(d);
    HIP_CHECK(hipGetLastError());
  }
  if (d.sink && d.rRho) k_update_wc<true, true><<<rb, kThreads, 0, stream>>>(d);
  else if (d.sink) k_update_wc<true, false><<<rb, kThreads, 0, stream>>>(d);
  else if (d.rRho) k_update_wc<false, true><<<rb, kThreads, 0, stream>>>(d);
  else k_update_wc<false, false><<<rb, kThreads, 0, stream>>>(d);
  HIP_CHECK(hipGetLastError());
  // post-allocation, final clamp, boundary inflow and dt inputs: one column
  // kernel writing the second buffer pair, which then becomes the state
  const int nh = 2 * (g.nx2 + g.ny);
  const int64_t nhc = static_cast<int64_t>(nh) * g.nz;
  k_sat_halo<<<static_cast<int>(std::min<int64_t>((nhc + kThreads - 1) / kThreads, 1024)),
               kThreads, 0, stream>>>(d);
  HIP_CHECK(hipGetLastError());
  postAllocation();
  k_boundary_in<<<std::min(rb, rbCol * 4), kThreads, 0, stream>>>(d);
  HIP_CHECK(hipGetLastError());
  if (d.tw) {
    k_transport_window<<<rb, kThreads, 0, stream>>>(d);
    HIP_CHECK(hipGetLastError());
  }
  if (dtMin < dtMax) {
    k_adapt_inputs<<<rb, kThreads, 0, stream>>>(d);
    HIP_CHECK(hipGetLastError());
  }
}

void GwSolver::step(double /*t*/, double dtg) {
  if (!(dtg > 0.0)) fail(FREHG_EINVAL, "dtg must be positive");
  dtgCurrent = dtg;
  audit = frehg_gw_audit{};
  *dtgPinned = dtg;
  if (stepMode == FREHG_STEP_HOST || timing || !budgetReady) hostStep();
  else deviceStep();
}

// Host-polled step: the solve's decisions are taken on the host between
// launches.
void GwSolver::hostStep() {
  if (timing) HIP_CHECK(hipEventRecord(ev0, stream));
  enqueueDtg();
  enqueuePre();
  if (timing) HIP_CHECK(hipEventRecord(evs0, stream));
  Stencil<double> A{d.Ad, d.Axp, d.Ayp, d.Azp};
  PcgStats ps;
  const int rc = pcg->solve(A, d.rhs, d.x, stepOptions(), &ps);
  if (timing) HIP_CHECK(hipEventRecord(evs1, stream));
  last.iterations = ps.iterations;
  last.refinements = ps.refinements;
  last.converged = ps.converged;
  last.fp64_iterations = ps.fp64_iterations;
  last.residual_norm = ps.residual_norm;
  last.rhs_norm = ps.rhs_norm;
  if (rc != FREHG_OK)
    fail(FREHG_EDIVERGED, "gw linear solve did not converge in " + std::to_string(ps.iterations) +
                              " iterations (residual " + std::to_string(ps.residual_norm) +
                              ", ||b|| " + std::to_string(ps.rhs_norm) + ")");
  if (stepMode != FREHG_STEP_HOST && !budgetSet && !budgetReady) {
    // the first step of a device-controlled mode sizes the budget: the first
    // round (or cycle) from its own count, the later ones from theirs
    const int cap = popt.mixed ? 400 : 4096;
    auto grow = [](int v) { return v + v / 2 + 4; };
    budget.iterations = std::min(cap, std::max(8, grow(ps.first_inner)));
    if (popt.mixed) {
      budget.later = std::min(cap, std::max(6, grow(ps.later_inner)));
      budget.rounds = std::max(2, ps.rounds + 1);
    } else {
      budget.later = std::max(8, budget.iterations / 4);
      budget.rounds = std::max(1, ps.rounds);
    }
    budgetReady = true;
    recentFirst = recentLater = recentRounds = recentSteps = 0;
  }
  enqueuePost();
  if (timing) HIP_CHECK(hipEventRecord(ev1, stream));
  reduceSlots();
  finishStep(ps);
}

// The whole device-controlled step, launches only: no host synchronization,
// so the sequence can be captured.  The kernels after the solve carry the
// solve's gate.
void GwSolver::captureStep(int p) {
  hipGraph_t graph = nullptr;
  HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  const Dev before = d;
  try {
    enqueueDtg();
    enqueuePre();
    pcg->enqueueDevice(Stencil<double>{d.Ad, d.Axp, d.Ayp, d.Azp}, d.rhs, d.x, stepOptions(),
                       budget);
    d.gate = pcg->gateFlag();
    enqueuePost();
    enqueueSlots();
    d.gate = nullptr;
  } catch (...) {
    d = before;
    hipGraph_t junk = nullptr;
    (void)hipStreamEndCapture(stream, &junk);
    if (junk) (void)hipGraphDestroy(junk);
    throw;
  }
  HIP_CHECK(hipStreamEndCapture(stream, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIP_CHECK(e);
  if (p >= 0) {
    gexec[p] = exec;
    capturedDev[p] = before;
    gstats.captures += 1;
  } else {
    (void)hipGraphExecDestroy(exec);
  }
}

void GwSolver::deviceStep() {
  const PcgOptions o = stepOptions();
  pcg->prepareDevice(o);
  const Stencil<double> A{d.Ad, d.Axp, d.Ayp, d.Azp};
  if (stepMode == FREHG_STEP_GRAPH) {
    const int p = parity();
    if (gexec[p] && std::memcmp(&capturedDev[p], &d, sizeof(Dev)) != 0) invalidateGraphs();
    if (!gexec[p]) {
      captureStep(p);  // runs the host side of the step: the parity flips
    } else {
      flipParity();
    }
    HIP_CHECK(hipGraphLaunch(gexec[p], stream));
    gstats.replays += 1;
  } else {
    enqueueDtg();
    enqueuePre();
    pcg->enqueueDevice(A, d.rhs, d.x, o, budget);
    d.gate = pcg->gateFlag();
    enqueuePost();
    enqueueSlots();
    d.gate = nullptr;
  }
  HIP_CHECK(hipStreamSynchronize(stream));
  gstats.device_steps += 1;
  PcgStats ps;
  DeviceSummary sum;
  const int gate = pcg->deviceResult(&ps, &sum);
  if (gate) {
    // The solve did not converge inside the budget and the kernels after it
    // returned at once: finish the solve host-polled from where it stopped,
    // then run the rest of the step.
    gstats.host_finishes += 1;
    flipParity();
    const int rc = pcg->resumeDevice(A, d.rhs, d.x, o, &ps);
    last.iterations = ps.iterations;
    last.refinements = ps.refinements;
    last.converged = ps.converged;
    last.fp64_iterations = ps.fp64_iterations;
    last.residual_norm = ps.residual_norm;
    last.rhs_norm = ps.rhs_norm;
    if (rc != FREHG_OK)
      fail(FREHG_EDIVERGED, "gw linear solve did not converge in " +
                                std::to_string(ps.iterations) + " iterations (residual " +
                                std::to_string(ps.residual_norm) + ", ||b|| " +
                                std::to_string(ps.rhs_norm) + ")");
    if (!budgetSet) {
      // a larger budget for the next steps (graphs are captured again): the
      // rounds that ran out of launches get half as many more, else one
      // more round
      const int cap = popt.mixed ? 400 : 4096;
      auto grow = [](int v) { return v + v / 2 + 4; };
      bool grew = false;
      if (sum.first_inner >= budget.iterations) {
        budget.iterations = std::min(cap, grow(budget.iterations));
        grew = true;
      }
      const int later = budget.later > 0 ? budget.later : budget.iterations;
      if (sum.rounds > 1 && sum.later_inner >= later) {
        budget.later = std::min(cap, grow(later));
        grew = true;
      }
      if (!grew) budget.rounds = std::min(16, budget.rounds + 1);
      invalidateGraphs();
      recentFirst = recentLater = recentRounds = recentSteps = 0;
    }
    enqueuePost();
    reduceSlots();
  }
  last.iterations = ps.iterations;
  last.refinements = ps.refinements;
  last.converged = ps.converged;
  last.fp64_iterations = ps.fp64_iterations;
  last.residual_norm = ps.residual_norm;
  last.rhs_norm = ps.rhs_norm;
  if (!budgetSet && !gate) shrinkBudget(sum);
  finishStep(ps);
}

// Every 8 device steps the budget comes down to what those steps used plus a
// quarter, if that is well below it: a budget sized on a transient would
// otherwise leave launches that only return at once (each still costs its
// dispatch, a few microseconds, in a replayed graph too).
void GwSolver::shrinkBudget(const DeviceSummary& sum) {
  recentFirst = std::max(recentFirst, sum.first_inner);
  recentLater = std::max(recentLater, sum.later_inner);
  recentRounds = std::max(recentRounds, sum.rounds);
  if (++recentSteps < 8) return;
  auto fit = [](int v) { return v + v / 4 + 2; };
  PcgBudget b = budget;
  const int first = std::max(4, fit(recentFirst));
  if (first * 5 / 4 < b.iterations) b.iterations = first;
  const int later = std::max(4, fit(recentLater));
  if (b.later > 0 && later * 5 / 4 < b.later) b.later = later;
  const int rounds = popt.mixed ? std::max(2, recentRounds + 1) : std::max(1, recentRounds);
  if (rounds < b.rounds) b.rounds = rounds;
  if (b.iterations != budget.iterations || b.later != budget.later || b.rounds != budget.rounds) {
    budget = b;
    invalidateGraphs();
  }
  recentFirst = recentLater = recentRounds = recentSteps = 0;
}

void GwSolver::setStepMode(int mode, int rounds, int iterations) {
  if (mode < FREHG_STEP_HOST || mode > FREHG_STEP_GRAPH) fail(FREHG_EINVAL, "unknown step mode");
  if (rounds < 0 || iterations < 0 || rounds > 64 || iterations > 4096)
    fail(FREHG_EINVAL, "step budget out of range (rounds <= 64, iterations <= 4096)");
  if ((rounds > 0) != (iterations > 0))
    fail(FREHG_EINVAL, "give both budget values, or neither (sized from the first step)");
  if (popt.mixed && iterations > 400) fail(FREHG_EINVAL, "mixed cycles are capped at 400 iterations");
  invalidateGraphs();
  stepMode = mode;
  budgetSet = rounds > 0;
  budgetReady = budgetSet;
  if (budgetSet) {
    budget.rounds = rounds;
    budget.iterations = iterations;
    budget.later = 0;
  }
  gstats.budget_rounds = budget.rounds;
  gstats.budget_iterations = budget.iterations;
  gstats.budget_later = budget.later > 0 ? budget.later : budget.iterations;
}

// Audit and the next dt from the step's sums (slotsHost).
void GwSolver::finishStep(const PcgStats& ps) {
  const double dtg = dtgCurrent;
  const double* s = slotsHost;
  audit.ss_storage = s[S_SS];
  if (d.reallocOff) {
    audit.realloc_adjust = 0.0;
    audit.realloc_dropped = 0.0;
    audit.cpl_vent = 0.0;
    audit.cpl_vent_thin = 0.0;
  } else {
    audit.realloc_adjust = s[S_ADJ] + s[S_LOST];
    audit.realloc_dropped = s[S_DISC] - s[S_LOST];
    audit.cpl_vent = s[S_VENT];
    audit.cpl_vent_thin = s[S_THIN];
  }
  audit.vloss = s[S_VLOSS];
  audit.boundary_in = s[S_BIN];
  audit.cpl_exchanged = s[S_CEX];
  audit.cpl_bounce = s[S_CBO];
  audit.cpl_evap = s[S_CEV];
  audit.pc_clamp = s[S_CLAMP];
  audit.tr_applied = d.sink ? s[S_TR] : 0.0;
  // adaptTimeStep (AdaptiveStep.cpp:100-117)
  double next = dtg;
  if (s[S_DQMAX] > dqShrink) next = dtg * 0.75;
  else if (s[S_DQMAX] < dqGrow) next = dtg * 1.25;
  if (next > s[S_COMIN]) next = s[S_COMIN];
  if (next > dtMax) next = dtMax;
  if (next < dtMin) next = dtMin;
  dtgNext = next;
  gstats.budget_rounds = budget.rounds;
  gstats.budget_iterations = budget.iterations;
  gstats.budget_later = budget.later > 0 ? budget.later : budget.iterations;
  if (timing) {
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
    tm.step_ms += ms;
    HIP_CHECK(hipEventElapsedTime(&ms, evs0, evs1));
    tm.solve_ms += ms;
    tm.spmv_ms += ps.spmv_ms;
    tm.spmv_launches += ps.spmv_launches;
    tm.steps += 1;
    tm.pcg_iterations += ps.iterations;
  }
}

double* GwSolver::fieldPtr(int f, int64_t* n) {
  *n = g.cellsH();
  switch (f) {
    case FREHG_F_HEAD: return d.h;
    case FREHG_F_WC: return d.wc;
    case FREHG_F_HEAD_N: return d.hn;
    case FREHG_F_WC_N: return d.wcn;
    case FREHG_F_KX: return d.kx;
    case FREHG_F_KY: return d.ky;
    case FREHG_F_KZF: *n = g.cellsF(); return d.kzF;
    case FREHG_F_QX: return d.qx;
    case FREHG_F_QY: return d.qy;
    case FREHG_F_QZF: *n = g.cellsF(); return d.qzF;
    case FREHG_F_VLOSS: return d.vloss;
    case FREHG_F_DZ3D: return d.dz3d;
    case FREHG_F_BOT3D: return d.bot3d;
    case FREHG_F_AREA_X: return d.ax;
    case FREHG_F_AREA_Y: return d.ay;
    case FREHG_F_SIN_X: return d.sinx;
    case FREHG_F_COS_X: return d.cosx;
    case FREHG_F_SIN_Y: return d.siny;
    case FREHG_F_COS_Y: return d.cosy;
    case FREHG_F_CAPACITY: return d.ch;
    case FREHG_F_SINK:
      if (!d.sink) fail(FREHG_EINVAL, "no root-uptake sink is set");
      return d.sink;
    case FREHG_F_WIN_X:
    case FREHG_F_WIN_Y:
    case FREHG_F_WIN_Z:
    case FREHG_F_WIN_SINK:
    case FREHG_F_QZ_PATH_CORR:
    case FREHG_F_WIN_VENT:
    case FREHG_F_WIN_VAP:
      if (!d.tw) fail(FREHG_EINVAL, "the transport window is not enabled");
      if (f == FREHG_F_WIN_Z || f == FREHG_F_QZ_PATH_CORR) *n = g.cellsF();
      if (f == FREHG_F_WIN_VENT || f == FREHG_F_WIN_VAP)
        *n = static_cast<int64_t>(g.ny2) * g.nx2;
      return f == FREHG_F_WIN_X      ? d.winX
             : f == FREHG_F_WIN_Y    ? d.winY
             : f == FREHG_F_WIN_Z    ? d.winZ
             : f == FREHG_F_WIN_SINK ? d.winSink
             : f == FREHG_F_QZ_PATH_CORR ? d.pathCorr
             : f == FREHG_F_WIN_VENT ? d.winVent
                                     : d.winVap;
    default: fail(FREHG_EINVAL, "unknown field id");
  }
}

namespace {
template <class Fn>
int guard(Fn fn) {
  try {
    fn();
    return FREHG_OK;
  } catch (const ApiError& e) {
    g_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_error = e.what();
    return FREHG_EDEVICE;
  }
}
GwSolver* S(frehg_gw_solver* s) {
  if (!s) fail(FREHG_EINVAL, "null solver handle");
  return reinterpret_cast<GwSolver*>(s);
}
const GwSolver* S(const frehg_gw_solver* s) {
  if (!s) fail(FREHG_EINVAL, "null solver handle");
  return reinterpret_cast<const GwSolver*>(s);
}
}  // namespace

}  // namespace frehg

using frehg::fail;
using frehg::guard;
using frehg::GwSolver;

extern "C" {

int frehg_abi_version(void) { return FREHG_GW_ABI_VERSION; }
const char* frehg_last_error(void) { return frehg::g_error.c_str(); }

int frehg_gw_create(const frehg_gw_desc* desc, frehg_gw_solver** out) {
  return guard([&] {
    if (!desc || !out) frehg::fail(FREHG_EINVAL, "null argument");
    auto s = std::make_unique<GwSolver>();
    s->create(*desc);
    *out = reinterpret_cast<frehg_gw_solver*>(s.release());
  });
}
int frehg_gw_destroy(frehg_gw_solver* s) {
  delete reinterpret_cast<GwSolver*>(s);
  return FREHG_OK;
}
int frehg_gw_step(frehg_gw_solver* s, double t, double dtg) {
  return guard([&] { frehg::S(s)->step(t, dtg); });
}
int frehg_gw_next_dt(const frehg_gw_solver* s, double* dt) {
  return guard([&] { *dt = frehg::S(s)->dtgNext; });
}
int frehg_gw_get_audit(const frehg_gw_solver* s, frehg_gw_audit* out) {
  return guard([&] { *out = frehg::S(s)->audit; });
}
int frehg_gw_owned_volume(frehg_gw_solver* s, double* v) {
  return guard([&] { *v = frehg::S(s)->volume(); });
}
int frehg_gw_last_solve(const frehg_gw_solver* s, frehg_gw_solve_stats* out) {
  return guard([&] { *out = frehg::S(s)->last; });
}
int frehg_gw_active_cells(const frehg_gw_solver* s, int64_t* n) {
  return guard([&] { *n = frehg::S(s)->nActive; });
}
int frehg_gw_field_size(const frehg_gw_solver* s, int field, int64_t* n) {
  return guard([&] { (void)const_cast<GwSolver*>(frehg::S(s))->fieldPtr(field, n); });
}
int frehg_gw_get_field(frehg_gw_solver* s, int field, double* host_out) {
  return guard([&] {
    GwSolver* g = frehg::S(s);
    int64_t n = 0;
    const double* p = g->fieldPtr(field, &n);
    HIP_CHECK(hipMemcpyAsync(host_out, p, static_cast<size_t>(n) * sizeof(double),
                             hipMemcpyDeviceToHost, g->stream));
    HIP_CHECK(hipStreamSynchronize(g->stream));
  });
}
int frehg_gw_set_field(frehg_gw_solver* s, int field, const double* host_in) {
  return guard([&] {
    GwSolver* g = frehg::S(s);
    if (field == FREHG_F_SINK) {
      g->setSink(host_in);
      return;
    }
    if (field != FREHG_F_HEAD && field != FREHG_F_WC && field != FREHG_F_VLOSS)
      frehg::fail(FREHG_EINVAL, "only head, water content, vloss and the sink are settable");
    if (!host_in) frehg::fail(FREHG_EINVAL, "null argument");
    int64_t n = 0;
    double* p = g->fieldPtr(field, &n);
    HIP_CHECK(hipMemcpyAsync(p, host_in, static_cast<size_t>(n) * sizeof(double),
                             hipMemcpyHostToDevice, g->stream));
    if (field == FREHG_F_VLOSS) {  // a cumulative diagnostic: nothing derives from it
      HIP_CHECK(hipStreamSynchronize(g->stream));
      return;
    }
    g->ghosts(true, true);
    g->copy(g->d.hn, g->d.h, g->g.cellsH());
    g->copy(g->d.wcn, g->d.wc, g->g.cellsH());
    HIP_CHECK(hipStreamSynchronize(g->stream));
  });
}
int frehg_gw_set_et(frehg_gw_solver* s, const int* columns, const double* qvap) {
  return guard([&] { frehg::S(s)->setEt(columns, qvap); });
}
int frehg_gw_set_density(frehg_gw_solver* s, const double* r_rho, const double* r_visc,
                         const double* r_rho_top, const double* r_visc_top,
                         int reallocation_gradient) {
  return guard([&] {
    frehg::S(s)->setDensity(r_rho, r_visc, r_rho_top, r_visc_top, reallocation_gradient);
  });
}
int frehg_gw_enable_transport_window(frehg_gw_solver* s) {
  return guard([&] { frehg::S(s)->enableTransportWindow(); });
}
int frehg_gw_reset_transport_window(frehg_gw_solver* s) {
  return guard([&] { frehg::S(s)->resetTransportWindow(); });
}
int frehg_gw_set_bc_values(frehg_gw_solver* s, int which, const double* values) {
  return guard([&] {
    GwSolver* g = frehg::S(s);
    if (which < 0 || which > 5) frehg::fail(FREHG_EINVAL, "bc selector out of range");
    double* dst[6] = {g->d.topValue, g->d.botValue, g->d.vxm, g->d.vxp, g->d.vym, g->d.vyp};
    const int nx = g->g.nx, ny = g->g.ny, nx2 = g->g.nx2;
    std::vector<double> vv(static_cast<size_t>(g->g.ny2) * nx2, 0.0);
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const size_t c = static_cast<size_t>(j) * nx2 + i;
        const int code = g->codesHost[which][c];
        const bool member = which < 2 || (which == 2 && i == 1) || (which == 3 && i == nx) ||
                            (which == 4 && j == 1) || (which == 5 && j == ny);
        if (!member) continue;
        double v = values[(j - 1) * nx + (i - 1)];
        if (code == FREHG_BC_HEAD_HYDROSTATIC) v += g->offset;
        vv[c] = v;
      }
    HIP_CHECK(hipMemcpyAsync(dst[which], vv.data(), vv.size() * sizeof(double),
                             hipMemcpyHostToDevice, g->stream));
    HIP_CHECK(hipStreamSynchronize(g->stream));
  });
}
int frehg_gw_attach_coupling(frehg_gw_solver* s, double* eta, double* dept, double* seep,
                             double* avail, double* gain, double* bounce, double min_depth,
                             int vent_all) {
  return guard([&] {
    GwSolver* g = frehg::S(s);
    if (!eta || !dept || !seep || !avail || !gain)
      frehg::fail(FREHG_EINVAL, "coupling arrays eta, dept, seep_accum, avail, gain are required");
    g->d.eta = eta;
    g->d.dept = dept;
    g->d.seep = seep;
    g->d.avail = avail;
    g->d.gain = gain;
    g->d.bounce = bounce;
    g->d.minDepth = min_depth;
    g->d.ventAll = vent_all ? 1 : 0;
    g->d.coupled = 1;
  });
}
int frehg_cpl_begin_window(frehg_gw_solver* s) {
  return guard([&] { frehg::S(s)->beginWindow(); });
}
int frehg_cpl_refresh_depth(frehg_gw_solver* s, const double* bottom) {
  return guard([&] { frehg::S(s)->refreshDepth(bottom); });
}
int frehg_cpl_apply_seepage(frehg_gw_solver* s, const double* bottom, double dt, double* qss,
                            double* applied, double* discarded) {
  return guard([&] { frehg::S(s)->applySeepage(bottom, dt, qss, applied, discarded); });
}
int frehg_cpl_reduce_audit(frehg_gw_solver* s, const double* bottom, double* surface_gain,
                           double* accum_volume, double* surface_volume) {
  return guard([&] {
    if (!surface_gain || !accum_volume || (surface_volume && !bottom))
      frehg::fail(FREHG_EINVAL, "frehg_cpl_reduce_audit: outputs (and bottom for the volume) "
                                "are required");
    frehg::S(s)->reduceAudit(bottom, surface_gain, accum_volume, surface_volume);
  });
}
int frehg_gw_set_pcg_shapes(frehg_gw_solver* s, int march64_waves_per_cu,
                            int march32_waves_per_cu, int column_waves_per_cu) {
  return guard([&] {
    for (int v : {march64_waves_per_cu, march32_waves_per_cu, column_waves_per_cu})
      if (v < 0 || v > 256)
        frehg::fail(FREHG_EINVAL, "frehg_gw_set_pcg_shapes: waves per CU must be in [1, 256] "
                                  "(0 keeps a value)");
    GwSolver* g = frehg::S(s);
    g->invalidateGraphs();
    g->pcg->setShapes(march64_waves_per_cu, march32_waves_per_cu, column_waves_per_cu);
  });
}
int frehg_gw_get_column_tiles(const frehg_gw_solver* s, int* rows_per_lane, int* columns,
                              int* columns_per_wave, int capacity, int* buckets) {
  return guard([&] {
    if (!buckets) frehg::fail(FREHG_EINVAL, "frehg_gw_get_column_tiles: buckets is required");
    *buckets = frehg::S(s)->pcg->columnTiles(rows_per_lane, columns, columns_per_wave,
                                             capacity < 0 ? 0 : capacity);
  });
}
int frehg_gw_set_chebyshev(frehg_gw_solver* s, int degree, double lower_ratio) {
  return guard([&] {
    GwSolver* g = frehg::S(s);
    if (!g->popt.cheb)
      frehg::fail(FREHG_EINVAL, "frehg_gw_set_chebyshev: the solver's preconditioner is not "
                                "FREHG_PC_CHEBYSHEV");
    if (degree < 3 || degree > 15 || !(degree & 1) || !(lower_ratio > 0.0 && lower_ratio < 1.0))
      frehg::fail(FREHG_EINVAL, "frehg_gw_set_chebyshev: odd degree in [3, 15] and a lower "
                                "ratio in (0, 1) are required");
    g->popt.cheb_degree = degree;
    g->popt.cheb_lower = lower_ratio;
    g->invalidateGraphs();
  });
}
int frehg_gw_enable_column_solve(frehg_gw_solver* s, int enable) {
  return guard([&] {
    GwSolver* g = frehg::S(s);
    if (enable && !g->pcg->columnSolveAvailable())
      frehg::fail(FREHG_EUNSUPPORTED, "the operator has lateral legs: no column solve");
    g->pcg->enableColumnSolve(enable != 0);
    g->invalidateGraphs();
  });
}
int frehg_gw_set_column_tiles(frehg_gw_solver* s, const int* columns_per_wave, int buckets) {
  return guard([&] {
    GwSolver* g = frehg::S(s);
    if (!columns_per_wave) frehg::fail(FREHG_EINVAL, "frehg_gw_set_column_tiles: tiles required");
    if (!g->pcg->columnSolveAvailable())
      frehg::fail(FREHG_EUNSUPPORTED, "the operator has lateral legs: no column solve");
    try {
      g->pcg->setColumnTiles(columns_per_wave, buckets);
    } catch (const std::runtime_error& e) {
      frehg::fail(FREHG_EINVAL, e.what());
    }
    g->invalidateGraphs();
  });
}
int frehg_gw_set_step_mode(frehg_gw_solver* s, int mode, int rounds, int iterations) {
  return guard([&] { frehg::S(s)->setStepMode(mode, rounds, iterations); });
}
int frehg_gw_get_graph_stats(const frehg_gw_solver* s, frehg_gw_graph_stats* out) {
  return guard([&] {
    if (!out) frehg::fail(FREHG_EINVAL, "frehg_gw_get_graph_stats: output is required");
    const GwSolver* g = frehg::S(s);
    *out = g->gstats;
    out->mode = g->stepMode;
  });
}
int frehg_gw_get_pcg_shapes(const frehg_gw_solver* s, int* march64_waves_per_cu,
                            int* march32_waves_per_cu, int* column_waves_per_cu) {
  return guard([&] {
    if (!march64_waves_per_cu || !march32_waves_per_cu || !column_waves_per_cu)
      frehg::fail(FREHG_EINVAL, "frehg_gw_get_pcg_shapes: outputs are required");
    frehg::S(s)->pcg->shapes(march64_waves_per_cu, march32_waves_per_cu, column_waves_per_cu);
  });
}
int frehg_gw_set_timing(frehg_gw_solver* s, int enable) {
  return guard([&] {
    GwSolver* g = frehg::S(s);
    g->timing = enable > 0;
    g->timingEvery = enable > 0 ? enable : 1;
    const double bytes = g->tm.spmv_bytes;
    g->tm = frehg_gw_timing{};
    g->tm.spmv_bytes = bytes;
  });
}
int frehg_gw_get_timing(frehg_gw_solver* s, frehg_gw_timing* out) {
  return guard([&] { *out = frehg::S(s)->tm; });
}
int frehg_device_malloc(int64_t bytes, void** ptr) {
  return guard([&] {
    if (!ptr || bytes < 0) fail(FREHG_EINVAL, "invalid arguments");
    HIP_CHECK(hipMalloc(ptr, static_cast<size_t>(bytes > 0 ? bytes : 1)));
    HIP_CHECK(hipMemset(*ptr, 0, static_cast<size_t>(bytes > 0 ? bytes : 1)));
  });
}
int frehg_device_free(void* ptr) {
  return guard([&] {
    if (ptr) HIP_CHECK(hipFree(ptr));
  });
}
int frehg_memcpy_htod(void* dst, const void* src, int64_t bytes) {
  return guard([&] {
    HIP_CHECK(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyHostToDevice));
  });
}
int frehg_memcpy_dtoh(void* dst, const void* src, int64_t bytes) {
  return guard([&] {
    HIP_CHECK(hipDeviceSynchronize());
    HIP_CHECK(hipMemcpy(dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost));
  });
}
int frehg_gw_synchronize(frehg_gw_solver* s) {
  return guard([&] { HIP_CHECK(hipStreamSynchronize(frehg::S(s)->stream)); });
}

// ---- test exports (include/frehg_gw_testing.h) ------------------------------
int frehg_test_vg(int closure, int64_t n, const double* soil, const double* x, const double* ks,
                  double* out) {
  return guard([&] {
    if (closure < FREHG_VG_THETA || closure > FREHG_VG_CONDUCTIVITY_C || n < 0 || !soil || !x ||
        !out)
      fail(FREHG_EINVAL, "frehg_test_vg: invalid arguments");
    const bool needKs = closure == FREHG_VG_CONDUCTIVITY || closure == FREHG_VG_DKDTHETA ||
                        closure == FREHG_VG_CONDUCTIVITY_C;
    if (needKs && !ks) fail(FREHG_EINVAL, "frehg_test_vg: this closure needs ks");
    if (n == 0) return;
    double *dS = nullptr, *dX = nullptr, *dK = nullptr, *dO = nullptr;
    auto release = [&] {
      for (double* p : {dS, dX, dK, dO})
        if (p) (void)hipFree(p);
    };
    try {
      const size_t b = static_cast<size_t>(n) * sizeof(double);
      HIP_CHECK(hipMalloc(&dS, 5 * b));
      HIP_CHECK(hipMalloc(&dX, b));
      HIP_CHECK(hipMalloc(&dK, b));
      HIP_CHECK(hipMalloc(&dO, b));
      HIP_CHECK(hipMemcpy(dS, soil, 5 * b, hipMemcpyHostToDevice));
      HIP_CHECK(hipMemcpy(dX, x, b, hipMemcpyHostToDevice));
      HIP_CHECK(hipMemset(dK, 0, b));
      if (ks) HIP_CHECK(hipMemcpy(dK, ks, b, hipMemcpyHostToDevice));
      const int blocks = static_cast<int>(std::min<int64_t>((n + 255) / 256, 1024));
      frehg::k_test_vg<<<blocks, 256>>>(closure, n, dS, dX, dK, dO);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpy(out, dO, b, hipMemcpyDeviceToHost));
    } catch (...) {
      release();
      throw;
    }
    release();
  });
}
int frehg_gw_test_post_allocate(frehg_gw_solver* s, double dtg) {
  return guard([&] { frehg::S(s)->testPostAllocate(dtg); });
}

}  // extern "C"

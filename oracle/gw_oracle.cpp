// gw_oracle.cpp — CPU FP64 restatement of Frehg2's groundwater step (the oracle).
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the
// MI355X path: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker or the CPU baseline.
// The product library never links or calls it.
//
// It restates, in plain serial/OpenMP C++ with IEEE double arithmetic and no
// FMA contraction (x86-64 baseline ISA), the reference's algorithm:
//   VanGenuchten.hpp:45-128        van Genuchten-Mualem closures
//   FaceConductivity.hpp:37-74     arithmetic / harmonic face rules
//   Predictor.cpp:91-386           face conductivities  <PC, HARM>
//   Predictor.cpp:392-698          capacity + 7-point assembly <PC>
//   Predictor.cpp:732-817          solve hand-off, scatter, ghost heads
//   Corrector.cpp:47-118           coupled-top classification <PC>
//   Corrector.cpp:123-533          Darcy fluxes <PC, HARM>
// PC = true is predictor_corrector: consistent (the default); the legacy pair
// <false, false> (pcLegacy) differs in four rules, each marked "legacy".
//   Corrector.cpp:539-627          coupled-top bookkeeping
//   Corrector.cpp:629-705          flux-divergence theta update, root-uptake sink
//   Predictor.cpp:700-726          the sink in the rhs
//   Corrector.cpp:709-792          moisture ghosts, final clamp
//   Reallocate.cpp:122-545         post-allocation (saturated_contact: head)
//   AdaptiveStep.cpp:44-118        dtg controller
//   RichardsSolver.cpp:331-427,729-794,892-1006  IC, step order, audits
//   TerrainMetric.cpp:69-430       regular / terrain-following geometry
//   Grid.cpp:160-245               compressed (j,i,k) global ids
//   LinearSystem.cpp:147-227,332-484  CG semantics (PETSc KSPCG replaced by
//                                  an explicit PCG; see the solver notes)
// Density/viscosity ratios (r_rho, r_visc and their face means, Baroclinic.cpp)
// enter every formula where the reference has them; they are 1 until a caller
// sets cell ratios (oracle_gw_set_density), and multiplying by 1.0 is exact.
//
// Parity pins (tests/test_oracle_*.py): the reference's hand values
// (tests/unit/test_vangenuchten.cpp:43-126), its gw invariants
// (tests/unit/test_gw.cpp), Poisson exact solutions
// (tests/unit/test_linearsystem.cpp:34-111), and closed forms / brute force
// (dense solves, symmetry, uniform-column invariance).

#include "../include/frehg_gw.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

thread_local std::string g_error;

struct OracleError : std::runtime_error {
  int code;
  OracleError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string& msg) { throw OracleError(code, msg); }

int g_threads = 1;

// ---------------------------------------------------------------------------
// van Genuchten-Mualem closures (VanGenuchten.hpp:45-128).
// ---------------------------------------------------------------------------
struct Soil {
  double vga, vgn, ths, thr, aev;
};

double theta_of_h(const Soil& s, double h) {
  const double m = 1.0 - 1.0 / s.vgn;
  double wc;
  if (h > s.aev) {
    wc = s.ths;
  } else {
    const double se = std::pow(1.0 + std::pow(std::fabs(s.vga * h), s.vgn), -m);
    wc = s.thr + (s.ths - s.thr) * se;
  }
  if (wc > s.ths) wc = s.ths;
  else if (wc < s.thr) wc = s.thr;
  return wc;
}

double h_of_theta(const Soil& s, double wc) {
  const double eps = 1.0e-7;
  const double m = 1.0 - 1.0 / s.vgn;
  if (wc - s.thr < eps) wc = s.thr + eps;
  if (wc < s.ths) {
    const double ratio = (s.ths - s.thr) / (wc - s.thr);
    return -(1.0 / s.vga) * std::pow(std::pow(ratio, 1.0 / m) - 1.0, 1.0 / s.vgn);
  }
  return 0.0;
}

double capacity_of_h(const Soil& s, double h) {
  if (h > 0.0) return 0.0;
  const double m = 1.0 - 1.0 / s.vgn;
  const double ah = std::fabs(s.vga * h);
  const double top = s.vga * s.vgn * m * (s.ths - s.thr) * std::pow(ah, s.vgn - 1.0);
  const double bot = std::pow(1.0 + std::pow(ah, s.vgn), m + 1.0);
  return top / bot;
}

double k_of_h(const Soil& s, double h, double ks) {
  const double m = 1.0 - 1.0 / s.vgn;
  const double se = std::pow(1.0 + std::pow(std::fabs(s.vga * h), s.vgn), -m);
  const double br = 1.0 - std::pow(1.0 - std::pow(se, 1.0 / m), m);
  double k = ks * std::pow(se, 0.5) * br * br;
  if (k > ks) k = ks;
  if (h > 0.0) k = ks;
  return k;
}

double dk_dtheta(const Soil& s, double wc, double ks) {
  if (wc > 0.9999 * s.ths && wc < s.ths) wc = 0.9999 * s.ths;
  const double m = 1.0 - 1.0 / s.vgn;
  const double se = (wc - s.thr) / (s.ths - s.thr);
  const double t0 = std::pow(1.0 - std::pow(se, 1.0 / m), m);
  const double t1 = 0.5 * ks * std::pow(se, -0.5) * (1.0 - t0) * (1.0 - t0);
  const double t2 = 2.0 * ks * std::pow(se, (2.0 - m) / (2.0 * m)) * (1.0 - t0) *
                    std::pow(1.0 - std::pow(se, 1.0 / m), m - 1.0);
  return (t1 + t2) / (s.ths - s.thr);
}

// FaceConductivity.hpp:37-58.
double face_k_harmonic(const Soil& sa, double ha, double ksa, double la, const Soil& sb,
                       double hb, double ksb, double lb) {
  if (ksa <= 0.0 || ksb <= 0.0) return 0.0;
  const double ksh = (la + lb) / (la / ksa + lb / ksb);
  const double kra = k_of_h(sa, ha, ksa) / ksa;
  const double krb = k_of_h(sb, hb, ksb) / ksb;
  return ksh * 0.5 * (kra + krb);
}

// ---------------------------------------------------------------------------
// Halo-layout storage (Types.hpp:35-43): (j, i, k), k unit stride.
// ---------------------------------------------------------------------------
struct F3 {
  int ny2 = 0, nx2 = 0, nk = 0;
  std::vector<double> v;
  void init(int ny2_, int nx2_, int nk_, double fill) {
    ny2 = ny2_;
    nx2 = nx2_;
    nk = nk_;
    v.assign(static_cast<size_t>(ny2) * nx2 * nk, fill);
  }
  double& operator()(int j, int i, int k) {
    return v[(static_cast<size_t>(j) * nx2 + i) * nk + k];
  }
  double operator()(int j, int i, int k) const {
    return v[(static_cast<size_t>(j) * nx2 + i) * nk + k];
  }
};
template <class T>
struct F2 {
  int ny2 = 0, nx2 = 0;
  std::vector<T> v;
  void init(int ny2_, int nx2_, T fill) {
    ny2 = ny2_;
    nx2 = nx2_;
    v.assign(static_cast<size_t>(ny2) * nx2, fill);
  }
  T& operator()(int j, int i) { return v[static_cast<size_t>(j) * nx2 + i]; }
  T operator()(int j, int i) const { return v[static_cast<size_t>(j) * nx2 + i]; }
};

// Deterministic reduction: one partial per j row, summed serially in j order.
template <class Fn>
double sum_rows(int j0, int j1, Fn fn) {
  std::vector<double> part(static_cast<size_t>(j1 - j0), 0.0);
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = j0; j < j1; ++j) part[j - j0] = fn(j);
  double s = 0.0;
  for (double p : part) s += p;
  return s;
}

// ---------------------------------------------------------------------------
// The oracle solver state (RichardsSolver.hpp:534-680, single rank).
// ---------------------------------------------------------------------------
struct Oracle {
  // grid
  int nx = 0, ny = 0, nz = 0;
  double dx = 0, dy = 0;
  double az = 0;
  std::vector<double> dzk;
  double totalDepth = 0;
  // options
  double ss = 0;
  bool full3d = true, harmonic = false, reallocOff = false, redistribute = false;
  bool pcLegacy = false;  // predictor_corrector: legacy, the <false, false> pair
  double dtMin = 0, dtMax = 0, dqGrow = 0.01, dqShrink = 0.02, courantMax = 2.0;
  double dtgNext = 0, dtgCurrent = 0;
  double rtol = 1e-8, atol = 1e-14;
  int maxIt = 1000, pc = FREHG_PC_JACOBI;
  bool guessPrevious = false;
  // geometry (TerrainMetric)
  bool followTerrain = false, uniformLayers = false;
  double offset = 0, boxTop = 0, stretch = 1.0;
  F2<double> bath;
  F3 dz3d, bot3d, ax, ay, sinx, cosx, siny, cosy;
  std::vector<int> ktop;  // ny*nx
  // ids (Grid::buildGlobalIds)
  std::vector<int64_t> gid;  // halo layout
  int64_t nActive = 0;
  // fields
  F3 h, hn, wc, wcn, ch, ksx, ksy, ksz, vga, vgn, wcs, wcr, aev;
  F3 sink;               // root-uptake sink S [1/s] (RichardsSolver.hpp:656)
  bool sinkActive = false;
  // density / viscosity ratios: cells (H, ghosts included), x+ / y+ face means
  // (H) and z face means (F, plane k = face above cell k); all 1 by default
  F3 rRho, rVisc, rXp, rYp, vXp, vYp, rZp, vZp;
  bool densityGradient = false;  // reallocation_gradient: density
  // the evapotranspiration soil channel (RichardsSolver.hpp:620-636): columns
  // marked in etCol take the applied flux etQ [m/s] at their top face
  std::vector<int> etCol;
  std::vector<double> etQ;
  bool etActive = false;
  void setDensity(const double* rr, const double* rv, const double* rrTop, const double* rvTop,
                  bool gradient);
  F3 kx, ky, kzF, qx, qy, qzF, vloss, room, sendUp, sendDown, wcSnap, hSnap;
  // the transport window (RichardsSolver.cpp:796-878): face volumes summed
  // over the steps since the last reset, and the walks' path corrections
  bool transportWindow = false;
  F3 winX, winY, winZ, winSink, pathCorr;
  F2<double> winVent, winVap;
  void enableTransportWindow();
  void resetTransportWindow();
  void accumulateTransportWindow(double dtg);
  // boundary markers
  F2<int> topCode, botCode, sxm, sxp, sym, syp, cplMode;
  F2<double> topValue, botValue, vxm, vxp, vym, vyp;
  // coupling (GwCoupling): host arrays aliased from the caller
  bool coupled = false, ventAll = false;
  double minDepth = 0;
  double *eta = nullptr, *dept = nullptr, *seep = nullptr, *avail = nullptr, *gain = nullptr,
         *bounce = nullptr;
  // linear system (COO order [diag,xm,xp,ym,yp,zm,zp], RichardsSolver.cpp:680-727)
  std::vector<int64_t> cols;
  std::vector<double> vals, rhs, sol;
  std::vector<int> rowJ, rowI, rowK;  // row -> cell
  frehg_gw_solve_stats last{};
  frehg_gw_audit audit{};
  double pcClampTotal = 0;
  double tSolve = 0, tStep = 0;
  int64_t itTotal = 0;

  int64_t g(int j, int i, int k) const {
    return gid[(static_cast<size_t>(j) * (nx + 2) + i) * nz + k];
  }
  Soil soil(int j, int i, int k) const {
    return Soil{vga(j, i, k), vgn(j, i, k), wcs(j, i, k), wcr(j, i, k), aev(j, i, k)};
  }
  double kh(int j, int i, int k, const F3& ks) const {
    return k_of_h(soil(j, i, k), h(j, i, k), ks(j, i, k));
  }
  // EdgeTopology (Grid.hpp:101-124), non-periodic, single rank: local i = gi+1.
  bool west(int i) const { return i == 1; }
  bool east(int i) const { return i == nx; }
  bool south(int j) const { return j == 1; }
  bool north(int j) const { return j == ny; }
  bool outX(int i) const { return i < 1 || i > nx; }
  bool outY(int j) const { return j < 1 || j > ny; }
  static bool isHead(int c) { return c == FREHG_BC_HEAD || c == FREHG_BC_HEAD_HYDROSTATIC; }

  void create(const frehg_gw_desc& d);
  void buildGeometry(const frehg_gw_desc& d);
  void buildIds();
  void applyIC(const frehg_gw_desc& d);
  void enforceHeadBc();
  void enforceMoistureBc();
  void classifyCoupledTop(double dtg);
  void faceConductivity();
  void assemble(double dtg);
  void solve();
  void fluxes(double dtg);
  void coupledBookkeeping(double dtg);
  void updateWaterContent(double dtg);
  void reallocate();
  void finalize();
  void boundaryFlux(double dtg);
  void adaptTimeStep(double dtg);
  double ownedVolume() const;
  void step(double t, double dtg);
};

// --- geometry: TerrainMetric.cpp:101-430 ---------------------------------
void Oracle::buildGeometry(const frehg_gw_desc& d) {
  const int ny2 = ny + 2, nx2 = nx + 2;
  bath.init(ny2, nx2, 0.0);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      bath(j, i) = d.bottom_elevation ? d.bottom_elevation[(j - 1) * nx + (i - 1)]
                                      : d.bottom_elevation_const;
  // lift into the non-negative frame (:128-148)
  double lo = 0.0, hi = 0.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      lo = bath(j, i) < lo ? bath(j, i) : lo;
      hi = bath(j, i) > hi ? bath(j, i) : hi;
    }
  offset = lo < 0.0 ? -lo : 0.0;
  boxTop = hi + offset;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) bath(j, i) += offset;
  // edge ghosts of the raster (:16-37)
  for (int j = 1; j <= ny; ++j) {
    bath(j, 0) = bath(j, 1);
    bath(j, nx + 1) = bath(j, nx);
  }
  for (int i = 0; i <= nx + 1; ++i) {
    bath(0, i) = bath(1, i);
    bath(ny + 1, i) = bath(ny, i);
  }

  dz3d.init(ny2, nx2, nz, 0.0);
  bot3d.init(ny2, nx2, nz, 0.0);
  ktop.assign(static_cast<size_t>(ny) * nx, 0);
  if (!followTerrain) {
    // regular mesh (:158-230)
    std::vector<double> bot1d(nz);
    double level = boxTop;
    for (int k = 0; k < nz; ++k) {
      level -= dzk[k];
      bot1d[k] = level;
    }
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double b = bath(j, i);
        std::vector<int> act(nz, 0);
        for (int k = 0; k < nz; ++k) {
          bot3d(j, i, k) = bot1d[k];
          dz3d(j, i, k) = dzk[k];
        }
        for (int k = 0; k < nz; ++k) {
          if (bot3d(j, i, k) >= b - 1.0e-5) {
            act[k] = 0;
            if (bot3d(j, i, k) - b < dzk[k]) bot3d(j, i, k) = b;
          } else {
            act[k] = 1;
            if (b - bot3d(j, i, k) <= dzk[k]) {
              if (b - bot3d(j, i, k) >= 0.25 * dzk[k]) {
                dz3d(j, i, k) = b - bot3d(j, i, k);
              } else if (k + 1 < nz) {
                act[k] = 0;
                dz3d(j, i, k + 1) += b - bot3d(j, i, k);
              }
            }
          }
        }
        int top = nz;
        for (int k = 0; k < nz; ++k)
          if (act[k] == 1) {
            top = k;
            break;
          }
        ktop[(j - 1) * nx + (i - 1)] = top;
      }
  } else {
    // terrain-following (:232-293)
    const double boxBottom = boxTop - totalDepth;
    double increSum = 0.0, w = 1.0;
    for (int k = 0; k < nz; ++k) {
      increSum += w;
      w *= stretch;
    }
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const double b = bath(j, i);
        if (b <= boxBottom && !uniformLayers)
          fail(FREHG_EINVAL, "follow_terrain: a bed sits at or below the subsurface box bottom");
        double dzv = uniformLayers ? dzk[0] : (b - boxBottom) / increSum;
        double bot = b;
        for (int k = 0; k < nz; ++k) {
          if (uniformLayers) dzv = dzk[k];
          bot -= dzv;
          dz3d(j, i, k) = dzv;
          bot3d(j, i, k) = bot;
          if (!uniformLayers) dzv *= stretch;
        }
        ktop[(j - 1) * nx + (i - 1)] = 0;
      }
  }
  // edge ghosts of dz3d / bot3d (:39-65)
  for (F3* f : {&dz3d, &bot3d})
    for (int k = 0; k < nz; ++k) {
      for (int j = 1; j <= ny; ++j) {
        (*f)(j, 0, k) = (*f)(j, 1, k);
        (*f)(j, nx + 1, k) = (*f)(j, nx, k);
      }
      for (int i = 0; i <= nx + 1; ++i) {
        (*f)(0, i, k) = (*f)(1, i, k);
        (*f)(ny + 1, i, k) = (*f)(ny, i, k);
      }
    }
  // metrics (:318-408)
  ax.init(ny2, nx2, nz, 0.0);
  ay.init(ny2, nx2, nz, 0.0);
  sinx.init(ny2, nx2, nz, 0.0);
  cosx.init(ny2, nx2, nz, 1.0);
  siny.init(ny2, nx2, nz, 0.0);
  cosy.init(ny2, nx2, nz, 1.0);
  if (followTerrain) {
    for (int j = 0; j <= ny + 1; ++j)
      for (int i = 0; i <= nx; ++i) {
        if (outX(i) || east(i)) continue;
        for (int k = 0; k < nz; ++k) {
          const double zc = bot3d(j, i, k) + 0.5 * dz3d(j, i, k);
          const double zp = bot3d(j, i + 1, k) + 0.5 * dz3d(j, i + 1, k);
          const double hd = std::fabs(zp - zc);
          const double dist = std::sqrt(hd * hd + dx * dx);
          sinx(j, i, k) = hd / dist;
          cosx(j, i, k) = dx / dist;
        }
      }
    for (int j = 0; j <= ny; ++j) {
      if (outY(j) || north(j)) continue;
      for (int i = 0; i <= nx + 1; ++i)
        for (int k = 0; k < nz; ++k) {
          const double zc = bot3d(j, i, k) + 0.5 * dz3d(j, i, k);
          const double zp = bot3d(j + 1, i, k) + 0.5 * dz3d(j + 1, i, k);
          const double hd = std::fabs(zp - zc);
          const double dist = std::sqrt(hd * hd + dy * dy);
          siny(j, i, k) = hd / dist;
          cosy(j, i, k) = dy / dist;
        }
    }
  }
  for (int j = 0; j <= ny + 1; ++j)
    for (int i = 0; i <= nx + 1; ++i)
      for (int k = 0; k < nz; ++k) {
        if (followTerrain) {
          const double dzE = (i <= nx) ? dz3d(j, i + 1, k) : dz3d(j, i, k);
          const double dzN = (j <= ny) ? dz3d(j + 1, i, k) : dz3d(j, i, k);
          ax(j, i, k) = 0.5 * (dz3d(j, i, k) + dzE) * cosx(j, i, k) * dy;
          ay(j, i, k) = 0.5 * (dz3d(j, i, k) + dzN) * cosy(j, i, k) * dx;
        } else {
          ax(j, i, k) = dy * dz3d(j, i, k);
          ay(j, i, k) = dx * dz3d(j, i, k);
        }
      }
}

// --- Grid::buildGlobalIds (Grid.cpp:160-245), single rank -----------------
void Oracle::buildIds() {
  gid.assign(static_cast<size_t>(ny + 2) * (nx + 2) * nz, -1);
  int64_t next = 0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const int kt = ktop[(j - 1) * nx + (i - 1)];
      for (int k = kt; k < nz; ++k) gid[(static_cast<size_t>(j) * (nx + 2) + i) * nz + k] = next++;
    }
  nActive = next;
  rowJ.resize(nActive);
  rowI.resize(nActive);
  rowK.resize(nActive);
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const int64_t r = g(j, i, k);
        if (r >= 0) {
          rowJ[r] = j;
          rowI[r] = i;
          rowK[r] = k;
        }
      }
  // COO columns (RichardsSolver.cpp:696-721)
  cols.assign(7 * nActive, -1);
  for (int64_t r = 0; r < nActive; ++r) {
    const int j = rowJ[r], i = rowI[r], k = rowK[r];
    cols[7 * r + 0] = r;
    cols[7 * r + 1] = g(j, i - 1, k);
    cols[7 * r + 2] = g(j, i + 1, k);
    cols[7 * r + 3] = g(j - 1, i, k);
    cols[7 * r + 4] = g(j + 1, i, k);
    cols[7 * r + 5] = (k > 0) ? g(j, i, k - 1) : -1;
    cols[7 * r + 6] = (k + 1 < nz) ? g(j, i, k + 1) : -1;
  }
  vals.assign(7 * nActive, 0.0);
  rhs.assign(nActive, 0.0);
  sol.assign(nActive, 0.0);
}

// --- initial conditions (RichardsSolver.cpp:331-427) ----------------------
void Oracle::applyIC(const frehg_gw_desc& d) {
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const Soil s = soil(j, i, k);
        const double zc = bot3d(j, i, k) + 0.5 * dz3d(j, i, k);
        double value = d.ic_const;
        if (d.ic_values) {
          const size_t col = static_cast<size_t>(j - 1) * nx + (i - 1);
          value = (d.ic_form == FREHG_IC_WATER_TABLE) ? d.ic_values[col]
                                                      : d.ic_values[col * nz + k];
        }
        double hv = 0.0, w = 0.0;
        if (d.ic_form == FREHG_IC_MOISTURE) {
          w = value;
          hv = (w >= s.ths) ? (bath(j, i) - zc) : h_of_theta(s, w);
        } else if (d.ic_form == FREHG_IC_HEAD) {
          hv = value;
          w = theta_of_h(s, hv);
        } else {
          hv = (value + offset) - zc;
          w = theta_of_h(s, hv);
        }
        h(j, i, k) = hv;
        wc(j, i, k) = w;
      }
  enforceHeadBc();
  enforceMoistureBc();
  hn = h;
  wcn = wc;
}

void Oracle::create(const frehg_gw_desc& d) {
  if (d.nx < 1 || d.ny < 1 || d.nz < 1) fail(FREHG_EINVAL, "extents must be positive");
  if (!(d.dx > 0) || !(d.dy > 0) || !(d.dz > 0)) fail(FREHG_EINVAL, "spacings must be positive");
  if (d.predictor_corrector != 0 && d.predictor_corrector != 1)
    fail(FREHG_EINVAL, "predictor_corrector: 1 consistent or 0 legacy");
  if (d.saturated_contact != 1)
    fail(FREHG_EUNSUPPORTED, "saturated_contact: legacy is order-dependent; head only");
  if (!d.ksx || !d.ksy || !d.ksz || !d.vg_alpha || !d.vg_n || !d.theta_s || !d.theta_r || !d.aev)
    fail(FREHG_EINVAL, "all eight per-cell soil arrays are required");
  nx = d.nx;
  ny = d.ny;
  nz = d.nz;
  dx = d.dx;
  dy = d.dy;
  az = dx * dy;
  stretch = d.dz_stretch > 0 ? d.dz_stretch : 1.0;
  dzk.resize(nz);
  totalDepth = 0;
  for (int k = 0; k < nz; ++k) {
    dzk[k] = d.dz * std::pow(stretch, k);  // Grid.cpp:44-49
    totalDepth += dzk[k];
  }
  followTerrain = d.follow_terrain != 0;
  uniformLayers = d.terrain_layers == 1;
  ss = d.specific_storage;
  full3d = d.use_full3d != 0;
  pcLegacy = d.predictor_corrector == 0;
  // the legacy pair has no harmonic instantiation (Predictor.cpp:94-102,
  // Corrector.cpp:124-132)
  harmonic = d.face_conductivity == 1 && !pcLegacy;
  reallocOff = d.reallocation == 1;
  redistribute = d.reallocation_surplus == 1;
  dtMin = d.dt_min;
  dtMax = d.dt_max;
  dqGrow = d.dq_grow;
  dqShrink = d.dq_shrink;
  courantMax = d.courant_max;
  dtgNext = d.dt_init;
  rtol = d.rtol;
  atol = d.atol;
  maxIt = d.max_iterations > 0 ? d.max_iterations : 1000;
  pc = d.preconditioner;
  if (pc != FREHG_PC_JACOBI && pc != FREHG_PC_ZLINE && pc != FREHG_PC_ICC0)
    fail(FREHG_EUNSUPPORTED, "the oracle's preconditioners are jacobi, zline and icc0");
  guessPrevious = d.initial_guess_previous != 0;

  buildGeometry(d);
  buildIds();

  const int ny2 = ny + 2, nx2 = nx + 2;
  for (F3* f : {&h, &hn, &wc, &wcn, &ch, &ksx, &ksy, &ksz, &vga, &vgn, &wcs, &wcr, &aev, &kx, &ky,
                &qx, &qy, &vloss, &room, &sendUp, &sendDown, &wcSnap, &hSnap})
    f->init(ny2, nx2, nz, 0.0);
  kzF.init(ny2, nx2, nz + 1, 0.0);
  for (F3* f : {&rRho, &rVisc, &rXp, &rYp, &vXp, &vYp}) f->init(ny2, nx2, nz, 1.0);
  rZp.init(ny2, nx2, nz + 1, 1.0);
  vZp.init(ny2, nx2, nz + 1, 1.0);
  qzF.init(ny2, nx2, nz + 1, 0.0);
  // soil staging: interior fill + mirrored edge ghosts (RichardsSolver.cpp:207-329, :24-57)
  const double* src[8] = {d.ksx, d.ksy, d.ksz, d.vg_alpha, d.vg_n, d.theta_s, d.theta_r, d.aev};
  F3* dst[8] = {&ksx, &ksy, &ksz, &vga, &vgn, &wcs, &wcr, &aev};
  for (int a = 0; a < 8; ++a) {
    F3& f = *dst[a];
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        for (int k = 0; k < nz; ++k)
          f(j, i, k) = src[a][(static_cast<size_t>(j - 1) * nx + (i - 1)) * nz + k];
    for (int j = 0; j <= ny + 1; ++j)
      for (int k = 0; k < nz; ++k) {
        f(j, 0, k) = f(j, 1, k);
        f(j, nx + 1, k) = f(j, nx, k);
      }
    for (int i = 0; i <= nx + 1; ++i)
      for (int k = 0; k < nz; ++k) {
        f(0, i, k) = f(1, i, k);
        f(ny + 1, i, k) = f(ny, i, k);
      }
  }
  // boundary markers (RichardsSolver.cpp:429-566)
  for (F2<int>* c : {&topCode, &botCode, &sxm, &sxp, &sym, &syp, &cplMode})
    c->init(ny2, nx2, FREHG_BC_NOFLUX);
  for (F2<double>* v : {&topValue, &botValue, &vxm, &vxp, &vym, &vyp}) v->init(ny2, nx2, 0.0);
  const int* cs[6] = {d.top_code, d.bot_code, d.xm_code, d.xp_code, d.ym_code, d.yp_code};
  const double* vs[6] = {d.top_value, d.bot_value, d.xm_value, d.xp_value, d.ym_value, d.yp_value};
  F2<int>* cd[6] = {&topCode, &botCode, &sxm, &sxp, &sym, &syp};
  F2<double>* vd[6] = {&topValue, &botValue, &vxm, &vxp, &vym, &vyp};
  for (int a = 0; a < 6; ++a)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        const bool member = a < 2 || (a == 2 && i == 1) || (a == 3 && i == nx) ||
                            (a == 4 && j == 1) || (a == 5 && j == ny);
        if (!member) continue;
        const size_t c = static_cast<size_t>(j - 1) * nx + (i - 1);
        const int code = cs[a] ? cs[a][c] : FREHG_BC_NOFLUX;
        if (code < 0 || code > 4) fail(FREHG_EINVAL, "boundary code out of range");
        (*cd[a])(j, i) = code;
        double v = vs[a] ? vs[a][c] : 0.0;
        if (code == FREHG_BC_HEAD_HYDROSTATIC) v += offset;  // :529-533
        (*vd[a])(j, i) = v;
      }
  applyIC(d);
}

// --- Baroclinic.cpp:46-187: face means of the density / viscosity ratios ----
// rr, rv: cell ratios over the halo box (H); rrTop, rvTop: the ratios of the
// ghost above each column's top cell (2D, (ny+2)(nx+2); the surface scalar's
// in coupled runs), null for the top cell's own.  rr = null resets every
// ratio to 1.
void Oracle::setDensity(const double* rr, const double* rv, const double* rrTop,
                        const double* rvTop, bool gradient) {
  for (F3* f : {&rRho, &rVisc, &rXp, &rYp, &vXp, &vYp}) std::fill(f->v.begin(), f->v.end(), 1.0);
  std::fill(rZp.v.begin(), rZp.v.end(), 1.0);
  std::fill(vZp.v.begin(), vZp.v.end(), 1.0);
  densityGradient = gradient && rr != nullptr;
  if (!rr) return;
  std::copy(rr, rr + rRho.v.size(), rRho.v.begin());
  if (rv) std::copy(rv, rv + rVisc.v.size(), rVisc.v.begin());
  const auto headCode = [](int c) { return isHead(c); };
  // x faces: slot i is the face east of cell i, slot 0 the west boundary face
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        double r = 1.0, v = 1.0;
        if (i == 0) {
          if (g(j, 1, k) >= 0) {
            if (headCode(sxm(j, 1))) {
              r = rRho(j, 0, k);
              v = rVisc(j, 0, k);
            } else {
              r = 0.5 * (rRho(j, 1, k) + rRho(j, 0, k));
              v = 0.5 * (rVisc(j, 1, k) + rVisc(j, 0, k));
            }
          }
        } else if (i == nx) {
          if (g(j, i, k) >= 0 && headCode(sxp(j, i))) {
            r = rRho(j, i + 1, k);
            v = rVisc(j, i + 1, k);
          }
        } else if (g(j, i, k) >= 0 && g(j, i + 1, k) >= 0) {
          r = 0.5 * (rRho(j, i, k) + rRho(j, i + 1, k));
          v = 0.5 * (rVisc(j, i, k) + rVisc(j, i + 1, k));
        }
        rXp(j, i, k) = r;
        vXp(j, i, k) = v;
      }
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        double r = 1.0, v = 1.0;
        if (j == 0) {
          if (g(1, i, k) >= 0) {
            if (headCode(sym(1, i))) {
              r = rRho(0, i, k);
              v = rVisc(0, i, k);
            } else {
              r = 0.5 * (rRho(1, i, k) + rRho(0, i, k));
              v = 0.5 * (rVisc(1, i, k) + rVisc(0, i, k));
            }
          }
        } else if (j == ny) {
          if (g(j, i, k) >= 0 && headCode(syp(j, i))) {
            r = rRho(j + 1, i, k);
            v = rVisc(j + 1, i, k);
          }
        } else if (g(j, i, k) >= 0 && g(j + 1, i, k) >= 0) {
          r = 0.5 * (rRho(j, i, k) + rRho(j + 1, i, k));
          v = 0.5 * (rVisc(j, i, k) + rVisc(j + 1, i, k));
        }
        rYp(j, i, k) = r;
        vYp(j, i, k) = v;
      }
  // z faces: plane k above cell k; the bottom plane nz stays 1
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k <= nz; ++k) {
        const bool below = k < nz && g(j, i, k) >= 0;
        const bool above = k >= 1 && g(j, i, k - 1) >= 0;
        double r = 1.0, v = 1.0;
        if (below && above) {
          r = 0.5 * (rRho(j, i, k - 1) + rRho(j, i, k));
          v = 0.5 * (rVisc(j, i, k - 1) + rVisc(j, i, k));
        } else if (below) {
          const size_t c2 = static_cast<size_t>(j) * (nx + 2) + i;
          r = 0.5 * (rRho(j, i, k) + (rrTop ? rrTop[c2] : rRho(j, i, k)));
          v = 0.5 * (rVisc(j, i, k) + (rvTop ? rvTop[c2] : rVisc(j, i, k)));
        }
        rZp(j, i, k) = r;
        vZp(j, i, k) = v;
      }
}

// --- ghost heads (Predictor.cpp:66-87, :757-817) --------------------------
// rFace: the boundary face's density ratio; the plus edges scale the depth
// term by it, the minus edges do not (Predictor.cpp:56-62)
static double ghost_head(int code, double value, double interior, double zc, bool coupled,
                         double bed, double depth, double rFace, bool plus) {
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

void Oracle::enforceHeadBc() {
  auto dep = [&](int j, int i) { return coupled ? dept[static_cast<size_t>(j) * (nx + 2) + i] : 0.0; };
  for (int j = 1; j <= ny; ++j)
    for (int k = 0; k < nz; ++k) {
      double zc = bot3d(j, 1, k) + 0.5 * dz3d(j, 1, k);
      h(j, 0, k) = ghost_head(sxm(j, 1), vxm(j, 1), h(j, 1, k), zc, coupled, bath(j, 1), dep(j, 1),
                              rXp(j, 0, k), false);
      zc = bot3d(j, nx, k) + 0.5 * dz3d(j, nx, k);
      h(j, nx + 1, k) =
          ghost_head(sxp(j, nx), vxp(j, nx), h(j, nx, k), zc, coupled, bath(j, nx), dep(j, nx),
                     rXp(j, nx, k), true);
    }
  for (int i = 1; i <= nx; ++i)
    for (int k = 0; k < nz; ++k) {
      double zc = bot3d(1, i, k) + 0.5 * dz3d(1, i, k);
      h(0, i, k) = ghost_head(sym(1, i), vym(1, i), h(1, i, k), zc, coupled, bath(1, i), dep(1, i),
                              rYp(0, i, k), false);
      zc = bot3d(ny, i, k) + 0.5 * dz3d(ny, i, k);
      h(ny + 1, i, k) =
          ghost_head(syp(ny, i), vyp(ny, i), h(ny, i, k), zc, coupled, bath(ny, i), dep(ny, i),
                     rYp(ny, i, k), true);
    }
}

// Corrector.cpp:709-756.
void Oracle::enforceMoistureBc() {
  for (int j = 1; j <= ny; ++j)
    for (int k = 0; k < nz; ++k) {
      wc(j, 0, k) = isHead(sxm(j, 1)) ? wcs(j, 1, k) : wc(j, 1, k);
      wc(j, nx + 1, k) = isHead(sxp(j, nx)) ? wcs(j, nx, k) : wc(j, nx, k);
    }
  for (int i = 1; i <= nx; ++i)
    for (int k = 0; k < nz; ++k) {
      wc(0, i, k) = isHead(sym(1, i)) ? wcs(1, i, k) : wc(1, i, k);
      wc(ny + 1, i, k) = isHead(syp(ny, i)) ? wcs(ny, i, k) : wc(ny, i, k);
    }
}

// --- Corrector.cpp:55-118 <PC> --------------------------------------------
void Oracle::classifyCoupledTop(double dtg) {
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      int kt = -1;
      for (int k = 0; k < nz; ++k)
        if (g(j, i, k) >= 0) {
          kt = k;
          break;
        }
      const size_t c2 = static_cast<size_t>(j) * (nx + 2) + i;
      if (kt < 0 || dept[c2] <= 0.0) {
        // legacy: no pressurized seepage face (mode 3), Corrector.cpp:96-106
        if (!pcLegacy && kt >= 0 && h(j, i, kt) > rZp(j, i, kt) * 0.5 * dz3d(j, i, kt)) {
          cplMode(j, i) = 3;
          continue;
        }
        cplMode(j, i) = 0;
        continue;
      }
      const double ks = ksz(j, i, kt);
      const double qEst = 2.0 * ks * (h(j, i, kt) - dept[c2]) / dz3d(j, i, kt) - ks;
      cplMode(j, i) = (qEst < 0.0 && -qEst * dtg > avail[c2]) ? 2 : 1;
    }
}

// --- Predictor.cpp:105-386 <PC, HARM> --------------------------------------
void Oracle::faceConductivity() {
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        if (harmonic)
          kx(j, i, k) = face_k_harmonic(soil(j, i, k), h(j, i, k), ksx(j, i, k), 1.0,
                                        soil(j, i + 1, k), h(j, i + 1, k), ksx(j, i + 1, k), 1.0);
        else
          kx(j, i, k) = 0.5 * (kh(j, i, k, ksx) + kh(j, i + 1, k, ksx));
      }
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        if (harmonic)
          ky(j, i, k) = face_k_harmonic(soil(j, i, k), h(j, i, k), ksy(j, i, k), 1.0,
                                        soil(j + 1, i, k), h(j + 1, i, k), ksy(j + 1, i, k), 1.0);
        else
          ky(j, i, k) = 0.5 * (kh(j, i, k, ksy) + kh(j + 1, i, k, ksy));
      }
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k <= nz; ++k) {
        double v;
        if (k == 0) {
          v = 0.0;
        } else if (k == nz) {
          v = (botCode(j, i) == FREHG_BC_NOFLUX) ? 0.0 : kh(j, i, nz - 1, ksz);
        } else if (g(j, i, k) >= 0 && g(j, i, k - 1) < 0) {
          v = kh(j, i, k, ksz);
        } else if (g(j, i, k - 1) < 0) {
          v = 0.0;
        } else if (harmonic) {
          v = face_k_harmonic(soil(j, i, k - 1), h(j, i, k - 1), ksz(j, i, k - 1), dz3d(j, i, k - 1),
                              soil(j, i, k), h(j, i, k), ksz(j, i, k), dz3d(j, i, k));
        } else {
          v = 0.5 * (kh(j, i, k - 1, ksz) + kh(j, i, k, ksz));
        }
        kzF(j, i, k) = v;
      }
  // domain-edge overrides (:236-268)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        auto apply = [&](F3& faces, int fj, int fi, int code, const F3& ks) {
          if (code == FREHG_BC_NOFLUX) faces(fj, fi, k) = 0.0;
          else if (isHead(code)) faces(fj, fi, k) = kh(j, i, k, ks);
        };
        if (west(i)) apply(kx, j, i - 1, sxm(j, i), ksx);
        if (east(i)) apply(kx, j, i, sxp(j, i), ksx);
        if (south(j)) apply(ky, j - 1, i, sym(j, i), ksy);
        if (north(j)) apply(ky, j, i, syp(j, i), ksy);
      }
  // top-face rule (:280-313)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const bool isTop = g(j, i, k) >= 0 && (k == 0 || g(j, i, k - 1) < 0);
        if (!isTop) continue;
        const double own = kh(j, i, k, ksz);
        if (coupled) {
          // a wet top at Ksz; legacy at 0.5 (Ksz + K(h)) (Predictor.cpp:293-302)
          const double wet = pcLegacy ? 0.5 * (ksz(j, i, k) + own) : ksz(j, i, k);
          kzF(j, i, k) = (dept[static_cast<size_t>(j) * (nx + 2) + i] > 0.0) ? wet : own;
          continue;
        }
        const int code = topCode(j, i);
        if (code == FREHG_BC_HEAD) kzF(j, i, k) = ksz(j, i, k);
        else if (code == FREHG_BC_NOFLUX) kzF(j, i, k) = 0.0;
        else kzF(j, i, k) = own;
      }
  if (!full3d) {  // (:318-343)
    auto unsat = [&](int jj, int ii, int k) { return g(jj, ii, k) >= 0 && wc(jj, ii, k) < wcs(jj, ii, k); };
    for (int j = 1; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i)
        for (int k = 0; k < nz; ++k)
          if (unsat(j, i, k) || unsat(j, i + 1, k)) kx(j, i, k) = 0.0;
    for (int j = 0; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        for (int k = 0; k < nz; ++k)
          if (unsat(j, i, k) || unsat(j + 1, i, k)) ky(j, i, k) = 0.0;
  }
  // seal faces next to inactive interior cells, last (:349-385)
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const bool a = !outX(i) && g(j, i, k) < 0;
        const bool b = !outX(i + 1) && g(j, i + 1, k) < 0;
        if (a || b) kx(j, i, k) = 0.0;
      }
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const bool a = !outY(j) && g(j, i, k) < 0;
        const bool b = !outY(j + 1) && g(j + 1, i, k) < 0;
        if (a || b) ky(j, i, k) = 0.0;
      }
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k <= nz; ++k) {
        const bool above = (k >= 1) && g(j, i, k - 1) < 0;
        const bool below = (k < nz) && g(j, i, k) < 0;
        if (above || below) kzF(j, i, k) = 0.0;
      }
}

// --- Predictor.cpp:400-698 <PC=true> ----------------------------------------
void Oracle::assemble(double dtg) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) ch(j, i, k) = capacity_of_h(soil(j, i, k), hn(j, i, k));
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const int64_t row = g(j, i, k);
        if (row < 0) continue;
        const size_t c2 = static_cast<size_t>(j) * (nx + 2) + i;
        const bool isTop = (k == 0) || (g(j, i, k - 1) < 0);
        const double volume = az * dz3d(j, i, k);
        double gxp = -kx(j, i, k) * dtg * rXp(j, i, k) * vXp(j, i, k) * cosx(j, i, k) / (dx * dx) *
                     (ax(j, i, k) * dx);
        double gxm = -kx(j, i - 1, k) * dtg * rXp(j, i - 1, k) * vXp(j, i - 1, k) *
                     cosx(j, i - 1, k) / (dx * dx) * (ax(j, i - 1, k) * dx);
        double gyp = -ky(j, i, k) * dtg * rYp(j, i, k) * vYp(j, i, k) * cosy(j, i, k) / (dy * dy) *
                     (ay(j, i, k) * dy);
        double gym = -ky(j - 1, i, k) * dtg * rYp(j - 1, i, k) * vYp(j - 1, i, k) *
                     cosy(j - 1, i, k) / (dy * dy) * (ay(j - 1, i, k) * dy);
        if (east(i) && isHead(sxp(j, i))) gxp *= 2.0;
        if (west(i) && isHead(sxm(j, i))) gxm *= 2.0;
        if (north(j) && isHead(syp(j, i))) gyp *= 2.0;
        if (south(j) && isHead(sym(j, i))) gym *= 2.0;
        double dzf = (k == nz - 1) ? 0.5 * dz3d(j, i, k) : 0.5 * (dz3d(j, i, k) + dz3d(j, i, k + 1));
        const double gzp = -kzF(j, i, k + 1) * dtg * rZp(j, i, k + 1) * vZp(j, i, k + 1) /
                           (dz3d(j, i, k) * dzf) * volume;
        dzf = isTop ? 0.5 * dz3d(j, i, k) : 0.5 * (dz3d(j, i, k) + dz3d(j, i, k - 1));
        const double gzm = -kzF(j, i, k) * dtg * rZp(j, i, k) * vZp(j, i, k) /
                           (dz3d(j, i, k) * dzf) * volume;

        const double storage = (ch(j, i, k) + ss * wcn(j, i, k) / wcs(j, i, k)) * rRho(j, i, k);
        double gct = storage * volume - (gxp + gxm + gyp + gym);
        if (east(i) && sxp(j, i) == FREHG_BC_FLUX) gct += gxp;
        if (west(i) && sxm(j, i) == FREHG_BC_FLUX) gct += gxm;
        if (north(j) && syp(j, i) == FREHG_BC_FLUX) gct += gyp;
        if (south(j) && sym(j, i) == FREHG_BC_FLUX) gct += gym;
        if (!(k == nz - 1 && botCode(j, i) != FREHG_BC_HEAD)) gct -= gzp;
        if (isTop) {
          bool dirichletTop = coupled ? (cplMode(j, i) == 1) : (topCode(j, i) == FREHG_BC_HEAD);
          if (!pcLegacy) dirichletTop = dirichletTop || (coupled && cplMode(j, i) == 3);
          if (dirichletTop) gct -= gzm;
        } else {
          gct -= gzm;
        }

        double b = storage * hn(j, i, k) * volume;
        b -= dtg * volume * kzF(j, i, k + 1) * rZp(j, i, k + 1) * rZp(j, i, k + 1) *
             vZp(j, i, k + 1) / dz3d(j, i, k);
        b += dtg * volume * kzF(j, i, k) * rZp(j, i, k) * rZp(j, i, k) * vZp(j, i, k) / dz3d(j, i, k);
        if (followTerrain) {
          // the along-layer gravity term: K v r^2 sin (consistent), K sin (legacy)
          const auto rv2 = [&](const F3& r, const F3& v, int jj, int ii) {
            return pcLegacy ? 1.0 : r(jj, ii, k) * r(jj, ii, k) * v(jj, ii, k);
          };
          double sgn = (bot3d(j, i + 1, k) > bot3d(j, i, k)) ? 1.0 : -1.0;
          b += sgn * dtg * ax(j, i, k) * kx(j, i, k) * rv2(rXp, vXp, j, i) * sinx(j, i, k);
          sgn = (bot3d(j, i, k) > bot3d(j, i - 1, k)) ? -1.0 : 1.0;
          b += sgn * dtg * ax(j, i - 1, k) * kx(j, i - 1, k) * rv2(rXp, vXp, j, i - 1) * sinx(j, i - 1, k);
          sgn = (bot3d(j + 1, i, k) > bot3d(j, i, k)) ? 1.0 : -1.0;
          b += sgn * dtg * ay(j, i, k) * ky(j, i, k) * rv2(rYp, vYp, j, i) * siny(j, i, k);
          sgn = (bot3d(j, i, k) > bot3d(j - 1, i, k)) ? -1.0 : 1.0;
          b += sgn * dtg * ay(j - 1, i, k) * ky(j - 1, i, k) * rv2(rYp, vYp, j - 1, i) * siny(j - 1, i, k);
        }
        if (east(i)) {
          if (sxp(j, i) == FREHG_BC_FLUX) b += rXp(j, i, k) * ax(j, i, k) * vxp(j, i) * dtg;
          else b -= gxp * hn(j, i + 1, k);
        }
        if (west(i)) {
          if (sxm(j, i) == FREHG_BC_FLUX) b += rXp(j, i - 1, k) * ax(j, i - 1, k) * vxm(j, i) * dtg;
          else b -= gxm * hn(j, i - 1, k);
        }
        if (north(j)) {
          if (syp(j, i) == FREHG_BC_FLUX) b += rYp(j, i, k) * ay(j, i, k) * vyp(j, i) * dtg;
          else b -= gyp * hn(j + 1, i, k);
        }
        if (south(j)) {
          if (sym(j, i) == FREHG_BC_FLUX) b += rYp(j - 1, i, k) * ay(j - 1, i, k) * vym(j, i) * dtg;
          else b -= gym * hn(j - 1, i, k);
        }
        if (k == nz - 1) {
          if (botCode(j, i) == FREHG_BC_FLUX)
            b += dtg * rZp(j, i, nz) * volume *
                 (botValue(j, i) + kzF(j, i, nz) * rZp(j, i, nz) * vZp(j, i, nz)) / dz3d(j, i, k);
          else if (botCode(j, i) == FREHG_BC_HEAD)
            b -= gzp * botValue(j, i);
        }
        if (isTop) {
          if (coupled) {
            const int mode = cplMode(j, i);
            if (mode == 1) {
              b -= gzm * dept[c2];
            } else if (mode == 3) {
              // Dirichlet ghost head 0: the leg sits in the diagonal only.
            } else if (mode == 2) {
              const double roomTop = std::fmax(0.0, (0.9999 * wcs(j, i, k) - wcn(j, i, k)) * dz3d(j, i, k));
              const double take = std::fmin(avail[c2], roomTop);
              const double qSupply = -take / dtg;
              b += -dtg * rZp(j, i, k) * volume * (qSupply + kzF(j, i, k) * rZp(j, i, k) * vZp(j, i, k)) /
                   dz3d(j, i, k);
            } else {
              double qtop = 0.0;
              if (etActive && etCol[c2]) {
                qtop = etQ[c2];  // the channel's capped flux, no guard (Predictor.cpp:660-664)
              } else if (topCode(j, i) == FREHG_BC_FLUX) {
                const double v = topValue(j, i);
                if ((v > 0.0 && wcn(j, i, k) < wcs(j, i, k)) || (v < 0.0 && wcn(j, i, k) > wcr(j, i, k)))
                  qtop = v;
              }
              b += -dtg * rZp(j, i, k) * volume * (qtop + kzF(j, i, k) * rZp(j, i, k) * vZp(j, i, k)) /
                   dz3d(j, i, k);
            }
          } else if (etActive && etCol[c2]) {  // Predictor.cpp:675-679
            b += -dtg * rZp(j, i, k) * volume * (etQ[c2] + kzF(j, i, k) * rZp(j, i, k) * vZp(j, i, k)) /
                 dz3d(j, i, k);
          } else if (topCode(j, i) == FREHG_BC_HEAD) {
            b -= gzm * topValue(j, i);
          } else if (topCode(j, i) == FREHG_BC_FLUX) {
            b += -dtg * rZp(j, i, k) * volume *
                 (topValue(j, i) + kzF(j, i, k) * rZp(j, i, k) * vZp(j, i, k)) / dz3d(j, i, k);
          }
        }
        const size_t base = 7 * static_cast<size_t>(row);
        vals[base + 0] = gct;
        vals[base + 1] = gxm;
        vals[base + 2] = gxp;
        vals[base + 3] = gym;
        vals[base + 4] = gyp;
        vals[base + 5] = gzm;
        vals[base + 6] = gzp;
        rhs[row] = b;
      }
  if (sinkActive) {  // Predictor.cpp:700-726, after the assembly
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        for (int k = 0; k < nz; ++k) {
          const int64_t row = g(j, i, k);
          if (row < 0 || sink(j, i, k) == 0.0) continue;
          rhs[row] -= dtg * rRho(j, i, k) * sink(j, i, k) * (az * dz3d(j, i, k));
        }
  }
}

// ---------------------------------------------------------------------------
// PCG on the COO system (replaces KSPCG + bjacobi/ICC(0); LinearSystem.cpp).
// The test is on the 2-norm of the residual, ||b - A x|| <= max(rtol ||b||,
// atol); the recursive residual drives iteration and the true residual is
// recomputed at convergence (restart if they disagree).  Dot products are
// summed in fixed chunks, so results do not depend on the thread count.
// ---------------------------------------------------------------------------
struct Csr7 {
  const std::vector<int64_t>* cols;
  const std::vector<double>* vals;
  int64_t n;
};

void spmv7(const Csr7& A, const double* x, double* y) {
  const int64_t* c = A.cols->data();
  const double* v = A.vals->data();
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int64_t r = 0; r < A.n; ++r) {
    double s = 0.0;
    for (int l = 0; l < 7; ++l) {
      const int64_t cc = c[7 * r + l];
      if (cc >= 0) s += v[7 * r + l] * x[cc];
    }
    y[r] = s;
  }
}

double dot(const double* a, const double* b, int64_t n) {
  const int64_t chunk = 4096;
  const int64_t nc = (n + chunk - 1) / chunk;
  std::vector<double> part(static_cast<size_t>(nc), 0.0);
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int64_t c = 0; c < nc; ++c) {
    double s = 0.0;
    const int64_t e = (c + 1) * chunk < n ? (c + 1) * chunk : n;
    for (int64_t i = c * chunk; i < e; ++i) s += a[i] * b[i];
    part[c] = s;
  }
  double s = 0.0;
  for (double p : part) s += p;
  return s;
}

struct Precond {
  int kind = FREHG_PC_JACOBI;
  int64_t n = 0;
  std::vector<double> dinv;
  // z-line: per column start/len and Thomas factors
  std::vector<int64_t> colStart;
  std::vector<int> colLen;
  std::vector<double> lower, upper, piv;
  // ICC(0): pivots d_i of M = (D + L) D^-1 (D + L^T)
  std::vector<double> icd;
  const std::vector<int64_t>* cols = nullptr;
  const std::vector<double>* vals = nullptr;

  void setup(const Oracle& o) {
    n = o.nActive;
    cols = &o.cols;
    vals = &o.vals;
    const double* v = o.vals.data();
    if (kind == FREHG_PC_JACOBI) {
      dinv.resize(n);
      for (int64_t r = 0; r < n; ++r) dinv[r] = 1.0 / v[7 * r];
    } else if (kind == FREHG_PC_ZLINE) {
      colStart.clear();
      colLen.clear();
      for (int64_t r = 0; r < n;) {
        const int j = o.rowJ[r], i = o.rowI[r];
        int64_t e = r;
        while (e < n && o.rowJ[e] == j && o.rowI[e] == i) ++e;
        colStart.push_back(r);
        colLen.push_back(static_cast<int>(e - r));
        r = e;
      }
      lower.assign(n, 0.0);
      upper.assign(n, 0.0);
      piv.assign(n, 0.0);
      for (size_t c = 0; c < colStart.size(); ++c) {
        const int64_t s = colStart[c];
        const int L = colLen[c];
        // Thomas LU of the column block: sub = zm(k), diag, sup = zp(k)
        double prevU = 0.0, prevPinv = 0.0;
        for (int q = 0; q < L; ++q) {
          const int64_t r = s + q;
          const double sub = q > 0 ? v[7 * r + 5] : 0.0;
          const double sup = q + 1 < L ? v[7 * r + 6] : 0.0;
          const double l = q > 0 ? sub * prevPinv : 0.0;
          const double d = v[7 * r] - l * prevU;
          lower[r] = l;
          piv[r] = 1.0 / d;
          upper[r] = sup;
          prevU = sup;
          prevPinv = piv[r];
        }
      }
    } else if (kind == FREHG_PC_ICC0) {
      icd.assign(n, 0.0);
      const int64_t* c = o.cols.data();
      // d_i = a_ii - sum_{l<i in pattern} a_il^2 / d_l  (zero-fill IC for 7-point order)
      for (int64_t r = 0; r < n; ++r) {
        double d = v[7 * r];
        for (int l = 1; l < 7; ++l) {
          const int64_t cc = c[7 * r + l];
          if (cc >= 0 && cc < r) d -= v[7 * r + l] * v[7 * r + l] / icd[cc];
        }
        if (!(d > 0.0)) fail(FREHG_EDIVERGED, "ICC(0) breakdown: nonpositive pivot");
        icd[r] = d;
      }
    }
  }

  void apply(const double* r, double* z) const {
    if (kind == FREHG_PC_JACOBI) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
      for (int64_t q = 0; q < n; ++q) z[q] = dinv[q] * r[q];
    } else if (kind == FREHG_PC_ZLINE) {
#pragma omp parallel for schedule(static) num_threads(g_threads)
      for (size_t c = 0; c < colStart.size(); ++c) {
        const int64_t s = colStart[c];
        const int L = colLen[c];
        double prev = 0.0;
        for (int q = 0; q < L; ++q) {
          const double y = r[s + q] - lower[s + q] * prev;
          z[s + q] = y;
          prev = y;
        }
        double next = 0.0;
        for (int q = L - 1; q >= 0; --q) {
          const double x = (z[s + q] - upper[s + q] * next) * piv[s + q];
          z[s + q] = x;
          next = x;
        }
      }
    } else {
      const int64_t* c = cols->data();
      const double* v = vals->data();
      std::vector<double> y(n);
      for (int64_t q = 0; q < n; ++q) {
        double s = r[q];
        for (int l = 1; l < 7; ++l) {
          const int64_t cc = c[7 * q + l];
          if (cc >= 0 && cc < q) s -= v[7 * q + l] * y[cc];
        }
        y[q] = s / icd[q];
      }
      // Backward solve with the transposed lower part: (I + D^-1 L^T) z = y,
      // L^T entry (q, u) = a_uq for u > q; with the mirrored 7-point legs
      // a_uq is row u's leg pointing back at q.
      for (int64_t q = n - 1; q >= 0; --q) {
        double s = 0.0;
        for (int l = 1; l < 7; ++l) {
          const int64_t u = c[7 * q + l];
          if (u > q) {
            // row u's leg toward q: xm<->xp, ym<->yp, zm<->zp
            const int back = (l % 2 == 1) ? l + 1 : l - 1;
            s += v[7 * u + back] * z[u];
          }
        }
        z[q] = y[q] - s / icd[q];
      }
    }
  }
};

frehg_gw_solve_stats pcg(const Csr7& A, const Precond& M, const double* b, double* x, double rtol,
                         double atol, int maxIt, bool x0Given) {
  const int64_t n = A.n;
  std::vector<double> r(n), z(n), p(n), q(n);
  if (!x0Given)
    for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
  spmv7(A, x, q.data());
  for (int64_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
  const double bnorm = std::sqrt(dot(b, b, n));
  const double tol = std::fmax(rtol * bnorm, atol);
  frehg_gw_solve_stats st{};
  st.rhs_norm = bnorm;
  double rnorm = std::sqrt(dot(r.data(), r.data(), n));
  int it = 0;
  bool restart = true;
  double rz = 0.0;
  while (true) {
    if (rnorm <= tol) {
      // confirm on the true residual
      spmv7(A, x, q.data());
      for (int64_t i = 0; i < n; ++i) q[i] = b[i] - q[i];
      const double tn = std::sqrt(dot(q.data(), q.data(), n));
      if (tn <= tol) {
        st.converged = 1;
        st.residual_norm = tn;
        break;
      }
      r = q;
      rnorm = tn;
      restart = true;
    }
    if (it >= maxIt) {
      st.converged = 0;
      st.residual_norm = rnorm;
      break;
    }
    M.apply(r.data(), z.data());
    const double rzNew = dot(r.data(), z.data(), n);
    if (restart) {
      p = z;
      restart = false;
    } else {
      const double beta = rzNew / rz;
#pragma omp parallel for schedule(static) num_threads(g_threads)
      for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    rz = rzNew;
    spmv7(A, p.data(), q.data());
    const double alpha = rz / dot(p.data(), q.data(), n);
#pragma omp parallel for schedule(static) num_threads(g_threads)
    for (int64_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
    }
    rnorm = std::sqrt(dot(r.data(), r.data(), n));
    ++it;
  }
  st.iterations = it;
  return st;
}

void Oracle::solve() {
  Csr7 A{&cols, &vals, nActive};
  Precond M;
  M.kind = pc;
  M.setup(*this);
  if (guessPrevious)
    for (int64_t r = 0; r < nActive; ++r) sol[r] = hn(rowJ[r], rowI[r], rowK[r]);
  const auto t0 = std::chrono::steady_clock::now();
  last = pcg(A, M, rhs.data(), sol.data(), rtol, atol, maxIt, guessPrevious);
  tSolve += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  itTotal += last.iterations;
  if (!last.converged)
    fail(FREHG_EDIVERGED, "gw linear solve did not converge in " + std::to_string(last.iterations) +
                              " iterations (residual " + std::to_string(last.residual_norm) + ")");
  // scatter (Predictor.cpp:745-754)
  for (int64_t r = 0; r < nActive; ++r) h(rowJ[r], rowI[r], rowK[r]) = sol[r];
}

// --- Corrector.cpp:135-533 <PC, HARM> -------------------------------------
void Oracle::fluxes(double dtg) {
  const auto faceKx = [&](int j, int i, int k, int jn, int in, const F3& ks, double la,
                          double lb) {
    // legacy: the minus cell's Ks enters both flanks (Corrector.cpp:224-236)
    if (pcLegacy)
      return 0.5 * (kh(j, i, k, ks) + k_of_h(soil(jn, in, k), h(jn, in, k), ks(j, i, k)));
    if (harmonic)
      return face_k_harmonic(soil(j, i, k), h(j, i, k), ks(j, i, k), la, soil(jn, in, k),
                             h(jn, in, k), ks(jn, in, k), lb);
    return 0.5 * (kh(j, i, k, ks) + kh(jn, in, k, ks));
  };
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = 1; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const bool westFace = outX(i);
        const bool eastFace = east(i);
        const double A = ax(j, i, k);
        double q = 0.0;
        if (westFace) {
          const double kf = kh(j, 1, k, ksx);
          q = A * kf * vXp(j, 0, k) * cosx(j, 0, k) * (h(j, 1, k) - h(j, 0, k)) / (0.5 * dx);
        } else {
          const double kc = kh(j, i, k, ksx);
          double kf, delta = dx;
          const int sgn = (bot3d(j, i + 1, k) > bot3d(j, i, k)) ? 1 : -1;
          if (eastFace) {
            kf = kc;
            delta = 0.5 * dx;
          } else {
            kf = faceKx(j, i, k, j, i + 1, ksx, 1.0, 1.0);
            if (g(j, i, k) < 0 || g(j, i + 1, k) < 0) kf = 0.0;
          }
          // the along-layer term carries r v (consistent) or neither (legacy)
          const double rvx = pcLegacy ? 1.0 : rXp(j, i, k) * vXp(j, i, k);
          q = A * (kf * vXp(j, i, k) * cosx(j, i, k) * (h(j, i + 1, k) - h(j, i, k)) / delta +
                   static_cast<double>(sgn) * kf * rvx * sinx(j, i, k));
          if (!full3d && ((g(j, i, k) >= 0 && wc(j, i, k) < wcs(j, i, k)) ||
                          (g(j, i + 1, k) >= 0 && wc(j, i + 1, k) < wcs(j, i + 1, k))))
            q = 0.0;
        }
        if (westFace) {
          const int code = sxm(j, 1);
          if (code == FREHG_BC_NOFLUX) q = 0.0;
          else if (code == FREHG_BC_FLUX) q = -vxm(j, 1) * A;
        } else if (eastFace) {
          const int code = sxp(j, i);
          if (code == FREHG_BC_NOFLUX) q = 0.0;
          else if (code == FREHG_BC_FLUX) q = vxp(j, i) * A;
        }
        qx(j, i, k) = q;
      }
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = 0; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const bool southFace = outY(j);
        const bool northFace = north(j);
        const double A = ay(j, i, k);
        double q = 0.0;
        if (southFace) {
          const double kf = kh(1, i, k, ksy);
          q = A * kf * vYp(0, i, k) * cosy(0, i, k) * (h(1, i, k) - h(0, i, k)) / (0.5 * dy);
        } else {
          const double kc = kh(j, i, k, ksy);
          double kf, delta = dy;
          const int sgn = (bot3d(j + 1, i, k) > bot3d(j, i, k)) ? 1 : -1;
          if (northFace) {
            kf = kc;
            delta = 0.5 * dy;
          } else {
            kf = faceKx(j, i, k, j + 1, i, ksy, 1.0, 1.0);
            if (g(j, i, k) < 0 || g(j + 1, i, k) < 0) kf = 0.0;
          }
          const double rvy = pcLegacy ? 1.0 : rYp(j, i, k) * vYp(j, i, k);
          q = A * (kf * vYp(j, i, k) * cosy(j, i, k) * (h(j + 1, i, k) - h(j, i, k)) / delta +
                   static_cast<double>(sgn) * kf * rvy * siny(j, i, k));
          if (!full3d && ((g(j, i, k) >= 0 && wc(j, i, k) < wcs(j, i, k)) ||
                          (g(j + 1, i, k) >= 0 && wc(j + 1, i, k) < wcs(j + 1, i, k))))
            q = 0.0;
        }
        if (southFace) {
          const int code = sym(1, i);
          if (code == FREHG_BC_NOFLUX) q = 0.0;
          else if (code == FREHG_BC_FLUX) q = -vym(1, i) * A;
        } else if (northFace) {
          const int code = syp(j, i);
          if (code == FREHG_BC_NOFLUX) q = 0.0;
          else if (code == FREHG_BC_FLUX) q = vyp(j, i) * A;
        }
        qy(j, i, k) = q;
      }
  double clamp = 0.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      const size_t c2 = static_cast<size_t>(j) * (nx + 2) + i;
      for (int k = 0; k <= nz; ++k) {
        const bool below = (k < nz) && g(j, i, k) >= 0;
        const bool above = (k >= 1) && g(j, i, k - 1) >= 0;
        double q = 0.0;
        if (below && !above) {
          if (coupled) {
            const double ks = ksz(j, i, k);
            const Soil s = soil(j, i, k);
            const double depth = dept[c2];
            const int mode = cplMode(j, i);
            if (mode == 2) {
              const double roomTop = std::fmax(0.0, (0.9999 * wcs(j, i, k) - wc(j, i, k)) * dz3d(j, i, k));
              const double take = std::fmin(avail[c2], roomTop);
              q = -take * az / dtg;
              avail[c2] -= take;
            } else if (mode == 1) {
              const double delta = 0.5 * dz3d(j, i, k);
              q = az * ks * vZp(j, i, k) * ((h(j, i, k) - depth) / delta - rZp(j, i, k));
              if (q < 0.0) {
                const double need = -q * dtg / az;
                if (need > avail[c2]) {
                  q = -avail[c2] * az / dtg;
                  avail[c2] = 0.0;
                } else {
                  avail[c2] -= need;
                }
              }
            } else if (mode == 3) {
              const double delta = 0.5 * dz3d(j, i, k);
              q = az * ks * vZp(j, i, k) * ((h(j, i, k) - depth) / delta - rZp(j, i, k));
              if (q <= 0.0) {
                clamp += -q * dtg;
                q = 0.0;
              }
            } else {
              const double kG = k_of_h(s, depth, ks);
              const double kC = k_of_h(s, h(j, i, k), ks);
              const double kf = 0.5 * (kG + kC);
              const double delta = 0.5 * dz3d(j, i, k);
              q = az * kf * vZp(j, i, k) * ((h(j, i, k) - depth) / delta - rZp(j, i, k));
              if (q <= 0.0) q = 0.0;
              if (etActive && etCol[c2]) {
                q += etQ[c2] * az;  // Corrector.cpp:458-462
              } else if (topCode(j, i) == FREHG_BC_FLUX) {
                const double v = topValue(j, i);
                if ((v > 0.0 && wc(j, i, k) > wcr(j, i, k)) || v < 0.0) q += v * az;
              }
            }
          } else {
            const int code = topCode(j, i);
            if (etActive && etCol[c2]) {
              q = etQ[c2] * az;  // Corrector.cpp:471-472
            } else if (code == FREHG_BC_FLUX) {
              q = topValue(j, i) * az;
            } else if (code == FREHG_BC_HEAD) {
              const double delta = 0.5 * dz3d(j, i, k);
              q = az * ksz(j, i, k) * vZp(j, i, k) * ((h(j, i, k) - topValue(j, i)) / delta - rZp(j, i, k));
            } else {
              q = 0.0;
            }
          }
        } else if (k == nz && above) {
          const int code = botCode(j, i);
          const double kf = kh(j, i, k - 1, ksz);
          if (code == FREHG_BC_FLUX) q = botValue(j, i) * az;
          else if (code == FREHG_BC_GRAVITY) q = -kf * vZp(j, i, k) * rZp(j, i, k) * az;
          else if (code == FREHG_BC_HEAD) {
            const double delta = 0.5 * dz3d(j, i, k - 1);
            q = az * kf * vZp(j, i, k) * ((botValue(j, i) - h(j, i, k - 1)) / delta - rZp(j, i, k));
          } else q = 0.0;
        } else if (above && below) {
          double kf;
          if (pcLegacy)  // the lower cell's Ks in both flanks (Corrector.cpp:512-521)
            kf = 0.5 * (k_of_h(soil(j, i, k - 1), h(j, i, k - 1), ksz(j, i, k)) + kh(j, i, k, ksz));
          else if (harmonic)
            kf = face_k_harmonic(soil(j, i, k - 1), h(j, i, k - 1), ksz(j, i, k - 1), dz3d(j, i, k - 1),
                                 soil(j, i, k), h(j, i, k), ksz(j, i, k), dz3d(j, i, k));
          else
            kf = 0.5 * (kh(j, i, k - 1, ksz) + kh(j, i, k, ksz));
          const double delta = 0.5 * (dz3d(j, i, k - 1) + dz3d(j, i, k));
          q = az * kf * vZp(j, i, k) * ((h(j, i, k) - h(j, i, k - 1)) / delta - rZp(j, i, k));
        }
        qzF(j, i, k) = q;
      }
    }
  audit.pc_clamp = clamp;
  pcClampTotal += clamp;
}

// --- Corrector.cpp:539-627 ---------------------------------------------------
void Oracle::coupledBookkeeping(double dtg) {
  double ex = 0.0, bo = 0.0, ev = 0.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) {
      int kt = -1;
      for (int k = 0; k < nz; ++k)
        if (g(j, i, k) >= 0) {
          kt = k;
          break;
        }
      if (kt < 0) continue;
      const size_t c2 = static_cast<size_t>(j) * (nx + 2) + i;
      const double q = qzF(j, i, kt);
      ex += q * dtg;
      if (q < 0.0 && dept[c2] <= minDepth) {
        const double volume = az * dz3d(j, i, kt);
        if ((wcs(j, i, kt) - wc(j, i, kt)) < -q * dtg / volume) {
          const double d = -q * dtg / az;
          eta[c2] += d;
          dept[c2] += d;
          avail[c2] += d;
          gain[c2] += d;
          if (bounce) bounce[c2] += d;
          bo += d * az;
        }
      }
      seep[c2] += q * dtg / az;
      if (etActive && etCol[c2]) {  // the channel's exact volume (Corrector.cpp:609-615)
        seep[c2] -= etQ[c2] * dtg;
        ev += etQ[c2] * dtg * az;
      } else if (topCode(j, i) == FREHG_BC_FLUX && topValue(j, i) > 0.0 && wc(j, i, kt) > wcr(j, i, kt) &&
                 dept[c2] <= minDepth) {
        seep[c2] -= topValue(j, i) * dtg;
        ev += topValue(j, i) * dtg * az;
      }
    }
  audit.cpl_exchanged = ex;
  audit.cpl_bounce = bo;
  audit.cpl_evap = ev;
}

// --- Corrector.cpp:629-672 ---------------------------------------------------
void Oracle::updateWaterContent(double dtg) {
  audit.ss_storage = sum_rows(1, ny + 1, [&](int j) {
    double s = 0.0;
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        if (g(j, i, k) < 0) continue;
        const double volume = az * dz3d(j, i, k);
        const double r = rRho(j, i, k);
        const double coeff = volume * (r + r * ss * (h(j, i, k) - hn(j, i, k)) / wcs(j, i, k));
        const double dqx = dtg * (qx(j, i, k) * rXp(j, i, k) - qx(j, i - 1, k) * rXp(j, i - 1, k));
        const double dqy = dtg * (qy(j, i, k) * rYp(j, i, k) - qy(j - 1, i, k) * rYp(j - 1, i, k));
        const double dqz = dtg * (qzF(j, i, k + 1) * rZp(j, i, k + 1) - qzF(j, i, k) * rZp(j, i, k));
        const double upd = (wcn(j, i, k) * r * volume + dqx + dqy + dqz) / coeff;
        wc(j, i, k) = upd;
        s += upd * volume * r * ss * (h(j, i, k) - hn(j, i, k)) / wcs(j, i, k);
      }
    return s;
  });
  if (!sinkActive) return;
  // the root-uptake sink after the flux update (Corrector.cpp:674-705); the
  // applied volume S V dtg is the vegetation module's trApplied
  // (Evapotranspiration.cpp:619-633)
  double ssDelta = 0.0, applied = 0.0;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        if (g(j, i, k) < 0 || sink(j, i, k) == 0.0) continue;
        const double volume = az * dz3d(j, i, k);
        const double r = rRho(j, i, k);
        const double coeff = volume * (r + r * ss * (h(j, i, k) - hn(j, i, k)) / wcs(j, i, k));
        const double dTheta = dtg * r * sink(j, i, k) * volume / coeff;
        wc(j, i, k) -= dTheta;
        ssDelta -= dTheta * volume * r * ss * (h(j, i, k) - hn(j, i, k)) / wcs(j, i, k);
        applied += sink(j, i, k) * volume * dtg;
      }
  audit.ss_storage += ssDelta;
  audit.tr_applied = applied;
}

// --- Reallocate.cpp:122-545, saturated_contact: head ------------------------
void Oracle::reallocate() {
  const double dtg = dtgCurrent;
  wcSnap = wc;
  hSnap = h;
  const F3& wR = wcSnap;
  const F3& hR = hSnap;
  struct Split {
    double up, down;
  };
  const auto gradientSplit = [&](int jc, int ic, int kc) -> Split {
    double d6[6];
    const auto act = [&](int jj, int ii, int kk) { return kk >= 0 && kk < nz && g(jj, ii, kk) >= 0; };
    d6[0] = (act(jc, ic + 1, kc) && kx(jc, ic, kc) > 0.0) ? (hR(jc, ic + 1, kc) - hR(jc, ic, kc)) / dx : 0.0;
    d6[1] = (act(jc, ic - 1, kc) && kx(jc, ic - 1, kc) > 0.0) ? (hR(jc, ic, kc) - hR(jc, ic - 1, kc)) / dx : 0.0;
    d6[2] = (act(jc + 1, ic, kc) && ky(jc, ic, kc) > 0.0) ? (hR(jc + 1, ic, kc) - hR(jc, ic, kc)) / dy : 0.0;
    d6[3] = (act(jc - 1, ic, kc) && ky(jc - 1, ic, kc) > 0.0) ? (hR(jc, ic, kc) - hR(jc - 1, ic, kc)) / dy : 0.0;
    if (act(jc, ic, kc + 1) && kzF(jc, ic, kc + 1) > 0.0) {
      const double dzp = 0.5 * (dz3d(jc, ic, kc) + dz3d(jc, ic, kc + 1));
      d6[4] = (hR(jc, ic, kc + 1) - hR(jc, ic, kc)) / dzp - (densityGradient ? rZp(jc, ic, kc + 1) : 1.0);
    } else {
      d6[4] = 0.0;
    }
    if (act(jc, ic, kc - 1) && kzF(jc, ic, kc) > 0.0) {
      const double dzm = 0.5 * (dz3d(jc, ic, kc) + dz3d(jc, ic, kc - 1));
      d6[5] = (hR(jc, ic, kc) - hR(jc, ic, kc - 1)) / dzm - (densityGradient ? rZp(jc, ic, kc) : 1.0);
    } else {
      d6[5] = 0.0;
    }
    if (!full3d) d6[0] = d6[1] = d6[2] = d6[3] = 0.0;
    double tot = 0.0, gx, gy, gz;
    if (d6[0] * d6[1] >= 0.0) {
      gx = 0.5 * (d6[0] + d6[1]);
      tot += std::fabs(gx);
    } else {
      gx = 0.0;
      tot += std::fabs(d6[0]) + std::fabs(d6[1]);
    }
    if (d6[2] * d6[3] >= 0.0) {
      gy = 0.5 * (d6[2] + d6[3]);
      tot += std::fabs(gy);
    } else {
      gy = 0.0;
      tot += std::fabs(d6[2]) + std::fabs(d6[3]);
    }
    if (d6[4] * d6[5] >= 0.0) {
      gz = 0.5 * (d6[4] + d6[5]);
      tot += std::fabs(gz);
    } else {
      gz = 0.0;
      tot += std::fabs(d6[4]) + std::fabs(d6[5]);
    }
    (void)gx;
    (void)gy;
    Split sp{0.0, 0.0};
    if (tot > 0.0) {
      if (gz > 0.0) sp.up = gz / tot;
      else if (gz < 0.0) sp.down = -gz / tot;
      else {
        sp.down = std::fabs(d6[4]) / tot;
        sp.up = std::fabs(d6[5]) / tot;
      }
    }
    return sp;
  };

  std::vector<double> adjRow(ny + 2, 0.0), dropRow(ny + 2, 0.0);
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = 1; j <= ny; ++j) {
    double sum = 0.0, drop = 0.0;
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        if (g(j, i, k) < 0) {
          room(j, i, k) = 0.0;
          sendUp(j, i, k) = 0.0;
          sendDown(j, i, k) = 0.0;
          continue;
        }
        const Soil s = soil(j, i, k);
        const double volume = az * dz3d(j, i, k);
        const double before = wc(j, i, k);
        sendUp(j, i, k) = 0.0;
        sendDown(j, i, k) = 0.0;
        if (before >= s.ths) {
          if (reallocOff) {
            room(j, i, k) = 0.0;
            continue;
          }
          const double excess = (before - s.ths) * volume;
          const Split sp = gradientSplit(j, i, k);
          sendUp(j, i, k) = excess * sp.up;
          sendDown(j, i, k) = excess * sp.down;
          sum -= excess - sendUp(j, i, k) - sendDown(j, i, k);
          drop += excess - sendUp(j, i, k) - sendDown(j, i, k);
          wc(j, i, k) = s.ths;
          room(j, i, k) = 0.0;
          continue;
        }
        room(j, i, k) = (s.ths - before) * volume;
        const auto sat = [&](int jj, int ii, int kk) {
          return wR(jj, ii, kk) >= wcs(jj, ii, kk) || hR(jj, ii, kk) > aev(jj, ii, kk);
        };
        bool adj = false;
        if (k + 1 < nz && g(j, i, k + 1) >= 0 && sat(j, i, k + 1)) adj = true;
        else if (kx(j, i, k) != 0.0 && sat(j, i + 1, k)) adj = true;
        else if (kx(j, i - 1, k) != 0.0 && sat(j, i - 1, k)) adj = true;
        else if (ky(j, i, k) != 0.0 && sat(j + 1, i, k)) adj = true;
        else if (ky(j - 1, i, k) != 0.0 && sat(j - 1, i, k)) adj = true;
        else {
          const bool isTop = (k == 0) || (g(j, i, k - 1) < 0);
          if (isTop) {
            if (coupled) {
              if (dept[static_cast<size_t>(j) * (nx + 2) + i] > 0.0) adj = true;
            } else if (topCode(j, i) == FREHG_BC_HEAD && topValue(j, i) >= 0.0) {
              adj = true;
            }
          } else if (g(j, i, k - 1) >= 0 && sat(j, i, k - 1)) {
            adj = true;
          }
        }
        if (!adj) {
          if (before < 0.9999 * s.ths) h(j, i, k) = h_of_theta(s, before);
        } else if (reallocOff) {
          continue;
        } else {
          const double after = theta_of_h(s, h(j, i, k));
          wc(j, i, k) = after;
          room(j, i, k) = (s.ths - after) * volume;
          if (redistribute && after < before) {
            const double surplus = (before - after) * volume;
            const Split sp = gradientSplit(j, i, k);
            sendUp(j, i, k) = surplus * sp.up;
            sendDown(j, i, k) = surplus * sp.down;
            sum -= surplus - sendUp(j, i, k) - sendDown(j, i, k);
            drop += surplus - sendUp(j, i, k) - sendDown(j, i, k);
          } else {
            sum += (after - before) * volume;
            if (after < before) drop += (before - after) * volume;
          }
        }
      }
    adjRow[j] = sum;
    dropRow[j] = drop;
  }
  double adjusted = 0.0, discarded = 0.0;
  for (int j = 1; j <= ny; ++j) {
    adjusted += adjRow[j];
    discarded += dropRow[j];
  }
  if (reallocOff) {
    audit.realloc_adjust = 0.0;
    audit.realloc_dropped = 0.0;
    audit.cpl_vent = 0.0;
    audit.cpl_vent_thin = 0.0;
    return;
  }
  // Phase 2: column send walks (Reallocate.cpp:410-540)
  std::vector<double> lostRow(ny + 2, 0.0), ventRow(ny + 2, 0.0), thinRow(ny + 2, 0.0);
#pragma omp parallel for schedule(static) num_threads(g_threads)
  for (int j = 1; j <= ny; ++j) {
    double lost = 0.0, vent = 0.0, thin = 0.0;
    for (int i = 1; i <= nx; ++i) {
      const size_t c2 = static_cast<size_t>(j) * (nx + 2) + i;
      for (int k = 0; k < nz; ++k) {
        if (g(j, i, k) < 0) continue;
        const bool topIsHead = topCode(j, i) == FREHG_BC_HEAD;
        // Path recording (transport window only, Reallocate.cpp:425-445):
        // the legacy mirror of a transfer from cell `from` to cell `to`
        // (the cell above the column top for a vent), recorded at face
        // `legacyFace`, is replaced by the flux on every face crossed.
        const auto record = [&](int from, int to, double vol, int legacyFace, double legacySign) {
          if (!transportWindow) return;
          const double rate = vol / dtg;
          pathCorr(j, i, legacyFace) -= legacySign * rate;
          if (to < from) {
            for (int f = to + 1; f <= from; ++f) pathCorr(j, i, f) += rate;
          } else {
            for (int f = from + 1; f <= to; ++f) pathCorr(j, i, f) -= rate;
          }
        };
        double dV = sendUp(j, i, k);
        if (dV > 0.0) {
          int ll = k;
          while (ll > 0 && g(j, i, ll - 1) >= 0 && dV > 0.0) {
            --ll;
            if (room(j, i, ll) > 0.0) {
              const double take = (room(j, i, ll) > dV) ? dV : room(j, i, ll);
              wc(j, i, ll) += take / (az * dz3d(j, i, ll));
              room(j, i, ll) -= take;
              qzF(j, i, ll + 1) += take / dtg;
              record(k, ll, take, ll + 1, 1.0);
              dV -= take;
            }
          }
          if (dV > 0.0) {
            if (coupled) {
              const double d = dV / az;
              if (ventAll || dept[c2] > 0.0 || d > minDepth) {
                if (!(dept[c2] > 0.0 || d > minDepth)) thin += dV;
                eta[c2] += d;
                dept[c2] += d;
                avail[c2] += d;
                gain[c2] += d;
                qzF(j, i, ll) += dV / dtg;
                record(k, ll - 1, dV, ll, 1.0);
                if (transportWindow) winVent(j, i) += dV;
                vent += dV;
              } else {
                lost -= dV;
              }
              dV = 0.0;
            } else if (topIsHead) {
              qzF(j, i, ll) += dV / dtg;
              record(k, ll - 1, dV, ll, 1.0);
              dV = 0.0;
            } else {
              while (ll + 1 < nz && dV > 0.0) {
                ++ll;
                if (room(j, i, ll) > 0.0) {
                  const double take = (room(j, i, ll) > dV) ? dV : room(j, i, ll);
                  wc(j, i, ll) += take / (az * dz3d(j, i, ll));
                  room(j, i, ll) -= take;
                  qzF(j, i, ll) -= take / dtg;
                  record(k, ll, take, ll, -1.0);
                  dV -= take;
                }
              }
              lost -= dV;
              dV = 0.0;
            }
          }
        }
        dV = sendDown(j, i, k);
        if (dV > 0.0) {
          int ll = k;
          while (ll + 1 < nz && dV > 0.0) {
            ++ll;
            if (room(j, i, ll) > 0.0) {
              const double take = (room(j, i, ll) > dV) ? dV : room(j, i, ll);
              wc(j, i, ll) += take / (az * dz3d(j, i, ll));
              room(j, i, ll) -= take;
              qzF(j, i, ll) -= take / dtg;
              record(k, ll, take, ll, -1.0);
              dV -= take;
            }
          }
          lost -= dV;
        }
      }
    }
    lostRow[j] = lost;
    ventRow[j] = vent;
    thinRow[j] = thin;
  }
  double dropped = 0.0, vented = 0.0, thin = 0.0;
  for (int j = 1; j <= ny; ++j) {
    dropped += lostRow[j];
    vented += ventRow[j];
    thin += thinRow[j];
  }
  audit.realloc_adjust = adjusted + dropped;
  audit.realloc_dropped = discarded - dropped;
  audit.cpl_vent = vented;
  audit.cpl_vent_thin = thin;
}

// --- RichardsSolver.cpp:796-878: the transport window --------------------------
void Oracle::enableTransportWindow() {
  if (transportWindow) return;
  transportWindow = true;
  winX = qx;
  winY = qy;
  winSink = qx;
  winZ = qzF;
  pathCorr = qzF;
  winVent = F2<double>();
  winVent.init(ny + 2, nx + 2, 0.0);
  winVap = winVent;
  resetTransportWindow();
  std::fill(pathCorr.v.begin(), pathCorr.v.end(), 0.0);
}

void Oracle::resetTransportWindow() {
  if (!transportWindow) return;
  for (F3* f : {&winX, &winY, &winZ, &winSink}) std::fill(f->v.begin(), f->v.end(), 0.0);
  std::fill(winVent.v.begin(), winVent.v.end(), 0.0);
  std::fill(winVap.v.begin(), winVap.v.end(), 0.0);
}

// Every face slot, halos included (RichardsSolver.cpp:826-878).
void Oracle::accumulateTransportWindow(double dtg) {
  for (size_t t = 0; t < winX.v.size(); ++t) {
    winX.v[t] += qx.v[t] * dtg;
    winY.v[t] += qy.v[t] * dtg;
  }
  for (size_t t = 0; t < winZ.v.size(); ++t) winZ.v[t] += (qzF.v[t] + pathCorr.v[t]) * dtg;
  if (etActive)
    for (int j = 0; j < ny + 2; ++j)
      for (int i = 0; i < nx + 2; ++i)
        winVap(j, i) += etQ[static_cast<size_t>(j) * (nx + 2) + i] * az * dtg;
  if (sinkActive)
    for (int j = 0; j < ny + 2; ++j)
      for (int i = 0; i < nx + 2; ++i)
        for (int k = 0; k < nz; ++k)
          winSink(j, i, k) += sink(j, i, k) * az * dz3d(j, i, k) * dtg;
}

// --- Corrector.cpp:758-792 ---------------------------------------------------
void Oracle::finalize() {
  audit.vloss = sum_rows(1, ny + 1, [&](int j) {
    double s = 0.0;
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        if (g(j, i, k) < 0) continue;
        const double volume = az * dz3d(j, i, k);
        if (wc(j, i, k) > wcs(j, i, k)) {
          const double ex = (wc(j, i, k) - wcs(j, i, k)) * volume;
          vloss(j, i, k) += ex;
          s += ex;
          wc(j, i, k) = wcs(j, i, k);
        } else if (wc(j, i, k) < wcr(j, i, k)) {
          const double de = (wcr(j, i, k) - wc(j, i, k)) * volume;
          vloss(j, i, k) -= de;
          s -= de;
          wc(j, i, k) = wcr(j, i, k);
        }
      }
    return s;
  });
}

// --- RichardsSolver.cpp:961-1006 ---------------------------------------------
void Oracle::boundaryFlux(double dtg) {
  audit.boundary_in = sum_rows(1, ny + 1, [&](int j) {
    double s = 0.0;
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        if (g(j, i, k) < 0) continue;
        const bool top = (k == 0) || (g(j, i, k - 1) < 0);
        if (top) s -= qzF(j, i, k) * dtg;
        if (k == nz - 1) s += qzF(j, i, k + 1) * dtg;
        if (west(i)) s -= qx(j, i - 1, k) * dtg;
        if (east(i)) s += qx(j, i, k) * dtg;
        if (south(j)) s -= qy(j - 1, i, k) * dtg;
        if (north(j)) s += qy(j, i, k) * dtg;
      }
    return s;
  });
}

// --- AdaptiveStep.cpp:44-118 -------------------------------------------------
void Oracle::adaptTimeStep(double dtg) {
  double dqMax = 0.0;
  double dtCoMin = 1.0e8;
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k) {
        const bool isTop = (k == 0) || (g(j, i, k - 1) < 0);
        if (g(j, i, k) < 0 || isTop) continue;
        const double qin = qx(j, i, k) / ax(j, i, k) + qy(j, i, k) / ay(j, i, k) + qzF(j, i, k + 1) / az;
        const double qou = qx(j, i - 1, k) / ax(j, i, k) + qy(j - 1, i, k) / ay(j, i, k) + qzF(j, i, k) / az;
        const double dq = std::fabs(qin - qou) * dtg / dz3d(j, i, k);
        if (dq > dqMax) dqMax = dq;
        if (wc(j, i, k) < wcs(j, i, k)) {
          const double dk = dk_dtheta(soil(j, i, k), wc(j, i, k), ksz(j, i, k));
          const double dtCo = courantMax * dz3d(j, i, k) / dk;
          if (dtCo < dtCoMin) dtCoMin = dtCo;
        }
      }
  double next = dtg;
  if (dqMax > dqShrink) next = dtg * 0.75;
  else if (dqMax < dqGrow) next = dtg * 1.25;
  if (next > dtCoMin) next = dtCoMin;
  if (next > dtMax) next = dtMax;
  if (next < dtMin) next = dtMin;
  dtgNext = next;
}

double Oracle::ownedVolume() const {
  return sum_rows(1, ny + 1, [&](int j) {
    double s = 0.0;
    for (int i = 1; i <= nx; ++i)
      for (int k = 0; k < nz; ++k)
        if (g(j, i, k) >= 0) s += wc(j, i, k) * az * dz3d(j, i, k);
    return s;
  });
}

// --- RichardsSolver::step (RichardsSolver.cpp:729-794) ------------------------
void Oracle::step(double /*t*/, double dtg) {
  if (!(dtg > 0.0)) fail(FREHG_EINVAL, "dtg must be positive");
  const auto t0 = std::chrono::steady_clock::now();
  audit = frehg_gw_audit{};
  dtgCurrent = dtg;
  enforceHeadBc();
  enforceMoistureBc();
  hn = h;
  wcn = wc;
  if (coupled) classifyCoupledTop(dtg);
  faceConductivity();
  assemble(dtg);
  solve();
  enforceHeadBc();
  faceConductivity();
  fluxes(dtg);
  if (coupled) coupledBookkeeping(dtg);
  updateWaterContent(dtg);
  if (transportWindow) std::fill(pathCorr.v.begin(), pathCorr.v.end(), 0.0);
  reallocate();
  finalize();
  enforceMoistureBc();
  boundaryFlux(dtg);
  if (transportWindow) accumulateTransportWindow(dtg);
  adaptTimeStep(dtg);
  tStep += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------------------
// Generic COO system (the LinearSystem seam), CSR with merged duplicates.
// ---------------------------------------------------------------------------
struct OracleLs {
  int64_t n = 0;
  double rtol = 1e-8, atol = 1e-14;
  int maxIt = 500, pc = FREHG_PC_JACOBI;
  std::vector<int64_t> rowPtr, colIdx, cooSlot;  // cooSlot: COO entry -> CSR slot (-1 ignored)
  std::vector<double> val;
  bool pattern = false;
};

template <class Fn>
int guard(Fn fn) {
  try {
    fn();
    return FREHG_OK;
  } catch (const OracleError& e) {
    g_error = e.what();
    return e.code;
  } catch (const std::exception& e) {
    g_error = e.what();
    return FREHG_EINVAL;
  }
}

const F3* field_of(Oracle& o, int f) {
  switch (f) {
    case FREHG_F_HEAD: return &o.h;
    case FREHG_F_WC: return &o.wc;
    case FREHG_F_HEAD_N: return &o.hn;
    case FREHG_F_WC_N: return &o.wcn;
    case FREHG_F_KX: return &o.kx;
    case FREHG_F_KY: return &o.ky;
    case FREHG_F_KZF: return &o.kzF;
    case FREHG_F_QX: return &o.qx;
    case FREHG_F_QY: return &o.qy;
    case FREHG_F_QZF: return &o.qzF;
    case FREHG_F_VLOSS: return &o.vloss;
    case FREHG_F_DZ3D: return &o.dz3d;
    case FREHG_F_BOT3D: return &o.bot3d;
    case FREHG_F_AREA_X: return &o.ax;
    case FREHG_F_AREA_Y: return &o.ay;
    case FREHG_F_SIN_X: return &o.sinx;
    case FREHG_F_COS_X: return &o.cosx;
    case FREHG_F_SIN_Y: return &o.siny;
    case FREHG_F_COS_Y: return &o.cosy;
    case FREHG_F_CAPACITY: return &o.ch;
    case FREHG_F_SINK:
      if (!o.sinkActive) fail(FREHG_EINVAL, "no root-uptake sink is set");
      return &o.sink;
    case FREHG_F_WIN_X:
    case FREHG_F_WIN_Y:
    case FREHG_F_WIN_Z:
    case FREHG_F_WIN_SINK:
    case FREHG_F_QZ_PATH_CORR:
      if (!o.transportWindow) fail(FREHG_EINVAL, "the transport window is not enabled");
      return f == FREHG_F_WIN_X   ? &o.winX
             : f == FREHG_F_WIN_Y ? &o.winY
             : f == FREHG_F_WIN_Z ? &o.winZ
             : f == FREHG_F_WIN_SINK ? &o.winSink
                                     : &o.pathCorr;
    default: fail(FREHG_EINVAL, "unknown field id");
  }
}

// Data and element count of any field, the 2D window planes included.
const double* field_data(Oracle& o, int f, size_t* n) {
  if (f == FREHG_F_WIN_VENT || f == FREHG_F_WIN_VAP) {
    if (!o.transportWindow) fail(FREHG_EINVAL, "the transport window is not enabled");
    const F2<double>& p = f == FREHG_F_WIN_VENT ? o.winVent : o.winVap;
    *n = p.v.size();
    return p.v.data();
  }
  const F3* fld = field_of(o, f);
  *n = fld->v.size();
  return fld->v.data();
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_error.c_str(); }
void oracle_set_threads(int n) { g_threads = n > 0 ? n : 1; }
int oracle_get_threads(void) { return g_threads; }

double oracle_vg_theta(double a, double n, double ts, double tr, double aev, double h) {
  return theta_of_h(Soil{a, n, ts, tr, aev}, h);
}
double oracle_vg_head(double a, double n, double ts, double tr, double aev, double wc) {
  return h_of_theta(Soil{a, n, ts, tr, aev}, wc);
}
double oracle_vg_capacity(double a, double n, double ts, double tr, double aev, double h) {
  return capacity_of_h(Soil{a, n, ts, tr, aev}, h);
}
double oracle_vg_conductivity(double a, double n, double ts, double tr, double aev, double h,
                              double ks) {
  return k_of_h(Soil{a, n, ts, tr, aev}, h, ks);
}
double oracle_vg_dkdtheta(double a, double n, double ts, double tr, double aev, double wc,
                          double ks) {
  return dk_dtheta(Soil{a, n, ts, tr, aev}, wc, ks);
}
double oracle_face_k_harmonic(double a1, double n1, double ts1, double tr1, double aev1, double h1,
                              double ks1, double l1, double a2, double n2, double ts2, double tr2,
                              double aev2, double h2, double ks2, double l2) {
  return face_k_harmonic(Soil{a1, n1, ts1, tr1, aev1}, h1, ks1, l1, Soil{a2, n2, ts2, tr2, aev2},
                         h2, ks2, l2);
}

int oracle_gw_create(const frehg_gw_desc* d, void** out) {
  return guard([&] {
    if (!d || !out) fail(FREHG_EINVAL, "null argument");
    auto* o = new Oracle();
    try {
      o->create(*d);
    } catch (...) {
      delete o;
      throw;
    }
    *out = o;
  });
}
int oracle_gw_destroy(void* p) {
  delete static_cast<Oracle*>(p);
  return FREHG_OK;
}
int oracle_gw_step(void* p, double t, double dtg) {
  return guard([&] { static_cast<Oracle*>(p)->step(t, dtg); });
}
int oracle_gw_next_dt(void* p, double* dt) {
  *dt = static_cast<Oracle*>(p)->dtgNext;
  return FREHG_OK;
}
int oracle_gw_audit(void* p, frehg_gw_audit* a) {
  *a = static_cast<Oracle*>(p)->audit;
  return FREHG_OK;
}
int oracle_gw_owned_volume(void* p, double* v) {
  *v = static_cast<Oracle*>(p)->ownedVolume();
  return FREHG_OK;
}
int oracle_gw_last_solve(void* p, frehg_gw_solve_stats* s) {
  *s = static_cast<Oracle*>(p)->last;
  return FREHG_OK;
}
int oracle_gw_active_cells(void* p, int64_t* n) {
  *n = static_cast<Oracle*>(p)->nActive;
  return FREHG_OK;
}
int oracle_gw_field_size(void* p, int f, int64_t* n) {
  return guard([&] {
    size_t m = 0;
    (void)field_data(*static_cast<Oracle*>(p), f, &m);
    *n = static_cast<int64_t>(m);
  });
}
int oracle_gw_get_field(void* p, int f, double* outp) {
  return guard([&] {
    size_t m = 0;
    const double* d = field_data(*static_cast<Oracle*>(p), f, &m);
    std::memcpy(outp, d, m * sizeof(double));
  });
}
int oracle_gw_enable_transport_window(void* p) {
  return guard([&] { static_cast<Oracle*>(p)->enableTransportWindow(); });
}
int oracle_gw_reset_transport_window(void* p) {
  return guard([&] { static_cast<Oracle*>(p)->resetTransportWindow(); });
}
int oracle_gw_set_field(void* p, int f, const double* in) {
  return guard([&] {
    Oracle& o = *static_cast<Oracle*>(p);
    if (f == FREHG_F_SINK) {  // NULL detaches the sink
      o.sinkActive = in != nullptr;
      if (in) {
        o.sink = o.h;  // halo-layout shape
        std::memcpy(o.sink.v.data(), in, o.sink.v.size() * sizeof(double));
      }
      return;
    }
    if (f == FREHG_F_VLOSS) {  // the checkpoint's cumulative vloss_cum
      if (!in) fail(FREHG_EINVAL, "null argument");
      std::memcpy(o.vloss.v.data(), in, o.vloss.v.size() * sizeof(double));
      return;
    }
    if (f != FREHG_F_HEAD && f != FREHG_F_WC)
      fail(FREHG_EINVAL, "only head, water content, vloss and the sink are settable");
    if (!in) fail(FREHG_EINVAL, "null argument");
    F3& fld = (f == FREHG_F_HEAD) ? o.h : o.wc;
    std::memcpy(fld.v.data(), in, fld.v.size() * sizeof(double));
    o.enforceHeadBc();
    o.enforceMoistureBc();
    o.hn = o.h;
    o.wcn = o.wc;
  });
}
int oracle_gw_set_et(void* p, const int* columns, const double* qvap) {
  return guard([&] {
    Oracle& o = *static_cast<Oracle*>(p);
    if (!columns != !qvap) fail(FREHG_EINVAL, "set_et: columns and qvap are given together");
    o.etActive = columns != nullptr && qvap != nullptr;
    if (!o.etActive) return;
    const size_t n2 = static_cast<size_t>(o.ny + 2) * (o.nx + 2);
    o.etCol.assign(columns, columns + n2);
    o.etQ.assign(qvap, qvap + n2);
  });
}
int oracle_gw_set_density(void* p, const double* r_rho, const double* r_visc,
                          const double* r_rho_top, const double* r_visc_top, int gradient) {
  return guard([&] {
    Oracle& o = *static_cast<Oracle*>(p);
    o.setDensity(r_rho, r_visc, r_rho_top, r_visc_top, gradient != 0);
    o.enforceHeadBc();  // hydrostatic ghosts carry the boundary-face ratio
  });
}
int oracle_gw_set_bc_values(void* p, int which, const double* values) {
  return guard([&] {
    Oracle& o = *static_cast<Oracle*>(p);
    if (which < 0 || which > 5) fail(FREHG_EINVAL, "bc selector out of range");
    F2<double>* vd[6] = {&o.topValue, &o.botValue, &o.vxm, &o.vxp, &o.vym, &o.vyp};
    F2<int>* cd[6] = {&o.topCode, &o.botCode, &o.sxm, &o.sxp, &o.sym, &o.syp};
    for (int j = 1; j <= o.ny; ++j)
      for (int i = 1; i <= o.nx; ++i) {
        double v = values[(j - 1) * o.nx + (i - 1)];
        if ((*cd[which])(j, i) == FREHG_BC_HEAD_HYDROSTATIC) v += o.offset;
        (*vd[which])(j, i) = v;
      }
  });
}
int oracle_gw_attach_coupling(void* p, double* eta, double* dept, double* seep, double* avail,
                              double* gain, double* bounce, double minDepth, int ventAll) {
  return guard([&] {
    Oracle& o = *static_cast<Oracle*>(p);
    if (!eta || !dept || !seep || !avail || !gain) fail(FREHG_EINVAL, "coupling arrays required");
    o.coupled = true;
    o.eta = eta;
    o.dept = dept;
    o.seep = seep;
    o.avail = avail;
    o.gain = gain;
    o.bounce = bounce;
    o.minDepth = minDepth;
    o.ventAll = ventAll != 0;
  });
}
// Accumulated wall seconds of the PCG solves and of whole steps, and the
// total PCG iteration count (the CPU baseline's clock).
int oracle_gw_timing(void* p, double* out3) {
  Oracle& o = *static_cast<Oracle*>(p);
  out3[0] = o.tSolve;
  out3[1] = o.tStep;
  out3[2] = static_cast<double>(o.itTotal);
  return FREHG_OK;
}
int oracle_gw_ktop(void* p, int* out) {
  Oracle& o = *static_cast<Oracle*>(p);
  std::memcpy(out, o.ktop.data(), o.ktop.size() * sizeof(int));
  return FREHG_OK;
}
int oracle_gw_geometry_scalars(void* p, double* out3) {
  Oracle& o = *static_cast<Oracle*>(p);
  out3[0] = o.offset;
  out3[1] = o.boxTop;
  out3[2] = o.pcClampTotal;
  return FREHG_OK;
}
// The assembled COO values [diag,xm,xp,ym,yp,zm,zp] (7N), rhs (N) and
// columns (7N) of the last predictor solve.
int oracle_gw_system(void* p, double* vals, double* rhs, int64_t* cols) {
  Oracle& o = *static_cast<Oracle*>(p);
  if (vals) std::memcpy(vals, o.vals.data(), o.vals.size() * sizeof(double));
  if (rhs) std::memcpy(rhs, o.rhs.data(), o.rhs.size() * sizeof(double));
  if (cols) std::memcpy(cols, o.cols.data(), o.cols.size() * sizeof(int64_t));
  return FREHG_OK;
}
// Run only the predictor's assembly on the current state for dtg (no solve).
int oracle_gw_assemble_only(void* p, double dtg) {
  return guard([&] {
    Oracle& o = *static_cast<Oracle*>(p);
    o.enforceHeadBc();
    o.enforceMoistureBc();
    o.hn = o.h;
    o.wcn = o.wc;
    if (o.coupled) o.classifyCoupledTop(dtg);
    o.faceConductivity();
    o.assemble(dtg);
  });
}

// The post-allocation sweep alone on the installed state, as the
// reference's tests run RichardsSolver::reallocateWaterContent
// (tests/unit/test_gw.cpp:1037), followed by the final clamp
// (Corrector.cpp:758-792) and the moisture ghosts, the same sequence the
// device path's test export runs.
int oracle_gw_test_post_allocate(void* p, double dtg) {
  return guard([&] {
    Oracle& o = *static_cast<Oracle*>(p);
    if (!(dtg > 0.0)) fail(FREHG_EINVAL, "dtg must be positive");
    o.audit = frehg_gw_audit{};
    o.dtgCurrent = dtg;
    if (o.transportWindow) std::fill(o.pathCorr.v.begin(), o.pathCorr.v.end(), 0.0);
    o.reallocate();
    o.finalize();
    o.enforceMoistureBc();
  });
}

// ---- generic COO linear system ---------------------------------------------
int oracle_ls_create(int64_t n, double rtol, double atol, int maxIt, int pc, void** out) {
  return guard([&] {
    if (n < 0) fail(FREHG_EINVAL, "negative row count");
    auto* ls = new OracleLs();
    ls->n = n;
    ls->rtol = rtol;
    ls->atol = atol;
    ls->maxIt = maxIt;
    ls->pc = pc;
    *out = ls;
  });
}
int oracle_ls_destroy(void* p) {
  delete static_cast<OracleLs*>(p);
  return FREHG_OK;
}
int oracle_ls_set_pattern(void* p, int64_t nCoo, const int64_t* rows, const int64_t* cols) {
  return guard([&] {
    OracleLs& ls = *static_cast<OracleLs*>(p);
    std::vector<std::vector<std::pair<int64_t, int64_t>>> perRow(ls.n);  // (col, coo index)
    for (int64_t e = 0; e < nCoo; ++e) {
      if (rows[e] < 0 || cols[e] < 0) continue;
      if (rows[e] >= ls.n || cols[e] >= ls.n) fail(FREHG_EINVAL, "COO index out of range");
      perRow[rows[e]].push_back({cols[e], e});
    }
    ls.rowPtr.assign(ls.n + 1, 0);
    ls.colIdx.clear();
    ls.cooSlot.assign(nCoo, -1);
    for (int64_t r = 0; r < ls.n; ++r) {
      auto& v = perRow[r];
      std::sort(v.begin(), v.end());
      int64_t last = -1;
      for (auto& pr : v) {
        if (pr.first != last) {
          ls.colIdx.push_back(pr.first);
          last = pr.first;
        }
        ls.cooSlot[pr.second] = static_cast<int64_t>(ls.colIdx.size()) - 1;
      }
      ls.rowPtr[r + 1] = static_cast<int64_t>(ls.colIdx.size());
    }
    ls.val.assign(ls.colIdx.size(), 0.0);
    ls.pattern = true;
  });
}
int oracle_ls_set_values(void* p, const double* values) {
  return guard([&] {
    OracleLs& ls = *static_cast<OracleLs*>(p);
    if (!ls.pattern) fail(FREHG_EINVAL, "setValues before setPattern");
    std::fill(ls.val.begin(), ls.val.end(), 0.0);
    for (size_t e = 0; e < ls.cooSlot.size(); ++e)
      if (ls.cooSlot[e] >= 0) ls.val[ls.cooSlot[e]] += values[e];
  });
}
int oracle_ls_solve(void* p, const double* b, double* x, frehg_gw_solve_stats* stats) {
  return guard([&] {
    OracleLs& ls = *static_cast<OracleLs*>(p);
    const int64_t n = ls.n;
    std::vector<double> r(n), z(n), pp(n), q(n), dinv(n, 1.0);
    auto mv = [&](const double* in, double* o) {
      for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t e = ls.rowPtr[i]; e < ls.rowPtr[i + 1]; ++e) s += ls.val[e] * in[ls.colIdx[e]];
        o[i] = s;
      }
    };
    for (int64_t i = 0; i < n; ++i)
      for (int64_t e = ls.rowPtr[i]; e < ls.rowPtr[i + 1]; ++e)
        if (ls.colIdx[e] == i && ls.val[e] != 0.0) dinv[i] = 1.0 / ls.val[e];
    for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
    for (int64_t i = 0; i < n; ++i) r[i] = b[i];
    const double bn = std::sqrt(dot(b, b, n));
    const double tol = std::fmax(ls.rtol * bn, ls.atol);
    double rn = bn, rz = 0.0;
    int it = 0;
    bool restart = true;
    frehg_gw_solve_stats st{};
    st.rhs_norm = bn;
    while (true) {
      if (rn <= tol) {
        mv(x, q.data());
        for (int64_t i = 0; i < n; ++i) q[i] = b[i] - q[i];
        const double tn = std::sqrt(dot(q.data(), q.data(), n));
        if (tn <= tol) {
          st.converged = 1;
          st.residual_norm = tn;
          break;
        }
        r = q;
        rn = tn;
        restart = true;
      }
      if (it >= ls.maxIt) {
        st.residual_norm = rn;
        break;
      }
      for (int64_t i = 0; i < n; ++i) z[i] = dinv[i] * r[i];
      const double rzn = dot(r.data(), z.data(), n);
      if (restart) {
        pp = z;
        restart = false;
      } else {
        const double beta = rzn / rz;
        for (int64_t i = 0; i < n; ++i) pp[i] = z[i] + beta * pp[i];
      }
      rz = rzn;
      mv(pp.data(), q.data());
      const double alpha = rz / dot(pp.data(), q.data(), n);
      for (int64_t i = 0; i < n; ++i) {
        x[i] += alpha * pp[i];
        r[i] -= alpha * q[i];
      }
      rn = std::sqrt(dot(r.data(), r.data(), n));
      ++it;
    }
    st.iterations = it;
    if (stats) *stats = st;
    if (!st.converged) fail(FREHG_EDIVERGED, "linear solve did not converge");
  });
}
int oracle_ls_is_symmetric(void* p, double tol, int* sym) {
  OracleLs& ls = *static_cast<OracleLs*>(p);
  auto at = [&](int64_t i, int64_t j) {
    for (int64_t e = ls.rowPtr[i]; e < ls.rowPtr[i + 1]; ++e)
      if (ls.colIdx[e] == j) return ls.val[e];
    return 0.0;
  };
  bool ok = true;
  for (int64_t i = 0; i < ls.n && ok; ++i)
    for (int64_t e = ls.rowPtr[i]; e < ls.rowPtr[i + 1]; ++e)
      if (std::fabs(ls.val[e] - at(ls.colIdx[e], i)) > tol) {
        ok = false;
        break;
      }
  *sym = ok ? 1 : 0;
  return FREHG_OK;
}

}  // extern "C"

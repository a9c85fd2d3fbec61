// gw_pcg.h — the structured-grid PCG that replaces PETSc KSPCG for the
// Richards head system (src/core/LinearSystem.cpp:147-227, :332-484).
//
// The operator is the symmetric 7-point stencil stored as four halo-layout
// arrays: Ad (diagonal), Axp / Ayp / Azp (coupling of a cell with its i+1 /
// j+1 / k+1 neighbour, zero where that neighbour is missing or inactive).
// Inactive interior rows are identity rows with a zero right-hand side.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>
#include <vector>

#include "gw_device.h"

namespace frehg {

template <class T>
struct Stencil {
  T *Ad = nullptr, *Axp = nullptr, *Ayp = nullptr, *Azp = nullptr;
};

// j-march launch shape: each wave walks S consecutive j rows of one i.
struct March {
  int S = 1;
  int nslab = 1;
  int blocks = 8;             // multiple of 8 (one share per XCD)
};

struct PcgStats {
  int iterations = 0;
  int refinements = 0;
  int converged = 0;
  double residual_norm = 0.0;
  double rhs_norm = 0.0;
  int fp64_iterations = 0;    // mixed: iterations finished in FP64 after an FP32 stall
  int rounds = 0;             // recursion rounds (FP64) or refinement cycles (mixed)
  int first_inner = 0;        // iterations of the first of them
  int later_inner = 0;        // most iterations of a later one
  double spmv_ms = 0.0;       // accumulated SpMV-kernel time (timing mode)
  int64_t spmv_launches = 0;
};

struct PcgOptions {
  double rtol = 1e-8, atol = 1e-14;
  int max_iterations = 1000;
  bool zline = true;          // z-line block Jacobi, else point Jacobi
  bool mixed = false;         // FP32 inner PCG inside FP64 refinement
  // mixed form: the FP32 inner PCG preconditioned by a Jacobi-Chebyshev
  // polynomial of odd degree cheb_degree on [cheb_lower * 2, 2] (zline off)
  bool cheb = false;
  int cheb_degree = 3;
  double cheb_lower = 0.05;
  double inner_rtol = 1e-4;   // FP32 residual reduction between reliable updates
  int max_refinements = 40;
  bool timing = false;        // bracket SpMV launches with HIP events
  int timing_every = 1;       // ... one launch in timing_every
  bool stencil32_ready = false;  // mixed: the caller wrote the FP32 stencil (stencil32())
};

// Laterally decoupled systems (no x or y legs): the columns with active rows
// in buckets of equal R = ceil(active rows / 64), deepest first, and the
// launch tile of each bucket (columns per wave).
struct ColumnTable {
  int buckets = 0;        // R = 1 .. 8 and 16 (513 .. 1024 rows): at most 9
  int R[9] = {};
  int cpw[9] = {};
  int colStart[10] = {};  // bucket b's columns: order[colStart[b] .. colStart[b + 1])
  int waveStart[10] = {}; // ... and its waves
};

// Launches of a device-controlled solve (Pcg::enqueueDevice): `rounds` FP64
// rounds (restarted recursions, each confirmed on the true residual) or mixed
// refinement cycles, of `iterations` direction/update launch pairs each.
struct PcgBudget {
  int rounds = 1;
  int iterations = 32;  // launch pairs of the first round / cycle
  int later = 0;        // ... of each later one (0: as the first)
};

// What a device-controlled solve left for the budget's next sizing.
struct DeviceSummary {
  int max_inner = 0;  // most iterations one round / cycle took
  int first_inner = 0;  // iterations of the first round / cycle
  int later_inner = 0;  // most iterations of a later one
  int rounds = 0;       // rounds / cycles that ran
  int stopped = 0;    // the solve ended inside the budget (converged or not)
  int fallback = 0;   // mixed: a cycle failed to halve the residual
};

// Caller-owned memory for FP64 solves (Pcg's external mode, which allocates
// nothing): device doubles for the work vectors, partial sums and state, and
// host buffers for the reduced values and the polled state.
struct PcgWorkspace {
  double* device = nullptr;    // >= Pcg::workspaceDoubles(g) doubles, 256-byte aligned
  int64_t doubles = 0;
  double* hostRed = nullptr;   // 4 doubles
  void* hostState = nullptr;   // Pcg::hostStateBytes() bytes
};

// Device-resident PCG workspace for one grid (allocated once).
class Pcg {
 public:
  Pcg(const Grid& g, hipStream_t stream);
  // External mode: the work vectors live in ws (nothing is allocated or
  // freed); FP64 solves only (the mixed form's FP32 arrays are allocated).
  Pcg(const Grid& g, hipStream_t stream, const PcgWorkspace& ws);
  ~Pcg();
  // Device doubles the external mode needs for grid g at the default launch
  // shapes on the current device, and the host state size.
  static int64_t workspaceDoubles(const Grid& g);
  static size_t hostStateBytes();
  // Zero r, z and both p buffers (a fill kernel): the stencil kernels read z
  // and p at the ghost cells around the grid, which a reused workspace may
  // hold from an earlier, larger grid.
  void clearVectors();
  Pcg(const Pcg&) = delete;
  Pcg& operator=(const Pcg&) = delete;

  // Solve A x = b; x holds the initial guess on entry (zero on inactive
  // rows) and the solution on exit.  Returns 0, or FREHG_EDIVERGED.
  int solve(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
            PcgStats* stats);

  // Device-controlled solve: the same launches and decisions as solve(), as
  // a fixed sequence with no host synchronization (capturable in a graph).
  // It ends by copying the control state to its host mirror; after the
  // stream has reached that point, deviceResult() reads it.  A solve that
  // did not converge inside the budget leaves `gate` (gateFlag(), device
  // memory) at 1 and resumeDevice() finishes it the host-polled way from
  // where it stopped.  prepareDevice() allocates what the form needs
  // (before any capture).
  void prepareDevice(const PcgOptions& opt);
  void enqueueDevice(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                     const PcgBudget& budget);
  const int* gateFlag() const;
  int deviceResult(PcgStats* stats, DeviceSummary* summary) const;  // returns the gate
  int resumeDevice(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                   PcgStats* stats);

  // Laterally decoupled operators: solve(), with the z-line preconditioner,
  // becomes one direct solve of every column (k_column_solve), checked on
  // its true residual (the PCG finishes a solve that misses the tolerance).
  // kt: first active row of each column (device, 2D halo layout), ktHost
  // the same on the host.  Tiles: columns per wave of each depth bucket.
  void setColumnSolve(const int* kt, const std::vector<int>& ktHost);
  bool columnSolve() const { return direct_; }
  bool columnSolveAvailable() const { return !bucketR_.empty() || kt_ != nullptr; }
  void enableColumnSolve(bool on) { direct_ = on && kt_ != nullptr; }
  void setColumnTiles(const int* cpw, int n);
  int columnTiles(int* R, int* counts, int* cpw, int cap) const;

  // The FP32 stencil of the mixed form (allocated on the first call).  A
  // caller that writes it as the rounded FP64 stencil sets
  // PcgOptions::stencil32_ready, and the solve skips its own cast.
  Stencil<float> stencil32();

  // Algorithmic bytes of one launch of the stencil kernel (the dominant one).
  double spmvBytesPerLaunch(bool mixed) const;

  // Launch shapes, as waves per CU: the FP64 and FP32 direction kernels and
  // the column kernels (0 keeps a value).  Defaults 4 / 4 / 16 where the
  // direction kernels run over column pairs (nx and nz even), else 16 / 32
  // / 16 (the fastest of sweeps at 512 x 512 x 100).  The FP64 residual and
  // refinement sweeps keep 16 waves per CU.  Results depend on the shapes at
  // rounding level (the order of the partial sums) and are deterministic for
  // given shapes.
  void setShapes(int march64WavesPerCu, int march32WavesPerCu, int columnWavesPerCu);
  void shapes(int* march64WavesPerCu, int* march32WavesPerCu, int* columnWavesPerCu) const {
    *march64WavesPerCu = wpc64_;
    *march32WavesPerCu = wpc32_;
    *columnWavesPerCu = wpcCol_;
  }

 private:
  int solveFp64(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                PcgStats* st);
  int solveMixed(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                 PcgStats* st);
  // mixed-loop state carried between cycles (and handed over by resumeDevice)
  struct MixedCarry {
    double* xa = nullptr;  // current iterate
    double* xb = nullptr;  // the other x buffer
    double rr = 0.0, tol2 = 0.0, s = 1.0;
    int total = 0, n0 = 0;
    bool restart = true;
    float* pOld = nullptr;
    float* pNew = nullptr;
  };
  int solveDirect(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                  PcgStats* st);
  void launchColumnSolve(const Stencil<double>& A, const double* b, double* x);
  void buildColumnTable();
  int fp64Loop(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
               PcgStats* st, double rr, double tol2, int total, int stalls);
  void mixedSetup(const Stencil<double>& A, const PcgOptions& opt);
  int mixedLoop(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                PcgStats* st, MixedCarry c);
  int fp64Finish(const Stencil<double>& A, const double* b, double* x, const PcgOptions& opt,
                 PcgStats* st, const double* best, int total);
  std::pair<double, double> residualHost(const Stencil<double>& A, const double* b,
                                         const double* x, bool zline);
  std::pair<double, double> refineHost(const Stencil<double>& A, const double* b, double* xi,
                                       double* xo, double sc, bool hasD);
  template <class T>
  void launchDir(const Stencil<T>& A, const T* z, const T* pOld, T* pNew, T* q, int np, int n,
                 bool restart, const March& mj);
  template <class T>
  void launchUpdate(T* x, const T* pPrev, const T* p, T* r, const T* q, T* z, const T* invd,
                    const T* w, bool zline, bool two, int n, const March& mj);
  void launchFactor(const Stencil<double>& A, bool zline, bool fp32);
  void launchResidual(const Stencil<double>& A, const double* b, const double* x, bool zline,
                      const void* ctl);
  void launchRefine(const Stencil<double>& A, const double* b, double* xa, double* xb, double sc,
                    bool hasD, const void* ctl);
  void launchInnerInit(bool zline, double s, const void* ctl);
  void launchChebSteps(const PcgOptions& opt, const int* skip);
  template <class T>
  int innerIterations(const Stencil<T>& A, T* x, T*& pOld, T*& pNew, T* r, T* z, T* q,
                      const T* invd, const T* w, bool zline, double tol2, int maxIt,
                      const March& mj, int nprev0, int n0, bool restart,
                      const PcgOptions& opt, PcgStats* st, int* itDone, double rr0 = -1.0);
  void reduceHost(int nparts, int nvals, double* out);
  void applyShapes();
  void initCommon();

  struct Shapes {
    int blocks = 0, pstride = 0;
    March march64, march32, march32w, march64w, sweep64w;
  };
  static Shapes shapesFor(const Grid& g, int cus, int wpcCol, int wpc64, int wpc32);
  // default direction-kernel shapes of grid g
  static void defaultShapes(const Grid& g, int* wpc64, int* wpc32);
  static March makeMarch(const Grid& g, int cus, int wavesPerCu, int columnsPerWave = 1);
  // block count of the direction kernel (whose partial sums the update reduces)
  int dirBlocks(bool fp32, const March& mj) const {
    if (fp32) return pairs_ && cw2_ ? march32w_.blocks : mj.blocks;
    return pairs64_ && cw2(false) ? march64w_.blocks : mj.blocks;
  }
  // the direction kernel runs over column pairs (FP64: columns of at most
  // kCw2Rows64 rows; deeper pairs spill registers)
  bool cw2(bool fp32) const { return cw2_ && (fp32 || g_.nz <= kCw2Rows64); }
  // the FP64 residual / refinement sweeps run over column pairs, and the
  // block count of their partial sums
  bool sweepPairs() const { return pairs64_ && cw2(false); }
  int sweepBlocks() const { return sweepPairs() ? sweep64w_.blocks : march64_.blocks; }
  static constexpr int kCw2Rows64 = 256;
  // grids whose FP64 iteration kernels can take the row-pair layout (nz even,
  // 64 < nz <= 512), and whose FP64 direction kernel then runs over column
  // pairs (nx even, nz <= kCw2Rows64)
  static bool pairForm64(const Grid& g) { return g.nz % 2 == 0 && g.nz > 64 && g.nz <= 512; }
  static bool colPairs64(const Grid& g) {
    return pairForm64(g) && g.nx % 2 == 0 && g.nz <= kCw2Rows64;
  }
  // the FP64 operands of a solve allow the pair layout (16-byte aligned)
  void choosePairs64(const Stencil<double>& A, const double* b, const double* x);

  Grid g_;
  hipStream_t stream_;
  int blocks_ = 0;            // grid size of the column kernels (multiple of 8)
  March march64_, march32_;   // stencil-kernel shapes, FP64 and FP32 operands
  March march32w_;            // FP32 direction kernel over column pairs (cw2_)
  March march64w_;            // FP64 direction kernel over column pairs (cw2_, pairs64_)
  March sweep64w_;            // FP64 residual / refinement sweeps over column pairs
  int pstride_ = 0;           // slot stride of parts_ (largest grid)
  int partsCap_ = 0;          // slot stride parts_ was allocated for
  int cus_ = 256;
  int wpc64_ = 16, wpc32_ = 32, wpcCol_ = 16;  // waves per CU (setShapes)
  bool pairs_ = false;        // FP32 kernels in the row-pair layout (nz even)
  bool pairs64_ = false;      // FP64 iteration kernels likewise (nz even, operands aligned)
  bool cw2_ = false;          // ... and the direction kernel over column pairs (nx even)
  bool quads_ = false;        // FP32 direction kernel in the quad layout (nz % 4 == 0, <= 128)
  // FP64 vectors (halo layout)
  double *r_ = nullptr, *z_ = nullptr, *p0_ = nullptr, *p1_ = nullptr, *q_ = nullptr;
  double *invd_ = nullptr, *w_ = nullptr;
  // FP32 copies for the mixed form
  float *Ad32_ = nullptr, *Axp32_ = nullptr, *Ayp32_ = nullptr, *Azp32_ = nullptr;
  float *r32_ = nullptr, *z32_ = nullptr, *p032_ = nullptr, *p132_ = nullptr, *q32_ = nullptr;
  float *d32_ = nullptr, *invd32_ = nullptr, *w32_ = nullptr;
  float* chebD32_ = nullptr;  // the Chebyshev recurrence's d (allocated on first use)
  double* parts_ = nullptr;   // 4 x blocks_ partial sums
  double* red_ = nullptr;     // reduced values (device)
  double* redHost_ = nullptr; // pinned mirror
  void* state_ = nullptr;     // PcgState (device)
  void* stateHost_ = nullptr; // pinned mirror (external mode: the caller's host buffer)
  bool external_ = false;     // memory owned by the caller (PcgWorkspace)
  // timing mode's event pairs around sampled direction launches, created on
  // first use and kept (creating them per solve leaves the device idle)
  std::vector<hipEvent_t> evs_;
  // direct column solve
  bool direct_ = false;
  const int* kt_ = nullptr;
  int* order_ = nullptr;          // columns, bucket by bucket (device)
  std::vector<int> bucketCols_;   // column indices of each bucket, concatenated (host)
  std::vector<int> bucketCount_, bucketR_, bucketCpw_;
  ColumnTable tab_;
  int colBlocks_ = 0;
};

}  // namespace frehg

// Internal layout of a gomc_box: one simulation box resident on a gfx950
// device, plus the host-side state of the single-box MC driver.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/gomc_hip.h"
#include "device_common.hpp"
#include "host_rng.hpp"

namespace gomc {

constexpr int kMaxMolAtoms = 16;   // atoms per molecule in a trial launch
constexpr int kBlock = 256;        // threads per workgroup, plain kernels

// Per-block partial of a fused trial: LJ, real, reciprocal, overlap.
struct Partial4 {
  double lj, real, recip, overlap;
};

// MoveSettings state of the driver (src/MoveSettings.cpp:19-258).
struct DriverMoves {
  std::vector<double> scale[2];  // [displace|rotate][kind]
  std::vector<uint64_t> tries[2], accepted[2], tempTries[2], tempAccepted[2];
};

// Running energies, src/EnergyTypes.h:75-129 order.
struct Energies {
  double inter = 0, tc = 0, real = 0, recip = 0, self = 0, corr = 0, total = 0;
  double Total() {
    total = inter + tc + real + recip + self + corr;
    return total;
  }
};

}  // namespace gomc

struct gomc_box {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  // ---- topology and force field (host) ----
  int natoms = 0, nmols = 0, ntypes = 0, nkinds = 0, maxMolAtoms = 0;
  std::vector<std::vector<int>> kindType;
  std::vector<std::vector<double>> kindCharge;
  std::vector<int> molKind, molStart, molLen, atomType, atomMol, kindCount;
  std::vector<double> atomCharge;
  std::vector<double> sigSq, epsCn, epsCn6, nOver6, nHalf, nRep;  // [ntypes^2]
  std::vector<double> pairEnCorr;                                 // [nkinds^2]
  double rCutLJ = 0, rCutCoul = 0, rCutLow = 0, tol = 0, T = 0, beta = 0;
  double alpha = 0, recipRcut = 0, recipRcutSq = 0, volume = 0, volInv = 0;
  bool ewald = true, elect = true, lrc = true;
  gomc::BoxGeom geom{};

  // ---- authoritative host copy of the coordinates, COMs ----
  std::vector<double> hx, hy, hz, comx, comy, comz;

  // ---- device state (HBM resident) ----
  double *x = nullptr, *y = nullptr, *z = nullptr, *q = nullptr;
  int *type = nullptr, *mol = nullptr;
  double *tSigSq = nullptr, *tEpsCn = nullptr, *tNHalf = nullptr, *tEpsCn6 = nullptr,
         *tNOver6 = nullptr;
  // cell list
  int ncell = 0, edge[3] = {0, 0, 0};
  double cellSize[3] = {0, 0, 0};
  int *cellCount = nullptr, *cellCursor = nullptr, *cellStart = nullptr, *cellVec = nullptr,
      *atomCell = nullptr, *nbr = nullptr;
  // Ewald
  int K = 0;
  double *kx = nullptr, *ky = nullptr, *kz = nullptr, *pref = nullptr;
  double *sumRref = nullptr, *sumIref = nullptr, *sumRnew = nullptr, *sumInew = nullptr;
  bool recipReady = false;
  // reductions: per-block partials, a completion counter, a mapped result
  gomc::Partial4 *partials = nullptr;
  int maxPartials = 0;
  unsigned *counter = nullptr;
  double *hostRes = nullptr, *devRes = nullptr;  // 8 doubles, host-mapped
  // a molecule whose accepted coordinates the next trial launch writes
  int pendingMol = -1;
  double pend[3][gomc::kMaxMolAtoms];

  // ---- single-box MC driver state ----
  gomc::MTRand mt{5489u};
  gomc::DriverMoves moves;
  gomc::Energies pot;
  double movePerc[3] = {0.5, 0.5, 0.0};
  uint64_t perAdjust = 1000;
};

namespace gomc {
int Fail(gomc_box *b, int code, const std::string &msg);
int HipFail(gomc_box *b, hipError_t e, const char *what);
// Launch helpers implemented in box.hip.
int LaunchTrial(gomc_box *b, int mol, const double *ox, const double *oy, const double *oz,
                const double *nx, const double *ny, const double *nz, bool doPair,
                bool doRecip, double out[4]);
void HostCorrectionSelfLRC(gomc_box *b, double &self, double &corr, double &lrc);
int RecipSums(gomc_box *b, double *erecip);
int BoxInter(gomc_box *b, double *lj, double *real);
}  // namespace gomc

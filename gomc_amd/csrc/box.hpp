// Internal layout of a gomc_box: one simulation box resident on a gfx950
// device, plus the host-side state of the single-box MC driver.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/gomc_hip.h"
#include "device_common.hpp"
#include "host_rng.hpp"

#define HIP_TRY(b, call)                                      \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return gomc::HipFail(b, e_, #call); \
  } while (0)

namespace gomc {

struct DevDriver;  // buffers of the device move loop (gcmc_dev.hip)
constexpr int kMaxMolAtoms = 16;   // atoms per molecule in a trial launch
constexpr int kBlock = 256;        // threads per workgroup, plain kernels

// Structure factors and reciprocal forces walk the k list in segments:
// pieces of at most kSeg consecutive-nz vectors of one (nx, ny) run, built
// on the device by one workgroup of kSegThreads (box.hip).
constexpr int kSeg = 16;
constexpr int kSegThreads = 1024;
// First structure-factor pass: a fixed grid of kSumsGrid workgroups loops
// over (segment block, atom chunk) items; chunks x K stays within
// kSumsPartCap so the per-chunk sums fit a bounded buffer.
constexpr int kSumsGrid = 1024;
constexpr size_t kSumsPartCap = size_t(1) << 21;

// Atom chunks of the first structure-factor pass: about kSumsGrid items
// over the segment blocks, chunks of at least 32 atoms, within the cap.
__host__ __device__ inline int SumsChunks(int nseg, int natoms, int K) {
  const int segBlocks = nseg > kBlock ? (nseg + kBlock - 1) / kBlock : 1;
  int c = (kSumsGrid + segBlocks - 1) / segBlocks;
  const int byAtoms = natoms > 32 ? (natoms + 31) / 32 : 1;
  if (c > byAtoms) c = byAtoms;
  const int byCap = K > 0 ? int(kSumsPartCap / size_t(K)) : 1;
  if (c > byCap) c = byCap;
  return c > 1 ? c : 1;
}

// CellList::PositionToCell (src/CellList.h:88-101), on the host and the
// device alike: both sides of the trial cell table must put a position in
// the same cell.
__host__ __device__ inline int CellOf(double px, double py, double pz, const double *cs,
                                      const int *edge) {
  int ix = int(px / cs[0]), iy = int(py / cs[1]), iz = int(pz / cs[2]);
  ix -= (ix == edge[0] ? 1 : 0);
  iy -= (iy == edge[1] ? 1 : 0);
  iz -= (iz == edge[2] ? 1 : 0);
  // Coordinates are validated to [0, L) on upload; the clamp only keeps a
  // bad value from indexing outside the cell arrays.
  ix = ix < 0 ? 0 : (ix > edge[0] - 1 ? edge[0] - 1 : ix);
  iy = iy < 0 ? 0 : (iy > edge[1] - 1 ? edge[1] - 1 : iy);
  iz = iz < 0 ? 0 : (iz > edge[2] - 1 ? edge[2] - 1 : iz);
  return ix * edge[1] * edge[2] + iy * edge[2] + iz;
}

// Per-block partial of a fused trial: LJ, real, reciprocal, overlap.
struct Partial4 {
  double lj, real, recip, overlap;
};

// Completion of a multi-block reduction: each block publishes its partial,
// the last block to arrive sums all partials -- thread t the partials t,
// t + blockDim, ... in that order, then the threads' sums by BlockSum's
// fixed tree, so the order depends on the grid and block sizes only -- and
// writes the four sums to the host-mapped result, then re-arms the counter.
// The partials are read by every thread at once: one round of loads rather
// than one dependent load per block.  Returns true in every thread of that
// last block.
__device__ inline bool FinishPartials(const Partial4 &mine, Partial4 *partials,
                                      unsigned *counter, double *result) {
  __shared__ bool last;
  __shared__ double red[1024 / kWave];
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = mine;
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (last) {
    __threadfence();
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    volatile Partial4 *vp = partials;
    for (unsigned i = threadIdx.x; i < gridDim.x; i += blockDim.x) {
      s0 += vp[i].lj;
      s1 += vp[i].real;
      s2 += vp[i].recip;
      s3 += vp[i].overlap;
    }
    s0 = BlockSum(s0, red);
    s1 = BlockSum(s1, red);
    s2 = BlockSum(s2, red);
    s3 = BlockSum(s3, red);
    if (threadIdx.x == 0) {
      result[0] = s0;
      result[1] = s1;
      result[2] = s2;
      result[3] = s3;
      *counter = 0u;
      __threadfence_system();
    }
  }
  return last;
}

// MultiParticle state (src/moves/MultiParticle.h:37-58 and the
// MultiParticle members of src/MoveSettings.h): step sizes, interval
// counters, and the device force sets [ref|trial].
struct MPState {
  bool singleAccepted = true;  // MoveSettings::isSingleMoveAccepted
  double tMax = 0.02, rMax = 0.01 * M_PI;
  uint64_t tries[2] = {0, 0}, acc[2] = {0, 0}, itries[2] = {0, 0}, iacc[2] = {0, 0};
  int ref = 0;                 // index of the reference force set
  bool ready = false;
  double *xn = nullptr, *yn = nullptr, *zn = nullptr;  // trial coordinates
  double *com[2][3] = {};                               // [current|trial][xyz]
  double *aF[2][3] = {}, *aFR[2][3] = {};               // atom forces [set][xyz]
  double *mF[2][3] = {}, *mFR[2][3] = {}, *tq[2][3] = {};  // molecule terms
  double *tk[3] = {};                                   // t_k or r_k
  int *inRange = nullptr, *fail = nullptr, *molStart = nullptr;
  int *chunks = nullptr;                                // force work list
  int nchunks = 0, maxChunks = 0;
  double *wpart = nullptr;                              // per-block products
  unsigned *wcounter = nullptr;
  double *fpart = nullptr;  // per-k-chunk reciprocal forces [chunk][xyz][natoms]
  size_t fpartCap = 0;
  double *spart = nullptr;  // per-slice pair forces [slice][xyz][natoms]
  size_t spartCap = 0;
};

// MoveSettings state of the driver (src/MoveSettings.cpp:19-258).
struct DriverMoves {
  std::vector<double> scale[2];  // [displace|rotate][kind]
  std::vector<uint64_t> tries[2], accepted[2], tempTries[2], tempAccepted[2];
  // VOL_TRANSFER (NPT): the reference keeps identical counts for every
  // kind (MoveSettings.cpp:130-145); one set stands for all of them
  double volScale = 500.0;  // MoveSettings.cpp:58-64
  uint64_t volTries = 0, volAccepted = 0, volTempTries = 0, volTempAccepted = 0;
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
  int precision = 0;  // 1: trial kernels with FP32 pair terms and phases, FP64 sums
  int devTeam = 0;    // workgroups per box in the device move loop (0: by k-vector count)
  std::string err;

  // ---- topology and force field (host) ----
  int natoms = 0, nmols = 0, ntypes = 0, nkinds = 0, maxMolAtoms = 0;
  std::vector<std::vector<int>> kindType;
  std::vector<std::vector<double>> kindCharge;
  std::vector<int> molKind, molStart, molLen, atomType, atomMol, kindCount;
  std::vector<double> atomCharge;
  std::vector<double> sigSq, epsCn, epsCn6, nOver6, nHalf, nRep;  // [ntypes^2]
  std::vector<double> pairEnCorr, pairVirCorr;                    // [nkinds^2]
  double rCutLJ = 0, rCutCoul = 0, rCutLow = 0, tol = 0, T = 0, beta = 0;
  double alpha = 0, recipRcut = 0, recipRcutSq = 0, volume = 0, volInv = 0;
  bool ewald = true, elect = true, lrc = true;
  gomc::BoxGeom geom{};

  // ---- authoritative host copy of the coordinates, COMs ----
  std::vector<double> hx, hy, hz, comx, comy, comz;

  // ---- device state (HBM resident) ----
  double *x = nullptr, *y = nullptr, *z = nullptr, *q = nullptr;
  int *type = nullptr, *mol = nullptr;
  int *molStartDev = nullptr;  // [nmols + 1] first atom of each molecule
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

  std::vector<int> hCellCount;  // atoms per cell at the last build
  // Trial cell table (plain trial path, boxes of more than 27 cells): per
  // cell up to tcCap atom slots of the live atoms, in no particular order.
  // The host copy (hTc*) is authoritative; the device copy trails it by the
  // edits in tcEdits -- (address, value) pairs, an address below
  // ncell * tcCap being a slot and above it a count -- which the next
  // trial launch applies after its last reader is done, or a small edit
  // kernel applies before anything else reads the table.
  bool tcValid = false;
  int tcCap = 0;
  int *tcCount = nullptr, *tcSlots = nullptr;  // device [ncell], [ncell * tcCap]
  size_t tcCountAlloc = 0, tcSlotsAlloc = 0;
  std::vector<int> hTcCount, hTcSlots, hTcAtomCell, hTcAtomSlot, tcMark;
  std::vector<int> tcEdits;
  // EwaldCached mode (src/EwaldCached.cpp): per-molecule structure factors
  // cacheC/S [nmols][K] of the current positions, and the trial molecule's
  // new terms staged until the move is accepted.
  bool cached = false;
  double *cacheC = nullptr, *cacheS = nullptr, *stageC = nullptr, *stageS = nullptr;
  int stageMol = -1;
  // staging for trial positions and their energies (ParticleInter)
  double *trialBuf = nullptr;
  int trialCap = 0;
  // per-block partial sums of the virial kernels
  double *virBuf = nullptr;
  size_t virCap = 0;
  // per-chunk structure factors of RecipSums [chunks][K][2]
  double *sumPart = nullptr;
  size_t sumPartCap = 0;
  // k segments (box.hip): starts [nseg + 1] and the count, on the device;
  // nseg also on the host for launch shapes
  int *kseg = nullptr, *nsegDev = nullptr;
  int ksegCap = 0, nseg = 0;

  // ---- GCMC (ensemble 1): box 0 interacts, box 1 is the reservoir ----
  // Reservoir atoms carry molecule id -1 on the device and are skipped by
  // every pair, cell and structure-factor kernel.
  int ensemble = 0;
  std::vector<int> molBox;
  // MoleculeLookup (src/MoleculeLookup.cpp:24-119): molecule indices by
  // [box][kind], section starts, swappable counts, kinds in order of first
  // appearance.
  std::vector<uint32_t> lookup, bkStart, swapCount;
  std::vector<int> canSwapKind, canMoveKind;
  std::vector<double> fugacity;  // K / A^3
  int cbmcFirst = 16, cbmcNth = 8, cbmcAng = 100;
  struct Hedron {  // rigid three-site kind grown from atom 0
    double bPrev = 0, bBonded = 0, theta = 0;
  };
  std::vector<Hedron> hedron;
  double moveW[13] = {0};  // StaticVals movePerc by move kind
  std::vector<uint64_t> swapTries, swapAcc;  // [dest box][kind]
  uint64_t seed = 0;            // Random_Seed (MT and Random123)
  gomc::MPState mp;

  // ---- single-box MC driver state ----
  gomc::MTRand mt{5489u};
  gomc::DriverMoves moves;
  gomc::Energies pot;
  double movePerc[3] = {0.5, 0.5, 0.0};
  uint64_t perAdjust = 1000;
  gomc::DevDriver *dev = nullptr;  // device move loop, allocated on first use
  double pressure = 0.0;             // NPT: K / A^3
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
// Cuts the box's k list into segments (after the k-vectors change).
int BuildKSegments(gomc_box *b);
int BoxInter(gomc_box *b, double *lj, double *real);
int EnsurePartials(gomc_box *b, int nblocks);
int Sync(gomc_box *b);
int BuildCells(gomc_box *b);
int FlushPending(gomc_box *b);
// Trial cell table (box.hip).  Valid only after TrialCellsEnsure; any
// change of the coordinates or of which molecules live in the box other
// than gomc_commit_mol calls TrialCellsInvalidate.
inline bool TrialCellsUseful(const gomc_box *b) { return b->ncell > 27; }
inline void TrialCellsInvalidate(gomc_box *b) {
  b->tcValid = false;
  b->tcEdits.clear();
}
int TrialCellsEnsure(gomc_box *b);
int TrialCellsFlushEdits(gomc_box *b);
// MultiParticle pieces (mp.hip).
int MPEnsure(gomc_box *b);
void MPFree(gomc_box *b);
int MPUploadCom(gomc_box *b);
int MPPairForce(gomc_box *b, int set, double *lj, double *real);
int MPRecipForce(gomc_box *b, int set);
int MPMolTerms(gomc_box *b, int set, int comSet);
int MPMove(gomc_box *b, uint64_t step, int &moveType, int &accepted, int &failed, double dE[3]);
// GCMC pieces (gcmc.cpp / box.hip).
int ParticleInterImpl(gomc_box *b, int atomType, double charge, int skipMol, int n,
                      const double *tx, const double *ty, const double *tz, double *en,
                      double *real, int32_t *overlap);
int SetMolLive(gomc_box *b, int m, bool live);
void LookupInit(gomc_box *b);
uint32_t NumKindInBox(const gomc_box *b, int k, int box);
int GetMolNum(const gomc_box *b, uint32_t sub, int k, int box);
int TransferMove(gomc_box *b, double subDraw, int &mol, int &accepted, int &overlap,
                 double dE[3]);
// Ewald swap delta; with the molecule cache on, an insertion of molecule
// `mol` stages its new terms.
int SwapRecipImpl(gomc_box *b, int kind, int insert, const double *x, const double *y,
                  const double *z, int mol, double *enew);
// Rebuilds the per-molecule cache from the current coordinates.
int CacheBuild(gomc_box *b);
// Frees the device move loop's buffers.
void DevFree(gomc_box *b);
// k-vectors of the current geometry, structure factors of the current
// coordinates, made the reference sums; *erecip = BoxReciprocal.
int RecipInitImpl(gomc_box *b, double *erecip);
// VolumeTransfer of an NPT box (volume.cpp).
int VolumeMove(gomc_box *b, int &accepted, double dE[3]);
// Replaces the cell grid by one for the box's current geometry (volume.cpp).
int Regrid(gomc_box *b);
// Half-space k-vectors of an orthogonal box in the reference's order
// (box_setup.cpp); `packed` (may be null) gets their index triples.
void KVectorsOrth(const double L[3], double alphaSq, double recipRcut, double volume,
                  std::vector<double> &kx, std::vector<double> &ky, std::vector<double> &kz,
                  std::vector<double> &pref, std::vector<int32_t> *packed);
}  // namespace gomc

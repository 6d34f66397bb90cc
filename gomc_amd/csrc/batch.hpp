// Batched persistent MC over many independent boxes: shared layout of the
// per-box parameter block used by the host (batch_setup.cpp) and the
// kernel (batch.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/gomc_hip.h"
#include "host_rng.hpp"

namespace gomc {

constexpr int kBatchThreads = 1024;  // one workgroup (16 waves) per box
constexpr int kBatchMaxAtoms = 4;    // atoms per molecule (water, CO2, TIP4P)
constexpr int kBatchMaxTypes = 4;    // LJ atom types per box
constexpr int kBatchMaxN = 32;       // largest |n| of a k-vector component
constexpr int kDrawChunk = 64;       // moves whose Philox draws are made at once
constexpr uint64_t kPhiloxTag = 0x474F4D43ull;  // key word 1 of every box stream
// Teams: a large box may be run by T = 2, 4 or 8 workgroups at once, each
// taking a share of the pair items and k-vector rounds; per move the
// members exchange their partial energy sums.
constexpr int kTeamMax = 8;
// Exchange block per box (64-bit words): tagged partial sums
// [2 parities][8 members][4].
constexpr int kXchgWords = 2 * kTeamMax * 4;
// Polls of a member's counter before a team gives up and flags an error.
constexpr int64_t kTeamSpinLimit = int64_t(1) << 22;
// Molecules whose FP64 centres fit in the LDS left beside the static
// tables, pair queues and near list (160 KiB per CU on gfx950, ~60 KiB
// static): 4096 x 24 B.
constexpr int kBatchLdsMolCap = 4096;
// Capacity of the per-move list of molecules near the moving one; boxes of
// more molecules skip the centre cull.
constexpr int kBatchNearCap = 4096;
// Pair-phase waves per box and the near-list entries each wave may hold:
// a wave tests at most ceil(4096 / (64 w)) chunks of 64 molecules.
constexpr int kPairWavesMax = 12;
constexpr int kNearListLen = 4608;  // 12 x 384, 8 x 576, 4 x 1152

// Read-only description of one box (host-filled, 8-byte aligned fields).
struct BatchBox {
  double L[3], half[3];
  double rCutSq, rCutLJSq, rCutCoulSq, rCutLowSq;
  double alpha, beta, pDisp;
  double kc[3];                       // (1/L) * 2pi, the k-vector spacing
  double nearCutSq;                   // (rc + 2 r_max + 1e-6)^2: centres farther apart
                                      // than this hold no pair inside the cutoff
  double sigSq[kBatchMaxAtoms * kBatchMaxAtoms];  // per (atom p, atom q) of the kind
  double epsCn[kBatchMaxAtoms * kBatchMaxAtoms];
  double nHalf[kBatchMaxAtoms * kBatchMaxAtoms];
  double qq[kBatchMaxAtoms * kBatchMaxAtoms];     // q_p q_q qqFact
  double charge[kBatchMaxAtoms];
  int64_t atomOff, molOff, kOff;      // offsets into the flat arrays
  int64_t perAdjust;
  uint64_t seed;
  uint64_t magicLen, magicLen2;       // ceil(2^32 / len), ceil(2^32 / len^2)
  int natoms, nmols, molLen;
  int K;       // k-vector slots of the box: kChunk * kSeg * kBatchThreads
  int kChunk;  // rounds: segments per thread
  int kSeg;    // S: k-vectors per segment (1, 2, 4 or 8)
  int nmax[3], ewald, elect;
  int cull;  // 1: cull far molecules by their centres before atom pairs
  int team;  // workgroups running the box (1, 2, 4, 8)
  int live;  // bit p * kBatchMaxAtoms + q: atoms p, q interact (charge or LJ)
  int pairWaves;  // waves 0 .. pairWaves-1 run the pair phase (4, 8 or 12)
  int kWidth;     // W = 1024 - 64 pairWaves: threads of the reciprocal phase and
                  // columns of the k-vector layout
};

static_assert(sizeof(BatchBox) % 8 == 0, "BatchBox is copied as 64-bit words");

// Mutable per-box state kept in HBM between launches.
struct BatchState {
  double energy[7];   // inter, tc, real, recip, self, correction, total
  double scale[2];    // displacement, rotation
  int64_t step;       // moves made so far (the Philox counter)
  int64_t tries[2], accepted[2], tempTries[2], tempAccepted[2];
  int64_t pad;
};

struct BatchTrace {
  int box;
  int64_t cap;
  int32_t *acc;
  double *dlj, *dreal, *drecip;
};

// One entry of a workgroup's task list: run `box` as member `rank`.
struct BatchTask {
  int box, rank;
};

struct BatchArgs {
  const BatchTask *tasks;     // task lists of all workgroups, back to back
  const int *wgTask;          // [grid + 1] start of each workgroup's list
  uint64_t *xchg;             // [nboxes * kXchgWords] team exchange blocks
  int *err;                   // set to 1 when a team member's wait runs out
  uint64_t *boxTicks;         // [nboxes] device wall-clock ticks of each box's moves
  const BatchBox *boxes;
  BatchState *state;
  double *x, *y, *z;          // [sum atoms]
  double *cx, *cy, *cz;       // [sum mols] centres of mass
  // k-vector arrays, [sum K] slots; per box, segments of S consecutive nz
  // at fixed (nx, ny): column t, round r, step j at slot (r*S + j)*W + t
  const int32_t *kn;          // packed (nx, ny+128, nz+128) bytes, -1 unused slot
  const double *pref;
  double *sumR, *sumI;        // reference structure factors
  double *dR, *dI;            // structure-factor change of the last trial
  const int64_t *moves;       // [nboxes] moves to make in this launch
  int ldsMols;                // per-box centre capacity of dynamic LDS
  BatchTrace trace;
};

}  // namespace gomc

struct gomc_batch {
  int device = 0;
  int precision = 0;
  int nboxes = 0;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double lastKernelMs = 0.0;
  std::vector<gomc::BatchBox> hboxes;
  gomc::BatchBox *boxes = nullptr;
  gomc::BatchState *state = nullptr;
  double *x = nullptr, *y = nullptr, *z = nullptr, *cx = nullptr, *cy = nullptr, *cz = nullptr;
  int32_t *kn = nullptr;
  double *pref = nullptr, *sumR = nullptr, *sumI = nullptr, *dR = nullptr, *dI = nullptr;
  int64_t *moves = nullptr;
  // workgroup task lists (device) and their capacity; team exchange blocks
  gomc::BatchTask *tasks = nullptr;
  int *wgTask = nullptr;
  uint64_t *xchg = nullptr;
  int *errFlag = nullptr;
  int grid = 0;        // workgroups per launch
  int gridCap = 0;     // workgroups that fit on the device at once
  bool teams = true;   // false: every box runs on one workgroup (serving)
  gomc_batch_tuning tune{{1800, 2700, 3500}, 1024};
  uint64_t *boxTicks = nullptr;
  std::vector<double> usPerMove;  // measured per-box cost of the last launch (0: none)
  std::vector<double> lastUs;     // per-box time of the last launch
  int64_t totalAtoms = 0, totalMols = 0, totalK = 0;
  int ldsMols = 0;
  gomc::BatchTrace trace{-1, 0, nullptr, nullptr, nullptr, nullptr};
  int64_t traceLen = 0;
  // one MC cycle (N_i moves per box) captured as a graph, replayed per cycle
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  // replica exchange: the stream of exchange draws (seeded at attempt 0)
  gomc::MTRand rexMt{5489u};
};

namespace gomc {
// Launch the MC kernel at the given precision (0 FP64, 1 FP32 pair terms);
// setAttr = false when PrepareBatch already sized its dynamic LDS (inside a
// stream capture).
hipError_t LaunchBatch(int precision, int grid, const BatchArgs &a, hipStream_t s,
                       bool setAttr = true);
hipError_t PrepareBatch(int precision, int ldsMols);
// Workgroups of the MC kernel that one CU holds at once with this LDS size.
hipError_t BatchResidentPerCU(int precision, int ldsMols, int *perCU);
// LDS centre capacity for a batch whose largest box has maxMols molecules
// (boxes beyond it read centres from HBM).
int BatchLdsMols(int maxMols);
void KVectorsOrth(const double L[3], double alphaSq, double recipRcut, double volume,
                  std::vector<double> &kx, std::vector<double> &ky, std::vector<double> &kz,
                  std::vector<double> &pref, std::vector<int32_t> *packed);

// One box packed for the batched kernel: its parameters, initial state and
// the contents of its flat-array extents.
struct PackedBox {
  BatchBox P;
  BatchState S;
  std::vector<double> x, y, z, cx, cy, cz, pref, sr, si;
  std::vector<int32_t> kn;
};
// Builds the box once through the plain path (initial energies, structure
// factors, centres) and packs it at the given offsets; errors go to bt->err.
int PackBox(gomc_batch *bt, const gomc_system_desc &d, int64_t atomOff, int64_t molOff,
            int64_t kOff, PackedBox &out);
// Device buffers for `nboxes` boxes and the given flat-array lengths.
int AllocBatch(gomc_batch *bt, int nboxes, int64_t atoms, int64_t mols, int64_t kvec);
// Team size of a box of n molecules (a function of n alone, so a box's
// results do not depend on the other boxes of its batch).
int TeamSize(const gomc_batch *bt, int nmols);
// Builds and uploads the workgroup task lists for the given move counts.
int BuildTasks(gomc_batch *bt, const std::vector<int64_t> &moves);
// Uploads a packed box: parameters and state at index `slot`, arrays at
// the box's offsets.
int UploadBox(gomc_batch *bt, int slot, const PackedBox &pb);
}  // namespace gomc

// Device-resident move loop: the single-box driver's moves -- Translate,
// Rotate and (GCMC) MoleculeTransfer with its configurational-bias growth --
// run on the device, one workgroup per box and many moves per launch, with
// each box's MT19937 stream held in LDS.  The draw order, the CBMC builders
// and the bookkeeping are those of the host driver (driver.cpp, gcmc.cpp),
// which follow src/System.cpp:262-302, src/moves/Translate.h:52-138,
// src/moves/Rotation.h:41-130, src/moves/MoleculeTransfer.h:85-246,
// src/cbmc/DCSingle.cpp:23-100, src/cbmc/DCFreeHedron.cpp:48-240,
// src/MoleculeLookup.cpp:165-198 and src/MoveSettings.cpp:205-258, so a box
// run here takes the same moves, in the same order, with the same outcomes,
// as on the host driver or the CPU restatement of the reference.
//
// Work split inside a move: thread 0 of the workgroup runs the serial part
// (random draws, trial geometry, Metropolis, molecule lookup), the whole
// workgroup runs the energies:
//   translate / rotate  every atom against the old and new positions of the
//                       molecule, and every k-vector's structure-factor delta,
//                       reduced in one pass;
//   CBMC stage          (trial position, atom) pairs spread over the waves,
//                       each wave summing one pair over every atom;
//   swap reciprocal     every k-vector's new structure factor.
// Energies are summed in a fixed order, so a run does not depend on timing.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "box.hpp"

namespace gomc {

namespace {

constexpr int kDevThreads = 1024;
constexpr int kDevWaves = kDevThreads / kWave;
constexpr int kDevKinds = 4;      // molecule kinds of a device-driven box
constexpr int kDevKindAtoms = 3;  // atoms per molecule (1- and 3-site rigid kinds)
constexpr int kDevTrials = 32;    // CBMC trials per growth stage
constexpr int kDevCombos = 2 * kDevTrials;
constexpr int kMtN = 624, kMtM = 397;
constexpr int kDevNMax = 40;             // largest |n| of a k-vector component (FP64 phases)
constexpr int kTab = 2 * kDevNMax + 1;   // phase-table entries per axis
constexpr int kDevPos = 2 * kDevKindAtoms;  // positions of one trial: new, then old
constexpr int kDisplace = 0, kRotate = 1, kMolTransfer = 10;
constexpr double kSmallWeight = 1.0e-38;  // lib/NumLib.h:19
// A box's moves may run on a team of workgroups: every member runs the whole
// move (draws, pair terms, CBMC stages, the Metropolis test) on the same
// data and so takes the same decisions, and the k-vectors are split among
// the members, whose partial reciprocal energies are exchanged once per move
// through device-coherent atomics (TeamExchange).
constexpr int kDevTeamMax = 8;
constexpr int kDevXchgWords = 2 * kDevTeamMax * 2;  // [parity][member]{value, seq}
constexpr int64_t kDevSpinLimit = int64_t(1) << 22;  // polls before a wait fails
constexpr int kDevCellMinAtoms = 512;  // live atoms from which the energy phases walk cells

// Read-only description of one box for a launch.
struct DevBox {
  BoxGeom g;
  double beta, volume, volInv;
  double pairEnCorr[kDevKinds * kDevKinds], fugacity[kDevKinds], selfKind[kDevKinds];
  double kindQ[kDevKinds][kDevKindAtoms];
  int kindT[kDevKinds][kDevKindAtoms], kindLen[kDevKinds];
  double hedB[kDevKinds], hedP[kDevKinds], hedTh[kDevKinds];
  double moveW[13], movePerc[2];
  int nMoveTypes, lrc, nkinds, natoms, nmols, ntypes, K;
  int cbmcFirst, cbmcNth, cbmcAng;
  int nSwapKinds, nMoveKinds, swapKind[kDevKinds], moveKind[kDevKinds];
  uint64_t perAdjust;
  double *x, *y, *z;
  const double *q;
  const int *type;
  int *mol;
  const double *sigSq, *epsCn, *nHalf;
  const double *kx, *ky, *kz, *pref;
  double *sumR, *sumI, *newR, *newI;
  const int *kpack;  // k-vector index triples: ix | (iy + 128) << 8 | (iz + 128) << 16
  int nmax[3];
  const int *molStart, *molKind;
  int *molBox;
  int *live, *livePos;  // atoms of the box's molecules (any order), each atom's place or -1
  int nLive0;
  // The same atoms by cell (cells = 1; this member's copy): counts [ncell],
  // slots [ncell][ccap], each atom's cell and slot (-1 outside the box);
  // nbr the 27 cells around each cell.  cells = 0: the energy phases scan
  // the live list.
  int cells, ncell, ccap;
  int edge[3];
  double cs[3];
  const int *nbr;
  int *cCount, *cSlots, *aCell, *aSlot;
  uint32_t *lookup;     // live, livePos and lookup are this member's own copies
  // team: members, this member's rank, its k-vectors [k0, k1), the box's
  // exchange block (kDevXchgWords words, zero at the launch's start)
  int team, rank, k0, k1;
  uint64_t *xchg;
  double *com;  // [3][nmols]
  struct DevState *st;
  int64_t step0, nsteps;
  int cycles;
  int32_t *tKind, *tMol, *tAcc, *tOvl;
  double *tE;  // [nsteps][3]
};

// Mutable state of a box's move loop, copied to LDS for a launch.
struct DevState {
  // MT19937: the current block of 624 state words is mt[mtCur]; the next
  // block, once the workgroup has computed it (nextReady), is mt[mtCur ^ 1]
  uint32_t mt[2][kMtN];
  int mtIdx, mtCur, nextReady, status;
  int recipSwapped;  // the launch ended with the two structure-factor arrays swapped
  double pot[7];  // inter, tc, real, recip, self, corr, total
  double scale[2][kDevKinds];
  uint64_t tries[2][kDevKinds], acc[2][kDevKinds], tTries[2][kDevKinds], tAcc[2][kDevKinds];
  uint64_t swapTries[2 * kDevKinds], swapAcc[2 * kDevKinds];
  uint32_t bkStart[2 * kDevKinds + 1], swapCount[2 * kDevKinds];
  int kindCount[kDevKinds];
  double sumN[2];
  int64_t samples, moves, accepted;
};

enum : int { kStErr = 1, kStEmptyReservoir = 2, kStWeights = 3, kStMoveKind = 4, kStTeam = 5 };

// One trial molecule (cbmc::TrialMol).
struct DTrial {
  double px[kDevKindAtoms], py[kDevKindAtoms], pz[kDevKindAtoms];
  double weight, inter, real;
  int overlap, box;
};

// Per-move control block shared by the workgroup.
struct Ctl {
  int op, mol, len, st, commit, doRecip, sign, found, abort, section;
  int src, dst, kind, mv, mk, tried, pick;
  double subDraw, pr, ncom[3], thetaW;
  double tcLose, tcGain, wTc, corrNew, corrOld, selfNew, selfOld, recipLose, recipGain;
  double ox[kDevKindAtoms], oy[kDevKindAtoms], oz[kDevKindAtoms];
  double nx[kDevKindAtoms], ny[kDevKindAtoms], nz[kDevKindAtoms];
  double mq[kDevKindAtoms];
  int mt_[kDevKindAtoms];
  // CBMC stage: nc (position, atom) combinations and their energies
  int nc, skip;
  double cx[kDevCombos], cy[kDevCombos], cz[kDevCombos], cq[kDevCombos];
  int ct[kDevCombos];
  double cE[kDevCombos], cR[kDevCombos], cO[kDevCombos];
  double w[kDevTrials], sI[kDevTrials], sR[kDevTrials];  // stage weights and sums
  double u[3][kDevTrials], hc[3], hb[3], hp[3];          // hedron spins: draws, frame
  int uPeek;                                             // spins read ahead in the stream
  int nLive;                                             // atoms in the box (live list)
  int cellsOff;  // a cell ran full: the rest of the launch scans the live list
  DTrial nm, om;
  int64_t done, cycleLeft;
  int cyclesDone;
  uint64_t xseq;  // the team's next exchange (1, 2, ... in every member)
  double red[kDevWaves][4];
  double res[4];
};

enum : int { kOpStop = 0, kOpNone = 1, kOpMol = 2, kOpTransfer = 3 };

#define NOINLINE __attribute__((noinline))

// Global-memory views of the box arrays, so the hot loops use global (not
// flat) loads and stores.
typedef const double __attribute__((address_space(1))) GD;
typedef double __attribute__((address_space(1))) GDW;
typedef const int __attribute__((address_space(1))) GI;
__device__ __forceinline__ GD *G(const double *p) { return (GD *)p; }
__device__ __forceinline__ GI *G(const int *p) { return (GI *)p; }
__device__ __forceinline__ GDW *GW(double *p) { return (GDW *)p; }

// ---------------------------------------------------------------- MT19937
// The generator of host_rng.hpp over the state in LDS.  Thread 0 draws;
// the workgroup computes the next block of state words ahead of need
// (MtPrepareNext), so thread 0 does not twist serially.
__device__ __forceinline__ uint32_t MtTwistWord(uint32_t a, uint32_t b, uint32_t m) {
  const uint32_t y = (a & 0x80000000u) | (b & 0x7fffffffu);
  uint32_t v = m ^ (y >> 1);
  if (b & 1u) v ^= 0x9908b0dfu;
  return v;
}
__device__ __forceinline__ uint32_t MtTemper(uint32_t v) {
  v ^= v >> 11;
  v ^= (v << 7) & 0x9d2c5680u;
  v ^= (v << 15) & 0xefc60000u;
  return v ^ (v >> 18);
}
// Serial twist of the current block in place (thread 0; only when the next
// block was not prepared).
NOINLINE __device__ void MtTwist(DevState &s) {
  uint32_t *mt = s.mt[s.mtCur];
  for (int i = 0; i < kMtN; ++i) {
    const int i1 = i + 1 < kMtN ? i + 1 : i + 1 - kMtN;
    const int im = i + kMtM < kMtN ? i + kMtM : i + kMtM - kMtN;
    mt[i] = MtTwistWord(mt[i], mt[i1], mt[im]);
  }
}
// The in-place twist, as three data-parallel passes into the other block:
// word i < 227 reads old words i, i+1, i+397; 227 <= i < 623 reads old i,
// i+1 and new i-227; word 623 reads old 623, new 0 and new 396.
__device__ void MtPrepareNext(DevState &s) {
  const uint32_t *o = s.mt[s.mtCur];
  uint32_t *nw = s.mt[s.mtCur ^ 1];
  const int t = threadIdx.x;
  if (t < kMtN - kMtM) nw[t] = MtTwistWord(o[t], o[t + 1], o[t + kMtM]);
  __syncthreads();
  if (t < kMtN - kMtM) {
    const int i = t + (kMtN - kMtM);
    nw[i] = MtTwistWord(o[i], o[i + 1], nw[i - (kMtN - kMtM)]);
  }
  __syncthreads();
  const int i = t + 2 * (kMtN - kMtM);
  if (i < kMtN - 1) nw[i] = MtTwistWord(o[i], o[i + 1], nw[i - (kMtN - kMtM)]);
  if (i == kMtN - 1) nw[i] = MtTwistWord(o[i], nw[0], nw[kMtM - 1]);
  __syncthreads();
  if (t == 0) s.nextReady = 1;
  __syncthreads();
}
__device__ __forceinline__ uint32_t MtNext(DevState &s) {
  if (s.mtIdx == kMtN) {
    if (s.nextReady) {
      s.mtCur ^= 1;
      s.nextReady = 0;
    } else {
      MtTwist(s);
    }
    s.mtIdx = 0;
  }
  return MtTemper(s.mt[s.mtCur][s.mtIdx++]);
}
// Words still unread, counting a prepared next block.
__device__ __forceinline__ int MtAvail(const DevState &s) {
  return kMtN - s.mtIdx + (s.nextReady ? kMtN : 0);
}
// The k-th next tempered word without consuming it (k < MtAvail).
__device__ __forceinline__ uint32_t MtPeek(const DevState &s, int k) {
  const int i = s.mtIdx + k;
  return MtTemper(i < kMtN ? s.mt[s.mtCur][i] : s.mt[s.mtCur ^ 1][i - kMtN]);
}
// Consumes k words (k <= MtAvail), as k MtNext calls would.
__device__ __forceinline__ void MtSkip(DevState &s, int k) {
  for (int r = k; r > 0;) {
    if (s.mtIdx == kMtN) {
      s.mtCur ^= 1;
      s.nextReady = 0;
      s.mtIdx = 0;
    }
    const int step = min(r, kMtN - s.mtIdx);
    s.mtIdx += step;
    r -= step;
  }
}
__device__ __forceinline__ uint32_t MtBelow(DevState &s, uint32_t n) {
  uint32_t mask = n;
  for (int sh = 1; sh < 32; sh <<= 1) mask |= mask >> sh;
  uint32_t v;
  do {
    v = MtNext(s) & mask;
  } while (v > n);
  return v;
}
__device__ __forceinline__ double U32Closed(uint32_t v) { return double(v) * (1.0 / 4294967295.0); }
__device__ __forceinline__ double U32HalfOpen(uint32_t v) { return double(v) * (1.0 / 4294967296.0); }
__device__ __forceinline__ double MtClosed(DevState &s) {
  return double(MtNext(s)) * (1.0 / 4294967295.0);
}
__device__ __forceinline__ double MtHalfOpen(DevState &s) {
  return double(MtNext(s)) * (1.0 / 4294967296.0);
}

struct DM3 {
  double m[3][3];
};
__device__ void Apply(const DM3 &r, double x, double y, double z, double &ox, double &oy,
                      double &oz) {
  ox = r.m[0][0] * x + r.m[0][1] * y + r.m[0][2] * z;
  oy = r.m[1][0] * x + r.m[1][1] * y + r.m[1][2] * z;
  oz = r.m[2][0] * x + r.m[2][1] * y + r.m[2][2] * z;
}
// TransformMatrix::FromAxisAngle, as AxisAngle in host_rng.hpp.
NOINLINE __device__ DM3 AxisAngleD(double theta, double ax, double ay, double az) {
  const double av[3] = {ax, ay, az};
  const double cross[3][3] = {{0.0, -az, ay}, {az, 0.0, -ax}, {-ay, ax, 0.0}};
  const double c = cos(theta), s = sin(theta);
  DM3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double tensor = av[i] * av[j];
      r.m[i][j] = (i == j ? c : 0.0) + (s * cross[i][j] + (1 - c) * tensor);
    }
  return r;
}
// TransformMatrix::UniformRandom, as UniformRotation in host_rng.hpp.
NOINLINE __device__ DM3 UniformRotationD(double u1, double u2, double u3) {
  u1 *= 2.0 * M_PI;
  u1 -= M_PI;
  u2 *= 2.0 * M_PI;
  u3 *= 2.0;
  const double r = sqrt(u3);
  const double vx = sin(u2) * r, vy = cos(u2) * r, vz = sqrt(2.0 - u3);
  const double sinTh = sin(u1), cosTh = cos(u1);
  const double sx = vx * cosTh - vy * sinTh;
  const double sy = vx * sinTh + vy * cosTh;
  DM3 m;
  m.m[0][0] = vx * sx - cosTh;
  m.m[0][1] = vx * sy - sinTh;
  m.m[0][2] = vx * vz;
  m.m[1][0] = vy * sx + sinTh;
  m.m[1][1] = vy * sy - cosTh;
  m.m[1][2] = vy * vz;
  m.m[2][0] = vz * sx;
  m.m[2][1] = vz * sy;
  m.m[2][2] = 1.0 - u3;
  return m;
}

__device__ uint32_t NumKindInBox(const DevBox &B, const DevState &s, int k, int box) {
  return s.bkStart[box * B.nkinds + k + 1] - s.bkStart[box * B.nkinds + k];
}

// PRNG::PickWeighted (src/PRNG.h:167-203).
NOINLINE __device__ int PickWeightedD(DevState &s, const double *w, int n, double total) {
  const double draw = MtClosed(s) * total;
  double sum = 0.0;
  for (int i = 0; i < n; ++i) {
    sum += w[i];
    if (sum >= draw) return i;
  }
  int lastZero = n;
  for (int i = n - 1; i > 0 && w[i] == 0.0; --i) lastZero = i;
  lastZero--;
  if (fabs(draw - total) < 0.001) return lastZero;
  return -1;
}

// Four sums over the workgroup, valid in c.res after the call.
__device__ void BlockSum4(Ctl &c, double a, double b, double d, double e) {
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  a = WaveSum(a);
  b = WaveSum(b);
  d = WaveSum(d);
  e = WaveSum(e);
  if (lane == 0) {
    c.red[wave][0] = a;
    c.red[wave][1] = b;
    c.red[wave][2] = d;
    c.red[wave][3] = e;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double r = 0.0;
    for (int w = 0; w < kDevWaves; ++w) r += c.red[w][threadIdx.x];
    c.res[threadIdx.x] = r;
  }
  __syncthreads();
}

// The team's sum of one value per member, in member order, into c.res[slot]
// of every member (all threads call it at the same point of a move; with a
// team of one it does nothing).  Each member's thread 0 publishes its value,
// then the exchange's sequence number; lanes 0..team-1 of wave 0 wait for
// the other members' numbers and read their values.  Every move of a team
// makes exactly one exchange, so the numbers run 1, 2, ... alike in every
// member and a member is at most one exchange ahead of the slowest: two
// parities of slots suffice.  A wait that outlasts kDevSpinLimit polls fails
// the launch (status kStTeam) instead of hanging.
NOINLINE __device__ void TeamExchange(const DevBox &B, DevState &s, Ctl &c, int slot) {
  if (B.team <= 1) return;
  if (threadIdx.x < kWave) {
    const int lane = threadIdx.x;
    const uint64_t seq = c.xseq;
    uint64_t *blk = B.xchg + (seq & 1) * (kDevTeamMax * 2);
    const double v = slot >= 0 ? c.res[slot] : 0.0;  // slot -1: a move without energies
    if (lane == 0) {
      __hip_atomic_exchange(blk + 2 * B.rank, __builtin_bit_cast(uint64_t, v), __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_exchange(blk + 2 * B.rank + 1, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    double mine = 0.0;
    int ok = 1;
    if (lane < B.team) {
      if (lane == B.rank) {
        mine = v;
      } else {
        for (int64_t n = 0;; ++n) {
          if (__hip_atomic_load(blk + 2 * lane + 1, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) ==
              seq)
            break;
          if (n > kDevSpinLimit) {
            ok = 0;
            break;
          }
        }
        if (ok)
          mine = __builtin_bit_cast(
              double, __hip_atomic_load(blk + 2 * lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      }
    }
    const bool all = __all(ok) != 0;
    double total = 0.0;
    for (int q = 0; q < B.team; ++q) total += __shfl(mine, q, kWave);
    if (lane == 0) {
      if (all) {
        if (slot >= 0) c.res[slot] = total;
      } else {
        s.status = kStTeam;
        c.abort = 1;
      }
      c.xseq = seq + 1;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------- cell table --
// The box's live atoms by cell, walked as CellList::EnumerateLocal walks the
// 27 cells around a position (CalculateEnergy.cpp:595,745).  A wave takes
// the 27 cells' counts in its lanes, their running sum gives each cell's
// first index, and item idx of the neighbourhood is found by a binary
// search of those starts.

__device__ __forceinline__ bool UseCells(const DevBox &B, const Ctl &c) {
  return B.cells && !c.cellsOff;
}

// Wave-wide: the 27 cells around (px, py, pz) into cells[27] and their
// first indices into start[28] (start[27] the total), in LDS; returns the
// total.  Called by every lane of the wave.
__device__ __forceinline__ int WaveCells(const DevBox &B, double px, double py, double pz, int lane,
                                         int *cells, int *start) {
  const int c0 = CellOf(px, py, pz, B.cs, B.edge);
  int cell = 0, n = 0;
  if (lane < 27) {
    cell = G(B.nbr)[c0 * 27 + lane];
    n = G(B.cCount)[cell];
  }
  int incl = n;
#pragma unroll
  for (int off = 1; off < 32; off <<= 1) {
    const int t = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += t;
  }
  __builtin_amdgcn_wave_barrier();  // the wave's last reads of these arrays are done
  if (lane < 27) {
    cells[lane] = cell;
    start[lane + 1] = incl;
  }
  if (lane == 0) start[0] = 0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  return __shfl(incl, 26, kWave);
}

// Atom idx (0 <= idx < start[27]) of a neighbourhood from WaveCells.
__device__ __forceinline__ int CellAtom(const DevBox &B, const int *cells, const int *start,
                                        int idx) {
  int lo = 0, hi = 27;  // the last cell whose first index is <= idx
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= idx)
      lo = mid;
    else
      hi = mid;
  }
  return G(B.cSlots)[cells[lo] * B.ccap + (idx - start[lo])];
}

// Thread 0: CellList::RemoveMol / AddMol for one atom (CellList.h:28-29):
// a leaving atom's slot goes to its cell's last atom, an arriving one takes
// its cell's next slot; a full cell turns the cell walk off for the rest of
// the launch (the live list stays exact).
NOINLINE __device__ void CellRemove(const DevBox &B, int a) {
  const int c0 = B.aCell[a];
  if (c0 < 0) return;
  const int slot = B.aSlot[a], last = --B.cCount[c0];
  const int moved = B.cSlots[c0 * B.ccap + last];
  B.cSlots[c0 * B.ccap + slot] = moved;
  B.aSlot[moved] = slot;
  B.aCell[a] = -1;
  B.aSlot[a] = -1;
}

NOINLINE __device__ void CellPlace(const DevBox &B, Ctl &c, int a, double x, double y, double z) {
  if (!UseCells(B, c)) return;
  const int c1 = CellOf(x, y, z, B.cs, B.edge);
  if (c1 == B.aCell[a]) return;
  CellRemove(B, a);
  const int n = B.cCount[c1];
  if (n >= B.ccap) {
    c.cellsOff = 1;
    return;
  }
  B.cSlots[c1 * B.ccap + n] = a;
  B.cCount[c1] = n + 1;
  B.aCell[a] = c1;
  B.aSlot[a] = n;
}

// ------------------------------------------------- workgroup energy phases --
__attribute__((noinline)) __device__ double ErfcOutOfLine(double x) { return erfc(x); }

// The erfcx polynomial's coefficients, copied to LDS at the kernel's start:
// read from there in the loop, they do not occupy registers.
__shared__ double sErfcx[kErfcxDeg + 1];

// ErfcPoly (device_common.hpp) with the coefficients from LDS and libm's
// erfc out of line for x > 4.
__device__ __forceinline__ double ErfcLds(double x) {
  if (!(x <= kErfcxUse)) return ErfcOutOfLine(x);
  const double t = (x - kErfcxMid) * kErfcxInvHalf;
  double p = sErfcx[kErfcxDeg];
#pragma unroll
  for (int k = kErfcxDeg - 1; k >= 0; --k) p = fma(p, t, sErfcx[k]);
  return exp(-(x * x)) * p;
}

// erfc-Coulomb at precision P; FP64 through the erfcx polynomial.
template <int P>
__device__ __forceinline__ double CoulDev(const BoxGeom &g, double d2, double qiqj) {
  if (P) return CoulombEnergyF32(g, d2, qiqj);
  const double dist = sqrt(d2);
  if (g.ewald) return qiqj * ErfcLds(g.alpha * dist) / dist;
  return qiqj / dist;
}

// The Mie repulsion for exponents other than 12, out of line so that its
// pow does not weigh on the registers of the pair loops.
__attribute__((noinline)) __device__ double PowOutOfLine(double x, double y) { return pow(x, y); }

// LJEnergy / LJEnergyF32 (device_common.hpp) with that out-of-line pow.
template <int P>
__device__ __forceinline__ double LJDev(double d2, double sigmaSq, double epsCn, double nHalf) {
  if (P) return LJEnergyF32(d2, sigmaSq, epsCn, nHalf);
  const double rRat2 = sigmaSq / d2;
  const double rRat4 = rRat2 * rRat2;
  const double attract = rRat4 * rRat2;
  const double repulse = (nHalf == 6.0) ? attract * attract : PowOutOfLine(rRat2, nHalf);
  return epsCn * (repulse - attract);
}

// FP64 phases e^{i k.r} of a trial's positions as products of per-axis
// tables e^{i (2 pi n / L_a) r_a}, |n| <= nmax_a: `np` positions given by
// (px, py, pz) in LDS, tab[p][axis][n + kDevNMax] = (cos, sin).
// With P = 1 the same LDS holds FP32 entries (float2, same indexing) and
// the products are FP32 (the mixed mode's phases).
typedef double2 PhaseTab[kDevPos][3][kTab];
template <int P>
__device__ void BuildPhases(const DevBox &B, const double *px, const double *py,
                            const double *pz, int np, PhaseTab &tab) {
  for (int e = threadIdx.x; e < np * 3 * kTab; e += kDevThreads) {
    const int p = e / (3 * kTab), a = (e / kTab) % 3, n = e % kTab - kDevNMax;
    if (n < -B.nmax[a] || n > B.nmax[a]) continue;
    const double r = a == 0 ? px[p] : (a == 1 ? py[p] : pz[p]);
    const double ca = (1.0 / B.g.L[a]) * (2.0 * M_PI);
    double sn, cs;
    sincos((ca * n) * r, &sn, &cs);
    if (P)
      reinterpret_cast<float2 *>(&tab[0][0][0])[e] = make_float2(float(cs), float(sn));
    else
      tab[p][a][n + kDevNMax] = make_double2(cs, sn);
  }
  __syncthreads();
}

__device__ __forceinline__ double2 CMulD(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 CMulF(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// e^{i k.r} of table position p for the k-vector with index triple (ix, iy, iz).
template <int P>
__device__ __forceinline__ double2 Phase(const PhaseTab &tab, int p, int ix, int iy, int iz) {
  if (P) {
    const float2 *t = reinterpret_cast<const float2 *>(&tab[0][0][0]) + p * 3 * kTab;
    const float2 e = CMulF(CMulF(t[ix + kDevNMax], t[kTab + iy + kDevNMax]),
                           t[2 * kTab + iz + kDevNMax]);
    return make_double2(e.x, e.y);
  }
  return CMulD(CMulD(tab[p][0][ix + kDevNMax], tab[p][1][iy + kDevNMax]),
               tab[p][2][iz + kDevNMax]);
}
// In-cutoff pairs found by a wave wait in its queue in LDS until 64 are
// there; the wave then evaluates 64 pair terms at once, one per lane, so
// the erfc path runs with full waves however sparse the cutoff sphere is.
constexpr int kQCap = 2 * kWave;
struct WaveQ {
  double d2[kQCap], qq[kQCap];
  int code[kQCap];  // LJ table index * 2 + (1 if the term is added, 0 if subtracted)
};
__shared__ WaveQ sQ[kDevWaves];

template <int P>
__device__ __forceinline__ void DrainQ(const BoxGeom &g, WaveQ &q, int n, int lane,
                                       const GD *sigSq, const GD *epsCn, const GD *nHalf,
                                       double &lj, double &real) {
  if (lane < n) {
    const double d2 = q.d2[lane], qq = q.qq[lane];
    const int code = q.code[lane], idx = code >> 1;
    const double sign = (code & 1) ? 1.0 : -1.0;
    if (g.elect && qq != 0.0 && !(g.rCutCoulSq < d2)) real += sign * CoulDev<P>(g, d2, qq);
    if (!(g.rCutLJSq < d2)) lj += sign * LJDev<P>(d2, sigSq[idx], epsCn[idx], nHalf[idx]);
  }
}

// Appends this lane's pair (when `in`) to the wave's queue; drains 64 when
// the queue holds that many.  Called by every lane of the wave.
template <int P>
__device__ __forceinline__ void PushQ(const BoxGeom &g, WaveQ &q, int &cnt, int lane, bool in,
                                      double d2, double qq, int code, const GD *sigSq,
                                      const GD *epsCn, const GD *nHalf, double &lj,
                                      double &real) {
  const uint64_t m = __ballot(in);
  if (in) {
    const int pos = cnt + __popcll(m & ((uint64_t(1) << lane) - 1));
    q.d2[pos] = d2;
    q.qq[pos] = qq;
    q.code[pos] = code;
  }
  cnt += __popcll(m);
  __builtin_amdgcn_wave_barrier();
  if (cnt >= kWave) {
    DrainQ<P>(g, q, kWave, lane, sigSq, epsCn, nHalf, lj, real);
    __builtin_amdgcn_wave_barrier();
    if (lane < cnt - kWave) {
      q.d2[lane] = q.d2[kWave + lane];
      q.qq[lane] = q.qq[kWave + lane];
      q.code[lane] = q.code[kWave + lane];
    }
    cnt -= kWave;
    __builtin_amdgcn_wave_barrier();
  }
}

// Each wave's neighbourhood from WaveCells.
__shared__ int sWCell[kDevWaves][27], sWStart[kDevWaves][28];

// ParticleInter (src/CalculateEnergy.cpp:727-785) for the stage's nc
// (position, atom) combinations: one wave per combination, the lanes over
// the atoms of the 27 cells around the position (or, with the cell walk
// off, over the live list), then the wave's sum.
template <int P>
NOINLINE __device__ void ComboEnergies(const DevBox &B, Ctl &c) {
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  // live atoms only; the atoms of molecule `skip` (c.st, c.len) are left out
  const int nc = c.nc, nLive = c.nLive, ntypes = B.ntypes;
  const int skipSt = c.skip >= 0 ? c.st : 0, skipEnd = c.skip >= 0 ? c.st + c.len : 0;
  const GI *live = G(B.live);
  const BoxGeom g = B.g;
  const GD *x = G(B.x), *y = G(B.y), *z = G(B.z), *q = G(B.q);
  const GI *type = G(B.type);
  const GD *sigSq = G(B.sigSq), *epsCn = G(B.epsCn), *nHalf = G(B.nHalf);
  for (int ci = wave; ci < nc; ci += kDevWaves) {
    const double px = c.cx[ci], py = c.cy[ci], pz = c.cz[ci], charge = c.cq[ci];
    const int atype = c.ct[ci];
    double lj = 0.0, real = 0.0, ov = 0.0;
    WaveQ &wq = sQ[wave];
    int cnt = 0;
    const bool walk = UseCells(B, c);
    const int total = walk ? WaveCells(B, px, py, pz, lane, sWCell[wave], sWStart[wave]) : nLive;
    for (int base = 0; base < total; base += kWave) {
      const int li = base + lane;
      bool in = false;
      double d2 = 0.0, qiqj = 0.0;
      int code = 0;
      if (li < total) {
        const int j = walk ? CellAtom(B, sWCell[wave], sWStart[wave], li) : live[li];
        if (!(j >= skipSt && j < skipEnd)) {
          d2 = DistSq(g, px, py, pz, x[j], y[j], z[j]);
          in = g.rCutSq > d2;
          if (in && d2 < g.rCutLowSq) ov = 1.0;
          qiqj = q[j] * charge * kQQFact;
          code = (atype + type[j] * ntypes) * 2 + 1;
        }
      }
      PushQ<P>(g, wq, cnt, lane, in, d2, qiqj, code, sigSq, epsCn, nHalf, lj, real);
    }
    DrainQ<P>(g, wq, cnt, lane, sigSq, epsCn, nHalf, lj, real);
    __builtin_amdgcn_wave_barrier();
    lj = WaveSum(lj);
    real = WaveSum(real);
    ov = WaveSum(ov);
    if (lane == 0) {
      c.cE[ci] = lj;
      c.cR[ci] = real;
      c.cO[ci] = ov;
    }
  }
  __syncthreads();
}

// Waves of a translate/rotate trial that take the pair terms while the rest
// take the k-vectors, in proportion to 2 N_live distance tests per atom
// against kRecipWeight K phase products per atom (at least two waves each).
constexpr double kRecipWeight = 1.0;
__device__ __forceinline__ int PairWaves(int nLive, int K) {
  const double wp = 2.0 * double(nLive), wr = kRecipWeight * double(K);
  const int pw = int(double(kDevWaves) * wp / (wp + wr) + 0.5);
  return pw < 2 ? 2 : (pw > kDevWaves - 2 ? kDevWaves - 2 : pw);
}

// One translate / rotate trial's energies: every atom against the old and
// new positions (CalculateEnergy::MoleculeInter, CalculateEnergy.cpp:581-686)
// and every k-vector's delta (Ewald::MolReciprocal, Ewald.cpp:409-473), the
// pair waves and the k-vector waves running side by side.
template <int P>
NOINLINE __device__ void MolTrialEnergies(const DevBox &B, DevState &s, Ctl &c, PhaseTab &tab) {
  double lj = 0.0, real = 0.0, ov = 0.0, e = 0.0;
  const int st = c.st, len = c.len, ntypes = B.ntypes;
  const BoxGeom g = B.g;
  const bool doRecip = c.doRecip;
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  // cell walk: wave sp takes position sp (atom sp >> 1; old when even, new
  // when odd) and finds the atoms of the 27 cells around it
  __shared__ int sPosTot[kDevPos];
  const bool walk = UseCells(B, c);
  if (walk && wave < 2 * len) {
    const int p = wave >> 1;
    const bool isNew = wave & 1;
    const int tot = WaveCells(B, isNew ? c.nx[p] : c.ox[p], isNew ? c.ny[p] : c.oy[p],
                              isNew ? c.nz[p] : c.oz[p], lane, sWCell[wave], sWStart[wave]);
    if (lane == 0) sPosTot[wave] = tot;
  }
  if (doRecip) {  // new positions are table rows [0, len), old ones [kDevKindAtoms, +len)
    __shared__ double tpx[kDevPos], tpy[kDevPos], tpz[kDevPos];
    if (threadIdx.x < kDevKindAtoms) {
      const int p = threadIdx.x;
      tpx[p] = c.nx[p];
      tpy[p] = c.ny[p];
      tpz[p] = c.nz[p];
      tpx[kDevKindAtoms + p] = c.ox[p];
      tpy[kDevKindAtoms + p] = c.oy[p];
      tpz[kDevKindAtoms + p] = c.oz[p];
    }
    __syncthreads();
    BuildPhases<P>(B, tpx, tpy, tpz, kDevPos, tab);  // ends with a barrier
  } else if (walk) {
    __syncthreads();
  }
  // (position, neighbour atom) items of the cell walk: one distance test
  // each, weighted twice a live-list test in the wave split
  int items = 0;
  if (walk)
    for (int sp = 0; sp < 2 * len; ++sp) items += sPosTot[sp];
  // the same split in every member of a team (its pair sums, computed by
  // each member in full, must agree to the last bit)
  const int nPair = walk ? (items + len - 1) / len : c.nLive;
  const int pw = doRecip ? PairWaves(nPair, (B.K + B.team - 1) / B.team) : kDevWaves;
  if (wave < pw && walk) {
    const GD *x = G(B.x), *y = G(B.y), *z = G(B.z), *q = G(B.q);
    const GI *type = G(B.type);
    const GD *sigSq = G(B.sigSq), *epsCn = G(B.epsCn), *nHalf = G(B.nHalf);
    WaveQ &wq = sQ[wave];
    int cnt = 0;
    for (int base = wave * kWave; base < items; base += pw * kWave) {
      const int it = base + lane;
      bool in = false;
      double d2 = 0.0, qq = 0.0;
      int code = 0;
      if (it < items) {
        int sp = 0, first = 0;
        while (it >= first + sPosTot[sp]) first += sPosTot[sp++];
        const int j = CellAtom(B, sWCell[sp], sWStart[sp], it - first);
        if (!(j >= st && j < st + len)) {
          const int p = sp >> 1;
          const bool isNew = sp & 1;
          const double xj = x[j], yj = y[j], zj = z[j];
          d2 = isNew ? DistSq(g, c.nx[p], c.ny[p], c.nz[p], xj, yj, zj)
                     : DistSq(g, c.ox[p], c.oy[p], c.oz[p], xj, yj, zj);
          in = g.rCutSq > d2;
          if (isNew && in && d2 < g.rCutLowSq) ov = 1.0;
          qq = c.mq[p] * q[j] * kQQFact;
          code = (c.mt_[p] + type[j] * ntypes) * 2 + (isNew ? 1 : 0);
        }
      }
      PushQ<P>(g, wq, cnt, lane, in, d2, qq, code, sigSq, epsCn, nHalf, lj, real);
    }
    DrainQ<P>(g, wq, cnt, lane, sigSq, epsCn, nHalf, lj, real);
    __builtin_amdgcn_wave_barrier();
  } else if (wave < pw) {
    // (atom p, old) then (atom p, new) for each p, the molecule's data read
    // from LDS; in-cutoff pairs go through the wave's queue
    const GD *x = G(B.x), *y = G(B.y), *z = G(B.z), *q = G(B.q);
    const GI *type = G(B.type);
    const GD *sigSq = G(B.sigSq), *epsCn = G(B.epsCn), *nHalf = G(B.nHalf);
    WaveQ &wq = sQ[wave];
    int cnt = 0;
    const int nLive = c.nLive;
    const GI *live = G(B.live);
    for (int base = wave * kWave; base < nLive; base += pw * kWave) {
      const int li = base + lane;
      const int j = li < nLive ? live[li] : 0;
      const bool valid = li < nLive && !(j >= st && j < st + len);
      double xj = 0.0, yj = 0.0, zj = 0.0, qj = 0.0;
      int tj = 0;
      if (valid) {
        xj = x[j];
        yj = y[j];
        zj = z[j];
        qj = q[j];
        tj = type[j];
      }
      for (int sp = 0; sp < 2 * len; ++sp) {
        const int p = sp >> 1;
        const bool isNew = sp & 1;
        double d2 = 0.0;
        bool in = false;
        if (valid) {
          d2 = isNew ? DistSq(g, c.nx[p], c.ny[p], c.nz[p], xj, yj, zj)
                     : DistSq(g, c.ox[p], c.oy[p], c.oz[p], xj, yj, zj);
          in = g.rCutSq > d2;
          if (isNew && in && d2 < g.rCutLowSq) ov = 1.0;
        }
        PushQ<P>(g, wq, cnt, lane, in, d2, c.mq[p] * qj * kQQFact,
                 (c.mt_[p] + tj * ntypes) * 2 + (isNew ? 1 : 0), sigSq, epsCn, nHalf, lj, real);
      }
    }
    DrainQ<P>(g, wq, cnt, lane, sigSq, epsCn, nHalf, lj, real);
    __builtin_amdgcn_wave_barrier();
  } else {
    // every k-vector's new structure factor and energy term, beside the pairs
    const GD *pref = G(B.pref);
    const GD *sumR = G(B.sumR), *sumI = G(B.sumI);
    const GI *kpack = G(B.kpack);
    GDW *newR = GW(B.newR), *newI = GW(B.newI);
    const int r0 = pw * kWave, rn = kDevThreads - r0;
    for (int k = B.k0 + int(threadIdx.x) - r0; k < B.k1; k += rn) {
      double sRN = 0.0, sIN = 0.0, sRO = 0.0, sIO = 0.0;
      const int code = kpack[k];
      const int ix = code & 0xff, iy = ((code >> 8) & 0xff) - 128, iz = ((code >> 16) & 0xff) - 128;
#pragma unroll 1
      for (int p = 0; p < len; ++p) {
        const double qp = c.mq[p];
        if (fabs(qp) < 0.000000001) continue;
        const double2 en = Phase<P>(tab, p, ix, iy, iz);
        const double2 eo = Phase<P>(tab, kDevKindAtoms + p, ix, iy, iz);
        sRN += qp * en.x;
        sIN += qp * en.y;
        sRO += qp * eo.x;
        sIO += qp * eo.y;
      }
      const double R = sumR[k] + (sRN - sRO);
      const double I = sumI[k] + (sIN - sIO);
      newR[k] = R;
      newI[k] = I;
      e += (R * R + I * I) * pref[k];
    }
  }
  BlockSum4(c, lj, real, e, ov);
  TeamExchange(B, s, c, 2);  // the members' k-vector ranges
}

// Ewald::SwapDestRecip / SwapSourceRecip (Ewald.cpp:478-531,657-710).
template <int P>
NOINLINE __device__ void SwapRecipEnergy(const DevBox &B, DevState &s, Ctl &c, PhaseTab &tab) {
  double e = 0.0;
  const int len = c.len, sign = c.sign;
  const GD *pref = G(B.pref);
  const GD *sumR = G(B.sumR), *sumI = G(B.sumI);
  const GI *kpack = G(B.kpack);
  GDW *newR = GW(B.newR), *newI = GW(B.newI);
  BuildPhases<P>(B, c.nx, c.ny, c.nz, kDevKindAtoms, tab);
  for (int k = B.k0 + int(threadIdx.x); k < B.k1; k += kDevThreads) {
    double sR = 0.0, sI = 0.0;
    const int code = kpack[k];
    const int ix = code & 0xff, iy = ((code >> 8) & 0xff) - 128, iz = ((code >> 16) & 0xff) - 128;
#pragma unroll 1
    for (int p = 0; p < len; ++p) {
      const double qp = c.mq[p];
      if (fabs(qp) < 0.000000001) continue;
      const double2 e1 = Phase<P>(tab, p, ix, iy, iz);
      sR += qp * e1.x;
      sI += qp * e1.y;
    }
    const double R = (sign > 0) ? sumR[k] + sR : sumR[k] - sR;
    const double I = (sign > 0) ? sumI[k] + sI : sumI[k] - sI;
    newR[k] = R;
    newI[k] = I;
    e += (R * R + I * I) * pref[k];
  }
  BlockSum4(c, e, 0.0, 0.0, 0.0);
  TeamExchange(B, s, c, 0);
}

// An accepted trial's structure factors become the box's (thread 0): the
// trial phases wrote every k of newR/newI, so the two arrays trade places
// instead of K values being copied.
__device__ void SwapRecip(DevBox &B) {
  double *t = B.sumR;
  B.sumR = B.newR;
  B.newR = t;
  t = B.sumI;
  B.sumI = B.newI;
  B.newI = t;
}

// ------------------------------------------------- thread-0 serial pieces --
__device__ void Total(DevState &s) {
  s.pot[6] = s.pot[0] + s.pot[1] + s.pot[2] + s.pot[3] + s.pot[4] + s.pot[5];
}

NOINLINE __device__ void Record(const DevBox &B, int64_t i, int kind, int mol, int acc, int ovl,
                                double e0, double e1, double e2) {
  if (!B.tKind) return;
  B.tKind[i] = kind;
  B.tMol[i] = mol;
  B.tAcc[i] = acc;
  B.tOvl[i] = ovl;
  B.tE[3 * i] = e0;
  B.tE[3 * i + 1] = e1;
  B.tE[3 * i + 2] = e2;
}

// MoveSettings::AdjustMoves / Adjust (src/MoveSettings.cpp:148-162,205-258).
NOINLINE __device__ void AdjustMovesD(const DevBox &B, DevState &s, uint64_t step) {
  if ((step + 1) % B.perAdjust != 0) return;
  const double lmin = fmin(B.g.L[0], fmin(B.g.L[1], B.g.L[2]));
  for (int mv = 0; mv < 2; ++mv)
    for (int k = 0; k < B.nkinds; ++k) {
      double sc = s.scale[mv][k];
      if (s.tTries[mv][k] > 0) {
        const double acc = double(s.tAcc[mv][k]) / double(s.tTries[mv][k]);
        const double fract = acc / 0.5;
        sc *= (fract > 0.0) ? fract : 0.5;
      }
      const double lo = (mv == kDisplace) ? 1.0e-10 : 1.0e-6;
      const double hi = (mv == kDisplace) ? (lmin * 0.5) - 1.0e-7 : M_PI - 1.0e-7;
      if (sc < lo) sc = lo;
      if (sc > hi) sc = hi;
      s.scale[mv][k] = sc;
      s.tAcc[mv][k] = 0;
      s.tTries[mv][k] = 0;
    }
}

// Top of the move loop: cycle bookkeeping, the adjustment, and
// PRNG::PickArbDist over the move kinds (src/PRNG.h:153-163).
NOINLINE __device__ void LoopTop(const DevBox &B, DevState &S, Ctl &c) {
  c.op = kOpStop;
  c.commit = 0;
  bool go = !c.abort && c.done < B.nsteps && S.status == 0;
  if (go && B.cycles > 0) {
    if (c.cycleLeft == 0) {  // a cycle ended: sample, maybe stop
      S.sumN[0] += B.nkinds > 0 ? double(NumKindInBox(B, S, 0, 0)) : 0.0;
      S.sumN[1] += B.nkinds > 1 ? double(NumKindInBox(B, S, 1, 0)) : 0.0;
      S.samples++;
      c.cyclesDone++;
      c.cycleLeft = -1;
      go = c.cyclesDone < B.cycles;
    }
    if (go && c.cycleLeft < 0) {
      const int64_t nIn = int64_t(S.bkStart[B.nkinds] - S.bkStart[0]);
      c.cycleLeft = nIn > 1 ? nIn : 1;
      if (c.cycleLeft > B.nsteps - c.done) c.cycleLeft = B.nsteps - c.done;
    }
  }
  if (!go) return;
  AdjustMovesD(B, S, uint64_t(B.step0 + c.done));
  const double draw0 = MtClosed(S) * 1.0;
  double sum = 0.0, prev = 0.0;
  uint32_t pick = 0;
  for (pick = 0; pick < uint32_t(B.nMoveTypes) && sum < draw0; ++pick) {
    prev = sum;
    sum += B.moveW[pick];
  }
  if (pick != 0) --pick;
  c.subDraw = draw0 - prev;
  c.pick = int(pick);
  if (pick == kMolTransfer && B.nMoveTypes == 13)
    c.op = kOpTransfer;
  else if (pick == kDisplace || pick == kRotate)
    c.op = kOpMol;
  else
    S.status = kStMoveKind;
}

// DCSingle::BuildNew / BuildOld (DCSingle.cpp:23-100): cbmcFirst positions
// uniform in the box (the old build's first is the actual atom).
NOINLINE __device__ void SinglePrep(const DevBox &B, DevState &s, Ctl &c, const DTrial &t,
                                    int atom, bool isNew) {
  const int n = B.cbmcFirst, kind = c.kind;
  for (int i = 0; i < n; ++i) {
    const double xx = MtHalfOpen(s) * B.g.L[0];
    const double yy = MtHalfOpen(s) * B.g.L[1];
    const double zz = MtHalfOpen(s) * B.g.L[2];
    c.cx[i] = xx;
    c.cy[i] = yy;
    c.cz[i] = zz;
    c.ct[i] = B.kindT[kind][atom];
    c.cq[i] = B.kindQ[kind][atom];
  }
  if (!isNew) {
    c.cx[0] = Wrap(t.px[atom], B.g.L[0]);
    c.cy[0] = Wrap(t.py[atom], B.g.L[1]);
    c.cz[0] = Wrap(t.pz[atom], B.g.L[2]);
  }
  c.nc = t.box == 0 ? n : 0;
  for (int i = c.nc; i < n; ++i) c.cE[i] = c.cR[i] = c.cO[i] = 0.0;
}

// Boltzmann weights of a stage's n trials, one trial per thread.
NOINLINE __device__ void StageWeights(const DevBox &B, Ctl &c, int n, bool hedron) {
  const int i = threadIdx.x;
  if (i < n) {
    if (hedron) {  // atom 2's terms first, then atom 1's, as the host accumulates them
      c.sI[i] = (0.0 + c.cE[i]) + c.cE[n + i];
      c.sR[i] = (0.0 + c.cR[i]) + c.cR[n + i];
      c.w[i] = exp(-B.beta * (c.sI[i] + c.sR[i]));
    } else {
      const double inter = 0.0 + c.cE[i], real = 0.0 + c.cR[i];
      c.w[i] = exp(-1 * B.beta * (inter + real));
    }
  }
  __syncthreads();
}

NOINLINE __device__ void SingleFinish(const DevBox &B, DevState &s, Ctl &c, DTrial &t, int atom,
                                      bool isNew) {
  const int n = B.cbmcFirst;
  double stepWeight = 0.0;
  for (int i = 0; i < n; ++i) stepWeight += c.w[i];
  int win = 0;
  if (isNew) {
    win = PickWeightedD(s, c.w, n, stepWeight);
    if (win < 0) {
      s.status = kStWeights;
      c.abort = 1;
      win = 0;
    }
    t.px[atom] = c.cx[win];
    t.py[atom] = c.cy[win];
    t.pz[atom] = c.cz[win];
  }
  t.overlap = t.overlap || (c.cO[win] > 0.0);
  t.weight *= stepWeight / n;
  t.inter += 0.0 + c.cE[win];
  t.real += 0.0 + c.cR[win];
}

// SinglePrep with the trial positions computed one per thread from words
// read ahead in the stream (when the stream holds them).
NOINLINE __device__ void SinglePrepParallel(const DevBox &B, DevState &s, Ctl &c,
                                            const DTrial &t, int atom, bool isNew) {
  const int n = B.cbmcFirst, kind = c.kind, i = threadIdx.x;
  if (i < n) {
    c.cx[i] = U32HalfOpen(MtPeek(s, 3 * i)) * B.g.L[0];
    c.cy[i] = U32HalfOpen(MtPeek(s, 3 * i + 1)) * B.g.L[1];
    c.cz[i] = U32HalfOpen(MtPeek(s, 3 * i + 2)) * B.g.L[2];
    c.ct[i] = B.kindT[kind][atom];
    c.cq[i] = B.kindQ[kind][atom];
  }
  __syncthreads();
  if (i == 0) {
    MtSkip(s, 3 * n);
    if (!isNew) {
      c.cx[0] = Wrap(t.px[atom], B.g.L[0]);
      c.cy[0] = Wrap(t.py[atom], B.g.L[1]);
      c.cz[0] = Wrap(t.pz[atom], B.g.L[2]);
    }
    c.nc = t.box == 0 ? n : 0;
    for (int k = c.nc; k < n; ++k) c.cE[k] = c.cR[k] = c.cO[k] = 0.0;
  }
}

template <int P>
__device__ void SingleStage(const DevBox &B, DevState &s, Ctl &c, DTrial &t, int atom,
                            bool isNew) {
  if (MtAvail(s) >= 3 * B.cbmcFirst)
    SinglePrepParallel(B, s, c, t, atom, isNew);
  else if (threadIdx.x == 0)
    SinglePrep(B, s, c, t, atom, isNew);
  __syncthreads();
  ComboEnergies<P>(B, c);
  StageWeights(B, c, B.cbmcFirst, false);
  if (threadIdx.x == 0) SingleFinish(B, s, c, t, atom, isNew);
  __syncthreads();
}

// DCFreeHedron (DCFreeHedron.cpp:48-240) for a rigid three-site kind grown
// from atom 0: the fixed angle's trials (all of weight 1, the winner still
// drawn when growing), the seed atom, then cbmcNth spins of the rigid frame.
NOINLINE __device__ void HedronTheta(const DevBox &B, DevState &s, Ctl &c, bool isNew) {
  double tw = 0.0;
  const int n = isNew ? B.cbmcAng : B.cbmcAng - 1;
  for (int i = 0; i < n; ++i) tw += 1.0;
  c.thetaW = tw;
  if (isNew) (void)MtClosed(s);  // DCHedron::FreeAnglesNew's PickWeighted, equal weights
}

// Thread 0: the frame to spin (the bonded and previous atoms relative to
// atom 0) and the spins' draws, in the host's order (trial n-1 first).
NOINLINE __device__ void HedronDraw(const DevBox &B, DevState &s, Ctl &c, const DTrial &t,
                                    bool isNew) {
  const int n = B.cbmcNth, kind = c.kind;
  const double cx = t.px[0], cy = t.py[0], cz = t.pz[0];
  c.hc[0] = cx;
  c.hc[1] = cy;
  c.hc[2] = cz;
  if (isNew) {
    const double th = B.hedTh[kind];
    const double sinTh = sin(th);
    double pbx = B.hedB[kind], pby = B.hedB[kind], pbz = B.hedB[kind];
    pbx *= sinTh * cos(0.0);
    pby *= sinTh * sin(0.0);
    pbz *= cos(th);
    double ppx = B.hedP[kind], ppy = B.hedP[kind], ppz = B.hedP[kind];
    ppx *= sin(0.0) * cos(0.0);
    ppy *= sin(0.0) * sin(0.0);
    ppz *= cos(0.0);
    c.hb[0] = pbx;
    c.hb[1] = pby;
    c.hb[2] = pbz;
    c.hp[0] = ppx;
    c.hp[1] = ppy;
    c.hp[2] = ppz;
  } else {
    c.thetaW += exp(-1 * B.beta * (0.0 + 0.0));  // DCHedron::IncorporateOld
    c.hb[0] = Unwrap(t.px[2], cx, B.g.L[0], B.g.half[0]) + (-cx);
    c.hb[1] = Unwrap(t.py[2], cy, B.g.L[1], B.g.half[1]) + (-cy);
    c.hb[2] = Unwrap(t.pz[2], cz, B.g.L[2], B.g.half[2]) + (-cz);
    c.hp[0] = Unwrap(t.px[1], cx, B.g.L[0], B.g.half[0]) + (-cx);
    c.hp[1] = Unwrap(t.py[1], cy, B.g.L[1], B.g.half[1]) + (-cy);
    c.hp[2] = Unwrap(t.pz[1], cz, B.g.L[2], B.g.half[2]) + (-cz);
  }
  const int lo = isNew ? 0 : 1;
  c.uPeek = MtAvail(s) >= 3 * (n - lo);
  if (!c.uPeek)
    for (int lj = n; lj-- > lo;) {
      c.u[0][lj] = MtClosed(s);
      c.u[1][lj] = MtClosed(s);
      c.u[2][lj] = MtClosed(s);
    }
  c.nc = t.box == 0 ? 2 * n : 0;
}

// Trial i (one per thread): spin the frame, place atoms 2 and 1 at
// c[i] and c[n + i]; the old build's trial 0 is the actual frame.
NOINLINE __device__ void HedronSpin(const DevBox &B, DevState &s, Ctl &c, bool isNew) {
  const int n = B.cbmcNth, kind = c.kind, i = threadIdx.x;
  if (i < n) {
    const double cx = c.hc[0], cy = c.hc[1], cz = c.hc[2];
    double bx = c.hb[0], by = c.hb[1], bz = c.hb[2], qx = c.hp[0], qy = c.hp[1], qz = c.hp[2];
    if (isNew || i > 0) {
      double u1 = c.u[0][i], u2 = c.u[1][i], u3 = c.u[2][i];
      if (c.uPeek) {  // trial n-1 draws first, three words each
        const int k = 3 * (n - 1 - i);
        u1 = U32Closed(MtPeek(s, k));
        u2 = U32Closed(MtPeek(s, k + 1));
        u3 = U32Closed(MtPeek(s, k + 2));
      }
      const DM3 spin = UniformRotationD(u1, u2, u3);
      double ax, ay, az;
      Apply(spin, bx, by, bz, ax, ay, az);
      bx = ax;
      by = ay;
      bz = az;
      Apply(spin, qx, qy, qz, ax, ay, az);
      qx = ax;
      qy = ay;
      qz = az;
    }
    c.cx[i] = Wrap(bx + cx, B.g.L[0]);
    c.cy[i] = Wrap(by + cy, B.g.L[1]);
    c.cz[i] = Wrap(bz + cz, B.g.L[2]);
    c.ct[i] = B.kindT[kind][2];
    c.cq[i] = B.kindQ[kind][2];
    c.cx[n + i] = Wrap(qx + cx, B.g.L[0]);
    c.cy[n + i] = Wrap(qy + cy, B.g.L[1]);
    c.cz[n + i] = Wrap(qz + cz, B.g.L[2]);
    c.ct[n + i] = B.kindT[kind][1];
    c.cq[n + i] = B.kindQ[kind][1];
    if (c.nc == 0) c.cE[i] = c.cR[i] = c.cO[i] = c.cE[n + i] = c.cR[n + i] = c.cO[n + i] = 0.0;
  }
  __syncthreads();
  if (i == 0 && c.uPeek) MtSkip(s, 3 * (isNew ? n : n - 1));
}

NOINLINE __device__ void HedronFinish(const DevBox &B, DevState &s, Ctl &c, DTrial &t,
                                      bool isNew) {
  const int n = B.cbmcNth;
  double stepWeight = 0.0;
  for (int i = 0; i < n; ++i) stepWeight += c.w[i];
  int win = 0;
  if (isNew) {
    win = PickWeightedD(s, c.w, n, stepWeight);
    if (win < 0) {
      s.status = kStWeights;
      c.abort = 1;
      win = 0;
    }
    t.px[2] = c.cx[win];
    t.py[2] = c.cy[win];
    t.pz[2] = c.cz[win];
    t.px[1] = c.cx[n + win];
    t.py[1] = c.cy[n + win];
    t.pz[1] = c.cz[n + win];
  }
  t.overlap = t.overlap || (c.cO[win] > 0.0 || c.cO[n + win] > 0.0);
  t.inter += c.sI[win];
  t.real += c.sR[win];
  t.weight *= c.thetaW * 1.0;
  t.weight *= stepWeight / n;
}

template <int P>
__device__ void HedronStage(const DevBox &B, DevState &s, Ctl &c, DTrial &t, bool isNew) {
  if (threadIdx.x == 0) HedronTheta(B, s, c, isNew);
  __syncthreads();  // the stream position is read by every thread next
  SingleStage<P>(B, s, c, t, 0, isNew);
  if (c.abort) return;
  if (threadIdx.x == 0) HedronDraw(B, s, c, t, isNew);
  __syncthreads();
  HedronSpin(B, s, c, isNew);
  ComboEnergies<P>(B, c);
  StageWeights(B, c, B.cbmcNth, true);
  if (threadIdx.x == 0) HedronFinish(B, s, c, t, isNew);
  __syncthreads();
}

// Ewald::SwapCorrection (src/Ewald.cpp:1311-1335) of a trial molecule.
NOINLINE __device__ double SwapCorrectionD(const DevBox &B, const DTrial &t, int kind) {
  const int len = B.kindLen[kind];
  double correction = 0.0;
  for (int i = 0; i < len; ++i)
    for (int j = i + 1; j < len; ++j) {
      double dx = t.px[i] - t.px[j], dy = t.py[i] - t.py[j], dz = t.pz[i] - t.pz[j];
      if (dx > B.g.half[0]) dx -= B.g.L[0];
      else if (dx < -B.g.half[0]) dx += B.g.L[0];
      if (dy > B.g.half[1]) dy -= B.g.L[1];
      else if (dy < -B.g.half[1]) dy += B.g.L[1];
      if (dz > B.g.half[2]) dz -= B.g.L[2];
      else if (dz < -B.g.half[2]) dz += B.g.L[2];
      const double dist = sqrt(dx * dx + dy * dy + dz * dz);
      correction -= (B.kindQ[kind][i] * B.kindQ[kind][j] * erf(B.g.alpha * dist) / dist);
    }
  return kQQFact * correction;
}

// CalculateEnergy::MoleculeTailChange (CalculateEnergy.cpp:788-808).
__device__ double TailChange(const DevBox &B, const DevState &s, int kind, bool add) {
  const double sign = add ? 1.0 : -1.0;
  double delta = 0.0;
  for (int j = 0; j < B.nkinds; ++j)
    delta += B.pairEnCorr[j * B.nkinds + kind] * (sign * 2.0 * double(s.kindCount[j]) * B.volInv);
  delta += B.pairEnCorr[kind * B.nkinds + kind] * B.volInv;
  return delta;
}

// MoleculeTransfer: PickMolAndBoxPair2 (PRNG.h:450-455) and Prep.
NOINLINE __device__ void TransferPick(const DevBox &B, DevState &s, Ctl &c, int64_t i) {
  const int K = B.nkinds;
  c.op = kOpTransfer;
  double subDraw = c.subDraw;
  const double boxDiv = B.moveW[kMolTransfer] / 2;
  uint32_t src = uint32_t(subDraw / boxDiv);
  if (src >= 2) src--;
  subDraw -= src * boxDiv;
  const int dst = src == 0 ? 1 : 0;
  const uint32_t mkTot = uint32_t(B.nSwapKinds);
  const double molDiv = boxDiv / mkTot;
  uint32_t k = uint32_t(subDraw / molDiv);
  if (k == mkTot) k = mkTot - 1;
  const int kind = B.swapKind[k];
  c.src = int(src);
  c.dst = dst;
  c.kind = kind;
  c.mol = -1;
  if (NumKindInBox(B, s, kind, int(src)) == 0) {
    if (src == 1) {
      s.status = kStEmptyReservoir;
      c.abort = 1;
    } else {
      s.swapTries[dst * K + kind]++;
      Record(B, i, kMolTransfer, -1, 0, 0, 0.0, 0.0, 0.0);
    }
    c.op = kOpNone;
    return;
  }
  const uint32_t off = MtBelow(s, NumKindInBox(B, s, kind, int(src)) - 1);
  const int m = int(B.lookup[s.bkStart[int(src) * K + kind] + off]);
  c.mol = m;
  c.st = B.molStart[m];
  c.len = B.molStart[m + 1] - c.st;
  c.skip = m;
  c.nm.box = dst;
  c.om.box = int(src);
  c.nm.weight = c.om.weight = 1.0;
  c.nm.overlap = c.om.overlap = 0;
  c.nm.inter = c.nm.real = c.om.inter = c.om.real = 0.0;
  for (int p = 0; p < kDevKindAtoms; ++p) {
    c.nm.px[p] = c.nm.py[p] = c.nm.pz[p] = 0.0;
    c.om.px[p] = p < c.len ? B.x[c.st + p] : 0.0;
    c.om.py[p] = p < c.len ? B.y[c.st + p] : 0.0;
    c.om.pz[p] = p < c.len ? B.z[c.st + p] : 0.0;
  }
  if (c.len == 1)
    (void)MtBelow(s, 1);  // DCLinear::Build: forward or backward
  else
    (void)MtBelow(s, 0);  // DCGraph::Build: its only node
}

// MoleculeTransfer::CalcEn (MoleculeTransfer.h:106-133) up to the
// reciprocal change.
NOINLINE __device__ void TransferCalc(const DevBox &B, DevState &s, Ctl &c) {
  const int src = c.src, dst = c.dst, kind = c.kind;
  c.tcLose = c.tcGain = 0.0;
  c.wTc = 1.0;
  if (B.lrc) {
    c.tcLose = src == 0 ? TailChange(B, s, kind, false) : 0.0;
    c.tcGain = dst == 0 ? TailChange(B, s, kind, true) : 0.0;
    c.wTc = exp(-1.0 * B.beta * (c.tcGain + c.tcLose));
  }
  c.tried = c.nm.weight > kSmallWeight && !c.nm.overlap;
  c.corrNew = c.corrOld = c.selfNew = c.selfOld = 0.0;
  c.doRecip = c.tried && B.g.ewald;
  if (c.doRecip) {
    c.corrNew = dst == 0 ? SwapCorrectionD(B, c.nm, kind) : 0.0;
    c.corrOld = src == 0 ? SwapCorrectionD(B, c.om, kind) : 0.0;
    c.selfNew = dst == 0 ? B.selfKind[kind] : 0.0;
    c.selfOld = src == 0 ? B.selfKind[kind] : 0.0;
    const DTrial &t = dst == 0 ? c.nm : c.om;
    for (int p = 0; p < kDevKindAtoms; ++p) {
      c.nx[p] = t.px[p];
      c.ny[p] = t.py[p];
      c.nz[p] = t.pz[p];
      c.mq[p] = p < c.len ? B.kindQ[kind][p] : 0.0;
    }
    c.sign = dst == 0 ? 1 : -1;
  }
}

// MoleculeTransfer::Accept (MoleculeTransfer.h:167-246) with GetCoeff's
// fugacities (143-165): the decision and the running energies.
NOINLINE __device__ void TransferAccept(const DevBox &B, DevState &s, Ctl &c) {
  const int K = B.nkinds, src = c.src, dst = c.dst, kind = c.kind;
  c.recipLose = c.recipGain = 0.0;
  double wRecip = 1.0;
  if (c.doRecip) {
    const double enew = c.res[0];
    if (dst == 0)
      c.recipGain = enew - s.pot[3];
    else
      c.recipLose = enew - s.pot[3];
    wRecip = exp(-1.0 * B.beta *
                 (c.recipGain + c.recipLose + c.corrNew - c.corrOld + c.selfNew - c.selfOld));
  }
  double coeff;
  if (src == 0)
    coeff = double(s.swapCount[0 * K + kind]) * B.volInv / (B.beta * B.fugacity[kind]);
  else
    coeff = B.volume / double(s.swapCount[dst * K + kind] + 1) * (B.beta * B.fugacity[kind]);
  const double Wrat = c.nm.weight / c.om.weight * c.wTc * wRecip;
  bool result = false;
  if (c.tried) result = MtClosed(s) < coeff * Wrat;
  c.commit = result;
  if (!result) return;
  if (src == 0) {
    s.pot[1] += c.tcLose;
    s.pot[0] -= c.om.inter;
    s.pot[2] -= c.om.real;
    s.pot[3] += c.recipLose;
    s.pot[5] -= c.corrOld;
    s.pot[4] -= c.selfOld;
  } else {
    s.pot[1] += c.tcGain;
    s.pot[0] += c.nm.inter;
    s.pot[2] += c.nm.real;
    s.pot[3] += c.recipGain;
    s.pot[5] += c.corrNew;
    s.pot[4] += c.selfNew;
  }
  c.section = src * K + kind;
  c.found = -1;
}

// The accepted transfer's bookkeeping: COM::SetNew (src/COM.h:63-88),
// MoleculeLookup::Shift (src/MoleculeLookup.cpp:165-198), counts; and the
// move's record.
NOINLINE __device__ void TransferFinish(const DevBox &B, DevState &s, Ctl &c, int64_t i) {
  const int K = B.nkinds, src = c.src, dst = c.dst, kind = c.kind, m = c.mol, len = c.len;
  if (c.commit) {
    const double rx = c.nm.px[0], ry = c.nm.py[0], rz = c.nm.pz[0];
    double cx = 0, cy = 0, cz = 0;
    for (int p = 0; p < len; ++p) {
      cx += Unwrap(c.nm.px[p], rx, B.g.L[0], B.g.half[0]);
      cy += Unwrap(c.nm.py[p], ry, B.g.L[1], B.g.half[1]);
      cz += Unwrap(c.nm.pz[p], rz, B.g.L[2], B.g.half[2]);
    }
    const double inv = 1.0 / double(len);
    B.com[m] = Wrap(cx * inv, B.g.L[0]);
    B.com[B.nmols + m] = Wrap(cy * inv, B.g.L[1]);
    B.com[2 * B.nmols + m] = Wrap(cz * inv, B.g.L[2]);
    uint32_t section = uint32_t(c.section);
    uint32_t oldIndex = uint32_t(c.found), newIndex;
    const uint32_t target = uint32_t(dst * K + kind);
    s.swapCount[section]--;
    s.swapCount[target]++;
    if (src >= dst) {
      while (section != target) {
        newIndex = s.bkStart[section]++;
        const uint32_t tmp = B.lookup[oldIndex];
        B.lookup[oldIndex] = B.lookup[newIndex];
        B.lookup[newIndex] = tmp;
        oldIndex = newIndex;
        --section;
      }
    } else {
      while (section != target) {
        newIndex = --s.bkStart[++section];
        const uint32_t tmp = B.lookup[oldIndex];
        B.lookup[oldIndex] = B.lookup[newIndex];
        B.lookup[newIndex] = tmp;
        oldIndex = newIndex;
      }
    }
    B.molBox[m] = dst;
    if (dst == 0)
      s.kindCount[kind]++;
    else
      s.kindCount[kind]--;
    const int st = c.st;
    for (int p = 0; p < len; ++p) {  // the live-atom list
      const int a = st + p;
      if (dst == 0) {
        B.live[c.nLive] = a;
        B.livePos[a] = c.nLive;
        c.nLive++;
        CellPlace(B, c, a, c.nm.px[p], c.nm.py[p], c.nm.pz[p]);
      } else {
        if (UseCells(B, c)) CellRemove(B, a);
        const int pos = B.livePos[a];
        if (pos < 0 || c.nLive <= 0) {  // list out of step with the molecule: stop
          s.status = kStErr;
          c.abort = 1;
          break;
        }
        const int last = B.live[c.nLive - 1];
        B.live[pos] = last;
        B.livePos[last] = pos;
        B.livePos[a] = -1;
        c.nLive--;
      }
    }
    const uint32_t left = s.bkStart[K] - s.bkStart[0];
    if (src == 0 && left == 0) {
      for (int e = 0; e < 7; ++e) s.pot[e] = 0.0;
    } else if (src == 0 && left == 1) {
      s.pot[0] = 0;
      s.pot[2] = 0;
    }
    Total(s);
    s.accepted++;
  }
  s.swapTries[dst * K + kind]++;
  if (c.commit) s.swapAcc[dst * K + kind]++;
  const double e0 = dst == 0 ? c.nm.inter : -c.om.inter;
  const double e1 = dst == 0 ? c.nm.real : -c.om.real;
  Record(B, i, kMolTransfer, m, c.commit, c.nm.overlap, e0, e1, c.recipGain + c.recipLose);
}

// The MoleculeTransfer move after PickArbDist (gcmc.cpp TransferMove).
template <int P>
__device__ void TransferMoveD(DevBox &B, DevState &s, Ctl &c, PhaseTab &tab, int64_t i) {
  if (threadIdx.x == 0) TransferPick(B, s, c, i);
  __syncthreads();
  if (c.op != kOpTransfer) {  // nothing to transfer: the move's exchange alone
    if (!c.abort) TeamExchange(B, s, c, -1);
    return;
  }
  if (c.len == 1) {
    SingleStage<P>(B, s, c, c.nm, 0, true);
    if (c.abort) return;
    SingleStage<P>(B, s, c, c.om, 0, false);
  } else {
    HedronStage<P>(B, s, c, c.nm, true);
    if (c.abort) return;
    HedronStage<P>(B, s, c, c.om, false);
  }
  if (c.abort) return;
  if (threadIdx.x == 0) TransferCalc(B, s, c);
  __syncthreads();
  if (c.doRecip)
    SwapRecipEnergy<P>(B, s, c, tab);
  else
    TeamExchange(B, s, c, -1);  // one exchange per move
  if (c.abort) return;
  if (threadIdx.x == 0) TransferAccept(B, s, c);
  __syncthreads();
  if (c.commit) {
    // the molecule's index in its lookup section, and its new coordinates
    const int m = c.mol, st = c.st;
    const int lo = int(s.bkStart[c.section]), hi = int(s.bkStart[c.section + 1]);
    for (int idx = lo + threadIdx.x; idx < hi; idx += kDevThreads)
      if (int(B.lookup[idx]) == m) c.found = idx;
    if (threadIdx.x < c.len) {
      const int p = threadIdx.x;
      B.x[st + p] = c.nm.px[p];
      B.y[st + p] = c.nm.py[p];
      B.z[st + p] = c.nm.pz[p];
      B.mol[st + p] = c.dst == 0 ? m : -1;
    }
    if ((c.recipLose != 0.0 || c.recipGain != 0.0) && threadIdx.x == 0) SwapRecip(B);
  }
  __syncthreads();
  if (threadIdx.x == 0 && c.commit && c.found < 0) {  // lookup out of step: stop here
    s.status = kStErr;
    c.abort = 1;
  }
  __syncthreads();
  if (c.abort) return;
  if (threadIdx.x == 0) TransferFinish(B, s, c, i);
  __syncthreads();
}

// Translate / Rotate: PickMolAndBox, Prep and Transform (driver.cpp
// OneMove; src/Coordinates.cpp:83-128).
NOINLINE __device__ void MolPick(const DevBox &B, DevState &s, Ctl &c, int64_t i) {
  c.op = kOpMol;
  c.skip = -1;
  double subDraw = c.subDraw;
  const int pick = c.pick;
  const double boxDiv = B.movePerc[pick] / 1;
  uint32_t box = uint32_t(subDraw / boxDiv);
  if (box >= 1) box--;
  subDraw -= box * boxDiv;
  const uint32_t mkTot = uint32_t(B.nMoveKinds);
  uint32_t k = uint32_t(subDraw / (boxDiv / mkTot));
  if (k == mkTot) k = mkTot - 1;
  const int mk = B.moveKind[k];
  int mv = pick;
  c.mk = mk;
  c.mv = mv;
  if (NumKindInBox(B, s, mk, 0) == 0) {
    s.tries[mv][mk]++;
    s.tTries[mv][mk]++;
    Record(B, i, mv, -1, 0, 0, 0.0, 0.0, 0.0);
    c.op = kOpNone;
    return;
  }
  const uint32_t off = MtBelow(s, NumKindInBox(B, s, mk, 0) - 1);
  const int m = int(B.lookup[s.bkStart[0 * B.nkinds + mk] + off]);
  if (mv == kRotate && B.kindLen[mk] <= 1) mv = kDisplace;  // Rotate::Prep
  c.mv = mv;
  c.mol = m;
  c.st = B.molStart[m];
  c.len = B.molStart[m + 1] - c.st;
  const double *L = B.g.L, *H = B.g.half;
  c.ncom[0] = B.com[m];
  c.ncom[1] = B.com[B.nmols + m];
  c.ncom[2] = B.com[2 * B.nmols + m];
  for (int p = 0; p < kDevKindAtoms; ++p) {
    const bool in = p < c.len;
    c.ox[p] = in ? B.x[c.st + p] : 0.0;
    c.oy[p] = in ? B.y[c.st + p] : 0.0;
    c.oz[p] = in ? B.z[c.st + p] : 0.0;
    c.mq[p] = in ? B.q[c.st + p] : 0.0;
    c.mt_[p] = in ? B.type[c.st + p] : 0;
    c.nx[p] = c.ny[p] = c.nz[p] = 0.0;
  }
  if (mv == kDisplace) {
    const double bb = s.scale[kDisplace][mk], b2 = 2.0 * bb;
    const double sx = MtClosed(s) * b2 - bb;
    const double sy = MtClosed(s) * b2 - bb;
    const double sz = MtClosed(s) * b2 - bb;
    for (int p = 0; p < c.len; ++p) {
      c.nx[p] = Wrap(c.ox[p] + sx, L[0]);
      c.ny[p] = Wrap(c.oy[p] + sy, L[1]);
      c.nz[p] = Wrap(c.oz[p] + sz, L[2]);
    }
    c.ncom[0] = Wrap(c.ncom[0] + sx, L[0]);
    c.ncom[1] = Wrap(c.ncom[1] + sy, L[1]);
    c.ncom[2] = Wrap(c.ncom[2] + sz, L[2]);
  } else {
    const double bb = s.scale[kRotate][mk];
    const double theta = MtClosed(s) * (2.0 * bb) - bb;
    const double u = MtClosed(s) * (2.0 * 1.0) - 1.0;
    const double phi = MtHalfOpen(s) * (2.0 * M_PI);
    const double rr = sqrt(1.0 - u * u);
    const DM3 R = AxisAngleD(theta, rr * cos(phi), rr * sin(phi), u);
    const double cx = c.ncom[0], cy = c.ncom[1], cz = c.ncom[2];
    for (int p = 0; p < c.len; ++p) {
      const double vx = Unwrap(c.ox[p], cx, L[0], H[0]) + (-cx);
      const double vy = Unwrap(c.oy[p], cy, L[1], H[1]) + (-cy);
      const double vz = Unwrap(c.oz[p], cz, L[2], H[2]) + (-cz);
      double ax, ay, az;
      Apply(R, vx, vy, vz, ax, ay, az);
      c.nx[p] = Wrap(ax + cx, L[0]);
      c.ny[p] = Wrap(ay + cy, L[1]);
      c.nz[p] = Wrap(az + cz, L[2]);
    }
  }
  c.pr = MtClosed(s);  // Translate/Rotate::Accept draw
  c.doRecip = B.g.ewald;
}

// Translate / Rotate::Accept and the move's record.
NOINLINE __device__ void MolAccept(const DevBox &B, DevState &s, Ctl &c, int64_t i) {
  const int mv = c.mv, mk = c.mk;
  const bool overlap = c.res[3] > 0.0;
  const double dlj = c.res[0], dreal = c.res[1];
  const double drecip = (!overlap && B.g.ewald) ? c.res[2] - s.pot[3] : 0.0;
  const bool res = c.pr < exp(-B.beta * (dlj + dreal + drecip));
  const bool result = res && !overlap;
  c.commit = result;
  if (result) {
    s.pot[0] += dlj;
    s.pot[2] += dreal;
    s.pot[3] += drecip;
    if (mv == kDisplace) {
      B.com[c.mol] = c.ncom[0];
      B.com[B.nmols + c.mol] = c.ncom[1];
      B.com[2 * B.nmols + c.mol] = c.ncom[2];
    }
    Total(s);
    s.accepted++;
  }
  s.tries[mv][mk]++;
  s.tTries[mv][mk]++;
  if (result) {
    s.tAcc[mv][mk]++;
    s.acc[mv][mk]++;
  }
  Record(B, i, mv, c.mol, result, overlap, dlj, dreal, drecip);
}

template <int P>
__device__ void MolMoveD(DevBox &B, DevState &s, Ctl &c, PhaseTab &tab, int64_t i) {
  if (threadIdx.x == 0) MolPick(B, s, c, i);
  __syncthreads();
  if (c.op != kOpMol) {  // no molecule of the kind: the move's exchange alone
    TeamExchange(B, s, c, -1);
    return;
  }
  MolTrialEnergies<P>(B, s, c, tab);
  if (c.abort) return;
  if (threadIdx.x == 0) MolAccept(B, s, c, i);
  __syncthreads();
  if (c.commit) {
    if (threadIdx.x < c.len) {
      const int p = threadIdx.x;
      B.x[c.st + p] = c.nx[p];
      B.y[c.st + p] = c.ny[p];
      B.z[c.st + p] = c.nz[p];
    }
    if (threadIdx.x == 0) {
      if (B.g.ewald) SwapRecip(B);
      for (int p = 0; p < c.len; ++p) CellPlace(B, c, c.st + p, c.nx[p], c.ny[p], c.nz[p]);
    }
  }
  __syncthreads();
}

// One workgroup per box: the box's moves from step0, either nsteps moves
// or, with cycles > 0, up to `cycles` cycles of N_box moves each (N_box
// counted at the cycle's start, at least 1; never more than nsteps in all),
// sampling the molecule count of kinds 0 and 1 after every cycle.
template <int P>
__global__ void __launch_bounds__(kDevThreads) DeviceMovesKernel(const DevBox *boxes) {
  __shared__ DevBox B;
  __shared__ DevState S;
  __shared__ Ctl c;
  __shared__ PhaseTab tab;
  {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&boxes[blockIdx.x]);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&B);
    for (int w = threadIdx.x; w < int(sizeof(DevBox) / 4); w += kDevThreads) dst[w] = src[w];
  }
  __syncthreads();
  {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(B.st);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&S);
    for (int w = threadIdx.x; w < int(sizeof(DevState) / 4); w += kDevThreads) dst[w] = src[w];
  }
  if (threadIdx.x <= kErfcxDeg) sErfcx[threadIdx.x] = kErfcxPoly[threadIdx.x];
  if (threadIdx.x == 0) {
    c.nLive = B.nLive0;
    c.cellsOff = 0;
    c.abort = 0;
    c.done = 0;
    c.cyclesDone = 0;
    c.cycleLeft = -1;
    c.xseq = 1;
  }
  __syncthreads();
  while (true) {
    // keep the next block of the MT19937 stream ready (a move draws < 400 words)
    if (!S.nextReady && kMtN - S.mtIdx < 400) MtPrepareNext(S);
    if (threadIdx.x == 0) LoopTop(B, S, c);
    __syncthreads();
    const int op = c.op;
    if (op == kOpStop) break;
    const int64_t i = c.done;
    if (op == kOpTransfer)
      TransferMoveD<P>(B, S, c, tab, i);
    else
      MolMoveD<P>(B, S, c, tab, i);
    if (threadIdx.x == 0 && !c.abort) {
      c.done++;
      S.moves++;
      if (B.cycles > 0) c.cycleLeft--;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && B.cycles > 0 && c.cycleLeft == 0 && !c.abort && S.status == 0) {
    S.sumN[0] += B.nkinds > 0 ? double(NumKindInBox(B, S, 0, 0)) : 0.0;
    S.sumN[1] += B.nkinds > 1 ? double(NumKindInBox(B, S, 1, 0)) : 0.0;
    S.samples++;
  }
  if (threadIdx.x == 0) S.recipSwapped = B.sumR != boxes[blockIdx.x].sumR;
  __syncthreads();
  if (B.rank == 0) {  // the members of a team hold the same state
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&S);
    uint32_t *dst = reinterpret_cast<uint32_t *>(B.st);
    for (int w = threadIdx.x; w < int(sizeof(DevState) / 4); w += kDevThreads) dst[w] = src[w];
  }
}

static_assert(sizeof(DevBox) % 4 == 0 && sizeof(DevState) % 4 == 0, "word copies");

}  // namespace

// Device buffers of a box's device move loop, allocated on first use.
struct DevDriver {
  DevBox hbox;                  // this box's description for the next launch
  DevState *state = nullptr;    // its state (device)
  int *molKind = nullptr, *molBox = nullptr, *live = nullptr, *livePos = nullptr;
  uint32_t *lookup = nullptr;
  double *com = nullptr;
  int32_t *kpack = nullptr;     // k-vector index triples
  int nmax[3] = {0, 0, 0};
  int32_t *trace = nullptr;     // [4][cap] ints then [3][cap] doubles
  int64_t traceCap = 0;
  DevBox *list = nullptr;       // launch list when this box leads a launch
  int listCap = 0;
  DevState hstate;
  int team = 0;                 // members; live, livePos and lookup hold one copy each
  uint64_t *xchg = nullptr;     // the team's exchange block
  // the cell table, one copy per member: counts [team][ncell], slots
  // [team][ncell][ccap], atom cell and slot [team][natoms]
  int *cCount = nullptr, *cSlots = nullptr, *aCell = nullptr, *aSlot = nullptr;
  size_t cCountCap = 0, cSlotsCap = 0, aCellCap = 0, aSlotCap = 0;
  int ccap = 0;
};

void DevFree(gomc_box *b) {
  DevDriver *d = b->dev;
  if (!d) return;
  for (void *p : {(void *)d->state, (void *)d->kpack, (void *)d->molKind, (void *)d->molBox,
                  (void *)d->live, (void *)d->livePos,
                  (void *)d->lookup, (void *)d->com, (void *)d->trace, (void *)d->list,
                  (void *)d->xchg, (void *)d->cCount, (void *)d->cSlots, (void *)d->aCell,
                  (void *)d->aSlot})
    if (p) (void)hipFree(p);
  delete d;
  b->dev = nullptr;
}

namespace {

// Members of the box's team: set by gomc_set_device_team, else by the
// k-vector count (the k-vectors are what the members share out).
int DevTeamSize(const gomc_box *b) {
  if (b->devTeam > 0) return std::min(b->devTeam, kDevTeamMax);
  if (!b->ewald) return 1;
  return b->K >= 16384 ? 4 : (b->K >= 8192 ? 2 : 1);
}

int DevCheck(gomc_box *b) {
  if (b->cached) return Fail(b, GOMC_ERR_STATE, "device move loop: EwaldCached mode is not supported");
  if (b->moveW[2] > 0.0)
    return Fail(b, GOMC_ERR_STATE, "device move loop: MultiParticle moves run on the host driver");
  if (b->ensemble == 2)
    return Fail(b, GOMC_ERR_STATE, "device move loop: NPT volume moves run on the host driver");
  if (b->nkinds > kDevKinds) return Fail(b, GOMC_ERR_ARG, "device move loop: more than 4 molecule kinds");
  for (int k = 0; k < b->nkinds; ++k) {
    const int len = int(b->kindType[k].size());
    if (len > kDevKindAtoms) return Fail(b, GOMC_ERR_ARG, "device move loop: molecules of more than 3 atoms");
    if (b->ensemble == 1 && len != 1 && len != 3)
      return Fail(b, GOMC_ERR_ARG, "device move loop: transferred kinds must have 1 or 3 atoms");
  }
  if (b->cbmcFirst > kDevTrials || b->cbmcNth > kDevTrials)
    return Fail(b, GOMC_ERR_ARG, "device move loop: more than 32 CBMC trials per stage");
  return GOMC_OK;
}

int DevEnsure(gomc_box *b, int64_t traceMoves) {
  if (!b->dev) {
    b->dev = new DevDriver();
    DevDriver *d = b->dev;
    HIP_TRY(b, hipMalloc(&d->state, sizeof(DevState)));
    HIP_TRY(b, hipMalloc(&d->molKind, sizeof(int) * std::max(1, b->nmols)));
    HIP_TRY(b, hipMalloc(&d->molBox, sizeof(int) * std::max(1, b->nmols)));
    HIP_TRY(b, hipMalloc(&d->com, sizeof(double) * 3 * std::max(1, b->nmols)));
    HIP_TRY(b, hipMalloc(&d->xchg, sizeof(uint64_t) * kDevXchgWords));
    HIP_TRY(b, hipMemcpy(d->molKind, b->molKind.data(), sizeof(int) * b->nmols, hipMemcpyHostToDevice));
    // index triples of the box's k-vectors, in their order
    std::vector<double> kx, ky, kz, pref;
    std::vector<int32_t> packed;
    KVectorsOrth(b->geom.L, b->alpha * b->alpha, b->recipRcut, b->volume, kx, ky, kz, pref, &packed);
    if (b->ewald && int(packed.size()) != b->K)
      return Fail(b, GOMC_ERR_STATE, "device move loop: k-vector count mismatch");
    for (int32_t code : packed) {
      d->nmax[0] = std::max(d->nmax[0], code & 0xff);
      d->nmax[1] = std::max(d->nmax[1], std::abs(((code >> 8) & 0xff) - 128));
      d->nmax[2] = std::max(d->nmax[2], std::abs(((code >> 16) & 0xff) - 128));
    }
    if (d->nmax[0] > kDevNMax || d->nmax[1] > kDevNMax || d->nmax[2] > kDevNMax)
      return Fail(b, GOMC_ERR_ARG, "device move loop: more than 40 k-vectors per axis");
    HIP_TRY(b, hipMalloc(&d->kpack, sizeof(int32_t) * std::max<size_t>(1, packed.size())));
    if (!packed.empty())
      HIP_TRY(b, hipMemcpy(d->kpack, packed.data(), sizeof(int32_t) * packed.size(),
                           hipMemcpyHostToDevice));
  }
  DevDriver *d = b->dev;
  const int team = DevTeamSize(b);
  if (team != d->team) {  // one copy of the lookup and the live-atom lists per member
    for (void *p : {(void *)d->lookup, (void *)d->live, (void *)d->livePos})
      if (p) (void)hipFree(p);
    d->lookup = nullptr;
    d->live = d->livePos = nullptr;
    d->team = 0;
    HIP_TRY(b, hipMalloc(&d->lookup, sizeof(uint32_t) * team * std::max(1, b->nmols)));
    HIP_TRY(b, hipMalloc(&d->live, sizeof(int) * team * std::max(1, b->natoms)));
    HIP_TRY(b, hipMalloc(&d->livePos, sizeof(int) * team * std::max(1, b->natoms)));
    d->team = team;
  }
  if (traceMoves > d->traceCap) {
    if (d->trace) (void)hipFree(d->trace);
    d->trace = nullptr;
    d->traceCap = 0;
    HIP_TRY(b, hipMalloc(&d->trace, size_t(traceMoves) * (4 * sizeof(int32_t) + 3 * sizeof(double))));
    d->traceCap = traceMoves;
  }
  return GOMC_OK;
}

// Host state -> device: the box description and the move loop's state.
int DevUpload(gomc_box *b, int64_t step0, int64_t nsteps, int cycles, bool trace) {
  DevDriver *d = b->dev;
  DevBox &h = d->hbox;
  std::memset(&h, 0, sizeof(h));
  h.g = b->geom;
  h.beta = b->beta;
  h.volume = b->volume;
  h.volInv = b->volInv;
  const int K = b->nkinds;
  for (int i = 0; i < K * K; ++i) h.pairEnCorr[i] = b->pairEnCorr.empty() ? 0.0 : b->pairEnCorr[i];
  for (int k = 0; k < K; ++k) {
    h.fugacity[k] = b->fugacity.empty() ? 0.0 : b->fugacity[k];
    double en = 0.0;  // Ewald::SwapSelf (src/Ewald.cpp:1375-1391)
    for (double q : b->kindCharge[k]) en -= (q * q);
    h.selfKind[k] = (en * b->alpha * kQQFact * M_2_SQRTPI * 0.5);
    h.kindLen[k] = int(b->kindType[k].size());
    for (int a = 0; a < h.kindLen[k]; ++a) {
      h.kindT[k][a] = b->kindType[k][a];
      h.kindQ[k][a] = b->kindCharge[k][a];
    }
    if (!b->hedron.empty()) {
      h.hedB[k] = b->hedron[k].bBonded;
      h.hedP[k] = b->hedron[k].bPrev;
      h.hedTh[k] = b->hedron[k].theta;
    }
  }
  for (int i = 0; i < 13; ++i) h.moveW[i] = b->moveW[i];
  h.movePerc[0] = b->movePerc[0];
  h.movePerc[1] = b->movePerc[1];
  h.nMoveTypes = b->ensemble == 1 ? 13 : 9;
  h.lrc = b->lrc;
  h.nkinds = K;
  h.natoms = b->natoms;
  h.nmols = b->nmols;
  h.ntypes = b->ntypes;
  h.K = b->ewald ? b->K : 0;
  h.cbmcFirst = b->cbmcFirst;
  h.cbmcNth = b->cbmcNth;
  h.cbmcAng = b->cbmcAng;
  h.nSwapKinds = int(b->canSwapKind.size());
  h.nMoveKinds = int(b->canMoveKind.size());
  for (int i = 0; i < h.nSwapKinds; ++i) h.swapKind[i] = b->canSwapKind[i];
  for (int i = 0; i < h.nMoveKinds; ++i) h.moveKind[i] = b->canMoveKind[i];
  h.perAdjust = b->perAdjust;
  h.x = b->x;
  h.y = b->y;
  h.z = b->z;
  h.q = b->q;
  h.type = b->type;
  h.mol = b->mol;
  h.sigSq = b->tSigSq;
  h.epsCn = b->tEpsCn;
  h.nHalf = b->tNHalf;
  h.kx = b->kx;
  h.ky = b->ky;
  h.kz = b->kz;
  h.pref = b->pref;
  h.sumR = b->sumRref;
  h.sumI = b->sumIref;
  h.newR = b->sumRnew;
  h.newI = b->sumInew;
  h.kpack = d->kpack;
  for (int a = 0; a < 3; ++a) h.nmax[a] = d->nmax[a];
  h.molStart = b->molStartDev;
  h.molKind = d->molKind;
  h.molBox = d->molBox;
  h.live = d->live;  // rank 0's copies; Launch points each member at its own
  h.livePos = d->livePos;
  h.lookup = d->lookup;
  h.team = d->team;
  h.rank = 0;
  h.k0 = 0;
  h.k1 = b->ewald ? b->K : 0;
  h.xchg = d->xchg;
  h.com = d->com;
  h.st = d->state;
  h.step0 = step0;
  h.nsteps = nsteps;
  h.cycles = cycles;
  if (trace) {
    h.tKind = d->trace;
    h.tMol = d->trace + d->traceCap;
    h.tAcc = d->trace + 2 * d->traceCap;
    h.tOvl = d->trace + 3 * d->traceCap;
    h.tE = reinterpret_cast<double *>(d->trace + 4 * d->traceCap);
  }
  DevState &s = d->hstate;
  std::memset(&s, 0, sizeof(s));
  b->mt.GetState(s.mt[0], &s.mtIdx);
  s.mtCur = 0;
  s.nextReady = 0;
  s.pot[0] = b->pot.inter;
  s.pot[1] = b->pot.tc;
  s.pot[2] = b->pot.real;
  s.pot[3] = b->pot.recip;
  s.pot[4] = b->pot.self;
  s.pot[5] = b->pot.corr;
  s.pot[6] = b->pot.total;
  for (int mv = 0; mv < 2; ++mv)
    for (int k = 0; k < K; ++k) {
      s.scale[mv][k] = b->moves.scale[mv][k];
      s.tries[mv][k] = b->moves.tries[mv][k];
      s.acc[mv][k] = b->moves.accepted[mv][k];
      s.tTries[mv][k] = b->moves.tempTries[mv][k];
      s.tAcc[mv][k] = b->moves.tempAccepted[mv][k];
    }
  for (int i = 0; i < 2 * K; ++i) {
    s.swapTries[i] = b->swapTries.empty() ? 0 : b->swapTries[i];
    s.swapAcc[i] = b->swapAcc.empty() ? 0 : b->swapAcc[i];
    s.swapCount[i] = b->swapCount[i];
  }
  for (int i = 0; i <= 2 * K; ++i) s.bkStart[i] = b->bkStart[i];
  for (int k = 0; k < K; ++k) s.kindCount[k] = b->kindCount[k];
  std::vector<double> com(3 * size_t(b->nmols));
  for (int m = 0; m < b->nmols; ++m) {
    com[m] = b->comx[m];
    com[b->nmols + m] = b->comy[m];
    com[2 * size_t(b->nmols) + m] = b->comz[m];
  }
  std::vector<int> molBox(b->nmols, 0);
  for (int m = 0; m < b->nmols && !b->molBox.empty(); ++m) molBox[m] = b->molBox[m];
  std::vector<int> live, livePos(b->natoms, -1);
  live.reserve(b->natoms);
  for (int a = 0; a < b->natoms; ++a)
    if (molBox[b->atomMol[a]] == 0) {
      livePos[a] = int(live.size());
      live.push_back(a);
    }
  h.nLive0 = int(live.size());
  // The live atoms by cell, for a grid of more than 27 cells holding at
  // least kDevCellMinAtoms live atoms (below that a wave's scan of the live
  // list is shorter than the walk's set-up; measured on a 60 A GCMC box: 41
  // molecules 28.8 us per move with the scan, 30.7 with the walk; 1081
  // molecules 50.1 and 35.1): capacity twice the fullest cell plus 8, and at
  // least 0.12 atoms per cubic Angstrom (a dense liquid) of the cell's
  // volume; a cell that still runs full turns the cell walk off for the rest
  // of that launch.
  std::vector<int> cCount, cSlots, aCell, aSlot;
  h.cells = 0;
  if (b->ncell > 27 && live.size() >= size_t(kDevCellMinAtoms)) {
    const int nc = b->ncell;
    cCount.assign(nc, 0);
    aCell.assign(b->natoms, -1);
    aSlot.assign(b->natoms, -1);
    for (int a : live) {
      const int cl = CellOf(b->hx[a], b->hy[a], b->hz[a], b->cellSize, b->edge);
      aCell[a] = cl;
      aSlot[a] = cCount[cl]++;
    }
    const int mx = *std::max_element(cCount.begin(), cCount.end());
    const double vol = b->cellSize[0] * b->cellSize[1] * b->cellSize[2];
    const int cap = std::max(2 * mx + 8, int(std::ceil(0.12 * vol)));
    cSlots.assign(size_t(nc) * cap, -1);
    for (int a : live) cSlots[size_t(aCell[a]) * cap + aSlot[a]] = a;
    const size_t tm = size_t(d->team);
    auto grow = [&](int *&p, size_t &have, size_t need) -> hipError_t {
      if (need <= have) return hipSuccess;
      if (p) (void)hipFree(p);
      p = nullptr;
      have = 0;
      const hipError_t e = hipMalloc(&p, sizeof(int) * need);
      if (e == hipSuccess) have = need;
      return e;
    };
    HIP_TRY(b, grow(d->cCount, d->cCountCap, tm * nc));
    HIP_TRY(b, grow(d->cSlots, d->cSlotsCap, tm * cSlots.size()));
    HIP_TRY(b, grow(d->aCell, d->aCellCap, tm * b->natoms));
    HIP_TRY(b, grow(d->aSlot, d->aSlotCap, tm * b->natoms));
    d->ccap = cap;
    for (size_t r = 0; r < tm; ++r) {
      HIP_TRY(b, hipMemcpyAsync(d->cCount + r * nc, cCount.data(), sizeof(int) * nc,
                                hipMemcpyHostToDevice, b->stream));
      HIP_TRY(b, hipMemcpyAsync(d->cSlots + r * cSlots.size(), cSlots.data(),
                                sizeof(int) * cSlots.size(), hipMemcpyHostToDevice, b->stream));
      HIP_TRY(b, hipMemcpyAsync(d->aCell + r * b->natoms, aCell.data(), sizeof(int) * b->natoms,
                                hipMemcpyHostToDevice, b->stream));
      HIP_TRY(b, hipMemcpyAsync(d->aSlot + r * b->natoms, aSlot.data(), sizeof(int) * b->natoms,
                                hipMemcpyHostToDevice, b->stream));
    }
    h.cells = 1;
    h.ncell = nc;
    h.ccap = cap;
    for (int i = 0; i < 3; ++i) {
      h.edge[i] = b->edge[i];
      h.cs[i] = b->cellSize[i];
    }
    h.nbr = b->nbr;
  }
  HIP_TRY(b, hipMemcpyAsync(d->state, &s, sizeof(s), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(d->com, com.data(), sizeof(double) * com.size(), hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(d->molBox, molBox.data(), sizeof(int) * b->nmols, hipMemcpyHostToDevice, b->stream));
  for (int r = 0; r < d->team; ++r)
    HIP_TRY(b, hipMemcpyAsync(d->lookup + size_t(r) * b->nmols, b->lookup.data(),
                              sizeof(uint32_t) * b->nmols, hipMemcpyHostToDevice, b->stream));
  if (!live.empty())
    for (int r = 0; r < d->team; ++r)
      HIP_TRY(b, hipMemcpyAsync(d->live + size_t(r) * b->natoms, live.data(),
                                sizeof(int) * live.size(), hipMemcpyHostToDevice, b->stream));
  for (int r = 0; r < d->team; ++r)
    HIP_TRY(b, hipMemcpyAsync(d->livePos + size_t(r) * b->natoms, livePos.data(),
                              sizeof(int) * b->natoms, hipMemcpyHostToDevice, b->stream));
  // the copies read host vectors that go out of scope: complete them now
  HIP_TRY(b, hipStreamSynchronize(b->stream));
  return GOMC_OK;
}

// Device -> host after a launch: the driver's state, coordinates, COMs,
// molecule boxes and lookup, so the host driver and every query continue
// from where the device loop stopped.
int DevDownload(gomc_box *b) {
  DevDriver *d = b->dev;
  DevState &s = d->hstate;
  std::vector<double> com(3 * size_t(b->nmols));
  std::vector<int> molBox(b->nmols);
  const size_t bytes = sizeof(double) * b->natoms;
  HIP_TRY(b, hipMemcpyAsync(&s, d->state, sizeof(s), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(b, hipMemcpyAsync(com.data(), d->com, sizeof(double) * com.size(), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(b, hipMemcpyAsync(molBox.data(), d->molBox, sizeof(int) * b->nmols, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->lookup.data(), d->lookup, sizeof(uint32_t) * b->nmols, hipMemcpyDeviceToHost, b->stream));
  TrialCellsInvalidate(b);  // the loop moved molecules and changed which ones live
  HIP_TRY(b, hipMemcpyAsync(b->hx.data(), b->x, bytes, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->hy.data(), b->y, bytes, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->hz.data(), b->z, bytes, hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(b, hipStreamSynchronize(b->stream));
  if (s.recipSwapped) {  // the box's structure factors ended in the other array
    std::swap(b->sumRref, b->sumRnew);
    std::swap(b->sumIref, b->sumInew);
  }
  const int K = b->nkinds;
  b->mt.SetState(s.mt[s.mtCur], s.mtIdx);
  b->pot.inter = s.pot[0];
  b->pot.tc = s.pot[1];
  b->pot.real = s.pot[2];
  b->pot.recip = s.pot[3];
  b->pot.self = s.pot[4];
  b->pot.corr = s.pot[5];
  b->pot.total = s.pot[6];
  for (int mv = 0; mv < 2; ++mv)
    for (int k = 0; k < K; ++k) {
      b->moves.scale[mv][k] = s.scale[mv][k];
      b->moves.tries[mv][k] = s.tries[mv][k];
      b->moves.accepted[mv][k] = s.acc[mv][k];
      b->moves.tempTries[mv][k] = s.tTries[mv][k];
      b->moves.tempAccepted[mv][k] = s.tAcc[mv][k];
    }
  for (int i = 0; i < 2 * K; ++i) {
    if (!b->swapTries.empty()) b->swapTries[i] = s.swapTries[i];
    if (!b->swapAcc.empty()) b->swapAcc[i] = s.swapAcc[i];
    b->swapCount[i] = s.swapCount[i];
  }
  for (int i = 0; i <= 2 * K; ++i) b->bkStart[i] = s.bkStart[i];
  for (int k = 0; k < K; ++k) b->kindCount[k] = s.kindCount[k];
  for (int m = 0; m < b->nmols; ++m) {
    b->comx[m] = com[m];
    b->comy[m] = com[b->nmols + m];
    b->comz[m] = com[2 * size_t(b->nmols) + m];
    if (!b->molBox.empty()) b->molBox[m] = molBox[m];
  }
  b->mp.singleAccepted = true;  // forces are recomputed by a later MultiParticle move
  switch (s.status) {
    case 0:
      return GOMC_OK;
    case kStEmptyReservoir:
      return Fail(b, GOMC_ERR_STATE, "no molecules of this kind left in the reservoir");
    case kStWeights:
      return Fail(b, GOMC_ERR_STATE, "CBMC weights inconsistent");
    case kStMoveKind:
      return Fail(b, GOMC_ERR_STATE, "move kind not served by this driver");
    case kStTeam:
      return Fail(b, GOMC_ERR_STATE, "device move loop: a team member did not answer");
    default:
      return Fail(b, GOMC_ERR_STATE, "device move loop failed");
  }
}

// Workgroups of the device move loop that the device holds at once.
int ResidentWorkgroups(gomc_box *b, int *out) {
  int perCU = 0, cus = 0;
  hipError_t e = b->precision == 1
                     ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, DeviceMovesKernel<1>,
                                                                    kDevThreads, 0)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, DeviceMovesKernel<0>,
                                                                    kDevThreads, 0);
  if (e != hipSuccess) return Fail(b, GOMC_ERR_HIP, std::string("occupancy: ") + hipGetErrorString(e));
  HIP_TRY(b, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
  *out = perCU * cus;
  return GOMC_OK;
}

// The boxes' moves, each box on its team of workgroups.  A launch holds
// whole teams and no more workgroups than the device holds at once, so a
// member never waits for one that cannot be scheduled; boxes beyond that go
// to the next launch (a box's result does not depend on its launch).
int Launch(gomc_box *lead, const std::vector<gomc_box *> &boxes) {
  DevDriver *d = lead->dev;
  int cap = 0;
  int rc = ResidentWorkgroups(lead, &cap);
  if (rc) return rc;
  if (cap > d->listCap) {
    if (d->list) (void)hipFree(d->list);
    d->list = nullptr;
    d->listCap = 0;
    HIP_TRY(lead, hipMalloc(&d->list, sizeof(DevBox) * cap));
    d->listCap = cap;
  }
  std::vector<DevBox> v;
  size_t i = 0;
  while (i < boxes.size()) {
    v.clear();
    for (; i < boxes.size(); ++i) {
      gomc_box *b = boxes[i];
      DevDriver *bd = b->dev;
      const int team = bd->team;
      if (team > cap)
        return Fail(lead, GOMC_ERR_STATE, "device move loop: a team larger than the device holds");
      if (int(v.size()) + team > cap) break;
      if (team > 1)  // exchange numbers start again at 1
        HIP_TRY(lead, hipMemsetAsync(bd->xchg, 0, sizeof(uint64_t) * kDevXchgWords, lead->stream));
      const int K = bd->hbox.k1;
      for (int r = 0; r < team; ++r) {
        DevBox m = bd->hbox;
        m.rank = r;
        m.k0 = int(int64_t(K) * r / team);
        m.k1 = int(int64_t(K) * (r + 1) / team);
        m.live = bd->live + size_t(r) * b->natoms;
        m.livePos = bd->livePos + size_t(r) * b->natoms;
        m.lookup = bd->lookup + size_t(r) * b->nmols;
        if (m.cells) {
          m.cCount = bd->cCount + size_t(r) * m.ncell;
          m.cSlots = bd->cSlots + size_t(r) * m.ncell * m.ccap;
          m.aCell = bd->aCell + size_t(r) * b->natoms;
          m.aSlot = bd->aSlot + size_t(r) * b->natoms;
        }
        if (r > 0) {  // the leader records the moves
          m.tKind = m.tMol = m.tAcc = m.tOvl = nullptr;
          m.tE = nullptr;
        }
        v.push_back(m);
      }
    }
    const int n = int(v.size());
    HIP_TRY(lead, hipMemcpyAsync(d->list, v.data(), sizeof(DevBox) * n, hipMemcpyHostToDevice,
                                 lead->stream));
    if (lead->precision == 1)
      hipLaunchKernelGGL(DeviceMovesKernel<1>, dim3(n), dim3(kDevThreads), 0, lead->stream, d->list);
    else
      hipLaunchKernelGGL(DeviceMovesKernel<0>, dim3(n), dim3(kDevThreads), 0, lead->stream, d->list);
    HIP_TRY(lead, hipGetLastError());
    HIP_TRY(lead, hipStreamSynchronize(lead->stream));
  }
  return GOMC_OK;
}

}  // namespace
}  // namespace gomc

using namespace gomc;

extern "C" {

int64_t gomc_run_moves_device(gomc_box *b, int64_t start_step, int64_t nsteps, int32_t *kind,
                              int32_t *mol, int32_t *acc, int32_t *ovl, double *dlj,
                              double *dreal, double *drecip) {
  if (!b || nsteps < 0) return GOMC_ERR_ARG;
  if (nsteps == 0) return 0;
  (void)hipSetDevice(b->device);
  int rc = DevCheck(b);
  if (!rc) rc = FlushPending(b);
  if (!rc) rc = DevEnsure(b, nsteps);
  if (!rc) rc = DevUpload(b, start_step, nsteps, 0, true);
  if (!rc) rc = Launch(b, {b});
  if (!rc) rc = DevDownload(b);
  if (rc) return rc;
  DevDriver *d = b->dev;
  const int64_t n = nsteps, cap = d->traceCap;
  std::vector<int32_t> ti(4 * size_t(n));
  std::vector<double> te(3 * size_t(n));
  for (int r = 0; r < 4; ++r)
    HIP_TRY(b, hipMemcpyAsync(ti.data() + r * n, d->trace + r * cap, sizeof(int32_t) * n,
                              hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(b, hipMemcpyAsync(te.data(), d->trace + 4 * cap, sizeof(double) * 3 * n,
                            hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(b, hipStreamSynchronize(b->stream));
  int64_t accepted = 0;
  for (int64_t i = 0; i < n; ++i) {
    accepted += ti[2 * n + i];
    if (kind) kind[i] = ti[i];
    if (mol) mol[i] = ti[n + i];
    if (acc) acc[i] = ti[2 * n + i];
    if (ovl) ovl[i] = ti[3 * n + i];
    if (dlj) dlj[i] = te[3 * i];
    if (dreal) dreal[i] = te[3 * i + 1];
    if (drecip) drecip[i] = te[3 * i + 2];
  }
  return accepted;
}

int gomc_set_device_team(gomc_box *b, int team) {
  if (!b || team < 0 || team > kDevTeamMax) return GOMC_ERR_ARG;
  b->devTeam = team;
  return GOMC_OK;
}

int gomc_run_moves_device_multi(gomc_box *const *boxes, int n, const int64_t *start_step,
                                const int64_t *nsteps, const int32_t *cycles,
                                int64_t *moves_done, int64_t *accepted, double *count_sum,
                                int64_t *count_samples, int32_t *status) {
  if (!boxes || n <= 0 || !start_step || !nsteps) return GOMC_ERR_ARG;
  gomc_box *lead = boxes[0];
  if (!lead) return GOMC_ERR_ARG;
  (void)hipSetDevice(lead->device);
  std::vector<gomc_box *> list(boxes, boxes + n);
  if (status)
    for (int i = 0; i < n; ++i) status[i] = GOMC_OK;
  for (int i = 0; i < n; ++i) {
    gomc_box *b = list[i];
    if (!b || b->device != lead->device || nsteps[i] < 0) return GOMC_ERR_ARG;
    if (b->precision != lead->precision)
      return Fail(lead, GOMC_ERR_ARG, "boxes of one device launch must share a precision");
    int rc = DevCheck(b);
    if (!rc) rc = FlushPending(b);
    if (!rc) rc = DevEnsure(b, 0);
    if (!rc) rc = DevUpload(b, start_step[i], nsteps[i], cycles ? cycles[i] : 0, false);
    if (rc) {
      if (status) status[i] = rc;
      return rc;
    }
  }
  int rc = Launch(lead, list);
  if (rc) {
    if (status)
      for (int i = 0; i < n; ++i) status[i] = rc;
    return rc;
  }
  int first = GOMC_OK;
  for (int i = 0; i < n; ++i) {
    gomc_box *b = list[i];
    rc = DevDownload(b);
    if (status) status[i] = rc;
    if (rc && first == GOMC_OK) first = rc;
    const DevState &s = b->dev->hstate;  // counters start at zero in each launch
    if (moves_done) moves_done[i] = s.moves;
    if (accepted) accepted[i] = s.accepted;
    if (count_sum) {
      count_sum[2 * i] = s.sumN[0];
      count_sum[2 * i + 1] = s.sumN[1];
    }
    if (count_samples) count_samples[i] = s.samples;
  }
  return first;
}

}  // extern "C"

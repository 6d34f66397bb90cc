// Plain single-box path on gfx950: box state resident in HBM, the device
// cell list, the full-box pair sum, the fused single-molecule trial
// (MoleculeInter + MolReciprocal in one launch), the structure factors,
// swap deltas and trial-position energies, and their C entry points.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "box.hpp"

using namespace gomc;

namespace gomc {

int Fail(gomc_box *b, int code, const std::string &msg) {
  if (b) b->err = msg;
  return code;
}

int HipFail(gomc_box *b, hipError_t e, const char *what) {
  return Fail(b, GOMC_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace gomc

namespace {

// ---------------------------------------------------------------- trial --
// Neighbour cells a trial walks when the box has more than 27 cells, and
// edits of the trial cell table one launch carries.
constexpr int kTrialCells = 96;
constexpr int kTrialEdits = 64;
constexpr uint32_t kNewBits = 16;  // mask bit kNewBits + p: new position p

struct TrialArgs {
  BoxGeom g;
  const double *x, *y, *z, *q;
  const int *type, *mol;
  const double *sigSq, *epsCn, *nHalf;
  int ntypes, natoms, mstart, mlen;
  int nPairBlocks, doPair, doRecip, K;
  double ox[kMaxMolAtoms], oy[kMaxMolAtoms], oz[kMaxMolAtoms];
  double nx[kMaxMolAtoms], ny[kMaxMolAtoms], nz[kMaxMolAtoms];
  double mq[kMaxMolAtoms];
  int mtype[kMaxMolAtoms];
  const double *kx, *ky, *kz, *pref, *sumRref, *sumIref;
  double *sumRnew, *sumInew;
  // EwaldCached: the moving molecule's cached old terms, and where its new
  // terms are staged (both null without the cache)
  const double *oldC, *oldS;
  double *stC, *stS;
  // accepted coordinates of molecule [cstart, cstart+clen) not yet in HBM
  int cstart, clen;
  double cx[kMaxMolAtoms], cy[kMaxMolAtoms], cz[kMaxMolAtoms];
  double *wx, *wy, *wz;
  // cell walk (nlist > 0): the cells within one cell of any old or new
  // position, each with the mask of the positions it neighbours
  int nlist, tcCap;
  int *tcCount, *tcSlots;
  int lcell[kTrialCells];
  uint32_t lmask[kTrialCells];
  // trial cell table edits, applied by the last block
  int nedits, tcSlotsN;
  int edits[2 * kTrialEdits];
  Partial4 *partials;
  unsigned *counter;
  double *result;
};
static_assert(sizeof(TrialArgs) <= 4096, "trial arguments must fit the kernel argument buffer");

// Atom j at (xj, yj, zj) against the molecule's old positions p with mask
// bit p set (their energies subtracted) and new positions with bit
// kNewBits + p set (added), p ascending, old before new
// (CalculateEnergy.cpp:593-679).
template <int P>
__device__ __forceinline__ void TrialAtom(const TrialArgs &a, int j, double xj, double yj,
                                          double zj, uint32_t mask, double &lj, double &real,
                                          double &ov) {
  const double qj = a.q[j];
  const int tj = a.type[j];
  for (int p = 0; p < a.mlen; ++p) {
    const int idx = a.mtype[p] + tj * a.ntypes;
    const double qiqj = a.mq[p] * qj * kQQFact;
    if (mask >> p & 1u) {
      const double d2 = DistSq(a.g, a.ox[p], a.oy[p], a.oz[p], xj, yj, zj);
      if (a.g.rCutSq > d2) {
        if (a.g.elect && qiqj != 0.0 && !(a.g.rCutCoulSq < d2)) real += -Coul<P>(a.g, d2, qiqj);
        if (!(a.g.rCutLJSq < d2)) lj += -LJ<P>(d2, a.sigSq[idx], a.epsCn[idx], a.nHalf[idx]);
      }
    }
    if (mask >> (kNewBits + p) & 1u) {
      const double d2 = DistSq(a.g, a.nx[p], a.ny[p], a.nz[p], xj, yj, zj);
      if (a.g.rCutSq > d2) {
        if (d2 < a.g.rCutLowSq) ov = 1.0;
        if (a.g.elect && qiqj != 0.0 && !(a.g.rCutCoulSq < d2)) real += Coul<P>(a.g, d2, qiqj);
        if (!(a.g.rCutLJSq < d2)) lj += LJ<P>(d2, a.sigSq[idx], a.epsCn[idx], a.nHalf[idx]);
      }
    }
  }
}

// Blocks [0, nPairBlocks): the pair part.  Without a cell list one atom j
// per thread against every old and new position; with one, one slot of a
// listed cell per thread against the positions that cell neighbours
// (CellList::EnumerateLocal, as MoleculeInter walks it), then one thread per
// atom of the molecule accepted last (its cells are edited at the end of
// this launch, so the walk skips it and takes it from the arguments).
// Blocks after that: one k-vector per thread, the structure-factor delta and
// the new per-k energy (Ewald.cpp:436-467).
// P = 1: pair terms and phases in FP32, sums in FP64 (the mixed mode).
template <int P>
__global__ void __launch_bounds__(kBlock) TrialKernel(const TrialArgs a) {
  __shared__ double red[kBlock / kWave];
  Partial4 part{0, 0, 0, 0};
  const int tid = threadIdx.x;
  if (blockIdx.x == 0 && tid < a.clen) {
    a.wx[a.cstart + tid] = a.cx[tid];
    a.wy[a.cstart + tid] = a.cy[tid];
    a.wz[a.cstart + tid] = a.cz[tid];
  }
  if (int(blockIdx.x) < a.nPairBlocks) {
    double lj = 0.0, real = 0.0, ov = 0.0;
    const int t = blockIdx.x * kBlock + tid;
    const uint32_t all = ~0u;
    if (a.doPair && a.nlist == 0) {
      const int j = t;
      if (j < a.natoms && (j < a.mstart || j >= a.mstart + a.mlen) && a.mol[j] >= 0) {
        const int cj = j - a.cstart;
        if (cj >= 0 && cj < a.clen)
          TrialAtom<P>(a, j, a.cx[cj], a.cy[cj], a.cz[cj], all, lj, real, ov);
        else
          TrialAtom<P>(a, j, a.x[j], a.y[j], a.z[j], all, lj, real, ov);
      }
    } else if (a.doPair) {
      const int nslots = a.nlist * a.tcCap;
      if (t < nslots) {
        const int e = t / a.tcCap, s = t - e * a.tcCap;
        const int c = a.lcell[e];
        if (s < a.tcCount[c]) {
          const int j = a.tcSlots[c * a.tcCap + s];
          const bool self = j >= a.mstart && j < a.mstart + a.mlen;
          const bool pend = j >= a.cstart && j < a.cstart + a.clen;
          if (!self && !pend) TrialAtom<P>(a, j, a.x[j], a.y[j], a.z[j], a.lmask[e], lj, real, ov);
        }
      } else if (t - nslots < a.clen) {
        const int cj = t - nslots, j = a.cstart + cj;
        if ((j < a.mstart || j >= a.mstart + a.mlen) && a.mol[j] >= 0)
          TrialAtom<P>(a, j, a.cx[cj], a.cy[cj], a.cz[cj], all, lj, real, ov);
      }
    }
    part.lj = BlockSum(lj, red);
    part.real = BlockSum(real, red);
    part.overlap = BlockSum(ov, red);
  } else {
    double e = 0.0;
    const int k = (blockIdx.x - a.nPairBlocks) * kBlock + tid;
    if (a.doRecip && k < a.K) {
      const double kxk = a.kx[k], kyk = a.ky[k], kzk = a.kz[k];
      double sRN = 0.0, sIN = 0.0, sRO = 0.0, sIO = 0.0;
      if (a.oldC) {  // EwaldCached::MolReciprocal (EwaldCached.cpp:244-295)
        for (int p = 0; p < a.mlen; ++p) {
          if (fabs(a.mq[p]) < 0.000000001) continue;
          double dNew = a.nx[p] * kxk + a.ny[p] * kyk + a.nz[p] * kzk;
          double sn, cn;
          Sincos<P>(dNew, &sn, &cn);
          sRN += a.mq[p] * cn;
          sIN += a.mq[p] * sn;
        }
        sRO = a.oldC[k];
        sIO = a.oldS[k];
        a.stC[k] = sRN;
        a.stS[k] = sIN;
      } else {
        for (int p = 0; p < a.mlen; ++p) {
          if (fabs(a.mq[p]) < 0.000000001) continue;
          double dNew = a.nx[p] * kxk + a.ny[p] * kyk + a.nz[p] * kzk;
          double dOld = a.ox[p] * kxk + a.oy[p] * kyk + a.oz[p] * kzk;
          double sn, cn, so, co;
          Sincos<P>(dNew, &sn, &cn);
          Sincos<P>(dOld, &so, &co);
          sRN += a.mq[p] * cn;
          sIN += a.mq[p] * sn;
          sRO += a.mq[p] * co;
          sIO += a.mq[p] * so;
        }
      }
      double R = a.sumRref[k] + (sRN - sRO);
      double I = a.sumIref[k] + (sIN - sIO);
      a.sumRnew[k] = R;
      a.sumInew[k] = I;
      e = (R * R + I * I) * a.pref[k];
    }
    part.recip = BlockSum(e, red);
  }
  // Every other block has published its partial, after its last read of
  // the cell table: the edits can no longer be seen by this launch.
  if (FinishPartials(part, a.partials, a.counter, a.result) && tid == 0)
    for (int e = 0; e < a.nedits; ++e) {
      const int at = a.edits[2 * e], v = a.edits[2 * e + 1];
      if (at < a.tcSlotsN)
        a.tcSlots[at] = v;
      else
        a.tcCount[at - a.tcSlotsN] = v;
    }
}

// Trial cell table edits that no trial launch carries, in order.
constexpr int kEditBatch = 256;
struct EditArgs {
  int *tcCount, *tcSlots;
  int n, tcSlotsN;
  int edits[2 * kEditBatch];
};

__global__ void CellEditKernel(const EditArgs a) {
  if (threadIdx.x != 0) return;
  for (int e = 0; e < a.n; ++e) {
    const int at = a.edits[2 * e], v = a.edits[2 * e + 1];
    if (at < a.tcSlotsN)
      a.tcSlots[at] = v;
    else
      a.tcCount[at - a.tcSlotsN] = v;
  }
}

// ------------------------------------------------------------ cell list --
struct CellArgs {
  const double *x, *y, *z;
  const int *mol;  // < 0: reservoir atom, not gridded
  int natoms, ncell;
  double cs[3];
  int edge[3];
  int *cellCount, *cellCursor, *cellStart, *cellVec, *atomCell;
};

__global__ void __launch_bounds__(kBlock) CellCountKernel(const CellArgs a) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.natoms) return;
  if (a.mol[i] < 0) {
    a.atomCell[i] = -1;
    return;
  }
  int c = CellOf(a.x[i], a.y[i], a.z[i], a.cs, a.edge);
  a.atomCell[i] = c;
  atomicAdd(&a.cellCount[c], 1);
}

// The same two passes for grids of at most kCellLdsMax cells: a workgroup of
// 1024 atoms counts its atoms per cell in LDS and adds each nonzero count to
// the global one with a single atomic, and the scatter reserves one range
// per (workgroup, cell) the same way and places its atoms inside it by LDS
// atomics.  Global atomics per cell drop from one per atom to one per
// workgroup; the order inside a cell is made ascending by CellSortKernel
// either way.
constexpr int kCellLdsMax = 6144;
constexpr int kCellLdsBlock = 1024;

__global__ void __launch_bounds__(kCellLdsBlock) CellCountLdsKernel(const CellArgs a) {
  extern __shared__ int cnt[];
  for (int c = threadIdx.x; c < a.ncell; c += kCellLdsBlock) cnt[c] = 0;
  __syncthreads();
  const int i = blockIdx.x * kCellLdsBlock + threadIdx.x;
  if (i < a.natoms) {
    if (a.mol[i] < 0) {
      a.atomCell[i] = -1;
    } else {
      const int c = CellOf(a.x[i], a.y[i], a.z[i], a.cs, a.edge);
      a.atomCell[i] = c;
      atomicAdd(&cnt[c], 1);
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < a.ncell; c += kCellLdsBlock)
    if (cnt[c]) atomicAdd(&a.cellCount[c], cnt[c]);
}

__global__ void __launch_bounds__(kCellLdsBlock) CellScatterLdsKernel(const CellArgs a) {
  extern __shared__ int cnt[];  // [ncell] counts, then [ncell] range starts
  int *base = cnt + a.ncell;
  for (int c = threadIdx.x; c < a.ncell; c += kCellLdsBlock) cnt[c] = 0;
  __syncthreads();
  const int i = blockIdx.x * kCellLdsBlock + threadIdx.x;
  const int c = i < a.natoms ? a.atomCell[i] : -1;
  const int r = c >= 0 ? atomicAdd(&cnt[c], 1) : 0;
  __syncthreads();
  for (int k = threadIdx.x; k < a.ncell; k += kCellLdsBlock)
    if (cnt[k]) base[k] = atomicAdd(&a.cellCursor[k], cnt[k]);
  __syncthreads();
  if (c >= 0) a.cellVec[base[c] + r] = i;
}

// Exclusive scan of the per-cell counts (one workgroup; ncell is small).
__global__ void __launch_bounds__(1024) CellScanKernel(const CellArgs a) {
  __shared__ int chunkSum[1024];
  const int t = threadIdx.x, n = a.ncell;
  const int per = (n + 1023) / 1024;
  const int lo = min(n, t * per), hi = min(n, lo + per);
  int s = 0;
  for (int c = lo; c < hi; ++c) s += a.cellCount[c];
  chunkSum[t] = s;
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int i = 0; i < 1024; ++i) {
      int v = chunkSum[i];
      chunkSum[i] = run;
      run += v;
    }
    a.cellStart[n] = run;
  }
  __syncthreads();
  int run = chunkSum[t];
  for (int c = lo; c < hi; ++c) {
    a.cellStart[c] = run;
    a.cellCursor[c] = run;
    run += a.cellCount[c];
  }
}

__global__ void __launch_bounds__(kBlock) CellScatterKernel(const CellArgs a) {
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.natoms || a.atomCell[i] < 0) return;
  int pos = atomicAdd(&a.cellCursor[a.atomCell[i]], 1);
  a.cellVec[pos] = i;
}

// Ascending order inside each cell (CellList.cpp:278-280) by rank: each
// atom's slot is the number of smaller indices in its cell.
__global__ void __launch_bounds__(kBlock) CellSortKernel(const CellArgs a) {
  extern __shared__ int idx[];
  const int c = blockIdx.x, s = a.cellStart[c], n = a.cellStart[c + 1] - s;
  for (int i = threadIdx.x; i < n; i += kBlock) idx[i] = a.cellVec[s + i];
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += kBlock) {
    int v = idx[i], r = 0;
    for (int j = 0; j < n; ++j) r += (idx[j] < v);
    a.cellVec[s + r] = v;
  }
}

// ------------------------------------------------------------- box inter --
// One workgroup per (cell, neighbour slot); pairs i < j of different
// molecules (CalculateEnergy.cpp:205-250).  The neighbour cell's atoms are
// staged in LDS; threads stride over the (i, j) product.
struct BoxInterArgs {
  BoxGeom g;
  const double *x, *y, *z, *q;
  const int *type, *mol;
  const double *sigSq, *epsCn, *nHalf;
  int ntypes;
  const int *cellStart, *cellVec, *nbr;
  Partial4 *partials;
  unsigned *counter;
  double *result;
};

__global__ void __launch_bounds__(kBlock) BoxInterKernel(const BoxInterArgs a) {
  __shared__ double sx[kBlock], sy[kBlock], sz[kBlock], sq[kBlock];
  __shared__ int sidx[kBlock], stype[kBlock], smol[kBlock];
  __shared__ double red[kBlock / kWave];
  const int cell = blockIdx.x / 27, nb = a.nbr[blockIdx.x];
  const int cs = a.cellStart[cell], ce = a.cellStart[cell + 1];
  const int ns = a.cellStart[nb], ne = a.cellStart[nb + 1];
  const int ni = ce - cs;
  double lj = 0.0, real = 0.0;
  for (int tile = ns; tile < ne; tile += kBlock) {
    const int tn = min(kBlock, ne - tile);
    if (threadIdx.x < tn) {
      int j = a.cellVec[tile + threadIdx.x];
      sidx[threadIdx.x] = j;
      sx[threadIdx.x] = a.x[j];
      sy[threadIdx.x] = a.y[j];
      sz[threadIdx.x] = a.z[j];
      sq[threadIdx.x] = a.q[j];
      stype[threadIdx.x] = a.type[j];
      smol[threadIdx.x] = a.mol[j];
    }
    __syncthreads();
    for (int p = threadIdx.x; p < ni * tn; p += kBlock) {
      const int i = a.cellVec[cs + p / tn], jj = p % tn, j = sidx[jj];
      if (i < j && a.mol[i] != smol[jj]) {
        double d2 = DistSq(a.g, a.x[i], a.y[i], a.z[i], sx[jj], sy[jj], sz[jj]);
        if (a.g.rCutSq > d2) {
          if (a.g.elect) {
            double qiqj = a.q[i] * sq[jj] * kQQFact;
            if (qiqj != 0.0 && !(a.g.rCutCoulSq < d2)) real += CoulombEnergy(a.g, d2, qiqj);
          }
          if (!(a.g.rCutLJSq < d2)) {
            int idx = a.type[i] + stype[jj] * a.ntypes;
            lj += LJEnergy(d2, a.sigSq[idx], a.epsCn[idx], a.nHalf[idx]);
          }
        }
      }
    }
    __syncthreads();
  }
  Partial4 part{0, 0, 0, 0};
  part.lj = BlockSum(lj, red);
  part.real = BlockSum(real, red);
  FinishPartials(part, a.partials, a.counter, a.result);
}

// ------------------------------------------------------ structure factor --
// S(k) = sum_j q_j (cos, sin)(k . r_j) over all charged atoms
// (Ewald.cpp:193-274), then the per-k energy prefact (Re^2 + Im^2)
// (Ewald.cpp:375-406).
//
// The k list comes in GOMC's x, y, z loop order, so it is made of runs of
// consecutive nz at fixed (nx, ny).  KSegmentsKernel cuts it into segments:
// pieces of a run of at most kSeg vectors.  Along a segment the phase of an
// atom advances by one complex product, e^{i k_{n+1} . r} =
// e^{i k_n . r} e^{i dz z}, with dz = 2 pi / L_z for every segment, so a
// segment costs one sincos per atom (its first vector; the step's phase is
// made once per atom) and one complex product per vector.  The first pass
// takes one segment per thread and one chunk of atoms per work item
// (per-chunk sums into part[chunk][k][2]); the second adds the chunks in
// order, one k per thread.  Both passes read the segment count from device
// memory, so a launch never waits on the host.
struct SumsArgs {
  const double *x, *y, *z, *q;
  const int *mol;  // may be null: every atom counts
  int natoms, K;
  const double *kx, *ky, *kz, *pref;
  const int *kseg;   // segment starts [nseg + 1], kseg[nseg] = K
  const int *nseg;   // device
  double cx, cy, cz;  // KVectorsOrth's spacings 2 pi / L_d (cz: every segment's nz step)
  int nmax[3];        // KVectorsOrth's |n_d| bounds: the phase tables' ranges
  int tabAtoms;       // atoms per LDS tile (their phase tables fill the dynamic LDS)
  double *sumR, *sumI;
  double *part;      // [chunks][K][2]
  Partial4 *partials;
  unsigned *counter;
  double *result;
};

__device__ __forceinline__ bool RunStart(const double *kx, const double *ky, const double *kz,
                                         int k) {
  return k == 0 || kx[k] != kx[k - 1] || ky[k] != ky[k - 1] || !(kz[k] > kz[k - 1]);
}

// One workgroup of kSegThreads: each thread takes a contiguous range of the
// K vectors; a max-scan carries the run start in effect into each range, a
// sum-scan places each range's segments.
__global__ void __launch_bounds__(kSegThreads) KSegmentsKernel(const double *kx, const double *ky,
                                                               const double *kz, int K, int *kseg,
                                                               int *nseg) {
  __shared__ int sRun[kSegThreads], sCnt[kSegThreads];
  const int t = threadIdx.x;
  const int per = (K + kSegThreads - 1) / kSegThreads;
  const int lo = min(K, t * per), hi = min(K, lo + per);
  int last = -1;
  for (int k = lo; k < hi; ++k)
    if (RunStart(kx, ky, kz, k)) last = k;
  sRun[t] = last;
  __syncthreads();
  for (int off = 1; off < kSegThreads; off <<= 1) {  // inclusive max-scan
    const int v = t >= off ? sRun[t - off] : -1;
    __syncthreads();
    sRun[t] = max(sRun[t], v);
    __syncthreads();
  }
  const int runAtLo = t > 0 ? sRun[t - 1] : -1;
  int run = runAtLo, cnt = 0;
  for (int k = lo; k < hi; ++k) {
    if (RunStart(kx, ky, kz, k)) run = k;
    if ((k - run) % kSeg == 0) ++cnt;
  }
  sCnt[t] = cnt;
  __syncthreads();
  for (int off = 1; off < kSegThreads; off <<= 1) {  // inclusive sum-scan
    const int v = t >= off ? sCnt[t - off] : 0;
    __syncthreads();
    sCnt[t] += v;
    __syncthreads();
  }
  int out = sCnt[t] - cnt;
  run = runAtLo;
  for (int k = lo; k < hi; ++k) {
    if (RunStart(kx, ky, kz, k)) run = k;
    if ((k - run) % kSeg == 0) kseg[out++] = k;
  }
  if (t == kSegThreads - 1) {
    kseg[sCnt[t]] = K;
    *nseg = sCnt[t];
  }
}

constexpr int kSumsTabBytes = 48 * 1024;  // dynamic LDS of the phase tables
constexpr int kSumsTabMaxAtoms = 64;

__host__ __device__ inline int SumsTabEntries(const int nmax[3]) {
  return nmax[0] + nmax[1] + nmax[2] + 3;
}

__global__ void __launch_bounds__(kBlock) RecipSumsPartKernel(const SumsArgs a) {
  // per atom of the tile: e^{i n c_d r_d} for n = 0..nmax_d on each axis (a
  // negative n is the conjugate), so the first phase of a segment at
  // (nx, ny, nz0) is two complex products instead of a sincos; the nz step
  // is the z table's n = 1
  extern __shared__ double2 tab[];
  __shared__ double sq[kSumsTabMaxAtoms];
  const int nseg = *a.nseg;
  const int segBlocks = max(1, (nseg + kBlock - 1) / kBlock);
  const int chunks = SumsChunks(nseg, a.natoms, a.K);
  const int aChunk = (a.natoms + chunks - 1) / chunks;
  const int E = SumsTabEntries(a.nmax), oy = a.nmax[0] + 1, oz = oy + a.nmax[1] + 1;
  const int T = a.tabAtoms;
  for (int item = blockIdx.x; item < segBlocks * chunks; item += gridDim.x) {
    const int seg = (item % segBlocks) * kBlock + threadIdx.x, chunk = item / segBlocks;
    const bool active = seg < nseg;
    const int k0 = active ? a.kseg[seg] : 0;
    const int len = active ? a.kseg[seg + 1] - k0 : 0;
    // the segment's integer triple (k = c n, one rounding, so rint recovers n)
    const int nx = active ? min(a.nmax[0], max(0, int(rint(a.kx[k0] / a.cx)))) : 0;
    const int ny = active ? min(a.nmax[1], max(-a.nmax[1], int(rint(a.ky[k0] / a.cy)))) : 0;
    const int nz = active ? min(a.nmax[2], max(-a.nmax[2], int(rint(a.kz[k0] / a.cz)))) : 0;
    const double syn = ny < 0 ? -1.0 : 1.0, szn = nz < 0 ? -1.0 : 1.0;
    const int iy = oy + abs(ny), iz = oz + abs(nz);
    double sR[kSeg], sI[kSeg];
#pragma unroll
    for (int i = 0; i < kSeg; ++i) sR[i] = sI[i] = 0.0;
    const int a0 = min(a.natoms, chunk * aChunk), a1 = min(a.natoms, a0 + aChunk);
    for (int tile = a0; tile < a1; tile += T) {
      const int tn = min(T, a1 - tile);
      __syncthreads();
      if (int(threadIdx.x) < tn) {
        const int j = tile + threadIdx.x;
        sq[threadIdx.x] = (!a.mol || a.mol[j] >= 0) ? a.q[j] : 0.0;  // reservoir adds nothing
      }
      // each (atom, axis) table by one thread: n = 1 by sincos, then one
      // complex product per n
      for (int w = threadIdx.x; w < 3 * tn; w += kBlock) {
        const int at = w / 3, ax = w - 3 * at, j = tile + at;
        const double r = ax == 0 ? a.x[j] : (ax == 1 ? a.y[j] : a.z[j]);
        const double c = ax == 0 ? a.cx : (ax == 1 ? a.cy : a.cz);
        const int o = ax == 0 ? 0 : (ax == 1 ? oy : oz), n1 = a.nmax[ax];
        double2 *t = tab + size_t(at) * E + o;
        double s1, c1;
        sincos(c * r, &s1, &c1);
        t[0] = make_double2(1.0, 0.0);
        double re = c1, im = s1;
        for (int n = 1; n <= n1; ++n) {
          t[n] = make_double2(re, im);
          const double nr = fma(re, c1, -(im * s1));
          im = fma(re, s1, im * c1);
          re = nr;
        }
      }
      __syncthreads();
      if (active)
        for (int jj = 0; jj < tn; ++jj) {
          const double qj = sq[jj];
          if (fabs(qj) < 0.000000001) continue;
          const double2 *t = tab + size_t(jj) * E;
          const double2 ex = t[nx], ey = t[iy], ez = t[iz], st = t[oz + 1];
          const double eyi = syn * ey.y, ezi = szn * ez.y;
          const double pr = fma(ex.x, ey.x, -(ex.y * eyi)), pi = fma(ex.x, eyi, ex.y * ey.x);
          double re = fma(pr, ez.x, -(pi * ezi)), im = fma(pr, ezi, pi * ez.x);
          const double ss = st.y, cs = st.x;
#pragma unroll
          for (int i = 0; i < kSeg; ++i) {
            if (i < len) {
              sR[i] = fma(qj, re, sR[i]);
              sI[i] = fma(qj, im, sI[i]);
            }
            const double nr = fma(re, cs, -(im * ss));
            im = fma(re, ss, im * cs);
            re = nr;
          }
        }
    }
    if (active) {
      double *out = a.part + (size_t(chunk) * a.K + k0) * 2;
#pragma unroll
      for (int i = 0; i < kSeg; ++i)
        if (i < len) {
          out[2 * i] = sR[i];
          out[2 * i + 1] = sI[i];
        }
    }
  }
}

__global__ void __launch_bounds__(kBlock) RecipSumsKernel(const SumsArgs a) {
  __shared__ double red[kBlock / kWave];
  const int k = blockIdx.x * kBlock + threadIdx.x;
  const bool active = k < a.K;
  const int chunks = SumsChunks(*a.nseg, a.natoms, a.K);
  double sR = 0.0, sI = 0.0, e = 0.0;
  if (active) {
    for (int c = 0; c < chunks; ++c) {
      sR += a.part[(size_t(c) * a.K + k) * 2];
      sI += a.part[(size_t(c) * a.K + k) * 2 + 1];
    }
    a.sumR[k] = sR;
    a.sumI[k] = sI;
    e = (sR * sR + sI * sI) * a.pref[k];
  }
  Partial4 part{0, 0, 0, 0};
  part.recip = BlockSum(e, red);
  FinishPartials(part, a.partials, a.counter, a.result);
}

// Insertion (+) or removal (-) of one molecule: SwapDestRecip /
// SwapSourceRecip (Ewald.cpp:478-531, 657-710).
struct SwapArgs {
  int K, n;
  double sign;
  double px[kMaxMolAtoms], py[kMaxMolAtoms], pz[kMaxMolAtoms], pq[kMaxMolAtoms];
  const double *kx, *ky, *kz, *pref, *sumRref, *sumIref;
  double *sumRnew, *sumInew;
  double *stC, *stS;  // EwaldCached insertion: the new molecule's terms
  Partial4 *partials;
  unsigned *counter;
  double *result;
};

template <int P>
__global__ void __launch_bounds__(kBlock) SwapRecipKernel(const SwapArgs a) {
  __shared__ double red[kBlock / kWave];
  const int k = blockIdx.x * kBlock + threadIdx.x;
  double e = 0.0;
  if (k < a.K) {
    double sR = 0.0, sI = 0.0;
    for (int p = 0; p < a.n; ++p) {
      if (fabs(a.pq[p]) < 0.000000001) continue;
      double dot = a.px[p] * a.kx[k] + a.py[p] * a.ky[k] + a.pz[p] * a.kz[k];
      double s, c;
      Sincos<P>(dot, &s, &c);
      sR += a.pq[p] * c;
      sI += a.pq[p] * s;
    }
    if (a.stC) {
      a.stC[k] = sR;
      a.stS[k] = sI;
    }
    double R = (a.sign > 0) ? a.sumRref[k] + sR : a.sumRref[k] - sR;
    double I = (a.sign > 0) ? a.sumIref[k] + sI : a.sumIref[k] - sI;
    a.sumRnew[k] = R;
    a.sumInew[k] = I;
    e = (R * R + I * I) * a.pref[k];
  }
  Partial4 part{0, 0, 0, 0};
  part.recip = BlockSum(e, red);
  FinishPartials(part, a.partials, a.counter, a.result);
}

// EwaldCached::BoxReciprocalSetup's per-molecule cache (EwaldCached.cpp:
// 92-141): for each k (one per thread) and each molecule in the box,
// sum q (cos, sin)(k . r) over its charged atoms in atom order -- the same
// arithmetic the trial kernel uses for a molecule's old terms.
struct CacheArgs {
  const double *x, *y, *z, *q;
  const int *mol, *molStart;
  int nmols, K;
  const double *kx, *ky, *kz;
  double *cacheC, *cacheS;
};

__global__ void __launch_bounds__(kBlock) CacheBuildKernel(const CacheArgs a) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.K) return;
  const double kxk = a.kx[k], kyk = a.ky[k], kzk = a.kz[k];
  for (int m = 0; m < a.nmols; ++m) {
    const int s = a.molStart[m], e = a.molStart[m + 1];
    double c = 0.0, sn = 0.0;
    if (a.mol[s] >= 0)
      for (int p = s; p < e; ++p) {
        if (fabs(a.q[p]) < 0.000000001) continue;
        double dot = a.x[p] * kxk + a.y[p] * kyk + a.z[p] * kzk;
        double sv, cv;
        sincos(dot, &sv, &cv);
        c += a.q[p] * cv;
        sn += a.q[p] * sv;
      }
    a.cacheC[int64_t(m) * a.K + k] = c;
    a.cacheS[int64_t(m) * a.K + k] = sn;
  }
}

// Trial positions of one atom against the live atoms
// (CalculateEnergy.cpp:741-783): one workgroup per trial.  With the trial
// cell table (tcCap > 0) a workgroup walks the 27 cells around its
// position's cell, as ParticleInter's EnumerateLocal does; without it, every
// atom.
struct ParticleArgs {
  BoxGeom g;
  const double *x, *y, *z, *q;
  const int *type, *mol;
  const double *sigSq, *epsCn, *nHalf;
  int ntypes, natoms, atype, skipMol;
  double charge;
  const double *tx, *ty, *tz;
  double *en, *real, *overlap;
  int tcCap;
  const int *tcCount, *tcSlots, *nbr;
  double cs[3];
  int edge[3];
};

template <int P>
__device__ __forceinline__ void ParticleAtom(const ParticleArgs &a, int j, double px, double py,
                                             double pz, double &lj, double &real, double &ov) {
  const double d2 = DistSq(a.g, px, py, pz, a.x[j], a.y[j], a.z[j]);
  if (a.g.rCutSq > d2) {
    if (d2 < a.g.rCutLowSq) ov = 1.0;
    if (!(a.g.rCutLJSq < d2)) {
      const int idx = a.atype + a.type[j] * a.ntypes;
      lj += LJ<P>(d2, a.sigSq[idx], a.epsCn[idx], a.nHalf[idx]);
    }
    if (a.g.elect) {
      const double qiqj = a.q[j] * a.charge * kQQFact;
      if (qiqj != 0.0 && !(a.g.rCutCoulSq < d2)) real += Coul<P>(a.g, d2, qiqj);
    }
  }
}

template <int P>
__global__ void __launch_bounds__(kBlock) ParticleInterKernel(const ParticleArgs a) {
  __shared__ double red[kBlock / kWave];
  const int t = blockIdx.x;
  const double px = a.tx[t], py = a.ty[t], pz = a.tz[t];
  double lj = 0.0, real = 0.0, ov = 0.0;
  if (a.tcCap > 0) {
    const int c0 = CellOf(px, py, pz, a.cs, a.edge);
    for (int i = threadIdx.x; i < 27 * a.tcCap; i += kBlock) {
      const int n = i / a.tcCap, s = i - n * a.tcCap;
      const int c = a.nbr[c0 * 27 + n];
      if (s >= a.tcCount[c]) continue;
      const int j = a.tcSlots[c * a.tcCap + s];
      if (a.mol[j] == a.skipMol) continue;
      ParticleAtom<P>(a, j, px, py, pz, lj, real, ov);
    }
  } else {
    for (int j = threadIdx.x; j < a.natoms; j += kBlock) {
      if (a.mol[j] == a.skipMol || a.mol[j] < 0) continue;
      ParticleAtom<P>(a, j, px, py, pz, lj, real, ov);
    }
  }
  lj = BlockSum(lj, red);
  real = BlockSum(real, red);
  ov = BlockSum(ov, red);
  if (threadIdx.x == 0) {
    a.en[t] = lj;
    a.real[t] = real;
    a.overlap[t] = ov;
  }
}

}  // namespace

namespace gomc {

int EnsurePartials(gomc_box *b, int nblocks) {
  if (nblocks <= b->maxPartials) return GOMC_OK;
  if (b->partials) (void)hipFree(b->partials);
  b->partials = nullptr;
  int n = std::max(nblocks, 1024);
  HIP_TRY(b, hipMalloc(&b->partials, sizeof(Partial4) * n));
  b->maxPartials = n;
  return GOMC_OK;
}

int Sync(gomc_box *b) {
  HIP_TRY(b, hipStreamSynchronize(b->stream));
  HIP_TRY(b, hipGetLastError());
  return GOMC_OK;
}

// ------------------------------------------------------ trial cell table --
// A position the cell walk can take: inside the box, as the reference's
// moves leave every trial position (WrapPBC).
static bool InBox(const gomc_box *b, double x, double y, double z) {
  return x >= 0.0 && x <= b->geom.L[0] && y >= 0.0 && y <= b->geom.L[1] && z >= 0.0 &&
         z <= b->geom.L[2];
}

// The 27 cells within one cell of cell c (periodic; edges of at least 3
// make them distinct).
static void CellNeighbours(const gomc_box *b, int c, int out[27]) {
  const int *E = b->edge;
  const int ix = c / (E[1] * E[2]), iy = (c / E[2]) % E[1], iz = c % E[2];
  int n = 0;
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz)
        out[n++] = ((ix + dx + E[0]) % E[0]) * E[1] * E[2] + ((iy + dy + E[1]) % E[1]) * E[2] +
                   (iz + dz + E[2]) % E[2];
}

// Builds the table from the host coordinates of the live molecules and
// uploads it: capacity the fullest cell plus a quarter (at least 8 more).
int TrialCellsEnsure(gomc_box *b) {
  if (b->tcValid) return GOMC_OK;
  const int nc = b->ncell;
  b->tcEdits.clear();
  b->hTcCount.assign(nc, 0);
  b->hTcAtomCell.assign(b->natoms, -1);
  b->hTcAtomSlot.assign(b->natoms, -1);
  for (int m = 0; m < b->nmols; ++m) {
    if (!b->molBox.empty() && b->molBox[m] != 0) continue;
    for (int a = b->molStart[m]; a < b->molStart[m] + b->molLen[m]; ++a) {
      const int c = CellOf(b->hx[a], b->hy[a], b->hz[a], b->cellSize, b->edge);
      b->hTcAtomCell[a] = c;
      b->hTcAtomSlot[a] = b->hTcCount[c]++;
    }
  }
  const int mx = nc > 0 ? *std::max_element(b->hTcCount.begin(), b->hTcCount.end()) : 0;
  const int cap = mx + std::max(8, mx / 4);
  if (size_t(nc) * size_t(cap) + size_t(nc) > size_t(INT32_MAX))
    return Fail(b, GOMC_ERR_ARG, "trial cell table beyond 2^31 entries");
  b->tcCap = cap;
  b->hTcSlots.assign(size_t(nc) * cap, -1);
  for (int a = 0; a < b->natoms; ++a)
    if (b->hTcAtomCell[a] >= 0) b->hTcSlots[size_t(b->hTcAtomCell[a]) * cap + b->hTcAtomSlot[a]] = a;
  if (size_t(nc) > b->tcCountAlloc) {
    if (b->tcCount) (void)hipFree(b->tcCount);
    b->tcCount = nullptr;
    b->tcCountAlloc = 0;
    HIP_TRY(b, hipMalloc(&b->tcCount, sizeof(int) * nc));
    b->tcCountAlloc = nc;
  }
  if (b->hTcSlots.size() > b->tcSlotsAlloc) {
    if (b->tcSlots) (void)hipFree(b->tcSlots);
    b->tcSlots = nullptr;
    b->tcSlotsAlloc = 0;
    HIP_TRY(b, hipMalloc(&b->tcSlots, sizeof(int) * b->hTcSlots.size()));
    b->tcSlotsAlloc = b->hTcSlots.size();
  }
  HIP_TRY(b, hipMemcpyAsync(b->tcCount, b->hTcCount.data(), sizeof(int) * nc,
                            hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->tcSlots, b->hTcSlots.data(), sizeof(int) * b->hTcSlots.size(),
                            hipMemcpyHostToDevice, b->stream));
  int rc = Sync(b);
  if (rc) return rc;
  b->tcValid = true;
  return GOMC_OK;
}

// Applies the queued edits on the box's stream, ahead of whatever is
// launched next; no wait.
int TrialCellsFlushEdits(gomc_box *b) {
  const int n = int(b->tcEdits.size() / 2);
  for (int e0 = 0; e0 < n; e0 += kEditBatch) {
    EditArgs a;
    a.tcCount = b->tcCount;
    a.tcSlots = b->tcSlots;
    a.tcSlotsN = b->ncell * b->tcCap;
    a.n = std::min(kEditBatch, n - e0);
    std::copy(b->tcEdits.begin() + 2 * e0, b->tcEdits.begin() + 2 * (e0 + a.n), a.edits);
    hipLaunchKernelGGL(CellEditKernel, dim3(1), dim3(kWave), 0, b->stream, a);
    HIP_TRY(b, hipGetLastError());
  }
  b->tcEdits.clear();
  return GOMC_OK;
}

// CellList::RemoveMol / AddMol (src/CellList.h:28-29) for an accepted
// displacement of molecule m:
// an atom that changes cell leaves its slot to its old cell's last atom and
// takes the next slot of its new cell.  A full cell makes the table invalid
// (rebuilt, larger, before its next use).
static void TrialCellsCommit(gomc_box *b, int m, const double *x, const double *y,
                             const double *z) {
  if (!b->tcValid || !TrialCellsUseful(b)) return;
  const int cap = b->tcCap, slotsN = b->ncell * cap;
  std::vector<int> &E = b->tcEdits;
  for (int p = 0; p < b->molLen[m]; ++p) {
    const int a = b->molStart[m] + p;
    const int c0 = b->hTcAtomCell[a], c1 = CellOf(x[p], y[p], z[p], b->cellSize, b->edge);
    if (c0 == c1) continue;
    if (c0 < 0 || b->hTcCount[c1] >= cap) {
      TrialCellsInvalidate(b);
      return;
    }
    const int slot = b->hTcAtomSlot[a], last = --b->hTcCount[c0];
    const int moved = b->hTcSlots[size_t(c0) * cap + last];
    b->hTcSlots[size_t(c0) * cap + slot] = moved;
    b->hTcAtomSlot[moved] = slot;
    b->hTcSlots[size_t(c0) * cap + last] = -1;
    E.insert(E.end(), {c0 * cap + slot, moved, slotsN + c0, last});
    const int ns = b->hTcCount[c1]++;
    b->hTcSlots[size_t(c1) * cap + ns] = a;
    b->hTcAtomSlot[a] = ns;
    b->hTcAtomCell[a] = c1;
    E.insert(E.end(), {c1 * cap + ns, a, slotsN + c1, ns + 1});
  }
}

int LaunchTrial(gomc_box *b, int m, const double *ox, const double *oy, const double *oz,
                const double *nx, const double *ny, const double *nz, bool doPair,
                bool doRecip, double out[4]) {
  if (m < 0 || m >= b->nmols) return Fail(b, GOMC_ERR_ARG, "molecule index out of range");
  const int len = b->molLen[m];
  if (len > kMaxMolAtoms) return Fail(b, GOMC_ERR_ARG, "molecule has more than 16 atoms");
  if (doRecip && !b->recipReady) return Fail(b, GOMC_ERR_STATE, "k-vectors not initialised");
  TrialArgs a;
  std::memset(&a, 0, sizeof(a));
  a.g = b->geom;
  a.x = b->x;
  a.y = b->y;
  a.z = b->z;
  a.q = b->q;
  a.type = b->type;
  a.mol = b->mol;
  a.sigSq = b->tSigSq;
  a.epsCn = b->tEpsCn;
  a.nHalf = b->tNHalf;
  a.ntypes = b->ntypes;
  a.natoms = b->natoms;
  a.mstart = b->molStart[m];
  a.mlen = len;
  const int kind = b->molKind[m];
  for (int p = 0; p < len; ++p) {
    a.ox[p] = ox[p];
    a.oy[p] = oy[p];
    a.oz[p] = oz[p];
    a.nx[p] = nx[p];
    a.ny[p] = ny[p];
    a.nz[p] = nz[p];
    a.mq[p] = b->kindCharge[kind][p];
    a.mtype[p] = b->kindType[kind][p];
  }
  a.doPair = doPair;
  a.doRecip = doRecip && b->ewald;
  const int nRecip = a.doRecip ? (b->K + kBlock - 1) / kBlock : 0;
  a.K = b->K;
  a.kx = b->kx;
  a.ky = b->ky;
  a.kz = b->kz;
  a.pref = b->pref;
  a.sumRref = b->sumRref;
  a.sumIref = b->sumIref;
  a.sumRnew = b->sumRnew;
  a.sumInew = b->sumInew;
  b->stageMol = -1;
  if (b->cached && a.doRecip) {
    a.oldC = b->cacheC + int64_t(m) * b->K;
    a.oldS = b->cacheS + int64_t(m) * b->K;
    a.stC = b->stageC;
    a.stS = b->stageS;
    b->stageMol = m;
  }
  a.cstart = 0;
  a.clen = 0;
  if (b->pendingMol >= 0) {
    a.cstart = b->molStart[b->pendingMol];
    a.clen = b->molLen[b->pendingMol];
    for (int p = 0; p < a.clen; ++p) {
      a.cx[p] = b->pend[0][p];
      a.cy[p] = b->pend[1][p];
      a.cz[p] = b->pend[2][p];
    }
  }
  a.wx = b->x;
  a.wy = b->y;
  a.wz = b->z;
  // the cells within one cell of every old and new position, when walking
  // them is less work than a scan of every atom
  a.nlist = 0;
  if (doPair && TrialCellsUseful(b)) {
    int rc = TrialCellsEnsure(b);
    if (rc) return rc;
    if (b->tcMark.size() != size_t(b->ncell)) b->tcMark.assign(b->ncell, -1);
    bool fits = true;
    for (int side = 0; side < 2 && fits; ++side)
      for (int p = 0; p < len && fits; ++p) {
        const double *px = side ? nx : ox, *py = side ? ny : oy, *pz = side ? nz : oz;
        if (!InBox(b, px[p], py[p], pz[p])) {  // only the scan of every atom takes any image
          fits = false;
          break;
        }
        int nb[27];
        CellNeighbours(b, CellOf(px[p], py[p], pz[p], b->cellSize, b->edge), nb);
        const uint32_t bit = 1u << (side ? kNewBits + p : p);
        for (int n = 0; n < 27; ++n) {
          int &mk = b->tcMark[nb[n]];
          if (mk < 0) {
            if (a.nlist == kTrialCells) {
              fits = false;
              break;
            }
            mk = a.nlist;
            a.lcell[a.nlist] = nb[n];
            a.lmask[a.nlist++] = 0;
          }
          a.lmask[mk] |= bit;
        }
      }
    for (int e = 0; e < a.nlist; ++e) b->tcMark[a.lcell[e]] = -1;
    if (!fits || size_t(a.nlist) * b->tcCap + a.clen >= size_t(b->natoms)) a.nlist = 0;
  }
  a.tcCap = b->tcCap;
  a.tcCount = b->tcCount;
  a.tcSlots = b->tcSlots;
  a.tcSlotsN = b->ncell * b->tcCap;
  a.nedits = 0;
  if (!b->tcEdits.empty()) {
    if (b->tcEdits.size() <= size_t(2 * kTrialEdits)) {
      a.nedits = int(b->tcEdits.size() / 2);
      std::copy(b->tcEdits.begin(), b->tcEdits.end(), a.edits);
    } else {
      int rc = TrialCellsFlushEdits(b);
      if (rc) return rc;
    }
  }
  if (doPair)
    a.nPairBlocks = a.nlist > 0 ? int((size_t(a.nlist) * b->tcCap + a.clen + kBlock - 1) / kBlock)
                                : (b->natoms + kBlock - 1) / kBlock;
  else
    a.nPairBlocks = 0;
  int nblocks = a.nPairBlocks + nRecip;
  if (nblocks == 0) nblocks = 1;
  int rc = EnsurePartials(b, nblocks);
  if (rc) return rc;
  a.partials = b->partials;
  a.counter = b->counter;
  a.result = b->devRes;
  if (b->precision == 1)
    hipLaunchKernelGGL(TrialKernel<1>, dim3(nblocks), dim3(kBlock), 0, b->stream, a);
  else
    hipLaunchKernelGGL(TrialKernel<0>, dim3(nblocks), dim3(kBlock), 0, b->stream, a);
  HIP_TRY(b, hipGetLastError());
  rc = Sync(b);
  if (rc) {
    TrialCellsInvalidate(b);
    return rc;
  }
  b->pendingMol = -1;
  if (a.nedits) b->tcEdits.clear();
  for (int i = 0; i < 3; ++i) out[i] = b->hostRes[i];
  out[3] = b->hostRes[3] > 0.0 ? 1.0 : 0.0;  // any block saw a pair inside RcutLow
  return GOMC_OK;
}

int BuildCells(gomc_box *b) {
  CellArgs a;
  a.x = b->x;
  a.y = b->y;
  a.z = b->z;
  a.mol = b->mol;
  a.natoms = b->natoms;
  a.ncell = b->ncell;
  for (int i = 0; i < 3; ++i) {
    a.cs[i] = b->cellSize[i];
    a.edge[i] = b->edge[i];
  }
  a.cellCount = b->cellCount;
  a.cellCursor = b->cellCursor;
  a.cellStart = b->cellStart;
  a.cellVec = b->cellVec;
  a.atomCell = b->atomCell;
  HIP_TRY(b, hipMemsetAsync(b->cellCount, 0, sizeof(int) * b->ncell, b->stream));
  if (b->ncell <= kCellLdsMax) {
    const int g = (b->natoms + kCellLdsBlock - 1) / kCellLdsBlock;
    const size_t lds = sizeof(int) * size_t(b->ncell);
    hipLaunchKernelGGL(CellCountLdsKernel, dim3(g), dim3(kCellLdsBlock), lds, b->stream, a);
    hipLaunchKernelGGL(CellScanKernel, dim3(1), dim3(1024), 0, b->stream, a);
    hipLaunchKernelGGL(CellScatterLdsKernel, dim3(g), dim3(kCellLdsBlock), 2 * lds, b->stream, a);
  } else {
    const int g = (b->natoms + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(CellCountKernel, dim3(g), dim3(kBlock), 0, b->stream, a);
    hipLaunchKernelGGL(CellScanKernel, dim3(1), dim3(1024), 0, b->stream, a);
    hipLaunchKernelGGL(CellScatterKernel, dim3(g), dim3(kBlock), 0, b->stream, a);
  }
  // largest cell bounds the LDS of the sort
  std::vector<int> counts(b->ncell);
  HIP_TRY(b, hipMemcpyAsync(counts.data(), b->cellCount, sizeof(int) * b->ncell,
                            hipMemcpyDeviceToHost, b->stream));
  int rc = Sync(b);
  if (rc) return rc;
  b->hCellCount = counts;
  int mx = *std::max_element(counts.begin(), counts.end());
  if (mx * sizeof(int) > 60 * 1024) return Fail(b, GOMC_ERR_ARG, "cell too full for device sort");
  hipLaunchKernelGGL(CellSortKernel, dim3(b->ncell), dim3(kBlock), std::max(mx, 1) * sizeof(int),
                     b->stream, a);
  HIP_TRY(b, hipGetLastError());
  return GOMC_OK;
}

int BoxInter(gomc_box *b, double *lj, double *real) {
  int rc = BuildCells(b);
  if (rc) return rc;
  BoxInterArgs a;
  a.g = b->geom;
  a.x = b->x;
  a.y = b->y;
  a.z = b->z;
  a.q = b->q;
  a.type = b->type;
  a.mol = b->mol;
  a.sigSq = b->tSigSq;
  a.epsCn = b->tEpsCn;
  a.nHalf = b->tNHalf;
  a.ntypes = b->ntypes;
  a.cellStart = b->cellStart;
  a.cellVec = b->cellVec;
  a.nbr = b->nbr;
  const int nblocks = b->ncell * 27;
  rc = EnsurePartials(b, nblocks);
  if (rc) return rc;
  a.partials = b->partials;
  a.counter = b->counter;
  a.result = b->devRes;
  hipLaunchKernelGGL(BoxInterKernel, dim3(nblocks), dim3(kBlock), 0, b->stream, a);
  HIP_TRY(b, hipGetLastError());
  rc = Sync(b);
  if (rc) return rc;
  *lj = b->hostRes[0];
  *real = b->hostRes[1];
  return GOMC_OK;
}

// Both passes on stream s.  a carries the arrays: the segments of the k
// list, part [SumsChunks(nseg) x K x 2], partials [ceil(K / kBlock)],
// counter (zero, re-armed by the last block) and result [4] (result[2]
// receives the energy).
hipError_t LaunchRecipSums(const SumsArgs &a, hipStream_t s) {
  const size_t tabBytes = size_t(a.tabAtoms) * SumsTabEntries(a.nmax) * sizeof(double2);
  hipLaunchKernelGGL(RecipSumsPartKernel, dim3(kSumsGrid), dim3(kBlock), tabBytes, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int nblocks = std::max(1, (a.K + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(RecipSumsKernel, dim3(nblocks), dim3(kBlock), 0, s, a);
  return hipGetLastError();
}

int BuildKSegments(gomc_box *b) {
  if (b->K + 1 > b->ksegCap) {
    if (b->kseg) (void)hipFree(b->kseg);
    b->kseg = nullptr;
    b->ksegCap = 0;
    HIP_TRY(b, hipMalloc(&b->kseg, sizeof(int) * (b->K + 1)));
    b->ksegCap = b->K + 1;
  }
  if (!b->nsegDev) HIP_TRY(b, hipMalloc(&b->nsegDev, sizeof(int)));
  hipLaunchKernelGGL(KSegmentsKernel, dim3(1), dim3(kSegThreads), 0, b->stream, b->kx, b->ky,
                     b->kz, b->K, b->kseg, b->nsegDev);
  HIP_TRY(b, hipGetLastError());
  HIP_TRY(b, hipMemcpyAsync(&b->nseg, b->nsegDev, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  return Sync(b);
}

// The box's arrays for the two structure-factor passes; the scratch
// (part, partials, counter) is filled in by the caller.
static SumsArgs BoxSumsArgs(const gomc_box *b) {
  SumsArgs a;
  a.x = b->x;
  a.y = b->y;
  a.z = b->z;
  a.q = b->q;
  a.mol = b->mol;
  a.natoms = b->natoms;
  a.K = b->K;
  a.kx = b->kx;
  a.ky = b->ky;
  a.kz = b->kz;
  a.pref = b->pref;
  a.kseg = b->kseg;
  a.nseg = b->nsegDev;
  a.cx = (1.0 / b->geom.L[0]) * (2.0 * M_PI);  // KVectorsOrth's spacings and ranges
  a.cy = (1.0 / b->geom.L[1]) * (2.0 * M_PI);
  a.cz = (1.0 / b->geom.L[2]) * (2.0 * M_PI);
  for (int i = 0; i < 3; ++i) a.nmax[i] = int(b->recipRcut * b->geom.L[i] / (2.0 * M_PI)) + 1;
  a.tabAtoms = std::max(1, std::min(kSumsTabMaxAtoms, kSumsTabBytes / int(SumsTabEntries(a.nmax) *
                                                                           sizeof(double2))));
  a.sumR = b->sumRnew;
  a.sumI = b->sumInew;
  a.result = b->devRes;
  return a;
}

// BoxReciprocalSums into the "new" sums; returns the reciprocal energy.
int RecipSums(gomc_box *b, double *erecip) {
  if (!b->recipReady) return Fail(b, GOMC_ERR_STATE, "k-vectors not initialised");
  SumsArgs a = BoxSumsArgs(b);
  const int nblocks = std::max(1, (b->K + kBlock - 1) / kBlock);
  int rc = EnsurePartials(b, nblocks);
  if (rc) return rc;
  a.partials = b->partials;
  a.counter = b->counter;
  const size_t need = size_t(SumsChunks(b->nseg, b->natoms, b->K)) * size_t(b->K) * 2;
  if (need > b->sumPartCap) {
    if (b->sumPart) (void)hipFree(b->sumPart);
    b->sumPart = nullptr;
    b->sumPartCap = 0;
    HIP_TRY(b, hipMalloc(&b->sumPart, sizeof(double) * need));
    b->sumPartCap = need;
  }
  a.part = b->sumPart;
  HIP_TRY(b, LaunchRecipSums(a, b->stream));
  rc = Sync(b);
  if (rc) return rc;
  *erecip = b->hostRes[2];
  return GOMC_OK;
}

// ------------------------------------------- allocation-free variant --
// gomc_recip_sums_noalloc runs the same two passes as RecipSums with its
// scratch -- the block counter, the second pass's per-block partials and
// the first pass's per-chunk sums -- in `scratch`, one file-scope device
// array that each GPU holds until the process ends.  Layout: counter at
// byte 0, partials at byte 256 (kNoAllocMaxK / kBlock entries), per-chunk
// sums after them (SumsChunks keeps chunks x K within kSumsPartCap, so
// 2 kSumsPartCap doubles hold them for K up to kNoAllocMaxK).  Every call on
// one GPU shares the array: calls that overlap on one GPU race.
constexpr int kNoAllocMaxK = 1 << 20;
constexpr size_t kNoAllocPartialsOff = 256;
constexpr size_t kNoAllocPartOff =
    kNoAllocPartialsOff + sizeof(Partial4) * size_t(kNoAllocMaxK / kBlock);
constexpr size_t kNoAllocPartDoubles = 2 * kSumsPartCap;
constexpr size_t kScratchBytes = kNoAllocPartOff + sizeof(double) * kNoAllocPartDoubles;
static_assert(kNoAllocPartOff % 256 == 0, "per-chunk sums must stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kScratchDevices = 64;
static std::mutex gScratchMu;
static unsigned char *gScratch[kScratchDevices] = {};

// Device address of this GPU's `scratch`, looked up once per device with
// that device current (the array's host-side name is not a device address).
static int ScratchBase(gomc_box *b, unsigned char **base) {
  if (b->device < 0 || b->device >= kScratchDevices)
    return Fail(b, GOMC_ERR_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(gScratchMu);
  if (!gScratch[b->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return Fail(b, GOMC_ERR_HIP,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    gScratch[b->device] = static_cast<unsigned char *>(p);
  }
  *base = gScratch[b->device];
  return GOMC_OK;
}

// The block counter must read zero when the second pass starts; a fill
// kernel, not a memset, so that a captured stream replays it.
__global__ void ZeroCounterKernel(unsigned *counter) { *counter = 0u; }

int RecipSumsNoAlloc(gomc_box *b, double *erecip) {
  if (!b->recipReady) return Fail(b, GOMC_ERR_STATE, "k-vectors not initialised");
  if (b->K > kNoAllocMaxK)
    return Fail(b, GOMC_ERR_ARG,
                "gomc_recip_sums_noalloc takes at most 1048576 k-vectors (use gomc_recip_sums)");
  const size_t need = size_t(SumsChunks(b->nseg, b->natoms, b->K)) * size_t(b->K) * 2;
  if (need > kNoAllocPartDoubles)
    return Fail(b, GOMC_ERR_ARG, "structure-factor chunks exceed the scratch array");
  unsigned char *base = nullptr;
  int rc = ScratchBase(b, &base);
  if (rc) return rc;
  SumsArgs a = BoxSumsArgs(b);
  a.counter = reinterpret_cast<unsigned *>(base);
  a.partials = reinterpret_cast<Partial4 *>(base + kNoAllocPartialsOff);
  a.part = reinterpret_cast<double *>(base + kNoAllocPartOff);
  hipLaunchKernelGGL(ZeroCounterKernel, dim3(1), dim3(1), 0, b->stream, a.counter);
  HIP_TRY(b, hipGetLastError());
  HIP_TRY(b, LaunchRecipSums(a, b->stream));
  rc = Sync(b);
  if (rc) return rc;
  *erecip = b->hostRes[2];
  return GOMC_OK;
}

static double BoxHostMinImage(double d, int axis, const gomc_box *b) {
  const double ax = b->geom.L[axis], half = b->geom.half[axis];
  if (d > half)
    d -= ax;
  else if (d < -half)
    d += ax;
  return d;
}

// Ewald::BoxSelf, MolCorrection summed over molecules, and
// CalculateEnergy::EnergyCorrection, from the host copy (O(molecules)).
void HostCorrectionSelfLRC(gomc_box *b, double &self, double &corr, double &lrc) {
  self = 0.0;
  for (int k = 0; k < b->nkinds; ++k) {
    double molSelf = 0.0;
    for (double c : b->kindCharge[k]) molSelf += (c * c);
    self += (molSelf * unsigned(b->kindCount[k]));
  }
  self *= -1.0 * b->alpha * kQQFact * M_2_SQRTPI * 0.5;
  corr = 0.0;
  for (int m = 0; m < b->nmols; ++m) {
    if (!b->molBox.empty() && b->molBox[m] != 0) continue;
    const int s = b->molStart[m], len = b->molLen[m], kind = b->molKind[m];
    double c = 0.0;
    for (int i = 0; i < len; ++i) {
      if (std::fabs(b->kindCharge[kind][i]) < 0.000000001) continue;
      for (int j = i + 1; j < len; ++j) {
        double dx = BoxHostMinImage(b->hx[s + i] - b->hx[s + j], 0, b);
        double dy = BoxHostMinImage(b->hy[s + i] - b->hy[s + j], 1, b);
        double dz = BoxHostMinImage(b->hz[s + i] - b->hz[s + j], 2, b);
        double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
        c += (b->kindCharge[kind][i] * b->kindCharge[kind][j] * std::erf(b->alpha * dist) / dist);
      }
    }
    corr += -1.0 * kQQFact * c * 1.0 * 1.0;
  }
  lrc = 0.0;
  for (int i = 0; i < b->nkinds; ++i)
    for (int j = 0; j < b->nkinds; ++j)
      lrc += b->pairEnCorr[i * b->nkinds + j] * unsigned(b->kindCount[i]) *
             unsigned(b->kindCount[j]) * b->volInv;
}

}  // namespace gomc

// ====================================================================== C ABI
extern "C" {

const char *gomc_last_error(const gomc_box *b) { return b ? b->err.c_str() : "null box"; }

static void FreeAll(gomc_box *b) {
  void *ptrs[] = {b->x,        b->y,       b->z,        b->q,        b->type,   b->mol,   b->molStartDev,
                  b->tSigSq,   b->tEpsCn,  b->tNHalf,   b->tEpsCn6,  b->tNOver6, b->cellCount,
                  b->cellCursor, b->cellStart, b->cellVec, b->atomCell, b->nbr,   b->kx,
                  b->ky,       b->kz,      b->pref,     b->sumRref,  b->sumIref, b->sumRnew,
                  b->sumInew,  b->partials, b->counter, b->kseg, b->nsegDev};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  MPFree(b);
  DevFree(b);
  if (b->trialBuf) (void)hipFree(b->trialBuf);
  if (b->virBuf) (void)hipFree(b->virBuf);
  if (b->sumPart) (void)hipFree(b->sumPart);
  b->sumPart = nullptr;
  for (void *p : {(void *)b->tcCount, (void *)b->tcSlots})
    if (p) (void)hipFree(p);
  b->sumPartCap = 0;
  for (void *p : {(void *)b->cacheC, (void *)b->cacheS, (void *)b->stageC, (void *)b->stageS})
    if (p) (void)hipFree(p);
  if (b->hostRes) (void)hipHostFree(b->hostRes);
  if (b->stream) (void)hipStreamDestroy(b->stream);
}

int gomc_box_destroy(gomc_box *b) {
  if (!b) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  FreeAll(b);
  delete b;
  return GOMC_OK;
}

int gomc_box_info(const gomc_box *b, int64_t out[7]) {
  if (!b) return GOMC_ERR_ARG;
  out[0] = b->natoms;
  out[1] = b->nmols;
  out[2] = b->K;
  out[3] = b->ncell;
  out[4] = b->edge[0];
  out[5] = b->edge[1];
  out[6] = b->edge[2];
  return GOMC_OK;
}

int gomc_box_dims(const gomc_box *b, double out[4]) {
  if (!b || !out) return GOMC_ERR_ARG;
  for (int i = 0; i < 3; ++i) out[i] = b->geom.L[i];
  out[3] = b->volume;
  return GOMC_OK;
}

int gomc_set_coords(gomc_box *b, const double *x, const double *y, const double *z) {
  if (!b || !x || !y || !z) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  std::copy(x, x + b->natoms, b->hx.begin());
  std::copy(y, y + b->natoms, b->hy.begin());
  std::copy(z, z + b->natoms, b->hz.begin());
  const size_t bytes = sizeof(double) * b->natoms;
  HIP_TRY(b, hipMemcpyAsync(b->x, x, bytes, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->y, y, bytes, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->z, z, bytes, hipMemcpyHostToDevice, b->stream));
  b->pendingMol = -1;
  TrialCellsInvalidate(b);
  int rc = Sync(b);
  if (!rc && b->cached) rc = CacheBuild(b);
  return rc;
}

int gomc_get_coords(const gomc_box *b, double *x, double *y, double *z) {
  if (!b) return GOMC_ERR_ARG;
  std::copy(b->hx.begin(), b->hx.end(), x);
  std::copy(b->hy.begin(), b->hy.end(), y);
  std::copy(b->hz.begin(), b->hz.end(), z);
  return GOMC_OK;
}

int gomc_commit_mol(gomc_box *b, int m, const double *x, const double *y, const double *z) {
  if (!b || m < 0 || m >= b->nmols) return GOMC_ERR_ARG;
  const int s = b->molStart[m], len = b->molLen[m];
  if (len > kMaxMolAtoms) return Fail(b, GOMC_ERR_ARG, "molecule has more than 16 atoms");
  if (b->pendingMol >= 0 && b->pendingMol != m) {
    // flush the earlier pending molecule first
    (void)hipSetDevice(b->device);
    const int ps = b->molStart[b->pendingMol], pl = b->molLen[b->pendingMol];
    HIP_TRY(b, hipMemcpyAsync(b->x + ps, b->pend[0], sizeof(double) * pl, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->y + ps, b->pend[1], sizeof(double) * pl, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->z + ps, b->pend[2], sizeof(double) * pl, hipMemcpyHostToDevice, b->stream));
    int rc = TrialCellsFlushEdits(b);
    if (!rc) rc = Sync(b);
    if (rc) return rc;
  }
  TrialCellsCommit(b, m, x, y, z);
  for (int p = 0; p < len; ++p) {
    b->hx[s + p] = b->pend[0][p] = x[p];
    b->hy[s + p] = b->pend[1][p] = y[p];
    b->hz[s + p] = b->pend[2][p] = z[p];
  }
  b->pendingMol = m;  // written by the next launch that reads coordinates
  return GOMC_OK;
}

}  // extern "C"

namespace gomc {
// Writes any pending accepted molecule to HBM before a kernel that reads
// all coordinates without the trial's override.
int FlushPending(gomc_box *b) {
  if (b->pendingMol < 0) return GOMC_OK;
  const int ps = b->molStart[b->pendingMol], pl = b->molLen[b->pendingMol];
  HIP_TRY(b, hipMemcpyAsync(b->x + ps, b->pend[0], sizeof(double) * pl, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->y + ps, b->pend[1], sizeof(double) * pl, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->z + ps, b->pend[2], sizeof(double) * pl, hipMemcpyHostToDevice, b->stream));
  int rc = TrialCellsFlushEdits(b);
  if (!rc) rc = Sync(b);
  if (rc) return rc;
  b->pendingMol = -1;
  return GOMC_OK;
}

// Energies of `n` trial positions of one atom against the live atoms
// (CalculateEnergy::ParticleInter, src/CalculateEnergy.cpp:727-785), through
// the trial cell table when 27 cells of it hold fewer slots than the box
// has atoms; the trial staging buffer is the box's and grows only when `n`
// does.
int ParticleInterImpl(gomc_box *b, int atomType, double charge, int skipMol, int n,
                      const double *tx, const double *ty, const double *tz, double *en,
                      double *real, int32_t *overlap) {
  if (n > b->trialCap) {
    if (b->trialBuf) (void)hipFree(b->trialBuf);
    b->trialBuf = nullptr;
    b->trialCap = 0;
    HIP_TRY(b, hipMalloc(&b->trialBuf, sizeof(double) * 6 * size_t(n)));
    b->trialCap = n;
  }
  double *buf = b->trialBuf;
  const size_t nb = sizeof(double) * n;
  HIP_TRY(b, hipMemcpyAsync(buf, tx, nb, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(buf + n, ty, nb, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(buf + 2 * n, tz, nb, hipMemcpyHostToDevice, b->stream));
  ParticleArgs a;
  a.g = b->geom;
  a.x = b->x;
  a.y = b->y;
  a.z = b->z;
  a.q = b->q;
  a.type = b->type;
  a.mol = b->mol;
  a.sigSq = b->tSigSq;
  a.epsCn = b->tEpsCn;
  a.nHalf = b->tNHalf;
  a.ntypes = b->ntypes;
  a.natoms = b->natoms;
  a.atype = atomType;
  a.skipMol = skipMol;
  a.charge = charge;
  a.tx = buf;
  a.ty = buf + n;
  a.tz = buf + 2 * n;
  a.en = buf + 3 * n;
  a.real = buf + 4 * n;
  a.overlap = buf + 5 * n;
  a.tcCap = 0;
  if (TrialCellsUseful(b)) {
    int rc = TrialCellsEnsure(b);
    if (!rc) rc = TrialCellsFlushEdits(b);
    if (rc) return rc;
    bool inBox = true;
    for (int t = 0; t < n; ++t) inBox = inBox && InBox(b, tx[t], ty[t], tz[t]);
    if (inBox && size_t(27) * b->tcCap < size_t(b->natoms)) {
      a.tcCap = b->tcCap;
      a.tcCount = b->tcCount;
      a.tcSlots = b->tcSlots;
      a.nbr = b->nbr;
      for (int i = 0; i < 3; ++i) {
        a.cs[i] = b->cellSize[i];
        a.edge[i] = b->edge[i];
      }
    }
  }
  if (b->precision == 1)
    hipLaunchKernelGGL(ParticleInterKernel<1>, dim3(n), dim3(kBlock), 0, b->stream, a);
  else
    hipLaunchKernelGGL(ParticleInterKernel<0>, dim3(n), dim3(kBlock), 0, b->stream, a);
  HIP_TRY(b, hipGetLastError());
  std::vector<double> host(3 * size_t(n));
  HIP_TRY(b, hipMemcpyAsync(host.data(), buf + 3 * n, 3 * nb, hipMemcpyDeviceToHost, b->stream));
  int rc = Sync(b);
  if (rc) return rc;
  for (int t = 0; t < n; ++t) {
    en[t] = host[t];
    real[t] = host[n + t];
    overlap[t] = host[2 * n + t] > 0.0 ? 1 : 0;
  }
  return GOMC_OK;
}

// Ewald::SwapDestRecip / SwapSourceRecip (src/Ewald.cpp:478-531,657-710)
// and, with the molecule cache, EwaldCached::SwapDestRecip's cache
// staging (src/EwaldCached.cpp:300-352).
int SwapRecipImpl(gomc_box *b, int kind, int insert, const double *x, const double *y,
                  const double *z, int mol, double *enew) {
  if (!b->recipReady) return Fail(b, GOMC_ERR_STATE, "k-vectors not initialised");
  const int n = int(b->kindType[kind].size());
  if (n > kMaxMolAtoms) return Fail(b, GOMC_ERR_ARG, "molecule has more than 16 atoms");
  SwapArgs a;
  std::memset(&a, 0, sizeof(a));
  a.K = b->K;
  a.n = n;
  a.sign = insert ? 1.0 : -1.0;
  for (int p = 0; p < n; ++p) {
    a.px[p] = x[p];
    a.py[p] = y[p];
    a.pz[p] = z[p];
    a.pq[p] = b->kindCharge[kind][p];
  }
  a.kx = b->kx;
  a.ky = b->ky;
  a.kz = b->kz;
  a.pref = b->pref;
  a.sumRref = b->sumRref;
  a.sumIref = b->sumIref;
  a.sumRnew = b->sumRnew;
  a.sumInew = b->sumInew;
  b->stageMol = -1;
  if (b->cached && insert && mol >= 0) {
    a.stC = b->stageC;
    a.stS = b->stageS;
    b->stageMol = mol;
  }
  const int nblocks = std::max(1, (b->K + kBlock - 1) / kBlock);
  int rc = EnsurePartials(b, nblocks);
  if (rc) return rc;
  a.partials = b->partials;
  a.counter = b->counter;
  a.result = b->devRes;
  if (b->precision == 1)
    hipLaunchKernelGGL(SwapRecipKernel<1>, dim3(nblocks), dim3(kBlock), 0, b->stream, a);
  else
    hipLaunchKernelGGL(SwapRecipKernel<0>, dim3(nblocks), dim3(kBlock), 0, b->stream, a);
  HIP_TRY(b, hipGetLastError());
  rc = Sync(b);
  if (rc) return rc;
  *enew = b->hostRes[2];
  return GOMC_OK;
}

// EwaldCached::BoxReciprocalSetup's per-molecule terms for the current
// coordinates.
int CacheBuild(gomc_box *b) {
  if (!b->cached) return GOMC_OK;
  CacheArgs a;
  a.x = b->x;
  a.y = b->y;
  a.z = b->z;
  a.q = b->q;
  a.mol = b->mol;
  a.molStart = b->molStartDev;
  a.nmols = b->nmols;
  a.K = b->K;
  a.kx = b->kx;
  a.ky = b->ky;
  a.kz = b->kz;
  a.cacheC = b->cacheC;
  a.cacheS = b->cacheS;
  const int g = std::max(1, (b->K + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(CacheBuildKernel, dim3(g), dim3(kBlock), 0, b->stream, a);
  HIP_TRY(b, hipGetLastError());
  b->stageMol = -1;
  return Sync(b);
}

// Moves molecule m's atoms in or out of the interacting set on the device.
int SetMolLive(gomc_box *b, int m, bool live) {
  TrialCellsInvalidate(b);
  const int s = b->molStart[m], len = b->molLen[m];
  std::vector<int> ids(len, live ? m : -1);
  HIP_TRY(b, hipMemcpyAsync(b->mol + s, ids.data(), sizeof(int) * len, hipMemcpyHostToDevice,
                            b->stream));
  return Sync(b);
}
}  // namespace gomc

extern "C" {

int gomc_set_mol_box(gomc_box *b, int mol, int box) {
  if (!b || mol < 0 || mol >= b->nmols || (box != 0 && box != 1)) return GOMC_ERR_ARG;
  if (box == 1 && b->ensemble != 1)
    return Fail(b, GOMC_ERR_STATE, "only a GCMC box has a reservoir");
  (void)hipSetDevice(b->device);
  int rc = FlushPending(b);
  if (rc) return rc;
  if (b->molBox[mol] == box) return GOMC_OK;
  DevFree(b);  // the device move loop keeps its own copy of the lookup
  b->kindCount[b->molKind[mol]] += box == 0 ? 1 : -1;
  b->molBox[mol] = box;
  LookupInit(b);
  return SetMolLive(b, mol, box == 0);
}

int gomc_build_cells(gomc_box *b, int32_t *cell_vector, int32_t *cell_start, int32_t *atom_cell) {
  if (!b) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = FlushPending(b);
  if (rc) return rc;
  rc = BuildCells(b);
  if (rc) return rc;
  if (cell_vector)
    HIP_TRY(b, hipMemcpyAsync(cell_vector, b->cellVec, sizeof(int) * b->natoms, hipMemcpyDeviceToHost, b->stream));
  if (cell_start)
    HIP_TRY(b, hipMemcpyAsync(cell_start, b->cellStart, sizeof(int) * (b->ncell + 1), hipMemcpyDeviceToHost, b->stream));
  if (atom_cell)
    HIP_TRY(b, hipMemcpyAsync(atom_cell, b->atomCell, sizeof(int) * b->natoms, hipMemcpyDeviceToHost, b->stream));
  return Sync(b);
}

int gomc_trial_cells(gomc_box *b, int32_t *cap, int32_t *count, int32_t *slots) {
  if (!b || !cap) return GOMC_ERR_ARG;
  if (!TrialCellsUseful(b))
    return Fail(b, GOMC_ERR_STATE, "a box of 27 cells has no trial cell table");
  (void)hipSetDevice(b->device);
  int rc = FlushPending(b);  // applies the queued edits
  if (!rc) rc = TrialCellsEnsure(b);
  if (rc) return rc;
  *cap = b->tcCap;
  if (count)
    HIP_TRY(b, hipMemcpyAsync(count, b->tcCount, sizeof(int) * b->ncell, hipMemcpyDeviceToHost,
                              b->stream));
  if (slots)
    HIP_TRY(b, hipMemcpyAsync(slots, b->tcSlots, sizeof(int) * size_t(b->ncell) * b->tcCap,
                              hipMemcpyDeviceToHost, b->stream));
  return Sync(b);
}

int gomc_box_inter(gomc_box *b, double *lj, double *real) {
  if (!b || !lj || !real) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = FlushPending(b);
  if (rc) return rc;
  return BoxInter(b, lj, real);
}

int gomc_mol_inter(gomc_box *b, int m, const double *nx, const double *ny, const double *nz,
                   double *dlj, double *dreal, int32_t *overlap) {
  if (!b || m < 0 || m >= b->nmols) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  const int s = b->molStart[m];
  double out[4];
  int rc = LaunchTrial(b, m, &b->hx[s], &b->hy[s], &b->hz[s], nx, ny, nz, true, false, out);
  if (rc) return rc;
  *dlj = out[0];
  *dreal = out[1];
  *overlap = out[3] > 0.0 ? 1 : 0;
  return GOMC_OK;
}

int gomc_mol_recip(gomc_box *b, int m, const double *nx, const double *ny, const double *nz,
                   double *enew) {
  if (!b || m < 0 || m >= b->nmols) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  const int s = b->molStart[m];
  double out[4];
  int rc = LaunchTrial(b, m, &b->hx[s], &b->hy[s], &b->hz[s], nx, ny, nz, false, true, out);
  if (rc) return rc;
  *enew = out[2];
  return GOMC_OK;
}

int gomc_mol_trial(gomc_box *b, int m, const double *nx, const double *ny, const double *nz,
                   double out[4]) {
  if (!b || m < 0 || m >= b->nmols) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  const int s = b->molStart[m];
  return LaunchTrial(b, m, &b->hx[s], &b->hy[s], &b->hz[s], nx, ny, nz, true, true, out);
}

int gomc_swap_recip(gomc_box *b, int kind, int mol, int insert, const double *x, const double *y,
                    const double *z, double *enew) {
  if (!b || kind < 0 || kind >= b->nkinds || mol >= b->nmols || !x || !y || !z || !enew)
    return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  return SwapRecipImpl(b, kind, insert, x, y, z, mol, enew);
}

int gomc_set_precision(gomc_box *b, int precision) {
  if (!b || precision < 0 || precision > 1) return GOMC_ERR_ARG;
  b->precision = precision;
  return GOMC_OK;
}

int gomc_recip_cache(gomc_box *b, int enable) {
  if (!b) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  if (!enable) {
    void *ptrs[] = {b->cacheC, b->cacheS, b->stageC, b->stageS};
    for (void *p : ptrs)
      if (p) (void)hipFree(p);
    b->cacheC = b->cacheS = b->stageC = b->stageS = nullptr;
    b->cached = false;
    b->stageMol = -1;
    return GOMC_OK;
  }
  if (!b->ewald || !b->recipReady) return Fail(b, GOMC_ERR_STATE, "k-vectors not initialised");
  if (!b->cached) {
    const size_t row = sizeof(double) * size_t(std::max(b->K, 1));
    HIP_TRY(b, hipMalloc(&b->cacheC, row * b->nmols));
    HIP_TRY(b, hipMalloc(&b->cacheS, row * b->nmols));
    HIP_TRY(b, hipMalloc(&b->stageC, row));
    HIP_TRY(b, hipMalloc(&b->stageS, row));
    b->cached = true;
  }
  int rc = FlushPending(b);
  if (rc) return rc;
  return CacheBuild(b);
}

int gomc_particle_inter(gomc_box *b, int atom_type, double charge, int skip_mol, int ntrials,
                        const double *tx, const double *ty, const double *tz, double *en,
                        double *real, int32_t *overlap) {
  if (!b || ntrials <= 0 || atom_type < 0 || atom_type >= b->ntypes || !tx || !ty || !tz ||
      !en || !real || !overlap)
    return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = FlushPending(b);
  if (rc) return rc;
  return ParticleInterImpl(b, atom_type, charge, skip_mol, ntrials, tx, ty, tz, en, real,
                           overlap);
}

int gomc_recip_sums(gomc_box *b, double *erecip) {
  if (!b || !erecip) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = FlushPending(b);
  if (rc) return rc;
  return RecipSums(b, erecip);
}

int gomc_recip_sums_noalloc(gomc_box *b, double *erecip) {
  if (!b || !erecip) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = FlushPending(b);
  if (rc) return rc;
  return RecipSumsNoAlloc(b, erecip);
}

int gomc_recip_accept(gomc_box *b) {
  if (!b) return GOMC_ERR_ARG;
  std::swap(b->sumRnew, b->sumRref);
  std::swap(b->sumInew, b->sumIref);
  if (b->cached && b->stageMol >= 0) {  // the trial molecule's terms become its cache
    (void)hipSetDevice(b->device);
    const size_t row = sizeof(double) * b->K;
    HIP_TRY(b, hipMemcpyAsync(b->cacheC + int64_t(b->stageMol) * b->K, b->stageC, row,
                              hipMemcpyDeviceToDevice, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->cacheS + int64_t(b->stageMol) * b->K, b->stageS, row,
                              hipMemcpyDeviceToDevice, b->stream));
    b->stageMol = -1;
  }
  return GOMC_OK;
}

int gomc_recip_copy_ref_to_new(gomc_box *b) {
  if (!b) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  HIP_TRY(b, hipMemcpyAsync(b->sumRnew, b->sumRref, sizeof(double) * b->K, hipMemcpyDeviceToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->sumInew, b->sumIref, sizeof(double) * b->K, hipMemcpyDeviceToDevice, b->stream));
  return Sync(b);
}

int gomc_recip_get(const gomc_box *b_, double *kx, double *ky, double *kz, double *prefact,
                   double *sum_r, double *sum_i) {
  gomc_box *b = const_cast<gomc_box *>(b_);
  if (!b) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  const size_t n = sizeof(double) * b->K;
  if (kx) HIP_TRY(b, hipMemcpyAsync(kx, b->kx, n, hipMemcpyDeviceToHost, b->stream));
  if (ky) HIP_TRY(b, hipMemcpyAsync(ky, b->ky, n, hipMemcpyDeviceToHost, b->stream));
  if (kz) HIP_TRY(b, hipMemcpyAsync(kz, b->kz, n, hipMemcpyDeviceToHost, b->stream));
  if (prefact) HIP_TRY(b, hipMemcpyAsync(prefact, b->pref, n, hipMemcpyDeviceToHost, b->stream));
  if (sum_r) HIP_TRY(b, hipMemcpyAsync(sum_r, b->sumRref, n, hipMemcpyDeviceToHost, b->stream));
  if (sum_i) HIP_TRY(b, hipMemcpyAsync(sum_i, b->sumIref, n, hipMemcpyDeviceToHost, b->stream));
  return Sync(b);
}

}  // extern "C"

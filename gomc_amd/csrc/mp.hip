// MultiParticle move on the plain path (src/moves/MultiParticle.h): the
// pair forces of CalculateEnergy::BoxForce, the reciprocal forces of
// Ewald::BoxForceReciprocal, molecule forces and torques
// (CalculateEnergy::CalculateTorque), the force-biased transform of every
// molecule (MultiParticle::CalculateTrialDistRot) and the acceptance
// coefficient (MultiParticle::GetCoeff), all on device-resident state.
//
// Forces are summed per atom over the full neighbourhood (each pair is
// evaluated from both ends) instead of scattered with atomics, so every
// sum has a fixed order and a run is reproducible bit for bit.
#include <algorithm>
#include <cmath>

#include "box.hpp"

using namespace gomc;

namespace {

constexpr int kChunk = 64;  // home atoms per force workgroup
constexpr double kMinForce = 1E-12, kMaxForce = 30;  // MultiParticle.h:18-19
constexpr double kLambda = 0.5;                      // MultiParticle.h:95

struct Vec3D {
  double x, y, z;
};

// r123::uneg11<double> for a 64-bit word (lib/Random123/uniform.hpp:196-214).
__device__ __forceinline__ double Uneg11(uint64_t v) {
  return double(int64_t(v)) * 0x1p-63 + 0x1p-64;
}

// ------------------------------------------------------------ pair force --
// CalculateEnergy::BoxForce (CalculateEnergy.cpp:268-406).  A chunk is a
// cell's atoms 64 at a time (one per lane); its 27 neighbour cells, taken
// 64 atoms at a time, make a list of tiles.  `slices` workgroups of G waves
// share a chunk and its waves take the tiles round robin: wave w of slice s
// takes tiles s G + w, s G + w + slices G, ...  A wave stages each tile
// (position, charge, type, molecule of 64 neighbour atoms) in its LDS, one
// coalesced load per tile instead of a chain of dependent loads per pair,
// and marks each lane's in-cutoff partners of the tile in a bit mask.  It
// then lists the tile's pairs in (home atom, tile entry) order and
// evaluates them 64 at a time, so the pair math runs about (pairs / 64)
// times per tile rather than as often as the busiest lane has partners;
// the force contributions of a batch are summed per home atom by a
// segmented scan and added to that atom's sum in the wave's LDS, an order
// fixed by the positions alone.  A workgroup adds its waves' per-atom
// forces in wave order through LDS; with one slice it writes the chunk's
// forces, with several each slice writes its own and
// PairForceSumKernel adds them in slice order.  Energies count each pair
// once (i < j).  The host picks G and the slice count so that a box of any
// size puts a few thousand waves on the device.
constexpr int kPairChunksFor4 = 1024;  // from this many chunks: 4 waves per workgroup
constexpr int kPairChunksFor8 = 256;   // from this many: 8; below: 8 and slices
constexpr int kPairSmallWaves = 2048;  // waves a small box is split into, about
constexpr int kPairMaxSlices = 8;

struct PairForceArgs {
  BoxGeom g;
  const double *x, *y, *z, *q;
  const int *type, *mol;
  const double *sigSq, *epsCn, *nHalf, *epsCn6, *nOver6;
  int ntypes;
  const int *cellStart, *cellVec, *nbr;
  const int *chunks;  // [3 * nchunks]: cell, first atom within the cell, count
  int slices, natoms;
  double *fx, *fy, *fz;
  double *spart;  // slices > 1: [slice][xyz][natoms]
  Partial4 *partials;
  unsigned *counter;
  double *result;
};

struct PairTile {
  double x[kChunk], y[kChunk], z[kChunk], q[kChunk];
  int type[kChunk], mol[kChunk], idx[kChunk];
};

// In-cutoff pairs of one tile listed per wave, evaluated 64 at a time.
constexpr int kPairList = 512;  // list entries per wave and window

// Sum of v over the lanes of each run of equal, non-decreasing keys, in a
// fixed tree order (inclusive scan; the run's last lane holds the total).
__device__ __forceinline__ double SegScan(double v, int key, int lane) {
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const double u = __shfl_up(v, d);
    const int ku = __shfl_up(key, d);
    if (lane >= d && ku == key) v += u;
  }
  return v;
}

template <int kGroups>
__global__ void __launch_bounds__(kGroups * kChunk) PairForceKernel(const PairForceArgs a) {
  static_assert(kChunk == kWave, "one tile entry per lane");
  __shared__ double sf[kGroups][3][kChunk];
  __shared__ double red[kGroups * kChunk / kWave];
  __shared__ PairTile tile[kGroups];
  __shared__ PairTile home;
  __shared__ uint16_t plist[kGroups][kPairList];
  __shared__ double poly[kErfcxDeg + 1];
  __shared__ int nbS[27], nbE[27], tileS[28];
  const int chunkId = blockIdx.x / a.slices, slice = blockIdx.x - chunkId * a.slices;
  const int cell = a.chunks[3 * chunkId], off = a.chunks[3 * chunkId + 1],
            cnt = a.chunks[3 * chunkId + 2];
  const int lane = threadIdx.x % kChunk, grp = threadIdx.x / kChunk;
  const bool valid = lane < cnt;
  const int i = valid ? a.cellVec[a.cellStart[cell] + off + lane] : 0;
  if (threadIdx.x <= kErfcxDeg) poly[threadIdx.x] = kErfcxPoly[threadIdx.x];
  if (threadIdx.x < 27) {
    const int nb = a.nbr[cell * 27 + threadIdx.x];
    nbS[threadIdx.x] = a.cellStart[nb];
    nbE[threadIdx.x] = a.cellStart[nb + 1];
  }
  if (grp == 0) {
    home.x[lane] = a.x[i];
    home.y[lane] = a.y[i];
    home.z[lane] = a.z[i];
    home.q[lane] = a.q[i];
    home.type[lane] = a.type[i];
    home.idx[lane] = i;
  }
  sf[grp][0][lane] = 0.0;
  sf[grp][1][lane] = 0.0;
  sf[grp][2][lane] = 0.0;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int c = 0; c < 27; ++c) {
      tileS[c] = t;
      t += (nbE[c] - nbS[c] + kChunk - 1) / kChunk;
    }
    tileS[27] = t;
  }
  __syncthreads();
  const double xi = home.x[lane], yi = home.y[lane], zi = home.z[lane];
  const int mi = valid ? a.mol[i] : -1;
  double lj = 0.0, real = 0.0;
  PairTile &t = tile[grp];
  uint16_t *list = plist[grp];
  const int ntiles = tileS[27], stride = a.slices * kGroups;
  int c = 0;
  for (int tt = slice * kGroups + grp; tt < ntiles; tt += stride) {
    while (tileS[c + 1] <= tt) ++c;
    const int base = nbS[c] + (tt - tileS[c]) * kChunk;
    const int n = min(kChunk, nbE[c] - base);
    if (lane < n) {
      const int j = a.cellVec[base + lane];
      t.idx[lane] = j;
      t.x[lane] = a.x[j];
      t.y[lane] = a.y[j];
      t.z[lane] = a.z[j];
      t.q[lane] = a.q[j];
      t.type[lane] = a.type[j];
      t.mol[lane] = a.mol[j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint64_t near = 0;
    if (valid)
      for (int k = 0; k < n; ++k) {
        const double dx = MinImage(xi - t.x[k], a.g.L[0], a.g.half[0]);
        const double dy = MinImage(yi - t.y[k], a.g.L[1], a.g.half[1]);
        const double dz = MinImage(zi - t.z[k], a.g.L[2], a.g.half[2]);
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (t.mol[k] != mi && a.g.rCutSq > d2) near |= uint64_t(1) << k;
      }
    // the wave's pairs in (lane, k) order: lane's pairs start at `pos`
    const int mine = __popcll(near);
    int incl = mine;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int u = __shfl_up(incl, d);
      if (lane >= d) incl += u;
    }
    const int total = __shfl(incl, kWave - 1);
    int pos = incl - mine;
    for (int b = 0; b < total; b += kPairList) {
      while (near && pos < b + kPairList) {
        list[pos - b] = uint16_t((lane << 6) | __builtin_ctzll(near));
        near &= near - 1;
        ++pos;
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const int nb = min(kPairList, total - b);
      for (int r = 0; r < nb; r += kWave) {
        const bool on = r + lane < nb;
        const int e = on ? list[r + lane] : 0;
        const int owner = on ? e >> 6 : kWave, k = e & 63;
        double fx = 0.0, fy = 0.0, fz = 0.0;
        if (on) {
          const double dx = MinImage(home.x[owner] - t.x[k], a.g.L[0], a.g.half[0]);
          const double dy = MinImage(home.y[owner] - t.y[k], a.g.L[1], a.g.half[1]);
          const double dz = MinImage(home.z[owner] - t.z[k], a.g.L[2], a.g.half[2]);
          const double d2 = dx * dx + dy * dy + dz * dz;
          const int idx = home.type[owner] + t.type[k] * a.ntypes;
          const PairTerms p = PairMath(a.g, poly, d2, home.q[owner] * t.q[k] * kQQFact,
                                       a.sigSq[idx], a.epsCn[idx], a.nHalf[idx], a.epsCn6[idx],
                                       a.nOver6[idx], home.idx[owner] < t.idx[k]);
          real += p.eReal;
          lj += p.eLJ;
          fx = dx * p.vLJ + dx * p.vReal;
          fy = dy * p.vLJ + dy * p.vReal;
          fz = dz * p.vLJ + dz * p.vReal;
        }
        fx = SegScan(fx, owner, lane);
        fy = SegScan(fy, owner, lane);
        fz = SegScan(fz, owner, lane);
        const int nextOwner = __shfl_down(owner, 1);
        if (on && (lane == kWave - 1 || nextOwner != owner)) {
          sf[grp][0][owner] += fx;
          sf[grp][1][owner] += fy;
          sf[grp][2][owner] += fz;
        }
      }
      // the list is read before the next window overwrites it
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // every lane is done with the tile before the next one overwrites it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  Partial4 part{0, 0, 0, 0};
  part.lj = BlockSum(lj, red);  // synchronises the LDS writes above
  part.real = BlockSum(real, red);
  if (grp == 0 && valid) {
    double sx = sf[0][0][lane], sy = sf[0][1][lane], sz = sf[0][2][lane];
    for (int g = 1; g < kGroups; ++g) {
      sx += sf[g][0][lane];
      sy += sf[g][1][lane];
      sz += sf[g][2][lane];
    }
    if (a.slices == 1) {
      a.fx[i] = sx;
      a.fy[i] = sy;
      a.fz[i] = sz;
    } else {
      double *sp = a.spart + size_t(slice) * 3 * a.natoms;
      sp[i] = sx;
      sp[a.natoms + i] = sy;
      sp[2 * size_t(a.natoms) + i] = sz;
    }
  }
  FinishPartials(part, a.partials, a.counter, a.result);
}

// The slices' forces of every chunk atom, added in slice order.
__global__ void __launch_bounds__(kChunk) PairForceSumKernel(const PairForceArgs a) {
  const int cell = a.chunks[3 * blockIdx.x], off = a.chunks[3 * blockIdx.x + 1],
            cnt = a.chunks[3 * blockIdx.x + 2];
  if (int(threadIdx.x) >= cnt) return;
  const int i = a.cellVec[a.cellStart[cell] + off + threadIdx.x];
  const size_t n = a.natoms;
  double sx = a.spart[i], sy = a.spart[n + i], sz = a.spart[2 * n + i];
  for (int s = 1; s < a.slices; ++s) {
    const double *sp = a.spart + size_t(s) * 3 * n;
    sx += sp[i];
    sy += sp[n + i];
    sz += sp[2 * n + i];
  }
  a.fx[i] = sx;
  a.fy[i] = sy;
  a.fz[i] = sz;
}

// ------------------------------------------------------ reciprocal force --
// Ewald::BoxForceReciprocal (Ewald.cpp:1496-1596), one atom per thread:
// the k-space force F_j = sum_k 2 q_j prefact_k (sin ReS - cos ImS) k over
// the current structure factors, walked segment by segment (pieces of at
// most kSeg consecutive-nz k-vectors, built with the box's k-vectors): one
// sincos per run of consecutive nz (a segment continuing a run starts at
// the phase the previous segment rotated to; the nz step's phase is the
// atom's own), then one complex product and six FMAs per vector against
// pref S_R k and pref S_I k, staged per tile of segments in LDS; the 2 q
// factor is applied once at the end.  Then the intramolecular erf correction is subtracted and
// the chunks added in order.
constexpr int kForceSegTile = 32;  // segments per LDS tile (at most 512 k)

struct RecipForceArgs {
  BoxGeom g;
  const double *x, *y, *z, *q;
  const int *mol, *molStart;
  int natoms, K;
  const double *kx, *ky, *kz, *pref, *sumR, *sumI;
  const int *kseg;  // segment starts [nseg + 1]
  int nseg;
  double cz;  // the nz step of every segment, 2 pi / L_z
  double *fx, *fy, *fz;
  // segment chunks: RecipForcePartKernel writes chunk c's k-space force of
  // every atom (segments [c segChunk, (c + 1) segChunk)) to
  // part[c][xyz][natoms]; RecipForceKernel adds the chunks in order after
  // the intramolecular term.
  int segChunk, kChunks;
  double *part;
};

__global__ void __launch_bounds__(kBlock) RecipForcePartKernel(const RecipForceArgs a) {
  // per k of the tile: pref S_R k_d and pref S_I k_d (d = x, y, z); per
  // segment: its first k-vector
  __shared__ double2 sk[3][kForceSegTile * kSeg];  // (pref S_R k_d, pref S_I k_d), one 16-byte read
  __shared__ double sf[3][kForceSegTile];
  __shared__ int ss[kForceSegTile + 1];
  const int chunk = blockIdx.x % a.kChunks;
  const int p = blockIdx.x / a.kChunks * kBlock + threadIdx.x;
  const bool active = p < a.natoms;
  double xp = 0.0, yp = 0.0, zp = 0.0, qp = 0.0;
  if (active) {
    xp = a.x[p];
    yp = a.y[p];
    zp = a.z[p];
    qp = a.q[p];
  }
  const bool on = active && !(fabs(qp) < 0.000000001);
  double X = 0.0, Y = 0.0, Z = 0.0;
  // the atom's phase step e^{i c_z z}, the same for every segment
  double sn = 0.0, cn = 1.0;
  if (on) sincos(a.cz * zp, &sn, &cn);
  const int s0 = min(a.nseg, chunk * a.segChunk), s1 = min(a.nseg, s0 + a.segChunk);
  // a segment that continues the previous one's run (same nx, ny after a
  // full segment) starts at the phase the previous one rotated to
  double im = 0.0, re = 1.0, pkx = 0.0, pky = 0.0;
  int plen = 0;
  for (int t = s0; t < s1; t += kForceSegTile) {
    const int tn = min(kForceSegTile, s1 - t);
    const int kA = a.kseg[t], kn = a.kseg[t + tn] - kA;
    __syncthreads();
    for (int i = threadIdx.x; i <= tn; i += kBlock) {
      const int k = a.kseg[t + i];
      ss[i] = k - kA;
      if (i < tn) {
        sf[0][i] = a.kx[k];
        sf[1][i] = a.ky[k];
        sf[2][i] = a.kz[k];
      }
    }
    for (int i = threadIdx.x; i < kn; i += kBlock) {
      const int k = kA + i;
      const double pr = a.pref[k] * a.sumR[k], pi = a.pref[k] * a.sumI[k];
      sk[0][i] = make_double2(pr * a.kx[k], pi * a.kx[k]);
      sk[1][i] = make_double2(pr * a.ky[k], pi * a.ky[k]);
      sk[2][i] = make_double2(pr * a.kz[k], pi * a.kz[k]);
    }
    __syncthreads();
    if (on)
      for (int j = 0; j < tn; ++j) {
        const int k0 = ss[j], len = ss[j + 1] - k0;
        const double kx0 = sf[0][j], ky0 = sf[1][j];
        if (!(plen == kSeg && kx0 == pkx && ky0 == pky))
          sincos(xp * kx0 + yp * ky0 + zp * sf[2][j], &im, &re);
        pkx = kx0;
        pky = ky0;
        plen = len;
#pragma unroll
        for (int i = 0; i < kSeg; ++i) {
          if (i < len) {
            const int k = k0 + i;
            const double2 vx = sk[0][k], vy = sk[1][k], vz = sk[2][k];
            X = fma(im, vx.x, X);
            X = fma(-re, vx.y, X);
            Y = fma(im, vy.x, Y);
            Y = fma(-re, vy.y, Y);
            Z = fma(im, vz.x, Z);
            Z = fma(-re, vz.y, Z);
          }
          const double nr = fma(re, cn, -(im * sn));
          im = fma(re, sn, im * cn);
          re = nr;
        }
      }
  }
  if (active) {
    const double q2 = 2.0 * qp;
    a.part[(size_t(chunk) * 3 + 0) * a.natoms + p] = q2 * X;
    a.part[(size_t(chunk) * 3 + 1) * a.natoms + p] = q2 * Y;
    a.part[(size_t(chunk) * 3 + 2) * a.natoms + p] = q2 * Z;
  }
}

__global__ void __launch_bounds__(kBlock) RecipForceKernel(const RecipForceArgs a) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= a.natoms) return;
  double X = 0.0, Y = 0.0, Z = 0.0;
  const double xp = a.x[p], yp = a.y[p], zp = a.z[p], qp = a.q[p];
  if (!(fabs(qp) < 0.000000001)) {
    const int m = a.mol[p], s = a.molStart[m], e = a.molStart[m + 1];
    const double constValue = a.g.alpha * M_2_SQRTPI;
    const double alphaSq = a.g.alpha * a.g.alpha;
    for (int j = s; j < e; ++j) {
      if (j == p) continue;
      const double dx = MinImage(xp - a.x[j], a.g.L[0], a.g.half[0]);
      const double dy = MinImage(yp - a.y[j], a.g.L[1], a.g.half[1]);
      const double dz = MinImage(zp - a.z[j], a.g.L[2], a.g.half[2]);
      const double distSq = dx * dx + dy * dy + dz * dz;
      const double dist = sqrt(distSq);
      const double expConstValue = exp(-1.0 * alphaSq * distSq);
      const double qiqj = qp * a.q[j] * kQQFact;
      double intraForce = qiqj * 1.0 * 1.0 / distSq;
      intraForce *= ((erf(a.g.alpha * dist) / dist) - constValue * expConstValue);
      X -= intraForce * dx;
      Y -= intraForce * dy;
      Z -= intraForce * dz;
    }
  }
  for (int c = 0; c < a.kChunks; ++c) {
    X += a.part[(size_t(c) * 3 + 0) * a.natoms + p];
    Y += a.part[(size_t(c) * 3 + 1) * a.natoms + p];
    Z += a.part[(size_t(c) * 3 + 2) * a.natoms + p];
  }
  a.fx[p] = X;
  a.fy[p] = Y;
  a.fz[p] = Z;
}

// ------------------------------------------------ molecule terms, torque --
// One molecule per thread: the molecule sums of the atom forces (the
// reference accumulates them pair by pair, CalculateEnergy.cpp:373-378,
// and atom by atom, Ewald.cpp:1586) and the torque about the molecule's
// centre (CalculateEnergy.cpp:1384-1398).  `mode` bit 0: pair sums,
// bit 1: reciprocal sums, bit 2: torque.
struct MolTermsArgs {
  BoxGeom g;
  const double *x, *y, *z, *cx, *cy, *cz;
  const int *molStart;
  int nmols, mode;
  const double *aF[3], *aFR[3];
  double *mF[3], *mFR[3], *tq[3];
};

__global__ void __launch_bounds__(kBlock) MolTermsKernel(const MolTermsArgs a) {
  const int m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= a.nmols) return;
  const int s = a.molStart[m], e = a.molStart[m + 1];
  if (a.mode & 1)
    for (int d = 0; d < 3; ++d) {
      double v = 0.0;
      for (int p = s; p < e; ++p) v += a.aF[d][p];
      a.mF[d][m] = v;
    }
  if (a.mode & 2)
    for (int d = 0; d < 3; ++d) {
      double v = 0.0;
      for (int p = s; p < e; ++p) v += a.aFR[d][p];
      a.mFR[d][m] = v;
    }
  if (a.mode & 4) {
    const double c[3] = {a.cx[m], a.cy[m], a.cz[m]};
    double tx = 0.0, ty = 0.0, tz = 0.0;
    for (int p = s; p < e; ++p) {
      const double dx = MinImage(a.x[p] - c[0], a.g.L[0], a.g.half[0]);
      const double dy = MinImage(a.y[p] - c[1], a.g.L[1], a.g.half[1]);
      const double dz = MinImage(a.z[p] - c[2], a.g.L[2], a.g.half[2]);
      const double fx = a.aF[0][p] + a.aFR[0][p], fy = a.aF[1][p] + a.aFR[1][p],
                   fz = a.aF[2][p] + a.aFR[2][p];
      tx += dy * fz - dz * fy;
      ty += dz * fx - dx * fz;
      tz += dx * fy - dy * fx;
    }
    a.tq[0][m] = tx;
    a.tq[1][m] = ty;
    a.tq[2][m] = tz;
  }
}

// ------------------------------------------------------------- transform --
// MultiParticle::CalculateTrialDistRot (MultiParticle.h:565-611) with
// CalcRandomTransform (546-563), Rotate/TranslateForceBiased (613-665) and
// Rotate/TranslateRandom (667-712): one molecule per thread, Philox4x64-10
// with key {step, seed} and counter {molecule, 0, 0, 0}
// (Random123Wrapper.cpp:16-22).  A displacement beyond half the box in any
// component, or a non-finite transform, sets `fail` (MultiParticle.h:366-406).
struct TransformArgs {
  BoxGeom g;
  int nmols, rotate;
  uint64_t step, seed;
  double beta, max;
  const double *x, *y, *z, *cx, *cy, *cz;
  double *xn, *yn, *zn, *ncx, *ncy, *ncz;
  const int *molStart;
  const double *mF[3], *mFR[3], *tq[3];
  double *tk[3];
  int *inRange, *fail;
};

// TransformMatrix::FromAxisAngle (src/TransformMatrix.h:186-214).
__device__ void AxisAngleDev(double theta, const double ax[3], double r[3][3]) {
  const double cross[3][3] = {{0.0, -ax[2], ax[1]}, {ax[2], 0.0, -ax[0]}, {-ax[1], ax[0], 0.0}};
  const double c = cos(theta), s = sin(theta);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double tensor = ax[i] * ax[j];
      r[i][j] = (i == j ? c : 0.0) + (s * cross[i][j] + (1 - c) * tensor);
    }
}

__global__ void __launch_bounds__(kBlock) TransformKernel(const TransformArgs a) {
  const int m = blockIdx.x * kBlock + threadIdx.x;
  if (m >= a.nmols) return;
  double lb[3], lbmax[3];
  for (int d = 0; d < 3; ++d) {
    lb[d] = a.rotate ? (a.tq[d][m] * kLambda) * a.beta : ((a.mF[d][m] + a.mFR[d][m]) * kLambda) * a.beta;
    lbmax[d] = lb[d] * a.max;
  }
  bool in = true;
  for (int d = 0; d < 3; ++d) in = in && fabs(lbmax[d]) > kMinForce && fabs(lbmax[d]) < kMaxForce;
  const U64x4 r = Philox(uint64_t(uint32_t(m)), 0, 0, 0, a.step, a.seed);
  double val[3] = {0.0, 0.0, 0.0};
  if (in)
    for (int d = 0; d < 3; ++d)
      val[d] = log(exp(-1.0 * lbmax[d]) + 2.0 * U01(r.v[d]) * sinh(lbmax[d])) / lb[d];
  for (int d = 0; d < 3; ++d) a.tk[d][m] = val[d];
  a.inRange[m] = in ? 1 : 0;
  bool bad = !isfinite(sqrt(val[0] * val[0] + val[1] * val[1] + val[2] * val[2]));
  if (!a.rotate && (val[0] > a.g.half[0] || val[1] > a.g.half[1] || val[2] > a.g.half[2])) bad = true;
  if (bad) atomicOr(a.fail, 1);

  const int s = a.molStart[m], e = a.molStart[m + 1];
  const double c[3] = {a.cx[m], a.cy[m], a.cz[m]};
  if (!a.rotate) {
    double sh[3];
    for (int d = 0; d < 3; ++d) sh[d] = in ? val[d] : a.max * Uneg11(r.v[d]);
    for (int p = s; p < e; ++p) {
      a.xn[p] = Wrap(a.x[p] + sh[0], a.g.L[0]);
      a.yn[p] = Wrap(a.y[p] + sh[1], a.g.L[1]);
      a.zn[p] = Wrap(a.z[p] + sh[2], a.g.L[2]);
    }
    a.ncx[m] = Wrap(c[0] + sh[0], a.g.L[0]);
    a.ncy[m] = Wrap(c[1] + sh[1], a.g.L[1]);
    a.ncz[m] = Wrap(c[2] + sh[2], a.g.L[2]);
    return;
  }
  double theta, ax[3];
  if (in) {
    theta = sqrt(val[0] * val[0] + val[1] * val[1] + val[2] * val[2]);
    const double inv = 1.0 / theta;
    for (int d = 0; d < 3; ++d) ax[d] = val[d] * inv;
  } else {
    theta = a.max * Uneg11(r.v[0]);
    const double u = Uneg11(r.v[1]);
    const double th = 2.0 * M_PI * U01(r.v[2]);
    const double root = sqrt(1.0 - u * u);
    ax[0] = root * cos(th);
    ax[1] = root * sin(th);
    ax[2] = u;
  }
  double R[3][3];
  AxisAngleDev(theta, ax, R);
  for (int p = s; p < e; ++p) {
    double v[3] = {Unwrap(a.x[p], c[0], a.g.L[0], a.g.half[0]) + (-c[0]),
                   Unwrap(a.y[p], c[1], a.g.L[1], a.g.half[1]) + (-c[1]),
                   Unwrap(a.z[p], c[2], a.g.L[2], a.g.half[2]) + (-c[2])};
    double w[3];
    for (int i = 0; i < 3; ++i) w[i] = R[i][0] * v[0] + R[i][1] * v[1] + R[i][2] * v[2];
    a.xn[p] = Wrap(w[0] + c[0], a.g.L[0]);
    a.yn[p] = Wrap(w[1] + c[1], a.g.L[1]);
    a.zn[p] = Wrap(w[2] + c[2], a.g.L[2]);
  }
  a.ncx[m] = c[0];
  a.ncy[m] = c[1];
  a.ncz[m] = c[2];
}

// ----------------------------------------------------------- coefficient --
// MultiParticle::GetCoeff / CalculateWRatio (MultiParticle.h:445-511): the
// product over in-range molecules of the forward/backward weight ratios,
// a fixed tree within each workgroup, then workgroups in index order.
struct CoeffArgs {
  int nmols, rotate;
  double lBeta, max;
  const double *tk[3];
  const int *inRange;
  const double *nF[3], *nFR[3], *nT[3], *oF[3], *oFR[3], *oT[3];
  double *parts;
  unsigned *counter;
  double *result;
};

__device__ double WRatio(const double n[3], const double o[3], const double k[3], double max) {
  double w = 1.0;
  for (int d = 0; d < 3; ++d) {
    w *= n[d] * exp(-n[d] * k[d]) / (2.0 * sinh(n[d] * max));
    w /= o[d] * exp(o[d] * k[d]) / (2.0 * sinh(o[d] * max));
  }
  return w;
}

__global__ void __launch_bounds__(kBlock) CoeffKernel(const CoeffArgs a) {
  __shared__ double sw[kBlock];
  __shared__ bool last;
  const int m = blockIdx.x * kBlock + threadIdx.x;
  double w = 1.0;
  if (m < a.nmols && a.inRange[m]) {
    double n[3], o[3], k[3];
    for (int d = 0; d < 3; ++d) {
      k[d] = a.tk[d][m];
      if (a.rotate) {
        n[d] = a.nT[d][m] * a.lBeta;
        o[d] = a.oT[d][m] * a.lBeta;
      } else {
        n[d] = (a.nF[d][m] + a.nFR[d][m]) * a.lBeta;
        o[d] = (a.oF[d][m] + a.oFR[d][m]) * a.lBeta;
      }
    }
    w = WRatio(n, o, k, a.max);
  }
  sw[threadIdx.x] = w;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if (int(threadIdx.x) < s) sw[threadIdx.x] *= sw[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.parts[blockIdx.x] = sw[0];
    __threadfence();
    last = (atomicAdd(a.counter, 1u) == gridDim.x - 1);
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    __threadfence();
    volatile double *vp = a.parts;
    double p = 1.0;
    for (unsigned i = 0; i < gridDim.x; ++i) p *= vp[i];
    *a.result = p;
    *a.counter = 0u;
    __threadfence_system();
  }
}

}  // namespace

namespace gomc {

void MPFree(gomc_box *b) {
  MPState &mp = b->mp;
  std::vector<void *> ptrs = {mp.xn, mp.yn, mp.zn, mp.inRange, mp.fail, mp.chunks, mp.wpart,
                              mp.wcounter, mp.fpart, mp.spart};
  for (int s = 0; s < 2; ++s)
    for (int d = 0; d < 3; ++d) {
      ptrs.push_back(mp.com[s][d]);
      ptrs.push_back(mp.aF[s][d]);
      ptrs.push_back(mp.aFR[s][d]);
      ptrs.push_back(mp.mF[s][d]);
      ptrs.push_back(mp.mFR[s][d]);
      ptrs.push_back(mp.tq[s][d]);
    }
  for (int d = 0; d < 3; ++d) ptrs.push_back(mp.tk[d]);
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  mp = MPState{};
}

// Allocates the MultiParticle state once, on first use.
int MPEnsure(gomc_box *b) {
  MPState &mp = b->mp;
  if (mp.ready) return GOMC_OK;
  const size_t na = sizeof(double) * std::max(b->natoms, 1), nm = sizeof(double) * std::max(b->nmols, 1);
  bool ok = true;
  ok &= hipMalloc(&mp.xn, na) == hipSuccess;
  ok &= hipMalloc(&mp.yn, na) == hipSuccess;
  ok &= hipMalloc(&mp.zn, na) == hipSuccess;
  for (int s = 0; s < 2; ++s)
    for (int d = 0; d < 3; ++d) {
      ok &= hipMalloc(&mp.com[s][d], nm) == hipSuccess;
      ok &= hipMalloc(&mp.aF[s][d], na) == hipSuccess;
      ok &= hipMalloc(&mp.aFR[s][d], na) == hipSuccess;
      ok &= hipMalloc(&mp.mF[s][d], nm) == hipSuccess;
      ok &= hipMalloc(&mp.mFR[s][d], nm) == hipSuccess;
      ok &= hipMalloc(&mp.tq[s][d], nm) == hipSuccess;
    }
  for (int d = 0; d < 3; ++d) ok &= hipMalloc(&mp.tk[d], nm) == hipSuccess;
  ok &= hipMalloc(&mp.inRange, sizeof(int) * std::max(b->nmols, 1)) == hipSuccess;
  ok &= hipMalloc(&mp.fail, sizeof(int)) == hipSuccess;

  ok &= hipMalloc(&mp.wpart, sizeof(double) * ((b->nmols + kBlock - 1) / kBlock + 1)) == hipSuccess;
  ok &= hipMalloc(&mp.wcounter, sizeof(unsigned)) == hipSuccess;
  if (!ok) {
    MPFree(b);
    return Fail(b, GOMC_ERR_HIP, "MultiParticle allocation failed");
  }
  mp.molStart = b->molStartDev;  // owned by the box
  HIP_TRY(b, hipMemset(mp.wcounter, 0, sizeof(unsigned)));
  mp.ready = true;
  return GOMC_OK;
}

// The driver keeps the authoritative centres on the host; a MultiParticle
// step starts from them.
int MPUploadCom(gomc_box *b) {
  const size_t bytes = sizeof(double) * b->nmols;
  HIP_TRY(b, hipMemcpyAsync(b->mp.com[0][0], b->comx.data(), bytes, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->mp.com[0][1], b->comy.data(), bytes, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->mp.com[0][2], b->comz.data(), bytes, hipMemcpyHostToDevice, b->stream));
  return Sync(b);
}

// Pair forces of the current coordinates into force set `set`; energies out.
int MPPairForce(gomc_box *b, int set, double *lj, double *real) {
  MPState &mp = b->mp;
  int rc = BuildCells(b);
  if (rc) return rc;
  std::vector<int> ch;
  for (int c = 0; c < b->ncell; ++c)
    for (int o = 0; o < b->hCellCount[c]; o += kChunk) {
      ch.push_back(c);
      ch.push_back(o);
      ch.push_back(std::min(kChunk, b->hCellCount[c] - o));
    }
  const int n = int(ch.size() / 3);
  if (n > mp.maxChunks) {
    if (mp.chunks) (void)hipFree(mp.chunks);
    mp.chunks = nullptr;
    mp.maxChunks = 0;
    HIP_TRY(b, hipMalloc(&mp.chunks, sizeof(int) * 3 * std::max(n, 1)));
    mp.maxChunks = n;
  }
  mp.nchunks = n;
  if (n == 0) {
    *lj = *real = 0.0;
    return GOMC_OK;
  }
  HIP_TRY(b, hipMemcpyAsync(mp.chunks, ch.data(), sizeof(int) * ch.size(), hipMemcpyHostToDevice,
                            b->stream));
  // waves per workgroup and slices per chunk: a large box fills the device
  // with chunks alone, a small one splits each chunk's tiles over slices
  int G = 8, slices = 1;
  if (n >= kPairChunksFor4) {
    G = 4;
  } else if (n < kPairChunksFor8) {
    slices = std::min(kPairMaxSlices, std::max(1, (kPairSmallWaves + 8 * n - 1) / (8 * n)));
  }
  if (slices > 1) {
    const size_t need = size_t(slices) * 3 * size_t(b->natoms);
    if (need > mp.spartCap) {
      if (mp.spart) (void)hipFree(mp.spart);
      mp.spart = nullptr;
      mp.spartCap = 0;
      HIP_TRY(b, hipMalloc(&mp.spart, sizeof(double) * need));
      mp.spartCap = need;
    }
  }
  rc = EnsurePartials(b, n * slices);
  if (rc) return rc;
  PairForceArgs a;
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
  a.epsCn6 = b->tEpsCn6;
  a.nOver6 = b->tNOver6;
  a.ntypes = b->ntypes;
  a.cellStart = b->cellStart;
  a.cellVec = b->cellVec;
  a.nbr = b->nbr;
  a.chunks = mp.chunks;
  a.fx = mp.aF[set][0];
  a.fy = mp.aF[set][1];
  a.fz = mp.aF[set][2];
  a.partials = b->partials;
  a.counter = b->counter;
  a.result = b->devRes;
  a.slices = slices;
  a.natoms = b->natoms;
  a.spart = mp.spart;
  const dim3 grid(n * slices);
  if (G == 8)
    hipLaunchKernelGGL(PairForceKernel<8>, grid, dim3(8 * kChunk), 0, b->stream, a);
  else
    hipLaunchKernelGGL(PairForceKernel<4>, grid, dim3(4 * kChunk), 0, b->stream, a);
  HIP_TRY(b, hipGetLastError());
  if (slices > 1) {
    hipLaunchKernelGGL(PairForceSumKernel, dim3(n), dim3(kChunk), 0, b->stream, a);
    HIP_TRY(b, hipGetLastError());
  }
  rc = Sync(b);
  if (rc) return rc;
  *lj = b->hostRes[0];
  *real = b->hostRes[1];
  return GOMC_OK;
}

// Reciprocal forces of the current coordinates from the "new" sums.
int MPRecipForce(gomc_box *b, int set) {
  MPState &mp = b->mp;
  if (!b->ewald || !b->recipReady) {
    for (int d = 0; d < 3; ++d)
      HIP_TRY(b, hipMemsetAsync(mp.aFR[set][d], 0, sizeof(double) * b->natoms, b->stream));
    return Sync(b);
  }
  RecipForceArgs a;
  a.g = b->geom;
  a.x = b->x;
  a.y = b->y;
  a.z = b->z;
  a.q = b->q;
  a.mol = b->mol;
  a.molStart = mp.molStart;
  a.natoms = b->natoms;
  a.K = b->K;
  a.kx = b->kx;
  a.ky = b->ky;
  a.kz = b->kz;
  a.pref = b->pref;
  a.sumR = b->sumRnew;
  a.sumI = b->sumInew;
  a.fx = mp.aFR[set][0];
  a.fy = mp.aFR[set][1];
  a.fz = mp.aFR[set][2];
  a.kseg = b->kseg;
  a.cz = (1.0 / b->geom.L[2]) * (2.0 * M_PI);  // KVectorsOrth's z spacing
  a.nseg = b->nseg;
  const int g = (b->natoms + kBlock - 1) / kBlock;
  // segment chunks, about 1024 workgroups in the first pass
  int chunks = std::min(std::max(1, b->nseg), std::max(1, (1024 + g - 1) / g));
  a.segChunk = std::max(1, (b->nseg + chunks - 1) / chunks);
  chunks = std::max(1, (b->nseg + a.segChunk - 1) / a.segChunk);
  a.kChunks = chunks;
  const size_t need = size_t(chunks) * 3 * size_t(std::max(b->natoms, 1));
  if (need > mp.fpartCap) {
    if (mp.fpart) (void)hipFree(mp.fpart);
    mp.fpart = nullptr;
    mp.fpartCap = 0;
    HIP_TRY(b, hipMalloc(&mp.fpart, sizeof(double) * need));
    mp.fpartCap = need;
  }
  a.part = mp.fpart;
  hipLaunchKernelGGL(RecipForcePartKernel, dim3(g * chunks), dim3(kBlock), 0, b->stream, a);
  HIP_TRY(b, hipGetLastError());
  hipLaunchKernelGGL(RecipForceKernel, dim3(g), dim3(kBlock), 0, b->stream, a);
  HIP_TRY(b, hipGetLastError());
  return Sync(b);
}

static int MolTerms(gomc_box *b, int set, int comSet, int mode) {
  MPState &mp = b->mp;
  MolTermsArgs a;
  a.g = b->geom;
  a.x = b->x;
  a.y = b->y;
  a.z = b->z;
  a.cx = mp.com[comSet][0];
  a.cy = mp.com[comSet][1];
  a.cz = mp.com[comSet][2];
  a.molStart = mp.molStart;
  a.nmols = b->nmols;
  a.mode = mode;
  for (int d = 0; d < 3; ++d) {
    a.aF[d] = mp.aF[set][d];
    a.aFR[d] = mp.aFR[set][d];
    a.mF[d] = mp.mF[set][d];
    a.mFR[d] = mp.mFR[set][d];
    a.tq[d] = mp.tq[set][d];
  }
  const int g = std::max(1, (b->nmols + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(MolTermsKernel, dim3(g), dim3(kBlock), 0, b->stream, a);
  HIP_TRY(b, hipGetLastError());
  return Sync(b);
}

int MPMolTerms(gomc_box *b, int set, int comSet) { return MolTerms(b, set, comSet, 7); }

// Transform -> CalcEn -> GetCoeff (MultiParticle.h:332-511) on the
// reference force set; out = {inter, real, recip, w_ratio, fail}.  On
// success the trial coordinates are the box's current ones until
// MPAccept decides.
static int MPTrial(gomc_box *b, int rotate, uint64_t step, double max, double out[5]) {
  MPState &mp = b->mp;
  const int R = mp.ref, N = 1 - mp.ref;
  b->stageMol = -1;
  HIP_TRY(b, hipMemsetAsync(mp.fail, 0, sizeof(int), b->stream));
  TransformArgs t;
  t.g = b->geom;
  t.nmols = b->nmols;
  t.rotate = rotate;
  t.step = step;
  t.seed = b->seed;
  t.beta = b->beta;
  t.max = max;
  t.x = b->x;
  t.y = b->y;
  t.z = b->z;
  t.cx = mp.com[0][0];
  t.cy = mp.com[0][1];
  t.cz = mp.com[0][2];
  t.xn = mp.xn;
  t.yn = mp.yn;
  t.zn = mp.zn;
  t.ncx = mp.com[1][0];
  t.ncy = mp.com[1][1];
  t.ncz = mp.com[1][2];
  t.molStart = mp.molStart;
  for (int d = 0; d < 3; ++d) {
    t.mF[d] = mp.mF[R][d];
    t.mFR[d] = mp.mFR[R][d];
    t.tq[d] = mp.tq[R][d];
    t.tk[d] = mp.tk[d];
  }
  t.inRange = mp.inRange;
  t.fail = mp.fail;
  const int g = std::max(1, (b->nmols + kBlock - 1) / kBlock);
  hipLaunchKernelGGL(TransformKernel, dim3(g), dim3(kBlock), 0, b->stream, t);
  HIP_TRY(b, hipGetLastError());
  int fail = 0;
  HIP_TRY(b, hipMemcpyAsync(&fail, mp.fail, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  int rc = Sync(b);
  if (rc) return rc;
  out[0] = b->pot.inter;
  out[1] = b->pot.real;
  out[2] = b->pot.recip;
  out[3] = 1.0;
  out[4] = fail ? 1.0 : 0.0;
  if (fail) return GOMC_OK;
  // CalcEn on the trial coordinates: they become the box's coordinates.
  std::swap(b->x, mp.xn);
  std::swap(b->y, mp.yn);
  std::swap(b->z, mp.zn);
  for (int d = 0; d < 3; ++d) std::swap(mp.com[0][d], mp.com[1][d]);
  if (b->ewald) {
    rc = RecipSums(b, &out[2]);
    if (rc) return rc;
  }
  rc = MPPairForce(b, N, &out[0], &out[1]);
  if (!rc) rc = MPRecipForce(b, N);
  if (!rc) rc = MolTerms(b, N, 0, 7);
  if (rc) return rc;
  CoeffArgs c;
  c.nmols = b->nmols;
  c.rotate = rotate;
  c.lBeta = kLambda * b->beta;
  c.max = max;
  for (int d = 0; d < 3; ++d) {
    c.tk[d] = mp.tk[d];
    c.nF[d] = mp.mF[N][d];
    c.nFR[d] = mp.mFR[N][d];
    c.nT[d] = mp.tq[N][d];
    c.oF[d] = mp.mF[R][d];
    c.oFR[d] = mp.mFR[R][d];
    c.oT[d] = mp.tq[R][d];
  }
  c.inRange = mp.inRange;
  c.parts = mp.wpart;
  c.counter = mp.wcounter;
  c.result = b->devRes + 4;
  hipLaunchKernelGGL(CoeffKernel, dim3(g), dim3(kBlock), 0, b->stream, c);
  HIP_TRY(b, hipGetLastError());
  rc = Sync(b);
  if (rc) return rc;
  double w = b->hostRes[4];
  out[3] = std::isfinite(w) ? w : 0.0;
  return GOMC_OK;
}

// MultiParticle::Accept's state change (MultiParticle.h:522-538).
static int MPAccept(gomc_box *b, bool result, bool evaluated, const double e[3]) {
  MPState &mp = b->mp;
  if (result) {
    b->pot.inter = e[0];
    b->pot.real = e[1];
    b->pot.recip = e[2];
    b->pot.Total();
    mp.ref = 1 - mp.ref;
    TrialCellsInvalidate(b);  // every molecule moved
    if (b->ewald) gomc_recip_accept(b);
    const size_t na = sizeof(double) * b->natoms, nm = sizeof(double) * b->nmols;
    HIP_TRY(b, hipMemcpyAsync(b->hx.data(), b->x, na, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->hy.data(), b->y, na, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->hz.data(), b->z, na, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->comx.data(), mp.com[0][0], nm, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->comy.data(), mp.com[0][1], nm, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->comz.data(), mp.com[0][2], nm, hipMemcpyDeviceToHost, b->stream));
    int rc = Sync(b);
    // every molecule moved: EwaldCached::BoxReciprocalSums refilled the cache
    if (!rc && b->cached) rc = CacheBuild(b);
    return rc;
  }
  if (evaluated) {
    std::swap(b->x, mp.xn);
    std::swap(b->y, mp.yn);
    std::swap(b->z, mp.zn);
    for (int d = 0; d < 3; ++d) std::swap(mp.com[0][d], mp.com[1][d]);
  }
  return GOMC_OK;
}

// Prep's force evaluation (MultiParticle.h:207-230) for the current state.
static int MPPrepForces(gomc_box *b) {
  const int R = b->mp.ref;
  int rc = GOMC_OK;
  if (b->ewald) rc = gomc_recip_copy_ref_to_new(b);
  if (!rc) rc = MPRecipForce(b, R);
  double lj = 0, real = 0;
  if (!rc) rc = MPPairForce(b, R, &lj, &real);
  if (!rc) rc = MolTerms(b, R, 0, 7);
  if (!rc) b->pot.Total();
  return rc;
}

// One MultiParticle move of the driver: Prep (the move type from MT19937,
// MultiParticle.h:176-182) -> Transform -> CalcEn -> Accept, with the
// acceptance draw from Philox counter N_mol (MultiParticle.h:521).
int MPMove(gomc_box *b, uint64_t step, int &moveType, int &accepted, int &failed, double dE[3]) {
  int rc = MPEnsure(b);
  if (!rc) rc = FlushPending(b);
  if (!rc) rc = MPUploadCom(b);
  if (rc) return rc;
  MPState &mp = b->mp;
  bool allTranslate = true;
  for (const auto &k : b->kindType) allTranslate = allTranslate && k.size() == 1;
  moveType = allTranslate ? 0 : int(b->mt.Below(1));
  if (mp.singleAccepted) {
    rc = MPPrepForces(b);
    if (rc) return rc;
  }
  const double max = moveType ? mp.rMax : mp.tMax;
  double out[5];
  rc = MPTrial(b, moveType, step, max, out);
  if (rc) return rc;
  const bool fail = out[4] != 0.0;
  Energies pn = b->pot;
  if (!fail) {
    pn.inter = out[0];
    pn.real = out[1];
    pn.recip = out[2];
  }
  const double uBoltz = std::exp(-b->beta * (pn.Total() - b->pot.Total()));
  const double accept = (fail ? 1.0 : out[3]) * uBoltz;
  const U64x4 r = Philox(uint64_t(uint32_t(b->nmols)), 0, 0, 0, step, b->seed);
  const double pr = U01(r.v[0]);
  const bool result = !fail && pr < accept;
  for (int i = 0; i < 3; ++i) dE[i] = fail ? 0.0 : out[i] - (i == 0 ? b->pot.inter : i == 1 ? b->pot.real : b->pot.recip);
  rc = MPAccept(b, result, !fail, out);
  if (rc) return rc;
  mp.itries[moveType]++;
  if (result) mp.iacc[moveType]++;
  mp.singleAccepted = false;
  accepted = result;
  failed = fail;
  return GOMC_OK;
}

}  // namespace gomc

// ====================================================================== C ABI
extern "C" {

static int CopyOut(gomc_box *b, double *const *dst, double *const *src, int n) {
  for (int d = 0; d < 3; ++d)
    if (dst[d]) HIP_TRY(b, hipMemcpyAsync(dst[d], src[d], sizeof(double) * n, hipMemcpyDeviceToHost, b->stream));
  return Sync(b);
}

int gomc_box_force(gomc_box *b, double *lj, double *real, double *afx, double *afy, double *afz,
                   double *mfx, double *mfy, double *mfz) {
  if (!b || !lj || !real) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = MPEnsure(b);
  if (!rc) rc = FlushPending(b);
  if (!rc) rc = MPPairForce(b, b->mp.ref, lj, real);
  if (!rc) rc = MolTerms(b, b->mp.ref, 0, 1);
  if (rc) return rc;
  double *da[3] = {afx, afy, afz}, *dm[3] = {mfx, mfy, mfz};
  rc = CopyOut(b, da, b->mp.aF[b->mp.ref], b->natoms);
  if (!rc) rc = CopyOut(b, dm, b->mp.mF[b->mp.ref], b->nmols);
  return rc;
}

int gomc_box_force_recip(gomc_box *b, double *afx, double *afy, double *afz, double *mfx,
                         double *mfy, double *mfz) {
  if (!b) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = MPEnsure(b);
  if (!rc) rc = FlushPending(b);
  if (!rc) rc = MPRecipForce(b, b->mp.ref);
  if (!rc) rc = MolTerms(b, b->mp.ref, 0, 2);
  if (rc) return rc;
  double *da[3] = {afx, afy, afz}, *dm[3] = {mfx, mfy, mfz};
  rc = CopyOut(b, da, b->mp.aFR[b->mp.ref], b->natoms);
  if (!rc) rc = CopyOut(b, dm, b->mp.mFR[b->mp.ref], b->nmols);
  return rc;
}

int gomc_torque(gomc_box *b, double *tx, double *ty, double *tz) {
  if (!b) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = MPEnsure(b);
  if (!rc) rc = FlushPending(b);
  if (!rc) rc = MPUploadCom(b);
  if (!rc) rc = MolTerms(b, b->mp.ref, 0, 4);
  if (rc) return rc;
  double *dt[3] = {tx, ty, tz};
  return CopyOut(b, dt, b->mp.tq[b->mp.ref], b->nmols);
}

int gomc_mp_trial(gomc_box *b, int move_type, uint64_t step, double max, double out[5]) {
  if (!b || !out || (move_type != 0 && move_type != 1) || !(max > 0.0)) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = MPEnsure(b);
  if (!rc) rc = FlushPending(b);
  if (!rc) rc = MPUploadCom(b);
  if (!rc) rc = MPTrial(b, move_type, step, max, out);
  return rc;
}

int gomc_mp_accept(gomc_box *b, int accept, const double out[5]) {
  if (!b || !out || !b->mp.ready) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  const bool fail = out[4] != 0.0;
  if (accept && fail) return Fail(b, GOMC_ERR_STATE, "cannot accept a failed transform");
  int rc = MPAccept(b, accept != 0, !fail, out);
  if (!rc) b->mp.singleAccepted = false;
  return rc;
}

int gomc_mp_settings(const gomc_box *b, double out[2]) {
  if (!b || !out) return GOMC_ERR_ARG;
  out[0] = b->mp.tMax;
  out[1] = b->mp.rMax;
  return GOMC_OK;
}

}  // extern "C"

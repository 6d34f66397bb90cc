// Batched persistent MC on gfx950: a 1024-thread workgroup runs the boxes
// of its task list one after another, making all of a box's moves of a
// launch without leaving the kernel; a large box is run by a team of 2, 4
// or 8 workgroups that split its pair items and k-vectors and exchange
// partial sums once per move.  Per move: Philox draws (made 64 moves ahead, one lane each),
// the trial coordinates (Translate/Rotate, src/Coordinates.cpp:83-128), the
// pair-energy change over every other molecule (MoleculeInter,
// src/CalculateEnergy.cpp:581-686), the structure-factor change per
// k-vector (MolReciprocal, src/Ewald.cpp:409-473), the Metropolis test
// (src/moves/Translate.h:97-138) and, on acceptance, the update of
// coordinates and S(k).
//
// Phases e^{i k.r} come from per-dimension tables e^{i (c_d n) r_d},
// n = 0..nmax, built each move for the moving molecule's old and new
// positions (one sincos per table entry), and advanced along nz.
#include "batch.hpp"
#include "device_common.hpp"
#include "erfcx_poly.hpp"

namespace gomc {
namespace {

// Pointers into global memory (HBM) and LDS.  The phases are out-of-line
// functions; a plain pointer parameter there is a generic (flat) address,
// and every flat access also counts against the LDS wait counter, so a
// wait for an LDS read would stall on outstanding HBM loads.
#define GMEM __attribute__((address_space(1)))
#define LMEM __attribute__((address_space(3)))
template <typename T>
__device__ __forceinline__ GMEM T *Gp(T *p) {
  return (GMEM T *)p;
}
template <typename T>
__device__ __forceinline__ LMEM T *Lp(T *p) {
  return (LMEM T *)p;
}

// Phase-table entries: FP64 in the FP64 mode, FP32 in the mixed mode.
template <typename T>
struct CplxT {
  T re, im;
};
using Cplx = CplxT<double>;

template <typename T>
__device__ __forceinline__ CplxT<T> CMul(CplxT<T> a, CplxT<T> b) {
  return {fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}

// Precision traits: P = 0 FP64 everywhere; P = 1 FP32 pair terms and phase
// products (distances, accumulation, S(k), totals and the Metropolis
// exponent stay FP64).
template <int P>
struct Prec {
  using T = double;
};
template <>
struct Prec<1> {
  using T = float;
};

struct alignas(16) MoveShared {
  double oldp[3][kBatchMaxAtoms], newp[3][kBatchMaxAtoms];
  double shift[3], rot[9], com[3], newCom[3];
  int mol, type, accept;
  int abort;  // a team exchange ran out of time: the box stops
};

// Reduction of four doubles over the workgroup in a fixed order; result in
// thread 0.
__device__ __forceinline__ void BlockSum4(double v[4], double (*lds)[4]) {
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
#pragma unroll
  for (int i = 0; i < 4; ++i) v[i] = WaveSum(v[i]);
  if (lane == 0)
#pragma unroll
    for (int i = 0; i < 4; ++i) lds[wave][i] = v[i];
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double s = 0.0;
      for (int w = 0; w < kBatchThreads / kWave; ++w) s += lds[w][i];
      v[i] = s;
    }
  }
}

// One phase-table entry: e^{i (c_d n) r} (kept out of line: the FP64
// sincos argument reduction is register-hungry and runs once per entry).
template <typename T>
__device__ __forceinline__ CplxT<T> PhaseEntry(double kc, int n, double r) {
  const double ang = (kc * double(n)) * r;
  if constexpr (sizeof(T) == 8) {
    double sn, c;
    sincos(ang, &sn, &c);
    return {c, sn};
  } else {
    // FP64 reduction to [-pi, pi), FP32 sincos of the remainder.
    const double turns = ang * (0.5 * M_1_PI);
    const float red = float((turns - rint(turns)) * (2.0 * M_PI));
    float sn, c;
    sincosf(red, &sn, &c);
    return {c, sn};
  }
}

// Thread 0's proposal: move settings, move kind, molecule, and the shift or
// rotation matrix (out of line; it runs on one lane).
__device__ __noinline__ void Propose(const BatchBox &sb, BatchState &st, MoveShared &ms,
                                     const double *u, int64_t step, const double *cx,
                                     const double *cy, const double *cz) {
  const double lmin = fmin(sb.L[0], fmin(sb.L[1], sb.L[2]));
  // MoveSettings::AdjustMoves (src/MoveSettings.cpp:148-258).
  if ((uint64_t(step) + 1) % uint64_t(sb.perAdjust) == 0) {
    for (int t = 0; t < 2; ++t) {
      if (st.tempTries[t] > 0) {
        const double acc = double(st.tempAccepted[t]) / double(st.tempTries[t]);
        const double fract = acc / 0.5;
        st.scale[t] *= (fract > 0.0) ? fract : 0.5;
      }
      const double lo = t == 0 ? 1.0e-10 : 1.0e-6;
      const double hi = t == 0 ? (lmin * 0.5) - 1.0e-7 : M_PI - 1.0e-7;
      st.scale[t] = fmin(fmax(st.scale[t], lo), hi);
      st.tempTries[t] = 0;
      st.tempAccepted[t] = 0;
    }
  }
  const int nmols = sb.nmols;
  int t = (u[0] < sb.pDisp) ? 0 : 1;
  int m = int(u[1] * nmols);
  if (m >= nmols) m = nmols - 1;
  if (sb.molLen <= 1) t = 0;
  ms.mol = m;
  ms.type = t;
  ms.com[0] = cx[m];
  ms.com[1] = cy[m];
  ms.com[2] = cz[m];
  if (t == 0) {
    const double bb = st.scale[0], b2 = 2.0 * bb;
    ms.shift[0] = u[2] * b2 - bb;
    ms.shift[1] = u[3] * b2 - bb;
    ms.shift[2] = u[4] * b2 - bb;
#pragma unroll
    for (int d = 0; d < 3; ++d) ms.newCom[d] = Wrap(ms.com[d] + ms.shift[d], sb.L[d]);
  } else {
#pragma unroll
    for (int d = 0; d < 3; ++d) ms.newCom[d] = ms.com[d];  // a rotation keeps the centre
    const double bb = st.scale[1];
    const double theta = u[2] * (2.0 * bb) - bb;
    const double cu = u[3] * 2.0 - 1.0;
    const double root = sqrt(1.0 - cu * cu);
    const double sp = u[6], cp = u[7];  // sincos(2 pi u[4]), made with the draws
    const double ax[3] = {root * cp, root * sp, cu};
    double sn, c;
    sincos(theta, &sn, &c);
    const double cr[3][3] = {{0.0, -ax[2], ax[1]}, {ax[2], 0.0, -ax[0]}, {-ax[1], ax[0], 0.0}};
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        ms.rot[i * 3 + j] = (i == j ? c : 0.0) + (sn * cr[i][j] + (1 - c) * (ax[i] * ax[j]));
  }
}

// Philox draws of one move: u[0..5] (two blocks, counter {step, 0|1}), and
// sin, cos of a rotation's axis azimuth 2 pi u[4] in u[6], u[7] (a function
// of the draw alone, so made here, 64 moves ahead, not by thread 0 within
// the move).
__device__ __noinline__ void MoveDraws(uint64_t seed, uint64_t step, double *out) {
  const U64x4 r0 = Philox(step, 0, 0, 0, seed, kPhiloxTag);
  const U64x4 r1 = Philox(step, 1, 0, 0, seed, kPhiloxTag);
  out[0] = U01(r0.v[0]);
  out[1] = U01(r0.v[1]);
  out[2] = U01(r0.v[2]);
  out[3] = U01(r0.v[3]);
  out[4] = U01(r1.v[0]);
  out[5] = U01(r1.v[1]);
  const double phi = out[4] * (2.0 * M_PI);
  sincos(phi, &out[6], &out[7]);
}

// Phase tables e^{i (c_d n) r_d}, n = 0..nmax_d, for the moving molecule's
// old (pos < len) and new positions; one entry per thread-iteration.
template <typename T>
__device__ __noinline__ void BuildTables(const BatchBox &sb, const MoveShared &ms, int len, int nTab,
                                         CplxT<T> (*tab)[3][kBatchMaxN + 1]) {
  const int per = (sb.nmax[0] + 1) + (sb.nmax[1] + 1) + (sb.nmax[2] + 1);
  for (int e = threadIdx.x; e < nTab; e += kBatchThreads) {
    const int pos = e / per;
    int rem = e - pos * per;
    int d = 0;
    while (rem > sb.nmax[d]) {
      rem -= sb.nmax[d] + 1;
      ++d;
    }
    const double r = pos < len ? ms.oldp[d][pos] : ms.newp[d][pos - len];
    tab[pos][d][rem] = PhaseEntry<T>(sb.kc[d], rem, r);
  }
}

// Per-wave queue of in-cutoff pairs: squared distance and a code
// (pair-table index | new-position flag << 8).
struct PairQueue {
  double d2[2 * kWave];
  int code[2 * kWave];
};

// Energy terms of one queued pair (FFParticle::CalcEn / CalcCoulomb,
// src/FFParticle.cpp:295-399), signed: + new position, - old position.
template <typename T>
__device__ __forceinline__ void PairTerm(const BatchBox &sb, double d2, int code, double &lj,
                                         double &real) {
  const int pq = code & 0xff;
  const double sign = (code >> 8) ? 1.0 : -1.0;
  const double qiqj = sb.qq[pq], eps = sb.epsCn[pq];
  if (sb.elect && qiqj != 0.0 && !(sb.rCutCoulSq < d2)) {
    const T dist = sqrt(T(d2));
    const T ec = sb.ewald ? T(qiqj) * erfc(T(sb.alpha) * dist) / dist : T(qiqj) / dist;
    real += sign * double(ec);
  }
  if (eps != 0.0 && !(sb.rCutLJSq < d2)) {
    // n = 12 (checked at setup): repulsion = attraction^2
    const T rRat2 = T(sb.sigSq[pq]) / T(d2);
    const T att = rRat2 * rRat2 * rRat2;
    lj += sign * double(T(eps) * (att * att - att));
  }
}

struct LjReal {
  double lj, real, ov;
};

// FP64 pair terms: one reciprocal square root serves the Coulomb distance
// and 1/r^2 of the Lennard-Jones ratio.
template <>
__device__ __forceinline__ void PairTerm<double>(const BatchBox &sb, double d2, int code,
                                                 double &lj, double &real) {
  const int pq = code & 0xff;
  const double sign = (code >> 8) ? 1.0 : -1.0;
  const double qiqj = sb.qq[pq], eps = sb.epsCn[pq];
  const double ir = Rsqrt64(d2);
  if (sb.elect && qiqj != 0.0 && !(sb.rCutCoulSq < d2)) {
    const double ec = sb.ewald ? qiqj * ErfcPoly(sb.alpha * (d2 * ir)) * ir : qiqj * ir;
    real += sign * ec;
  }
  if (eps != 0.0 && !(sb.rCutLJSq < d2)) {
    const double rRat2 = sb.sigSq[pq] * (ir * ir);
    const double att = rRat2 * rRat2 * rRat2;
    lj += sign * (eps * (att * att - att));
  }
}

// Evaluates `n` queued pairs (one per lane) and adds them to acc.
template <typename T>
__device__ __noinline__ LjReal DrainQueue(const BatchBox &sb, const LMEM PairQueue *q, int n,
                                          LjReal acc) {
  const int lane = threadIdx.x & (kWave - 1);
  if (lane < n) PairTerm<T>(sb, q->d2[lane], q->code[lane], acc.lj, acc.real);
  return acc;
}

// The pair phase and the reciprocal phase of a move are independent once
// the trial coordinates and phase tables exist, and both are bound by memory
// and pipeline latency: waves 0 .. sb.pairWaves-1 run the pair phase while
// the other sb.kWidth threads run the reciprocal phase, with no barrier
// inside either.

// Pair-energy change of the move (MoleculeInter, src/CalculateEnergy.cpp:
// 581-686), in two levels, by the waves of the pair group.  Level 1
// (boxes with sb.cull): a molecule whose centre is at least rc + 2 r_max
// from both the old and the new centre of the moving molecule (minimum
// image) holds no atom pair inside the cutoff (triangle inequality); each
// wave tests its own chunks of 64 molecules and lists the near ones in its
// own part of nearList (kNearListLen / pairWaves entries: no more molecules
// than that reach a wave, so the list cannot overflow).  Level 2: one atom of a near molecule
// (of every other molecule without the cull) per lane, tested against the
// old and new positions of the moving atoms; distances inside the cutoff are
// compacted into the wave's queue (ballot + prefix count) and their energies
// evaluated 64 at a time, so the erfc path runs with full waves.  A team
// member culls its share of the molecules (or, without the cull, takes its
// share of the items).  CX/CY/CZ point to LDS or HBM (ComPtr).  Returns this
// thread's {lj, real, overlap}: overlap = 1 when a new position lies inside
// rCutLow of an atom.
template <typename T, int LEN, typename ComPtr>
__device__ __noinline__ LjReal PairPhase(const BatchBox &sb, const MoveShared &ms,
                                         const GMEM double *X, const GMEM double *Y,
                                         const GMEM double *Z, ComPtr CX, ComPtr CY, ComPtr CZ,
                                         int m, int nmols, int rank, int team, PairQueue *queues,
                                         int *nearList) {
  constexpr int len = LEN;
  BoxGeom g;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    g.L[d] = sb.L[d];
    g.half[d] = sb.half[d];
  }
  const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;  // wave < pairWaves
  const int pt = sb.pairWaves * kWave;
  PairQueue &q = queues[wave];
  int *myNear = nearList + wave * (kNearListLen / sb.pairWaves);
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int count = 0;  // wave-uniform
  LjReal acc{0.0, 0.0, 0.0};
  auto isNear = [&](int j) {
    if (j == m) return false;
    const double cx = CX[j], cy = CY[j], cz = CZ[j];
    const double d2o = DistSq(g, ms.com[0], ms.com[1], ms.com[2], cx, cy, cz);
    const double d2n = DistSq(g, ms.newCom[0], ms.newCom[1], ms.newCom[2], cx, cy, cz);
    return d2o < sb.nearCutSq || d2n < sb.nearCutSq;
  };
  // one level-2 item: atom qa of molecule jm against the moving atoms
  auto item = [&](bool valid, int jm, int qa) {
    const int jj = jm * len + qa;
    const double xj = X[jj], yj = Y[jj], zj = Z[jj];
#pragma unroll
    for (int p = 0; p < len; ++p) {
      const int pq = p * kBatchMaxAtoms + qa;
      const bool live = valid && ((sb.live >> pq) & 1);
#pragma unroll
      for (int isNew = 0; isNew < 2; ++isNew) {
        const double px = isNew ? ms.newp[0][p] : ms.oldp[0][p];
        const double py = isNew ? ms.newp[1][p] : ms.oldp[1][p];
        const double pz = isNew ? ms.newp[2][p] : ms.oldp[2][p];
        const double d2 = DistSq(g, px, py, pz, xj, yj, zj);
        const bool inCut = valid && sb.rCutSq > d2;
        if (isNew && inCut && d2 < sb.rCutLowSq) acc.ov = 1.0;
        const bool push = live && inCut;
        const uint64_t mask = __ballot(push);
        if (push) {
          const int slot = count + __popcll(mask & below);
          q.d2[slot] = d2;
          q.code[slot] = pq | (isNew << 8);
        }
        count += __popcll(mask);
        if (count >= kWave) {  // drain one full wave of pairs
          __builtin_amdgcn_wave_barrier();
          acc = DrainQueue<T>(sb, Lp(&q), kWave, acc);
          const double d2n = q.d2[kWave + lane];
          const int cn = q.code[kWave + lane];
          __builtin_amdgcn_wave_barrier();
          q.d2[lane] = d2n;
          q.code[lane] = cn;
          count -= kWave;
          __builtin_amdgcn_wave_barrier();
        }
      }
    }
  };
  if (sb.cull) {
    const int mper = (nmols + team - 1) / team;
    const int m0 = rank * mper, m1 = min(nmols, m0 + mper);
    int nnear = 0;  // wave-uniform
    for (int base = m0 + wave * kWave; base < m1; base += pt) {
      const int j = base + lane;
      const bool near = j < m1 && isNear(j);
      const uint64_t nmask = __ballot(near);
      if (near) myNear[nnear + __popcll(nmask & below)] = j;
      nnear += __popcll(nmask);
    }
    __builtin_amdgcn_wave_barrier();
    const int items = nnear * len;
    for (int ib = 0; ib < items; ib += kWave) {
      const int it = ib + lane;
      const bool valid = it < items;
      const int itc = valid ? it : 0;
      const uint32_t k = uint32_t((uint64_t(uint32_t(itc)) * sb.magicLen) >> 32);  // itc / len
      item(valid, myNear[k], itc - int(k) * len);
    }
  } else {
    const int items = (nmols - 1) * len;
    const int per = (items + team - 1) / team;
    const int lo = rank * per, hi = min(items, lo + per);
    for (int ib = lo + wave * kWave; ib < hi; ib += pt) {
      const int it = ib + lane;
      const bool valid = it < hi;
      const int itc = valid ? it : 0;
      const uint32_t k = uint32_t((uint64_t(uint32_t(itc)) * sb.magicLen) >> 32);  // itc / len
      item(valid, int(k) < m ? int(k) : int(k) + 1, itc - int(k) * len);
    }
  }
  __builtin_amdgcn_wave_barrier();
  return DrainQueue<T>(sb, Lp(&q), count, acc);
}

template <typename T, typename ComPtr>
__device__ __forceinline__ LjReal PairDispatch(const BatchBox &sb, const MoveShared &ms,
                                               const GMEM double *X, const GMEM double *Y,
                                               const GMEM double *Z, ComPtr CX, ComPtr CY,
                                               ComPtr CZ, int m, int len, int nmols, int rank,
                                               int team, PairQueue *queues, int *nearList) {
  switch (len) {
    case 1:
      return PairPhase<T, 1>(sb, ms, X, Y, Z, CX, CY, CZ, m, nmols, rank, team, queues, nearList);
    case 2:
      return PairPhase<T, 2>(sb, ms, X, Y, Z, CX, CY, CZ, m, nmols, rank, team, queues, nearList);
    case 3:
      return PairPhase<T, 3>(sb, ms, X, Y, Z, CX, CY, CZ, m, nmols, rank, team, queues, nearList);
    default:
      return PairPhase<T, 4>(sb, ms, X, Y, Z, CX, CY, CZ, m, nmols, rank, team, queues, nearList);
  }
}

// Reciprocal-space energy change of the move (MolReciprocal,
// src/Ewald.cpp:409-473): per k-vector the change (dR, dI) of the
// structure factor and prefact * ((R+dR)^2 + (I+dI)^2 - R^2 - I^2),
// written as prefact * (dR (2R + dR) + dI (2I + dI)) so no large terms
// cancel.  (dR, dI) are kept in dRs/dIs for the commit.
//
// A thread takes one segment of S k-vectors per round: consecutive nz at
// fixed (nx, ny).  The phase of each of the 2 LEN positions is formed from
// the tables at the segment start, (ex * ey) * ez, and advanced along nz by
// one complex product with e^{i c_z z} per step, so a k-vector costs no
// table reads after the first of its segment.  The loads of two steps are
// issued before they are used.
// pend: the previous move was accepted and its (dR, dI), still in the
// thread's slots, are added to S(k) here (Ewald::UpdateRecip,
// src/Ewald.cpp:1420) before the slots are reused.  Returns this thread's
// energy sum.  Unused slots (kn -1, pref 0, S(k) 0) add zero and are not
// written.
constexpr int kZeroCode = (128 << 8) | (128 << 16);  // n = (0, 0, 0)

template <typename T>
__device__ __forceinline__ CplxT<T> TablePhase(const CplxT<T> (*row)[kBatchMaxN + 1], int code) {
  const int nx = code & 0xff;
  const int ny = ((code >> 8) & 0xff) - 128;
  const int nz = ((code >> 16) & 0xff) - 128;
  CplxT<T> ey = row[1][ny < 0 ? -ny : ny], ez = row[2][nz < 0 ? -nz : nz];
  if (ny < 0) ey.im = -ey.im;
  if (nz < 0) ez.im = -ez.im;
  return CMul(CMul(row[0][nx], ey), ez);
}

// A value every lane holds alike, moved to scalar registers.
__device__ __forceinline__ double Uniform(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(u));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(u >> 32));
  return __builtin_bit_cast(double, (uint64_t(hi) << 32) | lo);
}
__device__ __forceinline__ float Uniform(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(uint32_t, v)));
}

__device__ __forceinline__ int64_t UniformI64(int64_t v) {
  const uint64_t u = uint64_t(v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(u));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(u >> 32));
  return int64_t((uint64_t(hi) << 32) | lo);
}

template <typename T, int S, int LEN>
__device__ __noinline__ double RecipPhase(const BatchBox &sb, const CplxT<T> (*tab)[3][kBatchMaxN + 1],
                                          const GMEM int32_t *__restrict__ kn,
                                          const GMEM double *__restrict__ pref,
                                          GMEM double *__restrict__ sumR, GMEM double *__restrict__ sumI,
                                          GMEM double *__restrict__ dRs, GMEM double *__restrict__ dIs,
                                          int rank, int team, bool pend) {
  // steps whose loads are issued together (registers leave room for two)
  constexpr int G = S < 2 ? S : 2;
  T q[LEN];
  CplxT<T> stepO[LEN], stepN[LEN];
#pragma unroll
  for (int p = 0; p < LEN; ++p) {  // the same for every lane: kept in scalar registers
    q[p] = Uniform(fabs(sb.charge[p]) < 0.000000001 ? T(0) : T(sb.charge[p]));
    const CplxT<T> so = tab[p][2][1], sn = tab[LEN + p][2][1];
    stepO[p] = {Uniform(so.re), Uniform(so.im)};
    stepN[p] = {Uniform(sn.re), Uniform(sn.im)};
  }
  double e = 0.0;
  // thread t of the reciprocal group owns column t of the W-wide layout
  const int W = sb.kWidth, t = int(threadIdx.x) - sb.pairWaves * kWave;
  for (int r = rank; r < sb.kChunk; r += team) {  // a team member's rounds
    CplxT<T> eO[LEN], eN[LEN];
#pragma unroll 1
    for (int jb = 0; jb < S; jb += G) {
      int code[G];
      double r0[G], i0[G], pf[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int s = ((r * S) + jb + g) * W + t;
        code[g] = kn[s];
        r0[g] = sumR[s];
        i0[g] = sumI[s];
        pf[g] = pref[s];
      }
      if (pend) {
        double ar[G], ai[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int s = ((r * S) + jb + g) * W + t;
          ar[g] = dRs[s];
          ai[g] = dIs[s];
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
          // unused slots hold no change (their dRs/dIs are never written)
          if (code[g] >= 0) {
            const int s = ((r * S) + jb + g) * W + t;
            r0[g] += ar[g];
            i0[g] += ai[g];
            sumR[s] = r0[g];
            sumI[s] = i0[g];
          }
        }
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int j = jb + g;
        const bool used = code[g] >= 0;
        if (j == 0) {
          const int c0 = used ? code[g] : kZeroCode;
          // one position at a time: the scheduler would otherwise hoist all
          // 6 LEN table reads and run out of registers
#pragma unroll
          for (int p = 0; p < LEN; ++p) {
            eO[p] = TablePhase<T>(tab[p], c0);
            __builtin_amdgcn_sched_barrier(0);
            eN[p] = TablePhase<T>(tab[LEN + p], c0);
            __builtin_amdgcn_sched_barrier(0);
          }
        } else {
#pragma unroll
          for (int p = 0; p < LEN; ++p) {
            eO[p] = CMul(eO[p], stepO[p]);
            eN[p] = CMul(eN[p], stepN[p]);
          }
        }
        T sRN = 0, sIN = 0, sRO = 0, sIO = 0;
#pragma unroll
        for (int p = 0; p < LEN; ++p) {
          sRN += q[p] * eN[p].re;
          sIN += q[p] * eN[p].im;
          sRO += q[p] * eO[p].re;
          sIO += q[p] * eO[p].im;
        }
        const double dR = used ? double(sRN - sRO) : 0.0;
        const double dI = used ? double(sIN - sIO) : 0.0;
        if (used) {
          const int s = ((r * S) + j) * W + t;
          dRs[s] = dR;
          dIs[s] = dI;
        }
        e += pf[g] * (dR * (2.0 * r0[g] + dR) + dI * (2.0 * i0[g] + dI));
      }
    }
  }
  return e;
}

template <typename T, int S>
__device__ __forceinline__ double RecipPhaseLen(const BatchBox &sb,
                                                const CplxT<T> (*tab)[3][kBatchMaxN + 1], int len,
                                                const GMEM int32_t *kn, const GMEM double *pref,
                                                GMEM double *sumR, GMEM double *sumI, GMEM double *dRs,
                                                GMEM double *dIs, int rank, int team, bool pend) {
  switch (len) {
    case 1:
      return RecipPhase<T, S, 1>(sb, tab, kn, pref, sumR, sumI, dRs, dIs, rank, team, pend);
    case 2:
      return RecipPhase<T, S, 2>(sb, tab, kn, pref, sumR, sumI, dRs, dIs, rank, team, pend);
    case 3:
      return RecipPhase<T, S, 3>(sb, tab, kn, pref, sumR, sumI, dRs, dIs, rank, team, pend);
    default:
      return RecipPhase<T, S, 4>(sb, tab, kn, pref, sumR, sumI, dRs, dIs, rank, team, pend);
  }
}

template <typename T>
__device__ __forceinline__ double RecipDispatch(const BatchBox &sb,
                                                const CplxT<T> (*tab)[3][kBatchMaxN + 1], int len,
                                                const GMEM int32_t *kn, const GMEM double *pref,
                                                GMEM double *sumR, GMEM double *sumI, GMEM double *dRs,
                                                GMEM double *dIs, int rank, int team, bool pend) {
  if (sb.kSeg == 8) return RecipPhaseLen<T, 8>(sb, tab, len, kn, pref, sumR, sumI, dRs, dIs, rank, team, pend);
  if (sb.kSeg == 4) return RecipPhaseLen<T, 4>(sb, tab, len, kn, pref, sumR, sumI, dRs, dIs, rank, team, pend);
  if (sb.kSeg == 2) return RecipPhaseLen<T, 2>(sb, tab, len, kn, pref, sumR, sumI, dRs, dIs, rank, team, pend);
  return RecipPhaseLen<T, 1>(sb, tab, len, kn, pref, sumR, sumI, dRs, dIs, rank, team, pend);
}

// Adds a pending accepted change to S(k) at the end of a launch (the
// member's own rounds).
__device__ __noinline__ void RecipFold(const BatchBox &sb, int rank, int team, const GMEM int32_t *kn,
                                       GMEM double *sumR, GMEM double *sumI, const GMEM double *dRs,
                                       const GMEM double *dIs) {
  const int per = sb.kSeg * sb.kWidth;  // slots of a round
  for (int r = rank; r < sb.kChunk; r += team)
    for (int s = r * per + int(threadIdx.x); s < (r + 1) * per; s += kBatchThreads) {
      if (kn[s] >= 0) {
        sumR[s] += dRs[s];
        sumI[s] += dIs[s];
      }
    }
}

// Team exchange, run by wave 0 of each member (acc valid in lane 0).  Each
// partial sum {lj, real, recip, overlap} is published with its 8 lowest
// mantissa bits replaced by the move tag (seq & 255), so a reader knows a
// word is this move's without a separate counter: lane 0 writes the
// member's four tagged words for move `seq`; lane q (q < team, q != rank)
// polls member q's words until all four carry the tag.  Slots alternate
// with the move parity, so a stale word is two moves old and carries
// another tag.  Every member clears the tag bits and sums the same values
// in member order (a partial loses at most 2^-44 of itself), so all
// members hold the same bits.  Words go through device-coherent atomics.  A wait
// that outlasts kTeamSpinLimit polls sets *err and ok = false.
struct Sum4 {
  double v[4];
  bool ok;
};

__device__ __forceinline__ uint64_t Tagged(double v, uint64_t tag) {
  return (__builtin_bit_cast(uint64_t, v) & ~uint64_t(0xff)) | tag;
}

__device__ __noinline__ Sum4 TeamSum(GMEM uint64_t *xb, int team, int rank, uint64_t seq, Sum4 acc,
                                     GMEM int *err) {
  const int lane = int(threadIdx.x) & (kWave - 1);
  const uint64_t tag = seq & 0xff;
  GMEM uint64_t *data = xb + (seq & 1) * kTeamMax * 4;
  uint64_t w[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) w[i] = Tagged(__shfl(acc.v[i], 0), tag);
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
      __hip_atomic_exchange(data + rank * 4 + i, w[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  int ok = 1;
  if (lane < team && lane != rank) {
    GMEM uint64_t *theirs = data + lane * 4;
    for (int64_t n = 0;; ++n) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
        w[i] = __hip_atomic_load(theirs + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if ((w[0] & 0xff) == tag && (w[1] & 0xff) == tag && (w[2] & 0xff) == tag &&
          (w[3] & 0xff) == tag)
        break;
      if (n > kTeamSpinLimit) {
        ok = 0;
        break;
      }
    }
  }
  Sum4 out{{0.0, 0.0, 0.0, 0.0}, __all(ok) != 0};
  if (!out.ok) {
    if (lane == 0) __hip_atomic_exchange(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return out;
  }
  double p[4];  // the tag bits cleared: the partials truncated alike in every member
#pragma unroll
  for (int i = 0; i < 4; ++i) p[i] = __builtin_bit_cast(double, w[i] & ~uint64_t(0xff));
  for (int q = 0; q < team; ++q) {
    out.v[0] += __shfl(p[0], q);
    out.v[1] += __shfl(p[1], q);
    out.v[2] += __shfl(p[2], q);
    out.v[3] = fmax(out.v[3], __shfl(p[3], q));
  }
  return out;
}

// Metropolis test and bookkeeping (Translate/Rotate::Accept,
// src/moves/Translate.h:97-138; MoveSettings::Update,
// src/MoveSettings.cpp:94-146), run by thread 0.
__device__ __noinline__ void Metropolis(const BatchBox &sb, BatchState &st, MoveShared &ms, Sum4 acc,
                                        double pr, int mtype) {
  const bool overlap = acc.v[3] > 0.0;
  const double dE = acc.v[0] + acc.v[1] + acc.v[2];
  const bool res = pr < exp(-sb.beta * dE);
  const bool accept = res && !overlap;
  st.tries[mtype]++;
  st.tempTries[mtype]++;
  if (accept) {
    st.accepted[mtype]++;
    st.tempAccepted[mtype]++;
    st.energy[0] += acc.v[0];
    st.energy[2] += acc.v[1];
    st.energy[3] += acc.v[2];
    st.energy[6] = st.energy[0] + st.energy[1] + st.energy[2] + st.energy[3] + st.energy[4] +
                   st.energy[5];
  }
  ms.accept = accept ? 1 : 0;
}

template <int PREC>
__global__ void __launch_bounds__(kBatchThreads) McBatchKernel(const BatchArgs a) {
  using T = typename Prec<PREC>::T;
  __shared__ BatchBox sb;
  __shared__ double draws[kDrawChunk][8];
  __shared__ CplxT<T> tab[2 * kBatchMaxAtoms][3][kBatchMaxN + 1];
  __shared__ MoveShared ms;
  __shared__ double red[kBatchThreads / kWave][4];
  __shared__ BatchState st;  // running energies and move settings
  __shared__ PairQueue queues[kPairWavesMax];
  __shared__ int nearList[kNearListLen];

  const int tid = threadIdx.x;
  extern __shared__ double ldsCom[];
  for (int ti = a.wgTask[blockIdx.x]; ti < a.wgTask[blockIdx.x + 1]; ++ti) {
    const BatchTask task = a.tasks[ti];
    const int b = __builtin_amdgcn_readfirstlane(task.box);
    const int rank = __builtin_amdgcn_readfirstlane(task.rank);
    {
      const uint64_t *src = reinterpret_cast<const uint64_t *>(&a.boxes[b]);
      uint64_t *dst = reinterpret_cast<uint64_t *>(&sb);
      for (int i = tid; i < int(sizeof(BatchBox) / 8); i += kBatchThreads) dst[i] = src[i];
    }
    if (tid == 0) ms.abort = 0;
    __syncthreads();
    // box constants read from LDS are the same in every lane: kept in
    // scalar registers, with the pointers derived from them
    const int len = __builtin_amdgcn_readfirstlane(sb.molLen);
    const int team = __builtin_amdgcn_readfirstlane(sb.team > 1 ? sb.team : 1);
    // Team members share the box's coordinates and centres.  Every member
    // writes the same accepted values at the same point of a move, after the
    // exchange that needs all members' pair phases of that move, and reads
    // them only after its own write and a barrier: a read sees its own
    // member's write or another's identical one.
    const int64_t aoff = UniformI64(sb.atomOff);
    const int64_t moff = UniformI64(sb.molOff);
    const int64_t koff = UniformI64(sb.kOff);
    double *X = a.x + aoff, *Y = a.y + aoff, *Z = a.z + aoff;
    double *CXg = a.cx + moff, *CYg = a.cy + moff, *CZg = a.cz + moff;
    // The box's molecule centres live in LDS for the launch when they fit
    // (a.ldsMols per workgroup); CXr/CYr/CZr are where the pair phase reads them.
    const bool inLds = __builtin_amdgcn_readfirstlane(sb.nmols) <= a.ldsMols;
    const bool ewald = __builtin_amdgcn_readfirstlane(sb.ewald) != 0;
    const int pairThreads = __builtin_amdgcn_readfirstlane(sb.pairWaves) * kWave;
    double *CXr = CXg, *CYr = CYg, *CZr = CZg;
    if (inLds) {
      CXr = ldsCom;
      CYr = ldsCom + a.ldsMols;
      CZr = ldsCom + 2 * a.ldsMols;
      for (int i = tid; i < sb.nmols; i += kBatchThreads) {
        CXr[i] = CXg[i];
        CYr[i] = CYg[i];
        CZr[i] = CZg[i];
      }
    }
    const int32_t *kn = a.kn + koff;
    const double *pref = a.pref + koff;
    GMEM uint64_t *xb = Gp(a.xchg + int64_t(b) * kXchgWords);

    if (tid == 0) st = a.state[b];
    int64_t step = UniformI64(a.state[b].step);
    const int64_t nMoves = UniformI64(a.moves[b]);
    const int nTab = 2 * len * ((sb.nmax[0] + 1) + (sb.nmax[1] + 1) + (sb.nmax[2] + 1));
    const int64_t traceBox = rank == 0 ? a.trace.box : -1;

    bool pend = false;  // the last move was accepted and S(k) awaits its change
    const uint64_t t0 = wall_clock64();
    for (int64_t mv = 0; mv < nMoves; ++mv, ++step) {
      // ---- Philox draws for the next kDrawChunk moves ----
      if ((mv % kDrawChunk) == 0) {
        if (tid < kDrawChunk) MoveDraws(sb.seed, uint64_t(step + tid), draws[tid]);
        __syncthreads();
      }
      const double *u = draws[mv % kDrawChunk];

      // ---- proposal (thread 0): move settings, move kind, molecule ----
      if (tid == 0) Propose(sb, st, ms, u, step, CXr, CYr, CZr);
      __syncthreads();
      const int m = __builtin_amdgcn_readfirstlane(ms.mol);
      const int mtype = __builtin_amdgcn_readfirstlane(ms.type);
      if (tid < len) {
        const int64_t ai = int64_t(m) * len + tid;
        const double o[3] = {X[ai], Y[ai], Z[ai]};
        double nw[3];
        if (mtype == 0) {
#pragma unroll
          for (int d = 0; d < 3; ++d) nw[d] = Wrap(o[d] + ms.shift[d], sb.L[d]);
        } else {
          double v[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) v[d] = Unwrap(o[d], ms.com[d], sb.L[d], sb.half[d]) + (-ms.com[d]);
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double r = ms.rot[d * 3 + 0] * v[0] + ms.rot[d * 3 + 1] * v[1] + ms.rot[d * 3 + 2] * v[2];
            nw[d] = Wrap(r + ms.com[d], sb.L[d]);
          }
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          ms.oldp[d][tid] = o[d];
          ms.newp[d][tid] = nw[d];
        }
      }
      __syncthreads();

      // ---- phase tables for the 2*len positions ----
      if (ewald) BuildTables<T>(sb, ms, len, nTab, tab);
      __syncthreads();  // tables complete

      // ---- pair phase beside the reciprocal phase ----
      double acc4[4] = {0.0, 0.0, 0.0, 0.0};  // lj, real, recip, overlap
      if (tid < pairThreads) {
        LjReal pr;
        if (inLds)
          pr = PairDispatch<T>(sb, ms, Gp(X), Gp(Y), Gp(Z), Lp(CXr), Lp(CYr), Lp(CZr), m, len,
                               sb.nmols, rank, team, queues, nearList);
        else
          pr = PairDispatch<T>(sb, ms, Gp(X), Gp(Y), Gp(Z), Gp(CXr), Gp(CYr), Gp(CZr), m, len,
                               sb.nmols, rank, team, queues, nearList);
        acc4[0] = pr.lj;
        acc4[1] = pr.real;
        acc4[3] = pr.ov;
      } else if (ewald) {
        acc4[2] = RecipDispatch<T>(sb, tab, len, Gp(kn), Gp(pref), Gp(a.sumR + koff),
                                   Gp(a.sumI + koff), Gp(a.dR + koff), Gp(a.dI + koff), rank, team,
                                   pend);
      }
      BlockSum4(acc4, red);

      // ---- team total (wave 0), then Metropolis (thread 0) ----
      if (team > 1 && tid < kWave) {
        const Sum4 t = TeamSum(xb, team, rank, uint64_t(step) + 1,
                               Sum4{{acc4[0], acc4[1], acc4[2], acc4[3]}, true}, Gp(a.err));
        if (tid == 0) {
          if (!t.ok) ms.abort = 1;
#pragma unroll
          for (int i = 0; i < 4; ++i) acc4[i] = t.v[i];
        }
      }
      if (tid == 0 && !ms.abort) {
        const Sum4 tot{{acc4[0], acc4[1], acc4[2], acc4[3]}, true};
        Metropolis(sb, st, ms, tot, u[5], mtype);
        if (b == traceBox && mv < a.trace.cap) {
          a.trace.acc[mv] = ms.accept;
          a.trace.dlj[mv] = acc4[0];
          a.trace.dreal[mv] = acc4[1];
          a.trace.drecip[mv] = acc4[3] > 0.0 ? 0.0 : acc4[2];
        }
      }
      __syncthreads();
      if (__builtin_amdgcn_readfirstlane(ms.abort)) break;

      // ---- commit (S(k) in the next move's reciprocal phase) ----
      pend = __builtin_amdgcn_readfirstlane(ms.accept) != 0;
      if (pend) {
        if (tid < len) {
          const int64_t ai = int64_t(m) * len + tid;
          X[ai] = ms.newp[0][tid];
          Y[ai] = ms.newp[1][tid];
          Z[ai] = ms.newp[2][tid];
        }
        if (tid == 0 && mtype == 0) {
          CXg[m] = ms.newCom[0];
          CYg[m] = ms.newCom[1];
          CZg[m] = ms.newCom[2];
          if (inLds) {
            CXr[m] = ms.newCom[0];
            CYr[m] = ms.newCom[1];
            CZr[m] = ms.newCom[2];
          }
        }
      }
      // No barrier here: every read of these writes by another thread comes
      // after the next move's first barrier (after the proposal; the trial
      // coordinates are read by the same threads < len that write them
      // here), and nothing before that barrier overwrites what the commit
      // reads (ms.newp is next written after it, ms.newCom by this thread 0).
    }

    if (pend && ewald)
      RecipFold(sb, rank, team, Gp(kn), Gp(a.sumR + koff), Gp(a.sumI + koff), Gp(a.dR + koff),
                Gp(a.dI + koff));
    if (tid == 0 && rank == 0) {
      st.step = step;
      a.state[b] = st;
      a.boxTicks[b] = wall_clock64() - t0;
    }
    __syncthreads();  // LDS is reused by the next task
  }
}

}  // namespace

template <int PREC>
static hipError_t SetLds(int ldsMols) {
  const size_t dyn = sizeof(double) * 3 * size_t(ldsMols);
  return hipFuncSetAttribute(reinterpret_cast<const void *>(&McBatchKernel<PREC>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, int(dyn));
}

template <int PREC>
static hipError_t LaunchOne(int grid, const BatchArgs &a, hipStream_t s, bool setAttr) {
  const size_t dyn = sizeof(double) * 3 * size_t(a.ldsMols);
  if (setAttr) {
    hipError_t e = SetLds<PREC>(a.ldsMols);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(McBatchKernel<PREC>, dim3(grid), dim3(kBatchThreads), dyn, s, a);
  return hipGetLastError();
}

template <int PREC>
static hipError_t ResidentOne(int ldsMols, int *perCU) {
  hipError_t e = SetLds<PREC>(ldsMols);
  if (e != hipSuccess) return e;
  const size_t dyn = sizeof(double) * 3 * size_t(ldsMols);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(perCU, McBatchKernel<PREC>, kBatchThreads,
                                                      dyn);
}

hipError_t BatchResidentPerCU(int precision, int ldsMols, int *perCU) {
  return precision == 1 ? ResidentOne<1>(ldsMols, perCU) : ResidentOne<0>(ldsMols, perCU);
}

hipError_t PrepareBatch(int precision, int ldsMols) {
  return precision == 1 ? SetLds<1>(ldsMols) : SetLds<0>(ldsMols);
}

int BatchLdsMols(int maxMols) { return maxMols < kBatchLdsMolCap ? maxMols : kBatchLdsMolCap; }

hipError_t LaunchBatch(int precision, int grid, const BatchArgs &a, hipStream_t s, bool setAttr) {
  return precision == 1 ? LaunchOne<1>(grid, a, s, setAttr) : LaunchOne<0>(grid, a, s, setAttr);
}

}  // namespace gomc

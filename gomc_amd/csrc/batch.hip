// Batched persistent MC on gfx950: one 1024-thread workgroup owns one box
// and makes all of that box's moves of a launch without leaving the
// kernel.  Per move: Philox draws (made 64 moves ahead, one lane each),
// the trial coordinates (Translate/Rotate, src/Coordinates.cpp:83-128), the
// pair-energy change over every other molecule (MoleculeInter,
// src/CalculateEnergy.cpp:581-686), the structure-factor change per
// k-vector (MolReciprocal, src/Ewald.cpp:409-473), the Metropolis test
// (src/moves/Translate.h:97-138) and, on acceptance, the update of
// coordinates and S(k).  S(k) of the box's k-vectors stays in registers for
// the whole launch (KPT k-vectors per thread), so a move reads only the
// read-only k data and the coordinates.
//
// Phases e^{i k.r} come from per-dimension tables e^{i (c_d n) r_d},
// n = 0..nmax, built each move for the moving molecule's old and new
// positions (one sincos per table entry), and multiplied per k-vector.
#include "batch.hpp"
#include "device_common.hpp"

namespace gomc {
namespace {

struct Cplx {
  double re, im;
};

__device__ __forceinline__ Cplx CMul(Cplx a, Cplx b) {
  return {fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}

struct alignas(16) MoveShared {
  double oldp[3][kBatchMaxAtoms], newp[3][kBatchMaxAtoms];
  double shift[3], rot[9], com[3], newCom[3];
  int mol, type, accept, pad;
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

// Structure-factor change of one k-vector for the moving molecule: sums of
// q e^{i k.r} over its new and old positions, from the phase tables.
__device__ __forceinline__ void KDelta(const Cplx (*tab)[3][kBatchMaxN + 1], const double *charge,
                                       int len, int packed, double &dR, double &dI) {
  const int nx = packed & 0xff;
  const int ny = ((packed >> 8) & 0xff) - 128;
  const int nz = ((packed >> 16) & 0xff) - 128;
  const int ay = ny < 0 ? -ny : ny, az = nz < 0 ? -nz : nz;
  const double sy = ny < 0 ? -1.0 : 1.0, sz = nz < 0 ? -1.0 : 1.0;
  double sRN = 0.0, sIN = 0.0, sRO = 0.0, sIO = 0.0;
  for (int p = 0; p < len; ++p) {
    const double q = charge[p];
    if (fabs(q) < 0.000000001) continue;
    {
      Cplx ex = tab[len + p][0][nx], ey = tab[len + p][1][ay], ez = tab[len + p][2][az];
      ey.im *= sy;
      ez.im *= sz;
      const Cplx e = CMul(CMul(ex, ey), ez);
      sRN += q * e.re;
      sIN += q * e.im;
    }
    {
      Cplx ex = tab[p][0][nx], ey = tab[p][1][ay], ez = tab[p][2][az];
      ey.im *= sy;
      ez.im *= sz;
      const Cplx e = CMul(CMul(ex, ey), ez);
      sRO += q * e.re;
      sIO += q * e.im;
    }
  }
  dR = sRN - sRO;
  dI = sIN - sIO;
}

// Real-space Coulomb term qq erfc(alpha r) / r (FFParticle::CalcCoulomb,
// src/FFParticle.cpp:387-399), out of line to keep the FP64 erfc's
// registers out of the pair loop's budget.
__device__ __noinline__ double CoulombTerm(double d2, double qiqj, double alpha) {
  const double dist = sqrt(d2);
  return qiqj * erfc(alpha * dist) / dist;
}

// One phase-table entry: e^{i (c_d n) r} (kept out of line: the FP64
// sincos argument reduction is register-hungry and runs once per entry).
__device__ __noinline__ Cplx PhaseEntry(double kc, int n, double r) {
  const double ang = (kc * double(n)) * r;
  double s, c;
  sincos(ang, &s, &c);
  return {c, s};
}

// Thread 0's proposal: move settings, move kind, molecule, and the shift or
// rotation matrix (out of line; it runs on one lane).
__device__ __noinline__ void Propose(const BatchBox &sb, BatchState &st, MoveShared &ms,
                                     const double *u, int64_t step, const double *cx,
                                     const double *cy, const double *cz, int64_t moff) {
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
  ms.com[0] = cx[moff + m];
  ms.com[1] = cy[moff + m];
  ms.com[2] = cz[moff + m];
  if (t == 0) {
    const double bb = st.scale[0], b2 = 2.0 * bb;
    ms.shift[0] = u[2] * b2 - bb;
    ms.shift[1] = u[3] * b2 - bb;
    ms.shift[2] = u[4] * b2 - bb;
#pragma unroll
    for (int d = 0; d < 3; ++d) ms.newCom[d] = Wrap(ms.com[d] + ms.shift[d], sb.L[d]);
  } else {
    const double bb = st.scale[1];
    const double theta = u[2] * (2.0 * bb) - bb;
    const double cu = u[3] * 2.0 - 1.0;
    const double phi = u[4] * (2.0 * M_PI);
    const double root = sqrt(1.0 - cu * cu);
    double sp, cp;
    sincos(phi, &sp, &cp);
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

// Philox draws of one move: u[0..5] (two blocks, counter {step, 0|1}).
__device__ __noinline__ void MoveDraws(uint64_t seed, uint64_t step, double *out) {
  const U64x4 r0 = Philox(step, 0, 0, 0, seed, kPhiloxTag);
  const U64x4 r1 = Philox(step, 1, 0, 0, seed, kPhiloxTag);
  out[0] = U01(r0.v[0]);
  out[1] = U01(r0.v[1]);
  out[2] = U01(r0.v[2]);
  out[3] = U01(r0.v[3]);
  out[4] = U01(r1.v[0]);
  out[5] = U01(r1.v[1]);
}

template <int KPT>
__global__ void __launch_bounds__(kBatchThreads) McBatchKernel(const BatchArgs a) {
  __shared__ BatchBox sb;
  __shared__ double draws[kDrawChunk][6];
  __shared__ Cplx tab[2 * kBatchMaxAtoms][3][kBatchMaxN + 1];
  __shared__ MoveShared ms;
  __shared__ double red[kBatchThreads / kWave][4];
  __shared__ BatchState st;  // running energies and move settings

  const int b = blockIdx.x, tid = threadIdx.x;
  {
    const uint64_t *src = reinterpret_cast<const uint64_t *>(&a.boxes[b]);
    uint64_t *dst = reinterpret_cast<uint64_t *>(&sb);
    for (int i = tid; i < int(sizeof(BatchBox) / 8); i += kBatchThreads) dst[i] = src[i];
  }
  __syncthreads();
  const int len = sb.molLen, nmols = sb.nmols, K = sb.K;
  const int64_t aoff = sb.atomOff, moff = sb.molOff, koff = sb.kOff;
  double *X = a.x + aoff, *Y = a.y + aoff, *Z = a.z + aoff;
  const int32_t *kn = a.kn + koff;
  const double *pref = a.pref + koff;

  // S(k) of this thread's k-vectors: k = tid + i * kBatchThreads.
  double R[KPT > 0 ? KPT : 1], I[KPT > 0 ? KPT : 1];
  if (KPT > 0) {
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const int k = tid + i * kBatchThreads;
      R[i] = (k < K) ? a.sumR[koff + k] : 0.0;
      I[i] = (k < K) ? a.sumI[koff + k] : 0.0;
    }
  }

  if (tid == 0) st = a.state[b];
  int64_t step = a.state[b].step;
  const int64_t nMoves = a.moves[b];
  const int nTab = 2 * len * ((sb.nmax[0] + 1) + (sb.nmax[1] + 1) + (sb.nmax[2] + 1));

  for (int64_t mv = 0; mv < nMoves; ++mv, ++step) {
    // ---- Philox draws for the next kDrawChunk moves ----
    if ((mv % kDrawChunk) == 0) {
      if (tid < kDrawChunk) MoveDraws(sb.seed, uint64_t(step + tid), draws[tid]);
      __syncthreads();
    }
    const double *u = draws[mv % kDrawChunk];

    // ---- proposal (thread 0): move settings, move kind, molecule ----
    if (tid == 0) Propose(sb, st, ms, u, step, a.cx, a.cy, a.cz, moff);
    __syncthreads();
    const int m = ms.mol, mtype = ms.type;
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
    if (sb.ewald) {
      for (int e = tid; e < nTab; e += kBatchThreads) {
        int rem = e;
        const int pos = rem / ((sb.nmax[0] + 1) + (sb.nmax[1] + 1) + (sb.nmax[2] + 1));
        rem -= pos * ((sb.nmax[0] + 1) + (sb.nmax[1] + 1) + (sb.nmax[2] + 1));
        int d = 0;
        while (rem > sb.nmax[d]) {
          rem -= sb.nmax[d] + 1;
          ++d;
        }
        const double r = pos < len ? ms.oldp[d][pos] : ms.newp[d][pos - len];
        tab[pos][d][rem] = PhaseEntry(sb.kc[d], rem, r);
      }
    }

    // ---- pair phase: every other molecule, old and new positions ----
    double acc4[4] = {0.0, 0.0, 0.0, 0.0};  // lj, real, recip, overlap
    {
      BoxGeom g;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        g.L[d] = sb.L[d];
        g.half[d] = sb.half[d];
      }
      g.rCutSq = sb.rCutSq;
      g.alpha = sb.alpha;
      g.ewald = sb.ewald;
      for (int j = tid; j < nmols; j += kBatchThreads) {
        if (j == m) continue;
        for (int qa = 0; qa < len; ++qa) {
          const int64_t ai = int64_t(j) * len + qa;
          const double xj = X[ai], yj = Y[ai], zj = Z[ai];
          // set 0: old positions (subtracted), set 1: new positions (added)
          for (int ps = 0; ps < 2 * len; ++ps) {
            const int p = ps < len ? ps : ps - len;
            const bool isNew = ps >= len;
            const int pq = p * kBatchMaxAtoms + qa;
            const double px = isNew ? ms.newp[0][p] : ms.oldp[0][p];
            const double py = isNew ? ms.newp[1][p] : ms.oldp[1][p];
            const double pz = isNew ? ms.newp[2][p] : ms.oldp[2][p];
            const double d2 = DistSq(g, px, py, pz, xj, yj, zj);
            if (sb.rCutSq > d2) {
              if (isNew && d2 < sb.rCutLowSq) acc4[3] = 1.0;
              const double qiqj = sb.qq[pq];
              double ec = 0.0, el = 0.0;
              if (sb.elect && qiqj != 0.0 && !(sb.rCutCoulSq < d2)) ec = sb.ewald ? CoulombTerm(d2, qiqj, sb.alpha) : qiqj / sqrt(d2);
              const double eps = sb.epsCn[pq];
              if (eps != 0.0 && !(sb.rCutLJSq < d2)) {
                // n = 12 (checked at setup): repulsion = attraction^2
                const double rRat2 = sb.sigSq[pq] / d2;
                const double att = rRat2 * rRat2 * rRat2;
                el = eps * (att * att - att);
              }
              if (isNew) {
                acc4[1] += ec;
                acc4[0] += el;
              } else {
                acc4[1] -= ec;
                acc4[0] -= el;
              }
            }
          }
        }
      }
    }
    __syncthreads();  // tables complete

    // ---- reciprocal phase: energy change per k-vector ----
    if (sb.ewald) {
      if (KPT > 0) {
#pragma unroll
        for (int i = 0; i < KPT; ++i) {
          const int k = tid + i * kBatchThreads;
          if (k < K) {
            double dR, dI;
            KDelta(tab, sb.charge, len, kn[k], dR, dI);
            acc4[2] += pref[k] * (dR * (2.0 * R[i] + dR) + dI * (2.0 * I[i] + dI));
          }
        }
      } else {
        for (int k = tid; k < K; k += kBatchThreads) {
          double dR, dI;
          KDelta(tab, sb.charge, len, kn[k], dR, dI);
          const double r0 = a.sumR[koff + k], i0 = a.sumI[koff + k];
          acc4[2] += pref[k] * (dR * (2.0 * r0 + dR) + dI * (2.0 * i0 + dI));
        }
      }
    }
    BlockSum4(acc4, red);

    // ---- Metropolis (thread 0) ----
    if (tid == 0) {
      const bool overlap = acc4[3] > 0.0;
      const double dE = acc4[0] + acc4[1] + acc4[2];
      const bool res = u[5] < exp(-sb.beta * dE);
      const bool accept = res && !overlap;
      st.tries[mtype]++;
      st.tempTries[mtype]++;
      if (accept) {
        st.accepted[mtype]++;
        st.tempAccepted[mtype]++;
        st.energy[0] += acc4[0];
        st.energy[2] += acc4[1];
        st.energy[3] += acc4[2];
        st.energy[6] = st.energy[0] + st.energy[1] + st.energy[2] + st.energy[3] + st.energy[4] +
                       st.energy[5];
      }
      if (b == a.trace.box && mv < a.trace.cap) {
        a.trace.acc[mv] = accept ? 1 : 0;
        a.trace.dlj[mv] = acc4[0];
        a.trace.dreal[mv] = acc4[1];
        a.trace.drecip[mv] = overlap ? 0.0 : acc4[2];
      }
      ms.accept = accept ? 1 : 0;
    }
    __syncthreads();

    // ---- commit ----
    if (ms.accept) {
      if (sb.ewald) {
        if (KPT > 0) {
#pragma unroll
          for (int i = 0; i < KPT; ++i) {
            const int k = tid + i * kBatchThreads;
            if (k < K) {
              double dR, dI;
              KDelta(tab, sb.charge, len, kn[k], dR, dI);
              R[i] += dR;
              I[i] += dI;
            }
          }
        } else {
          for (int k = tid; k < K; k += kBatchThreads) {
            double dR, dI;
            KDelta(tab, sb.charge, len, kn[k], dR, dI);
            a.sumR[koff + k] += dR;
            a.sumI[koff + k] += dI;
          }
        }
      }
      if (tid < len) {
        const int64_t ai = int64_t(m) * len + tid;
        X[ai] = ms.newp[0][tid];
        Y[ai] = ms.newp[1][tid];
        Z[ai] = ms.newp[2][tid];
      }
      if (tid == 0 && mtype == 0) {
        a.cx[moff + m] = ms.newCom[0];
        a.cy[moff + m] = ms.newCom[1];
        a.cz[moff + m] = ms.newCom[2];
      }
    }
    __syncthreads();
  }

  if (KPT > 0) {
#pragma unroll
    for (int i = 0; i < KPT; ++i) {
      const int k = tid + i * kBatchThreads;
      if (k < K) {
        a.sumR[koff + k] = R[i];
        a.sumI[koff + k] = I[i];
      }
    }
  }
  if (tid == 0) {
    st.step = step;
    a.state[b] = st;
  }
}

}  // namespace

static const int kKpts[] = {1, 2, 4};

int BatchMaxKpt() { return 4; }

int BatchRoundKpt(int kpt) {
  for (int v : kKpts)
    if (kpt <= v) return v;
  return 0;  // S(k) in global memory
}

hipError_t LaunchBatch(int kpt, int precision, int nboxes, const BatchArgs &a, hipStream_t s) {
  (void)precision;
  const dim3 grid(nboxes), block(kBatchThreads);
  switch (kpt) {
    case 1: hipLaunchKernelGGL(McBatchKernel<1>, grid, block, 0, s, a); break;
    case 2: hipLaunchKernelGGL(McBatchKernel<2>, grid, block, 0, s, a); break;
    case 4: hipLaunchKernelGGL(McBatchKernel<4>, grid, block, 0, s, a); break;
    default: hipLaunchKernelGGL(McBatchKernel<0>, grid, block, 0, s, a); break;
  }
  return hipGetLastError();
}

}  // namespace gomc

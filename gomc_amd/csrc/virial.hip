// Molecular virial and pressure of a box (CalculateEnergy::VirialCalc,
// src/CalculateEnergy.cpp:411-579; VirialCorrection 1317-1336;
// Ewald::VirialReciprocal, src/Ewald.cpp:1168-1305; the pressure of
// OutputVars.cpp:147-172), diagonal tensor terms as in the reference.
//
// PairVirialKernel: one workgroup per (cell, neighbour slot), the same
// enumeration as BoxInterKernel: for every inter-molecular pair inside the
// box cutoff, F/r of the LJ and erfc-Coulomb terms times the atom
// separation's and the minimum-image centre separation's components.  A
// wave takes (64 cell atoms, 64 neighbour atoms) items and queues their
// in-cutoff pairs in its LDS, so the pair terms run once per 64 queued
// pairs, not once per 64 tested.
// RecipVirialKernel: blocks [0, kBlocks) one k-vector per thread (the
// k-space term); the others one atom per thread over one chunk of LDS-staged
// k data (the intramolecular term: q 2 pref (S_I cos - S_R sin) k_d times
// the atom's offset from its molecule's centre).  Each block writes its partial sums;
// the host adds them in block order.
#include <algorithm>
#include <cmath>

#include "box.hpp"

namespace {
using namespace gomc;

constexpr int kVirSlots = 8;

struct PairVirArgs {
  BoxGeom g;
  const double *x, *y, *z, *q;
  const int *type, *mol;
  const double *sigSq, *epsCn6, *nOver6, *nHalf;
  int ntypes;
  const int *cellStart, *cellVec, *nbr;
  const double *cx, *cy, *cz;  // molecule centres
  double *part;                // [blocks][kVirSlots]
};

constexpr int kUnroll = 4;  // neighbour entries tested per step of the pair virial

// 64 atoms of a cell staged in a wave's LDS, with their molecules' centres.
struct VirTile {
  double x[kWave], y[kWave], z[kWave], q[kWave], cx[kWave], cy[kWave], cz[kWave];
  int idx[kWave], type[kWave], mol[kWave];
};

__device__ __forceinline__ void StageVirTile(const PairVirArgs &a, VirTile &t, int first, int n,
                                             int lane) {
  if (lane < n) {
    const int j = a.cellVec[first + lane], m = a.mol[j];
    t.idx[lane] = j;
    t.x[lane] = a.x[j];
    t.y[lane] = a.y[j];
    t.z[lane] = a.z[j];
    t.q[lane] = a.q[j];
    t.type[lane] = a.type[j];
    t.mol[lane] = m;
    t.cx[lane] = a.cx[m];
    t.cy[lane] = a.cy[m];
    t.cz[lane] = a.cz[m];
  }
}

// The virial terms of one in-cutoff pair (entry u of tile I, entry k of
// tile J) added to v: FFParticle::CalcCoulombVir and CalcVir
// (FFParticle.cpp:295-359, 401-446) times the atom and centre separations'
// components, the Coulomb term with erfc(a r) = exp(-(a r)^2) erfcx(a r)
// where the reference writes 1 - erf.
__device__ __forceinline__ void PairVirial(const PairVirArgs &a, const double *poly,
                                           const VirTile &I, int u, const VirTile &J, int k,
                                           double v[6]) {
  const double dx = MinImage(I.x[u] - J.x[k], a.g.L[0], a.g.half[0]);
  const double dy = MinImage(I.y[u] - J.y[k], a.g.L[1], a.g.half[1]);
  const double dz = MinImage(I.z[u] - J.z[k], a.g.L[2], a.g.half[2]);
  const double d2 = dx * dx + dy * dy + dz * dz;
  const double cdx = MinImage(I.cx[u] - J.cx[k], a.g.L[0], a.g.half[0]);
  const double cdy = MinImage(I.cy[u] - J.cy[k], a.g.L[1], a.g.half[1]);
  const double cdz = MinImage(I.cz[u] - J.cz[k], a.g.L[2], a.g.half[2]);
  const int idx = I.type[u] + J.type[k] * a.ntypes;
  const PairTerms t = PairMath(a.g, poly, d2, I.q[u] * J.q[k], a.sigSq[idx], 0.0, a.nHalf[idx],
                               a.epsCn6[idx], a.nOver6[idx], false);
  v[0] += t.vLJ * (dx * cdx);
  v[1] += t.vLJ * (dy * cdy);
  v[2] += t.vLJ * (dz * cdz);
  v[3] += t.vReal * (dx * cdx);
  v[4] += t.vReal * (dy * cdy);
  v[5] += t.vReal * (dz * cdz);
}

__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The block's (cell chunk of 64, neighbour tile of 64) items go to its
// waves round robin.  A wave stages both tiles in its LDS, each lane holds
// one cell atom and walks the neighbour tile kUnroll entries at a time, and
// the in-cutoff pairs are queued; the pair terms run once per 64 queued
// pairs (and once for an item's rest).  The sums' order is fixed by the
// positions.
__global__ void __launch_bounds__(kBlock, 4) PairVirialKernel(const PairVirArgs a) {
  constexpr int kWaves = kBlock / kWave;
  __shared__ VirTile tileI[kWaves], tileJ[kWaves];
  __shared__ int pairList[kWaves][(kUnroll + 1) * kWave];
  __shared__ double red[kWaves];
  __shared__ double poly[kErfcxDeg + 1];
  if (threadIdx.x <= kErfcxDeg) poly[threadIdx.x] = kErfcxPoly[threadIdx.x];
  __syncthreads();
  const int cell = blockIdx.x / 27, nb = a.nbr[blockIdx.x];
  const int cs = a.cellStart[cell], ce = a.cellStart[cell + 1];
  const int ns = a.cellStart[nb], ne = a.cellStart[nb + 1];
  const int ic = (ce - cs + kWave - 1) / kWave, jc = (ne - ns + kWave - 1) / kWave;
  const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
  VirTile &I = tileI[w], &J = tileJ[w];
  int *q = pairList[w];
  const uint64_t below = (uint64_t(1) << lane) - 1;
  double v[6] = {0, 0, 0, 0, 0, 0};  // LJ xx yy zz, real xx yy zz
  for (int item = w; item < ic * jc; item += kWaves) {
    const int i0 = cs + (item / jc) * kWave, j0 = ns + (item % jc) * kWave;
    const int ni = min(kWave, ce - i0), nj = min(kWave, ne - j0);
    WaveSync();  // the previous item's reads are done
    StageVirTile(a, I, i0, ni, lane);
    StageVirTile(a, J, j0, nj, lane);
    WaveSync();
    const bool iv = lane < ni;
    const int i = iv ? I.idx[lane] : 0, mi = iv ? I.mol[lane] : -1;
    const double xi = iv ? I.x[lane] : 0.0, yi = iv ? I.y[lane] : 0.0, zi = iv ? I.z[lane] : 0.0;
    // kUnroll tile entries per step, each lane testing its atom against
    // them; the wave queues the in-cutoff pairs (i < j, other molecule)
    int queued = 0;
    for (int k0 = 0; k0 < nj || queued > 0; k0 += kUnroll) {
      if (k0 < nj) {
        bool hit[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const int k = k0 + u;
          hit[u] = false;
          if (iv && k < nj && i < J.idx[k] && mi != J.mol[k]) {
            const double dx = MinImage(xi - J.x[k], a.g.L[0], a.g.half[0]);
            const double dy = MinImage(yi - J.y[k], a.g.L[1], a.g.half[1]);
            const double dz = MinImage(zi - J.z[k], a.g.L[2], a.g.half[2]);
            hit[u] = a.g.rCutSq > dx * dx + dy * dy + dz * dz;
          }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
          const uint64_t m = __ballot(hit[u]);
          if (hit[u]) q[queued + __popcll(m & below)] = (lane << 6) | (k0 + u);
          queued += __popcll(m);
        }
      }
      // full batches, or the item's last pairs
      while (queued >= kWave || (queued > 0 && k0 + kUnroll >= nj)) {
        WaveSync();
        const int n = min(queued, kWave);
        if (lane < n) {
          const int e = q[queued - n + lane];
          PairVirial(a, poly, I, e >> 6, J, e & 63, v);
        }
        queued -= n;
        WaveSync();
      }
    }
  }
  for (int c = 0; c < 6; ++c) {
    const double s = BlockSum(v[c], red);
    if (threadIdx.x == 0) a.part[size_t(blockIdx.x) * kVirSlots + c] = s;
  }
}

// Segments of the k list (consecutive nz, at most kSeg vectors) staged in
// LDS this many at a time by the intramolecular term.
constexpr int kVirSegTile = 32;

struct RecipVirArgs {
  BoxGeom g;
  const double *x, *y, *z, *q;
  const int *mol;
  int natoms, K, kBlocks;
  int segChunk, kChunks;  // segments per chunk, chunks per atom block
  const int *kseg;        // segment starts [nseg + 1]
  int nseg;
  double kzStep;          // the nz step of every segment, 2 pi / L_z
  const double *kx, *ky, *kz, *pref, *sumR, *sumI;
  const double *cx, *cy, *cz;
  double constVal;  // 1 / (4 alpha^2)
  double *part;
};

__global__ void __launch_bounds__(kBlock) RecipVirialKernel(const RecipVirArgs a) {
  __shared__ double red[kBlock / kWave];
  __shared__ double2 sk[3][kVirSegTile * kSeg];  // (pref S_R k_d, pref S_I k_d) per k
  __shared__ double sf[3][kVirSegTile];           // each segment's first k-vector
  __shared__ int ss[kVirSegTile + 1];
  double w[3] = {0, 0, 0};
  if (int(blockIdx.x) < a.kBlocks) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < a.K) {
      const double kx = a.kx[i], ky = a.ky[i], kz = a.kz[i];
      const double hsqr = kx * kx + ky * ky + kz * kz;
      const double factor = a.pref[i] * (a.sumR[i] * a.sumR[i] + a.sumI[i] * a.sumI[i]);
      w[0] = factor * (1.0 - 2.0 * (a.constVal + 1.0 / hsqr) * kx * kx);
      w[1] = factor * (1.0 - 2.0 * (a.constVal + 1.0 / hsqr) * ky * ky);
      w[2] = factor * (1.0 - 2.0 * (a.constVal + 1.0 / hsqr) * kz * kz);
    }
  } else {
    const int p = (blockIdx.x - a.kBlocks) / a.kChunks * kBlock + threadIdx.x;
    bool on = false;
    double xp = 0, yp = 0, zp = 0, qp = 0, ddx = 0, ddy = 0, ddz = 0;
    if (p < a.natoms) {
      qp = a.q[p];
      const int m = a.mol[p];
      on = m >= 0 && !(fabs(qp) < 0.000000001);
      if (on) {
        xp = a.x[p];
        yp = a.y[p];
        zp = a.z[p];
        const double cx = a.cx[m], cy = a.cy[m], cz = a.cz[m];
        ddx = Unwrap(xp, cx, a.g.L[0], a.g.half[0]) - cx;
        ddy = Unwrap(yp, cy, a.g.L[1], a.g.half[1]) - cy;
        ddz = Unwrap(zp, cz, a.g.L[2], a.g.half[2]) - cz;
      }
    }
    // this chunk's segments, kVirSegTile at a time: one sincos per run of
    // consecutive nz, for its first phase (the nz step's phase is the
    // atom's own), then one complex product and six FMAs per k-vector
    // against pref S_R k and pref S_I k; 2 q times the offset at the end
    const int chunk = (blockIdx.x - a.kBlocks) % a.kChunks;
    const int s0 = min(a.nseg, chunk * a.segChunk), s1 = min(a.nseg, s0 + a.segChunk);
    double sn = 0.0, cn = 1.0;  // the atom's phase step e^{i c_z z}, every segment's
    if (on) sincos(a.kzStep * zp, &sn, &cn);
    double W[3] = {0.0, 0.0, 0.0};
    // a segment that continues the previous one's run (same nx, ny after a
    // full segment) starts at the phase the previous one rotated to
    double s = 0.0, c = 1.0, pkx = 0.0, pky = 0.0;
    int plen = 0;
    for (int t = s0; t < s1; t += kVirSegTile) {
      const int tn = min(kVirSegTile, s1 - t);
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
            sincos(xp * kx0 + yp * ky0 + zp * sf[2][j], &s, &c);
          pkx = kx0;
          pky = ky0;
          plen = len;
#pragma unroll
          for (int i = 0; i < kSeg; ++i) {
            if (i < len) {
              const int k = k0 + i;
              const double2 vx = sk[0][k], vy = sk[1][k], vz = sk[2][k];
              W[0] = fma(c, vx.y, W[0]);
              W[0] = fma(-s, vx.x, W[0]);
              W[1] = fma(c, vy.y, W[1]);
              W[1] = fma(-s, vy.x, W[1]);
              W[2] = fma(c, vz.y, W[2]);
              W[2] = fma(-s, vz.x, W[2]);
            }
            const double nc = fma(c, cn, -(s * sn));
            s = fma(c, sn, s * cn);
            c = nc;
          }
        }
    }
    const double q2 = 2.0 * qp;
    w[0] = q2 * (W[0] * ddx);
    w[1] = q2 * (W[1] * ddy);
    w[2] = q2 * (W[2] * ddz);
  }
  for (int c = 0; c < 3; ++c) {
    const double s = BlockSum(w[c], red);
    if (threadIdx.x == 0) a.part[size_t(blockIdx.x) * kVirSlots + c] = s;
  }
}

}  // namespace

extern "C" int gomc_box_virial(gomc_box *b, double out[9]) {
  if (!b || !out) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = FlushPending(b);
  if (!rc) rc = MPEnsure(b);
  if (!rc) rc = MPUploadCom(b);  // molecule centres of the current coordinates
  if (!rc) rc = BuildCells(b);
  if (rc) return rc;
  const int pairBlocks = b->ncell * 27;
  const int kBlocks = b->ewald ? (b->K + kBlock - 1) / kBlock : 0;
  const int atomBlocks = b->ewald ? (b->natoms + kBlock - 1) / kBlock : 0;
  // The intramolecular term splits the k list's segments into chunks as
  // well as the atoms into blocks, so that a large box fills the chip (about
  // 2048 workgroups).
  int kChunks = 1, segChunk = std::max(b->nseg, 1);
  if (atomBlocks > 0) {
    if (b->nseg <= 0) return Fail(b, GOMC_ERR_STATE, "k-vector segments not built");
    kChunks = std::min(b->nseg, std::max(1, (2048 + atomBlocks - 1) / atomBlocks));
    segChunk = (b->nseg + kChunks - 1) / kChunks;
    kChunks = (b->nseg + segChunk - 1) / segChunk;
  }
  const int nblocks = pairBlocks + kBlocks + atomBlocks * kChunks;
  const size_t need = size_t(kVirSlots) * std::max(nblocks, 1);
  if (need > b->virCap) {  // grows only with the box's block count
    if (b->virBuf) (void)hipFree(b->virBuf);
    b->virBuf = nullptr;
    b->virCap = 0;
    HIP_TRY(b, hipMalloc(&b->virBuf, sizeof(double) * need));
    b->virCap = need;
  }
  double *part = b->virBuf;
  PairVirArgs a;
  a.g = b->geom;
  a.x = b->x;
  a.y = b->y;
  a.z = b->z;
  a.q = b->q;
  a.type = b->type;
  a.mol = b->mol;
  a.sigSq = b->tSigSq;
  a.epsCn6 = b->tEpsCn6;
  a.nOver6 = b->tNOver6;
  a.nHalf = b->tNHalf;
  a.ntypes = b->ntypes;
  a.cellStart = b->cellStart;
  a.cellVec = b->cellVec;
  a.nbr = b->nbr;
  a.cx = b->mp.com[0][0];
  a.cy = b->mp.com[0][1];
  a.cz = b->mp.com[0][2];
  a.part = part;
  hipLaunchKernelGGL(PairVirialKernel, dim3(pairBlocks), dim3(kBlock), 0, b->stream, a);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess && b->ewald) {
    RecipVirArgs r;
    r.g = b->geom;
    r.x = b->x;
    r.y = b->y;
    r.z = b->z;
    r.q = b->q;
    r.mol = b->mol;
    r.natoms = b->natoms;
    r.K = b->K;
    r.kBlocks = kBlocks;
    r.segChunk = segChunk;
    r.kChunks = kChunks;
    r.kseg = b->kseg;
    r.nseg = b->nseg;
    r.kzStep = (1.0 / b->geom.L[2]) * (2.0 * M_PI);  // KVectorsOrth's z spacing
    r.kx = b->kx;
    r.ky = b->ky;
    r.kz = b->kz;
    r.pref = b->pref;
    r.sumR = b->sumRref;
    r.sumI = b->sumIref;
    r.cx = a.cx;
    r.cy = a.cy;
    r.cz = a.cz;
    r.constVal = 1.0 / (4.0 * (b->alpha * b->alpha));
    r.part = part + size_t(pairBlocks) * kVirSlots;
    hipLaunchKernelGGL(RecipVirialKernel, dim3(kBlocks + atomBlocks * kChunks), dim3(kBlock), 0,
                       b->stream, r);
    e = hipGetLastError();
  }
  std::vector<double> h(size_t(kVirSlots) * std::max(nblocks, 1), 0.0);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), part, sizeof(double) * kVirSlots * nblocks, hipMemcpyDeviceToHost,
                       b->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return HipFail(b, e, "virial kernels");
  double vT[3] = {0, 0, 0}, rT[3] = {0, 0, 0}, wT[3] = {0, 0, 0};
  for (int blk = 0; blk < pairBlocks; ++blk)
    for (int d = 0; d < 3; ++d) {
      vT[d] += h[size_t(blk) * kVirSlots + d];
      rT[d] += h[size_t(blk) * kVirSlots + 3 + d];
    }
  for (int blk = pairBlocks; blk < nblocks; ++blk)
    for (int d = 0; d < 3; ++d) wT[d] += h[size_t(blk) * kVirSlots + d];
  const double inter = vT[0] + vT[1] + vT[2];
  const double real = b->elect ? (rT[0] + rT[1] + rT[2]) * kQQFact : 0.0;
  double tail = 0.0;  // VirialCorrection
  if (b->lrc)
    for (int i = 0; i < b->nkinds; ++i)
      for (int j = 0; j < b->nkinds; ++j)
        tail += b->pairVirCorr[i * b->nkinds + j] * unsigned(b->kindCount[i]) *
                unsigned(b->kindCount[j]) * b->volInv;
  const double recip = wT[0] + wT[1] + wT[2];
  out[0] = inter;
  out[1] = tail;
  out[2] = real;
  out[3] = recip;
  out[4] = inter + tail + real + recip;
  // the pressure: each term over the dimensionality and the volume, the
  // ideal part from the number densities, K molecule / A^3 to bar
  const double V = b->volume;
  const double total = inter / 3.0 / V + tail / 3.0 / V + real / 3.0 / V + recip / 3.0 / V;
  double raw = 0.0;
  for (int k = 0; k < b->nkinds; ++k) raw += b->kindCount[k] * b->volInv;
  raw *= b->T;
  raw += total;
  out[5] = raw * 138.0649;
  for (int d = 0; d < 3; ++d) out[6 + d] = vT[d] + (b->elect ? rT[d] * kQQFact : 0.0) + wT[d];
  return GOMC_OK;
}

// Device helpers shared by the gfx950 kernels: minimum image, the Mie/LJ
// and erfc-Coulomb pair terms, Philox4x64-10, and wave64/workgroup
// reductions in a fixed order (so a sum does not depend on timing).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "erfcx_poly.hpp"

namespace gomc {

constexpr double kQQFact = 167103.208067979;  // lib/NumLib.h:22
constexpr int kWave = 64;

// BoxDimensions::MinImageSigned (src/BoxDimensions.h:169-175): the same
// single add or subtract, chosen by selects rather than branches.
__device__ __forceinline__ double MinImage(double d, double ax, double half) {
  const double down = d - ax, up = d + ax;
  return d > half ? down : (d < -half ? up : d);
}

// BoxDimensions::WrapPBC (src/BoxDimensions.cpp:261-296).
__host__ __device__ __forceinline__ double Wrap(double v, double ax) {
  if (v >= ax)
    v -= ax;
  else if (v < 0)
    v += ax;
  return v;
}

// BoxDimensions::UnwrapPBC (src/BoxDimensions.cpp:298-320).
__host__ __device__ __forceinline__ double Unwrap(double v, double ref, double ax,
                                                  double half) {
  if (fabs(ref - v) > half) {
    if (ref < half)
      v -= ax;
    else
      v += ax;
  }
  return v;
}

// Orthogonal box geometry and cutoffs passed by value to every kernel.
struct BoxGeom {
  double L[3], half[3];
  double rCutSq;      // box cutoff max(Rcut, RcutCoulomb) squared
  double rCutLJSq;    // Rcut squared (FFParticle::CalcEn gate)
  double rCutCoulSq;  // RcutCoulomb squared (FFParticle::CalcCoulomb gate)
  double rCutLowSq;   // RcutLow squared (overlap test)
  double alpha;       // Ewald alpha
  int ewald, elect;
};

// Squared minimum-image distance in the reference's operation order
// (InRcut, src/BoxDimensions.h:231-261): each component min-imaged, then
// x*x + y*y + z*z left to right.  Files are built without FMA contraction.
__device__ __forceinline__ double DistSq(const BoxGeom &g, double ax, double ay, double az,
                                         double bx, double by, double bz) {
  double dx = MinImage(ax - bx, g.L[0], g.half[0]);
  double dy = MinImage(ay - by, g.L[1], g.half[1]);
  double dz = MinImage(az - bz, g.L[2], g.half[2]);
  return dx * dx + dy * dy + dz * dz;
}

// FFParticle::CalcEn (src/FFParticle.cpp:295-325, lambda = 1):
// eps_cn * ((sigma^2/r^2)^(n/2) - (sigma^2/r^2)^3).  For n = 12 the
// repulsion is the attraction squared (pow(x, 6) within an ulp or two).
__device__ __forceinline__ double LJEnergy(double distSq, double sigmaSq, double epsCn,
                                           double nHalf) {
  double rRat2 = sigmaSq / distSq;
  double rRat4 = rRat2 * rRat2;
  double attract = rRat4 * rRat2;
  double repulse = (nHalf == 6.0) ? attract * attract : pow(rRat2, nHalf);
  return epsCn * (repulse - attract);
}

// FFParticle::CalcVir (src/FFParticle.cpp:327-359): F.r / r^2.
__device__ __forceinline__ double LJVirial(double distSq, double sigmaSq, double epsCn6,
                                           double nOver6, double nHalf) {
  double rNeg2 = 1.0 / distSq;
  double rRat2 = rNeg2 * sigmaSq;
  double rRat4 = rRat2 * rRat2;
  double attract = rRat4 * rRat2;
  double repulse = (nHalf == 6.0) ? attract * attract : pow(rRat2, nHalf);
  return epsCn6 * (nOver6 * repulse - attract) * rNeg2;
}

// FFParticle::CalcCoulomb (src/FFParticle.cpp:387-399).
__device__ __forceinline__ double CoulombEnergy(const BoxGeom &g, double distSq,
                                                double qiqj) {
  double dist = sqrt(distSq);
  if (g.ewald) return qiqj * erfc(g.alpha * dist) / dist;
  return qiqj / dist;
}

// Mixed precision (FP32 pair terms and phases, FP64 accumulation): the
// same terms in FP32 from the FP64 squared distance.
__device__ __forceinline__ double CoulombEnergyF32(const BoxGeom &g, double distSq, double qiqj) {
  const float d = sqrtf(float(distSq));
  if (g.ewald) return double(float(qiqj) * erfcf(float(g.alpha) * d) / d);
  return double(float(qiqj) / d);
}

__device__ __forceinline__ double LJEnergyF32(double distSq, double sigmaSq, double epsCn,
                                              double nHalf) {
  const float rRat2 = float(sigmaSq) / float(distSq);
  const float att = rRat2 * rRat2 * rRat2;
  const float rep = (nHalf == 6.0) ? att * att : powf(rRat2, float(nHalf));
  return double(float(epsCn) * (rep - att));
}

// sincos(ang) in FP32 after reducing the FP64 argument to [-pi, pi).
__device__ __forceinline__ void SincosF32(double ang, double *s, double *c) {
  const double turns = ang * (0.5 * M_1_PI);
  const float red = float((turns - rint(turns)) * (2.0 * M_PI));
  float sf, cf;
  sincosf(red, &sf, &cf);
  *s = sf;
  *c = cf;
}

// FFParticle::CalcCoulombVir (src/FFParticle.cpp:432-446).
__device__ __forceinline__ double CoulombVirial(const BoxGeom &g, double distSq,
                                                double qiqj) {
  double dist = sqrt(distSq);
  if (g.ewald) {
    double constValue = g.alpha * M_2_SQRTPI;
    double expConst = exp(-1.0 * (g.alpha * g.alpha) * distSq);
    double temp = 1.0 - erf(g.alpha * dist);
    return qiqj * (temp / dist + constValue * expConst) / distSq;
  }
  return qiqj / (distSq * dist);
}

// ---------------------------------------------------------- Philox -------
// Philox4x64-10 (Salmon et al. 2011; the generator the reference vendors as
// lib/Random123/philox.h): 10 rounds of two 64x64->128 multiplies with the
// published multipliers, key bumped by the Weyl constants between rounds.
struct U64x4 {
  uint64_t v[4];
};

__host__ __device__ __forceinline__ void MulHiLo(uint64_t a, uint64_t b, uint64_t &hi,
                                                 uint64_t &lo) {
#if defined(__HIP_DEVICE_COMPILE__)
  lo = a * b;
  hi = __umul64hi(a, b);
#else
  unsigned __int128 p = (unsigned __int128)a * b;
  hi = uint64_t(p >> 64);
  lo = uint64_t(p);
#endif
}

__host__ __device__ __forceinline__ U64x4 Philox(uint64_t c0, uint64_t c1, uint64_t c2,
                                                 uint64_t c3, uint64_t k0, uint64_t k1) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t hi0, lo0, hi1, lo1;
    MulHiLo(M0, c0, hi0, lo0);
    MulHiLo(M1, c2, hi1, lo1);
    uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
  }
  U64x4 o;
  o.v[0] = c0;
  o.v[1] = c1;
  o.v[2] = c2;
  o.v[3] = c3;
  return o;
}

// r123::u01<double> for a 64-bit word (lib/Random123/uniform.hpp:174-183):
// x * 2^-64 + 2^-65, the product exact, one rounding in the add.
__host__ __device__ __forceinline__ double U01(uint64_t v) {
  return double(v) * 0x1p-64 + 0x1p-65;
}

// erfc for 0 <= x <= 4 as exp(-x^2) erfcx(x), erfcx from the fitted
// polynomial of erfcx_poly.hpp (relative error about 3e-15); libm beyond.
__device__ __forceinline__ double ErfcPoly(double x) {
  if (!(x <= kErfcxUse)) return erfc(x);
  const double t = (x - kErfcxMid) * kErfcxInvHalf;
  double p = kErfcxPoly[kErfcxDeg];
#pragma unroll
  for (int k = kErfcxDeg - 1; k >= 0; --k) p = fma(p, t, kErfcxPoly[k]);
  return exp(-(x * x)) * p;
}

// erfcx(x) = exp(x^2) erfc(x) for x >= 4 from Laplace's continued fraction
// erfc(x) = exp(-x^2) / sqrt(pi) / (x + (1/2) / (x + 1 / (x + (3/2) / ...))),
// evaluated from the tail, 24 levels: relative error about 4e-16 for every
// x >= 4.  Two live values, where libm's erfc holds dozens of registers.
__device__ __forceinline__ double ErfcxTail(double x) {
  double f = x;
#pragma unroll 1
  for (int n = 24; n >= 1; --n) f = x + (0.5 * n) / f;
  return 1.0 / (1.7724538509055160273 * f);
}

// x^h for the Mie repulsion (sigma^2/r^2)^(n/2): whole and half-integer h
// up to 32 by repeated squaring (and one sqrt), within a few ulps of pow;
// any other h as exp(h log x).
__device__ __forceinline__ double MiePow(double x, double h) {
  const double twice = 2.0 * h;
  if (h > 0.0 && h <= 32.0 && twice == floor(twice)) {
    int k = int(h);
    double r = 1.0, b = x;
    while (k) {
      if (k & 1) r *= b;
      b *= b;
      k >>= 1;
    }
    return twice == 2.0 * floor(h) ? r : r * sqrt(x);
  }
  return exp(h * log(x));
}

// 1/sqrt(d2) to double precision: the hardware estimate and two Newton
// steps y <- y (3/2 - d2 y^2 / 2).
__device__ __forceinline__ double Rsqrt64(double d2) {
  const double h = 0.5 * d2;
  double y = __builtin_amdgcn_rsq(d2);
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}

// The energies (counted when `count`) and virial factors of one in-cutoff
// pair: FFParticle::CalcCoulomb / CalcCoulombVir (src/FFParticle.cpp:
// 387-399, 432-446) and CalcEn / CalcVir (295-359).  erfc(a r) is
// exp(-(a r)^2) erfcx(a r) with erfcx from the fitted polynomial of
// erfcx_poly.hpp (coefficients `poly`, in LDS), and the same exponential is the virial's Gaussian: one
// exp and one polynomial per pair instead of libm's erfc, erf and exp, whose
// range branches split a wave.  (1 - erf is erfc up to rounding: the
// reference's form differs from this one by about 1e-16 absolute.)
struct PairTerms {
  double eLJ, eReal, vLJ, vReal;
};
__device__ __forceinline__ PairTerms PairMath(const BoxGeom &g, const double *poly, double d2,
                                              double qiqj, double sig, double eps, double nh,
                                              double eps6, double nov6, bool count) {
  PairTerms p{0.0, 0.0, 0.0, 0.0};
  if (g.elect && qiqj != 0.0 && !(g.rCutCoulSq < d2)) {
    const double dist = sqrt(d2);
    if (g.ewald) {
      const double x = g.alpha * dist;
      double ec, gauss;
      if (x <= kErfcxUse) {
        const double t = (x - kErfcxMid) * kErfcxInvHalf;
        double px = poly[kErfcxDeg];
        // read from LDS each step, eight steps' reads issued together: fully
        // unrolled, the coefficients would be held in 58 VGPRs across the
        // pair loop
#pragma unroll 8
        for (int k = kErfcxDeg - 1; k >= 0; --k) px = fma(px, t, poly[k]);
        gauss = exp(-(x * x));
        ec = gauss * px;
      } else {
        gauss = exp(-(x * x));
        ec = gauss * ErfcxTail(x);
      }
      if (count) p.eReal = qiqj * ec / dist;
      p.vReal = qiqj * (ec / dist + g.alpha * M_2_SQRTPI * gauss) / d2;
    } else {
      if (count) p.eReal = qiqj / dist;
      p.vReal = qiqj / (d2 * dist);
    }
  }
  if (!(g.rCutLJSq < d2)) {
    // FFParticle::CalcEn and CalcVir in their operation orders, the
    // repulsion by MiePow rather than libm's pow
    if (count) {
      const double rRat2 = sig / d2;
      const double attract = rRat2 * rRat2 * rRat2;
      const double repulse = nh == 6.0 ? attract * attract : MiePow(rRat2, nh);
      p.eLJ = eps * (repulse - attract);
    }
    const double rNeg2 = 1.0 / d2;
    const double rRat2 = rNeg2 * sig;
    const double attract = rRat2 * rRat2 * rRat2;
    const double repulse = nh == 6.0 ? attract * attract : MiePow(rRat2, nh);
    p.vLJ = eps6 * (nov6 * repulse - attract) * rNeg2;
  }
  return p;
}

// The pair terms and phases at precision P: 0 is FP64 throughout, 1 the
// FP32 forms above (sums stay FP64 in every caller).
template <int P>
__device__ __forceinline__ double Coul(const BoxGeom &g, double d2, double qiqj) {
  return P ? CoulombEnergyF32(g, d2, qiqj) : CoulombEnergy(g, d2, qiqj);
}
template <int P>
__device__ __forceinline__ double LJ(double d2, double sig, double eps, double nh) {
  return P ? LJEnergyF32(d2, sig, eps, nh) : LJEnergy(d2, sig, eps, nh);
}
template <int P>
__device__ __forceinline__ void Sincos(double ang, double *s, double *c) {
  if (P)
    SincosF32(ang, s, c);
  else
    sincos(ang, s, c);
}

// ------------------------------------------------------- reductions -------
template <typename T>
__device__ __forceinline__ T WaveSum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;  // valid in lane 0
}

// Sum over the workgroup; the result is valid in thread 0.  `lds` holds one
// slot per wave.  Order: lanes by tree within a wave, then waves in index
// order.  Contains __syncthreads().
template <typename T>
__device__ __forceinline__ T BlockSum(T v, T *lds) {
  const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
  const int nwaves = (blockDim.x + kWave - 1) / kWave;
  v = WaveSum(v);
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  T r = T(0);
  if (threadIdx.x == 0)
    for (int w = 0; w < nwaves; ++w) r += lds[w];
  __syncthreads();
  return r;
}

}  // namespace gomc

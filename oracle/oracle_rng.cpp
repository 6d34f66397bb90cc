// CPU oracle -- random number generation.  TEST INFRASTRUCTURE ONLY (see
// oracle.hpp).
#include "oracle.hpp"

namespace orc {

int g_plant = 0;


// src/MersenneTwister.h:194-207 (initialize) then 226-231 (seed -> reload).
void MTRand::seed(uint32_t s) {
  state_[0] = s;
  for (int i = 1; i < N; ++i)
    state_[i] = 1812433253u * (state_[i - 1] ^ (state_[i - 1] >> 30)) + uint32_t(i);
  reload();
}

// src/MersenneTwister.h:209-224: standard MT19937 generation of N words.
void MTRand::reload() {
  auto twist = [](uint32_t m, uint32_t s0, uint32_t s1) {
    uint32_t y = (s0 & 0x80000000u) | (s1 & 0x7fffffffu);
    return m ^ (y >> 1) ^ ((s1 & 1u) ? 0x9908b0dfu : 0u);
  };
  for (int i = 0; i < N - M; ++i) state_[i] = twist(state_[i + M], state_[i], state_[i + 1]);
  for (int i = N - M; i < N - 1; ++i)
    state_[i] = twist(state_[i + M - N], state_[i], state_[i + 1]);
  state_[N - 1] = twist(state_[M - 1], state_[N - 1], state_[0]);
  left_ = N;
  next_ = state_;
}

// src/MersenneTwister.h:320-331.
uint32_t MTRand::randInt() {
  if (left_ == 0) reload();
  --left_;
  ++draws;
  uint32_t s1 = *next_++;
  s1 ^= (s1 >> 11);
  s1 ^= (s1 << 7) & 0x9d2c5680u;
  s1 ^= (s1 << 15) & 0xefc60000u;
  return s1 ^ (s1 >> 18);
}

// src/MersenneTwister.h:333-350: draw in [0,n] by masking to the bit width
// of n and rejecting values above n.
uint32_t MTRand::randInt(uint32_t n) {
  uint32_t used = n;
  used |= used >> 1;
  used |= used >> 2;
  used |= used >> 4;
  used |= used >> 8;
  used |= used >> 16;
  uint32_t i;
  do {
    i = randInt() & used;
  } while (i > n);
  return i;
}

// Philox4x64 round and key schedule (published algorithm; the vendored
// lib/Random123/philox.h:228-275 uses the same multipliers and Weyl
// increments).
static inline void MulHiLo64(uint64_t a, uint64_t b, uint64_t &hi, uint64_t &lo) {
  unsigned __int128 p = (unsigned __int128)a * b;
  hi = uint64_t(p >> 64);
  lo = uint64_t(p);
}

void Philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const uint64_t W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
  uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint64_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t hi0, lo0, hi1, lo1;
    MulHiLo64(M0, c0, hi0, lo0);
    MulHiLo64(M1, c2, hi1, lo1);
    uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0;
    c1 = n1;
    c2 = n2;
    c3 = n3;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// src/Random123Wrapper.cpp:16-22: ctr = {counter, key2, 0, 0}, key = {step, seed}.
void Random123Wrapper::Raw(uint32_t counter, uint64_t out[4]) const {
  uint64_t c[4] = {counter, key2_, 0, 0};
  uint64_t k[2] = {step_, seed_};
  Philox4x64_10(c, k, out);
}

double Random123Wrapper::GetRandomNumber(uint32_t counter) const {
  uint64_t r[4];
  Raw(counter, r);
  return U01(r[0]);
}

// src/Random123Wrapper.cpp:46-53.
XYZ Random123Wrapper::GetRandomCoords(uint32_t counter) const {
  uint64_t r[4];
  Raw(counter, r);
  return {U01(r[0]), U01(r[1]), U01(r[2])};
}

XYZ Random123Wrapper::GetSymRandomCoords(uint32_t counter, double bound) const {
  uint64_t r[4];
  Raw(counter, r);
  return {bound * Uneg11(r[0]), bound * Uneg11(r[1]), bound * Uneg11(r[2])};
}

double Random123Wrapper::GetSymRandom(uint32_t counter, double bound) const {
  uint64_t r[4];
  Raw(counter, r);
  return bound * Uneg11(r[0]);
}

// src/Random123Wrapper.cpp:69-82.
XYZ Random123Wrapper::GetRandomCoordsOnSphere(uint32_t counter) const {
  uint64_t r[4];
  Raw(counter, r);
  double u = Uneg11(r[1]);
  double theta = 2.0 * M_PI * U01(r[2]);
  double rootTerm = std::sqrt(1.0 - u * u);
  return {rootTerm * std::cos(theta), rootTerm * std::sin(theta), u};
}

// src/PRNG.h:72-78.
XYZ PRNG::SymXYZ(double bound) {
  double bound2 = 2.0 * bound;
  double xx = gen->rand(bound2) - bound;
  double yy = gen->rand(bound2) - bound;
  double zz = gen->rand(bound2) - bound;
  return {xx, yy, zz};
}

// src/PRNG.h:143-150.
XYZ PRNG::PickOnUnitSphere() {
  double u = Sym(1.0);
  double theta = gen->randExc(2.0 * M_PI);
  double rootTerm = std::sqrt(1.0 - u * u);
  return {rootTerm * std::cos(theta), rootTerm * std::sin(theta), u};
}

// src/PRNG.h:153-163.
void PRNG::PickArbDist(uint32_t &pick, double &subDraw, const double *w,
                       double Wt, uint32_t len) {
  double sum = 0.0, prevSum = 0.0, draw = rand(Wt);
  for (pick = 0; pick < len && sum < draw; ++pick) {
    prevSum = sum;
    sum += w[pick];
  }
  if (pick != 0) --pick;
  subDraw = draw - prevSum;
}

// src/PRNG.h:167-203 (the error exit becomes -1).
int PRNG::PickWeighted(const double *w, int n, double total) {
  double draw = rand(total);
  double sum = 0.0;
  for (int i = 0; i < n; ++i) {
    sum += w[i];
    if (sum >= draw) return i;
  }
  int lastZero = n;
  for (int i = n - 1; i > 0 && w[i] == 0.0; --i) lastZero = i;
  lastZero--;
  if (std::fabs(draw - total) < 0.001) return lastZero;
  return -1;
}

// src/PRNG.h:92-98 for an orthogonal box (TransformSlant is the identity).
XYZ PRNG::FillWithRandom(const XYZ &axis) {
  double xx = randExc(axis.x);
  double yy = randExc(axis.y);
  double zz = randExc(axis.z);
  return {xx, yy, zz};
}

// src/TransformMatrix.h:160-200: Rodrigues' formula through the cross and
// tensor product matrices, accumulated in the reference's order.
RotationMatrix RotationMatrix::FromAxisAngle(double theta, const XYZ &a) {
  double cross[3][3] = {{0.0, -a.z, a.y}, {a.z, 0.0, -a.x}, {-a.y, a.x, 0.0}};
  double tens[3][3];
  const double av[3] = {a.x, a.y, a.z};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) tens[i][j] = av[i];
  for (int i = 0; i < 3; ++i) {
    tens[i][0] *= a.x;
    tens[i][1] *= a.y;
    tens[i][2] *= a.z;
  }
  double c = std::cos(theta);
  RotationMatrix r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] = (i == j) ? c : 0.0;
  double s = std::sin(theta);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) r.m[i][j] += s * cross[i][j] + (1 - c) * tens[i][j];
  return r;
}

// src/TransformMatrix.h:204-260 (Arvo 1992 uniform random rotation).
RotationMatrix RotationMatrix::UniformRandom(double u1, double u2, double u3) {
  u1 *= 2.0 * M_PI;
  u1 -= M_PI;
  u2 *= 2.0 * M_PI;
  u3 *= 2.0;
  double r = std::sqrt(u3);
  XYZ V(std::sin(u2) * r, std::cos(u2) * r, std::sqrt(2.0 - u3));
  double sinTh = std::sin(u1), cosTh = std::cos(u1);
  double Sx = V.x * cosTh - V.y * sinTh;
  double Sy = V.x * sinTh + V.y * cosTh;
  RotationMatrix res;
  res.m[0][0] = V.x * Sx - cosTh;
  res.m[0][1] = V.x * Sy - sinTh;
  res.m[0][2] = V.x * V.z;
  res.m[1][0] = V.y * Sx + sinTh;
  res.m[1][1] = V.y * Sy - cosTh;
  res.m[1][2] = V.y * V.z;
  res.m[2][0] = V.z * Sx;
  res.m[2][1] = V.z * Sy;
  res.m[2][2] = 1.0 - u3;
  return res;
}

}  // namespace orc

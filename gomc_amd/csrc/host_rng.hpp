// Host-side random draws of the single-box driver, consumed in the order
// the reference consumes them: MT19937 with the MTRand interface
// (src/MersenneTwister.h:194-366) and the PRNG helpers of src/PRNG.h.
#pragma once

#include <cmath>
#include <cstdint>

namespace gomc {

class MTRand {
 public:
  explicit MTRand(uint32_t s) { Seed(s); }
  // init_genrand (Knuth multiplier 1812433253), then one full twist
  // (MersenneTwister.h:194-231).
  void Seed(uint32_t s) {
    mt_[0] = s;
    for (int i = 1; i < kN; ++i) mt_[i] = 1812433253u * (mt_[i - 1] ^ (mt_[i - 1] >> 30)) + uint32_t(i);
    Twist();
  }
  // Tempered 32-bit word (MersenneTwister.h:320-331).
  uint32_t Next() {
    if (idx_ == kN) Twist();
    uint32_t v = mt_[idx_++];
    v ^= v >> 11;
    v ^= (v << 7) & 0x9d2c5680u;
    v ^= (v << 15) & 0xefc60000u;
    return v ^ (v >> 18);
  }
  // [0, n] by masking to n's bit width and rejecting (333-350).
  uint32_t Below(uint32_t n) {
    uint32_t mask = n;
    for (int s = 1; s < 32; s <<= 1) mask |= mask >> s;
    uint32_t v;
    do {
      v = Next() & mask;
    } while (v > n);
    return v;
  }
  double Closed() { return double(Next()) * (1.0 / 4294967295.0); }    // rand()
  double HalfOpen() { return double(Next()) * (1.0 / 4294967296.0); }  // randExc()
  // The generator's state (624 words and the read position), so that the
  // same stream can continue elsewhere (the device move loop).
  void GetState(uint32_t *words, int *index) const {
    for (int i = 0; i < kN; ++i) words[i] = mt_[i];
    *index = idx_;
  }
  void SetState(const uint32_t *words, int index) {
    for (int i = 0; i < kN; ++i) mt_[i] = words[i];
    idx_ = index;
  }

 private:
  static constexpr int kN = 624, kM = 397;
  void Twist() {
    for (int i = 0; i < kN; ++i) {
      uint32_t y = (mt_[i] & 0x80000000u) | (mt_[(i + 1) % kN] & 0x7fffffffu);
      uint32_t v = mt_[(i + kM) % kN] ^ (y >> 1);
      if (mt_[(i + 1) % kN] & 1u) v ^= 0x9908b0dfu;
      mt_[i] = v;
    }
    idx_ = 0;
  }
  uint32_t mt_[kN];
  int idx_ = kN;
};

struct Vec3 {
  double x, y, z;
};

// PRNG::Sym (src/PRNG.h:63): rand(2b) - b.
inline double Sym(MTRand &g, double b) { return g.Closed() * (2.0 * b) - b; }

// PRNG::SymXYZ (src/PRNG.h:72-78).
inline Vec3 SymXYZ(MTRand &g, double b) {
  double b2 = 2.0 * b;
  double sx = g.Closed() * b2 - b;
  double sy = g.Closed() * b2 - b;
  double sz = g.Closed() * b2 - b;
  return {sx, sy, sz};
}

// PRNG::PickOnUnitSphere (src/PRNG.h:143-150): u = cos(polar) uniform.
inline Vec3 OnUnitSphere(MTRand &g) {
  double u = Sym(g, 1.0);
  double theta = g.HalfOpen() * (2.0 * M_PI);
  double r = std::sqrt(1.0 - u * u);
  return {r * std::cos(theta), r * std::sin(theta), u};
}

// TransformMatrix::FromAxisAngle (src/TransformMatrix.h:186-200):
// R = cos(t) I + sin(t) [a]x + (1 - cos(t)) a a^T, summed in that order.
struct Mat3 {
  double m[3][3];
  Vec3 Apply(const Vec3 &p) const {
    return {m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z,
            m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z,
            m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z};
  }
};

inline Mat3 AxisAngle(double theta, const Vec3 &a) {
  const double av[3] = {a.x, a.y, a.z};
  const double cross[3][3] = {{0.0, -a.z, a.y}, {a.z, 0.0, -a.x}, {-a.y, a.x, 0.0}};
  double c = std::cos(theta), s = std::sin(theta);
  Mat3 r;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double tensor = av[i] * av[j];
      r.m[i][j] = (i == j ? c : 0.0) + (s * cross[i][j] + (1 - c) * tensor);
    }
  return r;
}

// TransformMatrix::UniformRandom (src/TransformMatrix.h:216-262): Arvo's
// (1992) uniform rotation -- a rotation about the pole followed by an
// inverted Householder reflection that moves the pole to a random point.
inline Mat3 UniformRotation(double u1, double u2, double u3) {
  u1 *= 2.0 * M_PI;
  u1 -= M_PI;
  u2 *= 2.0 * M_PI;
  u3 *= 2.0;
  const double r = std::sqrt(u3);
  const double vx = std::sin(u2) * r, vy = std::cos(u2) * r, vz = std::sqrt(2.0 - u3);
  const double sinTh = std::sin(u1), cosTh = std::cos(u1);
  const double sx = vx * cosTh - vy * sinTh;
  const double sy = vx * sinTh + vy * cosTh;
  Mat3 m;
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

}  // namespace gomc

// Driver that exercises the reference's own RNG sources (compiled from
// src/Random123Wrapper.cpp, src/MersenneTwister.h, lib/Random123 and
// src/TransformMatrix.h where they lie) and prints their outputs as JSON.
// Built by oracle/Makefile into oracle/_ref/; used only to write and check
// the golden vectors under tests/golden/.
#include <cinttypes>
#include <cstdio>

#include "MersenneTwister.h"
#include "Random123Wrapper.h"
#include "Random123/philox.h"
#include "Random123/uniform.hpp"
#include "TransformMatrix.h"

static void philox_case(const uint64_t c[4], const uint64_t k[2], bool last) {
  r123::Philox4x64 rng;
  r123::Philox4x64::ctr_type ctr = {{c[0], c[1], c[2], c[3]}};
  r123::Philox4x64::key_type key = {{k[0], k[1]}};
  r123::Philox4x64::ctr_type r = rng(ctr, key);
  std::printf("    {\"ctr\": [\"%016" PRIx64 "\", \"%016" PRIx64 "\", \"%016" PRIx64
              "\", \"%016" PRIx64 "\"], \"key\": [\"%016" PRIx64 "\", \"%016" PRIx64
              "\"], \"out\": [\"%016" PRIx64 "\", \"%016" PRIx64 "\", \"%016" PRIx64
              "\", \"%016" PRIx64 "\"], \"u01\": [%.17g, %.17g], \"uneg11\": [%.17g, %.17g]}%s\n",
              c[0], c[1], c[2], c[3], k[0], k[1], (uint64_t)r[0], (uint64_t)r[1],
              (uint64_t)r[2], (uint64_t)r[3], r123::u01<double>((uint64_t)r[0]),
              r123::u01<double>((uint64_t)r[1]), r123::uneg11<double>((uint64_t)r[2]),
              r123::uneg11<double>((uint64_t)r[3]), last ? "" : ",");
}

int main() {
  std::printf("{\n  \"philox4x64_10\": [\n");
  const uint64_t cases[][6] = {
      {0, 0, 0, 0, 0, 0},
      {~0ull, ~0ull, ~0ull, ~0ull, ~0ull, ~0ull},
      {0x243f6a8885a308d3ull, 0x13198a2e03707344ull, 0xa4093822299f31d0ull,
       0x082efa98ec4e6c89ull, 0x452821e638d01377ull, 0xbe5466cf34e90c6cull},
      {7, 0, 0, 0, 12345, 0x474F4D43ull},
      {123456789, 1, 0, 0, 987654321, 0x474F4D43ull},
      {42, 3, 0, 0, 100, 12345},
  };
  const int nc = sizeof(cases) / sizeof(cases[0]);
  for (int i = 0; i < nc; ++i) philox_case(cases[i], cases[i] + 4, i == nc - 1);
  std::printf("  ],\n");

  // Random123Wrapper: key = {step, seed}, counter = {id, 0}.
  std::printf("  \"wrapper\": [\n");
  const unsigned long steps[] = {0, 1, 999, 123456};
  for (int s = 0; s < 4; ++s) {
    Random123Wrapper w;
    w.SetRandomSeed(12345);
    w.SetStep(steps[s]);
    for (unsigned c = 0; c < 3; ++c) {
      XYZ sym = w.GetSymRandomCoords(c * 17 + 1, 0.75);
      XYZ sph = w.GetRandomCoordsOnSphere(c * 17 + 1);
      std::printf("    {\"seed\": 12345, \"step\": %lu, \"counter\": %u, \"rand\": %.17g, "
                  "\"sym\": [%.17g, %.17g, %.17g], \"sphere\": [%.17g, %.17g, %.17g], "
                  "\"symr\": %.17g}%s\n",
                  steps[s], c * 17 + 1, w.GetRandomNumber(c * 17 + 1), sym.x, sym.y, sym.z,
                  sph.x, sph.y, sph.z, w.GetSymRandom(c * 17 + 1, 2.5),
                  (s == 3 && c == 2) ? "" : ",");
    }
  }
  std::printf("  ],\n");

  // MTRand: raw 32-bit draws, rand(), randExc(), randInt(n).
  std::printf("  \"mtrand\": [\n");
  const unsigned seeds[] = {5489u, 12345u, 4294967295u, 1u};
  for (int s = 0; s < 4; ++s) {
    MTRand g(seeds[s]), g2(seeds[s]), g3(seeds[s]), g4(seeds[s]);
    std::printf("    {\"seed\": %u, \"raw\": [", seeds[s]);
    for (int i = 0; i < 700; ++i) std::printf("%u%s", g.randInt(), i == 699 ? "" : ", ");
    std::printf("], \"rand\": [");
    for (int i = 0; i < 8; ++i) std::printf("%.17g%s", g2.rand(), i == 7 ? "" : ", ");
    std::printf("], \"randExc\": [");
    for (int i = 0; i < 8; ++i) std::printf("%.17g%s", g3.randExc(), i == 7 ? "" : ", ");
    std::printf("], \"randInt_999\": [");
    for (int i = 0; i < 64; ++i) std::printf("%u%s", g4.randInt(999u), i == 63 ? "" : ", ");
    std::printf("]}%s\n", s == 3 ? "" : ",");
  }
  std::printf("  ],\n");
  {
    MTRand g(5489u);
    unsigned v = 0;
    for (int i = 0; i < 10000; ++i) v = g.randInt();
    std::printf("  \"mt_default_10000th\": %u,\n", v);
  }

  // TransformMatrix: axis-angle and Arvo uniform rotations.
  std::printf("  \"rotation\": [\n");
  const double rc[][4] = {{0.3, 0.0, 0.0, 1.0}, {-1.2, 0.6, 0.0, 0.8},
                          {2.9, 0.2672612419124244, 0.5345224838248488, 0.8017837257372732}};
  for (int i = 0; i < 3; ++i) {
    RotationMatrix m = RotationMatrix::FromAxisAngle(rc[i][0], XYZ(rc[i][1], rc[i][2], rc[i][3]));
    std::printf("    {\"theta\": %.17g, \"axis\": [%.17g, %.17g, %.17g], \"m\": [", rc[i][0],
                rc[i][1], rc[i][2], rc[i][3]);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        std::printf("%.17g%s", m.matrix[a][b], (a == 2 && b == 2) ? "" : ", ");
    std::printf("]},\n");
  }
  const double ur[][3] = {{0.1, 0.2, 0.3}, {0.9, 0.5, 0.05}};
  for (int i = 0; i < 2; ++i) {
    RotationMatrix m = RotationMatrix::UniformRandom(ur[i][0], ur[i][1], ur[i][2]);
    std::printf("    {\"uniform\": [%.17g, %.17g, %.17g], \"m\": [", ur[i][0], ur[i][1], ur[i][2]);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        std::printf("%.17g%s", m.matrix[a][b], (a == 2 && b == 2) ? "" : ", ");
    std::printf("]}%s\n", i == 1 ? "" : ",");
  }
  std::printf("  ]\n}\n");
  return 0;
}

// Synthetic initial configurations: random sequential insertion of rigid
// molecules (uniform centre, uniform orientation, overlap rejection) into a
// cubic box.  Host code; it stands in for the PDB coordinates the reference
// reads at start-up and is not on the timed path.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/gomc_hip.h"

namespace {

// splitmix64 (Steele, Lea, Flood 2014): a fully specified 64-bit stream, so
// the same seed gives the same box on every machine.
struct SplitMix {
  uint64_t s;
  uint64_t Next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double Uniform() { return double(Next() >> 11) * 0x1p-53; }
};

// Uniform random rotation from a uniform unit quaternion (Shoemake 1992).
void RandomRotation(SplitMix &g, double R[3][3]) {
  const double u1 = g.Uniform(), u2 = g.Uniform() * 2.0 * M_PI, u3 = g.Uniform() * 2.0 * M_PI;
  const double a = std::sqrt(1.0 - u1) * std::sin(u2), b = std::sqrt(1.0 - u1) * std::cos(u2);
  const double c = std::sqrt(u1) * std::sin(u3), d = std::sqrt(u1) * std::cos(u3);
  R[0][0] = a * a + b * b - c * c - d * d;
  R[0][1] = 2 * (b * c - a * d);
  R[0][2] = 2 * (b * d + a * c);
  R[1][0] = 2 * (b * c + a * d);
  R[1][1] = a * a - b * b + c * c - d * d;
  R[1][2] = 2 * (c * d - a * b);
  R[2][0] = 2 * (b * d - a * c);
  R[2][1] = 2 * (c * d + a * b);
  R[2][2] = a * a - b * b - c * c + d * d;
}

}  // namespace

extern "C" int gomc_place_molecules(int n_kinds, const int32_t *kind_natoms, const double *kind_geom,
                                    int n_mols, const int32_t *mol_kind, double box, uint64_t seed,
                                    double min_heavy, double min_any, int max_tries, double *x,
                                    double *y, double *z) {
  if (n_kinds <= 0 || n_mols <= 0 || !(box > 0) || !kind_natoms || !kind_geom || !mol_kind)
    return GOMC_ERR_ARG;
  std::vector<int> goff(n_kinds + 1, 0);
  for (int k = 0; k < n_kinds; ++k) {
    if (kind_natoms[k] <= 0) return GOMC_ERR_ARG;
    goff[k + 1] = goff[k] + kind_natoms[k];
  }
  // body-frame geometry centred on its centroid
  std::vector<double> geo(3 * goff[n_kinds]);
  for (int k = 0; k < n_kinds; ++k) {
    double c[3] = {0, 0, 0};
    for (int a = goff[k]; a < goff[k + 1]; ++a)
      for (int d = 0; d < 3; ++d) c[d] += kind_geom[3 * a + d];
    for (int a = goff[k]; a < goff[k + 1]; ++a)
      for (int d = 0; d < 3; ++d) geo[3 * a + d] = kind_geom[3 * a + d] - c[d] / kind_natoms[k];
  }
  const double minLen = std::fmax(min_heavy, min_any);
  int nc = std::max(3, int(box / std::fmax(minLen, 2.0)));
  const double csz = box / nc;
  std::vector<std::vector<int>> cells(size_t(nc) * nc * nc);
  std::vector<double> px, py, pz;
  std::vector<char> heavy;
  SplitMix g{seed};
  auto cellOf = [&](double v) {
    int c = int(v / csz);
    return c >= nc ? nc - 1 : (c < 0 ? 0 : c);
  };
  auto mi = [&](double d) { return d - box * std::nearbyint(d / box); };
  double tx[64], ty[64], tz[64];
  for (int m = 0; m < n_mols; ++m) {
    const int k = mol_kind[m];
    if (k < 0 || k >= n_kinds) return GOMC_ERR_ARG;
    const int na = kind_natoms[k];
    if (na > 64) return GOMC_ERR_ARG;
    bool placed = false;
    for (int t = 0; t < max_tries && !placed; ++t) {
      const double cx = g.Uniform() * box, cy = g.Uniform() * box, cz = g.Uniform() * box;
      double R[3][3];
      RandomRotation(g, R);
      for (int a = 0; a < na; ++a) {
        const double *v = &geo[3 * (goff[k] + a)];
        double p[3];
        for (int d = 0; d < 3; ++d) p[d] = R[d][0] * v[0] + R[d][1] * v[1] + R[d][2] * v[2];
        tx[a] = std::fmod(cx + p[0] + box, box);
        ty[a] = std::fmod(cy + p[1] + box, box);
        tz[a] = std::fmod(cz + p[2] + box, box);
        if (tx[a] >= box) tx[a] -= box;
        if (ty[a] >= box) ty[a] -= box;
        if (tz[a] >= box) tz[a] -= box;
      }
      bool ok = true;
      for (int a = 0; a < na && ok; ++a) {
        const int ix = cellOf(tx[a]), iy = cellOf(ty[a]), iz = cellOf(tz[a]);
        for (int dx = -1; dx <= 1 && ok; ++dx)
          for (int dy = -1; dy <= 1 && ok; ++dy)
            for (int dz = -1; dz <= 1 && ok; ++dz) {
              const size_t c = (size_t((ix + dx + nc) % nc) * nc + (iy + dy + nc) % nc) * nc +
                               (iz + dz + nc) % nc;
              for (int j : cells[c]) {
                const double ddx = mi(tx[a] - px[j]), ddy = mi(ty[a] - py[j]), ddz = mi(tz[a] - pz[j]);
                const double lim = (a == 0 && heavy[j]) ? min_heavy : min_any;
                if (ddx * ddx + ddy * ddy + ddz * ddz < lim * lim) {
                  ok = false;
                  break;
                }
              }
            }
      }
      if (!ok) continue;
      for (int a = 0; a < na; ++a) {
        const int id = int(px.size());
        px.push_back(tx[a]);
        py.push_back(ty[a]);
        pz.push_back(tz[a]);
        heavy.push_back(a == 0);
        cells[(size_t(cellOf(tx[a])) * nc + cellOf(ty[a])) * nc + cellOf(tz[a])].push_back(id);
      }
      placed = true;
    }
    if (!placed) return GOMC_ERR_STATE;
  }
  for (size_t i = 0; i < px.size(); ++i) {
    x[i] = px[i];
    y[i] = py[i];
    z[i] = pz[i];
  }
  return GOMC_OK;
}

// Grand-canonical molecule transfer on the plain path: GOMC's
// MoleculeTransfer move (src/moves/MoleculeTransfer.h:85-246) between the
// simulation box (0) and the non-interacting reservoir (1), with the
// configurational-bias growth the reference uses for rigid kinds
// (src/CBMC.cpp:18-36):
//   one-atom kinds   DCLinear -> DCSingle (src/cbmc/DCLinear.cpp:44-55,
//                    src/cbmc/DCSingle.cpp:23-100);
//   three-site kinds DCGraph's single node -> DCFreeHedron grown from atom 0
//                    (src/cbmc/DCGraph.cpp:188-203,
//                    src/cbmc/DCFreeHedron.cpp:85-240) with the angle trials
//                    of DCHedron (src/cbmc/DCHedron.cpp:95-206,301-316).
// Bonds and angles are fixed and 1-3 pairs excluded, so their energies are
// zero, while every MT19937 draw those routines take is still taken.  The
// trial energies of each growth stage come from one ParticleInter launch;
// the reciprocal change from one SwapRecip launch.  Also the molecule
// lookup the move pickers read (src/MoleculeLookup.cpp:24-198).
#include <algorithm>
#include <cmath>

#include "box.hpp"

namespace gomc {

namespace {
constexpr double kSmallWeight = 1.0e-38;  // lib/NumLib.h:19

struct V3 {
  double x, y, z;
};
inline V3 Add(const V3 &a, const V3 &b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 Neg(const V3 &a) { return {-a.x, -a.y, -a.z}; }
inline V3 Apply(const Mat3 &r, const V3 &p) {
  const Vec3 v = r.Apply(Vec3{p.x, p.y, p.z});
  return {v.x, v.y, v.z};
}
inline V3 WrapV(const gomc_box *b, const V3 &p) {
  return {Wrap(p.x, b->geom.L[0]), Wrap(p.y, b->geom.L[1]), Wrap(p.z, b->geom.L[2])};
}
inline V3 UnwrapV(const gomc_box *b, const V3 &p, const V3 &ref) {
  return {Unwrap(p.x, ref.x, b->geom.L[0], b->geom.half[0]),
          Unwrap(p.y, ref.y, b->geom.L[1], b->geom.half[1]),
          Unwrap(p.z, ref.z, b->geom.L[2], b->geom.half[2])};
}
inline double MinImageH(double d, int i, const gomc_box *b) {
  if (d > b->geom.half[i])
    d -= b->geom.L[i];
  else if (d < -b->geom.half[i])
    d += b->geom.L[i];
  return d;
}

// One trial molecule (cbmc::TrialMol): its box, atom positions, Rosenbluth
// weight, overlap flag and the inter/real energies it collected.
struct Trial {
  int box = 0;
  std::vector<V3> pos;
  double weight = 1.0;
  bool overlap = false;
  double inter = 0.0, real = 0.0;
};

// ParticleInter of one atom's trial positions, accumulated; the reservoir
// has no interactions (CalculateEnergy.cpp:731-732).
int TrialEnergies(gomc_box *b, const Trial &t, int kind, int atom, int mol,
                  const std::vector<V3> &p, std::vector<double> &inter,
                  std::vector<double> &real, std::vector<char> &ov) {
  if (t.box != 0) return GOMC_OK;
  const int n = int(p.size());
  std::vector<double> x(n), y(n), z(n), en(n), re(n);
  std::vector<int32_t> o(n);
  for (int i = 0; i < n; ++i) {
    x[i] = p[i].x;
    y[i] = p[i].y;
    z[i] = p[i].z;
  }
  int rc = ParticleInterImpl(b, b->kindType[kind][atom], b->kindCharge[kind][atom], mol, n,
                             x.data(), y.data(), z.data(), en.data(), re.data(), o.data());
  if (rc) return rc;
  for (int i = 0; i < n; ++i) {
    inter[i] += en[i];
    real[i] += re[i];
    ov[i] = ov[i] || o[i];
  }
  return GOMC_OK;
}

// PRNG::PickWeighted (src/PRNG.h:167-203).
int PickWeighted(gomc_box *b, const double *w, int n, double total) {
  const double draw = b->mt.Closed() * total;
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

// PRNG::FillWithRandom for an orthogonal box (src/PRNG.h:81-90).
V3 RandomInBox(gomc_box *b) {
  const double xx = b->mt.HalfOpen() * b->geom.L[0];
  const double yy = b->mt.HalfOpen() * b->geom.L[1];
  const double zz = b->mt.HalfOpen() * b->geom.L[2];
  return {xx, yy, zz};
}

// DCSingle::BuildNew (DCSingle.cpp:60-99).
int SingleNew(gomc_box *b, Trial &t, int kind, int atom, int mol) {
  const int n = b->cbmcFirst;
  std::vector<V3> p(n);
  for (int i = 0; i < n; ++i) p[i] = RandomInBox(b);
  std::vector<double> inter(n, 0.0), real(n, 0.0), w(n, 0.0);
  std::vector<char> ov(n, 0);
  int rc = TrialEnergies(b, t, kind, atom, mol, p, inter, real, ov);
  if (rc) return rc;
  double stepWeight = 0;
  for (int i = 0; i < n; ++i) {
    w[i] = std::exp(-1 * b->beta * (inter[i] + real[i]));
    stepWeight += w[i];
  }
  int win = PickWeighted(b, w.data(), n, stepWeight);
  if (win < 0) return Fail(b, GOMC_ERR_STATE, "CBMC weights inconsistent");
  t.overlap = t.overlap || ov[win];
  t.weight *= stepWeight / n;
  t.inter += inter[win];
  t.real += real[win];
  t.pos[atom] = p[win];
  return GOMC_OK;
}

// DCSingle::BuildOld (DCSingle.cpp:23-58).
int SingleOld(gomc_box *b, Trial &t, int kind, int atom, int mol) {
  const int n = b->cbmcFirst;
  std::vector<V3> p(n);
  for (int i = 0; i < n; ++i) p[i] = RandomInBox(b);
  p[0] = WrapV(b, t.pos[atom]);
  std::vector<double> inter(n, 0.0), real(n, 0.0);
  std::vector<char> ov(n, 0);
  int rc = TrialEnergies(b, t, kind, atom, mol, p, inter, real, ov);
  if (rc) return rc;
  double stepWeight = 0;
  for (int i = 0; i < n; ++i) stepWeight += std::exp(-1 * b->beta * (inter[i] + real[i]));
  t.overlap = t.overlap || ov[0];
  t.weight *= stepWeight / n;
  t.inter += inter[0];
  t.real += real[0];
  return GOMC_OK;
}

// DCFreeHedron::PrepareNew + BuildNew (DCFreeHedron.cpp:48-58,85-158):
// the fixed angle's trials all weigh 1 and the winner is still drawn
// (DCHedron::FreeAnglesNew); then the seed atom, then nLJTrialsNth uniform
// spins of the rigid frame about it.
int FreeHedronNew(gomc_box *b, Trial &t, int kind, int mol) {
  const gomc_box::Hedron &h = b->hedron[kind];
  std::vector<double> aw(b->cbmcAng, 1.0);
  double thetaWeight = 0.0;
  for (double v : aw) thetaWeight += v;
  PickWeighted(b, aw.data(), b->cbmcAng, thetaWeight);
  int rc = SingleNew(b, t, kind, 0, mol);
  if (rc) return rc;
  const int n = b->cbmcNth;
  const V3 center = t.pos[0];
  // TrialMol::RawRectCoords in the identity basis (TrialMol.cpp:123-130):
  // the bonded atom at (theta, phi = 0), the anchor on the z axis.
  const double sinTh = std::sin(h.theta);
  V3 protoB{h.bBonded, h.bBonded, h.bBonded};
  protoB.x *= sinTh * std::cos(0.0);
  protoB.y *= sinTh * std::sin(0.0);
  protoB.z *= std::cos(h.theta);
  V3 protoP{h.bPrev, h.bPrev, h.bPrev};
  protoP.x *= std::sin(0.0) * std::cos(0.0);
  protoP.y *= std::sin(0.0) * std::sin(0.0);
  protoP.z *= std::cos(0.0);
  std::vector<V3> pb(n), pp(n);
  for (int lj = n; lj-- > 0;) {
    const double u1 = b->mt.Closed(), u2 = b->mt.Closed(), u3 = b->mt.Closed();
    const Mat3 spin = UniformRotation(u1, u2, u3);
    pb[lj] = Add(Apply(spin, protoB), center);
    pp[lj] = Add(Apply(spin, protoP), center);
  }
  for (int i = 0; i < n; ++i) {
    pb[i] = WrapV(b, pb[i]);
    pp[i] = WrapV(b, pp[i]);
  }
  std::vector<double> inter(n, 0.0), real(n, 0.0), w(n, 0.0);
  std::vector<char> ov(n, 0);
  rc = TrialEnergies(b, t, kind, 2, mol, pb, inter, real, ov);
  if (!rc) rc = TrialEnergies(b, t, kind, 1, mol, pp, inter, real, ov);
  if (rc) return rc;
  double stepWeight = 0;
  for (int i = 0; i < n; ++i) {
    w[i] = std::exp(-b->beta * (inter[i] + real[i]));
    stepWeight += w[i];
  }
  int win = PickWeighted(b, w.data(), n, stepWeight);
  if (win < 0) return Fail(b, GOMC_ERR_STATE, "CBMC weights inconsistent");
  t.pos[2] = pb[win];
  t.pos[1] = pp[win];
  t.overlap = t.overlap || ov[win];
  t.inter += inter[win];
  t.real += real[win];
  t.weight *= thetaWeight * 1.0;  // DCHedron::GetWeight: theta x phi weight
  t.weight *= stepWeight / n;
  return GOMC_OK;
}

// DCFreeHedron::PrepareOld + BuildOld (DCFreeHedron.cpp:60-70,160-238):
// the actual frame is trial 0, nLJTrialsNth - 1 spins are the others.
int FreeHedronOld(gomc_box *b, Trial &t, int kind, int mol) {
  double thetaWeight = 0.0;
  for (int i = 0; i < b->cbmcAng - 1; ++i) thetaWeight += 1.0;  // FreeAnglesOld
  int rc = SingleOld(b, t, kind, 0, mol);
  if (rc) return rc;
  thetaWeight += std::exp(-1 * b->beta * (0.0 + 0.0));  // IncorporateOld
  const int n = b->cbmcNth;
  const V3 center = t.pos[0];
  std::vector<V3> pb(n), pp(n);
  pb[0] = Add(UnwrapV(b, t.pos[2], center), Neg(center));
  pp[0] = Add(UnwrapV(b, t.pos[1], center), Neg(center));
  for (int lj = n; lj-- > 1;) {
    const double u1 = b->mt.Closed(), u2 = b->mt.Closed(), u3 = b->mt.Closed();
    const Mat3 spin = UniformRotation(u1, u2, u3);
    pb[lj] = Add(Apply(spin, pb[0]), center);
    pp[lj] = Add(Apply(spin, pp[0]), center);
  }
  pb[0] = Add(pb[0], center);
  pp[0] = Add(pp[0], center);
  for (int i = 0; i < n; ++i) {
    pb[i] = WrapV(b, pb[i]);
    pp[i] = WrapV(b, pp[i]);
  }
  std::vector<double> inter(n, 0.0), real(n, 0.0);
  std::vector<char> ov(n, 0);
  rc = TrialEnergies(b, t, kind, 2, mol, pb, inter, real, ov);
  if (!rc) rc = TrialEnergies(b, t, kind, 1, mol, pp, inter, real, ov);
  if (rc) return rc;
  double stepWeight = 0;
  for (int i = 0; i < n; ++i) stepWeight += std::exp(-b->beta * (inter[i] + real[i]));
  t.overlap = t.overlap || ov[0];
  t.inter += inter[0];
  t.real += real[0];
  t.weight *= thetaWeight * 1.0;
  t.weight *= stepWeight / n;
  return GOMC_OK;
}

// Ewald::SwapCorrection (src/Ewald.cpp:1311-1335) of a trial molecule.
double SwapCorrection(const gomc_box *b, const Trial &t, int kind) {
  const std::vector<double> &q = b->kindCharge[kind];
  double correction = 0.0;
  for (size_t i = 0; i < t.pos.size(); ++i)
    for (size_t j = i + 1; j < t.pos.size(); ++j) {
      const double dx = MinImageH(t.pos[i].x - t.pos[j].x, 0, b);
      const double dy = MinImageH(t.pos[i].y - t.pos[j].y, 1, b);
      const double dz = MinImageH(t.pos[i].z - t.pos[j].z, 2, b);
      const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
      correction -= (q[i] * q[j] * std::erf(b->alpha * dist) / dist);
    }
  return kQQFact * correction;
}

// Ewald::SwapSelf (src/Ewald.cpp:1375-1391).
double SwapSelf(const gomc_box *b, int kind) {
  double en = 0.0;
  for (double q : b->kindCharge[kind]) en -= (q * q);
  return (en * b->alpha * kQQFact * M_2_SQRTPI * 0.5);
}

// MoleculeLookup::Shift (src/MoleculeLookup.cpp:165-198): the molecule
// swaps its way across the section boundaries between the two lists.
void ShiftMolBox(gomc_box *b, int m, int from, int to, int k) {
  const int K = b->nkinds;
  uint32_t section = uint32_t(from * K + k);
  uint32_t index = b->bkStart[section];
  while (int(b->lookup[index]) != m) ++index;
  uint32_t oldIndex = index, newIndex;
  b->swapCount[section]--;
  b->swapCount[to * K + k]++;
  if (from >= to) {
    while (section != uint32_t(to * K + k)) {
      newIndex = b->bkStart[section]++;
      std::swap(b->lookup[oldIndex], b->lookup[newIndex]);
      std::swap(oldIndex, newIndex);
      --section;
    }
  } else {
    while (section != uint32_t(to * K + k)) {
      newIndex = --b->bkStart[++section];
      std::swap(b->lookup[oldIndex], b->lookup[newIndex]);
      std::swap(oldIndex, newIndex);
    }
  }
}

// COM::SetNew (src/COM.h:63-88).
void SetNewCOM(gomc_box *b, int m) {
  const int s = b->molStart[m], e = s + b->molLen[m];
  const V3 ref{b->hx[s], b->hy[s], b->hz[s]};
  double cx = 0, cy = 0, cz = 0;
  for (int p = s; p < e; ++p) {
    const V3 u = UnwrapV(b, V3{b->hx[p], b->hy[p], b->hz[p]}, ref);
    cx += u.x;
    cy += u.y;
    cz += u.z;
  }
  const double inv = 1.0 / double(b->molLen[m]);
  const V3 c = WrapV(b, V3{cx * inv, cy * inv, cz * inv});
  b->comx[m] = c.x;
  b->comy[m] = c.y;
  b->comz[m] = c.z;
}

}  // namespace

// MoleculeLookup::Init (src/MoleculeLookup.cpp:24-119).
void LookupInit(gomc_box *b) {
  const int K = b->nkinds;
  std::vector<std::vector<uint32_t>> idx(2 * K);
  b->canSwapKind.clear();
  b->canMoveKind.clear();
  b->swapCount.assign(2 * K, 0);
  for (int m = 0; m < b->nmols; ++m) {
    const int bx = b->molBox[m], k = b->molKind[m];
    idx[bx * K + k].push_back(uint32_t(m));
    if (std::find(b->canSwapKind.begin(), b->canSwapKind.end(), k) == b->canSwapKind.end())
      b->canSwapKind.push_back(k);
    if (std::find(b->canMoveKind.begin(), b->canMoveKind.end(), k) == b->canMoveKind.end())
      b->canMoveKind.push_back(k);
    b->swapCount[bx * K + k]++;
  }
  b->lookup.clear();
  b->bkStart.assign(2 * K + 1, 0);
  for (int s = 0; s < 2 * K; ++s) {
    b->bkStart[s] = uint32_t(b->lookup.size());
    b->lookup.insert(b->lookup.end(), idx[s].begin(), idx[s].end());
  }
  b->bkStart[2 * K] = uint32_t(b->nmols);
}

uint32_t NumKindInBox(const gomc_box *b, int k, int box) {
  return b->bkStart[box * b->nkinds + k + 1] - b->bkStart[box * b->nkinds + k];
}

int GetMolNum(const gomc_box *b, uint32_t sub, int k, int box) {
  return int(b->lookup[b->bkStart[box * b->nkinds + k] + sub]);
}

// One MoleculeTransfer move (PRNG::PickMolAndBoxPair2, PRNG.h:450-455,
// then Prep -> Transform -> CalcEn -> Accept).
int TransferMove(gomc_box *b, double subDraw, int &mol, int &accepted, int &overlapOut,
                 double dE[3]) {
  const int K = b->nkinds;
  const double movPerc = b->moveW[10];
  const double boxDiv = movPerc / 2;
  uint32_t src = uint32_t(subDraw / boxDiv);
  if (src >= 2) src--;
  subDraw -= src * boxDiv;
  const int dst = src == 0 ? 1 : 0;
  const uint32_t mkTot = uint32_t(b->canSwapKind.size());
  const double molDiv = boxDiv / mkTot;
  uint32_t k = uint32_t(subDraw / molDiv);
  if (k == mkTot) k = mkTot - 1;
  const int kind = b->canSwapKind[k];
  mol = -1;
  accepted = 0;
  overlapOut = 0;
  dE[0] = dE[1] = dE[2] = 0.0;
  if (NumKindInBox(b, kind, int(src)) == 0) {
    if (src == 1) return Fail(b, GOMC_ERR_STATE, "no molecules of this kind left in the reservoir");
    b->swapTries[dst * K + kind]++;
    return GOMC_OK;
  }
  const uint32_t off = b->mt.Below(NumKindInBox(b, kind, int(src)) - 1);
  const int m = GetMolNum(b, off, kind, int(src));
  mol = m;
  const int st = b->molStart[m], len = b->molLen[m];
  int rc = FlushPending(b);
  if (rc) return rc;
  Trial nm, om;
  nm.box = dst;
  om.box = int(src);
  nm.pos.assign(len, V3{0, 0, 0});
  om.pos.resize(len);
  for (int p = 0; p < len; ++p) om.pos[p] = V3{b->hx[st + p], b->hy[st + p], b->hz[st + p]};
  // Transform: MoleculeKind::Build through the kind's CBMC graph.
  if (len == 1) {
    b->mt.Below(1);  // DCLinear::Build: forward or backward (randInt(1))
    rc = SingleNew(b, nm, kind, 0, m);
    if (!rc) rc = SingleOld(b, om, kind, 0, m);
  } else {
    b->mt.Below(0);  // DCGraph::Build: its only node (randIntExc(1))
    rc = FreeHedronNew(b, nm, kind, m);
    if (!rc) rc = FreeHedronOld(b, om, kind, m);
  }
  if (rc) return rc;
  const bool overlap = nm.overlap;
  // CalcEn (MoleculeTransfer.h:106-133).
  double W_tc = 1.0, W_recip = 1.0, correct_old = 0, correct_new = 0, self_old = 0, self_new = 0,
         tcLose = 0, tcGain = 0, recipLose = 0, recipGain = 0;
  if (b->lrc) {
    // CalculateEnergy::MoleculeTailChange (CalculateEnergy.cpp:788-808).
    auto tail = [&](bool add) {
      const double sign = add ? 1.0 : -1.0;
      double delta = 0.0;
      for (int j = 0; j < K; ++j)
        delta += b->pairEnCorr[j * K + kind] * (sign * 2.0 * double(b->kindCount[j]) * b->volInv);
      delta += b->pairEnCorr[kind * K + kind] * b->volInv;
      return delta;
    };
    tcLose = src == 0 ? tail(false) : 0.0;
    tcGain = dst == 0 ? tail(true) : 0.0;
    W_tc = std::exp(-1.0 * b->beta * (tcGain + tcLose));
  }
  const bool tried = nm.weight > kSmallWeight && !overlap;
  if (tried && b->ewald) {
    correct_new = dst == 0 ? SwapCorrection(b, nm, kind) : 0.0;
    correct_old = src == 0 ? SwapCorrection(b, om, kind) : 0.0;
    self_new = dst == 0 ? SwapSelf(b, kind) : 0.0;
    self_old = src == 0 ? SwapSelf(b, kind) : 0.0;
    const Trial &t = dst == 0 ? nm : om;
    std::vector<double> x(len), y(len), z(len);
    for (int p = 0; p < len; ++p) {
      x[p] = t.pos[p].x;
      y[p] = t.pos[p].y;
      z[p] = t.pos[p].z;
    }
    double enew = 0.0;
    rc = SwapRecipImpl(b, kind, dst == 0 ? 1 : 0, x.data(), y.data(), z.data(), m, &enew);
    if (rc) return rc;
    if (dst == 0)
      recipGain = enew - b->pot.recip;
    else
      recipLose = enew - b->pot.recip;
    W_recip = std::exp(-1.0 * b->beta *
                       (recipGain + recipLose + correct_new - correct_old + self_new - self_old));
  }
  // Accept (MoleculeTransfer.h:167-246), GetCoeff with fugacities (143-165).
  double coeff;
  if (src == 0)
    coeff = double(b->swapCount[0 * K + kind]) * b->volInv / (b->beta * b->fugacity[kind]);
  else
    coeff = b->volume / double(b->swapCount[dst * K + kind] + 1) * (b->beta * b->fugacity[kind]);
  const double Wrat = nm.weight / om.weight * W_tc * W_recip;
  bool result = false;
  if (tried) result = b->mt.Closed() < coeff * Wrat;
  if (result) {
    if (src == 0) {
      b->pot.tc += tcLose;
      b->pot.inter -= om.inter;
      b->pot.real -= om.real;
      b->pot.recip += recipLose;
      b->pot.corr -= correct_old;
      b->pot.self -= self_old;
    } else {
      b->pot.tc += tcGain;
      b->pot.inter += nm.inter;
      b->pot.real += nm.real;
      b->pot.recip += recipGain;
      b->pot.corr += correct_new;
      b->pot.self += self_new;
    }
    for (int p = 0; p < len; ++p) {
      b->hx[st + p] = nm.pos[p].x;
      b->hy[st + p] = nm.pos[p].y;
      b->hz[st + p] = nm.pos[p].z;
    }
    const size_t bytes = sizeof(double) * len;
    HIP_TRY(b, hipMemcpyAsync(b->x + st, &b->hx[st], bytes, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->y + st, &b->hy[st], bytes, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(b, hipMemcpyAsync(b->z + st, &b->hz[st], bytes, hipMemcpyHostToDevice, b->stream));
    rc = SetMolLive(b, m, dst == 0);  // synchronises the copies above
    if (rc) return rc;
    SetNewCOM(b, m);
    ShiftMolBox(b, m, int(src), dst, kind);
    b->molBox[m] = dst;
    if (dst == 0)
      b->kindCount[kind]++;
    else
      b->kindCount[kind]--;
    const uint32_t left = b->bkStart[K] - b->bkStart[0];
    if (src == 0 && left == 0) {
      b->pot = Energies();
    } else if (src == 0 && left == 1) {
      b->pot.inter = 0;
      b->pot.real = 0;
    }
    if (recipLose != 0.0 || recipGain != 0.0) gomc_recip_accept(b);
    b->pot.Total();
    if (dst == 0) b->mp.singleAccepted = true;  // MoveSettings::Update marks box dst
  }
  b->swapTries[dst * K + kind]++;
  if (result) b->swapAcc[dst * K + kind]++;
  accepted = result;
  overlapOut = overlap;
  dE[0] = dst == 0 ? nm.inter : -om.inter;
  dE[1] = dst == 0 ? nm.real : -om.real;
  dE[2] = recipGain + recipLose;
  return GOMC_OK;
}

}  // namespace gomc

extern "C" int gomc_mol_boxes(const gomc_box *b, int32_t *box) {
  if (!b || !box) return GOMC_ERR_ARG;
  for (int m = 0; m < b->nmols; ++m) box[m] = b->molBox[m];
  return int(b->bkStart[b->nkinds] - b->bkStart[0]);  // molecules in box 0
}

// CPU oracle -- GCMC molecule transfer between the simulation box (0) and
// the non-interacting reservoir (1) with configurational-bias growth.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
//
// Follows src/moves/MoleculeTransfer.h:85-246 with the CBMC builders the
// reference picks for rigid kinds (src/CBMC.cpp:18-36): one-atom kinds grow
// through DCLinear -> DCSingle (src/cbmc/DCLinear.cpp:44-55,
// src/cbmc/DCSingle.cpp:23-100); three-site kinds through DCGraph's single
// node -> DCFreeHedron (src/cbmc/DCGraph.cpp:188-203,
// src/cbmc/DCFreeHedron.cpp:85-240) with the angle trials of DCHedron
// (src/cbmc/DCHedron.cpp:95-206,301-316).  Bonds and the angle are fixed
// and 1-3 pairs excluded, so bond, bend and 1-3 energies are zero
// (FFBonds.h:23-25, FFAngles.h:29-31, CalculateEnergy.cpp:1196-1200) while
// every random draw of those routines is still taken.
#include <algorithm>
#include <cmath>

#include "oracle.hpp"

namespace orc {

namespace {
constexpr double kSmallWeight = 1.0e-38;  // lib/NumLib.h:19
constexpr int kMolTransfer = 10;

// One trial molecule (cbmc::TrialMol): box, atom positions, Rosenbluth
// weight, overlap flag and the inter/real energies it collected.
struct TrialMol {
  int box = 0;
  std::vector<XYZ> pos;
  double weight = 1.0;
  bool overlap = false;
  double inter = 0.0, real = 0.0;
};
}  // namespace

// MoleculeLookup::Shift (MoleculeLookup.cpp:165-198).
void System::ShiftMolBox(int m, int from, int to, int k) {
  const int K = int(kinds.size());
  uint32_t section = uint32_t(from * K + k);
  uint32_t index = bkStart[section];
  while (int(lookup[index]) != m) ++index;
  uint32_t oldIndex = index, newIndex;
  swapCount[section]--;
  swapCount[to * K + k]++;
  if (from >= to) {
    while (section != uint32_t(to * K + k)) {
      newIndex = bkStart[section]++;
      std::swap(lookup[oldIndex], lookup[newIndex]);
      std::swap(oldIndex, newIndex);
      --section;
    }
  } else {
    while (section != uint32_t(to * K + k)) {
      newIndex = --bkStart[++section];
      std::swap(lookup[oldIndex], lookup[newIndex]);
      std::swap(oldIndex, newIndex);
    }
  }
}

// COM::SetNew (COM.h:63-88): centre of the molecule unwrapped around its
// first atom, wrapped.
void System::SetNewCOM(int m) {
  int s = molStart[m], e = s + molLen[m];
  XYZ ref = Pos(s);
  double cx = 0, cy = 0, cz = 0;
  for (int p = s; p < e; ++p) {
    XYZ u = box.Unwrap(Pos(p), ref);
    cx += u.x;
    cy += u.y;
    cz += u.z;
  }
  double inv = 1.0 / double(molLen[m]);
  XYZ c = box.Wrap(XYZ(cx * inv, cy * inv, cz * inv));
  comx[m] = c.x;
  comy[m] = c.y;
  comz[m] = c.z;
}

namespace {

// CalculateEnergy::ParticleInter for the trial positions of one atom,
// accumulated into inter/real; the reservoir has no interactions.
void TrialEnergies(const System &s, int box, int kind, int atom, const std::vector<XYZ> &p,
                   std::vector<double> &inter, std::vector<double> &real,
                   std::vector<char> &overlap) {
  if (box != 0) return;
  bool *ovp = new bool[p.size()];
  for (size_t i = 0; i < p.size(); ++i) ovp[i] = false;
  s.ParticleInter(inter.data(), real.data(), p, ovp, s.kinds[kind].atomType[atom],
                  s.kinds[kind].charge[atom], -1);
  for (size_t i = 0; i < p.size(); ++i) overlap[i] = overlap[i] || ovp[i];
  delete[] ovp;
}

// DCSingle::BuildNew (DCSingle.cpp:60-99).
void SingleNew(System &s, TrialMol &t, int kind, int atom) {
  const int n = s.cbmcFirst;
  std::vector<XYZ> p(n);
  for (int i = 0; i < n; ++i) p[i] = s.prng.FillWithRandom(s.box.axis);
  std::vector<double> inter(n, 0.0), real(n, 0.0), w(n, 0.0);
  std::vector<char> ov(n, 0);
  TrialEnergies(s, t.box, kind, atom, p, inter, real, ov);
  double stepWeight = 0;
  for (int i = 0; i < n; ++i) {
    w[i] = std::exp(-1 * s.ff.beta * (inter[i] + real[i]));
    stepWeight += w[i];
  }
  int winner = s.prng.PickWeighted(w.data(), n, stepWeight);
  if (winner < 0) winner = 0;
  if (g_plant == 2) winner = (winner + 1) % n;
  t.overlap = t.overlap || ov[winner];
  t.weight *= stepWeight / n;
  t.inter += inter[winner];
  t.real += real[winner];
  t.pos[atom] = p[winner];
}

// DCSingle::BuildOld (DCSingle.cpp:23-58).
void SingleOld(System &s, TrialMol &t, int kind, int atom) {
  const int n = s.cbmcFirst;
  std::vector<XYZ> p(n);
  for (int i = 0; i < n; ++i) p[i] = s.prng.FillWithRandom(s.box.axis);
  p[0] = s.box.Wrap(t.pos[atom]);
  std::vector<double> inter(n, 0.0), real(n, 0.0);
  std::vector<char> ov(n, 0);
  TrialEnergies(s, t.box, kind, atom, p, inter, real, ov);
  double stepWeight = 0;
  for (int i = 0; i < n; ++i) stepWeight += std::exp(-1 * s.ff.beta * (inter[i] + real[i]));
  t.overlap = t.overlap || ov[0];
  t.weight *= stepWeight / n;
  t.inter += inter[0];
  t.real += real[0];
}

// DCHedron::PrepareNew for the one free angle (FreeAnglesNew,
// DCHedron.cpp:178-192): every trial is the fixed angle with weight 1; the
// winner is still drawn.  Returns the theta weight.
double HedronAnglesNew(System &s) {
  const int n = s.cbmcAng;
  std::vector<double> w(n, 1.0);
  double stepWeight = 0.0;
  for (int i = 0; i < n; ++i) stepWeight += w[i];
  s.prng.PickWeighted(w.data(), n, stepWeight);
  return stepWeight;
}

// DCFreeHedron::BuildNew (DCFreeHedron.cpp:85-158) for focus 0, prev 1,
// bonded 2, after PrepareNew.
void FreeHedronNew(System &s, TrialMol &t, int kind) {
  const System::Hedron &h = s.hedron[kind];
  const double thetaWeight = HedronAnglesNew(s);  // PrepareNew
  SingleNew(s, t, kind, 0);                        // seed
  const int n = s.cbmcNth;
  const XYZ center = t.pos[0];
  // TrialMol::RawRectCoords with the identity basis (TrialMol.cpp:123-130).
  const double sinTh = std::sin(h.theta);
  XYZ protoB(h.bBonded, h.bBonded, h.bBonded);
  protoB.x *= sinTh * std::cos(0.0);
  protoB.y *= sinTh * std::sin(0.0);
  protoB.z *= std::cos(h.theta);
  XYZ protoP(h.bPrev, h.bPrev, h.bPrev);
  protoP.x *= std::sin(0.0) * std::cos(0.0);
  protoP.y *= std::sin(0.0) * std::sin(0.0);
  protoP.z *= std::cos(0.0);
  std::vector<XYZ> pb(n), pp(n);
  for (int lj = n; lj-- > 0;) {
    double u1 = s.prng(), u2 = s.prng(), u3 = s.prng();
    RotationMatrix spin = RotationMatrix::UniformRandom(u1, u2, u3);
    pb[lj] = spin.Apply(protoB) + center;
    pp[lj] = spin.Apply(protoP) + center;
  }
  for (int i = 0; i < n; ++i) {
    pb[i] = s.box.Wrap(pb[i]);
    pp[i] = s.box.Wrap(pp[i]);
  }
  std::vector<double> inter(n, 0.0), real(n, 0.0), w(n, 0.0);
  std::vector<char> ov(n, 0);
  TrialEnergies(s, t.box, kind, 2, pb, inter, real, ov);
  TrialEnergies(s, t.box, kind, 1, pp, inter, real, ov);
  double stepWeight = 0;
  for (int i = 0; i < n; ++i) {
    w[i] = std::exp(-s.ff.beta * (inter[i] + real[i]));
    stepWeight += w[i];
  }
  int winner = s.prng.PickWeighted(w.data(), n, stepWeight);
  if (winner < 0) winner = 0;
  if (g_plant == 2) winner = (winner + 1) % n;
  t.pos[2] = pb[winner];
  t.pos[1] = pp[winner];
  t.overlap = t.overlap || ov[winner];
  t.inter += inter[winner];
  t.real += real[winner];
  t.weight *= thetaWeight * 1.0;  // DCHedron::GetWeight: theta x phi weight
  t.weight *= stepWeight / n;
}

// DCFreeHedron::BuildOld (DCFreeHedron.cpp:160-238) after PrepareOld.
void FreeHedronOld(System &s, TrialMol &t, int kind) {
  // PrepareOld: FreeAnglesOld over nAngleTrials - 1 fixed angles, then
  // IncorporateOld adds the actual angle's weight.
  double thetaWeight = 0.0;
  for (int i = 0; i < s.cbmcAng - 1; ++i) thetaWeight += 1.0;
  SingleOld(s, t, kind, 0);  // seed
  thetaWeight += std::exp(-1 * s.ff.beta * (0.0 + 0.0));
  const int n = s.cbmcNth;
  const XYZ center = t.pos[0];
  std::vector<XYZ> pb(n), pp(n);
  pb[0] = s.box.Unwrap(t.pos[2], center) + (-center);
  pp[0] = s.box.Unwrap(t.pos[1], center) + (-center);
  for (int lj = n; lj-- > 1;) {
    double u1 = s.prng(), u2 = s.prng(), u3 = s.prng();
    RotationMatrix spin = RotationMatrix::UniformRandom(u1, u2, u3);
    pb[lj] = spin.Apply(pb[0]) + center;
    pp[lj] = spin.Apply(pp[0]) + center;
  }
  pb[0] = pb[0] + center;
  pp[0] = pp[0] + center;
  for (int i = 0; i < n; ++i) {
    pb[i] = s.box.Wrap(pb[i]);
    pp[i] = s.box.Wrap(pp[i]);
  }
  std::vector<double> inter(n, 0.0), real(n, 0.0);
  std::vector<char> ov(n, 0);
  TrialEnergies(s, t.box, kind, 2, pb, inter, real, ov);
  TrialEnergies(s, t.box, kind, 1, pp, inter, real, ov);
  double stepWeight = 0;
  for (int i = 0; i < n; ++i) stepWeight += std::exp(-s.ff.beta * (inter[i] + real[i]));
  t.overlap = t.overlap || ov[0];
  t.inter += inter[0];
  t.real += real[0];
  t.weight *= thetaWeight * 1.0;
  t.weight *= stepWeight / n;
}

// MoleculeKind::Build through the kind's CBMC graph.
void BuildPair(System &s, int kind, TrialMol &oldMol, TrialMol &newMol) {
  if (s.kinds[kind].NumAtoms() == 1) {
    s.prng.randIntExc(2);  // DCLinear::Build picks forward/backward (randInt(1))
    SingleNew(s, newMol, kind, 0);
    SingleOld(s, oldMol, kind, 0);
  } else {
    s.prng.randIntExc(1);  // DCGraph::Build picks its only node
    FreeHedronNew(s, newMol, kind);
    FreeHedronOld(s, oldMol, kind);
  }
}

}  // namespace

// MoleculeTransfer Prep -> Transform -> CalcEn -> Accept
// (MoleculeTransfer.h:85-246), GCMC, with PRNG::PickMolAndBoxPair2
// (PRNG.h:450-455) for the box pair, kind and molecule.
int System::RunMoleculeTransfer(double subDraw, MoveRecord *rec) {
  const int K = int(kinds.size());
  const double movPerc = moveW[kMolTransfer];
  double boxDiv = movPerc / 2;
  uint32_t src = uint32_t(subDraw / boxDiv);
  if (src >= 2) src--;
  subDraw -= src * boxDiv;
  const int dst = src == 0 ? 1 : 0;
  uint32_t mkTot = uint32_t(canSwapKind.size());
  double molDiv = boxDiv / mkTot;
  uint32_t k = uint32_t(subDraw / molDiv);
  if (k == mkTot) k = mkTot - 1;
  const int kind = canSwapKind[k];
  if (rec) {
    rec->kind = kMolTransfer;
    rec->mol = -1;
    rec->accepted = 0;
    rec->overlap = 0;
    rec->dLJ = rec->dReal = rec->dRecip = 0.0;
  }
  if (NumKindInBox(kind, int(src)) == 0) {
    if (src == 1) return -2;  // the reference stops: reservoir exhausted
    ms.swapTries[dst * K + kind]++;
    return 0;
  }
  uint32_t mOff = prng.randIntExc(NumKindInBox(kind, int(src)));
  const int m = GetMolNum(mOff, kind, int(src));
  const int st = molStart[m], len = molLen[m];
  // Prep + Transform.
  TrialMol newMol, oldMol;
  newMol.box = dst;
  oldMol.box = int(src);
  newMol.pos.assign(len, XYZ());
  oldMol.pos.resize(len);
  for (int p = 0; p < len; ++p) oldMol.pos[p] = Pos(st + p);
  if (src == 0)
    for (int p = st; p < st + len; ++p) cells.RemoveAtom(p, Pos(p));
  BuildPair(*this, kind, oldMol, newMol);
  const bool overlap = newMol.overlap;
  // CalcEn.
  double W_tc = 1.0, W_recip = 1.0, correct_old = 0, correct_new = 0, self_old = 0,
         self_new = 0, tcLose = 0, tcGain = 0, recipLose = 0, recipGain = 0;
  if (ff.useLRC) {
    tcLose = src == 0 ? MoleculeTailChange(kind, false) : 0.0;
    tcGain = dst == 0 ? MoleculeTailChange(kind, true) : 0.0;
    W_tc = std::exp(-1.0 * ff.beta * (tcGain + tcLose));
  }
  const bool tried = newMol.weight > kSmallWeight && !overlap;
  if (tried && ff.ewald) {
    correct_new = dst == 0 ? ewald.SwapCorrection(newMol.pos, kind) : 0.0;
    correct_old = src == 0 ? ewald.SwapCorrection(oldMol.pos, kind) : 0.0;
    self_new = dst == 0 ? ewald.SwapSelf(kind) : 0.0;
    self_old = src == 0 ? ewald.SwapSelf(kind) : 0.0;
    recipGain = dst == 0 ? ewald.SwapDestRecip(newMol.pos, kind) : 0.0;
    recipLose = src == 0 ? ewald.SwapSourceRecip(oldMol.pos, kind) : 0.0;
    W_recip = std::exp(-1.0 * ff.beta *
                       (recipGain + recipLose + correct_new - correct_old + self_new - self_old));
  }
  // Accept: GetCoeff with fugacities (MoleculeTransfer.h:143-165).
  double coeff;
  if (src == 0)
    coeff = double(swapCount[0 * K + kind]) * box.volInv / (ff.beta * fugacity[kind]);
  else
    coeff = box.volume / double(swapCount[dst * K + kind] + (g_plant == 3 ? 0 : 1)) *
            (ff.beta * fugacity[kind]);
  const double Wrat = newMol.weight / oldMol.weight * W_tc * W_recip;
  bool result = false;
  if (tried) result = prng() < coeff * Wrat;
  if (result) {
    if (src == 0) {
      pot.tailCorrection += tcLose;
      pot.inter -= oldMol.inter;
      pot.real -= oldMol.real;
      pot.recip += recipLose;
      pot.correction -= correct_old;
      pot.self -= self_old;
    } else {
      pot.tailCorrection += tcGain;
      pot.inter += newMol.inter;
      pot.real += newMol.real;
      pot.recip += recipGain;
      pot.correction += correct_new;
      pot.self += self_new;
    }
    for (int p = 0; p < len; ++p) {
      x[st + p] = newMol.pos[p].x;
      y[st + p] = newMol.pos[p].y;
      z[st + p] = newMol.pos[p].z;
    }
    SetNewCOM(m);
    ShiftMolBox(m, int(src), dst, kind);
    molBox[m] = dst;
    if (dst == 0) {
      kindCount[kind]++;
      for (int p = st; p < st + len; ++p) cells.AddAtom(p, Pos(p));
    } else {
      kindCount[kind]--;
    }
    if (src == 0 && NumInBox(0) == 0) {
      pot = Energy();
    } else if (src == 0 && NumInBox(0) == 1) {
      pot.inter = 0;
      pot.real = 0;
    }
    if (recipLose != 0.0 || recipGain != 0.0) ewald.UpdateRecip();
    pot.Total();
    // MoveSettings::Update marks the destination box (MoveSettings.cpp:102-104).
    if (dst == 0) ms.singleMoveAccepted = true;
  } else if (src == 0) {
    for (int p = st; p < st + len; ++p) cells.AddAtom(p, Pos(p));
  }
  ms.swapTries[dst * K + kind]++;
  if (result) ms.swapAccepted[dst * K + kind]++;
  if (rec) {
    rec->mol = m;
    rec->accepted = result;
    rec->overlap = overlap;
    rec->dLJ = dst == 0 ? newMol.inter : -oldMol.inter;
    rec->dReal = dst == 0 ? newMol.real : -oldMol.real;
    rec->dRecip = recipGain + recipLose;
  }
  return result;
}

}  // namespace orc

// CPU oracle -- system setup, totals and the single-molecule move loop.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <algorithm>

#include "oracle.hpp"
#include "oracle_capi.h"

namespace orc {

namespace {
constexpr int kDisplace = 0, kRotate = 1, kMultiParticle = 2, kMolTransfer = 10;
constexpr int kVolTransfer = 9;  // NPT (MoveConst.h:52-54)
constexpr double kTinyAmount = 1.0e-7;      // MoveSettings.cpp:14
constexpr double kTargetAccept = 0.5;       // MoveSettings.cpp:13
constexpr uint64_t kPhiloxTag = 0x474F4D43ull;  // key word 1 of the batched streams

void Bound(double &v, double lo, double hi) {  // lib/NumLib.h:37-55
  if (v < lo) v = lo;
  if (v > hi) v = hi;
}
}  // namespace

void System::Build(const SystemDesc &d) {
  threads = d.threads > 0 ? d.threads : 1;
  ff.Init(d.n_atom_types, d.sigma, d.epsilon, d.n_rep, d.rcut, d.rcut_low,
          d.rcut_coulomb, d.tolerance, d.temperature, d.ewald != 0,
          d.electrostatic != 0, d.lrc != 0);
  // BoxDimensions.cpp:17-19: the box cutoff is max(Rcut, RcutCoulomb).
  box.Set(d.box[0], d.box[1], d.box[2], std::max(ff.rCut, ff.rCutCoulomb));
  kinds.clear();
  int off = 0;
  for (int k = 0; k < d.n_kinds; ++k) {
    MolKind mk;
    for (int a = 0; a < d.kind_natoms[k]; ++a) {
      mk.atomType.push_back(d.kind_atom_type[off + a]);
      mk.charge.push_back(d.kind_charge[off + a]);
    }
    off += d.kind_natoms[k];
    kinds.push_back(mk);
  }
  molKind.assign(d.mol_kind, d.mol_kind + d.n_mols);
  molStart.clear();
  molLen.clear();
  particleType.clear();
  particleMol.clear();
  particleCharge.clear();
  particleHasNoCharge.clear();
  kindCount.assign(kinds.size(), 0);
  ensemble = d.ensemble;
  molBox.assign(d.n_mols, 0);
  if (d.mol_box)
    for (int m = 0; m < d.n_mols; ++m) molBox[m] = d.mol_box[m] ? 1 : 0;
  int atom = 0;
  for (int m = 0; m < d.n_mols; ++m) {
    const MolKind &mk = kinds[molKind[m]];
    molStart.push_back(atom);
    molLen.push_back(mk.NumAtoms());
    if (molBox[m] == 0) kindCount[molKind[m]]++;
    for (int a = 0; a < mk.NumAtoms(); ++a) {
      particleType.push_back(mk.atomType[a]);
      particleMol.push_back(m);
      particleCharge.push_back(mk.charge[a]);
      particleHasNoCharge.push_back(std::abs(mk.charge[a]) < 0.000000001);
      ++atom;
    }
  }
  x.assign(d.x, d.x + atom);
  y.assign(d.y, d.y + atom);
  z.assign(d.z, d.z + atom);
  // Molecules.cpp:157-175: LRC pair table over molecule kinds.
  const int K = int(kinds.size());
  pairEnCorrections.assign(K * K, 0.0);
  pairVirCorrections.assign(K * K, 0.0);
  if (ff.useLRC) {
    for (int i = 0; i < K; ++i)
      for (int j = i; j < K; ++j) {
        for (int pI = 0; pI < kinds[i].NumAtoms(); ++pI)
          for (int pJ = 0; pJ < kinds[j].NumAtoms(); ++pJ) {
            pairEnCorrections[i * K + j] +=
                ff.EnergyLRC(kinds[i].atomType[pI], kinds[j].atomType[pJ]);
            pairVirCorrections[i * K + j] +=
                ff.VirialLRC(kinds[i].atomType[pI], kinds[j].atomType[pJ]);
          }
        pairEnCorrections[j * K + i] = pairEnCorrections[i * K + j];
        pairVirCorrections[j * K + i] = pairVirCorrections[i * K + j];
      }
  }
  // MoleculeLookup::Init (MoleculeLookup.cpp:24-119): [box][kind] lists in
  // molecule order; kinds in order of first appearance.
  {
    const int K = int(kinds.size());
    std::vector<std::vector<uint32_t>> idx(2 * K);
    canSwapKind.clear();
    canMoveKind.clear();
    swapCount.assign(2 * K, 0);
    for (int m = 0; m < d.n_mols; ++m) {
      int b = molBox[m], k = molKind[m];
      idx[b * K + k].push_back(uint32_t(m));
      if (std::find(canSwapKind.begin(), canSwapKind.end(), k) == canSwapKind.end())
        canSwapKind.push_back(k);
      if (std::find(canMoveKind.begin(), canMoveKind.end(), k) == canMoveKind.end())
        canMoveKind.push_back(k);
      swapCount[b * K + k]++;
    }
    lookup.clear();
    bkStart.assign(2 * K + 1, 0);
    for (int s = 0; s < 2 * K; ++s) {
      bkStart[s] = uint32_t(lookup.size());
      lookup.insert(lookup.end(), idx[s].begin(), idx[s].end());
    }
    bkStart[2 * K] = uint32_t(d.n_mols);
  }
  fugacity.assign(kinds.size(), 0.0);
  if (d.kind_fugacity)
    for (size_t k = 0; k < kinds.size(); ++k) fugacity[k] = d.kind_fugacity[k] * 0.007242971;
  if (d.cbmc_first > 0) cbmcFirst = d.cbmc_first;
  if (d.cbmc_nth > 0) cbmcNth = d.cbmc_nth;
  if (d.cbmc_ang > 0) cbmcAng = d.cbmc_ang;
  // System.cpp:250-259: COM, then cell list, then Ewald, then totals.
  CalcCOM();
  hedron.assign(kinds.size(), Hedron());
  for (size_t k = 0; k < kinds.size(); ++k) {
    if (kinds[k].NumAtoms() != 3) continue;
    for (int m = 0; m < NumMols(); ++m) {
      if (molKind[m] != int(k)) continue;
      int st = molStart[m];
      XYZ a = box.MinImage(Pos(st + 1) - Pos(st)), b = box.MinImage(Pos(st + 2) - Pos(st));
      XYZ c = Cross(a, b);
      hedron[k].bPrev = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
      hedron[k].bBonded = std::sqrt(b.x * b.x + b.y * b.y + b.z * b.z);
      hedron[k].theta = std::atan2(std::sqrt(c.x * c.x + c.y * c.y + c.z * c.z),
                                   a.x * b.x + a.y * b.y + a.z * b.z);
      break;
    }
  }
  cells.Resize(box, box.rCut);
  cells.Clear(atom);
  for (int m = 0; m < NumMols(); ++m)
    if (InBox0(m))
      for (int p = molStart[m]; p < molStart[m] + molLen[m]; ++p) cells.AddAtom(p, Pos(p));
  ewald.Init();
  pot = SystemTotal();
  // StaticVals.cpp:37-96: move percentages normalised to one, summed in
  // move-kind order.
  const double swapW = d.ensemble == 1 ? d.move_swap : 0.0;
  const double volW = d.ensemble == 2 ? d.move_vol : 0.0;
  double tot = d.move_disp + d.move_rot + d.move_mp + swapW + volW;
  movePerc[0] = d.move_disp / tot;
  movePerc[1] = d.move_rot / tot;
  movePerc[2] = d.move_mp / tot;
  for (double &w : moveW) w = 0.0;
  moveW[0] = movePerc[0];
  moveW[1] = movePerc[1];
  moveW[2] = movePerc[2];
  moveW[10] = swapW / tot;
  if (d.ensemble == 2) moveW[kVolTransfer] = volW / tot;
  pressure = d.pressure * 0.007242971;
  ms.swapTries.assign(2 * K, 0);
  ms.swapAccepted.assign(2 * K, 0);
  perAdjust = d.adj_steps > 0 ? uint64_t(d.adj_steps) : 1000;
  // MoveSettings.cpp:42-56: initial displacement L_min/4, rotation pi/4.
  double lmin = std::min(box.axis.x, std::min(box.axis.y, box.axis.z));
  for (int mv = 0; mv < 2; ++mv) {
    ms.scale[mv].assign(K, mv == kDisplace ? lmin * 0.25 : M_PI_4);
    ms.tries[mv].assign(K, 0);
    ms.accepted[mv].assign(K, 0);
    ms.tempTries[mv].assign(K, 0);
    ms.tempAccepted[mv].assign(K, 0);
  }
  mt.seed(d.seed);
  philoxSeed = d.seed;
}

// COM.h:41-62: geometric centre of the unwrapped molecule; the molecule's
// atoms are unwrapped around its first atom and wrapped back in place.
void System::CalcCOM() {
  comx.assign(NumMols(), 0.0);
  comy.assign(NumMols(), 0.0);
  comz.assign(NumMols(), 0.0);
  for (int m = 0; m < NumMols(); ++m) {
    int s = molStart[m], e = s + molLen[m];
    XYZ ref = Pos(s);
    for (int p = s; p < e; ++p) {
      XYZ u = box.Unwrap(Pos(p), ref);
      x[p] = u.x;
      y[p] = u.y;
      z[p] = u.z;
    }
    double cx = 0, cy = 0, cz = 0;
    for (int p = s; p < e; ++p) {
      cx += x[p];
      cy += y[p];
      cz += z[p];
    }
    for (int p = s; p < e; ++p) {
      XYZ w = box.Wrap(Pos(p));
      x[p] = w.x;
      y[p] = w.y;
      z[p] = w.z;
    }
    double inv = 1.0 / double(molLen[m]);
    XYZ c = box.Wrap(XYZ(cx * inv, cy * inv, cz * inv));
    comx[m] = c.x;
    comy[m] = c.y;
    comz[m] = c.z;
  }
}

// CalculateEnergy.cpp:83-138 + SystemInter (140-153) for one box.
Energy System::SystemTotal() {
  Energy e;
  BoxInter(e.inter, e.real);
  if (ff.useLRC) e.tailCorrection = EnergyCorrection();
  e.recip = ff.ewald ? ewald.BoxReciprocal(false) : 0.0;
  double correction = 0.0;
  if (ff.ewald)
    for (int m = 0; m < NumMols(); ++m)
      if (InBox0(m)) correction += ewald.MolCorrection(m);
  e.correction = correction;
  e.self = ff.ewald ? ewald.BoxSelf() : 0.0;
  e.Total();
  return e;
}

// MoveSettings.cpp:148-162, 205-258 for displace/rotate.
void System::AdjustMoves(uint64_t step) {
  if ((step + 1) % perAdjust != 0) return;
  double lmin = std::min(box.axis.x, std::min(box.axis.y, box.axis.z));
  for (int mv = 0; mv < 2; ++mv)
    for (size_t k = 0; k < kinds.size(); ++k) {
      if (ms.tempTries[mv][k] > 0) {
        double currentAccept = double(ms.tempAccepted[mv][k]) / double(ms.tempTries[mv][k]);
        double fract = currentAccept / kTargetAccept;
        if (fract > 0.0)
          ms.scale[mv][k] *= fract;
        else
          ms.scale[mv][k] *= 0.5;
      }
      if (mv == kDisplace)
        Bound(ms.scale[mv][k], 1.0e-10, (lmin * 0.5) - kTinyAmount);
      else
        Bound(ms.scale[mv][k], 1.0e-6, M_PI - kTinyAmount);
      ms.tempAccepted[mv][k] = 0;
      ms.tempTries[mv][k] = 0;
    }
  if (ensemble == 2) {  // MoveSettings::Adjust, VOL_TRANSFER (MoveSettings.cpp:241-259)
    if (ms.volTempTries > 0) {
      double currentAccept = double(ms.volTempAccepted) / double(ms.volTempTries);
      double fract = currentAccept / kTargetAccept;
      if (fract > 0.0)
        ms.volScale *= fract;
      else
        ms.volScale *= 0.5;
    }
    const double minVol = 8.0 * box.rCutSq * box.rCut + 0.001;  // BoxDimensions.cpp:19
    Bound(ms.volScale, 0.001, 0.34 * (box.volume - minVol));
    ms.volTempAccepted = 0;
    ms.volTempTries = 0;
  }
  // MoveSettings::AdjustMultiParticle (MoveSettings.cpp:164-200); the
  // MultiParticle target acceptance is mp::TARGET_ACCEPT_FRACT = 0.3
  // (MoveSettings.h:22), the tolerance 0.1 and both alphas 0.2.
  for (int t = 0; t < 2; ++t) {
    if (ms.mp_int_tries[t] == 0) continue;
    ms.mp_tries[t] += ms.mp_int_tries[t];
    ms.mp_accepted[t] += ms.mp_int_acc[t];
    double fractTotal = (double(ms.mp_accepted[t]) / double(ms.mp_tries[t])) / 0.3;
    double fractInterval = (double(ms.mp_int_acc[t]) / double(ms.mp_int_tries[t])) / 0.3;
    double &mx = (t == 0) ? ms.mp_t_max : ms.mp_r_max;
    if (fractInterval == 0.0)
      mx *= 0.5;
    else if (std::fabs(fractInterval - 0.3) > 0.1)
      mx *= ((1.0 - 0.2) * fractTotal + 0.2 * fractInterval);
    if (t == 0)
      Bound(mx, 0.001, (lmin * 0.5) - 0.001);
    else
      Bound(mx, 0.001, M_PI - 0.001);
    ms.mp_int_acc[t] = 0;
    ms.mp_int_tries[t] = 0;
  }
}

namespace {
// Coordinates.cpp:83-100.
void TranslateCoords(const System &s, int m, const XYZ &shift, std::vector<XYZ> &dest,
                     XYZ &newCOM) {
  int st = s.molStart[m], len = s.molLen[m];
  dest.resize(len);
  for (int p = 0; p < len; ++p) dest[p] = s.Pos(st + p);
  newCOM = XYZ(s.comx[m], s.comy[m], s.comz[m]);
  for (int p = 0; p < len; ++p) dest[p] = dest[p] + shift;
  newCOM = newCOM + shift;
  for (int p = 0; p < len; ++p) dest[p] = s.box.Wrap(dest[p]);
  newCOM = s.box.Wrap(newCOM);
}

// Coordinates.cpp:103-128.
void RotateCoords(const System &s, int m, double theta, const XYZ &axis,
                  std::vector<XYZ> &dest) {
  RotationMatrix mat = RotationMatrix::FromAxisAngle(theta, axis);
  XYZ center(s.comx[m], s.comy[m], s.comz[m]);
  int st = s.molStart[m], len = s.molLen[m];
  dest.resize(len);
  for (int p = 0; p < len; ++p) dest[p] = s.box.Unwrap(s.Pos(st + p), center);
  for (int p = 0; p < len; ++p) {
    dest[p] = dest[p] + (-center);
    dest[p] = mat.Apply(dest[p]);
    dest[p] = dest[p] + center;
  }
  for (int p = 0; p < len; ++p) dest[p] = s.box.Wrap(dest[p]);
}
}  // namespace

// Translate.h:82-138 / Rotation.h:74-130 after the trial coordinates are
// made: energies, Metropolis test, commit.  `pr` is the acceptance draw.
static int EvaluateAndAccept(System &s, int mv, int m, const std::vector<XYZ> &newPos,
                             const XYZ *newCOM, double pr, MoveRecord *rec) {
  int st = s.molStart[m], len = s.molLen[m];
  for (int p = st; p < st + len; ++p) s.cells.RemoveAtom(p, s.Pos(p));
  double lj = 0, real = 0, recip = 0;
  bool overlap = s.MoleculeInter(lj, real, newPos, m);
  if (!overlap && s.ff.ewald) recip = s.ewald.MolReciprocal(newPos, m);
  bool res = pr < std::exp(-s.ff.beta * (lj + real + recip));
  bool result = res && !overlap;
  if (result) {
    s.pot.inter += lj;
    s.pot.real += real;
    s.pot.recip += recip;
    for (int p = 0; p < len; ++p) {
      s.x[st + p] = newPos[p].x;
      s.y[st + p] = newPos[p].y;
      s.z[st + p] = newPos[p].z;
    }
    if (newCOM) {
      s.comx[m] = newCOM->x;
      s.comy[m] = newCOM->y;
      s.comz[m] = newCOM->z;
    }
    if (s.ff.ewald) s.ewald.UpdateRecip();
    s.pot.Total();
  }
  for (int p = st; p < st + len; ++p) s.cells.AddAtom(p, s.Pos(p));
  int k = s.molKind[m];
  s.ms.tries[mv][k]++;
  s.ms.tempTries[mv][k]++;
  if (result) {
    s.ms.tempAccepted[mv][k]++;
    s.ms.accepted[mv][k]++;
    s.ms.singleMoveAccepted = true;
  }
  if (rec) {
    rec->kind = mv;
    rec->mol = m;
    rec->accepted = result;
    rec->overlap = overlap;
    rec->dLJ = lj;
    rec->dReal = real;
    rec->dRecip = recip;
  }
  return result;
}

// System.cpp:262-302 + PRNG.h:267-299,457-462 + MoveBase.h:118-134 for the
// NVT displace/rotate moves, MT19937 draws in the reference's order.
int System::RunStepMT(uint64_t step, MoveRecord *rec) {
  (void)step;
  uint32_t kind = 0;
  double draw = 0;
  // MOVE_KINDS_TOTAL: 9 kinds in NVT, 13 in GCMC, 10 in NPT (MoveConst.h:31-56).
  prng.PickArbDist(kind, draw, moveW, 1.0, ensemble == 1 ? 13 : (ensemble == 2 ? 10 : 9));
  if (kind == kMultiParticle) return ensemble == 1 ? -1 : RunMultiParticle(step, rec);
  if (kind == kMolTransfer) return RunMoleculeTransfer(draw, rec);
  if (kind == kVolTransfer && ensemble == 2) return RunVolumeMove(rec);
  if (kind != kDisplace && kind != kRotate) return -1;
  // PickMolAndBox with one box: boxDiv = movPerc, subDraw unchanged.
  double subPerc = movePerc[kind];
  double boxDiv = subPerc / 1;
  uint32_t b = uint32_t(draw / boxDiv);
  if (b >= 1) b--;
  draw -= b * boxDiv;
  // PickMol (PRNG.h:267-299) in box 0 through the molecule lookup.
  uint32_t mkTot = uint32_t(canMoveKind.size());
  double molDiv = boxDiv / mkTot;
  uint32_t k = uint32_t(draw / molDiv);
  if (k == mkTot) k = mkTot - 1;
  int mk = canMoveKind[k];
  if (NumKindInBox(mk, 0) == 0) {
    // NO_MOL_OF_KIND_IN_BOX: Accept draws nothing and counts a failed try.
    ms.tries[kind][mk]++;
    ms.tempTries[kind][mk]++;
    if (rec) {
      rec->kind = int(kind);
      rec->mol = -1;
    }
    return 0;
  }
  uint32_t mOff = prng.randIntExc(NumKindInBox(mk, 0));
  int m = GetMolNum(mOff, mk, 0);
  int mv = int(kind);
  if (mv == kRotate && kinds[mk].NumAtoms() <= 1) mv = kDisplace;  // Rotate::Prep
  std::vector<XYZ> newPos;
  if (mv == kDisplace) {
    XYZ shift = prng.SymXYZ(ms.scale[kDisplace][mk]);
    XYZ newCOM;
    TranslateCoords(*this, m, shift, newPos, newCOM);
    double pr = prng();
    return EvaluateAndAccept(*this, kDisplace, m, newPos, &newCOM, pr, rec);
  }
  double theta = prng.Sym(ms.scale[kRotate][mk]);
  XYZ axis = prng.PickOnUnitSphere();
  RotateCoords(*this, m, theta, axis, newPos);
  double pr = prng();
  return EvaluateAndAccept(*this, kRotate, m, newPos, nullptr, pr, rec);
}


// CellList::GridAll (src/CellList.cpp:220-240): every molecule re-added in
// molecule order.
void System::GridAll() {
  cells.Clear(NumAtoms());
  for (int m = 0; m < NumMols(); ++m)
    if (InBox0(m))
      for (int p = molStart[m]; p < molStart[m] + molLen[m]; ++p) cells.AddAtom(p, Pos(p));
}

namespace {
constexpr double kMinForce = 1E-12, kMaxForce = 30;  // MultiParticle.h:18-19
constexpr double kMPLambda = 0.5;                     // MultiParticle.h:95

// MultiParticle::CalcRandomTransform (MultiParticle.h:546-563).
XYZ CalcRandomTransform(bool &inRange, const XYZ &lb, double max, uint32_t m,
                        const Random123Wrapper &r) {
  XYZ lbmax = lb * max;
  XYZ val;
  inRange = std::abs(lbmax.x) > kMinForce && std::abs(lbmax.x) < kMaxForce &&
            std::abs(lbmax.y) > kMinForce && std::abs(lbmax.y) < kMaxForce &&
            std::abs(lbmax.z) > kMinForce && std::abs(lbmax.z) < kMaxForce;
  if (inRange) {
    XYZ u = r.GetRandomCoords(m);
    val.x = std::log(std::exp(-1.0 * lbmax.x) + 2.0 * u.x * std::sinh(lbmax.x)) / lb.x;
    val.y = std::log(std::exp(-1.0 * lbmax.y) + 2.0 * u.y * std::sinh(lbmax.y)) / lb.y;
    val.z = std::log(std::exp(-1.0 * lbmax.z) + 2.0 * u.z * std::sinh(lbmax.z)) / lb.z;
  }
  return val;
}

// MultiParticle::CalculateWRatio (MultiParticle.h:445-462).
double CalculateWRatio(const XYZ &n, const XYZ &o, const XYZ &k, double max) {
  double w = 1.0;
  w *= n.x * std::exp(-n.x * k.x) / (2.0 * std::sinh(n.x * max));
  w /= o.x * std::exp(o.x * k.x) / (2.0 * std::sinh(o.x * max));
  w *= n.y * std::exp(-n.y * k.y) / (2.0 * std::sinh(n.y * max));
  w /= o.y * std::exp(o.y * k.y) / (2.0 * std::sinh(o.y * max));
  w *= n.z * std::exp(-n.z * k.z) / (2.0 * std::sinh(n.z * max));
  w /= o.z * std::exp(o.z * k.z) / (2.0 * std::sinh(o.z * max));
  return w;
}
}  // namespace

// MultiParticle Prep -> Transform -> CalcEn -> Accept for the one box
// (MultiParticle.h:164-235, 332-413, 415-544, 565-712), CPU branch.
int System::RunMultiParticle(uint64_t step, MoveRecord *rec) {
  const int nm = NumMols(), na = NumAtoms();
  bool allTranslate = true;  // MultiParticle.h:99-107
  for (const MolKind &k : kinds) allTranslate = allTranslate && k.NumAtoms() == 1;
  const int moveType = allTranslate ? 0 : int(prng.randIntExc(2));
  Random123Wrapper r;  // System.cpp:262-269: key {step, seed}
  r.SetStep(step);
  r.SetRandomSeed(philoxSeed);
  // Prep: forces of the current state when a single-molecule move was
  // accepted since the last MultiParticle move.
  if (ms.singleMoveAccepted) {
    if (ff.ewald) {
      ewald.CopyRecip();
      ewald.BoxForceReciprocal(x, y, z, mpRef.aFR, mpRef.mFR);
    } else {
      mpRef.aFR.assign(na, XYZ());
      mpRef.mFR.assign(nm, XYZ());
    }
    double lj = 0, real = 0;
    BoxForce(lj, real, mpRef.aF, mpRef.mF);
    CalculateTorque(mpRef.aF, mpRef.aFR, mpRef.torque);
    pot.Total();
  }
  // Transform: CalculateTrialDistRot (MultiParticle.h:565-611).
  const double tmax = ms.mp_t_max, rmax = ms.mp_r_max;
  std::vector<double> nx = x, ny = y, nz = z, ncx = comx, ncy = comy, ncz = comz;
  std::vector<XYZ> tk(nm), rk(nm);
  std::vector<char> inRange(nm, 0);
  for (int m = 0; m < nm; ++m) {
    const int st = molStart[m], len = molLen[m];
    const XYZ com(comx[m], comy[m], comz[m]);
    bool in = false;
    if (moveType == 1) {
      XYZ lbt = mpRef.torque[m] * kMPLambda * ff.beta;
      XYZ val = CalcRandomTransform(in, lbt, rmax, uint32_t(m), r);
      rk[m] = val;
      RotationMatrix mat;
      if (in) {  // RotateForceBiased (613-644)
        double rotLen = std::sqrt(val.x * val.x + val.y * val.y + val.z * val.z);
        XYZ axis = val * (1.0 / rotLen);
        mat = RotationMatrix::FromAxisAngle(rotLen, axis);
      } else {  // RotateRandom (667-692)
        double symRand = r.GetSymRandom(uint32_t(m), rmax);
        XYZ sph = r.GetRandomCoordsOnSphere(uint32_t(m));
        mat = RotationMatrix::FromAxisAngle(symRand, sph);
      }
      for (int p = st; p < st + len; ++p) {
        XYZ t = box.Unwrap(Pos(p), com);
        t = t + (-com);
        t = mat.Apply(t);
        t = t + com;
        t = box.Wrap(t);
        nx[p] = t.x;
        ny[p] = t.y;
        nz[p] = t.z;
      }
    } else {
      XYZ lbf = (mpRef.mF[m] + mpRef.mFR[m]) * kMPLambda * ff.beta;
      XYZ val = CalcRandomTransform(in, lbf, tmax, uint32_t(m), r);
      tk[m] = val;
      // TranslateForceBiased (646-665) / TranslateRandom (694-712)
      XYZ shift = in ? val : r.GetSymRandomCoords(uint32_t(m), tmax);
      for (int p = st; p < st + len; ++p) {
        XYZ t = box.Wrap(Pos(p) + shift);
        nx[p] = t.x;
        ny[p] = t.y;
        nz[p] = t.z;
      }
      XYZ c = box.Wrap(com + shift);
      ncx[m] = c.x;
      ncy[m] = c.y;
      ncz[m] = c.z;
    }
    inRange[m] = in;
  }
  // Transform's checks (MultiParticle.h:366-406): a displacement beyond
  // half the box in any component, or a non-finite transform, fails.
  bool fail = false;
  for (int m = 0; m < nm && !fail; ++m) {
    XYZ num = (moveType == 1) ? rk[m] : tk[m];
    if (moveType == 0 && (num.x > box.halfAx.x || num.y > box.halfAx.y || num.z > box.halfAx.z))
      fail = true;
    else if (!std::isfinite(std::sqrt(num.x * num.x + num.y * num.y + num.z * num.z)))
      fail = true;
  }
  Energy potNew = pot;
  double wRatio = 1.0;
  if (!fail) {
    // CalcEn (415-443) on the trial coordinates.
    std::swap(x, nx);
    std::swap(y, ny);
    std::swap(z, nz);
    std::swap(comx, ncx);
    std::swap(comy, ncy);
    std::swap(comz, ncz);
    GridAll();
    if (ff.ewald) ewald.BoxReciprocalSums();
    BoxForce(potNew.inter, potNew.real, mpNew.aF, mpNew.mF);
    if (ff.ewald) {
      potNew.recip = ewald.BoxReciprocal(false);
      ewald.BoxForceReciprocal(x, y, z, mpNew.aFR, mpNew.mFR);
    } else {
      mpNew.aFR.assign(na, XYZ());
      mpNew.mFR.assign(nm, XYZ());
    }
    CalculateTorque(mpNew.aF, mpNew.aFR, mpNew.torque);
    // GetCoeff (464-511), molecules in index order.
    const double lBeta = kMPLambda * ff.beta;
    for (int m = 0; m < nm; ++m) {
      if (!inRange[m]) continue;
      const System::MPForces &fn = g_plant == 1 ? mpRef : mpNew, &fo = g_plant == 1 ? mpNew : mpRef;
      if (moveType == 1)
        wRatio *= CalculateWRatio(fn.torque[m] * lBeta, fo.torque[m] * lBeta, rk[m], rmax);
      else
        wRatio *= CalculateWRatio((fn.mF[m] + fn.mFR[m]) * lBeta,
                                  (fo.mF[m] + fo.mFR[m]) * lBeta, tk[m], tmax);
    }
    if (!std::isfinite(wRatio)) wRatio = 0.0;
  }
  // Accept (513-544).
  const double uBoltz = std::exp(-ff.beta * (potNew.Total() - pot.Total()));
  const double accept = wRatio * uBoltz;
  const double pr = r.GetRandomNumber(uint32_t(nm));
  const bool result = !fail && pr < accept;
  if (rec) {
    rec->kind = kMultiParticle;
    rec->mol = moveType;
    rec->accepted = result;
    rec->overlap = fail;
    rec->dLJ = fail ? 0.0 : potNew.inter - pot.inter;
    rec->dReal = fail ? 0.0 : potNew.real - pot.real;
    rec->dRecip = fail ? 0.0 : potNew.recip - pot.recip;
  }
  if (result) {
    pot = potNew;
    std::swap(mpRef, mpNew);
    if (ff.ewald) ewald.UpdateRecip();
  } else if (!fail) {
    std::swap(x, nx);
    std::swap(y, ny);
    std::swap(z, nz);
    std::swap(comx, ncx);
    std::swap(comy, ncy);
    std::swap(comz, ncz);
    GridAll();
  }
  ms.mp_int_tries[moveType]++;  // UpdateMoveSettingMultiParticle (202-208)
  if (result) ms.mp_int_acc[moveType]++;
  ms.singleMoveAccepted = false;  // MoveSettings::Update (MoveSettings.cpp:117-118)
  return result;
}

// The batched streams: one Philox4x64-10 stream per box, two blocks per move
// (counter {step, 0} and {step, 1}), key {seed, tag}.  The draws are used as
// the MT draws of RunStepMT are, with molecules picked uniformly.
int System::RunStepPhilox(uint64_t step, MoveRecord *rec) {
  uint64_t key[2] = {philoxSeed, kPhiloxTag};
  uint64_t c0[4] = {step, 0, 0, 0}, c1[4] = {step, 1, 0, 0}, r0[4], r1[4];
  Philox4x64_10(c0, key, r0);
  Philox4x64_10(c1, key, r1);
  double u[8];
  for (int i = 0; i < 4; ++i) {
    u[i] = U01(r0[i]);
    u[4 + i] = U01(r1[i]);
  }
  double pDisp = movePerc[kDisplace] / (movePerc[kDisplace] + movePerc[kRotate]);
  int mv = (u[0] < pDisp) ? kDisplace : kRotate;
  int nm = NumMols();
  int m = int(u[1] * nm);
  if (m >= nm) m = nm - 1;
  int mk = molKind[m];
  if (mv == kRotate && kinds[mk].NumAtoms() <= 1) mv = kDisplace;
  std::vector<XYZ> newPos;
  if (mv == kDisplace) {
    double b = ms.scale[kDisplace][mk], b2 = 2.0 * b;
    XYZ shift(u[2] * b2 - b, u[3] * b2 - b, u[4] * b2 - b);
    XYZ newCOM;
    TranslateCoords(*this, m, shift, newPos, newCOM);
    return EvaluateAndAccept(*this, kDisplace, m, newPos, &newCOM, u[5], rec);
  }
  double b = ms.scale[kRotate][mk];
  double theta = u[2] * (2.0 * b) - b;
  double cu = u[3] * 2.0 - 1.0;
  double phi = u[4] * (2.0 * M_PI);
  double root = std::sqrt(1.0 - cu * cu);
  XYZ axis(root * std::cos(phi), root * std::sin(phi), cu);
  RotateCoords(*this, m, theta, axis, newPos);
  return EvaluateAndAccept(*this, kRotate, m, newPos, nullptr, u[5], rec);
}

}  // namespace orc

namespace orc {

// VolumeTransfer for NPT (src/moves/VolumeTransfer.h:74-310) with one box:
// PickBox draws nothing; Transform draws delta = Sym(scale) and scales the
// box (BoxDimensions::ShiftVolume/SetVolume, BoxDimensions.cpp:101-123,
// 215-227) and every molecule's centre (Coordinates::TranslateOneBox,
// Coordinates.cpp:149-180); CalcEn regrids, rebuilds the k-vectors and the
// structure factors for the new box and recomputes BoxInter and the tail
// term; Accept uses (V'/V)^N exp(-beta P dV) exp(-beta dU).
int System::RunVolumeMove(MoveRecord *rec) {
  const double delta = prng.Sym(ms.volScale);
  const double oldVol = box.volume, newVol = box.volume + delta;
  BoxDims nb = box;
  const double ratio = std::cbrt(newVol / box.volume);
  nb.axis = {box.axis.x * ratio, box.axis.y * ratio, box.axis.z * ratio};
  nb.halfAx = {box.halfAx.x * ratio, box.halfAx.y * ratio, box.halfAx.z * ratio};
  nb.volume = newVol;
  nb.volInv = 1.0 / newVol;
  const double minVol = 8.0 * box.rCutSq * box.rCut + 0.001;
  if (nb.halfAx.x < box.rCut || nb.halfAx.y < box.rCut || nb.halfAx.z < box.rCut ||
      newVol < minVol)
    return -1;  // the reference stops: box shrunk below 2 Rcut
  const XYZ scale(nb.axis.x / box.axis.x, nb.axis.y / box.axis.y, nb.axis.z / box.axis.z);
  // keep the current state for a rejection
  const std::vector<double> ox = x, oy = y, oz = z, ocx = comx, ocy = comy, ocz = comz;
  const BoxDims ob = box;
  const std::vector<double> okx = ewald.kx, oky = ewald.ky, okz = ewald.kz, oh = ewald.hsqr,
                            op = ewald.prefact, osr = ewald.sumRnew, osi = ewald.sumInew;
  const int oImage = ewald.imageSize, oKmax = ewald.kmax;
  for (int m = 0; m < NumMols(); ++m) {
    if (!InBox0(m)) continue;
    const XYZ oldC(comx[m], comy[m], comz[m]);
    const XYZ newC(oldC.x * scale.x, oldC.y * scale.y, oldC.z * scale.z);
    const XYZ shift(newC.x - oldC.x, newC.y - oldC.y, newC.z - oldC.z);
    comx[m] = newC.x;
    comy[m] = newC.y;
    comz[m] = newC.z;
    for (int p = molStart[m]; p < molStart[m] + molLen[m]; ++p) {
      XYZ u = ob.Unwrap(Pos(p), oldC);
      u = {u.x + shift.x, u.y + shift.y, u.z + shift.z};
      const XYZ w = nb.Wrap(u);
      x[p] = w.x;
      y[p] = w.y;
      z[p] = w.z;
    }
  }
  box = nb;
  cells.Resize(box, box.rCut);
  GridAll();
  Energy potNew = pot;
  if (ff.ewald) {
    ewald.RecipInitOrth();
    ewald.BoxReciprocalSetup();
    potNew.recip = ewald.BoxReciprocal(true);
  }
  BoxInter(potNew.inter, potNew.real);
  if (ff.useLRC) potNew.tailCorrection = EnergyCorrection();
  potNew.Total();
  const double coeff = std::pow(newVol / oldVol, double(NumInBox(0))) *
                       std::exp(-ff.beta * pressure * (newVol - oldVol));
  const double uBoltz = std::exp(-ff.beta * (potNew.Total() - pot.Total()));
  const double accept = coeff * uBoltz;
  const bool result = prng() < accept;
  if (rec) {
    rec->kind = kVolTransfer;
    rec->mol = -1;
    rec->accepted = result;
    rec->overlap = 0;
    rec->dLJ = potNew.inter - pot.inter;
    rec->dReal = potNew.real - pot.real;
    rec->dRecip = potNew.recip - pot.recip;
  }
  if (result) {
    pot = potNew;
    if (ff.ewald) {  // UpdateRecip + UpdateRecipVec
      ewald.UpdateRecip();
      ewald.kxRef = ewald.kx;
      ewald.kyRef = ewald.ky;
      ewald.kzRef = ewald.kz;
      ewald.hsqrRef = ewald.hsqr;
      ewald.prefactRef = ewald.prefact;
      ewald.imageSizeRef = ewald.imageSize;
      // the reference sizes these arrays for the largest k count
      // (Ewald.cpp:166); here the trial arrays take the new count
      ewald.sumRnew.resize(ewald.imageSizeRef);
      ewald.sumInew.resize(ewald.imageSizeRef);
    }
    ms.singleMoveAccepted = true;
  } else {
    x = ox;
    y = oy;
    z = oz;
    comx = ocx;
    comy = ocy;
    comz = ocz;
    box = ob;
    cells.Resize(box, box.rCut);
    GridAll();
    ewald.kx = okx;
    ewald.ky = oky;
    ewald.kz = okz;
    ewald.hsqr = oh;
    ewald.prefact = op;
    ewald.sumRnew = osr;
    ewald.sumInew = osi;
    ewald.imageSize = oImage;
    ewald.kmax = oKmax;
  }
  ms.volTries++;
  ms.volTempTries++;
  if (result) {
    ms.volTempAccepted++;
    ms.volAccepted++;
  }
  return int(result);
}

}  // namespace orc

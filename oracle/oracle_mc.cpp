// CPU oracle -- system setup, totals and the single-molecule move loop.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <algorithm>

#include "oracle.hpp"
#include "oracle_capi.h"

namespace orc {

namespace {
constexpr int kDisplace = 0, kRotate = 1, kMultiParticle = 2;
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
  int atom = 0;
  for (int m = 0; m < d.n_mols; ++m) {
    const MolKind &mk = kinds[molKind[m]];
    molStart.push_back(atom);
    molLen.push_back(mk.NumAtoms());
    kindCount[molKind[m]]++;
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
  if (ff.useLRC) {
    for (int i = 0; i < K; ++i)
      for (int j = i; j < K; ++j) {
        for (int pI = 0; pI < kinds[i].NumAtoms(); ++pI)
          for (int pJ = 0; pJ < kinds[j].NumAtoms(); ++pJ)
            pairEnCorrections[i * K + j] +=
                ff.EnergyLRC(kinds[i].atomType[pI], kinds[j].atomType[pJ]);
        pairEnCorrections[j * K + i] = pairEnCorrections[i * K + j];
      }
  }
  // System.cpp:250-259: COM, then cell list, then Ewald, then totals.
  CalcCOM();
  cells.Resize(box, box.rCut);
  cells.Clear(atom);
  for (int m = 0; m < NumMols(); ++m)
    for (int p = molStart[m]; p < molStart[m] + molLen[m]; ++p) cells.AddAtom(p, Pos(p));
  ewald.Init();
  pot = SystemTotal();
  // StaticVals.cpp:37-96: move percentages normalised to one.
  double tot = d.move_disp + d.move_rot + d.move_mp;
  movePerc[0] = d.move_disp / tot;
  movePerc[1] = d.move_rot / tot;
  movePerc[2] = d.move_mp / tot;
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
    for (int m = 0; m < NumMols(); ++m) correction += ewald.MolCorrection(m);
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
  const double w[9] = {movePerc[0], movePerc[1], movePerc[2], 0, 0, 0, 0, 0, 0};
  uint32_t kind = 0;
  double draw = 0;
  prng.PickArbDist(kind, draw, w, 1.0, 9);
  if (kind != kDisplace && kind != kRotate) return -1;  // MP not in this loop
  // PickMolAndBox with one box: boxDiv = movPerc, subDraw unchanged.
  double subPerc = movePerc[kind];
  double boxDiv = subPerc / 1;
  uint32_t b = uint32_t(draw / boxDiv);
  if (b >= 1) b--;
  draw -= b * boxDiv;
  // PickMol: kinds in order of first appearance with molecules.
  std::vector<int> canMove;
  for (int m = 0; m < NumMols(); ++m)
    if (std::find(canMove.begin(), canMove.end(), molKind[m]) == canMove.end())
      canMove.push_back(molKind[m]);
  uint32_t mkTot = uint32_t(canMove.size());
  double molDiv = boxDiv / mkTot;
  uint32_t k = uint32_t(draw / molDiv);
  if (k == mkTot) k = mkTot - 1;
  int mk = canMove[k];
  if (kindCount[mk] == 0) {
    // NO_MOL_OF_KIND_IN_BOX: Accept draws nothing and counts a failed try.
    ms.tries[kind][mk]++;
    ms.tempTries[kind][mk]++;
    return 0;
  }
  uint32_t mOff = prng.randIntExc(uint32_t(kindCount[mk]));
  int m = -1, seen = 0;
  for (int i = 0; i < NumMols(); ++i)
    if (molKind[i] == mk && seen++ == int(mOff)) {
      m = i;
      break;
    }
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

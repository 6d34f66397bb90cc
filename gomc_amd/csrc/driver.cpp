// Single-box MC driver over the plain device path: the reference's move loop
// (Simulation::RunSimulation -> MoveSettings::AdjustMoves ->
// System::ChooseAndRunMove, src/Simulation.cpp:94-96, src/System.cpp:262-302)
// for Translate, Rotate and MultiParticle, with each single-molecule trial's
// energies from one fused launch (gomc_mol_trial), the MultiParticle move
// from the force/transform kernels, and the accepted state kept in HBM.
#include <algorithm>
#include <cmath>

#include "box.hpp"

using namespace gomc;

namespace {

constexpr int kDisplace = 0, kRotate = 1, kMultiParticle = 2, kVolTransfer = 9, kMolTransfer = 10;

// MoveSettings::AdjustMoves / Adjust (src/MoveSettings.cpp:148-162,205-258).
void AdjustMoves(gomc_box *b, uint64_t step) {
  if ((step + 1) % b->perAdjust != 0) return;
  const double lmin = std::min(b->geom.L[0], std::min(b->geom.L[1], b->geom.L[2]));
  for (int mv = 0; mv < 2; ++mv)
    for (int k = 0; k < b->nkinds; ++k) {
      double &sc = b->moves.scale[mv][k];
      if (b->moves.tempTries[mv][k] > 0) {
        const double acc = double(b->moves.tempAccepted[mv][k]) / double(b->moves.tempTries[mv][k]);
        const double fract = acc / 0.5;
        sc *= (fract > 0.0) ? fract : 0.5;
      }
      const double lo = (mv == kDisplace) ? 1.0e-10 : 1.0e-6;
      const double hi = (mv == kDisplace) ? (lmin * 0.5) - 1.0e-7 : M_PI - 1.0e-7;
      if (sc < lo) sc = lo;
      if (sc > hi) sc = hi;
      b->moves.tempAccepted[mv][k] = 0;
      b->moves.tempTries[mv][k] = 0;
    }
  if (b->ensemble == 2) {  // MoveSettings::Adjust, VOL_TRANSFER (MoveSettings.cpp:241-259)
    DriverMoves &mv = b->moves;
    if (mv.volTempTries > 0) {
      const double acc = double(mv.volTempAccepted) / double(mv.volTempTries);
      const double fract = acc / 0.5;
      mv.volScale *= (fract > 0.0) ? fract : 0.5;
    }
    const double rc = std::max(b->rCutLJ, b->rCutCoul);
    const double minVol = 8.0 * (rc * rc) * rc + 0.001;  // BoxDimensions.cpp:19
    const double hi = 0.34 * (b->volume - minVol);
    if (mv.volScale < 0.001) mv.volScale = 0.001;
    if (mv.volScale > hi) mv.volScale = hi;
    mv.volTempAccepted = 0;
    mv.volTempTries = 0;
  }
  // MoveSettings::AdjustMultiParticle (src/MoveSettings.cpp:164-200): the
  // MultiParticle target is mp::TARGET_ACCEPT_FRACT = 0.3 (MoveSettings.h:22),
  // the tolerance 0.1 and the averaging weight 0.2.
  MPState &mp = b->mp;
  for (int t = 0; t < 2; ++t) {
    if (mp.itries[t] == 0) continue;
    mp.tries[t] += mp.itries[t];
    mp.acc[t] += mp.iacc[t];
    const double total = (double(mp.acc[t]) / double(mp.tries[t])) / 0.3;
    const double interval = (double(mp.iacc[t]) / double(mp.itries[t])) / 0.3;
    double &mx = t == 0 ? mp.tMax : mp.rMax;
    if (interval == 0.0)
      mx *= 0.5;
    else if (std::fabs(interval - 0.3) > 0.1)
      mx *= ((1.0 - 0.2) * total + 0.2 * interval);
    const double lo = 0.001, hi = t == 0 ? (lmin * 0.5) - 0.001 : M_PI - 0.001;
    if (mx < lo) mx = lo;
    if (mx > hi) mx = hi;
    mp.iacc[t] = 0;
    mp.itries[t] = 0;
  }
}

struct Record {
  int kind = -1, mol = -1, acc = 0, ovl = 0;
  double dlj = 0, dreal = 0, drecip = 0;
};

// One move: PickArbDist -> PickMolAndBox -> TranslateRand/RotateRand ->
// CalcEn -> Accept, drawing from MT19937 exactly where the reference does.
int OneMove(gomc_box *b, uint64_t step, Record &r) {
  // PRNG::PickArbDist over the move kinds (src/PRNG.h:153-163): nine in
  // NVT, thirteen in GCMC (src/MoveConst.h:31-40).
  const double *w = b->moveW;
  const uint32_t nKinds = b->ensemble == 1 ? 13 : (b->ensemble == 2 ? 10 : 9);
  const double draw0 = b->mt.Closed() * 1.0;
  double sum = 0.0, prev = 0.0;
  uint32_t pick = 0;
  for (pick = 0; pick < nKinds && sum < draw0; ++pick) {
    prev = sum;
    sum += w[pick];
  }
  if (pick != 0) --pick;
  double subDraw = draw0 - prev;
  if (pick == kMultiParticle) {
    // MultiParticle::Prep picks the box without a draw (one box).
    int type = 0, acc = 0, failed = 0;
    double dE[3];
    int rc = MPMove(b, step, type, acc, failed, dE);
    if (rc) return rc;
    r.kind = kMultiParticle;
    r.mol = type;
    r.acc = acc;
    r.ovl = failed;
    r.dlj = dE[0];
    r.dreal = dE[1];
    r.drecip = dE[2];
    return 0;
  }
  if (pick == kVolTransfer && b->ensemble == 2) {
    int acc = 0;
    double dE[3];
    int rc = VolumeMove(b, acc, dE);
    if (rc) return rc;
    r.kind = kVolTransfer;
    r.mol = -1;
    r.acc = acc;
    r.ovl = 0;
    r.dlj = dE[0];
    r.dreal = dE[1];
    r.drecip = dE[2];
    return 0;
  }
  if (pick == kMolTransfer && b->ensemble == 1) {
    int m = -1, acc = 0, ovl = 0;
    double dE[3];
    int rc = TransferMove(b, subDraw, m, acc, ovl, dE);
    if (rc) return rc;
    r.kind = kMolTransfer;
    r.mol = m;
    r.acc = acc;
    r.ovl = ovl;
    r.dlj = dE[0];
    r.dreal = dE[1];
    r.drecip = dE[2];
    return 0;
  }
  if (pick != kDisplace && pick != kRotate) return Fail(b, GOMC_ERR_STATE, "move kind not served by this driver");
  // PRNG::PickMolAndBox with one box, then PickMol (src/PRNG.h:267-299).
  const double boxDiv = b->movePerc[pick] / 1;
  uint32_t box = uint32_t(subDraw / boxDiv);
  if (box >= 1) box--;
  subDraw -= box * boxDiv;
  // the molecule through the lookup of box 0 (PRNG.h:283-296)
  const uint32_t mkTot = uint32_t(b->canMoveKind.size());
  uint32_t k = uint32_t(subDraw / (boxDiv / mkTot));
  if (k == mkTot) k = mkTot - 1;
  const int mk = b->canMoveKind[k];
  int mv = int(pick);
  if (NumKindInBox(b, mk, 0) == 0) {
    b->moves.tries[mv][mk]++;
    b->moves.tempTries[mv][mk]++;
    r.kind = mv;
    return 0;
  }
  const uint32_t off = b->mt.Below(NumKindInBox(b, mk, 0) - 1);
  const int m = GetMolNum(b, off, mk, 0);
  if (mv == kRotate && b->kindType[mk].size() <= 1) mv = kDisplace;  // Rotate::Prep
  const int s = b->molStart[m], len = b->molLen[m];
  const double *L = b->geom.L, *H = b->geom.half;
  double nx[kMaxMolAtoms], ny[kMaxMolAtoms], nz[kMaxMolAtoms];
  double ncx = b->comx[m], ncy = b->comy[m], ncz = b->comz[m];
  if (mv == kDisplace) {
    // Coordinates::TranslateRand (src/Coordinates.cpp:83-100).
    const Vec3 sh = SymXYZ(b->mt, b->moves.scale[kDisplace][mk]);
    for (int p = 0; p < len; ++p) {
      nx[p] = Wrap(b->hx[s + p] + sh.x, L[0]);
      ny[p] = Wrap(b->hy[s + p] + sh.y, L[1]);
      nz[p] = Wrap(b->hz[s + p] + sh.z, L[2]);
    }
    ncx = Wrap(ncx + sh.x, L[0]);
    ncy = Wrap(ncy + sh.y, L[1]);
    ncz = Wrap(ncz + sh.z, L[2]);
  } else {
    // Coordinates::RotateRand (src/Coordinates.cpp:103-128).
    const double theta = Sym(b->mt, b->moves.scale[kRotate][mk]);
    const Vec3 axis = OnUnitSphere(b->mt);
    const Mat3 R = AxisAngle(theta, axis);
    const double cx = b->comx[m], cy = b->comy[m], cz = b->comz[m];
    for (int p = 0; p < len; ++p) {
      Vec3 v{Unwrap(b->hx[s + p], cx, L[0], H[0]) + (-cx), Unwrap(b->hy[s + p], cy, L[1], H[1]) + (-cy),
             Unwrap(b->hz[s + p], cz, L[2], H[2]) + (-cz)};
      v = R.Apply(v);
      nx[p] = Wrap(v.x + cx, L[0]);
      ny[p] = Wrap(v.y + cy, L[1]);
      nz[p] = Wrap(v.z + cz, L[2]);
    }
  }
  const double pr = b->mt.Closed();  // Translate/Rotate::Accept draw
  double out[4];
  int rc = LaunchTrial(b, m, &b->hx[s], &b->hy[s], &b->hz[s], nx, ny, nz, true, b->ewald, out);
  if (rc) return rc;
  const bool overlap = out[3] > 0.0;
  const double dlj = out[0], dreal = out[1];
  const double drecip = (!overlap && b->ewald) ? out[2] - b->pot.recip : 0.0;
  const bool res = pr < std::exp(-b->beta * (dlj + dreal + drecip));
  const bool result = res && !overlap;
  if (result) {
    b->pot.inter += dlj;
    b->pot.real += dreal;
    b->pot.recip += drecip;
    rc = gomc_commit_mol(b, m, nx, ny, nz);
    if (rc) return rc;
    if (mv == kDisplace) {
      b->comx[m] = ncx;
      b->comy[m] = ncy;
      b->comz[m] = ncz;
    }
    if (b->ewald) gomc_recip_accept(b);
    b->pot.Total();
    b->mp.singleAccepted = true;  // MoveSettings::Update (MoveSettings.cpp:102-104)
  }
  b->moves.tries[mv][mk]++;
  b->moves.tempTries[mv][mk]++;
  if (result) {
    b->moves.tempAccepted[mv][mk]++;
    b->moves.accepted[mv][mk]++;
  }
  r.kind = mv;
  r.mol = m;
  r.acc = result;
  r.ovl = overlap;
  r.dlj = dlj;
  r.dreal = dreal;
  r.drecip = drecip;
  return 0;
}

}  // namespace

extern "C" {

int64_t gomc_run_moves(gomc_box *b, int64_t start_step, int64_t nsteps, int32_t *kind,
                       int32_t *mol, int32_t *acc, int32_t *ovl, double *dlj, double *dreal,
                       double *drecip) {
  if (!b || nsteps < 0) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int64_t accepted = 0;
  for (int64_t i = 0; i < nsteps; ++i) {
    AdjustMoves(b, uint64_t(start_step + i));
    Record r;
    int rc = OneMove(b, uint64_t(start_step + i), r);
    if (rc) return rc;
    accepted += r.acc;
    if (kind) kind[i] = r.kind;
    if (mol) mol[i] = r.mol;
    if (acc) acc[i] = r.acc;
    if (ovl) ovl[i] = r.ovl;
    if (dlj) dlj[i] = r.dlj;
    if (dreal) dreal[i] = r.dreal;
    if (drecip) drecip[i] = r.drecip;
  }
  return accepted;
}

int gomc_energies(const gomc_box *b, double out[7]) {
  if (!b || !out) return GOMC_ERR_ARG;
  out[0] = b->pot.inter;
  out[1] = b->pot.tc;
  out[2] = b->pot.real;
  out[3] = b->pot.recip;
  out[4] = b->pot.self;
  out[5] = b->pot.corr;
  out[6] = b->pot.total;
  return GOMC_OK;
}

int gomc_get_scales(const gomc_box *b, double *disp, double *rot) {
  if (!b) return GOMC_ERR_ARG;
  for (int k = 0; k < b->nkinds; ++k) {
    if (disp) disp[k] = b->moves.scale[0][k];
    if (rot) rot[k] = b->moves.scale[1][k];
  }
  return GOMC_OK;
}

}  // extern "C"

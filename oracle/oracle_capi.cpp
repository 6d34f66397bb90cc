// C entry points of the CPU oracle.  TEST INFRASTRUCTURE ONLY (see
// oracle.hpp).
#include "oracle_capi.h"

#include <omp.h>

#include <chrono>
#include <cstring>
#include <vector>

#include "oracle.hpp"

using orc::System;
using orc::XYZ;

static System *S(void *h) { return static_cast<System *>(h); }

static void EnergyOut(const orc::Energy &e, double out[7]) {
  out[0] = e.inter;
  out[1] = e.tailCorrection;
  out[2] = e.real;
  out[3] = e.recip;
  out[4] = e.self;
  out[5] = e.correction;
  out[6] = e.total;
}

static std::vector<XYZ> Gather(int n, const double *x, const double *y, const double *z) {
  std::vector<XYZ> v(n);
  for (int i = 0; i < n; ++i) v[i] = XYZ(x[i], y[i], z[i]);
  return v;
}

extern "C" {

void *orc_create(const orc_system_desc *d) {
  System *s = new System();
  s->Build(*d);
  return s;
}

void orc_destroy(void *h) { delete S(h); }
void orc_plant(int which) { orc::g_plant = which; }
int orc_natoms(void *h) { return S(h)->NumAtoms(); }
int orc_nmols(void *h) { return S(h)->NumMols(); }
int orc_nkvec(void *h) { return S(h)->ewald.imageSizeRef; }

void orc_cutoffs(void *h, double out[6]) {
  System *s = S(h);
  out[0] = s->ff.alpha;
  out[1] = s->box.rCut;
  out[2] = s->ff.rCutLow;
  out[3] = s->ff.recip_rcut;
  out[4] = s->box.volume;
  out[5] = s->ff.beta;
}

// Simulation::RecalculateAndCheck (src/Simulation.cpp:170-172): refresh the
// k-vectors and structure factors from the coordinates, then SystemTotal.
void orc_system_total(void *h, double out[7]) {
  System *s = S(h);
  if (s->ff.ewald) s->ewald.Init();
  EnergyOut(s->SystemTotal(), out);
}
void orc_energies(void *h, double out[7]) { EnergyOut(S(h)->pot, out); }
void orc_box_inter(void *h, double *lj, double *real) { S(h)->BoxInter(*lj, *real); }
void orc_box_inter_brute(void *h, double *lj, double *real) { S(h)->BoxInterBrute(*lj, *real); }

int orc_mol_inter(void *h, int mol, const double *nx, const double *ny, const double *nz,
                  double *lj, double *real) {
  System *s = S(h);
  int st = s->molStart[mol], len = s->molLen[mol];
  for (int p = st; p < st + len; ++p) s->cells.RemoveAtom(p, s->Pos(p));
  bool ov = s->MoleculeInter(*lj, *real, Gather(len, nx, ny, nz), mol);
  for (int p = st; p < st + len; ++p) s->cells.AddAtom(p, s->Pos(p));
  return ov ? 1 : 0;
}

double orc_mol_recip(void *h, int mol, const double *nx, const double *ny, const double *nz,
                     double *sumRnew, double *sumInew) {
  System *s = S(h);
  double d = s->ewald.MolReciprocal(Gather(s->molLen[mol], nx, ny, nz), mol);
  if (sumRnew)
    std::memcpy(sumRnew, s->ewald.sumRnew.data(), sizeof(double) * s->ewald.imageSizeRef);
  if (sumInew)
    std::memcpy(sumInew, s->ewald.sumInew.data(), sizeof(double) * s->ewald.imageSizeRef);
  return d;
}

double orc_swap_recip(void *h, int kind, int insert, const double *x, const double *y,
                      const double *z) {
  System *s = S(h);
  auto pos = Gather(s->kinds[kind].NumAtoms(), x, y, z);
  return insert ? s->ewald.SwapDestRecip(pos, kind) : s->ewald.SwapSourceRecip(pos, kind);
}

void orc_particle_inter(void *h, int atom_type, double charge, int ntrials, const double *tx,
                        const double *ty, const double *tz, double *en, double *real,
                        int32_t *overlap) {
  System *s = S(h);
  bool *ov = new bool[ntrials];
  for (int t = 0; t < ntrials; ++t) {
    ov[t] = false;
    en[t] = 0.0;
    real[t] = 0.0;
  }
  s->ParticleInter(en, real, Gather(ntrials, tx, ty, tz), ov, atom_type, charge, -1);
  for (int t = 0; t < ntrials; ++t) overlap[t] = ov[t] ? 1 : 0;
  delete[] ov;
}

void orc_kvectors(void *h, double *kx, double *ky, double *kz, double *prefact) {
  System *s = S(h);
  int n = s->ewald.imageSizeRef;
  std::memcpy(kx, s->ewald.kxRef.data(), sizeof(double) * n);
  std::memcpy(ky, s->ewald.kyRef.data(), sizeof(double) * n);
  std::memcpy(kz, s->ewald.kzRef.data(), sizeof(double) * n);
  std::memcpy(prefact, s->ewald.prefactRef.data(), sizeof(double) * n);
}

void orc_recip_sums(void *h, double *sumR, double *sumI) {
  System *s = S(h);
  int n = s->ewald.imageSizeRef;
  std::memcpy(sumR, s->ewald.sumRref.data(), sizeof(double) * n);
  std::memcpy(sumI, s->ewald.sumIref.data(), sizeof(double) * n);
}

void orc_self_corr_lrc(void *h, double out[3]) {
  System *s = S(h);
  out[0] = s->ewald.BoxSelf();
  double c = 0.0;
  for (int m = 0; m < s->NumMols(); ++m) c += s->ewald.MolCorrection(m);
  out[1] = c;
  out[2] = s->EnergyCorrection();
}

double orc_tail_change(void *h, int kind, int add) { return S(h)->MoleculeTailChange(kind, add != 0); }

void orc_get_coords(void *h, double *x, double *y, double *z) {
  System *s = S(h);
  std::memcpy(x, s->x.data(), sizeof(double) * s->NumAtoms());
  std::memcpy(y, s->y.data(), sizeof(double) * s->NumAtoms());
  std::memcpy(z, s->z.data(), sizeof(double) * s->NumAtoms());
}

void orc_get_com(void *h, double *x, double *y, double *z) {
  System *s = S(h);
  std::memcpy(x, s->comx.data(), sizeof(double) * s->NumMols());
  std::memcpy(y, s->comy.data(), sizeof(double) * s->NumMols());
  std::memcpy(z, s->comz.data(), sizeof(double) * s->NumMols());
}

void orc_get_scales(void *h, double *disp, double *rot) {
  System *s = S(h);
  for (size_t k = 0; k < s->kinds.size(); ++k) {
    disp[k] = s->ms.scale[0][k];
    rot[k] = s->ms.scale[1][k];
  }
}

void orc_set_scales(void *h, const double *disp, const double *rot, double mp_t_max,
                    double mp_r_max) {
  System *s = S(h);
  for (size_t k = 0; k < s->kinds.size(); ++k) {
    s->ms.scale[0][k] = disp[k];
    s->ms.scale[1][k] = rot[k];
  }
  s->ms.mp_t_max = mp_t_max;
  s->ms.mp_r_max = mp_r_max;
}

int orc_cell_csr(void *h, int32_t *cellVector, int32_t *cellStart, int32_t *map,
                 int32_t edge[3]) {
  System *s = S(h);
  std::vector<int> cv, cs, mp;
  s->cells.GetCellListNeighbor(s->NumAtoms(), cv, cs, mp);
  if (cellVector) std::copy(cv.begin(), cv.end(), cellVector);
  if (cellStart) std::copy(cs.begin(), cs.end(), cellStart);
  if (map) std::copy(mp.begin(), mp.end(), map);
  for (int i = 0; i < 3; ++i) edge[i] = s->cells.edge[i];
  return s->cells.NumCells();
}

int orc_cells_exhaustive(void *h) { return S(h)->cells.IsExhaustive(S(h)->NumAtoms()) ? 1 : 0; }

// Simulation.cpp:94-96: AdjustMoves(step) then ChooseAndRunMove(step).
int64_t orc_run(void *h, int64_t start_step, int64_t nsteps, int mode, int32_t *kind,
                int32_t *mol, int32_t *acc, int32_t *ovl, double *dlj, double *dreal,
                double *drecip) {
  System *s = S(h);
  int64_t accepted = 0;
  for (int64_t i = 0; i < nsteps; ++i) {
    uint64_t step = uint64_t(start_step + i);
    s->AdjustMoves(step);
    orc::MoveRecord rec{-1, -1, 0, 0, 0, 0, 0};
    int r = mode == 0 ? s->RunStepMT(step, &rec) : s->RunStepPhilox(step, &rec);
    if (r <= -2) return r;  // the reference stops the run (empty reservoir)
    if (r > 0) accepted++;
    if (kind) kind[i] = rec.kind;
    if (mol) mol[i] = rec.mol;
    if (acc) acc[i] = rec.accepted;
    if (ovl) ovl[i] = rec.overlap;
    if (dlj) dlj[i] = rec.dLJ;
    if (dreal) dreal[i] = rec.dReal;
    if (drecip) drecip[i] = rec.dRecip;
  }
  return accepted;
}

void orc_box_force(void *h, double *lj, double *real, double *afx, double *afy, double *afz,
                   double *mfx, double *mfy, double *mfz) {
  System *s = S(h);
  std::vector<XYZ> aF, mF;
  s->BoxForce(*lj, *real, aF, mF);
  for (int i = 0; i < s->NumAtoms(); ++i) {
    afx[i] = aF[i].x;
    afy[i] = aF[i].y;
    afz[i] = aF[i].z;
  }
  for (int m = 0; m < s->NumMols(); ++m) {
    mfx[m] = mF[m].x;
    mfy[m] = mF[m].y;
    mfz[m] = mF[m].z;
  }
}

void orc_box_force_recip(void *h, double *afx, double *afy, double *afz, double *mfx,
                         double *mfy, double *mfz) {
  System *s = S(h);
  std::vector<XYZ> aF, mF;
  s->ewald.CopyRecip();
  s->ewald.BoxForceReciprocal(s->x, s->y, s->z, aF, mF);
  for (int i = 0; i < s->NumAtoms(); ++i) {
    afx[i] = aF[i].x;
    afy[i] = aF[i].y;
    afz[i] = aF[i].z;
  }
  for (int m = 0; m < s->NumMols(); ++m) {
    mfx[m] = mF[m].x;
    mfy[m] = mF[m].y;
    mfz[m] = mF[m].z;
  }
}

// CalculateEnergy::CalculateTorque (CalculateEnergy.cpp:1365-1406) from the
// pair forces and the reciprocal forces of the current coordinates.
void orc_torque(void *h, double *tx, double *ty, double *tz) {
  System *s = S(h);
  std::vector<XYZ> aF, mF, aFR, mFR, t;
  double lj = 0, real = 0;
  s->BoxForce(lj, real, aF, mF);
  if (s->ff.ewald) {
    s->ewald.CopyRecip();
    s->ewald.BoxForceReciprocal(s->x, s->y, s->z, aFR, mFR);
  } else {
    aFR.assign(s->NumAtoms(), XYZ());
  }
  s->CalculateTorque(aF, aFR, t);
  for (int m = 0; m < s->NumMols(); ++m) {
    tx[m] = t[m].x;
    ty[m] = t[m].y;
    tz[m] = t[m].z;
  }
}

// CalculateEnergy::VirialCalc and the pressure of OutputVars:
// out = {inter, tail, real, recip, total, pressure (bar), total tensor xx, yy, zz}.
void orc_virial(void *h, double *out) {
  System *s = S(h);
  const orc::VirialD v = s->VirialCalc();
  out[0] = v.inter;
  out[1] = v.tc;
  out[2] = v.real;
  out[3] = v.recip;
  out[4] = v.total;
  out[5] = s->Pressure(v);
  for (int d = 0; d < 3; ++d) out[6 + d] = v.totalT[d];
}

// ParallelTemperingUtilities::evaluateExchangeCriteria, neighbour branch
// (src/ParallelTemperingUtilities.cpp:108-240), over systems in one process:
// pairs (s-1, s) with s % 2 == attempt % 2, one draw each from MT19937
// seeded with `seed` after `drawsBefore` earlier draws; an accepted pair
// exchanges the temperature-bound state (beta and T, move scales and
// counters, the Philox key) of the two systems, which is the same as
// exchanging their configurations, and swaps the ladder entries.  Returns
// the draws made.
int orc_exchange(void *const *hs, int32_t *ladder, int n, int64_t attempt, uint32_t seed,
                 int64_t drawsBefore, int32_t *accepted) {
  orc::MTRand g(seed);
  for (int64_t i = 0; i < drawsBefore; ++i) (void)g.randInt();
  std::vector<double> beta(n), energy(n);
  for (int s = 0; s < n; ++s) {
    beta[s] = S(hs[ladder[s]])->ff.beta;
    energy[s] = S(hs[ladder[s]])->pot.total;
  }
  const int parity = int(attempt % 2);
  int draws = 0;
  for (int s = 1; s < n; ++s) {
    int ok = 0;
    if (s % 2 == parity) {
      const double uBoltz = std::exp((beta[s] - beta[s - 1]) * (energy[s] - energy[s - 1]));
      ok = g.rand() < uBoltz;
      ++draws;
    }
    accepted[s] = ok;
    if (!ok) continue;
    System *a = S(hs[ladder[s - 1]]), *b = S(hs[ladder[s]]);
    std::swap(a->ff.beta, b->ff.beta);
    std::swap(a->ff.T, b->ff.T);
    std::swap(a->philoxSeed, b->philoxSeed);
    for (int mv = 0; mv < 2; ++mv) {
      std::swap(a->ms.scale[mv], b->ms.scale[mv]);
      std::swap(a->ms.tries[mv], b->ms.tries[mv]);
      std::swap(a->ms.accepted[mv], b->ms.accepted[mv]);
      std::swap(a->ms.tempTries[mv], b->ms.tempTries[mv]);
      std::swap(a->ms.tempAccepted[mv], b->ms.tempAccepted[mv]);
    }
    std::swap(ladder[s - 1], ladder[s]);
  }
  accepted[0] = 0;
  return draws;
}

int orc_mol_boxes(void *h, int32_t *box) {
  System *s = S(h);
  for (int m = 0; m < s->NumMols(); ++m) box[m] = s->molBox[m];
  return int(s->NumInBox(0));
}

void orc_mp_settings(void *h, double out[2]) {
  out[0] = S(h)->ms.mp_t_max;
  out[1] = S(h)->ms.mp_r_max;
}

double orc_batch_run(int nboxes, void *const *handles, const int64_t *moves, int threads) {
  for (int b = 0; b < nboxes; ++b) S(handles[b])->threads = 1;
  auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int b = 0; b < nboxes; ++b) orc_run(handles[b], 0, moves[b], 1, nullptr, nullptr,
                                           nullptr, nullptr, nullptr, nullptr, nullptr);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

void orc_mt_raw(uint32_t seed, int n, uint32_t *out) {
  orc::MTRand g(seed);
  for (int i = 0; i < n; ++i) out[i] = g.randInt();
}

void orc_mt_rand_int(uint32_t seed, int n, uint32_t bound, uint32_t *out) {
  orc::MTRand g(seed);
  for (int i = 0; i < n; ++i) out[i] = g.randInt(bound);
}

void orc_philox(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  orc::Philox4x64_10(ctr, key, out);
}
double orc_u01(uint64_t v) { return orc::U01(v); }
double orc_uneg11(uint64_t v) { return orc::Uneg11(v); }

void orc_r123(uint64_t seed, uint64_t step, uint32_t counter, double bound, double out[8]) {
  orc::Random123Wrapper w;
  w.SetRandomSeed(seed);
  w.SetStep(step);
  out[0] = w.GetRandomNumber(counter);
  XYZ s = w.GetSymRandomCoords(counter, bound);
  XYZ p = w.GetRandomCoordsOnSphere(counter);
  out[1] = s.x;
  out[2] = s.y;
  out[3] = s.z;
  out[4] = p.x;
  out[5] = p.y;
  out[6] = p.z;
  out[7] = w.GetSymRandom(counter, 2.5);
}

void orc_rotation(double theta, double ax, double ay, double az, double m[9]) {
  orc::RotationMatrix r = orc::RotationMatrix::FromAxisAngle(theta, XYZ(ax, ay, az));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m[i * 3 + j] = r.m[i][j];
}

void orc_uniform_rotation(double u1, double u2, double u3, double m[9]) {
  orc::RotationMatrix r = orc::RotationMatrix::UniformRandom(u1, u2, u3);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) m[i * 3 + j] = r.m[i][j];
}

void orc_pair_terms(const orc_system_desc *d, int t1, int t2, double distSq, double qiqj,
                    double out[5]) {
  orc::ForceField ff;
  ff.Init(d->n_atom_types, d->sigma, d->epsilon, d->n_rep, d->rcut, d->rcut_low,
          d->rcut_coulomb, d->tolerance, d->temperature, d->ewald != 0, d->electrostatic != 0,
          d->lrc != 0);
  out[0] = ff.CalcEn(distSq, t1, t2);
  out[1] = ff.CalcCoulomb(distSq, qiqj);
  out[2] = ff.CalcVir(distSq, t1, t2);
  out[3] = ff.CalcCoulombVir(distSq, qiqj);
  out[4] = ff.EnergyLRC(t1, t2);
}

}  // extern "C"

// Creation of a gomc_box (force-field blending, cell grid, k-vectors, HBM
// allocation) and the whole-box totals.  Host code; kernels are in box.hip.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "box.hpp"

using namespace gomc;

#define HIP_TRY(b, call)                                      \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return gomc::HipFail(b, e_, #call); \
  } while (0)

namespace gomc {
int FlushPending(gomc_box *b);

// Coordinates must lie in the wrapped box [0, L) (the reference keeps them
// wrapped, BoxDimensions::WrapPBC).
static bool CoordsInBox(const gomc_box *b, const double *x, const double *y, const double *z) {
  for (int i = 0; i < b->natoms; ++i)
    if (!(x[i] >= 0.0 && x[i] < b->geom.L[0] && y[i] >= 0.0 && y[i] < b->geom.L[1] &&
          z[i] >= 0.0 && z[i] < b->geom.L[2]))
      return false;
  return true;
}

// Ewald::RecipInitOrth (src/Ewald.cpp:847-903): half-space integer triples
// in x, y, z loop order, kept while |k|^2 < recip_rcut^2.
void KVectorsOrth(const double L[3], double alphaSq, double recipRcut, double volume,
                  std::vector<double> &kx, std::vector<double> &ky, std::vector<double> &kz,
                  std::vector<double> &pref, std::vector<int32_t> *packed) {
  const double recipRcutSq = recipRcut * recipRcut;
  const double alpsqr4 = 1.0 / (4.0 * alphaSq);
  double c[3];
  for (int i = 0; i < 3; ++i) c[i] = (1.0 / L[i]) * (2.0 * M_PI);
  const double vol = volume / (4.0 * M_PI);
  int nmax[3];
  for (int i = 0; i < 3; ++i) nmax[i] = int(recipRcut * L[i] / (2.0 * M_PI)) + 1;
  kx.clear();
  ky.clear();
  kz.clear();
  pref.clear();
  if (packed) packed->clear();
  for (int ix = 0; ix <= nmax[0]; ix++) {
    const int ylo = (ix == 0) ? 0 : -nmax[1];
    for (int iy = ylo; iy <= nmax[1]; iy++) {
      const int zlo = (ix == 0 && iy == 0) ? 1 : -nmax[2];
      for (int iz = zlo; iz <= nmax[2]; iz++) {
        double kX = c[0] * ix, kY = c[1] * iy, kZ = c[2] * iz;
        double ksqr = kX * kX + kY * kY + kZ * kZ;
        if (ksqr < recipRcutSq) {
          kx.push_back(kX);
          ky.push_back(kY);
          kz.push_back(kZ);
          pref.push_back(kQQFact * std::exp(-ksqr * alpsqr4) / (ksqr * vol));
          if (packed) packed->push_back(int32_t(ix | ((iy + 128) << 8) | ((iz + 128) << 16)));
        }
      }
    }
  }
}

// Upload k-vectors, structure factors from the coordinates, and make them
// the reference sums (UpdateVectorsAndRecipTerms, Ewald.cpp:129-139).
int RecipInitImpl(gomc_box *b, double *erecip) {
  std::vector<double> kx, ky, kz, pref;
  KVectorsOrth(b->geom.L, b->alpha * b->alpha, b->recipRcut, b->volume, kx, ky, kz, pref, nullptr);
  const int K = int(kx.size());
  if (K != b->K) {
    double **arrs[] = {&b->kx, &b->ky, &b->kz, &b->pref, &b->sumRref, &b->sumIref, &b->sumRnew,
                       &b->sumInew};
    for (double **a : arrs) {
      if (*a) (void)hipFree(*a);
      *a = nullptr;
      HIP_TRY(b, hipMalloc(a, sizeof(double) * std::max(K, 1)));
    }
    b->K = K;
  }
  const size_t n = sizeof(double) * K;
  HIP_TRY(b, hipMemcpyAsync(b->kx, kx.data(), n, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->ky, ky.data(), n, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->kz, kz.data(), n, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->pref, pref.data(), n, hipMemcpyHostToDevice, b->stream));
  int rc = BuildKSegments(b);
  if (rc) return rc;
  b->recipReady = true;
  rc = RecipSums(b, erecip);
  if (rc) return rc;
  HIP_TRY(b, hipMemcpyAsync(b->sumRref, b->sumRnew, n, hipMemcpyDeviceToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->sumIref, b->sumInew, n, hipMemcpyDeviceToDevice, b->stream));
  HIP_TRY(b, hipStreamSynchronize(b->stream));
  return GOMC_OK;
}

// COM.h:41-62: geometric centre of each molecule, computed on its unwrapped
// atoms, which are written back wrapped (unwrap-then-wrap may move a
// boundary atom by an ulp, as in the reference).
static void CalcCOM(gomc_box *b) {
  b->comx.assign(b->nmols, 0.0);
  b->comy.assign(b->nmols, 0.0);
  b->comz.assign(b->nmols, 0.0);
  const double *L = b->geom.L, *H = b->geom.half;
  for (int m = 0; m < b->nmols; ++m) {
    const int s = b->molStart[m], e = s + b->molLen[m];
    const double rx = b->hx[s], ry = b->hy[s], rz = b->hz[s];
    double cx = 0, cy = 0, cz = 0;
    for (int p = s; p < e; ++p) {
      double ux = Unwrap(b->hx[p], rx, L[0], H[0]);
      double uy = Unwrap(b->hy[p], ry, L[1], H[1]);
      double uz = Unwrap(b->hz[p], rz, L[2], H[2]);
      cx += ux;
      cy += uy;
      cz += uz;
      b->hx[p] = Wrap(ux, L[0]);
      b->hy[p] = Wrap(uy, L[1]);
      b->hz[p] = Wrap(uz, L[2]);
    }
    const double inv = 1.0 / double(b->molLen[m]);
    b->comx[m] = Wrap(cx * inv, L[0]);
    b->comy[m] = Wrap(cy * inv, L[1]);
    b->comz[m] = Wrap(cz * inv, L[2]);
  }
}

}  // namespace gomc

extern "C" {

int gomc_box_create(const gomc_system_desc *d, int device, gomc_box **out) {
  if (!d || !out) return GOMC_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0)
    return GOMC_ERR_NODEV;
  gomc_box *b = new gomc_box();
  b->device = device;
  auto fail = [&](int code, const char *msg) {
    b->err = msg;
    gomc_box_destroy(b);
    return code;
  };
  if (d->n_mols <= 0 || d->n_kinds <= 0 || d->n_atom_types <= 0)
    return fail(GOMC_ERR_ARG, "empty system");
  // ---- force field: Forcefield::InitBasicVals + FFParticle::Blend ----
  b->ntypes = d->n_atom_types;
  b->rCutLJ = d->rcut;
  b->rCutCoul = d->rcut_coulomb;
  b->rCutLow = d->rcut_low;
  b->tol = d->tolerance;
  b->T = d->temperature;
  b->beta = 1 / b->T;
  b->ewald = d->ewald != 0;
  b->elect = d->electrostatic != 0;
  b->lrc = d->lrc != 0;
  b->alpha = std::sqrt(-std::log(b->tol)) / b->rCutCoul;
  b->recipRcut = -2.0 * std::log(b->tol) / b->rCutCoul;
  b->recipRcutSq = b->recipRcut * b->recipRcut;
  const int T2 = b->ntypes * b->ntypes;
  b->sigSq.assign(T2, 0);
  b->epsCn.assign(T2, 0);
  b->epsCn6.assign(T2, 0);
  b->nOver6.assign(T2, 0);
  b->nHalf.assign(T2, 0);
  b->nRep.assign(T2, 0);
  for (int i = 0; i < b->ntypes; ++i)
    for (int j = 0; j < b->ntypes; ++j) {
      const int idx = i + j * b->ntypes;
      const double n = (d->n_rep[i] + d->n_rep[j]) * 0.5;
      if (!(n > 6.0)) return fail(GOMC_ERR_ARG, "Mie exponent must exceed 6");
      const double cn = n / (n - 6.0) * std::pow(n / 6.0, (6.0 / (n - 6.0)));
      const double sig = (d->sigma[i] + d->sigma[j]) * 0.5;
      b->nRep[idx] = n;
      b->nHalf[idx] = n * 0.5;
      b->nOver6[idx] = n / 6;
      b->sigSq[idx] = sig * sig;
      b->epsCn[idx] = cn * std::sqrt(d->epsilon[i] * d->epsilon[j]);
      b->epsCn6[idx] = b->epsCn[idx] * 6;
    }
  // ---- geometry (BoxDimensions::Init, src/BoxDimensions.cpp:17-111) ----
  const double rCutBox = std::max(b->rCutLJ, b->rCutCoul);
  for (int i = 0; i < 3; ++i) {
    b->geom.L[i] = d->box[i];
    b->geom.half[i] = d->box[i] * 0.5;
    if (d->box[i] < 2.0 * rCutBox)
      return fail(GOMC_ERR_ARG, "Cutoff value is larger than half of minimum box length");
  }
  b->geom.rCutSq = rCutBox * rCutBox;
  b->geom.rCutLJSq = b->rCutLJ * b->rCutLJ;
  b->geom.rCutCoulSq = b->rCutCoul * b->rCutCoul;
  b->geom.rCutLowSq = b->rCutLow * b->rCutLow;
  b->geom.alpha = b->alpha;
  b->geom.ewald = b->ewald;
  b->geom.elect = b->elect;
  b->volume = d->box[0] * d->box[1] * d->box[2];
  b->volInv = 1.0 / b->volume;
  // ---- molecules ----
  b->nkinds = d->n_kinds;
  b->kindType.assign(b->nkinds, {});
  b->kindCharge.assign(b->nkinds, {});
  int off = 0;
  for (int k = 0; k < b->nkinds; ++k) {
    const int na = d->kind_natoms[k];
    if (na <= 0 || na > kMaxMolAtoms) return fail(GOMC_ERR_ARG, "molecule kind size out of range");
    for (int a = 0; a < na; ++a) {
      if (d->kind_atom_type[off + a] < 0 || d->kind_atom_type[off + a] >= b->ntypes)
        return fail(GOMC_ERR_ARG, "atom type out of range");
      b->kindType[k].push_back(d->kind_atom_type[off + a]);
      b->kindCharge[k].push_back(d->kind_charge[off + a]);
    }
    off += na;
    b->maxMolAtoms = std::max(b->maxMolAtoms, na);
  }
  b->nmols = d->n_mols;
  b->kindCount.assign(b->nkinds, 0);
  b->ensemble = d->ensemble;
  if (d->ensemble < 0 || d->ensemble > 2) return fail(GOMC_ERR_ARG, "ensemble must be 0, 1 or 2");
  b->molBox.assign(b->nmols, 0);
  if (b->ensemble == 1 && d->mol_box)
    for (int m = 0; m < b->nmols; ++m) {
      if (d->mol_box[m] != 0 && d->mol_box[m] != 1) return fail(GOMC_ERR_ARG, "mol_box must be 0 or 1");
      b->molBox[m] = d->mol_box[m];
    }
  for (int m = 0; m < b->nmols; ++m) {
    const int k = d->mol_kind[m];
    if (k < 0 || k >= b->nkinds) return fail(GOMC_ERR_ARG, "molecule kind out of range");
    b->molKind.push_back(k);
    b->molStart.push_back(b->natoms);
    b->molLen.push_back(int(b->kindType[k].size()));
    if (b->molBox[m] == 0) b->kindCount[k]++;
    for (size_t a = 0; a < b->kindType[k].size(); ++a) {
      b->atomType.push_back(b->kindType[k][a]);
      b->atomCharge.push_back(b->kindCharge[k][a]);
      b->atomMol.push_back(m);
    }
    b->natoms += int(b->kindType[k].size());
  }
  if (!CoordsInBox(b, d->x, d->y, d->z)) return fail(GOMC_ERR_ARG, "coordinates outside [0, L)");
  b->hx.assign(d->x, d->x + b->natoms);
  b->hy.assign(d->y, d->y + b->natoms);
  b->hz.assign(d->z, d->z + b->natoms);
  CalcCOM(b);
  // LRC table over molecule kinds (src/Molecules.cpp:157-175 with
  // FFParticle::EnergyLRC, src/FFParticle.cpp:96-116).
  b->pairEnCorr.assign(b->nkinds * b->nkinds, 0.0);
  b->pairVirCorr.assign(b->nkinds * b->nkinds, 0.0);
  if (b->lrc) {
    for (int i = 0; i < b->nkinds; ++i)
      for (int j = i; j < b->nkinds; ++j) {
        for (int ti : b->kindType[i])
          for (int tj : b->kindType[j]) {
            const int idx = ti + tj * b->ntypes;
            double tc = 1.0;
            const double sigma = std::sqrt(b->sigSq[idx]);
            const double rRat = sigma / b->rCutLJ;
            const double N = b->nRep[idx];
            tc *= sigma * b->sigSq[idx];
            tc *= 2.0 * M_PI * b->epsCn[idx] / (N - 3.0);
            tc *= (std::pow(rRat, b->nRep[idx] - 3.0) - ((N - 3.0) / 3.0) * rRat * rRat * rRat);
            b->pairEnCorr[i * b->nkinds + j] += tc;
            // FFParticle::VirialLRC (src/FFParticle.cpp:118-138)
            double tv = 1.0;
            tv *= sigma * b->sigSq[idx];
            tv *= 2.0 * M_PI * b->epsCn[idx] / (N - 3.0);
            tv *= (N * std::pow(rRat, b->nRep[idx] - 3.0) - (N - 3.0) * 2.0 * rRat * rRat * rRat);
            b->pairVirCorr[i * b->nkinds + j] += tv;
          }
        b->pairEnCorr[j * b->nkinds + i] = b->pairEnCorr[i * b->nkinds + j];
        b->pairVirCorr[j * b->nkinds + i] = b->pairVirCorr[i * b->nkinds + j];
      }
  }
  // ---- cell grid (CellList::ResizeGrid/RebuildNeighbors) ----
  for (int i = 0; i < 3; ++i) {
    b->edge[i] = std::max(int(std::floor(d->box[i] / rCutBox)), 3);
    b->cellSize[i] = d->box[i] / b->edge[i];
  }
  b->ncell = b->edge[0] * b->edge[1] * b->edge[2];
  std::vector<int> nbr(b->ncell * 27);
  const int *E = b->edge;
  for (int ix = 0; ix < E[0]; ++ix)
    for (int iy = 0; iy < E[1]; ++iy)
      for (int iz = 0; iz < E[2]; ++iz) {
        const int cell = ix * E[2] * E[1] + iy * E[2] + iz;
        int slot = 0;
        for (int dx = -1; dx <= 1; ++dx)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz)
              nbr[cell * 27 + slot++] = ((ix + dx + E[0]) % E[0]) * E[2] * E[1] +
                                        ((iy + dy + E[1]) % E[1]) * E[2] + ((iz + dz + E[2]) % E[2]);
      }
  // ---- device ----
  if (hipSetDevice(device) != hipSuccess) return fail(GOMC_ERR_NODEV, "hipSetDevice failed");
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(GOMC_ERR_HIP, "stream create failed");
  const size_t nd = sizeof(double) * b->natoms, ni = sizeof(int) * b->natoms;
  bool ok = true;
  ok &= hipMalloc(&b->x, nd) == hipSuccess;
  ok &= hipMalloc(&b->y, nd) == hipSuccess;
  ok &= hipMalloc(&b->z, nd) == hipSuccess;
  ok &= hipMalloc(&b->q, nd) == hipSuccess;
  ok &= hipMalloc(&b->type, ni) == hipSuccess;
  ok &= hipMalloc(&b->mol, ni) == hipSuccess;
  ok &= hipMalloc(&b->molStartDev, sizeof(int) * (b->nmols + 1)) == hipSuccess;
  ok &= hipMalloc(&b->tSigSq, sizeof(double) * T2) == hipSuccess;
  ok &= hipMalloc(&b->tEpsCn, sizeof(double) * T2) == hipSuccess;
  ok &= hipMalloc(&b->tNHalf, sizeof(double) * T2) == hipSuccess;
  ok &= hipMalloc(&b->tEpsCn6, sizeof(double) * T2) == hipSuccess;
  ok &= hipMalloc(&b->tNOver6, sizeof(double) * T2) == hipSuccess;
  ok &= hipMalloc(&b->cellCount, sizeof(int) * b->ncell) == hipSuccess;
  ok &= hipMalloc(&b->cellCursor, sizeof(int) * b->ncell) == hipSuccess;
  ok &= hipMalloc(&b->cellStart, sizeof(int) * (b->ncell + 1)) == hipSuccess;
  ok &= hipMalloc(&b->cellVec, ni) == hipSuccess;
  ok &= hipMalloc(&b->atomCell, ni) == hipSuccess;
  ok &= hipMalloc(&b->nbr, sizeof(int) * b->ncell * 27) == hipSuccess;
  ok &= hipMalloc(&b->counter, sizeof(unsigned)) == hipSuccess;
  ok &= hipHostMalloc(&b->hostRes, sizeof(double) * 8, hipHostMallocMapped) == hipSuccess;
  if (!ok) return fail(GOMC_ERR_HIP, "device allocation failed");
  if (hipHostGetDevicePointer((void **)&b->devRes, b->hostRes, 0) != hipSuccess)
    return fail(GOMC_ERR_HIP, "mapped result pointer");
  bool cp = true;
  cp &= hipMemcpy(b->x, b->hx.data(), nd, hipMemcpyHostToDevice) == hipSuccess;
  cp &= hipMemcpy(b->y, b->hy.data(), nd, hipMemcpyHostToDevice) == hipSuccess;
  cp &= hipMemcpy(b->z, b->hz.data(), nd, hipMemcpyHostToDevice) == hipSuccess;
  cp &= hipMemcpy(b->q, b->atomCharge.data(), nd, hipMemcpyHostToDevice) == hipSuccess;
  cp &= hipMemcpy(b->type, b->atomType.data(), ni, hipMemcpyHostToDevice) == hipSuccess;
  std::vector<int> liveMol(b->atomMol);
  for (int i = 0; i < b->natoms; ++i)
    if (b->molBox[b->atomMol[i]] != 0) liveMol[i] = -1;
  cp &= hipMemcpy(b->mol, liveMol.data(), ni, hipMemcpyHostToDevice) == hipSuccess;
  std::vector<int> msd(b->molStart);
  msd.push_back(b->natoms);
  cp &= hipMemcpy(b->molStartDev, msd.data(), sizeof(int) * msd.size(), hipMemcpyHostToDevice) ==
        hipSuccess;
  cp &= hipMemcpy(b->tSigSq, b->sigSq.data(), sizeof(double) * T2, hipMemcpyHostToDevice) == hipSuccess;
  cp &= hipMemcpy(b->tEpsCn, b->epsCn.data(), sizeof(double) * T2, hipMemcpyHostToDevice) == hipSuccess;
  cp &= hipMemcpy(b->tNHalf, b->nHalf.data(), sizeof(double) * T2, hipMemcpyHostToDevice) == hipSuccess;
  cp &= hipMemcpy(b->tEpsCn6, b->epsCn6.data(), sizeof(double) * T2, hipMemcpyHostToDevice) == hipSuccess;
  cp &= hipMemcpy(b->tNOver6, b->nOver6.data(), sizeof(double) * T2, hipMemcpyHostToDevice) == hipSuccess;
  cp &= hipMemcpy(b->nbr, nbr.data(), sizeof(int) * b->ncell * 27, hipMemcpyHostToDevice) == hipSuccess;
  cp &= hipMemset(b->counter, 0, sizeof(unsigned)) == hipSuccess;
  if (!cp) return fail(GOMC_ERR_HIP, "device upload failed");
  // ---- Ewald and the running energies (System::Init, System.cpp:250-259) ----
  double erecip = 0.0;
  if (b->ewald) {
    int rc = RecipInitImpl(b, &erecip);
    if (rc) {
      std::string m = b->err;
      gomc_box_destroy(b);
      (void)m;
      return rc;
    }
  }
  double lj = 0, real = 0;
  int rc = BoxInter(b, &lj, &real);
  if (rc) {
    gomc_box_destroy(b);
    return rc;
  }
  double self = 0, corr = 0, lrc = 0;
  HostCorrectionSelfLRC(b, self, corr, lrc);
  b->pot.inter = lj;
  b->pot.real = real;
  b->pot.recip = erecip;
  b->pot.tc = b->lrc ? lrc : 0.0;
  b->pot.self = b->ewald ? self : 0.0;
  b->pot.corr = b->ewald ? corr : 0.0;
  b->pot.Total();
  // ---- driver: move mix (StaticVals.cpp:37-96), MoveSettings::Init ----
  const double swapW = b->ensemble == 1 ? d->move_swap : 0.0;
  const double volW = b->ensemble == 2 ? d->move_vol : 0.0;
  const double tot = d->move_disp + d->move_rot + d->move_mp + swapW + volW;
  if (!(tot > 0)) return fail(GOMC_ERR_ARG, "no moves enabled");
  if (b->ensemble == 1 && d->move_mp > 0)
    return fail(GOMC_ERR_ARG, "MultiParticle moves are not supported with GCMC");
  b->movePerc[0] = d->move_disp / tot;
  b->movePerc[1] = d->move_rot / tot;
  b->movePerc[2] = d->move_mp / tot;
  for (double &w : b->moveW) w = 0.0;
  b->moveW[0] = b->movePerc[0];
  b->moveW[1] = b->movePerc[1];
  b->moveW[2] = b->movePerc[2];
  b->moveW[10] = swapW / tot;
  if (b->ensemble == 2) {  // NPT: VOL_TRANSFER is move kind 9 (MoveConst.h:52-54)
    if (b->cached) return fail(GOMC_ERR_ARG, "volume moves are not supported with the molecule cache");
    b->moveW[9] = volW / tot;
  }
  b->pressure = d->pressure * 0.007242971;  // bar -> K / A^3 (ConfigSetup.cpp:387-389)
  // GCMC: fugacities (bar -> K/A^3, UnitConst.h:44), CBMC trial counts, the
  // molecule lookup and the growth geometry of three-site kinds.
  b->fugacity.assign(b->nkinds, 0.0);
  if (d->kind_fugacity)
    for (int k = 0; k < b->nkinds; ++k) b->fugacity[k] = d->kind_fugacity[k] * 0.007242971;
  if (d->cbmc_first > 0) b->cbmcFirst = d->cbmc_first;
  if (d->cbmc_nth > 0) b->cbmcNth = d->cbmc_nth;
  if (d->cbmc_ang > 0) b->cbmcAng = d->cbmc_ang;
  LookupInit(b);
  b->hedron.assign(b->nkinds, gomc_box::Hedron());
  for (int k = 0; k < b->nkinds; ++k) {
    const int na = int(b->kindType[k].size());
    if (b->ensemble == 1 && na != 1 && na != 3)
      return fail(GOMC_ERR_ARG, "GCMC kinds must have one or three atoms");
    if (b->ensemble == 1 && !(b->fugacity[k] > 0))
      return fail(GOMC_ERR_ARG, "GCMC needs a positive fugacity for every kind");
    if (na != 3) continue;
    for (int m = 0; m < b->nmols; ++m) {
      if (b->molKind[m] != k) continue;
      const int st = b->molStart[m];
      double a[3], c[3];
      for (int i = 0; i < 3; ++i) {
        const double *P = i == 0 ? b->hx.data() : i == 1 ? b->hy.data() : b->hz.data();
        double u = P[st + 1] - P[st], v = P[st + 2] - P[st];
        if (u > b->geom.half[i]) u -= b->geom.L[i];
        else if (u < -b->geom.half[i]) u += b->geom.L[i];
        if (v > b->geom.half[i]) v -= b->geom.L[i];
        else if (v < -b->geom.half[i]) v += b->geom.L[i];
        a[i] = u;
        c[i] = v;
      }
      const double cx = a[1] * c[2] - a[2] * c[1], cy = a[2] * c[0] - a[0] * c[2],
                   cz = a[0] * c[1] - a[1] * c[0];
      b->hedron[k].bPrev = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
      b->hedron[k].bBonded = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      b->hedron[k].theta = std::atan2(std::sqrt(cx * cx + cy * cy + cz * cz),
                                      a[0] * c[0] + a[1] * c[1] + a[2] * c[2]);
      break;
    }
  }
  b->swapTries.assign(2 * b->nkinds, 0);
  b->swapAcc.assign(2 * b->nkinds, 0);
  b->perAdjust = d->adj_steps > 0 ? uint64_t(d->adj_steps) : 1000;
  const double lmin = std::min(d->box[0], std::min(d->box[1], d->box[2]));
  for (int mv = 0; mv < 2; ++mv) {
    b->moves.scale[mv].assign(b->nkinds, mv == 0 ? lmin * 0.25 : M_PI_4);
    b->moves.tries[mv].assign(b->nkinds, 0);
    b->moves.accepted[mv].assign(b->nkinds, 0);
    b->moves.tempTries[mv].assign(b->nkinds, 0);
    b->moves.tempAccepted[mv].assign(b->nkinds, 0);
  }
  b->mt.Seed(uint32_t(d->seed));
  b->seed = d->seed;
  *out = b;
  return GOMC_OK;
}

int gomc_recip_init(gomc_box *b, double *erecip) {
  if (!b || !erecip) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = FlushPending(b);
  if (rc) return rc;
  return RecipInitImpl(b, erecip);
}

int gomc_system_total(gomc_box *b, double out[7]) {
  if (!b || !out) return GOMC_ERR_ARG;
  (void)hipSetDevice(b->device);
  int rc = FlushPending(b);
  if (rc) return rc;
  double erecip = 0.0;
  if (b->ewald) {
    rc = RecipInitImpl(b, &erecip);
    if (rc) return rc;
  }
  double lj = 0, real = 0;
  rc = BoxInter(b, &lj, &real);
  if (rc) return rc;
  double self = 0, corr = 0, lrc = 0;
  HostCorrectionSelfLRC(b, self, corr, lrc);
  Energies e;
  e.inter = lj;
  e.tc = b->lrc ? lrc : 0.0;
  e.real = real;
  e.recip = erecip;
  e.self = b->ewald ? self : 0.0;
  e.corr = b->ewald ? corr : 0.0;
  e.Total();
  out[0] = e.inter;
  out[1] = e.tc;
  out[2] = e.real;
  out[3] = e.recip;
  out[4] = e.self;
  out[5] = e.corr;
  out[6] = e.total;
  return GOMC_OK;
}

}  // extern "C"

// NPT volume move on the plain device path: VolumeTransfer with one box
// (src/moves/VolumeTransfer.h:74-310).  Transform scales the box
// (BoxDimensions::ShiftVolume/SetVolume, BoxDimensions.cpp:101-123,215-227)
// and every molecule's centre (Coordinates::TranslateOneBox,
// Coordinates.cpp:149-180); CalcEn regrids, builds the k-vectors and the
// structure factors of the new box on the device and recomputes BoxInter
// and the tail term; Accept weighs (V'/V)^N exp(-beta P dV) exp(-beta dU).
// A rejected move puts back the old geometry, coordinates, cell grid and
// k-space arrays (kept aside, not recomputed).
#include <algorithm>
#include <cmath>

#include "box.hpp"

namespace gomc {

namespace {

// The cell grid of the current geometry (CellList::ResizeGridBox,
// CellList.cpp:166-200, and the 27-cell table of RebuildNeighbors).
struct Grid {
  int edge[3], ncell;
  double cellSize[3];
  int *cellCount, *cellCursor, *cellStart, *nbr;
};

Grid SaveGrid(const gomc_box *b) {
  Grid g;
  for (int i = 0; i < 3; ++i) {
    g.edge[i] = b->edge[i];
    g.cellSize[i] = b->cellSize[i];
  }
  g.ncell = b->ncell;
  g.cellCount = b->cellCount;
  g.cellCursor = b->cellCursor;
  g.cellStart = b->cellStart;
  g.nbr = b->nbr;
  return g;
}

void PutGrid(gomc_box *b, const Grid &g) {
  for (int i = 0; i < 3; ++i) {
    b->edge[i] = g.edge[i];
    b->cellSize[i] = g.cellSize[i];
  }
  b->ncell = g.ncell;
  b->cellCount = g.cellCount;
  b->cellCursor = g.cellCursor;
  b->cellStart = g.cellStart;
  b->nbr = g.nbr;
}

void FreeGridArrays(const Grid &g) {
  for (void *p : {(void *)g.cellCount, (void *)g.cellCursor, (void *)g.cellStart, (void *)g.nbr})
    if (p) (void)hipFree(p);
}

// New grid arrays for the box's current geometry (the old ones stay with
// the caller).
int NewGrid(gomc_box *b) {
  TrialCellsInvalidate(b);
  const double rCutBox = std::max(b->rCutLJ, b->rCutCoul);
  for (int i = 0; i < 3; ++i) {
    b->edge[i] = std::max(int(std::floor(b->geom.L[i] / rCutBox)), 3);
    b->cellSize[i] = b->geom.L[i] / b->edge[i];
  }
  b->ncell = b->edge[0] * b->edge[1] * b->edge[2];
  std::vector<int> nbr(size_t(b->ncell) * 27);
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
  b->cellCount = b->cellCursor = b->cellStart = b->nbr = nullptr;
  HIP_TRY(b, hipMalloc(&b->cellCount, sizeof(int) * b->ncell));
  HIP_TRY(b, hipMalloc(&b->cellCursor, sizeof(int) * b->ncell));
  HIP_TRY(b, hipMalloc(&b->cellStart, sizeof(int) * (b->ncell + 1)));
  HIP_TRY(b, hipMalloc(&b->nbr, sizeof(int) * b->ncell * 27));
  HIP_TRY(b, hipMemcpy(b->nbr, nbr.data(), sizeof(int) * nbr.size(), hipMemcpyHostToDevice));
  return GOMC_OK;
}

struct KArrays {
  int K;
  double *kx, *ky, *kz, *pref, *sumRref, *sumIref, *sumRnew, *sumInew;
  int *kseg, *nsegDev;  // the k list's segments (box.hip)
  int ksegCap, nseg;
};

KArrays TakeK(gomc_box *b) {
  KArrays k{b->K,       b->kx,       b->ky,      b->kz,      b->pref,    b->sumRref,
            b->sumIref, b->sumRnew,  b->sumInew, b->kseg,    b->nsegDev, b->ksegCap,
            b->nseg};
  b->K = 0;
  b->kx = b->ky = b->kz = b->pref = b->sumRref = b->sumIref = b->sumRnew = b->sumInew = nullptr;
  b->kseg = b->nsegDev = nullptr;
  b->ksegCap = b->nseg = 0;
  return k;
}

void FreeK(const KArrays &k) {
  for (double *p : {k.kx, k.ky, k.kz, k.pref, k.sumRref, k.sumIref, k.sumRnew, k.sumInew})
    if (p) (void)hipFree(p);
  for (int *p : {k.kseg, k.nsegDev})
    if (p) (void)hipFree(p);
}

void PutK(gomc_box *b, const KArrays &k) {
  b->kseg = k.kseg;
  b->nsegDev = k.nsegDev;
  b->ksegCap = k.ksegCap;
  b->nseg = k.nseg;
  b->K = k.K;
  b->kx = k.kx;
  b->ky = k.ky;
  b->kz = k.kz;
  b->pref = k.pref;
  b->sumRref = k.sumRref;
  b->sumIref = k.sumIref;
  b->sumRnew = k.sumRnew;
  b->sumInew = k.sumInew;
}

int UploadCoords(gomc_box *b) {
  const size_t bytes = sizeof(double) * b->natoms;
  HIP_TRY(b, hipMemcpyAsync(b->x, b->hx.data(), bytes, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->y, b->hy.data(), bytes, hipMemcpyHostToDevice, b->stream));
  HIP_TRY(b, hipMemcpyAsync(b->z, b->hz.data(), bytes, hipMemcpyHostToDevice, b->stream));
  return Sync(b);
}

}  // namespace

int Regrid(gomc_box *b) {
  const Grid old = SaveGrid(b);
  int rc = NewGrid(b);
  if (rc) {
    FreeGridArrays(SaveGrid(b));
    PutGrid(b, old);
    return rc;
  }
  FreeGridArrays(old);
  return GOMC_OK;
}

int VolumeMove(gomc_box *b, int &accepted, double dE[3]) {
  accepted = 0;
  dE[0] = dE[1] = dE[2] = 0.0;
  if (b->cached) return Fail(b, GOMC_ERR_STATE, "volume moves are not supported with the molecule cache");
  int rc = FlushPending(b);
  if (rc) return rc;
  TrialCellsInvalidate(b);  // coordinates scale, and the grid may change
  const double delta = Sym(b->mt, b->moves.volScale);
  const double oldVol = b->volume, newVol = b->volume + delta;
  const double rCutBox = std::max(b->rCutLJ, b->rCutCoul);
  const double ratio = std::cbrt(newVol / b->volume);
  BoxGeom ng = b->geom;
  for (int i = 0; i < 3; ++i) {
    ng.L[i] = b->geom.L[i] * ratio;
    ng.half[i] = b->geom.half[i] * ratio;
  }
  const double minVol = 8.0 * (rCutBox * rCutBox) * rCutBox + 0.001;  // BoxDimensions.cpp:19
  if (ng.half[0] < rCutBox || ng.half[1] < rCutBox || ng.half[2] < rCutBox || newVol < minVol)
    return Fail(b, GOMC_ERR_STATE, "box shrunk below 2 Rcut");
  double scale[3];
  for (int i = 0; i < 3; ++i) scale[i] = ng.L[i] / b->geom.L[i];
  // keep the current state for a rejection
  const BoxGeom og = b->geom;
  const double oldVolInv = b->volInv;
  const std::vector<double> ox = b->hx, oy = b->hy, oz = b->hz;
  const std::vector<double> ocx = b->comx, ocy = b->comy, ocz = b->comz;
  const Grid oldGrid = SaveGrid(b);
  // Coordinates::TranslateOneBox
  for (int m = 0; m < b->nmols; ++m) {
    if (!b->molBox.empty() && b->molBox[m] != 0) continue;
    const double oc[3] = {b->comx[m], b->comy[m], b->comz[m]};
    const double nc[3] = {oc[0] * scale[0], oc[1] * scale[1], oc[2] * scale[2]};
    const double sh[3] = {nc[0] - oc[0], nc[1] - oc[1], nc[2] - oc[2]};
    b->comx[m] = nc[0];
    b->comy[m] = nc[1];
    b->comz[m] = nc[2];
    for (int p = b->molStart[m]; p < b->molStart[m] + b->molLen[m]; ++p) {
      b->hx[p] = Wrap(Unwrap(b->hx[p], oc[0], og.L[0], og.half[0]) + sh[0], ng.L[0]);
      b->hy[p] = Wrap(Unwrap(b->hy[p], oc[1], og.L[1], og.half[1]) + sh[1], ng.L[1]);
      b->hz[p] = Wrap(Unwrap(b->hz[p], oc[2], og.L[2], og.half[2]) + sh[2], ng.L[2]);
    }
  }
  b->geom = ng;
  b->volume = newVol;
  b->volInv = 1.0 / newVol;
  bool gridReplaced = false, kTaken = false;
  KArrays oldK{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
               nullptr, nullptr, 0, 0};
  // Puts back the kept geometry, coordinates, grid and k-space arrays: the
  // rejection of the move, and the way out of a failure part-way through
  // its energy so that the box stays as it was before the move.
  auto putBack = [&]() -> int {
    b->geom = og;
    b->volume = oldVol;
    b->volInv = oldVolInv;
    b->hx = ox;
    b->hy = oy;
    b->hz = oz;
    b->comx = ocx;
    b->comy = ocy;
    b->comz = ocz;
    if (gridReplaced) {
      FreeGridArrays(SaveGrid(b));
      PutGrid(b, oldGrid);
    }
    if (kTaken) {
      FreeK(TakeK(b));
      PutK(b, oldK);
    }
    return UploadCoords(b);
  };
  rc = UploadCoords(b);
  if (!rc) {
    gridReplaced = true;  // NewGrid replaces the grid pointers even when it fails
    rc = NewGrid(b);
  }
  Energies potNew = b->pot;
  if (!rc && b->ewald) {
    oldK = TakeK(b);
    kTaken = true;
    rc = RecipInitImpl(b, &potNew.recip);
  }
  if (!rc) rc = BoxInter(b, &potNew.inter, &potNew.real);
  if (rc) {
    const std::string why = b->err;
    putBack();
    b->err = why;
    if (!why.empty()) b->err += " (volume move undone)";
    return rc;
  }
  if (b->lrc) {  // EnergyCorrection of the new volume
    double tc = 0.0;
    for (int i = 0; i < b->nkinds; ++i)
      for (int j = 0; j < b->nkinds; ++j)
        tc += b->pairEnCorr[i * b->nkinds + j] * unsigned(b->kindCount[i]) *
              unsigned(b->kindCount[j]) * b->volInv;
    potNew.tc = tc;
  }
  potNew.Total();
  // VolumeTransfer::GetCoeff / Accept (VolumeTransfer.h:191-245)
  const uint32_t nIn = b->bkStart[b->nkinds] - b->bkStart[0];
  const double coeff = std::pow(newVol / oldVol, double(nIn)) *
                       std::exp(-b->beta * b->pressure * (newVol - oldVol));
  const double uBoltz = std::exp(-b->beta * (potNew.Total() - b->pot.Total()));
  const double accept = coeff * uBoltz;
  const bool result = b->mt.Closed() < accept;
  dE[0] = potNew.inter - b->pot.inter;
  dE[1] = potNew.real - b->pot.real;
  dE[2] = potNew.recip - b->pot.recip;
  if (result) {
    b->pot = potNew;
    if (kTaken) FreeK(oldK);
    FreeGridArrays(oldGrid);
    DevFree(b);  // the device loop's k-vector tables belong to the old box
    b->mp.singleAccepted = true;
  } else {
    rc = putBack();
    if (rc) return rc;
  }
  b->moves.volTries++;
  b->moves.volTempTries++;
  if (result) {
    b->moves.volTempAccepted++;
    b->moves.volAccepted++;
  }
  accepted = result;
  return GOMC_OK;
}

}  // namespace gomc

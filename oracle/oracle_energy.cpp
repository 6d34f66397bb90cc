// CPU oracle -- force field, geometry, cell list, pair energies and Ewald.
// TEST INFRASTRUCTURE ONLY (see oracle.hpp).
#include <algorithm>
#include <cstring>

#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.hpp"

namespace orc {

// ------------------------------------------------------------ ForceField --
void ForceField::Init(int ntypes, const double *sigma, const double *eps,
                      const double *nrep, double rcut, double rcutLow,
                      double rcutCoul, double tol, double temperature,
                      bool ewaldOn, bool elect, bool lrc) {
  // src/Forcefield.cpp:37-83
  rCut = rcut;
  rCutSq = rCut * rCut;
  rCutLow = rcutLow;
  rCutLowSq = rCutLow * rCutLow;
  T = temperature;
  beta = 1 / T;
  ewald = ewaldOn;
  electrostatic = elect;
  useLRC = lrc;
  tolerance = tol;
  rCutCoulomb = rcutCoul;
  rCutCoulombSq = rCutCoulomb * rCutCoulomb;
  alpha = std::sqrt(-std::log(tolerance)) / rCutCoulomb;
  alphaSq = alpha * alpha;
  recip_rcut = -2.0 * std::log(tolerance) / rCutCoulomb;
  recip_rcut_Sq = recip_rcut * recip_rcut;
  // src/FFParticle.cpp:153-201 (Mie, arithmetic sigma, geometric epsilon).
  count = ntypes;
  int size = count * count;
  sigmaSq.assign(size, 0);
  epsilon.assign(size, 0);
  epsilon_cn.assign(size, 0);
  epsilon_cn_6.assign(size, 0);
  n.assign(size, 0);
  nOver6.assign(size, 0);
  for (int i = 0; i < count; ++i) {
    for (int j = 0; j < count; ++j) {
      int idx = Flat(i, j);
      n[idx] = (nrep[i] + nrep[j]) * 0.5;
      double cn = n[idx] / (n[idx] - 6.0) * std::pow(n[idx] / 6.0, (6.0 / (n[idx] - 6.0)));
      double sig = (sigma[i] + sigma[j]) * 0.5;
      sigmaSq[idx] = sig * sig;
      epsilon[idx] = std::sqrt(eps[i] * eps[j]);
      epsilon_cn[idx] = cn * epsilon[idx];
      epsilon_cn_6[idx] = epsilon_cn[idx] * 6;
      nOver6[idx] = n[idx] / 6;
    }
  }
}

// src/FFParticle.cpp:295-325 (lambda = 1 branch).
double ForceField::CalcEn(double distSq, int k1, int k2) const {
  if (rCutSq < distSq) return 0.0;
  int index = Flat(k1, k2);
  double rRat2 = sigmaSq[index] / distSq;
  double rRat4 = rRat2 * rRat2;
  double attract = rRat4 * rRat2;
  double repulse = std::pow(rRat2, (n[index] * 0.5));
  return epsilon_cn[index] * (repulse - attract);
}

// src/FFParticle.cpp:327-359.
double ForceField::CalcVir(double distSq, int k1, int k2) const {
  if (rCutSq < distSq) return 0.0;
  int index = Flat(k1, k2);
  double rNeg2 = 1.0 / distSq;
  double rRat2 = rNeg2 * sigmaSq[index];
  double rRat4 = rRat2 * rRat2;
  double attract = rRat4 * rRat2;
  double repulse = std::pow(rRat2, (n[index] * 0.5));
  return epsilon_cn_6[index] * (nOver6[index] * repulse - attract) * rNeg2;
}

// src/FFParticle.cpp:361-399.
double ForceField::CalcCoulomb(double distSq, double qiqj) const {
  if (rCutCoulombSq < distSq) return 0.0;
  double dist = std::sqrt(distSq);
  if (ewald) {
    double val = alpha * dist;
    return qiqj * std::erfc(val) / dist;
  }
  return qiqj / dist;
}

// src/FFParticle.cpp:401-446.
double ForceField::CalcCoulombVir(double distSq, double qiqj) const {
  if (rCutCoulombSq < distSq) return 0.0;
  double dist = std::sqrt(distSq);
  if (ewald) {
    double constValue = alpha * M_2_SQRTPI;
    double expConstValue = std::exp(-1.0 * alphaSq * distSq);
    double temp = 1.0 - std::erf(alpha * dist);
    return qiqj * (temp / dist + constValue * expConstValue) / distSq;
  }
  return qiqj / (distSq * dist);
}

// src/FFParticle.cpp:96-116.
double ForceField::EnergyLRC(int k1, int k2) const {
  int idx = Flat(k1, k2);
  double tc = 1.0;
  double sigma = std::sqrt(sigmaSq[idx]);
  double rRat = sigma / rCut;
  double N = n[idx];
  tc *= sigma * sigmaSq[idx];
  tc *= 2.0 * M_PI * epsilon_cn[idx] / (N - 3.0);
  tc *= (std::pow(rRat, n[idx] - 3.0) - ((N - 3.0) / 3.0) * rRat * rRat * rRat);
  return tc;
}

// src/FFParticle.cpp:118-138.
double ForceField::VirialLRC(int k1, int k2) const {
  int idx = Flat(k1, k2);
  double tc = 1.0;
  double sigma = std::sqrt(sigmaSq[idx]);
  double rRat = sigma / rCut;
  double N = n[idx];
  tc *= sigma * sigmaSq[idx];
  tc *= 2.0 * M_PI * epsilon_cn[idx] / (N - 3.0);
  tc *= (N * std::pow(rRat, n[idx] - 3.0) - (N - 3.0) * 2.0 * rRat * rRat * rRat);
  return tc;
}

// ----------------------------------------------------------------- BoxDims --
void BoxDims::Set(double lx, double ly, double lz, double rcut) {
  axis = {lx, ly, lz};
  halfAx = {lx * 0.5, ly * 0.5, lz * 0.5};
  volume = lx * ly * lz;
  volInv = 1.0 / volume;
  rCut = rcut;
  rCutSq = rCut * rCut;
}

// ---------------------------------------------------------------- CellList --
// src/CellList.cpp:138-218.
void CellList::Resize(const BoxDims &b, double cutoff) {
  edge[0] = std::max(int(std::floor(b.axis.x / cutoff)), 3);
  edge[1] = std::max(int(std::floor(b.axis.y / cutoff)), 3);
  edge[2] = std::max(int(std::floor(b.axis.z / cutoff)), 3);
  cellSize = {b.axis.x / edge[0], b.axis.y / edge[1], b.axis.z / edge[2]};
  int nCells = edge[0] * edge[1] * edge[2];
  head.assign(nCells, -1);
  neighbors.assign(nCells, {});
  for (int xx = 0; xx < edge[0]; ++xx)
    for (int yy = 0; yy < edge[1]; ++yy)
      for (int zz = 0; zz < edge[2]; ++zz) {
        int cell = xx * edge[2] * edge[1] + yy * edge[2] + zz;
        for (int dx = -1; dx <= 1; ++dx)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz)
              neighbors[cell].push_back(((xx + dx + edge[0]) % edge[0]) * edge[2] * edge[1] +
                                        ((yy + dy + edge[1]) % edge[1]) * edge[2] +
                                        ((zz + dz + edge[2]) % edge[2]));
      }
}

void CellList::Clear(int natoms) {
  list.assign(natoms, -1);
  std::fill(head.begin(), head.end(), -1);
}

// src/CellList.h:88-101.
int CellList::PositionToCell(const XYZ &p) const {
  int xx = int(p.x / cellSize.x);
  int yy = int(p.y / cellSize.y);
  int zz = int(p.z / cellSize.z);
  xx -= (xx == edge[0] ? 1 : 0);
  yy -= (yy == edge[1] ? 1 : 0);
  zz -= (zz == edge[2] ? 1 : 0);
  return xx * edge[1] * edge[2] + yy * edge[2] + zz;
}

// src/CellList.cpp:110-135 (one atom).
void CellList::AddAtom(int p, const XYZ &pos) {
  int cell = PositionToCell(pos);
  list[p] = head[cell];
  head[cell] = p;
}

// src/CellList.cpp:81-108 (one atom).
void CellList::RemoveAtom(int p, const XYZ &pos) {
  int cell = PositionToCell(pos);
  int at = head[cell];
  if (at == p) {
    head[cell] = list[p];
  } else {
    while (at != -1) {
      if (list[at] == p) {
        list[at] = list[p];
        break;
      }
      at = list[at];
    }
  }
}

// src/CellList.h:121-196: walk the 27 cells in table order, each cell's
// linked list from its head.
void CellList::NeighborsOfCell(int cell, std::vector<int> &out) const {
  out.clear();
  for (int nb : neighbors[cell])
    for (int at = head[nb]; at != -1; at = list[at]) out.push_back(at);
}

void CellList::Neighbors(const XYZ &pos, std::vector<int> &out) const {
  NeighborsOfCell(PositionToCell(pos), out);
}

// src/CellList.cpp:258-285.
void CellList::GetCellListNeighbor(int natoms, std::vector<int> &cellVector,
                                   std::vector<int> &cellStartIndex,
                                   std::vector<int> &mapParticleToCell) const {
  cellVector.assign(natoms, 0);
  cellStartIndex.assign(head.size(), 0);
  mapParticleToCell.assign(natoms, 0);
  int vi = 0;
  for (size_t cell = 0; cell < head.size(); ++cell) {
    cellStartIndex[cell] = vi;
    for (int p = head[cell]; p != -1; p = list[p]) {
      cellVector[vi] = p;
      mapParticleToCell[p] = int(cell);
      vi++;
    }
    std::sort(cellVector.begin() + cellStartIndex[cell], cellVector.begin() + vi);
  }
  cellStartIndex.push_back(vi);
  cellVector.resize(vi);
}

// src/CellList.cpp:66-79.
bool CellList::IsExhaustive(int natoms) const {
  std::vector<int> particles(list);
  particles.insert(particles.end(), head.begin(), head.end());
  particles.erase(std::remove(particles.begin(), particles.end(), -1), particles.end());
  std::sort(particles.begin(), particles.end());
  if (int(particles.size()) != natoms) return false;
  for (int i = 0; i < int(particles.size()); ++i)
    if (i != particles[i]) return false;
  return true;
}

// --------------------------------------------------------- CalculateEnergy --
// src/CalculateEnergy.cpp:157-266 (CPU branch): loop over the CSR form of
// the cell list, each unordered pair once (currParticle < nParticle), pairs
// within a molecule skipped.
void System::BoxInter(double &lj, double &real) {
  std::vector<int> cellVector, cellStart, map;
  cells.GetCellListNeighbor(NumAtoms(), cellVector, cellStart, map);
  double tempREn = 0.0, tempLJEn = 0.0;
  const int nv = int(cellVector.size());
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : tempREn, tempLJEn)
  for (int ci = 0; ci < nv; ci++) {
    int cur = cellVector[ci];
    int curCell = map[cur];
    for (int nc = 0; nc < kNeighborCells; nc++) {
      int nb = cells.neighbors[curCell][nc];
      int end = cellStart[nb + 1];
      for (int ni = cellStart[nb]; ni < end; ni++) {
        int n = cellVector[ni];
        if (cur < n && particleMol[cur] != particleMol[n]) {
          double distSq;
          XYZ d;
          if (box.InRcut(distSq, d, Pos(cur), Pos(n))) {
            if (ff.electrostatic) {
              double qiqj = particleCharge[cur] * particleCharge[n] * kQQFact;
              if (qiqj != 0.0) tempREn += ff.CalcCoulomb(distSq, qiqj);
            }
            tempLJEn += ff.CalcEn(distSq, particleType[cur], particleType[n]);
          }
        }
      }
    }
  }
  lj = tempLJEn;
  real = tempREn;
}

// Independent O(N^2) restatement of the same sum with no cell list: the
// check that the cell-list enumeration is exhaustive.
void System::BoxInterBrute(double &lj, double &real) const {
  double tempREn = 0.0, tempLJEn = 0.0;
  const int na = NumAtoms();
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64) reduction(+ : tempREn, tempLJEn)
  for (int i = 0; i < na; i++) {
    if (!InBox0(particleMol[i])) continue;
    for (int j = i + 1; j < na; j++) {
      if (particleMol[i] == particleMol[j] || !InBox0(particleMol[j])) continue;
      double distSq;
      XYZ d;
      if (box.InRcut(distSq, d, Pos(i), Pos(j))) {
        if (ff.electrostatic) {
          double qiqj = particleCharge[i] * particleCharge[j] * kQQFact;
          if (qiqj != 0.0) tempREn += ff.CalcCoulomb(distSq, qiqj);
        }
        tempLJEn += ff.CalcEn(distSq, particleType[i], particleType[j]);
      }
    }
  }
  lj = tempLJEn;
  real = tempREn;
}

// src/CalculateEnergy.cpp:581-686.  Precondition (Translate.h:84): the
// molecule has been removed from the cell list.
bool System::MoleculeInter(double &lj, double &real, const std::vector<XYZ> &newPos,
                           int mol) const {
  double tempREn = 0.0, tempLJEn = 0.0;
  bool overlap = false;
  int length = molLen[mol], start = molStart[mol];
  std::vector<int> nIndex;
  for (int p = 0; p < length; ++p) {
    int atom = start + p;
    cells.Neighbors(Pos(atom), nIndex);
    const int nn = int(nIndex.size());
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : tempREn, tempLJEn) if (nn > 4096)
    for (int i = 0; i < nn; i++) {
      double distSq = 0.0;
      XYZ d;
      if (box.InRcut(distSq, d, Pos(atom), Pos(nIndex[i]))) {
        if (ff.electrostatic) {
          double qiqj = particleCharge[atom] * particleCharge[nIndex[i]] * kQQFact;
          if (qiqj != 0.0) tempREn += -ff.CalcCoulomb(distSq, qiqj);
        }
        tempLJEn += -ff.CalcEn(distSq, particleType[atom], particleType[nIndex[i]]);
      }
    }
    cells.Neighbors(newPos[p], nIndex);
    const int nn2 = int(nIndex.size());
    bool ov = false;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : tempREn, tempLJEn) reduction(|| : ov) if (nn2 > 4096)
    for (int i = 0; i < nn2; i++) {
      double distSq = 0.0;
      XYZ d;
      if (box.InRcut(distSq, d, newPos[p], Pos(nIndex[i]))) {
        if (distSq < ff.rCutLowSq) ov = true;
        if (ff.electrostatic) {
          double qiqj = particleCharge[atom] * particleCharge[nIndex[i]] * kQQFact;
          if (qiqj != 0.0) tempREn += ff.CalcCoulomb(distSq, qiqj);
        }
        tempLJEn += ff.CalcEn(distSq, particleType[atom], particleType[nIndex[i]]);
      }
    }
    overlap = overlap || ov;
  }
  lj = tempLJEn;
  real = tempREn;
  return overlap;
}

// src/CalculateEnergy.cpp:727-785: trial positions of one atom against the
// atoms in the cell list (the molecule being grown is not in it).
void System::ParticleInter(double *en, double *real, const std::vector<XYZ> &trial,
                           bool *overlap, int atomType, double charge, int mol) const {
  (void)mol;
  std::vector<int> nIndex;
  for (size_t t = 0; t < trial.size(); ++t) {
    double tempReal = 0.0, tempLJ = 0.0;
    cells.Neighbors(trial[t], nIndex);
    bool ov = false;
    for (size_t i = 0; i < nIndex.size(); i++) {
      double distSq = 0.0;
      XYZ d;
      if (box.InRcut(distSq, d, trial[t], Pos(nIndex[i]))) {
        if (distSq < ff.rCutLowSq) ov = true;
        tempLJ += ff.CalcEn(distSq, atomType, particleType[nIndex[i]]);
        if (ff.electrostatic) {
          double qiqj = particleCharge[nIndex[i]] * charge * kQQFact;
          if (qiqj != 0.0) tempReal += ff.CalcCoulomb(distSq, qiqj);
        }
      }
    }
    overlap[t] = overlap[t] || ov;
    en[t] += tempLJ;
    real[t] += tempReal;
  }
}

// src/CalculateEnergy.cpp:1254-1272.
double System::EnergyCorrection() const {
  double en = 0.0;
  const int K = int(kinds.size());
  for (int i = 0; i < K; ++i) {
    unsigned numI = unsigned(kindCount[i]);
    for (int j = 0; j < K; ++j) {
      unsigned numJ = unsigned(kindCount[j]);
      en += pairEnCorrections[i * K + j] * numI * numJ * box.volInv;
    }
  }
  return en;
}

// src/CalculateEnergy.cpp:788-808.
double System::MoleculeTailChange(int kind, bool add) const {
  double delta = 0.0;
  const int K = int(kinds.size());
  double sign = (add ? 1.0 : -1.0);
  int mkIdxII = kind * K + kind;
  for (int j = 0; j < K; ++j) {
    int mkIdxIJ = j * K + kind;
    double rhoDeltaIJ_2 = sign * 2.0 * double(kindCount[j]) * box.volInv;
    delta += pairEnCorrections[mkIdxIJ] * rhoDeltaIJ_2;
  }
  delta += pairEnCorrections[mkIdxII] * box.volInv;
  return delta;
}

// src/CalculateEnergy.cpp:268-406 (CPU branch).
void System::BoxForce(double &lj, double &real, std::vector<XYZ> &aF,
                      std::vector<XYZ> &mF) {
  const int na = NumAtoms(), nm = NumMols();
  std::vector<int> cellVector, cellStart, map;
  cells.GetCellListNeighbor(na, cellVector, cellStart, map);
  const int nv = int(cellVector.size());
  // The reference reduces OpenMP thread copies of the force arrays
  // (CalculateEnergy.cpp:320-327); here each thread owns a contiguous
  // static slice and the copies are summed in thread order.
  const int nt = std::max(1, threads);
  std::vector<std::vector<XYZ>> tA(nt, std::vector<XYZ>(na)), tM(nt, std::vector<XYZ>(nm));
  std::vector<double> tR(nt, 0.0), tL(nt, 0.0);
#pragma omp parallel num_threads(nt)
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num(), T = omp_get_num_threads();
#else
    const int t = 0, T = 1;
#endif
    const int lo = int(int64_t(nv) * t / T), hi = int(int64_t(nv) * (t + 1) / T);
    std::vector<XYZ> &A = tA[t];
    std::vector<XYZ> &M = tM[t];
    double tempREn = 0.0, tempLJEn = 0.0;
    for (int ci = lo; ci < hi; ci++) {
      int cur = cellVector[ci];
      int curCell = map[cur];
      for (int nc = 0; nc < kNeighborCells; nc++) {
        int nb = cells.neighbors[curCell][nc];
        for (int ni = cellStart[nb]; ni < cellStart[nb + 1]; ni++) {
          int n = cellVector[ni];
          if (cur < n && particleMol[cur] != particleMol[n]) {
            double distSq;
            XYZ vc, fLJ, fReal;
            if (box.InRcut(distSq, vc, Pos(cur), Pos(n))) {
              if (ff.electrostatic) {
                double qiqj = particleCharge[cur] * particleCharge[n] * kQQFact;
                if (qiqj != 0.0) {
                  tempREn += ff.CalcCoulomb(distSq, qiqj);
                  fReal = vc * ff.CalcCoulombVir(distSq, qiqj);
                }
              }
              tempLJEn += ff.CalcEn(distSq, particleType[cur], particleType[n]);
              fLJ = vc * ff.CalcVir(distSq, particleType[cur], particleType[n]);
              XYZ f = {fLJ.x + fReal.x, fLJ.y + fReal.y, fLJ.z + fReal.z};
              A[cur] = A[cur] + f;
              A[n] = A[n] + (-f);
              M[particleMol[cur]] = M[particleMol[cur]] + f;
              M[particleMol[n]] = M[particleMol[n]] + (-f);
            }
          }
        }
      }
    }
    tR[t] = tempREn;
    tL[t] = tempLJEn;
  }
  aF.assign(na, XYZ());
  mF.assign(nm, XYZ());
  double tempREn = 0.0, tempLJEn = 0.0;
  for (int t = 0; t < nt; ++t) {
    for (int i = 0; i < na; ++i) aF[i] = aF[i] + tA[t][i];
    for (int m = 0; m < nm; ++m) mF[m] = mF[m] + tM[t][m];
    tempREn += tR[t];
    tempLJEn += tL[t];
  }
  lj = tempLJEn;
  real = tempREn;
}

// src/CalculateEnergy.cpp:1365-1406.
void System::CalculateTorque(const std::vector<XYZ> &aF, const std::vector<XYZ> &aFR,
                             std::vector<XYZ> &torque) const {
  torque.assign(NumMols(), XYZ());
  for (int m = 0; m < NumMols(); m++) {
    double tx = 0, ty = 0, tz = 0;
    XYZ com(comx[m], comy[m], comz[m]);
    for (int p = molStart[m]; p < molStart[m] + molLen[m]; p++) {
      XYZ d = box.MinImage(Pos(p) - com);
      XYZ t = Cross(d, aF[p] + aFR[p]);
      tx += t.x;
      ty += t.y;
      tz += t.z;
    }
    torque[m] = {tx, ty, tz};
  }
}

// ------------------------------------------------------------------- Ewald --
void Ewald::Init() {
  RecipInitOrth();
  BoxReciprocalSetup();
  SetRecipRef();
}

// src/Ewald.cpp:847-903.
void Ewald::RecipInitOrth() {
  const ForceField &ff = sys.ff;
  const BoxDims &b = sys.box;
  double alpsqr4 = 1.0 / (4.0 * ff.alphaSq);
  XYZ cv = b.axis;
  cv.x = 1.0 / cv.x;
  cv.y = 1.0 / cv.y;
  cv.z = 1.0 / cv.z;
  cv = {cv.x * (2.0 * M_PI), cv.y * (2.0 * M_PI), cv.z * (2.0 * M_PI)};
  double vol = b.volume / (4.0 * M_PI);
  int nkx_max = int(ff.recip_rcut * b.axis.x / (2.0 * M_PI)) + 1;
  int nky_max = int(ff.recip_rcut * b.axis.y / (2.0 * M_PI)) + 1;
  int nkz_max = int(ff.recip_rcut * b.axis.z / (2.0 * M_PI)) + 1;
  kmax = std::max(std::max(nkx_max, nky_max), std::max(nky_max, nkz_max));
  kx.clear();
  ky.clear();
  kz.clear();
  hsqr.clear();
  prefact.clear();
  for (int xx = 0; xx <= nkx_max; xx++) {
    int nky_min = (xx == 0) ? 0 : -nky_max;
    for (int yy = nky_min; yy <= nky_max; yy++) {
      int nkz_min = (xx == 0 && yy == 0) ? 1 : -nkz_max;
      for (int zz = nkz_min; zz <= nkz_max; zz++) {
        double kX = cv.x * xx, kY = cv.y * yy, kZ = cv.z * zz;
        double ksqr = kX * kX + kY * kY + kZ * kZ;
        if (ksqr < ff.recip_rcut_Sq) {
          kx.push_back(kX);
          ky.push_back(kY);
          kz.push_back(kZ);
          hsqr.push_back(ksqr);
          prefact.push_back(kQQFact * std::exp(-ksqr * alpsqr4) / (ksqr * vol));
        }
      }
    }
  }
  imageSize = int(kx.size());
  sumRnew.assign(imageSize, 0.0);
  sumInew.assign(imageSize, 0.0);
}

// src/Ewald.cpp:193-274 (CPU branch): molecules in box order, per k the sum
// over that molecule's charged atoms, then added into the box total.
void Ewald::BoxReciprocalSetup() {
  std::fill(sumRnew.begin(), sumRnew.end(), 0.0);
  std::fill(sumInew.begin(), sumInew.end(), 0.0);
  for (int m = 0; m < sys.NumMols(); ++m) {
    if (!sys.InBox0(m)) continue;
    const MolKind &kind = sys.kinds[sys.molKind[m]];
    const double lambdaCoef = 1.0;
    int start = sys.molStart[m];
#pragma omp parallel for num_threads(sys.threads) schedule(static) if (imageSize > 2048)
    for (int i = 0; i < imageSize; i++) {
      double sumReal = 0.0, sumImaginary = 0.0;
      for (int j = 0; j < kind.NumAtoms(); j++) {
        int a = start + j;
        if (sys.particleHasNoCharge[a]) continue;
        double dot = sys.x[a] * kx[i] + sys.y[a] * ky[i] + sys.z[a] * kz[i];
        sumReal += (kind.charge[j] * std::cos(dot));
        sumImaginary += (kind.charge[j] * std::sin(dot));
      }
      sumRnew[i] += (lambdaCoef * sumReal);
      sumInew[i] += (lambdaCoef * sumImaginary);
    }
  }
}

// src/Ewald.cpp:281-361: the same sums over the Ref k-vectors.
void Ewald::BoxReciprocalSums() {
  // The reference loops molecules outside and k inside, adding each
  // molecule's term to sumRnew[i]; looping k outside and molecules inside
  // performs the same additions in the same order for every k.
  const int nm = sys.NumMols();
#pragma omp parallel for num_threads(sys.threads) schedule(static)
  for (int i = 0; i < imageSizeRef; i++) {
    double R = 0.0, I = 0.0;
    for (int m = 0; m < nm; ++m) {
      if (!sys.InBox0(m)) continue;
      const MolKind &kind = sys.kinds[sys.molKind[m]];
      int start = sys.molStart[m];
      double sumReal = 0.0, sumImaginary = 0.0;
      for (int j = 0; j < kind.NumAtoms(); j++) {
        int a = start + j;
        if (sys.particleHasNoCharge[a]) continue;
        double dot = sys.x[a] * kxRef[i] + sys.y[a] * kyRef[i] + sys.z[a] * kzRef[i];
        sumReal += (kind.charge[j] * std::cos(dot));
        sumImaginary += (kind.charge[j] * std::sin(dot));
      }
      R += sumReal;
      I += sumImaginary;
    }
    sumRnew[i] = R;
    sumInew[i] = I;
  }
}

// src/Ewald.cpp:375-406.
double Ewald::BoxReciprocal(bool isNewVolume) const {
  const double *pf = isNewVolume ? prefact.data() : prefactRef.data();
  int n = isNewVolume ? imageSize : imageSizeRef;
  double e = 0.0;
#pragma omp parallel for num_threads(sys.threads) schedule(static) reduction(+ : e) if (n > 2048)
  for (int i = 0; i < n; i++)
    e += ((sumRnew[i] * sumRnew[i] + sumInew[i] * sumInew[i]) * pf[i]);
  return e;
}

// src/Ewald.cpp:409-473 (CPU branch).  Returns E_new - sysPotRef.recip.
double Ewald::MolReciprocal(const std::vector<XYZ> &newPos, int mol) {
  const MolKind &kind = sys.kinds[sys.molKind[mol]];
  int length = kind.NumAtoms(), start = sys.molStart[mol];
  const double lambdaCoef = 1.0;
  double energyRecipNew = 0.0;
#pragma omp parallel for num_threads(sys.threads) schedule(static) reduction(+ : energyRecipNew) if (imageSizeRef > 2048)
  for (int i = 0; i < imageSizeRef; i++) {
    double sRN = 0.0, sIN = 0.0, sRO = 0.0, sIO = 0.0;
    for (int p = 0; p < length; ++p) {
      int a = start + p;
      if (sys.particleHasNoCharge[a]) continue;
      double dotNew = newPos[p].x * kxRef[i] + newPos[p].y * kyRef[i] + newPos[p].z * kzRef[i];
      double dotOld = sys.x[a] * kxRef[i] + sys.y[a] * kyRef[i] + sys.z[a] * kzRef[i];
      sRN += (kind.charge[p] * std::cos(dotNew));
      sIN += (kind.charge[p] * std::sin(dotNew));
      sRO += (kind.charge[p] * std::cos(dotOld));
      sIO += (kind.charge[p] * std::sin(dotOld));
    }
    sumRnew[i] = sumRref[i] + lambdaCoef * (sRN - sRO);
    sumInew[i] = sumIref[i] + lambdaCoef * (sIN - sIO);
    energyRecipNew += (sumRnew[i] * sumRnew[i] + sumInew[i] * sumInew[i]) * prefactRef[i];
  }
  return energyRecipNew - sys.pot.recip;
}

// src/Ewald.cpp:478-531: insertion of a molecule of `kindIdx` at `pos`.
double Ewald::SwapDestRecip(const std::vector<XYZ> &pos, int kindIdx) {
  const MolKind &kind = sys.kinds[kindIdx];
  int length = kind.NumAtoms();
  double energyRecipNew = 0.0;
#pragma omp parallel for num_threads(sys.threads) schedule(static) reduction(+ : energyRecipNew) if (imageSizeRef > 2048)
  for (int i = 0; i < imageSizeRef; i++) {
    double sR = 0.0, sI = 0.0;
    for (int p = 0; p < length; ++p) {
      if (std::fabs(kind.charge[p]) < 0.000000001) continue;
      double dot = pos[p].x * kxRef[i] + pos[p].y * kyRef[i] + pos[p].z * kzRef[i];
      sR += (kind.charge[p] * std::cos(dot));
      sI += (kind.charge[p] * std::sin(dot));
    }
    sumRnew[i] = sumRref[i] + sR;
    sumInew[i] = sumIref[i] + sI;
    energyRecipNew += (sumRnew[i] * sumRnew[i] + sumInew[i] * sumInew[i]) * prefactRef[i];
  }
  return energyRecipNew - sys.pot.recip;
}

// src/Ewald.cpp:657-710: removal of a molecule at `pos`.
double Ewald::SwapSourceRecip(const std::vector<XYZ> &pos, int kindIdx) {
  const MolKind &kind = sys.kinds[kindIdx];
  int length = kind.NumAtoms();
  double energyRecipNew = 0.0;
#pragma omp parallel for num_threads(sys.threads) schedule(static) reduction(+ : energyRecipNew) if (imageSizeRef > 2048)
  for (int i = 0; i < imageSizeRef; i++) {
    double sR = 0.0, sI = 0.0;
    for (int p = 0; p < length; ++p) {
      if (std::fabs(kind.charge[p]) < 0.000000001) continue;
      double dot = pos[p].x * kxRef[i] + pos[p].y * kyRef[i] + pos[p].z * kzRef[i];
      sR += (kind.charge[p] * std::cos(dot));
      sI += (kind.charge[p] * std::sin(dot));
    }
    sumRnew[i] = sumRref[i] - sR;
    sumInew[i] = sumIref[i] - sI;
    energyRecipNew += (sumRnew[i] * sumRnew[i] + sumInew[i] * sumInew[i]) * prefactRef[i];
  }
  return energyRecipNew - sys.pot.recip;
}

// src/Ewald.cpp:1125-1163 (no fractional molecules).
double Ewald::BoxSelf() const {
  double self = 0.0;
  for (size_t i = 0; i < sys.kinds.size(); i++) {
    const MolKind &k = sys.kinds[i];
    unsigned molNum = unsigned(sys.kindCount[i]);
    double molSelf = 0.0;
    for (int j = 0; j < k.NumAtoms(); j++) molSelf += (k.charge[j] * k.charge[j]);
    self += (molSelf * molNum);
  }
  self *= -1.0 * sys.ff.alpha * kQQFact * M_2_SQRTPI * 0.5;
  return self;
}

// src/Ewald.cpp:1056-1085.
double Ewald::MolCorrection(int mol) const {
  const MolKind &k = sys.kinds[sys.molKind[mol]];
  int start = sys.molStart[mol];
  double correction = 0.0;
  for (int i = 0; i < k.NumAtoms(); i++) {
    if (sys.particleHasNoCharge[start + i]) continue;
    for (int j = i + 1; j < k.NumAtoms(); j++) {
      double distSq;
      XYZ d;
      sys.box.InRcut(distSq, d, sys.Pos(start + i), sys.Pos(start + j));
      double dist = std::sqrt(distSq);
      correction += (k.charge[i] * k.charge[j] * std::erf(sys.ff.alpha * dist) / dist);
    }
  }
  const double lambdaCoef = 1.0;
  return -1.0 * kQQFact * correction * lambdaCoef * lambdaCoef;
}

// src/Ewald.cpp:1311-1335.
double Ewald::SwapCorrection(const std::vector<XYZ> &pos, int kindIdx) const {
  const MolKind &k = sys.kinds[kindIdx];
  double correction = 0.0;
  for (int i = 0; i < k.NumAtoms(); i++)
    for (int j = i + 1; j < k.NumAtoms(); j++) {
      double distSq;
      XYZ d;
      sys.box.InRcut(distSq, d, pos[i], pos[j]);
      double dist = std::sqrt(distSq);
      correction -= (k.charge[i] * k.charge[j] * std::erf(sys.ff.alpha * dist) / dist);
    }
  return kQQFact * correction;
}

// src/Ewald.cpp:1375-1391.
double Ewald::SwapSelf(int kindIdx) const {
  const MolKind &k = sys.kinds[kindIdx];
  double en_self = 0.0;
  for (int i = 0; i < k.NumAtoms(); i++) en_self -= (k.charge[i] * k.charge[i]);
  return (en_self * sys.ff.alpha * kQQFact * M_2_SQRTPI * 0.5);
}

// src/Ewald.cpp:1420-1434: the new sums become the reference sums.
void Ewald::UpdateRecip() {
  std::swap(sumRnew, sumRref);
  std::swap(sumInew, sumIref);
}

// src/Ewald.cpp:1441-1460.
void Ewald::CopyRecip() {
  std::copy(sumRref.begin(), sumRref.begin() + imageSizeRef, sumRnew.begin());
  std::copy(sumIref.begin(), sumIref.begin() + imageSizeRef, sumInew.begin());
}

// src/Ewald.cpp:1021-1053.
void Ewald::SetRecipRef() {
  sumRref = sumRnew;
  sumIref = sumInew;
  kxRef = kx;
  kyRef = ky;
  kzRef = kz;
  hsqrRef = hsqr;
  prefactRef = prefact;
  imageSizeRef = imageSize;
}

// src/Ewald.cpp:1496-1596 (CPU branch), using the current sumRnew/sumInew.
void Ewald::BoxForceReciprocal(const std::vector<double> &x, const std::vector<double> &y,
                               const std::vector<double> &z, std::vector<XYZ> &aFR,
                               std::vector<XYZ> &mFR) {
  const ForceField &ff = sys.ff;
  double constValue = ff.alpha * M_2_SQRTPI;
  aFR.assign(sys.NumAtoms(), XYZ());
  mFR.assign(sys.NumMols(), XYZ());
  // Molecules are independent; each atom's k-sum keeps the serial order.
#pragma omp parallel for num_threads(sys.threads) schedule(dynamic, 4)
  for (int m = 0; m < sys.NumMols(); m++) {
    if (!sys.InBox0(m)) continue;
    int start = sys.molStart[m], length = sys.molLen[m];
    const double lambdaCoef = 1.0;
    for (int p = start; p < start + length; p++) {
      double X = 0.0, Y = 0.0, Z = 0.0;
      if (!sys.particleHasNoCharge[p]) {
        for (int j = start; j < start + length; j++) {
          if (p != j) {
            double distSq;
            XYZ dv;
            sys.box.InRcut(distSq, dv, XYZ(x[p], y[p], z[p]), XYZ(x[j], y[j], z[j]));
            double dist = std::sqrt(distSq);
            double expConstValue = std::exp(-1.0 * ff.alphaSq * distSq);
            double qiqj = sys.particleCharge[p] * sys.particleCharge[j] * kQQFact;
            double intraForce = qiqj * lambdaCoef * lambdaCoef / distSq;
            intraForce *= ((std::erf(ff.alpha * dist) / dist) - constValue * expConstValue);
            X -= intraForce * dv.x;
            Y -= intraForce * dv.y;
            Z -= intraForce * dv.z;
          }
        }
        for (int i = 0; i < imageSizeRef; i++) {
          double dot = x[p] * kxRef[i] + y[p] * kyRef[i] + z[p] * kzRef[i];
          double factor = 2.0 * sys.particleCharge[p] * prefactRef[i] * lambdaCoef *
                          (std::sin(dot) * sumRnew[i] - std::cos(dot) * sumInew[i]);
          X += factor * kxRef[i];
          Y += factor * kyRef[i];
          Z += factor * kzRef[i];
        }
      }
      aFR[p] = {X, Y, Z};
      mFR[m] = mFR[m] + XYZ(X, Y, Z);
    }
  }
}

}  // namespace orc

namespace orc {

// src/CalculateEnergy.cpp:411-579 (CPU branch, lambda = 1): the molecular
// virial of every inter-molecular pair inside the box cutoff, the atom
// separation times the minimum-image centre separation, diagonal terms;
// then VirialCorrection (1317-1336) and VirialReciprocal.
VirialD System::VirialCalc() {
  VirialD v;
  std::vector<int> cellVector, cellStart, map;
  cells.GetCellListNeighbor(NumAtoms(), cellVector, cellStart, map);
  double vT11 = 0.0, vT22 = 0.0, vT33 = 0.0, rT11 = 0.0, rT22 = 0.0, rT33 = 0.0;
  const int nv = int(cellVector.size());
#pragma omp parallel for num_threads(threads) schedule(static) \
    reduction(+ : vT11, vT22, vT33, rT11, rT22, rT33)
  for (int ci = 0; ci < nv; ci++) {
    int cur = cellVector[ci];
    int curCell = map[cur];
    for (int nc = 0; nc < kNeighborCells; nc++) {
      int nb = cells.neighbors[curCell][nc];
      for (int ni = cellStart[nb]; ni < cellStart[nb + 1]; ni++) {
        int n = cellVector[ni];
        if (cur < n && particleMol[cur] != particleMol[n]) {
          double distSq;
          XYZ virC;
          if (box.InRcut(distSq, virC, Pos(cur), Pos(n))) {
            const int mc = particleMol[cur], mn = particleMol[n];
            XYZ comC = box.MinImage(XYZ(comx[mc] - comx[mn], comy[mc] - comy[mn],
                                        comz[mc] - comz[mn]));
            if (ff.electrostatic) {
              double qi_qj = particleCharge[cur] * particleCharge[n];
              if (qi_qj != 0.0) {
                double pRF = ff.CalcCoulombVir(distSq, qi_qj);
                rT11 += pRF * (virC.x * comC.x);
                rT22 += pRF * (virC.y * comC.y);
                rT33 += pRF * (virC.z * comC.z);
              }
            }
            double pVF = ff.CalcVir(distSq, particleType[cur], particleType[n]);
            vT11 += pVF * (virC.x * comC.x);
            vT22 += pVF * (virC.y * comC.y);
            vT33 += pVF * (virC.z * comC.z);
          }
        }
      }
    }
  }
  v.interT[0] = vT11;
  v.interT[1] = vT22;
  v.interT[2] = vT33;
  if (ff.electrostatic) {
    v.realT[0] = rT11 * kQQFact;
    v.realT[1] = rT22 * kQQFact;
    v.realT[2] = rT33 * kQQFact;
  }
  v.inter = vT11 + vT22 + vT33;
  v.real = (rT11 + rT22 + rT33) * kQQFact;
  if (ff.useLRC) {  // VirialCorrection
    const int K = int(kinds.size());
    double vir = 0.0;
    for (int i = 0; i < K; ++i) {
      const unsigned numI = unsigned(kindCount[i]);
      for (int j = 0; j < K; ++j) {
        const unsigned numJ = unsigned(kindCount[j]);
        vir += pairVirCorrections[i * K + j] * numI * numJ * box.volInv;
      }
    }
    v.tc = vir;
  }
  if (ff.ewald) ewald.VirialReciprocal(v);
  // Virial::Total (EnergyTypes.h:181-191)
  v.total = v.inter + v.tc + v.real + v.recip;
  for (int d = 0; d < 3; ++d) v.totalT[d] = v.interT[d] + v.realT[d] + v.recipT[d];
  return v;
}

// OutputVars::CalcAndConvert (OutputVars.cpp:147-172): every virial term
// divided by the dimensionality and the volume, the ideal part from the
// number densities, K molecule / A^3 to bar.
double System::Pressure(const VirialD &v) const {
  const double inter = v.inter / 3.0 / box.volume, tc = v.tc / 3.0 / box.volume;
  const double real = v.real / 3.0 / box.volume, recip = v.recip / 3.0 / box.volume;
  const double total = inter + tc + real + recip;
  double raw = 0.0;
  for (size_t k = 0; k < kinds.size(); ++k) raw += kindCount[k] * box.volInv;
  raw *= ff.T;
  raw += total;
  return raw * 138.0649;  // UnitConst.h:31
}

// src/Ewald.cpp:1168-1305 (CPU branch): the k-space term of every k, then
// the intramolecular term of every charged atom (its offset from the
// molecule's centre, the atom unwrapped about it).
void Ewald::VirialReciprocal(VirialD &v) const {
  const ForceField &ff = sys.ff;
  double wT11 = 0.0, wT22 = 0.0, wT33 = 0.0;
  const double constVal = 1.0 / (4.0 * ff.alphaSq);
  for (int i = 0; i < imageSizeRef; i++) {
    double factor = prefactRef[i] * (sumRref[i] * sumRref[i] + sumIref[i] * sumIref[i]);
    wT11 += factor * (1.0 - 2.0 * (constVal + 1.0 / hsqrRef[i]) * kxRef[i] * kxRef[i]);
    wT22 += factor * (1.0 - 2.0 * (constVal + 1.0 / hsqrRef[i]) * kyRef[i] * kyRef[i]);
    wT33 += factor * (1.0 - 2.0 * (constVal + 1.0 / hsqrRef[i]) * kzRef[i] * kzRef[i]);
  }
  for (int m = 0; m < sys.NumMols(); m++) {
    if (!sys.InBox0(m)) continue;
    const XYZ comC(sys.comx[m], sys.comy[m], sys.comz[m]);
    for (int p = sys.molStart[m]; p < sys.molStart[m] + sys.molLen[m]; p++) {
      if (sys.particleHasNoCharge[p]) continue;
      XYZ atomC(BoxDims::Unwrap1(sys.x[p], comC.x, sys.box.axis.x, sys.box.halfAx.x),
                BoxDims::Unwrap1(sys.y[p], comC.y, sys.box.axis.y, sys.box.halfAx.y),
                BoxDims::Unwrap1(sys.z[p], comC.z, sys.box.axis.z, sys.box.halfAx.z));
      const XYZ diffC(atomC.x - comC.x, atomC.y - comC.y, atomC.z - comC.z);
      const double charge = sys.particleCharge[p] * 1.0;
      for (int i = 0; i < imageSizeRef; i++) {
        double arg = sys.x[p] * kxRef[i] + sys.y[p] * kyRef[i] + sys.z[p] * kzRef[i];
        double factor = prefactRef[i] * 2.0 *
                        (sumIref[i] * std::cos(arg) - sumRref[i] * std::sin(arg)) * charge;
        wT11 += factor * (kxRef[i] * diffC.x);
        wT22 += factor * (kyRef[i] * diffC.y);
        wT33 += factor * (kzRef[i] * diffC.z);
      }
    }
  }
  v.recipT[0] = wT11;
  v.recipT[1] = wT22;
  v.recipT[2] = wT33;
  v.recip = wT11 + wT22 + wT33;
}

}  // namespace orc

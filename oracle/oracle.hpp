// CPU oracle for the GOMC energy hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a from-scratch C++17/OpenMP restatement
// of the reference CPU path (GOMC 2.80) that the test suite, smoke() and the
// bench's cpu_baseline leg use as the checker.  The product library
// (gomc_amd/) never links, loads or calls anything in this directory.
//
// Each function cites the reference file:line whose arithmetic order it
// follows.  How each part is pinned (golden vectors, closed forms, brute
// force, invariants) is listed in tests/test_oracle_*.py.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

namespace orc {

// Planted mistakes, for the tests that show a statistical check detects
// them (0, the default, plants nothing): 1 swaps the new and old forces in
// MultiParticle's weight ratio, 2 makes a CBMC builder keep the trial after
// the one PickWeighted chose, 3 drops the +1 of N+1 in the insertion
// coefficient of MoleculeTransfer.  Set only through orc_plant.
extern int g_plant;

// lib/NumLib.h:22 -- Coulomb constant in K*A/e^2.
constexpr double kQQFact = 167103.208067979;
constexpr int kNeighborCells = 27;  // CellList.cpp:199-214 (3x3x3 stencil)

struct XYZ {
  double x = 0.0, y = 0.0, z = 0.0;
  XYZ() = default;
  XYZ(double a, double b, double c) : x(a), y(b), z(c) {}
  XYZ operator+(const XYZ &o) const { return {x + o.x, y + o.y, z + o.z}; }
  XYZ operator-(const XYZ &o) const { return {x - o.x, y - o.y, z - o.z}; }
  XYZ operator*(double s) const { return {x * s, y * s, z * s}; }
  XYZ operator-() const { return {-x, -y, -z}; }
};
inline XYZ Cross(const XYZ &a, const XYZ &b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---------------------------------------------------------------- RNG -----
// MT19937 with the MTRand interface of src/MersenneTwister.h:194-366.
class MTRand {
 public:
  static constexpr int N = 624, M = 397;
  explicit MTRand(uint32_t s = 5489u) { seed(s); }
  void seed(uint32_t s);                    // initialize() + reload()
  uint32_t randInt();                       // tempered 32-bit output
  uint32_t randInt(uint32_t n);             // [0,n] by bit-mask rejection
  double rand() { return double(randInt()) * (1.0 / 4294967295.0); }
  double rand(double n) { return rand() * n; }
  double randExc() { return double(randInt()) * (1.0 / 4294967296.0); }
  double randExc(double n) { return randExc() * n; }
  uint64_t draws = 0;                       // count of 32-bit outputs taken

 private:
  void reload();
  uint32_t state_[N];
  uint32_t *next_ = nullptr;
  int left_ = 0;
};

// Philox4x64-10 (Salmon et al., SC'11), the generator vendored as
// lib/Random123/philox.h:228-275 and used through Random123Wrapper.
void Philox4x64_10(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
// lib/Random123/uniform.hpp:174-214 for 64-bit inputs.
inline double U01(uint64_t v) { return double(v) * 0x1p-64 + 0x1p-65; }
inline double Uneg11(uint64_t v) { return double(int64_t(v)) * 0x1p-63 + 0x1p-64; }

// src/Random123Wrapper.cpp:16-84: key = {step, seed}, counter = {id, key2}.
class Random123Wrapper {
 public:
  void SetStep(uint64_t s) { step_ = s; }
  void SetRandomSeed(uint64_t s) { seed_ = s; }
  void SetKey(uint32_t k) { key2_ = k; }
  void Raw(uint32_t counter, uint64_t out[4]) const;
  double GetRandomNumber(uint32_t counter) const;
  XYZ GetRandomCoords(uint32_t counter) const;
  XYZ GetSymRandomCoords(uint32_t counter, double bound) const;
  XYZ GetRandomCoordsOnSphere(uint32_t counter) const;
  double GetSymRandom(uint32_t counter, double bound) const;

 private:
  uint64_t step_ = 0, seed_ = 0;
  uint32_t key2_ = 0;
};

// src/PRNG.h:48-203 (the members the hot-path moves use).
class PRNG {
 public:
  explicit PRNG(MTRand *g) : gen(g) {}
  double operator()() { return gen->rand(); }
  double rand(double b) { return gen->rand(b); }
  double randExc(double b) { return gen->randExc(b); }
  uint32_t randIntExc(uint32_t b) { return gen->randInt(b - 1); }
  double Sym(double b) { return gen->rand(2.0 * b) - b; }
  XYZ SymXYZ(double b);
  XYZ PickOnUnitSphere();
  void PickArbDist(uint32_t &pick, double &subDraw, const double *w, double Wt,
                   uint32_t len);
  int PickWeighted(const double *w, int n, double total);
  XYZ FillWithRandom(const XYZ &axis);
  MTRand *gen;
};

// ----------------------------------------------------------- geometry -----
// Orthogonal box: src/BoxDimensions.h:61-66,169-175,231-261 and
// BoxDimensions.cpp:17-19,261-320.
struct BoxDims {
  XYZ axis, halfAx;
  double volume = 0, volInv = 0, rCut = 0, rCutSq = 0;
  void Set(double lx, double ly, double lz, double rcut);
  static double MinImageSigned(double raw, double ax, double half) {
    if (raw > half)
      raw -= ax;
    else if (raw < -half)
      raw += ax;
    return raw;
  }
  XYZ MinImage(XYZ v) const {
    v.x = MinImageSigned(v.x, axis.x, halfAx.x);
    v.y = MinImageSigned(v.y, axis.y, halfAx.y);
    v.z = MinImageSigned(v.z, axis.z, halfAx.z);
    return v;
  }
  bool InRcut(double &distSq, XYZ &dist, const XYZ &a, const XYZ &b) const {
    dist = MinImage(a - b);
    distSq = dist.x * dist.x + dist.y * dist.y + dist.z * dist.z;
    return rCutSq > distSq;
  }
  static double Wrap1(double v, double ax) {
    if (v >= ax)
      v -= ax;
    else if (v < 0)
      v += ax;
    return v;
  }
  static double Unwrap1(double v, double ref, double ax, double half) {
    if (std::fabs(ref - v) > half) {
      if (ref < half)
        v -= ax;
      else
        v += ax;
    }
    return v;
  }
  XYZ Wrap(XYZ p) const {
    return {Wrap1(p.x, axis.x), Wrap1(p.y, axis.y), Wrap1(p.z, axis.z)};
  }
  XYZ Unwrap(XYZ p, const XYZ &ref) const {
    return {Unwrap1(p.x, ref.x, axis.x, halfAx.x),
            Unwrap1(p.y, ref.y, axis.y, halfAx.y),
            Unwrap1(p.z, ref.z, axis.z, halfAx.z)};
  }
};

// src/TransformMatrix.h:23-27,117-125,186-260.
struct RotationMatrix {
  double m[3][3];
  static RotationMatrix FromAxisAngle(double theta, const XYZ &axis);
  static RotationMatrix UniformRandom(double u1, double u2, double u3);
  XYZ Apply(const XYZ &p) const {
    return {m[0][0] * p.x + m[0][1] * p.y + m[0][2] * p.z,
            m[1][0] * p.x + m[1][1] * p.y + m[1][2] * p.z,
            m[2][0] * p.x + m[2][1] * p.y + m[2][2] * p.z};
  }
};

// -------------------------------------------------------- force field -----
// Forcefield::InitBasicVals (src/Forcefield.cpp:37-83) + FFParticle::Blend
// (src/FFParticle.cpp:153-201) for the standard Mie/LJ kind, arithmetic-mean
// sigma, geometric-mean epsilon.
struct ForceField {
  int count = 0;
  std::vector<double> sigmaSq, epsilon, epsilon_cn, epsilon_cn_6, n, nOver6;
  double rCut = 0, rCutSq = 0, rCutLow = 0, rCutLowSq = 0;
  double rCutCoulomb = 0, rCutCoulombSq = 0, tolerance = 0;
  double alpha = 0, alphaSq = 0, recip_rcut = 0, recip_rcut_Sq = 0;
  double T = 0, beta = 0;
  bool ewald = true, electrostatic = true, useLRC = true;

  void Init(int ntypes, const double *sigma, const double *eps,
            const double *nrep, double rcut, double rcutLow, double rcutCoul,
            double tol, double temperature, bool ewaldOn, bool elect, bool lrc);
  int Flat(int i, int j) const { return i + j * count; }
  double CalcEn(double distSq, int k1, int k2) const;          // 295-325
  double CalcVir(double distSq, int k1, int k2) const;         // 327-359
  double CalcCoulomb(double distSq, double qiqj) const;        // 361-399
  double CalcCoulombVir(double distSq, double qiqj) const;     // 401-446
  double EnergyLRC(int k1, int k2) const;                      // 96-116
  double VirialLRC(int k1, int k2) const;                      // 118-138
};

// The diagonal of the virial tensors and the scalar virials (EnergyTypes.h
// Virial; the off-diagonal terms are left out there too,
// CalculateEnergy.cpp:408-410).
struct VirialD {
  double inter = 0, tc = 0, real = 0, recip = 0, total = 0;
  double interT[3] = {0, 0, 0}, realT[3] = {0, 0, 0}, recipT[3] = {0, 0, 0}, totalT[3] = {0, 0, 0};
};

// ---------------------------------------------------------- molecules -----
struct MolKind {
  std::string name;
  std::vector<int> atomType;     // LJ type of each atom
  std::vector<double> charge;    // e
  int NumAtoms() const { return int(atomType.size()); }
};

// ---------------------------------------------------------- cell list -----
// src/CellList.cpp:81-285 and CellList.h:88-196: per-cell singly linked
// lists (head/list) with a 27-cell wrapped neighbour table.
class CellList {
 public:
  void Resize(const BoxDims &b, double cutoff);       // ResizeGrid+Rebuild
  int PositionToCell(const XYZ &p) const;             // CellList.h:88-101
  void AddAtom(int p, const XYZ &pos);                // AddMol body
  void RemoveAtom(int p, const XYZ &pos);             // RemoveMol body
  void Clear(int natoms);
  // Atoms of the 27 cells around `pos`, in GOMC's enumeration order.
  void Neighbors(const XYZ &pos, std::vector<int> &out) const;
  void NeighborsOfCell(int cell, std::vector<int> &out) const;
  void GetCellListNeighbor(int natoms, std::vector<int> &cellVector,
                           std::vector<int> &cellStartIndex,
                           std::vector<int> &mapParticleToCell) const;
  bool IsExhaustive(int natoms) const;                // CellList.cpp:66-79
  int NumCells() const { return int(head.size()); }
  int edge[3] = {0, 0, 0};
  XYZ cellSize;
  std::vector<int> head, list;
  std::vector<std::vector<int>> neighbors;
};

struct Energy {
  double inter = 0, tailCorrection = 0, real = 0, recip = 0, self = 0,
         correction = 0, total = 0;
  // src/EnergyTypes.h:88-91 (intra terms are zero for rigid molecules).
  double Total() {
    total = inter + tailCorrection + real + recip + self + correction;
    return total;
  }
};

struct MoveCounters;
class System;

// -------------------------------------------------------------- Ewald -----
// src/Ewald.cpp (non-cached, the reference default, ConfigSetup.cpp:34).
class Ewald {
 public:
  explicit Ewald(System &s) : sys(s) {}
  void Init();                                     // Ewald.cpp:100-139
  void RecipInitOrth();                            // 847-903
  void BoxReciprocalSetup();                       // 193-274 (CPU branch)
  void BoxReciprocalSums();                        // 281-361
  double BoxReciprocal(bool isNewVolume) const;    // 375-406
  double MolReciprocal(const std::vector<XYZ> &newPos, int mol);  // 409-473
  double SwapDestRecip(const std::vector<XYZ> &pos, int kind);    // 478-531
  double SwapSourceRecip(const std::vector<XYZ> &pos, int kind);  // 657-710
  double BoxSelf() const;                          // 1125-1163
  double MolCorrection(int mol) const;             // 1056-1085
  double SwapCorrection(const std::vector<XYZ> &pos, int kind) const; // 1311-1335
  double SwapSelf(int kind) const;                 // 1375-1391
  void UpdateRecip();                              // 1420-1434
  void CopyRecip();                                // 1441-1460
  void SetRecipRef();                              // 1021-1053
  void BoxForceReciprocal(const std::vector<double> &x,
                          const std::vector<double> &y,
                          const std::vector<double> &z,
                          std::vector<XYZ> &atomForceRec,
                          std::vector<XYZ> &molForceRec);  // 1496-1596
  void VirialReciprocal(VirialD &v) const;                 // 1168-1305 (CPU branch)
  int imageSize = 0, imageSizeRef = 0, kmax = 0;
  std::vector<double> kx, ky, kz, hsqr, prefact;
  std::vector<double> kxRef, kyRef, kzRef, hsqrRef, prefactRef;
  std::vector<double> sumRnew, sumInew, sumRref, sumIref;

 private:
  System &sys;
};

// ---------------------------------------------------------- the system -----
}  // namespace orc
struct orc_system_desc;  // C layout, oracle_capi.h
namespace orc {
using SystemDesc = ::orc_system_desc;

struct MoveSettings {  // src/MoveSettings.cpp:19-258 (displace/rotate)
  // [move][molecule kind]
  std::vector<double> scale[2];
  std::vector<uint64_t> tries[2], accepted[2], tempTries[2], tempAccepted[2];
  std::vector<uint64_t> swapTries, swapAccepted;  // MOL_TRANSFER [dest box][kind]
  double mp_t_max = 0.02, mp_r_max = 0.01 * M_PI;
  uint64_t mp_tries[2] = {0, 0}, mp_accepted[2] = {0, 0};
  uint64_t mp_int_tries[2] = {0, 0}, mp_int_acc[2] = {0, 0};
  bool singleMoveAccepted = true;
  // VOL_TRANSFER (NPT): the same counts are kept for every kind
  // (MoveSettings.cpp:130-145), so one set stands for all of them.
  double volScale = 500.0;  // MoveSettings.cpp:58-64
  uint64_t volTries = 0, volAccepted = 0, volTempTries = 0, volTempAccepted = 0;
};

// Record of one MC move, for per-move parity.
struct MoveRecord {
  int kind;          // 0 displace, 1 rotate, 2 multiparticle
  int mol;           // molecule moved (-1 for MP)
  int accepted;      // 0/1
  int overlap;       // 0/1
  double dLJ, dReal, dRecip;  // energy change components of the trial
};

class System {
 public:
  // Setup.
  void Build(const SystemDesc &d);
  // CalculateEnergy (src/CalculateEnergy.cpp).
  Energy SystemTotal();                                    // 83-138
  void BoxInter(double &lj, double &real);                 // 157-266
  void BoxInterBrute(double &lj, double &real) const;      // O(N^2) check
  bool MoleculeInter(double &lj, double &real,
                     const std::vector<XYZ> &newPos, int mol) const;  // 581-686
  void ParticleInter(double *en, double *real, const std::vector<XYZ> &trial,
                     bool *overlap, int atomType, double charge,
                     int mol) const;                       // 727-785
  double EnergyCorrection() const;                         // 1254-1315
  double MoleculeTailChange(int kind, bool add) const;     // 788-808
  void BoxForce(double &lj, double &real, std::vector<XYZ> &atomForce,
                std::vector<XYZ> &molForce);               // 268-406
  void CalculateTorque(const std::vector<XYZ> &aF, const std::vector<XYZ> &aFR,
                       std::vector<XYZ> &torque) const;    // 1365-1406
  VirialD VirialCalc();                                    // 411-579
  double Pressure(const VirialD &v) const;                 // OutputVars.cpp:120-185
  // Moves.
  void CalcCOM();                                          // COM.h:41-62
  int RunStepMT(uint64_t step, MoveRecord *rec);           // System.cpp:262-302
  int RunStepPhilox(uint64_t step, MoveRecord *rec);       // batched variant
  void AdjustMoves(uint64_t step);                         // MoveSettings.cpp:148-258
  // MultiParticle (src/moves/MultiParticle.h:164-712), NVT, one box.
  int RunMultiParticle(uint64_t step, MoveRecord *rec);
  void GridAll();                                          // CellList.cpp:220-240

  // Forces of the MultiParticle move (MultiParticle.h:37-58): the
  // reference set and the trial set.
  struct MPForces {
    std::vector<XYZ> aF, aFR, mF, mFR, torque;
  };
  MPForces mpRef, mpNew;

  XYZ Pos(int a) const { return {x[a], y[a], z[a]}; }
  int NumAtoms() const { return int(x.size()); }
  int NumMols() const { return int(molKind.size()); }

  ForceField ff;
  BoxDims box;
  std::vector<MolKind> kinds;
  std::vector<int> molKind, molStart, molLen;
  std::vector<int> particleType, particleMol;
  std::vector<double> particleCharge;
  std::vector<char> particleHasNoCharge;
  std::vector<double> x, y, z, comx, comy, comz;
  std::vector<int> kindCount;          // molecules of each kind in the box
  std::vector<double> pairEnCorrections;  // Molecules.cpp:157-175
  std::vector<double> pairVirCorrections;
  CellList cells;
  Ewald ewald{*this};
  Energy pot;
  MTRand mt;
  PRNG prng{&mt};
  uint64_t philoxSeed = 0;
  MoveSettings ms;
  double movePerc[3] = {0.5, 0.5, 0.0};  // displace, rotate, multiparticle
  // StaticVals movePerc over the reference's move kinds (MoveConst.h:22-35):
  // DISPLACE 0, ROTATE 1, MULTIPARTICLE 2, MOL_TRANSFER 10 (GCMC).
  double moveW[13] = {0};

  // ---- GCMC (ensemble 1): box 0 interacts, box 1 is the reservoir ----
  int ensemble = 0;
  std::vector<int> molBox;
  // MoleculeLookup (MoleculeLookup.cpp:24-119,165-198): molecule indices
  // grouped by [box][kind], section starts, swappable counts, and the
  // kinds in order of first appearance.
  std::vector<uint32_t> lookup, bkStart, swapCount;
  std::vector<int> canSwapKind, canMoveKind;
  std::vector<double> fugacity;  // K / A^3 (bar * 0.007242971, UnitConst.h:44)
  int cbmcFirst = 16, cbmcNth = 8, cbmcAng = 100;
  // Rigid three-site kinds grow from atom 0 (bonded to atoms 1 and 2):
  // bond lengths and the 1-0-2 angle of the kind.
  struct Hedron {
    double bPrev = 0, bBonded = 0, theta = 0;
  };
  std::vector<Hedron> hedron;
  bool InBox0(int m) const { return molBox.empty() || molBox[m] == 0; }
  uint32_t NumKindInBox(int k, int b) const {
    const int K = int(kinds.size());
    return bkStart[b * K + k + 1] - bkStart[b * K + k];
  }
  uint32_t NumInBox(int b) const {
    const int K = int(kinds.size());
    return bkStart[(b + 1) * K] - bkStart[b * K];
  }
  int GetMolNum(uint32_t sub, int k, int b) const {
    return int(lookup[bkStart[b * int(kinds.size()) + k] + sub]);
  }
  void ShiftMolBox(int m, int from, int to, int k);  // MoleculeLookup.cpp:165-198
  void SetNewCOM(int m);                              // COM.h:63-88
  int RunMoleculeTransfer(double subDraw, MoveRecord *rec);  // MoleculeTransfer.h
  int RunVolumeMove(MoveRecord *rec);                         // VolumeTransfer.h (NPT)
  double pressure = 0.0;  // K / A^3 (bar * 0.007242971, ConfigSetup.cpp:387-389)
  uint64_t perAdjust = 1000;
  int threads = 1;

 private:
  int TranslateMove(int m, const XYZ &shift, MoveRecord *rec, double pr,
                    bool haveDraw);
};

}  // namespace orc

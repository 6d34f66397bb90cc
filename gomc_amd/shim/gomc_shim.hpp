// gomc_shim.hpp -- C++ face of the gfx950 energy path with the method
// names, argument meanings and error behaviour of GOMC's CalculateEnergy,
// Ewald / EwaldCached and CellList (src/CalculateEnergy.h:48-180,
// src/Ewald.h:46-180, src/EwaldCached.h:11-68, src/CellList.h:22-68), over
// the C ABI of include/gomc_hip.h.  Header-only, host C++17; link
// libgomc_hip.so.
//
// One gomc_hip::Box is one GOMC simulation box resident on a device.  Its
// methods take what the reference's take, with XYZArray replaced by three
// coordinate pointers and a count, and return what the reference's return:
//   * MolReciprocal / SwapDestRecip / SwapSourceRecip return the change of
//     the reciprocal energy against the box's reference sums, as
//     Ewald::MolReciprocal returns energyRecipNew - sysPotRef...recip
//     (Ewald.cpp:409-473); UpdateRecip makes the last trial the reference;
//   * ParticleInter adds into en[], real[] and ORs overlap[]
//     (CalculateEnergy.cpp:727-785);
//   * MoleculeInter returns the overlap flag and fills the two
//     Intermolecular energies (CalculateEnergy.cpp:581-686).
// Errors: the reference exits the process on a fatal error (gpuErrchk,
// GPU/VariablesCUDA.cuh:20-39); here a failing call throws gomc_hip::Error
// with the library's message, or, with OnError::Exit, prints it and exits
// with EXIT_FAILURE as the reference does.
//
// The reference-side classes that derive from Ewald / EwaldCached and stand
// in for CalculateEnergy and CellList forward to this class (see
// gomc_amd/shim/reference/).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gomc_hip.h"

namespace gomc_hip {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

enum class OnError { Throw, Exit };

// SystemPotential's per-box Energy (src/EnergyTypes.h:75-129) of a rigid box.
struct Energy {
  double inter = 0, tailCorrection = 0, real = 0, recip = 0, self = 0, correction = 0,
         total = 0;
};

// Virial of a box (src/EnergyTypes.h:176-230): components, the diagonal of
// the total tensor, and the pressure OutputVars forms from them (bar).
struct Virial {
  double inter = 0, tailCorrection = 0, real = 0, recip = 0, total = 0, pressure = 0;
  double totalTens[3] = {0, 0, 0};
};

class Box {
 public:
  Box(const gomc_system_desc &desc, int device = 0, OnError onError = OnError::Throw)
      : onError_(onError) {
    Check(gomc_box_create(&desc, device, &b_), "gomc_box_create", true);
    int64_t info[7];
    Check(gomc_box_info(b_, info), "gomc_box_info");
    natoms_ = int(info[0]);
    nmols_ = int(info[1]);
  }
  ~Box() {
    if (b_) gomc_box_destroy(b_);
  }
  Box(const Box &) = delete;
  Box &operator=(const Box &) = delete;

  gomc_box *handle() { return b_; }
  int NumAtoms() const { return natoms_; }
  int NumMols() const { return nmols_; }

  // ---- coordinates (Coordinates / XYZArray) ----
  void SetCoords(const double *x, const double *y, const double *z) {
    Check(gomc_set_coords(b_, x, y, z), "gomc_set_coords");
  }
  void GetCoords(double *x, double *y, double *z) {
    Check(gomc_get_coords(b_, x, y, z), "gomc_get_coords");
  }
  // XYZArray::CopyRange of an accepted molecule (Translate.h:114).
  void CommitMol(int molIndex, const double *x, const double *y, const double *z) {
    Check(gomc_commit_mol(b_, molIndex, x, y, z), "gomc_commit_mol");
  }
  // MoleculeLookup::ShiftMolBox of an accepted transfer: box 0 or the
  // reservoir (1).
  void ShiftMolBox(int molIndex, int box) {
    Check(gomc_set_mol_box(b_, molIndex, box), "gomc_set_mol_box");
  }

  // ---- CellList (CellList.cpp:220-289) ----
  // GridAll + GetCellListNeighbor: the CSR form, atoms ascending per cell.
  void GetCellListNeighbor(std::vector<int> &cellVector, std::vector<int> &cellStartIndex,
                           std::vector<int> &mapParticleToCell) {
    int64_t info[7];
    Check(gomc_box_info(b_, info), "gomc_box_info");
    cellVector.resize(size_t(natoms_));
    cellStartIndex.resize(size_t(info[3]) + 1);
    mapParticleToCell.resize(size_t(natoms_));
    Check(gomc_build_cells(b_, cellVector.data(), cellStartIndex.data(),
                           mapParticleToCell.data()),
          "gomc_build_cells");
  }

  // ---- CalculateEnergy ----
  // SystemTotal of this box (CalculateEnergy.cpp:83-138), rigid molecules.
  Energy SystemTotal() {
    double e[7];
    Check(gomc_system_total(b_, e), "gomc_system_total");
    return Energy{e[0], e[1], e[2], e[3], e[4], e[5], e[6]};
  }
  // BoxInter's pair sums (CalculateEnergy.cpp:157-266): LJ and real-space
  // Coulomb of the current coordinates.
  void BoxInter(double &inter, double &real) {
    Check(gomc_box_inter(b_, &inter, &real), "gomc_box_inter");
  }
  bool MoleculeInter(double &interLJ, double &interCoulomb, const double *x, const double *y,
                     const double *z, int molIndex) {
    int32_t ov = 0;
    Check(gomc_mol_inter(b_, molIndex, x, y, z, &interLJ, &interCoulomb, &ov),
          "gomc_mol_inter");
    return ov != 0;
  }
  // ParticleInter for `trials` positions of one atom (kind's atom type and
  // charge); molIndex >= 0 leaves that molecule's atoms out.
  void ParticleInter(double *en, double *real, const double *x, const double *y,
                     const double *z, bool *overlap, int atomType, double charge, int molIndex,
                     int trials) {
    const size_t n = trials > 0 ? static_cast<size_t>(trials) : 0;
    std::vector<double> e(n), r(n);
    std::vector<int32_t> o(n);
    Check(gomc_particle_inter(b_, atomType, charge, molIndex, trials, x, y, z, e.data(),
                              r.data(), o.data()),
          "gomc_particle_inter");
    for (int t = 0; t < trials; ++t) {
      en[t] += e[size_t(t)];
      real[t] += r[size_t(t)];
      overlap[t] = overlap[t] || o[size_t(t)] != 0;
    }
  }
  // BoxForce (CalculateEnergy.cpp:268-406): pair energies, atom and molecule
  // forces (any force pointer may be null: the forces stay on the device).
  void BoxForce(double &inter, double &real, double *afx, double *afy, double *afz, double *mfx,
                double *mfy, double *mfz) {
    Check(gomc_box_force(b_, &inter, &real, afx, afy, afz, mfx, mfy, mfz), "gomc_box_force");
  }
  // CalculateTorque (CalculateEnergy.cpp:1365-1406) of the device forces.
  void CalculateTorque(double *tx, double *ty, double *tz) {
    Check(gomc_torque(b_, tx, ty, tz), "gomc_torque");
  }
  // VirialCalc with VirialReciprocal and the pressure (bar).
  Virial VirialCalc() {
    double v[9];
    Check(gomc_box_virial(b_, v), "gomc_box_virial");
    Virial out;
    out.inter = v[0];
    out.tailCorrection = v[1];
    out.real = v[2];
    out.recip = v[3];
    out.total = v[4];
    out.pressure = v[5];
    for (int i = 0; i < 3; ++i) out.totalTens[i] = v[6 + i];
    return out;
  }

  // ---- Ewald ----
  // RecipInit + BoxReciprocalSetup + SetRecipRef; returns BoxReciprocal.
  double BoxReciprocalSetup() {
    Check(gomc_recip_init(b_, &recipRef_), "gomc_recip_init");
    recipNew_ = recipRef_;
    return recipRef_;
  }
  // BoxReciprocalSums of the current coordinates into the new sums;
  // returns BoxReciprocal(box, false) of them.
  double BoxReciprocalSums() {
    Check(gomc_recip_sums(b_, &recipNew_), "gomc_recip_sums");
    return recipNew_;
  }
  // BoxReciprocal of the reference sums, or (isNewVolume) of the last
  // computed ones.
  double BoxReciprocal(bool isNewVolume = false) const {
    return isNewVolume ? recipNew_ : recipRef_;
  }
  double MolReciprocal(const double *x, const double *y, const double *z, int molIndex) {
    Check(gomc_mol_recip(b_, molIndex, x, y, z, &recipNew_), "gomc_mol_recip");
    return recipNew_ - recipRef_;
  }
  double SwapDestRecip(int kind, int molIndex, const double *x, const double *y,
                       const double *z) {
    Check(gomc_swap_recip(b_, kind, molIndex, 1, x, y, z, &recipNew_), "gomc_swap_recip");
    return recipNew_ - recipRef_;
  }
  double SwapSourceRecip(int kind, int molIndex, const double *x, const double *y,
                         const double *z) {
    Check(gomc_swap_recip(b_, kind, molIndex, 0, x, y, z, &recipNew_), "gomc_swap_recip");
    return recipNew_ - recipRef_;
  }
  // UpdateRecip: the last trial's sums become the reference.
  void UpdateRecip() {
    Check(gomc_recip_accept(b_), "gomc_recip_accept");
    recipRef_ = recipNew_;
  }
  // CopyRecip: the reference sums are copied to the new ones.
  void CopyRecip() {
    Check(gomc_recip_copy_ref_to_new(b_), "gomc_recip_copy_ref_to_new");
    recipNew_ = recipRef_;
  }
  // BoxForceReciprocal of the new sums (CopyRecip first, as
  // MultiParticle::Prep does); any pointer may be null.
  void BoxForceReciprocal(double *afx, double *afy, double *afz, double *mfx, double *mfy,
                          double *mfz) {
    Check(gomc_box_force_recip(b_, afx, afy, afz, mfx, mfy, mfz), "gomc_box_force_recip");
  }
  // EwaldCached mode (CachedFourier true) on or off.
  void SetCached(bool on) { Check(gomc_recip_cache(b_, on ? 1 : 0), "gomc_recip_cache"); }
  // k-vectors and reference sums, arrays of NumKVectors(); any may be null.
  int NumKVectors() {
    int64_t info[7];
    Check(gomc_box_info(b_, info), "gomc_box_info");
    return int(info[2]);
  }
  void GetRecip(double *kx, double *ky, double *kz, double *prefact, double *sumR,
                double *sumI) {
    Check(gomc_recip_get(b_, kx, ky, kz, prefact, sumR, sumI), "gomc_recip_get");
  }

 private:
  void Check(int rc, const char *what, bool creating = false) {
    if (rc >= 0) return;
    std::string msg = std::string(what) + " failed (" + std::to_string(rc) + ")";
    if (!creating && b_) msg += ": " + std::string(gomc_last_error(b_));
    if (onError_ == OnError::Exit) {
      std::fprintf(stderr, "%s\n", msg.c_str());
      std::exit(EXIT_FAILURE);
    }
    throw Error(rc, msg);
  }

  gomc_box *b_ = nullptr;
  OnError onError_;
  int natoms_ = 0, nmols_ = 0;
  double recipRef_ = 0.0, recipNew_ = 0.0;
};

}  // namespace gomc_hip

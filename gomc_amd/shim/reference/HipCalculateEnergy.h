// HipCalculateEnergy -- GOMC's CalculateEnergy (src/CalculateEnergy.h:48-180)
// with its pair energies, forces and virial on the gfx950 path.  Drop-in for
// the reference's src/ tree: System (src/System.h) declares
//     HipCalculateEnergy calcEnergy;
// in place of CalculateEnergy calcEnergy;.  CalculateEnergy's methods are
// not virtual, so the ones below hide the base class's with the same
// signatures, and every caller, which reaches them through System's member,
// calls these.  The tail corrections, intramolecular (bonded) terms and
// free-energy terms stay the base class's host code.
//
// The device box is made in Init from the same Forcefield, Molecules and
// Coordinates the base class reads (HipMakeDesc), and it holds the
// coordinates across calls: an accepted move tells it what changed
// (AcceptMol for Translate / Rotate / a CBMC insertion, AcceptAll for
// MultiParticle and volume moves, ShiftMolBox for a transfer), from the
// moves' Accept methods.
//
// Compiled in the reference's build (its headers need GOMC_Config.h).
#pragma once

#include <memory>

#include "CalculateEnergy.h"
#include "gomc_shim.hpp"

class HipCalculateEnergy : public CalculateEnergy {
public:
  HipCalculateEnergy(StaticVals &stat, System &sys);

  // CalculateEnergy::Init (CalculateEnergy.cpp:60-81), then the device box.
  void Init(System &sys);
  gomc_hip::Box &Device() { return *dev; }

  SystemPotential SystemTotal();
  SystemPotential BoxInter(SystemPotential potential, XYZArray const &coords,
                           BoxDimensions const &boxAxes, const uint box);
  SystemPotential BoxForce(SystemPotential potential, XYZArray const &coords,
                           XYZArray &atomForce, XYZArray &molForce,
                           BoxDimensions const &boxAxes, const uint box);
  Virial VirialCalc(const uint box);
  bool MoleculeInter(Intermolecular &inter_LJ, Intermolecular &inter_coulomb,
                     XYZArray const &molCoords, const uint molIndex,
                     const uint box) const;
  void ParticleInter(double *en, double *real, XYZArray const &trialPos,
                     bool *overlap, const uint partIndex, const uint molIndex,
                     const uint box, const uint trials) const;
  void CalculateTorque(std::vector<uint> &moleculeIndex,
                       XYZArray const &coordinates, XYZArray const &com,
                       XYZArray const &atomForce, XYZArray const &atomForceRec,
                       XYZArray &molTorque, const uint box);

  // What an accepted move changed, for the device's copy of box 0.
  void AcceptMol(const uint molIndex, XYZArray const &newPos) const;
  void AcceptAll(XYZArray const &coords) const;
  void ShiftMolBox(const uint molIndex, const uint box) const;

private:
  const Forcefield &ffRef;
  const Molecules &molsRef;
  const Coordinates &coordsRef;
  const BoxDimensions &axesRef;
  const MoleculeLookup &lookupRef;
  std::unique_ptr<gomc_hip::Box> dev;
};

// The C ABI's description of box 0 built from the reference's classes:
// box edges, temperature, cutoffs, Ewald settings, per-type Mie parameters,
// molecule kinds (atom types and charges), molecules and coordinates; in
// GCMC the molecules of box 1 go to the reservoir.  `arrays` keeps the
// arrays the descriptor points into alive.
struct HipDescArrays {
  std::vector<double> sigma, epsilon, n, charge, x, y, z, fugacity;
  std::vector<int32_t> natoms, atomType, molKind, molBox;
};
gomc_system_desc HipMakeDesc(const Forcefield &ff, const Molecules &mols,
                             const Coordinates &coords, const BoxDimensions &axes,
                             const MoleculeLookup &lookup, HipDescArrays &arrays);

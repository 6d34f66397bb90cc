// HipCalculateEnergy: CalculateEnergy's pair energies on the gfx950 path
// (see HipCalculateEnergy.h).  Compiled in the reference's build.
#include "HipCalculateEnergy.h"

#include <cstdio>
#include <cstdlib>

#include "BoxDimensions.h"
#include "Coordinates.h"
#include "FFParticle.h"
#include "Forcefield.h"
#include "MoleculeKind.h"
#include "MoleculeLookup.h"
#include "Molecules.h"
#include "StaticVals.h"
#include "System.h"

namespace {
[[noreturn]] void OffDevice(const char *what, uint box) {
  std::fprintf(stderr, "HipCalculateEnergy::%s: box %u is not on the device (box 0 is)\n",
               what, box);
  std::exit(EXIT_FAILURE);
}
} // namespace

gomc_system_desc HipMakeDesc(const Forcefield &ff, const Molecules &mols,
                             const Coordinates &coords, const BoxDimensions &axes,
                             const MoleculeLookup &lookup, HipDescArrays &a) {
  gomc_system_desc d{};
  d.box[0] = axes.axis.Get(0).x;
  d.box[1] = axes.axis.Get(0).y;
  d.box[2] = axes.axis.Get(0).z;
  d.temperature = ff.T_in_K;
  d.rcut = ff.rCut;
  d.rcut_low = ff.rCutLow;
  d.rcut_coulomb = ff.rCutCoulomb[0];
  d.tolerance = ff.tolerance;
  d.ewald = ff.ewald;
  d.electrostatic = ff.electrostatic;
  d.lrc = ff.useLRC;
  d.threads = 1;
  const uint ntypes = ff.particles->NumKinds();
  for (uint t = 0; t < ntypes; ++t) {
    a.sigma.push_back(ff.particles->GetSigma(t, t));
    a.epsilon.push_back(ff.particles->GetEpsilon(t, t));
    a.n.push_back(ff.particles->GetN(t, t));
  }
  d.n_atom_types = int32_t(ntypes);
  d.sigma = a.sigma.data();
  d.epsilon = a.epsilon.data();
  d.n_rep = a.n.data();
  const uint nkinds = mols.GetKindsCount();
  for (uint k = 0; k < nkinds; ++k) {
    const MoleculeKind &mk = mols.kinds[k];
    a.natoms.push_back(int32_t(mk.NumAtoms()));
    for (uint p = 0; p < mk.NumAtoms(); ++p) {
      a.atomType.push_back(int32_t(mk.AtomKind(p)));
      a.charge.push_back(mk.AtomCharge(p));
    }
  }
  d.n_kinds = int32_t(nkinds);
  d.kind_natoms = a.natoms.data();
  d.kind_atom_type = a.atomType.data();
  d.kind_charge = a.charge.data();
  const uint nmols = mols.count;
  for (uint m = 0; m < nmols; ++m) {
    a.molKind.push_back(int32_t(mols.GetMolKind(m)));
    const uint k = mols.GetMolKind(m);
    a.molBox.push_back(int32_t(lookup.IsMoleculeInBox(m, k, 0) ? 0 : 1));
  }
  d.n_mols = int32_t(nmols);
  d.mol_kind = a.molKind.data();
  a.x.assign(coords.x, coords.x + coords.Count());
  a.y.assign(coords.y, coords.y + coords.Count());
  a.z.assign(coords.z, coords.z + coords.Count());
  d.x = a.x.data();
  d.y = a.y.data();
  d.z = a.z.data();
#if ENSEMBLE == GCMC
  d.ensemble = 1;
  d.mol_box = a.molBox.data();
  a.fugacity.assign(nkinds, 0.0);  // the host's moves make the transfers
  d.kind_fugacity = a.fugacity.data();
#else
  d.ensemble = 0;  // NVT and NPT: the host's moves change the volume
#endif
  return d;
}

HipCalculateEnergy::HipCalculateEnergy(StaticVals &stat, System &sys)
    : CalculateEnergy(stat, sys), ffRef(stat.forcefield), molsRef(stat.mol),
      coordsRef(sys.coordinates), axesRef(sys.boxDimRef), lookupRef(sys.molLookupRef) {}

void HipCalculateEnergy::Init(System &sys) {
  CalculateEnergy::Init(sys);
  HipDescArrays arrays;
  const gomc_system_desc d =
      HipMakeDesc(ffRef, molsRef, coordsRef, axesRef, lookupRef, arrays);
  dev = std::make_unique<gomc_hip::Box>(d, 0, gomc_hip::OnError::Exit);
}

// SystemTotal (CalculateEnergy.cpp:83-138): box 0's inter-molecular and
// Ewald terms from the device; the base class adds nothing else for rigid
// molecules (their intramolecular energies are zero).
SystemPotential HipCalculateEnergy::SystemTotal() {
  SystemPotential pot = CalculateEnergy::SystemTotal();
  const gomc_hip::Energy e = dev->SystemTotal();
  pot.boxEnergy[0].inter = e.inter;
  pot.boxEnergy[0].tailCorrection = e.tailCorrection;
  pot.boxEnergy[0].real = e.real;
  pot.boxEnergy[0].recip = e.recip;
  pot.boxEnergy[0].self = e.self;
  pot.boxEnergy[0].correction = e.correction;
  pot.Total();
  return pot;
}

SystemPotential HipCalculateEnergy::BoxInter(SystemPotential potential,
                                             XYZArray const &coords,
                                             BoxDimensions const &boxAxes,
                                             const uint box) {
  if (box >= BOXES_WITH_U_NB) return potential;
  if (box != 0) OffDevice("BoxInter", box);
  dev->SetCoords(coords.x, coords.y, coords.z);
  double lj = 0.0, real = 0.0;
  dev->BoxInter(lj, real);
  potential.boxEnergy[box].inter = lj;
  potential.boxEnergy[box].real = real;
  // EnergyCorrection (CalculateEnergy.cpp:1254-1298) of the box's counts
  potential.boxEnergy[box].tailCorrection = dev->SystemTotal().tailCorrection;
  potential.Total();
  return potential;
}

SystemPotential HipCalculateEnergy::BoxForce(SystemPotential potential,
                                             XYZArray const &coords,
                                             XYZArray &atomForce, XYZArray &molForce,
                                             BoxDimensions const &boxAxes,
                                             const uint box) {
  if (box >= BOXES_WITH_U_NB) return potential;
  if (box != 0) OffDevice("BoxForce", box);
  dev->SetCoords(coords.x, coords.y, coords.z);
  double lj = 0.0, real = 0.0;
  dev->BoxForce(lj, real, atomForce.x, atomForce.y, atomForce.z, molForce.x, molForce.y,
                molForce.z);
  potential.boxEnergy[box].inter = lj;
  potential.boxEnergy[box].real = real;
  potential.Total();
  return potential;
}

Virial HipCalculateEnergy::VirialCalc(const uint box) {
  Virial v;
  if (box >= BOXES_WITH_U_NB) return v;
  if (box != 0) OffDevice("VirialCalc", box);
  const gomc_hip::Virial d = dev->VirialCalc();
  v.inter = d.inter;
  v.tailCorrection = d.tailCorrection;
  v.real = d.real;
  v.recip = d.recip;
  for (int i = 0; i < 3; ++i) v.totalTens[i][i] = d.totalTens[i];
  v.Total();
  return v;
}

bool HipCalculateEnergy::MoleculeInter(Intermolecular &inter_LJ,
                                       Intermolecular &inter_coulomb,
                                       XYZArray const &molCoords, const uint molIndex,
                                       const uint box) const {
  if (box >= BOXES_WITH_U_NB) return false;
  if (box != 0) OffDevice("MoleculeInter", box);
  return dev->MoleculeInter(inter_LJ.energy, inter_coulomb.energy, molCoords.x, molCoords.y,
                            molCoords.z, int(molIndex));
}

// ParticleInter: the molecule's own atoms are left out (the reference has
// removed the molecule from its cell list before growing it).
void HipCalculateEnergy::ParticleInter(double *en, double *real, XYZArray const &trialPos,
                                       bool *overlap, const uint partIndex,
                                       const uint molIndex, const uint box,
                                       const uint trials) const {
  if (box >= BOXES_WITH_U_NB) return;
  if (box != 0) OffDevice("ParticleInter", box);
  const MoleculeKind &kind = molsRef.GetKind(molIndex);
  dev->ParticleInter(en, real, trialPos.x, trialPos.y, trialPos.z, overlap,
                     int(kind.AtomKind(partIndex)), kind.AtomCharge(partIndex), int(molIndex),
                     int(trials));
}

// CalculateTorque of the forces the device holds from the last BoxForce and
// BoxForceReciprocal, for every molecule of the box (moleculeIndex lists
// them all in the reference's MultiParticle use).
void HipCalculateEnergy::CalculateTorque(std::vector<uint> &moleculeIndex,
                                         XYZArray const &coordinates, XYZArray const &com,
                                         XYZArray const &atomForce,
                                         XYZArray const &atomForceRec, XYZArray &molTorque,
                                         const uint box) {
  if (box >= BOXES_WITH_U_NB) return;
  if (box != 0) OffDevice("CalculateTorque", box);
  dev->CalculateTorque(molTorque.x, molTorque.y, molTorque.z);
}

void HipCalculateEnergy::AcceptMol(const uint molIndex, XYZArray const &newPos) const {
  dev->CommitMol(int(molIndex), newPos.x, newPos.y, newPos.z);
}

void HipCalculateEnergy::AcceptAll(XYZArray const &coords) const {
  dev->SetCoords(coords.x, coords.y, coords.z);
}

void HipCalculateEnergy::ShiftMolBox(const uint molIndex, const uint box) const {
  dev->ShiftMolBox(int(molIndex), box == 0 ? 0 : 1);
}

// HipEwald: Ewald's reciprocal space on the gfx950 path (see HipEwald.h).
// Compiled in the reference's build (its headers need GOMC_Config.h).
#include "HipEwald.h"

#include <cstdio>
#include <cstdlib>

#include "TrialMol.h"

HipEwald::HipEwald(StaticVals &stat, System &sys, gomc_hip::Box &dev, bool cached)
    : Ewald(stat, sys), dev(dev), cached(cached) {}

void HipEwald::OnDevice(uint box, const char *what) const {
  if (box != 0) {
    std::fprintf(stderr, "HipEwald::%s: box %u is not on the device (box 0 is)\n", what,
                 box);
    std::exit(EXIT_FAILURE);
  }
}

void HipEwald::NotBuilt(const char *what) const {
  std::fprintf(stderr, "HipEwald::%s is not on the gfx950 path\n", what);
  std::exit(EXIT_FAILURE);
}

// Ewald::Init (Ewald.cpp:72-127) fills the host tables the closed-form terms
// use (charges, molecule starts); the k-vectors and sums live on the device.
void HipEwald::Init() {
  Ewald::Init();
  dev.SetCached(cached);
}

// The host k-vector arrays are not used: the device builds its own
// (RecipInitOrth's loop order, bit for bit).
void HipEwald::AllocMem() {}

void HipEwald::RecipInit(uint box, BoxDimensions const &boxAxes) { OnDevice(box, "RecipInit"); }

void HipEwald::UpdateRecipVec(uint box) { OnDevice(box, "UpdateRecipVec"); }

// RecipInit + BoxReciprocalSetup + SetRecipRef of the device's coordinates
// (the caller's molCoords are the same system's current coordinates).
void HipEwald::BoxReciprocalSetup(uint box, XYZArray const &molCoords) {
  OnDevice(box, "BoxReciprocalSetup");
  dev.SetCoords(molCoords.x, molCoords.y, molCoords.z);
  dev.BoxReciprocalSetup();
}

void HipEwald::BoxReciprocalSums(uint box, XYZArray const &molCoords) {
  OnDevice(box, "BoxReciprocalSums");
  dev.SetCoords(molCoords.x, molCoords.y, molCoords.z);
  dev.BoxReciprocalSums();
}

double HipEwald::BoxReciprocal(uint box, bool isNewVolume) const {
  if (box >= BOXES_WITH_U_NB) return 0.0;
  OnDevice(box, "BoxReciprocal");
  return dev.BoxReciprocal(isNewVolume);
}

// VirialReciprocal (Ewald.cpp:1168-1305): the device computes the whole
// virial; its reciprocal component is taken here.  The per-component
// reciprocal tensor is not returned by the device (the pressure needs the
// total's diagonal, which gomc_box_virial gives).
Virial HipEwald::VirialReciprocal(Virial &virial, uint box) const {
  if (box >= BOXES_WITH_U_NB) return virial;
  OnDevice(box, "VirialReciprocal");
  const gomc_hip::Virial v = dev.VirialCalc();
  Virial out = virial;
  out.recip = v.recip;
  return out;
}

double HipEwald::MolReciprocal(XYZArray const &molCoords, const uint molIndex,
                               const uint box) {
  if (box >= BOXES_WITH_U_NB) return 0.0;
  OnDevice(box, "MolReciprocal");
  return dev.MolReciprocal(molCoords.x, molCoords.y, molCoords.z, int(molIndex));
}

double HipEwald::SwapDestRecip(const cbmc::TrialMol &newMol, const uint box,
                               const int molIndex) {
  if (box >= BOXES_WITH_U_NB) return 0.0;
  OnDevice(box, "SwapDestRecip");
  const XYZArray &c = newMol.GetCoords();
  return dev.SwapDestRecip(int(newMol.GetKind().kindIndex), molIndex, c.x, c.y, c.z);
}

double HipEwald::SwapSourceRecip(const cbmc::TrialMol &oldMol, const uint box,
                                 const int molIndex) {
  if (box >= BOXES_WITH_U_NB) return 0.0;
  OnDevice(box, "SwapSourceRecip");
  const XYZArray &c = oldMol.GetCoords();
  return dev.SwapSourceRecip(int(oldMol.GetKind().kindIndex), molIndex, c.x, c.y, c.z);
}

// SetRecipRef / UpdateRecip: the new sums become the reference (a pointer
// swap on the device).
void HipEwald::SetRecipRef(uint box) {
  if (box < BOXES_WITH_U_NB) dev.UpdateRecip();
}

void HipEwald::UpdateRecip(uint box) {
  if (box < BOXES_WITH_U_NB) dev.UpdateRecip();
}

void HipEwald::CopyRecip(uint box) {
  if (box < BOXES_WITH_U_NB) dev.CopyRecip();
}

// EwaldCached's per-molecule cache is staged on the device by the trial
// and committed by UpdateRecip; a rejected trial leaves it as it was, so
// the restores have nothing to undo.
void HipEwald::RestoreMol(int molIndex) {}
void HipEwald::exgMolCache() {}
void HipEwald::backupMolCache() {}

void HipEwald::UpdateVectorsAndRecipTerms(bool output) {
  dev.BoxReciprocalSetup();
  if (output)
    std::printf("%-40s %d\n", "Info: k-vectors of box 0 on the device", dev.NumKVectors());
}

// BoxForceReciprocal of the current new sums (MultiParticle::Prep calls
// CopyRecip first).  The forces are copied into the reference's arrays.
void HipEwald::BoxForceReciprocal(XYZArray const &molCoords, XYZArray &atomForceRec,
                                  XYZArray &molForceRec, uint box) {
  if (box >= BOXES_WITH_U_NB) return;
  OnDevice(box, "BoxForceReciprocal");
  dev.BoxForceReciprocal(atomForceRec.x, atomForceRec.y, atomForceRec.z, molForceRec.x,
                         molForceRec.y, molForceRec.z);
}

double HipEwald::ChangeLambdaRecip(XYZArray const &, const double, const double, const uint,
                                   const uint) {
  NotBuilt("ChangeLambdaRecip");
}

double HipEwald::MolExchangeReciprocal(const std::vector<cbmc::TrialMol> &,
                                       const std::vector<cbmc::TrialMol> &,
                                       const std::vector<uint> &, const std::vector<uint> &,
                                       bool) {
  NotBuilt("MolExchangeReciprocal");
}

void HipEwald::ChangeRecip(Energy *, Energy &, const std::vector<double> &, const uint,
                           const uint, const uint) const {
  NotBuilt("ChangeRecip");
}

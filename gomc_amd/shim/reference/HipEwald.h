// HipEwald -- the Ewald class of GOMC (src/Ewald.h:46-180) with its
// reciprocal-space work on the gfx950 path, and its EwaldCached mode
// (src/EwaldCached.h:11-68) as a flag.  Drop-in for the reference's src/
// tree: System::Init (src/System.cpp:132-148) constructs
//     calcEwald = new HipEwald(statV, *this, gpuBox, cachedFourier);
// instead of Ewald / EwaldCached.  gpuBox is the gomc_hip::Box of box 0
// (HipCalculateEnergy::Device()).
//
// Every override below forwards to the C ABI through gomc_hip::Box; the
// structure factors stay in HBM (no sumRnew / sumInew copies back).  The
// closed-form terms (BoxSelf, MolCorrection, SwapCorrection, SwapSelf,
// ChangeSelf, ChangeCorrection) stay the base class's host code: O(atoms of
// a molecule) per call.  The free-energy and MEMC reciprocal terms
// (ChangeLambdaRecip, ChangeRecip, MolExchangeReciprocal) are not on the
// device path and stop the run with an error, as a GPU build of the
// reference does for what it does not support.
//
// This file includes the reference's headers, which need the CMake-generated
// GOMC_Config.h: it is compiled in the reference's build, not in this
// repository.
#pragma once

#include "Ewald.h"
#include "gomc_shim.hpp"

class HipEwald : public Ewald {
public:
  HipEwald(StaticVals &stat, System &sys, gomc_hip::Box &dev, bool cached);
  ~HipEwald() override = default;

  void Init() override;
  void AllocMem() override;
  void RecipInit(uint box, BoxDimensions const &boxAxes) override;
  void BoxReciprocalSetup(uint box, XYZArray const &molCoords) override;
  void BoxReciprocalSums(uint box, XYZArray const &molCoords) override;
  double BoxReciprocal(uint box, bool isNewVolume) const override;
  Virial VirialReciprocal(Virial &virial, uint box) const override;
  double MolReciprocal(XYZArray const &molCoords, const uint molIndex,
                       const uint box) override;
  double SwapDestRecip(const cbmc::TrialMol &newMol, const uint box,
                       const int molIndex) override;
  double SwapSourceRecip(const cbmc::TrialMol &oldMol, const uint box,
                         const int molIndex) override;
  void SetRecipRef(uint box) override;
  void UpdateRecip(uint box) override;
  void CopyRecip(uint box) override;
  void UpdateRecipVec(uint box) override;
  void RestoreMol(int molIndex) override;
  void exgMolCache() override;
  void backupMolCache() override;
  void UpdateVectorsAndRecipTerms(bool output) override;
  void BoxForceReciprocal(XYZArray const &molCoords, XYZArray &atomForceRec,
                          XYZArray &molForceRec, uint box) override;

  double ChangeLambdaRecip(XYZArray const &molCoords, const double lambdaOld,
                           const double lambdaNew, const uint molIndex,
                           const uint box) override;
  double MolExchangeReciprocal(const std::vector<cbmc::TrialMol> &newMol,
                               const std::vector<cbmc::TrialMol> &oldMol,
                               const std::vector<uint> &molIndexNew,
                               const std::vector<uint> &molIndexOld,
                               bool first_call) override;
  void ChangeRecip(Energy *energyDiff, Energy &dUdL_Coul,
                   const std::vector<double> &lambda_Coul, const uint iState,
                   const uint molIndex, const uint box) const override;

private:
  // The device holds box 0 (the box with interactions in NVT, NPT and GCMC).
  void OnDevice(uint box, const char *what) const;
  [[noreturn]] void NotBuilt(const char *what) const;

  gomc_hip::Box &dev;
  bool cached;
};

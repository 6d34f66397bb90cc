// HipCellList -- GOMC's CellList (src/CellList.h:22-68) whose CSR view is
// built on the device.  Drop-in for the reference's src/ tree where a host
// caller still asks for GetCellListNeighbor (the energy calls that used it
// are HipCalculateEnergy's): System declares HipCellList cellList; and
// calls SetDevice once the device box exists.  The linked-list cells that
// the reference's moves maintain (AddMol / RemoveMol / GridAll) stay the
// base class's.  CellList's methods are not virtual: GetCellListNeighbor
// below hides the base class's with the same signature.
//
// Compiled in the reference's build (its headers need GOMC_Config.h).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "CellList.h"
#include "gomc_shim.hpp"

class HipCellList : public CellList {
public:
  using CellList::CellList;

  void SetDevice(gomc_hip::Box *d) { dev = d; }

  // GetCellListNeighbor (CellList.cpp:258-285) of box 0: the counting sort
  // of the device's coordinates, atoms ascending within each cell, the same
  // cell numbering and the same CSR as the host builds.
  void GetCellListNeighbor(uint box, int coordinateSize, std::vector<int> &cellVector,
                           std::vector<int> &cellStartIndex,
                           std::vector<int> &mapParticleToCell) const {
    if (!dev || box != 0) {
      CellList::GetCellListNeighbor(box, coordinateSize, cellVector, cellStartIndex,
                                    mapParticleToCell);
      return;
    }
    dev->GetCellListNeighbor(cellVector, cellStartIndex, mapParticleToCell);
  }

private:
  gomc_hip::Box *dev = nullptr;
};

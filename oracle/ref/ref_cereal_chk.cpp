// Writes GOMC checkpoint archives with the reference's own cereal (the
// header-only library vendored under lib/cereal) and the reference's own
// MolSetup serializers (src/MolSetup.h), then prints the values it wrote as
// JSON.  Built by oracle/Makefile into oracle/_ref/; used only to write the
// golden archives under tests/golden/ that pin this build's reader and
// writer of the format (gomc_amd/gomc_chk.py).
//
// The reference's Checkpoint (src/Checkpoint.h:118-167), MoleculeLookup
// (src/MoleculeLookup.h:194-204) and pdb_setup::Atoms (src/PDBSetup.h:
// 117-136) headers pull in the CMake-generated GOMC_Config.h, so they cannot
// be compiled here; the three classes below restate their member lists and
// serialize() member order (the order is the format).  CheckpointOutput::
// saveCheckpointFile (src/CheckpointOutput.cpp:69-90) appends the lookup's
// raw arrays after the object; main() does the same.
//
// usage: ref_cereal_chk <box_total 1|2> <archive path>  (JSON on stdout)
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include <cereal/archives/binary.hpp>
#include <cereal/types/map.hpp>
#include <cereal/types/string.hpp>
#include <cereal/types/vector.hpp>

#include "MolSetup.h"

namespace mirror {

struct Lookup {  // src/MoleculeLookup.h:157-204
  uint32_t molLookupCount = 0, atomCount = 0, boxAndKindStartLength = 0,
           boxAndKindSwappableLength = 0, numKinds = 0;
  std::vector<uint32_t> fixedMolecule, canSwapKind, canMoveKind;
  std::vector<uint32_t> molLookup, boxAndKindStart, boxAndKindSwappableCounts;
  std::vector<int32_t> molIndex, atomIndex, molKind, atomKind;
  std::vector<double> atomCharge;
  template <class Archive> void serialize(Archive &ar, const unsigned int) {
    ar & molLookupCount;
    ar & atomCount;
    ar & boxAndKindStartLength;
    ar & boxAndKindSwappableLength;
    ar & numKinds;
    ar & fixedMolecule;
    ar & canSwapKind;
    ar & canMoveKind;
  }
};

template <int NB> struct PdbAtoms {  // src/PDBSetup.h:97-136
  std::vector<char> chainLetter;
  std::vector<double> x, y, z, beta, occ;
  std::vector<uint> box;
  std::vector<std::string> resNames;
  bool restart = false, firstResInFile = false, recalcTrajectory = false;
  uint currBox = 0, count = 0;
  uint numAtomsInBox[NB] = {};
  template <class Archive> void serialize(Archive &ar, const unsigned int) {
    ar & chainLetter;
    ar & x;
    ar & y;
    ar & z;
    ar & beta;
    ar & occ;
    ar & box;
    ar & resNames;
    ar & restart;
    ar & firstResInFile;
    ar & recalcTrajectory;
    ar & currBox;
    ar & count;
    ar & numAtomsInBox;
  }
};

template <int NB> struct Checkpoint {  // src/Checkpoint.h:71-167
  char gomc_version[5] = {'2', '.', '8', '0', 0};
  uint64_t stepNumber = 0, trueStepNumber = 0;
  std::vector<uint32_t> originalStartVec, restartedStartVec;
  uint32_t saveArray[625] = {};
  uint32_t seedLocation = 0, seedLeft = 0, seedValue = 0;
  std::vector<std::vector<std::vector<double>>> scaleVec, acceptPercentVec;
  std::vector<std::vector<std::vector<uint32_t>>> acceptedVec, triesVec, tempAcceptedVec,
      tempTriesVec;
  std::vector<std::vector<uint32_t>> mp_acceptedVec, mp_triesVec, mp_interval_acceptedVec,
      mp_interval_triesVec;
  std::vector<double> mp_r_maxVec, mp_t_maxVec;
  std::vector<bool> isSingleMoveAcceptedVec;
  MolSetup originalMolSetup;
  PdbAtoms<NB> originalAtoms;
  Lookup originalMoleculeLookup;
  template <class Archive> void serialize(Archive &ar, const unsigned int) {
    ar & gomc_version;
    ar & stepNumber;
    ar & trueStepNumber;
    ar & saveArray;
    ar & seedLocation;
    ar & seedLeft;
    ar & seedValue;
    ar & scaleVec;
    ar & acceptPercentVec;
    ar & acceptedVec;
    ar & triesVec;
    ar & tempAcceptedVec;
    ar & tempTriesVec;
    ar & mp_triesVec;
    ar & mp_acceptedVec;
    ar & mp_interval_acceptedVec;
    ar & mp_interval_triesVec;
    ar & mp_t_maxVec;
    ar & mp_r_maxVec;
    ar & isSingleMoveAcceptedVec;
    ar & originalStartVec;
    ar & restartedStartVec;
    ar & originalMoleculeLookup;
    ar & originalMolSetup;
    ar & originalAtoms;
  }
};

}  // namespace mirror

// ---------------------------------------------------------------- values --
// A small deterministic generator so the fixture holds varied numbers.
static uint64_t gState = 0x9E3779B97F4A7C15ull;
static uint32_t Next() {
  gState ^= gState << 13;
  gState ^= gState >> 7;
  gState ^= gState << 17;
  return uint32_t(gState >> 11);
}
static double Real() { return double(Next()) / 4294967296.0 * 10.0 - 5.0; }

template <int NB> static void Fill(mirror::Checkpoint<NB> &c) {
  const int moveKinds = NB == 2 ? 13 : 9, kinds = 2;
  c.stepNumber = 123457;
  c.trueStepNumber = 223457;
  for (int i = 0; i < 624; ++i) c.saveArray[i] = Next();
  c.saveArray[624] = 417;
  c.seedLocation = 207;
  c.seedLeft = 417;
  c.seedValue = 12345;
  auto dims3 = [&](auto &v, auto gen) {
    v.resize(NB);
    for (auto &b : v) {
      b.resize(moveKinds);
      for (auto &m : b) {
        m.resize(kinds);
        for (auto &e : m) e = gen();
      }
    }
  };
  dims3(c.scaleVec, [] { return Real(); });
  dims3(c.acceptPercentVec, [] { return Real(); });
  dims3(c.acceptedVec, [] { return Next() % 1000; });
  dims3(c.triesVec, [] { return Next() % 5000; });
  dims3(c.tempAcceptedVec, [] { return Next() % 100; });
  dims3(c.tempTriesVec, [] { return Next() % 500; });
  for (auto *v : {&c.mp_acceptedVec, &c.mp_triesVec, &c.mp_interval_acceptedVec,
                  &c.mp_interval_triesVec}) {
    v->resize(NB);
    for (auto &b : *v) b = {Next() % 100, Next() % 100};
  }
  for (int b = 0; b < NB; ++b) {
    c.mp_t_maxVec.push_back(Real());
    c.mp_r_maxVec.push_back(Real());
  }
  c.isSingleMoveAcceptedVec = {true};
  // five molecules: three of kind MOLA (3 atoms), two of MOLB (1 atom)
  const std::vector<int> molKind = {0, 1, 0, 1, 0};
  uint32_t at = 0;
  c.originalStartVec.push_back(0);
  for (int k : molKind) c.originalStartVec.push_back(at += (k == 0 ? 3 : 1));
  c.restartedStartVec = {0, 3, 6, 9, 10, 11};
  mirror::Lookup &L = c.originalMoleculeLookup;
  L.molLookupCount = 5;
  L.atomCount = at;
  L.numKinds = kinds;
  L.boxAndKindStartLength = NB * kinds + 1;
  L.boxAndKindSwappableLength = NB * kinds;
  L.fixedMolecule = {0, 0, 0, 0, 0};
  L.canSwapKind = {0, 1};
  L.canMoveKind = {0, 1};
  L.molLookup = {0, 2, 4, 1, 3};
  L.boxAndKindStart = NB == 2 ? std::vector<uint32_t>{0, 2, 3, 4, 5}
                              : std::vector<uint32_t>{0, 3, 5};
  L.boxAndKindSwappableCounts =
      NB == 2 ? std::vector<uint32_t>{2, 1, 1, 1} : std::vector<uint32_t>{3, 2};
  for (int m = 0; m < 5; ++m)
    for (int a = 0; a < (molKind[m] == 0 ? 3 : 1); ++a) {
      L.molIndex.push_back(m);
      L.atomIndex.push_back(a);
      L.molKind.push_back(molKind[m]);
      L.atomKind.push_back(molKind[m] == 0 ? (a == 0 ? 0 : 1) : 2);
      L.atomCharge.push_back(molKind[m] == 0 ? (a == 0 ? -0.834 : 0.417) : 0.0);
    }
  // MolSetup: kind MOLA with bonds and an angle, kind MOLB with none, so the
  // class versions of Bond and Angle are first written inside MOLA and those
  // of Dihedral/Improper never
  mol_setup::MolKind a, b;
  a.atoms = {mol_setup::Atom("OH2", "TIP3", 1, "WAT", "OT", -0.834, 15.9994),
             mol_setup::Atom("H1", "TIP3", 1, "WAT", "HT", 0.417, 1.008),
             mol_setup::Atom("H2", "TIP3", 1, "WAT", "HT", 0.417, 1.008)};
  for (size_t i = 0; i < a.atoms.size(); ++i) a.atoms[i].kind = i == 0 ? 0 : 1;
  a.bonds = {mol_setup::Bond(0, 2), mol_setup::Bond(0, 1)};
  a.bonds[0].kind = 0;
  a.bonds[1].kind = 0;
  a.angles = {mol_setup::Angle(1, 0, 2)};
  a.angles[0].kind = 3;
  a.kindIndex = 0;
  a.firstAtomID = 1;
  a.firstMolID = 1;
  a.incomplete = true;
  a.isMultiResidue = false;
  a.moleculeName = "TIP3";
  b.atoms = {mol_setup::Atom("C1", "CH4", 2, "MET", "CH4", 0.0, 16.043)};
  b.atoms[0].kind = 2;
  b.kindIndex = 1;
  b.firstAtomID = 4;
  b.firstMolID = 2;
  b.incomplete = false;
  b.isMultiResidue = false;
  b.intraMoleculeResIDs = {0};
  b.moleculeName = "CH4";
  c.originalMolSetup.kindMap["MOLA"] = a;
  c.originalMolSetup.kindMap["MOLB"] = b;
  mol_setup::MoleculeVariables &mv = c.originalMolSetup.molVars;
  mv.startIdxMolecules = {0, 3, 4, 7, 8};
  mv.moleculeKinds = {0, 1, 0, 1, 0};
  mv.moleculeNames = {"TIP3", "CH4", "TIP3", "CH4", "TIP3"};
  mv.moleculeKindNames = {"TIP3", "CH4"};
  mv.moleculeSegmentNames = {"WAT", "MET", "WAT", "MET", "WAT"};
  mv.uniqueMapKeys = {"MOLA", "MOLB"};
  mv.lastAtomIndexInBox0 = 8;
  mv.numberMolsInBox0 = 5;
  mv.molKindIndex = 2;
  mv.stringSuffixMultiResidue = 0;
  mv.stringSuffixNonMultiResidue = 2;
  mv.moleculeIteration = 5;
  auto &P = c.originalAtoms;
  for (uint32_t i = 0; i < at; ++i) {
    P.chainLetter.push_back(char('A' + i % 3));
    P.x.push_back(Real());
    P.y.push_back(Real());
    P.z.push_back(Real());
    P.beta.push_back(0.0);
    P.occ.push_back(i < 8 || NB == 1 ? 0.0 : 1.0);
    P.box.push_back(i < 8 || NB == 1 ? 0 : 1);
    P.resNames.push_back(i < 9 ? (i % 4 == 3 ? "CH4" : "TIP3") : "CH4");
  }
  P.restart = false;
  P.firstResInFile = true;
  P.recalcTrajectory = false;
  P.currBox = NB - 1;
  P.count = at;
  P.numAtomsInBox[0] = NB == 2 ? 8 : at;
  if (NB == 2) P.numAtomsInBox[NB - 1] = at - 8;
}

// ------------------------------------------------------------------ JSON --
static void U32s(const char *name, const std::vector<uint32_t> &v, bool comma = true) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%" PRIu32, i ? ", " : "", v[i]);
  std::printf("]%s\n", comma ? "," : "");
}
static void I32s(const char *name, const std::vector<int32_t> &v) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%" PRId32, i ? ", " : "", v[i]);
  std::printf("],\n");
}
static void F64s(const char *name, const std::vector<double> &v, bool comma = true) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
  std::printf("]%s\n", comma ? "," : "");
}
static void Strs(const char *name, const std::vector<std::string> &v, bool comma = true) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf("%s\"%s\"", i ? ", " : "", v[i].c_str());
  std::printf("]%s\n", comma ? "," : "");
}
template <class T> static void Nest3(const char *name, const T &v, bool real) {
  std::printf("\"%s\": [", name);
  for (size_t b = 0; b < v.size(); ++b) {
    std::printf("%s[", b ? ", " : "");
    for (size_t m = 0; m < v[b].size(); ++m) {
      std::printf("%s[", m ? ", " : "");
      for (size_t k = 0; k < v[b][m].size(); ++k) {
        if (real)
          std::printf("%s%.17g", k ? ", " : "", double(v[b][m][k]));
        else
          std::printf("%s%" PRIu64, k ? ", " : "", uint64_t(v[b][m][k]));
      }
      std::printf("]");
    }
    std::printf("]");
  }
  std::printf("],\n");
}
static void Nest2(const char *name, const std::vector<std::vector<uint32_t>> &v) {
  std::printf("\"%s\": [", name);
  for (size_t b = 0; b < v.size(); ++b) {
    std::printf("%s[", b ? ", " : "");
    for (size_t k = 0; k < v[b].size(); ++k) std::printf("%s%" PRIu32, k ? ", " : "", v[b][k]);
    std::printf("]");
  }
  std::printf("],\n");
}

template <class B> static void Bonds(const char *name, const std::vector<B> &v, int n) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) {
    const uint *p = &v[i].a0;
    std::printf("%s[", i ? ", " : "");
    for (int j = 0; j < n; ++j) std::printf("%u, ", p[j]);
    std::printf("%u]", v[i].kind);
  }
  std::printf("],\n");
}

template <int NB> static void Dump(const mirror::Checkpoint<NB> &c, const char *path) {
  std::printf("{\n\"box_total\": %d,\n\"archive\": \"%s\",\n", NB, path);
  std::printf("\"gomc_version\": [%d, %d, %d, %d, %d],\n", c.gomc_version[0], c.gomc_version[1],
              c.gomc_version[2], c.gomc_version[3], c.gomc_version[4]);
  std::printf("\"step\": %" PRIu64 ",\n\"true_step\": %" PRIu64 ",\n", c.stepNumber,
              c.trueStepNumber);
  U32s("save_array", std::vector<uint32_t>(c.saveArray, c.saveArray + 625));
  std::printf("\"seed_location\": %u,\n\"seed_left\": %u,\n\"seed_value\": %u,\n", c.seedLocation,
              c.seedLeft, c.seedValue);
  Nest3("scale", c.scaleVec, true);
  Nest3("accept_percent", c.acceptPercentVec, true);
  Nest3("accepted", c.acceptedVec, false);
  Nest3("tries", c.triesVec, false);
  Nest3("temp_accepted", c.tempAcceptedVec, false);
  Nest3("temp_tries", c.tempTriesVec, false);
  Nest2("mp_tries", c.mp_triesVec);
  Nest2("mp_accepted", c.mp_acceptedVec);
  Nest2("mp_interval_accepted", c.mp_interval_acceptedVec);
  Nest2("mp_interval_tries", c.mp_interval_triesVec);
  F64s("mp_t_max", c.mp_t_maxVec);
  F64s("mp_r_max", c.mp_r_maxVec);
  std::printf("\"is_single_move_accepted\": [");
  for (size_t i = 0; i < c.isSingleMoveAcceptedVec.size(); ++i)
    std::printf("%s%s", i ? ", " : "", c.isSingleMoveAcceptedVec[i] ? "true" : "false");
  std::printf("],\n");
  U32s("original_start", c.originalStartVec);
  U32s("restarted_start", c.restartedStartVec);
  const mirror::Lookup &L = c.originalMoleculeLookup;
  std::printf("\"lookup\": {\"mol_lookup_count\": %u, \"atom_count\": %u, "
              "\"box_and_kind_start_length\": %u, \"box_and_kind_swappable_length\": %u, "
              "\"num_kinds\": %u,\n",
              L.molLookupCount, L.atomCount, L.boxAndKindStartLength,
              L.boxAndKindSwappableLength, L.numKinds);
  U32s("fixed_molecule", L.fixedMolecule);
  U32s("can_swap_kind", L.canSwapKind);
  U32s("can_move_kind", L.canMoveKind);
  U32s("mol_lookup", L.molLookup);
  U32s("box_and_kind_start", L.boxAndKindStart);
  U32s("box_and_kind_swappable_counts", L.boxAndKindSwappableCounts);
  I32s("mol_index", L.molIndex);
  I32s("atom_index", L.atomIndex);
  I32s("mol_kind", L.molKind);
  I32s("atom_kind", L.atomKind);
  F64s("atom_charge", L.atomCharge, false);
  std::printf("},\n\"mol_setup\": {\"kind_map\": [\n");
  bool first = true;
  for (const auto &kv : c.originalMolSetup.kindMap) {
    const mol_setup::MolKind &k = kv.second;
    std::printf("%s[\"%s\", {\"atoms\": [", first ? "" : ",\n", kv.first.c_str());
    first = false;
    for (size_t i = 0; i < k.atoms.size(); ++i) {
      const mol_setup::Atom &a = k.atoms[i];
      std::printf("%s{\"name\": \"%s\", \"type\": \"%s\", \"residue\": \"%s\", \"segment\": "
                  "\"%s\", \"charge\": %.17g, \"mass\": %.17g, \"residue_id\": %u, "
                  "\"kind\": %u}",
                  i ? ", " : "", a.name.c_str(), a.type.c_str(), a.residue.c_str(),
                  a.segment.c_str(), a.charge, a.mass, a.residueID, a.kind);
    }
    std::printf("],\n");
    Bonds("bonds", k.bonds, 2);
    Bonds("angles", k.angles, 3);
    Bonds("dihedrals", k.dihedrals, 4);
    Bonds("impropers", k.impropers, 4);
    Bonds("donors", k.donors, 2);
    Bonds("acceptors", k.acceptors, 2);
    std::printf("\"kind_index\": %u, \"first_atom_id\": %u, \"first_mol_id\": %u, "
                "\"incomplete\": %s, \"is_multi_residue\": %s,\n",
                k.kindIndex, k.firstAtomID, k.firstMolID, k.incomplete ? "true" : "false",
                k.isMultiResidue ? "true" : "false");
    U32s("intra_molecule_res_ids", std::vector<uint32_t>(k.intraMoleculeResIDs.begin(),
                                                          k.intraMoleculeResIDs.end()));
    std::printf("\"molecule_name\": \"%s\"}]", k.moleculeName.c_str());
  }
  const mol_setup::MoleculeVariables &mv = c.originalMolSetup.molVars;
  std::printf("],\n\"mol_vars\": {\n");
  U32s("start_idx_molecules",
       std::vector<uint32_t>(mv.startIdxMolecules.begin(), mv.startIdxMolecules.end()));
  U32s("molecule_kinds", std::vector<uint32_t>(mv.moleculeKinds.begin(), mv.moleculeKinds.end()));
  Strs("molecule_names", mv.moleculeNames);
  Strs("molecule_kind_names", mv.moleculeKindNames);
  Strs("molecule_segment_names", mv.moleculeSegmentNames);
  Strs("unique_map_keys", mv.uniqueMapKeys);
  std::printf("\"last_atom_index_in_box0\": %u, \"number_mols_in_box0\": %u, "
              "\"mol_kind_index\": %u, \"string_suffix_multi_residue\": %u, "
              "\"string_suffix_non_multi_residue\": %u, \"molecule_iteration\": %u}},\n",
              mv.lastAtomIndexInBox0, mv.numberMolsInBox0, mv.molKindIndex,
              mv.stringSuffixMultiResidue, mv.stringSuffixNonMultiResidue, mv.moleculeIteration);
  const auto &P = c.originalAtoms;
  std::printf("\"pdb_atoms\": {\"chain_letter\": \"");
  for (char ch : P.chainLetter) std::printf("%c", ch);
  std::printf("\",\n");
  F64s("x", P.x);
  F64s("y", P.y);
  F64s("z", P.z);
  F64s("beta", P.beta);
  F64s("occ", P.occ);
  U32s("box", std::vector<uint32_t>(P.box.begin(), P.box.end()));
  Strs("res_names", P.resNames);
  std::printf("\"restart\": %s, \"first_res_in_file\": %s, \"recalc_trajectory\": %s, "
              "\"curr_box\": %u, \"count\": %u,\n",
              P.restart ? "true" : "false", P.firstResInFile ? "true" : "false",
              P.recalcTrajectory ? "true" : "false", P.currBox, P.count);
  U32s("num_atoms_in_box", std::vector<uint32_t>(P.numAtomsInBox, P.numAtomsInBox + NB), false);
  std::printf("}\n}\n");
}

template <int NB> static int Run(const char *path) {
  mirror::Checkpoint<NB> c;
  Fill(c);
  {
    std::ofstream ofs(path, std::ios::binary);
    if (!ofs) return 1;
    cereal::BinaryOutputArchive oa(ofs);
    oa << c;
    const mirror::Lookup &L = c.originalMoleculeLookup;
    oa(cereal::binary_data(L.molLookup.data(), sizeof(uint32_t) * L.molLookup.size()));
    oa(cereal::binary_data(L.boxAndKindStart.data(), sizeof(uint32_t) * L.boxAndKindStart.size()));
    oa(cereal::binary_data(L.boxAndKindSwappableCounts.data(),
                           sizeof(uint32_t) * L.boxAndKindSwappableCounts.size()));
    oa(cereal::binary_data(L.molIndex.data(), sizeof(int32_t) * L.molIndex.size()));
    oa(cereal::binary_data(L.atomIndex.data(), sizeof(int32_t) * L.atomIndex.size()));
    oa(cereal::binary_data(L.molKind.data(), sizeof(int32_t) * L.molKind.size()));
    oa(cereal::binary_data(L.atomKind.data(), sizeof(int32_t) * L.atomKind.size()));
    oa(cereal::binary_data(L.atomCharge.data(), sizeof(double) * L.atomCharge.size()));
  }
  // read it back through cereal: the same values come out
  mirror::Checkpoint<NB> r;
  {
    std::ifstream ifs(path, std::ios::binary);
    cereal::BinaryInputArchive ia(ifs);
    ia >> r;
    if (r.stepNumber != c.stepNumber || r.originalMolSetup.kindMap.size() != 2 ||
        r.originalAtoms.count != c.originalAtoms.count)
      return 2;
  }
  Dump(c, path);
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <box_total 1|2> <archive path>\n", argv[0]);
    return 1;
  }
  const int nb = std::atoi(argv[1]);
  if (nb == 1) return Run<1>(argv[2]);
  if (nb == 2) return Run<2>(argv[2]);
  return 1;
}

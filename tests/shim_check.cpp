// Driver for the C++ shim test (tests/test_shim.py): builds a box from a
// text description written by the test, calls the shim's CalculateEnergy /
// Ewald / CellList methods and prints their results as JSON.  Without a
// device it prints the error the shim raised.
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "../gomc_amd/shim/gomc_shim.hpp"

template <class T> std::vector<T> ReadVec(std::ifstream &f) {
  size_t n = 0;
  f >> n;
  std::vector<T> v(n);
  for (auto &x : v) f >> x;
  return v;
}

template <class T> void PrintVec(const char *name, const std::vector<T> &v, bool last = false) {
  std::printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); ++i) std::printf(i ? ", %.17g" : "%.17g", double(v[i]));
  std::printf("]%s\n", last ? "" : ",");
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  std::ifstream f(argv[1]);
  gomc_system_desc d{};
  auto box = ReadVec<double>(f);      // Lx Ly Lz T rcut rcut_low rcut_coulomb tol
  auto flags = ReadVec<int>(f);       // ewald electrostatic lrc
  auto sig = ReadVec<double>(f), eps = ReadVec<double>(f), nrep = ReadVec<double>(f);
  auto kn = ReadVec<int32_t>(f), kt = ReadVec<int32_t>(f);
  auto kq = ReadVec<double>(f);
  auto mk = ReadVec<int32_t>(f);
  auto x = ReadVec<double>(f), y = ReadVec<double>(f), z = ReadVec<double>(f);
  auto trial = ReadVec<double>(f);    // mol, then x, y, z of its atoms
  auto part = ReadVec<double>(f);     // atom type, charge, then trial x, y, z triples
  for (int i = 0; i < 3; ++i) d.box[i] = box[size_t(i)];
  d.temperature = box[3];
  d.rcut = box[4];
  d.rcut_low = box[5];
  d.rcut_coulomb = box[6];
  d.tolerance = box[7];
  d.ewald = flags[0];
  d.electrostatic = flags[1];
  d.lrc = flags[2];
  d.threads = 1;
  d.n_atom_types = int32_t(sig.size());
  d.sigma = sig.data();
  d.epsilon = eps.data();
  d.n_rep = nrep.data();
  d.n_kinds = int32_t(kn.size());
  d.kind_natoms = kn.data();
  d.kind_atom_type = kt.data();
  d.kind_charge = kq.data();
  d.n_mols = int32_t(mk.size());
  d.mol_kind = mk.data();
  d.x = x.data();
  d.y = y.data();
  d.z = z.data();
  d.seed = 1;
  d.move_disp = d.move_rot = 0.5;
  d.adj_steps = 1000;
  d.cbmc_first = 16;
  d.cbmc_nth = 8;
  d.cbmc_ang = 100;
  try {
    gomc_hip::Box b(d, 0);
    std::printf("{\n");
    const gomc_hip::Energy e = b.SystemTotal();
    PrintVec("total", std::vector<double>{e.inter, e.tailCorrection, e.real, e.recip, e.self,
                                          e.correction, e.total});
    const double erecip = b.BoxReciprocalSetup();
    std::printf("\"recip_setup\": %.17g,\n", erecip);
    const int mol = int(trial[0]);
    const size_t na = (trial.size() - 1) / 3;
    std::vector<double> tx(na), ty(na), tz(na);
    for (size_t a = 0; a < na; ++a) {
      tx[a] = trial[1 + 3 * a];
      ty[a] = trial[2 + 3 * a];
      tz[a] = trial[3 + 3 * a];
    }
    double lj = 0, coul = 0;
    const bool ov = b.MoleculeInter(lj, coul, tx.data(), ty.data(), tz.data(), mol);
    const double drecip = b.MolReciprocal(tx.data(), ty.data(), tz.data(), mol);
    std::printf("\"mol_inter\": [%.17g, %.17g, %d],\n\"mol_recip\": %.17g,\n", lj, coul,
                int(ov), drecip);
    // the trial accepted: the reference sums follow, and a second delta of
    // the same coordinates is zero
    b.UpdateRecip();
    b.CommitMol(mol, tx.data(), ty.data(), tz.data());
    const double again = b.MolReciprocal(tx.data(), ty.data(), tz.data(), mol);
    std::printf("\"recip_after_accept\": [%.17g, %.17g],\n", b.BoxReciprocal(), again);
    const int type = int(part[0]);
    const double q = part[1];
    const size_t nt = (part.size() - 2) / 3;
    std::vector<double> px(nt), py(nt), pz(nt), en(nt, 1.0), re(nt, 2.0);
    std::vector<char> pov(nt, 0);
    for (size_t t = 0; t < nt; ++t) {
      px[t] = part[2 + 3 * t];
      py[t] = part[3 + 3 * t];
      pz[t] = part[4 + 3 * t];
    }
    bool ovb[64] = {false};
    b.ParticleInter(en.data(), re.data(), px.data(), py.data(), pz.data(), ovb, type, q, -1,
                    int(nt));
    PrintVec("particle_en_plus_1", en);
    PrintVec("particle_real_plus_2", re);
    std::vector<int> cv, cs, mp;
    b.GetCellListNeighbor(cv, cs, mp);
    PrintVec("cell_vector", cv);
    PrintVec("cell_start", cs);
    const gomc_hip::Energy e2 = b.SystemTotal();
    PrintVec("total_after", std::vector<double>{e2.inter, e2.tailCorrection, e2.real, e2.recip,
                                                e2.self, e2.correction, e2.total});
    int code = 0;
    try {
      b.MolReciprocal(tx.data(), ty.data(), tz.data(), int(mk.size()) + 5);
    } catch (const gomc_hip::Error &err) {
      code = err.code();
    }
    std::printf("\"bad_mol_error\": %d\n}\n", code);
  } catch (const gomc_hip::Error &err) {
    std::printf("{\"error_code\": %d, \"error\": \"%s\"}\n", err.code(), err.what());
  }
  return 0;
}

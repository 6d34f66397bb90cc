"""GOMC input files (control file, PDB, PSF, CHARMM and Mie parameter files)
read into systems, checked against the fixtures the reference's own example
inputs hold (tests/golden/gomc_inputs: the AR_KR Gibbs/GCMC example and the
TIP3 water molecule with its CHARMM stream file) and against closed forms:

* AR_KR: the values the files state (Vrabec 2001 Mie parameters in K and A,
  34 A box, 143.15 K, Rcut 10, LRC on), the three box-0 atoms all farther
  apart than Rcut under the minimum image (LJ energy exactly zero), and the
  LJ tail energy equal to the textbook mixture tail
  sum_ij N_i N_j / (2V) 16 pi eps_ij [sig^12/(9 rc^9) - sig^6/(3 rc^3)];
  with Rcut 13 the one pair inside the cutoff gives 4 eps [(s/r)^12-(s/r)^6];
* TIP3: CHARMM NONBONDED values converted with the reference's constants
  (eps_K = -eps * 503.21959899, sigma = Rmin/2 * 1.7817974362807), the two
  PSF bonds making one 3-atom molecule whose body frame is the PDB's
  coordinates relative to the oxygen;
* a system written with write_gomc and read back has the same force field,
  kinds and ensemble settings, coordinates rounded to the PDB's 3 decimals,
  and the oracle's energies of the re-read system equal those of the
  original with its coordinates rounded the same way.
"""
import math
import os
import shutil
from dataclasses import replace

import numpy as np
import pytest

from gomc_amd import gomc_io, systems
from oracle.pyoracle import Oracle

HERE = os.path.dirname(__file__)
ARKR = os.path.join(HERE, "golden", "gomc_inputs", "AR_KR")
TIP3 = os.path.join(HERE, "golden", "gomc_inputs", "TIP3")
CONF = os.path.join(ARKR, "Base", "in.conf")


def lj(eps, sig, r):
    x = (sig / r) ** 6
    return 4.0 * eps * (x * x - x)


def tail(eps, sig, rc):
    return 16.0 * math.pi * eps * (sig ** 12 / (9 * rc ** 9) - sig ** 6 / (3 * rc ** 3))


def test_ar_kr_control_file_values():
    c = gomc_io.read_conf(CONF)
    assert not c.charmm and c.parameters[0].endswith("Par_Vrabec01_Noble_Gases_Exotic.inp")
    assert c.temperature == 143.15 and c.rcut == 10.0 and c.lrc and not c.ewald
    assert not c.electrostatic and c.seed == 123 and c.adj_steps == 10
    assert np.array_equal(c.cell[0], np.diag([34.0] * 3))
    assert np.array_equal(c.cell[1], np.diag([65.0] * 3))
    assert c.freq == {"disp": 0.87, "rot": 0.129, "vol": 0.0, "swap": 0.001}
    assert c.chem_pot == {"AR": -968.0, "KR": -968.0}
    assert c.cbmc == {"first": 10, "nth": 8, "ang": 100, "dih": 30}


def test_mie_parameter_file():
    p = gomc_io.read_parameters(os.path.join(ARKR, "Par_Vrabec01_Noble_Gases_Exotic.inp"),
                                charmm=False)
    assert (p["AR"].epsilon, p["AR"].sigma, p["AR"].n) == (116.79, 3.3952, 12.0)
    assert (p["KR"].epsilon, p["KR"].sigma, p["KR"].n) == (162.58, 3.6274, 12.0)
    assert (p["XE"].epsilon, p["XE"].sigma) == (227.55, 3.9011)
    with pytest.raises(gomc_io.GomcFormatError):
        gomc_io.read_parameters(os.path.join(ARKR, "Par_Vrabec01_Noble_Gases_Exotic.inp"),
                                charmm=True)


def arkr_nvt(tmp_path, rcut=None):
    """The AR_KR example set up for the NVT binary: GOMC_CPU_NVT does not
    read SwapFreq, so the example's displacement frequency takes its 0.001
    for the frequencies to sum to one (ConfigSetup.cpp:2077-2085)."""
    base = tmp_path / "Base"
    shutil.copytree(os.path.join(ARKR, "Base"), base)
    shutil.copy(os.path.join(ARKR, "Par_Vrabec01_Noble_Gases_Exotic.inp"), tmp_path)
    text = (base / "in.conf").read_text().replace("DisFreq               0.87",
                                                  "DisFreq               0.871")
    if rcut is not None:
        text = text.replace("Rcut\t\t10", f"Rcut\t\t{rcut}")
    (base / "in.conf").write_text(text)
    return str(base / "in.conf")


def test_ar_kr_nvt_energy_is_the_lj_tail_only(tmp_path):
    # read as the GCMC example it is, the frequencies sum to one; the NVT
    # binary does not read SwapFreq, so there they fall short
    with pytest.raises(gomc_io.GomcFormatError, match="sum of move frequencies"):
        gomc_io.system_from_gomc(CONF, ensemble="NVT")
    s = gomc_io.system_from_gomc(arkr_nvt(tmp_path), ensemble="NVT")
    assert s.ensemble == 0 and s.nmols == 3 and s.box == (34.0, 34.0, 34.0)
    assert [k.name for k in s.kinds] == ["AR", "KR"] and list(s.mol_kind) == [0, 1, 0]
    e = Oracle(s).system_total()
    assert e[0] == 0.0 and e[2] == 0.0 and e[3] == 0.0
    V = 34.0 ** 3
    sig_ak, eps_ak = (3.3952 + 3.6274) / 2, math.sqrt(116.79 * 162.58)
    expect = (2 * 2 * tail(116.79, 3.3952, 10.0) + 2 * 2 * 1 * tail(eps_ak, sig_ak, 10.0)
              + 1 * 1 * tail(162.58, 3.6274, 10.0)) / (2 * V)
    assert e[1] == pytest.approx(expect, rel=1e-12)
    assert e[6] == pytest.approx(expect, rel=1e-12)


def test_ar_kr_one_pair_inside_a_longer_cutoff(tmp_path):
    s = gomc_io.system_from_gomc(arkr_nvt(tmp_path, rcut=13), ensemble="NVT")
    assert s.rcut == 13.0
    d = np.array([8.340, 31.348, 31.671]) - np.array([13.326, 21.201, 26.910])
    r = float(np.sqrt(d @ d))  # 12.27 A; the other two pairs are 13.30 and 20.98 A
    e = Oracle(s).system_total()
    assert e[0] == pytest.approx(lj(math.sqrt(116.79 * 162.58), (3.3952 + 3.6274) / 2, r),
                                 rel=1e-12)


def test_ar_kr_gcmc_reservoir_and_chemical_potential():
    s = gomc_io.system_from_gomc(CONF)
    assert s.ensemble == 1 and list(s.mol_box) == [0, 0, 0, 1, 1, 1]
    # beta f = exp(beta mu): the transfer coefficient of ChemPot -968 K
    for f in s.fugacity:
        beta_f = f * gomc_io.BAR_TO_K_MOLECULE_PER_A3 / 143.15
        assert beta_f == pytest.approx(math.exp(-968.0 / 143.15), rel=1e-14)
    assert s.move_swap == 0.001 and s.cbmc_first == 10
    # reservoir atoms (box 1 is 65 A) are wrapped into box 0's cell
    assert np.all((s.xyz >= 0) & (s.xyz < 34.0))
    assert np.allclose(s.xyz[3:], np.array([[16.323, 61.619, 62.256], [26.139, 44.297, 52.883],
                                            [30.844, 0.499, 9.084]]) % 34.0)
    o = Oracle(s)
    assert o.mol_boxes()[0] == 3


def test_tip3_charmm_stream_and_structure():
    p = gomc_io.read_parameters(os.path.join(TIP3, "GOMC_toppar_water_ions_namd.str"),
                                charmm=True)
    assert p["OT"].epsilon == pytest.approx(0.1521 * 503.21959899, rel=1e-15)
    assert p["OT"].sigma == pytest.approx(1.7682 * 1.7817974362807, rel=1e-15)
    assert p["HT"].epsilon == pytest.approx(0.046 * 503.21959899, rel=1e-15)
    assert p["HT"].sigma == pytest.approx(0.2245 * 1.7817974362807, rel=1e-15)
    assert p["OT"].n == p["HT"].n == 12.0
    psf = gomc_io.read_psf(os.path.join(TIP3, "water_reservior_1_molecule.psf"))
    assert psf.type == ["OT", "HT", "HT"] and list(psf.charge) == [-0.834, 0.417, 0.417]
    assert psf.bonds.tolist() == [[0, 1], [0, 2]]
    pdb = gomc_io.read_pdb(os.path.join(TIP3, "water_reservior_1_molecule.pdb"))
    assert pdb.cell == (30.0, 30.0, 30.0) and pdb.resname == ["TIP3"] * 3
    assert pdb.name == ["OH2", "H1", "H2"]
    assert pdb.xyz.tolist() == [[6.763, 15.473, 16.671], [6.966, 16.434, 16.863],
                                [7.582, 14.921, 16.831]]


def tip3_lattice(tmp_path, n_side=4, L=18.0, ewald=True):
    """The TIP3 molecule of the fixture copied onto an n^3 lattice, written
    as PDB/PSF, with a control file that names the CHARMM stream file."""
    psf = gomc_io.read_psf(os.path.join(TIP3, "water_reservior_1_molecule.psf"))
    pdb = gomc_io.read_pdb(os.path.join(TIP3, "water_reservior_1_molecule.pdb"))
    mol = pdb.xyz - pdb.xyz[0]
    g = (np.arange(n_side) + 0.5) * (L / n_side)
    centres = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1).T
    xyz = np.concatenate([c + mol for c in centres]) % L
    nm = len(centres)
    rep = lambda a: [a[i % 3] for i in range(3 * nm)]  # noqa: E731
    resid = np.repeat(np.arange(1, nm + 1), 3)
    gomc_io.write_pdb(str(tmp_path / "w.pdb"), xyz, rep(pdb.name), rep(pdb.resname), resid,
                      cell=(L, L, L))
    bonds = np.concatenate([psf.bonds + 3 * m for m in range(nm)])
    gomc_io.write_psf(str(tmp_path / "w.psf"), rep(psf.segment), resid, rep(psf.resname),
                      rep(psf.name), rep(psf.type), rep(list(psf.charge)), rep(list(psf.mass)),
                      bonds)
    onoff = "true" if ewald else "false"
    (tmp_path / "in.conf").write_text(f"""
ParaTypeCHARMM on
Parameters {os.path.join(TIP3, 'GOMC_toppar_water_ions_namd.str')}
Coordinates 0 w.pdb
Structure 0 w.psf
Temperature 298.0
Potential VDW
LRC true
Rcut 8.0
RcutLow 1.0
Exclude 1-3
Ewald {onoff}
ElectroStatic {onoff}
Tolerance 1e-5
Random_Seed 77
AdjSteps 100
DisFreq 0.6
RotFreq 0.4
CellBasisVector1 0 {L} 0.0 0.0
CellBasisVector2 0 0.0 {L} 0.0
CellBasisVector3 0 0.0 0.0 {L}
""")
    return str(tmp_path / "in.conf")


def test_tip3_lattice_system(tmp_path):
    s = gomc_io.system_from_gomc(tip3_lattice(tmp_path))
    assert s.ensemble == 0 and s.nmols == 64 and len(s.kinds) == 1
    k = s.kinds[0]
    assert k.name == "TIP3" and k.atom_types == [0, 1, 1] and k.charges == [-0.834, 0.417, 0.417]
    pdb = gomc_io.read_pdb(os.path.join(TIP3, "water_reservior_1_molecule.pdb"))
    assert np.allclose(k.geometry, pdb.xyz - pdb.xyz[0], rtol=0, atol=1e-12)
    assert s.epsilon == [0.1521 * 503.21959899, 0.046 * 503.21959899]
    assert s.ewald and s.electrostatic and s.lrc and s.rcut_coulomb == 8.0
    assert s.move_disp == 0.6 and s.move_rot == 0.4 and s.seed == 77
    e = Oracle(s).system_total()
    assert np.all(np.isfinite(e)) and e[0] != 0.0 and e[2] != 0.0 and e[3] > 0.0
    # Ewald on implies electrostatics on (ConfigSetup::fillDefaults)
    conf = tmp_path / "in.conf"
    conf.write_text(conf.read_text().replace("ElectroStatic true", "ElectroStatic false"))
    s2 = gomc_io.system_from_gomc(str(conf))
    assert s2.ewald and s2.electrostatic
    # the tail correction is on unless LRC says otherwise (ConfigSetup.cpp:73)
    conf.write_text(conf.read_text().replace("LRC true\n", ""))
    assert gomc_io.system_from_gomc(str(conf)).lrc


def _rounded(s):
    xyz = np.array([[float(f"{v:8.3f}") for v in p] for p in s.xyz])
    return replace(s, xyz=xyz, _keep=[])


@pytest.mark.parametrize("which", ["water", "gcmc", "npt"])
def test_write_and_read_back(tmp_path, which):
    if which == "water":
        s = systems.water_box(60, seed=4, rcut=6.0)
    elif which == "gcmc":
        s = systems.gcmc_mixture(20, 15, 24.0, fugacity_bar=(30.0, 12.0), reservoir=25, seed=9)
    else:
        s = replace(systems.water_box(60, seed=4, rcut=6.0), ensemble=2, pressure=3.5,
                    move_vol=0.02, _keep=[])
    conf = gomc_io.write_gomc(str(tmp_path), s, prefix="rt")
    ens = {0: "NVT", 1: "GCMC", 2: "NPT"}[s.ensemble]
    r = gomc_io.system_from_gomc(conf, ensemble=ens)
    assert r.ensemble == s.ensemble and r.box == tuple(s.box)
    for f in ("sigma", "epsilon", "n_rep"):
        assert list(getattr(r, f)) == list(getattr(s, f)), f
    for f in ("temperature", "rcut", "rcut_low", "rcut_coulomb", "tolerance", "lrc", "ewald",
              "electrostatic", "seed", "adj_steps", "cbmc_first", "cbmc_nth", "cbmc_ang"):
        assert getattr(r, f) == getattr(s, f), f
    assert [k.atom_types for k in r.kinds] == [k.atom_types for k in s.kinds]
    assert [k.charges for k in r.kinds] == [k.charges for k in s.kinds]
    assert np.array_equal(r.mol_kind, s.mol_kind)
    tot = s.move_disp + s.move_rot + s.move_mp + s.move_swap + s.move_vol
    assert r.move_disp == pytest.approx(s.move_disp / tot, rel=1e-15)
    if s.ensemble == 1:
        assert np.array_equal(r.mol_box, s.mol_box) and r.fugacity == s.fugacity
    if s.ensemble == 2:
        assert r.pressure == s.pressure and r.move_vol == pytest.approx(s.move_vol / tot)
    assert np.max(np.abs(r.xyz - s.xyz)) <= 5e-4 + 1e-12
    e_read = Oracle(r).system_total()
    e_orig = Oracle(_rounded(s)).system_total()
    assert np.allclose(e_read, e_orig, rtol=1e-12, atol=1e-9)


def test_refusals(tmp_path):
    with pytest.raises(gomc_io.GomcFormatError):
        gomc_io.system_from_gomc(CONF, ensemble="GEMC")
    conf = tip3_lattice(tmp_path)
    text = open(conf).read()
    open(conf, "w").write(text.replace("CellBasisVector2 0 0.0", "CellBasisVector2 0 1.0"))
    with pytest.raises(gomc_io.GomcFormatError, match="orthogonal"):
        gomc_io.system_from_gomc(conf)
    open(conf, "w").write(text.replace("Tolerance 1e-5", ""))
    with pytest.raises(gomc_io.GomcFormatError, match="Tolerance"):
        gomc_io.system_from_gomc(conf)
    open(conf, "w").write(text.replace("Potential VDW", "Potential SHIFT"))
    with pytest.raises(gomc_io.GomcFormatError, match="VDW"):
        gomc_io.system_from_gomc(conf)
    par = tmp_path / "fix.par"
    par.write_text("NONBONDED\nOT 0.0 -0.1521 1.7682\nHT 0.0 -0.046 0.2245\n\n"
                   "NBFIX\nOT HT -0.1 2.0\n")
    open(conf, "w").write(text.replace(os.path.join(TIP3, "GOMC_toppar_water_ions_namd.str"),
                                       str(par)))
    with pytest.raises(gomc_io.GomcFormatError, match="NBFIX"):
        gomc_io.system_from_gomc(conf)
    lines = open(tmp_path / "w.pdb").read().splitlines()
    open(tmp_path / "w.pdb", "w").write("\n".join(lines[:-2] + lines[-1:]) + "\n")
    open(conf, "w").write(text)
    with pytest.raises(gomc_io.GomcFormatError, match="PDB atoms"):
        gomc_io.system_from_gomc(conf)


def test_move_frequencies_checked_as_the_reference_does(tmp_path):
    """ConfigSetup.cpp:2036-2085: the frequencies an ensemble reads must sum
    to one, and a move that is not built must not be silently dropped."""
    conf = tmp_path / "in.conf"
    base = gomc_io.system_from_gomc(tip3_lattice(tmp_path))
    text = conf.read_text()
    # a nonzero frequency of an unbuilt move (the BPTI example sets this one)
    conf.write_text(text.replace("DisFreq 0.6", "DisFreq 0.5\nRegrowthFreq 0.1"))
    with pytest.raises(gomc_io.GomcFormatError, match="RegrowthFreq"):
        gomc_io.system_from_gomc(str(conf))
    conf.write_text(text.replace("DisFreq 0.6", "DisFreq 0.5\nIntraSwapFreq 0.1"))
    with pytest.raises(gomc_io.GomcFormatError, match="IntraSwapFreq"):
        gomc_io.system_from_gomc(str(conf))
    # a zero frequency of an unbuilt move is harmless
    conf.write_text(text.replace("DisFreq 0.6", "DisFreq 0.6\nCrankShaftFreq 0.0"))
    assert gomc_io.system_from_gomc(str(conf)).move_disp == base.move_disp
    # frequencies that do not sum to one
    conf.write_text(text.replace("DisFreq 0.6", "DisFreq 0.7"))
    with pytest.raises(gomc_io.GomcFormatError, match="sum of move frequencies"):
        gomc_io.system_from_gomc(str(conf))
    # within the reference's 0.001 tolerance
    conf.write_text(text.replace("DisFreq 0.6", "DisFreq 0.6005"))
    assert gomc_io.system_from_gomc(str(conf)).move_disp == 0.6005
    # DisFreq is required
    conf.write_text(text.replace("DisFreq 0.6\n", "RotFreq 0.6\n"))
    with pytest.raises(gomc_io.GomcFormatError, match="DisFreq"):
        gomc_io.system_from_gomc(str(conf))
    # a volume frequency is read only by an ensemble whose volume varies
    conf.write_text(text.replace("DisFreq 0.6", "DisFreq 0.6\nVolFreq 0.2"))
    assert gomc_io.system_from_gomc(str(conf), ensemble="NVT").move_vol == 0.0
    with pytest.raises(gomc_io.GomcFormatError, match="sum of move frequencies"):
        gomc_io.system_from_gomc(str(conf), ensemble="NPT")


def test_keys_match_without_case(tmp_path):
    """CheckString upper-cases both strings (ConfigSetup.cpp:194-210)."""
    conf = tmp_path / "in.conf"
    base = gomc_io.system_from_gomc(tip3_lattice(tmp_path))
    text = conf.read_text()
    conf.write_text(text.replace("DisFreq 0.6", "DISFREQ 0.6").replace("Rcut 8.0", "rcut 8.0"))
    s = gomc_io.system_from_gomc(str(conf))
    assert s.move_disp == base.move_disp and s.rcut == base.rcut


def test_ar_kr_example_frequencies_sum_to_one():
    """The AR_KR example (DisFreq 0.87, RotFreq 0.129, SwapFreq 0.001; its
    VolFreq line is not read by the GCMC binary) reads as given."""
    s = gomc_io.system_from_gomc(os.path.join(ARKR, "Base", "in.conf"), ensemble="GCMC")
    assert (s.move_disp, s.move_rot, s.move_swap) == (0.87, 0.129, 0.001)

/* C entry points of the CPU oracle, for ctypes.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * liboracle.so, and only as the checker. */
#ifndef GOMC_ORACLE_CAPI_H
#define GOMC_ORACLE_CAPI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One box: force field, molecule kinds, molecules, coordinates, moves.
 * Same field layout as gomc_system_desc in include/gomc_hip.h. */
typedef struct orc_system_desc {
  double box[3];
  double temperature;
  double rcut, rcut_low, rcut_coulomb, tolerance;
  int32_t ewald, electrostatic, lrc, threads;
  int32_t n_atom_types, n_kinds;
  const double *sigma, *epsilon, *n_rep;      /* [n_atom_types] */
  const int32_t *kind_natoms;                 /* [n_kinds] */
  const int32_t *kind_atom_type;              /* [sum kind_natoms] */
  const double *kind_charge;                  /* [sum kind_natoms] */
  int32_t n_mols, _pad0;
  const int32_t *mol_kind;                    /* [n_mols] */
  const double *x, *y, *z;                    /* [n_atoms] */
  uint64_t seed;
  double move_disp, move_rot, move_mp;
  int64_t adj_steps;
  /* GCMC (ensemble 1): molecules with mol_box[m] == 0 are in the simulation
   * box; mol_box[m] == 1 puts a molecule in the non-interacting reservoir
   * (GOMC's second GCMC box), given the same edge lengths. */
  int32_t ensemble;                        /* 0 NVT, 1 GCMC, 2 NPT */
  int32_t cbmc_first, cbmc_nth, cbmc_ang;  /* CBMC_First, CBMC_Nth, CBMC_Ang */
  const int32_t *mol_box;                  /* [n_mols]; NULL: all in box 0 */
  const double *kind_fugacity;             /* [n_kinds] Fugacity, bar */
  double move_swap;                        /* SwapFreq */
  /* NPT (ensemble 2): Pressure (bar) and VolFreq. */
  double pressure, move_vol;
} orc_system_desc;

void *orc_create(const orc_system_desc *d);
void orc_destroy(void *h);
/* Plants a known mistake in the oracle for every box (see orc::g_plant);
 * 0 removes it.  Used by tests that show a statistical check fails on it. */
void orc_plant(int which);
/* Sets the displacement and rotation scales of every molecule kind and the
 * MultiParticle step sizes (test set-up: a run whose AdjSteps exceeds its
 * length then samples with fixed steps). */
void orc_set_scales(void *h, const double *disp, const double *rot, double mp_t_max,
                    double mp_r_max);
int orc_natoms(void *h);
int orc_nmols(void *h);
int orc_nkvec(void *h);
void orc_cutoffs(void *h, double out[6]); /* alpha, rCut(box), rCutLow, recip_rcut, volume, beta */
void orc_system_total(void *h, double out[7]);
void orc_energies(void *h, double out[7]);
void orc_box_inter(void *h, double *lj, double *real);
void orc_box_inter_brute(void *h, double *lj, double *real);
int orc_mol_inter(void *h, int mol, const double *nx, const double *ny,
                  const double *nz, double *lj, double *real);
double orc_mol_recip(void *h, int mol, const double *nx, const double *ny,
                     const double *nz, double *sumRnew, double *sumInew);
double orc_swap_recip(void *h, int kind, int insert, const double *x,
                      const double *y, const double *z);
void orc_particle_inter(void *h, int atom_type, double charge, int ntrials,
                        const double *tx, const double *ty, const double *tz,
                        double *en, double *real, int32_t *overlap);
void orc_kvectors(void *h, double *kx, double *ky, double *kz, double *prefact);
void orc_recip_sums(void *h, double *sumR, double *sumI);
void orc_self_corr_lrc(void *h, double out[3]);
double orc_tail_change(void *h, int kind, int add);
void orc_get_coords(void *h, double *x, double *y, double *z);
void orc_get_com(void *h, double *x, double *y, double *z);
void orc_get_scales(void *h, double *disp, double *rot);
int orc_cell_csr(void *h, int32_t *cellVector, int32_t *cellStart,
                 int32_t *map, int32_t edge[3]);
int orc_cells_exhaustive(void *h);
/* mode 0: MT19937 draws (the reference order); mode 1: per-box Philox. */
int64_t orc_run(void *h, int64_t start_step, int64_t nsteps, int mode,
                int32_t *kind, int32_t *mol, int32_t *acc, int32_t *ovl,
                double *dlj, double *dreal, double *drecip);
void orc_box_force(void *h, double *lj, double *real, double *afx,
                   double *afy, double *afz, double *mfx, double *mfy,
                   double *mfz);
void orc_box_force_recip(void *h, double *afx, double *afy, double *afz,
                         double *mfx, double *mfy, double *mfz);
void orc_torque(void *h, double *tx, double *ty, double *tz);
/* Box of every molecule (0 box, 1 reservoir); returns molecules in box 0. */
int orc_mol_boxes(void *h, int32_t *box);
/* MultiParticle step sizes: {t_max, r_max}. */
void orc_mp_settings(void *h, double out[2]);

/* Many boxes, Philox streams, one box per OpenMP thread.  Returns elapsed
 * seconds of the move loop only; moves[b] moves are made in box b. */
double orc_batch_run(int nboxes, void *const *handles, const int64_t *moves,
                     int threads);

/* Generators on their own. */
void orc_mt_raw(uint32_t seed, int n, uint32_t *out);
void orc_mt_rand_int(uint32_t seed, int n, uint32_t bound, uint32_t *out);
void orc_philox(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
double orc_u01(uint64_t v);
double orc_uneg11(uint64_t v);
/* Random123Wrapper draws: out = {rand, sym.x, sym.y, sym.z, sphere.x,
 * sphere.y, sphere.z, symr(2.5)} for GetSymRandomCoords(counter, bound). */
void orc_r123(uint64_t seed, uint64_t step, uint32_t counter, double bound, double out[8]);
void orc_rotation(double theta, double ax, double ay, double az, double m[9]);
void orc_uniform_rotation(double u1, double u2, double u3, double m[9]);
/* CalculateEnergy::VirialCalc (CalculateEnergy.cpp:411-579, with
 * VirialCorrection and Ewald::VirialReciprocal) and the pressure of
 * OutputVars.cpp:147-172: out = {inter, tail, real, recip, total,
 * pressure in bar, total tensor xx, yy, zz}. */
void orc_virial(void *h, double out[9]);
/* Replica exchange over systems by temperature slot (ParallelTempering-
 * Utilities.cpp:108-240, neighbour branch): see oracle_capi.cpp. */
int orc_exchange(void *const *hs, int32_t *ladder, int n, int64_t attempt, uint32_t seed,
                 int64_t drawsBefore, int32_t *accepted);
/* Pair terms on their own: [lj, coulomb, ljvir, coulvir, lrc]. */
void orc_pair_terms(const orc_system_desc *d, int t1, int t2, double distSq,
                    double qiqj, double out[5]);

#ifdef __cplusplus
}
#endif
#endif

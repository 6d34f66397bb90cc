/* gomc_hip.h -- C ABI of the MI355X (gfx950) energy hot path for GOMC.
 *
 * Every entry point takes plain pointers and sizes, returns 0 on success and
 * a negative code on failure (gomc_last_error() gives the text), and never
 * exits the process.  The reference's GPU seam is a set of host free
 * functions taking VariablesCUDA* (the src/GPU headers); each entry point below
 * names the reference function it replaces.  One gomc_box holds one
 * simulation box resident in HBM (coordinates, force-field tables, cell
 * list, Ewald k-vectors and structure factors) across calls, so no call
 * re-uploads the box.
 */
#ifndef GOMC_HIP_H
#define GOMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GOMC_OK 0
#define GOMC_ERR_ARG -1     /* bad argument / size */
#define GOMC_ERR_HIP -2     /* HIP runtime error */
#define GOMC_ERR_STATE -3   /* call out of order (e.g. no k-vectors yet) */
#define GOMC_ERR_NODEV -4   /* no usable gfx950 device */

/* One box: force field, molecule kinds, molecules, coordinates, moves.
 * Field meaning follows the reference's in.conf keywords:
 *   box        -- CellBasisVector1..3 lengths (orthogonal box), A
 *   temperature-- Temperature, K
 *   rcut, rcut_low, rcut_coulomb, tolerance -- Rcut, RcutLow,
 *                 RcutCoulomb, Tolerance (src/ConfigSetup.cpp:671-701)
 *   ewald, electrostatic, lrc -- Ewald, ElectroStatic, LRC
 *   sigma/epsilon/n_rep -- Mie parameters per atom type (A, K, exponent)
 *   kind_* -- molecule kinds: atom types and charges (e)
 *   mol_kind -- kind of each molecule; atoms of molecule m are contiguous
 *   x,y,z  -- wrapped coordinates of all atoms, A
 *   seed   -- Random_Seed; move_* -- DisFreq, RotFreq, MultiParticleFreq
 *   adj_steps -- AdjSteps
 *   ensemble, cbmc_*, mol_box, kind_fugacity, move_swap -- GCMC settings
 *                 (Fugacity, CBMC_First/Nth/Ang, SwapFreq)
 *   pressure, move_vol -- NPT settings (Pressure in bar, VolFreq)
 * The same layout is used by the CPU oracle's orc_system_desc. */
typedef struct gomc_system_desc {
  double box[3];
  double temperature;
  double rcut, rcut_low, rcut_coulomb, tolerance;
  int32_t ewald, electrostatic, lrc, threads;
  int32_t n_atom_types, n_kinds;
  const double *sigma, *epsilon, *n_rep;
  const int32_t *kind_natoms;
  const int32_t *kind_atom_type;
  const double *kind_charge;
  int32_t n_mols, _pad0;
  const int32_t *mol_kind;
  const double *x, *y, *z;
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
} gomc_system_desc;

typedef struct gomc_box gomc_box;

/* ---- lifetime -------------------------------------------------------- */
/* Replaces InitGPUForceField + InitCoordinatesCUDA + InitEwaldVariablesCUDA
 * (src/GPU/ConstantDefinitionsCUDAKernel.cuh:22-46), called from
 * FFParticle::Init (src/FFParticle.cpp:54-93) and
 * CalculateEnergy::Init (src/CalculateEnergy.cpp:60-81). */
int gomc_box_create(const gomc_system_desc *desc, int device, gomc_box **out);
/* Replaces DestroyCUDAVars (src/GPU/ConstantDefinitionsCUDAKernel.cuh:44). */
int gomc_box_destroy(gomc_box *b);
const char *gomc_last_error(const gomc_box *b);
/* Sizes: [n_atoms, n_mols, n_kvec, n_cells, cells_per_edge_x, y, z]. */
int gomc_box_info(const gomc_box *b, int64_t out[7]);
/* Current edge lengths and volume, A and A^3 (BoxDimensions::axis, volume;
 * they change with NPT volume moves): out = {Lx, Ly, Lz, V}. */
int gomc_box_dims(const gomc_box *b, double out[4]);

/* ---- coordinates ----------------------------------------------------- */
/* Replaces the per-call coordinate uploads inside every Call*GPU
 * (e.g. src/GPU/CalculateEnergyCUDAKernel.cu:70-87). */
int gomc_set_coords(gomc_box *b, const double *x, const double *y,
                    const double *z);
int gomc_get_coords(const gomc_box *b, double *x, double *y, double *z);
/* Writes the coordinates of one molecule (XYZArray::CopyRange at accept,
 * src/moves/Translate.h:114). */
int gomc_commit_mol(gomc_box *b, int mol, const double *x, const double *y,
                    const double *z);

/* ---- cell list (src/CellList.cpp:220-285) ---------------------------- */
/* GridAll + GetCellListNeighbor on the device: counting sort of atoms into
 * cells, each cell's atoms in ascending index order (the CSR form the
 * reference builds on the host every BoxInter call).  Outputs may be NULL;
 * cell_vector[n_atoms], cell_start[n_cells+1], atom_cell[n_atoms]. */
int gomc_build_cells(gomc_box *b, int32_t *cell_vector, int32_t *cell_start,
                     int32_t *atom_cell);
/* The trial cell table that gomc_trial and gomc_particle_inter walk in a
 * box of more than 27 cells (MoleculeInter and ParticleInter over
 * CellList::EnumerateLocal, src/CalculateEnergy.cpp:595,745): per cell up
 * to *cap slots of live atoms, kept up to date by the accepted
 * displacements (CellList::RemoveMol/AddMol, src/CellList.h:28-29).  Writes
 * *cap; count[n_cells] and slots[n_cells * cap] (each NULL or sized for the
 * cap from a call with both NULL) receive the device's copy.  A read-back
 * for checking; no counterpart in the reference.  GOMC_ERR_STATE in a box
 * of 27 cells, which the trials scan whole. */
int gomc_trial_cells(gomc_box *b, int32_t *cap, int32_t *count, int32_t *slots);

/* ---- real-space pair energies ---------------------------------------- */
/* CalculateEnergy::BoxInter (src/CalculateEnergy.cpp:157-266); replaces
 * CallBoxInterGPU (src/GPU/CalculateEnergyCUDAKernel.cuh:17-26).  Rebuilds
 * the cell list first. */
int gomc_box_inter(gomc_box *b, double *lj, double *real);
/* CalculateEnergy::MoleculeInter (src/CalculateEnergy.cpp:581-686): the
 * change of LJ and real-space Coulomb energy when molecule `mol` moves to
 * (nx,ny,nz); *overlap = 1 if a new pair is closer than RcutLow.  No GPU
 * counterpart exists in the reference (host OpenMP only). */
int gomc_mol_inter(gomc_box *b, int mol, const double *nx, const double *ny,
                   const double *nz, double *dlj, double *dreal,
                   int32_t *overlap);
/* CalculateEnergy::ParticleInter (src/CalculateEnergy.cpp:727-785): LJ and
 * real energies of `ntrials` trial positions of one atom (type, charge)
 * against every atom of the box; en/real/overlap are written (not added).
 * Atoms of molecule `skip_mol` (>=0) are left out. */
int gomc_particle_inter(gomc_box *b, int atom_type, double charge,
                        int skip_mol, int ntrials, const double *tx,
                        const double *ty, const double *tz, double *en,
                        double *real, int32_t *overlap);

/* Molecular virial and pressure of the current configuration:
 * CalculateEnergy::VirialCalc (CalculateEnergy.cpp:411-579) with
 * VirialCorrection (1317-1336) and Ewald::VirialReciprocal
 * (Ewald.cpp:1168-1305), the pressure as OutputVars.cpp:147-172 forms it;
 * replaces CallBoxInterForceGPU and CallVirialReciprocalGPU
 * (src/GPU/CalculateForceCUDAKernel.cuh:17-50).  out = {inter (LJ), tail,
 * real, recip, total, pressure in bar, total tensor xx, yy, zz} (virials
 * in K). */
int gomc_box_virial(gomc_box *b, double out[9]);

/* Checkpoint of the box's Monte Carlo state -- what Checkpoint::serialize
 * keeps (src/Checkpoint.h:118-167, written by CheckpointOutput.cpp:69-90,
 * read back by CheckpointSetup): the caller's step number, the MT19937
 * state, MoveSettings, MultiParticle settings, MoleculeLookup, coordinates
 * -- plus the geometry, centres, running energies and structure factors,
 * so that a restored box continues move for move.  The bytes are this
 * library's own versioned format, not a cereal archive.  Returns the size
 * in bytes (writing them only when cap >= size; buf may be NULL to ask for
 * the size) or a negative error code. */
int64_t gomc_box_checkpoint(gomc_box *b, int64_t step, void *buf, int64_t cap);
/* Restores a checkpoint taken from a box of the same system and returns
 * its step in *step (may be NULL).  A checkpoint of another system (other
 * atom, molecule or kind counts, ensemble or molecule kinds) or a truncated
 * one is refused with GOMC_ERR_ARG before the box is touched. */
int gomc_box_restore(gomc_box *b, const void *buf, int64_t size, int64_t *step);

/* ---- Ewald reciprocal space (src/Ewald.cpp) -------------------------- */
/* RecipInitOrth + BoxReciprocalSetup + SetRecipRef (Ewald.cpp:847-903,
 * 193-274, 1021-1053); replaces CallBoxReciprocalSetupGPU
 * (src/GPU/CalculateEwaldCUDAKernel.cuh:16-24).  *erecip = BoxReciprocal. */
int gomc_recip_init(gomc_box *b, double *erecip);
/* BoxReciprocalSums + BoxReciprocal (Ewald.cpp:281-406); replaces
 * CallBoxReciprocalSumsGPU.  Fills the "new" sums from the current
 * coordinates; gomc_recip_accept makes them the reference sums. */
int gomc_recip_sums(gomc_box *b, double *erecip);
/* gomc_recip_sums without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Same kernels and results; the scratch
 * (per-chunk sums, block partials, counter) lives in one 32.1 MiB
 * file-scope device array per GPU, held until the process ends, so calls
 * that overlap on one GPU race.  Refuses boxes of more than 2^20 k-vectors
 * (GOMC_ERR_ARG).  The box must have been set up by gomc_recip_init. */
int gomc_recip_sums_noalloc(gomc_box *b, double *erecip);
/* Ewald::MolReciprocal (Ewald.cpp:409-473); replaces CallMolReciprocalGPU
 * (src/GPU/CalculateEwaldCUDAKernel.cuh:31-36).  *enew = reciprocal energy
 * with molecule `mol` at (nx,ny,nz); the caller's delta is
 * enew - (current reciprocal energy).  No structure factors leave HBM. */
int gomc_mol_recip(gomc_box *b, int mol, const double *nx, const double *ny,
                   const double *nz, double *enew);
/* SwapDestRecip / SwapSourceRecip (Ewald.cpp:478-531, 657-710); replaces
 * CallSwapReciprocalGPU.  insert=1 adds, 0 removes the molecule of `kind`
 * at (x,y,z); *enew as above.  `mol` is the molecule's index (-1 if none):
 * with the molecule cache on, an inserted molecule's terms become its cache
 * entry when the move is accepted (EwaldCached::SwapDestRecip,
 * src/EwaldCached.cpp:300-352). */
int gomc_swap_recip(gomc_box *b, int kind, int mol, int insert, const double *x,
                    const double *y, const double *z, double *enew);
/* EwaldCached mode (src/EwaldCached.cpp, chosen by CachedFourier,
 * src/System.cpp:132-148): enable=1 keeps each molecule's structure-factor
 * terms [n_mols][n_kvec] in HBM, so a translate/rotate trial computes only
 * the new positions' terms; the cache follows accepted moves.  enable=0
 * frees it.  Energies are the same as without the cache. */
int gomc_recip_cache(gomc_box *b, int enable);
/* Trial precision (no counterpart in the reference): 1 evaluates the pair
 * terms and e^{ik.r} phases of MoleculeInter / MolReciprocal trials, CBMC
 * trial energies (ParticleInter) and SwapReciprocal in FP32 from FP64
 * distances and arguments, with FP64 sums, S(k), running energies and
 * Metropolis; 0 (default) is FP64 throughout.  Whole-box totals stay FP64. */
int gomc_set_precision(gomc_box *b, int precision);
/* UpdateRecip (Ewald.cpp:1420-1434): the new sums become the reference
 * sums (a pointer swap); replaces UpdateRecipCUDA. */
int gomc_recip_accept(gomc_box *b);
/* CopyRecip (Ewald.cpp:1441-1460); replaces CopyRefToNewCUDA. */
int gomc_recip_copy_ref_to_new(gomc_box *b);
/* Reference k-vectors and sums for inspection: arrays of n_kvec; any NULL. */
int gomc_recip_get(const gomc_box *b, double *kx, double *ky, double *kz,
                   double *prefact, double *sum_r, double *sum_i);

/* ---- fused single-molecule trial ------------------------------------- */
/* MoleculeInter + MolReciprocal of one displacement/rotation trial in one
 * launch (the pair blocks and the k-vector blocks share the grid), one
 * host read-back of out[4] = {dLJ, dReal, recip_new, overlap}. */
int gomc_mol_trial(gomc_box *b, int mol, const double *nx, const double *ny,
                   const double *nz, double out[4]);

/* ---- whole-box totals ------------------------------------------------ */
/* CalculateEnergy::SystemTotal for one box (src/CalculateEnergy.cpp:83-138)
 * without intramolecular bonded terms (rigid molecules):
 * out = {inter, tailCorrection, real, recip, self, correction, total}.
 * Recomputes the reciprocal sums from the coordinates. */
int gomc_system_total(gomc_box *b, double out[7]);

/* ---- forces and the MultiParticle move (src/moves/MultiParticle.h) --- */
/* CalculateEnergy::BoxForce (src/CalculateEnergy.cpp:268-406); replaces
 * CallBoxForceGPU (src/GPU/CalculateForceCUDAKernel.cuh:17-30).  Pair
 * energies and the atom / molecule forces of the current coordinates into
 * the device's reference force set; the six host arrays (n_atoms, n_mols
 * doubles) receive a copy and may each be NULL. */
int gomc_box_force(gomc_box *b, double *lj, double *real, double *afx,
                   double *afy, double *afz, double *mfx, double *mfy,
                   double *mfz);
/* Ewald::BoxForceReciprocal (src/Ewald.cpp:1496-1596); replaces
 * CallBoxForceReciprocalGPU (src/GPU/CalculateEwaldCUDAKernel.cuh:50-58).
 * Uses the current "new" sums: call gomc_recip_copy_ref_to_new first, as
 * MultiParticle::Prep does (MultiParticle.h:211-217). */
int gomc_box_force_recip(gomc_box *b, double *afx, double *afy, double *afz,
                         double *mfx, double *mfy, double *mfz);
/* CalculateEnergy::CalculateTorque (src/CalculateEnergy.cpp:1365-1406) of
 * the reference force set about the molecule centres; arrays of n_mols. */
int gomc_torque(gomc_box *b, double *tx, double *ty, double *tz);
/* MultiParticle Transform + CalcEn + GetCoeff (MultiParticle.h:332-511);
 * replaces CallTranslateParticlesGPU / CallRotateParticlesGPU
 * (src/GPU/TransformParticlesCUDAKernel.cuh:20-35) and the host force and
 * coefficient work.  move_type 0 displaces, 1 rotates every molecule with
 * the force-biased transform from the reference force set, Random123 key
 * {step, seed}; `max` is t_max or r_max.  out = {inter, real, recip of the
 * trial, w_ratio, fail (1 when the transform check of MultiParticle.h:366-406
 * fails)}.  Until gomc_mp_accept the trial coordinates are the box's. */
int gomc_mp_trial(gomc_box *b, int move_type, uint64_t step, double max,
                  double out[5]);
/* MultiParticle::Accept's state change (MultiParticle.h:522-538): accept
 * keeps the trial coordinates, energies (from `out` of gomc_mp_trial),
 * forces and sums; reject restores the reference state. */
int gomc_mp_accept(gomc_box *b, int accept, const double out[5]);
/* The driver's MultiParticle step sizes {t_max, r_max}. */
int gomc_mp_settings(const gomc_box *b, double out[2]);

/* ---- plain single-box MC driver (src/System.cpp:262-302) ------------- */
/* Runs `nsteps` GOMC moves (Translate / Rotate / MultiParticle, and in
 * GCMC MoleculeTransfer with CBMC growth, src/moves/MoleculeTransfer.h; MT19937
 * draws in the reference's order and Random123 draws for MultiParticle,
 * Metropolis acceptance, MoveSettings adaptation) with every
 * single-molecule energy evaluated by gomc_mol_trial.  Per-move records may
 * be NULL; a MultiParticle record has kind 2, mol = its move type and ovl =
 * its transform failure; a MoleculeTransfer record has kind 10, the
 * molecule, and energy changes of box 0.  Returns the number of accepted moves (>= 0) or
 * an error code. */
int64_t gomc_run_moves(gomc_box *b, int64_t start_step, int64_t nsteps,
                       int32_t *kind, int32_t *mol, int32_t *acc,
                       int32_t *ovl, double *dlj, double *dreal,
                       double *drecip);
/* The same move loop run on the device: one launch runs `nsteps` moves
 * (Translate, Rotate and, in GCMC, MoleculeTransfer with its CBMC growth,
 * src/System.cpp:262-302, src/moves/MoleculeTransfer.h:85-246,
 * src/cbmc/DCSingle.cpp:23-100, src/cbmc/DCFreeHedron.cpp:48-240) with the
 * box's MT19937 stream continued on the device, so the moves, their order
 * and their outcomes are those of gomc_run_moves; records and return value
 * as there.  Boxes with MultiParticle moves, the molecule cache, more than
 * 4 kinds, molecules of more than 3 atoms or more than 32 CBMC trials per
 * stage are refused (GOMC_ERR_STATE / GOMC_ERR_ARG).  No counterpart in the
 * reference, which runs the move loop on the host. */
int64_t gomc_run_moves_device(gomc_box *b, int64_t start_step, int64_t nsteps,
                              int32_t *kind, int32_t *mol, int32_t *acc,
                              int32_t *ovl, double *dlj, double *dreal,
                              double *drecip);
/* Workgroups per box in the device move loop: 1..8, or 0 for the default
 * (by k-vector count: 4 from 16384, 2 from 8192, else 1; 1 without Ewald).
 * Every member runs the box's whole move on the same data and takes the
 * same decisions; the k-vectors are split among the members, whose partial
 * reciprocal energies are summed in member order once per move, so a
 * box's results depend on its team size and on nothing else in a launch.
 * No counterpart in the reference. */
int gomc_set_device_team(gomc_box *b, int team);
/* One launch of the device move loop over `n` boxes of one device and one
 * precision (a team of workgroups per box, gomc_set_device_team; boxes
 * beyond what the device holds at once go to a further launch).  Box i
 * runs from start_step[i]:
 * nsteps[i] moves, or with cycles (may be NULL) up to cycles[i] cycles of
 * N_box moves each (N_box counted at the cycle's start, at least 1, never
 * more than nsteps[i] moves in all), adding its kind-0 and kind-1 molecule
 * counts to count_sum[2i], count_sum[2i+1] after each cycle
 * (count_samples[i] cycles); status[i] is box i's return code.
 * moves_done, accepted, count_sum, count_samples and status may be NULL.
 * Returns 0 or the first failing box's error code (gomc_last_error of
 * that box).  No counterpart in the reference. */
int gomc_run_moves_device_multi(gomc_box *const *boxes, int n,
                                const int64_t *start_step,
                                const int64_t *nsteps, const int32_t *cycles,
                                int64_t *moves_done, int64_t *accepted,
                                double *count_sum, int64_t *count_samples,
                                int32_t *status);
/* GCMC with the moves driven by the caller (MoleculeTransfer::Accept,
 * src/moves/MoleculeTransfer.h:168-246, through MoleculeLookup::ShiftMolBox):
 * puts molecule `mol` in box 0 (it interacts) or the reservoir (1), with its
 * coordinates as last set.  The box's own move driver's lookup is rebuilt in
 * molecule order.  GOMC_ERR_STATE for 1 outside GCMC. */
int gomc_set_mol_box(gomc_box *b, int mol, int box);
/* GCMC: box of every molecule (0 simulation box, 1 reservoir) into `box`
 * (n_mols); returns the number of molecules in box 0 or an error code.
 * MoleculeLookup::NumInBox / IsMoleculeInBox (src/MoleculeLookup.h:98-130). */
int gomc_mol_boxes(const gomc_box *b, int32_t *box);
/* Running energies of the driver (same order as gomc_system_total). */
int gomc_energies(const gomc_box *b, double out[7]);
/* Current displacement / rotation scales per molecule kind. */
int gomc_get_scales(const gomc_box *b, double *disp, double *rot);

/* ---- batched persistent MC (many independent boxes, one launch) ------ */
typedef struct gomc_batch gomc_batch;
/* Boxes hold one molecule kind each (rigid, up to 4 atoms).  Each box runs
 * its own Philox4x64-10 stream (key {seed, tag}, counter {move, block}).
 * precision: 0 = FP64 everywhere; 1 = FP32 pair terms and phases with FP64
 * accumulation, running totals and Metropolis exponent.
 * A box of 1800 molecules or more runs on a team of 2, 4 or 8 workgroups
 * (a function of its molecule count), unless the batch holds more large
 * boxes than the device fits, when the smallest teams are halved.  A box's
 * results depend only on the box and its team size. */
int gomc_batch_create(const gomc_system_desc *descs, int nboxes, int device,
                      int precision, gomc_batch **out);
/* Launch-shape choices of a batch (no counterpart in the reference; the
 * autotuner gomc_amd/autotune.py picks them by timing MC cycles):
 * team_min[i]: molecule count from which a box runs on 2, 4, 8 workgroups
 * (<= 0: never); pair12_max_k: a single-workgroup box with at most this many
 * k-vectors runs its pair phase on 12 waves beside a 4-wave reciprocal
 * phase, other boxes on 8 + 8.  Results depend on a box's team size. */
typedef struct gomc_batch_tuning {
  int32_t team_min[3];
  int32_t pair12_max_k;
} gomc_batch_tuning;
int gomc_batch_default_tuning(gomc_batch_tuning *out);
int gomc_batch_create_tuned(const gomc_system_desc *descs, int nboxes, int device,
                            int precision, const gomc_batch_tuning *tuning,
                            gomc_batch **out);
/* Wall time of each box's moves in the last launch, microseconds
 * (device clock; a team box counts its first member). */
int gomc_batch_box_times(const gomc_batch *bt, double *us);
int gomc_batch_destroy(gomc_batch *bt);
const char *gomc_batch_last_error(const gomc_batch *bt);
/* Advance box i by moves[i] moves (NULL: by its molecule count, one MC
 * cycle) in one persistent launch. */
int gomc_batch_run(gomc_batch *bt, const int64_t *moves);
/* `ncycles` MC cycles of every box (N_i moves each, one launch per cycle).
 * With use_graph the cycle is captured once as a HIP graph and replayed;
 * gomc_batch_last_kernel_ms then covers all cycles. */
int gomc_batch_run_cycles(gomc_batch *bt, int ncycles, int use_graph);
/* Replica exchange between boxes of a batch (parallel tempering on one
 * device): ParallelTemperingUtilities::evaluateExchangeCriteria's
 * neighbour exchange (src/ParallelTemperingUtilities.cpp:108-240) with the
 * MPI exchange of coordinates (406-448) replaced by an exchange of the
 * temperature-bound state (beta, move scales and counters, Philox stream)
 * between the two boxes.  ladder[0..n) lists the boxes running temperature
 * slots 0..n-1 (same molecules and box edges); pairs (s-1, s) with
 * s % 2 == attempt % 2 are tried with one draw each from the batch's
 * exchange stream (MT19937 seeded with `seed` at attempt 0).  accepted[s]
 * (may be NULL) is 1 when slots s-1 and s exchanged; ladder is updated so
 * that ladder[s] is again the box at slot s's temperature. */
int gomc_batch_exchange(gomc_batch *bt, int32_t *ladder, int n, int64_t attempt,
                        uint32_t seed, int32_t *accepted);
/* Per-box running energies [nboxes][7], accepted counts [nboxes], moves
 * made [nboxes]; any NULL. */
int gomc_batch_stats(const gomc_batch *bt, double *energies,
                     int64_t *accepted, int64_t *moves_done);
int gomc_batch_get_coords(const gomc_batch *bt, int box, double *x, double *y,
                          double *z);
/* How box `box` is laid out: out = {team size, k-vectors per segment,
 * segment rounds, centre cull on}.  Device-layout query with no
 * counterpart in the reference. */
int gomc_batch_box_layout(const gomc_batch *bt, int box, int32_t out[4]);
/* Per-move trace of box `box` over the last gomc_batch_run (first `cap`
 * moves): accepted flag and the three energy deltas. */
int gomc_batch_set_trace(gomc_batch *bt, int box, int64_t cap);
int gomc_batch_get_trace(const gomc_batch *bt, int32_t *acc, double *dlj,
                         double *dreal, double *drecip, int64_t *n);
/* Kernel-only time of the last gomc_batch_run in ms (HIP events on the
 * launch stream). */
int gomc_batch_last_kernel_ms(const gomc_batch *bt, double *ms);

/* ---- serving: continuous batching of state-point jobs ---------------- */
/* No counterpart in the reference (one simulation per process there).  A
 * request loop on its own thread keeps one resident batch of `slots`
 * workgroup slots; submitted jobs (NVT boxes as for gomc_batch_create) are
 * admitted in arrival order when a slot is free and their coordinates,
 * centres, k-vectors and structure factors fit the `budget_bytes` device
 * arenas (the NVT lane's arenas take the budget up to 64 MiB per slot);
 * every launch advances each resident job by the moves that cost about
 * `tick_ms`; finished jobs release their slot and memory.  A job's result
 * does not depend on the other jobs in the batch. */
typedef struct gomc_server gomc_server;
int gomc_server_create(int device, int precision, int slots, int64_t budget_bytes,
                       double tick_ms, gomc_server **out);
int gomc_server_destroy(gomc_server *sv);
const char *gomc_server_last_error(const gomc_server *sv);
/* Queues a job for `moves` MC moves; *job_id identifies it.  An NVT
 * descriptor's initial state is built on the caller's thread and the job
 * joins the resident batch.  A GCMC descriptor (ensemble 1) goes to the
 * GCMC lane: builder threads make each such job a device box in arrival
 * order while the budget holds its bytes, and every built job is advanced,
 * by whole MC cycles, in each launch of the device move loop
 * (gomc_run_moves_device_multi: a team of workgroups per job by its
 * k-vector count, the job's own MT19937 stream) at the server's
 * precision. */
int gomc_server_submit(gomc_server *sv, const gomc_system_desc *d,
                       int64_t moves, int64_t *job_id);
/* status 0 queued, 1 running, 2 done, 3 failed; times = {submitted,
 * admitted, done} in seconds since the server started; energies as
 * gomc_energies; any output may be NULL. */
int gomc_server_poll(gomc_server *sv, int64_t job_id, int32_t *status,
                     double times[3], double energies[7], int64_t *accepted,
                     int64_t *moves_done);
/* Mean number of molecules of kinds 0 and 1 in the simulation box over a
 * GCMC job's MC cycles (sampled after each cycle); for a batched NVT job,
 * {N, 0}.  Serving query with no counterpart in the reference (its
 * observable is the block average of BoxMolecules, src/BlockOutput.cpp). */
int gomc_server_poll_counts(gomc_server *sv, int64_t job_id, double mean_n[2]);
/* Waits until every submitted job has finished (GOMC_ERR_STATE on timeout). */
int gomc_server_drain(gomc_server *sv, double timeout_s);
/* {moves done, kernel seconds, launches, peak bytes in use, budget bytes,
 *  seconds since start, jobs queued, jobs outstanding}. */
int gomc_server_stats(gomc_server *sv, double out[8]);

/* ---- synthetic inputs (host only) ------------------------------------ */
/* Random sequential insertion of rigid molecules into a cubic box of edge
 * `box`: uniform centre, uniform orientation (splitmix64 stream from
 * `seed`), rejected if the first atoms of two molecules are closer than
 * min_heavy or any other inter-molecular pair closer than min_any.
 * kind_geom holds body-frame xyz triples of every kind's atoms in order.
 * No counterpart in the reference: it stands in for the PDB coordinates
 * the reference reads at start-up.  GOMC_ERR_STATE if a molecule cannot be
 * placed within max_tries. */
int gomc_place_molecules(int n_kinds, const int32_t *kind_natoms,
                         const double *kind_geom, int n_mols,
                         const int32_t *mol_kind, double box, uint64_t seed,
                         double min_heavy, double min_any, int max_tries,
                         double *x, double *y, double *z);

#ifdef __cplusplus
}
#endif
#endif /* GOMC_HIP_H */

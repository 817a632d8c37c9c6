/*
 * raspa3_amd — C ABI of the MI355X (gfx950) framework–adsorbate energy path.
 *
 * Plain C: opaque handles, plain pointers and sizes, status codes.  No exception crosses this
 * boundary.  The caller owns every host array it passes; the library owns device memory behind
 * handles.  One context per host thread; calls on a context are issued in order on the context's
 * HIP stream and the entry points below return after the device work they start has finished
 * unless stated otherwise.
 *
 * Units follow the reference: lengths in Angstrom, energies in the reference's internal unit
 * (1 internal unit = 1.2027242847 K, src/raspakit/units.cpp:98), charges in e.
 *
 * Each entry point names the reference interface it replaces (paths relative to the RASPA3 tree).
 */
#ifndef RASPA3_AMD_RG_H
#define RASPA3_AMD_RG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rg_status
{
  RG_OK = 0,
  RG_ERR_INVALID = 1,     /* bad argument (null pointer, size out of range, shape mismatch) */
  RG_ERR_UNSUPPORTED = 2, /* a configuration this build does not cover (message says which) */
  RG_ERR_DEVICE = 3,      /* a HIP call failed; rg_last_error holds the HIP error string */
  RG_ERR_NOMEM = 4        /* device allocation failed */
} rg_status;

typedef struct rg_ctx rg_ctx;
typedef struct rg_forcefield rg_forcefield;
typedef struct rg_framework rg_framework;
typedef struct rg_grid rg_grid;

/* ---------------------------------------------------------------------------------------------
 * Context
 * ------------------------------------------------------------------------------------------- */

/* Creates a context on HIP device 'device' with its own stream. */
rg_status rg_ctx_create(int device, rg_ctx** out);
/* The *_destroy functions wait for the device and do not use the context, so objects may be released after it. */
rg_status rg_ctx_destroy(rg_ctx* ctx);
/* Decisions of the online autotuner (launch shapes timed on the first calls of each problem size, see
 * raspa3_amd/csrc/autotune.hpp), one line per tuned key; writes at most len - 1 characters and a terminating NUL into
 * buf (if not NULL) and returns the full length.  No reference counterpart. */
size_t rg_autotune_summary(char* buf, size_t len);
/* Last error message of this context (empty string if none).  Valid until the next call. */
const char* rg_last_error(const rg_ctx* ctx);
/* Blocks until all work issued on the context's stream has finished. */
rg_status rg_sync(rg_ctx* ctx);
/* The context's hipStream_t, as void*, for callers that time or order work around it. */
void* rg_stream(rg_ctx* ctx);
/* Library version string. */
const char* rg_version(void);

/* ---------------------------------------------------------------------------------------------
 * Simulation box and force field set-up (host-side, no device work)
 * ------------------------------------------------------------------------------------------- */

/* SimulationBox(a, b, c, alpha, beta, gamma) — src/raspakit/simulationbox.cpp:40-60 — angles in degrees.
 * lengths_only != 0 follows the SimulationBox(a, b, c) constructor (src/raspakit/simulationbox.cpp:15-37).
 * cell9 = lattice vectors a, b, c as consecutive columns; inv9 likewise for the inverse.
 * widths3 = perpendicular widths (src/raspakit/simulationbox.cpp:137-148). */
rg_status rg_box(double a, double b, double c, double alpha_deg, double beta_deg, double gamma_deg, int lengths_only,
                 double* cell9, double* inv9, int* is_rectangular, double* volume, double* widths3);

/* ForceField::initializeEwaldParameters with automaticEwald (src/raspakit/forcefield/forcefield.cpp:1415-1438):
 * alpha and kmax3 from EwaldPrecision, the Coulomb cutoff and the perpendicular widths. */
rg_status rg_ewald_parameters(double precision, double rc_coulomb, const double* widths3, double* alpha, int* kmax3);

/* Grid dimensions from a spacing: n = (size_t)(w/h + 0.5) per axis,
 * src/raspakit/system_interpolation_grids.cpp:192-198,222-226. */
rg_status rg_grid_dims(const double* cell9, double spacing, int* n3);

/* ForceField(pseudoAtoms, selfInteractions, LorentzBerthelot, rcFramework, rcMolecule, rcCoulomb, shifted,
 * tailCorrections = false) — src/raspakit/forcefield/forcefield.cpp:104-131, mixing :924-1014, shift
 * :1038-1051.  eps_K[n] in Kelvin, sigma[n] in Angstrom, is_framework[n] selects the framework cutoff for
 * pairs that involve a framework pseudo-atom.  alpha is EwaldAlpha, use_charge the ForceField flag,
 * overlap the energyOverlapCriteria (internal units; the reference default is 1e6). */
rg_status rg_forcefield_create(rg_ctx* ctx, int n_types, const double* eps_K, const double* sigma,
                               const int* is_framework, int shifted, double rc_framework_vdw,
                               double rc_molecule_vdw, double rc_coulomb, double alpha, int use_charge,
                               double overlap, rg_forcefield** out);
rg_status rg_forcefield_destroy(rg_forcefield* ff);
/* The mixed pair table, row-major n*n: eps (internal units), sigma, shift (internal units), lj (1 = LJ, 0 = None). */
rg_status rg_forcefield_pairs(const rg_forcefield* ff, double* eps, double* sigma, double* shift, int* lj);

/* ---------------------------------------------------------------------------------------------
 * Framework: the rigid unit cell (Framework::unitCellAtoms, src/raspakit/framework.ixx) copied to the
 * device once.  pos is n*3 (x, y, z per atom), type the pseudo-atom index, charge in e.
 * ------------------------------------------------------------------------------------------- */
rg_status rg_framework_create(rg_ctx* ctx, const double* cell9, int is_rectangular, int n_atoms, const double* pos,
                              const int* type, const double* charge, rg_framework** out);
rg_status rg_framework_destroy(rg_framework* fw);

/* ---------------------------------------------------------------------------------------------
 * Interpolation grids: InterpolationEnergyGrid (src/raspakit/interpolation_energy_grid.ixx:19-63),
 * tricubic or triquintic scheme.  Device-resident, layout (V, nx, ny, nz) layout_left with V = 8 (tricubic)
 * or 27 (triquintic): value l of point (i, j, k) at index l + V*(i + nx*(j + ny*k))
 * (src/raspakit/interpolation_energy_grid.cpp:3807-3808, :3882-3924).
 * ------------------------------------------------------------------------------------------- */

/* ForceField::InterpolationScheme (src/raspakit/forcefield/forcefield.ixx:79-84); the value is the number of
 * stored values per grid point.  The Polynomial scheme is not provided. */
typedef enum rg_interpolation_scheme
{
  RG_SCHEME_TRICUBIC = 8,
  RG_SCHEME_TRIQUINTIC = 27
} rg_interpolation_scheme;
typedef enum rg_grid_type
{
  RG_GRID_LENNARD_JONES = 0, /* ForceField::InterpolationGridType::LennardJones */
  RG_GRID_EWALD_REAL = 3     /* ForceField::InterpolationGridType::EwaldReal (unit-charge probe) */
} rg_grid_type;

/* InterpolationEnergyGrid(unitCellBox, origin, numberOfGridPoints, Tricubic) —
 * src/raspakit/interpolation_energy_grid.cpp:3505-3514.  Allocates 8*nx*ny*nz doubles on the device. */
rg_status rg_grid_create(rg_ctx* ctx, const double* cell9, int is_rectangular, int nx, int ny, int nz, rg_grid** out);
/* InterpolationEnergyGrid(unitCellBox, origin, numberOfGridPoints, order) with the scheme given: allocates
 * scheme*nx*ny*nz doubles (src/raspakit/interpolation_energy_grid.cpp:3505-3514). */
rg_status rg_grid_create_scheme(rg_ctx* ctx, const double* cell9, int is_rectangular, int nx, int ny, int nz,
                                rg_interpolation_scheme scheme, rg_grid** out);
rg_status rg_grid_destroy(rg_grid* grid);
/* Device address of the grid data (double*), for callers that keep it resident. */
void* rg_grid_device_data(rg_grid* grid);

/* InterpolationEnergyGrid::makeFrameworkInterpolationGrid(stream, gridType, forceField, framework, cutOff,
 * pseudo_atom_index) — src/raspakit/interpolation_energy_grid.cpp:3800-3965, per-point analytic
 * src/raspakit/interactions/framework_molecule_grid.cpp:168-282 (tricubic) and :549-810 (triquintic; for a
 * rectangular cell the values are the fractional derivatives themselves, where the reference's rectangular branch
 * scales the y and z first derivatives by the cell's x length, :655-658 — equal for cubic cells).  The scheme is
 * the grid's own.  'cutoff_supercell' is the
 * cutOff argument (it only sizes the supercell); pair inclusion uses the force field's framework VDW
 * cutoff, as in the reference.  Requires the supercell to span at least twice that VDW cutoff
 * (RG_ERR_UNSUPPORTED otherwise). */
rg_status rg_grid_build(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fw, rg_grid_type type,
                        int probe_pseudo_atom, double cutoff_supercell, rg_grid* grid);

/* System::createFrameworkInterpolationGrids (src/raspakit/system_interpolation_grids.cpp:176-238) for grids
 * of equal dimensions: one fused device pass builds n_grids grids that share the framework and the
 * distance work.  types[g], probes[g] and cutoffs[g] are as in rg_grid_build. */
rg_status rg_grid_build_many(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fw, int n_grids,
                             const rg_grid_type* types, const int* probes, const double* cutoffs,
                             rg_grid* const* grids);

/* Same as rg_grid_build_many but returns right after the launch (asynchronous on the context stream). */
rg_status rg_grid_build_many_async(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fw, int n_grids,
                                   const rg_grid_type* types, const int* probes, const double* cutoffs,
                                   rg_grid* const* grids);

/* Arithmetic of the Ewald real-space pair terms of a tricubic grid build.  The reference computes every pair term in
 * FP64 (src/raspakit/potentials/interpolation/tricubic_derivative_real_ewald.ixx:32-56). */
typedef enum rg_grid_precision
{
  RG_GRID_FP64 = 0,      /* every pair term in FP64 (what rg_grid_build_many does) */
  RG_GRID_SPLIT_FP32 = 1 /* framework images at least split_radius from every point of a 64-point block: radial
                          * factors in FP32 (hardware rsq and exp2, erfc from a fitted FP32 polynomial), accumulated in
                          * FP64; nearer images in FP64.  Lennard-Jones grids and triquintic grids stay FP64. */
} rg_grid_precision;

/* System::createFrameworkInterpolationGrids (src/raspakit/system_interpolation_grids.cpp:176-238) as
 * rg_grid_build_many, with the precision of the Ewald real-space grid chosen; async != 0 returns after the launch.
 * RG_ERR_UNSUPPORTED when the FP32 erfc polynomial cannot reach 4e-8 relative on [alpha split_radius, alpha rc];
 * RG_ERR_INVALID for a split_radius below sqrt(0.1) A (inside the reference's clamp of r^2 at 0.1 A^2). */
rg_status rg_grid_build_many_precision(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fw, int n_grids,
                                       const rg_grid_type* types, const int* probes, const double* cutoffs,
                                       rg_grid* const* grids, rg_grid_precision precision, double split_radius,
                                       int async);

/* Copies 'count' doubles starting at 'offset' (in doubles) between the grid and host memory. */
rg_status rg_grid_download(rg_ctx* ctx, const rg_grid* grid, size_t offset, size_t count, double* host);
rg_status rg_grid_upload(rg_ctx* ctx, rg_grid* grid, size_t offset, size_t count, const double* host);

/* Grids on disk.  The reference rebuilds every grid at start and restart and archives only the grid's metadata
 * (unit cell, origin, point and cell counts, order; src/raspakit/interpolation_energy_grid.cpp:4793-4836,
 * src/raspakit/system.cpp:1382-1384).  rg_grid_save / rg_grid_load keep the built values in a binary file with that
 * metadata and a 64-bit key of the inputs they were built from, so that a restart loads instead of rebuilding:
 * rg_grid_load refuses (RG_ERR_INVALID) a file whose key, scheme, dimensions or cell differ from the grid's.
 * No reference counterpart for the values in the file. */
/* Key of a grid build: FNV-1a over the grid type, probe, supercell cutoff, precision and split radius, the framework's
 * cell, atoms, types and charges, and the force field's pair table, cutoffs, alpha and overlap criterion. */
rg_status rg_grid_key(const rg_forcefield* ff, const rg_framework* fw, rg_grid_type type, int probe_pseudo_atom,
                      double cutoff_supercell, int precision, double split_radius, unsigned long long* key);
rg_status rg_grid_save(rg_ctx* ctx, const rg_grid* grid, unsigned long long key, const char* path);
rg_status rg_grid_load(rg_ctx* ctx, rg_grid* grid, unsigned long long key, const char* path);
/* InterpolationEnergyGrid::writeOutput (src/raspakit/interpolation_energy_grid.cpp:4760-4790): the Gaussian-cube
 * header the reference writes (origin 0; point counts negated; voxel vectors = lattice vectors / point counts; one
 * dummy atom), then every stored value of the grid in memory order (all V values per point), in Kelvin, one per
 * line, each as the shortest decimal that reads back to the same double. */
rg_status rg_grid_write_cube(rg_ctx* ctx, const rg_grid* grid, const char* path);

/* InterpolationEnergyGrid::interpolate (tricubic and triquintic branches,
 * src/raspakit/interpolation_energy_grid.cpp:3980-4180) for n positions (pos = n*3, host).  energy[n] receives the
 * interpolated energies; gradient (n*3) receives d/dr (interpolateGradient, :4189-4440) when non-null. */
rg_status rg_grid_interpolate(rg_ctx* ctx, const rg_grid* grid, int n, const double* pos, double* energy,
                              double* gradient);

/* InterpolationEnergyGrid::interpolateHessian (src/raspakit/interpolation_energy_grid.cpp:4443-4760), either
 * scheme: energies (n), Cartesian gradients (n*3) and Cartesian Hessians (n*9, row-major) of the interpolant.
 * The Hessian is symmetric; the reference's triquintic branch fills its (z, y) entry with the (x, z) second
 * derivative (:4748), and its tricubic branch returns a zero Hessian (:4600) where this returns the second
 * derivatives of the tricubic interpolant. */
rg_status rg_grid_interpolate_hessian(rg_ctx* ctx, const rg_grid* grid, int n, const double* pos, double* energy,
                                      double* gradient, double* hessian);

/* Test of a built grid on random points, as System::createFrameworkInterpolationGrids does after building
 * (src/raspakit/system_interpolation_grids.cpp:240-840, numberOfGridTestPoints = 1e5 by default): the
 * interpolation of the Lennard-Jones grid of pseudo-atom 'probe' (and, when coulomb_grid is non-null, of the
 * unit-charge Ewald real-space grid scaled by probe_charge) against the analytic sums of
 * Interactions::calculateTricubicFractionalAtPosition / calculateTriquinticFractionalAtPosition at the same points
 * (both grids of one scheme; Hessians for the triquintic scheme).  Points are uniform in the unit cell:
 * frac (n_points * 3 fractional coordinates) when non-null, else drawn from a Philox4x32-10 stream keyed by seed.
 * Every point is weighted by exp(-beta E_LJ) with beta = 1/(k_B T) in inverse internal energy units and E_LJ in K,
 * as the reference computes it; the Ewald statistics take the points whose E_LJ (K) is below the overlap criterion.
 * Energies in K, gradients in K/A, Hessians in K/A^2 (Cartesian). */
typedef struct rg_grid_check_stats
{
  double weight;                      /* sum of the Boltzmann weights */
  double energy_table, energy_full;   /* Boltzmann averages, interpolated and analytic */
  double energy_relative_error;       /* sqrt(sum w (full - table)^2 / sum w full^2) */
  double gradient_table[3], gradient_full[3], gradient_relative_error[3];
  /* Cartesian Hessian components xx, xy, xz, yy, yz, zz in K/A^2: triquintic grids only, zero for tricubic ones
   * (the reference compares Hessians for the triquintic scheme only, system_interpolation_grids.cpp:694-744) */
  double hessian_table[6], hessian_full[6], hessian_relative_error[6];
} rg_grid_check_stats;

typedef struct rg_grid_check_t
{
  long long n_points;
  rg_grid_check_stats lennard_jones;
  long long n_coulomb_points;
  rg_grid_check_stats coulomb;        /* zero when coulomb_grid is NULL */
} rg_grid_check_t;

rg_status rg_grid_check(rg_ctx* ctx, const rg_forcefield* ff, const rg_framework* fw, const rg_grid* lj_grid, int probe,
                        const rg_grid* coulomb_grid, double probe_charge, double temperature, int n_points,
                        const double* frac, unsigned long long seed, rg_grid_check_t* out);

/* ---------------------------------------------------------------------------------------------
 * Systems: what CBMC::GrowContext (src/raspakit/cbmc/cbmc_growth_context.ixx:15-29) gives the trial-energy code —
 * box, force field, framework atoms or interpolation grids, adsorbate atoms — plus the Ewald state
 * (System::storedEik, src/raspakit/system.ixx:235-241), all device-resident.
 * ------------------------------------------------------------------------------------------- */
typedef struct rg_system rg_system;

/* fw: framework atoms for the direct minimum-image sum (may be NULL).  grids[g] is the LJ grid of pseudo-atom
 * grid_types[g] (System::interpolationGrids[type], src/raspakit/system.ixx:296); coulomb_grid the unit-charge
 * grid (interpolationGrids.back()) or NULL.  A site whose type has a grid uses it, as in
 * src/raspakit/cbmc/interactions_framework_molecule.cpp:57-68.  kmax3 = ForceField::numberOfWaveVectors;
 * recip_int_cut2 = reciprocalIntegerCutOffSquared (<= 0: none, the automaticEwald = false default).
 * capacity_atoms bounds the adsorbate atoms. */
rg_status rg_system_create(rg_ctx* ctx, const rg_forcefield* ff, const double* cell9, int is_rectangular,
                           const rg_framework* fw, int n_grids, const int* grid_types, rg_grid* const* grids,
                           const rg_grid* coulomb_grid, const int* kmax3, long long recip_int_cut2,
                           int capacity_atoms, rg_system** out);
rg_status rg_system_destroy(rg_system* sys);
rg_status rg_system_info(const rg_system* sys, int* n_kvectors, double* single_ion_fourier_sum, double* net_charge,
                         int* n_adsorbate_atoms);
/* Replaces the adsorbate atoms (pos n*3; molecule ids consecutive per molecule) and rebuilds the stored structure
 * factor (framework + adsorbates). */
rg_status rg_system_set_adsorbates(rg_ctx* ctx, rg_system* sys, int n, const double* pos, const int* type,
                                   const double* charge, const int* molecule);

/* Trial-site energies: CBMC::computeFrameworkMoleculeEnergy + CBMC::computeInterMolecularEnergy for n sites
 * (src/raspakit/cbmc/interactions.cpp:54-96, interactions_framework_molecule.cpp:31-111,
 * interactions_intermolecular.cpp:22-95).  Adsorbate atoms of the site's own molecule and of 'skip_molecule'
 * (skipBackgroundMolecule; -1 for none) are excluded.  energies[4*i + 0..3] = framework VDW, framework Coulomb,
 * inter-molecular VDW, inter-molecular Coulomb; overlap[i] = 1 where the reference returns std::nullopt. */
rg_status rg_trial_energies(rg_ctx* ctx, const rg_system* sys, int n, const double* pos, const int* type,
                            const double* charge, const int* molecule, int skip_molecule, double* energies,
                            int* overlap);

/* No reference counterpart: rg_trial_energies for callers that cannot allocate device memory (code under HIP graph
 * capture, tight Monte Carlo or serving loops, harnesses that forbid hipMalloc).  Same arguments, results and errors,
 * but the device staging of the inputs and results lives in a static array of the library (one per device, held until
 * the process ends) instead of a per-call allocation, so n may not exceed RG_TRIAL_NOALLOC_MAX_SITES (RG_ERR_INVALID
 * beyond it).  The array is shared by every call on a device: calls that overlap on one device (from several host
 * threads or contexts) race and must be serialised by the caller. */
#define RG_TRIAL_NOALLOC_MAX_SITES 16384
rg_status rg_trial_energies_noalloc(rg_ctx* ctx, const rg_system* sys, int n, const double* pos, const int* type,
                                    const double* charge, const int* molecule, int skip_molecule, double* energies,
                                    int* overlap);

/* Rosenbluth weight and selection for n_sets sets of k trials (src/raspakit/cbmc/multiple_first_bead.cpp:20-49,
 * src/raspakit/cbmc/util.cpp:15-48): log_weight[s] = log(sum_i exp(-beta U_i) / k) over non-overlapping trials,
 * selected[s] = the trial the cumulative search picks for uniforms[s] (-1 if all overlap). */
rg_status rg_rosenbluth_select(rg_ctx* ctx, int n_sets, int k, const double* energies, const int* overlap,
                               double beta, const double* uniforms, int* selected, double* log_weight);

/* Interactions::energyDifferenceEwaldFourier (src/raspakit/interactions/ewald.cpp:1043-1244) replacing the old
 * atoms by the new ones (charges already scaled): out[0] = ewald_fourier (incl. the net-charge term),
 * out[1] = ewald_self, out[2] = ewald_exclusion.  Writes the trial structure factor; at most 32 atoms. */
rg_status rg_ewald_delta(rg_ctx* ctx, rg_system* sys, int n_new, const double* new_pos, const double* new_charge,
                         const int* new_molecule, int n_old, const double* old_pos, const double* old_charge,
                         const int* old_molecule, double* out);
/* Interactions::acceptEwaldMove (src/raspakit/interactions/ewald.cpp:1565-1574): stored <- trial. */
rg_status rg_ewald_accept(rg_ctx* ctx, rg_system* sys);
/* Interactions::computeEwaldFourierEnergy (src/raspakit/interactions/ewald.cpp:352-563) of the adsorbates from the
 * stored structure factor: out[0] = fourier (incl. net charge), out[1] = self, out[2] = exclusion. */
rg_status rg_ewald_energy(rg_ctx* ctx, const rg_system* sys, double* out);

/* ---------------------------------------------------------------------------------------------
 * Monte Carlo on the device: Widom insertions and multi-system GCMC of rigid molecules and flexible linear chains
 * (src/raspakit/mc_moves/component/{widom,insertion_cbmc,deletion_cbmc,translation,rotation,reinsertion}.cpp; first
 * beads src/raspakit/cbmc/multiple_first_bead.cpp; orientations of the rigid body
 * src/raspakit/cbmc/operators.cpp:958-968; chain growth src/raspakit/cbmc/flexible_molecules/insertion.cpp).
 * ------------------------------------------------------------------------------------------- */

/* A component (Component, src/raspakit/component.ixx).  Rigid (flexible = 0): up to 4 atoms in a reference geometry;
 * the first bead is the atom grown first (Component::startingBead).  Flexible linear chain (flexible = 1, uncharged,
 * up to 8 united-atom beads bonded in index order, first_bead = 0): bonds of fixed length, harmonic bend
 * U = k (theta - theta0)^2 / 2, torsion U = c0 + c1 (1 + cos p) + c2 (1 - cos 2p) + c3 (1 + cos 3p) (trans at 180 deg),
 * Lennard-Jones between beads more than three bonds apart (src/raspakit/cbmc/flexible_molecules/insertion.cpp).
 * Bend and torsion are sampled exactly from their Boltzmann factors, so growth weights carry the external and
 * intra-molecular non-bonded energies; log_ideal_rosenbluth is log W_IG of that growth (Widom insertions into an
 * empty box give it), divided out in the acceptance rules. */
typedef struct rg_component_desc
{
  int n_atoms;          /* 1..4 rigid, 2..8 flexible */
  int first_bead;
  double position[24];  /* n_atoms * 3, reference geometry (A); unused for flexible chains */
  int type[8];          /* pseudo-atom indices */
  double charge[8];
  double mol_fraction;  /* Component::molFraction: the fugacity is mol_fraction x pressure (components are drawn
                         * uniformly per step, System::randomComponent, src/raspakit/system.ixx:372-375) */
  int flexible;
  double bond_length;       /* A */
  double bend_theta0_deg;
  double bend_k_K;          /* K / rad^2 */
  double torsion_K[4];      /* c0 .. c3 in K */
  double log_ideal_rosenbluth;
} rg_component_desc;

typedef struct rg_mc_params
{
  double temperature;       /* K */
  double pressure_pa;       /* Pa; fugacity = mole fraction * pressure (fugacity coefficient 1) */
  int k_first;              /* numberOfFirstBeadPositions (reference default 10), <= 64 */
  int k_orient;             /* numberOfTrialDirections (reference default 10), <= 64 */
  double p_translation;     /* move probabilities (with p_reinsertion below); the rest are CBMC swaps, half
                             * insertions, half deletions */
  double p_rotation;
  double max_translation;   /* A */
  double max_rotation;      /* rad */
  int min_steps_per_cycle;  /* steps per cycle = max(this, N) x number of components
                             * (src/raspakit/drivers/monte_carlo.cpp:301) */
  int max_molecules;        /* per system; each system's capacity_atoms must be >= max_molecules * max(n_atoms) */
  unsigned long long seed;  /* Philox4x32-10 key; trials are addressed by (system, step, slot) */
  long long production_cycles; /* cycles over which the 5 block averages are spread */
  double p_reinsertion;     /* probability of a CBMC reinsertion (MC_Moves::reinsertionMove,
                             * src/raspakit/mc_moves/component/reinsertion.cpp:34-214) */
} rg_mc_params;

typedef struct rg_mc_state_t
{
  int n_molecules;
  int n_per_component[2];
  double energy;            /* running adsorbate energy (internal units) */
  long long steps, cycles_sampled;
  long long trials[5], accepted[5];  /* insertion, deletion, translation, rotation, reinsertion */
  double block_count[5], block_energy[5], block_n[5][2], block_energy_n[5][2], block_nn[5][2][2];
} rg_mc_state_t;

typedef struct rg_mc rg_mc;

/* Widom insertions first..first+count-1 into a system (MC_Moves::WidomMove, widom.cpp:32-155): per insertion the
 * log Rosenbluth weight (-inf if every trial overlaps), the external energy of the grown molecule and the Ewald
 * Fourier energy difference.  W = exp(log_weight - beta ewald) / W_IG with W_IG = 1 for rigid bodies.  Host
 * output arrays; returns when they are filled. */
rg_status rg_widom(rg_ctx* ctx, const rg_system* sys, const rg_component_desc* component, const rg_mc_params* params,
                   long long first, int count, double* log_weight, double* energy, double* ewald);

/* Same as rg_widom with device-resident output arrays (count doubles each), for callers that keep the results on
 * the GPU; returns when the insertions have finished. */
rg_status rg_widom_device(rg_ctx* ctx, const rg_system* sys, const rg_component_desc* component,
                          const rg_mc_params* params, long long first, int count, double* d_log_weight,
                          double* d_energy, double* d_ewald);

/* GCMC over n_systems independent, initially empty systems sharing one force field; one wavefront per system. */
rg_status rg_mc_create(rg_ctx* ctx, int n_systems, rg_system* const* systems, int n_components,
                       const rg_component_desc* components, const rg_mc_params* params, rg_mc** out);
/* Releases the engine.  The systems hold the adsorbates, Ewald state and net charge of the last rg_mc_run, so they
 * may be released before or after the engine (but not used by the engine once released). */
rg_status rg_mc_destroy(rg_mc* mc);
/* Learned first-bead proposal for one system, the same for every component: probability[v] > 0 over
 * n3[0]*n3[1]*n3[2] voxels of the cell (x fastest); NULL restores the uniform proposal.  No reference counterpart
 * (the reference's first beads are uniform, src/raspakit/cbmc/multiple_first_bead.cpp:26-27). */
rg_status rg_mc_set_proposal(rg_ctx* ctx, rg_mc* mc, int system, const int* n3, const double* probability);
/* The same for one component of one system (each adsorbate its own proposal); no reference counterpart. */
rg_status rg_mc_set_component_proposal(rg_ctx* ctx, rg_mc* mc, int system, int component, const int* n3,
                                       const double* probability);
/* Runs 'cycles' MC cycles; sample != 0 accumulates block averages; use_graph != 0 replays one captured cycle.
 * On return the systems' adsorbate arrays, structure factors, adsorbate count and net charge are the final state. */
rg_status rg_mc_run(rg_ctx* ctx, rg_mc* mc, int cycles, int sample, int use_graph);
/* Launch shape of rg_mc_run (no counterpart in the reference): -1 (default) lets the autotuner choose; 0, 1, 2 run
 * one wavefront per system with 1, 2 or 4 systems per workgroup; 3 runs a team of four wavefronts per system.  Every
 * shape computes the same trajectory bit for bit, so this only pins the speed (for reproducible timings). */
rg_status rg_mc_set_launch_shape(rg_ctx* ctx, rg_mc* mc, int shape);
/* Balanced launches (no counterpart in the reference, whose systems run one after another): with enable != 0 each
 * of the 'cycles' launches of rg_mc_run gives every system at least one cycle and lets it run further whole cycles
 * while their summed cost stays within the largest single-cycle cost over the systems at the start of the launch,
 * a cycle's cost being its steps times (molecules + overhead_molecules).  Lightly loaded or small systems then do
 * more cycles in the time the most loaded one takes for its first, instead of leaving their compute units idle; each
 * system's trajectory is the same sequence of cycles as without balancing (only their number changes), and the
 * samples of a launch go to the block of the launch's index among the sampled launches.  Rigid components only
 * (RG_ERR_UNSUPPORTED with a flexible one); set it before the run it applies to. */
rg_status rg_mc_set_balance(rg_ctx* ctx, rg_mc* mc, int enable, double overhead_molecules);
rg_status rg_mc_state(rg_ctx* ctx, const rg_mc* mc, int system, rg_mc_state_t* out);
/* Adsorbate atoms of a system (max(n_atoms) of the components slots per molecule; unused slots have type -1)
 * and the component of each molecule. */
rg_status rg_mc_adsorbates(rg_ctx* ctx, const rg_mc* mc, int system, int capacity, double* pos, int* type,
                           int* component, int* n_out);
/* Training data of the learned proposal: the orientation-averaged Boltzmann factor of a rigid molecule with its
 * first bead at each of n points (fractional coordinates of the system's box), framework energy from the system's
 * interpolation grids (every atom type of the component needs one): log_weight[i] = log (1/M) sum_o exp(-U(s_i, R_o)
 * / kT) over M = n_orient (a power of two <= 64) uniformly random orientations drawn from the Philox stream
 * (point i, orientation o) of 'seed'; -inf where every orientation overlaps.  No reference counterpart: it is the
 * first-bead marginal of the orientation trials (src/raspakit/cbmc/operators.cpp:958-968). */
rg_status rg_molecule_weights(rg_ctx* ctx, const rg_system* sys, const rg_component_desc* component,
                              double temperature, int n, const double* frac, int n_orient, unsigned long long seed,
                              double* log_weight);
/* Adsorbate energy of every system recomputed from scratch (System::computeTotalEnergies, src/raspakit/system.cpp:690). */
rg_status rg_mc_total_energy(rg_ctx* ctx, const rg_mc* mc, double* out);

/* ---------------------------------------------------------------------------------------------
 * Statistics of a run, as the reference reports them (host-side, no device work).  Blocks are the reference's 5 block
 * averages (rg_mc_state_t carries their weighted sums); a confidence interval is the 95 % half-width
 * t_(n-1) s / sqrt(n) over the blocks whose weight exceeds half the first block's, 0 with fewer than 3 such blocks.
 * ------------------------------------------------------------------------------------------- */

/* blockErrorEstimate (src/raspakit/averages.ixx:144-169; Student-t deviates :28-34, 95 % :46): the confidence
 * interval of n <= 21 block averages about 'average'. */
rg_status rg_block_error(int n, const double* block_averages, double average, double* ci);

/* BlockAverage<double>::average (src/raspakit/properties/property_block_average.ixx:108-150) of a directly averaged
 * property: mean = sum(sums) / sum(weights), ci over the block averages sums[b] / weights[b]. */
rg_status rg_block_average(int n_blocks, const double* sums, const double* weights, double* mean, double* ci);

/* EnthalpyOfAdsorptionTerms (src/raspakit/properties/enthalpy.ixx:101-200) block-averaged as a composite property
 * (property_block_average.ixx:108-126): per block, from the weighted sums of U, N_i, U N_i and N_i N_j (row-major
 * [block][i] and [block][i][j] with n_components columns; rg_mc_state_t's arrays have this layout for two
 * components, for one component take column 0), Delta H_i = sum_j (<U N_j> - <U><N_j>) (M^-1)_ji - k_B T with
 * M_ij = <N_i N_j> - <N_i><N_j>; mean[i] = the mean over the blocks, ci[i] its interval.  Internal energy units
 * (the isosteric heat is -Delta H); temperature in K.  A singular block matrix gives NaN. */
rg_status rg_enthalpy_of_adsorption(int n_blocks, int n_components, const double* weight, const double* sum_u,
                                    const double* sum_n, const double* sum_un, const double* sum_nn,
                                    double temperature, double* mean, double* ci);

/* PropertyWidom::result / chemicalPotentialResult / fugacityResult (src/raspakit/properties/widom.ixx:176-230):
 * from the per-block weighted sums of the Widom Rosenbluth weight W and of the number density N/V (A^-3):
 * mu_excess = -kT ln<W>, mu_ideal_gas = kT ln<N/V>, mu_total = their sum, fugacity = kT exp(mu_total / kT), in
 * internal energy units (fugacity in internal energy / A^3); each with its block interval. */
typedef struct rg_widom_stats_t
{
  double rosenbluth_weight, rosenbluth_weight_ci;
  double mu_excess, mu_excess_ci;
  double mu_ideal_gas, mu_ideal_gas_ci;
  double mu_total, mu_total_ci;
  double fugacity, fugacity_ci;
} rg_widom_stats_t;
rg_status rg_widom_statistics(int n_blocks, const double* weight, const double* sum_w, const double* sum_density,
                              double temperature, rg_widom_stats_t* out);

/* ---------------------------------------------------------------------------------------------
 * Learned first-bead proposal.  The reference has no counterpart: its first-bead trials are uniform in the cell
 * (src/raspakit/cbmc/multiple_first_bead.cpp:26-27 for insertion, :56-75 for retrace).  An MLP on periodic Fourier
 * features of fractional coordinates, 192 -> 256 -> 256 -> 256 -> 1 with SiLU, gives an unnormalised log-density
 * f(s); BF16 weights and activations, FP32 accumulation (MFMA), FP32 master weights updated by Adam.
 * Trained by minibatch softmax cross-entropy against Boltzmann weights exp(-beta U_i).  rg_mlp_proposal tabulates
 * it at voxel centres with a uniform floor, the exactly normalised density handed to rg_mc_set_proposal.
 * ------------------------------------------------------------------------------------------- */
typedef struct rg_mlp rg_mlp;

/* Number of FP32 parameters: W1 (192 x 256), b1, W2 (256 x 256), b2, W3 (256 x 256), b3, w4 (256), b4, in this
 * order, W[in][out] row-major; input q = 64 a + 2 (m - 1) + c is sin (c = 0) or cos (c = 1) of 2 pi m s_a. */
size_t rg_mlp_num_params(void);
/* He-uniform weights from a splitmix64 stream of 'seed', zero biases. */
rg_status rg_mlp_create(rg_ctx* ctx, unsigned long long seed, rg_mlp** out);
rg_status rg_mlp_destroy(rg_mlp* mlp);
rg_status rg_mlp_get_params(rg_ctx* ctx, const rg_mlp* mlp, float* out);
rg_status rg_mlp_set_params(rg_ctx* ctx, rg_mlp* mlp, const float* in);
/* f(s) at n points (fractional coordinates, 3 per point). */
rg_status rg_mlp_forward(rg_ctx* ctx, rg_mlp* mlp, int n, const double* frac, double* logdensity);
/* Loss -sum_i p_i log softmax(f)_i with p = softmax(-beta U) over one batch (n a multiple of 64, <= 65536) and its
 * gradient with respect to every parameter; the parameters are not changed.  Points of energy +inf (overlaps) have
 * p_i = 0, as long as some point of the batch has a finite energy. */
rg_status rg_mlp_gradients(rg_ctx* ctx, rg_mlp* mlp, int n, const double* frac, const double* energy, double beta,
                           float* grads, double* loss);
/* Adam (beta1 0.9, beta2 0.999, eps 1e-8) over 'epochs' shuffled passes of minibatches of 'batch' points (a multiple
 * of 64, <= 65536 and <= n; the n % batch points left over in an epoch are skipped); epoch_loss[epochs] (may be
 * NULL) receives the mean minibatch loss of every epoch. */
rg_status rg_mlp_train(rg_ctx* ctx, rg_mlp* mlp, int n, const double* frac, const double* energy, double beta,
                       int batch, int epochs, double lr, unsigned long long seed, double* epoch_loss);
/* probability[v] = (1 - floor) exp(f(c_v)) / sum_w exp(f(c_w)) + floor / nvox at the centres c_v of
 * n3[0] x n3[1] x n3[2] voxels (x fastest); 0 < floor <= 1. */
rg_status rg_mlp_proposal(rg_ctx* ctx, rg_mlp* mlp, const int* n3, double floor_fraction, double* probability);

#ifdef __cplusplus
}
#endif

#endif

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
rg_status rg_ctx_destroy(rg_ctx* ctx);
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
 * tricubic scheme.  Device-resident, layout (8, nx, ny, nz) layout_left:
 * value l of point (i, j, k) at index l + 8*(i + nx*(j + ny*k)) (src/raspakit/interpolation_energy_grid.cpp:3807-3808).
 * ------------------------------------------------------------------------------------------- */
typedef enum rg_grid_type
{
  RG_GRID_LENNARD_JONES = 0, /* ForceField::InterpolationGridType::LennardJones */
  RG_GRID_EWALD_REAL = 3     /* ForceField::InterpolationGridType::EwaldReal (unit-charge probe) */
} rg_grid_type;

/* InterpolationEnergyGrid(unitCellBox, origin, numberOfGridPoints, Tricubic) —
 * src/raspakit/interpolation_energy_grid.cpp:3505-3514.  Allocates 8*nx*ny*nz doubles on the device. */
rg_status rg_grid_create(rg_ctx* ctx, const double* cell9, int is_rectangular, int nx, int ny, int nz, rg_grid** out);
rg_status rg_grid_destroy(rg_grid* grid);
/* Device address of the grid data (double*), for callers that keep it resident. */
void* rg_grid_device_data(rg_grid* grid);

/* InterpolationEnergyGrid::makeFrameworkInterpolationGrid(stream, gridType, forceField, framework, cutOff,
 * pseudo_atom_index), tricubic branch — src/raspakit/interpolation_energy_grid.cpp:3800-3875, per-point
 * analytic src/raspakit/interactions/framework_molecule_grid.cpp:168-282.  'cutoff_supercell' is the
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

/* Copies 'count' doubles starting at 'offset' (in doubles) between the grid and host memory. */
rg_status rg_grid_download(rg_ctx* ctx, const rg_grid* grid, size_t offset, size_t count, double* host);
rg_status rg_grid_upload(rg_ctx* ctx, rg_grid* grid, size_t offset, size_t count, const double* host);

/* InterpolationEnergyGrid::interpolate (tricubic branch, src/raspakit/interpolation_energy_grid.cpp:3980-4070)
 * for n positions (pos = n*3, host).  energy[n] receives the interpolated energies; gradient (n*3) receives
 * d/dr (interpolateGradient, :4189-4347) when non-null. */
rg_status rg_grid_interpolate(rg_ctx* ctx, const rg_grid* grid, int n, const double* pos, double* energy,
                              double* gradient);

#ifdef __cplusplus
}
#endif

#endif

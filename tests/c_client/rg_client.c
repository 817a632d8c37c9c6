/*
 * A plain C99 client of the raspa3_amd C ABI (include/raspa3_amd/rg.h): what a non-Python caller links against.
 *
 * Reads a whitespace-separated description from the file named by argv[1]:
 *   cell (9 numbers, lattice vectors as columns) is_rectangular
 *   n_types, then per type: eps_K sigma is_framework
 *   shifted rc_framework rc_molecule rc_coulomb alpha use_charge overlap
 *   n_atoms, then per atom: x y z type charge
 *   probe spacing                    (Lennard-Jones grid of pseudo-atom 'probe' at the given spacing)
 *   n_sites, then per site: x y z type charge
 * and prints, per site, the direct framework energies (rg_trial_energies on a system without grids) and the
 * tricubic interpolation of the grid (rg_grid_interpolate), in internal energy units:
 *   dims nx ny nz
 *   site i overlap e_vdw e_coulomb e_grid
 * Exit status 0 on success; on an error the entry point's status and rg_last_error are printed.
 */
#include <stdio.h>
#include <stdlib.h>

#include "raspa3_amd/rg.h"

static int check(rg_status st, rg_ctx* ctx, const char* what)
{
  if (st != RG_OK)
  {
    fprintf(stderr, "%s failed: status %d (%s)\n", what, (int)st, ctx ? rg_last_error(ctx) : "");
    return 1;
  }
  return 0;
}

static int read_doubles(FILE* f, double* v, int n)
{
  int i;
  for (i = 0; i < n; ++i)
    if (fscanf(f, "%lf", &v[i]) != 1) return 1;
  return 0;
}

int main(int argc, char** argv)
{
  FILE* f;
  double cell[9], ffp[7], spacing;
  int rect, n_types, n_atoms, n_sites, probe, i, n3[3], rc = 1;
  double *eps = NULL, *sig = NULL, *pos = NULL, *q = NULL, *site = NULL, *site_q = NULL, *energies = NULL;
  double* e_grid = NULL;
  int *is_fw = NULL, *typ = NULL, *site_t = NULL, *site_m = NULL, *overlap = NULL;
  rg_ctx* ctx = NULL;
  rg_forcefield* ff = NULL;
  rg_framework* fw = NULL;
  rg_grid* grid = NULL;
  rg_system* sys = NULL;
  rg_grid_type gtype = RG_GRID_LENNARD_JONES;
  double cutoff;
  int kmax3[3] = {0, 0, 0};

  if (argc != 2 || !(f = fopen(argv[1], "r")))
  {
    fprintf(stderr, "usage: rg_client input.txt\n");
    return 2;
  }
  if (read_doubles(f, cell, 9) || fscanf(f, "%d %d", &rect, &n_types) != 2 || n_types <= 0) goto bad_input;
  eps = malloc(sizeof(double) * n_types);
  sig = malloc(sizeof(double) * n_types);
  is_fw = malloc(sizeof(int) * n_types);
  for (i = 0; i < n_types; ++i)
    if (fscanf(f, "%lf %lf %d", &eps[i], &sig[i], &is_fw[i]) != 3) goto bad_input;
  if (read_doubles(f, ffp, 7) || fscanf(f, "%d", &n_atoms) != 1 || n_atoms <= 0) goto bad_input;
  pos = malloc(sizeof(double) * 3 * n_atoms);
  q = malloc(sizeof(double) * n_atoms);
  typ = malloc(sizeof(int) * n_atoms);
  for (i = 0; i < n_atoms; ++i)
    if (read_doubles(f, pos + 3 * i, 3) || fscanf(f, "%d %lf", &typ[i], &q[i]) != 2) goto bad_input;
  if (fscanf(f, "%d %lf %d", &probe, &spacing, &n_sites) != 3 || n_sites <= 0) goto bad_input;
  site = malloc(sizeof(double) * 3 * n_sites);
  site_q = malloc(sizeof(double) * n_sites);
  site_t = malloc(sizeof(int) * n_sites);
  site_m = malloc(sizeof(int) * n_sites);
  for (i = 0; i < n_sites; ++i)
  {
    if (read_doubles(f, site + 3 * i, 3) || fscanf(f, "%d %lf", &site_t[i], &site_q[i]) != 2) goto bad_input;
    site_m[i] = i;
  }
  fclose(f);
  f = NULL;

  if (check(rg_ctx_create(0, &ctx), NULL, "rg_ctx_create")) goto done;
  if (check(rg_forcefield_create(ctx, n_types, eps, sig, is_fw, (int)ffp[0], ffp[1], ffp[2], ffp[3], ffp[4], (int)ffp[5],
                                 ffp[6], &ff),
            ctx, "rg_forcefield_create"))
    goto done;
  if (check(rg_framework_create(ctx, cell, rect, n_atoms, pos, typ, q, &fw), ctx, "rg_framework_create")) goto done;
  if (check(rg_grid_dims(cell, spacing, n3), ctx, "rg_grid_dims")) goto done;
  if (check(rg_grid_create(ctx, cell, rect, n3[0], n3[1], n3[2], &grid), ctx, "rg_grid_create")) goto done;
  cutoff = ffp[1];
  if (check(rg_grid_build_many(ctx, ff, fw, 1, &gtype, &probe, &cutoff, &grid), ctx, "rg_grid_build_many")) goto done;
  if (check(rg_system_create(ctx, ff, cell, rect, fw, 0, NULL, NULL, NULL, kmax3, 0, 64, &sys), ctx, "rg_system_create"))
    goto done;
  energies = malloc(sizeof(double) * 4 * n_sites);
  overlap = malloc(sizeof(int) * n_sites);
  e_grid = malloc(sizeof(double) * n_sites);
  if (check(rg_trial_energies(ctx, sys, n_sites, site, site_t, site_q, site_m, -1, energies, overlap), ctx,
            "rg_trial_energies"))
    goto done;
  if (check(rg_grid_interpolate(ctx, grid, n_sites, site, e_grid, NULL), ctx, "rg_grid_interpolate")) goto done;
  printf("dims %d %d %d\n", n3[0], n3[1], n3[2]);
  for (i = 0; i < n_sites; ++i)
    printf("site %d %d %.17g %.17g %.17g\n", i, overlap[i], energies[4 * i], energies[4 * i + 1], e_grid[i]);
  rc = 0;
  goto done;

bad_input:
  fprintf(stderr, "rg_client: malformed input\n");
  rc = 2;
done:
  if (f) fclose(f);
  if (sys) rc |= check(rg_system_destroy(sys), NULL, "rg_system_destroy");
  if (grid) rc |= check(rg_grid_destroy(grid), NULL, "rg_grid_destroy");
  if (fw) rc |= check(rg_framework_destroy(fw), NULL, "rg_framework_destroy");
  if (ff) rc |= check(rg_forcefield_destroy(ff), NULL, "rg_forcefield_destroy");
  if (ctx) rc |= check(rg_ctx_destroy(ctx), NULL, "rg_ctx_destroy");
  free(eps);
  free(sig);
  free(is_fw);
  free(pos);
  free(q);
  free(typ);
  free(site);
  free(site_q);
  free(site_t);
  free(site_m);
  free(energies);
  free(overlap);
  free(e_grid);
  return rc;
}

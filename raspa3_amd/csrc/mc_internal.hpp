// Device-resident simulation-box state shared by the trial-energy, Ewald and Monte Carlo kernels.
#pragma once

#include "mc_device.hpp"
#include "rg_internal.hpp"

#define EW_MAXATOMS 32  // atoms in one Ewald difference (new + old)
#define EW_MAXK 15      // largest |k| per axis

namespace rg
{

// Plain-old-data view of one system, copied to the device; kernels take a const pointer to it.
struct SysDev
{
  rgd::Box box;
  double alpha, coulomb, overlap;
  double rcv2Fw, rcv2Mol, rcc2;
  int useCharge;
  int ntypes;
  const rgd::Pair* pairs;  // ntypes * ntypes
  // framework atoms for the direct sum (SoA)
  int nfw;
  const double *fwx, *fwy, *fwz, *fwq;
  const int* fwt;
  // interpolation grids: gridOfType[type] indexes grids (or -1); coulGrid indexes the unit-charge grid (or -1)
  const int* gridOfType;
  const rgd::GridView* grids;
  int coulGrid;
  // adsorbate atoms (SoA, capacity cap)
  int nads, cap;
  double *ax, *ay, *az, *aq;
  int *at, *amol;
  // Ewald
  int nk;
  int kmax[3];
  const int* kcode;  // kx | (ky + 128) << 8 | (kz + 128) << 16
  const double* fk;  // prefactor incl. the factor 2 for kx > 0
  double2* Sk;       // stored structure factor (framework + adsorbates)
  double2* St;       // trial structure factor
  const double2* Sfw;  // framework part
  double singleIon;    // sum_k f_k (a unit charge)
  double netCharge;    // framework + adsorbates
};

}  // namespace rg

struct rg_system
{
  rg_ctx* ctx = nullptr;
  int device = 0;
  const rg_forcefield* ff = nullptr;
  rg::SysDev h{};              // host copy of the device view
  rg::SysDev* d = nullptr;     // device copy
  std::vector<void*> owned;    // device buffers freed on destroy
  std::vector<rgd::GridView> gridViews;
  std::vector<int> gridOfType;  // host copy of h.gridOfType (-1: direct sums for that type)
  double fwNetCharge = 0.0;
  double pendingCharge = 0.0;  // charge change of the last rg_ewald_delta, applied by rg_ewald_accept
  long long intCut2 = 0;
};

namespace rg
{
rg_status syncSystem(rg_system* s);  // copies s->h to s->d
}

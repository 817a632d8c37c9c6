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
  // framework atoms for the direct sum (SoA), sorted by cell when cellStart is set: cell (c0, c1, c2) of ncell
  // cells per axis (fractional cells, each at least the cutoff wide) holds atoms cellStart[c] .. cellStart[c + 1] - 1,
  // c = (c2 ncell[1] + c1) ncell[0] + c0
  int nfw;
  const double *fwx, *fwy, *fwz, *fwq;
  const int* fwt;
  int ncell[3];
  const int* cellStart;
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

// Calls f(a) for the framework atoms that can lie within the cutoff of (px, py, pz), lanes striding over them:
// with a cell list the atoms of the 27 cells around the point's cell (every axis has at least three cells, each at
// least one cutoff wide, so the 27 are distinct and hold every atom within the cutoff under the minimum image),
// otherwise all atoms.  The caller applies the cutoff to the minimum-image distance, as for a full scan.
template <class F>
__device__ __forceinline__ void forNearFramework(const SysDev& S, double px, double py, double pz, int lane, F&& f)
{
  if (!S.cellStart)
  {
    for (int a = lane; a < S.nfw; a += 64) f(a);
    return;
  }
  int c[3];
#pragma unroll
  for (int ax = 0; ax < 3; ++ax)
  {
    double s = S.box.inv[ax] * px + S.box.inv[3 + ax] * py + S.box.inv[6 + ax] * pz;
    s -= floor(s);
    c[ax] = min(S.ncell[ax] - 1, int(s * S.ncell[ax]));
  }
  for (int d2 = -1; d2 <= 1; ++d2)
  {
    const int c2 = (c[2] + d2 + S.ncell[2]) % S.ncell[2];
    for (int d1 = -1; d1 <= 1; ++d1)
    {
      const int c1 = (c[1] + d1 + S.ncell[1]) % S.ncell[1];
      for (int d0 = -1; d0 <= 1; ++d0)
      {
        const int c0 = (c[0] + d0 + S.ncell[0]) % S.ncell[0];
        const int cell = (c2 * S.ncell[1] + c1) * S.ncell[0] + c0;
        const int end = S.cellStart[cell + 1];
        for (int a = S.cellStart[cell] + lane; a < end; a += 64) f(a);
      }
    }
  }
}

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

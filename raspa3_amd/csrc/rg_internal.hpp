// Internal declarations shared by the raspa3_amd HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "raspa3_amd/rg.h"

struct rg_ctx
{
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  void* d_table = nullptr;  // grid-build parameter tables (persistent, grown on demand)
  size_t table_bytes = 0;
};

struct rg_forcefield
{
  rg_ctx* ctx = nullptr;
  int n = 0;
  std::vector<double> eps, sigma, shift;  // n*n, internal units / Angstrom
  std::vector<int> lj;                    // n*n
  std::vector<int> is_framework;
  double rc_framework = 12.0, rc_molecule = 12.0, rc_coulomb = 12.0;
  double alpha = 0.265058;
  int use_charge = 1;
  double overlap = 1e6;
};

// Periodic images of the framework atoms around the unit cell, binned in fractional space (built once per
// framework and cutoff, kept on the device).
struct ImageSet
{
  double rc = -1.0;
  double binTarget = 0.0;  // minimum bin width (A) the bins were built with
  int n_images = 0;
  int nb[3] = {0, 0, 0};
  double lo[3] = {0, 0, 0}, bw[3] = {0, 0, 0};
  void* d_img = nullptr;    // Img[n_images]
  int* d_bin_start = nullptr;  // nb0*nb1*nb2 + 1
};

struct rg_framework
{
  rg_ctx* ctx = nullptr;  // the context that created it (must outlive work issued on it)
  int device = 0;
  double cell[9], inv[9];
  int rect = 1;
  int n = 0;
  std::vector<double> pos, charge;
  std::vector<int> type;
  double widths[3];
  ImageSet images;
};

struct rg_grid
{
  rg_ctx* ctx = nullptr;
  int device = 0;
  double cell[9], inv[9];
  int rect = 1;
  int n[3] = {0, 0, 0};
  int values = 8;    // per point: 8 (tricubic) or 27 (triquintic)
  size_t count = 0;  // doubles
  double* d = nullptr;
};

namespace rg
{

// Physical constants of the reference (src/foundationkit/units.ixx, src/raspakit/units.cpp:15-100).
struct Units
{
  double energyToKelvin, kelvinToEnergy, coulomb, kB;
};
const Units& units();

rg_status fail(rg_ctx* ctx, rg_status s, const char* fmt, ...);
rg_status hipfail(rg_ctx* ctx, hipError_t e, const char* what);

void invert3(const double* m, double* inv);  // column storage
double det3(const double* m);
void widths3(const double* m, double* w);

}  // namespace rg

#define RG_HIP(ctx, call)                                      \
  do                                                           \
  {                                                            \
    hipError_t rg_e_ = (call);                                 \
    if (rg_e_ != hipSuccess) return rg::hipfail(ctx, rg_e_, #call); \
  } while (0)

// Entry of an API call: makes the context's device current and reports an error that an earlier HIP call (another
// library, or an unchecked release path) left pending, instead of letting the hipGetLastError after this call's own
// launches attribute it to this call.  Reading the error clears it, so the caller may retry.
#define RG_ENTER(ctx)                                                                                  \
  do                                                                                                   \
  {                                                                                                    \
    RG_HIP(ctx, hipSetDevice((ctx)->device));                                                          \
    const hipError_t rg_pending_ = hipGetLastError();                                                  \
    if (rg_pending_ != hipSuccess)                                                                     \
      return rg::fail(ctx, RG_ERR_DEVICE, "%s: a HIP error was pending before this call: %s", __func__, \
                      hipGetErrorString(rg_pending_));                                                 \
  } while (0)

// Status of a release path: the first failing HIP call of a destroy function (the object is released anyway).
inline rg_status rgReleaseStatus(hipError_t e, rg_status st) { return (st == RG_OK && e != hipSuccess) ? RG_ERR_DEVICE : st; }

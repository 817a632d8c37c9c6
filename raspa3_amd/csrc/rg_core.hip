// Context, simulation-box, force-field and framework set-up of the raspa3_amd C ABI.
#include <cmath>
#include <cstring>

#include "rg_internal.hpp"

namespace rg
{

const Units& units()
{
  // src/foundationkit/units.ixx:25-60 (SI constants), src/raspakit/units.cpp:15-100 (derived values)
  static const Units u = []
  {
    const double L = 1e-10, T = 1e-12, M = 1.6605402e-27, Q = 1.60217733e-19;
    const double R = 8.314464919, NA = 6.0221419947e23, kB_SI = 1.380650324e-23, eps0 = 8.8541878176e-12;
    const double E = M * L * L / (T * T);
    Units x;
    x.energyToKelvin = E * NA / R;
    x.kelvinToEnergy = R / (E * NA);
    x.coulomb = Q * Q / (4.0 * M_PI * eps0 * L * E);
    x.kB = kB_SI / E;
    return x;
  }();
  return u;
}

rg_status fail(rg_ctx* ctx, rg_status s, const char* fmt, ...)
{
  if (ctx)
  {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    ctx->err = buf;
  }
  return s;
}

rg_status hipfail(rg_ctx* ctx, hipError_t e, const char* what)
{
  // the error is reported here: clear HIP's last-error state so that the next entry point does not report it again
  (void)hipGetLastError();
  return fail(ctx, e == hipErrorOutOfMemory ? RG_ERR_NOMEM : RG_ERR_DEVICE, "%s failed: %s", what,
              hipGetErrorString(e));
}

// Column storage: m[3*col + row].  Same cofactor formula as src/mathkit/double3x3.cpp:102-118.
void invert3(const double* m, double* inv)
{
  const double m11 = m[0], m21 = m[1], m31 = m[2];
  const double m12 = m[3], m22 = m[4], m32 = m[5];
  const double m13 = m[6], m23 = m[7], m33 = m[8];
  double d = m11 * (m22 * m33 - m23 * m32) - m12 * (m21 * m33 - m23 * m31) + m13 * (m21 * m32 - m22 * m31);
  inv[0] = (m22 * m33 - m32 * m23) / d;
  inv[1] = -(m21 * m33 - m31 * m23) / d;
  inv[2] = (m21 * m32 - m31 * m22) / d;
  inv[3] = -(m12 * m33 - m32 * m13) / d;
  inv[4] = (m11 * m33 - m31 * m13) / d;
  inv[5] = -(m11 * m32 - m31 * m12) / d;
  inv[6] = (m12 * m23 - m22 * m13) / d;
  inv[7] = -(m11 * m23 - m21 * m13) / d;
  inv[8] = (m11 * m22 - m21 * m12) / d;
}

double det3(const double* m)
{
  const double m11 = m[0], m21 = m[1], m31 = m[2];
  const double m12 = m[3], m22 = m[4], m32 = m[5];
  const double m13 = m[6], m23 = m[7], m33 = m[8];
  return m11 * (m22 * m33 - m23 * m32) - m12 * (m21 * m33 - m23 * m31) + m13 * (m21 * m32 - m22 * m31);
}

// src/raspakit/simulationbox.cpp:137-148
void widths3(const double* m, double* w)
{
  auto cross = [](const double* x, const double* y, double* o)
  {
    o[0] = x[1] * y[2] - x[2] * y[1];
    o[1] = x[2] * y[0] - x[0] * y[2];
    o[2] = x[0] * y[1] - x[1] * y[0];
  };
  double c1[3], c2[3], c3[3];
  cross(m + 0, m + 3, c1);
  cross(m + 3, m + 6, c2);
  cross(m + 0, m + 6, c3);
  double v = std::fabs(m[0] * c2[0] + m[1] * c2[1] + m[2] * c2[2]);
  auto n2 = [](const double* x) { return x[0] * x[0] + x[1] * x[1] + x[2] * x[2]; };
  w[0] = v / std::sqrt(n2(c2));
  w[1] = v / std::sqrt(n2(c3));
  w[2] = v / std::sqrt(n2(c1));
}

}  // namespace rg

extern "C"
{

const char* rg_version(void) { return "raspa3_amd 0.1 (gfx950)"; }

rg_status rg_ctx_create(int device, rg_ctx** out)
{
  if (!out) return RG_ERR_INVALID;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return RG_ERR_DEVICE;
  if (device < 0 || device >= count) return RG_ERR_INVALID;
  rg_ctx* c = new rg_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
  {
    delete c;
    return RG_ERR_DEVICE;
  }
  *out = c;
  return RG_OK;
}

rg_status rg_ctx_destroy(rg_ctx* ctx)
{
  if (!ctx) return RG_ERR_INVALID;
  rg_status st = rgReleaseStatus(hipSetDevice(ctx->device), RG_OK);
  if (ctx->stream)
  {
    st = rgReleaseStatus(hipStreamSynchronize(ctx->stream), st);
    st = rgReleaseStatus(hipStreamDestroy(ctx->stream), st);
  }
  if (ctx->d_table) st = rgReleaseStatus(hipFree(ctx->d_table), st);
  delete ctx;
  return st;
}

const char* rg_last_error(const rg_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

rg_status rg_sync(rg_ctx* ctx)
{
  if (!ctx) return RG_ERR_INVALID;
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

void* rg_stream(rg_ctx* ctx) { return ctx ? reinterpret_cast<void*>(ctx->stream) : nullptr; }

rg_status rg_box(double a, double b, double c, double alpha_deg, double beta_deg, double gamma_deg, int lengths_only,
                 double* cell9, double* inv9, int* is_rectangular, double* volume, double* widths3)
{
  if (!cell9 || !(a > 0.0) || !(b > 0.0) || !(c > 0.0)) return RG_ERR_INVALID;
  double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  int rect = 1;
  if (lengths_only)
  {
    m[0] = a;
    m[4] = b;
    m[8] = c;
  }
  else
  {
    const double al = M_PI * alpha_deg / 180.0, be = M_PI * beta_deg / 180.0, ga = M_PI * gamma_deg / 180.0;
    rect = (std::fabs(al - 0.5 * M_PI) < 1e-5) && (std::fabs(be - 0.5 * M_PI) < 1e-5) &&
           (std::fabs(ga - 0.5 * M_PI) < 1e-5);
    const double temp = (std::cos(al) - std::cos(ga) * std::cos(be)) / std::sin(ga);
    m[0] = a;
    m[3] = b * std::cos(ga);
    m[4] = b * std::sin(ga);
    m[6] = c * std::cos(be);
    m[7] = c * temp;
    m[8] = c * std::sqrt(1.0 - std::cos(be) * std::cos(be) - temp * temp);
  }
  std::memcpy(cell9, m, sizeof(m));
  if (inv9) rg::invert3(m, inv9);
  if (is_rectangular) *is_rectangular = rect;
  if (volume) *volume = rg::det3(m);
  if (widths3) rg::widths3(m, widths3);
  return RG_OK;
}

rg_status rg_ewald_parameters(double precision, double rc_coulomb, const double* widths3, double* alpha, int* kmax3)
{
  if (!widths3 || !alpha || !kmax3 || !(rc_coulomb > 0.0)) return RG_ERR_INVALID;
  const double eps = std::min(std::fabs(precision), 0.5);
  const double tol = std::sqrt(std::abs(std::log(eps * rc_coulomb)));
  *alpha = std::sqrt(std::abs(std::log(eps * rc_coulomb * tol))) / rc_coulomb;
  const double tol1 = std::sqrt(-std::log(eps * rc_coulomb * (2.0 * tol * *alpha) * (2.0 * tol * *alpha)));
  for (int a = 0; a < 3; ++a) kmax3[a] = static_cast<int>(std::rint(0.25 + widths3[a] * *alpha * tol1 / M_PI));
  return RG_OK;
}

rg_status rg_grid_dims(const double* cell9, double spacing, int* n3)
{
  if (!cell9 || !n3 || !(spacing > 0.0)) return RG_ERR_INVALID;
  double w[3];
  rg::widths3(cell9, w);
  for (int a = 0; a < 3; ++a) n3[a] = static_cast<int>(static_cast<size_t>(w[a] / spacing + 0.5));
  return RG_OK;
}

rg_status rg_forcefield_create(rg_ctx* ctx, int n, const double* eps_K, const double* sigma, const int* is_framework,
                               int shifted, double rc_framework_vdw, double rc_molecule_vdw, double rc_coulomb,
                               double alpha, int use_charge, double overlap, rg_forcefield** out)
{
  if (!out || n <= 0 || n > 4096 || !eps_K || !sigma || !is_framework)
    return rg::fail(ctx, RG_ERR_INVALID, "rg_forcefield_create: invalid arguments");
  rg_forcefield* ff = new rg_forcefield();
  ff->ctx = ctx;
  ff->n = n;
  ff->eps.assign(size_t(n) * n, 0.0);
  ff->sigma.assign(size_t(n) * n, 0.0);
  ff->shift.assign(size_t(n) * n, 0.0);
  ff->lj.assign(size_t(n) * n, 1);
  ff->is_framework.assign(is_framework, is_framework + n);
  ff->rc_framework = rc_framework_vdw;
  ff->rc_molecule = rc_molecule_vdw;
  ff->rc_coulomb = rc_coulomb;
  ff->alpha = alpha;
  ff->use_charge = use_charge;
  ff->overlap = overlap;
  const double k2e = rg::units().kelvinToEnergy;
  // VDWParameters(epsilon, sigma) stores epsilon in internal units (src/raspakit/forcefield/vdwparameters.ixx:113-118)
  for (int i = 0; i < n; ++i)
  {
    ff->eps[size_t(i) * n + i] = eps_K[i] * k2e;
    ff->sigma[size_t(i) * n + i] = sigma[i];
  }
  // Lorentz-Berthelot (src/raspakit/forcefield/forcefield.cpp:924-960)
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
    {
      const double e = std::sqrt(ff->eps[size_t(i) * n + i] * ff->eps[size_t(j) * n + j]);
      const double s = 0.5 * (ff->sigma[size_t(i) * n + i] + ff->sigma[size_t(j) * n + j]);
      ff->eps[size_t(i) * n + j] = ff->eps[size_t(j) * n + i] = e;
      ff->sigma[size_t(i) * n + j] = ff->sigma[size_t(j) * n + i] = s;
    }
  // pairs with a zero self-epsilon or self-sigma become type None (:995-1013)
  for (int i = 0; i < n; ++i)
    for (int j = i; j < n; ++j)
    {
      const bool zi = ff->eps[size_t(i) * n + i] == 0.0 || ff->sigma[size_t(i) * n + i] == 0.0;
      const bool zj = ff->eps[size_t(j) * n + j] == 0.0 || ff->sigma[size_t(j) * n + j] == 0.0;
      if (zi || zj) ff->lj[size_t(i) * n + j] = ff->lj[size_t(j) * n + i] = 0;
    }
  // shift = U_LJ(rc) with the framework cutoff if either pseudo-atom is a framework atom (:1016-1051,
  // src/raspakit/forcefield/vdwparameters.cpp:211-223)
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
    {
      const size_t k = size_t(i) * n + j;
      if (shifted && ff->lj[k])
      {
        const double rc = (is_framework[i] || is_framework[j]) ? rc_framework_vdw : rc_molecule_vdw;
        const double temp = (rc * rc) / (ff->sigma[k] * ff->sigma[k]);
        const double rri3 = 1.0 / (temp * temp * temp);
        ff->shift[k] = 4.0 * ff->eps[k] * (rri3 * (rri3 - 1.0));
      }
    }
  *out = ff;
  return RG_OK;
}

rg_status rg_forcefield_destroy(rg_forcefield* ff)
{
  if (!ff) return RG_ERR_INVALID;
  delete ff;
  return RG_OK;
}

rg_status rg_forcefield_pairs(const rg_forcefield* ff, double* eps, double* sigma, double* shift, int* lj)
{
  if (!ff) return RG_ERR_INVALID;
  const size_t nn = size_t(ff->n) * ff->n;
  if (eps) std::memcpy(eps, ff->eps.data(), nn * sizeof(double));
  if (sigma) std::memcpy(sigma, ff->sigma.data(), nn * sizeof(double));
  if (shift) std::memcpy(shift, ff->shift.data(), nn * sizeof(double));
  if (lj) std::memcpy(lj, ff->lj.data(), nn * sizeof(int));
  return RG_OK;
}

rg_status rg_framework_create(rg_ctx* ctx, const double* cell9, int is_rectangular, int n_atoms, const double* pos,
                              const int* type, const double* charge, rg_framework** out)
{
  if (!ctx || !out || !cell9 || n_atoms < 0 || (n_atoms > 0 && (!pos || !type || !charge)))
    return rg::fail(ctx, RG_ERR_INVALID, "rg_framework_create: invalid arguments");
  rg_framework* fw = new rg_framework();
  fw->ctx = ctx;
  fw->device = ctx->device;
  std::memcpy(fw->cell, cell9, 9 * sizeof(double));
  rg::invert3(fw->cell, fw->inv);
  fw->rect = is_rectangular;
  fw->n = n_atoms;
  fw->pos.assign(pos, pos + 3 * size_t(n_atoms));
  fw->type.assign(type, type + n_atoms);
  fw->charge.assign(charge, charge + n_atoms);
  rg::widths3(fw->cell, fw->widths);
  *out = fw;
  return RG_OK;
}

rg_status rg_framework_destroy(rg_framework* fw)
{
  if (!fw) return RG_ERR_INVALID;
  rg_status st = rgReleaseStatus(hipSetDevice(fw->device), RG_OK);
  if (fw->images.d_img) st = rgReleaseStatus(hipFree(fw->images.d_img), st);
  if (fw->images.d_bin_start) st = rgReleaseStatus(hipFree(fw->images.d_bin_start), st);
  delete fw;
  return st;
}

}  // extern "C"

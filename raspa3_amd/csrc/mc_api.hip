// Host side of the rg_system object and the operator entry points of the C ABI (trial energies, Rosenbluth
// selection, Ewald difference / accept / total).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "mc_internal.hpp"

namespace rg
{
__global__ void k_trial_energies(const SysDev* Sp, int n, const double* pos, const int* type, const double* q,
                                 const int* mol, int skipMol, double* out, int* overlap);
__global__ void k_rosenbluth(int nsets, int k, const double* U, const int* ov, double beta, const double* uniforms,
                             int* selected, double* logW);
__global__ void k_ewald_delta(const SysDev* Sp, int nn, const double* nw, const int* nmol, int no, const double* od,
                              const int* omol, double* out);
__global__ void k_copy_sk(double2* dst, const double2* src, int nk);
__global__ void k_structure_factor(const SysDev* Sp, int n, const double* x, const double* y, const double* z,
                                  const double* q, double2* out, int accumulate);
__global__ void k_fourier_energy(const SysDev* Sp, double* out);

rg_status syncSystem(rg_system* s)
{
  RG_HIP(s->ctx, hipMemcpyAsync(s->d, &s->h, sizeof(SysDev), hipMemcpyHostToDevice, s->ctx->stream));
  return RG_OK;
}

namespace
{

template <typename T>
rg_status devAlloc(rg_system* s, T*& p, size_t count)
{
  void* v = nullptr;
  RG_HIP(s->ctx, hipMalloc(&v, std::max<size_t>(1, count) * sizeof(T)));
  s->owned.push_back(v);
  p = static_cast<T*>(v);
  return RG_OK;
}

template <typename T>
rg_status devUpload(rg_system* s, T*& p, const T* host, size_t count)
{
  rg_status st = devAlloc(s, p, count);
  if (st != RG_OK) return st;
  if (count) RG_HIP(s->ctx, hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s->ctx->stream));
  return RG_OK;
}

// RAII scratch on the device for host-array entry points
struct Scratch
{
  rg_ctx* ctx;
  void* p = nullptr;
  ~Scratch()
  {
    if (p)
    {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipFree(p);
    }
  }
};

}  // namespace

}  // namespace rg

using namespace rg;

extern "C"
{

rg_status rg_system_create(rg_ctx* ctx, const rg_forcefield* ff, const double* cell9, int is_rectangular,
                           const rg_framework* fw, int n_grids, const int* grid_types, rg_grid* const* grids,
                           const rg_grid* coulomb_grid, const int* kmax3, long long recip_int_cut2,
                           int capacity_atoms, rg_system** out)
{
  if (!ctx || !ff || !cell9 || !out || !kmax3 || capacity_atoms < 0 || n_grids < 0 ||
      (n_grids > 0 && (!grid_types || !grids)))
    return fail(ctx, RG_ERR_INVALID, "rg_system_create: invalid arguments");
  for (int a = 0; a < 3; ++a)
    if (kmax3[a] < 0 || kmax3[a] > EW_MAXK)
      return fail(ctx, RG_ERR_UNSUPPORTED, "rg_system_create: kmax must lie in [0, %d]", EW_MAXK);
  RG_ENTER(ctx);
  rg_system* s = new rg_system();
  s->ctx = ctx;
  s->device = ctx->device;
  s->ff = ff;
  SysDev& h = s->h;
  std::memcpy(h.box.cell, cell9, sizeof(h.box.cell));
  invert3(h.box.cell, h.box.inv);
  h.box.rect = is_rectangular;
  h.box.volume = det3(h.box.cell);
  h.alpha = ff->alpha;
  h.coulomb = units().coulomb;
  h.overlap = ff->overlap;
  h.rcv2Fw = ff->rc_framework * ff->rc_framework;
  h.rcv2Mol = ff->rc_molecule * ff->rc_molecule;
  h.rcc2 = ff->rc_coulomb * ff->rc_coulomb;
  h.useCharge = ff->use_charge;
  h.ntypes = ff->n;
  rg_status st = RG_OK;
  auto bail = [&](rg_status e)
  {
    rg_system_destroy(s);
    return e;
  };
  {
    std::vector<rgd::Pair> pairs(size_t(ff->n) * ff->n);
    for (size_t k = 0; k < pairs.size(); ++k)
      pairs[k] = {4.0 * ff->eps[k], ff->sigma[k] * ff->sigma[k], ff->shift[k], double(ff->lj[k])};
    const rgd::Pair* dp;
    rgd::Pair* tmp;
    if ((st = devUpload(s, tmp, pairs.data(), pairs.size())) != RG_OK) return bail(st);
    dp = tmp;
    h.pairs = dp;
  }
  // framework atoms (direct path), sorted by cell when the box holds at least three cells of one cutoff width per
  // axis (perpendicular widths); the sums then visit only the 27 cells around a site
  if (fw)
  {
    const int n = fw->n;
    for (int t : fw->type)
      if (t < 0 || t >= ff->n) return bail(fail(ctx, RG_ERR_INVALID, "rg_system_create: framework type out of range"));
    const double rc = std::max(ff->rc_framework, ff->use_charge ? ff->rc_coulomb : 0.0);
    int nc[3];
    bool cells = rc > 0.0;
    for (int ax = 0; ax < 3; ++ax)
    {
      const double* v = h.box.inv;
      const double width = 1.0 / std::sqrt(v[ax] * v[ax] + v[3 + ax] * v[3 + ax] + v[6 + ax] * v[6 + ax]);
      // at most 128 cells per axis (2M cells); cells wider than the cutoff only visit more atoms, never fewer
      nc[ax] = int(std::min(128.0, std::floor(width / std::max(rc, 1e-12))));
      cells = cells && nc[ax] >= 3;
    }
    std::vector<int> cellOf(n, 0), start;
    if (cells)
    {
      for (int a = 0; a < n; ++a)
      {
        int c[3];
        for (int ax = 0; ax < 3; ++ax)
        {
          double sv = h.box.inv[ax] * fw->pos[3 * a] + h.box.inv[3 + ax] * fw->pos[3 * a + 1] +
                      h.box.inv[6 + ax] * fw->pos[3 * a + 2];
          sv -= std::floor(sv);
          c[ax] = std::min(nc[ax] - 1, int(sv * nc[ax]));
        }
        cellOf[a] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
      }
      start.assign(size_t(nc[0]) * nc[1] * nc[2] + 1, 0);
      for (int a = 0; a < n; ++a) start[cellOf[a] + 1] += 1;
      for (size_t c = 1; c < start.size(); ++c) start[c] += start[c - 1];
    }
    // atom order: by cell (stable), or as given
    std::vector<int> order(n);
    {
      std::vector<int> fill(start.begin(), start.end());
      for (int a = 0; a < n; ++a) order[cells ? fill[cellOf[a]]++ : a] = a;
    }
    std::vector<double> x(n), y(n), z(n), q(n);
    std::vector<int> type(n);
    double net = 0.0;
    for (int i = 0; i < n; ++i)
    {
      const int a = order[i];
      x[i] = fw->pos[3 * a];
      y[i] = fw->pos[3 * a + 1];
      z[i] = fw->pos[3 * a + 2];
      q[i] = fw->charge[a];
      type[i] = fw->type[a];
      net += fw->charge[a];
    }
    double *dx, *dy, *dz, *dq;
    int* dt;
    if ((st = devUpload(s, dx, x.data(), n)) != RG_OK || (st = devUpload(s, dy, y.data(), n)) != RG_OK ||
        (st = devUpload(s, dz, z.data(), n)) != RG_OK || (st = devUpload(s, dq, q.data(), n)) != RG_OK ||
        (st = devUpload(s, dt, type.data(), n)) != RG_OK)
      return bail(st);
    h.cellStart = nullptr;
    if (cells)
    {
      int* dstart;
      if ((st = devUpload(s, dstart, start.data(), start.size())) != RG_OK) return bail(st);
      h.cellStart = dstart;
      for (int ax = 0; ax < 3; ++ax) h.ncell[ax] = nc[ax];
    }
    h.nfw = n;
    h.fwx = dx;
    h.fwy = dy;
    h.fwz = dz;
    h.fwq = dq;
    h.fwt = dt;
    s->fwNetCharge = net;
  }
  // grids
  h.coulGrid = -1;
  h.gridOfType = nullptr;
  if (n_grids > 0 || coulomb_grid)
  {
    std::vector<int> got(ff->n, -1);
    auto view = [](const rg_grid* g)
    {
      rgd::GridView v{};
      v.data = g->d;
      std::memcpy(v.inv, g->inv, sizeof(v.inv));
      for (int a = 0; a < 3; ++a)
      {
        v.n[a] = g->n[a];
        v.ncd[a] = double(g->n[a] - 1);
      }
      v.values = g->values;
      return v;
    };
    for (int g = 0; g < n_grids; ++g)
    {
      if (!grids[g] || grid_types[g] < 0 || grid_types[g] >= ff->n)
        return bail(fail(ctx, RG_ERR_INVALID, "rg_system_create: grid %d invalid", g));
      got[grid_types[g]] = static_cast<int>(s->gridViews.size());
      s->gridViews.push_back(view(grids[g]));
    }
    if (coulomb_grid)
    {
      h.coulGrid = static_cast<int>(s->gridViews.size());
      s->gridViews.push_back(view(coulomb_grid));
    }
    int* dg;
    rgd::GridView* dv;
    if ((st = devUpload(s, dg, got.data(), got.size())) != RG_OK ||
        (st = devUpload(s, dv, s->gridViews.data(), s->gridViews.size())) != RG_OK)
      return bail(st);
    h.gridOfType = dg;
    h.grids = dv;
    s->gridOfType = got;
  }
  // adsorbates
  h.cap = capacity_atoms;
  h.nads = 0;
  if ((st = devAlloc(s, h.ax, capacity_atoms)) != RG_OK || (st = devAlloc(s, h.ay, capacity_atoms)) != RG_OK ||
      (st = devAlloc(s, h.az, capacity_atoms)) != RG_OK || (st = devAlloc(s, h.aq, capacity_atoms)) != RG_OK ||
      (st = devAlloc(s, h.at, capacity_atoms)) != RG_OK || (st = devAlloc(s, h.amol, capacity_atoms)) != RG_OK)
    return bail(st);
  // Ewald k-vectors in the reference's nvec order (src/raspakit/interactions/ewald.cpp:1093-1120)
  {
    std::vector<int> code;
    std::vector<double> fk;
    const double* inv = h.box.inv;
    const double ax[3] = {inv[0], inv[3], inv[6]}, ay[3] = {inv[1], inv[4], inv[7]}, az[3] = {inv[2], inv[5], inv[8]};
    const double prefactor = h.coulomb * (2.0 * M_PI / h.box.volume);
    const long long icut = recip_int_cut2 > 0 ? recip_int_cut2 : (1ll << 62);
    double single = 0.0;
    for (int kx = 0; kx <= kmax3[0]; ++kx)
    {
      const double factor = (kx == 0) ? prefactor : 2.0 * prefactor;
      for (int ky = -kmax3[1]; ky <= kmax3[1]; ++ky)
        for (int kz = -kmax3[2]; kz <= kmax3[2]; ++kz)
        {
          double kv[3];
          for (int c = 0; c < 3; ++c)
            kv[c] = 2.0 * M_PI * kx * ax[c] + 2.0 * M_PI * ky * ay[c] + 2.0 * M_PI * kz * az[c];
          const double rksq = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
          const long long ksq = (long long)kx * kx + (long long)ky * ky + (long long)kz * kz;
          if (ksq != 0 && ksq <= icut)
          {
            code.push_back(kx | ((ky + 128) << 8) | ((kz + 128) << 16));
            const double f = factor * std::exp((-0.25 / (h.alpha * h.alpha)) * rksq) / rksq;
            fk.push_back(f);
            single += f;
          }
        }
    }
    h.nk = static_cast<int>(code.size());
    for (int a = 0; a < 3; ++a) h.kmax[a] = kmax3[a];
    h.singleIon = single;
    s->intCut2 = recip_int_cut2;
    int* dcode;
    double* dfk;
    double2 *dS, *dSt, *dSfw;
    if ((st = devUpload(s, dcode, code.data(), code.size())) != RG_OK ||
        (st = devUpload(s, dfk, fk.data(), fk.size())) != RG_OK || (st = devAlloc(s, dS, code.size())) != RG_OK ||
        (st = devAlloc(s, dSt, code.size())) != RG_OK || (st = devAlloc(s, dSfw, code.size())) != RG_OK)
      return bail(st);
    h.kcode = dcode;
    h.fk = dfk;
    h.Sk = dS;
    h.St = dSt;
    h.Sfw = dSfw;
  }
  h.netCharge = s->fwNetCharge;
  {
    SysDev* d;
    if ((st = devAlloc(s, d, 1)) != RG_OK) return bail(st);
    s->d = d;
  }
  if ((st = syncSystem(s)) != RG_OK) return bail(st);
  // framework structure factor (the rigid part, src/raspakit/interactions/ewald.cpp:276-348)
  if (h.nk > 0)
  {
    double2* sfw = const_cast<double2*>(h.Sfw);
    if (h.nfw > 0)
      hipLaunchKernelGGL(k_structure_factor, dim3((h.nk + 255) / 256), dim3(256), 0, ctx->stream, s->d, h.nfw, h.fwx,
                         h.fwy, h.fwz, h.fwq, sfw, 0);
    else
      hipMemsetAsync(sfw, 0, h.nk * sizeof(double2), ctx->stream);
    hipLaunchKernelGGL(k_copy_sk, dim3((h.nk + 255) / 256), dim3(256), 0, ctx->stream, h.Sk, h.Sfw, h.nk);
  }
  hipError_t e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return bail(hipfail(ctx, e, "rg_system_create"));
  *out = s;
  return RG_OK;
}

rg_status rg_system_destroy(rg_system* s)
{
  if (!s) return RG_ERR_INVALID;
  rg_status st = rgReleaseStatus(hipSetDevice(s->device), RG_OK);
  // the context (and its stream) may already be gone: wait for the whole device instead
  st = rgReleaseStatus(hipDeviceSynchronize(), st);
  for (void* p : s->owned) st = rgReleaseStatus(hipFree(p), st);
  delete s;
  return st;
}

rg_status rg_system_info(const rg_system* s, int* nk, double* single_ion, double* net_charge, int* n_adsorbate_atoms)
{
  if (!s) return RG_ERR_INVALID;
  if (nk) *nk = s->h.nk;
  if (single_ion) *single_ion = s->h.singleIon;
  if (net_charge) *net_charge = s->h.netCharge;
  if (n_adsorbate_atoms) *n_adsorbate_atoms = s->h.nads;
  return RG_OK;
}

rg_status rg_system_set_adsorbates(rg_ctx* ctx, rg_system* s, int n, const double* pos, const int* type,
                                   const double* q, const int* molecule)
{
  if (!ctx || !s || n < 0 || n > s->h.cap || (n > 0 && (!pos || !type || !q || !molecule)))
    return fail(ctx, RG_ERR_INVALID, "rg_system_set_adsorbates: invalid arguments (capacity %d)", s ? s->h.cap : 0);
  RG_ENTER(ctx);
  std::vector<double> x(n), y(n), z(n);
  double net = s->fwNetCharge;
  for (int a = 0; a < n; ++a)
  {
    x[a] = pos[3 * a];
    y[a] = pos[3 * a + 1];
    z[a] = pos[3 * a + 2];
    net += q[a];
    if (type[a] < 0 || type[a] >= s->h.ntypes) return fail(ctx, RG_ERR_INVALID, "adsorbate type out of range");
  }
  SysDev& h = s->h;
  if (n)
  {
    RG_HIP(ctx, hipMemcpyAsync(h.ax, x.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(h.ay, y.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(h.az, z.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(h.aq, q, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(h.at, type, n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(h.amol, molecule, n * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
  }
  h.nads = n;
  h.netCharge = net;
  rg_status st = syncSystem(s);
  if (st != RG_OK) return st;
  if (h.nk > 0)
  {
    hipLaunchKernelGGL(k_copy_sk, dim3((h.nk + 255) / 256), dim3(256), 0, ctx->stream, h.Sk, h.Sfw, h.nk);
    if (n > 0)
      hipLaunchKernelGGL(k_structure_factor, dim3((h.nk + 255) / 256), dim3(256), 0, ctx->stream, s->d, n, h.ax,
                         h.ay, h.az, h.aq, h.Sk, 1);
  }
  RG_HIP(ctx, hipGetLastError());
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

}  // extern "C"

namespace
{

// Bytes of device staging per trial site: position (3 doubles), charge, the four energies, type, molecule, overlap.
constexpr size_t kTrialStageBytes = 3 * 8 + 8 + 4 * 8 + 4 + 4 + 4;

// Stages the host inputs of n trial sites into 'stage' (n * kTrialStageBytes bytes of device memory), evaluates their
// energies and copies the results back; returns when they are on the host.
rg_status trialEnergiesStaged(rg_ctx* ctx, const rg_system* s, int n, const double* pos, const int* type,
                              const double* q, const int* molecule, int skip_molecule, double* energies, int* overlap,
                              void* stage)
{
  char* p = static_cast<char*>(stage);
  double* dpos = reinterpret_cast<double*>(p);
  double* dq = dpos + 3 * size_t(n);
  double* dout = dq + n;
  int* dtype = reinterpret_cast<int*>(dout + 4 * size_t(n));
  int* dmol = dtype + n;
  int* dov = dmol + n;
  RG_HIP(ctx, hipMemcpyAsync(dpos, pos, 3 * size_t(n) * 8, hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(dq, q, size_t(n) * 8, hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(dtype, type, size_t(n) * 4, hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(dmol, molecule, size_t(n) * 4, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_trial_energies, dim3((n + 3) / 4), dim3(256), 0, ctx->stream, s->d, n, dpos, dtype, dq, dmol,
                     skip_molecule, dout, dov);
  RG_HIP(ctx, hipGetLastError());
  RG_HIP(ctx, hipMemcpyAsync(energies, dout, 4 * size_t(n) * 8, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(overlap, dov, size_t(n) * 4, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

rg_status trialArgs(rg_ctx* ctx, const rg_system* s, int n, const double* pos, const int* type, const double* q,
                    const int* molecule, double* energies, int* overlap, const char* who)
{
  if (!ctx || !s || n < 0 || (n > 0 && (!pos || !type || !q || !molecule || !energies || !overlap)))
    return fail(ctx, RG_ERR_INVALID, "%s: invalid arguments", who);
  for (int i = 0; i < n; ++i)
    if (type[i] < 0 || type[i] >= s->h.ntypes) return fail(ctx, RG_ERR_INVALID, "%s: type out of range", who);
  return RG_OK;
}

// Staging of rg_trial_energies_noalloc: one array per device, held until the process ends.  Host code never uses the
// name 'scratch' as an address (on the host it is only a handle for the symbol lookup); it uses the device address
// that hipGetSymbolAddress returns for the current device, looked up once per device.
static __device__ __align__(256) char scratch[RG_TRIAL_NOALLOC_MAX_SITES * kTrialStageBytes];

std::mutex gScratchMutex;
std::vector<void*> gScratchPtr;  // per device: device address of 'scratch', or null before the first lookup

rg_status scratchAddress(rg_ctx* ctx, void** out)
{
  std::lock_guard<std::mutex> lock(gScratchMutex);
  if (ctx->device >= int(gScratchPtr.size())) gScratchPtr.resize(ctx->device + 1, nullptr);
  if (!gScratchPtr[ctx->device])
  {
    void* ptr = nullptr;
    RG_HIP(ctx, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(ctx, RG_ERR_DEVICE, "rg_trial_energies_noalloc: no device address for the staging array");
    gScratchPtr[ctx->device] = ptr;
  }
  *out = gScratchPtr[ctx->device];
  return RG_OK;
}

}  // namespace

extern "C"
{

rg_status rg_trial_energies(rg_ctx* ctx, const rg_system* s, int n, const double* pos, const int* type,
                            const double* q, const int* molecule, int skip_molecule, double* energies, int* overlap)
{
  rg_status st = trialArgs(ctx, s, n, pos, type, q, molecule, energies, overlap, "rg_trial_energies");
  if (st != RG_OK || n == 0) return st;
  RG_ENTER(ctx);
  Scratch sc{ctx};
  RG_HIP(ctx, hipMalloc(&sc.p, size_t(n) * kTrialStageBytes));
  return trialEnergiesStaged(ctx, s, n, pos, type, q, molecule, skip_molecule, energies, overlap, sc.p);
}

rg_status rg_trial_energies_noalloc(rg_ctx* ctx, const rg_system* s, int n, const double* pos, const int* type,
                                    const double* q, const int* molecule, int skip_molecule, double* energies,
                                    int* overlap)
{
  rg_status st = trialArgs(ctx, s, n, pos, type, q, molecule, energies, overlap, "rg_trial_energies_noalloc");
  if (st != RG_OK) return st;
  if (n > RG_TRIAL_NOALLOC_MAX_SITES)
    return fail(ctx, RG_ERR_INVALID, "rg_trial_energies_noalloc: %d sites exceed the staging array's %d", n,
                RG_TRIAL_NOALLOC_MAX_SITES);
  if (n == 0) return RG_OK;
  RG_ENTER(ctx);
  void* stage = nullptr;
  st = scratchAddress(ctx, &stage);
  if (st != RG_OK) return st;
  return trialEnergiesStaged(ctx, s, n, pos, type, q, molecule, skip_molecule, energies, overlap, stage);
}

rg_status rg_rosenbluth_select(rg_ctx* ctx, int n_sets, int k, const double* energies, const int* overlap,
                               double beta, const double* uniforms, int* selected, double* log_weight)
{
  if (!ctx || n_sets < 0 || k <= 0 || (n_sets > 0 && (!energies || !overlap || !uniforms || !selected || !log_weight)))
    return fail(ctx, RG_ERR_INVALID, "rg_rosenbluth_select: invalid arguments");
  if (n_sets == 0) return RG_OK;
  RG_ENTER(ctx);
  Scratch sc{ctx};
  const size_t m = size_t(n_sets) * k;
  RG_HIP(ctx, hipMalloc(&sc.p, m * 12 + size_t(n_sets) * 20));
  double* dU = static_cast<double*>(sc.p);
  double* du = dU + m;
  double* dlw = du + n_sets;
  int* dov = reinterpret_cast<int*>(dlw + n_sets);
  int* dsel = dov + m;
  RG_HIP(ctx, hipMemcpyAsync(dU, energies, m * 8, hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(dov, overlap, m * 4, hipMemcpyHostToDevice, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(du, uniforms, size_t(n_sets) * 8, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_rosenbluth, dim3((n_sets + 255) / 256), dim3(256), 0, ctx->stream, n_sets, k, dU, dov, beta,
                     du, dsel, dlw);
  RG_HIP(ctx, hipGetLastError());
  RG_HIP(ctx, hipMemcpyAsync(selected, dsel, size_t(n_sets) * 4, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipMemcpyAsync(log_weight, dlw, size_t(n_sets) * 8, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

rg_status rg_ewald_delta(rg_ctx* ctx, rg_system* s, int n_new, const double* new_pos, const double* new_q,
                         const int* new_mol, int n_old, const double* old_pos, const double* old_q, const int* old_mol,
                         double* out)
{
  if (!ctx || !s || !out || n_new < 0 || n_old < 0 || n_new + n_old > EW_MAXATOMS ||
      (n_new > 0 && (!new_pos || !new_q || !new_mol)) || (n_old > 0 && (!old_pos || !old_q || !old_mol)))
    return fail(ctx, RG_ERR_INVALID, "rg_ewald_delta: invalid arguments (at most %d atoms)", EW_MAXATOMS);
  RG_ENTER(ctx);
  std::vector<double> at(4 * size_t(n_new + n_old));
  std::vector<int> mol(n_new + n_old);
  for (int a = 0; a < n_new; ++a)
  {
    for (int c = 0; c < 3; ++c) at[4 * a + c] = new_pos[3 * a + c];
    at[4 * a + 3] = new_q[a];
    mol[a] = new_mol[a];
  }
  for (int a = 0; a < n_old; ++a)
  {
    for (int c = 0; c < 3; ++c) at[4 * (n_new + a) + c] = old_pos[3 * a + c];
    at[4 * (n_new + a) + 3] = old_q[a];
    mol[n_new + a] = old_mol[a];
  }
  Scratch sc{ctx};
  RG_HIP(ctx, hipMalloc(&sc.p, at.size() * 8 + mol.size() * 4 + 4 * 8 + 64));
  double* dat = static_cast<double*>(sc.p);
  double* dout = dat + at.size();
  int* dmol = reinterpret_cast<int*>(dout + 4);
  if (!at.empty()) RG_HIP(ctx, hipMemcpyAsync(dat, at.data(), at.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  if (!mol.empty()) RG_HIP(ctx, hipMemcpyAsync(dmol, mol.data(), mol.size() * 4, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_ewald_delta, dim3(1), dim3(256), 0, ctx->stream, s->d, n_new, dat, dmol, n_old,
                     dat + 4 * size_t(n_new), dmol + n_new, dout);
  RG_HIP(ctx, hipGetLastError());
  RG_HIP(ctx, hipMemcpyAsync(out, dout, 3 * 8, hipMemcpyDeviceToHost, ctx->stream));
  double dq = 0.0;
  RG_HIP(ctx, hipMemcpyAsync(&dq, dout + 3, 8, hipMemcpyDeviceToHost, ctx->stream));
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  s->pendingCharge = dq;
  return RG_OK;
}

rg_status rg_ewald_accept(rg_ctx* ctx, rg_system* s)
{
  if (!ctx || !s) return fail(ctx, RG_ERR_INVALID, "rg_ewald_accept: invalid arguments");
  RG_ENTER(ctx);
  if (s->h.nk > 0)
    hipLaunchKernelGGL(k_copy_sk, dim3((s->h.nk + 255) / 256), dim3(256), 0, ctx->stream, s->h.Sk, s->h.St, s->h.nk);
  RG_HIP(ctx, hipGetLastError());
  s->h.netCharge += s->pendingCharge;
  s->pendingCharge = 0.0;
  rg_status st = syncSystem(s);
  if (st != RG_OK) return st;
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RG_OK;
}

rg_status rg_ewald_energy(rg_ctx* ctx, const rg_system* s, double* out)
{
  if (!ctx || !s || !out) return fail(ctx, RG_ERR_INVALID, "rg_ewald_energy: invalid arguments");
  RG_ENTER(ctx);
  const SysDev& h = s->h;
  const int nb = 64;
  Scratch sc{ctx};
  RG_HIP(ctx, hipMalloc(&sc.p, nb * 8));
  double* dpart = static_cast<double*>(sc.p);
  hipLaunchKernelGGL(k_fourier_energy, dim3(nb), dim3(256), 0, ctx->stream, s->d, dpart);
  RG_HIP(ctx, hipGetLastError());
  std::vector<double> part(nb);
  const int n = h.nads;
  std::vector<double> x(n), y(n), z(n), q(n);
  std::vector<int> mol(n);
  RG_HIP(ctx, hipMemcpyAsync(part.data(), dpart, nb * 8, hipMemcpyDeviceToHost, ctx->stream));
  if (n)
  {
    RG_HIP(ctx, hipMemcpyAsync(x.data(), h.ax, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(y.data(), h.ay, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(z.data(), h.az, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(q.data(), h.aq, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    RG_HIP(ctx, hipMemcpyAsync(mol.data(), h.amol, n * 4, hipMemcpyDeviceToHost, ctx->stream));
  }
  RG_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double ef = 0.0;
  for (double v : part) ef += v;
  // self and intra-molecular exclusion (src/raspakit/interactions/ewald.cpp:477-530), net charge (:536-560)
  const double C = h.coulomb, ps = C * h.alpha / std::sqrt(M_PI);
  double self = 0.0, excl = 0.0, netAds = 0.0;
  for (int i = 0; i < n; ++i)
  {
    self -= ps * q[i] * q[i];
    netAds += q[i];
    for (int j = i + 1; j < n && mol[j] == mol[i]; ++j)
    {
      double d[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
      // minimum image with the reference's rounding
      if (h.box.rect)
        for (int c = 0; c < 3; ++c)
        {
          double t = d[c] * h.box.inv[4 * c] + ((d[c] >= 0.0) ? 0.5 : -0.5);
          d[c] -= double((long long)t) * h.box.cell[4 * c];
        }
      else
      {
        double sfr[3];
        for (int c = 0; c < 3; ++c) sfr[c] = h.box.inv[c] * d[0] + h.box.inv[3 + c] * d[1] + h.box.inv[6 + c] * d[2];
        for (int c = 0; c < 3; ++c) sfr[c] -= double((long long)(sfr[c] + ((sfr[c] >= 0.0) ? 0.5 : -0.5)));
        for (int c = 0; c < 3; ++c)
          d[c] = h.box.cell[c] * sfr[0] + h.box.cell[3 + c] * sfr[1] + h.box.cell[6 + c] * sfr[2];
      }
      const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      excl -= C * q[i] * q[j] * std::erf(h.alpha * r) / r;
    }
  }
  const double uIon = -(h.singleIon - C * h.alpha / std::sqrt(M_PI));
  ef += uIon * (2.0 * s->fwNetCharge + netAds) * netAds;
  out[0] = ef;
  out[1] = self;
  out[2] = excl;
  return RG_OK;
}

}  // extern "C"

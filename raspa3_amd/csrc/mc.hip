// Trial energies, Rosenbluth weights and the Ewald Fourier difference on gfx950.
//
// An rg_system is one simulation box as the CBMC moves see it (src/raspakit/cbmc/cbmc_growth_context.ixx:15-29):
// box, pair table, framework atoms for the direct sum or interpolation grids per pseudo-atom type, the adsorbate
// atoms, and the Ewald state (k-vector table, prefactors, stored structure factor).  All of it is device-resident.
//
// Kernels:
//   k_trial_energies   one wavefront per trial site; lanes stride over framework atoms (direct path) or evaluate
//                      the tricubic grids (grid path), then over the adsorbate atoms; wavefront shuffles reduce.
//                      Replaces CBMC::computeFrameworkMoleculeEnergy (src/raspakit/cbmc/interactions_framework_molecule.cpp:31-111)
//                      and CBMC::computeInterMolecularEnergy (src/raspakit/cbmc/interactions_intermolecular.cpp:22-95).
//   k_rosenbluth       one lane per trial set: log-sum-exp weight and selection given a uniform
//                      (src/raspakit/cbmc/multiple_first_bead.cpp:20-49, src/raspakit/cbmc/util.cpp:15-48).
//   k_ewald_delta      one workgroup; lanes stride over k-vectors; FP64 eik recursions per atom in LDS
//                      (src/raspakit/interactions/ewald.cpp:1043-1244, ewald_kvector.ixx:25-131).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mc_device.hpp"
#include "mc_internal.hpp"

namespace rg
{

// ---------------------------------------------------------------------------------------------------------------
// Trial-site energies
// ---------------------------------------------------------------------------------------------------------------

// Energies of one site with the whole wavefront.  Returns in e[0..3] = framework VDW, framework Coulomb,
// inter-molecular VDW, inter-molecular Coulomb; 'overlap' is set if any pair (or the grid) exceeds the criterion.
__device__ void siteEnergyWave(const SysDev& S, double px, double py, double pz, int type, double q, int mol,
                               int skipMol, int lane, double e[4], bool& overlap)
{
  double evf = 0.0, ecf = 0.0, evi = 0.0, eci = 0.0;
  bool ov = false;
  const int gi = S.gridOfType ? S.gridOfType[type] : -1;
  if (gi >= 0)
  {
    // grid path (src/raspakit/cbmc/interactions_framework_molecule.cpp:57-68); every lane computes the same value
    const double ev = rgd::interpolate(S.grids[gi], px, py, pz);
    ov = ev > S.overlap;
    evf = (lane == 0) ? ev : 0.0;
    if (S.useCharge && S.coulGrid >= 0)
      ecf = (lane == 0) ? q * rgd::interpolate(S.grids[S.coulGrid], px, py, pz) : 0.0;
  }
  else
  {
    forNearFramework(S, px, py, pz, lane, [&](int a) {
      double dx = S.fwx[a] - px, dy = S.fwy[a] - py, dz = S.fwz[a] - pz;
      rgd::pbc(S.box, dx, dy, dz);
      const double rr = dx * dx + dy * dy + dz * dz;
      if (rr < S.rcv2Fw)
      {
        const double v = rgd::vdw0(S.pairs[S.fwt[a] * S.ntypes + type], 1.0, rr);
        ov = ov || v > S.overlap;
        evf += v;
      }
      if (S.useCharge && rr < S.rcc2)
      {
        const double r = sqrt(rr);
        ecf += rgd::coul0(S.alpha, 1.0, r, S.coulomb * S.fwq[a] * q);
      }
    });
  }
  for (int a = lane; a < S.nads; a += 64)
  {
    const int m = S.amol[a];
    if (m == mol || m == skipMol) continue;
    double dx = S.ax[a] - px, dy = S.ay[a] - py, dz = S.az[a] - pz;
    rgd::pbc(S.box, dx, dy, dz);
    const double rr = dx * dx + dy * dy + dz * dz;
    if (rr < S.rcv2Mol)
    {
      const double v = rgd::vdw0(S.pairs[S.at[a] * S.ntypes + type], 1.0, rr);
      ov = ov || v > S.overlap;
      evi += v;
    }
    if (S.useCharge && rr < S.rcc2)
    {
      const double r = sqrt(rr);
      eci += rgd::coul0(S.alpha, 1.0, r, S.coulomb * S.aq[a] * q);
    }
  }
  e[0] = rgd::waveSum(evf);
  e[1] = rgd::waveSum(ecf);
  e[2] = rgd::waveSum(evi);
  e[3] = rgd::waveSum(eci);
  overlap = __ballot(ov) != 0ull;
}

__global__ __launch_bounds__(256) void k_trial_energies(const SysDev* __restrict__ Sp, int n, const double* pos,
                                                        const int* type, const double* q, const int* mol,
                                                        int skipMol, double* out, int* overlap)
{
  const SysDev& S = *Sp;
  const int site = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (site >= n) return;
  double e[4];
  bool ov;
  siteEnergyWave(S, pos[3 * site], pos[3 * site + 1], pos[3 * site + 2], type[site], q[site], mol[site], skipMol,
                 lane, e, ov);
  if (lane == 0)
  {
    for (int c = 0; c < 4; ++c) out[4 * site + c] = e[c];
    overlap[site] = ov ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------------------------------------------
// Rosenbluth weight and trial selection
// ---------------------------------------------------------------------------------------------------------------

// lbf_i = -beta U_i over the non-overlapping trials; W = sum exp(lbf_i) / k evaluated as a shifted sum; the
// selection walks the shifted weights exactly as the reference's cumulative search.
__device__ void rosenbluthSelect(const double* U, const int* ov, int k, double beta, double u, int& selected,
                                 double& logW)
{
  double mx = -INFINITY;
  int any = 0;
  for (int i = 0; i < k; ++i)
    if (!ov[i])
    {
      mx = fmax(mx, -beta * U[i]);
      any = 1;
    }
  if (!any)
  {
    selected = -1;
    logW = -INFINITY;
    return;
  }
  double sum = 0.0;
  for (int i = 0; i < k; ++i)
    if (!ov[i]) sum += exp(-beta * U[i] - mx);
  const double ws = u * sum;
  int sel = -1;
  double cum = 0.0;
  for (int i = 0; i < k; ++i)
  {
    if (ov[i]) continue;
    if (sel < 0)
    {
      sel = i;
      cum = exp(-beta * U[i] - mx);
      continue;
    }
    if (!(cum < ws)) break;
    cum += exp(-beta * U[i] - mx);
    sel = i;
  }
  selected = sel;
  logW = mx + log(sum) - log(double(k));
}

__global__ void k_rosenbluth(int nsets, int k, const double* U, const int* ov, double beta, const double* uniforms,
                             int* selected, double* logW)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsets) return;
  rosenbluthSelect(U + size_t(s) * k, ov + size_t(s) * k, k, beta, uniforms[s], selected[s], logW[s]);
}

// ---------------------------------------------------------------------------------------------------------------
// Ewald Fourier
// ---------------------------------------------------------------------------------------------------------------

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 cconj(double2 a, bool c) { return make_double2(a.x, c ? -a.y : a.y); }

// eik tables of up to EW_MAXATOMS atoms in LDS: tab[atom][axis][k], k = 0..kmax (<= EW_MAXK)
__device__ void buildEik(const SysDev& S, int na, const double* pos, double2 (*tab)[3][EW_MAXK + 1])
{
  for (int t = threadIdx.x; t < na * 3; t += blockDim.x)
  {
    const int a = t / 3, ax = t % 3;
    const double px = pos[4 * a], py = pos[4 * a + 1], pz = pos[4 * a + 2];
    const double s = 2.0 * M_PI * (S.box.inv[ax] * px + S.box.inv[3 + ax] * py + S.box.inv[6 + ax] * pz);
    double sn, cs;
    sincos(s, &sn, &cs);
    const double2 e1 = make_double2(cs, sn);
    tab[a][ax][0] = make_double2(1.0, 0.0);
    double2 v = e1;
    for (int k = 1; k <= S.kmax[ax]; ++k)
    {
      tab[a][ax][k] = v;
      v = cmul(v, e1);
    }
  }
}

__device__ __forceinline__ double2 eikOf(const double2 (*tab)[3][EW_MAXK + 1], int a, int kx, int ky, int kz)
{
  const double2 ex = tab[a][0][kx];
  const double2 ey = cconj(tab[a][1][ky < 0 ? -ky : ky], ky < 0);
  const double2 ez = cconj(tab[a][2][kz < 0 ? -kz : kz], kz < 0);
  return cmul(cmul(ex, ey), ez);
}

__device__ double blockSum(double v, double* red)
{
  v = rgd::waveSum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < int(blockDim.x >> 6); ++i) t += red[i];
  __syncthreads();
  return t;
}

// Fourier part of the energy difference of replacing 'old' by 'new' atoms (each as x, y, z, q*scaling).
// Writes the trial structure factor into St and returns the Fourier difference without the net-charge term.
__device__ double ewaldDeltaBlock(const SysDev& S, int nn, const double* nw, int no, const double* od, double* red,
                                  double2 (*tab)[3][EW_MAXK + 1])
{
  buildEik(S, nn, nw, tab);
  buildEik(S, no, od, tab + nn);
  __syncthreads();
  double de = 0.0;
  for (int k = threadIdx.x; k < S.nk; k += blockDim.x)
  {
    const int code = S.kcode[k];
    const int kx = code & 0xff, ky = ((code >> 8) & 0xff) - 128, kz = ((code >> 16) & 0xff) - 128;
    double2 d = make_double2(0.0, 0.0);
    for (int a = 0; a < nn; ++a)
    {
      const double2 e = eikOf(tab, a, kx, ky, kz);
      d.x += nw[4 * a + 3] * e.x;
      d.y += nw[4 * a + 3] * e.y;
    }
    for (int a = 0; a < no; ++a)
    {
      const double2 e = eikOf(tab, nn + a, kx, ky, kz);
      d.x -= od[4 * a + 3] * e.x;
      d.y -= od[4 * a + 3] * e.y;
    }
    const double2 s = S.Sk[k];
    const double2 t = make_double2(s.x + d.x, s.y + d.y);
    de += S.fk[k] * (t.x * t.x + t.y * t.y) - S.fk[k] * (s.x * s.x + s.y * s.y);
    S.St[k] = t;
  }
  return blockSum(de, red);
}

// Self and intra-molecular exclusion terms of a set of atoms (src/raspakit/interactions/ewald.cpp:1165-1234):
// returns sum over atoms of C alpha/sqrt(pi) q^2 and sum over same-molecule pairs of C qi qj erf(alpha r)/r.
__device__ void selfExclusion(const SysDev& S, int n, const double* at, const int* mol, double& self, double& excl)
{
  self = 0.0;
  excl = 0.0;
  const double ps = S.coulomb * S.alpha / sqrt(M_PI);
  for (int i = 0; i < n; ++i)
  {
    self += ps * at[4 * i + 3] * at[4 * i + 3];
    for (int j = i + 1; j < n; ++j)
    {
      if (mol[i] != mol[j]) continue;
      double dx = at[4 * i] - at[4 * j], dy = at[4 * i + 1] - at[4 * j + 1], dz = at[4 * i + 2] - at[4 * j + 2];
      rgd::pbc(S.box, dx, dy, dz);
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      excl += S.coulomb * at[4 * i + 3] * at[4 * j + 3] * erf(S.alpha * r) / r;
    }
  }
}

__global__ __launch_bounds__(256) void k_ewald_delta(const SysDev* __restrict__ Sp, int nn, const double* nw,
                                                     const int* nmol, int no, const double* od, const int* omol,
                                                     double* out)
{
  const SysDev& S = *Sp;
  __shared__ double red[4];
  __shared__ double2 tab[EW_MAXATOMS][3][EW_MAXK + 1];
  const double df = ewaldDeltaBlock(S, nn, nw, no, od, red, tab);
  if (threadIdx.x == 0)
  {
    double sn, en, so, eo;
    selfExclusion(S, nn, nw, nmol, sn, en);
    selfExclusion(S, no, od, omol, so, eo);
    double dq = 0.0;
    for (int a = 0; a < nn; ++a) dq += nw[4 * a + 3];
    for (int a = 0; a < no; ++a) dq -= od[4 * a + 3];
    const double uIon = -(S.singleIon - S.coulomb * S.alpha / sqrt(M_PI));
    const double qn = S.netCharge + dq;
    out[0] = df + uIon * (qn * qn - S.netCharge * S.netCharge);
    out[1] = so - sn;
    out[2] = eo - en;
    out[3] = dq;
  }
}

__global__ void k_copy_sk(double2* dst, const double2* src, int nk)
{
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nk) dst[k] = src[k];
}

// S_k = sum_atoms q e^{i k.r} by direct phases (used when (re)building the stored structure factor).
__global__ void k_structure_factor(const SysDev* __restrict__ Sp, int n, const double* x, const double* y,
                                   const double* z, const double* q, double2* out, int accumulate)
{
  const SysDev& S = *Sp;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= S.nk) return;
  const int code = S.kcode[k];
  const double kx = code & 0xff, ky = ((code >> 8) & 0xff) - 128, kz = ((code >> 16) & 0xff) - 128;
  double re = 0.0, im = 0.0;
  for (int a = 0; a < n; ++a)
  {
    const double s0 = S.box.inv[0] * x[a] + S.box.inv[3] * y[a] + S.box.inv[6] * z[a];
    const double s1 = S.box.inv[1] * x[a] + S.box.inv[4] * y[a] + S.box.inv[7] * z[a];
    const double s2 = S.box.inv[2] * x[a] + S.box.inv[5] * y[a] + S.box.inv[8] * z[a];
    // reduce each fractional coordinate first so the phase stays small
    const double f0 = s0 - rint(s0), f1 = s1 - rint(s1), f2 = s2 - rint(s2);
    double sn, cs;
    sincos(2.0 * M_PI * (kx * f0 + ky * f1 + kz * f2), &sn, &cs);
    re += q[a] * cs;
    im += q[a] * sn;
  }
  if (accumulate)
  {
    out[k].x += re;
    out[k].y += im;
  }
  else
    out[k] = make_double2(re, im);
}

// Fourier energy of the adsorbates: sum_k f_k (|S_k|^2 - |S_fw,k|^2), block-reduced into out[blockIdx.x].
__global__ __launch_bounds__(256) void k_fourier_energy(const SysDev* __restrict__ Sp, double* out)
{
  const SysDev& S = *Sp;
  __shared__ double red[4];
  double e = 0.0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < S.nk; k += gridDim.x * blockDim.x)
  {
    const double2 s = S.Sk[k], f = S.Sfw[k];
    e += S.fk[k] * ((s.x * s.x + s.y * s.y) - (f.x * f.x + f.y * f.y));
  }
  e = blockSum(e, red);
  if (threadIdx.x == 0) out[blockIdx.x] = e;
}

}  // namespace rg

// Device-resident Monte Carlo for rigid molecules on gfx950: Widom insertion batches and multi-system GCMC.
//
// One wavefront owns one system (GCMC) or one insertion (Widom); four per workgroup.  Inside a move each lane takes
// one trial (first-bead position or orientation): grid-typed sites are interpolated lane-parallel, direct sums and
// inter-molecular sums loop cooperatively over atoms with wavefront-shuffle reductions, and the Ewald Fourier
// difference strides over k-vectors with FP64 eik recursions per atom in LDS.
//
// Moves (src/raspakit/mc_moves/component/): Widom (widom.cpp:32-155), CBMC swap insertion (insertion_cbmc.cpp:34-200)
// and deletion (deletion_cbmc.cpp:67-174) with multiple first beads (src/raspakit/cbmc/multiple_first_bead.cpp:20-76)
// followed by uniformly random orientations of the rigid body about the first bead (the rigid seed step of
// src/raspakit/cbmc/operators.cpp:958-968), translation and rotation.  The ideal-gas Rosenbluth weight of a rigid
// body with orientation trials is 1.  Trials come from a Philox4x32-10 stream addressed by (system, move, slot),
// which the oracle replays.
//
// A learned insertion proposal replaces the uniform first-bead distribution by a tabulated density q on voxels
// (sampled through an alias table, uniformly inside the voxel); each first-bead trial then carries the weight
// exp(-beta U) / (V q) in both growth and retrace, so the acceptance rule stays exact.
#include <cmath>
#include <cstring>

#include "mc_engine.hpp"

namespace rg
{

// ---------------------------------------------------------------------------------------------------------------
// random numbers: slot layout of one move
// ---------------------------------------------------------------------------------------------------------------
enum : uint32_t
{
  SLOT_MOVE = 0,       // (move type, component)
  SLOT_PICK = 1,       // (molecule, acceptance)
  SLOT_SELECT = 2,     // (first-bead selection, orientation selection)
  SLOT_DISP = 3,       // translation (2 slots) / rotation (2 slots)
  SLOT_FIRST = 16,     // first-bead trial j: slots 16 + 2j, 17 + 2j
  SLOT_ORIENT = 160,   // orientation trial j: slots 160 + 2j, 161 + 2j
  SLOT_VOXEL = 304,    // proposal trial j: slots 304 + 2j, 305 + 2j
  SLOT_REINS = 1u << 24  // offset of every slot of the retrace of a reinsertion (its trials are new draws)
};

// What a growth does.  retrace < 0 grows a new molecule; retrace = m retraces molecule m (its atoms are trial 0 of
// every step).  'own' is the molecule left out of the background (-2: none).  'so' is added to every Philox slot of
// the growth.  reins = 1 grows the new configuration of a reinsertion and sets logStoredR to the log of the summed
// weights of the first-bead trials it did not select; reins = 2 retraces the old configuration of a reinsertion,
// whose first-bead weight is that of its first bead alone plus exp(logStoredR) (Esselink et al., J. Chem. Phys. 101
// (1994) 9033, eqs. 16 and 18; src/raspakit/cbmc/multiple_first_bead.cpp:78-136).
struct GrowMode
{
  int retrace, own;
  uint32_t so;
  int reins;
  double logStoredR;
};

__device__ __forceinline__ void uni(const McRun& R, int sys, unsigned long long step, uint32_t slot, double& a,
                                    double& b)
{
  rgd::uniform2(R.seed, uint32_t(sys), uint32_t(step), uint32_t(step >> 32), slot, a, b);
}

// Ordering inside one wavefront: LDS (compiler barrier + wavefront fence) and global memory written by one lane
// and read by others (workgroup-scope fence: waits for the stores and keeps the per-CU cache coherent).
__device__ __forceinline__ void waveSyncLds()
{
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ void waveSyncGlobal()
{
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// Compiler-only memory barrier: values read from the per-wavefront LDS descriptors before it are read again after
// it.  Placed at the start of each move and of the heavy helpers, it keeps the compiler from holding every
// descriptor field in registers across the whole kernel (which spills them to scratch memory).
__device__ __forceinline__ void reloadBarrier() { asm volatile("" ::: "memory"); }

// Rotation matrix (row-major) of the unit quaternion from three uniforms (K. Shoemake, "Uniform random rotations",
// Graphics Gems III, 1992).
__device__ __forceinline__ void randomRotation(double u1, double u2, double u3, double Rm[9])
{
  const double s1 = sqrt(1.0 - u1), s2 = sqrt(u1);
  double sa, ca, sb, cb;
  sincos(2.0 * M_PI * u2, &sa, &ca);
  sincos(2.0 * M_PI * u3, &sb, &cb);
  const double x = s1 * sa, y = s1 * ca, z = s2 * sb, w = s2 * cb;
  Rm[0] = 1.0 - 2.0 * (y * y + z * z);
  Rm[1] = 2.0 * (x * y - z * w);
  Rm[2] = 2.0 * (x * z + y * w);
  Rm[3] = 2.0 * (x * y + z * w);
  Rm[4] = 1.0 - 2.0 * (x * x + z * z);
  Rm[5] = 2.0 * (y * z - x * w);
  Rm[6] = 2.0 * (x * z - y * w);
  Rm[7] = 2.0 * (y * z + x * w);
  Rm[8] = 1.0 - 2.0 * (x * x + y * y);
}

// ---------------------------------------------------------------------------------------------------------------
// energies of sets of sites
// ---------------------------------------------------------------------------------------------------------------

// External energy of one site with the whole wavefront (framework direct or grid, plus adsorbates except molecule
// 'own' and 'skip'); overlap as in the reference (any pair above the criterion, or the grid value above it).
// withGrid = false leaves out the grid part (the caller interpolates it lane-parallel).
__device__ double siteExternal(const SysDev& S, int nads, double px, double py, double pz, int type, double q, int own,
                               int skip, int lane, bool& ov, bool withGrid = true)
{
  double e = 0.0;
  bool o = false;
  const int gi = S.gridOfType ? S.gridOfType[type] : -1;
  if (!withGrid && gi >= 0 && nads == 0)
  {
    ov = false;
    return 0.0;
  }
  if (gi < 0)
  {
    forNearFramework(S, px, py, pz, lane, [&](int a) {
      double dx = S.fwx[a] - px, dy = S.fwy[a] - py, dz = S.fwz[a] - pz;
      rgd::pbc(S.box, dx, dy, dz);
      const double rr = dx * dx + dy * dy + dz * dz;
      if (rr < S.rcv2Fw)
      {
        const double v = rgd::vdw0(S.pairs[S.fwt[a] * S.ntypes + type], 1.0, rr);
        o = o || v > S.overlap;
        e += v;
      }
      if (S.useCharge && rr < S.rcc2) e += rgd::coul0(S.alpha, 1.0, sqrt(rr), S.coulomb * S.fwq[a] * q);
    });
  }
  // adsorbate atoms, four per lane and pass so that one memory latency covers their loads
  constexpr int U = 4;
  for (int a0 = 0; a0 < nads; a0 += 64 * U)
  {
    double ax[U], ay[U], az[U], aq[U];
    int am[U], at[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      const int a = a0 + u * 64 + lane;
      const bool in = a < nads;
      am[u] = in ? S.amol[a] : -3;
      at[u] = in ? S.at[a] : -1;
      ax[u] = in ? S.ax[a] : 0.0;
      ay[u] = in ? S.ay[a] : 0.0;
      az[u] = in ? S.az[a] : 0.0;
      aq[u] = in ? S.aq[a] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
    {
      if (am[u] == own || am[u] == skip || at[u] < 0) continue;
      double dx = ax[u] - px, dy = ay[u] - py, dz = az[u] - pz;
      rgd::pbc(S.box, dx, dy, dz);
      const double rr = dx * dx + dy * dy + dz * dz;
      if (rr < S.rcv2Mol)
      {
        const double v = rgd::vdw0(S.pairs[at[u] * S.ntypes + type], 1.0, rr);
        o = o || v > S.overlap;
        e += v;
      }
      if (S.useCharge && rr < S.rcc2) e += rgd::coul0(S.alpha, 1.0, sqrt(rr), S.coulomb * aq[u] * q);
    }
  }
  e = rgd::waveSum(e);
  if (withGrid && gi >= 0)
  {
    const double g = rgd::interpolate(S.grids[gi], px, py, pz);
    o = o || g > S.overlap;
    e += g;
    if (S.useCharge && S.coulGrid >= 0) e += q * rgd::interpolate(S.grids[S.coulGrid], px, py, pz);
  }
  ov = __ballot(o) != 0ull;
  return e;
}

// Pair table of the force field in LDS, filled by the kernels that use the batched site energies (every system of
// a launch shares the force field); types beyond MC_LDSTYPES read the global table.
__shared__ rgd::Pair g_pairL[MC_LDSTYPES * MC_LDSTYPES];

__device__ __forceinline__ void loadPairTable(const SysDev& S)
{
  if (S.ntypes <= MC_LDSTYPES)
    for (int i = threadIdx.x; i < S.ntypes * S.ntypes; i += blockDim.x) g_pairL[i] = S.pairs[i];
  __syncthreads();
}

// Framework atoms of a site whose type has no grid: direct sum over the whole wavefront.
__device__ double frameworkSiteEnergy(const SysDev& S, double px, double py, double pz, int type, double q, int lane,
                                      bool& ov)
{
  double e = 0.0;
  bool o = false;
  for (int a = lane; a < S.nfw; a += 64)
  {
    double dx = S.fwx[a] - px, dy = S.fwy[a] - py, dz = S.fwz[a] - pz;
    rgd::pbc(S.box, dx, dy, dz);
    const double rr = dx * dx + dy * dy + dz * dz;
    if (rr < S.rcv2Fw)
    {
      const double v = rgd::vdw0(S.pairs[S.fwt[a] * S.ntypes + type], 1.0, rr);
      o = o || v > S.overlap;
      e += v;
    }
    if (S.useCharge && rr < S.rcc2) e += rgd::coul0(S.alpha, 1.0, sqrt(rr), S.coulomb * S.fwq[a] * q);
  }
  ov = __ballot(o) != 0ull;
  return rgd::waveSum(e);
}

// Evaluates the candidate pairs W.pKey / W.pRR [0, fill) and adds each site's sum into W.siteT.  Keys are
// nondecreasing in the site, so an inclusive scan that only adds across equal keys leaves every run's total in its
// last lane; a site has at most one run per block of 64, and blocks are taken in order, so each site's sum has a
// fixed order whatever the launch shape.
__device__ __attribute__((always_inline)) void flushPairs(const SysDev& S, int fill, int lane, WaveScratch& W)
{
  waveSyncLds();
  const bool ldsPairs = S.ntypes <= MC_LDSTYPES;
  for (int p0 = 0; p0 < fill; p0 += 64)
  {
    const int p = p0 + lane;
    const bool live = p < fill;
    const int key = live ? W.pKey[p] : 0x7fffffff;
    const int s = key >> 8;
    double e = 0.0;
    bool ov = false;
    if (live)
    {
      const double rr = W.pRR[p];
      const int l = key & 0xff;
      const int t = W.chT[l], st = W.sType[s];
      if (rr < S.rcv2Mol)
      {
        const rgd::Pair& P = ldsPairs ? g_pairL[t * S.ntypes + st] : S.pairs[t * S.ntypes + st];
        const double v = rgd::vdw0(P, 1.0, rr);
        ov = v > S.overlap;
        e = v;
      }
      if (S.useCharge && rr < S.rcc2) e += rgd::coul0(S.alpha, 1.0, sqrt(rr), S.coulomb * W.chQ[l] * W.sQ[s]);
    }
    // runs are keyed by the site alone (dead lanes carry a site past every real one)
    for (int o = 1; o < 64; o <<= 1)
    {
      const double t = __shfl_up(e, o, 64);
      const int ks = __shfl_up(s, o, 64);
      if (lane >= o && ks == s) e += t;
    }
    const int next = __shfl_down(s, 1, 64);
    if (live && (lane == 63 || next != s)) W.siteT[s] += e;
    if (live && ov) W.ovT[s] = 1;
    waveSyncLds();
  }
}

// External energies of ns sites (positions pos[3 s .. 3 s + 2], types W.sType, charges W.sQ) without their grid
// part: framework atoms of types without a grid, and every adsorbate atom except those of molecules 'own' and
// 'skip'.  Results in W.siteE[s] and W.siteOv[s] (any pair above the overlap criterion).  The adsorbate atoms are
// read 64 at a time, once for all sites: each lane tests its atom against every site, the pairs within the cutoff
// are compacted into LDS and evaluated densely by flushPairs.  The chunks are split into MC_PARTS partitions (one
// per wavefront of a team of MC_PARTS, or all on one wavefront), summed separately and combined in order.
__device__ __attribute__((always_inline)) void siteSetEnergies(const SysDev& S, int nads, int ns, const double* pos, int own, int skip, int lane,
                                WaveScratch& W, const Team& T)
{
  reloadBarrier();
  for (int s = lane; s < ns; s += 64)
  {
    W.siteE[s] = 0.0;
    W.siteOv[s] = 0;
  }
  waveSyncLds();
  if (S.nfw > 0)
  {
    for (int s = 0; s < ns; ++s)
    {
      const int type = W.sType[s];
      if (S.gridOfType && S.gridOfType[type] >= 0) continue;
      bool o;
      const double e = frameworkSiteEnergy(S, pos[3 * s], pos[3 * s + 1], pos[3 * s + 2], type, W.sQ[s], lane, o);
      if (lane == 0)
      {
        W.siteE[s] = e;
        W.siteOv[s] = o ? 1 : 0;
      }
    }
    waveSyncLds();
  }
  const double cut2 = S.useCharge ? fmax(S.rcv2Mol, S.rcc2) : S.rcv2Mol;
  const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  // adsorbate chunks q, q + MC_PARTS, ... of 64 atoms into W.siteT / W.ovT
  auto partition = [&](int q)
  {
    for (int s = lane; s < ns; s += 64)
    {
      W.siteT[s] = 0.0;
      W.ovT[s] = 0;
    }
    waveSyncLds();
    for (int a0 = 64 * q; a0 < nads; a0 += 64 * MC_PARTS)
    {
      const int a = a0 + lane;
      double ax = 0.0, ay = 0.0, az = 0.0;
      bool valid = false;
      if (a < nads)
      {
        const int t = S.at[a], m = S.amol[a];
        valid = t >= 0 && m != own && m != skip;
        ax = S.ax[a];
        ay = S.ay[a];
        az = S.az[a];
        W.chT[lane] = t;
        W.chQ[lane] = S.aq[a];
      }
      int fill = 0;
      for (int s = 0; s < ns; ++s)
      {
        if (fill > MC_PAIRCAP - 64)
        {
          flushPairs(S, fill, lane, W);
          fill = 0;
        }
        double dx = ax - pos[3 * s], dy = ay - pos[3 * s + 1], dz = az - pos[3 * s + 2];
        rgd::pbc(S.box, dx, dy, dz);
        const double rr = dx * dx + dy * dy + dz * dz;
        const bool in = valid && rr < cut2;
        const unsigned long long m = __ballot(in);
        if (in)
        {
          const int idx = fill + __popcll(m & below);
          W.pKey[idx] = (s << 8) | lane;
          W.pRR[idx] = rr;
        }
        fill += __popcll(m);
      }
      flushPairs(S, fill, lane, W);  // before the next chunk replaces chT / chQ
    }
    waveSyncLds();
  };
  // site sums ((((framework + P0) + P1) + P2) + P3) in either team size
  if (T.nw == 1)
  {
    for (int q = 0; q < MC_PARTS && 64 * q < nads; ++q)
    {
      partition(q);
      for (int s = lane; s < ns; s += 64)
      {
        W.siteE[s] += W.siteT[s];
        W.siteOv[s] |= W.ovT[s];
      }
      waveSyncLds();
    }
  }
  else
  {
    partition(T.wi);
    __syncthreads();
    for (int s = lane; s < ns; s += 64)
    {
      double e = W.siteE[s];
      int o = W.siteOv[s];
      for (int q = 0; q < MC_PARTS; ++q)
      {
        e += T.ws[q].siteT[s];
        o |= T.ws[q].ovT[s];
      }
      W.siteE[s] = e;
      W.siteOv[s] = o;
    }
    __syncthreads();  // the partition arrays are read by every wavefront before any overwrites them
  }
}

// Energies of K trial sets of n sites each (sites in LDS as x, y, z per site, set-major).  Lane j < K receives the
// total of set j and its overlap flag.
__device__ void trialSetEnergies(const SysDev& S, int nads, const Comp& c, const int* atomIdx, int n, int K,
                                 const double* pos, int own, int skip, int lane, WaveScratch& W, const Team& T,
                                 double& Uj, bool& ovj)
{
  const int ns = K * n;
  for (int s = 0; s < n; ++s)  // uniform loop: atomIdx stays in registers
  {
    const int a = atomIdx[s];
    for (int j = lane; j < K; j += 64)
    {
      W.sType[j * n + s] = c.type[a];
      W.sQ[j * n + s] = c.q[a];
    }
  }
  waveSyncLds();
  // grid-typed sites: term t = 2 (j n + s) + g is term g of site s of trial set j (g = 0 the site type's grid, g = 1
  // its charge times the Coulomb grid); lane t % 64 of round t / 64 takes it, so a round's gathers are all in flight
  // at once.  Lane j adds its set's terms in increasing t (site by site, own grid before Coulomb: the order of one
  // lane interpolating its whole set), whatever round they fall in.
  double gpart = 0.0;
  bool gov = false;
  if (S.gridOfType)
  {
    const int nterm = 2 * ns;
    const int tj = 2 * n * lane;  // lane j's first term
    for (int r0 = 0; r0 < nterm; r0 += 64)
    {
      const int t = r0 + lane;
      double v = 0.0;
      bool o = false;
      if (t < nterm)
      {
        const int site = t >> 1;
        const bool coulTerm = (t & 1) != 0;
        const int gi = S.gridOfType[W.sType[site]];
        const int gsel = coulTerm ? ((S.useCharge && S.coulGrid >= 0) ? S.coulGrid : -1) : gi;
        if (gi >= 0 && gsel >= 0)
        {
          const double* p = pos + 3 * site;
          v = rgd::interpolate(S.grids[gsel], p[0], p[1], p[2]);
          if (coulTerm)
            v *= W.sQ[site];
          else
            o = v > S.overlap;
        }
      }
      const unsigned long long om = __ballot(o);
      for (int q = 0; q < 2 * n; ++q)
      {
        const int src = tj + q - r0;
        const double x = __shfl(v, min(max(src, 0), 63), 64);
        if (lane < K && src >= 0 && src < 64)
        {
          gpart += x;
          gov = gov || ((om >> src) & 1ull) != 0ull;
        }
      }
    }
  }
  // adsorbate atoms (and framework atoms of types without a grid): all K * n sites in one batch
  siteSetEnergies(S, nads, ns, pos, own, skip, lane, W, T);
  Uj = 0.0;
  ovj = false;
  if (lane < K)
  {
    double u = 0.0;
    bool any = false;
    for (int s = 0; s < n; ++s)
    {
      u += W.siteE[lane * n + s];
      any = any || W.siteOv[lane * n + s] != 0;
    }
    Uj = u + gpart;
    ovj = any || gov;
  }
  waveSyncLds();
}

// Rosenbluth weight over the lanes (lane j holds trial j): logW = log(sum_j w_j / K) with w_j = exp(-beta U_j) * bias_j
// over non-overlapping trials, and the trial the cumulative search picks for uniform u (-1 if none).
__device__ void waveRosenbluth(double Uj, bool ovj, double logBias, int K, double beta, double u, int lane, int& sel,
                               double& logW)
{
  const bool valid = lane < K && !ovj;
  const double lbf = valid ? -beta * Uj + logBias : -INFINITY;
  const double mx = rgd::waveMax(lbf);
  if (__ballot(valid) == 0ull)
  {
    sel = -1;
    logW = -INFINITY;
    return;
  }
  const double w = valid ? exp(lbf - mx) : 0.0;
  // inclusive prefix sum in lane order (the reference's accumulation order)
  double pre = w;
  for (int o = 1; o < 64; o <<= 1)
  {
    const double t = __shfl_up(pre, o, 64);
    if (lane >= o) pre += t;
  }
  const double sum = __shfl(pre, 63, 64);
  const double ws = u * sum;
  // first valid trial whose cumulative weight reaches ws; else the last valid trial
  const unsigned long long reach = __ballot(valid && !(pre < ws));
  const unsigned long long vmask = __ballot(valid);
  sel = reach ? __ffsll((long long)reach) - 1 : 63 - __clzll((long long)vmask);
  logW = mx + log(sum) - log(double(K));
}

// ---------------------------------------------------------------------------------------------------------------
// Ewald Fourier difference with one wavefront
// ---------------------------------------------------------------------------------------------------------------
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// atoms: x, y, z, q (q carries the sign: + new, - old); tab: per atom 3 x (EW_MAXK + 1) complex.
// Passes of 256 k-vectors (four per lane: their loads are issued together, so one memory latency covers four);
// pass p belongs to partition p % MC_PARTS, and a wavefront of a team of MC_PARTS runs its own partition only.  The
// trial structure factor of a pass is written, and later committed by waveAcceptEwald, by the same wavefront.
__device__ __attribute__((always_inline)) double waveEwaldDelta(const SysDev& S, int na, const double* at, double2* tab,
                                                                int lane, bool writeTrial, WaveScratch& W,
                                                                const Team& T)
{
  reloadBarrier();
  const int stride = 3 * (EW_MAXK + 1);
  for (int t = lane; t < na * 3; t += 64)
  {
    const int a = t / 3, ax = t % 3;
    const double s =
        2.0 * M_PI * (S.box.inv[ax] * at[4 * a] + S.box.inv[3 + ax] * at[4 * a + 1] + S.box.inv[6 + ax] * at[4 * a + 2]);
    double sn, cs;
    sincos(s, &sn, &cs);
    const double2 e1 = make_double2(cs, sn);
    double2* row = tab + a * stride + ax * (EW_MAXK + 1);
    row[0] = make_double2(1.0, 0.0);
    double2 v = e1;
    for (int k = 1; k <= S.kmax[ax]; ++k)
    {
      row[k] = v;
      v = cmul(v, e1);
    }
  }
  waveSyncLds();
  constexpr int U = 4;
  const int nk = S.nk;
  const int npass = (nk + 64 * U - 1) / (64 * U);
  auto partition = [&](int q)
  {
    double de = 0.0;
    for (int pass = q; pass < npass; pass += MC_PARTS)
    {
      const int k0 = pass * 64 * U;
      int code[U];
      double2 sk[U];
      double fk[U];
#pragma unroll
      for (int u = 0; u < U; ++u)
      {
        const int k = k0 + u * 64 + lane;
        const bool in = k < nk;
        code[u] = in ? S.kcode[k] : (128 << 8) | (128 << 16);  // padding decodes to k = (0, 0, 0)
        sk[u] = in ? S.Sk[k] : make_double2(0.0, 0.0);
        fk[u] = in ? S.fk[k] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
      {
        const int k = k0 + u * 64 + lane;
        const int kx = code[u] & 0xff, ky = ((code[u] >> 8) & 0xff) - 128, kz = ((code[u] >> 16) & 0xff) - 128;
        double2 d = make_double2(0.0, 0.0);
        for (int a = 0; a < na; ++a)
        {
          const double2* row = tab + a * stride;
          const double2 ex = row[kx];
          double2 ey = row[(EW_MAXK + 1) + (ky < 0 ? -ky : ky)];
          if (ky < 0) ey.y = -ey.y;
          double2 ez = row[2 * (EW_MAXK + 1) + (kz < 0 ? -kz : kz)];
          if (kz < 0) ez.y = -ez.y;
          const double2 e = cmul(cmul(ex, ey), ez);
          d.x += at[4 * a + 3] * e.x;
          d.y += at[4 * a + 3] * e.y;
        }
        if (k >= nk) continue;
        const double2 t = make_double2(sk[u].x + d.x, sk[u].y + d.y);
        de += fk[u] * ((t.x * t.x + t.y * t.y) - (sk[u].x * sk[u].x + sk[u].y * sk[u].y));
        if (writeTrial) S.St[k] = t;
      }
    }
    return rgd::waveSum(de);
  };
  // ((P0 + P1) + P2) + P3 in either team size
  if (T.nw == 1)
  {
    double tot = 0.0;
    for (int q = 0; q < MC_PARTS; ++q) tot += partition(q);
    return tot;
  }
  const double mine = partition(T.wi);
  if (lane == 0) W.deT = mine;
  __syncthreads();
  double tot = 0.0;
  for (int q = 0; q < MC_PARTS; ++q) tot += T.ws[q].deT;
  __syncthreads();
  return tot;
}

// Commits the trial structure factor: each wavefront copies the passes it computed.
__device__ void waveAcceptEwald(const SysDev& S, int lane, const Team& T)
{
  // four k-vectors per lane and pass, loaded before any is stored (indices clamped, so no load is conditional)
  const int nk = S.nk;
  const double2* __restrict__ src = S.St;
  double2* __restrict__ dst = S.Sk;
  const int npass = (nk + 255) / 256;
  for (int pass = (T.nw == 1 ? 0 : T.wi); pass < npass; pass += (T.nw == 1 ? 1 : MC_PARTS))
  {
    const int k = pass * 256 + lane;
    const double2 v0 = src[min(k, nk - 1)], v1 = src[min(k + 64, nk - 1)];
    const double2 v2 = src[min(k + 128, nk - 1)], v3 = src[min(k + 192, nk - 1)];
    if (k < nk) dst[k] = v0;
    if (k + 64 < nk) dst[k + 64] = v1;
    if (k + 128 < nk) dst[k + 128] = v2;
    if (k + 192 < nk) dst[k + 192] = v3;
  }
}


// ---------------------------------------------------------------------------------------------------------------
// insertion growth and deletion retrace (rigid molecules)
// ---------------------------------------------------------------------------------------------------------------

// First-bead position of trial j: uniform in the cell or drawn from the proposal (alias table over voxels, uniform
// inside the voxel).  Returns log(1 / (V q)) = -log(nvox p_v) (0 for the uniform distribution).
__device__ double firstBeadTrial(const McRun& R, const SysDev& S, const McSys& Y, int ci, int sys,
                                 unsigned long long step, int j, double p[3], uint32_t so = 0)
{
  double u0, u1, u2, u3;
  if (Y.prop[ci].nvox > 0)
  {
    const Proposal& P = Y.prop[ci];
    uni(R, sys, step, so + SLOT_VOXEL + 2 * j, u0, u1);
    uni(R, sys, step, so + SLOT_VOXEL + 2 * j + 1, u2, u3);
    int col = min(int(u0 * P.nvox), P.nvox - 1);
    const int v = (u1 < P.aliasProb[col]) ? col : P.alias[col];
    const int vx = v % P.n[0], vy = (v / P.n[0]) % P.n[1], vz = v / (P.n[0] * P.n[1]);
    double w0, w1;
    uni(R, sys, step, so + SLOT_FIRST + 2 * j, w0, w1);
    const double f[3] = {(vx + u2) / P.n[0], (vy + u3) / P.n[1], (vz + w0) / P.n[2]};
    for (int c = 0; c < 3; ++c) p[c] = S.box.cell[c] * f[0] + S.box.cell[3 + c] * f[1] + S.box.cell[6 + c] * f[2];
    return -P.logNvoxP[v];
  }
  uni(R, sys, step, so + SLOT_FIRST + 2 * j, u0, u1);
  uni(R, sys, step, so + SLOT_FIRST + 2 * j + 1, u2, u3);
  // SimulationBox::randomPosition: cell * (u, u, u) (src/raspakit/simulationbox.cpp:132-135)
  for (int c = 0; c < 3; ++c) p[c] = S.box.cell[c] * u0 + S.box.cell[3 + c] * u1 + S.box.cell[6 + c] * u2;
  return 0.0;
}

__device__ double proposalLogBias(const McSys& Y, int ci, const SysDev& S, const double p[3])
{
  if (Y.prop[ci].nvox <= 0) return 0.0;
  const Proposal& P = Y.prop[ci];
  int idx[3];
  for (int a = 0; a < 3; ++a)
  {
    double f = S.box.inv[a] * p[0] + S.box.inv[3 + a] * p[1] + S.box.inv[6 + a] * p[2];
    f -= floor(f);
    idx[a] = min(int(f * P.n[a]), P.n[a] - 1);
  }
  return -P.logNvoxP[idx[0] + P.n[0] * (idx[1] + P.n[1] * idx[2])];
}

// First-bead step of a growth: kFirst trials (uniform or from the proposal; the existing first bead is trial 0 of a
// retrace), Rosenbluth selection.  On success every lane holds the selected position, its energy and log weight,
// and the second selection uniform of the move.  The retrace of a reinsertion (G.reins == 2) evaluates the existing
// first bead alone and adds the stored weight of the growth's unselected trials.
__device__ bool firstBeadStep(const McRun& R, const SysDev& S, const McSys& Y, const Comp& c, int sys, int nads,
                              unsigned long long step, const GrowMode G, int lane, WaveScratch& W, const Team& T,
                              double& fx, double& fy, double& fz, double& Ufirst, double& logW1, double& uSel2,
                              double& logStoredR)
{
  const int K1 = R.kFirst;
  const int retrace = G.retrace;
  const int Keval = G.reins == 2 ? 1 : K1;
  const int fb = c.first;
  double p[3] = {0.0, 0.0, 0.0};
  double logBias = 0.0;
  if (lane < Keval)
  {
    if (retrace >= 0 && lane == 0)
    {
      const int a0 = retrace * R.slots + fb;
      p[0] = S.ax[a0];
      p[1] = S.ay[a0];
      p[2] = S.az[a0];
      logBias = proposalLogBias(Y, c.index, S, p);
    }
    else
      logBias = firstBeadTrial(R, S, Y, c.index, sys, step, lane, p, G.so);
    W.pos[3 * lane] = p[0];
    W.pos[3 * lane + 1] = p[1];
    W.pos[3 * lane + 2] = p[2];
  }
  waveSyncLds();
  double U1;
  bool ov1;
  const int fbIdx[1] = {fb};
  trialSetEnergies(S, nads, c, fbIdx, 1, Keval, W.pos, G.own, -2, lane, W, T, U1, ov1);
  double uSel1;
  uni(R, sys, step, G.so + SLOT_SELECT, uSel1, uSel2);
  int sel1;
  if (G.reins == 2)
  {
    // w(o) = exp(-beta u(o)) + r, divided by the number of trials (Esselink et al. eq. 18)
    if (__shfl(ov1 ? 1 : 0, 0, 64)) return false;
    const double a0 = __shfl(-R.beta * U1 + logBias, 0, 64);
    const double hi = fmax(a0, G.logStoredR);
    logW1 = hi + log(exp(a0 - hi) + exp(G.logStoredR - hi)) - log(double(K1));
    sel1 = 0;
  }
  else
    waveRosenbluth(U1, ov1, logBias, K1, R.beta, uSel1, lane, sel1, logW1);
  if (retrace >= 0) sel1 = 0;  // the retrace keeps the existing molecule
  if (sel1 < 0 || logW1 < R.logMinW) return false;
  if (G.reins == 1)
  {
    // r = w(n) - exp(-beta U[selected]) (Esselink et al. eq. 16), as a log
    const bool other = lane < K1 && !ov1 && lane != sel1;
    const double lbf = other ? -R.beta * U1 + logBias : -INFINITY;
    const double mx = rgd::waveMax(lbf);
    const double sum = rgd::waveSum(other ? exp(lbf - mx) : 0.0);
    logStoredR = (mx == -INFINITY) ? -INFINITY : mx + log(sum);
  }
  fx = __shfl(p[0], sel1, 64);
  fy = __shfl(p[1], sel1, 64);
  fz = __shfl(p[2], sel1, 64);
  Ufirst = __shfl(U1, sel1, 64);
  if (retrace >= 0 && __shfl(ov1 ? 1 : 0, 0, 64)) return false;
  return true;
}

// Grows a molecule of component c: K first-bead trials, then (for more than one atom) K orientation trials about
// the selected first bead.  On success writes the atom positions into W.mol and returns true with the log Rosenbluth
// weight and the external energy of the grown molecule.  'retrace' >= 0 retraces that existing molecule instead
// (trial 0 = its current first bead and orientation; its own atoms are excluded).
__device__ bool growRigid(const McRun& R, const SysDev& S, const McSys& Y, const Comp& c, int sys, int nads,
                          unsigned long long step, const GrowMode G, int lane, WaveScratch& W, const Team& T,
                          double& logW, double& Uext, double& logStoredR)
{
  const int K2 = R.kOrient;
  const int retrace = G.retrace;
  const int own = G.own;
  const int fb = c.first;
  double fx, fy, fz, Ufirst, logW1, uSel2;
  if (!firstBeadStep(R, S, Y, c, sys, nads, step, G, lane, W, T, fx, fy, fz, Ufirst, logW1, uSel2, logStoredR))
    return false;
  if (c.na == 1)
  {
    if (lane == 0)
    {
      W.mol[0] = fx;
      W.mol[1] = fy;
      W.mol[2] = fz;
    }
    logW = logW1;
    Uext = Ufirst;
    return true;
  }
  // --- orientations about the first bead
  int idx[MC_MAXATOMS - 1];
  int n = 0;
  for (int a = 0; a < c.na; ++a)
    if (a != fb) idx[n++] = a;
  if (lane < K2)
  {
    double Rm[9];
    if (retrace >= 0 && lane == 0)
    {
      for (int s = 0; s < n; ++s)
      {
        const int a0 = retrace * R.slots + idx[s];
        W.pos[3 * s] = S.ax[a0];
        W.pos[3 * s + 1] = S.ay[a0];
        W.pos[3 * s + 2] = S.az[a0];
      }
    }
    else
    {
      double u1, u2, u3, u4;
      uni(R, sys, step, G.so + SLOT_ORIENT + 2 * lane, u1, u2);
      uni(R, sys, step, G.so + SLOT_ORIENT + 2 * lane + 1, u3, u4);
      randomRotation(u1, u2, u3, Rm);
      for (int s = 0; s < n; ++s)
      {
        const double* r = c.ref[idx[s]];
        W.pos[3 * (lane * n + s)] = fx + Rm[0] * r[0] + Rm[1] * r[1] + Rm[2] * r[2];
        W.pos[3 * (lane * n + s) + 1] = fy + Rm[3] * r[0] + Rm[4] * r[1] + Rm[5] * r[2];
        W.pos[3 * (lane * n + s) + 2] = fz + Rm[6] * r[0] + Rm[7] * r[1] + Rm[8] * r[2];
      }
    }
  }
  waveSyncLds();
  double U2;
  bool ov2;
  trialSetEnergies(S, nads, c, idx, n, K2, W.pos, own, -2, lane, W, T, U2, ov2);
  int sel2;
  double logW2;
  waveRosenbluth(U2, ov2, 0.0, K2, R.beta, uSel2, lane, sel2, logW2);
  if (retrace >= 0) sel2 = 0;
  if (sel2 < 0) return false;
  if (retrace >= 0 && __shfl(ov2 ? 1 : 0, 0, 64)) return false;
  const double Uorient = __shfl(U2, sel2, 64);
  if (lane == 0)
  {
    W.mol[3 * fb] = fx;
    W.mol[3 * fb + 1] = fy;
    W.mol[3 * fb + 2] = fz;
    for (int s = 0; s < n; ++s)
      for (int d = 0; d < 3; ++d) W.mol[3 * idx[s] + d] = W.pos[3 * (sel2 * n + s) + d];
  }
  waveSyncLds();
  logW = logW1 + logW2;
  Uext = Ufirst + Uorient;
  return true;
}

// ---------------------------------------------------------------------------------------------------------------
// flexible linear chains (united-atom alkanes), the counterpart of CBMC::growFlexibleMoleculeChainInsertion
// (src/raspakit/cbmc/flexible_molecules/insertion.cpp:32-126) and retraceFlexibleMoleculeChainDeletion
// (src/raspakit/cbmc/flexible_molecules/deletion.cpp:31-104): bond geometry sampled exactly from the bonded
// Boltzmann factors (Frenkel & Smit, Understanding Molecular Simulation, 2nd ed., sec. 13.3), Rosenbluth weights
// over the external and intra-molecular non-bonded energies.  Lane j builds trial j of a bead; its rejection samplers draw from the
// Philox slots SLOT_CHAIN + ((bead * 64 + trial) * 4 + kind) * MC_MAXREJ + attempt, the bead's selection from
// SLOT_CHAIN - 16 + bead.
// ---------------------------------------------------------------------------------------------------------------
constexpr uint32_t SLOT_CHAIN = 1024;

__device__ __forceinline__ void uniChain(const McRun& R, int sys, unsigned long long step, int bead, int trial,
                                         int kind, int attempt, double& a, double& b, uint32_t so = 0)
{
  uni(R, sys, step, so + SLOT_CHAIN + uint32_t(((bead * 64 + trial) * 4 + kind) * MC_MAXREJ + attempt), a, b);
}

__device__ __forceinline__ double chainTorsionEnergy(const Comp& c, double phi)
{
  return c.tors[0] + c.tors[1] * (1.0 + cos(phi)) + c.tors[2] * (1.0 - cos(2.0 * phi)) +
         c.tors[3] * (1.0 + cos(3.0 * phi));
}

// p(theta) ~ sin(theta) exp(-beta k (theta - theta0)^2 / 2): Gaussian proposal (Box-Muller), accepted with sin(theta)
__device__ double sampleBend(const McRun& R, const Comp& c, int sys, unsigned long long step, int bead, int trial,
                             uint32_t so)
{
  if (c.kbend <= 0.0)
  {
    double u, v;
    uniChain(R, sys, step, bead, trial, 1, 0, u, v, so);
    return acos(1.0 - 2.0 * u);
  }
  const double sd = 1.0 / sqrt(R.beta * c.kbend);
  for (int t = 0; t < MC_MAXREJ; ++t)
  {
    double g1, g2, a, b;
    uniChain(R, sys, step, bead, trial, 1, t, g1, g2, so);
    uniChain(R, sys, step, bead, trial, 2, t, a, b, so);
    const double th = c.theta0 + sd * sqrt(-2.0 * log(1.0 - g1)) * cos(2.0 * M_PI * g2);
    if (th > 0.0 && th < M_PI && a < sin(th)) return th;
  }
  return c.theta0;
}

// p(phi) ~ exp(-beta U_tors(phi)) on [-pi, pi): uniform proposal accepted with exp(-beta (U - U_min))
__device__ double sampleTorsion(const McRun& R, const Comp& c, int sys, unsigned long long step, int bead, int trial,
                                uint32_t so)
{
  for (int t = 0; t < MC_MAXREJ; ++t)
  {
    double u, a;
    uniChain(R, sys, step, bead, trial, 3, t, u, a, so);
    const double phi = 2.0 * M_PI * u - M_PI;
    if (a < exp(-R.beta * (chainTorsionEnergy(c, phi) - c.torsMin))) return phi;
  }
  return M_PI;
}

__device__ __forceinline__ void unit3(double v[3])
{
  const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] /= n;
  v[1] /= n;
  v[2] /= n;
}

__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3])
{
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// trial position of bead i given beads 0 .. i-1 (mol, 3 per bead): bond, sampled bend, azimuth or torsion
__device__ void chainTrial(const McRun& R, const Comp& c, int sys, unsigned long long step, int i, int j,
                           const double* mol, double out[3], uint32_t so)
{
  if (i == 1)
  {
    double u, v;
    uniChain(R, sys, step, i, j, 0, 0, u, v, so);
    const double z = 2.0 * u - 1.0, r = sqrt(fmax(0.0, 1.0 - z * z)), ps = 2.0 * M_PI * v;
    out[0] = mol[0] + c.bond * r * cos(ps);
    out[1] = mol[1] + c.bond * r * sin(ps);
    out[2] = mol[2] + c.bond * z;
    return;
  }
  const double* B = mol + 3 * (i - 2);
  const double* Cp = mol + 3 * (i - 1);
  double e[3] = {Cp[0] - B[0], Cp[1] - B[1], Cp[2] - B[2]};
  unit3(e);
  double m[3], n[3];
  if (i == 2)
  {
    const double ax[3] = {fabs(e[0]) < 0.9 ? 1.0 : 0.0, fabs(e[0]) < 0.9 ? 0.0 : 1.0, 0.0};
    cross3(e, ax, m);
    unit3(m);
    cross3(e, m, n);
  }
  else
  {
    const double* A = mol + 3 * (i - 3);
    const double b1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    cross3(b1, e, n);
    unit3(n);
    cross3(n, e, m);
  }
  const double th = sampleBend(R, c, sys, step, i, j, so);
  double ph;
  if (i == 2)
  {
    double u, v;
    uniChain(R, sys, step, i, j, 0, 0, u, v, so);
    ph = 2.0 * M_PI * u;
  }
  else
    ph = sampleTorsion(R, c, sys, step, i, j, so);
  const double ct = cos(th), st = sin(th), cp = cos(ph), sp = sin(ph);
  for (int d = 0; d < 3; ++d) out[d] = Cp[d] + c.bond * (-ct * e[d] + st * cp * m[d] + st * sp * n[d]);
}

// bend and torsion energy of bead i against beads i-3 .. i-1 (mol, 3 per bead)
__device__ __attribute__((always_inline)) double chainBonded(const Comp& c, const double* mol, int i)
{
  double e = 0.0;
  if (i >= 2)
  {
    const double* B = mol + 3 * (i - 2);
    const double* Cp = mol + 3 * (i - 1);
    const double* D = mol + 3 * i;
    double u[3] = {B[0] - Cp[0], B[1] - Cp[1], B[2] - Cp[2]}, v[3] = {D[0] - Cp[0], D[1] - Cp[1], D[2] - Cp[2]};
    unit3(u);
    unit3(v);
    const double th = acos(fmax(-1.0, fmin(1.0, u[0] * v[0] + u[1] * v[1] + u[2] * v[2])));
    const double d = th - c.theta0;
    e += 0.5 * c.kbend * d * d;
  }
  if (i >= 3)
  {
    const double* A = mol + 3 * (i - 3);
    const double* B = mol + 3 * (i - 2);
    const double* Cp = mol + 3 * (i - 1);
    const double* D = mol + 3 * i;
    const double b1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, b2[3] = {Cp[0] - B[0], Cp[1] - B[1], Cp[2] - B[2]},
                 b3[3] = {D[0] - Cp[0], D[1] - Cp[1], D[2] - Cp[2]};
    double n1[3], n2[3], m1[3], e2[3] = {b2[0], b2[1], b2[2]};
    cross3(b1, b2, n1);
    cross3(b2, b3, n2);
    unit3(e2);
    cross3(n1, e2, m1);
    const double x = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
    const double y = m1[0] * n2[0] + m1[1] * n2[1] + m1[2] * n2[2];
    e += chainTorsionEnergy(c, atan2(y, x));
  }
  return e;
}

// intra-molecular LJ of a site of bead i against beads 0 .. i-4
__device__ double chainIntra(const SysDev& S, const Comp& c, const double* mol, int i, const double* p, bool& ov)
{
  double e = 0.0;
  for (int b = 0; b + 3 < i; ++b)
  {
    const double dx = p[0] - mol[3 * b], dy = p[1] - mol[3 * b + 1], dz = p[2] - mol[3 * b + 2];
    const double rr = dx * dx + dy * dy + dz * dz;
    if (rr < S.rcv2Mol)
    {
      const double v = rgd::vdw0(S.pairs[c.type[b] * S.ntypes + c.type[i]], 1.0, rr);
      ov = ov || v > S.overlap;
      e += v;
    }
  }
  return e;
}

// Chain growth (retrace >= 0: the existing molecule is trial 0 of every step): the first bead as for a rigid
// molecule, then kOrient trials for each of beads 1 .. na-1.  Writes the chain into W.mol; Uext is the external
// plus intra-molecular (non-bonded and bonded) energy.
__device__ bool growChain(const McRun& R, const SysDev& S, const McSys& Y, const Comp& c, int sys, int nads,
                          unsigned long long step, const GrowMode G, int lane, WaveScratch& W, const Team& T,
                          double& logW, double& Uext, double& logStoredR)
{
  const int K = R.kOrient;
  const int retrace = G.retrace;
  const int own = G.own;
  double fx, fy, fz, Ufirst, lw, uSel2;
  if (!firstBeadStep(R, S, Y, c, sys, nads, step, G, lane, W, T, fx, fy, fz, Ufirst, lw, uSel2, logStoredR))
    return false;
  if (lane == 0)
  {
    W.mol[0] = fx;
    W.mol[1] = fy;
    W.mol[2] = fz;
  }
  waveSyncLds();
  logW = lw;
  Uext = Ufirst;
  for (int i = 1; i < c.na; ++i)
  {
    double p[3] = {0.0, 0.0, 0.0};
    if (lane < K)
    {
      if (retrace >= 0 && lane == 0)
      {
        const int a0 = retrace * R.slots + i;
        p[0] = S.ax[a0];
        p[1] = S.ay[a0];
        p[2] = S.az[a0];
      }
      else
        chainTrial(R, c, sys, step, i, lane, W.mol, p, G.so);
      W.pos[3 * lane] = p[0];
      W.pos[3 * lane + 1] = p[1];
      W.pos[3 * lane + 2] = p[2];
    }
    waveSyncLds();
    double Ui;
    bool ovi;
    const int bi[1] = {i};
    trialSetEnergies(S, nads, c, bi, 1, K, W.pos, own, -2, lane, W, T, Ui, ovi);
    if (lane < K) Ui += chainIntra(S, c, W.mol, i, p, ovi);
    double us, unused;
    uni(R, sys, step, G.so + SLOT_CHAIN - 16 + i, us, unused);
    int sel;
    waveRosenbluth(Ui, ovi, 0.0, K, R.beta, us, lane, sel, lw);
    if (retrace >= 0) sel = 0;
    if (sel < 0) return false;
    if (retrace >= 0 && __shfl(ovi ? 1 : 0, 0, 64)) return false;
    const double sx = __shfl(p[0], sel, 64), sy = __shfl(p[1], sel, 64), sz = __shfl(p[2], sel, 64);
    const double Usel = __shfl(Ui, sel, 64);
    waveSyncLds();  // every lane has read W.pos / W.mol of this step
    if (lane == 0)
    {
      W.mol[3 * i] = sx;
      W.mol[3 * i + 1] = sy;
      W.mol[3 * i + 2] = sz;
    }
    waveSyncLds();
    logW += lw;
    if (logW < R.logMinW) return false;
    Uext += Usel + chainBonded(c, W.mol, i);
  }
  return true;
}

// CHAIN = false compiles the chain growth out of kernels whose components are all rigid (fewer registers).
template <bool CHAIN>
__device__ __forceinline__ bool growMolecule(const McRun& R, const SysDev& S, const McSys& Y, const Comp& c, int sys,
                                             int nads, unsigned long long step, const GrowMode G, int lane,
                                             WaveScratch& W, const Team& T, double& logW, double& Uext,
                                             double& logStoredR)
{
  if (CHAIN && c.flexible) return growChain(R, S, Y, c, sys, nads, step, G, lane, W, T, logW, Uext, logStoredR);
  return growRigid(R, S, Y, c, sys, nads, step, G, lane, W, T, logW, Uext, logStoredR);
}

// A growth of a new molecule (retrace < 0) or the retrace of molecule 'retrace' for an insertion or deletion.
__device__ __forceinline__ GrowMode swapMode(int retrace)
{
  GrowMode G;
  G.retrace = retrace;
  G.own = retrace >= 0 ? retrace : -2;
  G.so = 0;
  G.reins = 0;
  G.logStoredR = -INFINITY;
  return G;
}


// Ewald energy change of inserting (+1) or removing (-1) the atoms in W.mol of component c, incl. the constant
// self and intra-molecular exclusion terms of the rigid body.
__device__ __attribute__((always_inline)) double ewaldInsertRemove(const McRun& R, const SysDev& S, double netCharge, const Comp& c, int sign, int lane,
                                    WaveScratch& W, const Team& T, bool writeTrial)
{
  if (!S.useCharge || !c.charged) return 0.0;
  if (lane < c.na)
  {
    W.at[4 * lane] = W.mol[3 * lane];
    W.at[4 * lane + 1] = W.mol[3 * lane + 1];
    W.at[4 * lane + 2] = W.mol[3 * lane + 2];
    W.at[4 * lane + 3] = sign * c.q[lane];
  }
  waveSyncLds();
  const double df = waveEwaldDelta(S, c.na, W.at, W.tab, lane, writeTrial, W, T);
  // net charge: neutral molecules leave it unchanged; otherwise (Q + dq)^2 - Q^2 times the single-ion term
  const double dq = sign * c.qsum;
  const double uIon = -(S.singleIon - S.coulomb * S.alpha / sqrt(M_PI));
  const double qn = netCharge + dq;
  return df + uIon * (qn * qn - netCharge * netCharge) - sign * c.eIntra;
}

// ---------------------------------------------------------------------------------------------------------------
// Widom insertions: one wavefront per insertion i (the "step" of the stream), system state unchanged.
// ---------------------------------------------------------------------------------------------------------------
template <bool CHAIN>
__global__ __launch_bounds__(256) void k_widom(const McRun* __restrict__ Rp, const SysDev* __restrict__ Sp,
                                               const McSys* __restrict__ Yp, int comp, long long first, int count,
                                               double* outLogW, double* outU, double* outEwald)
{
  __shared__ WaveScratch scratch[4];
  const McRun& R = *Rp;
  // the wavefront index is uniform over the wavefront; readfirstlane tells the compiler so, and the per-system
  // structures are then read through the scalar cache instead of per-lane vector loads
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + w;
  loadPairTable(*Sp);
  if (i >= count) return;
  const SysDev& S = *Sp;
  const McSys& Y = *Yp;
  const Comp& c = R.comp[comp];
  double logW = -INFINITY, U = 0.0, de = 0.0;
  const Team T{1, 0, &scratch[w]};
  double unusedR;
  if (growMolecule<CHAIN>(R, S, Y, c, 0, S.nads, (unsigned long long)(first + i), swapMode(-1), lane, scratch[w], T,
                          logW, U, unusedR))
    de = ewaldInsertRemove(R, S, S.netCharge, c, +1, lane, scratch[w], T, false);
  else
  {
    // a failed growth (every trial of a step overlapping) has weight zero, whatever partial sums it left
    logW = -INFINITY;
    U = 0.0;
  }
  if (lane == 0)
  {
    outLogW[i] = logW;
    outU[i] = U;
    outEwald[i] = de;
  }
}

// Widom insertions of a single-atom, uncharged component whose pseudo-atom has an interpolation grid: 64 / kFirst
// insertions per wavefront, one first-bead trial per lane (the general kernel above keeps kFirst lanes of a
// wavefront busy).  Trial positions come from the same Philox streams as firstBeadStep (system 0, step first + i,
// trial slot k); the trial energy is the inter-molecular LJ against the system's adsorbates plus the grid value, as
// in trialSetEnergies; the Rosenbluth weight and the selection follow waveRosenbluth within each group of kFirst
// lanes, summing in trial order (src/raspakit/cbmc/util.cpp:15-48, multiple_first_bead.cpp:20-76).  The Ewald
// difference of an uncharged atom is zero.
__global__ __launch_bounds__(256) void k_widom_single(const McRun* __restrict__ Rp, const SysDev* __restrict__ Sp,
                                                      const McSys* __restrict__ Yp, int comp, long long first,
                                                      int count, double* outLogW, double* outU, double* outEwald)
{
  const McRun& R = *Rp;
  const SysDev& S = *Sp;
  const McSys& Y = *Yp;
  const Comp& c = R.comp[comp];
  const int K = R.kFirst;
  const int G = 64 / K;  // insertions per wavefront
  const int lane = threadIdx.x & 63;
  const long long wave = (static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const int j = lane / K, k = lane - j * K;
  const long long i = wave * G + j;
  const bool live = j < G && i < count;
  const int type = c.type[0];
  double U = 0.0, logBias = 0.0;
  bool ov = false;
  unsigned long long step = 0;
  if (live)
  {
    step = static_cast<unsigned long long>(first + i);
    double p[3];
    logBias = firstBeadTrial(R, S, Y, comp, 0, step, k, p);
    for (int a = 0; a < S.nads; ++a)  // the same atom for every lane: uniform loads
    {
      const int at = S.at[a];
      if (at < 0) continue;
      double dx = S.ax[a] - p[0], dy = S.ay[a] - p[1], dz = S.az[a] - p[2];
      rgd::pbc(S.box, dx, dy, dz);
      const double rr = dx * dx + dy * dy + dz * dz;
      if (rr < S.rcv2Mol)
      {
        const double v = rgd::vdw0(S.pairs[at * S.ntypes + type], 1.0, rr);
        ov = ov || v > S.overlap;
        U += v;
      }
    }
    const double g = rgd::interpolate(S.grids[S.gridOfType[type]], p[0], p[1], p[2]);
    ov = ov || g > S.overlap;
    U += g;
  }
  // Rosenbluth weight and selection of the group, with every lane of the wavefront active (shuffles read only
  // active lanes): the group's maximum, each trial's weight, then the sum and the cumulative search in trial order
  const bool valid = live && !ov;
  const double lbf = valid ? -R.beta * U + logBias : -INFINITY;
  const unsigned long long vmask = __ballot(valid);
  const int g0 = min(j, G - 1) * K;  // lanes past the last group follow a real group and discard the result
  double mx = -INFINITY;
  for (int t = 0; t < K; ++t) mx = fmax(mx, __shfl(lbf, g0 + t, 64));
  const double w = valid ? exp(lbf - mx) : 0.0;
  double uSel1 = 0.0, uSel2 = 0.0;
  if (live) uni(R, 0, step, SLOT_SELECT, uSel1, uSel2);
  (void)uSel2;
  double sum = 0.0;
  for (int t = 0; t < K; ++t) sum += __shfl(w, g0 + t, 64);
  const double ws = uSel1 * sum;
  double pre = 0.0;
  int sel = -1, lastValid = -1;
  for (int t = 0; t < K; ++t)
  {
    pre += __shfl(w, g0 + t, 64);
    if ((vmask >> (g0 + t)) & 1ull)
    {
      lastValid = t;
      if (sel < 0 && !(pre < ws)) sel = t;
    }
  }
  if (sel < 0) sel = lastValid;
  const double Usel = __shfl(U, g0 + max(sel, 0), 64);
  if (!live || k != 0) return;
  double logW = -INFINITY;
  if (lastValid >= 0) logW = mx + log(sum) - log(double(K));
  const bool ok = lastValid >= 0 && logW >= R.logMinW;
  outLogW[i] = ok ? logW : -INFINITY;
  outU[i] = ok ? Usel : 0.0;
  outEwald[i] = 0.0;
}

// ---------------------------------------------------------------------------------------------------------------
// GCMC: one wavefront per system, 'cycles' MC cycles per launch.  The wave-uniform state lives in registers (every
// lane holds the same values) and is written back once by lane 0; the adsorbate arrays, the molecule component
// table and the structure factors stay in global memory and are ordered with workgroup-scope fences.
// ---------------------------------------------------------------------------------------------------------------
struct Live
{
  int nmol;
  int ncomp[MC_MAXCOMP];
  int nads;
  double U;
  double netCharge;
  long long step;
  long long cycle;
  long long trials[5];
  long long accepted[5];
};

// Orientation-averaged Boltzmann factor of a rigid molecule with its first bead at each of n points (fractional
// coordinates of the box), from the interpolation grids only: log W(s) = log (1/M) sum_o exp(-beta U(s, R_o)) over M
// random orientations (Philox stream (point, orientation) of 'seed'); -inf when every orientation overlaps.  W(s) is
// the first-bead density that the orientation trials of growRigid reward, the training target of the learned
// proposal.  M is a power of two up to 64; a group of M lanes takes one point.
__global__ __launch_bounds__(256) void k_molecule_weights(const SysDev* __restrict__ Sp, Comp c, double beta, int n,
                                                          const double* __restrict__ frac, int M, unsigned long long seed,
                                                          double* __restrict__ logW)
{
  const SysDev& S = *Sp;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = t / M, o = t % M;
  const bool active = i < n;
  double e = 0.0;
  bool ov = false;
  if (active)
  {
    const double f[3] = {frac[3 * size_t(i)], frac[3 * size_t(i) + 1], frac[3 * size_t(i) + 2]};
    double p[3];
    for (int k = 0; k < 3; ++k) p[k] = S.box.cell[k] * f[0] + S.box.cell[3 + k] * f[1] + S.box.cell[6 + k] * f[2];
    double u1, u2, u3, u4, Rm[9];
    rgd::uniform2(seed, uint32_t(i), uint32_t(o), 0u, 0u, u1, u2);
    rgd::uniform2(seed, uint32_t(i), uint32_t(o), 0u, 1u, u3, u4);
    randomRotation(u1, u2, u3, Rm);
    for (int a = 0; a < c.na; ++a)
    {
      const double* r = c.ref[a];
      const double x = p[0] + Rm[0] * r[0] + Rm[1] * r[1] + Rm[2] * r[2];
      const double y = p[1] + Rm[3] * r[0] + Rm[4] * r[1] + Rm[5] * r[2];
      const double z = p[2] + Rm[6] * r[0] + Rm[7] * r[1] + Rm[8] * r[2];
      const double g = rgd::interpolate(S.grids[S.gridOfType[c.type[a]]], x, y, z);
      ov = ov || g > S.overlap;
      e += g;
      if (S.useCharge && S.coulGrid >= 0) e += c.q[a] * rgd::interpolate(S.grids[S.coulGrid], x, y, z);
    }
  }
  const double a = (active && !ov) ? -beta * e : -INFINITY;
  double m = a;
  for (int off = 1; off < M; off <<= 1) m = fmax(m, __shfl_xor(m, off, 64));
  double sum = (m == -INFINITY) ? 0.0 : exp(a - m);
  for (int off = 1; off < M; off <<= 1) sum += __shfl_xor(sum, off, 64);
  if (active && o == 0) logW[i] = (m == -INFINITY) ? -INFINITY : m + log(sum / M);
}

// Index of the kth molecule (in slot order) of component ci among the first nmol, or -1: 64 slots per ballot.
__device__ __forceinline__ int kthMoleculeOfComponent(const McSys& Y, int nmol, int ci, int kth, int lane)
{
  int seen = 0;
  for (int base = 0; base < nmol; base += 64)
  {
    const int q = base + lane;
    const unsigned long long b = __ballot(q < nmol && Y.comp[q] == ci);
    const int cnt = __popcll(b);
    if (kth < seen + cnt)
    {
      unsigned long long bb = b;
      for (int r = 0; r < kth - seen; ++r) bb &= bb - 1;
      return base + __ffsll((long long)bb) - 1;
    }
    seen += cnt;
  }
  return -1;
}

// Inlined into k_gcmc: as an outlined function (the compiler's choice once the step grew) it reached the kernel's LDS
// descriptors and scratch through generic pointers, one flat access per field.
// REINS: the kernel instance carries the reinsertion move (runs with p_reinsertion > 0; without it the move's code,
// and the registers it holds, are left out)
template <bool CHAIN, bool REINS>
__device__ __attribute__((always_inline)) void gcmcStep(const McRun& R, const SysDev& S, const McSys& Y, int sys, Live& L, int lane, WaveScratch& W,
                         const Team& T)
{
  reloadBarrier();
  const unsigned long long step = (unsigned long long)L.step;
  double uMove, uComp, uMol, uAcc;
  uni(R, sys, step, SLOT_MOVE, uMove, uComp);
  uni(R, sys, step, SLOT_PICK, uMol, uAcc);
  // the component is drawn uniformly (System::randomComponent, src/raspakit/system.ixx:372-375); the molecule that a
  // translation, rotation or deletion acts on is a random molecule of that component
  const int ci = min(int(uComp * R.ncomp), R.ncomp - 1);
  const Comp& c = R.comp[ci];
  const double pTrans = R.pTranslation, pRot = pTrans + R.pRotation;
  if (uMove < pRot)
  {
    // translation or rotation of a random molecule of the component
    const bool isTrans = uMove < pTrans;
    L.trials[isTrans ? 2 : 3] += 1;
    if (L.ncomp[ci] == 0) return;
    const int m = kthMoleculeOfComponent(Y, L.nmol, ci, min(int(uMol * L.ncomp[ci]), L.ncomp[ci] - 1), lane);
    if (m < 0) return;
    const Comp& cm = c;
    double d0, d1, d2, d3;
    uni(R, sys, step, SLOT_DISP, d0, d1);
    uni(R, sys, step, SLOT_DISP + 1, d2, d3);
    // old positions into W.pos[0 .. na), new ones into W.pos[na .. 2na)
    if (lane < cm.na)
    {
      const int a0 = m * R.slots + lane;
      const double ox = S.ax[a0], oy = S.ay[a0], oz = S.az[a0];
      W.pos[3 * lane] = ox;
      W.pos[3 * lane + 1] = oy;
      W.pos[3 * lane + 2] = oz;
      double nx, ny, nz;
      if (isTrans)
      {
        nx = ox + R.maxTranslation * (2.0 * d0 - 1.0);
        ny = oy + R.maxTranslation * (2.0 * d1 - 1.0);
        nz = oz + R.maxTranslation * (2.0 * d2 - 1.0);
      }
      else
      {
        // rotation by theta about a random unit axis through the first bead (Rodrigues' formula)
        const int f0 = m * R.slots + cm.first;
        const double cx = S.ax[f0], cy = S.ay[f0], cz = S.az[f0];
        const double zc = 2.0 * d0 - 1.0, ph = 2.0 * M_PI * d1, sr = sqrt(fmax(0.0, 1.0 - zc * zc));
        const double kx = sr * cos(ph), ky = sr * sin(ph), kz = zc;
        const double th = R.maxRotation * (2.0 * d2 - 1.0);
        const double ct = cos(th), st = sin(th);
        const double vx = ox - cx, vy = oy - cy, vz = oz - cz;
        const double kv = kx * vx + ky * vy + kz * vz;
        nx = cx + vx * ct + (ky * vz - kz * vy) * st + kx * kv * (1.0 - ct);
        ny = cy + vy * ct + (kz * vx - kx * vz) * st + ky * kv * (1.0 - ct);
        nz = cz + vz * ct + (kx * vy - ky * vx) * st + kz * kv * (1.0 - ct);
      }
      W.pos[3 * (cm.na + lane)] = nx;
      W.pos[3 * (cm.na + lane) + 1] = ny;
      W.pos[3 * (cm.na + lane) + 2] = nz;
    }
    waveSyncLds();
    // grid parts of the old (sites 0 .. na-1) and new (sites na .. 2na-1) sites: lane 2 site + g takes term g of the
    // site (g = 0 its type's grid, g = 1 its charge times the Coulomb grid), so all gathers are in flight at once;
    // lane 'site' then adds its two terms
    double gl = 0.0;
    bool gov = false;
    if (S.gridOfType)
    {
      double v = 0.0;
      bool o = false;
      if (lane < 4 * cm.na)
      {
        const int site = lane >> 1;
        const bool coulTerm = (lane & 1) != 0;
        const int a = site < cm.na ? site : site - cm.na;
        const int gi = S.gridOfType[cm.type[a]];
        const int gsel = coulTerm ? ((S.useCharge && S.coulGrid >= 0) ? S.coulGrid : -1) : gi;
        if (gi >= 0 && gsel >= 0)
        {
          const double* p = W.pos + 3 * site;
          v = rgd::interpolate(S.grids[gsel], p[0], p[1], p[2]);
          if (coulTerm)
            v *= cm.q[a];
          else
            o = v > S.overlap;
        }
      }
      const int src = min(2 * lane, 62);
      const double own = __shfl(v, src, 64), coul = __shfl(v, src + 1, 64);
      const bool oo = __shfl(o ? 1 : 0, src, 64) != 0;
      if (lane < 2 * cm.na)
      {
        gl = own + coul;
        gov = oo;
      }
    }
    const bool newLane = lane >= cm.na && lane < 2 * cm.na;
    double Uold = rgd::waveSum(lane < cm.na ? gl : 0.0);
    double Unew = rgd::waveSum(newLane ? gl : 0.0);
    bool ovNew = __ballot(newLane && gov) != 0ull;
    // the other molecules (and framework atoms of types without a grid): the 2 na sites in one batch
    if (lane < 2 * cm.na)
    {
      const int a = lane < cm.na ? lane : lane - cm.na;
      W.sType[lane] = cm.type[a];
      W.sQ[lane] = cm.q[a];
    }
    waveSyncLds();
    siteSetEnergies(S, L.nads, 2 * cm.na, W.pos, m, -2, lane, W, T);
    for (int a = 0; a < cm.na; ++a)
    {
      Uold += W.siteE[a];
      Unew += W.siteE[cm.na + a];
      ovNew = ovNew || W.siteOv[cm.na + a] != 0;
    }
    if (ovNew) return;
    double de = 0.0;
    const bool ew = S.useCharge && cm.charged;
    if (ew)
    {
      if (lane < 2 * cm.na)
      {
        const bool isNew = lane >= cm.na;
        const int a = isNew ? lane - cm.na : lane;
        W.at[4 * lane] = W.pos[3 * lane];
        W.at[4 * lane + 1] = W.pos[3 * lane + 1];
        W.at[4 * lane + 2] = W.pos[3 * lane + 2];
        W.at[4 * lane + 3] = isNew ? cm.q[a] : -cm.q[a];
      }
      waveSyncLds();
      de = waveEwaldDelta(S, 2 * cm.na, W.at, W.tab, lane, true, W, T);
    }
    const double dU = Unew - Uold + de;
    if (uAcc < exp(-R.beta * dU))
    {
      if (T.wi == 0 && lane < cm.na)  // adsorbate arrays: written by the team's first wavefront
      {
        const int a0 = m * R.slots + lane;
        S.ax[a0] = W.pos[3 * (cm.na + lane)];
        S.ay[a0] = W.pos[3 * (cm.na + lane) + 1];
        S.az[a0] = W.pos[3 * (cm.na + lane) + 2];
      }
      if (ew) waveAcceptEwald(S, lane, T);
      L.U += dU;
      L.accepted[isTrans ? 2 : 3] += 1;
    }
    return;
  }
  // CBMC insertion (src/raspakit/mc_moves/component/insertion_cbmc.cpp:34-200), deletion
  // (src/raspakit/mc_moves/component/deletion_cbmc.cpp:67-174) or reinsertion (MC_Moves::reinsertionMove,
  // src/raspakit/mc_moves/component/reinsertion.cpp:34-214) of a molecule of the component.  One growth call site
  // serves all three: an insertion grows, a deletion retraces molecule m, a reinsertion regrows m with its own atoms
  // left out of the background (pass 0) and then retraces its old configuration with the growth's unselected
  // first-bead trials (pass 1); a flexible chain is regrown and retraced bead by bead, its own atoms left out of
  // the background in both passes.
  const double pRe = pRot + R.pReinsertion;
  const bool reins = REINS && uMove < pRe;
  const bool insert = !reins && uMove < pRe + 0.5 * (1.0 - pRe);
  L.trials[reins ? 4 : (insert ? 0 : 1)] += 1;
  const int nc = L.ncomp[ci];
  int m = -1;
  if (insert)
  {
    if (L.nmol >= R.maxMolecules) return;
  }
  else
  {
    if (nc == 0) return;
    m = kthMoleculeOfComponent(Y, L.nmol, ci, min(int(uMol * nc), nc - 1), lane);
    if (m < 0) return;
  }
  double logW = 0.0, Uext = 0.0, logWn = 0.0, Un = 0.0, logStoredR = -INFINITY;
  double nx = 0.0, ny = 0.0, nz = 0.0;  // new configuration of a reinsertion (lane a holds atom a)
  for (int pass = 0; pass < (reins ? 2 : 1); ++pass)
  {
    GrowMode G = swapMode(insert ? -1 : m);
    if (reins)
    {
      G.retrace = pass == 0 ? -1 : m;
      G.so = pass == 0 ? 0u : uint32_t(SLOT_REINS);
      G.reins = pass + 1;
      G.logStoredR = logStoredR;
    }
    double stored = -INFINITY;
    if (!growMolecule<CHAIN>(R, S, Y, c, sys, L.nads, step, G, lane, W, T, logW, Uext, stored)) return;
    if (reins && pass == 0)
    {
      logStoredR = stored;
      logWn = logW;
      Un = Uext;
      if (lane < c.na)
      {
        nx = W.mol[3 * lane];
        ny = W.mol[3 * lane + 1];
        nz = W.mol[3 * lane + 2];
      }
      waveSyncLds();
    }
  }
  if (reins)
  {
    double de = 0.0;
    const bool ew = S.useCharge && c.charged;
    if (ew)
    {
      // Fourier difference of the old atoms (-q) and the new ones (+q), as for a translation
      const int src = lane >= c.na ? lane - c.na : lane;
      const double sx = __shfl(nx, src, 64), sy = __shfl(ny, src, 64), sz = __shfl(nz, src, 64);
      if (lane < 2 * c.na)
      {
        const bool isNew = lane >= c.na;
        const int a0 = m * R.slots + src;
        W.at[4 * lane] = isNew ? sx : S.ax[a0];
        W.at[4 * lane + 1] = isNew ? sy : S.ay[a0];
        W.at[4 * lane + 2] = isNew ? sz : S.az[a0];
        W.at[4 * lane + 3] = isNew ? c.q[src] : -c.q[src];
      }
      waveSyncLds();
      de = waveEwaldDelta(S, 2 * c.na, W.at, W.tab, lane, true, W, T);
    }
    // Pacc = exp(-beta dE_Fourier) W(new) / W(old) (reinsertion.cpp:189-190); logW, Uext: the old configuration
    if (log(uAcc) < -R.beta * de + logWn - logW)
    {
      if (ew) waveAcceptEwald(S, lane, T);
      if (T.wi == 0 && lane < c.na)
      {
        const int a0 = m * R.slots + lane;
        S.ax[a0] = nx;
        S.ay[a0] = ny;
        S.az[a0] = nz;
      }
      L.U += Un - Uext + de;
      L.accepted[4] += 1;
    }
    return;
  }
  const double fugVol = R.beta * c.fugacity * S.box.volume;
  const double de = ewaldInsertRemove(R, S, L.netCharge, c, insert ? +1 : -1, lane, W, T, true);
  if (insert)
  {
    // Pacc = exp(-beta dE_Ewald) beta f V / (N + 1) W / W_IG (insertion_cbmc.cpp:157-168); W_IG = 1 for rigid bodies
    const double logPacc = -R.beta * de + log(fugVol / double(nc + 1)) + logW - c.logWIG;
    if (log(uAcc) < logPacc)
    {
      const int mn = L.nmol;
      if (T.wi == 0 && lane < R.slots)
      {
        const int a0 = mn * R.slots + lane;
        const bool real = lane < c.na;
        S.ax[a0] = real ? W.mol[3 * lane] : 0.0;
        S.ay[a0] = real ? W.mol[3 * lane + 1] : 0.0;
        S.az[a0] = real ? W.mol[3 * lane + 2] : 0.0;
        S.aq[a0] = real ? c.q[lane] : 0.0;
        S.at[a0] = real ? c.type[lane] : -1;
        S.amol[a0] = mn;
      }
      if (T.wi == 0 && lane == 0) Y.comp[mn] = ci;
      if (S.useCharge && c.charged) waveAcceptEwald(S, lane, T);
      L.nmol = mn + 1;
      L.ncomp[ci] = nc + 1;
      L.nads = (mn + 1) * R.slots;
      L.netCharge += c.qsum;
      L.U += Uext + de;
      L.accepted[0] += 1;
    }
    return;
  }
  // Pacc = exp(-beta dE_Ewald) N / (beta f V) W_IG / W (deletion_cbmc.cpp:159-161)
  const double logPacc = -R.beta * de + log(double(nc) / fugVol) - logW + c.logWIG;
  if (log(uAcc) < logPacc)
  {
    if (S.useCharge && c.charged) waveAcceptEwald(S, lane, T);
    // move the last molecule into slot m
    const int last = L.nmol - 1;
    if (m != last)
    {
      if (T.wi == 0 && lane < R.slots)
      {
        const int d = m * R.slots + lane, s = last * R.slots + lane;
        S.ax[d] = S.ax[s];
        S.ay[d] = S.ay[s];
        S.az[d] = S.az[s];
        S.aq[d] = S.aq[s];
        S.at[d] = S.at[s];
      }
      if (T.wi == 0 && lane == 0) Y.comp[m] = Y.comp[last];
    }
    L.nmol = last;
    L.ncomp[ci] = nc - 1;
    L.nads = last * R.slots;
    L.netCharge -= c.qsum;
    L.U += -Uext + de;
    L.accepted[1] += 1;
  }
}

// The cost of one cycle of a system in a balanced launch: its steps times (molecules + the run's overhead constant),
// a model of a step's work (the inter-molecular energies grow with the molecules, the grid and Ewald parts do not).
__device__ __forceinline__ double balancedCycleCost(const McRun& R, int nmol)
{
  return double(max(R.minStepsPerCycle, nmol) * R.ncomp) * (double(nmol) + R.balanceOverhead);
}

// Before each balanced launch: the largest single-cycle cost over the systems, and the launch's sampled index.  One
// workgroup; thread 0 writes the control record.
__global__ __launch_bounds__(256) void k_mc_budget(const McRun* __restrict__ Rp, const McSys* __restrict__ Yarr,
                                                   int nsys, McCtl* __restrict__ ctl, int sample)
{
  __shared__ double red[4];
  const McRun& R = *Rp;
  double m = 0.0;
  for (int s = threadIdx.x; s < nsys; s += blockDim.x) m = fmax(m, balancedCycleCost(R, Yarr[s].nmol));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0)
  {
    for (int w = 1; w < int(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
    ctl->budget = m;
    if (sample)
    {
      ctl->sampledIndex = ctl->sampledCount;
      ctl->sampledCount += 1;
    }
  }
}

// One wavefront per system, blockDim.x / 64 (1, 2 or 4) systems per workgroup: with few systems, one per workgroup
// spreads them over all compute units.  Dynamic LDS: one WaveScratch per wavefront.
// BAL: balanced launch (rigid components only): after its 'cycles' cycles a system keeps running whole cycles while
// their summed cost stays within the budget k_mc_budget wrote, so lightly loaded systems do more cycles in the time
// the most loaded one takes for its first; the samples of the launch go to the block of its sampled index.
// Otherwise every system runs 'cycles' cycles and a sample goes to the block of the system's sampled-cycle count.
template <bool CHAIN, bool BAL, bool REINS>
__global__ __launch_bounds__(256) void k_gcmc(const McRun* __restrict__ Rp, SysDev* __restrict__ Sarr,
                                              McSys* __restrict__ Yarr, int nsys, int cycles, int sample, int nw,
                                              const McCtl* __restrict__ ctl)
{
  extern __shared__ WaveScratch scratch[];
  // each wavefront's system descriptors, copied into LDS: the moves read their fields constantly, and LDS reads
  // stay cheap across the fences that order the wavefront's LDS traffic (global reads would go back to memory)
  __shared__ SysDev sdev[4];
  __shared__ McSys ysys[4];
  const McRun& R = *Rp;
  // the wavefront index is uniform over the wavefront; readfirstlane tells the compiler so, and the per-system
  // structures are then read through the scalar cache instead of per-lane vector loads
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  // nw wavefronts (1 or MC_PARTS) per system, blockDim.x / (64 nw) systems per workgroup
  const int slot = w / nw, wi = w % nw;
  const int sys = blockIdx.x * (blockDim.x / (64 * nw)) + slot;
  loadPairTable(Sarr[0]);  // the systems of a launch share the force field
  if (sys >= nsys) return;  // a whole team returns: its wavefronts share sys
  const Team T{nw, wi, &scratch[slot * nw]};
  {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(&Sarr[sys]);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&sdev[w]);
    for (int i = lane; i < int(sizeof(SysDev) / 4); i += 64) dst[i] = src[i];
    const uint32_t* ysrc = reinterpret_cast<const uint32_t*>(&Yarr[sys]);
    uint32_t* ydst = reinterpret_cast<uint32_t*>(&ysys[w]);
    for (int i = lane; i < int(sizeof(McSys) / 4); i += 64) ydst[i] = ysrc[i];
    waveSyncLds();
  }
  const SysDev& S = sdev[w];
  const McSys& Y = ysys[w];
  Live L;
  L.nmol = Y.nmol;
  for (int q = 0; q < MC_MAXCOMP; ++q) L.ncomp[q] = Y.ncomp[q];
  L.nads = S.nads;
  L.U = Y.U;
  L.netCharge = S.netCharge;
  L.step = Y.step;
  L.cycle = Y.cycle;
  for (int q = 0; q < 5; ++q)
  {
    L.trials[q] = Y.stat.trials[q];
    L.accepted[q] = Y.stat.accepted[q];
  }
  const double budget = BAL ? ctl->budget : 0.0;
  double spent = 0.0;
  for (int cyc = 0;; ++cyc)
  {
    // steps per cycle = max(20, N) times the number of components (src/raspakit/drivers/monte_carlo.cpp:301)
    const int nsteps = max(R.minStepsPerCycle, L.nmol) * R.ncomp;
    if (cyc >= cycles)
    {
      // L.nmol is the same on every wavefront of a team, so the whole team leaves together
      if (!BAL || cyc >= MC_MAX_BALANCED_CYCLES) break;
      const double cost = balancedCycleCost(R, L.nmol);
      if (spent + cost > budget) break;
    }
    if (BAL) spent += balancedCycleCost(R, L.nmol);
    for (int s = 0; s < nsteps; ++s)
    {
      gcmcStep<CHAIN, REINS>(R, S, Y, sys, L, lane, scratch[w], T);
      // the team's first wavefront wrote the adsorbate arrays; the next step reads them on every wavefront
      if (nw > 1)
        __syncthreads();
      else
        waveSyncGlobal();
      L.step += 1;
    }
    if (sample)
    {
      // block sums live in global memory, read and written by lane 0 of the team's first wavefront only
      if (wi == 0 && lane == 0)
      {
        const long long ci = BAL ? (long long)ctl->sampledIndex : L.cycle;
        const int b = min(MC_BLOCKS - 1, int(ci * MC_BLOCKS / max(1LL, R.productionCycles)));
        McBlock& B = Yarr[sys].block[b];
        B.count += 1.0;
        B.U += L.U;
        for (int q = 0; q < R.ncomp; ++q)
        {
          const double n = L.ncomp[q];
          B.N[q] += n;
          B.UN[q] += L.U * n;
          for (int r = 0; r < R.ncomp; ++r) B.NN[q][r] += n * L.ncomp[r];
        }
      }
      L.cycle += 1;
    }
  }
  if (wi == 0 && lane == 0)
  {
    McSys& Yg = Yarr[sys];
    SysDev& Sg = Sarr[sys];
    Yg.nmol = L.nmol;
    for (int q = 0; q < MC_MAXCOMP; ++q) Yg.ncomp[q] = L.ncomp[q];
    Sg.nads = L.nads;
    Yg.U = L.U;
    Sg.netCharge = L.netCharge;
    Yg.step = L.step;
    Yg.cycle = L.cycle;
    for (int q = 0; q < 5; ++q)
    {
      Yg.stat.trials[q] = L.trials[q];
      Yg.stat.accepted[q] = L.accepted[q];
    }
  }
}

// Recomputes the adsorbate energy of every system from scratch (framework-molecule + molecule-molecule pairs, and
// the Ewald Fourier energy from the stored structure factor with the self and exclusion terms): the drift check of
// the reference's integration tests (tests/raspa3-tests/mc_muvt_grids_drift.cpp:93-115) on the device.
__global__ __launch_bounds__(256) void k_total_energy(const McRun* __restrict__ Rp, const SysDev* __restrict__ Sarr,
                                                      const McSys* __restrict__ Yarr, int nsys, double* out)
{
  const McRun& R = *Rp;
  // the wavefront index is uniform over the wavefront; readfirstlane tells the compiler so, and the per-system
  // structures are then read through the scalar cache instead of per-lane vector loads
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)), lane = threadIdx.x & 63;
  const int sys = blockIdx.x * 4 + w;
  if (sys >= nsys) return;
  const SysDev& S = Sarr[sys];
  const McSys& Y = Yarr[sys];
  double e = 0.0;
  for (int m = 0; m < Y.nmol; ++m)
  {
    const Comp& c = R.comp[Y.comp[m]];
    for (int a = 0; a < c.na; ++a)
    {
      const int a0 = m * R.slots + a;
      bool o;
      // pairs with molecules of lower index only, so each inter-molecular pair counts once
      double ef = siteExternal(S, 0, S.ax[a0], S.ay[a0], S.az[a0], c.type[a], c.q[a], m, -2, lane, o);
      double ei = 0.0;
      for (int b = lane; b < m * R.slots; b += 64)
      {
        const int t = S.at[b];
        if (t < 0) continue;
        double dx = S.ax[b] - S.ax[a0], dy = S.ay[b] - S.ay[a0], dz = S.az[b] - S.az[a0];
        rgd::pbc(S.box, dx, dy, dz);
        const double rr = dx * dx + dy * dy + dz * dz;
        if (rr < S.rcv2Mol) ei += rgd::vdw0(S.pairs[t * S.ntypes + c.type[a]], 1.0, rr);
        if (S.useCharge && rr < S.rcc2) ei += rgd::coul0(S.alpha, 1.0, sqrt(rr), S.coulomb * S.aq[b] * c.q[a]);
      }
      e += ef + rgd::waveSum(ei);
    }
    if (c.flexible)
    {
      // intra-molecular energy of the chain: bend, torsion and LJ between beads more than three bonds apart
      double mol[3 * MC_MAXATOMS];
      for (int a = 0; a < c.na; ++a)
      {
        mol[3 * a] = S.ax[m * R.slots + a];
        mol[3 * a + 1] = S.ay[m * R.slots + a];
        mol[3 * a + 2] = S.az[m * R.slots + a];
      }
      for (int i = 1; i < c.na; ++i)
      {
        bool o = false;
        e += chainIntra(S, c, mol, i, mol + 3 * i, o) + chainBonded(c, mol, i);
      }
    }
  }
  // Ewald Fourier energy of the adsorbates from the stored structure factor
  double ef = 0.0;
  if (S.useCharge)
  {
    for (int k = lane; k < S.nk; k += 64)
    {
      const double2 s = S.Sk[k], f = S.Sfw[k];
      ef += S.fk[k] * ((s.x * s.x + s.y * s.y) - (f.x * f.x + f.y * f.y));
    }
    ef = rgd::waveSum(ef);
    double netAds = 0.0;
    for (int m = 0; m < Y.nmol; ++m)
    {
      const Comp& c = R.comp[Y.comp[m]];
      ef -= c.eIntra;
      netAds += c.qsum;
    }
    const double fwNet = S.netCharge - netAds;
    const double uIon = -(S.singleIon - S.coulomb * S.alpha / sqrt(M_PI));
    ef += uIon * (2.0 * fwNet + netAds) * netAds;
  }
  if (lane == 0) out[sys] = e + ef;
}

template __global__ void k_widom<false>(const McRun*, const SysDev*, const McSys*, int, long long, int, double*,
                                        double*, double*);
template __global__ void k_widom<true>(const McRun*, const SysDev*, const McSys*, int, long long, int, double*,
                                       double*, double*);
template __global__ void k_gcmc<false, false, false>(const McRun*, SysDev*, McSys*, int, int, int, int, const McCtl*);
template __global__ void k_gcmc<false, false, true>(const McRun*, SysDev*, McSys*, int, int, int, int, const McCtl*);
template __global__ void k_gcmc<false, true, false>(const McRun*, SysDev*, McSys*, int, int, int, int, const McCtl*);
template __global__ void k_gcmc<false, true, true>(const McRun*, SysDev*, McSys*, int, int, int, int, const McCtl*);
template __global__ void k_gcmc<true, false, false>(const McRun*, SysDev*, McSys*, int, int, int, int, const McCtl*);
template __global__ void k_gcmc<true, false, true>(const McRun*, SysDev*, McSys*, int, int, int, int, const McCtl*);

}  // namespace rg
